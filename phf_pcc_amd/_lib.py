"""ctypes binding of libphf_hip.so (declarations in include/phf_hip.h).

The library is required: importing the product without it raises, there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PHF_HIP_LIB", os.path.join(HERE, "lib", "libphf_hip.so"))

PHF_OK = 0
PHF_ERR_ARG = 1
PHF_ERR_HIP = 2
PHF_ERR_SINGULAR = 3
PHF_ERR_NOT_HERMITIAN = 4
PHF_ERR_NO_HAMILTONIAN = 5
PHF_ERR_UNSUPPORTED = 6
PHF_ERR_NOMEM = 7

PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2, "f16x2": 3, "i8x4": 4}
PHF_HAM_COMPLEX = 1
PHF_HAM_DENSE = 2
PHF_HAM_BLOCKS_ONLY = 4

# every symbol declared in include/phf_hip.h
EXPORTS = ["phf_ctx_create", "phf_ctx_destroy", "phf_last_error", "phf_set_hamiltonian",
           "phf_buildhsg", "phf_localg", "phf_set_profiling", "phf_stage_times", "phf_set_grid",
           "phf_buildhsg_z", "phf_autotune", "phf_tuning_report", "phf_set_hamiltonian_factored",
           "phf_get_hone", "phf_eval_resident", "phf_localg_vo", "phf_batch_create", "phf_batch_destroy",
           "phf_batch_last_error", "phf_batch_set_instance", "phf_batch_eval", "phf_batch_set_graph",
           "phf_batch_autotune", "phf_batch_stats", "phf_batch_tuning_report", "phf_set_blocked", "phf_blocked",
           "phf_set_spinfree", "phf_spinfree",
           "phf_buildsc", "phf_sc_set_grid", "phf_sc_apply", "phf_sc_apply_packed", "phf_sc_apply_packed_dev",
           "phf_sd_dim", "phf_sc_set_points", "phf_eval_objective", "phf_invert_batched",
           "phf_invert_batched_noalloc", "phf_pcc_olap0", "phf_pcc_olap"]


class PhfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[phf error {code}] {msg}")
        self.code = code


class SingularOverlapError(PhfError, np.linalg.LinAlgError):
    pass


class NotHermitianError(PhfError):
    pass


_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libphf_hip.so not found at {LIB_PATH}; run phf_pcc_amd.build_lib.build() first")
    lib = C.CDLL(LIB_PATH)
    lib.phf_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    lib.phf_ctx_destroy.argtypes = [C.c_void_p]
    lib.phf_last_error.argtypes = [C.c_void_p]
    lib.phf_last_error.restype = C.c_char_p
    lib.phf_set_hamiltonian.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_int]
    lib.phf_buildhsg.argtypes = [C.c_void_p, _dp, _dp, _ip, C.c_int, C.c_int, C.c_int, C.c_int,
                                 _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int,
                                 _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                 _dp, _dp, _dp, _dp]
    lib.phf_localg.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int,
                               C.c_int, C.c_int, C.c_int, _dp]
    lib.phf_set_grid.argtypes = [C.c_void_p, _dp, _ip, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int,
                                 _dp, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int]
    lib.phf_buildhsg_z.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
    lib.phf_set_hamiltonian_factored.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_int, _dp, C.c_int]
    lib.phf_get_hone.argtypes = [C.c_void_p, _dp]
    lib.phf_eval_resident.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, _dp, _dp]
    lib.phf_localg_vo.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp]
    lib.phf_autotune.argtypes = [C.c_void_p, C.c_int]
    lib.phf_tuning_report.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
    lib.phf_set_profiling.argtypes = [C.c_void_p, C.c_int]
    lib.phf_stage_times.argtypes = [C.c_void_p, _dp, C.c_int]
    lib.phf_set_blocked.argtypes = [C.c_void_p, C.c_int]
    lib.phf_blocked.argtypes = [C.c_void_p]
    lib.phf_set_spinfree.argtypes = [C.c_void_p, C.c_int]
    lib.phf_spinfree.argtypes = [C.c_void_p]
    lib.phf_batch_create.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    lib.phf_batch_destroy.argtypes = [C.c_void_p]
    lib.phf_batch_last_error.argtypes = [C.c_void_p]
    lib.phf_batch_last_error.restype = C.c_char_p
    lib.phf_batch_set_instance.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp, C.c_int, _ip,
                                           _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int, C.c_int]
    lib.phf_batch_eval.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, _dp, _dp, _dp, _ip]
    lib.phf_batch_set_graph.argtypes = [C.c_void_p, C.c_int]
    lib.phf_batch_autotune.argtypes = [C.c_void_p, C.c_int]
    lib.phf_batch_stats.argtypes = [C.c_void_p, _dp, C.c_int]
    lib.phf_batch_tuning_report.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
    lib.phf_buildsc.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_int, _ip, C.c_int, _dp, C.c_int, _dp,
                                C.c_int, _dp, C.c_int, _dp, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, _dp]
    lib.phf_sc_set_grid.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp,
                                    C.c_int, _dp, _dp, C.c_int, _dp, C.c_int, _dp]
    lib.phf_sc_apply.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.phf_sc_apply_packed.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    lib.phf_sc_apply_packed_dev.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    lib.phf_sd_dim.argtypes = [C.c_int, C.c_int]
    lib.phf_sc_set_points.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp]
    lib.phf_eval_objective.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, _dp, _dp]
    lib.phf_invert_batched.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_int)]
    lib.phf_invert_batched_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_int)]
    lib.phf_pcc_olap0.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _ip,
                                  C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp,
                                  C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.phf_pcc_olap.argtypes = lib.phf_pcc_olap0.argtypes[:-2] + [_dp] * 6
    for name in EXPORTS:
        if name not in ("phf_last_error", "phf_batch_last_error"):
            getattr(lib, name).restype = C.c_int
    lib.phf_sd_dim.restype = C.c_long
    _lib = lib
    return lib


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def check(rc: int, ctx, batch: bool = False) -> None:
    """Raise the Python exception of status rc (message from the context, or
    from the batch handle when batch is True)."""
    if rc == PHF_OK:
        return
    lib = load()
    msg = (lib.phf_batch_last_error(ctx) if batch else lib.phf_last_error(ctx)).decode(errors="replace")
    if rc == PHF_ERR_SINGULAR:
        raise SingularOverlapError(rc, msg)
    if rc == PHF_ERR_NOT_HERMITIAN:
        raise NotHermitianError(rc, msg)
    if rc == PHF_ERR_ARG:
        raise ValueError(msg)
    raise PhfError(rc, msg)


class Context:
    """One GPU, one HIP stream, one cached device Hamiltonian."""

    def __init__(self, device: int = 0):
        lib = load()
        h = C.c_void_p()
        rc = lib.phf_ctx_create(int(device), C.byref(h))
        if rc != PHF_OK:
            raise PhfError(rc, lib.phf_last_error(None).decode(errors="replace"))
        self._h = h
        self.device = device
        self.ham_key = None
        self.nso = None

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            load().phf_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_hamiltonian_factored(self, h: np.ndarray, b: np.ndarray, ref: np.ndarray, key=None, flags: int = 0):
        """Build A on the device from (pq|rs) = sum_L b[L,p,q] b[L,r,s] and a real Ref
        (flags: PHF_HAM_DENSE, PHF_HAM_BLOCKS_ONLY)."""
        h = np.ascontiguousarray(h, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        ref = np.ascontiguousarray(ref, dtype=np.complex128)
        n = h.shape[0]
        if h.shape != (n, n) or b.ndim != 3 or b.shape[1:] != (n, n) or ref.shape != (2 * n, 2 * n):
            raise ValueError("h (n, n), b (L, n, n) and ref (2n, 2n) expected")
        check(load().phf_set_hamiltonian_factored(self._h, ptr(h), ptr(b), n, b.shape[0], ptr(ref), int(flags)),
              self._h)
        self.nso = 2 * n
        self.ham_key = key

    def invert_batched(self, m: np.ndarray):
        """Inverses, determinants and info of complex[batch][n][n] matrices
        (InvertC + determinantC on the kernel of the grid's overlaps)."""
        m = np.ascontiguousarray(m, dtype=np.complex128)
        if m.ndim != 3 or m.shape[1] != m.shape[2]:
            raise ValueError("invert_batched: m must be [batch][n][n]")
        b, n = m.shape[0], m.shape[1]
        inv = np.zeros_like(m)
        det = np.zeros(b, np.complex128)
        info = np.zeros(b, np.int32)
        check(load().phf_invert_batched(self._h, n, b, ptr(m), ptr(inv), ptr(det),
                                        info.ctypes.data_as(C.POINTER(C.c_int))), self._h)
        return inv, det, info

    def invert_batched_noalloc(self, m: np.ndarray):
        """invert_batched through phf_invert_batched_noalloc (no device
        allocation: staging in the library's 64 MiB scratch array)."""
        m = np.ascontiguousarray(m, dtype=np.complex128)
        if m.ndim != 3 or m.shape[1] != m.shape[2]:
            raise ValueError("invert_batched_noalloc: m must be [batch][n][n]")
        b, n = m.shape[0], m.shape[1]
        inv = np.zeros_like(m)
        det = np.zeros(b, np.complex128)
        info = np.zeros(b, np.int32)
        check(load().phf_invert_batched_noalloc(self._h, n, b, ptr(m), ptr(inv), ptr(det),
                                                info.ctypes.data_as(C.POINTER(C.c_int))), self._h)
        return inv, det, info

    def hone(self) -> np.ndarray:
        out = np.zeros((self.nso, self.nso), dtype=np.complex128)
        check(load().phf_get_hone(self._h, ptr(out)), self._h)
        return out

    def set_profiling(self, on: bool = True):
        check(load().phf_set_profiling(self._h, 1 if on else 0), self._h)

    def autotune(self, on: bool = True):
        check(load().phf_autotune(self._h, 1 if on else 0), self._h)

    def tuning_report(self) -> str:
        buf = C.create_string_buffer(1 << 20)
        n = load().phf_tuning_report(self._h, buf, len(buf))
        return buf.raw[:n].decode()

    def set_blocked(self, on: bool):
        """Evaluate through the spin blocks of A (True, the default when they
        exist) or through the dense A (False: the plain contraction)."""
        check(load().phf_set_blocked(self._h, int(bool(on))), self._h)

    def blocked(self) -> bool:
        return bool(load().phf_blocked(self._h))

    def set_spinfree(self, on: bool):
        """Evaluate through the spin-free (2J - K, K) form when it is built
        (True, the default) or through the spin blocks (False)."""
        check(load().phf_set_spinfree(self._h, int(bool(on))), self._h)

    def spinfree(self) -> bool:
        return bool(load().phf_spinfree(self._h))

    def stage_times(self) -> np.ndarray:
        out = np.zeros(11)
        n = load().phf_stage_times(self._h, ptr(out), 11)
        return out[:n].copy()

    # ------------------------------------------------------------------
    # projected SD-space overlap (BuildSC, PHFTools.f90:464-537)
    def sc_set_grid(self, nso: int, nocc: int, ngrid, qj: int, r1, r2, rpg, weightsp, weightpg, fsp, fpg):
        """Upload a projection grid and build its per-point factors on the device."""
        ngrid = np.ascontiguousarray(np.asarray(ngrid, dtype=np.int32).reshape(-1)[:2])
        nl, ny = int(ngrid[0]), int(ngrid[1])
        r1 = np.ascontiguousarray(r1, dtype=np.complex128)
        r2 = np.ascontiguousarray(r2, dtype=np.complex128)
        rpg = np.ascontiguousarray(rpg, dtype=np.complex128)
        weightsp = np.ascontiguousarray(np.asarray(weightsp, dtype=np.float64)[:nl, :ny])
        weightpg = np.ascontiguousarray(weightpg, dtype=np.float64)
        fsp = np.ascontiguousarray(np.asarray(fsp, dtype=np.complex128).reshape(-1))
        fpg = np.ascontiguousarray(np.asarray(fpg, dtype=np.complex128).reshape(-1))
        npg, ncipg = rpg.shape[0], weightpg.shape[1]
        if (r1.shape[1:] != (nso, nso) or r2.shape[1:] != (nso, nso) or rpg.shape[1:] != (nso, nso)
                or weightpg.shape != (npg, ncipg, ncipg) or fpg.size < ncipg or weightsp.shape != (nl, ny)):
            raise ValueError("sc_set_grid: inconsistent grid shapes")
        check(load().phf_sc_set_grid(self._h, nso, nocc, ngrid.ctypes.data_as(_ip), int(qj), ptr(r1), r1.shape[0], ptr(r2),
                                     r2.shape[0], ptr(rpg), npg, ptr(weightsp), ptr(weightpg), ncipg, ptr(fsp),
                                     fsp.size, ptr(fpg)), self._h)
        self.sc_shape = (nso, nocc)

    def sc_set_points(self, y, r=None, w=None):
        """Points of given rotation r, transition density y and weight w
        (stacked along a leading axis; r, w default to identities and ones)."""
        y = np.ascontiguousarray(y, dtype=np.complex128)
        if y.ndim == 2:
            y = y[None]
        npts, nso = y.shape[0], y.shape[1]
        if y.shape != (npts, nso, nso):
            raise ValueError("sc_set_points: y must be (npts, nso, nso)")
        if r is not None:
            r = np.ascontiguousarray(np.broadcast_to(r, y.shape), dtype=np.complex128)
        if w is not None:
            w = np.ascontiguousarray(np.asarray(w, dtype=np.complex128).reshape(npts))
        check(load().phf_sc_set_points(self._h, nso, npts, ptr(r), ptr(y), ptr(w)), self._h)
        self.sc_shape = (nso, nso)

    def sc_apply(self, c0, c1, c2):
        """(t0, t1, t2) for stacked inputs c0 [nv], c1 [nv, nso, nso], c2 [nv, nso^4]."""
        nso = self.sc_shape[0]
        c0 = np.ascontiguousarray(np.asarray(c0, dtype=np.complex128).reshape(-1))
        nv = c0.size
        c1 = np.ascontiguousarray(np.asarray(c1, dtype=np.complex128).reshape(nv, nso, nso))
        c2 = np.ascontiguousarray(np.asarray(c2, dtype=np.complex128).reshape((nv,) + (nso,) * 4))
        t0 = np.empty(nv, np.complex128)
        t1 = np.empty_like(c1)
        t2 = np.empty_like(c2)
        check(load().phf_sc_apply(self._h, nv, ptr(c0), ptr(c1), ptr(c2), ptr(t0), ptr(t1), ptr(t2)), self._h)
        return t0, t1, t2

    def sc_apply_packed(self, vecs: np.ndarray) -> np.ndarray:
        """Packed SD vectors [nv, dim] (LinDep.py:156-187 order) -> their products."""
        v = np.ascontiguousarray(vecs, dtype=np.complex128)
        one = v.ndim == 1
        v = v.reshape(1, -1) if one else v
        out = np.empty_like(v)
        check(load().phf_sc_apply_packed(self._h, v.shape[0], ptr(v), ptr(out)), self._h)
        return out[0] if one else out

    def set_hamiltonian(self, hone: np.ndarray, htwo: np.ndarray, force_complex: bool = False, key=None,
                        dense: bool = False):
        hone = np.ascontiguousarray(hone, dtype=np.complex128)
        htwo = np.ascontiguousarray(htwo, dtype=np.complex128)
        nso = hone.shape[0]
        if hone.shape != (nso, nso) or htwo.shape != (nso,) * 4:
            raise ValueError("hone must be (nso, nso) and htwo (nso, nso, nso, nso)")
        flags = (PHF_HAM_COMPLEX if force_complex else 0) | (PHF_HAM_DENSE if dense else 0)
        check(load().phf_set_hamiltonian(self._h, ptr(hone), ptr(htwo), nso, flags), self._h)
        self.nso = nso
        self.ham_key = key
