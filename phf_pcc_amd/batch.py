"""Batched projected-HF objective: many independent instances in one launch sequence.

Each instance is one EandG (PHF.py:55-91) at J = 0 (nci = 1) on its own
factored Hamiltonian ((pq|rs) = sum_L b[L,p,q] b[L,r,s], real h, real Ref), its
own electron count and its own spin grid.  ``ProjectedHFBatch.energy_and_gradient``
evaluates all of them with one call of ``phf_batch_eval``: per-instance
rotation/overlap/density stages, one persistent grouped ERI GEMM over every
instance (16-bit forms), per-instance epilogues, all replayed from a hipGraph
after the first call.  With nci = 1 the EigenSolver/VecDecomp step
(PHFTools.py:8-26) reduces to E0 = Re H / |Re S| and LocalG (PHFTools.f90:203-247)
to (Gradmat - E0 Rdmmat) / Smat, which the device finishes itself.

``BatchEnergies.apply(x, batch)`` is the differentiable form: x concatenates
[Re Z_i, Im Z_i] of every instance; the output is the vector of energies and
backward applies the per-instance adjoint gradients 2 [Re G_i, Im G_i].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class ProjectedHFBatch:
    def __init__(self, instances, storage: str = "fp64", device: int = 0, graph: bool = True, structure: bool = True):
        """instances: sequence of dicts with keys h (n x n), b (L x n x n), ref
        (2n x 2n real), nocc, and the J = 0 grid arguments ngrid, r1, r2, rpg,
        weightsp and optionally weightpg (npg values).  storage: "fp64" keeps
        the FP64 A of every instance (all precisions available); a reduced
        mode keeps only that copy of A (evaluations in that mode only).
        structure: store a spin-block structured A as its blocks (False: the
        dense A, the plain contraction)."""
        lib = _lib.load()
        self.n_inst = len(instances)
        h = C.c_void_p()
        rc = lib.phf_batch_create(int(device), self.n_inst, C.byref(h))
        if rc != _lib.PHF_OK:
            raise _lib.PhfError(rc, "phf_batch_create failed")
        self._h = h
        self.storage = storage
        self.nso, self.nocc = [], []
        for i, inst in enumerate(instances):
            hh = np.ascontiguousarray(inst["h"], dtype=np.float64)
            bb = np.ascontiguousarray(inst["b"], dtype=np.float64)
            n = hh.shape[0]
            ref = np.ascontiguousarray(inst["ref"], dtype=np.complex128)
            ngrid = np.ascontiguousarray(inst["ngrid"], dtype=np.int32)
            r1 = np.ascontiguousarray(inst["r1"], dtype=np.complex128)
            r2 = np.ascontiguousarray(inst["r2"], dtype=np.complex128)
            rpg = np.ascontiguousarray(inst["rpg"], dtype=np.complex128)
            wsp = np.ascontiguousarray(inst["weightsp"], dtype=np.float64)
            wpg = inst.get("weightpg")
            wpg = None if wpg is None else np.ascontiguousarray(np.reshape(wpg, -1), dtype=np.float64)
            if wpg is not None and wpg.size != rpg.shape[0]:
                raise ValueError("weightpg must hold one weight per Rpg matrix (ncipg = 1)")
            rc = lib.phf_batch_set_instance(self._h, i, _lib.ptr(hh), _lib.ptr(bb), n, bb.shape[0], _lib.ptr(ref),
                                            int(inst["nocc"]), ngrid.ctypes.data_as(_lib._ip), _lib.ptr(r1),
                                            r1.shape[0], _lib.ptr(r2), r2.shape[0], _lib.ptr(rpg), rpg.shape[0],
                                            _lib.ptr(wsp), _lib.ptr(wpg), _lib.PREC[storage],
                                            0 if structure else _lib.PHF_HAM_DENSE)
            _lib.check(rc, self._h, batch=True)
            self.nso.append(2 * n)
            self.nocc.append(int(inst["nocc"]))
        self.shapes = [(s - o, o) for s, o in zip(self.nso, self.nocc)]
        self.sizes = [a * b for a, b in self.shapes]
        self.set_graph(graph)

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().phf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph(self, on: bool):
        _lib.check(_lib.load().phf_batch_set_graph(self._h, int(bool(on))), self._h, batch=True)

    def autotune(self, on: bool):
        _lib.check(_lib.load().phf_batch_autotune(self._h, int(bool(on))), self._h, batch=True)

    def energy_and_gradient(self, zs, precision: str | None = None, cmplxconj: int = 1, strict: bool = True):
        """zs: list of complex (2n_i - N_i) x N_i arrays.  Returns (energies
        float64[n_inst], list of gradient blocks dE/dZ*_i, hs (n_inst, 4)).
        An instance that fails (singular overlap, Hermiticity test) raises
        when strict; otherwise self.status holds each instance's code."""
        if len(zs) != self.n_inst:
            raise ValueError(f"expected {self.n_inst} Z blocks")
        for z, shp in zip(zs, self.shapes):
            if np.shape(z) != shp:
                raise ValueError(f"Z block of shape {np.shape(z)}, expected {shp}")
        zcat = np.ascontiguousarray(np.concatenate([np.asarray(z, dtype=np.complex128).ravel() for z in zs]))
        prec = precision or (self.storage if self.storage != "fp64" else "fp64")
        e = np.zeros(self.n_inst)
        g = np.zeros(zcat.size, dtype=np.complex128)
        hs = np.zeros((self.n_inst, 4))
        self.status = np.zeros(self.n_inst, dtype=np.int32)
        rc = _lib.load().phf_batch_eval(self._h, _lib.ptr(zcat), int(cmplxconj), _lib.PREC[prec], _lib.ptr(e),
                                        _lib.ptr(g), _lib.ptr(hs), self.status.ctypes.data_as(_lib._ip))
        if strict or rc not in (_lib.PHF_ERR_NOT_HERMITIAN, _lib.PHF_ERR_SINGULAR):
            _lib.check(rc, self._h, batch=True)
        grads, off = [], 0
        for shp, sz in zip(self.shapes, self.sizes):
            gi = g[off:off + sz].reshape(shp)
            # CmplxConj = 0: the objective ignores Im Z (the kernel takes Re Z),
            # and EandG returns a real gradient (PHF.py:80-81)
            grads.append(gi.real.astype(np.complex128) if int(cmplxconj) == 0 else gi)
            off += sz
        return e, grads, hs

    def stats(self):
        """(device ms of the last evaluation, its ERI GEMM ms or -1 for a graph
        replay, 1.0 if it was a graph replay)."""
        out = np.zeros(4)
        _lib.load().phf_batch_stats(self._h, _lib.ptr(out), 4)
        return tuple(out[:3])

    def spinfree_instances(self) -> int:
        """Instances evaluated through the spin-free (2J - K, K) form."""
        out = np.zeros(4)
        _lib.load().phf_batch_stats(self._h, _lib.ptr(out), 4)
        return int(out[3])

    def tuning_report(self) -> str:
        buf = C.create_string_buffer(1 << 16)
        n = _lib.load().phf_batch_tuning_report(self._h, buf, len(buf))
        return buf.raw[:n].decode()


class BatchEnergies:
    """energies = BatchEnergies.apply(x, batch, precision) as a torch.autograd.Function."""
    _fn = None

    @classmethod
    def apply(cls, x, batch: ProjectedHFBatch, precision: str | None = None):
        if cls._fn is None:
            import torch

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, xx, bt, prec):
                    xv = xx.detach().cpu().numpy().astype(np.float64)
                    zs, off = [], 0
                    for shp, sz in zip(bt.shapes, bt.sizes):
                        zs.append((xv[off:off + sz] + 1j * xv[off + sz:off + 2 * sz]).reshape(shp))
                        off += 2 * sz
                    e, g, _ = bt.energy_and_gradient(zs, prec)
                    jac = np.concatenate([np.concatenate([2 * gi.real.ravel(), 2 * gi.imag.ravel()]) for gi in g])
                    ctx.save_for_backward(torch.as_tensor(jac, dtype=xx.dtype, device=xx.device))
                    ctx.sizes = bt.sizes
                    return torch.as_tensor(e, dtype=xx.dtype, device=xx.device)

                @staticmethod
                def backward(ctx, ge):
                    (jac,) = ctx.saved_tensors
                    reps = torch.as_tensor([2 * s for s in ctx.sizes], device=jac.device)
                    return jac * torch.repeat_interleave(ge, reps), None, None

            cls._fn = _Fn
        return cls._fn.apply(x, batch, precision)
