"""Drop-in for the f2py extension ``PHFTools_F`` (PHF_fort/makefile:7).

``from phf_pcc_amd.PHFTools_F import phftools`` replaces
``from PHFTools_F import phftools`` (PHF.py:10).  The calling convention is
the one f2py gives the Fortran module procedures (call sites PHF.py:70-71,76):

  hmat, smat, gradmat, rdmmat = phftools.buildhsg(
      hone, htwo, ref, z, ngrid, nci, cmplxconj, ncisp, nocc, r1, r2, rpg,
      weightsp, weightpg, roota, rootb, rooty, qj, sp, comm, [npg, ncipg, nso])
  g = phftools.localg(gradmat, rdmmat, smat, e0, fsp, fpg, fk, cmplxconj,
                      [nso, nci, ncisp, ncipg])
  t0, t1, t2 = phftools.buildsc(c0, c1, c2, nocc, ngrid, qj, r1, r2, rpg, rk,
      roota, rootb, rooty, weightsp, weightpg, fsp, fpg, fk,
      [nso, npg, ncisp, ncipg, ncik])                      (LinDep.py:73)
  t0, t1, t2 = buildsc_dr.buildsc_vec(c0, c1, c2, y, [nso])  (BuildSC_dr.f90:8)

Inputs are cast to complex128 / float64 / int as f2py would; a shape that
disagrees with the explicit dimension arguments raises ValueError.  Numerical
failures that STOP the reference process raise instead:
``SingularOverlapError`` (a numpy LinAlgError; Wrappers.f90:311) and
``NotHermitianError`` (PHFTools.f90:195-200).  ``comm`` (an MPI handle) is
accepted and ignored: one context drives one GPU.

The device keeps the last Hamiltonian (the reference re-reads HTwo on every
call; re-uploading a multi-GB array per call would cost more than the
evaluation).  It is re-uploaded when ``hone`` or ``htwo`` differs from the
cached pair in shape, dtype or content: arrays up to 256 MiB
(``PHF_CACHE_FULL_HASH_MAX``) are compared by a hash of every byte; larger
ones by address plus a hash of a sample of elements, unless
``PHF_STRICT_CACHE=1`` asks for the full hash on every call.  A caller that
edits a larger array in place without PHF_STRICT_CACHE must call
``phftools.set_hamiltonian`` (uploads explicitly) or ``phftools.invalidate()``
(drops the cache).

``phftools.precision`` selects the ERI-contraction precision (default
"fp64", the reference's arithmetic).
"""
from __future__ import annotations

import hashlib
import os

import numpy as np

from . import _lib


try:
    import xxhash as _xxhash
except ImportError:  # pragma: no cover - the image ships xxhash
    _xxhash = None

# Arrays up to this many bytes are hashed in full on every call (xxh3-128
# streams about 8 GB/s, so 256 MiB costs about 30 ms).  Larger arrays are
# hashed in full only with PHF_STRICT_CACHE=1; otherwise their key is the
# address plus a sample of elements, and a caller that edits such an array in
# place must call phftools.set_hamiltonian or phftools.invalidate().
FULL_HASH_MAX_BYTES = int(os.environ.get("PHF_CACHE_FULL_HASH_MAX", str(256 << 20)))


def _digest(buf) -> str:
    if _xxhash is not None:
        return _xxhash.xxh3_128_hexdigest(buf)
    return hashlib.blake2b(buf, digest_size=16).hexdigest()


def _fingerprint(a: np.ndarray) -> tuple:
    a = np.asarray(a)
    flat = a.reshape(-1)
    strict = os.environ.get("PHF_STRICT_CACHE", "0") not in ("", "0")
    if a.nbytes <= FULL_HASH_MAX_BYTES or strict:
        # exact: every byte of the content (the address does not matter)
        return (a.shape, a.dtype.str, _digest(memoryview(np.ascontiguousarray(flat)).cast("B")))
    n = flat.size
    h = hashlib.blake2b(digest_size=16)
    # sampled: 2048 evenly spaced elements plus both ends, keyed with the address
    h.update(np.ascontiguousarray(flat[:1024]).tobytes())
    h.update(np.ascontiguousarray(flat[n - 1024:]).tobytes())
    h.update(np.ascontiguousarray(flat[:: max(1, n // 2048)]).tobytes())
    return (a.__array_interface__["data"][0], a.shape, a.dtype.str, h.hexdigest())


class _PhfTools:
    def __init__(self):
        self._ctx = None
        self.precision = "fp64"
        self.device = int(os.environ.get("PHF_DEVICE", "0"))

    # ------------------------------------------------------------------
    def context(self) -> _lib.Context:
        if self._ctx is None:
            self._ctx = _lib.Context(self.device)
        return self._ctx

    def invalidate(self):
        if self._ctx is not None:
            self._ctx.ham_key = None
            self._ctx.grid_key = None
            self._ctx.sc_key = None

    def set_hamiltonian(self, hone, htwo):
        key = (_fingerprint(np.asarray(hone)), _fingerprint(np.asarray(htwo)))
        self.context().grid_key = None
        self.context().set_hamiltonian(hone, htwo, key=key)

    def _ensure_ham(self, hone, htwo):
        ctx = self.context()
        key = (_fingerprint(hone), _fingerprint(htwo))
        if ctx.ham_key != key:
            ctx.grid_key = None
            ctx.set_hamiltonian(hone, htwo, key=key)

    # ------------------------------------------------------------------
    def buildhsg(self, hone, htwo, ref, z, ngrid, nci, cmplxconj, ncisp, nocc, r1, r2, rpg,
                 weightsp, weightpg, roota, rootb, rooty, qj, sp, comm=None,
                 npg=None, ncipg=None, nso=None):
        """BuildHSG (PHFTools.f90:10-201) on the GPU."""
        hone = np.asarray(hone)
        htwo = np.asarray(htwo)
        nso = hone.shape[0] if nso is None else int(nso)
        nocc = int(nocc)
        rpg = np.ascontiguousarray(rpg, dtype=np.complex128)
        weightpg = np.ascontiguousarray(weightpg, dtype=np.float64)
        npg = rpg.shape[0] if npg is None else int(npg)
        ncipg = weightpg.shape[1] if ncipg is None else int(ncipg)
        ngrid = np.ascontiguousarray(ngrid, dtype=np.int32).reshape(-1)
        if ngrid.size != 2:
            raise ValueError("ngrid must have 2 entries")
        ref = np.ascontiguousarray(ref, dtype=np.complex128)
        z = np.ascontiguousarray(z, dtype=np.complex128)
        r1 = np.ascontiguousarray(r1, dtype=np.complex128)
        r2 = np.ascontiguousarray(r2, dtype=np.complex128)
        weightsp = np.ascontiguousarray(weightsp, dtype=np.float64)
        roota = np.ascontiguousarray(roota, dtype=np.float64).reshape(-1)
        rootb = np.ascontiguousarray(rootb, dtype=np.float64).reshape(-1)
        rooty = np.ascontiguousarray(rooty, dtype=np.float64).reshape(-1)
        nl, ny = int(ngrid[0]), int(ngrid[1])
        checks = [
            (hone.shape == (nso, nso), "hone must be (nso, nso)"),
            (htwo.shape == (nso,) * 4, "htwo must be (nso, nso, nso, nso)"),
            (ref.shape == (nso, nso), "ref must be (nso, nso)"),
            (z.shape == (nso - nocc, nocc), "z must be (nso-nocc, nocc)"),
            (r1.ndim == 3 and r1.shape[1:] == (nso, nso), "r1 must be (*, nso, nso)"),
            (r2.ndim == 3 and r2.shape[1:] == (nso, nso), "r2 must be (*, nso, nso)"),
            (rpg.shape == (npg, nso, nso), "rpg must be (npg, nso, nso)"),
            (weightpg.shape == (npg, ncipg, ncipg), "weightpg must be (npg, ncipg, ncipg)"),
            (weightsp.ndim == 2 and weightsp.shape[0] >= nl and weightsp.shape[1] >= ny,
             "weightsp must cover (ngrid[0], ngrid[1])"),
        ]
        for ok, msg in checks:
            if not ok:
                raise ValueError(f"buildhsg: {msg}")
        if weightsp.shape != (nl, ny):
            weightsp = np.ascontiguousarray(weightsp[:nl, :ny])
        self._ensure_ham(hone, htwo)
        ctx = self.context()
        nci = int(nci)
        lib = _lib.load()
        gkey = (ctx.ham_key, _fingerprint(ref), tuple(int(v) for v in ngrid), nci, int(ncisp), _fingerprint(r1),
                _fingerprint(r2), _fingerprint(rpg), _fingerprint(weightsp), _fingerprint(weightpg),
                _fingerprint(roota), _fingerprint(rootb), _fingerprint(rooty), int(qj), int(sp))
        if getattr(ctx, "grid_key", None) != gkey:
            ctx.grid_key = None
            _lib.check(lib.phf_set_grid(ctx.handle, _lib.ptr(ref), ngrid.ctypes.data_as(_lib._ip), nci, int(ncisp),
                                        _lib.ptr(r1), r1.shape[0], _lib.ptr(r2), r2.shape[0], _lib.ptr(rpg), npg,
                                        _lib.ptr(weightsp), _lib.ptr(weightpg), ncipg, _lib.ptr(roota), roota.size,
                                        _lib.ptr(rootb), rootb.size, _lib.ptr(rooty), rooty.size, int(qj),
                                        int(sp)), ctx.handle)
            ctx.grid_key = gkey
        hmat = np.zeros((nci, nci), dtype=np.complex128)
        smat = np.zeros((nci, nci), dtype=np.complex128)
        gradmat = np.zeros((nci, nci, nso, nso), dtype=np.complex128)
        rdmmat = np.zeros((nci, nci, nso, nso), dtype=np.complex128)
        prec = _lib.PREC[self.precision]
        rc = lib.phf_buildhsg_z(ctx.handle, _lib.ptr(z), int(cmplxconj), nocc, prec, _lib.ptr(hmat), _lib.ptr(smat),
                                _lib.ptr(gradmat), _lib.ptr(rdmmat))
        _lib.check(rc, ctx.handle)
        return hmat, smat, gradmat, rdmmat

    def localg(self, gradmat, rdmmat, smat, e0, fsp, fpg, fk, cmplxconj,
               nso=None, nci=None, ncisp=None, ncipg=None):
        """LocalG (PHFTools.f90:203-247) on the GPU."""
        gradmat = np.ascontiguousarray(gradmat, dtype=np.complex128)
        rdmmat = np.ascontiguousarray(rdmmat, dtype=np.complex128)
        smat = np.ascontiguousarray(smat, dtype=np.complex128)
        fsp = np.ascontiguousarray(fsp, dtype=np.complex128).reshape(-1)
        fpg = np.ascontiguousarray(fpg, dtype=np.complex128).reshape(-1)
        fk = np.ascontiguousarray(fk, dtype=np.complex128).reshape(-1)
        nso = gradmat.shape[2] if nso is None else int(nso)
        nci = gradmat.shape[0] if nci is None else int(nci)
        ncisp = fsp.size if ncisp is None else int(ncisp)
        ncipg = fpg.size if ncipg is None else int(ncipg)
        if gradmat.shape != (nci, nci, nso, nso) or rdmmat.shape != gradmat.shape or smat.shape != (nci, nci):
            raise ValueError("localg: gradmat/rdmmat must be (nci, nci, nso, nso) and smat (nci, nci)")
        if fsp.size < ncisp or fpg.size < ncipg:
            raise ValueError("localg: fsp/fpg shorter than ncisp/ncipg")
        e0a = np.array([complex(e0)], dtype=np.complex128)
        g = np.zeros((nso, nso), dtype=np.complex128)
        ctx = self.context()
        rc = _lib.load().phf_localg(ctx.handle, _lib.ptr(gradmat), _lib.ptr(rdmmat), _lib.ptr(smat),
                                    _lib.ptr(e0a), _lib.ptr(fsp), ncisp, _lib.ptr(fpg), ncipg, _lib.ptr(fk),
                                    fk.size, int(cmplxconj), nso, nci, _lib.ptr(g))
        _lib.check(rc, ctx.handle)
        return g

    # ------------------------------------------------------------------
    def buildsc(self, c0, c1, c2, nocc, ngrid, qj, *args, nso=None, npg=None, ncisp=None, ncipg=None, ncik=None):
        """BuildSC (PHFTools.f90:464-537) on the GPU: (t0, t1, t2).

        f2py order of the Fortran signature: c0, c1, c2, nocc, ngrid, qj, r1, r2,
        rpg, rk, roota, rootb, rooty, weightsp, weightpg, fsp, fpg, fk,
        [nso, npg, ncisp, ncipg, ncik].  The call at LinDep.py:73 passes one more
        integer (nci) after qj; an integer in that place is accepted and, like
        rk, fk and the roots (which the Fortran never reads), unused.
        """
        args = list(args)
        if args and np.ndim(args[0]) == 0:
            args.pop(0)
        names = ["r1", "r2", "rpg", "rk", "roota", "rootb", "rooty", "weightsp", "weightpg", "fsp", "fpg", "fk"]
        if len(args) < len(names):
            raise TypeError(f"buildsc: missing required argument {names[len(args)]!r}")
        opt = args[len(names):]
        if len(opt) > 5:
            raise TypeError("buildsc: too many arguments")
        kw = dict(zip(["nso", "npg", "ncisp", "ncipg", "ncik"], opt))
        nso = kw.get("nso", nso)
        npg = kw.get("npg", npg)
        ncisp = kw.get("ncisp", ncisp)
        ncipg = kw.get("ncipg", ncipg)
        r1, r2, rpg, _rk, _ra, _rb, _ry, weightsp, weightpg, fsp, fpg, _fk = args[:len(names)]
        c1 = np.ascontiguousarray(c1, dtype=np.complex128)
        c2 = np.ascontiguousarray(c2, dtype=np.complex128)
        nso = c1.shape[0] if nso is None else int(nso)
        nocc = int(nocc)
        r1 = np.ascontiguousarray(r1, dtype=np.complex128)
        r2 = np.ascontiguousarray(r2, dtype=np.complex128)
        rpg = np.ascontiguousarray(rpg, dtype=np.complex128)
        weightsp = np.ascontiguousarray(weightsp, dtype=np.float64)
        weightpg = np.ascontiguousarray(weightpg, dtype=np.float64)
        fsp = np.ascontiguousarray(fsp, dtype=np.complex128).reshape(-1)
        fpg = np.ascontiguousarray(fpg, dtype=np.complex128).reshape(-1)
        npg = rpg.shape[0] if npg is None else int(npg)
        ncisp = fsp.size if ncisp is None else int(ncisp)
        ncipg = fpg.size if ncipg is None else int(ncipg)
        ngrid = np.ascontiguousarray(ngrid, dtype=np.int32).reshape(-1)
        nl, ny = int(ngrid[0]), int(ngrid[1])
        checks = [
            (ngrid.size == 2, "ngrid must have 2 entries"),
            (c1.shape == (nso, nso), "c1 must be (nso, nso)"),
            (c2.shape == (nso,) * 4, "c2 must be (nso, nso, nso, nso)"),
            (r1.ndim == 3 and r1.shape[1:] == (nso, nso) and r1.shape[0] >= nl, "r1 must be (>=ngrid[0], nso, nso)"),
            (r2.ndim == 3 and r2.shape[1:] == (nso, nso) and r2.shape[0] >= ny, "r2 must be (>=ngrid[1], nso, nso)"),
            (rpg.shape == (npg, nso, nso), "rpg must be (npg, nso, nso)"),
            (weightpg.shape == (npg, ncipg, ncipg), "weightpg must be (npg, ncipg, ncipg)"),
            (weightsp.ndim == 2 and weightsp.shape[0] >= nl and weightsp.shape[1] >= ny,
             "weightsp must cover (ngrid[0], ngrid[1])"),
            (fsp.size >= ncisp and ncisp >= 2 * int(qj) + 1, "fsp must hold 2 qj + 1 entries"),
            (fpg.size >= ncipg, "fpg shorter than ncipg"),
        ]
        for ok, msg in checks:
            if not ok:
                raise ValueError(f"buildsc: {msg}")
        ctx = self.context()
        key = (nso, nocc, tuple(int(v) for v in ngrid), int(qj), _fingerprint(r1), _fingerprint(r2),
               _fingerprint(rpg), _fingerprint(weightsp), _fingerprint(weightpg), _fingerprint(fsp),
               _fingerprint(fpg))
        if getattr(ctx, "sc_key", None) != key:
            ctx.sc_key = None
            ctx.sc_set_grid(nso, nocc, ngrid, int(qj), r1, r2, rpg, weightsp[:nl, :ny], weightpg, fsp[:ncisp],
                            fpg[:ncipg])
            ctx.sc_key = key
        t0, t1, t2 = ctx.sc_apply(np.array([complex(np.asarray(c0).reshape(-1)[0])]), c1[None], c2[None])
        return t0[0], np.asfortranarray(t1[0]), np.asfortranarray(t2[0])


class _BuildSCDr:
    """The f2py module's ``buildsc_dr`` (BuildSC_dr.f90), imported at LinDep.py:6."""

    def __init__(self, tools: _PhfTools):
        self._tools = tools

    def buildsc_vec(self, c0, c1, c2, y, nso=None):
        """BuildSC_Vec (BuildSC_dr.f90:8-1244) on the GPU: (t0, t1, t2) for any Y."""
        c1 = np.ascontiguousarray(c1, dtype=np.complex128)
        c2 = np.ascontiguousarray(c2, dtype=np.complex128)
        y = np.ascontiguousarray(y, dtype=np.complex128)
        nso = c1.shape[0] if nso is None else int(nso)
        if c1.shape != (nso, nso) or c2.shape != (nso,) * 4 or y.shape != (nso, nso):
            raise ValueError("buildsc_vec: c1 and y must be (nso, nso), c2 (nso, nso, nso, nso)")
        ctx = self._tools.context()
        ctx.sc_key = None          # the grid of phftools.buildsc is replaced
        ctx.sc_set_points(y)
        t0, t1, t2 = ctx.sc_apply(np.array([complex(np.asarray(c0).reshape(-1)[0])]), c1[None], c2[None])
        return t0[0], np.asfortranarray(t1[0]), np.asfortranarray(t2[0])


phftools = _PhfTools()
buildsc_dr = _BuildSCDr(phftools)
