"""Projected-CC grid consumer on the GPU.

``pbar_hbar_olap0`` returns (Olap0, EOlap0) of PbarHbarOlap
(SUCCSD_Boson_cmplx/FPUCC.f90:14-272) with the argument meaning of the
Fortran routine (the Hamiltonian, the CC amplitudes T1, T2 and the rotations
in the MO basis of the reference, as Main_PCC.py:30-42 prepares them):

  Olap0  = sum_g w_g <Phi| R_g e^{T1+T2} |Phi>          (lambda0 of EvalKernel)
  EOlap0 = sum_g w_g <Phi| H R_g e^{T1+T2} |Phi>        (eps0, with ENuc)

with the rotated amplitudes of DSTT2 (FPUCC_Tools.f90:697-722) and the CC
energy of CCEnergy (CCRes.f90:9-29).  The per-point work runs in
``phf_pcc_olap0`` (csrc/pcc.hip).  ``pbar_hbar_olap`` returns every output,
(Olap0, OlapEx1, OlapEx2, EOlap0, EOlapEx1, EOlapEx2), adding EvalKernel's
singles and doubles (``phf_pcc_olap``): per point the CCSD residual CCRes12
(CCRes.f90:164-536) of the rotated amplitudes (U1t, U2t) and BuildKernSD's
fills (FPUCC_Tools.f90:291-436),

  OlapEx1  = sum_g w_g lambda0 U1t,  OlapEx2 = sum_g w_g lambda0 (U2t + U1t^U1t)
  EOlapEx1 = sum_g w_g lambda0 (R1 + U1t EOvl0)
  EOlapEx2 = sum_g w_g lambda0 (R2 + (U2t + U1t^U1t) EOvl0 + P(ij)P(ab) U1t R1)

restricted by ProjSz (FPUCC_Tools.f90:150-233) when sp = 2.
"""
from __future__ import annotations

import numpy as np

from . import _lib


def pbar_hbar_olap0(hone, htwo, t1, t2, nocc, enuc, qj, sp, npoints, r1, r2, rpg, rk, roota, rootb, rooty,
                    weightsp, weightpg, fsp, fpg, fk, ctx: _lib.Context | None = None):
    """Olap0, EOlap0 (complex) of PbarHbarOlap.  t1: (nso-nocc, nocc) =
    T1(NOcc+1:NSO, 1:NOcc); t2: (v, v, o, o); r1 (nl, nso, nso), r2 (ny, nso,
    nso), rpg (npg, nso, nso), rk (ncik, nso, nso) or None; weightsp (nl, ny);
    weightpg (npg, ncipg, ncipg); fsp (2 qj + 1); fpg (ncipg); fk (ncik)."""
    return _run(False, hone, htwo, t1, t2, nocc, enuc, qj, sp, npoints, r1, r2, rpg, rk, roota, rootb, rooty,
                weightsp, weightpg, fsp, fpg, fk, ctx)


def pbar_hbar_olap(hone, htwo, t1, t2, nocc, enuc, qj, sp, npoints, r1, r2, rpg, rk, roota, rootb, rooty,
                   weightsp, weightpg, fsp, fpg, fk, ctx: _lib.Context | None = None):
    """(Olap0, OlapEx1, OlapEx2, EOlap0, EOlapEx1, EOlapEx2) of PbarHbarOlap,
    arguments as pbar_hbar_olap0; OlapEx1 / EOlapEx1 (v, o), OlapEx2 /
    EOlapEx2 (v, v, o, o)."""
    return _run(True, hone, htwo, t1, t2, nocc, enuc, qj, sp, npoints, r1, r2, rpg, rk, roota, rootb, rooty,
                weightsp, weightpg, fsp, fpg, fk, ctx)


def _run(sd, hone, htwo, t1, t2, nocc, enuc, qj, sp, npoints, r1, r2, rpg, rk, roota, rootb, rooty,
         weightsp, weightpg, fsp, fpg, fk, ctx):
    c128 = lambda a: np.ascontiguousarray(a, dtype=np.complex128)
    f64 = lambda a: np.ascontiguousarray(np.reshape(a, -1) if np.ndim(a) <= 1 else a, dtype=np.float64)
    hone, htwo, t1, t2 = c128(hone), c128(htwo), c128(t1), c128(t2)
    nso = hone.shape[0]
    nocc = int(nocc)
    nv = nso - nocc
    npoints = np.ascontiguousarray(np.reshape(npoints, -1), dtype=np.int32)
    r1, r2, rpg = c128(r1), c128(r2), c128(rpg)
    fsp, fpg, fk = c128(np.reshape(fsp, -1)), c128(np.reshape(fpg, -1)), c128(np.reshape(fk, -1))
    weightsp, weightpg = f64(weightsp), f64(weightpg)
    roota, rootb, rooty = f64(roota), f64(rootb), f64(rooty)
    ncik = fk.size
    rk = c128(rk) if rk is not None else None
    checks = [
        (htwo.shape == (nso,) * 4, "htwo must be (nso,)*4"),
        (t1.shape == (nv, nocc), "t1 must be (nso-nocc, nocc)"),
        (t2.shape == (nv, nv, nocc, nocc), "t2 must be (v, v, o, o)"),
        (npoints.size == 2, "npoints must have 2 entries"),
        (r1.ndim == 3 and r1.shape[1:] == (nso, nso), "r1 must be (*, nso, nso)"),
        (r2.ndim == 3 and r2.shape[1:] == (nso, nso), "r2 must be (*, nso, nso)"),
        (rpg.ndim == 3 and rpg.shape[1:] == (nso, nso), "rpg must be (npg, nso, nso)"),
        (weightsp.shape == (int(npoints[0]), int(npoints[1])), "weightsp must be (npoints[0], npoints[1])"),
        (weightpg.ndim == 3 and weightpg.shape[0] == rpg.shape[0] and weightpg.shape[1] == weightpg.shape[2] ==
         fpg.size, "weightpg must be (npg, ncipg, ncipg)"),
        (fsp.size == 2 * int(qj) + 1, "fsp must have 2 qj + 1 entries"),
        (ncik in (1, 2), "fk must have 1 or 2 entries"),
        (ncik == 1 or (rk is not None and rk.shape == (2, nso, nso)), "ncik = 2 needs rk (2, nso, nso)"),
    ]
    for ok, msg in checks:
        if not ok:
            raise ValueError(msg)
    own = ctx is None
    ctx = ctx or _lib.Context(0)
    try:
        o = np.zeros(1, np.complex128)
        e = np.zeros(1, np.complex128)
        args = [ctx.handle, _lib.ptr(hone), _lib.ptr(htwo), _lib.ptr(t1), _lib.ptr(t2), nocc, nso, float(enuc),
                int(qj), int(sp), npoints.ctypes.data_as(_lib._ip), rpg.shape[0], fsp.size, fpg.size, ncik,
                _lib.ptr(r1), r1.shape[0], _lib.ptr(r2), r2.shape[0], _lib.ptr(rpg), _lib.ptr(rk), _lib.ptr(roota),
                roota.size, _lib.ptr(rootb), rootb.size, _lib.ptr(rooty), rooty.size, _lib.ptr(weightsp),
                _lib.ptr(weightpg), _lib.ptr(fsp), _lib.ptr(fpg), _lib.ptr(fk)]
        if not sd:
            rc = _lib.load().phf_pcc_olap0(*args, _lib.ptr(o), _lib.ptr(e))
            _lib.check(rc, ctx.handle)
            return complex(o[0]), complex(e[0])
        o1, e1 = np.zeros((nv, nocc), np.complex128), np.zeros((nv, nocc), np.complex128)
        o2, e2 = np.zeros((nv, nv, nocc, nocc), np.complex128), np.zeros((nv, nv, nocc, nocc), np.complex128)
        rc = _lib.load().phf_pcc_olap(*args, _lib.ptr(o), _lib.ptr(o1), _lib.ptr(o2), _lib.ptr(e), _lib.ptr(e1),
                                      _lib.ptr(e2))
        _lib.check(rc, ctx.handle)
        return complex(o[0]), o1, o2, complex(e[0]), e1, e2
    finally:
        if own:
            ctx.close()
