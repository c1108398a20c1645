"""ORACLE -- test infrastructure only: ctypes wrapper of oracle/lib/libphf_cref.so
(the C + OpenMP restatement of BuildHSG in oracle/buildhsg_cref.c)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libphf_cref.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} missing; run make -C oracle")
        _lib = C.CDLL(LIB)
        _lib.cref_buildhsg.restype = C.c_int
    return _lib


def buildhsg(hone, htwo, ref, z, ngrid, nci, cmplxconj, ncisp, nocc, r1, r2, rpg, weightsp, weightpg, roota,
             rootb, rooty, qj, sp, points=None):
    c = lambda a: np.ascontiguousarray(a, dtype=np.complex128)
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    hone, htwo, ref, z, r1, r2, rpg = map(c, (hone, htwo, ref, z, r1, r2, rpg))
    weightsp, weightpg, roota, rootb, rooty = map(f, (weightsp, weightpg, roota, rootb, rooty))
    nso = hone.shape[0]
    nl, ny = int(ngrid[0]), int(ngrid[1])
    npg, ncipg = rpg.shape[0], weightpg.shape[1]
    pts = np.ascontiguousarray(np.asarray(points if points is not None else np.zeros((0, 3)), dtype=np.int32))
    out = [np.zeros((nci, nci), complex), np.zeros((nci, nci), complex), np.zeros((nci, nci, nso, nso), complex),
           np.zeros((nci, nci, nso, nso), complex)]
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = load().cref_buildhsg(nso, p(hone), p(htwo), p(ref), p(z), nl, ny, npg, nci, int(cmplxconj), int(ncisp),
                              int(nocc), p(r1), p(r2), p(rpg), p(weightsp), p(weightpg), ncipg, p(roota), p(rootb),
                              p(rooty), int(qj), int(sp), p(pts), len(pts), *[p(o) for o in out])
    if rc:
        raise RuntimeError(f"singular overlap at work item {rc - 1}")
    return tuple(out)
