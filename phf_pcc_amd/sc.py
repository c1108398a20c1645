"""Projected SD-space overlap times vector on the GPU, for linear-dependency
solvers over the singles-and-doubles space (LinDep.py).

``ProjectedOverlap`` keeps one projection grid on the device and applies
t = BuildSC(c) (PHFTools.f90:464-537) to blocks of packed SD vectors in one
call: the product that LinDep.py:71-76 forms one vector at a time, through
Vec2Mat / Mat2Vec (LinDep.py:156-187), for its Lanczos, Davidson and
conjugate-gradient drivers.  ``as_linear_operator()`` gives the
scipy.sparse.linalg.LinearOperator of LinDep.py:79-82 (matvec = rmatvec, as
there).
"""
from __future__ import annotations

import numpy as np

from . import _lib


def sd_dim(nso: int, nocc: int) -> int:
    """1 + O V + O(O-1)/2 V(V-1)/2 (LinDep.py:7)."""
    return int(_lib.load().phf_sd_dim(int(nso), int(nocc)))


class ProjectedOverlap:
    def __init__(self, nso: int, nocc: int, ngrid, r1, r2, rpg=None, weightsp=None, weightpg=None, qj: int = 0,
                 fsp=None, fpg=None, device: int = 0, ctx: _lib.Context | None = None):
        self.nso, self.nocc = int(nso), int(nocc)
        self.dim = sd_dim(nso, nocc)
        if rpg is None:
            rpg = np.eye(nso, dtype=complex)[None]
        if weightpg is None:
            weightpg = np.ones((rpg.shape[0], 1, 1))
        if fsp is None:
            fsp = np.ones(2 * qj + 1, complex)
        if fpg is None:
            fpg = np.ones(np.asarray(weightpg).shape[1], complex)
        if weightsp is None:
            raise ValueError("weightsp is required")
        self.ctx = ctx if ctx is not None else _lib.Context(device)
        self.ctx.sc_set_grid(self.nso, self.nocc, ngrid, qj, r1, r2, rpg, weightsp, weightpg, fsp, fpg)

    def matmat(self, vecs: np.ndarray) -> np.ndarray:
        """Columns of ``vecs`` [dim, k] -> their products [dim, k]."""
        v = np.asarray(vecs, dtype=np.complex128)
        if v.ndim == 1:
            return self.matvec(v)
        if v.shape[0] != self.dim:
            raise ValueError(f"expected {self.dim} rows")
        return self.ctx.sc_apply_packed(np.ascontiguousarray(v.T)).T

    def matvec(self, vec: np.ndarray) -> np.ndarray:
        v = np.asarray(vec, dtype=np.complex128).reshape(-1)
        if v.size != self.dim:
            raise ValueError(f"expected a vector of length {self.dim}")
        return self.ctx.sc_apply_packed(v)

    def apply_full(self, c0, c1, c2):
        """(t0, t1, t2) for full c1 (nso x nso) and c2 (nso^4) inputs, stacked
        along a leading axis (BuildSC's own arguments)."""
        return self.ctx.sc_apply(c0, c1, c2)

    def as_linear_operator(self):
        from scipy.sparse.linalg import LinearOperator
        return LinearOperator((self.dim, self.dim), matvec=self.matvec, rmatvec=self.matvec, matmat=self.matmat,
                              dtype=np.complex128)
