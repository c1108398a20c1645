"""ORACLE -- test infrastructure only.

MakeGamma (PHF_fort/PHFTools.f90:310-337) for a Hamiltonian given by its
Cholesky-like factors, so that the oracle reaches the sizes of the n = 96 and
n = 128 workloads, whose dense <ij||kl> (21.7 GB and 68.7 GB as complex128)
does not fit a host.  The integrals are those of the dense construction:

  spatial (pq|rs) = sum_L B^L_pq B^L_rs                 (B^L symmetric)
  spin-orbital Mulliken blocks as Main_PHF.py:22-29, then
  <ab||cd> = (ac|bd) - (ad|bc)                           (makeH.py:63-67)
  V_MO = ao2mo(V_AO, Ref, 4)                             (ao2mo.py:4-10)

and Gamma_MO(i,k) = sum_{jl} V_MO(i,j,k,l) rho(l,j) is evaluated without V:
with rho_AO = Ref rho Ref^H split into spin blocks rho^{ss'} (n x n),

  Gamma_AO^{ss'} = delta_{ss'} J - K^{ss'},
  J = sum_L B^L tr(B^L (rho^{aa} + rho^{bb})),   K^{ss'} = sum_L B^L rho^{ss'} B^L,

and Gamma_MO = Ref^H Gamma_AO Ref.  tests/test_oracle.py pins it against the
dense ``phf_oracle.make_gamma`` on random (not only idempotent) rho, for
permutation and general complex Ref.
"""
from __future__ import annotations

import numpy as np


def gamma_factored(b: np.ndarray, ref: np.ndarray):
    """Return gamma(rho) -> Gamma_MO for factors b[L, n, n] and Ref[2n, 2n]."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    nl, n, _ = b.shape
    ref = np.asarray(ref, dtype=np.complex128)
    refh = ref.conj().T
    bflat = b.reshape(nl, n * n)

    def gamma(rho: np.ndarray) -> np.ndarray:
        rao = ref @ rho @ refh
        out = np.zeros((2 * n, 2 * n), dtype=np.complex128)
        for s in range(2):
            for t in range(2):
                blk = rao[s * n:(s + 1) * n, t * n:(t + 1) * n]
                # K^{st} = sum_L B^L blk B^L: one product over all L, then a batched one
                x = (b.reshape(nl * n, n) @ blk).reshape(nl, n, n)
                out[s * n:(s + 1) * n, t * n:(t + 1) * n] = -np.einsum('lpq,lqr->pr', x, b, optimize=True)
        tr = bflat @ (rao[:n, :n] + rao[n:, n:]).T.reshape(n * n)   # tr(B^L rho) = sum_pq B_pq rho_qp
        jmat = (tr @ bflat).reshape(n, n)
        out[:n, :n] += jmat
        out[n:, n:] += jmat
        return refh @ out @ ref

    return gamma
