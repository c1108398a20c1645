"""Shared problem builders for the tests (inputs only; no checking logic)."""
from __future__ import annotations

import numpy as np

from phf_pcc_amd.hamiltonian import (ao2mo, hubbard, hubbard_ring_problem, mulliken_to_dirac,  # noqa: F401
                                     random_unitary, spin_orbital_blocks, synthetic_problem, sort_orb, uhf_hubbard)
from phf_pcc_amd.spin_proj import grid_args  # noqa: F401


def hubbard_general_ref(nsite: int, u: float, seed: int, zscale: float = 0.1, complex_ref: bool = False):
    """Hubbard ring with a UHF reference optionally rotated by a random unitary
    (complex MO integrals when complex_ref)."""
    t0, u0 = hubbard(1, nsite, u)
    ca, cb, _ = uhf_hubbard(t0, u, nsite // 2, nsite // 2)
    ref = sort_orb(ca, cb, nsite // 2, nsite // 2).astype(complex)
    rng = np.random.default_rng(seed)
    nso = 2 * nsite
    if complex_ref:
        ref = ref @ random_unitary(nso, rng)
    hso, h2m = spin_orbital_blocks(t0, u0)
    vao = mulliken_to_dirac(h2m)
    h1 = ao2mo(hso, ref, 2)
    h2 = ao2mo(vao, ref, 4)
    z = rng.normal(0, zscale, (nso - nsite, nsite)) + 1j * rng.normal(0, zscale, (nso - nsite, nsite))
    return dict(hone=h1, htwo=h2, ref=ref, z=z, nocc=nsite, hso=hso, vao=vao)


def hubbard_cfg1():
    """Hubbard ring n=8, U=4, half filling, UHF + Z ~ N(0, 0.05^2) (seed 1001), 14 x 8 grid."""
    p = hubbard_ring_problem(8, 4.0, seed=1001, zscale=0.05, ngrid=(14, 8))
    return dict(hone=p.hone, htwo=p.htwo, ref=p.ref.astype(complex), z=p.z, nocc=p.nocc), p.n


def synthetic_dense(n: int, nocc: int, seed: int, zscale: float = 0.05):
    p = synthetic_problem(n, nocc, seed=seed, build_dense=True, zscale=zscale)
    return dict(hone=p.hone, htwo=p.htwo, ref=p.ref.astype(complex), z=p.z, nocc=nocc)
