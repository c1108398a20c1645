"""Pinning the CPU oracle (oracle/phf_oracle.py) without the reference binary.

Each test fixes what a plausible mistake would break: the grid/rotation/Wigner
conventions (exact Fock-space projection), the gradient (finite differences,
factor 2 on [Re Z, Im Z]), the LU determinant and inverse (numpy), and the
mean-field and closed-shell limits.
"""
import numpy as np
import pytest

from oracle import phf_oracle as O
from oracle.fock_proj import projected_energy
from phf_pcc_amd.hamiltonian import hubbard, make_ene, spin_orbital_blocks, mulliken_to_dirac, sort_orb, ao2mo
from tests._cases import grid_args, hubbard_general_ref


def _oracle_energy(case, ga, cmplxconj=1):
    h, s, gm, rm = O.buildhsg(case["hone"], case["htwo"], case["ref"], case["z"], cmplxconj=cmplxconj,
                              nocc=case["nocc"], **ga)
    ev, vec = O.eigen_solver(h, s)
    return ev[0].real, (h, s, gm, rm, ev, vec)


@pytest.mark.parametrize("nsite,u,ngrid,j,seed", [
    (4, 4.0, (14, 8), 0, 3),
    (4, 4.0, (6, 4), 0, 4),
    (6, 4.0, (14, 8), 0, 5),
    (4, 4.0, (26, 8), 1, 6),
    (4, 2.0, (50, 16), 1, 7),
    (6, 4.0, (38, 8), 2, 3),
])
def test_projected_energy_matches_fock_space(nsite, u, ngrid, j, seed):
    """Grid projection == exact Loewdin projection once the quadrature is exact
    (Lebedev degree >= N/2 and enough gamma points)."""
    case = hubbard_general_ref(nsite, u, seed, zscale=0.3, complex_ref=(j < 2))
    e_grid, _ = _oracle_energy(case, grid_args(nsite, ngrid, j))
    c_ao = case["ref"] @ np.vstack([np.eye(nsite), case["z"]])
    e_exact, _ = projected_energy(case["hso"], case["vao"], c_ao, j)
    assert abs(e_grid - e_exact) < 1e-11


def test_coarse_grid_is_not_exact():
    """Guard that the Fock-space check can fail: a too-small grid differs."""
    case = hubbard_general_ref(6, 4.0, 11, zscale=0.4)
    e_grid, _ = _oracle_energy(case, grid_args(6, (1, 8), 0))
    c_ao = case["ref"] @ np.vstack([np.eye(6), case["z"]])
    e_exact, _ = projected_energy(case["hso"], case["vao"], c_ao, 0)
    assert abs(e_grid - e_exact) > 1e-6


@pytest.mark.parametrize("j,ngrid", [(0, (14, 8)), (1, (26, 8))])
def test_gradient_matches_central_differences(j, ngrid):
    case = hubbard_general_ref(4, 4.0, 21, zscale=0.2)
    ga = grid_args(4, ngrid, j)
    x = np.concatenate([case["z"].real.ravel(), case["z"].imag.ravel()])
    e0, g = O.eandg(x, case["hone"], case["htwo"], case["ref"], case["nocc"], ga)
    rng = np.random.default_rng(1)
    h = 1e-5
    for _ in range(4):
        d = rng.normal(size=x.shape)
        ep = O.eandg(x + h * d, case["hone"], case["htwo"], case["ref"], case["nocc"], ga)[0]
        em = O.eandg(x - h * d, case["hone"], case["htwo"], case["ref"], case["nocc"], ga)[0]
        fd = (ep - em) / (2 * h)
        assert abs(fd - 2 * g @ d) < 1e-6 * max(1.0, abs(fd))


def test_identity_grid_gives_mean_field_energy():
    """One grid point with R = I: E = the mean-field energy (makeH.py:79-84)."""
    case = hubbard_general_ref(4, 4.0, 31, zscale=0.3, complex_ref=True)
    nso = 8
    ga = grid_args(4, (1, 1), 0, sp=0)
    e, _ = _oracle_energy(case, ga)
    c = np.vstack([np.eye(4), case["z"]])
    q, _ = np.linalg.qr(c)
    e_mf = make_ene(case["hone"], case["htwo"], q).real
    assert abs(e - e_mf) < 1e-12


def test_closed_shell_rhf_is_a_singlet():
    """U=0 ring, RHF orbitals, Z=0: every grid point has det=1 and the projected
    energy equals the closed-form RHF energy 2 sum_occ (-2 cos k)."""
    n = 6
    t0, u0 = hubbard(1, n, 0.0)
    e, c = np.linalg.eigh(t0)
    ref = sort_orb(c, c, n // 2, n // 2)
    hso, h2m = spin_orbital_blocks(t0, u0)
    h1, h2 = ao2mo(hso, ref, 2), ao2mo(mulliken_to_dirac(h2m), ref, 4)
    case = dict(hone=h1, htwo=h2, ref=ref.astype(complex), z=np.zeros((n, n), complex), nocc=n)
    ga = grid_args(n, (14, 8), 0)
    energy, (h, s, *_rest) = _oracle_energy(case, ga)
    ks = 2 * np.pi * np.arange(n) / n
    eps = np.sort(-2 * np.cos(ks))
    assert abs(energy - 2 * eps[: n // 2].sum()) < 1e-12
    assert abs(s[0, 0] - 2 * np.pi) < 1e-12       # sum of weights, det = 1


def test_lu_matches_numpy():
    rng = np.random.default_rng(5)
    for n in (1, 2, 5, 17):
        m = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        assert abs(O.determinant_c(m) - np.linalg.det(m)) < 1e-10 * abs(np.linalg.det(m)) + 1e-14
        assert np.allclose(O.invert_c(m), np.linalg.inv(m), atol=1e-10)
    # a permutation matrix: every pivot is a swap
    p = np.eye(6)[[3, 0, 5, 1, 2, 4]].astype(complex)
    assert abs(O.determinant_c(p) - np.linalg.det(p)) < 1e-14
    sing = np.ones((3, 3), complex)
    assert O.determinant_c(sing) == 0
    with pytest.raises(O.OracleError):
        O.invert_c(sing)


def test_wigner_closed_forms():
    b = 0.7
    assert abs(O.wignerfac(1, 0, 0, 0.3, b, 0.2) - np.cos(b)) < 1e-14
    assert abs(O.wignerfac(1, 1, 1, 0.0, b, 0.0) - (1 + np.cos(b)) / 2) < 1e-14
    assert abs(O.wignerfac(1, 1, -1, 0.0, b, 0.0) - (1 - np.cos(b)) / 2) < 1e-14
    assert abs(O.wignerfac(1, 1, 0, 0.0, b, 0.0) - (-np.sin(b) / np.sqrt(2))) < 1e-14
    assert abs(O.wignerfac(1, 1, 0, 0.5, b, 0.0) - np.exp(-0.5j) * (-np.sin(b) / np.sqrt(2))) < 1e-14
    assert abs(O.wignerfac(0, 0, 0, 1.0, 2.0, 3.0) - 1.0) < 1e-15


def test_not_hermitian_raises():
    case = hubbard_general_ref(4, 4.0, 41, zscale=0.3)
    ga = grid_args(4, (1, 1), 0, sp=1)
    from phf_pcc_amd.spin_proj import build_rot_mat
    ga["r2"] = build_rot_mat(0.0, 0.0, 0.7, 4)[None]
    with pytest.raises(O.OracleError):
        _oracle_energy(case, ga)


def test_c_restatement_matches_numpy_oracle():
    """oracle/buildhsg_cref.c (the CPU baseline) reproduces the numpy oracle."""
    from oracle import cref
    for j, ng in [(0, (14, 8)), (1, (26, 8))]:
        case = hubbard_general_ref(4, 4.0, 61 + j, zscale=0.3, complex_ref=True)
        ga = grid_args(4, ng, j)
        a = O.buildhsg(case["hone"], case["htwo"], case["ref"], case["z"], cmplxconj=1, nocc=4, **ga)
        b = cref.buildhsg(case["hone"], case["htwo"], case["ref"], case["z"], cmplxconj=1, nocc=4, **ga)
        for x, y in zip(a, b):
            assert np.abs(x - y).max() <= 1e-12 * np.abs(x).max()


def test_factored_gamma_equals_dense_make_gamma():
    """oracle.factored (MakeGamma from B factors, for the n = 96 / 128 sizes)
    against the dense MakeGamma of the restatement on the Dirac integrals
    built the reference's way (Main_PHF.py:22-33, makeH.py:63-67,
    ao2mo.py:4-10), for random non-Hermitian rho and three Ref: a sorted
    permutation, a random real orthogonal and a random complex unitary one."""
    from oracle.factored import gamma_factored
    from phf_pcc_amd.hamiltonian import (ao2mo, eri_from_factors, mulliken_to_dirac, random_unitary, sorted_ref,
                                         spin_orbital_blocks, synthetic_factors)
    n = 5
    rng = np.random.default_rng(77)
    h, b = synthetic_factors(n, 7, rng)
    eri = eri_from_factors(b)
    _, htwo_m = spin_orbital_blocks(h, eri)
    vao = mulliken_to_dirac(htwo_m)
    for ref in (sorted_ref(h, 4).astype(complex), random_unitary(2 * n, rng, real=True).astype(complex),
                random_unitary(2 * n, rng)):
        vmo = ao2mo(vao, ref, 4)
        g = gamma_factored(b, ref)
        for _ in range(3):
            rho = rng.normal(size=(2 * n, 2 * n)) + 1j * rng.normal(size=(2 * n, 2 * n))
            dense = O.make_gamma(vmo, rho)
            assert np.abs(g(rho) - dense).max() < 1e-12 * max(1.0, np.abs(dense).max())
