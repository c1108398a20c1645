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


# ----------------------------------------------------------------------------
# CmplxConj = 2 (PHFTools.f90:95-119,153-179 blocks; LocalG :227-241)
# ----------------------------------------------------------------------------

def _cc2(nsite, j, ngrid, seed):
    case = hubbard_general_ref(nsite, 4.0, seed, zscale=0.3, complex_ref=False)
    ga = grid_args(nsite, ngrid, j)
    ga["nci"] *= 2
    return case, ga


@pytest.mark.parametrize("nsite,ngrid,seed", [(4, (14, 8), 3), (6, (14, 8), 5)])
def test_cmplxconj2_j0_matches_fock_space(nsite, ngrid, seed):
    """J = 0, real Ref, complex Z, exact grid: every element of the 2 x 2
    Hmat / Smat equals sum(weightsp) * <Phi_a| H P |Phi_b> / <Phi_a| P |Phi_b>
    over {Phi, Phi*} (Phi* = the determinant of conj(C0), brute-force Fock
    space); E0 is the lowest root on span{P Phi, P Phi*}; the LocalG gradient
    (cross terms included) equals central differences of E (factor 2)."""
    from oracle.fock_proj import projected_energy_multi, projected_matrices
    case, ga = _cc2(nsite, 0, ngrid, seed)
    nocc = case["nocc"]
    h, s, gm, rm = O.buildhsg(case["hone"], case["htwo"], case["ref"], case["z"], cmplxconj=2, nocc=nocc, **ga)
    c0 = np.vstack([np.eye(nocc), case["z"]])
    cl = [case["ref"] @ c0, case["ref"] @ c0.conj()]
    hm, sm = projected_matrices(case["hso"], case["vao"], cl, 0)
    c = ga["weightsp"].sum()
    assert abs(c - 2 * np.pi) < 1e-12
    assert np.abs(h - c * hm).max() < 1e-11 * np.abs(h).max()
    assert np.abs(s - c * sm).max() < 1e-11 * np.abs(s).max()
    ev, _ = O.eigen_solver(h, s)
    assert abs(ev[0] - projected_energy_multi(case["hso"], case["vao"], cl, 0)[0]) < 1e-11
    x0 = np.concatenate([case["z"].real.ravel(), case["z"].imag.ravel()])
    f = lambda x: O.eandg(x, case["hone"], case["htwo"], case["ref"], nocc, ga, cmplxconj=2)
    e, g = f(x0)
    rng = np.random.default_rng(seed)
    for _ in range(4):
        d = rng.normal(size=x0.shape)
        d /= np.linalg.norm(d)
        fd = (f(x0 + 1e-5 * d)[0] - f(x0 - 1e-5 * d)[0]) / 2e-5
        assert abs(fd - 2 * g @ d) < 1e-8 * (1 + abs(fd))


def _wirtinger(fun, z0, dz, h=1e-5):
    """sum dF/dZ* conj(dz) of an array-valued F(Z) by central differences:
    (F'(dz) + i F'(i dz)) / 2."""
    d1 = (fun(z0 + h * dz) - fun(z0 - h * dz)) / (2 * h)
    d2 = (fun(z0 + 1j * h * dz) - fun(z0 - 1j * h * dz)) / (2 * h)
    return 0.5 * (d1 + 1j * d2)


def test_cmplxconj2_gradmat_blocks_are_the_derivatives_of_hmat():
    """J = 1, CmplxConj = 2: Gradmat / Rdmmat hold the Z* derivatives of the
    Hmat / Smat blocks.  Measured block by block with d = 2J+1:
      d H(m,n)/dZ*       = Gradmat(m,n)                      (bra)
      d H(m+d,n+d)/dZ*   = conj Gradmat(m+d,n+d)             (ket of Phi*)
      d H(m,n+d)/dZ*     = Gradmat(m,n+d) + conj Gradmat(n+d,m)
      d H(m+d,n)/dZ*     = 0
    and the same with Smat / Rdmmat.  The third line carries the CI indices
    (n+d, m); LocalG (PHFTools.f90:236-237) uses conj Gradmat(m+d, n)."""
    case, ga = _cc2(4, 1, (26, 8), 3)
    nocc, d = case["nocc"], ga["nci"] // 2
    run = lambda z: O.buildhsg(case["hone"], case["htwo"], case["ref"], z, cmplxconj=2, nocc=nocc, **ga)
    h, s, gm, rm = run(case["z"])
    rng = np.random.default_rng(9)
    dz = rng.normal(size=case["z"].shape) + 1j * rng.normal(size=case["z"].shape)
    proj = lambda m: np.sum(m[..., nocc:, :nocc] * np.conj(dz), axis=(-2, -1))
    for idx, dmat in ((0, gm), (1, rm)):
        b = _wirtinger(lambda z: run(z)[idx], case["z"], dz)
        scale = np.abs(b).max()
        g = proj(dmat)
        gc = proj(np.conj(dmat))
        assert np.abs(b[:d, :d] - g[:d, :d]).max() < 1e-7 * scale
        assert np.abs(b[d:, d:] - gc[d:, d:]).max() < 1e-7 * scale
        assert np.abs(b[:d, d:] - g[:d, d:] - gc[d:, :d].T).max() < 1e-7 * scale
        assert np.abs(b[d:, :d]).max() < 1e-7 * scale
        # LocalG's orientation differs wherever the exchange-block derivative is not symmetric
        assert np.abs(b[:d, d:] - g[:d, d:] - gc[d:, :d]).max() > 1e-3 * scale


def test_cmplxconj2_j1_reference_departures_from_the_projection():
    """J = 1, CmplxConj = 2, exact grid: what the restatement keeps from the
    reference, with the exact quantities beside it.
    (1) Hmat/Smat block (1,1) is a copy of (0,0) (PHFTools.f90:177-178); the
        projection on span{P_MK Phi, P_MK Phi*} has conj((0,0)) there (plain
        complex conjugation commutes with P for a real Hamiltonian): with that
        block, E0 equals the brute-force root; the copy gives a lower E0.
    (2) The gradient of the reference's own E (its Hmat, its EigenSolver) is
        f^H (dH - E dS) f / f^H S f with the block derivatives of the test
        above and the whole CI vector f: it matches central differences,
        while LocalG's (transposed exchange-block term, and f rebuilt as
        fsp x fk by VecDecomp, PHFTools.py:20-26) does not."""
    from oracle.fock_proj import projected_energy_multi
    case, ga = _cc2(4, 1, (26, 8), 3)
    nocc, d = case["nocc"], ga["nci"] // 2
    run = lambda z: O.buildhsg(case["hone"], case["htwo"], case["ref"], z, cmplxconj=2, nocc=nocc, **ga)
    h, s, gm, rm = run(case["z"])
    assert np.array_equal(h[d:, d:], h[:d, :d]) and np.array_equal(s[d:, d:], s[:d, :d])
    c0 = np.vstack([np.eye(nocc), case["z"]])
    e_exact = projected_energy_multi(case["hso"], case["vao"], [case["ref"] @ c0, case["ref"] @ c0.conj()], 1)[0]
    h2, s2 = h.copy(), s.copy()
    h2[d:, d:], s2[d:, d:] = h[:d, :d].conj(), s[:d, :d].conj()
    assert abs(O.eigen_solver(h2, s2)[0][0] - e_exact) < 1e-10
    e_ref = O.eigen_solver(h, s)[0][0]
    assert e_ref < e_exact - 1e-2

    def energy_and_exact_gradient(z):
        hh, ss, g_, r_ = run(z)
        ev, vec = O.eigen_solver(hh, ss)
        v = vec[:, 0]
        dg, dr = np.zeros_like(g_), np.zeros_like(r_)
        dg[:d, :d], dr[:d, :d] = g_[:d, :d], r_[:d, :d]
        dg[d:, d:], dr[d:, d:] = np.conj(g_[d:, d:]), np.conj(r_[d:, d:])
        dg[:d, d:] = g_[:d, d:] + np.conj(g_[d:, :d].transpose(1, 0, 2, 3))
        dr[:d, d:] = r_[:d, d:] + np.conj(r_[d:, :d].transpose(1, 0, 2, 3))
        num = np.einsum('m,n,mnpq->pq', v.conj(), v, dg - ev[0] * dr)
        grad = (num / (v.conj() @ ss @ v))[nocc:, :nocc]
        return ev[0].real, grad, (g_, r_, ss, ev, vec)

    e, g_exact, (g_, r_, ss, ev, vec) = energy_and_exact_gradient(case["z"])
    fsp, fpg, fk = O.vec_decomp(vec[:, 0], ga["ncisp"], 1, 2)
    g_localg = O.localg(g_, r_, ss, ev[0], fsp, fpg, fk, 2)[nocc:, :nocc]
    x0 = np.concatenate([case["z"].real.ravel(), case["z"].imag.ravel()])
    nv = case["z"].size
    f = lambda x: energy_and_exact_gradient((x[:nv] + 1j * x[nv:]).reshape(case["z"].shape))[0]
    rng = np.random.default_rng(4)
    worst_exact = worst_localg = 0.0
    for _ in range(4):
        dd = rng.normal(size=x0.shape)
        dd /= np.linalg.norm(dd)
        fd = (f(x0 + 1e-5 * dd) - f(x0 - 1e-5 * dd)) / 2e-5
        dz = (dd[:nv] + 1j * dd[nv:]).reshape(case["z"].shape)
        worst_exact = max(worst_exact, abs(fd - 2 * np.real(np.sum(np.conj(g_exact) * dz))))
        worst_localg = max(worst_localg, abs(fd - 2 * np.real(np.sum(np.conj(g_localg) * dz))))
    assert worst_exact < 1e-7
    assert worst_localg > 1e-2
