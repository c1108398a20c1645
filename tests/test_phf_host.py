"""Host side of the objective (phf_pcc_amd/phf.py) against the oracle's
restatement of PHFTools.py / PHF.py (no GPU)."""
import numpy as np

from oracle import phf_oracle as O
from phf_pcc_amd import phf
from tests._cases import grid_args, hubbard_general_ref


def test_eigen_solver_and_vec_decomp_match_oracle():
    rng = np.random.default_rng(3)
    a = rng.normal(size=(5, 5)) + 1j * rng.normal(size=(5, 5))
    h = a + a.conj().T
    b = rng.normal(size=(5, 5)) + 1j * rng.normal(size=(5, 5))
    s = b @ b.conj().T
    s[:, 0] *= 1e-4
    s[0, :] *= 1e-4
    e1, v1 = phf.eigen_solver(h, s)
    e2, v2 = O.eigen_solver(h, s)
    assert np.allclose(e1, e2)
    for k in range(v1.shape[1]):
        ph = np.vdot(v1[:, k], v2[:, k])
        assert np.allclose(v1[:, k] * ph / abs(ph), v2[:, k], atol=1e-10)
    vec = rng.normal(size=12) + 1j * rng.normal(size=12)
    for part1, part2 in zip(phf.vec_decomp(vec, 3, 2, 2), O.vec_decomp(vec, 3, 2, 2)):
        assert np.allclose(part1, part2)


def test_make_eandg_matches_oracle_objective():
    case = hubbard_general_ref(4, 4.0, 5, zscale=0.2)
    ga = grid_args(4, (14, 8), 1)
    cfg = phf.ObjectiveConfig(nso=8, nocc=4, cmplxconj=1, ncisp=ga["ncisp"])
    bh = lambda z: O.buildhsg(case["hone"], case["htwo"], case["ref"], z, cmplxconj=1, nocc=4, **ga)
    f = phf.make_eandg(cfg, bh, lambda gm, rm, s, e0, fsp, fpg, fk, cc: O.localg(gm, rm, s, e0, fsp, fpg, fk, cc))
    x = np.concatenate([case["z"].real.ravel(), case["z"].imag.ravel()])
    e1, g1 = f(x)
    e2, g2 = O.eandg(x, case["hone"], case["htwo"], case["ref"], 4, ga)
    assert abs(e1 - e2) < 1e-13 and np.abs(g1 - g2).max() < 1e-12


def test_z_masks():
    z = np.arange(12.0).reshape(4, 3) + 0j
    m = phf.sz_proj(z, 2, 1)
    assert m[0, 0] == z[0, 0] and m[0, 1] == 0 and m[2, 1] == z[2, 1] and m[2, 0] == 0
    r = phf.rhf_proj(z, 2, 1)
    assert r[2, 1] == z[0, 0] and r[3, 2] == z[1, 0]


def test_eigen_solver_indefinite_s_keeps_the_reference_columns():
    """PHFTools.py:10-14 scales the LAST dim eigenvectors of S by the inverse
    square roots of the eigenvalues that pass |lambda| > 1e-6, in ascending
    order.  With S = diag(-0.3, 1e-8, 2.0) the eigenvectors are the unit
    vectors, so X = [e1 / sqrt(0.3), e2 / sqrt(2.0)] and the projected H is
    the 2 x 2 matrix written out below; its lowest root is E0."""
    rng = np.random.default_rng(11)
    a = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    h = a + a.conj().T
    s = np.diag([-0.3, 1e-8, 2.0]).astype(complex)
    c1, c2 = 0.3 ** -0.5, 2.0 ** -0.5
    hh = np.array([[h[1, 1] * c1 * c1, h[1, 2] * c1 * c2], [h[2, 1] * c2 * c1, h[2, 2] * c2 * c2]])
    tr, det = (hh[0, 0] + hh[1, 1]).real, (hh[0, 0] * hh[1, 1] - hh[0, 1] * hh[1, 0]).real
    e_low = tr / 2 - np.sqrt(tr * tr / 4 - det)
    for solver in (phf.eigen_solver, O.eigen_solver):
        ev, vec = solver(h, s)
        assert ev.size == 2
        assert abs(ev[0] - e_low) < 1e-12
        # the CI vector lives in span(e1, e2): no weight on the negative eigenvector
        assert abs(vec[0, 0]) < 1e-14
    # a |lambda| > tol mask pairs e0 with 0.3**-0.5 instead and gives another E0
    keep = np.abs(np.diag(s).real) > 1e-6
    x = np.eye(3)[:, keep] / np.sqrt(np.abs(np.diag(s).real[keep]))
    e_mask = np.linalg.eigvalsh(x.conj().T @ h @ x)[0]
    assert abs(e_mask - e_low) > 1e-3


def test_eigen_solver_positive_definite_matches_generalized_eigh():
    """With S positive definite, canonical orthogonalisation keeps every
    direction and equals the generalized Hermitian eigenproblem."""
    from scipy.linalg import eigh
    rng = np.random.default_rng(12)
    a = rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6))
    h = a + a.conj().T
    b = rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6))
    s = b @ b.conj().T + 0.5 * np.eye(6)
    ev, vec = phf.eigen_solver(h, s)
    ref = eigh(h, s, eigvals_only=True)
    assert np.allclose(ev, ref, atol=1e-10)
    # S-orthonormal eigenvectors
    assert np.allclose(vec.conj().T @ s @ vec, np.eye(6), atol=1e-10)


def test_make_eandg_cmplxconj0_gradient_is_real():
    """CmplxConj = 0: EandG drops Im Z and returns a real gradient
    (PHF.py:62-63,80-81); the energy does not depend on Im Z."""
    case = hubbard_general_ref(4, 4.0, 6, zscale=0.2)
    ga = grid_args(4, (6, 4), 0)
    cfg = phf.ObjectiveConfig(nso=8, nocc=4, cmplxconj=0, ncisp=ga["ncisp"])
    bh = lambda z: O.buildhsg(case["hone"], case["htwo"], case["ref"], z, cmplxconj=0, nocc=4, **ga)
    f = phf.make_eandg(cfg, bh, lambda gm, rm, s, e0, fsp, fpg, fk, cc: O.localg(gm, rm, s, e0, fsp, fpg, fk, cc))
    x = np.concatenate([case["z"].real.ravel(), case["z"].imag.ravel()])
    e1, g1 = f(x)
    assert np.all(g1[16:] == 0)
    x2 = x.copy()
    x2[16:] += 0.3
    assert abs(f(x2)[0] - e1) < 1e-13


def test_hamiltonian_cache_key_sees_every_entry(monkeypatch):
    """The drop-in's Hamiltonian cache key: a change of one unsampled entry of
    htwo changes the key (full hash up to the size cap, and above it with
    PHF_STRICT_CACHE=1); a new array with equal content keeps the key."""
    from phf_pcc_amd import PHFTools_F as P
    rng = np.random.default_rng(13)
    a = rng.normal(size=(12, 12, 12, 12)) + 0j
    k = P._fingerprint(a)
    assert P._fingerprint(a.copy()) == k
    a[3, 5, 7, 9] = np.nextafter(a[3, 5, 7, 9].real, np.inf)
    assert P._fingerprint(a) != k
    monkeypatch.setattr(P, "FULL_HASH_MAX_BYTES", 1024)
    b = rng.normal(size=(20, 20, 20, 20))
    kb = P._fingerprint(b)
    b[1, 2, 3, 4] += 1.0      # not among the sampled elements
    assert P._fingerprint(b) == kb
    monkeypatch.setenv("PHF_STRICT_CACHE", "1")
    kb = P._fingerprint(b)
    b[1, 2, 3, 5] += 1.0
    assert P._fingerprint(b) != kb
