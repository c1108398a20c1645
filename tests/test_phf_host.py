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
