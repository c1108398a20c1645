"""Synthetic n = 64 workload at full size (NSO = 128, N = 32, L = 192, Lebedev
50 x 16 gamma = 800 grid points, seed 1002): the shapes the default bench line
is measured on.

Checks against the oracle's C restatement of BuildHSG (oracle/buildhsg_cref.c,
PHFTools.f90:10-201) on the whole grid, every element of Hmat, Smat, Gradmat
and Rdmmat to 1e-10 of the largest entry of each:
  * FP64 through the spin-free form (2J - K and K over spatial pairs, K = 4096),
  * FP64 through the spin blocks (D: K = 8192, X1/X2: K = 4096),
  * FP64 through the dense A (the plain form, K = 16384),
  * the PHFTools_F drop-in with the dense host htwo (4.3 GB complex128).
The reduced forms F16X2 and I8X4 give E and G = dE/dZ* within 1e-6 Eh and
1e-5 of the oracle's EandG; the round-to-nearest 16-bit tile at its longest
admitted contraction (K = 8192, the D block) stays within 1e-6 Eh.
"""
import numpy as np
import pytest

from oracle import cref
from oracle import phf_oracle as O
from phf_pcc_amd.hamiltonian import synthetic_problem
from phf_pcc_amd.spin_proj import grid_args

pytestmark = pytest.mark.gpu

N_SPATIAL, NOCC, NGRID, SEED = 64, 32, (50, 16), 1002


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def cfg2():
    from phf_pcc_amd.model import ProjectedHF
    prob = synthetic_problem(N_SPATIAL, NOCC, seed=SEED, build_dense=True, ngrid=NGRID)
    ga = grid_args(N_SPATIAL, NGRID, 0)
    print(f"[cfg2] oracle (C restatement) on {NGRID[0] * NGRID[1]} grid points", flush=True)
    ref_mats = cref.buildhsg(prob.hone, prob.htwo, prob.ref, prob.z, cmplxconj=1, nocc=NOCC, **ga)
    h, s, gm, rm = ref_mats
    ev, vec = O.eigen_solver(h, s)
    g_ref = O.localg(gm, rm, s, ev[0], vec[:, 0] / vec[0, 0], np.ones(1), np.ones(1), 1)[NOCC:, :NOCC]
    model = ProjectedHF(NOCC, prob.ref, ga, factors=(prob.h_spatial, prob.b))
    yield dict(prob=prob, ga=ga, ref=ref_mats, e_ref=float(ev[0].real), g_ref=g_ref, model=model)
    model.ctx.close()


def _structure(m, mode):
    """mode: "spin-free", "spin-blocks" or "dense"."""
    m.ctx.set_blocked(mode != "dense")
    m.ctx.set_spinfree(mode == "spin-free")
    assert m.ctx.blocked() == (mode != "dense") and m.ctx.spinfree() == (mode == "spin-free")


@pytest.mark.parametrize("mode", ["spin-free", "spin-blocks", "dense"])
def test_fp64_matches_c_oracle_full_grid(cfg2, mode):
    m = cfg2["model"]
    _structure(m, mode)
    m.precision = "fp64"
    try:
        gpu = m.buildhsg(cfg2["prob"].z)
    finally:
        _structure(m, "spin-free")
    for name, a, b in zip(("Hmat", "Smat", "Gradmat", "Rdmmat"), gpu, cfg2["ref"]):
        assert _rel(a, b) < 1e-10, name


def test_dropin_host_htwo_matches_c_oracle(cfg2):
    from phf_pcc_amd.PHFTools_F import phftools
    p, ga = cfg2["prob"], cfg2["ga"]
    out = phftools.buildhsg(p.hone, p.htwo, p.ref, p.z, ga["ngrid"], ga["nci"], 1, ga["ncisp"], NOCC, ga["r1"],
                            ga["r2"], ga["rpg"], ga["weightsp"], ga["weightpg"], ga["roota"], ga["rootb"],
                            ga["rooty"], 0, 1, 0)
    phftools.invalidate()
    for name, a, b in zip(("Hmat", "Smat", "Gradmat", "Rdmmat"), out, cfg2["ref"]):
        assert _rel(a, b) < 1e-10, name


def test_objective_matches_oracle_eandg(cfg2):
    m = cfg2["model"]
    m.precision = "fp64"
    e, g = m.energy_and_gradient(cfg2["prob"].z)
    assert abs(e - cfg2["e_ref"]) < 1e-10
    assert np.abs(g - cfg2["g_ref"]).max() < 1e-10


@pytest.mark.parametrize("mode", ["spin-free", "spin-blocks"])
@pytest.mark.parametrize("prec", ["f16x2", "i8x4"])
def test_reduced_forms_within_budget(cfg2, prec, mode):
    m = cfg2["model"]
    _structure(m, mode)
    m.precision = prec
    try:
        e, g = m.energy_and_gradient(cfg2["prob"].z)
    finally:
        m.precision = "fp64"
        _structure(m, "spin-free")
    print(f"[cfg2] {prec} {mode}: |dE| = {abs(e - cfg2['e_ref']):.3e}", flush=True)
    assert abs(e - cfg2["e_ref"]) <= 1e-6
    assert np.abs(g - cfg2["g_ref"]).max() <= 1e-5


def test_rn16_tile_at_its_contraction_limit(cfg2, monkeypatch):
    """The 256 x 256 round-to-nearest 16-bit tile (configuration 5) admits
    contractions up to K = 8192; the D block of n = 64 is exactly that long.
    Forced onto every block, F16X2 stays within 1e-6 Eh of the oracle."""
    m = cfg2["model"]
    monkeypatch.setenv("PHF_H16_CFG", "5")
    m.ctx.autotune(True)               # drops the tuned choices, so the forced one is taken
    _structure(m, "spin-blocks")       # the D block; the spin-free products contract over K = 4096
    m.precision = "f16x2"
    try:
        rep_before = len(m.ctx.tuning_report())
        e, _ = m.energy_and_gradient(cfg2["prob"].z)
        rep = m.ctx.tuning_report()[rep_before:]
    finally:
        m.precision = "fp64"
        monkeypatch.delenv("PHF_H16_CFG")
        m.ctx.autotune(True)
        _structure(m, "spin-free")
    assert '"k": 24576' in rep and '"chosen": {"cfg": 5' in rep   # D block: K = 8192 x three segments
    assert abs(e - cfg2["e_ref"]) <= 1e-6
