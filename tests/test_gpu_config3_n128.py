"""Synthetic n = 128 workload at full size (NSO = 256, N = 64, L = 384,
Lebedev 50 x 24 gamma = 1200 grid points, seed 1003), the mixed-precision
configuration.  Its dense <ij||kl> (68.7 GB as complex128) does not fit a
host, so:
  * on a 6 x 4 sub-grid at the full n (real Z), FP64 (spin-free form: K = 16384;
    spin blocks, D: K = 32768) is checked against the oracle's BuildHSG with
    MakeGamma from the factors (oracle/factored.py, pinned against the dense
    MakeGamma), every element of Hmat/Smat/Gradmat/Rdmmat to 1e-10 of its
    largest entry;
  * on the whole grid, FP64 through the spin-free form and through the spin
    blocks equals FP64 through the dense A (K = 65536, the plain form) to the
    same bar;
  * I8X4 and FP32 give E within 1e-6 Eh and G within 1e-5 of FP64; the
    16-bit forms carry the truncation bias of the 16-bit MFMA (1.3e-6 at this
    size over the blocks), which stays below 3e-6;
  * the contraction gates of the 256 x 256 tiles: the int8 tile runs the
    D block (K = 32768 <= 33285, the int32 class-3 bound) within 1e-6 Eh, and
    the round-to-nearest 16-bit tile refuses it (K > 8192).
"""
import numpy as np
import pytest

from oracle import phf_oracle as O
from oracle.factored import gamma_factored
from phf_pcc_amd.hamiltonian import synthetic_problem
from phf_pcc_amd.spin_proj import grid_args

pytestmark = pytest.mark.gpu


def _equator_subgrid(n):
    ga = grid_args(n, (6, 4), 0)
    idx = np.nonzero(np.abs(ga["rootb"] - np.pi / 2) < 1e-12)[0]
    assert idx.size == 4
    sub = dict(ga)
    sub.update(ngrid=np.array([idx.size, 4]), r1=ga["r1"][idx], roota=ga["roota"][idx], rootb=ga["rootb"][idx],
               weightsp=ga["weightsp"][idx])
    return sub

N_SPATIAL, NOCC, NGRID, SEED = 128, 64, (50, 24), 1003


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _structure(m, mode):
    """mode: "spin-free", "spin-blocks" or "dense"."""
    m.ctx.set_blocked(mode != "dense")
    m.ctx.set_spinfree(mode == "spin-free")
    assert m.ctx.blocked() == (mode != "dense") and m.ctx.spinfree() == (mode == "spin-free")


@pytest.fixture(scope="module")
def cfg3():
    from phf_pcc_amd.model import ProjectedHF
    prob = synthetic_problem(N_SPATIAL, NOCC, seed=SEED, ngrid=NGRID)
    ga = grid_args(N_SPATIAL, NGRID, 0)
    model = ProjectedHF(NOCC, prob.ref, ga, factors=(prob.h_spatial, prob.b))
    assert model.ctx.blocked()
    model.precision = "fp64"
    e64, g64 = model.energy_and_gradient(prob.z)
    yield dict(prob=prob, ga=ga, model=model, e64=e64, g64=g64)
    model.ctx.close()


@pytest.mark.parametrize("mode", ["spin-free", "spin-blocks"])
def test_fp64_matches_factored_oracle_on_subgrid(cfg3, mode):
    m, p = cfg3["model"], cfg3["prob"]
    _structure(m, mode)
    # a small grid is far from exact for N = 64 electrons: with complex Z, Im H
    # then exceeds the Hermiticity tolerance (PHFTools.f90:195-200) and the
    # evaluation stops, as the reference's would.  With real Z, H is real on
    # a grid closed under (alpha, gamma) -> (-alpha, -gamma): the four equator
    # points of the 6-point Lebedev rule (the poles sit at azimuth pi/2, whose
    # mirror is not in the rule) times the 4-point trapezoid in gamma.
    sub = _equator_subgrid(N_SPATIAL)
    z = p.z.real.astype(complex)
    m._set_grid(p.ref.astype(complex), sub)
    try:
        gpu = m.buildhsg(z)
    finally:
        m._set_grid(p.ref.astype(complex), cfg3["ga"])
        _structure(m, "spin-free")
    print("[cfg3] factored oracle on 24 grid points", flush=True)
    ref = O.buildhsg(p.hone, None, p.ref, z, cmplxconj=1, nocc=NOCC, gamma=gamma_factored(p.b, p.ref), **sub)
    for name, a, b in zip(("Hmat", "Smat", "Gradmat", "Rdmmat"), gpu, ref):
        assert _rel(a, b) < 1e-10, name


def test_fp64_structures_equal_dense_full_grid(cfg3):
    m, p = cfg3["model"], cfg3["prob"]
    out = {}
    try:
        for mode in ("spin-free", "spin-blocks", "dense"):
            _structure(m, mode)
            out[mode] = m.buildhsg(p.z)
    finally:
        _structure(m, "spin-free")
    for mode in ("spin-free", "spin-blocks"):
        for name, a, b in zip(("Hmat", "Smat", "Gradmat", "Rdmmat"), out[mode], out["dense"]):
            assert _rel(a, b) < 1e-10, (mode, name)


@pytest.mark.parametrize("mode", ["spin-free", "spin-blocks"])
@pytest.mark.parametrize("prec,tol_e", [("i8x4", 1e-6), ("fp32", 1e-6), ("f16x2", 3e-6), ("bf16x3", 3e-6)])
def test_reduced_forms(cfg3, prec, tol_e, mode):
    m = cfg3["model"]
    _structure(m, mode)
    m.precision = prec
    try:
        e, g = m.energy_and_gradient(cfg3["prob"].z)
    finally:
        m.precision = "fp64"
        _structure(m, "spin-free")
    print(f"[cfg3] {prec} {mode}: |dE| = {abs(e - cfg3['e64']):.3e}  "
          f"max|dG| = {np.abs(g - cfg3['g64']).max():.3e}", flush=True)
    assert abs(e - cfg3["e64"]) <= tol_e
    assert np.abs(g - cfg3["g64"]).max() <= 1e-5


def test_tile_contraction_gates(cfg3, monkeypatch):
    from phf_pcc_amd import _lib
    m, p = cfg3["model"], cfg3["prob"]
    _structure(m, "spin-blocks")   # the D block (K = 32768); the spin-free products have K = 16384
    try:
        # the int8 256 x 256 tile (configuration 4) at K = 32768, under its 33285 limit
        monkeypatch.setenv("PHF_H16_CFG", "4")
        m.ctx.autotune(True)
        m.precision = "i8x4"
        start = len(m.ctx.tuning_report())
        e, _ = m.energy_and_gradient(p.z)
        rep = m.ctx.tuning_report()[start:]
        assert '"k": 327680' in rep and '"chosen": {"cfg": 4' in rep   # D: K = 32768 x ten segments
        assert abs(e - cfg3["e64"]) <= 1e-6
        # the round-to-nearest tile (configuration 5) refuses K = 32768
        monkeypatch.setenv("PHF_H16_CFG", "5")
        m.ctx.autotune(True)
        m.precision = "f16x2"
        with pytest.raises(_lib.PhfError) as err:
            m.energy_and_gradient(p.z)
        assert err.value.code == _lib.PHF_ERR_UNSUPPORTED
    finally:
        m.precision = "fp64"
        monkeypatch.delenv("PHF_H16_CFG", raising=False)
        m.ctx.autotune(True)
        _structure(m, "spin-free")
    # the context still evaluates after the refusal
    e, _ = m.energy_and_gradient(p.z)
    assert abs(e - cfg3["e64"]) < 1e-12
