"""Every tile configuration of the reduced-precision ERI GEMM, pinned one at a
time (PHF_H16_CFG=<index>,<splitk>), against the FP64 contraction.

Configurations 0-3 walk the product segments one after another (every 16-bit
form, I8X4, and FP32 with its one segment of four 16x16x4 MFMAs per fragment;
2 and 3 fold into FP32 and are candidates only when pinned or allowed); 4 is the 256 x 256 exact int8 tile (I8X4 only; one exact int32 sum
per product class, added to Gamma in FP64); 5 is the 256 x 256 16-bit tile
with round-to-nearest accumulation of two-MFMA partials (no FP64 fold tile).  Checks, on a spin blocked and a
dense instance and on the grouped batch launch: |dE| <= 1e-6 Eh and max |dG|
<= 1e-5 against FP64; the int8 tile against the FP64-fold int8 segment walk,
|dE| <= 1e-9 and |dG| <= 1e-8 (both exact up to the dropped classes);
configurations refuse the forms they do not serve.
"""
import os

import numpy as np
import pytest

from phf_pcc_amd.hamiltonian import synthetic_problem
from tests._cases import grid_args

pytestmark = pytest.mark.gpu

SEGMENT_CFGS = [0, 1, 2, 3]
I8_TILE = 4
RN16_TILE = 5


@pytest.fixture(scope="module")
def prob():
    p = synthetic_problem(16, 16, seed=77, zscale=0.05)
    return p, grid_args(16, (14, 8), 0)


class _pinned:
    def __init__(self, cfg, sk):
        self.v = f"{cfg},{sk}"

    def __enter__(self):
        os.environ["PHF_H16_CFG"] = self.v

    def __exit__(self, *a):
        os.environ.pop("PHF_H16_CFG", None)


def _eval(prob, prec, structure=True, cfg=None, sk=1):
    """structure: True / "spinfree" (the spin-free form of the restricted Ref),
    "blocks" (the spin blocks D, X1, X2), False (the dense A)."""
    from phf_pcc_amd.model import ProjectedHF
    p, ga = prob
    m = ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b), structure=bool(structure), precision=prec)
    m.ctx.set_spinfree(structure != "blocks")
    assert m.ctx.spinfree() == (structure in (True, "spinfree"))
    if cfg is None:
        return m.energy_and_gradient(p.z)
    with _pinned(cfg, sk):
        out = m.energy_and_gradient(p.z)
    assert f'"cfg": {cfg}, "splitk": {sk}' in m.ctx.tuning_report()
    return out


@pytest.fixture(scope="module")
def fp64(prob):
    return {st: _eval(prob, "fp64", structure=st) for st in (True, "blocks", False)}


@pytest.mark.parametrize("structure", [True, "blocks", False])
@pytest.mark.parametrize("sk", [1, 2])
@pytest.mark.parametrize("cfg", SEGMENT_CFGS)
def test_f16x2_config(prob, fp64, cfg, sk, structure):
    e64, g64 = fp64[structure]
    e, g = _eval(prob, "f16x2", structure, cfg, sk)
    assert abs(e - e64) <= 1e-6, (cfg, sk, e - e64)
    assert np.abs(g - g64).max() <= 1e-5


@pytest.mark.parametrize("prec", ["bf16x3", "i8x4"])
@pytest.mark.parametrize("cfg", SEGMENT_CFGS)
def test_other_forms_config(prob, fp64, prec, cfg):
    e64, g64 = fp64[True]
    e, g = _eval(prob, prec, True, cfg, 1)
    assert abs(e - e64) <= 1e-6, (prec, cfg, e - e64)
    assert np.abs(g - g64).max() <= 1e-5


@pytest.mark.parametrize("structure", [True, False])
@pytest.mark.parametrize("sk", [1, 2])
@pytest.mark.parametrize("cfg", SEGMENT_CFGS)
def test_fp32_config(prob, fp64, cfg, sk, structure):
    e64, g64 = fp64[structure]
    e, g = _eval(prob, "fp32", structure, cfg, sk)
    assert abs(e - e64) <= 1e-6, (cfg, sk, e - e64)
    assert np.abs(g - g64).max() <= 1e-5


def test_configs_refuse_other_forms(prob):
    from phf_pcc_amd import _lib
    for prec, cfg in [("f16x2", I8_TILE), ("bf16x3", I8_TILE), ("i8x4", RN16_TILE), ("fp32", I8_TILE),
                      ("fp32", RN16_TILE)]:
        with pytest.raises(_lib.PhfError):
            _eval(prob, prec, True, cfg, 1)


@pytest.mark.parametrize("structure", [True, False])
@pytest.mark.parametrize("sk", [1, 2])
@pytest.mark.parametrize("prec", ["f16x2", "bf16x3"])
def test_rn16_tile(prob, fp64, prec, sk, structure):
    e64, g64 = fp64[structure]
    e, g = _eval(prob, prec, structure, RN16_TILE, sk)
    assert abs(e - e64) <= 1e-6, (prec, sk, e - e64)
    assert np.abs(g - g64).max() <= 1e-5


@pytest.mark.parametrize("structure", [True, False])
@pytest.mark.parametrize("sk", [1, 2])
def test_int8_tile(prob, fp64, sk, structure):
    e64, g64 = fp64[structure]
    e, g = _eval(prob, "i8x4", structure, I8_TILE, sk)
    assert abs(e - e64) <= 1e-6, (sk, e - e64)
    assert np.abs(g - g64).max() <= 1e-5
    e0, g0 = _eval(prob, "i8x4", structure, 0, 1)
    assert abs(e - e0) <= 1e-9
    assert np.abs(g - g0).max() <= 1e-8


@pytest.mark.parametrize("prec,cfg,sk", [("f16x2", 0, 1), ("f16x2", 0, 2), ("f16x2", 1, 1), ("bf16x3", 0, 1),
                                         ("i8x4", I8_TILE, 1), ("i8x4", I8_TILE, 2), ("f16x2", RN16_TILE, 1),
                                         ("f16x2", RN16_TILE, 2), ("fp32", 0, 1), ("fp32", 0, 2), ("fp32", 1, 1)])
def test_grouped_batch_config(prec, cfg, sk):
    from phf_pcc_amd.batch import ProjectedHFBatch
    inst, zs = [], []
    for n, ngrid, seed in [(8, (6, 4), 21), (12, (14, 8), 22), (16, (26, 8), 23)]:
        p = synthetic_problem(n, n // 2, seed=seed, zscale=0.05)
        ga = grid_args(n, ngrid, 0)
        inst.append(dict(h=p.h_spatial, b=p.b, ref=p.ref, nocc=p.nocc, ngrid=ga["ngrid"], r1=ga["r1"], r2=ga["r2"],
                         rpg=ga["rpg"], weightsp=ga["weightsp"]))
        zs.append(p.z)
    bt = ProjectedHFBatch(inst, storage="fp64")
    e64, g64, _ = bt.energy_and_gradient(zs, "fp64")
    with _pinned(cfg, sk):
        e, g, _ = bt.energy_and_gradient(zs, prec)
    assert f'"chosen": {{"cfg": {cfg}, "splitk": {sk}' in bt.tuning_report()
    assert np.abs(e - e64).max() <= 1e-6
    for a, b in zip(g, g64):
        assert np.abs(a - b).max() <= 1e-5


@pytest.mark.parametrize("structure", ["blocks", "spinfree"])
@pytest.mark.parametrize("prec", ["f16x2", "i8x4"])
def test_exchange_blocks_grouped_equal_two_launches(prob, fp64, prec, structure):
    """X1 and X2 (or the spin-free form's two products) as one grouped launch
    (PHF_XGROUP=1) against two launches (PHF_XGROUP=0): the same tiles and
    arithmetic, so results agree to 1e-9."""
    out = {}
    for g in ("0", "1"):
        os.environ["PHF_XGROUP"] = g
        try:
            out[g] = _eval(prob, prec, structure)
        finally:
            os.environ.pop("PHF_XGROUP", None)
    (e0, g0), (e1, g1) = out["0"], out["1"]
    assert abs(e0 - e1) <= 1e-9
    assert np.abs(g0 - g1).max() <= 1e-8
    e64, _ = fp64[True]
    assert abs(e1 - e64) <= 1e-6


@pytest.mark.parametrize("spinfree", [False, True])
@pytest.mark.parametrize("prec", ["f16x2", "i8x4"])
def test_exchange_blocks_grouped_256_tile(prec, spinfree):
    """X1 and X2 (or the spin-free form's two products) as one grouped launch
    of the 256 x 256 tile (PHF_XGROUP=2: the round-to-nearest 16-bit tile,
    the exact int8 tile) on an n = 32 instance, whose exchange blocks have
    256-row multiples: within the form's budget of FP64 (|dE| <= 1e-6,
    |dG| <= 1e-5)."""
    from phf_pcc_amd.model import ProjectedHF
    p = synthetic_problem(32, 32, seed=78, zscale=0.05)
    ga = grid_args(32, (50, 16), 0)
    m64 = ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b), structure=True, precision="fp64")
    m64.ctx.set_spinfree(False)
    e64, g64 = m64.energy_and_gradient(p.z)
    m64.ctx.close()
    os.environ["PHF_XGROUP"] = "2"
    try:
        m = ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b), structure=True, precision=prec)
        m.ctx.set_spinfree(spinfree)
        assert m.ctx.spinfree() == spinfree
        e, g = m.energy_and_gradient(p.z)
        rep = m.ctx.tuning_report()
        m.ctx.close()
    finally:
        os.environ.pop("PHF_XGROUP", None)
    assert '"chosen": "grouped_256x256"' in rep, rep
    assert abs(e - e64) <= 1e-6
    assert np.abs(g - g64).max() <= 1e-5
