"""The uneven batch workload at full size: 64 synthetic instances (n uniform on
{16, ..., 96}, N = n/2, L = 3n, grids Lebedev {6..50} x gamma {4, 8, 16},
seed 1004; 15144 grid points), evaluated in one batch (grouped persistent ERI
GEMM for the 16-bit and int8 forms, hipGraph replay).

Checks:
  * FP64 batch = the single-instance device path, instance by instance:
    E and G = dE/dZ* to 1e-10 where the instance passes the Hermiticity test
    of PHFTools.f90:195-200, and the same status (NotHermitianError from the
    single path) where it does not.  The single-instance path is checked
    against the oracle in the other GPU tests.
  * F16X2, I8X4 and FP32 batches within 1e-6 Eh and 1e-5 (gradient) of the
    FP64 batch on every instance (a failed Hermiticity test does not change
    the arithmetic: E0 = Re H / |Re S| is formed all the same).  The
    instances' Refs are spin-restricted, so these run the spin-free form; the
    F16X2 batch over the spin blocks (PHF_SPINFREE=0) is held to the same bar.
  * graph replays: a replay of the first Z after another Z reproduces it bit
    for bit.
"""
import numpy as np
import pytest

from phf_pcc_amd.hamiltonian import batch_instances
from phf_pcc_amd.spin_proj import grid_args

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg4():
    from phf_pcc_amd.batch import ProjectedHFBatch
    inst, zs, ns, gs = batch_instances(64, 1004)
    bt = ProjectedHFBatch(inst, storage="fp64")
    e, g, hs = bt.energy_and_gradient(zs, "fp64", strict=False)
    status = bt.status.copy()
    print(f"[cfg4] fp64 batch: {int((status != 0).sum())} of 64 instances fail the Hermiticity test", flush=True)
    yield dict(inst=inst, zs=zs, ns=ns, bt=bt, e=e.copy(), g=[x.copy() for x in g], status=status)
    bt.close()


def test_fp64_batch_equals_single_instances(cfg4):
    from phf_pcc_amd import _lib
    from phf_pcc_amd.model import ProjectedHF
    for i, (inst, z) in enumerate(zip(cfg4["inst"], cfg4["zs"])):
        n = inst["h"].shape[0]
        ga = grid_args(n, tuple(int(v) for v in inst["ngrid"]), 0)
        m = ProjectedHF(inst["nocc"], inst["ref"], ga, factors=(inst["h"], inst["b"]))
        try:
            if cfg4["status"][i] == 0:
                e1, g1 = m.energy_and_gradient(z)
                assert abs(cfg4["e"][i] - e1) < 1e-10, i
                assert np.abs(cfg4["g"][i] - g1).max() < 1e-10, i
            else:
                assert cfg4["status"][i] == _lib.PHF_ERR_NOT_HERMITIAN, i
                with pytest.raises(_lib.NotHermitianError):
                    m.energy_and_gradient(z)
        finally:
            m.ctx.close()


@pytest.mark.parametrize("prec,spinfree", [("f16x2", "1"), ("i8x4", "1"), ("fp32", "1"), ("f16x2", "0")])
def test_reduced_batch_forms(cfg4, prec, spinfree, monkeypatch):
    from phf_pcc_amd.batch import ProjectedHFBatch
    monkeypatch.setenv("PHF_SPINFREE", spinfree)   # read when the instances' contexts are created
    bt = ProjectedHFBatch(cfg4["inst"], storage=prec)
    try:
        e, g, _ = bt.energy_and_gradient(cfg4["zs"], prec, strict=False)
        assert np.array_equal(bt.status, cfg4["status"])
        de = np.abs(e - cfg4["e"])
        dg = max(float(np.abs(a - b).max()) for a, b in zip(g, cfg4["g"]))
        print(f"[cfg4] {prec}: max|dE| over 64 instances {de.max():.3e} (n = {cfg4['ns'][de.argmax()]}), "
              f"max|dG| {dg:.3e}", flush=True)
        assert de.max() <= 1e-6
        assert dg <= 1e-5
        assert '"grouped": true' in bt.tuning_report()   # one grouped launch, FP32 included
    finally:
        bt.close()


def test_graph_replay_reproduces_first_z(cfg4):
    bt = cfg4["bt"]
    rng = np.random.default_rng(5)
    e1, g1, _ = bt.energy_and_gradient(cfg4["zs"], "fp64", strict=False)
    other = [z + 0.01 * (rng.normal(size=z.shape) + 1j * rng.normal(size=z.shape)) for z in cfg4["zs"]]
    e2, _, _ = bt.energy_and_gradient(other, "fp64", strict=False)
    assert np.abs(e2 - e1).max() > 1e-8
    e3, g3, _ = bt.energy_and_gradient(cfg4["zs"], "fp64", strict=False)
    assert bt.stats()[2] == 1.0            # a graph replay
    assert np.array_equal(e3, e1)
    assert all(np.array_equal(a, b) for a, b in zip(g3, g1))
    assert np.abs(e1 - cfg4["e"]).max() < 1e-10
