"""Batched form: uneven instances in one launch sequence (grouped persistent
ERI GEMM, hipGraph replay) against the single-instance device path and the oracle.

Checks: FP64 batch = single-instance FP64 per instance (|dE| <= 1e-10,
|dG| <= 1e-10) and = the oracle's EandG on a dense instance (1e-10); 16-bit
batch forms within |dE| <= 1e-6 Eh, |dG| <= 1e-5 of FP64; graph replays with
new Z, with other device work between them, match eager single-instance
results, and a replay of the first Z reproduces it bit for bit; reduced-only
storage refuses FP64; autograd through the batch = central differences.
"""
import numpy as np
import pytest

from oracle import phf_oracle as O
from phf_pcc_amd.hamiltonian import synthetic_problem
from tests._cases import grid_args

pytestmark = pytest.mark.gpu

SPECS = [(8, (6, 4), 11), (12, (14, 8), 12), (16, (26, 8), 13), (20, (14, 16), 14), (24, (26, 8), 15)]


def _instances(dense_first=False):
    out, probs = [], []
    for k, (n, ngrid, seed) in enumerate(SPECS):
        p = synthetic_problem(n, n // 2, seed=seed, ngrid=ngrid, build_dense=dense_first and k == 0, zscale=0.05)
        ga = grid_args(n, ngrid, 0)
        out.append(dict(h=p.h_spatial, b=p.b, ref=p.ref, nocc=p.nocc, ngrid=ga["ngrid"], r1=ga["r1"], r2=ga["r2"],
                        rpg=ga["rpg"], weightsp=ga["weightsp"]))
        probs.append((p, ga))
    return out, probs


@pytest.fixture(scope="module")
def setup():
    from phf_pcc_amd.batch import ProjectedHFBatch
    from phf_pcc_amd.model import ProjectedHF
    inst, probs = _instances(dense_first=True)
    batch = ProjectedHFBatch(inst, storage="fp64")
    singles = [ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b)) for p, ga in probs]
    return batch, singles, probs


def _zs(probs, seed):
    rng = np.random.default_rng(seed)
    return [0.05 * (rng.normal(size=p.z.shape) + 1j * rng.normal(size=p.z.shape)) for p, _ in probs]


def test_fp64_batch_matches_single_instances(setup):
    batch, singles, probs = setup
    zs = [p.z for p, _ in probs]
    e, g, hs = batch.energy_and_gradient(zs, "fp64")
    for i, (m, z) in enumerate(zip(singles, zs)):
        e1, g1 = m.energy_and_gradient(z)
        assert abs(e[i] - e1) < 1e-10, i
        assert np.abs(g[i] - g1).max() < 1e-10, i
        assert hs[i, 2] > 0 and abs(hs[i, 3]) < 1e-5 * hs[i, 2]


def test_fp64_batch_matches_oracle(setup):
    batch, _, probs = setup
    p, ga = probs[0]
    e, g, _ = batch.energy_and_gradient([q.z for q, _ in probs], "fp64")
    x = np.concatenate([p.z.real.ravel(), p.z.imag.ravel()])
    e_ref, g_ref = O.eandg(x, p.hone, p.htwo, p.ref.astype(complex), p.nocc, ga)
    assert abs(e[0] - e_ref) < 1e-10
    assert np.abs(np.concatenate([g[0].real.ravel(), g[0].imag.ravel()]) - g_ref).max() < 1e-10


@pytest.mark.parametrize("prec", ["f16x2", "bf16x3"])
def test_grouped_16bit_batch(setup, prec):
    batch, _, probs = setup
    zs = [p.z for p, _ in probs]
    e64, g64, _ = batch.energy_and_gradient(zs, "fp64")
    e, g, _ = batch.energy_and_gradient(zs, prec)
    assert np.abs(e - e64).max() <= 1e-6
    for a, b in zip(g, g64):
        assert np.abs(a - b).max() <= 1e-5
    assert '"grouped": true' in batch.tuning_report()


@pytest.mark.parametrize("prec", ["fp64", "f16x2"])
def test_graph_replays(setup, prec):
    batch, singles, probs = setup
    tol_e, tol_g = (1e-10, 1e-10) if prec == "fp64" else (1e-6, 1e-5)
    z0, z1, z2 = _zs(probs, 1), _zs(probs, 2), _zs(probs, 3)
    batch.set_graph(True)
    batch.autotune(True)          # forces a fresh plan: the next call runs eagerly and captures
    e0, g0, _ = batch.energy_and_gradient(z0, prec)
    assert batch.stats()[2] == 0.0
    for zz in (z1, z2):
        e, g, _ = batch.energy_and_gradient(zz, prec)
        assert batch.stats()[2] == 1.0            # replayed from the graph
        for i, m in enumerate(singles):           # eager work on other streams between replays
            e1, g1 = m.energy_and_gradient(zz[i])
            assert abs(e[i] - e1) <= tol_e, (prec, i)
            assert np.abs(g[i] - g1).max() <= tol_g, (prec, i)
    e0b, g0b, _ = batch.energy_and_gradient(z0, prec)
    assert np.array_equal(e0, e0b)
    for a, b in zip(g0, g0b):
        assert np.array_equal(a, b)


def test_graph_off_equals_graph_on(setup):
    batch, _, probs = setup
    zs = _zs(probs, 4)
    batch.set_graph(False)
    e_a, g_a, _ = batch.energy_and_gradient(zs, "fp64")
    e_b, g_b, _ = batch.energy_and_gradient(zs, "fp64")
    batch.set_graph(True)
    assert np.array_equal(e_a, e_b)
    for a, b in zip(g_a, g_b):
        assert np.array_equal(a, b)


def test_reduced_storage_only_serves_its_mode():
    from phf_pcc_amd import _lib
    from phf_pcc_amd.batch import ProjectedHFBatch
    inst, probs = _instances()
    bt = ProjectedHFBatch(inst[:3], storage="f16x2")
    zs = [p.z for p, _ in probs[:3]]
    e, _, _ = bt.energy_and_gradient(zs)
    ref = ProjectedHFBatch(inst[:3], storage="fp64").energy_and_gradient(zs, "fp64")[0]
    assert np.abs(e - ref).max() <= 1e-6
    with pytest.raises(_lib.PhfError):
        bt.energy_and_gradient(zs, "fp64")


def test_batch_autograd_central_differences(setup):
    """Backward of a weighted sum of the energies against central differences
    of each instance's energy along a direction inside that instance's block
    (differencing the sum itself would lose ~1e-16 |sum E| / h to rounding)."""
    import torch
    from phf_pcc_amd.batch import BatchEnergies
    batch, _, probs = setup
    x0 = np.concatenate([np.concatenate([p.z.real.ravel(), p.z.imag.ravel()]) for p, _ in probs])
    rng = np.random.default_rng(9)
    w = rng.normal(size=batch.n_inst)
    x = torch.tensor(x0, requires_grad=True)
    (BatchEnergies.apply(x, batch, "fp64") * torch.tensor(w)).sum().backward()
    grad = x.grad.numpy()
    h = 1e-5
    off = 0
    for k, sz in enumerate(batch.sizes):
        d = np.zeros_like(x0)
        d[off:off + 2 * sz] = rng.normal(size=2 * sz)
        d /= np.linalg.norm(d)
        ek = lambda v: float(BatchEnergies.apply(torch.tensor(v), batch, "fp64")[k])
        fd = (ek(x0 + h * d) - ek(x0 - h * d)) / (2 * h)
        assert abs(w[k] * fd - grad @ d) <= 1e-6 * (1 + abs(w[k] * fd)), k
        off += 2 * sz


def test_batch_dense_and_blocked_agree():
    """The same instances stored dense (plain contraction) and as spin blocks."""
    from phf_pcc_amd.batch import ProjectedHFBatch
    inst, probs = _instances()
    zs = [p.z for p, _ in probs]
    e_b, g_b, _ = ProjectedHFBatch(inst, storage="fp64").energy_and_gradient(zs, "fp64")
    e_d, g_d, _ = ProjectedHFBatch(inst, storage="fp64", structure=False).energy_and_gradient(zs, "fp64")
    assert np.abs(e_b - e_d).max() < 1e-10
    for a, b in zip(g_b, g_d):
        assert np.abs(a - b).max() < 1e-10
    e_h, _, _ = ProjectedHFBatch(inst, storage="f16x2", structure=False).energy_and_gradient(zs)
    assert np.abs(e_h - e_d).max() <= 1e-6
