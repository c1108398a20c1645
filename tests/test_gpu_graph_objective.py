"""Graph replay of the device-resident objective (phf_eval_objective): after
an eager evaluation the launch sequence is captured and later evaluations
with the same shapes replay it, with only the pinned C0 block rewritten.

Every replay is checked bit for bit against an eager evaluation of the same Z
on the same context (profiling on forces the eager path; the tuned tiles are
the context's own), with other device work and synchronizations between the
replays; the capture is reported in the tuning log; a setter (a new grid)
drops the graph and the next evaluation captures again."""
import numpy as np
import pytest

from phf_pcc_amd.hamiltonian import synthetic_problem
from tests._cases import grid_args

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,structure", [("fp64", True), ("f16x2", True), ("i8x4", True), ("fp64", False)])
def test_replays_equal_eager(prec, structure):
    from phf_pcc_amd.model import ProjectedHF
    p = synthetic_problem(16, 16, seed=91, zscale=0.05)
    ga = grid_args(16, (14, 8), 0)
    m = ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b), structure=structure, precision=prec)
    rng = np.random.default_rng(7)
    zs = [p.z + 0.01 * (rng.normal(size=p.z.shape) + 1j * rng.normal(size=p.z.shape)) for _ in range(4)]
    try:
        m.ctx.set_profiling(False)
        m.energy_and_gradient(p.z)                     # eager, then captured
        assert '"objective_graph": "captured"' in m.ctx.tuning_report()
        mats = rng.normal(size=(64, 24, 24)) + 1j * rng.normal(size=(64, 24, 24))
        for z in zs:
            m.ctx.set_profiling(False)
            e_r, g_r = m.energy_and_gradient(z)       # replay
            m.ctx.invert_batched(mats)                 # other device work (synchronised) between replays
            m.ctx.set_profiling(True)
            e_e, g_e = m.energy_and_gradient(z)       # eager
            assert e_r == e_e
            assert np.array_equal(g_r, g_e)
        # a different Z gives a different energy (the replay reads the new C0)
        m.ctx.set_profiling(False)
        assert m.energy_and_gradient(zs[0])[0] != m.energy_and_gradient(zs[1])[0]
    finally:
        m.ctx.close()


def test_setter_drops_the_graph():
    """A setter that changes what an evaluation launches (here set_blocked)
    bumps the context's generation: the next evaluation runs eagerly and
    captures again, and its result equals the earlier replays'."""
    from phf_pcc_amd.model import ProjectedHF
    p = synthetic_problem(16, 16, seed=92, zscale=0.05)
    m = ProjectedHF(p.nocc, p.ref, grid_args(16, (14, 8), 0), factors=(p.h_spatial, p.b), structure=True,
                    precision="fp64")
    try:
        m.ctx.set_profiling(False)
        e1, g1 = m.energy_and_gradient(p.z)
        e2, g2 = m.energy_and_gradient(p.z)
        assert m.ctx.tuning_report().count('"objective_graph": "captured"') == 1
        m.ctx.set_blocked(True)
        e3, g3 = m.energy_and_gradient(p.z)
        e4, g4 = m.energy_and_gradient(p.z)
        assert m.ctx.tuning_report().count('"objective_graph": "captured"') == 2
        assert e1 == e2 == e3 == e4
        assert np.array_equal(g1, g4)
    finally:
        m.ctx.close()
