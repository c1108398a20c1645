"""The density pack fused with the 16-bit/int8 split (one pass per grid point
for NSO <= 128) against the two-pass path over FP64 rows (PHF_FUSED_SPLIT=0),
through the device objective and buildhsg: the fused kernel runs the same
MFMA sequence and the same split arithmetic, so every result must be
bit-identical.  Cases: spin blocks at n = 32 (no padding columns) and n = 13
(padded blocks: zeroed padding columns), the dense contraction order
(blocks off), every reduced form, and padding rows (grid points not a
multiple of the tile height)."""
import os

import numpy as np
import pytest

from phf_pcc_amd.hamiltonian import synthetic_problem
from tests._cases import grid_args

pytestmark = pytest.mark.gpu


def _run(model, z, fused):
    old = os.environ.get("PHF_FUSED_SPLIT")
    os.environ["PHF_FUSED_SPLIT"] = "1" if fused else "0"
    try:
        e, g = model.energy_and_gradient(z)
        h, s, gm, rm = model.buildhsg(z)
    finally:
        if old is None:
            del os.environ["PHF_FUSED_SPLIT"]
        else:
            os.environ["PHF_FUSED_SPLIT"] = old
    return e, g, h, s, gm, rm


@pytest.mark.parametrize("n,ngrid,blocked", [(32, (14, 8), True), (13, (6, 5), True), (24, (14, 8), False)])
@pytest.mark.parametrize("prec", ["bf16x3", "f16x2", "i8x4"])
def test_fused_equals_two_pass(n, ngrid, blocked, prec):
    from phf_pcc_amd.model import ProjectedHF
    p = synthetic_problem(n, n // 2, seed=300 + n, zscale=0.05)
    ga = grid_args(n, ngrid, 0)
    model = ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b), precision=prec)
    model.ctx.set_blocked(blocked)
    model.ctx.set_spinfree(False)   # the spin blocks' column map (tests/test_gpu_spinfree.py covers the other)
    assert model.ctx.blocked() == blocked and not model.ctx.spinfree()
    print(f"fused split n={n} G={int(np.prod(ngrid))} blocked={blocked} {prec}", flush=True)
    a = _run(model, p.z, True)
    b = _run(model, p.z, False)
    assert a[0] == b[0]
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("n,nocc", [(64, 32), (48, 36)])
@pytest.mark.parametrize("spinfree", [False, True])
def test_fused_equals_two_pass_chunks(n, nocc, spinfree):
    """N = 32 stages both 16-deep chunks of the density product at once; N = 36
    takes the chunk loop (three chunks, the last one partial): both bit for bit
    against the two-pass split, through the spin blocks and the spin-free form."""
    from phf_pcc_amd.model import ProjectedHF
    p = synthetic_problem(n, nocc, seed=310 + n, zscale=0.05)
    ga = grid_args(n, (6, 5), 0)
    model = ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b), precision="f16x2")
    model.ctx.set_blocked(True)
    model.ctx.set_spinfree(spinfree)
    assert model.ctx.spinfree() == spinfree
    print(f"fused split chunks n={n} N={nocc} spinfree={spinfree}", flush=True)
    a = _run(model, p.z, True)
    b = _run(model, p.z, False)
    assert a[0] == b[0]
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)
