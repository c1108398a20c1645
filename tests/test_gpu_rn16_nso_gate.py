"""The round-to-nearest 16-bit tile's NSO limit.

Its products truncate within each 64-product partial, a bias that grows with
NSO^2: F16X2 on it is 4.0e-7 Eh from FP64 at n = 64 (NSO = 128) but 1.34e-6 for
an n = 80 instance of the 64-instance batch, although that instance's
spin-free contraction (K = 6400) is inside the tile's length limit.  At
n = 80 (NSO = 160) the tuners leave the tile out (no "grouped_256x256" and no
configuration 5 chosen) and F16X2 stays within 1e-6 Eh of FP64."""
import numpy as np
import pytest

from phf_pcc_amd.hamiltonian import synthetic_problem
from tests._cases import grid_args

pytestmark = pytest.mark.gpu


def test_rn16_tile_left_out_above_nso_128():
    from phf_pcc_amd.model import ProjectedHF
    n = 80
    p = synthetic_problem(n, n // 2, seed=980, zscale=0.05)
    ga = grid_args(n, (50, 24), 0)
    m = ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b), precision="fp64")
    try:
        assert m.ctx.spinfree()
        e64, g64 = m.energy_and_gradient(p.z)
        m.ctx.autotune(True)
        m.precision = "f16x2"
        start = len(m.ctx.tuning_report())
        e, g = m.energy_and_gradient(p.z)
        rep = m.ctx.tuning_report()[start:]
        print(f"n=80 f16x2: |dE| = {abs(e - e64):.3e}; tuning: {rep[:400]}", flush=True)
        assert '"chosen": "grouped_256x256"' not in rep
        assert '"chosen": {"cfg": 5' not in rep
        assert abs(e - e64) <= 1e-6
        assert np.abs(g - g64).max() <= 1e-5
    finally:
        m.ctx.close()
