"""The per-point kernels fused across launches, against the separate launches.

Epilogue: the per-point epilogue in one kernel (fock_adjoint_kernel: the Fock product
FX = (h + Gamma) X, Q = C0^H FX, the energy e = (tr(Hc X) + tr Q)/2 and the
adjoint Y = FX - X (Q - e I) for NSO <= 128, N <= 32, one density per point)
against the separate launches (PHF_FUSED_EPILOGUE=0).  Only the FP64
summation order differs: energies and gradients to 1e-12 relative, every
element of Hmat/Smat/Gradmat/Rdmmat to 1e-12 of its largest entry; both paths
are checked against the oracle in the other GPU tests.  Cases: n = 64 (NSO =
128, N = 32: the full tile), n = 13 and 24 (odd sizes: padded rows and
columns), the spin blocks and the dense A, FP64 and F16X2.

Inverse + solve: lu_inverse_solve_kernel (N <= 32, one density per point)
inverts M_g in LDS and forms X_g = T_g M_g^{-1} in the same workgroup,
against the inverse kernel and the solve GEMM (PHF_FUSED_SOLVE=0), to the
same bounds; the overlaps' determinants enter Hmat/Smat and the gradient.

Split spin-free GEMM: the direct grouped launch as two launches on two
streams against one launch (PHF_SPLIT_RJK=0).  Side-stream pre-stages: Hc,
the T_g combine and the determinant weights on side streams (PHF_SIDE_PRE=1)
against the main stream."""
import os

import numpy as np
import pytest

from phf_pcc_amd.hamiltonian import synthetic_problem
from tests._cases import grid_args

pytestmark = pytest.mark.gpu


def _eval(m, z, fused, var="PHF_FUSED_EPILOGUE"):
    old = os.environ.get(var)
    os.environ[var] = "1" if fused else "0"
    try:
        e, g = m.energy_and_gradient(z)
        res = m.buildhsg(z)
    finally:
        if old is None:
            del os.environ[var]
        else:
            os.environ[var] = old
    return e, g, res


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("n,nocc,ngrid", [(64, 32, (50, 16)), (13, 6, (14, 8)), (24, 12, (14, 8)), (18, 4, (14, 8))])
@pytest.mark.parametrize("prec", ["fp64", "f16x2"])
@pytest.mark.parametrize("structure", ["spin-free", "dense"])
def test_fused_epilogue_equals_launches(n, nocc, ngrid, prec, structure):
    from phf_pcc_amd.model import ProjectedHF
    p = synthetic_problem(n, nocc, seed=900 + n, zscale=0.05)
    ga = grid_args(n, ngrid, 0)
    m = ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b), precision=prec)
    m.ctx.set_blocked(structure != "dense")
    print(f"fused epilogue n={n} N={nocc} G={int(np.prod(ngrid))} {prec} {structure}", flush=True)
    e1, g1, r1 = _eval(m, p.z, True)
    e0, g0, r0 = _eval(m, p.z, False)
    m.ctx.close()
    assert abs(e1 - e0) <= 1e-12 * max(1.0, abs(e0))
    assert np.abs(g1 - g0).max() <= 1e-12 * max(1.0, np.abs(g0).max())
    for name, a, b in zip(["Hmat", "Smat", "Gradmat", "Rdmmat"], r1, r0):
        assert _rel(a, b) <= 1e-12, name


@pytest.mark.parametrize("n,nocc,ngrid", [(64, 32, (50, 16)), (13, 6, (14, 8)), (24, 12, (14, 8)), (18, 4, (14, 8))])
@pytest.mark.parametrize("prec", ["fp64", "f16x2"])
def test_fused_inverse_solve_equals_launches(n, nocc, ngrid, prec):
    from phf_pcc_amd.model import ProjectedHF
    p = synthetic_problem(n, nocc, seed=950 + n, zscale=0.05)
    ga = grid_args(n, ngrid, 0)
    m = ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b), precision=prec)
    print(f"fused inverse+solve n={n} N={nocc} G={int(np.prod(ngrid))} {prec}", flush=True)
    e1, g1, r1 = _eval(m, p.z, True, "PHF_FUSED_SOLVE")
    e0, g0, r0 = _eval(m, p.z, False, "PHF_FUSED_SOLVE")
    m.ctx.close()
    assert abs(e1 - e0) <= 1e-12 * max(1.0, abs(e0))
    assert np.abs(g1 - g0).max() <= 1e-12 * max(1.0, np.abs(g0).max())
    for name, a, b in zip(["Hmat", "Smat", "Gradmat", "Rdmmat"], r1, r0):
        assert _rel(a, b) <= 1e-12, name


@pytest.mark.parametrize("n,nocc,ngrid", [(64, 32, (50, 16)), (24, 12, (14, 8))])
@pytest.mark.parametrize("prec", ["f16x2", "bf16x3"])
def test_split_spinfree_gemm_equals_one_launch(n, nocc, ngrid, prec):
    """The direct grouped GEMM of the spin-free form as two launches on two
    streams (Gamma_aa / Gamma_bb product tiles, then the unpack, on the main
    stream; the exchange tiles on a side stream) against one launch followed
    by the unpack (PHF_SPLIT_RJK=0): the same tiles and the same unpack, so
    the results agree to the last bit; checked to 1e-12 like the cases above."""
    from phf_pcc_amd.model import ProjectedHF
    p = synthetic_problem(n, nocc, seed=970 + n, zscale=0.05)
    ga = grid_args(n, ngrid, 0)
    m = ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b), precision=prec)
    m.ctx.set_blocked(True)
    print(f"split spin-free GEMM n={n} N={nocc} G={int(np.prod(ngrid))} {prec}", flush=True)
    e1, g1, r1 = _eval(m, p.z, True, "PHF_SPLIT_RJK")
    e0, g0, r0 = _eval(m, p.z, False, "PHF_SPLIT_RJK")
    e2, g2, _ = _eval(m, p.z, True, "PHF_SPLIT_RJK")
    m.ctx.close()
    assert e1 == e2 and np.array_equal(g1, g2)   # repeated split launches
    assert abs(e1 - e0) <= 1e-12 * max(1.0, abs(e0))
    assert np.abs(g1 - g0).max() <= 1e-12 * max(1.0, np.abs(g0).max())
    for name, a, b in zip(["Hmat", "Smat", "Gradmat", "Rdmmat"], r1, r0):
        assert _rel(a, b) <= 1e-12, name


@pytest.mark.parametrize("n,nocc,ngrid,prec", [(64, 32, (50, 16), "f16x2"), (24, 12, (14, 8), "fp64")])
def test_side_stream_pre_stages_equal_inline(n, nocc, ngrid, prec):
    """Hc, the T_g combine and the determinant weights on side streams
    (PHF_SIDE_PRE=1) against the default, where they run on the main stream:
    the same kernels on the same inputs, to 1e-12."""
    from phf_pcc_amd.model import ProjectedHF
    p = synthetic_problem(n, nocc, seed=990 + n, zscale=0.05)
    ga = grid_args(n, ngrid, 0)
    m = ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b), precision=prec)
    print(f"side-stream pre-stages n={n} N={nocc} G={int(np.prod(ngrid))} {prec}", flush=True)
    e1, g1, r1 = _eval(m, p.z, True, "PHF_SIDE_PRE")
    e0, g0, r0 = _eval(m, p.z, False, "PHF_SIDE_PRE")
    m.ctx.close()
    assert abs(e1 - e0) <= 1e-12 * max(1.0, abs(e0))
    assert np.abs(g1 - g0).max() <= 1e-12 * max(1.0, np.abs(g0).max())
    for name, a, b in zip(["Hmat", "Smat", "Gradmat", "Rdmmat"], r1, r0):
        assert _rel(a, b) <= 1e-12, name
