"""Reduced-precision ERI contraction (FP32, BF16x3, F16x2, I8x4) against the FP64
oracle.  The overlap inverse, determinant, Fock contraction and grid sums stay
FP64, so the error comes from the Gamma = A rho products only.

Tolerances (stated per result): projected energy |dE| <= 1e-6 Eh; gradient
block max |dG| <= 1e-5; Hmat/Smat/Gradmat/Rdmmat: Smat and Rdmmat do not depend
on Gamma and must agree to 1e-10 relative, Hmat and Gradmat to 1e-6 relative.
"""
import numpy as np
import pytest

from oracle import phf_oracle as O
from tests._cases import grid_args, hubbard_cfg1, hubbard_general_ref, synthetic_dense

pytestmark = pytest.mark.gpu

PRECS = ["fp32", "bf16x3", "f16x2", "i8x4"]


@pytest.fixture(scope="module")
def tools():
    from phf_pcc_amd.PHFTools_F import phftools as t
    yield t
    t.precision = "fp64"


def _gpu(tools, case, ga, prec):
    tools.precision = prec
    args = (case["hone"], case["htwo"], case["ref"], case["z"], ga["ngrid"], ga["nci"], 1, ga["ncisp"],
            case["nocc"], ga["r1"], ga["r2"], ga["rpg"], ga["weightsp"], ga["weightpg"], ga["roota"], ga["rootb"],
            ga["rooty"], ga["qj"], ga["sp"], 0)
    return tools.buildhsg(*args)


def _e_g(res, nocc, localg):
    h, s, gm, rm = res
    ev, vec = O.eigen_solver(h, s)
    fsp, fpg, fk = O.vec_decomp(vec[:, 0], h.shape[0], 1, 1)
    return ev[0].real, localg(gm, rm, s, ev[0], fsp, fpg, fk, 1)[nocc:, :nocc]


@pytest.fixture(scope="module")
def synth():
    case = synthetic_dense(32, 16, seed=1234)
    ga = grid_args(32, (14, 8), 0)
    ref = O.buildhsg(case["hone"], case["htwo"], case["ref"], case["z"], cmplxconj=1, nocc=16, **ga)
    return case, ga, ref


@pytest.mark.parametrize("prec", PRECS)
def test_mixed_energy_within_1e6(tools, synth, prec):
    case, ga, cpu = synth
    gpu = _gpu(tools, case, ga, prec)
    e_ref, g_ref = _e_g(cpu, 16, O.localg)
    e, g = _e_g(gpu, 16, tools.localg)
    print(f"{prec}: |dE| = {abs(e - e_ref):.3e}  max|dG| = {np.abs(g - g_ref).max():.3e}")
    assert abs(e - e_ref) <= 1e-6
    assert np.abs(g - g_ref).max() <= 1e-5
    for name, a, b, tol in zip(["Hmat", "Smat", "Gradmat", "Rdmmat"], gpu, cpu, [1e-6, 1e-10, 1e-6, 1e-10]):
        err = np.abs(a - b).max() / np.abs(b).max()
        assert err < tol, f"{name} rel err {err:.3e}"


@pytest.mark.parametrize("prec", ["f16x2", "bf16x3", "i8x4"])
def test_mixed_complex_integrals(tools, prec):
    case = hubbard_general_ref(6, 3.0, 7, zscale=0.2, complex_ref=True)
    ga = grid_args(6, (14, 8), 0)
    cpu = O.buildhsg(case["hone"], case["htwo"], case["ref"], case["z"], cmplxconj=1, nocc=6, **ga)
    gpu = _gpu(tools, case, ga, prec)
    e_ref, _ = _e_g(cpu, 6, O.localg)
    e, _ = _e_g(gpu, 6, tools.localg)
    assert abs(e - e_ref) <= 1e-6


def test_mixed_hubbard_ring(tools):
    """On-site U = 4 integrals (a few large entries per row)."""
    case, nao = hubbard_cfg1()
    ga = grid_args(nao, (14, 8), 0)
    cpu = O.buildhsg(case["hone"], case["htwo"], case["ref"], case["z"], cmplxconj=1, nocc=8, **ga)
    e_ref, _ = _e_g(cpu, 8, O.localg)
    for prec in PRECS:
        e, _ = _e_g(_gpu(tools, case, ga, prec), 8, tools.localg)
        assert abs(e - e_ref) <= 1e-6, prec


def test_precision_switch_keeps_fp64_exact(tools, synth):
    """Switching back to FP64 after the reduced forms reproduces the oracle."""
    case, ga, cpu = synth
    _gpu(tools, case, ga, "f16x2")
    gpu = _gpu(tools, case, ga, "fp64")
    for a, b in zip(gpu, cpu):
        assert np.abs(a - b).max() / np.abs(b).max() < 1e-10
