"""The spin-free form of the ERI contraction (spin-restricted Ref: (2J - K)
applied to rho_aa + rho_bb and K applied to rho_aa - rho_bb, rho_ab, rho_ba)
against the spin blocks D, X1, X2 it replaces, through the device objective
and buildhsg.

- FP64: energies, gradients and every element of Hmat/Smat/Gradmat/Rdmmat
  equal the block path to 1e-12 relative to each output's largest entry (the
  two differ only in the order of the FP64 sums: 2J - K is formed once).
- Reduced forms: within 1e-6 Eh (energy) and 1e-5 (gradient) of the FP64
  block path, Smat/Rdmmat to 1e-10 relative (they do not depend on Gamma).
- The fused density pack + split writes the spin-free operand bit for bit
  as the two-pass path (density rows, rjk_pack_kernel, split_rows_kernel).
- Cases: n = 64 (every label of the 128-wide tile), n = 13 (padding columns:
  n^2 = 169 of 256; padding rows: 60 grid points), n = 24 and 32, CmplxConj = 2 (four densities per point,
  the two-pass operand), a point-group grid (npg = 2), the drop-in with a host
  HTwo; a UHF Hubbard reference is not spin-restricted and stays on the blocks.
"""
import os

import numpy as np
import pytest

from phf_pcc_amd.hamiltonian import synthetic_problem
from tests._cases import grid_args, hubbard_cfg1, synthetic_dense

pytestmark = pytest.mark.gpu


def _model(n, ngrid, prec, seed, cmplxconj=1, npg=1):
    from phf_pcc_amd.model import ProjectedHF
    p = synthetic_problem(n, n // 2, seed=seed, zscale=0.05)
    ga = grid_args(n, ngrid, 0)
    if cmplxconj == 2:
        ga = dict(ga, nci=2 * ga["nci"])
    if npg > 1:
        ga = _with_point_group(ga, 2 * n, seed)
    m = ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b), precision=prec, cmplxconj=cmplxconj)
    return p, m


def _with_point_group(ga, nso, seed):
    """The two-element group {I, -I} (it commutes with any Hamiltonian, so the
    projection stays Hermitian), weights 1/2: the npg = 2 code path."""
    ga = dict(ga)
    ga["rpg"] = np.stack([np.eye(nso, dtype=complex), -np.eye(nso, dtype=complex)])
    ga["weightpg"] = np.full((2, 1, 1), 0.5)
    return ga


def _eval(m, z, spinfree, fused=True):
    old = os.environ.get("PHF_FUSED_SPLIT")
    os.environ["PHF_FUSED_SPLIT"] = "1" if fused else "0"
    try:
        m.ctx.set_spinfree(spinfree)
        assert m.ctx.spinfree() == spinfree
        e, g = m.energy_and_gradient(z)
        res = m.buildhsg(z)
    finally:
        m.ctx.set_spinfree(True)
        if old is None:
            del os.environ["PHF_FUSED_SPLIT"]
        else:
            os.environ["PHF_FUSED_SPLIT"] = old
    return e, g, res


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("n,ngrid", [(64, (50, 16)), (13, (14, 8)), (24, (14, 8))])
def test_fp64_spinfree_equals_blocks(n, ngrid):
    p, m = _model(n, ngrid, "fp64", seed=700 + n)
    print(f"spin-free fp64 n={n} G={int(np.prod(ngrid))}", flush=True)
    e0, g0, r0 = _eval(m, p.z, False)
    e1, g1, r1 = _eval(m, p.z, True)
    print(f"  |dE| = {abs(e1 - e0):.2e}  max|dG| = {np.abs(g1 - g0).max():.2e}", flush=True)
    assert abs(e1 - e0) <= 1e-12 * max(1.0, abs(e0))
    assert np.abs(g1 - g0).max() <= 1e-12 * max(1.0, np.abs(g0).max())
    for name, a, b in zip(["Hmat", "Smat", "Gradmat", "Rdmmat"], r1, r0):
        assert _rel(a, b) <= 1e-12, name


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "f16x2", "i8x4"])
def test_reduced_spinfree_within_budget(prec):
    p, m = _model(32, (14, 8), "fp64", seed=733)
    e_ref, g_ref, r_ref = _eval(m, p.z, False)
    m.precision = prec
    e, g, r = _eval(m, p.z, True)
    print(f"spin-free {prec} n=32: |dE| = {abs(e - e_ref):.3e}  max|dG| = {np.abs(g - g_ref).max():.3e}", flush=True)
    assert abs(e - e_ref) <= 1e-6
    assert np.abs(g - g_ref).max() <= 1e-5
    assert _rel(r[1], r_ref[1]) <= 1e-10
    assert _rel(r[3], r_ref[3]) <= 1e-10
    assert _rel(r[0], r_ref[0]) <= 1e-6
    assert _rel(r[2], r_ref[2]) <= 1e-6


@pytest.mark.parametrize("n,ngrid", [(64, (50, 16)), (13, (6, 5)), (32, (14, 8))])
@pytest.mark.parametrize("prec", ["bf16x3", "f16x2", "i8x4"])
def test_fused_spinfree_operand_equals_two_pass(n, ngrid, prec):
    p, m = _model(n, ngrid, prec, seed=760 + n)
    print(f"spin-free fused split n={n} G={int(np.prod(ngrid))} {prec}", flush=True)
    e1, g1, r1 = _eval(m, p.z, True, fused=True)
    e0, g0, r0 = _eval(m, p.z, True, fused=False)
    assert e1 == e0
    assert np.array_equal(g1, g0)
    for a, b in zip(r1, r0):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["fp64", "f16x2"])
def test_cmplxconj2_spinfree(prec):
    p, m = _model(16, (14, 8), "fp64", seed=781, cmplxconj=2)
    print(f"spin-free CmplxConj=2 n=16 {prec}", flush=True)
    m.ctx.set_spinfree(False)
    r0 = m.buildhsg(p.z)
    m.ctx.set_spinfree(True)
    m.precision = prec
    assert m.ctx.spinfree()
    r1 = m.buildhsg(p.z)
    tol = 1e-12 if prec == "fp64" else 1e-6
    for name, a, b in zip(["Hmat", "Smat", "Gradmat", "Rdmmat"], r1, r0):
        t = tol if name in ("Hmat", "Gradmat") else (1e-12 if prec == "fp64" else 1e-10)
        assert _rel(a, b) <= t, name


def test_point_group_grid_spinfree():
    p, m = _model(16, (14, 8), "fp64", seed=790, npg=2)
    print("spin-free npg=2 n=16", flush=True)
    e0, g0, r0 = _eval(m, p.z, False)
    e1, g1, r1 = _eval(m, p.z, True)
    assert abs(e1 - e0) <= 1e-12 * max(1.0, abs(e0))
    for a, b in zip(r1, r0):
        assert _rel(a, b) <= 1e-12


def test_dropin_host_htwo_spinfree():
    """The drop-in with a host HTwo: the blocks and the spin-free form are built
    at the first call from the Ref it passes; FP64 results equal the blocks'."""
    from phf_pcc_amd.PHFTools_F import phftools as t
    case = synthetic_dense(12, 6, seed=795)
    ga = grid_args(12, (14, 8), 0)
    args = (case["hone"], case["htwo"], case["ref"], case["z"], ga["ngrid"], ga["nci"], 1, ga["ncisp"],
            case["nocc"], ga["r1"], ga["r2"], ga["rpg"], ga["weightsp"], ga["weightpg"], ga["roota"], ga["rootb"],
            ga["rooty"], ga["qj"], ga["sp"], 0)
    t.precision = "fp64"
    r1 = t.buildhsg(*args)
    assert t.context().spinfree()
    t.context().set_spinfree(False)
    try:
        r0 = t.buildhsg(*args)
    finally:
        t.context().set_spinfree(True)
    for a, b in zip(r1, r0):
        assert _rel(a, b) <= 1e-12


def test_uhf_reference_stays_on_blocks():
    from phf_pcc_amd.PHFTools_F import phftools as t
    case, nao = hubbard_cfg1()
    ga = grid_args(nao, (14, 8), 0)
    args = (case["hone"], case["htwo"], case["ref"], case["z"], ga["ngrid"], ga["nci"], 1, ga["ncisp"],
            case["nocc"], ga["r1"], ga["r2"], ga["rpg"], ga["weightsp"], ga["weightpg"], ga["roota"], ga["rootb"],
            ga["rooty"], ga["qj"], ga["sp"], 0)
    t.precision = "fp64"
    t.buildhsg(*args)
    assert t.context().blocked()
    assert not t.context().spinfree()
