"""Device-resident objective: factored Hamiltonian built on the GPU, the
differentiable form (autograd Function over the adjoint kernel) and L-BFGS.

Checks: factored A equals the host-assembled dense A (through buildhsg parity
with the oracle, 1e-10 relative); gradient = central finite differences
(h = 1e-5, 8 random directions, |dfd - dE.d| <= 1e-6 (1 + |dfd|)) and = the
oracle's analytic LocalG gradient (1e-10); f16x2 energy within 1e-6 Eh.
"""
import numpy as np
import pytest

from oracle import phf_oracle as O
from phf_pcc_amd.hamiltonian import synthetic_problem
from tests._cases import grid_args

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prob():
    p = synthetic_problem(16, 16, seed=55, build_dense=True, zscale=0.05)
    ga = grid_args(16, (14, 8), 0)
    return p, ga


@pytest.fixture(scope="module")
def model(prob):
    from phf_pcc_amd.model import ProjectedHF
    p, ga = prob
    return ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b))


def test_factored_hamiltonian_matches_dense(prob, model):
    p, ga = prob
    assert np.abs(model.ctx.hone() - p.hone).max() < 1e-12
    gpu = model.buildhsg(p.z)
    cpu = O.buildhsg(p.hone, p.htwo, p.ref.astype(complex), p.z, cmplxconj=1, nocc=p.nocc, **ga)
    for a, b in zip(gpu, cpu):
        assert np.abs(a - b).max() / np.abs(b).max() < 1e-10


def test_gradient_matches_oracle_localg(prob, model):
    p, ga = prob
    e, g = model.energy_and_gradient(p.z)
    x = np.concatenate([p.z.real.ravel(), p.z.imag.ravel()])
    e_ref, g_ref = O.eandg(x, p.hone, p.htwo, p.ref.astype(complex), p.nocc, ga)
    assert abs(e - e_ref) < 1e-10
    gv = np.concatenate([g.real.ravel(), g.imag.ravel()])
    assert np.abs(gv - g_ref).max() < 1e-10


def test_autograd_matches_central_differences(prob, model):
    import torch
    from phf_pcc_amd.model import ProjectedEnergy
    p, _ = prob
    zr = torch.tensor(p.z.real, requires_grad=True)
    zi = torch.tensor(p.z.imag, requires_grad=True)
    e = ProjectedEnergy.apply(zr, zi, model)
    e.backward()
    grad = np.concatenate([zr.grad.numpy().ravel(), zi.grad.numpy().ravel()])
    x0 = np.concatenate([p.z.real.ravel(), p.z.imag.ravel()])
    nv = p.z.size
    f = lambda x: model.energy_and_gradient((x[:nv] + 1j * x[nv:]).reshape(p.z.shape))[0]
    rng = np.random.default_rng(7)
    h = 1e-5
    for _ in range(8):
        d = rng.normal(size=x0.shape)
        d /= np.linalg.norm(d)
        fd = (f(x0 + h * d) - f(x0 - h * d)) / (2 * h)
        assert abs(fd - grad @ d) <= 1e-6 * (1 + abs(fd)), (fd, grad @ d)


def test_lbfgs_lowers_the_energy(prob, model):
    from phf_pcc_amd.model import lbfgs_optimize
    p, _ = prob
    z, energies = lbfgs_optimize(model, p.z, steps=10)
    e_final = model.energy_and_gradient(z)[0]
    assert e_final < energies[0] - 1e-4
    assert np.isfinite(energies).all()


def test_mixed_precision_model(prob, model):
    from phf_pcc_amd.model import ProjectedHF
    p, ga = prob
    e64, g64 = model.energy_and_gradient(p.z)
    model.precision = "f16x2"
    try:
        e16, g16 = model.energy_and_gradient(p.z)
    finally:
        model.precision = "fp64"
    assert abs(e16 - e64) <= 1e-6
    assert np.abs(g16 - g64).max() <= 1e-5


@pytest.mark.parametrize("nsite,seed,j,ngrid,npg,cmplxconj", [(4, 101, 1, (26, 8), 2, 1), (6, 3, 2, (38, 8), 1, 1),
                                                               (4, 9, 0, (14, 8), 2, 0)])
def test_resident_gradient_projected_blocks(nsite, seed, j, ngrid, npg, cmplxconj):
    """Device-resident E+grad (grid sums kept on the device, vo block formed there)
    against the oracle's EandG with spin (J > 0, nci > 1) and point-group
    (two irreps of {E, C2}) coefficients, and with CmplxConj = 0."""
    from phf_pcc_amd.model import ProjectedHF
    from tests._cases import hubbard_general_ref
    case = hubbard_general_ref(nsite, 4.0, seed, zscale=0.3)
    nso = 2 * nsite
    if npg == 2:
        perm = np.eye(nsite)[np.roll(np.arange(nsite), nsite // 2)]
        rpg = np.stack([np.eye(nso), np.kron(np.eye(2), perm)]).astype(complex)
        wpg = np.stack([np.diag([0.5, 0.5]), np.diag([0.5, -0.5])])
        ga = grid_args(nsite, ngrid, j, rpg=rpg, weightpg=wpg)
    else:
        ga = grid_args(nsite, ngrid, j)
    z = case["z"].real.astype(complex) if cmplxconj == 0 else case["z"]
    m = ProjectedHF(case["nocc"], case["ref"], ga, hone=case["hone"], htwo=case["htwo"], cmplxconj=cmplxconj)
    e, g = m.energy_and_gradient(z)
    x = np.concatenate([z.real.ravel(), z.imag.ravel()])
    e_ref, g_ref = O.eandg(x, case["hone"], case["htwo"], case["ref"], case["nocc"], ga, cmplxconj=cmplxconj)
    assert abs(e - e_ref) < 1e-10
    if cmplxconj == 0:
        # EandG's mask (PHF.py:80-81) is applied by the model itself
        assert np.all(g.imag == 0)
    gv = np.concatenate([g.real.ravel(), g.imag.ravel()])
    assert np.abs(gv - g_ref).max() < 1e-10
    # the same model through the full Gradmat/Rdmmat drop-in path
    h, s, gm, rm = m.buildhsg(z)
    from phf_pcc_amd.phf import eigen_solver, vec_decomp
    ev, vec = eigen_solver(h, s)
    fsp, fpg, fk = vec_decomp(vec[:, 0], m.ncisp, m.ncipg, 1)
    g2 = m.localg(gm, rm, s, ev[0], fsp, fpg, fk)[case["nocc"]:, :case["nocc"]]
    if cmplxconj == 0:
        g2 = g2.real.astype(complex)
    assert np.abs(g2 - g).max() < 1e-11


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 1e-6), ("bf16x3", 1e-6), ("f16x2", 1e-6), ("i8x4", 1e-7)])
def test_spin_blocks_match_dense(prob, prec, tol):
    """Spin-pure Ref: A is stored as its three spin blocks; evaluating through
    them equals the dense contraction (FP64: summation order only)."""
    from phf_pcc_amd.model import ProjectedHF
    p, ga = prob
    m = ProjectedHF(p.nocc, p.ref, ga, factors=(p.h_spatial, p.b), precision=prec)
    assert m.ctx.blocked()
    e1, g1 = m.energy_and_gradient(p.z)
    m.ctx.set_blocked(False)
    assert not m.ctx.blocked()
    e0, g0 = m.energy_and_gradient(p.z)
    m.ctx.set_blocked(True)
    assert abs(e1 - e0) <= tol
    assert np.abs(g1 - g0).max() <= 10 * tol
    if prec == "fp64":
        x = np.concatenate([p.z.real.ravel(), p.z.imag.ravel()])
        e_ref, _ = O.eandg(x, p.hone, p.htwo, p.ref.astype(complex), p.nocc, ga)
        assert abs(e1 - e_ref) < 1e-10


def test_structure_detection():
    """Blocks for a spin-pure Ref through the dense drop-in (detected at
    phf_set_grid from Ref and the zero pattern of A); none for a Ref that
    mixes spins; PHF_HAM_BLOCKS_ONLY releases the dense A."""
    from phf_pcc_amd import _lib
    from phf_pcc_amd.model import ProjectedHF
    from tests._cases import hubbard_general_ref
    case = hubbard_general_ref(4, 4.0, 13, zscale=0.3)
    ga = grid_args(4, (14, 8), 0)
    m = ProjectedHF(case["nocc"], case["ref"], ga, hone=case["hone"], htwo=case["htwo"])
    assert m.ctx.blocked()
    e_b, _ = m.energy_and_gradient(case["z"])
    x = np.concatenate([case["z"].real.ravel(), case["z"].imag.ravel()])
    e_ref, _ = O.eandg(x, case["hone"], case["htwo"], case["ref"], case["nocc"], ga)
    assert abs(e_b - e_ref) < 1e-10
    mix = hubbard_general_ref(4, 4.0, 13, zscale=0.3, complex_ref=True)
    m2 = ProjectedHF(mix["nocc"], mix["ref"], ga, hone=mix["hone"], htwo=mix["htwo"])
    assert not m2.ctx.blocked()
    p = synthetic_problem(8, 8, seed=3)
    ga8 = grid_args(8, (14, 8), 0)
    ctx = _lib.Context(0)
    ctx.set_hamiltonian_factored(p.h_spatial, p.b, p.ref, flags=_lib.PHF_HAM_BLOCKS_ONLY)
    assert ctx.blocked()
    with pytest.raises(_lib.PhfError):
        ctx.set_blocked(False)
    m3 = ProjectedHF(p.nocc, p.ref, ga8, factors=(p.h_spatial, p.b))
    dense = ProjectedHF(p.nocc, p.ref, ga8, factors=(p.h_spatial, p.b), structure=False)
    assert m3.ctx.blocked() and not dense.ctx.blocked()
    assert abs(m3.energy_and_gradient(p.z)[0] - dense.energy_and_gradient(p.z)[0]) < 1e-10


def test_device_objective_equals_host_eigensolver_path(prob, model):
    """phf_eval_objective (nci = 1: E0 = Re H / |Re S| and LocalG on the
    device, one copy back) against phf_eval_resident + the host EigenSolver /
    VecDecomp (PHFTools.py:8-26) + phf_localg_vo on the same Z; a grid with
    nci > 1 is refused."""
    from phf_pcc_amd import _lib
    from phf_pcc_amd.phf import eigen_solver, vec_decomp
    p, ga = prob
    lib = _lib.load()
    h = model.ctx.handle
    e0 = np.zeros(1)
    g = np.zeros((p.nso - p.nocc, p.nocc), complex)
    z = np.ascontiguousarray(p.z, dtype=np.complex128)
    _lib.check(lib.phf_eval_objective(h, _lib.ptr(z), 1, p.nocc, 0, _lib.ptr(e0), _lib.ptr(g)), h)
    hm = np.zeros((1, 1), complex)
    sm = np.zeros((1, 1), complex)
    _lib.check(lib.phf_eval_resident(h, _lib.ptr(z), 1, p.nocc, 0, _lib.ptr(hm), _lib.ptr(sm)), h)
    ev, vec = eigen_solver(hm, sm)
    fsp, fpg, fk = vec_decomp(vec[:, 0], 1, 1, 1)
    c = lambda a: np.ascontiguousarray(a, dtype=np.complex128).reshape(-1)
    fsp, fpg, fk = c(fsp), c(fpg), c(fk)
    e0a = np.array([complex(ev[0])])
    g2 = np.zeros_like(g)
    _lib.check(lib.phf_localg_vo(h, _lib.ptr(sm), _lib.ptr(e0a), _lib.ptr(fsp), 1, _lib.ptr(fpg), 1, _lib.ptr(fk), 1,
                                 1, _lib.ptr(g2)), h)
    assert abs(e0[0] - ev[0].real) <= 1e-12 * abs(ev[0].real)
    assert np.abs(g - g2).max() <= 1e-12 * np.abs(g2).max()


def test_device_objective_refuses_nci_above_one():
    from phf_pcc_amd import _lib
    from phf_pcc_amd.model import ProjectedHF
    from tests._cases import hubbard_general_ref
    case = hubbard_general_ref(4, 4.0, 9, zscale=0.2)
    ga = grid_args(4, (26, 8), 1)
    m = ProjectedHF(case["nocc"], case["ref"], ga, hone=case["hone"], htwo=case["htwo"])
    z = np.ascontiguousarray(case["z"], dtype=np.complex128)
    e0 = np.zeros(1)
    g = np.zeros_like(z)
    rc = _lib.load().phf_eval_objective(m.ctx.handle, _lib.ptr(z), 1, case["nocc"], 0, _lib.ptr(e0), _lib.ptr(g))
    assert rc == _lib.PHF_ERR_UNSUPPORTED
    e, gg = m.energy_and_gradient(case["z"])     # the model takes the host EigenSolver path there
    assert np.isfinite(e) and np.isfinite(gg).all()


@pytest.mark.parametrize("cmplxconj,sp,is_rhf", [(0, 1, False), (1, 2, False), (1, 1, True)])
def test_autograd_with_objective_masks(cmplxconj, sp, is_rhf):
    """The Z masks of EandG (PHF.py:58-68,78-85) in the differentiable form:
    CmplxConj = 0 (E ignores Im Z, so dE/dIm Z = 0), SP = 2 (SzProj: no
    spin-flip blocks) and RHFProj.  Autograd equals central differences of
    the masked objective along random directions of the full [Re Z, Im Z],
    and equals the oracle's EandG with the same masks.  RHFProj copies the
    alpha block of Z into the beta block (PHF.py:139-143) and of G likewise,
    so there G is not the derivative of the masked objective (that would be
    G_aa + G_bb): only the equality with EandG is checked."""
    import torch
    from phf_pcc_amd.model import ProjectedEnergy, ProjectedHF
    from phf_pcc_amd.phf import ObjectiveConfig, make_eandg
    from tests._cases import hubbard_general_ref
    case = hubbard_general_ref(4, 4.0, 21, zscale=0.2)
    ga = grid_args(4, (14, 8) if sp == 1 else (1, 8), 0, sp=sp)
    m = ProjectedHF(case["nocc"], case["ref"], ga, hone=case["hone"], htwo=case["htwo"], cmplxconj=cmplxconj,
                    is_rhf=is_rhf)
    z = case["z"]
    zr = torch.tensor(z.real, requires_grad=True)
    zi = torch.tensor(z.imag, requires_grad=True)
    e = ProjectedEnergy.apply(zr, zi, m)
    e.backward()
    grad = np.concatenate([zr.grad.numpy().ravel(), zi.grad.numpy().ravel()])
    if cmplxconj == 0:
        assert np.all(zi.grad.numpy() == 0)
    cfg = ObjectiveConfig(nso=8, nocc=4, cmplxconj=cmplxconj, sp=sp, is_rhf=is_rhf, ncisp=ga["ncisp"])
    bh = lambda zz: O.buildhsg(case["hone"], case["htwo"], case["ref"], zz, cmplxconj=cmplxconj, nocc=4, **ga)
    f_ref = make_eandg(cfg, bh, lambda gm, rm, s, e0, fsp, fpg, fk, cc: O.localg(gm, rm, s, e0, fsp, fpg, fk, cc))
    x0 = np.concatenate([z.real.ravel(), z.imag.ravel()])
    e_ref, g_ref = f_ref(x0)
    assert abs(float(e) - e_ref) < 1e-10
    assert np.abs(grad - 2 * g_ref).max() < 1e-9
    if is_rhf:
        return
    nv = z.size
    f = lambda x: m.energy_and_gradient((x[:nv] + 1j * x[nv:]).reshape(z.shape))[0]
    rng = np.random.default_rng(8)
    h = 1e-5
    for _ in range(6):
        d = rng.normal(size=x0.shape)
        d /= np.linalg.norm(d)
        fd = (f(x0 + h * d) - f(x0 - h * d)) / (2 * h)
        assert abs(fd - grad @ d) <= 1e-6 * (1 + abs(fd)), (fd, grad @ d)
