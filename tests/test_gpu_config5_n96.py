"""The differentiable form at full size: synthetic n = 96 (NSO = 192, N = 48,
L = 288, Lebedev 50 x 16 gamma = 800 grid points, seed 1005).

Checks:
  * autograd (ProjectedEnergy over the adjoint kernel) against central
    differences, h = 1e-5, on 32 random unit directions of [Re Z, Im Z]:
    |fd - grad.d| <= 1e-6 (1 + |fd|);
  * the device E and G = dE/dZ* against the oracle's EandG (PHF.py:55-91)
    with MakeGamma from the factors (oracle/factored.py), at the full n on a
    6 x 4 sub-grid with real Z: 1e-10;
  * on the whole grid, the device-resident gradient equals the drop-in
    LocalG (PHFTools.f90:203-247) applied to the full Gradmat/Rdmmat: 1e-11;
  * F16X2 and I8X4 within 1e-6 Eh and 1e-5 of FP64;
  * five L-BFGS iterations lower the energy.
"""
import numpy as np
import pytest

from oracle import phf_oracle as O
from oracle.factored import gamma_factored
from phf_pcc_amd.hamiltonian import synthetic_problem
from phf_pcc_amd.spin_proj import grid_args

pytestmark = pytest.mark.gpu


def _equator_subgrid(n):
    ga = grid_args(n, (6, 4), 0)
    idx = np.nonzero(np.abs(ga["rootb"] - np.pi / 2) < 1e-12)[0]
    assert idx.size == 4
    sub = dict(ga)
    sub.update(ngrid=np.array([idx.size, 4]), r1=ga["r1"][idx], roota=ga["roota"][idx], rootb=ga["rootb"][idx],
               weightsp=ga["weightsp"][idx])
    return sub

N_SPATIAL, NOCC, NGRID, SEED = 96, 48, (50, 16), 1005


@pytest.fixture(scope="module")
def cfg5():
    from phf_pcc_amd.model import ProjectedHF
    prob = synthetic_problem(N_SPATIAL, NOCC, seed=SEED, ngrid=NGRID)
    ga = grid_args(N_SPATIAL, NGRID, 0)
    model = ProjectedHF(NOCC, prob.ref, ga, factors=(prob.h_spatial, prob.b))
    e64, g64 = model.energy_and_gradient(prob.z)
    yield dict(prob=prob, ga=ga, model=model, e64=e64, g64=g64)
    model.ctx.close()


def test_autograd_matches_central_differences_32_directions(cfg5):
    import torch
    from phf_pcc_amd.model import ProjectedEnergy
    m, p = cfg5["model"], cfg5["prob"]
    zr = torch.tensor(p.z.real.copy(), requires_grad=True)
    zi = torch.tensor(p.z.imag.copy(), requires_grad=True)
    e = ProjectedEnergy.apply(zr, zi, m)
    e.backward()
    assert abs(float(e.detach()) - cfg5["e64"]) < 1e-12
    grad = np.concatenate([zr.grad.numpy().ravel(), zi.grad.numpy().ravel()])
    x0 = np.concatenate([p.z.real.ravel(), p.z.imag.ravel()])
    nv = p.z.size
    f = lambda x: m.energy_and_gradient((x[:nv] + 1j * x[nv:]).reshape(p.z.shape))[0]
    rng = np.random.default_rng(1005)
    h = 1e-5
    worst = 0.0
    for _ in range(32):
        d = rng.normal(size=x0.shape)
        d /= np.linalg.norm(d)
        fd = (f(x0 + h * d) - f(x0 - h * d)) / (2 * h)
        err = abs(fd - grad @ d) / (1 + abs(fd))
        worst = max(worst, err)
        assert err <= 1e-6, (fd, grad @ d)
    print(f"[cfg5] max relative FD error over 32 directions {worst:.2e}", flush=True)


def test_gradient_matches_factored_oracle_on_subgrid(cfg5):
    m, p = cfg5["model"], cfg5["prob"]
    # a small grid is far from exact for N = 48 electrons: with complex Z, Im H
    # then exceeds the Hermiticity tolerance (PHFTools.f90:195-200) and the
    # evaluation stops, as the reference's would.  With real Z, H is real on
    # a grid closed under (alpha, gamma) -> (-alpha, -gamma): the four equator
    # points of the 6-point Lebedev rule (the poles sit at azimuth pi/2, whose
    # mirror is not in the rule) times the 4-point trapezoid in gamma.
    sub = _equator_subgrid(N_SPATIAL)
    z = p.z.real.astype(complex)
    m._set_grid(p.ref.astype(complex), sub)
    try:
        e, g = m.energy_and_gradient(z)
    finally:
        m._set_grid(p.ref.astype(complex), cfg5["ga"])
    x = np.concatenate([z.real.ravel(), z.imag.ravel()])
    e_ref, g_ref = O.eandg(x, p.hone, None, p.ref.astype(complex), NOCC, sub, gamma=gamma_factored(p.b, p.ref))
    assert abs(e - e_ref) < 1e-10
    assert np.abs(np.concatenate([g.real.ravel(), g.imag.ravel()]) - g_ref).max() < 1e-10


def test_resident_gradient_equals_dropin_localg_full_grid(cfg5):
    from phf_pcc_amd.phf import eigen_solver, vec_decomp
    m, p = cfg5["model"], cfg5["prob"]
    h, s, gm, rm = m.buildhsg(p.z)
    ev, vec = eigen_solver(h, s)
    fsp, fpg, fk = vec_decomp(vec[:, 0], m.ncisp, m.ncipg, 1)
    g = m.localg(gm, rm, s, ev[0], fsp, fpg, fk)[NOCC:, :NOCC]
    assert abs(ev[0] - cfg5["e64"]) < 1e-11
    assert np.abs(g - cfg5["g64"]).max() < 1e-11


@pytest.mark.parametrize("prec", ["f16x2", "i8x4"])
def test_reduced_adjoint_within_budget(cfg5, prec):
    m = cfg5["model"]
    m.precision = prec
    try:
        e, g = m.energy_and_gradient(cfg5["prob"].z)
    finally:
        m.precision = "fp64"
    print(f"[cfg5] {prec}: |dE| = {abs(e - cfg5['e64']):.3e}", flush=True)
    assert abs(e - cfg5["e64"]) <= 1e-6
    assert np.abs(g - cfg5["g64"]).max() <= 1e-5


def test_lbfgs_steps_lower_the_energy(cfg5):
    from phf_pcc_amd.model import lbfgs_optimize
    m, p = cfg5["model"], cfg5["prob"]
    z, energies = lbfgs_optimize(m, p.z, steps=5)
    assert m.energy_and_gradient(z)[0] < energies[0] - 1e-6
