"""Pins of the BuildSC oracle (oracle/buildsc_oracle.py) on the CPU.

* buildsc_vec (BuildSC_dr.f90:8-1244) against a brute-force Fock-space
  evaluation on every index: with Psi = det(R C0), C0 the first NOcc unit
  vectors, Y the transition density of RDM (PHFTools.f90:249-263) and
  tau(c) = c0 + sum c1(p,q) a+_p a_q + sum c2(p,q,r,s) a+_p a+_q a_s a_r,
      t0         = <Phi| tau(c) |Psi> / <Phi|Psi>
      t1(a,b)    = <Phi| a+_b a_a tau(c) |Psi> / <Phi|Psi>
      t2(a,b,c,d)= <Phi| a+_c a+_d a_b a_a tau(c) |Psi> / <Phi|Psi>.
  The adjoint with the other order of a_c a_d (a plausible transposition)
  fails, so the check is not vacuous.
* buildsc at one identity grid point returns its packed input (the packed SD
  basis of LinDep.py:156-187 is orthonormal).
* buildsc over an exact spin grid equals the Fock-space projector
  <tau_mu Phi| P_S tau(c) |Phi> / <Phi| P_S |Phi> on every packed component.
* mat2vec(vec2mat(x)) == x and the packed dimension (LinDep.py:7).
"""
import itertools

import numpy as np
import pytest

from oracle import buildsc_oracle as B
from oracle.fock_proj import FockSpace
from oracle.phf_oracle import rdm
from tests._cases import grid_args, random_unitary


def _cr(rng, *s):
    return rng.normal(size=s) + 1j * rng.normal(size=s)


def _two(fs, p, q, r, s):
    """a+_p a+_q a_s a_r = E_pr E_qs - delta_qr E_ps."""
    m = fs.epq(p, r) @ fs.epq(q, s)
    if q == r:
        m = m - fs.epq(p, s)
    return m


def _tau_apply(fs, c0, c1, c2, vec):
    n = c1.shape[0]
    out = c0 * vec
    for p, q in itertools.product(range(n), repeat=2):
        if c1[p, q] != 0:
            out = out + c1[p, q] * (fs.epq(p, q) @ vec)
    for p, q, r, s in itertools.product(range(n), repeat=4):
        if c2[p, q, r, s] != 0:
            out = out + c2[p, q, r, s] * (_two(fs, p, q, r, s) @ vec)
    return out


@pytest.mark.parametrize("nso,nocc,seed", [(5, 2, 1), (6, 3, 2)])
def test_buildsc_vec_matches_fock_space(nso, nocc, seed):
    rng = np.random.default_rng(seed)
    r = 0.3 * _cr(rng, nso, nso) + np.eye(nso)
    c0mo = np.eye(nso)[:, :nocc].astype(complex)
    y, _ = rdm(r, c0mo)
    c0, c1, c2 = _cr(rng), _cr(rng, nso, nso), _cr(rng, nso, nso, nso, nso)
    t0, t1, t2 = B.buildsc_vec(c0, c1, c2, y)
    fs = FockSpace(nso, nocc)
    psi, phi = fs.slater(r[:, :nocc]), fs.slater(c0mo)
    norm = phi.conj() @ psi
    tv = _tau_apply(fs, c0, c1, c2, psi)
    assert abs(phi.conj() @ tv / norm - t0) < 1e-11 * abs(t0)
    f1 = np.array([[phi.conj() @ (fs.epq(b, a) @ tv) for b in range(nso)] for a in range(nso)]) / norm
    assert np.abs(f1 - t1).max() < 1e-11 * np.abs(t1).max()
    f2 = np.zeros_like(t2)
    f2_swapped = np.zeros_like(t2)
    for a, b, c, d in itertools.product(range(nso), repeat=4):
        f2[a, b, c, d] = phi.conj() @ (_two(fs, c, d, a, b) @ tv)
        f2_swapped[a, b, c, d] = phi.conj() @ (_two(fs, d, c, a, b) @ tv)
    scale = np.abs(t2).max()
    assert np.abs(f2 / norm - t2).max() < 1e-11 * scale
    assert np.abs(f2_swapped / norm - t2).max() > 1e-3 * scale


def test_packing_round_trip():
    nso, nocc = 9, 4
    dim = B.sd_dim(nso, nocc)
    assert dim == 1 + 4 * 5 + 6 * 10
    x = _cr(np.random.default_rng(0), dim)
    v0, v1, v2 = B.vec2mat(x, nso, nocc)
    assert np.count_nonzero(v1) == 20 and np.count_nonzero(v2) == 60
    assert np.array_equal(B.mat2vec(v0, v1, v2, nocc), x)


def _single_point(nso):
    eye = np.eye(nso, dtype=complex)[None]
    return dict(ngrid=np.array([1, 1]), qj=0, r1=eye, r2=eye, rpg=eye, weightsp=np.ones((1, 1)),
                weightpg=np.ones((1, 1, 1)), fsp=np.ones(1, complex), fpg=np.ones(1, complex))


def test_identity_point_returns_input():
    nso, nocc = 8, 3
    x = _cr(np.random.default_rng(4), B.sd_dim(nso, nocc))
    t = B.buildsc(*B.vec2mat(x, nso, nocc), nocc, **_single_point(nso))
    assert np.abs(B.mat2vec(*t, nocc) - x).max() < 1e-13


def _excited(fs, u, nocc, repl):
    """b+_a b_i (and b+_b b_j) on Phi = det(u[:, :nocc]) for MO operators b:
    the determinant with column i replaced by u_a (j by u_b); for i < j and
    a < b, b+_a b+_b b_j b_i = b+_a b_i b+_b b_j, so no sign arises."""
    cols = u[:, :nocc].copy()
    for i, a in repl:
        cols[:, i] = u[:, a]
    return fs.slater(cols)


def _packed_states(fs, u, nso, nocc):
    """tau_mu Phi for every packed component, in vec2mat's order."""
    out = [fs.slater(u[:, :nocc])]
    for i in range(nocc):
        for a in range(nocc, nso):
            out.append(_excited(fs, u, nocc, [(i, a)]))
    for j in range(1, nocc):
        for i in range(j):
            for b in range(nocc + 1, nso):
                for a in range(nocc, b):
                    out.append(_excited(fs, u, nocc, [(i, a), (j, b)]))
    return np.array(out)


@pytest.mark.parametrize("seed", [5, 6])
def test_buildsc_matches_fock_space_projector(seed):
    """J = 0 projection over an exact grid (Lebedev 14 x 8 gamma points for
    four electrons) against the Loewdin singlet projector in Fock space."""
    nao, nocc = 4, 4
    nso = 2 * nao
    rng = np.random.default_rng(seed)
    u = random_unitary(nso, rng)                  # MO basis: Phi = first nocc columns
    ga = grid_args(nao, (14, 8), 0)
    r1 = np.einsum('pa,gpq,qb->gab', u.conj(), ga["r1"], u)
    r2 = np.einsum('pa,gpq,qb->gab', u.conj(), ga["r2"], u)
    x = _cr(rng, B.sd_dim(nso, nocc))
    t = B.buildsc(*B.vec2mat(x, nso, nocc), nocc, ga["ngrid"], 0, r1, r2, ga["rpg"], ga["weightsp"],
                  ga["weightpg"], np.ones(1, complex), np.ones(1, complex))
    got = B.mat2vec(*t, nocc)

    fs = FockSpace(nso, nocc)
    states = _packed_states(fs, u, nso, nocc)     # rows: tau_mu Phi
    phi = states[0]
    ptau = fs.spin_projector_apply(0.0, x @ states)
    want = states.conj() @ ptau / (phi.conj() @ fs.spin_projector_apply(0.0, phi))
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()
    # the same check with a grid too small for four electrons fails
    gc = grid_args(nao, (1, 8), 0)
    r1c = np.einsum('pa,gpq,qb->gab', u.conj(), gc["r1"], u)
    r2c = np.einsum('pa,gpq,qb->gab', u.conj(), gc["r2"], u)
    tc = B.buildsc(*B.vec2mat(x, nso, nocc), nocc, gc["ngrid"], 0, r1c, r2c, gc["rpg"], gc["weightsp"],
                   gc["weightpg"], np.ones(1, complex), np.ones(1, complex))
    assert np.abs(B.mat2vec(*tc, nocc) - want).max() > 1e-6 * np.abs(want).max()
