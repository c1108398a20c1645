"""Pinning the projected-CC oracle (oracle/pcc_oracle.py, the Olap0 / EOlap0
part of PbarHbarOlap, FPUCC.f90:14-272) in brute-force Fock space.

TransT2 (PCCCI.f90:98-245) expands e^{V T2 V^-1} to second order.  When
T2^3 |Phi> = 0 (at most five electrons or holes) that is exact, and then
  lambda0 = <Phi| R e^{T1+T2} |Phi>,   eps0 = <Phi| H R e^{T1+T2} |Phi>
for every orbital rotation R (EvalKernel, FPUCC.f90:529-602), which fixes
ProjT1, BuildThoulessZ, IntTran4Sim, MKFock_MO, the second-order terms,
CI2CC, ASymm2 and CCEnergy together: a dropped term, a wrong sign or a
transposed index moves lambda0 or eps0.  Summed over an exact spin grid,
Olap0 and EOlap0 are 2 pi <Phi| P |Psi> and 2 pi <Phi| H P |Psi>.
"""
import numpy as np
import pytest

from oracle import pcc_oracle as P
from oracle.fock_proj import FockSpace, cc_state, one_body, rotation_operator
from phf_pcc_amd.hamiltonian import random_unitary
from tests._cases import grid_args, hubbard_general_ref


def _amplitudes(nso, nocc, rng, s1, s2):
    nv = nso - nocc
    t1 = s1 * (rng.normal(size=(nv, nocc)) + 1j * rng.normal(size=(nv, nocc)))
    t2 = s2 * (rng.normal(size=(nv, nv, nocc, nocc)) + 1j * rng.normal(size=(nv, nv, nocc, nocc)))
    t2 = P.asymm2(t2) / 4
    return t1, t2


def _exact(fs, case, t1, t2, r):
    psi = cc_state(fs, t1, t2, case["nocc"])
    rpsi = rotation_operator(fs, r) @ psi
    phi = np.zeros(fs.dim, complex)
    phi[fs.index(np.array([(1 << case["nocc"]) - 1]))[0]] = 1.0
    return np.vdot(phi, rpsi), np.vdot(phi, fs.apply_h(case["hone"], case["htwo"], rpsi)), psi, phi


@pytest.mark.parametrize("nsite,seed,s1,s2", [(4, 1, 0.1, 0.1), (4, 2, 0.3, 0.2), (2, 3, 0.4, 0.4), (3, 4, 0.2, 0.3)])
def test_eval_kernel0_is_the_rotated_cc_overlap(nsite, seed, s1, s2):
    """lambda0 and eps0 of EvalKernel for random complex unitaries, a spin
    rotation of the grid and the identity, against Fock space (MO basis of
    a UHF Hubbard reference; 2, 3 and 4 electrons)."""
    case = hubbard_general_ref(nsite, 4.0, seed, zscale=0.1)
    nso, nocc = 2 * nsite, case["nocc"]
    rng = np.random.default_rng(seed)
    t1, t2 = _amplitudes(nso, nocc, rng, s1, s2)
    fs = FockSpace(nso, nocc)
    ga = grid_args(nsite, (6, 4), 0)
    spin_rot = case["ref"].conj().T @ (ga["r1"][3] @ ga["r2"][1]) @ case["ref"]
    for r in (random_unitary(nso, rng), random_unitary(nso, rng), spin_rot, np.eye(nso, dtype=complex)):
        lam, eps = P.eval_kernel0(case["hone"], case["htwo"], t1, t2, 0.0, r, nocc)
        lam_x, eps_x, _, _ = _exact(fs, case, t1, t2, r)
        assert abs(lam - lam_x) < 1e-11 * max(1.0, abs(lam_x))
        assert abs(eps - eps_x) < 1e-11 * max(1.0, abs(eps_x))


def test_second_order_truncation_is_what_breaks_at_six_electrons():
    """Negative control: at six electrons T2^3 |Phi> != 0 and TransT2's
    second-order expansion no longer equals the rotated CC overlap."""
    case = hubbard_general_ref(6, 4.0, 5, zscale=0.1)
    nso, nocc = 12, 6
    rng = np.random.default_rng(5)
    t1, t2 = _amplitudes(nso, nocc, rng, 0.05, 0.3)
    fs = FockSpace(nso, nocc)
    r = random_unitary(nso, rng)
    lam, _ = P.eval_kernel0(case["hone"], case["htwo"], t1, t2, 0.0, r, nocc)
    lam_x, _, _, _ = _exact(fs, case, t1, t2, r)
    assert abs(lam - lam_x) > 1e-6 * abs(lam_x)


def test_pbar_hbar_olap0_is_the_spin_projection():
    """Olap0 = 2 pi <Phi| P_S=0 |Psi> and EOlap0 = 2 pi <Phi| H P |Psi> on an
    exact grid (4 electrons, Lebedev 14 x 8 gamma), with Psi = e^{T1+T2} Phi,
    the rotations given in the MO basis as the PCC driver passes them
    (Main_PCC.py:37-42), ENuc added to every eps0."""
    case = hubbard_general_ref(4, 4.0, 7, zscale=0.1)
    nso, nocc = 8, 4
    rng = np.random.default_rng(7)
    t1, t2 = _amplitudes(nso, nocc, rng, 0.2, 0.2)
    ga = grid_args(4, (14, 8), 0)
    ref = case["ref"]
    r1 = np.einsum('pa,gpq,qb->gab', ref.conj(), ga["r1"], ref)
    r2 = np.einsum('pa,gpq,qb->gab', ref.conj(), ga["r2"], ref)
    rpg = np.eye(nso, dtype=complex)[None]
    enuc = 0.75
    o0, e0 = P.pbar_hbar_olap0(case["hone"], case["htwo"], t1, t2, nocc, enuc, 0, 1, ga["ngrid"], r1, r2, rpg, None,
                               ga["roota"], ga["rootb"], ga["rooty"], ga["weightsp"], np.ones((1, 1, 1)),
                               np.ones(1, complex), np.ones(1, complex), np.ones(1, complex))
    fs = FockSpace(nso, nocc)
    sz = np.diag(np.r_[0.5 * np.ones(4), -0.5 * np.ones(4)])
    sp = np.zeros((nso, nso))
    sp[np.arange(4), np.arange(4, 8)] = 1.0
    # spin operators in the MO basis of Ref
    szm, spm = ref.conj().T @ sz @ ref, ref.conj().T @ sp @ ref
    s_p, s_z = one_body(fs, spm), one_body(fs, szm)
    s2 = s_p.conj().T @ s_p + s_z @ s_z + s_z
    psi = cc_state(fs, t1, t2, nocc)
    ppsi = psi.copy()
    for s in (1.0, 2.0):                                   # Loewdin: remove S = 1, 2
        ppsi = (s2 @ ppsi - s * (s + 1) * ppsi) / (0.0 - s * (s + 1))
    phi = np.zeros(fs.dim, complex)
    phi[fs.index(np.array([(1 << nocc) - 1]))[0]] = 1.0
    ov = np.vdot(phi, ppsi)
    hv = np.vdot(phi, fs.apply_h(case["hone"], case["htwo"], ppsi))
    assert abs(o0 - 2 * np.pi * ov) < 1e-11 * abs(o0)
    assert abs(e0 - 2 * np.pi * (hv + enuc * ov)) < 1e-11 * abs(e0)


# ---------------------------------------------------------------------------
# singles and doubles (CCRes12, BuildKernSD, ProjSz)
# ---------------------------------------------------------------------------

def _excitation_vectors(fs, nocc):
    """|Phi_i^a> = a+_a a_i |Phi> as [a, i] and |Phi_ij^ab> = a+_a a_i a+_b a_j |Phi>
    as [a, b, i, j] (dense vectors, zero where a = b or i = j)."""
    nv = fs.nso - nocc
    phi = np.zeros(fs.dim, complex)
    phi[fs.index(np.array([(1 << nocc) - 1]))[0]] = 1.0
    e1 = np.zeros((nv, nocc, fs.dim), complex)
    for a in range(nv):
        for i in range(nocc):
            e1[a, i] = fs.epq(nocc + a, i) @ phi
    e2 = np.zeros((nv, nv, nocc, nocc, fs.dim), complex)
    for a in range(nv):
        for b in range(nv):
            for i in range(nocc):
                for j in range(nocc):
                    if a != b and i != j:
                        e2[a, b, i, j] = fs.epq(nocc + a, i) @ e1[b, j]
    return phi, e1, e2


def _t_operator(fs, t1, t2, nocc):
    nv = fs.nso - nocc
    t = None
    for a in range(nv):
        for i in range(nocc):
            term = fs.epq(nocc + a, i) * t1[a, i]
            t = term if t is None else t + term
    for a in range(nv):
        for b in range(nv):
            for i in range(nocc):
                for j in range(nocc):
                    if a != b and i != j:
                        t = t + (fs.epq(nocc + a, i) @ fs.epq(nocc + b, j)) * (0.25 * t2[a, b, i, j])
    return t


def _expm_apply(t, vec, sign, kmax):
    out, term = vec.copy(), vec.copy()
    for k in range(1, kmax + 1):
        term = sign * (t @ term) / k
        out = out + term
    return out


@pytest.mark.parametrize("nsite,seed,cplx", [(4, 11, False), (4, 12, True), (3, 13, True)])
def test_cc_res12_is_the_similarity_transformed_projection(nsite, seed, cplx):
    """CCRes12 = <Phi_i^a| e^{-T} H e^T |Phi> and <Phi_ij^ab| e^{-T} H e^T |Phi>
    for random complex amplitudes (Hubbard MO integrals, complex with a
    rotated reference): every intermediate of CCItms enters, so a dropped
    term, sign or index fails."""
    case = hubbard_general_ref(nsite, 4.0, seed, zscale=0.1, complex_ref=cplx)
    nso, nocc = 2 * nsite, case["nocc"]
    rng = np.random.default_rng(seed)
    t1, t2 = _amplitudes(nso, nocc, rng, 0.3, 0.3)
    fock, _ = P.mkfock_mo(case["hone"], case["htwo"], nocc)
    r1, r2 = P.cc_res12(fock, case["htwo"], t1, t2, nocc)
    fs = FockSpace(nso, nocc)
    phi, e1, e2 = _excitation_vectors(fs, nocc)
    t = _t_operator(fs, t1, t2, nocc)
    hpsi = fs.apply_h(case["hone"], case["htwo"], _expm_apply(t, phi, 1.0, nocc))
    sim = _expm_apply(t, hpsi, -1.0, nocc)
    x1 = np.einsum('aix,x->ai', e1.conj(), sim)
    x2 = np.einsum('abijx,x->abij', e2.conj(), sim)
    scale = max(1.0, np.abs(x2).max())
    assert np.abs(r1 - x1).max() < 1e-11 * scale
    mask = (~np.eye(nso - nocc, dtype=bool))[:, :, None, None] & (~np.eye(nocc, dtype=bool))[None, None]
    assert np.abs(np.where(mask, r2 - x2, 0)).max() < 1e-11 * scale
    # the diagonal a = b or i = j entries vanish by antisymmetry
    assert np.abs(np.where(mask, 0, r2)).max() < 1e-11 * scale


def _block_unitary(nso, nocc, rng):
    """A unitary that mixes occupied with occupied and virtual with virtual
    orbitals only: R e^T |Phi> stays a CC state with singles and doubles."""
    r = np.zeros((nso, nso), complex)
    r[:nocc, :nocc] = random_unitary(nocc, rng)
    r[nocc:, nocc:] = random_unitary(nso - nocc, rng)
    return r


@pytest.mark.parametrize("nsite,seed", [(4, 21), (4, 22), (3, 23)])
def test_eval_kernel_singles_doubles_are_rotated_projections(nsite, seed):
    """lambda0, lambda1, lambda2 and eps0 of EvalKernel are <Phi|, <Phi_i^a|,
    <Phi_ij^ab| of R e^T |Phi> and <Phi| (H + ENuc) R e^T |Phi> for any
    unitary (at most five electrons: TransT2 exact).  eps1 and eps2 take
    (H + ENuc) on the rotated state truncated at doubles, U0 e^{U1+U2} |Phi>
    (TrunODE = 1, 'SD; SD'), whose U1 and U2 follow from the exact lambda1 and
    lambda2: checked against that in Fock space for random unitaries, and
    against the untruncated projection where R e^T |Phi> has no triples (R
    block diagonal in occupied / virtual, or the identity)."""
    case = hubbard_general_ref(nsite, 4.0, seed, zscale=0.1)
    nso, nocc = 2 * nsite, case["nocc"]
    rng = np.random.default_rng(seed)
    t1, t2 = _amplitudes(nso, nocc, rng, 0.2, 0.2)
    fs = FockSpace(nso, nocc)
    phi, e1, e2 = _excitation_vectors(fs, nocc)
    psi = cc_state(fs, t1, t2, nocc)
    enuc = 0.37
    mask = (~np.eye(nso - nocc, dtype=bool))[:, :, None, None] & (~np.eye(nocc, dtype=bool))[None, None]
    hop = lambda v: fs.apply_h(case["hone"], case["htwo"], v) + enuc * v
    rots = [(random_unitary(nso, rng), False), (random_unitary(nso, rng), False),
            (_block_unitary(nso, nocc, rng), True), (np.eye(nso, dtype=complex), True)]
    for r, no_triples in rots:
        lam0, eps0, lam1, lam2, eps1, eps2 = P.eval_kernel(case["hone"], case["htwo"], t1, t2, enuc, r, nocc)
        rpsi = rotation_operator(fs, r) @ psi
        l0 = np.vdot(phi, rpsi)
        x1 = np.einsum('aix,x->ai', e1.conj(), rpsi)
        x2 = np.einsum('abijx,x->abij', e2.conj(), rpsi)
        assert abs(lam0 - l0) < 1e-11 * max(1.0, abs(l0))
        assert abs(eps0 - np.vdot(phi, hop(rpsi))) < 1e-11 * max(1.0, abs(eps0))
        assert np.abs(lam1 - x1).max() < 1e-11 * max(1.0, np.abs(x1).max())
        assert np.abs(np.where(mask, lam2 - x2, 0)).max() < 1e-11 * max(1.0, np.abs(x2).max())
        u1 = x1 / l0
        u2 = x2 / l0 - np.einsum('ai,bj->abij', u1, u1) + np.einsum('bi,aj->abij', u1, u1)
        trunc = l0 * cc_state(fs, u1, np.where(mask, u2, 0), nocc)
        for target, exact in ((trunc, True), (rpsi, no_triples)):
            hv = hop(target)
            y1 = np.einsum('aix,x->ai', e1.conj(), hv)
            y2 = np.einsum('abijx,x->abij', e2.conj(), hv)
            s = max(1.0, np.abs(y2).max(), np.abs(y1).max())
            ok = (np.abs(eps1 - y1).max() < 1e-11 * s and np.abs(np.where(mask, eps2 - y2, 0)).max() < 1e-11 * s)
            assert ok == exact, (no_triples, exact)


def test_proj_sz_is_the_sz_rule():
    """ProjSz1 / ProjSz2VO keep exactly the elements whose S_z is conserved
    (Sz(a) = Sz(i); Sz(a) + Sz(b) = Sz(i) + Sz(j)) for the orbital order
    alpha occupied, beta occupied, alpha virtual, beta virtual: the rule of
    the reference's own earlier form (FPUCC_Tools.f90:123-148, 192-215)."""
    rng = np.random.default_rng(3)
    for nso, nocc in ((8, 4), (12, 6), (10, 4), (8, 2)):
        nv = nso - nocc
        noa = nocc // 2
        nva = nso // 2 - noa
        so = np.r_[np.ones(noa), -np.ones(nocc - noa)]
        sv = np.r_[np.ones(nva), -np.ones(nv - nva)]
        r1 = rng.normal(size=(nv, nocc)) + 1j * rng.normal(size=(nv, nocc))
        r2 = rng.normal(size=(nv, nv, nocc, nocc)) + 1j * rng.normal(size=(nv, nv, nocc, nocc))
        keep1 = sv[:, None] == so[None, :]
        keep2 = (sv[:, None, None, None] + sv[None, :, None, None]) == (so[None, None, :, None] + so[None, None, None, :])
        assert np.array_equal(P.proj_sz1(r1, nocc, nso), np.where(keep1, r1, 0))
        assert np.array_equal(P.proj_sz2(r2, nocc, nso), np.where(keep2, r2, 0))


def test_pbar_hbar_olap_singles_doubles_are_the_spin_projection():
    """OlapEx1/2 = 2 pi <Phi_i^a| P |Psi>, <Phi_ij^ab| P |Psi> on an exact
    grid (4 electrons, Lebedev 14 x 8), and the zeroth-order outputs equal
    pbar_hbar_olap0's.  EOlapEx1/2 sum the doubles-truncated eps1/eps2 of
    every point (pinned per point above) with the same weights, so they are
    not the projection of H P |Psi>: the spin rotations give R e^T |Phi>
    triples that 'SD; SD' drops (checked: they differ)."""
    case = hubbard_general_ref(4, 4.0, 7, zscale=0.1)
    nso, nocc = 8, 4
    rng = np.random.default_rng(7)
    t1, t2 = _amplitudes(nso, nocc, rng, 0.2, 0.2)
    ga = grid_args(4, (14, 8), 0)
    ref = case["ref"]
    r1 = np.einsum('pa,gpq,qb->gab', ref.conj(), ga["r1"], ref)
    r2 = np.einsum('pa,gpq,qb->gab', ref.conj(), ga["r2"], ref)
    rpg = np.eye(nso, dtype=complex)[None]
    enuc = 0.75
    args = (case["hone"], case["htwo"], t1, t2, nocc, enuc, 0, 1, ga["ngrid"], r1, r2, rpg, None, ga["roota"],
            ga["rootb"], ga["rooty"], ga["weightsp"], np.ones((1, 1, 1)), np.ones(1, complex), np.ones(1, complex),
            np.ones(1, complex))
    o0, o1, o2, e0, e1, e2 = P.pbar_hbar_olap(*args)
    z0, ze0 = P.pbar_hbar_olap0(*args)
    assert abs(o0 - z0) < 1e-13 * abs(z0) and abs(e0 - ze0) < 1e-13 * abs(ze0)
    fs = FockSpace(nso, nocc)
    sz = np.diag(np.r_[0.5 * np.ones(4), -0.5 * np.ones(4)])
    spm = np.zeros((nso, nso))
    spm[np.arange(4), np.arange(4, 8)] = 1.0
    s_p, s_z = one_body(fs, ref.conj().T @ spm @ ref), one_body(fs, ref.conj().T @ sz @ ref)
    s2 = s_p.conj().T @ s_p + s_z @ s_z + s_z
    ppsi = cc_state(fs, t1, t2, nocc)
    for s in (1.0, 2.0):
        ppsi = (s2 @ ppsi - s * (s + 1) * ppsi) / (0.0 - s * (s + 1))
    hp = fs.apply_h(case["hone"], case["htwo"], ppsi) + enuc * ppsi
    _, ex1, ex2 = _excitation_vectors(fs, nocc)
    mask = (~np.eye(nso - nocc, dtype=bool))[:, :, None, None] & (~np.eye(nocc, dtype=bool))[None, None]
    x1 = 2 * np.pi * np.einsum('aix,x->ai', ex1.conj(), ppsi)
    x2 = 2 * np.pi * np.einsum('abijx,x->abij', ex2.conj(), ppsi)
    y1 = 2 * np.pi * np.einsum('aix,x->ai', ex1.conj(), hp)
    y2 = 2 * np.pi * np.einsum('abijx,x->abij', ex2.conj(), hp)
    s = max(np.abs(y1).max(), np.abs(y2).max())
    assert np.abs(o1 - x1).max() < 1e-11 * np.abs(x1).max()
    assert np.abs(np.where(mask, o2 - x2, 0)).max() < 1e-11 * np.abs(x2).max()
    assert np.abs(e1 - y1).max() > 1e-6 * s
