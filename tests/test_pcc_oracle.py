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
