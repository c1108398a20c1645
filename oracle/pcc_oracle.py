"""ORACLE -- test infrastructure only (see phf_oracle.py header).

CPU restatement of the zeroth-order part of the projected-CC grid consumer,
statement by statement (paths relative to the reference repository):

  pbar_hbar_olap0  SUCCSD_Boson_cmplx/FPUCC.f90:14-272     (PbarHbarOlap: Olap0, EOlap0)
  eval_kernel0     SUCCSD_Boson_cmplx/FPUCC.f90:529-602    (EvalKernel: lambda0, eps0)
  dstt2            SUCCSD_Boson_cmplx/FPUCC_Tools.f90:697-722
  proj_t1          SUCCSD_Boson_cmplx/FPUCC_Tools.f90:676-695
  build_thouless_z SUCCSD_Boson_cmplx/FPUCC_Tools.f90:50-87
  build_w          SUCCSD_Boson_cmplx/PCCCI.f90:13-41
  trans_t2         SUCCSD_Boson_cmplx/PCCCI.f90:98-245
  ci2cc            SUCCSD_Boson_cmplx/PCCCI.f90:70-96
  asymm2           SUCCSD_Boson_cmplx/CCRes.f90:31-69
  cc_energy        SUCCSD_Boson_cmplx/CCRes.f90:9-29
  mkfock_mo        SUCCSD_Boson_cmplx/UHF.f90:461-497
  int_tran4_sim    SUCCSD_Boson_cmplx/IntTrans.f90:486-572
  wigner           SUCCSD_Boson_cmplx/Wrappers.f90:769-797 (the factorial sum of
                   phf_oracle.wignerfac)

Arrays are 0-based: t1[a, i] is T1(NOcc+1+a, 1+i), t2[a, b, i, j] is
T2(NOcc+1+a, NOcc+1+b, 1+i, 1+j); U and HTwo are NSO^4 with <pq||rs> at
[p, q, r, s].  Only what Olap0 and EOlap0 need is restated: the singles and
doubles kernels (CCRes12, the Ovl1/Ovl2 fills, ProjSz) are not.

Pinning (tests/test_pcc_oracle.py): with T2^3 |Phi> = 0 (at most five
electrons or five holes) the truncation of TransT2 at second order is exact,
so lambda0 = <Phi| R e^{T1+T2} |Phi> and eps0 = <Phi| H R e^{T1+T2} |Phi>
for any orbital rotation R, checked in brute-force Fock space; the grid sum
Olap0 / EOlap0 then equals the exact spin projection of the CC state.
"""
from __future__ import annotations

import numpy as np

from .phf_oracle import determinant_c, invert_c, wignerfac


def build_thouless_z(nocc, mo1, mo2):
    """|2> = e^Z |1> * Olap (FPUCC_Tools.f90:50-87): Tran = MO1^H MO2,
    Z = Tran[vo] Tran[oo]^-1, Olap = det Tran[oo]."""
    tran = mo1.conj().T @ mo2
    ell = tran[:nocc, :nocc]
    y = tran[nocc:, :nocc]
    zph = y @ invert_c(ell)
    return determinant_c(ell), zph


def proj_t1(r, t1, nocc):
    """ProjT1 (FPUCC_Tools.f90:676-695): W0, W1 of R e^{T1} |Phi> = W0 e^{W1} |Phi>."""
    nso = r.shape[0]
    mo2 = np.eye(nso, dtype=complex)
    mo2[nocc:, :nocc] = t1
    mo2 = r @ mo2
    return build_thouless_z(nocc, np.eye(nso, dtype=complex), mo2)


def int_tran4_sim(eri, rmat):
    """IntTran4Sim (IntTrans.f90:486-572): O' = R O R^-1 with bra indices
    (1, 2) transformed by R and ket indices (3, 4) by R^-1:
    O'(p,q,r,s) = sum R(p,a) R(q,b) O(a,b,c,d) Rinv(c,r) Rinv(d,s)."""
    rinv = invert_c(rmat)
    e2 = np.einsum('abpt,ts->abps', eri, rinv)          # index 4
    e1 = np.einsum('abqs,qr->abrs', e2, rinv)           # index 3
    e2 = np.einsum('abqp,rb->arqp', e1, rmat)           # index 2
    return np.einsum('ma,arqp->mrqp', rmat, e2)          # index 1


def mkfock_mo(oneh, eri, nocc):
    """MKFock_MO (UHF.f90:461-497): Fock = OneH + sum_i ERI(:,i,:,i),
    ESCF = sum_i OneH(i,i) + sum_ij ERI(i,j,i,j)/2 (Dirac, antisymmetrised)."""
    fock = oneh + np.einsum('piqi->pq', eri[:, :nocc, :, :nocc])
    escf = np.trace(oneh[:nocc, :nocc]) + 0.5 * np.einsum('ijij->', eri[:nocc, :nocc, :nocc, :nocc])
    return fock, complex(escf)


def asymm2(t2):
    """ASymm2 (CCRes.f90:31-69): X = T(a,b,i,j) - T(a,b,j,i); T = X(a,b,i,j) - X(b,a,i,j)."""
    x2 = t2 - t2.transpose(0, 1, 3, 2)
    return x2 - x2.transpose(1, 0, 2, 3)


def ci2cc(c0, c1, c2):
    """CI2CC (PCCCI.f90:70-96): T0 = C0, T1 = C1/C0, T2 = C2/C0 - T1(a,i)T1(b,j)/2
    (a != b, i != j), then ASymm2."""
    t1 = c1 / c0
    t2 = c2 / c0
    nv, no = t1.shape
    corr = 0.5 * np.einsum('ai,bj->abij', t1, t1)
    mask = (~np.eye(nv, dtype=bool))[:, :, None, None] & (~np.eye(no, dtype=bool))[None, None, :, :]
    t2 = t2 - np.where(mask, corr, 0.0)
    return c0, t1, asymm2(t2)


def trans_t2(r, t1, t2, nocc):
    """TransT2 (PCCCI.f90:98-245): the CI coefficients C0, C1, C2 of
    e^{V T2 V^-1} |Phi> to second order, V = e^{-T1} R."""
    nso = r.shape[0]
    o, v = slice(0, nocc), slice(nocc, nso)
    vmat = np.eye(nso, dtype=complex)
    vmat[v, o] = -t1
    vmat = vmat @ r
    u = np.zeros((nso,) * 4, dtype=complex)
    u[v, v, o, o] = t2
    u = int_tran4_sim(u, vmat)
    f, e0 = mkfock_mo(np.zeros((nso, nso), complex), u, nocc)
    # 1st order
    c0 = 1.0 + e0
    c1 = f[v, o].copy()
    c2 = u[v, v, o, o] / 4
    # 2nd order
    fov, fvo, foo, fvv = f[o, v], f[v, o], f[o, o], f[v, v]
    vvoo = u[v, v, o, o].copy()
    oovv = u[o, o, v, v].copy()
    c0 = c0 + e0 ** 2 / 2 + np.sum(fov * fvo.T) / 2
    c0 = c0 + np.einsum('ijab,abij->', oovv / 8, vvoo)                      # tA_dot_tB
    c1 = c1 + e0 * fvo
    c1 = c1 - fvo @ foo / 2 + fvv @ fvo / 2
    ooov = u[o, o, o, v]
    oovv = vvoo.transpose(2, 3, 0, 1).copy()                                # Transpose2(vvoo, oovv)
    c1 = c1 + np.einsum('ijab,ijkb->ak', -oovv / 4, ooov)                    # At_dot_B over b
    vvov = u[o, v, v, v].transpose(2, 3, 0, 1)                               # vvov(c,d,j,a) = U(j,a,c,d)
    c1 = c1 + np.einsum('cdja,cdij->ai', -vvov / 4, vvoo)                    # At_dot_B over j
    ovov = u[o, v, o, v]
    c1 = c1 + np.einsum('abij,jb->ai', vvoo, fov) / 2
    c1 = c1 - np.einsum('jaib,bj->ai', ovov, fvo) / 2
    c2 = c2 + e0 / 4 * vvoo
    c2 = c2 - np.einsum('aj,bi->abij', fvo, fvo) / 2
    c2 = c2 + np.einsum('jkac,kbic->abij', oovv, ovov) / 2
    c2tmp = np.zeros((nocc, nocc, nso - nocc, nso - nocc), dtype=complex)
    vvvv = u[v, v, v, v].transpose(2, 3, 0, 1)                               # Transpose1: vvvv(c,d,a,b) = U(a,b,c,d)
    c2tmp += np.einsum('cdij,cdab->ijab', vvoo / 16, vvvv)
    oooo = u[o, o, o, o]
    c2tmp += np.einsum('klij,klab->ijab', oooo / 16, oovv)
    c2tmp += np.einsum('ikab,kj->ijab', oovv, -foo / 4)
    c2tmp += np.einsum('ijcb,ac->ijab', oovv, fvv / 4)
    ooov2 = u[o, v, o, o].transpose(2, 3, 0, 1)                              # ooov(j,l,i,a) = U(i,a,j,l)
    c2tmp += np.einsum('ijkb,ak->ijab', ooov2, -fvo / 4)
    ovvv = u[v, v, o, v].transpose(2, 3, 0, 1)                               # ovvv(k,c,a,b) = U(a,b,k,c)
    c2tmp += np.einsum('icab,cj->ijab', ovvv, fvo) / 4
    c2 = c2 + c2tmp.transpose(2, 3, 0, 1)
    return complex(c0), c1, c2


def build_w(r, t1, t2, nocc):
    """BuildW (PCCCI.f90:13-41): TransT2, then CI2CC."""
    return ci2cc(*trans_t2(r, t1, t2, nocc))


def dstt2(r, t1, t2, nocc):
    """DSTT2 (FPUCC_Tools.f90:697-722): W0 = U0 * W0, W1 = U1 + W1 with
    (U0, U1) from ProjT1 and (W0, W1, W2) from BuildW over U1."""
    u0, u1 = proj_t1(r, t1, nocc)
    w0, w1, w2 = build_w(r, u1, t2, nocc)
    return w0 * u0, w1 + u1, w2


def cc_energy(t1, t2, fock, eri, nocc):
    """CCEnergy (CCRes.f90:9-29)."""
    o, v = slice(0, nocc), slice(nocc, fock.shape[0])
    e = np.sum(fock[o, v] * t1.T)
    tau = t2 + np.einsum('ai,bj->abij', t1, t1) - np.einsum('bi,aj->abij', t1, t1)
    e += 0.25 * np.einsum('ijab,abij->', eri[o, o, v, v], tau)
    return complex(e)


def eval_kernel0(hone, htwo, t1, t2, enuc, rotmat, nocc):
    """The lambda0 / eps0 part of EvalKernel (FPUCC.f90:529-602) with TrunODE = 1:
    lambda0 = U0t, eps0 = (CCEnergy(U1t, U2t) + ESCF + ENuc) * lambda0."""
    u0t, u1t, u2t = dstt2(rotmat, t1, t2, nocc)
    fock, escf = mkfock_mo(hone, htwo, nocc)
    escf = escf + enuc
    eovl0 = cc_energy(u1t, u2t, fock, htwo, nocc) + escf
    lam0 = u0t
    return lam0, eovl0 * lam0


def point_weight(il, iy, ipg, qj, sp, roota, rootb, rooty, weightsp, weightpg, fsp, fpg):
    """The sum over (M, N, P, Q) of WeightTot in PbarHbarOlap (FPUCC.f90:163-180):
    every term multiplies the same lambda0 / eps0."""
    w = 0j
    for m in range(-qj, qj + 1):
        for n in range(-qj, qj + 1):
            for p in range(weightpg.shape[1]):
                for q in range(weightpg.shape[2]):
                    if sp == 2:
                        wig = wignerfac(qj, m, n, roota[il], rootb[iy], rooty[il])
                    else:
                        wig = wignerfac(qj, m, n, roota[il], rootb[il], rooty[iy])
                    wt = weightsp[il, iy] * weightpg[ipg, p, q]
                    wt = wig * wt * np.conj(fsp[qj + m]) * fsp[qj + n]
                    w += wt * np.conj(fpg[p]) * fpg[q]
    return w


def pbar_hbar_olap0(hone, htwo, t1, t2, nocc, enuc, qj, sp, npoints, r1, r2, rpg, rk, roota, rootb, rooty,
                    weightsp, weightpg, fsp, fpg, fk):
    """Olap0 and EOlap0 of PbarHbarOlap (FPUCC.f90:14-272), one rank; ncik =
    len(fk): with ncik = 2 the second pass over the conjugated Hamiltonian
    and amplitudes with R Rk(2) and the extra factor fk(2) (:195-232)."""
    olap0 = 0j
    eolap0 = 0j
    npg = rpg.shape[0]
    passes = [(hone, htwo, t1, t2, None, 1.0)]
    if len(fk) == 2:
        passes.append((np.conj(hone), np.conj(htwo), np.conj(t1), np.conj(t2), rk[1], fk[1]))
    for ho, ht, a1, a2, rkk, fac in passes:
        for ia in range(int(npoints[0])):
            for iy in range(int(npoints[1])):
                for ipg in range(npg):
                    rot = r1[ia] @ r2[iy]
                    rot = rot @ rpg[ipg]
                    if rkk is not None:
                        rot = rot @ rkk
                    lam0, eps0 = eval_kernel0(ho, ht, a1, a2, enuc, rot, nocc)
                    w = point_weight(ia, iy, ipg, qj, sp, roota, rootb, rooty, weightsp, weightpg, fsp, fpg) * fac
                    eolap0 += w * eps0
                    olap0 += w * lam0
    return olap0, eolap0
