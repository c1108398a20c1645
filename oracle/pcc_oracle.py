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
  cc_itms          SUCCSD_Boson_cmplx/CCRes.f90:355-536   (CCItms)
  cc_res12         SUCCSD_Boson_cmplx/CCRes.f90:164-353   (CCRes12)
  proj_sz1/2       SUCCSD_Boson_cmplx/FPUCC_Tools.f90:171-233
  build_kern_sd    SUCCSD_Boson_cmplx/FPUCC_Tools.f90:291-436 (TrunODE 1 and 2)
  eval_kernel      SUCCSD_Boson_cmplx/FPUCC.f90:529-602   (every output)
  pbar_hbar_olap   SUCCSD_Boson_cmplx/FPUCC.f90:14-272    (every output)

Arrays are 0-based: t1[a, i] is T1(NOcc+1+a, 1+i), t2[a, b, i, j] is
T2(NOcc+1+a, NOcc+1+b, 1+i, 1+j); U and HTwo are NSO^4 with <pq||rs> at
[p, q, r, s].

Pinning (tests/test_pcc_oracle.py): with T2^3 |Phi> = 0 (at most five
electrons or five holes) the truncation of TransT2 at second order is exact,
so lambda0 = <Phi| R e^{T1+T2} |Phi> and eps0 = <Phi| H R e^{T1+T2} |Phi>
for any orbital rotation R, checked in brute-force Fock space; the grid sum
Olap0 / EOlap0 then equals the exact spin projection of the CC state.  The
singles and doubles outputs are pinned the same way: CCRes12 against
<Phi_i^a| e^{-T} H e^T |Phi> and <Phi_ij^ab| e^{-T} H e^T |Phi> for random
amplitudes, and lambda1/2, eps1/2 of EvalKernel against <Phi_i^a| R e^T |Phi>,
<Phi_ij^ab| R e^T |Phi> and the same with H (|Phi_i^a> = a+_a a_i |Phi>,
|Phi_ij^ab> = a+_a a_i a+_b a_j |Phi>); ProjSz against the S_z rule of the
reference's own commented-out form (FPUCC_Tools.f90:123-148).
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


# ---------------------------------------------------------------------------
# singles and doubles: CCRes12 / CCItms, the BuildKernSD fills, ProjSz
# ---------------------------------------------------------------------------

def _es(*args):
    """np.einsum with contraction-order optimisation (BLAS for the pair products)."""
    return np.einsum(*args, optimize=True)


def _aperm1(m):
    """APerm(M, 1) (CCRes.f90:120-156): M(I,J,:,:) - M(J,I,:,:)."""
    return m - m.transpose(1, 0, 2, 3)


def _aperm2(m):
    """APerm(M, 2): M(:,:,I,J) - M(:,:,J,I)."""
    return m - m.transpose(0, 1, 3, 2)


def cc_itms(fock, h2, t1, t2, nocc):
    """CCItms (CCRes.f90:355-536): tau, Fvv, Foo, Fov, Woooo, Wvvvv, Wovvo.
    t1 [a, i], t2 [a, b, i, j]; every intermediate in the reference's index
    order (Fov [i, a], Woooo [m, n, i, j], Wovvo [m, b, e, j])."""
    nso = fock.shape[0]
    o, v = slice(0, nocc), slice(nocc, nso)
    hoovv, hooov, hoovo = h2[o, o, v, v], h2[o, o, o, v], h2[o, o, v, o]
    hvovv, hovvv = h2[v, o, v, v], h2[o, v, v, v]
    itm1 = _es('ai,bj->abij', t1, t1)
    itm4 = t2 / 2 + itm1                                            # before the APerms
    itm1 = _aperm2(_aperm1(itm1))
    tau = t2 + itm1 / 2
    tb = t2 + itm1 / 4
    # Fvv: Transpose2(tb/2) then At_dot_B over the last index (:407-424)
    fvv = fock[v, v] - t1 @ fock[o, v] / 2
    fvv = fvv + _es('fm,ambf->ab', t1, hvovv)
    fvv = fvv - 0.5 * _es('ebmn,mnfb->ef', tb, hoovv)
    # Foo (:426-441)
    foo = fock[o, o] + fock[o, v] @ t1 / 2
    foo = foo + _es('em,imje->ij', t1, hooov)
    foo = foo + 0.5 * _es('ineb,ebjn->ij', hoovv, tb)
    # Fov (:443-452)
    fov = fock[o, v] + _es('em,imae->ia', t1, hoovv)
    # Woooo (:453-462)
    woooo = h2[o, o, o, o] + _es('mnef,efij->mnij', hoovv, tau / 4)
    woooo = woooo + _aperm2(_es('mnie,ej->mnij', hooov, t1))
    # Wvvvv (:463-485)
    wvvvv = h2[v, v, v, v] + _es('efmn,mncb->efcb', tau / 4, hoovv)
    wvvvv = wvvvv - _aperm1(_es('amcd,bm->abcd', hvovv, t1))
    # Wovvo (:487-519)
    wovvo = h2[o, v, v, o] + _es('mxcf,fj->mxcj', hovvv, t1)
    wovvo = wovvo - _es('mncj,bn->mbcj', hoovo, t1)
    wovvo = wovvo - _es('xbjn,incx->ibcj', itm4, hoovv)
    return tau, fvv, foo, fov, woooo, wvvvv, wovvo


def cc_res12(fock, h2, t1, t2, nocc):
    """CCRes12 (CCRes.f90:164-353): the CCSD residuals res1 [a, i] and
    res2 [a, b, i, j] of (t1, t2) with Fock and <pq||rs>.  r2tmp is kept in
    the reference's (i, j, a, b) order and transposed at the end."""
    nso = fock.shape[0]
    o, v = slice(0, nocc), slice(nocc, nso)
    tau, fvv, foo, fov, woooo, wvvvv, wovvo = cc_itms(fock, h2, t1, t2, nocc)
    t2tmp = t2.transpose(2, 3, 0, 1)                                 # (i, j, a, b)
    hvoov, hooov, hvovv = h2[v, o, o, v], h2[o, o, o, v], h2[v, o, v, v]
    # res1 (:215-236)
    r1 = fock[v, o] + fvv @ t1 - t1 @ foo
    r1 = r1 + _es('me,imae->ai', fov, t2tmp)
    r1 = r1 + _es('em,amie->ai', t1, hvoov)
    r1 = r1 - 0.5 * _es('mnab,mnib->ai', t2tmp, hooov)
    r1 = r1 + 0.5 * _es('ameb,imeb->ai', hvovv, t2tmp)
    # res2 (:238-352), r2tmp (i, j, a, b)
    r2tmp = np.zeros_like(t2tmp)
    itm1 = fvv - t1 @ fov / 2
    r2tmp = r2tmp + _aperm2(_es('ijae,be->ijab', t2tmp, itm1))
    itm3 = foo + fov @ t1 / 2
    r2tmp = r2tmp - _aperm1(_es('imab,mj->ijab', t2tmp, itm3))
    hvoov2 = h2[o, v, v, o].transpose(2, 3, 0, 1)                     # Transpose2: (e, j, m, b) = H(m, b, e, j)
    itm4 = _es('ei,ejmb->ijmb', t1, hvoov2)
    itm2 = _es('ijmb,am->ijab', itm4, -t1)
    itm2 = itm2 + _es('imae,mbej->ijab', t2tmp, wovvo)
    r2tmp = r2tmp + _aperm2(_aperm1(itm2))
    wt = wvvvv.transpose(2, 3, 0, 1)                                  # Transpose1: (e, f, a, b) = W(a, b, e, f)
    r2tmp = r2tmp + _es('efij,efab->ijab', tau / 2, wt)
    r2tmp = r2tmp + _es('mnij,mnab->ijab', woooo, tau.transpose(2, 3, 0, 1) / 2)
    hvovv2 = h2[v, v, v, o].transpose(2, 3, 0, 1)                     # (e, j, a, b) = H(a, b, e, j)
    hooov2 = h2[o, v, o, o].transpose(2, 3, 0, 1)                     # (i, j, m, b) = H(m, b, i, j)
    r2tmp = r2tmp + _aperm1(_es('ei,ejab->ijab', t1, hvovv2))
    r2tmp = r2tmp - _aperm2(_es('ijmb,am->ijab', hooov2, t1))
    res2 = r2tmp.transpose(2, 3, 0, 1) + h2[v, v, o, o]
    return r1, res2


def _spin_blocks(nocc, nso):
    """Occupied alpha / beta and virtual alpha / beta ranges in the order the
    ProjSz routines assume (NOA = NOcc/2 alpha occupied first; NVA = NAO - NOA)."""
    noa = nocc // 2
    nva = nso // 2 - noa
    return noa, nva


def proj_sz1(r1, nocc, nso):
    """ProjSz1 on the vo block (FPUCC_Tools.f90:171-180): r1 [a, i] with the
    spin-flip blocks (alpha virtual, beta occupied) and (beta virtual, alpha
    occupied) zeroed."""
    noa, nva = _spin_blocks(nocc, nso)
    out = r1.copy()
    out[:nva, noa:] = 0
    out[nva:, :noa] = 0
    return out


def proj_sz2(r2, nocc, nso):
    """ProjSz2VO (FPUCC_Tools.f90:217-233): keep the aaaa, bbbb, abab, abba,
    baba and baab blocks of r2 [a, b, i, j] (virtual spins a, b; occupied
    spins i, j), zero the rest."""
    noa, nva = _spin_blocks(nocc, nso)
    out = np.zeros_like(r2)
    A, B = slice(0, nva), slice(nva, None)
    a, b = slice(0, noa), slice(noa, None)
    for va, vb, oi, oj in ((A, A, a, a), (B, B, b, b), (A, B, a, b), (A, B, b, a), (B, A, b, a), (B, A, a, b)):
        out[va, vb, oi, oj] = r2[va, vb, oi, oj]
    return out


def build_kern_sd(fock, h2, escf, u1t, u2t, nocc, sp):
    """BuildKernSD (FPUCC_Tools.f90:291-436) for TrunODE 1 and 2, the blocks
    EvalKernel keeps: Ovl1[vo], Ovl2[vvoo], EOvl0, EOvl1[vo], EOvl2[vvoo]."""
    nso = fock.shape[0]
    ovl1 = u1t.copy()
    if sp == 2:
        ovl1 = proj_sz1(ovl1, nocc, nso)
    ovl2 = u2t + _es('ai,bj->abij', u1t, u1t) - _es('bi,aj->abij', u1t, u1t)
    if sp == 2:
        ovl2 = proj_sz2(ovl2, nocc, nso)
    eovl0 = cc_energy(u1t, u2t, fock, h2, nocc) + escf
    ex1, ex2 = cc_res12(fock, h2, u1t, u2t, nocc)
    eovl1 = ex1 + u1t * eovl0
    if sp == 2:
        eovl1 = proj_sz1(eovl1, nocc, nso)
    eovl2 = (ex2 + ovl2 * eovl0
             + _es('ai,bj->abij', u1t, ex1) - _es('bi,aj->abij', u1t, ex1)
             + _es('bj,ai->abij', u1t, ex1) - _es('aj,bi->abij', u1t, ex1))
    if sp == 2:
        eovl2 = proj_sz2(eovl2, nocc, nso)
    return ovl1, ovl2, eovl0, eovl1, eovl2


def eval_kernel(hone, htwo, t1, t2, enuc, rotmat, nocc, sp=0):
    """EvalKernel (FPUCC.f90:529-602), TrunODE 1: (lambda0, eps0, lambda1,
    lambda2, eps1, eps2) of one grid point; the per-point ProjSz1/2 of SP = 2
    (:583-588) repeat BuildKernSD's and change nothing."""
    nso = hone.shape[0]
    u0t, u1t, u2t = dstt2(rotmat, t1, t2, nocc)
    fock, escf = mkfock_mo(hone, htwo, nocc)
    escf = escf + enuc
    ovl1, ovl2, eovl0, eovl1, eovl2 = build_kern_sd(fock, htwo, escf, u1t, u2t, nocc, sp)
    if sp == 2:
        ovl1, eovl1 = proj_sz1(ovl1, nocc, nso), proj_sz1(eovl1, nocc, nso)
        ovl2, eovl2 = proj_sz2(ovl2, nocc, nso), proj_sz2(eovl2, nocc, nso)
    lam0 = u0t
    return lam0, eovl0 * lam0, ovl1 * lam0, ovl2 * lam0, eovl1 * lam0, eovl2 * lam0


def pbar_hbar_olap(hone, htwo, t1, t2, nocc, enuc, qj, sp, npoints, r1, r2, rpg, rk, roota, rootb, rooty,
                   weightsp, weightpg, fsp, fpg, fk):
    """Every output of PbarHbarOlap (FPUCC.f90:14-272), one rank: Olap0,
    OlapEx1 [a, i], OlapEx2 [a, b, i, j], EOlap0, EOlapEx1, EOlapEx2, with
    the final ProjSz of SP = 2 (:234-237)."""
    nso = hone.shape[0]
    nv = nso - nocc
    o0 = e0 = 0j
    o1, e1 = np.zeros((nv, nocc), complex), np.zeros((nv, nocc), complex)
    o2, e2 = np.zeros((nv, nv, nocc, nocc), complex), np.zeros((nv, nv, nocc, nocc), complex)
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
                    lam0, eps0, lam1, lam2, eps1, eps2 = eval_kernel(ho, ht, a1, a2, enuc, rot, nocc, sp)
                    w = point_weight(ia, iy, ipg, qj, sp, roota, rootb, rooty, weightsp, weightpg, fsp, fpg) * fac
                    o0 += w * lam0
                    e0 += w * eps0
                    o1 += w * lam1
                    e1 += w * eps1
                    o2 += w * lam2
                    e2 += w * eps2
    if sp == 2:
        o1, o2 = proj_sz1(o1, nocc, nso), proj_sz2(o2, nocc, nso)
        e1, e2 = proj_sz1(e1, nocc, nso), proj_sz2(e2, nocc, nso)
    return o0, o1, o2, e0, e1, e2
