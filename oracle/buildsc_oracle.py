"""ORACLE -- test infrastructure only (see phf_oracle.py header).

CPU restatement of the reference's projected SD-space overlap times vector
(the LinDep matrix-vector product), used solely as the checker for the HIP
path.  Each function follows the reference text it cites:

  buildsc_vec   PHF_fort/BuildSC_dr.f90:8-1244  (BuildSC_Vec, generated
                contraction sequence; restated statement by statement, the
                intermediates keep the generated names tau0..tau27)
  ao2mo4        PHF_fort/PHFTools.f90:400-462  (four one-index passes)
  buildsc       PHF_fort/PHFTools.f90:464-537  (BuildSC: grid loop, RDM with
                newMO = the first NOcc unit vectors, weights, 1/ovlp)
  vec2mat       LinDep.py:156-171  (packed SD vector -> c0, c1, c2)
  mat2vec       LinDep.py:173-187  (c0, c1, c2 -> packed SD vector)

Reference behaviour kept on purpose: the Wigner factor is overwritten by one
(PHFTools.f90:515), the weight ``w`` is a REAL variable so every complex
product assigned to it keeps only its real part (:484, :518-520), Rk/fk/ncik
are accepted and never used (the ik loop of LinDep.py:119-153 is not in the
Fortran), and the overlap sums s*w over the same points (:524).

Pinning (tests/test_buildsc_oracle.py): ``buildsc_vec`` against a brute-force
Fock-space evaluation of <Phi| O_out tau(c) |Psi_B> / <Phi|Psi_B> with the
transition density of Psi_B, on every index of t1 and t2; ``buildsc`` at a
single identity grid point returns the input vector (the SD basis is
orthonormal); ``buildsc`` over an exact spin grid equals the Fock-space
projector <Phi| tau_mu^+ P_S tau(c) |Phi> / <Phi|P_S|Phi>.
"""
from __future__ import annotations

import numpy as np

from oracle.phf_oracle import rdm


def buildsc_vec(c0, c1, c2, y):
    """BuildSC_Vec (BuildSC_dr.f90:8-1244): (t0, t1, t2) from (c0, c1, c2) and Y.

    Arrays are indexed as the Fortran ones (t2[a, b, c, d] is t2(a,b,c,d)).
    """
    c1 = np.asarray(c1, dtype=complex)
    c2 = np.asarray(c2, dtype=complex)
    Y = np.asarray(y, dtype=complex)
    ein = lambda s, *a: np.einsum(s, *a, optimize=True)
    tau0 = c2 - c2.transpose(1, 0, 2, 3)
    tau1 = ein('cd,dabc->ab', Y, tau0)
    tau2 = -tau1
    tau6 = -tau1
    tau8 = -ein('ab,ba->', Y, tau1)
    t2 = tau8 * (ein('ac,bd->abcd', Y, Y) - ein('ad,bc->abcd', Y, Y))
    tau13 = -tau1
    tau5 = ein('cd,adcb->ab', Y, tau0)
    tau6 = tau6 - tau5
    tau13 = tau13 - tau5
    tau14 = ein('ac,cb->ab', Y, tau13)
    tau15 = ein('ca,bc->ab', Y, tau14)
    tau16 = -ein('ac,db->abcd', Y, tau15)
    tau21 = ein('ca,bc->ab', Y, tau13)
    tau22 = ein('ab,cd->abcd', Y, tau21)
    tau2 = tau2 + c1
    tau3 = ein('ab,ba->', Y, tau2)
    t0 = tau3
    t1 = tau3 * Y
    tau4 = -c2 + c2.transpose(1, 0, 2, 3)
    tau7 = -ein('cd,dabc->ab', Y, tau4) - ein('cd,adcb->ab', Y, tau4)
    tau6 = tau6 + c1
    tau7 = tau7 + Y @ tau6 - c1
    t1 = t1 - ein('cb,ac->ab', Y, tau7)
    tau9 = ein('ea,bcde->abcd', Y, c2)
    tau10 = ein('ea,bcde->abcd', Y, tau9)
    tau11 = ein('ae,bcde->abcd', Y, tau10)
    tau12 = ein('ae,bcde->abcd', Y, tau11)
    tau16 = tau16 + tau12
    t2 = t2 + tau16 - tau16.transpose(0, 1, 3, 2) - tau16.transpose(1, 0, 2, 3) + tau16.transpose(1, 0, 3, 2)
    tau19 = -tau10 + tau10.transpose(0, 1, 3, 2)
    tau20 = ein('ae,bcde->abcd', Y, tau19)
    tau22 = tau22 - tau20
    # t2(a,b,c,d) += tau22(a,c,d,b) - tau22(a,d,c,b) - tau22(b,c,d,a) + tau22(b,d,c,a)
    t2 = t2 + (ein('acdb->abcd', tau22) - ein('adcb->abcd', tau22)
               - ein('bcda->abcd', tau22) + ein('bdca->abcd', tau22))
    # t2(a,b,c,d) += tau10(c,d,a,b) - tau10(c,d,b,a) - tau10(d,c,a,b) + tau10(d,c,b,a)
    t2 = t2 + (ein('cdab->abcd', tau10) - ein('cdba->abcd', tau10)
               - ein('dcab->abcd', tau10) + ein('dcba->abcd', tau10))
    tau17 = ein('ca,bc->ab', Y, c1)
    tau18 = ein('ac,bc->ab', Y, tau17)
    t2 = t2 - ein('ac,bd->abcd', Y, tau18) + ein('bc,ad->abcd', Y, tau18)
    t2 = t2 + ein('ac,db->abcd', Y, tau17) - ein('bc,da->abcd', Y, tau17)
    tau23 = ein('ba,ab->', Y, c1)
    tau27 = tau23 * Y.T
    tau24 = Y @ c1
    tau25 = tau24 - c1
    tau26 = ein('ca,bc->ab', Y, tau25)
    tau27 = tau27 - tau26 + c0 * Y.T
    t2 = t2 + ein('bd,ca->abcd', Y, tau27) - ein('ad,cb->abcd', Y, tau27)
    t0 = t0 + c0
    t1 = t1 + c0 * Y
    return t0, t1, t2


def ao2mo4(h, m):
    """ao2mo4 (PHFTools.f90:400-462): newH(i,j,k,l) =
    sum conj(M(p,i)) conj(M(q,j)) M(r,k) M(s,l) H(p,q,r,s), as its four passes."""
    h = np.asarray(h, dtype=complex)
    tmp = np.einsum('pqrs,sl->pqrl', h, m)            # :417 Dot_Product(Conjg(M(:,l)), H(p,q,r,:))
    new = np.einsum('pqrl,rk->pqkl', tmp, m)          # :429
    tmp = np.einsum('pqkl,qj->pjkl', new, m.conj())   # :441 Dot_Product(M(:,j), newH(p,:,k,l))
    return np.einsum('pjkl,pi->ijkl', tmp, m.conj())  # :453


def sc_weights(weightpg, fsp, fpg, qj):
    """The per-(ipg) factor sum_{m,n,p,q} w of PHFTools.f90:510-525 with the
    REAL ``w`` (each complex product keeps its real part on assignment)."""
    npg, ncipg, _ = weightpg.shape
    out = np.zeros(npg)
    for ipg in range(npg):
        acc = 0.0
        for m in range(-qj, qj + 1):
            for n in range(-qj, qj + 1):
                for p in range(ncipg):
                    for q in range(ncipg):
                        w = float(weightpg[ipg, p, q])                                   # * wig, wig = One
                        w = (w * np.conj(fsp[m + qj]) * fsp[n + qj]).real
                        w = (w * np.conj(fpg[p]) * fpg[q]).real
                        acc += w
        out[ipg] = acc
    return out


def buildsc(c0, c1, c2, nocc, ngrid, qj, r1, r2, rpg, weightsp, weightpg, fsp, fpg, points=None):
    """BuildSC (PHFTools.f90:464-537) on one rank: returns (t0, t1, t2).

    ``points`` optionally restricts the grid to (il, iy, ipg) triples (bounded
    CPU samples); the overlap then sums over the same points.
    """
    c1 = np.asarray(c1, dtype=complex)
    c2 = np.asarray(c2, dtype=complex)
    nso = c1.shape[0]
    npg = rpg.shape[0]
    newmo = np.zeros((nso, nocc), dtype=complex)
    newmo[np.arange(nocc), np.arange(nocc)] = 1.0
    wpg = sc_weights(np.asarray(weightpg, dtype=float), np.asarray(fsp, dtype=complex),
                     np.asarray(fpg, dtype=complex), qj)
    t0 = 0j
    t1 = np.zeros((nso, nso), dtype=complex)
    t2 = np.zeros((nso,) * 4, dtype=complex)
    ovlp = 0j
    if points is None:
        points = [(il, iy, ipg) for il in range(int(ngrid[0]))
                  for iy in range(int(ngrid[1])) for ipg in range(npg)]
    for il, iy, ipg in points:
        r = r1[il] @ r2[iy]
        r = r @ rpg[ipg]
        rho, s = rdm(r, newmo)
        rh = r.conj().T
        c1t = rh.conj().T @ c1 @ rh                         # ao2mo2(c1tmp, c1, R^H)
        c2t = ao2mo4(c2, rh)
        a0, a1, a2 = buildsc_vec(c0, c1t, c2t, rho)
        s = weightsp[il, iy] * s
        w = wpg[ipg]
        t0 += a0 * s * w
        t1 += a1 * s * w
        t2 += a2 * s * w
        ovlp += s * w
    return t0 / ovlp, t1 / ovlp, t2 / ovlp


# ----------------------------------------------------------------------------
# packed SD vectors (LinDep.py:156-187)
# ----------------------------------------------------------------------------

def sd_dim(nso, nocc):
    """1 + O V + O(O-1)/2 V(V-1)/2 (LinDep.py:7)."""
    nv = nso - nocc
    return 1 + nocc * nv + nocc * (nocc - 1) // 2 * (nv * (nv - 1) // 2)


def vec2mat(vec, nso, nocc):
    """Vec2Mat (LinDep.py:156-171)."""
    v0 = vec[0]
    v1 = np.zeros((nso, nso), dtype=complex)
    v2 = np.zeros((nso,) * 4, dtype=complex)
    k = 0
    for i in range(nocc):
        for a in range(nocc, nso):
            k += 1
            v1[a, i] = vec[k]
    for j in range(1, nocc):
        for i in range(j):
            for b in range(nocc + 1, nso):
                for a in range(nocc, b):
                    k += 1
                    v2[a, b, i, j] = vec[k]
    return v0, v1, v2


def mat2vec(v0, v1, v2, nocc):
    """Mat2Vec (LinDep.py:173-187)."""
    nso = v1.shape[0]
    out = np.zeros(sd_dim(nso, nocc), dtype=complex)
    out[0] = v0
    k = 0
    for i in range(nocc):
        for a in range(nocc, nso):
            k += 1
            out[k] = v1[a, i]
    for j in range(1, nocc):
        for i in range(j):
            for b in range(nocc + 1, nso):
                for a in range(nocc, b):
                    k += 1
                    out[k] = v2[a, b, i, j]
    return out
