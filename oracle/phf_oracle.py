"""ORACLE -- test infrastructure only.

CPU restatement of the reference's spin-projection grid kernel, used solely as
the checker for the HIP path: only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it.  The product never
routes through this module.

Each function follows the reference text it cites (paths relative to the
reference repository):

  buildhsg       PHF_fort/PHFTools.f90:10-201   (BuildHSG, incl. CmplxConj=2)
  localg         PHF_fort/PHFTools.f90:203-247  (LocalG)
  rdm/rdm_conj   PHF_fort/PHFTools.f90:249-279  (RDM, RDMConj)
  kernels        PHF_fort/PHFTools.f90:281-289
  gradient       PHF_fort/PHFTools.f90:291-308
  make_gamma     PHF_fort/PHFTools.f90:310-337
  makeh          PHF_fort/PHFTools.f90:339-348
  wignerfac      PHF_fort/PHFTools.f90:350-376
  ao2mo2         PHF_fort/PHFTools.f90:390-398
  zgetrf         LAPACK ZGETRF as called by Wrappers.f90:210,310 (partial
                 pivoting on max |Re|+|Im|, IZAMAX's measure)
  determinant_c  PHF_fort/Wrappers.f90:195-231
  invert_c       PHF_fort/Wrappers.f90:295-335
  eigen_solver   PHFTools.py:8-18
  vec_decomp     PHFTools.py:20-26
  eandg          PHF.py:55-91 (objective: E and the Z gradient block)

Pinning (tests/test_oracle_*.py): brute-force Fock-space spin projection on
small lattices (energies for J=0 and J=1, exact once the grid is exact),
central finite differences of E(Z) for the gradient (factor 2 on [Re, Im]),
CmplxConj = 2 against the Fock-space matrices over {Phi, Phi*} at J = 0 and
block by block as derivatives at J > 0 (where the reference's (1,1) copy,
LocalG's exchange-block orientation and VecDecomp's product form are kept
as the reference has them),
the R=I mean-field limit against ``make_ene``, numpy.linalg det/inv for the
LU, closed-form Wigner d for j<=1.  BuildSC / BuildSC_Vec / ao2mo4 are in
``buildsc_oracle.py``.
"""
from __future__ import annotations

import math

import numpy as np


class OracleError(RuntimeError):
    pass


# ----------------------------------------------------------------------------
# LAPACK-equivalent helpers
# ----------------------------------------------------------------------------

def zgetrf(a: np.ndarray):
    """Right-looking LU with row partial pivoting, pivot = argmax |Re|+|Im|.

    Returns (lu, piv, info) with 0-based piv[i] = row swapped with row i,
    info = 1-based index of the first exactly-zero pivot (0 if none), as ZGETRF.
    """
    lu = np.array(a, dtype=np.complex128, copy=True)
    n = lu.shape[0]
    piv = np.arange(n)
    info = 0
    for k in range(n):
        col = lu[k:, k]
        p = k + int(np.argmax(np.abs(col.real) + np.abs(col.imag)))
        piv[k] = p
        if lu[p, k] != 0:
            if p != k:
                lu[[k, p], :] = lu[[p, k], :]
            lu[k + 1:, k] /= lu[k, k]
        elif info == 0:
            info = k + 1
        if k + 1 < n:
            lu[k + 1:, k + 1:] -= np.outer(lu[k + 1:, k], lu[k, k + 1:])
    return lu, piv, info


def determinant_c(m: np.ndarray) -> complex:
    """determinantC (Wrappers.f90:195-231): prod U_ii with a sign flip per
    PIV(I) != I; an exactly zero pivot gives det = 0."""
    lu, piv, info = zgetrf(m)
    if info > 0:
        return 0j
    det = 1 + 0j
    for i in range(m.shape[0]):
        det = det * lu[i, i] if piv[i] == i else -det * lu[i, i]
    return det


def invert_c(m: np.ndarray) -> np.ndarray:
    """InvertC (Wrappers.f90:295-335); the reference STOPs on info != 0."""
    lu, piv, info = zgetrf(m)
    if info != 0:
        raise OracleError("Error in ZGetRF")
    n = m.shape[0]
    perm = np.arange(n)
    for i in range(n):
        perm[[i, piv[i]]] = perm[[piv[i], i]]
    lower = np.tril(lu, -1) + np.eye(n)
    upper = np.triu(lu)
    # A = P L U  ->  A^{-1} = U^{-1} L^{-1} P^T
    pmat = np.eye(n)[perm]           # pmat @ A = L U
    inv = np.linalg.solve(upper, np.linalg.solve(lower, pmat.astype(complex)))
    return inv


# ----------------------------------------------------------------------------
# per-grid-point pieces
# ----------------------------------------------------------------------------

def ao2mo2(h: np.ndarray, m: np.ndarray) -> np.ndarray:
    """newH = M^H H M (PHFTools.f90:390-398)."""
    return m.conj().T @ (h @ m)


def rdm(r: np.ndarray, mo: np.ndarray):
    """RDM (PHFTools.f90:249-263): rho = R C (C^H R C)^{-1} C^H, det = det(C^H R C)."""
    tmp = r @ mo
    m = mo.conj().T @ tmp
    minv = invert_c(m)
    det = determinant_c(m)
    tmp = tmp @ minv
    return tmp @ mo.conj().T, det


def rdm_conj(r: np.ndarray, mo: np.ndarray):
    """RDMConj (PHFTools.f90:265-279): as RDM with R C* in place of R C."""
    tmp = r @ mo.conj()
    m = mo.conj().T @ tmp
    minv = invert_c(m)
    det = determinant_c(m)
    tmp = tmp @ minv
    return tmp @ mo.conj().T, det


def make_gamma(h2: np.ndarray, rho: np.ndarray) -> np.ndarray:
    """Gamma(i,k) = sum_{j,l} V(i,j,k,l) rho(l,j) (PHFTools.f90:310-337)."""
    nso = rho.shape[0]
    a = h2.transpose(0, 2, 1, 3).reshape(nso * nso, nso * nso)   # A[(i,k),(j,l)]
    return (a @ rho.T.reshape(nso * nso)).reshape(nso, nso)


def makeh(h1, gama, rho) -> complex:
    """e = sum H1^T * rho + sum Gamma * rho^T / 2 (PHFTools.f90:339-348)."""
    return complex(np.sum(h1.T * rho) + np.sum(gama * rho.T) / 2)


def kernels(h1, gama, rho, det) -> complex:
    """h = Makeh * det (PHFTools.f90:281-289)."""
    return makeh(h1, gama, rho) * det


def gradient(h1, gama, rho, det) -> np.ndarray:
    """G = det [ (I - rho)(H1 + Gamma) rho + e rho ] (PHFTools.f90:291-308)."""
    nso = rho.shape[0]
    e = makeh(h1, gama, rho)
    g = (h1 + gama) @ rho
    g = (np.eye(nso) - rho) @ g
    g = g + e * rho
    return g * det


def wignerfac(j: int, n: int, m: int, alpha: float, beta: float, gama: float) -> complex:
    """Wigner D^j_{nm} via the factorial sum (PHFTools.f90:350-376)."""
    fac = math.sqrt(math.gamma(j + n + 1) * math.gamma(j - n + 1)
                    * math.gamma(j + m + 1) * math.gamma(j - m + 1))
    smax = min(j + m, j - n)
    smin = max(0, m - n)
    d = 0.0
    for s in range(smin, smax + 1):
        p1 = n - m + s
        p2 = 2 * j + m - n - 2 * s
        p3 = n - m + 2 * s
        numer = (-1) ** p1 * math.cos(beta / 2) ** p2 * math.sin(beta / 2) ** p3
        denom = (math.gamma(j + m - s + 1) * math.gamma(s + 1)
                 * math.gamma(n - m + s + 1) * math.gamma(j - n - s + 1))
        d += numer / denom
    d *= fac
    return complex(np.exp(n * complex(0, -alpha)) * d * np.exp(m * complex(0, -gama)))


# ----------------------------------------------------------------------------
# BuildHSG / LocalG
# ----------------------------------------------------------------------------

def buildhsg(hone, htwo, ref, z, ngrid, nci, cmplxconj, ncisp, nocc,
             r1, r2, rpg, weightsp, weightpg, roota, rootb, rooty, qj, sp,
             check_hermitian=True, points=None, gamma=None):
    """BuildHSG (PHFTools.f90:10-201) for one rank.

    Returns (Hmat, Smat, Gradmat, Rdmmat) with shapes (nci,nci), (nci,nci),
    (nci,nci,NSO,NSO), (nci,nci,NSO,NSO).  ``points`` optionally restricts the
    grid to a list of (il, iy, ipg) triples (used for bounded CPU samples).
    ``gamma`` optionally replaces MakeGamma over the dense htwo by a callable
    rho -> Gamma (``oracle.factored.gamma_factored``; htwo may then be None).
    """
    hone = np.asarray(hone, dtype=complex)
    ref = np.asarray(ref, dtype=complex)
    nso = hone.shape[0]
    npg = rpg.shape[0]
    ncipg = weightpg.shape[1]
    if gamma is not None:
        gamma_of = gamma
    else:
        a_mat = np.asarray(htwo, dtype=complex).transpose(0, 2, 1, 3).reshape(nso * nso, nso * nso)

        def gamma_of(rho):
            return (a_mat @ rho.T.reshape(nso * nso)).reshape(nso, nso)

    ztmp = np.array(z, dtype=complex)
    if cmplxconj == 0:
        ztmp = ztmp.real.astype(complex)
    newmo = np.zeros((nso, nocc), dtype=complex)
    newmo[np.arange(nocc), np.arange(nocc)] = 1.0
    newmo[nocc:, :] = ztmp

    hmat = np.zeros((nci, nci), dtype=complex)
    smat = np.zeros((nci, nci), dtype=complex)
    gradmat = np.zeros((nci, nci, nso, nso), dtype=complex)
    rdmmat = np.zeros((nci, nci, nso, nso), dtype=complex)
    d = ncisp * ncipg
    if points is None:
        points = [(il, iy, ipg) for il in range(int(ngrid[0]))
                  for iy in range(int(ngrid[1])) for ipg in range(npg)]
    for il, iy, ipg in points:
        rtmp = r1[il] @ r2[iy]
        rtmp = rtmp @ rpg[ipg]
        r = ao2mo2(rtmp, ref)
        rho0, s0 = rdm(r, newmo)
        gama0 = gamma_of(rho0)
        h0 = kernels(hone, gama0, rho0, s0)
        g0 = gradient(hone, gama0, rho0, s0)
        if cmplxconj == 2:
            rho1, s1 = rdm_conj(r, newmo)
            gama1 = gamma_of(rho1)
            h1 = kernels(hone, gama1, rho1, s1)
            g1 = gradient(hone, gama1, rho1, s1)
            rho0c, s0c = rdm(r, newmo.conj())
            g0c = gradient(hone, gamma_of(rho0c), rho0c, s0c)
            rho1c, s1c = rdm_conj(r, newmo.conj())
            g1c = gradient(hone, gamma_of(rho1c), rho1c, s1c)
        else:
            h1, g1, s1, rho1 = h0, g0, s0, rho0
            g0c, s0c, rho0c = g0, s0, rho0
            g1c, s1c, rho1c = g0, s0, rho0
        w = weightsp[il, iy]
        rho0 = rho0 * w * s0
        rho1 = rho1 * w * s1
        rho0c = rho0c * w * s0c
        rho1c = rho1c * w * s1c
        h0 = w * h0
        h1 = w * h1
        s0 = w * s0
        s1 = w * s1
        g0 = g0 * w
        g1 = g1 * w
        g0c = g0c * w
        g1c = g1c * w
        for ni in range(-qj, qj + 1):
            for nj in range(-qj, qj + 1):
                if sp == 2:
                    wig = wignerfac(qj, ni, nj, roota[il], rootb[iy], rooty[il])
                else:
                    wig = wignerfac(qj, ni, nj, roota[il], rootb[il], rooty[iy])
                for p in range(ncipg):
                    for q in range(ncipg):
                        ww = weightpg[ipg, p, q] * wig
                        m = ni + qj + p * ncisp
                        n = nj + qj + q * ncisp
                        hmat[m, n] += h0 * ww
                        smat[m, n] += s0 * ww
                        gradmat[m, n] += g0 * ww
                        rdmmat[m, n] += rho0 * ww
                        if cmplxconj == 2:
                            hmat[m, n + d] += h1 * ww
                            smat[m, n + d] += s1 * ww
                            gradmat[m, n + d] += g1 * ww
                            rdmmat[m, n + d] += rho1 * ww
                            gradmat[m + d, n] += g1c * ww
                            rdmmat[m + d, n] += rho1c * ww
                            gradmat[m + d, n + d] += g0c * ww
                            rdmmat[m + d, n + d] += rho0c * ww
    if cmplxconj == 2:
        hmat[d:, :d] = hmat[:d, d:].conj().T
        smat[d:, :d] = smat[:d, d:].conj().T
        hmat[d:, d:] = hmat[:d, :d]
        smat[d:, d:] = smat[:d, :d]
    if check_hermitian:
        tol = 1e-5
        w1 = np.abs((hmat.conj().T - hmat).imag).max()
        w2 = np.abs((smat.conj().T - smat).imag).max()
        if w1 > np.abs(hmat).max() * tol or w2 > np.abs(smat).max() * tol:
            raise OracleError("Not Hermitian")
    return hmat, smat, gradmat, rdmmat


def localg(gradmat, rdmmat, smat, e0, fsp, fpg, fk, cmplxconj):
    """LocalG (PHFTools.f90:203-247): G = (sum f*f Grad - E0 sum f*f Rdm) / sum f*f S."""
    ncisp, ncipg = len(fsp), len(fpg)
    nso = gradmat.shape[-1]
    g = np.zeros((nso, nso), dtype=complex)
    rdm_acc = np.zeros((nso, nso), dtype=complex)
    ovlp = 0j
    d = ncisp * ncipg
    for ni in range(ncisp):
        for nj in range(ncisp):
            for p in range(ncipg):
                for q in range(ncipg):
                    m = ni + p * ncisp
                    n = nj + q * ncisp
                    f = np.conj(fsp[ni]) * fsp[nj] * np.conj(fpg[p]) * fpg[q]
                    f11 = f * np.conj(fk[0]) * fk[0]
                    g += f11 * gradmat[m, n]
                    rdm_acc += f11 * rdmmat[m, n]
                    ovlp += f11 * smat[m, n]
                    if cmplxconj == 2:
                        g += f * np.conj(fk[1]) * fk[1] * np.conj(gradmat[m + d, n + d])
                        rdm_acc += f * np.conj(fk[1]) * fk[1] * np.conj(rdmmat[m + d, n + d])
                        ovlp += f * np.conj(fk[1]) * fk[1] * smat[m + d, n + d]
                        g += f * np.conj(fk[0]) * fk[1] * gradmat[m, n + d]
                        rdm_acc += f * np.conj(fk[0]) * fk[1] * rdmmat[m, n + d]
                        ovlp += f * np.conj(fk[0]) * fk[1] * smat[m, n + d]
                        g += f * np.conj(fk[0]) * fk[1] * np.conj(gradmat[m + d, n])
                        rdm_acc += f * np.conj(fk[0]) * fk[1] * np.conj(rdmmat[m + d, n])
                        ovlp += f * np.conj(fk[1]) * fk[0] * smat[m + d, n]
    return (g - e0 * rdm_acc) / ovlp


# ----------------------------------------------------------------------------
# host numerics of the objective
# ----------------------------------------------------------------------------

def eigen_solver(h, s, tol=1e-6):
    """Canonical-orthogonalised generalized eigenproblem (PHFTools.py:8-18)."""
    evals, evecs = np.linalg.eigh(s)
    dim_full = len(evals)
    inv_s = [1 / np.sqrt(abs(i)) for i in evals if abs(i) > tol]
    dim = len(inv_s)
    x = evecs[:, dim_full - dim:] @ np.diag(inv_s)
    hh = x.conj().T @ h @ x
    evals, evecs = np.linalg.eigh(hh)
    return evals, x @ evecs


def vec_decomp(vec, ncisp, ncipg, ncik):
    """CI vector -> (fsp, fpg, fk) (PHFTools.py:20-26)."""
    m = vec.reshape((ncisp * ncipg, ncik), order='F')
    fk = m[0, :] / m[0, 0]
    m = m[:, 0].reshape((ncisp, ncipg), order='F')
    fpg = m[0, :] / m[0, 0]
    fsp = m[:, 0]
    return fsp, fpg, fk


def eandg(x, hone, htwo, ref, nocc, grid_args, cmplxconj=1, gamma=None, points=None):
    """EandG (PHF.py:55-91) without the optional Z masks: x = [Re z, Im z] ->
    (E0.real, [Re G, Im G]) where G is the (NSO-N, N) block of LocalG."""
    nso = hone.shape[0]
    nvrt = nso - nocc
    z = (x[:nvrt * nocc] + 1j * x[nvrt * nocc:]).reshape(nvrt, nocc)
    if cmplxconj == 0:
        z = z.real.astype(complex)
    ga = dict(grid_args)
    hmat, smat, gradmat, rdmmat = buildhsg(hone, htwo, ref, z, cmplxconj=cmplxconj, nocc=nocc, gamma=gamma,
                                           points=points, **ga)
    evals, evecs = eigen_solver(hmat, smat)
    e0 = evals[0]
    ncik = 2 if cmplxconj == 2 else 1
    fsp, fpg, fk = vec_decomp(evecs[:, 0], ga["ncisp"], ga["weightpg"].shape[1], ncik)
    g = localg(gradmat, rdmmat, smat, e0, fsp, fpg, fk, cmplxconj)
    g = g[nocc:, :nocc]
    if cmplxconj == 0:
        g = g.real.astype(complex)
    g = g.reshape(nvrt * nocc)
    return float(np.real(e0)), np.concatenate((g.real, g.imag))
