"""ORACLE -- test infrastructure only (see phf_oracle.py header).

Brute-force Fock-space spin projection, independent of the grid quadrature,
the rotation matrices and the Wigner factors.  Used to pin the oracle's
projected energies on small lattices.

States live in the N-electron determinant space of NSO spin orbitals in the
AO order [alpha_1..alpha_n, beta_1..beta_n].  A determinant with orbital
matrix C (NSO x N) has coefficients det(C[S, :]) on the ascending occupied
set S.  H = sum h_pq a+_p a_q + 1/4 sum <pq||rs> a+_p a+_q a_s a_r with
a+_p a+_q a_s a_r = E_pr E_qs - delta_qr E_ps, E_pq = a+_p a_q.

Exact projectors: P_S = prod_{S' != S} (S^2 - S'(S'+1)) / (S(S+1) - S'(S'+1))
(Loewdin).  For J > 0 the projected energy is the lowest generalized
eigenvalue on span{ L_{K->M} Pi_{Sz=K} P_J Phi : K = -J..J }, with L the
ladder map S_+^{M-K} or S_-^{K-M}: this span is independent of every phase
convention of the Wigner-D matrices.
"""
from __future__ import annotations

import itertools

import numpy as np
import scipy.sparse as sps


def _popcount(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.int64)
    c = np.zeros_like(x)
    while np.any(x):
        c += x & 1
        x = x >> 1
    return c


class FockSpace:
    def __init__(self, nso: int, nelec: int):
        self.nso = nso
        self.nelec = nelec
        dets = []
        for occ in itertools.combinations(range(nso), nelec):
            m = 0
            for o in occ:
                m |= 1 << o
            dets.append(m)
        self.dets = np.array(sorted(dets), dtype=np.int64)
        self.dim = len(self.dets)
        self._epq = {}

    def index(self, masks):
        idx = np.searchsorted(self.dets, masks)
        return idx

    def epq(self, p: int, q: int) -> sps.csr_matrix:
        """Sparse matrix of a+_p a_q on this space."""
        key = (p, q)
        if key in self._epq:
            return self._epq[key]
        d = self.dets
        has_q = (d >> q) & 1 == 1
        if p == q:
            rows = np.nonzero(has_q)[0]
            mat = sps.csr_matrix((np.ones(len(rows)), (rows, rows)), shape=(self.dim, self.dim))
        else:
            ok = has_q & (((d >> p) & 1) == 0)
            src = np.nonzero(ok)[0]
            m = d[src]
            # sign of a_q: bits below q; then a+_p on the result: bits below p
            below_q = _popcount(m & ((1 << q) - 1))
            m2 = m & ~(1 << q)
            below_p = _popcount(m2 & ((1 << p) - 1))
            sign = np.where((below_q + below_p) % 2 == 0, 1.0, -1.0)
            new = m2 | (1 << p)
            dst = self.index(new)
            mat = sps.csr_matrix((sign, (dst, src)), shape=(self.dim, self.dim))
        self._epq[key] = mat
        return mat

    def slater(self, c: np.ndarray) -> np.ndarray:
        """Coefficient vector of the determinant with orbital columns c."""
        out = np.empty(self.dim, dtype=complex)
        for k, m in enumerate(self.dets):
            rows = [i for i in range(self.nso) if (m >> i) & 1]
            out[k] = np.linalg.det(c[rows, :])
        return out

    # ------------------------------------------------------------------
    def apply_h(self, h1: np.ndarray, v: np.ndarray, vec: np.ndarray) -> np.ndarray:
        nso = self.nso
        out = np.zeros(self.dim, dtype=complex)
        ev = np.zeros((nso, nso, self.dim), dtype=complex)
        for p in range(nso):
            for q in range(nso):
                ev[p, q] = self.epq(p, q) @ vec
        out += np.einsum('pq,pqd->d', h1, ev)
        # 1/4 sum V_pqrs (E_pr E_qs - d_qr E_ps)
        w = np.einsum('pqrs,qsd->prd', v, ev, optimize=True)
        two = np.zeros(self.dim, dtype=complex)
        for p in range(nso):
            for r in range(nso):
                two += self.epq(p, r) @ w[p, r]
        corr = np.einsum('pqqs,psd->d', v, ev, optimize=True)
        out += 0.25 * (two - corr)
        return out

    def spin_ops(self):
        n = self.nso // 2
        sp = sum(self.epq(p, p + n) for p in range(n))
        sm = sp.T.tocsr()
        na = sum(self.epq(p, p) for p in range(n))
        nb = sum(self.epq(p + n, p + n) for p in range(n))
        sz = 0.5 * (na - nb)
        s2 = sm @ sp + sz @ sz + sz
        return sp.tocsr(), sm, sz.tocsr(), s2.tocsr()

    def spin_projector_apply(self, s: float, vec: np.ndarray) -> np.ndarray:
        _, _, _, s2 = self.spin_ops()
        smax = self.nelec / 2.0
        out = vec.copy()
        sprime = self.nelec % 2 / 2.0
        while sprime <= smax + 1e-9:
            if abs(sprime - s) > 1e-9:
                lam = sprime * (sprime + 1)
                out = (s2 @ out - lam * out) / (s * (s + 1) - lam)
            sprime += 1.0
        return out

    def sz_values(self) -> np.ndarray:
        n = self.nso // 2
        amask = (1 << n) - 1
        na = _popcount(self.dets & amask)
        nb = _popcount(self.dets >> n)
        return 0.5 * (na - nb)


def projected_energy(h1_ao, v_ao, c_ao, j: int = 0, m: int = 0, tol: float = 1e-8):
    """Exact spin-projected energy of the determinant with AO orbitals c_ao."""
    nso, nelec = c_ao.shape
    fs = FockSpace(nso, nelec)
    phi = fs.slater(c_ao)
    pphi = fs.spin_projector_apply(float(j), phi)
    if j == 0:
        num = np.vdot(phi, fs.apply_h(h1_ao, v_ao, pphi))
        den = np.vdot(phi, pphi)
        return float((num / den).real), float(den.real)
    return _lowest_root(fs, h1_ao, v_ao, _projected_span(fs, [pphi], j, m), tol)


def _projected_span(fs, pvecs, j, m):
    """Columns L_{K->M} Pi_{Sz=K} v for K = -j..j and every spin-projected v."""
    sp, sm, _, _ = fs.spin_ops()
    szv = fs.sz_values()
    vecs = []
    for pv in pvecs:
        for k in range(-j, j + 1):
            vk = np.where(np.abs(szv - k) < 1e-9, pv, 0.0)
            for _ in range(abs(m - k)):
                vk = (sp @ vk) if m > k else (sm @ vk)
            vecs.append(vk)
    return np.stack(vecs, 1)


def projected_energy_multi(h1_ao, v_ao, c_list, j: int = 0, m: int = 0, tol: float = 1e-8):
    """Lowest energy on the span of the spin-projected determinants c_list
    (AO orbital matrices): span{ L_{K->M} Pi_{Sz=K} P_J Phi_i }.  With
    c_list = [C, conj(C)] this is the complex-conjugation-restored projection
    (the space CmplxConj = 2 diagonalises over)."""
    nso, nelec = c_list[0].shape
    fs = FockSpace(nso, nelec)
    pvecs = [fs.spin_projector_apply(float(j), fs.slater(c)) for c in c_list]
    return _lowest_root(fs, h1_ao, v_ao, _projected_span(fs, pvecs, j, m), tol)


def projected_matrices(h1_ao, v_ao, c_list, j: int = 0):
    """<Phi_a| H P_J |Phi_b> and <Phi_a| P_J |Phi_b> for J = 0 (Loewdin P_J),
    over the determinants of c_list: the matrices the grid kernel's J = 0
    Hmat / Smat approximate up to one common normalisation."""
    nso, nelec = c_list[0].shape
    fs = FockSpace(nso, nelec)
    phis = [fs.slater(c) for c in c_list]
    pphis = [fs.spin_projector_apply(float(j), p) for p in phis]
    hp = [fs.apply_h(h1_ao, v_ao, p) for p in pphis]
    k = len(c_list)
    hm = np.array([[np.vdot(phis[a], hp[b]) for b in range(k)] for a in range(k)])
    sm = np.array([[np.vdot(phis[a], pphis[b]) for b in range(k)] for a in range(k)])
    return hm, sm


def _lowest_root(fs, h1_ao, v_ao, vmat, tol):
    hv = np.stack([fs.apply_h(h1_ao, v_ao, vmat[:, a]) for a in range(vmat.shape[1])], 1)
    hh = vmat.conj().T @ hv
    ss = vmat.conj().T @ vmat
    ev, u = np.linalg.eigh(ss)
    keep = ev > tol * ev.max()
    x = u[:, keep] / np.sqrt(ev[keep])
    e = np.linalg.eigvalsh(x.conj().T @ hh @ x)
    return float(e[0]), float(ev.max())


# ----------------------------------------------------------------------------
# helpers for the projected-CC checks (any orthonormal spin-orbital basis)
# ----------------------------------------------------------------------------

def one_body(fs: FockSpace, mat: np.ndarray):
    """sum_pq mat[p, q] a+_p a_q as a sparse matrix."""
    out = None
    for p in range(fs.nso):
        for q in range(fs.nso):
            if mat[p, q] != 0:
                term = fs.epq(p, q) * mat[p, q]
                out = term if out is None else out + term
    return out.tocsr() if out is not None else sps.csr_matrix((fs.dim, fs.dim))


def rotation_operator(fs: FockSpace, r: np.ndarray) -> np.ndarray:
    """The N-electron action of the orbital map a+_p -> sum_q r[q, p] a+_q:
    |det(C)> -> |det(r C)>, so <D'| R |D> = det(r[D', D])."""
    occ = [[i for i in range(fs.nso) if (m >> i) & 1] for m in fs.dets]
    out = np.empty((fs.dim, fs.dim), dtype=complex)
    for a, rows in enumerate(occ):
        for b, cols in enumerate(occ):
            out[a, b] = np.linalg.det(r[np.ix_(rows, cols)])
    return out


def cc_state(fs: FockSpace, t1: np.ndarray, t2: np.ndarray, nocc: int) -> np.ndarray:
    """e^{T1 + T2} |Phi> with |Phi> the determinant of orbitals 0..nocc-1,
    T1 = sum t1[a,i] a+_{nocc+a} a_i, T2 = 1/4 sum t2[a,b,i,j] a+_{nocc+a} a+_{nocc+b} a_j a_i
    (the Taylor series terminates: T^k = 0 beyond k = nocc)."""
    nv = fs.nso - nocc
    t = None
    for a in range(nv):
        for i in range(nocc):
            if t1[a, i] != 0:
                term = fs.epq(nocc + a, i) * t1[a, i]
                t = term if t is None else t + term
    for a in range(nv):
        for b in range(nv):
            for i in range(nocc):
                for j in range(nocc):
                    if t2[a, b, i, j] != 0 and a != b and i != j:
                        term = (fs.epq(nocc + a, i) @ fs.epq(nocc + b, j)) * (0.25 * t2[a, b, i, j])
                        t = term if t is None else t + term
    phi = np.zeros(fs.dim, dtype=complex)
    phi[fs.index(np.array([(1 << nocc) - 1]))[0]] = 1.0
    out, term = phi.copy(), phi.copy()
    for k in range(1, nocc + 1):
        term = (t @ term) / k
        out += term
    return out
