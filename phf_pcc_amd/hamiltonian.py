"""Hamiltonian and orbital inputs for the spin-projection grid kernel.

Reference behaviour followed:
  * ``hubbard``            makeH.py:5-33 (BuildHamitonian)
  * ``spin_orbital_blocks`` Main_PHF.py:22-29 (spin-orbital block assembly)
  * ``mulliken_to_dirac``  makeH.py:63-67 (Mulliken -> antisymmetrised Dirac)
  * ``ao2mo``              ao2mo.py:4-10
  * ``sort_orb``           PHFTools.py:230-242 (order [occ a, occ b, vir a, vir b])
  * ``make_ene``           makeH.py:79-84 (mean-field energy, used in checks)

Spin-orbital order in the AO basis is [alpha_1..alpha_n, beta_1..beta_n].
``HTwo[i,j,k,l]`` is <ij||kl> = (ik|jl) - (il|jk).

Synthetic factored Hamiltonians (used for the n=64/96/128 workloads):
  rng = numpy.random.default_rng(seed); then, in this order,
  h    : symmetric; diagonal ~ U(-5,-0.5), upper triangle ~ N(0, 0.1^2)
  B^L  : for L in range(nL): upper triangle ~ N(0, 0.05^2), mirrored, times
         exp(-|p-q|/8)
  Z    : Re ~ N(0, 0.05^2) then Im ~ N(0, 0.05^2), shape (2n-N, N)
  (pq|rs) = sum_L B^L_pq B^L_rs; Ref = the permutation that puts the N/2
  spatial orbitals of lowest h_pp first for both spins.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


# ----------------------------------------------------------------------------
# reference input builders
# ----------------------------------------------------------------------------

def hubbard(nx: int, ny: int, u: float, pbcx: bool = True, pbcy: bool = True):
    """Hubbard lattice hopping T0 (n x n) and on-site (pp|pp)=u ERI (makeH.py:5-33)."""
    n = nx * ny
    t0 = np.zeros((n, n))
    u0 = np.zeros((n, n, n, n))
    for i in range(n):
        u0[i, i, i, i] = u
    for i in range(nx):
        for j in range(ny - 1):
            a, b = ny * i + j, ny * i + j + 1
            t0[a, b] = t0[b, a] = -1.0
    for i in range(nx - 1):
        for j in range(ny):
            a, b = ny * i + j, ny * (i + 1) + j
            t0[a, b] = t0[b, a] = -1.0
    if pbcy and ny > 1:
        for i in range(nx):
            a, b = ny * i, ny * i + ny - 1
            t0[a, b] += -1.0
            t0[b, a] += -1.0
    if pbcx and nx > 1:
        for j in range(ny):
            a, b = j, ny * (nx - 1) + j
            t0[a, b] += -1.0
            t0[b, a] += -1.0
    return t0, u0


def spin_orbital_blocks(h1: np.ndarray, eri: np.ndarray):
    """Spin-orbital one-body and Mulliken two-body arrays (Main_PHF.py:22-29)."""
    n = h1.shape[0]
    nso = 2 * n
    hone = np.zeros((nso, nso), dtype=h1.dtype)
    hone[:n, :n] = h1
    hone[n:, n:] = h1
    htwo = np.zeros((nso, nso, nso, nso), dtype=eri.dtype)
    htwo[:n, :n, :n, :n] = eri
    htwo[n:, n:, n:, n:] = eri
    htwo[n:, n:, :n, :n] = eri
    htwo[:n, :n, n:, n:] = eri
    return hone, htwo


def mulliken_to_dirac(h2: np.ndarray) -> np.ndarray:
    """<ab||cd> = (ac|bd) - (ad|bc) from Mulliken (pq|rs) (makeH.py:63-67)."""
    tmp = np.ascontiguousarray(h2.transpose(0, 2, 1, 3))   # tmp[a,b,c,d] = (ac|bd)
    return tmp - tmp.transpose(0, 1, 3, 2)


def ao2mo(mat: np.ndarray, mos: np.ndarray, dim: int) -> np.ndarray:
    """C^H mat C (dim=2) or the four-index analogue (dim=4) (ao2mo.py:4-10)."""
    ch = mos.conj().T
    if dim == 2:
        return ch @ mat @ mos
    if dim == 4:
        t = np.einsum('ia,abcd->ibcd', ch, mat, optimize=True)
        t = np.einsum('jb,ibcd->ijcd', ch, t, optimize=True)
        t = np.einsum('ijcd,ck->ijkd', t, mos, optimize=True)
        return np.einsum('ijkd,dl->ijkl', t, mos, optimize=True)
    raise ValueError("dim must be 2 or 4")


def make_ene(h1, h2, mo_occ):
    """Mean-field energy of the determinant mo_occ (makeH.py:79-84)."""
    rho = mo_occ @ mo_occ.conj().T
    e = np.einsum('ij,ji', h1, rho)
    e += 0.25 * np.einsum('ijkl,ki,lj', h2, rho, rho)
    e -= 0.25 * np.einsum('ijkl,kj,li', h2, rho, rho)
    return e


def sort_orb(mo_a: np.ndarray, mo_b: np.ndarray, nocc_a: int, nocc_b: int) -> np.ndarray:
    """Spin-orbital coefficient matrix in the order [occ a, occ b, vir a, vir b]
    (the SortOrb type-1 layout, PHFTools.py:230-242) from spatial UHF/RHF orbitals."""
    n = mo_a.shape[0]
    nso = 2 * n
    ref = np.zeros((nso, nso), dtype=np.result_type(mo_a, mo_b))
    col = 0
    for k in range(nocc_a):
        ref[:n, col] = mo_a[:, k]; col += 1
    for k in range(nocc_b):
        ref[n:, col] = mo_b[:, k]; col += 1
    for k in range(nocc_a, n):
        ref[:n, col] = mo_a[:, k]; col += 1
    for k in range(nocc_b, n):
        ref[n:, col] = mo_b[:, k]; col += 1
    return ref


def uhf_hubbard(t0: np.ndarray, u: float, nocc_a: int, nocc_b: int,
                max_iter: int = 2000, tol: float = 1e-12, damp: float = 0.5):
    """Unrestricted HF for an on-site Hubbard model with an antiferromagnetic guess.

    F_a = T + U diag(n_b), F_b = T + U diag(n_a).  Returns (mo_a, mo_b, energy).
    """
    n = t0.shape[0]
    na = np.array([0.5 + 0.4 * (-1) ** i for i in range(n)])
    nb = np.array([0.5 - 0.4 * (-1) ** i for i in range(n)])
    e_old = None
    for _ in range(max_iter):
        ea, ca = np.linalg.eigh(t0 + u * np.diag(nb))
        eb, cb = np.linalg.eigh(t0 + u * np.diag(na))
        na_new = (ca[:, :nocc_a] ** 2).sum(1)
        nb_new = (cb[:, :nocc_b] ** 2).sum(1)
        e = (np.trace(t0 @ (ca[:, :nocc_a] @ ca[:, :nocc_a].T))
             + np.trace(t0 @ (cb[:, :nocc_b] @ cb[:, :nocc_b].T))
             + u * np.dot(na_new, nb_new))
        if e_old is not None and abs(e - e_old) < tol and np.abs(na_new - na).max() < 1e-10:
            break
        e_old = e
        na = damp * na + (1 - damp) * na_new
        nb = damp * nb + (1 - damp) * nb_new
    return ca, cb, float(e)


# ----------------------------------------------------------------------------
# synthetic factored Hamiltonians
# ----------------------------------------------------------------------------

def synthetic_factors(n: int, nl: int, rng: np.random.Generator):
    """Random h (n x n) and B (nl x n x n) per the construction in the module docstring."""
    h = np.zeros((n, n))
    h[np.diag_indices(n)] = rng.uniform(-5.0, -0.5, size=n)
    iu = np.triu_indices(n, 1)
    h[iu] = rng.normal(0.0, 0.1, size=len(iu[0]))
    h[(iu[1], iu[0])] = h[iu]
    iu0 = np.triu_indices(n)
    decay = np.exp(-np.abs(np.subtract.outer(np.arange(n), np.arange(n))) / 8.0)
    b = np.zeros((nl, n, n))
    for l in range(nl):
        v = rng.normal(0.0, 0.05, size=len(iu0[0]))
        b[l][iu0] = v
        b[l][(iu0[1], iu0[0])] = v
        b[l] *= decay
    return h, b


def sorted_ref(h: np.ndarray, nocc: int) -> np.ndarray:
    """Permutation Ref = [occ a, occ b, vir a, vir b] with spatial orbitals sorted by h_pp."""
    n = h.shape[0]
    order = np.argsort(np.diag(h), kind="stable")
    eye = np.eye(n)[:, order]
    return sort_orb(eye, eye, nocc // 2, nocc // 2)


def eri_from_factors(b: np.ndarray) -> np.ndarray:
    """(pq|rs) = sum_L B^L_pq B^L_rs as an n^4 array."""
    nl, n, _ = b.shape
    flat = b.reshape(nl, n * n)
    return (flat.T @ flat).reshape(n, n, n, n)


def dirac_mo(h: np.ndarray, eri: np.ndarray, ref: np.ndarray):
    """MO-basis spin-orbital HOne / <ij||kl> from spatial h, Mulliken eri and Ref
    (Main_PHF.py:22-33).  Uses a direct construction when Ref is a permutation."""
    n = h.shape[0]
    nso = 2 * n
    is_perm = (np.isrealobj(ref) or np.abs(ref.imag).max() == 0) and \
        np.all(np.isin(ref.real, (0.0, 1.0))) and np.all(ref.real.sum(0) == 1)
    hone, _ = spin_orbital_blocks(h, np.zeros((1, 1, 1, 1)))
    if not is_perm:
        hone_so, htwo_m = spin_orbital_blocks(h, eri)
        htwo = mulliken_to_dirac(htwo_m)
        return ao2mo(hone_so, ref, 2), ao2mo(htwo, ref, 4)
    perm = np.argmax(ref.real, axis=0)          # MO index -> AO spin-orbital index
    sp = perm // n                                  # spin of each MO
    p = perm % n                                    # spatial index of each MO
    h1 = hone[np.ix_(perm, perm)]
    # (ik|jl) delta(si,sk) delta(sj,sl) - (il|jk) delta(si,sl) delta(sj,sk)
    coul = eri[np.ix_(p, p, p, p)]                 # coul[i,k,j,l] = (ik|jl)
    same = (sp[:, None] == sp[None, :])
    h2 = coul.transpose(0, 2, 1, 3) * (same[:, None, :, None] & same[None, :, None, :])
    h2 = h2 - h2.transpose(0, 1, 3, 2)
    return h1, np.ascontiguousarray(h2)


@dataclass
class Problem:
    """One projected-HF evaluation problem in the reference's MO-basis form."""
    name: str
    n: int
    nocc: int
    h_spatial: np.ndarray
    hone: np.ndarray                  # [NSO, NSO] MO basis
    ref: np.ndarray                   # [NSO, NSO]
    z: np.ndarray                     # [NSO-N, N] complex
    ngrid: tuple
    htwo: np.ndarray | None = None    # [NSO]^4 MO basis <ij||kl>, built on demand
    b: np.ndarray | None = None       # factored ERI [L, n, n] (synthetic problems)
    eri: np.ndarray | None = None     # spatial Mulliken n^4 (small problems)
    extra: dict = field(default_factory=dict)

    @property
    def nso(self):
        return 2 * self.n


def hubbard_ring_problem(nsite: int = 8, u: float = 4.0, seed: int = 1001,
                         zscale: float = 0.05, ngrid=(14, 8)) -> Problem:
    """1D Hubbard ring at half filling, UHF reference, random complex Thouless Z."""
    t0, u0 = hubbard(1, nsite, u, pbcx=True, pbcy=True)
    nocc = nsite
    ca, cb, _ = uhf_hubbard(t0, u, nocc // 2, nocc // 2)
    ref = sort_orb(ca, cb, nocc // 2, nocc // 2)
    hone_so, htwo_m = spin_orbital_blocks(t0, u0)
    htwo_ao = mulliken_to_dirac(htwo_m)
    h1 = ao2mo(hone_so, ref, 2)
    h2 = ao2mo(htwo_ao, ref, 4)
    rng = np.random.default_rng(seed)
    nso = 2 * nsite
    z = rng.normal(0, zscale, size=(nso - nocc, nocc)) + 1j * rng.normal(0, zscale, size=(nso - nocc, nocc))
    return Problem(f"hubbard{nsite}_U{u:g}", nsite, nocc, t0, h1, ref, z, tuple(ngrid), htwo=h2, eri=u0,
                   extra={"htwo_ao": htwo_ao, "hone_ao": hone_so})


def synthetic_problem(n: int, nocc: int, nl: int | None = None, seed: int = 1002,
                      ngrid=(50, 16), build_dense: bool = False, zscale: float = 0.05) -> Problem:
    """Synthetic molecular-like factored Hamiltonian (construction in the module docstring)."""
    nl = 3 * n if nl is None else nl
    rng = np.random.default_rng(seed)
    h, b = synthetic_factors(n, nl, rng)
    nso = 2 * n
    z = rng.normal(0, zscale, size=(nso - nocc, nocc)) + 1j * rng.normal(0, zscale, size=(nso - nocc, nocc))
    ref = sorted_ref(h, nocc)
    hone_so, _ = spin_orbital_blocks(h, np.zeros((1, 1, 1, 1)))
    h1 = ao2mo(hone_so, ref, 2)
    prob = Problem(f"synthetic_n{n}_N{nocc}_L{nl}", n, nocc, h, h1, ref, z, tuple(ngrid), b=b)
    if build_dense:
        prob.eri = eri_from_factors(b)
        _, prob.htwo = dirac_mo(h, prob.eri, ref)
    return prob


def random_unitary(n: int, rng, real: bool = False) -> np.ndarray:
    """Haar-distributed unitary (orthogonal when real) from QR with the phase of R's diagonal removed."""
    a = rng.normal(size=(n, n)) + (0 if real else 1j * rng.normal(size=(n, n)))
    q, r = np.linalg.qr(a)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def batch_instances(ninst: int = 64, seed: int = 1004):
    """The uneven batch workload: n uniform on {16, 24, ..., 96}, N = n/2,
    L = 3n, grids Lebedev {6, 14, 26, 38, 50} x gamma {4, 8, 16} drawn per
    instance (instance i seeded with seed * 1000 + i).  Returns (instances,
    zs, ns, grid points per instance) with instances in the form
    ProjectedHFBatch takes (keys h, b, ref, nocc, ngrid, r1, r2, rpg, weightsp)."""
    from .spin_proj import grid_args
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.arange(16, 97, 8), size=ninst)
    nls = rng.choice([6, 14, 26, 38, 50], size=ninst)
    nys = rng.choice([4, 8, 16], size=ninst)
    inst, zs = [], []
    for i in range(ninst):
        n, ngrid = int(ns[i]), (int(nls[i]), int(nys[i]))
        p = synthetic_problem(n, n // 2, seed=seed * 1000 + i, ngrid=ngrid)
        ga = grid_args(n, ngrid, 0)
        inst.append(dict(h=p.h_spatial, b=p.b, ref=p.ref, nocc=n // 2, ngrid=ga["ngrid"], r1=ga["r1"], r2=ga["r2"],
                         rpg=ga["rpg"], weightsp=ga["weightsp"]))
        zs.append(p.z)
    return inst, zs, ns, nls * nys
