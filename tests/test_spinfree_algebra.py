"""Host check of the spin-free form's index conventions (no GPU): for a
spin-restricted Ref of the synthetic Hamiltonian, A (the antisymmetrised
spin-orbital ERIs in the Ref MO basis, A[(i,k),(j,l)] = <ij||kl>) applied to
density columns P[(j,l)] = rho(l,j) equals the spin-free evaluation that the
device kernels perform: the pairing of alpha and beta MOs by their spatial
coefficients, J and K gathered from A, the operand sets P_aa + P_bb,
P_aa - P_bb, P_x1, P_x2, the products (2J - K) and K, and the unpack
Gamma_aa = (Cs - Cd)/2, Gamma_bb = (Cs + Cd)/2, Gamma_ab = -C_x1,
Gamma_ba = -C_x2.  A UHF Hubbard Ref has no pairing."""
import numpy as np

from tests._cases import hubbard_cfg1, synthetic_dense


def _pairing(ref, n):
    nso = 2 * n
    spin = [0 if np.all(ref[n:, p] == 0) else (1 if np.all(ref[:n, p] == 0) else -1) for p in range(nso)]
    alpha = [p for p in range(nso) if spin[p] == 0]
    beta = [p for p in range(nso) if spin[p] == 1]
    if len(alpha) != n or len(beta) != n:
        return None
    mo = alpha + [-1] * n
    used = set()
    for s, p in enumerate(alpha):
        for q in beta:
            if q not in used and np.array_equal(ref[:n, p], ref[n:, q]):
                mo[n + s] = q
                used.add(q)
                break
        if mo[n + s] < 0:
            return None
    return np.array(mo)


def _spinfree_gamma(a4, mo, n, rho):
    """Gamma(i,k) for density rho by the spin-free route (a4[i,k,j,l] = A)."""
    al, be = mo[:n], mo[n:]
    J = a4[np.ix_(al, al, be, be)].reshape(n * n, n * n)
    K = -a4[np.ix_(al, be, be, al)].reshape(n * n, n * n)
    P = rho.T     # P[j, l] = rho[l, j]
    paa, pbb = P[np.ix_(al, al)].ravel(), P[np.ix_(be, be)].ravel()
    px1, px2 = P[np.ix_(be, al)].ravel(), P[np.ix_(al, be)].ravel()
    cs = (2 * J - K) @ (paa + pbb)
    cd, c1, c2 = K @ (paa - pbb), K @ px1, K @ px2
    nso = 2 * n
    g = np.zeros((nso, nso), complex)
    g[np.ix_(al, al)] = (0.5 * (cs - cd)).reshape(n, n)
    g[np.ix_(be, be)] = (0.5 * (cs + cd)).reshape(n, n)
    g[np.ix_(al, be)] = (-c1).reshape(n, n)
    g[np.ix_(be, al)] = (-c2).reshape(n, n)
    return g


def test_spinfree_gamma_equals_dense_contraction():
    n = 6
    case = synthetic_dense(n, 3, seed=11)
    nso = 2 * n
    # scramble the MO order (keeps the pairing nontrivial) and rotate the
    # spatial orbitals so Ref is no longer a permutation
    rng = np.random.default_rng(5)
    q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    ref = np.zeros((nso, nso))
    perm = rng.permutation(nso)
    for p in range(nso):
        s, sp = perm[p] % n, perm[p] // n
        ref[sp * n:(sp + 1) * n, p] = q[:, s]
    mo = _pairing(ref, n)
    assert mo is not None
    # A in the rotated basis: A' = ref^T A_ref ref on every index (real), from the
    # case's MO-basis htwo (Dirac <ij||kl> in Ref basis; Ref here is a permutation)
    h4 = case["htwo"].reshape(nso, nso, nso, nso)
    r0 = case["ref"].real
    vao = np.einsum("pi,qj,rk,sl,ijkl->pqrs", r0, r0, r0, r0, h4)   # back to the AO spin-orbital basis
    v = np.einsum("pi,qj,rk,sl,pqrs->ijkl", ref, ref, ref, ref, vao)  # <ij||kl> in the new MO basis
    a4 = np.transpose(v, (0, 2, 1, 3))                                  # A[(i,k),(j,l)] = <ij||kl>
    rho = rng.normal(size=(nso, nso)) + 1j * rng.normal(size=(nso, nso))
    dense = np.einsum("ikjl,lj->ik", a4, rho)
    np.testing.assert_allclose(_spinfree_gamma(a4, mo, n, rho), dense, atol=1e-12 * np.abs(dense).max())


def test_uhf_reference_has_no_pairing():
    case, nao = hubbard_cfg1()
    assert _pairing(case["ref"].real, nao) is None
