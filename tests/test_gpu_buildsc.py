"""GPU parity of the projected SD-space overlap (BuildSC, PHFTools.f90:464-537)
against the oracle (oracle/buildsc_oracle.py), through the C ABI.

Floating-point tolerance: FP64 on both sides with a different summation order
(closed-form pair transforms on the GPU, the generated contraction sequence in
the oracle): max |GPU - oracle| <= 1e-10 x max |oracle| per output.  Checked:
full c1/c2 inputs (phftools.buildsc, f2py order and the LinDep.py:73 form),
packed LinDep blocks, point-group and J > 0 weights with complex fsp/fpg (the
REAL weight of :518-520), batches against single vectors, linearity at
NSO = 16, the identity point (t = c, the SD basis is orthonormal) at NSO = 20,
and the singular-overlap error.
"""
import numpy as np
import pytest

from oracle import buildsc_oracle as B
from tests._cases import grid_args, random_unitary

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _cr(rng, *s):
    return rng.normal(size=s) + 1j * rng.normal(size=s)


def _mo_grid(nao, ngrid, seed, qj=0, npg=1):
    """Spin grid in an MO basis (Phi = the first nocc MOs of a random unitary)."""
    rng = np.random.default_rng(seed)
    nso = 2 * nao
    u = random_unitary(nso, rng)
    ga = grid_args(nao, ngrid, 0)
    r1 = np.einsum('pa,gpq,qb->gab', u.conj(), ga["r1"], u)
    r2 = np.einsum('pa,gpq,qb->gab', u.conj(), ga["r2"], u)
    if npg == 1:
        rpg = np.eye(nso, dtype=complex)[None]
        weightpg = np.ones((1, 1, 1))
        fpg = np.ones(1, complex)
    else:
        # a second "point-group" element: a unitary rotation, with 2 x 2 irrep weights
        rpg = np.stack([np.eye(nso, dtype=complex), random_unitary(nso, rng)])
        weightpg = rng.normal(size=(npg, 2, 2))
        fpg = _cr(rng, 2)
    fsp = _cr(rng, 2 * qj + 1) if qj else np.ones(1, complex)
    return dict(ngrid=ga["ngrid"], qj=qj, r1=r1, r2=r2, rpg=rpg, weightsp=ga["weightsp"], weightpg=weightpg,
                fsp=fsp, fpg=fpg)


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("nao,nocc,ngrid,qj,npg,seed", [
    (4, 4, (14, 8), 0, 1, 1),
    (3, 2, (6, 4), 0, 1, 2),
    (4, 3, (6, 4), 1, 2, 3),
    (5, 5, (6, 4), 0, 1, 4),
])
def test_full_tensors_match_oracle(nao, nocc, ngrid, qj, npg, seed):
    from phf_pcc_amd.PHFTools_F import phftools
    g = _mo_grid(nao, ngrid, seed, qj, npg)
    nso = 2 * nao
    rng = np.random.default_rng(100 + seed)
    c0, c1, c2 = _cr(rng), _cr(rng, nso, nso), _cr(rng, nso, nso, nso, nso)
    want = B.buildsc(c0, c1, c2, nocc, g["ngrid"], qj, g["r1"], g["r2"], g["rpg"], g["weightsp"], g["weightpg"],
                     g["fsp"], g["fpg"])
    rk = np.eye(nso, dtype=complex)[None]
    ga = grid_args(nao, ngrid, 0)
    got = phftools.buildsc(c0, c1, c2, nocc, g["ngrid"], qj, g["r1"], g["r2"], g["rpg"], rk, ga["roota"],
                           ga["rootb"], ga["rooty"], g["weightsp"], g["weightpg"], g["fsp"], g["fpg"],
                           np.ones(1, complex))
    print(f"full nso={nso} nocc={nocc} G={np.prod(g['ngrid']) * npg}", flush=True)
    assert abs(got[0] - want[0]) <= TOL * abs(want[0])
    assert _rel(got[1], want[1]) <= TOL
    assert _rel(got[2], want[2]) <= TOL
    # the call form of LinDep.py:73 (an extra nci after qj, then the optional dimensions)
    got2 = phftools.buildsc(c0, c1, c2, nocc, g["ngrid"], qj, 1, g["r1"], g["r2"], g["rpg"], rk, ga["roota"],
                            ga["rootb"], ga["rooty"], g["weightsp"], g["weightpg"], g["fsp"], g["fpg"],
                            np.ones(1, complex), nso, g["rpg"].shape[0], g["fsp"].size, g["fpg"].size, 1)
    assert np.array_equal(got2[2], got[2]) and got2[0] == got[0]


@pytest.fixture(scope="module")
def packed8():
    from phf_pcc_amd.sc import ProjectedOverlap
    g = _mo_grid(4, (14, 8), 7)
    op = ProjectedOverlap(8, 3, g["ngrid"], g["r1"], g["r2"], g["rpg"], g["weightsp"], g["weightpg"])
    return g, op


def test_packed_block_matches_oracle(packed8):
    g, op = packed8
    nso, nocc = 8, 3
    rng = np.random.default_rng(8)
    x = _cr(rng, op.dim, 5)
    got = op.matmat(x)
    for k in range(x.shape[1]):
        t = B.buildsc(*B.vec2mat(x[:, k], nso, nocc), nocc, g["ngrid"], 0, g["r1"], g["r2"], g["rpg"],
                      g["weightsp"], g["weightpg"], np.ones(1, complex), np.ones(1, complex))
        assert _rel(got[:, k], B.mat2vec(*t, nocc)) <= TOL


def test_packed_batch_equals_single_vectors(packed8):
    _, op = packed8
    x = _cr(np.random.default_rng(9), op.dim, 37)
    block = op.matmat(x)
    for k in (0, 17, 36):
        assert _rel(block[:, k], op.matvec(x[:, k])) <= 1e-12
    lin = op.as_linear_operator()
    assert _rel(lin @ x[:, 3], block[:, 3]) <= 1e-12


def test_packed_linearity_nso16():
    from phf_pcc_amd.sc import ProjectedOverlap
    g = _mo_grid(8, (14, 8), 11)
    op = ProjectedOverlap(16, 8, g["ngrid"], g["r1"], g["r2"], g["rpg"], g["weightsp"], g["weightpg"])
    assert op.dim == 1 + 64 + 28 * 28
    rng = np.random.default_rng(12)
    x, y = _cr(rng, op.dim, 3), _cr(rng, op.dim, 3)
    a, b = 0.7 - 0.2j, -1.3 + 0.4j
    lhs = op.matmat(a * x + b * y)
    rhs = a * op.matmat(x) + b * op.matmat(y)
    assert _rel(lhs, rhs) <= 1e-11
    print(f"linearity nso=16 dim={op.dim}", flush=True)


def test_identity_point_returns_input_nso20():
    from phf_pcc_amd.sc import ProjectedOverlap
    nso, nocc = 20, 10
    eye = np.eye(nso, dtype=complex)[None]
    op = ProjectedOverlap(nso, nocc, [1, 1], eye, eye, eye, np.ones((1, 1)), np.ones((1, 1, 1)))
    x = _cr(np.random.default_rng(13), op.dim, 4)
    assert _rel(op.matmat(x), x) <= 1e-13


def test_singular_overlap_raises():
    from phf_pcc_amd import _lib
    from phf_pcc_amd.sc import ProjectedOverlap
    nso, nocc = 6, 3
    perm = np.eye(nso, dtype=complex)[[3, 1, 2, 0, 4, 5]]     # swaps an occupied and a virtual orbital
    eye = np.eye(nso, dtype=complex)[None]
    with pytest.raises(_lib.SingularOverlapError):
        ProjectedOverlap(nso, nocc, [2, 1], np.stack([eye[0], perm]), eye, eye, np.ones((2, 1)),
                         np.ones((1, 1, 1)))


def test_full_and_packed_paths_agree_nso48():
    """NSO = 48, NOcc = 40 on 2 grid points: the full-tensor path (pair spaces
    over all 48 orbitals; 256-thread t0/t1 workgroups with 68 KB of LDS) and
    the packed path (V = 8) on the same vectors, a size-independent check of
    both against each other (the oracle would take minutes here)."""
    from phf_pcc_amd.sc import ProjectedOverlap
    nso, nocc = 48, 40
    g = _mo_grid(24, (1, 2), 15)
    op = ProjectedOverlap(nso, nocc, g["ngrid"], g["r1"], g["r2"], g["rpg"], g["weightsp"], g["weightpg"])
    rng = np.random.default_rng(16)
    x = _cr(rng, op.dim, 2)
    packed = op.matmat(x)
    c0 = np.array([x[0, k] for k in range(2)])
    mats = [B.vec2mat(x[:, k], nso, nocc) for k in range(2)]
    t0, t1, t2 = op.apply_full(c0, np.stack([m[1] for m in mats]), np.stack([m[2] for m in mats]))
    for k in range(2):
        full = B.mat2vec(t0[k], t1[k], t2[k], nocc)
        assert _rel(full, packed[:, k]) <= 1e-11
    print(f"nso=48 dim={op.dim}", flush=True)


@pytest.mark.parametrize("nso,seed", [(6, 21), (9, 22)])
def test_buildsc_vec_any_density(nso, seed):
    """buildsc_dr.buildsc_vec (BuildSC_Vec for an arbitrary Y, every column
    used) against the oracle's statement-by-statement restatement."""
    from phf_pcc_amd.PHFTools_F import buildsc_dr
    rng = np.random.default_rng(seed)
    c0, c1, c2, y = _cr(rng), _cr(rng, nso, nso), _cr(rng, nso, nso, nso, nso), 0.3 * _cr(rng, nso, nso)
    want = B.buildsc_vec(c0, c1, c2, y)
    got = buildsc_dr.buildsc_vec(c0, c1, c2, y)
    assert abs(got[0] - want[0]) <= TOL * abs(want[0])
    assert _rel(got[1], want[1]) <= TOL
    assert _rel(got[2], want[2]) <= TOL
