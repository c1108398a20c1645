"""GPU parity of phftools.buildhsg / localg (HIP path, through the C ABI)
against the CPU oracle on the same inputs.

Tolerance: FP64 path, energies to 1e-10 Eh (the plain-form target), every
matrix element of Hmat/Smat/Gradmat/Rdmmat to 1e-10 relative to its largest
entry.
"""
import numpy as np
import pytest

from oracle import phf_oracle as O
from tests._cases import grid_args, hubbard_cfg1, hubbard_general_ref, random_unitary, synthetic_dense

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def phftools():
    from phf_pcc_amd.PHFTools_F import phftools as t
    t.precision = "fp64"
    return t


def _run_both(phftools, case, ga, cmplxconj=1):
    args = (case["hone"], case["htwo"], case["ref"], case["z"], ga["ngrid"], ga["nci"], cmplxconj, ga["ncisp"],
            case["nocc"], ga["r1"], ga["r2"], ga["rpg"], ga["weightsp"], ga["weightpg"], ga["roota"], ga["rootb"],
            ga["rooty"], ga["qj"], ga["sp"], 0)
    gpu = phftools.buildhsg(*args)
    cpu = O.buildhsg(case["hone"], case["htwo"], case["ref"], case["z"], cmplxconj=cmplxconj, nocc=case["nocc"],
                     **ga)
    return gpu, cpu


def _assert_close(gpu, cpu, tol=1e-10):
    names = ["Hmat", "Smat", "Gradmat", "Rdmmat"]
    for n, a, b in zip(names, gpu, cpu):
        scale = max(np.abs(b).max(), 1e-300)
        err = np.abs(a - b).max() / scale
        assert err < tol, f"{n}: rel err {err:.3e}"


def _energy_and_grad(phftools, gpu, cpu, nocc, cmplxconj=1):
    out = []
    for res, is_gpu in ((gpu, True), (cpu, False)):
        h, s, gm, rm = res
        ev, vec = O.eigen_solver(h, s)
        fsp, fpg, fk = O.vec_decomp(vec[:, 0], h.shape[0], 1, 1)
        if is_gpu:
            g = phftools.localg(gm, rm, s, ev[0], fsp, fpg, fk, cmplxconj)
        else:
            g = O.localg(gm, rm, s, ev[0], fsp, fpg, fk, cmplxconj)
        out.append((ev[0].real, g[nocc:, :nocc]))
    return out


def test_hubbard_ring_plain_fp64(phftools):
    """Hubbard ring n=8, U=4, UHF + random Z, Lebedev 14 x 8 gamma (112 points)."""
    case, nao = hubbard_cfg1()
    ga = grid_args(nao, (14, 8), 0)
    gpu, cpu = _run_both(phftools, case, ga)
    _assert_close(gpu, cpu)
    (eg, gg), (ec, gc) = _energy_and_grad(phftools, gpu, cpu, case["nocc"])
    assert abs(eg - ec) < 1e-10
    assert np.abs(gg - gc).max() < 1e-10


@pytest.mark.parametrize("nsite,seed,j,ngrid", [(4, 101, 1, (26, 8)), (6, 3, 2, (38, 8))])
def test_wigner_blocks(phftools, nsite, seed, j, ngrid):
    case = hubbard_general_ref(nsite, 4.0, seed, zscale=0.3)
    ga = grid_args(nsite, ngrid, j)
    gpu, cpu = _run_both(phftools, case, ga)
    _assert_close(gpu, cpu)


def test_complex_integrals(phftools):
    """General complex Ref: complex A (four-way re/im contraction)."""
    case = hubbard_general_ref(6, 3.0, 7, zscale=0.2, complex_ref=True)
    assert np.abs(case["htwo"].imag).max() > 1e-3
    ga = grid_args(6, (14, 8), 0)
    gpu, cpu = _run_both(phftools, case, ga)
    _assert_close(gpu, cpu)


def test_real_z_cmplxconj0(phftools):
    case = hubbard_general_ref(4, 4.0, 9, zscale=0.3)
    ga = grid_args(4, (14, 8), 0)
    gpu, cpu = _run_both(phftools, case, ga, cmplxconj=0)
    _assert_close(gpu, cpu)


def test_point_group_and_sp2(phftools):
    """npg = 2 (identity + translation by two sites, weights 1/2), and the SP=2
    (beta-only) grid with an S_z-conserving Z."""
    case = hubbard_general_ref(4, 4.0, 13, zscale=0.3)
    nso = 8
    p4 = np.eye(4)[np.roll(np.arange(4), 2)]
    rpg = np.stack([np.eye(nso), np.kron(np.eye(2), p4)]).astype(complex)
    ga = grid_args(4, (14, 8), 0, rpg=rpg, weightpg=np.full((2, 1, 1), 0.5))
    gpu, cpu = _run_both(phftools, case, ga)
    _assert_close(gpu, cpu)
    z = case["z"].copy()
    z[:2, 2:] = 0
    z[2:, :2] = 0
    case_sz = dict(case, z=z)
    ga2 = grid_args(4, (1, 6), 0, sp=2)
    gpu, cpu = _run_both(phftools, case_sz, ga2)
    _assert_close(gpu, cpu)


def test_odd_sizes_and_padding(phftools):
    """n=9 spatial orbitals (NSO=18, NSO^2 not a multiple of 16), N=8."""
    case = synthetic_dense(9, 8, seed=77, zscale=0.1)
    ga = grid_args(9, (6, 5), 0)
    gpu, cpu = _run_both(phftools, case, ga)
    _assert_close(gpu, cpu)


def test_large_grid_tile_path(phftools):
    """2G > 768 rows takes the 128-row ERI tile."""
    case = hubbard_general_ref(4, 4.0, 17, zscale=0.2)
    ga = grid_args(4, (50, 16), 0)
    gpu, cpu = _run_both(phftools, case, ga)
    _assert_close(gpu, cpu)


def test_synthetic_n32(phftools):
    """Synthetic factored Hamiltonian n=32 (NSO=64), N=16, 14 x 8 grid."""
    case = synthetic_dense(32, 16, seed=1234)
    ga = grid_args(32, (14, 8), 0)
    gpu, cpu = _run_both(phftools, case, ga)
    _assert_close(gpu, cpu)
    (eg, gg), (ec, gc) = _energy_and_grad(phftools, gpu, cpu, case["nocc"])
    assert abs(eg - ec) < 1e-10
    assert np.abs(gg - gc).max() < 1e-10


def test_singular_overlap_raises(phftools):
    """A rotation that maps the occupied MOs onto virtual MOs makes M = 0 exactly
    (permutation Ref, Z = 0); the reference STOPs in InvertC (Wrappers.f90:311),
    the drop-in raises a LinAlgError."""
    from phf_pcc_amd._lib import SingularOverlapError
    case = synthetic_dense(4, 4, seed=3, zscale=0.0)
    nso = 8
    ref = case["ref"].real
    r_mo = np.zeros((nso, nso))
    r_mo[4:, :4] = np.eye(4)
    r_mo[:4, 4:] = np.eye(4)
    ga = grid_args(4, (1, 1), 0, sp=0)
    ga["r1"] = (ref @ r_mo @ ref.T)[None].astype(complex)
    with pytest.raises(O.OracleError):
        O.buildhsg(case["hone"], case["htwo"], case["ref"], case["z"], cmplxconj=1, nocc=4, **ga)
    with pytest.raises(SingularOverlapError):
        _run_both(phftools, case, ga)


def test_not_hermitian_raises(phftools):
    from phf_pcc_amd._lib import NotHermitianError
    from phf_pcc_amd.spin_proj import build_rot_mat
    case = hubbard_general_ref(4, 4.0, 41, zscale=0.3)
    ga = grid_args(4, (1, 1), 0, sp=1)
    ga["r2"] = build_rot_mat(0.0, 0.0, 0.7, 4)[None]
    args = (case["hone"], case["htwo"], case["ref"], case["z"], ga["ngrid"], ga["nci"], 1, ga["ncisp"],
            case["nocc"], ga["r1"], ga["r2"], ga["rpg"], ga["weightsp"], ga["weightpg"], ga["roota"], ga["rootb"],
            ga["rooty"], ga["qj"], ga["sp"], 0)
    with pytest.raises(NotHermitianError):
        phftools.buildhsg(*args)


def test_repeated_calls_and_hamiltonian_switch(phftools):
    """The device Hamiltonian cache follows the htwo argument."""
    c1, nao = hubbard_cfg1()
    c2 = hubbard_general_ref(8, 2.0, 5, zscale=0.1)
    ga = grid_args(8, (14, 8), 0)
    for case in (c1, c2, c1):
        gpu, cpu = _run_both(phftools, case, ga)
        _assert_close(gpu, cpu)


def test_grid_cache_follows_grid_arguments(phftools):
    """The device-resident grid is rebuilt when R1/R2/weights change."""
    case = hubbard_general_ref(4, 4.0, 23, zscale=0.3)
    for ngrid, j in [((14, 8), 0), ((26, 8), 1), ((14, 8), 0), ((50, 16), 0)]:
        ga = grid_args(4, ngrid, j)
        gpu, cpu = _run_both(phftools, case, ga)
        _assert_close(gpu, cpu)


def test_direct_c_abi_one_shot_matches_drop_in(phftools):
    """phf_buildhsg (one call: grid upload + evaluation) through ctypes."""
    from phf_pcc_amd import _lib
    case = hubbard_general_ref(4, 4.0, 29, zscale=0.3)
    ga = grid_args(4, (14, 8), 0)
    gpu, cpu = _run_both(phftools, case, ga)
    ctx = phftools.context()
    lib = _lib.load()
    C = lambda a, dt=np.complex128: np.ascontiguousarray(a, dtype=dt)
    ngrid = np.ascontiguousarray(ga["ngrid"], dtype=np.int32)
    out = [np.zeros((1, 1), complex), np.zeros((1, 1), complex), np.zeros((1, 1, 8, 8), complex),
           np.zeros((1, 1, 8, 8), complex)]
    rc = lib.phf_buildhsg(ctx.handle, _lib.ptr(C(case["ref"])), _lib.ptr(C(case["z"])), ngrid.ctypes.data_as(_lib._ip),
                          1, 1, 1, 4, _lib.ptr(C(ga["r1"])), ga["r1"].shape[0], _lib.ptr(C(ga["r2"])), ga["r2"].shape[0],
                          _lib.ptr(C(ga["rpg"])), 1, _lib.ptr(C(ga["weightsp"], np.float64)),
                          _lib.ptr(C(ga["weightpg"], np.float64)), 1, _lib.ptr(C(ga["roota"], np.float64)),
                          len(ga["roota"]), _lib.ptr(C(ga["rootb"], np.float64)), len(ga["rootb"]),
                          _lib.ptr(C(ga["rooty"], np.float64)), len(ga["rooty"]), 0, 1, 0,
                          *[_lib.ptr(o) for o in out])
    _lib.check(rc, ctx.handle)
    phftools.invalidate()
    _assert_close(out, cpu)


@pytest.mark.parametrize("nsite,seed,j,ngrid,cplx", [(4, 5, 0, (14, 8), False), (6, 7, 0, (14, 8), False),
                                                     (4, 11, 1, (26, 8), False), (4, 5, 0, (14, 8), True)])
def test_cmplxconj2_against_oracle(phftools, nsite, seed, j, ngrid, cplx):
    """CmplxConj = 2: four density variants per grid point, nci = 2 ncisp ncipg,
    block assembly of Hmat/Smat/Gradmat/Rdmmat and the LocalG cross terms."""
    case = hubbard_general_ref(nsite, 4.0, seed, zscale=0.3, complex_ref=cplx)
    ga = grid_args(nsite, ngrid, j)
    ga = dict(ga, nci=2 * ga["nci"])
    gpu, cpu = _run_both(phftools, case, ga, cmplxconj=2)
    _assert_close(gpu, cpu)
    out = []
    for res, is_gpu in ((gpu, True), (cpu, False)):
        h, s, gm, rm = res
        ev, vec = O.eigen_solver(h, s)
        fsp, fpg, fk = O.vec_decomp(vec[:, 0], ga["ncisp"], 1, 2)
        fn = phftools.localg if is_gpu else O.localg
        g = fn(gm, rm, s, ev[0], fsp, fpg, fk, 2)
        out.append((ev[0].real, g[case["nocc"]:, :case["nocc"]]))
    (eg, gg), (ec, gc) = out
    assert abs(eg - ec) < 1e-10
    assert np.abs(gg - gc).max() < 1e-10


def test_cmplxconj2_model_and_mixed(phftools):
    """The device-resident model falls back to buildhsg + localg for CmplxConj = 2;
    its E and gradient equal the oracle's EandG, and f16x2 stays within 1e-6 Eh."""
    from phf_pcc_amd.model import ProjectedHF
    case = hubbard_general_ref(6, 4.0, 7, zscale=0.3)
    ga = grid_args(6, (14, 8), 0)
    ga = dict(ga, nci=2 * ga["nci"])
    m = ProjectedHF(case["nocc"], case["ref"], ga, hone=case["hone"], htwo=case["htwo"], cmplxconj=2)
    e, g = m.energy_and_gradient(case["z"])
    x = np.concatenate([case["z"].real.ravel(), case["z"].imag.ravel()])
    e_ref, g_ref = O.eandg(x, case["hone"], case["htwo"], case["ref"], case["nocc"], ga, cmplxconj=2)
    assert abs(e - e_ref) < 1e-10
    assert np.abs(np.concatenate([g.real.ravel(), g.imag.ravel()]) - g_ref).max() < 1e-10
    m.precision = "f16x2"
    e16, g16 = m.energy_and_gradient(case["z"])
    assert abs(e16 - e) <= 1e-6
    assert np.abs(g16 - g).max() <= 1e-5


def _last_grid_rotation(phftools):
    lines = [ln for ln in phftools.context().tuning_report().splitlines() if '"grid"' in ln]
    return lines[-1] if lines else ""


def test_general_rotations_match_spin_rotation_path(phftools):
    """Spin grids (every R1, R2 exactly u (x) I_n) take the four-product
    rotation path; conjugating R1, R2, Rpg and Ref by a random unitary W of
    the spin-orbital space leaves Ref^H R Ref, hence every output, unchanged
    but breaks that structure, so the same problem runs the general chain.
    Both equal the oracle to 1e-10."""
    case = hubbard_general_ref(4, 4.0, 33, zscale=0.2)
    ga = grid_args(4, (14, 8), 0)
    gpu_su2, cpu = _run_both(phftools, case, ga)
    assert '"rotation": "su2 x identity"' in _last_grid_rotation(phftools)
    w = random_unitary(8, np.random.default_rng(34))
    conj = lambda r: np.einsum('pq,gqr,sr->gps', w, r, w.conj())
    ga2 = dict(ga, r1=conj(ga["r1"]), r2=conj(ga["r2"]), rpg=conj(ga["rpg"]))
    case2 = dict(case, ref=w @ case["ref"])
    gpu_gen, cpu2 = _run_both(phftools, case2, ga2)
    assert '"rotation": "general"' in _last_grid_rotation(phftools)
    _assert_close(gpu_su2, cpu)
    _assert_close(gpu_gen, cpu2)
    _assert_close(gpu_gen, gpu_su2)
