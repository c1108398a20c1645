"""Grid builders (phf_pcc_amd/spin_proj.py) against the reference's Lebedev
tables (tests/golden/lebedev_reference.npz) and quadrature exactness."""
import math
import os

import numpy as np
import pytest

from phf_pcc_amd import spin_proj as SP

GOLD = os.path.join(os.path.dirname(__file__), "golden", "lebedev_reference.npz")


def _sorted(m):
    return m[np.lexsort((m[:, 1].round(8), m[:, 0].round(8)))]


@pytest.mark.parametrize("npts", [1, 6, 14, 26, 38, 50])
def test_lebedev_matches_reference_tables(npts):
    ref = np.load(GOLD)[f"n{npts}"]
    a, b, w = SP.lebedev_angles(npts)
    mine = np.stack([np.degrees(a), np.degrees(b), w], 1)
    assert mine.shape == ref.shape
    assert np.abs(_sorted(mine) - _sorted(ref)).max() < 1e-12


def _sphere_mean(a, b, c):
    """Average of x^a y^b z^c over the unit sphere."""
    if a % 2 or b % 2 or c % 2:
        return 0.0
    df = lambda k: math.prod(range(k, 0, -2)) if k > 0 else 1
    return df(a - 1) * df(b - 1) * df(c - 1) / df(a + b + c + 1)


@pytest.mark.parametrize("npts,deg", [(6, 3), (14, 5), (26, 7), (38, 9), (50, 11)])
def test_lebedev_exact_to_degree(npts, deg):
    xyzw = SP.lebedev_xyzw(npts)
    x, y, z, w = xyzw.T
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                q = np.sum(w * x ** a * y ** b * z ** c)
                assert abs(q - _sphere_mean(a, b, c)) < 1e-13
    # and not exact one degree higher (the rule is not over-specified)
    assert abs(np.sum(w * (x ** (deg + 1) + y ** (deg + 1) + z ** (deg + 1))) - 3 * _sphere_mean(deg + 1, 0, 0)) > 1e-6


def test_gamma_trapezoid():
    r, w = SP.gamma_trapezoid(8)
    assert abs(w.sum() - 2 * np.pi) < 1e-14 and r[0] == 0 and abs(r[-1] - 2 * np.pi) < 1e-14
    assert w[0] == w[-1] == w[1] / 2


def test_rotation_composition_and_unitarity():
    nao = 3
    a, b, y = 0.3, 1.1, 2.0
    r = SP.build_rot_mat(a, b, y, nao)
    assert np.allclose(r @ r.conj().T, np.eye(6))
    assert np.allclose(SP.build_rot_mat(a, b, 0, nao) @ SP.build_rot_mat(0, 0, y, nao), r)
    # 2pi rotation is -1 on every spin orbital
    assert np.allclose(SP.build_rot_mat(0, 0, 2 * np.pi, nao), -np.eye(6))


def test_wigner_unitarity_and_product_rule():
    j = 2
    def dmat(a, b, g):
        return np.array([[SP.wigner_d(j, n, m, a, b, g) for m in range(-j, j + 1)] for n in range(-j, j + 1)])
    d = dmat(0.4, 1.2, -0.3)
    assert np.allclose(d @ d.conj().T, np.eye(2 * j + 1))
    # D(alpha, 0, 0) D(0, 0, gamma) = D(alpha + gamma, 0, 0)
    assert np.allclose(dmat(0.3, 0, 0) @ dmat(0, 0, 0.5), dmat(0.8, 0, 0))


def test_root_weight_sy_is_gauss_legendre():
    """SpinSymm.py:5-30 maps the Gauss-Legendre nodes linearly to beta = (x+1) pi/2
    and multiplies the weights by pi/2 sin(beta)."""
    rb, wb = SP.root_weight_sy(6)
    x, wx = np.polynomial.legendre.leggauss(6)
    order = np.argsort(rb)
    assert np.allclose(rb[order], (x + 1) * np.pi / 2)
    assert np.allclose(wb[order], wx * np.pi / 2 * np.sin((x + 1) * np.pi / 2))
    # divided by sin(beta) it is an exact rule for polynomials in beta up to degree 11
    assert abs(np.sum(wb / np.sin(rb) * rb ** 11) - np.pi ** 12 / 12) < 1e-9 * np.pi ** 12
