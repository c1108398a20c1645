"""Spin-projection quadrature grids and rotation matrices.

Host-side builders of the inputs that ``phftools.buildhsg`` consumes: the
Lebedev (alpha, beta) rule times a trapezoid gamma rule, the spin-rotation
matrices R1/R2 and the Wigner-D factors.

Reference behaviour followed:
  * ``Spin_Proj``      Spin_Proj.py:3-39
  * ``RootWeight``     SpinSymm.py:32-68  (gamma trapezoid with both ends,
                       Lebedev table picked by point count)
  * ``RootWeightSy``   SpinSymm.py:5-30   (Golub-Welsch for SP=2)
  * ``BuildRotMat``    SpinSymm.py:70-89
  * ``WignerMat``      WignerMat.py:8-27 / PHFTools.f90:350-376

The Lebedev rules are generated from their octahedral orbit generators
(Lebedev 1975/1976) instead of being read from tables.  Every weight and
generator below has a closed form:

  degree 3  (6)   a1: 1/6
  degree 5  (14)  a1: 1/15,  a3: 3/40
  degree 7  (26)  a1: 1/21,  a2: 4/105,  a3: 9/280
  degree 9  (38)  a1: 1/105, a3: 9/280,  c1: 1/35 with p^2=(1+1/sqrt3)/2,
                  q^2=(1-1/sqrt3)/2  (from exactness on the octahedral
                  invariants 1, sum x^4, x^2y^2z^2, (sum x^4)^2)
  degree 11 (50)  a1: 4/315, a2: 64/2835, a3: 27/1280,
                  b1: 14641/725760 with l=1/sqrt11, m=3/sqrt11

Points are converted to (azimuth alpha, polar beta) in radians.  At the two
poles the azimuth is set to pi/2, the value the reference tables carry
(lebedev_005.txt:5-6), because alpha at a pole still shifts the sampled
gamma points and so changes the quadrature result.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

LEBEDEV_POINTS = {1: 0, 6: 3, 14: 5, 26: 7, 38: 9, 50: 11}  # npts -> degree


def _orbit_a1():
    pts = []
    for ax in range(3):
        for s in (1.0, -1.0):
            v = [0.0, 0.0, 0.0]
            v[ax] = s
            pts.append(v)
    return pts


def _orbit_a2():
    r = 1.0 / math.sqrt(2.0)
    pts = []
    for zero_ax in range(3):
        others = [a for a in range(3) if a != zero_ax]
        for s1 in (1.0, -1.0):
            for s2 in (1.0, -1.0):
                v = [0.0, 0.0, 0.0]
                v[others[0]] = s1 * r
                v[others[1]] = s2 * r
                pts.append(v)
    return pts


def _orbit_a3():
    r = 1.0 / math.sqrt(3.0)
    return [[sx * r, sy * r, sz * r] for sx in (1.0, -1.0) for sy in (1.0, -1.0) for sz in (1.0, -1.0)]


def _orbit_b(l, m):
    """(+-l, +-l, +-m) and permutations: 24 points."""
    pts = []
    for m_ax in range(3):
        others = [a for a in range(3) if a != m_ax]
        for s1 in (1.0, -1.0):
            for s2 in (1.0, -1.0):
                for s3 in (1.0, -1.0):
                    v = [0.0, 0.0, 0.0]
                    v[others[0]] = s1 * l
                    v[others[1]] = s2 * l
                    v[m_ax] = s3 * m
                    pts.append(v)
    return pts


def _orbit_c(p, q):
    """(+-p, +-q, 0) and permutations: 24 points."""
    pts = []
    for zero_ax in range(3):
        others = [a for a in range(3) if a != zero_ax]
        for (i, j) in ((others[0], others[1]), (others[1], others[0])):
            for s1 in (1.0, -1.0):
                for s2 in (1.0, -1.0):
                    v = [0.0, 0.0, 0.0]
                    v[i] = s1 * p
                    v[j] = s2 * q
                    pts.append(v)
    return pts


def lebedev_xyzw(npts: int) -> np.ndarray:
    """Cartesian Lebedev rule: array (npts, 4) of x, y, z, weight (weights sum to 1)."""
    if npts == 1:
        return np.array([[0.0, 0.0, 1.0, 1.0]])
    rule = []
    if npts == 6:
        rule = [(_orbit_a1(), 1.0 / 6.0)]
    elif npts == 14:
        rule = [(_orbit_a1(), 1.0 / 15.0), (_orbit_a3(), 3.0 / 40.0)]
    elif npts == 26:
        rule = [(_orbit_a1(), 1.0 / 21.0), (_orbit_a2(), 4.0 / 105.0), (_orbit_a3(), 9.0 / 280.0)]
    elif npts == 38:
        p = math.sqrt((1.0 + 1.0 / math.sqrt(3.0)) / 2.0)
        q = math.sqrt((1.0 - 1.0 / math.sqrt(3.0)) / 2.0)
        rule = [(_orbit_a1(), 1.0 / 105.0), (_orbit_a3(), 9.0 / 280.0), (_orbit_c(p, q), 1.0 / 35.0)]
    elif npts == 50:
        l = 1.0 / math.sqrt(11.0)
        m = 3.0 / math.sqrt(11.0)
        rule = [(_orbit_a1(), 4.0 / 315.0), (_orbit_a2(), 64.0 / 2835.0),
                (_orbit_a3(), 27.0 / 1280.0), (_orbit_b(l, m), 14641.0 / 725760.0)]
    else:
        raise ValueError(
            f"Lebedev rule with {npts} points is not supported; use one of {sorted(LEBEDEV_POINTS)}")
    out = []
    for pts, w in rule:
        for v in pts:
            out.append([v[0], v[1], v[2], w])
    return np.array(out, dtype=np.float64)


def lebedev_angles(npts: int):
    """(alpha, beta, weight) in radians for the npts-point rule (SpinSymm.py:85-107)."""
    if npts == 1:
        # lebedev_000.txt holds the single point (0, 0) with weight 1
        return np.zeros(1), np.zeros(1), np.ones(1)
    xyzw = lebedev_xyzw(npts)
    alpha = np.empty(npts)
    beta = np.empty(npts)
    for i, (x, y, z, _) in enumerate(xyzw):
        x = 0.0 if abs(x) < 1e-14 else x
        y = 0.0 if abs(y) < 1e-14 else y
        if x == 0.0 and y == 0.0:
            a = 0.5 * math.pi
        else:
            a = math.atan2(y, x)
        alpha[i] = a
        beta[i] = math.acos(max(-1.0, min(1.0, z)))
    return alpha, beta, xyzw[:, 3].copy()


def gamma_trapezoid(ny: int):
    """gamma roots/weights on [0, 2pi] with both end points (SpinSymm.py:34-45)."""
    if ny == 1:
        return np.zeros(1), np.ones(1)
    dx = 2.0 * math.pi / (ny - 1)
    roots = np.array([i * dx for i in range(ny)])
    w = np.full(ny, dx)
    w[0] = dx / 2.0
    w[-1] = dx / 2.0
    return roots, w


def root_weight(ngrid):
    """RootWeight (SpinSymm.py:32-68): roota, rootb, rooty, weight[nl, ny]."""
    nl, ny = int(ngrid[0]), int(ngrid[1])
    rooty, wy = gamma_trapezoid(ny)
    roota, rootb, wl = lebedev_angles(nl)
    weight = np.outer(wl, wy)
    return roota, rootb, rooty, weight


def root_weight_sy(ngrid: int):
    """RootWeightSy (SpinSymm.py:5-30): Gauss-Legendre in cos(beta) via the
    symmetric tridiagonal Jacobi matrix (Golub-Welsch), mapped to beta in [0, pi]
    and multiplied by sin(beta)."""
    n = int(ngrid)
    if n == 1:
        return np.zeros(1), np.ones(1)
    k = np.arange(1, n + 1, dtype=np.float64)
    a = 2.0 - 1.0 / k
    c = 1.0 - 1.0 / k
    off = np.sqrt(c[1:] / (a[:-1] * a[1:]))
    jac = np.diag(off, 1) + np.diag(off, -1)
    vals, vecs = np.linalg.eigh(jac)
    w = 2.0 * vecs[0, :] ** 2
    rootb = (vals + 1.0) * math.pi / 2.0
    weightb = w * math.pi / 2.0 * np.sin(rootb)
    return rootb, weightb


def build_rot_mat(a: float, b: float, y: float, nao: int) -> np.ndarray:
    """BuildRotMat (SpinSymm.py:70-89): spin rotation in the [alpha..., beta...]
    spin-orbital basis: diag(e^{ia/2}, e^{-ia/2}) . [[cb, sb], [-sb, cb]] .
    diag(e^{iy/2}, e^{-iy/2}) on every spatial orbital."""
    ca, sa = math.cos(a / 2), math.sin(a / 2)
    cb, sb = math.cos(b / 2), math.sin(b / 2)
    cy, sy = math.cos(y / 2), math.sin(y / 2)
    ea = np.array([complex(ca, sa), complex(ca, -sa)])
    ey = np.array([complex(cy, sy), complex(cy, -sy)])
    blk = np.diag(ea) @ np.array([[cb, sb], [-sb, cb]], dtype=complex) @ np.diag(ey)
    return np.kron(blk, np.eye(nao))


def wigner_d(j: int, n: int, m: int, alpha: float, beta: float, gamma: float) -> complex:
    """Wigner D^j_{nm}(alpha, beta, gamma) = e^{-i n alpha} d^j_{nm}(beta) e^{-i m gamma}
    with the explicit factorial sum of WignerMat.py:8-27 / PHFTools.f90:350-376."""
    pre = math.sqrt(math.factorial(j + n) * math.factorial(j - n)
                    * math.factorial(j + m) * math.factorial(j - m))
    smax = min(j + m, j - n)
    smin = max(0, m - n)
    acc = 0.0
    cb, sb = math.cos(beta / 2), math.sin(beta / 2)
    for s in range(smin, smax + 1):
        p1 = n - m + s
        p2 = 2 * j + m - n - 2 * s
        p3 = n - m + 2 * s
        num = (-1) ** p1 * cb ** p2 * sb ** p3
        den = (math.factorial(j + m - s) * math.factorial(s)
               * math.factorial(n - m + s) * math.factorial(j - n - s))
        acc += num / den
    return complex(np.exp(-1j * n * alpha) * (pre * acc) * np.exp(-1j * m * gamma))


@dataclass
class SpinGrid:
    """Everything ``buildhsg`` takes about the spin grid (Spin_Proj.py:3-39)."""
    sp: int
    ngrid: np.ndarray          # int[2] = (nl, ny)
    nci: int                   # 2J+1 for SP=1, else 1
    roota: np.ndarray
    rootb: np.ndarray
    rooty: np.ndarray
    weightsp: np.ndarray       # float64 [nl, ny]
    r1: np.ndarray             # complex128 [nl, NSO, NSO]
    r2: np.ndarray             # complex128 [ny, NSO, NSO]


def spin_proj(sp: int, ngrid, nao: int, j: int = 0) -> SpinGrid:
    """Spin_Proj (Spin_Proj.py:3-39) for SP in {0, 1, 2, 3}."""
    nso = 2 * nao
    ngrid = np.asarray(ngrid, dtype=np.int64).copy()
    if sp == 0:
        eye = np.eye(nso, dtype=complex)[None]
        z = np.zeros(1)
        return SpinGrid(0, np.array([1, 1]), 1, z, z.copy(), z.copy(), np.ones((1, 1)), eye, eye.copy())
    if sp == 1:
        roota, rootb, rooty, w = root_weight(ngrid)
        r1 = np.stack([build_rot_mat(roota[i], rootb[i], 0.0, nao) for i in range(ngrid[0])])
        r2 = np.stack([build_rot_mat(0.0, 0.0, rooty[i], nao) for i in range(ngrid[1])])
        return SpinGrid(1, ngrid, 2 * j + 1, roota, rootb, rooty, w, r1, r2)
    if sp == 2:
        rootb, wb = root_weight_sy(ngrid[1])
        r1 = np.eye(nso, dtype=complex)[None]
        r2 = np.stack([build_rot_mat(0.0, rootb[i], 0.0, nao) for i in range(ngrid[1])])
        return SpinGrid(2, ngrid, 1, np.zeros(1), rootb, np.zeros(1), wb[None, :].copy(), r1, r2)
    if sp == 3:
        # Spin_Proj.py:31-38 keeps a [nl, ...] R1 with only R1[0]=I filled in
        roota, rootb, rooty, w = root_weight(ngrid)
        r1 = np.zeros((ngrid[0], nso, nso), dtype=complex)
        r1[0] = np.eye(nso)
        r2 = np.stack([build_rot_mat(0.0, 0.0, rooty[i], nao) for i in range(ngrid[1])])
        return SpinGrid(3, ngrid, 1, roota, rootb, rooty, w, r1, r2)
    raise ValueError(f"unsupported SP={sp}")


def grid_args(nao: int, ngrid=(14, 8), j: int = 0, sp: int = 1, rpg=None, weightpg=None) -> dict:
    """The grid arguments of phftools.buildhsg for a spin grid (Spin_Proj.py:3-39),
    as keywords: ngrid, nci, ncisp, r1, r2, rpg, weightsp, weightpg, roota,
    rootb, rooty, qj, sp.  Without rpg/weightpg there is no point group (one
    identity Rpg, weight 1); nci = ncisp * ncipg."""
    g = spin_proj(sp, ngrid, nao, j)
    nso = 2 * nao
    if rpg is None:
        rpg = np.eye(nso, dtype=complex)[None]
    if weightpg is None:
        weightpg = np.ones((rpg.shape[0], 1, 1))
    ncipg = weightpg.shape[1]
    return dict(ngrid=np.array(g.ngrid), nci=g.nci * ncipg, ncisp=g.nci, r1=g.r1, r2=g.r2, rpg=rpg,
                weightsp=g.weightsp, weightpg=weightpg, roota=g.roota, rootb=g.rootb, rooty=g.rooty,
                qj=j, sp=sp)
