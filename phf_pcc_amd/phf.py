"""Host numerics of the projected-HF objective around the grid kernel.

  eigen_solver  PHFTools.py:8-18   canonical orthogonalisation, |eig(S)| > 1e-6
  vec_decomp    PHFTools.py:20-26  CI vector -> (fsp, fpg, fk)
  sz_proj       PHF.py:133-137     Z mask for SP=2 (no spin-flip blocks)
  rhf_proj      PHF.py:139-143     Z mask for restricted orbitals
  frozen_proj   PHF.py:146-151     Z mask for frozen core orbitals
  make_eandg    PHF.py:55-91       the objective x -> (E, gradient)

These are the caller side of phftools.buildhsg/localg: O(nci^3) host work on
the small projected matrices.  The grid kernel itself always runs on the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

TOL = 1e-6


def eigen_solver(h: np.ndarray, s: np.ndarray, tol: float = TOL):
    """Lowest states of H c = E S c by canonical orthogonalisation (PHFTools.py:8-18).

    The column selection is the reference's: the inverse square roots of the
    eigenvalues of S with |lambda| > tol, in eigh's ascending order, scale the
    LAST dim eigenvectors, dim being how many eigenvalues pass.  When S has an
    eigenvalue below -tol this pairs the columns with other eigenvalues than a
    |lambda| > tol mask would (e.g. eigenvalues (-0.3, 1e-8, 2.0): the
    reference keeps the eigenvectors of 1e-8 and 2.0, scaled by 0.3**-0.5 and
    2.0**-0.5); it is kept exactly so that E0 and the CI vector agree with
    the reference on indefinite S."""
    sval, svec = np.linalg.eigh(s)
    inv_s = np.array([1.0 / np.sqrt(abs(v)) for v in sval if abs(v) > tol])
    dim = inv_s.size
    x = svec[:, sval.size - dim:] * inv_s[None, :]
    hh = x.conj().T @ h @ x
    ev, vec = np.linalg.eigh(hh)
    return ev, x @ vec


def vec_decomp(vec: np.ndarray, ncisp: int, ncipg: int, ncik: int):
    """Split the CI vector into spin, point-group and complex-conjugation parts."""
    m = vec.reshape((ncisp * ncipg, ncik), order="F")
    fk = m[0, :] / m[0, 0]
    m = m[:, 0].reshape((ncisp, ncipg), order="F")
    fpg = m[0, :] / m[0, 0]
    fsp = m[:, 0]
    return fsp, fpg, fk


def sz_proj(z: np.ndarray, nvrt_a: int, nocc_a: int) -> np.ndarray:
    out = np.zeros_like(z)
    out[:nvrt_a, :nocc_a] = z[:nvrt_a, :nocc_a]
    out[nvrt_a:, nocc_a:] = z[nvrt_a:, nocc_a:]
    return out


def rhf_proj(z: np.ndarray, nvrt_a: int, nocc_a: int) -> np.ndarray:
    out = np.zeros_like(z)
    out[:nvrt_a, :nocc_a] = z[:nvrt_a, :nocc_a]
    out[nvrt_a:, nocc_a:] = z[:nvrt_a, :nocc_a]
    return out


def frozen_proj(z: np.ndarray, nfo: int, nocc_ao: int) -> np.ndarray:
    out = np.array(z, copy=True)
    out[:, :nfo] = 0
    out[:, nocc_ao:nocc_ao + nfo] = 0
    return out


@dataclass
class ObjectiveConfig:
    """The module-level switches EandG reads from parameter_fc.py / PHF.py."""
    nso: int
    nocc: int
    cmplxconj: int = 1
    sp: int = 1
    is_rhf: bool = False
    nfo: int = 0
    ncisp: int = 1
    ncipg: int = 1

    @property
    def nvrt(self):
        return self.nso - self.nocc

    def masks(self, z):
        nao = self.nso // 2
        nocc_a = self.nocc // 2
        nvrt_a = nao - nocc_a
        if self.nfo > 0:
            z = frozen_proj(z, self.nfo, nocc_a)
        if self.cmplxconj == 0:
            z = z.real.astype(complex)
        if self.sp == 2:
            z = sz_proj(z, nvrt_a, nocc_a)
        if self.is_rhf:
            z = rhf_proj(z, nvrt_a, nocc_a)
        return z


def make_eandg(cfg: ObjectiveConfig, buildhsg, localg):
    """EandG (PHF.py:55-91) over any buildhsg(z) -> (H, S, Grad, Rdm) and the
    phftools.localg calling convention.  Returns f(x) -> (E, [Re G, Im G]) with
    G the (NSO-N, N) block of LocalG, exactly as the reference returns it (half
    the gradient with respect to x = [Re Z, Im Z])."""
    def eandg(x):
        nv, no = cfg.nvrt, cfg.nocc
        z = (x[:nv * no] + 1j * x[nv * no:]).reshape(nv, no)
        z = cfg.masks(z)
        h, s, gm, rm = buildhsg(z)
        ev, vec = eigen_solver(h, s)
        ncik = 2 if cfg.cmplxconj == 2 else 1
        fsp, fpg, fk = vec_decomp(vec[:, 0], cfg.ncisp, cfg.ncipg, ncik)
        g = localg(gm, rm, s, ev[0], fsp, fpg, fk, cfg.cmplxconj)[no:, :no]
        g = cfg.masks(g).reshape(nv * no)
        return float(ev[0].real), np.concatenate((g.real, g.imag))
    return eandg
