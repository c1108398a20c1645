"""Device-resident projected-HF objective and its differentiable form.

``ProjectedHF`` keeps the Hamiltonian (dense drop-in arrays or the factored
form built on the device) and the spin grid on one GPU; each evaluation takes
a Thouless Z and returns the projected energy and dE/dZ* (the (NSO-N, N)
block of LocalG, PHFTools.f90:203-247) from one fused pass of the grid kernel.

``ProjectedEnergy`` is the torch.autograd.Function of E(Re Z, Im Z).  Its
gradient is the hand-written adjoint of the grid kernel: the epilogue that
forms Y_g = (h + Gamma_g) X_g - X_g C0^H (h + Gamma_g) X_g + e_g X_g and the
Wigner-weighted grid sum of d_g Y_g.  Per-grid-point densities, Fock matrices
and Y_g are transient device buffers of the one pass; only the (NSO-N) x N
gradient is kept for backward.  With x = [Re Z, Im Z] the gradient is
dE/dx = 2 [Re G, Im G] (the reference's EandG returns half of it, PHF.py:91).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .phf import ObjectiveConfig, eigen_solver, vec_decomp


class ProjectedHF:
    def __init__(self, nocc: int, ref: np.ndarray, grid: dict, hone=None, htwo=None, factors=None,
                 precision: str = "fp64", device: int = 0, cmplxconj: int = 1, ctx: _lib.Context | None = None,
                 structure: bool = True, is_rhf: bool = False, nfo: int = 0):
        """grid: keys ngrid, nci, ncisp, r1, r2, rpg, weightsp, weightpg, roota, rootb, rooty, qj, sp
        (the buildhsg grid arguments).  Either hone/htwo (MO basis, as the
        drop-in takes them) or factors=(h, b) with (pq|rs) = sum_L b[L,p,q] b[L,r,s].
        structure: use the spin-block structure of A when Ref has spin-pure MOs
        (False: always contract the dense A, the plain form).
        is_rhf, nfo: the RHFProj and FrozenProj masks of EandG (PHF.py:62-68,
        78-85); with the grid's sp == 2 SzProj and with cmplxconj == 0 the
        real-Z mask apply too, to Z before the evaluation and to G after it."""
        self.ctx = ctx or _lib.Context(device)
        self.nocc = int(nocc)
        self.cmplxconj = int(cmplxconj)
        self.precision = precision
        ref = np.ascontiguousarray(ref, dtype=np.complex128)
        if factors is not None:
            self.ctx.set_hamiltonian_factored(factors[0], factors[1], ref,
                                              flags=0 if structure else _lib.PHF_HAM_DENSE)
        else:
            self.ctx.set_hamiltonian(hone, htwo, dense=not structure)
        self.nso = ref.shape[0]
        self.grid = grid
        self.masks = ObjectiveConfig(nso=self.nso, nocc=self.nocc, cmplxconj=self.cmplxconj, sp=int(grid["sp"]),
                                     is_rhf=bool(is_rhf), nfo=int(nfo)).masks
        self._set_grid(ref, grid)
        self.nevals = 0

    def _set_grid(self, ref, g):
        lib = _lib.load()
        c128 = lambda a: np.ascontiguousarray(a, dtype=np.complex128)
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(-1) if np.ndim(a) <= 1 else \
            np.ascontiguousarray(a, dtype=np.float64)
        self._keep = dict(ref=ref, ngrid=np.ascontiguousarray(g["ngrid"], dtype=np.int32), r1=c128(g["r1"]),
                          r2=c128(g["r2"]), rpg=c128(g["rpg"]), wsp=f64(g["weightsp"]), wpg=f64(g["weightpg"]),
                          ra=f64(g["roota"]), rb=f64(g["rootb"]), ry=f64(g["rooty"]))
        k = self._keep
        self.nci = int(g["nci"])
        self.ncisp = int(g["ncisp"])
        self.ncipg = int(k["wpg"].shape[1])
        _lib.check(lib.phf_set_grid(self.ctx.handle, _lib.ptr(k["ref"]), k["ngrid"].ctypes.data_as(_lib._ip),
                                    self.nci, self.ncisp, _lib.ptr(k["r1"]), k["r1"].shape[0], _lib.ptr(k["r2"]),
                                    k["r2"].shape[0], _lib.ptr(k["rpg"]), k["rpg"].shape[0], _lib.ptr(k["wsp"]),
                                    _lib.ptr(k["wpg"]), self.ncipg, _lib.ptr(k["ra"]), k["ra"].size,
                                    _lib.ptr(k["rb"]), k["rb"].size, _lib.ptr(k["ry"]), k["ry"].size,
                                    int(g["qj"]), int(g["sp"])), self.ctx.handle)

    # ------------------------------------------------------------------
    def buildhsg(self, z: np.ndarray):
        z = np.ascontiguousarray(z, dtype=np.complex128)
        if z.shape != (self.nso - self.nocc, self.nocc):
            raise ValueError(f"z must be ({self.nso - self.nocc}, {self.nocc})")
        nci, nso = self.nci, self.nso
        h = np.zeros((nci, nci), complex)
        s = np.zeros((nci, nci), complex)
        gm = np.zeros((nci, nci, nso, nso), complex)
        rm = np.zeros((nci, nci, nso, nso), complex)
        rc = _lib.load().phf_buildhsg_z(self.ctx.handle, _lib.ptr(z), self.cmplxconj, self.nocc,
                                        _lib.PREC[self.precision], _lib.ptr(h), _lib.ptr(s), _lib.ptr(gm),
                                        _lib.ptr(rm))
        _lib.check(rc, self.ctx.handle)
        self.nevals += 1
        return h, s, gm, rm

    def localg(self, gm, rm, s, e0, fsp, fpg, fk):
        nso, nci = self.nso, self.nci
        g = np.zeros((nso, nso), complex)
        e0a = np.array([complex(e0)])
        c = lambda a: np.ascontiguousarray(a, dtype=np.complex128).reshape(-1)
        fsp, fpg, fk = c(fsp), c(fpg), c(fk)
        rc = _lib.load().phf_localg(self.ctx.handle, _lib.ptr(gm), _lib.ptr(rm), _lib.ptr(s), _lib.ptr(e0a),
                                    _lib.ptr(fsp), fsp.size, _lib.ptr(fpg), fpg.size, _lib.ptr(fk), fk.size,
                                    self.cmplxconj, nso, nci, _lib.ptr(g))
        _lib.check(rc, self.ctx.handle)
        return g

    def energy_and_gradient(self, z: np.ndarray):
        """Projected energy E0 and G = dE/dZ* (shape (NSO-N, N)) as EandG forms
        them (PHF.py:55-91): the Z masks are applied to Z before the grid pass
        and to G after it (with cmplxconj == 0, G's imaginary part is zero,
        PHF.py:80-81).  The grid sums stay on the device; only Hmat/Smat
        (nci x nci) and G cross PCIe."""
        z = np.ascontiguousarray(z, dtype=np.complex128)
        if z.shape != (self.nso - self.nocc, self.nocc):
            raise ValueError(f"z must be ({self.nso - self.nocc}, {self.nocc})")
        z = np.ascontiguousarray(self.masks(z))
        e0, g = self._energy_and_gradient(z)
        return e0, np.ascontiguousarray(self.masks(g))

    def _energy_and_gradient(self, z: np.ndarray):
        if self.cmplxconj == 2:
            h, s, gm, rm = self.buildhsg(z)
            ev, vec = eigen_solver(h, s)
            fsp, fpg, fk = vec_decomp(vec[:, 0], self.ncisp, self.ncipg, 2)
            return float(ev[0].real), self.localg(gm, rm, s, ev[0], fsp, fpg, fk)[self.nocc:, :self.nocc]
        lib = _lib.load()
        nci = self.nci
        if nci == 1:
            # one call: grid sums, the 1 x 1 EigenSolver and LocalG on the device
            e0 = np.zeros(1)
            g = np.zeros((self.nso - self.nocc, self.nocc), complex)
            _lib.check(lib.phf_eval_objective(self.ctx.handle, _lib.ptr(z), self.cmplxconj, self.nocc,
                                              _lib.PREC[self.precision], _lib.ptr(e0), _lib.ptr(g)), self.ctx.handle)
            self.nevals += 1
            return float(e0[0]), g
        h = np.zeros((nci, nci), complex)
        s = np.zeros((nci, nci), complex)
        _lib.check(lib.phf_eval_resident(self.ctx.handle, _lib.ptr(z), self.cmplxconj, self.nocc,
                                         _lib.PREC[self.precision], _lib.ptr(h), _lib.ptr(s)), self.ctx.handle)
        self.nevals += 1
        ev, vec = eigen_solver(h, s)
        fsp, fpg, fk = vec_decomp(vec[:, 0], self.ncisp, self.ncipg, 1)
        c = lambda a: np.ascontiguousarray(a, dtype=np.complex128).reshape(-1)
        fsp, fpg, fk = c(fsp), c(fpg), c(fk)
        e0a = np.array([complex(ev[0])])
        g = np.zeros((self.nso - self.nocc, self.nocc), complex)
        _lib.check(lib.phf_localg_vo(self.ctx.handle, _lib.ptr(s), _lib.ptr(e0a), _lib.ptr(fsp), fsp.size,
                                     _lib.ptr(fpg), fpg.size, _lib.ptr(fk), fk.size, self.cmplxconj, _lib.ptr(g)),
                   self.ctx.handle)
        return float(ev[0].real), g


def _torch():
    import torch
    return torch


class ProjectedEnergy:
    """E = ProjectedEnergy.apply(z_re, z_im, model) as a torch.autograd.Function."""
    _fn = None

    @classmethod
    def apply(cls, z_re, z_im, model: ProjectedHF):
        if cls._fn is None:
            torch = _torch()

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, zr, zi, mdl):
                    z = zr.detach().cpu().numpy().astype(np.float64) + 1j * zi.detach().cpu().numpy()
                    e, g = mdl.energy_and_gradient(z)
                    ctx.save_for_backward(torch.as_tensor(2.0 * g.real, dtype=zr.dtype, device=zr.device),
                                          torch.as_tensor(2.0 * g.imag, dtype=zr.dtype, device=zr.device))
                    return torch.tensor(e, dtype=zr.dtype, device=zr.device)

                @staticmethod
                def backward(ctx, ge):
                    gr, gi = ctx.saved_tensors
                    return ge * gr, ge * gi, None

            cls._fn = _Fn
        return cls._fn.apply(z_re, z_im, model)


def lbfgs_optimize(model: ProjectedHF, z0: np.ndarray, steps: int = 50, lr: float = 1.0, history: int = 20,
                   callback=None):
    """Orbital optimisation of Z with torch.optim.LBFGS (strong-Wolfe line
    search) over E(Z); `steps` optimiser iterations.  Returns (z, energies)."""
    torch = _torch()
    zr = torch.tensor(np.ascontiguousarray(z0.real), dtype=torch.float64, requires_grad=True)
    zi = torch.tensor(np.ascontiguousarray(z0.imag), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.LBFGS([zr, zi], lr=lr, max_iter=steps, history_size=history, tolerance_grad=1e-9,
                            tolerance_change=1e-14, line_search_fn="strong_wolfe")
    energies = []

    def closure():
        opt.zero_grad()
        e = ProjectedEnergy.apply(zr, zi, model)
        e.backward()
        energies.append(float(e.detach()))
        if callback:
            callback(len(energies), float(e))
        return e

    opt.step(closure)
    z = zr.detach().numpy() + 1j * zi.detach().numpy()
    return z, energies
