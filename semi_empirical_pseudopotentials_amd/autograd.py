"""Differentiable filter diagonalization.

Forward: per-species pseudopotential parameters (A, sigma, B, tau) of
v(r) = A exp(-r^2/sigma^2) + B exp(-r^2/tau^2) and atom positions
-> tabulated, shifted radial tables (filter/read.c:43-122, shift at :106)
-> V(r) (sep_build_potential, filter/init.c:161-182)
-> spectral bound, Newton filter, ortho, Rayleigh-Ritz, sigma
   (sep_energy_range, sep_filter, sep_ortho, sep_rayleigh_ritz, sep_sigma;
   filter/main.c:67-142)
-> Ritz values E_n.

Backward: one sep_hf_adjoint call on the kept Ritz vectors, which stay
resident on the device after the forward (nothing of the filter itself is
stored: a new forward recomputes it).  For converged eigenpairs
dE_n/dp = <psi_n|dV/dp|psi_n> (Hellmann-Feynman); states whose energy
variance sigma_n exceeds the gate get no gradient (with a warning), since
the identity does not hold for them.  Degenerate levels: only symmetric
functions of a whole degenerate group (e.g. its sum) have a derivative.

The reference has no derivative path; its nearest relative is the
non-adiabatic coupling program (nonAdiabaticCoupling/nonadiabatic.c:9-301),
which evaluates <psi_m|dV/dR_a|psi_n> on the CPU.  Positions are used as
given (the reference recentres conf.par on the mean, filter/read.c:23-30).
"""
import warnings

import numpy as np
import torch

from ._lib import (SEP_INTERP_CUBIC, SEP_INTERP_LINEAR, SEP_PREC_FP64, Context, init_psi,
                   leja_points, newton_coeffs)


def radial_tables(params, rmax=12.0, dr=0.01):
    """(vr, pot, npot, drs) for params (nspecies x 4); pot shifted to 0 at rmax."""
    params = np.asarray(params, dtype=np.float64).reshape(-1, 4)
    npot = int(round(rmax / dr)) + 1
    r = np.arange(npot) * dr
    ns = params.shape[0]
    vr = np.tile(r, (ns, 1))
    pot = np.empty((ns, npot))
    for s, (A, sg, B, tau) in enumerate(params):
        v = A * np.exp(-r ** 2 / sg ** 2) + B * np.exp(-r ** 2 / tau ** 2)
        pot[s] = v - v[-1]
    return vr, pot, np.full(ns, npot, dtype=np.int64), np.full(ns, dr)


def radial_jacobian(params, rmax=12.0, dr=0.01):
    """d pot[s][k] / d (A, sigma, B, tau)_s as (nspecies, 4, npot)."""
    params = np.asarray(params, dtype=np.float64).reshape(-1, 4)
    npot = int(round(rmax / dr)) + 1
    r = np.arange(npot) * dr
    J = np.empty((params.shape[0], 4, npot))
    for s, (A, sg, B, tau) in enumerate(params):
        e1 = np.exp(-r ** 2 / sg ** 2)
        e2 = np.exp(-r ** 2 / tau ** 2)
        cols = [e1, A * e1 * 2 * r ** 2 / sg ** 3, e2, B * e2 * 2 * r ** 2 / tau ** 3]
        for q, f in enumerate(cols):
            J[s, q] = f - f[-1]
    return J


class Problem:
    """One nanocrystal box and filter setup, holding the libsep context.

    species_of_atom: species index per atom (rows of params); grid: (nx, ny,
    nz); box: (dx, dy, dz, xmin, ymin, zmin) in bohr; targets: energies in Ha.
    interp: "linear" builds V(r) with the reference's piecewise-linear table
    interpolation (filter/interpolate.c:11-23), whose kinks at every table
    row make E piecewise smooth in the atom positions; "cubic" uses the C^1
    cubic Hermite interpolant of the same tables (SEP_INTERP_CUBIC), which
    makes the position derivatives agree with finite differences.
    """

    def __init__(self, species_of_atom, grid, box, nc, ns, targets, seed=-1001,
                 rmax=12.0, dr=0.01, sigma_gate=1e-6, cutoff=1e-10, interp="linear"):
        self.types = np.ascontiguousarray(species_of_atom, dtype=np.int32)
        self.grid = tuple(int(g) for g in grid)
        self.box = tuple(float(b) for b in box)
        self.nc, self.ns = int(nc), int(ns)
        self.targets = np.asarray(targets, dtype=np.float64)
        self.seed, self.rmax, self.dr = int(seed), float(rmax), float(dr)
        self.sigma_gate, self.cutoff = float(sigma_gate), float(cutoff)
        nx, ny, nz = self.grid
        dx, dy, dz, x0, y0, z0 = self.box
        self.ctx = Context(nx, ny, nz, dx, dy, dz, x0, y0, z0)
        if interp not in ("linear", "cubic"):
            raise ValueError(f"interp must be 'linear' or 'cubic', not {interp!r}")
        self.ctx.set_interp(SEP_INTERP_CUBIC if interp == "cubic" else SEP_INTERP_LINEAR)
        self.dv = dx * dy * dz
        self.generation = 0
        self._psi0 = None

    def close(self):
        self.ctx.close()

    def _random_states(self):
        """Seeded random states (filter/init.c:236-251), kept on the device."""
        if self._psi0 is None:
            N = int(np.prod(self.grid))
            psi0 = np.empty((self.ns, N))
            idum, rng = self.seed, None
            for j in range(self.ns):
                psi0[j], idum, rng = init_psi(idum, N, self.dv, rng)
            self._psi0 = torch.from_numpy(psi0).to(f"cuda:{self.ctx.device}")
        return self._psi0

    def solve(self, params, xyz):
        """Ritz values and sigma of the filtered subspace; Ritz vectors stay resident."""
        vr, pot, npot, drs = radial_tables(params, self.rmax, self.dr)
        self.ctx.build_potential(xyz, self.types, npot, vr, pot, drs)
        vmin = self.ctx.vmin
        dE, _, _ = self.ctx.energy_range()
        zn, _ = leja_points(self.nc)
        an = newton_coeffs(zn, self.targets, dE, vmin, (self.nc / (2.5 * dE)) ** 2)
        self.ctx.filter_dev(self._random_states().data_ptr(), self.ns, an, zn, vmin, dE,
                            SEP_PREC_FP64)
        kept, _ = self.ctx.ortho(self.ns * self.targets.size, self.cutoff)
        ev = self.ctx.rayleigh_ritz(kept)
        sig = self.ctx.sigma(kept)
        self.generation += 1
        return ev, sig

    def adjoint(self, params, xyz, weights, sig):
        """(dL/dparams [nspecies, 4], dL/dxyz [natom, 3]) for dL/dE_n = weights."""
        w = np.array(weights, dtype=np.float64)
        bad = (sig >= self.sigma_gate) & (w != 0.0)
        if bad.any():
            warnings.warn(f"{int(bad.sum())} Ritz values with sigma >= {self.sigma_gate:g} get no "
                          "gradient (Hellmann-Feynman needs converged states)")
            w[bad] = 0.0
        vr, pot, npot, drs = radial_tables(params, self.rmax, self.dr)
        pairs = np.repeat(np.arange(w.size, dtype=np.int64)[:, None], 2, axis=1)
        dR, _, dtab = self.ctx.hf_adjoint(pairs, xyz, self.types, npot, vr, pot, drs, weights=w)
        J = radial_jacobian(params, self.rmax, self.dr)
        dparams = np.einsum("sk,sqk->sq", dtab[0][:, :J.shape[2]], J)
        return dparams, dR[0]


class FilterEigen(torch.autograd.Function):
    """(params [nspecies, 4], xyz [natom, 3]) -> (Ritz values, sigma)."""

    @staticmethod
    def forward(ctx, params, xyz, problem):
        p = params.detach().cpu().double().numpy()
        x = xyz.detach().cpu().double().numpy()
        ev, sig = problem.solve(p, x)
        ctx.problem, ctx.p, ctx.x, ctx.sig = problem, p, x, sig
        ctx.generation = problem.generation
        out_sig = torch.from_numpy(sig)
        ctx.mark_non_differentiable(out_sig)
        return torch.from_numpy(ev), out_sig

    @staticmethod
    def backward(ctx, g_ev, g_sig):
        pr = ctx.problem
        if pr.generation != ctx.generation:
            raise RuntimeError("the Ritz vectors of this forward were replaced by a later forward "
                               "on the same Problem; run backward before the next forward")
        dparams, dxyz = pr.adjoint(ctx.p, ctx.x, g_ev.detach().cpu().double().numpy(), ctx.sig)
        return torch.from_numpy(dparams), torch.from_numpy(np.ascontiguousarray(dxyz)), None


def ritz_values(params, xyz, problem):
    """Differentiable Ritz values (and their sigma) of the filter diagonalization."""
    return FilterEigen.apply(params, xyz, problem)
