"""Synthetic input decks in the reference `filter` program's file formats.

A zincblende CdSe sphere (Cd on the fcc sublattice, Se at +a(1/4,1/4,1/4)),
surface bonds passivated by pseudo-ligand sites P1 (on Cd) / P2 (on Se) at a
fraction of the bond length, and radial tables v(r) = A exp(-r^2/s^2) +
B exp(-r^2/t^2) tabulated like the reference's pot*.par files
(filter/read.c:43-122: rows "r v", shifted later by the reader).
The reference ships no structure builder (its common/ has none), so this is
the project's own generator; positions are in bohr as conf.par expects.
"""
import os

import numpy as np

BOHR_PER_ANGSTROM = 1.0 / 0.529177210903

# per-species (A, sigma, B, tau) in Ha / bohr
DEFAULT_POTS = {
    "Cd": (-1.2, 1.4, 0.3, 3.0),
    "Se": (-0.9, 1.8, 0.2, 3.5),
    "P1": (-0.5, 1.0, 0.0, 1.0),
    "P2": (-0.5, 1.0, 0.0, 1.0),
}

HARTREE_EV = 27.211386245988

# Energy window of the synthetic decks.  With the tables above the dot binds
# no state: both species' radial potentials integrate to a positive value
# (Cd: -1.2 (pi 1.4^2)^1.5 + 0.3 (pi 3.0^2)^1.5 = +26.8 Ha bohr^3, Se: +18.5),
# so the lowest levels (E > 0, the lowest at 0.004-0.007 Ha for 1.5-4 nm dots)
# live in the box around the dot and the spectrum has no gap to centre the
# +-2 eV window on.  It is centred just below the lower edge of the spectrum
# instead, so that the targets pull out the lowest levels; "band-edge"
# comparisons use the levels within 0.5 eV of the lowest converged one.
WINDOW_CENTER_HA = -0.03
WINDOW_HALF_HA = 2.0 / HARTREE_EV
EDGE_BAND_HA = 0.5 / HARTREE_EV


def cdse_dot(diameter_nm, lattice_A=6.08, ligand_frac=0.55, center=(0.0, 0.0, 0.0)):
    """Atoms of a zincblende CdSe sphere; returns (symbols, xyz[bohr])."""
    a = lattice_A * BOHR_PER_ANGSTROM
    R = 0.5 * diameter_nm * 10.0 * BOHR_PER_ANGSTROM
    bond = a * np.sqrt(3.0) / 4.0
    fcc = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]]) * a
    shift = np.array([.25, .25, .25]) * a
    nmax = int(np.ceil(R / a)) + 2
    c = np.asarray(center, dtype=float)
    sites = []
    for i in range(-nmax, nmax + 1):
        for j in range(-nmax, nmax + 1):
            for k in range(-nmax, nmax + 1):
                o = np.array([i, j, k]) * a
                for f in fcc:
                    for sym, off in (("Cd", 0.0), ("Se", shift)):
                        p = o + f + off
                        if np.linalg.norm(p - c) <= R:
                            sites.append((sym, p))
    # tetrahedral bond directions of a cation; anions use the opposite ones
    dirs = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / np.sqrt(3.0)
    syms = [s for s, _ in sites]
    pos = np.array([p for _, p in sites])
    key = {tuple(np.round(p / (0.125 * a)).astype(int)): n for n, p in enumerate(pos)}

    def neighbours(n):
        sgn = 1.0 if syms[n] == "Cd" else -1.0
        out = []
        for d in dirs:
            q = pos[n] + sgn * d * bond
            out.append((d * sgn, key.get(tuple(np.round(q / (0.125 * a)).astype(int)))))
        return out

    # drop atoms bonded to fewer than two others (repeat until stable)
    alive = np.ones(len(pos), dtype=bool)
    changed = True
    while changed:
        changed = False
        for n in range(len(pos)):
            if alive[n] and sum(1 for _, m in neighbours(n) if m is not None and alive[m]) < 2:
                alive[n] = False
                changed = True
    out_sym, out_xyz = [], []
    lig_sym, lig_xyz = [], []
    for n in range(len(pos)):
        if not alive[n]:
            continue
        out_sym.append(syms[n])
        out_xyz.append(pos[n])
        for d, m in neighbours(n):
            if m is None or not alive[m]:
                lig_sym.append("P1" if syms[n] == "Cd" else "P2")
                lig_xyz.append(pos[n] + d * bond * ligand_frac)
    return out_sym + lig_sym, np.array(out_xyz + lig_xyz)


def pot_table(A, s, B, t, rmax=12.0, dr=0.01):
    r = np.arange(0.0, rmax + 0.5 * dr, dr)
    v = A * np.exp(-r * r / (s * s)) + (B * np.exp(-r * r / (t * t)) if B != 0.0 else 0.0)
    return r, v


def write_deck(path, syms, xyz, nx, ny, nz, ms, ns, nc, elmin, elmax, nthreads=1, pots=None,
               seed=None):
    """Write input.par, conf.par, pot<Sym>.par (+ seed.par) into `path`."""
    os.makedirs(path, exist_ok=True)
    pots = DEFAULT_POTS if pots is None else pots
    with open(os.path.join(path, "input.par"), "w") as f:
        f.write(f"{nx}\n{ny}\n{nz}\n{ms}\n{ns}\n{nc}\n{elmin}\n{elmax}\n{nthreads}\n")
    with open(os.path.join(path, "conf.par"), "w") as f:
        f.write(f"{len(syms)}\n")
        for s, p in zip(syms, xyz):
            f.write(f"{s} {p[0]:.8f} {p[1]:.8f} {p[2]:.8f}\n")
    for sym, prm in pots.items():
        r, v = pot_table(*prm)
        with open(os.path.join(path, f"pot{sym}.par"), "w") as f:
            for ri, vi in zip(r, v):
                f.write(f"{ri:.4f} {vi:.10e}\n")
    if seed is not None:
        with open(os.path.join(path, "seed.par"), "w") as f:
            f.write(f"{seed}\n")
    return path


def write_spinor_deck(path, syms, xyz, nx, ny, nz, dx, ms, ns, nc, elmin, elmax, pots,
                      nthreads=1, flagkb=0, seed=None):
    """A deck of the spin-orbit filter (spinors/filterSpinors): input.par as
    its init_size reads it (nx ny nz dx ms ns nc Elmin Elmax nthreads flagkb,
    init.c:5-21), conf.par and pot<Sym>.par (+ seed.par)."""
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "input.par"), "w") as f:
        f.write(f"{nx}\n{ny}\n{nz}\n{dx}\n{ms}\n{ns}\n{nc}\n{elmin!r}\n{elmax!r}\n{nthreads}\n"
                f"{flagkb}\n")
    with open(os.path.join(path, "conf.par"), "w") as f:
        f.write(f"{len(syms)}\n")
        for s, p in zip(syms, xyz):
            f.write(f"{s} {p[0]:.8f} {p[1]:.8f} {p[2]:.8f}\n")
    for sym, prm in pots.items():
        r, v = pot_table(*prm)
        with open(os.path.join(path, f"pot{sym}.par"), "w") as f:
            for ri, vi in zip(r, v):
                f.write(f"{ri:.4f} {vi:.10e}\n")
    if seed is not None:
        with open(os.path.join(path, "seed.par"), "w") as f:
            f.write(f"{seed}\n")
    return path


def box_from_atoms(xyz, nx, ny, nz):
    """Box and spacing by the reference rule (filter/init.c:90-122)."""
    xyz = np.asarray(xyz, dtype=float)
    xyz = xyz - xyz.mean(axis=0)
    half = np.rint(0.5 * (xyz.max(axis=0) - xyz.min(axis=0)) + 5.0)
    n = np.array([nx, ny, nz], dtype=float)
    d = 2.0 * half / n
    return half, d
