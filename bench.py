#!/usr/bin/env python
"""H.psi applications/s of the filter-diagonalization hot loop on one MI355X.

Workload (the production configuration the metric is quoted on): a 4.0 nm
zincblende CdSe dot (1667 atoms incl. pseudo-ligands), 128^3 grid, 16 seeded
random states x 32 energy targets, Newton length 4096.  One step = one full
filter pass (filter/filter.c:59-83 for all 16 states: 16 x 4095 H.psi with
the Newton recurrence and the 32-target accumulation, then normalize_all)
with the random states resident in HBM.  The headline is the complex64
filter; the FP64 filter is timed the same way beside it.

value = states x (nc - 1) x steps / time, i.e. reference-equivalent H.psi
applications (each real state's application counts once; the kernels carry
two real states per complex vector because H is real).

Prints one JSON line (rank 0, N = 1).
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("H·ψ applications/s at 128³; FP32+refine vs FP64 GPU at "
          "≤1 meV error")
PEAK_HBM_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--grid", type=int, default=128)
    p.add_argument("--diameter", type=float, default=4.0)
    p.add_argument("--ns", type=int, default=16)
    p.add_argument("--ms", type=int, default=32)
    p.add_argument("--nc", type=int, default=4096)
    p.add_argument("--lanes", type=int, default=0,
                   help="filter lanes (sep_set_filter_lanes): 0 = automatic, 1 = one stream")
    p.add_argument("--no-fp64", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-steps", type=int, default=24,
                   help="Newton steps of the CPU baseline sample, all random states in "
                        "parallel (about 20 s on 16 threads)")
    p.add_argument("--pipeline", action=argparse.BooleanOptionalAction, default=True,
                   help="time the whole filter-diagonalization pipeline of both precisions")
    p.add_argument("--no-spinor", action="store_true",
                   help="skip the spin-orbit spinor filter workload")
    p.add_argument("--no-batched", action="store_true",
                   help="skip the batched uneven-grid measurement (8 dots, 64-160^3)")
    p.add_argument("--no-check", action="store_true",
                   help="skip the FP32-pipeline accuracy check (2 nm dot, 64^3)")
    p.add_argument("--gemm-only", action="store_true",
                   help="only time the subspace-stage matrix-core products (Gram and rotation, "
                        "FP64 and FP32, m = 512 states at the grid); for rocprofv3 --pmc runs")
    p.add_argument("--kernels-only", action="store_true",
                   help="only launch the H.psi passes (FP32, 5 times) and a 256 MiB device "
                        "copy that calibrates the HBM byte counters; for rocprofv3 --pmc runs")
    p.add_argument("--spinor-only", action="store_true",
                   help="only run the spin-orbit spinor filter workload (InAs 3.0 nm dot, "
                        "96^3); for rocprofv3 runs")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hpsi_fp32_128_traffic.md")


def measured_traffic(kernel, n, nvec):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel factor),
    or None when the summary does not cover this workload."""
    if n != 128 or nvec != 8 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch") and kernel in line:
            return int(line.split()[-1])
    return None


def subspace_gemms(sep, np, torch, n, m=512, reps=3):
    """The subspace stage's N x m products on the library's MFMA kernels
    (filter/Hmat.c:30-46 projected H / ortho.c:32 Gram, Hmat.c:61-78
    rotation) at m states of N = n^3 points: the Gram C = A B^T
    (sep_gram_dev) and the rotation Out = M^T A (sep_rotate_dev), FP64 and
    FP32 operands, 2 m^2 N flop each.  Times include the m x m transfer."""
    out = {"m": m, "N": n ** 3, "flop_per_product": 2.0 * m * m * n ** 3}
    with sep.Context(n, n, n, 0.5, 0.5, 0.5) as ctx:
        N = ctx.N
        g = torch.Generator(device="cuda:0").manual_seed(7)
        for name, dt, P in (("fp64", torch.float64, sep.SEP_PREC_FP64),
                            ("fp32", torch.float32, sep.SEP_PREC_FP32)):
            A = torch.randn(m, N, dtype=dt, device="cuda:0", generator=g)
            B = torch.randn(m, N, dtype=dt, device="cuda:0", generator=g)
            O = torch.empty(m, N, dtype=dt, device="cuda:0")
            M = np.random.default_rng(1).standard_normal((m, m))
            ctx.gram_dev(A.data_ptr(), m, B.data_ptr(), m, P)
            ctx.rotate_dev(A.data_ptr(), m, M, O.data_ptr(), P)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                ctx.gram_dev(A.data_ptr(), m, B.data_ptr(), m, P)
            tg = (time.perf_counter() - t0) / reps
            t0 = time.perf_counter()
            for _ in range(reps):
                ctx.rotate_dev(A.data_ptr(), m, M, O.data_ptr(), P)
            tr = (time.perf_counter() - t0) / reps
            fl = out["flop_per_product"]
            peak = 78.6 if name == "fp64" else 157.3
            out[name] = {"gram_ms": round(tg * 1e3, 3), "gram_TFLOPs": round(fl / tg / 1e12, 1),
                         "rotate_ms": round(tr * 1e3, 3),
                         "rotate_TFLOPs": round(fl / tr / 1e12, 1),
                         "peak_TFLOPs": peak,
                         "gram_frac": round(fl / tg / 1e12 / peak, 3),
                         "rotate_frac": round(fl / tr / 1e12 / peak, 3)}
            del A, B, O
    return out


def precision_check(sep, decks, read_deck, np):
    """FP32 pipeline vs FP64 pipeline on the 2.0 nm CdSe dot at 64^3 with 8
    targets: the error of every level within 0.5 eV of the lowest converged
    level.  16 random states and Newton length 2048 (4 and 1024 leave no
    level of this gapless spectrum converged, sigma > 1e-3, in either
    precision)."""
    from semi_empirical_pseudopotentials_amd import levels
    import tempfile
    n, nc, ms, ns = 64, 2048, 8, 16
    syms, xyz = decks.cdse_dot(2.0)
    tmp = tempfile.mkdtemp(prefix="sep_check_")
    c, h = decks.WINDOW_CENTER_HA, decks.WINDOW_HALF_HA
    decks.write_deck(tmp, syms, xyz, n, n, n, ms, ns, nc, c - h, c + h)
    d = read_deck(tmp)
    ctx = sep.Context(n, n, n, d["dx"], d["dy"], d["dz"], d["xmin"], d["ymin"], d["zmin"])
    ctx.build_potential(d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"],
                        expot=d["expot"])
    vmin = ctx.vmin
    dE, _, _ = ctx.energy_range()
    zn, _ = sep.leja_points(nc)
    an = sep.newton_coeffs(zn, np.linspace(d["Elmin"], d["Elmax"], ms), dE, vmin,
                           (nc / (2.5 * dE)) ** 2)
    N = n ** 3
    psi0 = np.empty((ns, N))
    idum, rng = -1001, None
    for j in range(ns):
        psi0[j], idum, rng = sep.init_psi(idum, N, d["dv"], rng)
    res = {}
    for name, prec in (("fp64", sep.SEP_PREC_FP64), ("fp32", sep.SEP_PREC_FP32)):
        t0 = time.perf_counter()
        ctx.filter(psi0, an, zn, vmin, dE, prec)
        if prec == sep.SEP_PREC_FP64:
            k, _ = ctx.ortho(ns * ms)
            ev = ctx.rayleigh_ritz(k)
            sg = ctx.sigma(k)
        else:
            run = levels.fp32_pipeline(ctx, ns * ms)
            k, ev, sg = run["kept"], run["ev_w"], run["sg_w"]
        res[name] = (ev, sg, k, time.perf_counter() - t0)
    ctx.close()
    return dict(config="CdSe 2.0 nm dot, 64^3, nc 2048, 8 targets x 16 states, seed -1001",
                kept_fp64=res["fp64"][2], kept_fp32=res["fp32"][2],
                pipeline_s={"fp64": round(res["fp64"][3], 3), "fp32": round(res["fp32"][3], 3)},
                **edge_gate(np, decks, levels, res["fp64"][0], res["fp64"][1], res["fp32"][0],
                            res["fp32"][1], conv=1e-3))


def physical_check(sep, decks, read_deck, np):
    """The 1 meV gate at a real band edge: the 4.0 nm CdSe dot at 128^3 on the
    reference's own tables (potCd.par, potSe.par of its In13_P16 test deck;
    analytic P1/P2 passivation), window [-0.26, -0.12] Ha around its gap, 16
    random states x 16 targets, nc 4096.  The FP32 pipeline finds the
    conduction edge from its own Ritz values (largest gap among pairs with
    sigma < 1e-2), refines the pairs from 0.1 eV below to 1.1 eV above it
    (10 FP64 passes, in place), and its levels within 1 eV of the edge are
    compared one to one with the FP64 pipeline's (sigma < 1e-4)."""
    from semi_empirical_pseudopotentials_amd import levels
    import shutil
    import tempfile
    ev_ha = 1.0 / 27.211386245988
    n, ns, ms, nc, lo, hi = 128, 16, 16, 4096, -0.26, -0.12
    tmp = tempfile.mkdtemp(prefix="sep_phys_")
    syms, xyz = decks.cdse_dot(4.0)
    decks.write_deck(tmp, syms, xyz, n, n, n, ms, ns, nc, lo, hi, pots={})
    for s in ("Cd", "Se"):
        shutil.copy(os.path.join(ROOT, "tests", "golden", "In13_P16", f"pot{s}.par"), tmp)
    d = read_deck(tmp)
    ctx = sep.Context(n, n, n, d["dx"], d["dy"], d["dz"], d["xmin"], d["ymin"], d["zmin"])
    ctx.build_potential(d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"],
                        expot=d["expot"])
    dE, _, _ = ctx.energy_range()
    zn, _ = sep.leja_points(nc)
    an = sep.newton_coeffs(zn, np.linspace(lo, hi, ms), dE, ctx.vmin, (nc / (2.5 * dE)) ** 2)
    N = n ** 3
    psi0 = np.empty((ns, N))
    idum, rng = -1001, None
    for j in range(ns):
        psi0[j], idum, rng = sep.init_psi(idum, N, d["dv"], rng)

    def edge(ev, sg):
        ok = np.sort(ev[(sg < 1e-2) & (ev >= lo) & (ev <= hi)])
        i = int(np.argmax(np.diff(ok)))
        return float(ok[i]), float(ok[i + 1])

    # an accuracy gate: not timed (the first filter allocates the ring, the
    # second reuses it)
    ctx.filter(psi0, an, zn, ctx.vmin, dE, sep.SEP_PREC_FP64, want_states=False)
    k64, _ = ctx.ortho(ns * ms)
    e64 = ctx.rayleigh_ritz(k64)
    s64 = ctx.sigma(k64)
    ctx.filter(psi0, an, zn, ctx.vmin, dE, sep.SEP_PREC_FP32, want_states=False)
    k32, _ = ctx.ortho_fp32(ns * ms)
    ev = ctx.rayleigh_ritz(k32)
    sg = ctx.sigma(k32)
    _, cbm = edge(ev, sg)
    ev32, sg32 = ctx.refine_window(ev, sg, cbm - 0.1 * ev_ha, cbm + 1.1 * ev_ha, passes=10,
                                   full=True)
    ctx.close()
    vbm64, cbm64 = edge(e64, s64)
    out = {"config": "CdSe 4.0 nm dot on the reference's potCd/potSe tables, 128^3, window "
                     "[-0.26, -0.12] Ha, 16 targets x 16 states, nc 4096, seed -1001",
           "gap_eV": round((cbm64 - vbm64) / ev_ha, 4), "cbm_Ha": {"fp64": round(cbm64, 8),
                                                                   "fp32": round(cbm, 8)}}
    try:
        nclu, nl, dmax = levels.compare_levels(e64, s64, ev32, sg32, cbm64 - levels.MEV,
                                               cbm64 + ev_ha, levels.MEV)
        out.update(conduction_levels=nl, multiplets=nclu, max_abs_err_meV=round(dmax / levels.MEV, 6),
                   passed=nl > 0, matching="one to one per multiplet")
    except AssertionError as exc:
        out.update(passed=False, failure=str(exc)[:300])
    return out


def spinor_bench(sep, decks, np, torch, nc=1024, ns=8, ms=16, cpu_s=10.0, cpu=True):
    """The spin-orbit filter (spinors/filterSpinors, FP64) on a 3.0 nm InAs
    dot at 96^3 (0.67 bohr): In and As on zincblende sites (the reference's Vso values,
    read.c:36,51), P1/P2 passivation (Vso 0), synthetic radial tables,
    Ekinmax 20 Ha; ns random spinors x ms targets, Newton length nc.
    Throughput in spinor H.psi applications per second (ns (nc - 1) / t,
    inputs resident), and the oracle's spinor filter (numpy FFT, one thread)
    timed on a bounded sample of the same Newton steps as the CPU figure."""
    n = 96
    syms, xyz = decks.cdse_dot(3.0, lattice_A=6.057)
    syms = [{"Cd": "In", "Se": "As"}.get(s, s) for s in syms]
    a = 6.057 / 0.529177
    vso = np.array([0.48 * a ** 3 / 8.0 / 27.2114 if s == "In" else
                    (3.15164354 if s == "As" else 0.0) for s in syms])
    xyz = xyz - xyz.mean(axis=0)
    half, d = decks.box_from_atoms(xyz, n, n, n)
    tabs = {"In": (-3.0, 1.5, 0.0, 1.0), "As": (-2.5, 1.4, 0.0, 1.0), "P": (-0.5, 1.0, 0.0, 1.0)}
    stride = 8192
    vr = np.zeros((3, stride))
    pot = np.zeros((3, stride))
    npot, dr = [], []
    for t, key in enumerate(("In", "As", "P")):
        r, v = decks.pot_table(*tabs[key])
        vr[t, :r.size], pot[t, :r.size] = r, v - v[-1]
        npot.append(r.size)
        dr.append(r[1] - r[0])
    types = np.array([0 if s == "In" else (1 if s == "As" else 2) for s in syms])
    ctx = sep.Context(n, n, n, d[0], d[1], d[2], -half[0], -half[1], -half[2], ekinmax=20.0)
    V = ctx.build_potential(xyz, types, npot, vr, pot, dr)
    nent = ctx.so_setup(xyz, vso)
    vmin, dE, its = ctx.energy_range_spinor()
    zn, _ = sep.leja_points(nc)
    e0 = vmin / 1.2 if vmin < 0 else vmin / 0.8
    an = sep.newton_coeffs(zn, np.linspace(e0, e0 + 0.4, ms), dE, vmin, (nc / (2.5 * dE)) ** 2)
    N = n ** 3
    g = torch.Generator(device="cuda:0").manual_seed(3)
    psi0 = torch.zeros((ns, 4, N), dtype=torch.float64, device="cuda:0")
    psi0[:, 0].uniform_(-1.0, 1.0, generator=g)
    psi0[:, 2].uniform_(-1.0, 1.0, generator=g)
    psi0 *= (1.0 / torch.sqrt((psi0[:, 0] ** 2).sum(dim=1) * float(np.prod(d))))[:, None, None]
    ctx.filter_spinor_dev(psi0.data_ptr(), ns, an, zn, vmin, dE)  # warm-up (ring allocation)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.filter_spinor_dev(psi0.data_ptr(), ns, an, zn, vmin, dE)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    ctx.close()
    out = {"config": f"InAs 3.0 nm dot, {len(syms)} sites ({int((vso > 0).sum())} with Vso), "
                     f"{n}^3, Ekinmax 20, {ns} spinors x {ms} targets, nc {nc}, FP64",
           "so_entries": int(nent), "dE": round(float(dE), 4),
           "filter_s": round(t, 4), "spinor_hpsi_per_s": round(ns * (nc - 1) / t, 1)}
    if not cpu:
        return out
    # CPU: the oracle's spinor H.psi on one spinor, Newton steps until cpu_s
    from oracle import oracle as O
    from oracle import spinor as S
    grid = (n, n, n)
    ksq = O.ksqr(n, n, n, d[0], d[1], d[2], 20.0)
    lists = S.so_lists(grid, tuple(-half), tuple(d), xyz, vso)
    psi = psi0[0].cpu().numpy()
    z = np.stack([psi[0] + 1j * psi[1], psi[2] + 1j * psi[3]])
    dv = float(np.prod(d))
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < cpu_s:
        z = S.hamiltonian(grid, ksq, V, lists, z, dv, d[0]) * (4.0 / dE) - z
        steps += 1
    tc = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": round(steps / tc, 3), "unit": "spinor H.psi/s", "cores": 1,
                           "kind": "port",
                           "sample": f"oracle spinor H (numpy FFT + matrix-form V_so), one spinor, "
                                     f"{steps} Newton-step H applications in {tc:.1f} s"}
    return out


def edge_gate(np, decks, levels, e64, s64, e32, s32, conv=1e-4):
    """The 1 meV gate, one to one: every FP64 level with sigma < conv within
    0.5 eV of the lowest such level, multiplet by multiplet, against the
    FP32 pipeline's levels (levels.compare_levels)."""
    ok = s64 < conv
    if not ok.any():
        return {"edge_levels": 0, "passed": False, "max_abs_err_meV": None}
    e0 = float(e64[ok].min())
    out = {"edge_band_Ha": [round(e0, 8), round(e0 + decks.EDGE_BAND_HA, 8)]}
    try:
        nc, nl, d = levels.compare_levels(e64, s64, e32, s32, e0, e0 + decks.EDGE_BAND_HA,
                                          levels.MEV, conv=conv)
        out.update(edge_levels=nl, multiplets=nc, max_abs_err_meV=round(d / levels.MEV, 6),
                   passed=True, matching="one to one per multiplet")
    except AssertionError as exc:
        out.update(edge_levels=int((ok & (e64 <= e0 + decks.EDGE_BAND_HA)).sum()), passed=False,
                   max_abs_err_meV=None, failure=str(exc)[:300])
    return out


def batched_uneven(sep, decks, read_deck, np, torch, steps=1):
    """Batched uneven grids: 8 CdSe dots of 2.0-5.8 nm, two each on 64, 96,
    128 and 160^3 (the smallest grid with spacing <= 0.85 bohr), Newton
    length 2048, 16 targets x 8 random states each, complex64 filter, kernel
    family per dot from the autotuner (sep_autotune_ex).  Times one filter
    pass of all dots: one dot after another, interleaved on their streams,
    and as one captured graph per dot."""
    import tempfile
    diams = [2.0, 2.4, 2.8, 3.4, 4.0, 4.6, 5.2, 5.8]
    nc, ms, ns = 2048, 16, 8
    dots = []
    for k, dnm in enumerate(diams):
        syms, xyz = decks.cdse_dot(float(dnm))
        n = 160
        for cand in (64, 96, 128, 160):
            half, dsp = decks.box_from_atoms(xyz, cand, cand, cand)
            if dsp.max() <= 0.85:
                n = cand
                break
        tmp = tempfile.mkdtemp(prefix="sep_c4_")
        c, h = decks.WINDOW_CENTER_HA, decks.WINDOW_HALF_HA
        decks.write_deck(tmp, syms, xyz, n, n, n, ms, ns, nc, c - h, c + h)
        d = read_deck(tmp)
        ctx = sep.Context(n, n, n, d["dx"], d["dy"], d["dz"], d["xmin"], d["ymin"], d["zmin"])
        ctx.build_potential(d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"],
                            expot=d["expot"])
        tune_ms, fam, lanes = ctx.autotune_ex(sep.SEP_PREC_FP32, (ns + 1) // 2)
        names = ("plane", "plane_chunked", "lines", "plane_swizzled")  # SEP_HPSI_* order
        tune = {"family": names[fam], "lanes": lanes,
                "ms_per_step": {f: round(float(t), 4) for f, t in
                                zip(names, tune_ms) if t > 0}}
        dE, _, _ = ctx.energy_range()
        zn, _ = sep.leja_points(nc)
        an = sep.newton_coeffs(zn, np.linspace(d["Elmin"], d["Elmax"], ms), dE, ctx.vmin,
                               (nc / (2.5 * dE)) ** 2)
        N = n ** 3
        psi0 = np.empty((ns, N))
        idum, r = -2001 - k, None
        for j in range(ns):
            psi0[j], idum, r = sep.init_psi(idum, N, d["dv"], r)
        dots.append(dict(ctx=ctx, n=n, d_nm=round(float(dnm), 3), natom=d["natom"],
                         psi0=torch.from_numpy(psi0).to("cuda:0"), an=an, zn=zn, dE=dE,
                         vmin=ctx.vmin, tune=tune))
        del psi0
    hpsi = sum(ns * (nc - 1) for _ in dots) * steps

    def seq():
        for o in dots:
            o["ctx"].filter_dev(o["psi0"].data_ptr(), ns, o["an"], o["zn"], o["vmin"], o["dE"],
                                sep.SEP_PREC_FP32)

    def bat(mode):
        sep.filter_batch([o["ctx"] for o in dots], [o["psi0"].data_ptr() for o in dots],
                         [ns] * len(dots), [o["an"] for o in dots], [o["zn"] for o in dots],
                         [o["vmin"] for o in dots], [o["dE"] for o in dots],
                         sep.SEP_PREC_FP32, mode)

    out = {"dots": [{"d_nm": o["d_nm"], "grid": o["n"], "atoms": o["natom"],
                     "autotune": o["tune"]} for o in dots]}
    for name, fn in (("sequential", seq), ("interleaved", lambda: bat(sep.SEP_BATCH_INTERLEAVE)),
                     ("graph", lambda: bat(sep.SEP_BATCH_GRAPH))):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        out[name] = {"s": round(t / steps, 3), "hpsi_per_s": round(hpsi / t, 1)}
        log(f"batched uneven grids {name}: {t / steps:.3f} s, {hpsi / t:.0f} H.psi/s")
    for o in dots:
        o["ctx"].close()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the filter "
              "would shard its random states across GPUs, one filter per GPU with no "
              "communication", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import semi_empirical_pseudopotentials_amd as sep
    from semi_empirical_pseudopotentials_amd import decks, levels
    from semi_empirical_pseudopotentials_amd._lib import read_deck

    n, ns, ms, nc = args.grid, args.ns, args.ms, args.nc
    if args.gemm_only:
        print(json.dumps({"gemm_only": True, **subspace_gemms(sep, np, torch, n)}), flush=True)
        return
    if args.spinor_only:
        print(json.dumps({"spinor_only": True,
                          **spinor_bench(sep, decks, np, torch, cpu=not args.no_cpu)}), flush=True)
        return
    # ---------------------------------------------------------------- setup
    t_setup = time.time()
    syms, xyz = decks.cdse_dot(args.diameter)
    tmp = tempfile.mkdtemp(prefix="sep_bench_")
    c, h = decks.WINDOW_CENTER_HA, decks.WINDOW_HALF_HA
    decks.write_deck(tmp, syms, xyz, n, n, n, ms, ns, nc, c - h, c + h)
    d = read_deck(tmp)
    ctx = sep.Context(n, n, n, d["dx"], d["dy"], d["dz"], d["xmin"], d["ymin"], d["zmin"])
    ctx.build_potential(d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"],
                        expot=d["expot"])
    vmin = ctx.vmin
    dE, emax, iters = ctx.energy_range()
    zn, _ = sep.leja_points(nc)
    el = np.linspace(d["Elmin"], d["Elmax"], ms)
    dt = (nc / (2.5 * dE)) ** 2
    an = sep.newton_coeffs(zn, el, dE, vmin, dt)
    N = n ** 3
    psi0 = np.empty((ns, N))
    idum, rng = -1002, None
    for j in range(ns):
        psi0[j], idum, rng = sep.init_psi(idum, N, d["dv"], rng)
    psi0_dev = torch.from_numpy(psi0).to("cuda:0")
    del psi0
    torch.cuda.synchronize()
    log(f"setup {time.time() - t_setup:.1f}s: natom={d['natom']} grid={n}^3 dx={d['dx']:.4f} "
        f"dE={dE:.4f} (power iterations {iters})")

    # ------------------------------------------------ per-pass kernel timing
    nvec = (ns + 1) // 2
    if args.kernels_only:
        k = ctx.bench_hpsi(sep.SEP_PREC_FP32, nvec, ms, 5)
        # the fused kernel's previous (xor-swizzled LDS) form, for comparison
        ctx.set_hpsi_family(sep.SEP_PREC_FP32, sep.SEP_HPSI_PLANE_SWIZZLED)
        k_swz = ctx.bench_hpsi(sep.SEP_PREC_FP32, nvec, ms, 5)
        ctx.set_hpsi_family(sep.SEP_PREC_FP32, sep.SEP_HPSI_PLANE)
        a = torch.empty(64 << 20, dtype=torch.float32, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            torch.add(a, 1.0, out=b)  # plain copy kernel: 256 MiB read + 256 MiB written
        torch.cuda.synchronize()
        print(json.dumps({"kernels_only": True, "pass_ms": list(k), "pass_ms_swizzled": list(k_swz),
                          "copy_bytes": 2 * a.numel() * 4}), flush=True)
        ctx.close()
        return
    # ------------------------------------------------------------ timed loop
    # The timed passes run as the library runs them by default (filter lanes:
    # groups of random states on concurrent streams).  One more pass samples
    # the three H.psi passes of one Newton step per ring block, and that
    # block's accumulation, with HIP events (sep_set_timing); live timing
    # keeps the loop on one stream, so those are the kernels' own durations.
    ctx.set_filter_lanes(args.lanes)

    def run(prec):
        for _ in range(args.warmup):
            ctx.filter_dev(psi0_dev.data_ptr(), ns, an, zn, vmin, dE, prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.filter_dev(psi0_dev.data_ptr(), ns, an, zn, vmin, dE, prec)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        ctx.set_timing(True)
        t1 = time.perf_counter()
        ctx.filter_dev(psi0_dev.data_ptr(), ns, an, zn, vmin, dE, prec)
        torch.cuda.synchronize()
        t_one = time.perf_counter() - t1
        live, samples = ctx.get_timing()
        ctx.set_timing(False)
        return t, live, samples, t_one

    t32, live32, smp32, one32 = run(sep.SEP_PREC_FP32)
    hpsi = ns * (nc - 1) * args.steps
    v32 = hpsi / t32
    log(f"fp32 filter: {t32 / args.steps:.3f} s/step, {v32:.0f} H.psi/s "
        f"(one stream, timed: {one32:.3f} s); live kernel ms "
        f"A {live32[0]:.4f} B {live32[1]:.4f} C {live32[2]:.4f} accum {live32[3]:.3f} "
        f"(samples {smp32})")
    v64 = live64 = one64 = None
    if not args.no_fp64:
        t64, live64, smp64, one64 = run(sep.SEP_PREC_FP64)
        v64 = hpsi / t64
        log(f"fp64 filter: {t64 / args.steps:.3f} s/step, {v64:.0f} H.psi/s "
            f"(one stream, timed: {one64:.3f} s); live kernel ms "
            f"A {live64[0]:.4f} B {live64[1]:.4f} C {live64[2]:.4f} accum {live64[3]:.3f}")

    # ------------------------------------- whole pipeline at this workload
    # after one more filter pass of each precision: ortho (FP64: double-double
    # Gram, the reference's SVD cutoff; FP32 pipeline: FP64 Gram eigenbasis),
    # Rayleigh-Ritz, and sigma (FP64) or the FP64 preconditioned correction
    # pass (FP32 pipeline); wall times per stage
    pipe = {}
    if args.pipeline:
        # one untimed pass of each pipeline's dense stages at the timed sizes
        # (on the states the timed filters left) loads rocBLAS / rocSOLVER's
        # kernels, which are chosen by size and loaded on first use once per
        # process; both precisions are warmed alike
        t0 = time.perf_counter()
        k, _ = ctx.ortho_fp32(ns * ms)
        ev = ctx.rayleigh_ritz(k)
        sg = ctx.sigma(k)
        lo, hi, _ = levels.edge_window(ev, sg)
        ctx.refine_window(ev, sg, lo, hi, passes=2)
        if not args.no_fp64:
            k, _ = ctx.ortho(ns * ms)
            ctx.rayleigh_ritz(k)
            ctx.sigma(k)
        pipe["library_load_s"] = round(time.perf_counter() - t0, 3)
        for name, prec in (("fp32", sep.SEP_PREC_FP32), ("fp64", sep.SEP_PREC_FP64)):
            if name == "fp64" and args.no_fp64:
                continue
            st = {}
            t0 = time.perf_counter()
            ctx.filter_dev(psi0_dev.data_ptr(), ns, an, zn, vmin, dE, prec)
            torch.cuda.synchronize()
            st["filter"] = time.perf_counter() - t0
            t0 = time.perf_counter()
            if prec == sep.SEP_PREC_FP32:
                kept, _ = ctx.ortho_fp32(ns * ms)
            else:
                kept, _ = ctx.ortho(ns * ms)
            st["ortho"] = time.perf_counter() - t0
            t0 = time.perf_counter()
            ev = ctx.rayleigh_ritz(kept)
            st["rayleigh_ritz"] = time.perf_counter() - t0
            t0 = time.perf_counter()
            sg = ctx.sigma(kept)
            st["sigma"] = time.perf_counter() - t0
            extra = {}
            if prec == sep.SEP_PREC_FP32:
                # FP64 correction passes on the pairs of the band-edge window
                t0 = time.perf_counter()
                lo, hi, e0 = levels.edge_window(ev, sg)
                ev, sg = ctx.refine_window(ev, sg, lo, hi, passes=10)
                st["window_refine"] = time.perf_counter() - t0
                extra = {"window_Ha": [round(lo, 8), round(hi, 8)], "window_pairs": int(ev.size),
                         "window_passes": 10}
            pipe[name] = {"kept": int(kept), "s": {k: round(v, 3) for k, v in st.items()},
                          "total_s": round(sum(st.values()), 3),
                          "converged_sigma_lt_1e-4": int((sg < 1e-4).sum()), **extra}
            pipe["_ev_" + name] = (ev, sg)
            log(f"{name} pipeline: {pipe[name]}")
        if "_ev_fp64" in pipe:
            e64, s64 = pipe.pop("_ev_fp64")
            e32, s32 = pipe.pop("_ev_fp32")
            pipe["edge_levels_fp32_vs_fp64"] = edge_gate(np, decks, levels, e64, s64, e32, s32)
            log(f"128^3 edge levels: {pipe['edge_levels_fp32_vs_fp64']}")
        else:
            pipe.pop("_ev_fp32", None)

    # ------------------------------------------------------------- roofline
    # Algorithmic bytes per launch, nvec complex vectors of N points
    # (two real states per vector), c = complex bytes, r = real bytes:
    #   pass A: read psi (re, im planes) + write T          2c N nvec
    #   pass B: read T + write T                             2c N nvec
    #   pass C: read T, psi, V + write psi_j                 (3c + r) N nvec
    # i.e. N (7c + r) per complex vector and H application.  Where the filter
    # loop fuses pass C with the next step's pass A (the live timing then
    # reports pass A as 0), that kernel reads T, psi_{j-1}, V and writes
    # psi_j and the next T: (4c + r) N nvec, and a step moves N (6c + r).
    def roof(live, c, r, lines=False):
        # lines: the row + strided-line family, whose fused step's spectral
        # part is y lines, z pencils and y lines (3 x 2c per point)
        if live[0] <= 0.0:
            per = {"B": (6 if lines else 2) * c * N * nvec, "CA": (4 * c + r) * N * nvec}
            tms = {"B": live[1], "CA": live[2]}
        else:
            per = {"A": 2 * c * N * nvec, "B": 2 * c * N * nvec, "C": (3 * c + r) * N * nvec}
            tms = {"A": live[0], "B": live[1], "C": live[2]}
        out = {}
        for key in per:
            gbs = per[key] / (tms[key] * 1e-3) / 1e9
            out[key] = {"ms": round(tms[key], 5), "bytes": per[key], "GBs": round(gbs, 1),
                        "frac": round(gbs / PEAK_HBM_GBS, 4)}
        tot, ts = sum(per.values()), sum(tms.values())
        out["step"] = {"ms": round(ts, 5), "bytes": tot, "GBs": round(tot / (ts * 1e-3) / 1e9, 1),
                       "frac": round(tot / (ts * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
        # the same step against the algorithmic bytes of the three-pass
        # schedule, N (7c + r) per complex vector and H application: what the
        # step would move with whole-plane kernels, whatever family runs
        alg = (7 * c + r) * N * nvec
        out["step_on_3pass_bytes"] = {"bytes": alg,
                                      "GBs": round(alg / (ts * 1e-3) / 1e9, 1),
                                      "frac": round(alg / (ts * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
        return out

    r32 = roof(live32, 8, 4)
    dom = max((q for q in r32 if not q.startswith("step")), key=lambda q: r32[q]["ms"])
    fused = {sep.SEP_HPSI_PLANE: "k_plane_bwd_fwd_s<1024>",
             sep.SEP_HPSI_PLANE_SWIZZLED: "k_plane_bwd_fwd_w<float,128>",
             sep.SEP_HPSI_PLANE_CHUNKED: "k_plane_bwd_fwd_h<float,128>"}.get(
                 ctx.hpsi_family(sep.SEP_PREC_FP32), "k_rows_bwd_fwd<float,128>")
    kname = {"A": "k_plane_fwd_w<float,128>", "B": "k_pencil<float,128>",
             "C": "k_plane_bwd_w<float,128>", "CA": fused}[dom]
    traffic = measured_traffic(kname, n, nvec)
    probe = round(ctx.bench_copy(4 << 30, 5), 1)  # plain 16 B/lane copy, best of 4 grids
    roofline = {"bound": "hbm", "kernel": kname, "achieved": r32[dom]["GBs"],
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": r32[dom]["frac"],
                "traffic": traffic, "algorithmic_bytes": r32[dom]["bytes"],
                "hpsi_passes_frac": r32["step"]["frac"],
                "copy_probe_GBs": probe,
                "frac_of_copy_probe": round(r32[dom]["GBs"] / probe, 4),
                "hpsi_passes_frac_of_copy_probe": round(r32["step"]["GBs"] / probe, 4)}

    # --------------------------------------- Hellmann-Feynman adjoint kernel
    # one backward of the differentiable path: the ns*ms filtered states of the
    # last filter pass (resident) contracted as one weighted density, and the
    # 16 diagonal densities of the first states one by one
    m = ns * ms
    rng_w = np.random.default_rng(0)
    w = rng_w.standard_normal(m)
    diag = np.repeat(np.arange(m, dtype=np.int64)[:, None], 2, axis=1)
    adj = {}
    for label, pr, ww in (("weighted_density", diag, w), ("16_pairs", diag[:16], None)):
        ctx.hf_adjoint(pr, d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"], weights=ww)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.hf_adjoint(pr, d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"], weights=ww)
        adj[label + "_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    log(f"hf adjoint: {adj}")

    check = phys = None
    # the timed filters are done: the main workload's ring of Newton vectors
    # (up to 60 % of the free memory) would otherwise shrink the checks' rings
    ctx.trim()
    if not args.no_check:
        check = precision_check(sep, decks, read_deck, np)
        log(f"fp32 pipeline check: {check}")
        phys = physical_check(sep, decks, read_deck, np)
        log(f"fp32 pipeline at a real band edge: {phys}")

    # ---------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        grid = (n, n, n)
        ksq = O.ksqr(*grid, d["dx"], d["dy"], d["dz"])
        V = np.empty(N)
        V[:] = ctx.build_potential(d["xyz"], d["type"], d["npot"], d["vr"], d["pot"], d["dr"],
                                   expot=d["expot"])
        steps = args.cpu_steps
        p0 = psi0_dev.cpu().numpy()
        t0 = time.perf_counter()
        O.filter_states(grid, ksq, V, an[:, :steps + 1], zn[:steps + 1], 4.0 / dE, vmin, p0)
        tc = time.perf_counter() - t0
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        model = "unknown"
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": round(ns * steps / tc, 3), "unit": "H·ψ/s", "cores": threads,
               "kind": "port", "cpu_model": model, "nproc": os.cpu_count(),
               "sample": f"oracle filter (own C FFT), all {ns} random states in parallel "
                         f"(one per OpenMP thread) x {steps} Newton steps x {ms} targets at "
                         f"{n}^3: {ns * steps} H.psi in {tc:.1f} s"}
        log(f"cpu oracle: {ns * steps / tc:.2f} H.psi/s on {threads} threads ({model})")

    batched = None
    if not args.no_batched:
        ctx.close()  # free the 128^3 workload's device memory first
        ctx = None
        del psi0_dev
        torch.cuda.empty_cache()
        batched = batched_uneven(sep, decks, read_deck, np, torch)
    if ctx is not None:
        ctx.close()
        ctx = None
    spinor = None
    if not args.no_spinor:
        spinor = spinor_bench(sep, decks, np, torch, cpu=not args.no_cpu)
        log(f"spinor filter: {spinor}")
    gemm = subspace_gemms(sep, np, torch, n)
    log(f"subspace gemms: {gemm}")

    res = {
        "metric": METRIC,
        "value": round(v32, 1),
        "unit": "H·ψ/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t32 / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"CdSe {args.diameter} nm dot ({d['natom']} atoms), {n}^3 grid, "
                               f"{ns} random states x {ms} targets, Newton length {nc}; "
                               "complex64 filter (FP64 accumulators)",
                   "grid": n, "states": ns, "targets": ms, "newton_length": nc,
                   "filter_lanes": args.lanes or "automatic"},
        "fp64_value": round(v64, 1) if v64 else None,
        "one_stream_filter_s": {"fp32": round(one32, 4),
                                "fp64": round(one64, 4) if one64 is not None else None},
        "fp32_over_fp64": round(v32 / v64, 3) if v64 else None,
        "complex_hpsi_per_s": round(v32 / 2, 1),
        "roofline": roofline,
        "roofline_passes_fp32": r32,
        # complex128 plane kernels exist up to 96 points per line; 128 runs the
        # row + line family
        "roofline_passes_fp64": (roof(live64, 16, 8, lines=n not in (16, 24, 32, 36, 48, 64, 96))
                                 if live64 is not None else None),
        "accum_ms_per_block": {"fp32": round(live32[3], 4),
                               "fp64": round(live64[3], 4) if live64 is not None else None},
        "fp32_pipeline_check": check,
        "fp32_physical_edge_check": phys,
        "pipeline_128": pipe or None,
        # the whole pipeline at this workload (filter, ortho, Rayleigh-Ritz,
        # sigma; FP32: plus the FP64 correction passes of the band-edge window)
        "end_to_end_fp32_over_fp64": (round(pipe["fp64"]["total_s"] / pipe["fp32"]["total_s"], 3)
                                      if pipe and "fp64" in pipe and "fp32" in pipe else None),
        "gate_1meV_one_to_one": (pipe.get("edge_levels_fp32_vs_fp64", {}).get("passed")
                                 if pipe else None),
        "batched_uneven": batched,
        "spinor_filter": spinor,
        "hf_adjoint": dict(adj, atoms=d["natom"], states=m),
        "subspace_gemm": gemm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    if ctx is not None:
        ctx.close()


if __name__ == "__main__":
    main()
