#!/usr/bin/env python3
"""Throughput of ExaMiniMD's time step on one MI355X.

Workload (default): LJ melt, 4 000 000 atoms (FCC 100^3 cells, density
0.8442, cutoff 2.5, skin 0.3, rebuild every 20 steps, dt 0.005, T* 1.44,
full CSR neighbour list, serial periodic images), run in FP64 and in mixed
precision (FP32 pair arithmetic and accumulation, FP64 positions/velocities).
A step is one velocity-Verlet step of the whole system; the timed region is
K steps with the state resident in HBM.  `value` is the mixed-precision
atom-steps/s; the FP64 figure and the force error of the mixed path stand
beside it.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload lj4m|lj32k|snap2k]
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6     # MI355X vector FP64
FP32_PEAK_TFLOPS = 157.3    # MI355X vector FP32 (published)


_T0 = time.perf_counter()


def progress(msg):
    """one line per finished leg on stderr (the JSON line stays alone on stdout)"""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="lj4m", choices=["lj4m", "lj32k"])
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--snap-steps", type=int, default=100)
    ap.add_argument("--no-snap", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--loop-steps", type=int, default=100)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-forcematch", action="store_true")
    ap.add_argument("--no-idial", action="store_true",
                    help="skip the lj/cut/idial arithmetic-intensity sweep")
    ap.add_argument("--neigh", default="2D", choices=["CSR", "2D"],
                    help="neighbour list layout of the LJ runs: Neighbor2D (the reference's "
                         "default, src/input.cpp:132) or NeighborCSR")
    ap.add_argument("--no-alt-layout", action="store_true",
                    help="skip the mixed LJ run with the other neighbour list layout")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device (profiler calibration)")
    return ap.parse_args()


def _profile_table(L, ctx):
    buf = C.create_string_buffer(1 << 16)
    L.emd_profile_report(ctx, buf, len(buf))
    prof = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.split()
        prof[name] = (int(cnt), float(ms))
    return prof


def _breakdown(app, L, ctx, steps):
    """per-kernel milliseconds per step over `steps` further steps with every
    launch timed (outside the timed region: the event records add gaps)"""
    if steps <= 0:
        return {}
    L.emd_profile_select(ctx, None)
    L.emd_profile_enable(ctx, 1)
    L.emd_profile_reset(ctx)
    app.run(steps)
    L.emd_synchronize(ctx)
    prof = _profile_table(L, ctx)
    L.emd_profile_enable(ctx, 0)
    return {k: v[1] / steps for k, v in prof.items()}


def lj_run(emd, decks, tmpdir, n, precision, warmup, steps, chunk=100, breakdown_steps=40,
           neigh="CSR"):
    """the timed LJ run, in chunks of `chunk` steps (a multiple of the
    rebuild interval, so the rebuild schedule is that of one run); E/N is
    sampled between chunks outside the timed region.  Inside the timed
    region only the force kernel carries HIP events (its mean duration is
    the roofline's denominator); a per-kernel breakdown follows over
    `breakdown_steps` untimed steps."""
    deck = decks.write(os.path.join(tmpdir, f"in.lj.{precision}"),
                       decks.lj_deck(n=n, steps=steps))
    app = emd.App(["-il", deck, "--neigh-type", neigh, "--comm-type", "SERIAL",
                   "--precision", precision]).init()
    app.set_phase_sync(False)
    L = emd.hip_lib()
    ctx = app.ctx
    if warmup:
        app.run(warmup)
    L.emd_synchronize(ctx)
    samples = [(0, app.thermo()[2])]
    kname = "lj_force_mixed" if precision == "mixed" else "lj_force_fp64"
    L.emd_profile_select(ctx, kname.encode())
    L.emd_profile_enable(ctx, 1)
    L.emd_profile_reset(ctx)
    wall, done = 0.0, 0
    while done < steps:
        k = min(chunk, steps - done)
        L.emd_synchronize(ctx)
        t0 = time.perf_counter()
        app.run(k)
        L.emd_synchronize(ctx)
        wall += time.perf_counter() - t0
        done += k
        L.emd_profile_enable(ctx, 0)
        samples.append((done, app.thermo()[2]))
        L.emd_profile_enable(ctx, 1)
    app.energy_samples = samples
    prof = _profile_table(L, ctx)
    L.emd_profile_enable(ctx, 0)
    app.thermo_end = list(app.thermo())
    nl = app.L.emd_app_neigh(app.h).contents
    num = emd.download_ints(ctx, nl.num, nl.n)
    natoms = app.natoms
    nlist = float(num.sum()) / natoms
    app.breakdown = _breakdown(app, L, ctx, breakdown_steps)
    return app, wall, prof, nlist


def energy_drift(samples, dt):
    """least-squares slope of E/N against time (reduced units per tau) and
    the largest excursion from the first sample; no slope from fewer than
    three samples"""
    import numpy as np
    st = np.array([s for s, _ in samples], dtype=float) * dt
    e = np.array([v for _, v in samples], dtype=float)
    slope = float(np.polyfit(st, e, 1)[0]) if len(e) >= 3 else None
    interval = int(samples[1][0] - samples[0][0]) if len(samples) > 1 else None
    return {"slope_per_tau": slope, "max_excursion": float(np.abs(e - e[0]).max()),
            "samples": len(e), "sample_interval_steps": interval, "e0": float(e[0])}


def drift_leg(emd, decks, tmpdir, n, precision, steps=1000, every=100, neigh="2D"):
    """the LJ run's energy drift over a fixed 1000 steps (independent of
    --steps), E/N sampled every 100 steps (11 samples).  Every sample step is
    a rebuild step (100 is a multiple of the rebuild interval 20).  For the
    mixed run, at each sample the list the run rebuilt incrementally (one-pass
    fill into the previous build's capacity, 2D layout) is compared entry for
    entry with a CSR list built from scratch (count, scan, fill) from the
    same positions in a second context."""
    import numpy as np
    from examinimd_amd.abi import EMD_NEIGH_CSR
    from examinimd_amd.pipeline import Pipeline
    deck = decks.write(os.path.join(tmpdir, f"in.lj.drift.{precision}"),
                       decks.lj_deck(n=n, steps=steps))
    app = emd.App(["-il", deck, "--neigh-type", neigh, "--precision", precision]).init()
    app.set_phase_sync(False)
    box = tuple(app.L.emd_app_system(app.h).contents.domain)
    samples = [(0, app.thermo()[2])]
    checks = []
    for k in range(steps // every):
        app.run(every)
        samples.append(((k + 1) * every, app.thermo()[2]))
        if precision == "mixed":
            a = app.arrays()
            row, _, ent = app.neigh()
            pl = Pipeline()
            pl.load(a["x"], box)
            pl.build(2.8, layout=EMD_NEIGH_CSR)
            frow, _, fent = emd.download_neigh(pl.ctx.h, pl.nl)
            # the positions are loaded in the run's (sorted) order with ids
            # 1..n: a fresh sort that keeps that order keeps the ids ascending
            same_order = np.array_equal(pl._device_ids(), np.arange(1, len(a["x"]) + 1))
            pl.close()
            checks.append(bool(same_order and np.array_equal(row, frow)
                               and np.array_equal(ent, fent)))
    app.close()
    ident = None
    if checks:
        ident = {"rebuilds_checked": len(checks), "identical": all(checks),
                 "steps": [every * (k + 1) for k in range(len(checks))],
                 "against": "CSR list built from scratch from the same positions"}
    return energy_drift(samples, 0.005), ident


def melt_force_error(emd, app):
    """force RMS error of the mixed path on its own trajectory: the FP32
    forces of its last step against FP64 forces re-evaluated at the same
    positions with the same neighbour list (trajectories of the two paths
    diverge, so errors are only comparable on identical positions)"""
    import numpy as np
    from examinimd_amd.abi import EMD_FP64, EmdLJParams
    L, ctx = emd.hip_lib(), app.ctx
    f32 = app.arrays()["f"]
    s = app.L.emd_app_system(app.h)
    nl = app.L.emd_app_neigh(app.h)
    p = EmdLJParams()
    p.ntypes = 1
    p.lj1[0][0], p.lj2[0][0], p.cutsq[0][0] = 48.0, 24.0, 6.25
    emd.check(ctx, L.emd_force_zero(ctx, s, EMD_FP64), "force_zero")
    emd.check(ctx, L.emd_lj_force(ctx, s, nl, C.byref(p), EMD_FP64, 0), "lj_force")
    n = app.nlocal
    f64 = np.zeros((n, 3))
    emd.check(ctx, L.emd_system_download(ctx, s, n, None, None,
                                         f64.ctypes.data_as(C.POINTER(C.c_double)), None, None,
                                         None), "download")
    df = f32 - f64
    return {"rms_error": float(np.sqrt(np.mean(np.sum(df * df, axis=1)))),
            "force_rms": float(np.sqrt(np.mean(np.sum(f64 * f64, axis=1))))}


def force_error(emd, decks, tmpdir, n):
    """force RMS error of the mixed path against FP64 on identical positions"""
    import numpy as np
    res = {}
    for prec in ("fp64", "mixed"):
        deck = decks.write(os.path.join(tmpdir, f"in.lj.err.{prec}"), decks.lj_deck(n=n, steps=0))
        app = emd.App(["-il", deck, "--neigh-type", "CSR", "--precision", prec]).init()
        res[prec] = app.arrays()
        app.close()
    df = res["mixed"]["f"] - res["fp64"]["f"]
    return float(np.sqrt(np.mean(np.sum(df * df, axis=1))))


def snap_files(emd, decks, tmpdir):
    """W-like SNAP potential: beta_1..55 ~ N(0, 0.05^2) (seed 2024), beta_0 set
    so that the perfect BCC lattice (a = 3.1803 A) has E/atom = -8.9 eV"""
    beta = decks.snap_beta(seed=2024, sigma=0.05)
    param = decks.write(os.path.join(tmpdir, "W.snapparam"), decks.SNAP_PARAM)
    coeff = decks.write(os.path.join(tmpdir, "W.snapcoeff"), decks.snap_coeff_text(beta))
    deck = decks.write(os.path.join(tmpdir, "in.snap.fit"),
                       decks.snap_deck(n=10, steps=0, coeff=coeff, param=param))
    app = emd.App(["-il", deck, "--neigh-type", "CSR"]).init()
    pe0 = app.thermo()[1]
    app.close()
    beta[0] = -8.9 - pe0
    decks.write(coeff, decks.snap_coeff_text(beta))
    return coeff, param, beta


def snap_run(emd, decks, tmpdir, coeff, param, precision, warmup, steps, breakdown_steps=20,
             repeats=5):
    """timed SNAP runs with no events (at 2000 atoms an event pair is a
    noticeable share of a kernel's dispatch gap): `repeats` consecutive runs
    of `steps` steps, the median one reported (a run of 100 steps lasts
    20-30 ms, where a single hiccup shows), then an untimed per-kernel
    breakdown whose SNAP kernel times feed the roofline"""
    deck = decks.write(os.path.join(tmpdir, f"in.snap.{precision}"),
                       decks.snap_deck(n=10, steps=repeats * steps + breakdown_steps,
                                       coeff=coeff, param=param))
    app = emd.App(["-il", deck, "--neigh-type", "CSR", "--precision", precision]).init()
    app.set_phase_sync(False)
    L = emd.hip_lib()
    ctx = app.ctx
    if warmup:
        app.run(warmup)
    L.emd_synchronize(ctx)
    L.emd_profile_enable(ctx, 0)
    walls = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        app.run(steps)
        L.emd_synchronize(ctx)
        walls.append(time.perf_counter() - t0)
    wall = sorted(walls)[len(walls) // 2]
    natoms = app.natoms
    e_per_atom = app.thermo()[2]
    breakdown = _breakdown(app, L, ctx, breakdown_steps)
    app.close()
    return natoms, wall, e_per_atom, breakdown


def snap_force_error(emd, beta):
    """FP32 U/Y/dU force RMS error against FP64 on a thermally displaced BCC
    configuration (0.1 A Gaussian displacements, 2000 atoms)"""
    import numpy as np
    from examinimd_amd.pipeline import Pipeline, lattice
    a = 3.1803
    x = lattice("bcc", 10, a) + np.random.default_rng(3).normal(0, 0.1, (2000, 3))
    pl = Pipeline()
    pl.load(x, (10 * a,) * 3, mass=183.84)
    pl.build(4.73442 + 1.0, newton=1)
    pl.snap_setup(list(beta))
    f64 = pl.snap_force(emd.EMD_FP64)
    f32 = pl.snap_force(emd.EMD_MIXED)
    pl.close()
    df = f32 - f64
    return float(np.sqrt(np.mean(np.sum(df * df, axis=1))))


def lj_force_error_disordered(emd):
    """FP32 LJ force RMS error against FP64 on a displaced FCC lattice"""
    import numpy as np
    from examinimd_amd.pipeline import Pipeline, lattice
    a = (4 / 0.8442) ** (1 / 3)
    x = lattice("fcc", 20, a) + np.random.default_rng(5).normal(0, 0.05, (32000, 3))
    pl = Pipeline()
    pl.load(x, (20 * a,) * 3)
    pl.build(2.8)
    f64 = pl.lj(emd.EMD_FP64)
    f32 = pl.lj(emd.EMD_MIXED)
    pl.close()
    df = f32 - f64
    return float(np.sqrt(np.mean(np.sum(df * df, axis=1))))


def idial_leg(emd, n=100, dials=(1, 4, 16, 64, 256, 1024), launches=10):
    """pair_style lj/cut/idial as an arithmetic-intensity dial: the full-list
    force kernel on the LJ 4M configuration (FCC 100^3 displaced by 0.05
    sigma, CSR list at rc + skin = 2.8), FP64 and mixed, at each dial.  Each
    pair inside rc = 2.5 adds its part `dial` times; the extra time per
    extra repeat between the two largest dials gives the rate of those
    dependent additions, compared with the published vector peak counted
    in additions (one add = one flop, half the FMA-based TFLOP/s figure)."""
    import numpy as np
    from examinimd_amd.pipeline import Pipeline, lattice
    a = (4 / 0.8442) ** (1 / 3)
    x = lattice("fcc", n, a) + np.random.default_rng(3).normal(0, 0.05, (4 * n ** 3, 3))
    pl = Pipeline()
    pl.load(x, (n * a,) * 3)
    del x
    L, h = pl.L, pl.ctx.h
    pl.build(2.5)
    n_in = float(emd.download_ints(h, pl.nl.num, pl.nl.n).astype(np.int64).sum())
    pl.build(2.8)
    natoms = pl.s.n_local
    p = emd.abi.EmdLJIDialParams()
    p.lj.ntypes = 1
    p.lj.lj1[0][0], p.lj.lj2[0][0], p.lj.cutsq[0][0] = 48.0, 24.0, 6.25
    res = {}
    for prec, name in ((emd.EMD_FP64, "lj_idial_fp64"), (emd.EMD_MIXED, "lj_idial_mixed")):
        key = "fp64" if prec == emd.EMD_FP64 else "mixed"
        res[key] = {}
        for d in dials:
            p.intensity[0][0] = d
            pl.chk(L.emd_lj_idial_force_store(h, C.byref(pl.s), C.byref(pl.nl), C.byref(p), prec),
                   "idial warm-up")
            L.emd_profile_select(h, name.encode())
            L.emd_profile_enable(h, 1)
            L.emd_profile_reset(h)
            for _ in range(launches):
                pl.chk(L.emd_lj_idial_force_store(h, C.byref(pl.s), C.byref(pl.nl), C.byref(p),
                                                  prec), "idial")
            L.emd_synchronize(h)
            cnt, ms = _profile_table(L, h).get(name, (0, 0.0))
            L.emd_profile_enable(h, 0)
            res[key][str(d)] = ms / max(cnt, 1)
    pl.close()
    peaks = {"fp64": FP64_PEAK_TFLOPS * 1e12 / 2, "mixed": FP32_PEAK_TFLOPS * 1e12 / 2}
    out = {"workload": f"FCC {n}^3 ({natoms} atoms) displaced 0.05 sigma, CSR list, rc 2.5",
           "atoms": natoms, "pairs_inside_rc_per_atom": n_in / natoms,
           "kernel_avg_ms": res, "add_peak_per_s": peaks}
    hi, lo = dials[-1], dials[-2]
    for key in res:
        slope = (res[key][str(hi)] - res[key][str(lo)]) * 1e-3 / (hi - lo)
        rate = n_in / slope if slope > 0 else None
        out[key + "_repeat_adds_per_s"] = rate
        out[key + "_repeat_adds_frac_of_peak"] = rate / peaks[key] if rate else None
        out[key + "_atom_steps_per_s_force_only"] = {d: natoms / (ms * 1e-3)
                                                     for d, ms in res[key].items()}
    out["mixed_over_fp64_force_time"] = {d: res["fp64"][d] / res["mixed"][d] for d in res["fp64"]}
    return out


def host_description():
    """nproc, the CPU model (lscpu) and the thread share this process may use"""
    import subprocess
    model = None
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return {"nproc": os.cpu_count(), "lscpu_model": model,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(n, steps, threads):
    """the CPU oracle timed on this host's cores.  The thread count is the
    CPU share given to this one-GPU process (OMP_NUM_THREADS, set to 16 on the
    GPU hosts, which share their cores between GPUs), else every core"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    if threads <= 0:
        env = os.environ.get("OMP_NUM_THREADS", "")
        threads = int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)
    cfg = pyoracle.make_config(n=(n, n, n), gaussian=True, nthreads=threads)
    m = pyoracle.MD(cfg)
    m.init()
    t0 = time.perf_counter()
    m.run(steps)
    dt = time.perf_counter() - t0
    natoms = m.system.N
    m.close()
    return {"value": natoms * steps / dt, "unit": "atom-steps/s", "cores": threads,
            "kind": "port", "host": host_description(),
            "sample": f"CPU oracle (C/OpenMP restatement of the reference), same LJ "
                      f"{natoms}-atom workload, {steps} steps incl. {steps // 20} "
                      f"neighbour rebuilds, {dt:.1f} s"}


def batch_leg(emd, steps, loop_steps):
    """256 uneven periodic LJ replicas (N log-uniform in [256, 8192] rounded
    to FCC n^3 cells, rho in [0.70, 0.95], T* in [0.8, 2.0]), mixed
    precision: batched launches with graph-replayed rebuild periods and an
    autotuned force tile, against the same ensemble run replica after
    replica through the per-system C ABI"""
    import numpy as np
    from examinimd_amd import batch
    from examinimd_amd.abi import EMD_FP64, EMD_MIXED
    e = batch.ensemble(nrep=256, seed=4)
    b = batch.Batch(e["natoms"], e["box"], precision=EMD_MIXED)
    b.upload(e["x"], e["v"])
    b.setup()
    lanes, block, tune_ms = b.autotune(trial_steps=10)
    # force error of the tuned tile: FP32 forces of the mixed batch (set-up
    # force recomputed with that tile) against an FP64 batch at the same
    # positions
    b.setup()
    x0, v0, f32 = b.download()
    b64 = batch.Batch(e["natoms"], e["box"], precision=EMD_FP64)
    b64.set_tile(lanes, block)
    b64.upload(x0, v0)
    b64.setup()
    _, _, f64 = b64.download()
    b64.close()
    df = f32 - f64
    tile_err = float(np.sqrt(np.mean(np.sum(df * df, axis=1))))
    b.run(20)                      # warm-up, captures the period graph
    ke0, pe0 = b.thermo()
    t0 = time.perf_counter()
    b.run(steps)                   # returns after a stream synchronisation
    wall = time.perf_counter() - t0
    ke1, pe1 = b.thermo()
    total = b.total
    b.close()
    sec, _, _ = batch.replica_loop(e["natoms"], e["box"], e["x"], e["v"], loop_steps,
                                   precision=EMD_MIXED)
    value = total * steps / wall
    loop_value = total * loop_steps / sec
    return {"value": value, "unit": "atom-steps/s", "ms_per_step": 1e3 * wall / steps,
            "steps": steps, "replicas": 256, "atoms": total,
            "atoms_min_max": [int(e["natoms"].min()), int(e["natoms"].max())],
            "tile": {"lanes": lanes, "block": block, "tuned_ms_per_step": tune_ms,
                     "mixed_force_rms_error_vs_fp64": tile_err},
            "max_energy_change_per_atom": float(np.max(np.abs((ke1 + pe1) - (ke0 + pe0))
                                                       / e["natoms"])),
            "replica_loop": {"value": loop_value, "ms_per_step": 1e3 * sec / loop_steps,
                             "steps": loop_steps},
            "batched_over_loop": value / loop_value,
            "workload": "256 uneven periodic LJ replicas, FP32 forces / FP64 state, rc 2.5, "
                        "skin 0.3, rebuild every 20, dt 0.005"}


def forcematch_leg(emd, repeats=3):
    """force-matching gradients: 64 SNAP configurations (BCC W, 128..1024
    atoms) w.r.t. the 56 beta, and 32 LJ boxes (500..4000 atoms) w.r.t.
    (eps, sigma); mixed-precision adjoint timed, checked against the FP64
    adjoint and a central difference of the FP64 loss"""
    import numpy as np
    from examinimd_amd import forcematch as fm
    from examinimd_amd.abi import EMD_FP64, EMD_MIXED
    out = {}
    bt = fm.snap_beta_true()
    b = fm.snap_beta_perturbed(bt)
    cfgs = fm.snap_set()

    def timed(obj):
        obj.evaluate(b, EMD_MIXED)
        t0 = time.perf_counter()
        for _ in range(repeats):
            res = obj.evaluate(b, EMD_MIXED)
        return (time.perf_counter() - t0) / repeats, res

    # the set merged into one system (one launch sequence per evaluation)
    s = fm.SnapForceMatchMerged(cfgs, bt)
    wall, (L32, g32) = timed(s)
    L64, g64 = s.evaluate(b, EMD_FP64)
    h, k = 1e-6, 29
    bp, bm = b.copy(), b.copy()
    bp[k] += h
    bm[k] -= h
    fd = (s.evaluate(bp, EMD_FP64, grad=False)[0] - s.evaluate(bm, EMD_FP64, grad=False)[0]) / (2 * h)
    natoms = sum(len(x) for x, _ in cfgs)
    ctx = s.ctx
    # one state and launch sequence per configuration, for comparison
    per = fm.SnapForceMatch(cfgs, bt, ctx=ctx)
    wall_per, (_, g_per) = timed(per)
    per.close()
    s.close()
    out["snap"] = {"configs": len(cfgs), "atoms": natoms, "ms_per_gradient": 1e3 * wall,
                   "atom_gradients_per_s": natoms / wall, "loss": L32,
                   "grad_rel_err_mixed_vs_fp64": float(np.abs(g32 - g64).max() / np.abs(g64).max()),
                   "fd_check": {"component": k, "fd": fd, "adjoint_fp64": float(g64[k])},
                   "per_configuration": {"ms_per_gradient": 1e3 * wall_per,
                                         "grad_rel_diff_vs_merged":
                                             float(np.abs(g_per - g32).max() / np.abs(g32).max())},
                   "method": "64 configurations merged into one system"}
    cfgs = fm.lj_set()
    th = np.array([1.04, 0.98])

    def timed_lj(obj):
        obj.evaluate(th, EMD_MIXED)
        t0 = time.perf_counter()
        for _ in range(repeats):
            res = obj.evaluate(th, EMD_MIXED)
        return (time.perf_counter() - t0) / repeats, res

    lj = fm.LJForceMatchMerged(cfgs)
    wall, (L32, g32) = timed_lj(lj)
    L64, g64 = lj.evaluate(th, EMD_FP64)
    per = fm.LJForceMatch(cfgs, ctx=lj.ctx)
    wall_per, (_, g_per) = timed_lj(per)
    per.close()
    natoms = sum(len(x) for x, _ in cfgs)
    lj.close()
    out["lj"] = {"configs": len(cfgs), "atoms": natoms, "ms_per_gradient": 1e3 * wall,
                 "atom_gradients_per_s": natoms / wall,
                 "grad_rel_err_mixed_vs_fp64": float(np.abs(g32 - g64).max() / np.abs(g64).max()),
                 "per_configuration": {"ms_per_gradient": 1e3 * wall_per,
                                       "grad_rel_diff_vs_merged":
                                           float(np.abs(g_per - g32).max() / np.abs(g32).max())},
                 "method": "32 boxes merged into one system"}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only)"}))
        sys.exit(2)
    import examinimd_amd as emd
    from examinimd_amd import decks

    n = 100 if a.workload == "lj4m" else 20
    tmpdir = tempfile.mkdtemp(prefix="emd_bench_")
    out = {}
    for prec in ("fp64", "mixed"):
        app, wall, prof, nlist = lj_run(emd, decks, tmpdir, n, prec, a.warmup, a.steps,
                                        neigh=a.neigh)
        natoms = app.natoms
        kname = "lj_force_mixed" if prec == "mixed" else "lj_force_fp64"
        cnt, ms = prof.get(kname, (0, 0.0))
        out[prec] = {"wall": wall, "natoms": natoms, "prof": prof, "nlist": nlist,
                     "kernel": kname, "kernel_launches": cnt, "kernel_ms": ms,
                     "thermo_end": app.thermo_end, "breakdown": app.breakdown}
        if prec == "mixed":
            out[prec]["melt_error"] = melt_force_error(emd, app)
        app.close()
        progress(f"LJ {natoms} atoms {prec} ({a.neigh}): {natoms * a.steps / wall:.4g} atom-steps/s")
    for prec in ("fp64", "mixed"):
        out[prec]["drift"], ident = drift_leg(emd, decks, tmpdir, n, prec, neigh=a.neigh)
        if ident:
            out[prec]["neigh_identity"] = ident
        progress(f"LJ {prec} 1000-step drift leg done")

    # the mixed run again with the other list layout (same workload)
    alt = None
    if not a.no_alt_layout:
        lay = "CSR" if a.neigh == "2D" else "2D"
        app, wall, prof, nlist = lj_run(emd, decks, tmpdir, n, "mixed", a.warmup, a.steps,
                                        neigh=lay, breakdown_steps=0)
        cnt, ms = prof.get("lj_force_mixed", (0, 0.0))
        progress(f"LJ mixed ({lay}): {app.natoms * a.steps / wall:.4g} atom-steps/s")
        alt = {"neigh": lay, "value": app.natoms * a.steps / wall,
               "ms_per_step": 1e3 * wall / a.steps,
               "lj_force_mixed_avg_ms": ms / max(cnt, 1)}
        app.close()

    mixed, fp64 = out["mixed"], out["fp64"]
    natoms = mixed["natoms"]
    value = natoms * a.steps / mixed["wall"]
    value_fp64 = natoms * a.steps / fp64["wall"]
    # roofline of the dominant kernel (LJ force, mixed): algorithmic bytes per
    # atom = 4*nlist (row entries) + row bookkeeping (2D: length 4 B; CSR:
    # offset and length 8 B) + 24 (x_i) + 24 (each x_j fetched once) + 12
    # (the f32 force, stored once)
    bytes_per_atom = 4.0 * mixed["nlist"] + (4 if a.neigh == "2D" else 8) + 24 + 24 + 12
    avg_ms = mixed["kernel_ms"] / max(mixed["kernel_launches"], 1)
    achieved = bytes_per_atom * natoms / (avg_ms * 1e-3) / 1e9
    ferr = force_error(emd, decks, tmpdir, 20)
    ferr_dis = lj_force_error_disordered(emd)
    # HBM bytes per launch of the same kernel from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes (scaled by a 256 MiB copy calibration)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "lj4m_traffic.json")
    if n == 100 and os.path.exists(tpath):
        t = json.load(open(tpath))
        if (t.get("kernel") == mixed["kernel"] and t.get("atoms") == natoms
                and t.get("neigh", "CSR") == a.neigh):
            traffic = t["fetch_bytes"] + t["write_bytes"]

    # SNAP 2000 atoms (BCC W), FP64 and mixed
    snap = {}
    if not a.no_snap:
        coeff, param, beta = snap_files(emd, decks, tmpdir)
        for prec in ("fp64", "mixed"):
            sn, sw, se, sb = snap_run(emd, decks, tmpdir, coeff, param, prec, 10, a.snap_steps)
            progress(f"SNAP {sn} atoms {prec}: {sn * a.snap_steps / sw:.4g} atom-steps/s")
            snap[prec] = {"value": sn * a.snap_steps / sw, "ms_per_step": 1e3 * sw / a.snap_steps,
                          "snap_kernels_ms_per_step": {k: v for k, v in sb.items()
                                                       if k in ("snap_neigh", "snap_ui", "snap_y",
                                                                "snap_fij")},
                          "kernels_ms_per_step_untimed_breakdown": sb,
                          "etot_per_atom_end": se}
        snap["mixed_over_fp64"] = snap["mixed"]["value"] / snap["fp64"]["value"]
        # VALU roofline of the SNAP kernels: loop-bound FLOPs of the Y form
        # with the reverse-mode force (examinimd_amd/flopcount.py; 26
        # neighbours inside the cutoff of BCC W at a = 3.1803) over the summed
        # time per step of the four SNAP kernels (compaction with the
        # neighbour parameters, U, Y, force)
        from examinimd_amd.flopcount import snap_flops
        fl = snap_flops(8, 26)["total"]
        for prec, peak in (("fp64", FP64_PEAK_TFLOPS), ("mixed", FP32_PEAK_TFLOPS)):
            kms = sum(snap[prec]["snap_kernels_ms_per_step"].values())
            ach = fl * 2000 / (kms * 1e-3) / 1e12 if kms > 0 else None
            snap[prec]["roofline"] = {"bound": "valu", "achieved": ach, "peak": peak,
                                      "unit": "TFLOP/s",
                                      "frac": ach / peak if ach else None,
                                      "flops_per_atom_step": fl,
                                      "kernel_ms_per_step": kms}
        snap["mixed_force_rms_error_eV_per_A"] = snap_force_error(emd, beta)
        snap["workload"] = (f"SNAP W BCC 10^3 (2000 atoms), twojmax 8, rc 4.73442, skin 1.0, "
                            f"rebuild every 10, dt 0.5 fs, {a.snap_steps} steps "
                            f"(median of 5 consecutive runs)")

    batched = None if a.no_batch else batch_leg(emd, a.batch_steps, a.loop_steps)
    if batched:
        progress("batched replicas done")
    gradients = None if a.no_forcematch else forcematch_leg(emd)
    if gradients:
        progress("force-matching gradients done")
    idial = None
    if not a.no_idial and n == 100:
        idial = idial_leg(emd)
        progress("lj/cut/idial intensity sweep done")

    line = {
        "metric": "atom-steps/s on 1 MI355X, LJ 4M & SNAP 2k atoms: mixed vs FP64 at fixed error",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * mixed["wall"] / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair math / f64 state",
        "data": "synthetic FCC lattice, seeded Gaussian velocities",
        "config": {"workload": f"LJ melt {natoms} atoms (FCC {n}^3, rho 0.8442, rc 2.5, "
                               f"skin 0.3, rebuild every 20), mixed precision",
                   "atoms": natoms, "neighbors_per_atom": round(mixed["nlist"], 3),
                   "neigh": a.neigh},
        "other_layout": alt,
        "fp64": {"value": value_fp64, "ms_per_step": 1e3 * fp64["wall"] / a.steps,
                 "dtype": "f64"},
        "mixed_over_fp64": value / value_fp64,
        "mixed_force_rms_error_vs_fp64": ferr,
        "mixed_force_rms_error_on_melt": mixed["melt_error"],
        "thermo_end": {"fp64": fp64["thermo_end"], "mixed": mixed["thermo_end"],
                       "fields": ["T", "PE/N", "E/N"]},
        "energy_drift": {"fp64": fp64["drift"], "mixed": mixed["drift"],
                         "unit": "E/N per tau (LJ reduced units)",
                         "run": "1000 steps from set-up, E/N every 100 steps, separate from "
                                "the timed run"},
        "neighbour_identity_at_rebuilds": mixed.get("neigh_identity"),
        "mixed_force_rms_error_vs_fp64_displaced_lattice": ferr_dis,
        "snap2k": snap,
        "batched": batched,
        "force_matching": gradients,
        "lj_idial_sweep": idial,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": bytes_per_atom * natoms,
                     "kernel": mixed["kernel"], "avg_ms": avg_ms,
                     "algorithmic_bytes_per_atom": bytes_per_atom},
        "kernels_ms_per_step_untimed_breakdown": mixed["breakdown"],
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(n, a.cpu_steps, a.cpu_threads)
        progress("CPU baseline done")
    print(json.dumps(line), flush=True)
    if a.calib_copy_mib:
        # a device-to-device copy of known size, for scaling profiler byte counters
        ctx = emd.Context()
        L = emd.hip_lib()
        nbytes = a.calib_copy_mib << 20
        src, dst = C.c_void_p(), C.c_void_p()
        ctx.check(L.emd_malloc(ctx.h, C.byref(src), nbytes), "malloc")
        ctx.check(L.emd_malloc(ctx.h, C.byref(dst), nbytes), "malloc")
        ctx.check(L.emd_memcpy_d2d(ctx.h, dst, src, nbytes), "copy")
        ctx.sync()
        L.emd_free(ctx.h, src)
        L.emd_free(ctx.h, dst)


if __name__ == "__main__":
    main()
