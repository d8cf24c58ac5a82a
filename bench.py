"""Benchmark of the LEE time advance on one MI355X.

Workload (the configuration the headline metric is quoted on): 2-D
curvilinear single block 1025x1025, x = xi + 0.05 L sin(2 pi eta / L),
y = eta + 0.05 L sin(2 pi xi / L) on [-50, 50]^2 (L = 100), uniform mean
flow M = 0.5 in x, time-harmonic monopole (omega = 0.2 pi, Gaussian
half-width 4 dx, amplitude 1e-3) at a position drawn by
numpy.random.default_rng(2) in the central 60 %, 40-point Freund sponges on
all faces, RK4 at CFL 0.5 with the 8th-order filter.  Synthetic inputs,
built here; state resident in HBM for the timed region.

A step is one call of the time advance (4 RK stages + filter); the metric
counts grid-point RK-stage updates: 1025^2 * 4 per step.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
N1D = 1025


def c2_inputs(n=N1D):
    L = 100.0
    d = 100.0 / (n - 1)
    xi = -50.0 + d * np.arange(n)
    X, E = np.meshgrid(xi, xi)
    x = X + 0.05 * L * np.sin(2.0 * np.pi * E / L)
    y = E + 0.05 * L * np.sin(2.0 * np.pi * X / L)
    mean = np.zeros((5, n, n))
    mean[1] = 0.5
    mean[4] = 1.0 / 1.4
    rng = np.random.default_rng(2)
    xs, ys = rng.uniform(-30.0, 30.0, 2)
    w = 40 - 1
    bc = [(0, "sponge_freund_ambient", "+xi", "0", str(w), "0", "-0", "0", "0"),
          (0, "sponge_freund_ambient", "-xi", f"-{w}", "-0", "0", "-0", "0", "0"),
          (0, "sponge_freund_ambient", "+eta", "0", "-0", "0", str(w), "0", "0"),
          (0, "sponge_freund_ambient", "-eta", "0", "-0", f"-{w}", "-0", "0", "0")]
    src = (1e-3, 0.2 * math.pi, float(xs), float(ys), 0.0, 4.0 * d)
    return dict(x=x, y=y, mean=mean, bc=bc, source=src, ic=np.zeros((5, n, n)), dx=d)


def c5_inputs(nsolve=64, seed=5, dx=0.5, sponge=20):
    """Batched ensemble: `nsolve` independent Cartesian solves (spacing dx),
    sizes drawn by numpy.random.default_rng(seed) with nx in [129, 1025] and
    ny in [129, 513], each with a monopole (amplitude 1e-3, half-width 4 dx)
    in the central 60 % at omega in [0.1 pi, 0.5 pi], uniform mean flow
    M in [0, 0.6] along x and Freund sponges `sponge` points wide on every
    face; RK4 at CFL 0.5 with the 8th-order filter."""
    rng = np.random.default_rng(seed)
    w = sponge - 1
    bc = [(0, "sponge_freund_ambient", "+xi", "0", str(w), "0", "-0", "0", "0"),
          (0, "sponge_freund_ambient", "-xi", f"-{w}", "-0", "0", "-0", "0", "0"),
          (0, "sponge_freund_ambient", "+eta", "0", "-0", "0", str(w), "0", "0"),
          (0, "sponge_freund_ambient", "-eta", "0", "-0", f"-{w}", "-0", "0", "0")]
    out = []
    for _ in range(nsolve):
        nx = int(rng.integers(129, 1026))
        ny = int(rng.integers(129, 514))
        fx, fy = rng.uniform(0.2, 0.8, 2)
        omega = float(rng.uniform(0.1, 0.5)) * math.pi
        mach = float(rng.uniform(0.0, 0.6))
        x, y = np.meshgrid(dx * np.arange(nx), dx * np.arange(ny))
        mean = np.zeros((5, ny, nx))
        mean[1] = mach
        mean[4] = 1.0 / 1.4
        src = (1e-3, omega, fx * dx * (nx - 1), fy * dx * (ny - 1), 0.0, 4.0 * dx)
        out.append(dict(x=x, y=y, mean=mean, boundaries=bc, simulation="CASE_MONOPOLE", source_params=src,
                        mode="FIX_CFL", value=0.5))
    return out


def c5_leg(steps, warmup, single_steps):
    """Times the batched ensemble (FP32 and FP64 bulk), the FP32 ensemble run
    one handle per solve for comparison, and each solve's FP32 pressure
    error against its FP64 run."""
    from camille_amd import FP32, FP64, GRID_RESIDENT, PATH_FUSED, Batch, Solver
    solves = c5_inputs()
    npts = sum(sv["x"].size for sv in solves)
    res = {"solves": len(solves), "points": npts, "steps": steps, "warmup": warmup}
    finals = {}
    for label, prec in (("fp32", FP32), ("fp64", FP64)):
        print(f"[bench] C5 batch {label}: {len(solves)} solves, {npts} points, {steps} steps", file=sys.stderr,
              flush=True)
        b = Batch(solves, precision=prec)
        for k, sv in enumerate(solves):
            b.set_solution(np.zeros((5,) + sv["x"].shape), k)
        t0 = time.perf_counter()
        b.solve(warmup)  # includes the autotuner and the graph capture
        b.synchronize()
        t_warm = time.perf_counter() - t0
        t0 = time.perf_counter()
        b.solve(steps, sync=False)
        b.synchronize()
        wall = time.perf_counter() - t0
        ups = npts * 4 * steps / wall
        finals[label] = [b.solution(k)[4] for k in range(len(solves))]
        picked = dict(zip(("graph", "grid_stage", "grid_filter"), b.launch()))
        # the form the workload names: one persistent launch per RK stage
        # (and filter), one resident wave of workgroups each, the step
        # replayed from a HIP graph
        b.set_launch(True, GRID_RESIDENT, GRID_RESIDENT)
        b.solve(warmup)
        b.synchronize()
        t0 = time.perf_counter()
        b.solve(steps, sync=False)
        b.synchronize()
        wall_p = time.perf_counter() - t0
        configured = dict(updates_per_s=npts * 4 * steps / wall_p, ms_per_step=wall_p / steps * 1e3,
                          launch=dict(zip(("graph", "grid_stage", "grid_filter"), b.launch())))
        b.set_launch(picked["graph"], picked["grid_stage"], picked["grid_filter"])
        kt = None
        if label == "fp32":
            g, gs, gf = b.launch()
            b.set_launch(False, gs, gf)  # per-kernel events need direct launches
            b.set_kernel_timing(True)
            b.solve(20)
            kt = b.kernel_timing()
            b.set_kernel_timing(False)
        res[label] = dict(updates_per_s=ups, ms_per_step=wall / steps * 1e3, warmup_s=t_warm,
                          tile_shape=list(b.tile_shape()), launch=picked,
                          persistent_graph_form=configured)
        if kt:
            res[label]["kernel_ms"] = {k: [round(t, 4), c] for k, (t, c) in kt.items()}
            per_point = {"stage_first": 92, "stage_mid": 124, "stage_last": 92, "filter": 32}
            res[label]["kernel_hbm_frac"] = {
                c: per_point[c] * npts / (kt[c][0] / kt[c][1] * 1e-3) / HBM_PEAK for c in per_point if kt[c][1]}
        b.close()
        print(f"[bench] C5 batch {label}: {ups:.3e} updates/s", file=sys.stderr, flush=True)
    errs = []
    for p32, p64 in zip(finals["fp32"], finals["fp64"]):
        errs.append(float(np.max(np.abs(p32 - p64)) / np.max(np.abs(p64))))
    res["p_rel_err_fp32_vs_fp64"] = {"max": max(errs), "median": float(np.median(errs))}
    # the same FP32 ensemble one handle per solve, launched one after another
    print(f"[bench] C5 one handle per solve: {single_steps} steps", file=sys.stderr, flush=True)
    hs = []
    for sv in solves:
        s = Solver([dict(x=sv["x"], y=sv["y"], mean=sv["mean"])], boundaries=sv["boundaries"], mode="FIX_CFL",
                   value=0.5, simulation="CASE_MONOPOLE", source_params=sv["source_params"], precision=FP32,
                   path=PATH_FUSED)
        s.set_solution(np.zeros((5,) + sv["x"].shape))
        s.solve(2)
        hs.append(s)
    for s in hs:
        s.synchronize()
    t0 = time.perf_counter()
    for _ in range(single_steps):
        for s in hs:
            s.solve(1, sync=False)
    for s in hs:
        s.synchronize()
    wall = time.perf_counter() - t0
    res["fp32_one_handle_per_solve_updates_per_s"] = npts * 4 * single_steps / wall
    for s in hs:
        s.close()
    return res


def c3_inputs():
    """Overset cylinder: background Cartesian 2049x1025 on [-16, 16] x [-8, 8]
    and an O-grid 513x129 from r = 0.5 to 2 around the cylinder
    (camille_amd.overset_gen: holes for r < 1.5, 5-cell fringe, O-grid seam
    from its own interior, ORDER-4 donor stencils); potential-flow mean at
    M = 0.3 with p from Bernoulli; wall-slip on the cylinder; 64-point
    Freund sponges on the background; Gaussian pressure pulse (1e-3,
    half-width 0.2) upstream at (-3, 0); RK4 at CFL 0.5 with the 8th-order
    filter.  Probes: 64 fluid background points in [-8, 8] x [-4, 4] drawn
    by numpy.random.default_rng(3)."""
    from camille_amd import overset_gen as og
    geo = og.cylinder()
    blocks = []
    for x, y in geo["blocks"]:
        ic = np.zeros((5,) + x.shape)
        ic[4] = 1e-3 * np.exp(-math.log(2.0) * ((x + 3.0) ** 2 + y * y) / 0.2 ** 2)
        blocks.append(dict(x=x, y=y, mean=og.potential_flow(x, y, 0.3, 0.5), ic=ic))
    w = 64 - 1
    bc = [(0, "sponge_freund_ambient", "+xi", "0", str(w), "0", "-0", "0", "0"),
          (0, "sponge_freund_ambient", "-xi", f"-{w}", "-0", "0", "-0", "0", "0"),
          (0, "sponge_freund_ambient", "+eta", "0", "-0", "0", str(w), "0", "0"),
          (0, "sponge_freund_ambient", "-eta", "0", "-0", f"-{w}", "-0", "0", "0"),
          (1, "wall_slip_kinematic", "+eta", "0", "-0", "0", "0", "0", "0")]
    return dict(geo=geo, blocks=blocks, bc=bc, tables=og.abi_tables(geo))


def c3_leg(steps, sample_every, window=1000, paths=("fp32", "fp64", "plain")):
    """Overset cylinder in FP32 bulk, FP64 fused and FP64 plain (the
    reference's arithmetic): throughput and probe time series.

    The reference's discretization of this configuration is not stable for
    long times: a wall mode grows without bound (the reference binary's own
    long run of the small cylinder case, tests/golden/cylinder_long.json,
    grows 2.7x per 100 steps).  The `steps` steps therefore run as restarts
    of a `window`-step run from the initial condition, inside which the
    state stays bounded; each window is checked (every probe sample finite
    and below 10x the pulse amplitude) and a violation fails the leg.  The
    fused paths run steps / window windows, the plain path one; errors are
    over the first window's probe samples (when the plain path runs).
    `paths` selects the legs (a profiling run may leave some out)."""
    from camille_amd import FP32, FP64, PATH_FUSED, PATH_PLAIN, Solver
    inp = c3_inputs()
    npts = sum(b["x"].size for b in inp["blocks"])
    nwin = max(1, steps // window)
    res = {"points": npts, "steps": nwin * window, "window_steps": window, "windows": nwin,
           "sample_every": sample_every, "receivers": int(len(inp["tables"]["recv_l"])),
           "holes": int(len(inp["tables"]["hole_l"]))}
    series = {}
    probes = None
    bound = 10.0 * 1e-3
    for label, prec, path in (("fp32", FP32, PATH_FUSED), ("fp64", FP64, PATH_FUSED), ("plain", FP64, PATH_PLAIN)):
        if label not in paths:
            continue
        runs = nwin if path == PATH_FUSED else 1
        print(f"[bench] C3 {label}: {npts} points, {runs} x {window} steps", file=sys.stderr, flush=True)
        s = Solver([dict(x=b["x"], y=b["y"], mean=b["mean"]) for b in inp["blocks"]], boundaries=inp["bc"],
                   overset=inp["tables"], precision=prec, path=path, mode="FIX_CFL", value=0.5)
        if probes is None:
            ib0 = s.iblank(0)
            x, y = inp["blocks"][0]["x"], inp["blocks"][0]["y"]
            cand = np.flatnonzero((ib0 == 0) & (np.abs(x) <= 8.0) & (np.abs(y) <= 4.0))
            pick = np.random.default_rng(3).choice(cand, 64, replace=False)
            ni = x.shape[1]
            probes = [(0, int(p % ni), int(p // ni), 1.0) for p in pick]
        s.set_observer(probes)
        s.autotune()  # the launch forms are chosen before the clock starts
        wall = 0.0
        for r in range(runs):
            for ib, b in enumerate(inp["blocks"]):
                s.set_solution(b["ic"], ib)
            s.set_time(0, 0.0, 1e-6)
            s.solve(0)
            samples = []
            t0 = time.perf_counter()
            for k in range(window // sample_every):
                s.solve(sample_every, sync=False)
                samples.append(s.sample_observers())
            s.synchronize()
            wall += time.perf_counter() - t0
            smp = np.array(samples)
            if not (np.isfinite(smp).all() and np.max(np.abs(smp)) <= bound):
                s.close()
                raise RuntimeError(f"C3 {label}: window {r} left the bounded range "
                                   f"(max |p| {np.nanmax(np.abs(smp)):.3e})")
            if r == 0:
                series[label] = smp
        res[label] = dict(updates_per_s=npts * 4 * runs * window / wall, ms_per_step=wall / (runs * window) * 1e3,
                          time=s.clock()[1], dt=s.clock()[2], windows=runs)
        s.close()
        print(f"[bench] C3 {label}: {res[label]['updates_per_s']:.3e} updates/s", file=sys.stderr, flush=True)
    if "plain" not in series:
        return res
    ref = series["plain"]
    scale = float(np.max(np.abs(ref)))
    res["probe_p_max"] = scale
    res["p_rel_err_fp32_vs_plain"] = float(np.max(np.abs(series["fp32"] - ref)) / scale)
    res["p_rel_err_fp64_fused_vs_plain"] = float(np.max(np.abs(series["fp64"] - ref)) / scale)
    return res


def c4_inputs(n=256, sponge=16):
    """3-D Cartesian n^3 block on [-2.5, 2.5]^3 with the parallel shear mean
    flow u(r) = 0.45 (1 - tanh((r - 0.5) / 0.05)), r = sqrt(y^2 + z^2),
    p = 1/gamma, s = 0 (so the mean-gradient LEE terms act); Freund sponges
    `sponge` points wide on all six faces; observer points within half a
    cell of the cylinder r = 1.5, weight h^2 each.

    The faces need the sponges: with the one-sided closures alone, the
    points where three closure directions meet (the box's corners) carry a
    growing mode of the scheme, which takes p' from the 1e-3 source level
    to O(1) within 500 steps in FP64 and amplifies FP32 round-off there."""
    w = sponge - 1
    spec = {"+xi": ("0", str(w), "0", "-0", "0", "-0"), "-xi": (f"-{w}", "-0", "0", "-0", "0", "-0"),
            "+eta": ("0", "-0", "0", str(w), "0", "-0"), "-eta": ("0", "-0", f"-{w}", "-0", "0", "-0"),
            "+zeta": ("0", "-0", "0", "-0", "0", str(w)), "-zeta": ("0", "-0", "0", "-0", f"-{w}", "-0")}
    bc = [(0, "sponge_freund_ambient", side) + spec[side] for side in spec] if sponge else []
    h = 5.0 / (n - 1)
    c = -2.5 + h * np.arange(n)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(y * y + z * z)
    mean = np.zeros((5, n, n, n))
    mean[1] = 0.45 * (1.0 - np.tanh((r - 0.5) / 0.05))
    mean[4] = 1.0 / 1.4
    obs = np.flatnonzero(np.abs(r - 1.5).reshape(-1) < 0.5 * h)
    return dict(x=x, y=y, z=z, mean=mean, obs=obs, h=h, n=n, bc=bc, sponge=sponge)


def c4_leg(nsteps=500, window=400, every=50):
    """Gradient of J = sum over the last 100 of 500 steps of the observer
    energy with respect to a 32^3 source-amplitude control (trilinear
    embedding, p-source a(x) sin(omega t)), through the discrete adjoint
    with checkpoints every 50 steps, in FP32 and FP64; central finite
    difference (eps 1e-4) of the FP64 forward along one random direction."""
    import torch

    from camille_amd import FP32, FP64, Solver
    from camille_amd.adjoint import embed_control3, observer_energy
    inp = c4_inputs()
    n = inp["n"]
    shape = (n, n, n)
    npts = n ** 3
    rng = np.random.default_rng(4)
    a0 = torch.tensor(rng.uniform(0.5, 1.5, (32, 32, 32)), dtype=torch.float64)
    e = torch.tensor(rng.standard_normal((32, 32, 32)), dtype=torch.float64)
    y0 = torch.zeros((5,) + shape, dtype=torch.float64)
    res = {"grid": [n, n, n], "steps": nsteps, "window": [window + 1, nsteps], "checkpoint_every": every,
           "observers": int(inp["obs"].size), "control": [32, 32, 32], "sponge_points": inp["sponge"]}
    grads = {}
    solvers = {}
    p_end = {}
    for label, prec in (("fp32", FP32), ("fp64", FP64)):
        print(f"[bench] C4 {label}: {n}^3, {nsteps} steps forward + adjoint", file=sys.stderr, flush=True)
        s = Solver([dict(x=inp["x"], y=inp["y"], z=inp["z"], mean=inp["mean"])], num_dim=3, precision=prec,
                   simulation="SOURCE_FIELD", source_params=(1e-3, 2.0 * math.pi * 0.5), mode="FIX_CFL", value=0.5,
                   boundaries=inp["bc"])
        s.set_observer_points(np.zeros(inp["obs"].size), inp["obs"], np.full(inp["obs"].size, inp["h"] ** 2))
        s.set_solution(np.zeros((5,) + shape))
        s.autotune()  # the launch forms are chosen before the clock starts
        a = a0.clone().requires_grad_(True)
        t0 = time.perf_counter()
        J = observer_energy(s, embed_control3(a, shape), y0, nsteps, window, every)
        t1 = time.perf_counter()
        J.backward()
        t2 = time.perf_counter()
        grads[label] = a.grad.detach().numpy().copy()
        p_end[label] = s.solution()[4].copy()  # the state after the forward run, as the reverse pass leaves it
        res[label] = dict(J=float(J.item()), forward_s=t1 - t0, backward_s=t2 - t1,
                          forward_updates_per_s=npts * 4 * nsteps / (t1 - t0),
                          dt=s.clock()[2])
        solvers[label] = s
        print(f"[bench] C4 {label}: J {J.item():.6e}, forward {t1 - t0:.2f} s, backward {t2 - t1:.2f} s",
              file=sys.stderr, flush=True)
    g64, g32 = grads["fp64"], grads["fp32"]
    res["grad_rel_err_fp32_vs_fp64"] = float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64)))
    res["J_rel_err_fp32_vs_fp64"] = abs(res["fp32"]["J"] - res["fp64"]["J"]) / res["fp64"]["J"]
    w = inp["sponge"]
    inner = (slice(w, -w),) * 3 if w else (slice(None),) * 3
    scale = float(np.max(np.abs(p_end["fp64"])))
    res["p_rel_err_fp32_vs_fp64_inside_sponges"] = float(np.max(np.abs(p_end["fp32"][inner] - p_end["fp64"][inner])) / scale)
    res["p_rel_err_fp32_vs_fp64_whole_box"] = float(np.max(np.abs(p_end["fp32"] - p_end["fp64"])) / scale)
    del p_end
    s = solvers["fp64"]
    eps = 1e-4
    with torch.no_grad():
        jp = observer_energy(s, embed_control3(a0 + eps * e, shape), y0, nsteps, window, every).item()
        jm = observer_energy(s, embed_control3(a0 - eps * e, shape), y0, nsteps, window, every).item()
    fd = (jp - jm) / (2 * eps)
    ad = float(np.sum(g64 * e.numpy()))
    res["fd_check"] = dict(eps=eps, fd=fd, adjoint=ad, rel_err=abs(fd - ad) / abs(fd),
                           adjoint_fp32=float(np.sum(g32 * e.numpy())),
                           rel_err_fp32=abs(fd - float(np.sum(g32 * e.numpy()))) / abs(fd))
    for v in solvers.values():
        v.close()
    return res


def make_solver(inp, precision, path):
    from camille_amd import Solver
    s = Solver([dict(x=inp["x"], y=inp["y"], mean=inp["mean"])], simulation="CASE_MONOPOLE",
               source_params=inp["source"], precision=precision, path=path, boundaries=inp["bc"],
               mode="FIX_CFL", value=0.5)
    s.set_solution(inp["ic"])
    return s


def timed(s, warmup, steps, kernel_timing=False):
    """Wall time of `steps` steps after `warmup` untimed ones; then, when
    kernel_timing, the same number of steps again with HIP events around
    every kernel (the events themselves slow the stream down, so they stay
    out of the throughput figure)."""
    s.solve(warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.solve(steps, sync=False)
    s.synchronize()
    wall = time.perf_counter() - t0
    kt = None
    if kernel_timing:
        s.set_kernel_timing(True)
        s.solve(steps)
        kt = s.kernel_timing()
        s.set_kernel_timing(False)
    return wall, kt


def cpu_baseline(inp, nsteps):
    """The CPU oracle (FP64 restatement, 1 thread) on the same workload for a
    bounded number of steps."""
    from oracle.pyoracle import Oracle
    scheme = dict(fd_order=4, filter_order=8, filter_blend=1.0, rk=4, sponge_order=2, sponge_constant=5.0)
    o = Oracle([dict(x=inp["x"], y=inp["y"], mean=inp["mean"], ic=inp["ic"])], scheme, inp["bc"],
               ("CASE_MONOPOLE", inp["source"]), "FIX_CFL", 0.5)
    t0 = time.perf_counter()
    o.run(nsteps)
    wall = time.perf_counter() - t0
    return wall


def _cpu_worker(nsteps):
    return cpu_baseline(c2_inputs(), nsteps)


def cpu_baseline_all_cores(nproc, nsteps):
    """The single-threaded oracle on every host core at once: `nproc`
    processes each run the same workload for `nsteps` steps; the aggregate
    rate is nproc * nsteps steps over the slowest process's time.  Worker
    processes start fresh (spawn), so none inherits this process's GPU
    state."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    with ctx.Pool(nproc) as pool:
        walls = pool.map(_cpu_worker, [nsteps] * nproc)
    return max(walls)


def host_cpu():
    """(usable cores, CPU model) of this host."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, model


def copy_calibration():
    """A plain device-to-device copy of known size: 256 MiB read and 256 MiB
    written per launch.  FETCH_SIZE / WRITE_SIZE of this launch give the
    factors that turn those counters into bytes."""
    import torch
    n = 64 * 1024 * 1024
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4, "launches": 5}))


def copy_ceiling():
    """Device-to-device copy bandwidth (bytes read + written per second) at a
    size the 256 MB Infinity Cache can hold (128 MiB) and at one it cannot
    (2 GiB): the practical HBM ceilings of in-cache and out-of-cache
    working sets.  Runs in a process of its own (torch's HIP runtime is not
    mixed with the library's)."""
    import torch
    out = {}
    for label, nbytes in (("128MiB", 128 << 20), ("2GiB", 2 << 30)):
        a = torch.ones(nbytes // 4, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 20
        e0.record()
        for _ in range(reps):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[label] = 2.0 * nbytes / (ms * 1e-3) / 1e9
        del a, b
    print(json.dumps({"copy_GBps": out}))


def measured_copy_ceiling():
    import subprocess
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-ceiling"], capture_output=True,
                           text=True, timeout=300)
        return json.loads(r.stdout.strip().splitlines()[-1])["copy_GBps"]
    except Exception as e:  # reported, not fatal: the headline does not depend on it
        return {"error": str(e)[:200]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--accuracy-steps", type=int, default=2000)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--cpu-procs", type=int, default=16, help="processes of the all-cores CPU leg (at most the host's cores)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--c5-steps", type=int, default=1000)
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--c3-steps", type=int, default=5000)
    ap.add_argument("--c3-window", type=int, default=1000)
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--c3-paths", default="fp32,fp64,plain",
                    help="with --only c3: the legs to run, comma-separated (fp32, fp64, plain)")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--c4-steps", type=int, default=500,
                    help="3-D adjoint leg: steps (window: the last fifth; checkpoints every tenth)")
    ap.add_argument("--only", choices=["fp32", "fp64", "plain", "c5", "c3", "c4"], default=None,
                    help="time one path only (for profiling)")
    ap.add_argument("--copy-ceiling", action="store_true",
                    help="only measure device-copy bandwidth at 128 MiB and 2 GiB (one JSON line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 256 MiB (HBM counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.copy_ceiling:
        copy_ceiling()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (N = 1); larger N is not supported",
              file=sys.stderr)
        sys.exit(2)

    ceiling = None if args.only else measured_copy_ceiling()
    from camille_amd import FP32, FP64, PATH_FUSED, PATH_PLAIN
    inp = c2_inputs()
    npts = N1D * N1D
    upd_per_step = npts * 4

    if args.only == "c5":
        print(json.dumps({"c5": c5_leg(args.steps, args.warmup, 10)}))
        return
    if args.only == "c3":
        print(json.dumps({"c3": c3_leg(args.steps, 50, min(args.steps, args.c3_window),
                                       tuple(args.c3_paths.split(",")))}))
        return
    if args.only == "c4":
        n4 = args.c4_steps
        print(json.dumps({"c4": c4_leg(n4, n4 * 4 // 5, max(1, n4 // 10))}))
        return
    res = {}
    paths = [("fp32", FP32, PATH_FUSED, args.steps), ("fp64", FP64, PATH_FUSED, args.steps),
             ("plain", FP64, PATH_PLAIN, max(10, args.steps // 20))]
    if args.only:
        paths = [p for p in paths if p[0] == args.only]
    for label, prec, path, steps in paths:
        print(f"[bench] timing {label}: {args.warmup} warmup + {steps} steps", file=sys.stderr, flush=True)
        s = make_solver(inp, prec, path)
        wall, kt = timed(s, args.warmup if label != "plain" else 2, steps, kernel_timing=(path == PATH_FUSED))
        res[label] = dict(steps=steps, wall=wall, ups=upd_per_step * steps / wall, kt=kt,
                          tiles=s.tile_shape() if path == PATH_FUSED else None)
        s.close()
        print(f"[bench] {label}: {res[label]['ups']:.3e} updates/s", file=sys.stderr, flush=True)
    if args.only:
        print(json.dumps({k: dict(ups=v["ups"], kt=v["kt"], tiles=v["tiles"]) for k, v in res.items()}))
        return

    # accuracy leg: the full 2000-step configuration in FP32 and FP64 (fused)
    # and FP64 plain (bit-identical arithmetic to the CPU reference)
    print(f"[bench] accuracy leg: {args.accuracy_steps} steps x 3 paths", file=sys.stderr, flush=True)
    finals = {}
    for label, prec, path in [("fp32", FP32, PATH_FUSED), ("fp64", FP64, PATH_FUSED), ("plain", FP64, PATH_PLAIN)]:
        s = make_solver(inp, prec, path)
        s.solve(args.accuracy_steps)
        finals[label] = s.solution()
        clock = s.clock()
        s.close()
        print(f"[bench] accuracy {label} done (t = {clock[1]:.6f})", file=sys.stderr, flush=True)
    p64 = finals["plain"][4]
    pmax = float(np.max(np.abs(p64)))
    err32 = float(np.max(np.abs(finals["fp32"][4] - p64)) / pmax)
    err64 = float(np.max(np.abs(finals["fp64"][4] - p64)) / pmax)

    # dominant kernel: the class with the most time in the FP32 timed region
    kt = res["fp32"]["kt"]
    stage_classes = ["stage_first", "stage_mid", "stage_last", "filter"]
    dom = max(stage_classes, key=lambda c: kt[c][0])
    avg_ms = kt[dom][0] / kt[dom][1]
    b = 4  # FP32 bulk
    V, D = 4, 2
    per_point = {  # algorithmic bytes per point per launch: state arrays, FP64 metrics + J, sigma
        "stage_first": 3 * V * b + 8 * (D * D + 1) + b,      # read y_n; write acc, next input
        "stage_mid": 5 * V * b + 8 * (D * D + 1) + b,        # read in, y_n, acc; write acc, next input
        "stage_last": 3 * V * b + 8 * (D * D + 1) + b,       # read in, acc; write acc
        "filter": 2 * V * b,                                 # read y_{n+1}; write filtered
    }[dom]
    achieved = per_point * npts / (avg_ms * 1e-3)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_c2_fp32.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get(dom)
        except Exception:
            traffic = None

    c5 = None if args.no_c5 else c5_leg(args.c5_steps, args.warmup, 50)
    c3 = None if args.no_c3 else c3_leg(args.c3_steps, 50, args.c3_window)
    c4 = None if args.no_c4 else c4_leg(args.c4_steps, args.c4_steps * 4 // 5, max(1, args.c4_steps // 10))

    cpu = None
    if not args.no_cpu_baseline:
        print(f"[bench] CPU oracle baseline: {args.cpu_steps} steps, 1 thread", file=sys.stderr, flush=True)
        wall1 = cpu_baseline(inp, args.cpu_steps)
        one = upd_per_step * args.cpu_steps / wall1
        cores, model = host_cpu()
        nproc = max(1, min(cores, args.cpu_procs))
        steps_all = max(1, args.cpu_steps // 2)
        print(f"[bench] CPU oracle baseline: {nproc} processes x {steps_all} steps", file=sys.stderr, flush=True)
        wall_all = cpu_baseline_all_cores(nproc, steps_all)
        cpu = {"value": upd_per_step * steps_all * nproc / wall_all, "unit": "grid-point RK-stage updates/s",
               "cores": nproc, "kind": "port",
               "sample": f"the same 1025x1025 workload in FP64 by the oracle (oracle/lee_oracle.c, single-threaded): "
                         f"{nproc} processes at once, {steps_all} RK4 steps each ({wall_all:.1f} s); "
                         f"alone on one core {args.cpu_steps} steps ({wall1:.1f} s)",
               "single_core_value": one, "host_cores": cores, "cpu_model": model}

    # algorithmic bytes per update of the whole step, FP32 bulk, uniform mean:
    # 5 V b state + 8 (D^2 + 1) metrics + b sigma + V b / 2 filter
    algo_step_bytes = 5 * 4 * 4 + 8 * 5 + 4 + 8
    v = res["fp32"]["ups"]
    out = {
        "metric": "grid-point RK-stage updates/s + % HBM peak; FP32 vs FP64 at <=1e-5 rel p-error",
        "value": v,
        "unit": "grid-point RK-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["fp32"]["wall"] / res["fp32"]["steps"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 metrics/J/time)",
        "data": "synthetic (grid, base state and source built in bench.py)",
        "config": {"workload": "2-D curvilinear 1025x1025, M=0.5, monopole omega=0.2pi, RK4 CFL 0.5, "
                               "8th-order filter, 40-pt sponges (C2)",
                   "grid": [N1D, N1D], "steps_accuracy": args.accuracy_steps, "path": "fused",
                   "working_set_MB_per_stage": round(npts * (5 * 4 * 4 + 8 * 5) / 1e6, 1),
                   "working_set_note": "the stage kernels touch ~130 MB, which the 256 MB Infinity Cache can "
                                       "largely hold; batched_ensemble_c5 (12.4M points) gives the out-of-cache "
                                       "kernel fractions and copy_GBps the measured copy ceilings"},
        "hbm_frac_step": v * algo_step_bytes / HBM_PEAK,
        "fp64_fused_updates_per_s": res["fp64"]["ups"],
        "fp64_plain_updates_per_s": res["plain"]["ups"],
        "speedup_fp32_vs_fp64_fused": v / res["fp64"]["ups"],
        "speedup_fp32_vs_fp64_plain": v / res["plain"]["ups"],
        "p_rel_err_fp32_vs_fp64": err32,
        "p_rel_err_fp64_fused_vs_plain": err64,
        "kernel_ms": {k: [round(t, 4), c] for k, (t, c) in kt.items()},
        "tile_shape": {"stage_rows_per_thread": res["fp32"]["tiles"][0], "filter_tile_rows": res["fp32"]["tiles"][1]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "traffic_source": "profiles/traffic_c2_fp32.json: rocprofv3 FETCH_SIZE and WRITE_SIZE of "
                                       "this kernel in KiB per dispatch, scaled by a 256 MiB device copy "
                                       "(profiles/traffic_c2_fp32.md); not measured in this run",
                     "algorithmic_bytes_per_point": per_point, "avg_launch_ms": avg_ms,
                     "out_of_cache_frac": (c5 or {}).get("fp32", {}).get("kernel_hbm_frac", {}).get(dom)},
        "copy_GBps": ceiling,
        "cpu_baseline": cpu,
        "batched_ensemble_c5": c5,
        "overset_cylinder_c3": c3,
        "adjoint_3d_c4": c4,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
