"""Benchmark: Yee Gcell-updates/s of the forward FDTD step on one MI355X.

Default workload (the dielectric-box configuration): 256^3 physical cells (281^3
Yee nodes with CPML 12 on all faces), 40 seeded dielectric boxes (eps ~ U[1,12],
edge 8-64 cells), +z soft plane wave with a Gaussian pulse, two z-plane DFT flux
monitors at 8 frequencies over f0 +- 20%, Courant 0.99.  A "step" is one FDTD
time step of that scene (H+CPML, soft-source terms, E+CPML, monitor DFT).

Prints ONE JSON line.  ``value`` is node-updates per second over exactly
``--steps`` timed steps with all inputs resident in HBM; ``protocol`` adds the
two-point slope (T(2S) - T(S)) / S as the median of several rounds, twice (A/A),
so the spread of the two medians shows the noise floor; ``roofline`` prices the
step's algorithmic bytes against 8 TB/s; ``cpu_baseline`` times the float32
OpenMP transcription of the same update equations (oracle/fdtd_oracle.c) on a
bounded sample of the same scene on the host cores.

Other workloads: ``--scene vacuum|dispersive`` (forward), ``--workload gradient``
(one checkpointed design-region gradient), ``--workload batch`` (a design
gradient for each of 32 scenes: batched forward and stream-overlapped reverse
sweeps against one scene at a time).  ``--summarize stats DIR`` and ``--summarize
traffic FETCH_DIR WRITE_DIR`` turn rocprofv3 CSV output into the per-kernel
tables and HBM bytes per dispatch kept under profiles/.
"""

from __future__ import annotations

import argparse
import csv
import ctypes
import glob
import json
import os
import platform
import sys
import time
from collections import defaultdict

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
COPY_PROBE_BYTES = 256 * 1024 * 1024  # --copy-probe copies this many bytes per dispatch


# ------------------------------------------------------------------ host facts
def host_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": avail, "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


# ------------------------------------------------------------ algorithmic bytes
def algorithmic_bytes(plan, runner):
    """Compulsory float32 bytes per step and per E/H launch.

    Per step: every field component read and written once (8 B per sample), each
    non-uniform coefficient array read once, each psi slab sample read and
    written once, ADE pole state read+written where the pole acts, full-field DFT
    accumulators read+written.  Per launch: the same restricted to what that
    kernel must touch (its own family R+W, the other family read once)."""
    import numpy as np
    from witwin_maxwell_amd.fdtd.plan import COMPONENTS
    numel = {c: int(np.prod(plan.field_shape(c))) for c in COMPONENTS}
    e_n = sum(numel[c] for c in COMPONENTS[:3])
    h_n = sum(numel[c] for c in COMPONENTS[3:])
    coef_e = sum(4 * numel[c] for c in COMPONENTS[:3] for k in ("decay", "curl") if runner.uniform[(c, k)] is None)
    coef_h = sum(4 * numel[c] for c in COMPONENTS[3:] for k in ("decay", "curl") if runner.uniform[(c, k)] is None)
    psi_e = sum(8 * t.numel() for n, t in runner.psi.items() if n.startswith("psi_e"))
    psi_h = sum(8 * t.numel() for n, t in runner.psi.items() if n.startswith("psi_h"))
    # ADE: pole state R+W over the samples a pole acts on (Drude J: 8 B; Debye /
    # Lorentz P and J: 16 B), counted on the float32 drives the device reads;
    # drive and 1/eps_inf are per-material constants, so 0.  Dense pole arrays
    # show up as traffic above this.
    ade = 0
    for c in COMPONENTS[:3]:
        for pole in plan.ade[c]["poles"]:
            nnz = int(np.count_nonzero(np.asarray(pole["drive"], dtype=np.float32)))
            ade += nnz * (8 if pole["kind"] == "drude" else 16)
    dft = sum(2 * 8 * len(plan.dft_frequencies) * numel[c] for c in COMPONENTS[:3]) if plan.dft_frequencies else 0
    step = 8 * (e_n + h_n) + coef_e + coef_h + psi_e + psi_h + ade + dft
    e_launch = 8 * e_n + 4 * h_n + coef_e + psi_e + ade
    h_launch = 8 * h_n + 4 * e_n + coef_h + psi_h
    return step, e_launch, h_launch


def state_bytes(runner):
    """Bytes of one scene's whole state (fields, psi slabs, pole state)."""
    n = sum(t.numel() for t in runner.fields.values()) + sum(t.numel() for t in runner.psi.values())
    for states in runner.ade_state.values():
        for s in states:
            n += sum(v.numel() for v in s.values() if v is not None)
    return 4 * n


# ------------------------------------------------------------------- timing
def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def timed_run(runner, steps):
    """Seconds for ``runner.run(steps)`` between HIP events on the launch stream."""
    import torch
    start, end = _events()
    torch.cuda.synchronize()
    start.record()
    runner.run(steps)
    end.record()
    end.synchronize()
    return start.elapsed_time(end) * 1e-3


def slope_protocol(runner, span, rounds):
    """Two-point slope (T(2S) - T(S)) / S per round (run from a reset state each
    round), median over ``rounds``; done twice, the second set is the A/A run
    whose median shows the noise floor."""
    import numpy as np
    medians, all_slopes = [], []
    for _ in range(2):
        slopes = []
        for _ in range(rounds):
            runner.reset()
            t1 = timed_run(runner, span)
            runner.reset()
            t2 = timed_run(runner, 2 * span)
            slopes.append((t2 - t1) / span)
        all_slopes.append(slopes)
        medians.append(float(np.median(slopes)))
    mid = 0.5 * (medians[0] + medians[1])
    return {"method": "two-point slope (T(2S)-T(S))/S, median of rounds, A/A", "S": span, "rounds": rounds,
            "slope_ms_per_step": mid * 1e3, "median_a_ms": medians[0] * 1e3, "median_b_ms": medians[1] * 1e3,
            "aa_spread": abs(medians[0] - medians[1]) / mid if mid > 0 else None,
            "round_ms_a": [s * 1e3 for s in all_slopes[0]], "round_ms_b": [s * 1e3 for s in all_slopes[1]]}


def time_fused(runner, reps):
    """Average duration of one single-pass step launch: the program is copied with
    its monitors detached, so each step is exactly the step kernels on torch's
    current stream (with z_tiling 4 the z-end launch runs on a second stream that
    each step joins back into it), bracketed by HIP events."""
    from witwin_maxwell_amd.fdtd import native as nat
    prog = nat.MwProgram()
    ctypes.memmove(ctypes.byref(prog), ctypes.byref(runner.program), ctypes.sizeof(prog))
    prog.n_planes = 0
    prog.n_points = 0
    first = runner.step_index
    reps = min(reps, runner.plan.time_steps - first - 2)
    if reps < 1:
        raise RuntimeError(f"no steps left in the program for the launch timing (at step {first})")
    stream = nat.current_stream()
    nat.check(runner.lib.mw_run_steps(ctypes.byref(prog), first, 2, stream), runner.lib)
    start, end = _events()
    import torch
    torch.cuda.synchronize()
    start.record()
    nat.check(runner.lib.mw_run_steps(ctypes.byref(prog), first + 2, reps, stream), runner.lib)
    end.record()
    end.synchronize()
    return start.elapsed_time(end) * 1e-3 / reps


def time_kernel(runner, which, reps):
    """Average duration of one E or H launch, measured with HIP events on the
    stream the kernel is launched on (torch's current stream)."""
    import torch
    from witwin_maxwell_amd.fdtd import native as nat
    lib, prog = runner.lib, runner.program
    stream = nat.current_stream()
    ade = ctypes.byref(prog.ade) if prog.has_ade else None

    def launch():
        if which == "E":
            rc = lib.mw_update_electric(ctypes.byref(prog.grid), ctypes.byref(prog.coeffs), ctypes.byref(prog.cpml),
                                        ade, ctypes.byref(prog.fields), 7, stream)
        else:
            rc = lib.mw_update_magnetic(ctypes.byref(prog.grid), ctypes.byref(prog.coeffs), ctypes.byref(prog.cpml),
                                        ctypes.byref(prog.fields), 7, stream)
        nat.check(rc, lib)

    for _ in range(3):
        launch()
    start, end = _events()
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        launch()
    end.record()
    end.synchronize()
    return start.elapsed_time(end) * 1e-3 / reps


# --------------------------------------------------------------- CPU baseline
def _threads():
    return int(os.environ.get("OMP_NUM_THREADS", host_info()["nproc"]))


def cpu_baseline(plan, steps=8, warm=2):
    """Float32 OpenMP transcription (oracle/fdtd_oracle.c) on the host cores, run
    for warm + steps steps of the same plan."""
    import numpy as np
    from oracle.driver import OracleRun
    run = OracleRun(plan, precision="f32")
    for _ in range(warm):
        run.step()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.step()
    dt = time.perf_counter() - t0
    nodes = int(np.prod(plan.shape))
    threads = _threads()
    return {"value": nodes * steps / dt / 1e9, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "cpu_model": host_info()["cpu_model"],
            "sample": f"same {plan.shape[0]}x{plan.shape[1]}x{plan.shape[2]}-node scene, {steps} steps after "
                      f"{warm} warm-up steps, float32 OpenMP ({threads} threads), {dt:.2f} s"}


def cpu_gradient_rates(plan, steps=4):
    """Seconds per forward step and per reverse step of the float32 OpenMP
    transcription on the host cores (the reverse step without the window
    gradient, which is O(window))."""
    from oracle.driver import OracleAdjoint, OracleRun
    run = OracleRun(plan, precision="f32", record_monitors=False)
    run.step()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.step()
    t_f = (time.perf_counter() - t0) / steps
    adj = OracleAdjoint(run)
    adj.reverse_step(0, gradients=False)
    t0 = time.perf_counter()
    for _ in range(steps):
        adj.reverse_step(0, gradients=False)
    t_r = (time.perf_counter() - t0) / steps
    return t_f, t_r


# ----------------------------------------------------------------- workloads
def _flux_objective(monitors):
    e = monitors["objective"]
    return e["flux"] if "flux" in e else sum((e[c].abs() ** 2).sum() for c in ("Ex", "Ey") if c in e)


def bench_gradient(args):
    """Design-region gradient: checkpointed forward + native reverse sweep of the
    design-region configuration (225x225x145 nodes, 64x64x32 design voxels,
    3000 steps, checkpoints every 100), loss = |E|^2 of the central 40x40 of the
    objective plane at f0.

    Roofline: algorithmic bytes of the whole gradient (forward steps, the segment
    replay steps, reverse steps at the forward step's bytes -- adjoint fields and
    psi R+W, coefficients read -- and each checkpoint's snapshot and restore)
    over its wall time."""
    import numpy as np
    import torch
    from witwin_maxwell_amd import configs
    from witwin_maxwell_amd.fdtd.adjoint import differentiable_run
    from witwin_maxwell_amd.scene import MaterialRegion
    scene, kw = configs.design_region(steps=args.grad_steps)
    theta0 = kw.pop("theta")
    region = scene.material_regions[0]
    stride = 100

    def one_gradient():
        theta = theta0.clone().requires_grad_(True)
        scene.material_regions[0] = MaterialRegion(region.name, region.geometry, torch.sigmoid(theta),
                                                   eps_bounds=region.eps_bounds)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan, mons, session = differentiable_run(scene, stride=stride, tune=True, **kw)
        e = mons["objective"]
        loss = 0.0
        for c in ("Ex", "Ey"):
            d = e[c]
            a0, b0 = (d.shape[-2] - 40) // 2, (d.shape[-1] - 40) // 2
            loss = loss + (d[..., a0:a0 + 40, b0:b0 + 40].abs() ** 2).sum()
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        loss.backward()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, t_fwd, plan, session, theta

    one_gradient()  # warm-up (library load, first launches)
    total, t_fwd, plan, session, theta = one_gradient()
    T = plan.time_steps
    run = session.run
    nodes = int(np.prod(plan.shape))
    step_bytes, _, _ = algorithmic_bytes(plan, run)
    n_ckpt = len(range(stride, T, stride))
    alg = T * 3 * step_bytes + n_ckpt * 2 * 2 * state_bytes(run)
    roofline = {"bound": "hbm", "achieved": alg / total / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": alg / total / HBM_PEAK, "traffic": None,
                "kernel": "whole gradient (forward, replay and reverse step kernels)",
                "alg_bytes_per_gradient": alg, "step_alg_bytes_per_node": step_bytes / nodes}
    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline (forward and reverse steps of the float32 port)", file=sys.stderr, flush=True)
        t_f, t_r = cpu_gradient_rates(plan, steps=args.cpu_grad_steps)
        per_grad = T * (2 * t_f + t_r)
        cpu = {"value": 1.0 / per_grad, "unit": "gradients/s", "cores": _threads(), "kind": "port",
               "cpu_model": host_info()["cpu_model"],
               "sample": f"{args.cpu_grad_steps} forward steps ({t_f:.3f} s each) and {args.cpu_grad_steps} reverse "
                         f"steps ({t_r:.3f} s each) of the same {plan.shape[0]}x{plan.shape[1]}x{plan.shape[2]}-node "
                         f"scene, float32 OpenMP; gradient = {T} x (2 forward + 1 reverse) steps"}
    line = {"metric": "design gradients/s", "value": 1.0 / total, "unit": "gradients/s", "n_gpus": 1,
            "steps": T, "warmup": 1, "ms_per_step": total * 1e3 / T,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "design region 200x200x120 cells (+2x12 CPML), 64x64x32 design voxels, Ez dipole, "
                                   "|E|^2 of a 40x40 objective plane at f0, checkpoints every 100 steps",
                       "nodes": list(plan.shape)},
            "roofline": roofline, "cpu_baseline": cpu, "host": host_info(),
            "detail": {"forward_s": t_fwd, "backward_s": total - t_fwd,
                       "tiling": [int(run.program.xchunk), int(run.program.shell_chunk), int(run.program.z_tiling)],
                       "forward_gnodes_per_s": nodes * T / t_fwd / 1e9,
                       "grad_norm": float(theta.grad.norm())}}
    print(json.dumps(line), flush=True)


def bench_batch(args):
    """Batched forward+adjoint over the 32-scene batch configuration (seeded
    uneven sizes 64^3 .. 192x192x128 cells, CPML 10, 32x32x16 design region,
    dipole, single-frequency flux plane, ``--batch-steps`` steps, checkpoints
    every sqrt(T) steps): one design gradient per scene.

    Plain: one scene at a time (``GradientSession``: checkpointed forward, then
    the reverse sweep).  Batched (``BatchGradientSession``): the forward advances
    all scenes in the batched launches (autotuned x-chunk, graph length and z
    tiling), then each scene's reverse sweep runs on one of ``--batch-streams``
    streams.  Both take the gradient of sum |o|^2 over each scene's flux-plane
    sums; the per-scene window gradients of the two are compared.

    Roofline: algorithmic bytes of every scene's gradient (forward, replay and
    reverse steps at the forward step's bytes, and each checkpoint's snapshot
    and restore) over the batched wall time."""
    import numpy as np
    import torch
    from witwin_maxwell_amd import configs
    from witwin_maxwell_amd.fdtd.adjoint import GradientSession, default_stride
    from witwin_maxwell_amd.fdtd.batch import BatchGradientSession
    from witwin_maxwell_amd.fdtd.plan import compile_plan

    lo_cells, hi_cells = (tuple(int(v) for v in c.split("x")) for c in args.batch_cells.split(","))
    scs, kw = configs.batched_scenes(count=args.batch_count, steps=args.batch_steps, smallest=lo_cells,
                                     largest=hi_cells, design=tuple(min(d, s // 2) for d, s in zip((32, 32, 16),
                                                                                                      lo_cells)))
    plans = [compile_plan(sc, region_densities=[r.density.detach() for r in sc.material_regions], **kw)
             for sc, _ in scs]
    T = kw["time_steps"]
    K = default_stride(T)
    n = len(plans)
    nodes = sum(int(np.prod(p.shape)) for p in plans)

    def seeds(outs):  # d/do of sum o^2
        return [2.0 * o.detach() for o in outs]

    def window_grads(res):
        g_decay, g_curl = res
        return [g.detach().cpu() for g in list(g_decay) + list(g_curl)]

    # plain: one scene at a time (the first scene once untimed: library load, first launches)
    warm = GradientSession(plans[0], "cuda", stride=K)
    warm.backward(seeds(warm.forward()))
    del warm
    lat, plain, alg = [], [], 0
    for p in plans:
        sess = GradientSession(p, "cuda", stride=K)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = sess.backward(seeds(sess.forward()))
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        plain.append(window_grads(res))
        step_bytes = algorithmic_bytes(p, sess.run)[0]
        alg += T * 3 * step_bytes + len(range(K, T, K)) * 2 * 2 * state_bytes(sess.run)
        del sess, res
    torch.cuda.empty_cache()
    t_plain = sum(lat)
    print(f"[bench] plain: {n} gradients in {t_plain:.2f}s", file=sys.stderr, flush=True)
    # batched
    session = BatchGradientSession(plans, "cuda", stride=K, streams=args.batch_streams)
    choice = session.batch.autotune(trial_steps=16)
    tune_table = _fastest(getattr(session.batch, "autotune_timings", None))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = session.forward()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    res = session.backward([seeds(o) for o in outs])
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    batched = [window_grads(r) for r in res]
    blat = [t_fwd + d for d in session.scene_done_s]
    err = 0.0
    for a, b in zip(plain, batched):
        num = sum(float((x.double() - y.double()).norm() ** 2) for x, y in zip(a, b)) ** 0.5
        den = sum(float(x.double().norm() ** 2) for x in a) ** 0.5
        err = max(err, num / max(den, 1e-300))
    del session, res, outs
    torch.cuda.empty_cache()
    roofline = {"bound": "hbm", "achieved": alg / t_batch / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": alg / t_batch / HBM_PEAK, "traffic": None,
                "kernel": "every scene's gradient (batched forward, per-scene replay and reverse step kernels)",
                "alg_bytes_per_batch": alg}
    cpu = None
    if not args.no_cpu:
        mid = plans[int(np.argsort([int(np.prod(p.shape)) for p in plans])[len(plans) // 2])]
        print("[bench] cpu baseline (forward and reverse steps of the float32 port)", file=sys.stderr, flush=True)
        t_f, t_r = cpu_gradient_rates(mid, steps=args.cpu_grad_steps)
        mid_nodes = int(np.prod(mid.shape))
        per_node_step = (2 * t_f + t_r) / mid_nodes
        per_batch = per_node_step * nodes * T
        cpu = {"value": n / per_batch, "unit": "gradients/s", "cores": _threads(), "kind": "port",
               "cpu_model": host_info()["cpu_model"],
               "sample": f"{args.cpu_grad_steps} forward steps ({t_f:.3f} s each) and {args.cpu_grad_steps} reverse "
                         f"steps ({t_r:.3f} s each) of the median-size scene {mid.shape}, float32 OpenMP; scaled "
                         f"per node to the batch's {nodes} nodes x {T} x (2 forward + 1 reverse) steps"}
    line = {"metric": "batched fwd+adjoint grads/s vs plain", "value": n / t_batch, "unit": "grads/s", "n_gpus": 1,
            "steps": T, "warmup": 1, "ms_per_step": t_batch * 1e3 / T, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"{n} scenes {args.batch_cells.replace(',', '..')} cells (+2x10 CPML), design "
                                   f"region, dipole, single-frequency flux plane, {T} steps, checkpoints every {K} "
                                   f"steps, one gradient per scene",
                       "total_nodes": nodes, "xchunk_graph_steps_z_tiling": list(choice),
                       "autotune_ms_per_step": tune_table,
                       "reverse_streams": args.batch_streams},
            "roofline": roofline, "cpu_baseline": cpu, "host": host_info(),
            "detail": {"batched_s": t_batch, "batched_forward_s": t_fwd, "plain_s": t_plain,
                       "speedup_vs_plain": t_plain / t_batch, "plain_grads_per_s": n / t_plain,
                       "batched_latency_p50_s": float(np.percentile(blat, 50)),
                       "batched_latency_p99_s": float(np.percentile(blat, 99)),
                       "plain_latency_p50_s": float(np.percentile(lat, 50)),
                       "plain_latency_p99_s": float(np.percentile(lat, 99)),
                       "max_rel_gradient_diff_batched_vs_plain": err}}
    print(json.dumps(line), flush=True)


# ----------------------------------------------------- rocprofv3 summaries
def _find(directory, suffix):
    hits = sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {directory}")
    return hits


def kernel_stats(directory):
    """Per-kernel calls, average and total duration from a ``rocprofv3
    --kernel-trace --stats --output-format csv`` run."""
    rows = []
    for path in _find(directory, "kernel_stats.csv"):
        with open(path) as fh:
            rows.extend(csv.DictReader(fh))
    out = {}
    for r in rows:
        name = r.get("Name") or r.get("KernelName") or r.get("Kernel_Name")
        calls = int(float(r.get("Calls", 0)))
        avg = float(r.get("AverageNs", r.get("Average", 0)))
        tot = float(r.get("TotalDurationNs", r.get("TotalDuration", 0)))
        out[name] = {"calls": calls, "avg_us": avg / 1e3, "total_ms": tot / 1e6}
    return out


def _short(name):
    """Kernel family of a rocprofv3 kernel name (template arguments kept for the
    step kernels, whose variants are timed separately)."""
    if "__amd_rocclr_copyBuffer" in name:
        return "copy"
    for key in ("lean_step_kernel", "lean_batch_kernel", "zedge_step_kernel", "zedge_batch_kernel",
                "zends_step_kernel"):
        if key in name:
            return name[name.index(key):].split("(")[0]
    return name.split("(")[0][:80]


def counter_per_dispatch(directory, counter):
    """Average counter value per dispatch of each kernel family; for the copy
    probe, the largest dispatch (the probe's 256 MiB copy, not the small
    device-to-device copies a run also makes)."""
    vals = defaultdict(list)
    for path in _find(directory, "counter_collection.csv"):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") != counter:
                    continue
                vals[_short(r.get("Kernel_Name", ""))].append(float(r["Counter_Value"]))
    return {k: (max(v) if k == "copy" else sum(v) / len(v)) for k, v in vals.items()}


def traffic_summary(fetch_dir, write_dir):
    """HBM bytes per dispatch: FETCH_SIZE + WRITE_SIZE (KiB) of each kernel, each
    scaled by bytes / counter reading of the copy probe, which reads and writes
    COPY_PROBE_BYTES."""
    fetch = counter_per_dispatch(fetch_dir, "FETCH_SIZE")
    write = counter_per_dispatch(write_dir, "WRITE_SIZE")
    if not fetch.get("copy") or not write.get("copy"):
        raise RuntimeError("the runs hold no copy-probe dispatch (run bench.py with --copy-probe)")
    f_scale = COPY_PROBE_BYTES / (fetch["copy"] * 1024.0)
    w_scale = COPY_PROBE_BYTES / (write["copy"] * 1024.0)
    out = {"_calibration": {"fetch_scale": f_scale, "write_scale": w_scale}}
    for k in fetch:
        if k == "copy":
            continue
        f = fetch[k] * 1024.0 * f_scale
        w = write.get(k, 0.0) * 1024.0 * w_scale
        out[k] = {"fetch_bytes": f, "write_bytes": w, "total_bytes": f + w}
    return out


def step_busy(directory):
    """(steps, union, sum): the single-pass step kernels of a kernel trace, steps
    counted by the shell-E dispatches; ``union`` is the time in microseconds per
    step during which at least one of them runs (launches that overlap on two
    streams count once), ``sum`` the sum of their durations per step."""
    spans, steps = [], 0
    for path in _find(directory, "kernel_trace.csv"):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                name = r.get("Kernel_Name", "")
                if "_batch_kernel" in name or not any(f in name for f in _STEP_FAMILIES):
                    continue
                steps += "shell_e_kernel" in name
                spans.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    if not steps:
        return None
    spans.sort()
    union, total, end = 0, 0, None
    for a, b in spans:
        total += b - a
        if end is None or a >= end:
            union += b - a
            end = b
        elif b > end:
            union += b - end
            end = b
    return steps, union / steps / 1e3, total / steps / 1e3


def summarize(argv):
    if argv[0] == "stats":
        for k, v in sorted(kernel_stats(argv[1]).items(), key=lambda kv: -kv[1]["total_ms"]):
            print(f"{v['calls']:8d} calls  avg {v['avg_us']:10.2f} us  total {v['total_ms']:10.2f} ms  {k}")
        try:
            busy = step_busy(argv[1])
        except FileNotFoundError:
            busy = None
        if busy:
            print(f"step kernels over {busy[0]} steps: {busy[1]:.2f} us per step with at least one running "
                  f"(sum of their durations {busy[2]:.2f} us)")
    elif argv[0] == "traffic":
        summary = traffic_summary(argv[1], argv[2])
        if len(argv) > 3:
            # the committed form that the forward workload reads back as roofline.traffic:
            # the step kernels' bytes per step, keyed by the tiling they were measured at
            calib = summary.pop("_calibration")
            step = step_traffic(summary)
            summary = {"by_tiling": {argv[3]: {STEP_KERNEL: step, "_per_kernel_bytes": summary,
                                               "_calibration": calib}}}
        print(json.dumps(summary, indent=1))
    else:
        raise SystemExit("--summarize stats DIR | traffic FETCH_DIR WRITE_DIR [XCHUNK,SHELL_CHUNK,Z_TILING]")


STEP_KERNEL = "single-pass step (interior, z-end / z-edge, PML-shell H and E, term and aux-line kernels)"
_STEP_FAMILIES = ("lean_step_kernel", "shell_h_lean_kernel", "zends_step_kernel", "zedge_step_kernel",
                  "shell_h_kernel", "shell_e_kernel", "shifted_values_kernel", "aux_line_h_kernel",
                  "aux_line_e_kernel")


def step_traffic(per_kernel):
    """HBM bytes of one single-pass step: every kernel mw_run_steps launches for
    the step itself (the monitors' gather/fold excluded, as in the step timing)."""
    return sum(v["total_bytes"] for k, v in per_kernel.items() if any(f in k for f in _STEP_FAMILIES))


def _fastest(timings, n=6):
    """The n fastest autotuner candidates as {"a,b,c": ms per step} (None if none)."""
    if not timings:
        return None
    return {",".join(str(v) for v in c): round(t, 4) for c, t in sorted(timings.items(), key=lambda kv: kv[1])[:n]}


def _traffic_table(path):
    """{tiling tuple: entry} of a committed counter summary (profiles/traffic_*.json)."""
    if not path or not os.path.exists(path):
        return {}
    try:
        data = json.load(open(path))
    except (OSError, ValueError):
        return {}
    return {tuple(int(v) for v in k.split(",")): e for k, e in data.get("by_tiling", {}).items()}


def _traffic_for(tiling, path, kernel):
    """Measured HBM bytes of the step from a committed counter summary, only when
    it was measured at the tiling this run uses.  z_tiling 4 launches the kernels
    of z_tiling 3 (the z-end launch on a second stream) and moves the same bytes
    per launch."""
    if tiling is None:
        return None
    table = _traffic_table(path)
    entry = table.get(tuple(tiling))
    if entry is None and len(tiling) == 3 and int(tiling[2]) == 4:
        entry = table.get((tiling[0], tiling[1], 3))
    return None if entry is None else entry.get(kernel)


# --------------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cells", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--cpu-grad-steps", type=int, default=4)
    ap.add_argument("--slope-steps", type=int, default=100, help="S of the two-point slope (0: skip the protocol)")
    ap.add_argument("--slope-rounds", type=int, default=5)
    ap.add_argument("--two-kernel", action="store_true", help="use the two-launch (H, E) step instead of the single pass")
    ap.add_argument("--no-tune", action="store_true", help="keep the library's default single-pass tiling")
    ap.add_argument("--plane-stage-depth", type=int, default=32,
                    help="steps per block of staged z-normal plane monitors (0: sample them after every step)")
    ap.add_argument("--tiling", default="",
                    help="XCHUNK,SHELL_CHUNK[,Z_TILING]: fixed single-pass tiling instead of autotuning")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run torch copies of 256 MiB (counter calibration for rocprofv3 --pmc runs)")
    ap.add_argument("--workload", choices=("forward", "gradient", "batch"), default="forward",
                    help="forward: Yee step of the dielectric-box scene (default); gradient: one design-region "
                         "gradient; batch: batched vs one-scene-per-launch forward runs")
    ap.add_argument("--scene", choices=("boxes", "vacuum", "dispersive", "design"), default="boxes",
                    help="forward workload scene: dielectric boxes 281^3 nodes (default), vacuum dipole 85^3 with "
                         "full-field DFT, dispersive spheres 409x409x281 with TFSF, design-region scene "
                         "225x225x145 (the gradient workload's forward step)")
    ap.add_argument("--grad-steps", type=int, default=3000)
    ap.add_argument("--batch-count", type=int, default=32)
    ap.add_argument("--batch-cells", default="64x64x64,192x192x128",
                    help="SMALLEST,LARGEST scene cells of the batch workload (default: the 32-scene configuration)")
    ap.add_argument("--batch-steps", type=int, default=1500)
    ap.add_argument("--batch-streams", type=int, default=8, help="streams of the batched reverse sweeps")
    ap.add_argument("--summarize", nargs="+", metavar="ARG",
                    help="stats DIR | traffic FETCH_DIR WRITE_DIR: summarise rocprofv3 CSV output and exit")
    args = ap.parse_args()
    if args.summarize:
        return summarize(args.summarize)
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the FDTD path runs on a single device).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "gradient":
        return bench_gradient(args)
    if args.workload == "batch":
        return bench_batch(args)

    import numpy as np
    import torch
    from witwin_maxwell_amd import configs
    from witwin_maxwell_amd.fdtd.plan import compile_plan
    from witwin_maxwell_amd.fdtd.runtime import DeviceRun

    total = args.warmup + args.steps
    if args.scene == "boxes":
        scene, kw = configs.dielectric_boxes(cells=args.cells, pml=12, n_boxes=40, steps=max(2000, total))
        workload = (f"dielectric boxes {args.cells}^3 cells (+2x12 CPML), 40 boxes eps U[1,12], "
                    f"soft +z plane wave, 2 flux monitors x 8 freq")
        cells = args.cells ** 3
        traffic_file = "traffic_dielectric_boxes_256.json" if args.cells == 256 else None
    elif args.scene == "design":
        scene, kw = configs.design_region(steps=max(3000, total))
        kw.pop("theta")
        workload = ("design region 200x200x120 cells (+2x12 CPML), 64x64x32 design voxels, Ez dipole, objective "
                    "plane at f0, forward step")
        cells = 200 * 200 * 120
        traffic_file = None
    elif args.scene == "vacuum":
        scene, kw = configs.vacuum_dipole(cells=64, pml=10, steps=max(500, total))
        workload = ("vacuum 64^3 cells (+2x10 CPML), centred Ez ideal dipole, full-field DFT at 3 frequencies "
                    "+ point monitor")
        cells = 64 ** 3
        traffic_file = None
    else:
        scene, kw = configs.dispersive_spheres(steps=max(4000, total))
        workload = ("dispersive spheres 384x384x256 cells (+2x12 CPML), 30% Drude + 2-pole Lorentz, eps U[1,4] "
                    "background, TFSF slab plane wave, flux at 16 frequencies")
        cells = 384 * 384 * 256
        traffic_file = "traffic_dispersive_spheres_409.json"
    kw = dict(kw)
    # room for the event-timed single launches after the timed run and for the
    # slope rounds (each from a reset state)
    reps = max(20, min(200, args.steps // 5))
    kw["time_steps"] = max(int(kw["time_steps"]), total + reps + 2, 2 * args.slope_steps + 2)
    t0 = time.perf_counter()
    plan = compile_plan(scene, **kw)
    compile_s = time.perf_counter() - t0
    runner = DeviceRun(plan, "cuda:0", fused=not args.two_kernel, plane_stage_depth=args.plane_stage_depth)
    nodes = int(np.prod(plan.shape))
    tiling = None
    if runner.fused and args.tiling:
        tiling = tuple(int(v) for v in args.tiling.split(","))
        if len(tiling) == 2:
            tiling = tiling + (0,)
        runner.set_tiling(*tiling)
    elif runner.fused and not args.no_tune:
        # launch geometry of the single-pass step, timed on this run and reset
        # ties (1%) go to a tiling whose HBM traffic is on record (z_tiling 4 moves the
        # bytes of z_tiling 3)
        prefer = sorted(_traffic_table(os.path.join(ROOT, "profiles", traffic_file))) if traffic_file else []
        prefer = [(x, sc, 4) for x, sc, z in prefer if z == 3] + prefer
        tiling = runner.autotune(trial_steps=12, cache=False, prefer=prefer)
        print(f"[bench] tiling (xchunk, shell_chunk, z_tiling) = {tiling}", file=sys.stderr, flush=True)
    print(f"[bench] plan {plan.shape} compiled in {compile_s:.1f}s; warmup {args.warmup} steps", file=sys.stderr,
          flush=True)
    runner.run(args.warmup)
    torch.cuda.synchronize()
    elapsed = timed_run(runner, args.steps)
    torch.cuda.synchronize()
    ms_per_step = elapsed * 1e3 / args.steps
    value = nodes * args.steps / elapsed / 1e9
    print(f"[bench] {args.steps} steps in {elapsed:.3f}s ({ms_per_step:.4f} ms/step)", file=sys.stderr, flush=True)

    step_bytes, e_bytes, h_bytes = algorithmic_bytes(plan, runner)
    if runner.fused:
        # the single-pass step: PML-shell H, lean interior, PML-shell E (+ the tiny
        # per-step term prep); rocprof splits the time between them
        dominant, t_dom, b_dom = (STEP_KERNEL, time_fused(runner, reps), step_bytes)
    else:
        t_e = time_kernel(runner, "E", reps)
        t_h = time_kernel(runner, "H", reps)
        dominant, t_dom, b_dom = ("update_e_kernel", t_e, e_bytes) if t_e >= t_h else ("update_h_kernel", t_h, h_bytes)
    achieved = b_dom / t_dom
    traffic = None
    if runner.fused and traffic_file:
        traffic = _traffic_for(tiling, os.path.join(ROOT, "profiles", traffic_file), dominant)
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": dominant,
                "kernel_avg_us": t_dom * 1e6, "alg_bytes_per_launch": b_dom,
                "step_alg_bytes_per_node": step_bytes / nodes,
                "step_frac": step_bytes * args.steps / elapsed / HBM_PEAK}
    protocol = None
    if args.slope_steps > 0:
        protocol = slope_protocol(runner, args.slope_steps, args.slope_rounds)
        print(f"[bench] slope {protocol['slope_ms_per_step']:.4f} ms/step, A/A spread {protocol['aa_spread']:.3%}",
              file=sys.stderr, flush=True)
    if args.copy_probe:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(plan, steps=args.cpu_steps)
    line = {"metric": "Yee Gcell-updates/s (% HBM peak)", "value": value, "unit": "Gcell-updates/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": workload,
                       "nodes": list(plan.shape), "courant": 0.99,
                       "step": "single-pass" if runner.fused else "two-kernel",
                       "tiling": None if tiling is None else dict(zip(("xchunk", "shell_chunk", "z_tiling"), tiling)),
                       "autotune_ms_per_step": _fastest(getattr(runner, "autotune_timings", None))},
            "physical_gcell_updates_per_s": cells * args.steps / elapsed / 1e9,
            "roofline": roofline, "protocol": protocol, "cpu_baseline": cpu, "host": host_info()}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
