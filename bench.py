"""Time-step throughput of the MI355X path: cell-updates/s and the HBM roofline of the dominant kernel.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload arcade2048|arcade|periodic|periodic2048|moc2048|batch2f|adjoint|adjoint_arcade|copy]

One "step" is one PlasmaDomain::advanceTime (evolution.cpp:59-82) over the whole grid, ghost cells
included: every integrator stage, every module subcycle, the ghost fill and the dt reduction.
Inputs are synthetic (spruce_amd.workloads) and resident in HBM before timing starts; the timed
region is steps W..W+K-1, bracketed by device synchronisation.  Rank 0 prints one JSON line.

The default workload is the 2048x2048 solar arcade (ideal MHD + thermal conduction + radiative
losses, RK2, FP64).  The CPU baseline is the reference executable compiled from its own sources
(oracle/_ref/run), started from the GPU's state after W steps and timed over step W+1 between its
own per-step output lines (setup and file output excluded).
"""
from __future__ import annotations

import argparse
import json
import os
import queue
import subprocess
import sys
import tempfile
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# Algorithmic bytes per cell and per stage launch (FP64 = 8 B): the stage reads the evolved fields
# E of its source set (with halo), the E base values (later stages), the S static fields, and
# writes E.  IdealMHD: E = 8 (rho, mom_x/y/z, thermal_energy, bi_x/y/z), S = 5 (be_x/y/z, grav_x/y).
E_MHD, S_MHD = 8, 5


def measured_traffic(workload: str, kernel: str, rows: str, summary: str = "hbm_traffic_mhd_stage.md"):
    """HBM bytes per stage launch from a committed FETCH_SIZE/WRITE_SIZE counter summary under
    profiles/ (lines `traffic <workload> <kernel> [<rows>] <bytes>`), for the same workload and
    launch shape; None otherwise."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    want = [workload, kernel] + ([rows] if rows else [])
    for line in open(path):
        f = line.split()
        if len(f) == len(want) + 2 and f[0] == "traffic" and f[1:-1] == want:
            return float(f[-1])
    return None


def resolved_stage_kernel(option: str, nx: int, ny: int) -> str:
    """The stage kernel the library runs for --stage-kernel auto: strips when 4-row strips of 60
    output columns give at least 1024 wavefronts, else 8x8 tiles."""
    if option != "auto":
        return option
    return "strip" if ((ny + 59) // 60) * ((nx + 3) // 4) >= 1024 else "tile"


def stage_bytes_per_cell(integrator: str) -> float:
    """Average algorithmic bytes per cell of one stage launch."""
    if integrator == "rk2":     # (E+S read, E write) + (2E+S read, E write) over 2 launches
        return 8.0 * (5 * E_MHD + 2 * S_MHD) / 2
    if integrator == "euler":
        return 8.0 * (2 * E_MHD + S_MHD)
    raise ValueError(integrator)


def kernel_bytes_per_cell(name: str, integrator: str):
    """Algorithmic bytes per cell of one launch of a kernel of the IdealMHD step (FP64), or None."""
    table = {
        "mhd_stage": stage_bytes_per_cell(integrator),
        "tc_step": 40.0,          # one Euler conduction subcycle: te, n, b_hat_x, b_hat_y read, te written
        "rl_iterate": 24.0,       # all radiative subcycles of a step: te, n read, te written
        "mhd_dt": 88.0,           # 8 evolved + be_x/y/z read
        "bhat": 64.0,             # be_x/y/z, bi_x/y/z read, b_hat_x/y written
        "propagate_live": 48.0,   # rho, te read; n, rho, te written
    }
    return table.get(name)


def make_workload(name: str):
    from spruce_amd import workloads as W
    from spruce_amd.spec import spec_from_files
    if name == "periodic":
        st = W.periodic_wave_state(256, 256, seed=0)
        cfg = W.periodic_wave_config(steps=1000)
        desc = {"workload": "periodic ideal-MHD 256x256 (RK2, FP64)", "grid": [256, 256],
                "equation_set": "ideal_mhd", "integrator": "rk2", "modules": []}
    elif name == "periodic2048":
        st = W.periodic_wave_state(2048, 2048, d=1.25e6, seed=0)
        cfg = W.periodic_wave_config(steps=1000)
        desc = {"workload": "periodic ideal-MHD 2048x2048 (RK2, FP64)", "grid": [2048, 2048],
                "equation_set": "ideal_mhd", "integrator": "rk2", "modules": []}
    elif name == "moc2048":
        # the same wave with open characteristic (open_moc) boundaries on both y sides
        st = W.periodic_wave_state(2048, 2048, d=1.25e6, seed=0)
        cfg = W.base_config("ideal_mhd", ("periodic", "periodic", "open_moc", "open_moc"), "rk2", epsilon=0.2,
                            max_iterations=1000, density_min=1.0, temp_min=1.0, thermal_energy_min=1e-12)
        desc = {"workload": "ideal-MHD 2048x2048 wave, periodic x, open_moc y (RK2, FP64)", "grid": [2048, 2048],
                "equation_set": "ideal_mhd", "integrator": "rk2", "modules": []}
    elif name in ("arcade", "arcade2048"):
        # 2048 rows stretched by 1.002 per row would reach 1.5e11 cm: the explicit conduction
        # subcycles, whose time scale uses d_x*d_y, go unstable in the tall, flat top cells (NaN in
        # the first step, in the reference's scheme as much as here).  The 2048x2048 arcade
        # stretches by 1.0005 per row instead, which spans the 512x1024 arcade's height (1.8e10 cm),
        # and keeps that arcade's dipole depth (6.4e8 cm) under a domain 4x as wide.
        if name == "arcade":
            nx, ny, stretch, depth = 512, 1024, 1.002, None
        else:
            nx, ny, stretch, depth = 2048, 2048, 1.0005, 6.4e8
        st = W.arcade_state(nx, ny, stretch=stretch, depth=depth)
        cfg = W.arcade_config(steps=2000)
        desc = {"workload": "solar arcade %dx%d ideal-MHD + thermal conduction + radiative losses (RK2, FP64)"
                % (nx, ny), "grid": [nx, ny], "dy_stretch": stretch, "equation_set": "ideal_mhd", "integrator": "rk2",
                "modules": ["thermal_conduction", "radiative_losses"]}
    else:
        raise SystemExit("unknown workload " + name)
    return st, cfg, spec_from_files(cfg, st), desc


def host_cpu() -> dict:
    """CPU model, logical CPUs this process may run on, and threads per core, from lscpu."""
    info = {"model": "unknown", "cpus_available": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
            elif k == "Socket(s)":
                info["sockets"] = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def cpu_threads() -> int:
    """OpenMP threads of the CPU leg: SPRUCE_CPU_THREADS, else OMP_NUM_THREADS, else the CPUs this
    process may use (at most 16)."""
    for k in ("SPRUCE_CPU_THREADS", "OMP_NUM_THREADS"):
        if os.environ.get(k, "").isdigit() and int(os.environ[k]) > 0:
            return int(os.environ[k])
    return max(1, min(16, len(os.sched_getaffinity(0))))


# glibc serves each 32 MiB grid of a 2048x2048 run with a fresh mmap by default, so every whole-grid
# operation of the reference first takes page faults on 32 MiB; these tunables keep freed grids in
# the heap instead.  Both settings are the reference binary unchanged.
MALLOC_HEAP_ENV = {"MALLOC_MMAP_MAX_": "0", "MALLOC_TRIM_THRESHOLD_": "137438953472", "MALLOC_TOP_PAD_": "1073741824"}


def reference_step_times(st, cfg, nsteps: int, threads: int, env_extra=None, line_timeout: float = 170.0):
    """Runs the reference executable on (st, cfg) for nsteps + 1 steps with one progress line per step
    and stops it after the last one.  Returns [(seconds, line)] for steps 1..nsteps: the wall time
    between consecutive per-step lines, i.e. one advanceTime each, without set-up or file output
    (evolution.cpp:26-77 prints the line at the end of every step).  A run that prints nothing for
    line_timeout seconds is stopped and reported as an error."""
    exe = os.path.join(ROOT, "oracle", "_ref", "run")
    if not os.path.exists(exe):
        raise RuntimeError("reference executable oracle/_ref/run is not built")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    env.update(env_extra or {})
    c = cfg
    c.settings["max_iterations"] = str(nsteps + 1)
    c.settings["iter_output_interval"] = "-1"
    c.settings["time_output_interval"] = "-1"
    c.settings["write_interval"] = "-1"
    c.settings["std_out_interval"] = "1"
    with tempfile.TemporaryDirectory() as tmp:
        # the state goes in as <output>/init.state, which input mode reads without rewriting it
        # (main.cpp:44-56; plasmadomain.cpp:62-65)
        out = os.path.join(tmp, "o")
        os.makedirs(out)
        print("cpu baseline: writing the %dx%d state" % (st.xdim, st.ydim), file=sys.stderr, flush=True)
        st.write(os.path.join(out, "init.state"))
        cp = os.path.join(tmp, "w.config")
        c.write(cp)
        print("cpu baseline: reference run of %d steps, %d threads" % (nsteps + 1, threads), file=sys.stderr,
              flush=True)
        p = subprocess.Popen([exe, "-m", "input", "-o", out, "-c", cp],
                             stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env, cwd=tmp)
        lines = queue.Queue()

        def reader():
            for ln in p.stdout:
                lines.put((time.perf_counter(), ln))
            lines.put((time.perf_counter(), None))

        threading.Thread(target=reader, daemon=True).start()
        stamps, tail = [], []
        try:
            while True:
                try:
                    now, line = lines.get(timeout=line_timeout)
                except queue.Empty:
                    raise RuntimeError("the reference printed nothing for %d s (step %d); stopped"
                                       % (line_timeout, len(stamps)))
                if line is None:
                    break
                line = line.rstrip()
                tail = (tail + [line])[-5:]
                print("  reference: " + line[:200], file=sys.stderr, flush=True)
                if line.startswith("Iter: "):
                    stamps.append((now, line))
                    if len(stamps) == nsteps + 1:
                        break
        finally:
            if p.poll() is None:
                p.kill()
            p.wait()
    if len(stamps) < nsteps + 1:
        raise RuntimeError("the reference stopped after %d of %d steps (rc %s): %s"
                           % (len(stamps), nsteps + 1, p.returncode, " | ".join(tail)))
    for k, (_, line) in enumerate(stamps):
        if not line.startswith("Iter: %d/" % k):
            raise RuntimeError("unexpected reference progress line %r" % line)
    return [(stamps[k][0] - stamps[k - 1][0], stamps[k][1]) for k in range(1, nsteps + 1)]


def module_count(line: str, label: str):
    """The integer after `label: ` in a reference progress line (e.g. Thermal Subcycles), or None."""
    for part in line.split("|"):
        k, _, v = part.partition(":")
        if k.strip() == label:
            try:
                return int(v.strip())
            except ValueError:
                return None
    return None


def state_of(dom, base):
    """The GPU domain's current state as a .state file on the mesh and statics of `base`."""
    from spruce_amd.stateio import StateFile
    grids = dict(base.grids)
    for v in ("rho", "temp", "mom_x", "mom_y", "mom_z", "bi_x", "bi_y", "bi_z"):
        grids[v] = dom.grid(v)
    return StateFile(base.xdim, base.ydim, base.ion_mass, base.adiabatic_index, dom.time, grids,
                     list(base.comments))


def tf_stage_bytes_per_cell(curl: bool) -> float:
    """Two-fluid RK2: 14 evolved fields, statics grav_x/y (+ be_x/y with the Maxwell curl terms).
    Stage 1 reads E + S and writes E, stage 2 reads 2E + S and writes E; average of the two."""
    e, st = 14, 4 if curl else 2
    return 8.0 * (5 * e + 2 * st) / 2


def bench_batch(args) -> dict:
    """64 two-fluid expansions of uneven sizes stepped as one batch (graph replay, autotuned block),
    against the same domains stepped one after another through the single-domain entry point."""
    from spruce_amd import workloads as W
    from spruce_amd.gpu import GpuBatch, GpuDomain
    from spruce_amd.spec import spec_from_files

    members = W.batch_2f_members(64)
    specs = [spec_from_files(c, s) for s, c in members]
    cells = sum(s.xdim * s.ydim for s in specs)
    doms = [GpuDomain(s) for s in specs]
    b = GpuBatch(doms, graph=True)
    shape, tune_ms = b.autotune(3)
    b.step(args.warmup)
    t0 = time.perf_counter()
    b.step(args.steps)          # returns after a device synchronisation
    wall = time.perf_counter() - t0
    value = cells * args.steps / wall

    b.set_profiling(True)
    b.step(max(5, min(args.steps, 20)))
    kt = b.kernel_times()
    b.set_profiling(False)

    # the same batch with all steps of the call in one cooperative launch
    b.set_option("persistent", "true")
    b.step(args.warmup)
    t0 = time.perf_counter()
    b.step(args.steps)
    pwall = time.perf_counter() - t0
    b.close()

    # sequential: each domain advanced alone (one launch sequence per domain and step), with the
    # default launch shape and with the batch's tuned shape; the faster of the two is the baseline
    seq_steps = max(5, min(args.steps, 100))
    seq = {}
    for sh in ("64x4", shape):
        if sh in seq:
            continue
        seq_wall = 0.0
        for d in doms:
            d.set_option("block", sh)
            d.step(2)
            t0 = time.perf_counter()
            d.step(seq_steps)
            seq_wall += time.perf_counter() - t0
        seq[sh] = cells * seq_steps / seq_wall
    seq_value = max(seq.values())
    for d in doms:
        d.close()

    stage_ms = kt.get("tf_stage", (float("nan"), 0))[0]
    bpc = tf_stage_bytes_per_cell(False)
    achieved = bpc * cells / (stage_ms * 1e-3) / 1e9
    edges = {}
    for s in specs:
        edges[s.xdim] = edges.get(s.xdim, 0) + 1
    out = {
        "metric": "cell-updates/s and % HBM roofline (FP64); adjoint/forward cost ratio",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "64 two-fluid plasma expansions, uneven square grids (ideal_2F, RK2, open_ucnp, "
                               "FP64), one batch, graph replay, autotuned block",
                   "domains": len(specs), "cells": cells, "edges": edges, "block": shape},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("batch2f", shape, "", "tf_stage_batch64.md"), "kernel": "tf_stage",
                     "kernel_ms": stage_ms, "stage_bytes_per_cell": bpc},
        "sequential": {"value": seq_value, "steps": seq_steps, "by_shape": seq,
                       "batched_over_sequential": value / seq_value},
        "persistent": {"value": cells * args.steps / pwall, "ms_per_step": pwall / args.steps * 1e3,
                       "over_graph": (cells * args.steps / pwall) / value},
        "autotune_ms_per_step": tune_ms,
        "kernels_ms": {k: v[0] for k, v in kt.items()},
    }
    if not args.no_cpu:
        threads = cpu_threads()
        k = next(i for i, s in enumerate(specs) if s.xdim == 256)
        st, cfg = members[k]
        n = args.cpu_steps or 10
        times = reference_step_times(st, cfg, n, threads)
        secs = sum(t for t, _ in times)
        cpu = host_cpu()
        out["cpu_baseline"] = {
            "value": st.xdim * st.ydim * n / secs, "unit": "cell-updates/s", "cores": threads, "kind": "reference",
            "sample": "member %d (256x256) only, steps 1..%d of its own run, reference executable (g++ -O3 -fopenmp), "
                      "timed between its per-step output lines (%.2f s); %d OpenMP threads on %s"
                      % (k, n, secs, threads, cpu["model"]),
            "host": cpu}
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    return out


def bench_adjoint(args) -> dict:
    """The periodic-wave adjoint: 256x256 ideal MHD, RK2, 500 steps, checkpoint every 50; gradient of
    sum((rho_final - rho_target)^2) with respect to the initial temperature grid and v0.  Reports the
    cost ratio (checkpointed forward + reverse sweep) / forward and a finite-difference probe."""
    import numpy as np
    from spruce_amd.adjoint import PeriodicWaveProblem

    steps = args.steps or 500
    prob = PeriodicWaveProblem(256, 256, nsteps=steps, checkpoint=50)
    prob.forward(prob.temp0, prob.v0)                       # warm-up (module load, first allocations)
    loss, rho, t_fwd = prob.forward(prob.temp0, prob.v0)
    g_temp, g_v0, t_adj = prob.gradient(prob.temp0, prob.v0, rho)
    rng = np.random.default_rng(1)
    dt = np.zeros((256, 256))
    dt.reshape(-1)[rng.integers(0, 65536, 256)] = rng.standard_normal(256) / 16.0
    h = 1e-6
    ad = float(np.sum(g_temp * dt * prob.temp0))
    lp, _, _ = prob.forward(prob.temp0 * (1 + h * dt), prob.v0)
    lm, _, _ = prob.forward(prob.temp0 * (1 - h * dt), prob.v0)
    fd = (lp - lm) / (2 * h)
    cells = 256 * 256
    return {
        "metric": "cell-updates/s and % HBM roofline (FP64); adjoint/forward cost ratio",
        "value": t_adj / t_fwd, "unit": "adjoint/forward time ratio", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": t_adj / steps * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "adjoint of the periodic ideal-MHD wave 256x256 (RK2, FP64), %d steps, "
                               "checkpoint every 50; gradient w.r.t. the initial temperature grid and v0" % steps},
        "forward_s": t_fwd, "adjoint_s": t_adj, "forward_cell_updates_per_s": cells * steps / t_fwd,
        "loss": loss, "dJ_dv0": g_v0,
        "fd_probe": {"adjoint": ad, "central_difference": fd, "relative_difference": abs(ad - fd) / (abs(ad) + abs(fd))},
    }


def bench_adjoint_arcade(args) -> dict:
    """The gradient workload with modules: the 2048x2048 arcade (conduction, radiative losses and
    the 16-coefficient parametrised heating), RK2, checkpoint every 50 steps; gradient of the
    coronal temperature misfit with respect to the 16 heating coefficients.  Reports the cost ratio
    (checkpointed forward + per-step replays + reverse sweep) / forward, the device memory the
    adjoint allocated, and central differences on 4 coefficients."""
    import numpy as np
    from spruce_amd.adjoint import ArcadeHeatingProblem

    steps = args.steps or 1000
    n = args.grid or 2048
    opts = dict(kv.partition("=")[::2] for kv in args.option)
    prob = ArcadeHeatingProblem(n, n, nsteps=steps, checkpoint=50, options=opts)
    th = prob.theta0
    print("adjoint_arcade: forward warm-up", file=sys.stderr, flush=True)
    prob.forward(th)
    loss, final, t_fwd = prob.forward(th)
    print("adjoint_arcade: forward %.2f s, loss %.6e; adjoint" % (t_fwd, loss), file=sys.stderr, flush=True)
    g, t_adj, mem = prob.gradient(th, final)
    print("adjoint_arcade: adjoint %.2f s" % t_adj, file=sys.stderr, flush=True)
    # Central differences need a loss that is smooth in theta over the difference interval.  Over
    # tens of steps this one is not: the conduction subcycle count is an integer set by the
    # smallest time scale over the whole grid, which the cells cooling at the open top move, so the
    # loss jumps wherever that count changes.  The probes use the same grid over 5 steps.
    short = ArcadeHeatingProblem(n, n, nsteps=5, checkpoint=50, options=opts)
    _, fin5, _ = short.forward(th)
    g5, _, _ = short.gradient(th, fin5)
    fd = []
    for k in (0, 5, 10, 15):
        ds = []
        for rel in (1e-5, 1e-6):
            e = rel * th[k]
            tp, tm = th.copy(), th.copy()
            tp[k] += e
            tm[k] -= e
            ds.append((short.forward(tp)[0] - short.forward(tm)[0]) / (tp[k] - tm[k]))
        fd.append({"k": k, "adjoint": float(g5[k]), "central_difference": ds[0], "central_difference_1e-6": ds[1],
                   "relative_difference": abs(g5[k] - ds[0]) / (abs(g5[k]) + abs(ds[0]))})
        print("adjoint_arcade: 5 steps, theta_%d adjoint %.9e fd %.9e %.9e" % (k, g5[k], ds[0], ds[1]),
              file=sys.stderr, flush=True)
    cells = n * n
    return {
        "metric": "cell-updates/s and % HBM roofline (FP64); adjoint/forward cost ratio",
        "value": t_adj / t_fwd, "unit": "adjoint/forward time ratio", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": t_adj / steps * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "adjoint of the solar arcade %dx%d (ideal MHD + thermal conduction + radiative "
                               "losses + 16-coefficient parametric heating, RK2, FP64), %d steps, checkpoint "
                               "every 50; gradient of the coronal temperature misfit w.r.t. the heating "
                               "coefficients" % (n, n, steps)},
        "forward_s": t_fwd, "adjoint_s": t_adj, "forward_cell_updates_per_s": cells * steps / t_fwd,
        "adjoint_device_bytes": mem, "loss": loss, "dJ_dtheta": [float(x) for x in g],
        "fd_probes": {"steps": 5, "relative_step": 1e-5, "probes": fd},
    }


def bench_copy(args) -> dict:
    """HBM calibration: one elementwise kernel (torch.neg) reading and writing 1 GiB of FP64.  Its
    FETCH_SIZE / WRITE_SIZE counters per dispatch, against the known 1 GiB each way, give the factor
    that turns those counters into bytes for the stage kernels; its rate is the achievable copy
    bandwidth of this GPU."""
    import torch

    n = 1 << 27
    a = torch.rand(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        torch.neg(a, out=b)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    gbs = 2.0 * n * 8 / (ms * 1e-3) / 1e9
    return {"metric": "HBM copy bandwidth (calibration)", "value": gbs, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "elementwise FP64 negate, 1 GiB read + 1 GiB write per launch"},
            "bytes_read": n * 8, "bytes_written": n * 8, "frac_of_peak": gbs / HBM_PEAK_GBS}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: 20; adjoint 500, "
                    "adjoint_arcade 1000, batch2f 200)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="arcade2048")
    ap.add_argument("--cpu-steps", type=int, default=0, help="reference steps for the CPU baseline (0: auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stage-kernel", default="auto", choices=["auto", "strip", "cell", "tile"])
    ap.add_argument("--strip-rows", type=int, default=0)
    ap.add_argument("--grid", type=int, default=0, help="edge of the square arcade of the adjoint_arcade workload")
    ap.add_argument("--tc-kernel", default="auto", choices=["auto", "strip", "cell"])
    ap.add_argument("--tc-rows", type=int, default=0)
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="execution option of the GPU domain (spruce_gpu_set_option), repeatable")
    ap.add_argument("--cpu-malloc", default="heap", choices=["heap", "default"],
                    help="glibc malloc tunables of the reference CPU leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from spruce_amd.gpu import GpuDomain

    if args.steps is None:
        args.steps = 200 if args.workload == "batch2f" else (20 if args.workload not in ("adjoint", "adjoint_arcade")
                                                             else 0)
    if args.workload == "batch2f":
        print(json.dumps(bench_batch(args)))
        return
    if args.workload == "adjoint":
        print(json.dumps(bench_adjoint(args)))
        return
    if args.workload == "adjoint_arcade":
        print(json.dumps(bench_adjoint_arcade(args)))
        return
    if args.workload == "copy":
        print(json.dumps(bench_copy(args)))
        return
    st, cfg, spec, desc = make_workload(args.workload)
    cells = spec.xdim * spec.ydim
    kernel = resolved_stage_kernel(args.stage_kernel, spec.xdim, spec.ydim)
    opts = {"stage_kernel": args.stage_kernel, "tc_kernel": args.tc_kernel}
    if args.strip_rows:
        opts["strip_rows"] = args.strip_rows
    if args.tc_rows:
        opts["tc_rows"] = args.tc_rows
    for kv in args.option:
        k, _, v = kv.partition("=")
        opts[k] = v
    dom = GpuDomain(spec, options=opts)
    dom.step(args.warmup)
    snapshot = None
    if not args.no_cpu:
        snapshot = state_of(dom, st)          # the CPU leg starts from the state of step W
    t0 = time.perf_counter()
    _, subs = dom.step(args.steps, subcycles=True)   # returns after a device synchronisation
    wall = time.perf_counter() - t0
    value = cells * args.steps / wall
    mods = desc["modules"]
    mean_sub = {m: float(np.mean(subs[:, k])) for k, m in enumerate(mods)}

    # Per-kernel durations from HIP events recorded on the domain's own stream, over as many steps
    # as were timed (at most 50), starting where the timed steps ended.
    dom.set_profiling(True)
    prof_steps = max(5, min(args.steps, 50))
    _, psubs = dom.step(prof_steps, subcycles=True)
    kt = dom.kernel_times()
    dom.set_profiling(False)
    integ = desc["integrator"]
    per_step = {k: v[0] * v[1] / prof_steps for k, v in kt.items()}
    dominant = max(per_step, key=per_step.get)
    kms = kt[dominant][0]
    bpc = kernel_bytes_per_cell(dominant, integ)
    achieved = bpc * cells / (kms * 1e-3) / 1e9 if bpc else None
    rows = str(args.tc_rows or "auto") if dominant == "tc_step" else str(args.strip_rows or "auto")
    traffic = measured_traffic(args.workload, dominant, rows,
                               "hbm_traffic_tc_step.md" if dominant == "tc_step" else "hbm_traffic_mhd_stage.md")
    # algorithmic bytes of one whole step: the stages + 40 B per conduction subcycle + 24 B per
    # radiative subcycle + 64 B of subcycle-count passes when those modules are present
    b_alg = 8.0 * (5 * E_MHD + 2 * S_MHD) if integ == "rk2" else stage_bytes_per_cell(integ)
    if mods:
        b_alg += 40.0 * mean_sub.get("thermal_conduction", 0.0) + 24.0 * mean_sub.get("radiative_losses", 0.0) + 64.0
    ms_per_step = wall / args.steps * 1e3

    out = {
        "metric": "cell-updates/s and % HBM roofline (FP64); adjoint/forward cost ratio",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": dict(desc, stage_kernel=kernel, strip_rows=args.strip_rows or "auto", tc_kernel=args.tc_kernel,
                       tc_rows=args.tc_rows or "auto", options=args.option,
                       timed_steps=[args.warmup, args.warmup + args.steps - 1]),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "kernel": dominant, "kernel_ms": kms, "bytes_per_cell": bpc,
                     "kernel_share_of_step": per_step[dominant] * prof_steps / sum(per_step.values()) / prof_steps
                     if per_step else None,
                     "step_bytes_per_cell_update": b_alg,
                     "step_frac": b_alg * value / 1e9 / HBM_PEAK_GBS},
        "mean_subcycles": mean_sub,
        "kernels_ms": {k: v[0] for k, v in kt.items()},
        "kernels_ms_per_step": per_step,
        "profiled_steps": [args.warmup + args.steps, args.warmup + args.steps + prof_steps - 1],
        "profiled_mean_subcycles": {m: float(np.mean(psubs[:, k])) for k, m in enumerate(mods)},
    }
    if not args.no_cpu:
        threads = cpu_threads()
        n = args.cpu_steps or (1 if cells >= 1 << 20 else 5)
        env = MALLOC_HEAP_ENV if args.cpu_malloc == "heap" else {}
        times = reference_step_times(snapshot, cfg, n, threads, env)
        secs = sum(t for t, _ in times)
        cpu = host_cpu()
        first = args.warmup + 1
        counts = [module_count(line, "Thermal Subcycles") for _, line in times]
        gpu_counts = [int(c) for c in subs[1:1 + n, 0]] if "thermal_conduction" in mods and args.steps > n else []
        out["cpu_baseline"] = {
            "value": cells * n / secs, "unit": "cell-updates/s", "cores": threads, "kind": "reference",
            "sample": "reference executable (g++ -O3 -fopenmp, its own sources), steps %d..%d of this workload, "
                      "started from the GPU's state after %d steps and timed between its own per-step output lines "
                      "(%.2f s, set-up and file output excluded); %d OpenMP threads (OMP_PLACES=cores) on %s; "
                      "malloc: %s" % (first, first + n - 1, args.warmup, secs, threads, cpu["model"],
                                      "freed grids kept in the heap" if env else "glibc default"),
            "step_seconds": [t for t, _ in times], "steps": [first, first + n - 1],
            "conduction_subcycles": counts, "gpu_conduction_subcycles_same_steps": gpu_counts,
            "host": cpu}
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
