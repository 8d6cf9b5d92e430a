"""Benchmark of the l() / rk3_ssp() hot path on one MI355X.

Workload: the smooth-mode periodic box [0,1)^2 at 2048 x 2048, 6 FP64
fields, seeded sum of 8 Fourier modes per primitive field, SSP-RK3 with
dt = 0.05*dt_cfl computed on the device every step (the reference's time
loop body, src/main.rs:459-467).  A "step" is one RK3 step (3 l() + 3
stage updates); throughput counts fluid cells x 3 stages x steps / wall.

Prints ONE JSON line.  The timed region runs the device time loop as
production runs it (rh3t_advance replaying the captured hipGraph of one
step); a second run of the same steps, launched eagerly with hipEvents
recorded around every kernel on the solver's own stream, gives the
per-kernel averages and the roofline entry for the dominant kernel (the
fused y-direction WENO flux + assembly pass; the x-direction flux kernel
gets its own entry beside it).  The CPU baseline is the oracle (C restatement of
the reference algorithm) timed on a bounded sample of the same workload:
whole RK3 steps of the same 2048^2 box on the host's cores, about 10 s.
"""
import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 cell-stage updates/s, %HBM/FP64 peak; batched vs plain; adjoint/fwd time"
PEAK_FP64 = 78.6e12    # MI355X dense FP64 (vector and matrix), flop/s
PEAK_HBM = 8.0e12      # MI355X HBM3E, B/s
# Algorithmic FP64 operation counts (+ - * / sqrt pow each 1, powi(n) as n-1
# multiplies, constant subexpressions 0), regenerated by oracle/opcount.cpp
# from the oracle: one characteristic WENO5 interface flux of
# src/weno.rs:685-802 and one interior cell-stage (x and y fluxes, noncon
# x/y, derived, assembly, mean stage combination, admissibility).
FLOP_PER_INTERFACE = 2252
FLOP_PER_CELL_STAGE = 5250
# the assembly part of a cell-stage: noncon x and y (2 x 308), assembly 56,
# stage combination 24, admissibility 12
FLOP_PER_CELL_ASSEMBLY = 2 * 308 + 56 + 24 + 12
# Derived::from_state of the new state, formed by the same pass (38)
FLOP_PER_CELL_DERIVED = 38
BYTES_PER_CELL_STAGE = 128  # stage 1: read U^k, write U (96 B); stages 2-3: + U^n (144 B)
# HBM traffic of one k_flux<0> launch on the 2048^2 box, from two rocprofv3
# --pmc passes over this bench (FETCH_SIZE and WRITE_SIZE, KiB per dispatch,
# mean of 15 dispatches, four significant digits), each scaled by what the
# same counter reads for a plain 1 GiB device-to-device copy
# (`--copy-calibration`: FETCH_SIZE reads 524,304 KiB -> x2.0, WRITE_SIZE
# 1,048,576 KiB -> x1.0).
FLUX_X_FETCH_KIB, FLUX_X_WRITE_KIB = 236946, 196704
COPY_FETCH_SCALE, COPY_WRITE_SCALE = 2.0, 1.0
FLUX_X_TRAFFIC_BYTES = (FLUX_X_FETCH_KIB * COPY_FETCH_SCALE + FLUX_X_WRITE_KIB * COPY_WRITE_SCALE) * 1024
TRAFFIC_SOURCE = ("not measured in this run: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of k_flux<0> at 2048^2, "
                  "copy-calibrated (profiles/flux_x_counters_2048.md)")
# the same for one k_fy_assemble<true> launch (profiles/hbm_per_stage_2048.md);
# it writes the new state's 8 Derived planes beside the state
FY_AS_FETCH_KIB, FY_AS_WRITE_KIB = 400295, 459561
FY_AS_TRAFFIC_BYTES = (FY_AS_FETCH_KIB * COPY_FETCH_SCALE + FY_AS_WRITE_KIB * COPY_WRITE_SCALE) * 1024
FY_AS_TRAFFIC_SOURCE = ("not measured in this run: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of k_fy_assemble<true> "
                        "at 2048^2, copy-calibrated (profiles/hbm_per_stage_2048.md)")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _host_facts():
    """nproc, the affinity mask, OMP_NUM_THREADS and the CPU model."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    return {"nproc": os.cpu_count(), "affinity_cores": affinity, "omp_num_threads": omp, "cpu_model": model}


def cpu_baseline(field, u, workloads, min_seconds=10.0):
    """The oracle (C restatement of the reference algorithm, OpenMP at the
    reference's Rayon sites) on whole RK3 steps of the same box: on the
    host's CPU share -- every core of the affinity mask, capped by
    OMP_NUM_THREADS where the host sets it (the share of one GPU on a shared
    node) -- for about min_seconds, then one step on one core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bridge as ob
    facts = _host_facts()
    cores = facts["affinity_cores"]
    if facts["omp_num_threads"] and facts["omp_num_threads"].isdigit():
        cores = min(cores, int(facts["omp_num_threads"]))
    cores = max(1, cores)
    prob = ob.OracleProblem(field)
    n_fluid = int(field.fluid.sum())

    def run(nthreads, min_s, max_steps):
        uu, steps, wall = u, 0, 0.0
        while steps == 0 or (wall < min_s and steps < max_steps):
            dt = 0.05 * ob.global_dt(prob, uu)
            t0 = time.perf_counter()
            uu, _ = ob.rk3_step(prob, uu, dt, nthreads=nthreads)
            wall += time.perf_counter() - t0
            steps += 1
        return steps, wall

    steps, wall = run(cores, min_seconds, 64)
    steps1, wall1 = run(1, 0.0, 1)
    box = f"{field.grid.nx}x{field.grid.ny}"
    return {"value": n_fluid * 3 * steps / wall, "unit": "cell-stage/s", "cores": cores, "kind": "port",
            "sample": f"{steps} SSP-RK3 steps (3 l() + stages each) of the same {box} box, C oracle with OpenMP "
                      f"on {cores} threads, {wall:.2f} s",
            "ms_per_step": wall / steps * 1e3,
            "one_core": {"value": n_fluid * 3 * steps1 / wall1, "ms_per_step": wall1 / steps1 * 1e3,
                         "sample": f"{steps1} SSP-RK3 step of the {box} box on 1 thread, {wall1:.2f} s"},
            "host": facts}


def batched_leg(rh3t, n_solves=64, steps=100, seed=4, tune=True, arith=0):
    """Config of the batched form: n_solves independent half-cylinder solves
    with grids drawn from {121x481, 161x641, 201x801, 241x961} and inflow Mach
    ~ U[2, 4] (seeded LCG), `steps` SSP-RK3 steps each with per-solve dt =
    0.05 dt_cfl; one batched launch sequence per step (a hipGraph of the
    step's ~20 batched kernels, replayed from the host once per step)
    against the same solves run back to back with the plain solver (one
    step graph per solve), both in the same arithmetic."""
    workloads = importlib.import_module("3t_rh_amd.workloads")
    autotune = importlib.import_module("3t_rh_amd.autotune")
    draws = workloads.batched_draw(n_solves, seed)
    log(f"batched (arith={arith}): building {n_solves} cylinders")
    fields, states = workloads.batched_cylinders(draws)
    cells = sum(int(f.fluid.sum()) for f in fields)
    batch = rh3t.Batch(fields, arith=arith)

    def reset():
        for k, u in enumerate(states):
            batch.set_state(k, u, 0.0)

    tuned = None
    if tune:
        reset()
        key = autotune.signature([(f.grid.nx, f.grid.ny) for f in fields], arith)
        tuned = autotune.tune(batch, steps=3, log=log)
        tuned["key"] = key
    reset()
    batch.advance(1)
    reset()
    t0 = time.perf_counter()
    batch.advance(steps)
    t_batch = time.perf_counter() - t0
    # per-kernel breakdown of a few profiled (eager) batched steps
    reset()
    batch.profile(True)
    batch.advance(3)
    kern = {k: round(v[0] * v[1] / 3, 4) for k, v in batch.profile_read().items() if v[1]}
    batch.profile(False)
    batch.close()
    # the same solves back to back with the plain solver (graph per solve)
    solvers = [f.solver(arith=arith) for f in fields]
    for s, u in zip(solvers, states):
        s.set_state(u, 0.0)
        s.advance(1)
        s.set_state(u, 0.0)
    t0 = time.perf_counter()
    for s in solvers:
        s.advance(steps)
    t_plain = time.perf_counter() - t0
    for s in solvers:
        s.close()
    return {"arith": "exact" if arith == 0 else "fast", "solves": n_solves, "steps": steps, "fluid_cells": cells,
            "grids": {str(nx): sum(1 for d in draws if d[0] == nx) for nx in workloads.CYLINDER_CLASSES},
            "mach_range": [min(d[1] for d in draws), max(d[1] for d in draws)],
            "launch": "per-step hipGraph of the batched kernels, replayed from the host",
            "batched_s": t_batch, "plain_back_to_back_s": t_plain, "batched_vs_plain": t_plain / t_batch,
            "batched_cell_stage_per_s": cells * 3 * steps / t_batch,
            "plain_cell_stage_per_s": cells * 3 * steps / t_plain,
            "autotune": None if tuned is None else {k: tuned[k] for k in ("bx", "by", "x_bx", "x_by",
                                                                          "ghosts_per_block",
                                                                          "ms_per_step")},
            "kernels_ms_per_step": kern}


def double_mach_leg(rh3t, h=1.0 / 320.0, steps=50):
    """The reference's second benchmark, init_double_mach (src/init.rs:26-113)
    at its own h = 1/320 (1270 x 892): the Mach-10 ramp with Wall (ILW)
    ghosts, `steps` SSP-RK3 steps with dt = 0.05 dt_cfl, hipGraph replay."""
    f = rh3t.Field.init_double_mach(h)
    n_fluid = int(f.fluid.sum())
    s = f.solver()
    s.set_state(f.double_mach_state(), 0.0)
    s.advance(2, 0.05)
    s.profile(True)
    s.advance(3, 0.05)
    kern = {k: round(v[0] * v[1] / 3, 4) for k, v in s.profile_read().items() if v[1]}
    s.profile(False)
    s.advance(1, 0.05)  # graph captured again outside the timed region
    t0 = time.perf_counter()
    s.advance(steps, 0.05)
    wall = time.perf_counter() - t0
    counts = f.counts()
    s.close()
    return {"workload": f"init_double_mach h=1/{round(1 / h)}", "grid": [f.grid.nx, f.grid.ny],
            "fluid_cells": n_fluid, "ghosts": counts[1], "wall_ghosts": counts[2], "steps": steps,
            "ms_per_step": wall / steps * 1e3, "cell_stage_per_s": n_fluid * 3 * steps / wall,
            "kernels_ms_per_step": kern}


def plain_leg(rh3t, field, u0, arith, steps, warmup):
    """The 2048^2 box in one arithmetic: warmup steps (which capture the
    step's hipGraph), the timed region of `steps` steps of the production
    loop (rh3t_advance replaying the graph; it returns after a stream sync),
    then the same steps launched eagerly with hipEvents around every kernel
    on the solver's stream (the per-kernel breakdown and the roofline)."""
    log(f"plain 2048 box, arith={arith}")
    solver = field.solver(arith=rh3t.ARITH_FAST if arith == "fast" else rh3t.ARITH_EXACT)
    solver.set_state(u0, 0.0)
    solver.advance(warmup, 0.05)
    t0 = time.perf_counter()
    solver.advance(steps, 0.05)
    wall = time.perf_counter() - t0
    log(f"timed {steps} steps in {wall:.3f} s")
    solver.profile(True)
    t0 = time.perf_counter()
    solver.advance(steps, 0.05)
    wall_profiled = time.perf_counter() - t0
    prof = solver.profile_read()
    solver.profile(False)
    u_end, _ = solver.get_state()
    solver.close()
    n_fluid = int(field.fluid.sum())
    return {"rate": n_fluid * 3 * steps / wall, "wall": wall, "wall_profiled": wall_profiled, "prof": prof,
            "finite": bool(abs(u_end[field.fluid]).max() < 1e300)}


def cylinder_leg(rh3t, nx=121, mach=3.0, steps=200, arith="exact"):
    """init_cylinder (src/init.rs:116-510) at 121x481, Mach 3: `steps` SSP-RK3
    steps of dt = 0.05 dt_cfl on the device (graph replay), from the
    uniform freestream; the per-kernel breakdown of 3 eagerly launched
    steps with hipEvents."""
    f = rh3t.Field.init_cylinder(nx, mach)
    n_fluid = int(f.fluid.sum())
    s = f.solver(arith=rh3t.ARITH_FAST if arith == "fast" else rh3t.ARITH_EXACT)
    u = f.cylinder_state()
    s.set_state(u, 0.0)
    s.advance(2, 0.05)
    s.profile(True)
    s.advance(3, 0.05)
    kern = {k: round(v[0] * v[1] / 3, 4) for k, v in s.profile_read().items() if v[1]}
    s.profile(False)
    s.set_state(u, 0.0)
    s.advance(1, 0.05)  # graph captured again outside the timed region
    s.set_state(u, 0.0)
    t0 = time.perf_counter()
    s.advance(steps, 0.05)
    wall = time.perf_counter() - t0
    counts = f.counts()
    s.close()
    return {"workload": f"init_cylinder {nx}x{f.grid.ny} Mach {mach}", "arith": arith, "fluid_cells": n_fluid,
            "ghosts": counts[1], "heavy_ghosts": counts[2], "steps": steps, "ms_per_step": wall / steps * 1e3,
            "cell_stage_per_s": n_fluid * 3 * steps / wall, "kernels_ms_per_step": kern}


def _gpu_ms(profile):
    """Summed kernel time (ms) of a solver's per-kernel profile."""
    return sum(avg * cnt for avg, cnt in profile.values())


def adjoint_leg(rh3t, n=512, nsteps=20, every=5, fd_dirs=32, reps=5):
    """The differentiation workload: the smooth-mode periodic box at n^2
    with the exchange terms on, nsteps RK3 steps with the dt schedule of a
    base run frozen, J = sum(w * E_r(u^N)) with a seeded Gaussian window.
    Times the checkpointing forward (state kept every `every` steps) and the
    backward (segment recompute + reverse sweep), by wall clock (medians of
    `reps` pairs; the headline ratio) and by summed per-kernel GPU time, and
    checks the gradient against central
    differences (h = 1e-6) of the forward on fd_dirs seeded smooth
    directions over (u0, omega_ei, omega_er)."""
    import numpy as np

    workloads = importlib.import_module("3t_rh_amd.workloads")
    f = workloads.periodic_box(n)
    u = workloads.smooth_modes_state(f, seed=3)
    phys = rh3t.Physics(**workloads.ADJOINT_PHYSICS)
    s = f.solver(physics=phys)
    s.set_state(u)
    dts = []
    for _ in range(nsteps):
        dts.append(0.05 * s.calc_global_dt())
        s.rk3_ssp(dts[-1])
    lam = workloads.window_loss_cotangent(f, seed=5)
    adj = s.adjoint(every)
    s.set_state(u)
    adj.forward(dts)
    adj.backward(lam)  # warm-up (allocations, code objects)
    t_f, t_b, d_f, d_b = [], [], [], []
    for _ in range(reps):  # medians of reps forward + backward pairs
        s.set_state(u)
        t0 = time.perf_counter()
        adj.forward(dts)
        t_f.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        gu, gp = adj.backward(lam)
        t_b.append(time.perf_counter() - t0)
        df, db = adj.device_ms()
        d_f.append(df)
        d_b.append(db)
    t_fwd, t_bwd = float(np.median(t_f)), float(np.median(t_b))
    dev_fwd, dev_bwd = float(np.median(d_f)), float(np.median(d_b))
    # GPU-side time of the same two calls
    s.set_state(u)
    s.profile(True)
    adj.forward(dts)
    prof_f = s.profile_read()
    s.profile(True)
    adj.backward(lam)
    prof_b = s.profile_read()
    s.profile(False)
    g_fwd, g_bwd = _gpu_ms(prof_f), _gpu_ms(prof_b)
    w = lam[..., 5]

    def J(x, pv):
        s.set_physics(phys.with_params(pv))
        s.set_state(x)
        for d in dts:
            s.rk3_ssp(d)
        y, _ = s.get_state()
        return float(np.sum(w * y[..., 5]))

    h, worst = 1e-6, 0.0
    p0 = phys.params()
    for seed in range(5001, 5001 + fd_dirs):
        rng = np.random.default_rng(seed)
        du = workloads.smooth_direction(f, u, seed)
        dp = np.zeros(5)
        dp[:2] = rng.standard_normal(2) * np.array([1.0, 0.1])
        fd = (J(u + h * du, p0 + h * dp) - J(u - h * du, p0 - h * dp)) / (2 * h)
        ad = float(np.sum(gu * du) + np.sum(gp * dp))
        worst = max(worst, abs(fd - ad) / abs(ad))
    adj.close()
    s.close()
    kern = {k: round(v[0] * v[1], 4) for k, v in prof_b.items() if v[1]}
    return {"workload": f"smooth-mode periodic box {n}x{n}, omega_ei=1, omega_er=0.1, {nsteps} SSP-RK3 steps, "
                        f"J = sum(gaussian window * E_r)",
            "checkpoint_every": every, "timing_reps": reps,
            # device-resident: HIP events from after the host inputs are on
            # the device to before the results leave it (the backward's 12.6
            # MB cotangent upload and gradient download excluded, as the
            # forward's state is already on the device)
            "forward_ms": dev_fwd, "backward_ms": dev_bwd, "adjoint_over_forward": dev_bwd / dev_fwd,
            # wall clock around the two calls, host transfers included
            "forward_wall_ms": t_fwd * 1e3, "backward_wall_ms": t_bwd * 1e3,
            "adjoint_over_forward_wall": t_bwd / t_fwd,
            # summed per-kernel times of one profiled pair: the backward's
            # recompute runs on a second stream concurrently with the
            # reverse sweep, so its kernels overlap and this sum exceeds the
            # elapsed time
            "forward_kernel_sum_ms": g_fwd, "backward_kernel_sum_ms": g_bwd, "backward_kernels_ms": kern,
            "grad_params": [float(gp[0]), float(gp[1])], "fd_directions": fd_dirs, "fd_h": h,
            "fd_max_rel_err": worst}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--arith", choices=["fast", "exact"], default="fast")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain leg")
    ap.add_argument("--batched-solves", type=int, default=64)
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiation leg")
    ap.add_argument("--no-double-mach", action="store_true", help="skip the init_double_mach leg")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the differentiation leg (for profiling it); prints its JSON object")
    ap.add_argument("--fd-dirs", type=int, default=32, help="central-difference directions of the adjoint leg")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched-vs-plain leg (for profiling it); prints its JSON object")
    ap.add_argument("--double-mach-only", action="store_true",
                    help="run only the init_double_mach leg (for profiling it); prints its JSON object")
    ap.add_argument("--cylinder-only", action="store_true",
                    help="run only the init_cylinder 121x481 leg (for profiling it); prints its JSON object")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (three launches), the reference for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        n = 1 << 27  # doubles: 1 GiB
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n * 8, "launches": 3}), flush=True)
        return
    rh3t = importlib.import_module("3t_rh_amd")
    workloads = importlib.import_module("3t_rh_amd.workloads")
    if args.batched_only:
        print(json.dumps({"batched": batched_leg(rh3t, args.batched_solves, args.batched_steps,
                                                 arith=0 if args.arith == "exact" else 1)}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_leg(rh3t, fd_dirs=args.fd_dirs)}), flush=True)
        return
    if args.double_mach_only:
        print(json.dumps({"double_mach": double_mach_leg(rh3t)}), flush=True)
        return
    if args.cylinder_only:
        print(json.dumps({"cylinder": cylinder_leg(rh3t, steps=args.steps, arith=args.arith)}), flush=True)
        return
    n = args.size
    log(f"building {n}x{n} periodic box")
    field = workloads.periodic_box(n)
    u0 = workloads.smooth_modes_state(field, seed=3)
    n_fluid = int(field.fluid.sum())
    head = args.arith
    other = "exact" if head == "fast" else "fast"
    legs = {a: plain_leg(rh3t, field, u0, a, args.steps, args.warmup) for a in (head, other)}
    h = legs[head]
    rate = h["rate"]
    log(f"plain: {head} {rate:.4e}, {other} {legs[other]['rate']:.4e} cell-stage/s")
    prof = h["prof"]
    fx_ms, fx_n = prof["flux_x"]
    interfaces = (n + 1) * n
    achieved = FLOP_PER_INTERFACE * interfaces / (fx_ms * 1e-3)
    # the y fluxes with the assembly (one pass) when the launch shape fuses
    # them, else the y flux kernel
    if "flux_y_assemble" in prof and prof["flux_y_assemble"][1]:
        dom_name, (dom_ms, dom_n) = "flux_y_assemble", prof["flux_y_assemble"]
        dom_flop = FLOP_PER_INTERFACE * interfaces + (FLOP_PER_CELL_ASSEMBLY + FLOP_PER_CELL_DERIVED) * n_fluid
        dom_traffic, dom_src = (FY_AS_TRAFFIC_BYTES, FY_AS_TRAFFIC_SOURCE) if (n == 2048 and head == "fast") \
            else (None, None)
    else:
        dom_name, (dom_ms, dom_n) = "flux_y", prof["flux_y"]
        dom_flop = FLOP_PER_INTERFACE * interfaces
        dom_traffic, dom_src = None, None
    roof_fx = {"bound": "valu", "pipe": "fp64 valu", "kernel": "flux_x", "achieved": achieved / 1e12,
               "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64,
               "traffic": FLUX_X_TRAFFIC_BYTES if n == 2048 else None,
               "traffic_source": TRAFFIC_SOURCE if n == 2048 else None,
               "flop_per_interface": FLOP_PER_INTERFACE, "interfaces": interfaces,
               "avg_launch_ms": fx_ms, "launches_timed": fx_n}
    dom_ach = dom_flop / (dom_ms * 1e-3)
    roof_dom = {"bound": "valu", "pipe": "fp64 valu", "kernel": dom_name, "achieved": dom_ach / 1e12,
                "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": dom_ach / PEAK_FP64,
                "traffic": dom_traffic, "traffic_source": dom_src,
                "flop_per_launch": dom_flop, "flop_per_interface": FLOP_PER_INTERFACE,
                "flop_per_cell_assembly": FLOP_PER_CELL_ASSEMBLY + FLOP_PER_CELL_DERIVED,
                "interfaces": interfaces, "cells": n_fluid,
                "avg_launch_ms": dom_ms, "launches_timed": dom_n}
    if fx_ms > dom_ms:
        roof_dom, roof_fx = roof_fx, roof_dom
    line = {
        "metric": METRIC,
        "value": rate,
        "unit": "cell-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"smooth-mode periodic box {n}x{n}, 6 FP64 fields, SSP-RK3, dt=0.05*dt_cfl",
                   "grid": [n, n], "fluid_cells": n_fluid, "arith": head},
        # agreement with the oracle measured by the GPU tests at this
        # workload (tests/test_gpu_parity.py, 2048^2 l() and RK3 step) and
        # where FAST misses the 1e-12 bar
        "arith_parity": {"bar": 1e-12, "fast_2048_l": 6.4e-13, "exact_2048": "<= 1e-13",
                         "fast_misses": "cylinder (121x481) first full RK3 step: 1.4e-12; every other config "
                                        "<= 1e-12",
                         "exact_misses": "none"},
        "fp64_frac_of_peak": FLOP_PER_CELL_STAGE * rate / PEAK_FP64,
        "hbm_frac_of_peak": BYTES_PER_CELL_STAGE * rate / PEAK_HBM,
        "ms_per_step_profiled": h["wall_profiled"] / args.steps * 1e3,
        "kernels_ms": {k: round(v[0], 4) for k, v in prof.items() if v[1]},
        "state_finite": h["finite"],
        other: {"value": legs[other]["rate"], "ms_per_step": legs[other]["wall"] / args.steps * 1e3,
                "kernels_ms": {k: round(v[0], 4) for k, v in legs[other]["prof"].items() if v[1]},
                "state_finite": legs[other]["finite"]},
        # the dominant kernel (the longest launch) and the other flux pass,
        # both bound by FP64 vector (VALU) issue, priced against the dense
        # FP64 peak (78.6 TFLOP/s: MI355X's vector and matrix FP64 peaks
        # are equal)
        "roofline": roof_dom,
        "roofline_other": roof_fx,
    }
    if not args.no_batched:
        log("batched leg")
        line["batched"] = batched_leg(rh3t, args.batched_solves, args.batched_steps, arith=0)
        log(f"batched exact: {json.dumps(line['batched'])}")
        line["batched_fast"] = batched_leg(rh3t, args.batched_solves, args.batched_steps, arith=1)
        log(f"batched fast: {json.dumps(line['batched_fast'])}")
    log("cylinder leg")
    line["cylinder"] = cylinder_leg(rh3t)
    log(f"cylinder: {json.dumps(line['cylinder'])}")
    if not args.no_double_mach:
        log("double Mach leg")
        line["double_mach"] = double_mach_leg(rh3t)
        log(f"double Mach: {json.dumps(line['double_mach'])}")
    if not args.no_adjoint:
        log("adjoint leg")
        line["adjoint"] = adjoint_leg(rh3t, fd_dirs=args.fd_dirs)
        log(f"adjoint: {json.dumps(line['adjoint'])}")
    if not args.no_cpu:
        log("cpu baseline (oracle, ~10 s of RK3 steps)")
        line["cpu_baseline"] = cpu_baseline(field, u0, workloads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
