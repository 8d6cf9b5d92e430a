"""Launch-shape autotuner for the batched and plain solvers.

Two knobs are tuned on the actual problem set:

* the workgroup shape (bx, by) of the WENO flux and assembly kernels --
  bx lanes along y (contiguous in memory), by rows along x; it trades
  coalescing width against reuse of the 6-point x stencils in L1/L2;
* the heavy-ghost schedule: 1, 2 or 4 ghosts per workgroup (one wavefront
  each), trading launch granularity against LDS footprint.

Search: coordinate descent (shape first with one ghost per workgroup, then
the ghost schedule at the best shape), each candidate timed over a few RK3
steps with hipEvents after one untimed step that rebuilds the graph.  The
caller restores its state afterwards (tuning advances the solution).
Results can be cached in a JSON file keyed by the problem signature.
"""
import json
import os
import time

SHAPES = [(64, 4), (32, 8), (128, 2), (64, 2), (32, 4), (16, 16), (64, 1)]
GHOSTS_PER_BLOCK = [1, 2, 4]


def signature(grids, arith):
    return "|".join(f"{nx}x{ny}" for nx, ny in sorted(grids)) + f"|arith={arith}"


def _time(runner, steps):
    runner.advance(1)  # rebuilds the graph for the new launch shape
    t0 = time.perf_counter()
    runner.advance(steps)
    return (time.perf_counter() - t0) / steps


def tune(runner, steps=3, shapes=SHAPES, ghost_options=GHOSTS_PER_BLOCK, cache_path=None, key=None, log=None):
    """runner: object with set_launch(bx, by, ghosts_per_block) and
    advance(nsteps).  Returns {"bx", "by", "ghosts_per_block", "ms_per_step",
    "trials"}."""
    if cache_path and key and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            c = cache[key]
            runner.set_launch(c["bx"], c["by"], c["ghosts_per_block"])
            return dict(c, cached=True)
    trials = []
    best = None
    for bx, by in shapes:
        runner.set_launch(bx, by, 1)
        t = _time(runner, steps)
        trials.append({"bx": bx, "by": by, "ghosts_per_block": 1, "ms_per_step": t * 1e3})
        if log:
            log(f"autotune shape {bx}x{by}: {t * 1e3:.3f} ms/step")
        if best is None or t < best[0]:
            best = (t, bx, by, 1)
    _, bx, by, _ = best
    for g in ghost_options:
        if g == 1:
            continue
        runner.set_launch(bx, by, g)
        t = _time(runner, steps)
        trials.append({"bx": bx, "by": by, "ghosts_per_block": g, "ms_per_step": t * 1e3})
        if log:
            log(f"autotune ghosts/wg {g}: {t * 1e3:.3f} ms/step")
        if t < best[0]:
            best = (t, bx, by, g)
    t, bx, by, g = best
    runner.set_launch(bx, by, g)
    res = {"bx": bx, "by": by, "ghosts_per_block": g, "ms_per_step": t * 1e3, "trials": trials}
    if cache_path and key:
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
        cache[key] = {k: res[k] for k in ("bx", "by", "ghosts_per_block", "ms_per_step")}
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return res
