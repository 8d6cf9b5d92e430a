"""Launch-shape autotuner for the batched and plain solvers.

Three knobs are tuned on the actual problem set:

* the workgroup shape (bx, by) of the fused y-flux + assembly pass -- bx
  lanes along y (contiguous in memory), by rows along x; rows of 16, 32 or
  64 lanes keep the y fluxes inside the pass (other widths split it into a
  y-flux kernel and an assembly kernel);
* the x-flux kernel's own shape (which follows the first unless set): it has
  no row constraint, and its 6-row stencils favour other shapes than the
  fused pass's 9-row noncon lines;
* the heavy-ghost schedule: 1, 2 or 4 ghosts per workgroup (one wavefront
  each), trading launch granularity against LDS footprint.

Search: coordinate descent -- the fused shape with the x flux following it
and one ghost per workgroup, then the x-flux shape, then the ghost
schedule -- each candidate timed by the wall clock around `advance(steps)`
after one untimed step that rebuilds the step graph for the new shapes.
`advance` replays the captured step graph and returns after the GPU is
done, so the figure is the production loop's cost per step.  The caller
restores its state afterwards (tuning advances the solution).  Results can
be cached in a JSON file keyed by the problem signature.
"""
import json
import os
import time

SHAPES = [(64, 4), (32, 8), (128, 2), (64, 2), (32, 4), (16, 16), (64, 1)]
X_SHAPES = [(0, 0), (64, 4), (32, 8), (64, 2), (16, 16), (128, 2), (64, 1)]
GHOSTS_PER_BLOCK = [1, 2, 4]


def signature(grids, arith):
    return "|".join(f"{nx}x{ny}" for nx, ny in sorted(grids)) + f"|arith={arith}"


def _time(runner, steps):
    runner.advance(1)  # rebuilds the graph for the new launch shape
    t0 = time.perf_counter()
    runner.advance(steps)
    return (time.perf_counter() - t0) / steps


def tune(runner, steps=3, shapes=SHAPES, x_shapes=X_SHAPES, ghost_options=GHOSTS_PER_BLOCK, cache_path=None,
         key=None, log=None):
    """runner: object with set_launch(bx, by, ghosts_per_block, x_shape) and
    advance(nsteps).  Returns {"bx", "by", "x_bx", "x_by",
    "ghosts_per_block", "ms_per_step", "trials"} ((x_bx, x_by) = (0, 0): the
    x flux follows (bx, by))."""
    fields = ("bx", "by", "x_bx", "x_by", "ghosts_per_block", "ms_per_step")
    if cache_path and key and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            c = cache[key]
            runner.set_launch(c["bx"], c["by"], c["ghosts_per_block"], (c["x_bx"], c["x_by"]))
            return dict(c, cached=True)
    trials = []

    def trial(bx, by, xs, g, what):
        runner.set_launch(bx, by, g, xs)
        t = _time(runner, steps)
        trials.append({"bx": bx, "by": by, "x_bx": xs[0], "x_by": xs[1], "ghosts_per_block": g,
                       "ms_per_step": t * 1e3})
        if log:
            log(f"autotune {what}: {t * 1e3:.3f} ms/step")
        return t

    best = None
    for bx, by in shapes:
        t = trial(bx, by, (0, 0), 1, f"shape {bx}x{by}")
        if best is None or t < best[0]:
            best = (t, bx, by, (0, 0), 1)
    _, bx, by, _, _ = best
    for xs in x_shapes:
        if xs == (0, 0) or xs == (bx, by):
            continue
        t = trial(bx, by, xs, 1, f"x-flux shape {xs[0]}x{xs[1]}")
        if t < best[0]:
            best = (t, bx, by, xs, 1)
    _, bx, by, xs, _ = best
    for g in ghost_options:
        if g == 1:
            continue
        t = trial(bx, by, xs, g, f"ghosts/wg {g}")
        if t < best[0]:
            best = (t, bx, by, xs, g)
    t, bx, by, xs, g = best
    runner.set_launch(bx, by, g, xs)
    res = {"bx": bx, "by": by, "x_bx": xs[0], "x_by": xs[1], "ghosts_per_block": g, "ms_per_step": t * 1e3,
           "trials": trials}
    if cache_path and key:
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
        cache[key] = {k: res[k] for k in fields}
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return res
