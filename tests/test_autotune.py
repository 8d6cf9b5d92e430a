"""The launch-shape autotuner's search and cache, with a stand-in runner
whose cost per step is a known function of the shapes (no GPU)."""
import importlib
import json

autotune = importlib.import_module("3t_rh_amd.autotune")


class FakeRunner:
    """Cost: the fused shape (32, 8) is best, the x flux's own (16, 16) is
    better than following it, two ghosts per workgroup beat one and four."""

    def __init__(self):
        self.shape = None
        self.calls = []

    def set_launch(self, bx, by, g, x_shape):
        self.shape = (bx, by, g, tuple(x_shape))
        self.calls.append(self.shape)

    def cost(self):
        bx, by, g, xs = self.shape
        fused = {(32, 8): 5.0}.get((bx, by), 6.0 + 0.01 * bx)
        xflux = {(16, 16): 1.0, (0, 0): 1.5}.get(xs, 2.0)
        ghosts = {1: 0.5, 2: 0.2, 4: 0.7}[g]
        return fused + xflux + ghosts


def test_coordinate_descent_picks_each_minimum(monkeypatch):
    r = FakeRunner()
    monkeypatch.setattr(autotune, "_time", lambda runner, steps: runner.cost() * 1e-3)
    res = autotune.tune(r, steps=3)
    assert (res["bx"], res["by"]) == (32, 8)
    assert (res["x_bx"], res["x_by"]) == (16, 16)
    assert res["ghosts_per_block"] == 2
    assert abs(res["ms_per_step"] - (5.0 + 1.0 + 0.2)) < 1e-12
    assert r.shape == (32, 8, 2, (16, 16))  # left at the choice
    # every fused shape is tried once with the x flux following it
    assert sum(1 for t in res["trials"] if (t["x_bx"], t["x_by"]) == (0, 0)) == len(autotune.SHAPES)


def test_cache_round_trip(monkeypatch, tmp_path):
    path = str(tmp_path / "tune.json")
    r = FakeRunner()
    monkeypatch.setattr(autotune, "_time", lambda runner, steps: runner.cost() * 1e-3)
    first = autotune.tune(r, cache_path=path, key="k")
    with open(path) as f:
        stored = json.load(f)["k"]
    assert stored == {k: first[k] for k in ("bx", "by", "x_bx", "x_by", "ghosts_per_block", "ms_per_step")}
    r2 = FakeRunner()
    again = autotune.tune(r2, cache_path=path, key="k")
    assert again["cached"] and r2.calls == [(32, 8, 2, (16, 16))]
