"""The autotuner's search (selas_amd/autotune.py) on a simulated device: a probe whose cost is a
known function of the launch shape, on a simulated clock."""
import pytest

from selas_amd import autotune as A


class SimCtx:
    def __init__(self):
        self.t = {"vcm_wavefront_paths": 1 << 20, "trace_blocks_per_cu": 8}
        self.now = 0.0
        self.sets = 0

    def tuning(self):
        return dict(self.t)

    def set_tuning(self, vcm_wavefront_paths=None, trace_blocks_per_cu=None):
        self.sets += 1
        self.t = {"vcm_wavefront_paths": vcm_wavefront_paths, "trace_blocks_per_cu": trace_blocks_per_cu}

    def device_info(self):
        return {"name": "sim", "free_bytes": 100 << 30}


class SimCtxLds(SimCtx):
    def __init__(self):
        super().__init__()
        self.t["trace_lds_nodes"] = 0

    def set_tuning(self, vcm_wavefront_paths=None, trace_blocks_per_cu=None, trace_lds_nodes=None):
        self.sets += 1
        lds = self.t["trace_lds_nodes"] if trace_lds_nodes is None else trace_lds_nodes
        self.t = {"vcm_wavefront_paths": vcm_wavefront_paths, "trace_blocks_per_cu": trace_blocks_per_cu,
                  "trace_lds_nodes": lds}


def cost(t):
    # falls with the wavefront size until 2^22, then rises; best grid 16; LDS top nodes save 1 %
    import math
    k = math.log2(t["vcm_wavefront_paths"])
    c = 1.0 + 0.05 * (k - 22) ** 2 + 0.01 * abs(t["trace_blocks_per_cu"] - 16) / 4
    return c * (0.99 if t.get("trace_lds_nodes") else 1.0)


def run(ctx, budget=None, footprint=lambda c, s, L: s * 1000):
    def probe(c):
        c.now += cost(c.t)
    return A.tune(ctx, probe, memory_budget=budget, footprint=footprint, timer=lambda: ctx.now, repeats=2)


def test_finds_the_best_shape_and_sets_it():
    ctx = SimCtx()
    r = run(ctx)
    assert r["best"] == {"vcm_wavefront_paths": 1 << 22, "trace_blocks_per_cu": 16}
    assert ctx.tuning() == r["best"]
    assert r["seconds"] == pytest.approx(1.0)
    assert r["default_seconds"] == pytest.approx(cost({"vcm_wavefront_paths": 1 << 20, "trace_blocks_per_cu": 8}))
    # every shape measured once
    shapes = [(t["vcm_wavefront_paths"], t["trace_blocks_per_cu"]) for t in r["trials"]]
    assert len(shapes) == len(set(shapes))


def test_memory_budget_bounds_the_wavefront():
    ctx = SimCtx()
    r = run(ctx, budget=(1 << 21) * 1000)
    assert r["best"]["vcm_wavefront_paths"] == 1 << 21
    assert all(t["vcm_wavefront_paths"] <= 1 << 21 for t in r["trials"])
    with pytest.raises(RuntimeError):
        run(SimCtx(), budget=10)


def test_cache_roundtrip(tmp_path):
    p = str(tmp_path / "t.json")
    assert A.load(p, "k") is None
    A.save(p, "k", {"vcm_wavefront_paths": 8, "trace_blocks_per_cu": 4})
    A.save(p, "j", {"vcm_wavefront_paths": 16, "trace_blocks_per_cu": 8})
    assert A.load(p, "k") == {"vcm_wavefront_paths": 8, "trace_blocks_per_cu": 4}
    assert A.load(p, "j")["trace_blocks_per_cu"] == 8


def test_lds_nodes_tried_at_the_best_shape():
    ctx = SimCtxLds()
    r = run(ctx)
    assert r["best"] == {"vcm_wavefront_paths": 1 << 22, "trace_blocks_per_cu": 16, "trace_lds_nodes": 9}
    assert ctx.tuning() == r["best"]
    assert r["seconds"] == pytest.approx(0.99)
