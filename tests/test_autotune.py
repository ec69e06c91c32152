"""Host logic of the autotuner (no GPU): candidates whose integrals differ from
the default configuration's are never chosen, the fastest accepted value wins
per knob, and choices round-trip through the JSON cache."""

import numpy as np

from giaoint_amd import autotune


class FakeCtx:
    def __init__(self):
        self.values = {}
        self.allocs = 0

    def set_fast_waves(self, v):
        self.values["fast_waves"] = v

    def set_eri_queue(self, v):
        self.values["eri_queue"] = v

    def set_cholesky_block(self, v):
        self.values["cholesky_block"] = v

    def alloc(self, nbytes):
        self.allocs += 1
        return 1

    def free(self, ptr):
        self.allocs -= 1

    def synchronize(self):
        pass


class FakeBasis:
    """Integrals depend on the workgroup size (by `drift`) and, wrongly, on
    the 'single' launch form; run time is a table lookup."""

    def __init__(self, seconds, drift, wrong_form="single"):
        self.ctx = FakeCtx()
        self.seconds = seconds
        self.drift = drift
        self.wrong_form = wrong_form
        self.clock = [0.0]

    def quartet_sizes(self, q):
        return np.full(len(q), 4, dtype=np.int64)

    def eri_batch(self, q, field6, device_out=None):
        v = self.ctx.values
        self.clock[0] += self.seconds.get(("fast_waves", v.get("fast_waves")), 0.0)
        self.clock[0] += self.seconds.get(("eri_queue", v.get("eri_queue")), 0.0)
        if device_out is not None:
            return None
        vals = np.arange(4 * len(q), dtype=np.complex128)
        vals = vals + self.drift.get(v.get("fast_waves"), 0.0)
        if v.get("eri_queue") == self.wrong_form:
            vals = vals + 1e-15
        return vals


def test_tune_eri_takes_fastest_accepted(monkeypatch):
    seconds = {("fast_waves", 0): 3.0, ("fast_waves", 4): 1.0, ("fast_waves", 5): 2.0, ("fast_waves", 6): 0.5,
               ("fast_waves", 7): 2.5, ("eri_queue", "none"): 2.0, ("eri_queue", "buckets"): 1.5,
               ("eri_queue", "single"): 0.1}
    # 6 waves is fastest but its integrals drift by 1e-9: rejected
    basis = FakeBasis(seconds, drift={6: 1e-9, 4: 1e-13})
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: basis.clock[0])
    q = np.zeros((10, 4), dtype=np.uint32)
    choice = autotune.tune_eri(basis, q, np.zeros(6), reps=2)
    assert choice.values == {"fast_waves": 4, "eri_queue": "buckets"}
    by = {(t.knob, t.value): t for t in choice.trials}
    assert not by[("fast_waves", 6)].accepted and by[("fast_waves", 6)].seconds == float("inf")
    # the launch forms must be bitwise equal: 'single' is off by 1e-15 and loses although fastest
    assert not by[("eri_queue", "single")].accepted
    assert by[("fast_waves", 4)].accepted and by[("fast_waves", 4)].max_abs_diff <= 1e-11
    assert basis.ctx.values == {"fast_waves": 4, "eri_queue": "buckets"}
    assert basis.ctx.allocs == 0  # the tuner's output buffer is freed


def test_cache_round_trip(tmp_path):
    choice = autotune.Choice(workload="w", values={"fast_waves": 5, "eri_queue": "none"})
    p = autotune.save("gfx950", choice, tmp_path / "t.json")
    assert autotune.load("gfx950", "w", p) == {"fast_waves": 5, "eri_queue": "none"}
    assert autotune.load("gfx950", "other", p) is None
    assert autotune.load("gfx942", "w", p) is None
