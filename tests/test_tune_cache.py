"""The autotuner's JSON cache (suqa_amd/tune_cache.py), on the CPU with stand-in
engines: a miss runs the autotuner and stores its choice, a hit applies the
stored choice without tuning, and the key separates devices, workloads, state
sizes and library builds."""
import json

from suqa_amd import tune_cache


class FakeEngine:
    def __init__(self, nq):
        self.nq = nq
        self.tuned = 0
        self.bits = None

    def autotune_tile_bits(self, gates, candidates):
        self.tuned += 1
        self.bits = 11
        return 11, {c: 1.0 / (1 + c) for c in candidates}

    def set_tile_bits(self, b):
        self.bits = b


class FakeF32:
    def __init__(self, nq):
        self.nq = nq
        self.passes = 0
        self.tile_bits = 12
        self.tuned = 0

    def autotune(self, sample):
        self.tuned += 1
        self.passes, self.tile_bits = 3, 12
        return 4, {"blocks4": 0.7, "gates12": 0.19}


def test_tile_bits_miss_then_hit(tmp_path):
    path = str(tmp_path / "tc.json")
    e = FakeEngine(30)
    bits, tms, how = tune_cache.tile_bits(e, None, (0, 8, 11), "circuit", path)
    assert (bits, how, e.tuned) == (11, "miss", 1)
    e2 = FakeEngine(30)
    bits2, tms2, how2 = tune_cache.tile_bits(e2, None, (0, 8, 11), "circuit", path)
    assert (bits2, how2, e2.tuned, e2.bits) == (11, "hit", 0, 11)
    assert tms2 == tms
    # another state size or workload tunes again
    e3 = FakeEngine(28)
    assert tune_cache.tile_bits(e3, None, (0, 8, 11), "circuit", path)[2] == "miss"
    assert tune_cache.tile_bits(FakeEngine(30), None, (0, 8, 11), "trotter", path)[2] == "miss"
    # a stored choice outside the candidates is not applied
    e4 = FakeEngine(30)
    assert tune_cache.tile_bits(e4, None, (0, 8, 9), "circuit", path)[2] == "miss"
    with open(path) as f:
        assert len(json.load(f)) == 3


def test_f32_config_round_trip(tmp_path):
    path = str(tmp_path / "tc.json")
    f = FakeF32(30)
    assert tune_cache.f32_config(f, None, "circuit", path)[2] == "miss"
    g = FakeF32(30)
    width, tms, how = tune_cache.f32_config(g, None, "circuit", path)
    assert (width, how, g.tuned, g.passes, g.tile_bits) == (4, "hit", 0, 3, 12)


def test_key_parts():
    k1 = tune_cache.make_key("w", 30, "f64", device="gpuA", lib_sig="1-2")
    assert k1 != tune_cache.make_key("w", 30, "f64", device="gpuB", lib_sig="1-2")
    assert k1 != tune_cache.make_key("w", 30, "f64", device="gpuA", lib_sig="1-3")
    assert k1 != tune_cache.make_key("w", 29, "f64", device="gpuA", lib_sig="1-2")


def test_unreadable_cache_is_empty(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text("{not json")
    assert tune_cache.load(str(p)) == {}
    tune_cache.put("k", {"tile_bits": 9}, str(p))
    assert tune_cache.get("k", str(p)) == {"tile_bits": 9}
