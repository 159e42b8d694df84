"""suqa_hip_prob_filter_noalloc (the entry point that allocates nothing: its
reduction partials live in a fixed device array of the library) against
suqa_hip_prob_filter on the same state, through the C-ABI: bit-identical
probabilities on two calls in a row, the second on a smaller state, over
several filters each; sizes beyond the array are rejected on the host."""
import ctypes as C

import numpy as np
import pytest

from suqa_amd import capi

pytestmark = pytest.mark.gpu

PD = C.POINTER(C.c_double)
PU32 = C.POINTER(C.c_uint32)


def hip_runtime():
    rt = C.CDLL("libamdhip64.so.7")   # the runtime libsuqa_hip.so is linked against (already loaded)
    rt.hipMemcpy.restype = C.c_int
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.restype = C.c_int
    return rt


def prepared(n, seed):
    """A state with every amplitude non-zero: H layer and a random circuit."""
    from suqa_amd import workloads as wl
    eng = capi.Engine(n)
    for q in range(n):
        eng.h(q)
    gates, _, _ = wl.random_circuit(n, 3, seed)
    eng.run_gates(capi.gates_from(gates))
    eng.synchronize()
    return eng


def noalloc(eng, qs, vals, d_out, rt):
    nq = C.c_uint32()
    N = C.c_uint64()
    re, im = PD(), PD()
    capi.check(capi.lib().suqa_hip_state_info(eng.ctx, C.byref(nq), C.byref(N), C.byref(re), C.byref(im)))
    qa = (C.c_uint32 * max(len(qs), 1))(*qs)
    va = (C.c_uint32 * max(len(vals), 1))(*vals)
    capi.check(capi.lib().suqa_hip_prob_filter_noalloc(re, im, nq.value, qa, va, len(qs), C.cast(d_out, PD), None))
    assert rt.hipDeviceSynchronize() == 0
    out = C.c_double()
    assert rt.hipMemcpy(C.byref(out), d_out, 8, 2) == 0   # hipMemcpyDeviceToHost
    return out.value


@pytest.mark.parametrize("sizes", [(22, 12), (18, 7)])
def test_noalloc_matches_prob_filter(sizes):
    rt = hip_runtime()
    d_out = C.c_void_p()
    capi.check(capi.lib().suqa_hip_device_alloc(8, C.byref(d_out)))
    try:
        for call, n in enumerate(sizes):
            print(f"prob_filter_noalloc call {call}: n={n}", flush=True)
            eng = prepared(n, 40 + n)
            try:
                rng = np.random.default_rng(n)
                filters = [([], []), ([0], [1]), ([1, 3], [0, 1]), ([n - 1], [0]), ([2, n - 2, 0], [1, 1, 0])]
                for _ in range(3):
                    k = int(rng.integers(1, min(n, 6) + 1))
                    qs = [int(q) for q in rng.choice(n, size=k, replace=False)]
                    filters.append((qs, [int(v) for v in rng.integers(0, 2, size=k)]))
                for qs, vals in filters:
                    want = eng.prob_filter(qs, vals)
                    got = noalloc(eng, qs, vals, d_out, rt)
                    assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64), (n, qs, vals, got, want)
                assert abs(noalloc(eng, [], [], d_out, rt) - 1.0) < 1e-12
            finally:
                eng.close()
    finally:
        capi.lib().suqa_hip_device_free(d_out)


def test_noalloc_rejects_sizes_beyond_the_array():
    """Rejected on the host before any device work (the pointers are never used)."""
    dummy = C.cast(C.c_void_p(0x1000), PD)
    rc = capi.lib().suqa_hip_prob_filter_noalloc(dummy, dummy, 27, None, None, 0, dummy, None)
    assert rc == -1
    assert b"26" in capi.lib().suqa_hip_last_error()
