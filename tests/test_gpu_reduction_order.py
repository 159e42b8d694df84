"""The order of the small-state measurement sums, pinned bit for bit.

States of at most 2^14 amplitudes measure in one workgroup (the measurement
kernel, and the programs of states of one pass tile).  Their probability sums
follow one fixed order: thread t < 256 adds |psi_i|^2 = re*re + im*im (no
fused multiply-add) for i = t, t + 256, ... in increasing i, separately for
the amplitudes with the measured bit 0 and 1; each wavefront of 64 such
partials is reduced by the lane-xor butterfly v(l) += v(l ^ o) for
o = 32, 16, 8, 4, 2, 1; the four wavefronts' sums are added in order from 0.0.
The outcome is rd < p1 (suqa.cu:690) and the kept half is multiplied by
1 / sqrt(p) (a reciprocal multiply, suqa.cu:694-695).  This test restates that
order in numpy (IEEE float64 throughout) and compares the device's collapsed
state with it bit for bit, so the device's lane moves (DPP, permlane swaps)
and LDS staging are pinned to the stated pairing, not only to each other."""
import numpy as np
import pytest

from tests.common import assert_bitwise

pytestmark = pytest.mark.gpu


ORDER = (32, 16, 8, 4, 2, 1)


def butterfly(v, order=ORDER):
    lanes = np.arange(64)
    v = v.copy()
    for o in order:
        v = v + v[lanes ^ o]
    return v[0]


def measured(re, im, q, rd, order=ORDER):
    N = re.size
    p = re * re + im * im
    idx = np.arange(N)
    s0 = np.zeros(256)
    s1 = np.zeros(256)
    for t in range(min(N, 256)):
        for i in range(t, N, 256):   # increasing i
            if (i >> q) & 1:
                s1[t] += p[i]
            else:
                s0[t] += p[i]
    p0 = 0.0
    p1 = 0.0
    for w in range(4):
        p0 += butterfly(s0[64 * w:64 * w + 64], order)
        p1 += butterfly(s1[64 * w:64 * w + 64], order)
    c = 1 if rd < p1 else 0
    scale = 1.0 / np.sqrt(p1 if c else p0)
    keep = ((idx >> q) & 1) == c
    return c, np.where(keep, re * scale, 0.0), np.where(keep, im * scale, 0.0)


def measured_seq(re, im, qs, rds, order=ORDER):
    """Measurements of qs in order, each on the state the previous one left."""
    cs = []
    for q, rd in zip(qs, rds):
        c, re, im = measured(re, im, q, rd, order)
        cs.append(c)
    return cs, re, im


def random_case(rng, n):
    N = 1 << n
    re = rng.standard_normal(N)
    im = rng.standard_normal(N)
    nrm = np.sqrt(np.sum(re * re + im * im))
    qs = [int(q) for q in rng.permutation(n)[:3]]
    rds = [float(x) for x in rng.uniform(0.2, 0.8, 3)]
    return re / nrm, im / nrm, qs, rds


@pytest.mark.parametrize("n,tile_bits", [(6, None), (8, None), (8, 0), (10, None), (12, None), (14, None)])
def test_measurement_sums_follow_the_stated_order(n, tile_bits):
    from suqa_amd import capi
    print(f"reduction order n={n} tile_bits={tile_bits}", flush=True)
    rng = np.random.default_rng(500 + n)
    eng = capi.Engine(n, tile_bits=tile_bits)
    for rep in range(12):
        re, im, qs, rds = random_case(rng, n)
        eng.set(re, im)
        r0, i0 = eng.get()
        cs = eng.measure_qbits(qs, rds)
        cs_ref, r_ref, i_ref = measured_seq(r0, i0, qs, rds)
        assert cs == cs_ref
        r1, i1 = eng.get()
        assert_bitwise(r1, r_ref, f"re n={n} rep={rep}")
        assert_bitwise(i1, i_ref, f"im n={n} rep={rep}")
    eng.close()
