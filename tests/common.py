"""Shared helpers for the parity tests: fixture loading and the gate-case
replayer that drives any engine exposing the reference's API names (the
oracle's State, or the HIP engine through its C-ABI)."""
import json
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def hexarr(xs):
    return np.array([float.fromhex(x) for x in xs], dtype=np.float64)


def golden_state(d):
    return hexarr(d["re"]), hexarr(d["im"])


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def assert_bitwise(got, want, what=""):
    g, w = bits(got), bits(want)
    if not np.array_equal(g, w):
        k = int(np.nonzero(g != w)[0][0])
        raise AssertionError(f"{what}: first difference at {k}: got {got[k]!r} want {want[k]!r}")


def prep(eng, n):
    """The fixed preparation used by oracle/ref_harness.cpp (prep())."""
    eng.init_state()
    for q in range(n):
        eng.h(q)
    for q in range(n):
        eng.pauli([q], [2], 0.31 * (q + 1))
    for q in range(n):
        eng.u1(q, 0.57 * (q + 1))
    eng.cx(0, 2)
    eng.cx(3, 1)
    eng.pauli([1, 4], [1, 3], 0.77)
    eng.pauli([0, 3], [2, 2], -0.41)


def run_case(eng, c, n):
    """Replays one golden gate case; returns (error, out, outc)."""
    prep(eng, n)
    gc = c["gc"]
    if gc:
        eng.activate_gc_mask(gc)
    q, aux, val = c["qs"], c["aux"], float.fromhex(c["val"])
    op = c["op"]
    err, out, outc = False, 0.0, 0
    if op == "x":
        eng.x(q[0])
    elif op == "y":
        eng.y(q[0])
    elif op == "z":
        eng.u1(q[0], math.pi)
    elif op == "h":
        eng.h(q[0])
    elif op == "s":
        eng.u1(q[0], math.pi / 2.0)
    elif op == "t":
        eng.u1(q[0], math.pi / 4.0)
    elif op == "tdg":
        eng.u1(q[0], -math.pi / 4.0)
    elif op == "u1":
        eng.u1(q[0], val, aux[0] if aux else 1)
    elif op == "cx":
        eng.cx(q[0], q[1], aux[0] if aux else 1)
    elif op == "mcx":
        eng.mcx(q[:-1], q[-1], aux if aux else None)
    elif op == "cu1":
        eng.cu1(q[0], q[1], val, aux[0] if aux else 1)
    elif op == "mcu1":
        eng.mcu1(q[:-1], q[-1], val, aux if aux else None)
    elif op == "swap":
        eng.swap(q[0], q[1])
    elif op == "pauli":
        err = eng.pauli(q, aux, val) != 0
    elif op == "qft":
        eng.qft(q)
    elif op == "qft_inverse":
        eng.qft_inverse(q)
    elif op == "measure":
        outc = eng.measure(q[0], val)
    elif op == "reset":
        eng.reset(q[0], val)
    elif op == "prob_filter":
        out = eng.prob_filter(q, aux)
    elif op == "vnorm":
        out = eng.vnorm()
    elif op == "vnormalize":
        eng.vnormalize()
    else:
        raise ValueError(op)
    if gc:
        eng.deactivate_gc_mask(gc)
    return err, out, outc


# (beta, steps, reset_each, r, seed, emin, emax, pe, therm, maxrev, recrev) of each QMS fixture
def qms_params(d):
    return dict(beta=float.fromhex(d["beta"]), steps=d["steps"], reset_each=d["reset_each"], r=d["r"],
                seed=d["seed"], ene_min=float.fromhex(d["ene_min"]), ene_max=float.fromhex(d["ene_max"]),
                pe_steps=d["pe_steps"], thermalization=d["thermalization"], max_reverse=d["max_reverse"],
                record_reverse=bool(d["record_reverse"]))


QMS_FIXTURES = ["qms_c1.json", "qms_r3_b05.json", "qms_r3_maxrev1.json", "qms_r5_b3.json", "qms_r6_b01.json"]


def pairwise_sum(x):
    """The harness's pairwise sum (oracle/ref_harness.cpp: recursive halving,
    ranges of at most 8 summed left to right) of a power-of-two-sized array,
    in the same order, so bit-identical for identical inputs."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    if n <= 8:
        s = 0.0
        for v in x:
            s += float(v)
        return s
    a = x.reshape(-1, 8)
    s = np.zeros(a.shape[0])
    for k in range(8):
        s = s + a[:, k]
    while s.size > 1:
        s = s[0::2] + s[1::2]
    return float(s[0])


def SAMPLE_INDEX(k, N):
    return (k * 2654435761 + 12345) % N


def state_summary(re, im):
    """(sum re, sum im, sum i |psi_i|^2 / N) by pairwise sums and the 64
    sample amplitudes, as oracle/ref_harness.cpp's circuit_summary writes them."""
    N = re.size
    idx = np.arange(N, dtype=np.float64)
    sums = (pairwise_sum(re), pairwise_sum(im), pairwise_sum(idx * (re * re + im * im)) / N)
    samples = [(SAMPLE_INDEX(k, N), re[SAMPLE_INDEX(k, N)], im[SAMPLE_INDEX(k, N)]) for k in range(64)]
    return sums, samples


def golden_summary(d):
    sums = tuple(float.fromhex(d[k]) for k in ("sum_re", "sum_im", "sum_index_prob"))
    samples = [(int(i), float.fromhex(r), float.fromhex(m)) for (i, r, m) in d["samples"]]
    return sums, samples
