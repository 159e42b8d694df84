"""Parity of the HIP engine's one-sweep-per-gate kernels (through the C-ABI,
tile passes off) with the reference's golden vectors and the oracle.  Gates are bit-exact; anything downstream of a
reduction (measure, reset, vnormalize) is within 1e-12 with identical
outcomes; prob_filter/vnorm values within 1e-14 relative."""
import math
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.common import assert_bitwise, golden_state, load, prep, run_case

pytestmark = pytest.mark.gpu

REDUCING = {"measure", "reset", "vnormalize"}


@pytest.fixture(scope="module")
def capi():
    from suqa_amd import capi
    return capi


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_gate_cases_vs_golden(capi):
    g = load("gates_n5.json")
    n = g["n"]
    log(f"gate cases n={n}")
    eng = capi.Engine(n, tile_bits=0)
    eng.qft(list(range(5)))
    prep(eng, n)
    re, im = golden_state(g["prep"])
    gre, gim = eng.get()
    assert_bitwise(gre, re, "prep re")
    assert_bitwise(gim, im, "prep im")
    for c in g["cases"]:
        tag = f"{c['op']} qs={c['qs']} aux={c['aux']} gc={c['gc']}"
        err, out, outc = run_case(eng, c, n)
        assert err == c["error"], tag
        assert outc == c["outc"], tag
        want_out = float.fromhex(c["out"])
        assert abs(out - want_out) <= 1e-14 * max(1.0, abs(want_out)), tag
        re, im = golden_state(c["state"])
        gre, gim = eng.get()
        if c["op"] in REDUCING:
            np.testing.assert_allclose(gre, re, atol=1e-12, rtol=0, err_msg=tag)
            np.testing.assert_allclose(gim, im, atol=1e-12, rtol=0, err_msg=tag)
        else:
            assert_bitwise(gre, re, tag + " re")
            assert_bitwise(gim, im, tag + " im")
        assert eng.gc_mask == 0
    eng.close()


@pytest.mark.parametrize("fx", ["circuit_n6.json", "circuit_n10.json"])
def test_circuit_vs_golden(capi, fx):
    d = load(fx)
    n = d["n"]
    log(f"circuit golden {fx}")
    gates, ng, mq, mr = po.circuit(n, d["layers"], d["seed"])
    eng = capi.Engine(n, tile_bits=0)
    for q in range(n):
        eng.h(q)
    eng.run_gates(capi.gates_from(gates[:ng]))
    re, im = golden_state(d["pre_measure"])
    gre, gim = eng.get()
    assert_bitwise(gre, re, "pre re")
    assert_bitwise(gim, im, "pre im")
    for (q, r, c) in d["measure"]:
        assert eng.measure(q, float.fromhex(r)) == c
    re, im = golden_state(d["final"])
    gre, gim = eng.get()
    np.testing.assert_allclose(gre, re, atol=1e-12, rtol=0)
    np.testing.assert_allclose(gim, im, atol=1e-12, rtol=0)
    eng.close()


def test_circuit_n18_vs_oracle(capi):
    n, layers, seed = 18, 200, 7
    log(f"circuit n={n} layers={layers} vs oracle")
    gates, ng, mq, mr = po.circuit(n, layers, seed)
    st = po.State(n)
    eng = capi.Engine(n, tile_bits=0)
    for q in range(n):
        st.h(q)
        eng.h(q)
    for g in gates[:ng]:
        st.apply_gate(g)
    eng.run_gates(capi.gates_from(gates[:ng]))
    gre, gim = eng.get()
    assert_bitwise(gre, st.re(), "re")
    assert_bitwise(gim, st.im(), "im")
    for q, r in zip(mq, mr):
        assert eng.measure(q, r) == st.measure(q, r)
    gre, gim = eng.get()
    np.testing.assert_allclose(gre, st.re(), atol=1e-12, rtol=0)
    np.testing.assert_allclose(gim, st.im(), atol=1e-12, rtol=0)
    assert abs(eng.vnorm() - 1.0) < 1e-12
    eng.close()


def test_pauli_general_vs_oracle(capi):
    """Any-weight exp(i phi P) in one pass vs the oracle's CNOT-ladder circuit."""
    n = 9
    rng = np.random.default_rng(5)
    eng = capi.Engine(n, tile_bits=0)
    st = po.State(n)
    v = rng.normal(size=(2, 1 << n))
    v /= np.linalg.norm(v)
    for case in range(60):
        w = int(rng.integers(1, 7))
        qs = [int(x) for x in rng.choice(n, size=w, replace=False)]
        types = [int(x) for x in rng.integers(0, 4, size=w)]
        free = [q for q in range(n) if q not in qs]
        gc = [int(x) for x in rng.choice(free, size=int(rng.integers(0, 3)), replace=False)]
        phi = float(rng.uniform(-4, 4))
        eng.set(v[0], v[1])
        st.set(v[0], v[1])
        eng.gc_mask = 0
        eng.activate_gc_mask(gc)
        st.activate_gc_mask(gc)
        eng.pauli_general(qs, types, phi)
        assert st.pauli_general(qs, types, phi) == 0
        eng.deactivate_gc_mask(gc)
        st.deactivate_gc_mask(gc)
        gre, gim = eng.get()
        np.testing.assert_allclose(gre, st.re(), atol=1e-13, rtol=0, err_msg=f"case {case}")
        np.testing.assert_allclose(gim, st.im(), atol=1e-13, rtol=0, err_msg=f"case {case}")
    eng.close()


def test_reference_pauli_patterns_bitwise_vs_oracle(capi):
    n = 7
    rng = np.random.default_rng(11)
    eng = capi.Engine(n, tile_bits=0)
    st = po.State(n)
    v = rng.normal(size=(2, 1 << n))
    v /= np.linalg.norm(v)
    for w in (1, 2, 3):
        for _ in range(40):
            qs = [int(x) for x in rng.choice(n, size=w, replace=False)]
            types = [int(x) for x in rng.integers(1, 4, size=w)]
            gc = [q for q in range(n) if q not in qs][:int(rng.integers(0, 2))]
            eng.set(v[0], v[1])
            st.set(v[0], v[1])
            eng.activate_gc_mask(gc)
            st.activate_gc_mask(gc)
            r1 = eng.pauli(qs, types, 0.777)
            r2 = st.pauli(qs, types, 0.777)
            eng.deactivate_gc_mask(gc)
            st.deactivate_gc_mask(gc)
            assert (r1 == 0) == (r2 == 0)
            gre, gim = eng.get()
            assert_bitwise(gre, st.re(), f"{qs} {types} re")
            assert_bitwise(gim, st.im(), f"{qs} {types} im")
    eng.close()


def test_phase_list_and_swap_vs_oracle(capi):
    n = 8
    rng = np.random.default_rng(3)
    eng = capi.Engine(n, tile_bits=0)
    st = po.State(n)
    v = rng.normal(size=(2, 1 << n))
    v /= np.linalg.norm(v)
    for gc in ([], [7], [0, 6]):
        for (q0, qsz) in ((1, 3), (0, 1), (4, 4)):
            ph = [float(x) for x in rng.uniform(-3, 3, size=1 << qsz)]
            eng.set(v[0], v[1])
            st.set(v[0], v[1])
            eng.gc_mask = sum(1 << q for q in gc)
            st.activate_gc_mask(gc)
            if any(q0 <= q < q0 + qsz for q in gc):
                st.deactivate_gc_mask(gc)
                continue
            assert eng.phase_list(q0, qsz, ph) == 0
            assert st.phase_list(q0, qsz, ph) == 0
            st.deactivate_gc_mask(gc)
            gre, gim = eng.get()
            assert_bitwise(gre, st.re(), "phase_list re")
            assert_bitwise(gim, st.im(), "phase_list im")
        eng.gc_mask = 0
    for (a, b) in ((0, 7), (3, 1), (2, 2)):
        eng.set(v[0], v[1])
        st.set(v[0], v[1])
        eng.swap(a, b)
        st.swap(a, b)
        gre, gim = eng.get()
        assert_bitwise(gre, st.re(), "swap")
    eng.close()


def test_qms_W_bitwise_vs_oracle(capi):
    q = po.Qms(1.7, 4, 9, -1.0, 3.0)
    f = q.fields
    n = f.nqubits
    rng = np.random.default_rng(1)
    v = rng.normal(size=(2, 1 << n))
    eng = capi.Engine(n, tile_bits=0)
    st = f.st.contents
    N = 1 << n
    re = np.ctypeslib.as_array(st.re, shape=(N,))
    im = np.ctypeslib.as_array(st.im, shape=(N,))
    po.lib().so_qms_apply_W.argtypes = [po.C.POINTER(po._Qms)]
    for Eold in (0, 5, 15):
        re[:] = v[0]
        im[:] = v[1]
        eng.set(v[0], v[1])
        f.curr_E_old = Eold
        po.lib().so_qms_apply_W(q.p)
        eng.qms_apply_W(f.bm_acc, f.W_mask_Enew, f.bm_enes[0], f.c_factor, Eold)
        gre, gim = eng.get()
        assert_bitwise(gre, re, f"W re Eold={Eold}")
        assert_bitwise(gim, im, f"W im Eold={Eold}")
    eng.close()


def test_errors_mirror_reference(capi):
    eng = capi.Engine(5)
    assert eng.pauli([0, 1, 2], [1, 2, 3], 0.1) == capi.SUQA_ERR_UNSUPPORTED
    assert eng.pauli([0, 1, 2, 3], [3, 3, 3, 3], 0.1) == capi.SUQA_ERR_UNSUPPORTED
    with pytest.raises(capi.SuqaError):
        eng.x(9)
    assert eng.phase_list(0, 2, [0.0, 1.0, 2.0]) == capi.SUQA_ERR_ARG
    eng.close()


def test_full_size_30_qubits_reversible(capi):
    """30 qubits (16 GiB): H layer, one layer of the seed-7 circuit, its inverse, H
    layer again -> |0...0>; the norm and the probabilities stay exact to 1e-12."""
    n = 30
    log("full size n=30")
    gates, ng, mq, mr = po.circuit(n, 1, 7)
    eng = capi.Engine(n, tile_bits=0)
    for q in range(n):
        eng.h(q)
    eng.run_gates(capi.gates_from(gates[:ng]))
    assert abs(eng.vnorm() - 1.0) < 1e-12
    inv = []
    for g in reversed(gates[:ng]):
        h = po.Gate()
        h.kind, h.target, h.nctrl = g.kind, g.target, g.nctrl
        for j in range(3):
            h.ctrl[j] = g.ctrl[j]
        h.angle = -g.angle if g.kind in (2, 3, 4) else 0.0
        inv.append(h)
    eng.run_gates(capi.gates_from(inv))
    assert abs(eng.prob_filter([3, 17, 29], [1, 0, 1]) - 0.125) < 1e-12
    for q in range(n):
        eng.h(q)
    assert abs(eng.prob_filter(list(range(n)), [0] * n) - 1.0) < 1e-12
    c = eng.measure(29, 0.5)
    assert c == 0
    assert abs(eng.vnorm() - 1.0) < 1e-12
    eng.close()


def test_low_and_high_bit_combinations_bitwise(capi):
    """Every gate family with targets, controls, gc bits and Pauli strings on
    qubits below and above the 16-amplitude line boundary (dense-low mode and
    strided mode), bit-exact against the oracle."""
    n = 12
    rng = np.random.default_rng(21)
    eng = capi.Engine(n, tile_bits=0)
    st = po.State(n)
    v = rng.normal(size=(2, 1 << n))
    v /= np.linalg.norm(v)
    eng.set(v[0], v[1])
    st.set(v[0], v[1])
    qsets = [(0, 5, 9), (1, 2, 3), (3, 4, 11), (7, 0, 1), (10, 11, 2), (6, 8, 9), (2, 9, 0)]

    def both(fn):
        fn(eng)
        fn(st)

    for (a, b, c) in qsets:
        for gc in ([], [c]):
            eng.gc_mask = 0
            eng.activate_gc_mask(gc)
            st.activate_gc_mask(gc)
            both(lambda e: e.x(a))
            both(lambda e: e.y(b))
            both(lambda e: e.h(a))
            both(lambda e: e.u1(b, 0.37))
            both(lambda e: e.u1(a, -1.1, 0))
            both(lambda e: e.cx(a, b))
            both(lambda e: e.cx(b, a, 0))
            both(lambda e: e.cu1(b, a, 0.9))
            both(lambda e: e.swap(a, b))
            both(lambda e: e.pauli([a, b], [1, 2], 0.3))
            both(lambda e: e.pauli([b, a], [2, 3], -0.7))
            both(lambda e: e.pauli([a], [1], 1.3))
            eng.deactivate_gc_mask(gc)
            st.deactivate_gc_mask(gc)
        both(lambda e: e.mcx([a, c], b, [1, 0]))
        both(lambda e: e.mcu1([b, c], a, 0.4, [0, 1]))
        gre, gim = eng.get()
        assert_bitwise(gre, st.re(), f"re after {a},{b},{c}")
        assert_bitwise(gim, st.im(), f"im after {a},{b},{c}")
        for q in (a, b):
            ps = [[q, c], [1, 0]]
            assert abs(eng.prob_filter(*ps) - st.prob_filter(*ps)) < 1e-14
    r = 0.3
    for q in (0, 2, 5, 11):
        assert eng.measure(q, r) == st.measure(q, r)
        gre, gim = eng.get()
        np.testing.assert_allclose(gre, st.re(), atol=1e-14, rtol=0)
    eng.close()


@pytest.mark.parametrize("n", [6, 12, 14, 15])
def test_measure_qbits_and_resets_vs_oracle(capi, n):
    """suqa_hip_measure_qbits / suqa_hip_apply_resets (one launch on states of
    at most 2^14 amplitudes, the per-qubit path above) against the oracle's
    sequential measure_qbit / apply_reset: identical outcomes, states within
    1e-12; resets under an active gc_mask (their X is controlled) included."""
    log(f"measure_qbits / resets n={n}")
    rng = np.random.default_rng(40 + n)
    eng = capi.Engine(n)
    st = po.State(n)
    prep(eng, n)
    prep(st, n)
    for rnd in range(3):
        qs = [int(q) for q in rng.permutation(n)[:min(n, 18)]]
        rds = [float(x) for x in rng.random(len(qs))]
        cs = eng.measure_qbits(qs, rds)
        assert cs == [st.measure(q, r) for q, r in zip(qs, rds)]
        for q in range(n):   # mix the state again so later outcomes are not forced
            eng.h(q)
            st.h(q)
        rqs = [int(q) for q in rng.permutation(n)[:3]]
        rr = [float(x) for x in rng.random(3)]
        ctl = [q for q in range(n) if q not in rqs][:1] if rnd == 2 else []
        eng.activate_gc_mask(ctl)
        st.activate_gc_mask(ctl)
        eng.resets(rqs, rr)
        for q, r in zip(rqs, rr):
            st.reset(q, r)
        eng.deactivate_gc_mask(ctl)
        st.deactivate_gc_mask(ctl)
        for q in range(n):
            eng.h(q)
            st.h(q)
        gre, gim = eng.get()
        np.testing.assert_allclose(gre, st.re(), atol=1e-12, rtol=0)
        np.testing.assert_allclose(gim, st.im(), atol=1e-12, rtol=0)
    assert abs(eng.vnorm() - 1.0) < 1e-12
    eng.close()
