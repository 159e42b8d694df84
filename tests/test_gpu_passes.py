"""Gate passes (several gates per HBM round trip over LDS-resident tiles) on
the GPU: bit-identical to one sweep per gate and to the reference's golden
vectors, for every gate kind, gc_mask controls, tile sizes from 1 to 12
qubits (tiles smaller than, equal to and larger than the state), controls
inside and outside the tile, the QMS driver and the full 30-qubit state."""
import math
import os
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from suqa_amd import workloads as wl
from tests.common import QMS_FIXTURES, assert_bitwise, golden_state, hexarr, load, qms_params, run_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    from suqa_amd import capi
    return capi


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def engine(capi, n, bits):
    e = capi.Engine(n)
    e.set_tile_bits(bits)
    assert e.tile_bits() == bits
    return e


@pytest.mark.parametrize("bits", [1, 3, 5, 12])
def test_gate_cases_in_passes_vs_golden(capi, bits):
    g = load("gates_n5.json")
    n = g["n"]
    log(f"gate cases n={n} tile bits {bits}")
    eng = engine(capi, n, bits)
    eng.qft(list(range(5)))
    for c in g["cases"]:
        tag = f"{c['op']} qs={c['qs']} aux={c['aux']} gc={c['gc']}"
        err, out, outc = run_case(eng, c, n)
        assert err == c["error"], tag
        assert outc == c["outc"], tag
        want_out = float.fromhex(c["out"])
        assert abs(out - want_out) <= 1e-14 * max(1.0, abs(want_out)), tag
        re, im = golden_state(c["state"])
        gre, gim = eng.get()
        if c["op"] in ("measure", "reset", "vnormalize"):
            np.testing.assert_allclose(gre, re, atol=1e-12, rtol=0, err_msg=tag)
            np.testing.assert_allclose(gim, im, atol=1e-12, rtol=0, err_msg=tag)
        else:
            assert_bitwise(gre, re, tag + " re")
            assert_bitwise(gim, im, tag + " im")
    st = eng.pass_stats()
    if bits >= 3:
        assert st["passes"] > 0
    else:   # tiles under 3 bits cannot hold a register step: every gate is its own sweep
        assert st["passes"] == 0 and st["single_sweeps"] > 0
    eng.close()


@pytest.mark.parametrize("fx,bits", [("circuit_n6.json", 4), ("circuit_n10.json", 6), ("circuit_n10.json", 12)])
def test_circuit_in_passes_vs_golden(capi, fx, bits):
    d = load(fx)
    n = d["n"]
    log(f"circuit golden {fx} tile bits {bits}")
    gates, ng, mq, mr = po.circuit(n, d["layers"], d["seed"])
    eng = engine(capi, n, bits)
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


@pytest.mark.parametrize("bits", [8, 12])
def test_circuit_n18_in_passes_vs_oracle(capi, bits):
    n, layers, seed = 18, 200, 7
    log(f"circuit n={n} tile bits {bits} vs oracle")
    gates, ng, mq, mr = po.circuit(n, layers, seed)
    st = po.State(n)
    eng = engine(capi, n, bits)
    for q in range(n):
        st.h(q)
        eng.h(q)
    for g in gates[:ng]:
        st.apply_gate(g)
    eng.run_gates(capi.gates_from(gates[:ng]))
    gre, gim = eng.get()
    assert_bitwise(gre, st.re(), "re")
    assert_bitwise(gim, st.im(), "im")
    stats = eng.pass_stats()
    log(f"pass stats {stats}")
    assert stats["pass_gates"] / stats["passes"] > 3
    for q, r in zip(mq, mr):
        assert eng.measure(q, r) == st.measure(q, r)
    eng.close()


def random_program(n, count, seed):
    """A random mix of every gate call, with gc_mask controls."""
    rng = np.random.default_rng(seed)
    prog = []
    for _ in range(count):
        k = int(rng.integers(0, 13))
        qs = [int(q) for q in rng.permutation(n)]
        ang = float(rng.uniform(-math.pi, math.pi))
        if k == 0:
            prog.append(("x", qs[0]))
        elif k == 1:
            prog.append(("y", qs[0]))
        elif k == 2:
            prog.append(("h", qs[0]))
        elif k == 3:
            prog.append(("u1", qs[0], ang, int(rng.integers(0, 2))))
        elif k == 4:
            prog.append(("cx", qs[0], qs[1], int(rng.integers(0, 2))))
        elif k == 5:
            nc = int(rng.integers(1, 4))
            prog.append(("mcx", qs[:nc], qs[nc], [int(v) for v in rng.integers(0, 2, nc)]))
        elif k == 6:
            prog.append(("cu1", qs[0], qs[1], ang, int(rng.integers(0, 2))))
        elif k == 7:
            nc = int(rng.integers(1, 3))
            prog.append(("mcu1", qs[:nc], qs[nc], ang, [int(v) for v in rng.integers(0, 2, nc)]))
        elif k == 8:
            prog.append(("swap", qs[0], qs[1]))
        elif k == 9:
            w = int(rng.integers(1, 3))
            prog.append(("pauli", qs[:w], [int(t) for t in rng.integers(1, 4, w)], ang))
        elif k == 10:
            w = int(rng.integers(1, 6))
            prog.append(("pauli_general", qs[:w], [int(t) for t in rng.integers(0, 4, w)], ang))
        elif k == 11:
            w = int(rng.integers(2, 5))
            prog.append(("qft" if rng.integers(0, 2) else "qft_inverse", qs[:w]))
        else:
            prog.append(("gc", qs[:int(rng.integers(1, 3))]))
    return prog


def play(eng, prog):
    gc = []
    for op in prog:
        name = op[0]
        if name == "gc":
            if gc:
                eng.deactivate_gc_mask(gc)
                gc = []
            else:
                gc = list(op[1])
                eng.activate_gc_mask(gc)
            continue
        busy = set(gc)
        qs = [op[1]] if isinstance(op[1], int) else list(op[1])
        if name in ("cx", "cu1", "mcx", "mcu1", "swap"):
            qs = qs + [op[2]]
        if busy & set(qs):
            continue
        if name == "x":
            eng.x(op[1])
        elif name == "y":
            eng.y(op[1])
        elif name == "h":
            eng.h(op[1])
        elif name == "u1":
            eng.u1(op[1], op[2], op[3])
        elif name == "cx":
            eng.cx(op[1], op[2], op[3])
        elif name == "mcx":
            eng.mcx(op[1], op[2], op[3])
        elif name == "cu1":
            eng.cu1(op[1], op[2], op[3], op[4])
        elif name == "mcu1":
            eng.mcu1(op[1], op[2], op[3], op[4])
        elif name == "swap":
            eng.swap(op[1], op[2])
        elif name == "pauli":
            eng.pauli(op[1], op[2], op[3])
        elif name == "pauli_general":
            eng.pauli_general(op[1], op[2], op[3])
        elif name == "qft":
            eng.qft(op[1])
        else:
            eng.qft_inverse(op[1])
    if gc:
        eng.deactivate_gc_mask(gc)


@pytest.mark.parametrize("n,bits,seed", [(13, 4, 1), (13, 7, 2), (13, 12, 3), (9, 12, 4), (16, 10, 5)])
def test_every_gate_kind_in_passes_matches_one_sweep_per_gate(capi, n, bits, seed):
    log(f"random program n={n} tile bits {bits} seed {seed}")
    prog = random_program(n, 400, seed)
    ref = engine(capi, n, 0)
    tiled = engine(capi, n, bits)
    for e in (ref, tiled):
        for q in range(n):
            e.h(q)
    for chunk in range(4):
        part = prog[chunk * 100:(chunk + 1) * 100]
        play(ref, part)
        play(tiled, part)
        # a reduction in the middle flushes the queue
        assert ref.prob_filter([0, n - 1], [1, 0]) == tiled.prob_filter([0, n - 1], [1, 0])
    r0, i0 = ref.get()
    r1, i1 = tiled.get()
    assert_bitwise(r1, r0, "re")
    assert_bitwise(i1, i0, "im")
    assert tiled.pass_stats()["passes"] > 0
    ref.close()
    tiled.close()


def string_program(n, count, seed):
    """Pauli strings of weight 2 to 6 (their flip masks span the pass tiles),
    mixed with controlled gates whose controls and targets fall on qubits the
    strings flip, and gc_mask controls."""
    rng = np.random.default_rng(seed)
    prog = []
    for _ in range(count):
        k = int(rng.integers(0, 10))
        qs = [int(q) for q in rng.permutation(n)]
        ang = float(rng.uniform(-math.pi, math.pi))
        if k < 6:
            w = int(rng.integers(2, 7))
            prog.append(("pauli_general", qs[:w], [int(t) for t in rng.integers(0, 4, w)], ang))
        elif k == 6:
            prog.append(("cx", qs[0], qs[1], int(rng.integers(0, 2))))
        elif k == 7:
            nc = int(rng.integers(1, 3))
            prog.append(("mcu1", qs[:nc], qs[nc], ang, [int(v) for v in rng.integers(0, 2, nc)]))
        elif k == 8:
            prog.append(("h", qs[0]))
        else:
            prog.append(("gc", qs[:int(rng.integers(1, 3))]))
    return prog


def span_engine(capi, n, bits, span):
    os.environ["SUQA_HIP_SPAN_TILES"] = "1" if span else "0"
    try:
        return engine(capi, n, bits)
    finally:
        del os.environ["SUQA_HIP_SPAN_TILES"]


@pytest.mark.parametrize("n,bits,seed", [(16, 8, 21), (16, 11, 22), (18, 10, 23), (14, 6, 24), (20, 12, 25)])
def test_span_tiles_match_one_sweep_per_gate(capi, n, bits, seed):
    """Pass tiles spanned by the gates' flip masks (coordinates that are sums
    of qubits, selections on the qubits they mix given coordinates of their
    own): bit-identical to one sweep per gate and to tiles of whole qubits,
    with fewer passes than those."""
    log(f"span tiles n={n} tile bits {bits} seed {seed}")
    prog = string_program(n, 300, seed)
    ref = engine(capi, n, 0)
    span = span_engine(capi, n, bits, True)
    qubit = span_engine(capi, n, bits, False)
    for e in (ref, span, qubit):
        for q in range(n):
            e.h(q)
    for chunk in range(3):
        part = prog[chunk * 100:(chunk + 1) * 100]
        for e in (ref, span, qubit):
            play(e, part)
        p = ref.prob_filter([1, n - 2], [0, 1])
        assert span.prob_filter([1, n - 2], [0, 1]) == p
        assert qubit.prob_filter([1, n - 2], [0, 1]) == p
    r0, i0 = ref.get()
    for e in (span, qubit):
        r1, i1 = e.get()
        assert_bitwise(r1, r0, "re")
        assert_bitwise(i1, i0, "im")
    ss, sq = span.pass_stats(), qubit.pass_stats()
    log(f"  passes: span {ss['passes']} ({ss['single_sweeps']} single), qubit tiles {sq['passes']} ({sq['single_sweeps']} single)")
    assert ss["passes"] + ss["single_sweeps"] < sq["passes"] + sq["single_sweeps"]
    for e in (ref, span, qubit):
        e.close()


def qpe_10_4(capi, bits, reorder):
    nsys, nreg = 10, 4
    terms = wl.trotter_terms(nsys, 96, 11)
    os.environ["SUQA_HIP_PAULI_ORDER"] = "1" if reorder else "0"
    try:
        e = engine(capi, nsys + nreg, bits)
    finally:
        del os.environ["SUQA_HIP_PAULI_ORDER"]
    e.init_state()
    h = wl.qpe_trotter(e, nsys, nreg, terms, 0.05, 2)
    out = (h, e.get())
    e.close()
    return out


@pytest.mark.parametrize("bits", [8, 12])
def test_qpe_in_passes_matches_one_sweep_per_gate(capi, bits):
    """Controlled Pauli rotations whose control lies outside the tile (8 bits)
    or inside it (12 bits), in list order (SUQA_HIP_PAULI_ORDER=0): the FP64
    QPE histogram and state are identical."""
    log(f"qpe 10+4 tile bits {bits}, list order")
    ref = qpe_10_4(capi, 0, False)
    got = qpe_10_4(capi, bits, False)
    assert got[0] == ref[0]
    assert_bitwise(got[1][0], ref[1][0], "re")
    assert_bitwise(got[1][1], ref[1][1], "im")


@pytest.mark.parametrize("bits", [8, 12])
def test_qpe_in_reordered_passes_within_fp64_rounding(capi, bits):
    """run_paulis' pass packing (the default) moves commuting rotations ahead:
    the product is the same operator, the rounding differs, so the state and
    histogram agree with list order to 1e-12 (not bit for bit)."""
    log(f"qpe 10+4 tile bits {bits}, reordered")
    ref = qpe_10_4(capi, 0, False)
    got = qpe_10_4(capi, bits, True)
    np.testing.assert_allclose(got[0], ref[0], atol=1e-12, rtol=0)
    for k in (0, 1):
        np.testing.assert_allclose(got[1][k], ref[1][k], atol=1e-12, rtol=0)


@pytest.mark.parametrize("fx", QMS_FIXTURES[:2])
def test_plain_qms_in_passes_vs_golden(fx):
    from suqa_amd import qms
    d = load(fx)
    p = qms_params(d)
    log(f"plain QMS {fx} tile bits 12")
    os.environ["SUQA_HIP_TILE_BITS"] = "12"
    try:
        out = qms.run(p["beta"], p["steps"], p["reset_each"], p["r"], p["seed"], p["ene_min"], p["ene_max"],
                      p["pe_steps"], p["thermalization"], p["max_reverse"], p["record_reverse"], final_state=True)
    finally:
        del os.environ["SUQA_HIP_TILE_BITS"]
    assert list(out["rets"]) == d["rets"]
    assert out["accepted"] == d["accepted"]
    assert_bitwise(out["E"], hexarr(d["E"]), "E")
    assert_bitwise(out["X"], hexarr(d["X"]), "X")


def test_full_size_passes_match_one_sweep_per_gate(capi):
    """30 qubits, the H layer and 3 layers of the seed-7 circuit: identical
    probabilities (every reduction of identical states is identical)."""
    n = 30
    log("full size n=30 passes vs one sweep per gate")
    gates, _, _ = wl.random_circuit(n, 3, 7)
    arr = capi.gates_from([wl.G(wl.H, q, [], 0.0) for q in range(n)] + gates)
    vals = []
    for b in (0, 12):
        e = engine(capi, n, b)
        e.init_state()
        e.run_gates(arr)
        vals.append([e.prob_filter([q], [1]) for q in (0, 3, 4, 11, 12, 29)] + [e.vnorm()])
        if b:
            log(f"pass stats {e.pass_stats()}")
        e.close()
    assert vals[0] == vals[1]


@pytest.mark.parametrize("n,bits", [(7, 5), (6, 5), (8, 6), (12, 11), (5, 5), (9, 11)])
def test_every_gate_at_every_qubit_in_a_register_step(capi, n, bits):
    """Each gate kind on each qubit, followed by a gate that selects on another
    qubit (so both share a register step and the first one lands on every
    local position), bit-identical to one sweep per gate.  Covers the
    compile-time register layouts of every partner mask (xl = 1, 2, 4), and
    (n = 5, 9: the state is one tile of at most 2^9 amplitudes) the passes
    that give every gate an LDS step of its own."""
    log(f"gate x qubit sweep n={n} tile bits {bits}")
    rng = np.random.default_rng(n)
    v = rng.normal(size=(2, 1 << n))
    v /= np.linalg.norm(v)
    ops = ("x", "y", "h", "pauliX", "pauliY", "pauliZ", "u1", "cx", "swap", "YY", "XZ")
    bad = []
    for op in ops:
        for q in range(n):
            q2 = (q + 1) % n
            out = []
            for b in (0, bits):
                e = engine(capi, n, b)
                e.set(v[0], v[1])
                if op == "x":
                    e.x(q)
                elif op == "y":
                    e.y(q)
                elif op == "h":
                    e.h(q)
                elif op == "pauliX":
                    e.pauli([q], [1], 0.37)
                elif op == "pauliY":
                    e.pauli([q], [2], 0.37)
                elif op == "pauliZ":
                    e.pauli([q], [3], 0.37)
                elif op == "u1":
                    e.u1(q, 0.37)
                elif op == "cx":
                    e.cx(q2, q)
                elif op == "swap":
                    e.swap(q, q2)
                elif op == "YY":
                    e.pauli([q, q2], [2, 2], 0.37)
                else:
                    e.pauli([q, q2], [1, 3], 0.37)
                e.u1(q2, 0.11)
                e.cx(q, q2)
                out.append(e.get())
                e.close()
            if not (np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])):
                bad.append((op, q))
    assert not bad, f"passes differ from one sweep per gate for {bad}"
