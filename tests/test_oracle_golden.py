"""Pins the oracle to the reference: every golden vector written by the
reference engine compiled from source (tests/golden/make_golden.py) must be
reproduced bit for bit."""
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.common import (QMS_FIXTURES, assert_bitwise, golden_state, golden_summary, hexarr, load, prep,
                          qms_params, run_case, state_summary)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pcg32_golden():
    d = load("pcg32.json")
    for c in d["cases"]:
        r = po.Pcg(c["seed"])
        assert [r.next() for _ in range(64)] == c["u32"]
        w = po.Pcg(c["seed"])
        assert_bitwise(np.array([w.doub() for _ in range(32)]), hexarr(c["doub"]), "doub")
        assert [w.randint(0, 30) for _ in range(32)] == c["randint_0_30"]
        assert w.s.state == int(c["state_after"])


def test_pcg32_published_vector():
    """pcg32 demo of the PCG reference implementation: srandom(42, 54)."""
    s = po._Pcg()
    po.lib().so_pcg32_seed(po.C.byref(s), 42, 54)
    got = [po.lib().so_pcg32_next(po.C.byref(s)) for _ in range(6)]
    assert got == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


@pytest.fixture(scope="module")
def gates():
    return load("gates_n5.json")


def test_prep_matches(gates):
    st = po.State(gates["n"])
    st.qft(list(range(5)))
    prep(st, gates["n"])
    re, im = golden_state(gates["prep"])
    assert_bitwise(st.re(), re, "prep re")
    assert_bitwise(st.im(), im, "prep im")


def test_gate_cases_bitwise(gates):
    n = gates["n"]
    st = po.State(n)
    st.qft(list(range(5)))   # same one-shot QFT table as the harness
    assert len(gates["cases"]) > 200
    for c in gates["cases"]:
        err, out, outc = run_case(st, c, n)
        tag = f"{c['op']} qs={c['qs']} aux={c['aux']} gc={c['gc']}"
        assert err == c["error"], tag
        assert outc == c["outc"], tag
        assert_bitwise(np.array([out]), np.array([float.fromhex(c["out"])]), tag + " out")
        re, im = golden_state(c["state"])
        assert_bitwise(st.re(), re, tag + " re")
        assert_bitwise(st.im(), im, tag + " im")
        assert st.gc_mask == 0


@pytest.mark.parametrize("fx", QMS_FIXTURES)
def test_qms_trace_bitwise(fx):
    d = load(fx)
    p = qms_params(d)
    q = po.Qms(p["beta"], p["r"], p["seed"], p["ene_min"], p["ene_max"], p["pe_steps"], p["max_reverse"],
               p["record_reverse"])
    f = q.fields
    assert f.t_PE_factor == float.fromhex(d["t_PE_factor"])
    assert f.t_phase_estimation == float.fromhex(d["t_phase_estimation"])
    assert f.c_factor == float.fromhex(d["c_factor"])
    # the driver loop of qms.cu:132-149, one step at a time
    L = po.lib()
    L.so_qms_init_state(q.p)
    s0, acc = 0, 0
    for s in range(p["steps"]):
        take = (s > s0 + p["thermalization"]) and ((s - s0) % p["reset_each"] == 0)
        ret = L.so_qms_metro_step(q.p, 1 if take else 0)
        if ret < 0:
            L.so_qms_init_state(q.p)
            s0 = s + 1
        if ret in (1, 2):
            acc += 1
        assert ret == d["rets"][s], f"step {s}"
        assert q.rng_state() == int(d["rng_after"][s]), f"rng after step {s}"
    assert acc == d["accepted"]
    E, X = q.measures()
    assert_bitwise(np.array(E), hexarr(d["E"]), "E")
    assert_bitwise(np.array(X), hexarr(d["X"]), "X")
    if p["record_reverse"]:
        assert q.reverse_counters() == d["reverse_counters"]
    re, im = q.state()
    gre, gim = golden_state(d["final_state"])
    assert_bitwise(re, gre, "final re")
    assert_bitwise(im, gim, "final im")


def test_qms_fixtures_cover_all_return_codes():
    seen = set()
    for fx in QMS_FIXTURES:
        seen |= set(load(fx)["rets"])
    assert {1, 2, 3, 4, -1} <= seen


@pytest.mark.parametrize("fx", ["circuit_n6.json", "circuit_n10.json"])
def test_circuit_bitwise(fx):
    d = load(fx)
    n, layers, seed = d["n"], d["layers"], d["seed"]
    gates, ng, mq, mr = po.circuit(n, layers, seed)
    assert ng == len(d["gates"])
    for g, want in zip(gates[:ng], d["gates"]):
        assert [g.kind, g.target, g.ctrl[0], g.ctrl[1], g.ctrl[2]] == want[:5]
        assert g.angle == float.fromhex(want[5])
    st = po.State(n)
    for q in range(n):
        st.h(q)
    for g in gates[:ng]:
        st.apply_gate(g)
    re, im = golden_state(d["pre_measure"])
    assert_bitwise(st.re(), re, "pre-measure re")
    assert_bitwise(st.im(), im, "pre-measure im")
    for (q, r, c), gq, gr in zip(d["measure"], mq, mr):
        assert q == gq and float.fromhex(r) == gr
        assert st.measure(q, gr) == c
    re, im = golden_state(d["final"])
    assert_bitwise(st.re(), re, "final re")
    assert_bitwise(st.im(), im, "final im")


def test_circuit_summary_definition():
    """The checksum and sample amplitudes tests compute from a state
    (tests/common.state_summary) equal, bit for bit, what the reference harness
    writes for the same state (the 10-qubit circuit, both stored in full);
    the 24-qubit summary fixture is for the same circuit family and seed."""
    full, summ = load("circuit_n10.json"), load("circuit_n10_summary.json")
    assert (summ["n"], summ["layers"], summ["seed"]) == (full["n"], full["layers"], full["seed"])
    for key in ("pre_measure", "final"):
        sums, samples = state_summary(*golden_state(full[key]))
        gsums, gsamples = golden_summary(summ[key])
        assert sums == gsums, key
        assert samples == gsamples, key
    assert summ["measure"] == full["measure"]
    big = load("circuit_n24_summary.json")
    assert (big["n"], big["layers"], big["seed"]) == (24, 200, 7)
    _, _, mq, mr = po.circuit(24, 200, 7)
    assert [(q, float.fromhex(r)) for (q, r, _) in big["measure"]] == list(zip(mq, mr))


def test_qms_driver_output_file_identical():
    """oracle/qms_oracle writes the reference driver's file byte for byte."""
    exe = os.path.join(ROOT, "oracle", "qms_oracle")
    if not os.path.exists(exe):
        po.build()
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.txt")
        subprocess.run([exe, "1.0", "1000", "1", "4", out, "--seed", "42", "--ene-min", "-1", "--ene-max", "3",
                        "--PE-steps", "10"], check=True, stdout=subprocess.DEVNULL)
        with open(out, "rb") as f, open(os.path.join(ROOT, "tests", "golden", "qms_c1_out.txt"), "rb") as g:
            assert f.read() == g.read()


def test_evo_table():
    """evo.cu:22-103 replayed on the oracle reproduces the reference file."""
    with open(os.path.join(ROOT, "tests", "golden", "evo_n3.txt")) as f:
        want = f.read().splitlines()
    st = po.State(3)
    masks = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]]
    lines = []
    for ii in range(21):
        t = ii * 0.1
        st.init_state()
        st.x(1)
        st.h(1)
        st.cx(1, 0)
        for iii in range(3):
            st.pauli([(iii + 1) % 3, (iii + 2) % 3], [1, 1], -t)
        ps = [st.prob_filter([0, 1, 2], m) for m in masks]
        lines.append(" ".join("%.12g" % v for v in [t] + ps))
    assert lines == want
    assert not math.isnan(st.vnorm())
