"""Small states run their queued gates, measurements, resets and apply_W as
one-workgroup programs.  Checked through the C-ABI: measurements in a program
give the outcomes and the state, bit for bit, of the one-measurement kernel
(tile passes off); deferred measurements (outcomes left on the device) match
the ones read back; the conditional X and the apply_W that read outcome slots
on the device match their host-driven forms, also when the measurement ran in
an earlier program of the same queue; states of 11-14 qubits run them as
launches of their own; larger states refuse them."""
import numpy as np
import pytest

from tests.common import assert_bitwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    from suqa_amd import capi
    return capi


def log(m):
    print(m, flush=True)


def prepared(capi, n, seed, tile_bits=None):
    """Every amplitude nonzero: an H layer, then X rotations, phases and CX."""
    eng = capi.Engine(n, tile_bits=tile_bits)
    rng = np.random.default_rng(seed)
    for q in range(n):
        eng.h(q)
    for _ in range(3 * n):
        eng.pauli([int(rng.integers(n))], [1], float(rng.uniform(-3, 3)))
        eng.u1(int(rng.integers(n)), float(rng.uniform(-3, 3)))
        a, b = rng.choice(n, 2, replace=False)
        eng.cx(int(a), int(b))
    return eng


def same_state(a, b, tag):
    ra, ia = a.get()
    rb, ib = b.get()
    assert_bitwise(ra, rb, tag + " re")
    assert_bitwise(ia, ib, tag + " im")


@pytest.mark.parametrize("n", [3, 5, 8, 9, 10])
def test_program_measurements_match_the_measurement_kernel(capi, n):
    log(f"program measurements n={n}")
    rng = np.random.default_rng(100 + n)
    qs = [int(q) for q in rng.permutation(n)[:min(n, 4)]]
    for rep in range(3):
        rd = [float(x) for x in rng.uniform(0, 1, len(qs))]
        prog = prepared(capi, n, rep)          # one pass tile: measurements join the queue
        plain = prepared(capi, n, rep, 0)      # tile passes off: k_measure_small
        defer = prepared(capi, n, rep)
        c_prog = prog.measure_qbits(qs, rd)
        c_plain = plain.measure_qbits(qs, rd)
        slot = defer.measure_qbits_deferred(qs, rd)
        c_defer = defer.read_outcomes(slot, len(qs))
        assert c_prog == c_plain == c_defer
        same_state(prog, plain, "program vs kernel")
        same_state(prog, defer, "program vs deferred")
        # resets (X over the gc bits when the outcome is 1) likewise
        prog.resets(qs, rd[::-1])
        plain.resets(qs, rd[::-1])
        same_state(prog, plain, "resets")
        for e in (prog, plain, defer):
            e.close()


@pytest.mark.parametrize("n", [6, 8])
def test_conditional_x_matches_host_flips(capi, n):
    log(f"x_if n={n}")
    rng = np.random.default_rng(7 + n)
    qs = [1, 3, n - 1]
    rd = [float(x) for x in rng.uniform(0, 1, len(qs))]
    host = prepared(capi, n, 3)
    dev = prepared(capi, n, 3)
    cs = host.measure_qbits(qs, rd)
    for q, c in zip(qs, cs):
        if c:
            host.x(q)
    host.h(0)
    slot = dev.measure_qbits_deferred(qs, rd)
    dev.x_if(qs, slot)
    dev.h(0)
    same_state(host, dev, "x_if")
    # the measurement in one program, the flips in a later one (more queued
    # operations than a program holds between them: the flips read the slots)
    rd2 = [float(x) for x in rng.uniform(0, 1, len(qs))]
    cs = host.measure_qbits(qs, rd2)
    slot = dev.measure_qbits_deferred(qs, rd2)
    for k in range(150):
        host.u1(k % n, 0.01 * k)
        dev.u1(k % n, 0.01 * k)
    for q, c in zip(qs, cs):
        if c:
            host.x(q)
    dev.x_if(qs, slot)
    same_state(host, dev, "x_if across programs")
    assert dev.read_outcomes(slot, len(qs)) == cs
    host.close()
    dev.close()


def test_deferred_apply_W_matches_host_E_old(capi):
    r = 4
    n = 3 + r + 1
    wmask = ((1 << r) - 1) << 3
    log("deferred apply_W, r=4")
    for rep, beta in enumerate((0.3, 1.0, 4.0)):
        cf = 0.25 * beta
        rng = np.random.default_rng(rep)
        rd = [float(x) for x in rng.uniform(0, 1, r)]
        host = prepared(capi, n, 10 + rep)
        plain = prepared(capi, n, 10 + rep, 0)
        dev = prepared(capi, n, 10 + rep)
        ene = list(range(3, 3 + r))
        cs = host.measure_qbits(ene, rd)
        assert plain.measure_qbits(ene, rd) == cs
        e_old = sum(c << k for k, c in enumerate(cs))
        slot = dev.measure_qbits_deferred(ene, rd)
        for e in (host, plain, dev):
            e.h(0)
            e.h(5)
        host.qms_apply_W(n - 1, wmask, 3, cf, e_old)
        plain.qms_apply_W(n - 1, wmask, 3, cf, e_old)
        dev.qms_apply_W_deferred(n - 1, wmask, 3, cf, slot)
        same_state(host, plain, "W program vs kernel")
        same_state(host, dev, "W deferred")
        for e in (host, plain, dev):
            e.close()


@pytest.mark.parametrize("lead", [88, 91, 92, 93, 95])
def test_measurements_at_program_boundaries(capi, lead):
    """A register's deferred measurements queued where a 96-operation program
    ends (after `lead` gates), then apply_W reading them: the same state and
    outcomes as the one-operation kernels."""
    r, n = 4, 8
    wmask = ((1 << r) - 1) << 3
    log(f"program boundary, {lead} gates first")
    rng = np.random.default_rng(lead)
    rd = [float(x) for x in rng.uniform(0, 1, r)]
    dev = prepared(capi, n, 30)
    plain = prepared(capi, n, 30, 0)
    dev.synchronize()
    for k in range(lead):
        dev.u1(k % n, 0.01 * (k + 1))
        plain.u1(k % n, 0.01 * (k + 1))
    ene = list(range(3, 3 + r))
    slot = dev.measure_qbits_deferred(ene, rd)
    cs = plain.measure_qbits(ene, rd)
    dev.x_if(ene, slot)
    for q, c in zip(ene, cs):
        if c:
            plain.x(q)
    dev.qms_apply_W_deferred(n - 1, wmask, 3, 0.7, slot)
    plain.qms_apply_W(n - 1, wmask, 3, 0.7, sum(c << k for k, c in enumerate(cs)))
    same_state(dev, plain, "boundary")
    assert dev.read_outcomes(slot, r) == cs
    dev.close()
    plain.close()


@pytest.mark.parametrize("n", [12, 13])
def test_deferred_on_states_beyond_one_program(capi, n):
    """11-14 qubits: the measurements launch at once with their outcomes in the
    ring, the conditional X is its own launch reading them."""
    log(f"deferred launches n={n}")
    rng = np.random.default_rng(40 + n)
    qs = [2, 7, n - 1]
    rd = [float(x) for x in rng.uniform(0, 1, len(qs))]
    host = prepared(capi, n, 5)
    dev = prepared(capi, n, 5)
    cs = host.measure_qbits(qs, rd)
    slot = dev.measure_qbits_deferred(qs, rd)
    for q, c in zip(qs, cs):
        if c:
            host.x(q)
    dev.x_if(qs, slot)
    host.h(3)
    dev.h(3)
    same_state(host, dev, "deferred launches")
    assert dev.read_outcomes(slot, len(qs)) == cs
    host.close()
    dev.close()


def test_larger_states_refuse_deferred_operations(capi):
    log("deferred on 15 qubits")
    eng = capi.Engine(15)
    with pytest.raises(capi.SuqaError) as ei:
        eng.measure_qbits_deferred([0, 1], [0.5, 0.5])
    assert ei.value.code == -2
    with pytest.raises(capi.SuqaError):
        eng.x_if([0], 0)
    eng.close()


@pytest.mark.parametrize("n", [8, 12])
def test_outcome_ring_wraps(capi, n):
    """More deferred measurements than the ring has slots: a group of slots never
    straddles the ring's end, the conditional X reads each group after the wrap
    as before it, and the last groups still read back their outcomes."""
    log(f"outcome ring wrap n={n}")
    rng = np.random.default_rng(70 + n)
    host = prepared(capi, n, 9)
    dev = prepared(capi, n, 9)
    groups = []
    for k in range(450):   # 1350 slots: the ring (1024) wraps once
        qs = [k % n, (k + 3) % n, (k + 5) % n]
        rd = [float(x) for x in rng.uniform(0, 1, 3)]
        cs = host.measure_qbits(qs, rd)
        slot = dev.measure_qbits_deferred(qs, rd)
        assert 0 <= slot and slot + 3 <= capi.OUTCOME_SLOTS
        for q, c in zip(qs, cs):
            if c:
                host.x(q)
        dev.x_if(qs, slot)
        for e in (host, dev):
            for q in qs:
                e.h(q)
            e.u1(k % n, 0.1 + 0.01 * k)
        groups.append((slot, cs))
    same_state(host, dev, "after the wrap")
    starts = [s for s, _ in groups]
    assert min(starts[1:]) == 0 and starts.index(0, 1) > 300   # wrapped once, after ~341 groups
    for slot, cs in groups[-300:]:   # not yet overwritten
        assert dev.read_outcomes(slot, 3) == cs
    host.close()
    dev.close()


@pytest.mark.parametrize("n", [1, 2, 4, 8])
def test_tiny_states_and_calls_longer_than_a_program(capi, n):
    """States of 1-8 qubits, and calls of more measurements than one program
    holds (split across launches): outcomes and amplitudes as the
    one-measurement kernel's."""
    log(f"tiny state n={n}, 150-measurement calls")
    rng = np.random.default_rng(90 + n)
    prog, plain = capi.Engine(n), capi.Engine(n, tile_bits=0)
    for rep in range(3):
        for e in (prog, plain):
            for q in range(n):
                e.h(q)
                e.pauli([q], [1], 0.3 + 0.2 * q + rep)
                e.u1(q, 0.7 - 0.1 * q)
        qs = [int(q) for q in rng.integers(0, n, 150)]
        rd = [float(x) for x in rng.uniform(0, 1, 150)]
        assert prog.measure_qbits(qs, rd) == plain.measure_qbits(qs, rd)
        same_state(prog, plain, "measure_qbits")
        for e in (prog, plain):
            for q in range(n):
                e.h(q)
        prog.resets(qs, rd[::-1])
        plain.resets(qs, rd[::-1])
        same_state(prog, plain, "resets")
    prog.close()
    plain.close()


@pytest.mark.parametrize("n,gc", [(8, 0), (12, 0), (13, 0), (12, 1 << 11)])
def test_resets_then_x_if_match_the_calls_it_stands_for(capi, n, gc):
    """One call for a register's resets interleaved with the flips of another
    register's measured qubits: the state of the separate calls, bit for bit."""
    log(f"resets_then_x_if n={n} gc={gc}")
    rng = np.random.default_rng(60 + n)
    r = 2 if n < 12 else 4
    old = [3 + r + k for k in range(r)]   # a two-register layout: E_new at 3.., E_old after it
    new = [3 + k for k in range(r)]
    for rep in range(3):
        loop = prepared(capi, n, 20 + rep)
        merged = prepared(capi, n, 20 + rep)
        rd_m = [float(x) for x in rng.uniform(0, 1, r)]
        rd_r = [float(x) for x in rng.uniform(0, 1, r)]
        s_loop = loop.measure_qbits_deferred(new, rd_m)
        s_merged = merged.measure_qbits_deferred(new, rd_m)
        loop.gc_mask = gc
        merged.gc_mask = gc
        for k in range(r):
            loop.reset(old[k], rd_r[k])
            loop.x_if([new[k]], s_loop + k)
        merged.resets_then_x_if(old, rd_r, new, s_merged)
        loop.gc_mask = 0
        merged.gc_mask = 0
        loop.h(0)
        merged.h(0)
        same_state(loop, merged, "resets then x_if")
        loop.close()
        merged.close()
