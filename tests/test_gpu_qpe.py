"""Trotter + phase estimation on the GPU (FP64 engine, one sweep per
controlled Pauli rotation; FP32 engine, rotations fused into MFMA blocks or,
in the sparse mode, applied as pair updates inside block passes) against
the oracle at a reduced size (10 system + 4 register qubits, 96 terms of the
seed-11 generator, dt 0.05, 2 Trotter steps per power), and FP32 against
FP64 at the full 24 + 6 qubits on one controlled Trotter step.
Tolerances: FP64 histogram 1e-12 and state 1e-11 (the oracle decomposes
each rotation into a CNOT ladder, the GPU applies it in one pass); FP32
histogram 2e-5 absolute per bin."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from suqa_amd import workloads as wl

pytestmark = pytest.mark.gpu

NSYS, NREG, DT, STEPS = 10, 4, 0.05, 2


@pytest.fixture(scope="module")
def capi():
    from suqa_amd import capi
    return capi


@pytest.fixture(scope="module")
def oracle_result():
    terms = po.trotter_terms(NSYS, 96, 11)
    st = po.State(NSYS + NREG)
    po.qpe_trotter(st, NSYS, NREG, terms, DT, STEPS)
    reg = list(range(NSYS, NSYS + NREG))
    hist = [st.prob_filter(reg, [(k >> j) & 1 for j in range(NREG)]) for k in range(1 << NREG)]
    return np.array(hist), st.re().copy(), st.im().copy()


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("graphs", [True, False], ids=["graphs", "calls"])
def test_qpe_fp64_vs_oracle(capi, oracle_result, graphs):
    """The controlled Trotter steps captured once per register qubit and
    replayed, or applied call by call: both within 1e-12 of the oracle."""
    log(f"qpe fp64 10+4 graphs={graphs}")
    hist, ore, oim = oracle_result
    eng = capi.Engine(NSYS + NREG)
    eng.init_state()
    got = wl.qpe_trotter(eng, NSYS, NREG, wl.trotter_terms(NSYS, 96, 11), DT, STEPS, graphs=graphs)
    np.testing.assert_allclose(got, hist, atol=1e-12, rtol=0)
    re, im = eng.get()
    np.testing.assert_allclose(re, ore, atol=1e-11, rtol=0)
    np.testing.assert_allclose(im, oim, atol=1e-11, rtol=0)
    eng.close()


@pytest.mark.parametrize("passes", [0, 1, 2], ids=["blocks", "passes", "sparse"])
@pytest.mark.parametrize("width", [5, 6])
def test_qpe_fp32_fused_vs_oracle(capi, oracle_result, width, passes):
    log(f"qpe fp32 10+4 width {width} passes={passes}")
    hist = oracle_result[0]
    f32 = capi.F32Engine(NSYS + NREG)
    f32.passes = passes
    f32.init_state()
    got = wl.qpe_trotter(f32, NSYS, NREG, wl.trotter_terms(NSYS, 96, 11), DT, STEPS, width=width)
    assert np.max(np.abs(np.array(got) - hist)) < 2e-5
    f32.close()


def test_fp32_rejects_ops_wider_than_width(capi):
    f32 = capi.F32Engine(12)
    ops = capi.pauli_ops([([0, 1, 2, 3, 4], [1, 2, 3, 1, 2], 0.1, 1 << 11)])   # 5 target qubits
    with pytest.raises(capi.SuqaError):
        f32.run_paulis(ops, 4)
    with pytest.raises(capi.SuqaError):   # control on a qubit of the string
        f32.run_paulis(capi.pauli_ops([([0, 1], [1, 3], 0.1, 1 << 1)]), 5)
    with pytest.raises(capi.SuqaError):
        f32.run_paulis(capi.pauli_ops([([0, 0], [1, 1], 0.1, 0)]), 5)      # repeated qubit
    for ctrl in (0, 1):                   # controls on qubits 0, 1: rejected in every mode, nothing launched
        for passes in (0, 1, 2):
            f32.passes = passes
            with pytest.raises(capi.SuqaError):
                f32.run_paulis(capi.pauli_ops([([4, 5], [1, 3], 0.1, 1 << ctrl)]), 5)
    f32.close()


@pytest.mark.parametrize("passes", [0, 1, 2], ids=["blocks", "passes", "sparse"])
def test_controlled_trotter_step_full_size(capi, passes):
    """24 + 6 qubits: H on register qubit 29, one Trotter step of the 96
    terms controlled on it, FP32 fused vs FP64 (fidelity >= 1 - 1e-6), and
    the controlled half untouched where the control is 0."""
    n, nsys = 30, 24
    log(f"controlled trotter step n=30 passes={passes}")
    terms = wl.trotter_terms(nsys, 96, 11)
    ops = capi.pauli_ops(wl.trotter_step_ops(terms, DT, 1 << 29))
    e64 = capi.Engine(n)
    e64.init_state()
    e64.h(29)
    e64.run_paulis(ops)
    f32 = capi.F32Engine(n)
    f32.passes = passes
    f32.init_state()
    f32.run_gates(capi.gates_from([wl.G(wl.H, 29, [], 0.0)]), 5)
    nb = f32.run_paulis(ops, 6)
    log(f"{nb} blocks for 96 rotations")
    F = f32.fidelity(e64)
    log(f"fidelity {F!r}")
    assert F > 1 - 1e-6
    assert abs(e64.prob_filter([29], [0]) - 0.5) < 1e-12
    assert abs(f32.prob_filter([29], [0]) - 0.5) < 1e-6
    f32.close()
    e64.close()


def test_sparse_rotations_with_z_qubits_outside_the_tile(capi):
    """FP32 sparse mode at 16 qubits (12-qubit tiles): Pauli strings with X/Y
    on tile qubits and Z on qubits 12-15, some controlled on qubit 15.  The Z
    qubits a pass keeps outside its tile enter as one parity per tile; the
    result follows the FP64 engine to FP32 rounding, as in the mode that
    brings every qubit of a string into the tile."""
    n = 16
    log("sparse rotations, Z qubits outside the tile, n=16")
    rg = wl.Pcg32(23)
    specs = []
    for _ in range(60):
        qs, types = [], []
        for _ in range(1 + rg.randint(0, 2)):          # 1-2 flipped qubits among 2..11
            q = 2 + rg.randint(0, 10)
            if q not in qs:
                qs.append(q)
                types.append(1 + rg.randint(0, 2))     # X or Y
        ctrl = 1 << 15 if rg.randint(0, 3) == 0 else 0
        for _ in range(1 + rg.randint(0, 3)):          # 1-3 Z qubits among 12..14 (15 when uncontrolled)
            q = 12 + rg.randint(0, 3 if ctrl else 4)
            if q not in qs:
                qs.append(q)
                types.append(3)
        specs.append((qs, types, 2.0 * rg.doub() - 1.0, ctrl))
    ops = capi.pauli_ops(specs)
    e64 = capi.Engine(n)
    e64.init_state()
    for q in range(n):
        e64.h(q)
    e64.run_paulis(ops)
    fids = {}
    for mode in (1, 2):
        f32 = capi.F32Engine(n)
        f32.passes = mode
        f32.init_state()
        f32.run_gates(capi.gates_from([wl.G(wl.H, q, [], 0.0) for q in range(n)]), 5)
        f32.run_paulis(ops, 5)
        fids[mode] = f32.fidelity(e64)
        for qs in ([12, 13], [14, 15], [2, 12]):
            for v in range(4):
                mask = [(v >> j) & 1 for j in range(2)]
                assert abs(f32.prob_filter(qs, mask) - e64.prob_filter(qs, mask)) < 1e-6
        f32.close()
    log(f"fidelity {fids!r}")
    assert fids[1] > 1 - 1e-9 and fids[2] > 1 - 1e-9
    e64.close()
