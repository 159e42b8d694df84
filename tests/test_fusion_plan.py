"""The fusion plan of the precision pattern (host only, no GPU): blocks built
in FP64 and applied with numpy reproduce the oracle's FP64 state, every gate
lands in exactly one block, and block qubits are sorted and distinct."""
import numpy as np
import pytest

from oracle import pyoracle as po
from suqa_amd import capi, workloads as wl


def apply_block(psi, n, qubits, M):
    """psi (2^n complex) <- M on the block index sum_j bit(qubits[j]) 2^j."""
    K = len(qubits)
    t = psi.reshape([2] * n)          # axis a <-> qubit n-1-a
    axes = [n - 1 - q for q in reversed(qubits)]   # block index most significant first
    t = np.moveaxis(t, axes, list(range(K)))
    sh = t.shape
    t = (M @ t.reshape(1 << K, -1)).reshape(sh)
    t = np.moveaxis(t, list(range(K)), axes)
    return t.reshape(-1)


@pytest.mark.parametrize("width", [4, 5, 6])
def test_plan_reproduces_oracle(width):
    n, layers = 11, 30
    gates, mq, mr = wl.random_circuit(n, layers, 7)
    allg = [wl.G(wl.H, q, [], 0.0) for q in range(n)] + gates
    arr = capi.gates_from(allg)
    plan = capi.fuse_plan(arr, n, width)
    assert sum(g for _, g, _ in plan) == len(allg)
    assert len(plan) < len(allg) / 2
    for qs, _, M in plan:
        assert qs == sorted(qs) and len(set(qs)) == width and max(qs) < n
        assert np.allclose(M.conj().T @ M, np.eye(1 << width), atol=1e-12)
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1
    for qs, _, M in plan:
        psi = apply_block(psi, n, qs, M)
    st = po.State(n)
    for q in range(n):
        st.h(q)
    og, ng, _, _ = po.circuit(n, layers, 7)
    for g in og[:ng]:
        st.apply_gate(g)
    ref = st.re() + 1j * st.im()
    np.testing.assert_allclose(psi, ref, atol=1e-12, rtol=0)


def test_plan_fuses_more_with_width():
    n = 30
    gates, _, _ = wl.random_circuit(n, 20, 7)
    arr = capi.gates_from(gates)
    counts = [len(capi.fuse_plan(arr, n, w)) for w in (4, 5, 6)]
    assert counts[0] > counts[1] > counts[2]
