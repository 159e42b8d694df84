"""Captured launch sequences (suqa_hip_graph_*): a recorded gate sequence,
replayed, gives bit for bit what the gate calls give when run, on every
replay, with eager work, read-backs and synchronizations between replays
(gate passes on and off, a small state whose passes are one-workgroup
launches and a 20-qubit one); calls that read a value back fail while a
capture is open and leave the engine usable."""
import numpy as np
import pytest

from suqa_amd import capi, workloads as wl
from tests.common import assert_bitwise

pytestmark = pytest.mark.gpu


def apply_seq(eng, gates, n):
    eng.run_gates(capi.gates_from(gates))
    eng.pauli([0, n - 1], [1, 2], 0.37)
    eng.u1(n // 2, 1.1)
    eng.cx(1, n - 2)


@pytest.mark.parametrize("n,tile_bits", [(8, 11), (8, 0), (20, 11), (20, 0)])
def test_replays_match_eager(n, tile_bits):
    print(f"graph replay n={n} tile bits {tile_bits}", flush=True)
    gates, _, _ = wl.random_circuit(n, 4, 100 + n)
    other, _, _ = wl.random_circuit(n, 1, 200 + n)
    a, b = capi.Engine(n), capi.Engine(n)
    try:
        for e in (a, b):
            e.set_tile_bits(tile_bits)
            for q in range(n):
                e.h(q)
        a.graph_begin()
        apply_seq(a, gates, n)
        g = a.graph_end()
        for rep in range(4):
            a.graph_launch(g)
            apply_seq(b, gates, n)
            ra, ia = a.get()
            rb, ib = b.get()
            assert_bitwise(ra, rb, f"replay {rep} re")
            assert_bitwise(ia, ib, f"replay {rep} im")
            # eager work between replays, the same on both
            for e in (a, b):
                e.run_gates(capi.gates_from(other))
                e.synchronize()
            assert a.vnorm() == b.vnorm()
    finally:
        a.close()
        b.close()


def test_read_back_fails_during_capture():
    e = capi.Engine(8)
    try:
        e.h(0)
        e.graph_begin()
        e.h(1)
        with pytest.raises(capi.SuqaError):
            e.vnorm()
        with pytest.raises(capi.SuqaError):
            e.measure(0, 0.5)
        g = e.graph_end()
        e.graph_launch(g)
        re, im = e.get()
        np.testing.assert_allclose(re[:4], [0.5, 0.5, 0.5, 0.5], atol=1e-15)
        with pytest.raises(capi.SuqaError):
            e.graph_launch(g + 1)
    finally:
        e.close()
