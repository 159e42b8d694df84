"""The gate-throughput workload's circuit family at 24 qubits (all 200 layers
of seed 7, 6000 gates, then 8 measurements) against the summary the
reference engine wrote for it (tests/golden/circuit_n24_summary.json:
pairwise checksum and 64 sample amplitudes before and after the
measurements), in one-sweep-per-gate mode and in gate passes:
* before the measurements the state is bit-identical to the reference's
  (every sample amplitude and the pairwise checksum of the downloaded state,
  bit for bit); the device checksum (another summation tree) within 1e-12;
* the 8 outcomes are identical; after them (probabilities summed in another
  order than the reference's sequential loop) the samples and the checksum
  agree within 1e-12.
With the passes' commuting gates packed (suqa_hip_set_gate_order), the same
circuit with another rounding: samples and checksums within 1e-12 before and
after the measurements, the same outcomes, and fewer passes."""
import numpy as np
import pytest

from suqa_amd import capi, workloads as wl
from tests.common import golden_summary, load, state_summary

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tile_bits,packed", [(0, False), (11, False), (11, True)])
def test_circuit_n24_vs_reference_summary(tile_bits, packed):
    d = load("circuit_n24_summary.json")
    n, layers, seed = d["n"], d["layers"], d["seed"]
    print(f"circuit n={n} layers={layers} seed={seed} tile bits {tile_bits} packed {packed}", flush=True)
    gates, mq, mr = wl.random_circuit(n, layers, seed)
    eng = capi.Engine(n)
    try:
        eng.set_tile_bits(tile_bits)
        eng.set_gate_order(packed)
        for q in range(n):
            eng.h(q)
        p0 = eng.pass_stats()["passes"]
        eng.run_gates(capi.gates_from(gates))
        gsums, gsamples = golden_summary(d["pre_measure"])
        dev = eng.checksum()
        np.testing.assert_allclose(dev, gsums, rtol=0, atol=1e-12)
        sums, samples = state_summary(*eng.get())
        npass = eng.pass_stats()["passes"] - p0
        print(f"passes {npass}", flush=True)
        if packed:
            assert npass < 268   # 335 passes in call order at 11 tile bits, 211 packed
            np.testing.assert_allclose(sums, gsums, rtol=0, atol=1e-12)
            for (i, re, im), (gi, gre, gim) in zip(samples, gsamples):
                assert i == gi
                assert abs(re - gre) <= 1e-12 and abs(im - gim) <= 1e-12
        else:
            assert samples == gsamples
            assert sums == gsums
        for (q, r, c), gq, gr in zip(d["measure"], mq, mr):
            assert q == gq and float.fromhex(r) == gr
            assert eng.measure(q, gr) == c
        gsums, gsamples = golden_summary(d["final"])
        np.testing.assert_allclose(eng.checksum(), gsums, rtol=0, atol=1e-12)
        sums, samples = state_summary(*eng.get())
        np.testing.assert_allclose(sums, gsums, rtol=0, atol=1e-12)
        for (i, re, im), (gi, gre, gim) in zip(samples, gsamples):
            assert i == gi
            assert abs(re - gre) <= 1e-12 and abs(im - gim) <= 1e-12
    finally:
        eng.close()
