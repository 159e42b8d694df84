"""The two-register QMS layout (E_old in its own register, 3 + 2r + 1 qubits;
the configuration the c1/c5 workloads describe) on the GPU: the C++ qms::
mirror over the engine, with suqa_hip_qms_apply_W2 reading E_old from the
register, reproduces the oracle's two-register chain bit for bit (return
codes, final pcg32 state, E/X measurements, reverse counters; final state
within 1e-12), with gate passes on and off.  The oracle's two-register mode
is pinned against the reference-pinned one-register path in
tests/test_qms_two_registers.py."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.common import assert_bitwise

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# beta, r, seed, steps, reset_each, thermalization, max_reverse, E range
CASES = [
    (1.0, 4, 42, 300, 1, 50, 100, (-1.0, 3.0)),    # c1 as written: 12 qubits
    (2.0, 3, 123, 300, 1, 10, 1, (-1.0, 3.0)),     # max_reverse 1: failed restorations (-1)
    (0.5, 5, 7, 120, 2, 20, 100, (0.0, 1.0)),
]


@pytest.fixture(scope="module")
def qms():
    from suqa_amd import qms
    return qms


@pytest.mark.parametrize("tile_bits", ["0", "11"])
@pytest.mark.parametrize("case", CASES, ids=[f"b{c[0]}-r{c[1]}-s{c[2]}" for c in CASES])
def test_two_register_chain_vs_oracle(qms, case, tile_bits, monkeypatch):
    beta, r, seed, steps, reset_each, therm, max_rev, (emin, emax) = case
    print(f"two-register QMS {case} tile bits {tile_bits}", flush=True)
    monkeypatch.setenv("SUQA_HIP_TILE_BITS", tile_bits)
    out = qms.run(beta, steps, reset_each, r, seed, emin, emax, 10, therm, max_rev, True, final_state=True,
                  two_registers=True)
    o = po.Qms(beta, r, seed, emin, emax, 10, max_rev, True, two_registers=True)
    acc, rets = o.run(steps, reset_each, therm)
    assert list(out["rets"]) == rets
    assert out["accepted"] == acc
    assert int(out["rng_after"][-1]) == o.rng_state()
    E, X = o.measures()
    assert_bitwise(out["E"], np.array(E), "E")
    assert_bitwise(out["X"], np.array(X), "X")
    assert list(out["reverse"]) == o.reverse_counters()
    re, im = o.state()
    assert out["final"][0].shape == re.shape == (1 << (2 * r + 4),)
    np.testing.assert_allclose(out["final"][0], re, atol=1e-12, rtol=0)
    np.testing.assert_allclose(out["final"][1], im, atol=1e-12, rtol=0)
    if max_rev == 1:
        assert -1 in rets


def test_driver_option(qms):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.txt")
        r = subprocess.run([qms.DRIVER, "1.0", "100", "1", "3", out, "--seed", "9", "--ene-min", "-1", "--ene-max",
                            "3", "--thermalization", "10", "--ene-registers", "2"], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
        assert r.stdout.splitlines()[0] == "10"    # qubits: 3 + 2 * 3 + 1
        bad = subprocess.run([qms.DRIVER, "1.0", "10", "1", "3", out, "--ene-registers", "3"], capture_output=True,
                             text=True, timeout=120)
        assert bad.returncode != 0 and "ene registers" in (bad.stdout + bad.stderr)


# beta, r, seed: one launch of the batched kernel with uneven widths (the
# widest, r = 4, is a 12-qubit chain: 64 KiB of LDS in FP64)
BATCH = [(1.0, 4, 42), (2.0, 3, 123), (0.5, 2, 7), (3.0, 4, 99), (0.2, 3, 5)]


@pytest.mark.parametrize("chains", [BATCH, [(1.5, 3, 11), (0.7, 2, 13), (2.5, 3, 17)]], ids=["r2-4", "r2-3"])
def test_two_register_batch_vs_oracle(qms, chains):
    """The batched kernel with two_registers: every chain reproduces the
    oracle's two-register chain bit for bit (return codes, final pcg32 state,
    acceptances, E/X).  The 12-qubit chains (r = 4) run a wavefront per chain
    with the state in LDS, the others a workgroup per chain with one
    amplitude per thread."""
    steps, reset_each, therm, max_rev, emin, emax = 160, 1, 20, 100, -1.0, 3.0
    print(f"two-register batch {chains}", flush=True)
    out = qms.batch_run(chains, steps, reset_each, max_rev, emin, emax, 10, therm, rng_after=True, two_registers=True)
    for c, (beta, r, seed) in enumerate(chains):
        o = po.Qms(beta, r, seed, emin, emax, 10, max_rev, two_registers=True)
        acc, rets = o.run(steps, reset_each, therm)
        assert list(out["rets"][c]) == rets, f"chain {c} (beta={beta}, r={r}, seed={seed})"
        assert int(out["accepted"][c]) == acc
        assert int(out["rng_after"][c][-1]) == o.rng_state()
        E, X = o.measures()
        nm = int(out["n_meas"][c])
        assert nm == len(E)
        assert_bitwise(out["E"][c][:nm], np.array(E), f"E chain {c}")
        assert_bitwise(out["X"][c][:nm], np.array(X), f"X chain {c}")


def test_two_register_batch_fp32_and_limits(qms):
    """FP32 two-register chains follow the FP64 ones (decisions differ only when
    a draw falls within FP32 rounding of a probability); a chain whose state
    does not fit a workgroup's LDS (r = 6 in FP64: 16 qubits, 1 MiB) is
    rejected before anything is launched."""
    steps, emin, emax = 120, -1.0, 3.0
    print("two-register batch fp32", flush=True)
    o64 = qms.batch_run(BATCH, steps, 1, 100, emin, emax, 10, 20, two_registers=True)
    o32 = qms.batch_run(BATCH, steps, 1, 100, emin, emax, 10, 20, two_registers=True, fp32=True)
    same = sum(1 for c in range(len(BATCH)) if np.array_equal(o64["rets"][c], o32["rets"][c]))
    assert same >= len(BATCH) - 1
    from suqa_amd import capi
    with pytest.raises(capi.SuqaError):
        qms.batch_run([(1.0, 6, 3)], 10, 1, 100, emin, emax, 10, 2, two_registers=True)
