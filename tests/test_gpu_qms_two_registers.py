"""The two-register QMS layout (E_old in its own register, 3 + 2r + 1 qubits;
the configuration the c1/c5 workloads describe) on the GPU: the C++ qms::
mirror over the engine, with suqa_hip_qms_apply_W2 reading E_old from the
register, reproduces the oracle's two-register chain bit for bit (return
codes, final pcg32 state, E/X measurements, reverse counters; final state
within 1e-12), with gate passes on and off.  The oracle's two-register mode
is pinned against the reference-pinned one-register path in
tests/test_qms_two_registers.py (identical return codes to the one-register
chain that makes the reference's "comparison with 2enereg" draws)."""
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



def infidelity(a, b):
    """1 - |<a|b>|^2 / (<a|a> <b|b>)."""
    return 1.0 - abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


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


# beta, r, seed: one launch of the batched kernel with uneven widths, r = 2..6
# (the reference layout's 8 to 16 qubits per chain)
BATCH = [(1.0, 4, 42), (2.0, 3, 123), (0.5, 2, 7), (3.0, 6, 99), (0.2, 5, 5), (1.3, 6, 31), (0.8, 5, 77)]


def check_chain_vs_oracle(out, c, beta, r, seed, steps, reset_each, therm, max_rev, emin, emax):
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
    return rets


@pytest.mark.parametrize("max_rev", [100, 2])
def test_two_register_batch_vs_oracle(qms, max_rev):
    """The batched kernel with two_registers: every chain (r = 2..6, 8 to 16
    qubits in the full layout, stored in 2^(r+4) amplitudes) reproduces the
    oracle's full-state two-register chain bit for bit (return codes, final
    pcg32 state, acceptances, E/X).  Seven chains run a workgroup per chain
    with one amplitude per thread; max_rev 2 makes restorations fail (-1)."""
    steps, reset_each, therm, emin, emax = 60, 1, 5, -1.0, 3.0
    print(f"two-register batch {BATCH} max_rev {max_rev}", flush=True)
    out = qms.batch_run(BATCH, steps, reset_each, max_rev, emin, emax, 10, therm, rng_after=True, two_registers=True)
    fails = 0
    for c, (beta, r, seed) in enumerate(BATCH):
        rets = check_chain_vs_oracle(out, c, beta, r, seed, steps, reset_each, therm, max_rev, emin, emax)
        fails += rets.count(-1)
    if max_rev == 2:
        assert fails > 0


def test_two_register_batch_lds_layout_vs_oracle(qms):
    """300 two-register chains with r = 2..6 run in LDS (1 to 4 wavefronts
    per chain); every 25th chain is checked against the oracle's full-state
    two-register chain, bit for bit, over an inexact energy range ([0, 1])."""
    nch, steps, reset_each, therm, max_rev, emin, emax = 300, 50, 2, 5, 30, 0.0, 1.0
    chains = [(float(0.1 + 4.9 * k / (nch - 1)), 2 + k % 5, 500 + k) for k in range(nch)]
    print("two-register batch, 300 chains", flush=True)
    out = qms.batch_run(chains, steps, reset_each, max_rev, emin, emax, 10, therm, rng_after=True, two_registers=True)
    for c in range(0, nch, 25):
        beta, r, seed = chains[c]
        check_chain_vs_oracle(out, c, beta, r, seed, steps, reset_each, therm, max_rev, emin, emax)


def test_two_register_batch_fp32_follows_fp64(qms):
    """FP32 two-register chains (r = 2..6) follow the FP64 ones (bit-exact with
    the oracle above): decisions differ only when a draw falls within FP32
    rounding of a probability, so nearly all 48 chains x 120 steps agree."""
    chains = [(float(0.1 + 4.9 * k / 47), 2 + k % 5, 2000 + k) for k in range(48)]
    steps, emin, emax = 120, -1.0, 3.0
    print("two-register batch fp32", flush=True)
    o64 = qms.batch_run(chains, steps, 1, 100, emin, emax, 10, 20, two_registers=True, final_state=True)
    o32 = qms.batch_run(chains, steps, 1, 100, emin, emax, 10, 20, two_registers=True, fp32=True, final_state=True)
    same = [c for c in range(len(chains)) if np.array_equal(o64["rets"][c], o32["rets"][c])]
    print(f"{len(same)} of {len(chains)} FP32 chains identical to FP64", flush=True)
    assert len(same) >= 44
    # the state after the last step: infidelity against the FP64 chain within
    # the precision pattern's bound (1e-5) wherever the decisions agree
    worst = max(infidelity(o64["final"][c], o32["final"][c]) for c in same)
    print(f"largest FP32 state infidelity {worst:.3g}", flush=True)
    assert worst <= 1e-5
    from suqa_amd import capi
    with pytest.raises(capi.SuqaError):   # register widths are 2..6
        qms.batch_run([(1.0, 7, 3)], 10, 1, 100, emin, emax, 10, 2, two_registers=True)
