"""QMS parity on the GPU: the plain FP64 path (suqa:: / qms:: C++ mirror over
the engine) and the batched-chain kernel reproduce the reference's per-step
accept/reject codes, pcg32 states and E/X measurements bit for bit."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.common import QMS_FIXTURES, assert_bitwise, golden_state, hexarr, load, qms_params

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def qms():
    from suqa_amd import qms
    return qms


def log(m):
    print(m, flush=True)


@pytest.mark.parametrize("mode", ["graphs", "deferred", "deferred-values-per-step", "synchronous"])
@pytest.mark.parametrize("fx", QMS_FIXTURES)
def test_plain_qms_vs_golden(qms, fx, mode, monkeypatch):
    """With the step's fixed launch sequences captured once and replayed; with
    deferred readouts (the default), their E/X values left for the next wait
    (the library driver's default) or resolved before every step returns (the
    default for other callers of qms::metro_step); with every readout waited for."""
    monkeypatch.setenv("SUQA_QMS_GRAPHS", "1" if mode == "graphs" else "0")
    monkeypatch.setenv("SUQA_QMS_DEFERRED", "0" if mode == "synchronous" else "1")
    monkeypatch.setenv("SUQA_QMS_DEFER_VALUES", "0" if mode == "deferred-values-per-step" else "1")
    d = load(fx)
    p = qms_params(d)
    log(f"plain QMS {fx} {mode}")
    out = qms.run(p["beta"], p["steps"], p["reset_each"], p["r"], p["seed"], p["ene_min"], p["ene_max"],
                  p["pe_steps"], p["thermalization"], p["max_reverse"], p["record_reverse"], final_state=True)
    assert list(out["rets"]) == d["rets"]
    assert [int(x) for x in out["rng_after"]] == [int(x) for x in d["rng_after"]]
    assert out["accepted"] == d["accepted"]
    assert_bitwise(out["E"], hexarr(d["E"]), "E")
    assert_bitwise(out["X"], hexarr(d["X"]), "X")
    if p["record_reverse"]:
        assert list(out["reverse"]) == d["reverse_counters"]
    re, im = golden_state(d["final_state"])
    np.testing.assert_allclose(out["final"][0], re, atol=1e-12, rtol=0)
    np.testing.assert_allclose(out["final"][1], im, atol=1e-12, rtol=0)



def infidelity(a, b):
    """1 - |<a|b>|^2 / (<a|a> <b|b>)."""
    return 1.0 - abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


def test_driver_output_file_matches_reference(qms):
    log("suqa_qms driver file")
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.txt")
        r = subprocess.run([qms.DRIVER, "1.0", "1000", "1", "4", out, "--seed", "42", "--ene-min", "-1", "--ene-max",
                            "3", "--PE-steps", "10"], capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr
        assert "acceptance" in r.stdout
        with open(out, "rb") as f, open(os.path.join(ROOT, "tests", "golden", "qms_c1_out.txt"), "rb") as g:
            assert f.read() == g.read()


@pytest.mark.parametrize("fx", QMS_FIXTURES)
def test_batch_single_chain_vs_golden(qms, fx):
    d = load(fx)
    p = qms_params(d)
    log(f"batch QMS {fx}")
    out = qms.batch_run([(p["beta"], p["r"], p["seed"])], p["steps"], p["reset_each"], p["max_reverse"], p["ene_min"],
                        p["ene_max"], p["pe_steps"], p["thermalization"], rng_after=True)
    assert list(out["rets"][0]) == d["rets"]
    assert [int(x) for x in out["rng_after"][0]] == [int(x) for x in d["rng_after"]]
    assert int(out["accepted"][0]) == d["accepted"]
    nm = int(out["n_meas"][0])
    assert nm == len(d["E"])
    assert_bitwise(out["E"][0][:nm], hexarr(d["E"]), "E")
    assert_bitwise(out["X"][0][:nm], hexarr(d["X"]), "X")


@pytest.mark.parametrize("nch,widths", [(48, (2, 3, 4, 5, 6)), (300, (2, 3, 4, 5, 6)), (300, (2, 3)), (300, (5,))],
                         ids=["48-chains", "300-chains", "300-chains-r2-3", "300-chains-r5"])
def test_batch_mixed_chains_vs_oracle(qms, nch, widths):
    """Chains with the given register widths, betas and seeds all different,
    in one call, bit-exact.  48 chains (no more than the CUs) run a workgroup
    per chain with one amplitude per thread, launched per register width; 300
    chains run in LDS with 4 (widest r = 6: 16 KiB per chain), 1 (r <= 3) or 2
    (r = 5) wavefronts per chain.  The 300-chain cases check every 6th chain."""
    rng = np.random.default_rng(9)
    chains = [(float(0.1 + 4.9 * k / (nch - 1)), int(widths[k % len(widths)]), int(1000 + k)) for k in range(nch)]
    rng.shuffle(chains)
    steps, reset_each, therm, maxrev, emin, emax = 120, 2, 10, 30, -1.0, 3.0
    log("batch mixed chains")
    out = qms.batch_run(chains, steps, reset_each, maxrev, emin, emax, 10, therm, rng_after=True)
    for c, (beta, r, seed) in enumerate(chains):
        if nch > 48 and c % 6:
            continue
        q = po.Qms(beta, r, seed, emin, emax, 10, maxrev)
        L = po.lib()
        L.so_qms_init_state(q.p)
        s0 = 0
        for s in range(steps):
            take = (s > s0 + therm) and ((s - s0) % reset_each == 0)
            ret = L.so_qms_metro_step(q.p, 1 if take else 0)
            if ret < 0:
                L.so_qms_init_state(q.p)
                s0 = s + 1
            assert out["rets"][c][s] == ret, f"chain {c} (beta={beta}, r={r}, seed={seed}) step {s}"
            assert int(out["rng_after"][c][s]) == q.rng_state()
        E, X = q.measures()
        nm = int(out["n_meas"][c])
        assert nm == len(E)
        assert_bitwise(out["E"][c][:nm], np.array(E), f"E chain {c}")
        assert_bitwise(out["X"][c][:nm], np.array(X), f"X chain {c}")


def oracle_rets(beta, r, seed, steps, reset_each, therm, maxrev, emin, emax):
    q = po.Qms(beta, r, seed, emin, emax, 10, maxrev)
    L = po.lib()
    L.so_qms_init_state(q.p)
    s0, rets = 0, []
    for s in range(steps):
        take = (s > s0 + therm) and ((s - s0) % reset_each == 0)
        ret = L.so_qms_metro_step(q.p, 1 if take else 0)
        if ret < 0:
            L.so_qms_init_state(q.p)
            s0 = s + 1
        rets.append(ret)
    E, X = q.measures()
    return rets, E, X


@pytest.mark.parametrize("widths", [5, 2], ids=["r2-6", "r2-3"])
def test_batch_fp32_chains_follow_fp64(qms, widths):
    """FP32 amplitudes with FP64 probabilities: a chain's decisions differ from
    the FP64 oracle only when a uniform draw falls within FP32 rounding of a
    probability, so nearly all of 48 chains x 120 steps reproduce the oracle's
    return codes exactly, and those that do carry the same E and X values.
    r in 2..6 runs two wavefronts per chain (8 KiB of LDS for the widest);
    r in 2..3 (1 KiB) one wavefront per chain."""
    chains = [(float(0.1 + 4.9 * k / 47), int(2 + k % widths), int(1000 + k)) for k in range(48)]
    steps, reset_each, therm, maxrev, emin, emax = 120, 2, 10, 30, -1.0, 3.0
    log("batch fp32 chains")
    out = qms.batch_run(chains, steps, reset_each, maxrev, emin, emax, 10, therm, fp32=True, final_state=True)
    o64 = qms.batch_run(chains, steps, reset_each, maxrev, emin, emax, 10, therm, final_state=True)
    same, worst = 0, 0.0
    for c, (beta, r, seed) in enumerate(chains):
        rets, E, X = oracle_rets(beta, r, seed, steps, reset_each, therm, maxrev, emin, emax)
        if list(out["rets"][c]) == rets:
            same += 1
            nm = int(out["n_meas"][c])
            assert nm == len(E)
            assert_bitwise(out["E"][c][:nm], np.array(E), f"E chain {c}")
            assert_bitwise(out["X"][c][:nm], np.array(X), f"X chain {c}")
            assert list(o64["rets"][c]) == rets
            worst = max(worst, infidelity(o64["final"][c], out["final"][c]))
    log(f"{same} of {len(chains)} FP32 chains identical to the FP64 oracle, largest state infidelity {worst:.3g}")
    assert same >= 44
    assert worst <= 1e-5   # the precision pattern's bound on the state


@pytest.mark.parametrize("fp32", [False, True])
def test_batch_final_state_vs_plain_path(qms, fp32):
    """The batch's final-state output against the plain FP64 path's state
    after the same chain (same amplitude order): within 1e-12 in FP64 (the
    measurement sums are summed in another order), infidelity <= 1e-5 in FP32;
    both the LDS layout (a batch larger than the CU count) and, in FP64, the
    one-amplitude-per-thread layout (a single chain)."""
    steps, reset_each, therm, maxrev, emin, emax = 60, 1, 5, 30, -1.0, 3.0
    picks = [(0.7, 3, 77), (2.5, 5, 91)]
    filler = [(1.0, 2, 5000 + k) for k in range(300)]   # more chains than CUs: the LDS layout
    for beta, r, seed in picks:
        plain = qms.run(beta, steps, reset_each, r, seed, emin, emax, 10, therm, maxrev, final_state=True)
        ref = plain["final"][0] + 1j * plain["final"][1]
        for chains in ([(beta, r, seed)] + filler, [(beta, r, seed)]):
            if fp32 and len(chains) == 1:
                continue
            log(f"final state beta {beta} r {r} seed {seed}, {len(chains)} chains, fp32 {fp32}")
            out = qms.batch_run(chains, steps, reset_each, maxrev, emin, emax, 10, therm, fp32=fp32,
                                final_state=True)
            assert list(out["rets"][0]) == list(plain["rets"])
            got = out["final"][0]
            assert got.shape == ref.shape
            if fp32:
                assert infidelity(ref, got) <= 1e-5
            else:
                np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)
