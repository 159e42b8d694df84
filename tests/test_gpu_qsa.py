"""QSA on the GPU: the C++ qsa:: mirror over the engine (suqa_qsa_run)
reproduces the reference's quantum simulated annealing loop (qsa.cu) bit for
bit, against traces the reference build wrote (tests/golden/qsa_*.json): per
annealing step the sample and restart indices, the move drawn, the Szegedy
register's readout and the pcg32 state after it; per sample the energy, X,
restart count and pcg32 state; the final pcg32 state; the final state within
1e-12.  The reference's bitmap puts the tilde spins on the energy and Szegedy
registers' qubits, so these runs apply H, swaps and XX rotations on qubits of
the active gc_mask (the engine's reproduction of those loops is pinned gate by
gate in gates_n5.json).  With one sweep per gate and with gate passes; and the
suqa_qsa command line's output file equal, byte for byte, to the stock
reference driver's."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests.common import golden_state, load

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURES = ["qsa_e2s2.json", "qsa_e1s2.json", "qsa_e2s1.json", "qsa_e2s3_default_seq.json"]


@pytest.fixture(scope="module")
def qsa():
    from suqa_amd import qsa
    return qsa


@pytest.mark.parametrize("tile_bits", ["0", "11"])
@pytest.mark.parametrize("fx", FIXTURES)
def test_qsa_trace_vs_reference(qsa, fx, tile_bits, monkeypatch):
    monkeypatch.setenv("SUQA_HIP_TILE_BITS", tile_bits)
    d = load(fx)
    h = float.fromhex
    print(f"qsa {fx} tile bits {tile_bits}", flush=True)
    out = qsa.run(h(d["beta"]), d["sampling"], d["ene_qbits"], d["szegedy_qbits"], d["seed"], h(d["ene_min"]),
                  h(d["ene_max"]), h(d["ene_min_szegedy"]), h(d["ene_max_szegedy"]), d["pe_steps"],
                  d["annealing_sequences"], d["max_annealing_steps"], final_state=True)
    assert out["n_steps"] == d["annealing_steps"] == len(d["steps"])
    for k, (got, rng, want) in enumerate(zip(out["steps"], out["rng_after"], d["steps"])):
        assert [int(x) for x in got] == want[:5], f"annealing step {k}"
        assert int(rng) == int(want[5]), f"pcg32 state after annealing step {k}"
    assert len(out["E"]) == len(d["samples"])
    for k, (E, X, rej, srng) in enumerate(d["samples"]):
        assert out["E"][k] == h(E) and out["X"][k] == h(X), f"sample {k}"
        assert int(out["rejections"][k]) == rej and int(out["sample_rng"][k]) == int(srng), f"sample {k}"
    assert int(out["final_rng"]) == int(d["final_rng"])
    re, im = golden_state(d["final_state"])
    np.testing.assert_allclose(out["final"][0], re, atol=1e-12, rtol=0)
    np.testing.assert_allclose(out["final"][1], im, atol=1e-12, rtol=0)


def test_qsa_driver_output_file_matches_reference(qsa):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.txt")
        r = subprocess.run([qsa.DRIVER, "0.25", "3", "2", "2", out, "--seed", "7", "--annealing-sequences", "3"],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        assert "all fine :)" in r.stdout
        with open(out, "rb") as f, open(os.path.join(ROOT, "tests", "golden", "qsa_out_e2s2.txt"), "rb") as g:
            assert f.read() == g.read()
