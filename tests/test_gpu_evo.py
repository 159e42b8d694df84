"""The reference's `evo` driver (evo.cu:22-103) as `suqa_evo` on the GPU,
against the reference binary's own output file (tests/golden/evo_n3.txt,
20 steps of dt 0.1): same number of lines, identical times, probabilities
within 1e-12 (the engine sums in another order than the reference's
sequential loop), with gate passes on and off."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EVO = os.path.join(ROOT, "suqa_amd", "bin", "suqa_evo")


@pytest.mark.parametrize("tile_bits", ["0", "11"])
def test_evo_table_vs_reference_file(tile_bits):
    print(f"suqa_evo, tile bits {tile_bits}", flush=True)
    want = np.loadtxt(os.path.join(ROOT, "tests", "golden", "evo_n3.txt"))
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "evo.txt")
        env = dict(os.environ, SUQA_HIP_TILE_BITS=tile_bits)
        r = subprocess.run([EVO, "20", "0.1", out], capture_output=True, text=True, timeout=300, env=env)
        assert r.returncode == 0, r.stderr
        assert "total_steps = 20" in r.stdout
        got = np.loadtxt(out)
    assert got.shape == want.shape == (21, 9)
    assert np.array_equal(got[:, 0], want[:, 0])
    np.testing.assert_allclose(got[:, 1:], want[:, 1:], atol=1e-12, rtol=0)
    np.testing.assert_allclose(got[:, 1:].sum(axis=1), 1.0, atol=1e-12)


def test_evo_usage_message():
    r = subprocess.run([EVO], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stdout
