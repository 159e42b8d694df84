"""Gate-count instrumentation (the reference's GATECOUNT build: gatecounters.hpp,
suqa.cu's per-gate increments, qms.cuh's metro-step records) as
`suqa_qms_gatecount` on the GPU, against the table the reference `qms`
binary built with -DGATECOUNT prints for the same arguments
(tests/golden/gatecount_*.txt): identical text, which needs the same gate
calls in the same order, the same control counts and the same record
boundaries.  Run with gate passes on and off."""
import os
import subprocess
import sys
import tempfile

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, GOLDEN)
import make_golden  # noqa: E402  (the cases and the table extraction)

DRIVER = os.path.join(ROOT, "suqa_amd", "bin", "suqa_qms_gatecount")


@pytest.mark.parametrize("tile_bits", ["0", "11"])
@pytest.mark.parametrize("case", make_golden.GATECOUNT_CASES, ids=[c[0] for c in make_golden.GATECOUNT_CASES])
def test_gate_count_table_matches_reference(case, tile_bits):
    name, pos, opts = case
    print(f"suqa_qms_gatecount {name}, tile bits {tile_bits}", flush=True)
    with open(os.path.join(GOLDEN, name + ".txt")) as f:
        want = f.read()
    with tempfile.TemporaryDirectory() as td:
        env = dict(os.environ, SUQA_HIP_TILE_BITS=tile_bits)
        r = subprocess.run([DRIVER] + pos + [os.path.join(td, "out.txt")] + opts, capture_output=True, text=True,
                           timeout=300, env=env)
    assert r.returncode == 0, r.stderr
    assert make_golden.gate_count_table(r.stdout) == want


def test_plain_driver_prints_no_counts():
    with tempfile.TemporaryDirectory() as td:
        r = subprocess.run([os.path.join(ROOT, "suqa_amd", "bin", "suqa_qms"), "1.0", "20", "1", "3",
                            os.path.join(td, "out.txt"), "--seed", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Gate counts" not in r.stdout
