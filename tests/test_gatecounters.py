"""The gate-count model (suqa_amd/host/gatecounters.hpp, the interface of the
reference's gatecounters.hpp:10-163) on the host, without a GPU: the
n-control Toffoli costs against the textbook decomposition the reference
cites (a 2-control Toffoli is 9 one-qubit gates and 6 CNOTs; n > 2 controls
take 2n - 3 of them and one CNOT), the per-counter statistics against
Python's statistics module, activation and record boundaries, and the
printed table's layout against the reference's own printout
(tests/golden/gatecount_c1.txt)."""
import math
import os
import statistics
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include <cstdio>
#include "gatecounters.hpp"
int main() {
    for (uint n = 0; n <= 6; ++n) {
        GateRecord g = GateCounter::n_ctrl_toffoli_gates(n);
        printf("toffoli %u %u %u\n", n, g.ng1, g.ng2);
    }
    GateCounter a("alpha"), b("beta"), idle("idle");
    GateCounterList L;
    L.add_counter(&a); L.add_counter(&b); L.add_counter(&idle);
    L.increment_g1g2(5, 5);              // nothing active: dropped
    a.new_record();
    L.increment_g1g2(3, 1);
    b.new_record();
    L.increment_g1g2(GateRecord(10, 2));
    a.new_record();
    L.increment_g1g2(7, 0);
    b.deactivate();
    L.increment_g1g2(1, 4);
    a.new_record();
    L.increment_g1g2(70000, 3);          // 70000^2 overflows 32 bits
    L.update_cmask_setbits(0x2d);
    printf("cmask %u\n", L.gc_mask_set_qbits);
    for (GateCounter* c : {&a, &b, &idle})
        for (const GateRecord& r : c->grecords) printf("rec %s %u %u\n", c->name.c_str(), r.ng1, r.ng2);
    L.print_counts();
    return 0;
}
'''


@pytest.fixture(scope="module")
def output():
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "t.cpp")
        exe = os.path.join(td, "t")
        with open(src, "w") as f:
            f.write(PROGRAM)
        subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "suqa_amd", "host"), src, "-o", exe],
                       check=True)
        return subprocess.run([exe], check=True, capture_output=True, text=True).stdout


def test_toffoli_costs(output):
    got = {int(p[1]): (int(p[2]), int(p[3])) for p in (l.split() for l in output.splitlines() if l.startswith("toffoli"))}
    assert got[0] == (1, 0)           # the gate itself
    assert got[1] == (0, 1)           # a CNOT
    assert got[2] == (9, 6)           # textbook Toffoli: 9 single-qubit gates + 6 CNOTs
    for n in range(3, 7):             # 2n - 3 Toffolis plus one CNOT
        assert got[n] == (9 * (2 * n - 3), 6 * (2 * n - 3) + 1)


def test_records_and_activation(output):
    recs = [l.split()[1:] for l in output.splitlines() if l.startswith("rec ")]
    a = [(int(x), int(y)) for n, x, y in recs if n == "alpha"]
    b = [(int(x), int(y)) for n, x, y in recs if n == "beta"]
    assert a == [(3 + 10, 1 + 2), (7 + 1, 0 + 4), (70000, 3)]
    assert b == [(10 + 7, 2 + 0)]      # deactivated before the last two increments
    assert not [r for r in recs if r[0] == "idle"]
    assert "cmask 4" in output          # 0x2d has four set bits


def stats_row(name, recs):
    """The table row recomputed independently (population variance / (n - 1))."""
    n = len(recs)
    cols = [name, str(n)]
    for k in (0, 1):
        xs = [r[k] for r in recs]
        mean = sum(xs) / n
        if n > 1:
            # the reference squares the counts in 32-bit unsigned arithmetic
            sq = [(x * x) % 2**32 for x in xs]
            var = abs(sum(sq) / n - mean * mean)
            err = math.sqrt(var / (n - 1))
            if all(x * x < 2**32 for x in xs):
                assert math.isclose(err, statistics.pstdev(xs) / math.sqrt(n - 1), rel_tol=1e-9, abs_tol=1e-9)
            err = 0.0 if err < 1e-10 else err
        else:
            err = 0.0
        cols += [str(round(mean * n)), "%.3g" % mean, "%.3g" % err]
    return cols


def test_table(output):
    lines = output.splitlines()
    i = lines.index("Gate counts")
    with open(os.path.join(ROOT, "tests", "golden", "gatecount_c1.txt")) as f:
        ref = f.read().splitlines()
    assert lines[i:i + 3] == ref[0:3]   # title, blank line, column header: the reference's layout
    rows = {l.split()[0]: l.split() for l in lines[i + 3:] if l.strip()}
    assert rows["alpha"] == stats_row("alpha", [(13, 3), (8, 4), (70000, 3)])
    assert rows["beta"] == stats_row("beta", [(17, 2)])
    assert rows["idle"] == ["idle", "0", "0", "0", "0", "0", "0", "0"]
    # fixed-width columns as the reference prints them (15 + 2 then 9 + 2 per column)
    row = next(l for l in lines if l.startswith("alpha"))
    assert row[17:26].strip() == "3" and row[:15].rstrip() == "alpha"
