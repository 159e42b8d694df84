"""Writes the golden fixtures in tests/golden/ from the reference engine.

The reference's dense CPU path is compiled from its own sources by
oracle/ref.mk into oracle/_ref/ (never copied into this repository); the
harness oracle/ref_harness.cpp drives it through the reference API and this
script stores what it wrote.  Run from the repository root, in a container
where the reference sources are mounted:

    python tests/golden/make_golden.py [--ref /root/reference]

Fixtures (all JSON; doubles are C99 hex floats, exact):
  pcg32.json            pcg32 / pcg wrapper sequences for six seeds
  gates_n5.json         every gate API case at n=5, with and without gc_mask
  qms_*.json            per-step metro_step return codes, pcg32 state after
                        each step, E/X measurements, reverse counters
  qms_c1_out.txt        the stock reference `qms` output file (E A lines)
                        for beta=1, 1000 steps, r=4, seed 42, E in [-1,3]
  evo_n3.txt            the stock reference `evo` output, 20 steps, dt=0.1
  gatecount_*.txt       the gate-count table printed by the reference `qms`
                        built with -DGATECOUNT (GATECOUNT_CASES below)
  circuit_*.json        the random-circuit family (gate list, state before
                        the measurements, outcomes, final state)
  qsa_*.json            QSA traces: qsa.cu's loop on the reference qsa:: functions,
                        bounded by a number of annealing steps (per step the move,
                        the Szegedy readout and the pcg32 state; per sample E, X,
                        restarts; the final state)
  qsa_out_*.txt         the stock reference `qsa` output file (E X restarts lines)
  circuit_*_summary.json  the same family at sizes too large to store
                        (n = 24: the 30-qubit workload's circuit family):
                        checksum + 64 sample amplitudes before and after
                        the measurements, outcomes
"""
import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

# (name, beta, steps, reset_each, r, seed, ene_min, ene_max, pe, therm, max_rev, record_rev)
QMS_CASES = [
    ("qms_c1", 1.0, 1000, 1, 4, 42, -1.0, 3.0, 10, 100, 100, 1),
    ("qms_r3_b05", 0.5, 300, 2, 3, 7, 0.0, 1.0, 10, 20, 100, 1),
    ("qms_r3_maxrev1", 2.0, 300, 1, 3, 123, -1.0, 3.0, 10, 10, 1, 1),
    ("qms_r5_b3", 3.0, 200, 1, 5, 99, -1.0, 3.0, 10, 10, 100, 0),
    ("qms_r6_b01", 0.1, 60, 3, 6, 5, -1.0, 3.0, 10, 5, 100, 1),
]

CIRCUITS = [("circuit_n10", 10, 200, 7), ("circuit_n6", 6, 20, 3)]
# the same family where the state is too large to store: checksum (pairwise
# sums of re, im and i |psi_i|^2 / N) and 64 sample amplitudes, before and
# after the measurements (the n = 24 one takes ~3 minutes on one core)
SUMMARY_CIRCUITS = [("circuit_n10_summary", 10, 200, 7), ("circuit_n24_summary", 24, 200, 7)]

# (name, beta, sampling, ene qbits, szegedy qbits, seed, ene_min, ene_max, ene_min_sz, ene_max_sz,
#  PE steps, annealing sequences, bound on annealing steps): QSA traces (qsa.cu's loop, bounded)
QSA_CASES = [
    ("qsa_e2s2", 0.25, 3, 2, 2, 7, -4.0, 8.0, -1.0, 3.0, 1, 3, 80),
    ("qsa_e1s2", 1.0, 4, 1, 2, 11, -4.0, 8.0, -1.0, 3.0, 1, 2, 100),
    ("qsa_e2s1", 0.5, 3, 2, 1, 5, -4.0, 8.0, -1.0, 3.0, 2, 2, 80),
    ("qsa_e2s3_default_seq", 0.25, 1, 2, 3, 3, -4.0, 8.0, -1.0, 3.0, 1, 100, 25),
]
# the stock qsa driver's output file: (name, arguments after the output file)
QSA_OUT_CASES = [("qsa_out_e2s2", ["0.25", "3", "2", "2"], ["--seed", "7", "--annealing-sequences", "3"])]

# (name, qms command-line arguments after the output file)
GATECOUNT_CASES = [
    ("gatecount_c1", ["1.0", "1000", "1", "4"], ["--seed", "42", "--ene-min", "-1", "--ene-max", "3",
                                                 "--PE-steps", "10"]),
    ("gatecount_r3", ["2.0", "300", "2", "3"], ["--seed", "123", "--ene-min", "-1", "--ene-max", "3",
                                                "--max-reverse", "3", "--thermalization", "10"]),
]


def gate_count_table(stdout):
    """The lines from 'Gate counts' to the blank line after the table."""
    lines = stdout.splitlines()
    i = lines.index("Gate counts")
    j = i + 3   # title, blank, header
    while j < len(lines) and lines[j].strip():
        j += 1
    return "\n".join(lines[i:j]) + "\n"


def run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    run(["make", "-s", "-f", "oracle/ref.mk", "REF=" + a.ref, "-j8"], cwd=ROOT)
    h = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    out = lambda n: os.path.join(HERE, n)
    run([h, "pcg", out("pcg32.json")], stdout=subprocess.DEVNULL)
    run([h, "gates", out("gates_n5.json")], stdout=subprocess.DEVNULL)
    for (name, beta, steps, re_, r, seed, emin, emax, pe, th, mr, rr) in QMS_CASES:
        run([h, "qms", repr(beta), str(steps), str(re_), str(r), str(seed), repr(emin),
             repr(emax), str(pe), str(th), str(mr), str(rr), out(name + ".json")],
            stdout=subprocess.DEVNULL)
    for (name, n, layers, seed) in CIRCUITS:
        run([h, "circuit", str(n), str(layers), str(seed), out(name + ".json")],
            stdout=subprocess.DEVNULL)
    for (name, n, layers, seed) in SUMMARY_CIRCUITS:
        run([h, "circuit_summary", str(n), str(layers), str(seed), out(name + ".json")],
            stdout=subprocess.DEVNULL)
    for (name, beta, sampling, e, sz, seed, emin, emax, emin_s, emax_s, pe, ann, bound) in QSA_CASES:
        run([h, "qsa", repr(beta), str(sampling), str(e), str(sz), str(seed), repr(emin), repr(emax), repr(emin_s),
             repr(emax_s), str(pe), str(ann), str(bound), out(name + ".json")], stdout=subprocess.DEVNULL)
    for (name, pos, opts) in QSA_OUT_CASES:
        tmp = out(name + ".txt")
        if os.path.exists(tmp):
            os.remove(tmp)
        run([os.path.join(ROOT, "oracle", "_ref", "qsa")] + pos + [tmp] + opts, stdout=subprocess.DEVNULL)
    # stock drivers: output files, byte for byte
    tmp = out("qms_c1_out.txt")
    if os.path.exists(tmp):
        os.remove(tmp)
    run([os.path.join(ROOT, "oracle", "_ref", "qms"), "1.0", "1000", "1", "4", tmp, "--seed", "42",
         "--ene-min", "-1", "--ene-max", "3", "--PE-steps", "10"], stdout=subprocess.DEVNULL)
    tmp = out("evo_n3.txt")
    if os.path.exists(tmp):
        os.remove(tmp)
    run([os.path.join(ROOT, "oracle", "_ref", "evo"), "20", "0.1", tmp], stdout=subprocess.DEVNULL)
    for (name, pos, opts) in GATECOUNT_CASES:
        tmp = out(name + "_out.tmp")
        cmd = [os.path.join(ROOT, "oracle", "_ref", "qms_gatecount")] + pos + [tmp] + opts
        print("+", " ".join(cmd), flush=True)
        r = subprocess.run(cmd, check=True, capture_output=True, text=True)
        os.remove(tmp)
        with open(out(name + ".txt"), "w") as f:
            f.write(gate_count_table(r.stdout))
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    sys.exit(main())
