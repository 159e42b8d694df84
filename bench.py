"""Benchmark: the five workloads of the SUQA state-vector engine on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 30] [--no-cpu]

Headline (c2): 30 qubits, complex FP64 (16 GiB in HBM), |0> then an H layer,
then 200 layers of 30 random gates from {H, X, RX, RZ, U1, CX, MCX(3)} drawn
with pcg32 seed 7 (suqa_amd/workloads.py); one step = one layer.  The FP64
engine applies the gates in passes (runs of gates whose flipped qubits fit
one tile, applied to LDS-resident tiles in one HBM round trip); by default a
pass may take a later gate ahead of earlier ones it commutes with (the same
product, amplitudes within 1e-12 of call order, extra.gate_order_sample
compares both orders on 40 layers); --gate-order call keeps call order,
bit-identical to one sweep per gate.  The tile size is autotuned first on 5
layers, then the state is re-initialised.  W warmup layers run untimed,
exactly K layers are timed between stream synchronizations, and the rest of
the 200 layers run untimed after them, so the checksum and the 8 final
measure_qbit outcomes are those of the whole circuit whatever K and W are.

value = gate-equivalent GB/s: the algorithmic bytes of the timed layers, each
gate counted as its own sweep (32 B per amplitude it modifies: H/X/RX/RZ 32N,
U1/CX 16N, MCX3 4N), over the wall time.  With passes holding several gates
this exceeds the HBM roof; the bytes the kernels actually move are in
extra.c2_hbm (their sum over the same wall time, and its fraction of 8 TB/s).
roofline: the dominant kernel (the tile pass), the HBM bytes it must move
(every amplitude of the tiles it touches read and written once, 32 B) over
the sum of its per-launch HIP-event durations, against the 8 TB/s HBM3E peak;
traffic = those bytes x the PMC ratio measured for it (rocprofv3 FETCH_SIZE /
WRITE_SIZE, scaled by a copy kernel's counts).  cpu_baseline: the oracle (a
single-threaded restatement of the reference CPU path, gcc -O3
-ffp-contract=off) on the first layer at 27 qubits.

extra, per workload (each with its roofline and its CPU leg):
  fused (c3)       -- the same 200 layers on an FP32 state, gates fused into
                      K-qubit blocks (K and the mode autotuned on 10 layers)
                      applied by FP32 MFMA; fidelity against the FP64 state,
                      outcomes, the dominant block kernel's fraction of its
                      binding roof (HBM bytes / 8 TB/s or MFMA flops / 157.3
                      TFLOP/s, whichever takes longer), and the oracle's FP64
                      state at 24 qubits (first 2 layers, 1 core) beside it;
  trotter_qpe (c4) -- 24 system + 6 register qubits, 96 Pauli strings of
                      weight 2-4 (pcg32 seed 11), dt 0.05, 20 Trotter steps
                      per power: a timed sample of controlled steps in FP64
                      passes and FP32 fused form with the dominant kernels'
                      HBM fractions, then the whole QPE (120,960 controlled
                      rotations) measured in both precisions with the
                      histogram error; the oracle's QPE at 10 + 4 qubits
                      (same generator, 20 steps per power, 1 core) beside it;
  qms_plain (c1)   -- the frustrated-triangle QMS (beta 1, 1000 steps, r = 4,
                      E in [-1,3], seed 42) on the plain FP64 GPU path, also
                      in the two-register layout and as a device-resident
                      batch of one; the oracle on 1 core beside it;
  qms_batch (c5)   -- 4096 chains in the two-register layout the workload
                      describes (r uniform in 3..6: 10 to 16 qubits per chain;
                      pcg32 seed 5; betas 0.1..5.0; seeds 1000+i; 500 steps),
                      FP64 and FP32, chain-steps/s, the LDS-bandwidth fraction
                      (algorithmic LDS bytes / kernel time against 256 CUs x
                      128 B/clk x 2.4 GHz), per-beta <E> of FP32 against FP64,
                      the FP32 chains' final-state infidelity against FP64,
                      sample chains against the oracle, and the oracle's
                      two-register chains over the host cores beside it; the
                      same batch in the master one-register layout too.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CIRCUIT_LAYERS = 200
# HBM bytes moved per launch relative to the algorithmic bytes, per kernel family, from two
# rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) on this workload, scaled by the counts a
# 4 GiB copy kernel gives (reads x2.000, writes x1.000): profiles/tile_pass_n30_fp64.md
# (tile passes) and profiles/gate_sweep_n30_fp64.md (one sweep per gate).
PMC_TRAFFIC_RATIO = {"tile_pass": 1.000, "h_pair": 1.000, "pauli_diag": 1.000, "pauli_pair": 1.000,
                     "phase": 1.154, "swap_pair": 1.023}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def hbm_copy_calibration(capi, eng, nbytes=4 << 30, reps=5):
    """Measured copy bandwidth (read + write bytes / time) of a plain copy kernel."""
    import ctypes as C
    a, b = C.c_void_p(), C.c_void_p()
    capi.check(capi.lib().suqa_hip_device_alloc(nbytes, C.byref(a)))
    capi.check(capi.lib().suqa_hip_device_alloc(nbytes, C.byref(b)))
    try:
        eng.timing(True)
        eng.timing_reset()
        for _ in range(reps):
            capi.check(capi.lib().suqa_hip_copy_kernel(eng.ctx, b, a, nbytes))
        rep = eng.timing_report()["copy"]
        eng.timing(False)
        eng.timing_reset()
        return rep["bytes"] / (rep["ms"] * 1e-3) / 1e9
    finally:
        capi.lib().suqa_hip_device_free(a)
        capi.lib().suqa_hip_device_free(b)


def cpu_baseline(n=27):
    """Oracle (reference CPU path restated, 1 thread) on layer 0 of the same circuit at n qubits."""
    from oracle import pyoracle as po
    from suqa_amd import workloads as wl
    gates, _, _ = wl.random_circuit(n, 1, 7)
    og, ng, _, _ = po.circuit(n, 1, 7)
    st = po.State(n)
    for q in range(n):
        st.h(q)
    t0 = time.perf_counter()
    for g in og[:ng]:
        st.apply_gate(g)
    dt = time.perf_counter() - t0
    nbytes = sum(wl.gate_bytes(g, n) for g in gates)
    del st
    return dict(value=round(nbytes / dt / 1e9, 2), unit="GB/s (gate-equivalent)", cores=1, kind="port",
                sample=f"first layer (30 gates) of the seed-7 circuit at {n} qubits, "
                       f"{dt:.1f} s on 1 host core (oracle/suqa_oracle.c, gcc -O3 -ffp-contract=off)")


def gate_order_sample(eng, arr, n, order, layers=40):
    """The first layers of the circuit in both gate orders of the FP64 passes: ms per
    layer, passes, and the largest checksum difference between the two states."""
    out = {}
    for name in ("call", "packed"):
        eng.set_gate_order(name == "packed")
        eng.init_state()
        for q in range(n):
            eng.h(q)
        eng.synchronize()
        p0 = eng.pass_stats()["passes"]
        t0 = time.perf_counter()
        eng.run_gates(arr, layers * n, 0)
        eng.synchronize()
        dt = time.perf_counter() - t0
        out[name] = {"ms_per_layer": round(dt / layers * 1e3, 3), "passes": eng.pass_stats()["passes"] - p0,
                     "checksum": eng.checksum()}
    eng.set_gate_order(order == "packed")
    diff = max(abs(x - y) for x, y in zip(out["call"]["checksum"], out["packed"]["checksum"]))
    return {"layers": layers, "call_order": {k: v for k, v in out["call"].items() if k != "checksum"},
            "packed": {k: v for k, v in out["packed"].items() if k != "checksum"},
            "checksum_max_abs_diff": diff}


def cpu_fp64_state_c3(n=24, layers=2):
    """The FP64 CPU reference state of the fused workload's circuit family: the
    oracle on the first layers at n qubits, 1 core (ms per layer)."""
    from oracle import pyoracle as po
    og, ng, _, _ = po.circuit(n, layers, 7)
    st = po.State(n)
    for q in range(n):
        st.h(q)
    t0 = time.perf_counter()
    for g in og[:ng]:
        st.apply_gate(g)
    dt = time.perf_counter() - t0
    del st
    return {"value": round(dt / layers * 1e3, 1), "unit": "ms per layer", "cores": 1, "kind": "port",
            "sample": f"first {layers} layers of the seed-7 circuit at {n} qubits on the oracle (FP64), {dt:.1f} s"}


def cpu_qpe_c4(nsys=10, nreg=4, steps=20):
    """The oracle's Trotter + QPE at a reduced size (same generator), 1 core."""
    from oracle import pyoracle as po
    terms = po.trotter_terms(nsys, 96, 11)
    st = po.State(nsys + nreg)
    t0 = time.perf_counter()
    po.qpe_trotter(st, nsys, nreg, terms, 0.05, steps)
    dt = time.perf_counter() - t0
    rot = 96 * steps * ((1 << nreg) - 1)
    return {"value": round(rot / dt, 1), "unit": "controlled rotations/s", "cores": 1, "kind": "port",
            "amplitude_rotations_per_s": round(rot * (1 << (nsys + nreg - 1)) / dt, 0),
            "sample": f"the whole QPE at {nsys} + {nreg} qubits, 96 strings, 20 steps per power "
                      f"({rot} controlled rotations) on the oracle, {dt:.1f} s"}


def _oracle_chain(args):
    """One chain on the oracle, in a worker process: (beta, r, seed, steps, emin, emax, two) -> rets."""
    from oracle import pyoracle as po
    beta, r, seed, steps, emin, emax, two = args
    _, rets = po.Qms(beta, r, seed, emin, emax, 10, 100, two_registers=two).run(steps, 1, 100)
    return rets


def oracle_chains_all_cores(chains, steps, emin, emax, two=False):
    """The oracle on a sample of chains, one single-threaded process per chain over
    the host cores the box gives a GPU (at most 16): (chain-steps/s, workers, rets)."""
    import concurrent.futures as cf
    import multiprocessing as mp
    workers = max(1, min(16, os.cpu_count() or 1, len(chains)))
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        list(ex.map(_oracle_chain, [(1.0, 2, 1, 1, emin, emax, False)] * workers))   # start every worker
        t0 = time.perf_counter()
        rets = list(ex.map(_oracle_chain, [(b, r, s, steps, emin, emax, two) for (b, r, s) in chains]))
        dt = time.perf_counter() - t0
    return len(chains) * steps / dt, workers, rets


# LDS roof of the batched QMS kernels: 256 CUs x 128 B/clk x 2.4 GHz (peak engine clock)
LDS_PEAK_GBS = 256 * 128 * 2.4


def batch_chains(nch, seed=5):
    """The c5 chains: r = 3 + randint(0, 4) from pcg32(seed), betas on the grid, seeds 1000+i."""
    from suqa_amd import workloads as wl
    rg = wl.Pcg32(seed)
    return [(0.1 + 4.9 * i / (nch - 1), 3 + rg.randint(0, 4), 1000 + i) for i in range(nch)]


def qms_batch_only():
    """The 4096-chain batches of the QMS legs alone, FP64 then FP32, two-register then
    one-register layout, one untimed run each first (a profiling mode)."""
    from suqa_amd import qms
    emin, emax = -1.0, 3.0
    chains, bsteps = batch_chains(4096), 500
    out = {}
    for two in (True, False):
        for fp32 in (False, True):
            name = f"qms_batch{'' if two else '_one_register'}{'_fp32' if fp32 else ''}"
            qms.batch_run(chains, bsteps, 1, 100, emin, emax, 10, 100, fp32=fp32, two_registers=two)
            b = qms.batch_run(chains, bsteps, 1, 100, emin, emax, 10, 100, fp32=fp32, two_registers=two,
                              lds_bytes=True)
            fig = batch_figures(b, len(chains), bsteps, 1.0)
            out[name] = {"chain_steps_per_s": fig["chain_steps_per_s"], "lds_roofline": fig["lds_roofline"]}
            log(f"{name}: {fig['chain_steps_per_s']:.0f} chain-steps/s, LDS {fig['lds_roofline']}")
    return out


def batch_figures(b, nch, bsteps, plain_rate):
    rate = nch * bsteps / b["seconds"]
    lds = float(np.sum(b["lds_bytes"])) if b["lds_bytes"] is not None else 0.0
    lds_gbs = lds / b["seconds"] / 1e9
    return {"chain_steps_per_s": round(rate, 0), "chains": nch, "steps": bsteps,
            "kernel_seconds": round(b["seconds"], 4),
            "speedup_vs_plain_fp64_gpu": round(rate / plain_rate, 1),
            "mean_acceptance": round(float(b["accepted"].mean()) / bsteps, 4),
            "lds_roofline": {"bound": "lds", "achieved": round(lds_gbs, 1), "peak": round(LDS_PEAK_GBS, 1),
                             "unit": "GB/s", "frac": round(lds_gbs / LDS_PEAK_GBS, 4),
                             "lds_bytes_per_chain_step": round(lds / (nch * bsteps), 1)}}


def per_beta(chains, b64, b32):
    """<E> of FP32 against FP64 chains in 8 beta bins, in units of the FP64 standard error."""
    nch = len(chains)
    bins = []
    for k in range(8):
        idx = range(k * nch // 8, (k + 1) * nch // 8)

        def chain_means(out):
            return np.array([out["E"][c][:int(out["n_meas"][c])].mean() for c in idx if int(out["n_meas"][c])])
        m64, m32 = chain_means(b64), chain_means(b32)
        se = m64.std(ddof=1) / np.sqrt(len(m64))
        bins.append({"beta_min": round(chains[idx[0]][0], 3), "beta_max": round(chains[idx[-1]][0], 3),
                     "E_fp64": round(float(m64.mean()), 6), "E_fp32": round(float(m32.mean()), 6),
                     "fp64_stderr": round(float(se), 6),
                     "diff_in_stderr": round(float(abs(m32.mean() - m64.mean()) / se), 3)})
    return bins


def state_infidelity(o64, o32, bound=1e-5):
    """Each chain's state after its last step, FP32 against FP64, over the chains
    whose decisions agree: 1 - |<psi64|psi32>|^2 / (|psi64|^2 |psi32|^2)."""
    inf = []
    for c in range(len(o64["final"])):
        if np.array_equal(o64["rets"][c], o32["rets"][c]):
            a, b = o64["final"][c], o32["final"][c]
            inf.append(max(0.0, 1.0 - abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)))
    inf = np.array(inf)
    return {"chains": int(len(inf)), "max": float(inf.max()), "mean": float(inf.mean()),
            "median": float(np.median(inf)), "above_1e-5": int((inf > bound).sum())}


def qms_legs(no_cpu):
    """Plain FP64 GPU QMS (c1), batched chains (c5), their oracle checks and CPU rates."""
    from suqa_amd import qms
    beta, steps, r, seed, emin, emax = 1.0, 1000, 4, 42, -1.0, 3.0
    plain = qms.run(beta, steps, 1, r, seed, emin, emax, 10, 100, 100)
    plain_rate = steps / plain["seconds"]
    res = {"qms_plain": {"steps_per_s": round(plain_rate, 1), "steps": steps,
                         "accepted": int(plain["accepted"]), "seconds": round(plain["seconds"], 4),
                         "config": "beta=1.0, 1000 steps, r=4 (8 qubits), E in [-1,3], seed 42, reset_each 1"}}
    log(f"qms plain: {plain_rate:.1f} steps/s")
    # the same chain as the only chain of the batched kernel: one workgroup runs the
    # whole chain on the device (no launches, no host round trips per measurement)
    qms.batch_run([(beta, r, seed)], steps, 1, 100, emin, emax, 10, 100)   # untimed, as the batches below
    one = qms.batch_run([(beta, r, seed)], steps, 1, 100, emin, emax, 10, 100)
    res["qms_single_chain_device"] = {
        "steps_per_s": round(steps / one["seconds"], 1), "kernel_seconds": round(one["seconds"], 5),
        "rets_match_plain": bool(np.array_equal(one["rets"][0], np.asarray(plain["rets"]))),
        "config": "as qms_plain, run as a batch of one chain (one workgroup, one amplitude per thread)"}
    log(f"qms single chain on the device: {steps / one['seconds']:.1f} steps/s")
    # the same chain in the two-register layout the workload describes (E_old in
    # its own register: 3 + 2 * 4 + 1 = 12 qubits)
    two = qms.run(beta, steps, 1, r, seed, emin, emax, 10, 100, 100, two_registers=True)
    res["qms_plain_two_registers"] = {"steps_per_s": round(steps / two["seconds"], 1), "steps": steps,
                                      "accepted": int(two["accepted"]), "seconds": round(two["seconds"], 4),
                                      "mean_E": round(float(np.mean(two["E"])), 6),
                                      "mean_E_one_register": round(float(np.mean(plain["E"])), 6),
                                      "config": "as qms_plain, two 4-qubit energy registers (12 qubits)"}
    log(f"qms plain, two registers: {steps / two['seconds']:.1f} steps/s")
    nch, bsteps = 4096, 500
    chains = batch_chains(nch)
    runs, finals = {}, {}
    # each batch is run once untimed first (with the chains' final states, for
    # the FP32 infidelity): the GPU's clocks after the previous leg (a different
    # kernel, another load) would otherwise set the figure
    for two_reg in (True, False):
        for fp32 in (False, True):
            finals[(two_reg, fp32)] = qms.batch_run(chains, bsteps, 1, 100, emin, emax, 10, 100, fp32=fp32,
                                                    two_registers=two_reg, final_state=True)
            runs[(two_reg, fp32)] = qms.batch_run(chains, bsteps, 1, 100, emin, emax, 10, 100, fp32=fp32,
                                                  two_registers=two_reg, lds_bytes=True)
    for two_reg, key in ((True, "qms_batch"), (False, "qms_batch_one_register")):
        b, b32 = runs[(two_reg, False)], runs[(two_reg, True)]
        res[key] = dict(batch_figures(b, nch, bsteps, plain_rate), dtype="f64",
                        layout=("two registers (E_old in its own register): 3 + 2r + 1 = 10..16 qubits per "
                                "chain, stored exactly in 2^(r+4) amplitudes" if two_reg else
                                "one register (master): r + 4 = 7..10 qubits per chain"),
                        register_widths=[3, 4, 5, 6])
        same = int(sum(1 for c in range(nch) if np.array_equal(b32["rets"][c], b["rets"][c])))
        res[key + "_fp32"] = dict(batch_figures(b32, nch, bsteps, plain_rate),
                                  dtype="f32 amplitudes, f64 probabilities", chains_identical_to_fp64=same,
                                  state_infidelity=state_infidelity(finals[(two_reg, False)],
                                                                    finals[(two_reg, True)]),
                                  per_beta_E=per_beta(chains, b, b32))
        log(f"{key}: {res[key]['chain_steps_per_s']:.0f} chain-steps/s, fp32 "
            f"{res[key + '_fp32']['chain_steps_per_s']:.0f} ({same}/{nch} chains identical to FP64, state "
            f"infidelity max {res[key + '_fp32']['state_infidelity']['max']:.2g}), "
            f"LDS {res[key]['lds_roofline']['frac']:.3f} / {res[key + '_fp32']['lds_roofline']['frac']:.3f} of peak")
    if no_cpu:
        return res
    from oracle import pyoracle as po
    q2 = po.Qms(beta, r, seed, emin, emax, 10, 100, two_registers=True)
    _, rets2 = q2.run(steps, 1, 100)
    res["qms_plain_two_registers"]["rets_match_oracle"] = bool(list(two["rets"]) == rets2)
    q = po.Qms(beta, r, seed, emin, emax, 10, 100)
    t0 = time.perf_counter()
    _, rets = q.run(steps, 1, 100)
    dt = time.perf_counter() - t0
    res["qms_plain"]["rets_match_oracle"] = bool(list(plain["rets"]) == rets)
    res["cpu_baseline_qms"] = {"value": round(steps / dt, 1), "unit": "steps/s", "cores": 1, "kind": "port",
                               "sample": "the same 1000-step chain on the oracle (oracle/suqa_oracle.c, 1 core)"}
    # sample chains against the oracle: the two-register batch against the oracle's
    # one-register chains with the reference's comparison draws (identical return
    # codes to its two-register chains, tests/test_qms_two_registers.py; a 16-qubit
    # full-state chain would take ~30 s here), the one-register batch directly
    b2, b1 = runs[(True, False)], runs[(False, False)]
    ok2 = ok1 = True
    for c in range(0, nch, nch // 12):
        bb, rr, ss = chains[c]
        _, o2 = po.Qms(bb, rr, ss, emin, emax, 10, 100, compare_draws=True).run(bsteps, 1, 100)
        ok2 &= list(b2["rets"][c]) == o2
        _, o1 = po.Qms(bb, rr, ss, emin, emax, 10, 100).run(bsteps, 1, 100)
        ok1 &= list(b1["rets"][c]) == o1
    res["qms_batch"]["sample_chains_match_oracle"] = bool(ok2)
    res["qms_batch_one_register"]["sample_chains_match_oracle"] = bool(ok1)
    # the CPU baseline as the workload defines it: one oracle process per chain over
    # the host cores; two registers: the oracle's full-state chains on a 16-chain
    # sample (4 of each width), 100 steps each
    try:
        s2 = [chains[c] for c in range(0, nch, nch // 64)]
        s2 = [ch for w in (3, 4, 5, 6) for ch in [x for x in s2 if x[1] == w][:4]]
        rate_cpu, workers, orets = oracle_chains_all_cores(s2, 100, emin, emax, two=True)
        res["qms_batch"]["cpu_baseline"] = {
            "value": round(rate_cpu, 1), "unit": "chain-steps/s", "cores": workers, "kind": "port",
            "sample": f"{len(s2)} two-register chains (4 of each width 3..6), 100 steps, on the oracle's "
                      f"full-state two-register chain, one process per chain"}
        res["qms_batch"]["speedup_vs_cpu_baseline"] = round(res["qms_batch"]["chain_steps_per_s"] / rate_cpu, 1)
        res["qms_batch"]["sample16_first100_match_oracle"] = bool(all(
            list(b2["rets"][chains.index(ch)][:100]) == o for ch, o in zip(s2, orets)))
        log(f"oracle, {len(s2)} two-register chains on {workers} processes: {rate_cpu:.0f} chain-steps/s")
        s1 = [chains[c] for c in range(0, nch, nch // 64)]
        rate1, workers, orets = oracle_chains_all_cores(s1, bsteps, emin, emax)
        res["qms_batch_one_register"]["cpu_baseline"] = {
            "value": round(rate1, 1), "unit": "chain-steps/s", "cores": workers, "kind": "port",
            "sample": f"{len(s1)} chains, 500 steps, on the oracle, one process per chain"}
        res["qms_batch_one_register"]["speedup_vs_cpu_baseline"] = round(
            res["qms_batch_one_register"]["chain_steps_per_s"] / rate1, 1)
        res["qms_batch_one_register"]["sample64_chains_match_oracle"] = bool(all(
            list(b1["rets"][chains.index(ch)]) == o for ch, o in zip(s1, orets)))
        log(f"oracle, {len(s1)} chains on {workers} processes: {rate1:.0f} chain-steps/s")
    except Exception as ex:   # worker processes unavailable
        res["qms_batch"]["oracle_all_cores_error"] = repr(ex)
    return res


def kernel_roofline(rep, fam, peak_gbs=HBM_PEAK_GBS, mfma_tflops=None):
    """Fraction of the binding roof of kernel family fam from a timing report: HBM
    time = bytes / peak, MFMA time = flops / peak; frac = the longer / measured."""
    r = rep[fam]
    t = r["ms"] * 1e-3
    t_hbm = r["bytes"] / (peak_gbs * 1e9)
    t_mfma = r.get("flops", 0.0) / (mfma_tflops * 1e12) if mfma_tflops else 0.0
    bound = "mfma" if t_mfma > t_hbm else "hbm"
    return {"kernel": fam, "bound": bound, "launches": r["launches"], "avg_launch_ms": round(r["ms"] / r["launches"], 4),
            "hbm_gbs": round(r["bytes"] / t / 1e9, 1), "hbm_frac": round(t_hbm / t, 4),
            "mfma_tflops": round(r.get("flops", 0.0) / t / 1e12, 2) if mfma_tflops else None,
            "mfma_frac": round(t_mfma / t, 4) if mfma_tflops else None,
            "frac": round(max(t_hbm, t_mfma) / t, 4)}


def fused_leg(capi, wl, eng, gates, n, layers, no_cpu, use_cache=True):
    """FP32 fused-block run of the whole circuit; eng holds the FP64 state after all layers."""
    from suqa_amd import tune_cache
    f32 = capi.F32Engine(n)
    try:
        sample = capi.gates_from(gates[:10 * n])
        if use_cache:
            best, tms, how = tune_cache.f32_config(f32, sample, "random-circuit-seed7")
        else:
            (best, tms), how = f32.autotune(sample), "off"
        log(f"fused: autotuned width {best}, passes {f32.passes} (cache {how}; {tms})")
        arr = capi.gates_from(gates[:layers * n])
        t0 = time.perf_counter()
        capi.fuse_plan(arr, n, best)
        t_plan = time.perf_counter() - t0
        f32.init_state()
        f32.run_gates(capi.gates_from([wl.G(wl.H, q, [], 0.0) for q in range(n)]), best)
        f32.synchronize()
        f32.timing(True)
        f32.timing_reset()
        t0 = time.perf_counter()
        nb = f32.run_gates(arr, best)
        f32.synchronize()
        dt = time.perf_counter() - t0
        rep = f32.timing_report()
        f32.timing(False)
        fid = f32.fidelity(eng)
        norm = f32.vnorm()
        log(f"fused: {nb} blocks for {layers * n} gates in {dt:.3f} s, fidelity {fid!r}")
        fam = max(rep, key=lambda k: rep[k]["ms"])
        roof = kernel_roofline(rep, fam, mfma_tflops=157.3)
        log(f"fused: {fam} {roof}")
        out = {
            "width": None if f32.passes == 3 else best, "mode": ["blocks", "passes", "sparse", "gates"][f32.passes],
            "tile_bits": f32.tile_bits if f32.passes == 3 else None,
            "ms_per_gate_by_mode_width": {k: (None if v is None or v < 0 else round(v, 4)) for k, v in (tms or {}).items()},
            "tune_cache": how,
            "layers": layers, "gates": layers * n, "blocks": nb, "gates_per_block": round(layers * n / nb, 2),
            "seconds": round(dt, 4), "layers_per_s": round(layers / dt, 3),
            "ms_per_layer": round(dt / layers * 1e3, 3), "host_fusion_seconds": round(t_plan, 4),
            "roofline": roof,
            "kernel_families": {k: dict(launches=v["launches"], ms=round(v["ms"], 3),
                                        gbs=round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1),
                                        tflops=round(v["flops"] / max(v["ms"], 1e-9) / 1e9, 2))
                                for k, v in rep.items()},
            "fidelity_vs_fp64": fid, "norm": norm, "dtype": "f32 state, f32 MFMA, f64 reductions"}
        if not no_cpu:
            out["cpu_fp64_reference"] = cpu_fp64_state_c3()
        return f32, out
    except Exception:
        f32.close()
        raise


def trotter_leg(capi, wl, no_fused, full_qpe=True, no_cpu=False):
    """Controlled Trotter steps of the 24 + 6 qubit QPE (FP64 passes, FP32 fused), then the whole QPE."""
    nsys, nreg, dt, steps = 24, 6, 0.05, 20
    n = nsys + nreg
    terms = wl.trotter_terms(nsys, 96, 11)
    top = n - 1
    ops = capi.pauli_ops(wl.trotter_step_ops(terms, dt, 1 << top))
    e = capi.Engine(n)
    try:
        e.init_state()
        e.h(top)
        e.run_paulis(ops)          # warm-up step
        e.synchronize()
        ps0 = e.pass_stats()
        e.timing(True)
        e.timing_reset()
        t0 = time.perf_counter()
        for _ in range(steps - 1):
            e.run_paulis(ops)
        e.synchronize()
        dt64 = time.perf_counter() - t0
        rep = e.timing_report()
        e.timing(False)
        ps1 = e.pass_stats()
        rot = 96 * (steps - 1)
        fam = max(rep, key=lambda k: rep[k]["ms"])
        res = {"qubits": n, "terms": 96, "dt": dt, "steps_per_power": steps, "sample_rotations": rot,
               "fp64_seconds": round(dt64, 4), "fp64_rotations_per_s": round(rot / dt64, 1),
               "fp64_gate_equivalent_gbs": round(rot * 16 * (1 << n) / dt64 / 1e9, 1),
               "fp64_gates_per_pass": round((ps1["pass_gates"] - ps0["pass_gates"]) /
                                            max(1, ps1["passes"] - ps0["passes"]), 2),
               "fp64_roofline": kernel_roofline(rep, fam)}
        log(f"trotter fp64: {rot / dt64:.0f} rotations/s, {res['fp64_roofline']}")
        f = None
        width = None
        if not no_fused:
            f = capi.F32Engine(n)
            # mode and block width: one controlled Trotter step timed at each
            tw, cand = {}, {}
            for name, mode in (("blocks", 0), ("passes", 1), ("sparse", 2)):
                for wdt in (4, 5, 6):
                    cand[f"{name}{wdt}"] = (mode, wdt, 12)
            for tb in (12, 13):   # gate passes: no fusion, the width is ignored
                cand[f"gates{tb}"] = (3, 4, tb)
            for key, (mode, wdt, tb) in cand.items():
                f.passes = mode
                f.tile_bits = tb
                f.init_state()
                f.run_paulis(ops, wdt)
                f.synchronize()
                t0 = time.perf_counter()
                f.run_paulis(ops, wdt)
                f.synchronize()
                tw[key] = time.perf_counter() - t0
            best = min(tw, key=tw.get)
            f.passes, width, f.tile_bits = cand[best]
            f.init_state()
            f.run_gates(capi.gates_from([wl.G(wl.H, top, [], 0.0)]), 6)
            f.run_paulis(ops, width)   # warm-up step, as in FP64
            f.synchronize()
            f.timing(True)
            f.timing_reset()
            t0 = time.perf_counter()
            nb = f.run_paulis(capi.pauli_ops(wl.trotter_step_ops(terms, dt, 1 << top) * (steps - 1)), width)
            f.synchronize()
            dt32 = time.perf_counter() - t0
            rep32 = f.timing_report()
            f.timing(False)
            fam32 = max(rep32, key=lambda k: rep32[k]["ms"])
            res.update({"fp32_width": None if f.passes == 3 else width, "fp32_mode": best,
                        "fp32_step_seconds_by_mode_width": {k: round(v, 4) for k, v in tw.items()},
                        "fp32_fused_seconds": round(dt32, 4), "fp32_blocks": nb,
                        "fp32_rotations_per_s": round(rot / dt32, 1),
                        "fp32_roofline": kernel_roofline(rep32, fam32, mfma_tflops=157.3),
                        "fp32_fidelity_vs_fp64": f.fidelity(e)})
            log(f"trotter fp32: {rot / dt32:.0f} rotations/s at {best}, fidelity {res['fp32_fidelity_vs_fp64']!r}")
        # the same controlled steps replayed from one capture, as the QPE below
        # applies them (after the fidelity above: this changes the state)
        e.graph_begin()
        e.run_paulis(ops)
        g = e.graph_end()
        e.graph_launch(g)
        e.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps - 1):
            e.graph_launch(g)
        e.synchronize()
        res["fp64_graph_replay_rotations_per_s"] = round(rot / (time.perf_counter() - t0), 1)
        # the whole QPE as specified (20 Trotter steps per power, 120,960 controlled
        # rotations), or with 2 steps per power when full_qpe is off
        qs = steps if full_qpe else 2
        e.init_state()
        t0 = time.perf_counter()
        h64 = wl.qpe_trotter(e, nsys, nreg, terms, dt, qs)
        t64 = time.perf_counter() - t0
        log(f"qpe fp64 ({qs} steps/power): {t64:.1f} s")
        q = {"steps_per_power": qs, "rotations": 96 * qs * 63, "fp64_seconds": round(t64, 2),
             "fp64_rotations_per_s": round(96 * qs * 63 / t64, 1),
             "fp64_mode_bin": max(range(64), key=lambda k: h64[k]), "fp64_hist_sum": sum(h64),
             "fp64_hist": [round(x, 12) for x in h64]}
        if f is not None:
            f.init_state()
            t0 = time.perf_counter()
            h32 = wl.qpe_trotter(f, nsys, nreg, terms, dt, qs, width=width)
            t32 = time.perf_counter() - t0
            diff = [abs(x - y) for x, y in zip(h32, h64)]
            q.update({"fp32_seconds": round(t32, 2), "fp32_rotations_per_s": round(96 * qs * 63 / t32, 1),
                      "hist_l1_error": sum(diff), "hist_linf_error": max(diff),
                      "fp32_mode_bin": max(range(64), key=lambda k: h32[k])})
            log(f"qpe fp32: {t32:.1f} s, L1 {sum(diff):.3g}, Linf {max(diff):.3g}")
        res["qpe_full" if full_qpe else "qpe_steps2"] = q
        if f is not None:
            f.close()
        if not no_cpu:
            res["cpu_baseline"] = cpu_qpe_c4()
            log(f"oracle qpe 10+4: {res['cpu_baseline']['value']:.0f} rotations/s")
        return res
    finally:
        e.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=198)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-qms", action="store_true", help="skip the QMS legs (profiling runs)")
    ap.add_argument("--no-fused", action="store_true", help="skip the FP32 fused leg (profiling runs)")
    ap.add_argument("--no-trotter", action="store_true", help="skip the Trotter + QPE leg (profiling runs)")
    ap.add_argument("--qms-batch-only", action="store_true",
                    help="only the 4096-chain QMS batches (profiling runs)")
    ap.add_argument("--short-qpe", action="store_true",
                    help="the whole QPE with 2 Trotter steps per power instead of 20 (profiling runs)")
    ap.add_argument("--gate-order", choices=("packed", "call"), default="packed",
                    help="FP64 passes: commuting gates packed (amplitudes within 1e-12 of call order) or call order")
    ap.add_argument("--tile-bits", type=int, default=-1,
                    help="FP64 tile size (0: one sweep per gate); default: autotuned")
    ap.add_argument("--no-tune-cache", action="store_true",
                    help="autotune every run instead of reusing the JSON cache's choices (suqa_amd/tune_cache.py)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {a.gpus}); the engine has no multi-GPU "
              f"sharding", file=sys.stderr)
        sys.exit(2)
    if a.qms_batch_only:
        print(json.dumps({"mode": "qms_batch_only", **qms_batch_only()}), flush=True)
        return
    from suqa_amd import capi, tune_cache, workloads as wl
    n = a.n
    layers = max(CIRCUIT_LAYERS, a.warmup + a.steps)
    gates, mq, mr = wl.random_circuit(n, layers, 7)
    arr = capi.gates_from(gates)
    log(f"bench: n={n} warmup={a.warmup} steps={a.steps} layers={layers}")
    eng = capi.Engine(n)
    peak_copy = hbm_copy_calibration(capi, eng)
    log(f"copy-kernel bandwidth {peak_copy:.0f} GB/s")
    tune = None
    eng.set_gate_order(a.gate_order == "packed")
    if a.tile_bits < 0:
        if a.no_tune_cache:
            best, tms = eng.autotune_tile_bits(capi.gates_from(gates[:5 * n]), (0, 8, 9, 10, 11, 12))
            tms, how = {str(k): round(v, 4) for k, v in tms.items()}, "off"
        else:
            best, tms, how = tune_cache.tile_bits(eng, capi.gates_from(gates[:5 * n]), (0, 8, 9, 10, 11, 12),
                                                  f"random-circuit-seed7-{a.gate_order}")
        tune = {"tile_bits": best, "ms_per_gate": tms, "cache": how}
        log(f"autotuned tile bits {best} (cache {how}): {tms}")
    else:
        eng.set_tile_bits(a.tile_bits)
    eng.init_state()
    for q in range(n):
        eng.h(q)
    eng.synchronize()
    ps0 = eng.pass_stats()
    per = n
    for s in range(a.warmup):
        eng.run_gates(arr, per, s * per)
    eng.synchronize()
    eng.timing(True)
    eng.timing_reset()
    t0 = time.perf_counter()
    for s in range(a.warmup, a.warmup + a.steps):
        eng.run_gates(arr, per, s * per)
        if (s - a.warmup) % 50 == 49:
            log(f"  layer {s + 1}/{layers}")
    eng.synchronize()
    t1 = time.perf_counter()
    rep = eng.timing_report()
    eng.timing(False)
    ps1 = eng.pass_stats()
    wall = t1 - t0
    # the rest of the circuit, untimed: the checksum and outcomes are the whole circuit's
    for s in range(a.warmup + a.steps, layers):
        eng.run_gates(arr, per, s * per)
    eng.synchronize()
    checksum = eng.checksum()
    alg = sum(wl.gate_bytes(g, n) for g in gates[a.warmup * per:(a.warmup + a.steps) * per])
    value = alg / wall / 1e9
    fam = max(rep, key=lambda k: rep[k]["ms"])
    achieved = rep[fam]["bytes"] / (rep[fam]["ms"] * 1e-3) / 1e9
    alg_per_launch = rep[fam]["bytes"] / rep[fam]["launches"]
    traffic = alg_per_launch * PMC_TRAFFIC_RATIO[fam] if fam in PMC_TRAFFIC_RATIO else None
    hbm_all = sum(v["bytes"] for v in rep.values()) / wall / 1e9
    eng_tile_bits = eng.tile_bits()
    npass = ps1["passes"] - ps0["passes"]
    pass_info = {"passes": npass, "gates_in_passes": ps1["pass_gates"] - ps0["pass_gates"],
                 "single_sweeps": ps1["single_sweeps"] - ps0["single_sweeps"],
                 "gates_per_pass": round((ps1["pass_gates"] - ps0["pass_gates"]) / npass, 2) if npass else None}
    f32, fres = (None, None) if a.no_fused else fused_leg(capi, wl, eng, gates, n, layers, a.no_cpu,
                                                          not a.no_tune_cache)
    # the final measurements of the workload (not in the timed region)
    tm = time.perf_counter()
    outcomes = [eng.measure(q, r) for q, r in zip(mq, mr)]
    tm = time.perf_counter() - tm
    norm = eng.vnorm()
    checksum_final = eng.checksum()
    order_cmp = gate_order_sample(eng, arr, n, a.gate_order)
    eng.close()
    if f32 is not None:
        fres["measure_outcomes"] = [f32.measure(q, r) for q, r in zip(mq, mr)]
        fres["outcomes_match_fp64"] = fres["measure_outcomes"] == outcomes
        fres["speedup_vs_fp64_per_layer"] = round((wall / max(a.steps, 1)) / (fres["seconds"] / layers), 2)
        f32.close()
    cpu = None if a.no_cpu else cpu_baseline(min(n, 27))
    qres = {} if a.no_qms else qms_legs(a.no_cpu)
    tres = None if a.no_trotter else trotter_leg(capi, wl, a.no_fused, not a.short_qpe, a.no_cpu)
    fams = {k: dict(launches=v["launches"], ms=round(v["ms"], 3),
                    gbs=round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)) for k, v in rep.items()}
    out = {
        "metric": "QMS Metropolis steps/s at fixed FP64 error; gate-sweep GB/s vs HBM roof",
        "value": round(value, 2),
        "unit": "GB/s (gate-equivalent: each gate counted as its own sweep)",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / max(a.steps, 1) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: pcg32 seed-7 random circuit",
        "config": {"workload": "FP64 gate sweeps, random circuit {H,X,RX,RZ,U1,CX,MCX3}",
                   "qubits": n, "gates_per_step": per, "layers_total": layers, "seed": 7,
                   "tile_bits": eng_tile_bits, "gate_order": a.gate_order},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else round(traffic), "kernel": fam,
                     "alg_bytes_per_launch": round(alg_per_launch),
                     "avg_launch_ms": round(rep[fam]["ms"] / rep[fam]["launches"], 4)},
        "cpu_baseline": cpu,
        "extra": dict({"c2_hbm": {"gbs": round(hbm_all, 1), "frac": round(hbm_all / HBM_PEAK_GBS, 4),
                                  "what": "HBM bytes all kernels of the timed layers must move / wall time"},
                       "checksum": {"layers": layers, "sum_re": checksum[0], "sum_im": checksum[1],
                                    "sum_index_prob": checksum[2], "after_measurements": list(checksum_final)},
                       "gate_order_sample": order_cmp,
                       "copy_kernel_gbs": round(peak_copy, 1), "kernel_families": fams,
                       "tile_autotune": tune, "passes": pass_info,
                       "measure_outcomes": outcomes, "measure_seconds": round(tm, 4), "final_norm": norm, "fused": fres,
                       "trotter_qpe": tres}, **qres),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
