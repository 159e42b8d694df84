"""Benchmark: plain FP64 gate sweeps of the 30-qubit random-circuit workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 30] [--no-cpu]

Workload: 30 qubits, complex FP64 (16 GiB in HBM), |0> then an H layer, then
layers of 30 random gates from {H, X, RX, RZ, U1, CX, MCX(3)} drawn with
pcg32 seed 7 (suqa_amd/workloads.py).  One step = one layer (30 gates).
The FP64 engine applies the gates in passes: runs of consecutive gates whose
flipped qubits fit in one tile are applied to LDS-resident tiles in one HBM
round trip, bit-identical to one sweep per gate.  The tile size is autotuned
first (each candidate, and one sweep per gate, timed on 5 layers; the state
is then re-initialised).  W warmup layers are applied untimed, then exactly
K layers are timed between stream synchronizations; the default W=2, K=198
applies all 200 layers.  The 8 final measure_qbit calls run after the timed
region.

value = algorithmic bytes of the timed layers / wall time (GB/s); a gate's
algorithmic bytes are 32 B per amplitude it modifies (H/X/RX/RZ 32N,
U1/CX 16N, MCX3 4N), as if each gate were its own sweep, so value exceeds
the HBM roof when passes hold several gates.  roofline: the dominant kernel
(the tile pass), the HBM bytes it must move (every amplitude of the tiles it
touches read and written once, 32 B) over the sum of its per-launch HIP-event
durations, against the 8 TB/s HBM3E peak.  cpu_baseline: the oracle (a
single-threaded restatement of the reference CPU path) on the first layer at
27 qubits.

The QMS half of the metric is reported under "extra":
  qms_plain  -- the reference's frustrated-triangle QMS (beta 1, 1000 steps,
                4-qubit energy register, E in [-1,3], pcg32 seed 42) on the
                plain FP64 GPU path (suqa_qms_run), Metropolis steps/s, with
                its accept/reject sequence checked against the oracle;
  qms_plain_two_registers -- the same chain with E_old in its own register
                (two 4-qubit registers, 12 qubits), checked against the
                oracle's two-register chain;
  qms_batch  -- 4096 independent chains (register widths 3..6 drawn with
                pcg32 seed 5, betas 0.1..5.0, seeds 1000+i, 500 steps) in a
                few launches, chain-steps/s, a sample of chains checked
                against the oracle; every batch (and the single chain below)
                is run once untimed before its timed run;
  qms_batch_fp32 -- the same 4096 chains with FP32 amplitudes and FP64
                probabilities: chain-steps/s, chains whose return codes equal
                the FP64 run's, and per-beta <E> (8 bins of 512 chains)
                against the FP64 run in units of its standard error;
  qms_batch_two_registers -- the same kind of batch in the two-register
                layout (3 + 2r + 1 qubits per chain) for r in {3, 4}, FP64
                and FP32, 4 sample chains checked against the oracle;
  qms_single_chain_device -- the c1 chain as a batch of one (one workgroup
                 runs it on the device), its steps/s and rets against qms_plain;
  cpu_baseline_qms -- the oracle on the same single chain, 1 core; for the
                 batch, a 64-chain sample on one oracle process per chain over
                 the host cores (at most 16), its rets against the batch's.

The Trotter + phase-estimation workload under "extra.trotter_qpe": 24 system
qubits + a 6-qubit register (30 qubits), 96 Pauli strings of weight 2-4
(pcg32 seed 11), dt 0.05, 20 Trotter steps per power.  Timed sample: H on
the top register qubit, then its 20 controlled Trotter steps (1920
controlled rotations), in FP64 (gate passes) and FP32 fused form, with the
fidelity between them (FP32 block width timed at 4, 5, 6 on one step); the
full QPE applies 63 such step groups (120,960 rotations), projected from
the sample.  Then the whole QPE at full size with 2 Trotter steps per power
(12,096 rotations) in both precisions: the 64-bin register histogram's L1
and Linf error of FP32 against FP64; with --full-qpe also the whole QPE with
20 steps per power (120,960 rotations, about 3 minutes) the same way.

and the precision pattern under "extra.fused": the same circuit on an FP32
state with gates fused into K-qubit blocks (K autotuned over 4..6 on the
first 10 layers) applied by FP32 MFMA, timed over all layers from |0> + H,
with its fidelity |<psi64|psi32>|^2 against the FP64 run and the 8 final
measurement outcomes compared.
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
    return dict(value=nbytes / dt / 1e9, unit="GB/s", cores=1, kind="port",
                sample=f"first layer (30 gates) of the seed-7 circuit at {n} qubits, "
                       f"{dt:.1f} s on 1 host core (oracle/suqa_oracle.c, gcc -O2)")


def _oracle_chain(args):
    """One chain on the oracle, in a worker process: (beta, r, seed, steps, emin, emax) -> rets."""
    from oracle import pyoracle as po
    beta, r, seed, steps, emin, emax = args
    _, rets = po.Qms(beta, r, seed, emin, emax, 10, 100).run(steps, 1, 100)
    return rets


def oracle_chains_all_cores(chains, steps, emin, emax):
    """The oracle on a sample of chains, one single-threaded process per chain over
    the host cores the box gives a GPU (at most 16): (chain-steps/s, workers, rets)."""
    import concurrent.futures as cf
    import multiprocessing as mp
    workers = max(1, min(16, os.cpu_count() or 1, len(chains)))
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        list(ex.map(_oracle_chain, [(1.0, 2, 1, 1, emin, emax)] * workers))   # start every worker
        t0 = time.perf_counter()
        rets = list(ex.map(_oracle_chain, [(b, r, s, steps, emin, emax) for (b, r, s) in chains]))
        dt = time.perf_counter() - t0
    return len(chains) * steps / dt, workers, rets


def po_two_register_chain(beta, r, seed, steps, emin, emax):
    """One two-register chain on the oracle: (accepted, rets)."""
    from oracle import pyoracle as po
    return po.Qms(beta, r, seed, emin, emax, 10, 100, two_registers=True).run(steps, 1, 100)


def qms_batch_only():
    """The 4096-chain batches of the QMS legs alone, FP64 then FP32, one untimed
    run each first (a profiling mode: no gate sweeps, no plain chain)."""
    from suqa_amd import qms, workloads as wl
    emin, emax = -1.0, 3.0
    rg = wl.Pcg32(5)
    nch, bsteps = 4096, 500
    chains = [(0.1 + 4.9 * i / (nch - 1), 3 + rg.randint(0, 4), 1000 + i) for i in range(nch)]
    out = {}
    for name, fp32 in (("qms_batch", False), ("qms_batch_fp32", True)):
        qms.batch_run(chains, bsteps, 1, 100, emin, emax, 10, 100, fp32=fp32)
        b = qms.batch_run(chains, bsteps, 1, 100, emin, emax, 10, 100, fp32=fp32)
        out[name] = round(nch * bsteps / b["seconds"], 0)
        log(f"{name}: {out[name]:.0f} chain-steps/s")
    return out


def qms_legs(no_cpu):
    """Plain FP64 GPU QMS (c1), batched chains (c5), their oracle checks and CPU rates."""
    from suqa_amd import qms, workloads as wl
    beta, steps, r, seed, emin, emax = 1.0, 1000, 4, 42, -1.0, 3.0
    plain = qms.run(beta, steps, 1, r, seed, emin, emax, 10, 100, 100)
    res = {"qms_plain": {"steps_per_s": round(steps / plain["seconds"], 1), "steps": steps,
                         "accepted": int(plain["accepted"]), "seconds": round(plain["seconds"], 4),
                         "config": "beta=1.0, 1000 steps, r=4 (8 qubits), E in [-1,3], seed 42, reset_each 1"}}
    log(f"qms plain: {steps / plain['seconds']:.1f} steps/s")
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
    rg = wl.Pcg32(5)
    nch, bsteps = 4096, 500
    chains = [(0.1 + 4.9 * i / (nch - 1), 3 + rg.randint(0, 4), 1000 + i) for i in range(nch)]
    # each batch is run once untimed first: the GPU's clocks after the previous
    # leg (a different kernel, another load) would otherwise set the figure
    qms.batch_run(chains, bsteps, 1, 100, emin, emax, 10, 100)
    b = qms.batch_run(chains, bsteps, 1, 100, emin, emax, 10, 100)
    rate = nch * bsteps / b["seconds"]
    res["qms_batch"] = {"chain_steps_per_s": round(rate, 0), "chains": nch, "steps": bsteps,
                        "kernel_seconds": round(b["seconds"], 4),
                        "speedup_vs_plain_fp64_gpu": round(rate / (steps / plain["seconds"]), 1),
                        "mean_acceptance": round(float(b["accepted"].mean()) / bsteps, 4), "dtype": "f64"}
    log(f"qms batch: {rate:.0f} chain-steps/s")
    # the same chains with FP32 amplitudes (FP64 probabilities): rate and per-beta
    # thermal averages against the FP64 run
    qms.batch_run(chains, bsteps, 1, 100, emin, emax, 10, 100, fp32=True)
    b32 = qms.batch_run(chains, bsteps, 1, 100, emin, emax, 10, 100, fp32=True)
    rate32 = nch * bsteps / b32["seconds"]
    same = int(sum(1 for c in range(nch) if np.array_equal(b32["rets"][c], b["rets"][c])))
    bins = []
    for k in range(8):   # 8 beta bins of 512 chains
        idx = range(k * nch // 8, (k + 1) * nch // 8)

        def chain_means(out):
            return np.array([out["E"][c][:int(out["n_meas"][c])].mean() for c in idx if int(out["n_meas"][c])])
        m64, m32 = chain_means(b), chain_means(b32)
        se = m64.std(ddof=1) / np.sqrt(len(m64))
        bins.append({"beta_min": round(chains[idx[0]][0], 3), "beta_max": round(chains[idx[-1]][0], 3),
                     "E_fp64": round(float(m64.mean()), 6), "E_fp32": round(float(m32.mean()), 6),
                     "fp64_stderr": round(float(se), 6), "diff_in_stderr": round(float(abs(m32.mean() - m64.mean()) / se), 3)})
    res["qms_batch_fp32"] = {"chain_steps_per_s": round(rate32, 0), "kernel_seconds": round(b32["seconds"], 4),
                             "speedup_vs_plain_fp64_gpu": round(rate32 / (steps / plain["seconds"]), 1),
                             "chains_identical_to_fp64": same, "per_beta_E": bins,
                             "dtype": "f32 amplitudes, f64 probabilities"}
    log(f"qms batch fp32: {rate32:.0f} chain-steps/s, {same}/{nch} chains identical to FP64")
    # the batch in the two-register layout (E_old in its own register, 3 + 2r + 1
    # qubits per chain) for the widths whose FP64 state fits a workgroup's LDS
    rg2 = wl.Pcg32(5)
    chains2 = [(0.1 + 4.9 * i / (nch - 1), 3 + rg2.randint(0, 2), 1000 + i) for i in range(nch)]
    qms.batch_run(chains2, bsteps, 1, 100, emin, emax, 10, 100, two_registers=True)
    b2 = qms.batch_run(chains2, bsteps, 1, 100, emin, emax, 10, 100, two_registers=True)
    qms.batch_run(chains2, bsteps, 1, 100, emin, emax, 10, 100, two_registers=True, fp32=True)
    b2f = qms.batch_run(chains2, bsteps, 1, 100, emin, emax, 10, 100, two_registers=True, fp32=True)
    res["qms_batch_two_registers"] = {
        "chain_steps_per_s": round(nch * bsteps / b2["seconds"], 0),
        "fp32_chain_steps_per_s": round(nch * bsteps / b2f["seconds"], 0),
        "fp32_chains_identical_to_fp64": int(sum(1 for c in range(nch) if np.array_equal(b2f["rets"][c], b2["rets"][c]))),
        "chains": nch, "steps": bsteps, "register_widths": [3, 4], "qubits_per_chain": [10, 12],
        "config": "as qms_batch, E_old in its own register; widths 5 and 6 (14 and 16 qubits) exceed a "
                  "workgroup's LDS in FP64"}
    log(f"qms batch, two registers: {nch * bsteps / b2['seconds']:.0f} chain-steps/s "
        f"(fp32 {nch * bsteps / b2f['seconds']:.0f})")
    if no_cpu:
        return res
    okb2 = True
    for c in range(0, nch, nch // 4):
        bb, rr, ss = chains2[c]
        _, orets = po_two_register_chain(bb, rr, ss, bsteps, emin, emax)
        okb2 &= list(b2["rets"][c]) == orets
    res["qms_batch_two_registers"]["sample_chains_match_oracle"] = bool(okb2)
    from oracle import pyoracle as po
    q2 = po.Qms(beta, r, seed, emin, emax, 10, 100, two_registers=True)
    _, rets2 = q2.run(steps, 1, 100)
    res["qms_plain_two_registers"]["rets_match_oracle"] = bool(list(two["rets"]) == rets2)
    q = po.Qms(beta, r, seed, emin, emax, 10, 100)
    t0 = time.perf_counter()
    acc, rets = q.run(steps, 1, 100)
    dt = time.perf_counter() - t0
    res["qms_plain"]["rets_match_oracle"] = bool(list(plain["rets"]) == rets)
    res["cpu_baseline_qms"] = {"value": round(steps / dt, 1), "unit": "steps/s", "cores": 1, "kind": "port",
                               "sample": "the same 1000-step chain on the oracle (oracle/suqa_oracle.c, 1 core)"}
    ok, t0, nchk = True, time.perf_counter(), 12
    for c in range(0, nch, nch // nchk):
        bb, rr, ss = chains[c]
        qq = po.Qms(bb, rr, ss, emin, emax, 10, 100)
        _, orets = qq.run(bsteps, 1, 100)
        ok &= list(b["rets"][c]) == orets
    dt = time.perf_counter() - t0
    res["qms_batch"]["sample_chains_match_oracle"] = bool(ok)
    res["qms_batch"]["oracle_chain_steps_per_s_1core"] = round(nchk * bsteps / dt, 1)
    # the CPU baseline as the workload defines it: one oracle process per chain over
    # the host cores, on a sample of 64 chains spread over the beta grid and widths
    sample = list(range(0, nch, nch // 64))
    try:
        rate_cpu, workers, orets = oracle_chains_all_cores([chains[c] for c in sample], bsteps, emin, emax)
    except Exception as ex:   # worker processes unavailable: keep the 1-core figure
        res["qms_batch"]["oracle_all_cores_error"] = repr(ex)
        return res
    res["qms_batch"]["oracle_chain_steps_per_s_all_cores"] = round(rate_cpu, 1)
    res["qms_batch"]["oracle_cores"] = workers
    res["qms_batch"]["speedup_vs_oracle_all_cores"] = round(rate / rate_cpu, 1)
    res["qms_batch"]["sample64_chains_match_oracle"] = bool(all(list(b["rets"][c]) == o for c, o in zip(sample, orets)))
    log(f"oracle, {len(sample)} chains on {workers} processes: {rate_cpu:.0f} chain-steps/s")
    return res


def fused_leg(capi, wl, eng, gates, n, layers):
    """FP32 fused-block run of the whole circuit; eng holds the FP64 state after all layers."""
    f32 = capi.F32Engine(n)
    try:
        sample = capi.gates_from(gates[:10 * n])
        best, tms = f32.autotune(sample)
        log(f"fused: autotuned width {best}, passes {f32.passes} ({tms})")
        arr = capi.gates_from(gates[:layers * n])
        t0 = time.perf_counter()
        plan = capi.fuse_plan(arr, n, best)
        t_plan = time.perf_counter() - t0
        f32.init_state()
        f32.run_gates(capi.gates_from([wl.G(wl.H, q, [], 0.0) for q in range(n)]), best)
        f32.synchronize()
        t0 = time.perf_counter()
        nb = f32.run_gates(arr, best)
        f32.synchronize()
        dt = time.perf_counter() - t0
        fid = f32.fidelity(eng)
        norm = f32.vnorm()
        log(f"fused: {nb} blocks for {layers * n} gates in {dt:.3f} s, fidelity {fid!r}")
        amp_bytes = 16 * (1 << n)     # an FP32 complex amplitude read and written per block
        return f32, {
            "width": best, "mode": ["blocks", "passes", "sparse"][f32.passes],
            "ms_per_gate_by_mode_width": {k: (None if v < 0 else round(v, 4)) for k, v in tms.items()},
            "gates": layers * n, "blocks": nb, "gates_per_block": round(layers * n / nb, 2),
            "seconds": round(dt, 4), "layers_per_s": round(layers / dt, 3),
            "ms_per_layer": round(dt / layers * 1e3, 3), "host_fusion_seconds": round(t_plan, 4),
            "block_kernel_gbs": round(nb * amp_bytes / dt / 1e9, 1),
            "fidelity_vs_fp64": fid, "norm": norm, "dtype": "f32 state, f32 MFMA, f64 reductions"}
    except Exception:
        f32.close()
        raise


def trotter_leg(capi, wl, no_fused, full_qpe=False):
    """Controlled Trotter steps of the 24 + 6 qubit QPE: FP64 passes and FP32 fused, one sample."""
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
        t0 = time.perf_counter()
        for _ in range(steps - 1):
            e.run_paulis(ops)
        e.synchronize()
        dt64 = time.perf_counter() - t0
        ps1 = e.pass_stats()
        rot = 96 * (steps - 1)
        res = {"qubits": n, "terms": 96, "dt": dt, "steps_per_power": steps, "sample_rotations": rot,
               "fp64_seconds": round(dt64, 4), "fp64_rotations_per_s": round(rot / dt64, 1),
               "fp64_alg_gbs": round(rot * 16 * (1 << n) / dt64 / 1e9, 1),
               "fp64_gates_per_pass": round((ps1["pass_gates"] - ps0["pass_gates"]) /
                                            max(1, ps1["passes"] - ps0["passes"]), 2),
               "fp64_full_qpe_projected_s": round(dt64 / rot * 96 * steps * 63, 1)}
        log(f"trotter fp64: {rot / dt64:.0f} rotations/s")
        if not no_fused:
            f = capi.F32Engine(n)
            try:
                # mode and block width: one controlled Trotter step timed at each
                tw = {}
                modes = {"blocks": 0, "passes": 1, "sparse": 2}
                for mode in modes:
                    f.passes = modes[mode]
                    for wdt in (4, 5, 6):
                        f.init_state()
                        f.run_paulis(ops, wdt)
                        f.synchronize()
                        t0 = time.perf_counter()
                        f.run_paulis(ops, wdt)
                        f.synchronize()
                        tw[f"{mode}{wdt}"] = time.perf_counter() - t0
                best = min(tw, key=tw.get)
                f.passes = modes[best[:-1]]
                width = int(best[-1])
                f.init_state()
                f.run_gates(capi.gates_from([wl.G(wl.H, top, [], 0.0)]), 6)
                f.run_paulis(ops, width)   # warm-up step, as in FP64
                f.synchronize()
                t0 = time.perf_counter()
                nb = f.run_paulis(capi.pauli_ops(wl.trotter_step_ops(terms, dt, 1 << top) * (steps - 1)), width)
                f.synchronize()
                dt32 = time.perf_counter() - t0
                res.update({"fp32_width": width, "fp32_mode": best[:-1],
                            "fp32_step_seconds_by_mode_width": {k: round(v, 4) for k, v in tw.items()},
                            "fp32_fused_seconds": round(dt32, 4), "fp32_blocks": nb,
                            "fp32_rotations_per_s": round(rot / dt32, 1),
                            "fp32_fidelity_vs_fp64": f.fidelity(e)})
                log(f"trotter fp32: {rot / dt32:.0f} rotations/s at {best}, "
                    f"fidelity {res['fp32_fidelity_vs_fp64']!r}")
                # the whole QPE at full size with 2 Trotter steps per power: histograms
                e.init_state()
                t0 = time.perf_counter()
                h64 = wl.qpe_trotter(e, nsys, nreg, terms, dt, 2)
                t64 = time.perf_counter() - t0
                f.init_state()
                t0 = time.perf_counter()
                h32 = wl.qpe_trotter(f, nsys, nreg, terms, dt, 2, width=width)
                t32 = time.perf_counter() - t0
                diff = [abs(x - y) for x, y in zip(h32, h64)]
                res["qpe_steps2"] = {"rotations": 96 * 2 * 63, "fp64_seconds": round(t64, 2),
                                     "fp32_seconds": round(t32, 2), "hist_l1_error": sum(diff),
                                     "hist_linf_error": max(diff), "fp64_mode": max(range(64), key=lambda k: h64[k]),
                                     "fp32_mode": max(range(64), key=lambda k: h32[k]), "fp64_hist_sum": sum(h64)}
                log(f"qpe (2 steps/power): fp64 {t64:.1f} s, fp32 {t32:.1f} s, L1 {sum(diff):.3g}, Linf {max(diff):.3g}")
                if full_qpe:   # the whole QPE as specified: 20 Trotter steps per power, 120,960 rotations
                    e.init_state()
                    t0 = time.perf_counter()
                    h64 = wl.qpe_trotter(e, nsys, nreg, terms, dt, steps)
                    t64 = time.perf_counter() - t0
                    log(f"full qpe fp64: {t64:.1f} s")
                    f.init_state()
                    t0 = time.perf_counter()
                    h32 = wl.qpe_trotter(f, nsys, nreg, terms, dt, steps, width=width)
                    t32 = time.perf_counter() - t0
                    diff = [abs(x - y) for x, y in zip(h32, h64)]
                    res["qpe_full"] = {"rotations": 96 * steps * 63, "fp64_seconds": round(t64, 2),
                                       "fp32_seconds": round(t32, 2), "fp32_mode": best[:-1],
                                       "hist_l1_error": sum(diff), "hist_linf_error": max(diff),
                                       "fp64_mode": max(range(64), key=lambda k: h64[k]),
                                       "fp32_mode_bin": max(range(64), key=lambda k: h32[k]),
                                       "fp64_hist": [round(x, 12) for x in h64]}
                    log(f"full qpe: fp64 {t64:.1f} s, fp32 {t32:.1f} s, L1 {sum(diff):.3g}, Linf {max(diff):.3g}")
            finally:
                f.close()
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
                    help="only the 4096-chain QMS batches, FP64 and FP32 (profiling runs)")
    ap.add_argument("--full-qpe", action="store_true",
                    help="also run the whole 24 + 6 qubit QPE (20 Trotter steps per power) in FP64 and FP32")
    ap.add_argument("--tile-bits", type=int, default=-1,
                    help="FP64 tile size (0: one sweep per gate); default: autotuned")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {a.gpus}); the engine has no multi-GPU "
              f"sharding", file=sys.stderr)
        sys.exit(2)
    if a.qms_batch_only:
        print(json.dumps({"mode": "qms_batch_only", **qms_batch_only()}), flush=True)
        return
    from suqa_amd import capi, workloads as wl
    n = a.n
    layers = a.warmup + a.steps
    gates, mq, mr = wl.random_circuit(n, max(layers, 1), 7)
    arr = capi.gates_from(gates)
    log(f"bench: n={n} warmup={a.warmup} steps={a.steps}")
    eng = capi.Engine(n)
    peak_copy = hbm_copy_calibration(capi, eng)
    log(f"copy-kernel bandwidth {peak_copy:.0f} GB/s")
    tune = None
    if a.tile_bits < 0:
        best, tms = eng.autotune_tile_bits(capi.gates_from(gates[:5 * n]), (0, 8, 9, 10, 11, 12))
        tune = {"tile_bits": best, "ms_per_gate": {str(k): round(v, 4) for k, v in tms.items()}}
        log(f"autotuned tile bits {best}: {tms}")
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
    for s in range(a.warmup, layers):
        eng.run_gates(arr, per, s * per)
        if (s - a.warmup) % 50 == 49:
            log(f"  layer {s + 1}/{layers}")
    eng.synchronize()
    t1 = time.perf_counter()
    rep = eng.timing_report()
    eng.timing(False)
    ps1 = eng.pass_stats()
    wall = t1 - t0
    alg = sum(wl.gate_bytes(g, n) for g in gates[a.warmup * per:layers * per])
    value = alg / wall / 1e9
    fam = max(rep, key=lambda k: rep[k]["ms"])
    achieved = rep[fam]["bytes"] / (rep[fam]["ms"] * 1e-3) / 1e9
    alg_per_launch = rep[fam]["bytes"] / rep[fam]["launches"]
    traffic = alg_per_launch * PMC_TRAFFIC_RATIO[fam] if fam in PMC_TRAFFIC_RATIO else None
    eng_tile_bits = eng.tile_bits()
    npass = ps1["passes"] - ps0["passes"]
    pass_info = {"passes": npass, "gates_in_passes": ps1["pass_gates"] - ps0["pass_gates"],
                 "single_sweeps": ps1["single_sweeps"] - ps0["single_sweeps"],
                 "gates_per_pass": round((ps1["pass_gates"] - ps0["pass_gates"]) / npass, 2) if npass else None}
    f32, fres = (None, None) if a.no_fused else fused_leg(capi, wl, eng, gates, n, layers)
    # the final measurements of the workload (not in the timed region)
    tm = time.perf_counter()
    outcomes = [eng.measure(q, r) for q, r in zip(mq, mr)]
    tm = time.perf_counter() - tm
    norm = eng.vnorm()
    eng.close()
    if f32 is not None:
        fres["measure_outcomes"] = [f32.measure(q, r) for q, r in zip(mq, mr)]
        fres["outcomes_match_fp64"] = fres["measure_outcomes"] == outcomes
        fres["speedup_vs_fp64_per_layer"] = round((wall / max(a.steps, 1)) / (fres["seconds"] / layers), 2)
        f32.close()
    cpu = None if a.no_cpu else cpu_baseline(min(n, 27))
    qres = {} if a.no_qms else qms_legs(a.no_cpu)
    tres = None if a.no_trotter else trotter_leg(capi, wl, a.no_fused, a.full_qpe)
    fams = {k: dict(launches=v["launches"], ms=round(v["ms"], 3),
                    gbs=round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)) for k, v in rep.items()}
    out = {
        "metric": "QMS Metropolis steps/s at fixed FP64 error; gate-sweep GB/s vs HBM roof",
        "value": round(value, 2),
        "unit": "GB/s",
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
                   "tile_bits": eng_tile_bits},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else round(traffic), "kernel": fam,
                     "alg_bytes_per_launch": round(alg_per_launch),
                     "avg_launch_ms": round(rep[fam]["ms"] / rep[fam]["launches"], 4)},
        "cpu_baseline": cpu,
        "extra": dict({"copy_kernel_gbs": round(peak_copy, 1), "kernel_families": fams,
                       "tile_autotune": tune, "passes": pass_info,
                       "measure_outcomes": outcomes, "measure_seconds": round(tm, 4), "final_norm": norm, "fused": fres,
                       "trotter_qpe": tres}, **qres),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
