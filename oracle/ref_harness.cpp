// ref_harness.cpp -- writes golden vectors by driving the reference SUQA
// engine (compiled from the reference sources into oracle/_ref by
// oracle/ref.mk).  Test infrastructure only: it is never part of the product
// and never runs on the GPU host.  It contains no reference source; it calls
// the reference API the way the reference drivers do (qms.cu, evo.cu).
//
// Modes (output goes to the file named by the last argument, as JSON):
//   pcg     <out>                      -- pcg32 sequences for several seeds
//   gates   <out>                      -- per-API golden amplitudes at n=5
//   qms     <beta> <steps> <reset_each> <r> <seed> <ene_min> <ene_max>
//           <pe_steps> <therm> <max_rev> <record_rev> <out>
//   circuit <n> <layers> <seed> <out>  -- random-circuit family (see mode_circuit)
//   qsa     <beta> <sampling> <ene qbits> <szegedy qbits> <seed> <ene_min> <ene_max>
//           <ene_min_szegedy> <ene_max_szegedy> <pe_steps> <annealing sequences>
//           <max annealing steps> <out>    -- the QSA loop of qsa.cu, bounded
//   circuit_summary <n> <layers> <seed> <out>
//                                      -- the same circuit: checksum and 64 sample
//                                         amplitudes before and after the measurements
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <cmath>
#include <stdexcept>
#include "Rand.hpp"
#include "suqa.cuh"
#include "system.cuh"
#include "qms.cuh"
#include "qsa.cuh"

static FILE* g_out = nullptr;

static void dump_state(const char* key) {
    fprintf(g_out, "\"%s\":{\"re\":[", key);
    for (uint i = 0; i < suqa::state.size(); ++i)
        fprintf(g_out, "%s\"%a\"", i ? "," : "", suqa::state.data_re[i]);
    fprintf(g_out, "],\"im\":[");
    for (uint i = 0; i < suqa::state.size(); ++i)
        fprintf(g_out, "%s\"%a\"", i ? "," : "", suqa::state.data_im[i]);
    fprintf(g_out, "]}");
}

// ------------------------------------------------------------------ pcg
static void mode_pcg() {
    const unsigned long long seeds[] = {1ULL, 7ULL, 11ULL, 42ULL, 12345ULL, 0xdeadbeefcafeULL};
    fprintf(g_out, "{\"cases\":[");
    bool first = true;
    for (auto sd : seeds) {
        pcg32 r;
        r.seed(sd);
        fprintf(g_out, "%s{\"seed\":%llu,\"u32\":[", first ? "" : ",", sd);
        first = false;
        for (int k = 0; k < 64; ++k) fprintf(g_out, "%s%u", k ? "," : "", r.nextUInt());
        pcg w;
        w.set_seed(sd);
        fprintf(g_out, "],\"doub\":[");
        for (int k = 0; k < 32; ++k) fprintf(g_out, "%s\"%a\"", k ? "," : "", w.doub());
        fprintf(g_out, "],\"randint_0_30\":[");
        for (int k = 0; k < 32; ++k) fprintf(g_out, "%s%d", k ? "," : "", w.randint(0, 30));
        fprintf(g_out, "],\"state_after\":\"%llu\"}", (unsigned long long)w.get_state().state);
    }
    fprintf(g_out, "]}\n");
}

// ------------------------------------------------------------------ gates
// A fixed preparation that leaves every amplitude non-zero and complex.
static void prep(uint n) {
    suqa::init_state();
    for (uint q = 0; q < n; ++q) suqa::apply_h(q);
    for (uint q = 0; q < n; ++q) suqa::apply_pauli_TP_rotation({q}, {PAULI_Y}, 0.31 * (q + 1));
    for (uint q = 0; q < n; ++q) suqa::apply_u1(q, 0.57 * (q + 1));
    suqa::apply_cx(0, 2);
    suqa::apply_cx(3, 1);
    suqa::apply_pauli_TP_rotation({1, 4}, {PAULI_X, PAULI_Z}, 0.77);
    suqa::apply_pauli_TP_rotation({0, 3}, {PAULI_Y, PAULI_Y}, -0.41);
}

struct Case {
    std::string op;
    std::vector<uint> qs;     // qubits (controls first, target last where relevant)
    std::vector<uint> aux;    // q_mask values / pauli types
    double val;               // phase or rdoub
    std::vector<uint> gc;     // gc_mask qubits
};

static bool g_first_case = true;

static void run_case(const Case& c, uint n) {
    prep(n);
    suqa::activate_gc_mask(c.gc);
    bool err = false;
    double out = 0.0;
    uint outc = 0;
    try {
        const auto& q = c.qs;
        if (c.op == "x") suqa::apply_x(q[0]);
        else if (c.op == "y") suqa::apply_y(q[0]);
        else if (c.op == "z") suqa::apply_z(q[0]);
        else if (c.op == "h") suqa::apply_h(q[0]);
        else if (c.op == "s") suqa::apply_s(q[0]);
        else if (c.op == "t") suqa::apply_t(q[0]);
        else if (c.op == "tdg") suqa::apply_tdg(q[0]);
        else if (c.op == "u1") suqa::apply_u1(q[0], c.aux.empty() ? 1U : c.aux[0], c.val);
        else if (c.op == "cx") suqa::apply_cx(q[0], q[1], c.aux.empty() ? 1U : c.aux[0]);
        else if (c.op == "mcx") {
            bmReg ctr(q.begin(), q.end() - 1);
            if (c.aux.empty()) suqa::apply_mcx(ctr, q.back());
            else suqa::apply_mcx(ctr, c.aux, q.back());
        } else if (c.op == "cu1") suqa::apply_cu1(q[0], q[1], c.val, c.aux.empty() ? 1U : c.aux[0]);
        else if (c.op == "mcu1") {
            bmReg ctr(q.begin(), q.end() - 1);
            if (c.aux.empty()) suqa::apply_mcu1(ctr, q.back(), c.val);
            else suqa::apply_mcu1(ctr, c.aux, q.back(), c.val);
        } else if (c.op == "swap") suqa::apply_swap(q[0], q[1]);
        else if (c.op == "pauli") suqa::apply_pauli_TP_rotation(q, c.aux, c.val);
        else if (c.op == "qft") suqa::apply_qft(q);
        else if (c.op == "qft_inverse") suqa::apply_qft_inverse(q);
        else if (c.op == "measure") suqa::measure_qbit(q[0], outc, c.val);
        else if (c.op == "reset") suqa::apply_reset(q[0], c.val);
        else if (c.op == "prob_filter") suqa::prob_filter(q, c.aux, out);
        else if (c.op == "vnorm") out = suqa::vnorm();
        else if (c.op == "vnormalize") suqa::vnormalize();
        else throw std::runtime_error("unknown op");
    } catch (const std::runtime_error&) {
        err = true;
    }
    suqa::deactivate_gc_mask(c.gc);
    fprintf(g_out, "%s{\"op\":\"%s\",\"qs\":[", g_first_case ? "" : ",\n", c.op.c_str());
    g_first_case = false;
    for (size_t k = 0; k < c.qs.size(); ++k) fprintf(g_out, "%s%u", k ? "," : "", c.qs[k]);
    fprintf(g_out, "],\"aux\":[");
    for (size_t k = 0; k < c.aux.size(); ++k) fprintf(g_out, "%s%u", k ? "," : "", c.aux[k]);
    fprintf(g_out, "],\"val\":\"%a\",\"gc\":[", c.val);
    for (size_t k = 0; k < c.gc.size(); ++k) fprintf(g_out, "%s%u", k ? "," : "", c.gc[k]);
    fprintf(g_out, "],\"error\":%s,\"out\":\"%a\",\"outc\":%u,", err ? "true" : "false", out, outc);
    dump_state("state");
    fprintf(g_out, "}");
}

static void mode_gates() {
    const uint n = 5;
    suqa::setup(n);
    // Fill the one-shot QFT phase table with the largest register used below
    // (suqa.cu:621-625), so every case sees the same table.
    suqa::apply_qft({0, 1, 2, 3, 4});
    prep(n);
    fprintf(g_out, "{\"n\":%u,", n);
    dump_state("prep");
    fprintf(g_out, ",\"cases\":[\n");
    std::vector<Case> cs;
    const std::vector<std::vector<uint>> gcs = {{}, {1}, {1, 3}};
    for (const auto& gc : gcs) {
        for (uint q : {0U, 2U, 4U}) {
            bool clash = false;
            for (uint g : gc) clash |= (g == q);
            if (clash) continue;
            for (const char* op : {"x", "y", "z", "h", "s", "t", "tdg"}) cs.push_back({op, {q}, {}, 0.0, gc});
            cs.push_back({"u1", {q}, {}, 0.9137, gc});
            cs.push_back({"u1", {q}, {0}, -2.271, gc});
            for (uint p = 1; p <= 3; ++p) cs.push_back({"pauli", {q}, {p}, 0.6180339, gc});
        }
    }
    for (const auto& gc : std::vector<std::vector<uint>>{{}, {3}}) {
        cs.push_back({"cx", {0, 2}, {}, 0.0, gc});
        cs.push_back({"cx", {4, 1}, {0}, 0.0, gc});
        cs.push_back({"cx", {2, 0}, {1}, 0.0, gc});
        cs.push_back({"mcx", {0, 2, 4}, {}, 0.0, gc});
        cs.push_back({"mcx", {4, 0, 1}, {1, 0}, 0.0, gc});
        cs.push_back({"cu1", {0, 4}, {}, 1.234, gc});
        cs.push_back({"cu1", {4, 2}, {0}, -0.777, gc});
        cs.push_back({"mcu1", {1, 2, 0}, {}, 2.5, gc});
        cs.push_back({"mcu1", {0, 4, 2}, {0, 1}, 0.3, gc});
        cs.push_back({"swap", {0, 4}, {}, 0.0, gc});
        cs.push_back({"swap", {2, 1}, {}, 0.0, gc});
    }
    for (const auto& gc : std::vector<std::vector<uint>>{{}, {2}}) {
        for (uint a = 1; a <= 3; ++a)
            for (uint b = 1; b <= 3; ++b) {
                cs.push_back({"pauli", {1, 4}, {a, b}, 0.4321, gc});
                cs.push_back({"pauli", {3, 0}, {a, b}, -1.111, gc});
            }
        const std::vector<std::vector<uint>> t3 = {{3, 1, 1}, {1, 3, 1}, {1, 1, 3}, {2, 3, 2},
                                                   {3, 2, 2}, {3, 3, 3}, {1, 2, 3}, {1, 1, 1}};
        for (const auto& t : t3) {
            cs.push_back({"pauli", {0, 1, 4}, t, 0.8765, gc});
            cs.push_back({"pauli", {4, 3, 1}, t, -0.25, gc});
        }
        cs.push_back({"pauli", {0, 1, 3, 4}, {1, 2, 3, 1}, 0.5, gc});
    }
    cs.push_back({"qft", {1, 2, 4}, {}, 0.0, {}});
    cs.push_back({"qft", {0, 1, 2, 3, 4}, {}, 0.0, {}});
    cs.push_back({"qft_inverse", {0, 1, 2, 3}, {}, 0.0, {}});
    cs.push_back({"qft_inverse", {3, 1}, {}, 0.0, {4}});
    cs.push_back({"qft", {0, 2, 3}, {}, 0.0, {1}});
    for (uint q : {0U, 2U, 4U})
        for (double r : {0.1, 0.5, 0.93}) {
            cs.push_back({"measure", {q}, {}, r, {}});
            cs.push_back({"reset", {q}, {}, r, {}});
        }
    cs.push_back({"prob_filter", {0, 1, 2}, {1, 0, 1}, 0.0, {}});
    cs.push_back({"prob_filter", {4}, {1}, 0.0, {}});
    cs.push_back({"prob_filter", {3, 0}, {0, 0}, 0.0, {}});
    cs.push_back({"prob_filter", {0, 1, 2, 3, 4}, {1, 1, 0, 0, 1}, 0.0, {}});
    cs.push_back({"vnorm", {}, {}, 0.0, {}});
    cs.push_back({"vnormalize", {}, {}, 0.0, {}});
    // gates on qubits of the active gc_mask (the QSA layout applies such gates):
    // the reference's loops then update coinciding slots
    for (const auto& gc : std::vector<std::vector<uint>>{{1}, {1, 3}}) {
        for (uint q : {1U, 3U}) {
            for (const char* op : {"x", "y", "h", "z", "s"}) cs.push_back({op, {q}, {}, 0.0, gc});
            for (uint p = 1; p <= 3; ++p) cs.push_back({"pauli", {q}, {p}, 0.6180339, gc});
        }
        cs.push_back({"swap", {1, 4}, {}, 0.0, gc});
        cs.push_back({"swap", {4, 1}, {}, 0.0, gc});
        cs.push_back({"swap", {3, 1}, {}, 0.0, gc});
        cs.push_back({"swap", {0, 3}, {}, 0.0, gc});
        for (uint a = 1; a <= 3; ++a)
            for (uint b = 1; b <= 3; ++b) {
                cs.push_back({"pauli", {1, 4}, {a, b}, 0.4321, gc});
                cs.push_back({"pauli", {4, 1}, {a, b}, -1.111, gc});
                cs.push_back({"pauli", {3, 1}, {a, b}, 0.777, gc});
            }
        cs.push_back({"cx", {1, 3}, {}, 0.0, gc});
        cs.push_back({"cu1", {3, 1}, {}, 1.234, gc});
    }
    for (const auto& c : cs) run_case(c, n);
    fprintf(g_out, "]}\n");
    suqa::clear();
}

// ------------------------------------------------------------------ qms
// Mirrors the driver loop of qms.cu:132-149 and logs each step's return code
// and the pcg32 state after it.
static void mode_qms(int argc, char** argv) {
    double beta = atof(argv[2]);
    uint steps = (uint)atoi(argv[3]);
    uint reset_each = (uint)atoi(argv[4]);
    uint r = (uint)atoi(argv[5]);
    unsigned long long seed = strtoull(argv[6], nullptr, 10);
    double emin = atof(argv[7]), emax = atof(argv[8]);
    int pe = atoi(argv[9]);
    int therm = atoi(argv[10]);
    uint maxrev = (uint)atoi(argv[11]);
    bool recrev = atoi(argv[12]) != 0;
    (void)argc;

    qms::metro_steps = steps;
    qms::reset_each = reset_each;
    qms::ene_qbits = r;
    qms::max_reverse_attempts = maxrev;
    qms::n_phase_estimation = pe;
    qms::record_reverse = recrev;
    qms::iseed = seed;
    qms::rangen.set_seed(qms::iseed);
    qms::syst_qbits = (uint)syst_qbits;
    qms::nqubits = qms::syst_qbits + qms::ene_qbits + 1;
    qms::Dim = (1U << qms::nqubits);
    qms::ene_levels = (1U << qms::ene_qbits);
    qms::syst_levels = (1U << qms::syst_qbits);
    qms::t_PE_shift = emin;
    qms::t_PE_factor = (qms::ene_levels - 1) / (double)(qms::ene_levels * (emax - emin));
    qms::t_phase_estimation = qms::t_PE_factor * 8. * atan(1.0);
    suqa::setup(qms::nqubits);
    qms::setup(beta);
    init_state();

    std::vector<int> rets;
    std::vector<unsigned long long> rng_after;
    uint s0 = 0U, acc = 0U;
    for (uint s = 0U; s < steps; ++s) {
        bool take = (s > s0 + (uint)therm and (s - s0) % qms::reset_each == 0U);
        int ret = qms::metro_step(take);
        if (ret < 0) { init_state(); s0 = s + 1; }
        if (ret == 1 or ret == 2) acc++;
        rets.push_back(ret);
        rng_after.push_back((unsigned long long)qms::rangen.get_state().state);
    }
    fprintf(g_out, "{\"beta\":\"%a\",\"steps\":%u,\"reset_each\":%u,\"r\":%u,\"seed\":%llu,"
                   "\"ene_min\":\"%a\",\"ene_max\":\"%a\",\"pe_steps\":%d,\"thermalization\":%d,"
                   "\"max_reverse\":%u,\"record_reverse\":%d,\"accepted\":%u,",
            beta, steps, reset_each, r, seed, emin, emax, pe, therm, maxrev, recrev ? 1 : 0, acc);
    fprintf(g_out, "\"t_PE_factor\":\"%a\",\"t_phase_estimation\":\"%a\",\"c_factor\":\"%a\",",
            qms::t_PE_factor, qms::t_phase_estimation, qms::c_factor);
    fprintf(g_out, "\"rets\":[");
    for (size_t k = 0; k < rets.size(); ++k) fprintf(g_out, "%s%d", k ? "," : "", rets[k]);
    fprintf(g_out, "],\"rng_after\":[");
    for (size_t k = 0; k < rng_after.size(); ++k) fprintf(g_out, "%s\"%llu\"", k ? "," : "", rng_after[k]);
    fprintf(g_out, "],\"E\":[");
    for (size_t k = 0; k < qms::E_measures.size(); ++k) fprintf(g_out, "%s\"%a\"", k ? "," : "", qms::E_measures[k]);
    fprintf(g_out, "],\"X\":[");
    for (size_t k = 0; k < qms::X_measures.size(); ++k) fprintf(g_out, "%s\"%a\"", k ? "," : "", qms::X_measures[k]);
    fprintf(g_out, "],\"reverse_counters\":[");
    for (size_t k = 0; k < qms::reverse_counters.size(); ++k)
        fprintf(g_out, "%s%u", k ? "," : "", qms::reverse_counters[k]);
    fprintf(g_out, "],");
    dump_state("final_state");
    fprintf(g_out, "}\n");
    qms::clear();
    suqa::clear();
}

// ---------------------------------------------------------------- circuit
// Random-circuit family of the gate-throughput workload: |0>, H on every
// qubit, `layers` x n gates drawn uniformly from {H,X,RX,RZ,U1,CX,MCX(3)}
// with pcg(seed); targets and controls distinct; angles 2*pi*doub().
// Then measure_qbit on min(8,n) distinct random qubits, one doub() each.
// RX(a) = exp(-i a X/2) = pauli({q},{X},-a/2); RZ likewise with Z.
// Applies the random layers with pcg `rng` (after the H layer); with `log`
// each gate is written to g_out as [kind, target, c0, c1, c2, "angle"].
static void apply_random_layers(uint n, uint layers, pcg& rng, bool log) {
    bool first = true;
    for (uint l = 0; l < layers; ++l) {
        for (uint g = 0; g < n; ++g) {
            int kind = rng.randint(0, 7);
            uint t = (uint)rng.randint(0, (int)n);
            uint nc = kind == 5 ? 1 : (kind == 6 ? 3 : 0);
            std::vector<uint> ctr;
            while (ctr.size() < nc) {
                uint c = (uint)rng.randint(0, (int)n);
                bool dup = (c == t);
                for (uint x : ctr) dup |= (x == c);
                if (!dup) ctr.push_back(c);
            }
            double ang = (kind >= 2 && kind <= 4) ? 2.0 * M_PI * rng.doub() : 0.0;
            switch (kind) {
            case 0: suqa::apply_h(t); break;
            case 1: suqa::apply_x(t); break;
            case 2: suqa::apply_pauli_TP_rotation({t}, {PAULI_X}, -ang / 2.0); break;
            case 3: suqa::apply_pauli_TP_rotation({t}, {PAULI_Z}, -ang / 2.0); break;
            case 4: suqa::apply_u1(t, ang); break;
            case 5: suqa::apply_cx(ctr[0], t); break;
            case 6: suqa::apply_mcx(ctr, t); break;
            }
            if (log)
                fprintf(g_out, "%s[%d,%u,%u,%u,%u,\"%a\"]", first ? "" : ",", kind, t,
                        nc > 0 ? ctr[0] : 0U, nc > 1 ? ctr[1] : 0U, nc > 2 ? ctr[2] : 0U, ang);
            first = false;
        }
    }
}

// The measurement plan: min(8,n) distinct random qubits, one doub() each.
static void measure_random(uint n, pcg& rng) {
    uint nm = n < 8 ? n : 8;
    std::vector<uint> mq;
    while (mq.size() < nm) {
        uint q = (uint)rng.randint(0, (int)n);
        bool dup = false;
        for (uint x : mq) dup |= (x == q);
        if (!dup) mq.push_back(q);
    }
    fprintf(g_out, "\"measure\":[");
    for (uint k = 0; k < nm; ++k) {
        double rd = rng.doub();
        uint c;
        suqa::measure_qbit(mq[k], c, rd);
        fprintf(g_out, "%s[%u,\"%a\",%u]", k ? "," : "", mq[k], rd, c);
    }
    fprintf(g_out, "]");
}

// Random-circuit family of the gate-throughput workload: |0>, H on every
// qubit, `layers` x n gates drawn uniformly from {H,X,RX,RZ,U1,CX,MCX(3)}
// with pcg(seed); targets and controls distinct; angles 2*pi*doub().
// Then measure_qbit on min(8,n) distinct random qubits, one doub() each.
// RX(a) = exp(-i a X/2) = pauli({q},{X},-a/2); RZ likewise with Z.
static void mode_circuit(char** argv) {
    uint n = (uint)atoi(argv[2]);
    uint layers = (uint)atoi(argv[3]);
    unsigned long long seed = strtoull(argv[4], nullptr, 10);
    pcg rng;
    rng.set_seed(seed);
    suqa::setup(n);
    for (uint q = 0; q < n; ++q) suqa::apply_h(q);
    fprintf(g_out, "{\"n\":%u,\"layers\":%u,\"seed\":%llu,\"gates\":[", n, layers, seed);
    apply_random_layers(n, layers, rng, true);
    fprintf(g_out, "],");
    dump_state("pre_measure");
    fprintf(g_out, ",");
    measure_random(n, rng);
    fprintf(g_out, ",");
    dump_state("final");
    fprintf(g_out, "}\n");
    suqa::clear();
}

// Pairwise (recursive halving) sum of f(i) over [lo, hi).
template <class F>
static double pairwise(uint64_t lo, uint64_t hi, const F& f) {
    if (hi - lo <= 8) {
        double s = 0.0;
        for (uint64_t i = lo; i < hi; ++i) s += f(i);
        return s;
    }
    uint64_t mid = lo + (hi - lo) / 2;
    return pairwise(lo, mid, f) + pairwise(mid, hi, f);
}

// Checksum of the state (sum re, sum im, sum i |psi_i|^2 / N; pairwise sums)
// and the amplitudes at 64 fixed indices i_k = (k * 2654435761 + 12345) mod N.
static void dump_summary(const char* key) {
    const uint64_t N = suqa::state.size();
    const double* re = suqa::state.data_re;
    const double* im = suqa::state.data_im;
    double sre = pairwise(0, N, [&](uint64_t i) { return re[i]; });
    double sim = pairwise(0, N, [&](uint64_t i) { return im[i]; });
    double sidx = pairwise(0, N, [&](uint64_t i) { return (double)i * (re[i] * re[i] + im[i] * im[i]); }) / (double)N;
    fprintf(g_out, "\"%s\":{\"sum_re\":\"%a\",\"sum_im\":\"%a\",\"sum_index_prob\":\"%a\",\"samples\":[", key, sre,
            sim, sidx);
    for (uint64_t k = 0; k < 64; ++k) {
        uint64_t i = (k * 2654435761ULL + 12345ULL) % N;
        fprintf(g_out, "%s[%llu,\"%a\",\"%a\"]", k ? "," : "", (unsigned long long)i, re[i], im[i]);
    }
    fprintf(g_out, "]}");
}

// The same circuit at sizes too large to store whole: the checksum and 64
// sample amplitudes of the state before the measurements and after them.
static void mode_circuit_summary(char** argv) {
    uint n = (uint)atoi(argv[2]);
    uint layers = (uint)atoi(argv[3]);
    unsigned long long seed = strtoull(argv[4], nullptr, 10);
    pcg rng;
    rng.set_seed(seed);
    suqa::setup(n);
    for (uint q = 0; q < n; ++q) suqa::apply_h(q);
    fprintf(g_out, "{\"n\":%u,\"layers\":%u,\"seed\":%llu,", n, layers, seed);
    apply_random_layers(n, layers, rng, false);
    dump_summary("pre_measure");
    fprintf(g_out, ",");
    measure_random(n, rng);
    fprintf(g_out, ",");
    dump_summary("final");
    fprintf(g_out, "}\n");
    suqa::clear();
}

// ------------------------------------------------------------------ qsa
// The QSA driver loop of qsa.cu:115-262 on the reference's qsa:: functions,
// bounded: it stops once `max_ann` annealing steps have run (the stock
// driver's samples can take ~10^5 restarts).  Records, per annealing step,
// the sample, restart and step indices, the move drawn, the Szegedy
// register's readout and the pcg32 state after its measurement; per finished
// sample the energy, the X value and the restart count; the final state.
static void mode_qsa(char** argv) {
    const double beta = atof(argv[2]);
    const uint sampling = (uint)atoi(argv[3]);
    const uint ene_q = (uint)atoi(argv[4]), sz_q = (uint)atoi(argv[5]);
    const unsigned long long seed = strtoull(argv[6], nullptr, 10);
    const double emin = atof(argv[7]), emax = atof(argv[8]), emin_sz = atof(argv[9]), emax_sz = atof(argv[10]);
    const int pe = atoi(argv[11]);
    const uint ann_seq = (uint)atoi(argv[12]), max_ann = (uint)atoi(argv[13]);
    // qsa.cu:115-160
    qsa::beta = beta;
    qsa::ene_qbits = ene_q;
    qsa::szegedy_qbits = sz_q;
    qsa::n_phase_estimation = pe;
    qsa::iseed = seed;
    qsa::rangen.set_seed(qsa::iseed);
    qsa::iseed = qsa::rangen.get_seed();
    qsa::syst_qbits = (uint)syst_qbits;
    qsa::nqubits = qsa::syst_qbits + qsa::ene_qbits + qsa::szegedy_qbits + 1;
    qsa::Dim = (1U << qsa::nqubits);
    qsa::ene_levels = (1U << qsa::ene_qbits);
    double szegedy_levels = (1U << qsa::szegedy_qbits);
    qsa::t_PE_shift = emin;
    qsa::t_PE_factor = (qsa::ene_levels - 1) / (double)(qsa::ene_levels * (emax - emin));
    qsa::t_phase_estimation = qsa::t_PE_factor * 8. * atan(1.0);
    qsa::t_PE_shift_szegedy = emin_sz;
    qsa::t_PE_factor_szegedy = (szegedy_levels - 1) / (double)((szegedy_levels) * (emax_sz - emin_sz));
    qsa::t_phase_estimation_szegedy = qsa::t_PE_factor_szegedy * 8. * atan(1.0);
    suqa::setup(qsa::nqubits);
    qsa::setup();
    fprintf(g_out, "{\"beta\":\"%a\",\"sampling\":%u,\"ene_qbits\":%u,\"szegedy_qbits\":%u,\"seed\":%llu,"
                   "\"ene_min\":\"%a\",\"ene_max\":\"%a\",\"ene_min_szegedy\":\"%a\",\"ene_max_szegedy\":\"%a\","
                   "\"pe_steps\":%d,\"annealing_sequences\":%u,\"max_annealing_steps\":%u,\"steps\":[",
            beta, sampling, ene_q, sz_q, seed, emin, emax, emin_sz, emax_sz, pe, ann_seq, max_ann);
    int rejection1 = 0;
    uint nann = 0;
    bool first = true;
    double beta_i = qsa::beta / ((double)ann_seq);
    std::vector<std::string> samples;
    for (uint iiii = 0; iiii < sampling && nann < max_ann;) {
        std::vector<uint> c_Ene_test(qsa::szegedy_qbits, 0);
        std::vector<uint> c_Ene(qsa::szegedy_qbits, 0);
        int control = 0;
        uint attempt = 0;
        while (control == 0 && nann < max_ann) {
            int control2 = 0;
            uint move = qsa::draw_C();
            qsa::reset_non_syst_qbits();
            qsa_init_state();
            for (uint s = 0U; s < ann_seq; s++) {
                move = qsa::draw_C();
                qsa::apply_PE_szegedy(qsa::bm_states, qsa::bm_szegedy, (double)(s + 1) * beta_i, move);
                suqa::measure_qbits(qsa::bm_szegedy, c_Ene_test, qsa::extract_rands(qsa::szegedy_qbits));
                ++nann;
                fprintf(g_out, "%s[%u,%u,%u,%u,%u,\"%llu\"]", first ? "" : ",", iiii, attempt, s, move,
                        qsa::creg_to_uint(c_Ene_test), (unsigned long long)qsa::rangen.get_state().state);
                first = false;
                if (qsa::creg_to_uint(c_Ene_test) != 0U) {
                    control2 = 1;
                    rejection1++;
                    break;
                }
                if (nann >= max_ann) break;
            }
            if (control2 == 0 && nann <= max_ann) control = 1;
            ++attempt;
        }
        if (control == 0) break;   // bounded: stopped inside a sample
        qsa::apply_phase_estimation_measure(qsa::bm_states, qsa::bm_szegedy, qsa::t_phase_estimation_szegedy,
                                            qsa::n_phase_estimation);
        suqa::measure_qbits(qsa::bm_szegedy, c_Ene, qsa::extract_rands(qsa::szegedy_qbits));
        double Ene_measure = qsa::t_PE_shift_szegedy +
                             qsa::creg_to_uint(c_Ene) / (double)(qsa::t_PE_factor_szegedy * szegedy_levels);
        double X = measure_X(qsa::rangen);
        char buf[256];
        snprintf(buf, sizeof buf, "[\"%a\",\"%a\",%d,\"%llu\"]", Ene_measure, X, rejection1,
                 (unsigned long long)qsa::rangen.get_state().state);
        samples.push_back(buf);
        ++iiii;
    }
    fprintf(g_out, "],\"samples\":[");
    for (size_t k = 0; k < samples.size(); ++k) fprintf(g_out, "%s%s", k ? "," : "", samples[k].c_str());
    fprintf(g_out, "],\"annealing_steps\":%u,\"final_rng\":\"%llu\",", nann,
            (unsigned long long)qsa::rangen.get_state().state);
    dump_state("final_state");
    fprintf(g_out, "}\n");
    qsa::clear();
    suqa::clear();
}

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s pcg|gates|qms|circuit|circuit_summary ... <out.json>\n", argv[0]);
        return 2;
    }
    std::string mode = argv[1];
    g_out = fopen(argv[argc - 1], "w");
    if (!g_out) return 3;
    if (mode == "pcg") mode_pcg();
    else if (mode == "gates") mode_gates();
    else if (mode == "qms" && argc == 14) mode_qms(argc, argv);
    else if (mode == "circuit" && argc == 6) mode_circuit(argv);
    else if (mode == "circuit_summary" && argc == 6) mode_circuit_summary(argv);
    else if (mode == "qsa" && argc == 15) mode_qsa(argv);
    else { fprintf(stderr, "bad arguments\n"); return 2; }
    fclose(g_out);
    return 0;
}
