// suqa.cpp -- suqa:: API over the C-ABI (mirrors suqa.cu; see suqa.hpp).
#include "suqa.hpp"

#include <algorithm>
#include <cmath>
#include <string>

ComplexVec suqa::state;
uint suqa::gc_mask = 0U;
uint suqa::nq = 0U;
std::vector<double> suqa::rphase_m;
suqa_hip_ctx* suqa::ctx = nullptr;
GateCounterList suqa::gatecounters;

namespace {

void chk(int rc, const char* where) {
    if (rc != SUQA_OK) throw std::runtime_error(std::string("ERROR: in ") + where + ": " + suqa_hip_last_error());
}

void refresh_view() {
    uint32_t n = 0;
    uint64_t N = 0;
    double *re = nullptr, *im = nullptr;
    chk(suqa_hip_state_info(suqa::ctx, &n, &N, &re, &im), "suqa::state");
    suqa::state.vecsize = (uint)N;
    suqa::state.data = re;
    suqa::state.data_re = re;
    suqa::state.data_im = im;
}

// Gate counts of one gate with `extra` controls besides the gc_mask ones, as
// an n-control Toffoli, `times` times (the reference's GATECOUNT blocks).
void count_toffoli(uint extra, uint times = 1) {
    if (suqa::gatecounters.counters.empty()) return;
    const GateRecord gr = GateCounter::n_ctrl_toffoli_gates(suqa::gatecounters.gc_mask_set_qbits + extra);
    suqa::gatecounters.increment_g1g2(times * gr.ng1, times * gr.ng2);
}

// Pauli-string rotation on qtar qubits, cnz of them X or Y (suqa.cu:582-603):
// the rotation, a ladder of 2 (qtar - 1) controlled gates, and basis changes.
void count_pauli(uint qtar, uint cnz) {
    if (suqa::gatecounters.counters.empty() || qtar == 0) return;
    const uint n = suqa::gatecounters.gc_mask_set_qbits;
    const GateRecord g0 = GateCounter::n_ctrl_toffoli_gates(n), g1 = GateCounter::n_ctrl_toffoli_gates(n + 1);
    suqa::gatecounters.increment_g1g2(g0);
    suqa::gatecounters.increment_g1g2(2 * (qtar - 1) * g1.ng1, 2 * (qtar - 1) * g1.ng2);
    if (cnz > 0) {
        suqa::gatecounters.increment_g1g2(2 * g0.ng1, 2 * g0.ng2);
        suqa::gatecounters.increment_g1g2(0, 2 * (cnz - 1));
    }
}

uint count_non_z(const std::vector<uint>& types) {
    uint k = 0;
    for (uint t : types)
        if (t != PAULI_Z) ++k;
    return k;
}

}  // namespace

void suqa::print_banner() {
    printf("\n   SUQA -- Simulator for Universal Quantum Algorithms (MI355X engine)\n\n");
}

void suqa::activate_gc_mask(const bmReg& q_controls) {   // suqa.cu:45-52
    for (const auto& q : q_controls) suqa::gc_mask |= 1U << q;
    gatecounters.update_cmask_setbits(gc_mask);
}

void suqa::deactivate_gc_mask(const bmReg& q_controls) { // suqa.cu:54-61
    for (const auto& q : q_controls) suqa::gc_mask &= ~(1U << q);
    gatecounters.update_cmask_setbits(gc_mask);
}

double suqa::vnorm() {
    double v = 0.0;
    chk(suqa_hip_vnorm(ctx, &v), "suqa::vnorm()");
    return v;
}

void suqa::vnormalize() { chk(suqa_hip_vnormalize(ctx), "suqa::vnormalize()"); }

void suqa::init_state() { chk(suqa_hip_init_state(ctx), "suqa::init_state()"); }

void suqa::deallocate_state() {
    if (ctx) chk(suqa_hip_clear(ctx), "suqa::deallocate_state()");
    ctx = nullptr;
    state = ComplexVec();
}

void suqa::allocate_state(uint Dim) {   // suqa.cu:772-788 (Dim must be a power of two)
    uint n = 0;
    while ((1ULL << n) < (unsigned long long)Dim) ++n;
    if ((1ULL << n) != (unsigned long long)Dim) throw std::runtime_error("ERROR: state dimension must be 2^n");
    if (ctx) chk(suqa_hip_allocate_state(ctx, n), "suqa::allocate_state()");
    else chk(suqa_hip_setup(n, &ctx), "suqa::allocate_state()");
    refresh_view();
}

void suqa::apply_x(uint q) {
    chk(suqa_hip_apply_x(ctx, q, gc_mask), "suqa::apply_x()");
    count_toffoli(0);
}
void suqa::apply_x(const bmReg& qs) {
    for (const auto& q : qs) suqa::apply_x(q);
}
void suqa::apply_y(uint q) {
    chk(suqa_hip_apply_y(ctx, q, gc_mask), "suqa::apply_y()");
    count_toffoli(0);
}
void suqa::apply_y(const bmReg& qs) {
    for (const auto& q : qs) suqa::apply_y(q);
}
void suqa::apply_z(uint q) { suqa::apply_u1(q, M_PI); }
void suqa::apply_z(const bmReg& qs) {
    for (const auto& q : qs) suqa::apply_u1(q, M_PI);
}
void suqa::apply_h(uint q) {
    chk(suqa_hip_apply_h(ctx, q, gc_mask), "suqa::apply_h()");
    count_toffoli(0);
}
void suqa::apply_h(const bmReg& qs) {
    for (const auto& q : qs) suqa::apply_h(q);
}
void suqa::apply_t(uint q) { suqa::apply_u1(q, M_PI / 4.0); }
void suqa::apply_t(const bmReg& qs) {
    for (const auto& q : qs) suqa::apply_u1(q, M_PI / 4.0);
}
void suqa::apply_tdg(uint q) { suqa::apply_u1(q, -M_PI / 4.0); }
void suqa::apply_tdg(const bmReg& qs) {
    for (const auto& q : qs) suqa::apply_u1(q, -M_PI / 4.0);
}
void suqa::apply_s(uint q) { suqa::apply_u1(q, M_PI / 2.0); }
void suqa::apply_s(const bmReg& qs) {
    for (const auto& q : qs) suqa::apply_u1(q, M_PI / 2.0);
}

void suqa::apply_u1(uint q, double phase) { suqa::apply_u1(q, 1U, phase); }
void suqa::apply_u1(uint q, uint q_mask, double phase) {
    chk(suqa_hip_apply_u1(ctx, q, q_mask, phase, gc_mask), "suqa::apply_u1()");
    count_toffoli(0);
}

void suqa::apply_cx(const uint& q_control, const uint& q_target, const uint& q_mask) {
    chk(suqa_hip_apply_mcx(ctx, &q_control, &q_mask, 1, q_target, gc_mask), "suqa::apply_cx()");
    count_toffoli(1);
}
void suqa::apply_mcx(const bmReg& q_controls, const uint& q_target) {
    chk(suqa_hip_apply_mcx(ctx, q_controls.data(), nullptr, (uint32_t)q_controls.size(), q_target, gc_mask),
        "suqa::apply_mcx()");
    count_toffoli((uint)q_controls.size());
}
void suqa::apply_mcx(const bmReg& q_controls, const bmReg& q_mask, const uint& q_target) {
    if (q_mask.size() < q_controls.size()) throw std::runtime_error("ERROR: in suqa::apply_mcx(): mask size");
    chk(suqa_hip_apply_mcx(ctx, q_controls.data(), q_mask.data(), (uint32_t)q_controls.size(), q_target, gc_mask),
        "suqa::apply_mcx()");
    count_toffoli((uint)q_controls.size());
}
void suqa::apply_cu1(uint q_control, uint q_target, double phase, uint q_mask) {
    chk(suqa_hip_apply_mcu1(ctx, &q_control, &q_mask, 1, q_target, phase, gc_mask), "suqa::apply_cu1()");
    count_toffoli(1);
}
void suqa::apply_mcu1(const bmReg& q_controls, const uint& q_target, double phase) {
    chk(suqa_hip_apply_mcu1(ctx, q_controls.data(), nullptr, (uint32_t)q_controls.size(), q_target, phase, gc_mask),
        "suqa::apply_mcu1()");
    count_toffoli((uint)q_controls.size());
}
void suqa::apply_mcu1(const bmReg& q_controls, const bmReg& q_mask, const uint& q_target, double phase) {
    if (q_mask.size() < q_controls.size()) throw std::runtime_error("ERROR: in suqa::apply_mcu1(): mask size");
    chk(suqa_hip_apply_mcu1(ctx, q_controls.data(), q_mask.data(), (uint32_t)q_controls.size(), q_target, phase,
                            gc_mask),
        "suqa::apply_mcu1()");
    count_toffoli((uint)q_controls.size());
}
void suqa::apply_swap(const uint& q1, const uint& q2) {
    chk(suqa_hip_apply_swap(ctx, q1, q2, gc_mask), "suqa::apply_swap()");
    count_toffoli(1, 3);   // three CNOTs
}

void suqa::apply_phase_list(uint q0, uint q_size, const std::vector<double>& phases) {
    chk(suqa_hip_apply_phase_list(ctx, q0, q_size, phases.data(), (uint32_t)phases.size(), gc_mask),
        "suqa::apply_phase_list()");
    // 2^(k+1) - 3 controlled phases on k qubits, each an (n+1)- and an n-control
    // Toffoli (suqa.cu:442-453, the count its comment states)
    if (!gatecounters.counters.empty()) {
        const uint fact = (1U << (q_size + 1)) - 3U, n = gatecounters.gc_mask_set_qbits;
        const GateRecord a = GateCounter::n_ctrl_toffoli_gates(n), b = GateCounter::n_ctrl_toffoli_gates(n + 1);
        gatecounters.increment_g1g2(fact * (a.ng1 + b.ng1), fact * (a.ng2 + b.ng2));
    }
}

void suqa::apply_pauli_TP_rotation(const bmReg& q_apply, const std::vector<uint>& types, double phase) {
    if (q_apply.size() != types.size())   // suqa.cu:477-479
        throw std::runtime_error("ERROR: in suqa::apply_pauli_TP_rotation(): mismatch between qubits number and "
                                 "pauli types specified");
    chk(suqa_hip_apply_pauli_TP_rotation(ctx, q_apply.data(), types.data(), (uint32_t)q_apply.size(), phase,
                                         gc_mask),
        "suqa::apply_pauli_TP_rotation()");
    count_pauli((uint)q_apply.size(), count_non_z(types));
}

void suqa::apply_pauli_TP_rotation_general(const bmReg& q_apply, const std::vector<uint>& types, double phase) {
    if (q_apply.size() != types.size())
        throw std::runtime_error("ERROR: in suqa::apply_pauli_TP_rotation_general(): size mismatch");
    chk(suqa_hip_apply_pauli_rotation_general(ctx, q_apply.data(), types.data(), (uint32_t)q_apply.size(), phase,
                                              gc_mask),
        "suqa::apply_pauli_TP_rotation_general()");
    count_pauli((uint)q_apply.size(), count_non_z(types));
}

void suqa::apply_rx(uint q, double theta) { suqa::apply_pauli_TP_rotation({q}, {PAULI_X}, -theta / 2.0); }
void suqa::apply_rz(uint q, double theta) { suqa::apply_pauli_TP_rotation({q}, {PAULI_Z}, -theta / 2.0); }

// QFT (suqa.cu:611-661): no final swaps.  The crm phase table is filled by
// the first call (suqa.cu:621-625); entry m is 2 pi / 2^m whatever the fill
// size, so a later, larger register extends the table (where the reference
// would read past its end) and every entry the reference defines is unchanged.
static void fill_rphase(const uint& nqubits) {
    suqa::rphase_m.resize(nqubits + 1);
    uint c = 1;
    for (uint i = 0; i < nqubits + 1; ++i) {
        suqa::rphase_m[i] = (2.0 * M_PI / (double)c);
        c <<= 1;
    }
}

static void qft_crm(const uint& q_control, const uint& q_target, const int& m, const int& nqubits) {
    const size_t idx = (size_t)(m > 0 ? m : -m);
    if (suqa::rphase_m.empty() || idx >= suqa::rphase_m.size())
        fill_rphase((uint)std::max<size_t>((size_t)nqubits, idx));
    double rphase = suqa::rphase_m[idx];
    if (m <= 0) rphase *= -1;
    suqa::apply_cu1(q_control, q_target, rphase);
}

void suqa::apply_qft(const std::vector<uint>& qact) {
    int qsize = (int)qact.size();
    for (int outer_i = qsize - 1; outer_i >= 0; outer_i--) {
        suqa::apply_h(qact[outer_i]);
        for (int inner_i = outer_i - 1; inner_i >= 0; inner_i--)
            qft_crm(qact[inner_i], qact[outer_i], +1 + (outer_i - inner_i), qsize);
    }
}

void suqa::apply_qft_inverse(const std::vector<uint>& qact) {
    int qsize = (int)qact.size();
    for (int outer_i = 0; outer_i < qsize; outer_i++) {
        for (int inner_i = 0; inner_i < outer_i; inner_i++)
            qft_crm(qact[inner_i], qact[outer_i], -1 - (outer_i - inner_i), qsize);
        suqa::apply_h(qact[outer_i]);
    }
}

void suqa::measure_qbit(uint q, uint& c, double rdoub) {
    uint32_t cc = 0;
    chk(suqa_hip_measure_qbit(ctx, q, rdoub, &cc), "suqa::measure_qbit()");
    c = cc;
}

void suqa::measure_qbits(const bmReg& qs, std::vector<uint>& cs, const std::vector<double>& rdoubs) {
    const uint32_t n = (uint32_t)qs.size();
    std::vector<uint32_t> out(n);
    chk(suqa_hip_measure_qbits(ctx, qs.data(), rdoubs.data(), n, out.data()), "suqa::measure_qbits()");
    for (uint32_t k = 0; k < n; ++k) cs[k] = out[k];
}

// With gate counters registered, the X of a reset is counted when the outcome
// is 1 (as the reference's reset counts its apply_x), so the host needs the
// outcome; otherwise the engine applies measurement and X without a wait.
void suqa::apply_reset(uint q, double rdoub) {   // suqa.cu:730-738
    if (!gatecounters.counters.empty()) {
        uint c;
        suqa::measure_qbit(q, c, rdoub);
        if (c) suqa::apply_x(q);
        return;
    }
    chk(suqa_hip_apply_reset(ctx, q, rdoub, gc_mask), "suqa::apply_reset()");
}

void suqa::apply_reset(const bmReg& qs, std::vector<double> rdoubs) {   // suqa.cu:752-757
    if (!gatecounters.counters.empty()) {
        for (uint i = 0; i < qs.size(); ++i) suqa::apply_reset(qs[i], rdoubs[i]);
        return;
    }
    chk(suqa_hip_apply_resets(ctx, qs.data(), rdoubs.data(), (uint32_t)qs.size(), gc_mask),
        "suqa::apply_reset()");
}

void suqa::prob_filter(const bmReg& qs, const std::vector<uint>& q_mask, double& prob) {
    const uint32_t n = (uint32_t)std::min(qs.size(), q_mask.size());
    chk(suqa_hip_prob_filter(ctx, qs.data(), q_mask.data(), n, &prob), "suqa::prob_filter()");
}

void suqa::setup(uint num_qubits) {   // suqa.cu:792-844
    suqa::nq = num_qubits;
    printf("Running on gpu (MI355X engine) in release mode\n");
    if (ctx) chk(suqa_hip_allocate_state(ctx, num_qubits), "suqa::setup()");
    else chk(suqa_hip_setup(num_qubits, &ctx), "suqa::setup()");
    refresh_view();
}

void suqa::clear() {   // suqa.cu:846-866: the gate counts are printed last
    suqa::deallocate_state();
    gatecounters.print_counts();
}
