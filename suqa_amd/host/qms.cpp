// qms.cpp -- qms:: on the MI355X engine.  Same control flow and random-draw
// order as qms.cuh:373-545, so a run with a fixed pcg32 seed makes the same
// accept/reject decisions as the reference CPU path.
#include "qms.hpp"

#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "system.hpp"

namespace qms {

uint syst_qbits, syst_levels, ene_qbits, ene_levels, nqubits, Dim;
uint max_reverse_attempts, metro_steps, reset_each;
unsigned long long iseed = 0ULL;
double t_phase_estimation, t_PE_factor, t_PE_shift;
int n_phase_estimation;
uint gCi, c_acc = 0;
bool record_reverse = false;
std::vector<uint> reverse_counters;
GateCounter gctr_global("global"), gctr_metrostep("metro step"), gctr_sample("sample"), gctr_measure("measure"),
    gctr_reverse("reverse");
pcg rangen(1ULL);
std::vector<double> X_measures, E_measures;
double c_factor;
std::vector<uint> bm_syst, bm_enes_new;
uint bm_acc;
uint W_mask, W_mask_Enew;
uint curr_E_old;
double curr_E_old_d;
bool two_registers = false;
std::vector<uint> bm_enes_old;
uint W_mask_Eold = 0U;
bool use_graphs = false;
bool deferred = false;
bool defer_values = false;

namespace {

// Captured launch sequences of the fixed parts of a step (suqa_hip_graph_*):
// Phi, Phi^dag and C_i Phi / Phi^dag C_i depend only on the run's parameters,
// so each is recorded the first time it is applied and replayed after that.
// Not with gate counters registered (a replay counts nothing).
struct Graph {
    bool ok = false;
    uint32_t id = 0;
};
Graph g_phi, g_phi_old, g_phi_inv, g_cphi[3], g_phi_inv_c[3];

void drop_graphs() {
    g_phi = g_phi_old = g_phi_inv = Graph{};
    for (auto& g : g_cphi) g = Graph{};
    for (auto& g : g_phi_inv_c) g = Graph{};
}

void chk_graph(int rc) {
    if (rc != SUQA_OK) throw std::runtime_error(std::string("ERROR: graph capture: ") + suqa_hip_last_error());
}

// measure_qbits whose outcomes the step never reads (the E re-collapse after
// the X measurement): the same measurements and draws, and nothing waits for
// the outcomes.  With gate counters the reference's call is kept.
void measure_discarded(const bmReg& reg, const std::vector<double>& rdoubs) {
    if (!suqa::gatecounters.counters.empty()) {
        std::vector<uint> cs(reg.size());
        suqa::measure_qbits(reg, cs, rdoubs);
        return;
    }
    if (suqa_hip_measure_qbits(suqa::ctx, reg.data(), rdoubs.data(), (uint32_t)reg.size(), nullptr) != SUQA_OK)
        throw std::runtime_error(std::string("ERROR: in suqa::measure_qbits(): ") + suqa_hip_last_error());
}

template <class F>
void replay(Graph& g, F apply) {
    if (!use_graphs || !suqa::gatecounters.counters.empty()) {
        apply();
        return;
    }
    if (!g.ok) {
        chk_graph(suqa_hip_graph_begin(suqa::ctx));
        apply();   // recorded, not run
        chk_graph(suqa_hip_graph_end(suqa::ctx, &g.id));
        g.ok = true;
    }
    chk_graph(suqa_hip_graph_launch(suqa::ctx, g.id));
}

// Deferred readouts: the outcome slots handed out since the last read (ring
// order), the host copy of what was read, and the E / X values that wait for
// their outcomes.
struct Deferred {
    bool pending = false;
    uint32_t begin = 0, end = 0;
    std::vector<uint32_t> ring = std::vector<uint32_t>(SUQA_HIP_OUTCOME_SLOTS, 0u);
    struct Rec {
        bool is_x;
        uint32_t slot;
        size_t idx;
    };
    std::vector<Rec> recs;
} D;

void chk(int rc, const char* where) {
    if (rc != SUQA_OK) throw std::runtime_error(std::string("ERROR: in ") + where + ": " + suqa_hip_last_error());
}

void note_slots(uint32_t slot, uint32_t n) {
    if (!D.pending) {
        D.begin = slot;
        D.pending = true;
    }
    D.end = slot + n;
}

uint32_t measure_deferred(const bmReg& qs, const std::vector<double>& rdoubs) {
    uint32_t slot = 0;
    chk(suqa_hip_measure_qbits_deferred(suqa::ctx, qs.data(), rdoubs.data(), (uint32_t)qs.size(), &slot),
        "suqa::measure_qbits()");
    note_slots(slot, (uint32_t)qs.size());
    return slot;
}

// The register value of n outcomes from slot (LSB first, creg_to_uint).
uint slot_value(uint32_t slot, uint32_t n) {
    uint v = 0;
    for (uint32_t k = 0; k < n; ++k) v += D.ring[(slot + k) % SUQA_HIP_OUTCOME_SLOTS] << k;
    return v;
}

double ene_value(uint e) { return t_PE_shift + e / (double)(t_PE_factor * ene_levels); }

// One wait: runs what is queued and reads every slot handed out since the
// last read, then fills in the values that waited for them.
void read_slots() {
    if (!D.pending) return;
    const uint32_t S = SUQA_HIP_OUTCOME_SLOTS;
    const uint32_t n = (D.end + S - D.begin) % S;
    if (n == 0 || n > S / 2) throw std::runtime_error("ERROR: deferred readouts: outcome ring overrun");
    std::vector<uint32_t> out(n);
    chk(suqa_hip_read_outcomes(suqa::ctx, D.begin, n, out.data()), "qms::metro_step()");
    for (uint32_t k = 0; k < n; ++k) D.ring[(D.begin + k) % S] = out[k];
    D.pending = false;
    for (const auto& r : D.recs) {
        if (r.is_x) X_measures[r.idx] = X_value(slot_value(r.slot, 3));
        else E_measures[r.idx] = ene_value(slot_value(r.slot, ene_qbits));
    }
    D.recs.clear();
}

void push_x_deferred() {
    const uint32_t sx = measure_X_deferred(rangen);
    note_slots(sx, 3);
    D.recs.push_back(Deferred::Rec{true, sx, X_measures.size()});
    X_measures.push_back(0.0);
}

}  // namespace

void sync_measures() {
    if (deferred) read_slots();
}

void fill_bitmap() {
    bm_syst.resize(syst_qbits);
    bm_enes_old.resize(two_registers ? ene_qbits : 0);
    bm_enes_new.resize(ene_qbits);
    uint c = 0;
    for (uint i = 0; i < syst_qbits; ++i) bm_syst[i] = c++;
    for (uint& q : bm_enes_old) q = c++;
    for (uint i = 0; i < ene_qbits; ++i) bm_enes_new[i] = c++;
    bm_acc = c;
}

void fill_W_utils(double beta, double t_PE_factor_) {
    c_factor = beta / (t_PE_factor_ * ene_levels);
    W_mask_Enew = 0U;
    for (uint i = 0; i < ene_qbits; ++i) W_mask_Enew |= (1U << bm_enes_new[i]);
    W_mask_Eold = 0U;
    for (uint q : bm_enes_old) W_mask_Eold |= 1U << q;
    W_mask = (1U << bm_acc) | W_mask_Enew;
}

void cevolution(const double& t, const int& n, const uint& q_control, const bmReg& qsyst) {
    if (qsyst.size() != syst_qbits)
        throw std::runtime_error("ERROR: controlled evolution has wrong number of state qbits");
    suqa::activate_gc_mask({q_control});
    evolution(t, n);
    suqa::deactivate_gc_mask({q_control});
}

uint creg_to_uint(const std::vector<uint>& c_reg) {
    if (c_reg.size() < 1) throw std::runtime_error("ERROR: size of register zero.");
    uint ret = c_reg[0];
    for (uint j = 1U; j < c_reg.size(); ++j) ret += c_reg[j] << j;
    return ret;
}

void reset_non_syst_qbits() {
    std::vector<double> rgenerates(ene_qbits);
    if (two_registers) {   // the E_old register first (qms.cuh:134-136)
        for (auto& el : rgenerates) el = rangen.doub();
        suqa::apply_reset(bm_enes_old, rgenerates);
    }
    for (auto& el : rgenerates) el = rangen.doub();
    // the new register, then the acceptance qubit with the next draw: one call
    std::vector<uint> qs(bm_enes_new);
    qs.push_back(bm_acc);
    rgenerates.push_back(rangen.doub());
    suqa::apply_reset(qs, rgenerates);
}

void apply_phase_estimation(const std::vector<uint>& q_syst, const std::vector<uint>& q_target, const double& t,
                            const uint& n) {
    suqa::apply_h(q_target);
    for (int trg = (int)q_target.size() - 1; trg > -1; --trg) {
        double powr = (double)(1U << (q_target.size() - 1 - trg));
        cevolution(-powr * t, powr * n, q_target[trg], q_syst);
        suqa::apply_u1(q_target[trg], -powr * t * t_PE_shift);
    }
    suqa::apply_qft_inverse(q_target);
}

void apply_phase_estimation_inverse(const std::vector<uint>& q_syst, const std::vector<uint>& q_target,
                                    const double& t, const uint& n) {
    suqa::apply_qft(q_target);
    for (uint trg = 0; trg < q_target.size(); ++trg) {
        double powr = (double)(1U << (q_target.size() - 1 - trg));
        cevolution(powr * t, powr * n, q_target[trg], q_syst);
        suqa::apply_u1(q_target[trg], powr * t * t_PE_shift);
    }
    suqa::apply_h(q_target);
}

void apply_Phi() {
    replay(g_phi, [] { apply_phase_estimation(bm_syst, bm_enes_new, t_phase_estimation, n_phase_estimation); });
}
void apply_Phi_old() {
    replay(g_phi_old, [] { apply_phase_estimation(bm_syst, bm_enes_old, t_phase_estimation, n_phase_estimation); });
}
void apply_Phi_inverse() {
    replay(g_phi_inv,
           [] { apply_phase_estimation_inverse(bm_syst, bm_enes_new, t_phase_estimation, n_phase_estimation); });
}

uint draw_C() {
    std::vector<double> C_weigthsums = get_C_weigthsums();
    double extract = rangen.doub();
    for (uint Ci = 0U; Ci < C_weigthsums.size(); ++Ci)
        if (extract < C_weigthsums[Ci]) return Ci;
    return (uint)C_weigthsums.size();
}

// qms.cuh:312-329: one engine kernel; exp/sqrt of the acceptance rotation are
// evaluated on the host with libm (as the reference CPU path does).
void apply_W() {
    int rc = two_registers ? suqa_hip_qms_apply_W2(suqa::ctx, bm_acc, W_mask_Enew, bm_enes_new[0], W_mask_Eold,
                                                   bm_enes_old[0], c_factor)
                           : suqa_hip_qms_apply_W(suqa::ctx, bm_acc, W_mask_Enew, bm_enes_new[0], c_factor, curr_E_old);
    if (rc != SUQA_OK) throw std::runtime_error(std::string("ERROR: in qms::apply_W(): ") + suqa_hip_last_error());
    // counted as L (L - 1) / 2 Toffolis controlled by both energy registers,
    // gc_mask ignored (qms.cuh:323-328)
    if (!suqa::gatecounters.counters.empty()) {
        const GateRecord gr = GateCounter::n_ctrl_toffoli_gates(2 * ene_qbits);
        const uint ntoff = (ene_levels * (ene_levels - 1)) / 2;
        suqa::gatecounters.increment_g1g2(ntoff * gr.ng1, ntoff * gr.ng2);
    }
}

void apply_W_inverse() { apply_W(); }

void apply_U() {
    if (gCi < 3) {
        const uint ci = gCi;
        replay(g_cphi[ci], [ci] {
            apply_C(ci);
            apply_phase_estimation(bm_syst, bm_enes_new, t_phase_estimation, n_phase_estimation);
        });
    } else {
        apply_C(gCi);
        apply_Phi();
    }
    apply_W();
}

void apply_U_inverse() {
    apply_W_inverse();
    if (gCi < 3) {
        const uint ci = gCi;
        replay(g_phi_inv_c[ci], [ci] {
            apply_phase_estimation_inverse(bm_syst, bm_enes_new, t_phase_estimation, n_phase_estimation);
            apply_C_inverse(ci);
        });
    } else {
        apply_Phi_inverse();
        apply_C_inverse(gCi);
    }
}

std::vector<double> extract_rands(uint n) {
    std::vector<double> ret(n);
    for (auto& el : ret) el = rangen.doub();
    return ret;
}

namespace {

// metro_step with deferred readouts (no counters, no graphs): the draws,
// gates and measurements of metro_step below in the same order.
int metro_step_deferred(bool take_measure) {
    int ret = 0;
    reset_non_syst_qbits();
    uint32_t s_eold;
    if (two_registers) {   // E_old into its own register, read by W there
        apply_Phi_old();
        s_eold = measure_deferred(bm_enes_old, extract_rands(ene_qbits));
        gCi = draw_C();
        apply_U();
    } else {
        apply_Phi();
        s_eold = measure_deferred(bm_enes_new, extract_rands(ene_qbits));
        chk(suqa_hip_apply_x_if(suqa::ctx, bm_enes_new.data(), ene_qbits, s_eold, 0u), "suqa::apply_x()");
        gCi = draw_C();
        apply_C(gCi);   // apply_U: C_i, Phi, W with E_old read on the device
        apply_Phi();
        chk(suqa_hip_qms_apply_W_deferred(suqa::ctx, bm_acc, W_mask_Enew, bm_enes_new[0], c_factor, s_eold),
            "qms::apply_W()");
    }
    const uint32_t s_acc = measure_deferred({bm_acc}, {rangen.doub()});
    read_slots();   // the step's one wait
    curr_E_old = slot_value(s_eold, ene_qbits);
    curr_E_old_d = ene_value(curr_E_old);
    c_acc = D.ring[s_acc];

    if (c_acc == 1U) {
        uint32_t s_enew = 0;
        if (take_measure) s_enew = measure_deferred(bm_enes_new, extract_rands(ene_qbits));
        else measure_discarded(bm_enes_new, extract_rands(ene_qbits));
        apply_Phi_inverse();
        if (take_measure) {
            D.recs.push_back(Deferred::Rec{false, s_enew, E_measures.size()});
            E_measures.push_back(0.0);
            if (two_registers && suqa::gc_mask == 0u) {   // each E_old qubit reset with its draw before its flip
                const std::vector<double> rds = extract_rands(ene_qbits);   // (qms.cuh:431-434), in one call
                chk(suqa_hip_resets_then_x_if(suqa::ctx, bm_enes_old.data(), rds.data(), bm_enes_new.data(), ene_qbits,
                                              s_enew, 0u),
                    "suqa::apply_reset()");
            } else if (two_registers) {
                for (uint ei = 0U; ei < ene_qbits; ++ei) {
                    suqa::apply_reset(bm_enes_old[ei], rangen.doub());
                    chk(suqa_hip_apply_x_if(suqa::ctx, &bm_enes_new[ei], 1, s_enew + ei, 0u), "suqa::apply_x()");
                }
            } else {
                chk(suqa_hip_apply_x_if(suqa::ctx, bm_enes_new.data(), ene_qbits, s_enew, 0u), "suqa::apply_x()");
            }
            push_x_deferred();
            suqa::apply_reset(bm_enes_new, extract_rands(ene_qbits));
            apply_Phi();
            measure_discarded(bm_enes_new, extract_rands(ene_qbits));   // E re-collapse
            apply_Phi_inverse();
            if (!defer_values) read_slots();
            ret = 2;
        } else {
            ret = 1;
        }
        return ret;
    }

    apply_U_inverse();   // E_old is known on the host now
    uint iters = 0;
    while (iters < max_reverse_attempts) {
        apply_Phi();
        if (two_registers) (void)extract_rands(ene_qbits);   // the E_old readout's draws (qms.cuh:478-479)
        const uint32_t s_en = measure_deferred(bm_enes_new, extract_rands(ene_qbits));
        read_slots();
        const uint Enew_meas = slot_value(s_en, ene_qbits);
        apply_Phi_inverse();
        if (curr_E_old == Enew_meas) {
            if (take_measure) {
                E_measures.push_back(curr_E_old_d);
                suqa::apply_reset(bm_enes_new, extract_rands(ene_qbits));
                push_x_deferred();
                suqa::apply_reset(bm_enes_new, extract_rands(ene_qbits));
                apply_Phi();
                measure_discarded(bm_enes_new, extract_rands(ene_qbits));   // E re-collapse
                apply_Phi_inverse();
                if (!defer_values) read_slots();
                ret = 4;
            } else {
                ret = 3;
            }
            break;
        }
        apply_U();
        measure_discarded({bm_acc}, {rangen.doub()});   // the outcome is not used
        apply_U_inverse();
        iters++;
    }
    if (record_reverse) reverse_counters.push_back(iters);
    if (iters == max_reverse_attempts) return -1;
    return ret;
}

}  // namespace

int metro_step(bool take_measure) {
    // 1/2: accepted (not measured/measured), 3/4: rejected and restored,
    // -1: rejected and not restored within max_reverse_attempts.
    if (deferred && suqa::gatecounters.counters.empty()) return metro_step_deferred(take_measure);
    const bool counting = !suqa::gatecounters.counters.empty();
    if (counting) gctr_metrostep.new_record();
    int ret = 0;
    reset_non_syst_qbits();
    std::vector<uint> c_E_news(ene_qbits, 0), c_E_olds(ene_qbits, 0);
    uint Enew_meas;
    if (two_registers) {   // E_old into its own register, which keeps the measured value
        apply_Phi_old();
        suqa::measure_qbits(bm_enes_old, c_E_olds, extract_rands(ene_qbits));
    } else {
        apply_Phi();
        suqa::measure_qbits(bm_enes_new, c_E_olds, extract_rands(ene_qbits));
    }

    curr_E_old = creg_to_uint(c_E_olds);
    curr_E_old_d = t_PE_shift + curr_E_old / (double)(t_PE_factor * ene_levels);
    if (!two_registers)
        for (uint ei = 0U; ei < ene_qbits; ++ei)
            if (c_E_olds[ei]) suqa::apply_x(bm_enes_new[ei]);

    gCi = draw_C();
    apply_U();

    suqa::measure_qbit(bm_acc, c_acc, rangen.doub());

    if (c_acc == 1U) {
        double Enew_meas_d;
        if (take_measure) suqa::measure_qbits(bm_enes_new, c_E_news, extract_rands(ene_qbits));
        else measure_discarded(bm_enes_new, extract_rands(ene_qbits));   // E_new is not recorded
        apply_Phi_inverse();
        if (take_measure) {
            if (counting) gctr_measure.new_record();
            Enew_meas_d = t_PE_shift + creg_to_uint(c_E_news) / (double)(t_PE_factor * ene_levels);
            E_measures.push_back(Enew_meas_d);
            for (uint ei = 0U; ei < ene_qbits; ++ei) {
                // two registers: E_old qubit ei reset with one draw before each
                // flip (the "comparison with 2enereg" draw of qms.cuh:431-434)
                if (two_registers) suqa::apply_reset(bm_enes_old[ei], rangen.doub());
                if (c_E_news[ei]) suqa::apply_x(bm_enes_new[ei]);
            }
            X_measures.push_back(measure_X(rangen));
            suqa::apply_reset(bm_enes_new, extract_rands(ene_qbits));   // the same draws, in order
            apply_Phi();
            measure_discarded(bm_enes_new, extract_rands(ene_qbits));   // E re-collapse
            apply_Phi_inverse();
            if (counting) {
                gctr_measure.deactivate();
                gctr_sample.new_record();
            }
            ret = 2;
        } else {
            ret = 1;
        }
        return ret;
    }

    apply_U_inverse();
    uint iters = 0;
    if (counting) gctr_reverse.new_record();
    while (iters < max_reverse_attempts) {
        apply_Phi();
        // two registers: the E_old readout takes r draws (qms.cuh:478-479); the
        // register still holds the value measured at the start of the step
        if (two_registers) (void)extract_rands(ene_qbits);
        suqa::measure_qbits(bm_enes_new, c_E_news, extract_rands(ene_qbits));
        Enew_meas = creg_to_uint(c_E_news);
        apply_Phi_inverse();
        if (curr_E_old == Enew_meas) {
            if (take_measure) {
                E_measures.push_back(curr_E_old_d);
                suqa::apply_reset(bm_enes_new, extract_rands(ene_qbits));   // the same draws, in order
                X_measures.push_back(measure_X(rangen));
                suqa::apply_reset(bm_enes_new, extract_rands(ene_qbits));   // the same draws, in order
                apply_Phi();
                measure_discarded(bm_enes_new, extract_rands(ene_qbits));   // E re-collapse
                apply_Phi_inverse();
                ret = 4;
            } else {
                ret = 3;
            }
            break;
        }
        uint c_acc_trash;
        apply_U();
        suqa::measure_qbit(bm_acc, c_acc_trash, rangen.doub());
        apply_U_inverse();
        iters++;
    }
    if (counting) gctr_reverse.deactivate();
    if (record_reverse) reverse_counters.push_back(iters);
    if (iters == max_reverse_attempts) return -1;
    return ret;
}

void setup(double beta) {
    fill_bitmap();
    fill_W_utils(beta, t_PE_factor);
    drop_graphs();   // a new engine context holds none of them
    const char* e = std::getenv("SUQA_QMS_GRAPHS");
    use_graphs = e && std::strcmp(e, "1") == 0;
    const char* d = std::getenv("SUQA_QMS_DEFERRED");
    uint32_t probe = 0;
    deferred = !(d && std::strcmp(d, "0") == 0) && !use_graphs &&
               suqa::gatecounters.counters.empty() &&
               suqa_hip_measure_qbits_deferred(suqa::ctx, nullptr, nullptr, 0, &probe) == SUQA_OK;
    defer_values = false;   // a driver that syncs before reading turns it on after setup
    D = Deferred{};
}

void clear() {
    sync_measures();
    drop_graphs();
}

}  // namespace qms
