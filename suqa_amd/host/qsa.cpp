// qsa.cpp -- qsa:: on the MI355X engine (include/qsa.cuh:16-772).
//
// Every function issues the reference's suqa:: calls in the reference's order
// with the reference's arguments: the gates are the physics of the algorithm,
// and the pcg32 draws must come in the same order for the same outcomes.  Where
// the reference's bitmap puts a gate on a qubit of the active gc_mask (the
// tilde spins share qubits 3..5 with the registers), the engine reproduces the
// reference's loops on coinciding slots (suqa_hip.h).
#include "qsa.hpp"

#include <cmath>
#include <stdexcept>

#include "system.hpp"

namespace qsa {

uint syst_qbits, ene_qbits, szegedy_qbits, ene_levels, nqubits, Dim;
unsigned long long iseed = 0ULL;
double t_phase_estimation, t_PE_factor, t_PE_shift;
double t_phase_estimation_szegedy, t_PE_factor_szegedy, t_PE_shift_szegedy;
int n_phase_estimation;
double beta;
uint W_mask_E = 0U;
uint gc_mask_szegedy = 0U;
pcg rangen(1ULL);
std::vector<double> E_measures, X_measures;
std::vector<double> rphase_m, rphase_m_szegedy;
std::vector<uint> bm_states, bm_enes, bm_szegedy;
uint bm_acc;

static std::vector<double> rphase_table(uint n) {   // 2 pi / 2^m, m = 0..n-1
    std::vector<double> r(n);
    uint c = 1;
    for (uint i = 0; i < n; ++i) {
        r[i] = (2.0 * M_PI / (double)c);
        c <<= 1;
    }
    return r;
}
void fill_rphase(const uint& n) { rphase_m = rphase_table(n); }
void fill_rphase_szegedy(const uint& n) { rphase_m_szegedy = rphase_table(n); }

// gc masks: the Szegedy control replaces the global mask; "into" adds controls
// on top of it and restores it after
static void counted_mask() { suqa::gatecounters.update_cmask_setbits(suqa::gc_mask); }
void activate_gc_mask_szegedy(const bmReg& q_controls) {
    gc_mask_szegedy = 0U;
    for (const auto& q : q_controls) gc_mask_szegedy |= 1U << q;
    suqa::gc_mask = gc_mask_szegedy;
    counted_mask();
}
void deactivate_gc_mask_szegedy() {
    gc_mask_szegedy = 0U;
    suqa::gc_mask = 0U;
    counted_mask();
}
void activate_gc_mask_into_szegedy(const bmReg& q_controls) {
    for (const auto& q : q_controls) suqa::gc_mask |= 1U << q;
    counted_mask();
}
void deactivate_gc_mask_into_szegedy() {
    suqa::gc_mask = gc_mask_szegedy;
    counted_mask();
}

void fill_bitmap() {   // states 0.., then E, then Szegedy, then the acceptance qubit
    bm_states.resize(syst_qbits);
    bm_enes.resize(ene_qbits);
    bm_szegedy.resize(szegedy_qbits);
    uint c = 0;
    for (auto& q : bm_states) q = c++;
    for (auto& q : bm_enes) q = c++;
    for (auto& q : bm_szegedy) q = c++;
    bm_acc = c;
    for (uint q : bm_enes) W_mask_E |= 1U << q;
}

void cevolution_measure(const double& t, const int& n, const uint& q_control, const bmReg& qstate) {
    if (qstate.size() != syst_qbits)
        throw std::runtime_error("ERROR: controlled evolution has wrong number of state qbits");
    suqa::activate_gc_mask({q_control});
    evolution_measure(t, n);
    suqa::deactivate_gc_mask({q_control});
}

void cevolution_szegedy_PE(const double& t, const int& n, const uint& q_control, const bmReg& qstate) {
    if (qstate.size() != syst_qbits)
        throw std::runtime_error("ERROR: controlled evolution has wrong number of state qbits");
    activate_gc_mask_into_szegedy({q_control});
    evolution_szegedy(t, n);
    deactivate_gc_mask_into_szegedy();
}

uint creg_to_uint(const std::vector<uint>& c_reg) {
    if (c_reg.size() < 1) throw std::runtime_error("ERROR: size of register zero.");
    uint ret = c_reg[0];
    for (uint j = 1U; j < c_reg.size(); ++j) ret += c_reg[j] << j;
    return ret;
}

void reset_non_syst_qbits() {   // E register draws, its resets, Szegedy draws, its resets, the acceptance qubit
    std::vector<double> r_e(ene_qbits), r_s(szegedy_qbits);
    for (auto& el : r_e) el = rangen.doub();
    suqa::apply_reset(bm_enes, r_e);
    for (auto& el : r_s) el = rangen.doub();
    suqa::apply_reset(bm_szegedy, r_s);
    suqa::apply_reset(bm_acc, rangen.doub());
}

// QFTs of qsa.cuh:177-247: the reference's no-swap loops on the per-namespace
// phase tables (filled by setup()), controlled by the current gc_mask
static void crm(const std::vector<double>& tab, const uint& qc, const uint& qt, const int& m) {
    double rphase = (m > 0) ? tab[m] : tab[-m];
    if (m <= 0) rphase *= -1;
    suqa::apply_cu1(qc, qt, rphase);
}
static void qft_on(const std::vector<double>& tab, const std::vector<uint>& qact) {
    const int qsize = (int)qact.size();
    for (int o = qsize - 1; o >= 0; o--) {
        suqa::apply_h(qact[o]);
        for (int in = o - 1; in >= 0; in--) crm(tab, qact[in], qact[o], +1 + (o - in));
    }
}
static void qft_inverse_on(const std::vector<double>& tab, const std::vector<uint>& qact) {
    const int qsize = (int)qact.size();
    for (int o = 0; o < qsize; o++) {
        for (int in = 0; in < o; in++) crm(tab, qact[in], qact[o], -1 - (o - in));
        suqa::apply_h(qact[o]);
    }
}
void qsa_crm(const uint& qc, const uint& qt, const int& m) { crm(rphase_m, qc, qt, m); }
void qsa_qft(const std::vector<uint>& qact) { qft_on(rphase_m, qact); }
void qsa_qft_inverse(const std::vector<uint>& qact) { qft_inverse_on(rphase_m, qact); }
void qsa_crm_szegedy(const uint& qc, const uint& qt, const int& m) { crm(rphase_m_szegedy, qc, qt, m); }
void qsa_qft_szegedy(const std::vector<uint>& qact) { qft_on(rphase_m_szegedy, qact); }
void qsa_qft_inverse_szegedy(const std::vector<uint>& qact) { qft_inverse_on(rphase_m_szegedy, qact); }

// phase estimation of the Szegedy evolution into the energy register (qsa.cuh:249-298)
void apply_phase_estimation(const std::vector<uint>& q_state, const std::vector<uint>& q_target, const double& t,
                            const uint& n) {
    suqa::apply_h(q_target);
    for (int trg = (int)q_target.size() - 1; trg > -1; --trg) {
        const double powr = (double)(1U << (q_target.size() - 1 - trg));
        cevolution_szegedy_PE(-powr * t, powr * n, q_target[trg], q_state);
        suqa::apply_u1(q_target[trg], -powr * t * t_PE_shift);
    }
    qsa_qft_inverse(q_target);
}

void apply_phase_estimation_inverse(const std::vector<uint>& q_state, const std::vector<uint>& q_target,
                                    const double& t, const uint& n) {
    qsa_qft(q_target);
    for (uint trg = 0; trg < q_target.size(); ++trg) {
        const double powr = (double)(1U << (q_target.size() - 1 - trg));
        cevolution_szegedy_PE(powr * t, powr * n, q_target[trg], q_state);
        suqa::apply_u1(q_target[trg], powr * t * t_PE_shift);
    }
    suqa::apply_h(q_target);
}

// phase estimation of the system's own evolution into the Szegedy register (qsa.cuh:300-338)
void apply_phase_estimation_measure(const std::vector<uint>& q_state, const std::vector<uint>& q_target,
                                    const double& t, const uint& n) {
    suqa::apply_h(q_target);
    for (int trg = (int)q_target.size() - 1; trg > -1; --trg) {
        const double powr = (double)(1U << (q_target.size() - 1 - trg));
        cevolution_measure(-powr * t, powr * n, q_target[trg], q_state);
        suqa::apply_u1(q_target[trg], -powr * t * t_PE_shift_szegedy);
    }
    qsa_qft_inverse_szegedy(q_target);
}

void apply_phase_estimation_measure_inverse(const std::vector<uint>& q_state, const std::vector<uint>& q_target,
                                            const double& t, const uint& n) {
    qsa_qft_szegedy(q_target);
    for (uint trg = 0; trg < q_target.size(); ++trg) {   // the u1 before the evolution, as the reference has it
        const double powr = (double)(1U << (q_target.size() - 1 - trg));
        suqa::apply_u1(q_target[trg], powr * t * t_PE_shift_szegedy);
        cevolution_measure(powr * t, powr * n, q_target[trg], q_state);
    }
    suqa::apply_h(q_target);
}

uint draw_C() {
    const std::vector<double> w = get_C_weigthsums();
    const double extract = rangen.doub();
    for (uint Ci = 0U; Ci < w.size(); ++Ci)
        if (extract < w[Ci]) return Ci;
    return (uint)w.size();
}

std::vector<double> extract_rands(uint n) {
    std::vector<double> ret(n);
    for (auto& el : ret) el = rangen.doub();
    return ret;
}

void apply_ctrl_x(int x) {   // X on every energy qubit whose bit of x is 0 (qsa.cuh:366-374)
    for (uint i = 0; i < ene_qbits; i++) {
        const int r = x % 2;
        if (r == 0) suqa::apply_x(bm_enes[i]);
        x = (x - r) / 2;
    }
}

double eigenvalue_trans(int x_2d) { return t_PE_shift + x_2d / (double)(t_PE_factor * ene_levels); }

// Y rotations of the acceptance qubit controlled by the energy register
// reading 2 (the reference's loops run for x_2d = 2 only, qsa.cuh:432-465)
static void controlled_acc_rotation(const double& beta_j, double sign) {
    for (int x_2d = 2; x_2d > 1; x_2d--) {
        const double tmp = eigenvalue_trans(x_2d);
        apply_ctrl_x(x_2d);
        activate_gc_mask_into_szegedy(bm_enes);
        suqa::apply_pauli_TP_rotation({bm_acc}, {PAULI_Y}, sign * acos(exp(-0.5 * beta_j * tmp)));
        deactivate_gc_mask_into_szegedy();
        apply_ctrl_x(x_2d);
    }
}
void universal_rotation(const double& beta_j) { controlled_acc_rotation(beta_j, -1.0); }
void universal_rotation_inverse(const double& beta_j) { controlled_acc_rotation(beta_j, 1.0); }

void universal_reflection() {   // a -1 phase on the acceptance qubit's |0> where the register reads 1 (qsa.cuh:466-485)
    apply_ctrl_x(1);
    activate_gc_mask_into_szegedy(bm_enes);
    suqa::apply_x(bm_acc);
    suqa::apply_u1(bm_acc, M_PI);
    suqa::apply_x(bm_acc);
    deactivate_gc_mask_into_szegedy();
    apply_ctrl_x(1);
}

static void acc_rotation(const double& beta_j, double sign) {   // qsa.cuh:562-589
    suqa::apply_x(bm_enes[0]);
    activate_gc_mask_into_szegedy(bm_enes);
    suqa::apply_pauli_TP_rotation({bm_acc}, {PAULI_Y}, sign * acos(exp(-2 * beta_j)));
    deactivate_gc_mask_into_szegedy();
    suqa::apply_x(bm_enes[0]);
}
void rotation(const double& beta_j) { acc_rotation(beta_j, -1.0); }
void rotation_inverse(const double& beta_j) { acc_rotation(beta_j, 1.0); }

void apply_ux(const double& beta_j, const uint& Ci) {   // qsa.cuh:591-606
    qsa_apply_C(Ci);
    apply_phase_estimation(bm_states, bm_enes, t_phase_estimation, n_phase_estimation);
    universal_rotation(beta_j);
    apply_phase_estimation_inverse(bm_states, bm_enes, t_phase_estimation, n_phase_estimation);
}

void apply_ux_inverse(const double& beta_j, const uint& Ci) {   // qsa.cuh:608-633
    apply_phase_estimation(bm_states, bm_enes, t_phase_estimation, n_phase_estimation);
    universal_rotation_inverse(beta_j);
    apply_phase_estimation_inverse(bm_states, bm_enes, t_phase_estimation, n_phase_estimation);
    qsa_apply_C_inverse(Ci);
}

// the spins swapped with their tilde copies where the acceptance qubit reads 0
static void swap_spins_on_acc0() {
    suqa::apply_x(bm_acc);
    activate_gc_mask_into_szegedy({bm_acc});
    for (uint j = 0; j < bm_spin.size(); ++j) suqa::apply_swap(bm_spin[j], bm_spin_tilde[j]);
    deactivate_gc_mask_into_szegedy();
    suqa::apply_x(bm_acc);
}

void apply_uy(const double& beta_j, const uint& Ci) {   // qsa.cuh:635-660
    apply_ux(beta_j, Ci);
    swap_spins_on_acc0();
}

void apply_uy_inverse(const double& beta_j, const uint& Ci) {   // qsa.cuh:662-688
    swap_spins_on_acc0();
    apply_ux_inverse(beta_j, Ci);
}

void reflection() {   // qsa.cuh:690-699
    suqa::apply_mcx(bm_enes, {1U, 0U}, bm_acc);
    suqa::apply_mcu1(bm_enes, {1U, 0U}, bm_acc, M_PI);
    suqa::apply_mcx(bm_enes, {1U, 0U}, bm_acc);
}

void apply_lambda1() {   // qsa.cuh:701-709
    apply_phase_estimation(bm_states, bm_enes, t_phase_estimation, n_phase_estimation);
    universal_reflection();
    apply_phase_estimation_inverse(bm_states, bm_enes, t_phase_estimation, n_phase_estimation);
}

void apply_szegedy(const double& beta_j, const uint& Ci) {   // qsa.cuh:711-720: the walk operator
    apply_lambda1();
    apply_ux(beta_j, Ci);
    apply_uy_inverse(beta_j, Ci);
    apply_lambda1();
    apply_uy(beta_j, Ci);
    apply_ux_inverse(beta_j, Ci);
}

void cevolution_szegedy(const uint& q_control, const bmReg& qstate, const int& powr, const double& beta_j,
                        const uint& Ci) {   // qsa.cuh:722-733: powr walk steps controlled by q_control
    if (qstate.size() != syst_qbits)
        throw std::runtime_error("ERROR: controlled evolution has wrong number of state qbits");
    activate_gc_mask_szegedy({q_control});
    for (int mm = 0; mm < powr; ++mm) apply_szegedy(beta_j, Ci);
    deactivate_gc_mask_szegedy();
}

void apply_PE_szegedy(const std::vector<uint>& q_state, const std::vector<uint>& q_target, const double& beta_j,
                      const uint& Ci) {   // qsa.cuh:735-757
    suqa::apply_h(q_target);
    for (int trg = (int)q_target.size() - 1; trg > -1; --trg) {
        const double powr = (double)(1U << (q_target.size() - 1 - trg));
        cevolution_szegedy(q_target[trg], q_state, (int)powr, beta_j, Ci);
    }
    qsa_qft_inverse_szegedy(q_target);
}

void setup() {
    fill_rphase(ene_qbits + 1);
    fill_rphase_szegedy(szegedy_qbits + 1);
    fill_bitmap();
}

void clear() {}

}  // namespace qsa
