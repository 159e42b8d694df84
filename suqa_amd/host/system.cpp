// system.cpp -- frustrated-triangle plug-in on the suqa:: API (system.cu).
#include "system.hpp"

#include <cmath>
#include <stdexcept>
#include <string>

void init_state() {   // an E = -1 eigenstate: CX(1->0) H(1) X(1) |000>
    suqa::init_state();
    suqa::apply_x(bm_spin[1]);
    suqa::apply_h(bm_spin[1]);
    suqa::apply_cx(bm_spin[1], bm_spin[0]);
}

static void exp_it_id_x_x(const bmReg& q, uint pos_id, double phase_t) {
    suqa::apply_pauli_TP_rotation({q[(pos_id + 1) % 3], q[(pos_id + 2) % 3]}, {PAULI_X, PAULI_X}, phase_t);
}

void evolution(const double& t, const int& n) {   // the XX terms commute: exact for any n
    (void)n;
    for (uint iii = 0; iii < 3; ++iii) exp_it_id_x_x(bm_spin, iii, -t);
}

// Observable: basis change, measure the three spins, map to {2,0,2,0,-2,0,-2,0}
// (system.cu:70-119).  The "antirotation" applies the rotation again.
static const uint op_bits = 3;
static const std::vector<double> op_vals = {2.0, 0.0, 2.0, 0.0, -2.0, 0.0, -2.0, 0.0};

static void apply_measure_rotation() {
    suqa::apply_h(bm_spin[0]);
    suqa::apply_h(bm_spin[1]);
    suqa::apply_h(bm_spin[2]);
    suqa::apply_cx(bm_spin[0], bm_spin[1]);
    suqa::apply_cx(bm_spin[0], bm_spin[2]);
    suqa::apply_u1(bm_spin[2], M_PI * 0.5);
}

double measure_X(pcg& rgen) {
    std::vector<uint> classics(op_bits);
    apply_measure_rotation();
    std::vector<double> rdoubs(op_bits);
    for (auto& el : rdoubs) el = rgen.doub();
    suqa::measure_qbits(bm_spin, classics, rdoubs);
    apply_measure_rotation();
    uint meas = 0U;
    for (uint i = 0; i < op_bits; ++i) meas |= (classics[i] << i);
    return op_vals[meas];
}

uint32_t measure_X_deferred(pcg& rgen) {
    apply_measure_rotation();
    std::vector<double> rdoubs(op_bits);
    for (auto& el : rdoubs) el = rgen.doub();
    uint32_t slot = 0;
    if (suqa_hip_measure_qbits_deferred(suqa::ctx, bm_spin.data(), rdoubs.data(), op_bits, &slot) != SUQA_OK)
        throw std::runtime_error(std::string("ERROR: in measure_X(): ") + suqa_hip_last_error());
    apply_measure_rotation();
    return slot;
}

double X_value(uint meas) { return op_vals[meas]; }

static const std::vector<double> C_weigthsums = {1. / 3, 2. / 3, 1.0};

void apply_C(const uint& Ci) {
    if (Ci > 2) throw std::runtime_error("ERROR: wrong move selection");
    suqa::apply_h(bm_spin[Ci]);
}

void apply_C_inverse(const uint& Ci) { apply_C(Ci); }

std::vector<double> get_C_weigthsums() { return C_weigthsums; }

// ---- QSA hooks
void qsa_init_state() {   // spins in |+++>, copied onto the tilde spins by CX
    suqa::init_state();
    for (uint k = 0; k < 3; ++k) suqa::apply_h(bm_spin[k]);
    for (uint k = 0; k < 3; ++k) suqa::apply_cx(bm_spin[k], bm_spin_tilde[k]);
}

void evolution_szegedy(const double& t, const int& n) {   // tilde pairs by -t, spin pairs by +t, interleaved
    (void)n;
    for (uint i = 0; i < 3; i++) {
        suqa::apply_pauli_TP_rotation({bm_spin_tilde[i % 3], bm_spin_tilde[(1 + i) % 3]}, {PAULI_X, PAULI_X}, -t);
        suqa::apply_pauli_TP_rotation({bm_spin[i % 3], bm_spin[(1 + i) % 3]}, {PAULI_X, PAULI_X}, t);
    }
}

void evolution_measure(const double& t, const int& n) {
    (void)n;
    for (uint i = 0; i < 3; i++)
        suqa::apply_pauli_TP_rotation({bm_spin[i % 3], bm_spin[(1 + i) % 3]}, {PAULI_X, PAULI_X}, -t);
}

void evolution_tracing(const double& t, const int& n) {
    (void)n;
    for (uint i = 0; i < 3; i++)
        suqa::apply_pauli_TP_rotation({bm_spin_tilde[i % 3], bm_spin_tilde[(1 + i) % 3]}, {PAULI_X, PAULI_X}, -t);
}

void qsa_apply_C(const uint& Ci) {
    if (Ci > 2) throw std::runtime_error("ERROR: wrong move selection");
    suqa::apply_h(bm_spin_tilde[Ci]);
}

void qsa_apply_C_inverse(const uint& Ci) { qsa_apply_C(Ci); }
