// suqa.hpp -- the reference's suqa:: API (include/suqa.cuh:56-178), same
// names and signatures, on the MI355X engine of include/suqa_hip.h.
//
// Drop-in notes
//  - One global state, as in the reference (suqa.cu:17); suqa::state is a
//    ComplexVec view of the device SoA buffer (data_re = data,
//    data_im = data + N), so code that reaches into it (qms::apply_W in the
//    reference) keeps working on device pointers.
//  - Errors: every failing engine call throws std::runtime_error, the
//    reference's exception type (suqa.cu:421, 478, 555, 563, 579).
//  - Gates whose qubits lie in the active gc_mask give the reference's
//    results (its loops on coinciding slots), as the QSA layout needs;
//    three-qubit Pauli rotations there throw.
#pragma once
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "gatecounters.hpp"
#include "suqa_hip.h"

typedef unsigned int uint;
typedef std::vector<uint> bmReg;

#ifndef TWOSQINV
#define TWOSQINV 0.7071067811865475
#endif
#define PAULI_ID 0
#define PAULI_X 1
#define PAULI_Y 2
#define PAULI_Z 3

// complex_defines.cuh:22-45: SoA view, here of device memory.
struct ComplexVec {
    uint vecsize = 0U;
    double* data = nullptr;
    double* data_re = nullptr;
    double* data_im = nullptr;
    size_t size() const { return vecsize; }
};

namespace suqa {

extern ComplexVec state;
extern uint gc_mask;
extern uint nq;
extern std::vector<double> rphase_m;   // QFT phases, filled once (suqa.cu:611-625)
extern suqa_hip_ctx* ctx;              // the engine context behind the singleton
extern GateCounterList gatecounters;   // gate-count instrumentation (suqa.cuh:89-91)

void print_banner();

void activate_gc_mask(const bmReg& q_controls);
void deactivate_gc_mask(const bmReg& q_controls);

double vnorm();
void vnormalize();

void init_state();
void deallocate_state();
void allocate_state(uint Dim);

void apply_x(uint q);
void apply_x(const bmReg& qs);
void apply_y(uint q);
void apply_y(const bmReg& qs);
void apply_z(uint q);
void apply_z(const bmReg& qs);
void apply_h(uint q);
void apply_h(const bmReg& qs);
void apply_t(uint q);
void apply_t(const bmReg& qs);
void apply_tdg(uint q);
void apply_tdg(const bmReg& qs);
void apply_s(uint q);
void apply_s(const bmReg& qs);

void apply_u1(uint q, double phase);
void apply_u1(uint q, uint q_mask, double phase);

void apply_cx(const uint& q_control, const uint& q_target, const uint& q_mask = 1U);
void apply_mcx(const bmReg& q_controls, const uint& q_target);
void apply_mcx(const bmReg& q_controls, const bmReg& q_mask, const uint& q_target);
void apply_cu1(uint q_control, uint q_target, double phase, uint q_mask = 1U);
void apply_mcu1(const bmReg& q_controls, const uint& q_target, double phase);
void apply_mcu1(const bmReg& q_controls, const bmReg& q_mask, const uint& q_target, double phase);
void apply_swap(const uint& q1, const uint& q2);

void apply_phase_list(uint q0, uint q_size, const std::vector<double>& phases);

void apply_pauli_TP_rotation(const bmReg& q_apply, const std::vector<uint>& pauli_TPtype, double phase);

// Extensions (the rotation gates RX and RZ, and the reference's TODO at suqa.cu:459):
// RX(a) = exp(-i a X/2), RZ(a) = exp(-i a Z/2), any-weight Pauli rotation.
void apply_rx(uint q, double theta);
void apply_rz(uint q, double theta);
void apply_pauli_TP_rotation_general(const bmReg& q_apply, const std::vector<uint>& pauli_TPtype, double phase);

void apply_qft(const std::vector<uint>& qact);
void apply_qft_inverse(const std::vector<uint>& qact);

void measure_qbit(uint q, uint& c, double rdoub);
void measure_qbits(const bmReg& qs, std::vector<uint>& cs, const std::vector<double>& rdoubs);

void apply_reset(uint q, double rdoub);
void apply_reset(const bmReg& qs, std::vector<double> rdoubs);

void prob_filter(const bmReg& qs, const std::vector<uint>& q_mask, double& prob);

void setup(uint nq);
void clear();

}  // namespace suqa
