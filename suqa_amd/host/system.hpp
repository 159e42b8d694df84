// system.hpp -- the system plug-in of the master branch (frustrated triangle),
// same interface as the reference's include/system.cuh:17-30 and the hooks
// qms.cuh:8-17 declares: H = X0X1 + X1X2 + X0X2 on spins 0..2.
#pragma once
#include <vector>

#include "pcg.hpp"
#include "suqa.hpp"

const int syst_qbits = 3;
const bmReg bm_spin = {0, 1, 2};

void init_state();                               // system.cu:6-11
void evolution(const double& t, const int& n);   // system.cu:21-26, exp(-iHt)
double measure_X(pcg& rgen);                     // system.cu:99-119
// measure_X with the spins' outcomes left on the device (extension): the same
// gates and draws; returns the outcome slot of spin 0 (spins 1, 2 follow,
// suqa_hip_measure_qbits_deferred); X_value maps the outcomes (bit i = spin i)
// to the observable as measure_X does.
uint32_t measure_X_deferred(pcg& rgen);
double X_value(uint meas);
void apply_C(const uint& Ci);                    // system.cu:125-129
void apply_C_inverse(const uint& Ci);            // system.cu:131-133
std::vector<double> get_C_weigthsums();          // system.cu:151

// QSA hooks (system.cuh:32-41, system.cu:29-66): the second copy of the spins
// and the Szegedy-walk evolutions.  The reference's QSA keeps syst_qbits = 3,
// so bm_spin_tilde shares qubits 3..5 with its energy and Szegedy registers.
const bmReg bm_spin_tilde = {3, 4, 5};
void qsa_init_state();                                  // system.cu:29-37
void evolution_szegedy(const double& t, const int& n);  // system.cu:39-52
void evolution_measure(const double& t, const int& n);  // system.cu:54-59
void evolution_tracing(const double& t, const int& n);  // system.cu:60-65
void qsa_apply_C(const uint& Ci);                       // system.cu:136-143
void qsa_apply_C_inverse(const uint& Ci);               // system.cu:145-150
