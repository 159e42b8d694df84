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
void apply_C(const uint& Ci);                    // system.cu:125-129
void apply_C_inverse(const uint& Ci);            // system.cu:131-133
std::vector<double> get_C_weigthsums();          // system.cu:151
