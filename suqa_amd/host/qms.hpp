// qms.hpp -- the reference's qms:: namespace (include/qms.cuh:30-566): the
// Metropolis step of quantum Metropolis sampling on the suqa:: API, master
// branch layout (one energy register).  Same globals, same functions.
#pragma once
#include <vector>

#include "pcg.hpp"
#include "suqa.hpp"

namespace qms {

extern uint syst_qbits, syst_levels, ene_qbits, ene_levels, nqubits, Dim;
extern uint max_reverse_attempts, metro_steps, reset_each;
extern unsigned long long iseed;
extern double t_phase_estimation, t_PE_factor, t_PE_shift;
extern int n_phase_estimation;
extern uint gCi, c_acc;
extern bool record_reverse;
extern std::vector<uint> reverse_counters;
extern pcg rangen;
extern std::vector<double> X_measures, E_measures;
extern double c_factor;
extern std::vector<uint> bm_syst, bm_enes_new;
extern uint bm_acc;
extern uint W_mask, W_mask_Enew;
extern uint curr_E_old;
extern double curr_E_old_d;
// Two-register layout (the commented bm_enes_old code of qms.cuh:72-85,
// 134-136, 201-202; extension, default off): E_old is phase-estimated into
// its own register (qubits 3..2+r, E_new moves to 3+r..2+2r) and stays there;
// the reset draws r doubles for it first; W reads E_old from the register.
extern bool two_registers;
extern std::vector<uint> bm_enes_old;
extern uint W_mask_Eold;
// Extension: the fixed launch sequences of a step (Phi, Phi^dag, C_i Phi,
// Phi^dag C_i) captured once per run and replayed (suqa_hip_graph_*); setup()
// turns it on when SUQA_QMS_GRAPHS=1.  Off by default: each sequence is one or
// two kernel launches, and a graph launch costs more host time than they do.
// Never with gate counters.
extern bool use_graphs;
// Extension: deferred readouts.  On a state of at most 2^14 amplitudes the
// step keeps the outcomes that only the device needs on the device: the E_old
// readout feeds the X flips and apply_W there (two registers: W reads the
// register), the flips of E_new read their outcomes there, and the E and X
// values of a measured step are read at a later wait, so a step waits once
// (the acceptance) plus once per reverse attempt.  The same draws, gates and
// measurements in the same order, so the same accept/reject sequence and
// values.  setup() turns it on unless SUQA_QMS_DEFERRED=0; never with gate
// counters or graphs.
extern bool deferred;
// With deferred readouts a step's E and X values (E_measures, X_measures)
// exist on the host only after a wait.  By default metro_step waits for them
// before it returns (a measured step then waits twice), so the vectors are
// complete after every step, as with the reference.  A driver that reads them
// only through sync_measures() (the bundled suqa_qms and suqa_qms_run) sets
// defer_values after setup() and lets them wait for the next step's wait.
extern bool defer_values;
// Resolves the E and X values of deferred steps (E_measures, X_measures).
void sync_measures();

// Gate-count records of the Metropolis step (qms.cuh:21-27, defined in
// qms.cu:33-39): used when a driver registers them in suqa::gatecounters.
extern GateCounter gctr_global, gctr_metrostep, gctr_sample, gctr_measure, gctr_reverse;

void fill_bitmap();                                                    // qms.cuh:79-88
void fill_W_utils(double beta, double t_PE_factor);                    // qms.cuh:96-103
void cevolution(const double& t, const int& n, const uint& q_control, const bmReg& qsyst);
uint creg_to_uint(const std::vector<uint>& c_reg);
void reset_non_syst_qbits();
void apply_phase_estimation(const std::vector<uint>& q_syst, const std::vector<uint>& q_target, const double& t,
                            const uint& n);
void apply_phase_estimation_inverse(const std::vector<uint>& q_syst, const std::vector<uint>& q_target,
                                    const double& t, const uint& n);
void apply_Phi();
void apply_Phi_inverse();
void apply_Phi_old();   // qms.cuh:201 (two-register layout)
uint draw_C();
void apply_W();
void apply_W_inverse();
void apply_U();
void apply_U_inverse();
std::vector<double> extract_rands(uint n);
int metro_step(bool take_measure);                                     // qms.cuh:373-545
void setup(double beta);
void clear();

}  // namespace qms
