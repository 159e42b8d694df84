// qsa.hpp -- the reference's qsa:: namespace (include/qsa.cuh:16-772):
// quantum simulated annealing on the suqa:: API, a Szegedy walk whose phase
// estimation reads an energy register, on the frustrated-triangle plug-in.
// Same globals and functions; the gate sequences are the reference's, so a run
// with a fixed pcg32 seed reproduces the reference's outcomes.
#pragma once
#include <vector>

#include "pcg.hpp"
#include "suqa.hpp"

namespace qsa {

extern uint syst_qbits, ene_qbits, szegedy_qbits, ene_levels, nqubits, Dim;
extern unsigned long long iseed;
extern double t_phase_estimation, t_PE_factor, t_PE_shift;
extern double t_phase_estimation_szegedy, t_PE_factor_szegedy, t_PE_shift_szegedy;
extern int n_phase_estimation;
extern double beta;
extern uint W_mask_E;
extern uint gc_mask_szegedy;
extern pcg rangen;
extern std::vector<double> E_measures, X_measures;
extern std::vector<double> rphase_m, rphase_m_szegedy;
extern std::vector<uint> bm_states, bm_enes, bm_szegedy;
extern uint bm_acc;

void fill_rphase(const uint& nqubits);                                 // qsa.cuh:48-55
void fill_rphase_szegedy(const uint& nqubits);                         // qsa.cuh:56-63
void activate_gc_mask_szegedy(const bmReg& q_controls);                // qsa.cuh:65-75
void deactivate_gc_mask_szegedy();                                     // qsa.cuh:77-83
void activate_gc_mask_into_szegedy(const bmReg& q_controls);           // qsa.cuh:85-92
void deactivate_gc_mask_into_szegedy();                                // qsa.cuh:94-99
void fill_bitmap();                                                    // qsa.cuh:109-123
void cevolution_measure(const double& t, const int& n, const uint& q_control, const bmReg& qstate);
void cevolution_szegedy_PE(const double& t, const int& n, const uint& q_control, const bmReg& qstate);
uint creg_to_uint(const std::vector<uint>& c_reg);
void reset_non_syst_qbits();                                           // qsa.cuh:162-175
void qsa_crm(const uint& q_control, const uint& q_target, const int& m);
void qsa_qft(const std::vector<uint>& qact);
void qsa_qft_inverse(const std::vector<uint>& qact);
void qsa_crm_szegedy(const uint& q_control, const uint& q_target, const int& m);
void qsa_qft_szegedy(const std::vector<uint>& qact);
void qsa_qft_inverse_szegedy(const std::vector<uint>& qact);
void apply_phase_estimation(const std::vector<uint>& q_state, const std::vector<uint>& q_target, const double& t,
                            const uint& n);
void apply_phase_estimation_inverse(const std::vector<uint>& q_state, const std::vector<uint>& q_target,
                                    const double& t, const uint& n);
void apply_phase_estimation_measure(const std::vector<uint>& q_state, const std::vector<uint>& q_target,
                                    const double& t, const uint& n);
void apply_phase_estimation_measure_inverse(const std::vector<uint>& q_state, const std::vector<uint>& q_target,
                                            const double& t, const uint& n);
uint draw_C();
std::vector<double> extract_rands(uint n);
void apply_ctrl_x(int x);
double eigenvalue_trans(int x_2d);
void universal_rotation(const double& beta_j);
void universal_rotation_inverse(const double& beta_j);
void universal_reflection();
void rotation(const double& beta_j);
void rotation_inverse(const double& beta_j);
void apply_ux(const double& beta_j, const uint& Ci);
void apply_ux_inverse(const double& beta_j, const uint& Ci);
void apply_uy(const double& beta_j, const uint& Ci);
void apply_uy_inverse(const double& beta_j, const uint& Ci);
void reflection();
void apply_lambda1();
void apply_szegedy(const double& beta_j, const uint& Ci);
void cevolution_szegedy(const uint& q_control, const bmReg& qstate, const int& powr, const double& beta_j,
                        const uint& Ci);
void apply_PE_szegedy(const std::vector<uint>& q_state, const std::vector<uint>& q_target, const double& beta_j,
                      const uint& Ci);
void setup();                                                          // qsa.cuh:760-765
void clear();

}  // namespace qsa
