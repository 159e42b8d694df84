/*
 * suqa_oracle.h -- CPU restatement of SUQA's dense FP64 state-vector engine,
 * its PCG32 random source and the QMS Metropolis step (frustrated triangle).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle the HIP engine is tested against:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (suqa_amd/) never links or calls it.
 *
 * Every function follows the reference file:line given in its comment
 * (paths relative to the reference repository root).  Arithmetic keeps the
 * reference's operation order; build with -ffp-contract=off so that no
 * multiply-add is fused (the reference CPU build targets baseline x86-64,
 * which has no FMA instruction).
 *
 * Parity pinning (see tests/test_oracle_*.py):
 *   - every gate, measurement and prob_filter: golden vectors written by the
 *     reference compiled from source (oracle/_ref), plus closed forms;
 *   - pcg32: golden sequence from the reference + published PCG32 outputs;
 *   - metro_step / QMS driver: golden per-step traces from the reference;
 *   - so_pauli_general (weights >= 4 and the weight-3 patterns the
 *     reference rejects): dense exp(i*phi*P) built independently in numpy
 *     and the reference's own special cases where they overlap.
 */
#ifndef SUQA_ORACLE_H
#define SUQA_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* PCG32 (pcg32.h:53-68, 121-130) and the `pcg` wrapper (Rand.hpp:249-291) */
typedef struct {
    uint64_t state;
    uint64_t inc;
} so_pcg32;

void     so_pcg32_seed(so_pcg32* r, uint64_t initstate, uint64_t initseq);
uint32_t so_pcg32_next(so_pcg32* r);
double   so_pcg32_doub(so_pcg32* r);
int      so_pcg32_randint(so_pcg32* r, int x0, int x1);

/* ------------------------------------------------------------------ */
/* State vector: SoA, data[0..N) = real parts, data[N..2N) = imaginary parts
 * (suqa.cu:772-788, complex_defines.cuh:22-45).                          */
typedef struct {
    uint32_t nq;
    uint64_t N;
    double*  data;
    double*  re;
    double*  im;
    uint32_t gc_mask;           /* suqa.cu:36 */
    int      rphase_filled;     /* suqa.cu:621 one-shot table */
    uint32_t rphase_len;
    double   rphase_m[65];      /* suqa.cu:611-618 */
} so_state;

so_state* so_create(uint32_t nq);                /* suqa.cu:792-844 */
void      so_destroy(so_state* s);               /* suqa.cu:846-866 */
void      so_init_state(so_state* s);            /* suqa_cpu.hpp:5-13 */
double    so_vnorm(so_state* s);                 /* suqa.cu:78-92 */
void      so_vnormalize(so_state* s);            /* suqa.cu:97-111 */
void      so_activate_gc_mask(so_state* s, const uint32_t* qs, uint32_t nqs);   /* suqa.cu:45-52 */
void      so_deactivate_gc_mask(so_state* s, const uint32_t* qs, uint32_t nqs); /* suqa.cu:54-61 */

void so_apply_x(so_state* s, uint32_t q);                          /* suqa.cu:115-127 */
void so_apply_y(so_state* s, uint32_t q);                          /* suqa.cu:136-148 */
void so_apply_h(so_state* s, uint32_t q);                          /* suqa.cu:207-218 */
void so_apply_u1(so_state* s, uint32_t q, uint32_t q_mask, double phase);  /* suqa.cu:271-285 */
void so_apply_cx(so_state* s, uint32_t qc, uint32_t qt, uint32_t q_mask);  /* suqa.cu:289-304 */
void so_apply_mcx(so_state* s, const uint32_t* qc, const uint32_t* qmask,
                  uint32_t nc, uint32_t qt);                        /* suqa.cu:307-344 */
void so_apply_cu1(so_state* s, uint32_t qc, uint32_t qt, double phase,
                  uint32_t q_mask);                                 /* suqa.cu:346-364 */
void so_apply_mcu1(so_state* s, const uint32_t* qc, const uint32_t* qmask,
                   uint32_t nc, uint32_t qt, double phase);         /* suqa.cu:366-393 */
void so_apply_swap(so_state* s, uint32_t q1, uint32_t q2);         /* suqa.cu:396-416 */
int  so_apply_phase_list(so_state* s, uint32_t q0, uint32_t q_size,
                         const double* phases, uint32_t nphases);   /* suqa.cu:419-454, documented intent */
/* returns 0, or -1 where the reference throws (suqa.cu:478,555,563,579) */
int  so_apply_pauli_TP_rotation(so_state* s, const uint32_t* qs, const uint32_t* types,
                                uint32_t n, double phase);          /* suqa.cu:467-604 */
/* any weight: basis change + CNOT ladder + Z rotation (Nielsen-Chuang 4.19) */
int  so_apply_pauli_general(so_state* s, const uint32_t* qs, const uint32_t* types,
                            uint32_t n, double phase);
void so_apply_qft(so_state* s, const uint32_t* qs, uint32_t n);          /* suqa.cu:635-647 */
void so_apply_qft_inverse(so_state* s, const uint32_t* qs, uint32_t n);  /* suqa.cu:649-661 */

double   so_prob1(so_state* s, uint32_t q);                        /* suqa_cpu.hpp:554-573 */
uint32_t so_measure_qbit(so_state* s, uint32_t q, double rdoub);   /* suqa.cu:673-696 */
void     so_apply_reset(so_state* s, uint32_t q, double rdoub);    /* suqa.cu:730-738 */
double   so_prob_filter(so_state* s, const uint32_t* qs, const uint32_t* q_mask,
                        uint32_t n);                                /* suqa.cu:705-727 */

/* ------------------------------------------------------------------ */
/* QMS (qms.cuh, system.cu, qms.cu) -- master branch, one energy register. */
typedef struct {
    so_state* st;
    so_pcg32  rng;
    uint64_t  seed;
    uint64_t  draws;          /* number of doub() calls so far (diagnostic) */
    uint32_t  syst_qbits, ene_qbits, ene_levels, nqubits;
    uint32_t  max_reverse_attempts;
    double    t_phase_estimation, t_PE_factor, t_PE_shift;
    int       n_phase_estimation;
    uint32_t  gCi, c_acc;
    uint32_t  bm_syst[3];
    uint32_t  bm_enes[32];
    uint32_t  bm_acc;
    uint32_t  W_mask, W_mask_Enew;
    /* two-register layout (the commented bm_enes_old code, qms.cuh:72-85):
     * E_old in its own register between the system and E_new */
    uint32_t  two_registers;
    uint32_t  bm_enes_old[32];
    uint32_t  W_mask_Eold;
    /* one register with the extra draws of the reference's "comparison with
     * 2enereg" debug lines (qms.cuh:134-136, 433, 479): the draws the
     * two-register chain makes, at the same points of metro_step */
    uint32_t  compare_draws;
    double    c_factor;
    uint32_t  curr_E_old;
    double    curr_E_old_d;
    int       record_reverse;
    /* measurements pushed by metro_step (qms.cuh:55-56) */
    double*   E_measures;
    double*   X_measures;
    size_t    n_meas, cap_meas;
    uint32_t* reverse_counters;
    size_t    n_rev, cap_rev;
} so_qms;

/* Builds the QMS context exactly as qms.cu:51-101 does (seed>0 required). */
so_qms* so_qms_create(double beta, uint32_t ene_qbits, uint32_t max_reverse,
                      uint64_t seed, double ene_min, double ene_max, int pe_steps,
                      int record_reverse);
/* two_registers == 1: two energy registers.  Bitmap syst 0..2, E_old
 * 3..2+r, E_new 3+r..2+2r, acceptance 3+2r.  Draws at the points of the
 * reference's "comparison with 2enereg" lines:
 *   - the reset draws r doubles for E_old first and resets it
 *     (qms.cuh:134-136);
 *   - E_old is phase-estimated into its register and stays there; W reads
 *     E_old from the register;
 *   - an accepted, measured step resets E_old qubit ei with one draw before
 *     the flip of E_new qubit ei (qms.cuh:431-434);
 *   - each restoration attempt reads E_old out after its Phi: r draws
 *     (qms.cuh:478-479); the register holds the value measured at the start
 *     of the step (only diagonal reads have acted on it since), so the
 *     readout is that value.
 * two_registers == SO_QMS_COMPARE_DRAWS: one register with those draws (the
 * debug lines as written: nothing reset at 134-136, E_new qubit ei reset at
 * 433, r draws discarded at 479). */
#define SO_QMS_COMPARE_DRAWS 2
so_qms* so_qms_create2(double beta, uint32_t ene_qbits, uint32_t max_reverse,
                       uint64_t seed, double ene_min, double ene_max, int pe_steps,
                       int record_reverse, int two_registers);
void    so_qms_destroy(so_qms* q);
void    so_qms_init_state(so_qms* q);                  /* system.cu:6-11 */
void    so_qms_evolution(so_qms* q, double t, int n);  /* system.cu:21-26 */
double  so_qms_measure_X(so_qms* q);                   /* system.cu:99-119 */
int     so_qms_metro_step(so_qms* q, int take_measure);/* qms.cuh:373-545 */
void    so_qms_apply_Phi(so_qms* q);                   /* qms.cuh:203 */
void    so_qms_apply_Phi_inverse(so_qms* q);           /* qms.cuh:204 */
void    so_qms_apply_W(so_qms* q);                     /* qms.cuh:245-309 dense */

/* Driver loop of qms.cu:132-149 without file output: rets[s] receives the
 * return code of step s; returns the number of accepted steps. */
uint32_t so_qms_run(so_qms* q, uint32_t metro_steps, uint32_t reset_each,
                    int thermalization, int* rets);

#ifdef __cplusplus
}
#endif
#endif
