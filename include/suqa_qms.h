/*
 * suqa_qms.h -- C entry points of the QMS drivers on the MI355X engine.
 *
 * suqa_qms_run: the reference's QMS driver loop (src/qms.cu:51-178 with
 * the argument defaults of include/parser_qms.hpp:9-24) run on the plain
 * FP64 GPU path: one engine call per suqa:: operation, decisions on the host.
 * It is what the reference binary `qms` does, minus the file output, with the
 * per-step results handed back.
 *
 * suqa_qms_batch_run: many independent chains (frustrated triangle, uneven
 * energy-register widths, betas and seeds) in a few GPU launches (one per
 * group of equal register width, chains taken from a device work queue; the
 * launches run side by side on streams the call creates and destroys), each
 * chain following the same metro_step control flow and pcg32 draw order.
 * The call returns when every chain has finished.
 * No counterpart in the reference, which runs one chain per process.
 */
#ifndef SUQA_QMS_H
#define SUQA_QMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double   beta;            /* <beta> */
    uint32_t metro_steps;     /* <metro steps> */
    uint32_t reset_each;      /* <reset each> */
    uint32_t ene_qbits;       /* <num ene qbits> */
    uint32_t max_reverse;     /* --max-reverse (100) */
    uint64_t seed;            /* --seed (must be > 0 here) */
    double   ene_min;         /* --ene-min (0.0) */
    double   ene_max;         /* --ene-max (1.0) */
    int32_t  pe_steps;        /* --PE-steps (10) */
    int32_t  thermalization;  /* --thermalization (100) */
    int32_t  record_reverse;  /* --record-reverse */
    int32_t  two_registers;   /* --ene-registers 2 (extension): E_old in its own register (r <= 8) */
} suqa_qms_params;

typedef struct {
    int32_t*  rets;           /* [metro_steps] metro_step return codes, or NULL */
    uint64_t* rng_after;      /* [metro_steps] pcg32 state after each step, or NULL */
    double*   E;              /* [cap_meas] energy measurements, or NULL */
    double*   X;              /* [cap_meas] observable measurements, or NULL */
    uint32_t  cap_meas;
    uint32_t  n_meas;         /* out: measurements taken (may exceed cap_meas) */
    uint32_t* reverse;        /* [cap_reverse] reverse-attempt counters, or NULL */
    uint32_t  cap_reverse;
    uint32_t  n_reverse;      /* out */
    uint32_t  accepted;       /* out: steps with return code 1 or 2 */
    double    seconds;        /* out: wall time of the Metropolis loop */
    double*   final_re;       /* [2^nq] final state, or NULL */
    double*   final_im;
} suqa_qms_result;

/* Runs the driver loop of qms.cu:132-149 on the GPU engine. */
int suqa_qms_run(const suqa_qms_params* p, suqa_qms_result* out);

/* ---- batched chains -------------------------------------------------- */
typedef struct {
    double   beta;
    uint32_t ene_qbits;       /* 2..6: register width of this chain (state 2^(r+4) amplitudes) */
    uint64_t seed;            /* pcg32 seed (> 0) */
} suqa_qms_chain;

typedef struct {
    uint32_t metro_steps, reset_each, max_reverse;
    double   ene_min, ene_max;
    int32_t  pe_steps, thermalization;
    uint32_t fp32;            /* 0: FP64 chains (bit-exact with the reference CPU path);
                                 1: FP32 amplitudes, FP64 probabilities and collapse scale */
    uint32_t two_registers;   /* 1: E_old in its own register (the commented-out layout of
                                 qms.cuh:72-85, 134-136, 201-202, with the draws of its
                                 "comparison with 2enereg" lines 134-136, 433, 479):
                                 2r + 4 qubits per chain, r in 2..6, stored exactly in
                                 2^(r+4) amplitudes (E_old is definite outside its own
                                 phase estimation) */
} suqa_qms_batch_params;

/* Per chain outputs (arrays of nchains * metro_steps for rets; measurements
 * packed per chain at offsets chain * cap_meas). */
typedef struct {
    int32_t*  rets;           /* [nchains * metro_steps] */
    uint64_t* rng_after;      /* [nchains * metro_steps] or NULL */
    double*   E;              /* [nchains * cap_meas] */
    double*   X;              /* [nchains * cap_meas] */
    uint32_t  cap_meas;
    uint32_t* n_meas;         /* [nchains] */
    uint32_t* accepted;       /* [nchains] */
    double    seconds;        /* out: kernel wall time */
    uint64_t* lds_accesses;   /* [nchains] or NULL: algorithmic LDS bytes of each chain (the
                                 amplitude reads and writes its gates and measurements must
                                 make); 0 for chains of the one-amplitude-per-thread layout */
    double*   final_state;    /* [2 * sum over chains of 2^(ene_qbits + 4)] or NULL: each chain's
                                 stored amplitudes after its last step as FP64 (re, im) pairs,
                                 chains in input order (two registers: the compact state, with
                                 E_old held by the value its register reads) */
} suqa_qms_batch_result;

int suqa_qms_batch_run(const suqa_qms_chain* chains, uint32_t nchains, const suqa_qms_batch_params* p,
                       suqa_qms_batch_result* out);

#ifdef __cplusplus
}
#endif
#endif
