/*
 * suqa_qsa.h -- C entry point of the QSA driver on the MI355X engine.
 *
 * suqa_qsa_run: the reference's quantum simulated annealing driver loop
 * (src/qsa.cu:115-262, argument defaults of include/parser_qsa.hpp:8-27) on
 * the plain FP64 GPU path, one engine call per suqa:: operation.  It is what
 * the reference binary `qsa` does, minus the file output and the printing,
 * with the per-step results handed back.  max_annealing_steps > 0 stops the
 * loop once that many annealing steps have run (the stock driver's samples
 * can take ~10^5 restarts); 0 runs until `sampling` samples are taken.
 */
#ifndef SUQA_QSA_H
#define SUQA_QSA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double   beta;                 /* <beta> */
    uint32_t sampling;             /* <sampling> */
    uint32_t ene_qbits;            /* <num ene qbits> */
    uint32_t szegedy_qbits;        /* <num szegedy qbits> */
    uint64_t seed;                 /* --seed (must be > 0 here) */
    double   ene_min, ene_max;     /* --ene-min (-4.0), --ene-max (8.0) */
    double   ene_min_szegedy;      /* --ene-min-szegedy (-1.0) */
    double   ene_max_szegedy;      /* --ene-max-szegedy (3.0) */
    int32_t  pe_steps;             /* --PE-steps (1) */
    uint32_t annealing_sequences;  /* --annealing-sequences (100) */
    uint32_t max_annealing_steps;  /* bound on annealing steps (0: none) */
} suqa_qsa_params;

typedef struct {
    /* per annealing step k < min(n_steps, cap_steps): steps[5k .. 5k+4] =
     * sample index, restart index, step s, move drawn, Szegedy register
     * readout; rng_after[k] = pcg32 state after its measurement */
    uint32_t* steps;
    uint64_t* rng_after;
    uint32_t  cap_steps;
    uint32_t  n_steps;             /* out */
    /* per finished sample: energy, X value, restarts so far, pcg32 state */
    double*   E;
    double*   X;
    int32_t*  rejections;
    uint64_t* sample_rng;
    uint32_t  cap_samples;
    uint32_t  n_samples;           /* out */
    uint64_t  final_rng;           /* out: pcg32 state at the end */
    double*   final_re;            /* [2^nqubits] final state, or NULL */
    double*   final_im;
    double    seconds;             /* out: wall time of the loop */
} suqa_qsa_result;

int suqa_qsa_run(const suqa_qsa_params* p, suqa_qsa_result* out);

#ifdef __cplusplus
}
#endif
#endif
