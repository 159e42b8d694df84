/* circuits.h -- synthetic workload generators on the oracle (test infrastructure only). */
#ifndef SUQA_ORACLE_CIRCUITS_H
#define SUQA_ORACLE_CIRCUITS_H
#include <stdint.h>
#include "suqa_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { SO_G_H = 0, SO_G_X = 1, SO_G_RX = 2, SO_G_RZ = 3, SO_G_U1 = 4, SO_G_CX = 5, SO_G_MCX = 6 };

typedef struct {
    int32_t  kind;
    uint32_t target;
    uint32_t nctrl;
    uint32_t ctrl[3];
    double   angle;
} so_gate;

typedef struct {
    uint32_t w;
    uint32_t qs[4];
    uint32_t types[4];
    double   coef;
} so_pauli_term;

/* out must hold layers*n gates; meas_q/meas_r must hold min(8,n) entries. */
uint32_t so_circuit_generate(uint32_t n, uint32_t layers, uint64_t seed, so_gate* out,
                             uint32_t* meas_q, double* meas_r);
void     so_circuit_apply(so_state* s, const so_gate* g);
void     so_trotter_terms_generate(uint32_t nsys, uint32_t nterms, uint64_t seed, so_pauli_term* out);
void     so_trotter_step(so_state* s, const so_pauli_term* terms, uint32_t nterms, double dt);
void     so_qpe_trotter(so_state* s, uint32_t nsys, uint32_t nreg, const so_pauli_term* terms,
                        uint32_t nterms, double dt, uint32_t steps);

#ifdef __cplusplus
}
#endif
#endif
