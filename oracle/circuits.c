/*
 * circuits.c -- the synthetic workloads' generators, restated on the oracle
 * (TEST INFRASTRUCTURE ONLY).  The product has its own generator
 * (suqa_amd/host/workloads.cpp); tests check that both emit the same gates.
 *
 * Random circuit (gate-throughput workload): pcg(seed), per layer n gates:
 *   kind   = randint(0,7) over {H,X,RX,RZ,U1,CX,MCX3}
 *   target = randint(0,n); controls (1 for CX, 3 for MCX) drawn by
 *            rejection with randint(0,n), distinct from target and each other
 *   angle  = 2*pi*doub() for RX, RZ, U1
 * then min(8,n) distinct measured qubits (rejection), one doub() each.
 * RX(a) = pauli({q},{X},-a/2), RZ(a) = pauli({q},{Z},-a/2) (exp(-i a P/2)).
 *
 * Trotter + QPE workload: pcg(seed), per term: weight = randint(2,5),
 * distinct qubits by rejection in [0,nsys), types randint(1,4), coefficient
 * 2*doub()-1.  U(dt) = prod_terms exp(-i c dt P); QPE as qms.cuh:153-176
 * without the energy shift, controlled by register qubits nsys..nsys+nreg-1.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include "suqa_oracle.h"
#include "circuits.h"

uint32_t so_circuit_generate(uint32_t n, uint32_t layers, uint64_t seed, so_gate* out,
                             uint32_t* meas_q, double* meas_r) {
    so_pcg32 r;
    so_pcg32_seed(&r, seed, 1u);
    uint32_t ng = 0;
    for (uint32_t l = 0; l < layers; ++l) {
        for (uint32_t g = 0; g < n; ++g) {
            so_gate* G = &out[ng++];
            G->kind = so_pcg32_randint(&r, 0, 7);
            G->target = (uint32_t)so_pcg32_randint(&r, 0, (int)n);
            G->nctrl = G->kind == SO_G_CX ? 1u : (G->kind == SO_G_MCX ? 3u : 0u);
            G->ctrl[0] = G->ctrl[1] = G->ctrl[2] = 0u;
            uint32_t have = 0;
            while (have < G->nctrl) {
                uint32_t c = (uint32_t)so_pcg32_randint(&r, 0, (int)n);
                int dup = (c == G->target);
                for (uint32_t k = 0; k < have; ++k) dup |= (G->ctrl[k] == c);
                if (!dup) G->ctrl[have++] = c;
            }
            G->angle = (G->kind >= SO_G_RX && G->kind <= SO_G_U1) ? 2.0 * M_PI * so_pcg32_doub(&r) : 0.0;
        }
    }
    uint32_t nm = n < 8 ? n : 8;
    uint32_t have = 0;
    while (have < nm) {
        uint32_t q = (uint32_t)so_pcg32_randint(&r, 0, (int)n);
        int dup = 0;
        for (uint32_t k = 0; k < have; ++k) dup |= (meas_q[k] == q);
        if (!dup) meas_q[have++] = q;
    }
    for (uint32_t k = 0; k < nm; ++k) meas_r[k] = so_pcg32_doub(&r);
    return ng;
}

void so_circuit_apply(so_state* s, const so_gate* G) {
    static const uint32_t tx = 1u, tz = 3u;
    switch (G->kind) {
    case SO_G_H: so_apply_h(s, G->target); break;
    case SO_G_X: so_apply_x(s, G->target); break;
    case SO_G_RX: so_apply_pauli_TP_rotation(s, &G->target, &tx, 1, -G->angle / 2.0); break;
    case SO_G_RZ: so_apply_pauli_TP_rotation(s, &G->target, &tz, 1, -G->angle / 2.0); break;
    case SO_G_U1: so_apply_u1(s, G->target, 1u, G->angle); break;
    case SO_G_CX: so_apply_cx(s, G->ctrl[0], G->target, 1u); break;
    case SO_G_MCX: so_apply_mcx(s, G->ctrl, NULL, 3u, G->target); break;
    default: break;
    }
}

void so_trotter_terms_generate(uint32_t nsys, uint32_t nterms, uint64_t seed, so_pauli_term* out) {
    so_pcg32 r;
    so_pcg32_seed(&r, seed, 1u);
    for (uint32_t t = 0; t < nterms; ++t) {
        so_pauli_term* T = &out[t];
        T->w = (uint32_t)so_pcg32_randint(&r, 2, 5);
        uint32_t have = 0;
        while (have < T->w) {
            uint32_t q = (uint32_t)so_pcg32_randint(&r, 0, (int)nsys);
            int dup = 0;
            for (uint32_t k = 0; k < have; ++k) dup |= (T->qs[k] == q);
            if (!dup) T->qs[have++] = q;
        }
        for (uint32_t k = 0; k < T->w; ++k) T->types[k] = (uint32_t)so_pcg32_randint(&r, 1, 4);
        T->coef = 2.0 * so_pcg32_doub(&r) - 1.0;
    }
}

void so_trotter_step(so_state* s, const so_pauli_term* terms, uint32_t nterms, double dt) {
    for (uint32_t t = 0; t < nterms; ++t)
        so_apply_pauli_general(s, terms[t].qs, terms[t].types, terms[t].w, -terms[t].coef * dt);
}

void so_qpe_trotter(so_state* s, uint32_t nsys, uint32_t nreg, const so_pauli_term* terms,
                    uint32_t nterms, double dt, uint32_t steps) {
    uint32_t reg[32];
    for (uint32_t k = 0; k < nreg; ++k) reg[k] = nsys + k;
    for (uint32_t k = 0; k < nreg; ++k) so_apply_h(s, reg[k]);
    for (int trg = (int)nreg - 1; trg > -1; --trg) {
        uint32_t powr = 1u << (nreg - 1 - (uint32_t)trg);
        so_activate_gc_mask(s, &reg[trg], 1);
        for (uint32_t p = 0; p < powr * steps; ++p) so_trotter_step(s, terms, nterms, dt);
        so_deactivate_gc_mask(s, &reg[trg], 1);
    }
    so_apply_qft_inverse(s, reg, nreg);
}
