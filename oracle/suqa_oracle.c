/*
 * suqa_oracle.c -- CPU restatement of the reference's dense FP64 path.
 * TEST INFRASTRUCTURE ONLY (the oracle tests compare against, never the product); see the header.
 *
 * Build: gcc -O3 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 * Loops keep the reference's order: full sweeps over i = 0..N-1 with a
 * bit-mask predicate, sequential FP64 sums, host sincos from libm.
 */
#define _GNU_SOURCE
#include "suqa_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* suqa.cuh:57 */
#define SO_TWOSQINV 0.7071067811865475
#define SO_PAULI_X 1u
#define SO_PAULI_Y 2u
#define SO_PAULI_Z 3u

/* ---------------------------------------------------------------- PCG32 */
#define SO_PCG32_MULT 0x5851f42d4c957f2dULL

uint32_t so_pcg32_next(so_pcg32* r) {            /* pcg32.h:62-68 */
    uint64_t old = r->state;
    r->state = old * SO_PCG32_MULT + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
}

void so_pcg32_seed(so_pcg32* r, uint64_t initstate, uint64_t initseq) { /* pcg32.h:53-59 */
    r->state = 0u;
    r->inc = (initseq << 1u) | 1u;
    so_pcg32_next(r);
    r->state += initstate;
    so_pcg32_next(r);
}

double so_pcg32_doub(so_pcg32* r) {              /* pcg32.h:121-130, Rand.hpp:271-273 */
    union { uint64_t u; double d; } x;
    x.u = ((uint64_t)so_pcg32_next(r) << 20) | 0x3ff0000000000000ULL;
    return x.d - 1.0;
}

int so_pcg32_randint(so_pcg32* r, int x0, int x1) { /* Rand.hpp:275-277 */
    return (int)((unsigned)x0 + so_pcg32_next(r) % (unsigned)(x1 - x0));
}

/* ------------------------------------------------------------ helpers */
/* suqa_cpu.hpp:49-54 */
static inline void rot2(double* a, double* b, double c, double s) {
    double cpy = *a;
    *a = cpy * c - (*b) * s;
    *b = (*b) * c + cpy * s;
}
/* suqa_cpu.hpp:56-66 */
static inline void rot4(double* a, double* b, double* c, double* d, double ct, double st) {
    rot2(a, b, ct, st);
    rot2(c, d, ct, st);
}
static inline void swapd(double* a, double* b) { double t = *a; *a = *b; *b = t; }

/* ------------------------------------------------------- lifecycle */
so_state* so_create(uint32_t nq) {
    so_state* s = (so_state*)calloc(1, sizeof(so_state));
    if (!s) return NULL;
    s->nq = nq;
    s->N = (uint64_t)1 << nq;
    s->data = (double*)malloc(2 * s->N * sizeof(double));
    if (!s->data) { free(s); return NULL; }
    s->re = s->data;
    s->im = s->data + s->N;
    s->gc_mask = 0u;
    so_init_state(s);
    return s;
}

void so_destroy(so_state* s) {
    if (!s) return;
    free(s->data);
    free(s);
}

void so_init_state(so_state* s) {                /* suqa_cpu.hpp:5-13 */
    for (uint64_t i = 1; i < 2 * s->N; ++i) s->data[i] = 0.0;
    s->data[0] = 1.0;
}

double so_vnorm(so_state* s) {                   /* suqa.cu:78-92, suqa_cpu.hpp:15-28 */
    double ret = 0.0;
    for (uint64_t i = 0; i < 2 * s->N; ++i) ret += s->data[i] * s->data[i];
    return sqrt(ret);
}

void so_vnormalize(so_state* s) {                /* suqa.cu:97-111, suqa_cpu.hpp:30-40 */
    double v = so_vnorm(s);
    double f = 1. / v;
    for (uint64_t i = 0; i < 2 * s->N; ++i) s->data[i] *= f;
}

void so_activate_gc_mask(so_state* s, const uint32_t* qs, uint32_t nqs) {
    for (uint32_t k = 0; k < nqs; ++k) s->gc_mask |= 1u << qs[k];
}
void so_deactivate_gc_mask(so_state* s, const uint32_t* qs, uint32_t nqs) {
    for (uint32_t k = 0; k < nqs; ++k) s->gc_mask &= ~(1u << qs[k]);
}

/* ------------------------------------------------------------ gates */
void so_apply_x(so_state* s, uint32_t q) {       /* suqa_cpu.hpp:69-94 */
    uint32_t gm = s->gc_mask | (1u << q);
    for (uint64_t i = 0; i < s->N; ++i) {
        if ((((uint32_t)i) & gm) == gm) {
            uint64_t j = i & ~(uint64_t)(1u << q);
            swapd(&s->re[i], &s->re[j]);
            swapd(&s->im[i], &s->im[j]);
        }
    }
}

void so_apply_y(so_state* s, uint32_t q) {       /* suqa_cpu.hpp:96-122 */
    uint32_t gm = s->gc_mask | (1u << q);
    for (uint64_t i = 0; i < s->N; ++i) {
        if ((((uint32_t)i) & gm) == gm) {
            uint64_t j = i & ~(uint64_t)(1u << q);
            double cpy = s->re[i];
            s->re[i] = s->im[j] * -1.0;
            s->im[j] = cpy * -1.0;
            cpy = s->im[i];
            s->im[i] = s->re[j];
            s->re[j] = cpy;
        }
    }
}

void so_apply_h(so_state* s, uint32_t q) {       /* suqa_cpu.hpp:150-206 dense */
    uint32_t gm = s->gc_mask;
    uint32_t lm = gm | (1u << q);
    for (uint64_t i0 = 0; i0 < s->N; ++i0) {
        if ((((uint32_t)i0) & lm) == gm) {
            uint64_t i1 = i0 | (uint64_t)(1u << q);
            double a0r = s->re[i0], a0i = s->im[i0];
            double a1r = s->re[i1], a1i = s->im[i1];
            s->re[i0] = SO_TWOSQINV * (a0r + a1r);
            s->im[i0] = SO_TWOSQINV * (a0i + a1i);
            s->re[i1] = SO_TWOSQINV * (a0r - a1r);
            s->im[i1] = SO_TWOSQINV * (a0i - a1i);
        }
    }
}

/* suqa_cpu.hpp:208-221 (predicate (i & (gc|1<<q)) == gc|(q_mask<<q)) */
static void so_u1_masked(so_state* s, uint32_t q, double c, double sn, uint32_t qmask) {
    uint32_t gm = s->gc_mask | (1u << q);
    for (uint64_t i = 0; i < s->N; ++i)
        if ((((uint32_t)i) & gm) == qmask) rot2(&s->re[i], &s->im[i], c, sn);
}

void so_apply_u1(so_state* s, uint32_t q, uint32_t q_mask, double phase) { /* suqa.cu:271-285 */
    uint32_t qmask = s->gc_mask | (q_mask << q);
    double sn, c;
    sincos(phase, &sn, &c);
    so_u1_masked(s, q, c, sn, qmask);
}

/* suqa_cpu.hpp:223-253 */
static void so_mcx_masks(so_state* s, uint32_t cmask, uint32_t mqs, uint32_t qt) {
    cmask |= 1u << qt;
    mqs |= 1u << qt;
    for (uint64_t i = 0; i < s->N; ++i) {
        if ((((uint32_t)i) & cmask) == mqs) {
            uint64_t j = i & ~(uint64_t)(1u << qt);
            swapd(&s->re[i], &s->re[j]);
            swapd(&s->im[i], &s->im[j]);
        }
    }
}

void so_apply_cx(so_state* s, uint32_t qc, uint32_t qt, uint32_t q_mask) { /* suqa.cu:289-304 */
    uint32_t mqs = s->gc_mask;
    uint32_t mask = mqs | (1u << qc);
    if (q_mask) mqs |= 1u << qc;
    so_mcx_masks(s, mask, mqs, qt);
}

void so_apply_mcx(so_state* s, const uint32_t* qc, const uint32_t* qmask, uint32_t nc,
                  uint32_t qt) {                 /* suqa.cu:307-344 */
    uint32_t mask = s->gc_mask, mqs = s->gc_mask;
    for (uint32_t k = 0; k < nc; ++k) {
        mask |= 1u << qc[k];
        if (!qmask || qmask[k]) mqs |= 1u << qc[k];
    }
    so_mcx_masks(s, mask, mqs, qt);
}

/* suqa_cpu.hpp:255-266 */
static void so_mcu1_masks(so_state* s, uint32_t cmask, uint32_t mqs, double c, double sn) {
    for (uint64_t i = 0; i < s->N; ++i)
        if ((((uint32_t)i) & cmask) == mqs) rot2(&s->re[i], &s->im[i], c, sn);
}

void so_apply_cu1(so_state* s, uint32_t qc, uint32_t qt, double phase, uint32_t q_mask) {
    /* suqa.cu:346-364 */
    uint32_t mqs = (1u << qt) | s->gc_mask;
    uint32_t mask = mqs | (1u << qc);
    if (q_mask) mqs |= 1u << qc;
    double sn, c;
    sincos(phase, &sn, &c);
    so_mcu1_masks(s, mask, mqs, c, sn);
}

void so_apply_mcu1(so_state* s, const uint32_t* qc, const uint32_t* qmask, uint32_t nc,
                   uint32_t qt, double phase) {  /* suqa.cu:366-393 */
    uint32_t mask = (1u << qt) | s->gc_mask;
    uint32_t mqs = (1u << qt) | s->gc_mask;
    for (uint32_t k = 0; k < nc; ++k) {
        mask |= 1u << qc[k];
        if (!qmask || qmask[k]) mqs |= 1u << qc[k];
    }
    double sn, c;
    sincos(phase, &sn, &c);
    so_mcu1_masks(s, mask, mqs, c, sn);
}

void so_apply_swap(so_state* s, uint32_t q1, uint32_t q2) {   /* suqa.cu:396-416, suqa_cpu.hpp:303-317 */
    uint32_t m00 = s->gc_mask;
    uint32_t m1 = 1u << q1, m2 = 1u << q2;
    uint32_t m11 = m00 | m1 | m2;
    for (uint64_t i = 0; i < s->N; ++i) {
        if ((((uint32_t)i) & m11) == m00) {
            uint64_t i1 = i | m1, i2 = i | m2;
            swapd(&s->re[i1], &s->re[i2]);
            swapd(&s->im[i1], &s->im[i2]);
        }
    }
}

/* suqa.cu:419-454 + suqa.cuh:157 (documented intent): phase phases[k] on every
 * basis state whose bits q0..q0+q_size-1 read k and whose gc_mask bits are all
 * set.  The reference GPU kernel tests (i & gc_mask) != 0 instead, which is a
 * no-op when gc_mask == 0; both agree when gc_mask has exactly one bit. */
int so_apply_phase_list(so_state* s, uint32_t q0, uint32_t q_size, const double* phases,
                        uint32_t nphases) {
    if (!(q_size > 0u && q_size <= 7u && nphases == (1u << q_size))) return -1;
    uint32_t size_mask = (1u << q_size) - 1u;
    double cs[128], sn[128];
    for (uint32_t k = 0; k < nphases; ++k) sincos(phases[k], &sn[k], &cs[k]);
    uint32_t gm = s->gc_mask;
    for (uint64_t i = 0; i < s->N; ++i) {
        if ((((uint32_t)i) & gm) == gm) {
            uint32_t k = (((uint32_t)i) >> q0) & size_mask;
            rot2(&s->re[i], &s->im[i], cs[k], sn[k]);
        }
    }
    return 0;
}

/* --------------------------------------------- Pauli tensor rotations */
/* suqa_cpu.hpp:338-389 (dense) */
static void so_pauli1(so_state* s, uint32_t p1, uint32_t m0, uint32_t m1s, uint32_t mq1,
                      double ct, double st) {
    double* re = s->re;
    double* im = s->im;
    for (uint64_t i0 = 0; i0 < s->N; ++i0) {
        if ((((uint32_t)i0) & m1s) != m0) continue;
        uint64_t i1 = i0 | mq1;
        switch (p1) {
        case 1: rot4(&re[i0], &im[i1], &re[i1], &im[i0], ct, st); break;
        case 2: rot4(&re[i0], &re[i1], &im[i0], &im[i1], ct, -st); break;
        case 3: rot4(&re[i0], &im[i0], &im[i1], &re[i1], ct, st); break;
        default: break;
        }
    }
}

/* suqa_cpu.hpp:392-463 (dense) */
static void so_pauli2(so_state* s, uint32_t p1, uint32_t p2, uint32_t m0, uint32_t m1s,
                      uint32_t mq1, uint32_t mq2, double ct, double st) {
    double* re = s->re;
    double* im = s->im;
    for (uint64_t i0 = 0; i0 < s->N; ++i0) {
        if ((((uint32_t)i0) & m1s) != m0) continue;
        uint64_t a = i0, b = i0 | mq1, c = i0 | mq2, d = b | c;
        if (p1 == 1 && p2 == 1) {
            rot4(&re[a], &im[d], &re[d], &im[a], ct, st);
            rot4(&re[b], &im[c], &re[c], &im[b], ct, st);
        } else if (p1 == 2 && p2 == 2) {
            rot4(&re[a], &im[d], &re[d], &im[a], ct, -st);
            rot4(&re[b], &im[c], &re[c], &im[b], ct, st);
        } else if (p1 == 3 && p2 == 3) {
            rot4(&re[a], &im[a], &re[d], &im[d], ct, st);
            rot4(&re[b], &im[b], &re[c], &im[c], ct, -st);
        } else if (p1 == 1 && p2 == 2) {
            rot4(&re[a], &re[d], &im[a], &im[d], ct, -st);
            rot4(&re[c], &re[b], &im[c], &im[b], ct, -st);
        } else if (p1 == 1 && p2 == 3) {
            rot4(&re[a], &im[b], &re[b], &im[a], ct, st);
            rot4(&re[c], &im[d], &re[d], &im[c], ct, -st);
        } else if (p1 == 2 && p2 == 3) {
            rot4(&re[a], &re[b], &im[a], &im[b], ct, -st);
            rot4(&re[c], &re[d], &im[c], &im[d], ct, st);
        }
    }
}

/* suqa_cpu.hpp:465-527 (dense) */
static void so_pauli3(so_state* s, uint32_t p1, uint32_t p2, uint32_t p3, uint32_t m0,
                      uint32_t m1s, uint32_t mq1, uint32_t mq2, uint32_t mq3, double ct,
                      double st) {
    double* re = s->re;
    double* im = s->im;
    for (uint64_t i0 = 0; i0 < s->N; ++i0) {
        if ((((uint32_t)i0) & m1s) != m0) continue;
        uint64_t k[8];
        k[0] = i0;
        k[1] = k[0] | mq1;
        k[2] = k[0] | mq2;
        k[3] = k[2] | k[1];
        k[4] = k[0] | mq3;
        k[5] = k[4] | k[1];
        k[6] = k[4] | k[2];
        k[7] = k[4] | k[3];
        if (p1 == 3 && p2 == 1 && p3 == 1) {
            rot4(&re[k[0]], &im[k[3]], &re[k[3]], &im[k[0]], ct, st);
            rot4(&re[k[1]], &im[k[2]], &re[k[2]], &im[k[1]], ct, st);
            rot4(&re[k[4]], &im[k[7]], &re[k[7]], &im[k[4]], ct, -st);
            rot4(&re[k[5]], &im[k[6]], &re[k[6]], &im[k[5]], ct, -st);
        } else if (p1 == 3 && p2 == 2 && p3 == 2) {
            rot4(&re[k[0]], &im[k[3]], &re[k[3]], &im[k[0]], ct, -st);
            rot4(&re[k[1]], &im[k[2]], &re[k[2]], &im[k[1]], ct, st);
            rot4(&re[k[4]], &im[k[7]], &re[k[7]], &im[k[4]], ct, st);
            rot4(&re[k[5]], &im[k[6]], &re[k[6]], &im[k[5]], ct, -st);
        } else if (p1 == 3 && p2 == 3 && p3 == 3) {
            rot4(&re[k[0]], &im[k[0]], &re[k[3]], &im[k[3]], ct, st);
            rot4(&re[k[5]], &im[k[5]], &re[k[6]], &im[k[6]], ct, st);
            rot4(&re[k[1]], &im[k[1]], &re[k[2]], &im[k[2]], ct, -st);
            rot4(&re[k[4]], &im[k[4]], &re[k[7]], &im[k[7]], ct, -st);
        }
    }
}

int so_apply_pauli_TP_rotation(so_state* s, const uint32_t* qs, const uint32_t* types,
                               uint32_t n, double phase) {   /* suqa.cu:467-604 */
    uint32_t m0 = s->gc_mask, m1s = m0;
    for (uint32_t k = 0; k < n; ++k) m1s |= 1u << qs[k];
    double sph, cph;
    sincos(phase, &sph, &cph);
    uint32_t t[3], q[3];
    if (n == 1) {
        so_pauli1(s, types[0], m0, m1s, 1u << qs[0], cph, sph);
        return 0;
    }
    if (n == 2) {
        t[0] = types[0]; t[1] = types[1]; q[0] = qs[0]; q[1] = qs[1];
        if (t[0] > t[1]) { uint32_t x = t[0]; t[0] = t[1]; t[1] = x; x = q[0]; q[0] = q[1]; q[1] = x; }
        so_pauli2(s, t[0], t[1], m0, m1s, 1u << q[0], 1u << q[1], cph, sph);
        return 0;
    }
    if (n == 3) {
        t[0] = types[0]; t[1] = types[1]; t[2] = types[2];
        q[0] = qs[0]; q[1] = qs[1]; q[2] = qs[2];
        uint32_t x;
        if (t[0] > t[1]) { x = t[0]; t[0] = t[1]; t[1] = x; x = q[0]; q[0] = q[1]; q[1] = x; }
        if (t[1] > t[2]) { x = t[1]; t[1] = t[2]; t[2] = x; x = q[1]; q[1] = q[2]; q[2] = x; }
        if (t[0] > t[1]) { x = t[0]; t[0] = t[1]; t[1] = x; x = q[0]; q[0] = q[1]; q[1] = x; }
        if (t[2] != SO_PAULI_Z) return -1;
        if (t[0] != t[1]) return -1;
        so_pauli3(s, t[2], t[0], t[1], m0, m1s, 1u << q[0], 1u << q[1], 1u << q[2], cph, sph);
        return 0;
    }
    return -1;
}

/* exp(i*phase*P) for any Pauli string by the textbook circuit: rotate each
 * factor to Z (X: H, Y: H.S^dag), CNOT ladder onto the last qubit,
 * exp(i*phase*Z) there, undo.  Uses only reference operations, so every
 * gc_mask control is honoured as the reference would. Identity factors are
 * skipped. */
int so_apply_pauli_general(so_state* s, const uint32_t* qs, const uint32_t* types, uint32_t n,
                           double phase) {
    uint32_t act[32];
    uint32_t na = 0;
    if (n > 32) return -1;
    for (uint32_t k = 0; k < n; ++k) {
        if (types[k] > 3u) return -1;
        if (types[k] != 0u) act[na++] = k;
    }
    if (na == 0) {
        /* exp(i phase) * I: a global phase, applied on the controlled subspace */
        double sn, c;
        sincos(phase, &sn, &c);
        uint32_t gm = s->gc_mask;
        for (uint64_t i = 0; i < s->N; ++i)
            if ((((uint32_t)i) & gm) == gm) rot2(&s->re[i], &s->im[i], c, sn);
        return 0;
    }
    const double half_pi = M_PI * 0.5;
    for (uint32_t a = 0; a < na; ++a) {
        uint32_t k = act[a];
        if (types[k] == SO_PAULI_X) so_apply_h(s, qs[k]);
        else if (types[k] == SO_PAULI_Y) { so_apply_u1(s, qs[k], 1u, -half_pi); so_apply_h(s, qs[k]); }
    }
    for (uint32_t a = 0; a + 1 < na; ++a) so_apply_cx(s, qs[act[a]], qs[act[a + 1]], 1u);
    {
        uint32_t zt = SO_PAULI_Z;
        uint32_t ql = qs[act[na - 1]];
        so_apply_pauli_TP_rotation(s, &ql, &zt, 1u, phase);
    }
    for (uint32_t a = na - 1; a >= 1; --a) so_apply_cx(s, qs[act[a - 1]], qs[act[a]], 1u);
    for (uint32_t a = 0; a < na; ++a) {
        uint32_t k = act[a];
        if (types[k] == SO_PAULI_X) so_apply_h(s, qs[k]);
        else if (types[k] == SO_PAULI_Y) { so_apply_h(s, qs[k]); so_apply_u1(s, qs[k], 1u, half_pi); }
    }
    return 0;
}

/* ------------------------------------------------------------- QFT */
static void so_fill_rphase(so_state* s, uint32_t nqubits) {   /* suqa.cu:611-618 */
    uint32_t c = 1;
    s->rphase_len = nqubits + 1;
    for (uint32_t i = 0; i < nqubits + 1 && i < 65; ++i) {
        s->rphase_m[i] = (2.0 * M_PI / (double)c);
        c <<= 1;
    }
}

static void so_qft_crm(so_state* s, uint32_t qc, uint32_t qt, int m, int nqubits) { /* suqa.cu:620-633 */
    if (!s->rphase_filled) {
        so_fill_rphase(s, (uint32_t)nqubits);
        s->rphase_filled = 1;
    }
    double rphase = (m > 0) ? s->rphase_m[m] : s->rphase_m[-m];
    if (m <= 0) rphase *= -1;
    so_apply_cu1(s, qc, qt, rphase, 1u);
}

void so_apply_qft(so_state* s, const uint32_t* qs, uint32_t n) {   /* suqa.cu:635-647 */
    int qsize = (int)n;
    for (int o = qsize - 1; o >= 0; o--) {
        so_apply_h(s, qs[o]);
        for (int in = o - 1; in >= 0; in--) so_qft_crm(s, qs[in], qs[o], +1 + (o - in), qsize);
    }
}

void so_apply_qft_inverse(so_state* s, const uint32_t* qs, uint32_t n) {  /* suqa.cu:649-661 */
    int qsize = (int)n;
    for (int o = 0; o < qsize; o++) {
        for (int in = 0; in < o; in++) so_qft_crm(s, qs[in], qs[o], -1 - (o - in), qsize);
        so_apply_h(s, qs[o]);
    }
}

/* ------------------------------------------------------- measurement */
double so_prob1(so_state* s, uint32_t q) {       /* suqa_cpu.hpp:554-573 */
    double ret = 0.0, t;
    for (uint64_t i = 0; i < s->N; ++i) {
        if (((uint32_t)i) & (1u << q)) {
            t = s->re[i];
            ret += t * t;
            t = s->im[i];
            ret += t * t;
        }
    }
    return ret;
}

static void so_set_ampl_to_zero(so_state* s, uint32_t q, uint32_t val) { /* suqa_cpu.hpp:532-551 */
    for (uint64_t i = 0; i < s->N; ++i) {
        if (((((uint32_t)i) >> q) & 1u) == val) {
            s->re[i] = 0.0;
            s->im[i] = 0.0;
        }
    }
}

uint32_t so_measure_qbit(so_state* s, uint32_t q, double rdoub) {  /* suqa.cu:673-696 */
    double p1 = so_prob1(s, q);
    uint32_t c = (uint32_t)(rdoub < p1);
    so_set_ampl_to_zero(s, q, c ^ 1u);
    so_vnormalize(s);
    return c;
}

void so_apply_reset(so_state* s, uint32_t q, double rdoub) {   /* suqa.cu:730-738 */
    if (so_measure_qbit(s, q, rdoub)) so_apply_x(s, q);
}

double so_prob_filter(so_state* s, const uint32_t* qs, const uint32_t* q_mask, uint32_t n) {
    /* suqa.cu:705-727, suqa_cpu.hpp:575-591 */
    uint32_t mqs = 0u, mask = 0u;
    for (uint32_t k = 0; k < n; ++k) mqs |= 1u << qs[k];
    for (uint32_t k = 0; k < n; ++k)
        if (q_mask[k]) mask |= q_mask[k] << qs[k];
    double ret = 0.0, t;
    for (uint64_t i = 0; i < s->N; ++i) {
        if ((((uint32_t)i) & mqs) == mask) {
            t = s->re[i];
            ret += t * t;
            t = s->im[i];
            ret += t * t;
        }
    }
    return ret;
}

/* =================================================================== QMS */
static double so_qms_doub(so_qms* q) {
    q->draws++;
    return so_pcg32_doub(&q->rng);
}

so_qms* so_qms_create(double beta, uint32_t ene_qbits, uint32_t max_reverse, uint64_t seed,
                      double ene_min, double ene_max, int pe_steps, int record_reverse) {
    return so_qms_create2(beta, ene_qbits, max_reverse, seed, ene_min, ene_max, pe_steps, record_reverse, 0);
}

so_qms* so_qms_create2(double beta, uint32_t ene_qbits, uint32_t max_reverse, uint64_t seed,
                       double ene_min, double ene_max, int pe_steps, int record_reverse, int two_registers) {
    if (seed == 0 || ene_qbits == 0 || ene_qbits > 28 || two_registers < 0 || two_registers > SO_QMS_COMPARE_DRAWS ||
        (two_registers == 1 && ene_qbits > 13))
        return NULL;
    so_qms* q = (so_qms*)calloc(1, sizeof(so_qms));
    if (!q) return NULL;
    /* qms.cu:56-84 */
    q->max_reverse_attempts = max_reverse;
    q->n_phase_estimation = pe_steps;
    q->record_reverse = record_reverse;
    q->seed = seed;
    so_pcg32_seed(&q->rng, seed, 1u);
    q->syst_qbits = 3;
    q->ene_qbits = ene_qbits;
    q->two_registers = two_registers == 1 ? 1u : 0u;
    q->compare_draws = two_registers == SO_QMS_COMPARE_DRAWS ? 1u : 0u;
    q->nqubits = q->syst_qbits + (1u + q->two_registers) * ene_qbits + 1;
    q->ene_levels = 1u << ene_qbits;
    q->t_PE_shift = ene_min;
    q->t_PE_factor = (double)(q->ene_levels - 1u) / ((double)q->ene_levels * (ene_max - ene_min));
    q->t_phase_estimation = q->t_PE_factor * 8. * atan(1.0);
    q->st = so_create(q->nqubits);
    if (!q->st) { free(q); return NULL; }
    /* qms.cuh:79-103 */
    uint32_t c = 0;
    for (uint32_t i = 0; i < q->syst_qbits; ++i) q->bm_syst[i] = c++;
    if (q->two_registers)
        for (uint32_t i = 0; i < ene_qbits; ++i) q->bm_enes_old[i] = c++;
    for (uint32_t i = 0; i < ene_qbits; ++i) q->bm_enes[i] = c++;
    q->bm_acc = c;
    q->W_mask_Eold = 0u;
    if (q->two_registers)
        for (uint32_t i = 0; i < ene_qbits; ++i) q->W_mask_Eold |= 1u << q->bm_enes_old[i];
    q->c_factor = beta / (q->t_PE_factor * q->ene_levels);
    q->W_mask_Enew = 0u;
    for (uint32_t i = 0; i < ene_qbits; ++i) q->W_mask_Enew |= 1u << q->bm_enes[i];
    q->W_mask = (1u << q->bm_acc) | q->W_mask_Enew;
    return q;
}

void so_qms_destroy(so_qms* q) {
    if (!q) return;
    so_destroy(q->st);
    free(q->E_measures);
    free(q->X_measures);
    free(q->reverse_counters);
    free(q);
}

static void so_qms_push_E(so_qms* q, double e) {
    if (q->n_meas >= q->cap_meas) {
        size_t nc = q->cap_meas ? 2 * q->cap_meas : 64;
        q->E_measures = (double*)realloc(q->E_measures, nc * sizeof(double));
        q->X_measures = (double*)realloc(q->X_measures, nc * sizeof(double));
        q->cap_meas = nc;
    }
    q->E_measures[q->n_meas] = e;
}
static void so_qms_push_X(so_qms* q, double x) { q->X_measures[q->n_meas++] = x; }

void so_qms_init_state(so_qms* q) {              /* system.cu:6-11 */
    so_init_state(q->st);
    so_apply_x(q->st, 1);
    so_apply_h(q->st, 1);
    so_apply_cx(q->st, 1, 0, 1u);
}

void so_qms_evolution(so_qms* q, double t, int n) {  /* system.cu:14-26 */
    (void)n;
    static const uint32_t xx[2] = {SO_PAULI_X, SO_PAULI_X};
    for (uint32_t iii = 0; iii < 3; ++iii) {
        uint32_t qq[2] = {q->bm_syst[(iii + 1) % 3], q->bm_syst[(iii + 2) % 3]};
        so_apply_pauli_TP_rotation(q->st, qq, xx, 2, -t);
    }
}

static void so_qms_measure_rotation(so_qms* q) {  /* system.cu:76-83 */
    so_apply_h(q->st, 0);
    so_apply_h(q->st, 1);
    so_apply_h(q->st, 2);
    so_apply_cx(q->st, 0, 1, 1u);
    so_apply_cx(q->st, 0, 2, 1u);
    so_apply_u1(q->st, 2, 1u, M_PI * 0.5);
}

double so_qms_measure_X(so_qms* q) {             /* system.cu:99-119 */
    static const double op_vals[8] = {2.0, 0.0, 2.0, 0.0, -2.0, 0.0, -2.0, 0.0};
    so_qms_measure_rotation(q);
    double rd[3];
    for (int k = 0; k < 3; ++k) rd[k] = so_qms_doub(q);
    uint32_t cl[3];
    for (int k = 0; k < 3; ++k) cl[k] = so_measure_qbit(q->st, (uint32_t)k, rd[k]);
    so_qms_measure_rotation(q);   /* apply_measure_antirotation == rotation (system.cu:86-88) */
    uint32_t meas = 0u;
    for (uint32_t i = 0; i < 3; ++i) meas |= cl[i] << i;
    return op_vals[meas];
}

static void so_qms_apply_C(so_qms* q, uint32_t Ci) { so_apply_h(q->st, q->bm_syst[Ci]); } /* system.cu:125-129 */

static uint32_t so_qms_draw_C(so_qms* q) {       /* qms.cuh:207-216, system.cu:123 */
    const double ws[3] = {1. / 3, 2. / 3, 1.0};
    double e = so_qms_doub(q);
    for (uint32_t ci = 0; ci < 3; ++ci)
        if (e < ws[ci]) return ci;
    return 3;
}

static void so_qms_cevolution(so_qms* q, double t, int n, uint32_t qc) { /* qms.cuh:106-115 */
    so_activate_gc_mask(q->st, &qc, 1);
    so_qms_evolution(q, t, n);
    so_deactivate_gc_mask(q->st, &qc, 1);
}

static void so_qms_phase_estimation(so_qms* q, const uint32_t* reg) {   /* qms.cuh:153-176 */
    uint32_t r = q->ene_qbits;
    double t = q->t_phase_estimation;
    int n = q->n_phase_estimation;
    for (uint32_t k = 0; k < r; ++k) so_apply_h(q->st, reg[k]);
    for (int trg = (int)r - 1; trg > -1; --trg) {
        double powr = (double)(1u << (r - 1 - (uint32_t)trg));
        so_qms_cevolution(q, -powr * t, (int)(powr * n), reg[trg]);
        so_apply_u1(q->st, reg[trg], 1u, -powr * t * q->t_PE_shift);
    }
    so_apply_qft_inverse(q->st, reg, r);
}

void so_qms_apply_Phi(so_qms* q) { so_qms_phase_estimation(q, q->bm_enes); }   /* qms.cuh:203 */

void so_qms_apply_Phi_inverse(so_qms* q) {       /* qms.cuh:179-198 */
    uint32_t r = q->ene_qbits;
    double t = q->t_phase_estimation;
    int n = q->n_phase_estimation;
    so_apply_qft(q->st, q->bm_enes, r);
    for (uint32_t trg = 0; trg < r; ++trg) {
        double powr = (double)(1u << (r - 1 - trg));
        so_qms_cevolution(q, powr * t, (int)(powr * n), q->bm_enes[trg]);
        so_apply_u1(q->st, q->bm_enes[trg], 1u, powr * t * q->t_PE_shift);
    }
    for (uint32_t k = 0; k < r; ++k) so_apply_h(q->st, q->bm_enes[k]);
}

void so_qms_apply_W(so_qms* q) {                 /* qms.cuh:286-307 (dense CPU) */
    so_state* s = q->st;
    double c = q->c_factor;
    uint32_t qa = q->bm_acc, wm = q->W_mask_Enew, b0 = q->bm_enes[0];
    double fs1, fs2;
    for (uint64_t i = s->N / 2; i < s->N; ++i) {
        uint64_t j = i & ~(uint64_t)(1u << qa);
        uint32_t Enew = (((uint32_t)i) & wm) >> b0;
        /* two registers: E_old read from the basis state */
        uint32_t Eold = q->two_registers ? (((uint32_t)i) & q->W_mask_Eold) >> q->bm_enes_old[0] : q->curr_E_old;
        double dE = (Enew * c - Eold * c);
        if (dE > 0) {
            fs1 = exp(-dE / 2.0);
            fs2 = sqrt(1.0 - fs1 * fs1);
        } else {
            fs1 = 1.0;
            fs2 = 0.0;
        }
        double tmp = s->re[j];
        s->re[j] = fs2 * s->re[j] + fs1 * s->re[i];
        s->re[i] = fs1 * tmp - fs2 * s->re[i];
        tmp = s->im[j];
        s->im[j] = fs2 * s->im[j] + fs1 * s->im[i];
        s->im[i] = fs1 * tmp - fs2 * s->im[i];
    }
}

static void so_qms_apply_U(so_qms* q) {          /* qms.cuh:335-348 */
    so_qms_apply_C(q, q->gCi);
    so_qms_apply_Phi(q);
    so_qms_apply_W(q);
}
static void so_qms_apply_U_inverse(so_qms* q) {  /* qms.cuh:350-360 */
    so_qms_apply_W(q);
    so_qms_apply_Phi_inverse(q);
    so_qms_apply_C(q, q->gCi);
}

static uint32_t so_creg_to_uint(const uint32_t* c, uint32_t n) {   /* qms.cuh:117-126 */
    uint32_t ret = c[0];
    for (uint32_t j = 1; j < n; ++j) ret += c[j] << j;
    return ret;
}

/* measure_qbits(bm_enes, cs, extract_rands(r)): all draws first (qms.cuh:363-369) */
static void so_qms_measure_reg(so_qms* q, const uint32_t* reg, uint32_t* cs) {
    double rd[32];
    for (uint32_t k = 0; k < q->ene_qbits; ++k) rd[k] = so_qms_doub(q);
    for (uint32_t k = 0; k < q->ene_qbits; ++k) cs[k] = so_measure_qbit(q->st, reg[k], rd[k]);
}
static void so_qms_measure_enes(so_qms* q, uint32_t* cs) { so_qms_measure_reg(q, q->bm_enes, cs); }

static void so_qms_reset_non_syst(so_qms* q) {   /* qms.cuh:128-150 */
    double rd[32];
    if (q->two_registers) {   /* the E_old register first (qms.cuh:134-136) */
        for (uint32_t k = 0; k < q->ene_qbits; ++k) rd[k] = so_qms_doub(q);
        for (uint32_t k = 0; k < q->ene_qbits; ++k) so_apply_reset(q->st, q->bm_enes_old[k], rd[k]);
    } else if (q->compare_draws) {   /* the same draws, nothing to reset */
        for (uint32_t k = 0; k < q->ene_qbits; ++k) (void)so_qms_doub(q);
    }
    for (uint32_t k = 0; k < q->ene_qbits; ++k) rd[k] = so_qms_doub(q);
    for (uint32_t k = 0; k < q->ene_qbits; ++k) so_apply_reset(q->st, q->bm_enes[k], rd[k]);
    so_apply_reset(q->st, q->bm_acc, so_qms_doub(q));
}

int so_qms_metro_step(so_qms* q, int take_measure) {   /* qms.cuh:373-545 */
    int ret = 0;
    uint32_t r = q->ene_qbits;
    uint32_t c_E_news[32] = {0}, c_E_olds[32] = {0};
    so_qms_reset_non_syst(q);
    if (q->two_registers) {
        /* E_old into its own register, which keeps the measured value */
        so_qms_phase_estimation(q, q->bm_enes_old);
        so_qms_measure_reg(q, q->bm_enes_old, c_E_olds);
    } else {
        so_qms_apply_Phi(q);
        so_qms_measure_enes(q, c_E_olds);
    }
    q->curr_E_old = so_creg_to_uint(c_E_olds, r);
    q->curr_E_old_d = q->t_PE_shift + q->curr_E_old / (double)(q->t_PE_factor * q->ene_levels);
    if (!q->two_registers)
        for (uint32_t ei = 0; ei < r; ++ei)
            if (c_E_olds[ei]) so_apply_x(q->st, q->bm_enes[ei]);
    q->gCi = so_qms_draw_C(q);
    so_qms_apply_U(q);
    q->c_acc = so_measure_qbit(q->st, q->bm_acc, so_qms_doub(q));
    if (q->c_acc == 1u) {
        so_qms_measure_enes(q, c_E_news);
        so_qms_apply_Phi_inverse(q);
        if (take_measure) {
            double e = q->t_PE_shift + so_creg_to_uint(c_E_news, r) / (double)(q->t_PE_factor * q->ene_levels);
            so_qms_push_E(q, e);
            for (uint32_t ei = 0; ei < r; ++ei) {
                /* qms.cuh:431-434: the comparison draw before each flip */
                if (q->two_registers) so_apply_reset(q->st, q->bm_enes_old[ei], so_qms_doub(q));
                else if (q->compare_draws) so_apply_reset(q->st, q->bm_enes[ei], so_qms_doub(q));
                if (c_E_news[ei]) so_apply_x(q->st, q->bm_enes[ei]);
            }
            so_qms_push_X(q, so_qms_measure_X(q));
            for (uint32_t ei = 0; ei < r; ++ei) so_apply_reset(q->st, q->bm_enes[ei], so_qms_doub(q));
            so_qms_apply_Phi(q);
            so_qms_measure_enes(q, c_E_news);
            so_qms_apply_Phi_inverse(q);
            ret = 2;
        } else {
            ret = 1;
        }
        return ret;
    }
    so_qms_apply_U_inverse(q);
    uint32_t iters = 0;
    while (iters < q->max_reverse_attempts) {
        so_qms_apply_Phi(q);
        /* qms.cuh:478-479: the E_old readout's r draws (two registers: the
         * register's value is the one measured at the start of the step) */
        if (q->two_registers || q->compare_draws)
            for (uint32_t k = 0; k < r; ++k) (void)so_qms_doub(q);
        so_qms_measure_enes(q, c_E_news);
        uint32_t Enew_meas = so_creg_to_uint(c_E_news, r);
        so_qms_apply_Phi_inverse(q);
        if (q->curr_E_old == Enew_meas) {
            if (take_measure) {
                so_qms_push_E(q, q->curr_E_old_d);
                for (uint32_t ei = 0; ei < r; ++ei) so_apply_reset(q->st, q->bm_enes[ei], so_qms_doub(q));
                so_qms_push_X(q, so_qms_measure_X(q));
                for (uint32_t ei = 0; ei < r; ++ei) so_apply_reset(q->st, q->bm_enes[ei], so_qms_doub(q));
                so_qms_apply_Phi(q);
                so_qms_measure_enes(q, c_E_news);
                so_qms_apply_Phi_inverse(q);
                ret = 4;
            } else {
                ret = 3;
            }
            break;
        }
        so_qms_apply_U(q);
        (void)so_measure_qbit(q->st, q->bm_acc, so_qms_doub(q));
        so_qms_apply_U_inverse(q);
        iters++;
    }
    if (q->record_reverse) {
        if (q->n_rev >= q->cap_rev) {
            size_t nc = q->cap_rev ? 2 * q->cap_rev : 64;
            q->reverse_counters = (uint32_t*)realloc(q->reverse_counters, nc * sizeof(uint32_t));
            q->cap_rev = nc;
        }
        q->reverse_counters[q->n_rev++] = iters;
    }
    if (iters == q->max_reverse_attempts) return -1;
    return ret;
}

uint32_t so_qms_run(so_qms* q, uint32_t metro_steps, uint32_t reset_each, int thermalization,
                    int* rets) {                 /* qms.cu:132-149 */
    so_qms_init_state(q);
    uint32_t s0 = 0u, acc = 0u;
    for (uint32_t s = 0u; s < metro_steps; ++s) {
        int take = (s > s0 + (uint32_t)thermalization && (s - s0) % reset_each == 0u);
        int ret = so_qms_metro_step(q, take);
        if (rets) rets[s] = ret;
        if (ret < 0) {
            so_qms_init_state(q);
            s0 = s + 1;
        }
        if (ret == 1 || ret == 2) acc++;
    }
    return acc;
}
