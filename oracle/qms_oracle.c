/*
 * qms_oracle.c -- the QMS driver (qms.cu:51-178) on the oracle, for CPU
 * timing beside the GPU (TEST INFRASTRUCTURE ONLY).  Same positional
 * arguments and flags as the reference driver; writes the same "# E A" file.
 * Prints one line "steps=<n> seconds=<t> accepted=<a>" at the end.
 */
#define _POSIX_C_SOURCE 199309L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>
#include "suqa_oracle.h"

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char** argv) {
    if (argc < 6) {
        fprintf(stderr, "usage: %s <beta> <metro steps> <reset each> <num ene qbits> <out> "
                        "[--max-reverse N] [--seed S] [--ene-min E] [--ene-max E] [--PE-steps N] "
                        "[--thermalization N] [--record-reverse]\n", argv[0]);
        return 1;
    }
    double beta = atof(argv[1]);
    int steps = atoi(argv[2]), reset_each = atoi(argv[3]), r = atoi(argv[4]);
    const char* out = argv[5];
    int max_rev = 100, pe = 10, therm = 100, rec = 0;
    unsigned long long seed = 0;
    double emin = 0.0, emax = 1.0;
    for (int i = 6; i < argc; ++i) {
        if (!strcmp(argv[i], "--record-reverse")) rec = 1;
        else if (i + 1 < argc && !strcmp(argv[i], "--max-reverse")) max_rev = atoi(argv[++i]);
        else if (i + 1 < argc && !strcmp(argv[i], "--seed")) seed = strtoull(argv[++i], NULL, 10);
        else if (i + 1 < argc && !strcmp(argv[i], "--ene-min")) emin = atof(argv[++i]);
        else if (i + 1 < argc && !strcmp(argv[i], "--ene-max")) emax = atof(argv[++i]);
        else if (i + 1 < argc && !strcmp(argv[i], "--PE-steps")) pe = atoi(argv[++i]);
        else if (i + 1 < argc && !strcmp(argv[i], "--thermalization")) therm = atoi(argv[++i]);
    }
    if (beta <= 0 || steps <= 0 || reset_each <= 0 || r <= 0 || max_rev <= 0 || pe <= 0 || seed == 0) {
        fprintf(stderr, "invalid arguments (seed must be > 0 here)\n");
        return 1;
    }
    so_qms* q = so_qms_create(beta, (uint32_t)r, (uint32_t)max_rev, seed, emin, emax, pe, rec);
    if (!q) return 1;
    /* qms.cu:120-149: header once, then each batch appended when s % perc == 0;
     * measurements taken after the last such step are never written. */
    FILE* f;
    if (access(out, F_OK) != 0) {
        f = fopen(out, "w");
        if (f) { fprintf(f, "# E A\n"); fclose(f); }
    }
    uint32_t perc = ((uint32_t)steps + 19u) / 20u;
    size_t saved = 0;
    double t0 = now_s();
    so_qms_init_state(q);
    uint32_t s0 = 0u, acc = 0u;
    for (uint32_t s = 0u; s < (uint32_t)steps; ++s) {
        int take = (s > s0 + (uint32_t)therm && (s - s0) % (uint32_t)reset_each == 0u);
        int ret = so_qms_metro_step(q, take);
        if (ret < 0) { so_qms_init_state(q); s0 = s + 1; }
        if (ret == 1 || ret == 2) acc++;
        if (s % perc == 0u) {
            f = fopen(out, "a");
            if (f) {
                for (size_t k = saved; k < q->n_meas; ++k)
                    fprintf(f, "%.16lg %.16lg\n", q->E_measures[k], q->X_measures[k]);
                fclose(f);
            }
            saved = q->n_meas;
        }
    }
    double t1 = now_s();
    printf("steps=%d seconds=%.9f accepted=%u\n", steps, t1 - t0, acc);
    so_qms_destroy(q);
    return 0;
}
