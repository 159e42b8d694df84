// qsa_run.cpp -- the QSA driver loop of src/qsa.cu:115-262 as a library call
// (suqa_qsa_run, include/suqa_qsa.h).  The CLI in qsa_main.cpp runs the same
// loop through qsadrv::run with file output.
#include "qsa_run.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "qsa.hpp"
#include "system.hpp"

namespace qsadrv {

void validate(const suqa_qsa_params& p) {   // parser_qsa.hpp:154-186
    if (p.beta <= 0.0) throw std::runtime_error("ERROR: argument <beta> invalid");
    if (p.ene_qbits == 0) throw std::runtime_error("ERROR: argument <num ene qbits> non positive");
    if (p.szegedy_qbits == 0) throw std::runtime_error("ERROR: argument <num ene qbits> non positive");
    if (p.pe_steps <= 0) throw std::runtime_error("ERROR: argument <steps of PE evolution> non positive");
    if (p.annealing_sequences == 0) throw std::runtime_error("ERROR: argument <annealing_sequences> non positive");
    if (p.sampling == 0) throw std::runtime_error("ERROR: argument <sampling> non positive");
    if (syst_qbits + p.ene_qbits + p.szegedy_qbits + 1 > 30)
        throw std::runtime_error("ERROR: too many qubits for this driver (at most 30)");
}

void configure(const suqa_qsa_params& p) {   // qsa.cu:115-146
    qsa::beta = p.beta;
    qsa::ene_qbits = p.ene_qbits;
    qsa::szegedy_qbits = p.szegedy_qbits;
    qsa::n_phase_estimation = p.pe_steps;
    qsa::iseed = p.seed;
    if (qsa::iseed > 0) qsa::rangen.set_seed(qsa::iseed);
    qsa::iseed = qsa::rangen.get_seed();
    qsa::syst_qbits = (uint)syst_qbits;
    qsa::nqubits = qsa::syst_qbits + qsa::ene_qbits + qsa::szegedy_qbits + 1;
    qsa::Dim = (1U << qsa::nqubits);
    qsa::ene_levels = (1U << qsa::ene_qbits);
    const double szegedy_levels = (1U << qsa::szegedy_qbits);
    qsa::t_PE_shift = p.ene_min;
    qsa::t_PE_factor = (qsa::ene_levels - 1) / (double)(qsa::ene_levels * (p.ene_max - p.ene_min));
    qsa::t_phase_estimation = qsa::t_PE_factor * 8. * atan(1.0);
    qsa::t_PE_shift_szegedy = p.ene_min_szegedy;
    qsa::t_PE_factor_szegedy = (szegedy_levels - 1) / (double)((szegedy_levels) * (p.ene_max_szegedy - p.ene_min_szegedy));
    qsa::t_phase_estimation_szegedy = qsa::t_PE_factor_szegedy * 8. * atan(1.0);
    qsa::W_mask_E = 0U;
    qsa::E_measures.clear();
    qsa::X_measures.clear();
}

// qsa.cu:175-255: for each sample, restart the annealing (move draw, resets,
// |+++> copied on the tilde spins) until every annealing step's Szegedy
// readout is 0, then phase-estimate the system's energy into the Szegedy
// register, measure it and the observable.
void run(const suqa_qsa_params& p, const Hooks& h) {
    const double szegedy_levels = (1U << qsa::szegedy_qbits);
    int rejection1 = 0;
    uint nann = 0;
    const uint bound = p.max_annealing_steps;
    const double beta_i = qsa::beta / ((double)p.annealing_sequences);
    for (uint iiii = 0; iiii < p.sampling && (bound == 0 || nann < bound);) {
        std::vector<uint> c_Ene_test(qsa::szegedy_qbits, 0);
        std::vector<uint> c_Ene(qsa::szegedy_qbits, 0);
        int control = 0;
        uint attempt = 0;
        while (control == 0 && (bound == 0 || nann < bound)) {
            int control2 = 0;
            uint move = qsa::draw_C();
            qsa::reset_non_syst_qbits();
            qsa_init_state();
            for (uint s = 0U; s < p.annealing_sequences; s++) {
                move = qsa::draw_C();
                qsa::apply_PE_szegedy(qsa::bm_states, qsa::bm_szegedy, (double)(s + 1) * beta_i, move);
                suqa::measure_qbits(qsa::bm_szegedy, c_Ene_test, qsa::extract_rands(qsa::szegedy_qbits));
                ++nann;
                const uint readout = qsa::creg_to_uint(c_Ene_test);
                if (h.step) h.step(iiii, attempt, s, move, readout);
                if (readout != 0U) {
                    control2 = 1;
                    rejection1++;
                    if (h.rejected) h.rejected(s, readout);
                    break;
                }
                if (bound && nann >= bound) break;
            }
            if (control2 == 0 && (bound == 0 || nann <= bound)) control = 1;
            ++attempt;
        }
        if (control == 0) break;   // stopped by the bound inside a sample
        qsa::apply_phase_estimation_measure(qsa::bm_states, qsa::bm_szegedy, qsa::t_phase_estimation_szegedy,
                                            qsa::n_phase_estimation);
        suqa::measure_qbits(qsa::bm_szegedy, c_Ene, qsa::extract_rands(qsa::szegedy_qbits));
        const double Ene_measure =
            qsa::t_PE_shift_szegedy + qsa::creg_to_uint(c_Ene) / (double)(qsa::t_PE_factor_szegedy * szegedy_levels);
        qsa::E_measures.push_back(Ene_measure);
        qsa::X_measures.push_back(measure_X(qsa::rangen));
        if (h.sample) h.sample(iiii, Ene_measure, qsa::X_measures.back(), rejection1);
        ++iiii;
    }
}

}  // namespace qsadrv

extern "C" int suqa_qsa_run(const suqa_qsa_params* p, suqa_qsa_result* out) {
    try {
        if (!p || !out) throw std::runtime_error("null argument");
        qsadrv::validate(*p);
        if (p->seed == 0) throw std::runtime_error("ERROR: suqa_qsa_run needs a seed > 0 (reproducible runs)");
        qsadrv::configure(*p);
        suqa::setup(qsa::nqubits);
        qsa::setup();
        out->n_steps = 0;
        out->n_samples = 0;
        qsadrv::Hooks h;
        h.step = [&](uint sample, uint attempt, uint s, uint move, uint readout) {
            const uint32_t k = out->n_steps++;
            if (k < out->cap_steps) {
                if (out->steps) {
                    uint32_t* r = out->steps + 5 * (size_t)k;
                    r[0] = sample;
                    r[1] = attempt;
                    r[2] = s;
                    r[3] = move;
                    r[4] = readout;
                }
                if (out->rng_after) out->rng_after[k] = qsa::rangen.get_state().state;
            }
        };
        h.sample = [&](uint, double E, double X, int rej) {
            const uint32_t k = out->n_samples++;
            if (k < out->cap_samples) {
                if (out->E) out->E[k] = E;
                if (out->X) out->X[k] = X;
                if (out->rejections) out->rejections[k] = rej;
                if (out->sample_rng) out->sample_rng[k] = qsa::rangen.get_state().state;
            }
        };
        auto t0 = std::chrono::steady_clock::now();
        qsadrv::run(*p, h);
        suqa_hip_synchronize(suqa::ctx);
        out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        out->final_rng = qsa::rangen.get_state().state;
        if (out->final_re || out->final_im) {
            if (suqa_hip_download(suqa::ctx, out->final_re, out->final_im) != SUQA_OK)
                throw std::runtime_error(suqa_hip_last_error());
        }
        qsa::clear();
        suqa::clear();
        return SUQA_OK;
    } catch (const std::exception& e) {
        fprintf(stderr, "suqa_qsa_run: %s\n", e.what());
        try {
            suqa::clear();
        } catch (...) {
        }
        return SUQA_ERR_ARG;
    }
}
