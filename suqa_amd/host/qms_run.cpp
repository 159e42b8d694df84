// qms_run.cpp -- the QMS driver loop of src/qms.cu:51-178 as a library call
// (suqa_qms_run, include/suqa_qms.h).  The CLI in qms_main.cpp uses the same
// setup and per-step code.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "qms.hpp"
#include "qms_run.hpp"
#include "suqa_qms.h"
#include "system.hpp"

namespace qmsdrv {

void configure(const suqa_qms_params& p) {   // qms.cu:251-276
    qms::metro_steps = p.metro_steps;
    qms::reset_each = p.reset_each;
    qms::ene_qbits = p.ene_qbits;
    qms::max_reverse_attempts = p.max_reverse;
    qms::n_phase_estimation = p.pe_steps;
    qms::record_reverse = p.record_reverse != 0;
    qms::two_registers = p.two_registers != 0;
    qms::iseed = p.seed;
    if (qms::iseed > 0) qms::rangen.set_seed(qms::iseed);
    qms::iseed = qms::rangen.get_seed();
    qms::syst_qbits = (uint)syst_qbits;
    qms::nqubits = qms::syst_qbits + (qms::two_registers ? 2 : 1) * qms::ene_qbits + 1;
    qms::Dim = (1U << qms::nqubits);
    qms::ene_levels = (1U << qms::ene_qbits);
    qms::syst_levels = (1U << qms::syst_qbits);
    qms::t_PE_shift = p.ene_min;
    qms::t_PE_factor = (qms::ene_levels - 1) / (double)(qms::ene_levels * (p.ene_max - p.ene_min));
    qms::t_phase_estimation = qms::t_PE_factor * 8. * atan(1.0);
    qms::E_measures.clear();
    qms::X_measures.clear();
    qms::reverse_counters.clear();
}

void validate(const suqa_qms_params& p) {   // parser_qms.hpp:130-157
    if (p.beta <= 0.0) throw std::runtime_error("ERROR: argument <beta> invalid");
    if (p.metro_steps == 0) throw std::runtime_error("ERROR: argument <metro steps> invalid");
    if (p.reset_each == 0) throw std::runtime_error("ERROR: argument <reset each> non positive");
    if (p.ene_qbits == 0) throw std::runtime_error("ERROR: argument <num ene qbits> non positive");
    if (p.ene_qbits > 28) throw std::runtime_error("ERROR: argument <num ene qbits> too large");
    if (p.two_registers && p.ene_qbits > 8)
        throw std::runtime_error("ERROR: argument <num ene qbits> too large for two energy registers (at most 8)");
    if (p.max_reverse == 0) throw std::runtime_error("ERROR: argument <max reverse attempts> non positive");
    if (p.pe_steps <= 0) throw std::runtime_error("ERROR: argument <steps of PE evolution> non positive");
}

}  // namespace qmsdrv

extern "C" int suqa_qms_run(const suqa_qms_params* p, suqa_qms_result* out) {
    try {
        if (!p || !out) throw std::runtime_error("null argument");
        qmsdrv::validate(*p);
        if (p->seed == 0) throw std::runtime_error("ERROR: suqa_qms_run needs a seed > 0 (reproducible runs)");
        qmsdrv::configure(*p);
        suqa::setup(qms::nqubits);
        qms::setup(p->beta);
        {   // E/X are read after sync_measures() below; SUQA_QMS_DEFER_VALUES=0 resolves them every step
            const char* e = std::getenv("SUQA_QMS_DEFER_VALUES");
            qms::defer_values = !(e && std::strcmp(e, "0") == 0);
        }
        init_state();
        out->accepted = 0;
        bool take_measure;
        uint s0 = 0U;
        const uint therm = (uint)p->thermalization;
        auto t0 = std::chrono::steady_clock::now();
        for (uint s = 0U; s < qms::metro_steps; ++s) {   // qms.cu:132-149
            take_measure = (s > s0 + therm and (s - s0) % qms::reset_each == 0U);
            int ret = qms::metro_step(take_measure);
            if (ret < 0) {
                init_state();
                s0 = s + 1;
            }
            if (ret == 1 or ret == 2) out->accepted++;
            if (out->rets) out->rets[s] = ret;
            if (out->rng_after) out->rng_after[s] = qms::rangen.get_state().state;
        }
        qms::sync_measures();   // the E and X values of the last steps
        suqa_hip_synchronize(suqa::ctx);
        auto t1 = std::chrono::steady_clock::now();
        out->seconds = std::chrono::duration<double>(t1 - t0).count();
        out->n_meas = (uint32_t)qms::E_measures.size();
        for (uint32_t k = 0; k < out->n_meas && k < out->cap_meas; ++k) {
            if (out->E) out->E[k] = qms::E_measures[k];
            if (out->X) out->X[k] = qms::X_measures[k];
        }
        out->n_reverse = (uint32_t)qms::reverse_counters.size();
        for (uint32_t k = 0; out->reverse && k < out->n_reverse && k < out->cap_reverse; ++k)
            out->reverse[k] = qms::reverse_counters[k];
        if (out->final_re || out->final_im) {
            if (suqa_hip_download(suqa::ctx, out->final_re, out->final_im) != SUQA_OK)
                throw std::runtime_error(suqa_hip_last_error());
        }
        qms::clear();
        suqa::clear();
        return SUQA_OK;
    } catch (const std::exception& e) {
        fprintf(stderr, "suqa_qms_run: %s\n", e.what());
        try {
            suqa::clear();
        } catch (...) {
        }
        return SUQA_ERR_ARG;
    }
}
