// qsa_run.hpp -- the QSA driver loop (src/qsa.cu:115-262) shared by
// suqa_qsa_run and the suqa_qsa command line.
#pragma once
#include <functional>

#include "suqa_qsa.h"

namespace qsadrv {

struct Hooks {   // called as the loop goes; any may be empty
    std::function<void(unsigned sample, unsigned restart, unsigned s, unsigned move, unsigned readout)> step;
    std::function<void(unsigned s, unsigned readout)> rejected;
    std::function<void(unsigned sample, double E, double X, int rejections)> sample;
};

void validate(const suqa_qsa_params& p);   // parser_qsa.hpp:154-186 (throws std::runtime_error)
void configure(const suqa_qsa_params& p);  // qsa.cu:115-146
void run(const suqa_qsa_params& p, const Hooks& h);

}  // namespace qsadrv
