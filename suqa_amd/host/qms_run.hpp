// qms_run.hpp -- shared setup of the QMS drivers (qms_run.cpp, qms_main.cpp).
#pragma once
#include "suqa_qms.h"

namespace qmsdrv {
void configure(const suqa_qms_params& p);   // qms.cu:251-276
void validate(const suqa_qms_params& p);    // parser_qms.hpp:130-157
}  // namespace qmsdrv
