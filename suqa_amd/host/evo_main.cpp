// evo_main.cpp -- suqa_evo: the real-time evolution table of the reference's
// `evo` driver (evo.cu:22-103) on this engine.  For t = ii * dt, ii = 0 ..
// total_steps: the system's initial state, exp(-iHt) (system evolution), and
// the probabilities of the 8 spin configurations, appended to <outfile> as
// "%.12lg" columns t p000 p100 p010 p110 p001 p101 p011 p111.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "suqa.hpp"
#include "system.hpp"

int main(int argc, char** argv) {
    if (argc < 4) {
        std::printf("usage: %s <total_steps> <trotter_stepsize> <outfile>\n", argv[0]);
        return 1;
    }
    const int total_steps = std::atoi(argv[1]);
    const double dt = std::stod(argv[2]);
    const std::string outname = argv[3];
    std::printf("arguments:\n total_steps = %d\n trotter_stepsize = %.16lg\n outfile = %s\n\n", total_steps, dt,
                outname.c_str());
    try {
        suqa::setup(syst_qbits);
        init_state();
        // the reference's order of configurations: bit k of the index is spin k
        const std::vector<std::vector<uint>> masks = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                                      {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
        for (uint ii = 0; ii <= (uint)total_steps; ++ii) {
            const double t = ii * dt;
            init_state();
            evolution(t, (int)ii);
            double p[8];
            for (int k = 0; k < 8; ++k) suqa::prob_filter(bm_spin, masks[k], p[k]);
            FILE* f = std::fopen(outname.c_str(), "a");
            if (!f) throw std::runtime_error("cannot open " + outname);
            std::fprintf(f, "%.12lg %.12lg %.12lg %.12lg %.12lg %.12lg %.12lg %.12lg %.12lg\n", t, p[0], p[1], p[2],
                         p[3], p[4], p[5], p[6], p[7]);
            std::fclose(f);
        }
        suqa::clear();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "suqa_evo: %s\n", e.what());
        return 2;
    }
    return 0;
}
