// qms_main.cpp -- `suqa_qms`, the QMS command line of the reference
// (src/qms.cu:51-178, include/parser_qms.hpp) on the MI355X engine:
//
//   suqa_qms <beta> <metro steps> <reset each> <num ene qbits> <output file path>
//            [--max-reverse N (100)] [--seed S (time)] [--ene-min E (0.0)]
//            [--ene-max E (1.0)] [--PE-steps N (10)] [--thermalization N (100)]
//            [--record-reverse] [--ene-registers 1|2 (1)]
//
// --ene-registers 2 (an extension) runs the two-register layout: E_old in its
// own register, 3 + 2r + 1 qubits.
// Output file: "# E A" header if the file is new, then "%.16lg %.16lg" lines
// appended every ceil(steps/20) steps (measurements after the last such step
// are not written, as in the reference); <out>_revcounts with
// --record-reverse.
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <string>

#include "qms.hpp"
#include "qms_run.hpp"
#include "system.hpp"

static void save_measures(const std::string& outfilename) {   // qms.cu:41-49
    qms::sync_measures();   // values of deferred readouts
    FILE* fil = fopen(outfilename.c_str(), "a");
    if (!fil) throw std::runtime_error("ERROR: cannot open " + outfilename);
    for (uint ei = 0; ei < qms::E_measures.size(); ++ei)
        fprintf(fil, "%.16lg %.16lg\n", qms::E_measures[ei], qms::X_measures[ei]);
    fclose(fil);
    qms::E_measures.clear();
    qms::X_measures.clear();
}

int main(int argc, char** argv) {
    if (argc < 6) {
        printf("usage: %s <beta> <metro steps> <reset each> <num ene qbits> <output file path> [--max-reverse <max "
               "reverse attempts> (100)] [--seed <seed> (random)] [--ene-min <min energy> (0.0)] [--ene-max <max "
               "energy> (1.0)] [--PE-steps <steps of PE evolution> (10)] [--thermalization <steps> (100)] "
               "[--record-reverse] [--ene-registers <1|2> (1)]\n",
               argv[0]);
        exit(1);
    }
    try {
        suqa_qms_params p{};
        p.beta = std::stod(argv[1]);
        int steps = atoi(argv[2]), reset_each = atoi(argv[3]), ene_qbits = atoi(argv[4]);
        std::string outfilename(argv[5]);
        int max_rev = 100, pe = 10, therm = 100;
        p.seed = 0;
        p.ene_min = 0.0;
        p.ene_max = 1.0;
        std::map<std::string, int> argmap;   // parser_qms.hpp:57-60: flags after the 5 fixed arguments
        for (int i = 6; i < argc; ++i) argmap[argv[i]] = i;
        auto val = [&](const char* flag) -> const char* {
            auto it = argmap.find(flag);
            if (it == argmap.end()) return nullptr;
            if (it->second + 1 >= argc) throw std::runtime_error(std::string("ERROR: set value after '") + flag + "' flag");
            return argv[it->second + 1];
        };
        p.record_reverse = argmap.count("--record-reverse") ? 1 : 0;
        if (const char* v = val("--max-reverse")) max_rev = atoi(v);
        if (const char* v = val("--seed")) p.seed = strtoull(v, nullptr, 10);
        if (const char* v = val("--ene-min")) p.ene_min = std::stod(v);
        if (const char* v = val("--ene-max")) p.ene_max = std::stod(v);
        if (const char* v = val("--PE-steps")) pe = (int)std::stod(v);
        if (const char* v = val("--thermalization")) therm = (int)std::stod(v);
        if (const char* v = val("--ene-registers")) {
            const int k = atoi(v);
            if (k != 1 && k != 2) throw std::runtime_error("ERROR: argument <ene registers> must be 1 or 2");
            p.two_registers = k == 2 ? 1 : 0;
        }
        if (steps <= 0) throw std::runtime_error("ERROR: argument <metro steps> invalid");
        if (reset_each <= 0) throw std::runtime_error("ERROR: argument <reset each> non positive");
        if (ene_qbits <= 0) throw std::runtime_error("ERROR: argument <num ene qbits> non positive");
        if (max_rev <= 0) throw std::runtime_error("ERROR: argument <max reverse attempts> non positive");
        if (outfilename.empty()) throw std::runtime_error("ERROR: argument <output file path> empty");
        p.metro_steps = (uint32_t)steps;
        p.reset_each = (uint32_t)reset_each;
        p.ene_qbits = (uint32_t)ene_qbits;
        p.max_reverse = (uint32_t)max_rev;
        p.pe_steps = pe;
        p.thermalization = therm;
        qmsdrv::validate(p);
        qmsdrv::configure(p);
        std::cout << qms::nqubits << std::endl;
        suqa::print_banner();
        std::cout << "arguments:\nbeta: " << p.beta << "\nmetro steps: " << steps << "\nreset each: " << reset_each
                  << "\nnum E qbits " << ene_qbits << "\nmax reverse attempts: " << max_rev
                  << "\nseed: " << qms::iseed << "\nout datafile: " << outfilename << "\nmin energy: " << p.ene_min
                  << "\nmax energy: " << p.ene_max << "\nsteps of PE evolution: " << pe
                  << "\nthermalization: " << therm << "\nrecord reverse: " << p.record_reverse << "\n"
                  << std::endl;

        auto t_start = std::chrono::high_resolution_clock::now();
        suqa::setup(qms::nqubits);
        qms::setup(p.beta);
        qms::defer_values = true;   // save_measures reads E/X after sync_measures()
#ifdef GATECOUNT
        // the gate-count build of the driver (qms.cu:97-105)
        suqa::gatecounters.add_counter(&qms::gctr_global);
        suqa::gatecounters.add_counter(&qms::gctr_metrostep);
        suqa::gatecounters.add_counter(&qms::gctr_sample);
        suqa::gatecounters.add_counter(&qms::gctr_measure);
        suqa::gatecounters.add_counter(&qms::gctr_reverse);
        qms::gctr_global.new_record();
        qms::gctr_sample.new_record();
#endif
        init_state();
        uint perc_mstep = (qms::metro_steps + 19) / 20;
        uint count_accepted = 0U;
        if (access(outfilename.c_str(), F_OK) == -1) {
            FILE* fil = fopen(outfilename.c_str(), "w");
            if (!fil) throw std::runtime_error("ERROR: cannot create " + outfilename);
            fprintf(fil, "# E A\n");
            fclose(fil);
        }
        bool take_measure;
        uint s0 = 0U;
        auto t_loop = std::chrono::high_resolution_clock::now();
        for (uint s = 0U; s < qms::metro_steps; ++s) {
            take_measure = (s > s0 + (uint)therm and (s - s0) % qms::reset_each == 0U);
            int ret = qms::metro_step(take_measure);
            if (ret < 0) {
                init_state();
                s0 = s + 1;
            }
            if (ret == 1 or ret == 2) count_accepted++;
            if (s % perc_mstep == 0) {
                std::cout << "iteration: " << s << "/" << qms::metro_steps << std::endl;
                save_measures(outfilename);
            }
        }
        auto t_loop_end = std::chrono::high_resolution_clock::now();
        std::cout << std::endl;
        printf("\n\tacceptance: %3.2lg%%\n", (count_accepted / static_cast<double>(qms::metro_steps)) * 100.0);
        qms::clear();
        suqa::clear();
        std::cout << "\nall fine :)\n" << std::endl;
        if (qms::record_reverse) {
            FILE* fil_rev = fopen((outfilename + "_revcounts").c_str(), "w");
            if (fil_rev) {
                for (uint i = 0; i < qms::reverse_counters.size(); ++i)
                    fprintf(fil_rev, "%d %d\n", i, static_cast<int>(qms::reverse_counters[i]));
                fclose(fil_rev);
            }
        }
        std::cout << "\n\tSuqa!\n" << std::endl;
        auto t_end = std::chrono::high_resolution_clock::now();
        double secs = std::chrono::duration<double>(t_end - t_start).count();
        double loop = std::chrono::duration<double>(t_loop_end - t_loop).count();
        std::cout << "All [DONE] in " << secs << " seconds" << std::endl;
        printf("Metropolis steps/s: %.1f\n", qms::metro_steps / loop);
    } catch (const std::exception& e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
