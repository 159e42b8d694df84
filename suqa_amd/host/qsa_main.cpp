// qsa_main.cpp -- `suqa_qsa`, the QSA command line of the reference
// (src/qsa.cu:82-275, include/parser_qsa.hpp) on the MI355X engine:
//
//   suqa_qsa <beta> <sampling> <num ene qbits> <num szegedy qbits> <output file path>
//            [--seed S (time)] [--ene-min E (-4)] [--ene-max E (8)]
//            [--ene-min-szegedy E (-1)] [--ene-max-szegedy E (3)]
//            [--PE-steps N (1)] [--annealing-sequences N (100)]
//
// Output file: "# E A" header if the file is new, then one
// "%.16lg\t%.16lg\t%d" line (energy, X, restarts so far) per sample, appended
// as each sample ends.
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <string>

#include "qsa.hpp"
#include "qsa_run.hpp"
#include "system.hpp"

static void save_measures(const std::string& outfilename, int rej) {   // qsa.cu:51-60
    FILE* fil = fopen(outfilename.c_str(), "a");
    if (!fil) throw std::runtime_error("ERROR: cannot open " + outfilename);
    for (uint ei = 0; ei < qsa::E_measures.size(); ++ei)
        fprintf(fil, "%.16lg\t%.16lg\t%d\n", qsa::E_measures[ei], qsa::X_measures[ei], rej);
    fclose(fil);
    qsa::E_measures.clear();
    qsa::X_measures.clear();
}

int main(int argc, char** argv) {
    if (argc < 6) {
        printf("usage: %s <beta> <sampling> <num ene qbits> <num szegedy qbits> <output file path>  [--seed <seed> "
               "(random)] [--annealing_sequences (100)]\n",
               argv[0]);
        exit(1);
    }
    try {
        suqa_qsa_params p{};
        p.beta = std::stod(argv[1]);
        const int sampling = atoi(argv[2]), ene_q = atoi(argv[3]), sz_q = atoi(argv[4]);
        const std::string outfilename(argv[5]);
        if (sampling <= 0) throw std::runtime_error("ERROR: argument <sampling> non positive");
        if (ene_q <= 0 || sz_q <= 0) throw std::runtime_error("ERROR: argument <num ene qbits> non positive");
        p.sampling = (uint32_t)sampling;
        p.ene_qbits = (uint32_t)ene_q;
        p.szegedy_qbits = (uint32_t)sz_q;
        p.ene_min = -4.0;
        p.ene_max = 8.0;
        p.ene_min_szegedy = -1.0;
        p.ene_max_szegedy = 3.0;
        p.pe_steps = 1;
        p.annealing_sequences = 100;
        std::map<std::string, int> argmap;   // parser_qsa.hpp: flags after the 5 fixed arguments
        for (int i = 6; i < argc; ++i) argmap[argv[i]] = i;
        auto val = [&](const char* flag) -> const char* {
            auto it = argmap.find(flag);
            if (it == argmap.end()) return nullptr;
            if (it->second + 1 >= argc) throw std::runtime_error(std::string("ERROR: set value after '") + flag + "' flag");
            return argv[it->second + 1];
        };
        if (const char* v = val("--seed")) p.seed = strtoull(v, nullptr, 10);
        if (const char* v = val("--ene-min")) p.ene_min = std::stod(v);
        if (const char* v = val("--ene-max")) p.ene_max = std::stod(v);
        if (const char* v = val("--ene-min-szegedy")) p.ene_min_szegedy = std::stod(v);
        if (const char* v = val("--ene-max-szegedy")) p.ene_max_szegedy = std::stod(v);
        if (const char* v = val("--PE-steps")) p.pe_steps = (int32_t)std::stod(v);
        if (const char* v = val("--annealing-sequences")) p.annealing_sequences = (uint32_t)std::stod(v);
        if (p.seed == 0)   // the reference seeds from the clock (Rand.hpp:258-262)
            p.seed = (uint64_t)std::chrono::high_resolution_clock::now().time_since_epoch().count();
        qsadrv::validate(p);
        suqa::print_banner();
        auto t_start = std::chrono::high_resolution_clock::now();
        qsadrv::configure(p);
        suqa::setup(qsa::nqubits);
        qsa::setup();
        if (access(outfilename.c_str(), F_OK) == -1) {
            FILE* fil = fopen(outfilename.c_str(), "w");
            if (!fil) throw std::runtime_error("ERROR: cannot open " + outfilename);
            fprintf(fil, "# E A\n");
            fclose(fil);
        }
        int itctr = 0, rej_total = 0;
        double mean = 0.0, stdv = 0.0;
        qsadrv::Hooks h;
        h.rejected = [&](unsigned s, unsigned readout) {
            ++rej_total;
            printf("break after %d steps   eigenvalue:%d   \n", (int)s, (int)readout);
        };
        h.sample = [&](unsigned iiii, double E, double, int rej) {
            printf("measured eigenvalue:%f\n", E);
            itctr++;
            mean += E;
            stdv += E * E;
            if (itctr > 1)
                printf("mean=%f, std=%f; expected mean=%f\n", mean / itctr,
                       sqrt((stdv - mean * mean / itctr) / ((itctr) * (itctr - 1))),
                       (6 - 6 * exp(4 * qsa::beta)) / (2 + 6 * exp(4 * qsa::beta)));
            std::cout << "iteration: " << iiii + 1 << "/" << p.sampling << std::endl;
            save_measures(outfilename, rej);
        };
        qsadrv::run(p, h);
        printf("n of rejection step cause by acceptance qubit= %d\n", 0);
        printf("n of rejection step cause by energy qubits= %d\n", rej_total);
        printf("n of total call of W= %d\n", (int)(p.annealing_sequences * p.sampling));
        std::cout << std::endl;
        qsa::clear();
        suqa::clear();
        std::cout << "\nall fine :)\n" << std::endl;
        std::cout << "\n\tSuqa!\n" << std::endl;
        const double secs = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t_start).count();
        std::cout << "All [DONE] in " << secs << " seconds" << std::endl;
    } catch (const std::exception& e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
