// gatecounters.hpp -- gate-count instrumentation of the suqa:: API, the
// interface of the reference's include/gatecounters.hpp:10-163 (GateRecord,
// GateCounter, GateCounterList; same names, same cost model, same printout).
//
// Each suqa:: gate adds the number of 1- and 2-qubit gates of a textbook
// decomposition to the last record of every active counter registered in
// suqa::gatecounters.  A gate controlled by n qubits (the set bits of
// gc_mask plus its own controls) counts as an n-control Toffoli
// (gatecounters.hpp:43-63): n = 0 -> one 1-qubit gate, n = 1 -> one CNOT,
// n = 2 -> 9 one-qubit + 6 CNOT, n > 2 -> 2n - 3 Toffolis and one CNOT.
// With no counter registered nothing is recorded (the reference compiles the
// counting only with -DGATECOUNT).
#pragma once
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

typedef unsigned int uint;

struct GateRecord {
    uint ng1 = 0;   // 1-qubit gates
    uint ng2 = 0;   // 2-qubit gates
    GateRecord(uint g1 = 0, uint g2 = 0) : ng1(g1), ng2(g2) {}
};

class GateCounter {
public:
    std::string name;
    std::vector<GateRecord> grecords;

    explicit GateCounter(std::string nname) : name(std::move(nname)) {}

    void increment_g1g2(uint ng1, uint ng2) {
        if (!active) return;
        grecords.back().ng1 += ng1;
        grecords.back().ng2 += ng2;
    }
    void increment_g1g2(const GateRecord& gr) { increment_g1g2(gr.ng1, gr.ng2); }

    // Mean and standard error of the per-record counts over the records
    // (errors below 1e-10 print as 0).  The squares are accumulated as the
    // reference does (gatecounters.hpp:84-89): the product of two uint counts,
    // modulo 2^32, then added to a double.
    void get_info(double& mean1, double& err1, double& mean2, double& err2, size_t& iters) const {
        iters = grecords.size();
        double s1 = 0.0, s2 = 0.0, q1 = 0.0, q2 = 0.0;
        for (const GateRecord& r : grecords) {
            s1 += r.ng1;
            s2 += r.ng2;
            q1 += (uint)(r.ng1 * r.ng1);
            q2 += (uint)(r.ng2 * r.ng2);
        }
        mean1 = iters ? s1 / (double)iters : 0.0;
        mean2 = iters ? s2 / (double)iters : 0.0;
        err1 = err2 = 0.0;
        if (iters > 1) {
            err1 = std::sqrt(std::fabs(q1 / (double)iters - mean1 * mean1) / ((double)iters - 1.0));
            err2 = std::sqrt(std::fabs(q2 / (double)iters - mean2 * mean2) / ((double)iters - 1.0));
            if (err1 < 1e-10) err1 = 0.0;
            if (err2 < 1e-10) err2 = 0.0;
        }
    }

    void activate() { new_record(); }
    void deactivate() { active = false; }

    // 1- and 2-qubit gates of an n-control Toffoli (n = 0: the gate itself).
    static GateRecord n_ctrl_toffoli_gates(uint n) {
        if (n == 0) return GateRecord(1, 0);
        if (n == 1) return GateRecord(0, 1);
        if (n == 2) return GateRecord(9, 6);
        const uint t = 2 * n - 3;   // 2-control Toffolis
        return GateRecord(9 * t, 6 * t + 1);
    }

    void new_record() {
        active = true;
        grecords.emplace_back();
    }

    void clear() {
        active = false;
        grecords.clear();
    }

    bool is_active() const { return active; }

private:
    bool active = false;
};

class GateCounterList {
public:
    uint gc_mask_set_qbits = 0;   // controls every gate inherits from gc_mask
    std::vector<GateCounter*> counters;

    void add_counter(GateCounter* gctr) { counters.push_back(gctr); }

    void increment_g1g2(uint ng1, uint ng2) {
        for (GateCounter* c : counters) c->increment_g1g2(ng1, ng2);
    }
    void increment_g1g2(const GateRecord& gr) { increment_g1g2(gr.ng1, gr.ng2); }

    void update_cmask_setbits(uint cmask) { gc_mask_set_qbits = (uint)__builtin_popcount(cmask); }

    void print_counts() const {
        if (counters.empty()) return;
        printf("\n\nGate counts\n");
        printf("\n%-15s  %-9s  %-9s  %-9s  %-9s  %-9s  %-9s  %-9s\n", "counter name", "records", "tot_ng1", "<ng1>",
               "d<ng1>", "tot_ng2", "<ng2>", "<dng2>");
        for (const GateCounter* c : counters) {
            size_t n;
            double m1, e1, m2, e2;
            c->get_info(m1, e1, m2, e2, n);
            printf("%-15s  %-9zu  %-9u  %-9.3lg  %-9.3lg  %-9u  %-9.3lg  %-9.3lg\n", c->name.c_str(), n,
                   (uint)std::round(m1 * (double)n), m1, e1, (uint)std::round(m2 * (double)n), m2, e2);
        }
        printf("\n");
    }

    void print_counts_detailed() const {
        if (counters.empty()) return;
        printf("\n\nDetailed gate counts\n");
        for (const GateCounter* c : counters) {
            printf("\ngate counter '%s' (%zu records):\n", c->name.c_str(), c->grecords.size());
            uint k = 0;
            for (const GateRecord& r : c->grecords) printf("record %u: ng1=%u, ng2=%u\n", k++, r.ng1, r.ng2);
        }
        printf("\n");
    }
};
