// qms_batch.hpp -- per-chain constants of the batched QMS kernel.
//
// Everything a chain needs that the reference evaluates with host libm is
// evaluated on the host here too, so the device applies the same doubles:
// sincos of every rotation phase (XX evolutions, energy-shift u1, QFT crm,
// the measure rotation's u1(pi/2)), the energy value of each register
// reading, and the acceptance rotation (fs1, fs2) for every (E_old, E_new).
#pragma once
#include <cstdint>

namespace qb {

constexpr int kMaxEne = 6;                  // energy qubits per chain: 2..6 (state 2^6..2^10)
constexpr int kMaxL = 1 << kMaxEne;

struct ChainDev {
    uint64_t seed;
    uint32_t r;                             // energy qubits
    uint32_t pad_;
    uint64_t state_off;                     // first double of the chain's final state in BatchDev::final_state
    double xx_c[2][kMaxEne], xx_s[2][kMaxEne];      // [Phi, Phi^dag][trg]
    double u1_c[2][kMaxEne], u1_s[2][kMaxEne];      // energy shift u1 on register qubit trg
    double qft_c[2][kMaxEne + 2], qft_s[2][kMaxEne + 2];  // [+, -][m]: crm phase +-2 pi / 2^m
    double hp_c, hp_s;                      // u1(pi/2) of the measure rotation
    double etab[kMaxL];                     // E value of register reading k
    const double* wtab;                     // [L][L][2]: (fs1, fs2) for (E_old, E_new)
};

struct BatchDev {
    const ChainDev* chains;
    const uint32_t* order;                  // chain indices; this launch runs order[first .. nchains)
    uint32_t first;
    uint32_t nchains;
    uint32_t steps, reset_each, max_reverse;
    int32_t thermalization;
    uint32_t cap_meas;
    uint32_t* next_chain;                   // work-queue counter (zeroed before launch)
    int32_t* rets;                          // [nchains][steps]
    uint64_t* rng_after;                    // [nchains][steps] or null
    double* E;                              // [nchains][cap_meas]
    double* X;
    uint32_t* n_meas;                       // [nchains]
    uint32_t* accepted;                     // [nchains]
    uint32_t two_registers;                 // E_old in its own register: 2r + 4 qubits per chain
    uint64_t* lds_amps;                     // [nchains] LDS amplitude accesses per chain, or null
    double* final_state;                    // the chains' stored amplitudes after the last step,
                                            // (re, im) in FP64, at ChainDev::state_off; or null
};

}  // namespace qb
