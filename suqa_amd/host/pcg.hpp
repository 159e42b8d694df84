// pcg.hpp -- the random source of the QMS step: PCG32 (O'Neill, PCG-XSH-RR
// 64->32) and the reference's `pcg` wrapper (Rand.hpp:249-291): seed with
// stream 1, doub() = 32 random mantissa bits in [0,1), randint(a,b) =
// a + next() % (b-a).  The default constructor of the reference seeds from
// the clock; here seed 0 does the same.
#pragma once
#include <sys/time.h>

#include <cstdint>

struct pcg32_engine {
    uint64_t state = 0x853c49e6748fea9bULL;
    uint64_t inc = 0xda3e39cb94b95bdbULL;

    void seed(uint64_t initstate, uint64_t initseq = 1u) {
        state = 0U;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    uint32_t next() {
        const uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
    }
    double next_double() {
        union {
            uint64_t u;
            double d;
        } x;
        x.u = ((uint64_t)next() << 20) | 0x3ff0000000000000ULL;
        return x.d - 1.0;
    }
};

struct pcg_state {
    uint64_t state;
};

class pcg {
    pcg32_engine rng;
    unsigned long long seed_ = 0;

public:
    pcg() : pcg(0ULL) {}
    explicit pcg(unsigned long long j) { set_seed(j); }
    void set_seed(unsigned long long j) {
        if (j == 0) {
            struct timeval tv;
            gettimeofday(&tv, nullptr);
            j = (unsigned long long)tv.tv_sec * 1000000ULL + (unsigned long long)tv.tv_usec;
        }
        seed_ = j;
        rng.seed(j);
    }
    double doub() { return rng.next_double(); }
    int randint(int x0, int x1) { return (int)((unsigned)x0 + rng.next() % (unsigned)(x1 - x0)); }
    pcg_state get_state() const { return pcg_state{rng.state}; }
    unsigned long long get_seed() const { return seed_; }
};
