// rng.hpp -- the reference's measurement RNG, kept on the host.
//
// The reference draws with rand_chacha 0.3.1 ChaCha8Rng seeded by
// rand_core 0.6.4 SeedableRng::seed_from_u64 and samples f64 with rand 0.8.5
// `Standard` (src/monolithic/statevector/simulator.rs:171-174,342-343;
// src/engine.rs:210,427; crate versions from Cargo.lock:376-400).  Restated
// from the published algorithms:
//  * seed_from_u64: PCG32 (XSH-RR) advanced before each 4-byte little-endian
//    chunk of the 32-byte seed (multiplier 6364136223846793005, increment
//    11634580027462260723);
//  * ChaCha with 8 rounds, key = seed, 64-bit block counter from 0 (words
//    12-13), 64-bit stream id 0 (words 14-15), buffered four blocks at a time;
//  * next_u64 = low word | high word << 32;  f64 = (next_u64 >> 11) * 2^-53.
#pragma once
#include <cstdint>
#include <cstring>
#include <random>

namespace dqh {

class ChaCha8Rng {
   public:
    static ChaCha8Rng from_seed(const uint8_t seed[32]) {
        ChaCha8Rng r;
        for (int i = 0; i < 8; ++i)
            r.key_[i] = uint32_t(seed[4 * i]) | (uint32_t(seed[4 * i + 1]) << 8) |
                        (uint32_t(seed[4 * i + 2]) << 16) | (uint32_t(seed[4 * i + 3]) << 24);
        return r;
    }
    static ChaCha8Rng seed_from_u64(uint64_t state) {
        uint8_t seed[32];
        for (int c = 0; c < 8; ++c) {
            state = state * 6364136223846793005ull + 11634580027462260723ull;
            const uint32_t xs = uint32_t(((state >> 18) ^ state) >> 27);
            const uint32_t rot = uint32_t(state >> 59);
            const uint32_t x = (xs >> rot) | (xs << ((32 - rot) & 31));
            for (int b = 0; b < 4; ++b) seed[4 * c + b] = uint8_t(x >> (8 * b));
        }
        return from_seed(seed);
    }
    // rand's ChaCha8Rng::from_entropy: 32 seed bytes from the OS.
    static ChaCha8Rng from_entropy() {
        std::random_device rd;
        uint8_t seed[32];
        for (int i = 0; i < 8; ++i) {
            const uint32_t w = rd();
            for (int b = 0; b < 4; ++b) seed[4 * i + b] = uint8_t(w >> (8 * b));
        }
        return from_seed(seed);
    }
    static uint64_t entropy_u64() {
        std::random_device rd;
        return (uint64_t(rd()) << 32) | rd();
    }

    uint64_t next_u64() {
        if (idx_ < 63) {
            const uint64_t v = uint64_t(buf_[idx_]) | (uint64_t(buf_[idx_ + 1]) << 32);
            idx_ += 2;
            return v;
        }
        if (idx_ >= 64) {
            refill();
            idx_ = 2;
            return uint64_t(buf_[0]) | (uint64_t(buf_[1]) << 32);
        }
        const uint64_t lo = buf_[63];
        refill();
        idx_ = 1;
        return lo | (uint64_t(buf_[0]) << 32);
    }
    double gen_f64() { return double(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

   private:
    uint32_t key_[8] = {};
    uint64_t counter_ = 0;
    uint32_t buf_[64] = {};
    int idx_ = 64;

    static uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
    void block(uint64_t ctr, uint32_t *out) const {
        uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        for (int i = 0; i < 8; ++i) in[4 + i] = key_[i];
        in[12] = uint32_t(ctr);
        in[13] = uint32_t(ctr >> 32);
        in[14] = 0;
        in[15] = 0;
        uint32_t x[16];
        std::memcpy(x, in, sizeof x);
        auto qr = [&](int a, int b, int c, int d) {
            x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 16);
            x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 12);
            x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 8);
            x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 7);
        };
        for (int r = 0; r < 8; r += 2) {
            qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
            qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
        }
        for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
    }
    void refill() {
        for (int b = 0; b < 4; ++b) block(counter_ + uint64_t(b), buf_ + 16 * b);
        counter_ += 4;
    }
};

}  // namespace dqh
