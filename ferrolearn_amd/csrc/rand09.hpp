// rand09.hpp -- the reference's random stream: `StdRng::seed_from_u64` of
// rand 0.9.4 (Cargo.lock:2413-2420), which is ChaCha12Rng of rand_chacha
// 0.9.0 seeded through rand_core 0.9.5's default `seed_from_u64`, with the
// two draws ferrolearn-cluster makes from it (kmeans.rs:178,201,209;
// mini_batch_kmeans.rs:239,268,280,332,457,467):
//
//   random_range(lo..hi) over usize -- rand 0.9's UniformUsize: when
//     hi - 1 fits in 32 bits it samples as u32, else as u64, each with
//     Canon's method (one widening multiply, a second draw only when the low
//     half of the product lies in the biased zone);
//   random::<f64>() -- StandardUniform: (next_u64 >> 11) * 2^-53.
//
// The generator itself: seed_from_u64 fills the 32-byte key with 8 outputs
// of a PCG32 step (state = state * 6364136223846793005 + 11634580027462260723,
// then XSH-RR: rotr(((s >> 18) ^ s) >> 27, s >> 59)), little-endian; the
// ChaCha state is "expand 32-byte k", the key, a 64-bit block counter from 0
// (words 12-13) and a zero stream id (words 14-15); each block is 12 rounds
// (6 column + diagonal double rounds) plus the input, and the u32 outputs are
// the blocks' words in order (next_u64: low word first).  Host code: the
// draws are few and inherently sequential.
#pragma once

#include <cstdint>

namespace fl {

inline uint32_t rotl32(uint32_t v, int k) { return (v << k) | (v >> (32 - k)); }

// One ChaCha block of `rounds` rounds (even) from state `in` into `out`.
inline void chacha_block(const uint32_t in[16], int rounds, uint32_t out[16]) {
    uint32_t x[16];
    for (int i = 0; i < 16; ++i) x[i] = in[i];
    auto qr = [&](int a, int b, int c, int d) {
        x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 16);
        x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 12);
        x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 8);
        x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 7);
    };
    for (int r = 0; r < rounds; r += 2) {
        qr(0, 4, 8, 12);
        qr(1, 5, 9, 13);
        qr(2, 6, 10, 14);
        qr(3, 7, 11, 15);
        qr(0, 5, 10, 15);
        qr(1, 6, 11, 12);
        qr(2, 7, 8, 13);
        qr(3, 4, 9, 14);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

class StdRng09 {
  public:
    explicit StdRng09(uint64_t seed) {
        uint64_t s = seed;
        for (int i = 0; i < 8; ++i) {
            s = s * 6364136223846793005ULL + 11634580027462260723ULL;
            const uint32_t xs = (uint32_t)(((s >> 18) ^ s) >> 27);
            const int rot = (int)(s >> 59);
            key_[i] = rot ? (xs >> rot) | (xs << (32 - rot)) : xs;
        }
    }
    uint32_t next_u32() {
        if (idx_ == 16) refill();
        return buf_[idx_++];
    }
    uint64_t next_u64() {
        const uint64_t lo = next_u32();
        const uint64_t hi = next_u32();
        return (hi << 32) | lo;
    }
    // random::<f64>()
    double f64() { return (double)(next_u64() >> 11) * 0x1.0p-53; }
    // random_range(lo..hi), lo < hi
    int64_t range_from(int64_t lo, int64_t hi) {
        const uint64_t low = (uint64_t)lo, high = (uint64_t)hi - 1;   // inclusive
        if ((high >> 32) == 0) {
            const uint32_t r = (uint32_t)high - (uint32_t)low + 1u;
            if (r == 0) return (int64_t)(uint32_t)((uint32_t)low + next_u32());
            const uint64_t m = (uint64_t)next_u32() * r;
            uint32_t result = (uint32_t)(m >> 32);
            const uint32_t lo_order = (uint32_t)m;
            if (lo_order > 0u - r) {
                const uint32_t nh = (uint32_t)(((uint64_t)next_u32() * r) >> 32);
                if ((uint64_t)lo_order + nh > 0xffffffffULL) result += 1u;
            }
            return (int64_t)(uint32_t)((uint32_t)low + result);
        }
        const uint64_t r = high - low + 1;
        if (r == 0) return (int64_t)(low + next_u64());
        const unsigned __int128 m = (unsigned __int128)next_u64() * r;
        uint64_t result = (uint64_t)(m >> 64);
        const uint64_t lo_order = (uint64_t)m;
        if (lo_order > 0ULL - r) {
            const uint64_t nh = (uint64_t)(((unsigned __int128)next_u64() * r) >> 64);
            if (lo_order + nh < lo_order) result += 1;   // the add overflowed
        }
        return (int64_t)(low + result);
    }
    // random_range(0..n)
    int64_t range(int64_t n) { return range_from(0, n); }

  private:
    void refill() {
        uint32_t st[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        for (int i = 0; i < 8; ++i) st[4 + i] = key_[i];
        st[12] = (uint32_t)ctr_;
        st[13] = (uint32_t)(ctr_ >> 32);
        st[14] = st[15] = 0;
        chacha_block(st, 12, buf_);
        ++ctr_;
        idx_ = 0;
    }
    uint32_t key_[8];
    uint64_t ctr_ = 0;
    uint32_t buf_[16] = {};
    int idx_ = 16;
};

}  // namespace fl
