#pragma once
#include <cstddef>
#include <cstdint>

#include "modarith.hpp"

namespace inspire {

// out: [2 primes][4][kN] u32: forward w, forward w' = floor(w*2^32/q),
// inverse w (halved), inverse w'.
void build_ntt_tables(uint32_t* out);

// In-place forward NTT of one residue polynomial (canonical in, canonical out).
void host_ntt_forward(uint32_t* a, int c, const uint32_t* tables);

// In-place inverse NTT of one residue polynomial, scaled by n^-1 (the
// halved inverse powers of build_ntt_tables); canonical in and out.
void host_ntt_inverse(uint32_t* a, int c, const uint32_t* tables);

// ChaCha20Rng::from_seed(seed) followed by n u64 draws (low word first).
void chacha20_u64(const uint8_t seed[32], uint64_t* out, size_t n);

// ChaCha20 block (RFC 8439 state with a 64-bit block counter in words 12-13
// and a zero stream id, as rand_chacha uses it): 16 output words.
void chacha20_block_words(const uint8_t key[32], uint64_t counter, uint32_t out[16]);

// out: [2][kN] = NTT(-a mod q_c) for the constant RGSW row-0 polynomial a.
void build_rgsw_neg_a(uint32_t* out, const uint32_t* tables);

// out: [2 primes][3][kN] u32 for the RGSW fold: the inverse NTT powers (not
// halved), their w' = floor(w*2^32/q), and n^-1 * NTT(-a) (neg_a as built by
// build_rgsw_neg_a); then [2 primes][kN]: the w' of n^-1 * NTT(-a), for its
// Shoup products (kFoldTableWords words in all).  With them the fold's
// inverse NTTs skip the per-stage halving: n^-1 is applied once, to the
// product rows before the transform.
constexpr int kFoldTableWords = 2 * 4 * kN;
void build_fold_tables(uint32_t* out, const uint32_t* tables, const uint32_t* neg_a);

// out: [N] discrete logs of the odd exponents mod 2N, (-1)^w 5^i -> i | w << 15.
void build_slot_dlog(uint16_t* out);

}  // namespace inspire
