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

// ChaCha20Rng::from_seed(seed) followed by n u64 draws (low word first).
void chacha20_u64(const uint8_t seed[32], uint64_t* out, size_t n);

// out: [2][kN] = NTT(-a mod q_c) for the constant RGSW row-0 polynomial a.
void build_rgsw_neg_a(uint32_t* out, const uint32_t* tables);

// Packing GEMM term order: kPackTerms = 6141 terms padded to 96 steps of 64.
constexpr int kPackTerms = 6144;
// Key rows: 6 planes of kN slots, then one zero row for the padding terms.
constexpr uint32_t kTermRowZero = 6 * kN;
// out: [kN][kPackTerms] key row read by term n at NTT slot z.
void build_packing_term_sources(uint16_t* out);

}  // namespace inspire
