// Modular arithmetic shared by the host code and the gfx950 kernels.
//
// Ring: Z_q[X]/(X^2048 + 1) with q = q0 * q1, q0 = 268369921, q1 = 249561089
// (src/pir/params.rs:9).  Residues are kept per CRT prime; a "condensed" u64
// holds (x mod q0) in bits 0..31 and (x mod q1) in bits 32..63
// (src/pir/packing.rs:867-884).
//
// Every routine returns the canonical representative, so results agree with the
// reference bit for bit even where the reference reduces lazily (its final
// values are canonical; see core/ntt.rs:154-158, core/arith.rs:133-141).
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace inspire {

constexpr int kN = 2048;               // poly_len (protocol/wire.rs:419-420)
constexpr int kLogN = 11;
constexpr int kTExp = 3;               // t_exp_left = t_gsw (pir/params.rs:86-88)
constexpr int kNumRot = kN / 2 - 1;    // 1023 rotations (gpu/packing_online.rs:113)
constexpr int kNInner = kNumRot * kTExp;  // 3069
constexpr uint32_t kQ0 = 268369921u;
constexpr uint32_t kQ1 = 249561089u;
constexpr uint64_t kQ = (uint64_t)kQ0 * kQ1;  // 66974689739603969
constexpr int kGadgetBits = 19;        // floor(56/3)+1 (core/gadget.rs:3-9)
constexpr int kRespBytes = 12288;      // ceil((28*2048 + 20*2048)/64)*8 (modulus_switch.rs:51-53)
constexpr uint64_t kQPrime1 = 1ull << 20;      // pir/params.rs:346-348
constexpr uint64_t kQPrime2 = 268369921ull;    // Q2_VALUES[28] (core/params.rs:37)

// floor(2^64 / q_i): barrett_cr_1 in core/params.rs:263.
constexpr uint64_t kCr1Q0 = 68736257792ull;
constexpr uint64_t kCr1Q1 = 73916747789ull;
// floor(2^58 / q_i) for the 32-bit Barrett below.
constexpr uint32_t kMu58Q0 = (uint32_t)((1ull << 58) / kQ0);
constexpr uint32_t kMu58Q1 = (uint32_t)((1ull << 58) / kQ1);
// 2^32 mod q_i.
constexpr uint32_t kC32Q0 = (uint32_t)((1ull << 32) % kQ0);
constexpr uint32_t kC32Q1 = (uint32_t)((1ull << 32) % kQ1);
// n^{-1} mod q_i for n = 2048.
constexpr uint32_t kNInvQ0 = 268238881u;
constexpr uint32_t kNInvQ1 = 249439233u;
// Garner constant: q0^{-1} mod q1.
constexpr uint32_t kQ0InvModQ1 = 97389680u;

__host__ __device__ constexpr uint32_t prime(int c) { return c == 0 ? kQ0 : kQ1; }
__host__ __device__ constexpr uint32_t mu58(int c) { return c == 0 ? kMu58Q0 : kMu58Q1; }
__host__ __device__ constexpr uint64_t cr1(int c) { return c == 0 ? kCr1Q0 : kCr1Q1; }

__host__ __device__ inline uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

__host__ __device__ inline uint32_t mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

// x mod m for any 64-bit x, with cr = floor(2^64/m) (core/arith.rs:133-141).
__host__ __device__ inline uint64_t barrett64(uint64_t x, uint64_t cr, uint64_t m) {
  uint64_t t = mulhi64(x, cr);
  uint64_t r = x - t * m;
  return r >= m ? r - m : r;
}

// x mod q_c for x < 2^58 using 32-bit operations only.
// t = floor(floor(x/2^26) * mu / 2^32) with mu = floor(2^58/q) satisfies
// floor(x/q) - 2 <= t <= floor(x/q), so x - t*q < 3q < 2^32.
template <int c>
__host__ __device__ inline uint32_t reduce58(uint64_t x) {
  uint32_t xs = (uint32_t)(x >> 26);
  uint32_t t = mulhi32(xs, mu58(c));
  uint32_t r = (uint32_t)x - t * prime(c);
  if (r >= prime(c)) r -= prime(c);
  if (r >= prime(c)) r -= prime(c);
  return r;
}

// reduce58 with the prime as data, without the final reduction: a
// representative of x mod q below 3q for x < 2^58, mu = floor(2^58/q).
__host__ __device__ inline uint32_t reduce58_lazy_rt(uint64_t x, uint32_t q, uint32_t mu) {
  return (uint32_t)x - mulhi32((uint32_t)(x >> 26), mu) * q;
}

// reduce58 with the prime as data: x mod q for x < 2^58.
__host__ __device__ inline uint32_t reduce58_rt(uint64_t x, uint32_t q, uint32_t mu) {
  uint32_t r = reduce58_lazy_rt(x, q, mu);  // < 3q
  r = r >= 2 * q ? r - 2 * q : r;
  return r >= q ? r - q : r;
}

template <int c>
__host__ __device__ inline uint32_t mulmod(uint32_t a, uint32_t b) {
  return reduce58<c>((uint64_t)a * b);
}

// Shoup product w*y mod q in [0, 2q) for y < 2^32, wp = floor(w*2^32/q)
// (the butterfly arithmetic of core/ntt.rs:143-149).
__host__ __device__ inline uint32_t shoup_lazy(uint32_t y, uint32_t w, uint32_t wp, uint32_t q) {
  uint32_t qt = mulhi32(y, wp);
  return w * y - qt * q;
}

// 2^-k mod q (q odd) and the Shoup companion floor(w*2^32/q), at compile time.
__host__ __device__ constexpr uint32_t inv_pow2_mod(uint32_t q, int k) {
  uint64_t r = 1;
  for (int i = 0; i < k; ++i) r = r * ((q + 1) / 2) % q;
  return (uint32_t)r;
}
__host__ __device__ constexpr uint32_t shoup_companion(uint32_t w, uint32_t q) {
  return (uint32_t)(((uint64_t)w << 32) / q);
}

// CRT compose (core/params.rs:203-215): the unique x < q with x = a mod q0 and
// x = b mod q1 (a < q0, b < q1), by Garner's formula
// x = a + q0 * ((b - a) * q0^{-1} mod q1).  Since a < q0 < 2 q1, b + 2 q1 - a
// lies in (0, 3 q1), and one Shoup product by the constant q0^{-1} brings it
// to [0, 2 q1).
__host__ __device__ inline uint64_t crt_compose(uint32_t a, uint32_t b) {
  static_assert(kQ0 < 2 * kQ1, "b + 2 q1 - a > 0");
  constexpr uint32_t kWp = shoup_companion(kQ0InvModQ1, kQ1);
  uint32_t t = shoup_lazy(b + 2 * kQ1 - a, kQ0InvModQ1, kWp, kQ1);
  t = t >= kQ1 ? t - kQ1 : t;
  return (uint64_t)a + (uint64_t)kQ0 * t;
}

// x mod q for x < 4q (sums of at most four canonical values).
__host__ __device__ inline uint64_t reduce_small_q(uint64_t x) {
  if (x >= 2 * kQ) x -= 2 * kQ;
  if (x >= kQ) x -= kQ;
  return x;
}

// rescale(a, q, out_mod) of core/arith.rs:436-451 for a canonical a < q:
// round(a_centered * out_mod / q) (ties away from zero), mapped to [0, out_mod).
__host__ __device__ inline uint64_t rescale_from_q(uint64_t a, uint64_t out_mod) {
  const int64_t q = (int64_t)kQ;
  int64_t v = (int64_t)a;
  if (v >= q / 2) v -= q;
  // |v| < 2^55 and out_mod <= 2^28: the product needs 83 bits.
  bool neg = v < 0;
  uint64_t av = neg ? (uint64_t)(-v) : (uint64_t)v;
  // floor((av*out_mod + q/2) / q), computed in 128 bits.
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t plo = av * out_mod;
  uint64_t phi = __umul64hi(av, out_mod);
  uint64_t half = (uint64_t)(q / 2);
  uint64_t slo = plo + half;
  phi += (slo < plo);
  // divide (phi:slo) by q; phi < 2^20 so the quotient fits in 64 bits.
  // Long division in two 32-bit digits of a 2^64-radix numerator.
  uint64_t quot = 0, rem = phi;
  for (int i = 63; i >= 0; --i) {
    rem = (rem << 1) | ((slo >> i) & 1);
    quot <<= 1;
    if (rem >= (uint64_t)q) { rem -= (uint64_t)q; quot |= 1; }
  }
#else
  unsigned __int128 num = (unsigned __int128)av * out_mod + (uint64_t)(q / 2);
  uint64_t quot = (uint64_t)(num / (uint64_t)q);
#endif
  int64_t r = neg ? -(int64_t)quot : (int64_t)quot;
  int64_t om = (int64_t)out_mod;
  r %= om;
  if (r < 0) r += om;
  return (uint64_t)r;
}

// Counter-based generator for synthetic server state and queries.  The CPU
// oracle implements the same function; tests compare device fills with it.
__host__ __device__ inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__host__ __device__ inline uint64_t synth_word(uint64_t seed, uint32_t stream, uint64_t idx) {
  return mix64(mix64(seed ^ ((uint64_t)stream << 40)) + idx);
}

}  // namespace inspire
