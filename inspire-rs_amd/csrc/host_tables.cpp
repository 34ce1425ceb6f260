// Host-side constant tables for the answer path:
//   - NTT twiddles for both CRT primes (core/ntt.rs:74-108 build_ntt_tables),
//   - the RGSW row-0 polynomial NTT(-a) (keyword.rs:546-555), where a is drawn
//     from ChaCha20Rng::from_seed([7;32]) (protocol/wire.rs:88-113,
//     core/poly.rs:211-231).
#include "host_tables.hpp"

#include <cstring>

namespace inspire {
namespace {

uint64_t mulm(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)(((unsigned __int128)a * b) % m); }

uint64_t powm(uint64_t b, uint64_t e, uint64_t m) {
  uint64_t r = 1;
  b %= m;
  while (e) {
    if (e & 1) r = mulm(r, b, m);
    b = mulm(b, b, m);
    e >>= 1;
  }
  return r;
}

uint64_t invm(uint64_t a, uint64_t m) { return powm(a, m - 2, m); }  // m prime

uint32_t rev11(uint32_t x) {
  uint32_t r = 0;
  for (int i = 0; i < kLogN; ++i) r |= ((x >> i) & 1u) << (kLogN - 1 - i);
  return r;
}

// number_theory.rs:41-55: the smallest primitive 2n-th root of unity mod m.
uint64_t min_primitive_root(uint64_t m) {
  const uint64_t deg = 2 * kN;
  uint64_t root = 0;
  for (uint64_t g = 2; g < m; ++g) {
    uint64_t r = powm(g, (m - 1) / deg, m);
    if (powm(r, deg / 2, m) == m - 1) {
      root = r;
      break;
    }
  }
  const uint64_t sq = mulm(root, root, m);
  uint64_t cur = root, best = root;
  for (uint64_t i = 0; i < deg; ++i) {
    if (cur < best) best = cur;
    cur = mulm(cur, sq, m);
  }
  return best;
}

// ntt.rs:6-17: table[rev(i)] = root^i.
void bitrev_powers(uint64_t root, uint64_t m, uint64_t* out) {
  uint64_t p = root;
  for (uint32_t i = 1; i < (uint32_t)kN; ++i) {
    out[rev11(i)] = p;
    p = mulm(p, root, m);
  }
  out[0] = 1;
}

void chacha_block(const uint8_t key[32], uint64_t counter, uint32_t out[16]) {
  auto rotl = [](uint32_t v, int c) { return (v << c) | (v >> (32 - c)); };
  uint32_t s[16];
  s[0] = 0x61707865u;
  s[1] = 0x3320646eu;
  s[2] = 0x79622d32u;
  s[3] = 0x6b206574u;
  for (int i = 0; i < 8; ++i) {
    uint32_t w;
    std::memcpy(&w, key + 4 * i, 4);  // little-endian host
    s[4 + i] = w;
  }
  s[12] = (uint32_t)counter;
  s[13] = (uint32_t)(counter >> 32);
  s[14] = 0;
  s[15] = 0;
  uint32_t x[16];
  std::memcpy(x, s, sizeof(s));
  auto qr = [&](int a, int b, int c, int d) {
    x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 16);
    x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 12);
    x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 8);
    x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 7);
  };
  for (int r = 0; r < 10; ++r) {
    qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
    qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
  }
  for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
}

}  // namespace

void build_ntt_tables(uint32_t* out) {
  for (int c = 0; c < 2; ++c) {
    const uint64_t m = prime(c);
    const uint64_t root = min_primitive_root(m);
    const uint64_t iroot = invm(root, m);
    uint64_t fw[kN], iw[kN];
    bitrev_powers(root, m, fw);
    bitrev_powers(iroot, m, iw);
    for (int i = 0; i < kN; ++i) iw[i] = (iw[i] & 1) ? (iw[i] + m) >> 1 : iw[i] >> 1;  // div2 mod m
    uint32_t* t = out + c * 4 * kN;
    for (int i = 0; i < kN; ++i) {
      t[0 * kN + i] = (uint32_t)fw[i];
      t[1 * kN + i] = (uint32_t)((fw[i] << 32) / m);
      t[2 * kN + i] = (uint32_t)iw[i];
      t[3 * kN + i] = (uint32_t)((iw[i] << 32) / m);
    }
  }
}

void host_ntt_forward(uint32_t* a, int c, const uint32_t* tables) {
  const uint32_t q = prime(c), q2 = 2 * q;
  const uint32_t* w_t = tables + (c * 4 + 0) * kN;
  const uint32_t* wp_t = tables + (c * 4 + 1) * kN;
  for (int mm = 0; mm < kLogN; ++mm) {
    const int m = 1 << mm, t = kN >> (mm + 1);
    for (int i = 0; i < m; ++i) {
      const uint32_t w = w_t[m + i], wp = wp_t[m + i];
      uint32_t* o = a + 2 * t * i;
      for (int j = 0; j < t; ++j) {
        uint32_t x = o[j];
        x = x >= q2 ? x - q2 : x;
        const uint32_t qn = shoup_lazy(o[t + j], w, wp, q);
        o[j] = x + qn;
        o[t + j] = x + q2 - qn;
      }
    }
  }
  for (int i = 0; i < kN; ++i) {
    uint32_t v = a[i];
    v = v >= q2 ? v - q2 : v;
    a[i] = v >= q ? v - q : v;
  }
}

void host_ntt_inverse(uint32_t* a, int c, const uint32_t* tables) {
  // Gentleman-Sande with the halved inverse powers: each stage halves, so the
  // result carries n^-1.  Values stay canonical.
  const uint32_t q = prime(c);
  const uint32_t* w_t = tables + (c * 4 + 2) * kN;
  for (int mm = kLogN - 1; mm >= 0; --mm) {
    const int m = 1 << mm, t = kN >> (mm + 1);
    for (int i = 0; i < m; ++i) {
      const uint64_t w = w_t[m + i];
      uint32_t* o = a + 2 * t * i;
      for (int j = 0; j < t; ++j) {
        const uint64_t x = o[j], y = o[t + j];
        uint64_t s = x + y;
        s = s >= q ? s - q : s;
        o[j] = (uint32_t)((s & 1) ? (s + q) >> 1 : s >> 1);  // (x + y) / 2
        o[t + j] = (uint32_t)(((x + q - y) * w) % q);         // (x - y) * w / 2
      }
    }
  }
}

void chacha20_block_words(const uint8_t key[32], uint64_t counter, uint32_t out[16]) {
  chacha_block(key, counter, out);
}

void chacha20_u64(const uint8_t seed[32], uint64_t* out, size_t n) {
  uint32_t blk[16];
  uint64_t counter = 0;
  size_t k = 0;
  while (k < n) {
    chacha_block(seed, counter++, blk);
    for (int i = 0; i < 8 && k < n; ++i) out[k++] = (uint64_t)blk[2 * i] | ((uint64_t)blk[2 * i + 1] << 32);
  }
}

void build_rgsw_neg_a(uint32_t* out, const uint32_t* tables) {
  uint8_t seed[32];
  std::memset(seed, 7, sizeof(seed));
  uint64_t a[kN];
  chacha20_u64(seed, a, kN);
  for (int c = 0; c < 2; ++c) {
    for (int i = 0; i < kN; ++i) {
      const uint64_t av = a[i] % kQ;           // random_rng: val % modulus
      const uint64_t neg = kQ - av;            // invert_poly: modulus - a (poly.rs:520-524)
      out[c * kN + i] = (uint32_t)(neg % prime(c));  // reduce_copy (poly.rs:857-863)
    }
    host_ntt_forward(out + c * kN, c, tables);
  }
}

void build_fold_tables(uint32_t* out, const uint32_t* tables, const uint32_t* neg_a) {
  for (int c = 0; c < 2; ++c) {
    const uint64_t m = prime(c);
    const uint64_t ninv = invm(kN, m);
    const uint32_t* iw = tables + (c * 4 + 2) * kN;  // inverse powers, halved
    uint32_t* t = out + c * 3 * kN;
    uint32_t* tp = out + 6 * kN + c * kN;
    for (int i = 0; i < kN; ++i) {
      const uint64_t w = (2 * (uint64_t)iw[i]) % m;  // the inverse power itself
      t[0 * kN + i] = (uint32_t)w;
      t[1 * kN + i] = (uint32_t)((w << 32) / m);
      t[2 * kN + i] = (uint32_t)mulm(neg_a[c * kN + i], ninv, m);
      tp[i] = (uint32_t)(((uint64_t)t[2 * kN + i] << 32) / m);
    }
  }
}

}  // namespace inspire

namespace inspire {

// Discrete logs for the slot-ordered packing GEMM: for the odd exponent
// e = (-1)^w 5^i mod 2N (i < N/2), out[e >> 1] = i | (w << 15).  The packing
// term of rotation i (packing.rs:1639-1652, automorphism X -> X^(5^i)) reads
// key slot sigma at NTT slot z exactly when the exponent of sigma equals the
// exponent of z times 5^i (w = 0) or minus it (w = 1).
void build_slot_dlog(uint16_t* out) {
  constexpr uint32_t kMask = 2 * kN - 1;
  uint32_t g = 1;
  for (int i = 0; i < kN / 2; ++i) {
    out[g >> 1] = (uint16_t)i;
    out[((2 * kN - g) & kMask) >> 1] = (uint16_t)(i | 0x8000);
    g = (g * 5u) & kMask;
  }
}

}  // namespace inspire
