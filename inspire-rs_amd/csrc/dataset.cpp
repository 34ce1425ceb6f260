// DatasetGenerator and the PQKEYV2 writer (src/dataset/generator.rs:61-193,
// kem.rs:27-53, writer.rs:17-79, dataset.rs:8-46) on the host: deterministic
// ML-KEM encapsulation keys, one per record, from a ChaCha20 stream seeded by a
// u64, each stored with its id SHA3-256("pq-key:v1:" || kem name || key).
//
// Record idx: the 64-byte seed is ChaCha20 keystream block idx (20 rounds,
// 64-bit block counter in words 12-13, stream 0, little-endian words: the
// rand_chacha 0.10 ChaCha20Rng at word position 16*idx, generator.rs:170-173)
// under the 32-byte key that rand_core's seed_from_u64 expands from the u64
// with PCG32 (multiplier 6364136223846793005, increment 11634580027462260723,
// state advanced before each output, XSH-RR output, little-endian words).
// Its halves are (d, z) of ML-KEM.KeyGen_internal (FIPS 203 Algorithm 16);
// the encapsulation key is K-PKE.KeyGen(d)'s ek (Algorithm 13):
// (rho, sigma) = SHA3-512(d || k), A-hat[i][j] = SampleNTT(SHAKE128(rho || j || i)),
// s, e from SamplePolyCBD_eta1(SHAKE256(sigma || N)), t-hat = A-hat o NTT(s) + NTT(e),
// ek = ByteEncode_12(t-hat) || rho.
// Keccak-f[1600] and its sponges follow FIPS 202; the round constants and the
// rotation offsets are derived by that standard's LFSR (Algorithm 5) and
// (x, y) walk (Algorithm 2) at start-up.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/inspire_hip.h"
#include "host_tables.hpp"
#include "runtime.hpp"

using namespace inspire;

namespace {

// ---------------------------------------------------------------------------
// FIPS 202: Keccak-f[1600], SHA3-256/512, SHAKE128/256
// ---------------------------------------------------------------------------

struct KeccakConsts {
  uint64_t rc[24];
  int rot[25];
  KeccakConsts() {
    // rc(t) of Algorithm 5: an LFSR over GF(2) with x^8 + x^6 + x^5 + x^4 + 1
    auto rcbit = [](int t) {
      if (t % 255 == 0) return 1;
      unsigned r = 1;  // bits r[0..7], r[0] first
      for (int i = 1; i <= t % 255; ++i) {
        unsigned r8 = r << 1;  // prepend a zero: R = 0 || R
        const unsigned b8 = (r8 >> 8) & 1;
        r8 ^= b8 | (b8 << 4) | (b8 << 5) | (b8 << 6);
        r = r8 & 0xff;
      }
      return (int)(r & 1);
    };
    for (int ir = 0; ir < 24; ++ir) {
      uint64_t c = 0;
      for (int j = 0; j <= 6; ++j)
        if (rcbit(j + 7 * ir)) c |= 1ull << ((1 << j) - 1);
      rc[ir] = c;
    }
    // rho offsets: (x, y) = (1, 0), then (y, (2x + 3y) mod 5), offset (t+1)(t+2)/2
    for (int i = 0; i < 25; ++i) rot[i] = 0;
    int x = 1, y = 0;
    for (int t = 0; t < 24; ++t) {
      rot[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
      const int nx = y, ny = (2 * x + 3 * y) % 5;
      x = nx;
      y = ny;
    }
  }
};

const KeccakConsts& keccak_consts() {
  static const KeccakConsts k;
  return k;
}

inline uint64_t rotl(uint64_t v, int r) { return r ? (v << r) | (v >> (64 - r)) : v; }

void keccak_f(uint64_t a[25]) {
  const KeccakConsts& k = keccak_consts();
  for (int round = 0; round < 24; ++round) {
    uint64_t c[5], d[5], b[25];
    for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];  // theta
    for (int x = 0; x < 5; ++x) d[x] = c[(x + 4) % 5] ^ rotl(c[(x + 1) % 5], 1);
    for (int i = 0; i < 25; ++i) a[i] ^= d[i % 5];
    for (int x = 0; x < 5; ++x)  // rho and pi: B[y][2x + 3y] = rot(A[x][y])
      for (int y = 0; y < 5; ++y) b[y + 5 * ((2 * x + 3 * y) % 5)] = rotl(a[x + 5 * y], k.rot[x + 5 * y]);
    for (int y = 0; y < 5; ++y)  // chi
      for (int x = 0; x < 5; ++x)
        a[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
    a[0] ^= k.rc[round];  // iota
  }
}

// Sponge with `rate` bytes and domain byte `pad` (0x06 SHA3, 0x1f SHAKE).
class Sponge {
 public:
  Sponge(int rate, uint8_t pad) : rate_(rate), pad_(pad) { std::memset(a_, 0, sizeof(a_)); }
  void absorb(const uint8_t* p, size_t n) {
    while (n) {
      const size_t take = std::min(n, (size_t)(rate_ - pos_));
      for (size_t i = 0; i < take; ++i) xor_byte(pos_ + i, p[i]);
      pos_ += (int)take;
      p += take;
      n -= take;
      if (pos_ == rate_) {
        keccak_f(a_);
        pos_ = 0;
      }
    }
  }
  void squeeze(uint8_t* out, size_t n) {
    if (!squeezing_) {
      xor_byte(pos_, pad_);
      xor_byte(rate_ - 1, 0x80);
      keccak_f(a_);
      pos_ = 0;
      squeezing_ = true;
    }
    while (n) {
      if (pos_ == rate_) {
        keccak_f(a_);
        pos_ = 0;
      }
      const size_t take = std::min(n, (size_t)(rate_ - pos_));
      for (size_t i = 0; i < take; ++i) out[i] = get_byte(pos_ + i);
      pos_ += (int)take;
      out += take;
      n -= take;
    }
  }

 private:
  void xor_byte(int i, uint8_t v) { a_[i / 8] ^= (uint64_t)v << (8 * (i % 8)); }
  uint8_t get_byte(int i) const { return (uint8_t)(a_[i / 8] >> (8 * (i % 8))); }
  uint64_t a_[25];
  int rate_, pos_ = 0;
  uint8_t pad_;
  bool squeezing_ = false;
};

Sponge sha3_256() { return Sponge(136, 0x06); }
Sponge sha3_512() { return Sponge(72, 0x06); }
Sponge shake128() { return Sponge(168, 0x1f); }
Sponge shake256() { return Sponge(136, 0x1f); }

// ---------------------------------------------------------------------------
// Record seeds: rand_core seed_from_u64 (PCG32) + ChaCha20 block idx
// ---------------------------------------------------------------------------

void seed_from_u64(uint64_t state, uint8_t key[32]) {
  const uint64_t mul = 6364136223846793005ull, inc = 11634580027462260723ull;
  for (int w = 0; w < 8; ++w) {
    state = state * mul + inc;
    const uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
    const uint32_t rot = (uint32_t)(state >> 59);
    const uint32_t x = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
    for (int b = 0; b < 4; ++b) key[4 * w + b] = (uint8_t)(x >> (8 * b));
  }
}

// RFC 8439 section 2.3 block function with a 64-bit counter in words 12-13
// and words 14-15 zero (the library's one ChaCha20 core, host_tables.cpp),
// serialized little-endian.
void chacha20_block(const uint8_t key[32], uint64_t counter, uint8_t out[64]) {
  uint32_t w[16];
  chacha20_block_words(key, counter, w);
  for (int i = 0; i < 16; ++i)
    for (int b = 0; b < 4; ++b) out[4 * i + b] = (uint8_t)(w[i] >> (8 * b));
}

// ---------------------------------------------------------------------------
// FIPS 203 K-PKE.KeyGen -> encapsulation key
// ---------------------------------------------------------------------------

constexpr int kQ = 3329;
constexpr int kPolyN = 256;

struct MlKemConsts {
  uint16_t zeta[128];   // 17^BitRev7(i) mod q
  MlKemConsts() {
    for (int i = 0; i < 128; ++i) {
      int r = 0;
      for (int b = 0; b < 7; ++b) r |= ((i >> b) & 1) << (6 - b);
      uint32_t v = 1;
      for (int e = 0; e < r; ++e) v = v * 17 % kQ;
      zeta[i] = (uint16_t)v;
    }
  }
};

const MlKemConsts& mlkem_consts() {
  static const MlKemConsts c;
  return c;
}

struct MlKemParams {
  int k, eta1;
  size_t ek_bytes;
};

bool mlkem_params(const std::string& name, MlKemParams* p) {
  if (name == "ml-kem-512") *p = {2, 3, 800};
  else if (name == "ml-kem-768") *p = {3, 2, 1184};
  else if (name == "ml-kem-1024") *p = {4, 2, 1568};
  else return false;
  return true;
}

// Algorithm 9 (NTT), in place, values in [0, q).
void ntt256(uint16_t f[kPolyN]) {
  const MlKemConsts& c = mlkem_consts();
  int i = 1;
  for (int len = 128; len >= 2; len >>= 1)
    for (int start = 0; start < kPolyN; start += 2 * len) {
      const uint32_t z = c.zeta[i++];
      for (int j = start; j < start + len; ++j) {
        const uint32_t t = z * f[j + len] % kQ;
        f[j + len] = (uint16_t)((f[j] + kQ - t) % kQ);
        f[j] = (uint16_t)((f[j] + t) % kQ);
      }
    }
}

// Algorithms 11-12 (MultiplyNTTs / BaseCaseMultiply), accumulated into acc.
void ntt_mul_acc(uint32_t acc[kPolyN], const uint16_t f[kPolyN], const uint16_t g[kPolyN]) {
  const MlKemConsts& c = mlkem_consts();
  for (int i = 0; i < 128; ++i) {
    // gamma = 17^(2 BitRev7(i) + 1) = zeta[64 + i/2] for even i, its negative for odd i
    const uint32_t z = c.zeta[64 + i / 2];
    const uint32_t gamma = (i & 1) ? (uint32_t)(kQ - z) % kQ : z;
    const uint32_t a0 = f[2 * i], a1 = f[2 * i + 1], b0 = g[2 * i], b1 = g[2 * i + 1];
    acc[2 * i] = (acc[2 * i] + a0 * b0 % kQ + a1 * b1 % kQ * gamma) % kQ;
    acc[2 * i + 1] = (acc[2 * i + 1] + a0 * b1 + a1 * b0) % kQ;
  }
}

// Algorithm 7 (SampleNTT) from SHAKE128(rho || j || i).
void sample_ntt(const uint8_t rho[32], int j, int i, uint16_t out[kPolyN]) {
  Sponge x = shake128();
  uint8_t in[34];
  std::memcpy(in, rho, 32);
  in[32] = (uint8_t)j;
  in[33] = (uint8_t)i;
  x.absorb(in, 34);
  int n = 0;
  uint8_t b[168];
  while (n < kPolyN) {
    x.squeeze(b, sizeof(b));
    for (int t = 0; t + 3 <= (int)sizeof(b) && n < kPolyN; t += 3) {
      const int d1 = b[t] | ((b[t + 1] & 15) << 8);
      const int d2 = (b[t + 1] >> 4) | (b[t + 2] << 4);
      if (d1 < kQ) out[n++] = (uint16_t)d1;
      if (d2 < kQ && n < kPolyN) out[n++] = (uint16_t)d2;
    }
  }
}

// Algorithm 8 (SamplePolyCBD_eta) of PRF_eta(sigma, N) = SHAKE256(sigma || N).
void sample_cbd(const uint8_t sigma[32], uint8_t nonce, int eta, uint16_t out[kPolyN]) {
  Sponge prf = shake256();
  uint8_t in[33];
  std::memcpy(in, sigma, 32);
  in[32] = nonce;
  prf.absorb(in, 33);
  uint8_t buf[64 * 3];
  prf.squeeze(buf, (size_t)64 * eta);
  auto bit = [&](int k) { return (buf[k >> 3] >> (k & 7)) & 1; };
  for (int i = 0; i < kPolyN; ++i) {
    int x = 0, y = 0;
    for (int j = 0; j < eta; ++j) {
      x += bit(2 * i * eta + j);
      y += bit(2 * i * eta + eta + j);
    }
    out[i] = (uint16_t)((x - y + kQ) % kQ);
  }
}

// ML-KEM.KeyGen_internal(d, z)'s encapsulation key (z only enters dk).
void mlkem_ek(const MlKemParams& P, const uint8_t d[32], uint8_t* ek) {
  uint8_t g_in[33], rs[64];
  std::memcpy(g_in, d, 32);
  g_in[32] = (uint8_t)P.k;
  Sponge g = sha3_512();
  g.absorb(g_in, 33);
  g.squeeze(rs, 64);
  const uint8_t* rho = rs;
  const uint8_t* sigma = rs + 32;
  uint16_t s[4][kPolyN], e[4][kPolyN], a[kPolyN];
  uint8_t nonce = 0;
  for (int i = 0; i < P.k; ++i) sample_cbd(sigma, nonce++, P.eta1, s[i]);
  for (int i = 0; i < P.k; ++i) sample_cbd(sigma, nonce++, P.eta1, e[i]);
  for (int i = 0; i < P.k; ++i) {
    ntt256(s[i]);
    ntt256(e[i]);
  }
  for (int i = 0; i < P.k; ++i) {
    uint32_t t[kPolyN];
    for (int z = 0; z < kPolyN; ++z) t[z] = e[i][z];
    for (int j = 0; j < P.k; ++j) {
      sample_ntt(rho, j, i, a);
      ntt_mul_acc(t, a, s[j]);
    }
    uint8_t* o = ek + (size_t)384 * i;  // ByteEncode_12
    for (int z = 0; z < kPolyN; z += 2) {
      o[3 * z / 2] = (uint8_t)(t[z] & 0xff);
      o[3 * z / 2 + 1] = (uint8_t)((t[z] >> 8) | ((t[z + 1] & 15) << 4));
      o[3 * z / 2 + 2] = (uint8_t)(t[z + 1] >> 4);
    }
  }
  std::memcpy(ek + (size_t)384 * P.k, rho, 32);
}

const char kIdDomain[] = "pq-key:v1:";  // dataset.rs:11

void record(const MlKemParams& P, const std::string& kem, const uint8_t key[32], uint64_t idx, uint8_t* id,
            uint8_t* pk) {
  uint8_t seed[64];
  chacha20_block(key, idx, seed);
  mlkem_ek(P, seed, pk);  // d = seed[0..32]; z = seed[32..64] enters only dk
  Sponge h = sha3_256();
  h.absorb(reinterpret_cast<const uint8_t*>(kIdDomain), sizeof(kIdDomain) - 1);
  h.absorb(reinterpret_cast<const uint8_t*>(kem.data()), kem.size());
  h.absorb(pk, P.ek_bytes);
  h.squeeze(id, 32);
}

int generate(const std::string& kem, uint64_t seed, int64_t start, int64_t count, uint8_t* ids, uint8_t* keys,
             int threads) {
  MlKemParams P;
  if (!mlkem_params(kem, &P)) return set_error(INSPIRE_ERR_INVALID, "unknown KEM '" + kem + "'");
  uint8_t key[32];
  seed_from_u64(seed, key);
  keccak_consts();
  mlkem_consts();
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  threads = (int)std::min<int64_t>(threads, std::max<int64_t>(1, count / 64));
  std::atomic<int64_t> next{0};
  auto work = [&]() {
    for (;;) {
      const int64_t b = next.fetch_add(256);
      if (b >= count) return;
      const int64_t e = std::min(count, b + 256);
      for (int64_t r = b; r < e; ++r)
        record(P, kem, key, (uint64_t)(start + r), ids + 32 * r, keys + P.ek_bytes * r);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  return INSPIRE_OK;
}

}  // namespace

extern "C" int inspire_keccak(int kind, const uint8_t* in, size_t len, uint8_t* out, size_t out_len) {
  if ((!in && len) || !out) return set_error(INSPIRE_ERR_INVALID, "null argument");
  Sponge s = kind == 0 ? sha3_256() : kind == 1 ? sha3_512() : kind == 2 ? shake128() : shake256();
  if (kind < 0 || kind > 3) return set_error(INSPIRE_ERR_INVALID, "unknown Keccak kind");
  if ((kind == 0 && out_len != 32) || (kind == 1 && out_len != 64))
    return set_error(INSPIRE_ERR_INVALID, "SHA3 output length");
  s.absorb(in, len);
  s.squeeze(out, out_len);
  return INSPIRE_OK;
}

extern "C" int inspire_dataset_record_seed(uint64_t seed, uint64_t index, uint8_t out64[64]) {
  if (!out64) return set_error(INSPIRE_ERR_INVALID, "null argument");
  uint8_t key[32];
  seed_from_u64(seed, key);
  chacha20_block(key, index, out64);
  return INSPIRE_OK;
}

extern "C" int inspire_mlkem_keygen(const char* kem_name, const uint8_t seed64[64], uint8_t* ek, size_t ek_cap) {
  MlKemParams P;
  if (!kem_name || !seed64 || !ek) return set_error(INSPIRE_ERR_INVALID, "null argument");
  if (!mlkem_params(kem_name, &P)) return set_error(INSPIRE_ERR_INVALID, std::string("unknown KEM '") + kem_name + "'");
  if (ek_cap < P.ek_bytes) return set_error(INSPIRE_ERR_CAPACITY, "encapsulation key buffer too small");
  mlkem_ek(P, seed64, ek);
  return INSPIRE_OK;
}

extern "C" int inspire_dataset_generate_records(const char* kem_name, uint64_t seed, int64_t start, int64_t count,
                                                uint8_t* ids, uint8_t* keys, int threads) {
  if (!kem_name || !ids || !keys) return set_error(INSPIRE_ERR_INVALID, "null argument");
  if (count <= 0 || start < 0) return set_error(INSPIRE_ERR_INVALID, "count must be greater than zero");
  return generate(kem_name, seed, start, count, ids, keys, threads);
}

extern "C" int inspire_dataset_generate(const char* path, const char* kem_name, int64_t count, uint64_t seed,
                                        int threads) {
  if (!path || !kem_name) return set_error(INSPIRE_ERR_INVALID, "null argument");
  MlKemParams P;
  const std::string kem = kem_name;
  if (!mlkem_params(kem, &P)) return set_error(INSPIRE_ERR_INVALID, "unknown KEM '" + kem + "'");
  if (count <= 0) return set_error(INSPIRE_ERR_INVALID, "count must be greater than zero");
  if (count > 0xffffffffll) return set_error(INSPIRE_ERR_INVALID, "count does not fit the u32 record count");
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_error(INSPIRE_ERR_INVALID, std::string("cannot create ") + path);
  // writer.rs:38-45: magic, u8 kem-name length + name, u32 LE count, u16 LE key length
  std::vector<uint8_t> hdr(8);
  std::memcpy(hdr.data(), "PQKEYV2\0", 8);
  hdr.push_back((uint8_t)kem.size());
  hdr.insert(hdr.end(), kem.begin(), kem.end());
  for (int b = 0; b < 4; ++b) hdr.push_back((uint8_t)((uint64_t)count >> (8 * b)));
  hdr.push_back((uint8_t)(P.ek_bytes & 0xff));
  hdr.push_back((uint8_t)(P.ek_bytes >> 8));
  bool ok = std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
  // batches of DEFAULT_DATASET_BATCH_SIZE (dataset.rs:12) generated in parallel, written in order
  const int64_t batch = 100000;
  std::vector<uint8_t> ids, keys, rec;
  for (int64_t b = 0; ok && b < count; b += batch) {
    const int64_t n = std::min(batch, count - b);
    ids.resize((size_t)n * 32);
    keys.resize((size_t)n * P.ek_bytes);
    int rc = generate(kem, seed, b, n, ids.data(), keys.data(), threads);
    if (rc) {
      std::fclose(f);
      return rc;
    }
    rec.resize((size_t)n * (32 + P.ek_bytes));
    for (int64_t r = 0; r < n; ++r) {
      std::memcpy(rec.data() + (size_t)r * (32 + P.ek_bytes), ids.data() + 32 * r, 32);
      std::memcpy(rec.data() + (size_t)r * (32 + P.ek_bytes) + 32, keys.data() + P.ek_bytes * r, P.ek_bytes);
    }
    ok = std::fwrite(rec.data(), 1, rec.size(), f) == rec.size();
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return set_error(INSPIRE_ERR_INVALID, std::string("write failed: ") + path);
  return INSPIRE_OK;
}
