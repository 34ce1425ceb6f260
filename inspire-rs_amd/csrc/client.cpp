// Keyword PIR client on the host: KeywordClient (src/pir/client/keyword.rs:45-311).
//
//   setup_from_public_params  keyword.rs:60-121: the scheme shape from the
//                             server's (num_items, item_size_bits, dim0)
//   query                     keyword.rs:198-206: packing keys
//                             (PackingKeys::init_full, packing.rs:1895-1944,
//                             generate_ksk_body 3272-3294) and, per cuckoo
//                             position, the query row (generate_query_b_values,
//                             client/online.rs:226-357) and the RGSW body
//                             (keyword.rs:150-178)
//   serialize_query           serialize_keyword_query (wire.rs:542-615)
//   extract                   keyword.rs:232-310: recover_how_many
//                             (modulus_switch.rs:102-134), decrypt
//                             (decrypt_ct_reg_measured, client/online.rs:118-136)
//
// Ring Z_q[X]/(X^2048 + 1), q = q0 q1.  Polynomials are kept as raw
// coefficients mod q (u64) or as NTT residues per prime ([2][2048] u32, the
// slot order of core/ntt.rs, which the server's kernels share).
//
// Public randomness is the reference's: ChaCha20Rng streams of fixed seeds
// (SEED_0 = all zeros for the query masks, W_SEED / V_SEED for the packing
// masks, [7;32] for the RGSW rows), one u64 per coefficient, reduced mod the
// modulus.  Secret randomness (the key and every noise sample) is a ChaCha20
// keystream under a 32-byte key from getrandom(2), as the reference draws
// rand::random seeds; a nonzero rng_seed replaces that key by a fixed
// expansion of the seed, for repeatable tests.  Noise is the discrete Gaussian
// of width 16.042421 (pir/params.rs:69-96) sampled by its CDF table
// (core/discrete_gaussian.rs:82-133).
//
// The reference's KeywordClient encrypts under Client::init's secret, which is
// the zero polynomial (client/base.rs:370-389); hamming = 0 keeps that.  With
// hamming = h > 0 the secret is ternary with h coefficients 1 and h
// coefficients -1 (gen_ternary_mat, client/base.rs:126-139), which the server
// path serves unchanged.
#include <sys/random.h>

#include <cerrno>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/inspire_hip.h"
#include "host_tables.hpp"
#include "modarith.hpp"
#include "runtime.hpp"

namespace inspire {
// params_rgswpir_given_input_size_and_dim0 (wire.rs:414-477), server.cpp.
int scheme_shape(int64_t num_items, int64_t item_size_bits, int64_t dim0, int64_t* db_rows, int64_t* instances,
                 int64_t* interpolate_degree);
}  // namespace inspire

using namespace inspire;

namespace {

constexpr uint64_t kP = 65535;               // plaintext modulus (wire.rs:421)
constexpr uint64_t kScaleK = kQ / kP;        // floor(q / p)
constexpr int kDgMax = 65;                   // ceil(16.042421 * 4), discrete_gaussian.rs:82
constexpr double kNoiseWidth = 16.042421;    // pir/params.rs:69-96
constexpr int kRowBits = 28;                 // log2_ceil of q0 and q1 (wire.rs:120-125)

using Ntt = std::vector<uint32_t>;  // [2][kN] residues

// rescale (core/arith.rs:436-451): round(centered(a) * out / in), mod out.
uint64_t rescale(uint64_t a, uint64_t in_mod, uint64_t out_mod) {
  const int64_t im = (int64_t)in_mod;
  int64_t v = (int64_t)(a % in_mod);
  if (v >= im / 2) v -= im;
  const int64_t sign = v >= 0 ? 1 : -1;
  const __int128 om = (__int128)out_mod;
  __int128 r = ((__int128)v * (__int128)out_mod + (__int128)(sign * (im / 2))) / (__int128)in_mod;
  r = (r + (__int128)((in_mod / out_mod) * out_mod) + 2 * om) % om;
  return (uint64_t)((r + om) % om);
}

// read_arbitrary_bits / write_arbitrary_bits (util.rs:289-321): LSB-first in
// little-endian 64-bit words.
uint64_t get_bits(const uint8_t* d, size_t len, size_t off, int nbits) {
  uint64_t v = 0;
  for (int i = 0; i < nbits; ++i) {
    const size_t b = off + i;
    if (b / 8 < len && ((d[b / 8] >> (b % 8)) & 1)) v |= 1ull << i;
  }
  return v;
}
void put_bits(uint8_t* d, size_t off, uint64_t v, int nbits) {
  size_t byte = off / 8;
  int bit = (int)(off % 8);
  while (nbits > 0) {
    const int take = std::min(8 - bit, nbits);
    const uint8_t m = (uint8_t)(((1u << take) - 1) << bit);
    d[byte] = (uint8_t)((d[byte] & ~m) | (((uint8_t)v << bit) & m));
    v >>= take;
    nbits -= take;
    bit = 0;
    ++byte;
  }
}

// Secret randomness: ChaCha20 keystream, two words per u64 (low first).
struct SecretRng {
  uint8_t key[32];
  uint64_t counter = 0;
  uint32_t buf[16];
  int pos = 16;
  uint64_t next() {
    if (pos >= 16) {
      chacha20_block_words(key, counter++, buf);
      pos = 0;
    }
    const uint64_t v = (uint64_t)buf[pos] | ((uint64_t)buf[pos + 1] << 32);
    pos += 2;
    return v;
  }
};

}  // namespace

struct inspire_client {
  inspire_public_params_t pp{};
  int64_t db_rows = 0, instances = 0, d = 0, c = 0, per = 0, ise = 0;
  std::vector<uint32_t> tables;  // build_ntt_tables
  std::vector<uint64_t> sk;      // raw secret, [kN] mod q
  Ntt sk_ntt;
  uint64_t cdf[2 * kDgMax + 1];
  SecretRng rng;
};

namespace {

Ntt to_ntt(const inspire_client* c, const uint64_t* raw) {
  Ntt out(2 * kN);
  for (int p = 0; p < 2; ++p) {
    for (int z = 0; z < kN; ++z) out[p * kN + z] = (uint32_t)(raw[z] % prime(p));
    host_ntt_forward(out.data() + p * kN, p, c->tables.data());
  }
  return out;
}

std::vector<uint64_t> from_ntt(const inspire_client* c, Ntt v) {
  host_ntt_inverse(v.data(), 0, c->tables.data());
  host_ntt_inverse(v.data() + kN, 1, c->tables.data());
  std::vector<uint64_t> raw(kN);
  for (int z = 0; z < kN; ++z) raw[z] = crt_compose(v[z], v[kN + z]);
  return raw;
}

// out (+)= a * b per prime.
void ntt_mul(Ntt& out, const Ntt& a, const Ntt& b, bool acc) {
  for (int p = 0; p < 2; ++p)
    for (int z = 0; z < kN; ++z) {
      const int i = p * kN + z;
      const uint64_t v = (uint64_t)a[i] * b[i] % prime(p);
      out[i] = (uint32_t)(acc ? (out[i] + v) % prime(p) : v);
    }
}

// One discrete Gaussian sample mod m: the smallest i with u <= cdf[i].
uint64_t noise_sample(inspire_client* c, uint64_t m) {
  const uint64_t u = c->rng.next();
  int i = 0;
  while (i < 2 * kDgMax && u > c->cdf[i]) ++i;
  const int64_t v = (int64_t)i - kDgMax;
  return v < 0 ? (uint64_t)((int64_t)m + v) : (uint64_t)v;
}

std::vector<uint64_t> noise_poly(inspire_client* c) {
  std::vector<uint64_t> e(kN);
  for (int z = 0; z < kN; ++z) e[z] = noise_sample(c, kQ);
  return e;
}

// PolyMatrixNTT::random_rng (poly.rs:418-442) from a ChaCha20Rng seed: per
// polynomial, prime 0 then prime 1, one u64 per residue mod q_p.
void random_ntt(const uint8_t seed[32], int npolys, std::vector<uint32_t>* out) {
  std::vector<uint64_t> w((size_t)npolys * 2 * kN);
  chacha20_u64(seed, w.data(), w.size());
  out->resize(w.size());
  for (size_t i = 0; i < w.size(); ++i) (*out)[i] = (uint32_t)(w[i] % prime((int)((i / kN) & 1)));
}

// automorph (poly.rs:632-653): X^i -> X^(i t mod 2n), negated past n.
std::vector<uint64_t> automorph(const std::vector<uint64_t>& a, uint32_t t) {
  std::vector<uint64_t> r(kN, 0);
  for (int i = 0; i < kN; ++i) {
    const uint32_t e = (uint32_t)(((uint64_t)i * t) % (2 * kN));
    const uint64_t v = a[i] % kQ;
    if (e < (uint32_t)kN)
      r[e] = v;
    else
      r[e - kN] = v ? kQ - v : 0;
  }
  return r;
}

// generate_ksk_body (packing.rs:3272-3294): for k < 3,
// body_k = NTT(tau_gen(s)) 2^(19k) - NTT(s) mask_k + NTT(e_k), condensed.
void ksk_body(inspire_client* c, uint32_t gen, const std::vector<uint32_t>& mask, uint64_t* out) {
  const Ntt tau = to_ntt(c, automorph(c->sk, gen).data());
  for (int k = 0; k < kTExp; ++k) {
    const Ntt e = to_ntt(c, noise_poly(c).data());
    for (int p = 0; p < 2; ++p) {
      const uint64_t m = prime(p), g = (1ull << (kGadgetBits * k)) % m;
      for (int z = 0; z < kN; ++z) {
        const int i = p * kN + z;
        const uint64_t neg_mask = (m - mask[((size_t)k * 2 + p) * kN + z]) % m;
        uint64_t v = (uint64_t)tau[i] * g % m;
        v = (v + (uint64_t)c->sk_ntt[i] * neg_mask) % m;
        const uint64_t r = (v + e[i]) % m;
        if (p == 0)
          out[(size_t)k * kN + z] = r;
        else
          out[(size_t)k * kN + z] |= r << 32;
      }
    }
  }
}

// The query row of a target row: db_rows / n RLWE b-values under the SEED_0
// masks (one mask polynomial per RLWE, in order), b = s a + e, plus floor(q/p)
// at the target coefficient; condensed.
void query_row(inspire_client* c, int64_t target_row, uint64_t* out) {
  const int64_t nrlwe = c->db_rows / kN;
  uint8_t seed0[32];
  std::memset(seed0, 0, sizeof(seed0));
  std::vector<uint64_t> a_all((size_t)nrlwe * kN);
  chacha20_u64(seed0, a_all.data(), a_all.size());
  for (int64_t k = 0; k < nrlwe; ++k) {
    std::vector<uint64_t> a(kN);
    for (int z = 0; z < kN; ++z) a[z] = a_all[(size_t)k * kN + z] % kQ;
    Ntt sa(2 * kN);
    ntt_mul(sa, c->sk_ntt, to_ntt(c, a.data()), false);
    const std::vector<uint64_t> b = from_ntt(c, sa);
    for (int z = 0; z < kN; ++z) {
      uint64_t v = (b[z] + noise_sample(c, kQ)) % kQ;
      if (k * kN + z == target_row) v = (v + kScaleK) % kQ;
      out[(size_t)k * kN + z] = (v % kQ0) | ((v % kQ1) << 32);
    }
  }
}

// ct_gsw_body for sigma = +-2^(19 j) X^(exponent mod n): column 2j+1 is the
// body of Enc(sigma), column 2j the body of Enc(s sigma), both under the RGSW
// mask a of seed [7;32]: body = NTT(s) NTT(a) + NTT(e) + msg.  out: [6][2][kN].
void rgsw_body(inspire_client* c, uint32_t exponent, uint32_t* out) {
  uint8_t seed[32];
  std::memset(seed, 7, sizeof(seed));
  std::vector<uint64_t> a(kN);
  chacha20_u64(seed, a.data(), kN);
  for (auto& v : a) v %= kQ;
  Ntt sa(2 * kN);
  ntt_mul(sa, c->sk_ntt, to_ntt(c, a.data()), false);
  exponent %= 2 * kN;
  for (int j = 0; j < kTExp; ++j) {
    std::vector<uint64_t> sig(kN, 0);
    const uint64_t g = 1ull << (kGadgetBits * j);
    sig[exponent % kN] = exponent < (uint32_t)kN ? g : kQ - g;
    const Ntt sig_ntt = to_ntt(c, sig.data());
    Ntt prod(2 * kN);
    ntt_mul(prod, c->sk_ntt, sig_ntt, false);
    for (int half = 0; half < 2; ++half) {
      const int col = 2 * j + (half ? 0 : 1);
      const Ntt& msg = half ? prod : sig_ntt;
      const Ntt e = to_ntt(c, noise_poly(c).data());
      for (int p = 0; p < 2; ++p)
        for (int z = 0; z < kN; ++z) {
          const int i = p * kN + z;
          out[((size_t)col * 2 + p) * kN + z] = (uint32_t)(((uint64_t)sa[i] + e[i] + msg[i]) % prime(p));
        }
    }
  }
}

size_t query_bytes(const inspire_client* c, int nq) {
  const size_t values = (size_t)kTExp * kN * 2 + (size_t)nq * (c->db_rows + 2 * kTExp * kN);
  return (8 + values * 2 * kRowBits + 7) / 8;
}

uint64_t pack_condensed(uint64_t v) { return (v & ((1ull << kRowBits) - 1)) | ((v >> 32) << kRowBits); }

}  // namespace

extern "C" int inspire_client_new(const inspire_public_params_t* pp, int32_t hamming, uint64_t rng_seed,
                                  inspire_client_t** out) {
  if (!pp || !out) return set_error(INSPIRE_ERR_INVALID, "null argument");
  *out = nullptr;
  if (pp->key_size <= 0) return set_error(INSPIRE_ERR_INVALID, "server keyword size must be non-zero");
  if (pp->value_size <= 0) return set_error(INSPIRE_ERR_INVALID, "server value size must be non-zero");
  if (pp->entry_size < pp->key_size + pp->value_size)
    return set_error(INSPIRE_ERR_INVALID, "server entry size " + std::to_string(pp->entry_size) +
                                              " is smaller than key_size " + std::to_string(pp->key_size) +
                                              " + value_size " + std::to_string(pp->value_size));
  if (pp->num_hashes < 1 || pp->num_hashes > 8) return set_error(INSPIRE_ERR_INVALID, "num_hashes must be in [1, 8]");
  if (pp->item_size_bits <= 0 || pp->item_size_bits % 16 != 0)
    return set_error(INSPIRE_ERR_INVALID, "item size must be a positive multiple of 16 bits");
  if (hamming < 0 || hamming > kN / 2) return set_error(INSPIRE_ERR_INVALID, "hamming weight must be in [0, 1024]");
  int64_t db_rows = 0, instances = 0, d = 0;
  int rc = scheme_shape(pp->num_items, pp->item_size_bits, pp->dim0, &db_rows, &instances, &d);
  if (rc) return rc;
  if (pp->num_items < db_rows)
    return set_error(INSPIRE_ERR_INVALID, "server bucket count " + std::to_string(pp->num_items) +
                                              " is too small for the PIR scheme (needs at least " +
                                              std::to_string(db_rows) + " records)");
  auto* c = new inspire_client();
  c->pp = *pp;
  c->db_rows = db_rows;
  c->instances = instances;
  c->d = d;
  c->c = instances / d;
  c->per = pp->num_items / db_rows;
  c->ise = pp->item_size_bits / 16;
  c->tables.resize(2 * 4 * kN);
  build_ntt_tables(c->tables.data());
  if (rng_seed == 0) {
    size_t got = 0;
    while (got < sizeof(c->rng.key)) {
      const ssize_t n = getrandom(c->rng.key + got, sizeof(c->rng.key) - got, 0);
      if (n < 0 && errno == EINTR) continue;
      if (n <= 0) {
        delete c;
        return set_error(INSPIRE_ERR_INVALID, "getrandom failed");
      }
      got += (size_t)n;
    }
  } else {
    for (int i = 0; i < 4; ++i) {
      const uint64_t w = mix64(rng_seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1));
      std::memcpy(c->rng.key + 8 * i, &w, 8);
    }
  }
  // discrete_gaussian.rs:82-109: CDF of p(i) ~ exp(-pi i^2 / w^2) on [-65, 65], times u64::MAX.
  {
    double p[2 * kDgMax + 1], total = 0;
    for (int i = -kDgMax; i <= kDgMax; ++i) {
      p[i + kDgMax] = std::exp(-M_PI * (double)i * (double)i / (kNoiseWidth * kNoiseWidth));
      total += p[i + kDgMax];
    }
    double cum = 0;
    for (int i = 0; i < 2 * kDgMax + 1; ++i) {
      cum += p[i] / total;
      const double v = std::round(cum * 18446744073709551615.0);
      c->cdf[i] = v >= 18446744073709551615.0 ? UINT64_MAX : (uint64_t)v;
    }
  }
  // gen_ternary_mat (client/base.rs:126-139): h ones, h minus ones, shuffled.
  c->sk.assign(kN, 0);
  for (int i = 0; i < hamming; ++i) c->sk[i] = 1;
  for (int i = hamming; i < 2 * hamming; ++i) c->sk[i] = kQ - 1;
  for (int i = kN - 1; i > 0; --i) {
    const uint64_t j = c->rng.next() % (uint64_t)(i + 1);
    std::swap(c->sk[i], c->sk[j]);
  }
  c->sk_ntt = to_ntt(c, c->sk.data());
  *out = c;
  return INSPIRE_OK;
}

extern "C" void inspire_client_free(inspire_client_t* c) { delete c; }

extern "C" int inspire_client_secret(const inspire_client_t* c, uint64_t* sk) {
  if (!c || !sk) return set_error(INSPIRE_ERR_INVALID, "null argument");
  std::memcpy(sk, c->sk.data(), sizeof(uint64_t) * kN);
  return INSPIRE_OK;
}

extern "C" size_t inspire_client_query_bytes(const inspire_client_t* c) {
  return c ? query_bytes(c, c->pp.num_hashes) : 0;
}

extern "C" int inspire_client_positions(const inspire_client_t* c, const uint8_t* key, size_t key_len,
                                        int64_t* positions) {
  if (!c || !key || !positions) return set_error(INSPIRE_ERR_INVALID, "null argument");
  if ((int64_t)key_len != c->pp.key_size)
    return set_error(INSPIRE_ERR_INVALID, "expected " + std::to_string(c->pp.key_size) + " keyword bytes, got " +
                                              std::to_string(key_len));
  // CuckooHash::new_from_seed(cuckoo_seed, num_hashes, num_items).all_positions
  // (cuckoo.rs:98-132): SipHash-2-4 keyed by the seed's two words, hash index
  // i folded into k0 and (i << 32) into k1.
  uint64_t s0 = 0, s1 = 0;
  std::memcpy(&s0, c->pp.cuckoo_seed, 8);
  std::memcpy(&s1, c->pp.cuckoo_seed + 8, 8);
  for (int h = 0; h < c->pp.num_hashes; ++h)
    positions[h] = (int64_t)(inspire_siphash24(s0 ^ (uint64_t)h, s1 ^ ((uint64_t)h << 32), key, key_len) %
                             (uint64_t)c->pp.num_items);
  return INSPIRE_OK;
}

extern "C" int inspire_client_query(inspire_client_t* c, const uint8_t* key, size_t key_len, int64_t* positions,
                                    uint8_t* out, size_t cap, size_t* len) {
  if (!c || !key || !positions || !out || !len) return set_error(INSPIRE_ERR_INVALID, "null argument");
  int rc = inspire_client_positions(c, key, key_len, positions);
  if (rc) return rc;
  const int nq = c->pp.num_hashes;
  const size_t need = query_bytes(c, nq);
  *len = need;
  if (cap < need) return set_error(INSPIRE_ERR_CAPACITY, "query buffer too small");
  // PackingKeys::init_full: y with generator 5 and mask W_SEED, z with 2n - 1
  // and mask V_SEED (pir/scheme.rs:46-51).
  uint8_t ws[32], vs[32];
  std::memset(ws, 7, 32);
  std::memset(vs, 7, 32);
  vs[0] = vs[1] = 8;
  std::vector<uint32_t> wmask, vmask;
  random_ntt(ws, kTExp, &wmask);
  random_ntt(vs, kTExp, &vmask);
  std::vector<uint64_t> y(kTExp * kN), z(kTExp * kN);
  ksk_body(c, 5, wmask, y.data());
  ksk_body(c, 2 * kN - 1, vmask, z.data());
  std::memset(out, 0, need);
  out[0] = (uint8_t)nq;
  size_t bo = 8;
  const int tb = 2 * kRowBits;
  for (int i = 0; i < kTExp; ++i)
    for (int k = 0; k < kN; ++k) {
      put_bits(out, bo, pack_condensed(y[(size_t)i * kN + k]), tb);
      bo += tb;
      put_bits(out, bo, pack_condensed(z[(size_t)i * kN + k]), tb);
      bo += tb;
    }
  std::vector<uint64_t> row(c->db_rows);
  std::vector<uint32_t> body(6 * 2 * kN);
  for (int h = 0; h < nq; ++h) {
    // build_query_for_index (keyword.rs:150-153)
    const int64_t bucket = positions[h];
    const int64_t target_row = bucket / c->per;
    const int64_t target_col = (bucket % c->per) * c->ise;
    const int64_t sub_col = (target_col % (c->d * kN)) / kN;
    const uint32_t exponent = (uint32_t)(((2 * kN * sub_col) / c->d) % (2 * kN));
    query_row(c, target_row, row.data());
    rgsw_body(c, exponent, body.data());
    for (int64_t r = 0; r < c->db_rows; ++r) {
      put_bits(out, bo, pack_condensed(row[r]), tb);
      bo += tb;
    }
    for (int col = 0; col < 6; ++col)
      for (int zz = 0; zz < kN; ++zz) {
        const uint64_t lo = body[((size_t)col * 2 + 0) * kN + zz], hi = body[((size_t)col * 2 + 1) * kN + zz];
        put_bits(out, bo, lo | (hi << kRowBits), tb);
        bo += tb;
      }
  }
  return INSPIRE_OK;
}

extern "C" int inspire_client_extract(const inspire_client_t* c, const uint8_t* key, size_t key_len,
                                      const int64_t* positions, const uint8_t* resp, size_t resp_len, uint8_t* value,
                                      size_t cap, size_t* value_len, int32_t* found) {
  if (!c || !key || !positions || !resp || !value_len || !found) return set_error(INSPIRE_ERR_INVALID, "null argument");
  *found = 0;
  *value_len = 0;
  if ((int64_t)key_len != c->pp.key_size) return INSPIRE_OK;  // extract returns None
  const size_t ks = (size_t)c->pp.key_size, vs = (size_t)c->pp.value_size, es = (size_t)c->pp.entry_size;
  // deserialize_keyword_response (wire.rs:728-760)
  size_t o = 0;
  auto u32_at = [&](uint32_t* v) {
    if (o + 4 > resp_len) return false;
    std::memcpy(v, resp + o, 4);
    o += 4;
    return true;
  };
  uint32_t nresp = 0;
  if (!u32_at(&nresp)) return set_error(INSPIRE_ERR_INVALID, "malformed keyword response");
  std::vector<std::pair<const uint8_t*, uint32_t>> blobs, stash;
  for (uint32_t i = 0; i < nresp; ++i) {
    uint32_t l = 0;
    if (!u32_at(&l) || o + l > resp_len) return set_error(INSPIRE_ERR_INVALID, "malformed keyword response");
    blobs.push_back({resp + o, l});
    o += l;
  }
  uint32_t nst = 0;
  if (!u32_at(&nst)) return set_error(INSPIRE_ERR_INVALID, "malformed keyword response");
  for (uint32_t i = 0; i < nst; ++i) {
    uint32_t l = 0;
    if (!u32_at(&l) || o + l > resp_len) return set_error(INSPIRE_ERR_INVALID, "malformed keyword response");
    stash.push_back({resp + o, l});
    o += l;
  }
  auto hit = [&](const uint8_t* entry, size_t n) {
    if (n < es || std::memcmp(entry, key, ks) != 0) return false;
    *value_len = vs;
    *found = 1;
    if (value && cap >= vs) std::memcpy(value, entry + ks, vs);
    return true;
  };
  const size_t chain_bytes = (size_t)kRespBytes;
  for (size_t i = 0; i < blobs.size() && i < (size_t)c->pp.num_hashes; ++i) {
    const uint8_t* blob = blobs[i].first;
    const size_t blen = blobs[i].second;
    if (blen != (size_t)c->c * chain_bytes) return set_error(INSPIRE_ERR_INVALID, "response blob size mismatch");
    // decrypt_response_values: every chain, concatenated
    std::vector<uint64_t> dec((size_t)c->c * kN);
    for (int64_t w = 0; w < c->c; ++w) {
      const uint8_t* chain = blob + (size_t)w * chain_bytes;
      std::vector<uint64_t> r0(kN), r1(kN);
      size_t bo = 0;
      for (int zz = 0; zz < kN; ++zz, bo += 28) r0[zz] = rescale(get_bits(chain, chain_bytes, bo, 28), kQPrime2, kQ);
      for (int zz = 0; zz < kN; ++zz, bo += 20) r1[zz] = rescale(get_bits(chain, chain_bytes, bo, 20), kQPrime1, kQ);
      Ntt m = to_ntt(c, r1.data());
      ntt_mul(m, c->sk_ntt, to_ntt(c, r0.data()), true);
      const std::vector<uint64_t> raw = from_ntt(c, m);
      for (int zz = 0; zz < kN; ++zz) dec[(size_t)w * kN + zz] = rescale(raw[zz], kQ, kP);
    }
    const int64_t bucket = positions[i];
    const int64_t target_col = (bucket % c->per) * c->ise;
    const int64_t off = (target_col / kN / c->d) * kN + target_col % kN;
    if (off + c->ise > (int64_t)dec.size()) continue;  // the item runs past the decrypted values
    std::vector<uint8_t> entry((size_t)c->ise * 2);
    for (int64_t k = 0; k < c->ise; ++k) {
      const uint16_t v = (uint16_t)dec[(size_t)(off + k)];
      entry[2 * k] = (uint8_t)(v >> 8);
      entry[2 * k + 1] = (uint8_t)v;
    }
    if (hit(entry.data(), entry.size())) break;
  }
  if (!*found)
    for (auto& e : stash)
      if (hit(e.first, e.second)) break;
  if (*found && value && cap < vs) return set_error(INSPIRE_ERR_CAPACITY, "value buffer too small");
  return INSPIRE_OK;
}
