// Keyword layer of the server setup (host side): the PQKEYV2 dataset file
// (dataset/dataset.rs:48-142), the cuckoo table over SipHash-2-4
// (pir/keyword/cuckoo.rs:86-353), and setup_keyword_server / public_params
// (dataset/keyword.rs:29-44, server/keyword.rs:29-46, 366-468, 505-531).  The
// table's column-major raw DB is encoded and precomputed on the device
// (inspire_server_encode_db, inspire_server_setup_offline).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/inspire_hip.h"
#include "runtime.hpp"

namespace inspire {
namespace {

// cuckoo.rs:27-29
constexpr uint8_t kCuckooSeed[16] = {0xDE, 0xAD, 0xBE, 0xEF, 0xCA, 0xFE, 0xBA, 0xBE,
                                     0x01, 0x23, 0x45, 0x67, 0x89, 0xAB, 0xCD, 0xEF};
constexpr int kNumHashes = 2;        // cuckoo.rs:56
constexpr int kMaxEvictions = 500;   // cuckoo.rs:57

inline uint64_t rotl(uint64_t x, int b) { return (x << b) | (x >> (64 - b)); }

inline uint64_t load_le64(const uint8_t* p) {
  uint64_t v = 0;
  for (int i = 7; i >= 0; --i) v = (v << 8) | p[i];
  return v;
}

// SipHash-2-4 (Aumasson and Bernstein, "SipHash: a fast short-input PRF",
// 2012), the function of siphasher's SipHasher::new_with_keys + write + finish.
uint64_t siphash24(uint64_t k0, uint64_t k1, const uint8_t* m, size_t len) {
  uint64_t v0 = k0 ^ 0x736f6d6570736575ull, v1 = k1 ^ 0x646f72616e646f6dull;
  uint64_t v2 = k0 ^ 0x6c7967656e657261ull, v3 = k1 ^ 0x7465646279746573ull;
  auto round = [&]() {
    v0 += v1; v1 = rotl(v1, 13); v1 ^= v0; v0 = rotl(v0, 32);
    v2 += v3; v3 = rotl(v3, 16); v3 ^= v2;
    v0 += v3; v3 = rotl(v3, 21); v3 ^= v0;
    v2 += v1; v1 = rotl(v1, 17); v1 ^= v2; v2 = rotl(v2, 32);
  };
  const size_t whole = len & ~(size_t)7;
  for (size_t i = 0; i < whole; i += 8) {
    const uint64_t mi = load_le64(m + i);
    v3 ^= mi;
    round();
    round();
    v0 ^= mi;
  }
  uint64_t b = (uint64_t)(len & 0xff) << 56;
  for (size_t i = 0; i < (len & 7); ++i) b |= (uint64_t)m[whole + i] << (8 * i);
  v3 ^= b;
  round();
  round();
  v0 ^= b;
  v2 ^= 0xff;
  for (int i = 0; i < 4; ++i) round();
  return v0 ^ v1 ^ v2 ^ v3;
}

// CuckooHash::hash (cuckoo.rs:98-124) with the deterministic seed.
int64_t cuckoo_hash(int idx, const uint8_t* key, size_t len, int64_t num_buckets) {
  const uint64_t k0 = load_le64(kCuckooSeed) ^ (uint64_t)idx;
  const uint64_t k1 = load_le64(kCuckooSeed + 8) ^ ((uint64_t)idx << 32);
  return (int64_t)(siphash24(k0, k1, key, len) % (uint64_t)num_buckets);
}

// CuckooTable (cuckoo.rs:141-244): flat buckets, occupancy, stash.  Eviction
// picks the hash index at random; the reference draws it from the OS
// generator (rand::rng), here from a fixed xorshift so that a build is
// repeatable.
struct Cuckoo {
  int64_t buckets;
  size_t ks, vs, es;
  std::vector<uint8_t> data;
  std::vector<uint8_t> occupied;
  std::vector<std::vector<uint8_t>> stash;
  uint64_t rng = 0x9E3779B97F4A7C15ull;

  Cuckoo(int64_t nb, size_t k, size_t v) : buckets(nb), ks(k), vs(v), es(k + v), data((size_t)nb * (k + v), 0),
                                           occupied((size_t)nb, 0) {}

  void put(int64_t pos, const std::vector<uint8_t>& e) {
    std::memcpy(&data[(size_t)pos * es], e.data(), es);
    occupied[(size_t)pos] = 1;
  }

  void insert(const uint8_t* key, const uint8_t* value) {
    std::vector<uint8_t> entry(es, 0);
    std::memcpy(entry.data(), key, ks);
    std::memcpy(entry.data() + ks, value, vs);
    for (int h = 0; h < kNumHashes; ++h) {
      const int64_t p = cuckoo_hash(h, key, ks, buckets);
      if (!occupied[(size_t)p]) return put(p, entry);
    }
    for (int it = 0; it < kMaxEvictions; ++it) {
      rng ^= rng << 13;
      rng ^= rng >> 7;
      rng ^= rng << 17;
      const int h = (int)(rng % kNumHashes);
      const int64_t pos = cuckoo_hash(h, entry.data(), ks, buckets);
      std::vector<uint8_t> evicted(data.begin() + (size_t)pos * es, data.begin() + (size_t)(pos + 1) * es);
      put(pos, entry);
      entry.swap(evicted);
      for (int g = 0; g < kNumHashes; ++g) {
        const int64_t p = cuckoo_hash(g, entry.data(), ks, buckets);
        if (!occupied[(size_t)p]) return put(p, entry);
      }
    }
    stash.push_back(entry);
  }

  // to_raw_db (cuckoo.rs:314-353): bucket row*per + i holds columns
  // [i*ise, (i+1)*ise) of DB row `row`, big-endian u16 words.
  void raw_db(uint16_t* out, int64_t db_rows, int64_t db_cols, int64_t per, int64_t ise) const {
    std::memset(out, 0, sizeof(uint16_t) * (size_t)db_rows * db_cols);
    for (int64_t i = 0; i < per; ++i)
      for (int64_t row = 0; row < db_rows; ++row) {
        const int64_t b = row * per + i;
        if (b >= buckets || !occupied[(size_t)b]) continue;
        const uint8_t* e = &data[(size_t)b * es];
        for (int64_t j = 0; j < ise; ++j)
          out[(i * ise + j) * db_rows + row] = (uint16_t)((e[2 * j] << 8) | e[2 * j + 1]);
      }
  }
};

int load_pqkeyv2(const char* path, std::vector<uint8_t>& ids, std::vector<uint8_t>& keys, uint32_t* count,
                 uint16_t* pk_len) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return set_error(INSPIRE_ERR_INVALID, std::string("cannot open dataset ") + path);
  auto rd = [&](void* p, size_t n) { return std::fread(p, 1, n, f) == n; };
  uint8_t magic[8], kl = 0;
  std::string err;
  if (!rd(magic, 8) || std::memcmp(magic, "PQKEYV2\0", 8) != 0) err = "invalid dataset magic (expected PQKEYV2)";
  char kem[256];
  uint8_t cnt[4], pl[2];
  if (err.empty() && !(rd(&kl, 1) && rd(kem, kl) && rd(cnt, 4) && rd(pl, 2))) err = "truncated dataset header";
  if (err.empty()) {
    *count = (uint32_t)cnt[0] | ((uint32_t)cnt[1] << 8) | ((uint32_t)cnt[2] << 16) | ((uint32_t)cnt[3] << 24);
    *pk_len = (uint16_t)(pl[0] | (pl[1] << 8));
    if (*pk_len == 0) err = "dataset public key length must be non-zero";
  }
  if (err.empty()) {
    ids.resize((size_t)*count * 32);
    keys.resize((size_t)*count * *pk_len);
    for (uint32_t r = 0; r < *count && err.empty(); ++r)
      if (!rd(&ids[(size_t)r * 32], 32) || !rd(&keys[(size_t)r * *pk_len], *pk_len)) err = "truncated dataset record";
  }
  std::fclose(f);
  return err.empty() ? INSPIRE_OK : set_error(INSPIRE_ERR_INVALID, err);
}

}  // namespace
}  // namespace inspire

using namespace inspire;

extern "C" int inspire_cuckoo_positions(const uint8_t* key, size_t key_len, int64_t num_buckets, int64_t* out2) {
  if (!key || !out2 || num_buckets <= 0) return set_error(INSPIRE_ERR_INVALID, "null argument or no buckets");
  for (int h = 0; h < kNumHashes; ++h) out2[h] = cuckoo_hash(h, key, key_len, num_buckets);
  return INSPIRE_OK;
}

extern "C" uint64_t inspire_siphash24(uint64_t k0, uint64_t k1, const uint8_t* msg, size_t len) {
  return siphash24(k0, k1, msg, len);
}

extern "C" int inspire_setup_keyword_records(const uint8_t* keys, uint32_t key_size, const uint8_t* values,
                                             uint32_t value_size, int64_t count, int64_t buckets, int64_t dim0,
                                             inspire_server_t** out, inspire_public_params_t* pp) {
  if (!keys || !values || !out) return set_error(INSPIRE_ERR_INVALID, "null argument");
  *out = nullptr;
  inspire_plan_t plan{};
  int rc = inspire_plan_shape(count, key_size, value_size, buckets, dim0, &plan);
  if (rc) return rc;
  // build_cuckoo_table (server/keyword.rs:33-46)
  Cuckoo table(plan.num_items, key_size, value_size);
  for (int64_t r = 0; r < count; ++r) table.insert(keys + (size_t)r * key_size, values + (size_t)r * value_size);
  const int64_t per = plan.num_items / plan.db_rows;
  const int64_t ise = plan.item_size_bits / 16;
  if (per * ise > plan.db_cols)  // the parameter derivation leaves room for every bucket of a row
    return set_error(INSPIRE_ERR_INVALID, "cuckoo buckets of a row do not fit the DB columns");
  std::vector<uint16_t> raw((size_t)plan.db_rows * plan.db_cols);
  table.raw_db(raw.data(), plan.db_rows, plan.db_cols, per, ise);
  inspire_shape_t shape{plan.db_rows, plan.instances, (int32_t)plan.interpolate_degree, 2};
  inspire_server_t* srv = nullptr;
  rc = inspire_server_create(&shape, &srv);
  if (rc) return rc;
  rc = inspire_server_encode_db(srv, raw.data(), 1);
  if (!rc) rc = inspire_server_setup_offline(srv);
  if (!rc) {
    std::vector<uint8_t> st;
    for (const auto& e : table.stash) st.insert(st.end(), e.begin(), e.end());
    rc = inspire_server_set_stash(srv, st.data(), (uint32_t)table.es, (uint32_t)table.stash.size());
  }
  if (rc) {
    inspire_server_destroy(srv);
    return rc;
  }
  if (pp) {
    // public_params (server/keyword.rs:505-531)
    int64_t occupied = 0;
    for (uint8_t o : table.occupied) occupied += o;
    pp->num_items = plan.num_items;
    pp->item_size_bits = plan.item_size_bits;
    pp->dim0 = plan.dim0;
    std::memcpy(pp->cuckoo_seed, kCuckooSeed, 16);
    pp->num_hashes = kNumHashes;
    pp->key_size = key_size;
    pp->value_size = value_size;
    pp->entry_size = key_size + value_size;
    pp->num_records = occupied + (int64_t)table.stash.size();
  }
  *out = srv;
  return INSPIRE_OK;
}

extern "C" int inspire_setup_keyword_server(const char* pqkeyv2_path, int64_t buckets, int64_t dim0,
                                            inspire_server_t** out, inspire_public_params_t* pp) {
  if (!pqkeyv2_path || !out) return set_error(INSPIRE_ERR_INVALID, "null argument");
  std::vector<uint8_t> ids, keys;
  uint32_t count = 0;
  uint16_t pk_len = 0;
  int rc = load_pqkeyv2(pqkeyv2_path, ids, keys, &count, &pk_len);
  if (rc) return rc;
  // dataset_to_keyword_records (dataset/keyword.rs:12-22): key = the 32-byte
  // public-key id, value = the public key.
  return inspire_setup_keyword_records(ids.data(), 32, keys.data(), pk_len, count, buckets, dim0, out, pp);
}
