// Handle-based answer path: KeywordServer::respond (src/pir/server/keyword.rs:534-628)
// with every stage on the GPU, plus the setup plan (keyword.rs:91-233) and the
// stage entry points declared in include/inspire_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/inspire_hip.h"
#include "host_tables.hpp"
#include "kernels.hpp"
#include "modarith.hpp"
#include "runtime.hpp"

using namespace inspire;

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------

namespace inspire {
thread_local std::string g_last_error;
int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
int hip_error(hipError_t e, const char* what) {
  return set_error(INSPIRE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace inspire

extern "C" const char* inspire_last_error(void) { return g_last_error.c_str(); }

#define HIP_TRY(call, what)                         \
  do {                                              \
    hipError_t e_ = (call);                         \
    if (e_ != hipSuccess) return hip_error(e_, what); \
  } while (0)

// ---------------------------------------------------------------------------
// Per-device constants: NTT twiddles and NTT(-a) for the RGSW row 0.
// ---------------------------------------------------------------------------

namespace inspire {

static std::mutex g_const_mu;
static DeviceConstants g_consts[64];

const DeviceConstants* device_constants() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(g_const_mu);
  DeviceConstants& dc = g_consts[dev];
  if (dc.tw) return &dc;
  std::vector<uint32_t> tw(2 * 4 * kN), na(2 * kN), fo(kFoldTableWords);
  std::vector<uint16_t> ts(kN);
  build_ntt_tables(tw.data());
  build_rgsw_neg_a(na.data(), tw.data());
  build_fold_tables(fo.data(), tw.data(), na.data());
  build_slot_dlog(ts.data());
  uint32_t *dtw = nullptr, *dna = nullptr, *dfo = nullptr;
  uint16_t* dts = nullptr;
  auto release = [&]() {
    if (dtw) (void)hipFree(dtw);
    if (dna) (void)hipFree(dna);
    if (dfo) (void)hipFree(dfo);
    if (dts) (void)hipFree(dts);
  };
  if (hipMalloc(&dtw, tw.size() * 4) != hipSuccess || hipMalloc(&dna, na.size() * 4) != hipSuccess ||
      hipMalloc(&dfo, fo.size() * 4) != hipSuccess || hipMalloc(&dts, ts.size() * 2) != hipSuccess) {
    release();
    return nullptr;
  }
  if (hipMemcpy(dtw, tw.data(), tw.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dna, na.data(), na.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dfo, fo.data(), fo.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dts, ts.data(), ts.size() * 2, hipMemcpyHostToDevice) != hipSuccess) {
    release();
    return nullptr;
  }
  dc.slot_dlog = dts;
  dc.neg_a = dna;
  dc.fold = dfo;
  dc.tw = dtw;
  return &dc;
}

}  // namespace inspire

// ---------------------------------------------------------------------------
// Setup plan (keyword.rs:145-233, wire.rs:386-477)
// ---------------------------------------------------------------------------

namespace {

// wire.rs:386-412
int64_t max_interpolate_degree(double modulus, double d0, double p, double t_exp, double t_rgsw,
                               double poly_len, bool* ok) {
  const double sigma_x = 6.4;
  const double modulus_len = std::log2(modulus);
  const double z1 = std::pow(2.0, std::ceil(modulus_len / t_exp));
  const double z2 = std::pow(2.0, std::ceil(modulus_len / t_rgsw));
  const double s2 = sigma_x * sigma_x;
  const double term1 = d0 * (p * p) * s2;
  const double term2 = t_exp * (poly_len * poly_len) * (z1 * z1) * s2 / 4.0;
  const double term3 = t_rgsw * poly_len * (z2 * z2) * s2 / 2.0;
  const double total = std::log2(std::sqrt(term1 + term2 + term3));
  const double ub = std::log2(modulus / (2. * 2. * p)) - std::log2(std::sqrt(2.0 * 41.0 * std::log(2.0)));
  const double maxlog = 2. * (ub - total);
  *ok = maxlog >= 0.;
  const int64_t m = (int64_t)std::floor(std::pow(2.0, maxlog));
  return m < (int64_t)poly_len ? m : (int64_t)poly_len;
}

int64_t ceil_div_f(double a, double b) { return (int64_t)std::ceil(a / b); }

struct RgswParams {
  int64_t db_rows, instances, interpolate_degree;
};

// wire.rs:414-477 params_rgswpir_given_input_size_and_dim0
int rgsw_params(int64_t num_items, int64_t item_size_bits, int64_t dim0, RgswParams* out) {
  const int64_t poly_len = 2048, log_p = 16;
  const double modulus = (double)kQ;
  bool ok = true;
  const int64_t max_interp = max_interpolate_degree(modulus, (double)dim0, 65535.0, 3.0, 3.0, 2048.0, &ok);
  if (!ok || max_interp <= 1) return set_error(INSPIRE_ERR_INVALID, "dim0 too large for the noise bound");
  const double bits_per_poly = (double)(poly_len * log_p);
  int64_t factor = (int64_t)std::floor(bits_per_poly / (double)item_size_bits);
  if (factor == 0) factor = 1;
  const int64_t n_items = ceil_div_f((double)num_items, (double)factor);
  const int64_t item_bits = item_size_bits * factor;
  const int64_t padded_bits = ceil_div_f((double)item_bits, bits_per_poly) * (int64_t)bits_per_poly;
  const int64_t padded_pts = ceil_div_f((double)padded_bits, (double)log_p);
  const int64_t dim1_lb = padded_pts * ceil_div_f((double)n_items, (double)dim0);
  int64_t cur = padded_pts, interp = 1;
  while (cur < dim1_lb) {
    interp *= 2;
    cur *= 2;
    if (2 * interp > max_interp) break;
  }
  const int64_t new_pts = ceil_div_f((double)dim1_lb, (double)cur) * cur;
  const int64_t db_rows_in = dim0 >= 8 ? dim0 : 8;
  const int64_t db_cols_poly = ceil_div_f((double)new_pts, (double)poly_len);
  int64_t np2 = 1;
  while (np2 < db_rows_in) np2 <<= 1;
  out->db_rows = np2;  // 1 << (nu_1 + 11) with nu_1 = log2(next_pow2(db_rows)) - 11
  if (np2 < 2048) return set_error(INSPIRE_ERR_INVALID, "dim0 must be at least 2048");
  out->instances = db_cols_poly;
  out->interpolate_degree = interp;
  return INSPIRE_OK;
}

}  // namespace

namespace inspire {
// The scheme shape of (num_items, item_size_bits, dim0), for the client.
int scheme_shape(int64_t num_items, int64_t item_size_bits, int64_t dim0, int64_t* db_rows, int64_t* instances,
                 int64_t* interpolate_degree) {
  RgswParams rp;
  const int rc = rgsw_params(num_items, item_size_bits, dim0, &rp);
  if (rc) return rc;
  *db_rows = rp.db_rows;
  *instances = rp.instances;
  *interpolate_degree = rp.interpolate_degree;
  return INSPIRE_OK;
}
}  // namespace inspire

extern "C" int inspire_plan_shape(int64_t record_count, int64_t key_size, int64_t value_size, int64_t buckets,
                                  int64_t dim0, inspire_plan_t* out) {
  if (!out) return set_error(INSPIRE_ERR_INVALID, "null plan");
  if (record_count <= 0)
    return set_error(INSPIRE_ERR_INVALID, "keyword record set must contain at least one record");
  if (key_size <= 0) return set_error(INSPIRE_ERR_INVALID, "keyword key size must be non-zero");
  if (value_size <= 0) return set_error(INSPIRE_ERR_INVALID, "keyword value size must be non-zero");
  int64_t num_items = buckets;
  if (num_items <= 0) {  // default_bucket_count (keyword.rs:29-31)
    int64_t m = (record_count > 512 ? record_count : 512) * 2, p = 1;
    while (p < m) p <<= 1;
    num_items = p;
  }
  if (num_items < record_count)
    return set_error(INSPIRE_ERR_INVALID, "bucket count " + std::to_string(num_items) +
                                              " must be >= record count " + std::to_string(record_count));
  const int64_t entry_size = key_size + value_size;
  const int64_t item_size_bits = entry_size * 8;
  if (item_size_bits % 16 != 0) return set_error(INSPIRE_ERR_INVALID, "entry size must be a multiple of 16 bits");
  RgswParams rp;
  if (dim0 <= 0) {
    const int64_t def = num_items >= 256000000 ? 32768 : 2048;
    int rc = rgsw_params(num_items, item_size_bits, def, &rp);
    if (rc) return rc;
    dim0 = def >= 8 ? def : 8;  // the derived db_rows of the default dim0
  }
  int rc = rgsw_params(num_items, item_size_bits, dim0, &rp);
  if (rc) return rc;
  const int64_t db_rows = rp.db_rows;
  const int64_t db_cols = rp.instances * 2048;
  if (num_items < db_rows) {
    const int64_t min_records = db_rows / 2 + 1;
    return set_error(INSPIRE_ERR_INVALID, "bucket count " + std::to_string(num_items) +
                                              " is too small for the PIR scheme (needs at least " +
                                              std::to_string(db_rows) + "); use at least " +
                                              std::to_string(min_records) + " records, or pass --buckets " +
                                              std::to_string(db_rows));
  }
  out->num_items = num_items;
  out->item_size_bits = item_size_bits;
  out->dim0 = dim0;
  out->db_rows = db_rows;
  out->db_cols = db_cols;
  out->instances = rp.instances;
  out->interpolate_degree = rp.interpolate_degree;
  out->c = rp.instances / rp.interpolate_degree;
  // KeywordServerGpuMemoryEstimate::from_plan (keyword.rs:104-142)
  const uint64_t poly_len = 2048, crt = 2, t = 3, nphalf = 1024;
  const uint64_t coeff_count = crt * poly_len;
  out->db_bytes = (uint64_t)db_rows * db_cols * 2;
  out->hint_bytes = (uint64_t)db_cols * poly_len * 8;
  out->prep_pack_bytes = (uint64_t)db_cols * coeff_count * 8;
  out->r_buffers_bytes = 2 * (uint64_t)rp.instances * nphalf * coeff_count * 8;
  out->bold_t_bytes = 2 * (uint64_t)rp.instances * (nphalf - 1) * t * poly_len * 8;
  out->peak_bytes = out->db_bytes + out->prep_pack_bytes + out->r_buffers_bytes + out->bold_t_bytes;
  return INSPIRE_OK;
}

// ---------------------------------------------------------------------------
// Server
// ---------------------------------------------------------------------------

// Per-request device buffers.  A server owns several so that consecutive
// requests overlap on the device (one request's fold with the next one's
// packing); a slot is reused only after its previous request has finished
// (its `done` event), whatever stream that request ran on.
struct Slot {
  uint8_t* qbytes = nullptr;
  uint64_t *y_body = nullptr, *z_body = nullptr, *rows = nullptr, *P = nullptr, *cvec = nullptr;
  uint32_t *bodies = nullptr, *S = nullptr;
  uint8_t* out = nullptr;
  void* gemv_scratch = nullptr;  // MFMA query fragments
  uint8_t* keys = nullptr;       // MFMA packing: client-key fragments
  hipEvent_t done = nullptr;
};

// Buffers for batched responds (inspire_respond_batch_device).
struct BatchGraph {
  int count = 0, nq = 0, variants = 0;
  const void* queries = nullptr;
  size_t q_stride = 0;
  const void* out = nullptr;
  size_t out_stride = 0;
  hipStream_t stream = nullptr;
  hipGraphExec_t exec = nullptr;  // nullptr: seen once, not captured yet
  bool same(const BatchGraph& o) const {
    return count == o.count && nq == o.nq && variants == o.variants && queries == o.queries &&
           q_stride == o.q_stride && out == o.out && out_stride == o.out_stride && stream == o.stream;
  }
};

struct Batch {
  int max = 0;
  uint64_t *y_body = nullptr, *z_body = nullptr, *rows = nullptr, *P = nullptr, *cvec = nullptr;
  uint32_t *bodies = nullptr, *S = nullptr;
  uint8_t* keys = nullptr;
  void* gemv_scratch = nullptr;
  hipEvent_t done = nullptr;
};

struct inspire_server {
  inspire_shape_t shape{};
  int64_t db_cols = 0;
  int c = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  const DeviceConstants* consts = nullptr;
  // resident state
  uint16_t* db = nullptr;
  void* dbfrag = nullptr;  // fragment-ordered DB for the MFMA GEMVs (nullptr: column-major kernels)
  uint64_t *bt = nullptr, *btb = nullptr, *bth = nullptr, *a_hat = nullptr;
  uint32_t* ahat_ntt = nullptr;  // the fold's row-0 table of a_hat (launch_ahat_ntt)
  bool have_db = false, have_precomp = false;
  // Lean servers keep only the fragment-ordered copy of bold_t (MFMA packing
  // only); the condensed copy exists only while a setter fills the tiles.
  bool lean = false;
  std::vector<uint8_t> stash;  // entries back to back
  uint32_t stash_entry = 0, stash_count = 0;
  // request slots
  size_t max_query_bytes = 0;
  std::vector<Slot> slots;
  size_t next_slot = 0;
  // Recorded after each single-query request's GEMV and packing; the next
  // request's GEMV waits for it.  Requests in flight then stream the DB and
  // the key-switch material one after the other at full bandwidth, and each
  // request's post-process and fold overlap the next one's GEMV and packing.
  // Two packings at once would each run at half speed and reach their folds
  // together, leaving HBM idle.
  hipEvent_t pack_done = nullptr;
  int gemv_kind = INSPIRE_GEMV_MFMA;
  int packing_kind = INSPIRE_PACKING_VALU;
  uint8_t* tiled = nullptr;      // fragment-ordered int8 limbs of bold_t (MFMA packing)
  // batched kernel variants (set by the autotuner or inspire_server_set_batch_kernels)
  int pack_variant = 0, gemv_variant = 0, fold_variant = 0;
  // Batched responds replayed as HIP graphs, one per launch key; a key is
  // captured on its second use (the first runs the launches directly).
  bool use_graphs = true;
  std::vector<BatchGraph> graphs;
  bool tiled_valid = false;
  Batch batch;
  std::vector<uint8_t> host_out;
  // optional per-stage timing
  bool profile = false;
  hipEvent_t ev[6] = {};
  double stage_ms[5] = {0, 0, 0, 0, 0};
  int64_t profiled = 0;
};

namespace {

size_t query_bytes(int64_t db_rows, int nq) {
  return 1 + 7 * (size_t)((int64_t)kTExp * kN * 2 + (int64_t)nq * (db_rows + 6 * kN));
}

template <typename T>
int dev_alloc(T** p, size_t count, const char* what) {
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(INSPIRE_ERR_NOMEM, std::string("hipMalloc failed for ") + what + ": " + hipGetErrorString(e));
  }
  return INSPIRE_OK;
}

}  // namespace

constexpr int kDefaultSlots = 2;

void free_graphs(inspire_server_t* s) {
  for (auto& g : s->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  s->graphs.clear();
}

void free_batch(inspire_server_t* s) {
  Batch& b = s->batch;
  if (b.done) (void)hipEventSynchronize(b.done);
  free_graphs(s);
  void* ptrs[] = {b.y_body, b.z_body, b.rows, b.P, b.cvec, b.bodies, b.S, b.keys, b.gemv_scratch};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (b.done) (void)hipEventDestroy(b.done);
  b = Batch{};
}

int alloc_batch(inspire_server_t* s, int max) {
  Batch& b = s->batch;
  const int64_t inst = s->shape.instances, rows = s->shape.db_rows, mq = s->shape.max_queries;
  int rc = INSPIRE_OK;
  b.max = max;
  if (!rc) rc = dev_alloc(&b.y_body, (size_t)max * kTExp * kN, "batch y_body");
  if (!rc) rc = dev_alloc(&b.z_body, (size_t)max * kTExp * kN, "batch z_body");
  if (!rc) rc = dev_alloc(&b.rows, (size_t)max * mq * rows, "batch rows");
  if (!rc) rc = dev_alloc(&b.bodies, (size_t)max * mq * 6 * 2 * kN, "batch bodies");
  if (!rc) rc = dev_alloc(&b.S, (size_t)max * inst * 2 * kN, "batch packing output");
  if (!rc) rc = dev_alloc(&b.P, (size_t)max * inst * kN, "batch post output");
  if (!rc) rc = dev_alloc(&b.cvec, (size_t)max * mq * s->db_cols, "batch gemv output");
  if (!rc) rc = dev_alloc(&b.keys, packing_afrag_bytes(max), "batch key fragments");
  if (!rc)
    rc = dev_alloc((uint8_t**)&b.gemv_scratch, gemv_mfma_batch_scratch_bytes((int)rows, max), "batch gemv scratch");
  if (!rc && hipEventCreateWithFlags(&b.done, hipEventDisableTiming) != hipSuccess)
    rc = set_error(INSPIRE_ERR_HIP, "hipEventCreate failed");
  return rc;
}

// Builds the fragment-ordered copy of bold_t used by the MFMA packing.
int ensure_tiled(inspire_server_t* s) {
  if (!s->tiled) {
    int rc = dev_alloc(&s->tiled, tiled_bold_t_bytes(s->shape.instances), "tiled bold_t");
    if (rc) return rc;
    s->tiled_valid = false;
  }
  if (!s->tiled_valid && s->have_precomp && s->bt && s->btb) {
    launch_tile_bold_t(s->tiled, s->bt, s->btb, s->consts->slot_dlog, s->shape.instances, s->stream);
    HIP_TRY(hipGetLastError(), "tile bold_t launch");
    HIP_TRY(hipStreamSynchronize(s->stream), "tile bold_t");
    s->tiled_valid = true;
  }
  return INSPIRE_OK;
}

void free_slots(inspire_server_t* s) {
  for (Slot& sl : s->slots) {
    if (sl.done) (void)hipEventSynchronize(sl.done);
    void* ptrs[] = {sl.qbytes, sl.y_body, sl.z_body, sl.rows,         sl.bodies,
                    sl.S,      sl.P,      sl.cvec,   sl.out, sl.gemv_scratch, sl.keys};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    if (sl.done) (void)hipEventDestroy(sl.done);
  }
  s->slots.clear();
  s->next_slot = 0;
}

int alloc_slots(inspire_server_t* s, int count) {
  const int64_t inst = s->shape.instances, rows = s->shape.db_rows, mq = s->shape.max_queries;
  s->slots.resize(count);
  for (Slot& sl : s->slots) {
    int rc = INSPIRE_OK;
    if (!rc) rc = dev_alloc(&sl.qbytes, s->max_query_bytes + 16, "query bytes");
    if (!rc) rc = dev_alloc(&sl.y_body, (size_t)kTExp * kN, "y_body");
    if (!rc) rc = dev_alloc(&sl.z_body, (size_t)kTExp * kN, "z_body");
    if (!rc) rc = dev_alloc(&sl.rows, (size_t)mq * rows, "query rows");
    if (!rc) rc = dev_alloc(&sl.bodies, (size_t)mq * 6 * 2 * kN, "rgsw bodies");
    if (!rc) rc = dev_alloc(&sl.S, (size_t)inst * 2 * kN, "packing output");
    if (!rc) rc = dev_alloc(&sl.P, (size_t)inst * kN, "post output");
    if (!rc) rc = dev_alloc(&sl.cvec, (size_t)mq * s->db_cols, "gemv output");
    if (!rc) rc = dev_alloc(&sl.out, (size_t)mq * s->c * kRespBytes, "responses");
    if (!rc) rc = dev_alloc((uint8_t**)&sl.gemv_scratch, gemv_mfma_scratch_bytes((int)rows), "gemv scratch");
    if (!rc) rc = dev_alloc(&sl.keys, packing_afrag_bytes(1), "packing key fragments");
    if (!rc && hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
      rc = set_error(INSPIRE_ERR_HIP, "hipEventCreate failed");
    if (rc) return rc;
  }
  return INSPIRE_OK;
}

static int server_create(const inspire_shape_t* shape, bool lean, inspire_server_t** out) {
  if (!shape || !out) return set_error(INSPIRE_ERR_INVALID, "null argument");
  *out = nullptr;
  if (shape->db_rows < 2048 || shape->db_rows % 2048 != 0 || shape->db_rows > (1 << 20))
    return set_error(INSPIRE_ERR_INVALID, "db_rows must be a multiple of 2048 (at most 2^20)");
  if (shape->instances <= 0 || shape->interpolate_degree <= 0 ||
      shape->instances % shape->interpolate_degree != 0)
    return set_error(INSPIRE_ERR_INVALID, "instances must be a positive multiple of interpolate_degree");
  // 65536 outputs are 2^27 DB columns, beyond one GPU's memory at any db_rows;
  // the packing kernels address a slot's key-switch fragments with 32-bit offsets.
  if (shape->instances > 65536) return set_error(INSPIRE_ERR_INVALID, "instances must be at most 65536");
  if (shape->max_queries <= 0 || shape->max_queries > 255)
    return set_error(INSPIRE_ERR_INVALID, "max_queries must be in [1, 255]");
  auto* s = new inspire_server();
  s->shape = *shape;
  s->db_cols = shape->instances * kN;
  s->c = (int)(shape->instances / shape->interpolate_degree);
  (void)hipGetDevice(&s->device);
  s->consts = device_constants();
  if (!s->consts) {
    delete s;
    return set_error(INSPIRE_ERR_HIP, "failed to build device constants");
  }
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return set_error(INSPIRE_ERR_HIP, "hipStreamCreate failed");
  }
  const int64_t inst = shape->instances, rows = shape->db_rows;
  s->max_query_bytes = query_bytes(rows, shape->max_queries);
  s->lean = lean;
  if (lean) s->packing_kind = INSPIRE_PACKING_MFMA;
  int rc = INSPIRE_OK;
  if (hipEventCreateWithFlags(&s->pack_done, hipEventDisableTiming) != hipSuccess)
    rc = set_error(INSPIRE_ERR_HIP, "hipEventCreate failed");
  if (!rc) rc = dev_alloc(&s->db, (size_t)rows * s->db_cols, "database");
  if (!rc && lean) rc = dev_alloc(&s->tiled, tiled_bold_t_bytes(inst), "tiled bold_t");
  if (!rc && !lean) rc = dev_alloc(&s->bt, (size_t)inst * kNInner * kN, "bold_t");
  if (!rc && !lean) rc = dev_alloc(&s->btb, (size_t)inst * kNInner * kN, "bold_t_bar");
  if (!rc) rc = dev_alloc(&s->bth, (size_t)inst * kTExp * kN, "bold_t_hat");
  if (!rc) rc = dev_alloc(&s->a_hat, (size_t)inst * kN, "a_hat");
  if (!rc) rc = dev_alloc(&s->ahat_ntt, (size_t)inst * 2 * kN, "NTT of a_hat");
  if (!rc) rc = alloc_slots(s, kDefaultSlots);
  // Optional, last: the fragment-ordered DB copy (the GEMVs read the
  // column-major DB when it does not fit).
  if (!rc && rows % 64 == 0 && s->db_cols % 16 == 0 &&
      hipMalloc(&s->dbfrag, tiled_db_bytes((int)rows, s->db_cols)) != hipSuccess) {
    s->dbfrag = nullptr;
    (void)hipGetLastError();
  }
  if (rc) {
    std::string msg = g_last_error;
    inspire_server_destroy(s);
    return set_error(rc, msg);
  }
  s->host_out.resize((size_t)shape->max_queries * s->c * kRespBytes);
  *out = s;
  return INSPIRE_OK;
}

extern "C" int inspire_server_create(const inspire_shape_t* shape, inspire_server_t** out) {
  return server_create(shape, false, out);
}

extern "C" int inspire_server_create_lean(const inspire_shape_t* shape, inspire_server_t** out) {
  return server_create(shape, true, out);
}

// Lean servers: a condensed bold_t / bold_t_bar pair for the duration of a
// setter (acquire), then tiled and released (release).
static int acquire_condensed(inspire_server_t* s) {
  if (!s->lean || s->bt) return INSPIRE_OK;
  const size_t n = (size_t)s->shape.instances * kNInner * kN;
  int rc = dev_alloc(&s->bt, n, "bold_t (transient)");
  if (!rc) rc = dev_alloc(&s->btb, n, "bold_t_bar (transient)");
  if (rc) {
    std::string msg = g_last_error;
    if (s->bt) (void)hipFree(s->bt);
    s->bt = nullptr;
    return set_error(rc, msg);
  }
  return INSPIRE_OK;
}

static int release_condensed(inspire_server_t* s, int rc) {
  if (!s->lean) return rc;
  if (!rc) {
    s->tiled_valid = false;
    rc = ensure_tiled(s);
  }
  (void)hipStreamSynchronize(s->stream);
  if (s->bt) (void)hipFree(s->bt);
  if (s->btb) (void)hipFree(s->btb);
  s->bt = s->btb = nullptr;
  return rc;
}

// Live DB update (encode_kernel.cu:453-516 fold + swap, packing_online_kernel.cu:
// 310-333 swap_precomp): `staged` was built (set_db / encode_db, set_precomp /
// setup_offline, stash) while `live` served; the two exchange their resident
// state.  Waits for both handles' device work; drops captured batch graphs,
// whose kernels hold the old buffers.
extern "C" int inspire_server_swap_state(inspire_server_t* live, inspire_server_t* staged) {
  if (!live || !staged || live == staged) return set_error(INSPIRE_ERR_INVALID, "two distinct servers needed");
  const inspire_shape_t &a = live->shape, &b = staged->shape;
  if (a.db_rows != b.db_rows || a.instances != b.instances || a.interpolate_degree != b.interpolate_degree ||
      live->device != staged->device)
    return set_error(INSPIRE_ERR_INVALID, "swap needs servers of the same shape on the same device");
  if (!staged->have_db || !staged->have_precomp)
    return set_error(INSPIRE_ERR_STATE, "staged server has no database or precomputation");
  for (inspire_server_t* s : {live, staged}) {
    HIP_TRY(hipStreamSynchronize(s->stream), "swap: drain server stream");
    for (Slot& sl : s->slots)
      if (sl.done) HIP_TRY(hipEventSynchronize(sl.done), "swap: drain request slot");
    if (s->batch.done) HIP_TRY(hipEventSynchronize(s->batch.done), "swap: drain batch");
    free_graphs(s);
  }
  std::swap(live->db, staged->db);
  std::swap(live->dbfrag, staged->dbfrag);
  std::swap(live->bt, staged->bt);
  std::swap(live->btb, staged->btb);
  std::swap(live->bth, staged->bth);
  std::swap(live->a_hat, staged->a_hat);
  std::swap(live->ahat_ntt, staged->ahat_ntt);
  std::swap(live->tiled, staged->tiled);
  std::swap(live->tiled_valid, staged->tiled_valid);
  std::swap(live->lean, staged->lean);
  std::swap(live->have_db, staged->have_db);
  std::swap(live->have_precomp, staged->have_precomp);
  std::swap(live->stash, staged->stash);
  std::swap(live->stash_entry, staged->stash_entry);
  std::swap(live->stash_count, staged->stash_count);
  for (inspire_server_t* s : {live, staged}) {
    if (s->lean || !s->bt) s->packing_kind = INSPIRE_PACKING_MFMA;
    if (s->packing_kind == INSPIRE_PACKING_MFMA && !s->tiled_valid) {
      int rc = ensure_tiled(s);
      if (rc) return rc;
    }
  }
  return INSPIRE_OK;
}

extern "C" void inspire_server_destroy(inspire_server_t* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  free_slots(s);
  free_batch(s);
  void* ptrs[] = {s->db, s->dbfrag, s->bt, s->btb, s->bth, s->a_hat, s->ahat_ntt, s->tiled};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  for (auto& e : s->ev)
    if (e) (void)hipEventDestroy(e);
  if (s->pack_done) (void)hipEventDestroy(s->pack_done);
  delete s;
}

extern "C" int inspire_server_set_slots(inspire_server_t* s, int count) {
  if (!s || count < 1 || count > 16) return set_error(INSPIRE_ERR_INVALID, "slot count must be in [1, 16]");
  HIP_TRY(hipDeviceSynchronize(), "drain before resizing slots");
  free_slots(s);
  return alloc_slots(s, count);
}

extern "C" void* inspire_server_stream(inspire_server_t* s) { return s ? (void*)s->stream : nullptr; }

extern "C" int inspire_server_set_packing_kernel(inspire_server_t* s, int kind) {
  if (s && s->lean && kind == INSPIRE_PACKING_VALU)
    return set_error(INSPIRE_ERR_STATE, "lean server: no condensed bold_t, MFMA packing only");
  if (!s || (kind != INSPIRE_PACKING_VALU && kind != INSPIRE_PACKING_MFMA))
    return set_error(INSPIRE_ERR_INVALID, "unknown packing kernel");
  if (kind == INSPIRE_PACKING_MFMA) {
    int rc = ensure_tiled(s);
    if (rc) return rc;
  }
  s->packing_kind = kind;
  return INSPIRE_OK;
}

extern "C" int inspire_server_set_max_batch(inspire_server_t* s, int max_batch) {
  if (!s || max_batch < 1 || max_batch > 1024) return set_error(INSPIRE_ERR_INVALID, "max_batch must be in [1, 1024]");
  HIP_TRY(hipDeviceSynchronize(), "drain before resizing the batch");
  free_batch(s);
  int rc = alloc_batch(s, max_batch);
  if (rc) {
    std::string msg = g_last_error;
    free_batch(s);
    return set_error(rc, msg);
  }
  return ensure_tiled(s);
}

extern "C" int inspire_server_set_batch_kernels(inspire_server_t* s, int packing_variant, int gemv_variant,
                                                int fold_variant) {
  if (!s || packing_variant < 0 || packing_variant >= packing_wide_variants() || gemv_variant < 0 ||
      gemv_variant >= gemv_batch_variants() || fold_variant < 0 || fold_variant > 2)
    return set_error(INSPIRE_ERR_INVALID, "unknown kernel variant");
  s->pack_variant = packing_variant;
  s->gemv_variant = gemv_variant;
  s->fold_variant = fold_variant;
  return INSPIRE_OK;
}

// Times each batched kernel variant on the device for `count` keyword queries
// of nq hash queries (on the batch buffers' contents: the kernels' work and
// memory traffic do not depend on the values) inside the whole stage
// sequence, keeps the fastest of each stage, then measures the batch-stage
// throughput at batch sizes 1, 2, 4, ...
// up to max_batch with those kernels and suggests the smallest batch within
// 10% of the best throughput.
extern "C" int inspire_server_autotune(inspire_server_t* s, int count, int nq, int reps, inspire_tune_t* out) {
  if (!s || !out) return set_error(INSPIRE_ERR_INVALID, "null argument");
  if (!s->have_db || !s->have_precomp) return set_error(INSPIRE_ERR_STATE, "database or precomputation not loaded");
  if (count < 1 || count > s->batch.max) return set_error(INSPIRE_ERR_INVALID, "count must be in [1, max_batch]");
  if (nq < 1 || nq > 2 || nq > s->shape.max_queries) return set_error(INSPIRE_ERR_INVALID, "hash queries per query");
  if (reps < 1) reps = 1;
  if (!s->tiled_valid) {
    int rc = ensure_tiled(s);
    if (rc) return rc;
  }
  *out = inspire_tune_t{};
  Batch& b = s->batch;
  hipStream_t st = s->stream;
  const int64_t rows = s->shape.db_rows, inst = s->shape.instances;
  const size_t blob = (size_t)nq * s->c * kRespBytes;
  uint8_t* d_out = nullptr;
  int rc = dev_alloc(&d_out, blob * s->batch.max, "autotune output");
  if (rc) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    (void)hipFree(d_out);
    return set_error(INSPIRE_ERR_HIP, "hipEventCreate failed");
  }
  auto pack = [&](int n, int v) {
    (void)launch_packing_mfma(b.S, inst * 2 * kN, s->tiled, b.keys, b.y_body, (int64_t)kTExp * kN, n, inst, st, v);
  };
  auto gemv = [&](int n, int v) {
    (void)launch_gemv_mfma_batch(b.cvec, (int64_t)nq * s->db_cols, b.rows, (int64_t)nq * rows, nq, s->db, s->dbfrag,
                                 (int)rows, s->db_cols, n, b.gemv_scratch, st, v);
  };
  auto fold = [&](int n, int v) {
    (void)launch_fold_batch(d_out, (int64_t)blob, s->a_hat, s->ahat_ntt, b.cvec, b.P, b.bodies, s->consts->fold,
                            s->consts->tw, inst, s->shape.interpolate_degree, nq, n, st, v);
  };
  // Average milliseconds of `reps` runs after one warm-up; -1 on a HIP error.
  auto timed = [&](auto&& fn) -> double {
    fn();
    if (hipEventRecord(e0, st) != hipSuccess) return -1;
    for (int r = 0; r < reps; ++r) fn();
    if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1;
    return (double)ms / reps;
  };
  double best;
  rc = INSPIRE_OK;
  (void)hipStreamWaitEvent(st, b.done, 0);
  // Uniform residues in the batch inputs: the device draws less power on
  // zero operands and runs those kernels faster than on real queries.
  {
    const uint64_t keys = (uint64_t)s->batch.max * kTExp * kN;
    launch_synth_fill(b.y_body, 0x5eed, 16, 0, keys, kSynthCondensed, st);
    launch_synth_fill(b.z_body, 0x5eed, 17, 0, keys, kSynthCondensed, st);
    launch_synth_fill(b.rows, 0x5eed, 18, 0, (uint64_t)s->batch.max * s->shape.max_queries * rows, kSynthCondensed,
                      st);
    launch_synth_fill(b.bodies, 0x5eed, 19, 0, (uint64_t)s->batch.max * s->shape.max_queries * 6 * kN,
                      kSynthCondensed, st);
  }
  // Each candidate is timed inside the whole stage sequence (packing, post,
  // GEMV, fold), stage by stage: a kernel timed back to back on its own ranks
  // the variants differently from the same kernel between the other stages.
  auto stages = [&](int n, int pv, int gv, int fv) {
    pack(n, pv);
    launch_post_ntt(b.P, b.S, b.z_body, (int64_t)kTExp * kN, s->bth, n, inst, st);
    gemv(n, gv);
    fold(n, fv);
  };
  // the packing variants differ only for groups of 64 clients
  out->packing_variants = batch_key_pad(count) >= 64 ? packing_wide_variants() : 1;
  best = 1e30;
  for (int v = 0; v < out->packing_variants && v < 16; ++v) {
    out->packing_ms[v] = timed([&] { stages(count, v, 0, 0); });
    if (out->packing_ms[v] < 0) rc = INSPIRE_ERR_HIP;
    if (out->packing_ms[v] >= 0 && out->packing_ms[v] < best) best = out->packing_ms[v], out->packing_variant = v;
  }
  out->gemv_variants = s->dbfrag ? gemv_batch_variants() : 1;
  best = 1e30;
  for (int v = 0; v < out->gemv_variants && v < 8; ++v) {
    out->gemv_ms[v] = timed([&] { stages(count, out->packing_variant, v, 0); });
    if (out->gemv_ms[v] < 0) rc = INSPIRE_ERR_HIP;
    if (out->gemv_ms[v] >= 0 && out->gemv_ms[v] < best) best = out->gemv_ms[v], out->gemv_variant = v;
  }
  best = 1e30;
  for (int v = 1; v <= 2; ++v) {
    out->fold_ms[v] = timed([&] { stages(count, out->packing_variant, out->gemv_variant, v); });
    if (out->fold_ms[v] < 0) rc = INSPIRE_ERR_HIP;
    if (out->fold_ms[v] >= 0 && out->fold_ms[v] < best) best = out->fold_ms[v], out->fold_variant = v;
  }
  if (rc == INSPIRE_OK) {
    s->pack_variant = out->packing_variant;
    s->gemv_variant = out->gemv_variant;
    s->fold_variant = out->fold_variant;
    // batch-size sweep with the chosen kernels (packing, post, GEMV, fold)
    double top = 0;
    for (int n = 1; n <= s->batch.max && out->nbatch < 16; n *= 2) {
      const double ms = timed([&] { stages(n, s->pack_variant, s->gemv_variant, s->fold_variant); });
      if (ms <= 0) {
        rc = INSPIRE_ERR_HIP;
        break;
      }
      out->batch_sizes[out->nbatch] = n;
      out->batch_qps[out->nbatch] = n / (ms * 1e-3);
      top = std::max(top, out->batch_qps[out->nbatch]);
      ++out->nbatch;
    }
    for (int i = 0; i < out->nbatch; ++i)
      if (out->batch_qps[i] >= 0.9 * top) {
        out->suggested_batch = out->batch_sizes[i];
        break;
      }
  }
  (void)hipEventRecord(b.done, st);
  (void)hipStreamSynchronize(st);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(d_out);
  if (rc != INSPIRE_OK) return set_error(rc, "autotune: a timed launch failed");
  HIP_TRY(hipGetLastError(), "autotune launch");
  return INSPIRE_OK;
}

static int launch_batch(inspire_server_t* s, const uint8_t* d_queries, size_t q_stride, int nq, int count,
                        uint8_t* d_out, size_t out_stride, hipStream_t st, int mark_events);

extern "C" int inspire_server_set_graphs(inspire_server_t* s, int on) {
  if (!s) return set_error(INSPIRE_ERR_INVALID, "null server");
  HIP_TRY(hipDeviceSynchronize(), "drain before changing graph use");
  free_graphs(s);
  s->use_graphs = on != 0;
  return INSPIRE_OK;
}

extern "C" int inspire_respond_batch_device(inspire_server_t* s, const uint8_t* d_queries, size_t q_stride,
                                            size_t len, int nq, int count, uint8_t* d_out, size_t out_stride,
                                            void* stream) {
  if (!s || !d_queries || !d_out) return set_error(INSPIRE_ERR_INVALID, "null argument");
  if (!s->have_db || !s->have_precomp) return set_error(INSPIRE_ERR_STATE, "database or precomputation not loaded");
  if (count < 1 || count > s->batch.max) return set_error(INSPIRE_ERR_INVALID, "batch larger than max_batch");
  if (nq < 1 || nq > s->shape.max_queries || nq > 2) return set_error(INSPIRE_ERR_INVALID, "hash queries per query");
  if (len < query_bytes(s->shape.db_rows, nq) || (count > 1 && q_stride < len))
    return set_error(INSPIRE_ERR_INVALID, "keyword query truncated");
  const size_t blob = (size_t)nq * s->c * kRespBytes;
  if (count > 1 && out_stride < blob) return set_error(INSPIRE_ERR_INVALID, "output stride too small");
  if (!s->tiled_valid) {
    int rc = ensure_tiled(s);
    if (rc) return rc;
  }
  hipStream_t st = stream ? (hipStream_t)stream : s->stream;
  Batch& b = s->batch;
  HIP_TRY(hipStreamWaitEvent(st, b.done, 0), "batch wait");
  // Graph replay: the same launch key (count, nq, buffers, stream, kernel
  // variants) replays the sequence captured on its second use.
  BatchGraph key;
  key.count = count, key.nq = nq, key.queries = d_queries, key.q_stride = q_stride, key.out = d_out;
  key.out_stride = out_stride, key.stream = st;
  key.variants = s->pack_variant | s->gemv_variant << 4 | s->fold_variant << 8;
  BatchGraph* g = nullptr;
  if (s->use_graphs && !s->profile) {
    for (auto& e : s->graphs)
      if (e.same(key)) g = &e;
    if (g && g->exec) {
      HIP_TRY(hipGraphLaunch(g->exec, st), "batch graph launch");
      HIP_TRY(hipEventRecord(b.done, st), "batch done");
      return INSPIRE_OK;
    }
    if (!g) {
      if (s->graphs.size() >= 256) free_graphs(s);
      s->graphs.push_back(key);  // first use: run directly below
    } else {
      hipGraph_t graph = nullptr;
      HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "batch capture");
      int rc = launch_batch(s, d_queries, q_stride, nq, count, d_out, out_stride, st, -1);
      hipError_t e = hipStreamEndCapture(st, &graph);
      if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
      if (graph) (void)hipGraphDestroy(graph);
      if (e != hipSuccess) {
        g->exec = nullptr;
        return set_error(INSPIRE_ERR_HIP, std::string("batch graph: ") + hipGetErrorString(e));
      }
      HIP_TRY(hipGraphLaunch(g->exec, st), "batch graph launch");
      HIP_TRY(hipEventRecord(b.done, st), "batch done");
      return INSPIRE_OK;
    }
  }
  int rc = launch_batch(s, d_queries, q_stride, nq, count, d_out, out_stride, st, s->profile ? 0 : -1);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(b.done, st), "batch done");
  if (s->profile) {
    HIP_TRY(hipEventSynchronize(s->ev[5]), "profile sync");
    for (int i = 0; i < 5; ++i) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]), "profile elapsed");
      s->stage_ms[i] += ms;
    }
    ++s->profiled;
  }
  return INSPIRE_OK;
}

// The launches of one batched respond on `st`; with mark_events >= 0 an event
// is recorded before and after each stage for the stage timer.
static int launch_batch(inspire_server_t* s, const uint8_t* d_queries, size_t q_stride, int nq, int count,
                        uint8_t* d_out, size_t out_stride, hipStream_t st, int mark_events) {
  Batch& b = s->batch;
  const int64_t rows = s->shape.db_rows, inst = s->shape.instances;
  auto mark = [&](int i) {
    if (mark_events >= 0) (void)hipEventRecord(s->ev[i], st);
  };
  mark(0);
  launch_unpack_query_batch(d_queries, (int64_t)q_stride, nq, (int)rows, count, b.y_body, b.z_body, b.rows, b.bodies,
                            st);
  mark(1);
  if (launch_packing_mfma(b.S, inst * 2 * kN, s->tiled, b.keys, b.y_body, (int64_t)kTExp * kN, count, inst, st,
                          s->pack_variant))
    return set_error(INSPIRE_ERR_INVALID, "packing: shape beyond the kernels' 32-bit fragment offsets");
  mark(2);
  launch_post_ntt(b.P, b.S, b.z_body, (int64_t)kTExp * kN, s->bth, count, inst, st);
  mark(3);
  if (launch_gemv_mfma_batch(b.cvec, (int64_t)nq * s->db_cols, b.rows, (int64_t)nq * rows, nq, s->db, s->dbfrag,
                             (int)rows, s->db_cols, count, b.gemv_scratch, st, s->gemv_variant))
    return set_error(INSPIRE_ERR_INVALID, "unsupported GEMV shape");
  mark(4);
  if (launch_fold_batch(d_out, (int64_t)out_stride, s->a_hat, s->ahat_ntt, b.cvec, b.P, b.bodies, s->consts->fold,
                        s->consts->tw, inst, s->shape.interpolate_degree, nq, count, st, s->fold_variant))
    return set_error(INSPIRE_ERR_INVALID, "unsupported fold shape");
  mark(5);
  HIP_TRY(hipGetLastError(), "batch launch");
  return INSPIRE_OK;
}

extern "C" int inspire_server_set_gemv_kernel(inspire_server_t* s, int kind) {
  if (!s || (kind != INSPIRE_GEMV_VALU && kind != INSPIRE_GEMV_MFMA))
    return set_error(INSPIRE_ERR_INVALID, "unknown GEMV kernel");
  s->gemv_kind = kind;
  return INSPIRE_OK;
}

extern "C" int inspire_server_profile_enable(inspire_server_t* s, int enable) {
  if (!s) return set_error(INSPIRE_ERR_INVALID, "null server");
  if (enable && !s->ev[0]) {
    for (auto& e : s->ev) HIP_TRY(hipEventCreate(&e), "hipEventCreate");
  }
  s->profile = enable != 0;
  return INSPIRE_OK;
}

extern "C" int inspire_server_profile_read(inspire_server_t* s, double* ms5, int64_t* count, int reset) {
  if (!s || !ms5 || !count) return set_error(INSPIRE_ERR_INVALID, "null argument");
  for (int i = 0; i < 5; ++i) ms5[i] = s->stage_ms[i];
  *count = s->profiled;
  if (reset) {
    for (double& v : s->stage_ms) v = 0;
    s->profiled = 0;
  }
  return INSPIRE_OK;
}

extern "C" int inspire_server_set_db(inspire_server_t* s, const uint16_t* db, int host) {
  if (!s || !db) return set_error(INSPIRE_ERR_INVALID, "null argument");
  const size_t bytes = (size_t)s->shape.db_rows * s->db_cols * sizeof(uint16_t);
  HIP_TRY(hipMemcpyAsync(s->db, db, bytes, host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s->stream),
          "copy database");
  if (s->dbfrag) launch_tile_db(s->dbfrag, s->db, (int)s->shape.db_rows, s->db_cols, s->stream);
  HIP_TRY(hipGetLastError(), "tile database");
  HIP_TRY(hipStreamSynchronize(s->stream), "copy database");
  s->have_db = true;
  return INSPIRE_OK;
}

extern "C" int inspire_server_set_precomp(inspire_server_t* s, const uint64_t* bt, const uint64_t* btb,
                                          const uint64_t* bth, const uint64_t* a_hat, int host) {
  if (!s || !bt || !btb || !bth || !a_hat) return set_error(INSPIRE_ERR_INVALID, "null argument");
  const size_t inst = (size_t)s->shape.instances;
  const hipMemcpyKind k = host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  if (s->lean) {
    int rc = acquire_condensed(s);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(s->bt, bt, inst * kNInner * kN * 8, k, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->btb, btb, inst * kNInner * kN * 8, k, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->bth, bth, inst * kTExp * kN * 8, k, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->a_hat, a_hat, inst * kN * 8, k, s->stream);
    if (e == hipSuccess) {
      launch_ahat_ntt(s->ahat_ntt, s->a_hat, s->consts->tw, s->shape.instances, s->stream);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    s->have_precomp = e == hipSuccess;
    return release_condensed(s, e == hipSuccess ? INSPIRE_OK : hip_error(e, "copy precomp"));
  }
  HIP_TRY(hipMemcpyAsync(s->bt, bt, inst * kNInner * kN * 8, k, s->stream), "copy bold_t");
  HIP_TRY(hipMemcpyAsync(s->btb, btb, inst * kNInner * kN * 8, k, s->stream), "copy bold_t_bar");
  HIP_TRY(hipMemcpyAsync(s->bth, bth, inst * kTExp * kN * 8, k, s->stream), "copy bold_t_hat");
  HIP_TRY(hipMemcpyAsync(s->a_hat, a_hat, inst * kN * 8, k, s->stream), "copy a_hat");
  launch_ahat_ntt(s->ahat_ntt, s->a_hat, s->consts->tw, s->shape.instances, s->stream);
  HIP_TRY(hipGetLastError(), "NTT of a_hat");
  HIP_TRY(hipStreamSynchronize(s->stream), "copy precomp");
  s->have_precomp = true;
  s->tiled_valid = false;
  return s->tiled ? ensure_tiled(s) : INSPIRE_OK;
}

// Work group of instances for the offline precomputation: as many as fit in
// free memory less 1 GiB, at most as many as keep all their collapse chains
// resident at once (384 on 256 CUs), so that the latency-bound collapse runs
// in as few passes as possible.
static int setup_group(int64_t instances, int64_t* group) {
  size_t free_b = 0, total_b = 0;
  HIP_TRY(hipMemGetInfo(&free_b, &total_b), "memory info");
  const size_t spare = free_b > (1ull << 30) ? free_b - (1ull << 30) : 0;
  const int64_t resident = setup_offline_resident_instances();
  if (resident <= 0) return set_error(INSPIRE_ERR_HIP, "offline precomputation: device occupancy query failed");
  *group = std::min<int64_t>({instances, resident, (int64_t)(spare / setup_offline_bytes_per_instance())});
  if (*group < 1) return set_error(INSPIRE_ERR_NOMEM, "no room for the offline precomputation of one instance");
  return INSPIRE_OK;
}

extern "C" int inspire_setup_offline_device(const uint16_t* d_db, int64_t db_rows, int64_t instances,
                                            uint64_t* d_bt, uint64_t* d_btb, uint64_t* d_bth, uint64_t* d_a_hat,
                                            void* stream) {
  if (!d_db || !d_bt || !d_btb || !d_bth || !d_a_hat) return set_error(INSPIRE_ERR_INVALID, "null argument");
  if (db_rows <= 0 || db_rows % kN || instances <= 0 || instances > (1 << 20))
    return set_error(INSPIRE_ERR_INVALID, "setup shape: db_rows a positive multiple of 2048, 0 < instances <= 2^20");
  const DeviceConstants* dc = device_constants();
  if (!dc) return set_error(INSPIRE_ERR_HIP, "device constants");
  int64_t group = 0;
  int rc = setup_group(instances, &group);
  if (rc) return rc;
  return launch_setup_offline(d_db, db_rows, instances, d_bt, d_btb, d_bth, d_a_hat, dc->tw, dc->fold, group,
                              (hipStream_t)stream);
}

extern "C" int inspire_server_setup_offline(inspire_server_t* s) {
  if (!s) return set_error(INSPIRE_ERR_INVALID, "null server");
  if (!s->have_db) return set_error(INSPIRE_ERR_STATE, "setup_offline needs the database (set_db or encode_db)");
  int64_t group = 0;
  int rc = setup_group(s->shape.instances, &group);
  if (rc) return rc;
  rc = acquire_condensed(s);
  if (rc) return rc;
  rc = launch_setup_offline(s->db, s->shape.db_rows, s->shape.instances, s->bt, s->btb, s->bth, s->a_hat,
                            s->consts->tw, s->consts->fold, group, s->stream);
  if (!rc) {
    launch_ahat_ntt(s->ahat_ntt, s->a_hat, s->consts->tw, s->shape.instances, s->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) rc = hip_error(e, "NTT of a_hat");
  }
  if (s->lean) {
    s->have_precomp = rc == INSPIRE_OK;
    return release_condensed(s, rc);
  }
  if (rc) return rc;
  s->have_precomp = true;
  s->tiled_valid = false;
  return s->tiled ? ensure_tiled(s) : INSPIRE_OK;
}

extern "C" int inspire_server_set_stash(inspire_server_t* s, const uint8_t* entries, uint32_t entry_size,
                                        uint32_t count) {
  if (!s || (count && !entries)) return set_error(INSPIRE_ERR_INVALID, "null argument");
  s->stash.assign(entries, entries + (size_t)entry_size * count);
  s->stash_entry = entry_size;
  s->stash_count = count;
  return INSPIRE_OK;
}

// Stream ids of the synthetic generator (the oracle uses the same numbers).
enum { kStreamDb = 1, kStreamBt = 2, kStreamBtb = 3, kStreamBth = 4, kStreamAhat = 5 };

// d^-1 mod 65535 for the encode's power-of-two degrees.
static uint32_t p16_inverse(int d) {
  for (uint32_t x = 1; x < 65535; ++x)
    if ((uint64_t)x * (uint32_t)d % 65535 == 1) return x;
  return 0;
}

// Encodes in place (or from d_in) with a work buffer of as many column groups
// as fit in free memory less 1 GiB (all of them when possible).
static int encode_db(uint16_t* d_out, const uint16_t* d_in, int64_t db_rows, int64_t db_cols, int d,
                     hipStream_t st) {
  if (d < 1 || d > 32 || (d & (d - 1)) || db_rows <= 0 || db_rows % 64 || db_cols <= 0 ||
      db_cols % ((int64_t)d * kN))
    return set_error(INSPIRE_ERR_INVALID, "encode shape: interp_degree a power of two <= 32, db_rows % 64 == 0, "
                                          "db_cols % (interp_degree * 2048) == 0");
  const int64_t c = db_cols / ((int64_t)d * kN);
  const size_t per_group = encode_work_bytes_per_group(db_rows, d);
  size_t free_b = 0, total_b = 0;
  HIP_TRY(hipMemGetInfo(&free_b, &total_b), "memory info");
  const size_t spare = free_b > (1ull << 30) ? free_b - (1ull << 30) : 0;
  const int64_t groups = std::min<int64_t>(c, (int64_t)(spare / per_group));
  if (groups < 1) return set_error(INSPIRE_ERR_NOMEM, "no room for one column group of encode work units");
  uint16_t* work = nullptr;
  int rc = dev_alloc(&work, (size_t)groups * per_group / sizeof(uint16_t), "encode work buffer");
  if (rc) return rc;
  if (launch_encode_db(d_out, d_in, db_rows, db_cols, d, p16_inverse(d), work, groups, st)) {
    (void)hipFree(work);
    return set_error(INSPIRE_ERR_INVALID, "unsupported encode shape");
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(work);
  if (e != hipSuccess) return set_error(INSPIRE_ERR_HIP, std::string("encode: ") + hipGetErrorString(e));
  return INSPIRE_OK;
}

extern "C" int inspire_encode_db_device(uint16_t* d_out, const uint16_t* d_in, int64_t db_rows, int64_t db_cols,
                                        int interp_degree, void* stream) {
  if (!d_out || !d_in) return set_error(INSPIRE_ERR_INVALID, "null argument");
  return encode_db(d_out, d_in, db_rows, db_cols, interp_degree, (hipStream_t)stream);
}

extern "C" int inspire_server_encode_db(inspire_server_t* s, const uint16_t* raw, int host) {
  if (!s || !raw) return set_error(INSPIRE_ERR_INVALID, "null argument");
  const size_t bytes = (size_t)s->shape.db_rows * s->db_cols * sizeof(uint16_t);
  HIP_TRY(hipMemcpyAsync(s->db, raw, bytes, host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s->stream),
          "copy raw database");
  int rc = encode_db(s->db, s->db, s->shape.db_rows, s->db_cols, s->shape.interpolate_degree, s->stream);
  if (rc) return rc;
  if (s->dbfrag) launch_tile_db(s->dbfrag, s->db, (int)s->shape.db_rows, s->db_cols, s->stream);
  HIP_TRY(hipGetLastError(), "tile database");
  HIP_TRY(hipStreamSynchronize(s->stream), "tile database");
  s->have_db = true;
  return INSPIRE_OK;
}

extern "C" int inspire_server_fill_synthetic(inspire_server_t* s, uint64_t seed) {
  if (!s) return set_error(INSPIRE_ERR_INVALID, "null server");
  const uint64_t inst = (uint64_t)s->shape.instances;
  launch_synth_fill(s->db, seed, kStreamDb, 0, (uint64_t)s->shape.db_rows * s->db_cols, kSynthU16, s->stream);
  if (s->lean) {
    launch_tile_bold_t_synth(s->tiled, seed, kStreamBt, kStreamBtb, s->consts->slot_dlog, s->shape.instances,
                             s->stream);
  } else {
    launch_synth_fill(s->bt, seed, kStreamBt, 0, inst * kNInner * kN, kSynthCondensed, s->stream);
    launch_synth_fill(s->btb, seed, kStreamBtb, 0, inst * kNInner * kN, kSynthCondensed, s->stream);
  }
  launch_synth_fill(s->bth, seed, kStreamBth, 0, inst * kTExp * kN, kSynthCondensed, s->stream);
  launch_synth_fill(s->a_hat, seed, kStreamAhat, 0, inst * kN, kSynthModQ, s->stream);
  launch_ahat_ntt(s->ahat_ntt, s->a_hat, s->consts->tw, s->shape.instances, s->stream);
  if (s->dbfrag) launch_tile_db(s->dbfrag, s->db, (int)s->shape.db_rows, s->db_cols, s->stream);
  HIP_TRY(hipGetLastError(), "synthetic fill launch");
  HIP_TRY(hipStreamSynchronize(s->stream), "synthetic fill");
  s->have_db = s->have_precomp = true;
  s->tiled_valid = s->lean;
  return s->tiled ? ensure_tiled(s) : INSPIRE_OK;
}

extern "C" size_t inspire_server_response_bytes(const inspire_server_t* s, int nq) {
  if (!s) return 0;
  size_t n = 4 + (size_t)nq * (4 + (size_t)s->c * kRespBytes) + 4;
  n += (size_t)s->stash_count * (4 + s->stash_entry);
  return n;
}

namespace {

// The respond pipeline on device (keyword.rs:534-628).  The packing inner
// product depends only on the client's packing keys
// (packing_online_kernel.cu:46-91), so it runs once per keyword query and
// serves every hash query; the reference repeats it per hash query with the
// same result.
int run_pipeline(inspire_server_t* s, Slot& sl, const uint8_t* d_query, int nq, uint8_t* d_out, hipStream_t st) {
  const int64_t rows = s->shape.db_rows, inst = s->shape.instances;
  // the slot's previous request (on whatever stream) must be finished; the
  // scratch-array slot of inspire_respond_device_noalloc has no event (its
  // callers order their calls themselves)
  if (sl.done) HIP_TRY(hipStreamWaitEvent(st, sl.done, 0), "slot wait");
  auto mark = [&](int i) {
    if (s->profile) (void)hipEventRecord(s->ev[i], st);
  };
  mark(0);
  launch_unpack_query(d_query, nq, (int)rows, sl.y_body, sl.z_body, sl.rows, sl.bodies, st);
  // The operand builds of the GEMV and the packing depend on this query alone
  // and write this slot's buffers: they run before the wait below, beside the
  // previous request's packing.  (With nq > 2 the GEMV pairs share one scratch,
  // so each pair builds its operands right before its stream.)
  const bool mfma_gemv = s->gemv_kind == INSPIRE_GEMV_MFMA, mfma_pack = s->packing_kind == INSPIRE_PACKING_MFMA;
  if (mfma_pack && !s->tiled_valid) return set_error(INSPIRE_ERR_STATE, "tiled key-switch material not built");
  if (mfma_gemv && nq <= 2 && launch_gemv_mfma_prep(sl.rows, nq, (int)rows, sl.gemv_scratch, st))
    return set_error(INSPIRE_ERR_INVALID, "unsupported GEMV shape");
  if (mfma_pack) launch_packing_mfma_prep(sl.keys, sl.y_body, (int64_t)kTExp * kN, 1, st);
  mark(1);
  // GEMV, then packing: the HBM-streaming stages run back to back (after the
  // previous request's), and the post-process waits off that critical chain.
  HIP_TRY(hipStreamWaitEvent(st, s->pack_done, 0), "wait for the previous request's GEMV and packing");
  if (mfma_gemv && nq <= 2) {
    if (launch_gemv_mfma_main(sl.cvec, nq, s->db, s->dbfrag, (int)rows, s->db_cols, sl.gemv_scratch, st))
      return set_error(INSPIRE_ERR_INVALID, "unsupported GEMV shape");
  } else if (mfma_gemv) {
    // two hash queries per MFMA tile; larger nq runs in pairs
    for (int h = 0; h < nq; h += 2) {
      const int m = nq - h < 2 ? nq - h : 2;
      if (launch_gemv_mfma(sl.cvec + (size_t)h * s->db_cols, sl.rows + (size_t)h * rows, m, s->db, s->dbfrag,
                           (int)rows, s->db_cols, sl.gemv_scratch, st))
        return set_error(INSPIRE_ERR_INVALID, "unsupported GEMV shape");
    }
  } else if (launch_gemv(sl.cvec, sl.rows, nq, s->db, (int)rows, s->db_cols, st)) {
    return set_error(INSPIRE_ERR_INVALID, "unsupported GEMV shape");
  }
  mark(2);
  if (mfma_pack) {
    if (launch_packing_mfma_main(sl.S, inst * 2 * kN, s->tiled, sl.keys, 1, inst, st))
      return set_error(INSPIRE_ERR_INVALID, "packing: shape beyond the kernels' 32-bit fragment offsets");
    HIP_TRY(hipEventRecord(s->pack_done, st), "GEMV and packing done");
    mark(3);
    launch_post_ntt(sl.P, sl.S, sl.z_body, (int64_t)kTExp * kN, s->bth, 1, inst, st);
  } else {
    launch_packing(sl.S, s->bt, s->btb, s->bth, sl.y_body, sl.z_body, (int)inst, st);
    HIP_TRY(hipEventRecord(s->pack_done, st), "GEMV and packing done");
    mark(3);
    launch_post(sl.P, sl.S, s->consts->tw, (int)inst, st, 1);
  }
  mark(4);
  if (launch_fold(d_out, s->a_hat, s->ahat_ntt, sl.cvec, sl.P, sl.bodies, s->consts->fold, s->consts->tw, inst,
                  s->shape.interpolate_degree, nq, st))
    return set_error(INSPIRE_ERR_INVALID, "unsupported fold shape");
  mark(5);
  HIP_TRY(hipGetLastError(), "respond launch");
  if (sl.done) HIP_TRY(hipEventRecord(sl.done, st), "slot done");
  if (s->profile) {
    // event pairs in launch order (unpack, GEMV, packing, post, fold) and the
    // stage each one times (stage order: unpack, packing, post, GEMV, fold)
    static constexpr int kStage[5] = {0, 3, 1, 2, 4};
    HIP_TRY(hipEventSynchronize(s->ev[5]), "profile sync");
    for (int i = 0; i < 5; ++i) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]), "profile elapsed");
      s->stage_ms[kStage[i]] += ms;
    }
    ++s->profiled;
  }
  return INSPIRE_OK;
}

Slot& take_slot(inspire_server_t* s) {
  Slot& sl = s->slots[s->next_slot % s->slots.size()];
  ++s->next_slot;
  return sl;
}

}  // namespace

extern "C" int inspire_respond_device(inspire_server_t* s, const uint8_t* d_query, size_t len, int nq,
                                      uint8_t* d_out, void* stream) {
  if (!s || !d_query || !d_out) return set_error(INSPIRE_ERR_INVALID, "null argument");
  if (!s->have_db || !s->have_precomp) return set_error(INSPIRE_ERR_STATE, "database or precomputation not loaded");
  if (nq < 0 || nq > s->shape.max_queries) return set_error(INSPIRE_ERR_INVALID, "too many hash queries");
  if (len < query_bytes(s->shape.db_rows, nq)) return set_error(INSPIRE_ERR_INVALID, "keyword query truncated");
  if (nq == 0) return INSPIRE_OK;
  hipStream_t st = stream ? (hipStream_t)stream : s->stream;
  return run_pipeline(s, take_slot(s), d_query, nq, d_out, st);
}

// ---------------------------------------------------------------------------
// Respond without allocating (inspire_respond_device_noalloc): the request's
// working buffers live in one file-scope device array per GPU instead of the
// server's hipMalloc'd request slots.
// ---------------------------------------------------------------------------

// Every call on a GPU uses this array; it keeps the last call's contents.
static __device__ __align__(256) uint8_t scratch[INSPIRE_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[64] = {};  // device address of `scratch` per device

constexpr size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// Byte offsets of the request buffers inside the array for a shape and nq.
struct ScratchLayout {
  size_t y_body, z_body, rows, bodies, S, P, cvec, gemv, keys, total;
};

ScratchLayout scratch_layout(int64_t db_rows, int64_t instances, int nq) {
  ScratchLayout l{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += align256(bytes);
    return at;
  };
  l.y_body = take((size_t)kTExp * kN * 8);
  l.z_body = take((size_t)kTExp * kN * 8);
  l.rows = take((size_t)nq * db_rows * 8);
  l.bodies = take((size_t)nq * 6 * 2 * kN * 4);
  l.S = take((size_t)instances * 2 * kN * 4);
  l.P = take((size_t)instances * kN * 8);
  l.cvec = take((size_t)nq * instances * kN * 8);
  l.gemv = take(gemv_mfma_scratch_bytes((int)db_rows));
  l.keys = take(packing_afrag_bytes(1));
  l.total = o;
  return l;
}

// The current device's address of `scratch`, looked up once per device (with
// that device current); host code passes this address, never the array's
// name, to kernels.
int scratch_address(int dev, uint8_t** out) {
  if (dev < 0 || dev >= 64) return set_error(INSPIRE_ERR_INVALID, "device index beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) return set_error(INSPIRE_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = static_cast<uint8_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return INSPIRE_OK;
}

}  // namespace

extern "C" size_t inspire_noalloc_scratch_bytes(const inspire_shape_t* shape, int nq) {
  if (!shape || nq < 1 || shape->db_rows <= 0 || shape->instances <= 0) return 0;
  return scratch_layout(shape->db_rows, shape->instances, nq).total;
}

extern "C" int inspire_respond_device_noalloc(inspire_server_t* s, const uint8_t* d_query, size_t len, int nq,
                                              uint8_t* d_out, void* stream) {
  if (!s || !d_query || !d_out) return set_error(INSPIRE_ERR_INVALID, "null argument");
  if (!s->have_db || !s->have_precomp) return set_error(INSPIRE_ERR_STATE, "database or precomputation not loaded");
  if (nq < 0 || nq > s->shape.max_queries) return set_error(INSPIRE_ERR_INVALID, "too many hash queries");
  if (len < query_bytes(s->shape.db_rows, nq)) return set_error(INSPIRE_ERR_INVALID, "keyword query truncated");
  if (nq == 0) return INSPIRE_OK;
  const ScratchLayout l = scratch_layout(s->shape.db_rows, s->shape.instances, nq);
  if (l.total > (size_t)INSPIRE_NOALLOC_SCRATCH_BYTES)
    return set_error(INSPIRE_ERR_INVALID,
                     "request needs " + std::to_string(l.total) + " scratch bytes, beyond the " +
                         std::to_string((size_t)INSPIRE_NOALLOC_SCRATCH_BYTES) +
                         "-byte no-allocation array (fewer hash queries or outputs per call)");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  if (dev != s->device) return set_error(INSPIRE_ERR_INVALID, "the server's device is not the current device");
  uint8_t* base = nullptr;
  int rc = scratch_address(dev, &base);
  if (rc) return rc;
  Slot sl;  // no event: calls that overlap on one GPU race on the array
  sl.y_body = reinterpret_cast<uint64_t*>(base + l.y_body);
  sl.z_body = reinterpret_cast<uint64_t*>(base + l.z_body);
  sl.rows = reinterpret_cast<uint64_t*>(base + l.rows);
  sl.bodies = reinterpret_cast<uint32_t*>(base + l.bodies);
  sl.S = reinterpret_cast<uint32_t*>(base + l.S);
  sl.P = reinterpret_cast<uint64_t*>(base + l.P);
  sl.cvec = reinterpret_cast<uint64_t*>(base + l.cvec);
  sl.gemv_scratch = base + l.gemv;
  sl.keys = base + l.keys;
  hipStream_t st = stream ? (hipStream_t)stream : s->stream;
  return run_pipeline(s, sl, d_query, nq, d_out, st);
}

extern "C" int inspire_respond_serialized(inspire_server_t* s, const uint8_t* query, size_t len, uint8_t* resp,
                                          size_t cap, size_t* resp_len) {
  if (!s || !query || !resp_len) return set_error(INSPIRE_ERR_INVALID, "null argument");
  if (!s->have_db || !s->have_precomp) return set_error(INSPIRE_ERR_STATE, "database or precomputation not loaded");
  if (len < 1) return set_error(INSPIRE_ERR_INVALID, "empty keyword query");
  const int nq = query[0];
  if (nq > s->shape.max_queries) return set_error(INSPIRE_ERR_INVALID, "too many hash queries");
  const size_t need = query_bytes(s->shape.db_rows, nq);
  if (len < need) return set_error(INSPIRE_ERR_INVALID, "keyword query truncated");
  const size_t out_bytes = inspire_server_response_bytes(s, nq);
  *resp_len = out_bytes;
  if (!resp || cap < out_bytes) return set_error(INSPIRE_ERR_CAPACITY, "response buffer too small");
  const size_t blob = (size_t)s->c * kRespBytes;
  if (nq > 0) {
    Slot& sl = take_slot(s);
    HIP_TRY(hipStreamWaitEvent(s->stream, sl.done, 0), "slot wait");
    HIP_TRY(hipMemcpyAsync(sl.qbytes, query, need, hipMemcpyHostToDevice, s->stream), "upload query");
    int rc = run_pipeline(s, sl, sl.qbytes, nq, sl.out, s->stream);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(s->host_out.data(), sl.out, (size_t)nq * blob, hipMemcpyDeviceToHost, s->stream),
            "download responses");
    HIP_TRY(hipStreamSynchronize(s->stream), "respond");
  }
  server_write_response(s, nq, s->host_out.data(), resp);
  return INSPIRE_OK;
}

namespace inspire {

size_t server_write_response(const inspire_server* s, int nq, const uint8_t* blobs, uint8_t* resp) {
  return write_response(nq, (size_t)s->c * kRespBytes, blobs, s->stash_entry, s->stash_count, s->stash.data(),
                        resp);
}

size_t response_bytes_with_stash(int nq, size_t blob, uint32_t entry, uint32_t count) {
  return 4 + (size_t)nq * (4 + blob) + 4 + (size_t)count * (4 + entry);
}

void server_stash(const inspire_server* s, uint32_t* entry, uint32_t* count, std::vector<uint8_t>* bytes) {
  *entry = s->stash_entry;
  *count = s->stash_count;
  *bytes = s->stash;
}

size_t write_response(int nq, size_t blob, const uint8_t* blobs, uint32_t stash_entry, uint32_t stash_count,
                      const uint8_t* stash, uint8_t* resp) {
  // serialize_keyword_response (wire.rs:708-726)
  size_t o = 0;
  auto put32 = [&](uint32_t v) {
    std::memcpy(resp + o, &v, 4);
    o += 4;
  };
  put32((uint32_t)nq);
  for (int h = 0; h < nq; ++h) {
    put32((uint32_t)blob);
    std::memcpy(resp + o, blobs + (size_t)h * blob, blob);
    o += blob;
  }
  put32(stash_count);
  for (uint32_t i = 0; i < stash_count; ++i) {
    put32(stash_entry);
    std::memcpy(resp + o, stash + (size_t)i * stash_entry, stash_entry);
    o += stash_entry;
  }
  return o;
}

int server_info(const inspire_server* s, int nq, ServerInfo* out) {
  if (!s || !out || nq < 1 || nq > s->shape.max_queries) return set_error(INSPIRE_ERR_INVALID, "bad server info request");
  out->device = s->device;
  out->query_bytes = query_bytes(s->shape.db_rows, nq);
  out->blob = (size_t)nq * s->c * kRespBytes;
  out->response = inspire_server_response_bytes(s, nq);
  out->batch_max = s->batch.max;
  return INSPIRE_OK;
}

size_t server_batch_bytes(const inspire_server* s, int max) {
  const size_t inst = s->shape.instances, rows = s->shape.db_rows, mq = s->shape.max_queries;
  size_t b = 0;
  b += 2 * (size_t)max * kTExp * kN * 8;        // y_body, z_body
  b += (size_t)max * mq * rows * 8;             // query rows
  b += (size_t)max * mq * 6 * 2 * kN * 4;       // RGSW bodies
  b += (size_t)max * inst * 2 * kN * 4;         // packing output
  b += (size_t)max * inst * kN * 8;             // post output
  b += (size_t)max * mq * s->db_cols * 8;       // GEMV output
  b += packing_afrag_bytes(max);
  b += gemv_mfma_batch_scratch_bytes((int)rows, max);
  return b;
}

}  // namespace inspire

// ---------------------------------------------------------------------------
// Stage entry points
// ---------------------------------------------------------------------------

extern "C" int inspire_synth_fill(void* d_out, uint64_t seed, uint32_t stream_id, uint64_t start, uint64_t count,
                                  int kind, void* stream) {
  if (!d_out || kind < 0 || kind > 4) return set_error(INSPIRE_ERR_INVALID, "bad synth_fill argument");
  launch_synth_fill(d_out, seed, stream_id, start, count, kind, (hipStream_t)stream);
  HIP_TRY(hipGetLastError(), "synth_fill launch");
  return INSPIRE_OK;
}

extern "C" int inspire_unpack_query_device(const uint8_t* d_query, int nq, int64_t db_rows, uint64_t* d_y_body,
                                           uint64_t* d_z_body, uint64_t* d_rows, uint32_t* d_bodies,
                                           void* stream) {
  if (!d_query || !d_y_body || !d_z_body || (nq > 0 && (!d_rows || !d_bodies)) || nq < 0 || db_rows <= 0)
    return set_error(INSPIRE_ERR_INVALID, "bad unpack argument");
  launch_unpack_query(d_query, nq, (int)db_rows, d_y_body, d_z_body, d_rows, d_bodies, (hipStream_t)stream);
  HIP_TRY(hipGetLastError(), "unpack launch");
  return INSPIRE_OK;
}

extern "C" int inspire_gemv_device(uint64_t* d_c, const uint64_t* d_rows, int nq, const uint16_t* d_db,
                                   int64_t db_rows, int64_t db_cols, void* stream) {
  if (!d_c || !d_rows || !d_db || nq <= 0) return set_error(INSPIRE_ERR_INVALID, "bad gemv argument");
  if (db_rows > (1 << 20)) return set_error(INSPIRE_ERR_INVALID, "db_rows too large");
  if (launch_gemv(d_c, d_rows, nq, d_db, (int)db_rows, db_cols, (hipStream_t)stream))
    return set_error(INSPIRE_ERR_INVALID, "db_rows must be a positive multiple of 2048");
  HIP_TRY(hipGetLastError(), "gemv launch");
  return INSPIRE_OK;
}

extern "C" int inspire_gemv_mfma_device(uint64_t* d_c, const uint64_t* d_rows, int nq, const uint16_t* d_db,
                                        int64_t db_rows, int64_t db_cols, void* d_scratch, void* stream) {
  if (!d_c || !d_rows || !d_db || !d_scratch || nq < 1 || nq > 2)
    return set_error(INSPIRE_ERR_INVALID, "bad gemv argument");
  if (launch_gemv_mfma(d_c, d_rows, nq, d_db, nullptr, (int)db_rows, db_cols, d_scratch, (hipStream_t)stream))
    return set_error(INSPIRE_ERR_INVALID, "db_rows must be a multiple of 2048 and at most 2^17");
  HIP_TRY(hipGetLastError(), "gemv launch");
  return INSPIRE_OK;
}

extern "C" size_t inspire_gemv_mfma_scratch_bytes(int64_t db_rows) {
  return db_rows > 0 && db_rows <= (1 << 17) ? gemv_mfma_scratch_bytes((int)db_rows) : 0;
}

extern "C" int inspire_packing_device(uint32_t* d_S, const uint64_t* d_bt, const uint64_t* d_btb,
                                      const uint64_t* d_bth, const uint64_t* d_y_body, const uint64_t* d_z_body,
                                      int64_t instances, void* stream) {
  if (!d_S || !d_bt || !d_btb || !d_bth || !d_y_body || !d_z_body || instances <= 0 || instances > (1 << 20))
    return set_error(INSPIRE_ERR_INVALID, "bad packing argument");
  launch_packing(d_S, d_bt, d_btb, d_bth, d_y_body, d_z_body, (int)instances, (hipStream_t)stream);
  HIP_TRY(hipGetLastError(), "packing launch");
  return INSPIRE_OK;
}

extern "C" int inspire_post_device(uint64_t* d_P, const uint32_t* d_S, int64_t count, void* stream) {
  if (!d_P || !d_S || count <= 0) return set_error(INSPIRE_ERR_INVALID, "bad post argument");
  const DeviceConstants* dc = device_constants();
  if (!dc) return set_error(INSPIRE_ERR_HIP, "device constants");
  launch_post(d_P, d_S, dc->tw, (int)count, (hipStream_t)stream);
  HIP_TRY(hipGetLastError(), "post launch");
  return INSPIRE_OK;
}

extern "C" int inspire_fold_device(uint8_t* d_out, const uint64_t* d_a_hat, const uint64_t* d_c,
                                   const uint64_t* d_P, const uint32_t* d_bodies, int64_t instances, int d,
                                   int nq, void* stream) {
  if (!d_out || !d_a_hat || !d_c || !d_P || !d_bodies || nq <= 0)
    return set_error(INSPIRE_ERR_INVALID, "bad fold argument");
  const DeviceConstants* dc = device_constants();
  if (!dc) return set_error(INSPIRE_ERR_HIP, "device constants");
  if (launch_fold(d_out, d_a_hat, nullptr, d_c, d_P, d_bodies, dc->fold, dc->tw, instances, d, nq,
                  (hipStream_t)stream))
    return set_error(INSPIRE_ERR_INVALID, "instances must be a positive multiple of d");
  HIP_TRY(hipGetLastError(), "fold launch");
  return INSPIRE_OK;
}
