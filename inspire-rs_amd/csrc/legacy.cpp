// The reference's native/cuda operator surface, re-implemented on the gfx950
// kernels.  Names, argument meaning and error behaviour (void return, message
// on stderr, outputs untouched on failure) follow the Rust FFI declarations in
// src/pir/gpu/{gemv,packing_online,encode}.rs so those wrappers link unchanged.
// Device state is process-global, as in the reference (gemv.rs:27-32,
// packing_online_kernel.cu:95-107, rotation_kernel.cu:61-68).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/inspire_hip.h"
#include "kernels.hpp"
#include "modarith.hpp"
#include "runtime.hpp"

using namespace inspire;

namespace {

std::mutex g_mu;
hipStream_t g_query_stream = nullptr;

// rotation state
uint32_t* d_tables = nullptr;
uint32_t* d_gen_pows = nullptr;
uint64_t* d_y_body = nullptr;
int saved_num_tables = 0;

// packing state
uint64_t *d_bold_t = nullptr, *d_bold_t_bar = nullptr, *d_bold_t_hat = nullptr;
bool adopted_precomp = false;
uint64_t *d_y_all = nullptr, *d_y_bar_all = nullptr, *d_z_body = nullptr;
uint32_t* d_S = nullptr;
int saved_outputs = 0, saved_n_inner = 0;

#define LEGACY_CHECK(tag, call)                                                              \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      std::fprintf(stderr, "[%s] HIP error at %s:%d: %s\n", tag, __FILE__, __LINE__,         \
                   hipGetErrorString(e_));                                                   \
      return;                                                                                \
    }                                                                                        \
  } while (0)

hipStream_t query_stream() {
  if (!g_query_stream) {
    if (hipStreamCreateWithFlags(&g_query_stream, hipStreamNonBlocking) != hipSuccess) g_query_stream = nullptr;
  }
  return g_query_stream;
}

bool shape_ok(const char* tag, int t_exp_left, int poly_len) {
  if (t_exp_left != kTExp || poly_len != kN) {
    std::fprintf(stderr, "[%s] unsupported shape t_exp_left=%d poly_len=%d (expected %d, %d)\n", tag, t_exp_left,
                 poly_len, kTExp, kN);
    return false;
  }
  return true;
}

void free_ptr(void* p) {
  if (p) (void)hipFree(p);
}

}  // namespace

extern "C" void gpu_init_memory_pool(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  (void)query_stream();
}

extern "C" uint64_t gpu_available_memory_bytes(void) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
  return (uint64_t)free_b;
}

// ---------------------------------------------------------------------------
// DB encode (encode.rs:3-32, encode_kernel.cu:386-510): a primary encoded DB
// and a second slot for a rebuild that swaps in when done.
// ---------------------------------------------------------------------------

namespace {
uint16_t* d_encoded_db = nullptr;
uint16_t* d_encoded_db_new = nullptr;

// Copies and encodes a host DB into a new device buffer; nullptr on failure.
uint16_t* encode_from_host(const char* tag, const uint16_t* h_db_in, int db_rows, int db_cols, int interp_degree,
                           int gamma, int c, uint32_t modulus, uint32_t num_inv) {
  if (!h_db_in || db_rows <= 0 || db_cols <= 0 || gamma != kN || modulus != 65535 || interp_degree < 1 ||
      (int64_t)interp_degree * gamma * c != db_cols || (uint64_t)num_inv * (uint32_t)interp_degree % modulus != 1) {
    std::fprintf(stderr, "[%s] unsupported arguments (db_rows=%d db_cols=%d interp_degree=%d gamma=%d c=%d "
                 "modulus=%u num_inv=%u)\n", tag, db_rows, db_cols, interp_degree, gamma, c, modulus, num_inv);
    return nullptr;
  }
  const size_t bytes = (size_t)db_rows * db_cols * sizeof(uint16_t);
  uint16_t* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e == hipSuccess) e = hipMemcpy(p, h_db_in, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    std::fprintf(stderr, "[%s] HIP error: %s\n", tag, hipGetErrorString(e));
    free_ptr(p);
    return nullptr;
  }
  if (inspire_encode_db_device(p, p, db_rows, db_cols, interp_degree, nullptr) != INSPIRE_OK) {
    std::fprintf(stderr, "[%s] %s\n", tag, inspire_last_error());
    free_ptr(p);
    return nullptr;
  }
  return p;
}
}  // namespace

extern "C" void gpu_encode_db_from_host(const uint16_t* h_db_in, uint16_t*, int db_rows, int db_cols,
                                        int interp_degree, int gamma, int c, uint32_t modulus, uint32_t num_inv) {
  std::lock_guard<std::mutex> lk(g_mu);
  uint16_t* p = encode_from_host("GPU Encode", h_db_in, db_rows, db_cols, interp_degree, gamma, c, modulus, num_inv);
  if (!p) return;
  if (d_encoded_db && d_encoded_db != d_encoded_db_new) free_ptr(d_encoded_db);
  d_encoded_db = p;
}

extern "C" const uint16_t* gpu_encode_get_device_db(void) { return d_encoded_db; }

extern "C" void gpu_encode_free_device_db(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  free_ptr(d_encoded_db);
  d_encoded_db = nullptr;
}

extern "C" void gpu_encode_db_from_host_fold(const uint16_t* h_db_in, int db_rows, int db_cols, int interp_degree,
                                             int gamma, int c, uint32_t modulus, uint32_t num_inv) {
  std::lock_guard<std::mutex> lk(g_mu);
  uint16_t* p =
      encode_from_host("GPU Encode Fold", h_db_in, db_rows, db_cols, interp_degree, gamma, c, modulus, num_inv);
  if (!p) return;
  if (d_encoded_db_new && d_encoded_db_new != d_encoded_db) free_ptr(d_encoded_db_new);
  d_encoded_db_new = p;
}

extern "C" const uint16_t* gpu_encode_get_device_db_fold(void) { return d_encoded_db_new; }

extern "C" void gpu_encode_swap_primary(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (d_encoded_db && d_encoded_db != d_encoded_db_new) free_ptr(d_encoded_db);
  d_encoded_db = d_encoded_db_new;
  d_encoded_db_new = nullptr;
}

extern "C" void gpu_trim_memory_pool(void) { (void)hipDeviceSynchronize(); }

// ---------------------------------------------------------------------------
// GEMV (gemv.rs:5-25, gemv_kernel.cu:105-194)
// ---------------------------------------------------------------------------

extern "C" void gpu_gemv_set_device_db(uint16_t** d_db_out, const uint16_t* d_db_in, size_t) {
  if (d_db_out) *d_db_out = const_cast<uint16_t*>(d_db_in);
}

extern "C" void gpu_gemv_upload_db(const uint16_t* h_db, uint16_t** d_db_out, size_t bytes) {
  if (!d_db_out) return;
  *d_db_out = nullptr;
  uint16_t* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    std::fprintf(stderr, "[GPU GEMV] hipMalloc failed (%.1f GB): %s\n", bytes / 1e9, hipGetErrorString(e));
    return;
  }
  e = hipMemcpy(p, h_db, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    std::fprintf(stderr, "[GPU GEMV] hipMemcpy failed: %s\n", hipGetErrorString(e));
    (void)hipFree(p);
    return;
  }
  *d_db_out = p;
}

extern "C" void gpu_gemv_free_db(uint16_t* d_db) { free_ptr(d_db); }

extern "C" void gpu_gemv_query(uint64_t* h_c, const uint64_t* h_a, const uint16_t* d_db, int db_rows, int b_cols,
                               uint64_t mod0, uint64_t mod1, uint64_t cr0, uint64_t cr1, uint64_t mod0_inv_mod1,
                               uint64_t mod1_inv_mod0, uint64_t modulus, uint64_t, uint64_t) {
  // The kernel hard-wires the reference moduli (src/pir/params.rs:9).
  if (mod0 != kQ0 || mod1 != kQ1 || cr0 != kCr1Q0 || cr1 != kCr1Q1 || modulus != kQ ||
      mod0_inv_mod1 != 26136460727815280ull || mod1_inv_mod0 != 40838229011788690ull) {
    std::fprintf(stderr, "[GPU GEMV] unsupported moduli\n");
    return;
  }
  if (!h_c || !h_a || !d_db || db_rows <= 0 || db_rows % 2048 != 0 || b_cols <= 0) {
    std::fprintf(stderr, "[GPU GEMV] unsupported shape db_rows=%d b_cols=%d\n", db_rows, b_cols);
    return;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  hipStream_t s = query_stream();
  uint64_t *d_a = nullptr, *d_c = nullptr;
  LEGACY_CHECK("GPU GEMV", hipMallocAsync((void**)&d_a, (size_t)db_rows * 8, s));
  hipError_t e = hipMallocAsync((void**)&d_c, (size_t)b_cols * 8, s);
  if (e != hipSuccess) {
    (void)hipFreeAsync(d_a, s);
    LEGACY_CHECK("GPU GEMV", e);
  }
  // every exit below releases both buffers
  e = hipMemcpyAsync(d_a, h_a, (size_t)db_rows * 8, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    launch_gemv(d_c, d_a, 1, d_db, db_rows, b_cols, s);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(h_c, d_c, (size_t)b_cols * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFreeAsync(d_a, s);
  (void)hipFreeAsync(d_c, s);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(s);
    std::fprintf(stderr, "[GPU GEMV] %s\n", hipGetErrorString(e));
  }
}

// ---------------------------------------------------------------------------
// Rotations (packing_online.rs:23-40, rotation_kernel.cu:61-144)
// ---------------------------------------------------------------------------

extern "C" void gpu_rotation_upload_tables(const uint32_t* h_tables, const uint32_t* h_gen_pows, int num_tables,
                                           int num_rotations, int poly_len) {
  std::lock_guard<std::mutex> lk(g_mu);
  free_ptr(d_tables);
  free_ptr(d_gen_pows);
  d_tables = nullptr;
  d_gen_pows = nullptr;
  saved_num_tables = num_tables;
  const size_t tb = (size_t)num_tables * poly_len * 4, gb = (size_t)num_rotations * 4;
  LEGACY_CHECK("GPU Rotation", hipMalloc(&d_tables, tb));
  LEGACY_CHECK("GPU Rotation", hipMalloc(&d_gen_pows, gb));
  LEGACY_CHECK("GPU Rotation", hipMemcpy(d_tables, h_tables, tb, hipMemcpyHostToDevice));
  LEGACY_CHECK("GPU Rotation", hipMemcpy(d_gen_pows, h_gen_pows, gb, hipMemcpyHostToDevice));
}

extern "C" void gpu_rotation_expand_keys(uint64_t* d_y_all_out, uint64_t* d_y_bar_all_out, const uint64_t* h_y_body,
                                         int num_rotations, int t_exp_left, int poly_len) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!d_tables || !d_gen_pows) {
    std::fprintf(stderr, "[GPU Rotation] Tables not uploaded!\n");
    return;
  }
  if (num_rotations <= 0 || num_rotations > poly_len / 2 || saved_num_tables < poly_len ||
      !shape_ok("GPU Rotation", t_exp_left, poly_len))
    return;
  hipStream_t s = query_stream();
  const size_t body = (size_t)t_exp_left * poly_len * 8;
  if (!d_y_body) LEGACY_CHECK("GPU Rotation", hipMalloc(&d_y_body, body));
  LEGACY_CHECK("GPU Rotation", hipMemcpyAsync(d_y_body, h_y_body, body, hipMemcpyHostToDevice, s));
  launch_rotation_legacy(d_y_all_out, d_y_bar_all_out, d_y_body, d_tables, d_gen_pows, num_rotations, t_exp_left,
                         poly_len, s);
  LEGACY_CHECK("GPU Rotation", hipGetLastError());
  LEGACY_CHECK("GPU Rotation", hipStreamSynchronize(s));
}

extern "C" void gpu_rotation_free(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  free_ptr(d_tables);
  free_ptr(d_gen_pows);
  free_ptr(d_y_body);
  d_tables = nullptr;
  d_gen_pows = nullptr;
  d_y_body = nullptr;
}

// ---------------------------------------------------------------------------
// Online packing (packing_online.rs:12-96, packing_online_kernel.cu:95-333)
// ---------------------------------------------------------------------------

namespace {

void free_keys() {
  free_ptr(d_y_all);
  free_ptr(d_y_bar_all);
  free_ptr(d_z_body);
  d_y_all = d_y_bar_all = d_z_body = nullptr;
}

void free_precomp() {
  if (!adopted_precomp) {
    free_ptr(d_bold_t);
    free_ptr(d_bold_t_bar);
    free_ptr(d_bold_t_hat);
  }
  d_bold_t = d_bold_t_bar = d_bold_t_hat = nullptr;
  adopted_precomp = false;
}

void realloc_output(int num_outputs) {
  free_ptr(d_S);
  d_S = nullptr;
  saved_outputs = 0;
  if (hipMalloc(&d_S, (size_t)num_outputs * 2 * kN * 4) == hipSuccess) saved_outputs = num_outputs;
}

}  // namespace

extern "C" void gpu_packing_online_ensure_key_buffers(int n_inner, int t_exp_left, int poly_len) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!shape_ok("GPU Packing Online", t_exp_left, poly_len)) return;
  if (d_y_all && n_inner == saved_n_inner) return;
  free_keys();
  const size_t y = (size_t)n_inner * poly_len * 8, z = (size_t)t_exp_left * poly_len * 8;
  LEGACY_CHECK("GPU Packing Online", hipMalloc(&d_y_all, y));
  LEGACY_CHECK("GPU Packing Online", hipMalloc(&d_y_bar_all, y));
  LEGACY_CHECK("GPU Packing Online", hipMalloc(&d_z_body, z));
  saved_n_inner = n_inner;
}

extern "C" uint64_t* gpu_packing_online_get_d_y_all(void) { return d_y_all; }
extern "C" uint64_t* gpu_packing_online_get_d_y_bar_all(void) { return d_y_bar_all; }

extern "C" void gpu_packing_online_upload_z_body(const uint64_t* h_z_body, int t_exp_left, int poly_len) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!d_z_body || !shape_ok("GPU Packing Online", t_exp_left, poly_len)) return;
  hipStream_t s = query_stream();
  LEGACY_CHECK("GPU Packing Online",
               hipMemcpyAsync(d_z_body, h_z_body, (size_t)t_exp_left * poly_len * 8, hipMemcpyHostToDevice, s));
  LEGACY_CHECK("GPU Packing Online", hipStreamSynchronize(s));
}

extern "C" void gpu_packing_online_upload_precomp(const uint64_t* h_bt, const uint64_t* h_btb, const uint64_t* h_bth,
                                                  int num_outputs, int n_inner, int t_exp_left, int poly_len) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!shape_ok("GPU Packing Online", t_exp_left, poly_len)) return;
  free_precomp();
  const size_t bt = (size_t)num_outputs * n_inner * poly_len * 8, bh = (size_t)num_outputs * t_exp_left * poly_len * 8;
  LEGACY_CHECK("GPU Packing Online", hipMalloc(&d_bold_t, bt));
  LEGACY_CHECK("GPU Packing Online", hipMalloc(&d_bold_t_bar, bt));
  LEGACY_CHECK("GPU Packing Online", hipMalloc(&d_bold_t_hat, bh));
  LEGACY_CHECK("GPU Packing Online", hipMemcpy(d_bold_t, h_bt, bt, hipMemcpyHostToDevice));
  LEGACY_CHECK("GPU Packing Online", hipMemcpy(d_bold_t_bar, h_btb, bt, hipMemcpyHostToDevice));
  LEGACY_CHECK("GPU Packing Online", hipMemcpy(d_bold_t_hat, h_bth, bh, hipMemcpyHostToDevice));
  saved_n_inner = saved_n_inner ? saved_n_inner : n_inner;
  realloc_output(num_outputs);
}

extern "C" void gpu_packing_online_adopt_precomp(uint64_t* bt, uint64_t* btb, uint64_t* bth, int num_outputs, int,
                                                 int t_exp_left, int poly_len) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!shape_ok("GPU Packing Online", t_exp_left, poly_len)) return;
  free_precomp();
  d_bold_t = bt;
  d_bold_t_bar = btb;
  d_bold_t_hat = bth;
  adopted_precomp = true;
  realloc_output(num_outputs);
}

extern "C" void gpu_packing_online_swap_precomp(uint64_t* bt, uint64_t* btb, uint64_t* bth, int num_outputs, int,
                                                int t_exp_left, int poly_len) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!shape_ok("GPU Packing Online", t_exp_left, poly_len)) return;
  d_bold_t = bt;
  d_bold_t_bar = btb;
  d_bold_t_hat = bth;
  adopted_precomp = true;
  realloc_output(num_outputs);
}

extern "C" void gpu_packing_online_upload_keys(const uint64_t* h_y_all, const uint64_t* h_y_bar_all,
                                               const uint64_t* h_z_body, int n_inner, int t_exp_left, int poly_len) {
  gpu_packing_online_ensure_key_buffers(n_inner, t_exp_left, poly_len);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!d_y_all) return;
  hipStream_t s = query_stream();
  const size_t y = (size_t)n_inner * poly_len * 8, z = (size_t)t_exp_left * poly_len * 8;
  LEGACY_CHECK("GPU Packing Online", hipMemcpyAsync(d_y_all, h_y_all, y, hipMemcpyHostToDevice, s));
  LEGACY_CHECK("GPU Packing Online", hipMemcpyAsync(d_y_bar_all, h_y_bar_all, y, hipMemcpyHostToDevice, s));
  LEGACY_CHECK("GPU Packing Online", hipMemcpyAsync(d_z_body, h_z_body, z, hipMemcpyHostToDevice, s));
  LEGACY_CHECK("GPU Packing Online", hipStreamSynchronize(s));
}

extern "C" void gpu_packing_online_compute(uint64_t* h_output, int num_outputs, int nphalf_minus_1, int t_exp_left,
                                           int poly_len, int, uint64_t mod0, uint64_t mod1, uint64_t cr0,
                                           uint64_t cr1) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!d_bold_t || !d_y_all) {
    std::fprintf(stderr, "[GPU Packing Online] Data not uploaded!\n");
    return;
  }
  if (!shape_ok("GPU Packing Online", t_exp_left, poly_len) || nphalf_minus_1 != kNumRot || mod0 != kQ0 ||
      mod1 != kQ1 || cr0 != kCr1Q0 || cr1 != kCr1Q1 || num_outputs <= 0 || num_outputs > saved_outputs) {
    std::fprintf(stderr, "[GPU Packing Online] unsupported arguments\n");
    return;
  }
  hipStream_t s = query_stream();
  launch_packing_materialized(d_S, d_bold_t, d_bold_t_bar, d_bold_t_hat, d_y_all, d_y_bar_all, d_z_body,
                              num_outputs, s);
  LEGACY_CHECK("GPU Packing Online", hipGetLastError());
  std::vector<uint32_t> tmp((size_t)num_outputs * 2 * kN);
  LEGACY_CHECK("GPU Packing Online",
               hipMemcpyAsync(tmp.data(), d_S, tmp.size() * 4, hipMemcpyDeviceToHost, s));
  LEGACY_CHECK("GPU Packing Online", hipStreamSynchronize(s));
  for (size_t i = 0; i < tmp.size(); ++i) h_output[i] = tmp[i];
}

extern "C" void gpu_packing_online_free(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  free_precomp();
  free_keys();
  free_ptr(d_S);
  d_S = nullptr;
  saved_outputs = 0;
  saved_n_inner = 0;
}
