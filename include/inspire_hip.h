/*
 * inspire_hip.h -- C ABI of the MI355X (gfx950) keyword-PIR answer path.
 *
 * Two groups of entry points:
 *
 *  1. The operator surface the reference's Rust FFI binds today
 *     (src/pir/gpu/{gemv,packing_online,encode}.rs extern "C" blocks,
 *     implemented in the native/cuda sources).  Same names, argument meaning and
 *     void/NULL error behaviour, so the Rust wrappers link unchanged.
 *
 *  2. Handle-based entry points for the parts of KeywordServer::respond
 *     (src/pir/server/keyword.rs:534-628) that the reference runs on the CPU
 *     (post-process, RGSW fold, modulus switch, response assembly) and for the
 *     whole respond path on device.  These return an int status
 *     (INSPIRE_OK = 0, negative on error; inspire_last_error() explains).
 *
 * All sizes follow the reference: n = poly_len = 2048, two CRT primes
 * q0 = 268369921, q1 = 249561089 (src/pir/params.rs:9), t_exp_left = 3,
 * "condensed" u64 = (x mod q0) | (x mod q1) << 32 (src/pir/packing.rs:867-884).
 */
#ifndef INSPIRE_HIP_H
#define INSPIRE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */
#define INSPIRE_OK 0
#define INSPIRE_ERR_INVALID (-1)  /* bad argument, shape or malformed query */
#define INSPIRE_ERR_HIP (-2)      /* HIP runtime error */
#define INSPIRE_ERR_NOMEM (-3)    /* device allocation failed */
#define INSPIRE_ERR_STATE (-4)    /* required server state not loaded */
#define INSPIRE_ERR_CAPACITY (-5) /* caller's output buffer too small */

/* Message for the last error on the calling thread ("" if none). */
const char* inspire_last_error(void);

/* ======================================================================== */
/* 1. Reference operator surface (native/cuda replacement)                  */
/* ======================================================================== */

/* encode_kernel.cu:233-255 (gpu_init_memory_pool): creates the query stream
 * (default priority) and warms the allocator.  Idempotent. */
void gpu_init_memory_pool(void);
/* encode_kernel.cu:258-260: free device memory in bytes (0 on failure). */
uint64_t gpu_available_memory_bytes(void);

/* encode.rs:7-18 / encode_kernel.cu:386-424 (gpu_encode_db_from_host):
 * encodes a raw host DB (u16, column-major [db_cols][db_rows], cuckoo.rs:
 * 314-353) by polynomial interpolation mod `modulus` (65535) and keeps it on
 * the device for gpu_encode_get_device_db; h_db_out is not written (the
 * reference leaves it untouched too).  gamma must be 2048, interp_degree a
 * power of two <= 32, num_inv = interp_degree^-1 mod modulus. */
void gpu_encode_db_from_host(const uint16_t* h_db_in, uint16_t* h_db_out, int db_rows, int db_cols,
                             int interp_degree, int gamma, int c, uint32_t modulus, uint32_t num_inv);
/* encode.rs:16 / encode_kernel.cu:427-430: the device DB of the last encode (NULL if none). */
const uint16_t* gpu_encode_get_device_db(void);
/* encode.rs:17 / encode_kernel.cu:432-436. */
void gpu_encode_free_device_db(void);
/* encode.rs:19-28 / encode_kernel.cu:439-494: the same encode into a second
 * slot, leaving the primary DB serving. */
void gpu_encode_db_from_host_fold(const uint16_t* h_db_in, int db_rows, int db_cols, int interp_degree, int gamma,
                                  int c, uint32_t modulus, uint32_t num_inv);
/* encode.rs:29 / encode_kernel.cu:497-499. */
const uint16_t* gpu_encode_get_device_db_fold(void);
/* encode.rs:30 / encode_kernel.cu:502-510: free the primary, promote the second slot. */
void gpu_encode_swap_primary(void);
/* encode.rs:5 / encode_kernel.cu:233-238: waits for the device (no pooled
 * allocator to trim here). */
void gpu_trim_memory_pool(void);

/* gemv.rs:6 / gemv_kernel.cu:116-136: copy a host DB (u16, column-major
 * [db_cols][db_rows]) to the device; *d_db_out = NULL on failure. */
void gpu_gemv_upload_db(const uint16_t* h_db, uint16_t** d_db_out, size_t db_size_bytes);
/* gemv.rs:7 / gemv_kernel.cu:107-113: adopt an existing device DB pointer. */
void gpu_gemv_set_device_db(uint16_t** d_db_out, const uint16_t* d_db_in, size_t db_size_bytes);
/* gemv.rs:8 / gemv_kernel.cu:139-141. */
void gpu_gemv_free_db(uint16_t* d_db);
/* gemv.rs:9-24 / gemv_kernel.cu:144-194: h_c[j] = sum_k h_a[k]*db[j][k] mod q,
 * synchronous, host in/out.  The modulus arguments must be the reference's
 * constants (they are checked; a mismatch leaves h_c untouched). */
void gpu_gemv_query(uint64_t* h_c, const uint64_t* h_a, const uint16_t* d_db, int db_rows, int b_cols,
                    uint64_t mod0, uint64_t mod1, uint64_t cr0, uint64_t cr1, uint64_t mod0_inv_mod1,
                    uint64_t mod1_inv_mod0, uint64_t modulus, uint64_t barrett_cr0_mod,
                    uint64_t barrett_cr1_mod);

/* packing_online.rs:23-29 / rotation_kernel.cu:71-94. */
void gpu_rotation_upload_tables(const uint32_t* h_tables, const uint32_t* h_gen_pows, int num_tables,
                                int num_rotations, int poly_len);
/* packing_online.rs:31-38 / rotation_kernel.cu:99-137. */
void gpu_rotation_expand_keys(uint64_t* d_y_all_out, uint64_t* d_y_bar_all_out, const uint64_t* h_y_body,
                              int num_rotations, int t_exp_left, int poly_len);
/* packing_online.rs:40 / rotation_kernel.cu:140-144. */
void gpu_rotation_free(void);

/* packing_online.rs:42-47 / packing_online_kernel.cu:176-211. */
void gpu_packing_online_ensure_key_buffers(int n_inner, int t_exp_left, int poly_len);
uint64_t* gpu_packing_online_get_d_y_all(void);
uint64_t* gpu_packing_online_get_d_y_bar_all(void);
void gpu_packing_online_upload_z_body(const uint64_t* h_z_body, int t_exp_left, int poly_len);
/* packing_online.rs:49-57 / packing_online_kernel.cu:110-143. */
void gpu_packing_online_upload_precomp(const uint64_t* h_bold_t, const uint64_t* h_bold_t_bar,
                                       const uint64_t* h_bold_t_hat, int num_outputs, int n_inner,
                                       int t_exp_left, int poly_len);
/* packing_online.rs:13-21 / packing_online_kernel.cu:258-288 (borrowed pointers). */
void gpu_packing_online_adopt_precomp(uint64_t* d_bold_t, uint64_t* d_bold_t_bar, uint64_t* d_bold_t_hat,
                                      int num_outputs, int n_inner, int t_exp_left, int poly_len);
/* packing_online.rs:59-66 / packing_online_kernel.cu:147-172. */
void gpu_packing_online_upload_keys(const uint64_t* h_y_all, const uint64_t* h_y_bar_all,
                                    const uint64_t* h_z_body, int n_inner, int t_exp_left, int poly_len);
/* packing_online.rs:68-79 / packing_online_kernel.cu:215-251: synchronous;
 * h_output [num_outputs][2][poly_len] (CRT0 slots, then CRT1 slots). */
void gpu_packing_online_compute(uint64_t* h_output, int num_outputs, int nphalf_minus_1, int t_exp_left,
                                int poly_len, int addition_capacity, uint64_t mod0, uint64_t mod1,
                                uint64_t cr0, uint64_t cr1);
/* packing_online.rs:83-91 / packing_online_kernel.cu:310-333. */
void gpu_packing_online_swap_precomp(uint64_t* new_bt, uint64_t* new_bt_bar, uint64_t* new_bt_hat,
                                     int num_outputs, int n_inner, int t_exp_left, int poly_len);
/* packing_online.rs:81 / packing_online_kernel.cu:291-306. */
void gpu_packing_online_free(void);

/* ======================================================================== */
/* 2. Handle-based answer path                                              */
/* ======================================================================== */

/* KeywordServerSetupPlan + KeywordServerGpuMemoryEstimate
 * (keyword.rs:91-233; parameters from protocol/wire.rs:386-477). */
typedef struct {
  int64_t num_items;          /* cuckoo buckets */
  int64_t item_size_bits;     /* 8 * (key_size + value_size) */
  int64_t dim0;
  int64_t db_rows;
  int64_t db_cols;
  int64_t instances;          /* db_cols_prime = packed RLWE outputs */
  int64_t interpolate_degree; /* d */
  int64_t c;                  /* fold chains = instances / d */
  uint64_t db_bytes, hint_bytes, prep_pack_bytes, r_buffers_bytes, bold_t_bytes, peak_bytes;
} inspire_plan_t;

/* keyword.rs:145-233 from_shape.  buckets / dim0 <= 0 select the reference
 * defaults.  On INSPIRE_ERR_INVALID the message (inspire_last_error) is the
 * reference's error text. */
int inspire_plan_shape(int64_t record_count, int64_t key_size, int64_t value_size, int64_t buckets,
                       int64_t dim0, inspire_plan_t* out);

typedef struct {
  int64_t db_rows;            /* 2^(db_dim_1 + 11), multiple of 2048 */
  int64_t instances;          /* db_cols / 2048 */
  int32_t interpolate_degree; /* d, divides instances */
  int32_t max_queries;        /* hash queries per keyword query (cuckoo num_hashes = 2) */
} inspire_shape_t;

typedef struct inspire_server inspire_server_t;

/* Allocates device state and per-request buffers on the current device,
 * including a fragment-ordered copy of the database (the database's size)
 * that the MFMA GEMVs read; if that copy cannot be allocated they read the
 * column-major database. */
int inspire_server_create(const inspire_shape_t* shape, inspire_server_t** out);
/* A server that keeps only the fragment-ordered copy of bold_t / bold_t_bar
 * (the MFMA packing's operand, the same size as the condensed pair) and not
 * the condensed pair itself: half the key-switch memory, so that the 4M-key
 * ML-KEM-512 state beside a 250k-key ML-KEM-1024 state, or a 32M-key state
 * at dim0 65536, fit on one GPU.  Only the MFMA packing runs
 * (INSPIRE_PACKING_VALU is refused).  The setters build the fragments from a
 * transient condensed copy (set_precomp, setup_offline: bold_t's size again
 * for the duration of the call) or straight from the generator
 * (fill_synthetic). */
int inspire_server_create_lean(const inspire_shape_t* shape, inspire_server_t** out);
void inspire_server_destroy(inspire_server_t* srv);

/* Encoded DB, column-major [db_cols][db_rows] u16 (cuckoo.rs:314-353 after
 * encode_kernel.cu).  host != 0: copy from host memory; host == 0: copy from
 * a device pointer. */
int inspire_server_set_db(inspire_server_t* srv, const uint16_t* db, int host);
/* Raw DB, column-major [db_cols][db_rows] u16 (cuckoo.rs:314-353), encoded on
 * the device (encode_kernel.cu:263-380 with interp_degree = the shape's d,
 * modulus 65535) into the server's DB: keyword.rs:404-424 in one call. */
int inspire_server_encode_db(inspire_server_t* srv, const uint16_t* raw, int host);
/* InspiRING precomputation (PrecompInsPIR, packing.rs:1718-1736) in the
 * layout of gpu_packing_online_upload_precomp: bold_t, bold_t_bar
 * [instances][3069][2048] condensed, bold_t_hat [instances][3][2048]
 * condensed, a_hat [instances][2048] raw mod q. */
int inspire_server_set_precomp(inspire_server_t* srv, const uint64_t* bold_t, const uint64_t* bold_t_bar,
                               const uint64_t* bold_t_hat, const uint64_t* a_hat, int host);
/* KeywordServer::setup's offline precomputation on the device
 * (keyword.rs:427-458 -> perform_offline_precomputation_simplepir,
 * server/engine.rs:678-1000 with the GPU prep-pack and collapse of
 * hint_kernel.cu, prep_pack_kernel.cu, packing_kernel.cu and
 * collapse_kernel.cu): from the server's encoded DB, computes the hint over the
 * SEED_0 query masks, prep-pack, and full_packing_with_preprocessing_offline
 * (packing.rs:2477-2750) with the W_SEED / V_SEED packing masks, and installs
 * bold_t, bold_t_bar, bold_t_hat and a_hat as inspire_server_set_precomp
 * would.  Needs the DB (set_db or encode_db) first.  Synchronous. */
int inspire_server_setup_offline(inspire_server_t* srv);
/* Exchanges the resident state (DB, precomputation, fragment copies, stash)
 * of two servers of the same shape on the same device, after waiting for
 * their device work; the live-update step without an engine.  Calls on
 * either handle must not overlap it. */
int inspire_server_swap_state(inspire_server_t* live, inspire_server_t* staged);
/* Cuckoo stash entries returned with every response (keyword.rs:624-627). */
int inspire_server_set_stash(inspire_server_t* srv, const uint8_t* entries, uint32_t entry_size,
                             uint32_t count);
/* Fills DB and precomputation on the device from the counter-based synthetic
 * generator (benchmarks and tests; see inspire_synth_fill). */
int inspire_server_fill_synthetic(inspire_server_t* srv, uint64_t seed);

/* Serialized response size for a query with nq hash queries. */
size_t inspire_server_response_bytes(const inspire_server_t* srv, int nq);

/* KeywordServer::respond_to_serialized_query (keyword.rs:636-640):
 * deserialize (wire.rs:618-705), respond, serialize (wire.rs:708-726). */
int inspire_respond_serialized(inspire_server_t* srv, const uint8_t* query, size_t len, uint8_t* resp,
                               size_t cap, size_t* resp_len);

/* Device-resident respond: d_query holds a serialized query in device memory
 * (len bytes, nq = its first byte, passed by the caller to avoid a read-back),
 * d_out receives nq * (instances/d) * 12288 response bytes.  Asynchronous on
 * `stream` (hipStream_t; NULL = the server's own stream). */
int inspire_respond_device(inspire_server_t* srv, const uint8_t* d_query, size_t len, int nq, uint8_t* d_out,
                           void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): inspire_respond_device with the request's working
 * buffers (unpacked query, packing and GEMV outputs, fragments) in one
 * file-scope device array of INSPIRE_NOALLOC_SCRATCH_BYTES per GPU instead of
 * the server's request slots.  Same arguments, same response bytes; the call
 * allocates nothing (no hipMalloc, hipFree or allocating library routine).
 * A request needing more than the array (inspire_noalloc_scratch_bytes; at
 * 2048 rows and 2 hash queries about 64 KiB per packed output, so up to the
 * 1024 outputs of a 4M-key ML-KEM-512 database) is refused with
 * INSPIRE_ERR_INVALID.  Every call on one GPU shares that GPU's array, which
 * keeps the last call's contents: calls that overlap on one GPU (other
 * streams, other servers, other threads) race, so the caller serializes them,
 * e.g. on one stream.  The array stays allocated on each GPU the library has
 * run on until the process ends. */
#define INSPIRE_NOALLOC_SCRATCH_BYTES 100663296u /* 96 MiB */
int inspire_respond_device_noalloc(inspire_server_t* srv, const uint8_t* d_query, size_t len, int nq,
                                   uint8_t* d_out, void* stream);
/* Scratch bytes inspire_respond_device_noalloc needs for a shape and nq hash
 * queries (0 for a bad argument).  Host only. */
size_t inspire_noalloc_scratch_bytes(const inspire_shape_t* shape, int nq);

/* Number of request slots (default 2, at most 16).  Consecutive respond calls
 * use the slots in turn, so requests issued on different streams overlap on
 * the device; a slot is reused only after its previous request completed. */
int inspire_server_set_slots(inspire_server_t* srv, int count);

/* Packing inner-product kernel of the single-query path: INSPIRE_PACKING_VALU
 * (default; reads bold_t in the reference's condensed layout) or
 * INSPIRE_PACKING_MFMA (int8-limb MFMA over a fragment-ordered copy of bold_t
 * that this call builds on the device, 8 bytes per coefficient). */
#define INSPIRE_PACKING_VALU 0
#define INSPIRE_PACKING_MFMA 1
int inspire_server_set_packing_kernel(inspire_server_t* srv, int kind);

/* Batched respond.  set_max_batch allocates buffers for up to max_batch
 * keyword queries per call (and builds the fragment-ordered bold_t).
 * respond_batch_device answers `count` keyword queries in one pass over the
 * database and the key-switch material: query i is serialized at
 * d_queries + i*q_stride (len bytes each, nq hash queries each), its
 * responses go to d_out + i*out_stride (nq*(instances/d)*12288 bytes), each
 * identical to what inspire_respond_device returns for that query alone.
 * Asynchronous on `stream` (NULL = the server's stream); a second batch waits
 * for the first one's buffers. */
int inspire_server_set_max_batch(inspire_server_t* srv, int max_batch);
/* Batched responds run as HIP graphs (default on): the launch sequence of a
 * (count, nq, query and output buffers, stream, kernel variants) key is
 * captured on its second use and replayed afterwards; up to 256 keys are kept.
 * Off: every call launches its kernels directly.  Results are identical. */
int inspire_server_set_graphs(inspire_server_t* srv, int on);
int inspire_respond_batch_device(inspire_server_t* srv, const uint8_t* d_queries, size_t q_stride, size_t len,
                                 int nq, int count, uint8_t* d_out, size_t out_stride, void* stream);

/* Batched kernel variants.  packing_variant: tile configuration of the
 * 64-client-group packing kernel (0..4); gemv_variant: 0 = query and DB
 * fragments shared through LDS, 1 = query fragments from L2 (0..1, when the
 * server holds the fragment-ordered DB); fold_variant: 0 = by launch size,
 * 1 = one 1024-thread workgroup per chain (both primes at once), 2 = 512
 * threads, primes in turn.  All variants give identical results. */
int inspire_server_set_batch_kernels(inspire_server_t* srv, int packing_variant, int gemv_variant, int fold_variant);

/* Autotuner.  Times every batched kernel variant on this device for `count`
 * keyword queries of nq hash queries, each inside the whole stage sequence
 * (packing, post, GEMV, fold; reps runs after a warm-up), stage by stage:
 * the packing variants first, then the GEMV variants with the chosen packing,
 * then the fold kernels.  Keeps the fastest for the server's batched respond
 * (the *_ms fields are those sequence times), then measures
 * the batch stages' throughput at batch sizes 1, 2, 4, ... up to max_batch and
 * suggests the smallest batch within 10% of the best.  Needs set_max_batch
 * >= count; fills the batch buffers with random residues first (the device
 * runs faster on zero operands than on real queries) and uses them, so no
 * other call on the server may run meanwhile. */
typedef struct {
  int packing_variant, gemv_variant, fold_variant; /* chosen */
  int packing_variants, gemv_variants;             /* variants timed */
  double packing_ms[16], gemv_ms[8], fold_ms[3];   /* per variant (fold: [1], [2]) */
  int nbatch;
  int batch_sizes[16];
  double batch_qps[16]; /* keyword queries per second of the batch stages */
  int suggested_batch;
} inspire_tune_t;
int inspire_server_autotune(inspire_server_t* srv, int count, int nq, int reps, inspire_tune_t* out);

/* First-pass kernel used by the server: INSPIRE_GEMV_MFMA (default, exact
 * int8-limb MFMA) or INSPIRE_GEMV_VALU (64-bit multiply-accumulate). */
#define INSPIRE_GEMV_VALU 0
#define INSPIRE_GEMV_MFMA 1
int inspire_server_set_gemv_kernel(inspire_server_t* srv, int kind);

/* The server's stream (hipStream_t). */
void* inspire_server_stream(inspire_server_t* srv);

/* Per-stage GPU timing of the respond pipeline: hipEvents recorded between
 * the kernels on the respond stream.  While enabled, every respond waits for
 * its own completion and adds its stage times.  profile_read returns the
 * summed milliseconds per stage [unpack, packing, post, gemv, fold] and the
 * number of profiled responds; reset != 0 clears them. */
int inspire_server_profile_enable(inspire_server_t* srv, int enable);
int inspire_server_profile_read(inspire_server_t* srv, double* ms5, int64_t* count, int reset);

/* --- Stage entry points (device pointers, asynchronous on `stream`) ------ */

/* Counter-based synthetic fill: kind 0 condensed residues, 1 residues mod q,
 * 2 u16 in [0, 65534], 3 per-CRT residues of [..][2][2048] polys, 4 small
 * residues (< 2^20, as from a client with sk = 0). */
int inspire_synth_fill(void* d_out, uint64_t seed, uint32_t stream_id, uint64_t start, uint64_t count, int kind,
                       void* stream);
/* wire.rs:618-705 on device: y_body, z_body [3][2048] condensed; rows
 * [nq][db_rows] condensed; bodies [nq][6][2][2048] u32. */
int inspire_unpack_query_device(const uint8_t* d_query, int nq, int64_t db_rows, uint64_t* d_y_body,
                                uint64_t* d_z_body, uint64_t* d_rows, uint32_t* d_bodies, void* stream);
/* encode_kernel.cu:263-380 (gpu_encode_db) on device buffers: d_out and d_in
 * column-major [db_cols][db_rows] u16 (d_out may equal d_in); interp_degree a
 * power of two <= 32; db_cols a multiple of interp_degree * 2048, db_rows of
 * 64.  Allocates its work buffer (up to the DB's size) and returns when done. */
/* The offline precomputation of inspire_server_setup_offline on caller-owned
 * device buffers: d_db encoded [instances*2048][db_rows] u16; outputs in the
 * layouts of inspire_server_set_precomp.  Work memory: 64 MiB per instance of
 * a pass; a pass takes as many instances as fit in free memory less 1 GiB, at
 * most as many as keep all their collapse chains resident (384 on 256 CUs).
 * Synchronous on `stream` (NULL: the null stream). */
int inspire_setup_offline_device(const uint16_t* d_db, int64_t db_rows, int64_t instances, uint64_t* d_bold_t,
                                 uint64_t* d_bold_t_bar, uint64_t* d_bold_t_hat, uint64_t* d_a_hat, void* stream);
int inspire_encode_db_device(uint16_t* d_out, const uint16_t* d_in, int64_t db_rows, int64_t db_cols,
                             int interp_degree, void* stream);
/* gemv_kernel.cu:46-103 for nq query rows at once: d_c [nq][db_cols]. */
int inspire_gemv_device(uint64_t* d_c, const uint64_t* d_rows, int nq, const uint16_t* d_db, int64_t db_rows,
                        int64_t db_cols, void* stream);
/* The same product on int8 MFMA for nq <= 2; d_scratch holds
 * inspire_gemv_mfma_scratch_bytes(db_rows) bytes (query fragments). */
int inspire_gemv_mfma_device(uint64_t* d_c, const uint64_t* d_rows, int nq, const uint16_t* d_db, int64_t db_rows,
                             int64_t db_cols, void* d_scratch, void* stream);
size_t inspire_gemv_mfma_scratch_bytes(int64_t db_rows);
/* packing_online_kernel.cu:18-92 with the rotations fused:
 * d_S [instances][2][2048] u32 canonical. */
int inspire_packing_device(uint32_t* d_S, const uint64_t* d_bold_t, const uint64_t* d_bold_t_bar,
                           const uint64_t* d_bold_t_hat, const uint64_t* d_y_body, const uint64_t* d_z_body,
                           int64_t instances, void* stream);
/* packing_online.rs:359-398: d_P [count][2048] = CRT(INTT(S)) mod q. */
int inspire_post_device(uint64_t* d_P, const uint32_t* d_S, int64_t count, void* stream);
/* keyword.rs:543-581 + modulus_switch.rs:45-73 for nq hash queries:
 * d_out [nq][instances/d][12288]. */
int inspire_fold_device(uint8_t* d_out, const uint64_t* d_a_hat, const uint64_t* d_c, const uint64_t* d_P,
                        const uint32_t* d_bodies, int64_t instances, int d, int nq, void* stream);

/* ---------------------------------------------------------------------- */
/* Keyword layer: PQKEYV2 datasets, cuckoo table, server setup             */
/* ---------------------------------------------------------------------- */

/* PublicParams with its KeywordPirHandshake (protocol/wire.rs, returned by
 * KeywordServer::public_params, server/keyword.rs:505-531). */
typedef struct {
  int64_t num_items;       /* cuckoo buckets */
  int64_t item_size_bits;  /* 8 * entry_size */
  int64_t dim0;
  uint8_t cuckoo_seed[16]; /* DETERMINISTIC_SEED (cuckoo.rs:27-29) */
  int32_t num_hashes;      /* 2 */
  int64_t key_size, value_size, entry_size;
  int64_t num_records;     /* occupied buckets + stash entries */
} inspire_public_params_t;

/* setup_keyword_server (dataset/keyword.rs:29-44 -> KeywordServer::setup,
 * server/keyword.rs:366-468): loads a PQKEYV2 dataset (dataset.rs:95-142:
 * magic "PQKEYV2\0", u8 kem-name length + name, u32 LE count, u16 LE key
 * length, then count x (32-byte id, key)), builds the cuckoo table (key = id,
 * value = public key, SipHash-2-4 positions, cuckoo.rs:98-244), lays it out as
 * the column-major raw DB (cuckoo.rs:314-353), and creates a server with the
 * DB encoded and the offline precomputation done on the device and the stash
 * installed.  buckets / dim0 <= 0 select the reference defaults.  pp (may be
 * NULL) receives public_params.  The cuckoo eviction order is seeded here
 * (the reference uses OS randomness), so tables are repeatable. */
int inspire_setup_keyword_server(const char* pqkeyv2_path, int64_t buckets, int64_t dim0, inspire_server_t** out,
                                 inspire_public_params_t* pp);
/* KeywordServer::setup from records in memory: count keys of key_size bytes
 * and count values of value_size bytes, back to back. */
int inspire_setup_keyword_records(const uint8_t* keys, uint32_t key_size, const uint8_t* values,
                                  uint32_t value_size, int64_t count, int64_t buckets, int64_t dim0,
                                  inspire_server_t** out, inspire_public_params_t* pp);
/* CuckooHash::all_positions with the deterministic seed (cuckoo.rs:98-132):
 * the client's two bucket indices for a key.  Host only. */
int inspire_cuckoo_positions(const uint8_t* key, size_t key_len, int64_t num_buckets, int64_t* out2);
/* SipHash-2-4 of msg under (k0, k1), as siphasher's SipHasher::new_with_keys,
 * write, finish.  Host only. */
uint64_t inspire_siphash24(uint64_t k0, uint64_t k1, const uint8_t* msg, size_t len);

/* ---------------------------------------------------------------------- */
/* Keyword client (host only): KeywordClient, src/pir/client/keyword.rs    */
/* ---------------------------------------------------------------------- */

typedef struct inspire_client inspire_client_t;

/* KeywordClient::setup_from_public_params (client/keyword.rs:60-121) from the
 * server's public_params (the handshake), with the reference's error texts.
 * hamming = 0 uses the reference client's secret, the zero polynomial of
 * Client::init (client/base.rs:370-389); hamming = h in [1, 1024] draws a
 * ternary secret with h coefficients 1 and h coefficients -1
 * (gen_ternary_mat, client/base.rs:126-139).  Secret randomness (key and
 * noise) is ChaCha20 under a key from getrandom(2); rng_seed != 0 derives that
 * key from rng_seed instead (repeatable, for tests).  A client is not
 * thread-safe: query advances its generator. */
int inspire_client_new(const inspire_public_params_t* pp, int32_t hamming, uint64_t rng_seed,
                       inspire_client_t** out);
void inspire_client_free(inspire_client_t* client);
/* The secret key, [2048] raw coefficients mod q (Client::get_sk_reg). */
int inspire_client_secret(const inspire_client_t* client, uint64_t* sk);
/* KeywordClient::positions (client/keyword.rs:133-145): num_hashes bucket
 * indices under the handshake's cuckoo seed. */
int inspire_client_positions(const inspire_client_t* client, const uint8_t* key, size_t key_len,
                             int64_t* positions);
/* Bytes of one serialized query (num_hashes hash queries). */
size_t inspire_client_query_bytes(const inspire_client_t* client);
/* KeywordClient::query + serialize_query (client/keyword.rs:198-210,
 * wire.rs:542-615): fills positions[num_hashes] and writes the query;
 * *len = its size (INSPIRE_ERR_CAPACITY if cap is smaller). */
int inspire_client_query(inspire_client_t* client, const uint8_t* key, size_t key_len, int64_t* positions,
                         uint8_t* out, size_t cap, size_t* len);
/* KeywordClient::extract_from_serialized_response (client/keyword.rs:232-310):
 * *found = 1 and the value_size bytes of the value in value[] when a chain
 * blob or a stash entry holds the key, *found = 0 otherwise (the reference's
 * None).  A malformed response is INSPIRE_ERR_INVALID. */
int inspire_client_extract(const inspire_client_t* client, const uint8_t* key, size_t key_len,
                           const int64_t* positions, const uint8_t* response, size_t response_len,
                           uint8_t* value, size_t cap, size_t* value_len, int32_t* found);

/* ---------------------------------------------------------------------- */
/* Datasets: DatasetGenerator and the PQKEYV2 writer (host only)           */
/* ---------------------------------------------------------------------- */

/* DatasetGenerator::stream_to_writer + DatasetWriter (generator.rs:112-169,
 * writer.rs:17-79): writes a PQKEYV2 file of `count` records for kem_name
 * ("ml-kem-512", "ml-kem-768" or "ml-kem-1024"), record i = (id, public key)
 * with the public key of DatasetGenerator::generate_record(i) for `seed`
 * (generator.rs:170-186, kem.rs:27-44) and id = SHA3-256("pq-key:v1:" ||
 * kem_name || key) (dataset.rs:36-46).  threads <= 0: one per hardware
 * thread. */
int inspire_dataset_generate(const char* path, const char* kem_name, int64_t count, uint64_t seed, int threads);
/* The records start .. start+count-1 of the same stream into memory:
 * ids [count][32], keys [count][public key length]. */
int inspire_dataset_generate_records(const char* kem_name, uint64_t seed, int64_t start, int64_t count,
                                     uint8_t* ids, uint8_t* keys, int threads);
/* The 64-byte ML-KEM seed of record `index` (generator.rs:170-173:
 * ChaCha20Rng::seed_from_u64(seed) at word position 16 * index). */
int inspire_dataset_record_seed(uint64_t seed, uint64_t index, uint8_t out64[64]);
/* KemVariant::generate_public_key_into (kem.rs:27-44): the encapsulation key
 * of ML-KEM.KeyGen_internal(d = seed64[0..32], z = seed64[32..64]) (FIPS 203),
 * 800 / 1184 / 1568 bytes; ek_cap too small: INSPIRE_ERR_CAPACITY. */
int inspire_mlkem_keygen(const char* kem_name, const uint8_t seed64[64], uint8_t* ek, size_t ek_cap);
/* FIPS 202 hashes used above: kind 0 SHA3-256 (out_len 32), 1 SHA3-512
 * (out_len 64), 2 SHAKE128, 3 SHAKE256 (any out_len). */
int inspire_keccak(int kind, const uint8_t* in, size_t len, uint8_t* out, size_t out_len);

/* ------------------------------------------------------------------------
 * Continuous-batching engine (no counterpart in the reference, whose server
 * loop answers one query per round trip, server.rs:116-158).  Callers submit
 * serialized keyword queries (nq hash queries each) from any thread; a
 * dispatcher packs the pending ones, up to max_batch, into one batched respond,
 * waiting up to window_ms for more arrivals when fewer than max_batch are
 * pending (0: dispatch at once).  A second batch goes in flight (its upload
 * overlapping the running one) only when a full batch is pending.
 * max_batch is capped at creation so the batch buffers fit in the free device
 * memory (less 1 GiB); inspire_engine_stats reports the cap.  capacity bounds
 * the submitted-but-not-waited queries (submit returns INSPIRE_ERR_CAPACITY
 * past it).  Every ticket is waited exactly once; wait writes the serialized
 * response (stash included) and the submit-to-ready latency.  When cap is too
 * small, wait returns INSPIRE_ERR_CAPACITY with the size needed in *resp_len
 * and the ticket stays valid (after a state swap an answer keeps the stash of
 * the state it ran on, which may be larger than the current one).  The server
 * must not be used by other callers while an engine runs on it.
 * ---------------------------------------------------------------------- */
typedef struct inspire_engine inspire_engine_t;
int inspire_engine_create(inspire_server_t* srv, int nq, int max_batch, double window_ms, int capacity,
                          inspire_engine_t** out);
int inspire_engine_submit(inspire_engine_t* eng, const uint8_t* query, size_t len, uint64_t* ticket);
int inspire_engine_wait(inspire_engine_t* eng, uint64_t ticket, uint8_t* resp, size_t cap, size_t* resp_len,
                        double* latency_ms);
int inspire_engine_stats(inspire_engine_t* eng, int64_t* batches, int64_t* queries, int* max_batch);
/* Live DB update (the reference's fold-then-swap of the encoded DB and the
 * precomputation: encode_kernel.cu:453-516, packing_online_kernel.cu:310-333).
 * Build the new state on a second server of the same shape (set_db or
 * encode_db, set_precomp or setup_offline, set_stash), then swap: new batches
 * are held back, the running ones finish, the engine's server and `staged`
 * exchange their resident state, and the held queries run on the new state.
 * `staged` then holds the old state (destroy it, or reuse it for the next
 * update).  Queries are never answered from a mix of the two states. */
int inspire_engine_swap_state(inspire_engine_t* eng, inspire_server_t* staged);
/* Finishes the queued and running batches, then frees the engine. */
void inspire_engine_destroy(inspire_engine_t* eng);

#ifdef __cplusplus
}
#endif

#endif /* INSPIRE_HIP_H */
