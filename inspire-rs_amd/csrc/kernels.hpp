// Launch wrappers for the gfx950 kernels (kernels.hip).  All pointers are
// device pointers; every launch is asynchronous on stream `s`.
#pragma once
#include <cstddef>
#include <cstdint>
#include <hip/hip_runtime.h>

namespace inspire {

enum SynthKind { kSynthCondensed = 0, kSynthModQ = 1, kSynthU16 = 2, kSynthCrt = 3, kSynthSmall = 4 };

void launch_synth_fill(void* out, uint64_t seed, uint32_t stream, uint64_t start, uint64_t count, int kind,
                       hipStream_t s);

void launch_unpack_query(const uint8_t* q, int nq, int db_rows, uint64_t* y_body, uint64_t* z_body,
                         uint64_t* rows, uint32_t* bodies, hipStream_t s);

__global__ void k_zero_u32(uint32_t* p, int n);

// Batched unpack: client b's query at q + b*q_stride; outputs per client
// y/z [3][N], rows [nq][db_rows], bodies [nq][6][2][N].
void launch_unpack_query_batch(const uint8_t* q, int64_t q_stride, int nq, int db_rows, int count, uint64_t* y_body,
                               uint64_t* z_body, uint64_t* rows, uint32_t* bodies, hipStream_t s);
// Fold for `clients` keyword queries; client b's responses at out + b*out_stride.
// fold_tab: DeviceConstants::fold.  ahat_ntt: launch_ahat_ntt's table of
// a_hat, or nullptr (the fold then transforms all six gadget digits per step).
int launch_fold_batch(uint8_t* out, int64_t out_stride, const uint64_t* a_hat, const uint32_t* ahat_ntt,
                      const uint64_t* cvec, const uint64_t* P, const uint32_t* bodies, const uint32_t* fold_tab,
                      const uint32_t* tw, int64_t instances, int d, int nq, int clients, hipStream_t s,
                      int variant = 0);
// ahat_ntt[j][c][z] = n^-1 NTT_c(a_hat_j mod q_c) at NTT slot z, for the
// fold's row-0 shortcut ([instances][2][N] u32).
void launch_ahat_ntt(uint32_t* ahat_ntt, const uint64_t* a_hat, const uint32_t* tw, int64_t instances,
                     hipStream_t s);

// DB encode (interpolation of each row's column groups, p = 65535): in and out
// column-major [db_cols][db_rows] (out may equal in); work holds
// groups_per_batch column groups of work units (encode_work_bytes_per_group
// bytes each).  d a power of two <= 32, db_rows % 64 == 0, db_cols % (d*2048)
// == 0; returns -1 otherwise.
size_t encode_work_bytes_per_group(int64_t db_rows, int d);
int launch_encode_db(uint16_t* out, const uint16_t* in, int64_t db_rows, int64_t db_cols, int d, uint32_t num_inv,
                     uint16_t* work, int64_t groups_per_batch, hipStream_t s);

// Fragment-ordered copy of the DB (signed byte planes in MFMA B-operand
// order, tiled_db_bytes bytes); needs db_rows % 64 == 0 and db_cols % 16 == 0.
size_t tiled_db_bytes(int db_rows, int64_t db_cols);
int launch_tile_db(void* out, const uint16_t* db, int db_rows, int64_t db_cols, hipStream_t s);

// Exact int8-limb MFMA GEMV for nq <= 2 hash queries; scratch holds the query
// fragments (gemv_mfma_scratch_bytes(db_rows) bytes, 16-byte aligned).  Reads
// dbfrag (launch_tile_db) when given, else the column-major DB.
size_t gemv_mfma_scratch_bytes(int db_rows);
int launch_gemv_mfma(uint64_t* c, const uint64_t* rows, int nq, const uint16_t* db, const void* dbfrag,
                     int db_rows, int64_t db_cols, void* scratch, hipStream_t s);
// The same in two halves: the query-fragment build into scratch (which
// depends on the query alone), then the DB stream that reads it.
int launch_gemv_mfma_prep(const uint64_t* rows, int nq, int db_rows, void* scratch, hipStream_t s);
int launch_gemv_mfma_main(uint64_t* c, int nq, const uint16_t* db, const void* dbfrag, int db_rows, int64_t db_cols,
                          const void* scratch, hipStream_t s);

// Batched: `count` keyword queries of nq <= 2 rows each; c[client][h][db_cols].
size_t gemv_mfma_batch_scratch_bytes(int db_rows, int count);
int launch_gemv_mfma_batch(uint64_t* c, int64_t c_client_stride, const uint64_t* rows, int64_t rows_client_stride,
                           int nq, const uint16_t* db, const void* dbfrag, int db_rows, int64_t db_cols, int count,
                           void* scratch, hipStream_t s, int variant = 0);
int gemv_batch_variants();
const char* gemv_batch_variant_name(int v);

// Tiled int8-limb key-switch material for the MFMA packing (packing_mfma.hip),
// rotation terms in key-slot order per NTT slot; dlog: DeviceConstants::slot_dlog.
size_t tiled_bold_t_bytes(int64_t instances);
void launch_tile_bold_t(uint8_t* out, const uint64_t* bt, const uint64_t* btb, const uint16_t* dlog,
                        int64_t instances, hipStream_t s);
// The same copy of the synthetic state of `seed` (what launch_synth_fill writes
// to bold_t / bold_t_bar with these stream ids), without the condensed copy.
void launch_tile_bold_t_synth(uint8_t* out, uint64_t seed, uint32_t stream_bt, uint32_t stream_btb,
                              const uint16_t* dlog, int64_t instances, hipStream_t s);
int batch_key_pad(int count);
// Device bytes of the client-key fragments and limb sums for up to `count` clients.
size_t packing_afrag_bytes(int count);
// Rotation-term sums of the packing, z-major: S[client][2][N][instances];
// y_body: [count][3][N] (key_client_stride elements apart); afrag: scratch of
// packing_afrag_bytes(count) bytes.  k_post_ntt adds the z_body terms.
int launch_packing_mfma(uint32_t* S, int64_t s_client_stride, const uint8_t* tiled, uint8_t* afrag,
                        const uint64_t* y_body, int64_t key_client_stride, int count, int64_t instances,
                        hipStream_t s, int variant = 0);
// The same in two halves: the client-key fragments and limb sums into afrag
// (from the keys alone), then the key-switch GEMM that reads them.
int launch_packing_mfma_prep(uint8_t* afrag, const uint64_t* y_body, int64_t key_client_stride, int count,
                             hipStream_t s);
int launch_packing_mfma_main(uint32_t* S, int64_t s_client_stride, const uint8_t* tiled, const uint8_t* afrag,
                             int count, int64_t instances, hipStream_t s, int variant = 0);
// Tile configurations of the 64-client-group packing kernel (variant index).
int packing_wide_variants();
const char* packing_wide_variant_name(int v);

// c: [nq][db_cols] canonical mod q; rows: [nq][db_rows] condensed; db: [db_cols][db_rows].
int launch_gemv(uint64_t* c, const uint64_t* rows, int nq, const uint16_t* db, int db_rows, int64_t db_cols,
                hipStream_t s);

// S: [instances][2][N] u32 canonical residues.
void launch_packing(uint32_t* S, const uint64_t* bt, const uint64_t* btb, const uint64_t* bth,
                    const uint64_t* y_body, const uint64_t* z_body, int instances, hipStream_t s);
void launch_packing_materialized(uint32_t* S, const uint64_t* bt, const uint64_t* btb, const uint64_t* bth,
                                 const uint64_t* y_all, const uint64_t* y_bar_all, const uint64_t* z_body,
                                 int instances, hipStream_t s);
void launch_rotation_legacy(uint64_t* y_all, uint64_t* y_bar_all, const uint64_t* y_body, const uint32_t* tables,
                            const uint32_t* gen_pows, int num_rotations, int t_exp, int poly_len,
                            hipStream_t s);

// P: [count][N] canonical mod q from S [count][2][N]; with ntt_out, the
// fold's form instead: n^-1 NTT_c(P_j) as two u32 rows [2][N] in the 8N bytes
// of row j.
void launch_post(uint64_t* P, const uint32_t* S, const uint32_t* tw, int count, hipStream_t s, int ntt_out = 0);
// The fold's form of P ([clients][instances] rows of two u32 rows [2][N],
// n^-1 NTT_c(P_j)) from the z-major rotation-term sums S
// [clients][2][N][instances] plus the z_body terms sum_k z_body[b][k] *
// bold_t_hat[j][k] (z_body: [clients][3][N], key_client_stride apart).
void launch_post_ntt(uint64_t* P, const uint32_t* S, const uint64_t* z_body, int64_t key_client_stride,
                     const uint64_t* bth, int clients, int64_t instances, hipStream_t s);


// Offline precomputation (setup.hip): bold_t, bold_t_bar, bold_t_hat (condensed)
// and a_hat of every instance from the encoded DB, `group` instances per pass
// with setup_offline_bytes_per_instance() bytes of work memory each.  Returns
// an INSPIRE_* status (the message set on failure); synchronous on s.
size_t setup_offline_bytes_per_instance();
// Instances per pass that keep every collapse chain of the pass resident (0 on
// a HIP failure).
int64_t setup_offline_resident_instances();
int launch_setup_offline(const uint16_t* db, int64_t db_rows, int64_t instances, uint64_t* bt, uint64_t* btb,
                         uint64_t* bth, uint64_t* a_hat, const uint32_t* tw, const uint32_t* fold_tab, int64_t group,
                         hipStream_t s);

size_t fold_lds_bytes(bool table);
// out: [nq][instances/d][12288] response blobs.
int launch_fold(uint8_t* out, const uint64_t* a_hat, const uint32_t* ahat_ntt, const uint64_t* cvec,
                const uint64_t* P, const uint32_t* bodies, const uint32_t* fold_tab, const uint32_t* tw,
                int64_t instances, int d, int nq, hipStream_t s);

}  // namespace inspire
