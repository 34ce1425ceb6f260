// First-pass GEMV on int8 MFMA (exact):  c[h][col] = sum_k a_h[k] * db[col][k] mod q
// (gemv_kernel.cu:46-103, gpu/kernel.rs:139-237), for up to two hash queries
// per launch.
//
// Limb split.  Each query residue (28 bits, one per CRT prime) becomes four
// 7-bit limbs in [0, 127]; each u16 DB value v = 256*hi + lo becomes two
// signed bytes (lo - 128, hi - 128).  With R_l = sum_k limb_l(a[k]):
//   sum_k a[k]*v[k] = sum_l 2^(7l) * (256*S_hi,l + S_lo,l + 32896*R_l)
// where S_x,l = sum_k limb_l(a[k]) * (x[k] - 128) is one int32 MFMA
// accumulator (|S| <= 127*128*K < 2^31 for K <= 2^17).  The sum is the
// exact non-negative integer, < 2^55 for K <= 65536, reduced mod q_p last.
//
// MFMA tile (v_mfma_i32_16x16x64_i8): rows i = 8h + 4p + l (hash query h,
// prime p, limb l), columns = 16 DB columns, K = 64 DB rows per step.  Lane
// (j, g) = (lane % 16, lane / 16) feeds DB column j with the 16 rows
// {8g..8g+7} and {32+8g..32+8g+7} of the step, in that byte order; the query
// fragments use the same row order, so the contraction is over matching k.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "modarith.hpp"

namespace inspire {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaChunkRows = 2048;                        // rows staged per LDS chunk
constexpr int kMfmaStepsPerChunk = kMfmaChunkRows / 64;     // 32
#ifndef GEMV_TPW
#define GEMV_TPW 1
#endif
#ifndef GEMV_D
#define GEMV_D 8
#endif
constexpr int kTilesPerWave = GEMV_TPW;                      // column tiles of 16 per wave
constexpr int kGemvRing = GEMV_D;                            // DB steps in flight per wave
constexpr int kColsPerBlock = 4 * kTilesPerWave * 16;

// Row k of step s held by lane group g at byte t.
__host__ __device__ inline int frag_row(int s, int g, int t) { return 64 * s + 8 * g + t + (t >= 8 ? 24 : 0); }

// One column's dot product mod q_p from the lane's four limb rows of its lo-
// and hi-plane tiles (p = the lane group's prime):
//   X = sum_r 2^(7r) (256 S_hi,r + S_lo,r + 32896 R_r) = sum_k a_k v_k,
// every limb term non-negative and X < K 2^44 <= 2^61 for K <= 2^17.  With
// X = 2^32 X1 + X0 (X1 < 2^29), X1 (2^32 mod q) + X0 < 2^58 is one 32 x 32 +
// 64 multiply-add, then reduce58 with the lane's prime: one chain of 32-bit
// operations per lane.
struct LanePrime {
  uint32_t q, mu, c32;  // q_p, floor(2^58 / q_p), 2^32 mod q_p
  __device__ explicit LanePrime(int p) : q(prime(p)), mu(mu58(p)), c32(p ? kC32Q1 : kC32Q0) {}
};

__device__ inline uint32_t gemv_residue(const v4i& lo, const v4i& hi, const int64_t (&rs)[4], const LanePrime& lp) {
  int64_t v = 256 * (int64_t)hi[3] + lo[3] + rs[3];
#pragma unroll
  for (int r = 2; r >= 0; --r) v = (v << 7) + (256 * (int64_t)hi[r] + lo[r] + rs[r]);
  const uint64_t x = (uint64_t)v;
  return reduce58_rt((uint64_t)(uint32_t)(x >> 32) * lp.c32 + (uint32_t)x, lp.q, lp.mu);
}

// afrag: [db_rows/64][64 lanes][16 B]; rsum: [16] int32 (row i = 8h+4p+l).
// Called by every thread of a 1-D grid (blockIdx.x, threadIdx.x).
__device__ void k_gemv_prep_body(const uint64_t* __restrict__ rows, int nq, int db_rows, int8_t* __restrict__ afrag,
                                 int32_t* __restrict__ rsum) {
  __shared__ int32_t part[16];
  if (threadIdx.x < 16) part[threadIdx.x] = 0;
  __syncthreads();
  const int steps = db_rows / 64;
  int32_t local[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) local[i] = 0;
  // One thread per (step, lane): 16 bytes.
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < steps * 64; e += gridDim.x * blockDim.x) {
    const int s = e / 64, lane = e % 64;
    const int i = lane % 16, g = lane / 16;
    const int h = i >> 3, p = (i >> 2) & 1, l = i & 3;
    uint32_t w[4] = {0, 0, 0, 0};
    if (h < nq) {
      for (int t = 0; t < 16; ++t) {
        const uint64_t v = rows[(int64_t)h * db_rows + frag_row(s, g, t)];
        const uint32_t r = (uint32_t)(p ? (v >> 32) : v);
        const uint32_t limb = (r >> (7 * l)) & 127u;
        w[t >> 2] |= limb << (8 * (t & 3));
        local[i] += (int32_t)limb;
      }
    }
    reinterpret_cast<uint4*>(afrag)[e] = make_uint4(w[0], w[1], w[2], w[3]);
  }
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (local[i]) atomicAdd(&part[i], local[i]);
  __syncthreads();
  if (threadIdx.x < 16) atomicAdd(&rsum[threadIdx.x], part[threadIdx.x]);
}

__global__ void k_gemv_prep(const uint64_t* __restrict__ rows, int nq, int db_rows, int8_t* __restrict__ afrag,
                            int32_t* __restrict__ rsum) {
  k_gemv_prep_body(rows, nq, db_rows, afrag, rsum);
}

__device__ inline uint32_t plane_lo(uint32_t a, uint32_t b) {
  return __builtin_amdgcn_perm(b, a, 0x06040200u) ^ 0x80808080u;
}
__device__ inline uint32_t plane_hi(uint32_t a, uint32_t b) {
  return __builtin_amdgcn_perm(b, a, 0x07050301u) ^ 0x80808080u;
}

// Fragment-ordered copy of the DB for the MFMA GEMVs:
//   dbfrag[tile T][step s][plane][lane (j, g)][16 B],
// plane 0 = low bytes, plane 1 = high bytes of the u16 values, each stored as
// the signed byte x - 128; byte t of lane (j, g) is DB row frag_row(s, g, t)
// of column 16T + j.  Same size as the DB; every MFMA B operand of a
// (tile, step) is 2 KiB contiguous.
__global__ void k_tile_db(u4* __restrict__ out, const uint16_t* __restrict__ db, int db_rows, int64_t db_cols) {
  const int steps = db_rows / 64;
  const int64_t total = db_cols / 16 * steps * 64;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(e % 64);
    const int64_t ts = e / 64;
    const int s = (int)(ts % steps);
    const int64_t tile = ts / steps;
    const int j = lane & 15, g = lane >> 4;
    const uint16_t* p = db + (tile * 16 + j) * (int64_t)db_rows + 64 * s + 8 * g;
    const u4 d0 = *reinterpret_cast<const u4*>(p);
    const u4 d1 = *reinterpret_cast<const u4*>(p + 32);
    out[ts * 128 + lane] = u4{plane_lo(d0.x, d0.y), plane_lo(d0.z, d0.w), plane_lo(d1.x, d1.y), plane_lo(d1.z, d1.w)};
    out[ts * 128 + 64 + lane] =
        u4{plane_hi(d0.x, d0.y), plane_hi(d0.z, d0.w), plane_hi(d1.x, d1.y), plane_hi(d1.z, d1.w)};
  }
}

size_t tiled_db_bytes(int db_rows, int64_t db_cols) { return (size_t)db_rows * (size_t)db_cols * 2; }

int launch_tile_db(void* out, const uint16_t* db, int db_rows, int64_t db_cols, hipStream_t s) {
  if (db_rows <= 0 || db_rows % 64 != 0 || db_cols <= 0 || db_cols % 16 != 0) return -1;
  const int64_t total = db_cols / 16 * (db_rows / 64) * 64;
  const int64_t blocks = std::min<int64_t>((total + 255) / 256, 1 << 20);
  hipLaunchKernelGGL(k_tile_db, dim3((unsigned)blocks), dim3(256), 0, s, static_cast<u4*>(out), db, db_rows,
                     db_cols);
  return 0;
}

// Block = 4 waves x TPW column tiles of 16.  The query fragments of one
// 2048-row chunk sit in LDS; the DB stream runs D steps ahead through a
// register ring with clamped (unconditional) loads, so every wave keeps
// about (D - 1) * TPW * 2 KiB of DB reads in flight.  FRAG: the DB is read
// from its fragment-ordered copy (k_tile_db), one contiguous KiB per load
// instruction; otherwise from the column-major DB.
template <int TPW, int D, bool FRAG>
__global__ __launch_bounds__(256) void k_gemv_mfma(uint64_t* __restrict__ c, const int8_t* __restrict__ afrag,
                                                   const int32_t* __restrict__ rsum,
                                                   const uint16_t* __restrict__ db, const u4* __restrict__ dbfrag,
                                                   int nq, int db_rows, int64_t db_cols) {
  static_assert(kMfmaStepsPerChunk % D == 0, "ring depth divides the chunk");
  __shared__ v4i a_lds[kMfmaStepsPerChunk][64];  // 32 KB
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int64_t col_base = (int64_t)blockIdx.x * (4 * TPW * 16) + wave * (TPW * 16);
  const int steps = db_rows / 64;

  v4i acc_lo[TPW], acc_hi[TPW];
  const uint16_t* colp[TPW];
  const u4* fragp[TPW];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    acc_lo[tt] = v4i{0, 0, 0, 0};
    acc_hi[tt] = v4i{0, 0, 0, 0};
    const int64_t col = min(col_base + tt * 16 + j, db_cols - 1);
    colp[tt] = db + col * (int64_t)db_rows + 8 * g;
    const int64_t tile = min(col_base / 16 + tt, db_cols / 16 - 1);
    fragp[tt] = FRAG ? dbfrag + tile * steps * 128 + lane : nullptr;
  }
  u4 ring[D][TPW][2];
  auto load = [&](int st, u4 (&r)[TPW][2]) {
    const int sc = min(st, steps - 1);
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) {
      if (FRAG) {
        r[tt][0] = __builtin_nontemporal_load(fragp[tt] + sc * 128);
        r[tt][1] = __builtin_nontemporal_load(fragp[tt] + sc * 128 + 64);
      } else {
        const uint16_t* p = colp[tt] + 64 * sc;
        r[tt][0] = __builtin_nontemporal_load(reinterpret_cast<const u4*>(p));
        r[tt][1] = __builtin_nontemporal_load(reinterpret_cast<const u4*>(p + 32));
      }
    }
  };

  // The DB ring fills while the first chunk's query fragments are staged.
#pragma unroll
  for (int r = 0; r < D; ++r) load(r, ring[r]);
  for (int k0 = 0; k0 < steps; k0 += kMfmaStepsPerChunk) {
    __syncthreads();
    const v4i* src = reinterpret_cast<const v4i*>(afrag) + (int64_t)k0 * 64;
    for (int e = threadIdx.x; e < kMfmaStepsPerChunk * 64; e += 256) a_lds[e / 64][e % 64] = src[e];
    __syncthreads();
    for (int s = 0; s < kMfmaStepsPerChunk; s += D) {
#pragma unroll
      for (int r = 0; r < D; ++r) {
        const v4i a = a_lds[s + r][lane];
        v4i lo[TPW], hi[TPW];
#pragma unroll
        for (int tt = 0; tt < TPW; ++tt) {
          const u4 d0 = ring[r][tt][0], d1 = ring[r][tt][1];
          if (FRAG) {
            lo[tt] = __builtin_bit_cast(v4i, d0);
            hi[tt] = __builtin_bit_cast(v4i, d1);
          } else {
            lo[tt] = v4i{(int)plane_lo(d0.x, d0.y), (int)plane_lo(d0.z, d0.w), (int)plane_lo(d1.x, d1.y),
                         (int)plane_lo(d1.z, d1.w)};
            hi[tt] = v4i{(int)plane_hi(d0.x, d0.y), (int)plane_hi(d0.z, d0.w), (int)plane_hi(d1.x, d1.y),
                         (int)plane_hi(d1.z, d1.w)};
          }
        }
        load(k0 + s + r + D, ring[r]);
#pragma unroll
        for (int tt = 0; tt < TPW; ++tt) {
          acc_lo[tt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, lo[tt], acc_lo[tt], 0, 0, 0);
          acc_hi[tt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, hi[tt], acc_hi[tt], 0, 0, 0);
        }
      }
    }
  }

  // Lane (j, G) holds rows 4G..4G+3 = limbs 0..3 of (h, p) = (G >> 1, G & 1).
  const int h = g >> 1, p = g & 1;
  const LanePrime lp(p);
  int64_t rs[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) rs[r] = (int64_t)rsum[4 * g + r] * 32896;
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    const uint32_t red = gemv_residue(acc_lo[tt], acc_hi[tt], rs, lp);
    const uint32_t other = __shfl_down(red, 16, 64);  // lane group 2h+1 -> 2h
    const int64_t col = col_base + tt * 16 + j;
    if (p == 0 && h < nq && col < db_cols) c[(int64_t)h * db_cols + col] = crt_compose(red, other);
  }
}

// ---------------------------------------------------------------------------
// Batched form: a block covers 16*TW DB columns for 4*CPW keyword queries
// (2 hash queries each, one 16-row M-tile per keyword query); wave w owns CPW
// of the queries and all TW column tiles.  The block stages each step's DB fragments once in LDS as the two
// signed byte planes (lo, hi) that every wave reads; each wave loads its own
// query fragments straight from L2.  Both streams run D steps ahead through
// register rings with clamped (unconditional) loads.  Query groups of one
// column block are consecutive on one XCD (block id mod 8), so their DB reads
// meet in L2 and the DB leaves HBM about once per batch.
// ---------------------------------------------------------------------------

template <int CPW, int TW, int D, bool FRAG>
__global__ __launch_bounds__(256) void k_gemv_wide(uint64_t* __restrict__ c, int64_t c_client_stride,
                                                   const int8_t* __restrict__ afrag, const int32_t* __restrict__ rsum,
                                                   const uint16_t* __restrict__ db, const u4* __restrict__ dbfrag,
                                                   int nq, int db_rows, int64_t db_cols, int count, int ngroups,
                                                   int64_t col_blocks) {
  static_assert(TW * 64 == 256, "one DB staging item per thread");
  __shared__ v4i planes[2][TW][2][64];
  const int64_t bid = blockIdx.x;
  const int xcd = (int)(bid % 8);
  const int64_t rest = bid / 8;
  const int grp = (int)(rest % ngroups);
  const int64_t cb = (rest / ngroups) * 8 + xcd;
  if (cb >= col_blocks) return;  // whole block: the grid rounds column blocks up to a multiple of 8
  const int64_t col0 = cb * (16 * TW);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int steps = db_rows / 64;
  // DB staging item: tile st = tid / 64, lane slot (column j, row group g) = tid % 64.
  const int st = tid >> 6, sj = lane & 15, sg = lane >> 4;
  const int64_t scol = min(col0 + 16 * st + sj, db_cols - 1);
  const uint16_t* dbp = db + scol * (int64_t)db_rows + 8 * sg;
  const u4* fp = FRAG ? dbfrag + min(col0 / 16 + st, db_cols / 16 - 1) * steps * 128 + lane : nullptr;
  // Query fragments of this wave's clients (clamped to the last real one).
  const int client0 = grp * 4 * CPW + wave * CPW;
  const u4* ap[CPW];
#pragma unroll
  for (int q = 0; q < CPW; ++q)
    ap[q] = reinterpret_cast<const u4*>(afrag) + (int64_t)min(client0 + q, count - 1) * steps * 64 + lane;

  auto load_db = [&](int s, u4 (&d)[2]) {
    if (FRAG) {
      d[0] = __builtin_nontemporal_load(fp + (int64_t)s * 128);
      d[1] = __builtin_nontemporal_load(fp + (int64_t)s * 128 + 64);
    } else {
      const uint16_t* p = dbp + 64 * s;
      d[0] = __builtin_nontemporal_load(reinterpret_cast<const u4*>(p));
      d[1] = __builtin_nontemporal_load(reinterpret_cast<const u4*>(p + 32));
    }
  };
  auto stage_db = [&](int buf, const u4 (&d)[2]) {
    if (FRAG) {
      planes[buf][st][0][lane] = __builtin_bit_cast(v4i, d[0]);
      planes[buf][st][1][lane] = __builtin_bit_cast(v4i, d[1]);
    } else {
      planes[buf][st][0][lane] = v4i{(int)plane_lo(d[0].x, d[0].y), (int)plane_lo(d[0].z, d[0].w),
                                     (int)plane_lo(d[1].x, d[1].y), (int)plane_lo(d[1].z, d[1].w)};
      planes[buf][st][1][lane] = v4i{(int)plane_hi(d[0].x, d[0].y), (int)plane_hi(d[0].z, d[0].w),
                                     (int)plane_hi(d[1].x, d[1].y), (int)plane_hi(d[1].z, d[1].w)};
    }
  };
  auto load_a = [&](int s, u4 (&a)[CPW]) {
#pragma unroll
    for (int q = 0; q < CPW; ++q) a[q] = ap[q][(int64_t)s * 64];
  };

  v4i acc[CPW][TW][2];
#pragma unroll
  for (int q = 0; q < CPW; ++q)
#pragma unroll
    for (int t = 0; t < TW; ++t) acc[q][t][0] = acc[q][t][1] = v4i{0, 0, 0, 0};

  // Rings: DB words of step t in dreg[t % D], query fragments in areg[t % D].
  u4 dreg[D][2];
  u4 areg[D][CPW];
  load_db(0, dreg[0]);
  stage_db(0, dreg[0]);
#pragma unroll
  for (int t = 0; t < D; ++t) {
    load_db(min(t + 1, steps - 1), dreg[(t + 1) % D]);
    load_a(min(t, steps - 1), areg[t]);
  }
  __syncthreads();

  for (int s0 = 0; s0 < steps; s0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int s = s0 + d;
      stage_db((d + 1) & 1, dreg[(d + 1) % D]);  // DB of step s+1 (unused past the end)
      load_db(min(s + 1 + D, steps - 1), dreg[(d + 1) % D]);
#pragma unroll
      for (int t = 0; t < TW; ++t) {
        const v4i lo = planes[d & 1][t][0][lane], hi = planes[d & 1][t][1][lane];
#pragma unroll
        for (int q = 0; q < CPW; ++q) {
          const v4i a = __builtin_bit_cast(v4i, areg[d][q]);
          acc[q][t][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, lo, acc[q][t][0], 0, 0, 0);
          acc[q][t][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, hi, acc[q][t][1], 0, 0, 0);
        }
      }
      load_a(min(s + D, steps - 1), areg[d]);
      __syncthreads();
      __builtin_amdgcn_sched_barrier(0);  // keep each step's loads in its own step
    }
  }

  // Lane (j, G) holds rows 4G..4G+3 = limbs 0..3 of (h, p) = (G >> 1, G & 1).
  const int j = lane & 15, g = lane >> 4;
  const int h = g >> 1, pr = g & 1;
  const LanePrime lp(pr);
#pragma unroll
  for (int q = 0; q < CPW; ++q) {
    const int client = client0 + q;
    const int rc = min(client, count - 1);
    int64_t rs[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) rs[r] = (int64_t)rsum[rc * 16 + 4 * g + r] * 32896;
#pragma unroll
    for (int t = 0; t < TW; ++t) {
      const uint32_t red = gemv_residue(acc[q][t][0], acc[q][t][1], rs, lp);
      const uint32_t other = __shfl_down(red, 16, 64);
      const int64_t col = col0 + 16 * t + j;
      if (pr == 0 && h < nq && col < db_cols && client < count)
        c[(int64_t)client * c_client_stride + (int64_t)h * db_cols + col] = crt_compose(red, other);
    }
  }
}

// ---------------------------------------------------------------------------
// Batched form over the fragment-ordered DB: a block covers QB = WQ*QW keyword
// queries x TB = WT*TW column tiles with WQ*WT waves; wave (wq, wt) owns
// queries wq*QW.. and tiles wt*TW...  Per K step the block's QB query
// fragments and 2*TB DB plane fragments (1 KiB each) are copied into an LDS
// ring by the LDS DMA path and read by every wave that needs them, so each
// query fragment feeds TB*2 MFMAs per block and each DB fragment QB.
// ---------------------------------------------------------------------------

typedef __attribute__((address_space(3))) void lds_void;

template <int QW, int TW, int WQ, int WT, int D>
__global__ __launch_bounds__(64 * WQ * WT) void k_gemv_lds(uint64_t* __restrict__ c, int64_t c_client_stride,
                                                           const int8_t* __restrict__ afrag,
                                                           const int32_t* __restrict__ rsum,
                                                           const u4* __restrict__ dbfrag, int nq, int db_rows,
                                                           int64_t db_cols, int count, int ngroups,
                                                           int64_t col_blocks) {
  constexpr int W = WQ * WT, QB = WQ * QW, TB = WT * TW;
  constexpr int F = QB + 2 * TB;                 // fragments per step
  constexpr int kPerWave = (F + W - 1) / W;      // copies per wave per step
  static_assert(kPerWave * (D - 1) < 64, "vector-memory count field");
  __shared__ u4 ring[D + 2][F][64];
  const int64_t bid = blockIdx.x;
  const int xcd = (int)(bid % 8);
  const int64_t rest = bid / 8;
  const int grp = (int)(rest % ngroups);
  const int64_t cb = (rest / ngroups) * 8 + xcd;
  if (cb >= col_blocks) return;  // whole block: the grid rounds column blocks up to a multiple of 8
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wq = wave % WQ, wt = wave / WQ;
  const int steps = db_rows / 64;
  const int64_t ntiles = db_cols / 16;
  const int client0 = grp * QB;

  // Copy sources of this wave (clamped: past-the-end queries and tiles repeat
  // the last real one).
  const u4* src[kPerWave];
  int64_t sstep[kPerWave];
  int dst[kPerWave];
#pragma unroll
  for (int i = 0; i < kPerWave; ++i) {
    const int f = min(wave * kPerWave + i, F - 1);
    dst[i] = f;
    if (f < QB) {
      src[i] = reinterpret_cast<const u4*>(afrag) + (int64_t)min(client0 + f, count - 1) * steps * 64 + lane;
      sstep[i] = 64;
    } else {
      const int t = (f - QB) >> 1, pl = (f - QB) & 1;
      const int64_t tile = min(cb * TB + t, ntiles - 1);
      src[i] = dbfrag + tile * steps * 128 + pl * 64 + lane;
      sstep[i] = 128;
    }
  }
  auto copy = [&](int s) {
    const int slot = s < steps ? s % (D + 1) : D + 1;
    const int ss = min(s, steps - 1);
#pragma unroll
    for (int i = 0; i < kPerWave; ++i)
      __builtin_amdgcn_global_load_lds((const void*)(src[i] + ss * sstep[i]), (lds_void*)&ring[slot][dst[i]][0], 16,
                                       0, 0);
  };

  v4i acc[QW][TW][2];
#pragma unroll
  for (int q = 0; q < QW; ++q)
#pragma unroll
    for (int t = 0; t < TW; ++t) acc[q][t][0] = acc[q][t][1] = v4i{0, 0, 0, 0};

  // Operands of one step, read from its ring slot.
  auto read = [&](int s, v4i (&a)[QW], v4i (&b)[TW][2]) {
    const u4* slot = &ring[s % (D + 1)][0][0];
#pragma unroll
    for (int q = 0; q < QW; ++q) a[q] = __builtin_bit_cast(v4i, slot[(wq * QW + q) * 64 + lane]);
#pragma unroll
    for (int t = 0; t < TW; ++t)
#pragma unroll
      for (int pl = 0; pl < 2; ++pl)
        b[t][pl] = __builtin_bit_cast(v4i, slot[(QB + 2 * (wt * TW + t) + pl) * 64 + lane]);
  };
  // Step s's operands are read from LDS during step s-1's MFMAs.  At the top
  // of step s a wave needs its own copies of step s+1 (at most kWait newer
  // copies outstanding) and, after the barrier, everyone else's; the copy
  // issued then refills the slot of step s-1, whose reads finished in step
  // s-2 before every wave reached this barrier.
  constexpr int kWait1 = kPerWave * (D - 1);
  constexpr int kWait2 = kPerWave * (D - 2);
  static_assert(D >= 2, "one step read ahead");
#pragma unroll
  for (int t = 0; t < D; ++t) copy(t);
  __builtin_amdgcn_s_waitcnt((kWait1 & 15) | (((kWait1 >> 4) & 3) << 14) | 0x3f70);
  __builtin_amdgcn_s_barrier();
  v4i a[QW], b[TW][2];
  read(0, a, b);
  for (int s = 0; s < steps; ++s) {
    __builtin_amdgcn_s_waitcnt((kWait2 & 15) | (((kWait2 >> 4) & 3) << 14) | 0x3f70);
    __builtin_amdgcn_s_barrier();
    copy(s + D);
    v4i an[QW], bn[TW][2];
    read(min(s + 1, steps - 1), an, bn);
#pragma unroll
    for (int t = 0; t < TW; ++t)
#pragma unroll
      for (int q = 0; q < QW; ++q)
#pragma unroll
        for (int pl = 0; pl < 2; ++pl)
          acc[q][t][pl] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[q], b[t][pl], acc[q][t][pl], 0, 0, 0);
#pragma unroll
    for (int q = 0; q < QW; ++q) a[q] = an[q];
#pragma unroll
    for (int t = 0; t < TW; ++t) b[t][0] = bn[t][0], b[t][1] = bn[t][1];
  }
  __builtin_amdgcn_s_waitcnt(0x3f70);  // vmcnt(0): the copies past the last step have landed
  __syncthreads();

  // Lane (j, G) holds rows 4G..4G+3 = limbs 0..3 of (h, p) = (G >> 1, G & 1).
  const int j = lane & 15, g = lane >> 4;
  const int h = g >> 1, pr = g & 1;
  const LanePrime lp(pr);
#pragma unroll
  for (int q = 0; q < QW; ++q) {
    const int client = client0 + wq * QW + q;
    const int rc = min(client, count - 1);
    int64_t rs[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) rs[r] = (int64_t)rsum[rc * 16 + 4 * g + r] * 32896;
#pragma unroll
    for (int t = 0; t < TW; ++t) {
      const uint32_t red = gemv_residue(acc[q][t][0], acc[q][t][1], rs, lp);
      const uint32_t other = __shfl_down(red, 16, 64);
      const int64_t col = (cb * TB + wt * TW + t) * 16 + j;
      if (pr == 0 && h < nq && col < db_cols && client < count)
        c[(int64_t)client * c_client_stride + (int64_t)h * db_cols + col] = crt_compose(red, other);
    }
  }
}

// Fragments of `count` clients, each with nq <= 2 rows of db_rows (stride
// rows_client_stride elements between clients).
__global__ void k_gemv_prep_batch(const uint64_t* __restrict__ rows, int64_t rows_client_stride, int nq, int db_rows,
                                  int count, int8_t* __restrict__ afrag, int32_t* __restrict__ rsum) {
  const int client = blockIdx.y;
  if (client >= count) return;
  k_gemv_prep_body(rows + (int64_t)client * rows_client_stride, nq, db_rows,
                   afrag + (int64_t)client * (db_rows / 64) * 64 * 16, rsum + client * 16);
}

size_t gemv_mfma_batch_scratch_bytes(int db_rows, int count) {
  return (size_t)count * ((size_t)(db_rows / 64) * 64 * 16 + 16 * sizeof(int32_t));
}

int gemv_batch_variants() { return 2; }
const char* gemv_batch_variant_name(int v) {
  return v == 0 ? "LDS-shared: 8 waves, 16 queries x 8 tiles" : v == 1 ? "L2 query fragments: 4 waves, 16 queries x 4 tiles" : "";
}

int launch_gemv_mfma_batch(uint64_t* c, int64_t c_client_stride, const uint64_t* rows, int64_t rows_client_stride,
                           int nq, const uint16_t* db, const void* dbfrag, int db_rows, int64_t db_cols, int count,
                           void* scratch, hipStream_t s, int variant) {
  if (nq < 1 || nq > 2 || count < 1 || db_rows <= 0 || db_rows % 256 != 0 || db_rows > (1 << 17) || db_cols <= 0)
    return -1;
  // One query: the single-query kernel (8 column tiles per wave, 8-step load
  // ring) streams the DB faster than a 16-query group that is 15/16 padding.
  // The layouts agree: rows [nq][db_rows], output [nq][db_cols], and the
  // scratch of one client is the single-query scratch.
  if (count == 1 && db_rows % kMfmaChunkRows == 0)
    return launch_gemv_mfma(c, rows, nq, db, dbfrag, db_rows, db_cols, scratch, s);
  int8_t* afrag = static_cast<int8_t*>(scratch);
  int32_t* rsum = reinterpret_cast<int32_t*>(afrag + (size_t)count * (db_rows / 64) * 64 * 16);
  hipLaunchKernelGGL(k_zero_u32, dim3((count * 16 + 255) / 256), dim3(256), 0, s, reinterpret_cast<uint32_t*>(rsum),
                     count * 16);
  const int prep_threads = (db_rows / 64) * 64;
  hipLaunchKernelGGL(k_gemv_prep_batch, dim3((prep_threads + 255) / 256, count), dim3(256), 0, s, rows,
                     rows_client_stride, nq, db_rows, count, afrag, rsum);
  // 16 queries per column-block group: 4 waves x 4 queries x 4 column tiles,
  // a 2-step ring (two blocks per CU).
  constexpr int kCpw = 4, kTw = 4;
  const int ngroups = (count + 4 * kCpw - 1) / (4 * kCpw);
  const int64_t col_blocks = (db_cols + 16 * kTw - 1) / (16 * kTw);
  const int64_t blocks = (col_blocks + 7) / 8 * 8 * ngroups;
  if (dbfrag && db_cols % 16 == 0 && variant == 1) {
    hipLaunchKernelGGL((k_gemv_wide<kCpw, kTw, 2, true>), dim3((unsigned)blocks), dim3(256), 0, s, c, c_client_stride,
                       afrag, rsum, db, static_cast<const u4*>(dbfrag), nq, db_rows, db_cols, count, ngroups,
                       col_blocks);
  } else if (dbfrag && db_cols % 16 == 0) {
    // 16 queries x 8 column tiles per block: 8 waves of 4 queries x 4 tiles,
    // a 3-step LDS ring of 32 KiB per step, one block per CU.
    constexpr int kQW = 4, kTW = 4, kWQ = 4, kWT = 2, kD = 3;
    const int ng = (count + kQW * kWQ - 1) / (kQW * kWQ);
    const int64_t cbs = (db_cols / 16 + kTW * kWT - 1) / (kTW * kWT);
    const int64_t nb = (cbs + 7) / 8 * 8 * ng;
    hipLaunchKernelGGL((k_gemv_lds<kQW, kTW, kWQ, kWT, kD>), dim3((unsigned)nb), dim3(64 * kWQ * kWT), 0, s, c,
                       c_client_stride, afrag, rsum, static_cast<const u4*>(dbfrag), nq, db_rows, db_cols, count, ng,
                       cbs);
  } else
    hipLaunchKernelGGL((k_gemv_wide<kCpw, kTw, 2, false>), dim3((unsigned)blocks), dim3(256), 0, s, c, c_client_stride,
                       afrag, rsum, db, static_cast<const u4*>(nullptr), nq, db_rows, db_cols, count, ngroups,
                       col_blocks);
  return 0;
}

size_t gemv_mfma_scratch_bytes(int db_rows) { return (size_t)(db_rows / 64) * 64 * 16 + 16 * sizeof(int32_t); }

static bool gemv_mfma_shape_ok(int nq, int db_rows, int64_t db_cols) {
  return nq >= 1 && nq <= 2 && db_rows > 0 && db_rows % kMfmaChunkRows == 0 && db_rows <= (1 << 17) && db_cols > 0;
}

int launch_gemv_mfma_prep(const uint64_t* rows, int nq, int db_rows, void* scratch, hipStream_t s) {
  if (!gemv_mfma_shape_ok(nq, db_rows, 1)) return -1;
  int8_t* afrag = static_cast<int8_t*>(scratch);
  int32_t* rsum = reinterpret_cast<int32_t*>(afrag + (size_t)(db_rows / 64) * 64 * 16);
  // rsum is accumulated with atomics: clear it with a kernel (graph-capture safe).
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t*>(rsum), 16);
  const int prep_threads = (db_rows / 64) * 64;
  hipLaunchKernelGGL(k_gemv_prep, dim3((prep_threads + 255) / 256), dim3(256), 0, s, rows, nq, db_rows, afrag, rsum);
  return 0;
}

int launch_gemv_mfma_main(uint64_t* c, int nq, const uint16_t* db, const void* dbfrag, int db_rows, int64_t db_cols,
                          const void* scratch, hipStream_t s) {
  if (!gemv_mfma_shape_ok(nq, db_rows, db_cols)) return -1;
  const int8_t* afrag = static_cast<const int8_t*>(scratch);
  const int32_t* rsum = reinterpret_cast<const int32_t*>(afrag + (size_t)(db_rows / 64) * 64 * 16);
  const unsigned blocks = (unsigned)((db_cols + kColsPerBlock - 1) / kColsPerBlock);
  if (dbfrag && db_cols % 16 == 0)
    hipLaunchKernelGGL((k_gemv_mfma<kTilesPerWave, kGemvRing, true>), dim3(blocks), dim3(256), 0, s, c, afrag, rsum,
                       db, static_cast<const u4*>(dbfrag), nq, db_rows, db_cols);
  else
    hipLaunchKernelGGL((k_gemv_mfma<kTilesPerWave, kGemvRing, false>), dim3(blocks), dim3(256), 0, s, c, afrag, rsum,
                       db, static_cast<const u4*>(nullptr), nq, db_rows, db_cols);
  return 0;
}

int launch_gemv_mfma(uint64_t* c, const uint64_t* rows, int nq, const uint16_t* db, const void* dbfrag,
                     int db_rows, int64_t db_cols, void* scratch, hipStream_t s) {
  if (!gemv_mfma_shape_ok(nq, db_rows, db_cols)) return -1;
  launch_gemv_mfma_prep(rows, nq, db_rows, scratch, s);
  return launch_gemv_mfma_main(c, nq, db, dbfrag, db_rows, db_cols, scratch, s);
}

}  // namespace inspire
