// gfx950 kernels for the keyword-PIR answer path (KeywordServer::respond,
// src/pir/server/keyword.rs:534-628).
//
//   k_unpack_query   wire.rs:618-705   bit-unpack of the serialized query
//   k_gemv           gemv_kernel.cu:46-103 / gpu/kernel.rs:139-237
//   k_packing        packing_online_kernel.cu:18-92 + rotation_kernel.cu:21-59
//   k_post           packing_online.rs:359-398 (INTT + CRT; or n^-1 NTT(P) for a server's fold)
//   k_post_ntt       packing_online_kernel.cu:70-88 + packing_online.rs:359-398 (n^-1 NTT(P) after the MFMA packing)
//   k_ahat_ntt       the server fold's n^-1 NTT(a_hat) table
//   k_fold           keyword.rs:543-581 + modulus_switch.rs:45-73 (k_fold_seq: the same, primes in turn)
//   k_rotation_legacy / k_packing (legacy form) for the native/cuda shims
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "modarith.hpp"
#include "ntt_device.hpp"

namespace inspire {

// ---------------------------------------------------------------------------
// Synthetic fills (benchmark and test inputs; same generator as the oracle)
// ---------------------------------------------------------------------------

__global__ void k_synth_fill(void* __restrict__ out, uint64_t seed, uint32_t stream, uint64_t start,
                             uint64_t count, int kind) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const uint64_t h = synth_word(seed, stream, start + i);
    switch (kind) {
      case kSynthCondensed: {
        uint64_t lo = ((h & 0xffffffffull) * kQ0) >> 32;
        uint64_t hi = ((h >> 32) * kQ1) >> 32;
        ((uint64_t*)out)[i] = lo | (hi << 32);
        break;
      }
      case kSynthModQ:
        ((uint64_t*)out)[i] = mulhi64(h, kQ);
        break;
      case kSynthU16:
        ((uint16_t*)out)[i] = (uint16_t)((h & 0xffffull) % 65535ull);
        break;
      case kSynthCrt: {
        const int c = (int)(((start + i) / kN) & 1);
        ((uint64_t*)out)[i] = ((h & 0xffffffffull) * prime(c)) >> 32;
        break;
      }
      case kSynthSmall: {
        uint64_t v = h & 0xfffffull;
        ((uint64_t*)out)[i] = (v % kQ0) | ((v % kQ1) << 32);
        break;
      }
      default:
        break;
    }
  }
}

void launch_synth_fill(void* out, uint64_t seed, uint32_t stream, uint64_t start, uint64_t count, int kind,
                       hipStream_t s) {
  if (count == 0) return;
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_synth_fill, dim3((unsigned)blocks), dim3(256), 0, s, out, seed, stream, start, count,
                     kind);
}

__global__ void k_zero_u32(uint32_t* __restrict__ p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Query unpack (wire.rs:618-705).  Every value is a 56-bit field (28 bits per
// CRT residue, wire.rs:120-125) starting at bit 8 + 56*f, i.e. at byte 1 + 7*f.
// ---------------------------------------------------------------------------

__global__ void k_unpack_query(const uint8_t* __restrict__ q, int nq, int db_rows, uint64_t* __restrict__ y_body,
                               uint64_t* __restrict__ z_body, uint64_t* __restrict__ rows,
                               uint32_t* __restrict__ bodies) {
  const int64_t keys = (int64_t)kTExp * kN * 2;
  const int64_t per_q = (int64_t)db_rows + 6 * kN;
  const int64_t total = keys + (int64_t)nq * per_q;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const uint64_t m28 = (1ull << 28) - 1;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += stride) {
    const uint8_t* p = q + 1 + 7 * f;
    uint64_t v = 0;
#pragma unroll
    for (int b = 0; b < 7; ++b) v |= (uint64_t)p[b] << (8 * b);
    const uint64_t lo = v & m28, hi = (v >> 28) & m28;
    if (f < keys) {
      const int64_t k = f >> 1;  // (i, k) flattened: i*2048 + k
      if (f & 1)
        z_body[k] = lo | (hi << 32);
      else
        y_body[k] = lo | (hi << 32);
    } else {
      const int64_t g = f - keys;
      const int64_t qi = g / per_q, r = g % per_q;
      if (r < db_rows) {
        rows[qi * db_rows + r] = lo | (hi << 32);
      } else {
        const int64_t e = r - db_rows;  // poly i = e / N, slot z = e % N
        const int64_t i = e / kN, z = e % kN;
        bodies[((qi * 6 + i) * 2 + 0) * kN + z] = (uint32_t)lo;
        bodies[((qi * 6 + i) * 2 + 1) * kN + z] = (uint32_t)hi;
      }
    }
  }
}

__global__ void k_unpack_query_batch(const uint8_t* __restrict__ q, int64_t q_stride, int nq, int db_rows,
                                     uint64_t* __restrict__ y_body, uint64_t* __restrict__ z_body,
                                     uint64_t* __restrict__ rows, uint32_t* __restrict__ bodies) {
  const int b = blockIdx.y;
  const int64_t keys = (int64_t)kTExp * kN;
  const uint8_t* qb = q + (int64_t)b * q_stride;
  const int64_t per = (int64_t)nq;
  // Same field walk as k_unpack_query, client b's buffers.
  const int64_t kf = (int64_t)kTExp * kN * 2;
  const int64_t per_q = (int64_t)db_rows + 6 * kN;
  const int64_t total = kf + per * per_q;
  const uint64_t m28 = (1ull << 28) - 1;
  uint64_t* yb = y_body + b * keys;
  uint64_t* zb = z_body + b * keys;
  uint64_t* rb = rows + b * per * db_rows;
  uint32_t* bb = bodies + b * per * 12 * kN;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (int64_t)gridDim.x * blockDim.x) {
    const uint8_t* p = qb + 1 + 7 * f;
    uint64_t v = 0;
#pragma unroll
    for (int i = 0; i < 7; ++i) v |= (uint64_t)p[i] << (8 * i);
    const uint64_t lo = v & m28, hi = (v >> 28) & m28;
    if (f < kf) {
      if (f & 1)
        zb[f >> 1] = lo | (hi << 32);
      else
        yb[f >> 1] = lo | (hi << 32);
    } else {
      const int64_t g = f - kf;
      const int64_t qi = g / per_q, r = g % per_q;
      if (r < db_rows) {
        rb[qi * db_rows + r] = lo | (hi << 32);
      } else {
        const int64_t e = r - db_rows;
        const int64_t i = e / kN, z = e % kN;
        bb[((qi * 6 + i) * 2 + 0) * kN + z] = (uint32_t)lo;
        bb[((qi * 6 + i) * 2 + 1) * kN + z] = (uint32_t)hi;
      }
    }
  }
}

void launch_unpack_query_batch(const uint8_t* q, int64_t q_stride, int nq, int db_rows, int count, uint64_t* y_body,
                               uint64_t* z_body, uint64_t* rows, uint32_t* bodies, hipStream_t s) {
  const int64_t total = (int64_t)kTExp * kN * 2 + (int64_t)nq * (db_rows + 6 * kN);
  int blocks = (int)((total + 255) / 256);
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_unpack_query_batch, dim3(blocks, count), dim3(256), 0, s, q, q_stride, nq, db_rows, y_body,
                     z_body, rows, bodies);
}

void launch_unpack_query(const uint8_t* q, int nq, int db_rows, uint64_t* y_body, uint64_t* z_body,
                         uint64_t* rows, uint32_t* bodies, hipStream_t s) {
  const int64_t total = (int64_t)kTExp * kN * 2 + (int64_t)nq * (db_rows + 6 * kN);
  int blocks = (int)((total + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_unpack_query, dim3(blocks), dim3(256), 0, s, q, nq, db_rows, y_body, z_body, rows,
                     bodies);
}

// ---------------------------------------------------------------------------
// First pass: c[h][col] = sum_k a_h[k] * db[col][k] mod q (gemv_kernel.cu:46-103).
// 8 lanes per DB column, 16 B per lane per load, 32 columns per 256-thread
// block; the query rows are staged in LDS in chunks of 2048 rows.
// ---------------------------------------------------------------------------

constexpr int kGemvChunk = 2048;

template <int NQ>
__global__ __launch_bounds__(256) void k_gemv(uint64_t* __restrict__ c, const uint64_t* __restrict__ rows,
                                              const uint16_t* __restrict__ db, int db_rows, int64_t db_cols) {
  __shared__ uint32_t a_lds[NQ][2][kGemvChunk];
  const int lane8 = threadIdx.x & 7;
  const int64_t col = (int64_t)blockIdx.x * 32 + (threadIdx.x >> 3);
  uint64_t acc[NQ][2];
#pragma unroll
  for (int h = 0; h < NQ; ++h) acc[h][0] = acc[h][1] = 0;

  for (int k0 = 0; k0 < db_rows; k0 += kGemvChunk) {
    __syncthreads();
    for (int t = threadIdx.x; t < NQ * kGemvChunk; t += 256) {
      const int h = t / kGemvChunk, k = t % kGemvChunk;
      const uint64_t v = rows[(int64_t)h * db_rows + k0 + k];
      a_lds[h][0][k] = (uint32_t)v;
      a_lds[h][1][k] = (uint32_t)(v >> 32);
    }
    __syncthreads();
    if (col < db_cols) {
      const uint4* p = reinterpret_cast<const uint4*>(db + col * db_rows + k0);
#pragma unroll 4
      for (int it = 0; it < kGemvChunk / 64; ++it) {
        const uint4 v = p[it * 8 + lane8];
        const int kb = it * 64 + lane8 * 8;
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int h = 0; h < NQ; ++h) {
          const uint4 alo0 = *reinterpret_cast<const uint4*>(&a_lds[h][0][kb]);
          const uint4 alo1 = *reinterpret_cast<const uint4*>(&a_lds[h][0][kb + 4]);
          const uint4 ahi0 = *reinterpret_cast<const uint4*>(&a_lds[h][1][kb]);
          const uint4 ahi1 = *reinterpret_cast<const uint4*>(&a_lds[h][1][kb + 4]);
          const uint32_t lo[8] = {alo0.x, alo0.y, alo0.z, alo0.w, alo1.x, alo1.y, alo1.z, alo1.w};
          const uint32_t hi[8] = {ahi0.x, ahi0.y, ahi0.z, ahi0.w, ahi1.x, ahi1.y, ahi1.z, ahi1.w};
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const uint32_t b = (w[e >> 1] >> (16 * (e & 1))) & 0xffffu;
            acc[h][0] += (uint64_t)lo[e] * b;
            acc[h][1] += (uint64_t)hi[e] * b;
          }
        }
      }
    }
  }
#pragma unroll
  for (int h = 0; h < NQ; ++h)
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int off = 4; off >= 1; off >>= 1) acc[h][r] += __shfl_xor(acc[h][r], off, 8);
  if (lane8 == 0 && col < db_cols) {
#pragma unroll
    for (int h = 0; h < NQ; ++h) {
      const uint32_t r0 = (uint32_t)barrett64(acc[h][0], kCr1Q0, kQ0);
      const uint32_t r1 = (uint32_t)barrett64(acc[h][1], kCr1Q1, kQ1);
      c[(int64_t)h * db_cols + col] = crt_compose(r0, r1);
    }
  }
}

int launch_gemv(uint64_t* c, const uint64_t* rows, int nq, const uint16_t* db, int db_rows, int64_t db_cols,
                hipStream_t s) {
  if (db_rows % kGemvChunk != 0 || db_rows <= 0 || db_cols <= 0) return -1;
  const unsigned blocks = (unsigned)((db_cols + 31) / 32);
  if (nq == 1)
    hipLaunchKernelGGL(k_gemv<1>, dim3(blocks), dim3(256), 0, s, c, rows, db, db_rows, db_cols);
  else if (nq == 2)
    hipLaunchKernelGGL(k_gemv<2>, dim3(blocks), dim3(256), 0, s, c, rows, db, db_rows, db_cols);
  else {
    for (int h = 0; h < nq; ++h)
      hipLaunchKernelGGL(k_gemv<1>, dim3(blocks), dim3(256), 0, s, c + (int64_t)h * db_cols,
                         rows + (int64_t)h * db_rows, db, db_rows, db_cols);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Packing inner product (packing_online_kernel.cu:18-92) with the key
// rotations (rotation_kernel.cu:21-59) fused in.
//
// S[j][c][z] = sum_{i<1023,k<3} y_k[perm_{g_i}(z)] * T_j[3i+k][z]
//            + y_k[perm_{2n-g_i}(z)] * Tbar_j[3i+k][z] + sum_k zb_k[z] * That_j[k][z]  (mod q_c)
//
// perm_t(z) is the automorphism table of packing.rs:1423-1468: the NTT slot
// with exponent (2*rev(z)+1)*t mod 2n.  For t' = 2n - t that exponent is the
// negation, whose slot is 2047 - perm_t(z), so both gathers share one index.
//
// Block = 4 waves; lanes own 64 consecutive slots z, the block owns JG outputs
// j, and the waves split the 1023 rotations.  Partial sums meet in LDS.
// ---------------------------------------------------------------------------

template <int JG, bool FUSED>
__global__ __launch_bounds__(256) void k_packing(uint32_t* __restrict__ S, const uint64_t* __restrict__ bt,
                                                 const uint64_t* __restrict__ btb,
                                                 const uint64_t* __restrict__ bth,
                                                 const uint64_t* __restrict__ y_body,
                                                 const uint64_t* __restrict__ z_body,
                                                 const uint64_t* __restrict__ y_all,
                                                 const uint64_t* __restrict__ y_bar_all) {
  // y_body (48 KB) while accumulating; the same LDS then holds the partial sums.
  constexpr int kPartWords = 4 * 2 * JG * 64;  // u32 residues
  constexpr int kSmemU64 = FUSED ? kTExp * kN : kPartWords / 2;
  static_assert(kPartWords / 2 <= kSmemU64, "partials must fit in the y_body buffer");
  __shared__ uint64_t smem[kSmemU64];
  uint64_t (*ylds)[kN] = reinterpret_cast<uint64_t (*)[kN]>(smem);
  uint32_t (*part)[2][JG][64] = reinterpret_cast<uint32_t (*)[2][JG][64]>(smem);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int z = blockIdx.x * 64 + lane;
  const int j0 = blockIdx.y * JG;

  if (FUSED) {
    for (int t = threadIdx.x; t < kTExp * kN; t += 256) ylds[t / kN][t % kN] = y_body[t];
    __syncthreads();
  }

  const int i_begin = wave * 256;
  const int i_end = min(i_begin + 256, kNumRot);
  // g = 5^i mod 4096 (packing.rs:1639-1652).
  uint32_t g = 1;
  for (int e = 0; e < i_begin; ++e) g = (g * 5u) & (2 * kN - 1);
  const uint32_t ez = 2u * (__brev((uint32_t)z) >> 21) + 1u;

  uint64_t acc_lo[JG], acc_hi[JG];
#pragma unroll
  for (int jj = 0; jj < JG; ++jj) acc_lo[jj] = acc_hi[jj] = 0;

  const size_t jstride = (size_t)kNInner * kN;
  const uint64_t* btj = bt + (size_t)j0 * jstride + z;
  const uint64_t* btbj = btb + (size_t)j0 * jstride + z;

  int since_reduce = 0;
  for (int i = i_begin; i < i_end; ++i) {
    uint64_t yv[kTExp], ybv[kTExp];
    if (FUSED) {
      const uint32_t e1 = (ez * g) & (2 * kN - 1);
      const uint32_t s1 = __brev(e1 >> 1) >> 21;
      const uint32_t s2 = (kN - 1) - s1;
#pragma unroll
      for (int k = 0; k < kTExp; ++k) {
        yv[k] = ylds[k][s1];
        ybv[k] = ylds[k][s2];
      }
    } else {
#pragma unroll
      for (int k = 0; k < kTExp; ++k) {
        yv[k] = y_all[((size_t)i * kTExp + k) * kN + z];
        ybv[k] = y_bar_all[((size_t)i * kTExp + k) * kN + z];
      }
    }
    g = (g * 5u) & (2 * kN - 1);
#pragma unroll
    for (int k = 0; k < kTExp; ++k) {
      const size_t off = ((size_t)i * kTExp + k) * kN;
      const uint32_t ylo = (uint32_t)yv[k], yhi = (uint32_t)(yv[k] >> 32);
      const uint32_t yblo = (uint32_t)ybv[k], ybhi = (uint32_t)(ybv[k] >> 32);
#pragma unroll
      for (int jj = 0; jj < JG; ++jj) {
        const uint64_t b = __builtin_nontemporal_load(btj + jj * jstride + off);
        const uint64_t bb = __builtin_nontemporal_load(btbj + jj * jstride + off);
        acc_lo[jj] += (uint64_t)ylo * (uint32_t)b + (uint64_t)yblo * (uint32_t)bb;
        acc_hi[jj] += (uint64_t)yhi * (uint32_t)(b >> 32) + (uint64_t)ybhi * (uint32_t)(bb >> 32);
      }
    }
    // 6 products < 2^56 per iteration: reduce well before 256 products.
    if (++since_reduce == 32) {
      since_reduce = 0;
#pragma unroll
      for (int jj = 0; jj < JG; ++jj) {
        acc_lo[jj] = barrett64(acc_lo[jj], kCr1Q0, kQ0);
        acc_hi[jj] = barrett64(acc_hi[jj], kCr1Q1, kQ1);
      }
    }
  }
  __syncthreads();  // every wave is done with y_body
#pragma unroll
  for (int jj = 0; jj < JG; ++jj) {
    part[wave][0][jj][lane] = (uint32_t)barrett64(acc_lo[jj], kCr1Q0, kQ0);
    part[wave][1][jj][lane] = (uint32_t)barrett64(acc_hi[jj], kCr1Q1, kQ1);
  }
  __syncthreads();
  // Wave w finalises outputs jj = w, w+4, ...
  for (int jj = wave; jj < JG; jj += 4) {
    const int j = j0 + jj;
    uint64_t lo = (uint64_t)part[0][0][jj][lane] + part[1][0][jj][lane] + part[2][0][jj][lane] +
                  part[3][0][jj][lane];
    uint64_t hi = (uint64_t)part[0][1][jj][lane] + part[1][1][jj][lane] + part[2][1][jj][lane] +
                  part[3][1][jj][lane];
#pragma unroll
    for (int k = 0; k < kTExp; ++k) {
      const uint64_t zv = z_body[k * kN + z];
      const uint64_t hv = bth[((size_t)j * kTExp + k) * kN + z];
      lo += (uint64_t)(uint32_t)zv * (uint32_t)hv;
      hi += (uint64_t)(uint32_t)(zv >> 32) * (uint32_t)(hv >> 32);
    }
    S[((size_t)j * 2 + 0) * kN + z] = (uint32_t)barrett64(lo, kCr1Q0, kQ0);
    S[((size_t)j * 2 + 1) * kN + z] = (uint32_t)barrett64(hi, kCr1Q1, kQ1);
  }
}

template <bool FUSED>
static void launch_packing_t(uint32_t* S, const uint64_t* bt, const uint64_t* btb, const uint64_t* bth,
                             const uint64_t* y_body, const uint64_t* z_body, const uint64_t* y_all,
                             const uint64_t* y_bar_all, int instances, hipStream_t s) {
  const dim3 block(256);
  if (instances % 8 == 0)
    hipLaunchKernelGGL((k_packing<8, FUSED>), dim3(kN / 64, instances / 8), block, 0, s, S, bt, btb, bth, y_body,
                       z_body, y_all, y_bar_all);
  else if (instances % 4 == 0)
    hipLaunchKernelGGL((k_packing<4, FUSED>), dim3(kN / 64, instances / 4), block, 0, s, S, bt, btb, bth, y_body,
                       z_body, y_all, y_bar_all);
  else if (instances % 2 == 0)
    hipLaunchKernelGGL((k_packing<2, FUSED>), dim3(kN / 64, instances / 2), block, 0, s, S, bt, btb, bth, y_body,
                       z_body, y_all, y_bar_all);
  else
    hipLaunchKernelGGL((k_packing<1, FUSED>), dim3(kN / 64, instances), block, 0, s, S, bt, btb, bth, y_body,
                       z_body, y_all, y_bar_all);
}

void launch_packing(uint32_t* S, const uint64_t* bt, const uint64_t* btb, const uint64_t* bth,
                    const uint64_t* y_body, const uint64_t* z_body, int instances, hipStream_t s) {
  launch_packing_t<true>(S, bt, btb, bth, y_body, z_body, nullptr, nullptr, instances, s);
}

void launch_packing_materialized(uint32_t* S, const uint64_t* bt, const uint64_t* btb, const uint64_t* bth,
                                 const uint64_t* y_all, const uint64_t* y_bar_all, const uint64_t* z_body,
                                 int instances, hipStream_t s) {
  launch_packing_t<false>(S, bt, btb, bth, nullptr, z_body, y_all, y_bar_all, instances, s);
}

// ---------------------------------------------------------------------------
// Legacy rotation (rotation_kernel.cu:21-59): materialise y_all / y_bar_all
// from uploaded automorphism tables.
// ---------------------------------------------------------------------------

__global__ void k_rotation_legacy(uint64_t* __restrict__ y_all, uint64_t* __restrict__ y_bar_all,
                                  const uint64_t* __restrict__ y_body, const uint32_t* __restrict__ tables,
                                  const uint32_t* __restrict__ gen_pows, int num_rotations, int t_exp,
                                  int poly_len) {
  const int rot = blockIdx.x, k = blockIdx.y;
  if (rot >= num_rotations || k >= t_exp) return;
  const uint32_t t = gen_pows[rot];
  const size_t t1 = (size_t)((t - 1) / 2) * poly_len;
  const size_t t2 = (size_t)((2 * poly_len - t - 1) / 2) * poly_len;
  const size_t ob = ((size_t)rot * t_exp + k) * poly_len;
  for (int z = threadIdx.x; z < poly_len; z += blockDim.x) {
    y_all[ob + z] = y_body[(size_t)k * poly_len + tables[t1 + z]];
    y_bar_all[ob + z] = y_body[(size_t)k * poly_len + tables[t2 + z]];
  }
}

void launch_rotation_legacy(uint64_t* y_all, uint64_t* y_bar_all, const uint64_t* y_body, const uint32_t* tables,
                            const uint32_t* gen_pows, int num_rotations, int t_exp, int poly_len,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_rotation_legacy, dim3(num_rotations, t_exp), dim3(256), 0, s, y_all, y_bar_all, y_body,
                     tables, gen_pows, num_rotations, t_exp, poly_len);
}

// ---------------------------------------------------------------------------
// Post-process (packing_online.rs:359-398): P_j = CRT(INTT(S_j)) mod q.
//
// The fold of a server adds P_j to its row-1 product before that row's
// inverse transform, so the server's post-process writes n^-1 NTT_c(P_j)
// per prime instead, two u32 rows [2][N] in the 8N bytes of P_j's row
// (ntt_out); the fold transforms the first P of each chain back itself.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(1024) void k_post(uint64_t* __restrict__ P, const uint32_t* __restrict__ S,
                                               const uint32_t* __restrict__ tw_tab, int ntt_out) {
  __shared__ uint32_t buf[2][kNPad];
  const NttTables tw{tw_tab};
  const int j = blockIdx.x;
  if (ntt_out) {
    uint32_t* P32 = reinterpret_cast<uint32_t*>(P) + (size_t)j * 2 * kN;
    for (int t = threadIdx.x; t < 2 * kN; t += 1024) {
      const uint32_t v = S[(size_t)j * 2 * kN + t];
      P32[t] = t < kN ? mulmod<0>(v, kNInvQ0) : mulmod<1>(v, kNInvQ1);
    }
    return;
  }
  for (int t = threadIdx.x; t < 2 * kN; t += 1024) buf[t / kN][lds_index(t % kN)] = S[(size_t)j * 2 * kN + t];
  __syncthreads();
  uint32_t* const polys[2] = {buf[0], buf[1]};
  const int pr[2] = {0, 1};
  ntt_inverse_r4<1024, 2>(polys, pr, tw);
  for (int z = threadIdx.x; z < kN; z += 1024) P[(size_t)j * kN + z] = crt_compose(buf[0][lds_index(z)], buf[1][lds_index(z)]);
}

void launch_post(uint64_t* P, const uint32_t* S, const uint32_t* tw, int count, hipStream_t s, int ntt_out) {
  hipLaunchKernelGGL(k_post, dim3(count), dim3(1024), 0, s, P, S, tw, ntt_out);
}

// The fold's row-0 table: out[j][c][z] = n^-1 NTT_c(a_hat_j mod q_c) at NTT
// slot z, one workgroup per instance j.
__global__ __launch_bounds__(512) void k_ahat_ntt(uint32_t* __restrict__ out, const uint64_t* __restrict__ a_hat,
                                                  const uint32_t* __restrict__ tw_tab) {
  __shared__ uint32_t buf[2][kNPad];
  const NttTables tw{tw_tab};
  const size_t j = blockIdx.x;
  for (int z = threadIdx.x; z < kN; z += 512) {
    const uint64_t x = a_hat[j * kN + z];
    buf[0][lds_index(z)] = (uint32_t)barrett64(x, cr1(0), kQ0);
    buf[1][lds_index(z)] = (uint32_t)barrett64(x, cr1(1), kQ1);
  }
  __syncthreads();
  uint32_t* const polys[2] = {buf[0], buf[1]};
  const int pr[2] = {0, 1};
  ntt_forward_r4m<512, 2>(polys, pr, tw);
  for (int z = threadIdx.x; z < kN; z += 512) {
    out[(j * 2 + 0) * kN + z] = mulmod<0>(buf[0][lds_index(z)], kNInvQ0);
    out[(j * 2 + 1) * kN + z] = mulmod<1>(buf[1][lds_index(z)], kNInvQ1);
  }
}

void launch_ahat_ntt(uint32_t* ahat_ntt, const uint64_t* a_hat, const uint32_t* tw, int64_t instances,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_ahat_ntt, dim3((unsigned)instances), dim3(512), 0, s, ahat_ntt, a_hat, tw);
}

// Post-process of the MFMA packing for the fold of a server.  S holds the
// rotation-term sums z-major, S[b][c][z][j]; this kernel adds the z_body terms
//     sum_k z_body_b[k][z] * bold_t_hat[j][k][z]   (packing_online_kernel.cu:70-88)
// and writes n^-1 NTT_c(P_j), P[b][j][c][z] (u32; see k_post).  One block per
// tile of 32 outputs j x 64 slots z of one client: S comes in along j (128-byte
// rows) through a padded LDS tile and leaves along z.  Block order: each XCD
// (block id mod 8) takes a contiguous range of (j tile, client, z tile), so the
// clients of a j tile meet its bold_t_hat rows in that XCD's L2.
constexpr int kPnJ = 32;
constexpr int kPnZ = 64;
constexpr int kPnThreads = 256;

__global__ __launch_bounds__(kPnThreads) void k_post_ntt(uint32_t* __restrict__ P, const uint32_t* __restrict__ S,
                                                         const uint64_t* __restrict__ z_body,
                                                         int64_t key_client_stride,
                                                         const uint64_t* __restrict__ bth, int64_t instances,
                                                         int clients) {
  __shared__ uint32_t tile[2][kPnZ][kPnJ + 1];
  constexpr int zt_n = kN / kPnZ;
  const int64_t jt_n = (instances + kPnJ - 1) / kPnJ;
  const int64_t total = jt_n * clients * zt_n;
  const int64_t per_xcd = (total + 7) / 8;
  const int64_t r = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (r >= total) return;
  const int zt = (int)(r % zt_n);
  const int b = (int)(r / zt_n % clients);
  const int64_t j0 = r / zt_n / clients * kPnJ;
  const int z0 = zt * kPnZ;
  const int nj = (int)(instances - j0 < kPnJ ? instances - j0 : kPnJ);
  const uint32_t* Sb = S + (size_t)b * 2 * kN * instances;
  for (int e = threadIdx.x; e < 2 * kPnZ * kPnJ; e += kPnThreads) {
    const int jj = e % kPnJ, zz = e / kPnJ % kPnZ, c = e / (kPnJ * kPnZ);
    tile[c][zz][jj] = jj < nj ? Sb[((size_t)c * kN + z0 + zz) * instances + j0 + jj] : 0u;
  }
  __syncthreads();
  const int zz = threadIdx.x % kPnZ, z = z0 + zz;
  const uint64_t* zb = z_body + b * key_client_stride + z;
  const uint64_t zv[3] = {zb[0], zb[kN], zb[2 * kN]};
  for (int jj = threadIdx.x / kPnZ; jj < nj; jj += kPnThreads / kPnZ) {
    const int64_t j = j0 + jj;
    const uint64_t* th = bth + (size_t)j * kTExp * kN + z;
    uint64_t h0 = 0, h1 = 0;
#pragma unroll
    for (int k = 0; k < kTExp; ++k) {
      const uint64_t tv = th[k * kN];
      h0 += (uint64_t)(uint32_t)zv[k] * (uint32_t)tv;
      h1 += (zv[k] >> 32) * (tv >> 32);
    }
    uint32_t v0 = (uint32_t)barrett64(h0, cr1(0), kQ0) + tile[0][zz][jj];
    uint32_t v1 = (uint32_t)barrett64(h1, cr1(1), kQ1) + tile[1][zz][jj];
    v0 = v0 >= kQ0 ? v0 - kQ0 : v0;
    v1 = v1 >= kQ1 ? v1 - kQ1 : v1;
    uint32_t* out = P + ((size_t)b * instances + j) * 2 * kN + z;
    out[0] = mulmod<0>(v0, kNInvQ0);
    out[kN] = mulmod<1>(v1, kNInvQ1);
  }
}

void launch_post_ntt(uint64_t* P, const uint32_t* S, const uint64_t* z_body, int64_t key_client_stride,
                     const uint64_t* bth, int clients, int64_t instances, hipStream_t s) {
  const int64_t total = (instances + kPnJ - 1) / kPnJ * clients * (kN / kPnZ);
  const int64_t blocks = (total + 7) / 8 * 8;
  hipLaunchKernelGGL(k_post_ntt, dim3((unsigned)blocks), dim3(kPnThreads), 0, s, reinterpret_cast<uint32_t*>(P), S,
                     z_body, key_client_stride, bth, instances, clients);
}

// ---------------------------------------------------------------------------
// RGSW fold + modulus switch (keyword.rs:543-581, modulus_switch.rs:45-73).
//
// One block per (hash query h, chain w).  Horner over i = d-1 .. 0:
//   sum <- raw(ct_gsw * NTT(G^-1(sum))) + packed[w*d+i]  (mod q)
// where ct_gsw row 0 is NTT(-a) in every column and row 1 is the client's
// RGSW body.  The first step has sum = 0, so it reduces to sum = packed[w*d+d-1].
// packed[j] = (a_hat_j, c_h[j*n..] + P_j).
//
// Two forms, the same bytes.  The server form (kNttAdd) takes a_hat_j and P_j
// in the NTT domain, n^-1 scaled (launch_ahat_ntt's table; launch_post_ntt or
// launch_post with ntt_out), adds them to the product rows before their
// inverse transforms and transforms the first P of its chain back at the
// start; the stateless form adds a_hat_j and P_j to the rows after them.
// ---------------------------------------------------------------------------

// Workgroup x runs logical chain xcd_chunk(x, n) of n: XCD k (block id mod 8)
// takes the contiguous range of n/8 logical chains (one more for k < n % 8)
// that starts at k*(n/8) + min(k, n%8), so the chains of one client, which read
// the same RGSW body at every Horner step and the same P rows, share an XCD's
// L2.
__device__ inline int xcd_chunk(int x, int n) {
  const int q = n / 8, r = n % 8, k = x % 8;
  return k * q + min(k, r) + x / 8;
}

constexpr int kFoldThreads = 1024;  // the two-prime NTT forms split the block into one half per prime
constexpr int kFoldSlots = kN / kFoldThreads;  // NTT slots owned per thread

struct FoldArgs {
  uint8_t* out;              // [clients] x out_stride bytes, each [nq][c][12288]
  const uint64_t* a_hat;     // [instances][N]
  const uint64_t* cvec;      // [clients][nq][instances*N]
  const uint64_t* P;         // [clients][instances][N]
  const uint32_t* bodies;    // [clients][nq][6][2][N]
  const uint32_t* fold;      // [2][3][N]: inverse powers, their w', n^-1 NTT(-a); then [2][N]: its w
  const uint32_t* tw;        // NTT tables
  // [instances][2][N]: n^-1 NTT_c(a_hat_j) per prime, or nullptr (fold row 0
  // from a_hat in the coefficient domain)
  const uint32_t* ahat_ntt;
  int64_t instances;
  int64_t out_stride;        // bytes between clients' outputs
  int d;
  int c;
  int nq;
};

// Bit-packs the rescaled rows r0 (28 bits) and r1 (20 bits), padded
// polynomial layout in LDS, into the 12288-byte chain blob of (client b, hash
// query h, chain w) (modulus_switch.rs:45-73, util.rs:303-321).
template <int NT>
__device__ inline void fold_write_response(const FoldArgs& args, int b, int h, int w, const uint32_t* r0,
                                           const uint32_t* r1) {
  uint32_t* out32 =
      reinterpret_cast<uint32_t*>(args.out + (size_t)b * args.out_stride + ((size_t)h * args.c + w) * kRespBytes);
  constexpr int kRow0Bits = 28 * kN;
  constexpr int kTotalBits = 28 * kN + 20 * kN;
  for (int wd = threadIdx.x; wd < kRespBytes / 4; wd += NT) {
    const int lo = wd * 32, hi = lo + 32;
    uint32_t word = 0;
    if (lo < kRow0Bits) {
      const int vlast = (min(hi, kRow0Bits) - 1) / 28;
      for (int v = lo / 28; v <= vlast; ++v) {
        const int sh = 28 * v - lo;
        const uint32_t val = r0[lds_index(v)];
        word |= sh >= 0 ? (val << sh) : (val >> (-sh));
      }
    }
    if (hi > kRow0Bits && lo < kTotalBits) {
      const int blo = max(lo, kRow0Bits) - kRow0Bits;
      const int bhi = min(hi, kTotalBits) - kRow0Bits;
      const int vlast = (bhi - 1) / 20;
      for (int v = blo / 20; v <= vlast; ++v) {
        const int sh = kRow0Bits + 20 * v - lo;
        const uint32_t val = r1[lds_index(v)];
        word |= sh >= 0 ? (val << sh) : (val >> (-sh));
      }
    }
    out32[wd] = word;
  }
}

// Row 0 of a Horner sum has a known transform.  With the table
// Ah_j = n^-1 NTT_c(a_hat_j), the fold adds Ah_j to its row-0 product before
// the inverse transform, so that product row S (n^-1 scaled) satisfies
// NTT_c(sum row 0) = n S = 2^11 S, and the transform of row 0's third gadget
// digit follows from S and the other two (third_gadget_digit_ntt<11>): it
// replaces one of the six digit transforms of a step per prime.

// LDS row of gadget digit row k (row r + 2 k' for digit k' of row r) among
// the transformed ones: in the server form, row 4 (row 0's third digit) is
// not transformed and row 5 takes its place.
template <bool kNttAdd>
__device__ constexpr int digit_row(int k) {
  return kNttAdd && k == 5 ? 4 : k;
}

template <bool kNttAdd>
__global__ __launch_bounds__(kFoldThreads) void k_fold(FoldArgs args) {
  constexpr int kDig = kNttAdd ? 5 : 6;  // transformed digit rows per prime
  // LDS (padded polynomial layout of ntt_device.hpp): the transformed gadget
  // digits for both primes and the two product rows of both primes.
  extern __shared__ uint32_t lds[];
  uint32_t* dig = lds;                     // [2 (prime)][kDig][kNPad]
  uint32_t* prod = lds + 2 * kDig * kNPad;  // [2 (prime)][2 (row)][kNPad]
  const NttTables tw{args.tw};
  const int per_client = args.nq * args.c;
  const int chain = xcd_chunk((int)blockIdx.x, (int)gridDim.x);
  const int b = chain / per_client;
  const int h = (chain % per_client) / args.c;
  const int w = chain % args.c;
  const int tid = threadIdx.x;
  const int d = args.d;
  const uint64_t* cv = args.cvec + ((size_t)b * args.nq + h) * args.instances * kN;
  const uint64_t* Pc = args.P + (size_t)b * args.instances * kN;
  const uint32_t* P32 = reinterpret_cast<const uint32_t*>(Pc);  // server form: n^-1 NTT_c(P_j) rows
  const uint32_t* body = args.bodies + ((size_t)b * args.nq + h) * 6 * 2 * kN;

  // sum[r][s] for slots z = tid + kFoldThreads*s; in the server form, S[c][s] is
  // n^-1 NTT_c(sum row 0) at NTT slot z
  uint64_t sum[2][kFoldSlots];
  uint32_t S[2][kFoldSlots];
  {
    const size_t j = (size_t)w * d + (d - 1);
#pragma unroll
    for (int s = 0; s < kFoldSlots; ++s) {
      const int z = tid + s * kFoldThreads;
      sum[0][s] = args.a_hat[j * kN + z];
      sum[1][s] = kNttAdd ? cv[j * kN + z] : reduce_small_q(cv[j * kN + z] + Pc[j * kN + z]);
      if constexpr (kNttAdd) {
        S[0][s] = args.ahat_ntt[(j * 2 + 0) * kN + z];
        S[1][s] = args.ahat_ntt[(j * 2 + 1) * kN + z];
        prod[0 * kNPad + lds_index(z)] = P32[(j * 2 + 0) * kN + z];
        prod[1 * kNPad + lds_index(z)] = P32[(j * 2 + 1) * kN + z];
      }
    }
    if constexpr (kNttAdd) {
      // the chain's first P arrives in the NTT domain too: transform it back
      __syncthreads();
      ntt_inverse_r4_halves_unscaled<1>(prod, args.fold);
#pragma unroll
      for (int s = 0; s < kFoldSlots; ++s) {
        const int lz = lds_index(tid + s * kFoldThreads);
        sum[1][s] = reduce_small_q(sum[1][s] + crt_compose(prod[lz], prod[kNPad + lz]));
      }
    }
  }

  const uint64_t mask = (1ull << kGadgetBits) - 1;

  for (int i = d - 2; i >= 0; --i) {
    const size_t j = (size_t)w * d + i;
    // G^-1(sum): digit k of row r is gadget row r + 2k (gadget.rs:35-61); the
    // digits (< 2^19) are canonical for both primes.
#pragma unroll
    for (int s = 0; s < kFoldSlots; ++s) {
      const int lz = lds_index(tid + s * kFoldThreads);
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          if (kNttAdd && r + 2 * k == 4) continue;
          const uint32_t dv = (uint32_t)((sum[r][s] >> (kGadgetBits * k)) & mask);
          dig[digit_row<kNttAdd>(r + 2 * k) * kNPad + lz] = dv;
          dig[(kDig + digit_row<kNttAdd>(r + 2 * k)) * kNPad + lz] = dv;
        }
    }
    __syncthreads();
    ntt_forward_r4_halves<kDig>(dig, tw);
    // ct_gsw * G: row 0 = NTT(-a) * sum_k G_k; row 1 = sum_k body_k * G_k.
#pragma unroll
    for (int s = 0; s < kFoldSlots; ++s) {
      const int z = tid + s * kFoldThreads;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const uint32_t q = prime(c);
        uint32_t g[6];
#pragma unroll
        for (int k = 0; k < 6; ++k)
          if (!kNttAdd || k != 4) g[k] = dig[(kDig * c + digit_row<kNttAdd>(k)) * kNPad + lds_index(z)];
        if constexpr (kNttAdd) g[4] = third_gadget_digit_ntt<kLogN>(S[c][s], g[0], g[2], c);
        uint32_t gsum = 0;
        uint64_t acc = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          gsum += g[k];
          acc += (uint64_t)body[(k * 2 + c) * kN + z] * g[k];
        }
        // both rows carry n^-1 (row 0 in the table, row 1 here): the inverse
        // transform below skips the per-stage halving
        // six canonical NTT values: gsum < 6q
        uint32_t gs = gsum >= 4 * q ? gsum - 4 * q : gsum;
        gs = gs >= 2 * q ? gs - 2 * q : gs;
        gs = gs >= q ? gs - q : gs;
        const uint32_t na = args.fold[(c * 3 + 2) * kN + z];
        const uint32_t r1 = (uint32_t)barrett64(acc, cr1(c), q);
        uint32_t p0 = c == 0 ? mulmod<0>(na, gs) : mulmod<1>(na, gs);
        uint32_t p1 = c == 0 ? mulmod<0>(r1, kNInvQ0) : mulmod<1>(r1, kNInvQ1);
        if constexpr (kNttAdd) {
          p0 += args.ahat_ntt[(j * 2 + c) * kN + z];
          p0 = p0 >= q ? p0 - q : p0;
          S[c][s] = p0;
          p1 += P32[(j * 2 + c) * kN + z];
          p1 = p1 >= q ? p1 - q : p1;
        }
        prod[(c * 2 + 0) * kNPad + lds_index(z)] = p0;
        prod[(c * 2 + 1) * kNPad + lds_index(z)] = p1;
      }
    }
    __syncthreads();
    ntt_inverse_r4_halves_unscaled<2>(prod, args.fold);
#pragma unroll
    for (int s = 0; s < kFoldSlots; ++s) {
      const int z = tid + s * kFoldThreads;
      const int lz = lds_index(z);
      const uint64_t x0 = crt_compose(prod[0 * kNPad + lz], prod[2 * kNPad + lz]);
      const uint64_t x1 = crt_compose(prod[1 * kNPad + lz], prod[3 * kNPad + lz]);
      sum[0][s] = kNttAdd ? x0 : reduce_small_q(x0 + args.a_hat[j * kN + z]);
      sum[1][s] = reduce_small_q(x1 + cv[j * kN + z] + (kNttAdd ? 0 : Pc[j * kN + z]));
    }
    // dig is rewritten next; prod only after two more barriers.
  }

  // switch_and_keep(2^20, 268369921, 2048): row 0 -> 28 bits, row 1 -> 20 bits.
  __syncthreads();
#pragma unroll
  for (int s = 0; s < kFoldSlots; ++s) {
    const int lz = lds_index(tid + s * kFoldThreads);
    dig[lz] = (uint32_t)rescale_from_q(sum[0][s], kQPrime2);
    dig[kNPad + lz] = (uint32_t)rescale_from_q(sum[1][s], kQPrime1);
  }
  __syncthreads();
  fold_write_response<kFoldThreads>(args, b, h, w, dig, dig + kNPad);
}

// The same fold with 512 threads and the two CRT primes taken one after the
// other: per Horner step and prime, the gadget digits are transformed (five in
// the server form, whose sixth follows from S), the two product rows are
// formed in place of digits 0 and 1 and transformed back;
// prime 0's result waits in registers for the CRT.  50 KiB of LDS, so three
// workgroups share a CU and every thread has a radix-4 unit in each pass.
constexpr int kFold2Threads = 512;
constexpr int kFoldSeqMinBlocks = 512;
constexpr int kFold2Slots = kN / kFold2Threads;
static_assert(kFold2Slots == 4, "a thread's slots are one radix-4 unit of the first and last passes");
static_assert(12ull * kQ0 < (1ull << 32) && 4ull * kQ0 * kQ0 < (1ull << 58) && kQ1 < kQ0,
              "pair stage bounds: six digits below 2q sum below 2^32; q times 4q fits reduce58");
static_assert(6ull * (1ull << 28) * 4 * kQ0 < (1ull << 61) && (1ull << 57) + (1ull << 32) < (1ull << 58),
              "the row-1 sum of six 28-bit body fields times values below 4q splits into hi < 2^29 and lo, "
              "and hi (2^32 mod q) + lo < 2^58");

template <bool kNttAdd>
__global__ __launch_bounds__(kFold2Threads) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_fold_seq(FoldArgs args) {
  constexpr int kDig = kNttAdd ? 5 : 6;  // transformed digit rows per prime (see third_gadget_digit_ntt)
  __shared__ uint32_t dig[6][kNPad];
  const NttTables tw{args.tw};
  const int per_client = args.nq * args.c;
  const int chain = xcd_chunk((int)blockIdx.x, (int)gridDim.x);
  // the hash query varies fastest, so the chains that read the same P rows
  // (one per hash query) run side by side and share them in L2
  const int b = chain / per_client;
  const int h = chain % args.nq;
  const int w = (chain % per_client) / args.nq;
  const int tid = threadIdx.x;
  const int d = args.d;
  const uint64_t* cv = args.cvec + ((size_t)b * args.nq + h) * args.instances * kN;
  const uint64_t* Pc = args.P + (size_t)b * args.instances * kN;
  const uint32_t* P32 = reinterpret_cast<const uint32_t*>(Pc);  // server form: n^-1 NTT_c(P_j) rows
  const uint32_t* body = args.bodies + ((size_t)b * args.nq + h) * 6 * 2 * kN;

  // sum[r][s] for slots z = tid + 512 s.  The product stage works on the slot
  // pairs (2v, 2v + 1), v = tid + 512 h, instead; in the server form S0 holds
  // n^-1 NTT_0(sum row 0) at those slots (index 2h + e for slot 2v + e), and
  // prime 1's value sits in LDS row 5, which this form does not transform
  // (each thread reads and writes only its own slots there).
  uint64_t sum[2][kFold2Slots];
  uint32_t S0[kFold2Slots];
  {
    const size_t j = (size_t)w * d + (d - 1);
#pragma unroll
    for (int s = 0; s < kFold2Slots; ++s) {
      const int z = tid + s * kFold2Threads;
      sum[0][s] = args.a_hat[j * kN + z];
      sum[1][s] = kNttAdd ? cv[j * kN + z] : reduce_small_q(cv[j * kN + z] + Pc[j * kN + z]);
      if constexpr (kNttAdd) {
        const int zp = 2 * (tid + (s >> 1) * kFold2Threads) + (s & 1);  // the pair slot of index s
        S0[s] = args.ahat_ntt[(j * 2 + 0) * kN + zp];
        dig[5][lds_index(z)] = args.ahat_ntt[(j * 2 + 1) * kN + z];
      }
    }
    if constexpr (kNttAdd) {
      // the chain's first P arrives in the NTT domain too: transform it back
      // (rows 0 and 1, one per prime)
#pragma unroll
      for (int s = 0; s < kFold2Slots; ++s) {
        const int z = tid + s * kFold2Threads;
        dig[0][lds_index(z)] = P32[(j * 2 + 0) * kN + z];
        dig[1][lds_index(z)] = P32[(j * 2 + 1) * kN + z];
      }
      __syncthreads();
      uint32_t* const ppolys[2] = {dig[0], dig[1]};
      const int pprime[2] = {0, 1};
      ntt_inverse_r4_unscaled<kFold2Threads, 2>(ppolys, pprime, args.fold);
#pragma unroll
      for (int s = 0; s < kFold2Slots; ++s) {
        const int lz = lds_index(tid + s * kFold2Threads);
        sum[1][s] = reduce_small_q(sum[1][s] + crt_compose(dig[0][lz], dig[1][lz]));
      }
    }
  }
  uint32_t* dpolys[kDig];
#pragma unroll
  for (int p = 0; p < kDig; ++p) dpolys[p] = dig[p];
  uint32_t* const ipolys[2] = {dig[0], dig[1]};
  const uint64_t mask = (1ull << kGadgetBits) - 1;

  for (int i = d - 2; i >= 0; --i) {
    const size_t j = (size_t)w * d + i;
    uint32_t res0[2][kFold2Slots];
#pragma unroll 1  // one prime at a time keeps the NTT within 80 VGPRs (6 waves per SIMD)
    for (int c = 0; c < 2; ++c) {
      // G^-1(sum): digit k of row r is gadget row r + 2k (gadget.rs:35-61),
      // and the first radix-4 pass of its transform in registers: the
      // thread's slots tid + 512 s are that pass's unit.  Each thread rewrites
      // only its own slots, which it alone read last.
#pragma unroll
      for (int p = 0; p < kDig; ++p) {
        const int row = kNttAdd && p == 4 ? 5 : p;  // gadget row held in LDS row p (digit_row)
        uint32_t a[kFold2Slots];
#pragma unroll
        for (int s = 0; s < kFold2Slots; ++s)
          a[s] = (uint32_t)((sum[row & 1][s] >> (kGadgetBits * (row >> 1))) & mask);
        ntt_forward_first_pass(a, c, tw);
#pragma unroll
        for (int s = 0; s < kFold2Slots; ++s) dig[p][lds_index(tid + s * kFold2Threads)] = a[s];
      }
      __syncthreads();
      ntt_forward_r4_1p<kFold2Threads, kDig, 2, false>(dpolys, c, tw);
      const uint32_t* ah_row = args.ahat_ntt + (j * 2 + c) * kN;  // server form: this step's table row
      const uint32_t* pn_row = P32 + (j * 2 + c) * kN;            // and P row
      // The last forward stage, the products (ct_gsw * G: row 0 = NTT(-a) *
      // sum_k G_k, row 1 = sum_k body_k * G_k) and the first inverse stage all
      // work on slot pairs (2v, 2v + 1): each thread runs them for its pairs
      // v = tid + 512 h in registers.
      const uint32_t q = prime(c), q2 = 2 * q, mu = mu58(c);
      const uint32_t c32 = c ? kC32Q1 : kC32Q0, ninv = c ? kNInvQ1 : kNInvQ0;
      const uint32_t ninvp = c ? shoup_companion(kNInvQ1, kQ1) : shoup_companion(kNInvQ0, kQ0);
      const uint32_t* fw = tw.fw(c);
      const uint32_t* fwp = tw.fwp(c);
      const uint32_t* iw = args.fold + c * 3 * kN;
      const uint32_t* iwp = iw + kN;
#pragma unroll 1  // one pair at a time keeps its loads and sums within the register budget
      for (int hh = 0; hh < 2; ++hh) {
        const int v = tid + hh * kFold2Threads;
        const int l0 = lds_index(2 * v), l1 = lds_index(2 * v + 1);
        // wave-uniform row bases plus a 32-bit lane offset, opaque so that the
        // compiler does not hoist 64-bit lane addresses out of the loops
        uint32_t off = 2u * (uint32_t)v;
        asm volatile("" : "+v"(off));
        // Each transformed digit row goes straight into the two slots' sums;
        // rows 0 and 2 stay for the derived row 4 (server form).  The pair's
        // per-slot operands are 8-byte loads.
        uint64_t acc[2] = {0, 0};
        uint32_t gsum[2] = {0, 0}, gsum_lo[2] = {0, 0}, g0[2] = {0, 0}, g2[2] = {0, 0};
        auto add_row = [&](int row, uint32_t x, uint32_t y) {
          const uint2 bv = *reinterpret_cast<const uint2*>(body + (row * 2 + c) * kN + off);
          uint32_t(&gs)[2] = row < 3 ? gsum_lo : gsum;
          gs[0] += x;
          gs[1] += y;
          acc[0] += (uint64_t)bv.x * x;
          acc[1] += (uint64_t)bv.y * y;
        };
        // The products take the transformed digits as the last stage leaves
        // them, below 4q: acc stays below 6 * 2^28 * 4q < 2^61 (the bodies are
        // 28-bit fields of the query), and the digit sum is
        // formed in two halves of three (each below 12q < 2^32), each brought
        // below 4q (gsum_lo, gsum).
#pragma unroll
        for (int p = 0; p < kDig; ++p) {
          uint32_t x = dig[p][l0], y = dig[p][l1];
          fwd_bf(x, y, fw[1024 + v], fwp[1024 + v], q);
          const int row = kNttAdd && p == 4 ? 5 : p;  // gadget row held in LDS row p (digit_row)
          add_row(row, x, y);
          if (row == 0) {
            g0[0] = x;
            g0[1] = y;
          } else if (row == 2) {
            g2[0] = x;
            g2[1] = y;
          }
        }
        if constexpr (kNttAdd) {
          // constant indices into S0 (hh is 0 or 1) keep it in registers
          const uint32_t s0 = c == 0 ? (hh ? S0[2] : S0[0]) : dig[5][l0];
          const uint32_t s1 = c == 0 ? (hh ? S0[3] : S0[1]) : dig[5][l1];
          add_row(4, third_gadget_digit_ntt<kLogN, false>(s0, g0[0], g2[0], c),
                  third_gadget_digit_ntt<kLogN, false>(s1, g0[1], g2[1], c));
        }
        // both rows carry n^-1 (see k_fold)
        const uint2 na = *reinterpret_cast<const uint2*>(args.fold + (c * 3 + 2) * kN + off);
        const uint2 nap = *reinterpret_cast<const uint2*>(args.fold + (6 + c) * kN + off);
        uint32_t p0[2], p1[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          // each half of the digit sum from below 12q to below 4q: the sum
          // of the halves (< 8q < 2^32) enters one Shoup product
          uint32_t lo = gsum_lo[e] >= 8 * q ? gsum_lo[e] - 8 * q : gsum_lo[e];
          lo = lo >= 4 * q ? lo - 4 * q : lo;
          uint32_t hi = gsum[e] >= 8 * q ? gsum[e] - 8 * q : gsum[e];
          hi = hi >= 4 * q ? hi - 4 * q : hi;
          uint32_t r0 = shoup_lazy(lo + hi, e ? na.y : na.x, e ? nap.y : nap.x, q);  // [0, 2q)
          p0[e] = r0 >= q ? r0 - q : r0;
          // acc < 2^61 is 2^32 hi + lo with hi < 2^29, so hi (2^32
          // mod q) + lo < 2^58; its residue below 3q enters a Shoup product
          // by n^-1
          const uint32_t r1 =
              reduce58_lazy_rt((uint64_t)(uint32_t)(acc[e] >> 32) * c32 + (uint32_t)acc[e], q, mu);
          const uint32_t s1 = shoup_lazy(r1, ninv, ninvp, q);  // [0, 2q)
          p1[e] = s1 >= q ? s1 - q : s1;
        }
        if constexpr (kNttAdd) {
          const uint2 ah = *reinterpret_cast<const uint2*>(ah_row + off);
          const uint2 pn = *reinterpret_cast<const uint2*>(pn_row + off);
          // below 2q, which the inverse butterfly and the next step's
          // derivation both take
          p0[0] += ah.x;
          p0[1] += ah.y;
          p1[0] += pn.x;
          p1[1] += pn.y;
          if (c == 0 && hh == 0) {
            S0[0] = p0[0];
            S0[1] = p0[1];
          } else if (c == 0) {
            S0[2] = p0[0];
            S0[3] = p0[1];
          } else {
            dig[5][l0] = p0[0];
            dig[5][l1] = p0[1];
          }
        }
        inv_bf_unscaled(p0[0], p0[1], iw[1024 + v], iwp[1024 + v], q);
        inv_bf_unscaled(p1[0], p1[1], iw[1024 + v], iwp[1024 + v], q);
        dig[0][l0] = p0[0];
        dig[0][l1] = p0[1];
        dig[1][l0] = p1[0];
        dig[1][l1] = p1[1];
      }
      __syncthreads();
      ntt_inverse_r4_1p_unscaled<kFold2Threads, 2, 3, false>(ipolys, c, args.fold);
      // the last inverse pass in registers, on the thread's own slots
      uint32_t y0[kFold2Slots], y1[kFold2Slots];
#pragma unroll
      for (int s = 0; s < kFold2Slots; ++s) {
        const int lz = lds_index(tid + s * kFold2Threads);
        y0[s] = dig[0][lz];
        y1[s] = dig[1][lz];
      }
      ntt_inverse_last_pass_unscaled(y0, c, args.fold);
      ntt_inverse_last_pass_unscaled(y1, c, args.fold);
      if (c == 0) {
#pragma unroll
        for (int s = 0; s < kFold2Slots; ++s) {
          res0[0][s] = y0[s];
          res0[1][s] = y1[s];
        }
      } else {
        const uint64_t* cv_row = cv + j * kN;
        const uint64_t* ahat_row = args.a_hat + j * kN;
        const uint64_t* p_row = Pc + j * kN;
#pragma unroll
        for (int s = 0; s < kFold2Slots; ++s) {
          uint32_t z = (uint32_t)(tid + s * kFold2Threads);
          asm volatile("" : "+v"(z));  // as off above
          const uint64_t x0 = crt_compose(res0[0][s], y0[s]);
          const uint64_t x1 = crt_compose(res0[1][s], y1[s]);
          if constexpr (kNttAdd) {
            sum[0][s] = x0;
            const uint64_t v1 = x1 + cv_row[z];  // < 2q
            sum[1][s] = v1 >= kQ ? v1 - kQ : v1;
          } else {
            sum[0][s] = reduce_small_q(x0 + ahat_row[z]);
            sum[1][s] = reduce_small_q(x1 + cv_row[z] + p_row[z]);
          }
        }
      }
    }
  }

  __syncthreads();
#pragma unroll
  for (int s = 0; s < kFold2Slots; ++s) {
    const int lz = lds_index(tid + s * kFold2Threads);
    dig[0][lz] = (uint32_t)rescale_from_q(sum[0][s], kQPrime2);
    dig[1][lz] = (uint32_t)rescale_from_q(sum[1][s], kQPrime1);
  }
  __syncthreads();
  fold_write_response<kFold2Threads>(args, b, h, w, dig[0], dig[1]);
}

size_t fold_lds_bytes(bool table) { return (size_t)(2 * (table ? 5 : 6) + 4) * kNPad * sizeof(uint32_t); }

int launch_fold(uint8_t* out, const uint64_t* a_hat, const uint32_t* ahat_ntt, const uint64_t* cvec,
                const uint64_t* P, const uint32_t* bodies, const uint32_t* fold_tab, const uint32_t* tw,
                int64_t instances, int d, int nq, hipStream_t s) {
  return launch_fold_batch(out, 0, a_hat, ahat_ntt, cvec, P, bodies, fold_tab, tw, instances, d, nq, 1, s, 0);
}

int launch_fold_batch(uint8_t* out, int64_t out_stride, const uint64_t* a_hat, const uint32_t* ahat_ntt,
                      const uint64_t* cvec, const uint64_t* P, const uint32_t* bodies, const uint32_t* fold_tab,
                      const uint32_t* tw, int64_t instances, int d, int nq, int clients, hipStream_t s,
                      int variant) {
  if (d <= 0 || instances % d != 0 || nq <= 0 || clients <= 0) return -1;
  FoldArgs a{out, a_hat, cvec, P, bodies, fold_tab, tw, ahat_ntt, instances, out_stride, d, (int)(instances / d), nq};
  static bool attr_set = false;
  if (!attr_set) {
    // Raises the dynamic-LDS ceiling of the 1024-thread forms (118 and 132 KiB
    // of the 160 KiB per CU).
    (void)hipFuncSetAttribute((const void*)k_fold<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)fold_lds_bytes(true));
    (void)hipFuncSetAttribute((const void*)k_fold<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)fold_lds_bytes(false));
    attr_set = true;
  }
  const int blocks = clients * nq * a.c;
  const bool table = ahat_ntt != nullptr;
  // Below two workgroups per CU the 1024-thread form (both primes at once)
  // has the shorter chain latency; above, three k_fold_seq per CU win.
  // variant 1 / 2 force k_fold / k_fold_seq.
  if (variant == 2 || (variant != 1 && blocks >= kFoldSeqMinBlocks)) {
    if (table)
      hipLaunchKernelGGL(k_fold_seq<true>, dim3(blocks), dim3(kFold2Threads), 0, s, a);
    else
      hipLaunchKernelGGL(k_fold_seq<false>, dim3(blocks), dim3(kFold2Threads), 0, s, a);
  } else if (table) {
    hipLaunchKernelGGL(k_fold<true>, dim3(blocks), dim3(kFoldThreads), fold_lds_bytes(true), s, a);
  } else {
    hipLaunchKernelGGL(k_fold<false>, dim3(blocks), dim3(kFoldThreads), fold_lds_bytes(false), s, a);
  }
  return 0;
}

}  // namespace inspire
