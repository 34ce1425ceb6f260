// DB encode (native/cuda/encode_kernel.cu:18-228, keyword.rs:404-420):
// polynomial interpolation of every DB row's column groups over
// Z_p[X]/(X^n + 1), p = 65535, n = 2048.
//
// For row r and column group wp the d polynomials P_j[k] = db[(wp*d + j)*n + k][r]
// (column-major DB) become E_t = d^-1 * sum_j P_j * X^(-2n*j*t/d): a radix-2
// transform with bit-reversed input whose twiddles are monomials, i.e. signed
// coefficient rotations.  Three passes over HBM:
//   k_encode_transpose<true>   64x64 tiles through LDS: column-major DB -> work
//                              units [row][group][slot bitrev(j)][k], both
//                              sides coalesced;
//   k_encode_levels            one workgroup per work unit (d*n u16, 128 KiB at
//                              d = 32): all log2(d) levels and the d^-1 scale
//                              in LDS;
//   k_encode_transpose<false>  the inverse transpose back to the DB.
// The reference runs one global-memory pass per level instead (about 2.3x the
// HBM traffic at d = 32).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "modarith.hpp"

namespace inspire {

constexpr uint32_t kP16 = 65535;
constexpr int kTile = 64;

// x mod 65535 for any 32-bit x (two end-around folds, then one subtraction).
__device__ inline uint32_t mod_p16(uint32_t x) {
  x = (x & 0xffffu) + (x >> 16);
  x = (x & 0xffffu) + (x >> 16);
  return x >= kP16 ? x - kP16 : x;
}

__device__ inline int rev_bits(int x, int bits) { return bits ? (int)(__brev((uint32_t)x) >> (32 - bits)) : 0; }

// Tile (row block tr, column block tc) of the groups [wp0, wp0 + nwp).
// TO_WORK: db -> work with bit-reversed slots; else work -> db, natural slots.
template <bool TO_WORK>
__global__ __launch_bounds__(256) void k_encode_transpose(uint16_t* __restrict__ dst, const uint16_t* __restrict__ src,
                                                          int64_t db_rows, int d, int logd, int64_t wp0, int64_t nwp) {
  __shared__ uint16_t tile[kTile][kTile + 2];
  const int64_t col_tiles = nwp * d * kN / kTile;
  const int64_t tr = blockIdx.x / col_tiles, tc = blockIdx.x % col_tiles;
  const int64_t row0 = tr * kTile, cc0 = tc * kTile;  // cc: column within the slab
  const int64_t col0 = wp0 * d * kN + cc0;
  // The 64 columns of a tile share (group, j) since 64 divides n.
  const int64_t lwp = cc0 / ((int64_t)d * kN);
  const int j = (int)((cc0 / kN) % d);
  const int k0 = (int)(cc0 % kN);
  const int slot = TO_WORK ? rev_bits(j, logd) : j;
  auto work_at = [&](int64_t row, int k) { return ((row * nwp + lwp) * d + slot) * (int64_t)kN + k; };
  for (int idx = threadIdx.x; idx < kTile * kTile; idx += 256) {
    const int a = idx / kTile, b = idx % kTile;
    if (TO_WORK)  // a = column, b = row: 128 B per column
      tile[b][a] = src[(col0 + a) * db_rows + row0 + b];
    else          // a = row, b = k: 128 B per row
      tile[a][b] = src[work_at(row0 + a, k0 + b)];
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < kTile * kTile; idx += 256) {
    const int a = idx / kTile, b = idx % kTile;
    if (TO_WORK)  // a = row, b = k
      dst[work_at(row0 + a, k0 + b)] = tile[a][b];
    else          // a = column, b = row
      dst[(col0 + a) * db_rows + row0 + b] = tile[b][a];
  }
}

// All levels of one work unit in LDS.  Level l pairs slots (even, even + 2^l)
// of groups of 2^(l+1); butterfly i of a group uses X^e, e = (2n - 2n*i/size)
// mod 2n (compute_twiddle_exponents, encode_kernel.cu:220-228):
//   even' = even + X^e odd,  odd' = even - X^e odd   (mod p, X^n = -1).
// Each thread keeps its D butterflies' results in registers between the read
// and write halves of a level (the rotation reads other threads' entries).
constexpr int kEncThreads = 1024;

template <int D>
__global__ __launch_bounds__(kEncThreads) void k_encode_levels(uint16_t* __restrict__ work, uint32_t num_inv) {
  constexpr int kItems = D >= 2 ? D / 2 * kN / kEncThreads : 1;  // butterflies per thread per level
  static_assert(D == 1 || (D / 2 * kN) % kEncThreads == 0, "whole butterflies per thread");
  extern __shared__ uint16_t buf[];  // [D][kN]
  uint16_t* unit = work + (size_t)blockIdx.x * D * kN;
  const int tid = threadIdx.x;
  for (int i = tid; i < D * kN / 2; i += kEncThreads)
    reinterpret_cast<uint32_t*>(buf)[i] = reinterpret_cast<const uint32_t*>(unit)[i];
  __syncthreads();
#pragma unroll 1
  for (int level = 0; (1 << level) < D; ++level) {
    const int half = 1 << level, size = half << 1;
    uint32_t res[kItems];
#pragma unroll
    for (int m = 0; m < kItems; ++m) {
      const int it = tid + m * kEncThreads;
      const int pi = it / kN, k = it % kN;
      const int i = pi % half;
      const int even = (pi / half) * size + i;
      const int e = (2 * kN - 2 * kN * i / size) & (2 * kN - 1);
      const int src = (k - e) & (2 * kN - 1);
      const uint32_t o = buf[(even + half) * kN + (src & (kN - 1))];
      const uint32_t t = src < kN ? o : (o ? kP16 - o : 0u);
      const uint32_t ev = buf[even * kN + k];
      res[m] = mod_p16(ev + t) | (mod_p16(ev + kP16 - t) << 16);
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kItems; ++m) {
      const int it = tid + m * kEncThreads;
      const int pi = it / kN, k = it % kN;
      const int i = pi % half;
      const int even = (pi / half) * size + i;
      buf[even * kN + k] = (uint16_t)(res[m] & 0xffffu);
      buf[(even + half) * kN + k] = (uint16_t)(res[m] >> 16);
    }
    __syncthreads();
  }
  for (int i = tid; i < D * kN; i += kEncThreads) unit[i] = (uint16_t)mod_p16((uint32_t)buf[i] * num_inv);
}

size_t encode_work_bytes_per_group(int64_t db_rows, int d) { return (size_t)db_rows * d * kN * sizeof(uint16_t); }

template <int D>
static void launch_levels(uint16_t* work, int64_t units, uint32_t num_inv, hipStream_t s) {
  const size_t lds = (size_t)D * kN * sizeof(uint16_t);
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)k_encode_levels<D>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr = true;
  }
  hipLaunchKernelGGL(k_encode_levels<D>, dim3((unsigned)units), dim3(kEncThreads), lds, s, work, num_inv);
}

int launch_encode_db(uint16_t* out, const uint16_t* in, int64_t db_rows, int64_t db_cols, int d, uint32_t num_inv,
                     uint16_t* work, int64_t groups_per_batch, hipStream_t s) {
  int logd = 0;
  while ((1 << logd) < d) ++logd;
  if ((1 << logd) != d || d > 32 || db_rows <= 0 || db_rows % kTile != 0 || db_cols % ((int64_t)d * kN) != 0 ||
      groups_per_batch < 1)
    return -1;
  const int64_t c = db_cols / ((int64_t)d * kN);
  for (int64_t wp0 = 0; wp0 < c; wp0 += groups_per_batch) {
    const int64_t nwp = std::min(groups_per_batch, c - wp0);
    const int64_t tiles = (db_rows / kTile) * (nwp * d * kN / kTile);
    hipLaunchKernelGGL((k_encode_transpose<true>), dim3((unsigned)tiles), dim3(256), 0, s, work, in, db_rows, d, logd,
                       wp0, nwp);
    const int64_t units = db_rows * nwp;
    switch (d) {
      case 1: launch_levels<1>(work, units, num_inv, s); break;  // no levels: the scale alone
      case 2: launch_levels<2>(work, units, num_inv, s); break;
      case 4: launch_levels<4>(work, units, num_inv, s); break;
      case 8: launch_levels<8>(work, units, num_inv, s); break;
      case 16: launch_levels<16>(work, units, num_inv, s); break;
      default: launch_levels<32>(work, units, num_inv, s); break;
    }
    hipLaunchKernelGGL((k_encode_transpose<false>), dim3((unsigned)tiles), dim3(256), 0, s, out, work, db_rows, d,
                       logd, wp0, nwp);
  }
  return 0;
}

}  // namespace inspire
