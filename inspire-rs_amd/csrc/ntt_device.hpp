// Workgroup-wide negacyclic NTT over one 28-bit prime, data resident in LDS.
//
// Same transform and slot order as core/ntt.rs (forward: Cooley-Tukey with the
// bit-reversed power table of the minimal primitive 2n-th root, so slot k holds
// the evaluation at psi^(2*rev(k)+1); inverse: Gentleman-Sande with the
// halved inverse powers, so the result is scaled by n^-1).  Values stay in
// [0, 4q) during the forward pass and [0, 2q) during the inverse pass, as in
// Harvey's lazy butterflies; both passes end with a canonical reduction.
#pragma once
#include "modarith.hpp"

namespace inspire {

// Twiddle tables per prime: [fwd w][fwd w'][inv w][inv w'] x 2048, u32.
struct NttTables {
  const uint32_t* __restrict__ t;  // [2][4][kN]
  __device__ const uint32_t* fw(int c) const { return t + (c * 4 + 0) * kN; }
  __device__ const uint32_t* fwp(int c) const { return t + (c * 4 + 1) * kN; }
  __device__ const uint32_t* iw(int c) const { return t + (c * 4 + 2) * kN; }
  __device__ const uint32_t* iwp(int c) const { return t + (c * 4 + 3) * kN; }
};

// One forward stage for butterfly b (0..1023) of polynomial a (LDS).
__device__ inline void fwd_butterfly(uint32_t* a, int mm, int b, uint32_t w, uint32_t wp, uint32_t q) {
  const int t = (kN / 2) >> mm;
  const int i = b >> (10 - mm);
  const int j = b & (t - 1);
  const int xi = 2 * t * i + j;
  uint32_t x = a[xi];
  uint32_t y = a[xi + t];
  const uint32_t q2 = 2 * q;
  x = x >= q2 ? x - q2 : x;
  uint32_t qn = shoup_lazy(y, w, wp, q);
  a[xi] = x + qn;
  a[xi + t] = x + q2 - qn;
}

__device__ inline void inv_butterfly(uint32_t* a, int mm, int b, uint32_t w, uint32_t wp, uint32_t q) {
  const int t = (kN / 2) >> mm;
  const int i = b >> (10 - mm);
  const int j = b & (t - 1);
  const int xi = 2 * t * i + j;
  uint32_t x = a[xi];
  uint32_t y = a[xi + t];
  const uint32_t q2 = 2 * q;
  uint32_t tt = q2 - y + x;
  uint32_t cx = x + y;
  cx = cx >= q2 ? cx - q2 : cx;
  a[xi] = (cx + ((cx & 1u) ? q : 0u)) >> 1;
  a[xi + t] = w * tt - mulhi32(tt, wp) * q;
}

// Forward NTT of `count` polynomials; polys[p] points into LDS and prime[p]
// selects the modulus.  Caller syncs before (inputs canonical) and after.
// Every thread of the block (blockDim.x must divide 1024*count evenly in
// units of 1024) participates.
template <int NT, int COUNT>
__device__ inline void ntt_forward_multi(uint32_t* const (&polys)[COUNT], const int (&prime_of)[COUNT],
                                         const NttTables& tw) {
  static_assert((1024 % NT) == 0, "thread count must divide 1024");
  const int tid = threadIdx.x;
  for (int mm = 0; mm < kLogN; ++mm) {
    const int m = 1 << mm;
#pragma unroll
    for (int p = 0; p < COUNT; ++p) {
      const int c = prime_of[p];
      const uint32_t q = prime(c);
      for (int b = tid; b < 1024; b += NT) {
        const int i = b >> (10 - mm);
        fwd_butterfly(polys[p], mm, b, tw.fw(c)[m + i], tw.fwp(c)[m + i], q);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < COUNT; ++p) {
    const uint32_t q = prime(prime_of[p]);
    for (int z = tid; z < kN; z += NT) {
      uint32_t v = polys[p][z];
      v = v >= 2 * q ? v - 2 * q : v;
      v = v >= q ? v - q : v;
      polys[p][z] = v;
    }
  }
  __syncthreads();
}

template <int NT, int COUNT>
__device__ inline void ntt_inverse_multi(uint32_t* const (&polys)[COUNT], const int (&prime_of)[COUNT],
                                         const NttTables& tw) {
  const int tid = threadIdx.x;
  for (int mm = kLogN - 1; mm >= 0; --mm) {
    const int h = 1 << mm;
#pragma unroll
    for (int p = 0; p < COUNT; ++p) {
      const int c = prime_of[p];
      const uint32_t q = prime(c);
      for (int b = tid; b < 1024; b += NT) {
        const int i = b >> (10 - mm);
        inv_butterfly(polys[p], mm, b, tw.iw(c)[h + i], tw.iwp(c)[h + i], q);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < COUNT; ++p) {
    const uint32_t q = prime(prime_of[p]);
    for (int z = tid; z < kN; z += NT) {
      uint32_t v = polys[p][z];
      v = v >= q ? v - q : v;
      polys[p][z] = v;
    }
  }
  __syncthreads();
}

}  // namespace inspire
