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

namespace inspire {

// Radix-4 forms of the same transforms: two butterfly stages per pass with the
// four values of a unit kept in registers (half the LDS traffic and barriers).
// All NP polynomials use prime c.  Every thread of the block participates.
template <int NT, int NP>
__device__ inline void ntt_forward_r4(uint32_t* const (&polys)[NP], int c, const NttTables& tw) {
  const int tid = threadIdx.x;
  const uint32_t q = prime(c), q2 = 2 * q;
  const uint32_t* fw = tw.fw(c);
  const uint32_t* fwp = tw.fwp(c);
  auto bf = [&](uint32_t& x, uint32_t& y, uint32_t w, uint32_t wp) {
    const uint32_t xr = x >= q2 ? x - q2 : x;
    const uint32_t qn = shoup_lazy(y, w, wp, q);
    x = xr + qn;
    y = xr + q2 - qn;
  };
  for (int mm = 0; mm + 1 < kLogN; mm += 2) {
    const int m = 1 << mm, t = (kN / 2) >> mm, th = t >> 1;
    for (int u = tid; u < kN / 4; u += NT) {
      const int i = u >> (9 - mm), j = u & (th - 1);
      const int x0 = 2 * t * i + j;
      const uint32_t w = fw[m + i], wp = fwp[m + i];
      const uint32_t w1 = fw[2 * m + 2 * i], wp1 = fwp[2 * m + 2 * i];
      const uint32_t w2 = fw[2 * m + 2 * i + 1], wp2 = fwp[2 * m + 2 * i + 1];
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        uint32_t* a = polys[p];
        uint32_t a0 = a[x0], a1 = a[x0 + th], a2 = a[x0 + t], a3 = a[x0 + t + th];
        bf(a0, a2, w, wp);
        bf(a1, a3, w, wp);
        bf(a0, a1, w1, wp1);
        bf(a2, a3, w2, wp2);
        a[x0] = a0;
        a[x0 + th] = a1;
        a[x0 + t] = a2;
        a[x0 + t + th] = a3;
      }
    }
    __syncthreads();
  }
  // last stage (mm = 10, t = 1) and the final reduction to [0, q)
  for (int u = tid; u < kN / 2; u += NT) {
    const uint32_t w = fw[1024 + u], wp = fwp[1024 + u];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      uint32_t* a = polys[p];
      uint32_t x = a[2 * u], y = a[2 * u + 1];
      bf(x, y, w, wp);
      x = x >= q2 ? x - q2 : x;
      x = x >= q ? x - q : x;
      y = y >= q2 ? y - q2 : y;
      y = y >= q ? y - q : y;
      a[2 * u] = x;
      a[2 * u + 1] = y;
    }
  }
  __syncthreads();
}

// Forward radix-4 over NP polynomials with per-polynomial primes.
template <int NT, int NP>
__device__ inline void ntt_forward_r4m(uint32_t* const (&polys)[NP], const int (&prime_of)[NP], const NttTables& tw) {
  const int tid = threadIdx.x;
  auto bf = [](uint32_t& x, uint32_t& y, uint32_t w, uint32_t wp, uint32_t q) {
    const uint32_t q2 = 2 * q;
    const uint32_t xr = x >= q2 ? x - q2 : x;
    const uint32_t qn = shoup_lazy(y, w, wp, q);
    x = xr + qn;
    y = xr + q2 - qn;
  };
  for (int mm = 0; mm + 1 < kLogN; mm += 2) {
    const int m = 1 << mm, t = (kN / 2) >> mm, th = t >> 1;
    for (int u = tid; u < kN / 4; u += NT) {
      const int i = u >> (9 - mm), j = u & (th - 1);
      const int x0 = 2 * t * i + j;
      uint32_t w[2][3], wp[2][3];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        w[c][0] = tw.fw(c)[m + i];
        wp[c][0] = tw.fwp(c)[m + i];
        w[c][1] = tw.fw(c)[2 * m + 2 * i];
        wp[c][1] = tw.fwp(c)[2 * m + 2 * i];
        w[c][2] = tw.fw(c)[2 * m + 2 * i + 1];
        wp[c][2] = tw.fwp(c)[2 * m + 2 * i + 1];
      }
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        const int c = prime_of[p];
        const uint32_t q = prime(c);
        uint32_t* a = polys[p];
        uint32_t a0 = a[x0], a1 = a[x0 + th], a2 = a[x0 + t], a3 = a[x0 + t + th];
        bf(a0, a2, w[c][0], wp[c][0], q);
        bf(a1, a3, w[c][0], wp[c][0], q);
        bf(a0, a1, w[c][1], wp[c][1], q);
        bf(a2, a3, w[c][2], wp[c][2], q);
        a[x0] = a0;
        a[x0 + th] = a1;
        a[x0 + t] = a2;
        a[x0 + t + th] = a3;
      }
    }
    __syncthreads();
  }
  for (int u = tid; u < kN / 2; u += NT) {
    uint32_t w[2], wp[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      w[c] = tw.fw(c)[1024 + u];
      wp[c] = tw.fwp(c)[1024 + u];
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int c = prime_of[p];
      const uint32_t q = prime(c), q2 = 2 * q;
      uint32_t* a = polys[p];
      uint32_t x = a[2 * u], y = a[2 * u + 1];
      bf(x, y, w[c], wp[c], q);
      x = x >= q2 ? x - q2 : x;
      x = x >= q ? x - q : x;
      y = y >= q2 ? y - q2 : y;
      y = y >= q ? y - q : y;
      a[2 * u] = x;
      a[2 * u + 1] = y;
    }
  }
  __syncthreads();
}

template <int NT, int NP>
__device__ inline void ntt_inverse_r4(uint32_t* const (&polys)[NP], const int (&prime_of)[NP], const NttTables& tw) {
  const int tid = threadIdx.x;
  auto bf = [](uint32_t& x, uint32_t& y, uint32_t w, uint32_t wp, uint32_t q) {
    const uint32_t q2 = 2 * q;
    const uint32_t tt = q2 - y + x;
    uint32_t cx = x + y;
    cx = cx >= q2 ? cx - q2 : cx;
    x = (cx + ((cx & 1u) ? q : 0u)) >> 1;
    y = w * tt - mulhi32(tt, wp) * q;
  };
  // first stage (mm = 10, t = 1)
  for (int u = tid; u < kN / 2; u += NT) {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int c = prime_of[p];
      uint32_t* a = polys[p];
      uint32_t x = a[2 * u], y = a[2 * u + 1];
      bf(x, y, tw.iw(c)[1024 + u], tw.iwp(c)[1024 + u], prime(c));
      a[2 * u] = x;
      a[2 * u + 1] = y;
    }
  }
  __syncthreads();
  // stage pairs (mm, mm-1) for mm = 9, 7, 5, 3, 1
  for (int mm = kLogN - 2; mm >= 1; mm -= 2) {
    const int h = 1 << mm, t = (kN / 2) >> mm;
    for (int u = tid; u < kN / 4; u += NT) {
      const int i = u >> (10 - mm), j = u & (t - 1);  // unit (i, j): base = 4t*i + j
      const int x0 = 4 * t * i + j;
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        const int c = prime_of[p];
        const uint32_t q = prime(c);
        const uint32_t* iw = tw.iw(c);
        const uint32_t* iwp = tw.iwp(c);
        uint32_t* a = polys[p];
        uint32_t a0 = a[x0], a1 = a[x0 + t], a2 = a[x0 + 2 * t], a3 = a[x0 + 3 * t];
        bf(a0, a1, iw[h + 2 * i], iwp[h + 2 * i], q);
        bf(a2, a3, iw[h + 2 * i + 1], iwp[h + 2 * i + 1], q);
        bf(a0, a2, iw[(h >> 1) + i], iwp[(h >> 1) + i], q);
        bf(a1, a3, iw[(h >> 1) + i], iwp[(h >> 1) + i], q);
        a[x0] = a0 >= q ? a0 - q : a0;
        a[x0 + t] = a1 >= q ? a1 - q : a1;
        a[x0 + 2 * t] = a2 >= q ? a2 - q : a2;
        a[x0 + 3 * t] = a3 >= q ? a3 - q : a3;
      }
    }
    __syncthreads();
  }
}

}  // namespace inspire
