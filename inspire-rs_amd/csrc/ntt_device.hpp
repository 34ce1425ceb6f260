// Workgroup-wide negacyclic NTTs over the 28-bit CRT primes, data resident in LDS.
//
// Same transform and slot order as core/ntt.rs (forward: Cooley-Tukey with the
// bit-reversed power table of the minimal primitive 2n-th root, so slot k holds
// the evaluation at psi^(2*rev(k)+1); inverse: Gentleman-Sande with the
// halved inverse powers, so the result is scaled by n^-1).  The inverse pass
// keeps values in [0, 2q), as in Harvey's lazy butterflies.  The forward pass
// is lazier still (fwd_stage_reduction): from inputs below 2q it skips the
// reduction of x in stages 1-7, where values stay below 16q < 2^32, restores
// x < 2q in stage 8 and runs Harvey's [0, 4q) butterflies after that.  Both
// passes end with a canonical reduction.
//
// Radix-4: two butterfly stages per pass with the four values of a unit in
// registers.  LDS layout: coefficient x of a polynomial lives at word
// lds_index(x) = x + x/32 (kNPad words per polynomial).  Passes whose units
// span at least 32 consecutive coefficients map lanes to consecutive
// coefficients; the small-stride passes map lanes to consecutive units, whose
// coefficients are 4*stride apart: with the pad every 32 words both mappings
// put the 32 lanes of a half-wave on 32 different banks.
#pragma once
#include <type_traits>

#include "modarith.hpp"

namespace inspire {

constexpr int kNPad = kN + kN / 32;  // 2112 words per polynomial in LDS

__host__ __device__ constexpr int lds_index(int x) { return x + (x >> 5); }

// t[i] for a workgroup-uniform table t and an unsigned 32-bit index: the byte
// offset stays 32-bit, so the load addresses from the scalar base (no 64-bit
// lane address arithmetic).
__device__ inline uint32_t tab_at(const uint32_t* __restrict__ t, uint32_t i) {
  return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(t) + (i << 2));
}

// Forward Cooley-Tukey butterfly with Harvey's lazy reduction: x, y in [0, 4q)
// -> x + w y, x - w y + 2q in [0, 4q), wp = floor(w 2^32 / q).  The Shoup
// product w y mod q lies in [0, 2q) for any y < 2^32.  It is formed negated,
// nqn = qt q - w y = -(w y mod q) mod 2^32, so that y' = xr + 2q + nqn is one
// three-input add (8 VALU operations per butterfly instead of 9).
//
// kRed sets how x is reduced first: 1 is Harvey's (x < 4q -> x < 2q); 0 skips
// it, so the outputs exceed x's bound by at most 2q (6 operations); 3
// subtracts 8q, 4q and 2q conditionally, taking any x < 16q below 2q.
template <int kRed>
__device__ inline void fwd_bf_r(uint32_t& x, uint32_t& y, uint32_t w, uint32_t wp, uint32_t q) {
  static_assert(kRed == 0 || kRed == 1 || kRed == 3, "reduction of x: none, 2q, or 8q, 4q, 2q");
  uint32_t xr = x;
  if constexpr (kRed == 3) {
    xr = xr >= 8 * q ? xr - 8 * q : xr;
    xr = xr >= 4 * q ? xr - 4 * q : xr;
  }
  if constexpr (kRed != 0) xr = xr >= 2 * q ? xr - 2 * q : xr;
  const uint32_t nqn = mulhi32(y, wp) * q - w * y;
  x = xr - nqn;
  y = xr + 2 * q + nqn;
}

__device__ inline void fwd_bf(uint32_t& x, uint32_t& y, uint32_t w, uint32_t wp, uint32_t q) {
  fwd_bf_r<1>(x, y, w, wp, q);
}

template <int V>
using IntC = std::integral_constant<int, V>;

// The reductions of a forward radix-4 pass over stages (mm + 1, mm + 2), 1-based,
// for inputs below 2q: stage s leaves values below 2q + 2sq without reducing x,
// so stages 1-7 skip it (below 16q after stage 7: 16 q0 = 4,293,918,736 <
// 2^32, and q1 < q0), stage 8 takes x from below 16q to below 2q, and stages
// 9-11 are Harvey's.  Calls f(IntC<kRed of stage mm + 1>, IntC<kRed of stage
// mm + 2>); mm is even and uniform across the workgroup.
template <class F>
__device__ inline void fwd_stage_reduction(int mm, F&& f) {
  static_assert(16ull * kQ0 < (1ull << 32) && kQ1 < kQ0, "stages 1-7 stay below 2^32");
  if (mm <= 4)
    f(IntC<0>{}, IntC<0>{});
  else if (mm == 6)
    f(IntC<0>{}, IntC<3>{});
  else
    f(IntC<1>{}, IntC<1>{});
}

// Twiddle tables per prime: [fwd w][fwd w'][inv w][inv w'] x 2048, u32.
struct NttTables {
  const uint32_t* __restrict__ t;  // [2][4][kN]
  __device__ const uint32_t* fw(int c) const { return t + (c * 4 + 0) * kN; }
  __device__ const uint32_t* fwp(int c) const { return t + (c * 4 + 1) * kN; }
  __device__ const uint32_t* iw(int c) const { return t + (c * 4 + 2) * kN; }
  __device__ const uint32_t* iwp(int c) const { return t + (c * 4 + 3) * kN; }
};

// Forward radix-4 over NP polynomials (padded LDS layout) with per-polynomial
// primes.  Caller syncs before (inputs canonical); returns after a sync with
// outputs canonical.  Every thread of the block participates.
// kFirstMM and kLastStage as in ntt_forward_r4_1p below.  With kOpaque the
// unit index is opaque to the compiler, so a caller's loop does not hoist the
// passes' lane addresses (and spill them) across its iterations.
template <int NT, int NP, int kFirstMM = 0, bool kLastStage = true, bool kOpaque = false>
__device__ inline void ntt_forward_r4m(uint32_t* const (&polys)[NP], const int (&prime_of)[NP], const NttTables& tw) {
  const int tid = threadIdx.x;
  auto bf = [](uint32_t& x, uint32_t& y, uint32_t w, uint32_t wp, uint32_t q) { fwd_bf(x, y, w, wp, q); };
  // Stages (mm, mm+1): t = N/2 >> mm, units (i < m, j < t/2) at base 2t*i + j.
  for (int mm = kFirstMM; mm + 1 < kLogN; mm += 2) {
    const int m = 1 << mm, t = (kN / 2) >> mm, th = t >> 1;
    for (int u = tid; u < kN / 4; u += NT) {
      if constexpr (kOpaque) asm volatile("" : "+v"(u));
      int i, j;
      if (th >= 32) {
        i = u / th;
        j = u % th;
      } else {
        i = u % m;
        j = u / m;
      }
      const int x0 = 2 * t * i + j;
      const int l0 = lds_index(x0), l1 = lds_index(x0 + th), l2 = lds_index(x0 + t), l3 = lds_index(x0 + t + th);
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
      fwd_stage_reduction(mm, [&](auto r1, auto r2) {
        constexpr int kR1 = decltype(r1)::value, kR2 = decltype(r2)::value;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          const int c = prime_of[p];
          const uint32_t q = prime(c);
          uint32_t* a = polys[p];
          uint32_t a0 = a[l0], a1 = a[l1], a2 = a[l2], a3 = a[l3];
          fwd_bf_r<kR1>(a0, a2, w[c][0], wp[c][0], q);
          fwd_bf_r<kR1>(a1, a3, w[c][0], wp[c][0], q);
          fwd_bf_r<kR2>(a0, a1, w[c][1], wp[c][1], q);
          fwd_bf_r<kR2>(a2, a3, w[c][2], wp[c][2], q);
          a[l0] = a0;
          a[l1] = a1;
          a[l2] = a2;
          a[l3] = a3;
        }
      });
    }
    __syncthreads();
  }
  if constexpr (!kLastStage) return;
  // last stage (mm = 10, t = 1) and the final reduction to [0, q)
  for (int u = tid; u < kN / 2; u += NT) {
    uint32_t w[2], wp[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      w[c] = tw.fw(c)[1024 + u];
      wp[c] = tw.fwp(c)[1024 + u];
    }
    const int l0 = lds_index(2 * u), l1 = lds_index(2 * u + 1);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int c = prime_of[p];
      const uint32_t q = prime(c), q2 = 2 * q;
      uint32_t* a = polys[p];
      uint32_t x = a[l0], y = a[l1];
      bf(x, y, w[c], wp[c], q);
      x = x >= q2 ? x - q2 : x;
      x = x >= q ? x - q : x;
      y = y >= q2 ? y - q2 : y;
      y = y >= q ? y - q : y;
      a[l0] = x;
      a[l1] = y;
    }
  }
  __syncthreads();
}

// Inverse radix-4 over NP polynomials (padded LDS layout) with per-polynomial
// primes.  Caller syncs before (inputs canonical); returns after a sync with
// outputs canonical.
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
    const int l0 = lds_index(2 * u), l1 = lds_index(2 * u + 1);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int c = prime_of[p];
      uint32_t* a = polys[p];
      uint32_t x = a[l0], y = a[l1];
      bf(x, y, tw.iw(c)[1024 + u], tw.iwp(c)[1024 + u], prime(c));
      a[l0] = x;
      a[l1] = y;
    }
  }
  __syncthreads();
  // stage pairs (mm, mm-1) for mm = 9, 7, 5, 3, 1: t = N/2 >> mm, units
  // (i < 512/t, j < t) at base 4t*i + j.
  for (int mm = kLogN - 2; mm >= 1; mm -= 2) {
    const int h = 1 << mm, t = (kN / 2) >> mm, ni = (kN / 4) / t;
    for (int u = tid; u < kN / 4; u += NT) {
      int i, j;
      if (t >= 32) {
        i = u / t;
        j = u % t;
      } else {
        i = u % ni;
        j = u / ni;
      }
      const int x0 = 4 * t * i + j;
      const int l0 = lds_index(x0), l1 = lds_index(x0 + t), l2 = lds_index(x0 + 2 * t), l3 = lds_index(x0 + 3 * t);
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        const int c = prime_of[p];
        const uint32_t q = prime(c);
        const uint32_t* iw = tw.iw(c);
        const uint32_t* iwp = tw.iwp(c);
        uint32_t* a = polys[p];
        uint32_t a0 = a[l0], a1 = a[l1], a2 = a[l2], a3 = a[l3];
        bf(a0, a1, iw[h + 2 * i], iwp[h + 2 * i], q);
        bf(a2, a3, iw[h + 2 * i + 1], iwp[h + 2 * i + 1], q);
        bf(a0, a2, iw[(h >> 1) + i], iwp[(h >> 1) + i], q);
        bf(a1, a3, iw[(h >> 1) + i], iwp[(h >> 1) + i], q);
        a[l0] = a0 >= q ? a0 - q : a0;
        a[l1] = a1 >= q ? a1 - q : a1;
        a[l2] = a2 >= q ? a2 - q : a2;
        a[l3] = a3 >= q ? a3 - q : a3;
      }
    }
    __syncthreads();
  }
}

// Single-prime forms: NP polynomials, all over prime c (twiddles of c only).
// Same transforms, layouts and sync contract as the forms above.  With
// kFirstMM = 2 the forward transform starts at its second radix-4 pass: the
// caller ran the first one (ntt_forward_first_pass) and synced.  Without
// kLastStage it returns (after a sync) before the last, radix-2 stage and
// its final reduction, which the caller runs on its own slot pairs.
template <int NT, int NP, int kFirstMM = 0, bool kLastStage = true>
__device__ inline void ntt_forward_r4_1p(uint32_t* const (&polys)[NP], int c, const NttTables& tw) {
  const int tid = threadIdx.x;
  const uint32_t q = prime(c), q2 = 2 * q;
  const uint32_t* fw = tw.fw(c);
  const uint32_t* fwp = tw.fwp(c);
  auto bf = [q](uint32_t& x, uint32_t& y, uint32_t w, uint32_t wp) { fwd_bf(x, y, w, wp, q); };
  for (int mm = kFirstMM; mm + 1 < kLogN; mm += 2) {
    const uint32_t m = 1u << mm, t = (kN / 2) >> mm, th = t >> 1;
    for (uint32_t u = tid; u < kN / 4; u += NT) {
      uint32_t i, j;
      if (th >= 32) {
        i = u / th;
        j = u % th;
      } else {
        i = u % m;
        j = u / m;
      }
      const int x0 = (int)(2 * t * i + j), ti = (int)t, thi = (int)th;
      const int l0 = lds_index(x0), l1 = lds_index(x0 + thi), l2 = lds_index(x0 + ti), l3 = lds_index(x0 + ti + thi);
      const uint32_t w0 = tab_at(fw, m + i), wp0 = tab_at(fwp, m + i);
      const uint32_t w1 = tab_at(fw, 2 * m + 2 * i), wp1 = tab_at(fwp, 2 * m + 2 * i);
      const uint32_t w2 = tab_at(fw, 2 * m + 2 * i + 1), wp2 = tab_at(fwp, 2 * m + 2 * i + 1);
      fwd_stage_reduction(mm, [&](auto r1, auto r2) {
        constexpr int kR1 = decltype(r1)::value, kR2 = decltype(r2)::value;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          uint32_t* a = polys[p];
          uint32_t a0 = a[l0], a1 = a[l1], a2 = a[l2], a3 = a[l3];
          fwd_bf_r<kR1>(a0, a2, w0, wp0, q);
          fwd_bf_r<kR1>(a1, a3, w0, wp0, q);
          fwd_bf_r<kR2>(a0, a1, w1, wp1, q);
          fwd_bf_r<kR2>(a2, a3, w2, wp2, q);
          a[l0] = a0;
          a[l1] = a1;
          a[l2] = a2;
          a[l3] = a3;
        }
      });
    }
    __syncthreads();
  }
  if constexpr (!kLastStage) return;
  for (int u = tid; u < kN / 2; u += NT) {
    const uint32_t w = fw[1024 + u], wp = fwp[1024 + u];
    const int l0 = lds_index(2 * u), l1 = lds_index(2 * u + 1);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      uint32_t* a = polys[p];
      uint32_t x = a[l0], y = a[l1];
      bf(x, y, w, wp);
      x = x >= q2 ? x - q2 : x;
      x = x >= q ? x - q : x;
      y = y >= q2 ? y - q2 : y;
      y = y >= q ? y - q : y;
      a[l0] = x;
      a[l1] = y;
    }
  }
  __syncthreads();
}

template <int NT, int NP>
__device__ inline void ntt_inverse_r4_1p(uint32_t* const (&polys)[NP], int c, const NttTables& tw) {
  const int tid = threadIdx.x;
  const uint32_t q = prime(c), q2 = 2 * q;
  const uint32_t* iw = tw.iw(c);
  const uint32_t* iwp = tw.iwp(c);
  auto bf = [q2, q](uint32_t& x, uint32_t& y, uint32_t w, uint32_t wp) {
    const uint32_t tt = q2 - y + x;
    uint32_t cx = x + y;
    cx = cx >= q2 ? cx - q2 : cx;
    x = (cx + ((cx & 1u) ? q : 0u)) >> 1;
    y = w * tt - mulhi32(tt, wp) * q;
  };
  for (int u = tid; u < kN / 2; u += NT) {
    const int l0 = lds_index(2 * u), l1 = lds_index(2 * u + 1);
    const uint32_t w = iw[1024 + u], wp = iwp[1024 + u];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      uint32_t* a = polys[p];
      uint32_t x = a[l0], y = a[l1];
      bf(x, y, w, wp);
      a[l0] = x;
      a[l1] = y;
    }
  }
  __syncthreads();
  for (int mm = kLogN - 2; mm >= 1; mm -= 2) {
    const int h = 1 << mm, t = (kN / 2) >> mm, ni = (kN / 4) / t;
    for (int u = tid; u < kN / 4; u += NT) {
      int i, j;
      if (t >= 32) {
        i = u / t;
        j = u % t;
      } else {
        i = u % ni;
        j = u / ni;
      }
      const int x0 = 4 * t * i + j;
      const int l0 = lds_index(x0), l1 = lds_index(x0 + t), l2 = lds_index(x0 + 2 * t), l3 = lds_index(x0 + 3 * t);
      const uint32_t wa = iw[h + 2 * i], wpa = iwp[h + 2 * i];
      const uint32_t wb = iw[h + 2 * i + 1], wpb = iwp[h + 2 * i + 1];
      const uint32_t wc = iw[(h >> 1) + i], wpc = iwp[(h >> 1) + i];
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        uint32_t* a = polys[p];
        uint32_t a0 = a[l0], a1 = a[l1], a2 = a[l2], a3 = a[l3];
        bf(a0, a1, wa, wpa);
        bf(a2, a3, wb, wpb);
        bf(a0, a2, wc, wpc);
        bf(a1, a3, wc, wpc);
        a[l0] = a0 >= q ? a0 - q : a0;
        a[l1] = a1 >= q ? a1 - q : a1;
        a[l2] = a2 >= q ? a2 - q : a2;
        a[l3] = a3 >= q ? a3 - q : a3;
      }
    }
    __syncthreads();
  }
}

// Inverse transforms without the per-stage halving, for callers that apply
// n^-1 to the input themselves (the RGSW fold folds it into its product
// rows).  Gentleman-Sande with the inverse powers w and w' = floor(w*2^32/q)
// (DeviceConstants::fold, [prime][3][kN]: w, w', -); values stay in [0, 2q)
// through the passes and end canonical.  Same layouts and sync contract as the
// forms above; the result is n times theirs.
__device__ inline void inv_bf_unscaled(uint32_t& x, uint32_t& y, uint32_t w, uint32_t wp, uint32_t q) {
  const uint32_t q2 = 2 * q;
  const uint32_t tt = x + q2 - y;  // (0, 4q)
  const uint32_t cx = x + y;       // [0, 4q)
  x = cx >= q2 ? cx - q2 : cx;
  y = w * tt - mulhi32(tt, wp) * q;
}

template <int NT, int NP, bool OnePrime, int kLastMM = 1, bool kFirstStage = true, bool kOpaque = false>
__device__ inline void ntt_inverse_unscaled_impl(uint32_t* const (&polys)[NP], const int* prime_of, int c1,
                                                 const uint32_t* __restrict__ fold_tab) {
  const int tid = threadIdx.x;
  auto prime_at = [&](int p) { return OnePrime ? c1 : prime_of[p]; };
  for (int u = tid; kFirstStage && u < kN / 2; u += NT) {
    const int l0 = lds_index(2 * u), l1 = lds_index(2 * u + 1);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int c = prime_at(p);
      const uint32_t* iw = fold_tab + c * 3 * kN;
      uint32_t* a = polys[p];
      uint32_t x = a[l0], y = a[l1];
      inv_bf_unscaled(x, y, iw[1024 + u], iw[kN + 1024 + u], prime(c));
      a[l0] = x;
      a[l1] = y;
    }
  }
  if constexpr (kFirstStage) __syncthreads();
  for (int mm = kLogN - 2; mm >= kLastMM; mm -= 2) {
    const uint32_t h = 1u << mm, t = (kN / 2) >> mm, ni = (kN / 4) / t;
    for (uint32_t u = tid; u < kN / 4; u += NT) {
      if constexpr (kOpaque) asm volatile("" : "+v"(u));  // as in ntt_forward_r4m
      uint32_t i, j;
      if (t >= 32) {
        i = u / t;
        j = u % t;
      } else {
        i = u % ni;
        j = u / ni;
      }
      const int x0 = (int)(4 * t * i + j), ti = (int)t;
      const int l0 = lds_index(x0), l1 = lds_index(x0 + ti), l2 = lds_index(x0 + 2 * ti), l3 = lds_index(x0 + 3 * ti);
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        const int c = prime_at(p);
        const uint32_t q = prime(c), q2 = 2 * q;
        const uint32_t* iw = fold_tab + c * 3 * kN;
        const uint32_t* iwp = iw + kN;
        uint32_t* a = polys[p];
        uint32_t a0 = a[l0], a1 = a[l1], a2 = a[l2], a3 = a[l3];
        inv_bf_unscaled(a0, a1, tab_at(iw, h + 2 * i), tab_at(iwp, h + 2 * i), q);
        inv_bf_unscaled(a2, a3, tab_at(iw, h + 2 * i + 1), tab_at(iwp, h + 2 * i + 1), q);
        inv_bf_unscaled(a0, a2, tab_at(iw, (h >> 1) + i), tab_at(iwp, (h >> 1) + i), q);
        inv_bf_unscaled(a1, a3, tab_at(iw, (h >> 1) + i), tab_at(iwp, (h >> 1) + i), q);
        if (mm == 1) {  // last pass: canonical
          a0 = a0 >= q2 ? a0 - q2 : a0;
          a1 = a1 >= q2 ? a1 - q2 : a1;
          a2 = a2 >= q2 ? a2 - q2 : a2;
          a3 = a3 >= q2 ? a3 - q2 : a3;
          a0 = a0 >= q ? a0 - q : a0;
          a1 = a1 >= q ? a1 - q : a1;
          a2 = a2 >= q ? a2 - q : a2;
          a3 = a3 >= q ? a3 - q : a3;
        }
        a[l0] = a0;
        a[l1] = a1;
        a[l2] = a2;
        a[l3] = a3;
      }
    }
    __syncthreads();
  }
}

// Two-prime forms for 1024-thread blocks over 2 NPH consecutive padded
// polynomials at `base`: threads [512 c, 512 c + 512) transform polynomials
// c NPH .. c NPH + NPH - 1, all over prime c, so every thread has a radix-4
// unit in every pass (the per-prime forms above leave threads 512.. idle in
// the radix-4 passes).  Same transforms, layouts and sync contract as
// ntt_forward_r4m and ntt_inverse_r4_unscaled.
template <int NPH>
__device__ inline void ntt_forward_r4_halves(uint32_t* base, const NttTables& tw) {
  int u = threadIdx.x & 511;
  // opaque to the compiler: the unit's addresses are recomputed per call
  // instead of being hoisted out of the caller's loop and spilled
  asm volatile("" : "+v"(u));
  const int c = threadIdx.x >> 9;
  const uint32_t q = prime(c), q2 = 2 * q;
  const uint32_t* fw = tw.fw(c);
  const uint32_t* fwp = tw.fwp(c);
  auto bf = [q](uint32_t& x, uint32_t& y, uint32_t w, uint32_t wp) { fwd_bf(x, y, w, wp, q); };
#pragma unroll 1
  for (int mm = 0; mm + 1 < kLogN; mm += 2) {
    const int m = 1 << mm, t = (kN / 2) >> mm, th = t >> 1;
    int i, j;
    if (th >= 32) {
      i = u / th;
      j = u % th;
    } else {
      i = u % m;
      j = u / m;
    }
    const int x0 = 2 * t * i + j;
    const int l0 = lds_index(x0), l1 = lds_index(x0 + th), l2 = lds_index(x0 + t), l3 = lds_index(x0 + t + th);
    const uint32_t w0 = fw[m + i], wp0 = fwp[m + i];
    const uint32_t w1 = fw[2 * m + 2 * i], wp1 = fwp[2 * m + 2 * i];
    const uint32_t w2 = fw[2 * m + 2 * i + 1], wp2 = fwp[2 * m + 2 * i + 1];
#pragma unroll
    for (int p = 0; p < NPH; ++p) {
      uint32_t* a = base + (c * NPH + p) * kNPad;
      uint32_t a0 = a[l0], a1 = a[l1], a2 = a[l2], a3 = a[l3];
      bf(a0, a2, w0, wp0);
      bf(a1, a3, w0, wp0);
      bf(a0, a1, w1, wp1);
      bf(a2, a3, w2, wp2);
      a[l0] = a0;
      a[l1] = a1;
      a[l2] = a2;
      a[l3] = a3;
    }
    __syncthreads();
  }
  // last stage (t = 1), two units per thread, and the final reduction
  for (int v = u; v < kN / 2; v += 512) {
    const uint32_t w = fw[1024 + v], wp = fwp[1024 + v];
    const int l0 = lds_index(2 * v), l1 = lds_index(2 * v + 1);
#pragma unroll
    for (int p = 0; p < NPH; ++p) {
      uint32_t* a = base + (c * NPH + p) * kNPad;
      uint32_t x = a[l0], y = a[l1];
      bf(x, y, w, wp);
      x = x >= q2 ? x - q2 : x;
      x = x >= q ? x - q : x;
      y = y >= q2 ? y - q2 : y;
      y = y >= q ? y - q : y;
      a[l0] = x;
      a[l1] = y;
    }
  }
  __syncthreads();
}

template <int NPH>
__device__ inline void ntt_inverse_r4_halves_unscaled(uint32_t* base, const uint32_t* __restrict__ fold_tab) {
  int u = threadIdx.x & 511;
  asm volatile("" : "+v"(u));  // as in ntt_forward_r4_halves
  const int c = threadIdx.x >> 9;
  const uint32_t q = prime(c), q2 = 2 * q;
  const uint32_t* iw = fold_tab + c * 3 * kN;
  const uint32_t* iwp = iw + kN;
  for (int v = u; v < kN / 2; v += 512) {
    const int l0 = lds_index(2 * v), l1 = lds_index(2 * v + 1);
#pragma unroll
    for (int p = 0; p < NPH; ++p) {
      uint32_t* a = base + (c * NPH + p) * kNPad;
      uint32_t x = a[l0], y = a[l1];
      inv_bf_unscaled(x, y, iw[1024 + v], iwp[1024 + v], q);
      a[l0] = x;
      a[l1] = y;
    }
  }
  __syncthreads();
#pragma unroll 1
  for (int mm = kLogN - 2; mm >= 1; mm -= 2) {
    const int h = 1 << mm, t = (kN / 2) >> mm, ni = (kN / 4) / t;
    int i, j;
    if (t >= 32) {
      i = u / t;
      j = u % t;
    } else {
      i = u % ni;
      j = u / ni;
    }
    const int x0 = 4 * t * i + j;
    const int l0 = lds_index(x0), l1 = lds_index(x0 + t), l2 = lds_index(x0 + 2 * t), l3 = lds_index(x0 + 3 * t);
    const uint32_t wa = iw[h + 2 * i], wpa = iwp[h + 2 * i];
    const uint32_t wb = iw[h + 2 * i + 1], wpb = iwp[h + 2 * i + 1];
    const uint32_t wc = iw[(h >> 1) + i], wpc = iwp[(h >> 1) + i];
#pragma unroll
    for (int p = 0; p < NPH; ++p) {
      uint32_t* a = base + (c * NPH + p) * kNPad;
      uint32_t a0 = a[l0], a1 = a[l1], a2 = a[l2], a3 = a[l3];
      inv_bf_unscaled(a0, a1, wa, wpa, q);
      inv_bf_unscaled(a2, a3, wb, wpb, q);
      inv_bf_unscaled(a0, a2, wc, wpc, q);
      inv_bf_unscaled(a1, a3, wc, wpc, q);
      if (mm == 1) {  // last pass: canonical
        a0 = a0 >= q2 ? a0 - q2 : a0;
        a1 = a1 >= q2 ? a1 - q2 : a1;
        a2 = a2 >= q2 ? a2 - q2 : a2;
        a3 = a3 >= q2 ? a3 - q2 : a3;
        a0 = a0 >= q ? a0 - q : a0;
        a1 = a1 >= q ? a1 - q : a1;
        a2 = a2 >= q ? a2 - q : a2;
        a3 = a3 >= q ? a3 - q : a3;
      }
      a[l0] = a0;
      a[l1] = a1;
      a[l2] = a2;
      a[l3] = a3;
    }
    __syncthreads();
  }
}

// The transform of the third unsigned base-2^19 gadget digit (gadget.rs:35-61)
// of a polynomial x with coefficients below q < 2^57, so that d0 + 2^19 d1 +
// 2^38 d2 = x exactly, from S = 2^-kM NTT_c(x) and the other two digits'
// transforms G0, G1, by linearity:
//     NTT_c(d2) = 2^(kM-38) S - 2^-38 G0 - 2^-19 G1   (mod q_c).
// Inputs any representatives below 2^32 (each enters one Shoup product);
// returns the canonical value, or one below 2q without kCanonical.
template <int kM, bool kCanonical = true>
__device__ inline uint32_t third_gadget_digit_ntt(uint32_t S, uint32_t g0, uint32_t g1, int c) {
  static_assert(3 * kGadgetBits >= 57, "three digits must cover q");
  constexpr int kE = 2 * kGadgetBits - kM;
  constexpr uint32_t kA0 = inv_pow2_mod(kQ0, kE), kA1 = inv_pow2_mod(kQ1, kE);
  constexpr uint32_t kB0 = inv_pow2_mod(kQ0, 2 * kGadgetBits), kB1 = inv_pow2_mod(kQ1, 2 * kGadgetBits);
  constexpr uint32_t kC0 = inv_pow2_mod(kQ0, kGadgetBits), kC1 = inv_pow2_mod(kQ1, kGadgetBits);
  const uint32_t q = prime(c);
  const uint32_t a = shoup_lazy(S, c ? kA1 : kA0, c ? shoup_companion(kA1, kQ1) : shoup_companion(kA0, kQ0), q);
  const uint32_t b = shoup_lazy(g0, c ? kB1 : kB0, c ? shoup_companion(kB1, kQ1) : shoup_companion(kB0, kQ0), q);
  const uint32_t e = shoup_lazy(g1, c ? kC1 : kC0, c ? shoup_companion(kC1, kQ1) : shoup_companion(kC0, kQ0), q);
  uint32_t x = a + 4 * q - b - e;  // (0, 6q)
  x = x >= 4 * q ? x - 4 * q : x;
  x = x >= 2 * q ? x - 2 * q : x;
  if constexpr (!kCanonical) return x;  // [0, 2q)
  return x >= q ? x - q : x;
}

// NP polynomials over prime c.  With kLastMM = 3 the transform stops before
// its last radix-4 pass (after a sync); the caller runs that pass on its own
// units (ntt_inverse_last_pass_unscaled).  Without kFirstStage it starts at
// its first radix-4 pass: the caller ran the first, radix-2 stage on its slot
// pairs and synced.
template <int NT, int NP, int kLastMM = 1, bool kFirstStage = true>
__device__ inline void ntt_inverse_r4_1p_unscaled(uint32_t* const (&polys)[NP], int c,
                                                  const uint32_t* __restrict__ fold_tab) {
  ntt_inverse_unscaled_impl<NT, NP, true, kLastMM, kFirstStage>(polys, nullptr, c, fold_tab);
}

// The first radix-4 pass of the forward transform (stages 1 and 2 in the
// numbering of fwd_stage_reduction, inputs below 2q) on one unit in registers:
// coefficients u, u + 512, u + 1024, u + 1536 of a polynomial over prime c,
// i.e. the unit of thread u in a 512-thread block.
__device__ inline void ntt_forward_first_pass(uint32_t (&a)[4], int c, const NttTables& tw) {
  const uint32_t q = prime(c);
  const uint32_t* fw = tw.fw(c);
  const uint32_t* fwp = tw.fwp(c);
  // fwd_stage_reduction(0, ..): no reduction of x
  fwd_bf_r<0>(a[0], a[2], fw[1], fwp[1], q);
  fwd_bf_r<0>(a[1], a[3], fw[1], fwp[1], q);
  fwd_bf_r<0>(a[0], a[1], fw[2], fwp[2], q);
  fwd_bf_r<0>(a[2], a[3], fw[3], fwp[3], q);
}

// The last radix-4 pass of the unscaled inverse (stages 1 and 0) on the same
// unit in registers; outputs canonical.
__device__ inline void ntt_inverse_last_pass_unscaled(uint32_t (&a)[4], int c, const uint32_t* __restrict__ fold_tab) {
  const uint32_t q = prime(c), q2 = 2 * q;
  const uint32_t* iw = fold_tab + c * 3 * kN;
  const uint32_t* iwp = iw + kN;
  inv_bf_unscaled(a[0], a[1], iw[2], iwp[2], q);
  inv_bf_unscaled(a[2], a[3], iw[3], iwp[3], q);
  inv_bf_unscaled(a[0], a[2], iw[1], iwp[1], q);
  inv_bf_unscaled(a[1], a[3], iw[1], iwp[1], q);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    a[i] = a[i] >= q2 ? a[i] - q2 : a[i];
    a[i] = a[i] >= q ? a[i] - q : a[i];
  }
}

// NP polynomials with per-polynomial primes.
template <int NT, int NP>
__device__ inline void ntt_inverse_r4_unscaled(uint32_t* const (&polys)[NP], const int (&prime_of)[NP],
                                               const uint32_t* __restrict__ fold_tab) {
  ntt_inverse_unscaled_impl<NT, NP, false>(polys, prime_of, 0, fold_tab);
}

}  // namespace inspire
