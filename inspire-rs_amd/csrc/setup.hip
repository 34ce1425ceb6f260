// Server offline precomputation on gfx950 (KeywordServer::setup,
// keyword.rs:427-458 -> perform_offline_precomputation_simplepir,
// server/engine.rs:678-1000): the hint over the SEED_0 query masks, prep-pack,
// and InspiRING's full_packing_with_preprocessing_offline (packing.rs:2477-2750),
// producing bold_t, bold_t_bar, bold_t_hat and a_hat for every output
// polynomial ("instance") of the encoded DB.
//
// 1. k_setup_atilde, one workgroup per DB column j: the packing input of the
//    column's LWE ciphertext is at_j = sum_k Dt_jk * abar_k, where abar_k is
//    the (negated) SEED_0 mask of query RLWE k and Dt_jk(X) = sum_i
//    DB[j][kN+i] X^-i; this equals negacyclic_perm of the hint column
//    (utils.rs:5-21, packing.rs:1032-1066).  Output: NTT residues.
// 2. k_setup_slot_ntt: per instance, prime and NTT slot z, a forward NTT over
//    the column index j of at_j[z].  With R_z[k] = sum_j at_j[z] psi^(e_k j)
//    (e_k = 2 rev(k) + 1, the evaluation point of slot k):
//      r_i[k]    = n^-1 R_{z}[k], e_z =  e_k g_i  (mod 2n),
//      rbar_i[k] = n^-1 R_{z}[k], e_z = -e_k g_i  (mod 2n),
//    which is the reference's r_i = tau_{g_i}(n^-1 sum_j mono(j g_-i) * at_j)
//    (and the bar form with 2n - j g_-i): the monomial sum in slot z is the
//    polynomial sum_j at_j[z] Y^j evaluated at psi^(e_z g_-i), and the
//    automorphism tau_{g_i} moves slot z to the slot k with e_k = e_z g_-i.
//    So the n^3 sum of the reference becomes n^2 log n.
// 3. k_setup_collapse, one workgroup per (instance, chain r or rbar): the
//    1023-step collapse (packing.rs:2694-2731), gathering r_i from R each step.
// 4. k_setup_ahat: a_hat = raw(r_0 + sum_k v_mask[k] * bold_t_hat[k]).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/inspire_hip.h"
#include "host_tables.hpp"
#include "kernels.hpp"
#include "ntt_device.hpp"
#include "runtime.hpp"

namespace inspire {
namespace {

constexpr int kAtThreads = 256;
constexpr int kSlotThreads = 512;
constexpr int kSlotBlock = 16;                      // NTT slots per k_setup_slot_ntt workgroup
constexpr int kColThreads = 512;
constexpr int kColSlots = kN / kColThreads;         // slots per thread in the collapse
constexpr uint32_t kFiveInv = 3277;                 // 5^-1 mod 4096
// n^-2 mod q_c: the scale of R in the collapse (n^-1 of r_i times the n^-1
// the chain carries)
constexpr uint32_t kNInv2Q0 = (uint32_t)((uint64_t)kNInvQ0 * kNInvQ0 % kQ0);
constexpr uint32_t kNInv2Q1 = (uint32_t)((uint64_t)kNInvQ1 * kNInvQ1 % kQ1);

__device__ inline uint32_t rev11d(uint32_t x) { return __brev(x) >> (32 - kLogN); }

// Slot of the evaluation point psi^(e_z * t) for slot z (NTT-domain automorphism
// X -> X^t: out[z] = in[auto_slot(t, z)], packing.rs:1500-1516).
__device__ inline int auto_slot(uint32_t t, int z) {
  const uint32_t e = ((2u * rev11d((uint32_t)z) + 1u) * t) & (2u * kN - 1u);
  return (int)rev11d((e - 1u) >> 1);
}

template <int c>
__device__ inline uint32_t neg_mod(uint32_t v) {
  return v ? prime(c) - v : 0u;
}

// kOne: one query RLWE row per column (db_rows = n), so the product is the
// output and no accumulator lives across the transform (8 waves per SIMD).
template <bool kOne>
__global__ __launch_bounds__(kAtThreads) __attribute__((amdgpu_waves_per_eu(kOne ? 8 : 4, 8))) void k_setup_atilde(
    uint32_t* __restrict__ at, const uint16_t* __restrict__ db, int db_rows, int64_t col0,
    const uint32_t* __restrict__ abar, const uint32_t* tw) {
  __shared__ uint32_t buf[2][kNPad];
  constexpr int kS = kN / kAtThreads;
  const NttTables T{tw};
  const int tid = threadIdx.x;
  const int64_t jl = blockIdx.x;
  const uint16_t* col = db + (col0 + jl) * (int64_t)db_rows;
  uint32_t* const polys[2] = {buf[0], buf[1]};
  const int primes[2] = {0, 1};
  uint32_t acc[2][kS];
#pragma unroll
  for (int s = 0; s < kS; ++s) acc[0][s] = acc[1][s] = 0;
  const int nrlwe = kOne ? 1 : db_rows / kN;
  for (int k = 0; k < nrlwe; ++k) {
    // Dt(X) = sum_i v_i X^-i: coefficient 0 is v_0, coefficient m >= 1 is -v_(n-m).
#pragma unroll
    for (int s = 0; s < kS; ++s) {
      const int x = tid + s * kAtThreads;
      const uint32_t v = col[(int64_t)k * kN + (x ? kN - x : 0)];
      buf[0][lds_index(x)] = x ? neg_mod<0>(v) : v;
      buf[1][lds_index(x)] = x ? neg_mod<1>(v) : v;
    }
    __syncthreads();
    ntt_forward_r4m<kAtThreads, 2>(polys, primes, T);
    const uint32_t* ab = abar + (size_t)k * 2 * kN;
    if constexpr (kOne) {
      uint32_t* o = at + jl * 2 * kN;
#pragma unroll
      for (int s = 0; s < kS; ++s) {
        const int x = tid + s * kAtThreads;
        o[x] = mulmod<0>(buf[0][lds_index(x)], ab[x]);
        o[kN + x] = mulmod<1>(buf[1][lds_index(x)], ab[kN + x]);
      }
      return;
    }
#pragma unroll
    for (int s = 0; s < kS; ++s) {
      const int x = tid + s * kAtThreads;
      acc[0][s] = reduce58<0>((uint64_t)buf[0][lds_index(x)] * ab[x] + acc[0][s]);
      acc[1][s] = reduce58<1>((uint64_t)buf[1][lds_index(x)] * ab[kN + x] + acc[1][s]);
    }
    __syncthreads();
  }
  uint32_t* o = at + jl * 2 * kN;
#pragma unroll
  for (int s = 0; s < kS; ++s) {
    const int x = tid + s * kAtThreads;
    o[x] = acc[0][s];
    o[kN + x] = acc[1][s];
  }
}

// R[g][c][z][k] = NTT over j of at[g*N + j][c][z]; kSlotBlock slots per workgroup.
__global__ __launch_bounds__(kSlotThreads) void k_setup_slot_ntt(uint32_t* __restrict__ R,
                                                                 const uint32_t* __restrict__ at, const uint32_t* tw) {
  extern __shared__ uint32_t lds[];  // [kSlotBlock][kNPad]
  const NttTables T{tw};
  const int tid = threadIdx.x;
  constexpr int kBlocksPerPrime = kN / kSlotBlock;
  const int zb = blockIdx.x % kBlocksPerPrime;
  const int c = (blockIdx.x / kBlocksPerPrime) & 1;
  const int64_t g = blockIdx.x / (2 * kBlocksPerPrime);
  const uint32_t* src = at + (g * kN * 2 + c) * kN + zb * kSlotBlock;
  // 16 consecutive slots (64 bytes) of each column j, as four 16-byte loads
  constexpr int kQuads = kSlotBlock / 4;
#pragma unroll 4
  for (int e = tid; e < kN * kQuads; e += kSlotThreads) {
    const int zq = (e % kQuads) * 4, j = e / kQuads;
    const uint4 v = *reinterpret_cast<const uint4*>(src + (int64_t)j * 2 * kN + zq);
    const int lj = lds_index(j);
    lds[(zq + 0) * kNPad + lj] = v.x;
    lds[(zq + 1) * kNPad + lj] = v.y;
    lds[(zq + 2) * kNPad + lj] = v.z;
    lds[(zq + 3) * kNPad + lj] = v.w;
  }
  __syncthreads();
  uint32_t* polys[kSlotBlock];
#pragma unroll
  for (int p = 0; p < kSlotBlock; ++p) polys[p] = lds + p * kNPad;
  ntt_forward_r4_1p<kSlotThreads, kSlotBlock>(polys, c, T);
  // Step order: the collapse of chain w reads, at step i and slot k, n^-2 R_z[k]
  // (r_i's n^-1 times the n^-1 its chain carries) with e_z = u e_k (mod 2n),
  // u = (-1)^w 5^i; it is stored at R[g][c][w n/2 +
  // i][k].  This workgroup's 16 slots z0.. are the exponents e = a (mod 256), and
  // multiplying by u maps that coset to the 16 consecutive slots k0(u).. with
  // e_k0 = a u^-1 (mod 256), so every (w, i) row gets one contiguous 64-byte run
  // from here and the collapse reads each step's row in order.
  const int z0 = zb * kSlotBlock;
  const uint32_t a = (2u * rev11d((uint32_t)z0) + 1u) & 255u;
  // each thread stores four consecutive slots (16 bytes) of rows rb + 128 it
  const int m4 = (tid % kQuads) * 4, rb = tid / kQuads;
  constexpr int kRowStep = kSlotThreads / kQuads;  // 128
  constexpr uint32_t kMask2n = 2u * kN - 1u;
  auto pow_mod_2n = [](uint32_t base, int e) {  // base^e mod 2n by squaring
    uint32_t r = 1;
    for (; e; e >>= 1, base = (base * base) & kMask2n)
      if (e & 1) r = (r * base) & kMask2n;
    return r;
  };
  const uint32_t p = pow_mod_2n(5u, rb), pinv = pow_mod_2n(kFiveInv, rb);
  const uint32_t step = pow_mod_2n(5u, kRowStep), step_inv = pow_mod_2n(kFiveInv, kRowStep);
  uint32_t* dst = R + (g * 2 + c) * (int64_t)kN * kN;
  for (int w = 0; w < 2; ++w) {
    uint32_t u5 = p, u5inv = pinv;  // 5^i and 5^-i for i = rb + 128 it
    for (int i = rb; i < kN / 2; i += kRowStep) {
      const uint32_t u = w ? (2u * kN - u5) & kMask2n : u5;
      const uint32_t uinv = w ? (2u * kN - u5inv) & kMask2n : u5inv;
      const uint32_t b = (a * uinv) & 255u;  // e_k0 mod 256 = 2 rev7(k0 / 16) + 1
      const int k0 = (int)(__brev((b - 1u) >> 1) >> 25) * kSlotBlock + m4;
      uint32_t o[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int k = k0 + t;
        const int zz = auto_slot(u, k) - z0;
        const uint32_t v = lds[zz * kNPad + lds_index(k)];
        o[t] = c ? mulmod<1>(v, kNInv2Q1) : mulmod<0>(v, kNInv2Q0);
      }
      *reinterpret_cast<uint4*>(dst + ((int64_t)w * (kN / 2) + i) * kN + k0) = uint4{o[0], o[1], o[2], o[3]};
      u5 = (u5 * step) & kMask2n;
      u5inv = (u5inv * step_inv) & kMask2n;
    }
  }
}

struct CollapseArgs {
  uint64_t* bt;         // [instances][3069][N] condensed (this group's first instance)
  uint64_t* btb;
  uint64_t* bth;        // [instances][3][N] condensed
  uint32_t* r0;         // [G][2][N] updated r_0 (NTT), read by k_setup_ahat
  const uint32_t* R;    // [G][2][N][N]: n^-2 R, collapse step order
  const uint32_t* w;    // n^-1 W mask [3][2][N] NTT residues
  const uint32_t* tw;
  const uint32_t* fold;  // unhalved inverse powers (build_fold_tables)
  uint32_t g_top;       // 5^(n/2 - 1) mod 2n
};

// One workgroup per (instance, chain r or rbar): the 1023-step collapse
// (packing.rs:2694-2731).  A step of chain r computes
//   bold_t[i][k] = NTT(G^-1(raw(r_(i+1))))[k],  r_i += sum_k w[i][k] bold_t[i][k]
// with r_i's own term gathered from R.  The chain carries S = n^-1 NTT_c(r)
// (R and W arrive scaled by n^-2 and n^-1), so the inverse transform runs
// without scaling, and the transform of the third gadget digit follows from S
// and the other two (third_gadget_digit_ntt<11>): 2 inverse and 4 forward
// transforms per step.
//
// Per step: the inverse passes of stages 2-9 over both primes in LDS; the
// last inverse pass, the CRT composition, the two digits and the first forward
// pass in registers (a thread's slots tid + 512 s are one unit of those
// passes); the forward passes of stages 3-10 in LDS; then, on slot pairs
// (2v, 2v + 1), v = tid + 512 h, in registers: the last forward stage, the
// third digit, the bold_t writes, the W products, and the first inverse stage
// of the next step.
__global__ __launch_bounds__(kColThreads) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_setup_collapse(
    CollapseArgs a) {
  // row 2k + c: digit k over prime c (forward); rows 0 and 1 also hold the
  // inverse transform of prime 0 and 1; row 4 + c: S over prime c (each
  // thread reads and writes only its own pair slots there)
  __shared__ uint32_t dig[6][kNPad];
  const NttTables T{a.tw};
  const int tid = threadIdx.x;
  const int64_t g = blockIdx.x >> 1;
  const int which = blockIdx.x & 1;  // 0: r chain (bold_t), 1: rbar chain (bold_t_bar)
  const uint32_t* Rg = a.R + g * 2 * (int64_t)kN * kN;
  uint64_t* out = (which ? a.btb : a.bt) + g * (int64_t)kNInner * kN;
  uint32_t* const fpolys[4] = {dig[0], dig[1], dig[2], dig[3]};
  const int fprime[4] = {0, 1, 0, 1};
  uint32_t* const ipolys[2] = {dig[0], dig[1]};
  const int iprime[2] = {0, 1};
  const uint64_t mask = (1ull << kGadgetBits) - 1;

  // rr[c][2h + e] at slot 2(tid + 512 h) + e (indices constant or selected by
  // h, so it stays in registers).
  uint32_t rr[2][4];
  auto load_r = [&](int i) {  // n^-2 R of step i at the pair slots
    const int64_t row = (int64_t)which * (kN / 2) + i;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const uint2 v2 = *reinterpret_cast<const uint2*>(Rg + ((int64_t)c * kN + row) * kN + 2 * (tid + h * kColThreads));
        rr[c][2 * h] = v2.x;
        rr[c][2 * h + 1] = v2.y;
      }
  };
  // First inverse stage of S at pair h: rows 0 and 1 (Harvey, inputs < 2q).
  auto first_inverse = [&](int h, const uint32_t (&s0)[2], const uint32_t (&s1)[2]) {
    int v = tid + h * kColThreads;
    asm volatile("" : "+v"(v));  // as in the passes (kOpaque)
    const int l0 = lds_index(2 * v), l1 = lds_index(2 * v + 1);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const uint32_t* iw = a.fold + c * 3 * kN;
      uint32_t x = c ? s1[0] : s0[0], y = c ? s1[1] : s0[1];
      inv_bf_unscaled(x, y, iw[1024 + v], iw[kN + 1024 + v], prime(c));
      dig[c][l0] = x;
      dig[c][l1] = y;
    }
  };
  // From rows 0 and 1 after the first inverse stage to the forward digit
  // transforms before their last stage (rows 0-3).
  auto transform = [&]() {
    __syncthreads();
    ntt_inverse_unscaled_impl<kColThreads, 2, false, 3, false, true>(ipolys, iprime, 0, a.fold);
    uint32_t y0[kColSlots], y1[kColSlots];
#pragma unroll
    for (int s = 0; s < kColSlots; ++s) {
      const int lz = lds_index(tid + s * kColThreads);
      y0[s] = dig[0][lz];
      y1[s] = dig[1][lz];
    }
    ntt_inverse_last_pass_unscaled(y0, 0, a.fold);
    ntt_inverse_last_pass_unscaled(y1, 1, a.fold);
    // G^-1: the first two unsigned 19-bit digits (< 2^19, canonical for both
    // primes); each thread rewrites only its own slots here
    uint32_t d[2][kColSlots];
#pragma unroll
    for (int s = 0; s < kColSlots; ++s) {
      const uint64_t raw = crt_compose(y0[s], y1[s]);
      d[0][s] = (uint32_t)(raw & mask);
      d[1][s] = (uint32_t)((raw >> kGadgetBits) & mask);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        uint32_t u[kColSlots];
#pragma unroll
        for (int s = 0; s < kColSlots; ++s) u[s] = d[k][s];
        ntt_forward_first_pass(u, c, T);
#pragma unroll
        for (int s = 0; s < kColSlots; ++s) dig[2 * k + c][lds_index(tid + s * kColThreads)] = u[s];
      }
    __syncthreads();
    ntt_forward_r4m<kColThreads, 4, 2, false, true>(fpolys, fprime, T);
  };
  // Last forward stage at pair h: the three digit transforms gk[k][c][e],
  // canonical (the third derived from S).
  auto digits_at = [&](int h, uint32_t (&gk)[3][2][2]) {
    int v = tid + h * kColThreads;
    asm volatile("" : "+v"(v));
    const int l0 = lds_index(2 * v), l1 = lds_index(2 * v + 1);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const uint32_t q = prime(c), q2 = 2 * q;
      const uint32_t w = T.fw(c)[1024 + v], wp = T.fwp(c)[1024 + v];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        uint32_t x = dig[2 * k + c][l0], y = dig[2 * k + c][l1];
        fwd_bf(x, y, w, wp, q);
        x = x >= q2 ? x - q2 : x;
        x = x >= q ? x - q : x;
        y = y >= q2 ? y - q2 : y;
        y = y >= q ? y - q : y;
        gk[k][c][0] = x;
        gk[k][c][1] = y;
      }
#pragma unroll
      for (int e = 0; e < 2; ++e)
        gk[2][c][e] = third_gadget_digit_ntt<kLogN>(dig[4 + c][e ? l1 : l0], gk[0][c][e], gk[1][c][e], c);
    }
  };

  load_r(kN / 2 - 1);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int v = tid + h * kColThreads;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      dig[4 + c][lds_index(2 * v)] = rr[c][2 * h];
      dig[4 + c][lds_index(2 * v + 1)] = rr[c][2 * h + 1];
    }
    first_inverse(h, {rr[0][2 * h], rr[0][2 * h + 1]}, {rr[1][2 * h], rr[1][2 * h + 1]});
  }
  uint32_t gi = a.g_top;  // g_(n/2 - 1)
  for (int i = kN / 2 - 2; i >= 0; --i) {
    gi = (gi * kFiveInv) & (2u * kN - 1u);  // g_i = 5^i mod 2n
    const uint32_t t = which ? 2u * kN - gi : gi;
    load_r(i);  // independent of the chain: in flight during the transforms
    transform();
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      const int v = tid + h * kColThreads;
      uint32_t off = 2u * (uint32_t)v;
      asm volatile("" : "+v"(off));  // keeps the lane offset 32-bit
      uint32_t gk[3][2][2];
      digits_at(h, gk);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        ulonglong2 pr;
        pr.x = (uint64_t)gk[k][0][0] | ((uint64_t)gk[k][1][0] << 32);
        pr.y = (uint64_t)gk[k][0][1] | ((uint64_t)gk[k][1][1] << 32);
        *reinterpret_cast<ulonglong2*>(out + ((int64_t)i * kTExp + k) * kN + off) = pr;
      }
      // S <- n^-1 NTT(r_i) = n^-2 R + sum_k n^-1 W(t_i) bold_t[i][k] (the W mask
      // rotated by t_i is a cache-resident gather)
      uint32_t sn[2][2];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int zs = auto_slot(t, 2 * v + e);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          uint64_t acc = h ? rr[c][2 + e] : rr[c][e];
#pragma unroll
          for (int k = 0; k < 3; ++k) acc += (uint64_t)a.w[(k * 2 + c) * kN + zs] * gk[k][c][e];
          sn[c][e] = c ? reduce58<1>(acc) : reduce58<0>(acc);
        }
      }
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        dig[4 + c][lds_index(2 * v)] = sn[c][0];
        dig[4 + c][lds_index(2 * v + 1)] = sn[c][1];
      }
      first_inverse(h, sn[0], sn[1]);
    }
  }
  if (which == 0) {
    // r_0 = n S
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int z = 2 * (tid + h * kColThreads) + e;
          const uint32_t sv = dig[4 + c][lds_index(z)];
          a.r0[(g * 2 + c) * kN + z] = c ? mulmod<1>(sv, (uint32_t)kN) : mulmod<0>(sv, (uint32_t)kN);
        }
    return;
  }
  // bold_t_hat = NTT(G^-1(raw(rbar_0)))  (packing.rs:2733-2734)
  transform();
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int v = tid + h * kColThreads;
    uint32_t gk[3][2][2];
    digits_at(h, gk);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      ulonglong2 pr;
      pr.x = (uint64_t)gk[k][0][0] | ((uint64_t)gk[k][1][0] << 32);
      pr.y = (uint64_t)gk[k][0][1] | ((uint64_t)gk[k][1][1] << 32);
      *reinterpret_cast<ulonglong2*>(a.bth + (g * kTExp + k) * kN + 2 * v) = pr;
    }
  }
}

// a_hat = raw(r_0 + sum_k v_mask[k] * bold_t_hat[k])  (packing.rs:2736-2749)
__global__ __launch_bounds__(kColThreads) void k_setup_ahat(uint64_t* __restrict__ a_hat, const uint32_t* __restrict__ r0,
                                                            const uint64_t* __restrict__ bth,
                                                            const uint32_t* __restrict__ v, const uint32_t* tw) {
  __shared__ uint32_t buf[2][kNPad];
  const NttTables T{tw};
  const int tid = threadIdx.x;
  const int64_t g = blockIdx.x;
#pragma unroll
  for (int s = 0; s < kColSlots; ++s) {
    const int z = tid + s * kColThreads;
    uint64_t acc0 = r0[(g * 2 + 0) * kN + z], acc1 = r0[(g * 2 + 1) * kN + z];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const uint64_t h = bth[(g * kTExp + k) * kN + z];
      acc0 += (uint64_t)v[(k * 2 + 0) * kN + z] * (uint32_t)h;
      acc1 += (uint64_t)v[(k * 2 + 1) * kN + z] * (uint32_t)(h >> 32);
    }
    buf[0][lds_index(z)] = reduce58<0>(acc0);
    buf[1][lds_index(z)] = reduce58<1>(acc1);
  }
  __syncthreads();
  uint32_t* const p[2] = {buf[0], buf[1]};
  const int pr[2] = {0, 1};
  ntt_inverse_r4<kColThreads, 2>(p, pr, T);
#pragma unroll
  for (int s = 0; s < kColSlots; ++s) {
    const int z = tid + s * kColThreads;
    a_hat[g * kN + z] = crt_compose(buf[0][lds_index(z)], buf[1][lds_index(z)]);
  }
}

uint32_t pow5_mod_2n(int e) {
  uint32_t r = 1;
  for (int i = 0; i < e; ++i) r = (r * 5u) & (2u * kN - 1u);
  return r;
}

}  // namespace

size_t setup_offline_bytes_per_instance() { return (size_t)2 * 2 * kN * kN * sizeof(uint32_t) + 2 * kN * 4; }

// Instances whose two collapse chains are all resident on the current device
// at once (workgroups per CU x CUs / 2).  Each chain's 1023 steps are
// dependent, so a pass takes about the same time for any number of chains up
// to that; 0 on a HIP failure.
int64_t setup_offline_resident_instances() {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_setup_collapse, kColThreads, 0) != hipSuccess)
    return 0;
  return (int64_t)cus * per_cu / 2;
}

int launch_setup_offline(const uint16_t* db, int64_t db_rows, int64_t instances, uint64_t* bt, uint64_t* btb,
                         uint64_t* bth, uint64_t* a_hat, const uint32_t* tw, const uint32_t* fold_tab, int64_t group,
                         hipStream_t s) {
  if (db_rows <= 0 || db_rows % kN || instances <= 0 || group <= 0) return set_error(INSPIRE_ERR_INVALID, "setup shape: db_rows % 2048 == 0, instances > 0");
  const int nrlwe = (int)(db_rows / kN);
  // Public masks (host): SEED_0 query masks abar_k = -a_k (client/online.rs:164-168,
  // 226-287: ChaCha20Rng(get_seed(0)), one random_rng polynomial per query RLWE);
  // W and V packing masks (OfflinePackingKeys::init_full, packing.rs:1838-1872).
  std::vector<uint32_t> tables(2 * 4 * kN);
  build_ntt_tables(tables.data());
  std::vector<uint32_t> abar((size_t)nrlwe * 2 * kN);
  {
    uint8_t seed[32] = {0};
    std::vector<uint64_t> a((size_t)nrlwe * kN);
    chacha20_u64(seed, a.data(), a.size());
    for (int k = 0; k < nrlwe; ++k)
      for (int c = 0; c < 2; ++c) {
        uint32_t* p = abar.data() + ((size_t)k * 2 + c) * kN;
        for (int z = 0; z < kN; ++z) {
          const uint64_t v = a[(size_t)k * kN + z] % kQ;
          p[z] = (uint32_t)((v ? kQ - v : 0) % prime(c));
        }
        host_ntt_forward(p, c, tables.data());
      }
  }
  std::vector<uint32_t> masks(2 * 3 * 2 * kN);
  for (int m = 0; m < 2; ++m) {
    uint8_t seed[32];
    for (int i = 0; i < 32; ++i) seed[i] = 7;
    if (m == 1) seed[0] = seed[1] = 8;  // V_SEED (pir/scheme.rs:46-51)
    std::vector<uint64_t> r(3 * 2 * kN);
    chacha20_u64(seed, r.data(), r.size());  // PolyMatrixNTT::random_rng, poly.rs:418-442
    for (size_t i = 0; i < r.size(); ++i) masks[(size_t)m * 3 * 2 * kN + i] = (uint32_t)(r[i] % prime((int)((i / kN) & 1)));
  }
  // the collapse takes W scaled by n^-1 (the scale its chain carries)
  for (size_t i = 0; i < (size_t)3 * 2 * kN; ++i) {
    const int c = (int)((i / kN) & 1);
    masks[i] = (uint32_t)((uint64_t)masks[i] * (c ? kNInvQ1 : kNInvQ0) % prime(c));
  }
  const int64_t G = std::min(group, instances);
  uint32_t *d_abar = nullptr, *d_masks = nullptr, *d_at = nullptr, *d_R = nullptr, *d_r0 = nullptr;
  auto cleanup = [&]() {
    for (void* p : {(void*)d_abar, (void*)d_masks, (void*)d_at, (void*)d_R, (void*)d_r0})
      if (p) (void)hipFree(p);
  };
  hipError_t e = hipMalloc((void**)&d_abar, abar.size() * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&d_masks, masks.size() * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&d_at, (size_t)G * kN * 2 * kN * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&d_R, (size_t)G * 2 * kN * kN * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&d_r0, (size_t)G * 2 * kN * 4);
  if (e != hipSuccess) {
    cleanup();
    return set_error(INSPIRE_ERR_NOMEM, std::string("setup: hipMalloc failed: ") + hipGetErrorString(e));
  }
  e = hipMemcpyAsync(d_abar, abar.data(), abar.size() * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d_masks, masks.data(), masks.size() * 4, hipMemcpyHostToDevice, s);
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)k_setup_slot_ntt, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(kSlotBlock * kNPad * sizeof(uint32_t)));
    attr = true;
  }
  for (int64_t t0 = 0; t0 < instances && e == hipSuccess; t0 += G) {
    const int64_t gn = std::min(G, instances - t0);
    if (nrlwe == 1)
      hipLaunchKernelGGL(k_setup_atilde<true>, dim3((unsigned)(gn * kN)), dim3(kAtThreads), 0, s, d_at, db,
                         (int)db_rows, t0 * kN, d_abar, tw);
    else
      hipLaunchKernelGGL(k_setup_atilde<false>, dim3((unsigned)(gn * kN)), dim3(kAtThreads), 0, s, d_at, db,
                         (int)db_rows, t0 * kN, d_abar, tw);
    hipLaunchKernelGGL(k_setup_slot_ntt, dim3((unsigned)(gn * 2 * (kN / kSlotBlock))), dim3(kSlotThreads),
                       kSlotBlock * kNPad * sizeof(uint32_t), s, d_R, d_at, tw);
    CollapseArgs ca{bt + t0 * (int64_t)kNInner * kN, btb + t0 * (int64_t)kNInner * kN, bth + t0 * (int64_t)kTExp * kN,
                    d_r0, d_R, d_masks, tw, fold_tab, pow5_mod_2n(kN / 2 - 1)};
    hipLaunchKernelGGL(k_setup_collapse, dim3((unsigned)(2 * gn)), dim3(kColThreads), 0, s, ca);
    hipLaunchKernelGGL(k_setup_ahat, dim3((unsigned)gn), dim3(kColThreads), 0, s, a_hat + t0 * kN, d_r0,
                       bth + t0 * (int64_t)kTExp * kN, d_masks + 3 * 2 * kN, tw);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  cleanup();
  if (e != hipSuccess) return set_error(INSPIRE_ERR_HIP, std::string("setup: ") + hipGetErrorString(e));
  return 0;
}

}  // namespace inspire
