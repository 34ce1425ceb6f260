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

__global__ __launch_bounds__(kAtThreads) void k_setup_atilde(uint32_t* __restrict__ at, const uint16_t* __restrict__ db,
                                                             int db_rows, int64_t col0,
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
  const int nrlwe = db_rows / kN;
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
  // 16 consecutive slots (64 bytes) of each column j
  for (int e = tid; e < kN * kSlotBlock; e += kSlotThreads) {
    const int zz = e % kSlotBlock, j = e / kSlotBlock;
    lds[zz * kNPad + lds_index(j)] = src[(int64_t)j * 2 * kN + zz];
  }
  __syncthreads();
  uint32_t* polys[kSlotBlock];
#pragma unroll
  for (int p = 0; p < kSlotBlock; ++p) polys[p] = lds + p * kNPad;
  ntt_forward_r4_1p<kSlotThreads, kSlotBlock>(polys, c, T);
  // Step order: the collapse of chain w reads, at step i and slot k, n^-1 R_z[k]
  // with e_z = u e_k (mod 2n), u = (-1)^w 5^i; it is stored at R[g][c][w n/2 +
  // i][k].  This workgroup's 16 slots z0.. are the exponents e = a (mod 256), and
  // multiplying by u maps that coset to the 16 consecutive slots k0(u).. with
  // e_k0 = a u^-1 (mod 256), so every (w, i) row gets one contiguous 64-byte run
  // from here and the collapse reads each step's row in order.
  const int z0 = zb * kSlotBlock;
  const uint32_t a = (2u * rev11d((uint32_t)z0) + 1u) & 255u;
  const int m = tid % kSlotBlock, rb = tid / kSlotBlock;  // rows rb + 32 it of each chain
  constexpr int kRowStep = kSlotThreads / kSlotBlock;     // 32
  constexpr uint32_t kMask2n = 2u * kN - 1u;
  uint32_t p = 1, pinv = 1;
  for (int x = 0; x < rb; ++x) {
    p = (p * 5u) & kMask2n;
    pinv = (pinv * kFiveInv) & kMask2n;
  }
  uint32_t step = 1, step_inv = 1;
  for (int x = 0; x < kRowStep; ++x) {
    step = (step * 5u) & kMask2n;
    step_inv = (step_inv * kFiveInv) & kMask2n;
  }
  uint32_t* dst = R + (g * 2 + c) * (int64_t)kN * kN;
  for (int w = 0; w < 2; ++w) {
    uint32_t u5 = p, u5inv = pinv;  // 5^i and 5^-i for i = rb + 32 it
    for (int i = rb; i < kN / 2; i += kRowStep) {
      const uint32_t u = w ? (2u * kN - u5) & kMask2n : u5;
      const uint32_t uinv = w ? (2u * kN - u5inv) & kMask2n : u5inv;
      const uint32_t b = (a * uinv) & 255u;  // e_k0 mod 256 = 2 rev7(k0 / 16) + 1
      const int k = (int)(__brev((b - 1u) >> 1) >> 25) * kSlotBlock + m;
      const int zz = auto_slot(u, k) - z0;
      const uint32_t v = lds[zz * kNPad + lds_index(k)];
      dst[((int64_t)w * (kN / 2) + i) * kN + k] = c ? mulmod<1>(v, kNInvQ1) : mulmod<0>(v, kNInvQ0);
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
  const uint32_t* R;    // [G][2][N][N]
  const uint32_t* w;    // W mask [3][2][N] NTT residues
  const uint32_t* tw;
  uint32_t g_top;       // 5^(n/2 - 1) mod 2n
};

// G^-1 of the canonical NTT-domain polynomial x in dig[0], dig[1] (both
// primes): INTT, CRT compose, the first two 19-bit digits into dig[2k + c]
// and their forward NTTs.  The third digit's transform follows from x and
// those two (third_gadget_digit_ntt<0>, with the copy of x the caller keeps in
// dig[4], dig[5]), so it is never transformed.  Every thread touches only its
// own slots between the two transforms.
__device__ inline void gadget_ntt(uint32_t (&dig)[6][kNPad], const NttTables& T) {
  uint32_t* const ip[2] = {dig[0], dig[1]};
  const int ipr[2] = {0, 1};
  uint32_t* const fp[4] = {dig[0], dig[1], dig[2], dig[3]};
  const int fpr[4] = {0, 1, 0, 1};
  ntt_inverse_r4<kColThreads, 2>(ip, ipr, T);
  const uint64_t mask = (1ull << kGadgetBits) - 1;
#pragma unroll
  for (int s = 0; s < kColSlots; ++s) {
    const int lz = lds_index(threadIdx.x + s * kColThreads);
    const uint64_t raw = crt_compose(dig[0][lz], dig[1][lz]);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const uint32_t d = (uint32_t)((raw >> (kGadgetBits * k)) & mask);  // < 2^19: canonical for both primes
      dig[2 * k][lz] = d;
      dig[2 * k + 1][lz] = d;
    }
  }
  __syncthreads();
  ntt_forward_r4m<kColThreads, 4>(fp, fpr, T);
}

__global__ __launch_bounds__(kColThreads) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_setup_collapse(
    CollapseArgs a) {
  // rows 0-3: the transformed digits; rows 4, 5: r_(i+1) per prime at each
  // thread's own slots (read back by the same thread for the third digit)
  __shared__ uint32_t dig[6][kNPad];
  const NttTables T{a.tw};
  const int tid = threadIdx.x;
  const int64_t g = blockIdx.x >> 1;
  const int which = blockIdx.x & 1;  // 0: r chain (bold_t), 1: rbar chain (bold_t_bar)
  const uint32_t* Rg = a.R + g * 2 * (int64_t)kN * kN;
  uint64_t* out = (which ? a.btb : a.bt) + g * (int64_t)kNInner * kN;

  // r_i (or rbar_i) at this thread's slots, in step order in R (see
  // k_setup_slot_ntt).
  auto load_r = [&](int i, uint32_t (&rr)[2][kColSlots]) {
    const int64_t row = (int64_t)which * (kN / 2) + i;
#pragma unroll
    for (int s = 0; s < kColSlots; ++s) {
      const int z = tid + s * kColThreads;
      rr[0][s] = Rg[row * kN + z];
      rr[1][s] = Rg[((int64_t)kN + row) * kN + z];
    }
  };

  uint32_t cur[2][kColSlots], rr[2][kColSlots];
  uint32_t gi = a.g_top;  // g_(n/2 - 1)
  load_r(kN / 2 - 1, rr);
#pragma unroll
  for (int s = 0; s < kColSlots; ++s) {
    cur[0][s] = rr[0][s];
    cur[1][s] = rr[1][s];
  }
  for (int i = kN / 2 - 2; i >= 0; --i) {
    gi = (gi * kFiveInv) & (2u * kN - 1u);  // g_i = 5^i mod 2n
    const uint32_t t = which ? 2u * kN - gi : gi;
    load_r(i, rr);  // independent of the chain: in flight during the transforms
#pragma unroll
    for (int s = 0; s < kColSlots; ++s) {
      const int lz = lds_index(tid + s * kColThreads);
      dig[0][lz] = dig[4][lz] = cur[0][s];
      dig[1][lz] = dig[5][lz] = cur[1][s];
    }
    __syncthreads();
    gadget_ntt(dig, T);
    // bold_t[i][k] = NTT(G^-1(raw(r_{i+1}))); r_i += sum_k w_all[i][k] * bold_t[i][k]
#pragma unroll
    for (int s = 0; s < kColSlots; ++s) {
      const int z = tid + s * kColThreads;
      const int lz = lds_index(z);
      const int zs = auto_slot(t, z);  // the W mask rotated by t_i (cache-resident)
      uint64_t acc0 = rr[0][s], acc1 = rr[1][s];
      uint32_t dk[3][2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        dk[k][0] = dig[2 * k][lz];
        dk[k][1] = dig[2 * k + 1][lz];
      }
#pragma unroll
      for (int c = 0; c < 2; ++c) dk[2][c] = third_gadget_digit_ntt<0>(dig[4 + c][lz], dk[0][c], dk[1][c], c);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        out[((int64_t)i * kTExp + k) * kN + z] = (uint64_t)dk[k][0] | ((uint64_t)dk[k][1] << 32);
        acc0 += (uint64_t)a.w[(k * 2 + 0) * kN + zs] * dk[k][0];
        acc1 += (uint64_t)a.w[(k * 2 + 1) * kN + zs] * dk[k][1];
      }
      cur[0][s] = reduce58<0>(acc0);
      cur[1][s] = reduce58<1>(acc1);
    }
  }
  if (which == 0) {
#pragma unroll
    for (int s = 0; s < kColSlots; ++s) {
      const int z = tid + s * kColThreads;
      a.r0[(g * 2 + 0) * kN + z] = cur[0][s];
      a.r0[(g * 2 + 1) * kN + z] = cur[1][s];
    }
    return;
  }
  // bold_t_hat = NTT(G^-1(raw(rbar_0)))  (packing.rs:2733-2734)
#pragma unroll
  for (int s = 0; s < kColSlots; ++s) {
    const int lz = lds_index(tid + s * kColThreads);
    dig[0][lz] = dig[4][lz] = cur[0][s];
    dig[1][lz] = dig[5][lz] = cur[1][s];
  }
  __syncthreads();
  gadget_ntt(dig, T);
#pragma unroll
  for (int s = 0; s < kColSlots; ++s) {
    const int z = tid + s * kColThreads;
    const int lz = lds_index(z);
    uint32_t dk[3][2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      dk[k][0] = dig[2 * k][lz];
      dk[k][1] = dig[2 * k + 1][lz];
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) dk[2][c] = third_gadget_digit_ntt<0>(dig[4 + c][lz], dk[0][c], dk[1][c], c);
#pragma unroll
    for (int k = 0; k < 3; ++k)
      a.bth[(g * kTExp + k) * kN + z] = (uint64_t)dk[k][0] | ((uint64_t)dk[k][1] << 32);
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
                         uint64_t* bth, uint64_t* a_hat, const uint32_t* tw, int64_t group, hipStream_t s) {
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
    hipLaunchKernelGGL(k_setup_atilde, dim3((unsigned)(gn * kN)), dim3(kAtThreads), 0, s, d_at, db, (int)db_rows,
                       t0 * kN, d_abar, tw);
    hipLaunchKernelGGL(k_setup_slot_ntt, dim3((unsigned)(gn * 2 * (kN / kSlotBlock))), dim3(kSlotThreads),
                       kSlotBlock * kNPad * sizeof(uint32_t), s, d_R, d_at, tw);
    CollapseArgs ca{bt + t0 * (int64_t)kNInner * kN, btb + t0 * (int64_t)kNInner * kN, bth + t0 * (int64_t)kTExp * kN,
                    d_r0, d_R, d_masks, tw, pow5_mod_2n(kN / 2 - 1)};
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
