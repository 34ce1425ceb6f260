// Packing inner product for a batch of clients on int8 MFMA (exact).
//
// For every CRT prime p and NTT slot z the packing of packing_online_kernel.cu:18-92
// is a small GEMM over the 6141 key-switch terms n:
//     S_b[j][p][z] = sum_n Y_b[p][z][n] * T[p][z][j][n]   (mod q_p)
// with Y_b the client's rotated packing keys (rotation_kernel.cu:21-59) and T the
// device-resident key-switch material (bold_t, bold_t_bar, bold_t_hat).
//
// Both operands are split into four 7-bit limbs (residues < 2^28), so every limb
// product is in [0, 127^2] and an int32 accumulator over 6144 terms stays below
// 2^27.  Rows of a 16x16 tile are (client, limb l) = 4 clients x 4 limbs, columns
// (output j, limb m) = 4 outputs x 4 limbs; the 16 partial sums of one
// (client, output) pair are recombined as sum_{l,m} 2^(7(l+m)) C_lm mod q_p.
//
// Term order n (K): n = 6i + 2k + w for rotation i < 1023, key row k < 3 and
// w = 0 (y with bold_t) or 1 (y-bar with bold_t_bar); n = 6138 + k for z_body with
// bold_t_hat; n = 6141..6143 are zero padding.  K step s covers n = 64s .. 64s+63
// and lane (row/col r, group g) holds n = 64s + 16g + t in byte t.
//
// T is stored fragment-ordered ("tiled"): [p][z][ntile][step][lane][16 B], one
// 1 KiB fragment per (16-column tile, K step), so a wave fetches a fragment with
// one fully contiguous 1 KiB load.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "modarith.hpp"

namespace inspire {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

constexpr int kKPad = 6144;            // 6141 terms padded to 96 steps of 64
constexpr int kKSteps = kKPad / 64;    // 96
constexpr int kTermsRot = 6 * kNumRot; // 6138

__host__ __device__ inline int64_t tiled_ntiles(int64_t instances) { return (instances + 3) / 4; }

// Bytes of the tiled key-switch material for `instances` outputs.
size_t tiled_bold_t_bytes(int64_t instances) {
  return (size_t)2 * kN * tiled_ntiles(instances) * kKSteps * 1024;
}

// Source of term n of output j at slot z: condensed u64 from the reference layout.
__device__ inline uint64_t term_value(const uint64_t* __restrict__ bt, const uint64_t* __restrict__ btb,
                                      const uint64_t* __restrict__ bth, int64_t j, int n, int z) {
  if (n < kTermsRot) {
    const int i = n / 6, r = n % 6, k = r >> 1;
    const uint64_t* src = (r & 1) ? btb : bt;
    return src[((size_t)j * kNInner + 3 * i + k) * kN + z];
  }
  if (n < kTermsRot + kTExp) return bth[((size_t)j * kTExp + (n - kTermsRot)) * kN + z];
  return 0;
}

// One thread per 16-byte fragment row: out[p][z][nt][s][lane][16].
__global__ void k_tile_bold_t(uint8_t* __restrict__ out, const uint64_t* __restrict__ bt,
                              const uint64_t* __restrict__ btb, const uint64_t* __restrict__ bth,
                              int64_t instances) {
  const int64_t ntiles = tiled_ntiles(instances);
  const int64_t total = 2 * (int64_t)kN * ntiles * kKSteps * 64;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(e % 64);
    int64_t r = e / 64;
    const int s = (int)(r % kKSteps);
    r /= kKSteps;
    const int64_t nt = r % ntiles;
    r /= ntiles;
    const int z = (int)(r % kN);
    const int p = (int)(r / kN);
    const int col = lane & 15, g = lane >> 4;
    const int64_t j = 4 * nt + (col >> 2);
    const int m = col & 3;
    uint32_t w[4] = {0, 0, 0, 0};
    if (j < instances) {
      for (int t = 0; t < 16; ++t) {
        const uint64_t v = term_value(bt, btb, bth, j, 64 * s + 16 * g + t, z);
        const uint32_t res = (uint32_t)(p ? (v >> 32) : v);
        w[t >> 2] |= ((res >> (7 * m)) & 127u) << (8 * (t & 3));
      }
    }
    reinterpret_cast<u4*>(out)[e] = u4{w[0], w[1], w[2], w[3]};
  }
}

// Same layout filled straight from the synthetic generator (streams of
// bold_t = 2, bold_t_bar = 3, bold_t_hat = 4 in condensed kind).
__device__ inline uint64_t synth_condensed(uint64_t seed, uint32_t stream, uint64_t idx) {
  const uint64_t h = synth_word(seed, stream, idx);
  const uint64_t lo = ((h & 0xffffffffull) * kQ0) >> 32;
  const uint64_t hi = ((h >> 32) * kQ1) >> 32;
  return lo | (hi << 32);
}

__global__ void k_tile_bold_t_synth(uint8_t* __restrict__ out, uint64_t seed, int64_t instances) {
  const int64_t ntiles = tiled_ntiles(instances);
  const int64_t total = 2 * (int64_t)kN * ntiles * kKSteps * 64;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(e % 64);
    int64_t r = e / 64;
    const int s = (int)(r % kKSteps);
    r /= kKSteps;
    const int64_t nt = r % ntiles;
    r /= ntiles;
    const int z = (int)(r % kN);
    const int p = (int)(r / kN);
    const int col = lane & 15, g = lane >> 4;
    const int64_t j = 4 * nt + (col >> 2);
    const int m = col & 3;
    uint32_t w[4] = {0, 0, 0, 0};
    if (j < instances) {
      for (int t = 0; t < 16; ++t) {
        const int n = 64 * s + 16 * g + t;
        uint64_t v = 0;
        if (n < kTermsRot) {
          const int i = n / 6, rr = n % 6, k = rr >> 1;
          v = synth_condensed(seed, (rr & 1) ? 3 : 2, ((uint64_t)j * kNInner + 3 * i + k) * kN + z);
        } else if (n < kTermsRot + kTExp) {
          v = synth_condensed(seed, 4, ((uint64_t)j * kTExp + (n - kTermsRot)) * kN + z);
        }
        const uint32_t res = (uint32_t)(p ? (v >> 32) : v);
        w[t >> 2] |= ((res >> (7 * m)) & 127u) << (8 * (t & 3));
      }
    }
    reinterpret_cast<u4*>(out)[e] = u4{w[0], w[1], w[2], w[3]};
  }
}

// Client keys for a batch, per prime and client-minor so the gathers of one
// term for all clients are contiguous:  ykeys[p][k][slot][bpad], zkeys[p][k][z][bpad].
// y_body, z_body: [count][3][N] condensed (k_unpack_query output of each client).
__global__ void k_batch_keys(uint32_t* __restrict__ ykeys, uint32_t* __restrict__ zkeys,
                             const uint64_t* __restrict__ y_body, const uint64_t* __restrict__ z_body,
                             int64_t client_stride, int count, int bpad) {
  const int64_t total = (int64_t)kTExp * kN * bpad;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e % bpad);
    const int64_t ks = e / bpad;  // k*N + slot
    uint64_t y = 0, zz = 0;
    if (b < count) {
      y = y_body[(int64_t)b * client_stride + ks];
      zz = z_body[(int64_t)b * client_stride + ks];
    }
    ykeys[e] = (uint32_t)y;
    ykeys[total + e] = (uint32_t)(y >> 32);
    zkeys[e] = (uint32_t)zz;
    zkeys[total + e] = (uint32_t)(zz >> 32);
  }
}

// ---------------------------------------------------------------------------
// The batched GEMM.  Block = WM x WN waves; the block owns MT M-tiles
// (4*MT clients) and NT N-tiles (4*NT outputs) of one (p, z).  Wave (wm, wn)
// owns M-tiles [wm*MT/WM, ...) x N-tiles [wn*NT/WN, ...).
// ---------------------------------------------------------------------------

struct PackMfmaArgs {
  uint32_t* S;               // [bpad][instances][2][N] (clients b < count written)
  const uint8_t* tiled;      // [2][N][ntiles][96][64][16]
  const uint32_t* ykeys;     // [2][3][N][bpad]
  const uint32_t* zkeys;     // [2][3][N][bpad]
  int64_t instances;
  int64_t ntiles;
  int bpad;                  // clients padded to a multiple of 4*MT
  int count;                 // real clients
  int ngroups;               // N-tile groups per (p, z)
  int mgroups;               // M-tile groups per (p, z)
  int64_t s_client_stride;   // elements between clients in S
};

template <int MT, int NT, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void k_packing_mfma(PackMfmaArgs a) {
  constexpr int kThreads = 64 * WM * WN;
  constexpr int MTW = MT / WM, NTW = NT / WN;
  // With one wave row (WM == 1) every B fragment feeds exactly one wave, which
  // loads it straight from HBM; otherwise B fragments are staged in LDS once
  // and read by the WM waves that share them.
  constexpr bool kDirect = WM == 1;
  constexpr int kBStage = kDirect ? 1 : NT;
  constexpr int kBLoads = kDirect ? 0 : (NT * 64 + kThreads - 1) / kThreads;
  static_assert(MT % WM == 0 && NT % WN == 0, "tiles must split evenly over waves");
  __shared__ u4 bfrag[2][kBStage][64];
  __shared__ u4 afrag[2][MT][64];
  __shared__ int src_slot[2][64];   // per term of the step: key slot (or -1 for zero)
  __shared__ int src_row[2][64];    // key row k, +3 for z_body

  // XCD-aware block order: consecutive z stay on one XCD (block id mod 8).
  const int64_t bid = blockIdx.x;
  const int xcd = (int)(bid % 8);
  const int64_t rest = bid / 8;
  const int zl = (int)(rest % (kN / 8));
  const int64_t r2 = rest / (kN / 8);
  const int z = xcd * (kN / 8) + zl;
  const int p = (int)(r2 % 2);
  const int64_t r3 = r2 / 2;
  const int ng = (int)(r3 % a.ngroups);
  const int mg = (int)(r3 / a.ngroups);
  const int nt0 = ng * NT, mt0 = mg * MT;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const uint32_t ez = 2u * (__brev((uint32_t)z) >> 21) + 1u;
  const size_t key_plane = (size_t)kTExp * kN * a.bpad;
  const uint32_t* yk = a.ykeys + (size_t)p * key_plane;
  const uint32_t* zk = a.zkeys + (size_t)p * key_plane;
  const uint8_t* tbase = a.tiled + (((size_t)p * kN + z) * a.ntiles) * (size_t)kKSteps * 1024;

  // Term sources of step s: NTT slot of the key value (or -1) and key row.
  auto stage_terms = [&](int s, int buf) {
    if (tid < 64) {
      const int n = 64 * s + tid;
      int slot = -1, row = 0;
      if (n < kTermsRot) {
        const int i = n / 6, rr = n % 6;
        uint32_t g = 1, b5 = 5, e = (uint32_t)i;  // g = 5^i mod 4096 (packing.rs:1639-1652)
        while (e) {
          if (e & 1) g = (g * b5) & (2 * kN - 1);
          b5 = (b5 * b5) & (2 * kN - 1);
          e >>= 1;
        }
        const uint32_t e1 = (ez * g) & (2 * kN - 1);
        const int s1 = (int)(__brev(e1 >> 1) >> 21);
        slot = (rr & 1) ? (kN - 1 - s1) : s1;
        row = rr >> 1;
      } else if (n < kTermsRot + kTExp) {
        slot = z;
        row = 3 + (n - kTermsRot);
      }
      src_slot[buf][tid] = slot;
      src_row[buf][tid] = row;
    }
  };
  auto frag_ptr = [&](int nt, int s) {
    return reinterpret_cast<const u4*>(tbase + ((size_t)nt * kKSteps + s) * 1024) + lane;
  };
  auto load_b_stage = [&](int s, u4 (&regs)[kBLoads > 0 ? kBLoads : 1]) {
#pragma unroll
    for (int r = 0; r < kBLoads; ++r) {
      const int e = tid + r * kThreads;
      u4 v = {0, 0, 0, 0};
      if (e < NT * 64) {
        const int nt = nt0 + e / 64;
        if (nt < a.ntiles)
          v = __builtin_nontemporal_load(reinterpret_cast<const u4*>(tbase + ((size_t)nt * kKSteps + s) * 1024) +
                                         (e % 64));
      }
      regs[r] = v;
    }
  };
  auto store_b_stage = [&](int buf, const u4 (&regs)[kBLoads > 0 ? kBLoads : 1]) {
#pragma unroll
    for (int r = 0; r < kBLoads; ++r) {
      const int e = tid + r * kThreads;
      if (e < NT * 64) bfrag[buf][e / 64][e % 64] = regs[r];
    }
  };
  auto load_b_direct = [&](int s, u4 (&regs)[NTW]) {
#pragma unroll
    for (int y = 0; y < NTW; ++y) {
      const int nt = nt0 + wn * NTW + y;
      regs[y] = nt < a.ntiles ? __builtin_nontemporal_load(frag_ptr(nt, s)) : u4{0, 0, 0, 0};
    }
  };
  // A fragment rows: lane = 4*bl + l + 16*g, dword tq holds terms n = 16g + 4tq .. +3.
  // Item e (< 16*MT) of a step: client quad cq = e % MT (= M-tile), term quad
  // tq4 = e / MT (g = tq4/4, tq = tq4%4).  Consecutive lanes read consecutive
  // client quads of one term, i.e. contiguous 16-byte key loads.
  constexpr int kAItems = (16 * MT + kThreads - 1) / kThreads;
  auto gather_a = [&](int buf, u4 (&vals)[kAItems][4]) {
#pragma unroll
    for (int r = 0; r < kAItems; ++r) {
      const int e = tid + r * kThreads;
      const int cq = e % MT, tq4 = e / MT;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        u4 v = {0, 0, 0, 0};
        if (e < 16 * MT) {
          const int tn = 4 * tq4 + u;
          const int slot = src_slot[buf][tn], row = src_row[buf][tn];
          if (slot >= 0) {
            const uint32_t* base = row < 3 ? yk : zk;
            v = *reinterpret_cast<const u4*>(base + ((size_t)(row % 3) * kN + slot) * a.bpad + 4 * (mt0 + cq));
          }
        }
        vals[r][u] = v;
      }
    }
  };
  auto store_a = [&](int buf, const u4 (&vals)[kAItems][4]) {
#pragma unroll
    for (int r = 0; r < kAItems; ++r) {
      const int e = tid + r * kThreads;
      if (e >= 16 * MT) continue;
      const int cq = e % MT, tq4 = e / MT;
      const int g = tq4 >> 2, tq = tq4 & 3;
#pragma unroll
      for (int bl = 0; bl < 4; ++bl)
#pragma unroll
        for (int l = 0; l < 4; ++l) {
          uint32_t w = 0;
#pragma unroll
          for (int u = 0; u < 4; ++u) w |= ((vals[r][u][bl] >> (7 * l)) & 127u) << (8 * u);
          reinterpret_cast<uint32_t*>(&afrag[buf][cq][4 * bl + l + 16 * g])[tq] = w;
        }
    }
  };

  v4i acc[MTW][NTW];
#pragma unroll
  for (int x = 0; x < MTW; ++x)
#pragma unroll
    for (int y = 0; y < NTW; ++y) acc[x][y] = v4i{0, 0, 0, 0};

  // Pipeline: while step s runs on the matrix cores, the key gathers and the
  // B fragments of step s+1 are in flight and the term sources of step s+2
  // are computed; one barrier per step.
  u4 bregs[kBLoads > 0 ? kBLoads : 1];
  u4 bcur[NTW], bnext[NTW];
  u4 avals[kAItems][4];
  stage_terms(0, 0);
  stage_terms(1, 1);
  if (kDirect) {
    load_b_direct(0, bcur);
  } else {
    load_b_stage(0, bregs);
    store_b_stage(0, bregs);
  }
  __syncthreads();
  gather_a(0, avals);
  store_a(0, avals);
  __syncthreads();

  for (int s = 0; s < kKSteps; ++s) {
    const int cur = s & 1, nxt = cur ^ 1;
    const bool more = s + 1 < kKSteps;
    if (more) {
      gather_a(nxt, avals);
      if (kDirect)
        load_b_direct(s + 1, bnext);
      else
        load_b_stage(s + 1, bregs);
    }
    if (s + 2 < kKSteps) stage_terms(s + 2, cur);
    v4i af[MTW];
#pragma unroll
    for (int x = 0; x < MTW; ++x) af[x] = __builtin_bit_cast(v4i, afrag[cur][wm * MTW + x][lane]);
#pragma unroll
    for (int y = 0; y < NTW; ++y) {
      const v4i bf = kDirect ? __builtin_bit_cast(v4i, bcur[y]) : __builtin_bit_cast(v4i, bfrag[cur][wn * NTW + y][lane]);
#pragma unroll
      for (int x = 0; x < MTW; ++x) acc[x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[x], bf, acc[x][y], 0, 0, 0);
    }
    if (more) {
      store_a(nxt, avals);
      if (!kDirect) store_b_stage(nxt, bregs);
    }
    __syncthreads();
    if (kDirect) {
#pragma unroll
      for (int y = 0; y < NTW; ++y) bcur[y] = bnext[y];
    }
  }

  // Epilogue.  Lane (col c = lane%16, G = lane/16) holds C[4G + l][c], l = 0..3:
  // client bl = G, output jl = c/4, output limb m = c%4.
  const int c = lane & 15, G = lane >> 4;
  const int m = c & 3, jl = c >> 2;
  const uint32_t q = prime(p);
  const uint32_t pw = (uint32_t)(((uint64_t)1 << (7 * m)) % q);  // 2^(7m) mod q
#pragma unroll
  for (int x = 0; x < MTW; ++x) {
#pragma unroll
    for (int y = 0; y < NTW; ++y) {
      uint64_t X = 0;
#pragma unroll
      for (int l = 0; l < 4; ++l) X += (uint64_t)(uint32_t)acc[x][y][l] << (7 * l);
      uint32_t v = (uint32_t)barrett64(X, cr1(p), q);
      v = p ? mulmod<1>(v, pw) : mulmod<0>(v, pw);
      uint32_t tot = v + __shfl_xor(v, 1, 64);
      tot = (tot >= q) ? tot - q : tot;
      uint32_t tot2 = tot + __shfl_xor(tot, 2, 64);
      tot2 = (tot2 >= q) ? tot2 - q : tot2;
      const int mt = mt0 + wm * MTW + x, nt = nt0 + wn * NTW + y;
      const int b = 4 * mt + G;
      const int64_t j = 4 * (int64_t)nt + jl;
      if (m == 0 && b < a.count && j < a.instances)
        a.S[(size_t)b * a.s_client_stride + ((size_t)j * 2 + p) * kN + z] = tot2;
    }
  }
}

// ---------------------------------------------------------------------------
// Wide batched GEMM: one block holds all 4*MT clients of its group, so each
// B fragment leaves HBM once per group.  The block's waves form WM rows x 4
// columns: row wm owns M-tiles [wm*MT/WM, (wm+1)*MT/WM), column wn owns N-tiles
// [wn*NTW, (wn+1)*NTW) of the group's NT (the last column possibly fewer).
// B fragments go straight from HBM into the wave's registers; the A fragments
// (built from the clients' keys by every thread of the block) are shared
// through LDS.  Both run D steps ahead of the matrix cores through register
// rings, and every load is unconditional (clamped addresses), so the only
// waits in the loop are the ones on the data a step consumes.  With WM = 2 two
// waves share each SIMD and cover each other's LDS and barrier waits.
//
// Keys are stored as limb bytes, keys[p][row][bpad] with row = plane*N + slot
// (planes 0-2: y rows, 3-5: z_body rows) and a final zero row; byte l of a key
// is bits 7l..7l+6 of its residue.  term_src[z][n] names the row term n reads.
// ---------------------------------------------------------------------------

constexpr int64_t kKeyRows = 6 * kN + 1;

__device__ inline uint32_t limb_bytes(uint32_t r) {
  return (r & 127u) | ((r << 1) & (127u << 8)) | ((r << 2) & (127u << 16)) | ((r << 3) & (127u << 24));
}

__global__ void k_batch_keys_limbs(uint32_t* __restrict__ keys, const uint64_t* __restrict__ y_body,
                                   const uint64_t* __restrict__ z_body, int64_t client_stride, int count, int bpad) {
  const int64_t total = kKeyRows * bpad;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e % bpad);
    const int64_t row = e / bpad;
    uint64_t v = 0;
    if (b < count && row < 6 * kN) {
      const int plane = (int)(row / kN);
      const int64_t off = (int64_t)b * client_stride + (plane % 3) * kN + row % kN;
      v = plane < 3 ? y_body[off] : z_body[off];
    }
    keys[e] = limb_bytes((uint32_t)v);
    keys[total + e] = limb_bytes((uint32_t)(v >> 32));
  }
}

struct PackWideArgs {
  uint32_t* S;                // [count][instances][2][N]
  const uint8_t* tiled;       // [2][N][ntiles][96][64][16]
  const uint32_t* keys;       // [2][kKeyRows][bpad] limb bytes
  const uint16_t* term_src;   // [N][kKPad]
  int64_t instances;
  int64_t ntiles;
  int bpad;
  int count;
  int ngroups;                // N-tile groups per (p, z)
  int mgroups;                // client groups (4*MT clients each)
  int64_t s_client_stride;
};

// Key vector of CPT consecutive clients.
template <int CPT>
struct KeyVec;
template <>
struct KeyVec<4> {
  typedef unsigned int type __attribute__((ext_vector_type(4)));
  __device__ static uint32_t get(type v, int c) { return v[c]; }
};
template <>
struct KeyVec<2> {
  typedef unsigned int type __attribute__((ext_vector_type(2)));
  __device__ static uint32_t get(type v, int c) { return v[c]; }
};
template <>
struct KeyVec<1> {
  typedef unsigned int type;
  __device__ static uint32_t get(type v, int) { return v; }
};

template <int MT, int NTW, int NT, int D, int WM>
__global__ __launch_bounds__(256 * WM) void k_packing_wide(PackWideArgs a) {
  constexpr int kThreads = 256 * WM;
  constexpr int kParts = kThreads / (16 * MT);  // threads sharing one (client quad, term quad) item
  constexpr int kCpt = 4 / kParts;              // clients per thread
  constexpr int MTW = MT / WM;
  static_assert(kParts * 16 * MT == kThreads && (kParts == 1 || kParts == 2 || kParts == 4),
                "the block's threads split the A items evenly");
  static_assert(MT % WM == 0, "M-tiles split evenly over wave rows");
  static_assert(NTW * 4 >= NT && NTW * 3 < NT, "four wave columns cover the N-tiles");
  static_assert(kKSteps % D == 0 && D % 2 == 0, "ring depth must divide the steps and be even");
  typedef typename KeyVec<kCpt>::type KV;
  constexpr int kTileStride = 65;  // u4 per A tile in LDS: one pad row spreads the limb writes over the banks
  __shared__ uint16_t tsrc[kKPad];
  __shared__ u4 afrag[2][MT * kTileStride];

  // Blocks of one (p, z) are consecutive on one XCD (block id mod 8), client
  // groups innermost so the groups' reads of a B fragment meet in L2.
  const int64_t bid = blockIdx.x;
  const int xcd = (int)(bid % 8);
  int64_t rest = bid / 8;
  const int mg = (int)(rest % a.mgroups);
  rest /= a.mgroups;
  const int ng = (int)(rest % a.ngroups);
  rest /= a.ngroups;
  const int z = xcd * (kN / 8) + (int)(rest % (kN / 8));
  const int p = (int)(rest / (kN / 8));

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave / 4, wn = wave % 4;

  {
    const u4* src = reinterpret_cast<const u4*>(a.term_src + (size_t)z * kKPad);
    for (int e = tid; e < kKPad / 8; e += kThreads) reinterpret_cast<u4*>(tsrc)[e] = src[e];
  }

  // A item part of this thread: clients part*kCpt.. of client quad cq, terms
  // 4*tq4 .. 4*tq4+3 of a step.  Consecutive threads read consecutive clients.
  const int part = tid % kParts, cq = (tid / kParts) % MT, tq4 = tid / (kParts * MT);
  const int g = tq4 >> 2, tq = tq4 & 3;
  const uint8_t* keys_p = reinterpret_cast<const uint8_t*>(a.keys + (size_t)p * kKeyRows * a.bpad +
                                                           4 * (mg * MT + cq) + part * kCpt);
  const uint32_t row_bytes = (uint32_t)a.bpad * 4u;

  // B tiles of this wave (clamped: tiles past the group or the matrix repeat
  // the last real one and their sums are dropped).
  const int nt0 = ng * NT;
  const uint8_t* tb = a.tiled + ((size_t)p * kN + z) * a.ntiles * kKSteps * 1024;
  const uint8_t* bbase[NTW];
#pragma unroll
  for (int y = 0; y < NTW; ++y) {
    const int ntl = min(wn * NTW + y, NT - 1);
    const int64_t nt = min((int64_t)(nt0 + ntl), a.ntiles - 1);
    bbase[y] = tb + (size_t)nt * kKSteps * 1024 + lane * 16;
  }

  auto gather = [&](int s, KV (&v)[4]) {
    const uint2 ts = *reinterpret_cast<const uint2*>(&tsrc[64 * s + 4 * tq4]);
    const uint32_t r[4] = {ts.x & 0xffffu, ts.x >> 16, ts.y & 0xffffu, ts.y >> 16};
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = *reinterpret_cast<const KV*>(keys_p + r[u] * row_bytes);
  };
  // 4x4 byte transposes: dword (client bl, limb l) gets limb l of the four terms.
  auto transform = [&](int buf, const KV (&v)[4]) {
    uint32_t* base = reinterpret_cast<uint32_t*>(&afrag[buf][cq * kTileStride + 16 * g]) + tq;
#pragma unroll
    for (int c = 0; c < kCpt; ++c) {
      const int bl = part * kCpt + c;
      const uint32_t x0 = KeyVec<kCpt>::get(v[0], c), x1 = KeyVec<kCpt>::get(v[1], c);
      const uint32_t x2 = KeyVec<kCpt>::get(v[2], c), x3 = KeyVec<kCpt>::get(v[3], c);
      const uint32_t t0 = __builtin_amdgcn_perm(x1, x0, 0x05010400u);
      const uint32_t t1 = __builtin_amdgcn_perm(x1, x0, 0x07030602u);
      const uint32_t t2 = __builtin_amdgcn_perm(x3, x2, 0x05010400u);
      const uint32_t t3 = __builtin_amdgcn_perm(x3, x2, 0x07030602u);
      base[(4 * bl + 0) * 4] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
      base[(4 * bl + 1) * 4] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
      base[(4 * bl + 2) * 4] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
      base[(4 * bl + 3) * 4] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
    }
  };
  auto load_b = [&](int s, u4 (&regs)[NTW]) {
#pragma unroll
    for (int y = 0; y < NTW; ++y)
      regs[y] = __builtin_nontemporal_load(reinterpret_cast<const u4*>(bbase[y] + (size_t)s * 1024));
  };

  v4i acc[MTW][NTW];
#pragma unroll
  for (int x = 0; x < MTW; ++x)
#pragma unroll
    for (int y = 0; y < NTW; ++y) acc[x][y] = v4i{0, 0, 0, 0};

  // Rings: A keys of step t in araw[t % D], B fragments of step t in breg[t % D].
  KV araw[D][4];
  u4 breg[D][NTW];
  __syncthreads();
  gather(0, araw[0]);
  transform(0, araw[0]);
#pragma unroll
  for (int t = 0; t < D; ++t) {
    gather(t + 1, araw[(t + 1) % D]);
    load_b(t, breg[t]);
  }
  __syncthreads();

  for (int s0 = 0; s0 < kKSteps; s0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int s = s0 + d;
      transform((d + 1) & 1, araw[(d + 1) % D]);  // A of step s+1 (unused past the end)
      gather(min(s + 1 + D, kKSteps - 1), araw[(d + 1) % D]);
      const u4* at = afrag[d & 1] + wm * MTW * kTileStride;
#pragma unroll
      for (int x = 0; x < MTW; ++x) {
        const v4i af = __builtin_bit_cast(v4i, at[x * kTileStride + lane]);
#pragma unroll
        for (int y = 0; y < NTW; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af, __builtin_bit_cast(v4i, breg[d][y]), acc[x][y], 0, 0, 0);
      }
      load_b(min(s + D, kKSteps - 1), breg[d]);
      __syncthreads();
      __builtin_amdgcn_sched_barrier(0);  // keep each step's loads in its own step
    }
  }

  // Epilogue as in k_packing_mfma: lane (col c, G) holds C[4G + l][c].
  const int c = lane & 15, G = lane >> 4;
  const int m = c & 3, jl = c >> 2;
  const uint32_t q = prime(p);
  const uint32_t pw = (uint32_t)(((uint64_t)1 << (7 * m)) % q);
#pragma unroll
  for (int x = 0; x < MTW; ++x) {
#pragma unroll
    for (int y = 0; y < NTW; ++y) {
      uint64_t X = 0;
#pragma unroll
      for (int l = 0; l < 4; ++l) X += (uint64_t)(uint32_t)acc[x][y][l] << (7 * l);
      uint32_t v = (uint32_t)barrett64(X, cr1(p), q);
      v = p ? mulmod<1>(v, pw) : mulmod<0>(v, pw);
      uint32_t tot = v + __shfl_xor(v, 1, 64);
      tot = (tot >= q) ? tot - q : tot;
      uint32_t tot2 = tot + __shfl_xor(tot, 2, 64);
      tot2 = (tot2 >= q) ? tot2 - q : tot2;
      const int ntl = wn * NTW + y;
      const int64_t j = 4 * (int64_t)(nt0 + ntl) + jl;
      const int b = 4 * (mg * MT + wm * MTW + x) + G;
      if (m == 0 && ntl < NT && b < a.count && j < a.instances)
        a.S[(size_t)b * a.s_client_stride + ((size_t)j * 2 + p) * kN + z] = tot2;
    }
  }
}

void launch_tile_bold_t(uint8_t* out, const uint64_t* bt, const uint64_t* btb, const uint64_t* bth, int64_t instances,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_tile_bold_t, dim3(65536), dim3(256), 0, s, out, bt, btb, bth, instances);
}

void launch_tile_bold_t_synth(uint8_t* out, uint64_t seed, int64_t instances, hipStream_t s) {
  hipLaunchKernelGGL(k_tile_bold_t_synth, dim3(65536), dim3(256), 0, s, out, seed, instances);
}

// Client padding: 4 and 16 clients for the k_packing_mfma configurations,
// then groups of 32 and 64 for k_packing_wide.
int batch_key_pad(int count) {
  const int unit = count <= 4 ? 4 : (count <= 16 ? 16 : (count <= 32 ? 32 : 64));
  return (count + unit - 1) / unit * unit;
}

size_t batch_keys_bytes(int bpad) { return (size_t)2 * kTExp * kN * bpad * sizeof(uint32_t); }

// Key buffer for up to `count` clients, large enough for either key layout.
size_t packing_keys_bytes(int count) { return (size_t)2 * kKeyRows * batch_key_pad(count) * sizeof(uint32_t); }

template <int MT, int NTW, int NT, int D, int WM>
static void launch_wide(PackWideArgs a, hipStream_t s) {
  a.ngroups = (int)((a.ntiles + NT - 1) / NT);
  a.mgroups = a.bpad / (4 * MT);
  const int64_t blocks = (int64_t)kN * 2 * a.ngroups * a.mgroups;
  hipLaunchKernelGGL((k_packing_wide<MT, NTW, NT, D, WM>), dim3((unsigned)blocks), dim3(256 * WM), 0, s, a);
}

template <int MT, int NT, int WM, int WN>
static void launch_cfg(PackMfmaArgs a, hipStream_t s) {
  a.ngroups = (int)((a.ntiles + NT - 1) / NT);
  a.mgroups = a.bpad / (4 * MT);
  const int64_t blocks = (int64_t)kN * 2 * a.ngroups * a.mgroups;
  hipLaunchKernelGGL((k_packing_mfma<MT, NT, WM, WN>), dim3((unsigned)blocks), dim3(64 * WM * WN), 0, s, a);
}

int launch_packing_mfma(uint32_t* S, int64_t s_client_stride, const uint8_t* tiled, const uint16_t* term_src,
                        uint32_t* keys, const uint64_t* y_body, const uint64_t* z_body, int64_t key_client_stride,
                        int count, int64_t instances, hipStream_t s) {
  if (count < 1) return -1;
  const int bpad = batch_key_pad(count);
  if (bpad > 16) {
    const int64_t total = kKeyRows * bpad;
    hipLaunchKernelGGL(k_batch_keys_limbs, dim3((unsigned)((total + 255) / 256 > 4096 ? 4096 : (total + 255) / 256)),
                       dim3(256), 0, s, keys, y_body, z_body, key_client_stride, count, bpad);
    PackWideArgs a{S, tiled, keys, term_src, instances, tiled_ntiles(instances), bpad, count, 0, 0, s_client_stride};
    if (bpad == 32)
      launch_wide<8, 6, 22, 4, 1>(a, s);   // 32 clients: 4 waves of 8 x 6 tiles over 22 N-tiles
    else
      launch_wide<16, 4, 16, 4, 2>(a, s);  // 64-client groups: 8 waves of 8 x 4 tiles over 16 N-tiles
    return 0;
  }
  uint32_t* ykeys = keys;
  uint32_t* zkeys = keys + batch_keys_bytes(bpad) / sizeof(uint32_t);
  const int64_t total = (int64_t)kTExp * kN * bpad;
  hipLaunchKernelGGL(k_batch_keys, dim3((unsigned)((total + 255) / 256 > 4096 ? 4096 : (total + 255) / 256)),
                     dim3(256), 0, s, ykeys, zkeys, y_body, z_body, key_client_stride, count, bpad);
  PackMfmaArgs a{S, tiled, ykeys, zkeys, instances, tiled_ntiles(instances), bpad, count, 0, 0, s_client_stride};
  if (bpad <= 4)
    launch_cfg<1, 44, 1, 4>(a, s);      // 1 M-tile: 4 waves x 11 N-tiles, B direct from HBM
  else
    launch_cfg<4, 44, 1, 4>(a, s);      // 4 M-tiles: 4 waves x (4 x 11) tiles, B direct
  return 0;
}

}  // namespace inspire
