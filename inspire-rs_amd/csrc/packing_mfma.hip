// Packing inner product for a batch of clients on int8 MFMA (exact).
//
// For every CRT prime p and NTT slot z the packing of packing_online_kernel.cu:18-92
// is a GEMM over the key-switch terms:
//     S_b[j][p][z] = sum_{i,k,w} y_b[k][slot_w(z,i)] * T_w[j][3i+k][z]
//                  + sum_k z_body_b[k][z] * bold_t_hat[j][k][z]          (mod q_p)
// The kernels here compute the first sum; k_post_ntt adds the second.
// with T_0 = bold_t, T_1 = bold_t_bar and slot_w(z, i) the slot that the
// automorphism X -> X^(+-5^i) (packing.rs:1639-1652, rotation_kernel.cu:21-59)
// maps z to: exponent (2 rev(z) + 1) * (+-5^i) mod 2N.
//
// Slot-ordered K.  For fixed z and row k the map (i, w) -> slot_w(z, i) is
// one-to-one and reaches every slot but two (the odd exponents mod 2N are
// exactly +-5^i, i < 1024, and i stops at 1022), so the rotation terms are
// reordered by key slot: K index kappa = 2048 k + sigma.  The A operand (client
// keys) is then the same for every z: A[(b, limb)][kappa] = limb of
// y_b[k][sigma], built once per batch in fragment order.  The z-dependence moves
// into the B operand, B_z[kappa][(j, limb)] = limb of T_w[j][3i+k][z] for the
// (i, w) that reaches sigma (zero for the two slots no rotation reaches),
// stored fragment-ordered per z ("tiled").
//
// Limbs.  A client-key residue (< 2^28) splits into four 7-bit limbs a_l in
// [0, 127]; a key-switch residue into limbs of 8, 8, 8 and 4 bits, b_m.  The
// byte limbs (m < 3) are stored as signed bytes b_m - 128, so the MFMA sums
// C'_lm = sum_k a_l (b_m - 128), and C_lm = C'_lm + 128 R_l with the client's
// limb sum R_l = sum_k a_l over the 6144 K entries; the 4-bit limb is stored
// as is.  Every |term| is at most 127 * 128 and every C_lm at most
// 127 * 255 * 6144 < 2^31.  Rows of a 16x16 tile are (client, limb l) = 4
// clients x 4 limbs, columns (output j, limb m) = 4 outputs x 4 limbs; the 16
// sums of one (client, output) pair recombine as sum_{l,m} 2^(7l + 8m) C_lm.
//
// Fragment order: K step s covers kappa = 64s .. 64s+63; lane (row/col r,
// group g) holds kappa = 64s + 16g + t in byte t (v_mfma_i32_16x16x64_i8).
//
// Packed B fragments.  A lane (K group g = lane / 16, column c = lane % 16,
// output o = c / 4, limb m = c % 4) holds 16 K entries of one limb.  A byte
// limb lane keeps its 16 bytes; a 4-bit limb lane keeps 16 nibbles in 8
// bytes, entries t and t + 8 in byte t of the pair of dwords (n0, n1), so the
// bytes come back as n0 & M, n1 & M, (n0 >> 4) & M, (n1 >> 4) & M with
// M = 0x0f0f0f0f.  K group g takes 224 bytes at 224 g, output o 56 of them
// at 56 o: its three byte-limb lanes (16 bytes each), then its nibble lane
// (8 bytes).  Every lane reads a 16-byte window; a nibble lane's window starts
// 8 bytes before its data (the first half is the previous lane's and is
// ignored), so no window leaves its fragment.  A fragment is 896 bytes and a
// residue costs its 28 bits.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "modarith.hpp"

namespace inspire {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;
typedef int i32x4 __attribute__((ext_vector_type(4)));

// buffer_load_dword... lds (the LDS DMA path with a buffer resource): the
// per-lane offset and the wave-uniform soffset address the source, so a copy
// whose step offset lives in an SGPR needs no VGPR address temporaries.
__device__ void raw_buffer_load_lds(i32x4 rsrc, __attribute__((address_space(3))) void* lds, int size, int voffset,
                                    int soffset, int offset, int aux) __asm("llvm.amdgcn.raw.buffer.load.lds");

// Raw buffer resource over `bytes` bytes at `base` (gfx9 word 3: 32-bit data
// format, no swizzle).  Wave-uniform.
__device__ inline i32x4 make_raw_rsrc(const void* base, uint32_t bytes) {
  const uint64_t a = reinterpret_cast<uint64_t>(base);
  i32x4 r;
  r.x = __builtin_amdgcn_readfirstlane((int)(uint32_t)a);
  r.y = __builtin_amdgcn_readfirstlane((int)(uint32_t)(a >> 32) & 0xffff);
  r.z = __builtin_amdgcn_readfirstlane((int)bytes);
  r.w = 0x00020000;
  return r;
}

constexpr int kK = kTExp * kN;    // 6144 rotation terms in slot order
constexpr int kKSteps = kK / 64;  // 96

constexpr int kFragBytes = 896;  // packed B fragment

__host__ __device__ inline int64_t tiled_ntiles(int64_t instances) { return (instances + 3) / 4; }

// Bytes of the tiled key-switch material for `instances` outputs.
size_t tiled_bold_t_bytes(int64_t instances) {
  return (size_t)2 * kN * tiled_ntiles(instances) * kKSteps * kFragBytes;
}

__device__ u4 raw_buffer_load_b128(i32x4 rsrc, int voffset, int soffset, int aux) __asm("llvm.amdgcn.raw.buffer.load.v4i32");
constexpr int kNonTemporal = 2;  // buffer-load cache policy: nt


// Byte offset of a lane's data in a packed fragment, and of its 16-byte window.
__host__ __device__ inline int b_lane_data(int lane) {
  return 224 * (lane >> 4) + 56 * ((lane >> 2) & 3) + 16 * (lane & 3);
}
__host__ __device__ inline int b_lane_offset(int lane) { return b_lane_data(lane) - ((lane & 3) == 3 ? 8 : 0); }

// The lane's 16-byte window of the fragment at byte `frag` (wave-uniform, an
// SGPR) of the buffer `rsrc`; `voff` = b_lane_offset(lane).  Non-temporal:
// each fragment is read once.
__device__ inline u4 load_b_packed(i32x4 rsrc, int voff, int frag) {
  return raw_buffer_load_b128(rsrc, voff, frag, kNonTemporal);
}

// Per-lane constants of unpack_b: a nibble lane (limb 3) masks and shifts.
struct BLane {
  uint32_t mask;   // 0x0f0f0f0f for a nibble lane, else all ones
  uint32_t shift;  // 4 for a nibble lane, else 0
  bool nib;
  __device__ explicit BLane(int lane)
      : mask((lane & 3) == 3 ? 0x0f0f0f0fu : 0xffffffffu), shift((lane & 3) == 3 ? 4u : 0u), nib((lane & 3) == 3) {}
};

// The 16 limb bytes (as an MFMA operand) from a lane's window (a nibble
// lane's data is its z, w).
__device__ inline v4i unpack_b(const u4 r, const BLane& bl) {
  return v4i{(int)((bl.nib ? r.z : r.x) & bl.mask), (int)((bl.nib ? r.w : r.y) & bl.mask),
             (int)((r.z >> bl.shift) & bl.mask), (int)((r.w >> bl.shift) & bl.mask)};
}

// Rotation term reaching slot sigma at NTT slot z: i | (w << 15), where i =
// 1023 means no rotation reaches it.  dlog[e >> 1] holds that value for the
// odd exponent e = (-1)^w 5^i mod 2N; ez_inv = (2 rev(z) + 1)^-1 mod 2N.
__device__ inline uint32_t slot_term(uint32_t ez_inv, int sigma, const uint16_t* __restrict__ dlog) {
  const uint32_t es = 2u * (__brev((uint32_t)sigma) >> 21) + 1u;
  return dlog[((es * ez_inv) & (2 * kN - 1)) >> 1];
}

__device__ inline uint32_t odd_inverse_mod_2n(uint32_t e) {
  uint32_t x = e;  // e * e = 1 mod 8 for odd e: 3 correct bits
#pragma unroll
  for (int it = 0; it < 3; ++it) x = x * (2u - e * x);  // 6, 12, 24 bits
  return x & (2 * kN - 1);
}

// Sources of the condensed key-switch terms the tiling reads: device memory,
// or the counter-based generator (synthetic state, the values that
// launch_synth_fill would have written, for servers without a condensed copy).
struct CondensedFromMemory {
  const uint64_t* bt;
  const uint64_t* btb;
  __device__ uint64_t operator()(uint32_t w, size_t idx) const { return (w ? btb : bt)[idx]; }
};
struct CondensedFromGenerator {
  uint64_t seed;
  uint32_t stream_bt, stream_btb;
  __device__ uint64_t operator()(uint32_t w, size_t idx) const {
    const uint64_t h = synth_word(seed, w ? stream_btb : stream_bt, idx);
    return (((h & 0xffffffffull) * kQ0) >> 32) | ((((h >> 32) * kQ1) >> 32) << 32);
  }
};

// One thread per lane of a fragment: out[p][z][nt][s] packed fragments.
template <class Src>
__global__ void k_tile_bold_t(uint8_t* __restrict__ out, Src src, const uint16_t* __restrict__ dlog,
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
    const uint32_t ez_inv = odd_inverse_mod_2n(2u * (__brev((uint32_t)z) >> 21) + 1u);
    // limb m of each of the 16 residues (zero where no rotation reaches the slot)
    uint32_t limb[16];
    for (int t = 0; t < 16; ++t) {
      uint32_t res = 0;
      const int kappa = 64 * s + 16 * g + t;
      const int k = kappa >> 11, sigma = kappa & (kN - 1);
      const uint32_t term = slot_term(ez_inv, sigma, dlog);
      const int i = (int)(term & 0x3ffu);
      if (j < instances && i < kNumRot) {
        const uint64_t v = src(term >> 15, ((size_t)j * kNInner + 3 * i + k) * kN + z);
        res = (uint32_t)(p ? (v >> 32) : v);
      }
      limb[t] = (res >> (8 * m)) & (m < 3 ? 255u : 15u);
    }
    uint8_t* dst = out + (size_t)(e / 64) * kFragBytes + b_lane_data(lane);
    if (m < 3) {
      uint32_t w[4] = {0, 0, 0, 0};
      for (int t = 0; t < 16; ++t) w[t >> 2] |= (limb[t] ^ 0x80u) << (8 * (t & 3));  // b - 128 as a signed byte
      *reinterpret_cast<u4*>(dst) = u4{w[0], w[1], w[2], w[3]};
    } else {
      uint32_t n0 = 0, n1 = 0;
      for (int t = 0; t < 4; ++t) {
        n0 |= (limb[t] | limb[t + 8] << 4) << (8 * t);
        n1 |= (limb[t + 4] | limb[t + 12] << 4) << (8 * t);
      }
      reinterpret_cast<uint32_t*>(dst)[0] = n0;
      reinterpret_cast<uint32_t*>(dst)[1] = n1;
    }
  }
}

// Client keys in A-fragment order for `count` clients padded to bpad:
// afrag[p][mt][s][lane][16], lane = (row 4*bl + l, group g), byte t holds limb
// l of y_{4 mt + bl}[k][sigma] for kappa = k*N + sigma = 64s + 16g + t.
__global__ void k_batch_afrag(uint8_t* __restrict__ afrag, const uint64_t* __restrict__ y_body,
                              int64_t client_stride, int count, int bpad) {
  const int mts = bpad / 4;
  const int64_t total = 2 * (int64_t)mts * kKSteps * 64;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(e % 64);
    int64_t r = e / 64;
    const int s = (int)(r % kKSteps);
    r /= kKSteps;
    const int mt = (int)(r % mts);
    const int p = (int)(r / mts);
    const int row = lane & 15, g = lane >> 4;
    const int b = 4 * mt + (row >> 2), l = row & 3;
    uint32_t w[4] = {0, 0, 0, 0};
    if (b < count) {
      const uint64_t* yb = y_body + (int64_t)b * client_stride + 64 * s + 16 * g;
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const uint64_t v = yb[t];
        const uint32_t res = (uint32_t)(p ? (v >> 32) : v);
        w[t >> 2] |= ((res >> (7 * l)) & 127u) << (8 * (t & 3));
      }
    }
    reinterpret_cast<u4*>(afrag)[e] = u4{w[0], w[1], w[2], w[3]};
  }
}

// R[p][b][l] = sum over the 6144 K entries of limb l of client b's keys
// (zero for padding clients): the correction of the signed byte limbs.
// One block per (p, b).
__global__ __launch_bounds__(256) void k_batch_rowsum(uint32_t* __restrict__ rsum, const uint64_t* __restrict__ y_body,
                                                      int64_t client_stride, int count, int bpad) {
  const int p = blockIdx.x / bpad, b = blockIdx.x % bpad;
  uint32_t part[4] = {0, 0, 0, 0};
  if (b < count) {
    const uint64_t* yb = y_body + (int64_t)b * client_stride;
    for (int kappa = threadIdx.x; kappa < kK; kappa += 256) {
      const uint64_t v = yb[kappa];
      const uint32_t res = (uint32_t)(p ? (v >> 32) : v);
#pragma unroll
      for (int l = 0; l < 4; ++l) part[l] += (res >> (7 * l)) & 127u;
    }
  }
  __shared__ uint32_t red[4][4];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    uint32_t x = part[l];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][l] = x;
  }
  __syncthreads();
  if (threadIdx.x < 4)
    rsum[((size_t)p * bpad + b) * 4 + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

struct PackArgs {
  uint32_t* S;               // [count][2][N][instances] (z-major)
  const uint8_t* tiled;      // [2][N][ntiles][96] packed fragments of 896 B
  const uint8_t* afrag;      // [2][bpad/4][96][64][16]
  const uint32_t* rsum;      // [2][bpad][4] client limb sums
  int64_t instances;
  int64_t ntiles;
  int bpad;
  int count;
  int ngroups;               // N-tile groups per (p, z)
  int mgroups;               // M-tile groups per (p, z)
  int64_t s_client_stride;   // elements between clients in S
};

// Epilogue of one 16x16 tile.  Lane (col c = lane%16, G = lane/16) holds
// C'[4G + l][c], l = 0..3: client 4*mt + G, output 4*nt + c/4, output limb
// m = c%4.  Corrects the signed byte limbs, recombines the limbs and stores
// the rotation-term sum in the z-major layout S[b][p][z][j], so a block's
// stores fill whole lines; the z_body terms are added by the post-process
// (k_post_ntt).  X = sum_l C_lm 2^(7l) < 2^53 is reduced with 32-bit
// operations only (reduce58), then multiplied by 2^(8m) < q as a Shoup
// product; the lane-invariant constants live in TileStore.
struct TileStore {
  uint32_t q, mu, w, wp, corr;
  __device__ TileStore(int p, int lane) {
    const int m = lane & 3;
    q = prime(p);
    mu = mu58(p);
    w = 1u << (8 * m);  // 2^(8m) < 2^24 < q
    wp = (uint32_t)(((uint64_t)w << 32) / q);
    corr = m < 3 ? 128u : 0u;
  }
};

__device__ inline void store_tile(const PackArgs& a, const TileStore& ts, const v4i& acc, int p, int z, int mt,
                                  int64_t nt) {
  const int lane = threadIdx.x & 63;
  const int c = lane & 15, G = lane >> 4;
  const int m = c & 3;
  const uint32_t q = ts.q;
  const int b = 4 * mt + G;
  const u4 r = *reinterpret_cast<const u4*>(a.rsum + ((size_t)p * a.bpad + b) * 4);
  const uint32_t rl[4] = {r.x, r.y, r.z, r.w};
  // Horner over the limbs: one 64-bit shift-add per limb
  uint64_t X = (uint32_t)acc[3] + ts.corr * rl[3];
#pragma unroll
  for (int l = 2; l >= 0; --l) X = (X << 7) + ((uint32_t)acc[l] + ts.corr * rl[l]);
  const uint32_t v0 = reduce58_rt(X, q, ts.mu);
  uint32_t v = shoup_lazy(v0, ts.w, ts.wp, q);  // [0, 2q)
  v = v >= q ? v - q : v;
  uint32_t tot = v + __shfl_xor(v, 1, 64);
  tot = (tot >= q) ? tot - q : tot;
  uint32_t tot2 = tot + __shfl_xor(tot, 2, 64);
  tot2 = (tot2 >= q) ? tot2 - q : tot2;
  const int64_t j = 4 * nt + (c >> 2);
  if (m == 0 && b < a.count && j < a.instances)
    a.S[(size_t)b * a.s_client_stride + ((size_t)p * kN + z) * a.instances + j] = tot2;
}

// ---------------------------------------------------------------------------
// Small batches (up to 16 clients): block = 4 waves, MT M-tiles (4*MT clients)
// and NT N-tiles of one (p, z); each wave owns NT/4 N-tiles and loads their B
// fragments straight from HBM; the MT A fragments of a step are staged in LDS.
// ---------------------------------------------------------------------------

template <int MT, int NT>
__global__ __launch_bounds__(256) void k_packing_mfma(PackArgs a) {
  constexpr int NTW = NT / 4;
  static_assert(NT % 4 == 0 && MT * 64 <= 256, "4 waves split the N-tiles; one A row per thread");
  __shared__ u4 afrag[2][MT][64];

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

  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // This (p, z)'s fragments as a buffer resource; a fragment's offset is an SGPR.
  const i32x4 brsrc = make_raw_rsrc(a.tiled + (((size_t)p * kN + z) * a.ntiles) * (size_t)kKSteps * kFragBytes,
                                    (uint32_t)(a.ntiles * kKSteps * kFragBytes));
  const int voff = b_lane_offset(lane);
  const BLane bl(lane);
  const u4* abase = reinterpret_cast<const u4*>(a.afrag) + ((size_t)p * (a.bpad / 4) + mt0) * kKSteps * 64;
  const bool a_thread = tid < MT * 64;
  const int ax = a_thread ? tid >> 6 : 0;  // A row (M-tile) this thread stages

  auto load_b = [&](int s, u4 (&regs)[NTW]) {
#pragma unroll
    for (int y = 0; y < NTW; ++y) {
      const int64_t nt = min((int64_t)(nt0 + wave * NTW + y), a.ntiles - 1);
      regs[y] = load_b_packed(brsrc, voff, __builtin_amdgcn_readfirstlane((int)((nt * kKSteps + s) * kFragBytes)));
    }
  };
  auto load_a = [&](int s) { return abase[((size_t)ax * kKSteps + s) * 64 + lane]; };

  v4i acc[MT][NTW];
#pragma unroll
  for (int x = 0; x < MT; ++x)
#pragma unroll
    for (int y = 0; y < NTW; ++y) acc[x][y] = v4i{0, 0, 0, 0};

  // Pipeline: while step s runs on the matrix cores, the A and B fragments of
  // step s+1 are in flight; one barrier per step.
  u4 bcur[NTW], bnext[NTW];
  load_b(0, bcur);
  u4 an = load_a(0);
  if (a_thread) afrag[0][ax][lane] = an;
  __syncthreads();
  for (int s = 0; s < kKSteps; ++s) {
    const int cur = s & 1;
    const int sn = min(s + 1, kKSteps - 1);
    an = load_a(sn);
    load_b(sn, bnext);
    v4i bf[NTW];
#pragma unroll
    for (int y = 0; y < NTW; ++y) bf[y] = unpack_b(bcur[y], bl);
#pragma unroll
    for (int x = 0; x < MT; ++x) {
      const v4i af = __builtin_bit_cast(v4i, afrag[cur][x][lane]);
#pragma unroll
      for (int y = 0; y < NTW; ++y) acc[x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af, bf[y], acc[x][y], 0, 0, 0);
    }
    if (a_thread) afrag[cur ^ 1][ax][lane] = an;
    __syncthreads();
#pragma unroll
    for (int y = 0; y < NTW; ++y) bcur[y] = bnext[y];
  }

  const TileStore ts(p, lane);
#pragma unroll
  for (int x = 0; x < MT; ++x)
#pragma unroll
    for (int y = 0; y < NTW; ++y) store_tile(a, ts, acc[x][y], p, z, mt0 + x, nt0 + wave * NTW + y);
}

// ---------------------------------------------------------------------------
// Large batches: block = W waves over MT M-tiles (a client group) x
// (W / WM) * NTW N-tiles; wave w owns M-tiles [(w % WM) * MT/WM, ...) and
// N-tiles (w / WM) * NTW ..., so every B fragment leaves HBM once per group
// (the WM waves that share it meet in L2) and feeds MT MFMAs.  Packed B
// fragments go through a D-step register ring and are unpacked after the
// step's barrier; the A fragments (MT KiB per step, the same for every slot)
// are copied into an LDS ring of D+1 steps, shared by all W waves, by the LDS
// DMA path (buffer_load_dwordx4 ... lds), so they take no registers.  Both
// are buffer loads whose per-step offset is an SGPR.  A step waits with an
// explicit vector-memory count for its own A copies (issued D steps earlier)
// and a barrier for everyone else's; each A fragment is read from LDS GA
// fragments' MFMAs ahead of its use.  WPE = waves per SIMD the register
// budget is sized for.
// ---------------------------------------------------------------------------

template <int MT, int NTW, int W, int D, int WPE, int WM, int GA = 4>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_packing_wide(PackArgs a) {
  constexpr int MTW = MT / WM;                 // M-tiles per wave
  constexpr int NTG = W / WM * NTW;            // N-tiles per block
  constexpr int kAPerWave = (MT + W - 1) / W;  // A fragments each wave copies per step
  constexpr int kGroups = MTW / GA;            // A reads GA fragments ahead
  static_assert(W % WM == 0 && MTW % GA == 0 && kKSteps % D == 0, "A reads in groups of GA; D divides the steps");
  // Vector-memory operations a wave issues after its A copies of step s:
  // that step's B loads and D-1 later steps.
  constexpr int kWait = (kAPerWave + NTW) * (D - 1) + NTW;
  static_assert(kWait < 64, "vector-memory count field");
  // Slots 0..D hold the ring; slot D+1 takes the copies issued past the last step.
  __shared__ u4 aring[D + 2][MT][64];
  // Epilogue rounds: all waves park kRoundX M-tiles x NTW tiles at once.
  constexpr int kRoundX = (W * MTW * NTW <= (D + 2) * MT) ? MTW
                          : (W * (MTW / 2) * NTW <= (D + 2) * MT) ? MTW / 2
                          : (W * (MTW / 4) * NTW <= (D + 2) * MT) ? MTW / 4
                                                                   : MTW / 8;
  static_assert(kRoundX >= 1 && MTW % kRoundX == 0 && W * kRoundX * NTW <= (D + 2) * MT, "epilogue fits the A ring");

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
  const int wm = wave % WM, wn = wave / WM;
  // Tiles past the matrix repeat the last real tile's loads and are not stored.
  const int nt0 = ng * NTG;
  // This (p, z)'s B fragments as a buffer resource; the wave's tiles start at
  // the wave-uniform byte offsets btile (SGPRs).
  const i32x4 brsrc = make_raw_rsrc(a.tiled + ((size_t)p * kN + z) * a.ntiles * kKSteps * kFragBytes,
                                    (uint32_t)(a.ntiles * kKSteps * kFragBytes));
  int btile[NTW];
#pragma unroll
  for (int y = 0; y < NTW; ++y) {
    const int64_t nt = min((int64_t)(nt0 + wn * NTW + y), a.ntiles - 1);
    btile[y] = __builtin_amdgcn_readfirstlane((int)(nt * kKSteps * kFragBytes));
  }
  const int voff = b_lane_offset(lane);
  const BLane bl(lane);
  // The client group's A fragments, [MT][96 steps][64 lanes][16 B], as a
  // buffer resource: a copy reads the lane's 16 bytes (VGPR offset) of
  // fragment (x, step) (wave-uniform offset, an SGPR).
  const i32x4 arsrc = make_raw_rsrc(a.afrag + ((size_t)p * (a.bpad / 4) + mg * MT) * kKSteps * 1024,
                                    (uint32_t)(MT * kKSteps * 1024));
  const int alane = lane * 16;

  auto load_b = [&](int s, u4 (&regs)[NTW]) {
#pragma unroll
    for (int y = 0; y < NTW; ++y) regs[y] = load_b_packed(brsrc, voff, btile[y] + s * kFragBytes);
  };
  // A fragments of step s (or, past the last step, a repeat into the spare
  // slot).  When W does not divide MT the last waves copy fragment MT-1 again:
  // the same bytes to the same place.
  auto copy_a = [&](int s) {
    const int slot = s < kKSteps ? s % (D + 1) : D + 1;
    const int ss = min(s, kKSteps - 1);
#pragma unroll
    for (int f = 0; f < kAPerWave; ++f) {
      const int x = min(wave * kAPerWave + f, MT - 1);
      raw_buffer_load_lds(arsrc, (lds_void*)&aring[slot][x][0], 16, alane,
                          __builtin_amdgcn_readfirstlane((x * kKSteps + ss) * 1024), 0, 0);
    }
  };

  v4i acc[MTW][NTW];
#pragma unroll
  for (int x = 0; x < MTW; ++x)
#pragma unroll
    for (int y = 0; y < NTW; ++y) acc[x][y] = v4i{0, 0, 0, 0};

  u4 breg[D][NTW];
#pragma unroll
  for (int t = 0; t < D; ++t) {
    copy_a(t);
    load_b(t, breg[t]);
  }

  for (int s0 = 0; s0 < kKSteps; s0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int s = s0 + d;
      // Nothing moves across step boundaries: hoisting a later step's unpack
      // into this one would wait for B loads issued only a step ago.
      __builtin_amdgcn_sched_barrier(0);
      // This wave's A copies of step s are done once at most kWait newer
      // vector-memory operations are outstanding; the barrier covers the other
      // waves' copies and their reads of the slot refilled below (step s-1's).
      __builtin_amdgcn_s_waitcnt((kWait & 15) | (((kWait >> 4) & 3) << 14) | 0x3f70);
      __builtin_amdgcn_s_barrier();
      const u4* at = &aring[s % (D + 1)][wm * MTW][0];
      v4i bf[NTW];
#pragma unroll
      for (int y = 0; y < NTW; ++y) bf[y] = unpack_b(breg[d][y], bl);
      // A fragments GA tiles at a time: fragment e of the next group is read
      // as soon as this group's MFMAs on fragment e have been issued.
      v4i ga[GA];
#pragma unroll
      for (int e = 0; e < GA; ++e) ga[e] = __builtin_bit_cast(v4i, at[e * 64 + lane]);
#pragma unroll
      for (int xg = 0; xg < kGroups; ++xg) {
#pragma unroll
        for (int e = 0; e < GA; ++e) {
#pragma unroll
          for (int y = 0; y < NTW; ++y)
            acc[xg * GA + e][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ga[e], bf[y], acc[xg * GA + e][y], 0, 0, 0);
          if (xg + 1 < kGroups) ga[e] = __builtin_bit_cast(v4i, at[((xg + 1) * GA + e) * 64 + lane]);
        }
      }
      __builtin_amdgcn_sched_group_barrier(0x100, GA, 0);
#pragma unroll
      for (int xg = 0; xg < kGroups; ++xg) {
#pragma unroll
        for (int e = 0; e < GA; ++e) {
          __builtin_amdgcn_sched_group_barrier(0x008, NTW, 0);
          if (xg + 1 < kGroups) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
      }
      copy_a(s + D);
      load_b(min(s + D, kKSteps - 1), breg[d]);
    }
  }

  // Epilogue through the A ring, now idle: each wave parks kRoundX x NTW of
  // its tiles there per round (static register indices), then reduces them in
  // a rolled loop.
  constexpr int kRound = kRoundX * NTW;
  // The copies issued past the last step write the spare slot: let them land
  // before the ring is reused.
  __builtin_amdgcn_s_waitcnt(0x3f70);  // vmcnt(0)
  v4i* park = reinterpret_cast<v4i*>(&aring[0][0][0]) + wave * kRound * 64 + lane;
  const TileStore ts(p, lane);
#pragma unroll
  for (int h = 0; h < MTW / kRoundX; ++h) {
    __syncthreads();
#pragma unroll
    for (int x = 0; x < kRoundX; ++x)
#pragma unroll
      for (int y = 0; y < NTW; ++y) park[(x * NTW + y) * 64] = acc[h * kRoundX + x][y];
    for (int t = 0; t < kRound; ++t) {
      const int x = h * kRoundX + t / NTW, y = t % NTW;
      store_tile(a, ts, park[t * 64], p, z, mg * MT + wm * MTW + x, nt0 + wn * NTW + y);
    }
  }
}

void launch_tile_bold_t(uint8_t* out, const uint64_t* bt, const uint64_t* btb, const uint16_t* dlog,
                        int64_t instances, hipStream_t s) {
  hipLaunchKernelGGL(k_tile_bold_t<CondensedFromMemory>, dim3(65536), dim3(256), 0, s, out,
                     CondensedFromMemory{bt, btb}, dlog, instances);
}

void launch_tile_bold_t_synth(uint8_t* out, uint64_t seed, uint32_t stream_bt, uint32_t stream_btb,
                              const uint16_t* dlog, int64_t instances, hipStream_t s) {
  hipLaunchKernelGGL(k_tile_bold_t<CondensedFromGenerator>, dim3(65536), dim3(256), 0, s, out,
                     CondensedFromGenerator{seed, stream_bt, stream_btb}, dlog, instances);
}

// Client padding: 4 and 16 clients for k_packing_mfma, then one group of 32 or
// groups of 64 for k_packing_wide.
int batch_key_pad(int count) {
  const int unit = count <= 4 ? 4 : (count <= 16 ? 16 : (count <= 32 ? 32 : 64));
  return (count + unit - 1) / unit * unit;
}

// Client-key fragments, then the limb sums R ([2][bpad][4] u32).
static size_t afrag_frag_bytes(int bpad) { return (size_t)2 * (bpad / 4) * kKSteps * 1024; }
size_t packing_afrag_bytes(int count) {
  const int bpad = batch_key_pad(count);
  return afrag_frag_bytes(bpad) + (size_t)2 * bpad * 4 * sizeof(uint32_t);
}

template <int MT, int NT>
static void launch_small(PackArgs a, hipStream_t s) {
  a.ngroups = (int)((a.ntiles + NT - 1) / NT);
  a.mgroups = a.bpad / (4 * MT);
  const int64_t blocks = (int64_t)kN * 2 * a.ngroups * a.mgroups;
  hipLaunchKernelGGL((k_packing_mfma<MT, NT>), dim3((unsigned)blocks), dim3(256), 0, s, a);
}

template <int MT, int NTW, int W, int D, int WPE, int WM, int GA = 4>
static void launch_wide(PackArgs a, hipStream_t s) {
  constexpr int NTG = W / WM * NTW;
  a.ngroups = (int)((a.ntiles + NTG - 1) / NTG);
  a.mgroups = a.bpad / (4 * MT);
  const int64_t blocks = (int64_t)kN * 2 * a.ngroups * a.mgroups;
  hipLaunchKernelGGL((k_packing_wide<MT, NTW, W, D, WPE, WM, GA>), dim3((unsigned)blocks), dim3(64 * W), 0, s, a);
}

// Tile configurations for 64-client groups (MT, NTW, W, D, WPE, WM); the
// server's autotuner times them on the device and keeps the fastest.
const char* const kWideVariantNames[] = {"16x2 tiles, 11 waves, ring 2, 3 waves/SIMD",
                                         "16x2 tiles, 4 waves, ring 3, 2 blocks/CU",
                                         "16x1 tiles, 8 waves, ring 3, 2 blocks/CU",
                                         "16x2 tiles, 8 waves, ring 6, 1 block/CU",
                                         "16x2 tiles, 11 waves, ring 3, A read 2 ahead, 3 waves/SIMD"};
int packing_wide_variants() { return 5; }
const char* packing_wide_variant_name(int v) {
  return v >= 0 && v < packing_wide_variants() ? kWideVariantNames[v] : "";
}

int launch_packing_mfma_prep(uint8_t* afrag, const uint64_t* y_body, int64_t key_client_stride, int count,
                             hipStream_t s) {
  if (count < 1) return -1;
  const int bpad = batch_key_pad(count);
  const int64_t total = 2 * (int64_t)(bpad / 4) * kKSteps * 64;
  hipLaunchKernelGGL(k_batch_afrag, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, afrag, y_body,
                     key_client_stride, count, bpad);
  uint32_t* rsum = reinterpret_cast<uint32_t*>(afrag + afrag_frag_bytes(bpad));
  hipLaunchKernelGGL(k_batch_rowsum, dim3((unsigned)(2 * bpad)), dim3(256), 0, s, rsum, y_body, key_client_stride,
                     count, bpad);
  return 0;
}

int launch_packing_mfma(uint32_t* S, int64_t s_client_stride, const uint8_t* tiled, uint8_t* afrag,
                        const uint64_t* y_body, int64_t key_client_stride, int count, int64_t instances,
                        hipStream_t s, int variant) {
  if (count < 1 || tiled_ntiles(instances) * kKSteps * kFragBytes > INT32_MAX) return -1;
  launch_packing_mfma_prep(afrag, y_body, key_client_stride, count, s);
  return launch_packing_mfma_main(S, s_client_stride, tiled, afrag, count, instances, s, variant);
}

int launch_packing_mfma_main(uint32_t* S, int64_t s_client_stride, const uint8_t* tiled, const uint8_t* afrag,
                             int count, int64_t instances, hipStream_t s, int variant) {
  if (count < 1) return -1;
  // Fragment offsets within one (prime, slot) are 32-bit buffer offsets.
  if (tiled_ntiles(instances) * kKSteps * kFragBytes > INT32_MAX) return -1;
  const int bpad = batch_key_pad(count);
  const uint32_t* rsum = reinterpret_cast<const uint32_t*>(afrag + afrag_frag_bytes(bpad));
  PackArgs a{S, tiled, afrag, rsum, instances, tiled_ntiles(instances), bpad, count, 0, 0, s_client_stride};
  if (bpad <= 4)
    launch_small<1, 44>(a, s);    // 1 M-tile x 44 N-tiles (11 per wave)
  else if (bpad <= 16)
    launch_small<4, 44>(a, s);    // 4 M-tiles x 44 N-tiles
  else if (bpad == 32)
    launch_wide<8, 4, 4, 6, 1, 1>(a, s);     // one 32-client group: 4 waves of 8 x 4 tiles
  else {
    // 64-client groups.  Variant 0: 11 waves of 16 x 2 tiles sharing one
    // client-key ring, 2-step B ring, 3 waves per SIMD; 88 N-tiles = 4 groups.
    switch (variant) {
      case 1: launch_wide<16, 2, 4, 3, 2, 1>(a, s); break;
      case 2: launch_wide<16, 1, 8, 3, 4, 1>(a, s); break;
      case 3: launch_wide<16, 2, 8, 6, 2, 1>(a, s); break;
      case 4: launch_wide<16, 2, 11, 3, 3, 1, 2>(a, s); break;
      default: launch_wide<16, 2, 11, 2, 3, 1>(a, s); break;
    }
  }
  return 0;
}

}  // namespace inspire
