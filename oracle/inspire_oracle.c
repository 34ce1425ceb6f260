/*
 * inspire_oracle.c -- CPU restatement of the inspire-rs keyword-PIR answer path
 * (KeywordServer::respond, src/pir/server/keyword.rs:534-628).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker and the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.  The product library (inspire-rs_amd/) never links or calls it.
 *
 * Every function cites the reference lines it restates.  The arithmetic is
 * written for clarity first; the loops that the reference parallelises with
 * rayon (packing outputs, fold chains) use OpenMP here.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   - Barrett constants: KAT arith.rs:484-497.
 *   - NTT tables: KAT ntt.rs:797-815 (entries and XOR of all tables).
 *   - NTT forward/inverse: invariants ntt.rs:818-840 and a naive O(n^2)
 *     evaluation at odd powers of the root (python big ints).
 *   - Automorphism tables: brute-force matching as in packing.rs:1423-1468.
 *   - GEMV: closed form sum a*b mod q (kernel.rs:239-262), python big ints.
 *   - Packing: closed-form per-slot sum, python big ints.
 *   - Gadget: KAT gadget.rs:79-96.
 *   - Rescale / bit I/O: exact rational definition, python.
 *   - Fold: one external product against a schoolbook negacyclic product.
 *   - DB encode: evaluating the encoded polynomials at Y = X^(2g*j/d)
 *     recovers the raw polynomials (python, direct O(d^2) evaluation); d = 1
 *     and d = 2 closed forms.
 *   - ChaCha20 block function: RFC 8439 section 2.3.2 / A.1 vectors.
 *     The rand_chacha word order (64-bit counter in words 12-13, u64 = lo|hi<<32)
 *     is parity unpinned: nothing in the reference's tests fixes it.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;
typedef __int128 i128;

#define N 2048
#define LOGN 11
#define TEXP 3                 /* t_exp_left = t_gsw = 3 (pir/params.rs:69-96) */
#define NROT (N / 2 - 1)       /* 1023 rotations (packing_online.rs:113) */
#define NINNER (NROT * TEXP)   /* 3069 */

static const uint64_t MODULI[2] = {268369921ULL, 249561089ULL}; /* pir/params.rs:9 */
#define MODQ (268369921ULL * 249561089ULL)

/* ------------------------------------------------------------------ */
/* arith.rs                                                            */
/* ------------------------------------------------------------------ */

/* arith.rs:5-7 */
static uint64_t mul_mod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)(((u128)a * b) % m); }

/* arith.rs:52-78 */
static uint64_t pow_mod(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    b %= m;
    while (e) {
        if (e & 1) r = mul_mod(r, b, m);
        b = mul_mod(b, b, m);
        e >>= 1;
    }
    return r;
}

/* arith.rs:80-87 */
static uint32_t rev_bits(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

/* arith.rs:89-100 */
static uint64_t div2_mod(uint64_t a, uint64_t m) {
    if (a & 1) {
        u128 s = (u128)a + m;
        return (uint64_t)(s >> 1);
    }
    return a >> 1;
}

/* arith.rs:117-122: floor(2^128 / m) split into (lo, hi). */
static void barrett_crs(uint64_t m, uint64_t* lo, uint64_t* hi) {
    u128 all = ~(u128)0;            /* 2^128 - 1; m odd so floor is unchanged */
    u128 qt = all / m;
    *lo = (uint64_t)qt;
    *hi = (uint64_t)(qt >> 64);
}

/* arith.rs:133-141 */
static uint64_t barrett_raw_u64(uint64_t x, uint64_t cr1, uint64_t m) {
    uint64_t t = (uint64_t)(((u128)x * cr1) >> 64);
    uint64_t r = x - t * m;
    return r >= m ? r - m : r;
}

/* arith.rs:172-209 (SEAL-style 128-bit Barrett, then one subtraction) */
static uint64_t barrett_u128(u128 v, uint64_t cr0, uint64_t cr1, uint64_t m) {
    uint64_t zx = (uint64_t)v, zy = (uint64_t)(v >> 64);
    uint64_t carry = (uint64_t)(((u128)zx * cr0) >> 64);
    u128 t2 = (u128)zx * cr1;
    uint64_t t2x = (uint64_t)t2, t2y = (uint64_t)(t2 >> 64);
    uint64_t tmp1 = t2x + carry;
    uint64_t tmp3 = t2y + (tmp1 < t2x);
    t2 = (u128)zy * cr0;
    t2x = (uint64_t)t2; t2y = (uint64_t)(t2 >> 64);
    uint64_t old = tmp1;
    tmp1 = old + t2x;
    carry = t2y + (tmp1 < old);
    tmp1 = zy * cr1 + tmp3 + carry;
    uint64_t r = zx - tmp1 * m;
    return r >= m ? r - m : r;
}

/* number_theory.rs:57-96 */
static uint64_t inv_mod(uint64_t a, uint64_t m) {
    int64_t t = 0, nt = 1;
    int64_t r = (int64_t)m, nr = (int64_t)(a % m);
    while (nr) {
        int64_t qq = r / nr, tmp;
        tmp = t - qq * nt; t = nt; nt = tmp;
        tmp = r - qq * nr; r = nr; nr = tmp;
    }
    if (t < 0) t += (int64_t)m;
    return (uint64_t)t;
}

/* Parameter constants, core/params.rs:258-270. */
typedef struct {
    uint64_t cr0[2], cr1[2];
    uint64_t cr0_mod, cr1_mod;
    uint64_t mod0_inv_mod1, mod1_inv_mod0;
} consts_t;

static consts_t C;
static uint64_t TABLES[2][4][N];  /* ntt.rs:74-108 layout: [crt][fwd, fwd', inv, inv'][i] */
static uint64_t ROOT[2];
static int g_init = 0;

/* number_theory.rs:6-55: the minimal primitive 2n-th root.  The reference
 * draws a random primitive root and then takes the minimum over all odd
 * powers, so the value does not depend on the random draw. */
static uint64_t minimal_root(uint64_t degree, uint64_t m) {
    uint64_t root = 0;
    for (uint64_t g = 2; g < m; g++) {
        uint64_t r = pow_mod(g, (m - 1) / degree, m);
        if (r && pow_mod(r, degree >> 1, m) == m - 1) { root = r; break; }
    }
    uint64_t gsq = mul_mod(root, root, m), cur = root, best = root;
    for (uint64_t i = 0; i < degree; i++) {
        if (cur < best) best = cur;
        cur = mul_mod(cur, gsq, m);
    }
    return best;
}

/* ntt.rs:6-17 */
static void powers_of_root(uint64_t root, uint64_t m, uint64_t* out) {
    uint64_t p = root;
    for (uint32_t i = 1; i < N; i++) {
        out[rev_bits(i, LOGN)] = p;
        p = mul_mod(p, root, m);
    }
    out[0] = 1;
}

static void init_all(void) {
    if (g_init) return;
    for (int i = 0; i < 2; i++) barrett_crs(MODULI[i], &C.cr0[i], &C.cr1[i]);
    barrett_crs(MODQ, &C.cr0_mod, &C.cr1_mod);
    C.mod0_inv_mod1 = MODULI[0] * inv_mod(MODULI[0], MODULI[1]);
    C.mod1_inv_mod0 = MODULI[1] * inv_mod(MODULI[1], MODULI[0]);
    for (int c = 0; c < 2; c++) {
        uint64_t m = MODULI[c];
        uint64_t root = minimal_root(2 * N, m);
        ROOT[c] = root;
        uint64_t iroot = inv_mod(root, m);
        powers_of_root(root, m, TABLES[c][0]);
        powers_of_root(iroot, m, TABLES[c][2]);
        for (int i = 0; i < N; i++) TABLES[c][2][i] = div2_mod(TABLES[c][2][i], m);
        /* ntt.rs:29-37 scale_powers_u32 */
        for (int i = 0; i < N; i++) {
            TABLES[c][1][i] = (uint64_t)(uint32_t)((TABLES[c][0][i] << 32) / m);
            TABLES[c][3][i] = (uint64_t)(uint32_t)((TABLES[c][2][i] << 32) / m);
        }
    }
    g_init = 1;
}

/* params.rs:203-215 */
static uint64_t crt_compose_2(uint64_t x, uint64_t y) {
    u128 v = (u128)x * C.mod1_inv_mod0 + (u128)y * C.mod0_inv_mod1;
    return barrett_u128(v, C.cr0_mod, C.cr1_mod, MODQ);
}

/* ------------------------------------------------------------------ */
/* NTT (ntt.rs:111-160 forward, ntt.rs:736-779 inverse, scalar forms)  */
/* ------------------------------------------------------------------ */

static void ntt_forward_one(uint64_t* op, int c) {
    const uint64_t* w_t = TABLES[c][0];
    const uint64_t* wp_t = TABLES[c][1];
    uint32_t q = (uint32_t)MODULI[c], q2 = 2 * q;
    for (int mm = 0; mm < LOGN; mm++) {
        int m = 1 << mm, t = N >> (mm + 1);
        for (int i = 0; i < m; i++) {
            uint64_t w = w_t[m + i], wp = wp_t[m + i];
            uint64_t* o = op + 2 * t * i;
            for (int j = 0; j < t; j++) {
                uint32_t x = (uint32_t)o[j], y = (uint32_t)o[t + j];
                uint32_t cx = x - (x >= q2 ? q2 : 0);
                uint64_t qt = ((uint64_t)y * wp) >> 32;
                uint64_t qn = w * (uint64_t)y - qt * q;
                o[j] = (uint64_t)cx + qn;
                o[t + j] = (uint64_t)cx + ((uint64_t)q2 - qn);
            }
        }
    }
    for (int i = 0; i < N; i++) {
        if (op[i] >= q2) op[i] -= q2;
        if (op[i] >= q) op[i] -= q;
    }
}

static void ntt_inverse_one(uint64_t* op, int c) {
    const uint64_t* w_t = TABLES[c][2];
    const uint64_t* wp_t = TABLES[c][3];
    uint64_t q = MODULI[c], q2 = 2 * q;
    for (int mm = LOGN - 1; mm >= 0; mm--) {
        int h = 1 << mm, t = N >> (mm + 1);
        for (int i = 0; i < h; i++) {
            uint64_t w = w_t[h + i], wp = wp_t[h + i];
            uint64_t* o = op + 2 * t * i;
            for (int j = 0; j < t; j++) {
                uint64_t x = o[j], y = o[t + j];
                uint64_t tt = q2 - y + x;
                uint64_t cx = x + y - (((x << 1) >= tt) ? q2 : 0);
                uint64_t ht = (tt * wp) >> 32;
                o[j] = (cx + ((tt & 1) ? q : 0)) >> 1;
                o[t + j] = w * tt - ht * q;
            }
        }
    }
    for (int i = 0; i < N; i++) {
        if (op[i] >= q2) op[i] -= q2;
        if (op[i] >= q) op[i] -= q;
    }
}

/* ------------------------------------------------------------------ */
/* Automorphism tables (packing.rs:1423-1468, closed form)             */
/* ------------------------------------------------------------------ */

/* NTT slot k holds the evaluation at psi^(2*rev(k)+1).  X -> X^t moves the
 * evaluation at psi^e to psi^(e*t), so table_t[j] = slot of exponent
 * (2*rev(j)+1)*t mod 2n.  The brute-force search of the reference finds
 * exactly this permutation (tests check it against a direct NTT match). */
static uint32_t auto_index(uint32_t t, uint32_t j) {
    uint32_t e = ((2u * rev_bits(j, LOGN) + 1u) * t) & (2u * N - 1u);
    return rev_bits((e - 1u) >> 1, LOGN);
}

/* ------------------------------------------------------------------ */
/* Exported primitives                                                 */
/* ------------------------------------------------------------------ */

void orc_constants(uint64_t* out) {
    init_all();
    out[0] = C.cr0[0]; out[1] = C.cr1[0]; out[2] = C.cr0[1]; out[3] = C.cr1[1];
    out[4] = C.cr0_mod; out[5] = C.cr1_mod;
    out[6] = C.mod0_inv_mod1; out[7] = C.mod1_inv_mod0;
    out[8] = MODQ; out[9] = ROOT[0]; out[10] = ROOT[1];
}

void orc_ntt_tables(uint64_t* out) {
    init_all();
    memcpy(out, TABLES, sizeof(TABLES));
}

void orc_barrett_raw_u64(const uint64_t* x, uint64_t* out, size_t n, uint64_t cr1, uint64_t m) {
    for (size_t i = 0; i < n; i++) out[i] = barrett_raw_u64(x[i], cr1, m);
}

void orc_barrett_u128(const uint64_t* lo, const uint64_t* hi, uint64_t* out, size_t n,
                      uint64_t cr0, uint64_t cr1, uint64_t m) {
    for (size_t i = 0; i < n; i++) out[i] = barrett_u128(((u128)hi[i] << 64) | lo[i], cr0, cr1, m);
}

void orc_crt_compose(const uint64_t* x, const uint64_t* y, uint64_t* out, size_t n) {
    init_all();
    for (size_t i = 0; i < n; i++) out[i] = crt_compose_2(x[i], y[i]);
}

/* poly: [2][N] (crt0 then crt1), in place. */
void orc_ntt_forward(uint64_t* poly, size_t count) {
    init_all();
    for (size_t p = 0; p < count; p++)
        for (int c = 0; c < 2; c++) ntt_forward_one(poly + p * 2 * N + c * N, c);
}

void orc_ntt_inverse(uint64_t* poly, size_t count) {
    init_all();
    for (size_t p = 0; p < count; p++)
        for (int c = 0; c < 2; c++) ntt_inverse_one(poly + p * 2 * N + c * N, c);
}

void orc_automorph_table(uint32_t t, uint32_t* out) {
    for (uint32_t j = 0; j < N; j++) out[j] = auto_index(t, j);
}

/* poly.rs:543-556 automorph_poly (coefficient domain, mod q) */
void orc_automorph_raw(const uint64_t* a, uint64_t* res, uint32_t t, uint64_t modulus) {
    for (uint64_t i = 0; i < N; i++) {
        uint64_t num = (i * t) / N, rem = (i * t) % N;
        res[rem] = (num % 2 == 0) ? a[i] : modulus - a[i];
    }
}

/* gadget.rs:3-9 */
static int bits_per(uint64_t modulus_log2, int dim) {
    if ((uint64_t)dim == modulus_log2) return 1;
    return (int)((double)modulus_log2 / (double)dim) + 1;
}

/* gadget.rs:35-61: inp [rdim][N] -> out [rdim*num_elems][N], row j + k*rdim. */
void orc_gadget_invert(const uint64_t* inp, uint64_t* out, int rdim, int mx, uint64_t modulus_log2) {
    int num_elems = mx / rdim;
    int bp = bits_per(modulus_log2, num_elems);
    uint64_t mask = (1ULL << bp) - 1;
    for (int j = 0; j < rdim; j++)
        for (int z = 0; z < N; z++) {
            uint64_t v = inp[(size_t)j * N + z];
            for (int k = 0; k < num_elems; k++) {
                int off = k * bp;
                out[(size_t)(j + k * rdim) * N + z] = off >= 64 ? 0 : (v >> off) & mask;
            }
        }
}

/* arith.rs:436-451 */
static uint64_t rescale(uint64_t a, uint64_t inp_mod, uint64_t out_mod) {
    int64_t im = (int64_t)inp_mod;
    i128 om = (i128)out_mod;
    int64_t v = (int64_t)(a % inp_mod);
    if (v >= im / 2) v -= im;
    int64_t sign = v >= 0 ? 1 : -1;
    i128 val = (i128)v * (i128)out_mod;
    i128 r = (val + (i128)(sign * (im / 2))) / (i128)inp_mod;
    r = (r + (i128)((inp_mod / out_mod) * out_mod) + 2 * om) % om;
    return (uint64_t)((r + om) % om);
}

void orc_rescale(const uint64_t* a, uint64_t* out, size_t n, uint64_t inp_mod, uint64_t out_mod) {
    for (size_t i = 0; i < n; i++) out[i] = rescale(a[i], inp_mod, out_mod);
}

/* util.rs:289-301 */
static uint64_t read_bits(const uint8_t* d, size_t bit_offs, int nbits) {
    size_t w = bit_offs / 64;
    int b = (int)(bit_offs % 64);
    if (b + nbits <= 64) {
        uint64_t v;
        memcpy(&v, d + w * 8, 8);
        return (v >> b) & ((nbits == 64) ? ~0ULL : ((1ULL << nbits) - 1));
    }
    u128 v;
    memcpy(&v, d + w * 8, 16);
    return (uint64_t)((v >> b) & (((u128)1 << nbits) - 1));
}

/* util.rs:303-321 */
static void write_bits(uint8_t* d, uint64_t val, size_t bit_offs, int nbits) {
    size_t w = bit_offs / 64;
    int b = (int)(bit_offs % 64);
    val &= (1ULL << nbits) - 1;
    if (b + nbits <= 64) {
        uint64_t cur;
        memcpy(&cur, d + w * 8, 8);
        cur &= ~(((1ULL << nbits) - 1) << b);
        cur |= val << b;
        memcpy(d + w * 8, &cur, 8);
    } else {
        u128 cur;
        memcpy(&cur, d + w * 8, 16);
        cur &= ~((((u128)1 << nbits) - 1) << b);
        cur |= (u128)val << b;
        memcpy(d + w * 8, &cur, 16);
    }
}

void orc_write_bits(uint8_t* d, uint64_t val, size_t bit_offs, int nbits) { write_bits(d, val, bit_offs, nbits); }
uint64_t orc_read_bits(const uint8_t* d, size_t bit_offs, int nbits) { return read_bits(d, bit_offs, nbits); }

/* ------------------------------------------------------------------ */
/* ChaCha20 (RFC 8439 block function; rand_chacha::ChaCha20Rng layout) */
/* ------------------------------------------------------------------ */

#define ROTL(a, b) (((a) << (b)) | ((a) >> (32 - (b))))
#define QR(a, b, c, d) (a += b, d ^= a, d = ROTL(d, 16), c += d, b ^= c, b = ROTL(b, 12), \
                        a += b, d ^= a, d = ROTL(d, 8), c += d, b ^= c, b = ROTL(b, 7))

/* One block: words 12..15 are the caller's (rand_chacha: 64-bit block counter
 * in words 12-13, 64-bit stream id in 14-15; RFC 8439: 32-bit counter in 12,
 * nonce in 13-15). */
void orc_chacha20_block(const uint8_t key[32], const uint32_t w12_15[4], uint32_t out[16]) {
    uint32_t s[16], x[16];
    s[0] = 0x61707865u; s[1] = 0x3320646eu; s[2] = 0x79622d32u; s[3] = 0x6b206574u;
    for (int i = 0; i < 8; i++)
        s[4 + i] = (uint32_t)key[4 * i] | ((uint32_t)key[4 * i + 1] << 8) |
                   ((uint32_t)key[4 * i + 2] << 16) | ((uint32_t)key[4 * i + 3] << 24);
    for (int i = 0; i < 4; i++) s[12 + i] = w12_15[i];
    memcpy(x, s, sizeof(s));
    for (int r = 0; r < 10; r++) {
        QR(x[0], x[4], x[8], x[12]); QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]); QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]); QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]); QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

/* ChaCha20Rng::from_seed(seed) then n draws of StandardUniform u64
 * (poly.rs:211-231 consumes them as `val % modulus`). */
void orc_chacha20_u64(const uint8_t seed[32], uint64_t* out, size_t n) {
    uint32_t blk[16];
    uint64_t counter = 0;
    size_t produced = 0;
    while (produced < n) {
        uint32_t w[4] = {(uint32_t)counter, (uint32_t)(counter >> 32), 0, 0};
        orc_chacha20_block(seed, w, blk);
        counter++;
        for (int i = 0; i < 8 && produced < n; i++)
            out[produced++] = (uint64_t)blk[2 * i] | ((uint64_t)blk[2 * i + 1] << 32);
    }
}

/* keyword.rs:546-555: ct_gsw row 0 column i = NTT(-a), a from ChaCha20Rng(RGSW_SEEDS[i]).
 * All six seeds are [7;32] (wire.rs:88-113), so the six columns are equal.
 * out: [2][N] NTT-domain residues of -a. Also returns a (raw mod q) in a_out if non-null. */
void orc_rgsw_neg_a_ntt(uint64_t* out, uint64_t* a_out) {
    init_all();
    uint8_t seed[32];
    memset(seed, 7, 32);
    uint64_t a[N];
    orc_chacha20_u64(seed, a, N);
    for (int i = 0; i < N; i++) a[i] %= MODQ;
    if (a_out) memcpy(a_out, a, sizeof(a));
    for (int c = 0; c < 2; c++) {
        for (int i = 0; i < N; i++) {
            uint64_t neg = MODQ - a[i];                 /* poly.rs:520-524 invert_poly */
            out[c * N + i] = barrett_raw_u64(neg, C.cr1[c], MODULI[c]); /* poly.rs:857-863 */
        }
        ntt_forward_one(out + c * N, c);
    }
}

/* ------------------------------------------------------------------ */
/* DB encode (encode_kernel.cu:18-228, driven by gpu_encode_db at       */
/* encode_kernel.cu:263-380; called from keyword.rs:404-420)           */
/* ------------------------------------------------------------------ */

/* Per DB row r and column group wp (c = db_cols / (d * gamma) groups), the d
 * polynomials P_j[k] = raw[col = (wp*d + j)*gamma + k][r] of Z_p[X]/(X^gamma + 1)
 * are interpolated: placed at bit-reversed positions (gather_tiled_kernel,
 * encode_kernel.cu:150-196), run through log2(d) radix-2 levels whose
 * twiddle is the monomial X^e, e = (2g - 2g*i/size) mod 2g for butterfly i of
 * a group of `size` (compute_twiddle_exponents, encode_kernel.cu:220-228;
 * butterfly_kernel, encode_kernel.cu:18-76), then scaled by num_inv
 * (scale_scatter_tiled_kernel, encode_kernel.cu:85-146) and written back to
 * the same columns, column-major [db_cols][db_rows].  p = modulus (65535);
 * every output is canonical in [0, p). */
void orc_encode(const uint16_t* raw, uint16_t* out, int64_t db_rows, int64_t db_cols, int d, int gamma,
                uint32_t modulus, uint32_t num_inv) {
    int logd = 0;
    while ((1 << logd) < d) logd++;
    const int64_t c = db_cols / ((int64_t)d * gamma);
    const int two_g = 2 * gamma;
#pragma omp parallel
    {
        uint32_t* a = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)d * gamma);
        uint32_t* b = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)d * gamma);
#pragma omp for schedule(dynamic, 1) collapse(2)
        for (int64_t r = 0; r < db_rows; r++) {
            for (int64_t wp = 0; wp < c; wp++) {
                for (int j = 0; j < d; j++) {
                    const int dj = (int)rev_bits((uint32_t)j, logd);
                    for (int k = 0; k < gamma; k++)
                        a[(size_t)dj * gamma + k] = raw[(((wp * d + j) * (int64_t)gamma) + k) * db_rows + r] % modulus;
                }
                for (int level = 0; level < logd; level++) {
                    const int size = 1 << (level + 1), half = 1 << level;
                    for (int g = 0; g < d / size; g++) {
                        for (int i = 0; i < half; i++) {
                            const int e = (two_g - two_g * i / size) % two_g;
                            const uint32_t* ev = a + (size_t)(g * size + i) * gamma;
                            const uint32_t* od = a + (size_t)(g * size + i + half) * gamma;
                            uint32_t* oe = b + (size_t)(g * size + i) * gamma;
                            uint32_t* oo = b + (size_t)(g * size + i + half) * gamma;
                            for (int k = 0; k < gamma; k++) {
                                /* coefficient k of X^e * odd in Z_p[X]/(X^gamma + 1) */
                                const int src = ((k - e) % two_g + two_g) % two_g;
                                const uint32_t t = src < gamma ? od[src] : (modulus - od[src - gamma]) % modulus;
                                oe[k] = (ev[k] + t) % modulus;
                                oo[k] = (ev[k] + modulus - t) % modulus;
                            }
                        }
                    }
                    uint32_t* tmp = a;
                    a = b;
                    b = tmp;
                }
                for (int j = 0; j < d; j++)
                    for (int k = 0; k < gamma; k++)
                        out[(((wp * d + j) * (int64_t)gamma) + k) * db_rows + r] =
                            (uint16_t)(((uint64_t)a[(size_t)j * gamma + k] * num_inv) % modulus);
            }
        }
        free(a);
        free(b);
    }
}

/* ------------------------------------------------------------------ */
/* First pass: GEMV (gpu/kernel.rs:139-237; gemv_kernel.cu:46-103)     */
/* ------------------------------------------------------------------ */

/* c[j] = CRT( sum_k lo32(a[k])*db[j][k] mod q0, sum_k hi32(a[k])*db[j][k] mod q1 ).
 * db is column-major [db_cols][db_rows] u16. */
void orc_gemv(uint64_t* c, const uint64_t* a, const uint16_t* db, int64_t db_rows, int64_t db_cols) {
    init_all();
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < db_cols; j++) {
        const uint16_t* col = db + j * db_rows;
        uint64_t lo = 0, hi = 0;
        for (int64_t k = 0; k < db_rows; k++) {
            lo += (a[k] & 0xffffffffULL) * col[k];
            hi += (a[k] >> 32) * col[k];
        }
        c[j] = crt_compose_2(barrett_raw_u64(lo, C.cr1[0], MODULI[0]),
                             barrett_raw_u64(hi, C.cr1[1], MODULI[1]));
    }
}

/* ------------------------------------------------------------------ */
/* InspiRING online packing (packing.rs:2941-3021 via                  */
/* packing_online_kernel.cu:18-92 and rotation_kernel.cu:21-59)        */
/* ------------------------------------------------------------------ */

/* gen_pows[i] = 5^i mod 2n (packing.rs:1639-1652). */
static uint32_t gen_pow(int i) { return (uint32_t)pow_mod(5, (uint64_t)i, 2 * N); }

/* S[j][crt][z] for outputs j in [j0, j1).  bold_t, bold_t_bar: [instances][3069][N]
 * condensed; bold_t_hat: [instances][3][N]; y_body, z_body: [3][N] condensed.
 * Output is canonical mod q_crt. */
void orc_packing(uint64_t* S, const uint64_t* bold_t, const uint64_t* bold_t_bar,
                 const uint64_t* bold_t_hat, const uint64_t* y_body, const uint64_t* z_body,
                 int64_t instances, int64_t j0, int64_t j1) {
    (void)instances;
    init_all();
    /* rotated keys y_all[i][k][z] = y_body[k][table_{g_i}[z]], y_bar uses 2n - g_i
     * (packing.rs:1756-1785, rotation_kernel.cu:21-59) */
    uint64_t* y_all = (uint64_t*)malloc(sizeof(uint64_t) * NINNER * N);
    uint64_t* y_bar = (uint64_t*)malloc(sizeof(uint64_t) * NINNER * N);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NROT; i++) {
        uint32_t g = gen_pow(i), gb = 2 * N - g;
        for (int z = 0; z < N; z++) {
            uint32_t s1 = auto_index(g, z), s2 = auto_index(gb, z);
            for (int k = 0; k < TEXP; k++) {
                y_all[((size_t)i * TEXP + k) * N + z] = y_body[(size_t)k * N + s1];
                y_bar[((size_t)i * TEXP + k) * N + z] = y_body[(size_t)k * N + s2];
            }
        }
    }
    const uint64_t addition_capacity = 1ULL << (64 - 2 * 28 - 1); /* packing.rs:2963 */
#pragma omp parallel for schedule(dynamic)
    for (int64_t j = j0; j < j1; j++) {
        const uint64_t* bt = bold_t + (size_t)j * NINNER * N;
        const uint64_t* bb = bold_t_bar + (size_t)j * NINNER * N;
        const uint64_t* bh = bold_t_hat + (size_t)j * TEXP * N;
        uint64_t* out = S + (size_t)(j - j0) * 2 * N;
        for (int z = 0; z < N; z++) {
            uint64_t lo = 0, hi = 0, added = 0;
            for (int i = 0; i < NROT; i++) {
                for (int k = 0; k < TEXP; k++) {
                    size_t idx = ((size_t)i * TEXP + k) * N + z;
                    uint64_t y = y_all[idx], b = bt[idx];
                    lo += (y & 0xffffffffULL) * (b & 0xffffffffULL);
                    hi += (y >> 32) * (b >> 32);
                    y = y_bar[idx]; b = bb[idx];
                    lo += (y & 0xffffffffULL) * (b & 0xffffffffULL);
                    hi += (y >> 32) * (b >> 32);
                }
                added += 2 * TEXP;
                if (added >= addition_capacity || i == NROT - 1) {
                    lo = barrett_raw_u64(lo, C.cr1[0], MODULI[0]);
                    hi = barrett_raw_u64(hi, C.cr1[1], MODULI[1]);
                    added = 0;
                }
            }
            for (int k = 0; k < TEXP; k++) {
                uint64_t y = z_body[(size_t)k * N + z], b = bh[(size_t)k * N + z];
                lo += (y & 0xffffffffULL) * (b & 0xffffffffULL);
                hi += (y >> 32) * (b >> 32);
            }
            out[z] = barrett_raw_u64(lo, C.cr1[0], MODULI[0]);
            out[N + z] = barrett_raw_u64(hi, C.cr1[1], MODULI[1]);
        }
    }
    free(y_all);
    free(y_bar);
}

/* Post-process (packing_online.rs:359-398): P_j = CRT(INTT(S_j)) mod q.
 * S: [count][2][N] canonical; P: [count][N]. */
void orc_post(uint64_t* P, const uint64_t* S, int64_t count) {
    init_all();
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < count; j++) {
        uint64_t tmp[2 * N];
        memcpy(tmp, S + (size_t)j * 2 * N, sizeof(tmp));
        ntt_inverse_one(tmp, 0);
        ntt_inverse_one(tmp + N, 1);
        for (int z = 0; z < N; z++) P[(size_t)j * N + z] = crt_compose_2(tmp[z], tmp[N + z]);
    }
}

/* ------------------------------------------------------------------ */
/* RGSW fold (keyword.rs:543-581) and modulus switch                   */
/* ------------------------------------------------------------------ */

/* One external product step: prod = raw(ct_gsw * NTT(G^-1(sum))), 2 rows.
 * neg_a: [2][N] NTT(-a); body: [6][2][N] NTT-domain (ct_gsw row 1). */
static void external_product(uint64_t* prod /*[2][N] mod q*/, const uint64_t* sum /*[2][N]*/,
                             const uint64_t* neg_a, const uint64_t* body) {
    static const int DIG = 3, BP = 19;          /* gadget.rs:3-9: floor(56/3)+1 */
    uint64_t ginv[6][2 * N];
    for (int r = 0; r < 2; r++)
        for (int k = 0; k < DIG; k++)
            for (int z = 0; z < N; z++) {
                uint64_t dgt = (sum[(size_t)r * N + z] >> (BP * k)) & ((1ULL << BP) - 1);
                /* to_ntt: reduce_copy mod q_c (poly.rs:857-876) */
                ginv[r + 2 * k][z] = barrett_raw_u64(dgt, C.cr1[0], MODULI[0]);
                ginv[r + 2 * k][N + z] = barrett_raw_u64(dgt, C.cr1[1], MODULI[1]);
            }
    for (int k = 0; k < 6; k++) {
        ntt_forward_one(ginv[k], 0);
        ntt_forward_one(ginv[k] + N, 1);
    }
    uint64_t pr[2][2 * N];
    for (int c = 0; c < 2; c++)
        for (int z = 0; z < N; z++) {
            uint64_t acc0 = 0, acc1 = 0;
            for (int k = 0; k < 6; k++) {
                uint64_t g = ginv[k][c * N + z];
                acc0 += (neg_a[c * N + z] & 0xffffffffULL) * (g & 0xffffffffULL);
                acc1 += (body[((size_t)k * 2 + c) * N + z] & 0xffffffffULL) * (g & 0xffffffffULL);
            }
            /* poly.rs:636-657 multiply (accumulate, then modular_reduce) */
            pr[0][c * N + z] = barrett_raw_u64(acc0, C.cr1[c], MODULI[c]);
            pr[1][c * N + z] = barrett_raw_u64(acc1, C.cr1[c], MODULI[c]);
        }
    for (int r = 0; r < 2; r++) {
        ntt_inverse_one(pr[r], 0);
        ntt_inverse_one(pr[r] + N, 1);
        for (int z = 0; z < N; z++) prod[(size_t)r * N + z] = crt_compose_2(pr[r][z], pr[r][N + z]);
    }
}

void orc_external_product(uint64_t* prod, const uint64_t* sum, const uint64_t* neg_a, const uint64_t* body) {
    init_all();
    external_product(prod, sum, neg_a, body);
}

/* modulus_switch.rs:45-73 with q_1 = 2^20, q_2 = 268369921 (pir/params.rs:337-357),
 * num_to_keep = gamma = 2048.  out: 12288 bytes. */
#define RESP_BYTES 12288
void orc_switch_and_keep(const uint64_t* ct /*[2][N] mod q*/, uint8_t* out) {
    const uint64_t q1p = 1ULL << 20, q2p = 268369921ULL;
    const int b1 = 20, b2 = 28;
    memset(out, 0, RESP_BYTES);
    size_t bo = 0;
    /* write_bits can touch up to 16 bytes at the last word; the final word never
     * straddles (12288 bytes is a whole number of words), see util.rs:303-321. */
    for (int z = 0; z < N; z++) { write_bits(out, rescale(ct[z], MODQ, q2p), bo, b2); bo += b2; }
    for (int z = 0; z < N; z++) { write_bits(out, rescale(ct[N + z], MODQ, q1p), bo, b1); bo += b1; }
}

/* Fold of one chain (keyword.rs:558-573): packed_a: [d][N] (a_hat rows),
 * packed_b: [d][N] (final_b rows, possibly in [0,2q)).  out_ct: [2][N]. */
void orc_fold_chain(uint64_t* out_ct, const uint64_t* packed_a, const uint64_t* packed_b, int d,
                    const uint64_t* neg_a, const uint64_t* body) {
    init_all();
    uint64_t sum[2 * N], prod[2 * N];
    memset(sum, 0, sizeof(sum));
    for (int i = d - 1; i >= 0; i--) {
        external_product(prod, sum, neg_a, body);
        for (int z = 0; z < N; z++) {
            sum[z] = (prod[z] + packed_a[(size_t)i * N + z]) % MODQ;
            sum[N + z] = (prod[N + z] + packed_b[(size_t)i * N + z]) % MODQ;
        }
    }
    memcpy(out_ct, sum, sizeof(sum));
}

/* ------------------------------------------------------------------ */
/* Wire format (wire.rs:618-705, 708-726)                              */
/* ------------------------------------------------------------------ */

/* Deserialize a keyword query.  y_body, z_body: [3][N] condensed; rows:
 * [nq][db_rows] condensed; bodies: [nq][6][2][N].  Returns num_queries, or -1
 * if the buffer is too short for the declared count. */
int orc_deserialize_query(const uint8_t* buf_in, size_t len, int64_t db_rows, int max_q,
                          uint64_t* y_body, uint64_t* z_body, uint64_t* rows, uint64_t* bodies) {
    if (len < 1) return -1;
    int nq = buf_in[0];
    const int b0 = 28, b1 = 28, tb = 56;   /* wire.rs:120-125 log2_ceil of both moduli */
    size_t need_bits = 8 + (size_t)(TEXP * N * 2 + (size_t)nq * (db_rows + 6 * N)) * tb;
    if ((need_bits + 7) / 8 > len || nq > max_q) return -1;
    uint8_t* buf = (uint8_t*)calloc(len + 16, 1);
    memcpy(buf, buf_in, len);
    size_t bo = 8;
    uint64_t m0 = (1ULL << b0) - 1, m1 = (1ULL << b1) - 1;
    for (int i = 0; i < TEXP; i++)
        for (int k = 0; k < N; k++) {
            uint64_t p = read_bits(buf, bo, tb); bo += tb;
            y_body[(size_t)i * N + k] = (p & m0) | (((p >> b0) & m1) << 32);
            p = read_bits(buf, bo, tb); bo += tb;
            z_body[(size_t)i * N + k] = (p & m0) | (((p >> b0) & m1) << 32);
        }
    for (int q = 0; q < nq; q++) {
        for (int64_t r = 0; r < db_rows; r++) {
            uint64_t p = read_bits(buf, bo, tb); bo += tb;
            rows[(size_t)q * db_rows + r] = (p & m0) | (((p >> b0) & m1) << 32);
        }
        for (int i = 0; i < 6; i++)
            for (int z = 0; z < N; z++) {
                uint64_t p = read_bits(buf, bo, tb); bo += tb;
                bodies[(((size_t)q * 6 + i) * 2 + 0) * N + z] = p & m0;
                bodies[(((size_t)q * 6 + i) * 2 + 1) * N + z] = (p >> b0) & m1;
            }
    }
    free(buf);
    return nq;
}

/* wire.rs:542-615 (the client side; used to build test queries). */
size_t orc_serialize_query(uint8_t* out, size_t cap, int nq, int64_t db_rows, const uint64_t* y_body,
                           const uint64_t* z_body, const uint64_t* rows, const uint64_t* bodies) {
    const int b0 = 28, tb = 56;
    size_t total_values = (size_t)TEXP * N * 2 + (size_t)nq * (db_rows + 6 * N);
    size_t buf_len = 1 + (total_values * tb + 63) / 64 * 8 + 16;
    uint8_t* buf = (uint8_t*)calloc(buf_len, 1);
    buf[0] = (uint8_t)nq;
    size_t bo = 8;
    uint64_t m0 = (1ULL << b0) - 1;
    for (int i = 0; i < TEXP; i++)
        for (int k = 0; k < N; k++) {
            uint64_t v = y_body[(size_t)i * N + k];
            write_bits(buf, (v & m0) | ((v >> 32) << b0), bo, tb); bo += tb;
            v = z_body[(size_t)i * N + k];
            write_bits(buf, (v & m0) | ((v >> 32) << b0), bo, tb); bo += tb;
        }
    for (int q = 0; q < nq; q++) {
        for (int64_t r = 0; r < db_rows; r++) {
            uint64_t v = rows[(size_t)q * db_rows + r];
            write_bits(buf, (v & m0) | ((v >> 32) << b0), bo, tb); bo += tb;
        }
        for (int i = 0; i < 6; i++)
            for (int z = 0; z < N; z++) {
                uint64_t lo = bodies[(((size_t)q * 6 + i) * 2 + 0) * N + z];
                uint64_t hi = bodies[(((size_t)q * 6 + i) * 2 + 1) * N + z];
                write_bits(buf, (lo & m0) | (hi << b0), bo, tb); bo += tb;
            }
    }
    size_t used = (bo + 7) / 8;
    if (used > cap) { free(buf); return 0; }
    memcpy(out, buf, used);
    free(buf);
    return used;
}

/* ------------------------------------------------------------------ */
/* Whole respond path (keyword.rs:534-628)                             */
/* ------------------------------------------------------------------ */

/* State: db [db_cols][db_rows] u16 (encoded); bold_t/bar/hat as in orc_packing;
 * a_hat [instances][N] raw mod q.  d = interpolate_degree, c = instances/d.
 * Output: for each hash query, c * 12288 bytes (the response blob before
 * serialization), concatenated.  Returns 0 on success. */
int orc_respond_raw(uint8_t* out, const uint16_t* db, int64_t db_rows, int64_t instances, int d,
                    const uint64_t* bold_t, const uint64_t* bold_t_bar, const uint64_t* bold_t_hat,
                    const uint64_t* a_hat, int nq, const uint64_t* y_body, const uint64_t* z_body,
                    const uint64_t* rows, const uint64_t* bodies) {
    init_all();
    if (instances % d) return -1;
    int64_t db_cols = instances * N;
    int c = (int)(instances / d);
    uint64_t neg_a[2 * N];
    orc_rgsw_neg_a_ntt(neg_a, NULL);
    /* The packing inner product depends only on the keys (packing_online_kernel.cu:46-91),
     * so one evaluation serves every hash query; the reference repeats it per query
     * with the same result. */
    uint64_t* S = (uint64_t*)malloc(sizeof(uint64_t) * instances * 2 * N);
    uint64_t* P = (uint64_t*)malloc(sizeof(uint64_t) * instances * N);
    uint64_t* cvec = (uint64_t*)malloc(sizeof(uint64_t) * db_cols);
    uint64_t* fb = (uint64_t*)malloc(sizeof(uint64_t) * instances * N);
    orc_packing(S, bold_t, bold_t_bar, bold_t_hat, y_body, z_body, instances, 0, instances);
    orc_post(P, S, instances);
    for (int q = 0; q < nq; q++) {
        orc_gemv(cvec, rows + (size_t)q * db_rows, db, db_rows, db_cols);
        /* add_raw(b_poly, sum_raw): unreduced, in [0, 2q) (poly.rs:492-496) */
        for (int64_t i = 0; i < instances * N; i++) fb[i] = cvec[i] + P[i];
        const uint64_t* body = bodies + (size_t)q * 6 * 2 * N;
        uint8_t* qout = out + (size_t)q * c * RESP_BYTES;
#pragma omp parallel for schedule(dynamic)
        for (int w = 0; w < c; w++) {
            uint64_t ct[2 * N];
            orc_fold_chain(ct, a_hat + (size_t)w * d * N, fb + (size_t)w * d * N, d, neg_a, body);
            orc_switch_and_keep(ct, qout + (size_t)w * RESP_BYTES);
        }
    }
    free(S); free(P); free(cvec); free(fb);
    return 0;
}

/* wire.rs:708-726 */
size_t orc_serialize_response(uint8_t* out, size_t cap, int nresp, const uint8_t* const* resp,
                              const uint32_t* resp_len, int nstash, const uint8_t* const* stash,
                              const uint32_t* stash_len) {
    size_t need = 4 + 4;
    for (int i = 0; i < nresp; i++) need += 4 + resp_len[i];
    for (int i = 0; i < nstash; i++) need += 4 + stash_len[i];
    if (need > cap) return 0;
    size_t o = 0;
    uint32_t v = (uint32_t)nresp;
    memcpy(out + o, &v, 4); o += 4;
    for (int i = 0; i < nresp; i++) {
        memcpy(out + o, &resp_len[i], 4); o += 4;
        memcpy(out + o, resp[i], resp_len[i]); o += resp_len[i];
    }
    v = (uint32_t)nstash;
    memcpy(out + o, &v, 4); o += 4;
    for (int i = 0; i < nstash; i++) {
        memcpy(out + o, &stash_len[i], 4); o += 4;
        memcpy(out + o, stash[i], stash_len[i]); o += stash_len[i];
    }
    return o;
}

/* ------------------------------------------------------------------ */
/* Synthetic state generator (shared definition with the HIP library) */
/* ------------------------------------------------------------------ */

/* Counter-based generator: value = mix(seed, stream, index).  The HIP
 * library implements the same function so that device-resident synthetic
 * state and the oracle's host copy agree without a 35 GB transfer. */
static uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
uint64_t orc_synth_word(uint64_t seed, uint32_t stream, uint64_t idx) {
    return mix64(mix64(seed ^ ((uint64_t)stream << 40)) + idx);
}

/* kind 0: condensed residue pair (q0 | q1<<32); 1: residue mod q; 2: u16 in [0, 65534];
 * 3: per-crt residue for [.. ][2][N] layouts (index parity of the crt position
 * is derived from (idx / N) & 1); 4: small "sk = 0 client" residues (value < 2^20). */
void orc_synth_fill(void* out, uint64_t seed, uint32_t stream, uint64_t start, uint64_t count, int kind) {
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)count; ii++) {
        uint64_t i = (uint64_t)ii;
        uint64_t h = orc_synth_word(seed, stream, start + i);
        switch (kind) {
        case 0: {
            uint64_t lo = ((h & 0xffffffffULL) * MODULI[0]) >> 32;
            uint64_t hi = ((h >> 32) * MODULI[1]) >> 32;
            ((uint64_t*)out)[i] = lo | (hi << 32);
            break;
        }
        case 1: ((uint64_t*)out)[i] = (uint64_t)(((u128)h * MODQ) >> 64); break;
        case 2: ((uint16_t*)out)[i] = (uint16_t)((h & 0xffffULL) % 65535ULL); break;
        case 3: {
            int crt = (int)(((start + i) / N) & 1);
            ((uint64_t*)out)[i] = ((h & 0xffffffffULL) * MODULI[crt]) >> 32;
            break;
        }
        case 4: {
            uint64_t v = h & 0xfffffULL;
            ((uint64_t*)out)[i] = (v % MODULI[0]) | ((v % MODULI[1]) << 32);
            break;
        }
        default: break;
        }
    }
}

/* orc_gemv for nq query rows over columns [j0, j0 + ncols) of a synthetic DB
 * (kind 2 values of [db_cols][db_rows] at index col * db_rows + row), each
 * value drawn from the generator where it is used rather than from a host
 * copy: c[h][j] = sum_k a[h][k] * db[j0 + j][k] mod q, as orc_gemv. */
void orc_gemv_synth(uint64_t* c, const uint64_t* a, int nq, uint64_t seed, uint32_t stream, int64_t db_rows,
                    int64_t j0, int64_t ncols) {
    init_all();
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ncols; j++) {
        uint64_t lo[8] = {0}, hi[8] = {0};
        const uint64_t base = (uint64_t)(j0 + j) * (uint64_t)db_rows;
        for (int64_t k = 0; k < db_rows; k++) {
            const uint64_t v = (orc_synth_word(seed, stream, base + (uint64_t)k) & 0xffffULL) % 65535ULL;
            for (int h = 0; h < nq; h++) {
                lo[h] += (a[h * db_rows + k] & 0xffffffffULL) * v;
                hi[h] += (a[h * db_rows + k] >> 32) * v;
            }
        }
        for (int h = 0; h < nq; h++)
            c[h * ncols + j] = crt_compose_2(barrett_raw_u64(lo[h], C.cr1[0], MODULI[0]),
                                             barrett_raw_u64(hi[h], C.cr1[1], MODULI[1]));
    }
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ================================================================== */
/* Client and server setup (the producers of the respond path's state) */
/* ================================================================== */
/* These restate the reference's client (client/keyword.rs, client/base.rs,
 * client/online.rs), its packing keys (packing.rs:1838-1945, 3272-3294) and
 * the server's offline precomputation (server/engine.rs:678-1000 CPU path:
 * hint, prep_pack_many_lwes, full_packing_with_preprocessing_offline).  They
 * exist so that tests can run real ciphertexts through the respond path and
 * check that the client decrypts the stored values: that end-to-end decrypt
 * is what pins the packing, post-process, fold and modulus switch semantics.
 *
 * Randomness.  Public randomness (SEED_0 query masks, W_SEED/V_SEED packing
 * masks, the RGSW seeds) comes from the ChaCha20 restatement above, exactly
 * as the reference consumes it (one u64 per coefficient, `% modulus`).  Secret
 * keys and noise come from the reference's OS randomness (rand::random), so no
 * vector of them exists; here they come from the counter-based generator with
 * the reference's distributions (ternary key of Hamming weight h, discrete
 * Gaussian of width 16.042421 by CDT, discrete_gaussian.rs:82-133). */

#define NOISE_WIDTH 16.042421          /* pir/params.rs:69-96 */
#define DG_MAX 65                      /* ceil(16.042421 * NUM_WIDTHS=4) */
static uint64_t DG_CDF[2 * DG_MAX + 1];
static int dg_init = 0;

/* discrete_gaussian.rs:82-109: CDF of p(i) ~ exp(-pi i^2 / w^2), i in [-max, max]. */
static void dg_setup(void) {
    if (dg_init) return;
    double table[2 * DG_MAX + 1], total = 0.0;
    for (int i = -DG_MAX; i <= DG_MAX; i++) {
        double p = __builtin_exp(-3.14159265358979323846 * (double)i * (double)i / (NOISE_WIDTH * NOISE_WIDTH));
        table[i + DG_MAX] = p;
        total += p;
    }
    double cum = 0.0;
    for (int i = 0; i < 2 * DG_MAX + 1; i++) {
        cum += table[i] / total;
        double v = cum * 18446744073709551615.0;
        DG_CDF[i] = v >= 18446744073709551615.0 ? UINT64_MAX : (uint64_t)(v + 0.5);
    }
    dg_init = 1;
}

/* discrete_gaussian.rs:111-133: the smallest i with u <= cdf[i]; value i - max mod m. */
static uint64_t dg_sample(uint64_t u, uint64_t m) {
    int out = 2 * DG_MAX;
    for (int i = 2 * DG_MAX; i >= 0; i--)
        if (u <= DG_CDF[i]) out = i;
    int64_t v = (int64_t)out - DG_MAX;
    return v < 0 ? (uint64_t)(v + (int64_t)m) : (uint64_t)v;
}

/* PolyMatrixRaw::noise (poly.rs:238-262): N coefficients mod q. */
static void noise_poly(uint64_t* e, uint64_t seed, uint32_t stream, uint64_t start) {
    dg_setup();
    for (int z = 0; z < N; z++) e[z] = dg_sample(orc_synth_word(seed, stream, start + z), MODQ);
}

/* raw poly mod q -> NTT residues [2][N] (poly.rs:857-876 to_ntt). */
static void raw_to_ntt(const uint64_t* raw, uint64_t* out) {
    for (int c = 0; c < 2; c++) {
        for (int z = 0; z < N; z++) out[c * N + z] = raw[z] % MODULI[c];
        ntt_forward_one(out + c * N, c);
    }
}

/* NTT residues [2][N] -> raw poly mod q (poly.rs:900-918 from_ntt: INTT + CRT). */
static void ntt_to_raw(const uint64_t* in, uint64_t* raw) {
    uint64_t t[2 * N];
    memcpy(t, in, sizeof(t));
    ntt_inverse_one(t, 0);
    ntt_inverse_one(t + N, 1);
    for (int z = 0; z < N; z++) raw[z] = crt_compose_2(t[z], t[N + z]);
}

/* out = a * b (+ out if acc), per CRT prime, canonical. */
static void ntt_mul(uint64_t* out, const uint64_t* a, const uint64_t* b, int acc) {
    for (int c = 0; c < 2; c++)
        for (int z = 0; z < N; z++) {
            uint64_t v = (a[c * N + z] * b[c * N + z]) % MODULI[c];
            out[c * N + z] = acc ? (out[c * N + z] + v) % MODULI[c] : v;
        }
}

static void ntt_add(uint64_t* out, const uint64_t* a) {
    for (int c = 0; c < 2; c++)
        for (int z = 0; z < N; z++) out[c * N + z] = (out[c * N + z] + a[c * N + z]) % MODULI[c];
}

static void condense(const uint64_t* ntt, uint64_t* out) { /* packing.rs:867-884 */
    for (int z = 0; z < N; z++) out[z] = ntt[z] | (ntt[N + z] << 32);
}

/* gen_ternary_mat (client/base.rs:126-139): h coefficients 1, h coefficients
 * q - 1, shuffled.  h = 0 is the keyword client's key: KeywordClient builds
 * its YClient from Client::init (client/keyword.rs:88), whose sk_reg is the
 * zero polynomial (client/base.rs:370-389). */
void orc_client_secret(uint64_t seed, int hamming, uint64_t* sk) {
    memset(sk, 0, sizeof(uint64_t) * N);
    for (int i = 0; i < hamming; i++) sk[i] = 1;
    for (int i = hamming; i < 2 * hamming; i++) sk[i] = MODQ - 1;
    for (int i = N - 1; i > 0; i--) { /* Fisher-Yates */
        uint64_t j = orc_synth_word(seed, 0x51, (uint64_t)i) % (uint64_t)(i + 1);
        uint64_t t = sk[i]; sk[i] = sk[j]; sk[j] = t;
    }
}

/* PolyMatrixNTT::random_rng (poly.rs:418-442): per poly, crt 0 then crt 1, one
 * u64 per residue, `% q_c`.  out: [npolys][2][N]. */
void orc_random_ntt(const uint8_t seed[32], uint64_t* out, int npolys) {
    size_t n = (size_t)npolys * 2 * N;
    orc_chacha20_u64(seed, out, n);
    for (size_t i = 0; i < n; i++) out[i] %= MODULI[(i / N) & 1];
}

static void packing_masks(uint64_t* w_mask /*[3][2][N]*/, uint64_t* v_mask /*[3][2][N]*/) {
    uint8_t ws[32], vs[32];
    memset(ws, 7, 32);             /* W_SEED, pir/scheme.rs:46-48 */
    memset(vs, 7, 32);
    vs[0] = 8; vs[1] = 8;          /* V_SEED, pir/scheme.rs:49-51 */
    orc_random_ntt(ws, w_mask, TEXP);
    orc_random_ntt(vs, v_mask, TEXP);
}

/* generate_ksk_body (packing.rs:3272-3294): for k < 3,
 *   body_k = NTT(tau_gen(s)) * 2^(19k) - NTT(s) * mask_k + NTT(e_k).
 * mask: [3][2][N]; out: [3][N] condensed (condense_matrix). */
static void ksk_body(uint64_t* out, const uint64_t* sk, uint32_t gen, const uint64_t* mask,
                     uint64_t seed, uint32_t stream) {
    uint64_t tau[N], s_ntt[2 * N], tau_ntt[2 * N], e[N], e_ntt[2 * N], acc[2 * N];
    orc_automorph_raw(sk, tau, gen, MODQ);
    raw_to_ntt(sk, s_ntt);
    raw_to_ntt(tau, tau_ntt);
    for (int k = 0; k < TEXP; k++) {
        noise_poly(e, seed, stream, (uint64_t)k * N);
        raw_to_ntt(e, e_ntt);
        for (int c = 0; c < 2; c++) {
            uint64_t m = MODULI[c], g = (1ULL << (19 * k)) % m;   /* build_gadget, gadget.rs:11-33 */
            for (int z = 0; z < N; z++) {
                uint64_t neg_mask = (m - mask[((size_t)k * 2 + c) * N + z]) % m;
                uint64_t v = (tau_ntt[c * N + z] * g) % m;
                v = (v + s_ntt[c * N + z] * neg_mask) % m;
                acc[c * N + z] = (v + e_ntt[c * N + z]) % m;
            }
        }
        condense(acc, out + (size_t)k * N);
    }
}

/* PackingKeys::init_full (packing.rs:1895-1944): y_body with generator
 * gen_pows[1] = 5 and mask W, z_body with 2n - 1 and mask V.  [3][N] condensed each. */
void orc_client_packing_keys(uint64_t* y_body, uint64_t* z_body, const uint64_t* sk, uint64_t seed) {
    init_all();
    uint64_t w_mask[TEXP * 2 * N], v_mask[TEXP * 2 * N];
    packing_masks(w_mask, v_mask);
    ksk_body(y_body, sk, 5, w_mask, seed, 0x61);
    ksk_body(z_body, sk, 2 * N - 1, v_mask, seed, 0x62);
}

/* Query row (client/keyword.rs:180-187 -> generate_query_b_values,
 * client/online.rs:226-287, 340-357, pack_query 26-52): db_rows/N RLWE
 * encryptions under the SEED_0 public masks (ChaCha20Rng(get_seed(0)) = the
 * all-zero seed, one mask polynomial per RLWE, consumed in order); the RLWE
 * holding the target row encrypts scale_k = floor(q/p) at coefficient
 * target_row % N.  Each b = s*a + e + m (get_scaled_regev_sample with factor 1,
 * client/base.rs:438-457).  out: [db_rows] condensed (b mod q0 | b mod q1 << 32). */
void orc_client_query_row(uint64_t* out, int64_t db_rows, int64_t target_row, const uint64_t* sk,
                          uint64_t seed, uint32_t stream) {
    init_all();
    const uint64_t scale_k = MODQ / 65535ULL;
    int64_t nrlwe = db_rows / N;
    uint8_t zero_seed[32];
    memset(zero_seed, 0, 32);
    uint64_t* a_all = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nrlwe * N);
    orc_chacha20_u64(zero_seed, a_all, (size_t)nrlwe * N);
    uint64_t s_ntt[2 * N];
    raw_to_ntt(sk, s_ntt);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nrlwe; k++) {
        uint64_t a[N], a_ntt[2 * N], e[N], b_ntt[2 * N], b[N];
        for (int z = 0; z < N; z++) a[z] = a_all[(size_t)k * N + z] % MODQ;
        raw_to_ntt(a, a_ntt);
        ntt_mul(b_ntt, s_ntt, a_ntt, 0);
        ntt_to_raw(b_ntt, b);
        noise_poly(e, seed, stream, (uint64_t)k * N);
        for (int z = 0; z < N; z++) {
            uint64_t v = (b[z] + e[z]) % MODQ;
            if (k * N + z == target_row) v = (v + scale_k) % MODQ;
            out[(size_t)k * N + z] = (v % MODULI[0]) | ((v % MODULI[1]) << 32);
        }
    }
    free(a_all);
}

/* ct_gsw_body (client/keyword.rs:150-178): for j < t_gsw = 3, with
 * sigma = +-2^(19j) X^(exponent mod n) (negative for exponent >= n),
 * column 2j+1 = body of Enc(sigma) and column 2j = body of Enc(s * sigma), both
 * under the RGSW public mask (all six RGSW seeds are [7;32], so the mask is the
 * one orc_rgsw_neg_a_ntt negates).  body = NTT(s)*NTT(a) + NTT(e) + msg
 * (encrypt_matrix_reg, client/base.rs:502-511).  out: [6][2][N] NTT residues. */
void orc_client_rgsw_body(uint64_t* out, const uint64_t* sk, uint32_t exponent, uint64_t seed,
                          uint32_t stream) {
    init_all();
    uint64_t neg_a[2 * N], a[N], a_ntt[2 * N], s_ntt[2 * N], sa[2 * N];
    orc_rgsw_neg_a_ntt(neg_a, a);
    raw_to_ntt(a, a_ntt);
    raw_to_ntt(sk, s_ntt);
    ntt_mul(sa, s_ntt, a_ntt, 0);
    exponent %= 2 * N;
    for (int j = 0; j < TEXP; j++) {
        uint64_t sig[N], sig_ntt[2 * N], prod[2 * N], e[N], e_ntt[2 * N];
        memset(sig, 0, sizeof(sig));
        uint64_t g = 1ULL << (19 * j);
        sig[exponent % N] = exponent < N ? g : MODQ - g;
        raw_to_ntt(sig, sig_ntt);
        ntt_mul(prod, s_ntt, sig_ntt, 0);
        for (int half = 0; half < 2; half++) {
            int col = 2 * j + (half ? 0 : 1);
            const uint64_t* msg = half ? prod : sig_ntt;
            noise_poly(e, seed, stream, (uint64_t)col * N);
            raw_to_ntt(e, e_ntt);
            uint64_t* o = out + (size_t)col * 2 * N;
            for (int c = 0; c < 2; c++)
                for (int z = 0; z < N; z++)
                    o[c * N + z] = (sa[c * N + z] + e_ntt[c * N + z] + msg[c * N + z]) % MODULI[c];
        }
    }
}

/* Hint + prep_pack (server/engine.rs:733-800 -> answer_hint_ring_gpu over the
 * SEED_0 masks, then prep_pack_many_lwes, packing.rs:1032-1104).
 *
 * LWE row r = kN + i of the query is coefficient i of RLWE k, whose row 0 is
 * abar_k = -a_k; its mask vector is negacyclic_perm(abar_k, i) (utils.rs:5-35),
 * i.e. <mask, s> = coef_i(abar_k * s).  The hint column of DB column j is
 * A_j = sum_r DB[j][r] * mask_r, and prep_pack turns it into the polynomial
 * at_j = negacyclic_perm(A_j, 0), the unique polynomial with
 * const(at_j * s) = <A_j, s> for every s.  Since coef_i(P) = const(X^-i P),
 *   at_j = sum_k Dt_jk * abar_k,  Dt_jk(X) = sum_i DB[j][kN+i] X^-i,
 * a negacyclic product computed here in the NTT domain.  (Any exact method
 * gives the same at_j mod q; the test decrypts through it.)
 * db: encoded, column-major [db_cols][db_rows] u16.  out: NTT residues
 * [ncols][2][N] for columns c0 .. c0+ncols-1. */
void orc_setup_atilde(uint64_t* out, const uint16_t* db, int64_t db_rows, int64_t c0, int64_t ncols) {
    init_all();
    int64_t nrlwe = db_rows / N;
    uint8_t zero_seed[32];
    memset(zero_seed, 0, 32);
    uint64_t* a_all = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nrlwe * N);
    uint64_t* abar = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nrlwe * 2 * N);
    orc_chacha20_u64(zero_seed, a_all, (size_t)nrlwe * N);
    for (int64_t k = 0; k < nrlwe; k++) {
        uint64_t t[N];
        for (int z = 0; z < N; z++) {
            uint64_t a = a_all[(size_t)k * N + z] % MODQ;
            t[z] = a ? MODQ - a : 0;
        }
        raw_to_ntt(t, abar + (size_t)k * 2 * N);
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t jj = 0; jj < ncols; jj++) {
        const uint16_t* col = db + (size_t)(c0 + jj) * db_rows;
        uint64_t* o = out + (size_t)jj * 2 * N;
        uint64_t dt[N], dt_ntt[2 * N];
        memset(o, 0, sizeof(uint64_t) * 2 * N);
        for (int64_t k = 0; k < nrlwe; k++) {
            const uint16_t* dk = col + k * N;
            dt[0] = dk[0];
            for (int m = 1; m < N; m++) dt[m] = dk[N - m] ? MODQ - dk[N - m] : 0;
            raw_to_ntt(dt, dt_ntt);
            ntt_mul(o, dt_ntt, abar + (size_t)k * 2 * N, 1);
        }
    }
    free(a_all);
    free(abar);
}

/* acc[c][m] += (X^e * a)[m] for a given as residues a[c][N] and their
 * negations na[c][N] (values < 2^28, so 2048 lazy additions fit in u64). */
static void add_monomial_times(uint64_t* acc, const uint32_t* a, const uint32_t* na, uint32_t e) {
    int neg = e >= N;
    uint32_t s = e % N;
    for (int c = 0; c < 2; c++) {
        const uint32_t* hi = (neg ? na : a) + c * N;   /* X^s a: coefficients s.. */
        const uint32_t* lo = (neg ? a : na) + c * N;   /* wrapped ones change sign */
        uint64_t* o = acc + c * N;
        for (uint32_t m = s; m < N; m++) o[m] += hi[m - s];
        for (uint32_t m = 0; m < s; m++) o[m] += lo[m + N - s];
    }
}

/* full_packing_with_preprocessing_offline_inner, CPU path
 * (packing.rs:2477-2750), for `instances` outputs of N columns each.
 * at: NTT residues [instances][N][2][N] from orc_setup_atilde.
 *   r_i    = tau_{g_i}     (n^-1 * sum_j mono(j*g_-i)        * at_j),
 *   rbar_i = tau_{2n-g_i}  (n^-1 * sum_j mono(2n - j*g_-i)   * at_j),
 * for i < n/2 (g_i = 5^i mod 2n, g_-i = gen_pows[(n-i) mod n]); the monomial
 * products are done as signed shifts in the coefficient domain, which equals
 * the reference's NTT-domain products with monomial_ntts.  Then the collapse
 * (2694-2749): for i = n/2-2 .. 0, bold_t[i] = NTT(G^-1(raw(r_{i+1}))),
 * r_i += sum_k w_all[i][k] * bold_t[i][k], likewise for the bars with w_bar_all;
 * bold_t_hat = NTT(G^-1(raw(rbar_0))), r_0 += sum_k v_mask[k] * bold_t_hat[k],
 * a_hat = raw(r_0).  w_all/w_bar_all are the W mask rotated by g_i and 2n - g_i
 * (OfflinePackingKeys::init_full, packing.rs:1838-1872 + generate_rotations_double).
 * Outputs: bold_t, bold_t_bar [instances][3069][N] and bold_t_hat [instances][3][N]
 * condensed; a_hat [instances][N] raw mod q. */
void orc_setup_precomp(const uint64_t* at, int64_t instances, uint64_t* bold_t, uint64_t* bold_t_bar,
                       uint64_t* bold_t_hat, uint64_t* a_hat) {
    init_all();
    const int HALF = N / 2;
    uint64_t* w_mask = (uint64_t*)malloc(sizeof(uint64_t) * TEXP * 2 * N);
    uint64_t* v_mask = (uint64_t*)malloc(sizeof(uint64_t) * TEXP * 2 * N);
    packing_masks(w_mask, v_mask);
    /* w_all[i][k][c][z] = w_mask[k][c][table_{g_i}[z]] (apply_automorph_ntt_double) */
    uint64_t* w_all = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)NROT * TEXP * 2 * N);
    uint64_t* w_bar = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)NROT * TEXP * 2 * N);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NROT; i++) {
        uint32_t g = gen_pow(i), gb = 2 * N - g;
        for (int k = 0; k < TEXP; k++)
            for (int c = 0; c < 2; c++)
                for (int z = 0; z < N; z++) {
                    size_t o = (((size_t)i * TEXP + k) * 2 + c) * N + z;
                    w_all[o] = w_mask[((size_t)k * 2 + c) * N + auto_index(g, z)];
                    w_bar[o] = w_mask[((size_t)k * 2 + c) * N + auto_index(gb, z)];
                }
    }
    const uint64_t ninv[2] = {inv_mod(N, MODULI[0]), inv_mod(N, MODULI[1])};
    uint32_t* a_res = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)N * 2 * N);
    uint32_t* na_res = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)N * 2 * N);
    uint64_t* r_all = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)HALF * 2 * N);
    uint64_t* r_bar = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)HALF * 2 * N);
    for (int64_t t = 0; t < instances; t++) {
        /* coefficient-domain residues of every at_j and of its negation */
#pragma omp parallel for schedule(static)
        for (int j = 0; j < N; j++) {
            uint64_t tmp[2 * N];
            memcpy(tmp, at + ((size_t)t * N + j) * 2 * N, sizeof(tmp));
            ntt_inverse_one(tmp, 0);
            ntt_inverse_one(tmp + N, 1);
            for (int c = 0; c < 2; c++)
                for (int z = 0; z < N; z++) {
                    uint32_t v = (uint32_t)tmp[c * N + z];
                    a_res[((size_t)j * 2 + c) * N + z] = v;
                    na_res[((size_t)j * 2 + c) * N + z] = v ? (uint32_t)MODULI[c] - v : 0;
                }
        }
#pragma omp parallel for schedule(dynamic, 4)
        for (int i = 0; i < HALF; i++) {
            uint64_t acc[2 * N], accb[2 * N], raw_c[2][N], rot[2 * N];
            memset(acc, 0, sizeof(acc));
            memset(accb, 0, sizeof(accb));
            uint32_t gmi = gen_pow((N - i) % N);
            for (uint32_t j = 0; j < N; j++) {
                uint32_t e = (uint32_t)(((uint64_t)j * gmi) % (2 * N));
                add_monomial_times(acc, a_res + (size_t)j * 2 * N, na_res + (size_t)j * 2 * N, e);
                add_monomial_times(accb, a_res + (size_t)j * 2 * N, na_res + (size_t)j * 2 * N,
                                   (2 * N - e) % (2 * N));
            }
            for (int which = 0; which < 2; which++) {
                uint64_t* src = which ? accb : acc;
                uint32_t tt = which ? 2 * N - gen_pow(i) : gen_pow(i);
                for (int c = 0; c < 2; c++) {
                    for (int z = 0; z < N; z++) raw_c[c][z] = ((src[c * N + z] % MODULI[c]) * ninv[c]) % MODULI[c];
                    /* automorph_poly per residue (poly.rs:543-556) */
                    for (uint64_t z = 0; z < N; z++) {
                        uint64_t num = (z * tt) / N, rem = (z * tt) % N;
                        uint64_t v = raw_c[c][z];
                        rot[c * N + rem] = (num % 2 == 0) ? v : (v ? MODULI[c] - v : 0);
                    }
                    ntt_forward_one(rot + c * N, c);
                }
                memcpy((which ? r_bar : r_all) + (size_t)i * 2 * N, rot, sizeof(rot));
            }
        }
        /* collapse: two independent chains (r and rbar) */
        uint64_t* bt = bold_t + (size_t)t * NINNER * N;
        uint64_t* btb = bold_t_bar + (size_t)t * NINNER * N;
#pragma omp parallel for schedule(static, 1) num_threads(2)
        for (int which = 0; which < 2; which++) {
            uint64_t* r = which ? r_bar : r_all;
            const uint64_t* w = which ? w_bar : w_all;
            uint64_t* outb = which ? btb : bt;
            for (int i = HALF - 2; i >= 0; i--) {
                uint64_t raw[N], dig[2 * N], res[2 * N];
                ntt_to_raw(r + (size_t)(i + 1) * 2 * N, raw);
                memset(res, 0, sizeof(res));
                for (int k = 0; k < TEXP; k++) {
                    for (int c = 0; c < 2; c++) {
                        for (int z = 0; z < N; z++) dig[c * N + z] = (raw[z] >> (19 * k)) & ((1ULL << 19) - 1);
                        ntt_forward_one(dig + c * N, c);
                    }
                    condense(dig, outb + ((size_t)i * TEXP + k) * N);
                    ntt_mul(res, w + (((size_t)i * TEXP + k) * 2) * N, dig, 1);
                }
                ntt_add(r + (size_t)i * 2 * N, res);
            }
        }
        {
            uint64_t raw[N], dig[2 * N], res[2 * N];
            ntt_to_raw(r_bar, raw);
            memset(res, 0, sizeof(res));
            for (int k = 0; k < TEXP; k++) {
                for (int c = 0; c < 2; c++) {
                    for (int z = 0; z < N; z++) dig[c * N + z] = (raw[z] >> (19 * k)) & ((1ULL << 19) - 1);
                    ntt_forward_one(dig + c * N, c);
                }
                condense(dig, bold_t_hat + ((size_t)t * TEXP + k) * N);
                ntt_mul(res, v_mask + (size_t)k * 2 * N, dig, 1);
            }
            ntt_add(r_all, res);
            ntt_to_raw(r_all, a_hat + (size_t)t * N);
        }
    }
    free(w_mask); free(v_mask); free(w_all); free(w_bar);
    free(a_res); free(na_res); free(r_all); free(r_bar);
}

/* recover_how_many (modulus_switch.rs:102-134) with how_many = gamma = N, then
 * decrypt_ct_reg_measured (client/online.rs:118-136): s * row0 + row1 in the
 * NTT domain, raw, rescaled q -> p = 65535.  blob: 12288 bytes; out: [N]. */
void orc_client_decrypt(const uint8_t* blob, const uint64_t* sk, uint64_t* out) {
    init_all();
    const uint64_t q1p = 1ULL << 20, q2p = 268369921ULL;
    uint64_t r0[N], r1[N], r0n[2 * N], r1n[2 * N], sn[2 * N], dec[2 * N], raw[N];
    size_t bo = 0;
    for (int z = 0; z < N; z++) { r0[z] = rescale(read_bits(blob, bo, 28), q2p, MODQ); bo += 28; }
    for (int z = 0; z < N; z++) { r1[z] = rescale(read_bits(blob, bo, 20), q1p, MODQ); bo += 20; }
    raw_to_ntt(r0, r0n);
    raw_to_ntt(r1, r1n);
    raw_to_ntt(sk, sn);
    ntt_mul(dec, sn, r0n, 0);
    ntt_add(dec, r1n);
    ntt_to_raw(dec, raw);
    for (int z = 0; z < N; z++) out[z] = rescale(raw[z], MODQ, 65535ULL);
}

/* Decrypt of an unswitched RLWE (a, b) [2][N] raw mod q (no modulus switch):
 * rescale(s*a + b, q, p).  Used to check the packing stage on its own. */
void orc_client_decrypt_raw(const uint64_t* ct, const uint64_t* sk, uint64_t* out) {
    init_all();
    uint64_t an[2 * N], bn[2 * N], sn[2 * N], dec[2 * N], raw[N], b[N];
    for (int z = 0; z < N; z++) b[z] = ct[N + z] % MODQ;
    raw_to_ntt(ct, an);
    raw_to_ntt(b, bn);
    raw_to_ntt(sk, sn);
    ntt_mul(dec, sn, an, 0);
    ntt_add(dec, bn);
    ntt_to_raw(dec, raw);
    for (int z = 0; z < N; z++) out[z] = rescale(raw[z], MODQ, 65535ULL);
}

/* SipHash-2-4 (Aumasson and Bernstein 2012; siphasher::sip::SipHasher with
 * new_with_keys, write, finish), as CuckooHash::hash uses it
 * (pir/keyword/cuckoo.rs:98-124). */
uint64_t orc_siphash24(uint64_t k0, uint64_t k1, const uint8_t* m, size_t len) {
#define SROTL(x, b) (((x) << (b)) | ((x) >> (64 - (b))))
#define SROUND() do { v0 += v1; v1 = SROTL(v1, 13); v1 ^= v0; v0 = SROTL(v0, 32); \
                      v2 += v3; v3 = SROTL(v3, 16); v3 ^= v2; \
                      v0 += v3; v3 = SROTL(v3, 21); v3 ^= v0; \
                      v2 += v1; v1 = SROTL(v1, 17); v1 ^= v2; v2 = SROTL(v2, 32); } while (0)
    uint64_t v0 = k0 ^ 0x736f6d6570736575ULL, v1 = k1 ^ 0x646f72616e646f6dULL;
    uint64_t v2 = k0 ^ 0x6c7967656e657261ULL, v3 = k1 ^ 0x7465646279746573ULL;
    size_t i = 0;
    for (; i + 8 <= len; i += 8) {
        uint64_t w = 0;
        for (int b = 0; b < 8; b++) w |= (uint64_t)m[i + b] << (8 * b);
        v3 ^= w; SROUND(); SROUND(); v0 ^= w;
    }
    uint64_t w = (uint64_t)(len & 0xff) << 56;
    for (int b = 0; i + b < len; b++) w |= (uint64_t)m[i + b] << (8 * b);
    v3 ^= w; SROUND(); SROUND(); v0 ^= w;
    v2 ^= 0xff;
    SROUND(); SROUND(); SROUND(); SROUND();
    return v0 ^ v1 ^ v2 ^ v3;
#undef SROUND
#undef SROTL
}
