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

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
