/*
 * fsmall_oracle.c -- CPU restatement of FLINT's fft_small multiplication
 * path, used as the checker for the GPU product.
 *
 * TEST INFRASTRUCTURE ONLY (see fsmall_oracle.h).
 *
 * Every function names the reference file:line it restates.  The transforms
 * follow the same mathematics as the reference (block-twisted radix-2
 * decimation with the bit-reversed root table w, van der Hoeven truncation)
 * but run on canonical residues with exact u128 arithmetic, one radix-2
 * level at a time.  Build: see oracle/Makefile (gcc -O2 -ffp-contract=off).
 */
#include "fsmall_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef int64_t s64;
typedef unsigned __int128 u128;

#define DEFAULT_PRIME UINT64_C(0x0003f00000000001)   /* default_ctx.c:17 */
#define BLK_SZ 256                                  /* fft_small.h:37-38 */
#define LG_BLK_SZ 8
#define NCRTS 8                                     /* fft_small.h:364 */
#define NP2_MAX_AN 40000                            /* fft_small.h:438 */

static int nbits64(u64 x) { return x ? 64 - __builtin_clzll(x) : 0; }
static u64 clog2(u64 x) { return (x <= 2) ? (x == 2) : (u64)(64 - __builtin_clzll(x - 1)); }
static u64 cdiv(u64 a, u64 b) { return (a + b - 1) / b; }
static u64 round_up(u64 a, u64 b) { return cdiv(a, b) * b; }

/* ------------------------------------------------------------------ */
/* exact modular arithmetic                                            */
/* ------------------------------------------------------------------ */

static u64 mulm(u64 a, u64 b, u64 n) { return (u64)(((u128)a * b) % n); }
/* a + b mod n for a, b < n < 2^64 (the sum may wrap past 2^64) */
static u64 addm(u64 a, u64 b, u64 n) { u64 s = a + b; return (s < a || s >= n) ? s - n : s; }
static u64 subm(u64 a, u64 b, u64 n) { return a >= b ? a - b : a + n - b; }

static u64 powm(u64 a, u64 e, u64 n)
{
    u64 r = 1 % n;
    a %= n;
    while (e) {
        if (e & 1) r = mulm(r, a, n);
        a = mulm(a, a, n);
        e >>= 1;
    }
    return r;
}

static u64 invm(u64 a, u64 n)   /* n prime */
{
    return powm(a, n - 2, n);
}

/* deterministic Miller-Rabin for 64-bit n (bases of Sinclair) */
int or_is_prime(u64 n)
{
    static const u64 bases[7] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    if (n < 2) return 0;
    for (u64 s = 2; s < 64 && s * s <= n; s++)
        if (s < 64 && n % s == 0) return n == s;
    u64 d = n - 1; int r = 0;
    while ((d & 1) == 0) { d >>= 1; r++; }
    for (int i = 0; i < 7; i++) {
        u64 a = bases[i] % n;
        if (a == 0) continue;
        u64 x = powm(a, d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int j = 1; j < r; j++) {
            x = mulm(x, x, n);
            if (x == n - 1) { comp = 0; break; }
        }
        if (comp) return 0;
    }
    return 1;
}

/* mpn_mul.c:1000-1016 next_fft_number */
u64 or_next_fft_number(u64 p)
{
    int bits = nbits64(p);
    int l = __builtin_ctzll(p - 1);
    u64 q = p - (UINT64_C(2) << l);
    if (bits < 15) return 0;
    if (nbits64(q) == bits) return q;
    if (l < 5) return (UINT64_C(1) << (bits - 2)) + 1;
    return (UINT64_C(1) << bits) - (UINT64_C(1) << (l - 1)) + 1;
}

/* mpn_ctx_init mpn_mul.c:1085-1091 with DEFAULT_PRIME (default_ctx.c:17) */
void or_default_primes(u64* out)
{
    u64 p = DEFAULT_PRIME;
    for (int i = 0; i < NCRTS; i++) {
        if (i > 0) p = or_next_fft_number(p);
        while (!or_is_prime(p)) p = or_next_fft_number(p);
        out[i] = p;
    }
}

static const u64* default_primes(void)
{
    static u64 P[NCRTS];
    static int init = 0;
    if (!init) { or_default_primes(P); init = 1; }
    return P;
}

/* mulmod_satisfies_bounds.c:120-150 */
int or_mulmod_satisfies_bounds(u64 nn)
{
    double n = (double)nn;
    double ninv = 1.0 / n;
    double t1 = fabs(fma(n, ninv, -1.0));
    u128 sq = (u128)nn * nn;
    u64 hi = (u64)(sq >> 64), lo = (u64)sq;
    int n1bits = nbits64(nn);
    int n2bits = hi ? 64 + nbits64(hi) : nbits64(lo);
    int B = 53 - n1bits - 1;
    if (B < 2) return 0;
    double limit2 = 2 * n * t1 + ldexp(ninv, 1 + n2bits - 53 - 1) + 0.5 + ldexp(1.0, -(B + 1));
    B -= 1;
    double limit4 = 4 * n * t1 + ldexp(ninv, 2 + n2bits - 53 - 1) + 0.5 + ldexp(1.0, -(B + 1));
    return (limit2 < 0.99) && (limit4 < 1.49);
}

/* machine_vectors.h:694-699 vec1d_mulmod */
double or_fp_mulmod(double a, double b, double n, double ninv)
{
    double h = a * b;
    double q = rint(h * ninv);
    double l = fma(a, b, -h);
    return fma(-q, n, h) + l;
}

/* machine_vectors.h:646-651 reduce_to_pm1n */
double or_fp_reduce_to_pm1n(double a, double n, double ninv)
{
    return fma(-rint(a * ninv), n, a);
}

/* machine_vectors.h:635-640,662-667 reduce_to_0n (blend on the sign bit) */
double or_fp_reduce_to_0n(double a, double n, double ninv)
{
    double r = or_fp_reduce_to_pm1n(a, n, ninv);
    return signbit(r) ? r + n : r;
}

static int jacobi(u64 a, u64 n)
{
    int t = 1;
    a %= n;
    while (a) {
        while ((a & 1) == 0) {
            a >>= 1;
            u64 r = n & 7;
            if (r == 3 || r == 5) t = -t;
        }
        u64 tmp = a; a = n; n = tmp;
        if ((a & 3) == 3 && (n & 3) == 3) t = -t;
        a %= n;
    }
    return n == 1 ? t : 0;
}

/* ulong_extras/primitive_root_prime.c:15-26 */
u64 or_quadratic_nonresidue(u64 n)
{
    if ((n & 7) == 3 || (n & 7) == 5) return 2;
    for (u64 a = 3; ; a += 2)
        if (jacobi(a, n) == -1) return a;
}

/* sd_fft_ctx.c:32-37 */
u64 or_primitive_2power_root(u64 p)
{
    int v = __builtin_ctzll(p - 1);
    return powm(or_quadratic_nonresidue(p), (p - 1) >> v, p);
}

static u64 brev(u64 x, int k)
{
    u64 r = 0;
    for (int i = 0; i < k; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

/* fft_small.h:124-199,291-312: w[j] = e(theta(j)), theta(j) = bit-reversed
   binary fraction of j; a primitive 2^v-th root g gives w[j] = g^brev_v(j) */
u64 or_w(u64 p, u64 j)
{
    int v = __builtin_ctzll(p - 1);
    if (nbits64(j) > v) return 0;
    return powm(or_primitive_2power_root(p), brev(j, v), p);
}

u64 or_w2(u64 p, u64 j) { return or_w(p, 2 * j); }

/* fft_small.h:247-253 */
u64 or_trunc_index(u64 L, u64 i)
{
    if (L >= 4)
        i = (i & ~UINT64_C(15)) | ((i >> 2) & 3) | ((i & 3) << 2);
    return i;
}

/* ------------------------------------------------------------------ */
/* transforms                                                          */
/* ------------------------------------------------------------------ */

typedef struct {
    u64 p;
    double pinv;  /* 1/p for the quotient estimate in mulp */
    u64* w2;      /* w2[q] = w[2q], q < 2^(L-1) */
    u64* w2inv;   /* w2[q]^-1 */
    u64 inv2;
} tw_t;

/* sd_fft_ctx.c:56-154 (w2tab layout): w2[2^(k-1) + r] = w2[2^(k-1)] * w2[r] */
static int tw_init(tw_t* T, u64 p, u64 L)
{
    u64 N = (L >= 1) ? (UINT64_C(1) << (L - 1)) : 1;
    int v = __builtin_ctzll(p - 1);
    if ((int)L > v) return 0;
    T->p = p;
    T->pinv = 1.0 / (double)p;
    T->w2 = (u64*)malloc(N * sizeof(u64));
    T->w2inv = (u64*)malloc(N * sizeof(u64));
    T->inv2 = (p + 1) / 2;
    u64 g = or_primitive_2power_root(p);
    T->w2[0] = 1; T->w2inv[0] = 1;
    for (u64 k = 1; (UINT64_C(1) << (k - 1)) < N; k++) {
        u64 base = UINT64_C(1) << (k - 1);
        /* w2[2^(k-1)] = w[2^k] = g^(2^(v-k-1)) */
        u64 root = powm(g, UINT64_C(1) << (v - k - 1), p);
        u64 rinv = invm(root, p);
        for (u64 r = 0; r < base && base + r < N; r++) {
            T->w2[base + r] = mulm(root, T->w2[r], p);
            T->w2inv[base + r] = mulm(rinv, T->w2inv[r], p);
        }
    }
    return 1;
}

static void tw_clear(tw_t* T) { free(T->w2); free(T->w2inv); }

static inline u64 addp(const tw_t* T, u64 a, u64 b) { return addm(a, b, T->p); }
static inline u64 subp(const tw_t* T, u64 a, u64 b) { return subm(a, b, T->p); }

/* a*b mod p for p < 2^51: the double quotient estimate is within one of
   the true quotient, so the u64 difference fits a signed word */
static inline u64 mulp(const tw_t* T, u64 a, u64 b)
{
    u64 p = T->p;
    if (p >= (UINT64_C(1) << 51)) return mulm(a, b, p);
    u64 q = (u64)((double)a * (double)b * T->pinv);
    s64 r = (s64)(a * b - q * p);
    while (r < 0) r += (s64)p;
    while (r >= (s64)p) r -= (s64)p;
    return (u64)r;
}

/* forward: block q of size 2^k holds a mod (X^(2^k) - w[q]);
   level split (lo, hi) -> (lo + w2[q] hi, lo - w2[q] hi)  (sd_fft.c:292-355) */
static void fft_rec(const tw_t* T, u64* x, u64 k, u64 q, u64 itrunc, u64 otrunc)
{
    if (otrunc == 0 || k == 0) return;
    u64 h = UINT64_C(1) << (k - 1);
    if (itrunc == 0) {
        for (u64 i = 0; i < otrunc; i++) x[i] = 0;
        return;
    }
    u64 r = T->w2[q];
    for (u64 i = 0; i < h; i++) {
        u64 lo = (i < itrunc) ? x[i] : 0;
        u64 hi = (i + h < itrunc) ? x[i + h] : 0;
        u64 t = mulp(T, r, hi);
        x[i] = addp(T, lo, t);
        if (otrunc > h) x[i + h] = subp(T, lo, t);
    }
    u64 it = itrunc < h ? itrunc : h;
    fft_rec(T, x, k - 1, 2 * q, it, otrunc < h ? otrunc : h);
    if (otrunc > h) fft_rec(T, x + h, k - 1, 2 * q + 1, it, otrunc - h);
}

/* unnormalized full inverse: outputs 2^k * coefficients */
static void ifft_full(const tw_t* T, u64* x, u64 k, u64 q)
{
    if (k == 0) return;
    u64 h = UINT64_C(1) << (k - 1);
    ifft_full(T, x, k - 1, 2 * q);
    ifft_full(T, x + h, k - 1, 2 * q + 1);
    u64 rinv = T->w2inv[q];
    for (u64 i = 0; i < h; i++) {
        u64 L = x[i], R = x[i + h];
        x[i] = addp(T, L, R);
        x[i + h] = mulp(T, subp(T, L, R), rinv);
    }
}

/*
   van der Hoeven inverse TFT (sd_ifft.c:1418-1493 and the truncated moths
   :79-230).  On entry x[0,n) hold outputs, x[n,z) hold inputs scaled by
   2^k, inputs at positions >= z are zero.  On exit x[0,n) hold the inputs
   scaled by 2^k; with f, x[n] holds the output at position n; every
   other position is left as it was.
*/
static void ifft_rec(const tw_t* T, u64* x, u64 k, u64 q, u64 z, u64 n, int f)
{
    if (k == 0 || (n == 0 && !f)) return;
    u64 N = UINT64_C(1) << k, h = N >> 1;
    if (n == N) { ifft_full(T, x, k, q); return; }
    u64 r = T->w2[q], rinv = T->w2inv[q];
    if (n >= h) {
        u64 m = n - h;
        u64 cnt = h - m;
        u64* save = (u64*)malloc((cnt ? cnt : 1) * sizeof(u64));
        for (u64 i = 0; i < cnt; i++) save[i] = x[h + m + i];
        ifft_full(T, x, k - 1, 2 * q);
        for (u64 i = m; i < h; i++) {
            u64 shi = (i + h < z) ? x[i + h] : 0;
            x[i + h] = subp(T, x[i], mulp(T, r, shi));
        }
        ifft_rec(T, x + h, k - 1, 2 * q + 1, h, m, f);
        for (u64 i = 0; i < m; i++) {
            u64 SL = x[i], SR = x[i + h];
            x[i] = addp(T, SL, SR);
            x[i + h] = mulp(T, subp(T, SL, SR), rinv);
        }
        for (u64 i = m; i < h; i++) {
            u64 shi = (i + h < z) ? save[i - m] : 0;
            x[i] = subp(T, addp(T, x[i], x[i]), mulp(T, r, shi));
            if (!(f && i == m)) x[i + h] = save[i - m];
        }
        free(save);
    } else {
        u64 zl = z < h ? z : h;
        u64 cnt = zl - n;
        u64* save = (u64*)malloc((cnt ? cnt : 1) * sizeof(u64));
        for (u64 i = n; i < zl; i++) {
            save[i - n] = x[i];
            u64 shi = (i + h < z) ? x[i + h] : 0;
            x[i] = mulp(T, addp(T, x[i], mulp(T, r, shi)), T->inv2);
        }
        ifft_rec(T, x, k - 1, 2 * q, zl, n, f);
        for (u64 i = 0; i < n; i++) {
            u64 shi = (i + h < z) ? x[i + h] : 0;
            x[i] = subp(T, addp(T, x[i], x[i]), mulp(T, r, shi));
        }
        for (u64 i = n; i < zl; i++)
            if (!(f && i == n)) x[i] = save[i - n];
        free(save);
    }
}

static void permute_trunc_index(u64* d, u64 L, u64 len)
{
    if (L < 4) return;
    for (u64 b = 0; b + 16 <= len; b += 16) {
        u64 t[16];
        for (u64 i = 0; i < 16; i++) t[or_trunc_index(L, i)] = d[b + i];
        memcpy(d + b, t, sizeof(t));
    }
}

/* sd_fft.c:1230-1274: zero-pads to the block-rounded itrunc, writes the
   first otrunc (block-rounded) outputs in sd_fft_ctx_trunc_index order */
int or_fft_trunc(u64 p, u64* d, u64 L, u64 itrunc, u64 otrunc)
{
    tw_t T;
    u64 N = UINT64_C(1) << L;
    if (itrunc > N || otrunc > N) return 0;
    if (!tw_init(&T, p, L)) return 0;
    u64 it, ot;
    if (L > LG_BLK_SZ) {
        it = round_up(itrunc, BLK_SZ);
        ot = round_up(otrunc, BLK_SZ);
    } else {
        it = itrunc;
        ot = N;
    }
    for (u64 i = itrunc; i < (L > LG_BLK_SZ ? it : N); i++) d[i] = 0;
    for (u64 i = 0; i < it; i++) d[i] %= p;
    fft_rec(&T, d, L, 0, it, ot);
    permute_trunc_index(d, L, ot);
    tw_clear(&T);
    return 1;
}

/* sd_ifft.c:1503-1533 */
int or_ifft_trunc(u64 p, u64* d, u64 L, u64 trunc)
{
    tw_t T;
    u64 N = UINT64_C(1) << L;
    if (trunc > N) return 0;
    if (!tw_init(&T, p, L)) return 0;
    u64 t = (L > LG_BLK_SZ) ? round_up(trunc, BLK_SZ) : N;
    permute_trunc_index(d, L, t);   /* trunc_index is an involution */
    for (u64 i = 0; i < t; i++) d[i] %= p;
    if (t > 0) ifft_rec(&T, d, L, 0, t, t, 0);
    tw_clear(&T);
    return 1;
}

/* plain bit-reversed-order transforms used by the product drivers */
static void conv_mod_p(u64 p, u64* A, u64 at, u64* B, u64 bt, u64 depth, u64 ztrunc,
                       u64 mult)
{
    tw_t T;
    tw_init(&T, p, depth);
    fft_rec(&T, A, depth, 0, at, ztrunc);
    fft_rec(&T, B, depth, 0, bt, ztrunc);
    for (u64 i = 0; i < ztrunc; i++)
        A[i] = mulp(&T, mulp(&T, A[i], B[i]), mult);
    ifft_rec(&T, A, depth, 0, ztrunc, ztrunc, 0);
    tw_clear(&T);
}

/* ------------------------------------------------------------------ */
/* small multi-limb helpers                                            */
/* ------------------------------------------------------------------ */

#define ML 10   /* limbs for CRT quantities (<= 400 bits + headroom) */

static void ml_zero(u64* x) { memset(x, 0, ML * sizeof(u64)); }
static void ml_mul_1(u64* r, const u64* a, u64 b)
{
    u64 c = 0;
    for (int i = 0; i < ML; i++) {
        u128 t = (u128)a[i] * b + c;
        r[i] = (u64)t; c = (u64)(t >> 64);
    }
}
static void ml_addmul_1(u64* r, const u64* a, u64 b)
{
    u64 c = 0;
    for (int i = 0; i < ML; i++) {
        u128 t = (u128)a[i] * b + r[i] + c;
        r[i] = (u64)t; c = (u64)(t >> 64);
    }
}
static int ml_cmp(const u64* a, const u64* b)
{
    for (int i = ML - 1; i >= 0; i--)
        if (a[i] != b[i]) return a[i] > b[i] ? 1 : -1;
    return 0;
}
static void ml_sub(u64* r, const u64* a, const u64* b)
{
    u64 br = 0;
    for (int i = 0; i < ML; i++) {
        u64 t = a[i] - b[i] - br;
        br = (a[i] < b[i]) || (a[i] == b[i] && br);
        r[i] = t;
    }
}
static u64 ml_divexact_1(u64* r, const u64* a, u64 d)
{
    u128 rem = 0;
    for (int i = ML - 1; i >= 0; i--) {
        u128 cur = (rem << 64) | a[i];
        r[i] = (u64)(cur / d);
        rem = cur % d;
    }
    return (u64)rem;
}
static u64 ml_mod_1(const u64* a, u64 d)
{
    u128 rem = 0;
    for (int i = ML - 1; i >= 0; i--) rem = ((rem << 64) | a[i]) % d;
    return (u64)rem;
}
static int ml_nbits(const u64* a)
{
    for (int i = ML - 1; i >= 0; i--)
        if (a[i]) return 64 * i + nbits64(a[i]);
    return 0;
}

/* M = prod_{i<np} p_{off+i}; cof_i = M/p_i; cop_i = cof_i mod p_i
   (mpn_mul.c:1093-1137) */
typedef struct {
    int np;
    u64 p[NCRTS];
    u64 M[ML];
    u64 cof[NCRTS][ML];
    u64 cop[NCRTS];
} crt_t;

static void crt_init(crt_t* C, const u64* primes, int np)
{
    C->np = np;
    ml_zero(C->M);
    C->M[0] = 1;
    for (int i = 0; i < np; i++) {
        C->p[i] = primes[i];
        ml_mul_1(C->M, C->M, primes[i]);
    }
    for (int i = 0; i < np; i++) {
        ml_divexact_1(C->cof[i], C->M, primes[i]);
        C->cop[i] = ml_mod_1(C->cof[i], primes[i]);
    }
}

/* M >= c * 2^e ? (flint_mpn_cmp_ui_2exp as used by plan.c:28-32) */
static int ml_ge_ui_2exp(const u64* M, u64 c, u64 e)
{
    u64 t[ML + 4];
    memset(t, 0, sizeof(t));
    u64 q = e / 64, r = e % 64;
    if (q + 2 >= ML + 4) return 0;
    t[q] = c << r;
    if (r) t[q + 1] = c >> (64 - r);
    for (int i = ML + 3; i >= 0; i--) {
        u64 mi = (i < ML) ? M[i] : 0;
        if (mi != t[i]) return mi > t[i];
    }
    return 1;
}

/* mpn_mul.c:152-184 crt_data_find_bn_bound */
u64 or_crt_bn_bound(int np, u64 bits)
{
    crt_t C;
    crt_init(&C, default_primes(), np);
    u64 x[ML + 1];
    u64 c = 0;
    for (int i = 0; i < ML; i++) {
        u128 t = (u128)C.M[i] * bits + c;
        x[i] = (u64)t; c = (u64)(t >> 64);
    }
    x[ML] = c;
    /* y = (M*bits) >> 2*bits */
    u64 sh = 2 * bits, q = sh / 64, r = sh % 64;
    u64 y[ML + 1];
    memset(y, 0, sizeof(y));
    for (u64 i = q; i <= ML; i++) {
        u64 lo = x[i] >> r;
        u64 hi = (r && i + 1 <= ML) ? (x[i + 1] << (64 - r)) : 0;
        y[i - q] = lo | hi;
    }
    /* y -= bits - 1, borrow => 0 */
    u64 sub = bits - 1, br = 0;
    for (int i = 0; i <= ML; i++) {
        u64 s = (i == 0) ? sub : 0;
        u64 t = y[i] - s - br;
        br = (y[i] < s) || (y[i] == s && br);
        y[i] = t;
    }
    if (br) return 0;
    /* >> 6 */
    for (int i = 0; i <= ML; i++)
        y[i] = (y[i] >> 6) | ((i < ML) ? (y[i + 1] << 58) : 0);
    for (int i = 1; i <= ML; i++)
        if (y[i]) return ~UINT64_C(0);
    return y[0];
}

/* mpn_mul.c:191-212 */
u64 or_crt_find_bits(int np, u64 bn)
{
    crt_t C;
    crt_init(&C, default_primes(), np);
    u64 pbits = (u64)ml_nbits(C.M);
    u64 b = (pbits - nbits64(bn)) / 2;
    u64 bits = b > 66 ? b : 66;
    if (bn > or_crt_bn_bound(np, bits)) {
        do { bits -= 1; } while (bits > 65 && bn > or_crt_bn_bound(np, bits));
    } else {
        while (bn <= or_crt_bn_bound(np, bits + 1)) bits += 1;
    }
    return bits;
}

/* the 15 profiles of mpn_mul.c:1171-1194 */
static const u64 prof_np[15]   = {4, 4, 4, 5, 5, 5, 6, 6, 6, 7, 7, 7, 8, 8, 8};
static const u64 prof_bits[15] = {84, 88, 92, 112, 116, 120, 136, 140, 144,
                                  160, 164, 168, 184, 188, 192};

/* mpn_mul.c:1248-1340 mpn_ctx_best_profile; nthreads = FLINT's thread count */
int or_best_profile(u64 an, u64 bn, u64 nthreads, u64* np_out, u64* bits_out)
{
    u64 bound[15];
    for (int i = 0; i < 15; i++) bound[i] = or_crt_bn_bound((int)prof_np[i], prof_bits[i]);
    u64 zn = an + bn, limit = 8;
    if (zn < 2048) limit = 1;
    else if (zn < 4096) limit = 4;
    else if (zn < 8192) limit = 5;
    else if (zn < 16384) limit = 6;
    else if (zn < 32768) limit = 7;
    u64 nt = nthreads < 1 ? 1 : nthreads;
    if (nt > limit) nt = limit;

    if (bn > bound[0]) {
        u64 np = nt < 8 ? nt : 8;
        if (np < 4) np = 4;
        *np_out = np;
        *bits_out = or_crt_find_bits((int)np, bn);
        return 0;   /* generic (slow) conversion path */
    }

    int i = 0, best_i = 0;
    double best_score = 100000000.0 * (double)(an + bn);
    for (;;) {
        while (i + 1 < 15 && bn <= bound[i + 1] && prof_np[i + 1] == prof_np[i]) i++;
        u64 np = prof_np[i];
        if (np % nt == 0) {
            u64 bits = prof_bits[i];
            u64 alen = cdiv(64 * an, bits), blen = cdiv(64 * bn, bits);
            u64 zlen = alen + blen - 1;
            u64 ztrunc = round_up(zlen, BLK_SZ);
            u64 depth = clog2(ztrunc) > LG_BLK_SZ ? clog2(ztrunc) : LG_BLK_SZ;
            double ratio = (double)ztrunc / (double)(UINT64_C(1) << depth);
            double score = (1 - 0.25 * ratio) * (1.0 / 1000000);
            score *= (double)(np * depth);
            score *= (double)ztrunc;
            if (score < best_score) { best_i = i; best_score = score; }
        }
        do {
            i++;
            if (i >= 15) {
                *np_out = prof_np[best_i];
                *bits_out = prof_bits[best_i];
                return 1;
            }
        } while (bn > bound[i]);
    }
}

/* mpn_mul_np2.c:46-70 */
int or_np2_choose(u64 an, u64 bn, u64 len_bound, u64* bits, u64* depth)
{
    const u64* P = default_primes();
    u128 pp = (u128)P[0] * P[1];
    u64 prodbits = 64 + nbits64((u64)(pp >> 64)) - 1;
    for (u64 d = LG_BLK_SZ; d <= 30; d++) {
        for (u64 b = 45; b >= 40; b--) {
            u64 alen = (64 * an + b - 1) / b, blen = (64 * bn + b - 1) / b;
            if (alen + blen - 1 > (UINT64_C(1) << d)) continue;
            u128 t = (u128)len_bound * blen;
            if ((u64)(t >> 64) != 0 || 2 * b + nbits64((u64)t) > prodbits) continue;
            *bits = b; *depth = d;
            return 1;
        }
    }
    return 0;
}

/* full-product plan of _mpn_ctx_mpn_mul_range (mpn_mul.c:1617-1721,
   mpn_mul_np2.c:247-265): out = {np, bits, depth, ztrunc, alen, blen} */
int or_mpn_plan(u64 an, u64 bn, u64 nthreads, u64* out)
{
    u64 np, bits, depth;
    if ((an > bn ? an : bn) <= NP2_MAX_AN) {
        if (an < bn) { u64 t = an; an = bn; bn = t; }
        if (!or_np2_choose(an, bn, 1, &bits, &depth)) return 0;
        np = 2;
    } else {
        or_best_profile(an, bn, nthreads, &np, &bits);
        depth = 0;
    }
    u64 alen = cdiv(64 * an, bits), blen = cdiv(64 * bn, bits);
    u64 zlen = alen + blen - 1;
    u64 ztrunc = round_up(zlen, BLK_SZ);
    if (np != 2) {
        depth = clog2(ztrunc);
        if (depth < LG_BLK_SZ) depth = LG_BLK_SZ;
    }
    out[0] = np; out[1] = bits; out[2] = depth; out[3] = ztrunc;
    out[4] = alen; out[5] = blen;
    return 1;
}

/* mpn_mul.c:422-481 slow_mpn_to_fft (the closed form the fast passes
   :595-695 are tested against): chunk i = bits [i*bits, (i+1)*bits) of a,
   reduced mod p; entries i >= ceil(64 an / bits) are zero */
void or_mpn_to_residues(u64 p, u64* d, u64 dn, const u64* a, u64 an, u64 bits)
{
    for (u64 i = 0; i < dn; i++) {
        u64 pos = i * bits;
        u64 r = 0;
        /* Horner over 32-bit pieces from the top down */
        u64 npieces = cdiv(bits, 32);
        for (s64 k = (s64)npieces - 1; k >= 0; k--) {
            u64 bp = pos + 32 * (u64)k;
            u64 take = bits - 32 * (u64)k;
            if (take > 32) take = 32;
            u64 w = bp / 64, s = bp % 64;
            u64 v = (w < an) ? (a[w] >> s) : 0;
            if (s > 32 && w + 1 < an) v |= a[w + 1] << (64 - s);
            v &= (take == 64) ? ~UINT64_C(0) : ((UINT64_C(1) << take) - 1);
            r = addm(mulm(r, (UINT64_C(1) << 32) % p, p), v % p, p);
        }
        d[i] = r;
    }
}

/* add the n-limb value s * 2^shift into z[0, zn) */
static void add_shifted(u64* z, u64 zn, const u64* s, int sn, u64 shift)
{
    u64 off = shift / 64, sh = shift % 64;
    u64 carry = 0;
    for (u64 i = 0; off + i < zn; i++) {
        u64 lo = (i < (u64)sn) ? s[i] : 0;
        u64 prev = (i >= 1 && i - 1 < (u64)sn) ? s[i - 1] : 0;
        u64 v = sh ? ((lo << sh) | (prev >> (64 - sh))) : lo;
        if (i > (u64)sn && carry == 0) break;
        u128 t = (u128)z[off + i] + v + carry;
        z[off + i] = (u64)t;
        carry = (u64)(t >> 64);
    }
}

/* residues of the coefficient stream -> integer coefficient (crt_helpers.h
   _big_mul/_big_addmul/_reduce_big_sum, mpn_mul.c:883-906): with
   x_l = c * cop_l^-1 mod p_l, sum_l x_l * (M/p_l) < np*M reduces to c mod M */
static void crt_combine(const crt_t* C, u64* const* res, u64 j, u64* s)
{
    ml_zero(s);
    for (int l = 0; l < C->np; l++) ml_addmul_1(s, C->cof[l], res[l][j]);
    while (ml_cmp(s, C->M) >= 0) ml_sub(s, s, C->M);
}

/* the coefficient stage of _mpn_ctx_mpn_mul_range (mpn_mul.c:1657-1858)
   and of the two-prime window (mpn_mul_np2.c:247-310): residues of the
   chunk convolution for each prime of the reference's plan, ready for
   crt_combine.  Returns the number of primes (0: no plan). */
typedef struct {
    int np;
    u64 bits, zlen;
    crt_t C;
    u64* res[NCRTS];
} coeffs_t;

static int mpn_coeffs(coeffs_t* K, const u64* a, u64 an, const u64* b, u64 bn, u64 nthreads)
{
    u64 plan[6];
    if (an == 0 || bn == 0) return 0;
    if (!or_mpn_plan(an, bn, nthreads, plan)) return 0;
    int np = (int)plan[0];
    u64 bits = plan[1], depth = plan[2], ztrunc = plan[3];
    u64 alen = cdiv(64 * an, bits), blen = cdiv(64 * bn, bits);
    u64 N = UINT64_C(1) << depth;
    const u64* P = default_primes();
    K->np = np;
    K->bits = bits;
    K->zlen = alen + blen - 1;
    crt_init(&K->C, P, np);
    u64* B = (u64*)malloc(N * sizeof(u64));
    for (int l = 0; l < np; l++) {
        u64 p = P[l];
        K->res[l] = (u64*)malloc(N * sizeof(u64));
        or_mpn_to_residues(p, K->res[l], N, a, an, bits);
        or_mpn_to_residues(p, B, N, b, bn, bits);
        /* normalizer m = (cop * 2^depth)^-1 (mpn_mul.c:1517-1518) */
        u64 m = invm(mulm(K->C.cop[l], powm(2, depth, p), p), p);
        conv_mod_p(p, K->res[l], round_up(alen, BLK_SZ), B, round_up(blen, BLK_SZ),
                   depth, ztrunc, m);
    }
    free(B);
    return np;
}

static void coeffs_clear(coeffs_t* K)
{
    for (int l = 0; l < K->np; l++) free(K->res[l]);
}

/* _mpn_ctx_mpn_mul_range with lo = 0, hi = an + bn (mpn_mul.c:1617-1978)
   and the two-prime window (mpn_mul_np2.c:247-310) */
int or_mpn_mul(u64* z, const u64* a, u64 an, const u64* b, u64 bn, u64 nthreads)
{
    coeffs_t K;
    if (!mpn_coeffs(&K, a, an, b, bn, nthreads)) return 0;
    u64 zn = an + bn;
    memset(z, 0, zn * sizeof(u64));
    u64 s[ML];
    for (u64 j = 0; j < K.zlen; j++) {
        crt_combine(&K.C, K.res, j, s);
        add_shifted(z, zn, s, ML, j * K.bits);
    }
    coeffs_clear(&K);
    return 1;
}

/* limbs of M * np, the reference's coeff_len (mpn_mul.c:1105-1117) */
static u64 coeff_len(const crt_t* C)
{
    u64 t[ML];
    ml_mul_1(t, C->M, (u64)C->np);
    u64 n = ML;
    while (n > 1 && t[n - 1] == 0) n--;
    return n;
}

/* s >> sh for an ML-limb s */
static void ml_rshift(u64* r, const u64* s, u64 sh)
{
    u64 q = sh / 64, b = sh % 64;
    for (u64 i = 0; i < ML; i++) {
        u64 lo = (i + q < ML) ? s[i + q] : 0;
        u64 hi = (i + q + 1 < ML) ? s[i + q + 1] : 0;
        r[i] = b ? ((lo >> b) | (hi << (64 - b))) : lo;
    }
}

/* _mpn_ctx_mpn_mul_range(R, z, lo, hi, a, an, b, bn) with FLINT running
   nthreads threads: z[0, hi - lo) = the reference's limbs [lo, hi).
   - two-prime plans (max(an, bn) <= 40000, mpn_mul_np2.c:247-433): every
     chunk coefficient c_j at bit j*bits contributes floor(c_j 2^(j bits) /
     2^(64 lo)); the low bits of the chunks straddling bit 64 lo are
     masked off one chunk at a time (mpn_mul_np2.c:380-402);
   - wider plans (mpn_mul.c:1669-1978): the coefficients j >= c_lo,
     c_lo = ceil((lo - n) 64 / bits) for lo > n (n = coeff_len, else 0),
     are summed exactly (the bottom band into a buffer anchored at limb
     floor(c_lo bits / 64), the rest in place) and limbs [lo, hi) of that
     sum are returned: carries of the coefficients below c_lo are lost;
   - limbs at or above an + bn are zero (mpn_mul.c:1636-1645). */
int or_mpn_mul_range(u64* z, u64 lo, u64 hi, const u64* a, u64 an, const u64* b, u64 bn,
                     u64 nthreads)
{
    if (an == 0 || bn == 0) return 0;
    if (lo >= hi) return 1;
    u64 zn = an + bn;
    memset(z, 0, (hi - lo) * sizeof(u64));
    if (lo >= zn) return 1;
    if (hi > zn) hi = zn;
    coeffs_t K;
    if (!mpn_coeffs(&K, a, an, b, bn, nthreads)) return 0;
    const u64 w = hi - lo, bits = K.bits;
    u64 s[ML], v[ML];
    if (K.np == 2) {
        for (u64 j = 0; j < K.zlen; j++) {
            u64 pos = j * bits;
            if (pos >= 64 * hi) break;
            crt_combine(&K.C, K.res, j, s);
            if (pos < 64 * lo) {
                ml_rshift(v, s, 64 * lo - pos);
                add_shifted(z, w, v, ML, 0);
            } else {
                add_shifted(z, w, s, ML, pos - 64 * lo);
            }
        }
    } else {
        u64 nn = coeff_len(&K.C);
        u64 c_lo = 0;
        if (lo > nn) {
            c_lo = cdiv((lo - nn) * 64, bits);
            if (c_lo > K.zlen) c_lo = K.zlen;
        }
        u64 c_hi = cdiv(64 * hi, bits);
        if (c_hi > K.zlen) c_hi = K.zlen;
        if (c_lo < c_hi) {
            u64 L0 = c_lo * bits / 64, tn = hi - L0;
            u64* t = (u64*)calloc(tn, sizeof(u64));
            for (u64 j = c_lo; j < c_hi; j++) {
                crt_combine(&K.C, K.res, j, s);
                add_shifted(t, tn, s, ML, j * bits - 64 * L0);
            }
            memcpy(z, t + (lo - L0), w * sizeof(u64));
            free(t);
        }
    }
    coeffs_clear(&K);
    return 1;
}

void or_mpn_mul_basecase(u64* z, const u64* a, u64 an, const u64* b, u64 bn)
{
    memset(z, 0, (an + bn) * sizeof(u64));
    for (u64 i = 0; i < an; i++) {
        u64 c = 0;
        for (u64 j = 0; j < bn; j++) {
            u128 t = (u128)a[i] * b[j] + z[i + j] + c;
            z[i + j] = (u64)t; c = (u64)(t >> 64);
        }
        z[i + bn] = c;
    }
}

/* convolution of residue vectors over np primes followed by CRT: writes the
   exact integer coefficients j in [0, zn) as ML-limb values into out */
static void conv_crt(const crt_t* C, u64* const* ares, u64 an, u64* const* bres, u64 bn,
                     u64 zn, u64* out)
{
    u64 depth = clog2(zn);
    if (depth < 4) depth = 4;
    u64 N = UINT64_C(1) << depth;
    u64* res[NCRTS];
    u64* B = (u64*)malloc(N * sizeof(u64));
    for (int l = 0; l < C->np; l++) {
        u64 p = C->p[l];
        res[l] = (u64*)calloc(N, sizeof(u64));
        memset(B, 0, N * sizeof(u64));
        memcpy(res[l], ares[l], an * sizeof(u64));
        memcpy(B, bres[l], bn * sizeof(u64));
        u64 m = (C->np == 1) ? invm(powm(2, depth, p), p)
                             : invm(mulm(C->cop[l], powm(2, depth, p), p), p);
        conv_mod_p(p, res[l], an, B, bn, depth, N, m);
    }
    free(B);
    for (u64 j = 0; j < zn; j++) {
        if (C->np == 1) { ml_zero(out + j * ML); out[j * ML] = res[0][j]; }
        else crt_combine(C, res, j, out + j * ML);
    }
    for (int l = 0; l < C->np; l++) free(res[l]);
}

/* nmod_poly_mul.c:904-972 with the np choice of plan.c:356-398,18-45 */
int or_nmod_poly_mul_mid(u64* z, s64 zl, s64 zh, const u64* a, s64 an,
                         const u64* b, s64 bn, or_nmod_t mod)
{
    if (an <= 0 || bn <= 0) return 0;
    if (zl >= zh) return 1;
    u64 n = mod.n;
    s64 zn = an + bn - 1;
    if (zh > zn) {
        s64 start = zl > zn ? zl : zn;
        for (s64 i = start; i < zh; i++) z[i - zl] = 0;
        if (zl >= zn) return 1;
        zh = zn;
    }
    const u64* P = default_primes();
    int modbits = nbits64(n);
    crt_t C;
    int found = 0;
    if (modbits == 50)
        for (int i = 0; i < NCRTS; i++)
            if (P[i] == n) { crt_init(&C, P + i, 1); found = 1; }
    if (!found) {
        int np;
        for (np = 1; np <= NCRTS; np++) {
            crt_t T;
            crt_init(&T, P, np);
            if (ml_ge_ui_2exp(T.M, (u64)bn, 2 * (u64)modbits)) break;
        }
        if (np > NCRTS) return 0;
        crt_init(&C, P, np);
    }
    u64* ar[NCRTS]; u64* br[NCRTS];
    for (int l = 0; l < C.np; l++) {
        ar[l] = (u64*)malloc(an * sizeof(u64));
        br[l] = (u64*)malloc(bn * sizeof(u64));
        for (s64 i = 0; i < an; i++) ar[l][i] = a[i] % C.p[l];
        for (s64 i = 0; i < bn; i++) br[l][i] = b[i] % C.p[l];
    }
    u64* co = (u64*)malloc((size_t)zn * ML * sizeof(u64));
    conv_crt(&C, ar, an, br, bn, zn, co);
    for (s64 i = zl; i < zh; i++) z[i - zl] = ml_mod_1(co + i * ML, n);
    free(co);
    for (int l = 0; l < C.np; l++) { free(ar[l]); free(br[l]); }
    return 1;
}

/* signed w-limb two's complement coefficient helpers */
static int coeff_neg(const u64* c, u64 w) { return (s64)c[w - 1] < 0; }

static u64 coeff_abs_bits(const u64* c, u64 w)
{
    u64 t[w];
    if (coeff_neg(c, w)) {
        u64 carry = 1;
        for (u64 i = 0; i < w; i++) { t[i] = ~c[i] + carry; carry = carry && t[i] == 0; }
    } else memcpy(t, c, w * sizeof(u64));
    for (s64 i = (s64)w - 1; i >= 0; i--)
        if (t[i]) return 64 * (u64)i + nbits64(t[i]);
    return 0;
}

static u64 coeff_mod(const u64* c, u64 w, u64 p)
{
    int neg = coeff_neg(c, w);
    u64 t[w];
    if (neg) {
        u64 carry = 1;
        for (u64 i = 0; i < w; i++) { t[i] = ~c[i] + carry; carry = carry && t[i] == 0; }
    } else memcpy(t, c, w * sizeof(u64));
    u128 rem = 0;
    for (s64 i = (s64)w - 1; i >= 0; i--) rem = ((rem << 64) | t[i]) % p;
    u64 r = (u64)rem;
    return neg ? (r ? p - r : 0) : r;
}

/* _fmpz_vec_max_bits semantics: max |bits|, negative if any entry < 0 */
static s64 vec_max_bits(const u64* a, u64 w, s64 n)
{
    u64 mb = 0; int neg = 0;
    for (s64 i = 0; i < n; i++) {
        u64 bb = coeff_abs_bits(a + i * w, w);
        if (bb > mb) mb = bb;
        if (coeff_neg(a + i * w, w)) neg = 1;
    }
    return neg ? -(s64)mb : (s64)mb;
}

/* prime count chosen by fmpz_poly_mul.c:354-379 (0 = declined) */
int or_fmpz_poly_nprimes(const u64* a, u64 aw, s64 an, const u64* b, u64 bw, s64 bn)
{
    s64 b1 = vec_max_bits(a, aw, an), b2 = vec_max_bits(b, bw, bn);
    u64 modbits = (u64)llabs(b1) + (u64)llabs(b2) + 1;
    const u64* P = default_primes();
    for (int np = 1; np <= NCRTS; np++) {
        crt_t T;
        crt_init(&T, P, np);
        if (ml_ge_ui_2exp(T.M, (u64)bn, modbits)) return np;
    }
    return 0;
}

/* fmpz_poly_mul.c:319-401, coefficients as w-limb two's complement */
int or_fmpz_poly_mul_mid(u64* z, u64 zw, s64 zl, s64 zh, const u64* a, u64 aw, s64 an,
                         const u64* b, u64 bw, s64 bn)
{
    if (an <= 0 || bn <= 0) return 0;
    if (zl >= zh) return 1;
    s64 zn = an + bn - 1;
    if (zh > zn) {
        s64 start = zl > zn ? zl : zn;
        for (s64 i = start; i < zh; i++) memset(z + (i - zl) * zw, 0, zw * sizeof(u64));
        if (zl >= zn) return 1;
        zh = zn;
    }
    int np = or_fmpz_poly_nprimes(a, aw, an, b, bw, bn);
    if (np == 0) return 0;
    crt_t C;
    crt_init(&C, default_primes(), np);
    u64* ar[NCRTS]; u64* br[NCRTS];
    for (int l = 0; l < np; l++) {
        ar[l] = (u64*)malloc(an * sizeof(u64));
        br[l] = (u64*)malloc(bn * sizeof(u64));
        for (s64 i = 0; i < an; i++) ar[l][i] = coeff_mod(a + i * aw, aw, C.p[l]);
        for (s64 i = 0; i < bn; i++) br[l][i] = coeff_mod(b + i * bw, bw, C.p[l]);
    }
    u64* co = (u64*)malloc((size_t)zn * ML * sizeof(u64));
    conv_crt(&C, ar, an, br, bn, zn, co);
    u64 Mhalf[ML];
    for (int i = 0; i < ML; i++) Mhalf[i] = (C.M[i] >> 1) | ((i + 1 < ML) ? (C.M[i + 1] << 63) : 0);
    for (s64 i = zl; i < zh; i++) {
        u64* s = co + i * ML;
        u64* out = z + (i - zl) * zw;
        if (ml_cmp(s, Mhalf) > 0) {
            /* CRT_EMIT fmpz_poly_mul.c:209-227: value - M, negative */
            u64 t[ML];
            ml_sub(t, s, C.M);     /* two's complement of M - s */
            for (u64 k = 0; k < zw; k++) out[k] = (k < ML) ? t[k] : ~UINT64_C(0);
        } else {
            for (u64 k = 0; k < zw; k++) out[k] = (k < ML) ? s[k] : 0;
        }
    }
    free(co);
    for (int l = 0; l < np; l++) { free(ar[l]); free(br[l]); }
    return 1;
}

/* ------------------------------------------------------------------ */
/* seeded synthetic inputs                                             */
/* ------------------------------------------------------------------ */

static u64 splitmix64(u64* x)
{
    u64 z = (*x += UINT64_C(0x9E3779B97F4A7C15));
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}

void or_rng_seed(or_rng* r, u64 seed)
{
    u64 x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static u64 rotl(u64 x, int k) { return (x << k) | (x >> (64 - k)); }

u64 or_rng_next(or_rng* r)
{
    u64* s = r->s;
    u64 result = rotl(s[1] * 5, 7) * 9;
    u64 t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return result;
}

u64 or_rng_uniform(or_rng* r, u64 lo, u64 hi)
{
    u64 range = hi - lo + 1;
    if (range == 0) return or_rng_next(r);
    u64 lim = ~UINT64_C(0) - (~UINT64_C(0) % range + 1) % range;
    for (;;) {
        u64 x = or_rng_next(r);
        if (x <= lim) return lo + x % range;
    }
}

u64 or_rng_random_prime(or_rng* r, u64 bits)
{
    if (bits < 2) bits = 2;
    u64 lo = UINT64_C(1) << (bits - 1);
    u64 hi = (bits == 64) ? ~UINT64_C(0) : (UINT64_C(1) << bits) - 1;
    for (;;) {
        u64 x = or_rng_uniform(r, lo, hi) | 1;
        if (bits == 2) x = or_rng_uniform(r, 2, 3);
        if (or_is_prime(x)) return x;
    }
}

void or_rng_limbs(or_rng* r, u64* a, u64 n, int top_nonzero)
{
    for (u64 i = 0; i < n; i++) a[i] = or_rng_next(r);
    if (top_nonzero && n > 0)
        while (a[n - 1] == 0) a[n - 1] = or_rng_next(r);
}

/* ------------------------------------------------------------------ */
/* request generator of the batched workloads, restated independently  */
/* of the library's (tests compare the two draw for draw)              */
/* ------------------------------------------------------------------ */

double or_rng_double(or_rng* r)
{
    return ldexp((double)(or_rng_next(r) >> 11), -53);
}

u64 or_rng_log_uniform(or_rng* r, double x, double y)
{
    double d = or_rng_double(r);
    return (u64)exp2(x + (y - x) * d);
}

void or_rng_fill_below(or_rng* r, u64* out, u64 n, u64 bound)
{
    for (u64 i = 0; i < n; i++) out[i] = bound ? or_rng_uniform(r, 0, bound - 1) : 0;
}

/* t uniform over the (bits+1)-bit values except the largest, c = t + 1 - 2^bits:
   add one, then subtract 2^bits by complementing from bit `bits` upwards
   with the borrow */
void or_rng_fill_signed(or_rng* r, u64* out, u64 n, int bits, u64 w)
{
    u64 tw = (u64)(bits + 64) / 64;
    int topb = bits + 1 - 64 * (int)(tw - 1);
    u64 mask = (topb == 64) ? ~(u64)0 : (((u64)1 << topb) - 1);
    for (u64 i = 0; i < n; i++) {
        u64* c = out + i * w;
        int again;
        do {
            again = 1;
            for (u64 k = 0; k < tw; k++) {
                c[k] = or_rng_next(r);
                if (k + 1 == tw) c[k] &= mask;
                if (c[k] != ((k + 1 == tw) ? mask : ~(u64)0)) again = 0;
            }
        } while (again);
        for (u64 k = tw; k < w; k++) c[k] = 0;
        /* t + 1 */
        for (u64 k = 0; k < w; k++) if (++c[k] != 0) break;
        /* - 2^bits */
        u64 k0 = (u64)bits / 64;
        u64 sub = (u64)1 << (bits % 64);
        for (u64 k = k0; k < w; k++) {
            u64 x = c[k];
            c[k] = x - sub;
            sub = (x < sub) ? 1 : 0;
            if (!sub) break;
        }
    }
}

/* ------------------------------------------------------------------ */
/* the full mixed stream: headers from the xoshiro256** stream, operand */
/* contents from SplitMix64 sequences keyed per operand                 */
/* ------------------------------------------------------------------ */

/* word j of the SplitMix64 sequence seeded with key: the state after j + 1
   increments, through the output function */
u64 or_stream_word(u64 key, u64 j)
{
    u64 x = key + j * UINT64_C(0x9E3779B97F4A7C15);
    return splitmix64(&x);
}

/* count headers, 7 words each: kind (0 integer, 1 nmod, 2 fmpz), bits, an,
   bn, n, key0, key1 */
void or_stream_headers(u64 seed, u64 count, u64* out)
{
    or_rng r;
    or_rng_seed(&r, seed);
    for (u64 k = 0; k < count; k++) {
        u64* h = out + 7 * k;
        memset(h, 0, 7 * sizeof(u64));
        double u = or_rng_double(&r);
        if (u < 0.5) {
            h[0] = 0;
            h[2] = or_rng_log_uniform(&r, 4, 22);
            h[3] = or_rng_log_uniform(&r, 4, 22);
        } else if (u < 0.8) {
            h[0] = 1;
            h[2] = or_rng_log_uniform(&r, 6, 16);
            h[3] = or_rng_log_uniform(&r, 6, 16);
            h[1] = or_rng_uniform(&r, 20, 63);
            h[4] = or_rng_random_prime(&r, h[1]);
        } else {
            h[0] = 2;
            h[2] = or_rng_uniform(&r, 1024, 16384);
            h[3] = or_rng_uniform(&r, 1024, 16384);
            h[1] = or_rng_uniform(&r, 64, 4096);
        }
        h[5] = or_rng_next(&r);
        h[6] = or_rng_next(&r);
    }
}

/* one operand of the stream: count limbs (kind 0: top limb 1 when its word
   is 0; kind 1: word * n / 2^64) or count coefficients of w limbs (kind 2:
   t of the words j tw .. j tw + tw - 1 cut to bits + 1 bits, the all-ones t
   replaced by 0, then t + 1 - 2^bits) */
void or_stream_fill(int kind, u64 key, u64 count, u64 n, u64 w, int bits, u64* out)
{
    if (kind == 0) {
        for (u64 j = 0; j < count; j++) out[j] = or_stream_word(key, j);
        if (count && out[count - 1] == 0) out[count - 1] = 1;
        return;
    }
    if (kind == 1) {
        for (u64 j = 0; j < count; j++) out[j] = (u64)(((u128)or_stream_word(key, j) * n) >> 64);
        return;
    }
    u64 tw = (u64)(bits + 64) / 64;
    int topb = bits + 1 - 64 * (int)(tw - 1);
    u64 mask = (topb == 64) ? ~(u64)0 : (((u64)1 << topb) - 1);
    for (u64 j = 0; j < count; j++) {
        u64* c = out + j * w;
        int ones = 1;
        for (u64 k = 0; k < w; k++) c[k] = 0;
        for (u64 k = 0; k < tw; k++) {
            c[k] = or_stream_word(key, j * tw + k);
            if (k + 1 == tw) c[k] &= mask;
            if (c[k] != ((k + 1 == tw) ? mask : ~(u64)0)) ones = 0;
        }
        if (ones)
            for (u64 k = 0; k < tw; k++) c[k] = 0;
        for (u64 k = 0; k < w; k++) if (++c[k] != 0) break;
        u64 sub = (u64)1 << (bits % 64);
        for (u64 k = (u64)bits / 64; k < w && sub; k++) {
            u64 x = c[k];
            c[k] = x - sub;
            sub = (x < sub) ? 1 : 0;
        }
    }
}

/* ------------------------------------------------------------------ */
/* checkers at full size: residues mod a word-size q                   */
/* ------------------------------------------------------------------ */

/* the integer with n little-endian limbs, mod q */
u64 or_mpn_mod_q(const u64* a, u64 n, u64 q)
{
    u64 r = 0;
    for (u64 i = n; i-- > 0;)
        r = (u64)((((u128)r << 64) | a[i]) % q);
    return r;
}

/* sum c_i x^i mod q over n coefficients (each reduced mod q first) */
u64 or_nmod_poly_eval(const u64* c, u64 n, u64 x, u64 q)
{
    u64 r = 0;
    x %= q;
    for (u64 i = n; i-- > 0;) r = addm(mulm(r, x, q), c[i] % q, q);
    return r;
}

/* sum c_i x^i mod q, c_i w-limb two's complement */
u64 or_fmpz_poly_eval_mod_q(const u64* c, u64 w, u64 n, u64 x, u64 q)
{
    u64 r = 0;
    x %= q;
    for (u64 i = n; i-- > 0;) {
        const u64* ci = c + i * w;
        u64 v = or_mpn_mod_q(ci, w, q);
        if ((s64)ci[w - 1] < 0) {
            /* value = limbs - 2^(64 w) */
            u64 t = powm(2, 64 * w, q);
            v = subm(v, t, q);
        }
        r = addm(mulm(r, x, q), v, q);
    }
    return r;
}
