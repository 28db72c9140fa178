// fa_window.hip -- the words of the reference's integer windows.
//
// _mpn_ctx_mpn_mul_range (mpn_mul.c:1617-1978, the body of
// flint_mpn_mulmid_fft_small) returns limbs [lo, hi) of a*b as a lower
// approximation whose words depend on the plan FLINT picks for (an, bn) on
// its thread count: with chunk coefficients c_j = sum_{i+k=j} A_i B_k
// (A_i = bits [i b, (i+1) b) of a, b the plan's chunk width), T = 2^(64 lo)
// and P = a*b,
//   - two-prime plans (max(an, bn) <= 40000, mpn_mul_np2.c:247-433): each
//     chunk contributes floor(c_j 2^(j b) / T), so
//         W = floor(P / T) - floor(F / T),
//         F = sum over j b < 64 lo of (c_j 2^(j b) mod T);
//   - wider plans (mpn_mul.c:1669-1978): the coefficients j >= c_lo are
//     summed exactly, c_lo = ceil((lo - n) 64 / b) for lo > n (n the limbs
//     of M np, mpn_mul.c:1105-1117), else 0, so
//         W = floor(P / T) - [ (P mod T) < X ],
//         X = (sum_{c_lo <= j, j b < 64 lo} c_j 2^(j b)) mod T
//     (the coefficients below c_lo add up to less than T).
// Both corrections need only the few coefficients whose bits reach from
// just below bit 64 lo to it, and limbs [L, lo) of P with L the limb of the
// first of them (the lower limbs of P and the rest of F or X are multiples
// of 2^(64 L), or below it, and cannot change the comparison).  The exact
// product comes from the ordinary device path; here:
//   fa_k_win_dot / fa_k_win_sum   the boundary coefficients exactly, as
//                                 512-bit sums of chunk products;
//   fa_k_win_delta                the correction d (one thread);
//   fa_k_win_sub                  window -= d with its borrow.
#include "fa_internal.h"
#include "fa_arith.h"

#include <string.h>

typedef unsigned __int128 u128;

#define FA_WIN_ACC 8        // limbs of a coefficient sum (< 2^414)
#define FA_WIN_MAXK 8       // boundary coefficients at most
#define FA_WIN_MAXM 14      // limbs [L, lo) of P at most
#define FA_WIN_TPB 256
#define FA_WIN_PER 8192     // chunk products per workgroup

// ------------------------------------------------------------------------
// the reference's plan for (an, bn) with FLINT on nthreads threads
// ------------------------------------------------------------------------

static int nbits_u64(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }
static uint64_t cdiv64(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// limbs of M = p_0 ... p_(np-1) times np: the reference's coeff_len
// (mpn_mul.c:1105-1117), and M itself in those limbs
static int coeff_len(int np, uint64_t* M)
{
    fa_crt_consts C;
    fa_crt_consts_init(&C, np, 0);
    uint64_t t[FA_CRT_MAXL + 2];
    memset(t, 0, sizeof(t));
    memset(M, 0, (FA_CRT_MAXL + 2) * sizeof(uint64_t));
    uint64_t c = 0;
    for (int i = 0; i < FA_CRT_MAXL; i++) {
        M[i] = C.M[i];
        const u128 v = (u128)C.M[i] * (uint64_t)np + c;
        t[i] = (uint64_t)v;
        c = (uint64_t)(v >> 64);
    }
    t[FA_CRT_MAXL] = c;
    int n = FA_CRT_MAXL + 1;
    while (n > 1 && t[n - 1] == 0) n--;
    return n;
}

int fa_ref_coeff_len(int np)
{
    uint64_t M[FA_CRT_MAXL + 2];
    return coeff_len(np, M);
}

// crt_data_find_bn_bound (mpn_mul.c:152-184): the largest bn with
// 64 bn <= bits M / 2^(2 bits) - (bits - 1), all in coeff_len limbs
static uint64_t ref_bn_bound(int np, uint64_t bits)
{
    uint64_t M[FA_CRT_MAXL + 2];
    const int n = coeff_len(np, M);
    uint64_t x[FA_CRT_MAXL + 3];
    memset(x, 0, sizeof(x));
    uint64_t c = 0;
    for (int i = 0; i < n; i++) {
        const u128 v = (u128)M[i] * bits + c;
        x[i] = (uint64_t)v;
        c = (uint64_t)(v >> 64);
    }
    x[n] = c;
    const uint64_t q = 2 * bits / 64, r = 2 * bits % 64;
    if (q >= (uint64_t)n + 1) return 0;
    uint64_t* y = x + q;
    const int yn = n + 1 - (int)q;
    if (r)
        for (int i = 0; i < yn; i++) y[i] = (y[i] >> r) | (i + 1 < yn ? y[i + 1] << (64 - r) : 0);
    // y -= bits - 1; a borrow out leaves the bound at 0
    uint64_t sub = bits - 1;
    for (int i = 0; i < yn && sub; i++) {
        const uint64_t before = y[i];
        y[i] = before - sub;
        sub = before < sub;
    }
    if (sub) return 0;
    for (int i = 0; i < yn; i++) y[i] = (y[i] >> 6) | (i + 1 < yn ? y[i + 1] << 58 : 0);
    for (int i = 1; i < yn; i++)
        if (y[i]) return ~UINT64_C(0);
    return y[0];
}

// crt_data_find_bits (mpn_mul.c:191-212)
static uint64_t ref_find_bits(int np, uint64_t bn)
{
    uint64_t M[FA_CRT_MAXL + 2];
    const int n = coeff_len(np, M);
    int pbits = 0;
    for (int i = n - 1; i >= 0; i--)
        if (M[i]) {
            pbits = 64 * i + nbits_u64(M[i]);
            break;
        }
    uint64_t bits = ((uint64_t)pbits - nbits_u64(bn)) / 2;
    if (bits < 66) bits = 66;
    if (bn > ref_bn_bound(np, bits)) {
        do bits--; while (bits > 65 && bn > ref_bn_bound(np, bits));
    } else {
        while (bn <= ref_bn_bound(np, bits + 1)) bits++;
    }
    return bits;
}

// the fifteen profiles of mpn_ctx_init (mpn_mul.c:1171-1194)
static const int ref_prof_np[15] = {4, 4, 4, 5, 5, 5, 6, 6, 6, 7, 7, 7, 8, 8, 8};
static const int ref_prof_bits[15] = {84, 88, 92, 112, 116, 120, 136, 140, 144,
                                      160, 164, 168, 184, 188, 192};

// mpn_ctx_best_profile (mpn_mul.c:1248-1340) for FLINT on nthreads threads
static void ref_best_profile(uint64_t an, uint64_t bn, int nthreads, int* np_out,
                             uint64_t* bits_out)
{
    uint64_t bound[15];
    for (int i = 0; i < 15; i++) bound[i] = ref_bn_bound(ref_prof_np[i], ref_prof_bits[i]);
    const uint64_t zn = an + bn;
    int limit = zn < 2048 ? 1 : zn < 4096 ? 4 : zn < 8192 ? 5 : zn < 16384 ? 6 : zn < 32768 ? 7 : 8;
    int nt = nthreads < 1 ? 1 : nthreads;
    if (nt > limit) nt = limit;
    if (bn > bound[0]) {
        int np = nt < 8 ? nt : 8;
        if (np < 4) np = 4;
        *np_out = np;
        *bits_out = ref_find_bits(np, bn);
        return;
    }
    int i = 0, best = 0;
    double best_score = 100000000.0 * (double)(an + bn);
    for (;;) {
        while (i + 1 < 15 && bn <= bound[i + 1] && ref_prof_np[i + 1] == ref_prof_np[i]) i++;
        const int np = ref_prof_np[i];
        if (np % nt == 0) {
            const uint64_t bits = ref_prof_bits[i];
            const uint64_t zlen = cdiv64(64 * an, bits) + cdiv64(64 * bn, bits) - 1;
            const uint64_t ztrunc = cdiv64(zlen, 256) * 256;
            int depth = 8;
            while ((UINT64_C(1) << depth) < ztrunc) depth++;
            const double ratio = (double)ztrunc / (double)(UINT64_C(1) << depth);
            double score = (1 - 0.25 * ratio) * (1.0 / 1000000);
            score *= (double)(np * depth);
            score *= (double)ztrunc;
            if (score < best_score) {
                best = i;
                best_score = score;
            }
        }
        do {
            if (++i >= 15) {
                *np_out = ref_prof_np[best];
                *bits_out = (uint64_t)ref_prof_bits[best];
                return;
            }
        } while (bn > bound[i]);
    }
}

// _fft_small_np2_choose with len_bound 1 (mpn_mul_np2.c:46-70), an >= bn
static uint64_t ref_np2_bits(uint64_t an, uint64_t bn)
{
    uint64_t P[8];
    fa_default_primes(P);
    const u128 pp = (u128)P[0] * P[1];
    const int prodbits = 64 + nbits_u64((uint64_t)(pp >> 64)) - 1;
    for (int d = 8; d <= 30; d++)
        for (uint64_t b = 45; b >= 40; b--) {
            const uint64_t alen = cdiv64(64 * an, b), blen = cdiv64(64 * bn, b);
            if (alen + blen - 1 > (UINT64_C(1) << d)) continue;
            if (2 * (int)b + nbits_u64(blen) > prodbits) continue;
            return b;
        }
    return 0;
}

struct fa_win_plan {
    int np2;                  // two-prime semantics
    int nk;                   // boundary coefficients k0 .. k0 + nk - 1 (0: exact window)
    uint64_t bits, alen, blen, k0, L, m;
    unsigned nblk;            // workgroups per coefficient in fa_k_win_dot
};

// nthreads = 0: the exact window (no correction)
static int win_plan(uint64_t an, uint64_t bn, uint64_t lo, int nthreads, fa_win_plan* W)
{
    memset(W, 0, sizeof(*W));
    if (nthreads <= 0 || lo == 0) return FA_OK;
    const uint64_t big = an > bn ? an : bn, small = an > bn ? bn : an;
    int np;
    if (big <= 40000) {
        W->np2 = 1;
        W->bits = ref_np2_bits(big, small);
        if (!W->bits) return fa_set_error(FA_E_RANGE, "no two-prime plan for %llu x %llu limbs",
                                          (unsigned long long)an, (unsigned long long)bn);
        np = 2;
    } else {
        ref_best_profile(an, bn, nthreads, &np, &W->bits);
    }
    const uint64_t b = W->bits;
    W->alen = cdiv64(64 * an, b);
    W->blen = cdiv64(64 * bn, b);
    const uint64_t zlen = W->alen + W->blen - 1;
    uint64_t E = cdiv64(64 * lo, b);
    if (E > zlen) E = zlen;
    uint64_t k0;
    if (W->np2) {
        // every coefficient below k0 lies below T, and so does their sum:
        // c_j < mlen 2^(2b) and j b + 2b + bits(mlen) <= 64 lo
        const uint64_t mlen = W->alen < W->blen ? W->alen : W->blen;
        const uint64_t guard = 2 * b + (uint64_t)nbits_u64(mlen);
        k0 = 64 * lo < guard ? 0 : (64 * lo - guard) / b + 1;
    } else {
        uint64_t M[FA_CRT_MAXL + 2];
        const uint64_t nn = (uint64_t)coeff_len(np, M);
        k0 = lo > nn ? cdiv64((lo - nn) * 64, b) : 0;
        if (k0 == 0) return FA_OK;      // c_lo = 0: nothing dropped
    }
    if (k0 >= E) return FA_OK;
    W->k0 = k0;
    W->nk = (int)(E - k0);
    W->L = k0 * b / 64;
    W->m = lo - W->L;
    if (W->nk > FA_WIN_MAXK || W->m > FA_WIN_MAXM || b > 192)
        return fa_set_error(FA_E_RANGE, "window correction outside its limits (nk %d, m %llu, "
                            "bits %llu)", W->nk, (unsigned long long)W->m,
                            (unsigned long long)b);
    const uint64_t mlen = W->alen < W->blen ? W->alen : W->blen;
    W->nblk = (unsigned)cdiv64(mlen, FA_WIN_PER);
    return FA_OK;
}

// ------------------------------------------------------------------------
// device side
// ------------------------------------------------------------------------

// chunk i (bits [i b, (i+1) b), b <= 64 NLC) of x[0, xn)
template <int NLC>
FA_DEV void win_chunk(const uint64_t* __restrict__ x, uint64_t xn, uint64_t bits, uint64_t i,
                      uint64_t* c)
{
    const uint64_t pos = i * bits, w = pos >> 6;
    const int s = (int)(pos & 63);
    uint64_t v[NLC + 1];
#pragma unroll
    for (int k = 0; k <= NLC; k++) v[k] = (w + k < xn) ? x[w + k] : 0;
#pragma unroll
    for (int k = 0; k < NLC; k++) c[k] = s ? (v[k] >> s) | (v[k + 1] << (64 - s)) : v[k];
    const int top = (int)bits - 64 * (NLC - 1);     // bits in the top limb, 1..64
    if (top < 64) c[NLC - 1] &= (UINT64_C(1) << top) - 1;
}

// acc += x * y (NLC limbs each), acc of FA_WIN_ACC limbs
template <int NLC>
FA_DEV void win_addmul(uint64_t* acc, const uint64_t* x, const uint64_t* y)
{
    uint64_t p[2 * NLC];
#pragma unroll
    for (int k = 0; k < 2 * NLC; k++) p[k] = 0;
#pragma unroll
    for (int i = 0; i < NLC; i++) {
        uint64_t carry = 0;
#pragma unroll
        for (int j = 0; j < NLC; j++) {
            const uint64_t lo = x[i] * y[j], hi = __umul64hi(x[i], y[j]);
            uint64_t t = p[i + j] + lo;
            uint64_t c1 = t < lo;
            t += carry;
            c1 += t < carry;
            p[i + j] = t;
            carry = hi + c1;    // p + x y + carry < 2^128: fits
        }
        p[i + NLC] = carry;
    }
    uint64_t c = 0;
#pragma unroll
    for (int k = 0; k < FA_WIN_ACC; k++) {
        const uint64_t v = k < 2 * NLC ? p[k] : 0;
        uint64_t t = acc[k] + v;
        uint64_t c1 = t < v;
        t += c;
        c1 += t < c;
        acc[k] = t;
        c = c1;
    }
}

FA_DEV void win_add(uint64_t* acc, const uint64_t* v)
{
    uint64_t c = 0;
#pragma unroll
    for (int k = 0; k < FA_WIN_ACC; k++) {
        uint64_t t = acc[k] + v[k];
        uint64_t c1 = t < v[k];
        t += c;
        c1 += t < c;
        acc[k] = t;
        c = c1;
    }
}

// one workgroup reduces its threads' sums (s: FA_WIN_TPB x FA_WIN_ACC)
FA_DEV void win_block_reduce(uint64_t* acc, uint64_t* s)
{
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < FA_WIN_ACC; k++) s[t * FA_WIN_ACC + k] = acc[k];
    __syncthreads();
    for (int h = FA_WIN_TPB / 2; h > 0; h >>= 1) {
        if (t < h) {
            uint64_t v[FA_WIN_ACC];
#pragma unroll
            for (int k = 0; k < FA_WIN_ACC; k++) v[k] = s[(t + h) * FA_WIN_ACC + k];
            win_add(s + t * FA_WIN_ACC, v);
        }
        __syncthreads();
    }
}

// workgroup (x, y): chunk products A_i B_(k-i) of coefficient k = k0 + y
// for i in its slice of FA_WIN_PER; partial sums to part[(y nblk + x) ACC]
template <int NLC>
__global__ void __launch_bounds__(FA_WIN_TPB) fa_k_win_dot(const uint64_t* __restrict__ a,
                                                            uint64_t an,
                                                            const uint64_t* __restrict__ b,
                                                            uint64_t bn, uint64_t bits,
                                                            uint64_t alen, uint64_t blen,
                                                            uint64_t k0, uint64_t* part)
{
    __shared__ uint64_t s[FA_WIN_TPB * FA_WIN_ACC];
    const uint64_t k = k0 + blockIdx.y;
    const uint64_t ilo = k >= blen - 1 ? k - (blen - 1) : 0;
    const uint64_t ihi = k < alen - 1 ? k : alen - 1;     // inclusive; ilo <= ihi for k < zlen
    uint64_t acc[FA_WIN_ACC];
#pragma unroll
    for (int q = 0; q < FA_WIN_ACC; q++) acc[q] = 0;
    const uint64_t start = ilo + (uint64_t)blockIdx.x * FA_WIN_PER;
    for (uint64_t i = start + threadIdx.x; i <= ihi && i < start + FA_WIN_PER; i += FA_WIN_TPB) {
        uint64_t x[NLC], y[NLC];
        win_chunk<NLC>(a, an, bits, i, x);
        win_chunk<NLC>(b, bn, bits, k - i, y);
        win_addmul<NLC>(acc, x, y);
    }
    win_block_reduce(acc, s);
    if (threadIdx.x < FA_WIN_ACC)
        part[((uint64_t)blockIdx.y * gridDim.x + blockIdx.x) * FA_WIN_ACC + threadIdx.x] =
            s[threadIdx.x];
}

// workgroup y: coefficient k0 + y = the sum of its nblk partial sums
__global__ void __launch_bounds__(FA_WIN_TPB) fa_k_win_sum(const uint64_t* __restrict__ part,
                                                            unsigned nblk, uint64_t* coef)
{
    __shared__ uint64_t s[FA_WIN_TPB * FA_WIN_ACC];
    uint64_t acc[FA_WIN_ACC];
#pragma unroll
    for (int q = 0; q < FA_WIN_ACC; q++) acc[q] = 0;
    for (unsigned x = threadIdx.x; x < nblk; x += FA_WIN_TPB)
        win_add(acc, part + ((uint64_t)blockIdx.y * nblk + x) * FA_WIN_ACC);
    win_block_reduce(acc, s);
    if (threadIdx.x < FA_WIN_ACC)
        coef[(uint64_t)blockIdx.y * FA_WIN_ACC + threadIdx.x] = s[threadIdx.x];
}

// dst[0, nd) += (v << sh) mod 2^(64 nd), v of FA_WIN_ACC limbs
FA_DEV void win_shl_add(uint64_t* dst, int nd, const uint64_t* v, uint64_t sh)
{
    const uint64_t q = sh >> 6;
    const int r = (int)(sh & 63);
    uint64_t c = 0;
    for (int d = 0; d < nd; d++) {
        uint64_t w = 0;
        if ((uint64_t)d >= q) {
            const uint64_t i = (uint64_t)d - q;
            const uint64_t lo = i < FA_WIN_ACC ? v[i] : 0;
            const uint64_t prev = (i >= 1 && i - 1 < FA_WIN_ACC) ? v[i - 1] : 0;
            w = r ? (lo << r) | (prev >> (64 - r)) : lo;
        }
        uint64_t t = dst[d] + w;
        uint64_t c1 = t < w;
        t += c;
        c1 += t < c;
        dst[d] = t;
        c = c1;
    }
}

// the correction d of the window (one thread): P = the full product,
// coef = the nk boundary coefficients from k0, limbs [L, L + m) of P
__global__ void fa_k_win_delta(const uint64_t* __restrict__ P, const uint64_t* __restrict__ coef,
                               int nk, uint64_t k0, uint64_t bits, uint64_t L, int m, int np2,
                               uint64_t* delta)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t Y[FA_WIN_MAXM + 1], R[FA_WIN_MAXM + 1], U[FA_WIN_MAXM + 1];
    for (int d = 0; d <= m; d++) Y[d] = R[d] = 0;
    for (int j = 0; j < nk; j++) {
        const uint64_t sh = (k0 + j) * bits - 64 * L;
        // Y: the sum of the coefficients' shifted values mod 2^(64 m)
        win_shl_add(Y, m, coef + j * FA_WIN_ACC, sh);
        if (np2) {
            // R: the sum of each shifted value mod 2^(64 m), in m + 1 limbs
            uint64_t t[FA_WIN_MAXM + 1];
            for (int d = 0; d <= m; d++) t[d] = 0;
            win_shl_add(t, m, coef + j * FA_WIN_ACC, sh);
            t[m] = 0;
            uint64_t c = 0;
            for (int d = 0; d <= m; d++) {
                uint64_t s = R[d] + t[d];
                uint64_t c1 = s < t[d];
                s += c;
                c1 += s < c;
                R[d] = s;
                c = c1;
            }
        }
    }
    // U = limbs [L, L + m) of P minus Y, mod 2^(64 m); lt: P's limbs < Y
    uint64_t br = 0;
    for (int d = 0; d < m; d++) {
        const uint64_t p = P[L + d];
        const uint64_t t = p - Y[d];
        const uint64_t b1 = p < Y[d];
        U[d] = t - br;
        br = b1 | (t < br);
    }
    uint64_t dv;
    if (!np2) {
        dv = br;     // [ (P mod T) < X ]
    } else {
        // floor((U + R) / 2^(64 m))
        U[m] = 0;
        uint64_t c = 0;
        for (int d = 0; d <= m; d++) {
            uint64_t s = U[d] + R[d];
            uint64_t c1 = s < R[d];
            s += c;
            c1 += s < c;
            U[d] = s;
            c = c1;
        }
        dv = U[m];
    }
    *delta = dv;
}

// z[0, w) -= *delta, the borrow rippling through zero limbs (one workgroup)
__global__ void __launch_bounds__(FA_WIN_TPB) fa_k_win_sub(uint64_t* z, uint64_t w,
                                                            const uint64_t* __restrict__ delta)
{
    __shared__ unsigned first;
    __shared__ int borrow;
    const uint64_t d = *delta;
    if (d == 0) return;
    if (threadIdx.x == 0) {
        const uint64_t v = z[0];
        z[0] = v - d;
        borrow = v < d;
    }
    __syncthreads();
    if (!borrow) return;
    for (uint64_t base = 1; base < w; base += FA_WIN_TPB) {
        if (threadIdx.x == 0) first = FA_WIN_TPB;
        __syncthreads();
        const uint64_t i = base + threadIdx.x;
        if (i < w && z[i] != 0) atomicMin(&first, threadIdx.x);
        __syncthreads();
        const unsigned f = first;
        if (i < w) {
            if (threadIdx.x < f) z[i] = ~UINT64_C(0);
            else if (threadIdx.x == f) z[i] -= 1;
        }
        if (f < FA_WIN_TPB) return;
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------

size_t fa_window_work_bytes(uint64_t an, uint64_t bn, uint64_t lo, int nthreads)
{
    fa_win_plan W;
    if (win_plan(an, bn, lo, nthreads, &W) != FA_OK || !W.nk) return 0;
    return ((size_t)W.nk * W.nblk + (size_t)W.nk) * FA_WIN_ACC * 8 + 64;
}

// zw: the window [lo, hi) of the exact product P (device, w limbs, may lie
// inside P); a, b: the device operands; work: fa_window_work_bytes bytes
int fa_window_fix(uint64_t* zw, const uint64_t* P, uint64_t lo, uint64_t hi, const uint64_t* a,
                  uint64_t an, const uint64_t* b, uint64_t bn, int nthreads, void* work,
                  hipStream_t s)
{
    fa_win_plan W;
    int rc = win_plan(an, bn, lo, nthreads, &W);
    if (rc != FA_OK) return rc;
    if (!W.nk) return FA_OK;
    uint64_t* part = (uint64_t*)work;
    uint64_t* coef = part + (size_t)W.nk * W.nblk * FA_WIN_ACC;
    uint64_t* delta = coef + (size_t)W.nk * FA_WIN_ACC;
    const dim3 gd(W.nblk, (unsigned)W.nk);
    const int nlc = (int)cdiv64(W.bits, 64);
    if (nlc == 1)
        FA_LAUNCH("fa_k_win_dot", fa_k_win_dot<1>, gd, dim3(FA_WIN_TPB), 0, s, a, an, b, bn,
                  W.bits, W.alen, W.blen, W.k0, part);
    else if (nlc == 2)
        FA_LAUNCH("fa_k_win_dot", fa_k_win_dot<2>, gd, dim3(FA_WIN_TPB), 0, s, a, an, b, bn,
                  W.bits, W.alen, W.blen, W.k0, part);
    else
        FA_LAUNCH("fa_k_win_dot", fa_k_win_dot<3>, gd, dim3(FA_WIN_TPB), 0, s, a, an, b, bn,
                  W.bits, W.alen, W.blen, W.k0, part);
    FA_LAUNCH("fa_k_win_sum", fa_k_win_sum, dim3(1, (unsigned)W.nk), dim3(FA_WIN_TPB), 0, s, part,
              W.nblk, coef);
    FA_LAUNCH("fa_k_win_delta", fa_k_win_delta, dim3(1), dim3(64), 0, s, P, coef, W.nk, W.k0,
              W.bits, W.L, (int)W.m, W.np2, delta);
    FA_LAUNCH("fa_k_win_sub", fa_k_win_sub, dim3(1), dim3(FA_WIN_TPB), 0, s, zw, hi - lo, delta);
    return hipGetLastError() == hipSuccess ? FA_OK
                                           : fa_set_error(FA_E_HIP, "window correction launch");
}

// the reference's plan as the window semantics see it: out = {two-prime,
// bits, k0, nk}
extern "C" int fa_mpn_window_plan(uint64_t an, uint64_t bn, uint64_t lo, int nthreads,
                                  int64_t* out4)
{
    if (!an || !bn || !out4) return fa_set_error(FA_E_ARG, "fa_mpn_window_plan: bad arguments");
    fa_win_plan W;
    const int rc = win_plan(an, bn, lo, nthreads, &W);
    if (rc != FA_OK) return rc;
    out4[0] = W.np2;
    out4[1] = (int64_t)W.bits;
    out4[2] = (int64_t)W.k0;
    out4[3] = W.nk;
    return FA_OK;
}
