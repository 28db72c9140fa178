// fa_ks.hip -- fmpz polynomial products beyond the eight-prime capacity, by
// Kronecker substitution onto the integer product (the route the
// reference's callers take when _fmpz_poly_mul_mid_default_mpn_ctx declines:
// fmpz_poly/mul.c:160, _fmpz_poly_mul_KS).
//
// Two routes.  The default packs signed coefficients into s-bit slots with
// borrows and needs one integer product (see the signed-slot kernels
// below).  The other packs without borrows: with a = a+ - a- (the positive
// part and the negated negative part, coefficient by coefficient)
//     a b = 2 (a+ b+ + a- b-) - |a| |b|,
// an identity of polynomials, so it holds coefficient by coefficient.  Each
// of the three products has non-negative coefficients below
// min(an, bn) 2^(|b1|+|b2|) < 2^(s-1) for the slot width
//     s = |b1| + |b2| + bitcount(min(an, bn)) + 1,
// so packing at s bits per coefficient puts every coefficient in its own
// slot with no carry between slots, at three times the product work.
// Packing is a gather per output limb; the combination is one thread per
// output coefficient.
#include <stdio.h>
#include <string.h>

#include "fa_internal.h"
#include "fa_arith.h"
#include "../../include/flint_amd.h"

enum { FA_KS_PLUS = 0, FA_KS_MINUS = 1, FA_KS_ABS = 2 };

// index of the lowest nonzero limb of each w-limb coefficient (w if zero)
__global__ void __launch_bounds__(256) fa_k_ks_low(const uint64_t* __restrict__ a, int w,
                                                   uint64_t n, int32_t* __restrict__ low)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t* c = a + i * (uint64_t)w;
    int m = 0;
    while (m < w && c[m] == 0) m++;
    low[i] = m;
}

// limb m of |c| for a w-limb two's complement coefficient c: for c < 0,
// -c = ~c + 1, whose limb m takes the carry iff limbs 0..m-1 of c are zero
FA_DEV uint64_t fa_ks_mag(const uint64_t* c, int w, int m, int neg, int low)
{
    if (m >= w) return 0;
    return neg ? ~c[m] + (m <= low ? 1u : 0u) : c[m];
}

// out[j] = limb j of sum_i part(c_i) 2^(s i), part = c+, c- or |c|
__global__ void __launch_bounds__(256) fa_k_ks_pack(const uint64_t* __restrict__ a, int w,
                                                    uint64_t n, const int32_t* __restrict__ low,
                                                    uint64_t s, int mode,
                                                    uint64_t* __restrict__ out, uint64_t L)
{
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= L) return;
    const uint64_t bit0 = 64 * j;
    uint64_t i0 = bit0 / s;
    uint64_t i1 = (bit0 + 63) / s;
    if (i1 >= n) i1 = n - 1;
    uint64_t acc = 0;
    for (uint64_t i = i0; i <= i1 && i < n; i++) {
        const uint64_t* c = a + i * (uint64_t)w;
        const int neg = (int)(c[w - 1] >> 63);
        if ((mode == FA_KS_PLUS && neg) || (mode == FA_KS_MINUS && !neg)) continue;
        const int lo = low[i];
        const int64_t rel = (int64_t)bit0 - (int64_t)(s * i);   // bit of c at bit0
        if (rel >= 0) {
            if ((uint64_t)rel >= s) continue;
            const int q = (int)(rel >> 6), r = (int)(rel & 63);
            uint64_t v = fa_ks_mag(c, w, q, neg, lo) >> r;
            if (r) v |= fa_ks_mag(c, w, q + 1, neg, lo) << (64 - r);
            acc |= v;
        } else {
            acc |= fa_ks_mag(c, w, 0, neg, lo) << (int)(-rel);
        }
    }
    out[j] = acc;
}

// limb l of slot k (bits [s k, s k + s)) of a product of nP limbs
FA_DEV uint64_t fa_ks_slot(const uint64_t* __restrict__ P, uint64_t nP, uint64_t s, uint64_t k,
                           int l)
{
    const uint64_t lb = 64 * (uint64_t)l;
    if (lb >= s) return 0;
    const uint64_t off = s * k + lb;
    const uint64_t q = off >> 6;
    const int r = (int)(off & 63);
    uint64_t v = q < nP ? P[q] >> r : 0;
    if (r && q + 1 < nP) v |= P[q + 1] << (64 - r);
    if (lb + 64 > s) v &= (UINT64_C(1) << (s - lb)) - 1;
    return v;
}

// z[t] = coefficient zl + t of a b = 2 (P1 + P4) - P5 as zw limbs (two's
// complement), zero past the product's length zn
__global__ void __launch_bounds__(256) fa_k_ks_combine(const uint64_t* __restrict__ P1,
                                                       const uint64_t* __restrict__ P4,
                                                       const uint64_t* __restrict__ P5,
                                                       uint64_t nP, uint64_t s, uint64_t zl,
                                                       uint64_t cnt, uint64_t zn, int zw,
                                                       uint64_t* __restrict__ z)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    uint64_t* o = z + t * (uint64_t)zw;
    const uint64_t k = zl + t;
    if (k >= zn) {
        for (int l = 0; l < zw; l++) o[l] = 0;
        return;
    }
    unsigned carry = 0, borrow = 0;
    for (int l = 0; l < zw; l++) {
        const uint64_t e1 = fa_ks_slot(P1, nP, s, k, l);
        const uint64_t e4 = fa_ks_slot(P4, nP, s, k, l);
        const uint64_t e5 = fa_ks_slot(P5, nP, s, k, l);
        // 2 (e1 + e4) + carry, as 64 bits and a carry of at most 3
        const uint64_t h = e1 + e4;
        unsigned c1 = h < e1;
        const uint64_t d = h << 1;
        unsigned c2 = (unsigned)(h >> 63) + 2 * c1;
        const uint64_t sum = d + carry;
        c2 += sum < d;
        carry = c2;
        // minus e5 and the borrow
        const uint64_t r1 = sum - e5;
        unsigned b1 = sum < e5;
        const uint64_t r2 = r1 - borrow;
        b1 += r1 < borrow;
        borrow = b1;
        o[l] = r2;
    }
}

// ---------------------------------------------------------------------------
// Signed slots, one product (the default).  With sigma = +1 or -1 chosen so
// that sigma a has a positive top nonzero coefficient, the packed integer
//     A = sum_i sigma a_i 2^(s i)
// is positive; its base-2^s digits are d_i = (sigma a_i - br_i) mod 2^s with
// the borrow br_0 = 0, br_(i+1) = [sigma a_i - br_i < 0], i.e. generate
// [sigma a_i < 0] and propagate [a_i = 0]: a 1-bit scan over coefficients.
// The product C = A B = sum_k c_k 2^(s k) with |c_k| < 2^(s-1) (the slot
// width above) has base-2^s digits e_k, and
//     c_k = e_k + t_(k-1) - 2^s t_k,  t_k = [e_k + t_(k-1) >= 2^(s-1)],
// again a 1-bit scan: generate [e_k >= 2^(s-1)], propagate
// [e_k = 2^(s-1) - 1].  So c_k is the s-bit field e_k + t_(k-1), read as a
// signed s-bit number, times sigma_a sigma_b.
// ---------------------------------------------------------------------------

// one workgroup per operand: the lowest nonzero limb of each coefficient,
// the sign sigma of the top nonzero coefficient (+1 for a zero operand) and
// the borrow into each coefficient's slot.  Each thread keeps one
// contiguous block of coefficients through both passes.
FA_DEV void ks_sflags_body(const uint64_t* __restrict__ a, int w, uint64_t n,
                           int32_t* __restrict__ low, uint8_t* __restrict__ br,
                           int32_t* __restrict__ sig, int64_t* stop, uint8_t* sg, uint8_t* sp)
{
    const int t = threadIdx.x;
    const uint64_t m = (n + 1023) / 1024;
    const uint64_t i0 = (uint64_t)t * m < n ? (uint64_t)t * m : n;
    const uint64_t i1 = i0 + m < n ? i0 + m : n;
    int64_t top = -1;
    for (uint64_t i = i0; i < i1; i++) {
        const uint64_t* c = a + i * (uint64_t)w;
        int l = 0;
        while (l < w && c[l] == 0) l++;
        low[i] = l;
        if (l < w) top = (int64_t)i;
    }
    stop[t] = top;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (t < h && stop[t + h] > stop[t]) stop[t] = stop[t + h];
        __syncthreads();
    }
    const int64_t ti = stop[0];
    const int negs = ti >= 0 ? (int)(a[(uint64_t)ti * w + w - 1] >> 63) : 0;
    if (t == 0) *sig = negs ? -1 : 1;
    // block aggregates of (generate, propagate), composed left to right
    uint8_t G = 0, P = 1;
    for (uint64_t i = i0; i < i1; i++) {
        const int z = low[i] >= w;
        const int neg = (int)(a[i * (uint64_t)w + w - 1] >> 63);
        const uint8_t gi = (uint8_t)(!z && neg != negs), pi = (uint8_t)z;
        G = (uint8_t)(gi | (pi & G));
        P = (uint8_t)(P & pi);
    }
    sg[t] = G;
    sp[t] = P;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        uint8_t g = sg[t], q = sp[t];
        if (t >= d) {
            g = (uint8_t)(g | (q & sg[t - d]));
            q = (uint8_t)(q & sp[t - d]);
        }
        __syncthreads();
        sg[t] = g;
        sp[t] = q;
        __syncthreads();
    }
    uint8_t cin = t ? sg[t - 1] : 0;
    for (uint64_t i = i0; i < i1; i++) {
        br[i] = cin;
        const int z = low[i] >= w;
        const int neg = (int)(a[i * (uint64_t)w + w - 1] >> 63);
        cin = (uint8_t)((!z && neg != negs) | (z & cin));
    }
}

__global__ void __launch_bounds__(1024) fa_k_ks_sflags(const uint64_t* __restrict__ a, int w,
                                                       uint64_t n, int32_t* __restrict__ low,
                                                       uint8_t* __restrict__ br,
                                                       int32_t* __restrict__ sig)
{
    __shared__ int64_t stop[1024];
    __shared__ uint8_t sg[1024], sp[1024];
    ks_sflags_body(a, w, n, low, br, sig, stop, sg, sp);
}

// limb m (any m >= 0, sign-extended past w) of v = sigma c - br for a w-limb
// two's complement coefficient c whose lowest nonzero limb is `low` (w if c
// is zero); negs: sigma = -1
FA_DEV uint64_t fa_ks_vlimb(const uint64_t* c, int w, int m, int negs, int br, int low)
{
    const bool zero = low >= w;
    const uint64_t cs = (c[w - 1] >> 63) ? ~UINT64_C(0) : 0;
    if (!negs) {
        if (!br) return m < w ? c[m] : cs;
        if (zero) return ~UINT64_C(0);                       // 0 - 1
        return m < w ? c[m] - (m <= low ? 1u : 0u) : cs;     // c - 1, same sign as c
    }
    if (br) return m < w ? ~c[m] : ~cs;                      // -c - 1 = ~c
    if (zero) return 0;
    return m < w ? ~c[m] + (m <= low ? 1u : 0u) : ~cs;       // -c
}

// out[j] = limb j of sum_i d_i 2^(s i), d_i = (sigma c_i - br_i) mod 2^s
FA_DEV uint64_t ks_spack_limb(const uint64_t* __restrict__ a, int w, uint64_t n,
                              const int32_t* __restrict__ low, const uint8_t* __restrict__ br,
                              int negs, uint64_t s, uint64_t j)
{
    const uint64_t bit0 = 64 * j;
    // floor(bit0 / s) from the correctly rounded double quotient (both
    // operands below 2^53, so exact as doubles), which is at most one above
    // it: a 64-bit integer division is a long emulated sequence here
    uint64_t i0 = (uint64_t)((double)bit0 / (double)s);
    if (i0 * s > bit0) i0--;
    uint64_t acc = 0;
    for (uint64_t i = i0; i < n && s * i < bit0 + 64; i++) {
        const uint64_t* c = a + i * (uint64_t)w;
        const int lo = low[i], b = br[i];
        const int64_t rel = (int64_t)bit0 - (int64_t)(s * i);   // slot bit at bit0
        uint64_t win;
        if (rel >= 0) {
            const int q = (int)(rel >> 6), r = (int)(rel & 63);
            win = fa_ks_vlimb(c, w, q, negs, b, lo) >> r;
            if (r) win |= fa_ks_vlimb(c, w, q + 1, negs, b, lo) << (64 - r);
        } else {
            win = fa_ks_vlimb(c, w, 0, negs, b, lo) << (int)(-rel);
        }
        // keep window positions p with 0 <= rel + p < s
        const int64_t plo = rel < 0 ? -rel : 0;
        const int64_t phi = (int64_t)s - rel < 64 ? (int64_t)s - rel : 64;
        if (phi <= plo) continue;
        uint64_t mask = phi == 64 ? ~UINT64_C(0) : ((UINT64_C(1) << phi) - 1);
        mask &= ~((UINT64_C(1) << plo) - 1);
        acc |= win & mask;
    }
    return acc;
}

// the same limb for slots of at least 64 bits: the window [64 j, 64 j + 64)
// meets slot i0 = floor(64 j / s) and at most slot i0 + 1, so both are read
// with loads that do not wait on each other (the second at a clamped index
// and used only when that slot starts inside the window)
FA_DEV uint64_t ks_spack_limb_wide(const uint64_t* __restrict__ a, int w, uint64_t n,
                                   const int32_t* __restrict__ low, const uint8_t* __restrict__ br,
                                   int negs, uint64_t s, uint64_t j)
{
    const uint64_t bit0 = 64 * j;
    uint64_t i0 = (uint64_t)((double)bit0 / (double)s);
    if (i0 * s > bit0) i0--;
    if (i0 >= n) return 0;
    const uint64_t i1 = i0 + 1 < n ? i0 + 1 : i0;
    const uint64_t rel = bit0 - s * i0;   // slot bit at bit0, in [0, s)
    const int q = (int)(rel >> 6), r = (int)(rel & 63);
    const uint64_t* c0 = a + i0 * (uint64_t)w;
    const uint64_t* c1 = a + i1 * (uint64_t)w;
    const int lo0 = low[i0], b0 = br[i0], lo1 = low[i1], b1 = br[i1];
    uint64_t win = fa_ks_vlimb(c0, w, q, negs, b0, lo0) >> r;
    if (r) win |= fa_ks_vlimb(c0, w, q + 1, negs, b0, lo0) << (64 - r);
    const uint64_t next = fa_ks_vlimb(c1, w, 0, negs, b1, lo1);
    // window positions below s - rel belong to slot i0, the rest (when
    // s - rel < 64) to slot i0 + 1, whose bit 0 sits at position s - rel
    const uint64_t phi = s - rel;
    if (phi >= 64) return win;
    uint64_t acc = win & ((UINT64_C(1) << phi) - 1);
    if (i0 + 1 < n) acc |= next << phi;
    return acc;
}

// packed limbs j0 + 256 k + threadIdx.x (k < K) below L: all computed
// before any is stored, so their loads can be in flight together.  A batch
// takes K = FA_KS_PACK_BLK / 256; one product alone takes K = 1, which
// leaves more waves per CU to hide the loads of a small grid.
template <int K>
FA_DEV void ks_spack_block(const uint64_t* __restrict__ a, int w, uint64_t n,
                           const int32_t* __restrict__ low, const uint8_t* __restrict__ br,
                           int negs, uint64_t s, uint64_t* __restrict__ out, uint64_t L,
                           uint64_t j0)
{
    uint64_t v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const uint64_t j = j0 + 256 * (uint64_t)k + threadIdx.x;
        v[k] = 0;
        if (j < L)
            v[k] = s >= 64 ? ks_spack_limb_wide(a, w, n, low, br, negs, s, j)
                           : ks_spack_limb(a, w, n, low, br, negs, s, j);
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const uint64_t j = j0 + 256 * (uint64_t)k + threadIdx.x;
        if (j < L) out[j] = v[k];
    }
}

__global__ void __launch_bounds__(256) fa_k_ks_spack(const uint64_t* __restrict__ a, int w,
                                                     uint64_t n, const int32_t* __restrict__ low,
                                                     const uint8_t* __restrict__ br,
                                                     const int32_t* __restrict__ sig, uint64_t s,
                                                     uint64_t* __restrict__ out, uint64_t L)
{
    ks_spack_block<1>(a, w, n, low, br, *sig < 0, s, out, L, (uint64_t)blockIdx.x * 256);
}

// generate and propagate of slot k of the product: the slot's top bit, and
// whether its other s - 1 bits are all ones with the top bit clear
FA_DEV void fa_ks_slot_gp(const uint64_t* __restrict__ C, uint64_t nP, uint64_t s, uint64_t k,
                          uint8_t* g, uint8_t* p)
{
    const uint64_t tb = s * k + s - 1;
    const uint64_t q = tb >> 6;
    const uint8_t top = q < nP ? (uint8_t)((C[q] >> (tb & 63)) & 1) : 0;
    *g = top;
    uint8_t ones = !top;
    for (uint64_t lb = 0; ones && lb < s - 1; lb += 64) {
        const uint64_t nb = s - 1 - lb < 64 ? s - 1 - lb : 64;
        const uint64_t want = nb == 64 ? ~UINT64_C(0) : ((UINT64_C(1) << nb) - 1);
        const uint64_t v = fa_ks_slot(C, nP, s, k, (int)(lb >> 6));
        if ((v & want) != want) ones = 0;
    }
    *p = ones;
}

// one workgroup: t_k for the slots k < K of the product (the borrow each
// slot passes up), as bytes
FA_DEV void ks_uscan_body(const uint64_t* __restrict__ C, uint64_t nP, uint64_t s, uint64_t K,
                          uint8_t* __restrict__ tk, uint8_t* sg, uint8_t* sp)
{
    const int t = threadIdx.x;
    const uint64_t m = (K + 1023) / 1024;
    const uint64_t k0 = (uint64_t)t * m < K ? (uint64_t)t * m : K;
    const uint64_t k1 = k0 + m < K ? k0 + m : K;
    uint8_t G = 0, P = 1;
    for (uint64_t k = k0; k < k1; k++) {
        uint8_t g, p;
        fa_ks_slot_gp(C, nP, s, k, &g, &p);
        tk[k] = (uint8_t)(g | (p << 1));
        G = (uint8_t)(g | (p & G));
        P = (uint8_t)(P & p);
    }
    sg[t] = G;
    sp[t] = P;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        uint8_t g = sg[t], q = sp[t];
        if (t >= d) {
            g = (uint8_t)(g | (q & sg[t - d]));
            q = (uint8_t)(q & sp[t - d]);
        }
        __syncthreads();
        sg[t] = g;
        sp[t] = q;
        __syncthreads();
    }
    uint8_t cin = t ? sg[t - 1] : 0;
    for (uint64_t k = k0; k < k1; k++) {
        const uint8_t f = tk[k];
        cin = (uint8_t)((f & 1) | ((f >> 1) & cin));
        tk[k] = cin;
    }
}

__global__ void __launch_bounds__(1024) fa_k_ks_uscan(const uint64_t* __restrict__ C,
                                                      uint64_t nP, uint64_t s, uint64_t K,
                                                      uint8_t* __restrict__ tk)
{
    __shared__ uint8_t sg[1024], sp[1024];
    ks_uscan_body(C, nP, s, K, tk, sg, sp);
}

// limbs of each coefficient staged through LDS per step of the unpack
#define FA_KS_UC 16
// the reciprocal divisions of the unpack are exact for divisors below 64
static_assert(FA_KS_UC + 1 < 64, "unpack staging width");

// z[t] = coefficient zl + t: the s-bit field e_k + t_(k-1) read as signed,
// times sigma_a sigma_b, as zw limbs; zero past the product's length zn.
// One workgroup of 256 threads takes t = t0 .. t0 + 255 (those below cnt),
// one coefficient per thread, FA_KS_UC output limbs at a time: the slot's
// limbs are read into LDS and the output limbs written from it as runs of
// consecutive limbs, so neither side accesses memory one limb per
// coefficient at a stride of zw limbs.  buf: 256 * (FA_KS_UC + 1) words.
FA_DEV void ks_unpack_tile(const uint64_t* __restrict__ C, uint64_t nP, uint64_t s,
                           const uint8_t* __restrict__ tk, const int32_t* __restrict__ sig,
                           uint64_t zl, uint64_t zn, uint64_t cnt, int zw,
                           uint64_t* __restrict__ z, uint64_t t0, uint64_t* __restrict__ buf)
{
    constexpr int W = FA_KS_UC + 1;
    const int c = threadIdx.x;
    const int nc = (int)(cnt - t0 < 256 ? cnt - t0 : 256);
    const uint64_t k = zl + t0 + (uint64_t)c;
    const bool mine = c < nc;
    const bool live = mine && k < zn;
    const int neg = (sig[0] < 0) != (sig[1] < 0);
    const int r = (int)((s * k) & 63);   // bit offset of the slot in its first limb
    uint64_t carry = (live && k) ? tk[k - 1] : 0;
    uint64_t fill = 0;
    uint64_t ncarry = 1;   // two's complement negation: ~v + 1
    for (int lc = 0; lc < zw; lc += FA_KS_UC) {
        const int nl = zw - lc < FA_KS_UC ? zw - lc : FA_KS_UC;
        // limbs lc .. lc + nl of each slot's span (limb lc + nl feeds the
        // shift of limb lc + nl - 1); only those that reach the field
        // i / d for i < 256 d and d <= FA_KS_UC + 1 as (i ceil(2^20 / d)) >> 20:
        // exact (error below 256 d / 2^20 < 1 / d for d < 64) and within 32 bits
        const uint32_t m1 = ((1u << 20) + (uint32_t)nl) / (uint32_t)(nl + 1);
        const uint32_t m0 = ((1u << 20) + (uint32_t)nl - 1) / (uint32_t)nl;
        for (int i = c; i < nc * (nl + 1); i += 256) {
            const int cc = (int)(((uint32_t)i * m1) >> 20);
            const int j = i - cc * (nl + 1);
            const uint64_t kk = zl + t0 + (uint64_t)cc;
            uint64_t v = 0;
            if (kk < zn && 64 * (uint64_t)(lc + j) < s + 64) {
                const uint64_t q = ((s * kk) >> 6) + (uint64_t)(lc + j);
                if (q < nP) v = C[q];
            }
            buf[cc * W + j] = v;
        }
        __syncthreads();
        if (mine) {
            uint64_t* row = buf + c * W;
            for (int j = 0; j < nl; j++) {
                const uint64_t lb = 64 * (uint64_t)(lc + j);
                uint64_t v;
                if (!live) {
                    v = 0;
                } else {
                    if (lb < s) {
                        uint64_t e = row[j] >> r;
                        if (r) e |= row[j + 1] << (64 - r);
                        if (lb + 64 > s) e &= (UINT64_C(1) << (s - lb)) - 1;
                        v = e + carry;
                        carry = (v < carry) ? 1 : 0;
                        if (lb + 64 >= s) {
                            const uint64_t nb = s - lb;   // 1..64 field bits in this limb
                            const uint64_t m = nb == 64 ? ~UINT64_C(0) : ((UINT64_C(1) << nb) - 1);
                            v &= m;
                            fill = ((v >> (nb - 1)) & 1) ? ~UINT64_C(0) : 0;
                            v |= fill & ~m;
                        }
                    } else {
                        v = fill;
                    }
                    if (neg) {
                        v = ~v + ncarry;
                        ncarry = (ncarry && v == 0) ? 1 : 0;
                    }
                }
                row[j] = v;   // limb j + 1 is still the slot's, read next
            }
        }
        __syncthreads();
        for (int i = c; i < nc * nl; i += 256) {
            const int cc = (int)(((uint32_t)i * m0) >> 20);
            const int j = i - cc * nl;
            z[(t0 + (uint64_t)cc) * (uint64_t)zw + (uint64_t)(lc + j)] = buf[cc * W + j];
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) fa_k_ks_unpack(const uint64_t* __restrict__ C,
                                                      uint64_t nP, uint64_t s,
                                                      const uint8_t* __restrict__ tk,
                                                      const int32_t* __restrict__ sig,
                                                      uint64_t zl, uint64_t cnt, uint64_t zn,
                                                      int zw, uint64_t* __restrict__ z)
{
    __shared__ uint64_t buf[256 * (FA_KS_UC + 1)];
    ks_unpack_tile(C, nP, s, tk, sig, zl, zn, cnt, zw, z, (uint64_t)blockIdx.x * 256, buf);
}

// ---------------------------------------------------------------------------
// batched signed-slot route: every KS product of a batch in one launch per
// step.  Flags and slot scans take one workgroup per operand or product;
// packing and unpacking take work items (record | operand << 31 | block of
// FA_KS_PACK_BLK packed limbs or 256 output coefficients << 32).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) fa_kb_ks_sflags(const fa_bks* __restrict__ R, unsigned nks)
{
    __shared__ int64_t stop[1024];
    __shared__ uint8_t sg[1024], sp[1024];
    if (blockIdx.x >= 2 * nks) return;
    const fa_bks& r = R[blockIdx.x >> 1];
    if (blockIdx.x & 1)
        ks_sflags_body(r.b, (int)r.bw, r.bn, r.lowb, r.brb, r.sig + 1, stop, sg, sp);
    else
        ks_sflags_body(r.a, (int)r.aw, r.an, r.lowa, r.bra, r.sig, stop, sg, sp);
}

__global__ void __launch_bounds__(256) fa_kb_ks_spack(const fa_bks* __restrict__ R,
                                                      const uint64_t* __restrict__ work,
                                                      unsigned nwork)
{
    if (blockIdx.x >= nwork) return;
    const uint64_t wi = work[blockIdx.x];
    const fa_bks& r = R[(unsigned)(wi & 0x7fffffffu)];
    const int op = (int)((wi >> 31) & 1);
    const uint64_t j0 = (wi >> 32) * FA_KS_PACK_BLK;
    if (op)
        ks_spack_block<FA_KS_PACK_BLK / 256>(r.b, (int)r.bw, r.bn, r.lowb, r.brb, r.sig[1] < 0, r.s, r.pb, r.Lb, j0);
    else
        ks_spack_block<FA_KS_PACK_BLK / 256>(r.a, (int)r.aw, r.an, r.lowa, r.bra, r.sig[0] < 0, r.s, r.pa, r.La, j0);
}

__global__ void __launch_bounds__(1024) fa_kb_ks_uscan(const fa_bks* __restrict__ R, unsigned nks)
{
    __shared__ uint8_t sg[1024], sp[1024];
    if (blockIdx.x >= nks) return;
    const fa_bks& r = R[blockIdx.x];
    if (r.top > 0) ks_uscan_body(r.pr, r.nP, r.s, r.top, r.tk, sg, sp);
}

__global__ void __launch_bounds__(256) fa_kb_ks_unpack(const fa_bks* __restrict__ R,
                                                       const uint64_t* __restrict__ work,
                                                       unsigned nwork)
{
    __shared__ uint64_t buf[256 * (FA_KS_UC + 1)];
    if (blockIdx.x >= nwork) return;
    const uint64_t wi = work[blockIdx.x];
    const fa_bks& r = R[(unsigned)(wi & 0x7fffffffu)];
    const uint64_t t0 = (wi >> 32) * 256;
    if (t0 < r.cnt) ks_unpack_tile(r.pr, r.nP, r.s, r.tk, r.sig, r.zl, r.zn, r.cnt, (int)r.zw, r.z, t0, buf);
}

int fa_batch_ks_pre(const fa_bks* R, unsigned nks, const uint64_t* work, unsigned nwork,
                    hipStream_t s)
{
    if (!nks) return FA_OK;
    FA_LAUNCH("fa_kb_ks_sflags", fa_kb_ks_sflags, dim3(2 * nks), dim3(1024), 0, s, R, nks);
    if (nwork)
        FA_LAUNCH("fa_kb_ks_spack", fa_kb_ks_spack, dim3(nwork), dim3(256), 0, s, R, work, nwork);
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "batch ks pack launch");
}

int fa_batch_ks_post(const fa_bks* R, unsigned nks, const uint64_t* work, unsigned nwork,
                     hipStream_t s)
{
    if (!nks) return FA_OK;
    FA_LAUNCH("fa_kb_ks_uscan", fa_kb_ks_uscan, dim3(nks), dim3(1024), 0, s, R, nks);
    if (nwork)
        FA_LAUNCH("fa_kb_ks_unpack", fa_kb_ks_unpack, dim3(nwork), dim3(256), 0, s, R, work, nwork);
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "batch ks unpack launch");
}

static uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
static int bitcount(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

// slot width for |b1|, |b2| bits and lengths an, bn
uint64_t fa_ks_slot_bits(int64_t b1, int64_t b2, uint64_t an, uint64_t bn)
{
    const uint64_t m = an < bn ? an : bn;
    return (uint64_t)(b1 < 0 ? -b1 : b1) + (uint64_t)(b2 < 0 ? -b2 : b2) + bitcount(m) + 1;
}

// 1: signed slots and one integer product (default); 0: the three products
// of non-negative parts
int fa_ks_variant = 1;
extern "C" FA_API void fa_set_ks_variant(int v) { fa_ks_variant = v ? 1 : 0; }
int fa_ks_current_variant() { return fa_ks_variant; }
int fa_ks_nprod(int variant) { return variant ? 1 : 3; }

// scratch layout of one KS product: the packed operands, the products, the
// low-limb indices; for signed slots also the borrows, the product's slot
// borrows and the two signs
struct ks_layout {
    uint64_t La, Lb, nP;
    uint64_t* pa[3];
    uint64_t* pb[3];
    uint64_t* pr[3];
    int32_t* lowa;
    int32_t* lowb;
    uint8_t* bra;
    uint8_t* brb;
    uint8_t* tk;
    int32_t* sig;
};

static ks_layout ks_lay(int variant, uint64_t sb, uint64_t an, uint64_t bn, uint64_t* scratch)
{
    ks_layout K;
    memset(&K, 0, sizeof(K));
    const int np = fa_ks_nprod(variant);
    K.La = cdiv(an * sb, 64);
    K.Lb = cdiv(bn * sb, 64);
    K.nP = K.La + K.Lb;
    uint64_t* w = scratch;
    for (int k = 0; k < np; k++) { K.pa[k] = w; w += K.La; }
    for (int k = 0; k < np; k++) { K.pb[k] = w; w += K.Lb; }
    for (int k = 0; k < np; k++) { K.pr[k] = w; w += K.nP; }
    K.lowa = (int32_t*)w;
    K.lowb = K.lowa + an;
    if (variant) {
        K.sig = K.lowb + bn;                  // two int32
        K.bra = (uint8_t*)(K.sig + 2);
        K.brb = K.bra + an;
        K.tk = K.brb + bn;                    // an + bn slots
    }
    return K;
}

// scratch words of a KS product
uint64_t fa_ks_scratch_words_v(int variant, int64_t b1, int64_t b2, uint64_t an, uint64_t bn)
{
    const uint64_t sb = fa_ks_slot_bits(b1, b2, an, bn);
    const uint64_t La = cdiv(an * sb, 64), Lb = cdiv(bn * sb, 64);
    const uint64_t np = (uint64_t)fa_ks_nprod(variant);
    uint64_t w = np * (La + Lb) + np * (La + Lb) + cdiv(an + bn, 2) + 8;
    if (variant) w += 1 + cdiv(2 * (an + bn), 8);
    return w;
}

uint64_t fa_ks_scratch_words(int64_t b1, int64_t b2, uint64_t an, uint64_t bn)
{
    return fa_ks_scratch_words_v(fa_ks_variant, b1, b2, an, bn);
}

// the integer products of a KS product: ptrs[3k .. 3k+2] = {a, b, z} device
// pointers of product k, lens = {La, Lb}; returns the number of products
int fa_ks_products_v(int variant, uint64_t sb, uint64_t an, uint64_t bn, uint64_t* scratch,
                     uint64_t* ptrs[9], uint64_t lens[2])
{
    ks_layout K = ks_lay(variant, sb, an, bn, scratch);
    const int np = fa_ks_nprod(variant);
    for (int k = 0; k < np; k++) {
        ptrs[3 * k] = K.pa[k];
        ptrs[3 * k + 1] = K.pb[k];
        ptrs[3 * k + 2] = K.pr[k];
    }
    lens[0] = K.La;
    lens[1] = K.Lb;
    return np;
}

// the device record of a signed-slot KS product (scratch of
// fa_ks_scratch_words_v(1, ...)); operands and output as the caller's
void fa_ks_record(uint64_t sb, const uint64_t* a, uint64_t aw, uint64_t an, const uint64_t* b,
                  uint64_t bw, uint64_t bn, uint64_t* scratch, uint64_t* z, uint64_t zw,
                  uint64_t zl, uint64_t cnt, fa_bks* r)
{
    ks_layout K = ks_lay(1, sb, an, bn, scratch);
    memset(r, 0, sizeof(*r));
    r->a = a;
    r->b = b;
    r->z = z;
    r->aw = aw;
    r->an = an;
    r->bw = bw;
    r->bn = bn;
    r->s = sb;
    r->La = K.La;
    r->Lb = K.Lb;
    r->nP = K.nP;
    r->zw = zw;
    r->zl = zl;
    r->cnt = cnt;
    r->zn = an + bn - 1;
    r->top = zl + cnt < r->zn ? zl + cnt : r->zn;
    r->lowa = K.lowa;
    r->lowb = K.lowb;
    r->bra = K.bra;
    r->brb = K.brb;
    r->tk = K.tk;
    r->sig = K.sig;
    r->pa = K.pa[0];
    r->pb = K.pb[0];
    r->pr = K.pr[0];
}

// packing before the products
int fa_ks_pre_v(int variant, const uint64_t* a, uint64_t aw, uint64_t an, const uint64_t* b,
                uint64_t bw, uint64_t bn, uint64_t sb, uint64_t* scratch, hipStream_t s)
{
    ks_layout K = ks_lay(variant, sb, an, bn, scratch);
    if (variant) {
        FA_LAUNCH("fa_k_ks_sflags", fa_k_ks_sflags, dim3(1), dim3(1024), 0, s, a, (int)aw, an,
                  K.lowa, K.bra, K.sig);
        FA_LAUNCH("fa_k_ks_sflags", fa_k_ks_sflags, dim3(1), dim3(1024), 0, s, b, (int)bw, bn,
                  K.lowb, K.brb, K.sig + 1);
        FA_LAUNCH("fa_k_ks_spack", fa_k_ks_spack, dim3((unsigned)cdiv(K.La, 256)), dim3(256), 0, s,
                  a, (int)aw, an, K.lowa, K.bra, K.sig, sb, K.pa[0], K.La);
        FA_LAUNCH("fa_k_ks_spack", fa_k_ks_spack, dim3((unsigned)cdiv(K.Lb, 256)), dim3(256), 0, s,
                  b, (int)bw, bn, K.lowb, K.brb, K.sig + 1, sb, K.pb[0], K.Lb);
        return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "ks pack launch");
    }
    FA_LAUNCH("fa_k_ks_low", fa_k_ks_low, dim3((unsigned)cdiv(an, 256)), dim3(256), 0, s, a,
              (int)aw, an, K.lowa);
    FA_LAUNCH("fa_k_ks_low", fa_k_ks_low, dim3((unsigned)cdiv(bn, 256)), dim3(256), 0, s, b,
              (int)bw, bn, K.lowb);
    const int modes[3] = {FA_KS_PLUS, FA_KS_MINUS, FA_KS_ABS};
    for (int k = 0; k < 3; k++) {
        FA_LAUNCH("fa_k_ks_pack", fa_k_ks_pack, dim3((unsigned)cdiv(K.La, 256)), dim3(256), 0, s,
                  a, (int)aw, an, K.lowa, sb, modes[k], K.pa[k], K.La);
        FA_LAUNCH("fa_k_ks_pack", fa_k_ks_pack, dim3((unsigned)cdiv(K.Lb, 256)), dim3(256), 0, s,
                  b, (int)bw, bn, K.lowb, sb, modes[k], K.pb[k], K.Lb);
    }
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "ks pack launch");
}

// recombination after the products
int fa_ks_post_v(int variant, uint64_t sb, uint64_t an, uint64_t bn, uint64_t* scratch,
                 uint64_t* z, uint64_t zw, uint64_t zl, uint64_t cnt, hipStream_t s)
{
    ks_layout K = ks_lay(variant, sb, an, bn, scratch);
    const uint64_t zn = an + bn - 1;
    if (variant) {
        const uint64_t top = zl + cnt < zn ? zl + cnt : zn;   // slots the window reads up to
        if (top > 0)
            FA_LAUNCH("fa_k_ks_uscan", fa_k_ks_uscan, dim3(1), dim3(1024), 0, s, K.pr[0], K.nP, sb,
                      top, K.tk);
        FA_LAUNCH("fa_k_ks_unpack", fa_k_ks_unpack, dim3((unsigned)cdiv(cnt, 256)), dim3(256), 0, s,
                  K.pr[0], K.nP, sb, K.tk, K.sig, zl, cnt, zn, (int)zw, z);
        return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "ks unpack launch");
    }
    FA_LAUNCH("fa_k_ks_combine", fa_k_ks_combine, dim3((unsigned)cdiv(cnt, 256)), dim3(256), 0, s,
              K.pr[0], K.pr[1], K.pr[2], K.nP, sb, zl, cnt, zn, (int)zw, z);
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "ks combine launch");
}

// device form: a (an x aw), b (bn x bw) resident; scratch of
// fa_ks_scratch_words; enqueued on s
int fa_ks_device(uint64_t* z, uint64_t zw, uint64_t zl, uint64_t cnt, const uint64_t* a,
                 uint64_t aw, uint64_t an, const uint64_t* b, uint64_t bw, uint64_t bn,
                 int64_t b1, int64_t b2, uint64_t* scratch, hipStream_t s)
{
    const int variant = fa_ks_variant;
    const uint64_t sb = fa_ks_slot_bits(b1, b2, an, bn);
    int rc = fa_ks_pre_v(variant, a, aw, an, b, bw, bn, sb, scratch, s);
    if (rc) return rc;
    uint64_t* ptrs[9];
    uint64_t lens[2];
    const int np = fa_ks_products_v(variant, sb, an, bn, scratch, ptrs, lens);
    for (int k = 0; k < np; k++) {
        rc = fa_mpn_mul_device(ptrs[3 * k + 2], ptrs[3 * k], (int64_t)lens[0], ptrs[3 * k + 1],
                               (int64_t)lens[1], s);
        if (rc) return rc;
    }
    return fa_ks_post_v(variant, sb, an, bn, scratch, z, zw, zl, cnt, s);
}

// host-pointer form: the product the reference's caller computes when the
// eight-prime path declines (fmpz_poly/mul.c:160); same output layout as
// fa_fmpz_poly_mul_mid, always computed (returns 1)
extern "C" int fa_fmpz_poly_mul_ks(uint64_t* z, uint64_t zw, int64_t zl, int64_t zh,
                                   const uint64_t* a, uint64_t aw, int64_t an,
                                   const uint64_t* b, uint64_t bw, int64_t bn)
{
    if (an <= 0 || bn <= 0 || !a || !b || aw == 0 || bw == 0 || zw == 0)
        return fa_set_error(FA_E_ARG, "fa_fmpz_poly_mul_ks: bad arguments");
    if (zl < 0 || zl >= zh) return 1;
    if (!z) return fa_set_error(FA_E_ARG, "fa_fmpz_poly_mul_ks: null output");
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    const int64_t b1 = fa_vec_max_bits(a, aw, an), b2 = fa_vec_max_bits(b, bw, bn);
    const uint64_t cnt = (uint64_t)(zh - zl);
    const uint64_t wa = (uint64_t)an * aw, wb = (uint64_t)bn * bw, wz = cnt * zw;
    const uint64_t ws = fa_ks_scratch_words(b1, b2, (uint64_t)an, (uint64_t)bn);
    // operands, output and KS scratch in the thread's grow-only staging
    // buffer (the integer product inside uses the separate scratch buffer)
    void* hb = nullptr;
    int rc0 = fa_hbuf_reserve(st, (wa + wb + wz + ws) * 8, &hb, st->stream);
    if (rc0) return rc0;
    uint64_t* d = (uint64_t*)hb;
    uint64_t *da = d, *db = d + wa, *dz = db + wb, *dw = dz + wz;
    hipStream_t s = st->stream;
    int rc = FA_OK;
    if (hipMemcpyAsync(da, a, wa * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(db, b, wb * 8, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = fa_set_error(FA_E_HIP, "fa_fmpz_poly_mul_ks: upload");
    if (!rc)
        rc = fa_ks_device(dz, zw, (uint64_t)zl, cnt, da, aw, (uint64_t)an, db, bw, (uint64_t)bn,
                          b1, b2, dw, s);
    if (!rc && hipMemcpyAsync(z, dz, wz * 8, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = fa_set_error(FA_E_HIP, "fa_fmpz_poly_mul_ks: download");
    if (hipStreamSynchronize(s) != hipSuccess && !rc)
        rc = fa_set_error(FA_E_HIP, "fa_fmpz_poly_mul_ks: %s", hipGetErrorString(hipGetLastError()));
    return rc ? rc : 1;
}
