// fa_ks.hip -- fmpz polynomial products beyond the eight-prime capacity, by
// Kronecker substitution onto the integer product (the route the
// reference's callers take when _fmpz_poly_mul_mid_default_mpn_ctx declines:
// fmpz_poly/mul.c:160, _fmpz_poly_mul_KS).
//
// Signed coefficients are packed without borrows: with a = a+ - a- (the
// positive part and the negated negative part, coefficient by coefficient)
//     a b = 2 (a+ b+ + a- b-) - |a| |b|,
// an identity of polynomials, so it holds coefficient by coefficient.  Each
// of the three products has non-negative coefficients below
// min(an, bn) 2^(|b1|+|b2|) < 2^(s-1) for the slot width
//     s = |b1| + |b2| + bitcount(min(an, bn)) + 1,
// so packing at s bits per coefficient puts every coefficient in its own
// slot with no carry between slots.  Packing is a gather per output limb;
// the combination is one thread per output coefficient.
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

static uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
static int bitcount(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

// slot width for |b1|, |b2| bits and lengths an, bn
uint64_t fa_ks_slot_bits(int64_t b1, int64_t b2, uint64_t an, uint64_t bn)
{
    const uint64_t m = an < bn ? an : bn;
    return (uint64_t)(b1 < 0 ? -b1 : b1) + (uint64_t)(b2 < 0 ? -b2 : b2) + bitcount(m) + 1;
}

// scratch layout of one KS product: the packed parts of a and b, the
// three products, the low-limb indices
struct ks_layout {
    uint64_t La, Lb, nP;
    uint64_t* pa[3];
    uint64_t* pb[3];
    uint64_t* pr[3];
    int32_t* lowa;
    int32_t* lowb;
};

static ks_layout ks_lay(uint64_t sb, uint64_t an, uint64_t bn, uint64_t* scratch)
{
    ks_layout K;
    K.La = cdiv(an * sb, 64);
    K.Lb = cdiv(bn * sb, 64);
    K.nP = K.La + K.Lb;
    uint64_t* w = scratch;
    for (int k = 0; k < 3; k++) { K.pa[k] = w; w += K.La; }
    for (int k = 0; k < 3; k++) { K.pb[k] = w; w += K.Lb; }
    for (int k = 0; k < 3; k++) { K.pr[k] = w; w += K.nP; }
    K.lowa = (int32_t*)w;
    K.lowb = K.lowa + an;
    return K;
}

// the three integer products of a KS product: out[k] = {a, b, z} device
// pointers and out_n = {La, Lb}
void fa_ks_products(uint64_t sb, uint64_t an, uint64_t bn, uint64_t* scratch,
                    uint64_t* ptrs[9], uint64_t lens[2])
{
    ks_layout K = ks_lay(sb, an, bn, scratch);
    for (int k = 0; k < 3; k++) {
        ptrs[3 * k] = K.pa[k];
        ptrs[3 * k + 1] = K.pb[k];
        ptrs[3 * k + 2] = K.pr[k];
    }
    lens[0] = K.La;
    lens[1] = K.Lb;
}

// packing before the products
int fa_ks_pre(const uint64_t* a, uint64_t aw, uint64_t an, const uint64_t* b, uint64_t bw,
              uint64_t bn, uint64_t sb, uint64_t* scratch, hipStream_t s)
{
    ks_layout K = ks_lay(sb, an, bn, scratch);
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
int fa_ks_post(uint64_t sb, uint64_t an, uint64_t bn, uint64_t* scratch, uint64_t* z,
               uint64_t zw, uint64_t zl, uint64_t cnt, hipStream_t s)
{
    ks_layout K = ks_lay(sb, an, bn, scratch);
    FA_LAUNCH("fa_k_ks_combine", fa_k_ks_combine, dim3((unsigned)cdiv(cnt, 256)), dim3(256), 0, s,
              K.pr[0], K.pr[1], K.pr[2], K.nP, sb, zl, cnt, an + bn - 1, (int)zw, z);
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "ks combine launch");
}

// device form: a (an x aw), b (bn x bw) resident; scratch holds the packed
// operands, the three products and the low-limb indices; enqueued on s
int fa_ks_device(uint64_t* z, uint64_t zw, uint64_t zl, uint64_t cnt, const uint64_t* a,
                 uint64_t aw, uint64_t an, const uint64_t* b, uint64_t bw, uint64_t bn,
                 int64_t b1, int64_t b2, uint64_t* scratch, hipStream_t s)
{
    const uint64_t sb = fa_ks_slot_bits(b1, b2, an, bn);
    int rc = fa_ks_pre(a, aw, an, b, bw, bn, sb, scratch, s);
    if (rc) return rc;
    uint64_t* ptrs[9];
    uint64_t lens[2];
    fa_ks_products(sb, an, bn, scratch, ptrs, lens);
    for (int k = 0; k < 3; k++) {
        rc = fa_mpn_mul_device(ptrs[3 * k + 2], ptrs[3 * k], (int64_t)lens[0], ptrs[3 * k + 1],
                               (int64_t)lens[1], s);
        if (rc) return rc;
    }
    return fa_ks_post(sb, an, bn, scratch, z, zw, zl, cnt, s);
}

// scratch words of fa_ks_device
uint64_t fa_ks_scratch_words(int64_t b1, int64_t b2, uint64_t an, uint64_t bn)
{
    const uint64_t sb = fa_ks_slot_bits(b1, b2, an, bn);
    const uint64_t La = cdiv(an * sb, 64), Lb = cdiv(bn * sb, 64);
    return 3 * La + 3 * Lb + 3 * (La + Lb) + cdiv(an + bn, 2) + 8;
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
    uint64_t* d = nullptr;
    if (hipMalloc((void**)&d, (wa + wb + wz + ws) * 8) != hipSuccess)
        return fa_set_error(FA_E_NOMEM, "fa_fmpz_poly_mul_ks: device memory");
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
    (void)hipFree(d);
    return rc ? rc : 1;
}
