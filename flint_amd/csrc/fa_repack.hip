// fa_repack.hip -- the small-modulus route of FLINT's nmod multiplication
// (nmod_poly/mullow_fft_small.c:24-429): for moduli n <= 23, pairs of
// coefficients are packed as a[2i] + a[2i+1] M (M = 2^16 or 2^17) into a
// polynomial of half the length over the FFT prime p = 1108307720798209,
// multiplied with that one prime (the modulus is an FFT prime, so no CRT),
// and unpacked: a packed product coefficient c = c0 + c1 M + c2 M^2 holds
// the even output 2i as c0 (plus c2 of the previous one) and the odd output
// 2i + 1 as c1.  The route applies when the coefficient bounds from the
// Cauchy-Schwarz norms of the even and odd parts keep c0, c1 < M and
// c2 M^2 < p (mullow_fft_small.c:360-410); the result is then the product
// mod n, the same words the plain route gives.
#include <math.h>
#include <string.h>

#include "fa_internal.h"

extern "C" int fa_nmod_poly_mul_mid_device(uint64_t* z, int64_t zl, int64_t zh,
                                           const uint64_t* a, int64_t an, const uint64_t* b,
                                           int64_t bn, fa_nmod_t mod, void* stream);

#define FA_REPACK_PRIME UINT64_C(1108307720798209)   // mullow_fft_small.c:43
#define FA_REPACK_MAX_N 23                           // mullow_fft_small.c:53

namespace {

// largest lengths at which packing was found possible, by modulus
// (mullow_fft_small.c:59-96): the norm test below is skipped beyond them
const int repack_limit[36] = {0,    264497, 256720, 77049, 36925, 21385, 14099, 9872, 7337,
                              5565, 4481,   3609,   3084,  2562,  2129,  1859,  1687, 1446,
                              1264, 1160,   1035,   942,   874,   725,   662,   626,  587,
                              562,  517,    490,    426,   410,   363,   342,   336,  321};

__global__ void __launch_bounds__(256) k_pack(uint64_t* __restrict__ o,
                                              const uint64_t* __restrict__ a, uint64_t an,
                                              int lgM)
{
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (2 * j >= an) return;
    o[j] = a[2 * j] | (2 * j + 1 < an ? a[2 * j + 1] << lgM : 0);
}

// output coefficient k in [zl, zh) from the packed product window
// p[i - pl], i in [pl, ph)
__global__ void __launch_bounds__(256) k_unpack(uint64_t* __restrict__ z,
                                                const uint64_t* __restrict__ p, int64_t zl,
                                                int64_t zh, int64_t pl, int64_t ph, int lgM,
                                                uint64_t n)
{
    const int64_t k = zl + (int64_t)((uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= zh) return;
    const int64_t i = k >> 1;
    const uint64_t mask = (UINT64_C(1) << lgM) - 1;
    uint64_t v;
    if (k & 1) {
        v = i < ph ? (p[i - pl] >> lgM) & mask : 0;
    } else {
        v = i < ph ? p[i - pl] & mask : 0;
        if (i >= 1 && i - 1 < ph) v += p[i - 1 - pl] >> (2 * lgM);
    }
    z[k - zl] = v % n;
}

// ||even part||, ||odd part|| bounds: sqrt of the sums of squares, plus one
void norms(const uint64_t* a, int64_t an, uint64_t* ev, uint64_t* od)
{
    uint64_t e = (an % 2) ? a[an - 1] * a[an - 1] : 0, o = 0;
    for (int64_t i = 0; i + 1 < an; i += 2) {
        e += a[i] * a[i];
        o += a[i + 1] * a[i + 1];
    }
    *ev = (uint64_t)sqrt((double)e) + 1;
    *od = (uint64_t)sqrt((double)o) + 1;
}

unsigned blocks(uint64_t len) { return (unsigned)((len + 255) / 256); }

}  // namespace

// _nmod_poly_mulmid_fft_small_repack on host arrays: 1 when the packed route
// computed z[0, zn - znlo) = coefficients [znlo, zn) of a*b mod n, 0 when its
// conditions do not hold (nothing written), negative on error
extern "C" int fa_nmod_poly_mulmid_repack(uint64_t* z, const uint64_t* a, int64_t an,
                                          const uint64_t* b, int64_t bn, int64_t znlo,
                                          int64_t zn, fa_nmod_t mod)
{
    const uint64_t n = mod.n;
    if (!z || !a || !b || an <= 0 || bn <= 0 || znlo < 0 || zn > an + bn - 1 || znlo >= zn)
        return 0;
    if (n < 1 || n > FA_REPACK_MAX_N) return 0;
    const int64_t lim = repack_limit[n];
    if (an > lim || bn > lim) return 0;
    const bool sqr = a == b && an == bn;
    uint64_t ae, ao, be, bo;
    norms(a, an, &ae, &ao);
    if (sqr) {
        be = ae;
        bo = ao;
    } else {
        norms(b, bn, &be, &bo);
    }
    const uint64_t c0 = ae * be, c1 = ae * bo + ao * be, c2 = ao * bo;
    int lgM = 0;
    for (int l = 16; l <= 17 && !lgM; l++) {
        const uint64_t M = UINT64_C(1) << l;
        if (c0 < M && c1 < M && c2 * M * M < FA_REPACK_PRIME) lgM = l;
    }
    if (!lgM) return 0;

    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = st->stream;
    const int64_t an2 = (an + 1) / 2, bn2 = (bn + 1) / 2;
    const int64_t zn2 = (zn + 1) / 2 < an2 + bn2 - 1 ? (zn + 1) / 2 : an2 + bn2 - 1;
    const int64_t znlo2 = znlo >= 1 ? (znlo - 1) / 2 : 0;
    const uint64_t cnt = (uint64_t)(zn - znlo), cnt2 = (uint64_t)(zn2 - znlo2);
    const uint64_t words = (uint64_t)an + (sqr ? 0 : (uint64_t)bn) + (uint64_t)an2 +
                           (sqr ? 0 : (uint64_t)bn2) + cnt2 + cnt;
    void* hb;
    int rc = fa_hbuf_reserve(st, words * 8, &hb, s);
    if (rc) return rc;
    uint64_t* da = (uint64_t*)hb;
    uint64_t* db = sqr ? da : da + an;
    uint64_t* pa = db + (sqr ? an : bn);
    uint64_t* pb = sqr ? pa : pa + an2;
    uint64_t* pz = pb + (sqr ? an2 : bn2);
    uint64_t* dz = pz + cnt2;
    if (hipMemcpyAsync(da, a, (uint64_t)an * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        (!sqr && hipMemcpyAsync(db, b, (uint64_t)bn * 8, hipMemcpyHostToDevice, s) != hipSuccess))
        return fa_set_error(FA_E_HIP, "repack: upload failed");
    hipLaunchKernelGGL(k_pack, dim3(blocks(an2)), dim3(256), 0, s, pa, da, (uint64_t)an, lgM);
    if (!sqr)
        hipLaunchKernelGGL(k_pack, dim3(blocks(bn2)), dim3(256), 0, s, pb, db, (uint64_t)bn, lgM);
    fa_nmod_t mp;
    mp.n = FA_REPACK_PRIME;
    mp.norm = 64 - 50;
    mp.ninv = 0;      // not read by the device path (it derives its own preinverse)
    rc = fa_nmod_poly_mul_mid_device(pz, znlo2, zn2, pa, an2, pb, bn2, mp, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_unpack, dim3(blocks(cnt)), dim3(256), 0, s, dz, pz, znlo, zn, znlo2, zn2,
                       lgM, n);
    if (hipGetLastError() != hipSuccess) return fa_set_error(FA_E_HIP, "repack launch");
    if (hipMemcpyAsync(z, dz, cnt * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "repack: download failed");
    return 1;
}

// _nmod_poly_mulmid_fft_small (mullow_fft_small.c:419-424): the packed route
// when it applies, the ordinary window product otherwise
extern "C" int fa_nmod_poly_mulmid_fft_small(uint64_t* z, const uint64_t* a, int64_t an,
                                             const uint64_t* b, int64_t bn, int64_t znlo,
                                             int64_t zn, fa_nmod_t mod)
{
    const int r = fa_nmod_poly_mulmid_repack(z, a, an, b, bn, znlo, zn, mod);
    if (r < 0) return r;
    if (r == 1) return FA_OK;
    return fa_nmod_poly_mul_mid(z, znlo, zn, a, an, b, bn, mod);
}
