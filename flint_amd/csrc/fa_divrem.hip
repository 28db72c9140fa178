// fa_divrem.hip -- polynomial division with remainder mod n through the
// device products (nmod_poly_mul.c:1130-1330): the inverse of rev(b) as a
// power series by Newton iteration, the quotient as a middle product, the
// remainder as a - b q on the low bn - 1 coefficients.
//
// Division with remainder by a b whose leading coefficient is invertible
// mod n has a unique result, so (q, r) equal the reference's.
#include <string.h>

#include <vector>

#include "fa_internal.h"

extern "C" int fa_nmod_poly_mul_mid_device(uint64_t* z, int64_t zl, int64_t zh,
                                           const uint64_t* a, int64_t an, const uint64_t* b,
                                           int64_t bn, fa_nmod_t mod, void* stream);

typedef unsigned __int128 u128;

namespace {

// (a - b) mod n, elementwise: r[i] = (a[i] - b[i]) mod n, or -b[i] when a is null
__global__ void __launch_bounds__(256) k_sub_mod(uint64_t* __restrict__ r,
                                                 const uint64_t* __restrict__ a,
                                                 const uint64_t* __restrict__ b, uint64_t len,
                                                 uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const uint64_t x = a ? a[i] : 0, y = b[i];
    r[i] = x >= y ? x - y : x + (n - y);
}

// r[i] = a[len - 1 - i]
__global__ void __launch_bounds__(256) k_reverse(uint64_t* __restrict__ r,
                                                 const uint64_t* __restrict__ a, uint64_t len)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) r[i] = a[len - 1 - i];
}

unsigned blocks(uint64_t len) { return (unsigned)((len + 255) / 256); }

// inverse of x mod n by the extended Euclidean algorithm; 0 when none
uint64_t invmod(uint64_t x, uint64_t n)
{
    if (n == 1) return 0;
    __int128 t = 0, nt = 1;
    uint64_t r = n, nr = x % n;
    while (nr) {
        const uint64_t q = r / nr;
        const __int128 tt = t - (__int128)q * nt;
        t = nt;
        nt = tt;
        const uint64_t rr = r - q * nr;
        r = nr;
        nr = rr;
    }
    if (r != 1) return 0;
    if (t < 0) t += n;
    return (uint64_t)t;
}

struct dev_buf {
    uint64_t* p = nullptr;
    ~dev_buf()
    {
        if (p) (void)hipFree(p);
    }
    int alloc(uint64_t words)
    {
        if (hipMalloc((void**)&p, (words ? words : 1) * 8) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return fa_set_error(FA_E_NOMEM, "divrem: device allocation failed");
        }
        return FA_OK;
    }
};

// g[0, m) = f^-1 mod x^m on the device (f[0] invertible mod n), Newton:
// with g f = 1 mod x^k, e = (f g)[k, 2k) and g[k, 2k) = -(g e) mod x^k
int inv_series_dev(uint64_t* g, const uint64_t* f, uint64_t fn, uint64_t m, fa_nmod_t mod,
                   uint64_t* e, uint64_t* t, hipStream_t s)
{
    uint64_t f0;
    if (hipMemcpyAsync(&f0, f, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "divrem: reading the constant term failed");
    const uint64_t g0 = invmod(f0, mod.n);
    if (!g0 && mod.n != 1)
        return fa_set_error(FA_E_ARG, "divrem: leading coefficient %llu not invertible mod %llu",
                            (unsigned long long)f0, (unsigned long long)mod.n);
    if (hipMemcpyAsync(g, &g0, 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "divrem: upload failed");
    for (uint64_t k = 1; k < m;) {
        const uint64_t k2 = 2 * k < m ? 2 * k : m, h = k2 - k;
        const uint64_t fl = fn < k2 ? fn : k2;
        int rc = fa_nmod_poly_mul_mid_device(e, (int64_t)k, (int64_t)k2, f, (int64_t)fl, g,
                                             (int64_t)k, mod, s);
        if (rc) return rc;
        rc = fa_nmod_poly_mul_mid_device(t, 0, (int64_t)h, g, (int64_t)k, e, (int64_t)h, mod, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_sub_mod, dim3(blocks(h)), dim3(256), 0, s, g + k, nullptr, t, h,
                           mod.n);
        k = k2;
    }
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "divrem launch");
}

// B = rev(rev(b)^-1 mod x^Bn), Bn coefficients (nmod_poly_mul.c:1246-1248)
int precision_inverse(uint64_t* B, const uint64_t* db, uint64_t bn, uint64_t Bn, fa_nmod_t mod,
                      hipStream_t s)
{
    dev_buf t, g, e, w;
    int rc;
    if ((rc = t.alloc(bn)) || (rc = g.alloc(Bn)) || (rc = e.alloc(Bn)) || (rc = w.alloc(Bn)))
        return rc;
    hipLaunchKernelGGL(k_reverse, dim3(blocks(bn)), dim3(256), 0, s, t.p, db, bn);
    rc = inv_series_dev(g.p, t.p, bn, Bn, mod, e.p, w.p, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_reverse, dim3(blocks(Bn)), dim3(256), 0, s, B, g.p, Bn);
    if (hipStreamSynchronize(s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "divrem: synchronize failed");
    return FA_OK;
}

// q, r (device) from a (device, an), b (device, bn), B (device, Bn >= qn)
int divrem_dev(uint64_t* q, uint64_t* r, const uint64_t* a, uint64_t an, const uint64_t* b,
               uint64_t bn, const uint64_t* B, uint64_t Bn, fa_nmod_t mod, uint64_t* p,
               hipStream_t s)
{
    const uint64_t qn = an - bn + 1;
    // q = coefficients [Bn - 1, Bn - 1 + qn) of a_top * B (nmod_poly_mul.c:1250)
    int rc = fa_nmod_poly_mul_mid_device(q, (int64_t)Bn - 1, (int64_t)(Bn - 1 + qn),
                                         a + an - qn, (int64_t)qn, B, (int64_t)Bn, mod, s);
    if (rc) return rc;
    if (bn < 2) return FA_OK;
    // r = a - b q on the low bn - 1 coefficients
    rc = fa_nmod_poly_mul_mid_device(p, 0, (int64_t)bn - 1, b, (int64_t)bn, q, (int64_t)qn, mod,
                                     s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sub_mod, dim3(blocks(bn - 1)), dim3(256), 0, s, r, a, p, bn - 1, mod.n);
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "divrem launch");
}

}  // namespace

// _nmod_poly_inv_series on host arrays: g[0, m) = f^-1 mod x^m
extern "C" int fa_nmod_poly_inv_series(uint64_t* g, const uint64_t* f, int64_t fn, int64_t m,
                                       fa_nmod_t mod)
{
    if (!g || !f || fn <= 0 || m <= 0 || mod.n == 0)
        return fa_set_error(FA_E_ARG, "fa_nmod_poly_inv_series: bad arguments");
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = st->stream;
    dev_buf df, dg, e, t;
    int rc;
    if ((rc = df.alloc((uint64_t)fn)) || (rc = dg.alloc((uint64_t)m)) ||
        (rc = e.alloc((uint64_t)m)) || (rc = t.alloc((uint64_t)m)))
        return rc;
    if (hipMemcpyAsync(df.p, f, (uint64_t)fn * 8, hipMemcpyHostToDevice, s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "upload failed");
    rc = inv_series_dev(dg.p, df.p, (uint64_t)fn, (uint64_t)m, mod, e.p, t.p, s);
    if (rc) return rc;
    if (hipMemcpyAsync(g, dg.p, (uint64_t)m * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "download failed");
    return FA_OK;
}

// _nmod_poly_divrem_mpn_ctx (nmod_poly_mul.c:1234-1262): q[0, an - bn + 1),
// r[0, bn - 1) with a = b q + r, an >= bn >= 2, lead(b) invertible mod n
extern "C" int fa_nmod_poly_divrem(uint64_t* q, uint64_t* r, const uint64_t* a, int64_t an,
                                   const uint64_t* b, int64_t bn, fa_nmod_t mod)
{
    if (!q || !r || !a || !b || bn < 2 || an < bn || mod.n == 0)
        return fa_set_error(FA_E_ARG, "fa_nmod_poly_divrem: need an >= bn >= 2");
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = st->stream;
    const uint64_t qn = (uint64_t)(an - bn + 1), Bn = qn;
    dev_buf da, db, dB, dq, dr, dp;
    int rc;
    if ((rc = da.alloc((uint64_t)an)) || (rc = db.alloc((uint64_t)bn)) || (rc = dB.alloc(Bn)) ||
        (rc = dq.alloc(qn)) || (rc = dr.alloc((uint64_t)bn)) || (rc = dp.alloc((uint64_t)bn)))
        return rc;
    if (hipMemcpyAsync(da.p, a, (uint64_t)an * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(db.p, b, (uint64_t)bn * 8, hipMemcpyHostToDevice, s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "upload failed");
    rc = precision_inverse(dB.p, db.p, (uint64_t)bn, Bn, mod, s);
    if (rc) return rc;
    rc = divrem_dev(dq.p, dr.p, da.p, (uint64_t)an, db.p, (uint64_t)bn, dB.p, Bn, mod, dp.p, s);
    if (rc) return rc;
    if (hipMemcpyAsync(q, dq.p, qn * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(r, dr.p, (uint64_t)(bn - 1) * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "download failed");
    return FA_OK;
}

// nmod_poly_divrem_precomp_struct (fft_small.h:800-812): the divisor and its
// precision-Bn inverse, kept on the device for repeated divisions
struct fa_divrem_precomp_s {
    fa_nmod_t mod;
    uint64_t bn, Bn;
    uint64_t* b;
    uint64_t* B;
};

extern "C" int fa_nmod_poly_divrem_precomp_init(fa_divrem_precomp* out, const uint64_t* b,
                                                int64_t bn, int64_t Bn, fa_nmod_t mod)
{
    if (!out || !b || bn < 2 || Bn < 1 || mod.n == 0)
        return fa_set_error(FA_E_ARG, "fa_nmod_poly_divrem_precomp_init: bad arguments");
    *out = nullptr;
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = st->stream;
    fa_divrem_precomp M = (fa_divrem_precomp)calloc(1, sizeof(fa_divrem_precomp_s));
    if (!M) return fa_set_error(FA_E_NOMEM, "host allocation failed");
    M->mod = mod;
    M->bn = (uint64_t)bn;
    M->Bn = (uint64_t)Bn;
    int rc = FA_OK;
    if (hipMalloc((void**)&M->b, (uint64_t)bn * 8) != hipSuccess ||
        hipMalloc((void**)&M->B, (uint64_t)Bn * 8) != hipSuccess) {
        (void)hipGetLastError();
        rc = fa_set_error(FA_E_NOMEM, "device allocation failed");
    } else if (hipMemcpyAsync(M->b, b, (uint64_t)bn * 8, hipMemcpyHostToDevice, s) !=
               hipSuccess) {
        rc = fa_set_error(FA_E_HIP, "upload failed");
    } else {
        rc = precision_inverse(M->B, M->b, (uint64_t)bn, (uint64_t)Bn, mod, s);
    }
    if (rc) {
        fa_nmod_poly_divrem_precomp_clear(M);
        return rc;
    }
    *out = M;
    return FA_OK;
}

extern "C" void fa_nmod_poly_divrem_precomp_clear(fa_divrem_precomp M)
{
    if (!M) return;
    (void)hipDeviceSynchronize();
    if (M->b) (void)hipFree(M->b);
    if (M->B) (void)hipFree(M->B);
    free(M);
}

// _nmod_poly_divrem_precomp (nmod_poly_mul.c:1292-1313): 1 on success, 0
// when the quotient is longer than the precomputed precision Bn
extern "C" int fa_nmod_poly_divrem_precomp(uint64_t* q, uint64_t* r, const uint64_t* a,
                                           int64_t an, fa_divrem_precomp M, fa_nmod_t mod)
{
    if (!M || !q || !r || !a || an < (int64_t)M->bn || mod.n != M->mod.n)
        return fa_set_error(FA_E_ARG, "fa_nmod_poly_divrem_precomp: bad arguments");
    const uint64_t qn = (uint64_t)an - M->bn + 1;
    if (qn > M->Bn) return 0;
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = st->stream;
    dev_buf da, dq, dr, dp;
    int rc;
    if ((rc = da.alloc((uint64_t)an)) || (rc = dq.alloc(qn)) || (rc = dr.alloc(M->bn)) ||
        (rc = dp.alloc(M->bn)))
        return rc;
    if (hipMemcpyAsync(da.p, a, (uint64_t)an * 8, hipMemcpyHostToDevice, s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "upload failed");
    // B has precision Bn >= qn: its top qn coefficients give the quotient
    // (rev(B) mod x^qn is the inverse to precision qn)
    rc = divrem_dev(dq.p, dr.p, da.p, (uint64_t)an, M->b, M->bn, M->B + (M->Bn - qn), qn, mod,
                    dp.p, s);
    if (rc) return rc;
    if (hipMemcpyAsync(q, dq.p, qn * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(r, dr.p, (M->bn - 1) * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "download failed");
    return 1;
}
