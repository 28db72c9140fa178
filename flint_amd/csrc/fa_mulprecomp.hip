// fa_mulprecomp.hip -- nmod middle products against a fixed operand whose
// transform is kept (mul_precomp_struct, _mul_precomp_init and
// _nmod_poly_mul_mid_precomp, fft_small.h:796-822, nmod_poly_mul.c:
// 1052-1140): b is transformed once on a cyclic plan of 2^depth points
// (which holds up to 2^depth accumulated products), and each middle
// product transforms a only.  A window [zl, zh) of a b with zn = an + bn - 1
// coefficients is computed when the transform holds it: zn <= 2^depth, or
// the wraparound case of _fft_small_plan_fit_window (plan.c:93-122),
// zh <= 2^depth <= zn <= zl + 2^depth, where the product mod x^(2^depth) - 1
// equals the product on [zl, zh); operands longer than 2^depth are folded
// mod x^(2^depth) - 1 first (_mod_red, nmod_poly_mul.c:33-168).  Otherwise
// the call returns 0 and computes nothing, as the reference does.  (Below
// 256 points the reference's fit also accepts a linear window longer than
// the transform, whose coefficients alias; this library returns 0 there.)
#include <string.h>

#include "fa_internal.h"

struct fa_mul_precomp_s {
    fa_plan_struct P;
    fa_op_struct bfft;
    uint64_t bn, btrunc;
    fa_nmod_t mod;
};

namespace {

uint64_t clog2u(uint64_t x)
{
    uint64_t d = 0;
    while ((UINT64_C(1) << d) < x) d++;
    return d;
}

// the reference's _len_trunc (nmod_poly_mul.c:26-28)
uint64_t len_trunc(uint64_t x)
{
    const uint64_t d = clog2u(x);
    return d < 8 ? UINT64_C(1) << (d < 4 ? 4 : d) : (x + 255) / 256 * 256;
}

// r[i] = sum of x[i + k N] over k, mod n (i < N): x mod x^N - 1
__global__ void __launch_bounds__(256) k_fold_cyclic(uint64_t* __restrict__ r,
                                                     const uint64_t* __restrict__ x, uint64_t xn,
                                                     uint64_t N, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    uint64_t s = 0;
    for (uint64_t j = i; j < xn; j += N) {
        const uint64_t v = x[j];
        s = s >= n - v ? s - (n - v) : s + v;
    }
    r[i] = s;
}

uint64_t* align256(void* p) { return (uint64_t*)(((uintptr_t)p + 255) & ~(uintptr_t)255); }

// upload x (host, xn) to dst (device) and fold it to at most N coefficients
// into fold; returns the device operand and its length
const uint64_t* stage(const uint64_t* x, uint64_t xn, uint64_t N, uint64_t n, uint64_t* dst,
                      uint64_t* fold, uint64_t* len, hipStream_t s, int* rc)
{
    *rc = FA_OK;
    if (hipMemcpyAsync(dst, x, xn * 8, hipMemcpyHostToDevice, s) != hipSuccess) {
        *rc = fa_set_error(FA_E_HIP, "mul precomp: upload failed");
        return nullptr;
    }
    if (xn <= N) {
        *len = xn;
        return dst;
    }
    hipLaunchKernelGGL(k_fold_cyclic, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, fold,
                       dst, xn, N, n);
    *len = N;
    return fold;
}

}  // namespace

// _mul_precomp_init (nmod_poly_mul.c:1052-1079): b (host, bn) transformed
// on a cyclic plan of 2^depth points
extern "C" int fa_nmod_poly_mul_precomp_init(fa_mul_precomp* out, const uint64_t* b, int64_t bn,
                                             int64_t btrunc, int depth, fa_nmod_t mod)
{
    if (!out || !b || bn <= 0 || btrunc < 0 || depth < 4 || depth > FA_MAX_DEPTH || mod.n < 2)
        return fa_set_error(FA_E_ARG, "fa_nmod_poly_mul_precomp_init: bad arguments");
    *out = nullptr;
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    const uint64_t N = UINT64_C(1) << depth;
    const uint64_t modbits = 64 - (uint64_t)mod.norm;
    fa_mul_precomp M = (fa_mul_precomp)calloc(1, sizeof(fa_mul_precomp_s));
    if (!M) return fa_set_error(FA_E_NOMEM, "fa_nmod_poly_mul_precomp_init: out of memory");
    int rc = fa_plan_init_nmod_cyclic(&M->P, (uint64_t)depth, N, N, 2 * modbits, mod);
    if (rc == FA_OK) rc = fa_op_init(&M->bfft, &M->P);
    if (rc) {
        free(M);
        return rc;
    }
    M->bn = (uint64_t)bn;
    M->btrunc = len_trunc((uint64_t)(btrunc ? btrunc : 1));
    M->mod = mod;
    hipStream_t s = st->stream;
    void* tb;
    rc = fa_tbuf_reserve(st, ((uint64_t)bn + N) * 8 + 512, &tb, s);
    uint64_t len = 0;
    const uint64_t* src = nullptr;
    if (rc == FA_OK) {
        uint64_t* db = align256(tb);
        src = stage(b, (uint64_t)bn, N, mod.n, db, align256(db + bn), &len, s, &rc);
    }
    if (rc == FA_OK) rc = fa_fft_nmod(&M->bfft, src, (int64_t)len, mod, &M->P, s);
    if (rc == FA_OK && hipStreamSynchronize(s) != hipSuccess)
        rc = fa_set_error(FA_E_HIP, "fa_nmod_poly_mul_precomp_init: synchronize failed");
    if (rc == FA_OK) rc = fa_buffers_leave(st, s);
    if (rc) {
        fa_nmod_poly_mul_precomp_clear(M);
        return rc;
    }
    *out = M;
    return FA_OK;
}

extern "C" void fa_nmod_poly_mul_precomp_clear(fa_mul_precomp M)
{
    if (!M) return;
    fa_op_clear(&M->bfft);
    free(M);
}

// _nmod_poly_mul_mid_precomp (nmod_poly_mul.c:1082-1140): z[0, zh - zl) =
// coefficients [zl, zh) of a b (a host, an); 1 when computed, 0 when the
// window does not fit the transform, < 0 on error
extern "C" int fa_nmod_poly_mul_mid_precomp(uint64_t* z, uint64_t zl, uint64_t zh,
                                            const uint64_t* a, int64_t an, fa_mul_precomp M,
                                            fa_nmod_t mod)
{
    if (!M || !a || an <= 0 || (zh > zl && !z) || mod.n != M->mod.n)
        return fa_set_error(FA_E_ARG, "fa_nmod_poly_mul_mid_precomp: bad arguments");
    if (zl >= zh) return 1;
    const uint64_t zn = (uint64_t)an + M->bn - 1;
    if (zh > zn) {
        if (zl >= zn) {
            memset(z, 0, (zh - zl) * 8);
            return 1;
        }
        memset(z + (zn - zl), 0, (zh - zn) * 8);
        zh = zn;
    }
    const uint64_t N = UINT64_C(1) << M->P.depth;
    uint64_t atrunc = len_trunc((uint64_t)an);
    if (atrunc > N) atrunc = N;
    const bool wrap = atrunc <= N && zh <= N && N <= zn && zn <= zl + N;
    if (!wrap && zn > N) return 0;
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = st->stream;
    const size_t opb = fa_op_sizeof_data(&M->P);
    void* tb;
    int rc = fa_tbuf_reserve(st, opb + ((uint64_t)an + N + (zh - zl)) * 8 + 1024, &tb, s);
    if (rc) return rc;
    fa_op_struct X;
    uint64_t* da = align256((char*)tb + opb);
    uint64_t* fold = align256(da + an);
    uint64_t* dz = align256(fold + N);
    uint64_t len = 0;
    const uint64_t* src = stage(a, (uint64_t)an, N, mod.n, da, fold, &len, s, &rc);
    if (rc || (rc = fa_op_init_borrowed(&X, &M->P, tb)) ||
        (rc = fa_fft_nmod(&X, src, (int64_t)len, mod, &M->P, s)) ||
        (rc = fa_op_mul(&X, &X, &M->bfft, &M->P, s)) || (rc = fa_ifft(&X, &M->P, s)) ||
        (rc = fa_export_nmod_range(dz, &X, zl, zh, mod, &M->P, s))) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    if (hipMemcpyAsync(z, dz, (zh - zl) * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "fa_nmod_poly_mul_mid_precomp: download failed");
    rc = fa_buffers_leave(st, s);
    return rc ? rc : 1;
}
