// fa_sdfft.hip -- the sd_fft_ctx plugin surface (fft_small.h:206-281):
// sd_fft_ctx_init_prime, sd_fft_trunc and sd_ifft_trunc on host data.
//
// The forward transform is computed in full (the first otrunc outputs of a
// full transform are the outputs of the truncated one); the inverse follows
// van der Hoeven's truncated inverse (sd_ifft.c:1418-1493), driven from the
// host as a sequence of device kernels: full inverse sub-transforms plus
// element-wise butterflies on the ranges the recursion touches.  Outputs
// are written in sd_fft_ctx_trunc_index order (fft_small.h:247-253) and as
// canonical representatives in [0, p).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "fa_internal.h"
#include "fa_ntt.h"

#define LG_BLK_SZ 8
#define BLK_SZ 256
#define FA_SD_MAX_SLOTS 64

namespace {
struct slot_t {
    int used;
    int device;
    fa_prime_host pr;
    int depth;
    double* w2;
    double* w2inv;
};
std::mutex slot_mutex;
slot_t slots[FA_SD_MAX_SLOTS];
}  // namespace

// mulmod_satisfies_bounds.c:120-150 (fft_small_mulmod_satisfies_bounds)
extern "C" int fa_mulmod_satisfies_bounds(uint64_t nn)
{
    double n = (double)nn;
    double ninv = 1.0 / n;
    double t1 = fabs(fma(n, ninv, -1.0));
    unsigned __int128 sq = (unsigned __int128)nn * nn;
    uint64_t hi = (uint64_t)(sq >> 64), lo = (uint64_t)sq;
    int n1bits = 64 - __builtin_clzll(nn);
    int n2bits = hi ? 128 - __builtin_clzll(hi) : 64 - __builtin_clzll(lo);
    int B = 53 - n1bits - 1;
    if (B < 2) return 0;
    double limit2 = 2 * n * t1 + ldexp(ninv, 1 + n2bits - 53 - 1) + 0.5 + ldexp(1.0, -(B + 1));
    B -= 1;
    double limit4 = 4 * n * t1 + ldexp(ninv, 2 + n2bits - 53 - 1) + 0.5 + ldexp(1.0, -(B + 1));
    return (limit2 < 0.99) && (limit4 < 1.49);
}

// ------------------------------------------------------------------------
// generic sub-transform kernels (root index R0)
// ------------------------------------------------------------------------

// blocks of 2^k <= 2^12 points; block b has root R0 + b
__global__ void __launch_bounds__(256) fa_k_gen_rows(double* __restrict__ x, int k, uint64_t R0,
                                                     int inverse, const double* __restrict__ w2,
                                                     const double* __restrict__ w2inv, double n,
                                                     double ninv)
{
    extern __shared__ double lds[];
    const int N = 1 << k;
    double* blk = x + (uint64_t)blockIdx.x * N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) lds[i] = blk[i];
    __syncthreads();
    if (inverse)
        fa_lds_inv(lds, k, 0, R0 + blockIdx.x, w2inv, n, ninv);
    else
        fa_lds_fwd(lds, k, 0, R0 + blockIdx.x, w2, n, ninv);
    for (int i = threadIdx.x; i < N; i += blockDim.x) blk[i] = lds[i];
}

// the first d1 levels of a 2^(d1+d2)-point sub-transform with root R0, on
// tiles of 2^lgC columns
__global__ void __launch_bounds__(256) fa_k_gen_cols(double* __restrict__ x, int d1, int d2,
                                                     int lgC, uint64_t R0, int inverse,
                                                     const double* __restrict__ w2,
                                                     const double* __restrict__ w2inv, double n,
                                                     double ninv)
{
    extern __shared__ double lds[];
    const int C = 1 << lgC;
    const uint64_t col0 = (uint64_t)blockIdx.x << lgC;
    const int total = (1 << d1) << lgC;
    for (int u = threadIdx.x; u < total; u += blockDim.x) {
        uint64_t row = (uint64_t)(u >> lgC), c = (uint64_t)(u & (C - 1));
        lds[u] = x[(row << d2) + col0 + c];
    }
    __syncthreads();
    if (inverse)
        fa_lds_inv(lds, d1, lgC, R0, w2inv, n, ninv);
    else
        fa_lds_fwd(lds, d1, lgC, R0, w2, n, ninv);
    for (int u = threadIdx.x; u < total; u += blockDim.x) {
        uint64_t row = (uint64_t)(u >> lgC), c = (uint64_t)(u & (C - 1));
        x[(row << d2) + col0 + c] = lds[u];
    }
}

// element-wise steps of the truncated inverse on x[0, 2h), i in [i0, i1)
enum {
    EW_REDUCE = 0,        // x[i] = x[i] mod n in [-n/2, n/2]
    EW_RIGHT_INPUT = 1,   // x[i+h] = x[i] - r*(i+h<z ? x[i+h] : 0)
    EW_COMBINE = 2,       // (x[i], x[i+h]) = (x[i]+x[i+h], (x[i]-x[i+h]) r^-1)
    EW_LO_FROM_L = 3,     // x[i] = 2 x[i] - r*shi, shi = i+h<z ? save[i-i0] : 0;
                          // x[i+h] = save[i-i0] unless i == keep
    EW_LEFT_INPUT = 4,    // x[i] = (x[i] + r*(i+h<z ? x[i+h] : 0)) / 2
    EW_LO_FROM_L2 = 5,    // x[i] = 2 x[i] - r*(i+h<z ? x[i+h] : 0)
    EW_RESTORE = 6,       // x[i] = save[i-i0] unless i == keep
    EW_SAVE = 7,          // save[i-i0] = x[i]
    EW_CANON = 8          // x[i] = x[i] mod n in [0, n)
};

__global__ void fa_k_ew(double* __restrict__ x, double* __restrict__ save, int op, uint64_t i0,
                        uint64_t i1, uint64_t h, uint64_t z, uint64_t keep, uint64_t q,
                        const double* __restrict__ w2, const double* __restrict__ w2inv,
                        double n, double ninv)
{
    uint64_t i = i0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= i1) return;
    double r = w2[q], rinv = w2inv[q];
    switch (op) {
    case EW_REDUCE:
        x[i] = fa_reduce_pm1n(x[i], n, ninv);
        break;
    case EW_RIGHT_INPUT: {
        double shi = (i + h < z) ? x[i + h] : 0.0;
        x[i + h] = fa_reduce_pm1n(x[i] - fa_mulmod(r, shi, n, ninv), n, ninv);
        break;
    }
    case EW_COMBINE: {
        double SL = x[i], SR = x[i + h];
        x[i] = fa_reduce_pm1n(SL + SR, n, ninv);
        x[i + h] = fa_mulmod(SL - SR, rinv, n, ninv);
        break;
    }
    case EW_LO_FROM_L: {
        double sv = save[i - i0];
        double shi = (i + h < z) ? sv : 0.0;
        x[i] = fa_reduce_pm1n(2.0 * x[i] - fa_mulmod(r, shi, n, ninv), n, ninv);
        if (i != keep) x[i + h] = sv;
        break;
    }
    case EW_LEFT_INPUT: {
        double shi = (i + h < z) ? x[i + h] : 0.0;
        double half = 0.5 * (n + 1.0) - n;   // (n+1)/2 as a signed representative
        double t = fa_reduce_pm1n(x[i] + fa_mulmod(r, shi, n, ninv), n, ninv);
        x[i] = fa_mulmod(t, half, n, ninv);
        break;
    }
    case EW_LO_FROM_L2: {
        double shi = (i + h < z) ? x[i + h] : 0.0;
        x[i] = fa_reduce_pm1n(2.0 * x[i] - fa_mulmod(r, shi, n, ninv), n, ninv);
        break;
    }
    case EW_RESTORE:
        if (i != keep) x[i] = save[i - i0];
        break;
    case EW_SAVE:
        save[i - i0] = x[i];
        break;
    case EW_CANON: {
        double v = fa_reduce_pm1n(x[i], n, ninv);
        x[i] = v < 0.0 ? v + n : v;
        break;
    }
    }
}

// ------------------------------------------------------------------------
// host drivers
// ------------------------------------------------------------------------

namespace {
struct sd_run {
    const slot_t* S;
    hipStream_t s;
    double* save_base;
};

int ew(const sd_run& R, double* x, double* save, int op, uint64_t i0, uint64_t i1, uint64_t h,
       uint64_t z, uint64_t keep, uint64_t q)
{
    if (i1 <= i0) return FA_OK;
    uint64_t cnt = i1 - i0;
    FA_LAUNCH("fa_k_ew", fa_k_ew, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, R.s, x, save,
              op, i0, i1, h, z, keep, q, R.S->w2, R.S->w2inv, R.S->pr.p, R.S->pr.pinv);
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "ew launch failed");
}

// full forward (inverse) transform of 2^k points at x with root R0
int sub_transform(const sd_run& R, double* x, int k, uint64_t R0, int inverse)
{
    const slot_t* S = R.S;
    if (k == 0) return FA_OK;
    if (k <= FA_ROW_LG) {
        FA_LAUNCH("fa_k_gen_rows", fa_k_gen_rows, dim3(1), dim3(256), ((size_t)1 << k) * 8, R.s, x,
                  k, R0, inverse, S->w2, S->w2inv, S->pr.p, S->pr.pinv);
        return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "launch failed");
    }
    int d2 = FA_ROW_LG, d1 = k - d2;
    int lgC = 0;
    while (((1 << d1) << (lgC + 1)) <= FA_LDS_COL_DOUBLES && lgC + 1 <= d2) lgC++;
    size_t col_lds = ((size_t)1 << d1 << lgC) * sizeof(double);
    if (!inverse)
        FA_LAUNCH("fa_k_gen_cols", fa_k_gen_cols, dim3(1u << (d2 - lgC)), dim3(256), col_lds, R.s, x,
                  d1, d2, lgC, R0, 0, S->w2, S->w2inv, S->pr.p, S->pr.pinv);
    FA_LAUNCH("fa_k_gen_rows", fa_k_gen_rows, dim3(1u << d1), dim3(256), ((size_t)1 << d2) * 8, R.s,
              x, d2, R0 << d1, inverse, S->w2, S->w2inv, S->pr.p, S->pr.pinv);
    if (inverse)
        FA_LAUNCH("fa_k_gen_cols", fa_k_gen_cols, dim3(1u << (d2 - lgC)), dim3(256), col_lds, R.s, x,
                  d1, d2, lgC, R0, 1, S->w2, S->w2inv, S->pr.p, S->pr.pinv);
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "launch failed");
}

// van der Hoeven truncated inverse on the block x[0, 2^k) with root q:
// x[0,n) outputs, x[n,z) inputs scaled by 2^k, zero inputs beyond z; with f
// also the output at position n.  save is scratch of at least 2^k doubles.
int itft(const sd_run& R, double* x, int k, uint64_t q, uint64_t z, uint64_t n, int f,
         double* save)
{
    if (k == 0 || (n == 0 && !f)) return FA_OK;
    uint64_t N = UINT64_C(1) << k, h = N >> 1;
    int rc;
    if (n == N) return sub_transform(R, x, k, q, 1);
    if (n >= h) {
        uint64_t m = n - h;
        rc = ew(R, x + h, save, EW_SAVE, m, h, 0, 0, 0, q);
        if (!rc) rc = sub_transform(R, x, k - 1, 2 * q, 1);
        if (!rc) rc = ew(R, x, NULL, EW_RIGHT_INPUT, m, h, h, z, 0, q);
        if (!rc) rc = itft(R, x + h, k - 1, 2 * q + 1, h, m, f, save + (h - m));
        if (!rc) rc = ew(R, x, NULL, EW_COMBINE, 0, m, h, z, 0, q);
        // save[] was taken from x + h starting at m: index i - m
        if (!rc) rc = ew(R, x, save - 0, EW_LO_FROM_L, m, h, h, z, f ? m : ~UINT64_C(0), q);
        return rc;
    }
    uint64_t zl = z < h ? z : h;
    rc = ew(R, x, save, EW_SAVE, n, zl, 0, 0, 0, q);
    if (!rc) rc = ew(R, x, NULL, EW_LEFT_INPUT, n, zl, h, z, 0, q);
    if (!rc) rc = itft(R, x, k - 1, 2 * q, zl, n, f, save + (zl - n));
    if (!rc) rc = ew(R, x, NULL, EW_LO_FROM_L2, 0, n, h, z, 0, q);
    if (!rc) rc = ew(R, x, save, EW_RESTORE, n, zl, 0, 0, f ? n : ~UINT64_C(0), q);
    return rc;
}

uint64_t trunc_index(uint64_t L, uint64_t i)
{
    if (L >= 4) i = (i & ~UINT64_C(15)) | ((i >> 2) & 3) | ((i & 3) << 2);
    return i;
}

void permute16(double* d, uint64_t L, uint64_t len)
{
    if (L < 4) return;
    for (uint64_t b = 0; b + 16 <= len; b += 16) {
        double t[16];
        for (uint64_t i = 0; i < 16; i++) t[trunc_index(L, i)] = d[b + i];
        memcpy(d + b, t, sizeof(t));
    }
}

int slot_fit_depth(slot_t* S, int L)
{
    std::lock_guard<std::mutex> g(slot_mutex);
    if (S->depth >= L && S->w2) return FA_OK;
    int want = L < 12 ? 12 : L;
    if (want > S->pr.v) want = S->pr.v;
    double *w2, *w2i;
    int rc = fa_build_prime_tables(&S->pr, want, &w2, &w2i);
    if (rc) return rc;
    // the old tables may still be read by launches in flight from other
    // threads: retire them instead of freeing
    fa_retire_device_table(S->w2);
    fa_retire_device_table(S->w2inv);
    S->w2 = w2;
    S->w2inv = w2i;
    S->depth = want;
    return FA_OK;
}

int prepare(fa_sd_fft_ctx_struct* Q, uint64_t L, sd_run* R, double** dbuf)
{
    if (!Q || Q->slot < 0 || Q->slot >= FA_SD_MAX_SLOTS || !slots[Q->slot].used)
        return fa_set_error(FA_E_ARG, "uninitialized sd_fft_ctx");
    slot_t* S = &slots[Q->slot];
    if ((int)L > S->pr.v || L > FA_MAX_DEPTH)
        return fa_set_error(FA_E_RANGE, "FFT prime %llu does not support depth %llu",
                            (unsigned long long)S->pr.n, (unsigned long long)L);
    int rc = slot_fit_depth(S, (int)L);
    if (rc) return rc;
    Q->depth = S->depth;
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    void* buf;
    rc = fa_scratch_reserve(st, (size_t)2 * ((size_t)1 << L) * sizeof(double) + 256, &buf, st->stream);
    if (rc) return rc;
    R->S = S;
    R->s = st->stream;
    *dbuf = (double*)buf;
    R->save_base = (double*)buf + ((size_t)1 << L);
    return FA_OK;
}
}  // namespace

extern "C" int fa_sd_fft_ctx_init_prime(fa_sd_fft_ctx_struct* Q, uint64_t pp)
{
    if (!Q) return fa_set_error(FA_E_ARG, "null ctx");
    if (!fa_mulmod_satisfies_bounds(pp))
        return fa_set_error(FA_E_ARG, "FFT prime %llu does not satisfy bounds for arithmetic",
                            (unsigned long long)pp);
    if (pp < 3 || !(pp & 1))
        return fa_set_error(FA_E_ARG, "Input %llu is either 2 or not a prime",
                            (unsigned long long)pp);
    if (__builtin_ctzll(pp - 1) < 4)
        return fa_set_error(FA_E_ARG, "Input %llu does not support FFT context initialization",
                            (unsigned long long)pp);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fa_set_error(FA_E_NODEV, "no HIP device");
    std::lock_guard<std::mutex> g(slot_mutex);
    int k;
    for (k = 0; k < FA_SD_MAX_SLOTS; k++)
        if (!slots[k].used) break;
    if (k == FA_SD_MAX_SLOTS) return fa_set_error(FA_E_NOMEM, "too many sd_fft contexts");
    slot_t* S = &slots[k];
    memset(S, 0, sizeof(*S));
    fa_prime_info(pp, &S->pr);
    S->used = 1;
    S->device = dev;
    Q->p = S->pr.p;
    Q->pinv = S->pr.pinv;
    Q->mod.n = pp;
    int norm = __builtin_clzll(pp);
    Q->mod.norm = (uint64_t)norm;
    unsigned __int128 d = (unsigned __int128)(pp << norm);
    Q->mod.ninv = (uint64_t)((~(unsigned __int128)0) / d - ((unsigned __int128)1 << 64));
    Q->primitive_2power_root = S->pr.g;
    Q->depth = 0;
    Q->slot = k;
    return FA_OK;
}

extern "C" void fa_sd_fft_ctx_clear(fa_sd_fft_ctx_struct* Q)
{
    if (!Q || Q->slot < 0 || Q->slot >= FA_SD_MAX_SLOTS) return;
    std::lock_guard<std::mutex> g(slot_mutex);
    slot_t* S = &slots[Q->slot];
    if (S->used) {
        if (S->w2) { (void)hipFree(S->w2); (void)hipFree(S->w2inv); }
        memset(S, 0, sizeof(*S));
    }
    Q->slot = -1;
}

// sd_fft_ctx_w (fft_small.h:308-312): w[j] = g^brev_v(j), as a signed double
extern "C" double fa_sd_fft_ctx_w(const fa_sd_fft_ctx_struct* Q, uint64_t j)
{
    if (!Q || Q->slot < 0 || Q->slot >= FA_SD_MAX_SLOTS) return 0.0;
    const slot_t* S = &slots[Q->slot];
    uint64_t r = 0, x = j;
    for (int i = 0; i < S->pr.v; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return fa_signed_rep(fa_powmod_u64(S->pr.g, r, S->pr.n), S->pr.n);
}

extern "C" int fa_sd_fft_trunc(fa_sd_fft_ctx_struct* Q, double* d, uint64_t L, uint64_t itrunc,
                               uint64_t otrunc)
{
    uint64_t N = UINT64_C(1) << L;
    if (!d || itrunc > N || otrunc > N) return fa_set_error(FA_E_ARG, "bad truncation");
    sd_run R;
    double* x;
    int rc = prepare(Q, L, &R, &x);
    if (rc) return rc;
    // sd_fft.c:1247-1261: inputs past itrunc (block-rounded above 2^8) are zero
    uint64_t it = (L > LG_BLK_SZ) ? ((itrunc + BLK_SZ - 1) / BLK_SZ) * BLK_SZ : itrunc;
    if (it > N) it = N;
    if (hipMemcpyAsync(x, d, itrunc * sizeof(double), hipMemcpyHostToDevice, R.s) != hipSuccess ||
        hipMemsetAsync(x + itrunc, 0, (N - itrunc) * sizeof(double), R.s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "upload failed");
    rc = ew(R, x, NULL, EW_REDUCE, 0, N, 0, 0, 0, 0);
    if (!rc) rc = sub_transform(R, x, (int)L, 0, 0);
    if (!rc) rc = ew(R, x, NULL, EW_CANON, 0, N, 0, 0, 0, 0);
    if (rc) return rc;
    if (hipMemcpyAsync(d, x, N * sizeof(double), hipMemcpyDeviceToHost, R.s) != hipSuccess ||
        hipStreamSynchronize(R.s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "download failed");
    (void)it;
    permute16(d, L, N);
    return FA_OK;
}

// a[i] = a[i] m b[i] mod n, or a[i]^2 m when b is null, canonical in [0, n)
__global__ void fa_k_sd_point(double* __restrict__ a, const double* __restrict__ b,
                              uint64_t npts, double m, double n, double ninv)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const double x = fa_mulmod(a[i], m, n, ninv);
    double v = fa_reduce_pm1n(fa_mulmod(x, b ? b[i] : a[i], n, ninv), n, ninv);
    a[i] = v < 0.0 ? v + n : v;
}

namespace {
// sd_fft_ctx_point_mul / _sqr (mpn_mul.c:1392-1452) on host arrays of
// 2^depth doubles: a = a m b (b null: a m a) pointwise, entries canonical
int sd_point(fa_sd_fft_ctx_struct* Q, double* a, const double* b, uint64_t m, uint64_t depth)
{
    if (!a || depth < 4 || depth > FA_MAX_DEPTH)
        return fa_set_error(FA_E_ARG, "sd_fft_ctx_point_mul: bad arguments");
    sd_run R;
    double* x;
    int rc = prepare(Q, depth, &R, &x);
    if (rc) return rc;
    const uint64_t N = UINT64_C(1) << depth;
    double* y = b ? R.save_base : nullptr;
    if (hipMemcpyAsync(x, a, N * sizeof(double), hipMemcpyHostToDevice, R.s) != hipSuccess ||
        (b && hipMemcpyAsync(y, b, N * sizeof(double), hipMemcpyHostToDevice, R.s) != hipSuccess))
        return fa_set_error(FA_E_HIP, "upload failed");
    const double n = R.S->pr.p;
    const double md = fa_signed_rep(m % R.S->pr.n, R.S->pr.n);
    FA_LAUNCH("fa_k_sd_point", fa_k_sd_point, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
              R.s, x, y, N, md, n, R.S->pr.pinv);
    if (hipGetLastError() != hipSuccess) return fa_set_error(FA_E_HIP, "point launch failed");
    if (hipMemcpyAsync(a, x, N * sizeof(double), hipMemcpyDeviceToHost, R.s) != hipSuccess ||
        hipStreamSynchronize(R.s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "download failed");
    return FA_OK;
}
}  // namespace

extern "C" int fa_sd_fft_ctx_point_mul(fa_sd_fft_ctx_struct* Q, double* a, const double* b,
                                       uint64_t m, uint64_t depth)
{
    if (!b) return fa_set_error(FA_E_ARG, "sd_fft_ctx_point_mul: bad arguments");
    return sd_point(Q, a, b, m, depth);
}

extern "C" int fa_sd_fft_ctx_point_sqr(fa_sd_fft_ctx_struct* Q, double* a, uint64_t m,
                                       uint64_t depth)
{
    return sd_point(Q, a, nullptr, m, depth);
}

extern "C" int fa_sd_ifft_trunc(fa_sd_fft_ctx_struct* Q, double* d, uint64_t L, uint64_t trunc)
{
    uint64_t N = UINT64_C(1) << L;
    if (!d || trunc > N) return fa_set_error(FA_E_ARG, "bad truncation");
    sd_run R;
    double* x;
    int rc = prepare(Q, L, &R, &x);
    if (rc) return rc;
    // sd_ifft.c:1510-1518: truncation in whole blocks above 2^8
    uint64_t t = (L > LG_BLK_SZ) ? ((trunc + BLK_SZ - 1) / BLK_SZ) * BLK_SZ : N;
    if (t > N) t = N;
    std::vector<double> tmp(d, d + t);
    permute16(tmp.data(), L, t);
    if (hipMemcpyAsync(x, tmp.data(), t * sizeof(double), hipMemcpyHostToDevice, R.s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "upload failed");
    if (t > 0) {
        rc = ew(R, x, NULL, EW_REDUCE, 0, t, 0, 0, 0, 0);
        if (!rc) rc = itft(R, x, (int)L, 0, t, t, 0, R.save_base);
        if (!rc) rc = ew(R, x, NULL, EW_CANON, 0, t, 0, 0, 0, 0);
        if (rc) return rc;
    }
    if (hipMemcpyAsync(tmp.data(), x, t * sizeof(double), hipMemcpyDeviceToHost, R.s) != hipSuccess ||
        hipStreamSynchronize(R.s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "download failed");
    memcpy(d, tmp.data(), t * sizeof(double));
    return FA_OK;
}
