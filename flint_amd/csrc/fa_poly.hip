// fa_poly.hip -- polynomial products over Z/nZ and Z on the convolution core.
//
//   fa_nmod_poly_mul_mid   restates _nmod_poly_mul_mid_mpn_ctx
//                          (nmod_poly_mul.c:904-972, plan.c:356-398)
//   fa_fmpz_poly_mul_mid   restates _fmpz_poly_mul_mid_mpn_ctx
//                          (fmpz_poly_mul.c:319-401)
//
// Both import coefficients into np residue rows, run the transform /
// pointwise / inverse core (fa_convolve) and recombine each output
// coefficient from its residues with the CRT of fa_crt_coeff; nmod then
// reduces the exact integer modulo n, fmpz emits it as a centred signed
// value.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <mutex>

#include "fa_internal.h"
#include "fa_ntt.h"

typedef unsigned __int128 u128;

// limbs that hold sum_i x_i (M/p_i) < np M with np <= 8: those of M plus
// one, at most FA_CRT_MAXL
__host__ __device__ inline int fa_crt_poly_limbs(const fa_crt_consts& C)
{
    return C.nl + 1 < FA_CRT_MAXL ? C.nl + 1 : FA_CRT_MAXL;
}

// exact integer coefficient from np residues (same recombination as the
// mpn path: x_i = c * cop_i^-1, c = sum x_i (M/p_i) - floor(sum x_i/p_i) M),
// on NL limbs: enough for sum x_i (M/p_i) < np M (fa_crt_poly_limbs), so
// the borrow after subtracting floor(.) M still tells a negative result
template <int NL>
FA_DEV void fa_crt_coeff_poly(const double* __restrict__ Q, uint64_t stride, uint64_t j,
                              const fa_crt_consts& C, uint64_t s[NL])
{
#pragma unroll
    for (int l = 0; l < NL; l++) s[l] = 0;
    double qe = 0.0;
    for (int i = 0; i < C.np; i++) {
        uint64_t x = fa_reduce_0n_u64(Q[(uint64_t)i * stride + fa_res_pos(j, stride)], C.p[i],
                                      C.pinv[i]);
        qe += (double)x * C.rp[i];
        uint64_t carry = 0;
#pragma unroll
        for (int l = 0; l < NL; l++) {
            uint64_t lo = x * C.cof[i][l];
            uint64_t hi = __umul64hi(x, C.cof[i][l]);
            uint64_t t = s[l] + lo;
            hi += (t < lo);
            t += carry;
            hi += (t < carry);
            s[l] = t;
            carry = hi;
        }
    }
    uint64_t q = (uint64_t)qe;
    uint64_t borrow = 0, carry = 0;
#pragma unroll
    for (int l = 0; l < NL; l++) {
        uint64_t lo = q * C.M[l];
        uint64_t hi = __umul64hi(q, C.M[l]);
        lo += carry;
        hi += (lo < carry);
        carry = hi;
        uint64_t t = s[l] - lo;
        uint64_t b1 = (s[l] < lo);
        uint64_t t2 = t - borrow;
        uint64_t b2 = (t < borrow);
        s[l] = t2;
        borrow = b1 | b2;
    }
    if (borrow) {
        uint64_t c2 = 0;
#pragma unroll
        for (int l = 0; l < NL; l++) {
            uint64_t t = s[l] + C.M[l];
            uint64_t cc = (t < s[l]);
            t += c2;
            cc |= (t < c2);
            s[l] = t;
            c2 = cc;
        }
    } else {
        int ge = 1;
#pragma unroll
        for (int l = NL - 1; l >= 0; l--)
            if (ge == 1 && s[l] != C.M[l]) ge = (s[l] > C.M[l]) ? 2 : 0;
        if (ge) {
            uint64_t br = 0;
#pragma unroll
            for (int l = 0; l < NL; l++) {
                uint64_t t = s[l] - C.M[l];
                uint64_t b1 = (s[l] < C.M[l]);
                uint64_t t2 = t - br;
                uint64_t b2 = (t < br);
                s[l] = t2;
                br = b1 | b2;
            }
        }
    }
}

// coefficient j modulo n from its NP residues, through the mixed-radix
// digits of c (Garner): t_0 = c mod p_0 and
//     t_k = (c - (t_0 + p_0 (t_1 + ... p_(k-2) t_(k-1)))) (p_0 ... p_(k-1))^-1 mod p_k,
// in FP64 mulmods, then c mod n by Horner's rule over the digits with
// 128-bit products.  c is the same integer in [0, M) as the limb-by-limb
// recombination gives, found without its rounded quotient: every step is
// exact.  Bounds: each mulmod sees |a b| < 2 p_k^2 (primes within a factor
// 2 of each other, the running value reduced to [-p_k/2, p_k/2] first), and
// v (p_k mod n) + t_k < n 2^64 keeps the high word below n.
// p_k mod n for the primes of C, computed once per workgroup into pm
// (threads 0..7; the caller puts a barrier after it)
FA_DEV void fa_crt_nmod_pm(const fa_crt_consts& C, uint64_t n, uint64_t dinv, int norm,
                           uint64_t* pm)
{
    const int k = threadIdx.x;
    if (k < FA_NPRIMES) pm[k] = k < C.np ? fa_ll_mod_preinv(0, (uint64_t)C.p[k], n, dinv, norm) : 0;
}

template <int NP>
FA_DEV uint64_t fa_crt_nmod_one(const double* __restrict__ Q, uint64_t stride, uint64_t j,
                                const fa_crt_consts& C, uint64_t n, uint64_t dinv, int norm,
                                const uint64_t* pm)
{
    const uint64_t pos = fa_res_pos(j, stride);
    double t[NP];
    {
        const double p = C.p[0], pinv = C.pinv[0];
        const double x = (double)fa_reduce_0n_u64(Q[pos], p, pinv);
        t[0] = (double)fa_reduce_0n_u64(fa_mulmod(x, C.gcm[0], p, pinv), p, pinv);
    }
#pragma unroll
    for (int k = 1; k < NP; k++) {
        const double p = C.p[k], pinv = C.pinv[k];
        const double x = (double)fa_reduce_0n_u64(Q[(uint64_t)k * stride + pos], p, pinv);
        const double r = fa_mulmod(x, C.gcm[k], p, pinv);   // c mod p_k, in (-p_k, p_k)
        double u = t[k - 1];
#pragma unroll
        for (int i = k - 2; i >= 0; i--)
            u = fa_reduce_pm1n(fa_mulmod(u, C.gpm[k][i], p, pinv) + t[i], p, pinv);
        u = fa_reduce_pm1n(u, p, pinv);
        t[k] = (double)fa_reduce_0n_u64(fa_mulmod(r - u, C.ginv[k], p, pinv), p, pinv);
    }
    uint64_t v = fa_ll_mod_preinv(0, (uint64_t)t[NP - 1], n, dinv, norm);
#pragma unroll
    for (int k = NP - 2; k >= 0; k--) {
        const uint64_t pk = pm[k];
        const uint64_t tk = (uint64_t)t[k];
        uint64_t lo = v * pk;
        uint64_t hi = __umul64hi(v, pk);
        lo += tk;
        hi += (lo < tk);
        v = fa_ll_mod_preinv(hi, lo, n, dinv, norm);
    }
    return v;
}

template <int NP>
__global__ void __launch_bounds__(256) fa_k_crt_nmod(const double* __restrict__ Q, uint64_t stride,
                                                     fa_crt_consts C, uint64_t zl, uint64_t cnt,
                                                     uint64_t n, uint64_t dinv, int norm,
                                                     uint64_t* __restrict__ z)
{
    __shared__ uint64_t pm[FA_NPRIMES];
    fa_crt_nmod_pm(C, n, dinv, norm, pm);
    __syncthreads();
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    z[t] = fa_crt_nmod_one<NP>(Q, stride, zl + t, C, n, dinv, norm, pm);
}

// NP = C.np (1 .. 8) as a template argument
#define FA_CRT_NP_SWITCH(np, F) \
    switch (np) {               \
    case 1: F(1); break;        \
    case 2: F(2); break;        \
    case 3: F(3); break;        \
    case 4: F(4); break;        \
    case 5: F(5); break;        \
    case 6: F(6); break;        \
    case 7: F(7); break;        \
    default: F(8); break;       \
    }

// centred lift (fmpz_poly_mul.c:209-227): s > M/2 -> s - M, written as a
// zw-limb two's complement value (limbs from NL on are the sign)
template <int NL>
FA_DEV void fa_emit_fmpz(uint64_t* s, const fa_crt_consts& C, uint64_t* o, uint64_t zw)
{
    int gt = 0;
    for (int l = NL - 1; l >= 0; l--) {
        uint64_t h = (C.M[l] >> 1) | ((l + 1 < FA_CRT_MAXL) ? (C.M[l + 1] << 63) : 0);
        if (s[l] != h) { gt = s[l] > h; break; }
    }
    uint64_t ext = 0;
    if (gt) {
        uint64_t br = 0;
#pragma unroll
        for (int l = 0; l < NL; l++) {
            uint64_t t1 = s[l] - C.M[l];
            uint64_t b1 = (s[l] < C.M[l]);
            uint64_t t2 = t1 - br;
            uint64_t b2 = (t1 < br);
            s[l] = t2;
            br = b1 | b2;
        }
        ext = ~UINT64_C(0);
    }
    for (uint64_t l = 0; l < zw; l++) {
        uint64_t v = ext;
#pragma unroll
        for (int k = 0; k < NL; k++)
            if ((uint64_t)k == l) v = s[k];
        o[l] = v;
    }
}

template <int NL>
FA_DEV void fa_crt_fmpz_one(const double* __restrict__ Q, uint64_t stride, uint64_t j,
                            const fa_crt_consts& C, uint64_t* o, uint64_t zw)
{
    uint64_t s[NL];
    fa_crt_coeff_poly<NL>(Q, stride, j, C, s);
    fa_emit_fmpz<NL>(s, C, o, zw);
}

template <int NL>
__global__ void __launch_bounds__(256) fa_k_crt_fmpz(const double* __restrict__ Q, uint64_t stride,
                                                     fa_crt_consts C, uint64_t zl, uint64_t cnt,
                                                     uint64_t zw, uint64_t* __restrict__ z)
{
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    fa_crt_fmpz_one<NL>(Q, stride, zl + t, C, z + t * zw, zw);
}

// NL = fa_crt_poly_limbs(C) (2 .. FA_CRT_MAXL) as a template argument
#define FA_CRT_NL_SWITCH(nl, F) \
    switch (nl) {               \
    case 2: F(2); break;        \
    case 3: F(3); break;        \
    case 4: F(4); break;        \
    case 5: F(5); break;        \
    case 6: F(6); break;        \
    case 7: F(7); break;        \
    default: F(8); break;       \
    }

// batched nmod / fmpz CRT: work item = job | chunk << 32, 256 outputs per
// chunk; outputs at or beyond the product length are zero
template <int FMPZ>
__global__ void __launch_bounds__(256) fa_kb_crt_poly(const fa_bpoly* __restrict__ P,
                                                      const uint64_t* __restrict__ work,
                                                      unsigned nwork)
{
    const unsigned nb = gridDim.x, bid = blockIdx.x;
    const unsigned L = (bid & 7) * (nb >> 3) + (bid >> 3);
    if (L >= nwork) return;
    const uint64_t w = work[L];
    const fa_bpoly& j = P[(unsigned)(w & 0xffffffffu)];
    __shared__ uint64_t pm[FA_NPRIMES];
    if (!FMPZ && j.Q) {         // the same job for the whole workgroup
        fa_crt_nmod_pm(*(const fa_crt_consts*)j.C, j.n, j.dinv, j.norm, pm);
        __syncthreads();
    }
    const uint64_t t = (w >> 32) * 256 + threadIdx.x;
    if (t >= j.cnt) return;
    const uint64_t idx = j.zl + t;
    if (FMPZ) {
        uint64_t* o = j.z + t * (uint64_t)j.zw;
        if (idx >= j.zn || !j.Q) {
            for (int l = 0; l < j.zw; l++) o[l] = 0;
            return;
        }
        const fa_crt_consts& C = *(const fa_crt_consts*)j.C;
#define FA_KB_FMPZ(NL) fa_crt_fmpz_one<NL>(j.Q, j.stride, idx, C, o, (uint64_t)j.zw)
        FA_CRT_NL_SWITCH(fa_crt_poly_limbs(C), FA_KB_FMPZ)
#undef FA_KB_FMPZ
    } else {
        if (idx >= j.zn || !j.Q) {
            j.z[t] = 0;
            return;
        }
        const fa_crt_consts& C = *(const fa_crt_consts*)j.C;
        uint64_t r = 0;
#define FA_KB_NMOD(NP) r = fa_crt_nmod_one<NP>(j.Q, j.stride, idx, C, j.n, j.dinv, j.norm, pm)
        FA_CRT_NP_SWITCH(C.np, FA_KB_NMOD)
#undef FA_KB_NMOD
        j.z[t] = r;
    }
}

int fa_batch_crt_poly(int fmpz, const fa_bpoly* P, const uint64_t* work, unsigned nwork,
                      hipStream_t s)
{
    if (!nwork) return FA_OK;
    const unsigned grid = (nwork + 7) & ~7u;
    if (fmpz)
        FA_LAUNCH("fa_kb_crt_fmpz", fa_kb_crt_poly<1>, dim3(grid), dim3(256), 0, s, P, work, nwork);
    else
        FA_LAUNCH("fa_kb_crt_nmod", fa_kb_crt_poly<0>, dim3(grid), dim3(256), 0, s, P, work, nwork);
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "batch crt launch");
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------

static int clog2_u64(uint64_t x) { return x <= 1 ? 0 : 64 - __builtin_clzll(x - 1); }
static int nbits_u64(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

// M_np >= c * 2^e (flint_mpn_cmp_ui_2exp as used by plan.c:28-32)
static int prod_ge(const fa_crt_consts& C, uint64_t c, uint64_t e)
{
    uint64_t t[FA_CRT_MAXL + 4];
    memset(t, 0, sizeof(t));
    uint64_t q = e / 64, r = e % 64;
    if (q + 2 > FA_CRT_MAXL + 3) return 0;
    t[q] = c << r;
    if (r) t[q + 1] = c >> (64 - r);
    for (int i = FA_CRT_MAXL + 3; i >= 0; i--) {
        uint64_t mi = (i < FA_CRT_MAXL) ? C.M[i] : 0;
        if (mi != t[i]) return mi > t[i];
    }
    return 1;
}

struct poly_job {
    int np, offset, depth;
    fa_crt_consts C;
};

// device buffers for one polynomial product: operands, residue rows, output
static int poly_buffers(fa_thread_state* st, uint64_t an_words, uint64_t bn_words, int np,
                        int depth, uint64_t zwords, uint64_t** da, uint64_t** db, double** A,
                        double** B, uint64_t** dz)
{
    size_t tr = (size_t)np * fa_res_stride(depth) * sizeof(double);
    size_t bytes = 2 * tr + (an_words + bn_words + zwords) * sizeof(uint64_t) + 256;
    void* buf;
    int rc = fa_scratch_reserve(st, bytes, &buf, st->stream);
    if (rc) return rc;
    *A = (double*)buf;
    *B = (double*)((char*)buf + tr);
    *da = (uint64_t*)((char*)buf + 2 * tr);
    *db = *da + an_words;
    *dz = *db + bn_words;
    return FA_OK;
}

// ------------------------------------------------------------------------
// direct transforms mod the modulus (plan.c:319-349): when n is itself an
// FFT prime whose 2-valuation covers the transform, the product is
// transformed mod n alone and needs no CRT (one transform set where the CRT
// needs two or more).  n's twiddle tables are built on first use and kept
// in a small per-device cache; a lease holds the cache lock from lookup to
// the last launch that reads the tables, and an eviction synchronizes the
// device before freeing them.  Not taken while the stream is captured (the
// tables could be evicted under a recorded graph).
// ------------------------------------------------------------------------
namespace {
struct direct_entry {
    int device = -1;
    uint64_t n = 0, used = 0;
    fa_tables T;
};
std::mutex direct_mu;
direct_entry direct_cache[4];
uint64_t direct_clock = 0;
std::atomic<uint64_t> direct_count{0};

struct direct_lease {
    std::unique_lock<std::mutex> lock;
    const fa_tables* T = nullptr;
    fa_crt_consts C;
};

int direct_enabled = 1;

bool direct_applies(uint64_t n, int depth, int crt_np)
{
    if (!direct_enabled || crt_np < 2 || n < (UINT64_C(1) << 20) || n > (UINT64_C(1) << 50) || !(n & 1)) return false;
    const int v = __builtin_ctzll(n - 1);
    const int tdepth = depth < 20 ? 20 : depth;
    if (v < tdepth) return false;
    return fa_mulmod_satisfies_bounds(n) && fa_is_prime_u64(n);
}

bool direct_acquire(uint64_t n, int depth, int crt_np, hipStream_t s, direct_lease* L)
{
    if (!direct_applies(n, depth, crt_np)) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
        return false;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    const int tdepth = depth < 20 ? 20 : depth;
    L->lock = std::unique_lock<std::mutex>(direct_mu);
    direct_entry* e = nullptr;
    for (auto& c : direct_cache)
        if (c.device == dev && c.n == n && c.T.depth >= tdepth) e = &c;
    if (!e) {
        // the least recently used slot (an empty one first)
        e = &direct_cache[0];
        for (auto& c : direct_cache)
            if (c.device < 0 || (e->device >= 0 && c.used < e->used)) e = &c;
        if (e->device >= 0) {
            int cur = dev;
            (void)hipSetDevice(e->device);
            (void)hipDeviceSynchronize();
            (void)hipFree(e->T.w2[0]);
            (void)hipFree(e->T.w2inv[0]);
            (void)hipSetDevice(cur);
            e->device = -1;
        }
        fa_tables T;
        memset(&T, 0, sizeof(T));
        if (fa_prime_info(n, &T.pr[0]) != FA_OK ||
            fa_build_prime_tables(&T.pr[0], tdepth, &T.w2[0], &T.w2inv[0]) != FA_OK) {
            (void)hipGetLastError();
            L->lock.unlock();
            return false;
        }
        T.device = dev;
        T.depth = tdepth;
        e->T = T;
        e->device = dev;
        e->n = n;
    }
    e->used = ++direct_clock;
    L->T = &e->T;
    memset(&L->C, 0, sizeof(L->C));
    L->C.np = 1;
    L->C.nl = 1;
    L->C.p[0] = (double)n;
    L->C.pinv[0] = 1.0 / (double)n;
    L->C.rp[0] = 1.0 / (double)n;
    L->C.M[0] = n;
    L->C.cof[0][0] = 1;
    L->C.gcm[0] = 1.0;
    L->C.ginv[0] = 1.0;
    direct_count++;
    return true;
}
}  // namespace

// products computed with direct transforms so far (a counter for tests)
extern "C" uint64_t fa_nmod_direct_count(void) { return direct_count.load(); }
// 0: always the CRT primes (comparisons of the two plans), 1: direct where it applies
extern "C" void fa_set_direct_fft(int on) { direct_enabled = on ? 1 : 0; }

#define HIPTRY(x, msg)                                                                  \
    do {                                                                                \
        if ((x) != hipSuccess) return fa_set_error(FA_E_HIP, "%s: %s", msg,             \
                                                   hipGetErrorString(hipGetLastError())); \
    } while (0)

// ------------------------------------------------------------------------
// polynomial products of more output coefficients than one 2^24-point
// transform holds: block products, each one's share of the window [zl, zh)
// added into the output on the device (mod n, or as zw-limb two's
// complement integers, coefficient by coefficient)
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(256) fa_k_poly_zero(uint64_t* __restrict__ z, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) z[i] = 0;
}

__global__ void __launch_bounds__(256) fa_k_addmod(uint64_t* __restrict__ z,
                                                   const uint64_t* __restrict__ t, uint64_t cnt,
                                                   uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= cnt) return;
    const uint64_t x = z[i];
    uint64_t s = x + t[i];
    if (s < x || s >= n) s -= n;   // both terms below n
    z[i] = s;
}

__global__ void __launch_bounds__(256) fa_k_add_coeffs(uint64_t* __restrict__ z,
                                                       const uint64_t* __restrict__ t,
                                                       uint64_t cnt, uint64_t zw)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= cnt) return;
    uint64_t* o = z + i * zw;
    const uint64_t* p = t + i * zw;
    uint64_t c = 0;
    for (uint64_t l = 0; l < zw; l++) {
        const uint64_t x = o[l], y = p[l];
        const uint64_t s1 = x + y;
        const uint64_t s2 = s1 + c;
        c = (uint64_t)(s1 < x) + (uint64_t)(s2 < s1);
        o[l] = s2;
    }
}

// output coefficients above which nmod/fmpz products go in blocks (tests
// force small blocks; 0 = only above one 2^24-point transform)
uint64_t fa_poly_block_cap = 0;
extern "C" FA_API void fa_set_poly_block_cap(uint64_t n) { fa_poly_block_cap = n; }

// transform depth of the window [zl, zh) of a product of zn coefficients
// whose operands have at most xmax coefficients (plan.c:47-89): the power of
// two 2^d = 2^floor(log2 zn) when the cyclic product mod X^(2^d) - 1 leaves
// the window untouched -- operands no longer than 2^d, zh <= 2^d and
// zn - 2^d <= zl, so the coefficients k + 2^d folded onto the window are
// zero -- otherwise the smallest 2^d >= zn.  At least 16 points.
int fa_poly_window_depth(uint64_t zl, uint64_t zh, uint64_t zn, uint64_t xmax)
{
    int d = 63 - __builtin_clzll(zn);
    const uint64_t i = UINT64_C(1) << d;
    if (!(xmax <= i && zh <= i && i <= zn && zn <= zl + i)) d = clog2_u64(zn);
    return d < 4 ? 4 : d;
}

static uint64_t poly_block_limit(uint64_t zn)
{
    const uint64_t full = UINT64_C(1) << FA_MAX_DEPTH;
    if (fa_poly_block_cap && zn > fa_poly_block_cap) return fa_poly_block_cap < full ? fa_poly_block_cap : full;
    return zn > full ? full : 0;   // 0: one product
}

// fmpz != 0: zw-limb signed coefficients (aw/bw limbs in), centred CRT;
// else nmod.  zl < zh <= an + bn - 1; np/offset and C for the whole product
static int poly_mul_blocked(int fmpz, uint64_t* z, uint64_t zw, int64_t zl, int64_t zh,
                            const uint64_t* a, uint64_t aw, int64_t an, const uint64_t* b,
                            uint64_t bw, int64_t bn, uint64_t n, int np, int offset,
                            const fa_crt_consts& C, uint64_t limit)
{
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = st->stream;
    // blocks: b whole beside blocks of a, or equal blocks, whichever needs
    // fewer block products; every block product has at most `limit` outputs
    const uint64_t A_ = (uint64_t)an, B_ = (uint64_t)bn;
    uint64_t la = B_ < limit ? limit - B_ + 1 : 0, lb = B_;
    if (la > A_) la = A_;
    uint64_t cost = la ? (A_ + la - 1) / la : UINT64_MAX;
    {
        const uint64_t h = (limit + 1) / 2;
        const uint64_t la2 = A_ < h ? A_ : h, lb2 = B_ < h ? B_ : h;
        const uint64_t c2 = ((A_ + la2 - 1) / la2) * ((B_ + lb2 - 1) / lb2);
        if (c2 < cost) { la = la2; lb = lb2; cost = c2; }
    }
    int dmax = clog2_u64(la + lb - 1);
    if (dmax < 4) dmax = 4;
    const uint64_t cnt = (uint64_t)(zh - zl);
    const uint64_t tcap = cnt < la + lb - 1 ? cnt : la + lb - 1;
    const uint64_t zwords = fmpz ? zw : 1;
    const size_t tr = (size_t)np * fa_res_stride(dmax) * sizeof(double);
    const size_t bytes = 2 * tr + (A_ * aw + B_ * bw + (cnt + tcap) * zwords) * 8 + 256;
    void* buf;
    int rc = fa_scratch_reserve(st, bytes, &buf, st->stream);
    if (rc) return rc;
    double* Ar = (double*)buf;
    double* Br = (double*)((char*)buf + tr);
    uint64_t* da = (uint64_t*)((char*)buf + 2 * tr);
    uint64_t* db = da + A_ * aw;
    uint64_t* dz = db + B_ * bw;
    uint64_t* dt = dz + cnt * zwords;
    HIPTRY(hipMemcpyAsync(da, a, A_ * aw * 8, hipMemcpyHostToDevice, s), "upload a");
    HIPTRY(hipMemcpyAsync(db, b, B_ * bw * 8, hipMemcpyHostToDevice, s), "upload b");
    FA_LAUNCH("fa_k_poly_zero", fa_k_poly_zero, dim3((unsigned)((cnt * zwords + 255) / 256)),
              dim3(256), 0, s, dz, cnt * zwords);
    int norm = 0;
    uint64_t dinv = 0;
    if (!fmpz) {
        norm = 64 - nbits_u64(n);
        const uint64_t d = n << norm;
        dinv = (uint64_t)(((u128)0 - 1) / d - ((u128)1 << 64));
    }
    for (uint64_t i = 0; i < A_; i += la) {
        const uint64_t na = A_ - i < la ? A_ - i : la;
        for (uint64_t j = 0; j < B_; j += lb) {
            const uint64_t nb = B_ - j < lb ? B_ - j : lb;
            const uint64_t off = i + j, blen = na + nb - 1;
            const uint64_t wlo = (uint64_t)zl > off ? (uint64_t)zl : off;
            const uint64_t whi = (uint64_t)zh < off + blen ? (uint64_t)zh : off + blen;
            if (wlo >= whi) continue;
            int depth = clog2_u64(blen);
            if (depth < 4) depth = 4;
            fa_tables* T = fa_get_tables(depth);
            if (!T) return FA_E_NODEV;
            fa_pack_desc pa = {fmpz ? FA_PACK_S64W : FA_PACK_U64, da + i * aw, na, 0, (int)aw};
            fa_pack_desc pb = {fmpz ? FA_PACK_S64W : FA_PACK_U64, db + j * bw, nb, 0, (int)bw};
            rc = fa_convolve(st, T, np, offset, depth, &pa, &pb, Ar, Br, s);
            if (rc) return rc;
            const uint64_t wc = whi - wlo;
            const unsigned g = (unsigned)((wc + 255) / 256);
            if (fmpz) {
#define FA_L(NL)                                                                            \
    FA_LAUNCH("fa_k_crt_fmpz", fa_k_crt_fmpz<NL>, dim3(g), dim3(256), 0, s, Ar,             \
              fa_res_stride(depth), C, wlo - off, wc, zw, dt)
                FA_CRT_NL_SWITCH(fa_crt_poly_limbs(C), FA_L)
#undef FA_L
                FA_LAUNCH("fa_k_add_coeffs", fa_k_add_coeffs, dim3(g), dim3(256), 0, s,
                          dz + (wlo - (uint64_t)zl) * zw, dt, wc, zw);
            } else {
#define FA_L(NP)                                                                            \
    FA_LAUNCH("fa_k_crt_nmod", fa_k_crt_nmod<NP>, dim3(g), dim3(256), 0, s, Ar,             \
              fa_res_stride(depth), C, wlo - off, wc, n, dinv, norm, dt)
                FA_CRT_NP_SWITCH(C.np, FA_L)
#undef FA_L
                FA_LAUNCH("fa_k_addmod", fa_k_addmod, dim3(g), dim3(256), 0, s,
                          dz + (wlo - (uint64_t)zl), dt, wc, n);
            }
            HIPTRY(hipGetLastError(), "block launch");
        }
    }
    HIPTRY(hipMemcpyAsync(z, dz, cnt * zwords * 8, hipMemcpyDeviceToHost, s), "download");
    HIPTRY(hipStreamSynchronize(s), "synchronize");
    return FA_OK;
}

extern "C" int fa_nmod_poly_mul_mid(uint64_t* z, int64_t zl, int64_t zh, const uint64_t* a,
                                    int64_t an, const uint64_t* b, int64_t bn, fa_nmod_t mod)
{
    if (an <= 0 || bn <= 0 || !a || !b || mod.n == 0)
        return fa_set_error(FA_E_ARG, "fa_nmod_poly_mul_mid: bad arguments");
    if (zl < 0 || zl >= zh) return FA_OK;
    if (!z) return fa_set_error(FA_E_ARG, "fa_nmod_poly_mul_mid: null output");
    int64_t zn = an + bn - 1;
    if (zh > zn) {
        int64_t s0 = zl > zn ? zl : zn;
        memset(z + (s0 - zl), 0, (size_t)(zh - s0) * sizeof(uint64_t));
        if (zl >= zn) return FA_OK;
        zh = zn;
    }
    uint64_t n = mod.n;
    if (n == 1) {
        memset(z, 0, (size_t)(zh - zl) * sizeof(uint64_t));
        return FA_OK;
    }
    // prime data from the tables of at most one transform's depth (a longer
    // product goes in blocks)
    const int tdepth = clog2_u64((uint64_t)zn) < FA_MAX_DEPTH ? clog2_u64((uint64_t)zn) : FA_MAX_DEPTH;
    fa_tables* T = fa_get_tables(tdepth);
    if (!T) return FA_E_NODEV;
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;

    int modbits = nbits_u64(n);
    poly_job J;
    int rc0 = fa_nmod_primes(T, (uint64_t)(an < bn ? an : bn), n, &J.np, &J.offset);
    if (rc0) return rc0;
    fa_crt_consts_init(&J.C, J.np, J.offset);
    J.depth = fa_poly_window_depth((uint64_t)zl, (uint64_t)zh, (uint64_t)zn,
                                   (uint64_t)(an > bn ? an : bn));
    if (const uint64_t lim = poly_block_limit((uint64_t)zn))
        return poly_mul_blocked(0, z, 1, zl, zh, a, 1, an, b, 1, bn, n, J.np, J.offset, J.C, lim);
    direct_lease D;
    if (direct_acquire(n, J.depth, J.np, st->stream, &D)) {
        T = const_cast<fa_tables*>(D.T);
        J.np = 1;
        J.offset = 0;
        J.C = D.C;
    }

    uint64_t cnt = (uint64_t)(zh - zl);
    uint64_t *da, *db, *dz;
    double *A, *B;
    int rc = poly_buffers(st, an, bn, J.np, J.depth, cnt, &da, &db, &A, &B, &dz);
    if (rc) return rc;
    hipStream_t s = st->stream;
    // a square keeps one device copy (nmod_poly_mul.c:939 squares when a == b)
    if (a == b && an == bn) db = da;
    HIPTRY(hipMemcpyAsync(da, a, an * sizeof(uint64_t), hipMemcpyHostToDevice, s), "upload a");
    if (db != da)
        HIPTRY(hipMemcpyAsync(db, b, bn * sizeof(uint64_t), hipMemcpyHostToDevice, s), "upload b");
    fa_pack_desc pa = {FA_PACK_U64, da, (uint64_t)an, 0, 1};
    fa_pack_desc pb = {FA_PACK_U64, db, (uint64_t)bn, 0, 1};
    rc = fa_convolve(st, T, J.np, J.offset, J.depth, &pa, &pb, A, B, s);
    if (rc) return rc;
    int norm = 64 - modbits;
    uint64_t d = n << norm;
    uint64_t dinv = (uint64_t)(((u128)0 - 1) / d - ((u128)1 << 64));
#define FA_L(NP)                                                                            \
    FA_LAUNCH("fa_k_crt_nmod", fa_k_crt_nmod<NP>, dim3((unsigned)((cnt + 255) / 256)),      \
              dim3(256), 0, s, A, fa_res_stride(J.depth), J.C, (uint64_t)zl, cnt, n, dinv,   \
              norm, dz)
    FA_CRT_NP_SWITCH(J.C.np, FA_L)
#undef FA_L
    HIPTRY(hipGetLastError(), "crt launch");
    if (D.lock.owns_lock()) D.lock.unlock();     // the tables' last reader is enqueued
    HIPTRY(hipMemcpyAsync(z, dz, cnt * sizeof(uint64_t), hipMemcpyDeviceToHost, s), "download");
    HIPTRY(hipStreamSynchronize(s), "synchronize");
    return FA_OK;
}

int fa_crt_nmod_launch(const double* A, uint64_t stride, const fa_crt_consts& C, uint64_t zl,
                       uint64_t cnt, uint64_t n, uint64_t* dz, hipStream_t s);

__global__ void __launch_bounds__(256) fa_k_fill_zero(uint64_t* __restrict__ z, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = 0;
}

// _nmod_poly_mul_mid on device pointers, enqueued on `stream` (NULL: the
// thread's): z[0, zh - zl) = coefficients [zl, zh) of a*b mod n, zeros past
// the product.  Products of more than one transform's length are refused.
extern "C" int fa_nmod_poly_mul_mid_device(uint64_t* z, int64_t zl, int64_t zh,
                                           const uint64_t* a, int64_t an, const uint64_t* b,
                                           int64_t bn, fa_nmod_t mod, void* stream)
{
    if (an <= 0 || bn <= 0 || !a || !b || mod.n == 0 || zl < 0)
        return fa_set_error(FA_E_ARG, "fa_nmod_poly_mul_mid_device: bad arguments");
    if (zl >= zh) return FA_OK;
    if (!z) return fa_set_error(FA_E_ARG, "fa_nmod_poly_mul_mid_device: null output");
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = stream ? (hipStream_t)stream : st->stream;
    const int64_t zn = an + bn - 1;
    const uint64_t n = mod.n;
    // zeros past the product, and everything when n = 1
    const int64_t zero_from = n == 1 ? zl : (zh > zn ? (zl > zn ? zl : zn) : zh);
    if (zero_from < zh) {
        const uint64_t cnt0 = (uint64_t)(zh - zero_from);
        FA_LAUNCH("fa_k_fill_zero", fa_k_fill_zero, dim3((unsigned)((cnt0 + 255) / 256)),
                  dim3(256), 0, s, z + (zero_from - zl), cnt0);
        zh = zero_from;
    }
    if (zl >= zh) return fa_buffers_leave(st, s);
    const int depth = fa_poly_window_depth((uint64_t)zl, (uint64_t)zh, (uint64_t)zn,
                                           (uint64_t)(an > bn ? an : bn));
    if (depth > FA_MAX_DEPTH)
        return fa_set_error(FA_E_RANGE, "fa_nmod_poly_mul_mid_device: %lld coefficients",
                            (long long)zn);
    fa_tables* T = fa_get_tables(depth < 20 ? 20 : depth);
    if (!T) return FA_E_NODEV;
    int np, offset;
    int rc = fa_nmod_primes(T, (uint64_t)(an < bn ? an : bn), n, &np, &offset);
    if (rc) return rc;
    fa_crt_consts C;
    fa_crt_consts_init(&C, np, offset);
    direct_lease D;
    if (direct_acquire(n, depth, np, s, &D)) {
        T = const_cast<fa_tables*>(D.T);
        np = 1;
        offset = 0;
        C = D.C;
    }
    const size_t tr = (size_t)np * fa_res_stride(depth) * sizeof(double);
    void* buf;
    rc = fa_scratch_reserve(st, 2 * tr + 256, &buf, s);
    if (rc) return rc;
    double* A = (double*)buf;
    double* B = (double*)((char*)buf + tr);
    fa_pack_desc pa = {FA_PACK_U64, a, (uint64_t)an, 0, 1};
    fa_pack_desc pb = {FA_PACK_U64, b, (uint64_t)bn, 0, 1};
    rc = fa_convolve(st, T, np, offset, depth, &pa, &pb, A, B, s);
    if (rc) return rc;
    rc = fa_crt_nmod_launch(A, fa_res_stride(depth), C, (uint64_t)zl, (uint64_t)(zh - zl), n, z,
                            s);
    if (rc) return rc;
    return fa_buffers_leave(st, s);
}

// _fmpz_poly_mul_mid on device arrays (w-limb two's complement
// coefficients, abits/bbits their _fmpz_vec_max_bits), enqueued on `stream`:
// 1 computed, 0 declined (more than eight primes, fmpz_poly_mul.c:377-379)
// unless `ks` asks for the Kronecker route then (fmpz_poly/mul.c:160)
extern "C" int fa_fmpz_poly_mul_mid_device(uint64_t* z, uint64_t zw, int64_t zl, int64_t zh,
                                           const uint64_t* a, uint64_t aw, int64_t an,
                                           const uint64_t* b, uint64_t bw, int64_t bn,
                                           int64_t abits, int64_t bbits, int ks, void* stream)
{
    if (an <= 0 || bn <= 0 || !a || !b || aw == 0 || bw == 0 || zw == 0 || zl < 0)
        return fa_set_error(FA_E_ARG, "fa_fmpz_poly_mul_mid_device: bad arguments");
    if (zl >= zh) return 1;
    if (!z) return fa_set_error(FA_E_ARG, "fa_fmpz_poly_mul_mid_device: null output");
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = stream ? (hipStream_t)stream : st->stream;
    const int64_t zn = an + bn - 1;
    if (zh > zn) {
        const int64_t s0 = zl > zn ? zl : zn;
        const uint64_t w = (uint64_t)(zh - s0) * zw;
        FA_LAUNCH("fa_k_fill_zero", fa_k_fill_zero, dim3((unsigned)((w + 255) / 256)), dim3(256),
                  0, s, z + (uint64_t)(s0 - zl) * zw, w);
        if (zl >= zn) return fa_buffers_leave(st, s) ? FA_E_HIP : 1;
        zh = zn;
    }
    const int np = fa_fmpz_primes((uint64_t)bn, abits, bbits);
    const uint64_t cnt = (uint64_t)(zh - zl);
    int rc;
    if (!np) {
        if (!ks) return 0;
        const uint64_t ws = fa_ks_scratch_words(abits, bbits, (uint64_t)an, (uint64_t)bn);
        void* hb;
        rc = fa_hbuf_reserve(st, ws * 8, &hb, s);
        if (rc) return rc;
        rc = fa_ks_device(z, zw, (uint64_t)zl, cnt, a, aw, (uint64_t)an, b, bw, (uint64_t)bn,
                          abits, bbits, (uint64_t*)hb, s);
        if (rc) return rc;
        rc = fa_buffers_leave(st, s);
        return rc ? rc : 1;
    }
    const int depth = fa_poly_window_depth((uint64_t)zl, (uint64_t)zh, (uint64_t)zn,
                                           (uint64_t)(an > bn ? an : bn));
    if (depth > FA_MAX_DEPTH)
        return fa_set_error(FA_E_RANGE, "fa_fmpz_poly_mul_mid_device: %lld coefficients",
                            (long long)zn);
    fa_tables* T = fa_get_tables(depth < 20 ? 20 : depth);
    if (!T) return FA_E_NODEV;
    fa_crt_consts C;
    fa_crt_consts_init(&C, np, 0);
    const size_t tr = (size_t)np * fa_res_stride(depth) * sizeof(double);
    void* buf;
    rc = fa_scratch_reserve(st, 2 * tr + 256, &buf, s);
    if (rc) return rc;
    double* A = (double*)buf;
    double* B = (double*)((char*)buf + tr);
    fa_pack_desc pa = {FA_PACK_S64W, a, (uint64_t)an, 0, (int)aw};
    fa_pack_desc pb = {FA_PACK_S64W, b, (uint64_t)bn, 0, (int)bw};
    rc = fa_convolve(st, T, np, 0, depth, &pa, &pb, A, B, s);
    if (rc) return rc;
#define FA_L(NL)                                                                            \
    FA_LAUNCH("fa_k_crt_fmpz", fa_k_crt_fmpz<NL>, dim3((unsigned)((cnt + 255) / 256)),      \
              dim3(256), 0, s, A, fa_res_stride(depth), C, (uint64_t)zl, cnt, zw, z)
    FA_CRT_NL_SWITCH(fa_crt_poly_limbs(C), FA_L)
#undef FA_L
    if (hipGetLastError() != hipSuccess) return fa_set_error(FA_E_HIP, "crt launch");
    rc = fa_buffers_leave(st, s);
    return rc ? rc : 1;
}

// _fmpz_vec_max_bits: max bit size of |c|, negated if some c < 0
int64_t fa_vec_max_bits(const uint64_t* v, uint64_t w, int64_t n)
{
    uint64_t mb = 0;
    int neg = 0;
    for (int64_t i = 0; i < n; i++) {
        const uint64_t* c = v + (uint64_t)i * w;
        int s = (int64_t)c[w - 1] < 0;
        neg |= s;
        uint64_t bits = 0;
        if (!s) {
            for (int64_t k = (int64_t)w - 1; k >= 0; k--)
                if (c[k]) { bits = 64 * (uint64_t)k + nbits_u64(c[k]); break; }
        } else {
            // |c| = ~c + 1, limb by limb from the bottom
            uint64_t cy = 1;
            for (uint64_t k = 0; k < w; k++) {
                uint64_t t = ~c[k] + cy;
                cy = (cy && t == 0);
                if (t) bits = 64 * k + nbits_u64(t);
            }
        }
        if (bits > mb) mb = bits;
    }
    return neg ? -(int64_t)mb : (int64_t)mb;
}

// _fmpz_vec_max_bits over w-limb two's complement coefficients
extern "C" int64_t fa_fmpz_vec_max_bits(const uint64_t* v, uint64_t w, int64_t n)
{
    return (v && w && n > 0) ? fa_vec_max_bits(v, w, n) : 0;
}

// primes for an fmpz product: the fewest np with prod >= bn * 2^(|b1|+|b2|+1)
// (fmpz_poly_mul.c:374-379); 0 when eight are not enough
int fa_fmpz_primes(uint64_t bn, int64_t b1, int64_t b2)
{
    uint64_t modbits = (uint64_t)(b1 < 0 ? -b1 : b1) + (uint64_t)(b2 < 0 ? -b2 : b2) + 1;
    for (int k = 1; k <= FA_NPRIMES; k++) {
        fa_crt_consts C;
        fa_crt_consts_init(&C, k, 0);
        if (prod_ge(C, bn, modbits)) return k;
    }
    return 0;
}

// prime set for an nmod product (plan.c:356-398): the modulus itself when it
// is one of the FFT primes, otherwise the fewest primes with
// prod >= len * 2^(2 modbits)
int fa_nmod_primes(const fa_tables* T, uint64_t len, uint64_t n, int* np, int* offset)
{
    const int modbits = nbits_u64(n);
    *np = 0;
    *offset = 0;
    if (modbits == 50)
        for (int i = 0; i < FA_NPRIMES; i++)
            if (T->pr[i].n == n) { *np = 1; *offset = i; }
    for (int k = 1; k <= FA_NPRIMES && !*np; k++) {
        fa_crt_consts C;
        fa_crt_consts_init(&C, k, 0);
        if (prod_ge(C, len, 2 * (uint64_t)modbits)) *np = k;
    }
    return *np ? FA_OK : fa_set_error(FA_E_RANGE, "no prime set for modulus");
}

// prime set of a plan whose output coefficients accumulate at most
// len_bound products of prod_bits-bit values (plan.c:356-398): the modulus
// itself when it is one of the FFT primes, otherwise the fewest primes with
// prod >= len_bound * 2^prod_bits
int fa_nmod_primes_bound(const fa_tables* T, uint64_t len_bound, uint64_t prod_bits, uint64_t n,
                         int* np, int* offset)
{
    *np = 0;
    *offset = 0;
    if (nbits_u64(n) == 50)
        for (int i = 0; i < FA_NPRIMES; i++)
            if (T->pr[i].n == n) { *np = 1; *offset = i; }
    for (int k = 1; k <= FA_NPRIMES && !*np; k++) {
        fa_crt_consts C;
        fa_crt_consts_init(&C, k, 0);
        if (prod_ge(C, len_bound, prod_bits)) *np = k;
    }
    return *np ? FA_OK : fa_set_error(FA_E_RANGE, "no prime set for %llu products of %llu bits",
                                      (unsigned long long)len_bound,
                                      (unsigned long long)prod_bits);
}

// coefficients [zl, zl + cnt) mod n of inverse-transformed residue rows
int fa_crt_nmod_launch(const double* A, uint64_t stride, const fa_crt_consts& C, uint64_t zl,
                       uint64_t cnt, uint64_t n, uint64_t* dz, hipStream_t s)
{
    if (!cnt) return FA_OK;
    const int norm = 64 - nbits_u64(n);
    const uint64_t d = n << norm;
    const uint64_t dinv = (uint64_t)(((u128)0 - 1) / d - ((u128)1 << 64));
#define FA_L(NP)                                                                            \
    FA_LAUNCH("fa_k_crt_nmod", fa_k_crt_nmod<NP>, dim3((unsigned)((cnt + 255) / 256)),      \
              dim3(256), 0, s, A, stride, C, zl, cnt, n, dinv, norm, dz)
    FA_CRT_NP_SWITCH(C.np, FA_L)
#undef FA_L
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "crt launch");
}

extern "C" int fa_fmpz_poly_mul_mid(uint64_t* z, uint64_t zw, int64_t zl, int64_t zh,
                                    const uint64_t* a, uint64_t aw, int64_t an, const uint64_t* b,
                                    uint64_t bw, int64_t bn)
{
    if (an <= 0 || bn <= 0 || !a || !b || aw == 0 || bw == 0 || zw == 0)
        return fa_set_error(FA_E_ARG, "fa_fmpz_poly_mul_mid: bad arguments");
    if (zl < 0 || zl >= zh) return 1;
    if (!z) return fa_set_error(FA_E_ARG, "fa_fmpz_poly_mul_mid: null output");
    int64_t zn = an + bn - 1;
    if (zh > zn) {
        int64_t s0 = zl > zn ? zl : zn;
        memset(z + (uint64_t)(s0 - zl) * zw, 0, (size_t)(zh - s0) * zw * sizeof(uint64_t));
        if (zl >= zn) return 1;
        zh = zn;
    }
    int64_t b1 = fa_vec_max_bits(a, aw, an);
    int64_t b2 = fa_vec_max_bits(b, bw, bn);
    int np = fa_fmpz_primes((uint64_t)bn, b1, b2);
    if (!np) return 0;
    const int depth = fa_poly_window_depth((uint64_t)zl, (uint64_t)zh, (uint64_t)zn,
                                           (uint64_t)(an > bn ? an : bn));
    fa_crt_consts C;
    fa_crt_consts_init(&C, np, 0);
    if (const uint64_t lim = poly_block_limit((uint64_t)zn)) {
        const int rcb = poly_mul_blocked(1, z, zw, zl, zh, a, aw, an, b, bw, bn, 0, np, 0, C, lim);
        return rcb ? rcb : 1;
    }
    fa_tables* T = fa_get_tables(depth);
    if (!T) return FA_E_NODEV;
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    uint64_t cnt = (uint64_t)(zh - zl);
    uint64_t *da, *db, *dz;
    double *A, *B;
    int rc = poly_buffers(st, (uint64_t)an * aw, (uint64_t)bn * bw, np, depth, cnt * zw, &da, &db,
                          &A, &B, &dz);
    if (rc) return rc;
    hipStream_t s = st->stream;
    // a square keeps one device copy (fmpz_poly_mul.c:352 squares when a == b)
    if (a == b && an == bn && aw == bw) db = da;
    HIPTRY(hipMemcpyAsync(da, a, (uint64_t)an * aw * sizeof(uint64_t), hipMemcpyHostToDevice, s),
           "upload a");
    if (db != da)
        HIPTRY(hipMemcpyAsync(db, b, (uint64_t)bn * bw * sizeof(uint64_t), hipMemcpyHostToDevice,
                              s),
               "upload b");
    fa_pack_desc pa = {FA_PACK_S64W, da, (uint64_t)an, 0, (int)aw};
    fa_pack_desc pb = {FA_PACK_S64W, db, (uint64_t)bn, 0, (int)bw};
    rc = fa_convolve(st, T, np, 0, depth, &pa, &pb, A, B, s);
    if (rc) return rc;
#define FA_L(NL)                                                                            \
    FA_LAUNCH("fa_k_crt_fmpz", fa_k_crt_fmpz<NL>, dim3((unsigned)((cnt + 255) / 256)),      \
              dim3(256), 0, s, A, fa_res_stride(depth), C, (uint64_t)zl, cnt, zw, dz)
    FA_CRT_NL_SWITCH(fa_crt_poly_limbs(C), FA_L)
#undef FA_L
    HIPTRY(hipGetLastError(), "crt launch");
    HIPTRY(hipMemcpyAsync(z, dz, cnt * zw * sizeof(uint64_t), hipMemcpyDeviceToHost, s),
           "download");
    HIPTRY(hipStreamSynchronize(s), "synchronize");
    return 1;
}
