// fa_stream.hip -- the mixed request stream at its full size, generated and
// checked on the device.
//
// The stream's requests (200,000 of them, half integer products of up to 2^22
// limbs per operand) hold about half a terabyte of operands, so they are
// produced where they are used:
//  - headers (kind, sizes, bit sizes, moduli) come from one xoshiro256**
//    stream seeded by SplitMix64, drawn on the host in a fixed order, with
//    two 64-bit keys per request drawn last;
//  - the contents of operand s of a request are the SplitMix64 sequence
//    seeded with key[s]: element j depends on (key, j) only, so any thread
//    computes any element, and a chunk of the stream is written in HBM by one
//    launch;
//  - every product is checked by evaluation: A(x) B(x) = Z(x) mod m for
//    integers (x = 2^64), nmod polynomials (m = n) and fmpz polynomials
//    (m a 61-bit prime), each evaluation a device reduction over the
//    operand's words.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fa_arith.h"
#include "fa_internal.h"

namespace {

constexpr uint64_t SM_GAMMA = UINT64_C(0x9E3779B97F4A7C15);

__host__ __device__ inline uint64_t sm64_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}

// output j (from 0) of SplitMix64 seeded with key
__host__ __device__ inline uint64_t sm64_at(uint64_t key, uint64_t j)
{
    return sm64_mix(key + (j + 1) * SM_GAMMA);
}

struct fill_job {
    uint64_t* dst;
    uint64_t count;          // elements: limbs (mpn, nmod) or coefficients (fmpz)
    uint64_t key;
    uint64_t n;              // nmod modulus
    uint64_t w;              // fmpz limbs per coefficient
    int32_t kind;
    int32_t bits;            // fmpz coefficient bits
    uint64_t tile0;          // first tile of the job
};

constexpr int FILL_TPB = 256;
constexpr int FILL_PER = 4;
constexpr uint64_t FILL_TILE = FILL_TPB * FILL_PER;

__device__ int find_job(const uint64_t* tile0, int nj, uint64_t tile)
{
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile0[mid] <= tile) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// an fmpz coefficient: t of tw = ceil((bits+1)/64) draws, the top one cut to
// bits+1-64(tw-1) bits, t = 2^(bits+1)-1 mapped to 0, c = t - (2^bits - 1)
// as w limbs of two's complement
__device__ void fill_fmpz(uint64_t* c, uint64_t key, uint64_t e, int bits, uint64_t w)
{
    const uint64_t tw = (uint64_t)(bits + 64) / 64;
    const int topb = bits + 1 - 64 * (int)(tw - 1);
    const uint64_t topmask = topb == 64 ? ~UINT64_C(0) : (UINT64_C(1) << topb) - 1;
    int all_ones = 1;
    for (uint64_t k = 0; k < tw; k++) {
        uint64_t v = sm64_at(key, e * tw + k);
        if (k == tw - 1) v &= topmask;
        all_ones &= (v == (k == tw - 1 ? topmask : ~UINT64_C(0)));
    }
    uint64_t borrow = 0;
    for (uint64_t k = 0; k < w; k++) {
        uint64_t x = 0;
        if (k < tw && !all_ones) {
            x = sm64_at(key, e * tw + k);
            if (k == tw - 1) x &= topmask;
        }
        const int lo = 64 * (int)k;
        const uint64_t m = bits >= lo + 64 ? ~UINT64_C(0) : bits <= lo ? 0
                                                                       : (UINT64_C(1) << (bits - lo)) - 1;
        const uint64_t d = x - m - borrow;
        borrow = (x < m) || (x - m < borrow);
        c[k] = d;
    }
}

__global__ void __launch_bounds__(FILL_TPB) k_stream_fill(const fill_job* __restrict__ J,
                                                          const uint64_t* __restrict__ tile0,
                                                          int nj, uint64_t ntiles)
{
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const fill_job job = J[find_job(tile0, nj, tile)];
        const uint64_t base = (tile - job.tile0) * FILL_TILE;
        for (int r = 0; r < FILL_PER; r++) {
            const uint64_t e = base + (uint64_t)r * FILL_TPB + threadIdx.x;
            if (e >= job.count) break;
            if (job.kind == FA_REQ_MPN) {
                uint64_t v = sm64_at(job.key, e);
                if (e == job.count - 1 && v == 0) v = 1;
                job.dst[e] = v;
            } else if (job.kind == FA_REQ_NMOD) {
                job.dst[e] = __umul64hi(sm64_at(job.key, e), job.n);
            } else {
                fill_fmpz(job.dst + e * job.w, job.key, e, job.bits, job.w);
            }
        }
    }
}

struct eval_job {
    const uint64_t* src;
    uint64_t count;          // coefficients
    uint64_t w;              // limbs per coefficient
    uint64_t x, m, dinv;     // point, modulus, its preinverse
    uint64_t R, Rw;          // 2^64 and 2^(64 w) mod m
    int32_t norm, is_signed;
    uint64_t tile0;
};

constexpr int EVAL_TPB = 256;
constexpr int EVAL_PER = 16;                    // coefficients per thread (Horner)
constexpr uint64_t EVAL_TILE = EVAL_TPB * EVAL_PER;

struct mod_t {
    uint64_t m, dinv;
    int norm;
};

__device__ uint64_t mulmod(uint64_t a, uint64_t b, const mod_t& M)
{
    return fa_mulmod_preinv(a, b, M.m, M.dinv, M.norm);
}

__device__ uint64_t addmod(uint64_t a, uint64_t b, const mod_t& M)
{
    const uint64_t s = a + b;
    return (s < a || s >= M.m) ? s - M.m : s;
}

// a x + c mod m for a, c < m
__device__ uint64_t muladd(uint64_t a, uint64_t x, uint64_t c, const mod_t& M)
{
    const uint64_t lo = a * x, hi = __umul64hi(a, x);
    const uint64_t l2 = lo + c;
    return fa_ll_mod_preinv(hi + (l2 < lo), l2, M.m, M.dinv, M.norm);
}

__device__ uint64_t powmod(uint64_t b, uint64_t e, const mod_t& M)
{
    uint64_t r = 1 % M.m;
    while (e) {
        if (e & 1) r = mulmod(r, b, M);
        b = mulmod(b, b, M);
        e >>= 1;
    }
    return r;
}

// coefficient value mod m: sum of limb_l R^l, less 2^(64 w) when signed and negative
__device__ uint64_t coef_mod(const eval_job& J, const mod_t& M, uint64_t i)
{
    const uint64_t* c = J.src + i * J.w;
    uint64_t r = 0;
    for (uint64_t l = J.w; l-- > 0;)
        r = muladd(r, J.R, fa_ll_mod_preinv(0, c[l], M.m, M.dinv, M.norm), M);
    if (J.is_signed && (c[J.w - 1] >> 63)) r = addmod(r, M.m - J.Rw, M);
    return r;
}

// tree sum of v[0..EVAL_TPB) with v[k] carrying x^(P k): v[0] + v[1] P + ...
__device__ uint64_t block_combine(uint64_t* v, uint64_t P, const mod_t& M)
{
    const int t = threadIdx.x;
    for (int s = 1; s < EVAL_TPB; s <<= 1) {
        __syncthreads();
        if ((t & (2 * s - 1)) == 0) v[t] = addmod(v[t], mulmod(v[t + s], P, M), M);
        P = mulmod(P, P, M);
    }
    __syncthreads();
    return v[0];
}

__global__ void __launch_bounds__(EVAL_TPB) k_eval_tiles(const eval_job* __restrict__ J,
                                                         const uint64_t* __restrict__ tile0,
                                                         int nj, uint64_t ntiles,
                                                         uint64_t* __restrict__ partial)
{
    __shared__ uint64_t v[EVAL_TPB];
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const eval_job job = J[find_job(tile0, nj, tile)];
        const mod_t M = {job.m, job.dinv, job.norm};
        const uint64_t first = (tile - job.tile0) * EVAL_TILE + (uint64_t)threadIdx.x * EVAL_PER;
        uint64_t p = 0;
        for (int j = EVAL_PER - 1; j >= 0; j--)
            if (first + j < job.count) p = muladd(p, job.x, coef_mod(job, M, first + j), M);
        v[threadIdx.x] = p;
        const uint64_t r = block_combine(v, powmod(job.x, EVAL_PER, M), M);
        if (threadIdx.x == 0) partial[tile] = r;
        __syncthreads();
    }
}

// job j's value: sum over its tiles of partial x^(EVAL_TILE k)
__global__ void __launch_bounds__(EVAL_TPB) k_eval_combine(const eval_job* __restrict__ J,
                                                           const uint64_t* __restrict__ partial,
                                                           uint64_t* __restrict__ out)
{
    __shared__ uint64_t v[EVAL_TPB];
    const eval_job job = J[blockIdx.x];
    const mod_t M = {job.m, job.dinv, job.norm};
    const uint64_t ntile = (job.count + EVAL_TILE - 1) / EVAL_TILE;
    // thread t sums tiles t, t + T, t + 2T, ... as a polynomial in x^(T EVAL_TILE)
    const uint64_t step = powmod(job.x, EVAL_TILE * EVAL_TPB, M);
    uint64_t acc = 0;
    const uint64_t mine = threadIdx.x < ntile ? (ntile - 1 - threadIdx.x) / EVAL_TPB + 1 : 0;
    for (uint64_t k = mine; k-- > 0;)
        acc = muladd(acc, step, partial[job.tile0 + threadIdx.x + k * EVAL_TPB], M);
    v[threadIdx.x] = acc;
    const uint64_t r = block_combine(v, powmod(job.x, EVAL_TILE, M), M);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}

uint64_t host_mulmod(uint64_t a, uint64_t b, uint64_t m)
{
    return (uint64_t)((unsigned __int128)a * b % m);
}

uint64_t host_powmod(uint64_t b, uint64_t e, uint64_t m)
{
    uint64_t r = 1 % m;
    b %= m;
    while (e) {
        if (e & 1) r = host_mulmod(r, b, m);
        b = host_mulmod(b, b, m);
        e >>= 1;
    }
    return r;
}

// a device copy of a host array for the length of one call
struct dev_copy {
    void* p = nullptr;
    ~dev_copy() { if (p) (void)hipFree(p); }
    int put(const void* src, size_t bytes, hipStream_t s)
    {
        if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) {
            p = nullptr;
            return fa_set_error(FA_E_NOMEM, "stream: device table of %zu bytes", bytes);
        }
        if (bytes && hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, s) != hipSuccess)
            return fa_set_error(FA_E_HIP, "stream: table upload");
        return 0;
    }
};

int grid_of(uint64_t ntiles)
{
    return (int)std::min<uint64_t>(ntiles, 256 * 64);
}

}  // namespace

// ------------------------------------------------------------------------
// headers
// ------------------------------------------------------------------------

extern "C" int fa_stream_headers(uint64_t seed, int64_t count, fa_stream_req* out)
{
    if (count < 0 || (count && !out)) return fa_set_error(FA_E_ARG, "stream headers: arguments");
    fa_rng r;
    fa_rng_seed(&r, seed);
    for (int64_t k = 0; k < count; k++) {
        fa_stream_req q;
        memset(&q, 0, sizeof q);
        const double u = fa_rng_double(&r);
        if (u < 0.5) {
            q.kind = FA_REQ_MPN;
            q.an = (int64_t)fa_rng_log_uniform(&r, 4, 22);
            q.bn = (int64_t)fa_rng_log_uniform(&r, 4, 22);
        } else if (u < 0.8) {
            q.kind = FA_REQ_NMOD;
            q.an = (int64_t)fa_rng_log_uniform(&r, 6, 16);
            q.bn = (int64_t)fa_rng_log_uniform(&r, 6, 16);
            q.bits = (int32_t)fa_rng_uniform(&r, 20, 63);
            q.n = fa_rng_random_prime(&r, q.bits);
        } else {
            q.kind = FA_REQ_FMPZ;
            q.an = (int64_t)fa_rng_uniform(&r, 1024, 16384);
            q.bn = (int64_t)fa_rng_uniform(&r, 1024, 16384);
            q.bits = (int32_t)fa_rng_uniform(&r, 64, 4096);
        }
        q.key[0] = fa_rng_next(&r);
        q.key[1] = fa_rng_next(&r);
        out[k] = q;
    }
    return 0;
}

// ------------------------------------------------------------------------
// operand contents
// ------------------------------------------------------------------------

extern "C" int fa_stream_fill_device(const fa_fill_spec* spec, int64_t n, void* stream)
{
    if (n < 0 || (n && !spec)) return fa_set_error(FA_E_ARG, "stream fill: arguments");
    if (n == 0) return 0;
    std::vector<fill_job> J;
    std::vector<uint64_t> T;
    uint64_t tiles = 0;
    for (int64_t k = 0; k < n; k++) {
        const fa_fill_spec& s = spec[k];
        if (s.count == 0) continue;
        if (!s.dst || s.kind < FA_REQ_MPN || s.kind > FA_REQ_FMPZ)
            return fa_set_error(FA_E_ARG, "stream fill: job %lld", (long long)k);
        if (s.kind == FA_REQ_NMOD && s.n == 0)
            return fa_set_error(FA_E_ARG, "stream fill: job %lld modulus 0", (long long)k);
        if (s.kind == FA_REQ_FMPZ && (s.bits < 1 || s.w < (uint64_t)s.bits / 64 + 1))
            return fa_set_error(FA_E_ARG, "stream fill: job %lld needs %d-bit coefficients in "
                                "%llu limbs", (long long)k, s.bits, (unsigned long long)s.w);
        fill_job j;
        j.dst = s.dst;
        j.count = s.count;
        j.key = s.key;
        j.n = s.n;
        j.w = s.kind == FA_REQ_FMPZ ? s.w : 1;
        j.kind = s.kind;
        j.bits = s.bits;
        j.tile0 = tiles;
        tiles += (s.count + FILL_TILE - 1) / FILL_TILE;
        J.push_back(j);
        T.push_back(j.tile0);
    }
    if (J.empty()) return 0;
    hipStream_t s = (hipStream_t)stream;
    dev_copy dj, dt;
    int rc;
    if ((rc = dj.put(J.data(), J.size() * sizeof(fill_job), s)) < 0) return rc;
    if ((rc = dt.put(T.data(), T.size() * sizeof(uint64_t), s)) < 0) return rc;
    hipLaunchKernelGGL(k_stream_fill, dim3(grid_of(tiles)), dim3(FILL_TPB), 0, s,
                       (const fill_job*)dj.p, (const uint64_t*)dt.p, (int)J.size(), tiles);
    if (hipGetLastError() != hipSuccess) return fa_set_error(FA_E_HIP, "stream fill: launch");
    if (hipStreamSynchronize(s) != hipSuccess) return fa_set_error(FA_E_HIP, "stream fill: run");
    return 0;
}

// ------------------------------------------------------------------------
// evaluation checks
// ------------------------------------------------------------------------

extern "C" int fa_eval_device(const fa_eval_spec* spec, int64_t n, uint64_t* out, void* stream)
{
    if (n < 0 || (n && (!spec || !out))) return fa_set_error(FA_E_ARG, "eval: arguments");
    if (n == 0) return 0;
    std::vector<eval_job> J((size_t)n);
    std::vector<uint64_t> T((size_t)n);
    uint64_t tiles = 0;
    for (int64_t k = 0; k < n; k++) {
        const fa_eval_spec& s = spec[k];
        if (s.m < 2 || s.w < 1 || (s.count && !s.src))
            return fa_set_error(FA_E_ARG, "eval: job %lld", (long long)k);
        eval_job& j = J[k];
        j.src = s.src;
        j.count = s.count;
        j.w = s.w;
        j.m = s.m;
        j.x = s.x % s.m;
        const int norm = __builtin_clzll(s.m);
        const uint64_t d = s.m << norm;
        j.norm = norm;
        j.dinv = (uint64_t)(~(unsigned __int128)0 / d - ((unsigned __int128)1 << 64));
        j.R = (uint64_t)(((unsigned __int128)1 << 64) % s.m);
        j.Rw = host_powmod(j.R, s.w, s.m);
        j.is_signed = s.is_signed;
        j.tile0 = tiles;
        T[k] = tiles;
        // at least one tile per job, so every job owns a partial
        tiles += std::max<uint64_t>(1, (s.count + EVAL_TILE - 1) / EVAL_TILE);
    }
    hipStream_t s = (hipStream_t)stream;
    dev_copy dj, dt, dp, dout;
    int rc;
    if ((rc = dj.put(J.data(), J.size() * sizeof(eval_job), s)) < 0) return rc;
    if ((rc = dt.put(T.data(), T.size() * sizeof(uint64_t), s)) < 0) return rc;
    if (hipMalloc(&dp.p, tiles * sizeof(uint64_t)) != hipSuccess) {
        dp.p = nullptr;
        return fa_set_error(FA_E_NOMEM, "eval: %llu partials", (unsigned long long)tiles);
    }
    if (hipMalloc(&dout.p, (size_t)n * sizeof(uint64_t)) != hipSuccess) {
        dout.p = nullptr;
        return fa_set_error(FA_E_NOMEM, "eval: results");
    }
    hipLaunchKernelGGL(k_eval_tiles, dim3(grid_of(tiles)), dim3(EVAL_TPB), 0, s,
                       (const eval_job*)dj.p, (const uint64_t*)dt.p, (int)n, tiles,
                       (uint64_t*)dp.p);
    hipLaunchKernelGGL(k_eval_combine, dim3((unsigned)n), dim3(EVAL_TPB), 0, s,
                       (const eval_job*)dj.p, (const uint64_t*)dp.p, (uint64_t*)dout.p);
    if (hipGetLastError() != hipSuccess) return fa_set_error(FA_E_HIP, "eval: launch");
    if (hipMemcpyAsync(out, dout.p, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "eval: run");
    return 0;
}

extern "C" uint64_t fa_stream_word(uint64_t key, uint64_t j) { return sm64_at(key, j); }
