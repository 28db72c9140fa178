// fa_ctx.hip -- primes, device twiddle tables, per-thread state, plans.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <new>
#include <vector>

#include "fa_arith.h"
#include "fa_internal.h"
#include "fa_crt_fp.h"

typedef unsigned __int128 u128;

#define DEFAULT_PRIME UINT64_C(0x0003f00000000001)   // default_ctx.c:17

// ------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------

static thread_local char fa_errbuf[512] = "";

int fa_set_error(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(fa_errbuf, sizeof(fa_errbuf), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char* fa_last_error(void) { return fa_errbuf; }

// ------------------------------------------------------------------------
// word arithmetic
// ------------------------------------------------------------------------

uint64_t fa_mulmod_u64(uint64_t a, uint64_t b, uint64_t n)
{
    return (uint64_t)(((u128)a * b) % n);
}

uint64_t fa_powmod_u64(uint64_t a, uint64_t e, uint64_t n)
{
    uint64_t r = 1 % n;
    a %= n;
    while (e) {
        if (e & 1) r = fa_mulmod_u64(r, a, n);
        a = fa_mulmod_u64(a, a, n);
        e >>= 1;
    }
    return r;
}

double fa_signed_rep(uint64_t x, uint64_t n)
{
    x %= n;
    return (x > n / 2) ? -(double)(n - x) : (double)x;
}

static int nbits64(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

static int is_prime_u64(uint64_t n)
{
    static const uint64_t bases[7] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    if (n < 2) return 0;
    for (uint64_t s = 2; s < 64 && s * s <= n; s++)
        if (n % s == 0) return n == s;
    uint64_t d = n - 1;
    int r = 0;
    while (!(d & 1)) { d >>= 1; r++; }
    for (int i = 0; i < 7; i++) {
        uint64_t a = bases[i] % n;
        if (!a) continue;
        uint64_t x = fa_powmod_u64(a, d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int j = 1; j < r && comp; j++) {
            x = fa_mulmod_u64(x, x, n);
            if (x == n - 1) comp = 0;
        }
        if (comp) return 0;
    }
    return 1;
}

// mpn_mul.c:1000-1016
static uint64_t next_fft_number(uint64_t p)
{
    int bits = nbits64(p);
    int l = __builtin_ctzll(p - 1);
    uint64_t q = p - (UINT64_C(2) << l);
    if (nbits64(q) == bits) return q;
    if (l < 5) return (UINT64_C(1) << (bits - 2)) + 1;
    return (UINT64_C(1) << bits) - (UINT64_C(1) << (l - 1)) + 1;
}

static int jacobi_u64(uint64_t a, uint64_t n)
{
    int t = 1;
    a %= n;
    while (a) {
        while (!(a & 1)) {
            a >>= 1;
            uint64_t r = n & 7;
            if (r == 3 || r == 5) t = -t;
        }
        uint64_t tmp = a; a = n; n = tmp;
        if ((a & 3) == 3 && (n & 3) == 3) t = -t;
        a %= n;
    }
    return n == 1 ? t : 0;
}

// ulong_extras/primitive_root_prime.c:15-26 then sd_fft_ctx.c:32-37
static uint64_t primitive_2power_root(uint64_t p)
{
    uint64_t a = 2;
    if (!((p & 7) == 3 || (p & 7) == 5))
        for (a = 3; jacobi_u64(a, p) != -1; a += 2) {}
    int v = __builtin_ctzll(p - 1);
    return fa_powmod_u64(a, (p - 1) >> v, p);
}

static void find_default_primes(uint64_t* out)
{
    uint64_t p = DEFAULT_PRIME;
    for (int i = 0; i < FA_NPRIMES; i++) {
        if (i > 0) p = next_fft_number(p);
        while (!is_prime_u64(p)) p = next_fft_number(p);
        out[i] = p;
    }
}

int fa_is_prime_u64(uint64_t n) { return is_prime_u64(n); }

// the primes are found once per process (the search runs Miller-Rabin on
// each candidate; plans and CRT constants ask for them on every product)
static void default_primes(uint64_t* out)
{
    static uint64_t P[FA_NPRIMES];
    static std::once_flag once;
    std::call_once(once, [] { find_default_primes(P); });
    memcpy(out, P, sizeof(P));
}

extern "C" int fa_default_primes(uint64_t* out8)
{
    default_primes(out8);
    return FA_OK;
}

// ------------------------------------------------------------------------
// device tables
// ------------------------------------------------------------------------

// w2[q] = hi[q >> lg] * lo[q & (2^lg - 1)] reduced to [-n/2, n/2]: q splits
// into disjoint bit ranges, so w[2q] = w[2(h << lg)] w[2l] (theta is additive
// over disjoint bits, fft_small.h:124-199).
__global__ void fa_expand_twiddles(double* __restrict__ out, const double* __restrict__ lo,
                                   const double* __restrict__ hi, uint64_t count, int lg,
                                   double n, double ninv)
{
    uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    double r = fa_mulmod(hi[q >> lg], lo[q & ((UINT64_C(1) << lg) - 1)], n, ninv);
    double h = 0.5 * (n - 1.0);
    if (r > h) r -= n;
    else if (r < -h) r += n;
    out[q] = r;
}

static std::mutex fa_table_mutex;
static fa_tables* fa_table_cache[16];

#define HIPCHECK(x)                                                                     \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess)                                                           \
            return fa_set_error(FA_E_HIP, "%s failed: %s", #x, hipGetErrorString(e_));  \
    } while (0)

static int check_device(int dev)
{
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, dev));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fa_set_error(FA_E_NODEV, "device %d is %s, this library targets gfx950", dev,
                            prop.gcnArchName);
    return FA_OK;
}

int fa_prime_info(uint64_t n, fa_prime_host* pr)
{
    if (n < 3 || !(n & 1)) return fa_set_error(FA_E_ARG, "%llu is not an odd prime",
                                               (unsigned long long)n);
    pr->n = n;
    pr->p = (double)n;
    pr->pinv = 1.0 / pr->p;
    pr->v = __builtin_ctzll(n - 1);
    pr->g = primitive_2power_root(n);
    static const int e[8] = {32, 48, 96, 144, 192, 0, 0, 0};
    for (int k = 0; k < 8; k++) pr->pow2[k] = fa_signed_rep(fa_powmod_u64(2, e[k], n), n);
    return FA_OK;
}

// device tables w2[q] = w[2q] and w2inv[q] = w[2q]^-1 for q < 2^(depth-1)
// (at least one entry); sd_fft_ctx.c:56-229 grows the same table lazily
int fa_build_prime_tables(const fa_prime_host* pr, int depth, double** w2_out,
                          double** w2inv_out)
{
    uint64_t n = pr->n;
    int v = pr->v;
    if (depth > v) return fa_set_error(FA_E_RANGE, "prime %llu supports depth <= %d",
                                       (unsigned long long)n, v);
    uint64_t count = depth >= 1 ? (UINT64_C(1) << (depth - 1)) : 1;
    int lg = 0;
    while ((UINT64_C(1) << lg) < count && lg < 12) lg++;
    uint64_t nlo = UINT64_C(1) << lg, nhi = (count + nlo - 1) / nlo;
    uint64_t gi = fa_powmod_u64(pr->g, n - 2, n);
    // w2[2^(j-1)] = w[2^j] = g^(2^(v-j-1))
    auto root = [&](int j, int inv) {
        return fa_powmod_u64(inv ? gi : pr->g, UINT64_C(1) << (v - j - 1), n);
    };
    double* hlo = (double*)malloc(2 * nlo * sizeof(double));
    double* hhi = (double*)malloc(2 * nhi * sizeof(double));
    uint64_t* lo = (uint64_t*)malloc(nlo * sizeof(uint64_t));
    uint64_t* li = (uint64_t*)malloc(nlo * sizeof(uint64_t));
    lo[0] = li[0] = 1;
    for (int j = 1; (UINT64_C(1) << (j - 1)) < nlo; j++) {
        uint64_t base = UINT64_C(1) << (j - 1);
        uint64_t r = root(j, 0), ri = root(j, 1);
        for (uint64_t t = 0; t < base; t++) {
            lo[base + t] = fa_mulmod_u64(r, lo[t], n);
            li[base + t] = fa_mulmod_u64(ri, li[t], n);
        }
    }
    for (uint64_t t = 0; t < nlo; t++) {
        hlo[t] = fa_signed_rep(lo[t], n);
        hlo[nlo + t] = fa_signed_rep(li[t], n);
    }
    for (uint64_t h = 0; h < nhi; h++) {
        uint64_t x = 1, xi = 1;
        for (int b = 0; b < 40; b++) {
            if (!((h >> b) & 1)) continue;
            int j = b + lg + 1;
            x = fa_mulmod_u64(x, root(j, 0), n);
            xi = fa_mulmod_u64(xi, root(j, 1), n);
        }
        hhi[h] = fa_signed_rep(x, n);
        hhi[nhi + h] = fa_signed_rep(xi, n);
    }
    free(lo);
    free(li);
    int rc = FA_OK;
    double *dlo = NULL, *dhi = NULL, *w2 = NULL, *w2i = NULL;
    if (hipMalloc(&dlo, 2 * nlo * sizeof(double)) != hipSuccess ||
        hipMalloc(&dhi, 2 * nhi * sizeof(double)) != hipSuccess ||
        hipMalloc(&w2, count * sizeof(double)) != hipSuccess ||
        hipMalloc(&w2i, count * sizeof(double)) != hipSuccess) {
        rc = fa_set_error(FA_E_NOMEM, "twiddle table allocation failed");
    } else if (hipMemcpy(dlo, hlo, 2 * nlo * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(dhi, hhi, 2 * nhi * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        rc = fa_set_error(FA_E_HIP, "twiddle upload failed");
    } else {
        unsigned grid = (unsigned)((count + 255) / 256);
        hipLaunchKernelGGL(fa_expand_twiddles, dim3(grid), dim3(256), 0, 0, w2, dlo, dhi, count, lg,
                           pr->p, pr->pinv);
        hipLaunchKernelGGL(fa_expand_twiddles, dim3(grid), dim3(256), 0, 0, w2i, dlo + nlo,
                           dhi + nhi, count, lg, pr->p, pr->pinv);
        if (hipDeviceSynchronize() != hipSuccess)
            rc = fa_set_error(FA_E_HIP, "twiddle expansion failed");
    }
    if (dlo) (void)hipFree(dlo);
    if (dhi) (void)hipFree(dhi);
    free(hlo);
    free(hhi);
    if (rc != FA_OK) {
        if (w2) (void)hipFree(w2);
        if (w2i) (void)hipFree(w2i);
        return rc;
    }
    *w2_out = w2;
    *w2inv_out = w2i;
    return FA_OK;
}

// Tables superseded by a deeper build stay allocated until the process
// ends: their addresses may sit in recorded batches and captured graphs
// (fa_batch_create copies them), or in the arguments of launches another
// thread is about to make from a fa_tables pointer it read before the
// growth.  A deeper table starts with the entries of the shallower one, so
// every holder of an old address still reads correct twiddles.  Growth
// happens at most a few times per process (depth 20, then up to 25).
static std::vector<void*> fa_retired_tables;
static std::mutex fa_retired_mutex;

void fa_retire_device_table(void* p)
{
    if (!p) return;
    std::lock_guard<std::mutex> g(fa_retired_mutex);
    fa_retired_tables.push_back(p);
}

static int build_tables(fa_tables* T, int depth)
{
    uint64_t P[FA_NPRIMES];
    default_primes(P);
    fa_prime_host pr[FA_NPRIMES];
    double* w2[FA_NPRIMES] = {};
    double* w2i[FA_NPRIMES] = {};
    for (int i = 0; i < FA_NPRIMES; i++) {
        fa_prime_info(P[i], &pr[i]);
        int rc = fa_build_prime_tables(&pr[i], depth, &w2[i], &w2i[i]);
        if (rc) {
            for (int j = 0; j < i; j++) {
                (void)hipFree(w2[j]);
                (void)hipFree(w2i[j]);
            }
            return rc;
        }
    }
    // publish the whole new set, then retire the old one
    for (int i = 0; i < FA_NPRIMES; i++) {
        fa_retire_device_table(T->w2[i]);
        fa_retire_device_table(T->w2inv[i]);
        T->pr[i] = pr[i];
        T->w2[i] = w2[i];
        T->w2inv[i] = w2i[i];
    }
    T->depth = depth;
    return FA_OK;
}

fa_tables* fa_get_tables(int depth)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        fa_set_error(FA_E_NODEV, "hipGetDevice failed");
        return NULL;
    }
    if (dev < 0 || dev >= 16) return NULL;
    if (depth < 13) depth = 13;
    if (depth > FA_MAX_DEPTH + 1) {
        fa_set_error(FA_E_RANGE, "transform depth %d above %d", depth, FA_MAX_DEPTH);
        return NULL;
    }
    std::lock_guard<std::mutex> g(fa_table_mutex);
    fa_tables* T = fa_table_cache[dev];
    if (T && T->depth >= depth) return T;
    if (!T) {
        if (check_device(dev) != FA_OK) return NULL;
        T = (fa_tables*)calloc(1, sizeof(fa_tables));
        T->device = dev;
    }
    // grow in steps so that repeated growth stays cheap
    int want = depth < 20 ? 20 : depth;
    if (build_tables(T, want) != FA_OK) {
        if (!fa_table_cache[dev]) free(T);
        return NULL;
    }
    fa_table_cache[dev] = T;
    return T;
}

// the current device's tables if they already hold `depth` (never grows
// them, so never allocates)
const fa_tables* fa_tables_ready(int depth)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return NULL;
    std::lock_guard<std::mutex> g(fa_table_mutex);
    const fa_tables* T = fa_table_cache[dev];
    return (T && T->depth >= depth) ? T : NULL;
}

int fa_noalloc_setup();

// builds the current device's tables and resolves the device address of
// the no-allocation entry point's scratch array
extern "C" int fa_init(void)
{
    if (!fa_get_tables(20)) return FA_E_NODEV;
    return fa_noalloc_setup();
}

// ------------------------------------------------------------------------
// per-thread state
// ------------------------------------------------------------------------

static thread_local fa_thread_state* fa_tls_state = NULL;
// an explicit context (fa_mpn_ctx_init) lends its own state to the calls
// made through it, in place of the thread's default one
static thread_local fa_thread_state* fa_tls_override = NULL;

fa_thread_state* fa_state_new(int dev)
{
    fa_thread_state* st = (fa_thread_state*)calloc(1, sizeof(fa_thread_state));
    if (!st) {
        fa_set_error(FA_E_NOMEM, "host allocation failed");
        return NULL;
    }
    st->device = dev;
    if (hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking) != hipSuccess) {
        fa_set_error(FA_E_HIP, "stream creation failed");
        free(st);
        return NULL;
    }
    return st;
}

void fa_state_free(fa_thread_state* st)
{
    if (!st) return;
    (void)hipStreamSynchronize(st->stream);
    if (st->last_stream && st->last_stream != st->stream) (void)hipEventSynchronize(st->last_ev);
    if (st->scratch) (void)hipFree(st->scratch);
    if (st->hbuf) (void)hipFree(st->hbuf);
    if (st->tbuf) (void)hipFree(st->tbuf);
    if (st->last_ev) (void)hipEventDestroy(st->last_ev);
    (void)hipStreamDestroy(st->stream);
    free(st);
}

fa_thread_state* fa_state_swap(fa_thread_state* st)
{
    fa_thread_state* prev = fa_tls_override;
    fa_tls_override = st;
    return prev;
}

fa_thread_state* fa_state()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        fa_set_error(FA_E_NODEV, "no HIP device");
        return NULL;
    }
    if (fa_tls_override && fa_tls_override->device == dev) return fa_tls_override;
    fa_thread_state* st = fa_tls_state;
    if (st && st->device == dev) return st;
    st = fa_state_new(dev);
    if (st) fa_tls_state = st;
    return st;
}

// The buffers of a state (scratch, hbuf, tbuf) serve every call made with
// it.  A device-pointer call returns as soon as its work is enqueued on the
// caller's stream, so the next call, perhaps on another stream, could
// overwrite buffers still being read.  fa_buffers_leave records an event
// after a call's last use; fa_buffers_enter makes a call on a different
// stream wait for it.  Streams being captured into a graph neither wait nor
// record: a graph's replays are ordered by the stream they are launched on.
static int stream_capturing(hipStream_t s)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

int fa_buffers_enter(fa_thread_state* st, hipStream_t s)
{
    if (!st->last_stream || st->last_stream == s || stream_capturing(s)) return FA_OK;
    HIPCHECK(hipStreamWaitEvent(s, st->last_ev, 0));
    return FA_OK;
}

int fa_buffers_leave(fa_thread_state* st, hipStream_t s)
{
    if (stream_capturing(s)) return FA_OK;
    if (!st->last_ev) HIPCHECK(hipEventCreateWithFlags(&st->last_ev, hipEventDisableTiming));
    HIPCHECK(hipEventRecord(st->last_ev, s));
    st->last_stream = s;
    return FA_OK;
}

// device bytes the grow-only buffers of all states may hold together
// (0: no limit); fa_set_device_memory_limit
static std::atomic<uint64_t> fa_mem_limit{0};
static std::atomic<uint64_t> fa_mem_held{0};

extern "C" void fa_set_device_memory_limit(uint64_t bytes) { fa_mem_limit = bytes; }

static int grow(void** buf, size_t* size, size_t bytes, hipStream_t s)
{
    if (*size >= bytes) return FA_OK;
    if (*buf) {
        // earlier calls may have used the buffer on a caller's stream (and
        // the streams forked from it), not only on the thread's own: wait
        // for the whole device before freeing (growth is rare)
        (void)s;
        (void)hipDeviceSynchronize();
        (void)hipFree(*buf);
        fa_mem_held -= *size;
        *buf = NULL;
        *size = 0;
    }
    size_t want = bytes + bytes / 8;
    want = (want + 0xFFFFF) & ~(size_t)0xFFFFF;
    const uint64_t lim = fa_mem_limit;
    if (lim && fa_mem_held + want > lim)
        return fa_set_error(FA_E_NOMEM,
                            "device allocation of %zu bytes exceeds the library limit of %llu "
                            "bytes (%llu held)",
                            want, (unsigned long long)lim, (unsigned long long)fa_mem_held.load());
    if (hipMalloc(buf, want) != hipSuccess) {
        *buf = NULL;
        (void)hipGetLastError();
        return fa_set_error(FA_E_NOMEM, "device allocation of %zu bytes failed", want);
    }
    fa_mem_held += want;
    *size = want;
    return FA_OK;
}

// s: the stream the caller is about to use the buffer on (null: the
// state's own); it first waits for the previous call's use on another stream
static int reserve(fa_thread_state* st, void** buf, size_t* size, size_t bytes, void** out,
                   hipStream_t s)
{
    int rc = fa_buffers_enter(st, s ? s : st->stream);
    if (rc == FA_OK) rc = grow(buf, size, bytes, s);
    *out = *buf;
    return rc;
}

int fa_scratch_reserve(fa_thread_state* st, size_t bytes, void** out, hipStream_t s)
{
    return reserve(st, &st->scratch, &st->scratch_size, bytes, out, s);
}

int fa_hbuf_reserve(fa_thread_state* st, size_t bytes, void** out, hipStream_t s)
{
    return reserve(st, &st->hbuf, &st->hbuf_size, bytes, out, s);
}

int fa_tbuf_reserve(fa_thread_state* st, size_t bytes, void** out, hipStream_t s)
{
    return reserve(st, &st->tbuf, &st->tbuf_size, bytes, out, s);
}

extern "C" int fa_sync(void)
{
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    HIPCHECK(hipStreamSynchronize(st->stream));
    return FA_OK;
}

extern "C" void* fa_stream(void)
{
    fa_thread_state* st = fa_state();
    return st ? (void*)st->stream : NULL;
}

// ------------------------------------------------------------------------
// CRT constants (mpn_mul.c:1093-1137)
// ------------------------------------------------------------------------

static void ml_mul_1(uint64_t* r, const uint64_t* a, int n, uint64_t b)
{
    uint64_t c = 0;
    for (int i = 0; i < n; i++) {
        u128 t = (u128)a[i] * b + c;
        r[i] = (uint64_t)t;
        c = (uint64_t)(t >> 64);
    }
}

static uint64_t ml_divexact_1(uint64_t* r, const uint64_t* a, int n, uint64_t d)
{
    u128 rem = 0;
    for (int i = n - 1; i >= 0; i--) {
        u128 cur = (rem << 64) | a[i];
        r[i] = (uint64_t)(cur / d);
        rem = cur % d;
    }
    return (uint64_t)rem;
}

static uint64_t ml_mod_1(const uint64_t* a, int n, uint64_t d)
{
    u128 rem = 0;
    for (int i = n - 1; i >= 0; i--) rem = ((rem << 64) | a[i]) % d;
    return (uint64_t)rem;
}

static void crt_consts_build(fa_crt_consts* C, int np, int offset);

// constants per (np, offset), built once per process
void fa_crt_consts_init(fa_crt_consts* C, int np, int offset)
{
    static fa_crt_consts cache[FA_NPRIMES + 1][FA_NPRIMES + 1];
    static bool have[FA_NPRIMES + 1][FA_NPRIMES + 1];
    static std::mutex mu;
    if (np < 1 || np > FA_NPRIMES || offset < 0 || offset + np > FA_NPRIMES) {
        crt_consts_build(C, np, offset);
        return;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!have[np][offset]) {
        crt_consts_build(&cache[np][offset], np, offset);
        have[np][offset] = true;
    }
    *C = cache[np][offset];
}

static void crt_consts_build(fa_crt_consts* C, int np, int offset)
{
    uint64_t P[FA_NPRIMES];
    default_primes(P);
    memset(C, 0, sizeof(*C));
    C->np = np;
    C->M[0] = 1;
    for (int i = 0; i < np; i++) {
        uint64_t p = P[offset + i];
        C->p[i] = (double)p;
        C->pinv[i] = 1.0 / (double)p;
        C->rp[i] = 1.0 / (double)p;
        ml_mul_1(C->M, C->M, FA_CRT_MAXL, p);
    }
    C->nl = FA_CRT_MAXL;
    while (C->nl > 1 && C->M[C->nl - 1] == 0) C->nl--;
    for (int i = 0; i < np; i++) ml_divexact_1(C->cof[i], C->M, FA_CRT_MAXL, P[offset + i]);
    for (int k = 0; k < np; k++) {
        const uint64_t pk = P[offset + k];
        C->gcm[k] = (double)ml_mod_1(C->cof[k], FA_CRT_MAXL, pk);
        uint64_t pre = 1 % pk;
        for (int j = 0; j < k; j++) {
            const uint64_t pj = P[offset + j] % pk;
            C->gpm[k][j] = (double)pj;
            pre = fa_mulmod_u64(pre, pj, pk);
        }
        C->ginv[k] = (double)fa_powmod_u64(pre, pk - 2, pk);
    }
}

// 24-bit digits of a multi-limb number
static void ml_digits24(double* out, int nd, const uint64_t* x, int nl)
{
    for (int d = 0; d < nd; d++) {
        int pos = 24 * d, l = pos / 64, b = pos % 64;
        uint64_t v = 0;
        if (l < nl) v = x[l] >> b;
        if (b > 40 && l + 1 < nl) v |= x[l + 1] << (64 - b);
        out[d] = (double)(v & 0xffffff);
    }
}

void fa_crt_fp_consts_init(fa_crt_fp_consts* F, const fa_crt_consts& C)
{
    memset(F, 0, sizeof(*F));
    for (int i = 0; i < C.np; i++) {
        F->p[i] = C.p[i];
        F->pinv[i] = C.pinv[i];
        F->rp[i] = C.rp[i];
        ml_digits24(F->cof[i], 16, C.cof[i], FA_CRT_MAXL);
    }
    ml_digits24(F->M, 18, C.M, FA_CRT_MAXL);
    uint64_t h[FA_CRT_MAXL];
    for (int l = 0; l < FA_CRT_MAXL; l++)
        h[l] = (C.M[l] >> 1) | (l + 1 < FA_CRT_MAXL ? C.M[l + 1] << 63 : 0);
    ml_digits24(F->Mh, 18, h, FA_CRT_MAXL);
}

uint64_t fa_crt_cop(int np, int i)
{
    fa_crt_consts C;
    uint64_t P[FA_NPRIMES];
    default_primes(P);
    fa_crt_consts_init(&C, np, 0);
    return ml_mod_1(C.cof[i], FA_CRT_MAXL, P[i]);
}

void fa_prime_set_init(fa_prime_set* S, const fa_tables* T, int np, int depth, int offset,
                       int with_norm)
{
    memset(S, 0, sizeof(*S));
    fa_crt_consts C;
    fa_crt_consts_init(&C, np, offset);
    for (int i = 0; i < np; i++) {
        const fa_prime_host* pr = &T->pr[offset + i];
        S->p[i] = pr->p;
        S->pinv[i] = pr->pinv;
        for (int k = 0; k < 8; k++) S->pow2[i][k] = pr->pow2[k];
        S->w2[i] = T->w2[offset + i];
        S->w2inv[i] = T->w2inv[offset + i];
        if (with_norm) {
            // m = (cop * 2^depth)^-1 mod p (mpn_mul.c:1517-1518, plan.c:127-149)
            uint64_t n = pr->n;
            uint64_t cop = (np == 1) ? 1 : ml_mod_1(C.cof[i], FA_CRT_MAXL, n);
            uint64_t m = fa_mulmod_u64(cop, fa_powmod_u64(2, (uint64_t)depth, n), n);
            m = fa_powmod_u64(m, n - 2, n);
            S->norm[i] = fa_signed_rep(m, n);
        }
    }
}

// ------------------------------------------------------------------------
// GPU plan for integer products
// ------------------------------------------------------------------------

static uint64_t cdiv_u64(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
static int clog2_u64(uint64_t x) { return x <= 1 ? 0 : 64 - __builtin_clzll(x - 1); }

// M >= c * 2^e
static int ml_ge_ui_2exp(const uint64_t* M, uint64_t c, uint64_t e)
{
    uint64_t t[FA_CRT_MAXL + 4];
    memset(t, 0, sizeof(t));
    uint64_t q = e / 64, r = e % 64;
    if (q + 2 > FA_CRT_MAXL + 3) return 0;
    t[q] = c << r;
    if (r) t[q + 1] = c >> (64 - r);
    for (int i = FA_CRT_MAXL + 3; i >= 0; i--) {
        uint64_t mi = (i < FA_CRT_MAXL) ? M[i] : 0;
        if (mi != t[i]) return mi > t[i];
    }
    return 1;
}

// Chooses (np, bits) with bits a multiple of 8 such that every coefficient
// of the chunked product, at most min(alen, blen) * 2^(2 bits), stays below
// M = p_0 ... p_(np-1) (the exactness bound of mpn_mul.c:79-95), minimizing
// the transform volume np * 2^depth.
int fa_mpn_choose_plan(uint64_t an, uint64_t bn, fa_mpn_plan_t* P)
{
    if (an == 0 || bn == 0) return fa_set_error(FA_E_ARG, "empty operand");
    double best = -1;
    for (int np = 2; np <= FA_NPRIMES; np++) {
        fa_crt_consts C;
        fa_crt_consts_init(&C, np, 0);
        int best_bits = 0;
        for (int bits = 8; bits <= 200; bits += 8) {
            uint64_t alen = cdiv_u64(64 * an, bits), blen = cdiv_u64(64 * bn, bits);
            uint64_t mlen = alen < blen ? alen : blen;
            // the CRT kernel keeps a group's overhang (< M / 2^bits) inside
            // the next group's bits/8 limbs
            if (9 * bits < 64 * C.nl + 64) continue;
            if (ml_ge_ui_2exp(C.M, mlen, 2 * (uint64_t)bits)) best_bits = bits;
        }
        if (!best_bits) continue;
        uint64_t alen = cdiv_u64(64 * an, best_bits), blen = cdiv_u64(64 * bn, best_bits);
        uint64_t zlen = alen + blen - 1;
        int depth = clog2_u64(zlen);
        if (depth < 4) depth = 4;
        if (depth > FA_MAX_DEPTH) continue;
        // transform volume weighted by the levels, plus the quadratic CRT
        double cost = (double)np * (double)(UINT64_C(1) << depth) * (depth + 6) +
                      0.5 * (double)zlen * np * np;
        if (best < 0 || cost < best) {
            best = cost;
            P->np = np;
            P->bits = best_bits;
            P->depth = depth;
            P->alen = alen;
            P->blen = blen;
            P->zlen = zlen;
        }
    }
    if (best < 0)
        return fa_set_error(FA_E_RANGE, "product of %llu x %llu limbs exceeds 2^%d-point plans",
                            (unsigned long long)an, (unsigned long long)bn, FA_MAX_DEPTH);
    P->d2 = P->depth <= FA_ROW_LG ? P->depth : FA_ROW_LG;
    P->d1 = P->depth - P->d2;
    return FA_OK;
}

// largest chunk size (a multiple of 8 up to 200) that the plan constraints
// allow for np primes, or 0
int fa_mpn_max_bits(uint64_t an, uint64_t bn, int np)
{
    if (an == 0 || bn == 0 || np < 2 || np > FA_NPRIMES) return 0;
    fa_crt_consts C;
    fa_crt_consts_init(&C, np, 0);
    for (int bits = 200; bits >= 8; bits -= 8) {
        if (9 * bits < 64 * C.nl + 64) break;
        uint64_t alen = cdiv_u64(64 * an, bits), blen = cdiv_u64(64 * bn, bits);
        uint64_t mlen = alen < blen ? alen : blen;
        if (!ml_ge_ui_2exp(C.M, mlen, 2 * (uint64_t)bits)) continue;
        int depth = clog2_u64(alen + blen - 1);
        if (depth > FA_MAX_DEPTH) return 0;
        return bits;
    }
    return 0;
}

// a caller-chosen (np, bits) checked against the same exactness and
// geometry constraints as the automatic choice
int fa_mpn_fixed_plan(uint64_t an, uint64_t bn, int np, int bits, fa_mpn_plan_t* P)
{
    if (an == 0 || bn == 0) return fa_set_error(FA_E_ARG, "empty operand");
    if (np < 2 || np > FA_NPRIMES || bits < 8 || bits > 200 || bits % 8)
        return fa_set_error(FA_E_ARG, "plan (np=%d, bits=%d) outside np in [2,8], bits in "
                            "8..200 step 8", np, bits);
    fa_crt_consts C;
    fa_crt_consts_init(&C, np, 0);
    if (9 * bits < 64 * C.nl + 64)
        return fa_set_error(FA_E_ARG, "bits=%d too small for %d primes", bits, np);
    uint64_t alen = cdiv_u64(64 * an, bits), blen = cdiv_u64(64 * bn, bits);
    uint64_t mlen = alen < blen ? alen : blen;
    if (!ml_ge_ui_2exp(C.M, mlen, 2 * (uint64_t)bits))
        return fa_set_error(FA_E_RANGE, "plan (np=%d, bits=%d) not exact for %llu x %llu limbs",
                            np, bits, (unsigned long long)an, (unsigned long long)bn);
    uint64_t zlen = alen + blen - 1;
    int depth = clog2_u64(zlen);
    if (depth < 4) depth = 4;
    if (depth > FA_MAX_DEPTH)
        return fa_set_error(FA_E_RANGE, "plan needs depth %d > %d", depth, FA_MAX_DEPTH);
    P->np = np;
    P->bits = bits;
    P->depth = depth;
    P->alen = alen;
    P->blen = blen;
    P->zlen = zlen;
    P->d2 = depth <= FA_ROW_LG ? depth : FA_ROW_LG;
    P->d1 = depth - P->d2;
    return FA_OK;
}

extern "C" int fa_mpn_plan(int64_t an, int64_t bn, int64_t* out)
{
    fa_mpn_plan_t P;
    if (an <= 0 || bn <= 0) return fa_set_error(FA_E_ARG, "an, bn must be >= 1");
    int rc = fa_mpn_choose_plan((uint64_t)an, (uint64_t)bn, &P);
    if (rc) return rc;
    out[0] = P.np; out[1] = P.bits; out[2] = P.depth; out[3] = P.d1; out[4] = P.d2;
    out[5] = (int64_t)P.zlen;
    return FA_OK;
}

// ------------------------------------------------------------------------
// synthetic inputs
// ------------------------------------------------------------------------

static uint64_t splitmix64(uint64_t* x)
{
    uint64_t z = (*x += UINT64_C(0x9E3779B97F4A7C15));
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

namespace {
struct synth_rng {
    uint64_t s[4];
    explicit synth_rng(uint64_t seed)
    {
        uint64_t x = seed;
        for (int i = 0; i < 4; i++) s[i] = splitmix64(&x);
    }
    uint64_t next()
    {
        uint64_t result = rotl64(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl64(s[3], 45);
        return result;
    }
    void limbs(uint64_t* a, uint64_t n, int top_nonzero)
    {
        for (uint64_t i = 0; i < n; i++) a[i] = next();
        if (top_nonzero && n > 0)
            while (a[n - 1] == 0) a[n - 1] = next();
    }
};
}  // namespace

extern "C" void fa_synth_limbs(uint64_t seed, uint64_t* a, uint64_t n, int top_nonzero)
{
    synth_rng r(seed);
    r.limbs(a, n, top_nonzero);
}

// both operands of one product from one stream: a then b, top limbs nonzero
extern "C" void fa_synth_mpn_pair(uint64_t seed, uint64_t* a, uint64_t an, uint64_t* b,
                                  uint64_t bn)
{
    synth_rng r(seed);
    r.limbs(a, an, 1);
    r.limbs(b, bn, 1);
}

// ------------------------------------------------------------------------
// the request generator of the batched workloads: one xoshiro256** stream
// (seeded by SplitMix64) per workload, drawn in a fixed order by the caller
// ------------------------------------------------------------------------

static synth_rng* rng_of(fa_rng* r) { return reinterpret_cast<synth_rng*>(r); }

extern "C" void fa_rng_seed(fa_rng* r, uint64_t seed)
{
    static_assert(sizeof(fa_rng) == sizeof(synth_rng), "rng state");
    new (rng_of(r)) synth_rng(seed);
}

extern "C" uint64_t fa_rng_next(fa_rng* r) { return rng_of(r)->next(); }

// uniform in [lo, hi] by rejection of the top partial range
extern "C" uint64_t fa_rng_uniform(fa_rng* r, uint64_t lo, uint64_t hi)
{
    const uint64_t range = hi - lo + 1;
    if (range == 0) return rng_of(r)->next();
    const uint64_t lim = ~UINT64_C(0) - (~UINT64_C(0) % range + 1) % range;
    for (;;) {
        const uint64_t x = rng_of(r)->next();
        if (x <= lim) return lo + x % range;
    }
}

// uniform double in [0, 1) from the top 53 bits
extern "C" double fa_rng_double(fa_rng* r)
{
    return (double)(rng_of(r)->next() >> 11) * 0x1p-53;
}

// floor(2^U), U uniform in [x, y)
extern "C" uint64_t fa_rng_log_uniform(fa_rng* r, double x, double y)
{
    const double u = x + (y - x) * fa_rng_double(r);
    return (uint64_t)exp2(u);
}

// odd draws in [2^(bits-1), 2^bits) until one is prime
extern "C" uint64_t fa_rng_random_prime(fa_rng* r, int bits)
{
    if (bits < 2) bits = 2;
    const uint64_t lo = UINT64_C(1) << (bits - 1);
    const uint64_t hi = bits == 64 ? ~UINT64_C(0) : (UINT64_C(1) << bits) - 1;
    for (;;) {
        uint64_t x = fa_rng_uniform(r, lo, hi) | 1;
        if (bits == 2) x = fa_rng_uniform(r, 2, 3);
        if (is_prime_u64(x)) return x;
    }
}

// n values uniform in [0, bound)
extern "C" void fa_rng_fill_below(fa_rng* r, uint64_t* out, uint64_t n, uint64_t bound)
{
    for (uint64_t i = 0; i < n; i++) out[i] = bound ? fa_rng_uniform(r, 0, bound - 1) : 0;
}

extern "C" void fa_rng_fill_limbs(fa_rng* r, uint64_t* out, uint64_t n, int top_nonzero)
{
    rng_of(r)->limbs(out, n, top_nonzero);
}

// n coefficients uniform in (-2^bits, 2^bits) as w-limb two's complement
// (w >= bits / 64 + 1): t uniform in [0, 2^(bits+1) - 2] by rejection of
// the all-ones (bits+1)-bit value, then c = t - (2^bits - 1)
extern "C" void fa_rng_fill_signed(fa_rng* r, uint64_t* out, uint64_t n, int bits, uint64_t w)
{
    const uint64_t tw = (uint64_t)(bits + 1 + 63) / 64;     // limbs of t
    const int topb = bits + 1 - 64 * (int)(tw - 1);          // bits in t's top limb
    const uint64_t topmask = topb == 64 ? ~UINT64_C(0) : (UINT64_C(1) << topb) - 1;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t* c = out + i * w;
        for (;;) {
            int all_ones = 1;
            for (uint64_t k = 0; k < tw; k++) {
                uint64_t v = rng_of(r)->next();
                if (k == tw - 1) v &= topmask;
                c[k] = v;
                all_ones &= (v == (k == tw - 1 ? topmask : ~UINT64_C(0)));
            }
            if (!all_ones) break;
        }
        for (uint64_t k = tw; k < w; k++) c[k] = 0;
        // subtract 2^bits - 1: limb k of the subtrahend is all ones below
        // bit `bits`, zero above
        uint64_t borrow = 0;
        for (uint64_t k = 0; k < w; k++) {
            const int lo = 64 * (int)k;
            uint64_t m;
            if (bits >= lo + 64) m = ~UINT64_C(0);
            else if (bits <= lo) m = 0;
            else m = (UINT64_C(1) << (bits - lo)) - 1;
            const uint64_t x = c[k];
            const uint64_t d = x - m - borrow;
            borrow = (x < m) || (x - m < borrow);
            c[k] = d;
        }
    }
}
