// fa_api.hip -- the reference-named drop-in entry points, their failure
// policy, and the mpn_ctx surface (explicit and thread-default contexts).
#include <stdio.h>
#include <string.h>

#include <atomic>

#include "fa_internal.h"

// ------------------------------------------------------------------------
// failure policy of the drop-ins
//
// The reference's entry points have no error return (default_ctx.c:41-57):
// they either produce the product or flint_throw.  The drop-ins here never
// abort.  On a device failure they
//   1. record the status for fa_last_status() on the calling thread, and
//   2. call the CPU implementation the caller registered with
//      fa_set_cpu_fallback (FLINT registers its own renamed CPU functions,
//      so the output is the reference's, limb for limb), or, when none is
//      registered, report the error on stderr and return with the output
//      unspecified.
// ------------------------------------------------------------------------

static std::atomic<fa_mpn_mul_fn> fb_mpn{nullptr};
static std::atomic<fa_nmod_poly_mul_mid_fn> fb_nmod{nullptr};
static thread_local int fa_status = FA_OK;

extern "C" void fa_set_cpu_fallback(fa_mpn_mul_fn mpn, fa_nmod_poly_mul_mid_fn nmod)
{
    fb_mpn = mpn;
    fb_nmod = nmod;
}

extern "C" int fa_last_status(void) { return fa_status; }

static void report(const char* fn, int rc)
{
    // a failed launch or copy leaves a sticky error in the runtime's
    // per-thread slot; clear it so the next call starts clean
    (void)hipGetLastError();
    fprintf(stderr, "%s: device path failed (%d): %s; output not computed\n", fn, rc,
            fa_last_error());
}

extern "C" void mpn_mul_default_mpn_ctx(uint64_t* r1, const uint64_t* i1, int64_t n1,
                                        const uint64_t* i2, int64_t n2)
{
    const int rc = fa_mpn_mul(r1, i1, n1, i2, n2);
    fa_status = rc;
    if (rc == FA_OK) return;
    if (fa_mpn_mul_fn f = fb_mpn.load()) {
        (void)hipGetLastError();
        f(r1, i1, n1, i2, n2);
        return;
    }
    report("mpn_mul_default_mpn_ctx", rc);
}

extern "C" void _nmod_poly_mul_mid_default_mpn_ctx(uint64_t* res, int64_t zl, int64_t zh,
                                                   const uint64_t* a, int64_t an,
                                                   const uint64_t* b, int64_t bn, fa_nmod_t mod)
{
    const int rc = fa_nmod_poly_mul_mid(res, zl, zh, a, an, b, bn, mod);
    fa_status = rc;
    if (rc == FA_OK) return;
    if (fa_nmod_poly_mul_mid_fn f = fb_nmod.load()) {
        (void)hipGetLastError();
        f(res, zl, zh, a, an, b, bn, mod);
        return;
    }
    report("_nmod_poly_mul_mid_default_mpn_ctx", rc);
}

// ------------------------------------------------------------------------
// mpn_ctx surface (mpn_mul.c:1076-1218, default_ctx.c:14-38)
//
// A context names the prime set (the eight primes found from p by
// next_fft_number, mpn_mul.c:1083-1091) and owns the execution state its
// calls use: a HIP stream and grow-only device buffers, as the reference's
// context owns its scratch buffer (mpn_ctx_fit_buffer, mpn_mul.c:1342).  The
// thread-default context uses the thread's own state.  The device twiddle
// tables are shared by all contexts of a device.  Like the reference's, one
// context serves one thread at a time.
// ------------------------------------------------------------------------

#define FA_DEFAULT_FIRST_PRIME UINT64_C(0x0003f00000000001)   // default_ctx.c:17

// runs f with the context's state standing in for the thread's
template <class F>
static int with_ctx(fa_mpn_ctx_struct* R, F f)
{
    if (!R || R->magic != FA_MPN_CTX_MAGIC)
        return fa_set_error(FA_E_ARG, "uninitialized mpn_ctx");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fa_set_error(FA_E_NODEV, "no HIP device");
    if (!R->state) return f();   // the thread-default context
    if (dev != R->device)
        return fa_set_error(FA_E_ARG, "mpn_ctx was made on device %d, the current one is %d",
                            R->device, dev);
    fa_thread_state* prev = fa_state_swap((fa_thread_state*)R->state);
    const int rc = f();
    fa_state_swap(prev);
    return rc;
}

extern "C" int fa_mpn_ctx_init(fa_mpn_ctx_struct* R, uint64_t p)
{
    if (!R) return fa_set_error(FA_E_ARG, "null mpn_ctx");
    memset(R, 0, sizeof(*R));
    uint64_t P[8];
    fa_default_primes(P);
    if (p != FA_DEFAULT_FIRST_PRIME)
        return fa_set_error(FA_E_ARG,
                            "mpn_ctx_init: first prime %#llx; the device tables are built for "
                            "the prime set of %#llx (every context in the reference uses it)",
                            (unsigned long long)p, (unsigned long long)FA_DEFAULT_FIRST_PRIME);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fa_set_error(FA_E_NODEV, "no HIP device");
    fa_thread_state* st = fa_state_new(dev);
    if (!st) return FA_E_NODEV;
    R->p = p;
    memcpy(R->primes, P, sizeof(P));
    R->device = dev;
    R->state = st;
    R->magic = FA_MPN_CTX_MAGIC;
    return FA_OK;
}

extern "C" void fa_mpn_ctx_clear(fa_mpn_ctx_struct* R)
{
    if (!R || R->magic != FA_MPN_CTX_MAGIC) return;
    if (R->state) fa_state_free((fa_thread_state*)R->state);
    memset(R, 0, sizeof(*R));
}

extern "C" fa_mpn_ctx_struct* fa_get_default_mpn_ctx(void)
{
    static thread_local fa_mpn_ctx_struct R;
    if (R.magic != FA_MPN_CTX_MAGIC) {
        R.p = FA_DEFAULT_FIRST_PRIME;
        fa_default_primes(R.primes);
        R.device = -1;     // follows the current device, like the thread's state
        R.state = NULL;
        R.magic = FA_MPN_CTX_MAGIC;
    }
    return &R;
}

extern "C" int fa_mpn_ctx_mpn_mul(fa_mpn_ctx_struct* R, uint64_t* z, const uint64_t* a,
                                  int64_t an, const uint64_t* b, int64_t bn)
{
    return with_ctx(R, [&] { return fa_mpn_mul(z, a, an, b, bn); });
}

extern "C" int fa_mpn_ctx_mpn_mul_device(fa_mpn_ctx_struct* R, uint64_t* z, const uint64_t* a,
                                         int64_t an, const uint64_t* b, int64_t bn,
                                         void* stream)
{
    return with_ctx(R, [&] { return fa_mpn_mul_device(z, a, an, b, bn, stream); });
}

extern "C" int fa_nmod_poly_mul_mid_mpn_ctx(uint64_t* z, int64_t zl, int64_t zh,
                                            const uint64_t* a, int64_t an, const uint64_t* b,
                                            int64_t bn, fa_nmod_t mod, fa_mpn_ctx_struct* R)
{
    return with_ctx(R, [&] { return fa_nmod_poly_mul_mid(z, zl, zh, a, an, b, bn, mod); });
}

extern "C" int fa_fmpz_poly_mul_mid_mpn_ctx(uint64_t* z, uint64_t zw, int64_t zl, int64_t zh,
                                            const uint64_t* a, uint64_t aw, int64_t an,
                                            const uint64_t* b, uint64_t bw, int64_t bn,
                                            fa_mpn_ctx_struct* R)
{
    return with_ctx(R, [&] { return fa_fmpz_poly_mul_mid(z, zw, zl, zh, a, aw, an, b, bw, bn); });
}
