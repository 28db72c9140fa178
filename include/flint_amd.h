/*
 * flint_amd.h -- C ABI of the MI355X (gfx950) implementation of FLINT's
 * fft_small multiplication path.
 *
 * Reference interfaces are cited as file:line of flintlib/flint 3.7.0-dev.
 * Plain pointers and sizes only.  Functions returning int report
 *   FA_OK (0) on success, a negative FA_E* code on failure
 * (fa_last_error() has the message); the fmpz entry points return 1 on
 * success and 0 exactly where the reference declines.
 *
 * Host-pointer entry points are synchronous: they return after the output
 * is in host memory.  Device-pointer entry points (suffix _device) enqueue
 * on the given HIP stream (NULL = the calling thread's stream) and return
 * immediately.  Each host thread gets its own stream and scratch, mirroring
 * the thread-local default mpn_ctx of default_ctx.c:14-15.
 */
#ifndef FLINT_AMD_H
#define FLINT_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FA_API __attribute__((visibility("default")))

#define FA_OK 0
#define FA_E_ARG (-1)       /* invalid argument */
#define FA_E_HIP (-2)       /* HIP runtime error */
#define FA_E_NOMEM (-3)     /* device allocation failed */
#define FA_E_RANGE (-4)     /* size outside the supported range */
#define FA_E_NODEV (-5)     /* no usable gfx950 device */

/* layout of FLINT's nmod_t (flint.h.in:441-447) */
typedef struct {
    uint64_t n;
    uint64_t ninv;
    uint64_t norm;
} fa_nmod_t;

/* ---------------------------------------------------------------------- */
/* the drop-in entry points, with the reference's exact signatures        */
/* ---------------------------------------------------------------------- */

/* replaces mpn_mul_default_mpn_ctx, fft_small.h:852, default_ctx.c:41-45:
   r1[0, n1+n2) = i1 * i2.  Never aborts: on a device failure it calls the
   registered CPU fallback (fa_set_cpu_fallback), or reports on stderr and
   leaves r1 unspecified; fa_last_status() tells which. */
FA_API void mpn_mul_default_mpn_ctx(uint64_t* r1, const uint64_t* i1, int64_t n1,
                             const uint64_t* i2, int64_t n2);

/* replaces _nmod_poly_mul_mid_default_mpn_ctx, fft_small.h:853,
   default_ctx.c:47-51: res[0, zh-zl) = coefficients [zl, zh) of a*b mod n.
   Failure policy as mpn_mul_default_mpn_ctx. */
FA_API void _nmod_poly_mul_mid_default_mpn_ctx(uint64_t* res, int64_t zl, int64_t zh,
                                        const uint64_t* a, int64_t an,
                                        const uint64_t* b, int64_t bn,
                                        fa_nmod_t mod);

/* Failure policy of the two drop-ins above (no counterpart in the reference,
   whose flint_throw aborts).  The caller registers the CPU implementations to
   run when the device path fails -- in FLINT, its own functions under another
   name -- so the output is still the reference's.  NULL: no fallback. */
typedef void (*fa_mpn_mul_fn)(uint64_t* r1, const uint64_t* i1, int64_t n1,
                              const uint64_t* i2, int64_t n2);
typedef void (*fa_nmod_poly_mul_mid_fn)(uint64_t* res, int64_t zl, int64_t zh,
                                        const uint64_t* a, int64_t an,
                                        const uint64_t* b, int64_t bn, fa_nmod_t mod);
FA_API void fa_set_cpu_fallback(fa_mpn_mul_fn mpn, fa_nmod_poly_mul_mid_fn nmod);
/* status of the calling thread's last drop-in call: FA_OK, or the FA_E_*
   code of the device failure (then the fallback ran, if one is registered) */
FA_API int fa_last_status(void);

/* ---------------------------------------------------------------------- */
/* the mpn_ctx plugin surface, fft_small.h:382-433, 769-774, 850-860      */
/* ---------------------------------------------------------------------- */

/* A context: the prime set found from p (mpn_mul.c:1076-1091) plus the
   execution state its calls use (a HIP stream and grow-only device buffers,
   as the reference's context owns its scratch buffer, mpn_mul.c:1342-1362).
   The layout differs from the reference's mpn_ctx_struct, whose fields are
   host tables; callers treat it as opaque.  One context serves one thread at
   a time, as in the reference. */
#define FA_MPN_CTX_MAGIC 0x66616d70636f6e74ULL
typedef struct {
    uint64_t p;              /* the first prime */
    uint64_t primes[8];      /* the eight FFT primes */
    int32_t device;          /* HIP device of the state (-1: the current one) */
    int32_t pad;
    void* state;             /* library-internal; NULL for the thread default */
    uint64_t magic;
} fa_mpn_ctx_struct;
typedef fa_mpn_ctx_struct fa_mpn_ctx_t[1];

/* mpn_ctx_init, mpn_mul.c:1076.  This build's device tables carry the
   prime set of p = 0x0003f00000000001 (default_ctx.c:17, the only p the
   reference's sources pass); another p returns FA_E_ARG. */
FA_API int fa_mpn_ctx_init(fa_mpn_ctx_struct* R, uint64_t p);
/* mpn_ctx_clear, mpn_mul.c:1220 */
FA_API void fa_mpn_ctx_clear(fa_mpn_ctx_struct* R);
/* get_default_mpn_ctx, default_ctx.c:29-38: the calling thread's context */
FA_API fa_mpn_ctx_struct* fa_get_default_mpn_ctx(void);
/* mpn_ctx_mpn_mul, fft_small.h:433, mpn_mul.c:1980 */
FA_API int fa_mpn_ctx_mpn_mul(fa_mpn_ctx_struct* R, uint64_t* z, const uint64_t* a,
                              int64_t an, const uint64_t* b, int64_t bn);
/* the same on device operands, enqueued on `stream` (NULL: the context's) */
FA_API int fa_mpn_ctx_mpn_mul_device(fa_mpn_ctx_struct* R, uint64_t* z, const uint64_t* a,
                                     int64_t an, const uint64_t* b, int64_t bn, void* stream);
/* _nmod_poly_mul_mid_mpn_ctx, fft_small.h:769-774, nmod_poly_mul.c:904 */
FA_API int fa_nmod_poly_mul_mid_mpn_ctx(uint64_t* z, int64_t zl, int64_t zh,
                                        const uint64_t* a, int64_t an, const uint64_t* b,
                                        int64_t bn, fa_nmod_t mod, fa_mpn_ctx_struct* R);
/* _fmpz_poly_mul_mid_mpn_ctx, fft_small.h:856-860, fmpz_poly_mul.c:319;
   coefficients as fa_fmpz_poly_mul_mid */
FA_API int fa_fmpz_poly_mul_mid_mpn_ctx(uint64_t* z, uint64_t zw, int64_t zl, int64_t zh,
                                        const uint64_t* a, uint64_t aw, int64_t an,
                                        const uint64_t* b, uint64_t bw, int64_t bn,
                                        fa_mpn_ctx_struct* R);

/* ---------------------------------------------------------------------- */
/* status-returning C ABI                                                 */
/* ---------------------------------------------------------------------- */

/* mpn_ctx_mpn_mul, mpn_mul.c:1980-1983 (full product, an, bn >= 1) */
FA_API int fa_mpn_mul(uint64_t* z, const uint64_t* a, int64_t an,
               const uint64_t* b, int64_t bn);

/* _mpn_ctx_mpn_mul_range, mpn_mul.c:1617-1978, as flint_mpn_mulmid_fft_small
   (mpn_extras/mulmid_fft_small.c:26-33) calls it: z[0, hi-lo) = limbs
   [lo, hi) of a*b, 0 <= lo < hi <= an+bn, word for word as the reference
   returns them with FLINT on one thread (its default): a lower
   approximation that drops carries from below lo as the reference's plan
   for (an, bn) does (mpn_mul_np2.c:380-402 per chunk for max(an, bn) <=
   40000, mpn_mul.c:1669-1978 below coefficient c_lo otherwise; deficit at
   most one carry, doc/source/mpn_extras.rst:221-231). */
FA_API int fa_mpn_mul_range(uint64_t* z, uint64_t lo, uint64_t hi, const uint64_t* a,
                            int64_t an, const uint64_t* b, int64_t bn);
/* fa_mpn_mul_range with the words of FLINT running nthreads threads (its
   plan, mpn_mul.c:1248-1340, depends on the thread count: pass
   flint_get_num_threads()); nthreads = 0 gives the exact window.  _host:
   host operands and output; _threads: device ones, enqueued on `stream` */
FA_API int fa_mpn_mul_range_host(uint64_t* z, uint64_t lo, uint64_t hi, const uint64_t* a,
                                 int64_t an, const uint64_t* b, int64_t bn, int nthreads);
FA_API int fa_mpn_mul_range_threads(uint64_t* z, uint64_t lo, uint64_t hi, const uint64_t* a,
                                    int64_t an, const uint64_t* b, int64_t bn, int nthreads,
                                    void* stream);
/* the reference's plan as the window sees it: out4 = {two-prime plan,
   chunk bits, first boundary coefficient, boundary coefficients (0: the
   window is exact)} */
FA_API int fa_mpn_window_plan(uint64_t an, uint64_t bn, uint64_t lo, int nthreads,
                              int64_t* out4);
/* sd_fft_mpn_mulmod_2expp1 (negacyclic.c:491-501), the pointwise stage of
   FLINT's Schoenhage-Strassen convolutions (fmpz_poly/mul_SS.c):
   z[0, n] = x y mod 2^(64 n) + 1 for x, y of n + 1 limbs with values at
   most 2^(64 n), the result normalized to [0, 2^(64 n)].  Host operands are
   checked; the device variant enqueues on `stream`.  n + 1 limbs must fit
   one transform (2^24 points). */
FA_API int fa_mpn_mulmod_2expp1(uint64_t* z, const uint64_t* x, const uint64_t* y, int64_t n);
FA_API int fa_mpn_mulmod_2expp1_device(uint64_t* z, const uint64_t* x, const uint64_t* y,
                                       int64_t n, void* stream);
/* fa_mpn_mul_range with device operands and output, enqueued on `stream`
   (FLINT's one-thread words) */
FA_API int fa_mpn_mul_range_device(uint64_t* z, uint64_t lo, uint64_t hi, const uint64_t* a,
                                   int64_t an, const uint64_t* b, int64_t bn, void* stream);

/* same with operands and output in device memory (no PCIe copies) */
FA_API int fa_mpn_mul_device(uint64_t* z, const uint64_t* a, int64_t an,
                      const uint64_t* b, int64_t bn, void* stream);

/* fa_mpn_mul_device for callers that cannot allocate (code under HIP graph
   capture, serving loops, harnesses that forbid hipMalloc).  No counterpart
   in the reference.  Same kernels and output as fa_mpn_mul_device, with the
   scratch in a static 64 MiB device array per GPU instead of allocated
   buffers: it calls no hipMalloc, hipFree or allocating library routine.
   Products whose scratch exceeds the array (fa_noalloc_scratch_bytes();
   about 2^20 limbs in total) return FA_E_RANGE.  fa_init() must have run on
   the device first (it builds the tables and resolves the array's address);
   otherwise FA_E_ARG.  All calls on one GPU share the array: calls that
   overlap on one GPU (other streams or threads) race. */
FA_API int fa_mpn_mul_device_noalloc(uint64_t* z, const uint64_t* a, int64_t an,
                                     const uint64_t* b, int64_t bn, void* stream);
FA_API uint64_t fa_noalloc_scratch_bytes(void);

/* the same with a caller-chosen plan: np primes, bits-bit chunks (checked
   against the exactness bound of mpn_mul.c:79-95); the autotuner's knob */
FA_API int fa_mpn_mul_plan(uint64_t* z, const uint64_t* a, int64_t an,
                           const uint64_t* b, int64_t bn, int np, int bits);
FA_API int fa_mpn_mul_device_plan(uint64_t* z, const uint64_t* a, int64_t an,
                                  const uint64_t* b, int64_t bn, int np, int bits,
                                  void* stream);

/* plan the device call would use: out = {np, bits, depth, d1, d2, zlen} */
FA_API int fa_mpn_plan(int64_t an, int64_t bn, int64_t* out);

/* nmod products computed with direct transforms mod the modulus so far
   (plan.c:319-349: n an FFT prime of 20..50 bits whose 2-valuation covers
   the transform, where the CRT would need two or more primes; the outputs
   are the same either way) */
FA_API uint64_t fa_nmod_direct_count(void);
/* 1 (default): direct transforms where they apply; 0: always the CRT primes */
FA_API void fa_set_direct_fft(int on);
/* _nmod_poly_mul_mid_mpn_ctx, nmod_poly_mul.c:904-972 */
FA_API int fa_nmod_poly_mul_mid(uint64_t* z, int64_t zl, int64_t zh,
                         const uint64_t* a, int64_t an,
                         const uint64_t* b, int64_t bn, fa_nmod_t mod);

/* the same on device arrays, enqueued on `stream` (NULL: the thread's);
   products of at most one transform (2^24 coefficients) */
FA_API int fa_nmod_poly_mul_mid_device(uint64_t* z, int64_t zl, int64_t zh,
                                       const uint64_t* a, int64_t an,
                                       const uint64_t* b, int64_t bn, fa_nmod_t mod,
                                       void* stream);

/* _nmod_poly_mulmid_fft_small_repack (nmod_poly/mullow_fft_small.c:
   340-411): for moduli n <= 23, coefficient pairs packed into one
   coefficient mod the FFT prime 1108307720798209 and multiplied at half
   length with that one prime.  z[0, zn-znlo) = coefficients [znlo, zn) of
   a*b mod n; returns 1 when the packed route applied, 0 when its bounds do
   not hold (z untouched), negative on error. */
FA_API int fa_nmod_poly_mulmid_repack(uint64_t* z, const uint64_t* a, int64_t an,
                                      const uint64_t* b, int64_t bn, int64_t znlo, int64_t zn,
                                      fa_nmod_t mod);
/* _nmod_poly_mulmid_fft_small (mullow_fft_small.c:419-424): the packed
   route when it applies, else _nmod_poly_mul_mid_default_mpn_ctx */
FA_API int fa_nmod_poly_mulmid_fft_small(uint64_t* z, const uint64_t* a, int64_t an,
                                         const uint64_t* b, int64_t bn, int64_t znlo,
                                         int64_t zn, fa_nmod_t mod);

/* _fmpz_poly_mul_mid on device arrays (two's complement coefficients as
   fa_fmpz_poly_mul_mid; abits, bbits = their _fmpz_vec_max_bits, which the
   caller knows or computes with fa_fmpz_vec_max_bits), enqueued on
   `stream`.  Returns 1 computed, 0 declined as the reference declines
   (fmpz_poly_mul.c:377-379); with ks != 0 the declined products are
   computed by Kronecker substitution instead (fmpz_poly/mul.c:160). */
FA_API int fa_fmpz_poly_mul_mid_device(uint64_t* z, uint64_t zw, int64_t zl, int64_t zh,
                                       const uint64_t* a, uint64_t aw, int64_t an,
                                       const uint64_t* b, uint64_t bw, int64_t bn,
                                       int64_t abits, int64_t bbits, int ks, void* stream);

/* _nmod_poly_inv_series as used by _nmod_poly_divrem_mpn_ctx
   (nmod_poly_mul.c:1246): g[0, m) = f^-1 mod x^m, f[0] invertible mod n;
   Newton iteration on the device products */
FA_API int fa_nmod_poly_inv_series(uint64_t* g, const uint64_t* f, int64_t fn, int64_t m,
                                   fa_nmod_t mod);
/* _nmod_poly_divrem_mpn_ctx, fft_small.h:776-782, nmod_poly_mul.c:1234-1262:
   q[0, an-bn+1) and r[0, bn-1) with a = b q + r, an >= bn >= 2, the leading
   coefficient of b invertible mod n (FA_E_ARG otherwise) */
FA_API int fa_nmod_poly_divrem(uint64_t* q, uint64_t* r, const uint64_t* a, int64_t an,
                               const uint64_t* b, int64_t bn, fa_nmod_t mod);
/* nmod_poly_divrem_precomp (fft_small.h:800-835, nmod_poly_mul.c:1265-1313):
   the divisor and its inverse to precision Bn kept on the device; divrem
   returns 1, or 0 when an - bn + 1 > Bn */
typedef struct fa_divrem_precomp_s* fa_divrem_precomp;
FA_API int fa_nmod_poly_divrem_precomp_init(fa_divrem_precomp* M, const uint64_t* b, int64_t bn,
                                            int64_t Bn, fa_nmod_t mod);
FA_API int fa_nmod_poly_divrem_precomp(uint64_t* q, uint64_t* r, const uint64_t* a, int64_t an,
                                       fa_divrem_precomp M, fa_nmod_t mod);
FA_API void fa_nmod_poly_divrem_precomp_clear(fa_divrem_precomp M);
/* mul_precomp_struct (fft_small.h:796-822, nmod_poly_mul.c:1052-1140): b
   (host) transformed once on a cyclic plan of 2^depth points, depth >= 4;
   fa_nmod_poly_mul_mid_precomp writes coefficients [zl, zh) of a b to z
   (host) and returns 1, or returns 0 without computing when the window
   does not fit the transform (zn = an + bn - 1 > 2^depth and not zh <=
   2^depth <= zn <= zl + 2^depth); a longer than 2^depth is folded mod
   x^(2^depth) - 1.  Calls on one precomp from several threads are safe. */
typedef struct fa_mul_precomp_s* fa_mul_precomp;
FA_API int fa_nmod_poly_mul_precomp_init(fa_mul_precomp* M, const uint64_t* b, int64_t bn,
                                         int64_t btrunc, int depth, fa_nmod_t mod);
FA_API int fa_nmod_poly_mul_mid_precomp(uint64_t* z, uint64_t zl, uint64_t zh, const uint64_t* a,
                                        int64_t an, fa_mul_precomp M, fa_nmod_t mod);
FA_API void fa_nmod_poly_mul_precomp_clear(fa_mul_precomp M);

/* _fmpz_poly_mul_mid_mpn_ctx, fmpz_poly_mul.c:319-401.  Coefficients are
   passed as fixed-width two's-complement limb vectors (aw, bw, zw limbs per
   coefficient), the form fmpz_get_signed_ui_array / fmpz_set_signed_ui_array
   (fmpz.h:238-239) convert to and from.  Returns 1 on success and 0 when
   the reference declines (more than 8 primes needed, fmpz_poly_mul.c:
   377-379), leaving z untouched; negative on error. */
FA_API int fa_fmpz_poly_mul_mid(uint64_t* z, uint64_t zw, int64_t zl, int64_t zh,
                         const uint64_t* a, uint64_t aw, int64_t an,
                         const uint64_t* b, uint64_t bw, int64_t bn);

/* ---------------------------------------------------------------------- */
/* plans and transformed operands, fft_small.h:530-700 (plan.c, op.c)      */
/* ---------------------------------------------------------------------- */

/* A plan: prime set, chunk size and transform depth for products of at
   most an_max x bn_max limbs (kind 0) or for the window [zl, zh) of a
   length-zn convolution mod n (kind 1), with up to len_bound products
   accumulated in transform space (fft_small_plan_struct, fft_small.h:530). */
typedef struct {
    int32_t kind;            /* 0: naturals (mpn), 1: polynomials mod n */
    int32_t np, offset;      /* primes offset .. offset + np - 1 */
    int32_t bits;            /* chunk bits (kind 0) */
    int32_t depth;           /* transform length 2^depth */
    int32_t pad;             /* kind 0: 1 when made for signed exports */
    uint64_t stride;         /* doubles per prime in an op */
    uint64_t zl, zh, zn;     /* output window; zn = coefficients of the convolution */
    uint64_t an_max, bn_max, len_bound;
    fa_nmod_t mod;
} fa_plan_struct;

#define FA_OP_PRIMAL 0       /* forward transforms (FFT_SMALL_OP_PRIMAL) */
#define FA_OP_PRODUCT 1      /* pointwise products, one normalization (FFT_SMALL_OP_PRODUCT) */
#define FA_OP_COEFF 2        /* after fa_ifft: ready for export */

/* A transformed operand in device memory: np arrays of `stride` doubles in
   this library's own layout (fft_small_op_struct, fft_small.h:620-640). */
typedef struct {
    double* data;
    uint64_t stride, np, offset, depth;
    int32_t domain;          /* FA_OP_* */
    int32_t owns_data;
} fa_op_struct;

/* fft_small_plan_init_mpn, plan.c:443: exact for up to len_bound
   accumulated products of an_max x bn_max limbs; is_signed leaves the spare
   factor of two the centred signed export needs (the plan's pad field
   records it) */
FA_API int fa_plan_init_mpn(fa_plan_struct* P, uint64_t an_max, uint64_t bn_max,
                            uint64_t len_bound, int is_signed);
/* fft_small_plan_init_mpn_np, plan.c:482-495: np = 0 the automatic choice,
   np = 2 forces two primes (fails when no two-prime plan is exact), any
   other value a plan of three or more primes */
FA_API int fa_plan_init_mpn_np(fa_plan_struct* P, uint64_t an_max, uint64_t bn_max,
                               uint64_t len_bound, int is_signed, uint64_t np);
/* fft_small_plan_init_nmod, plan.c:400: len_bound products of prod_bits-bit
   values per coefficient (2 * bits of n for a plain product) */
FA_API int fa_plan_init_nmod(fa_plan_struct* P, uint64_t zl, uint64_t zh, uint64_t zn,
                             uint64_t xtrunc_max, uint64_t len_bound, uint64_t prod_bits,
                             fa_nmod_t mod);
/* fft_small_plan_init_nmod_cyclic, plan.c:425-441: convolutions mod
   x^N - 1, N = 2^depth (depth >= 4), output coefficients [0, zh) */
FA_API int fa_plan_init_nmod_cyclic(fa_plan_struct* P, uint64_t depth, uint64_t zh,
                                    uint64_t len_bound, uint64_t prod_bits, fa_nmod_t mod);
/* fft_small_plan_init_mpn_cyclic, plan.c:497-647: products of naturals of
   at most *nn limbs mod 2^(64 *nn) - 1, *nn rounded up to a length with
   bits 2^depth = 64 nn and written back; fa_export_mpn gives a
   representative the caller folds (zn >= nn + (2 bits + depth +
   bits(len_bound)) / 64 + 1 limbs hold it) */
FA_API int fa_plan_init_mpn_cyclic(fa_plan_struct* P, uint64_t* nn, uint64_t len_bound);
FA_API void fa_plan_clear(fa_plan_struct* P);
/* fft_small_op_init / _init_borrowed / _sizeof_data / _clear, op.c:18-80
   (init allocates device memory; borrowed uses the caller's) */
FA_API int fa_op_init(fa_op_struct* X, const fa_plan_struct* P);
FA_API int fa_op_init_borrowed(fa_op_struct* X, const fa_plan_struct* P, void* data);
FA_API uint64_t fa_op_sizeof_data(const fa_plan_struct* P);
FA_API void fa_op_clear(fa_op_struct* X);
/* fft_small_fft_mpn / fft_small_fft_nmod (mpn_mul.c:2040, nmod_poly_mul.c:752):
   forward transform of a device operand into X.  An operand of the first
   factor has at most an_max limbs, of the second at most bn_max. */
FA_API int fa_fft_mpn(fa_op_struct* X, const uint64_t* a, int64_t an, const fa_plan_struct* P,
                      void* stream);
FA_API int fa_fft_nmod(fa_op_struct* X, const uint64_t* a, int64_t an, fa_nmod_t mod,
                       const fa_plan_struct* P, void* stream);
/* fft_small_op_mul/sqr/addmul/submul/add/sub/neg, op.c:272-418; Z may
   alias A and B */
FA_API int fa_op_mul(fa_op_struct* Z, const fa_op_struct* A, const fa_op_struct* B,
                     const fa_plan_struct* P, void* stream);
FA_API int fa_op_sqr(fa_op_struct* Z, const fa_op_struct* A, const fa_plan_struct* P,
                     void* stream);
FA_API int fa_op_addmul(fa_op_struct* Z, const fa_op_struct* A, const fa_op_struct* B,
                        const fa_plan_struct* P, void* stream);
FA_API int fa_op_submul(fa_op_struct* Z, const fa_op_struct* A, const fa_op_struct* B,
                        const fa_plan_struct* P, void* stream);
FA_API int fa_op_add(fa_op_struct* Z, const fa_op_struct* A, const fa_op_struct* B,
                     const fa_plan_struct* P, void* stream);
FA_API int fa_op_sub(fa_op_struct* Z, const fa_op_struct* A, const fa_op_struct* B,
                     const fa_plan_struct* P, void* stream);
FA_API int fa_op_neg(fa_op_struct* Z, const fa_op_struct* A, const fa_plan_struct* P,
                     void* stream);
/* fft_small_ifft, op.c:420: inverse transform in place */
FA_API int fa_ifft(fa_op_struct* X, const fa_plan_struct* P, void* stream);
/* fft_small_export_mpn (mpn_mul.c:2240): low zn limbs of the accumulated
   integer into device z */
FA_API int fa_export_mpn(uint64_t* z, int64_t zn, const fa_op_struct* X,
                         const fa_plan_struct* P, void* stream);
/* fft_small_export_mpn_signed (export_signed.c:608-612), signed plans:
   the first nslots coefficients centred to (-M/2, M/2] and summed at their
   bit offsets; z = zn limbs of |value| (64 zn >= nslots bits + 64 n + 2, n
   the limbs of M np), *sign (a device int32) = 1 when the value is
   negative */
FA_API int fa_export_mpn_signed(uint64_t* z, int64_t zn, int32_t* sign, const fa_op_struct* X,
                                uint64_t nslots, const fa_plan_struct* P, void* stream);
/* fft_small_export_mpn_signed_trunc (export_signed.c:367-606): limbs
   [lo, lo + zn) of |value| with the reference's truncation (64 (lo + zn) >=
   nslots bits + 64 n + 2): slots below the first that can reach limb lo
   are dropped (that slot index rounded down to a multiple of 256 on plans
   of more than two primes), and on those plans a negative value's window
   is rounded up, on two-prime plans down, as the reference's words are */
FA_API int fa_export_mpn_signed_trunc(uint64_t* z, int64_t zn, int32_t* sign,
                                      const fa_op_struct* X, uint64_t nslots, uint64_t lo,
                                      const fa_plan_struct* P, void* stream);
/* fft_small_export_nmod / _range (nmod_poly_mul.c:800-901): coefficients
   of the plan's window, or of [zl, zh), mod n into device z */
FA_API int fa_export_nmod(uint64_t* z, const fa_op_struct* X, fa_nmod_t mod,
                          const fa_plan_struct* P, void* stream);
FA_API int fa_export_nmod_range(uint64_t* z, const fa_op_struct* X, uint64_t zl, uint64_t zh,
                                fa_nmod_t mod, const fa_plan_struct* P, void* stream);

/* sd_fft_mpn_mulmod_2expp1_ctx (negacyclic.c:43-501): products in
   Z/(2^N + 1), N a multiple of 64, operands and results of n + 1 = N/64 + 1
   limbs with values in [0, 2^N], device memory.  m = 0 or a power of two
   >= 16 dividing N, checked as the reference checks it and kept; the
   device plan (the full product folded) does not depend on it. */
typedef struct {
    int64_t N, n, m;
    fa_plan_struct P;        /* transforms of one (n + 1)-limb operand */
} fa_mulmod_2expp1_ctx_struct;
/* sd_fft_mpn_mulmod_2expp1_ctx_init / _clear (negacyclic.c:65-230) */
FA_API int fa_mulmod_2expp1_ctx_init(fa_mulmod_2expp1_ctx_struct* C, int64_t N, int64_t m);
FA_API void fa_mulmod_2expp1_ctx_clear(fa_mulmod_2expp1_ctx_struct* C);
/* sd_fft_mpn_mulmod_2expp1 with a context (negacyclic.c:491-501) */
FA_API int fa_mulmod_2expp1(const fa_mulmod_2expp1_ctx_struct* C, uint64_t* z,
                            const uint64_t* x, const uint64_t* y, void* stream);
/* _transformed_size / _transform / _mul_cached (negacyclic.c:349-489): F
   (device, transformed_size doubles) caches y's transform for products
   z = x y mod 2^N + 1 */
FA_API int64_t fa_mulmod_2expp1_transformed_size(const fa_mulmod_2expp1_ctx_struct* C);
FA_API int fa_mulmod_2expp1_transform(const fa_mulmod_2expp1_ctx_struct* C, double* F,
                                      const uint64_t* y, void* stream);
FA_API int fa_mulmod_2expp1_mul_cached(const fa_mulmod_2expp1_ctx_struct* C, uint64_t* z,
                                       const uint64_t* x, const double* F, void* stream);

/* ---------------------------------------------------------------------- */
/* sd_fft_ctx plugin surface, fft_small.h:206-281                         */
/* ---------------------------------------------------------------------- */

typedef struct {
    double p;                          /* the FFT prime */
    double pinv;
    fa_nmod_t mod;
    uint64_t primitive_2power_root;    /* sd_fft_ctx.c:32-37 */
    int32_t depth;                     /* twiddle depth available on device */
    int32_t slot;                      /* library-internal table handle */
} fa_sd_fft_ctx_struct;

/* fft_small_mulmod_satisfies_bounds (mulmod_satisfies_bounds.c:120-150):
   1 when the FP64 mulmod's range claims hold for the odd modulus n */
FA_API int fa_mulmod_satisfies_bounds(uint64_t n);
/* sd_fft_ctx_init_prime, sd_fft_ctx.c:56-154 */
FA_API int fa_sd_fft_ctx_init_prime(fa_sd_fft_ctx_struct* Q, uint64_t pp);
/* sd_fft_ctx_clear, sd_fft_ctx.c:16-26 */
FA_API void fa_sd_fft_ctx_clear(fa_sd_fft_ctx_struct* Q);
/* sd_fft_trunc, sd_fft.c:1230-1274 (host data, in place) */
FA_API int fa_sd_fft_trunc(fa_sd_fft_ctx_struct* Q, double* d, uint64_t L,
                    uint64_t itrunc, uint64_t otrunc);
/* sd_ifft_trunc, sd_ifft.c:1503-1533 (host data, in place) */
FA_API int fa_sd_ifft_trunc(fa_sd_fft_ctx_struct* Q, double* d, uint64_t L, uint64_t trunc);
/* sd_fft_ctx_point_mul / _point_sqr (mpn_mul.c:1392-1452) on host arrays
   of 2^depth doubles (depth >= 4): a[i] = a[i] m b[i] (or a[i] m a[i]) mod
   p, entries returned in [0, p) */
FA_API int fa_sd_fft_ctx_point_mul(fa_sd_fft_ctx_struct* Q, double* a, const double* b,
                                   uint64_t m, uint64_t depth);
FA_API int fa_sd_fft_ctx_point_sqr(fa_sd_fft_ctx_struct* Q, double* a, uint64_t m,
                                   uint64_t depth);
/* sd_fft_ctx_w, fft_small.h:308-312 */
FA_API double fa_sd_fft_ctx_w(const fa_sd_fft_ctx_struct* Q, uint64_t j);

/* ---------------------------------------------------------------------- */
/* library state                                                          */
/* ---------------------------------------------------------------------- */

FA_API const char* fa_last_error(void);
FA_API int fa_init(void);                   /* optional: create device tables now */
FA_API int fa_default_primes(uint64_t* out8);
/* synchronize the calling thread's stream */
FA_API int fa_sync(void);
/* the calling thread's stream (hipStream_t) */
FA_API void* fa_stream(void);

/* seeded synthetic inputs (SplitMix64 -> xoshiro256**), identical to the
   generator the tests use, written on the host */
FA_API void fa_synth_limbs(uint64_t seed, uint64_t* a, uint64_t n, int top_nonzero);
/* both operands of one product from one stream: a, then b, top limbs nonzero */
FA_API void fa_synth_mpn_pair(uint64_t seed, uint64_t* a, uint64_t an, uint64_t* b, uint64_t bn);

/* the request generator of the batched workloads (the same xoshiro256**
   stream seeded by SplitMix64; every draw advances it) */
typedef struct { uint64_t s[4]; } fa_rng;
FA_API void fa_rng_seed(fa_rng* r, uint64_t seed);
FA_API uint64_t fa_rng_next(fa_rng* r);
/* uniform in [lo, hi], by rejection */
FA_API uint64_t fa_rng_uniform(fa_rng* r, uint64_t lo, uint64_t hi);
/* uniform in [0, 1): the top 53 bits of one draw */
FA_API double fa_rng_double(fa_rng* r);
/* floor(2^U) with U uniform in [x, y) */
FA_API uint64_t fa_rng_log_uniform(fa_rng* r, double x, double y);
/* a prime of `bits` bits: odd uniform draws until Miller-Rabin accepts */
FA_API uint64_t fa_rng_random_prime(fa_rng* r, int bits);
/* n values uniform in [0, bound) */
FA_API void fa_rng_fill_below(fa_rng* r, uint64_t* out, uint64_t n, uint64_t bound);
/* n limbs, the last redrawn until nonzero if top_nonzero */
FA_API void fa_rng_fill_limbs(fa_rng* r, uint64_t* out, uint64_t n, int top_nonzero);
/* n coefficients uniform in (-2^bits, 2^bits), w-limb two's complement */
FA_API void fa_rng_fill_signed(fa_rng* r, uint64_t* out, uint64_t n, int bits, uint64_t w);

/* The mixed stream at its full size, generated and checked on the device (no
   counterpart in the reference).  Request headers come from the stream above
   (seed 0xF1A70005), per request in this order: a double picks the kind
   (< 0.5 integer, < 0.8 nmod, else fmpz); integer: an, bn log-uniform in
   [2^4, 2^22]; nmod: an, bn log-uniform in [2^6, 2^16], modulus bits uniform
   in [20, 63], the prime; fmpz: an, bn uniform in [1024, 16384], b uniform in
   [64, 4096]; then key[0], key[1].  Operand s holds the SplitMix64 sequence
   seeded with key[s] (output j = fa_stream_word(key, j)):
     integer limb j       = word j (the top limb 1 if its word is 0)
     nmod coefficient j   = floor(word j * n / 2^64), in [0, n)
     fmpz coefficient j   = t - (2^b - 1), t the (b+1)-bit number of words
                            j tw .. j tw + tw - 1 (tw = ceil((b+1)/64), the
                            last cut to its low bits), 0 when all ones; in
                            (-2^b, 2^b) as b / 64 + 1 limbs of two's complement */
typedef struct {
    int32_t kind;          /* FA_REQ_* */
    int32_t bits;          /* nmod: modulus bits; fmpz: coefficient bits b */
    int64_t an, bn;        /* limbs (integers) or coefficients */
    uint64_t n;            /* nmod modulus */
    uint64_t key[2];       /* seeds of the operands' contents */
} fa_stream_req;
FA_API int fa_stream_headers(uint64_t seed, int64_t count, fa_stream_req* out);
FA_API uint64_t fa_stream_word(uint64_t key, uint64_t j);

/* one operand to write on the device: count limbs (integer, nmod) or
   coefficients of w limbs (fmpz) at dst, contents as above */
typedef struct {
    uint64_t* dst;
    uint64_t count;
    uint64_t key;
    uint64_t n;            /* nmod modulus */
    uint64_t w;            /* fmpz limbs per coefficient (>= bits / 64 + 1) */
    int32_t kind;          /* FA_REQ_* */
    int32_t bits;          /* fmpz coefficient bits */
} fa_fill_spec;
/* writes n operands in one launch on `stream` and waits for it */
FA_API int fa_stream_fill_device(const fa_fill_spec* spec, int64_t n, void* stream);

/* value at x mod m (2 <= m < 2^64) of count coefficients of w limbs each at
   the device address src (two's complement when is_signed): sum_i c_i x^i;
   an integer of count limbs is count coefficients of one limb at x = 2^64 */
typedef struct {
    const uint64_t* src;
    uint64_t count;
    uint64_t w;
    uint64_t x;
    uint64_t m;
    int32_t is_signed;
    int32_t pad;
} fa_eval_spec;
/* n evaluations on `stream`; out (host) receives the n values when it returns */
FA_API int fa_eval_device(const fa_eval_spec* spec, int64_t n, uint64_t* out, void* stream);

/* ---------------------------------------------------------------------- */
/* measurement support                                                    */
/* ---------------------------------------------------------------------- */

/* per-kernel timing with HIP events recorded around each launch on the
   launching stream (calling thread only) */
FA_API void fa_prof_enable(int on);
FA_API void fa_prof_reset(void);
/* names: max*64 chars; ms: total milliseconds; counts: launches */
FA_API int fa_prof_read(char* names, double* ms, int64_t* counts, int max);
/* selects the register radix-16 kernels (1, default) or the generic LDS
   radix-4 kernels (0) for transform lengths 2^16, 2^20, 2^24 */
FA_API void fa_set_fast_kernels(int on);
/* column-pass tiles at 2^24 points: 1 = 8 columns (default), 0 = 4 columns */
FA_API void fa_set_col_variant(int v);
/* truncation of transforms of 2^13 .. 2^23 points at the rows of the
   product's length (row pass skipped past it, truncated inverse column
   transform): 0 never; 1 (default) automatic: single products and batch
   requests with FA_TRUNC_AUTO truncate only at 2^23 points with at most 55%
   of the rows (elsewhere the truncated inverse, which runs in LDS, costs
   more than the skipped rows save on MI355X); 2 wherever it applies.
   Tuning and test switch; outputs are identical. */
FA_API void fa_set_truncation(int mode);
/* row pass variant of the two-level transforms: 1 = last-stage twiddles
   formed per radix-4 layer, three workgroups per CU (default); 0 = all 15
   held in registers, two workgroups per CU.  Tuning switch, no counterpart
   in the reference. */
FA_API void fa_set_row_variant(int v);
/* integer products whose operands together exceed `limbs` (0: only those no
   single 2^24-point plan covers, the default) are computed as block products
   added into the output on the device.  Lets tests exercise the block path
   at small sizes; no counterpart in the reference. */
FA_API void fa_set_mpn_block_cap(uint64_t limbs);
/* nmod and fmpz products of more than n output coefficients (0: more than
   one 2^24-point transform holds, the default) are computed as block
   products whose shares of the output window are added on the device.  Lets
   tests exercise the block path at small sizes; no counterpart in the
   reference. */
FA_API void fa_set_poly_block_cap(uint64_t n);
/* 2^24-point products: the primes split into v groups (2 or 3; default 2),
   each on a stream of its own forked from and joined back to the caller's,
   so one group's row and inverse passes run beside the next group's forward
   passes; 0 or 1 = one launch per pass over all primes.  Not used while the
   stream is being captured into a graph or kernel timing is on.  Tuning
   switch, no counterpart in the reference. */
FA_API void fa_set_prime_groups(int v);
/* squares (a == b, an == bn; mpn_mul.c:1665, nmod_poly_mul.c:939,
   fmpz_poly_mul.c:352): 1 = one forward transform per prime and a pointwise
   square (default), 0 = the general product.  Tuning switch for tests. */
FA_API void fa_set_square_route(int on);
/* CRT of integer products: 1 = radix-2^24 digits on the FP64 lanes
   (default), 0 = 64-bit limb arithmetic */
FA_API void fa_set_crt_variant(int v);
/* plain device-to-device copy kernel (nbytes multiple of 16), used to
   calibrate hardware byte counters */
FA_API int fa_copy_device(void* dst, const void* src, uint64_t nbytes, void* stream);
/* cap on the device bytes the library's grow-only buffers may hold
   together (0: no cap, the default).  A call that would grow past it fails
   with FA_E_NOMEM, as when HBM runs out.  No counterpart in the reference. */
FA_API void fa_set_device_memory_limit(uint64_t bytes);

/* ------------------------------------------------------------------------
   Batches of heterogeneous products (no counterpart in the reference: the
   reference multiplies one product per call).  Every request is planned as
   its single-product entry point would plan it (mpn_mul_default_mpn_ctx,
   _nmod_poly_mul_mid_default_mpn_ctx, _fmpz_poly_mul_mid_default_mpn_ctx)
   and gives the identical output; the batch runs as one recorded sequence of
   binned launches, optionally captured into a hipGraph and replayed.
   ------------------------------------------------------------------------ */
enum { FA_REQ_MPN = 0, FA_REQ_NMOD = 1, FA_REQ_FMPZ = 2 };
enum {
    FA_BATCH_DEVICE = 1,   /* a, b, z are device addresses; results stay in HBM */
    FA_BATCH_GRAPH = 2,    /* capture the launch sequence into a hipGraph */
    FA_BATCH_FMPZ_KS = 4   /* fmpz products the eight-prime rule declines are computed by
                              Kronecker substitution (status 1), as the reference's
                              callers do (fmpz_poly/mul.c:160), instead of returning 0 */
};

typedef struct {
    int32_t kind;          /* FA_REQ_* */
    int32_t status;        /* out: 1 done, 0 declined (fmpz needing more than 8
                              primes, as the reference), < 0 an FA_E_* error */
    const uint64_t* a;
    const uint64_t* b;
    int64_t an, bn;        /* limbs (mpn) or coefficients (nmod, fmpz) */
    int64_t aw, bw;        /* fmpz: limbs per two's complement coefficient */
    uint64_t* z;           /* mpn: an + bn limbs; nmod: zh - zl values;
                              fmpz: (zh - zl) * zw limbs */
    int64_t zl, zh, zw;    /* nmod / fmpz output window and width */
    fa_nmod_t mod;         /* nmod modulus (nmod_init) */
    int64_t abits, bbits;  /* fmpz: _fmpz_vec_max_bits of a and b; 0, 0 = measure
                              (host pointers only) */
    int32_t np, bits;      /* mpn plan override (0 = choose); out: plan used */
    int32_t depth;         /* out: transform depth */
    int32_t trunc;         /* FA_TRUNC_*: transforms of 2^13..2^23 points truncated at the
                              rows of the product (outputs identical either way) */
} fa_request;
enum { FA_TRUNC_AUTO = 0, FA_TRUNC_OFF = 1, FA_TRUNC_ON = 2 };

/* the fmpz product the reference's callers compute when the eight-prime
   path declines, by Kronecker substitution onto the integer product
   (fmpz_poly/mul.c:160, _fmpz_poly_mul_KS); arguments and output as
   fa_fmpz_poly_mul_mid; always computes (returns 1) */
/* Kronecker route of fa_fmpz_poly_mul_ks and of FA_BATCH_FMPZ_KS batches
   (fixed when a batch is created): 1 = signed slots, one integer product
   (default); 0 = three integer products of the non-negative parts.
   Tuning switch, no counterpart in the reference. */
FA_API void fa_set_ks_variant(int v);
FA_API int fa_fmpz_poly_mul_ks(uint64_t* z, uint64_t zw, int64_t zl, int64_t zh,
                               const uint64_t* a, uint64_t aw, int64_t an, const uint64_t* b,
                               uint64_t bw, int64_t bn);

/* _fmpz_vec_max_bits (fmpz_vec.h) of n w-limb two's complement
   coefficients: the largest bit size of |c|, negated if some c < 0 */
FA_API int64_t fa_fmpz_vec_max_bits(const uint64_t* v, uint64_t w, int64_t n);

typedef struct fa_batch_s* fa_batch;

/* plans and lays out a batch; with FA_BATCH_GRAPH records the launches.
   Statuses are written into the requests.  The request array and, for
   FA_BATCH_DEVICE, the buffers it names must outlive the batch. */
FA_API int fa_batch_create(fa_batch* out, fa_request* reqs, int64_t n, int flags);
/* runs the batch on `stream` (NULL: the batch's own stream).  Host-pointer
   batches copy in, run and copy out, and return when the outputs are in host
   memory; device batches given a stream return without waiting. */
FA_API int fa_batch_launch(fa_batch b, void* stream);
FA_API void fa_batch_destroy(fa_batch b);
/* create + launch + destroy for host-pointer requests */
FA_API int fa_batch_mul(fa_request* reqs, int64_t n, int flags);

/* ------------------------------------------------------------------------
   Plan space, autotuner and learned planner for integer products (no
   counterpart in the reference, whose fixed rule is mpn_ctx_best_profile,
   mpn_mul.c:1248-1340).  Candidate j < 56 of an x bn limbs is
   (np, bits, trunc) = (j / 8 + 2, bmax(np) - 8 ((j / 2) mod 4), j mod 2),
   bmax(np) the largest chunk size np primes carry exactly, trunc = 1 the
   transform truncated at the product's rows (valid only where that
   applies); candidates 56..63 and rejected sizes are (0, 0, 0).
   ------------------------------------------------------------------------ */
FA_API int fa_mpn_candidates(int64_t an, int64_t bn,
                             int32_t* out /* 64 x {np, bits, trunc} */);
/* 16 planner features of a shape (log sizes, the automatic plan, bmax) */
FA_API int fa_plan_features(int64_t an, int64_t bn, float* f16);
/* exhaustive autotuner: device time (ms) per product of every candidate,
   as batched products: a graph-captured device batch of K copies of the
   product per candidate (fresh SplitMix64 operands), the batches replayed
   `reps` times in turn (replay r of every candidate before replay r + 1 of
   any), the median replay per product (host launch overhead excluded); -1
   for invalid candidates */
FA_API int fa_autotune_mpn(int64_t an, int64_t bn, int reps, float* ms64);

/* FP32 MLP 16 -> 64 -> 64 -> 64 predicting log(ms) of the 64 candidates */
typedef struct fa_mlp_s* fa_mlp;
FA_API int fa_mlp_create(fa_mlp* out, uint64_t seed);
FA_API void fa_mlp_destroy(fa_mlp m);
FA_API int fa_mlp_param_count(void);
FA_API int fa_mlp_get_params(fa_mlp m, float* host);
FA_API int fa_mlp_set_params(fa_mlp m, const float* host);
FA_API int fa_mlp_forward(fa_mlp m, const float* X, int B, float* Y);
/* loss (masked mean square error) and its gradient at the current parameters */
FA_API int fa_mlp_loss_grad(fa_mlp m, const float* X, const float* T, const float* M, int B,
                            float* loss, float* grad);
/* minibatch Adam on the device; loss_out[e] = mean loss of epoch e */
FA_API int fa_mlp_train(fa_mlp m, const float* X, const float* T, const float* M, int64_t N,
                        int batch, int epochs, float lr, float* loss_out);
/* learned plans for n shapes (an, bn): out = n x {np, bits, trunc} */
FA_API int fa_mlp_plan_mpn(fa_mlp m, const int64_t* shapes, int n, int32_t* out);

#ifdef __cplusplus
}
#endif
#endif
