/*
 * fsmall_oracle.h -- CPU restatement of FLINT's fft_small multiplication
 * path (src/fft_small/ in flintlib/flint 3.7.0-dev).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the GPU
 * product: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in flint_amd/ links or calls it.
 *
 * Arithmetic: exact u64/u128 modular arithmetic on canonical residues for
 * the transforms, CRT and reductions, so the oracle does not share the
 * FP64 carrier trick with the GPU path; the FP64 primitives of the
 * reference (machine_vectors.h:646-704) are restated separately
 * (or_fp_*) and checked against u128 arithmetic.
 */
#ifndef FSMALL_ORACLE_H
#define FSMALL_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t n;
    uint64_t ninv;
    uint64_t norm;
} or_nmod_t;                       /* layout of FLINT's nmod_t (flint.h.in:441-447) */

/* primes and primitives */
int      or_is_prime(uint64_t n);
uint64_t or_next_fft_number(uint64_t p);
void     or_default_primes(uint64_t* out /* [8] */);
int      or_mulmod_satisfies_bounds(uint64_t n);
double   or_fp_mulmod(double a, double b, double n, double ninv);
double   or_fp_reduce_to_pm1n(double a, double n, double ninv);
double   or_fp_reduce_to_0n(double a, double n, double ninv);
uint64_t or_quadratic_nonresidue(uint64_t n);
uint64_t or_primitive_2power_root(uint64_t p);
uint64_t or_w(uint64_t p, uint64_t j);          /* sd_fft_ctx_w */
uint64_t or_w2(uint64_t p, uint64_t j);         /* sd_fft_ctx_w2 */
uint64_t or_trunc_index(uint64_t L, uint64_t i);

/* transforms on canonical residues (uint64 in [0,p)); data length 2^L */
int or_fft_trunc(uint64_t p, uint64_t* d, uint64_t L, uint64_t itrunc, uint64_t otrunc);
int or_ifft_trunc(uint64_t p, uint64_t* d, uint64_t L, uint64_t trunc);

/* CRT tables and plans */
uint64_t or_crt_bn_bound(int np, uint64_t bits);
uint64_t or_crt_find_bits(int np, uint64_t bn);
int      or_best_profile(uint64_t an, uint64_t bn, uint64_t nthreads,
                         uint64_t* np, uint64_t* bits);
int      or_np2_choose(uint64_t an, uint64_t bn, uint64_t len_bound,
                       uint64_t* bits, uint64_t* depth);
int      or_mpn_plan(uint64_t an, uint64_t bn, uint64_t nthreads, uint64_t* out /* [6] */);

/* limb splitting (chunk i = bits [i*bits, (i+1)*bits) of a) mod p */
void or_mpn_to_residues(uint64_t p, uint64_t* d, uint64_t dn,
                        const uint64_t* a, uint64_t an, uint64_t bits);

/* the three products */
int or_mpn_mul(uint64_t* z, const uint64_t* a, uint64_t an,
               const uint64_t* b, uint64_t bn, uint64_t nthreads);
void or_mpn_mul_basecase(uint64_t* z, const uint64_t* a, uint64_t an,
                         const uint64_t* b, uint64_t bn);
/* _mpn_ctx_mpn_mul_range's limbs [lo, hi) with FLINT on nthreads threads */
int or_mpn_mul_range(uint64_t* z, uint64_t lo, uint64_t hi, const uint64_t* a, uint64_t an,
                     const uint64_t* b, uint64_t bn, uint64_t nthreads);
int or_nmod_poly_mul_mid(uint64_t* z, int64_t zl, int64_t zh,
                         const uint64_t* a, int64_t an,
                         const uint64_t* b, int64_t bn, or_nmod_t mod);
int or_fmpz_poly_mul_mid(uint64_t* z, uint64_t zw, int64_t zl, int64_t zh,
                         const uint64_t* a, uint64_t aw, int64_t an,
                         const uint64_t* b, uint64_t bw, int64_t bn);
int or_fmpz_poly_nprimes(const uint64_t* a, uint64_t aw, int64_t an,
                         const uint64_t* b, uint64_t bw, int64_t bn);

/* seeded synthetic inputs shared by tests and bench (SplitMix64 -> xoshiro256**) */
typedef struct { uint64_t s[4]; } or_rng;
void     or_rng_seed(or_rng* r, uint64_t seed);
uint64_t or_rng_next(or_rng* r);
uint64_t or_rng_uniform(or_rng* r, uint64_t lo, uint64_t hi);   /* [lo, hi] */
uint64_t or_rng_random_prime(or_rng* r, uint64_t bits);
void     or_rng_limbs(or_rng* r, uint64_t* a, uint64_t n, int top_nonzero);
double   or_rng_double(or_rng* r);
uint64_t or_rng_log_uniform(or_rng* r, double x, double y);
void     or_rng_fill_below(or_rng* r, uint64_t* out, uint64_t n, uint64_t bound);
void     or_rng_fill_signed(or_rng* r, uint64_t* out, uint64_t n, int bits, uint64_t w);

/* the full mixed stream (headers: 7 words each; operand contents) */
uint64_t or_stream_word(uint64_t key, uint64_t j);
void     or_stream_headers(uint64_t seed, uint64_t count, uint64_t* out);
void     or_stream_fill(int kind, uint64_t key, uint64_t count, uint64_t n, uint64_t w, int bits,
                        uint64_t* out);

/* full-size checkers: residues mod a word-size q */
uint64_t or_mpn_mod_q(const uint64_t* a, uint64_t n, uint64_t q);
uint64_t or_nmod_poly_eval(const uint64_t* c, uint64_t n, uint64_t x, uint64_t q);
uint64_t or_fmpz_poly_eval_mod_q(const uint64_t* c, uint64_t w, uint64_t n, uint64_t x, uint64_t q);

#ifdef __cplusplus
}
#endif
#endif
