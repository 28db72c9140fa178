// fa_arith.h -- exact modular arithmetic on FP64 lanes for gfx950.
//
// Integers below 2^53 are carried exactly in doubles.  Every kernel that
// includes this header is compiled with -ffp-contract=off so that the
// explicit fma() calls below are the only fused operations.
//
// mulmod follows machine_vectors.h:694-699 (vec1d_mulmod):
//   h = a*b, q = rint(h/n), l = fma(a,b,-h), r = fma(-q,n,h) + l
// and, for a prime n accepted by fft_small_mulmod_satisfies_bounds
// (mulmod_satisfies_bounds.c:120-150), maps products in (-2n^2, 2n^2) to
// (-n, n) and products in (-4n^2, 4n^2) to (-3n/2, 3n/2).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FA_DEV __device__ __forceinline__

FA_DEV double fa_mulmod(double a, double b, double n, double ninv)
{
    double h = a * b;
    double q = __builtin_rint(h * ninv);
    double l = __builtin_fma(a, b, -h);
    return __builtin_fma(-q, n, h) + l;
}

// a mod n in [-n/2, n/2] (machine_vectors.h:646-651 reduce_to_pm1n)
FA_DEV double fa_reduce_pm1n(double a, double n, double ninv)
{
    return __builtin_fma(-__builtin_rint(a * ninv), n, a);
}

// a mod n in [0, n) as an integer (machine_vectors.h:662-667 reduce_to_0n)
FA_DEV uint64_t fa_reduce_0n_u64(double a, double n, double ninv)
{
    double r = fa_reduce_pm1n(a, n, ninv);
    if (r < 0.0) r += n;
    int64_t v = (int64_t)r;
    return (uint64_t)v;
}

// 64x64 -> 128 helpers
FA_DEV uint64_t fa_umulhi(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// (hi*2^64 + lo) mod n with hi < n, by the Moller-Granlund preinverse
// (udiv_qrnnd_preinv; dinv = floor((2^128-1)/d) - 2^64 for d = n << norm,
// the ninv of FLINT's nmod_t, nmod.h:284)
FA_DEV uint64_t fa_ll_mod_preinv(uint64_t hi, uint64_t lo, uint64_t n, uint64_t dinv, int norm)
{
    uint64_t d = n << norm;
    if (norm) {
        hi = (hi << norm) | (lo >> (64 - norm));
        lo <<= norm;
    }
    uint64_t p_lo = dinv * hi;
    uint64_t p_hi = __umul64hi(dinv, hi);
    uint64_t q0 = p_lo + lo;
    uint64_t q1 = p_hi + hi + 1 + (q0 < p_lo);
    uint64_t r = lo - q1 * d;
    if (r > q0) r += d;
    if (r >= d) r -= d;
    return r >> norm;
}

// a b mod n for a, b < n (n < 2^64 with its nmod_t preinverse)
FA_DEV uint64_t fa_mulmod_preinv(uint64_t a, uint64_t b, uint64_t n, uint64_t dinv, int norm)
{
    return fa_ll_mod_preinv(__umul64hi(a, b), a * b, n, dinv, norm);
}

struct fa_prime_dev {
    double p;
    double pinv;
};
