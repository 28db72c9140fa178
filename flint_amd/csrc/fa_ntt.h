// fa_ntt.h -- block-twisted number-theoretic transforms on LDS-resident data.
//
// Convention (fft_small.h:124-199, sd_fft.c:292-355): a block q of length
// 2^k at level s holds a polynomial modulo X^(2^k) - w[q]; one level splits
// it with the root r = w2[q] = w[2q] into
//     (lo + r*hi, lo - r*hi)   -> blocks 2q and 2q+1 of level s+1,
// so after all levels position i holds the evaluation at w[i] (bit-reversed
// order).  A sub-transform that starts at global block R uses, at its local
// level m and local block q, the table entry w2[(R << m) + q].
//
// Two levels are fused into one radix-4 butterfly (sd_fft.c:292-355 and
// sd_ifft.c:235-292); the value-range invariants are those proved in
// mulmod_satisfies_bounds.c:152-219:
//   forward: inputs in (-3n, 3n) -> outputs in (-3n, 3n)
//   inverse: inputs in (-2n, 2n) -> outputs in (-2n, 2n)
// Twiddle tables hold signed representatives in [-n/2, n/2].
#pragma once

#include "fa_arith.h"

// forward radix-4 over levels (m, m+1) of block Q:
// W2 = w2[Q], W = w2[2Q], IW = w2[2Q+1]
FA_DEV void fa_fwd4(double& a0, double& a1, double& a2, double& a3,
                    double W2, double W, double IW, double n, double ninv)
{
    a0 = fa_reduce_pm1n(a0, n, ninv);
    double t2 = fa_mulmod(a2, W2, n, ninv);
    double t3 = fa_mulmod(a3, W2, n, ninv);
    double u0 = a0 + t2, u2 = a0 - t2;
    double v1 = fa_mulmod(a1 + t3, W, n, ninv);
    double v3 = fa_mulmod(a1 - t3, IW, n, ninv);
    a0 = u0 + v1;
    a1 = u0 - v1;
    a2 = u2 + v3;
    a3 = u2 - v3;
}

// forward radix-2 (a single last level): (x, y) -> (x + r y, x - r y)
FA_DEV void fa_fwd2(double& a0, double& a1, double r, double n, double ninv)
{
    double x = fa_reduce_pm1n(a0, n, ninv);
    double t = fa_mulmod(a1, r, n, ninv);
    a0 = x + t;
    a1 = x - t;
}

// inverse radix-4 over levels (m+1, m) of block Q (unnormalized, factor 4):
// V2 = w2[Q]^-1, V = w2[2Q]^-1, IV = w2[2Q+1]^-1
FA_DEV void fa_inv4(double& b0, double& b1, double& b2, double& b3,
                    double V2, double V, double IV, double n, double ninv)
{
    double u = b0 + b1, v = b2 + b3;
    double s1 = fa_mulmod(b0 - b1, V, n, ninv);
    double s2 = fa_mulmod(b2 - b3, IV, n, ninv);
    b0 = fa_reduce_pm1n(u + v, n, ninv);
    b2 = fa_mulmod(u - v, V2, n, ninv);
    b1 = s1 + s2;
    b3 = fa_mulmod(s1 - s2, V2, n, ninv);
}

// inverse radix-2 (factor 2): (u, v) -> (u + v, (u - v) r^-1)
FA_DEV void fa_inv2(double& b0, double& b1, double rinv, double n, double ninv)
{
    double u = b0, v = b1;
    b0 = fa_reduce_pm1n(u + v, n, ninv);
    b1 = fa_mulmod(u - v, rinv, n, ninv);
}

// Forward transform of 2^k points for each of 2^lgC interleaved columns:
// element e of column c lives at s[(e << lgC) + c].  Root offset R.
// Must be called by all threads of the block; ends with a barrier.
FA_DEV void fa_lds_fwd(double* s, int k, int lgC, uint64_t R,
                       const double* __restrict__ w2, double n, double ninv)
{
    const int nt = blockDim.x, tid = threadIdx.x;
    const int C = 1 << lgC;
    int m = 0;
    for (; m + 2 <= k; m += 2) {
        const int shift = k - m - 2;
        const int qu = 1 << shift;
        const int nwork = (1 << (k - 2)) << lgC;
        for (int u = tid; u < nwork; u += nt) {
            int c = u & (C - 1);
            int t = u >> lgC;
            int ql = t >> shift;
            int i = t & (qu - 1);
            int e = (ql << (shift + 2)) + i;
            uint64_t Q = (R << m) + (uint64_t)ql;
            double W2 = w2[Q], W = w2[2 * Q], IW = w2[2 * Q + 1];
            double* p0 = s + ((size_t)e << lgC) + c;
            const size_t st = (size_t)qu << lgC;
            double a0 = p0[0], a1 = p0[st], a2 = p0[2 * st], a3 = p0[3 * st];
            fa_fwd4(a0, a1, a2, a3, W2, W, IW, n, ninv);
            p0[0] = a0; p0[st] = a1; p0[2 * st] = a2; p0[3 * st] = a3;
        }
        __syncthreads();
    }
    if (m < k) {
        const int nwork = (1 << (k - 1)) << lgC;
        for (int u = tid; u < nwork; u += nt) {
            int c = u & (C - 1);
            int t = u >> lgC;
            uint64_t Q = (R << m) + (uint64_t)t;
            double* p0 = s + ((size_t)(2 * t) << lgC) + c;
            double a0 = p0[0], a1 = p0[C];
            fa_fwd2(a0, a1, w2[Q], n, ninv);
            p0[0] = a0; p0[C] = a1;
        }
        __syncthreads();
    }
}

// Inverse of fa_lds_fwd (outputs scaled by 2^k).
FA_DEV void fa_lds_inv(double* s, int k, int lgC, uint64_t R,
                       const double* __restrict__ w2inv, double n, double ninv)
{
    const int nt = blockDim.x, tid = threadIdx.x;
    const int C = 1 << lgC;
    int m = k;
    if (k & 1) {
        m = k - 1;
        const int nwork = (1 << (k - 1)) << lgC;
        for (int u = tid; u < nwork; u += nt) {
            int c = u & (C - 1);
            int t = u >> lgC;
            uint64_t Q = (R << m) + (uint64_t)t;
            double* p0 = s + ((size_t)(2 * t) << lgC) + c;
            double a0 = p0[0], a1 = p0[C];
            fa_inv2(a0, a1, w2inv[Q], n, ninv);
            p0[0] = a0; p0[C] = a1;
        }
        __syncthreads();
    }
    for (m -= 2; m >= 0; m -= 2) {
        const int shift = k - m - 2;
        const int qu = 1 << shift;
        const int nwork = (1 << (k - 2)) << lgC;
        for (int u = tid; u < nwork; u += nt) {
            int c = u & (C - 1);
            int t = u >> lgC;
            int ql = t >> shift;
            int i = t & (qu - 1);
            int e = (ql << (shift + 2)) + i;
            uint64_t Q = (R << m) + (uint64_t)ql;
            double V2 = w2inv[Q], V = w2inv[2 * Q], IV = w2inv[2 * Q + 1];
            double* p0 = s + ((size_t)e << lgC) + c;
            const size_t st = (size_t)qu << lgC;
            double a0 = p0[0], a1 = p0[st], a2 = p0[2 * st], a3 = p0[3 * st];
            fa_inv4(a0, a1, a2, a3, V2, V, IV, n, ninv);
            p0[0] = a0; p0[st] = a1; p0[2 * st] = a2; p0[3 * st] = a3;
        }
        __syncthreads();
    }
}
