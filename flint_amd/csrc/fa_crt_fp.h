// fa_crt_fp.h -- CRT recombination of an integer product on the FP64 lanes.
//
// For a coefficient c < M = p_0 ... p_(np-1) whose residues after the
// inverse transforms are x_i = c (M/p_i)^-1 mod p_i (the normalizer of the
// pointwise product carries (M/p_i)^-1), the explicit CRT gives
//     c = sum_i x_i (M/p_i) - k M,   k = floor(sum_i x_i / p_i)
// (mpn_mul.c:1120-1180 restates it with 64-bit limbs).  Any representative
// of x_i serves: S = sum_i x_i (M/p_i) is c + k M for k = floor(S / M).  The
// kernel takes the one in [-p_i/2, p_i/2].  The products x_i (M/p_i) are
// formed in radix 2^24 in doubles: x_i = u_i + v_i 2^24 with 0 <= u_i < 2^24
// and |v_i| <= 2^25 (v_i = floor(x_i / 2^24)), and (M/p_i) has 24-bit digits
// e_(i,d); every term u e < 2^48 and |v e| < 2^49 is an exact FMA, and a
// column collects at most six primes' terms, 6 (2^48 + 2^49) < 2^53, before
// it is carried into 24-bit digits (signed floor division).
// k = rint(sum_i x_i / p_i) is k_true or k_true + 1 (the sum's rounding
// error is far below 1/2), so c' = S - k M lies in [-M, M) and one
// conditional + M fixes the sign.
#pragma once

#include "fa_internal.h"
#include "fa_arith.h"

template <int NP>
struct fa_crt_fp_dims {
    static constexpr int ND = (50 * NP + 23) / 24;         // digits of M
    static constexpr int NDC = (50 * (NP - 1) + 23) / 24;  // digits of M / p_i
    static constexpr int NL = (50 * NP + 63) / 64;         // 64-bit limbs of c < M
};

// constants of the radix-2^24 CRT, filled on the host
struct fa_crt_fp_consts {
    double p[FA_NPRIMES];
    double pinv[FA_NPRIMES];
    double rp[FA_NPRIMES];
    double cof[FA_NPRIMES][16];   // 24-bit digits of M / p_i
    double M[18];                 // 24-bit digits of M
    double Mh[18];                // 24-bit digits of floor(M / 2)
};

void fa_crt_fp_consts_init(fa_crt_fp_consts* F, const fa_crt_consts& C);

// carries the columns col[0..ND) into 24-bit digits (signed floor
// division); returns the carry out of the top digit
template <int ND>
FA_DEV double fa_crt_fp_carry(double* col)
{
    const double B = 16777216.0, Binv = 1.0 / 16777216.0;
    double carry = 0.0;
#pragma unroll
    for (int m = 0; m < ND; m++) {
        const double v = col[m] + carry;
        carry = __builtin_floor(v * Binv);
        col[m] = __builtin_fma(-carry, B, v);
    }
    return carry;
}

// coefficient j -> its NL little-endian 64-bit limbs.  MODE 0: c in
// [0, M); for the centred lift c~ = c - M when c > M/2 (the signed exports,
// export_signed.c:60-80), MODE 1 gives max(c~, 0) and MODE 2 max(-c~, 0).
template <int NP, int MODE = 0>
FA_DEV void fa_crt_fp_value(const double* __restrict__ Q, uint64_t stride, uint64_t j,
                            const fa_crt_fp_consts& F, uint64_t* L)
{
    typedef fa_crt_fp_dims<NP> D;
    constexpr int ND = D::ND, NDC = D::NDC, NL = D::NL;
    const double B = 16777216.0, Binv = 1.0 / 16777216.0;
    double col[ND];
#pragma unroll
    for (int m = 0; m < ND; m++) col[m] = 0.0;
    double qe = 0.0;
#pragma unroll
    for (int i = 0; i < NP; i++) {
        if (i == 6) {
            // keep every column below 2^53; the partial sum may be negative,
            // so its carry out of the top digit goes back into that digit
            const double top = fa_crt_fp_carry<ND>(col);
            col[ND - 1] = __builtin_fma(top, B, col[ND - 1]);
        }
        const double n = F.p[i];
        const double x = fa_reduce_pm1n(Q[(uint64_t)i * stride + fa_res_pos(j, stride)], n,
                                        F.pinv[i]);   // [-n/2, n/2]
        qe = __builtin_fma(x, F.rp[i], qe);
        const double v = __builtin_floor(x * Binv);
        const double u = __builtin_fma(-v, B, x);
#pragma unroll
        for (int d = 0; d < NDC; d++) {
            col[d] = __builtin_fma(u, F.cof[i][d], col[d]);
            col[d + 1] = __builtin_fma(v, F.cof[i][d], col[d + 1]);
        }
    }
    const double k = __builtin_rint(qe);
#pragma unroll
    for (int m = 0; m < ND; m++) col[m] = __builtin_fma(-k, F.M[m], col[m]);
    const double top = fa_crt_fp_carry<ND>(col);
    if (top < 0.0) {
        // c' = c - M: add M back
#pragma unroll
        for (int m = 0; m < ND; m++) col[m] += F.M[m];
        fa_crt_fp_carry<ND>(col);
    }
    if (MODE) {
        // c > floor(M/2): the first differing digit from the top decides
        int gt = 0, decided = 0;
#pragma unroll
        for (int m = ND - 1; m >= 0; m--) {
            if (!decided && col[m] != F.Mh[m]) {
                gt = col[m] > F.Mh[m];
                decided = 1;
            }
        }
        if (MODE == 1) {
            if (gt) {
#pragma unroll
                for (int m = 0; m < ND; m++) col[m] = 0.0;
            }
        } else {
#pragma unroll
            for (int m = 0; m < ND; m++) col[m] = gt ? F.M[m] - col[m] : 0.0;
            fa_crt_fp_carry<ND>(col);
        }
    }
#pragma unroll
    for (int l = 0; l < NL; l++) L[l] = 0;
#pragma unroll
    for (int m = 0; m < ND; m++) {
        const uint64_t dv = (uint64_t)(uint32_t)col[m];
        const int pos = 24 * m;
        if (pos / 64 < NL) L[pos / 64] |= dv << (pos % 64);
        if (pos % 64 > 40 && pos / 64 + 1 < NL) L[pos / 64 + 1] |= dv >> (64 - pos % 64);
    }
}
