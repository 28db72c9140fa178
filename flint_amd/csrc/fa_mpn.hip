// fa_mpn.hip -- integer multiplication: limb splitting, transforms,
// pointwise product, inverse transforms, CRT recombination with a carry
// scan.  Restates _mpn_ctx_mpn_mul_range (mpn_mul.c:1617-1978) for one GPU.
//
// Data layout in HBM for a product with np primes and depth D = d1 + d2:
//   A[i][x], B[i][x], i < np, x < 2^D doubles (prime-major rows).
// The transform is done in place in two levels: the first d1 levels act on
// columns (elements with equal x mod 2^d2, "column passes"), the last d2
// levels on contiguous rows of 2^d2 elements ("row pass").  After the row
// pass A holds evaluations in bit-reversed order, i.e. A[i][x] = a(w[x])
// mod p_i, the layout of sd_fft_trunc without its 4x4 block transposition.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fa_internal.h"
#include "fa_ntt.h"

#include "fa_pack.h"
#include "fa_crt_fp.h"

// ------------------------------------------------------------------------
// transform kernels
// ------------------------------------------------------------------------

// depth <= 12: one workgroup per prime does import, both forward
// transforms, the pointwise product and the inverse transform in LDS.
// sqr: a == b, one forward transform and a pointwise square
__global__ void __launch_bounds__(256) fa_k_small_conv(fa_operand a, fa_operand b, int depth,
                                                       double* __restrict__ out, uint64_t stride,
                                                       fa_prime_set S, int sqr)
{
    extern __shared__ double lds[];
    const int i = blockIdx.x;
    const double n = S.p[i], ninv = S.pinv[i];
    const int N = 1 << depth;
    double* sa = lds;
    double* sb = sqr ? lds : lds + N;
    for (int x = threadIdx.x; x < N; x += blockDim.x) {
        sa[x] = fa_operand_value(a, (uint64_t)x, S.pow2[i], n, ninv);
        if (!sqr) sb[x] = fa_operand_value(b, (uint64_t)x, S.pow2[i], n, ninv);
    }
    __syncthreads();
    fa_lds_fwd(sa, depth, 0, 0, S.w2[i], n, ninv);
    if (!sqr) fa_lds_fwd(sb, depth, 0, 0, S.w2[i], n, ninv);
    const double m = S.norm[i];
    for (int x = threadIdx.x; x < N; x += blockDim.x)
        sa[x] = fa_mulmod(fa_mulmod(sa[x], m, n, ninv), sb[x], n, ninv);
    __syncthreads();
    fa_lds_inv(sa, depth, 0, 0, S.w2inv[i], n, ninv);
    double* o = out + (uint64_t)i * stride;
    for (int x = threadIdx.x; x < N; x += blockDim.x) o[x] = sa[x];
}

// first d1 levels on a tile of 2^lgC columns: import + forward.
// blockIdx.x = column group, blockIdx.y = prime, blockIdx.z = operand (a/b)
__global__ void __launch_bounds__(256) fa_k_col_fwd(fa_operand a, fa_operand b, int d1, int d2,
                                                    int lgC, double* __restrict__ A,
                                                    double* __restrict__ B, uint64_t stride,
                                                    fa_prime_set S)
{
    extern __shared__ double lds[];
    const int i = blockIdx.y;
    const double n = S.p[i], ninv = S.pinv[i];
    const fa_operand& op = blockIdx.z ? b : a;
    double* dst = (blockIdx.z ? B : A) + (uint64_t)i * stride;
    const int C = 1 << lgC;
    const uint64_t col0 = (uint64_t)blockIdx.x << lgC;
    const int total = (1 << d1) << lgC;
    for (int u = threadIdx.x; u < total; u += blockDim.x) {
        uint64_t row = (uint64_t)(u >> lgC), c = (uint64_t)(u & (C - 1));
        uint64_t x = (row << d2) + col0 + c;
        lds[u] = fa_operand_value(op, x, S.pow2[i], n, ninv);
    }
    __syncthreads();
    fa_lds_fwd(lds, d1, lgC, 0, S.w2[i], n, ninv);
    for (int u = threadIdx.x; u < total; u += blockDim.x) {
        uint64_t row = (uint64_t)(u >> lgC), c = (uint64_t)(u & (C - 1));
        dst[row * FA_ROW_PITCH + col0 + c] = lds[u];
    }
}

// last d2 levels of both operands on row r, pointwise product with the
// normalizer, then the first d2 inverse levels, in place in A.
// blockIdx.x = row, blockIdx.y = prime
__global__ void __launch_bounds__(256) fa_k_row_conv(double* __restrict__ A,
                                                     const double* __restrict__ B, int d2,
                                                     uint64_t stride, fa_prime_set S)
{
    extern __shared__ double lds[];
    const int i = blockIdx.y;
    const double n = S.p[i], ninv = S.pinv[i];
    const int N2 = 1 << d2;
    const uint64_t r = blockIdx.x;
    double* arow = A + (uint64_t)i * stride + r * FA_ROW_PITCH;
    const double* brow = B + (uint64_t)i * stride + r * FA_ROW_PITCH;
    double* sa = lds;
    double* sb = lds + N2;
    for (int x = threadIdx.x; x < N2; x += blockDim.x) {
        sa[x] = arow[x];
        sb[x] = brow[x];
    }
    __syncthreads();
    fa_lds_fwd(sa, d2, 0, r, S.w2[i], n, ninv);
    fa_lds_fwd(sb, d2, 0, r, S.w2[i], n, ninv);
    const double m = S.norm[i];
    for (int x = threadIdx.x; x < N2; x += blockDim.x)
        sa[x] = fa_mulmod(fa_mulmod(sa[x], m, n, ninv), sb[x], n, ninv);
    __syncthreads();
    fa_lds_inv(sa, d2, 0, r, S.w2inv[i], n, ninv);
    for (int x = threadIdx.x; x < N2; x += blockDim.x) arow[x] = sa[x];
}

// last inverse levels (the first d1 levels of the forward) on columns
__global__ void __launch_bounds__(256) fa_k_col_inv(double* __restrict__ A, int d1, int d2,
                                                    int lgC, uint64_t stride, fa_prime_set S)
{
    extern __shared__ double lds[];
    const int i = blockIdx.y;
    const double n = S.p[i], ninv = S.pinv[i];
    double* dst = A + (uint64_t)i * stride;
    const int C = 1 << lgC;
    const uint64_t col0 = (uint64_t)blockIdx.x << lgC;
    const int total = (1 << d1) << lgC;
    for (int u = threadIdx.x; u < total; u += blockDim.x) {
        uint64_t row = (uint64_t)(u >> lgC), c = (uint64_t)(u & (C - 1));
        lds[u] = dst[row * FA_ROW_PITCH + col0 + c];
    }
    __syncthreads();
    fa_lds_inv(lds, d1, lgC, 0, S.w2inv[i], n, ninv);
    for (int u = threadIdx.x; u < total; u += blockDim.x) {
        uint64_t row = (uint64_t)(u >> lgC), c = (uint64_t)(u & (C - 1));
        dst[row * FA_ROW_PITCH + col0 + c] = lds[u];
    }
}

// ------------------------------------------------------------------------
// CRT recombination and carry propagation
// ------------------------------------------------------------------------

#define FA_WIN 9        // sliding window limbs: S (<= 7 limbs) shifted + pending
#define FA_CRT_WG 256

// x_i = c * cop_i^-1 mod p_i for the true coefficient c < M (the
// normalizer folded in the pointwise product), so
// c = sum_i x_i (M/p_i) - q M with q = floor(sum_i x_i / p_i)
// (crt_helpers.h _big_mul/_big_addmul/_reduce_big_sum, mpn_mul.c:883-906).
FA_DEV void fa_crt_coeff(const double* __restrict__ Q, uint64_t stride, uint64_t j,
                         const fa_crt_consts& C, uint64_t s[FA_CRT_MAXL])
{
#pragma unroll
    for (int l = 0; l < FA_CRT_MAXL; l++) s[l] = 0;
    double qe = 0.0;
    for (int i = 0; i < C.np; i++) {
        uint64_t x = fa_reduce_0n_u64(Q[(uint64_t)i * stride + fa_res_pos(j, stride)], C.p[i],
                                      C.pinv[i]);
        qe += (double)x * C.rp[i];
        uint64_t carry = 0;
#pragma unroll
        for (int l = 0; l < FA_CRT_MAXL; l++) {
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
    // subtract q*M; q may be off by one near integers: fix up with M
    uint64_t q = (uint64_t)qe;
    uint64_t borrow = 0, carry = 0;
#pragma unroll
    for (int l = 0; l < FA_CRT_MAXL; l++) {
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
        for (int l = 0; l < FA_CRT_MAXL; l++) {
            uint64_t t = s[l] + C.M[l];
            uint64_t cc = (t < s[l]);
            t += c2;
            cc |= (t < c2);
            s[l] = t;
            c2 = cc;
        }
    } else {
        // s >= M ?
        int ge = 1;
#pragma unroll
        for (int l = FA_CRT_MAXL - 1; l >= 0; l--) {
            if (ge == 1 && s[l] != C.M[l]) ge = (s[l] > C.M[l]) ? 2 : 0;
        }
        if (ge) {
            uint64_t br = 0;
#pragma unroll
            for (int l = 0; l < FA_CRT_MAXL; l++) {
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

// One thread per group of 8 consecutive coefficients (8*bits is a multiple
// of 64, so a group owns bits/8 whole limbs).  Thread 0 of each workgroup
// recomputes the group just before the workgroup's range so that its
// overhang is available without a second pass.
__global__ void __launch_bounds__(FA_CRT_WG) fa_k_crt_mpn(const double* __restrict__ Q,
                                                         uint64_t stride, fa_crt_consts C,
                                                         uint64_t zlen, int bits,
                                                         uint64_t* __restrict__ z, uint64_t zn,
                                                         uint64_t ngroups,
                                                         uint8_t* __restrict__ flags,
                                                         uint8_t* __restrict__ agg)
{
    __shared__ uint64_t sh_ovh[FA_CRT_WG][FA_WIN];
    __shared__ uint8_t sh_f[FA_CRT_WG];
    const int tid = threadIdx.x;
    const int64_t g = (int64_t)blockIdx.x * (FA_CRT_WG - 1) + tid - 1;
    const int own = bits >> 3;
    const bool active = g >= 0 && (uint64_t)g < ngroups;

    uint64_t win[FA_WIN];
#pragma unroll
    for (int l = 0; l < FA_WIN; l++) win[l] = 0;

    if (active) {
        uint64_t zbase = (uint64_t)g * own;
        uint64_t emitted = 0;
        uint64_t prev_limb = 0;
        for (int k = 0; k < 8; k++) {
            uint64_t j = (uint64_t)g * 8 + k;
            uint64_t bitoff = (uint64_t)k * bits;
            int sh = (int)(bitoff & 63);
            // limbs below floor(bitoff/64) are final: emit them
            uint64_t cur_limb = bitoff >> 6;
            while (prev_limb < cur_limb) {
                if (tid > 0 && zbase + emitted < zn) z[zbase + emitted] = win[0];
                emitted++;
#pragma unroll
                for (int l = 0; l < FA_WIN - 1; l++) win[l] = win[l + 1];
                win[FA_WIN - 1] = 0;
                prev_limb++;
            }
            if (j < zlen) {
                uint64_t s[FA_CRT_MAXL];
                fa_crt_coeff(Q, stride, j, C, s);
                uint64_t carry = 0;
#pragma unroll
                for (int l = 0; l < FA_WIN; l++) {
                    uint64_t cur = (l < FA_CRT_MAXL) ? s[l] : 0;
                    uint64_t prv = (l >= 1 && l - 1 < FA_CRT_MAXL) ? s[l - 1] : 0;
                    uint64_t v = sh ? ((cur << sh) | (prv >> (64 - sh))) : cur;
                    uint64_t t = win[l] + v;
                    uint64_t c1 = (t < v);
                    t += carry;
                    c1 |= (t < carry);
                    win[l] = t;
                    carry = c1;
                }
            }
        }
        uint64_t end_limb = (uint64_t)own;   // 8*bits/64
        while (prev_limb < end_limb) {
            if (tid > 0 && zbase + emitted < zn) z[zbase + emitted] = win[0];
            emitted++;
#pragma unroll
            for (int l = 0; l < FA_WIN - 1; l++) win[l] = win[l + 1];
            win[FA_WIN - 1] = 0;
            prev_limb++;
        }
    }
#pragma unroll
    for (int l = 0; l < FA_WIN; l++) sh_ovh[tid][l] = win[l];
    __syncthreads();

    uint8_t f = 0;   // bit0: carry out (generate), bit1: all limbs ones (propagate)
    if (active && tid > 0) {
        uint64_t zbase = (uint64_t)g * own;
        uint64_t carry = 0;
        int prop = 1;
        for (int l = 0; l < own; l++) {
            if (zbase + l >= zn) { prop = 0; break; }
            uint64_t add = (l < FA_WIN) ? sh_ovh[tid - 1][l] : 0;
            uint64_t v = z[zbase + l];
            uint64_t cin = carry;
            uint64_t t = v + add;
            uint64_t c1 = (t < add);
            t += cin;
            c1 |= (t < cin);
            carry = c1;
            if (add | cin) z[zbase + l] = t;
            if (t != ~UINT64_C(0)) prop = 0;
            if (l >= FA_WIN - 1 && carry == 0 && !prop) break;
        }
        f = (uint8_t)(carry | (prop << 1));
        flags[g] = f;
    }
    sh_f[tid] = (tid > 0 && active) ? f : 2;   // identity: propagate, no generate
    __syncthreads();
    if (tid == 0) {
        uint8_t gg = 0, pp = 1;
        for (int t = 1; t < FA_CRT_WG; t++) {
            uint8_t ft = sh_f[t];
            gg = (uint8_t)((ft & 1) | (((ft >> 1) & 1) & gg));
            pp = (uint8_t)(pp & ((ft >> 1) & 1));
        }
        agg[blockIdx.x] = (uint8_t)(gg | (pp << 1));
    }
}

// exclusive scan of workgroup (generate, propagate) pairs -> carry-in.
// One workgroup of 1024 threads: each thread folds a contiguous chunk of
// aggregates, the 1024 chunk aggregates are scanned in LDS, then each thread
// walks its chunk again writing the carry-ins.
__global__ void __launch_bounds__(1024) fa_k_carry_scan(const uint8_t* __restrict__ agg,
                                                        uint8_t* __restrict__ cin, uint64_t nblk)
{
    __shared__ uint8_t sg[1024], sp[1024];
    const int t = threadIdx.x;
    const uint64_t chunk = (nblk + 1023) / 1024;
    const uint64_t b0 = (uint64_t)t * chunk;
    const uint64_t b1 = (b0 + chunk < nblk) ? b0 + chunk : nblk;
    uint8_t g = 0, p = 1;
    for (uint64_t b = b0; b < b1; b++) {
        uint8_t a = agg[b];
        g = (uint8_t)((a & 1) | (((a >> 1) & 1) & g));
        p = (uint8_t)(p & ((a >> 1) & 1));
    }
    sg[t] = g;
    sp[t] = p;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint8_t g1 = 0, p1 = 1;
        if (t >= off) { g1 = sg[t - off]; p1 = sp[t - off]; }
        __syncthreads();
        uint8_t g2 = sg[t], p2 = sp[t];
        sg[t] = g2 | (p2 & g1);
        sp[t] = p1 & p2;
        __syncthreads();
    }
    uint8_t c = t ? sg[t - 1] : 0;
    for (uint64_t b = b0; b < b1; b++) {
        cin[b] = c;
        uint8_t a = agg[b];
        c = (uint8_t)((a & 1) | (((a >> 1) & 1) & c));
    }
}

// add the incoming carries: group g gets a carry iff the exclusive scan of
// the flags of the groups before it (within the workgroup) and the
// workgroup carry-in produce one
__global__ void __launch_bounds__(FA_CRT_WG) fa_k_carry_apply(const uint8_t* __restrict__ flags,
                                                             const uint8_t* __restrict__ cin,
                                                             uint64_t* __restrict__ z, uint64_t zn,
                                                             int own, uint64_t ngroups)
{
    __shared__ uint8_t sg[FA_CRT_WG], sp[FA_CRT_WG];
    const int tid = threadIdx.x;
    const int64_t g = (int64_t)blockIdx.x * (FA_CRT_WG - 1) + tid - 1;
    const bool active = tid > 0 && g >= 0 && (uint64_t)g < ngroups;
    uint8_t f = active ? flags[g] : 2;
    if (tid == 0) f = (uint8_t)(cin[blockIdx.x] & 1);   // seed with the block carry-in
    sg[tid] = f & 1;
    sp[tid] = (f >> 1) & 1;
    __syncthreads();
    for (int off = 1; off < FA_CRT_WG; off <<= 1) {
        uint8_t g1 = 0, p1 = 1;
        if (tid >= off) { g1 = sg[tid - off]; p1 = sp[tid - off]; }
        __syncthreads();
        uint8_t g2 = sg[tid], p2 = sp[tid];
        sg[tid] = g2 | (p2 & g1);
        sp[tid] = p1 & p2;
        __syncthreads();
    }
    if (!active) return;
    uint8_t carry_in = sg[tid - 1];
    if (!carry_in) return;
    uint64_t zbase = (uint64_t)g * own;
    for (int l = 0; l < own && zbase + l < zn; l++) {
        uint64_t v = z[zbase + l] + 1;
        z[zbase + l] = v;
        if (v != 0) break;
    }
}

// ------------------------------------------------------------------------
// coalesced CRT: one thread per coefficient computes its integer value into
// LDS, then one thread per group of 8 coefficients stitches the shifted
// values into the group's bits/8 limbs, the overhang of each group is added
// into the next one, and the workgroup writes its limbs contiguously.
// A workgroup covers 32 groups; its first group (the last one of the
// previous workgroup) is recomputed only for its overhang.
// ------------------------------------------------------------------------

template <int NP>
struct fa_crt_limbs {
    // limbs of sum_i x_i (M/p_i) < NP * M < 2^(50 NP + 3)
    static constexpr int N = (50 * NP + 3 + 63) / 64;
};

template <int NP>
FA_DEV void fa_crt_value(const double* __restrict__ Q, uint64_t stride, uint64_t j,
                         const fa_crt_consts& C, uint64_t* s)
{
    constexpr int NL = fa_crt_limbs<NP>::N;
#pragma unroll
    for (int l = 0; l < NL; l++) s[l] = 0;
    double qe = 0.0;
#pragma unroll
    for (int i = 0; i < NP; i++) {
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

#define FA_CRT16_GROUPS 32   // groups per workgroup, the first one a halo
#define FA_OUT_STRIDE 25     // LDS row stride (doubles) of a group's limbs

template <int NP>
__global__ void __launch_bounds__(256) fa_k_crt_mpn16(const double* __restrict__ Q,
                                                      uint64_t stride, fa_crt_consts C,
                                                      uint64_t zlen, int bits,
                                                      uint64_t* __restrict__ z, uint64_t zn,
                                                      uint64_t ngroups,
                                                      uint8_t* __restrict__ flags,
                                                      uint8_t* __restrict__ agg)
{
    constexpr int NL = fa_crt_limbs<NP>::N;
    __shared__ uint64_t sS[NL][264];
    __shared__ uint64_t sOut[FA_CRT16_GROUPS * FA_OUT_STRIDE];
    __shared__ uint64_t sOvh[FA_CRT16_GROUPS][FA_WIN];
    __shared__ uint8_t sF[FA_CRT16_GROUPS];
    const int t = threadIdx.x;
    const int own = bits >> 3;
    const int64_t g0 = (int64_t)blockIdx.x * (FA_CRT16_GROUPS - 1) - 1;

    {   // coefficient c = t of the workgroup, stored at (c mod 8)*33 + c/8
        int64_t j = g0 * 8 + t;
        uint64_t s[NL];
#pragma unroll
        for (int l = 0; l < NL; l++) s[l] = 0;
        if (j >= 0 && (uint64_t)j < zlen) fa_crt_value<NP>(Q, stride, (uint64_t)j, C, s);
        const int pos = (t & 7) * 33 + (t >> 3);
#pragma unroll
        for (int l = 0; l < NL; l++) sS[l][pos] = s[l];
    }
    __syncthreads();

    if (t < FA_CRT16_GROUPS) {
        uint64_t win[FA_WIN];
#pragma unroll
        for (int l = 0; l < FA_WIN; l++) win[l] = 0;
        int emitted = 0;
        uint64_t prev_limb = 0;
        uint64_t* out = sOut + t * FA_OUT_STRIDE;
        for (int k = 0; k < 8; k++) {
            uint64_t bitoff = (uint64_t)k * bits;
            int sh = (int)(bitoff & 63);
            uint64_t cur = bitoff >> 6;
            while (prev_limb < cur) {
                out[emitted++] = win[0];
#pragma unroll
                for (int l = 0; l < FA_WIN - 1; l++) win[l] = win[l + 1];
                win[FA_WIN - 1] = 0;
                prev_limb++;
            }
            const int pos = k * 33 + t;
            uint64_t carry = 0;
#pragma unroll
            for (int l = 0; l < FA_WIN; l++) {
                uint64_t cu = (l < NL) ? sS[l < NL ? l : 0][pos] : 0;
                uint64_t pv = (l >= 1 && l - 1 < NL) ? sS[(l >= 1 && l - 1 < NL) ? l - 1 : 0][pos] : 0;
                uint64_t v = sh ? ((cu << sh) | (pv >> (64 - sh))) : cu;
                uint64_t tt = win[l] + v;
                uint64_t c1 = (tt < v);
                tt += carry;
                c1 |= (tt < carry);
                win[l] = tt;
                carry = c1;
            }
        }
        while (prev_limb < (uint64_t)own) {
            out[emitted++] = win[0];
#pragma unroll
            for (int l = 0; l < FA_WIN - 1; l++) win[l] = win[l + 1];
            win[FA_WIN - 1] = 0;
            prev_limb++;
        }
#pragma unroll
        for (int l = 0; l < FA_WIN; l++) sOvh[t][l] = win[l];
    }
    __syncthreads();

    if (t >= 1 && t < FA_CRT16_GROUPS) {
        const int64_t g = g0 + t;
        uint8_t f = 2;   // identity of the carry operator: propagate, no generate
        if (g >= 0 && (uint64_t)g < ngroups) {
            uint64_t* out = sOut + t * FA_OUT_STRIDE;
            uint64_t carry = 0;
            int prop = 1;
            for (int l = 0; l < own; l++) {
                if ((uint64_t)g * own + l >= zn) { prop = 0; break; }
                uint64_t add = (l < FA_WIN) ? sOvh[t - 1][l < FA_WIN ? l : 0] : 0;
                uint64_t v = out[l];
                uint64_t tt = v + add;
                uint64_t c1 = (tt < add);
                tt += carry;
                c1 |= (tt < carry);
                carry = c1;
                out[l] = tt;
                if (tt != ~UINT64_C(0)) prop = 0;
            }
            f = (uint8_t)(carry | (prop << 1));
            flags[g] = f;
        }
        sF[t] = f;
    }
    __syncthreads();
    if (t == 0) {
        // bit 0: carry out with no carry in, bit 1: every group propagates,
        // bit 2: some group generates a carry
        uint8_t gg = 0, pp = 1, any = 0;
        for (int k = 1; k < FA_CRT16_GROUPS; k++) {
            uint8_t ft = sF[k];
            gg = (uint8_t)((ft & 1) | (((ft >> 1) & 1) & gg));
            pp = (uint8_t)(pp & ((ft >> 1) & 1));
            any = (uint8_t)(any | (ft & 1));
        }
        agg[blockIdx.x] = (uint8_t)(gg | (pp << 1) | (any << 2));
    }
    // contiguous write of the limbs of groups 1..31
    const uint64_t base = (uint64_t)(g0 + 1) * own;
    const int nl = (FA_CRT16_GROUPS - 1) * own;
    for (int e = t; e < nl; e += 256) {
        uint64_t za = base + e;
        if (za < zn) z[za] = sOut[(1 + e / own) * FA_OUT_STRIDE + e % own];
    }
}

// carry-in of every group of a 31-group workgroup: scan of the group flags
// seeded with the workgroup carry-in
__global__ void __launch_bounds__(64) fa_k_carry_apply16(const uint8_t* __restrict__ flags,
                                                         const uint8_t* __restrict__ cin,
                                                         uint64_t* __restrict__ z, uint64_t zn,
                                                         int own, uint64_t ngroups)
{
    __shared__ uint8_t sg[FA_CRT16_GROUPS], sp[FA_CRT16_GROUPS];
    const int t = threadIdx.x;
    const int64_t g = (int64_t)blockIdx.x * (FA_CRT16_GROUPS - 1) + t - 1;
    const bool active = t >= 1 && t < FA_CRT16_GROUPS && g >= 0 && (uint64_t)g < ngroups;
    if (t < FA_CRT16_GROUPS) {
        uint8_t f = active ? flags[g] : 2;
        if (t == 0) f = (uint8_t)(cin[blockIdx.x] & 1);
        sg[t] = f & 1;
        sp[t] = (f >> 1) & 1;
    }
    __syncthreads();
    if (t != 0) return;
    // serial inclusive scan over 32 entries, then apply
    uint8_t gg = sg[0];
    for (int k = 1; k < FA_CRT16_GROUPS; k++) {
        int64_t gk = (int64_t)blockIdx.x * (FA_CRT16_GROUPS - 1) + k - 1;
        if (gk < 0 || (uint64_t)gk >= ngroups) break;
        if (gg) {
            uint64_t zb = (uint64_t)gk * own;
            for (int l = 0; l < own && zb + l < zn; l++) {
                uint64_t v = z[zb + l] + 1;
                z[zb + l] = v;
                if (v != 0) break;
            }
        }
        gg = sg[k] | (sp[k] & gg);
    }
}

// adds 1 to the n limbs at p with carry (stops at the first limb that does
// not wrap); eight loads in flight at a time
FA_DEV void fa_add_one(uint64_t* __restrict__ p, int n)
{
    for (int l0 = 0; l0 < n; l0 += 8) {
        const int m = n - l0 < 8 ? n - l0 : 8;
        uint64_t v[8];
#pragma unroll
        for (int l = 0; l < 8; l++)
            if (l < m) v[l] = p[l0 + l];
#pragma unroll
        for (int l = 0; l < 8; l++) {
            if (l < m) {
                v[l] += 1;
                p[l0 + l] = v[l];
                if (v[l] != 0) return;
            }
        }
    }
}

FA_DEV uint8_t fa_gp_fold(uint8_t acc, uint8_t a)   // (generate, propagate) of acc then a
{
    const uint8_t g = (uint8_t)((a & 1) | (((a >> 1) & 1) & acc));
    const uint8_t p = (uint8_t)((acc >> 1) & (a >> 1) & 1);
    return (uint8_t)(g | (p << 1));
}

// carries between CRT workgroups ("blocks" of FA_CRT16_GROUPS - 1 groups of
// own limbs), by workgroup wg of nwg (1024 threads each) on its contiguous
// range of blocks.  A block's aggregate holds bit 0: carry out with no
// carry in, bit 1: every group propagates, bit 2: some group generates.
//  1. the carry into the range: each thread folds a slice of the aggregates
//     before it, the slices are combined in order in LDS;
//  2. the carry into each block of the range: a scan over the range in
//     tiles of 1024 blocks;
//  3. one wave per block: lane k takes group k of the block, its carry-in
//     composed from the flags of the groups before it, and a group with a
//     carry-in adds it to its limbs.  A group's addition stays inside its
//     own limbs except when they are all ones (then they become zeros and
//     the carry passes on, which the flags already give the next group), so
//     groups are independent once their carries-in are known.  Blocks that
//     receive no carry and generate none inside are skipped.
// No workgroup waits on another: each computes its own prefix.
FA_DEV void fa_carry_resolve(const uint8_t* __restrict__ flags, const uint8_t* __restrict__ agg,
                             uint64_t* __restrict__ z, uint64_t zn, int own, uint64_t ngroups,
                             uint64_t nblk, unsigned wg, unsigned nwg)
{
    __shared__ uint8_t sa[1024];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint64_t per = (nblk + nwg - 1) / nwg;
    const uint64_t r0 = (uint64_t)wg * per < nblk ? (uint64_t)wg * per : nblk;
    const uint64_t r1 = r0 + per < nblk ? r0 + per : nblk;
    // 1. aggregate of blocks [0, r0), identity (no generate, propagate) = 2
    {
        const uint64_t sl = (r0 + 1023) / 1024;
        const uint64_t s0 = (uint64_t)t * sl < r0 ? (uint64_t)t * sl : r0;
        const uint64_t s1 = s0 + sl < r0 ? s0 + sl : r0;
        uint8_t acc = 2;
        for (uint64_t b = s0; b < s1; b++) acc = fa_gp_fold(acc, agg[b]);
        sa[t] = acc;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            uint8_t v = 0;
            const bool act = (t & (2 * off - 1)) == 0;
            if (act) v = fa_gp_fold(sa[t], sa[t + off]);
            __syncthreads();
            if (act) sa[t] = v;
            __syncthreads();
        }
    }
    uint8_t c = sa[0] & 1;          // carry into block r0
    __syncthreads();
    const uint64_t gper = FA_CRT16_GROUPS - 1;
    for (uint64_t tb = r0; tb < r1; tb += 1024) {
        // 2. inclusive scan of this tile's aggregates, seeded with c
        const uint64_t b = tb + (uint64_t)t;
        const uint8_t a = b < r1 ? agg[b] : 2;
        sa[t] = a;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            uint8_t v = sa[t];
            if (t >= off) v = fa_gp_fold(sa[t - off], v);
            __syncthreads();
            sa[t] = v;
            __syncthreads();
        }
        // carry into block b: fold of c with the tile's blocks before b
        const uint8_t cin = t ? (uint8_t)(fa_gp_fold(c, sa[t - 1]) & 1) : c;
        const uint8_t cout_tile = (uint8_t)(fa_gp_fold(c, sa[1023]) & 1);
        __syncthreads();
        sa[t] = (uint8_t)(cin | (a & 4));
        __syncthreads();
        // 3. one wave per block of the tile
        const uint64_t nb = r1 - tb < 1024 ? r1 - tb : 1024;
        for (uint64_t k = (uint64_t)wave; k < nb; k += 16) {
            const uint8_t s = sa[k];
            if (!(s & 1) && !(s & 4)) continue;
            const uint64_t gs = (tb + k) * gper;
            const uint64_t ge = gs + gper < ngroups ? gs + gper : ngroups;
            const uint64_t gi = gs + (uint64_t)lane;
            const uint8_t f = gi < ge ? flags[gi] : 2;
            uint8_t cg = s & 1;
            // every lane takes part in every shuffle (the block test above
            // is the same for the whole wave)
            for (int j = 0; j < (int)gper; j++) {
                const uint8_t fj = (uint8_t)__shfl((int)f, j);
                if (j < lane) cg = (uint8_t)((fj & 1) | (((fj >> 1) & 1) & cg));
            }
            if (cg && gi < ge) {
                const uint64_t zb = gi * (uint64_t)own;
                const int n = zb + (uint64_t)own <= zn ? own : (int)(zn - zb);
                fa_add_one(z + zb, n);
            }
        }
        c = cout_tile;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(1024) fa_k_carry_fused(const uint8_t* __restrict__ flags,
                                                         const uint8_t* __restrict__ agg,
                                                         uint64_t* __restrict__ z, uint64_t zn,
                                                         int own, uint64_t ngroups, uint64_t nblk)
{
    fa_carry_resolve(flags, agg, z, zn, own, ngroups, nblk, blockIdx.x, gridDim.x);
}

// workgroups of the carry kernel for nblk CRT workgroups: one per 32 blocks,
// at most 256
static unsigned carry_grid(uint64_t nblk)
{
    const uint64_t g = nblk / 32;
    return (unsigned)(g < 1 ? 1 : g > 256 ? 256 : g);
}

// ------------------------------------------------------------------------
// CRT on the FP64 lanes (fa_crt_fp.h), 31 groups of 8 coefficients per
// workgroup of 256 threads plus one halo group in front (its overhang
// reaches the first group).
//   1. thread t: coefficient 8 g0 + t -> NL limbs, shifted left by
//      (bits * (t mod 8)) mod 64 into NLS = NL + 1 limbs in LDS;
//   2. lanes (group gl, limb q), 32 lanes per group: the sum of the <= 8
//      shifted coefficients overlapping limb q of the group, then the
//      group's carries: limb q gets the carry count of limb q-1, and the
//      single-bit overflows ripple through all-ones limbs, resolved at once
//      from the wave's ballot masks (generate G, propagate P: the carries
//      in are (G + (G|P)) ^ (G|P) ^ G);
//   3. lanes (group gl >= 1, limb q < own): owned limb q plus limb own + q
//      of the previous group (its overhang, which fits in own limbs), the
//      same mask carry resolution, the store, and the group's flag
//      (generate: carry out of the owned limbs, propagate: all ones).
// ------------------------------------------------------------------------
#define FA_CRTFP_TL 32   // max limbs of a group's sum: 7*200/64 + 8 + 1

FA_DEV uint32_t fa_seg_ballot(bool pred)
{
    const unsigned long long m = __ballot(pred);
    return (uint32_t)(m >> (threadIdx.x & 32));
}

template <int NP, int MODE = 0>
FA_DEV void fa_crt_fp_block(const double* __restrict__ Q, uint64_t stride,
                            const fa_crt_fp_consts& F, uint64_t zlen, int bits,
                            uint64_t* __restrict__ z, uint64_t zn, uint64_t ngroups,
                            uint8_t* __restrict__ flags, uint8_t* __restrict__ agg, uint64_t blk,
                            uint64_t jlo = 0)
{
    constexpr int NL = fa_crt_fp_dims<NP>::NL;
    constexpr int NLS = NL + 1;
    __shared__ uint64_t sC[256 * NLS];
    __shared__ uint64_t sR[FA_CRT16_GROUPS][FA_CRTFP_TL];
    __shared__ uint8_t sF[FA_CRT16_GROUPS];
    const int t = threadIdx.x;
    const int own = bits >> 3;
    const int TL = ((7 * bits) >> 6) + NLS;
    const int ovh = TL - own;
    const int64_t g0 = (int64_t)blk * (FA_CRT16_GROUPS - 1) - 1;

    {   // 1. the coefficient of thread t, shifted to its bit offset in the group
        //    (coefficients below jlo read as zero)
        const int64_t j = g0 * 8 + t;
        uint64_t L[NL];
        if (j >= (int64_t)jlo && (uint64_t)j < zlen) {
            fa_crt_fp_value<NP, MODE>(Q, stride, (uint64_t)j, F, L);
        } else {
#pragma unroll
            for (int l = 0; l < NL; l++) L[l] = 0;
        }
        const int sh = (bits * (t & 7)) & 63;
        uint64_t* o = sC + t * NLS;
        uint64_t prev = 0;
#pragma unroll
        for (int l = 0; l < NL; l++) {
            o[l] = (L[l] << sh) | (sh ? (prev >> (64 - sh)) : 0);
            prev = L[l];
        }
        o[NL] = sh ? (prev >> (64 - sh)) : 0;
    }

    // the coefficients k overlapping limb q: o_k <= q < o_k + NLS with
    // o_k = (bits k) >> 6, k in [kmin, kmax] (at most 8 of them)
    const int q = t & 31, seg = t >> 5;
    const bool qin = q < TL;
    int kmax = (64 * q + 63) / bits;
    if (kmax > 7) kmax = 7;
    int kmin = 0;
    if (q >= NLS) kmin = (64 * (q - NLS + 1) + bits - 1) / bits;
    int offs[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int k = kmin + e;
        offs[e] = k * NLS + q - ((bits * k) >> 6);
    }
    const int nk = qin ? kmax - kmin + 1 : 0;
    // no limb overlaps more than (64 NLS - 1) / bits + 1 coefficients: a
    // bound the whole workgroup shares, so the loop below stops there
    const int emax = (64 * NLS - 1) / bits + 1;
    __syncthreads();

    // 2. group sums and the carries inside each group
#pragma unroll
    for (int it = 0; it < 4; it++) {
        const int gl = seg + 8 * it;
        const uint64_t* base = sC + gl * 8 * NLS;
        uint64_t lo = 0;
        uint32_t hi = 0;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            if (e >= emax) break;
            if (e < nk) {
                const uint64_t v = base[offs[e]];
                lo += v;
                hi += (lo < v);
            }
        }
        uint32_t hprev = __shfl_up(hi, 1, 32);
        if (q == 0) hprev = 0;
        const uint64_t v = lo + hprev;
        const uint32_t G = fa_seg_ballot(qin && v < lo);
        const uint32_t P = fa_seg_ballot(qin && v == ~UINT64_C(0));
        const uint64_t X = (uint64_t)(G | P);
        const uint64_t cin = ((uint64_t)G + X) ^ X ^ (uint64_t)G;
        if (qin) sR[gl][q] = v + ((cin >> q) & 1);
    }
    __syncthreads();

    // 3. the previous group's overhang into the owned limbs, store, flags
#pragma unroll
    for (int it = 0; it < 4; it++) {
        const int gl = seg + 8 * it;
        const int64_t g = g0 + gl;
        const bool valid = gl >= 1 && g >= 0 && (uint64_t)g < ngroups;
        const uint64_t zi = (uint64_t)g * own + q;
        const bool lin = valid && q < own && zi < zn;
        uint64_t S = 0, add = 0;
        if (lin) {
            add = (q < ovh) ? sR[gl - 1][own + q] : 0;
            S = sR[gl][q] + add;
        }
        const uint32_t G = fa_seg_ballot(lin && S < add);
        const uint32_t P = fa_seg_ballot(lin && S == ~UINT64_C(0));
        const uint64_t X = (uint64_t)(G | P);
        const uint64_t cin = ((uint64_t)G + X) ^ X ^ (uint64_t)G;
        if (lin) z[zi] = S + ((cin >> q) & 1);
        if (q == 0 && gl >= 1) {
            uint8_t f = 2;   // identity of the carry operator: propagate, no generate
            if (valid) {
                const uint32_t full = (own >= 32) ? 0xffffffffu : ((1u << own) - 1);
                f = (uint8_t)(((cin >> own) & 1) | ((P == full) ? 2 : 0));
                flags[g] = f;
            }
            sF[gl] = f;
        }
    }
    __syncthreads();
    if (t == 0) {
        // bit 0: carry out with no carry in, bit 1: every group propagates,
        // bit 2: some group generates a carry
        uint8_t gg = 0, pp = 1, any = 0;
        for (int k = 1; k < FA_CRT16_GROUPS; k++) {
            const uint8_t ft = sF[k];
            gg = (uint8_t)((ft & 1) | (((ft >> 1) & 1) & gg));
            pp = (uint8_t)(pp & ((ft >> 1) & 1));
            any = (uint8_t)(any | (ft & 1));
        }
        agg[blk] = (uint8_t)(gg | (pp << 1) | (any << 2));
    }
}

template <int NP, int MODE = 0>
__global__ void __launch_bounds__(256) fa_k_crt_fp(const double* __restrict__ Q, uint64_t stride,
                                                   fa_crt_fp_consts F, uint64_t zlen, int bits,
                                                   uint64_t* __restrict__ z, uint64_t zn,
                                                   uint64_t ngroups, uint8_t* __restrict__ flags,
                                                   uint8_t* __restrict__ agg, uint64_t jlo)
{
    fa_crt_fp_block<NP, MODE>(Q, stride, F, zlen, bits, z, zn, ngroups, flags, agg, blockIdx.x,
                              jlo);
}

// batched CRT of integer products: work item = job | block << 32
FA_DEV unsigned fa_xcd_lin()
{
    const unsigned nb = gridDim.x, bid = blockIdx.x;
    return (bid & 7) * (nb >> 3) + (bid >> 3);
}

template <int NP>
__global__ void __launch_bounds__(256) fa_kb_crt_fp(const fa_bmpn* __restrict__ M,
                                                    const uint64_t* __restrict__ work,
                                                    unsigned nwork,
                                                    const fa_crt_fp_consts* __restrict__ F)
{
    const unsigned L = fa_xcd_lin();
    if (L >= nwork) return;
    const uint64_t w = work[L];
    const fa_bmpn& m = M[(unsigned)(w & 0xffffffffu)];
    fa_crt_fp_block<NP>(m.Q, m.stride, *F, m.zlen, m.bits, m.z, m.zn, m.ngroups, m.flags, m.agg,
                        w >> 32);
}

__global__ void __launch_bounds__(1024) fa_kb_carry(const fa_bmpn* __restrict__ M,
                                                    const unsigned* __restrict__ jobs,
                                                    unsigned njobs)
{
    if (blockIdx.x >= njobs) return;
    const fa_bmpn& m = M[jobs[blockIdx.x]];
    fa_carry_resolve(m.flags, m.agg, m.z, m.zn, m.bits / 8, m.ngroups, m.nblk, 0, 1);
}

int fa_batch_crt_mpn(int np, const fa_bmpn* M, const uint64_t* work, unsigned nwork,
                     const void* Fv, hipStream_t s)
{
    if (!nwork) return FA_OK;
    const fa_crt_fp_consts* F = (const fa_crt_fp_consts*)Fv;
    const unsigned grid = (nwork + 7) & ~7u;
    switch (np) {
    case 2: FA_LAUNCH("fa_kb_crt_fp", fa_kb_crt_fp<2>, dim3(grid), dim3(256), 0, s, M, work, nwork, F); break;
    case 3: FA_LAUNCH("fa_kb_crt_fp", fa_kb_crt_fp<3>, dim3(grid), dim3(256), 0, s, M, work, nwork, F); break;
    case 4: FA_LAUNCH("fa_kb_crt_fp", fa_kb_crt_fp<4>, dim3(grid), dim3(256), 0, s, M, work, nwork, F); break;
    case 5: FA_LAUNCH("fa_kb_crt_fp", fa_kb_crt_fp<5>, dim3(grid), dim3(256), 0, s, M, work, nwork, F); break;
    case 6: FA_LAUNCH("fa_kb_crt_fp", fa_kb_crt_fp<6>, dim3(grid), dim3(256), 0, s, M, work, nwork, F); break;
    case 7: FA_LAUNCH("fa_kb_crt_fp", fa_kb_crt_fp<7>, dim3(grid), dim3(256), 0, s, M, work, nwork, F); break;
    case 8: FA_LAUNCH("fa_kb_crt_fp", fa_kb_crt_fp<8>, dim3(grid), dim3(256), 0, s, M, work, nwork, F); break;
    default: return fa_set_error(FA_E_ARG, "batch crt: np %d", np);
    }
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "batch crt launch");
}

int fa_batch_carry(const fa_bmpn* M, const unsigned* jobs, unsigned njobs, hipStream_t s)
{
    if (!njobs) return FA_OK;
    FA_LAUNCH("fa_kb_carry", fa_kb_carry, dim3(njobs), dim3(1024), 0, s, M, jobs, njobs);
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "batch carry launch");
}

int fa_crt_fp_enabled = 1;
extern "C" FA_API void fa_set_crt_variant(int v) { fa_crt_fp_enabled = v; }

template <int NP>
static void launch_crt_fp(const double* Q, uint64_t stride, const fa_crt_fp_consts& F,
                          uint64_t zlen, int bits, uint64_t* z, uint64_t zn, uint64_t ngroups,
                          uint8_t* flags, uint8_t* agg, uint64_t nblk, hipStream_t s,
                          int mode = 0, uint64_t jlo = 0)
{
    if (mode == 1)
        FA_LAUNCH("fa_k_crt_fp", (fa_k_crt_fp<NP, 1>), dim3((unsigned)nblk), dim3(256), 0, s, Q,
                  stride, F, zlen, bits, z, zn, ngroups, flags, agg, jlo);
    else if (mode == 2)
        FA_LAUNCH("fa_k_crt_fp", (fa_k_crt_fp<NP, 2>), dim3((unsigned)nblk), dim3(256), 0, s, Q,
                  stride, F, zlen, bits, z, zn, ngroups, flags, agg, jlo);
    else
        FA_LAUNCH("fa_k_crt_fp", fa_k_crt_fp<NP>, dim3((unsigned)nblk), dim3(256), 0, s, Q,
                  stride, F, zlen, bits, z, zn, ngroups, flags, agg, jlo);
}

template <int NP>
static void launch_crt16(const double* Q, uint64_t stride, const fa_crt_consts& C, uint64_t zlen,
                         int bits, uint64_t* z, uint64_t zn, uint64_t ngroups, uint8_t* flags,
                         uint8_t* agg, uint64_t nblk, hipStream_t s)
{
    FA_LAUNCH("fa_k_crt_mpn16", fa_k_crt_mpn16<NP>, dim3((unsigned)nblk), dim3(256), 0, s, Q,
              stride, C, zlen, bits, z, zn, ngroups, flags, agg);
}

// mode: 0 the CRT value c in [0, M); 1, 2 the positive and negated
// negative parts of the centred value (FP64 route only)
static int crt16(int np, const double* Q, uint64_t stride, const fa_crt_consts& C, uint64_t zlen,
                 int bits, uint64_t* z, uint64_t zn, uint64_t ngroups, uint8_t* flags, uint8_t* agg,
                 uint8_t* cin, uint64_t nblk, hipStream_t s, int mode = 0, uint64_t jlo = 0)
{
    if (fa_crt_fp_enabled || mode || jlo) {
        fa_crt_fp_consts F;
        fa_crt_fp_consts_init(&F, C);
        switch (np) {
        case 2: launch_crt_fp<2>(Q, stride, F, zlen, bits, z, zn, ngroups, flags, agg, nblk, s, mode, jlo); break;
        case 3: launch_crt_fp<3>(Q, stride, F, zlen, bits, z, zn, ngroups, flags, agg, nblk, s, mode, jlo); break;
        case 4: launch_crt_fp<4>(Q, stride, F, zlen, bits, z, zn, ngroups, flags, agg, nblk, s, mode, jlo); break;
        case 5: launch_crt_fp<5>(Q, stride, F, zlen, bits, z, zn, ngroups, flags, agg, nblk, s, mode, jlo); break;
        case 6: launch_crt_fp<6>(Q, stride, F, zlen, bits, z, zn, ngroups, flags, agg, nblk, s, mode, jlo); break;
        case 7: launch_crt_fp<7>(Q, stride, F, zlen, bits, z, zn, ngroups, flags, agg, nblk, s, mode, jlo); break;
        case 8: launch_crt_fp<8>(Q, stride, F, zlen, bits, z, zn, ngroups, flags, agg, nblk, s, mode, jlo); break;
        default: return fa_set_error(FA_E_ARG, "crt: np %d", np);
        }
        FA_LAUNCH("fa_k_carry_fused", fa_k_carry_fused, dim3(carry_grid(nblk)), dim3(1024), 0, s,
                  flags, agg, z, zn, bits / 8, ngroups, nblk);
        return FA_OK;
    }
    switch (np) {
    case 2: launch_crt16<2>(Q, stride, C, zlen, bits, z, zn, ngroups, flags, agg, nblk, s); break;
    case 3: launch_crt16<3>(Q, stride, C, zlen, bits, z, zn, ngroups, flags, agg, nblk, s); break;
    case 4: launch_crt16<4>(Q, stride, C, zlen, bits, z, zn, ngroups, flags, agg, nblk, s); break;
    case 5: launch_crt16<5>(Q, stride, C, zlen, bits, z, zn, ngroups, flags, agg, nblk, s); break;
    case 6: launch_crt16<6>(Q, stride, C, zlen, bits, z, zn, ngroups, flags, agg, nblk, s); break;
    case 7: launch_crt16<7>(Q, stride, C, zlen, bits, z, zn, ngroups, flags, agg, nblk, s); break;
    case 8: launch_crt16<8>(Q, stride, C, zlen, bits, z, zn, ngroups, flags, agg, nblk, s); break;
    default: return fa_set_error(FA_E_ARG, "crt16: np %d", np);
    }
    (void)cin;
    FA_LAUNCH("fa_k_carry_fused", fa_k_carry_fused, dim3(carry_grid(nblk)), dim3(1024), 0, s,
              flags, agg, z, zn, bits / 8, ngroups, nblk);
    return FA_OK;
}

// ------------------------------------------------------------------------
// host drivers
// ------------------------------------------------------------------------

#define LAUNCH_CHECK()                                                               \
    do {                                                                             \
        hipError_t e_ = hipGetLastError();                                           \
        if (e_ != hipSuccess)                                                        \
            return fa_set_error(FA_E_HIP, "kernel launch failed: %s", hipGetErrorString(e_)); \
    } while (0)

static fa_operand to_operand(const fa_pack_desc* d)
{
    fa_operand o;
    o.kind = d->kind;
    o.src = d->src;
    o.n = d->n;
    o.bits = d->bits;
    o.width = d->width;
    return o;
}

int fa_convolve_fast(int depth, const fa_operand& oa, const fa_operand& ob, int np, double* A,
                     double* B, uint64_t stride, const fa_prime_set& S, hipStream_t s, int sqr,
                     uint64_t zlen);

// coefficients of an operand descriptor (chunks of an integer)
static uint64_t desc_len(const fa_pack_desc* d)
{
    if (d->kind == FA_PACK_MPN) return (64 * d->n + (uint64_t)d->bits - 1) / (uint64_t)d->bits;
    return d->n;
}

int fa_fast_enabled = 1;   // switched off by tests comparing the two kernel sets

extern "C" FA_API void fa_set_fast_kernels(int on) { fa_fast_enabled = on; }

// squares take the one-transform route (switched off by tests comparing it
// with the two-transform product)
int fa_sqr_enabled = 1;
extern "C" FA_API void fa_set_square_route(int on) { fa_sqr_enabled = on; }

static int col_lg(int d1, int d2)
{
    int lgC = 0;
    while (((1 << d1) << (lgC + 1)) <= FA_LDS_COL_DOUBLES && lgC + 1 <= d2) lgC++;
    return lgC;
}

int fa_convolve(fa_thread_state* st, const fa_tables* T, int np, int offset, int depth,
                const fa_pack_desc* a, const fa_pack_desc* b, double* A, double* B,
                hipStream_t stream)
{
    (void)st;
    fa_prime_set S;
    fa_prime_set_init(&S, T, np, depth, offset, 1);
    fa_operand oa = to_operand(a), ob = to_operand(b);
    const uint64_t stride = fa_res_stride(depth);
    // a square (the same operand twice, mpn_mul.c:1665, 1513-1521): one forward
    // transform per prime
    const int sqr = fa_sqr_enabled && a->kind == b->kind && a->src == b->src && a->n == b->n &&
                    a->bits == b->bits && a->width == b->width;
    if (depth <= FA_ROW_LG) {
        size_t lds = (sqr ? 1 : 2) * stride * sizeof(double);
        FA_LAUNCH("fa_k_small_conv", fa_k_small_conv, dim3(np), dim3(256), lds, stream, oa, ob,
                  depth, A, stride, S, sqr);
        LAUNCH_CHECK();
        return FA_OK;
    }
    const uint64_t la = desc_len(a), lb = desc_len(b);
    const uint64_t zlen = la && lb ? la + lb - 1 : 0;
    if (fa_fast_enabled && fa_convolve_fast(depth, oa, ob, np, A, B, stride, S, stream, sqr, zlen)) {
        LAUNCH_CHECK();
        return FA_OK;
    }
    const int d2 = FA_ROW_LG, d1 = depth - d2;
    const int lgC = col_lg(d1, d2);
    const size_t col_lds = ((size_t)1 << d1 << lgC) * sizeof(double);
    dim3 gcol((unsigned)(1u << (d2 - lgC)), (unsigned)np, 2);
    FA_LAUNCH("fa_k_col_fwd", fa_k_col_fwd, gcol, dim3(256), col_lds, stream, oa, ob, d1, d2, lgC, A, B,
                       stride, S);
    LAUNCH_CHECK();
    FA_LAUNCH("fa_k_row_conv", fa_k_row_conv, dim3(1u << d1, np), dim3(256),
                       2 * ((size_t)1 << d2) * sizeof(double), stream, A, B, d2, stride, S);
    LAUNCH_CHECK();
    dim3 gci((unsigned)(1u << (d2 - lgC)), (unsigned)np, 1);
    FA_LAUNCH("fa_k_col_inv", fa_k_col_inv, gci, dim3(256), col_lds, stream, A, d1, d2, lgC, stride, S);
    LAUNCH_CHECK();
    return FA_OK;
}

// scratch layout for one mpn product
struct mpn_scratch {
    double* A;
    double* B;
    uint8_t* flags;
    uint8_t* agg;
    uint8_t* cin;
};

static size_t mpn_scratch_bytes(const fa_mpn_plan_t& P, uint64_t zn, uint64_t* ngroups,
                                uint64_t* nblk)
{
    uint64_t own = (uint64_t)P.bits / 8;
    *ngroups = (zn + own - 1) / own;
    *nblk = (*ngroups + (FA_CRT_WG - 2)) / (FA_CRT_WG - 1);
    uint64_t nblk16 = (*ngroups + 29) / 30;   // covers the 31-group workgroups
    size_t tr = (size_t)P.np * fa_res_stride(P.depth) * sizeof(double);
    size_t bytes = 2 * tr + *ngroups + 2 * nblk16 + 4096;
    return bytes;
}

// ------------------------------------------------------------------------
// integer products above one 2^24-point plan: block products added into z.
// z[off, zn) += p[0, len) with the carry rippling to the top of z: a local
// pass (sums, per-limb generate/propagate, per-block aggregates), a scan of
// the block aggregates (one workgroup), and a pass applying the carries.
// Blocks are 1024 limbs, 4 per thread.
// ------------------------------------------------------------------------
#define FA_ADD_BLK 1024

__global__ void __launch_bounds__(256) fa_k_add_local(uint64_t* __restrict__ z, uint64_t n,
                                                      const uint64_t* __restrict__ p,
                                                      uint64_t len, uint8_t* __restrict__ lf,
                                                      uint8_t* __restrict__ bf)
{
    __shared__ uint8_t sg[256], sp[256];
    const int t = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * FA_ADD_BLK + 4 * (uint64_t)t;
    uint8_t G = 0, P = 1;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const uint64_t i = base + e;
        if (i >= n) break;
        const uint64_t x = z[i];
        const uint64_t s = x + (i < len ? p[i] : 0);
        const uint8_t g = s < x, q = s == ~UINT64_C(0);
        z[i] = s;
        lf[i] = (uint8_t)(g | (q << 1));
        G = (uint8_t)(g | (q & G));
        P = (uint8_t)(P & q);
    }
    sg[t] = G;
    sp[t] = P;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {   // fold the 256 thread aggregates in order
        uint8_t g = 0, q = 1;
        if (t < h) {
            const uint8_t ge = sg[2 * t], pe = sp[2 * t], gl = sg[2 * t + 1], pl = sp[2 * t + 1];
            g = (uint8_t)(gl | (pl & ge));
            q = (uint8_t)(pl & pe);
        }
        __syncthreads();
        if (t < h) {
            sg[t] = g;
            sp[t] = q;
        }
        __syncthreads();
    }
    if (t == 0) bf[blockIdx.x] = (uint8_t)(sg[0] | (sp[0] << 1));
}

// one workgroup: carry into each block
__global__ void __launch_bounds__(1024) fa_k_add_scan(uint8_t* __restrict__ bf, uint64_t nb)
{
    __shared__ uint8_t sg[1024], sp[1024];
    const int t = threadIdx.x;
    const uint64_t m = (nb + 1023) / 1024;
    const uint64_t b0 = (uint64_t)t * m < nb ? (uint64_t)t * m : nb;
    const uint64_t b1 = b0 + m < nb ? b0 + m : nb;
    uint8_t G = 0, P = 1;
    for (uint64_t b = b0; b < b1; b++) {
        const uint8_t f = bf[b];
        G = (uint8_t)((f & 1) | ((f >> 1) & G));
        P = (uint8_t)(P & (f >> 1));
    }
    sg[t] = G;
    sp[t] = P;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        uint8_t g = sg[t], q = sp[t];
        if (t >= d) {
            g = (uint8_t)(g | (q & sg[t - d]));
            q = (uint8_t)(q & sp[t - d]);
        }
        __syncthreads();
        sg[t] = g;
        sp[t] = q;
        __syncthreads();
    }
    uint8_t c = t ? sg[t - 1] : 0;
    for (uint64_t b = b0; b < b1; b++) {
        const uint8_t f = bf[b];
        bf[b] = c;   // carry into block b
        c = (uint8_t)((f & 1) | ((f >> 1) & c));
    }
}

__global__ void __launch_bounds__(256) fa_k_add_apply(uint64_t* __restrict__ z, uint64_t n,
                                                      const uint8_t* __restrict__ lf,
                                                      const uint8_t* __restrict__ bf)
{
    __shared__ uint8_t sg[256], sp[256];
    const int t = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * FA_ADD_BLK + 4 * (uint64_t)t;
    uint8_t G = 0, P = 1;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const uint64_t i = base + e;
        if (i >= n) break;
        const uint8_t f = lf[i];
        G = (uint8_t)((f & 1) | ((f >> 1) & G));
        P = (uint8_t)(P & (f >> 1));
    }
    sg[t] = G;
    sp[t] = P;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        uint8_t g = sg[t], q = sp[t];
        if (t >= d) {
            g = (uint8_t)(g | (q & sg[t - d]));
            q = (uint8_t)(q & sp[t - d]);
        }
        __syncthreads();
        sg[t] = g;
        sp[t] = q;
        __syncthreads();
    }
    const uint8_t cb = bf[blockIdx.x];
    // carry into this thread: the threads before it, after the block's carry
    uint8_t c = t ? (uint8_t)(sg[t - 1] | (sp[t - 1] & cb)) : cb;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const uint64_t i = base + e;
        if (i >= n) break;
        const uint8_t f = lf[i];
        if (c) z[i] += 1;
        c = (uint8_t)((f & 1) | ((f >> 1) & c));
    }
}

// z[off, zn) += p[0, len), len <= zn - off; flags: zn - off + (zn - off) /
// 1024 + 1 bytes of device scratch
static int mpn_add_device(uint64_t* z, uint64_t zn, uint64_t off, const uint64_t* p, uint64_t len,
                          uint8_t* flags, hipStream_t s)
{
    const uint64_t n = zn - off;
    if (!n) return FA_OK;
    const uint64_t nb = (n + FA_ADD_BLK - 1) / FA_ADD_BLK;
    uint8_t* lf = flags;
    uint8_t* bf = flags + n;
    FA_LAUNCH("fa_k_add_local", fa_k_add_local, dim3((unsigned)nb), dim3(256), 0, s, z + off, n, p,
              len, lf, bf);
    FA_LAUNCH("fa_k_add_scan", fa_k_add_scan, dim3(1), dim3(1024), 0, s, bf, nb);
    FA_LAUNCH("fa_k_add_apply", fa_k_add_apply, dim3((unsigned)nb), dim3(256), 0, s, z + off, n, lf,
              bf);
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "add launch");
}

__global__ void __launch_bounds__(256) fa_k_zero(uint64_t* __restrict__ z, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) z[i] = 0;
}

// longest x <= n with x + other <= cap and a plan for x x other limbs
static uint64_t fit_block(uint64_t n, uint64_t other, uint64_t cap)
{
    if (other >= cap) return 0;
    uint64_t lo = 0, hi = n < cap - other ? n : cap - other;
    fa_mpn_plan_t P;
    while (lo < hi) {
        const uint64_t mid = hi - (hi - lo) / 2;
        if (fa_mpn_choose_plan(mid, other, &P) == FA_OK) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// total limbs above which integer products go in blocks (tests force small
// blocks; 0 = only products that no single plan covers)
uint64_t fa_mpn_block_cap = UINT64_MAX;
extern "C" FA_API void fa_set_mpn_block_cap(uint64_t limbs)
{
    fa_mpn_block_cap = limbs ? limbs : UINT64_MAX;
}

static int mpn_mul_plan_device(uint64_t* z, const uint64_t* a, uint64_t an, const uint64_t* b,
                               uint64_t bn, hipStream_t stream, int np_force, int bits_force,
                               const fa_mpn_plan_t* chosen = nullptr);

// a x b above one plan (or above the block cap): the longer operand in
// blocks beside the whole shorter one when that leaves blocks at least as
// long as the shorter operand, else both in equal blocks; every block
// product goes through the ordinary path and is added into z
static int mpn_mul_blocked(uint64_t* z, const uint64_t* a, uint64_t an, const uint64_t* b,
                           uint64_t bn, hipStream_t stream)
{
    if (an < bn) {
        const uint64_t* t = a; a = b; b = t;
        const uint64_t tn = an; an = bn; bn = tn;
    }
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    if (!stream) stream = st->stream;   // the block products run there too
    const uint64_t cap = fa_mpn_block_cap;
    // (1) the shorter operand whole
    uint64_t lb = bn;
    uint64_t la = fit_block(an, lb, cap);
    uint64_t cost = la ? (an + la - 1) / la : UINT64_MAX;
    // (2) equal blocks: the longest s with a plan for s x s within the cap
    {
        uint64_t lo = 0, hi = bn;
        fa_mpn_plan_t Q;
        while (lo < hi) {
            const uint64_t mid = hi - (hi - lo) / 2;
            if (2 * mid <= cap && fa_mpn_choose_plan(mid, mid, &Q) == FA_OK) lo = mid;
            else hi = mid - 1;
        }
        const uint64_t la2 = lo ? fit_block(an, lo, cap) : 0;
        const uint64_t cost2 =
            la2 ? ((an + la2 - 1) / la2) * ((bn + lo - 1) / lo) : UINT64_MAX;
        if (cost2 < cost) {
            lb = lo;
            la = la2;
            cost = cost2;
        }
    }
    if (la == 0 || lb == 0)
        return fa_set_error(FA_E_RANGE, "no block plan for %llu x %llu limbs",
                            (unsigned long long)an, (unsigned long long)bn);
    const uint64_t zn = an + bn;
    const uint64_t plen = la + lb;
    const size_t flag_bytes = (size_t)zn + zn / FA_ADD_BLK + 64;
    void* tb;
    int rc = fa_tbuf_reserve(st, plen * 8 + flag_bytes, &tb, stream);
    if (rc) return rc;
    uint64_t* tmp = (uint64_t*)tb;
    uint8_t* flags = (uint8_t*)(tmp + plen);
    FA_LAUNCH("fa_k_zero", fa_k_zero, dim3((unsigned)((zn + 255) / 256)), dim3(256), 0, stream, z,
              zn);
    for (uint64_t i = 0; i < an; i += la) {
        const uint64_t na = an - i < la ? an - i : la;
        for (uint64_t j = 0; j < bn; j += lb) {
            const uint64_t nbk = bn - j < lb ? bn - j : lb;
            rc = mpn_mul_plan_device(tmp, a + i, na, b + j, nbk, stream, 0, 0);
            if (rc) return rc;
            rc = mpn_add_device(z, zn, i + j, tmp, na + nbk, flags, stream);
            if (rc) return rc;
        }
    }
    return FA_OK;
}

static int mpn_mul_device(uint64_t* z, const uint64_t* a, uint64_t an, const uint64_t* b,
                          uint64_t bn, hipStream_t stream, int np_force = 0, int bits_force = 0)
{
    if (!np_force) {
        fa_mpn_plan_t Q;
        const int rq = fa_mpn_choose_plan(an, bn, &Q);
        if (an + bn > fa_mpn_block_cap || rq == FA_E_RANGE)
            return mpn_mul_blocked(z, a, an, b, bn, stream);
        if (rq) return rq;
        return mpn_mul_plan_device(z, a, an, b, bn, stream, 0, 0, &Q);
    }
    return mpn_mul_plan_device(z, a, an, b, bn, stream, np_force, bits_force);
}

// CRT of zlen coefficients of `bits`-bit chunks into z[0, zn) (the low zn
// limbs of the sum; zeros above it), carry flags in the thread's scratch
int fa_crt_mpn_export(int np, const double* A, uint64_t stride, uint64_t zlen, int bits,
                      uint64_t* z, uint64_t zn, hipStream_t stream, int mode, uint64_t jlo)
{
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    if (!stream) stream = st->stream;
    const uint64_t own = (uint64_t)bits / 8;
    const uint64_t ngroups = (zn + own - 1) / own;
    const uint64_t nblk16 = (ngroups + (FA_CRT16_GROUPS - 2)) / (FA_CRT16_GROUPS - 1);
    const size_t bytes = ((ngroups + 7) & ~(size_t)7) + 2 * (nblk16 + 8) + 4096;
    void* buf;
    int rc = fa_scratch_reserve(st, bytes, &buf, stream);
    if (rc) return rc;
    uint8_t* flags = (uint8_t*)buf;
    uint8_t* agg = flags + ((ngroups + 7) & ~UINT64_C(7));
    uint8_t* cin = agg + ((nblk16 + 7) & ~UINT64_C(7));
    fa_crt_consts C;
    fa_crt_consts_init(&C, np, 0);
    rc = crt16(np, A, stride, C, zlen, bits, z, zn, ngroups, flags, agg, cin, nblk16, stream,
               mode, jlo);
    if (rc) return rc;
    rc = fa_buffers_leave(st, stream);
    if (rc) return rc;
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "crt launch");
}

// the launches of one product on plan P, with scratch `buf` of at least
// mpn_scratch_bytes(P, an + bn) bytes
static int mpn_mul_run(uint64_t* z, const uint64_t* a, uint64_t an, const uint64_t* b,
                       uint64_t bn, hipStream_t stream, const fa_mpn_plan_t& P,
                       const fa_tables* T, void* buf)
{
    uint64_t zn = an + bn, ngroups, nblk;
    (void)mpn_scratch_bytes(P, zn, &ngroups, &nblk);
    size_t tr = (size_t)P.np * fa_res_stride(P.depth) * sizeof(double);
    mpn_scratch W;
    W.A = (double*)buf;
    W.B = (double*)((char*)buf + tr);
    W.flags = (uint8_t*)((char*)buf + 2 * tr);
    // aggregates start on an 8-byte boundary: the carry kernel reads them
    // as 64-bit words
    W.agg = W.flags + ((ngroups + 7) & ~UINT64_C(7));
    W.cin = W.agg + (((ngroups + 29) / 30 + 7) & ~UINT64_C(7));

    fa_pack_desc da = {FA_PACK_MPN, a, an, P.bits, 1};
    fa_pack_desc db = {FA_PACK_MPN, b, bn, P.bits, 1};
    int rc = fa_convolve(NULL, T, P.np, 0, P.depth, &da, &db, W.A, W.B, stream);
    if (rc) return rc;

    fa_crt_consts C;
    fa_crt_consts_init(&C, P.np, 0);
    const uint64_t stride = fa_res_stride(P.depth);
    if (fa_fast_enabled) {
        uint64_t nblk16 = (ngroups + (FA_CRT16_GROUPS - 2)) / (FA_CRT16_GROUPS - 1);
        rc = crt16(P.np, W.A, stride, C, P.zlen, P.bits, z, zn, ngroups, W.flags, W.agg, W.cin,
                   nblk16, stream);
        if (rc) return rc;
        LAUNCH_CHECK();
        return FA_OK;
    }
    FA_LAUNCH("fa_k_crt_mpn", fa_k_crt_mpn, dim3((unsigned)nblk), dim3(FA_CRT_WG), 0, stream, W.A,
              stride, C, P.zlen, P.bits, z, zn, ngroups, W.flags, W.agg);
    LAUNCH_CHECK();
    FA_LAUNCH("fa_k_carry_scan", fa_k_carry_scan, dim3(1), dim3(1024), 0, stream, W.agg, W.cin, nblk);
    LAUNCH_CHECK();
    FA_LAUNCH("fa_k_carry_apply", fa_k_carry_apply, dim3((unsigned)nblk), dim3(FA_CRT_WG), 0, stream,
              W.flags, W.cin, z, zn, P.bits / 8, ngroups);
    LAUNCH_CHECK();
    return FA_OK;
}

// one product on one plan (`chosen`: the automatic plan, already made)
static int mpn_mul_plan_device(uint64_t* z, const uint64_t* a, uint64_t an, const uint64_t* b,
                               uint64_t bn, hipStream_t stream, int np_force, int bits_force,
                               const fa_mpn_plan_t* chosen)
{
    fa_mpn_plan_t P;
    int rc = FA_OK;
    if (chosen) P = *chosen;
    else rc = np_force ? fa_mpn_fixed_plan(an, bn, np_force, bits_force, &P)
                       : fa_mpn_choose_plan(an, bn, &P);
    if (rc) return rc;
    fa_tables* T = fa_get_tables(P.depth);
    if (!T) return FA_E_NODEV;
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    if (!stream) stream = st->stream;
    uint64_t ngroups, nblk;
    size_t bytes = mpn_scratch_bytes(P, an + bn, &ngroups, &nblk);
    void* buf;
    rc = fa_scratch_reserve(st, bytes, &buf, stream);
    if (rc) return rc;
    return mpn_mul_run(z, a, an, b, bn, stream, P, T, buf);
}

int fa_mpn_mul_prealloc(uint64_t* z, const uint64_t* a, uint64_t an, const uint64_t* b,
                        uint64_t bn, hipStream_t stream, void* buf, size_t cap)
{
    fa_mpn_plan_t P;
    int rc = fa_mpn_choose_plan(an, bn, &P);
    if (rc) return rc;
    uint64_t ngroups, nblk;
    const size_t bytes = mpn_scratch_bytes(P, an + bn, &ngroups, &nblk);
    if (bytes > cap)
        return fa_set_error(FA_E_RANGE,
                            "%llu x %llu limbs need %zu bytes of scratch, above the %zu available",
                            (unsigned long long)an, (unsigned long long)bn, bytes, cap);
    const fa_tables* T = fa_tables_ready(P.depth);
    if (!T)
        return fa_set_error(FA_E_ARG,
                            "twiddle tables of depth %d are not built on this device (fa_init)",
                            P.depth);
    return mpn_mul_run(z, a, an, b, bn, stream, P, T, buf);
}

// a device-pointer call: enqueued on the caller's stream, after which the
// thread's buffers are marked busy on that stream (fa_buffers_leave)
static int mpn_mul_device_call(uint64_t* z, const uint64_t* a, uint64_t an, const uint64_t* b,
                               uint64_t bn, hipStream_t stream, int np, int bits)
{
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = stream ? stream : st->stream;
    const int rc = mpn_mul_device(z, a, an, b, bn, s, np, bits);
    const int rc2 = fa_buffers_leave(st, s);
    return rc ? rc : rc2;
}

extern "C" int fa_mpn_mul_device(uint64_t* z, const uint64_t* a, int64_t an, const uint64_t* b,
                                 int64_t bn, void* stream)
{
    if (an <= 0 || bn <= 0 || !z || !a || !b)
        return fa_set_error(FA_E_ARG, "fa_mpn_mul_device: bad arguments");
    return mpn_mul_device_call(z, a, (uint64_t)an, b, (uint64_t)bn, (hipStream_t)stream, 0, 0);
}

static int mpn_mul_host(uint64_t* z, const uint64_t* a, int64_t an, const uint64_t* b, int64_t bn,
                        int np, int bits)
{
    if (an <= 0 || bn <= 0 || !z || !a || !b)
        return fa_set_error(FA_E_ARG, "fa_mpn_mul: bad arguments");
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    uint64_t zn = (uint64_t)(an + bn);
    size_t bytes = ((size_t)an + (size_t)bn + zn) * sizeof(uint64_t);
    void* hb;
    int rc = fa_hbuf_reserve(st, bytes, &hb, st->stream);
    if (rc) return rc;
    uint64_t* da = (uint64_t*)hb;
    uint64_t* db = da + an;
    uint64_t* dz = db + bn;
    // a square keeps one device copy, so the device path sees a == b
    const bool sqr = a == b && an == bn;
    if (sqr) db = da;
    if (hipMemcpyAsync(da, a, an * sizeof(uint64_t), hipMemcpyHostToDevice, st->stream) != hipSuccess ||
        (!sqr && hipMemcpyAsync(db, b, bn * sizeof(uint64_t), hipMemcpyHostToDevice,
                                st->stream) != hipSuccess))
        return fa_set_error(FA_E_HIP, "operand upload failed");
    rc = mpn_mul_device(dz, da, (uint64_t)an, db, (uint64_t)bn, st->stream, np, bits);
    if (rc) {
        (void)hipStreamSynchronize(st->stream);
        return rc;
    }
    if (hipMemcpyAsync(z, dz, zn * sizeof(uint64_t), hipMemcpyDeviceToHost, st->stream) != hipSuccess)
        return fa_set_error(FA_E_HIP, "product download failed");
    if (hipStreamSynchronize(st->stream) != hipSuccess)
        return fa_set_error(FA_E_HIP, "stream synchronize failed: %s",
                            hipGetErrorString(hipGetLastError()));
    return FA_OK;
}

extern "C" int fa_mpn_mul(uint64_t* z, const uint64_t* a, int64_t an, const uint64_t* b,
                          int64_t bn)
{
    return mpn_mul_host(z, a, an, b, bn, 0, 0);
}

// ------------------------------------------------------------------------
// products mod 2^N + 1, N = 64 n (sd_fft_mpn_mulmod_2expp1, negacyclic.c:
// 491-501): operands of n + 1 limbs with value <= 2^N, result normalized to
// [0, 2^N].  The full product P = x y <= 2^(2N) is formed by the ordinary
// path; with lo = P mod 2^N and hi = P >> N <= 2^N, x y = lo - hi (mod
// 2^N + 1): D = lo - hi over n + 1 limbs, and D + 2^N + 1 when D < 0.
// One workgroup, each thread a run of consecutive limbs, borrow and carry
// chains resolved by workgroup scans of (generate, propagate) bits.
// ------------------------------------------------------------------------

// exclusive scan of per-thread (g, p) chains over the workgroup: the carry
// into thread t given carry-in c0 at thread 0
FA_DEV uint8_t fa_wg_chain_in(uint8_t g, uint8_t p, uint8_t c0, uint8_t* sg, uint8_t* sp)
{
    const int t = threadIdx.x, nt = blockDim.x;
    sg[t] = g;
    sp[t] = p;
    __syncthreads();
    for (int d = 1; d < nt; d <<= 1) {
        uint8_t gg = sg[t], pp = sp[t];
        if (t >= d) {
            gg = (uint8_t)(gg | (pp & sg[t - d]));
            pp = (uint8_t)(pp & sp[t - d]);
        }
        __syncthreads();
        sg[t] = gg;
        sp[t] = pp;
        __syncthreads();
    }
    const uint8_t c = t ? (uint8_t)(sg[t - 1] | (sp[t - 1] & c0)) : c0;
    __syncthreads();
    return c;
}

__global__ void __launch_bounds__(1024) fa_k_fold_2expp1(uint64_t* __restrict__ z,
                                                         const uint64_t* __restrict__ P,
                                                         uint64_t n)
{
    __shared__ uint8_t sg[1024], sp[1024];
    __shared__ uint8_t sneg;
    const int t = threadIdx.x;
    const uint64_t L = (n + 1 + 1023) / 1024;        // limbs of d = lo - hi per thread
    const uint64_t i0 = (uint64_t)t * L < n + 1 ? (uint64_t)t * L : n + 1;
    const uint64_t i1 = i0 + L < n + 1 ? i0 + L : n + 1;
    // 1. d = lo - hi over n + 1 limbs (lo_n = 0): borrow generate / propagate
    uint8_t G = 0, Q = 1;
    for (uint64_t i = i0; i < i1; i++) {
        const uint64_t x = i < n ? P[i] : 0, y = P[n + i];
        const uint8_t g = x < y, q = x == y;
        G = (uint8_t)(g | (q & G));
        Q = (uint8_t)(Q & q);
    }
    uint8_t b = fa_wg_chain_in(G, Q, 0, sg, sp);
    for (uint64_t i = i0; i < i1; i++) {
        const uint64_t x = i < n ? P[i] : 0, y = P[n + i];
        z[i] = x - y - b;
        b = (uint8_t)(x < y || (x == y && b));
    }
    if (t == 1023) sneg = b;         // borrow out of the last limb: D < 0
    __syncthreads();
    const bool neg = sneg;
    // 2. D < 0: r = D + 2^N + 1, i.e. the low n limbs (D mod 2^N) plus one,
    // the carry out of them giving limb n; D >= 0: limb n is 0
    if (!neg) {
        if (t == 0) z[n] = 0;
        return;
    }
    const uint64_t m0 = i0 < n ? i0 : n, m1 = i1 < n ? i1 : n;
    uint8_t Pn = 1;
    for (uint64_t i = m0; i < m1; i++) Pn = (uint8_t)(Pn & (z[i] == ~UINT64_C(0)));
    uint8_t c = fa_wg_chain_in(0, Pn, 1, sg, sp);
    for (uint64_t i = m0; i < m1; i++) {
        const uint64_t v = z[i] + c;
        c = (uint8_t)(c && v == 0);
        z[i] = v;
    }
    if (m1 == n && m0 < m1) z[n] = c;   // the run holding limb n - 1
}

// z[0, zn) -= w[0, zn) as two's complement; then |z| and *sign = 1 when the
// difference is negative (the top bit; the caller's zn leaves room for it),
// 0 otherwise, and 0 for a zero difference (export_signed.c:588-599).  One
// workgroup, runs of consecutive limbs per thread, chains by workgroup scans.
__global__ void __launch_bounds__(1024) fa_k_signed_diff(uint64_t* __restrict__ z,
                                                         const uint64_t* __restrict__ w,
                                                         uint64_t zn, int32_t* __restrict__ sign)
{
    __shared__ uint8_t sg[1024], sp[1024];
    __shared__ uint8_t sneg, snz;
    const int t = threadIdx.x;
    const uint64_t L = (zn + 1023) / 1024;
    const uint64_t i0 = (uint64_t)t * L < zn ? (uint64_t)t * L : zn;
    const uint64_t i1 = i0 + L < zn ? i0 + L : zn;
    uint8_t G = 0, Q = 1;
    for (uint64_t i = i0; i < i1; i++) {
        const uint64_t x = z[i], y = w[i];
        G = (uint8_t)((x < y) | ((x == y) & G));
        Q = (uint8_t)(Q & (x == y));
    }
    uint8_t b = fa_wg_chain_in(G, Q, 0, sg, sp);
    for (uint64_t i = i0; i < i1; i++) {
        const uint64_t x = z[i], y = w[i];
        z[i] = x - y - b;
        b = (uint8_t)(x < y || (x == y && b));
    }
    if (t == 0) snz = 0;
    __syncthreads();
    if (i1 == zn && i0 < i1) sneg = (uint8_t)(z[zn - 1] >> 63);
    __syncthreads();
    const bool neg = sneg;
    if (neg) {
        // |D| = ~D + 1
        uint8_t P = 1;
        for (uint64_t i = i0; i < i1; i++) P = (uint8_t)(P & (z[i] == 0));   // ~z all ones
        uint8_t c = fa_wg_chain_in(0, P, 1, sg, sp);
        for (uint64_t i = i0; i < i1; i++) {
            const uint64_t v = ~z[i] + c;
            c = (uint8_t)(c && v == 0);
            z[i] = v;
        }
    }
    uint8_t nz = 0;
    for (uint64_t i = i0; i < i1; i++) nz |= z[i] != 0;
    if (nz) snz = 1;
    __syncthreads();
    if (t == 0) *sign = (neg && snz) ? 1 : 0;
}

int fa_signed_diff_launch(uint64_t* z, const uint64_t* w, uint64_t zn, int32_t* sign,
                          hipStream_t s)
{
    FA_LAUNCH("fa_k_signed_diff", fa_k_signed_diff, dim3(1), dim3(1024), 0, s, z, w, zn, sign);
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "signed diff launch");
}

// z[0, zn) = limbs [lo, lo + zn) of the magnitude t (one workgroup), plus
// one when `up` is set, the value is negative (*sign) and a limb of t below
// lo is nonzero: the wide-plan truncated signed export negates only the
// window's limbs of the two's-complement sum, ~W + 1, which is the ceiling
// of |value| / 2^(64 lo) (export_signed.c:592-604)
__global__ void __launch_bounds__(1024) fa_k_signed_window(uint64_t* __restrict__ z, uint64_t zn,
                                                           const uint64_t* __restrict__ t,
                                                           uint64_t lo, int up,
                                                           const int32_t* __restrict__ sign)
{
    __shared__ uint8_t sg[1024], sp[1024];
    __shared__ int snz;
    const int k = threadIdx.x;
    if (k == 0) snz = 0;
    __syncthreads();
    int nz = 0;
    if (up && *sign)
        for (uint64_t i = k; i < lo; i += 1024) nz |= t[i] != 0;
    if (nz) snz = 1;
    __syncthreads();
    const uint8_t c0 = (uint8_t)snz;
    const uint64_t L = (zn + 1023) / 1024;
    const uint64_t i0 = (uint64_t)k * L < zn ? (uint64_t)k * L : zn;
    const uint64_t i1 = i0 + L < zn ? i0 + L : zn;
    uint8_t P = 1;
    for (uint64_t i = i0; i < i1; i++) P = (uint8_t)(P & (t[lo + i] == ~UINT64_C(0)));
    uint8_t c = fa_wg_chain_in(0, P, c0, sg, sp);
    for (uint64_t i = i0; i < i1; i++) {
        const uint64_t v = t[lo + i] + c;
        c = (uint8_t)(c && v == 0);
        z[i] = v;
    }
}

int fa_signed_window_launch(uint64_t* z, uint64_t zn, const uint64_t* t, uint64_t lo, int up,
                            const int32_t* sign, hipStream_t s)
{
    FA_LAUNCH("fa_k_signed_window", fa_k_signed_window, dim3(1), dim3(1024), 0, s, z, zn, t, lo,
              up, sign);
    return hipGetLastError() == hipSuccess ? FA_OK
                                           : fa_set_error(FA_E_HIP, "signed window launch");
}

// z[0, n] = P mod 2^(64 n) + 1 for the 2n + 2 limbs of P (one workgroup)
int fa_fold_2expp1_launch(uint64_t* z, const uint64_t* P, uint64_t n, hipStream_t s)
{
    FA_LAUNCH("fa_k_fold_2expp1", fa_k_fold_2expp1, dim3(1), dim3(1024), 0, s, z, P, n);
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "fold launch failed");
}

// z[0, n] = x y mod 2^(64 n) + 1 on device operands of n + 1 limbs with
// values <= 2^(64 n), enqueued on `stream`
extern "C" int fa_mpn_mulmod_2expp1_device(uint64_t* z, const uint64_t* x, const uint64_t* y,
                                           int64_t n, void* stream)
{
    if (n <= 0 || !z || !x || !y)
        return fa_set_error(FA_E_ARG, "fa_mpn_mulmod_2expp1_device: bad arguments");
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = stream ? (hipStream_t)stream : st->stream;
    // one plan covers the product (block products would share the buffers)
    fa_mpn_plan_t Q;
    const uint64_t pn = 2 * (uint64_t)n + 2;
    if (pn > fa_mpn_block_cap || fa_mpn_choose_plan((uint64_t)n + 1, (uint64_t)n + 1, &Q))
        return fa_set_error(FA_E_RANGE, "mulmod 2^N+1: n = %lld above one transform",
                            (long long)n);
    void* hb;
    int rc = fa_hbuf_reserve(st, pn * sizeof(uint64_t), &hb, s);
    if (rc) return rc;
    uint64_t* P = (uint64_t*)hb;
    rc = mpn_mul_device(P, x, (uint64_t)n + 1, y, (uint64_t)n + 1, s);
    if (rc) return rc;
    FA_LAUNCH("fa_k_fold_2expp1", fa_k_fold_2expp1, dim3(1), dim3(1024), 0, s, z, P, (uint64_t)n);
    if (hipGetLastError() != hipSuccess) return fa_set_error(FA_E_HIP, "fold launch failed");
    return fa_buffers_leave(st, s);
}

// host operands: checked to be normalized (top limb 0, or 1 with the rest 0)
extern "C" int fa_mpn_mulmod_2expp1(uint64_t* z, const uint64_t* x, const uint64_t* y, int64_t n)
{
    if (n <= 0 || !z || !x || !y)
        return fa_set_error(FA_E_ARG, "fa_mpn_mulmod_2expp1: bad arguments");
    for (const uint64_t* v : {x, y}) {
        if (v[n] > 1) return fa_set_error(FA_E_ARG, "operand above 2^N: top limb %llu",
                                          (unsigned long long)v[n]);
        if (v[n] == 1)
            for (int64_t i = 0; i < n; i++)
                if (v[i]) return fa_set_error(FA_E_ARG, "operand above 2^N");
    }
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    const uint64_t m = (uint64_t)n + 1;
    void* tb;
    int rc = fa_tbuf_reserve(st, 3 * m * sizeof(uint64_t), &tb, st->stream);
    if (rc) return rc;
    uint64_t* dx = (uint64_t*)tb;
    uint64_t* dy = x == y ? dx : dx + m;
    uint64_t* dz = dx + 2 * m;
    if (hipMemcpyAsync(dx, x, m * 8, hipMemcpyHostToDevice, st->stream) != hipSuccess ||
        (dy != dx && hipMemcpyAsync(dy, y, m * 8, hipMemcpyHostToDevice, st->stream) != hipSuccess))
        return fa_set_error(FA_E_HIP, "operand upload failed");
    rc = fa_mpn_mulmod_2expp1_device(dz, dx, dy, n, st->stream);
    if (rc) {
        (void)hipStreamSynchronize(st->stream);
        return rc;
    }
    if (hipMemcpyAsync(z, dz, m * 8, hipMemcpyDeviceToHost, st->stream) != hipSuccess ||
        hipStreamSynchronize(st->stream) != hipSuccess)
        return fa_set_error(FA_E_HIP, "result download failed");
    return FA_OK;
}

// _mpn_ctx_mpn_mul_range (mpn_mul.c:1617-1978) and flint_mpn_mulmid_fft_small
// (mpn_extras/mulmid_fft_small.c:26-33): limbs [lo, hi) of a*b as the
// reference returns them, a lower approximation whose dropped carry
// depends on FLINT's plan for (an, bn) (doc/source/mpn_extras.rst:221-231).
// The exact product is formed on the device and the reference's words are
// recovered from it by an exact correction (fa_window.hip); nthreads is the
// FLINT thread count whose plan the words follow, 0 for the exact window.
static int mpn_window_device(uint64_t* z, uint64_t lo, uint64_t hi, const uint64_t* a,
                             uint64_t an, const uint64_t* b, uint64_t bn, int nthreads,
                             fa_thread_state* st, hipStream_t s)
{
    const uint64_t zn = an + bn;
    const size_t wb = fa_window_work_bytes(an, bn, lo, nthreads);
    // the window's own buffer (block products, mpn_mul_blocked, use tbuf):
    // the full product, then the correction's work space
    void* hb;
    int rc = fa_hbuf_reserve(st, zn * sizeof(uint64_t) + wb, &hb, s);
    if (rc) return rc;
    uint64_t* dz = (uint64_t*)hb;
    rc = mpn_mul_device(dz, a, an, b, bn, s);
    if (rc) return rc;
    if (wb) {
        rc = fa_window_fix(dz + lo, dz, lo, hi, a, an, b, bn, nthreads, dz + zn, s);
        if (rc) return rc;
    }
    if (hipMemcpyAsync(z, dz + lo, (hi - lo) * sizeof(uint64_t), hipMemcpyDeviceToDevice, s) !=
        hipSuccess)
        return fa_set_error(FA_E_HIP, "window copy failed");
    return fa_buffers_leave(st, s);
}

// host operands and output; the words of FLINT on nthreads threads (0: the
// exact window)
extern "C" int fa_mpn_mul_range_host(uint64_t* z, uint64_t lo, uint64_t hi, const uint64_t* a,
                                     int64_t an, const uint64_t* b, int64_t bn, int nthreads)
{
    if (an <= 0 || bn <= 0 || !a || !b || lo >= hi || hi > (uint64_t)(an + bn) || !z ||
        nthreads < 0)
        return fa_set_error(FA_E_ARG, "fa_mpn_mul_range: bad arguments");
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    const uint64_t zn = (uint64_t)(an + bn);
    const size_t wb = fa_window_work_bytes((uint64_t)an, (uint64_t)bn, lo, nthreads);
    void* hb;
    int rc = fa_hbuf_reserve(st, ((size_t)an + (size_t)bn + zn) * sizeof(uint64_t) + wb, &hb,
                             st->stream);
    if (rc) return rc;
    uint64_t* da = (uint64_t*)hb;
    uint64_t* db = da + an;
    uint64_t* dz = db + bn;
    const bool sqr = a == b && an == bn;
    if (sqr) db = da;
    if (hipMemcpyAsync(da, a, an * sizeof(uint64_t), hipMemcpyHostToDevice, st->stream) != hipSuccess ||
        (!sqr && hipMemcpyAsync(db, b, bn * sizeof(uint64_t), hipMemcpyHostToDevice,
                                st->stream) != hipSuccess))
        return fa_set_error(FA_E_HIP, "operand upload failed");
    rc = mpn_mul_device(dz, da, (uint64_t)an, db, (uint64_t)bn, st->stream);
    if (rc == FA_OK && wb)
        rc = fa_window_fix(dz + lo, dz, lo, hi, da, (uint64_t)an, db, (uint64_t)bn, nthreads,
                           dz + zn, st->stream);
    if (rc) {
        (void)hipStreamSynchronize(st->stream);
        return rc;
    }
    if (hipMemcpyAsync(z, dz + lo, (hi - lo) * sizeof(uint64_t), hipMemcpyDeviceToHost,
                       st->stream) != hipSuccess)
        return fa_set_error(FA_E_HIP, "window download failed");
    if (hipStreamSynchronize(st->stream) != hipSuccess)
        return fa_set_error(FA_E_HIP, "stream synchronize failed: %s",
                            hipGetErrorString(hipGetLastError()));
    return FA_OK;
}

extern "C" int fa_mpn_mul_range(uint64_t* z, uint64_t lo, uint64_t hi, const uint64_t* a,
                                int64_t an, const uint64_t* b, int64_t bn)
{
    return fa_mpn_mul_range_host(z, lo, hi, a, an, b, bn, 1);
}

// the same on device operands, enqueued on `stream`
extern "C" int fa_mpn_mul_range_device(uint64_t* z, uint64_t lo, uint64_t hi, const uint64_t* a,
                                       int64_t an, const uint64_t* b, int64_t bn, void* stream)
{
    return fa_mpn_mul_range_threads(z, lo, hi, a, an, b, bn, 1, stream);
}

extern "C" int fa_mpn_mul_range_threads(uint64_t* z, uint64_t lo, uint64_t hi,
                                        const uint64_t* a, int64_t an, const uint64_t* b,
                                        int64_t bn, int nthreads, void* stream)
{
    if (an <= 0 || bn <= 0 || !a || !b || lo >= hi || hi > (uint64_t)(an + bn) || !z ||
        nthreads < 0)
        return fa_set_error(FA_E_ARG, "fa_mpn_mul_range_threads: bad arguments");
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = stream ? (hipStream_t)stream : st->stream;
    const uint64_t zn = (uint64_t)(an + bn);
    if (lo == 0 && hi == zn) return mpn_mul_device_call(z, a, an, b, bn, s, 0, 0);
    return mpn_window_device(z, lo, hi, a, (uint64_t)an, b, (uint64_t)bn, nthreads, st, s);
}

extern "C" int fa_mpn_mul_plan(uint64_t* z, const uint64_t* a, int64_t an, const uint64_t* b,
                               int64_t bn, int np, int bits)
{
    if (np == 0) return fa_set_error(FA_E_ARG, "np must be given");
    return mpn_mul_host(z, a, an, b, bn, np, bits);
}

extern "C" int fa_mpn_mul_device_plan(uint64_t* z, const uint64_t* a, int64_t an,
                                      const uint64_t* b, int64_t bn, int np, int bits,
                                      void* stream)
{
    if (an <= 0 || bn <= 0 || !z || !a || !b || np == 0)
        return fa_set_error(FA_E_ARG, "fa_mpn_mul_device_plan: bad arguments");
    return mpn_mul_device_call(z, a, (uint64_t)an, b, (uint64_t)bn, (hipStream_t)stream, np,
                               bits);
}
