// fa_fast.hip -- register radix-16 kernels for the two-level transforms with
// a 4096-point row (d2 = 12) and a column of 2^d1 points, d1 = 1..12, i.e.
// transform lengths 2^13 .. 2^24.
//
// A 2^(4s)-point transform is s stages of radix 16: each thread holds 16
// elements in registers and applies four levels (two radix-4 butterflies,
// fa_fwd4 / fa_inv4 of fa_ntt.h) per stage; between stages the elements are
// exchanged through LDS.  At stage j a thread owns, inside one block of
// level 4j (block b, offset i0), the 16 positions
//     (b << (D - 4j)) + i0 + k * 2^(D - 4j - 4),   k < 16,
// so after the last stage thread t holds the 16 consecutive outputs
// 16t .. 16t+15 (bit-reversed evaluation order).  LDS rows are padded by
// one double per 16 so that all three access patterns are conflict-free;
// the column passes interleave their columns in LDS (element (c, r) of a
// tile at PADX(r) * columns + c) for the same reason.
#include "fa_internal.h"
#include "fa_ntt.h"
#include "fa_pack.h"

#define PADX(p) ((p) + ((p) >> 4))

// radix-16 = four levels from the block Q at level m0 (fft_small.h block
// convention): t[0..2] = (w2[Q], w2[2Q], w2[2Q+1]), t[3+3j..5+3j] the same
// for Q' = 4Q + j
FA_DEV void fa_fwd16(double* x, const double* t, double n, double ninv)
{
#pragma unroll
    for (int k = 0; k < 4; k++) fa_fwd4(x[k], x[k + 4], x[k + 8], x[k + 12], t[0], t[1], t[2], n, ninv);
#pragma unroll
    for (int j = 0; j < 4; j++)
        fa_fwd4(x[4 * j], x[4 * j + 1], x[4 * j + 2], x[4 * j + 3], t[3 + 3 * j], t[4 + 3 * j],
                t[5 + 3 * j], n, ninv);
}

// fa_fwd16 for inputs whose upper half x[8..15] is zero: the first radix-4
// layer needs only the products by x[k+4]
FA_DEV void fa_fwd16_hz(double* x, const double* t, double n, double ninv)
{
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double a0 = fa_reduce_pm1n(x[k], n, ninv);
        const double v1 = fa_mulmod(x[k + 4], t[1], n, ninv);
        const double v3 = fa_mulmod(x[k + 4], t[2], n, ninv);
        x[k] = a0 + v1;
        x[k + 4] = a0 - v1;
        x[k + 8] = a0 + v3;
        x[k + 12] = a0 - v3;
    }
#pragma unroll
    for (int j = 0; j < 4; j++)
        fa_fwd4(x[4 * j], x[4 * j + 1], x[4 * j + 2], x[4 * j + 3], t[3 + 3 * j], t[4 + 3 * j],
                t[5 + 3 * j], n, ninv);
}

FA_DEV void fa_inv16(double* x, const double* t, double n, double ninv)
{
#pragma unroll
    for (int j = 0; j < 4; j++)
        fa_inv4(x[4 * j], x[4 * j + 1], x[4 * j + 2], x[4 * j + 3], t[3 + 3 * j], t[4 + 3 * j],
                t[5 + 3 * j], n, ninv);
#pragma unroll
    for (int k = 0; k < 4; k++) fa_inv4(x[k], x[k + 4], x[k + 8], x[k + 12], t[0], t[1], t[2], n, ninv);
}

// twiddles of block Q from a table indexed by global block number
FA_DEV void fa_tw16(const double* __restrict__ w, uint64_t Q, double* t)
{
    t[0] = w[Q];
    t[1] = w[2 * Q];
    t[2] = w[2 * Q + 1];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        uint64_t Qp = 4 * Q + j;
        t[3 + 3 * j] = w[Qp];
        t[4 + 3 * j] = w[2 * Qp];
        t[5 + 3 * j] = w[2 * Qp + 1];
    }
}

// ---------------------------------------------------------------------------
// row twiddles: for row r (root offset r at level 12) the entry of local
// level m, block q is T(m, q) = w2[(r << m) + q] = rho_m * w2[q] with
// rho_m = w2[r << m] (the index bits of r << m and q are disjoint), reduced
// exactly to [-n/2, n/2].  Levels 0..7 (255 entries, shared by the first
// two stages) live in LDS at (2^m - 1) + q; the 15 entries of the last
// stage (levels 8..11 of block t) are used by thread t alone and are
// computed in registers.
// ---------------------------------------------------------------------------

// the mulmod lands in (-n, n); one rounding step brings it to [-n/2, n/2]
// (up to the last bit), the range the butterflies' bounds assume of a twiddle
FA_DEV double fa_row_entry(double rho, const double* __restrict__ w, int q, double n,
                           double ninv)
{
    return fa_reduce_pm1n(fa_mulmod(rho, w[q], n, ninv), n, ninv);
}

// rho[m] for m < 12 and the LDS entries of levels 0..7
FA_DEV void fa_build_row_table(double* tab, const double* __restrict__ w, uint64_t r,
                               double* rho, double n, double ninv)
{
    const int tid = threadIdx.x;
    if (tid < 12) rho[tid] = w[r << tid];
    __syncthreads();
    if (tid < 255) {
        const int m = 31 - __builtin_clz(tid + 1);
        tab[tid] = fa_row_entry(rho[m], w, tid + 1 - (1 << m), n, ninv);
    }
    __syncthreads();
}

// twiddles of stage j < 2 (levels 4j .. 4j+3) of block bq from the table
FA_DEV void fa_row_tw(const double* tab, int j, int bq, double* t)
{
    const int m = 4 * j;
    t[0] = tab[(1 << m) - 1 + bq];
    t[1] = tab[(2 << m) - 1 + 2 * bq];
    t[2] = tab[(2 << m) - 1 + 2 * bq + 1];
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
        const int Qp = 4 * bq + jj;
        t[3 + 3 * jj] = tab[(4 << m) - 1 + Qp];
        t[4 + 3 * jj] = tab[(8 << m) - 1 + 2 * Qp];
        t[5 + 3 * jj] = tab[(8 << m) - 1 + 2 * Qp + 1];
    }
}

// twiddles of the last stage (levels 8..11) of block t, in registers
FA_DEV void fa_row_tw_last(const double* rho, const double* __restrict__ w, int t, double* tw,
                           double n, double ninv)
{
    tw[0] = fa_row_entry(rho[8], w, t, n, ninv);
    tw[1] = fa_row_entry(rho[9], w, 2 * t, n, ninv);
    tw[2] = fa_row_entry(rho[9], w, 2 * t + 1, n, ninv);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int Qp = 4 * t + j;
        tw[3 + 3 * j] = fa_row_entry(rho[10], w, Qp, n, ninv);
        tw[4 + 3 * j] = fa_row_entry(rho[11], w, 2 * Qp, n, ninv);
        tw[5 + 3 * j] = fa_row_entry(rho[11], w, 2 * Qp + 1, n, ninv);
    }
}

// radix-16 stage j < 2 of block bq with each radix-4 layer's twiddles read
// from the LDS table just before the layer
FA_DEV void fa_fwd16_rowtab(double* x, const double* tab, int j, int bq, double n, double ninv)
{
    const int m = 4 * j;
    {
        const double T0 = tab[(1 << m) - 1 + bq];
        const double T1 = tab[(2 << m) - 1 + 2 * bq];
        const double T2 = tab[(2 << m) - 1 + 2 * bq + 1];
#pragma unroll
        for (int k = 0; k < 4; k++) fa_fwd4(x[k], x[k + 4], x[k + 8], x[k + 12], T0, T1, T2, n, ninv);
    }
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
        const int Qp = 4 * bq + jj;
        const double T0 = tab[(4 << m) - 1 + Qp];
        const double T1 = tab[(8 << m) - 1 + 2 * Qp];
        const double T2 = tab[(8 << m) - 1 + 2 * Qp + 1];
        fa_fwd4(x[4 * jj], x[4 * jj + 1], x[4 * jj + 2], x[4 * jj + 3], T0, T1, T2, n, ninv);
    }
}

FA_DEV void fa_inv16_rowtab(double* x, const double* tab, int j, int bq, double n, double ninv)
{
    const int m = 4 * j;
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
        const int Qp = 4 * bq + jj;
        const double T0 = tab[(4 << m) - 1 + Qp];
        const double T1 = tab[(8 << m) - 1 + 2 * Qp];
        const double T2 = tab[(8 << m) - 1 + 2 * Qp + 1];
        fa_inv4(x[4 * jj], x[4 * jj + 1], x[4 * jj + 2], x[4 * jj + 3], T0, T1, T2, n, ninv);
    }
    const double T0 = tab[(1 << m) - 1 + bq];
    const double T1 = tab[(2 << m) - 1 + 2 * bq];
    const double T2 = tab[(2 << m) - 1 + 2 * bq + 1];
#pragma unroll
    for (int k = 0; k < 4; k++) fa_inv4(x[k], x[k + 4], x[k + 8], x[k + 12], T0, T1, T2, n, ninv);
}

// the last radix-16 stage (levels 8..11 of block t) with its twiddles formed
// just before each radix-4 layer: fewer registers live across the row
// transforms than with all 15 held (three workgroups share a CU)
FA_DEV void fa_fwd16_rowlast(double* x, const double* rho, const double* __restrict__ w, int t,
                             double n, double ninv)
{
    // the table loads depend on this opaque copy of t, so they are issued
    // here rather than hoisted to the kernel's start (where they would hold
    // registers through both forward rows)
    asm volatile("" : "+v"(t));
    {
        const double T0 = fa_row_entry(rho[8], w, t, n, ninv);
        const double T1 = fa_row_entry(rho[9], w, 2 * t, n, ninv);
        const double T2 = fa_row_entry(rho[9], w, 2 * t + 1, n, ninv);
#pragma unroll
        for (int k = 0; k < 4; k++) fa_fwd4(x[k], x[k + 4], x[k + 8], x[k + 12], T0, T1, T2, n, ninv);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int Qp = 4 * t + j;
        const double T0 = fa_row_entry(rho[10], w, Qp, n, ninv);
        const double T1 = fa_row_entry(rho[11], w, 2 * Qp, n, ninv);
        const double T2 = fa_row_entry(rho[11], w, 2 * Qp + 1, n, ninv);
        fa_fwd4(x[4 * j], x[4 * j + 1], x[4 * j + 2], x[4 * j + 3], T0, T1, T2, n, ninv);
    }
}

FA_DEV void fa_inv16_rowlast(double* x, const double* rho, const double* __restrict__ w, int t,
                             double n, double ninv)
{
    asm volatile("" : "+v"(t));
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int Qp = 4 * t + j;
        const double T0 = fa_row_entry(rho[10], w, Qp, n, ninv);
        const double T1 = fa_row_entry(rho[11], w, 2 * Qp, n, ninv);
        const double T2 = fa_row_entry(rho[11], w, 2 * Qp + 1, n, ninv);
        fa_inv4(x[4 * j], x[4 * j + 1], x[4 * j + 2], x[4 * j + 3], T0, T1, T2, n, ninv);
    }
    const double T0 = fa_row_entry(rho[8], w, t, n, ninv);
    const double T1 = fa_row_entry(rho[9], w, 2 * t, n, ninv);
    const double T2 = fa_row_entry(rho[9], w, 2 * t + 1, n, ninv);
#pragma unroll
    for (int k = 0; k < 4; k++) fa_inv4(x[k], x[k + 4], x[k + 8], x[k + 12], T0, T1, T2, n, ninv);
}

// 4096-point forward transform of one row, 256 threads, input in the
// stage-0 layout (positions t + 256k), output in the stage-2 layout
// (positions 16t + k); twl = the thread's last-stage twiddles
FA_DEV void fa_fwd4096_row(double* x, double* xch, const double* tab, const double* twl,
                           double n, double ninv)
{
    const int t = threadIdx.x;
    double tw[15];
    fa_row_tw(tab, 0, 0, tw);
    fa_fwd16(x, tw, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) xch[PADX(t + 256 * k)] = x[k];
    __syncthreads();
    const int b = t >> 4, i0 = t & 15;
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = xch[PADX(256 * b + i0 + 16 * k)];
    fa_row_tw(tab, 1, b, tw);
    fa_fwd16(x, tw, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) xch[PADX(256 * b + i0 + 16 * k)] = x[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = xch[PADX(16 * t + k)];
    fa_fwd16(x, twl, n, ninv);
    __syncthreads();
}

FA_DEV void fa_inv4096_row(double* x, double* xch, const double* tab, const double* twl,
                           double n, double ninv)
{
    const int t = threadIdx.x;
    double tw[15];
    fa_inv16(x, twl, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) xch[PADX(16 * t + k)] = x[k];
    __syncthreads();
    const int b = t >> 4, i0 = t & 15;
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = xch[PADX(256 * b + i0 + 16 * k)];
    fa_row_tw(tab, 1, b, tw);
    fa_inv16(x, tw, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) xch[PADX(256 * b + i0 + 16 * k)] = x[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = xch[PADX(t + 256 * k)];
    fa_row_tw(tab, 0, 0, tw);
    fa_inv16(x, tw, n, ninv);
}

// row pass: both forward row transforms, pointwise product with the
// normalizer m, inverse row transform; in place in arow.  Row r of the
// two-level transform (root offset r at level 12).
FA_DEV void fa_row_conv_tile(double* __restrict__ arow, const double* __restrict__ brow,
                             uint64_t r, double n, double ninv, double m,
                             const double* __restrict__ w2, const double* __restrict__ w2inv,
                             double* xch, double* tab, double* rho)
{
    const int t = threadIdx.x;
    double x[16], y[16], twl[15];
#pragma unroll
    for (int k = 0; k < 16; k++) y[k] = brow[t + 256 * k];
    fa_build_row_table(tab, w2, r, rho, n, ninv);
    fa_row_tw_last(rho, w2, t, twl, n, ninv);
    fa_fwd4096_row(y, xch, tab, twl, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = arow[t + 256 * k];
    fa_fwd4096_row(x, xch, tab, twl, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = fa_mulmod(fa_mulmod(x[k], m, n, ninv), y[k], n, ninv);
    fa_build_row_table(tab, w2inv, r, rho, n, ninv);
    fa_row_tw_last(rho, w2inv, t, twl, n, ninv);
    fa_inv4096_row(x, xch, tab, twl, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) arow[t + 256 * k] = x[k];
}

// fa_row_conv_tile with the last-stage twiddles formed on the fly.
// SQR: a square (brow unused): one forward transform, pointwise x^2
// (mpn_mul.c:1425-1452, sd_fft_ctx_point_sqr)
template <bool SQR = false>
FA_DEV void fa_row_conv_tile_lr(double* __restrict__ arow, const double* __restrict__ brow,
                                uint64_t r, double n, double ninv, double m,
                                const double* __restrict__ w2, const double* __restrict__ w2inv,
                                double* xch, double* tab, double* rho)
{
    const int t = threadIdx.x;
    const int b = t >> 4, i0 = t & 15;
    double x[16], y[16];
    fa_build_row_table(tab, w2, r, rho, n, ninv);
#pragma unroll
    for (int z = SQR ? 1 : 0; z < 2; z++) {
        double* v = z ? x : y;
        const double* src = z ? arow : brow;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = src[t + 256 * k];
        fa_fwd16_rowtab(v, tab, 0, 0, n, ninv);
#pragma unroll
        for (int k = 0; k < 16; k++) xch[PADX(t + 256 * k)] = v[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = xch[PADX(256 * b + i0 + 16 * k)];
        fa_fwd16_rowtab(v, tab, 1, b, n, ninv);
#pragma unroll
        for (int k = 0; k < 16; k++) xch[PADX(256 * b + i0 + 16 * k)] = v[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = xch[PADX(16 * t + k)];
        __builtin_amdgcn_sched_barrier(0);
        fa_fwd16_rowlast(v, rho, w2, t, n, ninv);
        // the finished row is materialized here, not next to its use
#pragma unroll
        for (int k = 0; k < 16; k++) asm volatile("" : "+v"(v[k]));
        __builtin_amdgcn_sched_barrier(0);
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 16; k++)
        x[k] = fa_mulmod(fa_mulmod(x[k], m, n, ninv), SQR ? x[k] : y[k], n, ninv);
    fa_build_row_table(tab, w2inv, r, rho, n, ninv);
    __builtin_amdgcn_sched_barrier(0);
    fa_inv16_rowlast(x, rho, w2inv, t, n, ninv);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < 16; k++) xch[PADX(16 * t + k)] = x[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = xch[PADX(256 * b + i0 + 16 * k)];
    fa_inv16_rowtab(x, tab, 1, b, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) xch[PADX(256 * b + i0 + 16 * k)] = x[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = xch[PADX(t + 256 * k)];
    fa_inv16_rowtab(x, tab, 0, 0, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) arow[t + 256 * k] = x[k];
}

template <bool SQR>
__global__ void __launch_bounds__(256, 3) fa_k_row_conv16lr(double* __restrict__ A,
                                                             const double* __restrict__ B,
                                                             uint64_t stride, fa_prime_set S)
{
    __shared__ double xch[4352];
    __shared__ double tab[256];
    __shared__ double rho[16];
    const int i = blockIdx.y;
    const uint64_t r = blockIdx.x;
    fa_row_conv_tile_lr<SQR>(A + (uint64_t)i * stride + r * FA_ROW_PITCH,
                        B + (uint64_t)i * stride + r * FA_ROW_PITCH,
                        r, S.p[i], S.pinv[i], S.norm[i], S.w2[i], S.w2inv[i], xch, tab, rho);
}

// Row passes of the separate forward and inverse transforms (the
// transformed-operand interface): the last 12 forward levels of one row,
// stored back in place with point 16t + k of the transformed row at
// position 16t + k; and its inverse from that layout.
__global__ void __launch_bounds__(256) fa_k_row_fwd16(double* __restrict__ A, uint64_t stride,
                                                      fa_prime_set S)
{
    __shared__ double xch[4352];
    __shared__ double tab[256];
    __shared__ double rho[16];
    const int i = blockIdx.y;
    const uint64_t r = blockIdx.x;
    const double n = S.p[i], ninv = S.pinv[i];
    double* row = A + (uint64_t)i * stride + r * FA_ROW_PITCH;
    const int t = threadIdx.x;
    const int b = t >> 4, i0 = t & 15;
    double v[16];
    fa_build_row_table(tab, S.w2[i], r, rho, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = row[t + 256 * k];
    fa_fwd16_rowtab(v, tab, 0, 0, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) xch[PADX(t + 256 * k)] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = xch[PADX(256 * b + i0 + 16 * k)];
    fa_fwd16_rowtab(v, tab, 1, b, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) xch[PADX(256 * b + i0 + 16 * k)] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = xch[PADX(16 * t + k)];
    fa_fwd16_rowlast(v, rho, S.w2[i], t, n, ninv);
    __syncthreads();
    // through LDS so that the stores are coalesced
#pragma unroll
    for (int k = 0; k < 16; k++) xch[PADX(16 * t + k)] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) row[t + 256 * k] = xch[PADX(t + 256 * k)];
}

__global__ void __launch_bounds__(256) fa_k_row_inv16(double* __restrict__ A, uint64_t stride,
                                                      fa_prime_set S)
{
    __shared__ double xch[4352];
    __shared__ double tab[256];
    __shared__ double rho[16];
    const int i = blockIdx.y;
    const uint64_t r = blockIdx.x;
    const double n = S.p[i], ninv = S.pinv[i];
    double* row = A + (uint64_t)i * stride + r * FA_ROW_PITCH;
    const int t = threadIdx.x;
    const int b = t >> 4, i0 = t & 15;
    double x[16];
    fa_build_row_table(tab, S.w2inv[i], r, rho, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) xch[PADX(t + 256 * k)] = row[t + 256 * k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = xch[PADX(16 * t + k)];
    fa_inv16_rowlast(x, rho, S.w2inv[i], t, n, ninv);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) xch[PADX(16 * t + k)] = x[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = xch[PADX(256 * b + i0 + 16 * k)];
    fa_inv16_rowtab(x, tab, 1, b, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) xch[PADX(256 * b + i0 + 16 * k)] = x[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = xch[PADX(t + 256 * k)];
    fa_inv16_rowtab(x, tab, 0, 0, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) row[t + 256 * k] = x[k];
}

// blockIdx.x = row, blockIdx.y = prime
__global__ void __launch_bounds__(256) fa_k_row_conv16(double* __restrict__ A,
                                                          const double* __restrict__ B,
                                                          uint64_t stride, fa_prime_set S)
{
    __shared__ double xch[4352];
    __shared__ double tab[256];
    __shared__ double rho[16];
    const int i = blockIdx.y;
    const uint64_t r = blockIdx.x;
    fa_row_conv_tile(A + (uint64_t)i * stride + r * FA_ROW_PITCH,
                     B + (uint64_t)i * stride + r * FA_ROW_PITCH, r,
                     S.p[i], S.pinv[i], S.norm[i], S.w2[i], S.w2inv[i], xch, tab, rho);
}

// ---------------------------------------------------------------------------
// column passes for any column depth D1 in [1, 12].  With D1 = 4s + r:
//  - D1 >= 4: 2^(D1-4) threads per column, 16 elements each; a partial stage
//    of r levels (if r > 0) on the stage-0 layout, then s radix-16 stages
//    starting at levels r, r+4, ...;
//  - D1 < 4: one thread per column, all 2^D1 elements in registers.
// A stage starting at level m0 owns, in block b of level m0, the positions
//     (b << (D1 - m0)) + i0 + (k << (D1 - m0 - 4)),  tc = b * 2^(D1-m0-4) + i0.
// ---------------------------------------------------------------------------

template <int D1>
struct col_geom {
    static constexpr int R = (D1 >= 4) ? (D1 % 4) : D1;   // levels of the partial stage
    static constexpr int NS = (D1 >= 4) ? (D1 / 4) : 0;   // radix-16 stages
    static constexpr int TPC = (D1 >= 4) ? (1 << (D1 - 4)) : 1;
    static constexpr int CPW = 1024 / TPC;
    static constexpr int LEN = 1 << D1;
    static constexpr int PLEN = LEN + LEN / 16;
    static constexpr int NEL = (D1 >= 4) ? 16 : (1 << D1);  // elements per thread
    static constexpr int NSTAGE = NS + (R > 0 && D1 >= 4 ? 1 : 0);
    static constexpr int LDS = (NSTAGE > 1) ? CPW * PLEN : 1;

    // first level of stage j (stage 0 is the partial one when R > 0)
    static FA_DEV int m0(int j) { return (R > 0) ? (j == 0 ? 0 : R + 4 * (j - 1)) : 4 * j; }
    static FA_DEV int pos(int j, int tc, int k)
    {
        if (D1 < 4) return k;
        const int sh = D1 - m0(j) - 4;
        const int b = tc >> sh, i0 = tc & ((1 << sh) - 1);
        return (b << (D1 - m0(j))) + i0 + (k << sh);
    }
    static FA_DEV int block(int j, int tc)
    {
        if (D1 < 4) return 0;
        return tc >> (D1 - m0(j) - 4);
    }
};

// the first r (< 4) levels of a radix-16 group, block Q: level 0 pairs
// (k, k+8), level 1 (k, k+4), level 2 (k, k+2)
FA_DEV void fa_fwd16_partial(double* x, int r, const double* __restrict__ w, uint64_t Q,
                             double n, double ninv)
{
    if (r == 1) {
        const double W = w[Q];
#pragma unroll
        for (int k = 0; k < 8; k++) fa_fwd2(x[k], x[k + 8], W, n, ninv);
        return;
    }
    const double W2 = w[Q], W = w[2 * Q], IW = w[2 * Q + 1];
#pragma unroll
    for (int k = 0; k < 4; k++) fa_fwd4(x[k], x[k + 4], x[k + 8], x[k + 12], W2, W, IW, n, ninv);
    if (r == 3) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double Wj = w[4 * Q + j];
            fa_fwd2(x[4 * j], x[4 * j + 2], Wj, n, ninv);
            fa_fwd2(x[4 * j + 1], x[4 * j + 3], Wj, n, ninv);
        }
    }
}

FA_DEV void fa_inv16_partial(double* x, int r, const double* __restrict__ wi, uint64_t Q,
                             double n, double ninv)
{
    if (r == 1) {
        const double V = wi[Q];
#pragma unroll
        for (int k = 0; k < 8; k++) fa_inv2(x[k], x[k + 8], V, n, ninv);
        return;
    }
    if (r == 3) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double Vj = wi[4 * Q + j];
            fa_inv2(x[4 * j], x[4 * j + 2], Vj, n, ninv);
            fa_inv2(x[4 * j + 1], x[4 * j + 3], Vj, n, ninv);
        }
    }
    const double V2 = wi[Q], V = wi[2 * Q], IV = wi[2 * Q + 1];
#pragma unroll
    for (int k = 0; k < 4; k++) fa_inv4(x[k], x[k + 4], x[k + 8], x[k + 12], V2, V, IV, n, ninv);
}

// a whole column of 2^D1 <= 8 points in one thread (block 0 at level 0)
template <int D1>
FA_DEV void fa_fwd_small_col(double* x, const double* __restrict__ w, double n, double ninv)
{
#pragma unroll
    for (int m = 0; m < D1; m++) {
        const int half = 1 << (D1 - m - 1);
#pragma unroll
        for (int q = 0; q < (1 << m); q++) {
            const double W = w[q];
#pragma unroll
            for (int i = 0; i < half; i++) fa_fwd2(x[2 * half * q + i], x[2 * half * q + i + half], W, n, ninv);
        }
    }
}

template <int D1>
FA_DEV void fa_inv_small_col(double* x, const double* __restrict__ wi, double n, double ninv)
{
#pragma unroll
    for (int m = D1 - 1; m >= 0; m--) {
        const int half = 1 << (D1 - m - 1);
#pragma unroll
        for (int q = 0; q < (1 << m); q++) {
            const double V = wi[q];
#pragma unroll
            for (int i = 0; i < half; i++) fa_inv2(x[2 * half * q + i], x[2 * half * q + i + half], V, n, ninv);
        }
    }
}

// linear work index with consecutive indices on the same XCD (workgroups
// are dealt round-robin to the 8 XCDs); grids are multiples of 8
FA_DEV unsigned fa_xcd_linear()
{
    unsigned nb = gridDim.x, bid = blockIdx.x;
    return (bid & 7) * (nb >> 3) + (bid >> 3);
}

// first D1 levels of column group g (CPW columns) of one residue array:
// import, forward, store.  lds holds CPW * PLEN doubles.
template <int D1, int KIND>
FA_DEV void fa_col_fwd_tile(const fa_operand& op, double* __restrict__ dst, int d2, int g,
                            double n, double ninv, const double* __restrict__ w2,
                            const double* __restrict__ pow2, double* lds, int trows)
{
    typedef col_geom<D1> G;
    const int c = threadIdx.x % G::CPW, tc = threadIdx.x / G::CPW;
    const uint64_t col = (uint64_t)g * G::CPW + c;
    double* my = lds + c;   // column c of the interleaved LDS rows
    double x[16], tw[15];
#pragma unroll
    for (int k = 0; k < G::NEL; k++)
        x[k] = fa_operand_value_k<KIND>(op, ((uint64_t)G::pos(0, tc, k) << d2) + col, pow2, n, ninv);
    if (D1 < 4) {
        fa_fwd_small_col<(D1 < 4 ? D1 : 1)>(x, w2, n, ninv);
    } else {
#pragma unroll
        for (int j = 0; j < G::NSTAGE; j++) {
            if (j > 0) {
                // each position of a stage layout belongs to one thread: a
                // single barrier separates stage j-1's stores from stage j's loads
#pragma unroll
                for (int k = 0; k < 16; k++) my[PADX(G::pos(j - 1, tc, k)) * G::CPW] = x[k];
                __syncthreads();
#pragma unroll
                for (int k = 0; k < 16; k++) x[k] = my[PADX(G::pos(j, tc, k)) * G::CPW];
            }
            if (G::R > 0 && j == 0) {
                fa_fwd16_partial(x, G::R, w2, (uint64_t)G::block(0, tc), n, ninv);
            } else {
                fa_tw16(w2, (uint64_t)G::block(j, tc), tw);
                fa_fwd16(x, tw, n, ninv);
            }
        }
    }
    const int jl = (D1 < 4) ? 0 : G::NSTAGE - 1;
    // rows at or past trows are not read by a truncated product
#pragma unroll
    for (int k = 0; k < G::NEL; k++) {
        const int row = G::pos(jl, tc, k);
        if (row < trows) dst[(uint64_t)row * FA_ROW_PITCH + col] = x[k];
    }
}

// ---------------------------------------------------------------------------
// Truncated inverse column transform (van der Hoeven's inverse TFT in the
// block-twisted form; sd_ifft.c:1320-1493 is the reference's version).  A
// column of 2^D1 points holds, at rows [0, t), the evaluations computed by
// the row pass; its coefficients at rows >= t are zero (the product is
// shorter than t rows).  Block Q of size 2L splits X^2L - w into
// (X^L - r)(X^L + r), r = w2[Q], children 2Q (a = lo + r hi) and 2Q + 1
// (b = lo - r hi).  Values are carried at the scale of their block (a block
// of size M holds M times its coefficients, the scale of the full inverse):
//  - t >= L: a is known; invert it (scale L), turn the known tail of hi into
//    the tail of b, v_j = u_j - r x_hi[j] for j >= t - L, solve b's block
//    with t - L evaluations, then (u + v, (u - v) r^-1) gives lo and hi at
//    scale 2L;
//  - t < L: b is unknown, lo's tail and all of hi are known; a's tail at
//    scale L is (x_lo + r x_hi) / 2; solve a's block with t evaluations, then
//    lo = 2 u - r x_hi.
// The descent is one chain of blocks (at most D1 of them); the way back
// applies each block's combination.  The tile sits in LDS as element e of
// column c at s[(e << LG) + c]; every step runs over all CPW columns.
// ---------------------------------------------------------------------------
template <int D1>
struct trunc_geom {
    static constexpr int CPW = (D1 >= 4) ? (1024 >> (D1 - 4)) : 1024;
    static constexpr int LG = (D1 >= 4) ? (14 - D1) : 10;   // log2 CPW
    static constexpr int LDS = CPW << D1;
};

FA_DEV void fa_trunc_step(double* s, int LG, int A, int Bo, int j0, int j1, int mode, double r,
                          double rinv, double n, double ninv)
{
    // mode 0: x_A[j] = (x_A[j] + r x_B[j]) / 2       (descent, t < L)
    // mode 1: x_B[j] = x_A[j] - r x_B[j]             (descent, t >= L)
    // mode 2: (x_A, x_B) = (x_A + x_B, (x_A - x_B) r^-1)   (way back, t >= L)
    // mode 3: x_A[j] = 2 x_A[j] - r x_B[j]           (way back, t < L)
    const int C = 1 << LG;
    const int total = (j1 - j0) << LG;
    const double half = 0.5 * (n + 1.0);
    for (int u = threadIdx.x; u < total; u += blockDim.x) {
        const int j = j0 + (u >> LG), c = u & (C - 1);
        double* pa = s + ((A + j) << LG) + c;
        double* pb = s + ((Bo + j) << LG) + c;
        const double xa = *pa, xb = *pb;
        if (mode == 0) {
            *pa = fa_reduce_pm1n(fa_mulmod(xa + fa_mulmod(xb, r, n, ninv), half, n, ninv), n, ninv);
        } else if (mode == 1) {
            *pb = fa_reduce_pm1n(xa - fa_mulmod(xb, r, n, ninv), n, ninv);
        } else if (mode == 2) {
            *pa = fa_reduce_pm1n(xa + xb, n, ninv);
            *pb = fa_reduce_pm1n(fa_mulmod(xa - xb, rinv, n, ninv), n, ninv);
        } else {
            *pa = fa_reduce_pm1n(2.0 * xa - fa_mulmod(xb, r, n, ninv), n, ninv);
        }
    }
    __syncthreads();
}

// evaluations at rows [0, t) of the tile's columns -> the coefficients of
// rows [0, t) (times 2^D1); rows >= t of s must hold zeros on entry
template <int D1>
FA_DEV void fa_trunc_inverse(double* s, int t, double n, double ninv,
                             const double* __restrict__ w2, const double* __restrict__ wi)
{
    constexpr int LG = trunc_geom<D1>::LG;
    int o = 0, k = D1, tt = t;
    uint64_t Q = 0;
    int po[D1 + 1], pk[D1 + 1], pt[D1 + 1];
    uint64_t pq[D1 + 1];
    int depth = 0;
    while (k > 0 && tt > 0 && tt < (1 << k)) {
        const int L = 1 << (k - 1);
        const double r = w2[Q];
        po[depth] = o; pk[depth] = k; pt[depth] = tt; pq[depth] = Q;
        depth++;
        if (tt >= L) {
            fa_lds_inv(s + ((size_t)o << LG), k - 1, LG, 2 * Q, wi, n, ninv);
            fa_trunc_step(s, LG, o, o + L, tt - L, L, 1, r, 0.0, n, ninv);
            o += L;
            tt -= L;
            Q = 2 * Q + 1;
        } else {
            fa_trunc_step(s, LG, o, o + L, tt, L, 0, r, 0.0, n, ninv);
            Q = 2 * Q;
        }
        k--;
    }
    if (k > 0 && tt == (1 << k)) fa_lds_inv(s + ((size_t)o << LG), k, LG, Q, wi, n, ninv);
    while (depth-- > 0) {
        const int L = 1 << (pk[depth] - 1);
        const uint64_t Qd = pq[depth];
        fa_trunc_step(s, LG, po[depth], po[depth] + L, 0, L, pt[depth] >= L ? 2 : 3, w2[Qd],
                      wi[Qd], n, ninv);
    }
}

template <int D1>
FA_DEV void fa_col_inv_trunc_tile(double* __restrict__ dst, int g, int t, double n, double ninv,
                                  const double* __restrict__ w2, const double* __restrict__ wi,
                                  double* lds)
{
    typedef trunc_geom<D1> G;
    const uint64_t col0 = (uint64_t)g * G::CPW;
    for (int u = threadIdx.x; u < (G::CPW << D1); u += blockDim.x) {
        const int e = u >> G::LG, c = u & (G::CPW - 1);
        lds[u] = e < t ? dst[(uint64_t)e * FA_ROW_PITCH + col0 + c] : 0.0;
    }
    __syncthreads();
    fa_trunc_inverse<D1>(lds, t, n, ninv, w2, wi);
    for (int u = threadIdx.x; u < (t << G::LG); u += blockDim.x) {
        const int e = u >> G::LG, c = u & (G::CPW - 1);
        dst[(uint64_t)e * FA_ROW_PITCH + col0 + c] = lds[u];
    }
}

template <int D1>
FA_DEV void fa_col_inv_tile(double* __restrict__ dst, int d2, int g, double n, double ninv,
                            const double* __restrict__ wi, double* lds)
{
    typedef col_geom<D1> G;
    const int c = threadIdx.x % G::CPW, tc = threadIdx.x / G::CPW;
    const uint64_t col = (uint64_t)g * G::CPW + c;
    double* my = lds + c;   // column c of the interleaved LDS rows
    double x[16], tw[15];
    const int jl = (D1 < 4) ? 0 : G::NSTAGE - 1;
#pragma unroll
    for (int k = 0; k < G::NEL; k++) x[k] = dst[(uint64_t)G::pos(jl, tc, k) * FA_ROW_PITCH + col];
    if (D1 < 4) {
        fa_inv_small_col<(D1 < 4 ? D1 : 1)>(x, wi, n, ninv);
    } else {
#pragma unroll
        for (int j = G::NSTAGE - 1; j >= 0; j--) {
            if (j < G::NSTAGE - 1) {
#pragma unroll
                for (int k = 0; k < 16; k++) my[PADX(G::pos(j + 1, tc, k)) * G::CPW] = x[k];
                __syncthreads();
#pragma unroll
                for (int k = 0; k < 16; k++) x[k] = my[PADX(G::pos(j, tc, k)) * G::CPW];
            }
            if (G::R > 0 && j == 0) {
                fa_inv16_partial(x, G::R, wi, (uint64_t)G::block(0, tc), n, ninv);
            } else {
                fa_tw16(wi, (uint64_t)G::block(j, tc), tw);
                fa_inv16(x, tw, n, ninv);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < G::NEL; k++) dst[(uint64_t)G::pos(0, tc, k) * FA_ROW_PITCH + col] = x[k];
}

// work index L -> (prime i, column group g, operand z) with the two column
// groups of one 128-byte line (g = 2h, 2h + 1) adjacent, then the primes,
// then the line pairs: concurrent workgroups of an XCD share every line they
// read or write, and the primes of one group follow each other (the limb
// reads of the import are shared too).  ngroups is even.
FA_DEV void fa_col_split(unsigned L, int np, int ngroups, int* i, int* g, int* z)
{
    const int lo = L & 1;
    const unsigned L2 = L >> 1;
    *i = (int)(L2 % (unsigned)np);
    const unsigned rest = L2 / (unsigned)np;
    const int half = ngroups >> 1;
    *g = 2 * (int)(rest % (unsigned)half) + lo;
    *z = (int)(rest / (unsigned)half);
}

template <int D1, int KIND>
__global__ void __launch_bounds__(1024) fa_k_col_fwd16(fa_operand a, fa_operand b, int d2,
                                                       double* __restrict__ A,
                                                       double* __restrict__ B, uint64_t stride,
                                                       fa_prime_set S, int np, int ngroups,
                                                       int nops, int trows)
{
    __shared__ double lds[col_geom<D1>::LDS];
    int i, g, z;
    fa_col_split(fa_xcd_linear(), np, ngroups, &i, &g, &z);
    if (z >= nops) return;
    fa_col_fwd_tile<D1, KIND>(z ? b : a, (z ? B : A) + (uint64_t)i * stride, d2, g, S.p[i],
                              S.pinv[i], S.w2[i], S.pow2[i], lds, trows);
}

template <int D1>
constexpr int fa_col_inv_lds()
{
    return col_geom<D1>::LDS > trunc_geom<D1>::LDS ? col_geom<D1>::LDS : trunc_geom<D1>::LDS;
}

// trows < 2^D1: the truncated inverse (the tile groups of both kernels
// cover the same 4096 columns: CPW columns per group)
template <int D1>
__global__ void __launch_bounds__(1024) fa_k_col_inv16(double* __restrict__ A, int d2,
                                                       uint64_t stride, fa_prime_set S, int np,
                                                       int ngroups, int trows)
{
    __shared__ double lds[fa_col_inv_lds<D1>()];
    int i, g, z;
    fa_col_split(fa_xcd_linear(), np, ngroups, &i, &g, &z);
    if (z >= 1) return;
    if (trows < (1 << D1))
        fa_col_inv_trunc_tile<D1>(A + (uint64_t)i * stride, g, trows, S.p[i], S.pinv[i], S.w2[i],
                                  S.w2inv[i], lds);
    else
        fa_col_inv_tile<D1>(A + (uint64_t)i * stride, d2, g, S.p[i], S.pinv[i], S.w2inv[i], lds);
}

// ---------------------------------------------------------------------------
// C-column tiles for D1 = 12 (C = 8: 1024 threads, C = 4: 512 threads).
// Lane l of a workgroup takes column c = l mod C and u = l / C < 128, and at
// every stage holds two radix-16 groups, tc = u and tc = u + 128 (x0, x1),
// so one wave instruction covers 64/C rows x 8C contiguous bytes.  A stage
// exchange runs in two halves of the column (rows [0, 2048) and
// [2048, 4096)), each C columns x 2048 rows in LDS:
//  - stage 0 -> 1: a stage-0 group (rows tc + 256k) has k < 8 in the first
//    half and k >= 8 in the second; the stage-1 group read back in half H is
//    tc = 128H + u (blocks 8H .. 8H+7 of level 4);
//  - stage 1 -> 2: stage-1 group u + 128H and stage-2 group u + 128H both
//    lie in half H.
// LDS rows interleave the C columns (element (c, r) at C PADX(r) + c), so the
// lanes of a wave touch consecutive doubles when their rows are consecutive,
// and the pad of one row per 16 separates the 16-row strides of the last
// stage: a half-wave of 64-bit accesses then covers 64 distinct banks for
// C = 8 (4 rows x 8 columns) and C = 4 (8 rows x 4 columns).  The C = 4 tile
// needs 69.6 KB of LDS, so two workgroups share a CU and one's loads and
// stores overlap the other's arithmetic.
// ---------------------------------------------------------------------------
#define FA_CW_LDS(C) ((C) * 2176)
#define FA_C8_LDS FA_CW_LDS(8)

template <int C>
FA_DEV int fa_cw_addr(int c, int row) { return C * PADX(row & 2047) + c; }

// stage-1 row of group tc, element k
FA_DEV int fa_c8_row1(int tc, int k) { return ((tc >> 4) << 8) + (tc & 15) + 16 * k; }

// stages 1 and 2 of a C-column tile after stage 0 (x0, x1 in the stage-0
// layout), and the stores of the stage-2 layout
template <int C>
FA_DEV void fa_col_fwdw_rest(double* x0, double* x1, double* __restrict__ dst, int g, double n,
                             double ninv, const double* __restrict__ w2, double* lds)
{
    const int c = threadIdx.x & (C - 1), u = threadIdx.x / C;
    const uint64_t col = (uint64_t)g * C + c;
    double tw[15];
    // exchange 0 -> 1 in two halves
    double y0[16], y1[16];
#pragma unroll
    for (int H = 0; H < 2; H++) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            lds[fa_cw_addr<C>(c, u + 256 * (8 * H + k))] = x0[8 * H + k];
            lds[fa_cw_addr<C>(c, u + 128 + 256 * (8 * H + k))] = x1[8 * H + k];
        }
        __syncthreads();
        const int tc = 128 * H + u;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const double v = lds[fa_cw_addr<C>(c, fa_c8_row1(tc, k))];
            if (H == 0) y0[k] = v; else y1[k] = v;
        }
        __syncthreads();
    }
    // stage 1: levels 4..7, block tc >> 4
    fa_tw16(w2, (uint64_t)(u >> 4), tw);
    fa_fwd16(y0, tw, n, ninv);
    fa_tw16(w2, (uint64_t)((u + 128) >> 4), tw);
    fa_fwd16(y1, tw, n, ninv);
    // exchange 1 -> 2, half H holds y_H entirely
#pragma unroll
    for (int H = 0; H < 2; H++) {
        const int tc = 128 * H + u;
#pragma unroll
        for (int k = 0; k < 16; k++)
            lds[fa_cw_addr<C>(c, fa_c8_row1(tc, k))] = H == 0 ? y0[k] : y1[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const double v = lds[fa_cw_addr<C>(c, 16 * tc + k)];
            if (H == 0) x0[k] = v; else x1[k] = v;
        }
        if (H == 0) __syncthreads();
    }
    // stage 2: levels 8..11, block tc
    fa_tw16(w2, (uint64_t)u, tw);
    fa_fwd16(x0, tw, n, ninv);
    fa_tw16(w2, (uint64_t)(u + 128), tw);
    fa_fwd16(x1, tw, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) {
        dst[(uint64_t)(16 * u + k) * FA_ROW_PITCH + col] = x0[k];
        dst[(uint64_t)(16 * (u + 128) + k) * FA_ROW_PITCH + col] = x1[k];
    }
}

// OWN > 0: the chunk width in bytes of a byte-aligned import, fixed at
// compile time (the import's word masks and partial sums fold away)
template <int C, int KIND, int OWN = 0>
FA_DEV void fa_col_fwdw_tile(const fa_operand& op, double* __restrict__ dst, int d2, int g,
                             double n, double ninv, const double* __restrict__ w2,
                             const double* __restrict__ pow2, double* lds)
{
    const int c = threadIdx.x & (C - 1), u = threadIdx.x / C;
    const uint64_t col = (uint64_t)g * C + c;
    double x0[16], x1[16], tw[15];
    // rows >= 2048 (elements >= 2^(d2+11)) all zero: the usual case of a
    // product's operands, each at most half the transform length
    const bool hz = fa_operand_len(op) <= (UINT64_C(1) << (d2 + 11));
    fa_tw16(w2, 0, tw);
    if constexpr (KIND == FA_PACK_MPN8) {
        // byte-aligned chunks: when the last chunk a lane reads (row u + 128
        // + 256 (nk - 1)) has its 32-byte window inside the limbs for every
        // lane of the wave, all of the wave's chunks do; they are then read
        // without per-chunk checks, at offsets stepping by 256 rows.  Each
        // import path runs the rest of the tile on its own (no join of
        // differently allocated registers).
        const int own = OWN ? OWN : op.bits >> 3;
        const uint64_t last = ((uint64_t)(u + 128 + 256 * (hz ? 7 : 15)) << d2) + col;
        if (__all(((last * (uint64_t)own) & ~UINT64_C(3)) + 32 <= op.n * 8)) {
            const char* base = (const char*)op.src;
            const uint64_t B0 = (((uint64_t)u << d2) + col) * (uint64_t)own;
            const uint64_t step = ((uint64_t)256 << d2) * (uint64_t)own;
            const uint64_t half = ((uint64_t)128 << d2) * (uint64_t)own;
            if (hz) {
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    x0[k] = fa_chunk_residue8_in(base, B0 + k * step, own, pow2, n, ninv);
                    x1[k] = fa_chunk_residue8_in(base, B0 + half + k * step, own, pow2, n, ninv);
                }
#pragma unroll
                for (int k = 8; k < 16; k++) x0[k] = x1[k] = 0.0;
                fa_fwd16_hz(x0, tw, n, ninv);
                fa_fwd16_hz(x1, tw, n, ninv);
            } else {
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    x0[k] = fa_chunk_residue8_in(base, B0 + k * step, own, pow2, n, ninv);
                    x1[k] = fa_chunk_residue8_in(base, B0 + half + k * step, own, pow2, n, ninv);
                }
                fa_fwd16(x0, tw, n, ninv);
                fa_fwd16(x1, tw, n, ninv);
            }
            fa_col_fwdw_rest<C>(x0, x1, dst, g, n, ninv, w2, lds);
            return;
        }
    }
    if (hz) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            x0[k] = fa_operand_value_k<KIND>(op, ((uint64_t)(u + 256 * k) << d2) + col, pow2, n,
                                             ninv);
            x1[k] = fa_operand_value_k<KIND>(op, ((uint64_t)(u + 128 + 256 * k) << d2) + col,
                                             pow2, n, ninv);
        }
#pragma unroll
        for (int k = 8; k < 16; k++) x0[k] = x1[k] = 0.0;
        // stage 0: levels 0..3, block 0
        fa_fwd16_hz(x0, tw, n, ninv);
        fa_fwd16_hz(x1, tw, n, ninv);
    } else {
#pragma unroll
        for (int k = 0; k < 16; k++) {
            x0[k] = fa_operand_value_k<KIND>(op, ((uint64_t)(u + 256 * k) << d2) + col, pow2, n,
                                             ninv);
            x1[k] = fa_operand_value_k<KIND>(op, ((uint64_t)(u + 128 + 256 * k) << d2) + col,
                                             pow2, n, ninv);
        }
        fa_fwd16(x0, tw, n, ninv);
        fa_fwd16(x1, tw, n, ninv);
    }
    fa_col_fwdw_rest<C>(x0, x1, dst, g, n, ninv, w2, lds);
}

template <int C>
FA_DEV void fa_col_invw_tile(double* __restrict__ dst, int d2, int g, double n, double ninv,
                             const double* __restrict__ wi, double* lds)
{
    const int c = threadIdx.x & (C - 1), u = threadIdx.x / C;
    const uint64_t col = (uint64_t)g * C + c;
    double x0[16], x1[16], tw[15];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        x0[k] = dst[(uint64_t)(16 * u + k) * FA_ROW_PITCH + col];
        x1[k] = dst[(uint64_t)(16 * (u + 128) + k) * FA_ROW_PITCH + col];
    }
    // stage 2 (levels 11..8), block tc
    fa_tw16(wi, (uint64_t)u, tw);
    fa_inv16(x0, tw, n, ninv);
    fa_tw16(wi, (uint64_t)(u + 128), tw);
    fa_inv16(x1, tw, n, ninv);
    // exchange 2 -> 1, half H holds x_H entirely
    double y0[16], y1[16];
#pragma unroll
    for (int H = 0; H < 2; H++) {
        const int tc = 128 * H + u;
#pragma unroll
        for (int k = 0; k < 16; k++) lds[fa_cw_addr<C>(c, 16 * tc + k)] = H == 0 ? x0[k] : x1[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const double v = lds[fa_cw_addr<C>(c, fa_c8_row1(tc, k))];
            if (H == 0) y0[k] = v; else y1[k] = v;
        }
        __syncthreads();
    }
    // stage 1 (levels 7..4), block tc >> 4
    fa_tw16(wi, (uint64_t)(u >> 4), tw);
    fa_inv16(y0, tw, n, ninv);
    fa_tw16(wi, (uint64_t)((u + 128) >> 4), tw);
    fa_inv16(y1, tw, n, ninv);
    // exchange 1 -> 0: half H receives stage-1 group 128H + u and returns
    // elements 8H .. 8H+7 of both stage-0 groups
#pragma unroll
    for (int H = 0; H < 2; H++) {
        const int tc = 128 * H + u;
#pragma unroll
        for (int k = 0; k < 16; k++)
            lds[fa_cw_addr<C>(c, fa_c8_row1(tc, k))] = H == 0 ? y0[k] : y1[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; k++) {
            x0[8 * H + k] = lds[fa_cw_addr<C>(c, u + 256 * (8 * H + k))];
            x1[8 * H + k] = lds[fa_cw_addr<C>(c, u + 128 + 256 * (8 * H + k))];
        }
        if (H == 0) __syncthreads();
    }
    // stage 0 (levels 3..0), block 0
    fa_tw16(wi, 0, tw);
    fa_inv16(x0, tw, n, ninv);
    fa_inv16(x1, tw, n, ninv);
#pragma unroll
    for (int k = 0; k < 16; k++) {
        dst[(uint64_t)(u + 256 * k) * FA_ROW_PITCH + col] = x0[k];
        dst[(uint64_t)(u + 128 + 256 * k) * FA_ROW_PITCH + col] = x1[k];
    }
}

template <int KIND>
FA_DEV void fa_col_fwd8_tile(const fa_operand& op, double* __restrict__ dst, int d2, int g,
                             double n, double ninv, const double* __restrict__ w2,
                             const double* __restrict__ pow2, double* lds)
{
    fa_col_fwdw_tile<8, KIND>(op, dst, d2, g, n, ninv, w2, pow2, lds);
}

FA_DEV void fa_col_inv8_tile(double* __restrict__ dst, int d2, int g, double n, double ninv,
                             const double* __restrict__ wi, double* lds)
{
    fa_col_invw_tile<8>(dst, d2, g, n, ninv, wi, lds);
}

// work index L -> (prime i, column group g, operand z) with the 16/C column
// groups of one 128-byte line adjacent, as fa_col_split does for two
template <int C>
FA_DEV void fa_colw_split(unsigned L, int np, int ngroups, int* i, int* g, int* z)
{
    constexpr int LG = (C == 8) ? 1 : 2;   // log2(16 / C)
    const int lo = L & ((1 << LG) - 1);
    const unsigned L2 = L >> LG;
    *i = (int)(L2 % (unsigned)np);
    const unsigned rest = L2 / (unsigned)np;
    const int lines = ngroups >> LG;
    *g = ((int)(rest % (unsigned)lines) << LG) + lo;
    *z = (int)(rest / (unsigned)lines);
}

template <int C, int KIND, int OWN = 0>
__global__ void __launch_bounds__(128 * C) fa_k_col_fwd16w(fa_operand a, fa_operand b, int d2,
                                                           double* __restrict__ A,
                                                           double* __restrict__ B,
                                                           uint64_t stride, fa_prime_set S,
                                                           int np, int ngroups, int nops)
{
    __shared__ double lds[FA_CW_LDS(C)];
    int i, g, z;
    fa_colw_split<C>(fa_xcd_linear(), np, ngroups, &i, &g, &z);
    if (z >= nops) return;
    fa_col_fwdw_tile<C, KIND, OWN>(z ? b : a, (z ? B : A) + (uint64_t)i * stride, d2, g, S.p[i],
                              S.pinv[i], S.w2[i], S.pow2[i], lds);
}

template <int C>
__global__ void __launch_bounds__(128 * C) fa_k_col_inv16w(double* __restrict__ A, int d2,
                                                           uint64_t stride, fa_prime_set S,
                                                           int np, int ngroups)
{
    __shared__ double lds[FA_CW_LDS(C)];
    int i, g, z;
    fa_colw_split<C>(fa_xcd_linear(), np, ngroups, &i, &g, &z);
    if (z >= 1) return;
    fa_col_invw_tile<C>(A + (uint64_t)i * stride, d2, g, S.p[i], S.pinv[i], S.w2inv[i], lds);
}

// ---------------------------------------------------------------------------
// batched kernels: the same tiles, each workgroup taking one work item
// ---------------------------------------------------------------------------
struct fa_witem {
    unsigned job;
    int prime, op, tile;
};

FA_DEV fa_witem fa_decode(uint64_t w)
{
    fa_witem r;
    r.job = (unsigned)(w & 0xffffffffu);
    r.prime = (int)((w >> 32) & 15);
    r.op = (int)((w >> 36) & 1);
    r.tile = (int)(w >> 40);
    return r;
}

FA_DEV fa_operand fa_job_operand(const fa_bjob& j, int z)
{
    fa_operand o;
    o.kind = j.kind;
    o.src = z ? j.b : j.a;
    o.n = z ? j.bn : j.an;
    o.bits = j.bits;
    o.width = z ? j.bw : j.aw;
    return o;
}

template <int D1, int KIND>
__global__ void __launch_bounds__(1024) fa_kb_col_fwd(const fa_bjob* __restrict__ J,
                                                      const uint64_t* __restrict__ work,
                                                      unsigned nwork, fa_prime_set S)
{
    __shared__ double lds[(D1 == 12) ? FA_C8_LDS : col_geom<D1>::LDS];
    const unsigned L = fa_xcd_linear();
    if (L >= nwork) return;
    const fa_witem w = fa_decode(work[L]);
    const fa_bjob& j = J[w.job];
    const int pi = j.offset + w.prime;
    const fa_operand op = fa_job_operand(j, w.op);
    double* dst = (w.op ? j.B : j.A) + (uint64_t)w.prime * j.stride;
    if constexpr (D1 == 12)
        fa_col_fwd8_tile<KIND>(op, dst, FA_ROW_LG, w.tile, S.p[pi], S.pinv[pi], S.w2[pi],
                               S.pow2[pi], lds);
    else
        fa_col_fwd_tile<D1, KIND>(op, dst, FA_ROW_LG, w.tile, S.p[pi], S.pinv[pi], S.w2[pi],
                                  S.pow2[pi], lds, j.trows);
}

template <int D1>
__global__ void __launch_bounds__(1024) fa_kb_col_inv(const fa_bjob* __restrict__ J,
                                                      const uint64_t* __restrict__ work,
                                                      unsigned nwork, fa_prime_set S)
{
    __shared__ double lds[(D1 == 12) ? FA_C8_LDS : fa_col_inv_lds<D1>()];
    const unsigned L = fa_xcd_linear();
    if (L >= nwork) return;
    const fa_witem w = fa_decode(work[L]);
    const fa_bjob& j = J[w.job];
    const int pi = j.offset + w.prime;
    double* dst = j.A + (uint64_t)w.prime * j.stride;
    if constexpr (D1 == 12)
        fa_col_inv8_tile(dst, FA_ROW_LG, w.tile, S.p[pi], S.pinv[pi], S.w2inv[pi], lds);
    else if (j.trows < (1 << D1))
        fa_col_inv_trunc_tile<D1>(dst, w.tile, j.trows, S.p[pi], S.pinv[pi], S.w2[pi],
                                  S.w2inv[pi], lds);
    else
        fa_col_inv_tile<D1>(dst, FA_ROW_LG, w.tile, S.p[pi], S.pinv[pi], S.w2inv[pi], lds);
}

__global__ void __launch_bounds__(256, 3) fa_kb_row_conv(const fa_bjob* __restrict__ J,
                                                         const uint64_t* __restrict__ work,
                                                         unsigned nwork, fa_prime_set S)
{
    __shared__ double xch[4352];
    __shared__ double tab[256];
    __shared__ double rho[16];
    const unsigned L = fa_xcd_linear();
    if (L >= nwork) return;
    const fa_witem w = fa_decode(work[L]);
    const fa_bjob& j = J[w.job];
    const int pi = j.offset + w.prime;
    const uint64_t r = (uint64_t)w.tile;
    const uint64_t off = (uint64_t)w.prime * j.stride + r * FA_ROW_PITCH;
    fa_row_conv_tile_lr(j.A + off, j.B + off, r, S.p[pi], S.pinv[pi], j.norm[w.prime], S.w2[pi],
                        S.w2inv[pi], xch, tab, rho);
}

// whole product of one prime in LDS (depth <= 12): import, both forward
// transforms, pointwise, inverse; the result goes to row `prime` of A
template <int KIND>
__global__ void __launch_bounds__(256) fa_kb_small_conv(const fa_bjob* __restrict__ J,
                                                        const uint64_t* __restrict__ work,
                                                        unsigned nwork, fa_prime_set S)
{
    extern __shared__ double lds[];
    const unsigned L = fa_xcd_linear();
    if (L >= nwork) return;
    const fa_witem w = fa_decode(work[L]);
    const fa_bjob& j = J[w.job];
    const int pi = j.offset + w.prime;
    const double n = S.p[pi], ninv = S.pinv[pi];
    const int depth = j.depth;
    const int N = 1 << depth;
    const fa_operand oa = fa_job_operand(j, 0), ob = fa_job_operand(j, 1);
    double* sa = lds;
    double* sb = lds + N;
    for (int x = threadIdx.x; x < N; x += blockDim.x) {
        sa[x] = fa_operand_value_k<KIND>(oa, (uint64_t)x, S.pow2[pi], n, ninv);
        sb[x] = fa_operand_value_k<KIND>(ob, (uint64_t)x, S.pow2[pi], n, ninv);
    }
    __syncthreads();
    fa_lds_fwd(sa, depth, 0, 0, S.w2[pi], n, ninv);
    fa_lds_fwd(sb, depth, 0, 0, S.w2[pi], n, ninv);
    const double m = j.norm[w.prime];
    for (int x = threadIdx.x; x < N; x += blockDim.x)
        sa[x] = fa_mulmod(fa_mulmod(sa[x], m, n, ninv), sb[x], n, ninv);
    __syncthreads();
    fa_lds_inv(sa, depth, 0, 0, S.w2inv[pi], n, ninv);
    double* o = j.A + (uint64_t)w.prime * j.stride;
    for (int x = threadIdx.x; x < N; x += blockDim.x) o[x] = sa[x];
}

static unsigned grid8(unsigned n) { return (n + 7) & ~7u; }

int fa_batch_col_groups(int depth)
{
    const int d1 = depth - FA_ROW_LG;
    if (d1 == 12) return 4096 / 8;
    if (d1 >= 4) return 4096 / (1024 >> (d1 - 4));
    return 4096 / 1024;
}

int fa_batch_small_bin(int depth, int kind, const fa_bjob* J, const uint64_t* work,
                       unsigned nwork, const fa_prime_set& S, hipStream_t s)
{
    if (!nwork) return FA_OK;
    const size_t lds = 2 * ((size_t)1 << depth) * sizeof(double);
    switch (kind) {
    case FA_PACK_MPN8:
        FA_LAUNCH("fa_kb_small_conv", fa_kb_small_conv<FA_PACK_MPN8>, dim3(grid8(nwork)),
                  dim3(256), lds, s, J, work, nwork, S);
        break;
    case FA_PACK_U64:
        FA_LAUNCH("fa_kb_small_conv", fa_kb_small_conv<FA_PACK_U64>, dim3(grid8(nwork)),
                  dim3(256), lds, s, J, work, nwork, S);
        break;
    case FA_PACK_S64W:
        FA_LAUNCH("fa_kb_small_conv", fa_kb_small_conv<FA_PACK_S64W>, dim3(grid8(nwork)),
                  dim3(256), lds, s, J, work, nwork, S);
        break;
    default:
        return fa_set_error(FA_E_ARG, "batch: import kind %d", kind);
    }
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "batch launch");
}

template <int D1, int KIND>
static void batch_bin(const fa_bjob* J, const uint64_t* wf, unsigned nf, const uint64_t* wr,
                      unsigned nr, const uint64_t* wi, unsigned ni, const fa_prime_set& S,
                      hipStream_t s)
{
    if (nf)
        FA_LAUNCH("fa_kb_col_fwd", (fa_kb_col_fwd<D1, KIND>), dim3(grid8(nf)), dim3(1024), 0, s,
                  J, wf, nf, S);
    if (nr)
        FA_LAUNCH("fa_kb_row_conv", fa_kb_row_conv, dim3(grid8(nr)), dim3(256), 0, s, J, wr, nr,
                  S);
    if (ni)
        FA_LAUNCH("fa_kb_col_inv", fa_kb_col_inv<D1>, dim3(grid8(ni)), dim3(1024), 0, s, J, wi,
                  ni, S);
}

template <int D1>
static int batch_bin_kind(int kind, const fa_bjob* J, const uint64_t* wf, unsigned nf,
                          const uint64_t* wr, unsigned nr, const uint64_t* wi, unsigned ni,
                          const fa_prime_set& S, hipStream_t s)
{
    switch (kind) {
    case FA_PACK_MPN8: batch_bin<D1, FA_PACK_MPN8>(J, wf, nf, wr, nr, wi, ni, S, s); break;
    case FA_PACK_U64: batch_bin<D1, FA_PACK_U64>(J, wf, nf, wr, nr, wi, ni, S, s); break;
    case FA_PACK_S64W: batch_bin<D1, FA_PACK_S64W>(J, wf, nf, wr, nr, wi, ni, S, s); break;
    default: return fa_set_error(FA_E_ARG, "batch: import kind %d", kind);
    }
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "batch launch");
}

int fa_batch_transform_bin(int depth, int kind, const fa_bjob* J, const uint64_t* wf,
                           unsigned nf, const uint64_t* wr, unsigned nr, const uint64_t* wi,
                           unsigned ni, const fa_prime_set& S, hipStream_t s)
{
    switch (depth - FA_ROW_LG) {
    case 1: return batch_bin_kind<1>(kind, J, wf, nf, wr, nr, wi, ni, S, s);
    case 2: return batch_bin_kind<2>(kind, J, wf, nf, wr, nr, wi, ni, S, s);
    case 3: return batch_bin_kind<3>(kind, J, wf, nf, wr, nr, wi, ni, S, s);
    case 4: return batch_bin_kind<4>(kind, J, wf, nf, wr, nr, wi, ni, S, s);
    case 5: return batch_bin_kind<5>(kind, J, wf, nf, wr, nr, wi, ni, S, s);
    case 6: return batch_bin_kind<6>(kind, J, wf, nf, wr, nr, wi, ni, S, s);
    case 7: return batch_bin_kind<7>(kind, J, wf, nf, wr, nr, wi, ni, S, s);
    case 8: return batch_bin_kind<8>(kind, J, wf, nf, wr, nr, wi, ni, S, s);
    case 9: return batch_bin_kind<9>(kind, J, wf, nf, wr, nr, wi, ni, S, s);
    case 10: return batch_bin_kind<10>(kind, J, wf, nf, wr, nr, wi, ni, S, s);
    case 11: return batch_bin_kind<11>(kind, J, wf, nf, wr, nr, wi, ni, S, s);
    case 12: return batch_bin_kind<12>(kind, J, wf, nf, wr, nr, wi, ni, S, s);
    default: return fa_set_error(FA_E_ARG, "batch: depth %d outside 13..24", depth);
    }
}

// truncated transforms (rows past the product's length skipped): 0 never,
// 1 automatic, 2 wherever they apply (tests compare them with the full
// transforms)
int fa_trunc_enabled = 1;
extern "C" FA_API void fa_set_truncation(int mode)
{
    fa_trunc_enabled = mode < 0 ? 0 : mode > 2 ? 2 : mode;
}
// the automatic choice: truncate at t of 2^d1 rows only where the row pass
// saved outweighs the slower truncated inverse (measured on MI355X: 2^23
// points with at most 55% of the rows; elsewhere the LDS truncated inverse
// costs more than the skipped rows, in single products and in batches)
int fa_trunc_rule(int d1, int t)
{
    return d1 == 11 && 20 * t <= 11 << d1;
}

// column tiles at d1 = 12: 1 = 8 columns, 1024 threads; 2 = 4 columns, 512
// threads, half-column LDS; 0 = 4 columns, 1024 threads, whole-column LDS
int fa_col_variant = 1;
extern "C" FA_API void fa_set_col_variant(int v) { fa_col_variant = v; }
// row pass: 1 = three workgroups per CU (twiddles of the last stage formed
// per layer, 168 registers), 0 = two workgroups per CU (all 15 held)
int fa_row_variant = 1;
extern "C" FA_API void fa_set_row_variant(int v) { fa_row_variant = v; }

// ---------------------------------------------------------------------------
// host launcher: transform lengths 2^13 .. 2^24 (row 2^12, column 2^1 ..
// 2^12).  Returns 1 if it handled the depth, 0 otherwise.
// ---------------------------------------------------------------------------

extern int fa_col_variant;
extern int fa_row_variant;

// the primes [h, np) of a prime set as a set of their own
static fa_prime_set prime_subset(const fa_prime_set& S, int h, int np)
{
    fa_prime_set T = S;
    for (int k = 0; k + h < np; k++) {
        T.p[k] = S.p[h + k];
        T.pinv[k] = S.pinv[h + k];
        T.norm[k] = S.norm[h + k];
        for (int e = 0; e < 8; e++) T.pow2[k][e] = S.pow2[h + k][e];
        T.w2[k] = S.w2[h + k];
        T.w2inv[k] = S.w2inv[h + k];
    }
    return T;
}

// further streams of the calling host thread (default priority) and the
// events that fork them from and join them to the caller's stream
#define FA_PIPE_MAX 3
struct fa_aux {
    int device = -1;
    hipStream_t s[FA_PIPE_MAX - 1] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t fwd[FA_PIPE_MAX] = {};
    hipEvent_t join[FA_PIPE_MAX - 1] = {};
};

static fa_aux* aux_streams()
{
    static thread_local fa_aux X;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    if (X.device == dev) return &X;
    if (X.device >= 0) return nullptr;   // created on another device: none here
    for (int k = 0; k < FA_PIPE_MAX - 1; k++)
        if (hipStreamCreateWithFlags(&X.s[k], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&X.join[k], hipEventDisableTiming) != hipSuccess)
            return nullptr;
    for (int k = 0; k < FA_PIPE_MAX; k++)
        if (hipEventCreateWithFlags(&X.fwd[k], hipEventDisableTiming) != hipSuccess) return nullptr;
    if (hipEventCreateWithFlags(&X.fork, hipEventDisableTiming) != hipSuccess) return nullptr;
    X.device = dev;
    return &X;
}

// prime groups of the 2^24-point passes on their own streams (0 or 1: one
// launch per pass over all primes)
int fa_pipe_groups = 2;
extern "C" FA_API void fa_set_prime_groups(int v) { fa_pipe_groups = v; }

// ng prime groups, group g on its own stream, each starting when the
// previous group's forward column pass is done: the forward and row passes
// of later groups (arithmetic-bound) run beside the row and inverse column
// passes of earlier ones (the latter bound by HBM).  Same kernels, same
// results as one launch per pass over all primes.
// nops = 1: a square (only a is transformed; the row pass squares)
template <int D1, int KIND>
static int launch_fast_pipe(const fa_operand& oa, const fa_operand& ob, int nops, int np, int ng,
                            double* A, double* B, uint64_t stride, const fa_prime_set& S,
                            hipStream_t s, fa_aux* X)
{
    const int d2 = FA_ROW_LG;
    const int ngroups = (1 << d2) / 8;
    if (hipEventRecord(X->fork, s) != hipSuccess) return fa_set_error(FA_E_HIP, "pipeline fork");
    for (int g = 0; g < ng; g++) {
        const int lo = np * g / ng, hi = np * (g + 1) / ng, m = hi - lo;
        const fa_prime_set Sg = prime_subset(S, lo, np);
        double* Ag = A + (uint64_t)lo * stride;
        double* Bg = B + (uint64_t)lo * stride;
        hipStream_t t = g ? X->s[g - 1] : s;
        if (g && (hipStreamWaitEvent(t, X->fork, 0) != hipSuccess ||
                  hipStreamWaitEvent(t, X->fwd[g - 1], 0) != hipSuccess))
            return fa_set_error(FA_E_HIP, "pipeline fork");
        const unsigned gf = ((unsigned)(m * ngroups * nops) + 7) & ~7u;
        if (KIND == FA_PACK_MPN8 && oa.bits == 136 && ob.bits == 136)
            FA_LAUNCH("fa_k_col_fwd16w", (fa_k_col_fwd16w<8, KIND, 17>), dim3(gf), dim3(1024), 0,
                      t, oa, ob, d2, Ag, Bg, stride, Sg, m, ngroups, nops);
        else
            FA_LAUNCH("fa_k_col_fwd16w", (fa_k_col_fwd16w<8, KIND>), dim3(gf), dim3(1024), 0, t,
                      oa, ob, d2, Ag, Bg, stride, Sg, m, ngroups, nops);
        if (hipEventRecord(X->fwd[g], t) != hipSuccess)
            return fa_set_error(FA_E_HIP, "pipeline event");
        if (nops == 1)
            FA_LAUNCH("fa_k_row_conv16", fa_k_row_conv16lr<true>, dim3(1u << D1, m), dim3(256), 0,
                      t, Ag, Ag, stride, Sg);
        else
            FA_LAUNCH("fa_k_row_conv16", fa_k_row_conv16lr<false>, dim3(1u << D1, m), dim3(256), 0,
                      t, Ag, Bg, stride, Sg);
        const unsigned gi = ((unsigned)(m * ngroups) + 7) & ~7u;
        FA_LAUNCH("fa_k_col_inv16w", fa_k_col_inv16w<8>, dim3(gi), dim3(1024), 0, t, Ag, d2, stride,
                  Sg, m, ngroups);
        if (g && hipEventRecord(X->join[g - 1], t) != hipSuccess)
            return fa_set_error(FA_E_HIP, "pipeline join");
    }
    for (int g = 1; g < ng; g++)
        if (hipStreamWaitEvent(s, X->join[g - 1], 0) != hipSuccess)
            return fa_set_error(FA_E_HIP, "pipeline join");
    return 1;
}

// trows: rows of the product's coefficients (its length / 4096, rounded
// up); the row pass and the inverse column pass stop there (D1 < 12)
template <int D1, int KIND>
static int launch_fast(const fa_operand& oa, const fa_operand& ob, int nops, int np, double* A,
                       double* B, uint64_t stride, const fa_prime_set& S, hipStream_t s,
                       int trows)
{
    if (D1 == 12 || trows < 1 || trows > (1 << D1)) trows = 1 << D1;
    if constexpr (D1 == 12) {
        if (fa_pipe_groups >= 2 && np >= 2 && fa_col_variant == 1 && fa_row_variant == 1 &&
            !fa_prof_on()) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
                fa_aux* X = aux_streams();
                const int ng = fa_pipe_groups < np ? fa_pipe_groups : np;
                if (X && ng >= 2 && ng <= FA_PIPE_MAX)
                    return launch_fast_pipe<D1, KIND>(oa, ob, nops, np, ng, A, B, stride, S, s,
                                                      X);
            }
        }
    }
    typedef col_geom<D1> G;
    const int d2 = FA_ROW_LG;
    const int wide = (D1 == 12 && (fa_col_variant == 1 || fa_col_variant == 2))
                         ? (fa_col_variant == 1 ? 8 : 4) : 0;
    const int ngroups = (1 << d2) / (wide ? wide : G::CPW);
    unsigned total = (unsigned)(np * ngroups * nops);
    unsigned grid = (total + 7) & ~7u;
    if (wide == 8 && KIND == FA_PACK_MPN8 && oa.bits == 136 && ob.bits == 136)
        FA_LAUNCH("fa_k_col_fwd16w", (fa_k_col_fwd16w<8, KIND, 17>), dim3(grid), dim3(1024), 0, s,
                  oa, ob, d2, A, B, stride, S, np, ngroups, nops);
    else if (wide == 8)
        FA_LAUNCH("fa_k_col_fwd16w", (fa_k_col_fwd16w<8, KIND>), dim3(grid), dim3(1024), 0, s, oa,
                  ob, d2, A, B, stride, S, np, ngroups, nops);
    else if (wide == 4)
        FA_LAUNCH("fa_k_col_fwd16w", (fa_k_col_fwd16w<4, KIND>), dim3(grid), dim3(512), 0, s, oa,
                  ob, d2, A, B, stride, S, np, ngroups, nops);
    else
        FA_LAUNCH("fa_k_col_fwd16", (fa_k_col_fwd16<D1, KIND>), dim3(grid), dim3(1024), 0, s, oa, ob,
                  d2, A, B, stride, S, np, ngroups, nops, trows);
    const unsigned nrow = (unsigned)trows;
    if (fa_row_variant == 1 && nops == 1)
        FA_LAUNCH("fa_k_row_conv16", fa_k_row_conv16lr<true>, dim3(nrow, np), dim3(256), 0, s,
                  A, A, stride, S);
    else if (fa_row_variant == 1)
        FA_LAUNCH("fa_k_row_conv16", fa_k_row_conv16lr<false>, dim3(nrow, np), dim3(256), 0, s,
                  A, B, stride, S);
    else
        FA_LAUNCH("fa_k_row_conv16", fa_k_row_conv16, dim3(nrow, np), dim3(256), 0, s, A,
                  nops == 1 ? A : B, stride, S);
    unsigned total2 = (unsigned)(np * ngroups);
    unsigned grid2 = (total2 + 7) & ~7u;
    if (wide == 8)
        FA_LAUNCH("fa_k_col_inv16w", fa_k_col_inv16w<8>, dim3(grid2), dim3(1024), 0, s, A, d2,
                  stride, S, np, ngroups);
    else if (wide == 4)
        FA_LAUNCH("fa_k_col_inv16w", fa_k_col_inv16w<4>, dim3(grid2), dim3(512), 0, s, A, d2,
                  stride, S, np, ngroups);
    else
        FA_LAUNCH("fa_k_col_inv16", fa_k_col_inv16<D1>, dim3(grid2), dim3(1024), 0, s, A, d2, stride,
                  S, np, ngroups, trows);
    return 1;
}

template <int D1>
static int launch_fast_any(const fa_operand& oa, const fa_operand& ob, int nops, int np,
                           double* A, double* B, uint64_t stride, const fa_prime_set& S,
                           hipStream_t s, int trows)
{
    switch (fa_pack_kind(oa, ob)) {
    case FA_PACK_MPN8:
        return launch_fast<D1, FA_PACK_MPN8>(oa, ob, nops, np, A, B, stride, S, s, trows);
    case FA_PACK_U64:
        return launch_fast<D1, FA_PACK_U64>(oa, ob, nops, np, A, B, stride, S, s, trows);
    case FA_PACK_S64W:
        return launch_fast<D1, FA_PACK_S64W>(oa, ob, nops, np, A, B, stride, S, s, trows);
    default: return launch_fast<D1, FA_PACK_ANY>(oa, ob, nops, np, A, B, stride, S, s, trows);
    }
}

// forward transform of one operand into A (column pass with the import,
// then the row pass), for the transformed-operand interface
template <int D1, int KIND>
static int launch_forward(const fa_operand& oa, int np, double* A, uint64_t stride,
                          const fa_prime_set& S, hipStream_t s)
{
    typedef col_geom<D1> G;
    const int d2 = FA_ROW_LG;
    const int wide = D1 == 12 ? 8 : 0;
    const int ngroups = (1 << d2) / (wide ? wide : G::CPW);
    const unsigned grid = ((unsigned)(np * ngroups) + 7) & ~7u;
    if (wide)
        FA_LAUNCH("fa_k_col_fwd16w", (fa_k_col_fwd16w<8, KIND>), dim3(grid), dim3(1024), 0, s, oa,
                  oa, d2, A, A, stride, S, np, ngroups, 1);
    else
        FA_LAUNCH("fa_k_col_fwd16", (fa_k_col_fwd16<D1, KIND>), dim3(grid), dim3(1024), 0, s, oa,
                  oa, d2, A, A, stride, S, np, ngroups, 1, 1 << D1);
    FA_LAUNCH("fa_k_row_fwd16", fa_k_row_fwd16, dim3(1u << D1, np), dim3(256), 0, s, A, stride, S);
    return 1;
}

template <int D1>
static int launch_inverse(int np, double* A, uint64_t stride, const fa_prime_set& S,
                          hipStream_t s)
{
    typedef col_geom<D1> G;
    const int d2 = FA_ROW_LG;
    const int wide = D1 == 12 ? 8 : 0;
    const int ngroups = (1 << d2) / (wide ? wide : G::CPW);
    const unsigned grid = ((unsigned)(np * ngroups) + 7) & ~7u;
    FA_LAUNCH("fa_k_row_inv16", fa_k_row_inv16, dim3(1u << D1, np), dim3(256), 0, s, A, stride, S);
    if (wide)
        FA_LAUNCH("fa_k_col_inv16w", fa_k_col_inv16w<8>, dim3(grid), dim3(1024), 0, s, A, d2,
                  stride, S, np, ngroups);
    else
        FA_LAUNCH("fa_k_col_inv16", fa_k_col_inv16<D1>, dim3(grid), dim3(1024), 0, s, A, d2, stride,
                  S, np, ngroups, 1 << D1);
    return 1;
}

template <int D1>
static int launch_forward_any(const fa_operand& oa, int np, double* A, uint64_t stride,
                              const fa_prime_set& S, hipStream_t s)
{
    switch (fa_pack_kind(oa, oa)) {
    case FA_PACK_MPN8: return launch_forward<D1, FA_PACK_MPN8>(oa, np, A, stride, S, s);
    case FA_PACK_U64: return launch_forward<D1, FA_PACK_U64>(oa, np, A, stride, S, s);
    case FA_PACK_S64W: return launch_forward<D1, FA_PACK_S64W>(oa, np, A, stride, S, s);
    default: return launch_forward<D1, FA_PACK_ANY>(oa, np, A, stride, S, s);
    }
}

#define FA_D1_SWITCH(d1, F)                                                                \
    switch (d1) {                                                                          \
    case 1: return F(1); case 2: return F(2); case 3: return F(3); case 4: return F(4);    \
    case 5: return F(5); case 6: return F(6); case 7: return F(7); case 8: return F(8);    \
    case 9: return F(9); case 10: return F(10); case 11: return F(11);                     \
    case 12: return F(12);                                                                 \
    default: return 0;                                                                     \
    }

// depth 13..24: returns 1 when launched, 0 for other depths
int fa_forward_fast(int depth, const fa_operand& oa, int np, double* A, uint64_t stride,
                    const fa_prime_set& S, hipStream_t s)
{
#define FA_F(D) launch_forward_any<D>(oa, np, A, stride, S, s)
    FA_D1_SWITCH(depth - FA_ROW_LG, FA_F)
#undef FA_F
}

int fa_inverse_fast(int depth, int np, double* A, uint64_t stride, const fa_prime_set& S,
                    hipStream_t s)
{
#define FA_F(D) launch_inverse<D>(np, A, stride, S, s)
    FA_D1_SWITCH(depth - FA_ROW_LG, FA_F)
#undef FA_F
}

// sqr: a == b; only a is transformed and the pointwise stage squares.
// zlen: coefficients of the product (0: unknown, no truncation); the product
// must vanish from coefficient zlen on (a product shorter than the
// transform, not a wrapped one)
int fa_convolve_fast(int depth, const fa_operand& oa, const fa_operand& ob, int np, double* A,
                     double* B, uint64_t stride, const fa_prime_set& S, hipStream_t s, int sqr,
                     uint64_t zlen)
{
    const int nops = sqr ? 1 : 2;
    const int d1 = depth - FA_ROW_LG;
    int trows = d1 >= 1 && d1 <= 12 ? 1 << d1 : 0;
    if (fa_trunc_enabled && d1 >= 1 && d1 < 12 && zlen && zlen <= ((uint64_t)1 << depth)) {
        const int t = (int)((zlen + (1u << FA_ROW_LG) - 1) >> FA_ROW_LG);
        if (fa_trunc_enabled == 2 || fa_trunc_rule(d1, t)) trows = t;
    }
#define FA_F(D) launch_fast_any<D>(oa, ob, nops, np, A, B, stride, S, s, trows)
    FA_D1_SWITCH(d1, FA_F)
#undef FA_F
}
