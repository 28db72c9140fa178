// fa_pack.h -- operand import: the residue mod p of element x of an
// operand (a bit chunk of a limb array, a u64 coefficient, or a signed
// multi-limb coefficient).
#pragma once

#include "fa_internal.h"
#include "fa_arith.h"

// ------------------------------------------------------------------------
// limb splitting (mpn_mul.c:595-695, slow_mpn_to_fft :422-481)
// ------------------------------------------------------------------------

// 32-bit word of the little-endian limb array starting at byte offset
// `byte` (zero beyond the array)
FA_DEV uint32_t fa_load_u32_at(const uint32_t* __restrict__ w, uint64_t nwords, uint64_t byte)
{
    uint64_t i = byte >> 2;
    int sh = (int)(byte & 3) * 8;
    uint32_t lo = (i < nwords) ? w[i] : 0u;
    if (sh == 0) return lo;
    uint32_t hi = (i + 1 < nwords) ? w[i + 1] : 0u;
    return (lo >> sh) | (hi << (32 - sh));
}

// residue mod p of chunk j = bits [j*bits, (j+1)*bits) of the limb array
// (mpn_mul.c:595-695 splits into 32-bit pieces; here the aligned 64-bit
// words covering the chunk are read and split into 48-bit pieces, exact in
// a double): value = sum_k piece_k 2^(48k), reduced with 2^(48k) mod p.
// Result in [-p/2, p/2], or the plain value when bits <= 48.
FA_DEV uint64_t fa_word(const uint64_t* __restrict__ a, uint64_t n, uint64_t i)
{
    return i < n ? a[i] : 0;
}

FA_DEV double fa_chunk_residue(const uint64_t* __restrict__ limbs, uint64_t nlimbs, int bits,
                               uint64_t j, const double* __restrict__ pow2, double n,
                               double ninv)
{
    const uint64_t bit0 = j * (uint64_t)bits;
    const uint64_t w = bit0 >> 6;
    if (w >= nlimbs) return 0.0;
    const int sh = (int)(bit0 & 63);
    const int span = sh + bits;                  // <= 263 bits: at most five words
    uint64_t v[5];
#pragma unroll
    for (int k = 0; k < 5; k++) v[k] = (64 * k < span) ? fa_word(limbs, nlimbs, w + k) : 0;
    uint64_t u[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        u[k] = sh ? ((v[k] >> sh) | (v[k + 1] << (64 - sh))) : v[k];
        if (64 * k >= bits) u[k] = 0;
        else if (64 * (k + 1) > bits) u[k] &= (UINT64_C(1) << (bits - 64 * k)) - 1;
    }
    const uint64_t M48 = (UINT64_C(1) << 48) - 1;
    double p0 = (double)(u[0] & M48);
    if (bits <= 48) return p0;
    double p1 = (double)(((u[0] >> 48) | (u[1] << 16)) & M48);
    double p2 = (double)(((u[1] >> 32) | (u[2] << 32)) & M48);
    double p3 = (double)((u[2] >> 16) & M48);
    double p4 = (double)(u[3] & M48);
    double acc = p0 + fa_mulmod(p1, pow2[1], n, ninv);
    if (bits > 96) acc += fa_mulmod(p2, pow2[2], n, ninv);
    if (bits > 144) acc += fa_mulmod(p3, pow2[3], n, ninv);
    if (bits > 192) acc += fa_mulmod(p4, pow2[4], n, ninv);
    return fa_reduce_pm1n(acc, n, ninv);
}

// the same for bits = 8 own (own <= 25 bytes per chunk): the chunk's bytes
// [own j, own j + own) come from two 16-byte loads at the enclosing 4-byte
// boundary, realigned with v_alignbyte, masked to own bytes and cut into
// 48-bit pieces of 32-bit words.  The result is in (-2n, 2.25n) when
// own <= 18 (at most two products) and in [-n/2, n/2] otherwise -- inside
// the (-3n, 3n) input range of the forward transform.  The last chunks,
// whose 32-byte window reaches past the array, read it word by word.
typedef uint32_t fa_u32x4 __attribute__((ext_vector_type(4)));
typedef fa_u32x4 fa_u32x4_a4 __attribute__((aligned(4)));

FA_DEV double fa_chunk_residue8(const uint64_t* __restrict__ limbs, uint64_t nlimbs, int own,
                                uint64_t j, const double* __restrict__ pow2, double n,
                                double ninv)
{
    const uint64_t B = j * (uint64_t)own;
    const uint64_t nbytes = nlimbs * 8;
    if (B >= nbytes) return 0.0;
    const uint64_t B4 = B & ~UINT64_C(3);
    uint32_t w[8];
    if (B4 + 32 <= nbytes) {
        const fa_u32x4_a4* q = (const fa_u32x4_a4*)((const char*)limbs + B4);
        const fa_u32x4 lo4 = q[0], hi4 = q[1];
        w[0] = lo4.x; w[1] = lo4.y; w[2] = lo4.z; w[3] = lo4.w;
        w[4] = hi4.x; w[5] = hi4.y; w[6] = hi4.z; w[7] = hi4.w;
    } else {
        const uint32_t* q = (const uint32_t*)((const char*)limbs + B4);
#pragma unroll
        for (int m = 0; m < 8; m++) w[m] = (B4 + 4 * m < nbytes) ? q[m] : 0u;
    }
    const uint32_t s = (uint32_t)(B & 3);
    uint32_t c[7];
#pragma unroll
    for (int m = 0; m < 7; m++) {
        c[m] = __builtin_amdgcn_alignbyte(w[m + 1], w[m], s);
        const int vb = own - 4 * m;   // valid bytes of word m
        if (vb < 4) c[m] = (vb <= 0) ? 0u : (c[m] & ((1u << (8 * vb)) - 1u));
    }
    const double T16 = 65536.0, T32 = 4294967296.0;
    double acc = __builtin_fma((double)(c[1] & 0xffffu), T32, (double)c[0]);
    if (own > 6) {
        const double p1 = __builtin_fma((double)c[2], T16, (double)(c[1] >> 16));
        acc += fa_mulmod(p1, pow2[1], n, ninv);
    }
    if (own > 12) {
        const double p2 = __builtin_fma((double)(c[4] & 0xffffu), T32, (double)c[3]);
        acc += fa_mulmod(p2, pow2[2], n, ninv);
    }
    if (own > 18) {
        const double p3 = __builtin_fma((double)c[5], T16, (double)(c[4] >> 16));
        acc += fa_mulmod(p3, pow2[3], n, ninv);
        if (own > 24) acc += fa_mulmod((double)c[6], pow2[4], n, ninv);
        acc = fa_reduce_pm1n(acc, n, ninv);
    }
    return acc;
}

// fa_chunk_residue8 for a chunk whose 32-byte window [B & ~3, (B & ~3) + 32)
// lies inside the limb array (no bounds checks)
FA_DEV double fa_chunk_residue8_in(const char* __restrict__ base, uint64_t B, int own,
                                   const double* __restrict__ pow2, double n, double ninv)
{
    const fa_u32x4_a4* q = (const fa_u32x4_a4*)(base + (B & ~UINT64_C(3)));
    const fa_u32x4 lo4 = q[0], hi4 = q[1];
    const uint32_t w[8] = {lo4.x, lo4.y, lo4.z, lo4.w, hi4.x, hi4.y, hi4.z, hi4.w};
    const uint32_t s = (uint32_t)(B & 3);
    uint32_t c[7];
#pragma unroll
    for (int m = 0; m < 7; m++) {
        c[m] = __builtin_amdgcn_alignbyte(w[m + 1], w[m], s);
        const int vb = own - 4 * m;
        if (vb < 4) c[m] = (vb <= 0) ? 0u : (c[m] & ((1u << (8 * vb)) - 1u));
    }
    const double T16 = 65536.0, T32 = 4294967296.0;
    double acc = __builtin_fma((double)(c[1] & 0xffffu), T32, (double)c[0]);
    if (own > 6) {
        const double p1 = __builtin_fma((double)c[2], T16, (double)(c[1] >> 16));
        acc += fa_mulmod(p1, pow2[1], n, ninv);
    }
    if (own > 12) {
        const double p2 = __builtin_fma((double)(c[4] & 0xffffu), T32, (double)c[3]);
        acc += fa_mulmod(p2, pow2[2], n, ninv);
    }
    if (own > 18) {
        const double p3 = __builtin_fma((double)c[5], T16, (double)(c[4] >> 16));
        acc += fa_mulmod(p3, pow2[3], n, ninv);
        if (own > 24) acc += fa_mulmod((double)c[6], pow2[4], n, ninv);
        acc = fa_reduce_pm1n(acc, n, ninv);
    }
    return acc;
}

// ------------------------------------------------------------------------
// operand import: the value of element x (coefficient index) of an operand
// ------------------------------------------------------------------------

struct fa_operand {
    int kind;                  // FA_PACK_*
    const uint64_t* src;
    uint64_t n;                // limbs (mpn) or coefficients
    int bits;
    int width;
};

// signed w-limb two's complement coefficient mod p: Horner over the 32-bit
// words from the top, minus 2^(64 w) for a negative coefficient
FA_DEV double fa_coeff_residue_s(const uint64_t* __restrict__ c, int w,
                                 const double* __restrict__ pow2, double n, double ninv)
{
    const uint32_t* u = (const uint32_t*)c;
    const double c32 = pow2[0];
    double r = 0.0, t = 1.0;
    for (int k = 2 * w - 1; k >= 0; k--) {
        r = fa_reduce_pm1n(fa_mulmod(r, c32, n, ninv) + (double)u[k], n, ninv);
        t = fa_reduce_pm1n(fa_mulmod(t, c32, n, ninv), n, ninv);
    }
    if ((int64_t)c[w - 1] < 0) r = fa_reduce_pm1n(r - t, n, ninv);
    return r;
}

FA_DEV double fa_operand_value(const fa_operand& op, uint64_t x,
                               const double* __restrict__ pow2, double n, double ninv)
{
    if (op.kind == FA_PACK_MPN) {
        if ((op.bits & 7) == 0 && op.bits <= 200)
            return fa_chunk_residue8(op.src, op.n, op.bits >> 3, x, pow2, n, ninv);
        return fa_chunk_residue(op.src, op.n, op.bits, x, pow2, n, ninv);
    }
    if (x >= op.n) return 0.0;
    if (op.kind == FA_PACK_U64) {
        uint64_t v = op.src[x];
        double lo = (double)(uint32_t)v, hi = (double)(uint32_t)(v >> 32);
        if (v < (UINT64_C(1) << 50)) return (double)v;
        return fa_reduce_pm1n(lo + fa_mulmod(hi, pow2[0], n, ninv), n, ninv);
    }
    return fa_coeff_residue_s(op.src + x * (uint64_t)op.width, op.width, pow2, n, ninv);
}

// number of elements of an operand (chunks of a limb array, coefficients)
FA_DEV uint64_t fa_operand_len(const fa_operand& op)
{
    if (op.kind == FA_PACK_MPN) return (64 * op.n + (uint64_t)op.bits - 1) / (uint64_t)op.bits;
    return op.n;
}

// import specialized at compile time (KIND = FA_PACK_*), FA_PACK_ANY
// dispatching on op.kind
template <int KIND>
FA_DEV double fa_operand_value_k(const fa_operand& op, uint64_t x,
                                 const double* __restrict__ pow2, double n, double ninv)
{
    if constexpr (KIND == FA_PACK_MPN8) {
        return fa_chunk_residue8(op.src, op.n, op.bits >> 3, x, pow2, n, ninv);
    } else if constexpr (KIND == FA_PACK_MPN) {
        return fa_chunk_residue(op.src, op.n, op.bits, x, pow2, n, ninv);
    } else if constexpr (KIND == FA_PACK_U64) {
        if (x >= op.n) return 0.0;
        uint64_t v = op.src[x];
        double lo = (double)(uint32_t)v, hi = (double)(uint32_t)(v >> 32);
        if (v < (UINT64_C(1) << 50)) return (double)v;
        return fa_reduce_pm1n(lo + fa_mulmod(hi, pow2[0], n, ninv), n, ninv);
    } else if constexpr (KIND == FA_PACK_S64W) {
        if (x >= op.n) return 0.0;
        return fa_coeff_residue_s(op.src + x * (uint64_t)op.width, op.width, pow2, n, ninv);
    } else {
        return fa_operand_value(op, x, pow2, n, ninv);
    }
}

// the specialization a pair of operands can share
static inline int fa_pack_kind(const fa_operand& a, const fa_operand& b)
{
    if (a.kind != b.kind) return FA_PACK_ANY;
    if (a.kind == FA_PACK_MPN)
        return (a.bits == b.bits && (a.bits & 7) == 0 && a.bits <= 200) ? FA_PACK_MPN8
                                                                         : FA_PACK_MPN;
    return a.kind;
}
