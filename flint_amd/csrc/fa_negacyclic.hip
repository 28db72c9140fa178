// fa_negacyclic.hip -- the context form of sd_fft_mpn_mulmod_2expp1
// (negacyclic.c:43-501): products in Z/(2^N + 1) with one operand's
// transform cached across products, the shape FLINT's Schoenhage-Strassen
// convolutions use (fft/convolution_precache.c).  Operands and results
// have n + 1 limbs (n = N / 64), values in [0, 2^N].
//
// The device path forms the full product of the two (n + 1)-limb operands
// on a plan of the op interface and folds it mod 2^N + 1 (fa_k_fold_2expp1):
// the cached "transform" is that plan's forward transform of y, so a cached
// product costs one forward transform, the pointwise product, one inverse
// and the CRT.  The reference's negacyclic digit splitting (N = m b,
// twisted transforms of m points) is its CPU geometry; the digit count m
// is accepted and checked as the reference checks it, and kept in the
// context, but does not change the device plan.
#include "fa_internal.h"

#include <string.h>

int fa_fold_2expp1_launch(uint64_t* z, const uint64_t* P, uint64_t n, hipStream_t s);

// sd_fft_mpn_mulmod_2expp1_ctx_init (negacyclic.c:65-168): m = 0 lets the
// library choose; otherwise N = m b with m a power of two >= 16
extern "C" int fa_mulmod_2expp1_ctx_init(fa_mulmod_2expp1_ctx_struct* C, int64_t N, int64_t m)
{
    if (!C || N <= 0 || N % 64)
        return fa_set_error(FA_E_ARG, "fa_mulmod_2expp1_ctx_init: N must be a positive "
                            "multiple of 64");
    if (m != 0 && (m < 16 || (m & (m - 1)) || N % m))
        return fa_set_error(FA_E_ARG, "fa_mulmod_2expp1_ctx_init: need N = m b with m a power "
                            "of two >= 16");
    memset(C, 0, sizeof(*C));
    C->N = N;
    C->n = N / 64;
    C->m = m;
    return fa_plan_init_mpn(&C->P, (uint64_t)C->n + 1, (uint64_t)C->n + 1, 1, 0);
}

extern "C" void fa_mulmod_2expp1_ctx_clear(fa_mulmod_2expp1_ctx_struct* C)
{
    if (C) memset(C, 0, sizeof(*C));
}

// sd_fft_mpn_mulmod_2expp1_transformed_size (negacyclic.c:349-354): doubles
// of a cached transform
extern "C" int64_t fa_mulmod_2expp1_transformed_size(const fa_mulmod_2expp1_ctx_struct* C)
{
    return C ? (int64_t)(fa_op_sizeof_data(&C->P) / sizeof(double)) : 0;
}

// sd_fft_mpn_mulmod_2expp1_transform (negacyclic.c:356-362): F (device,
// transformed_size doubles) = the transform of y (device, n + 1 limbs)
extern "C" int fa_mulmod_2expp1_transform(const fa_mulmod_2expp1_ctx_struct* C, double* F,
                                          const uint64_t* y, void* stream)
{
    if (!C || !C->n || !F || !y)
        return fa_set_error(FA_E_ARG, "fa_mulmod_2expp1_transform: bad arguments");
    fa_op_struct Y;
    int rc = fa_op_init_borrowed(&Y, &C->P, F);
    if (rc) return rc;
    return fa_fft_mpn(&Y, y, C->n + 1, &C->P, stream);
}

// sd_fft_mpn_mulmod_2expp1_mul_cached (negacyclic.c:364-489): z = x y mod
// 2^N + 1 with y given by its cached transform F; z, x device, n + 1 limbs
extern "C" int fa_mulmod_2expp1_mul_cached(const fa_mulmod_2expp1_ctx_struct* C, uint64_t* z,
                                           const uint64_t* x, const double* F, void* stream)
{
    if (!C || !C->n || !z || !x || !F)
        return fa_set_error(FA_E_ARG, "fa_mulmod_2expp1_mul_cached: bad arguments");
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = stream ? (hipStream_t)stream : st->stream;
    const uint64_t n = (uint64_t)C->n;
    const size_t opb = fa_op_sizeof_data(&C->P);
    // the product's transform, then its 2n + 2 limbs, in the thread's block
    // buffer (the CRT keeps its flags in the scratch buffer)
    void* tb;
    int rc = fa_tbuf_reserve(st, opb + (2 * n + 2) * sizeof(uint64_t), &tb, s);
    if (rc) return rc;
    uint64_t* prod = (uint64_t*)((char*)tb + opb);
    fa_op_struct X, Y;
    if ((rc = fa_op_init_borrowed(&X, &C->P, tb)) ||
        (rc = fa_op_init_borrowed(&Y, &C->P, (void*)F)) ||
        (rc = fa_fft_mpn(&X, x, (int64_t)n + 1, &C->P, s)) ||
        (rc = fa_op_mul(&X, &X, &Y, &C->P, s)) || (rc = fa_ifft(&X, &C->P, s)) ||
        (rc = fa_export_mpn(prod, (int64_t)(2 * n + 2), &X, &C->P, s)) ||
        (rc = fa_fold_2expp1_launch(z, prod, n, s)))
        return rc;
    return fa_buffers_leave(st, s);
}

// sd_fft_mpn_mulmod_2expp1 with a context (negacyclic.c:491-501)
extern "C" int fa_mulmod_2expp1(const fa_mulmod_2expp1_ctx_struct* C, uint64_t* z,
                                const uint64_t* x, const uint64_t* y, void* stream)
{
    if (!C || !C->n) return fa_set_error(FA_E_ARG, "fa_mulmod_2expp1: bad context");
    return fa_mpn_mulmod_2expp1_device(z, x, y, C->n, stream);
}
