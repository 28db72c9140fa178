// fa_op.hip -- plans and transformed operands (fft_small.h:530-700,
// plan.c:400-656, op.c:18-471, mpn_mul.c:2040-2310, nmod_poly_mul.c:752-901).
//
// A transformed operand ("op") holds the forward transforms of one operand
// modulo each prime of a plan, in device memory: np residue arrays of
// `stride` doubles in the layout of the product path (rows of 4096 points at
// a pitch of 4224 doubles above depth 12).  Ops are combined pointwise
// (mul, sqr, addmul, submul, add, sub, neg), inverse-transformed, and
// exported by the CRT of the product path, so transforms of an operand that
// enters many products are computed once, and sums of products are
// accumulated in transform space before one inverse transform and one CRT.
// The layout is this library's own; ops are not interchangeable with the
// reference's host buffers.
//
// Every function here enqueues on the given stream (NULL: the calling
// thread's) and returns without waiting; operand and output pointers are
// device addresses.
#include <string.h>

#include "fa_arith.h"
#include "fa_internal.h"
#include "fa_ntt.h"
#include "fa_pack.h"

int fa_forward_fast(int depth, const fa_operand& oa, int np, double* A, uint64_t stride,
                    const fa_prime_set& S, hipStream_t s);
int fa_inverse_fast(int depth, int np, double* A, uint64_t stride, const fa_prime_set& S,
                    hipStream_t s);
int fa_crt_mpn_export(int np, const double* A, uint64_t stride, uint64_t zlen, int bits,
                      uint64_t* z, uint64_t zn, hipStream_t stream, int mode, uint64_t jlo = 0);
int fa_signed_diff_launch(uint64_t* z, const uint64_t* w, uint64_t zn, int32_t* sign,
                          hipStream_t s);
int fa_signed_window_launch(uint64_t* z, uint64_t zn, const uint64_t* t, uint64_t lo, int up,
                            const int32_t* sign, hipStream_t s);
int fa_crt_nmod_launch(const double* A, uint64_t stride, const fa_crt_consts& C, uint64_t zl,
                       uint64_t cnt, uint64_t n, uint64_t* dz, hipStream_t s);
int fa_nmod_primes_bound(const fa_tables* T, uint64_t len_bound, uint64_t prod_bits, uint64_t n,
                         int* np, int* offset);

namespace {

uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

int clog2(uint64_t x)
{
    int k = 0;
    while ((UINT64_C(1) << k) < x) k++;
    return k;
}

// M of the np primes from offset 0 >= c * 2^e
bool capacity(int np, uint64_t c, uint64_t e)
{
    fa_crt_consts C;
    fa_crt_consts_init(&C, np, 0);
    uint64_t t[FA_CRT_MAXL + 4];
    memset(t, 0, sizeof(t));
    const uint64_t q = e / 64, r = e % 64;
    if (q + 2 > FA_CRT_MAXL + 3) return false;
    t[q] = c << r;
    if (r) t[q + 1] = c >> (64 - r);
    for (int i = FA_CRT_MAXL + 3; i >= 0; i--) {
        const uint64_t mi = i < FA_CRT_MAXL ? C.M[i] : 0;
        if (mi != t[i]) return mi > t[i];
    }
    return true;
}

bool compatible(const fa_op_struct* X, const fa_plan_struct* P)
{
    return X && P && X->data && X->np == (uint64_t)P->np && X->offset == (uint64_t)P->offset &&
           X->depth == (uint64_t)P->depth && X->stride == P->stride;
}

hipStream_t stream_of(void* s)
{
    if (s) return (hipStream_t)s;
    fa_thread_state* st = fa_state();
    return st ? st->stream : (hipStream_t)0;
}

const fa_tables* tables_for(const fa_plan_struct* P)
{
    return fa_get_tables(P->depth < 20 ? 20 : P->depth);
}

enum { OP_MUL, OP_SQR, OP_ADDMUL, OP_SUBMUL, OP_ADD, OP_SUB, OP_NEG };

}  // namespace

// ------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------

// forward transform of one operand at depth <= 12 (one LDS-resident array
// per prime; blockIdx.x = prime)
__global__ void __launch_bounds__(256) fa_k_op_small_fwd(fa_operand a, int depth,
                                                         double* __restrict__ out,
                                                         uint64_t stride, fa_prime_set S)
{
    extern __shared__ double lds[];
    const int i = blockIdx.x;
    const double n = S.p[i], ninv = S.pinv[i];
    const int N = 1 << depth;
    for (int x = threadIdx.x; x < N; x += blockDim.x)
        lds[x] = fa_operand_value(a, (uint64_t)x, S.pow2[i], n, ninv);
    __syncthreads();
    fa_lds_fwd(lds, depth, 0, 0, S.w2[i], n, ninv);
    double* o = out + (uint64_t)i * stride;
    for (int x = threadIdx.x; x < N; x += blockDim.x) o[x] = lds[x];
}

__global__ void __launch_bounds__(256) fa_k_op_small_inv(double* __restrict__ A, int depth,
                                                         uint64_t stride, fa_prime_set S)
{
    extern __shared__ double lds[];
    const int i = blockIdx.x;
    const double n = S.p[i], ninv = S.pinv[i];
    const int N = 1 << depth;
    double* o = A + (uint64_t)i * stride;
    for (int x = threadIdx.x; x < N; x += blockDim.x) lds[x] = o[x];
    __syncthreads();
    fa_lds_inv(lds, depth, 0, 0, S.w2inv[i], n, ninv);
    for (int x = threadIdx.x; x < N; x += blockDim.x) o[x] = lds[x];
}

// pointwise operations over every transform point of np primes (op.c:92-271):
// products carry the normalizer m once (z = a b m), the containment step
// multiplies by m first; additive steps reduce back into [-n, n]
__global__ void __launch_bounds__(256) fa_k_op_point(int op, double* __restrict__ Z,
                                                     const double* __restrict__ A,
                                                     const double* __restrict__ B, int np,
                                                     int depth, uint64_t stride, fa_prime_set S)
{
    const uint64_t npts = UINT64_C(1) << depth;
    const uint64_t total = npts * (uint64_t)np;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const int i = (int)(g >> depth);
        const uint64_t x = g & (npts - 1);
        const uint64_t at = (uint64_t)i * stride + fa_res_pos(x, stride);
        const double n = S.p[i], ninv = S.pinv[i], m = S.norm[i];
        const double a = A[at];
        double r;
        switch (op) {
        case OP_MUL:
            r = fa_mulmod(fa_mulmod(a, m, n, ninv), B[at], n, ninv);
            break;
        case OP_SQR:
            r = fa_mulmod(fa_mulmod(a, m, n, ninv), a, n, ninv);
            break;
        case OP_ADDMUL:
        case OP_SUBMUL: {
            const double t = fa_mulmod(fa_mulmod(a, m, n, ninv), B[at], n, ninv);
            r = fa_reduce_pm1n(op == OP_ADDMUL ? Z[at] + t : Z[at] - t, n, ninv);
            break;
        }
        case OP_ADD: r = fa_reduce_pm1n(a + B[at], n, ninv); break;
        case OP_SUB: r = fa_reduce_pm1n(a - B[at], n, ninv); break;
        default: r = -a; break;
        }
        Z[at] = r;
    }
}

// ------------------------------------------------------------------------
// plans
// ------------------------------------------------------------------------

// the mpn plan over prime counts np_lo .. np_hi: among the exact (np, bits)
// -- len_bound * min(alen, blen) * 2^(2 bits) <= prod p_i, chunks a
// multiple of 8 bits -- the one with the fewest transform points np *
// 2^depth, then the fewest primes
static int plan_mpn(fa_plan_struct* P, uint64_t an_max, uint64_t bn_max, uint64_t len_bound,
                    int is_signed, int np_lo, int np_hi)
{
    if (!P || !an_max || !bn_max || !len_bound)
        return fa_set_error(FA_E_ARG, "fa_plan_init_mpn: bad arguments");
    // the centred lift of the signed exports needs one spare factor of two
    // (plan.c:443-446)
    if (is_signed) {
        if (len_bound > UINT64_MAX / 2)
            return fa_set_error(FA_E_RANGE, "fa_plan_init_mpn: len_bound too large");
        len_bound *= 2;
    }
    memset(P, 0, sizeof(*P));
    uint64_t best_cost = UINT64_MAX;
    for (int np = np_lo; np <= np_hi; np++) {
        fa_crt_consts C;
        fa_crt_consts_init(&C, np, 0);
        for (int bits = 200; bits >= 8; bits -= 8) {
            if (9 * bits < 64 * C.nl + 64) break;       // the CRT's limb layout
            const uint64_t alen = cdiv(64 * an_max, bits), blen = cdiv(64 * bn_max, bits);
            const uint64_t mlen = alen < blen ? alen : blen;
            if (mlen > UINT64_MAX / len_bound) continue;
            if (!capacity(np, mlen * len_bound, 2 * (uint64_t)bits)) continue;
            int depth = clog2(alen + blen - 1);
            if (depth < 4) depth = 4;
            if (depth > FA_MAX_DEPTH) break;
            const uint64_t cost = (uint64_t)np << depth;
            if (cost < best_cost) {
                best_cost = cost;
                P->np = np;
                P->bits = bits;
                P->depth = depth;
                P->zn = alen + blen - 1;
            }
            break;                                      // the largest exact chunk for this np
        }
    }
    if (!P->np)
        return fa_set_error(FA_E_RANGE, "no plan for %llu x %llu limbs, %llu accumulated, "
                            "primes %d..%d", (unsigned long long)an_max,
                            (unsigned long long)bn_max, (unsigned long long)len_bound, np_lo,
                            np_hi);
    P->kind = 0;
    P->offset = 0;
    P->stride = fa_res_stride(P->depth);
    P->zl = 0;
    P->zh = P->zn;
    P->an_max = an_max;
    P->bn_max = bn_max;
    P->len_bound = len_bound;
    P->pad = is_signed ? 1 : 0;         // signed: exports centre the coefficients
    return FA_OK;
}

// fft_small_plan_init_mpn (plan.c:443-459): products of naturals of at most
// an_max x bn_max limbs, with up to len_bound such products accumulated in
// transform space
extern "C" int fa_plan_init_mpn(fa_plan_struct* P, uint64_t an_max, uint64_t bn_max,
                                uint64_t len_bound, int is_signed)
{
    return plan_mpn(P, an_max, bn_max, len_bound, is_signed, 2, FA_NPRIMES);
}

// fft_small_plan_init_mpn_np (plan.c:482-495): np = 0 the automatic choice,
// np = 2 the two-prime plan (failing when it is not exact), any other value
// a plan of three or more primes
extern "C" int fa_plan_init_mpn_np(fa_plan_struct* P, uint64_t an_max, uint64_t bn_max,
                                   uint64_t len_bound, int is_signed, uint64_t np)
{
    if (np == 0) return plan_mpn(P, an_max, bn_max, len_bound, is_signed, 2, FA_NPRIMES);
    if (np == 2) return plan_mpn(P, an_max, bn_max, len_bound, is_signed, 2, 2);
    return plan_mpn(P, an_max, bn_max, len_bound, is_signed, 3, FA_NPRIMES);
}

// fft_small_plan_init_nmod (plan.c:400-417): the window [zl, zh) of a
// convolution of length zn over Z/nZ, at most len_bound products of
// prod_bits-bit values accumulating on one coefficient
extern "C" int fa_plan_init_nmod(fa_plan_struct* P, uint64_t zl, uint64_t zh, uint64_t zn,
                                 uint64_t xtrunc_max, uint64_t len_bound, uint64_t prod_bits,
                                 fa_nmod_t mod)
{
    if (!P || zl > zh || zh > zn || !zn || !len_bound || mod.n < 2 || xtrunc_max > zn)
        return fa_set_error(FA_E_ARG, "fa_plan_init_nmod: bad arguments");
    memset(P, 0, sizeof(*P));
    int depth = clog2(zn);
    if (depth < 4) depth = 4;
    if (depth > FA_MAX_DEPTH)
        return fa_set_error(FA_E_RANGE, "fa_plan_init_nmod: %llu coefficients need depth %d",
                            (unsigned long long)zn, depth);
    const fa_tables* T = fa_get_tables(depth < 20 ? 20 : depth);
    if (!T) return FA_E_NODEV;
    int np, offset;
    int rc = fa_nmod_primes_bound(T, len_bound, prod_bits, mod.n, &np, &offset);
    if (rc) return rc;
    P->kind = 1;
    P->np = np;
    P->offset = offset;
    P->depth = depth;
    P->stride = fa_res_stride(depth);
    P->zl = zl;
    P->zh = zh;
    P->zn = zn;
    P->len_bound = len_bound;
    P->an_max = xtrunc_max;
    P->mod = mod;
    return FA_OK;
}

// fft_small_plan_init_nmod_cyclic (plan.c:425-441): convolutions modulo
// x^N - 1, N = 2^depth, output coefficients [0, zh); operands of up to N
// coefficients (a full cyclic convolution accumulates N products, so a
// plain product passes len_bound = N)
extern "C" int fa_plan_init_nmod_cyclic(fa_plan_struct* P, uint64_t depth, uint64_t zh,
                                        uint64_t len_bound, uint64_t prod_bits, fa_nmod_t mod)
{
    if (!P || depth < 4 || depth > FA_MAX_DEPTH || zh > (UINT64_C(1) << depth))
        return fa_set_error(FA_E_ARG, "fa_plan_init_nmod_cyclic: bad arguments");
    const uint64_t N = UINT64_C(1) << depth;
    return fa_plan_init_nmod(P, 0, zh, N, N, len_bound, prod_bits, mod);
}

// fft_small_plan_init_mpn_cyclic (plan.c:497-647): products of naturals of
// at most *nn limbs modulo 2^(64 *nn) - 1.  The chunks tile the operand
// exactly, bits 2^depth = 64 nn, so the coefficient convolution's wrap mod
// x^(2^depth) - 1 is the integer wrap mod B^nn - 1; *nn is rounded up to
// the smallest such length whose plan is exact for len_bound 2^depth chunk
// products (bits a multiple of 8 here), ties to fewer primes, and written
// back.  fa_export_mpn then gives a representative below len_bound
// 2^(64 nn + 2 bits + depth), which the caller folds.
extern "C" int fa_plan_init_mpn_cyclic(fa_plan_struct* P, uint64_t* nn, uint64_t len_bound)
{
    if (!P || !nn || !*nn || !len_bound)
        return fa_set_error(FA_E_ARG, "fa_plan_init_mpn_cyclic: bad arguments");
    uint64_t best_nn = 0;
    int best_np = 0, best_bits = 0, best_depth = 0;
    for (int np = 2; np <= FA_NPRIMES; np++) {
        fa_crt_consts C;
        fa_crt_consts_init(&C, np, 0);
        for (int bits = 8; bits <= 200; bits += 8) {
            if (9 * bits < 64 * C.nl + 64) continue;   // the CRT's limb layout
            int depth = 4;
            while (depth <= FA_MAX_DEPTH && ((uint64_t)bits << depth) < 64 * *nn) depth++;
            if (depth > FA_MAX_DEPTH) continue;
            const uint64_t N = UINT64_C(1) << depth;
            if (len_bound > UINT64_MAX / N || !capacity(np, len_bound * N, 2 * (uint64_t)bits))
                continue;
            const uint64_t nn1 = ((uint64_t)bits << depth) / 64;
            if (!best_np || nn1 < best_nn) {
                best_nn = nn1;
                best_np = np;
                best_bits = bits;
                best_depth = depth;
            }
        }
    }
    if (!best_np)
        return fa_set_error(FA_E_RANGE, "no cyclic plan for %llu limbs, %llu accumulated",
                            (unsigned long long)*nn, (unsigned long long)len_bound);
    memset(P, 0, sizeof(*P));
    P->kind = 0;
    P->np = best_np;
    P->offset = 0;
    P->bits = best_bits;
    P->depth = best_depth;
    P->stride = fa_res_stride(best_depth);
    P->zn = UINT64_C(1) << best_depth;
    P->zl = 0;
    P->zh = P->zn;
    P->an_max = best_nn;
    P->bn_max = best_nn;
    P->len_bound = len_bound;
    *nn = best_nn;
    return FA_OK;
}

extern "C" void fa_plan_clear(fa_plan_struct* P)
{
    if (P) memset(P, 0, sizeof(*P));
}

// ------------------------------------------------------------------------
// ops
// ------------------------------------------------------------------------

extern "C" uint64_t fa_op_sizeof_data(const fa_plan_struct* P)
{
    return P ? (uint64_t)P->np * P->stride * sizeof(double) : 0;
}

static void op_fields(fa_op_struct* X, const fa_plan_struct* P)
{
    X->stride = P->stride;
    X->np = (uint64_t)P->np;
    X->offset = (uint64_t)P->offset;
    X->depth = (uint64_t)P->depth;
    X->domain = FA_OP_PRIMAL;
}

extern "C" int fa_op_init(fa_op_struct* X, const fa_plan_struct* P)
{
    if (!X || !P || !P->np) return fa_set_error(FA_E_ARG, "fa_op_init: bad arguments");
    memset(X, 0, sizeof(*X));
    op_fields(X, P);
    void* d = NULL;
    if (hipMalloc(&d, fa_op_sizeof_data(P)) != hipSuccess) {
        (void)hipGetLastError();
        return fa_set_error(FA_E_NOMEM, "fa_op_init: device allocation of %llu bytes failed",
                            (unsigned long long)fa_op_sizeof_data(P));
    }
    X->data = (double*)d;
    X->owns_data = 1;
    return FA_OK;
}

extern "C" int fa_op_init_borrowed(fa_op_struct* X, const fa_plan_struct* P, void* data)
{
    if (!X || !P || !P->np || !data)
        return fa_set_error(FA_E_ARG, "fa_op_init_borrowed: bad arguments");
    memset(X, 0, sizeof(*X));
    op_fields(X, P);
    X->data = (double*)data;
    X->owns_data = 0;
    return FA_OK;
}

extern "C" void fa_op_clear(fa_op_struct* X)
{
    if (!X) return;
    if (X->owns_data && X->data) {
        (void)hipDeviceSynchronize();   // launches in flight may still use it
        (void)hipFree(X->data);
    }
    memset(X, 0, sizeof(*X));
}

static int forward(fa_op_struct* X, const fa_operand& op, const fa_plan_struct* P, void* stream)
{
    const fa_tables* T = tables_for(P);
    if (!T) return FA_E_NODEV;
    hipStream_t s = stream_of(stream);
    fa_prime_set S;
    fa_prime_set_init(&S, T, P->np, P->depth, P->offset, 1);
    if (P->depth <= FA_ROW_LG) {
        FA_LAUNCH("fa_k_op_small_fwd", fa_k_op_small_fwd, dim3(P->np), dim3(256),
                  ((size_t)1 << P->depth) * sizeof(double), s, op, P->depth, X->data, X->stride,
                  S);
    } else if (!fa_forward_fast(P->depth, op, P->np, X->data, X->stride, S, s)) {
        return fa_set_error(FA_E_RANGE, "forward transform: depth %d", P->depth);
    }
    X->domain = FA_OP_PRIMAL;
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "forward launch");
}

// fft_small_fft_mpn (mpn_mul.c:2040-2100): an-limb natural at device address a
extern "C" int fa_fft_mpn(fa_op_struct* X, const uint64_t* a, int64_t an,
                          const fa_plan_struct* P, void* stream)
{
    const uint64_t amax = P->an_max > P->bn_max ? P->an_max : P->bn_max;
    if (!compatible(X, P) || P->kind != 0 || !a || an <= 0 || (uint64_t)an > amax)
        return fa_set_error(FA_E_ARG, "fa_fft_mpn: operand or op does not fit the plan");
    fa_operand op;
    op.kind = FA_PACK_MPN;
    op.src = a;
    op.n = (uint64_t)an;
    op.bits = P->bits;
    op.width = 1;
    return forward(X, op, P, stream);
}

// fft_small_fft_nmod (nmod_poly_mul.c:752-800): an reduced coefficients at
// device address a, zero-padded to the transform length
extern "C" int fa_fft_nmod(fa_op_struct* X, const uint64_t* a, int64_t an, fa_nmod_t mod,
                           const fa_plan_struct* P, void* stream)
{
    if (!compatible(X, P) || P->kind != 1 || !a || an <= 0 || mod.n != P->mod.n ||
        (uint64_t)an > ((uint64_t)1 << P->depth))
        return fa_set_error(FA_E_ARG, "fa_fft_nmod: operand or op does not fit the plan");
    fa_operand op;
    op.kind = FA_PACK_U64;
    op.src = a;
    op.n = (uint64_t)an;
    op.bits = 0;
    op.width = 1;
    return forward(X, op, P, stream);
}

static int point(int what, fa_op_struct* Z, const fa_op_struct* A, const fa_op_struct* B,
                 const fa_plan_struct* P, void* stream)
{
    if (!compatible(Z, P) || !compatible(A, P) || (B && !compatible(B, P)))
        return fa_set_error(FA_E_ARG, "pointwise op: operands do not match the plan");
    const bool product = what <= OP_SUBMUL;
    if (product && (A->domain != FA_OP_PRIMAL || (B && B->domain != FA_OP_PRIMAL)))
        return fa_set_error(FA_E_ARG, "pointwise product of an op that is already a product");
    if ((what == OP_ADDMUL || what == OP_SUBMUL) && Z->domain != FA_OP_PRODUCT)
        return fa_set_error(FA_E_ARG, "addmul/submul into an op that is not a product");
    if ((what == OP_ADD || what == OP_SUB) && A->domain != B->domain)
        return fa_set_error(FA_E_ARG, "add/sub of ops in different domains");
    const fa_tables* T = tables_for(P);
    if (!T) return FA_E_NODEV;
    fa_prime_set S;
    fa_prime_set_init(&S, T, P->np, P->depth, P->offset, 1);
    const uint64_t total = (uint64_t)P->np << P->depth;
    uint64_t grid = (total + 255) / 256;
    if (grid > 8192) grid = 8192;
    FA_LAUNCH("fa_k_op_point", fa_k_op_point, dim3((unsigned)grid), dim3(256), 0,
              stream_of(stream), what, Z->data, A->data, B ? B->data : NULL, P->np, P->depth,
              P->stride, S);
    const int dom = product ? FA_OP_PRODUCT : A->domain;
    Z->domain = dom;
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "pointwise launch");
}

extern "C" int fa_op_mul(fa_op_struct* Z, const fa_op_struct* A, const fa_op_struct* B,
                         const fa_plan_struct* P, void* stream)
{
    return point(OP_MUL, Z, A, B, P, stream);
}
extern "C" int fa_op_sqr(fa_op_struct* Z, const fa_op_struct* A, const fa_plan_struct* P,
                         void* stream)
{
    return point(OP_SQR, Z, A, NULL, P, stream);
}
extern "C" int fa_op_addmul(fa_op_struct* Z, const fa_op_struct* A, const fa_op_struct* B,
                            const fa_plan_struct* P, void* stream)
{
    return point(OP_ADDMUL, Z, A, B, P, stream);
}
extern "C" int fa_op_submul(fa_op_struct* Z, const fa_op_struct* A, const fa_op_struct* B,
                            const fa_plan_struct* P, void* stream)
{
    return point(OP_SUBMUL, Z, A, B, P, stream);
}
extern "C" int fa_op_add(fa_op_struct* Z, const fa_op_struct* A, const fa_op_struct* B,
                         const fa_plan_struct* P, void* stream)
{
    return point(OP_ADD, Z, A, B, P, stream);
}
extern "C" int fa_op_sub(fa_op_struct* Z, const fa_op_struct* A, const fa_op_struct* B,
                         const fa_plan_struct* P, void* stream)
{
    return point(OP_SUB, Z, A, B, P, stream);
}
extern "C" int fa_op_neg(fa_op_struct* Z, const fa_op_struct* A, const fa_plan_struct* P,
                         void* stream)
{
    return point(OP_NEG, Z, A, NULL, P, stream);
}

// fft_small_ifft (op.c:420-471): in place; the normalizer was applied by the
// pointwise product
extern "C" int fa_ifft(fa_op_struct* X, const fa_plan_struct* P, void* stream)
{
    if (!compatible(X, P) || X->domain != FA_OP_PRODUCT)
        return fa_set_error(FA_E_ARG, "fa_ifft: op is not a pointwise product of the plan");
    const fa_tables* T = tables_for(P);
    if (!T) return FA_E_NODEV;
    hipStream_t s = stream_of(stream);
    fa_prime_set S;
    fa_prime_set_init(&S, T, P->np, P->depth, P->offset, 1);
    if (P->depth <= FA_ROW_LG) {
        FA_LAUNCH("fa_k_op_small_inv", fa_k_op_small_inv, dim3(P->np), dim3(256),
                  ((size_t)1 << P->depth) * sizeof(double), s, X->data, P->depth, X->stride, S);
    } else if (!fa_inverse_fast(P->depth, P->np, X->data, X->stride, S, s)) {
        return fa_set_error(FA_E_RANGE, "inverse transform: depth %d", P->depth);
    }
    X->domain = FA_OP_COEFF;
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "inverse launch");
}

// fft_small_export_mpn (mpn_mul.c:2240-2310): the low zn limbs of the
// accumulated integer sum_i c_i 2^(bits i) into z (device)
extern "C" int fa_export_mpn(uint64_t* z, int64_t zn, const fa_op_struct* X,
                             const fa_plan_struct* P, void* stream)
{
    if (!compatible(X, P) || P->kind != 0 || X->domain != FA_OP_COEFF || !z || zn <= 0)
        return fa_set_error(FA_E_ARG, "fa_export_mpn: op not inverse-transformed, or bad output");
    return fa_crt_mpn_export(P->np, X->data, X->stride, P->zn, P->bits, z, (uint64_t)zn,
                             stream_of(stream), 0);
}

// fft_small_export_mpn_signed_trunc (export_signed.c:367-606): the first
// nslots coefficients, each centred to (-M/2, M/2], summed at their bit
// offsets; z[0, zn) = limbs [lo, lo + zn) of |value|, *sign (device int32)
// = 1 when the value is negative; 64 (lo + zn) >= nslots bits + 64 n + 2
// with n the limbs of M np, as the reference asserts.  lo > 0 keeps the
// reference's truncation: only slots from j0 = ceil((64 lo - 64 n - 1) /
// bits) on are summed (rounded down to a multiple of 256 on plans of more
// than two primes), and on those plans a negative value's window is the
// ceiling of |value| / 2^(64 lo) (the window of ~W + 1), on two-prime plans
// its floor.  The positive and the negated negative parts are two CRT
// passes of the product path; one workgroup subtracts them, another cuts
// the window.
extern "C" int fa_export_mpn_signed_trunc(uint64_t* z, int64_t zn, int32_t* sign,
                                          const fa_op_struct* X, uint64_t nslots, uint64_t lo,
                                          const fa_plan_struct* P, void* stream)
{
    if (!compatible(X, P) || P->kind != 0 || X->domain != FA_OP_COEFF || !z || zn <= 0 ||
        !sign || nslots > P->zn)
        return fa_set_error(FA_E_ARG, "fa_export_mpn_signed: bad arguments");
    if (!P->pad)
        return fa_set_error(FA_E_ARG, "fa_export_mpn_signed: the plan is not signed");
    const uint64_t n = (uint64_t)fa_ref_coeff_len(P->np);
    if (64 * (lo + (uint64_t)zn) < nslots * (uint64_t)P->bits + 64 * n + 2)
        return fa_set_error(FA_E_ARG, "fa_export_mpn_signed: %lld limbs from limb %llu cannot "
                            "hold %llu slots of %d bits", (long long)zn, (unsigned long long)lo,
                            (unsigned long long)nslots, P->bits);
    uint64_t j0 = 0;
    if (64 * lo > 64 * n + 1) j0 = cdiv(64 * lo - 64 * n - 1, (uint64_t)P->bits);
    if (j0 > nslots) j0 = nslots;
    if (P->np > 2) j0 &= ~UINT64_C(255);
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    hipStream_t s = stream_of(stream);
    const uint64_t L = lo + (uint64_t)zn;
    // the negative part (and, truncated, the whole magnitude) goes to the
    // thread's block buffer, kept apart from the CRT's scratch
    void* tb;
    int rc = fa_tbuf_reserve(st, (size_t)(lo ? 2 * L : L) * 8, &tb, s);
    if (rc) return rc;
    uint64_t* w = (uint64_t*)tb;
    uint64_t* t = lo ? w + L : z;
    rc = fa_crt_mpn_export(P->np, X->data, X->stride, nslots, P->bits, t, L, s, 1, j0);
    if (rc) return rc;
    rc = fa_crt_mpn_export(P->np, X->data, X->stride, nslots, P->bits, w, L, s, 2, j0);
    if (rc) return rc;
    rc = fa_signed_diff_launch(t, w, L, sign, s);
    if (rc) return rc;
    if (lo) {
        rc = fa_signed_window_launch(z, (uint64_t)zn, t, lo, P->np > 2, sign, s);
        if (rc) return rc;
    }
    return fa_buffers_leave(st, s);
}

// fft_small_export_mpn_signed (export_signed.c:608-612): the same with lo = 0
extern "C" int fa_export_mpn_signed(uint64_t* z, int64_t zn, int32_t* sign,
                                    const fa_op_struct* X, uint64_t nslots,
                                    const fa_plan_struct* P, void* stream)
{
    return fa_export_mpn_signed_trunc(z, zn, sign, X, nslots, 0, P, stream);
}

// fft_small_export_nmod_range (nmod_poly_mul.c:860-901): coefficients
// [zl, zh) mod n into z (device, zh - zl values)
extern "C" int fa_export_nmod_range(uint64_t* z, const fa_op_struct* X, uint64_t zl,
                                    uint64_t zh, fa_nmod_t mod, const fa_plan_struct* P,
                                    void* stream)
{
    if (!compatible(X, P) || P->kind != 1 || X->domain != FA_OP_COEFF || !z || zl > zh ||
        zh > P->zn || mod.n != P->mod.n)
        return fa_set_error(FA_E_ARG, "fa_export_nmod_range: bad window, op or modulus");
    fa_crt_consts C;
    fa_crt_consts_init(&C, P->np, P->offset);
    return fa_crt_nmod_launch(X->data, X->stride, C, zl, zh - zl, mod.n, z, stream_of(stream));
}

// fft_small_export_nmod: the plan's window [zl, zh)
extern "C" int fa_export_nmod(uint64_t* z, const fa_op_struct* X, fa_nmod_t mod,
                              const fa_plan_struct* P, void* stream)
{
    return fa_export_nmod_range(z, X, P ? P->zl : 0, P ? P->zh : 0, mod, P, stream);
}
