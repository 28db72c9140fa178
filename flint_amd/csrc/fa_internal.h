// fa_internal.h -- host-side structures shared by the flint_amd sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/flint_amd.h"

#define FA_NPRIMES 8
#define FA_MAX_DEPTH 24          // longest transform: 2^24 points
#define FA_ROW_LG 12             // row length of the two-level decomposition

// Residue arrays of a two-level transform (depth > 12) hold 2^(depth-12) rows
// of 4096 points at a pitch of FA_ROW_PITCH doubles.  The column passes read
// and write eight rows per wave instruction, one pitch apart; at a 32 KB
// pitch those accesses fall on the same few HBM channels, and the 1 KB pad
// spreads them (a column-pattern copy runs 1.5x faster with it).
// Coefficient j of such an array sits at (j >> 12) * FA_ROW_PITCH + (j & 4095);
// a one-level array (depth <= 12) is contiguous.  A two-level array's stride
// exceeds 4096 doubles, which is how fa_res_pos tells the two apart.
#ifndef FA_ROW_PITCH
#define FA_ROW_PITCH 4224
#endif
__host__ __device__ inline uint64_t fa_res_stride(int depth)
{
    return depth > FA_ROW_LG ? ((uint64_t)FA_ROW_PITCH << (depth - FA_ROW_LG))
                             : ((uint64_t)1 << depth);
}
__host__ __device__ inline uint64_t fa_res_pos(uint64_t j, uint64_t stride)
{
    return stride > ((uint64_t)1 << FA_ROW_LG)
               ? (j >> FA_ROW_LG) * FA_ROW_PITCH + (j & (((uint64_t)1 << FA_ROW_LG) - 1))
               : j;
}
#define FA_LDS_COL_DOUBLES 8192  // LDS budget of one column tile (64 KiB)
#define FA_CRT_MAXL 8            // limbs of M = prod p_i (np <= 8: < 400 bits)

// one of the 8 default primes with its device tables
struct fa_prime_host {
    uint64_t n;
    double p, pinv;
    uint64_t g;              // primitive 2^v-th root (sd_fft_ctx.c:32-37)
    int v;                   // 2-valuation of n-1
    double pow2[8];          // 2^32, 2^48, 2^96, 2^144, 2^192 mod n in [-n/2, n/2]
};

// device-wide tables (one set per device, shared by all host threads)
struct fa_tables {
    int device;
    int depth;                     // w2 holds 2^(depth-1) entries
    fa_prime_host pr[FA_NPRIMES];
    double* w2[FA_NPRIMES];        // w2[q] = w[2q]
    double* w2inv[FA_NPRIMES];     // w2[q]^-1
};

// CRT constants for np primes (mpn_mul.c:1093-1137)
struct fa_crt_consts {
    int np;
    int nl;                        // limbs of M
    double p[FA_NPRIMES];
    double pinv[FA_NPRIMES];
    double rp[FA_NPRIMES];         // 1/p_i for the quotient estimate
    uint64_t M[FA_CRT_MAXL];
    uint64_t cof[FA_NPRIMES][FA_CRT_MAXL];   // M / p_i
    // mixed radix (Garner): c = t_0 + p_0 (t_1 + p_1 (t_2 + ...)), t_k in [0, p_k)
    double gcm[FA_NPRIMES];                  // (M / p_k) mod p_k
    double ginv[FA_NPRIMES];                 // (p_0 ... p_(k-1))^-1 mod p_k
    double gpm[FA_NPRIMES][FA_NPRIMES];      // [k][j] = p_j mod p_k, j < k
};

// the kernel-side view of the per-prime tables
struct fa_prime_set {
    double p[FA_NPRIMES];
    double pinv[FA_NPRIMES];
    double norm[FA_NPRIMES];       // pointwise normalizer (cop*2^D)^-1
    double pow2[FA_NPRIMES][8];
    const double* w2[FA_NPRIMES];
    const double* w2inv[FA_NPRIMES];
};

// per-host-thread execution state (the analogue of the thread-local
// default mpn_ctx, default_ctx.c:14-15)
struct fa_thread_state {
    int device;
    hipStream_t stream;
    void* scratch;
    size_t scratch_size;
    void* hbuf;                    // device staging for host-pointer calls
    size_t hbuf_size;
    void* tbuf;                    // block products of integer products above one plan
    size_t tbuf_size;
    hipEvent_t last_ev;            // recorded after the last call's last use of the buffers
    hipStream_t last_stream;       // the stream it was recorded on (null: none pending)
};

// mpn plan chosen for the GPU
struct fa_mpn_plan_t {
    int np;
    int bits;            // multiple of 8
    int depth;           // transform length 2^depth
    int d1, d2;          // column / row split (d1 = 0: single LDS kernel)
    uint64_t alen, blen, zlen;
};

extern "C++" {
int fa_set_error(int code, const char* fmt, ...);
fa_tables* fa_get_tables(int depth);
const fa_tables* fa_tables_ready(int depth);
// one automatic-plan integer product in caller-provided device scratch of
// cap bytes; allocates nothing (the tables must already hold its depth)
int fa_mpn_mul_prealloc(uint64_t* z, const uint64_t* a, uint64_t an, const uint64_t* b,
                        uint64_t bn, hipStream_t stream, void* buf, size_t cap);
int fa_prime_info(uint64_t n, fa_prime_host* pr);
int fa_build_prime_tables(const fa_prime_host* pr, int depth, double** w2, double** w2inv);
// keeps a superseded device table allocated until the process ends
void fa_retire_device_table(void* p);
fa_thread_state* fa_state();
fa_thread_state* fa_state_new(int device);
void fa_state_free(fa_thread_state* st);
// makes fa_state() return st on this thread (null: the thread's own); returns the previous
fa_thread_state* fa_state_swap(fa_thread_state* st);
// order the state's buffers across streams: enter before the first enqueue
// of a call on stream s, leave after its last
int fa_buffers_enter(fa_thread_state* st, hipStream_t s);
int fa_buffers_leave(fa_thread_state* st, hipStream_t s);
// grow-only buffers of a state; s = the stream the caller uses them on
// (null: the state's own), which first waits for a call on another stream
int fa_scratch_reserve(fa_thread_state* st, size_t bytes, void** out, hipStream_t s);
int fa_hbuf_reserve(fa_thread_state* st, size_t bytes, void** out, hipStream_t s);
int fa_tbuf_reserve(fa_thread_state* st, size_t bytes, void** out, hipStream_t s);
void fa_crt_consts_init(fa_crt_consts* C, int np, int offset);

// the reference's integer windows from the exact product (fa_window.hip):
// work bytes for (an, bn, lo) under FLINT's plan on nthreads threads (0: no
// correction), and the correction of the window zw = [lo, hi) of P
size_t fa_window_work_bytes(uint64_t an, uint64_t bn, uint64_t lo, int nthreads);
// limbs of M np for the first np default primes (the reference's coeff_len)
int fa_ref_coeff_len(int np);
int fa_window_fix(uint64_t* zw, const uint64_t* P, uint64_t lo, uint64_t hi, const uint64_t* a,
                  uint64_t an, const uint64_t* b, uint64_t bn, int nthreads, void* work,
                  hipStream_t s);
uint64_t fa_crt_cop(int np, int i);         // (M/p_i) mod p_i
int fa_mpn_choose_plan(uint64_t an, uint64_t bn, fa_mpn_plan_t* P);
// automatic truncation of a 2^(12 + d1)-point transform whose product fills t rows
int fa_trunc_rule(int d1, int t);
int fa_mpn_fixed_plan(uint64_t an, uint64_t bn, int np, int bits, fa_mpn_plan_t* P);
int fa_mpn_max_bits(uint64_t an, uint64_t bn, int np);
void fa_prime_set_init(fa_prime_set* S, const fa_tables* T, int np, int depth, int offset,
                       int with_norm);
uint64_t fa_mulmod_u64(uint64_t a, uint64_t b, uint64_t n);
int fa_is_prime_u64(uint64_t n);                 // deterministic Miller-Rabin
uint64_t fa_powmod_u64(uint64_t a, uint64_t e, uint64_t n);
double fa_signed_rep(uint64_t x, uint64_t n);

// convolution core: residues of np primes, transforms, pointwise, inverse.
// Operands are given through a pack descriptor; the result (np rows of
// 2^depth doubles, coefficient order) is left in A.
struct fa_pack_desc {
    int kind;                  // FA_PACK_MPN, FA_PACK_U64, FA_PACK_S64W
    const uint64_t* src;
    uint64_t n;                // limbs (mpn) or coefficients
    int bits;                  // chunk bits (mpn)
    int width;                 // limbs per coefficient (signed coefficients)
};
enum { FA_PACK_MPN = 0, FA_PACK_U64 = 1, FA_PACK_S64W = 2 };
// kernel specializations of the import: a limb array cut at byte
// boundaries (bits = 8 own, own <= 25), and any kind chosen at run time
enum { FA_PACK_MPN8 = 3, FA_PACK_ANY = 4 };

int fa_convolve(fa_thread_state* st, const fa_tables* T, int np, int offset, int depth,
                const fa_pack_desc* a, const fa_pack_desc* b, double* A, double* B,
                hipStream_t stream);
}

// ---------------------------------------------------------------------------
// batched products: one device-side job record per product; the transform
// kernels of a bin (equal depth and import kind) walk a work list whose
// 64-bit items are  job | prime << 32 | operand << 36 | tile << 40
// (tile = column group or row)
// ---------------------------------------------------------------------------
struct fa_bjob {
    const uint64_t* a;         // operands (device)
    const uint64_t* b;
    uint64_t an, bn;           // limbs (mpn) or coefficients
    int aw, bw;                // limbs per coefficient (signed coefficients)
    int kind;                  // FA_PACK_*
    int bits;                  // chunk bits (mpn)
    double* A;                 // residue rows, np x stride each
    double* B;
    uint64_t stride;           // 2^depth
    int np, offset, depth, pad;
    int trows, pad2;           // rows of the product's coefficients (the row pass and the
                               // inverse column pass stop there)
    double norm[FA_NPRIMES];   // pointwise normalizer of prime offset + i
};

#define FA_WORK(job, prime, op, tile)                                                     \
    ((uint64_t)(job) | ((uint64_t)(prime) << 32) | ((uint64_t)(op) << 36) |              \
     ((uint64_t)(tile) << 40))

// per-product records of the batched CRT stages (device)
struct fa_bmpn {               // integer product
    const double* Q;           // residue rows (np x stride)
    uint64_t stride;
    uint64_t zlen;             // coefficients of the chunked product
    uint64_t* z;               // output limbs
    uint64_t zn;               // an + bn
    uint64_t ngroups, nblk;    // CRT groups of 8 coefficients, CRT workgroups
    uint8_t* flags;            // per group / per workgroup carry flags
    uint8_t* agg;
    int bits, np;
};
struct fa_bpoly {              // nmod or fmpz polynomial product
    const double* Q;           // residue rows; null: the zero product (n = 1)
    uint64_t stride;
    const void* C;             // fa_crt_consts of the prime set (device)
    uint64_t zl, cnt, zn;      // window [zl, zl + cnt) of a product of zn coefficients
    uint64_t n, dinv;          // nmod modulus and its preinverse
    int norm, zw;              // nmod shift, fmpz output limbs per coefficient
    uint64_t* z;
};

extern "C++" {
int fa_batch_crt_mpn(int np, const fa_bmpn* M, const uint64_t* work, unsigned nwork,
                     const void* F, hipStream_t s);
int fa_batch_carry(const fa_bmpn* M, const unsigned* jobs, unsigned njobs, hipStream_t s);
int fa_nmod_primes(const fa_tables* T, uint64_t len, uint64_t n, int* np, int* offset);
// transform depth of the window [zl, zh) of a zn-coefficient product, with
// the wraparound of plan.c:47-89 when it leaves the window exact
int fa_poly_window_depth(uint64_t zl, uint64_t zh, uint64_t zn, uint64_t xmax);
uint64_t fa_ks_slot_bits(int64_t b1, int64_t b2, uint64_t an, uint64_t bn);
// Kronecker route: variant 1 = signed slots, one integer product; 0 = three
// products of non-negative parts (fa_set_ks_variant)
#define FA_KS_PACK_BLK 1024      // packed limbs per workgroup of the signed-slot pack
struct fa_bks {                 // one signed-slot KS product of a batch
    const uint64_t* a;
    const uint64_t* b;
    uint64_t* z;
    uint64_t aw, an, bw, bn, s, La, Lb, nP, zw, zl, cnt, zn, top;
    int32_t* lowa;
    int32_t* lowb;
    uint8_t* bra;
    uint8_t* brb;
    uint8_t* tk;
    int32_t* sig;
    uint64_t* pa;
    uint64_t* pb;
    uint64_t* pr;
};
void fa_ks_record(uint64_t sb, const uint64_t* a, uint64_t aw, uint64_t an, const uint64_t* b,
                  uint64_t bw, uint64_t bn, uint64_t* scratch, uint64_t* z, uint64_t zw,
                  uint64_t zl, uint64_t cnt, fa_bks* r);
int fa_batch_ks_pre(const fa_bks* R, unsigned nks, const uint64_t* work, unsigned nwork,
                    hipStream_t s);
int fa_batch_ks_post(const fa_bks* R, unsigned nks, const uint64_t* work, unsigned nwork,
                     hipStream_t s);
int fa_ks_current_variant();
int fa_ks_nprod(int variant);
uint64_t fa_ks_scratch_words(int64_t b1, int64_t b2, uint64_t an, uint64_t bn);
uint64_t fa_ks_scratch_words_v(int variant, int64_t b1, int64_t b2, uint64_t an, uint64_t bn);
int fa_ks_products_v(int variant, uint64_t sb, uint64_t an, uint64_t bn, uint64_t* scratch,
                     uint64_t* ptrs[9], uint64_t lens[2]);
int fa_ks_pre_v(int variant, const uint64_t* a, uint64_t aw, uint64_t an, const uint64_t* b,
                uint64_t bw, uint64_t bn, uint64_t sb, uint64_t* scratch, hipStream_t s);
int fa_ks_post_v(int variant, uint64_t sb, uint64_t an, uint64_t bn, uint64_t* scratch,
                 uint64_t* z, uint64_t zw, uint64_t zl, uint64_t cnt, hipStream_t s);
int fa_ks_device(uint64_t* z, uint64_t zw, uint64_t zl, uint64_t cnt, const uint64_t* a,
                 uint64_t aw, uint64_t an, const uint64_t* b, uint64_t bw, uint64_t bn,
                 int64_t b1, int64_t b2, uint64_t* scratch, hipStream_t s);
int fa_fmpz_primes(uint64_t bn, int64_t b1, int64_t b2);
int64_t fa_vec_max_bits(const uint64_t* v, uint64_t w, int64_t n);
int fa_batch_crt_poly(int fmpz, const fa_bpoly* P, const uint64_t* work, unsigned nwork,
                      hipStream_t s);
// launches of one bin; return FA_OK or an error code
int fa_batch_small_bin(int depth, int kind, const fa_bjob* J, const uint64_t* work,
                       unsigned nwork, const fa_prime_set& S, hipStream_t s);
int fa_batch_transform_bin(int depth, int kind, const fa_bjob* J, const uint64_t* wfwd,
                           unsigned nfwd, const uint64_t* wrow, unsigned nrow,
                           const uint64_t* winv, unsigned ninv, const fa_prime_set& S,
                           hipStream_t s);
// column groups per residue array in the batched column passes at depth D
int fa_batch_col_groups(int depth);
}

// ---------------------------------------------------------------------------
// optional per-kernel timing with HIP events on the launching stream
// ---------------------------------------------------------------------------
extern "C++" {
int fa_prof_on();
void fa_prof_begin(const char* name, hipStream_t s);
void fa_prof_end(hipStream_t s);
}

#define FA_LAUNCH(name, kern, grid, block, lds, stream, ...)                          \
    do {                                                                              \
        if (fa_prof_on()) fa_prof_begin(name, stream);                                \
        hipLaunchKernelGGL(kern, grid, block, lds, stream, __VA_ARGS__);              \
        if (fa_prof_on()) fa_prof_end(stream);                                        \
    } while (0)
