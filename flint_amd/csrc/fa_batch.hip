// fa_batch.hip -- batches of heterogeneous products (integer, nmod, fmpz)
// executed as one recorded launch sequence.
//
// fa_batch_create plans every request (the same plans as the single-product
// entry points), lays out one device allocation for all residue rows, job
// records and work lists, and bins the products:
//   - depth <= 12: one workgroup per (product, prime) does the whole
//     convolution in LDS, binned by (depth, import kind);
//   - depth 13..24: the column / row / column passes of the two-level
//     transform, binned by (depth, import kind), every workgroup taking one
//     (product, prime, operand, tile) item of its bin's work list;
//   - CRT: integer products binned by prime count (then one carry
//     workgroup per product), nmod and fmpz windows in 256-coefficient chunks.
// The launch sequence is captured once into a hipGraph (FA_BATCH_GRAPH) and
// replayed by fa_batch_launch.  No memset node is recorded: every output
// word is written by a kernel (outputs past the product are written as 0).
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <utility>
#include <vector>

#include "fa_internal.h"

extern int fa_trunc_enabled;
#include "fa_crt_fp.h"
#include "../../include/flint_amd.h"

namespace {

struct copy_op {
    void* dst;
    const void* src;
    size_t bytes;
};

struct bin_key {
    int depth, kind;
    bool operator<(const bin_key& o) const
    {
        return depth != o.depth ? depth < o.depth : kind < o.kind;
    }
};

struct tbin {
    std::vector<uint64_t> fwd, row, inv;
};

// device-offset bookkeeping: everything lives in one allocation
struct layout {
    size_t size = 0;
    size_t take(size_t bytes)
    {
        size_t off = size;
        size += (bytes + 255) & ~(size_t)255;
        return off;
    }
};

int clog2_u64(uint64_t x) { return x <= 1 ? 0 : 64 - __builtin_clzll(x - 1); }
int nbits_u64(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

}  // namespace

// per-thread grow-only buffers lent to one-shot host batches (fa_batch_mul):
// device memory and pinned staging survive across calls
namespace {
struct lent_buffers {
    char* dmem = nullptr;
    size_t dsize = 0;
    char* hin = nullptr;
    size_t hin_size = 0;
    char* hout = nullptr;
    size_t hout_size = 0;
};
thread_local lent_buffers lent;

int lend_device(size_t bytes, char** out)
{
    if (bytes > lent.dsize) {
        if (lent.dmem) (void)hipFree(lent.dmem);
        lent.dmem = nullptr;
        lent.dsize = 0;
        if (hipMalloc((void**)&lent.dmem, bytes) != hipSuccess) return FA_E_NOMEM;
        lent.dsize = bytes;
    }
    *out = lent.dmem;
    return FA_OK;
}

int lend_pinned(char** buf, size_t* have, size_t bytes, char** out)
{
    if (bytes > *have) {
        if (*buf) (void)hipHostFree(*buf);
        *buf = nullptr;
        *have = 0;
        if (hipHostMalloc((void**)buf, bytes, 0) != hipSuccess) return FA_E_NOMEM;
        *have = bytes;
    }
    *out = *buf;
    return FA_OK;
}
}  // namespace

#define FA_BATCH_LENT 0x100   // internal: borrow the per-thread buffers

struct fa_batch_s {
    int flags;
    int lent;                          // buffers borrowed from the thread
    int64_t nreq;
    fa_request* reqs;
    hipStream_t stream;
    char* dmem;
    size_t dsize;
    hipGraph_t graph;
    hipGraphExec_t exec;
    std::vector<copy_op> h2d, d2h;     // host-pointer batches: (offset, host) pairs
    char* din;                         // device staging of operands / outputs
    char* dout;
    size_t nin, nout;
    char* hin;                         // pinned host staging
    char* hout;
    // recorded launch parameters (device addresses inside dmem)
    fa_prime_set S;
    std::vector<std::pair<bin_key, std::pair<size_t, unsigned>>> small;   // key -> (work, n)
    struct tlaunch {
        bin_key key;
        size_t fwd, row, inv;
        unsigned nf, nr, ni;
    };
    std::vector<tlaunch> trans;
    size_t jobs;                                  // fa_bjob[]
    size_t mpn_rec, fp_consts[FA_NPRIMES + 1];    // fa_bmpn[], fa_crt_fp_consts per np
    std::vector<std::pair<int, std::pair<size_t, unsigned>>> crt_mpn;     // np -> (work, n)
    size_t carry_jobs;
    unsigned ncarry;
    size_t poly_rec;
    size_t nmod_work, fmpz_work;
    unsigned nnmod, nfmpz;
    // fmpz products routed through Kronecker substitution: packing before
    // the transforms, recombination after the integer CRT (device addresses)
    struct ksrec {
        const uint64_t* a;
        const uint64_t* b;
        uint64_t* z;
        uint64_t* scratch;
        uint64_t aw, an, bw, bn, sb, zw, zl, cnt;
        int variant;
    };
    std::vector<ksrec> ks;     // three-product route: launches per product
    // signed-slot route: records and work lists of all KS products, one
    // launch per step
    size_t ks_rec = 0, ks_packw = 0, ks_unpackw = 0;
    unsigned nks = 0, nks_pack = 0, nks_unpack = 0;
};

static int batch_record(fa_batch_s* B, hipStream_t s)
{
    char* d = B->dmem;
    const fa_bjob* J = (const fa_bjob*)(d + B->jobs);
    int rc;
    for (auto& k : B->ks) {
        rc = fa_ks_pre_v(k.variant, k.a, k.aw, k.an, k.b, k.bw, k.bn, k.sb, k.scratch, s);
        if (rc) return rc;
    }
    rc = fa_batch_ks_pre((const fa_bks*)(d + B->ks_rec), B->nks,
                         (const uint64_t*)(d + B->ks_packw), B->nks_pack, s);
    if (rc) return rc;
    for (auto& e : B->small) {
        rc = fa_batch_small_bin(e.first.depth, e.first.kind, J, (const uint64_t*)(d + e.second.first),
                                e.second.second, B->S, s);
        if (rc) return rc;
    }
    for (auto& t : B->trans) {
        rc = fa_batch_transform_bin(t.key.depth, t.key.kind, J, (const uint64_t*)(d + t.fwd), t.nf,
                                    (const uint64_t*)(d + t.row), t.nr,
                                    (const uint64_t*)(d + t.inv), t.ni, B->S, s);
        if (rc) return rc;
    }
    const fa_bmpn* M = (const fa_bmpn*)(d + B->mpn_rec);
    for (auto& c : B->crt_mpn) {
        rc = fa_batch_crt_mpn(c.first, M, (const uint64_t*)(d + c.second.first), c.second.second,
                              d + B->fp_consts[c.first], s);
        if (rc) return rc;
    }
    rc = fa_batch_carry(M, (const unsigned*)(d + B->carry_jobs), B->ncarry, s);
    if (rc) return rc;
    const fa_bpoly* P = (const fa_bpoly*)(d + B->poly_rec);
    rc = fa_batch_crt_poly(0, P, (const uint64_t*)(d + B->nmod_work), B->nnmod, s);
    if (rc) return rc;
    rc = fa_batch_crt_poly(1, P, (const uint64_t*)(d + B->fmpz_work), B->nfmpz, s);
    if (rc) return rc;
    for (auto& k : B->ks) {
        rc = fa_ks_post_v(k.variant, k.sb, k.an, k.bn, k.scratch, k.z, k.zw, k.zl, k.cnt, s);
        if (rc) return rc;
    }
    rc = fa_batch_ks_post((const fa_bks*)(d + B->ks_rec), B->nks,
                          (const uint64_t*)(d + B->ks_unpackw), B->nks_unpack, s);
    if (rc) return rc;
    return FA_OK;
}

extern "C" void fa_batch_destroy(fa_batch_s* B)
{
    if (!B) return;
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    if (!B->lent) {
        if (B->dmem) (void)hipFree(B->dmem);
        if (B->hin) (void)hipHostFree(B->hin);
        if (B->hout) (void)hipHostFree(B->hout);
    }
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
}

extern "C" int fa_batch_create(fa_batch_s** out, fa_request* R, int64_t nreq, int flags)
{
    if (!out || (nreq > 0 && !R) || nreq < 0)
        return fa_set_error(FA_E_ARG, "fa_batch_create: bad arguments");
    *out = nullptr;
    const bool dev = (flags & FA_BATCH_DEVICE) != 0;
    // the Kronecker route the whole batch uses, fixed at creation
    const int ks_variant = fa_ks_current_variant();
    const int ks_nprod = fa_ks_nprod(ks_variant);
    fa_batch_s* B = new fa_batch_s();
    B->flags = flags;
    B->nreq = nreq;
    B->reqs = R;
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
        delete B;
        return fa_set_error(FA_E_HIP, "fa_batch_create: stream");
    }

    // ---- plan every request; fmpz products declined by the eight-prime rule
    // become, under FA_BATCH_FMPZ_KS, three internal integer products
    // appended after the caller's requests
    struct plan {
        int live;           // has device work
        int np, offset, depth, bits, kind;
        uint64_t zn, zlen, cnt, ngroups, nblk;
        int64_t b1, b2;
        int ks;             // caller's fmpz request routed through KS: index into ksv
        int virt;           // internal product k of KS record virt_ks
        int virt_ks;
    };
    struct ksplan {
        int64_t r;
        uint64_t sb, La, Lb;
    };
    std::vector<ksplan> ksv;
    std::vector<fa_request> V;          // internal integer products
    std::vector<plan> P((size_t)nreq);
    int maxdepth = 4;
    auto req = [&](int64_t r) -> fa_request& { return r < nreq ? R[r] : V[(size_t)(r - nreq)]; };
    for (int64_t r = 0; r < nreq + (int64_t)V.size(); r++) {
        if (r >= (int64_t)P.size()) P.resize((size_t)r + 1);
        fa_request& q = req(r);
        plan& p = P[(size_t)r];
        const int virt = r >= nreq ? p.virt : 0, virt_ks = r >= nreq ? p.virt_ks : 0;
        memset(&p, 0, sizeof(p));
        p.virt = virt;
        p.virt_ks = virt_ks;
        p.ks = -1;
        q.status = 1;
        if (q.an <= 0 || q.bn <= 0 || !q.a || !q.b || !q.z) {
            q.status = FA_E_ARG;
            continue;
        }
        if (q.kind == FA_REQ_MPN) {
            fa_mpn_plan_t M;
            int rc = (q.np || q.bits) ? fa_mpn_fixed_plan(q.an, q.bn, q.np, q.bits, &M)
                                      : fa_mpn_choose_plan(q.an, q.bn, &M);
            if (rc) { q.status = rc; continue; }
            p.np = M.np; p.bits = M.bits; p.depth = M.depth; p.zlen = M.zlen;
            p.kind = FA_PACK_MPN8;
            p.zn = (uint64_t)(q.an + q.bn);
            const uint64_t own = (uint64_t)M.bits / 8;
            p.ngroups = (p.zn + own - 1) / own;
            p.nblk = (p.ngroups + 30) / 31;
            p.live = 1;
        } else if (q.kind == FA_REQ_NMOD || q.kind == FA_REQ_FMPZ) {
            if (q.zl < 0 || q.zh < q.zl) { q.status = FA_E_ARG; continue; }
            p.zn = (uint64_t)(q.an + q.bn - 1);
            p.cnt = (uint64_t)(q.zh - q.zl);
            if (p.cnt == 0) { p.live = 0; continue; }
            p.depth = fa_poly_window_depth((uint64_t)q.zl,
                                           (uint64_t)q.zh < p.zn ? (uint64_t)q.zh : p.zn, p.zn,
                                           (uint64_t)(q.an > q.bn ? q.an : q.bn));
            if (q.kind == FA_REQ_NMOD) {
                if (q.mod.n == 0) { q.status = FA_E_ARG; continue; }
                p.kind = FA_PACK_U64;
                if (q.mod.n == 1) {
                    p.np = 0;        // zero product: only the CRT stage writes zeros
                } else {
                    fa_tables* T = fa_get_tables(p.depth < 20 ? 20 : p.depth);
                    if (!T) { q.status = FA_E_NODEV; continue; }
                    int rc = fa_nmod_primes(T, (uint64_t)(q.an < q.bn ? q.an : q.bn), q.mod.n,
                                            &p.np, &p.offset);
                    if (rc) { q.status = rc; continue; }
                }
            } else {
                if (q.aw <= 0 || q.bw <= 0 || q.zw <= 0) { q.status = FA_E_ARG; continue; }
                p.kind = FA_PACK_S64W;
                p.b1 = q.abits;
                p.b2 = q.bbits;
                if (p.b1 == 0 && p.b2 == 0) {
                    if (dev) { q.status = FA_E_ARG; continue; }   // bounds needed up front
                    p.b1 = fa_vec_max_bits(q.a, (uint64_t)q.aw, q.an);
                    p.b2 = fa_vec_max_bits(q.b, (uint64_t)q.bw, q.bn);
                }
                p.np = fa_fmpz_primes((uint64_t)q.bn, p.b1, p.b2);
                if (!p.np && (flags & FA_BATCH_FMPZ_KS)) {
                    // Kronecker substitution: integer products of La x Lb limbs (one
                    // with signed slots, three with the non-negative parts)
                    ksplan K;
                    K.r = r;
                    K.sb = fa_ks_slot_bits(p.b1, p.b2, (uint64_t)q.an, (uint64_t)q.bn);
                    K.La = (q.an * K.sb + 63) / 64;
                    K.Lb = (q.bn * K.sb + 63) / 64;
                    p.ks = (int)ksv.size();
                    ksv.push_back(K);
                    for (int k = 0; k < ks_nprod; k++) {
                        fa_request v;
                        memset(&v, 0, sizeof(v));
                        v.kind = FA_REQ_MPN;
                        v.an = (int64_t)K.La;
                        v.bn = (int64_t)K.Lb;
                        v.a = v.b = (const uint64_t*)&V;     // resolved to the KS scratch
                        v.z = (uint64_t*)&V;
                        V.push_back(v);
                        plan pv;
                        memset(&pv, 0, sizeof(pv));
                        pv.virt = k + 1;
                        pv.virt_ks = p.ks;
                        P.push_back(pv);
                    }
                    continue;                            // status stays 1
                }
                if (!p.np) { q.status = 0; continue; }   // declined, as the reference does
            }
            p.live = 1;
        } else {
            q.status = FA_E_ARG;
            continue;
        }
        if (p.depth > FA_MAX_DEPTH) {
            q.status = fa_set_error(FA_E_RANGE, "batch request %lld: depth %d", (long long)r,
                                    p.depth);
            p.live = 0;
            continue;
        }
        q.np = p.np;
        q.bits = p.bits;
        q.depth = p.depth;
        if (p.depth > maxdepth) maxdepth = p.depth;
    }
    const int64_t ntot = nreq + (int64_t)V.size();
    for (auto& K : ksv) {
        // a failed internal product fails its fmpz request
        for (int k = 0; k < ks_nprod; k++) {
            const int64_t v = nreq + ks_nprod * (int64_t)(&K - ksv.data()) + k;
            if (req(v).status != 1) R[K.r].status = req(v).status;
        }
    }
    fa_tables* T = fa_get_tables(maxdepth < 20 ? 20 : maxdepth);
    if (!T) { fa_batch_destroy(B); return FA_E_NODEV; }
    fa_prime_set_init(&B->S, T, FA_NPRIMES, maxdepth, 0, 0);

    // ---- layout
    // two regions: the records, constants and work lists (uploaded), then the
    // residue rows, staging and carry flags (no initial contents)
    layout L, Lb, Lin, Lout;
    std::vector<fa_bjob> jobs;
    std::vector<fa_bmpn> mrec;
    std::vector<fa_bpoly> prec;
    std::vector<int> job_of((size_t)ntot, -1), rec_of((size_t)ntot, -1);
    struct offs { size_t A, B, a, b, z, flags, agg; };
    std::vector<offs> O((size_t)ntot);
    std::vector<size_t> ks_scratch(ksv.size());
    for (size_t k = 0; k < ksv.size(); k++) {
        const fa_request& q = R[ksv[k].r];
        const plan& p = P[(size_t)ksv[k].r];
        ks_scratch[k] = Lb.take(
            fa_ks_scratch_words_v(ks_variant, p.b1, p.b2, (uint64_t)q.an, (uint64_t)q.bn) * 8);
        if (!dev) {
            offs& o = O[(size_t)ksv[k].r];
            o.a = Lin.take((size_t)q.an * (size_t)q.aw * 8);
            o.b = Lin.take((size_t)q.bn * (size_t)q.bw * 8);
            o.z = Lout.take((size_t)(q.zh - q.zl) * (size_t)q.zw * 8);
        }
    }
    std::map<std::pair<int, int>, size_t> pconsts;   // (np, offset) -> device offset
    for (int64_t r = 0; r < ntot; r++) {
        fa_request& q = req(r);
        plan& p = P[(size_t)r];
        if (!p.live || q.status != 1) continue;
        offs& o = O[(size_t)r];
        const uint64_t stride = fa_res_stride(p.depth);
        if (p.np) {
            o.A = Lb.take((size_t)p.np * stride * sizeof(double));
            if (p.depth > FA_ROW_LG) o.B = Lb.take((size_t)p.np * stride * sizeof(double));
        }
        if (!dev && !p.virt) {
            const size_t aw = q.kind == FA_REQ_FMPZ ? (size_t)q.aw : 1;
            const size_t bw = q.kind == FA_REQ_FMPZ ? (size_t)q.bw : 1;
            o.a = Lin.take((size_t)q.an * aw * 8);
            o.b = Lin.take((size_t)q.bn * bw * 8);
        }
        size_t zwords = q.kind == FA_REQ_MPN ? (size_t)p.zn
                        : q.kind == FA_REQ_NMOD ? (size_t)p.cnt : (size_t)p.cnt * (size_t)q.zw;
        if (!dev && !p.virt) o.z = Lout.take(zwords * 8);
        if (q.kind == FA_REQ_MPN) {
            o.flags = Lb.take((size_t)p.ngroups + 8);
            o.agg = Lb.take((size_t)p.nblk + 8);
        }
        if (q.kind != FA_REQ_MPN && p.np && !pconsts.count({p.np, p.offset}))
            pconsts[{p.np, p.offset}] = L.take(sizeof(fa_crt_consts));
    }
    size_t fpc[FA_NPRIMES + 1];
    for (int k = 0; k <= FA_NPRIMES; k++) fpc[k] = (k >= 2) ? L.take(sizeof(fa_crt_fp_consts)) : 0;

    // job records and work lists (host images first; offsets resolved after
    // the allocation)
    std::map<int, std::vector<double>> norms;
    std::map<bin_key, std::vector<uint64_t>> small_bins;
    std::map<bin_key, tbin> trans_bins;
    std::map<int, std::vector<uint64_t>> crt_bins;
    std::vector<unsigned> carry;
    std::vector<uint64_t> nmod_w, fmpz_w;
    for (int64_t r = 0; r < ntot; r++) {
        fa_request& q = req(r);
        plan& p = P[(size_t)r];
        if (!p.live || q.status != 1) continue;
        if (p.np) {
            const unsigned jid = (unsigned)jobs.size();
            job_of[(size_t)r] = (int)jid;
            fa_bjob j;
            memset(&j, 0, sizeof(j));
            j.an = (uint64_t)q.an;
            j.bn = (uint64_t)q.bn;
            j.aw = q.kind == FA_REQ_FMPZ ? (int)q.aw : 1;
            j.bw = q.kind == FA_REQ_FMPZ ? (int)q.bw : 1;
            j.kind = p.kind == FA_PACK_MPN8 ? FA_PACK_MPN : p.kind;
            j.bits = p.bits;
            j.stride = fa_res_stride(p.depth);
            j.np = p.np;
            j.offset = p.offset;
            j.depth = p.depth;
            // truncation at row granularity below 2^24 points, for products
            // that end inside the transform (not wrapped ones)
            j.trows = p.depth > FA_ROW_LG ? 1 << (p.depth - FA_ROW_LG) : 0;
            {
                const uint64_t plen = q.kind == FA_REQ_MPN ? p.zlen : p.zn;
                if (fa_trunc_enabled && p.depth > FA_ROW_LG && p.depth < 2 * FA_ROW_LG && plen &&
                    plen <= ((uint64_t)1 << p.depth)) {
                    const int d1 = p.depth - FA_ROW_LG;
                    const int t = (int)((plen + (1u << FA_ROW_LG) - 1) >> FA_ROW_LG);
                    const int want = q.trunc == FA_TRUNC_ON ||
                                     (q.trunc == FA_TRUNC_AUTO &&
                                      (fa_trunc_enabled == 2 || fa_trunc_rule(d1, t)));
                    if (want) j.trows = t;
                }
            }
            // normalizers depend on (np, depth, offset) only
            const int nkey = (p.np * 32 + p.depth) * 16 + p.offset;
            auto it = norms.find(nkey);
            if (it == norms.end()) {
                fa_prime_set Sj;
                fa_prime_set_init(&Sj, T, p.np, p.depth, p.offset, 1);
                std::vector<double> v(Sj.norm, Sj.norm + FA_NPRIMES);
                it = norms.emplace(nkey, v).first;
            }
            for (int i = 0; i < p.np; i++) j.norm[i] = it->second[(size_t)i];
            jobs.push_back(j);
            const bin_key key{p.depth, p.kind};
            if (p.depth <= FA_ROW_LG) {
                auto& w = small_bins[key];
                for (int i = 0; i < p.np; i++) w.push_back(FA_WORK(jid, i, 0, 0));
            } else {
                auto& tb = trans_bins[key];
                const int ng = fa_batch_col_groups(p.depth);
                const int nrow = 1 << (p.depth - FA_ROW_LG);
                // the two column groups of a 128-byte line adjacent, then
                // the primes (the order of the single-product kernels)
                for (int z = 0; z < 2; z++)
                    for (int h = 0; h < ng / 2; h++)
                        for (int i = 0; i < p.np; i++)
                            for (int lo = 0; lo < 2; lo++)
                                tb.fwd.push_back(FA_WORK(jid, i, z, 2 * h + lo));
                for (int i = 0; i < p.np; i++)
                    for (int row = 0; row < nrow && row < j.trows; row++)
                        tb.row.push_back(FA_WORK(jid, i, 0, row));
                for (int h = 0; h < ng / 2; h++)
                    for (int i = 0; i < p.np; i++)
                        for (int lo = 0; lo < 2; lo++)
                            tb.inv.push_back(FA_WORK(jid, i, 0, 2 * h + lo));
            }
        }
        if (q.kind == FA_REQ_MPN) {
            const unsigned mid = (unsigned)mrec.size();
            rec_of[(size_t)r] = (int)mid;
            fa_bmpn m;
            memset(&m, 0, sizeof(m));
            m.stride = fa_res_stride(p.depth);
            m.zlen = p.zlen;
            m.zn = p.zn;
            m.ngroups = p.ngroups;
            m.nblk = p.nblk;
            m.bits = p.bits;
            m.np = p.np;
            mrec.push_back(m);
            auto& w = crt_bins[p.np];
            for (uint64_t blk = 0; blk < p.nblk; blk++) w.push_back((uint64_t)mid | (blk << 32));
            carry.push_back(mid);
        } else {
            const unsigned pid = (unsigned)prec.size();
            rec_of[(size_t)r] = (int)pid;
            fa_bpoly b;
            memset(&b, 0, sizeof(b));
            b.stride = fa_res_stride(p.depth);
            b.zl = (uint64_t)q.zl;
            b.cnt = p.cnt;
            b.zn = p.zn;
            if (q.kind == FA_REQ_NMOD) {
                const int mb = nbits_u64(q.mod.n);
                b.n = q.mod.n;
                b.norm = 64 - mb;
                const unsigned __int128 d = (unsigned __int128)(q.mod.n << b.norm);
                b.dinv = (uint64_t)((~(unsigned __int128)0) / d - ((unsigned __int128)1 << 64));
            } else {
                b.zw = (int)q.zw;
            }
            prec.push_back(b);
            auto& w = q.kind == FA_REQ_NMOD ? nmod_w : fmpz_w;
            for (uint64_t c = 0; c < (p.cnt + 255) / 256; c++) w.push_back((uint64_t)pid | (c << 32));
        }
    }
    B->jobs = L.take(jobs.size() * sizeof(fa_bjob) + 8);
    B->mpn_rec = L.take(mrec.size() * sizeof(fa_bmpn) + 8);
    B->poly_rec = L.take(prec.size() * sizeof(fa_bpoly) + 8);
    for (auto& e : small_bins) {
        size_t w = L.take(e.second.size() * 8);
        B->small.push_back({e.first, {w, (unsigned)e.second.size()}});
    }
    for (auto& e : trans_bins) {
        fa_batch_s::tlaunch t;
        t.key = e.first;
        t.fwd = L.take(e.second.fwd.size() * 8);
        t.row = L.take(e.second.row.size() * 8);
        t.inv = L.take(e.second.inv.size() * 8);
        t.nf = (unsigned)e.second.fwd.size();
        t.nr = (unsigned)e.second.row.size();
        t.ni = (unsigned)e.second.inv.size();
        B->trans.push_back(t);
    }
    for (auto& e : crt_bins) B->crt_mpn.push_back({e.first, {L.take(e.second.size() * 8),
                                                             (unsigned)e.second.size()}});
    B->carry_jobs = L.take(carry.size() * 4 + 8);
    B->ncarry = (unsigned)carry.size();
    // signed-slot KS products: record index k of ks_live, work items in
    // blocks of FA_KS_PACK_BLK packed limbs (per operand) and 256 output
    // coefficients
    std::vector<size_t> ks_live;
    std::vector<uint64_t> ks_pw, ks_uw;
    if (ks_variant) {
        for (size_t k = 0; k < ksv.size(); k++) {
            const fa_request& q = R[ksv[k].r];
            if (q.status != 1) continue;
            const uint64_t idx = ks_live.size();
            ks_live.push_back(k);
            for (uint64_t j = 0; j < (ksv[k].La + FA_KS_PACK_BLK - 1) / FA_KS_PACK_BLK; j++)
                ks_pw.push_back(idx | (j << 32));
            for (uint64_t j = 0; j < (ksv[k].Lb + FA_KS_PACK_BLK - 1) / FA_KS_PACK_BLK; j++)
                ks_pw.push_back(idx | (UINT64_C(1) << 31) | (j << 32));
            const uint64_t cnt = (uint64_t)(q.zh - q.zl);
            for (uint64_t j = 0; j < (cnt + 255) / 256; j++) ks_uw.push_back(idx | (j << 32));
        }
    }
    B->nks = (unsigned)ks_live.size();
    B->nks_pack = (unsigned)ks_pw.size();
    B->nks_unpack = (unsigned)ks_uw.size();
    B->ks_rec = L.take(ks_live.size() * sizeof(fa_bks) + 8);
    B->ks_packw = L.take(ks_pw.size() * 8 + 8);
    B->ks_unpackw = L.take(ks_uw.size() * 8 + 8);
    B->nmod_work = L.take(nmod_w.size() * 8 + 8);
    B->fmpz_work = L.take(fmpz_w.size() * 8 + 8);
    B->nnmod = (unsigned)nmod_w.size();
    B->nfmpz = (unsigned)fmpz_w.size();
    for (int k = 0; k <= FA_NPRIMES; k++) B->fp_consts[k] = fpc[k];

    const size_t big = (L.size + 255) & ~(size_t)255;
    const size_t inb = big + Lb.size, outb = inb + Lin.size;
    B->dsize = outb + Lout.size + 256;
    B->lent = (flags & FA_BATCH_LENT) != 0;
    if (B->lent ? lend_device(B->dsize, &B->dmem) != FA_OK
                : hipMalloc((void**)&B->dmem, B->dsize) != hipSuccess) {
        fa_batch_destroy(B);
        return fa_set_error(FA_E_NOMEM, "fa_batch_create: %zu bytes of device memory", B->dsize);
    }
    char* d = B->dmem;
    char* db0 = d + big;   // residue rows and carry flags
    char* din = d + inb;   // staged operands (host-pointer batches)
    char* dout = d + outb; // staged outputs

    // KS records: the caller's operands and output, the scratch
    for (size_t k = 0; k < ksv.size(); k++) {
        const fa_request& q = R[ksv[k].r];
        if (q.status != 1) continue;
        const offs& o = O[(size_t)ksv[k].r];
        fa_batch_s::ksrec K;
        K.a = dev ? q.a : (const uint64_t*)(din + o.a);
        K.b = dev ? q.b : (const uint64_t*)(din + o.b);
        K.z = dev ? q.z : (uint64_t*)(dout + o.z);
        K.scratch = (uint64_t*)(db0 + ks_scratch[k]);
        K.aw = (uint64_t)q.aw;
        K.an = (uint64_t)q.an;
        K.bw = (uint64_t)q.bw;
        K.bn = (uint64_t)q.bn;
        K.sb = ksv[k].sb;
        K.zw = (uint64_t)q.zw;
        K.zl = (uint64_t)q.zl;
        K.cnt = (uint64_t)(q.zh - q.zl);
        K.variant = ks_variant;
        if (!ks_variant) B->ks.push_back(K);
        if (!dev) {
            B->h2d.push_back({(void*)(uintptr_t)o.a, q.a, (size_t)(q.an * q.aw) * 8});
            B->h2d.push_back({(void*)(uintptr_t)o.b, q.b, (size_t)(q.bn * q.bw) * 8});
            B->d2h.push_back({q.z, (const void*)(uintptr_t)o.z, (size_t)(K.cnt * K.zw) * 8});
        }
    }

    // resolve device addresses into the records
    for (int64_t r = 0; r < ntot; r++) {
        fa_request& q = req(r);
        plan& p = P[(size_t)r];
        if (!p.live || q.status != 1) continue;
        offs& o = O[(size_t)r];
        const uint64_t* da = dev ? q.a : (const uint64_t*)(din + o.a);
        const uint64_t* db = dev ? q.b : (const uint64_t*)(din + o.b);
        uint64_t* dz = dev ? q.z : (uint64_t*)(dout + o.z);
        if (p.virt) {
            // internal product k of a KS record: operands and product in its scratch
            const size_t kk = (size_t)p.virt_ks;
            uint64_t* ptrs[9];
            uint64_t lens[2];
            const fa_request& qs = R[ksv[kk].r];
            fa_ks_products_v(ks_variant, ksv[kk].sb, (uint64_t)qs.an, (uint64_t)qs.bn,
                             (uint64_t*)(db0 + ks_scratch[kk]), ptrs, lens);
            da = ptrs[3 * (p.virt - 1)];
            db = ptrs[3 * (p.virt - 1) + 1];
            dz = ptrs[3 * (p.virt - 1) + 2];
        }
        if (!dev && !p.virt) {
            // host side of the staging: offsets into the pinned buffers
            const size_t aw = q.kind == FA_REQ_FMPZ ? (size_t)q.aw : 1;
            const size_t bw = q.kind == FA_REQ_FMPZ ? (size_t)q.bw : 1;
            B->h2d.push_back({(void*)(uintptr_t)o.a, q.a, (size_t)q.an * aw * 8});
            B->h2d.push_back({(void*)(uintptr_t)o.b, q.b, (size_t)q.bn * bw * 8});
            size_t zwords = q.kind == FA_REQ_MPN ? (size_t)p.zn
                            : q.kind == FA_REQ_NMOD ? (size_t)p.cnt : (size_t)p.cnt * (size_t)q.zw;
            B->d2h.push_back({q.z, (const void*)(uintptr_t)o.z, zwords * 8});
        }
        if (job_of[(size_t)r] >= 0) {
            fa_bjob& j = jobs[(size_t)job_of[(size_t)r]];
            j.a = da;
            j.b = db;
            j.A = (double*)(db0 + o.A);
            j.B = p.depth > FA_ROW_LG ? (double*)(db0 + o.B) : nullptr;
        }
        if (q.kind == FA_REQ_MPN) {
            fa_bmpn& m = mrec[(size_t)rec_of[(size_t)r]];
            m.Q = (const double*)(db0 + o.A);
            m.z = dz;
            m.flags = (uint8_t*)(db0 + o.flags);
            m.agg = (uint8_t*)(db0 + o.agg);
        } else {
            fa_bpoly& b = prec[(size_t)rec_of[(size_t)r]];
            b.Q = p.np ? (const double*)(db0 + o.A) : nullptr;
            b.C = p.np ? (const void*)(d + pconsts[{p.np, p.offset}]) : nullptr;
            b.z = dz;
        }
    }

    // upload records, constants and work lists in one image
    std::vector<char> img(L.size, 0);
    auto put = [&](size_t off, const void* src, size_t bytes) {
        if (bytes) memcpy(img.data() + off, src, bytes);
    };
    put(B->jobs, jobs.data(), jobs.size() * sizeof(fa_bjob));
    put(B->mpn_rec, mrec.data(), mrec.size() * sizeof(fa_bmpn));
    put(B->poly_rec, prec.data(), prec.size() * sizeof(fa_bpoly));
    {
        size_t k = 0;
        for (auto& e : small_bins) { put(B->small[k].second.first, e.second.data(), e.second.size() * 8); k++; }
        k = 0;
        for (auto& e : trans_bins) {
            put(B->trans[k].fwd, e.second.fwd.data(), e.second.fwd.size() * 8);
            put(B->trans[k].row, e.second.row.data(), e.second.row.size() * 8);
            put(B->trans[k].inv, e.second.inv.data(), e.second.inv.size() * 8);
            k++;
        }
        k = 0;
        for (auto& e : crt_bins) { put(B->crt_mpn[k].second.first, e.second.data(), e.second.size() * 8); k++; }
    }
    put(B->carry_jobs, carry.data(), carry.size() * 4);
    {
        std::vector<fa_bks> recs(ks_live.size());
        for (size_t i = 0; i < ks_live.size(); i++) {
            const size_t k = ks_live[i];
            const fa_request& q = R[ksv[k].r];
            const offs& o = O[(size_t)ksv[k].r];
            fa_ks_record(ksv[k].sb, dev ? q.a : (const uint64_t*)(din + o.a), (uint64_t)q.aw,
                         (uint64_t)q.an, dev ? q.b : (const uint64_t*)(din + o.b), (uint64_t)q.bw,
                         (uint64_t)q.bn, (uint64_t*)(db0 + ks_scratch[k]),
                         dev ? q.z : (uint64_t*)(dout + o.z), (uint64_t)q.zw, (uint64_t)q.zl,
                         (uint64_t)(q.zh - q.zl), &recs[i]);
        }
        put(B->ks_rec, recs.data(), recs.size() * sizeof(fa_bks));
        put(B->ks_packw, ks_pw.data(), ks_pw.size() * 8);
        put(B->ks_unpackw, ks_uw.data(), ks_uw.size() * 8);
    }
    put(B->nmod_work, nmod_w.data(), nmod_w.size() * 8);
    put(B->fmpz_work, fmpz_w.data(), fmpz_w.size() * 8);
    for (auto& e : pconsts) {
        fa_crt_consts C;
        fa_crt_consts_init(&C, e.first.first, e.first.second);
        put(e.second, &C, sizeof(C));
    }
    for (int k = 2; k <= FA_NPRIMES; k++) {
        fa_crt_consts C;
        fa_crt_consts_init(&C, k, 0);
        fa_crt_fp_consts F;
        fa_crt_fp_consts_init(&F, C);
        put(fpc[k], &F, sizeof(F));
    }
    if (hipMemcpy(d, img.data(), L.size, hipMemcpyHostToDevice) != hipSuccess) {
        fa_batch_destroy(B);
        return fa_set_error(FA_E_HIP, "fa_batch_create: upload");
    }
    B->din = din;
    B->dout = dout;
    B->nin = Lin.size;
    B->nout = Lout.size;
    if (!dev && (Lin.size || Lout.size)) {
        const bool ok =
            B->lent ? (lend_pinned(&lent.hin, &lent.hin_size, Lin.size + 256, &B->hin) == FA_OK &&
                       lend_pinned(&lent.hout, &lent.hout_size, Lout.size + 256, &B->hout) == FA_OK)
                    : (hipHostMalloc((void**)&B->hin, Lin.size + 256, 0) == hipSuccess &&
                       hipHostMalloc((void**)&B->hout, Lout.size + 256, 0) == hipSuccess);
        if (!ok) {
            fa_batch_destroy(B);
            return fa_set_error(FA_E_NOMEM, "fa_batch_create: pinned staging");
        }
    }

    if (flags & FA_BATCH_GRAPH) {
        if (hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            fa_batch_destroy(B);
            return fa_set_error(FA_E_HIP, "fa_batch_create: begin capture");
        }
        const int prof = fa_prof_on();
        if (prof) fa_prof_enable(0);
        int rc = batch_record(B, B->stream);
        if (prof) fa_prof_enable(1);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(B->stream, &g);
        if (rc || e != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            fa_batch_destroy(B);
            return rc ? rc : fa_set_error(FA_E_HIP, "fa_batch_create: end capture");
        }
        B->graph = g;
        if (hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            fa_batch_destroy(B);
            return fa_set_error(FA_E_HIP, "fa_batch_create: instantiate");
        }
    }
    *out = B;
    return FA_OK;
}

extern "C" int fa_batch_launch(fa_batch_s* B, void* stream)
{
    if (!B) return fa_set_error(FA_E_ARG, "fa_batch_launch: null batch");
    hipStream_t s = stream ? (hipStream_t)stream : B->stream;
    if (!B->h2d.empty()) {
        // operands gathered into pinned memory, one copy up
        for (auto& c : B->h2d) memcpy(B->hin + (uintptr_t)c.dst, c.src, c.bytes);
        if (hipMemcpyAsync(B->din, B->hin, B->nin, hipMemcpyHostToDevice, s) != hipSuccess)
            return fa_set_error(FA_E_HIP, "fa_batch_launch: upload");
    }
    int rc = FA_OK;
    if (B->exec) {
        if (hipGraphLaunch(B->exec, s) != hipSuccess)
            return fa_set_error(FA_E_HIP, "fa_batch_launch: graph launch");
    } else {
        rc = batch_record(B, s);
        if (rc) return rc;
    }
    if (!B->d2h.empty() &&
        hipMemcpyAsync(B->hout, B->dout, B->nout, hipMemcpyDeviceToHost, s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "fa_batch_launch: download");
    if (!(B->flags & FA_BATCH_DEVICE) || !stream) {
        if (hipStreamSynchronize(s) != hipSuccess)
            return fa_set_error(FA_E_HIP, "fa_batch_launch: %s",
                                hipGetErrorString(hipGetLastError()));
    }
    for (auto& c : B->d2h) memcpy(c.dst, B->hout + (uintptr_t)c.src, c.bytes);
    return FA_OK;
}

extern "C" int fa_batch_mul(fa_request* R, int64_t nreq, int flags)
{
    fa_batch_s* B = nullptr;
    int rc = fa_batch_create(&B, R, nreq, (flags & ~FA_BATCH_DEVICE & ~FA_BATCH_GRAPH) | FA_BATCH_LENT);
    if (rc) return rc;
    rc = fa_batch_launch(B, nullptr);
    fa_batch_destroy(B);
    return rc;
}
