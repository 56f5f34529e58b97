// batch.hip -- many independent systems in single launches.
//
// A batch is built from assembled handles (same device).  Their work lists are
// merged phase by phase -- up passes aligned at the leaves, one M2L + near-field
// launch, down passes ending at the leaves, one launch of semi-nested column
// factors -- so a batched product is one launch per phase for all systems,
// whatever their depths.  Every merged item carries its system's index in
// flags >> 8; the kernel takes that system's operator pools from a table, so no
// operator is copied.  The systems' vector pools are concatenated (16
// right-hand sides interleaved), and solves run in lockstep over every
// (system, right-hand side) column.
//
// Caller vectors: for each system in order, an N_s x nrhs column-major block
// (leading dimension N_s), the blocks one after another.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "../../include/h2w.h"
#include "h2w_handle.hpp"

struct h2w_batch {
    std::vector<h2w_handle *> sys;
    std::vector<int> inner_gen;          // inner-operator generation of every system at creation
    int S = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n = 0;                       // lockstep slots, sum of N_s
    std::vector<int64_t> nstart;         // lockstep start of system s (S + 1 entries)
    std::vector<int64_t> vbase;          // pool slot base of system s
    int64_t vec_len = 0;
    bool inner = false;                  // every system has an inner operator
    h2w::DevBuf<double> vec64m;
    h2w::DevBuf<float> vec32m;
    h2w::DevBuf<const double *> tab64;   // [2 s]: mat64 (twice)
    h2w::DevBuf<const float *> tab32;    // [2 s]: mat32, [2 s + 1]: inner buffer
    h2w::LaunchSet lists[2];             // 0: stored operators, 1: inner operators
    h2w::DevBuf<int32_t> xmap, ymap, seg;
    h2w::DevBuf<int2> blk;
    h2w::DevBuf<int32_t> sblk;
    int nblk = 0;
    std::vector<int32_t> perm;           // lockstep slot -> row of its system's caller block
    hipGraphExec_t graph[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    h2w::LockBufs buf;
    bool sym16 = true;                   // fused launch reads each block once (k_bgemm_sym)
    // the systems' deflation spaces (h2w_set_deflation on each inner operator), row-major over the
    // lockstep slots and zero-padded to the largest k: U (n x k) then C (n x k); rebuilt when a
    // system's space changes
    h2w::DevBuf<double> defl;
    int defl_k = 0;
    std::vector<int> defl_src;           // the systems' space serials it was built from
};

namespace h2w {
namespace {

using Batch = h2w_batch;

// merge the systems' launch lists phase by phase
void merge_lists(Batch &B, int which) {
    struct Src {
        int s;
        const Launch *L;
        const LaunchSet *ls;  // null: the handle's main lists
        int role;
    };
    auto key_of = [](const Launch &L) {
        // up steps from the leaves, down steps ending at the leaves
        return std::make_pair(L.stage, L.stage == 2 ? -L.step : L.step);
    };
    std::map<std::pair<int, int>, std::vector<Src>> groups;
    for (int s = 0; s < B.S; ++s) {
        Handle &h = *B.sys[s];
        const std::vector<Launch> &Ls = which == 0 ? h.launches : h.inner.launches;
        const LaunchSet &tr = which == 0 ? h.trans : h.inner_trans;
        for (auto &L : Ls) {
            groups[key_of(L)].push_back(Src{s, &L, which == 0 ? nullptr : &h.inner, L.role});
            if (L.role == 1 && !B.sym16)  // the transposed products, after the direct ones
                for (auto &Lt : tr.launches) groups[key_of(Lt)].push_back(Src{s, &Lt, &tr, 3});
        }
    }
    LaunchSet &M = B.lists[which];
    M.launches.clear();
    M.elems.clear();
    std::vector<GItem> items;
    std::vector<GTerm> terms;
    for (auto &kv : groups) {
        Launch Lm;
        Lm.item0 = (int64_t)items.size();
        Lm.term0 = (int64_t)terms.size();
        Lm.role = kv.second.front().role;
        Lm.stage = kv.first.first;
        Lm.step = kv.second.front().L->step;
        int64_t e = 0;
        for (auto &src : kv.second) {
            Handle &h = *B.sys[src.s];
            const std::vector<GItem> &hi = src.ls ? src.ls->h_items : h.h_items;
            const std::vector<GTerm> &ht = src.ls ? src.ls->h_terms : h.h_terms;
            const int64_t vb = B.vbase[src.s];
            for (int64_t q = src.L->item0; q < src.L->item0 + src.L->nitems; ++q) {
                GItem it = hi[q];
                const int32_t t0 = (int32_t)terms.size();
                for (int32_t t = it.t0; t < it.t1; ++t) {
                    GTerm tm = ht[t];
                    tm.x += vb;
                    if (tm.t >= 0) tm.t += vb;  // (the partner's rows of a symmetric term)
                    e += (int64_t)it.rows * tm.cols;
                    terms.push_back(tm);
                }
                it.t0 = t0;
                it.t1 = (int32_t)terms.size();
                it.y += vb;
                it.xb += vb;
                it.flags = (it.flags & 3) | (src.s << 8);
                items.push_back(it);
            }
        }
        Lm.nitems = (int64_t)items.size() - Lm.item0;
        Lm.nterms = (int64_t)terms.size() - Lm.term0;
        int64_t sc = 0;
        for (int64_t t = Lm.term0; t < (int64_t)terms.size(); ++t) sc += terms[t].cols;
        Lm.avg_cols = Lm.nterms ? (int)(sc / Lm.nterms) : 0;
        if (Lm.nitems == 0) continue;
        Lm.avg_item_elems = e / Lm.nitems;
        M.launches.push_back(Lm);
        M.elems.push_back(e);
    }
    M.d_items.upload(items, B.stream);
    M.d_terms.upload(terms, B.stream);
}

void check_batch(Batch &B, int precision) {
    if (precision != H2W_FP64 && precision != H2W_FP32 && precision != H2W_FP32_INNER)
        throw Error{H2W_EINVAL, "batches run on the stored FP64, FP32 or FP32-inner operators"};
    for (int s = 0; s < B.S; ++s) {
        Handle &h = *B.sys[s];
        if (is32(precision) && !h.mat32.p) throw Error{H2W_ESTATE, "a system of the batch has no FP32 shadow"};
        if (precision == H2W_FP32_INNER && (!B.inner || h.inner_gen != B.inner_gen[s]))
            throw Error{H2W_ESTATE, "the batch's inner operators are missing or were rebuilt after h2w_batch_create"};
    }
}

void run_batch_launch(Batch &B, const Launch &L, int precision) {
    hipStream_t s = B.stream;
    const LaunchSet &M = B.lists[precision == H2W_FP32_INNER ? 1 : 0];
    if (L.role == 1 && B.sym16) {
        if (is32(precision))
            launch_bgemm_sym32(M.d_items.p + L.item0, (int)L.nitems, M.d_terms.p, nullptr, nullptr, B.tab32.p,
                               B.vec32m.p, s);
        else
            launch_bgemm_sym64(M.d_items.p + L.item0, (int)L.nitems, M.d_terms.p, nullptr, nullptr, B.tab64.p,
                               B.vec64m.p, s);
    } else if (is32(precision)) {
        if (L.role == 3)
            launch_bgemm_t32(M.d_items.p + L.item0, (int)L.nitems, M.d_terms.p, nullptr, nullptr, B.tab32.p,
                             B.vec32m.p, trans_narrow(L, true), s);
        else
            launch_bgemm32_tab(M.d_items.p + L.item0, (int)L.nitems, M.d_terms.p, B.tab32.p, B.vec32m.p, L.role,
                               per_wave_items(L), s);
    } else {
        if (L.role == 3)
            launch_bgemm_t64(M.d_items.p + L.item0, (int)L.nitems, M.d_terms.p, nullptr, nullptr, B.tab64.p,
                             B.vec64m.p, trans_narrow(L, false), s);
        else
            launch_bgemm64_tab(M.d_items.p + L.item0, (int)L.nitems, M.d_terms.p, B.tab64.p, B.vec64m.p, L.role,
                               per_wave_items(L), s);
    }
}

void run_batch_eager(Batch &B, int precision) {
    for (auto &L : B.lists[precision == H2W_FP32_INNER ? 1 : 0].launches) run_batch_launch(B, L, precision);
    H2W_HIP(hipGetLastError());
}

// the launches of one batched product, captured once per precision
void run_batch(Batch &B, int precision) {
    hipStream_t s = B.stream;
    if (!B.graph[precision]) {
        hipGraph_t g = nullptr;
        H2W_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            run_batch_eager(B, precision);
        } catch (...) {
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        H2W_HIP(hipStreamEndCapture(s, &g));
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) throw Error{H2W_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)};
        B.graph[precision] = ex;
    }
    H2W_HIP(hipGraphLaunch(B.graph[precision], s));
}

// Y = A X on lockstep vectors (device): zero the pool, map x in, the launches, map y out
void batch_matvec_lock(Batch &B, const double *X, double *Y, int precision) {
    check_batch(B, precision);
    hipStream_t s = B.stream;
    if (is32(precision)) {
        B.vec32m.ensure((size_t)B.vec_len * kNR);
        launch_fill32(B.vec32m.p, B.vec_len * kNR, 0.0f, s);
        launch_map_in32(X, B.xmap.p, B.n, B.vec32m.p, s);
        run_batch(B, precision);
        launch_map_out32(B.vec32m.p, B.ymap.p, B.n, Y, s);
    } else {
        B.vec64m.ensure((size_t)B.vec_len * kNR);
        launch_fill(B.vec64m.p, B.vec_len * kNR, 0.0, s);
        launch_map_in64(X, B.xmap.p, B.n, B.vec64m.p, s);
        run_batch(B, precision);
        launch_map_out64(B.vec64m.p, B.ymap.p, B.n, Y, s);
    }
    H2W_HIP(hipGetLastError());
}

// The systems' deflation spaces as one row-major array over the lockstep slots, when every system has one
// for its current inner operator (else none is used).
void batch_deflation(Batch &B, Lockstep &L) {
    int k = 0;
    std::vector<int> src;
    for (auto *h : B.sys) {
        if (!(h->defl_k > 0 && h->defl_prec == H2W_FP32_INNER && h->defl_gen == h->inner_gen)) return;
        k = std::max(k, h->defl_k);
        src.push_back(h->defl_serial);
    }
    hipStream_t s = B.stream;
    if (k != B.defl_k || src != B.defl_src) {
        B.defl.alloc((size_t)2 * k * B.n);
        launch_fill(B.defl.p, 2 * k * B.n, 0.0, s);
        for (int sy = 0; sy < B.S; ++sy) {
            const h2w_handle &h = *B.sys[sy];
            const int64_t Ns = B.nstart[sy + 1] - B.nstart[sy];
            for (int part = 0; part < 2; ++part)  // U, then C
                H2W_HIP(hipMemcpy2DAsync(B.defl.p + (int64_t)part * k * B.n + B.nstart[sy] * k, (size_t)k * 8,
                                         h.g_defl_rm.p + (int64_t)part * h.defl_k * Ns, (size_t)h.defl_k * 8,
                                         (size_t)h.defl_k * 8, (size_t)Ns, hipMemcpyDeviceToDevice, s));
        }
        B.defl_k = k;
        B.defl_src = src;
    }
    L.dU = B.defl.p;
    L.dC = B.defl.p + (int64_t)k * B.n;
    L.dk = k;
    L.dprec = H2W_FP32_INNER;
}

Lockstep lockstep_of(Batch &B) {
    Lockstep L;
    L.s = B.stream;
    L.n = B.n;
    L.S = B.S;
    L.seg = B.seg.p;
    L.blk = B.blk.p;
    L.nblk = B.nblk;
    L.sblk = B.sblk.p;
    for (auto *h : B.sys) L.ir_floor.push_back(std::pow(10.0, -std::max(1.0, h->ir_exp)));
    L.matvec = [&B](const double *x, double *y, int prec) { batch_matvec_lock(B, x, y, prec); };
    L.buf = &B.buf;
    batch_deflation(B, L);
    return L;
}

// caller blocks (columns r0 .. r0 + nc - 1 of every system's N_s x nrhs block) <-> lockstep vectors
struct BlockMap {
    DevBuf<int64_t> src;
    DevBuf<int32_t> ld;
};
void block_map(Batch &B, int nrhs, int r0, BlockMap &bm) {
    std::vector<int64_t> src(B.n);
    std::vector<int32_t> ld(B.n);
    for (int s = 0; s < B.S; ++s) {
        const int64_t Ns = B.nstart[s + 1] - B.nstart[s];
        const int64_t base = (int64_t)nrhs * B.nstart[s] + (int64_t)r0 * Ns;
        for (int64_t t = 0; t < Ns; ++t) {
            src[B.nstart[s] + t] = base + B.perm[B.nstart[s] + t];
            ld[B.nstart[s] + t] = (int32_t)Ns;
        }
    }
    bm.src.upload(src, B.stream);
    bm.ld.upload(ld, B.stream);
}

}  // namespace
}  // namespace h2w

// ------------------------------------------------------------------ C ABI
using h2w::Error;

#define B_TRY try {
#define B_CATCH                                   \
    }                                             \
    catch (const Error &e) {                      \
        h2w::set_error(e.msg);                    \
        return e.code;                            \
    }                                             \
    catch (const std::bad_alloc &) {              \
        h2w::set_error("host allocation failed"); \
        return H2W_ENOMEM;                        \
    }                                             \
    catch (const std::exception &e) {             \
        h2w::set_error(e.what());                 \
        return H2W_EINVAL;                        \
    }

extern "C" {

int h2w_batch_create(h2w_handle *const *systems, int nsys, h2w_batch **out) {
    B_TRY
    if (!systems || nsys < 1 || !out) throw Error{H2W_EINVAL, "bad arguments"};
    if (nsys > (1 << 22)) throw Error{H2W_EINVAL, "too many systems (system index lives in item flags >> 8)"};
    *out = nullptr;
    std::unique_ptr<h2w_batch> B(new h2w_batch);
    B->S = nsys;
    B->device = systems[0] ? systems[0]->opts.device : 0;
    B->inner = true;
    B->sym16 = std::getenv("H2W_NO_SYM16") == nullptr;
    for (int s = 0; s < nsys; ++s) {
        h2w_handle *h = systems[s];
        if (!h || !h->assembled) throw Error{H2W_ESTATE, "every system of a batch must be assembled"};
        if (h->cplx) throw Error{H2W_EINVAL, "complex kernels are not batched (one h2w_gmres per system)"};
        if (h->opts.device != B->device) throw Error{H2W_EINVAL, "the systems of a batch share one device"};
        B->sys.push_back(h);
        B->inner_gen.push_back(h->inner_gen);
        B->inner = B->inner && !h->inner.launches.empty();
    }
    H2W_HIP(hipSetDevice(B->device));
    H2W_HIP(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    // slot layout of the lockstep vectors and of the concatenated pool
    B->nstart.assign(nsys + 1, 0);
    B->vbase.assign(nsys, 0);
    for (int s = 0; s < nsys; ++s) {
        B->nstart[s + 1] = B->nstart[s] + B->sys[s]->N;
        B->vbase[s] = B->vec_len;
        B->vec_len += B->sys[s]->vec_len;
    }
    B->n = B->nstart[nsys];
    if (B->vec_len >= ((int64_t)1 << 31)) throw Error{H2W_EINVAL, "batch vector pool exceeds 2^31 slots"};
    std::vector<int32_t> xmap(B->n), ymap(B->n), seg(B->n);
    B->perm.assign(B->n, 0);
    for (int s = 0; s < nsys; ++s) {
        h2w_handle &h = *B->sys[s];
        std::vector<int32_t> perm(h.N);
        if (!h.plan.perm.empty())
            for (int64_t t = 0; t < h.N; ++t) perm[t] = (int32_t)h.plan.perm[t];
        for (int64_t t = 0; t < h.N; ++t) {
            const int64_t g = B->nstart[s] + t;
            xmap[g] = (int32_t)(B->vbase[s] + h.off_x + t);
            ymap[g] = (int32_t)(B->vbase[s] + h.off_y + t);
            seg[g] = s;
            B->perm[g] = perm[t];
        }
    }
    B->xmap.upload(xmap, B->stream);
    B->ymap.upload(ymap, B->stream);
    B->seg.upload(seg, B->stream);
    std::vector<int2> blk;
    std::vector<int32_t> sblk;
    h2w::lock_blocks(B->nstart, blk, sblk);
    B->blk.upload(blk, B->stream);
    B->sblk.upload(sblk, B->stream);
    B->nblk = (int)blk.size();
    std::vector<const double *> t64(2 * nsys);
    std::vector<const float *> t32(2 * nsys);
    for (int s = 0; s < nsys; ++s) {
        t64[2 * s] = t64[2 * s + 1] = B->sys[s]->mat64.p;
        t32[2 * s] = B->sys[s]->mat32.p;
        t32[2 * s + 1] = B->sys[s]->mat32in.p ? B->sys[s]->mat32in.p : B->sys[s]->mat32.p;
    }
    B->tab64.upload(t64, B->stream);
    B->tab32.upload(t32, B->stream);
    h2w::merge_lists(*B, 0);
    if (B->inner) h2w::merge_lists(*B, 1);
    H2W_HIP(hipStreamSynchronize(B->stream));
    *out = B.release();
    return H2W_OK;
    B_CATCH
}

int h2w_batch_matvec(h2w_batch *B, const double *X, double *Y, int nrhs, int precision, int where) {
    B_TRY
    if (!B || !X || !Y || nrhs < 1) throw Error{H2W_EINVAL, "bad arguments"};
    H2W_HIP(hipSetDevice(B->device));
    h2w::check_batch(*B, precision);
    hipStream_t s = B->stream;
    const size_t tot = (size_t)B->n * nrhs;
    h2w::DevBuf<double> hx, hy, lx, ly;
    const double *xs = X;
    double *ys = Y;
    if (where == H2W_HOST) {
        hx.alloc(tot);
        hy.alloc(tot);
        H2W_HIP(hipMemcpyAsync(hx.p, X, tot * 8, hipMemcpyHostToDevice, s));
        xs = hx.p;
        ys = hy.p;
    }
    lx.alloc((size_t)B->n * h2w::kNR);
    ly.alloc((size_t)B->n * h2w::kNR);
    h2w::BlockMap bm;
    for (int r0 = 0; r0 < nrhs; r0 += h2w::kNR) {
        const int nc = std::min(h2w::kNR, nrhs - r0);
        h2w::block_map(*B, nrhs, r0, bm);
        h2w::launch_gather_b(xs, bm.src.p, bm.ld.p, nc, B->n, lx.p, s);
        h2w::batch_matvec_lock(*B, lx.p, ly.p, precision);
        h2w::launch_scatter_b(ly.p, bm.src.p, bm.ld.p, nc, B->n, ys, s);
    }
    if (where == H2W_HOST) H2W_HIP(hipMemcpyAsync(Y, hy.p, tot * 8, hipMemcpyDeviceToHost, s));
    H2W_HIP(hipStreamSynchronize(s));
    return H2W_OK;
    B_CATCH
}

int h2w_batch_gmres(h2w_batch *B, const double *Bv, double *X, int nrhs, int max_iter, double tol, int mode,
                    int where, h2w_gmres_stats *st) {
    B_TRY
    if (!B || !Bv || !X || !st || nrhs < 1 || max_iter < 1) throw Error{H2W_EINVAL, "bad arguments"};
    if (mode != H2W_GMRES_FP64 && mode != H2W_GMRES_FP32_IR && mode != H2W_GMRES_FP32_INNER_IR)
        throw Error{H2W_EINVAL, "batches: modes FP64, FP32_IR and FP32_INNER_IR"};
    H2W_HIP(hipSetDevice(B->device));
    h2w::check_batch(*B, H2W_FP64);
    if (mode != H2W_GMRES_FP64) h2w::check_batch(*B, mode == H2W_GMRES_FP32_INNER_IR ? H2W_FP32_INNER : H2W_FP32);
    hipStream_t s = B->stream;
    const size_t tot = (size_t)B->n * nrhs;
    h2w::DevBuf<double> hb, hx, lb, lx;
    const double *bs = Bv;
    double *xs = X;
    if (where == H2W_HOST) {
        hb.alloc(tot);
        hx.alloc(tot);
        H2W_HIP(hipMemcpyAsync(hb.p, Bv, tot * 8, hipMemcpyHostToDevice, s));
        H2W_HIP(hipMemcpyAsync(hx.p, X, tot * 8, hipMemcpyHostToDevice, s));
        bs = hb.p;
        xs = hx.p;
    }
    lb.alloc((size_t)B->n * h2w::kNR);
    lx.alloc((size_t)B->n * h2w::kNR);
    h2w::Lockstep L = h2w::lockstep_of(*B);
    h2w::BlockMap bm;
    std::vector<h2w_gmres_stats> cs;
    bool all = true;
    for (int r0 = 0; r0 < nrhs; r0 += h2w::kNR) {
        const int nc = std::min(h2w::kNR, nrhs - r0);
        h2w::block_map(*B, nrhs, r0, bm);
        h2w::launch_gather_b(bs, bm.src.p, bm.ld.p, nc, B->n, lb.p, s);
        h2w::launch_gather_b(xs, bm.src.p, bm.ld.p, nc, B->n, lx.p, s);
        cs.assign((size_t)B->S * nc, h2w_gmres_stats{});
        h2w::lock_solve(L, lb.p, lx.p, nc, max_iter, tol, mode, cs.data());
        h2w::launch_scatter_b(lx.p, bm.src.p, bm.ld.p, nc, B->n, xs, s);
        for (int sy = 0; sy < B->S; ++sy)
            for (int r = 0; r < nc; ++r) {
                st[(size_t)sy * nrhs + r0 + r] = cs[(size_t)sy * nc + r];
                all = all && cs[(size_t)sy * nc + r].converged;
            }
    }
    if (where == H2W_HOST) H2W_HIP(hipMemcpyAsync(X, hx.p, tot * 8, hipMemcpyDeviceToHost, s));
    H2W_HIP(hipStreamSynchronize(s));
    return all ? H2W_OK : H2W_ENOCONV;
    B_CATCH
}

int h2w_batch_bench_matvec(h2w_batch *B, int nrep, int precision, h2w_bench *out) {
    B_TRY
    if (!B || !out || nrep < 1) throw Error{H2W_EINVAL, "bad arguments"};
    H2W_HIP(hipSetDevice(B->device));
    h2w::check_batch(*B, precision);
    std::memset(out, 0, sizeof(*out));
    hipStream_t s = B->stream;
    h2w::DevBuf<double> lx, ly;
    lx.alloc((size_t)B->n * h2w::kNR);
    ly.alloc((size_t)B->n * h2w::kNR);
    h2w::launch_fill(lx.p, B->n * h2w::kNR, 1.0, s);
    h2w::batch_matvec_lock(*B, lx.p, ly.p, precision);
    h2w::Event e0, e1;
    H2W_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < nrep; ++r) h2w::batch_matvec_lock(*B, lx.p, ly.p, precision);
    H2W_HIP(hipEventRecord(e1, s));
    out->ms_per_matvec = h2w::elapsed_ms(e0, e1) / nrep;
    const auto &M = B->lists[precision == H2W_FP32_INNER ? 1 : 0];
    out->n_launch = (int)std::min<size_t>(M.launches.size(), H2W_MAX_LAUNCHES);
    const int es = h2w::is32(precision) ? 4 : 8;
    // per-launch durations: the launches one by one between events (eager, same kernels as the graph)
    std::vector<h2w::Event> ev(out->n_launch + 1);
    std::vector<double> acc(out->n_launch, 0.0);
    for (int r = 0; r < nrep; ++r) {
        H2W_HIP(hipEventRecord(ev[0], s));
        for (int q = 0; q < out->n_launch; ++q) {
            h2w::run_batch_launch(*B, M.launches[q], precision);
            H2W_HIP(hipEventRecord(ev[q + 1], s));
        }
        for (int q = 0; q < out->n_launch; ++q) acc[q] += h2w::elapsed_ms(ev[q], ev[q + 1]);
    }
    for (int q = 0; q < out->n_launch; ++q) {
        out->launch_ms[q] = acc[q] / nrep;
        out->launch_role[q] = M.launches[q].role;
        out->launch_bytes[q] = M.elems[q] * es;
    }
    return H2W_OK;
    B_CATCH
}

void h2w_batch_destroy(h2w_batch *B) {
    if (!B) return;
    (void)hipSetDevice(B->device);
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    for (auto &g : B->graph)
        if (g) (void)hipGraphExecDestroy(g);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
}

}  // extern "C"
