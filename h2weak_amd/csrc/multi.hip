// multi.hip -- kNR-wide runtime of libh2w: the product with up to kNR
// right-hand sides per pass on the matrix cores, and kNR independent GMRES
// solves (or FP64 refinements) advanced in lockstep so that every Arnoldi step
// streams the compressed operator once for all of them.
//
// Vectors are N x kNR, interleaved: element (t, r) at t * kNR + r, t in tree
// order.  Columns beyond the caller's nrhs, and columns whose solve has
// converged, are held at zero, so they cost no bandwidth beyond their share of
// the one operator pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "h2w_handle.hpp"

namespace h2w {

namespace {

void ensure_pool_m(Handle &h, int precision) {
    if (is32(precision)) h.vec32m.ensure((size_t)h.vec_len * kNR);
    else h.vec64m.ensure((size_t)h.vec_len * kNR);
}

// The fused stage-1 launch applies every stored block K(b, p), p >= b, directly and its
// transpose K(b, p)^T X_b in the same read (k_bgemm_sym); with H2W_NO_SYM16 set it applies
// them directly only and the transposed products follow as the launches of the role-3 lists
// (h.trans, h.inner_trans).
// One step of the sequence: a launch of the precision's lists, or (trans) launch Lh of
// its role-3 lists.
void run_launch_m(Handle &h, const Launch &Lh, int precision, bool trans) {
    hipStream_t s = h.stream;
    if (trans) {
        const LaunchSet &T = precision == H2W_FP32_INNER ? h.inner_trans : h.trans;
        if (precision == H2W_FP32_INNER)
            launch_bgemm_t32(T.d_items.p + Lh.item0, (int)Lh.nitems, T.d_terms.p, h.mat32.p, h.mat32in.p, nullptr,
                             h.vec32m.p, trans_narrow(Lh, true), s);
        else if (precision == H2W_FP32)
            launch_bgemm_t32(T.d_items.p + Lh.item0, (int)Lh.nitems, T.d_terms.p, h.mat32.p, nullptr, nullptr,
                             h.vec32m.p, trans_narrow(Lh, true), s);
        else
            launch_bgemm_t64(T.d_items.p + Lh.item0, (int)Lh.nitems, T.d_terms.p, h.mat64.p, nullptr, nullptr,
                             h.vec64m.p, trans_narrow(Lh, false), s);
    } else if (Lh.role == 1 && h.sym16) {
        // direct and transposed products in one read of every block
        if (precision == H2W_FP32_INNER)
            launch_bgemm_sym32(h.inner.d_items.p + Lh.item0, (int)Lh.nitems, h.inner.d_terms.p, h.mat32.p,
                               h.mat32in.p, nullptr, h.vec32m.p, s);
        else if (precision == H2W_FP32)
            launch_bgemm_sym32(h.d_items.p + Lh.item0, (int)Lh.nitems, h.d_terms.p, h.mat32.p, nullptr, nullptr,
                               h.vec32m.p, s);
        else
            launch_bgemm_sym64(h.d_items.p + Lh.item0, (int)Lh.nitems, h.d_terms.p, h.mat64.p, nullptr, nullptr,
                               h.vec64m.p, s);
    } else if (precision == H2W_FP32_INNER) {
        launch_bgemm32(h.inner.d_items.p + Lh.item0, (int)Lh.nitems, h.inner.d_terms.p, h.mat32.p, h.mat32in.p,
                       h.vec32m.p, Lh.role, per_wave_items(Lh), s);
    } else if (precision == H2W_FP32) {
        launch_bgemm32(h.d_items.p + Lh.item0, (int)Lh.nitems, h.d_terms.p, h.mat32.p, nullptr, h.vec32m.p, Lh.role,
                       per_wave_items(Lh), s);
    } else {
        launch_bgemm64(h.d_items.p + Lh.item0, (int)Lh.nitems, h.d_terms.p, h.mat64.p, nullptr, h.vec64m.p, Lh.role,
                       per_wave_items(Lh), s);
    }
}

// the sequence as (launch, trans) steps, output zeroing not included
std::vector<std::pair<const Launch *, bool>> steps_m(Handle &h, int precision) {
    std::vector<std::pair<const Launch *, bool>> st;
    const LaunchSet &T = precision == H2W_FP32_INNER ? h.inner_trans : h.trans;
    for (auto &Lh : launches_for(h, precision)) {
        st.push_back({&Lh, false});
        if (Lh.role == 1 && !h.sym16)
            for (auto &Lt : T.launches) st.push_back({&Lt, true});
    }
    return st;
}

void zero_outputs_m(Handle &h, int precision) {
    const int64_t nout = (h.vec_len - h.off_y) * kNR;
    if (is32(precision)) launch_fill32(h.vec32m.p + h.off_y * kNR, nout, 0.0f, h.stream);
    else launch_fill(h.vec64m.p + h.off_y * kNR, nout, 0.0, h.stream);
}

void run_launches_m_eager(Handle &h, int precision) {
    zero_outputs_m(h, precision);
    for (auto &st : steps_m(h, precision)) run_launch_m(h, *st.first, precision, st.second);
    H2W_HIP(hipGetLastError());
}

// captured once per precision, like the single-vector sequence
void run_launches_m(Handle &h, int precision) {
    if (!h.use_graph) {
        run_launches_m_eager(h, precision);
        return;
    }
    if (!h.graph_m[precision]) {
        hipGraph_t g = nullptr;
        H2W_HIP(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
        try {
            run_launches_m_eager(h, precision);
        } catch (...) {
            (void)hipStreamEndCapture(h.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        H2W_HIP(hipStreamEndCapture(h.stream, &g));
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) throw Error{H2W_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)};
        h.graph_m[precision] = ex;
    }
    H2W_HIP(hipGraphLaunch(h.graph_m[precision], h.stream));
}

void check_precision_m(Handle &h, int precision) {
    if (precision != H2W_FP64 && precision != H2W_FP32 && precision != H2W_FP32_INNER)
        throw Error{H2W_EINVAL, "several right-hand sides run on the stored FP64, FP32 or FP32-inner operators only"};
    if (is32(precision) && !h.mat32.p) throw Error{H2W_ESTATE, "FP32 matvec requested but the FP32 shadow was not built"};
}

// Y = A X on interleaved tree-order vectors (device)
void matvec_m_tree(Handle &h, const double *X, double *Y, int precision) {
    check_precision_m(h, precision);
    ensure_pool_m(h, precision);
    hipStream_t s = h.stream;
    const int64_t n = h.N * kNR;
    if (is32(precision)) {
        launch_to_fp32(X, h.vec32m.p + h.off_x * kNR, n, s);
        run_launches_m(h, precision);
        launch_to_fp64(h.vec32m.p + h.off_y * kNR, Y, n, s);
    } else {
        H2W_HIP(hipMemcpyAsync(h.vec64m.p + h.off_x * kNR, X, n * 8, hipMemcpyDeviceToDevice, s));
        run_launches_m(h, precision);
        H2W_HIP(hipMemcpyAsync(Y, h.vec64m.p + h.off_y * kNR, n * 8, hipMemcpyDeviceToDevice, s));
    }
}

}  // namespace

// ------------------------------------------------------------------ lockstep solves
void lock_blocks(const std::vector<int64_t> &start, std::vector<int2> &blk, std::vector<int32_t> &sblk) {
    const int S = (int)start.size() - 1;
    const int64_t n = start[S] - start[0];
    const int64_t chunk = std::max<int64_t>(64, (n + 1023) / 1024);  // ~1024 blocks over the whole vector
    blk.clear();
    sblk.assign(S + 1, 0);
    for (int s = 0; s < S; ++s) {
        sblk[s] = (int32_t)blk.size();
        for (int64_t a = start[s]; a < start[s + 1]; a += chunk)
            blk.push_back(make_int2((int)a, (int)std::min(start[s + 1], a + chunk)));
    }
    sblk[S] = (int32_t)blk.size();
}

namespace {

// per-column dot products of two lockstep vectors, to the host (S x kNR values)
void ldots(Lockstep &L, const double *a, const double *b, double *out) {
    LockBufs &B = *L.buf;
    B.part.ensure((size_t)L.nblk * kNR);
    B.h.ensure((size_t)L.S * kNR);
    launch_bdot16(a, 0, 1, b, L.blk, L.nblk, B.part.p, L.s);
    launch_bdot16_reduce(B.part.p, L.nblk, 1, L.sblk, L.S, B.h.p, L.s);
    H2W_HIP(hipMemcpyAsync(out, B.h.p, (size_t)L.S * kNR * 8, hipMemcpyDeviceToHost, L.s));
    H2W_HIP(hipStreamSynchronize(L.s));
}

// dst(t, r) = sc(s(t), r) src(t, r) with sc from the host
void lscale(Lockstep &L, const std::vector<double> &sc, const double *src, double *dst) {
    LockBufs &B = *L.buf;
    B.sc.ensure(sc.size());
    H2W_HIP(hipMemcpyAsync(B.sc.p, sc.data(), sc.size() * 8, hipMemcpyHostToDevice, L.s));
    launch_bscale16(src, B.sc.p, dst, L.n, L.seg, L.s);
}

void ensure_V(Lockstep &L, int64_t cols) {
    LockBufs &B = *L.buf;
    if (cols <= B.Vcols) return;
    const int64_t len = L.n * kNR;
    int64_t nc = std::max<int64_t>(cols, std::max<int64_t>(8, B.Vcols * 2));
    double *np = nullptr;
    if (hipMalloc(&np, (size_t)nc * len * 8) != hipSuccess) throw Error{H2W_ENOMEM, "GMRES basis allocation failed"};
    if (B.Vcols > 0) H2W_HIP(hipMemcpyAsync(np, B.V.p, (size_t)B.Vcols * len * 8, hipMemcpyDeviceToDevice, L.s));
    H2W_HIP(hipStreamSynchronize(L.s));
    B.V.release();
    B.V.p = np;
    B.V.n = (size_t)nc * len;
    B.Vcols = nc;
}

struct LockOut {
    std::vector<int> iters;
    std::vector<double> resid;
    std::vector<char> conv;
    int matvecs = 0;
};

// Unrestarted GMRES on every active column in lockstep: the single-vector core's
// algorithm (CGS2 Arnoldi, Givens rotations, resid = |s_{i+1}| / normb, back-
// substitution) per column, one batched product per step for all columns.
// Column c stops at its own tolerance tol[c]; a stopped column's basis vectors
// are zero from then on.  x_zero: x0 = 0 and no initial product; normb_in[c] > 0
// overrides |b_c|.  kc > 0 (with x_zero): recycle the space kept by lock_keep_space
// (C in the first kc basis columns, U in buf->U), as the single-vector core does.
// hkeep: receives the Hessenberg columns before the rotations, and every running
// column's last basis vector is kept, for lock_keep_space.  shared (with x_zero, kc = 0):
// start from the deflation space L.dU, L.dC that all columns of a system share instead
// (x = U C^T r, r -= C C^T r, every Arnoldi vector orthogonalised against C as well).
LockOut lock_gmres(Lockstep &L, const double *b, double *x, const std::vector<char> &active, int max_iter,
                   const std::vector<double> &tol, int prec, bool x_zero, const double *normb_in, int kc = 0,
                   std::vector<std::vector<double>> *hkeep = nullptr, bool shared = false) {
    const int C = L.S * kNR;
    const int64_t len = L.n * kNR;
    hipStream_t s = L.s;
    LockBufs &B = *L.buf;
    LockOut out;
    out.iters.assign(C, 0);
    out.resid.assign(C, 0.0);
    out.conv.assign(C, 0);
    B.w.ensure(len);
    B.r.ensure(len);
    double *w = B.w.p, *r = B.r.p;
    std::vector<double> normb(C), beta(C), tmp(C);
    if (normb_in) std::copy(normb_in, normb_in + C, tmp.begin());
    else {
        ldots(L, b, b, tmp.data());
        for (double &v : tmp) v = std::sqrt(v);
    }
    for (int c = 0; c < C; ++c) normb[c] = tmp[c] > 0.0 ? tmp[c] : 1.0;
    H2W_HIP(hipMemcpyAsync(r, b, len * 8, hipMemcpyDeviceToDevice, s));
    if (!x_zero) {
        L.matvec(x, w, prec);
        out.matvecs++;
        launch_axpby(-1.0, w, 1.0, r, len, s);
        kc = 0;
    }
    const int m = max_iter;
    B.part.ensure((size_t)L.nblk * (kc + m + 2) * kNR);
    B.h.ensure((size_t)(2 * (kc + m + 2) + 2) * C);
    if (kc > 0) {
        // x = U C^T r, r -= C C^T r (per column)
        launch_bdot16(B.V.p, len, kc, r, L.blk, L.nblk, B.part.p, s);
        launch_bdot16_reduce(B.part.p, L.nblk, kc, L.sblk, L.S, B.h.p, s);
        launch_bgemv16(B.V.p, len, kc, B.h.p, L.S, -1.0, r, L.n, L.seg, s);
        launch_bgemv16(B.U.p, len, kc, B.h.p, L.S, 1.0, x, L.n, L.seg, s);
    }
    const int ks = shared && x_zero && kc == 0 ? L.dk : 0;
    if (ks > 0) {
        // the same with the shared space: x = U C^T r, r -= C C^T r
        B.spart.ensure((size_t)L.nblk * ks * kNR);
        B.sh.ensure((size_t)2 * ks * C);
        launch_sdot16(L.dC, ks, r, L.blk, L.nblk, B.spart.p, s);
        launch_bdot16_reduce(B.spart.p, L.nblk, ks, L.sblk, L.S, B.sh.p, s);
        launch_supd16(L.dC, ks, B.sh.p, L.S, -1.0, r, L.n, L.seg, s);
        launch_supd16(L.dU, ks, B.sh.p, L.S, 1.0, x, L.n, L.seg, s);
    }
    ldots(L, r, r, tmp.data());
    std::vector<char> run(C, 0);
    int nrun = 0;
    for (int c = 0; c < C; ++c) {
        beta[c] = std::sqrt(tmp[c]);
        out.resid[c] = beta[c] / normb[c];
        run[c] = active[c] && out.resid[c] > tol[c];
        out.conv[c] = active[c] && !run[c];
        nrun += run[c];
    }
    if (nrun == 0) {
        H2W_HIP(hipStreamSynchronize(s));
        return out;
    }
    const size_t ld = (size_t)m + 1;
    std::vector<std::vector<double>> H;  // Hessenberg column j: rows 0..j+1 of every column c at i * C + c
    std::vector<double> sv(ld * C, 0.0), cs(ld * C, 0.0), sn(ld * C, 0.0);
    auto Hh = [&](int c, int i, int j) -> double & { return H[j][(size_t)i * C + c]; };
    std::vector<double> Y;  // update coefficients (j, c) at j * C + c
    std::vector<std::vector<double>> Bst;  // (recycled) C^T A v_i: (q, c) at q * C + c
    ensure_V(L, kc + std::min<int64_t>(m + 1, 16));
    for (int c = 0; c < C; ++c) {
        tmp[c] = run[c] ? 1.0 / beta[c] : 0.0;
        sv[c * ld] = beta[c];
    }
    lscale(L, tmp, r, B.V.p + (int64_t)kc * len);
    double *hh = B.pin.ensure((size_t)(2 * (kc + m + 2) + 2) * C + (size_t)2 * ks * C);
    int last = 0;
    for (int i = 0; i < m && nrun > 0; ++i) {
        ensure_V(L, kc + i + 2);
        double *Cb = B.V.p, *V = B.V.p + (int64_t)kc * len;
        L.matvec(V + (int64_t)i * len, w, prec);
        out.matvecs++;
        const int k = kc + i + 1;  // vectors orthogonalised against: C, v_0 .. v_i (contiguous)
        // classical Gram-Schmidt twice; the shared space's dot products come from the same w
        for (int pass = 0; pass < 2; ++pass) {
            launch_bdot16(Cb, len, k, w, L.blk, L.nblk, B.part.p, s);
            launch_bdot16_reduce(B.part.p, L.nblk, k, L.sblk, L.S, B.h.p + (size_t)pass * k * C, s);
            if (ks > 0 && pass == 0) {
                // (the shared space spans approximate eigenvectors: one pass against it, as the
                // single-vector solve does)
                launch_sdot16(L.dC, ks, w, L.blk, L.nblk, B.spart.p, s);
                launch_bdot16_reduce(B.spart.p, L.nblk, ks, L.sblk, L.S, B.sh.p, s);
                launch_supd16(L.dC, ks, B.sh.p, L.S, -1.0, w, L.n, L.seg, s);
            }
            launch_bgemv16(Cb, len, k, B.h.p + (size_t)pass * k * C, L.S, -1.0, w, L.n, L.seg, s);
        }
        launch_bdot16(w, 0, 1, w, L.blk, L.nblk, B.part.p, s);
        launch_bdot16_reduce(B.part.p, L.nblk, 1, L.sblk, L.S, B.h.p + (size_t)2 * k * C, s);
        H2W_HIP(hipMemcpyAsync(hh, B.h.p, (size_t)(2 * k + 1) * C * 8, hipMemcpyDeviceToHost, s));
        double *shh = hh + (size_t)(2 * k + 1) * C;
        if (ks > 0) H2W_HIP(hipMemcpyAsync(shh, B.sh.p, (size_t)ks * C * 8, hipMemcpyDeviceToHost, s));
        H2W_HIP(hipStreamSynchronize(s));
        last = i;
        H.emplace_back((size_t)(i + 2) * C, 0.0);
        Y.resize((size_t)(i + 1) * C, 0.0);
        if (kc > 0 || ks > 0) Bst.emplace_back((size_t)(kc + ks) * C, 0.0);
        for (int c = 0; c < C; ++c) {
            tmp[c] = 0.0;
            if (!run[c]) continue;
            for (int q = 0; q < kc; ++q) Bst.back()[(size_t)q * C + c] = hh[(size_t)q * C + c] + hh[(size_t)(k + q) * C + c];
            for (int q = 0; q < ks; ++q) Bst.back()[(size_t)q * C + c] = shh[(size_t)q * C + c];
            for (int q = 0; q <= i; ++q)
                Hh(c, q, i) = hh[(size_t)(kc + q) * C + c] + hh[(size_t)(k + kc + q) * C + c];
            Hh(c, i + 1, i) = std::sqrt(hh[(size_t)2 * k * C + c]);
            if (Hh(c, i + 1, i) > 0.0) tmp[c] = 1.0 / Hh(c, i + 1, i);
        }
        if (hkeep) hkeep->push_back(H.back());  // (before the rotations)
        for (int c = 0; c < C; ++c) {
            if (!run[c]) continue;
            double *csc = &cs[c * ld], *snc = &sn[c * ld], *svc = &sv[c * ld];
            for (int q = 0; q < i; ++q) app_rot(Hh(c, q, i), Hh(c, q + 1, i), csc[q], snc[q]);
            gen_rot(Hh(c, i, i), Hh(c, i + 1, i), csc[i], snc[i]);
            app_rot(Hh(c, i, i), Hh(c, i + 1, i), csc[i], snc[i]);
            app_rot(svc[i], svc[i + 1], csc[i], snc[i]);
            out.resid[c] = std::fabs(svc[i + 1]) / normb[c];
            out.iters[c] = i + 1;
            if (out.resid[c] < tol[c] || i == m - 1 || tmp[c] == 0.0) {
                std::vector<double> y(svc, svc + i + 1);
                for (int q = i; q >= 0; q--) {
                    y[q] /= Hh(c, q, q);
                    for (int jj = q - 1; jj >= 0; jj--) y[jj] -= Hh(c, jj, q) * y[q];
                }
                for (int q = 0; q <= i; ++q) Y[(size_t)q * C + c] = y[q];
                out.conv[c] = out.resid[c] < tol[c];
                run[c] = 0;
                if (!hkeep) tmp[c] = 0.0;  // (kept: the column's last basis vector enters C)
                --nrun;
            }
        }
        if (nrun > 0 || hkeep) lscale(L, tmp, w, V + (int64_t)(i + 1) * len);
    }
    // x += V Y over the basis vectors the columns used, - U (B Y) with a recycled space
    H2W_HIP(hipMemcpyAsync(B.h.p, Y.data(), (size_t)(last + 1) * C * 8, hipMemcpyHostToDevice, s));
    launch_bgemv16(B.V.p + (int64_t)kc * len, len, last + 1, B.h.p, L.S, 1.0, x, L.n, L.seg, s);
    if (kc > 0 || ks > 0) {
        const int kk = kc + ks;
        std::vector<double> by((size_t)kk * C, 0.0);
        for (int c = 0; c < C; ++c)
            for (int jj = 0; jj < out.iters[c] && jj < (int)Bst.size(); ++jj)
                for (int q = 0; q < kk; ++q) by[(size_t)q * C + c] -= Bst[jj][(size_t)q * C + c] * Y[(size_t)jj * C + c];
        B.coef.ensure(by.size());
        H2W_HIP(hipMemcpyAsync(B.coef.p, by.data(), by.size() * 8, hipMemcpyHostToDevice, s));
        if (kc > 0) launch_bgemv16(B.U.p, len, kc, B.coef.p, L.S, 1.0, x, L.n, L.seg, s);
        else launch_supd16(L.dU, ks, B.coef.p, L.S, 1.0, x, L.n, L.seg, s);
    }
    H2W_HIP(hipStreamSynchronize(s));
    return out;
}

// The space of a finished lockstep correction kept for the next ones, per column as the
// single-vector refinement does (U = V_j R^-1, C = V_{j+1} Q from Hbar = QR of the column's
// Hessenberg matrix, j its iteration count).  Columns of fewer than two iterations or with a
// near-singular R keep nothing (zero C and U); k is the largest j kept.
void lock_keep_space(Lockstep &L, const std::vector<std::vector<double>> &Hk, const LockOut &o) {
    LockBufs &B = *L.buf;
    B.rk = 0;
    const int C = L.S * kNR;
    const int64_t len = L.n * kNR;
    hipStream_t s = L.s;
    std::vector<int> jc(C, 0);
    std::vector<std::vector<std::vector<double>>> Qc(C), Ric(C);
    int k = 0;
    for (int c = 0; c < C; ++c) {
        const int j = o.iters[c];
        if (j < 2 || j > (int)Hk.size()) continue;
        std::vector<std::vector<double>> Q(j, std::vector<double>(j + 1, 0.0)), R(j, std::vector<double>(j, 0.0));
        bool ok = true;
        for (int col = 0; col < j && ok; ++col) {
            std::vector<double> v(j + 1, 0.0);
            for (int q = 0; q <= col + 1; ++q) v[q] = Hk[col][(size_t)q * C + c];
            for (int pass = 0; pass < 2; ++pass)
                for (int p = 0; p < col; ++p) {
                    double d = 0.0;
                    for (int q = 0; q <= j; ++q) d += Q[p][q] * v[q];
                    for (int q = 0; q <= j; ++q) v[q] -= d * Q[p][q];
                    R[col][p] += d;
                }
            double nv = 0.0;
            for (int q = 0; q <= j; ++q) nv += v[q] * v[q];
            nv = std::sqrt(nv);
            const double ref = col == 0 ? nv : R[0][0];
            if (!std::isfinite(nv) || !(nv > 1e-12 * ref)) ok = false;
            else {
                R[col][col] = nv;
                for (int q = 0; q <= j; ++q) Q[col][q] = v[q] / nv;
            }
        }
        if (!ok) continue;
        std::vector<std::vector<double>> Ri(j, std::vector<double>(j, 0.0));
        for (int col = 0; col < j; ++col)
            for (int q = col; q >= 0; --q) {
                double v = q == col ? 1.0 : 0.0;
                for (int p = q + 1; p <= col; ++p) v -= R[p][q] * Ri[col][p];
                Ri[col][q] = v / R[q][q];
            }
        jc[c] = j;
        Qc[c] = std::move(Q);
        Ric[c] = std::move(Ri);
        k = std::max(k, j);
    }
    if (k == 0) return;
    // output vector p of C: coefficients (q, c) = Q_c[p][q] over basis vectors q <= k; of U: Ri_c[p][q]
    const size_t per = (size_t)(k + 1) * C;
    std::vector<double> coef((size_t)2 * k * per, 0.0);
    for (int c = 0; c < C; ++c)
        for (int p = 0; p < jc[c]; ++p)
            for (int q = 0; q <= jc[c]; ++q) {
                coef[(size_t)p * per + (size_t)q * C + c] = Qc[c][p][q];
                if (q < jc[c]) coef[(size_t)(k + p) * per + (size_t)q * C + c] = Ric[c][p][q];
            }
    B.coef.ensure(coef.size());
    H2W_HIP(hipMemcpyAsync(B.coef.p, coef.data(), coef.size() * 8, hipMemcpyHostToDevice, s));
    B.U.ensure((size_t)2 * k * len);  // U, then C before it moves into the basis
    launch_fill(B.U.p, 2 * k * len, 0.0, s);
    for (int p = 0; p < k; ++p) {
        launch_bgemv16(B.V.p, len, k + 1, B.coef.p + (size_t)p * per, L.S, 1.0, B.U.p + (int64_t)(k + p) * len, L.n,
                       L.seg, s);
        launch_bgemv16(B.V.p, len, k + 1, B.coef.p + (size_t)(k + p) * per, L.S, 1.0, B.U.p + (int64_t)p * len, L.n,
                       L.seg, s);
    }
    H2W_HIP(hipMemcpyAsync(B.V.p, B.U.p + (int64_t)k * len, (size_t)k * len * 8, hipMemcpyDeviceToDevice, s));
    H2W_HIP(hipStreamSynchronize(s));
    B.rk = k;
}

}  // namespace

// Y = A X for nrhs <= kNR columns of column-major X / Y (device pointers)
void matvec_multi_dev(Handle &h, const double *X, int64_t ldx, double *Y, int64_t ldy, int nrhs, int precision,
                      bool tree) {
    check_precision_m(h, precision);
    ensure_pool_m(h, precision);
    hipStream_t s = h.stream;
    const int32_t *perm = tree ? nullptr : h.d_perm.p;
    if (is32(precision)) {
        launch_gather_m32(X, ldx, nrhs, perm, h.N, h.vec32m.p + h.off_x * kNR, s);
        run_launches_m(h, precision);
        launch_scatter_m32(h.vec32m.p + h.off_y * kNR, h.N, Y, ldy, nrhs, perm, s);
    } else {
        launch_gather_m64(X, ldx, nrhs, perm, h.N, h.vec64m.p + h.off_x * kNR, s);
        run_launches_m(h, precision);
        launch_scatter_m64(h.vec64m.p + h.off_y * kNR, h.N, Y, ldy, nrhs, perm, s);
    }
    H2W_HIP(hipGetLastError());
}

void lock_solve(Lockstep &L, const double *B, double *X, int nrhs, int max_iter, double tol, int mode,
                h2w_gmres_stats *st) {
    hipStream_t s = L.s;
    const int S = L.S, C = S * kNR;
    const int64_t len = L.n * kNR;
    const int inner_prec = mode == H2W_GMRES_FP32_INNER_IR ? H2W_FP32_INNER : H2W_FP32;
    const double t0 = now_s();
    std::vector<char> active(C, 0);
    for (int sy = 0; sy < S; ++sy)
        for (int r = 0; r < nrhs; ++r) active[sy * kNR + r] = 1;
    auto stat = [&](int c) -> h2w_gmres_stats & { return st[(c / kNR) * nrhs + c % kNR]; };
    std::memset(st, 0, sizeof(h2w_gmres_stats) * S * nrhs);
    if (mode == H2W_GMRES_FP64) {
        LockOut o = lock_gmres(L, B, X, active, max_iter, std::vector<double>(C, tol), H2W_FP64, false, nullptr);
        for (int c = 0; c < C; ++c)
            if (active[c]) {
                stat(c).iterations = o.iters[c];
                stat(c).resid = o.resid[c];
                stat(c).matvecs_fp64 = o.matvecs;
                stat(c).converged = o.conv[c];
            }
    } else {
        // FP64 refinement per column: r = b - A x on the stored FP64 operators, corrections from
        // lockstep GMRES on the FP32 (inner) operators; converged columns drop out (zero residual).
        // The first residual is the FP32 operators' own (the first correction solves the FP32
        // systems themselves); convergence is only decided on FP64 residuals.
        LockBufs &Bf = *L.buf;
        Bf.d.ensure(len);
        Bf.b.ensure(len);
        double *d = Bf.d.p, *r = Bf.b.p;
        std::vector<double> nb2(C), rr(C), normb(C), rn(C), tmp(C), nr(C), eta(C), dd(C), db(C);
        ldots(L, B, B, nb2.data());
        for (int c = 0; c < C; ++c) normb[c] = nb2[c] > 0.0 ? std::sqrt(nb2[c]) : 1.0;
        int outer = 0, total = 0, mv64 = 0, mv32 = 0;
        std::vector<char> run(C, 0);
        const bool defl = L.dk > 0 && L.dprec == inner_prec;
        bool fp64_resid = max_iter < 1;  // (no correction allowed: report x0's FP64 residual)
        while (true) {
            L.matvec(X, d, fp64_resid ? H2W_FP64 : inner_prec);
            ++(fp64_resid ? mv64 : mv32);
            if (outer == 0) {
                // x0 -> alpha_c x0 with alpha_c minimising |b_c - alpha_c A x0_c|
                ldots(L, d, d, dd.data());
                ldots(L, d, B, db.data());
                for (int c = 0; c < C; ++c) {
                    const double a = dd[c] > 0.0 ? db[c] / dd[c] : 1.0;
                    tmp[c] = std::isfinite(a) ? a : 1.0;
                }
                lscale(L, tmp, X, X);
                lscale(L, tmp, d, d);
            }
            H2W_HIP(hipMemcpyAsync(r, B, len * 8, hipMemcpyDeviceToDevice, s));
            launch_axpby(-1.0, d, 1.0, r, len, s);
            ldots(L, r, r, rr.data());
            int nrun = 0;
            for (int c = 0; c < C; ++c) {
                rn[c] = std::sqrt(rr[c]) / normb[c];
                run[c] = active[c] && (!fp64_resid || (rn[c] > tol && total < max_iter));
                nrun += run[c];
                if (active[c] && fp64_resid) {
                    stat(c).resid = rn[c];
                    stat(c).converged = rn[c] <= tol;
                }
            }
            if (nrun == 0 || (fp64_resid && outer >= 50)) break;
            for (int c = 0; c < C; ++c) {
                double fl = mode == H2W_GMRES_FP32_INNER_IR ? L.ir_floor[c / kNR] : 1e-6;
                if (outer == 0) fl = std::max(fl, std::sqrt(tol));  // (as the single-vector refinement)
                nr[c] = std::sqrt(rr[c]);
                eta[c] = run[c] ? std::min(1e-2, std::max(fl, 0.1 * tol / rn[c])) : 1.0;
                tmp[c] = run[c] ? 1.0 : 0.0;
            }
            lscale(L, tmp, r, r);  // converged columns get a zero right-hand side
            launch_fill(d, len, 0.0, s);
            // every correction starts from the shared deflation space when one was built for this
            // inner operator; otherwise later corrections recycle the first one's Krylov space (per column)
            std::vector<std::vector<double>> hk;
            LockOut o = defl ? lock_gmres(L, r, d, run, max_iter - total, eta, inner_prec, true, nr.data(), 0,
                                          nullptr, true)
                             : lock_gmres(L, r, d, run, max_iter - total, eta, inner_prec, true, nr.data(),
                                          outer == 0 ? 0 : L.buf->rk, outer == 0 ? &hk : nullptr);
            if (outer == 0 && !defl) lock_keep_space(L, hk, o);
            int most = 0;
            for (int c = 0; c < C; ++c)
                if (run[c]) {
                    stat(c).iterations += o.iters[c];
                    stat(c).outer_iterations++;
                    most = std::max(most, o.iters[c]);
                }
            total += most;
            mv32 += o.matvecs;
            launch_axpby(1.0, d, 1.0, X, len, s);
            ++outer;
            // (no iteration and no recycled or deflation projection: the corrections were zero)
            if (most == 0 && fp64_resid && !defl && L.buf->rk == 0) break;
            fp64_resid = true;
        }
        for (int c = 0; c < C; ++c)
            if (active[c]) {
                stat(c).matvecs_fp64 = mv64;
                stat(c).matvecs_fp32 = mv32;
            }
    }
    H2W_HIP(hipStreamSynchronize(s));
    const double dt = now_s() - t0;
    for (int q = 0; q < S * nrhs; ++q) st[q].seconds = dt;
}

// nrhs <= kNR solves of one system on interleaved tree-order vectors: B (device,
// read) and X (device, x0 on entry, the solutions on return)
void solve_tree_m(Handle &h, const double *B, double *X, int nrhs, int max_iter, double tol, int mode,
                  h2w_gmres_stats *st) {
    if (!h.m_nblk) {
        std::vector<int2> blk;
        std::vector<int32_t> sblk;
        lock_blocks({0, h.N}, blk, sblk);
        h.m_blk.upload(blk, h.stream);
        h.m_sblk.upload(sblk, h.stream);
        h.m_nblk = (int)blk.size();
    }
    Lockstep L;
    L.s = h.stream;
    L.n = h.N;
    L.S = 1;
    L.blk = h.m_blk.p;
    L.nblk = h.m_nblk;
    L.sblk = h.m_sblk.p;
    L.ir_floor = {std::pow(10.0, -std::max(1.0, h.ir_exp))};
    L.matvec = [&h](const double *x, double *y, int prec) { matvec_m_tree(h, x, y, prec); };
    L.buf = &h.mbuf;
    if (h.defl_k > 0 && (h.defl_prec != H2W_FP32_INNER || h.defl_gen == h.inner_gen)) {
        L.dU = h.g_defl_rm.p;
        L.dC = h.g_defl_rm.p + (int64_t)h.defl_k * h.N;
        L.dk = h.defl_k;
        L.dprec = h.defl_prec;
    }
    lock_solve(L, B, X, nrhs, max_iter, tol, mode, st);
}

}  // namespace h2w

// ------------------------------------------------------------------ C ABI: timing of the 16-wide pass
using h2w::Error;

extern "C" int h2w_bench_matvec_multi(h2w_handle *h, int nrep, int precision, h2w_bench *out) {
    try {
        if (!h || !out || nrep < 1) throw Error{H2W_EINVAL, "bad argument"};
        if (!h->assembled) throw Error{H2W_ESTATE, "not assembled"};
        if (h->cplx) throw Error{H2W_EINVAL, "complex kernels: no 16-wide product"};
        h2w::check_precision_m(*h, precision);
        H2W_HIP(hipSetDevice(h->opts.device));
        h2w::ensure_pool_m(*h, precision);
        std::memset(out, 0, sizeof(*out));
        hipStream_t s = h->stream;
        const int64_t nx = h->N * h2w::kNR;
        if (h2w::is32(precision)) h2w::launch_fill32(h->vec32m.p + h->off_x * h2w::kNR, nx, 1.0f, s);
        else h2w::launch_fill(h->vec64m.p + h->off_x * h2w::kNR, nx, 1.0, s);
        h2w::run_launches_m(*h, precision);
        h2w::Event e0, e1;
        H2W_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < nrep; ++r) h2w::run_launches_m(*h, precision);
        H2W_HIP(hipEventRecord(e1, s));
        out->ms_per_matvec = h2w::elapsed_ms(e0, e1) / nrep;
        // entry 0: the whole pass and the operator bytes it multiplies (every stored element
        // once, the symmetric blocks twice: direct and transposed products); entries 1..: the
        // launches one by one between events (eager), role 4 for the transposed products
        const bool in = precision == H2W_FP32_INNER;
        const h2w::LaunchSet &tr = in ? h->inner_trans : h->trans;
        const int es = h2w::is32(precision) ? 4 : 8;
        int64_t el = in ? h->inner_elems : h->stored_elems;
        for (auto e : tr.elems) el += e;
        out->launch_ms[0] = out->ms_per_matvec;
        out->launch_role[0] = -2;
        out->launch_bytes[0] = el * es;
        const auto st = h2w::steps_m(*h, precision);
        const int nl = (int)std::min<size_t>(st.size(), H2W_MAX_LAUNCHES - 1);
        std::vector<h2w::Event> ev(nl + 1);
        std::vector<double> acc(nl, 0.0);
        for (int r = 0; r < nrep; ++r) {
            h2w::zero_outputs_m(*h, precision);
            H2W_HIP(hipEventRecord(ev[0], s));
            for (int q = 0; q < nl; ++q) {
                h2w::run_launch_m(*h, *st[q].first, precision, st[q].second);
                H2W_HIP(hipEventRecord(ev[q + 1], s));
            }
            for (int q = 0; q < nl; ++q) acc[q] += h2w::elapsed_ms(ev[q], ev[q + 1]);
        }
        const auto &L = h2w::launches_for(*h, precision);
        for (int q = 0; q < nl; ++q) {
            const h2w::Launch *Lp = st[q].first;
            out->launch_ms[q + 1] = acc[q] / nrep;
            out->launch_role[q + 1] = st[q].second ? 4 : Lp->role;
            int64_t e = 0;
            if (st[q].second) e = tr.elems[Lp - tr.launches.data()];
            else if (in) e = h->inner.elems[Lp - L.data()];
            else e = h->launch_elems[Lp - L.data()];
            out->launch_bytes[q + 1] = e * es;
        }
        out->n_launch = nl + 1;
        return H2W_OK;
    } catch (const Error &e) {
        h2w::set_error(e.msg);
        return e.code;
    } catch (const std::exception &e) {
        h2w::set_error(e.what());
        return H2W_EHIP;
    }
}
