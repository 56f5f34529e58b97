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

void run_launches_m_eager(Handle &h, int precision) {
    const auto &L = launches_for(h, precision);
    hipStream_t s = h.stream;
    const int64_t nout = (h.vec_len - h.off_y) * kNR;
    if (precision == H2W_FP32_INNER) {
        launch_fill32(h.vec32m.p + h.off_y * kNR, nout, 0.0f, s);
        for (auto &Lh : L)
            launch_bgemm32(h.inner.d_items.p + Lh.item0, (int)Lh.nitems, h.inner.d_terms.p, h.mat32.p, h.mat32in.p,
                           h.vec32m.p, Lh.role, s);
    } else if (precision == H2W_FP32) {
        launch_fill32(h.vec32m.p + h.off_y * kNR, nout, 0.0f, s);
        for (auto &Lh : L)
            launch_bgemm32(h.d_items.p + Lh.item0, (int)Lh.nitems, h.d_terms.p, h.mat32.p, nullptr, h.vec32m.p,
                           Lh.role, s);
    } else {
        launch_fill(h.vec64m.p + h.off_y * kNR, nout, 0.0, s);
        for (auto &Lh : L)
            launch_bgemm64(h.d_items.p + Lh.item0, (int)Lh.nitems, h.d_terms.p, h.mat64.p, nullptr, h.vec64m.p,
                           Lh.role, s);
    }
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

// per-column dot products of two interleaved vectors, to the host
void dots16(Handle &h, const double *a, const double *b, double *out) {
    const int nb = mdot16_blocks(h.N);
    h.m_part.ensure((size_t)nb * kNR);
    h.m_h.ensure(kNR);
    launch_mdot16(a, 0, 1, b, h.N, h.m_part.p, nb, h.stream);
    launch_mdot16_reduce(h.m_part.p, nb, 1, h.m_h.p, h.stream);
    H2W_HIP(hipMemcpyAsync(out, h.m_h.p, kNR * 8, hipMemcpyDeviceToHost, h.stream));
    H2W_HIP(hipStreamSynchronize(h.stream));
}

void set_scale(Handle &h, const double *sc) {
    h.m_s.ensure(kNR);
    H2W_HIP(hipMemcpyAsync(h.m_s.p, sc, kNR * 8, hipMemcpyHostToDevice, h.stream));
}

void ensure_Vm(Handle &h, int64_t cols) {
    if (cols <= h.m_Vcols) return;
    const int64_t len = h.N * kNR;
    int64_t nc = std::max<int64_t>(cols, std::max<int64_t>(8, h.m_Vcols * 2));
    double *np = nullptr;
    if (hipMalloc(&np, (size_t)nc * len * 8) != hipSuccess) throw Error{H2W_ENOMEM, "GMRES basis allocation failed"};
    if (h.m_Vcols > 0)
        H2W_HIP(hipMemcpyAsync(np, h.m_V.p, (size_t)h.m_Vcols * len * 8, hipMemcpyDeviceToDevice, h.stream));
    H2W_HIP(hipStreamSynchronize(h.stream));
    h.m_V.release();
    h.m_V.p = np;
    h.m_V.n = (size_t)nc * len;
    h.m_Vcols = nc;
}

struct GmresOutM {
    int iters[kNR] = {};
    double resid[kNR] = {};
    bool conv[kNR] = {};
    int matvecs = 0;
};

// kNR unrestarted GMRES solves of A x_r = b_r in lockstep (the algorithm of the
// single-vector core: CGS2 Arnoldi, Givens, resid = |s_{i+1}| / normb).  Column r
// stops at its own tolerance tol[r]; columns with active[r] false are skipped.
// x_zero: x0 = 0 and no initial product.  normb_in[r] > 0 overrides |b_r|.
GmresOutM gmres_core_m(Handle &h, const double *b, double *x, const bool *active, int max_iter, const double *tol,
                       int prec, bool x_zero, const double *normb_in) {
    const int64_t N = h.N, len = N * kNR;
    hipStream_t s = h.stream;
    GmresOutM out;
    h.m_w.ensure(len);
    h.m_r.ensure(len);
    double *w = h.m_w.p, *r = h.m_r.p;
    double normb[kNR], beta[kNR], tmp[kNR];
    if (normb_in) std::memcpy(tmp, normb_in, sizeof(tmp));
    else {
        dots16(h, b, b, tmp);
        for (double &v : tmp) v = std::sqrt(v);
    }
    for (int c = 0; c < kNR; ++c) normb[c] = tmp[c] > 0.0 ? tmp[c] : 1.0;
    H2W_HIP(hipMemcpyAsync(r, b, len * 8, hipMemcpyDeviceToDevice, s));
    if (!x_zero) {
        matvec_m_tree(h, x, w, prec);
        out.matvecs++;
        launch_axpby(-1.0, w, 1.0, r, len, s);
    }
    dots16(h, r, r, tmp);
    bool run[kNR];
    int nrun = 0;
    for (int c = 0; c < kNR; ++c) {
        beta[c] = std::sqrt(tmp[c]);
        out.resid[c] = beta[c] / normb[c];
        run[c] = active[c] && out.resid[c] > tol[c];
        out.conv[c] = active[c] && !run[c];
        nrun += run[c];
    }
    if (nrun == 0) return out;
    const int m = max_iter;
    const size_t ld = (size_t)m + 1;
    // Hessenberg columns as they are produced: column j holds rows 0..j+1 of every solve
    std::vector<std::vector<double>> H;
    std::vector<double> sv(ld * kNR, 0.0), cs(ld * kNR, 0.0), sn(ld * kNR, 0.0);
    auto Hh = [&](int c, int i, int j) -> double & { return H[j][(size_t)i * kNR + c]; };
    std::vector<double> Y((size_t)m * kNR, 0.0);  // update coefficients, (j, c) at j * kNR + c
    ensure_Vm(h, std::min<int64_t>(m + 1, 16));
    for (int c = 0; c < kNR; ++c) {
        tmp[c] = run[c] ? 1.0 / beta[c] : 0.0;
        sv[c * ld] = beta[c];
    }
    set_scale(h, tmp);
    launch_scale16(r, h.m_s.p, h.m_V.p, N, s);
    const int nb = mdot16_blocks(N);
    h.m_part.ensure((size_t)nb * (m + 2) * kNR);
    h.m_h.ensure((size_t)(2 * (m + 2) + 2) * kNR);
    std::vector<double> hh((size_t)(2 * (m + 2) + 2) * kNR);
    int last = 0;
    for (int i = 0; i < m && nrun > 0; ++i) {
        ensure_Vm(h, i + 2);
        double *V = h.m_V.p;
        matvec_m_tree(h, V + (int64_t)i * len, w, prec);
        out.matvecs++;
        const int k = i + 1;
        launch_mdot16(V, len, k, w, N, h.m_part.p, nb, s);
        launch_mdot16_reduce(h.m_part.p, nb, k, h.m_h.p, s);
        launch_mgemv16(V, len, k, h.m_h.p, -1.0, w, N, s);
        launch_mdot16(V, len, k, w, N, h.m_part.p, nb, s);
        launch_mdot16_reduce(h.m_part.p, nb, k, h.m_h.p + (size_t)k * kNR, s);
        launch_mgemv16(V, len, k, h.m_h.p + (size_t)k * kNR, -1.0, w, N, s);
        launch_mdot16(w, 0, 1, w, N, h.m_part.p, nb, s);
        launch_mdot16_reduce(h.m_part.p, nb, 1, h.m_h.p + (size_t)2 * k * kNR, s);
        H2W_HIP(hipMemcpyAsync(hh.data(), h.m_h.p, (size_t)(2 * k + 1) * kNR * 8, hipMemcpyDeviceToHost, s));
        H2W_HIP(hipStreamSynchronize(s));
        last = i;
        H.emplace_back((size_t)(i + 2) * kNR, 0.0);
        for (int c = 0; c < kNR; ++c) {
            tmp[c] = 0.0;
            if (!run[c]) continue;
            for (int q = 0; q < k; ++q) Hh(c, q, i) = hh[(size_t)q * kNR + c] + hh[(size_t)(k + q) * kNR + c];
            Hh(c, i + 1, i) = std::sqrt(hh[(size_t)2 * k * kNR + c]);
            if (Hh(c, i + 1, i) > 0.0) tmp[c] = 1.0 / Hh(c, i + 1, i);
            double *csc = &cs[c * ld], *snc = &sn[c * ld], *svc = &sv[c * ld];
            for (int q = 0; q < i; ++q) app_rot(Hh(c, q, i), Hh(c, q + 1, i), csc[q], snc[q]);
            gen_rot(Hh(c, i, i), Hh(c, i + 1, i), csc[i], snc[i]);
            app_rot(Hh(c, i, i), Hh(c, i + 1, i), csc[i], snc[i]);
            app_rot(svc[i], svc[i + 1], csc[i], snc[i]);
            out.resid[c] = std::fabs(svc[i + 1]) / normb[c];
            out.iters[c] = i + 1;
            if (out.resid[c] < tol[c] || i == m - 1 || tmp[c] == 0.0) {
                // back-substitution for this column; its later basis vectors stay zero
                std::vector<double> y(svc, svc + i + 1);
                for (int q = i; q >= 0; q--) {
                    y[q] /= Hh(c, q, q);
                    for (int jj = q - 1; jj >= 0; jj--) y[jj] -= Hh(c, jj, q) * y[q];
                }
                for (int q = 0; q <= i; ++q) Y[(size_t)q * kNR + c] = y[q];
                out.conv[c] = out.resid[c] < tol[c];
                run[c] = false;
                tmp[c] = 0.0;
                --nrun;
            }
        }
        if (nrun > 0) {
            set_scale(h, tmp);
            launch_scale16(w, h.m_s.p, V + (int64_t)(i + 1) * len, N, s);
        }
    }
    // x += V Y over the basis vectors every column used
    H2W_HIP(hipMemcpyAsync(h.m_h.p, Y.data(), (size_t)(last + 1) * kNR * 8, hipMemcpyHostToDevice, s));
    launch_mgemv16(h.m_V.p, len, last + 1, h.m_h.p, 1.0, x, N, s);
    H2W_HIP(hipStreamSynchronize(s));
    return out;
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

// nrhs <= kNR solves on interleaved tree-order vectors: B (device, read) and X
// (device, x0 on entry, the solutions on return).  Per-column statistics in st.
void solve_tree_m(Handle &h, const double *B, double *X, int nrhs, int max_iter, double tol, int mode,
                  h2w_gmres_stats *st) {
    hipStream_t s = h.stream;
    const int64_t N = h.N, len = N * kNR;
    const int inner_prec = mode == H2W_GMRES_FP32_INNER_IR ? H2W_FP32_INNER : H2W_FP32;
    const double t0 = now_s();
    bool active[kNR];
    double tolv[kNR];
    for (int c = 0; c < kNR; ++c) {
        active[c] = c < nrhs;
        tolv[c] = tol;
    }
    for (int c = 0; c < nrhs; ++c) std::memset(&st[c], 0, sizeof(h2w_gmres_stats));
    if (mode == H2W_GMRES_FP64) {
        GmresOutM o = gmres_core_m(h, B, X, active, max_iter, tolv, H2W_FP64, false, nullptr);
        for (int c = 0; c < nrhs; ++c) {
            st[c].iterations = o.iters[c];
            st[c].resid = o.resid[c];
            st[c].matvecs_fp64 = o.matvecs;
            st[c].converged = o.conv[c];
        }
    } else {
        // FP64 refinement per column: r = b - A x on the stored FP64 operators, corrections from
        // kNR-wide GMRES on the FP32 (inner) operator; converged columns drop out (zero residual)
        h.m_d.ensure(len);
        h.m_b.ensure(len);
        double *d = h.m_d.p, *r = h.m_b.p;
        double nb2[kNR], rr[kNR], normb[kNR], rn[kNR], tmp[kNR];
        dots16(h, B, B, nb2);
        for (int c = 0; c < kNR; ++c) normb[c] = nb2[c] > 0.0 ? std::sqrt(nb2[c]) : 1.0;
        int outer = 0, total = 0, mv64 = 0, mv32 = 0;
        const double floor_eta = mode == H2W_GMRES_FP32_INNER_IR ? std::pow(10.0, -1.0 * h.opts.ir_tol_pow) : 1e-6;
        bool run[kNR];
        while (true) {
            matvec_m_tree(h, X, d, H2W_FP64);
            ++mv64;
            if (outer == 0) {
                // x0 -> alpha_r x0 with alpha_r minimising |b_r - alpha_r A x0_r|
                double dd[kNR], db[kNR];
                dots16(h, d, d, dd);
                dots16(h, d, B, db);
                for (int c = 0; c < kNR; ++c) {
                    const double a = dd[c] > 0.0 ? db[c] / dd[c] : 1.0;
                    tmp[c] = std::isfinite(a) ? a : 1.0;
                }
                set_scale(h, tmp);
                launch_scale16(X, h.m_s.p, X, N, s);
                launch_scale16(d, h.m_s.p, d, N, s);
            }
            H2W_HIP(hipMemcpyAsync(r, B, len * 8, hipMemcpyDeviceToDevice, s));
            launch_axpby(-1.0, d, 1.0, r, len, s);
            dots16(h, r, r, rr);
            int nrun = 0;
            for (int c = 0; c < kNR; ++c) {
                rn[c] = std::sqrt(rr[c]) / normb[c];
                run[c] = active[c] && rn[c] > tol && total < max_iter;
                nrun += run[c];
                if (c < nrhs) {
                    st[c].resid = rn[c];
                    st[c].converged = rn[c] <= tol;
                }
            }
            if (nrun == 0 || outer >= 50) break;
            double nr[kNR], eta[kNR];
            for (int c = 0; c < kNR; ++c) {
                nr[c] = std::sqrt(rr[c]);
                eta[c] = run[c] ? std::min(1e-2, std::max(floor_eta, 0.1 * tol / rn[c])) : 1.0;
                tmp[c] = run[c] ? 1.0 : 0.0;
            }
            // converged columns get a zero right-hand side
            set_scale(h, tmp);
            launch_scale16(r, h.m_s.p, r, N, s);
            launch_fill(d, len, 0.0, s);
            GmresOutM o = gmres_core_m(h, r, d, run, max_iter - total, eta, inner_prec, true, nr);
            int most = 0;
            for (int c = 0; c < nrhs; ++c)
                if (run[c]) {
                    st[c].iterations += o.iters[c];
                    st[c].outer_iterations++;
                    most = std::max(most, o.iters[c]);
                }
            total += most;
            mv32 += o.matvecs;
            launch_axpby(1.0, d, 1.0, X, len, s);
            ++outer;
            if (most == 0) break;
        }
        for (int c = 0; c < nrhs; ++c) {
            st[c].matvecs_fp64 = mv64;
            st[c].matvecs_fp32 = mv32;
        }
    }
    H2W_HIP(hipStreamSynchronize(s));
    const double dt = now_s() - t0;
    for (int c = 0; c < nrhs; ++c) st[c].seconds = dt;
}

}  // namespace h2w

// ------------------------------------------------------------------ C ABI: timing of the 16-wide pass
using h2w::Error;

extern "C" int h2w_bench_matvec_multi(h2w_handle *h, int nrep, int precision, h2w_bench *out) {
    try {
        if (!h || !out || nrep < 1) throw Error{H2W_EINVAL, "bad argument"};
        if (!h->assembled) throw Error{H2W_ESTATE, "not assembled"};
        h2w::check_precision_m(*h, precision);
        H2W_HIP(hipSetDevice(h->opts.device));
        h2w::ensure_pool_m(*h, precision);
        std::memset(out, 0, sizeof(*out));
        hipStream_t s = h->stream;
        const int64_t nx = h->N * h2w::kNR;
        if (h2w::is32(precision)) h2w::launch_fill32(h->vec32m.p + h->off_x * h2w::kNR, nx, 1.0f, s);
        else h2w::launch_fill(h->vec64m.p + h->off_x * h2w::kNR, nx, 1.0, s);
        h2w::run_launches_m(*h, precision);
        hipEvent_t e0, e1;
        H2W_HIP(hipEventCreate(&e0));
        H2W_HIP(hipEventCreate(&e1));
        H2W_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < nrep; ++r) h2w::run_launches_m(*h, precision);
        H2W_HIP(hipEventRecord(e1, s));
        H2W_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        H2W_HIP(hipEventElapsedTime(&ms, e0, e1));
        out->ms_per_matvec = ms / nrep;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        out->n_launch = 1;
        out->launch_ms[0] = out->ms_per_matvec;
        out->launch_role[0] = -2;
        return H2W_OK;
    } catch (const Error &e) {
        h2w::set_error(e.msg);
        return e.code;
    } catch (const std::exception &e) {
        h2w::set_error(e.what());
        return H2W_EHIP;
    }
}
