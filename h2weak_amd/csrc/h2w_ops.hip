// h2w_ops.hip -- host orchestration of libh2w: assembly, matvec, GMRES and
// the C ABI declared in include/h2w.h.
//
// Device layout (one handle):
//   points  px/py/pz    tree-order coordinates (FP64)
//   mat64 / mat32       one pool with every stored operator, column-major,
//                       32-element aligned; mat32 is the FP32 shadow at the
//                       same element offsets
//   vec64 / vec32       one pool with x and y in tree order, then the outgoing
//                       (o) and incoming (u) skeleton vectors of every family,
//                       level and box, then the semi-nested pair vectors (t)
// A matvec is a fixed sequence of batched-GEMV launches over work lists built
// once at assembly (up pass per level, one M2L + near-field launch, down pass
// per level, leaf down, semi-nested CUR down).
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <thread>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "h2w_handle.hpp"

namespace h2w {

static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }

}  // namespace h2w

namespace h2w {
namespace {

int ref_nlevels(int dim, int64_t N, int n_max, bool grid) {
    if (grid && dim > 1) {
        int n = (int)std::llround(dim == 2 ? std::sqrt((double)N) : std::cbrt((double)N));
        int p = (int)std::llround(dim == 2 ? std::sqrt((double)n_max) : std::cbrt((double)n_max));
        if (dim == 2) return (int)std::ceil(2.0 * std::log(double(n) / p) / std::log(4));
        return (int)std::ceil(3 * std::log(double(n) / p) / std::log(8));
    }
    return (int)std::ceil(std::log(double(N) / n_max) / std::log((double)(1 << dim)));
}

// GL-16 nodes/weights by Newton on P_16, for the IE diagonal's Kii
// (*/codes/integral{1,2,3}d.hpp; Gauss_Legendre_Nodes_and_Weights.hpp:471-523)
void gl16(double *nodes, double *weights) {
    const int n = 16;
    double pos[8], wts[8];
    for (int i = 0; i < 8; ++i) {
        double x = std::cos(kPI * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-17) break;
        }
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= n; ++k) {
            double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (x * p1 - p0) / (x * x - 1.0);
        pos[i] = x;
        wts[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    for (int i = 0; i < 8; ++i) {
        int s = 7 - i;
        nodes[2 * i] = -pos[s];
        nodes[2 * i + 1] = pos[s];
        weights[2 * i] = weights[2 * i + 1] = wts[s];
    }
}

// Diagonal of the complex kernels.  3D Helmholtz (choice 2): 500 sqrt(N) (3D/codes/kernel.hpp:132-144).
// 2D Hankel (choice 10): (1 + 15 log N + Kii) / h^2 with h = N^(-1/2) and the complex
// Kii = 4 GL16^2 of (i/4) H0^(1)(R) = (i J0(R) - Y0(R)) / 4 on [0, h/2]^2 (2D/codes/kernel.hpp:126-135,
// 2D/codes/integral2d.hpp:16-57,60-90).
void diag_value_c(int dim, int choice, int64_t N, double &re, double &im) {
    if (choice == 2) {
        re = 500.0 * std::sqrt(double(N));
        im = 0.0;
        return;
    }
    double nd[16], w[16];
    gl16(nd, w);
    const double h = 1.0 / std::sqrt(double(N));
    const double cx = 0.5 * (0.5 * h), Lx = cx;
    double kr = 0.0, ki = 0.0;
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 16; ++j) {
            const double tx = cx * nd[i] + Lx, ty = cx * nd[j] + Lx;
            const double R = std::sqrt(tx * tx + ty * ty);
            kr += w[i] * w[j] * (-::y0(R) / 4.0);
            ki += w[i] * w[j] * (::j0(R) / 4.0);
        }
    kr *= cx * cx * 4;
    ki *= cx * cx * 4;
    const double h2 = std::pow(h, 2.0);
    re = (1.0 + 15.0 * std::log(double(N)) + kr) / h2;
    im = ki / h2;
    (void)dim;
}

// diagonal value of A (the i == j branch of getMatrixEntry):
// IE: (1 + Kii) / h^dim with Kii = 2^dim * GL16^dim of the integrand on [0, h/2]^dim,
// h = N^(-1/dim) (1D/codes/kernel.hpp:95-114, 2D/codes/kernel.hpp:101-122, 3D/codes/kernel.hpp:101-129);
// RBF: 1D sqrt(2N) (reference_compat) / 2D 1 / 3D sqrt(N).
double diag_value(int dim, int choice, int64_t N, int compat) {
    if (choice == 1) {
        if (dim == 1) return std::sqrt(double(compat ? 2 * N : N));
        if (dim == 2) return 1.0;
        return std::sqrt(double(N));
    }
    double nd[16], w[16];
    gl16(nd, w);
    double h = dim == 1 ? 1.0 / double(N) : (dim == 2 ? 1.0 / std::sqrt(double(N)) : 1.0 / std::cbrt(double(N)));
    double cx = 0.5 * (0.5 * h), Lx = cx;
    double res = 0.0;
    if (dim == 1) {
        for (int i = 0; i < 16; ++i) res += w[i] * std::log(std::fabs(1.0 * (cx * nd[i] + Lx)));
        res *= cx;
        res *= 2;
        return (1.0 + res) / h;
    }
    if (dim == 2) {
        double c = -1.0 / (2 * kPI);
        for (int i = 0; i < 16; ++i)
            for (int j = 0; j < 16; ++j) {
                double tx = cx * nd[i] + Lx, ty = cx * nd[j] + Lx;
                res += w[i] * w[j] * (c * 0.5 * std::log(tx * tx + ty * ty));
            }
        res *= (cx * cx);
        res *= 4;
        return (1.0 + res) / std::pow(h, 2.0);
    }
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 16; ++j)
            for (int k = 0; k < 16; ++k) {
                double tx = cx * nd[i] + Lx, ty = cx * nd[j] + Lx, tz = cx * nd[k] + Lx;
                res += (w[i] * w[j] * w[k] * (1.0 / ((4 * kPI) * std::sqrt(tx * tx + ty * ty + tz * tz))));
            }
    res *= (cx * cx * cx);
    res *= 8;
    return (1.0 + res) / std::pow(h, 3.0);
}

// ------------------------------------------------------------------ batched ACA driver
// Runs every block through the device ACA.  A batch shares one workspace pool
// from which each block takes one (U, V) column pair per step; batches are
// admitted by an estimated rank (rank_hint), and a block that runs out of pool
// is re-run in a later batch with a larger estimate (alone if needed).
struct AcaPool {
    DevBuf<double> ws;
    DevBuf<unsigned long long> top;
    int64_t cap = 0;  // doubles
};

// The workspace holds the U and V columns of every block of a batch.  It is sized
// to what the blocks at hand are estimated to need (within the budget: opts or
// half of the free device memory, at most 96 GB) and grows when a later call
// needs more, so small systems do not pay for a large allocation.
void ensure_pool(Handle &h, AcaPool &P, int64_t need_doubles) {
    int64_t budget = h.opts.workspace_bytes;
    if (budget <= 0) {
        size_t fr = 0, tot = 0;
        H2W_HIP(hipMemGetInfo(&fr, &tot));
        budget = std::min<int64_t>((int64_t)96 << 30, (int64_t)(fr / 2) + P.cap * 8);
        budget = std::max<int64_t>(budget, (int64_t)64 << 20);
    }
    const int64_t want = std::min<int64_t>(budget / 8, std::max<int64_t>(need_doubles + need_doubles / 4, 8 << 20));
    if (P.cap >= want) return;
    P.ws.release();
    P.cap = want;
    P.ws.alloc(P.cap);
    if (!P.top.p) P.top.alloc(1);
}

// Geometry features of an ACA block for the rank predictor: family, level, row and
// column counts, the box's and the partners' point counts, the interaction set's
// size and centre separations (in box half-widths), how much of their boxes the
// points fill, and the box centre.  All of order one.
void block_features(const Handle &h, int fam, int j, int b, const std::vector<int32_t> &partners, int64_t m,
                    int64_t n, AcaBlock &B) {
    const Plan &P = h.plan;
    const double hw = P.L / double(1 << j);
    auto centre = [&](int bx, int d) { return d < P.dim ? -P.L + (2.0 * P.icoord[j][bx][d] + 1.0) * hw : 0.0; };
    auto fill = [&](int bx) {
        const auto &bb = P.bbox[j][bx];
        if (bb[0] > bb[3]) return 0.0;
        double e = 0.0;
        for (int d = 0; d < P.dim; ++d) e += bb[3 + d] - bb[d];
        return e / (P.dim * 2.0 * hw);
    };
    double sum_d = 0.0, min_d = 0.0, sum_fill = 0.0;
    int64_t pcount = 0;
    for (size_t q = 0; q < partners.size(); ++q) {
        const int p = partners[q];
        double d2 = 0.0;
        for (int d = 0; d < 3; ++d) d2 += (centre(p, d) - centre(b, d)) * (centre(p, d) - centre(b, d));
        const double dist = std::sqrt(d2) / hw;
        sum_d += dist;
        min_d = q == 0 ? dist : std::min(min_d, dist);
        sum_fill += fill(p);
        pcount += P.bcount[j][p];
    }
    const double np = (double)partners.size();
    float *f = B.feat.data();
    f[0] = fam == 0;
    f[1] = fam == 1;
    f[2] = fam == 2;
    f[3] = (float)(j / 8.0);
    f[4] = (float)(std::log2(1.0 + m) / 16.0);
    f[5] = (float)(std::log2(1.0 + n) / 20.0);
    f[6] = (float)(std::log2(1.0 + P.bcount[j][b]) / 16.0);
    f[7] = (float)(std::log2(1.0 + pcount) / 20.0);
    f[8] = (float)(np / 64.0);
    f[9] = (float)(np > 0 ? sum_d / np / 8.0 : 0.0);
    f[10] = (float)(min_d / 8.0);
    f[11] = (float)fill(b);
    f[12] = (float)(np > 0 ? sum_fill / np : 0.0);
    for (int d = 0; d < 3; ++d) f[13 + d] = (float)(centre(b, d) / P.L);
    B.has_feat = true;
}

// The wide cross of the first nh jobs of a batch (device copies at djobs).  Blocks with a
// fixed-rank phase (AcaJob.pad = kfix > 1, from the rank model) first take kfix pivot steps
// without the normS update and stopping test -- every launch enqueued at once, no host reads
// in between -- and are then judged by the reference's own stopping rule, applied to all their
// steps at once: the normS sequence follows from the Gram products |<U_q,U_k>||<V_q,V_k>| of
// their terms, the block ends at the first step the rule stops (the pivots up to there are the
// ones the adaptive cross takes, since the rule never influences a pivot) or, when the rule
// holds for every step (rank under-predicted), it continues adaptively from there.  Adaptive
// steps are enqueued 16 at a time; between groups the host reads how many blocks are done.
// Every step advances a block's rank, moves its first-row search on, or finishes it, so the
// loop ends within max(m) + max(tcap) + 1 steps.
void run_wide(Handle &h, const std::vector<AcaJob> &jobs, size_t nh, const AcaJob *djobs, const int32_t *didx,
              AcaPool &pool, int64_t *dtab, uint8_t *dflags, int32_t *drank, int32_t *dstat, int32_t *drp,
              int32_t *dcp, double *dhist, double tol) {
    hipStream_t s = h.stream;
    const int nb = (int)nh;
    int maxm = 0, maxn = 0, maxt = 0, maxfix = 0;
    for (size_t q = 0; q < nh; ++q) {
        maxm = std::max(maxm, jobs[q].m);
        maxn = std::max(maxn, jobs[q].n);
        maxt = std::max(maxt, jobs[q].tcap);
        maxfix = std::max(maxfix, jobs[q].pad);
    }
    const int pstride = std::max(1, std::max((maxm + 2047) / 2048, (maxn + 2047) / 2048));
    const int dstride = std::max(maxt, 1);
    DevBuf<WideState> st;
    DevBuf<WidePart> part;
    DevBuf<double> dots;
    DevBuf<int32_t> ndone;
    st.alloc(nb);
    part.alloc((size_t)nb * pstride);
    dots.alloc((size_t)nb * dstride);
    ndone.alloc(1);
    launch_wide_init(djobs, nb, maxm + maxn, st.p, dflags, dtab, pool.top.p, pool.cap, s);
    const int64_t limit = (int64_t)maxm + maxt + 1;
    int64_t steps = 0;
    if (maxfix > 1) {
        // ---- fixed phase (a block whose first rows are zero spends steps on its first-row search:
        // enqueue more until every block has parked or ended)
        std::vector<WideState> hs(nb);
        int64_t todo = maxfix;
        while (true) {
            for (int64_t c = 0; c < todo; ++c, ++steps)
                launch_wide_fixed_step(djobs, nb, maxm, maxn, didx, pool.ws.p, dtab, dflags, st.p, part.p, pstride,
                                       drp, dcp, dhist, pool.top.p, pool.cap, h.kp, tol, s);
            H2W_HIP(hipGetLastError());
            H2W_HIP(hipMemcpyAsync(hs.data(), st.p, (size_t)nb * sizeof(WideState), hipMemcpyDeviceToHost, s));
            H2W_HIP(hipStreamSynchronize(s));
            int running = 0;
            for (auto &w : hs) running += !w.done;
            if (running == 0) break;
            if (steps > limit) throw Error{H2W_EHIP, "wide cross (fixed phase) did not finish within its step bound"};
            todo = 16;
        }
        // ---- Gram products of the judged blocks (status 2: parked at kfix, 3: ended inside the phase)
        std::vector<GramJob> gj;
        std::vector<int64_t> poff(nb, -1);
        int64_t plen = 0;
        for (int b = 0; b < nb; ++b) {
            if (hs[b].status != 2 && hs[b].status != 3) continue;
            const int K = hs[b].rank;
            poff[b] = plen;
            plen += (int64_t)K * (K + 1) / 2;
            const int nt = (K + 63) / 64;
            for (int ki = 0; ki < nt; ++ki)
                for (int qi = 0; qi <= ki; ++qi) gj.push_back(GramJob{b, qi, ki, 0, poff[b]});
        }
        std::vector<int32_t> fin(nb, -2);
        std::vector<double> nrm(nb, 0.0);
        if (!gj.empty()) {
            DevBuf<GramJob> dgj;
            DevBuf<double> dprod;
            dgj.upload(gj, s);
            dprod.alloc(std::max<int64_t>(plen, 1));
            launch_wide_gram(dgj.p, (int)gj.size(), djobs, st.p, pool.ws.p, dtab, dprod.p, s);
            H2W_HIP(hipGetLastError());
            std::vector<double> prod(plen);
            H2W_HIP(hipMemcpyAsync(prod.data(), dprod.p, plen * 8, hipMemcpyDeviceToHost, s));
            H2W_HIP(hipStreamSynchronize(s));
            for (int b = 0; b < nb; ++b) {
                if (poff[b] < 0) continue;
                const int K = hs[b].rank;
                const int64_t ho = 2 * jobs[b].piv;
                std::vector<double> hb(2 * (size_t)K);
                H2W_HIP(hipMemcpyAsync(hb.data(), dhist + ho, 2 * (size_t)K * 8, hipMemcpyDeviceToHost, s));
                H2W_HIP(hipStreamSynchronize(s));
                // the reference's recurrence (2D/codes/aca.hpp:404-409): the first term enters with
                // normS = 0; term k adds (|u_k||v_k|)^2 + 2 sum_{q <= k} |<U_q,u_k>||<V_q,v_k>|
                const int mn = std::min(jobs[b].m, jobs[b].n);
                double normS = 0.0;
                int stop = -1;
                for (int k = 0; k < K; ++k) {
                    const double uv = hb[2 * k];
                    if (k > 0) {
                        double S = 0.0;
                        const double *pk = prod.data() + poff[b] + (int64_t)k * (k + 1) / 2;
                        for (int q = 0; q <= k; ++q) S += pk[q];
                        normS = std::sqrt(normS * normS + uv * uv + 2.0 * S);
                    }
                    hb[2 * k + 1] = normS;
                    if (!(uv > tol * normS && k + 1 < mn)) {
                        stop = k + 1;
                        break;
                    }
                }
                const int kept = stop > 0 ? stop : K;
                if (stop > 0 || hs[b].status == 3) {
                    fin[b] = kept;
                    h.rank_fix_extra_steps += K - kept;
                } else {
                    fin[b] = -1;  // continue adaptively
                    nrm[b] = normS;
                    h.rank_fix_resumed++;
                }
                h.rank_fix_blocks++;
                H2W_HIP(hipMemcpyAsync(dhist + ho, hb.data(), 2 * (size_t)kept * 8, hipMemcpyHostToDevice, s));
                H2W_HIP(hipStreamSynchronize(s));
            }
            DevBuf<int32_t> dfin;
            DevBuf<double> dnrm;
            dfin.upload(fin, s);
            dnrm.upload(nrm, s);
            launch_wide_judge(djobs, nb, st.p, dfin.p, dnrm.p, dtab, pool.top.p, pool.cap, s);
            H2W_HIP(hipGetLastError());
            H2W_HIP(hipStreamSynchronize(s));
        }
    }
    while (true) {
        launch_wide_finish(st.p, nb, drank, dstat, ndone.p, s);
        H2W_HIP(hipGetLastError());
        int nd = 0;
        H2W_HIP(hipMemcpyAsync(&nd, ndone.p, 4, hipMemcpyDeviceToHost, s));
        H2W_HIP(hipStreamSynchronize(s));
        if (nd == nb) break;
        if (steps > limit) throw Error{H2W_EHIP, "wide cross did not finish within its step bound"};
        for (int c = 0; c < 16; ++c, ++steps)
            launch_wide_step(djobs, nb, maxm, maxn, (int)std::min<int64_t>(steps + 1, dstride), didx, pool.ws.p, dtab,
                             dflags, st.p, part.p, pstride, dots.p, dstride, drp, dcp, dhist, pool.top.p, pool.cap,
                             h.kp, tol, s);
    }
    // sampled entries of the fixed-rank blocks against their cross approximation
    if (maxfix > 1) {
        std::vector<int> sel;
        for (int b = 0; b < nb; ++b)
            if (jobs[b].pad > 1) sel.push_back(b);
        DevBuf<double> e2, a2;
        e2.alloc(nb);
        a2.alloc(nb);
        launch_aca_sample(djobs, nb, drank, didx, pool.ws.p, dtab, h.kp, 256, e2.p, a2.p, s);
        H2W_HIP(hipGetLastError());
        std::vector<double> he(nb), ha(nb);
        std::vector<int32_t> hst(nb);
        H2W_HIP(hipMemcpyAsync(he.data(), e2.p, nb * 8, hipMemcpyDeviceToHost, s));
        H2W_HIP(hipMemcpyAsync(ha.data(), a2.p, nb * 8, hipMemcpyDeviceToHost, s));
        H2W_HIP(hipMemcpyAsync(hst.data(), dstat, nb * 4, hipMemcpyDeviceToHost, s));
        H2W_HIP(hipStreamSynchronize(s));
        for (int b : sel)
            if (hst[b] == 0 && ha[b] > 0.0)
                h.rank_fix_sampled_err = std::max(h.rank_fix_sampled_err, std::sqrt(he[b] / ha[b]));
    }
}

void run_aca(Handle &h, std::vector<AcaBlock> &blocks, int rank_hint, AcaPool &pool) {
    if (blocks.empty()) return;
    const double tol = std::pow(10.0, -1.0 * h.tol_pow);
    hipStream_t s = h.stream;
    const bool verbose = std::getenv("H2W_VERBOSE") != nullptr;
    std::vector<int> est(blocks.size());
    for (size_t b = 0; b < blocks.size(); ++b) {
        int mn = (int)std::min(blocks[b].rows.size(), blocks[b].cols.size());
        est[b] = std::min(mn, std::max(1, rank_hint));
    }
    std::vector<size_t> todo(blocks.size());
    std::iota(todo.begin(), todo.end(), 0);
    // learned rank estimates: the workspace each block reserves, and rank-homogeneous batches
    // (blocks in descending predicted rank); the cross itself still stops at the tolerance
    bool pred = h.rank_model.p != nullptr;
    for (auto &B : blocks) pred = pred && B.has_feat;
    std::vector<float> plog;
    if (pred) {
        const int nb = (int)blocks.size();
        std::vector<float> F((size_t)nb * kRankFeatures);
        for (int b = 0; b < nb; ++b) std::copy(blocks[b].feat.begin(), blocks[b].feat.end(), &F[(size_t)b * kRankFeatures]);
        DevBuf<float> dF, dO;
        dF.upload(F, s);
        dO.alloc(nb);
        launch_rank_mlp(dF.p, nb, h.rank_model.p, dO.p, s);
        H2W_HIP(hipGetLastError());
        plog.resize(nb);
        H2W_HIP(hipMemcpyAsync(plog.data(), dO.p, (size_t)nb * 4, hipMemcpyDeviceToHost, s));
        H2W_HIP(hipStreamSynchronize(s));
        for (int b = 0; b < nb; ++b) {
            const int mn = (int)std::min(blocks[b].rows.size(), blocks[b].cols.size());
            const double pr = std::max(0.0, std::exp2((double)plog[b]) - 1.0);
            est[b] = std::max(1, std::min(mn, (int)std::ceil(1.2 * pr) + 4));
        }
        std::stable_sort(todo.begin(), todo.end(), [&](size_t a, size_t c) { return est[a] > est[c]; });
    }
    const std::vector<int> est0 = est;
    // fixed-rank phase of the wide cross: 3 % above the predicted rank plus 2 steps (H2W_NO_FIXED_RANK
    // keeps the adaptive cross throughout)
    const bool fixed_off = std::getenv("H2W_NO_FIXED_RANK") != nullptr;
    auto kfix_of = [&](size_t b) {
        const double pr = pred ? std::max(0.0, std::exp2((double)plog[b]) - 1.0) : 0.0;
        return std::max(2, (int)std::lround(1.03 * pr) + 2);
    };
    // complex kernels: workspace, L and R in complex elements (two doubles each), every block through
    // the one-workgroup cross
    const bool cx = h.cplx;
    const int esz = cx ? 2 : 1;
    int64_t need = 0;
    for (size_t b = 0; b < blocks.size(); ++b) need += (int64_t)(blocks[b].rows.size() + blocks[b].cols.size()) * est[b];
    ensure_pool(h, pool, need * esz);
    h.aca_pool_bytes = std::max<int64_t>(h.aca_pool_bytes, pool.cap * 8);
    DevBuf<uint8_t> flags;
    DevBuf<int32_t> didx, drank, dstat, drp, dcp;
    DevBuf<double> dhist;
    DevBuf<int64_t> dtab;
    DevBuf<AcaJob> djobs;
    DevBuf<LrJob> dlr;
    DevBuf<double> dlrout;
    const size_t kBig = 16384;    // m + n above which a block gets a 1024-thread workgroup
    // m + n above which a block's cross is spread over many workgroups (H2W_WIDE_MIN overrides)
    const size_t kHuge = cx ? SIZE_MAX
                            : std::getenv("H2W_WIDE_MIN") ? (size_t)std::atoll(std::getenv("H2W_WIDE_MIN")) : 8192;
    while (!todo.empty()) {
        std::vector<size_t> rerun;
        size_t pos = 0;
        while (pos < todo.size()) {
            std::vector<size_t> batch;
            int64_t need = 0, isz = 0;
            while (pos < todo.size()) {
                size_t b = todo[pos];
                const AcaBlock &B = blocks[b];
                int64_t nb = (int64_t)(B.rows.size() + B.cols.size()) * est[b];
                if (!batch.empty() && need + nb > pool.cap / esz) break;
                batch.push_back(b);
                need += nb;
                isz += (int64_t)(B.rows.size() + B.cols.size());
                ++pos;
                if (batch.size() >= 65535) break;
            }
            // huge blocks first (wide cross), then big ones (1024-thread workgroups), each group contiguous
            auto mn_of = [&](size_t b) { return blocks[b].rows.size() + blocks[b].cols.size(); };
            std::stable_partition(batch.begin(), batch.end(), [&](size_t b) { return mn_of(b) > kBig; });
            std::stable_partition(batch.begin(), batch.end(), [&](size_t b) { return mn_of(b) > kHuge; });
            size_t nhuge = 0, nbig = 0;
            while (nhuge < batch.size() && mn_of(batch[nhuge]) > kHuge) ++nhuge;
            nbig = nhuge;
            while (nbig < batch.size() && mn_of(batch[nbig]) > kBig) ++nbig;
            std::vector<AcaJob> jobs(batch.size());
            std::vector<int32_t> idx;
            idx.reserve(isz);
            int64_t fo = 0, po = 0, to = 0;
            for (size_t q = 0; q < batch.size(); ++q) {
                const AcaBlock &B = blocks[batch[q]];
                AcaJob &J = jobs[q];
                J.m = (int32_t)B.rows.size();
                J.n = (int32_t)B.cols.size();
                J.tcap = std::min(std::min(J.m, J.n), 8192);
                // wide blocks with a predicted rank: a fixed-rank phase of about that many steps
                J.pad = (pred && q < nhuge && !fixed_off) ? std::min(J.tcap, kfix_of(batch[q])) : 0;
                J.rows = (int64_t)idx.size();
                idx.insert(idx.end(), B.rows.begin(), B.rows.end());
                J.cols = (int64_t)idx.size();
                idx.insert(idx.end(), B.cols.begin(), B.cols.end());
                J.tab = to;
                to += 2LL * J.tcap;
                J.wsF = fo;
                fo += align_up((int64_t)(J.m + J.n), 64);
                J.piv = po;
                po += J.tcap;
            }
            flags.ensure(std::max<int64_t>(fo, 1));
            dtab.ensure(std::max<int64_t>(to, 1));
            didx.upload(idx, s);
            djobs.upload(jobs, s);
            drank.ensure(batch.size());
            dstat.ensure(batch.size());
            drp.ensure(std::max<int64_t>(po, 1));
            dcp.ensure(std::max<int64_t>(po, 1));
            dhist.ensure(std::max<int64_t>(2 * po, 2));
            H2W_HIP(hipMemsetAsync(pool.top.p, 0, sizeof(unsigned long long), s));
            ++h.aca_batches;
            double tb = now_s();
            if (cx) {
                double2 *wsc = reinterpret_cast<double2 *>(pool.ws.p);
                if (nbig > 0)
                    launch_aca_c(djobs.p, (int)nbig, didx.p, wsc, pool.cap / 2, pool.top.p, dtab.p, flags.p, drank.p,
                                 dstat.p, drp.p, dcp.p, h.kpc, tol, dhist.p, 1, s);
                if (batch.size() > nbig)
                    launch_aca_c(djobs.p + nbig, (int)(batch.size() - nbig), didx.p, wsc, pool.cap / 2, pool.top.p,
                                 dtab.p, flags.p, drank.p + nbig, dstat.p + nbig, drp.p, dcp.p, h.kpc, tol, dhist.p, 0, s);
            } else {
                if (nhuge > 0)
                    run_wide(h, jobs, nhuge, djobs.p, didx.p, pool, dtab.p, flags.p, drank.p, dstat.p, drp.p, dcp.p,
                             dhist.p, tol);
                if (nbig > nhuge)
                    launch_aca(djobs.p + nhuge, (int)(nbig - nhuge), didx.p, pool.ws.p, pool.cap, pool.top.p, dtab.p,
                               flags.p, drank.p + nhuge, dstat.p + nhuge, drp.p, dcp.p, h.kp, tol, dhist.p, 1, s);
                if (batch.size() > nbig)
                    launch_aca(djobs.p + nbig, (int)(batch.size() - nbig), didx.p, pool.ws.p, pool.cap, pool.top.p,
                               dtab.p, flags.p, drank.p + nbig, dstat.p + nbig, drp.p, dcp.p, h.kp, tol, dhist.p, 0, s);
            }
            H2W_HIP(hipGetLastError());
            std::vector<int32_t> rk(batch.size()), st(batch.size()), rp(po), cp(po);
            std::vector<double> hs(2 * po);
            H2W_HIP(hipMemcpyAsync(hs.data(), dhist.p, 2 * po * 8, hipMemcpyDeviceToHost, s));
            H2W_HIP(hipMemcpyAsync(rk.data(), drank.p, rk.size() * 4, hipMemcpyDeviceToHost, s));
            H2W_HIP(hipMemcpyAsync(st.data(), dstat.p, st.size() * 4, hipMemcpyDeviceToHost, s));
            H2W_HIP(hipMemcpyAsync(rp.data(), drp.p, po * 4, hipMemcpyDeviceToHost, s));
            H2W_HIP(hipMemcpyAsync(cp.data(), dcp.p, po * 4, hipMemcpyDeviceToHost, s));
            H2W_HIP(hipStreamSynchronize(s));
            if (verbose) {
                int mr = 0, nov = 0;
                for (size_t q = 0; q < batch.size(); ++q) {
                    mr = std::max(mr, rk[q]);
                    nov += st[q] != 0;
                }
                std::fprintf(stderr, "[h2w]   aca batch: %zu blocks (%zu big), max rank %d, %d re-run, %.3f s\n",
                             batch.size(), nbig, mr, nov, now_s() - tb);
            }
            std::vector<LrJob> lj;
            std::vector<size_t> lb;
            int64_t lo = 0;
            for (size_t q = 0; q < batch.size(); ++q) {
                size_t b = batch[q];
                if (st[q] != 0) {
                    int mn = (int)std::min(blocks[b].rows.size(), blocks[b].cols.size());
                    if (batch.size() == 1 && est[b] >= mn)
                        throw Error{H2W_ENOMEM, "ACA workspace pool too small for one block"};
                    est[b] = std::min(mn, std::max(est[b] * 4, rk[q] * 2));
                    rerun.push_back(b);
                    h.aca_reruns++;
                    continue;
                }
                AcaBlock &B = blocks[b];
                B.rank = rk[q];
                if (B.has_feat) {
                    h.sample_feat.push_back(B.feat);
                    h.sample_rank.push_back(rk[q]);
                }
                if (pred) {
                    h.rank_pred_blocks++;
                    h.rank_pred_under += rk[q] > est0[b];
                    h.rank_pred_abs_log2 += std::fabs((double)plog[b] - std::log2(1.0 + rk[q]));
                }
                B.rb.assign(rp.begin() + jobs[q].piv, rp.begin() + jobs[q].piv + rk[q]);
                B.cb.assign(cp.begin() + jobs[q].piv, cp.begin() + jobs[q].piv + rk[q]);
                B.hist.assign(hs.begin() + 2 * jobs[q].piv, hs.begin() + 2 * (jobs[q].piv + rk[q]));
                if (rk[q] > 0) {
                    LrJob J;
                    J.tab = jobs[q].tab;
                    J.piv = jobs[q].piv;
                    J.m = jobs[q].m;
                    J.n = jobs[q].n;
                    J.rank = rk[q];
                    J.pad = 0;
                    J.out = lo;
                    lo += 2LL * rk[q] * rk[q];
                    lj.push_back(J);
                    lb.push_back(b);
                }
            }
            if (!lj.empty()) {
                dlr.upload(lj, s);
                dlrout.ensure(lo * esz);
                if (cx)
                    launch_lr_extract_c(dlr.p, (int)lj.size(), reinterpret_cast<const double2 *>(pool.ws.p), dtab.p,
                                        drp.p, dcp.p, reinterpret_cast<double2 *>(dlrout.p), s);
                else launch_lr_extract(dlr.p, (int)lj.size(), pool.ws.p, dtab.p, drp.p, dcp.p, dlrout.p, s);
                H2W_HIP(hipGetLastError());
                std::vector<double> hl(lo * esz);
                H2W_HIP(hipMemcpyAsync(hl.data(), dlrout.p, lo * esz * 8, hipMemcpyDeviceToHost, s));
                H2W_HIP(hipStreamSynchronize(s));
                for (size_t q = 0; q < lj.size(); ++q) {
                    AcaBlock &B = blocks[lb[q]];
                    // (L then R, r x r each; complex entries as (re, im) pairs)
                    B.lr.assign(hl.begin() + lj[q].out * esz,
                                hl.begin() + (lj[q].out + 2LL * lj[q].rank * lj[q].rank) * esz);
                }
            }
        }
        todo.swap(rerun);
        if (!todo.empty()) {
            // re-runs go with larger estimates: let the workspace grow to hold them
            int64_t need2 = 0;
            for (size_t b : todo) need2 += (int64_t)(blocks[b].rows.size() + blocks[b].cols.size()) * est[b];
            ensure_pool(h, pool, need2 * esz);
        }
    }
}

std::vector<int32_t> range_list(int64_t start, int64_t count) {
    std::vector<int32_t> v(count);
    for (int64_t i = 0; i < count; ++i) v[i] = (int32_t)(start + i);
    return v;
}

// far field, bottom-up NCA (3D/codes/HODLR3Dtree.hpp:1872-1939)
void far_nca(Handle &h, AcaPool &pool) {
    const Plan &P = h.plan;
    auto &F = h.fam[0];
    F.assign(P.nL + 1, FamLevel());
    if (P.dim == 1) return;
    for (int j = P.nL; j >= 2; --j) {
        FamLevel &fl = F[j];
        fl.nb = P.nboxes[j];
        fl.S.resize(fl.nb);
        for (int b = 0; b < fl.nb; ++b) {
            if (j == P.nL) fl.S[b] = range_list(P.bstart[j][b], P.bcount[j][b]);
            else
                for (int c = 0; c < P.nchild; ++c) {
                    auto &ir = F[j + 1].Ir[b * P.nchild + c];
                    fl.S[b].insert(fl.S[b].end(), ir.begin(), ir.end());
                }
        }
        std::vector<AcaBlock> blocks(fl.nb);
        for (int b = 0; b < fl.nb; ++b) {
            blocks[b].rows = fl.S[b];
            const auto far = P.far_of(j, b);
            for (int f : far) blocks[b].cols.insert(blocks[b].cols.end(), fl.S[f].begin(), fl.S[f].end());
            block_features(h, 0, j, b, far, (int64_t)blocks[b].rows.size(), (int64_t)blocks[b].cols.size(), blocks[b]);
        }
        // the far rank is bounded by the row-set size (children's skeletons)
        int hint = 0;
        for (auto &B : blocks) hint = std::max<int>(hint, (int)B.rows.size());
        run_aca(h, blocks, hint, pool);
        fl.rank.resize(fl.nb);
        fl.Ir.resize(fl.nb);
        fl.Ic.resize(fl.nb);
        fl.lr.resize(fl.nb);
        fl.hist.resize(fl.nb);
        for (int b = 0; b < fl.nb; ++b) {
            AcaBlock &B = blocks[b];
            fl.rank[b] = B.rank;
            fl.hist[b] = std::move(B.hist);
            for (int q = 0; q < B.rank; ++q) {
                fl.Ir[b].push_back(B.rows[B.rb[q]]);
                fl.Ic[b].push_back(B.cols[B.cb[q]]);
            }
            fl.lr[b] = std::move(B.lr);
        }
        if (std::getenv("H2W_VERBOSE")) std::fprintf(stderr, "[h2w]  far level %d done\n", j);
    }
}

// vertex-sharing field, top-down NCA with the parent's column pivots appended
// (3D/codes/HODLR3Dtree.hpp:2062-2107; 1D/codes/HODLR1Dtree.hpp:231-276)
void vertex_nca(Handle &h, AcaPool &pool) {
    const Plan &P = h.plan;
    auto &F = h.fam[1];
    F.assign(P.nL + 1, FamLevel());
    for (int j = 1; j <= P.nL; ++j) {
        FamLevel &fl = F[j];
        fl.nb = P.nboxes[j];
        std::vector<AcaBlock> blocks(fl.nb);
        for (int b = 0; b < fl.nb; ++b) {
            blocks[b].rows = range_list(P.bstart[j][b], P.bcount[j][b]);
            const auto ver = P.vert_of(j, b);
            for (int v : ver) {
                auto r = range_list(P.bstart[j][v], P.bcount[j][v]);
                blocks[b].cols.insert(blocks[b].cols.end(), r.begin(), r.end());
            }
            if (j > 1) {
                auto &pic = F[j - 1].Ic[b / P.nchild];
                blocks[b].cols.insert(blocks[b].cols.end(), pic.begin(), pic.end());
            }
            block_features(h, 1, j, b, ver, (int64_t)blocks[b].rows.size(), (int64_t)blocks[b].cols.size(), blocks[b]);
        }
        // ranks shrink with the box size: start from twice the parent level's largest rank
        int hint = std::max(16, h.opts.rank_cap);
        if (j > 1) {
            hint = 0;
            for (int r : F[j - 1].rank) hint = std::max(hint, r);
            hint = std::max(16, 2 * hint);
        }
        run_aca(h, blocks, hint, pool);
        fl.rank.resize(fl.nb);
        fl.Ir.resize(fl.nb);
        fl.Ic.resize(fl.nb);
        fl.lr.resize(fl.nb);
        fl.hist.resize(fl.nb);
        for (int b = 0; b < fl.nb; ++b) {
            AcaBlock &B = blocks[b];
            fl.rank[b] = B.rank;
            fl.hist[b] = std::move(B.hist);
            for (int q = 0; q < B.rank; ++q) {
                fl.Ir[b].push_back(B.rows[B.rb[q]]);
                fl.Ic[b].push_back(B.cols[B.cb[q]]);
            }
            fl.lr[b] = std::move(B.lr);
        }
        if (std::getenv("H2W_VERBOSE")) std::fprintf(stderr, "[h2w]  vertex level %d done\n", j);
    }
    for (int j = 1; j <= P.nL; ++j) {
        FamLevel &fl = F[j];
        fl.S.assign(fl.nb, {});
        for (int b = 0; b < fl.nb; ++b) {
            if (j == P.nL) fl.S[b] = range_list(P.bstart[j][b], P.bcount[j][b]);
            else
                for (int c = 0; c < P.nchild; ++c) {
                    auto &ir = F[j + 1].Ir[b * P.nchild + c];
                    fl.S[b].insert(fl.S[b].end(), ir.begin(), ir.end());
                }
        }
    }
}

// semi-nested vertex blocks: one cross approximation per directed pair and level
// (3D/codes/HODLR3Dtree.hpp:2250-2321; 1D/codes/HODLR1Dtree.hpp:418-443)
void cur_pairs(Handle &h, AcaPool &pool) {
    const Plan &P = h.plan;
    h.pairs.clear();
    std::vector<AcaBlock> blocks;
    for (int j = 1; j <= P.nL; ++j)
        for (int b = 0; b < P.nboxes[j]; ++b)
            for (int v : (h.variant == H2W_NON_NESTED ? P.il[j][b] : P.vert_of(j, b))) {
                Pair p;
                p.level = j;
                p.b = b;
                p.v = v;
                p.rank = 0;
                h.pairs.push_back(p);
                AcaBlock B;
                B.rows = range_list(P.bstart[j][b], P.bcount[j][b]);
                B.cols = range_list(P.bstart[j][v], P.bcount[j][v]);
                block_features(h, 2, j, b, {v}, (int64_t)B.rows.size(), (int64_t)B.cols.size(), B);
                blocks.push_back(std::move(B));
            }
    run_aca(h, blocks, std::max(16, h.opts.rank_cap), pool);
    for (size_t q = 0; q < blocks.size(); ++q) {
        Pair &p = h.pairs[q];
        AcaBlock &B = blocks[q];
        p.rank = B.rank;
        for (int k = 0; k < B.rank; ++k) {
            p.Ir.push_back(B.rows[B.rb[k]]);
            p.Ic.push_back(B.cols[B.cb[k]]);
        }
        p.lr = std::move(B.lr);
    }
}

// The rows of a transfer (columns of a CUR row factor) are independent: sub-jobs of
// 64 rows (256 columns) spread one large box over many workgroups.
std::vector<XferJob> split_xfer(const std::vector<XferJob> &in) {
    std::vector<XferJob> out;
    for (const XferJob &J : in)
        for (int r0 = 0; r0 < J.m; r0 += 64) {
            XferJob K = J;
            K.m = std::min(64, J.m - r0);
            K.s = J.s + r0;
            K.e = J.e + r0;
            K.et = J.et + (int64_t)r0 * J.ldet;
            out.push_back(K);
        }
    return out;
}
std::vector<CurJob> split_cur(const std::vector<CurJob> &in) {
    std::vector<CurJob> out;
    for (const CurJob &J : in)
        for (int c0 = 0; c0 < J.n; c0 += 256) {
            CurJob K = J;
            K.n = std::min(256, J.n - c0);
            K.v = J.v + c0;
            K.a = J.a + (int64_t)c0 * J.lda;
            out.push_back(K);
        }
    return out;
}

// ------------------------------------------------------------------ work lists
constexpr int64_t kItemWork = 128 * 4096;  // elements of A per GEMV item before column chunking

void add_item(WorkList &W, int64_t y, int rows, const std::vector<TermSpec> &terms, bool atomic, int lda = -1,
              int row_base = 0, int pool = 0) {
    if (rows <= 0 || terms.empty()) return;
    if (lda < 0) lda = rows;
    int64_t C = 0;
    for (auto &t : terms) C += t.cols;
    if (C == 0) return;
    for (int r0 = 0; r0 < rows; r0 += kRowTile) {
        int tr = std::min(kRowTile, rows - r0);
        if ((int64_t)tr * C <= kItemWork) {
            GItem it;
            it.y = y + r0;
            it.xb = 0;
            it.rows = tr;
            it.lda = lda;
            it.row0 = row_base + r0;
            it.t0 = (int32_t)W.terms.size();
            for (auto &t : terms)
                if (t.cols > 0) W.terms.push_back(GTerm{t.a, t.x, -1, t.cols, 0});
            it.t1 = (int32_t)W.terms.size();
            it.flags = (atomic ? 1 : 0) | (pool << 1);
            W.items.push_back(it);
        } else {
            // column chunks accumulate atomically
            const int64_t maxc = std::max<int64_t>(1, kItemWork / kRowTile);
            GItem it;
            it.rows = tr;
            it.lda = lda;
            it.row0 = row_base + r0;
            it.y = y + r0;
            it.xb = 0;
            it.flags = 1 | (pool << 1);
            int64_t used = 0;
            it.t0 = (int32_t)W.terms.size();
            for (auto &t : terms) {
                for (int64_t c0 = 0; c0 < t.cols;) {
                    int64_t take = std::min<int64_t>(t.cols - c0, maxc - used);
                    W.terms.push_back(GTerm{t.a + c0 * lda, t.x + c0, -1, (int32_t)take, 0});
                    used += take;
                    c0 += take;
                    if (used == maxc) {
                        it.t1 = (int32_t)W.terms.size();
                        W.items.push_back(it);
                        it.t0 = it.t1;
                        used = 0;
                    }
                }
            }
            it.t1 = (int32_t)W.terms.size();
            if (it.t1 > it.t0) W.items.push_back(it);
        }
    }
}

// rank at which the ACA would have stopped for tolerance tau: the first accepted
// step k >= 1 with |u_k||v_k| <= tau * normS_k ends the cross (2D/codes/aca.hpp:363)
// Items whose terms are never split by columns: a row tile's terms are grouped into
// items of at most kItemWork elements (a tile split over several items accumulates
// atomically).  Terms keep their t (a symmetric term's transposed-product output).
void add_item_whole(WorkList &W, int64_t y, int64_t xb, int rows, const std::vector<TermSpec> &terms, int lda,
                    int pool) {
    if (rows <= 0) return;
    std::vector<const TermSpec *> ts;
    for (auto &t : terms)
        if (t.cols > 0) ts.push_back(&t);
    if (ts.empty()) return;
    for (int r0 = 0; r0 < rows; r0 += kRowTile) {
        const int tr = std::min(kRowTile, rows - r0);
        std::vector<std::pair<size_t, size_t>> groups;  // [first, last) of ts per item
        size_t g0 = 0;
        int64_t work = 0;
        for (size_t q = 0; q < ts.size(); ++q) {
            const int64_t w = (int64_t)tr * ts[q]->cols;
            if (q > g0 && work + w > kItemWork) {
                groups.push_back({g0, q});
                g0 = q;
                work = 0;
            }
            work += w;
        }
        groups.push_back({g0, ts.size()});
        for (auto &g : groups) {
            GItem it;
            it.y = y + r0;
            it.xb = xb + r0;
            it.rows = tr;
            it.lda = lda;
            it.row0 = r0;
            it.t0 = (int32_t)W.terms.size();
            for (size_t q = g.first; q < g.second; ++q) {
                const TermSpec &t = *ts[q];
                W.terms.push_back(GTerm{t.a, t.x, t.t, t.cols, t.ld});
            }
            it.t1 = (int32_t)W.terms.size();
            it.flags = (groups.size() > 1 ? 1 : 0) | (pool << 1);
            W.items.push_back(it);
        }
    }
}

// A launch of few large items (the transfer passes near the root) leaves most of the
// GPU idle while a handful of workgroups stream their items: split such items by
// columns (the pieces accumulate atomically) until the launch has about kMinItems
// items of at least kMinPiece elements.  Lists with symmetric terms are left alone.
constexpr int64_t kMinItems = 2048, kMinPiece = 128 * 64;
void rebalance(WorkList &W) {
    if ((int64_t)W.items.size() >= kMinItems) return;
    int64_t E = 0;
    for (auto &it : W.items)
        for (int32_t k = it.t0; k < it.t1; ++k) {
            if (W.terms[k].t >= 0) return;
            E += (int64_t)it.rows * W.terms[k].cols;
        }
    const int64_t target = std::max<int64_t>(E / kMinItems, kMinPiece);
    WorkList out;
    out.role = W.role;
    out.stage = W.stage;
    out.step = W.step;
    for (const GItem &it : W.items) {
        int64_t work = 0;
        for (int32_t k = it.t0; k < it.t1; ++k) work += (int64_t)it.rows * W.terms[k].cols;
        GItem nt = it;
        nt.t0 = (int32_t)out.terms.size();
        if (work <= 2 * target) {
            for (int32_t k = it.t0; k < it.t1; ++k) out.terms.push_back(W.terms[k]);
            nt.t1 = (int32_t)out.terms.size();
            out.items.push_back(nt);
            continue;
        }
        const int64_t maxc = std::max<int64_t>(1, target / it.rows);
        nt.flags |= 1;
        int64_t used = 0;
        for (int32_t k = it.t0; k < it.t1; ++k) {
            const GTerm &t = W.terms[k];
            for (int64_t c0 = 0; c0 < t.cols;) {
                const int64_t take = std::min<int64_t>(t.cols - c0, maxc - used);
                out.terms.push_back(GTerm{t.a + c0 * it.lda, t.x + c0, -1, (int32_t)take, t.ld});
                used += take;
                c0 += take;
                if (used == maxc) {
                    nt.t1 = (int32_t)out.terms.size();
                    out.items.push_back(nt);
                    nt.t0 = nt.t1;
                    used = 0;
                }
            }
        }
        nt.t1 = (int32_t)out.terms.size();
        if (nt.t1 > nt.t0) out.items.push_back(nt);
    }
    W = std::move(out);
}

// One group of boxes whose pairwise blocks are stored once per unordered pair (the
// M2L blocks of one family and level, or the near field): box b has output offset
// y[b], input offset x[b] and size r[b]; up[b] lists the stored blocks K(b, p), p > b
// (r[b] x r[p], leading dimension pad_ld(r[b])), self[b] the block K(b, b) or -1.
struct SymGroup {
    std::vector<int64_t> y, x;
    std::vector<int> r;
    std::vector<std::vector<std::pair<int, int64_t>>> up;
    std::vector<int64_t> self;
    int pool = 0;
};
// The three lists of such a group: fused items (box b: y_b += K(b, p) x_p over its stored
// blocks, and each block's K(b, p)^T x_b added to y_p; in the single-vector product every
// fused item accumulates atomically, since its rows also receive the transposed products of
// other items), the
// 16-wide path's transposed items (box p: y_p += K(b, p)^T X_b, b < p) and the on-the-fly
// path's directed items.
struct SymLists {
    WorkList fused, trans, otf;
};
void sym_lists(const SymGroup &G, SymLists &L) {
    const int nb = (int)G.r.size();
    std::vector<std::vector<std::pair<int, int64_t>>> down(nb);
    for (int b = 0; b < nb; ++b)
        for (auto &pe : G.up[b]) down[pe.first].push_back({b, pe.second});
    const size_t item0 = L.fused.items.size();
    for (int b = 0; b < nb; ++b) {
        if (G.r[b] <= 0) continue;
        std::vector<TermSpec> t;
        if (G.self[b] >= 0) t.push_back(TermSpec{G.self[b], G.x[b], G.r[b]});
        for (auto &pe : G.up[b]) {
            TermSpec ts{pe.second, G.x[pe.first], G.r[pe.first]};
            ts.t = G.y[pe.first];
            t.push_back(ts);
        }
        add_item_whole(L.fused, G.y[b], G.x[b], G.r[b], t, (int)pad_ld(G.r[b]), G.pool);
    }
    // (flag bit 2: the single-vector kernel adds y atomically; the 16-wide path applies the
    // transposes in a later launch and keeps exclusive writers)
    for (size_t q = item0; q < L.fused.items.size(); ++q) L.fused.items[q].flags |= 4;
    for (int p = 0; p < nb; ++p) {
        std::vector<TermSpec> t;
        for (auto &be : down[p]) {
            TermSpec ts{be.second, G.x[be.first], G.r[be.first]};
            ts.ld = (int32_t)pad_ld(G.r[be.first]);
            t.push_back(ts);
        }
        add_item_whole(L.trans, G.y[p], 0, G.r[p], t, (int)pad_ld(G.r[p]), G.pool);
    }
    for (int b = 0; b < nb; ++b) {
        std::vector<TermSpec> t;
        if (G.self[b] >= 0) t.push_back(TermSpec{0, G.x[b], G.r[b]});
        for (auto &pe : G.up[b]) t.push_back(TermSpec{0, G.x[pe.first], G.r[pe.first]});
        for (auto &be : down[b]) t.push_back(TermSpec{0, G.x[be.first], G.r[be.first]});
        add_item(L.otf, G.y[b], G.r[b], t, false, (int)pad_ld(G.r[b]));
    }
}

int inner_rank(const std::vector<double> &hist, int rank, double tau) {
    for (int k = 1; k < rank && 2 * k + 1 < (int)hist.size(); ++k)
        if (hist[2 * k] <= tau * hist[2 * k + 1]) return k + 1;
    return rank;
}

void upload_lists(std::vector<WorkList> &WL, LaunchSet &LS, hipStream_t s) {
    std::vector<GItem> items;
    std::vector<GTerm> terms;
    LS.launches.clear();
    LS.elems.clear();
    for (auto &W : WL) {
        if (W.items.empty()) continue;
        Launch Lh;
        Lh.item0 = (int64_t)items.size();
        Lh.nitems = (int64_t)W.items.size();
        Lh.term0 = (int64_t)terms.size();
        Lh.nterms = (int64_t)W.terms.size();
        Lh.role = W.role;
        Lh.stage = W.stage;
        Lh.step = W.step;
        int64_t e = 0, sc = 0;
        for (auto &tm : W.terms) sc += tm.cols;
        Lh.avg_cols = W.terms.empty() ? 0 : (int)(sc / (int64_t)W.terms.size());
        for (auto it : W.items) {
            for (int q = it.t0; q < it.t1; ++q) e += (int64_t)it.rows * W.terms[q].cols;
            it.t0 += (int32_t)Lh.term0;
            it.t1 += (int32_t)Lh.term0;
            items.push_back(it);
        }
        terms.insert(terms.end(), W.terms.begin(), W.terms.end());
        Lh.avg_item_elems = e / Lh.nitems;
        LS.launches.push_back(Lh);
        LS.elems.push_back(e);
    }
    LS.d_items.upload(items, s);
    LS.d_terms.upload(terms, s);
    LS.h_items = std::move(items);
    LS.h_terms = std::move(terms);
}

// ------------------------------------------------------------------ the refinement's inner operator
// Every box keeps the leading pivots at which its cross would have stopped for
// tolerance 10^-irp (from the ACA history); its M2L blocks are compact copies of
// the leading sub-blocks of the stored FP32 M2L blocks, its transfers are
// recomputed from the leading blocks of L and R.  Lives in its own FP32 buffer
// (work items flagged with pool 1); rebuilt by h2w_set_inner_tolerance and the
// autotuner.

void build_inner(Handle &h, double irp) {
    drop_graph(h, H2W_FP32_INNER);
    ++h.inner_gen;
    const Plan &P = h.plan;
    hipStream_t s = h.stream;
    // the inner buffers and lists are freed and reallocated below: nothing queued may still read them
    H2W_HIP(hipStreamSynchronize(s));
    const int nfam_lo[2] = {2, 1};
    const bool fam_on[2] = {P.dim >= 2 && h.variant != H2W_NON_NESTED, h.variant == H2W_NESTED};
    const int nleaf = P.nboxes[P.nL];
    auto &m2l = h.m2l;
    auto &nearb = h.nearb;
    h.inner.launches.clear();
    h.inner_elems = 0;
    h.mat32in.release();
    // ---- inner operator of the refinement: each box keeps the leading pivots at which its
    // cross would have stopped for tolerance 10^-ir_tol_pow; its M2L blocks are the leading
    // sub-blocks of the stored ones, its transfers are recomputed from the leading L, R blocks
    const bool inner_on = h.opts.fp32_shadow && irp > 0 && irp < h.tol_pow;
    int64_t in_len = 0;
    auto take_in = [&](int64_t elems) {
        int64_t o = in_len;
        in_len += align_up(elems, 32);
        return o;
    };
    // compact copies of the inner M2L sub-blocks: (index into m2l, offset in the inner region)
    std::vector<std::pair<size_t, int64_t>> m2l_in;
    if (inner_on) {
        const double tau = std::pow(10.0, -1.0 * irp);
        for (int f = 0; f < 2; ++f) {
            if (!fam_on[f]) continue;
            for (int j = nfam_lo[f]; j <= P.nL; ++j) {
                FamLevel &fl = h.fam[f][j];
                fl.kin.assign(fl.nb, 0);
                for (int b = 0; b < fl.nb; ++b)
                    fl.kin[b] = fl.rank[b] > 0 ? inner_rank(fl.hist[b], fl.rank[b], tau) : 0;
            }
            // inner transfers: rows are the children's leading pivots, each child's segment
            // padded to a multiple of 4 rows so every row tile stays aligned
            for (int j = nfam_lo[f]; j <= P.nL; ++j) {
                FamLevel &fl = h.fam[f][j];
                fl.ein_off.assign(fl.nb, -1);
                fl.eint_off.assign(fl.nb, -1);
                for (int b = 0; b < fl.nb; ++b) {
                    int64_t m = 0;
                    if (j == P.nL) m = P.bcount[j][b];
                    else
                        for (int c = 0; c < P.nchild; ++c) m += pad_ld(h.fam[f][j + 1].kin[b * P.nchild + c]);
                    if (m * fl.kin[b] == 0) continue;
                    fl.ein_off[b] = take_in(pad_ld(m) * fl.kin[b]);
                    fl.eint_off[b] = take_in(pad_ld(fl.kin[b]) * m);
                }
            }
        }
        for (size_t q = 0; q < m2l.size(); ++q) {
            const M2LBlk &m = m2l[q];
            const FamLevel &fl = h.fam[m.f][m.j];
            if (fl.kin[m.b] > 0 && fl.kin[m.p] > 0)
                m2l_in.push_back({q, take_in(pad_ld(fl.kin[m.b]) * fl.kin[m.p])});
        }
    }
    if (inner_on) h.mat32in.alloc(std::max<int64_t>(in_len, 32));
    if (inner_on && in_len > 0) H2W_HIP(hipMemsetAsync(h.mat32in.p, 0, in_len * 4, s));
    if (inner_on && in_len > 0) {
        std::vector<int32_t> idx;
        std::vector<XferJob> xj;
        std::vector<double> lr;
        for (int f = 0; f < 2; ++f) {
            if (!fam_on[f]) continue;
            for (int j = nfam_lo[f]; j <= P.nL; ++j) {
                FamLevel &fl = h.fam[f][j];
                for (int b = 0; b < fl.nb; ++b) {
                    if (fl.ein_off[b] < 0) continue;
                    const int k = fl.kin[b], r = fl.rank[b];
                    // leading k x k blocks of L and R
                    const int64_t lro = (int64_t)lr.size();
                    for (int part = 0; part < 2; ++part)
                        for (int cc = 0; cc < k; ++cc)
                            for (int rr = 0; rr < k; ++rr)
                                lr.push_back(fl.lr[b][(size_t)part * r * r + (size_t)cc * r + rr]);
                    const int64_t icq = (int64_t)idx.size();
                    idx.insert(idx.end(), fl.Ic[b].begin(), fl.Ic[b].begin() + k);
                    int64_t m = 0;
                    if (j == P.nL) m = P.bcount[j][b];
                    else
                        for (int c = 0; c < P.nchild; ++c) m += pad_ld(h.fam[f][j + 1].kin[b * P.nchild + c]);
                    const int lde = (int)pad_ld(m), ldet = (int)pad_ld(k);
                    auto add_job = [&](const int32_t *S, int ms, int64_t row) {
                        if (ms == 0) return;
                        XferJob J;
                        J.r = k;
                        J.m = ms;
                        J.s = (int64_t)idx.size();
                        idx.insert(idx.end(), S, S + ms);
                        J.ic = icq;
                        J.lr = lro;
                        J.e = fl.ein_off[b] + row;
                        J.et = fl.eint_off[b] + row * ldet;
                        J.lde = lde;
                        J.ldet = ldet;
                        xj.push_back(J);
                    };
                    if (j == P.nL) {
                        auto rl = range_list(P.bstart[j][b], P.bcount[j][b]);
                        add_job(rl.data(), (int)rl.size(), 0);
                    } else {
                        int64_t row = 0;
                        for (int c = 0; c < P.nchild; ++c) {
                            const int cc = b * P.nchild + c;
                            const int kc = h.fam[f][j + 1].kin[cc];
                            add_job(h.fam[f][j + 1].Ir[cc].data(), kc, row);
                            row += pad_ld(kc);
                        }
                    }
                }
            }
        }
        DevBuf<int32_t> didx;
        DevBuf<double> dlr, tmp;
        DevBuf<XferJob> dxj;
        didx.upload(idx, s);
        dlr.upload(lr, s);
        xj = split_xfer(xj);
        dxj.upload(xj, s);
        tmp.alloc(in_len);
        H2W_HIP(hipMemsetAsync(tmp.p, 0, in_len * 8, s));
        launch_xfer(dxj.p, (int)xj.size(), didx.p, dlr.p, tmp.p, h.kp, s);
        H2W_HIP(hipGetLastError());
        launch_to_fp32(tmp.p, h.mat32in.p, in_len, s);
        H2W_HIP(hipGetLastError());
        // compact copies of the leading M2L sub-blocks
        std::vector<CopyJob> cj;
        for (auto &pr : m2l_in) {
            const M2LBlk &m = m2l[pr.first];
            const FamLevel &fl = h.fam[m.f][m.j];
            CopyJob J;
            J.src = m.off;
            J.dst = pr.second;
            J.m = fl.kin[m.b];
            J.n = fl.kin[m.p];
            J.lds = (int32_t)pad_ld(fl.rank[m.b]);
            J.ldd = (int32_t)pad_ld(fl.kin[m.b]);
            cj.push_back(J);
        }
        DevBuf<CopyJob> dcj;
        dcj.upload(cj, s);
        launch_copy32(dcj.p, (int)cj.size(), h.mat32.p, h.mat32in.p, s);
        H2W_HIP(hipGetLastError());
        H2W_HIP(hipStreamSynchronize(s));
    }

    if (inner_on) {
        const int64_t base = 0;  // offsets into the inner buffer (items flagged with pool 1)
        std::vector<WorkList> IL;
        {
            WorkList W;
            for (int f = 0; f < 2; ++f) {
                if (!fam_on[f] || P.nL < nfam_lo[f]) continue;
                FamLevel &fl = h.fam[f][P.nL];
                for (int b = 0; b < fl.nb; ++b)
                    if (fl.eint_off[b] >= 0)
                        add_item(W, fl.o_off[b], fl.kin[b],
                                 {TermSpec{base + fl.eint_off[b], h.off_x + P.bstart[P.nL][b], (int32_t)P.bcount[P.nL][b]}},
                                 false, (int)pad_ld(fl.kin[b]), 0, 1);
            }
            for (auto &p : h.pairs)
                if (p.rank > 0)
                    add_item(W, p.t_off, p.rank,
                             {TermSpec{p.ar_off, h.off_x + P.bstart[p.level][p.v], (int32_t)P.bcount[p.level][p.v]}},
                             false, (int)pad_ld(p.rank));
            IL.push_back(std::move(W));
        }
        for (int j = P.nL - 1; j >= 1; --j) {
            WorkList W;
            W.step = P.nL - j;
            for (int f = 0; f < 2; ++f) {
                if (!fam_on[f] || j < nfam_lo[f]) continue;
                FamLevel &fl = h.fam[f][j];
                FamLevel &ch = h.fam[f][j + 1];
                for (int b = 0; b < fl.nb; ++b) {
                    if (fl.eint_off[b] < 0) continue;
                    const int64_t ldet = pad_ld(fl.kin[b]);
                    std::vector<TermSpec> t;
                    int64_t col = 0;
                    for (int c = 0; c < P.nchild; ++c) {
                        const int cc = b * P.nchild + c;
                        if (ch.kin[cc] > 0)
                            t.push_back(TermSpec{base + fl.eint_off[b] + col * ldet, ch.o_off[cc], (int32_t)ch.kin[cc]});
                        col += pad_ld(ch.kin[cc]);
                    }
                    add_item(W, fl.o_off[b], fl.kin[b], t, false, (int)ldet, 0, 1);
                }
            }
            IL.push_back(std::move(W));
        }
        {
            // the inner M2L blocks (compact leading sub-blocks, one per unordered pair, second
            // pool) and the stored near field, both directions in one fused launch
            SymLists SL;
            for (int f = 0; f < 2; ++f) {
                if (!fam_on[f]) continue;
                for (int j = nfam_lo[f]; j <= P.nL; ++j) {
                    FamLevel &fl = h.fam[f][j];
                    SymGroup G;
                    G.y.assign(fl.u_off.begin(), fl.u_off.end());
                    G.x.assign(fl.o_off.begin(), fl.o_off.end());
                    G.r = fl.kin;
                    G.up.assign(fl.nb, {});
                    G.self.assign(fl.nb, -1);
                    G.pool = 1;
                    for (auto &pr : m2l_in) {
                        const M2LBlk &m = m2l[pr.first];
                        if (m.f == f && m.j == j) G.up[m.b].push_back({m.p, base + pr.second});
                    }
                    sym_lists(G, SL);
                }
            }
            {
                SymGroup G;
                G.y.resize(nleaf);
                G.x.resize(nleaf);
                G.r.resize(nleaf);
                G.up.assign(nleaf, {});
                G.self.assign(nleaf, -1);
                for (int b = 0; b < nleaf; ++b) {
                    G.y[b] = h.off_y + P.bstart[P.nL][b];
                    G.x[b] = h.off_x + P.bstart[P.nL][b];
                    G.r[b] = (int)P.bcount[P.nL][b];
                }
                for (auto &nb : nearb) {
                    if (nb.n == nb.b) G.self[nb.b] = nb.off;
                    else G.up[nb.b].push_back({nb.n, nb.off});
                }
                sym_lists(G, SL);
            }
            SL.fused.role = 1;
            SL.fused.stage = 1;
            IL.push_back(std::move(SL.fused));
            SL.trans.role = 3;
            SL.trans.stage = 1;
            SL.trans.step = 1;
            std::vector<WorkList> tl;
            tl.push_back(std::move(SL.trans));
            upload_lists(tl, h.inner_trans, s);
        }
        for (int j = 1; j <= P.nL - 1; ++j) {
            WorkList W;
            W.stage = 2;
            W.step = P.nL - j;
            for (int f = 0; f < 2; ++f) {
                if (!fam_on[f] || j < nfam_lo[f]) continue;
                FamLevel &fl = h.fam[f][j];
                FamLevel &ch = h.fam[f][j + 1];
                for (int b = 0; b < fl.nb; ++b) {
                    if (fl.ein_off[b] < 0) continue;
                    int64_t m = 0;
                    for (int c = 0; c < P.nchild; ++c) m += pad_ld(ch.kin[b * P.nchild + c]);
                    int64_t row = 0;
                    for (int c = 0; c < P.nchild; ++c) {
                        const int cc = b * P.nchild + c;
                        if (ch.kin[cc] > 0)
                            add_item(W, ch.u_off[cc], ch.kin[cc],
                                     {TermSpec{base + fl.ein_off[b], fl.u_off[b], (int32_t)fl.kin[b]}}, false,
                                     (int)pad_ld(m), (int)row, 1);
                        row += pad_ld(ch.kin[cc]);
                    }
                }
            }
            IL.push_back(std::move(W));
        }
        {
            WorkList W;
            W.stage = 2;
            for (int b = 0; b < nleaf; ++b) {
                std::vector<TermSpec> t;
                for (int f = 0; f < 2; ++f) {
                    if (!fam_on[f] || P.nL < nfam_lo[f]) continue;
                    FamLevel &fl = h.fam[f][P.nL];
                    if (fl.ein_off[b] >= 0) t.push_back(TermSpec{base + fl.ein_off[b], fl.u_off[b], (int32_t)fl.kin[b]});
                }
                add_item(W, h.off_y + P.bstart[P.nL][b], (int)P.bcount[P.nL][b], t, false,
                         (int)pad_ld(P.bcount[P.nL][b]), 0, 1);
            }
            IL.push_back(std::move(W));
        }
        if (!h.pairs.empty()) {
            WorkList W;
            W.role = 2;
            W.stage = 3;
            std::vector<std::vector<std::vector<TermSpec>>> per(P.nL + 1);
            for (int j = 1; j <= P.nL; ++j) per[j].assign(P.nboxes[j], {});
            for (auto &p : h.pairs)
                if (p.rank > 0) per[p.level][p.b].push_back(TermSpec{p.ac_off, p.t_off, (int32_t)p.rank});
            for (int j = 1; j <= P.nL; ++j)
                for (int b = 0; b < P.nboxes[j]; ++b)
                    add_item(W, h.off_y + P.bstart[j][b], (int)P.bcount[j][b], per[j][b], true,
                             (int)pad_ld(P.bcount[j][b]));
            IL.push_back(std::move(W));
        }
        for (auto &W : IL)
            if (W.role == 0) rebalance(W);
        upload_lists(IL, h.inner, s);
        h.inner_elems = 0;
        for (auto e : h.inner.elems) h.inner_elems += e;
    } else {
        h.inner.launches.clear();
        h.inner_elems = 0;
    }
    H2W_HIP(hipStreamSynchronize(s));
    h.ir_exp = inner_on ? irp : 0.0;
}

// ------------------------------------------------------------------ assembly

void assemble_operators(Handle &h, double t0);

void assemble(Handle &h) {
    drop_graphs(h);
    double t0 = now_s();
    h.aca_reruns = h.aca_batches = 0;
    h.aca_pool_bytes = 0;
    h.rank_pred_blocks = h.rank_pred_under = 0;
    h.rank_pred_abs_log2 = 0.0;
    h.rank_fix_blocks = h.rank_fix_resumed = 0;
    h.rank_fix_extra_steps = 0;
    h.rank_fix_sampled_err = 0.0;
    h.sample_feat.clear();
    h.sample_rank.clear();
    const bool verbose = std::getenv("H2W_VERBOSE") != nullptr;
    auto phase = [&](const char *what) {
        if (verbose) {
            (void)hipStreamSynchronize(h.stream);
            std::fprintf(stderr, "[h2w] %-28s %8.3f s\n", what, now_s() - t0);
        }
    };
    const Plan &P = h.plan;
    {
        AcaPool pool;
        if (P.dim >= 2 && h.variant != H2W_NON_NESTED) far_nca(h, pool);
        else h.fam[0].assign(P.nL + 1, FamLevel());
        phase("far NCA");
        if (h.variant == H2W_NESTED) vertex_nca(h, pool);
        else {
            h.fam[1].assign(P.nL + 1, FamLevel());
            cur_pairs(h, pool);
        }
        phase("vertex NCA / CUR");
    }
    assemble_operators(h, t0);
}

// Everything after the skeletons: pools, transfer / M2L / near-field / CUR blocks,
// the FP32 shadow, work lists and the inner operator.  Reads h.fam (ranks, pivots,
// L and R, ACA histories, row sets S) and h.pairs only, so it also rebuilds an
// operator from saved skeletons.
void assemble_operators(Handle &h, double t0) {
    const bool verbose = std::getenv("H2W_VERBOSE") != nullptr;
    auto phase = [&](const char *what) {
        if (verbose) {
            (void)hipStreamSynchronize(h.stream);
            std::fprintf(stderr, "[h2w] %-28s %8.3f s\n", what, now_s() - t0);
        }
    };
    const Plan &P = h.plan;
    hipStream_t s = h.stream;
    const int nfam_lo[2] = {2, 1};  // first level of each family
    const bool fam_on[2] = {P.dim >= 2 && h.variant != H2W_NON_NESTED, h.variant == H2W_NESTED};

    // ---- vector pool: x, y, skeleton vectors
    int64_t vo = 0;
    h.off_x = vo;
    vo += align_up(P.N, 32);
    h.off_y = vo;
    vo += align_up(P.N, 32);
    for (int f = 0; f < 2; ++f) {
        if (!fam_on[f]) continue;
        for (int j = nfam_lo[f]; j <= P.nL; ++j) {
            FamLevel &fl = h.fam[f][j];
            fl.o_off.resize(fl.nb);
            fl.u_off.resize(fl.nb);
            // o segments of one parent's children are contiguous (the parent's up input)
            for (int b = 0; b < fl.nb; ++b) {
                fl.o_off[b] = vo;
                vo += fl.rank[b];
            }
            vo = align_up(vo, 32);
            for (int b = 0; b < fl.nb; ++b) {
                fl.u_off[b] = vo;
                vo += fl.rank[b];
            }
            vo = align_up(vo, 32);
        }
    }
    for (auto &p : h.pairs) {
        p.t_off = vo;
        vo += p.rank;
    }
    h.vec_len = align_up(std::max<int64_t>(vo, 1), 32);

    // ---- matrix pool offsets
    int64_t mo = 0;
    auto take = [&](int64_t elems) {
        int64_t o = mo;
        mo += align_up(elems, 32);
        return o;
    };
    double memory = 0.0;
    for (int f = 0; f < 2; ++f) {
        if (!fam_on[f]) continue;
        for (int j = nfam_lo[f]; j <= P.nL; ++j) {
            FamLevel &fl = h.fam[f][j];
            fl.e_off.assign(fl.nb, -1);
            fl.et_off.assign(fl.nb, -1);
            for (int b = 0; b < fl.nb; ++b) {
                int64_t m = (int64_t)fl.S[b].size(), r = fl.rank[b];
                memory += 2.0 * (double)(m * r);
                if (m * r == 0) continue;
                fl.e_off[b] = take(pad_ld(m) * r);
                fl.et_off[b] = take(pad_ld(r) * m);
            }
        }
    }
    auto &m2l = h.m2l;
    m2l.clear();
    for (int f = 0; f < 2; ++f) {
        if (!fam_on[f]) continue;
        for (int j = nfam_lo[f]; j <= P.nL; ++j) {
            FamLevel &fl = h.fam[f][j];
            for (int b = 0; b < fl.nb; ++b) {
                auto partners = f == 0 ? P.far_of(j, b) : P.vert_of(j, b);
                for (int p : partners) {
                    int64_t e = (int64_t)fl.rank[b] * fl.rank[p];
                    memory += (double)e;  // the reference counts both directions
                    // K(Ir_p, Ir_b) = K(Ir_b, Ir_p)^T entry for entry (the kernel depends on |x - y|
                    // only): one block per unordered pair, the lower-numbered box's
                    if (e == 0 || p < b) continue;
                    m2l.push_back(M2LBlk{f, j, b, p, take(pad_ld(fl.rank[b]) * fl.rank[p])});
                }
            }
        }
    }
    auto &nearb = h.nearb;
    nearb.clear();
    const int nleaf = P.nboxes[P.nL];
    const bool count_near = !(P.dim == 1 && h.variant == H2W_SEMI_NESTED);  // 1D findMemory_sn skips it
    for (int b = 0; b < nleaf; ++b) {
        std::vector<int32_t> lst = P.near[b];
        lst.push_back(b);  // self last (3D/codes/HODLR3Dtree.hpp:2388)
        for (int n : lst) {
            int64_t e = P.bcount[P.nL][b] * P.bcount[P.nL][n];
            if (count_near) memory += (double)e;
            if (e == 0 || n < b) continue;  // one block per unordered pair (self blocks kept)
            nearb.push_back(NearBlk{b, n, take(pad_ld(P.bcount[P.nL][b]) * P.bcount[P.nL][n])});
        }
    }
    for (auto &p : h.pairs) {
        int64_t mb = P.bcount[p.level][p.b], nv = P.bcount[p.level][p.v];
        if (P.dim == 1) memory += (double)(mb * p.rank + (int64_t)p.rank * nv);  // U + V
        else memory += 2.0 * p.rank * p.rank + (double)(mb * p.rank + (int64_t)p.rank * nv);  // L + R + Ac + Ar
        p.ac_off = p.ar_off = -1;
        if (p.rank == 0) continue;
        p.ac_off = take(pad_ld(mb) * p.rank);
        p.ar_off = take(pad_ld(p.rank) * nv);
    }
    h.mat_len = std::max<int64_t>(mo, 32);
    h.memory = memory;
    // (complex kernels: pools of (re, im) pairs at the same element offsets)
    const int esz = h.cplx ? 2 : 1;
    h.mat64.alloc(h.mat_len * esz);
    H2W_HIP(hipMemsetAsync(h.mat64.p, 0, h.mat_len * esz * 8, s));
    h.vec64.alloc(h.vec_len * esz);
    if (h.cplx) h.vec32.release();
    else h.vec32.alloc(h.vec_len);
    {
        // the point of every vector slot, for on-the-fly near-field / M2L evaluation
        std::vector<int32_t> pid(h.vec_len, -1);
        for (int64_t t = 0; t < P.N; ++t) pid[h.off_x + t] = pid[h.off_y + t] = (int32_t)t;
        for (int f = 0; f < 2; ++f) {
            if (!fam_on[f]) continue;
            for (int j = nfam_lo[f]; j <= P.nL; ++j) {
                FamLevel &fl = h.fam[f][j];
                for (int b = 0; b < fl.nb; ++b)
                    for (int k = 0; k < fl.rank[b]; ++k) pid[fl.o_off[b] + k] = pid[fl.u_off[b] + k] = fl.Ir[b][k];
            }
        }
        std::vector<float> fx(h.vec_len, 0.f), fy(h.vec_len, 0.f), fz(h.vec_len, 0.f);
        std::vector<float> lx(h.vec_len, 0.f), ly(h.vec_len, 0.f), lz(h.vec_len, 0.f);
        std::vector<double> dx(h.vec_len, 0.0), dy(h.vec_len, 0.0), dz(h.vec_len, 0.0);
        for (int64_t v = 0; v < h.vec_len; ++v) {
            if (pid[v] < 0) continue;
            dx[v] = P.px[pid[v]];
            dy[v] = P.py[pid[v]];
            dz[v] = P.pz[pid[v]];
            fx[v] = (float)dx[v];
            fy[v] = (float)dy[v];
            fz[v] = (float)dz[v];
            lx[v] = (float)(dx[v] - (double)fx[v]);
            ly[v] = (float)(dy[v] - (double)fy[v]);
            lz[v] = (float)(dz[v] - (double)fz[v]);
        }
        h.spid.upload(pid, s);
        h.sx32.upload(fx, s);
        h.sy32.upload(fy, s);
        h.sz32.upload(fz, s);
        h.sxl.upload(lx, s);
        h.syl.upload(ly, s);
        h.szl.upload(lz, s);
        h.sx64.upload(dx, s);
        h.sy64.upload(dy, s);
        h.sz64.upload(dz, s);
        H2W_HIP(hipStreamSynchronize(s));
        h.sp.x = h.sx32.p;
        h.sp.y = h.sy32.p;
        h.sp.z = h.sz32.p;
        h.sp.xl = h.sxl.p;
        h.sp.yl = h.syl.p;
        h.sp.zl = h.szl.p;
        h.sp.xd = h.sx64.p;
        h.sp.yd = h.sy64.p;
        h.sp.zd = h.sz64.p;
        h.sp.pid = h.spid.p;
        h.sp.dim = P.dim;
        h.sp.choice = h.choice;
        h.sp.diag = h.diag;
    }

    // ---- evaluate operators on the device
    {
        // transfers
        std::vector<int32_t> idx;
        std::vector<XferJob> xj;
        std::vector<double> lr;
        for (int f = 0; f < 2; ++f) {
            if (!fam_on[f]) continue;
            for (int j = nfam_lo[f]; j <= P.nL; ++j) {
                FamLevel &fl = h.fam[f][j];
                for (int b = 0; b < fl.nb; ++b) {
                    if (fl.e_off[b] < 0) continue;
                    XferJob J;
                    J.m = (int32_t)fl.S[b].size();
                    J.r = fl.rank[b];
                    J.s = (int64_t)idx.size();
                    idx.insert(idx.end(), fl.S[b].begin(), fl.S[b].end());
                    J.ic = (int64_t)idx.size();
                    idx.insert(idx.end(), fl.Ic[b].begin(), fl.Ic[b].end());
                    J.lr = (int64_t)lr.size() / esz;
                    lr.insert(lr.end(), fl.lr[b].begin(), fl.lr[b].end());
                    J.e = fl.e_off[b];
                    J.et = fl.et_off[b];
                    J.lde = (int32_t)pad_ld(J.m);
                    J.ldet = (int32_t)pad_ld(J.r);
                    xj.push_back(J);
                }
            }
        }
        // semi-nested row factors
        std::vector<CurJob> cj;
        for (auto &p : h.pairs) {
            if (p.rank == 0) continue;
            CurJob J;
            J.r = p.rank;
            J.n = (int32_t)P.bcount[p.level][p.v];
            J.ir = (int64_t)idx.size();
            idx.insert(idx.end(), p.Ir.begin(), p.Ir.end());
            J.v = (int64_t)idx.size();
            auto vr = range_list(P.bstart[p.level][p.v], J.n);
            idx.insert(idx.end(), vr.begin(), vr.end());
            J.lr = (int64_t)lr.size() / esz;
            lr.insert(lr.end(), p.lr.begin(), p.lr.end());
            J.a = p.ar_off;
            J.lda = (int32_t)pad_ld(J.r);
            J.pad = 0;
            cj.push_back(J);
        }
        DevBuf<int32_t> didx;
        DevBuf<double> dlr;
        DevBuf<XferJob> dxj;
        DevBuf<CurJob> dcj;
        didx.upload(idx, s);
        dlr.upload(lr, s);
        xj = split_xfer(xj);
        dxj.upload(xj, s);
        cj = split_cur(cj);
        dcj.upload(cj, s);
        if (h.cplx) {
            launch_xfer_c(dxj.p, (int)xj.size(), didx.p, reinterpret_cast<const double2 *>(dlr.p),
                          reinterpret_cast<double2 *>(h.mat64.p), h.kpc, s);
            launch_cur_c(dcj.p, (int)cj.size(), didx.p, reinterpret_cast<const double2 *>(dlr.p),
                         reinterpret_cast<double2 *>(h.mat64.p), h.kpc, s);
        } else {
            launch_xfer(dxj.p, (int)xj.size(), didx.p, dlr.p, h.mat64.p, h.kp, s);
            launch_cur(dcj.p, (int)cj.size(), didx.p, dlr.p, h.mat64.p, h.kp, s);
        }
        H2W_HIP(hipGetLastError());
        H2W_HIP(hipStreamSynchronize(s));
    }
    phase("transfer operators");
    {
        // plain entry blocks: M2L, near field, semi-nested column factors
        std::vector<int32_t> idx;
        std::vector<EvalJob> ej;
        auto add = [&](const std::vector<int32_t> &r, const std::vector<int32_t> &c, int64_t off) {
            EvalJob J;
            J.m = (int32_t)r.size();
            J.n = (int32_t)c.size();
            J.lda = (int32_t)pad_ld(J.m);
            J.pad = 0;
            J.rows = (int64_t)idx.size();
            idx.insert(idx.end(), r.begin(), r.end());
            J.cols = (int64_t)idx.size();
            idx.insert(idx.end(), c.begin(), c.end());
            J.a = off;
            ej.push_back(J);
        };
        DevBuf<int32_t> didx;
        DevBuf<EvalJob> dej;
        auto flush = [&]() {
            if (ej.empty()) return;
            didx.upload(idx, s);
            dej.upload(ej, s);
            if (h.cplx) launch_eval_c(dej.p, (int)ej.size(), didx.p, reinterpret_cast<double2 *>(h.mat64.p), h.kpc, s);
            else launch_eval(dej.p, (int)ej.size(), didx.p, h.mat64.p, h.kp, s);
            H2W_HIP(hipGetLastError());
            H2W_HIP(hipStreamSynchronize(s));
            idx.clear();
            ej.clear();
        };
        for (auto &m : m2l) {
            add(h.fam[m.f][m.j].Ir[m.b], h.fam[m.f][m.j].Ir[m.p], m.off);
            if (idx.size() > ((size_t)64 << 20)) flush();
        }
        for (auto &nb : nearb) {
            add(range_list(P.bstart[P.nL][nb.b], P.bcount[P.nL][nb.b]),
                range_list(P.bstart[P.nL][nb.n], P.bcount[P.nL][nb.n]), nb.off);
            if (idx.size() > ((size_t)64 << 20)) flush();
        }
        for (auto &p : h.pairs) {
            if (p.rank == 0) continue;
            add(range_list(P.bstart[p.level][p.b], P.bcount[p.level][p.b]), p.Ic, p.ac_off);
            if (idx.size() > ((size_t)64 << 20)) flush();
        }
        flush();
    }
    phase("M2L / near / CUR blocks");
    if (h.opts.fp32_shadow) {
        h.mat32.alloc(h.mat_len);
        launch_to_fp32(h.mat64.p, h.mat32.p, h.mat_len, s);
        H2W_HIP(hipGetLastError());
    } else {
        h.mat32.release();
    }
    // ---- GEMV work lists, one per launch of the matvec
    std::vector<WorkList> WL;
    auto fam_up_item = [&](WorkList &W, int f, int j, int b, int64_t src) {
        FamLevel &fl = h.fam[f][j];
        if (fl.et_off[b] < 0) return;
        add_item(W, fl.o_off[b], fl.rank[b], {TermSpec{fl.et_off[b], src, (int32_t)fl.S[b].size()}}, false,
                 (int)pad_ld(fl.rank[b]));
    };
    // up, leaf level (and semi-nested row factors, which read x directly)
    {
        WorkList W;
        for (int f = 0; f < 2; ++f) {
            if (!fam_on[f] || P.nL < nfam_lo[f]) continue;
            for (int b = 0; b < P.nboxes[P.nL]; ++b) fam_up_item(W, f, P.nL, b, h.off_x + P.bstart[P.nL][b]);
        }
        for (auto &p : h.pairs)
            if (p.rank > 0)
                add_item(W, p.t_off, p.rank,
                         {TermSpec{p.ar_off, h.off_x + P.bstart[p.level][p.v], (int32_t)P.bcount[p.level][p.v]}},
                         false, (int)pad_ld(p.rank));
        WL.push_back(std::move(W));
    }
    // up, levels nL-1 .. 1
    for (int j = P.nL - 1; j >= 1; --j) {
        WorkList W;
        W.step = P.nL - j;
        for (int f = 0; f < 2; ++f) {
            if (!fam_on[f] || j < nfam_lo[f]) continue;
            for (int b = 0; b < P.nboxes[j]; ++b)
                fam_up_item(W, f, j, b, h.fam[f][j + 1].o_off[b * P.nchild]);
        }
        WL.push_back(std::move(W));
    }
    // M2L (every level, both families) + near field, one block per unordered pair, both
    // directions in one fused launch
    {
        SymLists SL;
        for (int f = 0; f < 2; ++f) {
            if (!fam_on[f]) continue;
            for (int j = nfam_lo[f]; j <= P.nL; ++j) {
                FamLevel &fl = h.fam[f][j];
                SymGroup G;
                G.y.assign(fl.u_off.begin(), fl.u_off.end());
                G.x.assign(fl.o_off.begin(), fl.o_off.end());
                G.r = fl.rank;
                G.up.assign(fl.nb, {});
                G.self.assign(fl.nb, -1);
                for (auto &m : m2l)
                    if (m.f == f && m.j == j) G.up[m.b].push_back({m.p, m.off});
                sym_lists(G, SL);
            }
        }
        {
            SymGroup G;
            G.y.resize(nleaf);
            G.x.resize(nleaf);
            G.r.resize(nleaf);
            G.up.assign(nleaf, {});
            G.self.assign(nleaf, -1);
            for (int b = 0; b < nleaf; ++b) {
                G.y[b] = h.off_y + P.bstart[P.nL][b];
                G.x[b] = h.off_x + P.bstart[P.nL][b];
                G.r[b] = (int)P.bcount[P.nL][b];
            }
            for (auto &nb : nearb) {
                if (nb.n == nb.b) G.self[nb.b] = nb.off;
                else G.up[nb.b].push_back({nb.n, nb.off});
            }
            sym_lists(G, SL);
        }
        SL.fused.role = 1;
        SL.fused.stage = 1;
        WL.push_back(std::move(SL.fused));
        SL.trans.role = 3;
        SL.trans.stage = 1;
        SL.trans.step = 1;
        SL.otf.role = 1;
        SL.otf.stage = 1;
        std::vector<WorkList> tl, ol;
        tl.push_back(std::move(SL.trans));
        ol.push_back(std::move(SL.otf));
        upload_lists(tl, h.trans, s);
        upload_lists(ol, h.otf, s);
    }
    // down, levels 1 .. nL-1
    for (int j = 1; j <= P.nL - 1; ++j) {
        WorkList W;
        W.stage = 2;
        W.step = P.nL - j;
        for (int f = 0; f < 2; ++f) {
            if (!fam_on[f] || j < nfam_lo[f]) continue;
            FamLevel &fl = h.fam[f][j];
            for (int b = 0; b < fl.nb; ++b) {
                if (fl.e_off[b] < 0) continue;
                int64_t m = (int64_t)fl.S[b].size();
                add_item(W, h.fam[f][j + 1].u_off[b * P.nchild], (int)m,
                         {TermSpec{fl.e_off[b], fl.u_off[b], (int32_t)fl.rank[b]}}, false, (int)pad_ld(m));
            }
        }
        WL.push_back(std::move(W));
    }
    // down, leaf level: far + vertex into y
    {
        WorkList W;
        W.stage = 2;
        for (int b = 0; b < nleaf; ++b) {
            std::vector<TermSpec> t;
            for (int f = 0; f < 2; ++f) {
                if (!fam_on[f] || P.nL < nfam_lo[f]) continue;
                FamLevel &fl = h.fam[f][P.nL];
                if (fl.e_off[b] >= 0) t.push_back(TermSpec{fl.e_off[b], fl.u_off[b], (int32_t)fl.rank[b]});
            }
            add_item(W, h.off_y + P.bstart[P.nL][b], (int)P.bcount[P.nL][b], t, false,
                     (int)pad_ld(P.bcount[P.nL][b]));
        }
        WL.push_back(std::move(W));
    }
    // semi-nested column factors into y (boxes of all levels overlap: atomic)
    if (!h.pairs.empty()) {
        WorkList W;
        W.role = 2;
        W.stage = 3;
        std::vector<std::vector<std::vector<TermSpec>>> per(P.nL + 1);
        for (int j = 1; j <= P.nL; ++j) per[j].assign(P.nboxes[j], {});
        for (auto &p : h.pairs)
            if (p.rank > 0) per[p.level][p.b].push_back(TermSpec{p.ac_off, p.t_off, (int32_t)p.rank});
        for (int j = 1; j <= P.nL; ++j)
            for (int b = 0; b < P.nboxes[j]; ++b)
                add_item(W, h.off_y + P.bstart[j][b], (int)P.bcount[j][b], per[j][b], true,
                         (int)pad_ld(P.bcount[j][b]));
        WL.push_back(std::move(W));
    }
    std::vector<GItem> items;
    std::vector<GTerm> terms;
    for (auto &W : WL)
        if (W.role == 0) rebalance(W);
    h.launches.clear();
    for (auto &W : WL) {
        if (W.items.empty()) continue;
        Launch Lh;
        Lh.item0 = (int64_t)items.size();
        Lh.nitems = (int64_t)W.items.size();
        Lh.term0 = (int64_t)terms.size();
        Lh.nterms = (int64_t)W.terms.size();
        Lh.role = W.role;
        Lh.stage = W.stage;
        Lh.step = W.step;
        for (auto it : W.items) {
            it.t0 += (int32_t)Lh.term0;
            it.t1 += (int32_t)Lh.term0;
            items.push_back(it);
        }
        terms.insert(terms.end(), W.terms.begin(), W.terms.end());
        h.launches.push_back(Lh);
    }
    h.d_items.upload(items, s);
    h.d_terms.upload(terms, s);
    h.h_items = items;
    h.h_terms = terms;
    int64_t se = 0;
    h.launch_elems.assign(h.launches.size(), 0);
    for (size_t q = 0; q < h.launches.size(); ++q) {
        const Launch &Lh = h.launches[q];
        for (int64_t i = Lh.item0; i < Lh.item0 + Lh.nitems; ++i)
            for (int t = items[i].t0; t < items[i].t1; ++t) h.launch_elems[q] += (int64_t)items[i].rows * terms[t].cols;
        h.launches[q].avg_item_elems = h.launch_elems[q] / Lh.nitems;
        se += h.launch_elems[q];
    }
    h.stored_elems = se;
    build_inner(h, (double)h.opts.ir_tol_pow);
    H2W_HIP(hipStreamSynchronize(s));
    phase("work lists");
    if (verbose)
        std::fprintf(stderr, "[h2w] memory %.0f elements (%.3f GB fp64), pool %.3f GB, %zu launches, %zu items\n",
                     h.memory, h.memory * 8e-9, h.mat_len * 8e-9, h.launches.size(), items.size());
    h.assembled = true;
    h.assembly_seconds = now_s() - t0;
}

// ------------------------------------------------------------------ matvec
// y_tree = A x_tree inside the vector pool (x already at off_x)
}  // namespace

// operator phase of the assembly from skeletons already in the handle (loading a saved operator)
void rebuild_operators(Handle &h) {
    drop_graphs(h);
    assemble_operators(h, now_s());
}

bool is32(int precision) { return precision == H2W_FP32 || precision == H2W_FP32_OTF || precision == H2W_FP32_INNER; }

// Stage 1 is one fused launch (role 1): every stored block K(b, p), p >= b, applied to
// x_p, and its transpose to x_b with the product added to y_p.  The on-the-fly
// precisions evaluate the blocks from the points instead, over the directed lists.
void run_launch(Handle &h, const Launch &Lh, int precision) {
    hipStream_t s = h.stream;
    if (h.cplx) {  // complex operators: FP64 only, every role through the complex GEMV
        launch_cgemv(h.d_items.p + Lh.item0, (int)Lh.nitems, h.d_terms.p, reinterpret_cast<const double2 *>(h.mat64.p),
                     reinterpret_cast<double2 *>(h.vec64.p), s);
        return;
    }
    const bool otf = (precision == H2W_FP32_OTF || precision == H2W_FP64_OTF) && Lh.role == 1;
    if (otf) {
        for (auto &Lo : h.otf.launches) {
            if (is32(precision))
                launch_otf32(h.otf.d_items.p + Lo.item0, (int)Lo.nitems, h.otf.d_terms.p, h.sp, h.vec32.p, s);
            else
                launch_otf64(h.otf.d_items.p + Lo.item0, (int)Lo.nitems, h.otf.d_terms.p, h.sp, h.vec64.p, s);
        }
        return;
    }
    const bool in = precision == H2W_FP32_INNER;
    const GItem *items = (in ? h.inner.d_items.p : h.d_items.p) + Lh.item0;
    const GTerm *terms = in ? h.inner.d_terms.p : h.d_terms.p;
    if (in) launch_bgemv32(items, (int)Lh.nitems, terms, h.mat32.p, h.mat32in.p, h.vec32.p, Lh.role, Lh.unr, s);
    else if (is32(precision))
        launch_bgemv32(items, (int)Lh.nitems, terms, h.mat32.p, nullptr, h.vec32.p, Lh.role, Lh.unr32, s);
    else launch_bgemv64(items, (int)Lh.nitems, terms, h.mat64.p, nullptr, h.vec64.p, Lh.role, Lh.unr, s);
}

void zero_outputs(Handle &h, int precision) {
    if (h.cplx) launch_cfill(reinterpret_cast<double2 *>(h.vec64.p) + h.off_y, h.vec_len - h.off_y, h.stream);
    else if (is32(precision)) launch_fill32(h.vec32.p + h.off_y, h.vec_len - h.off_y, 0.0f, h.stream);
    else launch_fill(h.vec64.p + h.off_y, h.vec_len - h.off_y, 0.0, h.stream);
}

const std::vector<Launch> &launches_for(Handle &h, int precision) {
    if (precision == H2W_FP32_INNER) {
        if (h.inner.launches.empty()) throw Error{H2W_ESTATE, "the inner operator was not built (opts.ir_tol_pow)"};
        return h.inner.launches;
    }
    return h.launches;
}

void run_launches_eager(Handle &h, int precision) {
    const auto &L = launches_for(h, precision);
    zero_outputs(h, precision);
    for (auto &Lh : L) run_launch(h, Lh, precision);
    H2W_HIP(hipGetLastError());
}

// The launch sequence of a matvec is fixed per handle and precision (pool pointers and work
// lists do not change after assembly), so it is captured once into a HIP graph and replayed.
// Every node is a kernel (the output zeroing included).
void run_launches(Handle &h, int precision) {
    if (!h.use_graph) {
        run_launches_eager(h, precision);
        return;
    }
    if (!h.graph[precision]) {
        hipGraph_t g = nullptr;
        H2W_HIP(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
        try {
            run_launches_eager(h, precision);
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
        h.graph[precision] = ex;
    }
    H2W_HIP(hipGraphLaunch(h.graph[precision], h.stream));
}

// graphs of the single-vector and the kNR-wide matvec.  The stream is drained first, so no
// replay of an exec is in flight when it is destroyed, and callers that go on to free or
// reallocate the buffers it read (build_inner) free idle memory.
void drop_graph(Handle &h, int precision) {
    if (!h.graph[precision] && !h.graph_m[precision]) return;
    if (h.stream) (void)hipStreamSynchronize(h.stream);  // (also on error paths: never throws)
    for (hipGraphExec_t *g : {&h.graph[precision], &h.graph_m[precision]})
        if (*g) {
            (void)hipGraphExecDestroy(*g);
            *g = nullptr;
        }
}

void drop_graphs(Handle &h) {
    for (int p = 0; p < 5; ++p) drop_graph(h, p);
}

// device-pointer matvec on one vector; tree-ordered in/out when tree == true
void matvec_dev(Handle &h, const double *x, double *y, int precision, bool tree) {
    hipStream_t s = h.stream;
    const int64_t N = h.N;
    if (h.cplx) {
        // complex vectors: N (re, im) pairs
        if (precision != H2W_FP64) throw Error{H2W_EINVAL, "complex kernels: precision H2W_FP64 only"};
        double2 *pool = reinterpret_cast<double2 *>(h.vec64.p);
        const double2 *xc = reinterpret_cast<const double2 *>(x);
        double2 *yc = reinterpret_cast<double2 *>(y);
        if (tree) H2W_HIP(hipMemcpyAsync(pool + h.off_x, xc, N * 16, hipMemcpyDeviceToDevice, s));
        else launch_cpermute_in(xc, pool + h.off_x, h.d_perm.p, N, s);
        run_launches(h, precision);
        if (tree) H2W_HIP(hipMemcpyAsync(yc, pool + h.off_y, N * 16, hipMemcpyDeviceToDevice, s));
        else launch_cpermute_out(pool + h.off_y, yc, h.d_perm.p, N, s);
        H2W_HIP(hipGetLastError());
        return;
    }
    if (is32(precision)) {
        if (!h.mat32.p) throw Error{H2W_ESTATE, "FP32 matvec requested but the FP32 shadow was not built"};
        if (tree) launch_to_fp32(x, h.vec32.p + h.off_x, N, s);
        else launch_permute_in32(x, N, h.vec32.p + h.off_x, N, h.d_perm.p, N, 1, s);
        run_launches(h, precision);
        if (tree) launch_to_fp64(h.vec32.p + h.off_y, y, N, s);
        else launch_permute_out32(h.vec32.p + h.off_y, N, y, N, h.d_perm.p, N, 1, s);
    } else {
        if (tree) H2W_HIP(hipMemcpyAsync(h.vec64.p + h.off_x, x, N * 8, hipMemcpyDeviceToDevice, s));
        else launch_permute_in(x, N, h.vec64.p + h.off_x, N, h.d_perm.p, N, 1, s);
        run_launches(h, precision);
        if (tree) H2W_HIP(hipMemcpyAsync(y, h.vec64.p + h.off_y, N * 8, hipMemcpyDeviceToDevice, s));
        else launch_permute_out(h.vec64.p + h.off_y, N, y, N, h.d_perm.p, N, 1, s);
    }
    H2W_HIP(hipGetLastError());
}

}  // namespace h2w

namespace h2w {
namespace {

// ------------------------------------------------------------------ GMRES (tree order, device vectors)
// Unrestarted GMRES as the reference's iterative_solver::GMRES (1D/codes/gmres.hpp:91-200):
// restart = max_iter, Givens rotations (:30-75), resid = |s_{i+1}| / ||b|| (:162),
// back-substitution update (:77-90).  The Arnoldi step orthogonalises with
// classical Gram-Schmidt applied twice (two fused passes over the basis instead
// of the reference's i+1 dependent modified Gram-Schmidt steps); all of it FP64.
struct GmresOut {
    int iters = 0;
    double resid = 0.0;
    bool conv = false;
    int matvecs = 0;
    std::vector<std::vector<double>> H;  // Hessenberg columns (rows 0..j+1) of a plain solve
};

// A space kept from one correction of the refinement for the next ones (same operator):
// C = A U with orthonormal C, k columns; C sits in the first k columns of the GMRES basis
// buffer, so that one fused pass orthogonalises against C and the new Arnoldi vectors.
struct Recycled {
    int k = 0;
    // C spans approximate eigenvectors (a deflation space): C^T A v is small for v orthogonal to C,
    // so the second Gram-Schmidt pass is taken against the new Arnoldi vectors only
    bool invariant = false;
};

void ensure_V(Handle &h, int64_t cols) {
    if (cols <= h.g_Vcols) return;
    int64_t nc = std::max<int64_t>(cols, std::max<int64_t>(16, h.g_Vcols * 2));
    double *np = nullptr;
    if (hipMalloc(&np, (size_t)nc * h.N * 8) != hipSuccess) throw Error{H2W_ENOMEM, "GMRES basis allocation failed"};
    if (h.g_Vcols > 0)
        H2W_HIP(hipMemcpyAsync(np, h.g_V.p, (size_t)h.g_Vcols * h.N * 8, hipMemcpyDeviceToDevice, h.stream));
    H2W_HIP(hipStreamSynchronize(h.stream));
    h.g_V.release();
    h.g_V.p = np;
    h.g_V.n = (size_t)nc * h.N;
    h.g_Vcols = nc;
}

double dev_norm(Handle &h, const double *v) {
    int nb = mdot_blocks(h.N);
    h.g_part.ensure((size_t)nb * 4);
    h.g_h.ensure(8);
    launch_mdot(v, h.N, 1, v, h.N, h.g_part.p, nb, h.stream);
    launch_mdot_reduce(h.g_part.p, nb, 1, h.g_h.p, h.stream);
    double r = 0.0;
    H2W_HIP(hipMemcpyAsync(&r, h.g_h.p, 8, hipMemcpyDeviceToHost, h.stream));
    H2W_HIP(hipStreamSynchronize(h.stream));
    return std::sqrt(r);
}

double dev_dot(Handle &h, const double *a, const double *b) {
    int nb = mdot_blocks(h.N);
    h.g_part.ensure((size_t)nb * 4);
    h.g_h.ensure(8);
    launch_mdot(a, h.N, 1, b, h.N, h.g_part.p, nb, h.stream);
    launch_mdot_reduce(h.g_part.p, nb, 1, h.g_h.p, h.stream);
    double r = 0.0;
    H2W_HIP(hipMemcpyAsync(&r, h.g_h.p, 8, hipMemcpyDeviceToHost, h.stream));
    H2W_HIP(hipStreamSynchronize(h.stream));
    return r;
}


// Solves A x = b (tree order, device).  x_zero: x0 = 0 and no initial matvec.
// normb_in > 0 overrides ||b|| in the stopping test.
GmresOut gmres_core(Handle &h, const double *b, double *x, int max_iter, double tol, int prec, bool x_zero,
                    double normb_in, const Recycled *rc = nullptr, bool keep_h = false) {
    const int64_t N = h.N;
    hipStream_t s = h.stream;
    GmresOut out;
    h.g_w.ensure(N);
    h.g_b.ensure(N);  // residual of this solve
    double *w = h.g_w.p, *r = h.g_b.p;
    double normb = normb_in > 0 ? normb_in : dev_norm(h, b);
    // recycled space (x starts at zero): x = U C^T r, r -= C C^T r, and every Arnoldi vector is
    // orthogonalised against C as well (its coefficients B = C^T A v_j enter the update of x)
    const int kc = rc && x_zero ? rc->k : 0;
    const bool once = kc > 0 && rc->invariant;
    if (!once) h.gv_defl = -1;  // (its basis overwrites the first columns of g_V)
    const int m = max_iter;
    const int nb = mdot_blocks(N);
    h.g_part.ensure((size_t)nb * (kc + m + 2));
    h.g_h.ensure((size_t)2 * (kc + m + 2) + 2);
    if (x_zero) {
        H2W_HIP(hipMemcpyAsync(r, b, N * 8, hipMemcpyDeviceToDevice, s));
    } else {
        matvec_dev(h, x, w, prec, true);
        out.matvecs++;
        H2W_HIP(hipMemcpyAsync(r, b, N * 8, hipMemcpyDeviceToDevice, s));
        launch_axpby(-1.0, w, 1.0, r, N, s);
    }
    if (kc > 0) {
        launch_mdot(h.g_V.p, N, kc, r, N, h.g_part.p, nb, s);
        launch_mdot_reduce(h.g_part.p, nb, kc, h.g_h.p, s);
        launch_mgemv_sub(h.g_V.p, N, kc, h.g_h.p, r, N, s);
        launch_vec_update(h.g_U.p, N, kc, h.g_h.p, x, N, s);
    }
    double beta = dev_norm(h, r);
    if (normb == 0.0) normb = 1;
    out.resid = beta / normb;
    if (out.resid <= tol) {
        out.conv = true;
        return out;
    }
    std::vector<std::vector<double>> H;  // Hessenberg columns as they are produced (rows 0..j+1)
    std::vector<std::vector<double>> Bc; // (recycled) C^T A v_j
    std::vector<double> sv(m + 1, 0.0), cs(m + 1, 0.0), sn(m + 1, 0.0);
    auto Hh = [&](int i, int j) -> double & { return H[j][i]; };
    ensure_V(h, kc + std::min<int64_t>(m + 1, 32));
    launch_axpby(1.0 / beta, r, 0.0, h.g_V.p + (int64_t)kc * N, N, s);
    sv[0] = beta;
    double *hh = h.g_pin.ensure(2 * (kc + m + 2) + 2);
    for (int i = 0; i < m; ++i) {
        ensure_V(h, kc + i + 2);
        double *C = h.g_V.p, *V = h.g_V.p + (int64_t)kc * N;
        matvec_dev(h, V + (int64_t)i * N, w, prec, true);
        out.matvecs++;
        const int k = kc + i + 1;  // vectors orthogonalised against: C, v_0 .. v_i (contiguous)
        launch_mdot(C, N, k, w, N, h.g_part.p, nb, s);
        launch_mdot_reduce(h.g_part.p, nb, k, h.g_h.p, s);
        launch_mgemv_sub(C, N, k, h.g_h.p, w, N, s);
        // second pass: against C as well, or (a deflation space) against v_0 .. v_i only
        const int k0 = once ? kc : 0;
        if (once) launch_fill(h.g_h.p + k, kc, 0.0, s);
        launch_mdot(C + (int64_t)k0 * N, N, k - k0, w, N, h.g_part.p, nb, s);
        launch_mdot_reduce(h.g_part.p, nb, k - k0, h.g_h.p + k + k0, s);
        launch_mgemv_sub(C + (int64_t)k0 * N, N, k - k0, h.g_h.p + k + k0, w, N, s);
        launch_mdot(w, N, 1, w, N, h.g_part.p, nb, s);
        launch_mdot_reduce(h.g_part.p, nb, 1, h.g_h.p + 2 * k, s);
        H2W_HIP(hipMemcpyAsync(hh, h.g_h.p, (2 * k + 1) * 8, hipMemcpyDeviceToHost, s));
        H2W_HIP(hipStreamSynchronize(s));
        if (kc > 0) {
            Bc.emplace_back(kc);
            for (int q = 0; q < kc; ++q) Bc.back()[q] = hh[q] + hh[k + q];
        }
        H.emplace_back(i + 2, 0.0);
        for (int q = 0; q <= i; ++q) Hh(q, i) = hh[kc + q] + hh[k + kc + q];
        Hh(i + 1, i) = std::sqrt(hh[2 * k]);
        if (keep_h) out.H.push_back(H.back());  // (before the rotations)
        // H(i+1, i) = 0: the Krylov space is invariant and the least-squares solution is exact
        // (the reference's gap, 1D/codes/gmres.hpp:153-154, divides by it); the next basis
        // vector is left zero, the rotation below leaves s_{i+1} = 0 and the solve ends here
        const bool breakdown = !(Hh(i + 1, i) > 0.0);
        if (breakdown) launch_fill(V + (int64_t)(i + 1) * N, N, 0.0, s);
        else launch_axpby(1.0 / Hh(i + 1, i), w, 0.0, V + (int64_t)(i + 1) * N, N, s);
        for (int q = 0; q < i; ++q) app_rot(Hh(q, i), Hh(q + 1, i), cs[q], sn[q]);
        gen_rot(Hh(i, i), Hh(i + 1, i), cs[i], sn[i]);
        app_rot(Hh(i, i), Hh(i + 1, i), cs[i], sn[i]);
        app_rot(sv[i], sv[i + 1], cs[i], sn[i]);
        out.resid = std::fabs(sv[i + 1]) / normb;
        out.iters = i + 1;
        bool done = out.resid < tol || i == m - 1 || breakdown;
        if (done) {
            std::vector<double> y(sv.begin(), sv.begin() + i + 1);
            for (int q = i; q >= 0; q--) {
                y[q] /= Hh(q, q);
                for (int jj = q - 1; jj >= 0; jj--) y[jj] -= Hh(jj, q) * y[q];
            }
            // x += V y - U (B y)
            std::vector<double> cy(y);
            if (kc > 0) {
                cy.resize(i + 1 + kc, 0.0);
                for (int jj = 0; jj <= i; ++jj)
                    for (int q = 0; q < kc; ++q) cy[i + 1 + q] -= Bc[jj][q] * y[jj];
            }
            H2W_HIP(hipMemcpyAsync(h.g_h.p, cy.data(), cy.size() * 8, hipMemcpyHostToDevice, s));
            launch_vec_update(V, N, i + 1, h.g_h.p, x, N, s);
            if (kc > 0) launch_vec_update(h.g_U.p, N, kc, h.g_h.p + i + 1, x, N, s);
            H2W_HIP(hipStreamSynchronize(s));
            out.conv = out.resid < tol || (breakdown && out.resid == 0.0);
            return out;
        }
    }
    return out;
}

// The space of a finished correction kept for the next ones: with the Arnoldi relation
// A V_j = V_{j+1} Hbar, a selection V_j S of its space (S: j x k, the whole space when S is null,
// S = I) has A V_j S = V_{j+1} Z with Z = Hbar S; with Z = Q R (thin QR, host) U = V_j S R^-1 and
// C = V_{j+1} Q satisfy A U = C with orthonormal C.  C replaces the first k basis columns, U goes
// to g_U.  Returns false (nothing kept) for fewer than two iterations or a near-singular R.
bool keep_space(Handle &h, const GmresOut &o, Recycled &rc, const std::vector<std::vector<double>> *S = nullptr) {
    rc.k = 0;
    h.gv_defl = -1;
    const int j = (int)o.H.size();
    if (j < 2 || o.iters != j) return false;
    const int kk = S ? (int)S->size() : j;
    if (kk < 1) return false;
    const int64_t N = h.N;
    hipStream_t s = h.stream;
    // Z columns (j+1 rows), Gram-Schmidt twice in double
    std::vector<std::vector<double>> Q(kk, std::vector<double>(j + 1, 0.0));
    std::vector<std::vector<double>> R(kk, std::vector<double>(kk, 0.0));  // R[col][row]
    for (int c = 0; c < kk; ++c) {
        std::vector<double> v(j + 1, 0.0);
        if (S) {
            for (int t = 0; t < j; ++t) {
                const double sc = (*S)[c][t];
                for (int q = 0; q < (int)o.H[t].size() && q <= j; ++q) v[q] += o.H[t][q] * sc;
            }
        } else {
            for (int q = 0; q < (int)o.H[c].size() && q <= j; ++q) v[q] = o.H[c][q];
        }
        for (int pass = 0; pass < 2; ++pass)
            for (int p = 0; p < c; ++p) {
                double d = 0.0;
                for (int q = 0; q <= j; ++q) d += Q[p][q] * v[q];
                for (int q = 0; q <= j; ++q) v[q] -= d * Q[p][q];
                R[c][p] += d;
            }
        double nv = 0.0;
        for (int q = 0; q <= j; ++q) nv += v[q] * v[q];
        nv = std::sqrt(nv);
        const double ref = c == 0 ? nv : R[0][0];
        if (!std::isfinite(nv) || !(nv > 1e-12 * ref)) return false;
        R[c][c] = nv;
        for (int q = 0; q <= j; ++q) Q[c][q] = v[q] / nv;
    }
    // Rinv (upper triangular), column by column: R Rinv[:, c] = e_c
    std::vector<std::vector<double>> Ri(kk, std::vector<double>(kk, 0.0));
    for (int c = 0; c < kk; ++c)
        for (int q = c; q >= 0; --q) {
            double v = q == c ? 1.0 : 0.0;
            for (int p = q + 1; p <= c; ++p) v -= R[p][q] * Ri[c][p];
            Ri[c][q] = v / R[q][q];
        }
    // coefficients: C_c = V Q[:, c] (j+1 terms), U_c = V S Ri[:, c] (j terms, padded to j+1)
    std::vector<double> coef((size_t)2 * kk * (j + 1), 0.0);
    for (int c = 0; c < kk; ++c) {
        for (int q = 0; q <= j; ++q) coef[(size_t)c * (j + 1) + q] = Q[c][q];
        double *u = coef.data() + (size_t)(kk + c) * (j + 1);
        for (int p = 0; p <= c; ++p) {
            if (S)
                for (int q = 0; q < j; ++q) u[q] += (*S)[p][q] * Ri[c][p];
            else
                u[p] = Ri[c][p];
        }
    }
    h.g_coef.ensure(coef.size());
    H2W_HIP(hipMemcpyAsync(h.g_coef.p, coef.data(), coef.size() * 8, hipMemcpyHostToDevice, s));
    h.g_U.ensure((size_t)2 * kk * N);  // U, then C before it moves into the basis
    launch_fill(h.g_U.p, 2 * kk * N, 0.0, s);
    for (int c = 0; c < kk; ++c) {
        launch_vec_update(h.g_V.p, N, j + 1, h.g_coef.p + (size_t)c * (j + 1), h.g_U.p + (int64_t)(kk + c) * N, N, s);
        launch_vec_update(h.g_V.p, N, j + 1, h.g_coef.p + (size_t)(kk + c) * (j + 1), h.g_U.p + (int64_t)c * N, N, s);
    }
    H2W_HIP(hipMemcpyAsync(h.g_V.p, h.g_U.p + (int64_t)kk * N, (size_t)kk * N * 8, hipMemcpyDeviceToDevice, s));
    H2W_HIP(hipGetLastError());
    // (the coefficients must outlive the copy: wait before the host vector goes)
    H2W_HIP(hipStreamSynchronize(s));
    rc.k = kk;
    return true;
}

// Eigen-decomposition of the symmetric n x n matrix a (row-major, destroyed): Householder reduction
// to tridiagonal form T = Q^T A Q, then implicit QL iterations with shifts on T, the rotations
// accumulated into Q.  Eigenvalues in w, eigenvectors as the columns of z (row-major).
void sym_eig(std::vector<double> &a, int n, std::vector<double> &w, std::vector<double> &z) {
    auto A = [&](int i, int j) -> double & { return a[(size_t)i * n + j]; };
    z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
    std::vector<double> v(n), p(n), e(n, 0.0);
    // Householder: column k below the subdiagonal is zeroed by H = I - 2 v v^T on rows/cols k+1..n-1
    for (int k = 0; k + 2 < n; ++k) {
        double nx = 0.0;
        for (int i = k + 1; i < n; ++i) nx += A(i, k) * A(i, k);
        nx = std::sqrt(nx);
        if (nx == 0.0) continue;
        const double alpha = A(k + 1, k) > 0 ? -nx : nx;
        for (int i = k + 1; i < n; ++i) v[i] = A(i, k);
        v[k + 1] -= alpha;
        double nv = 0.0;
        for (int i = k + 1; i < n; ++i) nv += v[i] * v[i];
        nv = std::sqrt(nv);
        if (nv == 0.0) continue;
        for (int i = k + 1; i < n; ++i) v[i] /= nv;
        // p = A_sub v, K = v^T p, q = p - K v; A_sub -= 2 (v q^T + q v^T)
        double K = 0.0;
        for (int i = k + 1; i < n; ++i) {
            double s = 0.0;
            for (int j = k + 1; j < n; ++j) s += A(i, j) * v[j];
            p[i] = s;
            K += v[i] * s;
        }
        for (int i = k + 1; i < n; ++i) p[i] -= K * v[i];
        for (int i = k + 1; i < n; ++i)
            for (int j = k + 1; j < n; ++j) A(i, j) -= 2.0 * (v[i] * p[j] + p[i] * v[j]);
        A(k + 1, k) = A(k, k + 1) = alpha;
        for (int i = k + 2; i < n; ++i) A(i, k) = A(k, i) = 0.0;
        // Q <- Q H: Q[:, k+1:] -= 2 (Q[:, k+1:] v) v^T
        for (int r = 0; r < n; ++r) {
            double s = 0.0;
            for (int j = k + 1; j < n; ++j) s += z[(size_t)r * n + j] * v[j];
            s *= 2.0;
            for (int j = k + 1; j < n; ++j) z[(size_t)r * n + j] -= s * v[j];
        }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) w[i] = A(i, i);
    for (int i = 0; i + 1 < n; ++i) e[i] = A(i + 1, i);
    e[n - 1] = 0.0;
    // implicit QL with Wilkinson-type shifts: eigenvalue l converges when e[l] vanishes
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < n; ++l) {
        for (int iter = 0; iter < 60; ++iter) {
            int m = l;
            for (; m + 1 < n; ++m) {
                const double dd = std::fabs(w[m]) + std::fabs(w[m + 1]);
                if (std::fabs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            // shift from the leading 2 x 2 block at l
            double g = (w[l + 1] - w[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = w[m] - w[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, pp = 0.0;
            int i = m - 1;
            bool under = false;
            for (; i >= l; --i) {
                double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {  // (underflow: split here)
                    w[i + 1] -= pp;
                    e[m] = 0.0;
                    under = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = w[i + 1] - pp;
                r = (w[i] - g) * s + 2.0 * c * b;
                pp = s * r;
                w[i + 1] = g + pp;
                g = c * r - b;
                for (int k = 0; k < n; ++k) {
                    const double zi1 = z[(size_t)k * n + i + 1], zi = z[(size_t)k * n + i];
                    z[(size_t)k * n + i + 1] = s * zi + c * zi1;
                    z[(size_t)k * n + i] = c * zi - s * zi1;
                }
            }
            if (under) continue;
            w[l] -= pp;
            e[l] = g;
            e[m] = 0.0;
        }
    }
}

// The deflation space of the refinement's inner operator: m Arnoldi steps of that operator from a
// fixed pseudo-random vector (independent of any right-hand side).  A Krylov space of a generic
// vector holds good approximations of the operator's outlying eigenvectors, so a correction that
// starts from it (x = U C^T r, every Arnoldi vector orthogonalised against C) needs fewer steps.
// k = 0 keeps the whole space (U = V_m R^-1, C = V_{m+1} Q, as keep_space keeps a correction's
// space); 0 < k < m keeps the k Ritz vectors (eigenvectors of the symmetric part of H_m) whose
// Ritz values lie farthest from the diagonal value d: the operators here are d I plus a kernel
// matrix, their spectrum clusters at d and GMRES spends its steps on the eigenvalues away from it,
// so those k vectors give nearly the reduction of the whole space at k/m of the orthogonalisation
// cost per step.  m = 0 drops the space.
void build_deflation(Handle &h, int m, int k, int prec) {
    const int64_t N = h.N;
    hipStream_t s = h.stream;
    H2W_HIP(hipStreamSynchronize(s));
    h.defl_k = 0;
    h.defl_prec = -1;
    ++h.defl_serial;
    h.g_defl.release();
    h.g_defl_rm.release();
    if (m <= 0) return;
    std::vector<double> v(N);
    uint64_t st = 0x9e3779b97f4a7c15ull;
    for (int64_t i = 0; i < N; ++i) {  // splitmix64, uniform in [-1, 1)
        uint64_t z = (st += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        v[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    h.g_r.ensure(N);
    h.g_d.ensure(N);
    H2W_HIP(hipMemcpyAsync(h.g_r.p, v.data(), N * 8, hipMemcpyHostToDevice, s));
    launch_fill(h.g_d.p, N, 0.0, s);
    // tol 0: exactly m steps unless the space becomes invariant first
    GmresOut o = gmres_core(h, h.g_r.p, h.g_d.p, m, 0.0, prec, true, -1.0, nullptr, true);
    Recycled rc;
    const int j = (int)o.H.size();
    std::vector<std::vector<double>> S;
    if (k > 0 && k < j) {
        // symmetric part of H_j (the operator is symmetric; its FP32 products are so up to rounding)
        std::vector<double> T((size_t)j * j, 0.0), w, Z;
        for (int c = 0; c < j; ++c)
            for (int q = 0; q < (int)o.H[c].size() && q < j; ++q) {
                T[(size_t)q * j + c] += 0.5 * o.H[c][q];
                T[(size_t)c * j + q] += 0.5 * o.H[c][q];
            }
        sym_eig(T, j, w, Z);
        std::vector<int> idx(j);
        for (int i = 0; i < j; ++i) idx[i] = i;
        const double d = h.diag;
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int a, int b) { return std::fabs(w[a] - d) > std::fabs(w[b] - d); });
        S.assign(k, std::vector<double>(j));
        for (int c = 0; c < k; ++c)
            for (int t = 0; t < j; ++t) S[c][t] = Z[(size_t)t * j + idx[c]];
    }
    if (!keep_space(h, o, rc, S.empty() ? nullptr : &S))
        throw Error{H2W_EINVAL, "deflation space: fewer than 2 steps, or a singular Arnoldi relation"};
    h.g_defl.alloc((size_t)2 * rc.k * N);
    H2W_HIP(hipMemcpyAsync(h.g_defl.p, h.g_U.p, (size_t)2 * rc.k * N * 8, hipMemcpyDeviceToDevice, s));
    h.g_defl_rm.alloc((size_t)2 * rc.k * N);
    launch_to_rowmajor(h.g_defl.p, N, rc.k, h.g_defl_rm.p, s);
    launch_to_rowmajor(h.g_defl.p + (int64_t)rc.k * N, N, rc.k, h.g_defl_rm.p + (int64_t)rc.k * N, s);
    H2W_HIP(hipStreamSynchronize(s));
    h.defl_k = rc.k;
    h.defl_prec = prec;
    h.defl_gen = h.inner_gen;
}

}  // namespace

// Complex GMRES (the reference's -DUSE_Complex / USE_Helm / USE_Hankel branch of
// iterative_solver::GMRES, 1D/codes/gmres.hpp:52-67 and :91-200) on the tree-ordered complex
// right-hand side bt (N (re, im) pairs, device), starting from and updating h.g_x.  Unrestarted,
// Arnoldi by classical Gram-Schmidt twice with the Hermitian inner product h_{k,i} = v_k^H w, complex
// Givens rotations G = [c, s; -conj(s), c] with real c = |dx| / n and s = (dx / |dx|) conj(dy) / n,
// resid = |s_{i+1}| / ||b||, back-substitution.  (The reference's branch forms h_{k,i} as
// w.dot(v_k) = w^H v_k, the conjugate, and applies -s instead of -conj(s); for complex data those
// leave the Krylov basis non-orthogonal and the subdiagonal unreduced, so this restatement keeps the
// textbook forms, which coincide with the reference's for real data.)
h2w_gmres_stats solve_tree_c(Handle &h, const double *bt, int max_iter, double tol) {
    using cd = std::complex<double>;
    const int64_t N = h.N;
    hipStream_t s = h.stream;
    const double t0 = now_s();
    h2w_gmres_stats S;
    std::memset(&S, 0, sizeof(S));
    double2 *x = reinterpret_cast<double2 *>(h.g_x.p);
    const double2 *b = reinterpret_cast<const double2 *>(bt);
    h.g_w.ensure(2 * N);
    h.g_r.ensure(2 * N);
    double2 *w = reinterpret_cast<double2 *>(h.g_w.p), *r = reinterpret_cast<double2 *>(h.g_r.p);
    const int nb = cmdot_blocks(N);
    const int m = max_iter;
    h.g_part.ensure((size_t)2 * nb * (m + 2));
    h.g_h.ensure((size_t)2 * (2 * (m + 2) + 2));
    double2 *part = reinterpret_cast<double2 *>(h.g_part.p), *hv = reinterpret_cast<double2 *>(h.g_h.p);
    auto norm2 = [&](const double2 *v) {
        launch_cmdot(v, N, 1, v, N, part, nb, s);
        launch_cmdot_reduce(part, nb, 1, hv, s);
        double2 q;
        H2W_HIP(hipMemcpyAsync(&q, hv, 16, hipMemcpyDeviceToHost, s));
        H2W_HIP(hipStreamSynchronize(s));
        return std::sqrt(std::max(0.0, q.x));
    };
    // basis: (m + 1) complex vectors in g_V
    const int64_t vcols = std::min<int64_t>(m + 1, 1024);
    if ((int64_t)h.g_V.n < vcols * 2 * N) {
        h.g_V.alloc((size_t)vcols * 2 * N);
        h.g_Vcols = 0;  // (the real basis layout is rebuilt on its next use)
        h.gv_defl = -1;
    }
    double2 *V = reinterpret_cast<double2 *>(h.g_V.p);
    double normb = norm2(b);
    if (normb == 0.0) normb = 1.0;
    matvec_dev(h, h.g_x.p, h.g_w.p, H2W_FP64, true);
    S.matvecs_fp64 = 1;
    H2W_HIP(hipMemcpyAsync(r, b, N * 16, hipMemcpyDeviceToDevice, s));
    launch_caxpby(make_double2(-1.0, 0.0), w, make_double2(1.0, 0.0), r, N, s);
    const double beta = norm2(r);
    S.resid = beta / normb;
    if (S.resid <= tol) {
        S.converged = 1;
        S.seconds = now_s() - t0;
        return S;
    }
    const int mm = (int)std::min<int64_t>(m, vcols - 1);
    std::vector<std::vector<cd>> H;
    std::vector<cd> sv(mm + 1, cd(0.0)), cs(mm + 1, cd(0.0)), sn(mm + 1, cd(0.0));
    launch_caxpby(make_double2(1.0 / beta, 0.0), r, make_double2(0.0, 0.0), V, N, s);
    sv[0] = beta;
    std::vector<double2> hh(2 * (mm + 2) + 2);
    auto rot = [](cd &dx, cd &dy, cd c, cd sg) {
        const cd t = c * dx + sg * dy;
        dy = -std::conj(sg) * dx + c * dy;
        dx = t;
    };
    for (int i = 0; i < mm; ++i) {
        matvec_dev(h, reinterpret_cast<double *>(V + (int64_t)i * N), h.g_w.p, H2W_FP64, true);
        S.matvecs_fp64++;
        const int k = i + 1;
        launch_cmdot(V, N, k, w, N, part, nb, s);
        launch_cmdot_reduce(part, nb, k, hv, s);
        launch_cmgemv(V, N, k, hv, w, N, -1.0, s);
        launch_cmdot(V, N, k, w, N, part, nb, s);
        launch_cmdot_reduce(part, nb, k, hv + k, s);
        launch_cmgemv(V, N, k, hv + k, w, N, -1.0, s);
        launch_cmdot(w, N, 1, w, N, part, nb, s);
        launch_cmdot_reduce(part, nb, 1, hv + 2 * k, s);
        H2W_HIP(hipMemcpyAsync(hh.data(), hv, (2 * k + 1) * 16, hipMemcpyDeviceToHost, s));
        H2W_HIP(hipStreamSynchronize(s));
        H.emplace_back(i + 2, cd(0.0));
        std::vector<cd> &Hc = H.back();
        for (int q = 0; q <= i; ++q) Hc[q] = cd(hh[q].x + hh[k + q].x, hh[q].y + hh[k + q].y);
        const double hn = std::sqrt(std::max(0.0, hh[2 * k].x));
        Hc[i + 1] = hn;
        const bool breakdown = !(hn > 0.0);
        if (breakdown) launch_cfill(V + (int64_t)(i + 1) * N, N, s);
        else launch_caxpby(make_double2(1.0 / hn, 0.0), w, make_double2(0.0, 0.0), V + (int64_t)(i + 1) * N, N, s);
        for (int q = 0; q < i; ++q) rot(Hc[q], Hc[q + 1], cs[q], sn[q]);
        {
            const cd dx = Hc[i], dy = Hc[i + 1];
            if (std::abs(dx) == 0.0) {
                cs[i] = 0.0;
                sn[i] = 1.0;
            } else {
                const double sc = std::abs(dx) + std::abs(dy);
                const double nrm = sc * std::sqrt(std::norm(dx / sc) + std::norm(dy / sc));
                const cd alpha = dx / std::abs(dx);
                cs[i] = std::abs(dx) / nrm;
                sn[i] = alpha * std::conj(dy) / nrm;
            }
        }
        rot(Hc[i], Hc[i + 1], cs[i], sn[i]);
        rot(sv[i], sv[i + 1], cs[i], sn[i]);
        S.resid = std::abs(sv[i + 1]) / normb;
        S.iterations = i + 1;
        if (S.resid < tol || i == mm - 1 || breakdown) {
            std::vector<cd> y(sv.begin(), sv.begin() + i + 1);
            for (int q = i; q >= 0; --q) {
                y[q] /= H[q][q];
                for (int jj = q - 1; jj >= 0; --jj) y[jj] -= H[q][jj] * y[q];
            }
            std::vector<double2> yd(i + 1);
            for (int q = 0; q <= i; ++q) yd[q] = make_double2(y[q].real(), y[q].imag());
            H2W_HIP(hipMemcpyAsync(hv, yd.data(), (i + 1) * 16, hipMemcpyHostToDevice, s));
            launch_cmgemv(V, N, i + 1, hv, x, N, 1.0, s);
            S.converged = S.resid < tol || (breakdown && S.resid == 0.0);
            break;
        }
    }
    H2W_HIP(hipStreamSynchronize(s));
    S.seconds = now_s() - t0;
    return S;
}

// GMRES on the tree-ordered right-hand side bt (device), starting from and
// updating h.g_x: plain FP64 GMRES, or FP64 iterative refinement whose
// corrections come from GMRES on an FP32 operator.
h2w_gmres_stats solve_tree(Handle &h, const double *bt, int max_iter, double tol, int mode) {
    hipStream_t s = h.stream;
    const int64_t N = h.N;
    const int inner_prec = mode == H2W_GMRES_FP32_OTF_IR ? H2W_FP32_OTF
                           : mode == H2W_GMRES_FP32_INNER_IR ? H2W_FP32_INNER
                                                              : H2W_FP32;
    double t0 = now_s();
    h2w_gmres_stats S;
    std::memset(&S, 0, sizeof(S));
    if (mode == H2W_GMRES_FP64) {
        auto o = gmres_core(h, bt, h.g_x.p, max_iter, tol, H2W_FP64, false, -1.0);
        S.iterations = o.iters;
        S.resid = o.resid;
        S.matvecs_fp64 = o.matvecs;
        S.converged = o.conv;
    } else {
        // FP64 iterative refinement: r = b - A x in FP64, correction from GMRES on the FP32 operator.
        // The first residual is the FP32 operator's own: the first correction then solves the
        // FP32 system itself (x1 = A32^-1 b to the inner tolerance, as a low-precision first
        // solve), which is as close to the solution as a correction of the FP64 residual of x0
        // would be, one FP64 product earlier.  Convergence is only ever decided on an FP64
        // residual.
        double normb = dev_norm(h, bt);
        if (normb == 0.0) normb = 1.0;
        int total = 0, outer = 0;
        double rn = 0.0;
        bool fp64_resid = max_iter < 1;  // (no correction allowed: report x0's FP64 residual)
        Recycled rc;
        // a deflation space built for this inner operator serves every correction (it goes where
        // keep_space puts a kept space: U in g_U, C in the first basis columns)
        const bool defl = h.defl_k > 0 && h.defl_prec == inner_prec &&
                          (inner_prec != H2W_FP32_INNER || h.defl_gen == h.inner_gen);
        if (defl) {
            const int k = h.defl_k;
            ensure_V(h, k + 2);  // (a larger basis buffer keeps its columns)
            const bool fresh = h.g_U.n < (size_t)k * N;  // (ensure reallocates: nothing kept)
            h.g_U.ensure((size_t)k * N);
            if (fresh || h.gv_defl != h.defl_serial) {  // (still there after an earlier deflated solve)
                H2W_HIP(hipMemcpyAsync(h.g_U.p, h.g_defl.p, (size_t)k * N * 8, hipMemcpyDeviceToDevice, s));
                H2W_HIP(hipMemcpyAsync(h.g_V.p, h.g_defl.p + (int64_t)k * N, (size_t)k * N * 8,
                                       hipMemcpyDeviceToDevice, s));
                h.gv_defl = h.defl_serial;
            }
            rc.k = k;
            rc.invariant = true;
        }
        while (true) {
            if (fp64_resid) {
                matvec_dev(h, h.g_x.p, h.g_d.p, H2W_FP64, true);
                S.matvecs_fp64++;
            } else {
                matvec_dev(h, h.g_x.p, h.g_d.p, inner_prec, true);
                S.matvecs_fp32++;
            }
            if (outer == 0) {
                // start from alpha x0 with alpha minimising |b - alpha A x0| (no extra product):
                // the first correction then starts from a residual no larger than b
                const double dd = dev_dot(h, h.g_d.p, h.g_d.p);
                const double alpha = dd > 0.0 ? dev_dot(h, h.g_d.p, bt) / dd : 1.0;
                if (std::isfinite(alpha) && alpha != 1.0) {
                    launch_scal(alpha, h.g_x.p, N, s);
                    launch_scal(alpha, h.g_d.p, N, s);
                }
            }
            H2W_HIP(hipMemcpyAsync(h.g_r.p, bt, N * 8, hipMemcpyDeviceToDevice, s));
            launch_axpby(-1.0, h.g_d.p, 1.0, h.g_r.p, N, s);
            double nr = dev_norm(h, h.g_r.p);
            rn = nr / normb;
            if (fp64_resid && (rn <= tol || total >= max_iter || outer >= 50)) break;
            // (the first correction, from the FP32 residual, stops at sqrt(tol): the FP64 residual
            // after it is bounded by the FP32 operator's error anyway, and the second correction
            // finishes from there)
            double floor_eta = mode == H2W_GMRES_FP32_INNER_IR ? std::pow(10.0, -h.ir_exp) : 1e-6;
            if (!fp64_resid) floor_eta = std::max(floor_eta, std::sqrt(tol));
            double eta = std::min(1e-2, std::max(floor_eta, 0.1 * tol / rn));
            launch_fill(h.g_d.p, N, 0.0, s);
            // later corrections (same FP32 operator) recycle the first one's Krylov space
            auto o = gmres_core(h, h.g_r.p, h.g_d.p, max_iter - total, eta, inner_prec, true, nr,
                                rc.k > 0 ? &rc : nullptr, outer == 0 && !defl);
            if (outer == 0 && !defl) keep_space(h, o, rc);
            total += o.iters;
            S.matvecs_fp32 += o.matvecs;
            launch_axpby(1.0, h.g_d.p, 1.0, h.g_x.p, N, s);
            ++outer;
            // (no iteration and no recycled projection: the correction was zero)
            if (o.iters == 0 && fp64_resid && rc.k == 0) break;
            fp64_resid = true;
        }
        S.iterations = total;
        S.outer_iterations = outer;
        S.resid = rn;
        S.converged = rn <= tol;
    }
    H2W_HIP(hipStreamSynchronize(s));
    S.seconds = now_s() - t0;
    return S;
}


// ------------------------------------------------------------------ autotuner
namespace {

struct EventPair {
    Event a, b;
    double ms(hipStream_t) { return elapsed_ms(a, b); }
};

double time_launch(Handle &h, const Launch &Lh, int precision, int nrep, EventPair &ev) {
    run_launch(h, Lh, precision);
    H2W_HIP(hipEventRecord(ev.a, h.stream));
    for (int r = 0; r < nrep; ++r) run_launch(h, Lh, precision);
    H2W_HIP(hipEventRecord(ev.b, h.stream));
    H2W_HIP(hipGetLastError());
    return ev.ms(h.stream) / nrep;
}

// unit input in the pool's x segment, zero outputs
void tune_inputs(Handle &h, int precision) {
    if (is32(precision)) launch_fill32(h.vec32.p + h.off_x, h.N, 1.0f, h.stream);
    else launch_fill(h.vec64.p + h.off_x, h.N, 1.0, h.stream);
    zero_outputs(h, precision);
}

}  // namespace

// one whole matvec (graph replay when enabled), ms
double time_matvec(Handle &h, int precision, int nrep) {
    EventPair ev;
    tune_inputs(h, precision);
    run_launches(h, precision);
    H2W_HIP(hipEventRecord(ev.a, h.stream));
    for (int r = 0; r < nrep; ++r) run_launches(h, precision);
    H2W_HIP(hipEventRecord(ev.b, h.stream));
    return ev.ms(h.stream) / nrep;
}

// Per launch of one precision's matvec: 2, 4 or 8 column loads in flight per lane,
// whichever runs fastest (4 unless another is at least 2 % faster).
void tune_unroll(Handle &h, int precision) {
    std::vector<Launch> &L = precision == H2W_FP32_INNER ? h.inner.launches : h.launches;
    if (L.empty()) return;
    EventPair ev;
    tune_inputs(h, precision);
    for (auto &Lh : L) {
        int &u = precision == H2W_FP32 ? Lh.unr32 : Lh.unr;
        const int64_t elems = precision == H2W_FP32_INNER ? h.inner.elems[&Lh - &L[0]] : h.launch_elems[&Lh - &L[0]];
        const int nrep = elems > (int64_t)1 << 26 ? 3 : 8;
        u = 4;
        double tb = time_launch(h, Lh, precision, nrep, ev);
        int best = 4;
        for (int c : {2, 8}) {
            u = c;
            double t = time_launch(h, Lh, precision, nrep, ev);
            if (t < 0.98 * tb) {
                tb = t;
                best = c;
            }
        }
        u = best;
    }
    H2W_HIP(hipStreamSynchronize(h.stream));
    drop_graph(h, precision);
}

// Inner tolerance of the refinement: 10^-p for p = 4, 4.5, ..., 8 below tol_pow.  Each
// candidate solves A x = bt from x0 = 0 and from x0 = 1 (the reference's start); its score
// is the slower of the two, raised by a fifth when a final residual sits within 3x of the
// tolerance (a rounding change could then cost another refinement step).  The best score
// is kept.
void tune_inner(Handle &h, const double *bt, double tol, int max_iter, bool unroll, h2w_tune *out) {
    const double p0 = h.ir_exp > 0 ? h.ir_exp : (double)h.opts.ir_tol_pow;
    double best = 0.0, sbest = 1e300;
    out->n_cand = 0;
    for (int q = 8; q <= 16 && 0.5 * q < h.tol_pow && out->n_cand < H2W_TUNE_MAX_CAND; ++q) {
        const double p = 0.5 * q;
        build_inner(h, p);
        if (h.inner.launches.empty()) continue;
        if (unroll) tune_unroll(h, H2W_FP32_INNER);
        time_matvec(h, H2W_FP32_INNER, 1);  // captures the graphs outside the timed solves
        time_matvec(h, H2W_FP64, 1);
        h2w_gmres_stats S{};
        double worst = 0.0, rmax = 0.0;
        bool conv = true;
        for (int start = 0; start < 2; ++start) {
            launch_fill(h.g_x.p, h.N, start == 0 ? 0.0 : 1.0, h.stream);
            h2w_gmres_stats S2 = solve_tree(h, bt, max_iter, tol, H2W_GMRES_FP32_INNER_IR);
            conv = conv && S2.converged;
            rmax = std::max(rmax, S2.resid);
            if (S2.seconds > worst) {
                worst = S2.seconds;
                S = S2;
            }
        }
        const double score = worst * (rmax > tol / 2.0 ? 1.2 : 1.0);
        const int c = out->n_cand++;
        out->cand_ir_tol_exp[c] = p;
        out->cand_seconds[c] = worst;
        out->cand_iterations[c] = S.iterations;
        out->cand_outer[c] = S.outer_iterations;
        out->cand_converged[c] = conv;
        out->cand_inner_bytes[c] = h.inner_elems * 4;
        if (conv && score < sbest) {
            sbest = score;
            best = p;
        }
    }
    if (best == 0.0) best = p0;
    if (best != h.ir_exp || h.inner.launches.empty()) {
        build_inner(h, best);
        if (unroll) tune_unroll(h, H2W_FP32_INNER);
    }
    out->ir_tol_exp = h.ir_exp;
}

// Deflation spaces for the chosen inner operator, scored as tune_inner scores tolerances (the
// slower of the solves from x0 = 0 and x0 = 1, a fifth more when a final residual exceeds tol / 2);
// then, with retune_inner, the inner tolerances half a decade either side under the chosen
// space.  Leaves the best (inner tolerance, space) built.
void tune_deflation(Handle &h, const double *bt, double tol, int max_iter, bool unroll, bool retune_inner,
                    h2w_tune *out) {
    static const int cand[][2] = {{0, 0}, {64, 32}, {128, 48}, {256, 64}, {256, 96}, {384, 128}};
    out->n_defl_cand = 0;
    double sbest = 1e300, pbest = h.ir_exp;
    int bm = 0, bk = 0;
    auto score = [&](int m, int k) {
        build_deflation(h, m, k, H2W_FP32_INNER);
        double worst = 0.0, rmax = 0.0;
        bool conv = true;
        h2w_gmres_stats S{};
        for (int start = 0; start < 2; ++start) {
            launch_fill(h.g_x.p, h.N, start == 0 ? 0.0 : 1.0, h.stream);
            h2w_gmres_stats S2 = solve_tree(h, bt, max_iter, tol, H2W_GMRES_FP32_INNER_IR);
            conv = conv && S2.converged;
            rmax = std::max(rmax, S2.resid);
            if (S2.seconds > worst) {
                worst = S2.seconds;
                S = S2;
            }
        }
        const double sc = worst * (rmax > tol / 2.0 ? 1.2 : 1.0);
        if (out->n_defl_cand < H2W_TUNE_MAX_CAND) {
            const int c = out->n_defl_cand++;
            out->cand_defl_m[c] = m;
            out->cand_defl_k[c] = k;
            out->cand_defl_ir_tol_exp[c] = h.ir_exp;
            out->cand_defl_seconds[c] = worst;
            out->cand_defl_iterations[c] = S.iterations;
            out->cand_defl_outer[c] = S.outer_iterations;
            out->cand_defl_converged[c] = conv;
        }
        if (conv && sc < sbest) {
            sbest = sc;
            bm = m;
            bk = k;
            pbest = h.ir_exp;
        }
    };
    time_matvec(h, H2W_FP32_INNER, 1);  // (graphs captured outside the timed solves)
    time_matvec(h, H2W_FP64, 1);
    for (auto &c : cand)
        if ((int64_t)c[0] < h.N) score(c[0], c[1]);
    if (retune_inner && bm > 0) {
        const double p0 = h.ir_exp;
        const int m = bm, k = bk;
        for (double p : {p0 - 0.5, p0 + 0.5}) {
            if (!(p >= 4.0) || !(p < h.tol_pow)) continue;
            build_inner(h, p);
            if (h.inner.launches.empty()) continue;
            if (unroll) tune_unroll(h, H2W_FP32_INNER);
            time_matvec(h, H2W_FP32_INNER, 1);
            score(m, k);
        }
        if (pbest != h.ir_exp) {
            build_inner(h, pbest);
            if (unroll) tune_unroll(h, H2W_FP32_INNER);
            time_matvec(h, H2W_FP32_INNER, 1);
        }
    }
    const double t0 = now_s();
    build_deflation(h, bm, bk, H2W_FP32_INNER);
    out->defl_build_seconds = now_s() - t0;
    out->defl_m = bm;
    out->defl_k = bk;
    out->ir_tol_exp = h.ir_exp;
}
}  // namespace h2w

namespace h2w {
// The device side of a handle whose plan is built: stream, tree-order points and
// permutation on the device.  Needs a visible HIP device.
void attach_device(Handle &h) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        throw Error{H2W_ENOLIB, "no HIP device visible; libh2w has no CPU fallback"};
    if (h.opts.device < 0 || h.opts.device >= ndev) throw Error{H2W_EINVAL, "device ordinal out of range"};
    h.opts.plan_only = 0;
    H2W_HIP(hipSetDevice(h.opts.device));
    H2W_HIP(hipStreamCreateWithFlags(&h.stream, hipStreamNonBlocking));
    h.d_px.upload(h.plan.px, h.stream);
    h.d_py.upload(h.plan.py, h.stream);
    h.d_pz.upload(h.plan.pz, h.stream);
    h.d_perm.upload(h.plan.perm, h.stream);
    std::vector<int32_t> iperm(h.N);
    for (int64_t t = 0; t < h.N; ++t) iperm[h.plan.perm[t]] = (int32_t)t;
    h.d_iperm.upload(iperm, h.stream);
    H2W_HIP(hipStreamSynchronize(h.stream));
    noalloc_setup(h.opts.device);
    h.kp.dim = h.dim;
    h.kp.choice = h.choice;
    h.kp.diag = h.diag;
    h.kp.px = h.d_px.p;
    h.kp.py = h.d_py.p;
    h.kp.pz = h.d_pz.p;
    h.kpc.dim = h.dim;
    h.kpc.choice = h.choice;
    h.kpc.diag = h.diag;
    h.kpc.diag_im = h.diag_im;
    h.kpc.px = h.d_px.p;
    h.kpc.py = h.d_py.p;
    h.kpc.pz = h.d_pz.p;
}
}  // namespace h2w

// ------------------------------------------------------------------ C ABI
using h2w::Error;

#define H2W_TRY try {
#define H2W_CATCH                                         \
    }                                                     \
    catch (const Error &e) {                              \
        h2w::set_error(e.msg);                            \
        return e.code;                                    \
    }                                                     \
    catch (const std::bad_alloc &) {                      \
        h2w::set_error("host allocation failed");         \
        return H2W_ENOMEM;                                \
    }                                                     \
    catch (const std::exception &e) {                     \
        h2w::set_error(e.what());                         \
        return H2W_EINVAL;                                \
    }

extern "C" {

void h2w_default_opts(h2w_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->compat = 1;
    o->nlevels = 0;
    o->fp32_shadow = 1;
    o->device = 0;
    o->rank_cap = 256;
    o->workspace_bytes = 0;
    o->plan_only = 0;
    o->ir_tol_pow = 6;
}

const char *h2w_last_error(void) { return h2w::g_err.c_str(); }

int h2w_create(int dim, int variant, int64_t N, const double *points, int n_max, double L, int tol_pow, int choice,
               const h2w_opts *opts, h2w_handle **out) {
    H2W_TRY
    if (!out) throw Error{H2W_EINVAL, "out is NULL"};
    *out = nullptr;
    if (dim < 1 || dim > 3) throw Error{H2W_EINVAL, "dim must be 1, 2 or 3"};
    if (variant != H2W_NESTED && variant != H2W_SEMI_NESTED && variant != H2W_NON_NESTED)
        throw Error{H2W_EINVAL, "unknown variant"};
    if (variant == H2W_NON_NESTED && dim == 1) throw Error{H2W_EINVAL, "H* is defined for 2D and 3D"};
    if (choice == H2W_CHOICE_HELMHOLTZ && dim != 3)
        throw Error{H2W_EINVAL, "choice 2 (Helmholtz, e^{iR}/(4 pi)) is the 3D complex kernel"};
    if (choice == H2W_CHOICE_HANKEL && dim != 2)
        throw Error{H2W_EINVAL, "choice 10 (Hankel, (i/4) H0^(1)(R)) is the 2D complex kernel"};
    if (choice != 0 && choice != 1 && choice != H2W_CHOICE_HELMHOLTZ && choice != H2W_CHOICE_HANKEL)
        throw Error{H2W_EINVAL, "Wrong choice! choice must be 0 (IE), 1 (RBF), 2 (3D Helmholtz) or 10 (2D Hankel)"};
    if (N < 2 || N > INT32_MAX) throw Error{H2W_EINVAL, "N out of range"};
    if (n_max < 1) throw Error{H2W_EINVAL, "n_max must be positive"};
    if (!(L > 0.0)) throw Error{H2W_EINVAL, "L must be positive"};
    if (tol_pow < 0 || tol_pow > 20) throw Error{H2W_EINVAL, "tol_pow out of range"};
    h2w_opts o;
    h2w_default_opts(&o);
    if (opts) o = *opts;
    int ndev = 0;
    if (!o.plan_only) {
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw Error{H2W_ENOLIB, "no HIP device visible; libh2w has no CPU fallback"};
        if (o.device < 0 || o.device >= ndev) throw Error{H2W_EINVAL, "device ordinal out of range"};
    }
    std::vector<double> grid;
    const bool is_grid = points == nullptr;
    if (is_grid) {
        // the reference's uniform grids (1D/codes/HODLR1Dtree.hpp:59-65, 2D/codes/HODLR2Dtree.hpp:75-91,
        // 3D/codes/HODLR3Dtree.hpp:82-102)
        grid.assign((size_t)dim * N, 0.0);
        if (dim == 1) {
            for (int64_t k = 0; k < N; ++k) grid[k] = -1.0 + 2.0 * (k + 1.0) / (N + 1.0);
        } else {
            int64_t n = std::llround(dim == 2 ? std::sqrt((double)N) : std::cbrt((double)N));
            if ((dim == 2 && n * n != N) || (dim == 3 && n * n * n != N))
                throw Error{H2W_EINVAL, "the uniform grid needs N to be a square (2D) or a cube (3D)"};
            std::vector<double> nd(n);
            for (int64_t k = 0; k < n; ++k) nd[k] = -L + 2.0 * L * (k + 1.0) / (n + 1.0);
            if (dim == 2) {
                for (int64_t j = 0; j < n; ++j)
                    for (int64_t k = 0; k < n; ++k) {
                        grid[j * n + k] = nd[k];
                        grid[N + j * n + k] = nd[j];
                    }
            } else {
                for (int64_t j = 0; j < n; ++j)
                    for (int64_t k = 0; k < n; ++k)
                        for (int64_t i = 0; i < n; ++i) {
                            int64_t p = (j * n + k) * n + i;
                            grid[p] = nd[k];
                            grid[N + p] = nd[j];
                            grid[2 * N + p] = nd[i];
                        }
            }
        }
        points = grid.data();
    }
    int nl = o.nlevels > 0 ? o.nlevels : h2w::ref_nlevels(dim, N, n_max, is_grid);
    if (nl < 1) throw Error{H2W_EINVAL, "nLevels < 1: N must exceed n_max"};
    if (nl > (dim == 1 ? 24 : (dim == 2 ? 12 : 8))) throw Error{H2W_EINVAL, "tree too deep"};
    std::unique_ptr<h2w_handle> h(new h2w_handle());
    h->dim = dim;
    h->variant = variant;
    h->choice = choice;
    h->tol_pow = tol_pow;
    h->n_max = n_max;
    h->N = N;
    h->L = L;
    h->opts = o;
    h->use_graph = std::getenv("H2W_NO_GRAPH") == nullptr;
    h->sym16 = std::getenv("H2W_NO_SYM16") == nullptr;
    h->plan.build(dim, N, points, L, nl);
    h->cplx = choice == H2W_CHOICE_HELMHOLTZ || choice == H2W_CHOICE_HANKEL;
    if (h->cplx) {
        // complex kernels: FP64 operators only (no FP32 copies, no refinement operator)
        h->opts.fp32_shadow = 0;
        h->opts.ir_tol_pow = 0;
        h2w::diag_value_c(dim, choice, N, h->diag, h->diag_im);
    } else {
        h->diag = h2w::diag_value(dim, choice, N, o.compat);
    }
    if (!o.plan_only) h2w::attach_device(*h);
    *out = h.release();
    return H2W_OK;
    H2W_CATCH
}

int h2w_assemble(h2w_handle *h) {
    H2W_TRY
    if (!h) throw Error{H2W_EINVAL, "NULL handle"};
    if (h->opts.plan_only) throw Error{H2W_ESTATE, "handle was created with plan_only"};
    H2W_HIP(hipSetDevice(h->opts.device));
    h2w::assemble(*h);
    return H2W_OK;
    H2W_CATCH
}

int h2w_assemble_many(h2w_handle *const *hs, int n, int nthreads) {
    H2W_TRY
    if (!hs || n < 0 || nthreads < 1) throw Error{H2W_EINVAL, "bad arguments"};
    for (int i = 0; i < n; ++i) {
        if (!hs[i]) throw Error{H2W_EINVAL, "NULL handle"};
        if (hs[i]->opts.plan_only) throw Error{H2W_ESTATE, "handle was created with plan_only"};
        for (int j = 0; j < i; ++j)
            if (hs[j] == hs[i]) throw Error{H2W_EINVAL, "a handle appears twice"};
    }
    // host threads take the handles in turn; each assembles on its handle's own stream, so the
    // small systems' level-by-level launches overlap on the device
    std::atomic<int> next{0};
    std::mutex mu;
    int code = H2W_OK;
    std::string msg;
    auto work = [&]() {
        for (int i = next++; i < n; i = next++) {
            const int rc = h2w_assemble(hs[i]);
            if (rc != H2W_OK) {
                std::lock_guard<std::mutex> lk(mu);
                if (code == H2W_OK) {
                    code = rc;
                    msg = std::string("system ") + std::to_string(i) + ": " + h2w_last_error();
                }
            }
        }
    };
    const int nt = std::min(nthreads, n);
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work);
    if (nt >= 1) work();
    for (auto &t : th) t.join();
    if (code != H2W_OK) throw Error{code, msg};
    return H2W_OK;
    H2W_CATCH
}

int h2w_matvec_ex(h2w_handle *h, const double *x, double *y, int nrhs, int64_t ld, int precision, int order,
                  int where) {
    H2W_TRY
    if (!h || !x || !y) throw Error{H2W_EINVAL, "NULL argument"};
    if (!h->assembled) throw Error{H2W_ESTATE, "h2w_matvec before h2w_assemble"};
    if (nrhs < 1 || ld < h->N) throw Error{H2W_EINVAL, "bad nrhs / ld"};
    if (precision < H2W_FP64 || precision > H2W_FP32_INNER) throw Error{H2W_EINVAL, "bad precision"};
    H2W_HIP(hipSetDevice(h->opts.device));
    const int64_t N = h->N;
    if (h->cplx) {
        // complex kernels: x, y hold N (re, im) pairs per right-hand side, ld counts complex entries
        if (precision != H2W_FP64) throw Error{H2W_EINVAL, "complex kernels: precision H2W_FP64 only"};
        for (int r = 0; r < nrhs; ++r) {
            const double *xr = x + 2 * r * ld;
            double *yr = y + 2 * r * ld;
            if (where == H2W_HOST) {
                h->d_stage_in.ensure(2 * N);
                h->d_stage_out.ensure(2 * N);
                H2W_HIP(hipMemcpyAsync(h->d_stage_in.p, xr, N * 16, hipMemcpyHostToDevice, h->stream));
                h2w::matvec_dev(*h, h->d_stage_in.p, h->d_stage_out.p, H2W_FP64, order == H2W_ORDER_TREE);
                H2W_HIP(hipMemcpyAsync(yr, h->d_stage_out.p, N * 16, hipMemcpyDeviceToHost, h->stream));
            } else {
                h2w::matvec_dev(*h, xr, yr, H2W_FP64, order == H2W_ORDER_TREE);
            }
        }
        H2W_HIP(hipStreamSynchronize(h->stream));
        return H2W_OK;
    }
    if (nrhs > 1 && (precision == H2W_FP64 || precision == H2W_FP32 || precision == H2W_FP32_INNER)) {
        // kNR right-hand sides per pass on the matrix cores
        const double *xs = x;
        double *ys = y;
        h2w::DevBuf<double> hx, hy;
        if (where == H2W_HOST) {
            hx.alloc((size_t)ld * nrhs);
            hy.alloc((size_t)ld * nrhs);
            H2W_HIP(hipMemcpyAsync(hx.p, x, (size_t)ld * nrhs * 8, hipMemcpyHostToDevice, h->stream));
            xs = hx.p;
            ys = hy.p;
        }
        for (int c0 = 0; c0 < nrhs; c0 += h2w::kNR)
            h2w::matvec_multi_dev(*h, xs + c0 * ld, ld, ys + c0 * ld, ld, std::min(h2w::kNR, nrhs - c0), precision,
                                  order == H2W_ORDER_TREE);
        if (where == H2W_HOST)
            for (int c = 0; c < nrhs; ++c)
                H2W_HIP(hipMemcpyAsync(y + c * ld, hy.p + c * ld, N * 8, hipMemcpyDeviceToHost, h->stream));
        H2W_HIP(hipStreamSynchronize(h->stream));
        return H2W_OK;
    }
    for (int r = 0; r < nrhs; ++r) {
        const double *xr = x + r * ld;
        double *yr = y + r * ld;
        if (where == H2W_HOST) {
            h->d_stage_in.ensure(N);
            h->d_stage_out.ensure(N);
            H2W_HIP(hipMemcpyAsync(h->d_stage_in.p, xr, N * 8, hipMemcpyHostToDevice, h->stream));
            h2w::matvec_dev(*h, h->d_stage_in.p, h->d_stage_out.p, precision, order == H2W_ORDER_TREE);
            H2W_HIP(hipMemcpyAsync(yr, h->d_stage_out.p, N * 8, hipMemcpyDeviceToHost, h->stream));
        } else {
            h2w::matvec_dev(*h, xr, yr, precision, order == H2W_ORDER_TREE);
        }
    }
    H2W_HIP(hipStreamSynchronize(h->stream));
    return H2W_OK;
    H2W_CATCH
}

int h2w_matvec(h2w_handle *h, const double *x, double *y, int where) {
    return h2w_matvec_ex(h, x, y, 1, h ? h->N : 0, H2W_FP64, H2W_ORDER_INPUT, where);
}

int h2w_gmres(h2w_handle *h, const double *b, double *x, int max_iter, double tol, int mode, int where,
              h2w_gmres_stats *st) {
    H2W_TRY
    if (!h || !b || !x) throw Error{H2W_EINVAL, "NULL argument"};
    if (!h->assembled) throw Error{H2W_ESTATE, "h2w_gmres before h2w_assemble"};
    if (max_iter < 1) throw Error{H2W_EINVAL, "max_iter must be positive"};
    if (mode < H2W_GMRES_FP64 || mode > H2W_GMRES_FP32_INNER_IR) throw Error{H2W_EINVAL, "bad mode"};
    if (h->cplx) {
        // complex kernels: b, x hold N (re, im) pairs; FP64 GMRES with the complex branch's rotations
        if (mode != H2W_GMRES_FP64) throw Error{H2W_EINVAL, "complex kernels: mode H2W_GMRES_FP64 only"};
        H2W_HIP(hipSetDevice(h->opts.device));
        const int64_t N = h->N;
        hipStream_t s = h->stream;
        h->g_x.ensure(2 * N);
        h->d_stage_in.ensure(2 * N);
        h->d_stage_out.ensure(2 * N);
        const double *bsrc = b, *xsrc = x;
        if (where == H2W_HOST) {
            H2W_HIP(hipMemcpyAsync(h->d_stage_in.p, b, N * 16, hipMemcpyHostToDevice, s));
            H2W_HIP(hipMemcpyAsync(h->d_stage_out.p, x, N * 16, hipMemcpyHostToDevice, s));
            bsrc = h->d_stage_in.p;
            xsrc = h->d_stage_out.p;
        }
        h2w::DevBuf<double> bt;
        bt.alloc(2 * N);
        h2w::launch_cpermute_in(reinterpret_cast<const double2 *>(bsrc), reinterpret_cast<double2 *>(bt.p), h->d_perm.p, N,
                                s);
        h2w::launch_cpermute_in(reinterpret_cast<const double2 *>(xsrc), reinterpret_cast<double2 *>(h->g_x.p),
                                h->d_perm.p, N, s);
        H2W_HIP(hipStreamSynchronize(s));
        h2w_gmres_stats S = h2w::solve_tree_c(*h, bt.p, max_iter, tol);
        double *xdst = where == H2W_HOST ? h->d_stage_out.p : x;
        h2w::launch_cpermute_out(reinterpret_cast<const double2 *>(h->g_x.p), reinterpret_cast<double2 *>(xdst),
                                 h->d_perm.p, N, s);
        if (where == H2W_HOST) H2W_HIP(hipMemcpyAsync(x, h->d_stage_out.p, N * 16, hipMemcpyDeviceToHost, s));
        H2W_HIP(hipStreamSynchronize(s));
        if (st) *st = S;
        return S.converged ? H2W_OK : H2W_ENOCONV;
    }
    if (mode != H2W_GMRES_FP64 && !h->mat32.p) throw Error{H2W_ESTATE, "FP32 shadow not built"};
    if (mode == H2W_GMRES_FP32_INNER_IR && h->inner.launches.empty())
        throw Error{H2W_ESTATE, "the inner operator was not built (opts.ir_tol_pow)"};
    H2W_HIP(hipSetDevice(h->opts.device));
    const int64_t N = h->N;
    hipStream_t s = h->stream;
    h->g_x.ensure(N);
    h->g_r.ensure(N);
    h->g_d.ensure(N);
    h->d_stage_in.ensure(N);
    h->d_stage_out.ensure(N);
    // b and x0 to tree order
    const double *bsrc = b, *xsrc = x;
    if (where == H2W_HOST) {
        H2W_HIP(hipMemcpyAsync(h->d_stage_in.p, b, N * 8, hipMemcpyHostToDevice, s));
        H2W_HIP(hipMemcpyAsync(h->d_stage_out.p, x, N * 8, hipMemcpyHostToDevice, s));
        bsrc = h->d_stage_in.p;
        xsrc = h->d_stage_out.p;
    }
    h2w::DevBuf<double> bt;
    bt.alloc(N);
    h2w::launch_permute_in(bsrc, N, bt.p, N, h->d_perm.p, N, 1, s);
    h2w::launch_permute_in(xsrc, N, h->g_x.p, N, h->d_perm.p, N, 1, s);
    H2W_HIP(hipStreamSynchronize(s));
    h2w_gmres_stats S = h2w::solve_tree(*h, bt.p, max_iter, tol, mode);
    double *xdst = where == H2W_HOST ? h->d_stage_out.p : x;
    h2w::launch_permute_out(h->g_x.p, N, xdst, N, h->d_perm.p, N, 1, s);
    if (where == H2W_HOST) H2W_HIP(hipMemcpyAsync(x, h->d_stage_out.p, N * 8, hipMemcpyDeviceToHost, s));
    H2W_HIP(hipStreamSynchronize(s));
    if (st) *st = S;
    return S.converged ? H2W_OK : H2W_ENOCONV;
    H2W_CATCH
}

int h2w_gmres_multi(h2w_handle *h, const double *B, int64_t ldb, double *X, int64_t ldx, int nrhs, int max_iter,
                    double tol, int mode, int where, h2w_gmres_stats *st) {
    H2W_TRY
    if (!h || !B || !X || !st) throw Error{H2W_EINVAL, "NULL argument"};
    if (!h->assembled) throw Error{H2W_ESTATE, "h2w_gmres_multi before h2w_assemble"};
    if (h->cplx) throw Error{H2W_EINVAL, "complex kernels: several right-hand sides go through h2w_gmres"};
    if (max_iter < 1 || nrhs < 1 || ldb < h->N || ldx < h->N) throw Error{H2W_EINVAL, "bad max_iter / nrhs / ld"};
    if (mode != H2W_GMRES_FP64 && mode != H2W_GMRES_FP32_IR && mode != H2W_GMRES_FP32_INNER_IR)
        throw Error{H2W_EINVAL, "several right-hand sides: modes FP64, FP32_IR and FP32_INNER_IR"};
    if (mode != H2W_GMRES_FP64 && !h->mat32.p) throw Error{H2W_ESTATE, "FP32 shadow not built"};
    if (mode == H2W_GMRES_FP32_INNER_IR && h->inner.launches.empty())
        throw Error{H2W_ESTATE, "the inner operator was not built (opts.ir_tol_pow)"};
    H2W_HIP(hipSetDevice(h->opts.device));
    const int64_t N = h->N;
    hipStream_t s = h->stream;
    const int64_t len = N * h2w::kNR;
    h2w::DevBuf<double> hb, hx, bt, xt;
    const double *bs = B;
    double *xs = X;
    if (where == H2W_HOST) {
        hb.alloc((size_t)ldb * nrhs);
        hx.alloc((size_t)ldx * nrhs);
        H2W_HIP(hipMemcpyAsync(hb.p, B, (size_t)ldb * nrhs * 8, hipMemcpyHostToDevice, s));
        H2W_HIP(hipMemcpyAsync(hx.p, X, (size_t)ldx * nrhs * 8, hipMemcpyHostToDevice, s));
        bs = hb.p;
        xs = hx.p;
    }
    bt.alloc(len);
    xt.alloc(len);
    for (int c0 = 0; c0 < nrhs; c0 += h2w::kNR) {
        const int nc = std::min(h2w::kNR, nrhs - c0);
        h2w::launch_gather_m64(bs + c0 * ldb, ldb, nc, h->d_perm.p, N, bt.p, s);
        h2w::launch_gather_m64(xs + c0 * ldx, ldx, nc, h->d_perm.p, N, xt.p, s);
        h2w::solve_tree_m(*h, bt.p, xt.p, nc, max_iter, tol, mode, st + c0);
        h2w::launch_scatter_m64(xt.p, N, xs + c0 * ldx, ldx, nc, h->d_perm.p, s);
    }
    if (where == H2W_HOST)
        for (int c = 0; c < nrhs; ++c)
            H2W_HIP(hipMemcpyAsync(X + c * ldx, hx.p + c * ldx, N * 8, hipMemcpyDeviceToHost, s));
    H2W_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < nrhs; ++c)
        if (!st[c].converged) return H2W_ENOCONV;
    return H2W_OK;
    H2W_CATCH
}

int h2w_dense_matvec(h2w_handle *h, const double *x, double *y, const int64_t *rows, int64_t nrows, int where) {
    H2W_TRY
    if (!h || !x || !y) throw Error{H2W_EINVAL, "NULL argument"};
    if (h->opts.plan_only) throw Error{H2W_ESTATE, "handle was created with plan_only"};
    H2W_HIP(hipSetDevice(h->opts.device));
    const int64_t N = h->N;
    hipStream_t s = h->stream;
    if (!rows) nrows = N;
    if (nrows < 0) throw Error{H2W_EINVAL, "nrows < 0"};
    // inverse permutation: input index -> tree index
    std::vector<int32_t> tr(nrows);
    {
        std::vector<int32_t> inv(N);
        for (int64_t t = 0; t < N; ++t) inv[h->plan.perm[t]] = (int32_t)t;
        for (int64_t q = 0; q < nrows; ++q) {
            int64_t i = rows ? rows[q] : q;
            if (i < 0 || i >= N) throw Error{H2W_EINVAL, "row index out of range"};
            tr[q] = inv[i];
        }
    }
    h2w::DevBuf<int32_t> drows;
    drows.upload(tr, s);
    h2w::DevBuf<double> xin, xt, yo;
    if (h->cplx) {
        // complex kernels: x (N) and y (nrows) hold (re, im) pairs
        xt.alloc(2 * N);
        yo.alloc(2 * std::max<int64_t>(nrows, 1));
        const double *xs = x;
        if (where == H2W_HOST) {
            xin.alloc(2 * N);
            H2W_HIP(hipMemcpyAsync(xin.p, x, N * 16, hipMemcpyHostToDevice, s));
            xs = xin.p;
        }
        h2w::launch_cpermute_in(reinterpret_cast<const double2 *>(xs), reinterpret_cast<double2 *>(xt.p), h->d_perm.p, N,
                                s);
        h2w::launch_cdense_rows(drows.p, nrows, reinterpret_cast<const double2 *>(xt.p), reinterpret_cast<double2 *>(yo.p),
                                N, h->kpc, s);
        H2W_HIP(hipGetLastError());
        H2W_HIP(hipMemcpyAsync(y, yo.p, nrows * 16, where == H2W_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                               s));
        H2W_HIP(hipStreamSynchronize(s));
        return H2W_OK;
    }
    xt.alloc(N);
    yo.alloc(std::max<int64_t>(nrows, 1));
    const double *xs = x;
    if (where == H2W_HOST) {
        xin.alloc(N);
        H2W_HIP(hipMemcpyAsync(xin.p, x, N * 8, hipMemcpyHostToDevice, s));
        xs = xin.p;
    }
    h2w::launch_permute_in(xs, N, xt.p, N, h->d_perm.p, N, 1, s);
    h2w::launch_dense_rows(drows.p, nrows, xt.p, yo.p, N, h->kp, s);
    H2W_HIP(hipGetLastError());
    H2W_HIP(hipMemcpyAsync(y, yo.p, nrows * 8, where == H2W_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s));
    H2W_HIP(hipStreamSynchronize(s));
    return H2W_OK;
    H2W_CATCH
}

int h2w_get_stats(const h2w_handle *h, h2w_stats *st) {
    if (!h || !st) {
        h2w::set_error("NULL argument");
        return H2W_EINVAL;
    }
    std::memset(st, 0, sizeof(*st));
    st->dim = h->dim;
    st->variant = h->variant;
    st->choice = h->choice;
    st->nlevels = h->plan.nL;
    st->N = h->N;
    st->memory = h->memory;
    st->diag = h->diag;
    st->assembly_seconds = h->assembly_seconds;
    st->aca_reruns = h->aca_reruns;
    st->aca_batches = h->aca_batches;
    st->inner_tol_exp = h->inner.launches.empty() ? 0.0 : h->ir_exp;
    st->aca_pool_bytes = h->aca_pool_bytes;
    st->rank_pred_blocks = h->rank_pred_blocks;
    st->rank_pred_under = h->rank_pred_under;
    st->rank_pred_mae_log2 = h->rank_pred_blocks ? h->rank_pred_abs_log2 / h->rank_pred_blocks : 0.0;
    st->rank_fixed_blocks = h->rank_fix_blocks;
    st->rank_fixed_resumed = h->rank_fix_resumed;
    st->rank_fixed_extra_steps = h->rank_fix_extra_steps;
    st->rank_fixed_sampled_err = h->rank_fix_sampled_err;
    for (int f = 0; f < 2; ++f) {
        int mx = 0;
        double sum = 0.0;
        int64_t cnt = 0;
        for (auto &fl : h->fam[f])
            for (int r : fl.rank) {
                mx = std::max(mx, r);
                sum += r;
                ++cnt;
            }
        if (f == 1)
            for (auto &p : h->pairs) {
                mx = std::max(mx, p.rank);
                sum += p.rank;
                ++cnt;
            }
        if (f == 0) {
            st->max_rank_far = mx;
            st->mean_rank_far = cnt ? sum / cnt : 0.0;
        } else {
            st->max_rank_vertex = mx;
            st->mean_rank_vertex = cnt ? sum / cnt : 0.0;
        }
    }
    st->stored_bytes_fp64 = h->stored_elems * (h->cplx ? 16 : 8);
    st->diag_im = h->diag_im;
    st->stored_bytes_fp32 = h->mat32.p ? h->stored_elems * 4 : 0;
    st->n_launches_matvec = (int)h->launches.size() + 3;
    st->stored_bytes_inner_fp32 = h->inner_elems * 4;
    return H2W_OK;
}

int h2w_tree_order(const h2w_handle *h, int32_t *perm) {
    if (!h || !perm) {
        h2w::set_error("NULL argument");
        return H2W_EINVAL;
    }
    std::copy(h->plan.perm.begin(), h->plan.perm.end(), perm);
    return H2W_OK;
}

int h2w_box_skeleton(const h2w_handle *h, int family, int level, int box, int32_t *ir, int32_t *ic, int cap) {
    if (!h || family < 0 || family > 1 || level < 1 || level > h->plan.nL || box < 0 ||
        box >= h->plan.nboxes[level]) {
        h2w::set_error("bad skeleton query");
        return -H2W_EINVAL;
    }
    if (!h->assembled) {
        h2w::set_error("not assembled");
        return -H2W_ESTATE;
    }
    const auto &F = h->fam[family];
    if ((int)F.size() <= level || F[level].rank.empty()) return 0;
    const auto &fl = F[level];
    int r = fl.rank[box];
    for (int q = 0; q < r && q < cap; ++q) {
        if (ir) ir[q] = h->plan.perm[fl.Ir[box][q]];
        if (ic) ic[q] = h->plan.perm[fl.Ic[box][q]];
    }
    return r;
}

int h2w_interaction_list(const h2w_handle *h, int level, int box, int32_t *il, int32_t *isv, int cap) {
    if (!h || level < 1 || level > h->plan.nL || box < 0 || box >= h->plan.nboxes[level]) {
        h2w::set_error("bad list query");
        return -H2W_EINVAL;
    }
    const auto &l = h->plan.il[level][box];
    const auto &v = h->plan.ilv[level][box];
    for (size_t q = 0; q < l.size() && (int)q < cap; ++q) {
        if (il) il[q] = l[q];
        if (isv) isv[q] = v[q];
    }
    return (int)l.size();
}

int h2w_near_list(const h2w_handle *h, int level, int box, int32_t *nb, int cap) {
    if (!h || level != h->plan.nL || box < 0 || box >= h->plan.nboxes[level]) {
        h2w::set_error("near lists exist at the leaf level only");
        return -H2W_EINVAL;
    }
    const auto &l = h->plan.near[box];
    for (size_t q = 0; q < l.size() && (int)q < cap; ++q)
        if (nb) nb[q] = l[q];
    return (int)l.size();
}

int h2w_bench_stream_read(int device, int64_t bytes, int nrep, double *gbps) {
    H2W_TRY
    if (!gbps || nrep < 1 || bytes < (1 << 20)) throw Error{H2W_EINVAL, "bad argument"};
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        throw Error{H2W_ENOLIB, "no HIP device visible; libh2w has no CPU fallback"};
    if (device < 0 || device >= ndev) throw Error{H2W_EINVAL, "device ordinal out of range"};
    H2W_HIP(hipSetDevice(device));
    bytes -= bytes % h2w::stream_chunk_bytes();
    int cus = 0;
    H2W_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const int grid = 16 * std::max(cus, 1);
    h2w::DevBuf<uint8_t> buf;
    h2w::DevBuf<uint32_t> sink;
    buf.alloc(bytes);
    sink.alloc(grid);
    // (on the null stream, which waits for the device's other work before it starts)
    h2w::Event e0, e1;
    H2W_HIP(hipMemset(buf.p, 0x5A, bytes));
    h2w::launch_stream_read(buf.p, bytes, sink.p, grid, nullptr);  // warm-up
    H2W_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < nrep; ++r) h2w::launch_stream_read(buf.p, bytes, sink.p, grid, nullptr);
    H2W_HIP(hipEventRecord(e1, nullptr));
    H2W_HIP(hipGetLastError());
    const double ms_total = h2w::elapsed_ms(e0, e1);
    H2W_HIP(hipDeviceSynchronize());
    *gbps = (double)bytes * nrep / (ms_total * 1e-3) * 1e-9;
    return H2W_OK;
    H2W_CATCH
}

int h2w_bench_matvec(h2w_handle *h, int nrep, int precision, h2w_bench *out) {
    H2W_TRY
    if (!h || !out || nrep < 1) throw Error{H2W_EINVAL, "bad argument"};
    if (!h->assembled) throw Error{H2W_ESTATE, "not assembled"};
    if (precision < H2W_FP64 || precision > H2W_FP32_INNER) throw Error{H2W_EINVAL, "bad precision"};
    if (h2w::is32(precision) && !h->mat32.p) throw Error{H2W_ESTATE, "FP32 shadow not built"};
    if (h->cplx && precision != H2W_FP64) throw Error{H2W_EINVAL, "complex kernels: precision H2W_FP64 only"};
    const auto &LL = h2w::launches_for(*h, precision);
    H2W_HIP(hipSetDevice(h->opts.device));
    std::memset(out, 0, sizeof(*out));
    hipStream_t s = h->stream;
    const int64_t N = h->N;
    // deterministic input in the pool's x segment
    if (h->cplx) h2w::launch_fill(h->vec64.p + 2 * h->off_x, 2 * N, 1.0, s);
    else {
        h2w::launch_fill(h->vec64.p + h->off_x, N, 1.0, s);
        h2w::launch_fill32(h->vec32.p + h->off_x, N, 1.0f, s);
    }
    const int nl = std::min<int>((int)h2w::launches_for(*h, precision).size(), H2W_MAX_LAUNCHES - 1);
    std::vector<h2w::Event> ev(nl + 2);
    std::vector<double> acc(nl + 1, 0.0);
    // warm-up
    h2w::run_launches(*h, precision);
    H2W_HIP(hipStreamSynchronize(s));
    h2w::Event t0, t1;
    H2W_HIP(hipEventRecord(t0, s));
    for (int r = 0; r < nrep; ++r) h2w::run_launches(*h, precision);
    H2W_HIP(hipEventRecord(t1, s));
    out->ms_per_matvec = h2w::elapsed_ms(t0, t1) / nrep;
    // per-launch durations: fill + each bgemv launch
    for (int r = 0; r < nrep; ++r) {
        H2W_HIP(hipEventRecord(ev[0], s));
        h2w::zero_outputs(*h, precision);
        H2W_HIP(hipEventRecord(ev[1], s));
        for (int q = 0; q < nl; ++q) {
            h2w::run_launch(*h, LL[q], precision);
            H2W_HIP(hipEventRecord(ev[q + 2], s));
        }
        for (int q = 0; q <= nl; ++q) acc[q] += h2w::elapsed_ms(ev[q], ev[q + 1]);
    }
    out->n_launch = nl + 1;
    const int es = h->cplx ? 16 : h2w::is32(precision) ? 4 : 8;
    out->launch_ms[0] = acc[0] / nrep;
    out->launch_bytes[0] = (h->vec_len - h->off_y) * (int64_t)es;
    out->launch_role[0] = -1;
    for (int q = 0; q < nl; ++q) {
        out->launch_ms[q + 1] = acc[q + 1] / nrep;
        auto &Lh = LL[q];
        out->launch_role[q + 1] = Lh.role;
        out->launch_bytes[q + 1] =
            (precision == H2W_FP32_INNER ? h->inner.elems[q] : h->launch_elems[q]) * es;
        if ((precision == H2W_FP32_OTF || precision == H2W_FP64_OTF) && Lh.role == 1) {
            // the on-the-fly launches evaluate every entry of the directed lists (both
            // directions of each pair), element size times that count
            int64_t e = 0;
            for (auto v : h->otf.elems) e += v;
            out->launch_bytes[q + 1] = e * es;
        }
    }
    return H2W_OK;
    H2W_CATCH
}

int h2w_set_inner_tolerance(h2w_handle *h, double ir_tol_pow) {
    H2W_TRY
    if (!h) throw Error{H2W_EINVAL, "NULL handle"};
    if (!h->assembled) throw Error{H2W_ESTATE, "not assembled"};
    if (h->cplx) throw Error{H2W_EINVAL, "complex kernels: FP64 operators only (no inner operator)"};
    if (!h->mat32.p) throw Error{H2W_ESTATE, "the inner operator needs the FP32 shadow (opts.fp32_shadow)"};
    if (!(ir_tol_pow >= 0.0) || (ir_tol_pow > 0.0 && ir_tol_pow >= h->tol_pow))
        throw Error{H2W_EINVAL, "ir_tol_pow must be 0 or in (0, tol_pow)"};
    H2W_HIP(hipSetDevice(h->opts.device));
    h2w::build_inner(*h, ir_tol_pow);
    return H2W_OK;
    H2W_CATCH
}

int h2w_set_deflation(h2w_handle *h, int m, int k, int precision) {
    H2W_TRY
    if (!h) throw Error{H2W_EINVAL, "NULL handle"};
    if (!h->assembled) throw Error{H2W_ESTATE, "not assembled"};
    if (h->cplx) throw Error{H2W_EINVAL, "complex kernels: no refinement, no deflation space"};
    if (m < 0 || m > 512 || m == 1) throw Error{H2W_EINVAL, "m must be 0 or in [2, 512]"};
    if (k < 0 || k > m) throw Error{H2W_EINVAL, "k must be in [0, m]"};
    if (precision != H2W_FP32 && precision != H2W_FP32_INNER && precision != H2W_FP32_OTF)
        throw Error{H2W_EINVAL, "precision must be an inner precision (H2W_FP32, H2W_FP32_INNER, H2W_FP32_OTF)"};
    if ((precision == H2W_FP32 || precision == H2W_FP32_INNER) && !h->mat32.p)
        throw Error{H2W_ESTATE, "the FP32 operators need the FP32 shadow (opts.fp32_shadow)"};
    if (precision == H2W_FP32_INNER && !(h->ir_exp > 0.0)) throw Error{H2W_ESTATE, "no inner operator built"};
    if ((int64_t)m >= h->N) throw Error{H2W_EINVAL, "m must be below N"};
    H2W_HIP(hipSetDevice(h->opts.device));
    h2w::build_deflation(*h, m, k, precision);
    return H2W_OK;
    H2W_CATCH
}

int h2w_autotune(h2w_handle *h, const double *b, int where, double tol, int max_iter, int flags, h2w_tune *out) {
    H2W_TRY
    if (!h || !out) throw Error{H2W_EINVAL, "NULL argument"};
    if (!h->assembled) throw Error{H2W_ESTATE, "not assembled"};
    if (flags & ~(H2W_TUNE_UNROLL | H2W_TUNE_INNER | H2W_TUNE_DEFLATION)) throw Error{H2W_EINVAL, "unknown flags"};
    if (h->cplx) throw Error{H2W_EINVAL, "complex kernels: no autotuner (one complex GEMV kernel)"};
    const bool solves = (flags & (H2W_TUNE_INNER | H2W_TUNE_DEFLATION)) != 0;
    if (solves && (!b || max_iter < 1 || !(tol > 0.0)))
        throw Error{H2W_EINVAL, "inner and deflation tuning need b, tol > 0 and max_iter >= 1"};
    if (solves && !h->mat32.p) throw Error{H2W_ESTATE, "FP32 shadow not built"};
    if ((flags & H2W_TUNE_DEFLATION) && !(flags & H2W_TUNE_INNER) && h->inner.launches.empty())
        throw Error{H2W_ESTATE, "deflation tuning needs an inner operator"};
    H2W_HIP(hipSetDevice(h->opts.device));
    std::memset(out, 0, sizeof(*out));
    const int64_t N = h->N;
    hipStream_t s = h->stream;
    const int precs[3] = {H2W_FP64, H2W_FP32, H2W_FP32_INNER};
    auto have = [&](int p) {
        return p == H2W_FP64 || (h->mat32.p && (p != H2W_FP32_INNER || !h->inner.launches.empty()));
    };
    for (int q = 0; q < 3; ++q) out->matvec_ms_before[q] = have(precs[q]) ? h2w::time_matvec(*h, precs[q], 5) : 0.0;
    if (flags & H2W_TUNE_UNROLL)
        for (int q = 0; q < 2; ++q)
            if (have(precs[q])) h2w::tune_unroll(*h, precs[q]);
    h2w::DevBuf<double> bin, bt;
    if (solves) {
        h->g_x.ensure(N);
        h->g_r.ensure(N);
        h->g_d.ensure(N);
        bt.alloc(N);
        const double *bsrc = b;
        if (where == H2W_HOST) {
            bin.alloc(N);
            H2W_HIP(hipMemcpyAsync(bin.p, b, N * 8, hipMemcpyHostToDevice, s));
            bsrc = bin.p;
        }
        h2w::launch_permute_in(bsrc, N, bt.p, N, h->d_perm.p, N, 1, s);
        H2W_HIP(hipStreamSynchronize(s));
    }
    if (flags & H2W_TUNE_INNER) {
        h2w::tune_inner(*h, bt.p, tol, max_iter, (flags & H2W_TUNE_UNROLL) != 0, out);
    } else {
        if ((flags & H2W_TUNE_UNROLL) && have(H2W_FP32_INNER)) h2w::tune_unroll(*h, H2W_FP32_INNER);
        out->ir_tol_exp = h->inner.launches.empty() ? 0.0 : h->ir_exp;
    }
    if ((flags & H2W_TUNE_DEFLATION) && !h->inner.launches.empty())
        h2w::tune_deflation(*h, bt.p, tol, max_iter, (flags & H2W_TUNE_UNROLL) != 0, (flags & H2W_TUNE_INNER) != 0,
                            out);
    for (int q = 0; q < 3; ++q) out->matvec_ms_after[q] = have(precs[q]) ? h2w::time_matvec(*h, precs[q], 5) : 0.0;
    H2W_HIP(hipStreamSynchronize(s));
    return H2W_OK;
    H2W_CATCH
}

int h2w_set_rank_model(h2w_handle *h, const uint16_t *params, int64_t count) {
    H2W_TRY
    if (!h) throw Error{H2W_EINVAL, "NULL handle"};
    if (params && count != h2w::kRankParams)
        throw Error{H2W_EINVAL, "rank model: expected " + std::to_string(h2w::kRankParams) + " BF16 parameters"};
    if (h->opts.plan_only) throw Error{H2W_ESTATE, "handle was created with plan_only"};
    H2W_HIP(hipSetDevice(h->opts.device));
    if (!params) {
        h->rank_model.release();
        return H2W_OK;
    }
    std::vector<uint16_t> v(params, params + count);
    h->rank_model.upload(v, h->stream);
    H2W_HIP(hipStreamSynchronize(h->stream));
    return H2W_OK;
    H2W_CATCH
}

int64_t h2w_rank_samples(const h2w_handle *h, float *feat, int32_t *rank, int64_t cap) {
    if (!h) {
        h2w::set_error("NULL handle");
        return -H2W_EINVAL;
    }
    const int64_t n = (int64_t)h->sample_rank.size();
    for (int64_t q = 0; q < std::min(n, cap); ++q) {
        if (feat) std::copy(h->sample_feat[q].begin(), h->sample_feat[q].end(), feat + q * h2w::kRankFeatures);
        if (rank) rank[q] = h->sample_rank[q];
    }
    return n;
}

int h2w_rank_predict(h2w_handle *h, const float *feat, int64_t n, float *out) {
    H2W_TRY
    if (!h || (!feat && n > 0) || (!out && n > 0) || n < 0) throw Error{H2W_EINVAL, "bad arguments"};
    if (!h->rank_model.p) throw Error{H2W_ESTATE, "no rank model (h2w_set_rank_model)"};
    H2W_HIP(hipSetDevice(h->opts.device));
    if (n == 0) return H2W_OK;
    std::vector<float> F(feat, feat + n * h2w::kRankFeatures);
    h2w::DevBuf<float> dF, dO;
    dF.upload(F, h->stream);
    dO.alloc(n);
    h2w::launch_rank_mlp(dF.p, (int)n, h->rank_model.p, dO.p, h->stream);
    H2W_HIP(hipGetLastError());
    H2W_HIP(hipMemcpyAsync(out, dO.p, n * 4, hipMemcpyDeviceToHost, h->stream));
    H2W_HIP(hipStreamSynchronize(h->stream));
    return H2W_OK;
    H2W_CATCH
}

void h2w_destroy(h2w_handle *h) {
    if (!h) return;
    if (!h->opts.plan_only) (void)hipSetDevice(h->opts.device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h2w::drop_graphs(*h);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // extern "C"
