// kernels.hip -- gfx950 kernels of libh2w.
//
// This file is compiled with -ffp-contract=off so that the kernel entries
// and the ACA residual updates round like the reference's scalar code
// (no fused multiply-add); the streaming GEMV and BLAS-1 loops ask for fused
// multiply-adds explicitly with fma().
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "h2w_internal.hpp"

namespace h2w {

namespace {

constexpr double kPI = 3.141592653589793238462643383279502884;

// ------------------------------------------------------------------ kernel entry
// A(i,j) of the reference's userkernel::getMatrixEntry for choice 0 (IE) and
// 1 (RBF): 1D/codes/kernel.hpp:108-185, 2D/codes/kernel.hpp:118-228,299-318,
// 3D/codes/kernel.hpp:123-192.  Off-diagonal entries depend on the pair's
// distance; i == j returns the diagonal rule (kp.diag).
// Flop weights (add/mul = 1, sqrt/log/div = 1 each): 3D IE 9, 3D RBF 10.
__device__ __forceinline__ double koff(int dim, int choice, double dx, double dy, double dz) {
    const double a = 0.0001;
    if (dim == 1) {
        double R = fabs(dx);
        if (choice == 0) return log(R);
        return R < a ? R / a : a / R;
    } else if (dim == 2) {
        double R2 = dx * dx + dy * dy;
        if (choice == 0) return (-1.0 / (2 * kPI)) * 0.5 * log(R2);
        double R = sqrt(R2);
        return R < a ? R / a : a / R;
    } else {
        double R2 = dx * dx + dy * dy + dz * dz;
        double R = sqrt(R2);
        if (choice == 0) return 1.0 / (4.0 * kPI * R);
        if (R < a) return (R * log(R) - 1.0) / (a * log(a) - 1.0);
        return (1.0 * log(R)) / log(a);
    }
}

struct PointI {
    double x, y, z;
    int i;
};
__device__ __forceinline__ PointI load_point(const KParams &kp, int i) {
    PointI p;
    p.i = i;
    p.x = kp.px[i];
    p.y = kp.dim > 1 ? kp.py[i] : 0.0;
    p.z = kp.dim > 2 ? kp.pz[i] : 0.0;
    return p;
}
__device__ __forceinline__ double kentry(const KParams &kp, const PointI &p, int j) {
    if (p.i == j) return kp.diag;
    double dx = p.x - kp.px[j];
    double dy = kp.dim > 1 ? p.y - kp.py[j] : 0.0;
    double dz = kp.dim > 2 ? p.z - kp.pz[j] : 0.0;
    return koff(kp.dim, kp.choice, dx, dy, dz);
}
__device__ __forceinline__ double kentry2(const KParams &kp, int i, int j) {
    return kentry(kp, load_point(kp, i), j);
}

// ------------------------------------------------------------------ reductions (wave64)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// argmax of |v| with first-index tie break; (a = |v|, idx, sv = signed v)
struct AMax {
    double a;
    int idx;
    double sv;
};
__device__ __forceinline__ void amax_merge(AMax &x, const AMax &y) {
    if (y.a > x.a || (y.a == x.a && y.idx < x.idx)) x = y;
}
__device__ __forceinline__ AMax wave_amax(AMax v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        AMax w;
        w.a = __shfl_xor(v.a, o, 64);
        w.idx = __shfl_xor(v.idx, o, 64);
        w.sv = __shfl_xor(v.sv, o, 64);
        amax_merge(v, w);
    }
    return v;
}
__device__ __forceinline__ void amax_update(AMax &m, double v, int idx) {
    double a = fabs(v);
    if (a > m.a || (a == m.a && idx < m.idx)) {
        m.a = a;
        m.idx = idx;
        m.sv = v;
    }
}

template <int NT>
struct BlockRed {
    static constexpr int NW = NT / 64;
    double s0[NW], s1[NW];
    double a[NW], sv[NW];
    int idx[NW];
};

// block-wide (sum, sum, argmax); all threads receive the result
template <int NT>
__device__ __forceinline__ void block_reduce(BlockRed<NT> &sh, double &s0, double &s1, AMax &m) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    m = wave_amax(m);
    __syncthreads();
    if (lane == 0) {
        sh.s0[w] = s0;
        sh.s1[w] = s1;
        sh.a[w] = m.a;
        sh.idx[w] = m.idx;
        sh.sv[w] = m.sv;
    }
    __syncthreads();
    s0 = 0.0;
    s1 = 0.0;
    m.a = -1.0;
    m.idx = INT_MAX;
    m.sv = 0.0;
    for (int q = 0; q < BlockRed<NT>::NW; ++q) {
        s0 += sh.s0[q];
        s1 += sh.s1[q];
        AMax o{sh.a[q], sh.idx[q], sh.sv[q]};
        amax_merge(m, o);
    }
}

// ------------------------------------------------------------------ batched ACA
// One workgroup per block.  Restates LowRank::ACA_only_nodes / ACA_only_nodesCUR
// (2D/codes/aca.hpp:158-446) and ACA_only_nodesUV (1D/codes/aca.hpp:40-156),
// which share the pivot loop:
//   - first row (in order) whose max |entry| exceeds 1e-36 starts the cross,
//   - pivot column = first argmax of |residual row| over unused columns,
//   - u = residual column / pivot, next row = first argmax of |residual column|
//     over unused rows,
//   - stop when |u||v| <= tol * normS or rank = min(m, n), or when |u| or |v|
//     falls below 1e-36 (that step is dropped),
//   - normS^2 += (|u||v|)^2 + 2 sum_{l<=new} |<U_l,u><V_l,v>| (the sum includes
//     the newest term, as the reference's loop does, 2D/codes/aca.hpp:404-409).
// Residual rows/columns subtract the previous terms one by one in rank order
// with separate multiply and subtract, as the reference does.
// U and V columns live in a workspace pool shared by the batch: before each
// step the workgroup's leader reserves m + n doubles with an atomic bump; the
// column offsets go to the job's table.  status = 1 when the table or the pool
// is full before the cross converged (the host re-runs the block alone).
constexpr int kChunk = 512;

template <int NT>
__device__ __forceinline__ bool aca_reserve(int64_t *tab, int k, int m, int n, int tcap,
                                            unsigned long long *ws_top, int64_t ws_cap, int64_t *sh_off) {
    if (threadIdx.x == 0) {
        int64_t off = -1;
        if (k < tcap) {
            unsigned long long o = atomicAdd(ws_top, (unsigned long long)(m + n));
            if ((int64_t)o + m + n <= ws_cap) off = (int64_t)o;
        }
        *sh_off = off;
        if (off >= 0) {
            tab[2 * k] = off;
            tab[2 * k + 1] = off + m;
        }
    }
    __syncthreads();
    return *sh_off >= 0;
}

template <int NT>
__global__ __launch_bounds__(NT) void k_aca(const AcaJob *__restrict__ jobs, const int32_t *__restrict__ idx,
                                            double *__restrict__ ws, int64_t ws_cap, unsigned long long *ws_top,
                                            int64_t *__restrict__ tab_all, uint8_t *__restrict__ flags,
                                            int32_t *__restrict__ rank_out, int32_t *__restrict__ status_out,
                                            int32_t *__restrict__ rowpiv, int32_t *__restrict__ colpiv, KParams kp,
                                            double tol, double *__restrict__ hist_all) {
    const AcaJob jb = jobs[blockIdx.x];
    const int m = jb.m, n = jb.n, tcap = jb.tcap;
    double *hist = hist_all + 2 * jb.piv;  // per accepted step: |u||v|, normS after the step
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    constexpr int NW = NT / 64;
    const int32_t *rows = idx + jb.rows;
    const int32_t *cols = idx + jb.cols;
    int64_t *tab = tab_all + jb.tab;
    uint8_t *ru = flags + jb.wsF;
    uint8_t *cu = ru + m;
    int32_t *rp_out = rowpiv + jb.piv;
    int32_t *cp_out = colpiv + jb.piv;
    __shared__ double coef[kChunk];
    __shared__ int64_t cofs[kChunk];
    __shared__ BlockRed<NT> red;
    __shared__ double dacc[NW];
    __shared__ int64_t sh_off;

    if (m == 0 || n == 0 || tcap == 0) {
        if (tid == 0) {
            rank_out[blockIdx.x] = 0;
            status_out[blockIdx.x] = 0;
        }
        return;
    }
    for (int i = tid; i < m; i += NT) ru[i] = 0;
    for (int s = tid; s < n; s += NT) cu[s] = 0;
    const int mn = min(m, n);
    if (!aca_reserve<NT>(tab, 0, m, n, tcap, ws_top, ws_cap, &sh_off)) {
        if (tid == 0) {
            rank_out[blockIdx.x] = 0;
            status_out[blockIdx.x] = 1;
        }
        return;
    }
    double *U0 = ws + sh_off;
    double *V0 = U0 + m;

    // ---- first row with a nonzero entry
    int l = 0, cp = -1;
    double piv = 0.0;
    for (l = 0; l < m; ++l) {
        PointI pr = load_point(kp, rows[l]);
        AMax am{-1.0, INT_MAX, 0.0};
        for (int s = tid; s < n; s += NT) {
            double v = kentry(kp, pr, cols[s]);
            V0[s] = v;
            amax_update(am, v, s);
        }
        double d0 = 0.0, d1 = 0.0;
        block_reduce<NT>(red, d0, d1, am);
        if (am.idx != INT_MAX && am.a > 1e-36) {
            cp = am.idx;
            piv = am.sv;
            break;
        }
    }
    if (l == m) {
        if (tid == 0) {
            rank_out[blockIdx.x] = 0;
            status_out[blockIdx.x] = 0;
        }
        return;
    }
    int rp = l;
    // ---- first column
    double un2 = 0.0, vn2 = 0.0;
    AMax an{-1.0, INT_MAX, 0.0};
    {
        PointI pc = load_point(kp, cols[cp]);
        for (int t = tid; t < m; t += NT) {
            double c = kentry(kp, pc, rows[t]);  // A(rows[t], cols[cp]) (symmetric entry)
            double u = c / piv;
            U0[t] = u;
            un2 += u * u;
            if (t != rp) amax_update(an, c, t);
        }
        for (int s = tid; s < n; s += NT) vn2 += V0[s] * V0[s];
    }
    block_reduce<NT>(red, un2, vn2, an);
    if (tid == 0) {
        rp_out[0] = rp;
        cp_out[0] = cp;
        ru[rp] = 1;
        cu[cp] = 1;
    }
    int rank = 1;
    int status = 0;
    double normS = 0.0;
    double un = sqrt(un2), vn = sqrt(vn2);
    int next_rp = an.idx;
    if (tid == 0) {
        hist[0] = un * vn;
        hist[1] = 0.0;
    }
    __syncthreads();

    while (un * vn > tol * normS && rank < mn) {
        rp = next_rp;
        if (rp == INT_MAX) break;
        if (!aca_reserve<NT>(tab, rank, m, n, tcap, ws_top, ws_cap, &sh_off)) {
            status = 1;
            break;
        }
        double *Ur = ws + sh_off;
        double *Vr = Ur + m;
        const int nch = (rank + kChunk - 1) / kChunk;
        // ---- residual row -> V column `rank`
        PointI pr = load_point(kp, rows[rp]);
        AMax am{-1.0, INT_MAX, 0.0};
        vn2 = 0.0;
        for (int ch = 0; ch < nch; ++ch) {
            const int l0 = ch * kChunk, l1 = min(rank, l0 + kChunk);
            for (int q = l0 + tid; q < l1; q += NT) {
                coef[q - l0] = ws[tab[2 * q] + rp];
                cofs[q - l0] = tab[2 * q + 1];
            }
            __syncthreads();
            for (int s = tid; s < n; s += NT) {
                double r = ch == 0 ? kentry(kp, pr, cols[s]) : Vr[s];
                for (int q = l0; q < l1; ++q) r -= coef[q - l0] * ws[cofs[q - l0] + s];
                Vr[s] = r;
                if (ch == nch - 1) {
                    vn2 += r * r;
                    if (!cu[s]) amax_update(am, r, s);
                }
            }
            __syncthreads();
        }
        double dummy = 0.0;
        block_reduce<NT>(red, vn2, dummy, am);
        if (am.idx == INT_MAX) break;
        cp = am.idx;
        piv = am.sv;
        // ---- residual column -> U column `rank`
        PointI pc = load_point(kp, cols[cp]);
        an = AMax{-1.0, INT_MAX, 0.0};
        un2 = 0.0;
        for (int ch = 0; ch < nch; ++ch) {
            const int l0 = ch * kChunk, l1 = min(rank, l0 + kChunk);
            for (int q = l0 + tid; q < l1; q += NT) {
                coef[q - l0] = ws[tab[2 * q + 1] + cp];
                cofs[q - l0] = tab[2 * q];
            }
            __syncthreads();
            for (int t = tid; t < m; t += NT) {
                double c = ch == 0 ? kentry(kp, pc, rows[t]) : Ur[t];
                for (int q = l0; q < l1; ++q) c -= coef[q - l0] * ws[cofs[q - l0] + t];
                if (ch == nch - 1) {
                    double u = c / piv;
                    Ur[t] = u;
                    un2 += u * u;
                    if (t != rp && !ru[t]) amax_update(an, c, t);
                } else {
                    Ur[t] = c;
                }
            }
            __syncthreads();
        }
        block_reduce<NT>(red, un2, dummy, an);
        un = sqrt(un2);
        vn = sqrt(vn2);
        if (un < 1e-36 || vn < 1e-36) break;
        if (tid == 0) {
            rp_out[rank] = rp;
            cp_out[rank] = cp;
            ru[rp] = 1;
            cu[cp] = 1;
        }
        ++rank;
        next_rp = an.idx;
        // ---- normS update: dots of the new term with every term
        double acc = 0.0;
        for (int q = wv; q < rank; q += NW) {
            const double *Uq = ws + tab[2 * q];
            const double *Vq = ws + tab[2 * q + 1];
            double du = 0.0, dv = 0.0;
            for (int t = lane; t < m; t += 64) du += Uq[t] * Ur[t];
            for (int s = lane; s < n; s += 64) dv += Vq[s] * Vr[s];
            du = wave_sum(du);
            dv = wave_sum(dv);
            acc += fabs(du * dv);
        }
        if (lane == 0) dacc[wv] = acc;
        __syncthreads();
        double S = 0.0;
        for (int q = 0; q < NW; ++q) S += dacc[q];
        double uv = un * vn;
        normS = sqrt(normS * normS + uv * uv + 2.0 * S);
        if (tid == 0) {
            hist[2 * (rank - 1)] = uv;
            hist[2 * (rank - 1) + 1] = normS;
        }
        __syncthreads();
    }
    if (tid == 0) {
        rank_out[blockIdx.x] = rank;
        status_out[blockIdx.x] = status;
    }
}

// ------------------------------------------------------------------ wide ACA
// The cross of k_aca for a few very large blocks, each spread over many
// workgroups: one pivot step is a sequence of launches over all such blocks --
// residual row (wide), row reduction, residual column (wide), column reduction
// and acceptance, normS dot products (wide), normS update and the next step's
// reservation.  Residual entries subtract the earlier terms in rank order as in
// k_aca, so rows, columns and pivots are the same numbers; only the sums of
// squares and dot products are reduced in another order.  Per-block state lives
// in WideState; a finished block (converged, exhausted, or out of workspace:
// status 1) is skipped by every later launch.
constexpr int kWideCols = 2048;  // row / column entries one workgroup of the wide passes covers
constexpr int kWidePT = kWideCols / 256;  // of them per thread

__device__ __forceinline__ void wide_reserve(WideState &st, const AcaJob &jb, int64_t *tab,
                                             unsigned long long *ws_top, int64_t ws_cap) {
    int64_t off = -1;
    if (st.rank < jb.tcap) {
        unsigned long long o = atomicAdd(ws_top, (unsigned long long)(jb.m + jb.n));
        if ((int64_t)o + jb.m + jb.n <= ws_cap) off = (int64_t)o;
    }
    if (off < 0) {
        st.status = 1;
        st.done = 1;
        return;
    }
    st.off = off;
    tab[2 * st.rank] = off;
    tab[2 * st.rank + 1] = off + jb.m;
}

// first launch: clear the used-row/column flags, reserve step 0, start the first-row search
__global__ __launch_bounds__(256) void k_wide_init(const AcaJob *__restrict__ jobs, WideState *__restrict__ state,
                                                   uint8_t *__restrict__ flags, int64_t *__restrict__ tab_all,
                                                   unsigned long long *ws_top, int64_t ws_cap) {
    const AcaJob jb = jobs[blockIdx.y];
    const int64_t tot = (int64_t)jb.m + jb.n;
    for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < tot; e += (int64_t)gridDim.x * 256) flags[jb.wsF + e] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        WideState st{};
        st.rank = 0;
        st.rp = 0;
        st.l = 0;
        st.next_rp = INT_MAX;
        st.done = (jb.m == 0 || jb.n == 0 || jb.tcap == 0);
        st.kfix = jb.pad;
        if (!st.done) wide_reserve(st, jb, tab_all + jb.tab, ws_top, ws_cap);
        state[blockIdx.y] = st;
    }
}

// residual row rp -> V slot of the current step; partial sum of squares and argmax over unused columns
__global__ __launch_bounds__(256) void k_wide_row(const AcaJob *__restrict__ jobs, const int32_t *__restrict__ idx,
                                                  const double *__restrict__ wsr, double *__restrict__ ws,
                                                  const int64_t *__restrict__ tab_all, const uint8_t *__restrict__ flags,
                                                  const WideState *__restrict__ state, WidePart *__restrict__ part,
                                                  int pstride, KParams kp) {
    const AcaJob jb = jobs[blockIdx.y];
    const WideState st = state[blockIdx.y];
    const int s0 = blockIdx.x * kWideCols;
    if (st.done || s0 >= jb.n) return;
    const int s1 = min(jb.n, s0 + kWideCols);
    const int rank = st.rank, rp = st.rp;
    const int32_t *rows = idx + jb.rows, *cols = idx + jb.cols;
    const int64_t *tab = tab_all + jb.tab;
    const uint8_t *cu = flags + jb.wsF + jb.m;
    double *Vr = ws + st.off + jb.m;
    __shared__ double coef[kChunk];
    __shared__ int64_t cofs[kChunk];
    __shared__ BlockRed<256> red;
    PointI pr = load_point(kp, rows[rp]);
    AMax am{-1.0, INT_MAX, 0.0};
    double vn2 = 0.0;
    // a thread's kWidePT columns s0 + threadIdx.x + 256 i stay in registers over the residual's terms,
    // so every term issues kWidePT independent loads (the arithmetic per column is unchanged)
    double rv[kWidePT];
#pragma unroll
    for (int i = 0; i < kWidePT; ++i) {
        const int s = s0 + threadIdx.x + 256 * i;
        rv[i] = s < s1 ? kentry(kp, pr, cols[s]) : 0.0;
    }
    const int nch = (rank + kChunk - 1) / kChunk;
    for (int ch = 0; ch < nch; ++ch) {
        const int l0 = ch * kChunk, l1 = min(rank, l0 + kChunk);
        for (int q = l0 + threadIdx.x; q < l1; q += 256) {
            coef[q - l0] = wsr[tab[2 * q] + rp];
            cofs[q - l0] = tab[2 * q + 1];
        }
        __syncthreads();
        for (int q = l0; q < l1; ++q) {
            const double c = coef[q - l0];
            const double *v = wsr + cofs[q - l0];
#pragma unroll
            for (int i = 0; i < kWidePT; ++i) {
                const int s = s0 + threadIdx.x + 256 * i;
                if (s < s1) rv[i] -= c * v[s];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < kWidePT; ++i) {
        const int s = s0 + threadIdx.x + 256 * i;
        if (s < s1) {
            Vr[s] = rv[i];
            vn2 += rv[i] * rv[i];
            if (!cu[s]) amax_update(am, rv[i], s);
        }
    }
    double dummy = 0.0;
    block_reduce<256>(red, vn2, dummy, am);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.y * pstride + blockIdx.x] = WidePart{vn2, am.a, am.idx, 0, am.sv};
}

// reduce a wide pass's partials: sum of squares and the first argmax
__device__ __forceinline__ void wide_reduce(const WidePart *p, int np, double &ss, AMax &am) {
    __shared__ BlockRed<256> red;
    ss = 0.0;
    am = AMax{-1.0, INT_MAX, 0.0};
    for (int w = threadIdx.x; w < np; w += 256) {
        ss += p[w].ss;
        AMax o{p[w].a, p[w].idx, p[w].sv};
        amax_merge(am, o);
    }
    double dummy = 0.0;
    block_reduce<256>(red, ss, dummy, am);
}

// row reduction: the pivot column (first step: a row with a nonzero entry, else the next row)
__global__ __launch_bounds__(256) void k_wide_row_reduce(const AcaJob *__restrict__ jobs,
                                                         WideState *__restrict__ state,
                                                         const WidePart *__restrict__ part, int pstride) {
    const AcaJob jb = jobs[blockIdx.x];
    WideState st = state[blockIdx.x];
    if (st.done) return;
    const int np = (jb.n + kWideCols - 1) / kWideCols;
    double vn2;
    AMax am;
    wide_reduce(part + (int64_t)blockIdx.x * pstride, np, vn2, am);
    if (threadIdx.x != 0) return;
    if (st.rank == 0) {
        if (am.idx != INT_MAX && am.a > 1e-36) {
            st.cp = am.idx;
            st.piv = am.sv;
            st.vn2 = vn2;
            st.found = 1;
        } else if (++st.l >= jb.m) {
            st.done = 1;  // every row is zero: rank 0
        } else {
            st.rp = st.l;
        }
    } else if (am.idx == INT_MAX) {
        st.done = 1;  // no unused column left
    } else {
        st.cp = am.idx;
        st.piv = am.sv;
        st.vn2 = vn2;
    }
    state[blockIdx.x] = st;
}

// residual column cp -> U slot (divided by the pivot); partial sum of squares, argmax over unused rows != rp
__global__ __launch_bounds__(256) void k_wide_col(const AcaJob *__restrict__ jobs, const int32_t *__restrict__ idx,
                                                  const double *__restrict__ wsr, double *__restrict__ ws,
                                                  const int64_t *__restrict__ tab_all, const uint8_t *__restrict__ flags,
                                                  const WideState *__restrict__ state, WidePart *__restrict__ part,
                                                  int pstride, KParams kp) {
    const AcaJob jb = jobs[blockIdx.y];
    const WideState st = state[blockIdx.y];
    const int t0 = blockIdx.x * kWideCols;
    if (st.done || t0 >= jb.m || (st.rank == 0 && !st.found)) return;
    const int t1 = min(jb.m, t0 + kWideCols);
    const int rank = st.rank, rp = st.rp, cp = st.cp;
    const double piv = st.piv;
    const int32_t *rows = idx + jb.rows, *cols = idx + jb.cols;
    const int64_t *tab = tab_all + jb.tab;
    const uint8_t *ru = flags + jb.wsF;
    double *Ur = ws + st.off;
    __shared__ double coef[kChunk];
    __shared__ int64_t cofs[kChunk];
    __shared__ BlockRed<256> red;
    PointI pc = load_point(kp, cols[cp]);
    AMax an{-1.0, INT_MAX, 0.0};
    double un2 = 0.0;
    double cv[kWidePT];  // (as in k_wide_row)
#pragma unroll
    for (int i = 0; i < kWidePT; ++i) {
        const int t = t0 + threadIdx.x + 256 * i;
        cv[i] = t < t1 ? kentry(kp, pc, rows[t]) : 0.0;
    }
    const int nch = (rank + kChunk - 1) / kChunk;
    for (int ch = 0; ch < nch; ++ch) {
        const int l0 = ch * kChunk, l1 = min(rank, l0 + kChunk);
        for (int q = l0 + threadIdx.x; q < l1; q += 256) {
            coef[q - l0] = wsr[tab[2 * q + 1] + cp];
            cofs[q - l0] = tab[2 * q];
        }
        __syncthreads();
        for (int q = l0; q < l1; ++q) {
            const double c = coef[q - l0];
            const double *v = wsr + cofs[q - l0];
#pragma unroll
            for (int i = 0; i < kWidePT; ++i) {
                const int t = t0 + threadIdx.x + 256 * i;
                if (t < t1) cv[i] -= c * v[t];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < kWidePT; ++i) {
        const int t = t0 + threadIdx.x + 256 * i;
        if (t < t1) {
            const double u = cv[i] / piv;
            Ur[t] = u;
            un2 += u * u;
            if (t != rp && !ru[t]) amax_update(an, cv[i], t);
        }
    }
    double dummy = 0.0;
    block_reduce<256>(red, un2, dummy, an);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.y * pstride + blockIdx.x] = WidePart{un2, an.a, an.idx, 0, an.sv};
}

// column reduction and acceptance of the step; the first step also checks the loop condition
// and reserves the next step (later steps do that after the normS update)
__global__ __launch_bounds__(256) void k_wide_accept(const AcaJob *__restrict__ jobs, WideState *__restrict__ state,
                                                     const WidePart *__restrict__ part, int pstride,
                                                     uint8_t *__restrict__ flags, int32_t *__restrict__ rowpiv,
                                                     int32_t *__restrict__ colpiv, double *__restrict__ hist_all,
                                                     int64_t *__restrict__ tab_all, unsigned long long *ws_top,
                                                     int64_t ws_cap, double tol) {
    const AcaJob jb = jobs[blockIdx.x];
    WideState st = state[blockIdx.x];
    if (st.done || (st.rank == 0 && !st.found)) return;
    const int np = (jb.m + kWideCols - 1) / kWideCols;
    double un2;
    AMax an;
    wide_reduce(part + (int64_t)blockIdx.x * pstride, np, un2, an);
    if (threadIdx.x != 0) return;
    const double un = sqrt(un2), vn = sqrt(st.vn2);
    uint8_t *ru = flags + jb.wsF, *cu = ru + jb.m;
    double *hist = hist_all + 2 * jb.piv;
    const bool fixed = st.kfix > 1 && st.dots_rank >= 0;  // (dots_rank < 0: the fixed phase was judged)
    if (st.rank > 0 && (un < 1e-36 || vn < 1e-36)) {
        st.done = 1;  // the step is dropped
        if (fixed && st.rank >= 2) st.status = 3;
        state[blockIdx.x] = st;
        return;
    }
    rowpiv[jb.piv + st.rank] = st.rp;
    colpiv[jb.piv + st.rank] = st.cp;
    ru[st.rp] = 1;
    cu[st.cp] = 1;
    st.un = un;
    st.vn = vn;
    st.next_rp = an.idx;
    ++st.rank;
    if (fixed) {
        // fixed phase: no normS update and no stopping test until kfix steps are taken (the stopping
        // rule is applied afterwards to all of them at once, from their Gram products)
        hist[2 * (st.rank - 1)] = un * vn;
        hist[2 * (st.rank - 1) + 1] = 0.0;
        st.dots_rank = st.rank;
        const int mn = min(jb.m, jb.n);
        if (st.rank >= st.kfix || st.rank >= mn || st.next_rp == INT_MAX) {
            st.done = 1;
            st.status = st.rank >= st.kfix && st.rank < mn && st.next_rp != INT_MAX ? 2 : 3;
        } else {
            st.rp = st.next_rp;
            wide_reserve(st, jb, tab_all + jb.tab, ws_top, ws_cap);
        }
        state[blockIdx.x] = st;
        return;
    }
    if (st.rank == 1) {
        // the first term enters without a normS update (normS stays 0)
        hist[0] = un * vn;
        hist[1] = 0.0;
        st.normS = 0.0;
        const int mn = min(jb.m, jb.n);
        if (!(un * vn > tol * st.normS && st.rank < mn) || st.next_rp == INT_MAX) st.done = 1;
        else {
            st.rp = st.next_rp;
            wide_reserve(st, jb, tab_all + jb.tab, ws_top, ws_cap);
        }
    }
    state[blockIdx.x] = st;
}

// |<U_q, u><V_q, v>| of the newest term (u, v) with every term q <= newest, one workgroup per q
__global__ __launch_bounds__(256) void k_wide_dots(const AcaJob *__restrict__ jobs, const double *__restrict__ ws,
                                                   const int64_t *__restrict__ tab_all,
                                                   const WideState *__restrict__ state, double *__restrict__ dots,
                                                   int dstride) {
    const AcaJob jb = jobs[blockIdx.y];
    const WideState st = state[blockIdx.y];
    const int q = blockIdx.x;
    if (st.done || st.rank < 2 || q >= st.rank || st.dots_rank == st.rank) return;
    const int64_t *tab = tab_all + jb.tab;
    const double *Uq = ws + tab[2 * q], *Vq = ws + tab[2 * q + 1];
    const double *Uk = ws + tab[2 * (st.rank - 1)], *Vk = ws + tab[2 * (st.rank - 1) + 1];
    // four partial sums per thread: four independent load pairs in flight
    auto dot = [](const double *a, const double *b, int len) {
        double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
        int t = threadIdx.x;
        for (; t + 768 < len; t += 1024) {
            p0 += a[t] * b[t];
            p1 += a[t + 256] * b[t + 256];
            p2 += a[t + 512] * b[t + 512];
            p3 += a[t + 768] * b[t + 768];
        }
        for (; t < len; t += 256) p0 += a[t] * b[t];
        return (p0 + p1) + (p2 + p3);
    };
    double du = dot(Uq, Uk, jb.m), dv = dot(Vq, Vk, jb.n);
    __shared__ BlockRed<256> red;
    AMax dummy{-1.0, INT_MAX, 0.0};
    block_reduce<256>(red, du, dv, dummy);
    if (threadIdx.x == 0) dots[(int64_t)blockIdx.y * dstride + q] = fabs(du * dv);
}

// normS update of the newest term, loop condition, next step's reservation
__global__ __launch_bounds__(256) void k_wide_norm(const AcaJob *__restrict__ jobs, WideState *__restrict__ state,
                                                   const double *__restrict__ dots, int dstride,
                                                   double *__restrict__ hist_all, int64_t *__restrict__ tab_all,
                                                   unsigned long long *ws_top, int64_t ws_cap, double tol) {
    const AcaJob jb = jobs[blockIdx.x];
    WideState st = state[blockIdx.x];
    if (st.done || st.rank < 2 || st.dots_rank == st.rank) return;
    double S = 0.0, dummy = 0.0;
    for (int q = threadIdx.x; q < st.rank; q += 256) S += dots[(int64_t)blockIdx.x * dstride + q];
    __shared__ BlockRed<256> red;
    AMax am{-1.0, INT_MAX, 0.0};
    block_reduce<256>(red, S, dummy, am);
    if (threadIdx.x != 0) return;
    const double uv = st.un * st.vn;
    st.normS = sqrt(st.normS * st.normS + uv * uv + 2.0 * S);
    double *hist = hist_all + 2 * jb.piv;
    hist[2 * (st.rank - 1)] = uv;
    hist[2 * (st.rank - 1) + 1] = st.normS;
    st.dots_rank = st.rank;
    const int mn = min(jb.m, jb.n);
    if (!(uv > tol * st.normS && st.rank < mn) || st.next_rp == INT_MAX) st.done = 1;
    else {
        st.rp = st.next_rp;
        wide_reserve(st, jb, tab_all + jb.tab, ws_top, ws_cap);
    }
    state[blockIdx.x] = st;
}

__global__ void k_wide_finish(const WideState *__restrict__ state, int nb, int32_t *__restrict__ rank_out,
                              int32_t *__restrict__ status_out, int32_t *__restrict__ ndone) {
    int done = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        rank_out[b] = state[b].rank;
        status_out[b] = state[b].status;
        done += state[b].done;
    }
    done = (int)wave_sum((double)done);
    if (threadIdx.x == 0) *ndone = done;
}

// Gram products of the terms of a parked block: one workgroup per 64 x 64 tile of term pairs
// (q in tile qi, k in tile ki, qi <= ki); each thread accumulates 4 x 4 dot products of U
// columns, then of V columns, over 32-entry slabs staged in LDS.  prod = |<U_q,U_k>||<V_q,V_k>|
// for q <= k (packed lower-triangular by k).
constexpr int kGramT = 64, kGramS = 32;
__device__ __forceinline__ void gram_tile(const double *ws, const int64_t *tab, int part, int len, int q0, int nq,
                                          int k0, int nk, double (&acc)[4][4]) {
    __shared__ double sq[kGramS][kGramT + 1], sk[kGramS][kGramT + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // q = q0 + 4 tx + a, k = k0 + 4 ty + b
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int t0 = 0; t0 < len; t0 += kGramS) {
        __syncthreads();
        for (int e = threadIdx.x; e < kGramS * kGramT; e += 256) {
            const int c = e / kGramS, t = e % kGramS;  // column c of the tile, entry t of the slab
            const bool tin = t0 + t < len;
            sq[t][c] = (c < nq && tin) ? ws[tab[2 * (q0 + c) + part] + t0 + t] : 0.0;
            sk[t][c] = (c < nk && tin) ? ws[tab[2 * (k0 + c) + part] + t0 + t] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int t = 0; t < kGramS; ++t) {
            double x[4], y[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) x[a] = sq[t][4 * tx + a];
#pragma unroll
            for (int b = 0; b < 4; ++b) y[b] = sk[t][4 * ty + b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = fma(x[a], y[b], acc[a][b]);
        }
    }
}
__global__ __launch_bounds__(256) void k_wide_gram(const GramJob *__restrict__ gj, const AcaJob *__restrict__ jobs,
                                                   const WideState *__restrict__ state, const double *__restrict__ ws,
                                                   const int64_t *__restrict__ tab_all, double *__restrict__ prod) {
    const GramJob g = gj[blockIdx.x];
    const AcaJob jb = jobs[g.blk];
    const int K = state[g.blk].rank;
    const int64_t *tab = tab_all + jb.tab;
    const int q0 = g.qi * kGramT, k0 = g.ki * kGramT;
    const int nq = min(kGramT, K - q0), nk = min(kGramT, K - k0);
    double gu[4][4], gv[4][4];
    gram_tile(ws, tab, 0, jb.m, q0, nq, k0, nk, gu);
    gram_tile(ws, tab, 1, jb.n, q0, nq, k0, nk, gv);
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int q = q0 + 4 * tx + a, k = k0 + 4 * ty + b;
            if (q < K && k < K && q <= k) prod[g.off + (int64_t)k * (k + 1) / 2 + q] = fabs(gu[a][b]) * fabs(gv[a][b]);
        }
}

// decisions of the host on the judged blocks (one thread per block)
__global__ void k_wide_judge(const AcaJob *__restrict__ jobs, int nb, WideState *__restrict__ state,
                             const int32_t *__restrict__ final_rank, const double *__restrict__ normS_in,
                             int64_t *__restrict__ tab_all, unsigned long long *ws_top, int64_t ws_cap) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        const int fr = final_rank[b];
        if (fr == -2) continue;
        WideState st = state[b];
        st.dots_rank = -1;  // (the fixed phase is over)
        if (fr >= 0) {
            st.rank = fr;
            st.done = 1;
            st.status = 0;
        } else {
            // the stopping rule did not hold within the fixed phase: continue adaptively
            st.normS = normS_in[b];
            st.dots_rank = st.rank;
            st.done = 0;
            st.status = 0;
            st.kfix = 0;
            st.rp = st.next_rp;
            wide_reserve(st, jobs[b], tab_all + jobs[b].tab, ws_top, ws_cap);
        }
        state[b] = st;
    }
}

// sampled entries of finished blocks against their cross approximation
__global__ __launch_bounds__(64) void k_aca_sample(const AcaJob *__restrict__ jobs, const int32_t *__restrict__ rank,
                                                   const int32_t *__restrict__ idx, const double *__restrict__ ws,
                                                   const int64_t *__restrict__ tab_all, KParams kp, int ns,
                                                   double *__restrict__ err2, double *__restrict__ ref2) {
    const AcaJob jb = jobs[blockIdx.x];
    const int r = rank[blockIdx.x];
    const int64_t *tab = tab_all + jb.tab;
    double e2 = 0.0, a2 = 0.0;
    if (jb.m > 0 && jb.n > 0)
        for (int q = threadIdx.x; q < ns; q += 64) {
            // (i, j) from a 64-bit mix of (block, sample)
            uint64_t h = ((uint64_t)blockIdx.x << 32) ^ (uint64_t)q ^ 0x9E3779B97F4A7C15ull;
            h ^= h >> 33;
            h *= 0xff51afd7ed558ccdull;
            h ^= h >> 33;
            h *= 0xc4ceb9fe1a85ec53ull;
            h ^= h >> 33;
            const int i = (int)(h % (uint64_t)jb.m), j = (int)((h >> 32) % (uint64_t)jb.n);
            const double a = kentry2(kp, idx[jb.rows + i], idx[jb.cols + j]);
            double v = 0.0;
            for (int k = 0; k < r; ++k) v = fma(ws[tab[2 * k] + i], ws[tab[2 * k + 1] + j], v);
            e2 += (a - v) * (a - v);
            a2 += a * a;
        }
    e2 = wave_sum(e2);
    a2 = wave_sum(a2);
    if (threadIdx.x == 0) {
        err2[blockIdx.x] = e2;
        ref2[blockIdx.x] = a2;
    }
}

// L (unit lower) and R (upper) of the cross: L(i,j) = U_j[rb_i], R(j,i) = V_j[cb_i]
// (2D/codes/aca.hpp:419-445)
__global__ void k_lr_extract(const LrJob *__restrict__ jobs, const double *__restrict__ ws,
                             const int64_t *__restrict__ tab_all, const int32_t *__restrict__ rowpiv,
                             const int32_t *__restrict__ colpiv, double *__restrict__ out) {
    const LrJob jb = jobs[blockIdx.x];
    const int r = jb.rank;
    double *L = out + jb.out;
    double *R = L + (int64_t)r * r;
    const int64_t *tab = tab_all + jb.tab;
    for (int e = threadIdx.x; e < r * r; e += blockDim.x) {
        int i = e % r, j = e / r;  // row i, column j
        double lv = 0.0, rv = 0.0;
        if (i == j) lv = 1.0;
        else if (j < i) lv = ws[tab[2 * j] + rowpiv[jb.piv + i]];
        if (i <= j) rv = ws[tab[2 * i + 1] + colpiv[jb.piv + j]];
        L[e] = lv;
        R[e] = rv;
    }
}

// A = K(rows, cols), column-major (getMatrix, 3D/codes/kernel.hpp:58-71)
__global__ void k_eval(const EvalJob *__restrict__ jobs, const int32_t *__restrict__ idx, double *__restrict__ A,
                       KParams kp) {
    const EvalJob jb = jobs[blockIdx.x];
    const int64_t tot = (int64_t)jb.m * jb.n;
    const int32_t *rows = idx + jb.rows;
    const int32_t *cols = idx + jb.cols;
    for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
        int i = (int)(e % jb.m), j = (int)(e / jb.m);
        A[jb.a + (int64_t)j * jb.lda + i] = kentry2(kp, rows[i], cols[j]);
    }
}

// E = K(S, Ic) R^{-1} L^{-1} (m x r) and its transpose.  For the far field this
// is L2P (3D/codes/HODLR3Dtree.hpp:1926-1927); for the vertex-sharing field it
// is the transpose of M2M_ver = (L R)^{-T} K(Ic, S) (:2110-2133), the same
// matrix since the kernel is symmetric.  One thread per row of S.
__global__ void k_xfer(const XferJob *__restrict__ jobs, const int32_t *__restrict__ idx,
                       const double *__restrict__ lr, double *__restrict__ A, KParams kp) {
    const XferJob jb = jobs[blockIdx.x];
    const int m = jb.m, r = jb.r;
    const int64_t lde = jb.lde, ldet = jb.ldet;
    const double *L = lr + jb.lr;
    const double *R = L + (int64_t)r * r;
    double *E = A + jb.e;
    double *ET = A + jb.et;
    const int32_t *S = idx + jb.s;
    const int32_t *Ic = idx + jb.ic;
    for (int t = threadIdx.x; t < m; t += blockDim.x) {
        PointI p = load_point(kp, S[t]);
        // X R = K(S_t, Ic)
        for (int k = 0; k < r; ++k) {
            double s = kentry(kp, p, Ic[k]);
            for (int q = 0; q < k; ++q) s -= E[q * lde + t] * R[(int64_t)k * r + q];
            E[k * lde + t] = s / R[(int64_t)k * r + k];
        }
        // Y L = X
        for (int k = r - 1; k >= 0; --k) {
            double s = E[k * lde + t];
            for (int q = k + 1; q < r; ++q) s -= E[q * lde + t] * L[(int64_t)k * r + q];
            E[k * lde + t] = s / L[(int64_t)k * r + k];
        }
        for (int k = 0; k < r; ++k) ET[t * ldet + k] = E[k * lde + t];
    }
}

// A = R^{-1} L^{-1} K(Ir, V) (r x n), one thread per column: the semi-nested
// vertex block's row factor with the two triangular solves of the reference's
// apply (3D/codes/HODLR3Dtree.hpp:2341-2344) folded in at assembly.
__global__ void k_cur(const CurJob *__restrict__ jobs, const int32_t *__restrict__ idx,
                      const double *__restrict__ lr, double *__restrict__ A, KParams kp) {
    const CurJob jb = jobs[blockIdx.x];
    const int r = jb.r, n = jb.n;
    const double *L = lr + jb.lr;
    const double *R = L + (int64_t)r * r;
    const int32_t *Ir = idx + jb.ir;
    const int32_t *Vv = idx + jb.v;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        double *a = A + jb.a + (int64_t)t * jb.lda;
        PointI p = load_point(kp, Vv[t]);
        for (int k = 0; k < r; ++k) a[k] = kentry(kp, p, Ir[k]);
        for (int k = 0; k < r; ++k) {
            double s = a[k];
            for (int q = 0; q < k; ++q) s -= L[(int64_t)q * r + k] * a[q];
            a[k] = s / L[(int64_t)k * r + k];
        }
        for (int k = r - 1; k >= 0; --k) {
            double s = a[k];
            for (int q = k + 1; q < r; ++q) s -= R[(int64_t)q * r + k] * a[q];
            a[k] = s / R[(int64_t)k * r + k];
        }
    }
}

// ------------------------------------------------------------------ batched skinny GEMV
// One workgroup (4 waves) per item of <= 128 rows.  Each lane loads VEC
// consecutive rows of a column with one 16-byte load (VEC = 2 for FP64, 4 for
// FP32); ceil(rows / VEC) lanes, rounded up to a power of two, cover a column,
// so a wave load reads 64/lpc whole columns.  The waves split the columns of
// every term.  Partial sums are combined across the wave's column groups with
// cross-lane adds, then across the waves in LDS, and added to y exclusively
// or atomically.  Blocks are column-major with leading dimensions padded to a
// multiple of 4, so every VEC group is aligned and stays inside the column.
// ROLE 1 (the fused M2L + near-field launch) also forms the transposed products
// of the symmetric terms; 0 (transfer passes) and 2 (semi-nested column factors)
// only name the instance so that profiles separate the launches.  UNR is the
// number of columns each lane loads per iteration.  Bit 1 of an item's flags
// selects the operator pool A1 (the inner operator of the refinement) instead
// of A0.
// Streaming load through a global-address-space pointer: operator pointers taken from a
// table in memory are generic to the compiler, whose flat loads would also count against
// the LDS wait counter.
template <typename V>
__device__ __forceinline__ V ld_nt(const V *p) {
    return __builtin_nontemporal_load((const __attribute__((address_space(1))) V *)p);
}

// Value of lane (lane ^ O) for a power of two O < 64: quad permutes (DPP) for 1 and 2,
// ds_swizzle in its bit-mask mode for 4 .. 16, ds_bpermute for 32.  Every lane of the
// aligned group of 2 O lanes must be active.
template <int O>
__device__ __forceinline__ int xor_lane(int v) {
    if constexpr (O == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);
    else if constexpr (O == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);
    else if constexpr (O < 32) return __builtin_amdgcn_ds_swizzle(v, 0x1F | (O << 10));
    else return __builtin_amdgcn_ds_bpermute((int)(((threadIdx.x & 63) ^ O) << 2), v);
}
template <int O>
__device__ __forceinline__ float xor_lane(float v) {
    return __int_as_float(xor_lane<O>(__float_as_int(v)));
}
template <int O>
__device__ __forceinline__ double xor_lane(double v) {
    const int64_t b = __double_as_longlong(v);
    const int lo = xor_lane<O>((int)(uint32_t)b), hi = xor_lane<O>((int)(b >> 32));
    return __longlong_as_double((int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}
// the same with O known only after unrolling
template <typename T>
__device__ __forceinline__ T xor_lane_n(T v, int o) {
    switch (o) {
    case 1: return xor_lane<1>(v);
    case 2: return xor_lane<2>(v);
    case 4: return xor_lane<4>(v);
    case 8: return xor_lane<8>(v);
    case 16: return xor_lane<16>(v);
    default: return xor_lane<32>(v);
    }
}

// Raw buffer access (gfx9 resource word 3): a lane whose offset lies at or past the
// resource's size reads zeros and writes nothing, without a branch.
constexpr int kBufWord3 = 0x00020000;
constexpr uint32_t kOutOfRange = 0xFFFFFFF0u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, kBufWord3);
}
// An offset the compiler cannot see through: keeps a load whose offset is selected per
// lane a single unconditional load (instead of one load per alternative under branches).
// The asm is empty and loads nothing.
__device__ __forceinline__ uint32_t opaque_off(uint32_t v) {
    asm volatile("" : "+v"(v));
    return v;
}
template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <typename V>
__device__ __forceinline__ V buf_ld16_nt(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 2));
}
// one 4-, 8- or 16-byte vector, streamed (not kept in the caches)
template <typename V>
__device__ __forceinline__ V buf_ldv_nt(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (sizeof(V) == 16) return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 2));
    else if constexpr (sizeof(V) == 8) return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 2));
    else return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 2));
}
// N consecutive elements (N sizeof(T) = 8, 16, 32 or 64 bytes) in 8- or 16-byte loads
template <typename T, int N>
__device__ __forceinline__ void buf_ld_n(T (&out)[N], __amdgpu_buffer_rsrc_t r, uint32_t off) {
    constexpr int B = N * (int)sizeof(T);
    if constexpr (B == 8) {
        typedef unsigned int U2 __attribute__((ext_vector_type(2)));
        typedef T W __attribute__((ext_vector_type(N)));
        const W w = __builtin_bit_cast(W, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
#pragma unroll
        for (int k = 0; k < N; ++k) out[k] = w[k];
    } else {
        constexpr int E = 16 / sizeof(T);
        typedef T W __attribute__((ext_vector_type(E)));
#pragma unroll
        for (int p = 0; p < B / 16; ++p) {
            const W w = __builtin_bit_cast(W, __builtin_amdgcn_raw_buffer_load_b128(r, off + 16 * p, 0, 0));
#pragma unroll
            for (int k = 0; k < E; ++k) out[E * p + k] = w[k];
        }
    }
}

// Work lists are never written by a kernel: read through the constant address space, their
// wave-uniform loads stay scalar after the kernel's own stores (a vector load there would
// wait for every load in flight).
typedef const __attribute__((address_space(4))) GTerm CGTerm;
__device__ __forceinline__ GTerm term_at(const CGTerm *p, int k) {
    GTerm t;
    t.a = p[k].a;
    t.x = p[k].x;
    t.t = p[k].t;
    t.cols = p[k].cols;
    t.ld = p[k].ld;
    return t;
}

// The column loop of an item walks (term, column) iterations with a wave-uniform
// cursor; each iteration is UNR column loads per lane.  Two iterations are in flight:
// the loads of the next are issued before the current one is multiplied.  Every load
// is unconditional (lanes out of range address past the buffer's end), so the
// compiler's waits count exactly the outstanding operations.
template <typename T, int VEC, int ROLE, int UNR>
__global__ __launch_bounds__(256) void k_bgemv(const GItem *__restrict__ items, const GTerm *__restrict__ terms,
                                               const T *__restrict__ A0, const T *__restrict__ A1, T *vec) {
    typedef T V __attribute__((ext_vector_type(VEC)));
    const GItem it = items[blockIdx.x];
    const CGTerm *const tms = (const CGTerm *)terms;
    const T *__restrict__ A = (it.flags & 2) ? A1 : A0;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int rows = it.rows;
    // (the fused launch gives a column at least UNR lanes, so that its transposed sums
    // end on UNR distinct lanes; the extra lanes of a short tile are inactive)
    int lpc = ROLE == 1 ? UNR : 1;
    while (lpc * VEC < rows) lpc <<= 1;
    const int cpw = 64 / lpc;
    const int q = lane / lpc, r0 = (lane % lpc) * VEC;
    const bool act = r0 < rows;
    __shared__ T part[4][kRowTile];
    V acc = (V)0;
    // this tile's own input rows: the x of the transposed products of symmetric terms
    T xb[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) xb[v] = (ROLE == 1 && act && r0 + v < rows) ? vec[it.xb + r0 + v] : T(0);
    const int64_t lda = it.lda;
    // a wave takes chunks of cpw UNR consecutive columns, lane group q the UNR columns
    // c0 + q UNR + u (one vector load of their x); the four waves take every fourth chunk
    const int chunk = cpw * UNR;
    const uint32_t roff = (uint32_t)((it.row0 + (act ? r0 : 0)) * sizeof(T));  // this lane's rows in a column
    struct Iter {
        V av[UNR];
        T xv[UNR];
        int cols, cb;
        int64_t t;
    };
    int tt = it.t0, c0 = wv * chunk;  // next iteration: term, first column of the wave's chunk
    auto settle = [&]() {
        while (tt < it.t1 && c0 >= tms[tt].cols) {
            ++tt;
            c0 = wv * chunk;
        }
    };
    // loads of the next iteration (past the last one: a term's resources with no valid lane)
    auto load = [&](Iter &B) -> bool {
        const bool has = tt < it.t1;
        const GTerm tm = term_at(tms, has ? tt : it.t1 - 1);
        B.cols = has ? tm.cols : 0;
        B.cb = c0 + q * UNR;
        B.t = tm.t;
        const __amdgpu_buffer_rsrc_t ra = buf_rsrc(A + tm.a, (uint32_t)((int64_t)tm.cols * lda * sizeof(T)));
        const __amdgpu_buffer_rsrc_t rx = buf_rsrc(vec + tm.x, (uint32_t)(B.cols * sizeof(T)));
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int cc = B.cb + u;
            const bool ok = act && cc < B.cols;
            B.av[u] = buf_ld16_nt<V>(ra, ok ? (uint32_t)(cc * lda * sizeof(T)) + roff : kOutOfRange);
        }
        // (columns past the term's end within the group read neighbouring values or zeros:
        // their A entries are zero and they are dropped below)
        buf_ld_n<T, UNR>(B.xv, rx, act && B.cb < B.cols ? (uint32_t)(B.cb * sizeof(T)) : kOutOfRange);
        if (has) {
            c0 += 4 * chunk;
            settle();
        }
        return has;
    };
    auto compute = [&](const Iter &B) {
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const T xu = B.cb + u < B.cols ? B.xv[u] : T(0);
#pragma unroll
            for (int v = 0; v < VEC; ++v) acc[v] = fma(B.av[u][v], xu, acc[v]);
        }
        if (ROLE == 1) {
            // columns cb + u of the block dotted with this tile's rows of x, summed over the
            // lpc lanes that hold them, and added to the partner box's output y_p (offset t)
            // with atomics.  A butterfly first halves the UNR values per exchange (lane bit o
            // keeps one half, sends the other), so that after log2 UNR exchanges lane l holds
            // column ui(l) summed over its UNR-lane group; plain exchanges then finish the
            // sums, and the first UNR lanes of every column group add them (non-symmetric
            // terms, t < 0, add nothing).
            // (a vector, not an array: the selects below then pick values, where selects
            // between array elements become a dynamically indexed array)
            typedef T VU __attribute__((ext_vector_type(UNR)));
            VU tp;
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                tp[u] = T(0);
#pragma unroll
                for (int v = 0; v < VEC; ++v) tp[u] = fma(B.av[u][v], xb[v], tp[u]);
            }
            int ui = 0;
#pragma unroll
            for (int k = UNR, o = 1; k > 1; k >>= 1, o <<= 1) {
                const bool hi = lane & o;
#pragma unroll
                for (int i = 0; i < k / 2; ++i) {
                    const T snd = hi ? tp[i] : tp[i + k / 2];
                    const T keep = hi ? tp[i + k / 2] : tp[i];
                    tp[i] = keep + xor_lane_n(snd, o);
                }
                ui += hi ? k / 2 : 0;
            }
#pragma unroll
            for (int o = UNR; o < 64; o <<= 1)
                if (o < lpc) tp[0] += xor_lane_n(tp[0], o);
            const int cc = B.cb + ui;
            if ((lane % lpc) < UNR && cc < B.cols && B.t >= 0) atomicAdd(vec + B.t + cc, tp[0]);
        }
    };
    // (an iteration loaded past the last one reads nothing and adds nothing: both halves
    // run unconditionally, so the loop has a single path for the waits)
    Iter B0, B1;
    settle();
    bool h0 = load(B0);
    load(B1);
    while (h0) {
        compute(B0);
        h0 = load(B0);
        compute(B1);
        load(B1);
    }
#pragma unroll
    for (int v = 0; v < VEC; ++v)
        for (int o = lpc; o < 64; o <<= 1) acc[v] += __shfl_xor(acc[v], o, 64);
    if (q == 0 && act)
#pragma unroll
        for (int v = 0; v < VEC; ++v)
            if (r0 + v < kRowTile) part[threadIdx.x >> 6][r0 + v] = acc[v];
    __syncthreads();
    if (threadIdx.x < rows) {
        const int t = threadIdx.x;
        T s = part[0][t] + part[1][t] + part[2][t] + part[3][t];
        T *y = vec + it.y + t;
        if (it.flags & 5) atomicAdd(y, s);  // split items; fused items (transposes land in y too)
        else *y += s;
    }
}

// ------------------------------------------------------------------ batched skinny GEMM, kNR right-hand sides
// The same work items as k_bgemv, on a vector pool holding kNR right-hand
// sides per slot (element (slot, rhs) at slot * kNR + rhs): every item is a
// small dense product Y[rows x kNR] += A[rows x cols] X[cols x kNR] on the
// matrix cores (16x16x4 MFMA, FP32 or FP64).
//
// Lane l (m = l % 16, q = l / 16) reads VEC consecutive rows of column c0 + q
// with one 16-byte load: rows VEC m .. VEC m + VEC-1 of a 16 VEC-row group (a
// quarter-wave reads 256 contiguous bytes of a column).  Component u of that
// vector is the MFMA's A(m, q) for the strided row set {VEC m + u}, so VEC
// MFMAs consume one load and X(c0 + q, m) -- kNR contiguous elements per
// column -- is the B operand of all of them.  An MFMA's D(row(q, v), m) then
// belongs to item row  group * 16 VEC + VEC row(q, v) + u, where row(q, v) =
// 4q + v (FP32) or q + 4v (FP64) are the two instructions' result layouts,
// checked on gfx950.  An item's columns, term after term, form 4-column chunks;
// the four waves take every fourth chunk (a scalar cursor walks the terms), four
// chunks in flight per wave whatever the terms' widths, and their partial
// products are summed through LDS by wave 0.
template <typename T>
struct Mfma16;
template <>
struct Mfma16<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int q, int v) { return 4 * q + v; }
};
template <>
struct Mfma16<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int q, int v) { return q + 4 * v; }
};

// One item with VEC rows per lane load, NG load groups (rows <= 16 VEC NG) and UNR
// 4-column chunks per iteration, two iterations in flight (the next one's loads are
// issued before the current one is multiplied).  Loads are raw buffer loads whose
// out-of-range lanes read zeros, so no branch guards them and the waits before the
// MFMAs count only the current iteration's loads.  WPI waves share the item: 4 (the
// waves take every fourth chunk and sum their products through red, the workgroup's
// LDS reduction buffer) or 1 (the wave takes every chunk and writes its rows itself;
// red is unused and no barrier is reached).
template <typename T, int VEC, int NG, int UNR, int WPI>
__device__ __forceinline__ void bgemm_item(const GItem &it, const GTerm *__restrict__ terms, const T *__restrict__ A,
                                           T *vec, T (*red)[kRowTile][kNR]) {
    typedef typename Mfma16<T>::v4 V4;
    typedef T VL __attribute__((ext_vector_type(VEC)));
    constexpr int RG = 16 * VEC;  // rows of one load group
    constexpr int NA = NG * VEC;  // accumulators
    const CGTerm *const tms = (const CGTerm *)terms;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int rows = it.rows;
    const int ng = (rows + RG - 1) / RG;
    const int m = lane & 15, q = lane >> 4;
    V4 acc[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a] = (V4)0;
    const int64_t lda = it.lda;
    // The item's columns, term after term, in 4-column chunks; wave wv takes chunks
    // wv, wv + 4, ... through a cursor (term tt, chunk ck of that term).
    int tt = it.t0, ck = WPI == 4 ? wv : 0;
    auto settle = [&]() {
        while (tt < it.t1) {
            const int nk = (tms[tt].cols + 3) >> 2;
            if (ck < nk) break;
            ck -= nk;
            ++tt;
        }
    };
    struct Iter {
        VL av[UNR][NG];
        T bx[UNR];
        bool has;
    };
    // this lane's rows of every load group (out of range: past the end of any resource)
    uint32_t roff[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g)
        roff[g] = (g < ng && RG * g + VEC * m < rows) ? (uint32_t)((it.row0 + VEC * m + RG * g) * sizeof(T)) : kOutOfRange;
    auto load = [&](Iter &B) -> bool {
        B.has = tt < it.t1;
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const bool has = tt < it.t1;
            const GTerm tm = term_at(tms, has ? tt : it.t1 - 1);
            const int c = 4 * ck + q;
            const bool ok = has && c < tm.cols;
            const __amdgpu_buffer_rsrc_t ra = buf_rsrc(A + tm.a, (uint32_t)((int64_t)tm.cols * lda * sizeof(T)));
            const __amdgpu_buffer_rsrc_t rx = buf_rsrc(vec + tm.x * kNR, (uint32_t)((int64_t)tm.cols * kNR * sizeof(T)));
            B.bx[u] = buf_ld<T>(rx, opaque_off(ok ? (uint32_t)((c * kNR + m) * sizeof(T)) : kOutOfRange));
            const uint32_t coff = ok ? (uint32_t)(c * lda * sizeof(T)) : kOutOfRange;
#pragma unroll
            for (int g = 0; g < NG; ++g)
                B.av[u][g] =
                    buf_ld16_nt<VL>(ra, opaque_off(ok && roff[g] != kOutOfRange ? coff + roff[g] : kOutOfRange));
            if (has) {
                ck += WPI;
                settle();
            }
        }
        return B.has;
    };
    auto compute = [&](const Iter &B) {
#pragma unroll
        for (int u = 0; u < UNR; ++u)
#pragma unroll
            for (int g = 0; g < NG; ++g)
                if (g < ng)
#pragma unroll
                    for (int e = 0; e < VEC; ++e)
                        acc[g * VEC + e] = Mfma16<T>::op(B.av[u][g][e], B.bx[u], acc[g * VEC + e]);
    };
    settle();
    // (a buffer loaded past the last chunk holds zeros: multiplying it adds nothing, so
    // both halves run unconditionally and the loop has a single path for the waits)
    Iter B0, B1;
    bool h0 = load(B0);
    load(B1);
    while (h0) {
        compute(B0);
        h0 = load(B0);
        compute(B1);
        load(B1);
    }
    // item row held by accumulator (g, e), result element v of this lane
    auto row_of = [&](int g, int e, int v) { return RG * g + VEC * Mfma16<T>::row(q, v) + e; };
    // sum the four waves' partial products: 3 and 2 into 1 and 0, then 1 into 0
    auto put = [&](int slot) {
#pragma unroll
        for (int g = 0; g < NG; ++g)
            if (g < ng)
#pragma unroll
                for (int e = 0; e < VEC; ++e)
#pragma unroll
                    for (int v = 0; v < 4; ++v) red[slot][row_of(g, e, v)][m] = acc[g * VEC + e][v];
    };
    auto add = [&](int slot) {
#pragma unroll
        for (int g = 0; g < NG; ++g)
            if (g < ng)
#pragma unroll
                for (int e = 0; e < VEC; ++e)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[g * VEC + e][v] += red[slot][row_of(g, e, v)][m];
    };
    if constexpr (WPI == 4) {
        if (wv >= 2) put(wv - 2);
        __syncthreads();
        if (wv < 2) add(wv);
        __syncthreads();
        if (wv == 1) put(0);
        __syncthreads();
        if (wv == 0) add(0);
    }
    if (WPI == 1 || wv == 0) {
        T *y = vec + it.y * kNR + m;
#pragma unroll
        for (int g = 0; g < NG; ++g)
            if (g < ng)
#pragma unroll
                for (int e = 0; e < VEC; ++e)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = row_of(g, e, v);
                        if (r < rows) {
                            if (it.flags & 1) atomicAdd(y + (int64_t)r * kNR, acc[g * VEC + e][v]);
                            else y[(int64_t)r * kNR] += acc[g * VEC + e][v];
                        }
                    }
    }
}

// WPI 4: one item per workgroup; WPI 1: one item per wave (launches of many small items:
// no cross-wave reduction, four items in flight per workgroup), nitems bounds the last one.
template <typename T, int ROLE, int WPI>
__global__ __launch_bounds__(256) void k_bgemm(const GItem *__restrict__ items, int nitems,
                                               const GTerm *__restrict__ terms, const T *__restrict__ A0,
                                               const T *__restrict__ A1, const T *const *__restrict__ tab, T *vec) {
    const int idx = WPI == 4 ? (int)blockIdx.x : 4 * (int)blockIdx.x + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (idx >= nitems) return;  // (WPI 1 only: whole waves, and no barrier follows)
    const GItem it = items[idx];
    // batched systems: pool pair of system flags >> 8 from the table
    const T *__restrict__ A = tab ? tab[2 * (it.flags >> 8) + ((it.flags >> 1) & 1)] : ((it.flags & 2) ? A1 : A0);
    constexpr int VEC = 16 / sizeof(T);
    if constexpr (WPI == 4) {
        __shared__ T red[2][kRowTile][kNR];
        bgemm_item<T, VEC, kRowTile / (16 * VEC), 2, 4>(it, terms, A, vec, red);
    } else {
        bgemm_item<T, VEC, kRowTile / (16 * VEC), 2, 1>(it, terms, A, vec, nullptr);
    }
}

// Direct and transposed products of the stored blocks in one read, kNR right-hand sides
// (the fused M2L + near-field launch of the multi-vector pass).  One item (box b, rows <= 128)
// per workgroup.  The item's terms K(b, p) are cut into 16-column groups (never across a term);
// wave wv takes groups wv, wv + 4, ... and walks each group's row groups of RG = 16 VEC rows as
// slabs, the next slab in flight while one is multiplied.  A slab is loaded in the direct
// product's layout (lane (m, q): VEC consecutive rows of column 4 j + q, j < 4: 16-byte loads,
// 256-byte runs per column) and multiplied at once into the item's rows (as k_bgemm: A(m, q) =
// K(VEC m + e, 4 j + q),
// B(q, m) = X_p(4 j + q, m)).  For a symmetric term (t >= 0) the wave also writes the slab to one
// of its two LDS buffers; one slab later it reads it back with a column per lane % 16 and a row
// per lane / 16, the A operand of K^T: with B(q, m) = X_b(4 s + q, m) from the item's input rows
// (staged in LDS once) it accumulates the 16 x kNR block Y_p(c0 .. c0 + 15, :) += K(rows, c0 ..)^T
// X_b over the row groups, and adds it to the partner's output with atomics when the group ends.
// (Reading a slab one step after writing it overlaps the LDS round trip with the next slab's
// direct products.)  The item's own rows are added atomically too (other items' transposes land
// in them).  Each stored element is read once per pass, where k_bgemm + k_bgemm_tl read it twice.
// FP64: 64 MFMA flops per stored element (direct and transposed, 16 right-hand sides) against
// 8 bytes, so the launch needs both HBM and the FP64 matrix cores (78.6 TFLOP/s) at once;
// launch bounds cap it at 256 registers a lane (two waves per SIMD).
template <typename T, int VEC>
__device__ __forceinline__ void bgemm_sym_item(const GItem it, const GTerm *__restrict__ terms,
                                               const T *__restrict__ A0, const T *__restrict__ A1,
                                               const T *const *__restrict__ tab, T *vec, T *smem) {
    typedef typename Mfma16<T>::v4 V4;
    constexpr int RG = 16 * VEC;           // rows of a slab
    constexpr int NG = kRowTile / RG;      // row groups of a full item
    constexpr int NA = NG * VEC;           // direct accumulators
    constexpr int NS = RG / 4;             // MFMA k-steps of a slab's transposed product
    constexpr int kPipe = 2;               // slabs in the load rotation (3 measured no faster)
    constexpr int LDR = RG + VEC;          // LDS column of a slab, padded by 16 bytes
    constexpr int XS = kRowTile * kNR;     // staged input rows
    constexpr int SLB = 16 * LDR;          // one slab buffer
    typedef T VL __attribute__((ext_vector_type(VEC)));
    T *const xs = smem;                    // [kRowTile][kNR]
    const T *__restrict__ A = tab ? tab[2 * (it.flags >> 8) + ((it.flags >> 1) & 1)] : ((it.flags & 2) ? A1 : A0);
    const CGTerm *const tms = (const CGTerm *)terms;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    T *const slab0 = smem + XS + 2 * wv * SLB;  // this wave's two buffers, [16][LDR] each
    const int rows = it.rows;
    const int ng = (rows + RG - 1) / RG;
    const int m = lane & 15, q = lane >> 4;
    const int64_t lda = it.lda;
    // the item's input rows X_b (zero beyond rows, up to the last row group)
    for (int e = threadIdx.x; e < ng * RG * kNR; e += 256) {
        const int r = e / kNR;
        xs[e] = r < rows ? vec[(it.xb + r) * kNR + (e % kNR)] : T(0);
    }
    __syncthreads();
    V4 acc[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a] = (V4)0;
    V4 acct[2] = {(V4)0, (V4)0};  // (two chains of the transposed MFMAs)
    // this lane's rows of every row group (out of range: past the end of any resource)
    uint32_t roff[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g)
        roff[g] = (g < ng && RG * g + VEC * m < rows) ? (uint32_t)((it.row0 + VEC * m + RG * g) * sizeof(T)) : kOutOfRange;
    // cursor: term tt, 16-column group cg of that term, row group rg
    int tt = it.t0, cg = wv, rg = 0;
    auto settle = [&]() {
        while (tt < it.t1) {
            const int ngr = (tms[tt].cols + 15) >> 4;
            if (cg < ngr) break;
            cg -= ngr;
            ++tt;
        }
    };
    struct Slab {
        VL av[4];
        T bx[4];
        int64_t t;
        int cols, c0, g;
        bool has;
    };
    auto load = [&](Slab &S) {
        S.has = tt < it.t1;
        const GTerm tm = term_at(tms, S.has ? tt : it.t1 - 1);
        S.t = tm.t;
        S.cols = tm.cols;
        S.c0 = 16 * cg;
        S.g = rg;
        const __amdgpu_buffer_rsrc_t ra = buf_rsrc(A + tm.a, (uint32_t)((int64_t)tm.cols * lda * sizeof(T)));
        const __amdgpu_buffer_rsrc_t rx = buf_rsrc(vec + tm.x * kNR, (uint32_t)((int64_t)tm.cols * kNR * sizeof(T)));
        uint32_t ro = kOutOfRange;
#pragma unroll
        for (int g = 0; g < NG; ++g) ro = g == rg ? roff[g] : ro;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = S.c0 + 4 * j + q;
            const bool ok = S.has && c < tm.cols;
            // (the group's X_p rows only with its first row group; later slabs keep them)
            S.bx[j] = buf_ld<T>(rx, opaque_off(ok && rg == 0 ? (uint32_t)((c * kNR + m) * sizeof(T)) : kOutOfRange));
            S.av[j] = buf_ldv_nt<VL>(
                ra, opaque_off(ok && ro != kOutOfRange ? (uint32_t)(c * lda * sizeof(T)) + ro : kOutOfRange));
        }
        if (S.has) {
            if (++rg == ng) {
                rg = 0;
                cg += 4;
                settle();
            }
        }
    };
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    // the slab waiting in LDS for its transposed product
    bool pend = false;
    int64_t pt = 0;
    int pcols = 0, pc0 = 0, pg = 0, pbuf = 0, cur = 0;
    T bxc[4] = {T(0), T(0), T(0), T(0)};
    auto compute = [&](const Slab &S) {
        // read back the pending slab (its LDS round trip overlaps the direct products below)
        T ta[NS], tb[NS];
        if (pend) {
            const T *sl = slab0 + pbuf * SLB;
            const T *xr = xs + (RG * pg) * kNR;
#pragma unroll
            for (int st = 0; st < NS; ++st) {
                ta[st] = sl[m * LDR + 4 * st + q];
                tb[st] = xr[(4 * st + q) * kNR + m];
            }
        }
        if (S.has) {
            if (S.g == 0) {
#pragma unroll
                for (int j = 0; j < 4; ++j) bxc[j] = S.bx[j];
            }
            // direct: rows of row group S.g
#pragma unroll
            for (int g = 0; g < NG; ++g)
                if (g == S.g)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
#pragma unroll
                        for (int e = 0; e < VEC; ++e)
                            acc[g * VEC + e] = Mfma16<T>::op(S.av[j][e], bxc[j], acc[g * VEC + e]);
        }
        if (pend) {
#pragma unroll
            for (int st = 0; st < NS; ++st) acct[st & 1] = Mfma16<T>::op(ta[st], tb[st], acct[st & 1]);
            if (pg == ng - 1) {
                T *yp = vec + pt * kNR + m;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int c = pc0 + Mfma16<T>::row(q, v);
                    if (c < pcols) atomicAdd(yp + (int64_t)c * kNR, acct[0][v] + acct[1][v]);
                }
                acct[0] = acct[1] = (V4)0;
            }
        }
        pend = S.has && S.t >= 0;
        if (pend) {
            // (the buffer written here was last read two slabs ago, and those reads are complete)
            T *sl = slab0 + cur * SLB;
#pragma unroll
            for (int j = 0; j < 4; ++j) *(VL *)&sl[(4 * j + q) * LDR + VEC * m] = S.av[j];
            pt = S.t;
            pcols = S.cols;
            pc0 = S.c0;
            pg = S.g;
            pbuf = cur;
            cur ^= 1;
        }
        wave_sync();
    };
    settle();
    // kPipe slabs rotate: while one is multiplied the next kPipe - 1 are in flight (a slab
    // loaded past the last one has has = false and adds nothing)
    Slab S[kPipe];
#pragma unroll
    for (int i = 0; i < kPipe; ++i) load(S[i]);
    while (S[0].has) {
#pragma unroll
        for (int i = 0; i < kPipe; ++i) {
            compute(S[i]);
            load(S[i]);
        }
    }
    if (pend) compute(S[0]);  // (S[0].has is false: the last pending transposed product only)
    // the four waves' direct products: 3 and 2 into 1 and 0, then 1 into 0 (LDS reused)
    T(*red)[kRowTile][kNR] = (T(*)[kRowTile][kNR])smem;
    auto row_of = [&](int g, int e, int v) { return RG * g + VEC * Mfma16<T>::row(q, v) + e; };
    auto put = [&](int slot) {
#pragma unroll
        for (int g = 0; g < NG; ++g)
            if (g < ng)
#pragma unroll
                for (int e = 0; e < VEC; ++e)
#pragma unroll
                    for (int v = 0; v < 4; ++v) red[slot][row_of(g, e, v)][m] = acc[g * VEC + e][v];
    };
    auto add = [&](int slot) {
#pragma unroll
        for (int g = 0; g < NG; ++g)
            if (g < ng)
#pragma unroll
                for (int e = 0; e < VEC; ++e)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[g * VEC + e][v] += red[slot][row_of(g, e, v)][m];
    };
    __syncthreads();  // (xs and the slabs are no longer read)
    if (wv >= 2) put(wv - 2);
    __syncthreads();
    if (wv < 2) add(wv);
    __syncthreads();
    if (wv == 1) put(0);
    __syncthreads();
    if (wv == 0) {
        T *y = vec + it.y * kNR + m;
#pragma unroll
        for (int g = 0; g < NG; ++g)
            if (g < ng)
#pragma unroll
                for (int e = 0; e < VEC; ++e)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = row_of(g, e, v);
                        if (r < rows) atomicAdd(y + (int64_t)r * kNR, acc[g * VEC + e][v] + red[0][r][m]);
                    }
    }
}
template <typename T, int VEC>
constexpr int sym_smem() {
    constexpr int RG = 16 * VEC, SLB = 16 * (RG + VEC), XS = kRowTile * kNR, RED = 2 * kRowTile * kNR;
    return XS + 8 * SLB > RED ? XS + 8 * SLB : RED;
}
// One item per workgroup.  An item that fits in one narrower row group (16 VEC rows: 8- or 4-byte
// loads per lane) runs in it: a box of 20 rows takes one 32-row group where a 64-row group would
// leave two thirds of every MFMA on zero rows.  Taller items keep the 16-byte loads (measured: more
// row groups of narrower loads are slower even when they pad less).
template <typename T>
__global__ __launch_bounds__(256, 2) void k_bgemm_sym(const GItem *__restrict__ items, const GTerm *__restrict__ terms,
                                                      const T *__restrict__ A0, const T *__restrict__ A1,
                                                      const T *const *__restrict__ tab, T *vec) {
    constexpr int VMAX = 16 / sizeof(T);
    __shared__ __attribute__((aligned(16))) T smem[sym_smem<T, VMAX>()];
    const GItem it = items[blockIdx.x];
    const int best = it.rows <= 16 ? 1 : it.rows <= 32 ? 2 : VMAX;
    if (best == VMAX) bgemm_sym_item<T, VMAX>(it, terms, A0, A1, tab, vec, smem);
    else if (best == 1) bgemm_sym_item<T, 1>(it, terms, A0, A1, tab, vec, smem);
    else if constexpr (VMAX == 4) bgemm_sym_item<T, 2>(it, terms, A0, A1, tab, vec, smem);
}

// Transposed products of the symmetric blocks on the 16-wide path:
// Y[j, n] += sum_i K[i, row0 + j] X[i, n] for an item's output rows j (box p) and
// terms K(b, p) (leading dimension tm.ld, inner dimension tm.cols = r_b), through LDS.
// The MFMA's A operand wants an output row (a column of K) per lane % 16 and the inner
// index per lane / 16, while K's contiguous direction is the inner one: loading that
// layout directly reads 64-byte runs per column.  Instead a slab is 16 VEC inner
// indices (256 bytes of every column); wave w owns output groups w and w + 4 (16 rows each) of the item and reads
// its 16-column groups slab by slab with 16 lanes per column (whole 256-byte runs),
// writes them to its own LDS region and reads them back in the MFMA's A layout (lane
// (m, q): VEC inner indices of column 16 g + m).  The next slab's loads are in flight
// while a slab is multiplied: the slab's B operands are loaded before them, since
// vector loads retire in order.  No wave reads another wave's region, and every wave
// writes only its own output rows.
template <typename T, int LPC>
__global__ __launch_bounds__(256) void k_bgemm_tl(const GItem *__restrict__ items, const GTerm *__restrict__ terms,
                                                  const T *__restrict__ A0, const T *__restrict__ A1,
                                                  const T *const *__restrict__ tab, T *vec) {
    typedef typename Mfma16<T>::v4 V4;
    constexpr int VEC = 16 / sizeof(T);
    constexpr int SL = LPC * VEC;      // inner indices of one slab (LPC lanes per column)
    constexpr int LD = SL + VEC;       // LDS row: one column's slab, padded by 16 bytes
    constexpr int CPL = 64 / LPC;      // columns per load instruction
    constexpr int NJ = 16 / CPL;       // load instructions per 16-column group
    constexpr int NSUB = LPC / 4;      // MFMA k-steps (4 VEC inner indices) per slab
    constexpr int U = 16 / LPC;        // slabs per iteration
    typedef T VL __attribute__((ext_vector_type(VEC)));
    __shared__ __attribute__((aligned(16))) T lds[4][U][32][LD];
    const GItem it = items[blockIdx.x];
    const CGTerm *const tms = (const CGTerm *)terms;
    const T *__restrict__ A = tab ? tab[2 * (it.flags >> 8) + ((it.flags >> 1) & 1)] : ((it.flags & 2) ? A1 : A0);
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int rows = it.rows;
    const int ng = (rows + 15) >> 4;
    const int m = lane & 15, q = lane >> 4;       // MFMA operand lane
    const int il = lane % LPC, cl = lane / LPC;   // load lane: inner chunk, column
    // Items of one or two groups split their slabs over the waves instead: wave w takes
    // group w % ng and every ns-th slab from the (w / ng)-th, and the partial products are
    // summed through LDS at the end.
    const int ns = ng <= 2 ? 4 / ng : 1;
    const int g0 = ns > 1 ? wv % ng : wv;  // first group of this wave (the second is wv + 4)
    const int nown = ns > 1 ? 1 : (wv < ng) + (wv + 4 < ng);
    if (nown == 0) return;  // (ns == 1 only: no barrier follows)
    V4 acc[2] = {(V4)0, (V4)0};
    int lt = it.t0, lk = ns > 1 ? wv / ng : 0;  // next slab to load (term, slab of the term)
    auto settle = [&]() {
        while (lt < it.t1) {
            const int nk = (tms[lt].cols + SL - 1) / SL;
            if (lk < nk) break;
            lk -= nk;
            ++lt;
        }
    };
    struct Slab {
        int64_t x;         // input offset
        int cols, i0;      // inner size, first inner index
        int okm, vlim;     // lanes' valid loads, valid inner indices from this lane's first
    };
    Slab S[U];
    VL L[U][2][NJ];
    // load the next slab into L[u] (past the last slab it still loads, from the last term,
    // and the loads are dropped: loads under a branch, or values used at once, would make
    // the next wait drain all of them)
    auto next = [&](int u) -> bool {
        settle();
        const bool has = lt < it.t1;
        const GTerm tm = term_at(tms, has ? lt : it.t1 - 1);
        S[u].x = tm.x;
        S[u].cols = tm.cols;
        S[u].i0 = has ? lk * SL : tm.cols;
        const int io = S[u].i0 + VEC * il;
        S[u].okm = 0;
        S[u].vlim = tm.cols - io;
#pragma unroll
        for (int gi = 0; gi < 2; ++gi)
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const int c = 16 * (gi ? wv + 4 : g0) + cl + CPL * j;  // item row = column of the block
                const bool ok = gi < nown && c < rows && io < tm.cols;
                S[u].okm |= ok ? 1 << (NJ * gi + j) : 0;
                L[u][gi][j] = ld_nt((const VL *)(A + tm.a + (ok ? (int64_t)(it.row0 + c) * tm.ld + io : 0)));
            }
        lk += has ? ns : 0;
        return has;
    };
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    bool more = false;
#pragma unroll
    for (int u = 0; u < U; ++u) more |= next(u);
    while (more) {
        // staged with every entry outside the block set to zero, so that the B operands of
        // inner indices beyond the block (clamped loads of a valid x) multiply zeros
        T bx[U][NSUB][VEC];
        Slab C[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int gi = 0; gi < 2; ++gi)
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    VL v = L[u][gi][j];
#pragma unroll
                    for (int e = 0; e < VEC; ++e)
                        v[e] = ((S[u].okm >> (NJ * gi + j)) & 1) && e < S[u].vlim ? v[e] : T(0);
                    *(VL *)&lds[wv][u][16 * gi + cl + CPL * j][VEC * il] = v;
                }
#pragma unroll
            for (int s = 0; s < NSUB; ++s) {
                const int i = S[u].i0 + 4 * VEC * s + VEC * q;
#pragma unroll
                for (int e = 0; e < VEC; ++e) bx[u][s][e] = vec[(S[u].x + min(i + e, S[u].cols - 1)) * kNR + m];
            }
            C[u] = S[u];
        }
        more = false;
#pragma unroll
        for (int u = 0; u < U; ++u) more |= next(u);
        wave_sync();
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (C[u].i0 < C[u].cols)  // (a slab past the last one multiplies nothing)
#pragma unroll
                for (int s = 0; s < NSUB; ++s)
#pragma unroll
                    for (int gi = 0; gi < 2; ++gi)
                        if (gi < nown) {
                            const VL a = *(const VL *)&lds[wv][u][16 * gi + m][4 * VEC * s + VEC * q];
#pragma unroll
                            for (int e = 0; e < VEC; ++e) acc[gi] = Mfma16<T>::op(a[e], bx[u][s][e], acc[gi]);
                        }
        wave_sync();
    }
    if (ns > 1) {
        // partial products of group g0 from waves g0, g0 + ng, ..: summed by wave g0
        T(*red)[kNR] = (T(*)[kNR])&lds[wv][0][0][0];
#pragma unroll
        for (int v = 0; v < 4; ++v) red[Mfma16<T>::row(q, v)][m] = acc[0][v];
        __syncthreads();
        if (wv >= ng) return;
        for (int o = wv + ng; o < 4; o += ng) {
            T(*ro)[kNR] = (T(*)[kNR])&lds[o][0][0][0];
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[0][v] += ro[Mfma16<T>::row(q, v)][m];
        }
    }
#pragma unroll
    for (int gi = 0; gi < 2; ++gi)
        if (gi < nown) {
            T *y = vec + it.y * kNR + m;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 16 * (gi ? wv + 4 : g0) + Mfma16<T>::row(q, v);
                if (r < rows) {
                    if (it.flags & 1) atomicAdd(y + (int64_t)r * kNR, acc[gi][v]);
                    else y[(int64_t)r * kNR] += acc[gi][v];
                }
            }
        }
}

// leading sub-block copies from the FP32 pool into the inner-operator buffer
__global__ void k_copy32(const CopyJob *__restrict__ jobs, const float *__restrict__ src, float *dst) {
    const CopyJob jb = jobs[blockIdx.x];
    const int64_t tot = (int64_t)jb.m * jb.n;
    for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
        int i = (int)(e % jb.m), j = (int)(e / jb.m);
        dst[jb.dst + (int64_t)j * jb.ldd + i] = src[jb.src + (int64_t)j * jb.lds + i];
    }
}

// ------------------------------------------------------------------ on-the-fly P2P / M2L
// The near-field and M2L terms of a work list are blocks of plain kernel entries
// K(target points, source points); this kernel evaluates them instead of reading
// them.  Target row r of an item is vector slot it.y + r, source column c of a
// term is slot tm.x + c; the slot tables give each slot's point.  VALU-bound:
// per entry 3 sub + 3 mul/fma + sqrt + log (3D RBF) and the diagonal test.
__device__ __forceinline__ float koff_f(int dim, int choice, float dx, float dy, float dz) {
    const float a = 1.0e-4f;
    if (dim == 1) {
        float R = fabsf(dx);
        if (choice == 0) return __logf(R);
        return R < a ? R * 1.0e4f : a / R;
    } else if (dim == 2) {
        float R2 = dx * dx + dy * dy;
        if (choice == 0) return -0.0795774715459476679f * __logf(R2);  // -(1/2pi) * 0.5 * log R^2
        float R = sqrtf(R2);
        return R < a ? R * 1.0e4f : a / R;
    } else {
        float R2 = dx * dx + dy * dy + dz * dz;
        if (choice == 0) return 0.0795774715459476679f * rsqrtf(R2);  // 1 / (4 pi R)
        float R = sqrtf(R2);
        float lr = __logf(R);
        // (R log R - 1) / (a log a - 1) below a, log R / log a above
        return R < a ? (R * lr - 1.0f) * -0.9990798135f : lr * -0.1085736205f;
    }
}

template <typename T>
__device__ __forceinline__ T koff_t(int dim, int choice, T dx, T dy, T dz);
template <>
__device__ __forceinline__ float koff_t<float>(int dim, int choice, float dx, float dy, float dz) {
    return koff_f(dim, choice, dx, dy, dz);
}
template <>
__device__ __forceinline__ double koff_t<double>(int dim, int choice, double dx, double dy, double dz) {
    return koff(dim, choice, dx, dy, dz);
}

// Coordinate differences: FP64 directly; FP32 from (hi, lo) pairs, where the
// hi difference of two nearby points is exact or nearly so and the lo
// difference restores the digits plain FP32 coordinates would lose.
template <typename T>
struct Coords {
    const T *x, *y, *z;
    const float *xl, *yl, *zl;
};

// FP32 entries from the squared distance R2 (1D: the distance R): the common case of each
// kernel (otf_val32) and the case of points closer than a = 1e-4 (otf_near32, RBF only;
// otf_rare32 says whether it applies).  log R = log2(R^2) ln2 / 2, so the common 3D and 2D
// cases need no square root; hardware log2 / rsq / rcp (about 1 ulp; arguments are never
// denormal: R >= 1e-4 in the common case).
template <int DIM, int CH>
__device__ __forceinline__ float otf_val32(float R2) {
    constexpr float kLn2 = 0.69314718055994531f;
    if constexpr (DIM == 3) {
        if constexpr (CH == 0) return 0.0795774715459476679f * __builtin_amdgcn_rsqf(R2);  // 1 / (4 pi R)
        return __builtin_amdgcn_logf(R2) * (float)(0.5 * 0.69314718055994531 / -9.2103403719761836);  // log R / log a
    } else if constexpr (DIM == 2) {
        if constexpr (CH == 0) return (-0.0795774715459476679f * kLn2) * __builtin_amdgcn_logf(R2);  // -(1/4pi) log R^2
        return 1.0e-4f * __builtin_amdgcn_rsqf(R2);  // a / R
    } else {
        if constexpr (CH == 0) return kLn2 * __builtin_amdgcn_logf(R2);  // log R (R2 holds R)
        return 1.0e-4f * __builtin_amdgcn_rcpf(R2);                     // a / R
    }
}
template <int DIM, int CH>
__device__ __forceinline__ bool otf_rare32(float R2) {
    if constexpr (CH == 0) return false;
    return DIM == 1 ? R2 < 1.0e-4f : R2 < 1.0e-8f;
}
template <int DIM>
__device__ __forceinline__ float otf_near32(float R2) {
    const float R = DIM == 1 ? R2 : sqrtf(R2);
    if constexpr (DIM == 3) return (R * logf(R) - 1.0f) * -0.9990798135f;  // (R log R - 1) / (a log a - 1)
    return R * 1.0e4f;                                                       // R / a
}

// One source point of a tile: coordinates (FP32: hi and lo parts), its value and point id.
template <typename T>
struct OtfSrc;
template <>
struct OtfSrc<float> {
    float x, y, z, xl, yl, zl, v;
    int pid;
};
template <>
struct OtfSrc<double> {
    double x, y, z, v;
    int pid, pad;
};

template <typename T>
struct OtfTgt {
    T x, y, z;
    float xl, yl, zl;
    int pid;
};

template <typename T>
__device__ __forceinline__ OtfTgt<T> otf_target(const Coords<T> &C, const int32_t *pid, int dim, int64_t ts) {
    OtfTgt<T> t;
    t.x = C.x[ts];
    t.y = dim > 1 ? C.y[ts] : T(0);
    t.z = dim > 2 ? C.z[ts] : T(0);
    t.xl = t.yl = t.zl = 0.f;
    if (C.xl) {
        t.xl = C.xl[ts];
        t.yl = dim > 1 ? C.yl[ts] : 0.f;
        t.zl = dim > 2 ? C.zl[ts] : 0.f;
    }
    t.pid = pid[ts];
    return t;
}

// Entries K(t0, q) and K(t1, q) of two target rows and one source, accumulated with the
// source's value.  DIAG: the diagonal rule where the points are the same (a tile holding one
// of the item's own points: the near-field self blocks).  The FP32 case for points closer
// than a is a branch the wave takes only when one of its lanes needs it.
template <typename T, int DIM, int CH, bool DIAG>
__device__ __forceinline__ void otf_pair(const OtfTgt<T> &t0, const OtfTgt<T> &t1, const OtfSrc<T> &q, T diag,
                                         T &acc0, T &acc1) {
    T v0, v1;
    if constexpr (sizeof(T) == 4) {
        float r0, r1;
        if constexpr (DIM == 1) {
            r0 = fabsf((t0.x - q.x) + (t0.xl - q.xl));
            r1 = fabsf((t1.x - q.x) + (t1.xl - q.xl));
        } else {
            const float ax = (t0.x - q.x) + (t0.xl - q.xl), bx = (t1.x - q.x) + (t1.xl - q.xl);
            const float ay = (t0.y - q.y) + (t0.yl - q.yl), by = (t1.y - q.y) + (t1.yl - q.yl);
            r0 = fmaf(ay, ay, ax * ax);
            r1 = fmaf(by, by, bx * bx);
            if constexpr (DIM == 3) {
                const float az = (t0.z - q.z) + (t0.zl - q.zl), bz = (t1.z - q.z) + (t1.zl - q.zl);
                r0 = fmaf(az, az, r0);
                r1 = fmaf(bz, bz, r1);
            }
        }
        v0 = otf_val32<DIM, CH>(r0);
        v1 = otf_val32<DIM, CH>(r1);
        if constexpr (CH != 0) {
            const bool n0 = otf_rare32<DIM, CH>(r0), n1 = otf_rare32<DIM, CH>(r1);
            if (__ballot(n0 || n1)) {
                if (n0) v0 = otf_near32<DIM>(r0);
                if (n1) v1 = otf_near32<DIM>(r1);
            }
        }
    } else {
        // the stored operator's own formula (koff), so the FP64 products agree to rounding
        v0 = koff(DIM, CH, t0.x - q.x, t0.y - q.y, t0.z - q.z);
        v1 = koff(DIM, CH, t1.x - q.x, t1.y - q.y, t1.z - q.z);
    }
    if constexpr (DIAG) {
        v0 = q.pid == t0.pid ? diag : v0;
        v1 = q.pid == t1.pid ? diag : v1;
    }
    acc0 = fma(v0, q.v, acc0);
    acc1 = fma(v1, q.v, acc1);
}

// P2P / M2L on the fly, tiled.  One workgroup per item (<= 128 target rows).  Source
// points of the item's terms are staged 256 at a time in LDS (coordinates, value, id: one
// coalesced load per source instead of one broadcast load per source and target row).
// Thread t owns target rows g and g + h (h = ceil(rows / 2), g = t mod h) in registers and
// the sources k = sub, sub + nsub, ... of each tile (sub = t / h, nsub = 256 / h), so every
// source read from LDS feeds two entries; the partial sums of the nsub source subsets are
// added through LDS at the end.  A tile is checked once for points of the item itself (the
// near-field self blocks); only such tiles test every entry for the diagonal.
constexpr int kOtfTile = 256;
template <typename T, int DIM, int CH>
__global__ __launch_bounds__(256) void k_otf(const GItem *__restrict__ items, const GTerm *__restrict__ terms,
                                             Coords<T> C, const int32_t *__restrict__ pid, T diag, T *vec) {
    constexpr int dim = DIM;
    __shared__ OtfSrc<T> tile[kOtfTile];
    __shared__ T part[2 * 256];
    __shared__ int tpmin, tpmax;
    const GItem it = items[blockIdx.x];
    const int R = it.rows;
    const int h = (R + 1) >> 1;
    const int nsub = 256 / h;
    const int t = threadIdx.x;
    const bool act = t < nsub * h;
    const int g = act ? t % h : 0, sub = act ? t / h : 0;
    const bool two = g + h < R;
    const OtfTgt<T> t0 = otf_target<T>(C, pid, dim, it.y + g);
    const OtfTgt<T> t1 = otf_target<T>(C, pid, dim, it.y + (two ? g + h : g));
    if (t == 0) {
        tpmin = INT_MAX;
        tpmax = INT_MIN;
    }
    __syncthreads();
    if (t < R) {
        const int p = pid[it.y + t];
        atomicMin(&tpmin, p);
        atomicMax(&tpmax, p);
    }
    __syncthreads();
    const int pmin = tpmin, pmax = tpmax;
    T acc0 = 0, acc1 = 0;
    for (int tt = it.t0; tt < it.t1; ++tt) {
        const GTerm tm = terms[tt];
        for (int c0 = 0; c0 < tm.cols; c0 += kOtfTile) {
            const int n = min(kOtfTile, tm.cols - c0);
            __syncthreads();  // (the previous tile is no longer read)
            bool own = false;
            if (t < n) {
                const int64_t sl = tm.x + c0 + t;
                OtfSrc<T> q;
                q.x = C.x[sl];
                q.y = dim > 1 ? C.y[sl] : T(0);
                q.z = dim > 2 ? C.z[sl] : T(0);
                if constexpr (sizeof(T) == 4) {
                    q.xl = C.xl ? C.xl[sl] : 0.f;
                    q.yl = dim > 1 && C.yl ? C.yl[sl] : 0.f;
                    q.zl = dim > 2 && C.zl ? C.zl[sl] : 0.f;
                }
                q.v = vec[sl];
                q.pid = pid[sl];
                tile[t] = q;
                own = q.pid >= pmin && q.pid <= pmax;
            }
            const bool diag_tile = __syncthreads_or(own);
            if (!act) continue;
            if (!diag_tile) {
                for (int k = sub; k < n; k += nsub) otf_pair<T, DIM, CH, false>(t0, t1, tile[k], diag, acc0, acc1);
            } else {
                for (int k = sub; k < n; k += nsub) otf_pair<T, DIM, CH, true>(t0, t1, tile[k], diag, acc0, acc1);
            }
        }
    }
    // partial sums of the source subsets: part[sub * R + row]
    __syncthreads();
    if (act) {
        part[sub * R + g] = acc0;
        if (two) part[sub * R + g + h] = acc1;
    }
    __syncthreads();
    if (t < R) {
        T sum = 0;
        for (int q = 0; q < nsub; ++q) sum += part[q * R + t];
        T *y = vec + it.y + t;
        if (it.flags & 1) atomicAdd(y, sum);
        else *y += sum;
    }
}

// ------------------------------------------------------------------ small vector kernels
__global__ void k_to_fp32(const double *a, float *b, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}
__global__ void k_to_fp64(const float *a, double *b, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}
// Read-only stream (the read bandwidth a streaming kernel is compared with): chunks of
// kStreamUnr x 256 16-byte words, one chunk per workgroup and trip (lane t loads words
// t, t + 256, ... of the chunk, all nontemporal and issued before any is used), chunks
// strided over the grid.  The words are folded into one that is written only when it
// equals a value it practically never takes, so the loads stay live and nothing is written.
constexpr int kStreamUnr = 8;
typedef uint32_t Word16 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_stream_read(const Word16 *__restrict__ p, int64_t nchunks,
                                                     uint32_t *__restrict__ out) {
    uint32_t acc = 0;
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const Word16 *q = p + c * (kStreamUnr * 256) + threadIdx.x;
        Word16 w[kStreamUnr];
#pragma unroll
        for (int u = 0; u < kStreamUnr; ++u) w[u] = __builtin_nontemporal_load(q + u * 256);
#pragma unroll
        for (int u = 0; u < kStreamUnr; ++u) acc = (acc ^ w[u].x ^ w[u].y) + (w[u].z ^ w[u].w);
    }
    if (acc == 0x9E3779B9u) out[blockIdx.x] = acc;
}

template <typename T>
__global__ void k_fill(T *a, int64_t n, T v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}
template <typename T>
__global__ void k_permute_in(const double *x, int64_t ldx, T *xt, int64_t ldt, const int32_t *perm, int64_t N) {
    const int r = blockIdx.y;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < N; t += (int64_t)gridDim.x * blockDim.x)
        xt[r * ldt + t] = (T)x[r * ldx + perm[t]];
}
template <typename T>
__global__ void k_permute_out(const T *yt, int64_t ldt, double *y, int64_t ldy, const int32_t *perm, int64_t N) {
    const int r = blockIdx.y;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < N; t += (int64_t)gridDim.x * blockDim.x)
        y[r * ldy + perm[t]] = (double)yt[r * ldt + t];
}

// Dense rows of A x with entries evaluated on the fly (getTruePoten,
// 3D/codes/HODLR3Dtree.hpp:2630-2643).  64 rows per workgroup, 4 column
// groups; source coordinates and x staged through LDS.
__global__ __launch_bounds__(256) void k_dense_rows(const int32_t *__restrict__ rows, int64_t nrows,
                                                    const double *__restrict__ xt, double *__restrict__ y, int64_t N,
                                                    KParams kp) {
    __shared__ double sx[256], sy[256], sz[256], sq[256];
    __shared__ double part[4][64];
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int64_t r = blockIdx.x * 64LL + lane;
    const bool active = r < nrows;
    const int i = active ? (rows ? rows[r] : (int)r) : 0;
    PointI p = load_point(kp, i);
    double acc = 0.0;
    for (int64_t j0 = 0; j0 < N; j0 += 256) {
        int64_t j = j0 + threadIdx.x;
        __syncthreads();
        if (j < N) {
            sx[threadIdx.x] = kp.px[j];
            sy[threadIdx.x] = kp.dim > 1 ? kp.py[j] : 0.0;
            sz[threadIdx.x] = kp.dim > 2 ? kp.pz[j] : 0.0;
            sq[threadIdx.x] = xt[j];
        }
        __syncthreads();
        const int nj = (int)min<int64_t>(256, N - j0);
        for (int q = g; q < nj; q += 4) {
            int jj = (int)(j0 + q);
            double v = (jj == i) ? kp.diag : koff(kp.dim, kp.choice, p.x - sx[q], p.y - sy[q], p.z - sz[q]);
            acc = fma(v, sq[q], acc);
        }
    }
    part[g][lane] = acc;
    __syncthreads();
    if (g == 0 && active) y[r] = part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane];
}

// ------------------------------------------------------------------ BLAS-1 for Arnoldi (FP64)
constexpr int kDotThreads = 256;

// partial[l * nblk + b] = sum over block b's rows of V_l . w
__global__ __launch_bounds__(kDotThreads) void k_mdot(const double *__restrict__ V, int64_t ldv,
                                                      const double *__restrict__ w, int64_t n,
                                                      double *__restrict__ partial) {
    const int b = blockIdx.x, l = blockIdx.y, nblk = gridDim.x;
    const double *v = V + (int64_t)l * ldv;
    double acc = 0.0;
    for (int64_t i = (int64_t)b * kDotThreads + threadIdx.x; i < n; i += (int64_t)nblk * kDotThreads)
        acc = fma(v[i], w[i], acc);
    acc = wave_sum(acc);
    __shared__ double sh[kDotThreads / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = 0; q < kDotThreads / 64; ++q) s += sh[q];
        partial[(int64_t)l * nblk + b] = s;
    }
}
__global__ void k_mdot_reduce(const double *__restrict__ partial, int nblk, double *__restrict__ out) {
    const int l = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 64) acc += partial[(int64_t)l * nblk + b];
    acc = wave_sum(acc);
    if (threadIdx.x == 0) out[l] = acc;
}
// w -= V h  (sign = -1) or x += V c (sign = +1)
__global__ void k_mgemv(const double *__restrict__ V, int64_t ldv, int k, const double *__restrict__ h,
                        double *__restrict__ w, int64_t n, double sign) {
    __shared__ double hs[1024];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i - threadIdx.x < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int l0 = 0; l0 < k; l0 += 1024) {
            int l1 = min(k, l0 + 1024);
            __syncthreads();
            for (int l = l0 + threadIdx.x; l < l1; l += blockDim.x) hs[l - l0] = h[l];
            __syncthreads();
            if (i < n)
                for (int l = l0; l < l1; ++l) acc = fma(V[(int64_t)l * ldv + i], hs[l - l0], acc);
        }
        if (i < n) w[i] += sign * acc;
    }
}
__global__ void k_axpby(double a, const double *__restrict__ x, double b, double *__restrict__ y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = b == 0.0 ? a * x[i] : a * x[i] + b * y[i];
}

__global__ void k_scal(double a, double *y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] *= a;
}


// ------------------------------------------------------------------ kNR-wide vectors (interleaved N x kNR)
// element (t, r) at t * kNR + r; one dot product, norm or update per column r
template <typename T>
__global__ void k_gather_m(const double *__restrict__ X, int64_t ldx, int nrhs, const int32_t *__restrict__ perm,
                           int64_t N, T *__restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * kNR; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = e / kNR;
        const int r = (int)(e % kNR);
        const int64_t src = perm ? perm[t] : t;
        dst[e] = r < nrhs ? (T)X[src + r * ldx] : T(0);
    }
}
template <typename T>
__global__ void k_scatter_m(const T *__restrict__ src, int64_t N, double *__restrict__ Y, int64_t ldy, int nrhs,
                            const int32_t *__restrict__ perm) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * kNR; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = e / kNR;
        const int r = (int)(e % kNR);
        if (r < nrhs) Y[(perm ? perm[t] : t) + r * ldy] = (double)src[e];
    }
}
// ---- lockstep solves: S systems x kNR columns.  A lockstep vector holds the
// systems' interleaved tree-order vectors one after another (n slots in all);
// column (s, r) is right-hand side r of system s, seg[t] the system of slot t
// (NULL: a single system).  Dot products reduce per block of slots (blk[b] =
// first and one-past-last slot, all in one system), then per system over the
// system's blocks sblk[s] .. sblk[s+1]-1.
// partial[(j * nblk + b) * kNR + r] = sum over block b of V_j(t, r) w(t, r)
__global__ __launch_bounds__(256) void k_bdot16(const double *__restrict__ V, int64_t ldv,
                                                const double *__restrict__ w, const int2 *__restrict__ blk,
                                                double *__restrict__ partial) {
    const int b = blockIdx.x, j = blockIdx.y, nblk = gridDim.x;
    const double *v = V + (int64_t)j * ldv;
    const int2 rg = blk[b];
    const int r = threadIdx.x & 15;
    double acc = 0.0;
    for (int64_t e = (int64_t)rg.x * kNR + threadIdx.x; e < (int64_t)rg.y * kNR; e += 256) acc = fma(v[e], w[e], acc);
    acc += __shfl_xor(acc, 16, 64);
    acc += __shfl_xor(acc, 32, 64);
    __shared__ double sh[4][kNR];
    if ((threadIdx.x & 63) < 16) sh[threadIdx.x >> 6][r] = acc;
    __syncthreads();
    if (threadIdx.x < 16)
        partial[((int64_t)j * nblk + b) * kNR + r] = sh[0][r] + sh[1][r] + sh[2][r] + sh[3][r];
}
// out[(j * S + s) * kNR + r] = sum over system s's blocks
__global__ __launch_bounds__(256) void k_bdot16_reduce(const double *__restrict__ partial, int nblk,
                                                       const int32_t *__restrict__ sblk, double *__restrict__ out) {
    const int j = blockIdx.x, sy = blockIdx.y, S = gridDim.y;
    const int r = threadIdx.x & 15, q = threadIdx.x >> 4;
    double acc = 0.0;
#pragma unroll 8
    for (int b = sblk[sy] + q; b < sblk[sy + 1]; b += 16) acc += partial[((int64_t)j * nblk + b) * kNR + r];
    acc += __shfl_xor(acc, 16, 64);
    acc += __shfl_xor(acc, 32, 64);
    __shared__ double sh[4][kNR];
    if ((threadIdx.x & 63) < 16) sh[threadIdx.x >> 6][r] = acc;
    __syncthreads();
    if (threadIdx.x < 16) out[((int64_t)j * S + sy) * kNR + r] = sh[0][r] + sh[1][r] + sh[2][r] + sh[3][r];
}
// w(t, r) += sign * sum_j V_j(t, r) h(j, seg(t), r)
__global__ __launch_bounds__(256) void k_bgemv16(const double *__restrict__ V, int64_t ldv, int k,
                                                 const double *__restrict__ h, int S, double sign,
                                                 double *__restrict__ w, int64_t n, const int32_t *__restrict__ seg) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * kNR; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t col = (seg ? (int64_t)seg[e / kNR] * kNR : 0) + (e % kNR);
        double acc = 0.0;
        for (int l = 0; l < k; ++l) acc = fma(V[(int64_t)l * ldv + e], h[(int64_t)l * S * kNR + col], acc);
        w[e] += sign * acc;
    }
}
// A space shared by every column of a one-system lockstep vector (the deflation space): k vectors
// of n slots, stored row-major (Cs[t * k + q]), so that the k values of a slot are contiguous.
// Both products with it are tall-skinny GEMMs and run on the FP64 matrix cores (16x16x4 MFMA:
// lane l supplies A(l % 16, l / 16) and B(l / 16, l % 16) and holds D(l / 16 + 4 v, l % 16), v < 4).
//
// partial[(q * nblk + b) * kNR + r] = sum over block b's slots of Cs(t, q) w(t, r), for the
// kSdTiles 16-vector tiles q = qg .. qg + 16 kSdTiles - 1 of this workgroup (grid.y covers k):
// D(16 q x 16 r) += A(16 q x 4 t) B(4 t x 16 r) with A(m, j) = Cs(t0 + j, q0 + m), B(j, n) = w(t0 + j, n);
// the four waves take every fourth group of 4 slots and are summed through LDS (the reduction over
// blocks is k_bdot16_reduce's).
constexpr int kSdTiles = 8;
__global__ __launch_bounds__(256) void k_sdot16(const double *__restrict__ Cs, int k, const double *__restrict__ w,
                                                const int2 *__restrict__ blk, double *__restrict__ partial) {
    typedef double v4 __attribute__((ext_vector_type(4)));
    const int b = blockIdx.x, nblk = gridDim.x;
    const int qg = blockIdx.y * 16 * kSdTiles;
    const int nz = min(kSdTiles, (k - qg + 15) / 16);
    const int2 rg = blk[b];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = lane & 15, j = lane >> 4;
    v4 acc[kSdTiles];
#pragma unroll
    for (int z = 0; z < kSdTiles; ++z) acc[z] = (v4)0.0;
#pragma unroll 2
    for (int t0 = rg.x + 4 * wv; t0 < rg.y; t0 += 16) {
        const int t = t0 + j;
        const bool ok = t < rg.y;
        const double bv = ok ? w[(int64_t)t * kNR + m] : 0.0;
        const double *c = Cs + (int64_t)t * k + qg + m;
        double av[kSdTiles];
#pragma unroll
        for (int z = 0; z < kSdTiles; ++z) av[z] = (z < nz && ok && qg + 16 * z + m < k) ? c[16 * z] : 0.0;
#pragma unroll
        for (int z = 0; z < kSdTiles; ++z)
            if (z < nz) acc[z] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[z], bv, acc[z], 0, 0, 0);
    }
    // waves 2, 3 -> waves 0, 1 -> wave 0
    __shared__ double sh[2][kSdTiles][4][64];
    if (wv >= 2)
#pragma unroll
        for (int z = 0; z < kSdTiles; ++z)
#pragma unroll
            for (int v = 0; v < 4; ++v) sh[wv - 2][z][v][lane] = acc[z][v];
    __syncthreads();
    if (wv < 2)
#pragma unroll
        for (int z = 0; z < kSdTiles; ++z)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[z][v] += sh[wv][z][v][lane];
    __syncthreads();
    if (wv == 1)
#pragma unroll
        for (int z = 0; z < kSdTiles; ++z)
#pragma unroll
            for (int v = 0; v < 4; ++v) sh[0][z][v][lane] = acc[z][v];
    __syncthreads();
    if (wv == 0)
#pragma unroll
        for (int z = 0; z < kSdTiles; ++z)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int q = qg + 16 * z + j + 4 * v;
                if (z < nz && q < k) partial[((int64_t)q * nblk + b) * kNR + m] = acc[z][v] + sh[0][z][v][lane];
            }
}
// w(t, r) += sign * sum_q Cs(t, q) h(q, seg(t), r)   (the shared space, row-major as above; with several
// systems every system's slots carry its own space, h holds the coefficients of every system).
// A wave takes 64 slots, four tiles of 16: D_i(16 t x 16 r) += A_i(16 t x 4 q) B(4 q x 16 r) with
// A_i(m, j) = Cs(t0 + 16 i + m, q), B(j, n) = h(q, s, n), q = q0 + j, one B load serving the four tiles;
// 16 vectors per trip, their 20 loads issued before the 16 MFMAs.  Slots of several systems in one
// group of 64 are taken once per system with the other systems' rows zeroed.
__global__ __launch_bounds__(256) void k_supd16(const double *__restrict__ Cs, int k, const double *__restrict__ h,
                                                int S, double sign, double *__restrict__ w, int64_t n,
                                                const int32_t *__restrict__ seg) {
    typedef double v4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63;
    const int m = lane & 15, j = lane >> 4;
    const int64_t ngrp = (n + 63) / 64;
    for (int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); g < ngrp; g += (int64_t)gridDim.x * 4) {
        const int64_t t0 = g * 64;
        const int s_lo = seg ? seg[t0] : 0;
        const int s_hi = seg ? seg[min(t0 + 63, n - 1)] : 0;
        const double *c[4];
        bool rok[4];
        int s_row[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t ta = t0 + 16 * i + m;  // this lane's A row of tile i
            rok[i] = ta < n;
            s_row[i] = seg && rok[i] ? seg[ta] : 0;
            c[i] = Cs + (rok[i] ? ta : 0) * k;
        }
        v4 acc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = (v4)0.0;
        for (int sy = s_lo; sy <= s_hi; ++sy) {
            const double *hq = h + (int64_t)sy * kNR + m;
            const int64_t hs = (int64_t)S * kNR;
            bool mine[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) mine[i] = rok[i] && s_row[i] == sy;
            for (int q0 = 0; q0 < k; q0 += 16) {
                double av[4][4], bv[4];
                // (a tile's four loads read the same 16 lines: issued one after another)
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int q = q0 + 4 * u + j;
                        av[i][u] = (mine[i] && q < k) ? c[i][q] : 0.0;
                    }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = q0 + 4 * u + j;
                    bv[u] = q < k ? hq[q * hs] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i][u], bv[u], acc[i], 0, 0, 0);
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t t = t0 + 16 * i + j + 4 * v;
                if (t < n) w[t * kNR + m] += sign * acc[i][v];
            }
    }
}
// dst[t * k + q] = src[q * n + t]: k column-major vectors of n entries to row-major
__global__ void k_to_rowmajor(const double *__restrict__ src, int64_t n, int k, double *__restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * k; e += (int64_t)gridDim.x * blockDim.x)
        dst[e] = src[(e % k) * n + e / k];
}

// dst(t, r) = sc(seg(t), r) * src(t, r)
__global__ void k_bscale16(const double *__restrict__ src, const double *__restrict__ sc, double *dst, int64_t n,
                           const int32_t *__restrict__ seg) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * kNR; e += (int64_t)gridDim.x * blockDim.x)
        dst[e] = sc[(seg ? (int64_t)seg[e / kNR] * kNR : 0) + (e % kNR)] * src[e];
}


// batch vector pool <-> lockstep vectors: lockstep slot t lives at pool slot map[t]
template <typename T>
__global__ void k_map_in(const double *__restrict__ X, const int32_t *__restrict__ map, int64_t n, T *__restrict__ pool) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * kNR; e += (int64_t)gridDim.x * blockDim.x)
        pool[(int64_t)map[e / kNR] * kNR + e % kNR] = (T)X[e];
}
template <typename T>
__global__ void k_map_out(const T *__restrict__ pool, const int32_t *__restrict__ map, int64_t n, double *__restrict__ Y) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * kNR; e += (int64_t)gridDim.x * blockDim.x)
        Y[e] = (double)pool[(int64_t)map[e / kNR] * kNR + e % kNR];
}
// column-major caller blocks <-> lockstep vectors: lockstep slot t reads row src[t] of a block with
// leading dimension ld[t] (per-system blocks concatenated), columns r < nrhs
__global__ void k_gather_b(const double *__restrict__ X, const int64_t *__restrict__ src, const int32_t *__restrict__ ld,
                           int nrhs, int64_t n, double *__restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * kNR; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = e / kNR;
        const int r = (int)(e % kNR);
        dst[e] = r < nrhs ? X[src[t] + (int64_t)r * ld[t]] : 0.0;
    }
}
__global__ void k_scatter_b(const double *__restrict__ src_v, const int64_t *__restrict__ src,
                            const int32_t *__restrict__ ld, int nrhs, int64_t n, double *__restrict__ Y) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * kNR; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = e / kNR;
        const int r = (int)(e % kNR);
        if (r < nrhs) Y[src[t] + (int64_t)r * ld[t]] = src_v[e];
    }
}


// ------------------------------------------------------------------ rank predictor
// A 16-64-64-64-1 perceptron with GELU (erf form) activations predicting
// log2(1 + rank) of an ACA block from its geometry features.  BF16 parameters
// (expanded to FP32 in LDS), FP32 accumulation, one thread per block.
// Parameter layout: W1 (F x H, row-major: input-major), b1, W2 (H x H), b2,
// W3 (H x H), b3, w4 (H), b4.
constexpr int kRF = 16, kRH = 64;
constexpr int kRP = kRF * kRH + kRH + 2 * (kRH * kRH + kRH) + kRH + 1;
__device__ __forceinline__ float bf16_f32(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }
__device__ __forceinline__ float gelu_erf(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__global__ __launch_bounds__(256) void k_rank_mlp(const float *__restrict__ feat, int n,
                                                  const uint16_t *__restrict__ w, float *__restrict__ out) {
    __shared__ float sw[kRP];
    for (int q = threadIdx.x; q < kRP; q += blockDim.x) sw[q] = bf16_f32(w[q]);
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *W1 = sw, *b1 = W1 + kRF * kRH, *W2 = b1 + kRH, *b2 = W2 + kRH * kRH, *W3 = b2 + kRH;
    const float *b3 = W3 + kRH * kRH, *w4 = b3 + kRH, *b4 = w4 + kRH;
    float x[kRF], h1[kRH], h2[kRH];
#pragma unroll
    for (int f = 0; f < kRF; ++f) x[f] = feat[(int64_t)i * kRF + f];
#pragma unroll
    for (int o = 0; o < kRH; ++o) {
        float a = b1[o];
#pragma unroll
        for (int f = 0; f < kRF; ++f) a = fmaf(x[f], W1[f * kRH + o], a);
        h1[o] = gelu_erf(a);
    }
#pragma unroll
    for (int o = 0; o < kRH; ++o) {
        float a = b2[o];
#pragma unroll
        for (int k = 0; k < kRH; ++k) a = fmaf(h1[k], W2[k * kRH + o], a);
        h2[o] = gelu_erf(a);
    }
#pragma unroll
    for (int o = 0; o < kRH; ++o) {
        float a = b3[o];
#pragma unroll
        for (int k = 0; k < kRH; ++k) a = fmaf(h2[k], W3[k * kRH + o], a);
        h1[o] = gelu_erf(a);
    }
    float y = b4[0];
#pragma unroll
    for (int k = 0; k < kRH; ++k) y = fmaf(h1[k], w4[k], y);
    out[i] = y;
}

inline int grid_for(int64_t n, int nt = 256, int cap = 8192) {
    int64_t g = (n + nt - 1) / nt;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace

// ------------------------------------------------------------------ launch wrappers
void launch_aca(const AcaJob *jobs, int njobs, const int32_t *idx, double *ws, int64_t ws_cap,
                unsigned long long *ws_top, int64_t *tab, uint8_t *flags, int32_t *rank, int32_t *status,
                int32_t *rowpiv, int32_t *colpiv, KParams kp, double tol, double *hist, int big, hipStream_t s) {
    if (njobs <= 0) return;
    if (big) hipLaunchKernelGGL(k_aca<1024>, dim3(njobs), dim3(1024), 0, s, jobs, idx, ws, ws_cap, ws_top, tab, flags,
                                rank, status, rowpiv, colpiv, kp, tol, hist);
    else hipLaunchKernelGGL(k_aca<256>, dim3(njobs), dim3(256), 0, s, jobs, idx, ws, ws_cap, ws_top, tab, flags, rank,
                            status, rowpiv, colpiv, kp, tol, hist);
}
void launch_lr_extract(const LrJob *jobs, int njobs, const double *ws, const int64_t *tab, const int32_t *rowpiv,
                       const int32_t *colpiv, double *out, hipStream_t s) {
    if (njobs > 0)
        hipLaunchKernelGGL(k_lr_extract, dim3(njobs), dim3(256), 0, s, jobs, ws, tab, rowpiv, colpiv, out);
}
void launch_eval(const EvalJob *jobs, int njobs, const int32_t *idx, double *A, KParams kp, hipStream_t s) {
    if (njobs > 0) hipLaunchKernelGGL(k_eval, dim3(njobs), dim3(256), 0, s, jobs, idx, A, kp);
}
void launch_xfer(const XferJob *jobs, int njobs, const int32_t *idx, const double *lr, double *A, KParams kp,
                 hipStream_t s) {
    if (njobs > 0) hipLaunchKernelGGL(k_xfer, dim3(njobs), dim3(64), 0, s, jobs, idx, lr, A, kp);
}
void launch_cur(const CurJob *jobs, int njobs, const int32_t *idx, const double *lr, double *A, KParams kp,
                hipStream_t s) {
    if (njobs > 0) hipLaunchKernelGGL(k_cur, dim3(njobs), dim3(256), 0, s, jobs, idx, lr, A, kp);
}
template <typename T, int ROLE>
static void launch_bgemv_r(const GItem *items, int nitems, const GTerm *terms, const T *A0, const T *A1, T *vec,
                           int unr, hipStream_t s) {
    constexpr int VEC = sizeof(T) == 8 ? 2 : 4;
    if (unr <= 2)
        hipLaunchKernelGGL((k_bgemv<T, VEC, ROLE, 2>), dim3(nitems), dim3(256), 0, s, items, terms, A0, A1, vec);
    else if (unr >= 8)
        hipLaunchKernelGGL((k_bgemv<T, VEC, ROLE, 8>), dim3(nitems), dim3(256), 0, s, items, terms, A0, A1, vec);
    else
        hipLaunchKernelGGL((k_bgemv<T, VEC, ROLE, 4>), dim3(nitems), dim3(256), 0, s, items, terms, A0, A1, vec);
}
template <typename T>
static void launch_bgemv_t(const GItem *items, int nitems, const GTerm *terms, const T *A0, const T *A1, T *vec,
                           int role, int unr, hipStream_t s) {
    if (nitems <= 0) return;
    if (!A1) A1 = A0;
    if (role == 1) launch_bgemv_r<T, 1>(items, nitems, terms, A0, A1, vec, unr, s);
    else if (role == 2) launch_bgemv_r<T, 2>(items, nitems, terms, A0, A1, vec, unr, s);
    else launch_bgemv_r<T, 0>(items, nitems, terms, A0, A1, vec, unr, s);
}
void launch_copy32(const CopyJob *jobs, int njobs, const float *src, float *dst, hipStream_t s) {
    if (njobs > 0) hipLaunchKernelGGL(k_copy32, dim3(njobs), dim3(256), 0, s, jobs, src, dst);
}
void launch_bgemv64(const GItem *items, int nitems, const GTerm *terms, const double *A0, const double *A1,
                    double *vec, int role, int unr, hipStream_t s) {
    launch_bgemv_t<double>(items, nitems, terms, A0, A1, vec, role, unr, s);
}
void launch_bgemv32(const GItem *items, int nitems, const GTerm *terms, const float *A0, const float *A1,
                    float *vec, int role, int unr, hipStream_t s) {
    launch_bgemv_t<float>(items, nitems, terms, A0, A1, vec, role, unr, s);
}
template <typename T>
static void launch_otf_t(const GItem *items, int nitems, const GTerm *terms, Coords<T> C, const SlotPoints &sp, T *vec,
                         hipStream_t s) {
    if (nitems <= 0) return;
    const T diag = (T)sp.diag;
    const int key = sp.dim * 2 + (sp.choice != 0);
#define H2W_OTF(D, CH) \
    hipLaunchKernelGGL((k_otf<T, D, CH>), dim3(nitems), dim3(256), 0, s, items, terms, C, sp.pid, diag, vec)
    switch (key) {
        case 2: H2W_OTF(1, 0); break;
        case 3: H2W_OTF(1, 1); break;
        case 4: H2W_OTF(2, 0); break;
        case 5: H2W_OTF(2, 1); break;
        case 6: H2W_OTF(3, 0); break;
        default: H2W_OTF(3, 1); break;
    }
#undef H2W_OTF
}
void launch_otf32(const GItem *items, int nitems, const GTerm *terms, SlotPoints sp, float *vec, hipStream_t s) {
    launch_otf_t<float>(items, nitems, terms, Coords<float>{sp.x, sp.y, sp.z, sp.xl, sp.yl, sp.zl}, sp, vec, s);
}
void launch_otf64(const GItem *items, int nitems, const GTerm *terms, SlotPoints sp, double *vec, hipStream_t s) {
    launch_otf_t<double>(items, nitems, terms, Coords<double>{sp.xd, sp.yd, sp.zd, nullptr, nullptr, nullptr}, sp,
                         vec, s);
}
void launch_to_fp32(const double *a, float *b, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_to_fp32, dim3(grid_for(n)), dim3(256), 0, s, a, b, n);
}
void launch_to_fp64(const float *a, double *b, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_to_fp64, dim3(grid_for(n)), dim3(256), 0, s, a, b, n);
}
int64_t stream_chunk_bytes() { return (int64_t)kStreamUnr * 256 * 16; }
void launch_stream_read(const void *p, int64_t bytes, uint32_t *out, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(256), 0, s, (const Word16 *)p, bytes / stream_chunk_bytes(), out);
}
void launch_fill(double *a, int64_t n, double v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(n)), dim3(256), 0, s, a, n, v);
}
void launch_fill32(float *a, int64_t n, float v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n)), dim3(256), 0, s, a, n, v);
}
void launch_permute_in(const double *x, int64_t ldx, double *xt, int64_t ldt, const int32_t *perm, int64_t N,
                       int nrhs, hipStream_t s) {
    hipLaunchKernelGGL(k_permute_in<double>, dim3(grid_for(N, 256, 2048), nrhs), dim3(256), 0, s, x, ldx, xt, ldt,
                       perm, N);
}
void launch_permute_in32(const double *x, int64_t ldx, float *xt, int64_t ldt, const int32_t *perm, int64_t N,
                         int nrhs, hipStream_t s) {
    hipLaunchKernelGGL(k_permute_in<float>, dim3(grid_for(N, 256, 2048), nrhs), dim3(256), 0, s, x, ldx, xt, ldt,
                       perm, N);
}
void launch_permute_out(const double *yt, int64_t ldt, double *y, int64_t ldy, const int32_t *perm, int64_t N,
                        int nrhs, hipStream_t s) {
    hipLaunchKernelGGL(k_permute_out<double>, dim3(grid_for(N, 256, 2048), nrhs), dim3(256), 0, s, yt, ldt, y, ldy,
                       perm, N);
}
void launch_permute_out32(const float *yt, int64_t ldt, double *y, int64_t ldy, const int32_t *perm, int64_t N,
                          int nrhs, hipStream_t s) {
    hipLaunchKernelGGL(k_permute_out<float>, dim3(grid_for(N, 256, 2048), nrhs), dim3(256), 0, s, yt, ldt, y, ldy,
                       perm, N);
}
void launch_dense_rows(const int32_t *rows, int64_t nrows, const double *xt, double *y, int64_t N, KParams kp,
                       hipStream_t s) {
    if (nrows > 0)
        hipLaunchKernelGGL(k_dense_rows, dim3((unsigned)((nrows + 63) / 64)), dim3(256), 0, s, rows, nrows, xt, y, N,
                           kp);
}
int mdot_blocks(int64_t n) {
    int64_t b = (n + kDotThreads * 8 - 1) / (kDotThreads * 8);
    if (b < 1) b = 1;
    if (b > 512) b = 512;
    return (int)b;
}
void launch_mdot(const double *V, int64_t ldv, int k, const double *w, int64_t n, double *partial, int nblk,
                 hipStream_t s) {
    if (k > 0) hipLaunchKernelGGL(k_mdot, dim3(nblk, k), dim3(kDotThreads), 0, s, V, ldv, w, n, partial);
}
void launch_mdot_reduce(const double *partial, int nblk, int k, double *out, hipStream_t s) {
    if (k > 0) hipLaunchKernelGGL(k_mdot_reduce, dim3(k), dim3(64), 0, s, partial, nblk, out);
}
void launch_mgemv_sub(const double *V, int64_t ldv, int k, const double *h, double *w, int64_t n, hipStream_t s) {
    if (k > 0) hipLaunchKernelGGL(k_mgemv, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, V, ldv, k, h, w, n, -1.0);
}
void launch_vec_update(const double *V, int64_t ldv, int k, const double *c, double *x, int64_t n, hipStream_t s) {
    if (k > 0) hipLaunchKernelGGL(k_mgemv, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, V, ldv, k, c, x, n, 1.0);
}
void launch_axpby(double a, const double *x, double b, double *y, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(256), 0, s, a, x, b, y, n);
}
void launch_scal(double a, double *y, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_scal, dim3(grid_for(n)), dim3(256), 0, s, a, y, n);
}


template <typename T, int ROLE>
static void launch_bgemm_r(const GItem *items, int nitems, const GTerm *terms, const T *A0, const T *A1,
                           const T *const *tab, T *vec, bool per_wave, hipStream_t s) {
    if (per_wave)
        hipLaunchKernelGGL((k_bgemm<T, ROLE, 1>), dim3((nitems + 3) / 4), dim3(256), 0, s, items, nitems, terms, A0, A1,
                           tab, vec);
    else
        hipLaunchKernelGGL((k_bgemm<T, ROLE, 4>), dim3(nitems), dim3(256), 0, s, items, nitems, terms, A0, A1, tab,
                           vec);
}
template <typename T>
static void launch_bgemm_t(const GItem *items, int nitems, const GTerm *terms, const T *A0, const T *A1,
                           const T *const *tab, T *vec, int role, bool per_wave, hipStream_t s) {
    if (nitems <= 0) return;
    if (!A1) A1 = A0;
    if (role == 1) launch_bgemm_r<T, 1>(items, nitems, terms, A0, A1, tab, vec, per_wave, s);
    else if (role == 2) launch_bgemm_r<T, 2>(items, nitems, terms, A0, A1, tab, vec, per_wave, s);
    else launch_bgemm_r<T, 0>(items, nitems, terms, A0, A1, tab, vec, per_wave, s);
}
void launch_bgemm64(const GItem *items, int nitems, const GTerm *terms, const double *A0, const double *A1,
                    double *vec, int role, bool per_wave, hipStream_t s) {
    launch_bgemm_t<double>(items, nitems, terms, A0, A1, nullptr, vec, role, per_wave, s);
}
void launch_bgemm32(const GItem *items, int nitems, const GTerm *terms, const float *A0, const float *A1, float *vec,
                    int role, bool per_wave, hipStream_t s) {
    launch_bgemm_t<float>(items, nitems, terms, A0, A1, nullptr, vec, role, per_wave, s);
}
void launch_bgemm64_tab(const GItem *items, int nitems, const GTerm *terms, const double *const *tab, double *vec,
                        int role, bool per_wave, hipStream_t s) {
    launch_bgemm_t<double>(items, nitems, terms, nullptr, nullptr, tab, vec, role, per_wave, s);
}
void launch_bgemm32_tab(const GItem *items, int nitems, const GTerm *terms, const float *const *tab, float *vec,
                        int role, bool per_wave, hipStream_t s) {
    launch_bgemm_t<float>(items, nitems, terms, nullptr, nullptr, tab, vec, role, per_wave, s);
}
void launch_gather_m64(const double *X, int64_t ldx, int nrhs, const int32_t *perm, int64_t N, double *dst,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_gather_m<double>, dim3(grid_for(N * kNR)), dim3(256), 0, s, X, ldx, nrhs, perm, N, dst);
}
void launch_gather_m32(const double *X, int64_t ldx, int nrhs, const int32_t *perm, int64_t N, float *dst,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_gather_m<float>, dim3(grid_for(N * kNR)), dim3(256), 0, s, X, ldx, nrhs, perm, N, dst);
}
void launch_scatter_m64(const double *src, int64_t N, double *Y, int64_t ldy, int nrhs, const int32_t *perm,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_m<double>, dim3(grid_for(N * kNR)), dim3(256), 0, s, src, N, Y, ldy, nrhs, perm);
}
void launch_scatter_m32(const float *src, int64_t N, double *Y, int64_t ldy, int nrhs, const int32_t *perm,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_m<float>, dim3(grid_for(N * kNR)), dim3(256), 0, s, src, N, Y, ldy, nrhs, perm);
}
void launch_bdot16(const double *V, int64_t ldv, int k, const double *w, const int2 *blk, int nblk, double *partial,
                   hipStream_t s) {
    if (k > 0) hipLaunchKernelGGL(k_bdot16, dim3(nblk, k), dim3(256), 0, s, V, ldv, w, blk, partial);
}
void launch_bdot16_reduce(const double *partial, int nblk, int k, const int32_t *sblk, int S, double *out,
                          hipStream_t s) {
    if (k > 0) hipLaunchKernelGGL(k_bdot16_reduce, dim3(k, S), dim3(256), 0, s, partial, nblk, sblk, out);
}
void launch_bgemv16(const double *V, int64_t ldv, int k, const double *h, int S, double sign, double *w, int64_t n,
                    const int32_t *seg, hipStream_t s) {
    if (k > 0)
        hipLaunchKernelGGL(k_bgemv16, dim3(grid_for(n * kNR, 256, 4096)), dim3(256), 0, s, V, ldv, k, h, S, sign, w, n,
                           seg);
}
void launch_sdot16(const double *Cs, int k, const double *w, const int2 *blk, int nblk, double *partial, hipStream_t s) {
    if (k > 0)
        hipLaunchKernelGGL(k_sdot16, dim3(nblk, (k + 16 * kSdTiles - 1) / (16 * kSdTiles)), dim3(256), 0, s, Cs, k, w, blk,
                           partial);
}
void launch_supd16(const double *Cs, int k, const double *h, int S, double sign, double *w, int64_t n,
                   const int32_t *seg, hipStream_t s) {
    if (k > 0 && n > 0)
        hipLaunchKernelGGL(k_supd16, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 8192)), dim3(256), 0, s, Cs, k, h, S,
                           sign, w, n, seg);
}
void launch_to_rowmajor(const double *src, int64_t n, int k, double *dst, hipStream_t s) {
    if (k > 0) hipLaunchKernelGGL(k_to_rowmajor, dim3(grid_for(n * k, 256, 4096)), dim3(256), 0, s, src, n, k, dst);
}
void launch_bscale16(const double *src, const double *sc, double *dst, int64_t n, const int32_t *seg, hipStream_t s) {
    hipLaunchKernelGGL(k_bscale16, dim3(grid_for(n * kNR)), dim3(256), 0, s, src, sc, dst, n, seg);
}
void launch_map_in64(const double *X, const int32_t *map, int64_t n, double *pool, hipStream_t s) {
    hipLaunchKernelGGL(k_map_in<double>, dim3(grid_for(n * kNR)), dim3(256), 0, s, X, map, n, pool);
}
void launch_map_in32(const double *X, const int32_t *map, int64_t n, float *pool, hipStream_t s) {
    hipLaunchKernelGGL(k_map_in<float>, dim3(grid_for(n * kNR)), dim3(256), 0, s, X, map, n, pool);
}
void launch_map_out64(const double *pool, const int32_t *map, int64_t n, double *Y, hipStream_t s) {
    hipLaunchKernelGGL(k_map_out<double>, dim3(grid_for(n * kNR)), dim3(256), 0, s, pool, map, n, Y);
}
void launch_map_out32(const float *pool, const int32_t *map, int64_t n, double *Y, hipStream_t s) {
    hipLaunchKernelGGL(k_map_out<float>, dim3(grid_for(n * kNR)), dim3(256), 0, s, pool, map, n, Y);
}
void launch_gather_b(const double *X, const int64_t *src, const int32_t *ld, int nrhs, int64_t n, double *dst,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_gather_b, dim3(grid_for(n * kNR)), dim3(256), 0, s, X, src, ld, nrhs, n, dst);
}
void launch_scatter_b(const double *v, const int64_t *src, const int32_t *ld, int nrhs, int64_t n, double *Y,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_b, dim3(grid_for(n * kNR)), dim3(256), 0, s, v, src, ld, nrhs, n, Y);
}
void launch_rank_mlp(const float *feat, int n, const uint16_t *w, float *out, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_rank_mlp, dim3((n + 255) / 256), dim3(256), 0, s, feat, n, w, out);
}
void launch_wide_init(const AcaJob *jobs, int nb, int maxmn, WideState *state, uint8_t *flags, int64_t *tab,
                      unsigned long long *ws_top, int64_t ws_cap, hipStream_t s) {
    hipLaunchKernelGGL(k_wide_init, dim3(std::max(1, std::min(1024, (maxmn + 2047) / 2048)), nb), dim3(256), 0, s,
                       jobs, state, flags, tab, ws_top, ws_cap);
}
void launch_wide_step(const AcaJob *jobs, int nb, int maxm, int maxn, int maxrank, const int32_t *idx, double *ws,
                      int64_t *tab, uint8_t *flags, WideState *state, WidePart *part, int pstride, double *dots,
                      int dstride, int32_t *rowpiv, int32_t *colpiv, double *hist, unsigned long long *ws_top,
                      int64_t ws_cap, KParams kp, double tol, hipStream_t s) {
    const int gr = std::max(1, (maxn + kWideCols - 1) / kWideCols), gc = std::max(1, (maxm + kWideCols - 1) / kWideCols);
    hipLaunchKernelGGL(k_wide_row, dim3(gr, nb), dim3(256), 0, s, jobs, idx, ws, ws, tab, flags, state, part, pstride,
                       kp);
    hipLaunchKernelGGL(k_wide_row_reduce, dim3(nb), dim3(256), 0, s, jobs, state, part, pstride);
    hipLaunchKernelGGL(k_wide_col, dim3(gc, nb), dim3(256), 0, s, jobs, idx, ws, ws, tab, flags, state, part, pstride,
                       kp);
    hipLaunchKernelGGL(k_wide_accept, dim3(nb), dim3(256), 0, s, jobs, state, part, pstride, flags, rowpiv, colpiv,
                       hist, tab, ws_top, ws_cap, tol);
    if (maxrank >= 2) {
        hipLaunchKernelGGL(k_wide_dots, dim3(std::max(1, maxrank), nb), dim3(256), 0, s, jobs, ws, tab, state, dots,
                           dstride);
        hipLaunchKernelGGL(k_wide_norm, dim3(nb), dim3(256), 0, s, jobs, state, dots, dstride, hist, tab, ws_top,
                           ws_cap, tol);
    }
}
void launch_wide_finish(const WideState *state, int nb, int32_t *rank_out, int32_t *status_out, int32_t *ndone,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_wide_finish, dim3(1), dim3(64), 0, s, state, nb, rank_out, status_out, ndone);
}
void launch_wide_fixed_step(const AcaJob *jobs, int nb, int maxm, int maxn, const int32_t *idx, double *ws,
                            int64_t *tab, uint8_t *flags, WideState *state, WidePart *part, int pstride,
                            int32_t *rowpiv, int32_t *colpiv, double *hist, unsigned long long *ws_top, int64_t ws_cap,
                            KParams kp, double tol, hipStream_t s) {
    const int gr = std::max(1, (maxn + kWideCols - 1) / kWideCols), gc = std::max(1, (maxm + kWideCols - 1) / kWideCols);
    hipLaunchKernelGGL(k_wide_row, dim3(gr, nb), dim3(256), 0, s, jobs, idx, ws, ws, tab, flags, state, part, pstride,
                       kp);
    hipLaunchKernelGGL(k_wide_row_reduce, dim3(nb), dim3(256), 0, s, jobs, state, part, pstride);
    hipLaunchKernelGGL(k_wide_col, dim3(gc, nb), dim3(256), 0, s, jobs, idx, ws, ws, tab, flags, state, part, pstride,
                       kp);
    hipLaunchKernelGGL(k_wide_accept, dim3(nb), dim3(256), 0, s, jobs, state, part, pstride, flags, rowpiv, colpiv,
                       hist, tab, ws_top, ws_cap, tol);
}
void launch_wide_gram(const GramJob *gj, int njobs, const AcaJob *jobs, const WideState *state, const double *ws,
                      const int64_t *tab, double *prod, hipStream_t s) {
    if (njobs > 0) hipLaunchKernelGGL(k_wide_gram, dim3(njobs), dim3(256), 0, s, gj, jobs, state, ws, tab, prod);
}
void launch_wide_judge(const AcaJob *jobs, int nb, WideState *state, const int32_t *final_rank, const double *normS_in,
                       int64_t *tab, unsigned long long *ws_top, int64_t ws_cap, hipStream_t s) {
    if (nb > 0)
        hipLaunchKernelGGL(k_wide_judge, dim3((nb + 63) / 64), dim3(64), 0, s, jobs, nb, state, final_rank, normS_in,
                           tab, ws_top, ws_cap);
}
void launch_aca_sample(const AcaJob *jobs, int nb, const int32_t *rank, const int32_t *idx, const double *ws,
                       const int64_t *tab, KParams kp, int ns, double *err2, double *ref2, hipStream_t s) {
    if (nb > 0) hipLaunchKernelGGL(k_aca_sample, dim3(nb), dim3(64), 0, s, jobs, rank, idx, ws, tab, kp, ns, err2, ref2);
}
template <typename T>
static void launch_bgemm_tl(const GItem *items, int nitems, const GTerm *terms, const T *A0, const T *A1,
                            const T *const *tab, T *vec, int narrow, hipStream_t s) {
    if (nitems <= 0) return;
    if (narrow)
        hipLaunchKernelGGL((k_bgemm_tl<T, 4>), dim3(nitems), dim3(256), 0, s, items, terms, A0, A1 ? A1 : A0, tab, vec);
    else
        hipLaunchKernelGGL((k_bgemm_tl<T, 16>), dim3(nitems), dim3(256), 0, s, items, terms, A0, A1 ? A1 : A0, tab, vec);
}
void launch_bgemm_sym64(const GItem *items, int nitems, const GTerm *terms, const double *A0, const double *A1,
                        const double *const *tab, double *vec, hipStream_t s) {
    if (nitems > 0)
        hipLaunchKernelGGL(k_bgemm_sym<double>, dim3(nitems), dim3(256), 0, s, items, terms, A0, A1 ? A1 : A0, tab, vec);
}
void launch_bgemm_sym32(const GItem *items, int nitems, const GTerm *terms, const float *A0, const float *A1,
                        const float *const *tab, float *vec, hipStream_t s) {
    if (nitems > 0)
        hipLaunchKernelGGL(k_bgemm_sym<float>, dim3(nitems), dim3(256), 0, s, items, terms, A0, A1 ? A1 : A0, tab, vec);
}
void launch_bgemm_t64(const GItem *items, int nitems, const GTerm *terms, const double *A0, const double *A1,
                      const double *const *tab, double *vec, int narrow, hipStream_t s) {
    launch_bgemm_tl<double>(items, nitems, terms, A0, A1, tab, vec, narrow, s);
}
void launch_bgemm_t32(const GItem *items, int nitems, const GTerm *terms, const float *A0, const float *A1,
                      const float *const *tab, float *vec, int narrow, hipStream_t s) {
    launch_bgemm_tl<float>(items, nitems, terms, A0, A1, tab, vec, narrow, s);
}
}  // namespace h2w
