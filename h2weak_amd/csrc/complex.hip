// complex.hip -- the complex-kernel path of libh2w: the reference's compile-time complex
// builds (-DUSE_Helm: 3D Helmholtz, 3D/codes/kernel.hpp:132-144; -DUSE_Hankel: 2D Hankel
// H0^(1), 2D/codes/kernel.hpp:126-145) as a run-time choice.  Every operator block,
// vector and Krylov basis is complex FP64, stored as interleaved (re, im) pairs (double2).
// The kernels are complex symmetric (K(x, y) = K(y, x)): transfers, the vertex M2M and the
// M2L blocks use plain transposes, as the reference's do, so the work lists and the pool
// layout of the real path serve unchanged, with complex elements.
//
// Compiled with -ffp-contract=off like kernels.hip (the cross rounds like scalar code).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "h2w_internal.hpp"

namespace h2w {

namespace {

constexpr double kPIc = 3.141592653589793238462643383279502884;

// ------------------------------------------------------------------ complex arithmetic
__device__ __forceinline__ double2 cmk(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return cmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return cmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) { return cmk(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    // Smith's algorithm (avoids overflow in |b|^2)
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, d = b.x + b.y * r;
        return cmk((a.x + a.y * r) / d, (a.y - a.x * r) / d);
    }
    const double r = b.x / b.y, d = b.y + b.x * r;
    return cmk((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double cabsd(double2 a) { return hypot(a.x, a.y); }

// ------------------------------------------------------------------ kernel entries
// 3D Helmholtz (choice 2): e^{iR} / (4 pi) off the diagonal, 500 sqrt(N) on it
// (3D/codes/kernel.hpp:132-144).  2D Hankel (choice 10): (i/4) H0^(1)(R) = (i J0(R) - Y0(R)) / 4 off
// the diagonal, (1 + 15 log N + Kii) / h^2 on it with the complex Kii of the same integrand
// (2D/codes/kernel.hpp:126-145, 2D/codes/integral2d.hpp:16-57); kp.diag / kp.diag_im hold it.
struct PointC {
    double x, y, z;
    int i;
};
__device__ __forceinline__ PointC load_point_c(const KParamsC &kp, int i) {
    PointC p;
    p.i = i;
    p.x = kp.px[i];
    p.y = kp.py[i];
    p.z = kp.dim > 2 ? kp.pz[i] : 0.0;
    return p;
}
__device__ __forceinline__ double2 koff_c(int dim, double dx, double dy, double dz) {
    if (dim == 3) {
        const double R = sqrt(dx * dx + dy * dy + dz * dz);
        double s, c;
        sincos(R, &s, &c);
        return cmk(c / (4.0 * kPIc), s / (4.0 * kPIc));
    }
    const double R = sqrt(dx * dx + dy * dy);
    return cmk(-y0(R) / 4.0, j0(R) / 4.0);
}
__device__ __forceinline__ double2 kentry_c(const KParamsC &kp, const PointC &p, int j) {
    if (p.i == j) return cmk(kp.diag, kp.diag_im);
    return koff_c(kp.dim, p.x - kp.px[j], p.y - kp.py[j], kp.dim > 2 ? p.z - kp.pz[j] : 0.0);
}
__device__ __forceinline__ double2 kentry2_c(const KParamsC &kp, int i, int j) {
    return kentry_c(kp, load_point_c(kp, i), j);
}

// ------------------------------------------------------------------ reductions
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// first argmax of |v| (tie: lower index); sv the entry itself
struct AMaxC {
    double a;
    int idx;
    double2 sv;
};
__device__ __forceinline__ void amaxc_merge(AMaxC &x, const AMaxC &y) {
    if (y.a > x.a || (y.a == x.a && y.idx < x.idx)) x = y;
}
__device__ __forceinline__ void amaxc_update(AMaxC &m, double2 v, int idx) {
    const double a = cabsd(v);
    if (a > m.a || (a == m.a && idx < m.idx)) {
        m.a = a;
        m.idx = idx;
        m.sv = v;
    }
}
template <int NT>
struct BlockRedC {
    static constexpr int NW = NT / 64;
    double s0[NW], s1[NW], a[NW];
    double2 sv[NW];
    int idx[NW];
};
template <int NT>
__device__ __forceinline__ void block_reduce_c(BlockRedC<NT> &sh, double &s0, double &s1, AMaxC &m) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    s0 = wsum(s0);
    s1 = wsum(s1);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        AMaxC y;
        y.a = __shfl_xor(m.a, o, 64);
        y.idx = __shfl_xor(m.idx, o, 64);
        y.sv.x = __shfl_xor(m.sv.x, o, 64);
        y.sv.y = __shfl_xor(m.sv.y, o, 64);
        amaxc_merge(m, y);
    }
    __syncthreads();
    if (lane == 0) {
        sh.s0[w] = s0;
        sh.s1[w] = s1;
        sh.a[w] = m.a;
        sh.idx[w] = m.idx;
        sh.sv[w] = m.sv;
    }
    __syncthreads();
    s0 = s1 = 0.0;
    m.a = -1.0;
    m.idx = INT_MAX;
    m.sv = cmk(0.0, 0.0);
    for (int q = 0; q < BlockRedC<NT>::NW; ++q) {
        s0 += sh.s0[q];
        s1 += sh.s1[q];
        AMaxC o{sh.a[q], sh.idx[q], sh.sv[q]};
        amaxc_merge(m, o);
    }
}

// ------------------------------------------------------------------ complex cross (one workgroup per block)
// The pivot loop of LowRank::ACA_only_nodes (2D/codes/aca.hpp:296-446) with complex entries:
// pivots by |entry| (first argmax over remaining indices), u = residual column / pivot, stop when
// |u||v| <= tol normS (2-norms of complex vectors) or rank = min(m, n), normS^2 += (|u||v|)^2 +
// 2 sum_{l <= new} |<U_l, u>||<V_l, v>| (Eigen's dot conjugates its first argument; only the
// moduli enter).  The workspace holds complex U, V columns (offsets in complex elements).
constexpr int kChunkC = 256;
template <int NT>
__device__ __forceinline__ bool aca_reserve_c(int64_t *tab, int k, int m, int n, int tcap,
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
__global__ __launch_bounds__(NT) void k_aca_c(const AcaJob *__restrict__ jobs, const int32_t *__restrict__ idx,
                                              double2 *__restrict__ ws, int64_t ws_cap, unsigned long long *ws_top,
                                              int64_t *__restrict__ tab_all, uint8_t *__restrict__ flags,
                                              int32_t *__restrict__ rank_out, int32_t *__restrict__ status_out,
                                              int32_t *__restrict__ rowpiv, int32_t *__restrict__ colpiv, KParamsC kp,
                                              double tol, double *__restrict__ hist_all) {
    const AcaJob jb = jobs[blockIdx.x];
    const int m = jb.m, n = jb.n, tcap = jb.tcap;
    double *hist = hist_all + 2 * jb.piv;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    constexpr int NW = NT / 64;
    const int32_t *rows = idx + jb.rows;
    const int32_t *cols = idx + jb.cols;
    int64_t *tab = tab_all + jb.tab;
    uint8_t *ru = flags + jb.wsF;
    uint8_t *cu = ru + m;
    int32_t *rp_out = rowpiv + jb.piv;
    int32_t *cp_out = colpiv + jb.piv;
    __shared__ double2 coef[kChunkC];
    __shared__ int64_t cofs[kChunkC];
    __shared__ BlockRedC<NT> red;
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
    if (!aca_reserve_c<NT>(tab, 0, m, n, tcap, ws_top, ws_cap, &sh_off)) {
        if (tid == 0) {
            rank_out[blockIdx.x] = 0;
            status_out[blockIdx.x] = 1;
        }
        return;
    }
    double2 *U0 = ws + sh_off;
    double2 *V0 = U0 + m;
    int l = 0, cp = -1;
    double2 piv = cmk(0.0, 0.0);
    for (l = 0; l < m; ++l) {
        PointC pr = load_point_c(kp, rows[l]);
        AMaxC am{-1.0, INT_MAX, cmk(0.0, 0.0)};
        for (int s = tid; s < n; s += NT) {
            const double2 v = kentry_c(kp, pr, cols[s]);
            V0[s] = v;
            amaxc_update(am, v, s);
        }
        double d0 = 0.0, d1 = 0.0;
        block_reduce_c<NT>(red, d0, d1, am);
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
    double un2 = 0.0, vn2 = 0.0;
    AMaxC an{-1.0, INT_MAX, cmk(0.0, 0.0)};
    {
        PointC pc = load_point_c(kp, cols[cp]);
        for (int t = tid; t < m; t += NT) {
            const double2 c = kentry_c(kp, pc, rows[t]);  // A(rows[t], cols[cp]) (symmetric kernel)
            const double2 u = cdiv(c, piv);
            U0[t] = u;
            un2 += cabs2(u);
            if (t != rp) amaxc_update(an, c, t);
        }
        for (int s = tid; s < n; s += NT) vn2 += cabs2(V0[s]);
    }
    block_reduce_c<NT>(red, un2, vn2, an);
    if (tid == 0) {
        rp_out[0] = rp;
        cp_out[0] = cp;
        ru[rp] = 1;
        cu[cp] = 1;
    }
    int rank = 1, status = 0;
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
        if (!aca_reserve_c<NT>(tab, rank, m, n, tcap, ws_top, ws_cap, &sh_off)) {
            status = 1;
            break;
        }
        double2 *Ur = ws + sh_off;
        double2 *Vr = Ur + m;
        const int nch = (rank + kChunkC - 1) / kChunkC;
        // residual row -> V column `rank`
        PointC pr = load_point_c(kp, rows[rp]);
        AMaxC am{-1.0, INT_MAX, cmk(0.0, 0.0)};
        vn2 = 0.0;
        for (int ch = 0; ch < nch; ++ch) {
            const int l0 = ch * kChunkC, l1 = min(rank, l0 + kChunkC);
            for (int q = l0 + tid; q < l1; q += NT) {
                coef[q - l0] = ws[tab[2 * q] + rp];
                cofs[q - l0] = tab[2 * q + 1];
            }
            __syncthreads();
            for (int s = tid; s < n; s += NT) {
                double2 r = ch == 0 ? kentry_c(kp, pr, cols[s]) : Vr[s];
                for (int q = l0; q < l1; ++q) r = csub(r, cmul(coef[q - l0], ws[cofs[q - l0] + s]));
                Vr[s] = r;
                if (ch == nch - 1) {
                    vn2 += cabs2(r);
                    if (!cu[s]) amaxc_update(am, r, s);
                }
            }
            __syncthreads();
        }
        double dummy = 0.0;
        block_reduce_c<NT>(red, vn2, dummy, am);
        if (am.idx == INT_MAX) break;
        cp = am.idx;
        piv = am.sv;
        // residual column -> U column `rank`
        PointC pc = load_point_c(kp, cols[cp]);
        an = AMaxC{-1.0, INT_MAX, cmk(0.0, 0.0)};
        un2 = 0.0;
        for (int ch = 0; ch < nch; ++ch) {
            const int l0 = ch * kChunkC, l1 = min(rank, l0 + kChunkC);
            for (int q = l0 + tid; q < l1; q += NT) {
                coef[q - l0] = ws[tab[2 * q + 1] + cp];
                cofs[q - l0] = tab[2 * q];
            }
            __syncthreads();
            for (int t = tid; t < m; t += NT) {
                double2 c = ch == 0 ? kentry_c(kp, pc, rows[t]) : Ur[t];
                for (int q = l0; q < l1; ++q) c = csub(c, cmul(coef[q - l0], ws[cofs[q - l0] + t]));
                if (ch == nch - 1) {
                    const double2 u = cdiv(c, piv);
                    Ur[t] = u;
                    un2 += cabs2(u);
                    if (t != rp && !ru[t]) amaxc_update(an, c, t);
                } else {
                    Ur[t] = c;
                }
            }
            __syncthreads();
        }
        block_reduce_c<NT>(red, un2, dummy, an);
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
        // normS update: |<U_q, u>||<V_q, v>| of the new term with every term
        double acc = 0.0;
        for (int q = wv; q < rank; q += NW) {
            const double2 *Uq = ws + tab[2 * q];
            const double2 *Vq = ws + tab[2 * q + 1];
            double2 du = cmk(0.0, 0.0), dv = cmk(0.0, 0.0);
            for (int t = lane; t < m; t += 64) du = cadd(du, cmulc(Uq[t], Ur[t]));
            for (int s = lane; s < n; s += 64) dv = cadd(dv, cmulc(Vq[s], Vr[s]));
            du.x = wsum(du.x);
            du.y = wsum(du.y);
            dv.x = wsum(dv.x);
            dv.y = wsum(dv.y);
            acc += cabsd(du) * cabsd(dv);
        }
        if (lane == 0) dacc[wv] = acc;
        __syncthreads();
        double S = 0.0;
        for (int q = 0; q < NW; ++q) S += dacc[q];
        const double uv = un * vn;
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

// L (unit lower) and R (upper) of the cross: L(i,j) = U_j[rb_i], R(j,i) = V_j[cb_i]
__global__ void k_lr_extract_c(const LrJob *__restrict__ jobs, const double2 *__restrict__ ws,
                               const int64_t *__restrict__ tab_all, const int32_t *__restrict__ rowpiv,
                               const int32_t *__restrict__ colpiv, double2 *__restrict__ out) {
    const LrJob jb = jobs[blockIdx.x];
    const int r = jb.rank;
    double2 *L = out + jb.out;
    double2 *R = L + (int64_t)r * r;
    const int64_t *tab = tab_all + jb.tab;
    for (int e = threadIdx.x; e < r * r; e += blockDim.x) {
        const int i = e % r, j = e / r;
        double2 lv = cmk(0.0, 0.0), rv = cmk(0.0, 0.0);
        if (i == j) lv = cmk(1.0, 0.0);
        else if (j < i) lv = ws[tab[2 * j] + rowpiv[jb.piv + i]];
        if (i <= j) rv = ws[tab[2 * i + 1] + colpiv[jb.piv + j]];
        L[e] = lv;
        R[e] = rv;
    }
}

// A = K(rows, cols)
__global__ void k_eval_c(const EvalJob *__restrict__ jobs, const int32_t *__restrict__ idx, double2 *__restrict__ A,
                         KParamsC kp) {
    const EvalJob jb = jobs[blockIdx.x];
    const int64_t tot = (int64_t)jb.m * jb.n;
    const int32_t *rows = idx + jb.rows;
    const int32_t *cols = idx + jb.cols;
    for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
        const int i = (int)(e % jb.m), j = (int)(e / jb.m);
        A[jb.a + (int64_t)j * jb.lda + i] = kentry2_c(kp, rows[i], cols[j]);
    }
}

// E = K(S, Ic) R^{-1} L^{-1} and E^T (plain transpose): the far L2P and the transpose of the vertex
// M2M = (L R)^{-T} K(Ic, S) (3D/codes/HODLR3Dtree.hpp:1926-1927, 2110-2133)
__global__ void k_xfer_c(const XferJob *__restrict__ jobs, const int32_t *__restrict__ idx,
                         const double2 *__restrict__ lr, double2 *__restrict__ A, KParamsC kp) {
    const XferJob jb = jobs[blockIdx.x];
    const int m = jb.m, r = jb.r;
    const int64_t lde = jb.lde, ldet = jb.ldet;
    const double2 *L = lr + jb.lr;
    const double2 *R = L + (int64_t)r * r;
    double2 *E = A + jb.e;
    double2 *ET = A + jb.et;
    const int32_t *S = idx + jb.s;
    const int32_t *Ic = idx + jb.ic;
    for (int t = threadIdx.x; t < m; t += blockDim.x) {
        PointC p = load_point_c(kp, S[t]);
        for (int k = 0; k < r; ++k) {
            double2 s = kentry_c(kp, p, Ic[k]);
            for (int q = 0; q < k; ++q) s = csub(s, cmul(E[q * lde + t], R[(int64_t)k * r + q]));
            E[k * lde + t] = cdiv(s, R[(int64_t)k * r + k]);
        }
        for (int k = r - 1; k >= 0; --k) {
            double2 s = E[k * lde + t];
            for (int q = k + 1; q < r; ++q) s = csub(s, cmul(E[q * lde + t], L[(int64_t)k * r + q]));
            E[k * lde + t] = cdiv(s, L[(int64_t)k * r + k]);
        }
        for (int k = 0; k < r; ++k) ET[t * ldet + k] = E[k * lde + t];
    }
}

// A = R^{-1} L^{-1} K(Ir, V): the semi-nested row factor (3D/codes/HODLR3Dtree.hpp:2341-2344)
__global__ void k_cur_c(const CurJob *__restrict__ jobs, const int32_t *__restrict__ idx,
                        const double2 *__restrict__ lr, double2 *__restrict__ A, KParamsC kp) {
    const CurJob jb = jobs[blockIdx.x];
    const int r = jb.r, n = jb.n;
    const double2 *L = lr + jb.lr;
    const double2 *R = L + (int64_t)r * r;
    const int32_t *Ir = idx + jb.ir;
    const int32_t *V = idx + jb.v;
    double2 *Acol = A + jb.a;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        double2 *col = Acol + (int64_t)c * jb.lda;
        PointC pv = load_point_c(kp, V[c]);
        for (int k = 0; k < r; ++k) {  // L y = K(Ir, v)  (unit lower)
            double2 s = kentry_c(kp, pv, Ir[k]);
            for (int q = 0; q < k; ++q) s = csub(s, cmul(L[(int64_t)q * r + k], col[q]));
            col[k] = cdiv(s, L[(int64_t)k * r + k]);
        }
        for (int k = r - 1; k >= 0; --k) {  // R z = y  (upper)
            double2 s = col[k];
            for (int q = k + 1; q < r; ++q) s = csub(s, cmul(R[(int64_t)q * r + k], col[q]));
            col[k] = cdiv(s, R[(int64_t)k * r + k]);
        }
    }
}

// ------------------------------------------------------------------ complex batched GEMV
// The real path's work lists over complex pools: item rows y (+)= sum_t A_t x_t; symmetric
// terms (t >= 0) also add A_t^T x_item to vec[t ..] (plain transpose, atomics per real part).
// One workgroup per item; lane = row (rows > 64: two passes), 16-byte loads of whole complex
// entries.  Wave wv takes groups of kCU consecutive columns (wv, wv + 4, ...), all kCU columns'
// loads issued before they are used.  The transposed sums of a group are reduced together: a
// butterfly halves the kCU values per exchange (lane bit o keeps one half, sends the other), so
// log2(64) exchanges serve kCU columns and lanes 0 .. kCU-1 end with one column sum each.  Role 2
// items and atomic items (flags bit 0 or 2) add their rows atomically.
__device__ __forceinline__ void catomic_add(double2 *p, double2 v) {
    double *d = reinterpret_cast<double *>(p);
    atomicAdd(d, v.x);
    atomicAdd(d + 1, v.y);
}
__device__ __forceinline__ double2 cshfl_xor(double2 v, int o) {
    return cmk(__shfl_xor(v.x, o, 64), __shfl_xor(v.y, o, 64));
}
constexpr int kCU = 8;  // columns in flight per wave
__global__ __launch_bounds__(256) void k_cgemv(const GItem *__restrict__ items, const GTerm *__restrict__ terms,
                                               const double2 *__restrict__ A, double2 *__restrict__ vec) {
    const GItem it = items[blockIdx.x];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __shared__ double2 part[4][kRowTile];
    __shared__ double2 xb[kRowTile];
    const int rows = it.rows;
    // (zero beyond the item's rows: the transposed sums then need no row test)
    for (int r = threadIdx.x; r < kRowTile; r += 256) xb[r] = r < rows ? vec[it.xb + r] : cmk(0.0, 0.0);
    __syncthreads();
    const bool two = rows > 64;
    const bool in0 = lane < rows, in1 = two && 64 + lane < rows;
    const double2 xb0 = xb[lane], xb1 = xb[64 + lane];
    double2 acc0 = cmk(0.0, 0.0), acc1 = cmk(0.0, 0.0);
    for (int tt = it.t0; tt < it.t1; ++tt) {
        const GTerm tm = terms[tt];
        const double2 *Ab = A + tm.a + it.row0;
        for (int c0 = kCU * wv; c0 < tm.cols; c0 += 4 * kCU) {
            double2 v0[kCU], v1[kCU], xv[kCU];
#pragma unroll
            for (int u = 0; u < kCU; ++u) {
                const int c = c0 + u;
                const bool ok = c < tm.cols;
                xv[u] = ok ? vec[tm.x + c] : cmk(0.0, 0.0);
                const double2 *col = Ab + (int64_t)(ok ? c : c0) * it.lda;
                v0[u] = ok && in0 ? col[lane] : cmk(0.0, 0.0);
                v1[u] = ok && in1 ? col[64 + lane] : cmk(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < kCU; ++u) {
                acc0 = cadd(acc0, cmul(v0[u], xv[u]));
                if (two) acc1 = cadd(acc1, cmul(v1[u], xv[u]));
            }
            if (tm.t >= 0) {
                double2 d[kCU];
#pragma unroll
                for (int u = 0; u < kCU; ++u) d[u] = cadd(cmul(v0[u], xb0), cmul(v1[u], xb1));
                int ui = 0;
#pragma unroll
                for (int k = kCU, o = 1; k > 1; k >>= 1, o <<= 1) {
                    const bool hi = lane & o;
#pragma unroll
                    for (int i = 0; i < k / 2; ++i) {
                        const double2 snd = hi ? d[i] : d[i + k / 2];
                        const double2 keep = hi ? d[i + k / 2] : d[i];
                        d[i] = cadd(keep, cshfl_xor(snd, o));
                    }
                    ui += hi ? k / 2 : 0;
                }
#pragma unroll
                for (int o = kCU; o < 64; o <<= 1) d[0] = cadd(d[0], cshfl_xor(d[0], o));
                if (lane < kCU && c0 + ui < tm.cols) catomic_add(vec + tm.t + c0 + ui, d[0]);
            }
        }
    }
    part[wv][lane] = acc0;
    part[wv][64 + lane] = acc1;
    __syncthreads();
    for (int r = threadIdx.x; r < rows; r += 256) {
        const double2 sum = cadd(cadd(part[0][r], part[1][r]), cadd(part[2][r], part[3][r]));
        double2 *y = vec + it.y + r;
        if (it.flags & 5) catomic_add(y, sum);
        else *y = cadd(*y, sum);
    }
}

// ------------------------------------------------------------------ vectors
__global__ void k_cfill(double2 *a, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = cmk(0.0, 0.0);
}
__global__ void k_cpermute_in(const double2 *x, double2 *xt, const int32_t *perm, int64_t N) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < N; t += (int64_t)gridDim.x * blockDim.x)
        xt[t] = x[perm[t]];
}
__global__ void k_cpermute_out(const double2 *yt, double2 *y, const int32_t *perm, int64_t N) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < N; t += (int64_t)gridDim.x * blockDim.x)
        y[perm[t]] = yt[t];
}
// dense rows of A x with entries on the fly (getTruePoten, 3D/codes/HODLR3Dtree.hpp:2630-2643)
__global__ __launch_bounds__(256) void k_cdense_rows(const int32_t *__restrict__ rows, int64_t nrows,
                                                     const double2 *__restrict__ xt, double2 *__restrict__ y,
                                                     int64_t N, KParamsC kp) {
    __shared__ double sx[256], sy[256], sz[256];
    __shared__ double2 sq[256];
    __shared__ double2 part[4][64];
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int64_t r = blockIdx.x * 64LL + lane;
    const bool active = r < nrows;
    const int i = active ? (rows ? rows[r] : (int)r) : 0;
    PointC p = load_point_c(kp, i);
    double2 acc = cmk(0.0, 0.0);
    for (int64_t j0 = 0; j0 < N; j0 += 256) {
        const int64_t j = j0 + threadIdx.x;
        __syncthreads();
        if (j < N) {
            sx[threadIdx.x] = kp.px[j];
            sy[threadIdx.x] = kp.py[j];
            sz[threadIdx.x] = kp.dim > 2 ? kp.pz[j] : 0.0;
            sq[threadIdx.x] = xt[j];
        }
        __syncthreads();
        const int nj = (int)min<int64_t>(256, N - j0);
        for (int q = g; q < nj; q += 4) {
            const int jj = (int)(j0 + q);
            const double2 v = (jj == i) ? cmk(kp.diag, kp.diag_im) : koff_c(kp.dim, p.x - sx[q], p.y - sy[q], p.z - sz[q]);
            acc = cadd(acc, cmul(v, sq[q]));
        }
    }
    part[g][lane] = acc;
    __syncthreads();
    if (g == 0 && active) y[r] = cadd(cadd(part[0][lane], part[1][lane]), cadd(part[2][lane], part[3][lane]));
}

// ------------------------------------------------------------------ BLAS-1 / 2 for the complex Arnoldi
constexpr int kDotC = 256;
// partial[l * nblk + b] = block b's share of V_l^H w
__global__ __launch_bounds__(kDotC) void k_cmdot(const double2 *__restrict__ V, int64_t ldv,
                                                 const double2 *__restrict__ w, int64_t n,
                                                 double2 *__restrict__ partial) {
    const int b = blockIdx.x, l = blockIdx.y, nblk = gridDim.x;
    const double2 *v = V + (int64_t)l * ldv;
    double2 acc = cmk(0.0, 0.0);
    for (int64_t i = (int64_t)b * kDotC + threadIdx.x; i < n; i += (int64_t)nblk * kDotC) acc = cadd(acc, cmulc(v[i], w[i]));
    acc.x = wsum(acc.x);
    acc.y = wsum(acc.y);
    __shared__ double2 sh[kDotC / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 s = cmk(0.0, 0.0);
        for (int q = 0; q < kDotC / 64; ++q) s = cadd(s, sh[q]);
        partial[(int64_t)l * nblk + b] = s;
    }
}
__global__ void k_cmdot_reduce(const double2 *__restrict__ partial, int nblk, double2 *__restrict__ out) {
    const int l = blockIdx.x;
    double2 acc = cmk(0.0, 0.0);
    for (int b = threadIdx.x; b < nblk; b += 64) acc = cadd(acc, partial[(int64_t)l * nblk + b]);
    acc.x = wsum(acc.x);
    acc.y = wsum(acc.y);
    if (threadIdx.x == 0) out[l] = acc;
}
// w += sign * V h
__global__ void k_cmgemv(const double2 *__restrict__ V, int64_t ldv, int k, const double2 *__restrict__ h,
                         double2 *__restrict__ w, int64_t n, double sign) {
    __shared__ double2 hs[512];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i - threadIdx.x < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double2 acc = cmk(0.0, 0.0);
        for (int l0 = 0; l0 < k; l0 += 512) {
            const int l1 = min(k, l0 + 512);
            __syncthreads();
            for (int l = l0 + threadIdx.x; l < l1; l += blockDim.x) hs[l - l0] = h[l];
            __syncthreads();
            if (i < n)
                for (int l = l0; l < l1; ++l) acc = cadd(acc, cmul(V[(int64_t)l * ldv + i], hs[l - l0]));
        }
        if (i < n) w[i] = cmk(w[i].x + sign * acc.x, w[i].y + sign * acc.y);
    }
}
// y = a x + b y (complex a, b; b = 0 writes y = a x)
__global__ void k_caxpby(double2 a, const double2 *__restrict__ x, double2 b, double2 *__restrict__ y, int64_t n) {
    const bool zb = b.x == 0.0 && b.y == 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = zb ? cmul(a, x[i]) : cadd(cmul(a, x[i]), cmul(b, y[i]));
}

inline int cgrid(int64_t n, int nt = 256, int cap = 8192) {
    int64_t g = (n + nt - 1) / nt;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace

// ------------------------------------------------------------------ launch wrappers
void launch_aca_c(const AcaJob *jobs, int njobs, const int32_t *idx, double2 *ws, int64_t ws_cap,
                  unsigned long long *ws_top, int64_t *tab, uint8_t *flags, int32_t *rank, int32_t *status,
                  int32_t *rowpiv, int32_t *colpiv, KParamsC kp, double tol, double *hist, int big, hipStream_t s) {
    if (njobs <= 0) return;
    if (big)
        hipLaunchKernelGGL(k_aca_c<1024>, dim3(njobs), dim3(1024), 0, s, jobs, idx, ws, ws_cap, ws_top, tab, flags, rank,
                           status, rowpiv, colpiv, kp, tol, hist);
    else
        hipLaunchKernelGGL(k_aca_c<256>, dim3(njobs), dim3(256), 0, s, jobs, idx, ws, ws_cap, ws_top, tab, flags, rank,
                           status, rowpiv, colpiv, kp, tol, hist);
}
void launch_lr_extract_c(const LrJob *jobs, int njobs, const double2 *ws, const int64_t *tab, const int32_t *rowpiv,
                         const int32_t *colpiv, double2 *out, hipStream_t s) {
    if (njobs > 0) hipLaunchKernelGGL(k_lr_extract_c, dim3(njobs), dim3(256), 0, s, jobs, ws, tab, rowpiv, colpiv, out);
}
void launch_eval_c(const EvalJob *jobs, int njobs, const int32_t *idx, double2 *A, KParamsC kp, hipStream_t s) {
    if (njobs > 0) hipLaunchKernelGGL(k_eval_c, dim3(njobs), dim3(256), 0, s, jobs, idx, A, kp);
}
void launch_xfer_c(const XferJob *jobs, int njobs, const int32_t *idx, const double2 *lr, double2 *A, KParamsC kp,
                   hipStream_t s) {
    if (njobs > 0) hipLaunchKernelGGL(k_xfer_c, dim3(njobs), dim3(64), 0, s, jobs, idx, lr, A, kp);
}
void launch_cur_c(const CurJob *jobs, int njobs, const int32_t *idx, const double2 *lr, double2 *A, KParamsC kp,
                  hipStream_t s) {
    if (njobs > 0) hipLaunchKernelGGL(k_cur_c, dim3(njobs), dim3(64), 0, s, jobs, idx, lr, A, kp);
}
void launch_cgemv(const GItem *items, int nitems, const GTerm *terms, const double2 *A, double2 *vec, hipStream_t s) {
    if (nitems > 0) hipLaunchKernelGGL(k_cgemv, dim3(nitems), dim3(256), 0, s, items, terms, A, vec);
}
void launch_cfill(double2 *a, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_cfill, dim3(cgrid(n)), dim3(256), 0, s, a, n);
}
void launch_cpermute_in(const double2 *x, double2 *xt, const int32_t *perm, int64_t N, hipStream_t s) {
    hipLaunchKernelGGL(k_cpermute_in, dim3(cgrid(N, 256, 2048)), dim3(256), 0, s, x, xt, perm, N);
}
void launch_cpermute_out(const double2 *yt, double2 *y, const int32_t *perm, int64_t N, hipStream_t s) {
    hipLaunchKernelGGL(k_cpermute_out, dim3(cgrid(N, 256, 2048)), dim3(256), 0, s, yt, y, perm, N);
}
void launch_cdense_rows(const int32_t *rows, int64_t nrows, const double2 *xt, double2 *y, int64_t N, KParamsC kp,
                        hipStream_t s) {
    if (nrows > 0)
        hipLaunchKernelGGL(k_cdense_rows, dim3((unsigned)((nrows + 63) / 64)), dim3(256), 0, s, rows, nrows, xt, y, N, kp);
}
int cmdot_blocks(int64_t n) {
    int64_t b = (n + kDotC * 8 - 1) / (kDotC * 8);
    return (int)std::max<int64_t>(1, std::min<int64_t>(512, b));
}
void launch_cmdot(const double2 *V, int64_t ldv, int k, const double2 *w, int64_t n, double2 *partial, int nblk,
                  hipStream_t s) {
    if (k > 0) hipLaunchKernelGGL(k_cmdot, dim3(nblk, k), dim3(kDotC), 0, s, V, ldv, w, n, partial);
}
void launch_cmdot_reduce(const double2 *partial, int nblk, int k, double2 *out, hipStream_t s) {
    if (k > 0) hipLaunchKernelGGL(k_cmdot_reduce, dim3(k), dim3(64), 0, s, partial, nblk, out);
}
void launch_cmgemv(const double2 *V, int64_t ldv, int k, const double2 *h, double2 *w, int64_t n, double sign,
                   hipStream_t s) {
    if (k > 0) hipLaunchKernelGGL(k_cmgemv, dim3(cgrid(n, 256, 4096)), dim3(256), 0, s, V, ldv, k, h, w, n, sign);
}
void launch_caxpby(double2 a, const double2 *x, double2 b, double2 *y, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_caxpby, dim3(cgrid(n)), dim3(256), 0, s, a, x, b, y, n);
}

}  // namespace h2w
