// h2w_handle.hpp -- the library handle and the host-side types shared by the
// translation units of libh2w (assembly and single-vector runtime in
// h2w_ops.hip, kNR-wide and batched runtime in multi.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <chrono>
#include <functional>
#include <cmath>
#include <string>
#include <vector>

#include "h2w_internal.hpp"

namespace h2w {


constexpr double kPI = 3.141592653589793238462643383279502884;

inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A HIP event owned by the scope: created at construction, destroyed on every exit path.
struct Event {
    hipEvent_t e = nullptr;
    Event() { H2W_HIP(hipEventCreate(&e)); }
    ~Event() {
        if (e) (void)hipEventDestroy(e);
    }
    Event(const Event &) = delete;
    Event &operator=(const Event &) = delete;
    Event(Event &&o) noexcept : e(o.e) { o.e = nullptr; }
    operator hipEvent_t() const { return e; }
};
// milliseconds between two recorded events (waits for the second)
inline double elapsed_ms(const Event &a, const Event &b) {
    H2W_HIP(hipEventSynchronize(b));
    float m = 0.f;
    H2W_HIP(hipEventElapsedTime(&m, a, b));
    return m;
}

// Page-locked host buffer (hipHostMalloc) for the small per-iteration read-backs of GMRES: a
// device-to-host copy into it is one DMA transfer, without the runtime's staging copy.
struct PinBuf {
    double *p = nullptr;
    size_t n = 0;
    PinBuf() {}
    PinBuf(const PinBuf &) = delete;
    PinBuf &operator=(const PinBuf &) = delete;
    ~PinBuf() { release(); }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
    double *ensure(size_t cnt) {
        if (cnt > n) {
            release();
            if (hipHostMalloc((void **)&p, cnt * sizeof(double), hipHostMallocDefault) != hipSuccess) {
                p = nullptr;
                throw Error{H2W_ENOMEM, "hipHostMalloc of " + std::to_string(cnt * sizeof(double)) + " bytes failed"};
            }
            n = cnt;
        }
        return p;
    }
};

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() {}
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t cnt) {
        release();
        if (cnt == 0) return;
        hipError_t e = hipMalloc(&p, cnt * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error{H2W_ENOMEM, "hipMalloc of " + std::to_string(cnt * sizeof(T)) + " bytes failed"};
        }
        n = cnt;
    }
    void ensure(size_t cnt) {
        if (cnt > n) alloc(cnt);
    }
    void upload(const std::vector<T> &v, hipStream_t s) {
        ensure(v.size());
        if (!v.empty()) H2W_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

inline int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// one ACA problem: rows/cols are tree-order point indices
constexpr int kRankFeatures = 16;  // geometry features of one ACA block (rank predictor input)
constexpr int kRankHidden = 64;    // hidden width of the rank predictor (three hidden layers)
// parameters: W1 (F x H), b1, W2 (H x H), b2, W3 (H x H), b3, w4 (H), b4 -- row-major, BF16
constexpr int kRankParams = kRankFeatures * kRankHidden + kRankHidden + 2 * (kRankHidden * kRankHidden + kRankHidden) +
                            kRankHidden + 1;

struct AcaBlock {
    std::vector<int32_t> rows, cols;
    std::array<float, kRankFeatures> feat{};
    bool has_feat = false;
    int rank = 0;
    std::vector<int32_t> rb, cb;  // local pivot indices
    std::vector<double> lr;       // L then R, rank x rank column-major each
    std::vector<double> hist;     // per step: |u||v| and normS after the step
};

// skeleton data of one family at one level
struct FamLevel {
    int nb = 0;
    std::vector<int> rank;
    std::vector<std::vector<int32_t>> Ir, Ic;  // tree-order indices
    std::vector<std::vector<double>> lr;
    std::vector<std::vector<double>> hist;     // ACA history (|u||v|, normS) per step
    std::vector<int> kin;                      // rank of the inner (truncated) operator
    std::vector<int64_t> ein_off, eint_off;    // inner transfers in the FP32 pool
    std::vector<std::vector<int32_t>> S;       // transfer row set (leaf: box points; else children's Ir)
    std::vector<int64_t> o_off, u_off;         // vec pool
    std::vector<int64_t> e_off, et_off;        // mat pool
};

struct Pair {                                  // semi-nested vertex pair (B <- V) at a level
    int level, b, v;
    int rank;
    std::vector<int32_t> Ir, Ic;
    std::vector<double> lr;
    int64_t ac_off, ar_off, t_off;
};

// Position of a launch in the matvec: stage 0 up pass (step = levels above the
// leaves), 1 M2L + near field, 2 down pass (step = levels above the leaves),
// 3 semi-nested column factors.  Batches of systems of different depths merge
// their launches by (stage, step): up steps from the leaves, down steps ending
// at the leaves.
struct Launch {
    int64_t item0, nitems, term0, nterms;
    int role;      // 0 transfer pass, 1 M2L + near field, 2 semi-nested column factors,
                   // 3 transposed products of the symmetric blocks (16-wide lists)
    int stage = 0, step = 0;
    int unr = 4;   // columns in flight per wave (autotuned), FP64 / inner FP32 lists
    int unr32 = 4; // the same for the stored FP32 copy
    int avg_cols = 0;  // mean inner size of the terms (picks the transposed 16-wide kernel's slabs)
    int64_t avg_item_elems = 0;  // mean operator elements per item
};
// one item per wave on the 16-wide path: launches of enough items that are small on average
// (the transfer passes at the leaves of small systems), where a workgroup's set-up and the
// reduction across its waves would outweigh the item
inline bool per_wave_items(const Launch &L) {
    return L.nitems >= 2048 && L.avg_item_elems > 0 && L.avg_item_elems <= 4096;
}
// narrow slabs for the transposed 16-wide products when the blocks have few rows
inline int trans_narrow(const Launch &L, bool fp32) { return L.avg_cols <= (fp32 ? 32 : 16); }

struct WorkList {
    std::vector<GItem> items;
    std::vector<GTerm> terms;
    int role = 0;
    int stage = 0, step = 0;
};

// Launch lists as uploaded to the device
struct LaunchSet {
    std::vector<Launch> launches;
    std::vector<int64_t> elems;
    std::vector<GItem> h_items;  // host copies (terms ranges absolute), for merging into batches
    std::vector<GTerm> h_terms;
    DevBuf<GItem> d_items;
    DevBuf<GTerm> d_terms;
};

struct TermSpec {
    int64_t a, x;
    int32_t cols;
    int64_t t = -1;  // symmetric term: output offset of its transposed product (the partner box's y)
    int32_t ld = 0;  // transposed term: leading dimension of the stored block
};

struct M2LBlk {  // one stored M2L block K(Ir_B, Ir_F): family, level, box, partner, pool offset
    int f, j, b, p;
    int64_t off;
};
struct NearBlk {  // one stored near-field block K(P(B), P(N))
    int b, n;
    int64_t off;
};


inline void gen_rot(double dx, double dy, double &cs, double &sn) {
    if (dy == 0.0) {
        cs = 1.0;
        sn = 0.0;
    } else if (std::fabs(dy) > std::fabs(dx)) {
        double t = dx / dy;
        sn = 1.0 / std::sqrt(1.0 + t * t);
        cs = t * sn;
    } else {
        double t = dy / dx;
        cs = 1.0 / std::sqrt(1.0 + t * t);
        sn = t * cs;
    }
}
inline void app_rot(double &dx, double &dy, double cs, double sn) {
    double t = cs * dx + sn * dy;
    dy = -sn * dx + cs * dy;
    dx = t;
}

// device work buffers of a lockstep solve
struct LockBufs {
    DevBuf<double> V, w, r, d, b, part, h, sc;
    DevBuf<double> spart, sh;  // shared-space dot products: per-block partials, reduced (2 passes)
    PinBuf pin;  // read-backs of the Hessenberg columns
    int64_t Vcols = 0;
    DevBuf<double> U, coef;  // recycled refinement space (U; C sits in the first rk basis columns)
    int rk = 0;
};

}  // namespace h2w

struct h2w_handle {
    int dim = 3, variant = 0, choice = 0, tol_pow = 10, n_max = 64;
    int64_t N = 0;
    double L = 1.0;
    h2w_opts opts;
    h2w::Plan plan;
    bool assembled = false;
    double diag = 0.0;
    double memory = 0.0;
    double assembly_seconds = 0.0;
    int aca_reruns = 0;
    int aca_batches = 0;                // k_aca launch batches of the last assembly
    int64_t aca_pool_bytes = 0;         // largest ACA workspace of the last assembly
    h2w::DevBuf<uint16_t> rank_model;   // BF16 rank-predictor parameters (empty: heuristic estimates)
    int rank_pred_blocks = 0, rank_pred_under = 0;  // predicted blocks / blocks whose rank exceeded the estimate
    double rank_pred_abs_log2 = 0.0;    // sum of |log2(1 + predicted) - log2(1 + rank)|
    int rank_fix_blocks = 0;            // wide blocks that ran a fixed-rank phase
    int rank_fix_resumed = 0;           // of those, blocks the stopping rule sent on adaptively
    int64_t rank_fix_extra_steps = 0;   // pivot steps taken beyond where the stopping rule ended a block
    double rank_fix_sampled_err = 0.0;  // largest sampled-entry relative error of those blocks
    std::vector<std::array<float, h2w::kRankFeatures>> sample_feat;  // harvest: every ACA block of the
    std::vector<int32_t> sample_rank;                                 // last assembly with its rank
    hipStream_t stream = nullptr;
    h2w::KParams kp;
    bool cplx = false;                  // complex kernel (choice 2 / 10): complex pools, vectors, GMRES
    double diag_im = 0.0;               // imaginary part of the diagonal value
    h2w::KParamsC kpc;
    h2w::DevBuf<double> d_px, d_py, d_pz;
    h2w::DevBuf<int32_t> d_perm;
    h2w::DevBuf<int32_t> d_iperm;          // input index -> tree index
    h2w::DevBuf<double> mat64;
    h2w::DevBuf<float> mat32;
    int64_t mat_len = 0;
    h2w::DevBuf<double> vec64;
    h2w::DevBuf<float> vec32;
    h2w::DevBuf<double> vec64m;            // the vector pool with kNR right-hand sides interleaved
    h2w::DevBuf<float> vec32m;
    h2w::DevBuf<float> sx32, sy32, sz32;   // slot coordinates (FP32)
    h2w::DevBuf<float> sxl, syl, szl;      // FP32 remainders of the slot coordinates
    h2w::DevBuf<double> sx64, sy64, sz64;  // slot coordinates (FP64)
    h2w::DevBuf<int32_t> spid;             // slot point ids
    h2w::SlotPoints sp;
    int64_t vec_len = 0, off_x = 0, off_y = 0;
    std::vector<h2w::FamLevel> fam[2];
    std::vector<h2w::Pair> pairs;
    std::vector<h2w::Launch> launches;
    std::vector<int64_t> launch_elems;  // operator elements each launch streams
    std::vector<h2w::M2LBlk> m2l;
    std::vector<h2w::NearBlk> nearb;
    h2w::DevBuf<float> mat32in;         // the inner operator's transfers and compact M2L blocks
    h2w::LaunchSet inner;               // the refinement's inner operator (ranks truncated)
    h2w::LaunchSet trans, inner_trans;  // 16-wide path: K(b, p)^T x_b of the stored blocks (role 3)
    h2w::LaunchSet otf;                 // on-the-fly path: the M2L + near-field items, every direction
    hipGraphExec_t graph[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // one per precision
    hipGraphExec_t graph_m[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // the same, kNR right-hand sides
    bool use_graph = true;
    bool sym16 = true;                  // 16-wide fused launch reads each block once (k_bgemm_sym)
    int64_t inner_elems = 0;
    int inner_gen = 0;                  // incremented whenever the inner operator is rebuilt
    double ir_exp = 0.0;                // the inner operator's tolerance exponent p (10^-p), 0: none
    h2w::DevBuf<h2w::GItem> d_items;
    h2w::DevBuf<h2w::GTerm> d_terms;
    std::vector<h2w::GItem> h_items;    // host copies of the main lists
    std::vector<h2w::GTerm> h_terms;
    int64_t stored_elems = 0;
    // staging for host-pointer calls and GMRES
    h2w::DevBuf<double> d_stage_in, d_stage_out, d_dense;
    h2w::DevBuf<double> g_V, g_w, g_r, g_b, g_x, g_h, g_part, g_d;
    h2w::DevBuf<double> g_U, g_coef;  // recycled refinement space: U (k columns), combination coefficients
    // deflation space built once for an inner precision (h2w_set_deflation): U (k columns), then
    // C = A U (k orthonormal columns); every correction of the refinement starts from it
    h2w::DevBuf<double> g_defl;
    h2w::DevBuf<double> g_defl_rm;      // the same row-major (N x k each: U then C), for the lockstep solves
    int defl_k = 0, defl_prec = -1, defl_gen = -1;
    int defl_serial = 0;                // incremented whenever the space is rebuilt or dropped
    int gv_defl = -1;                   // serial of the space whose U, C sit in g_U / g_V's first columns
    h2w::PinBuf g_pin;                // GMRES read-backs (Hessenberg columns)
    int64_t g_Vcols = 0;
    // kNR-wide lockstep GMRES (interleaved N x kNR vectors)
    h2w::LockBufs mbuf;
    h2w::DevBuf<int2> m_blk;      // dot-product blocks of the one-system lockstep vector
    h2w::DevBuf<int32_t> m_sblk;
    int m_nblk = 0;
};

namespace h2w {

using Handle = h2w_handle;

// assembly from skeletons already in h.fam / h.pairs (h2w_ops.hip)
void rebuild_operators(Handle &h);
// stream and device copies of the points of a handle created with plan_only (h2w_ops.hip)
void attach_device(Handle &h);
// device address of the per-GPU staging array of h2w_dense_matvec_noalloc, looked up once per
// device with that device current (noalloc.hip)
void noalloc_setup(int device);

// matvec runtime (h2w_ops.hip)
bool is32(int precision);
const std::vector<Launch> &launches_for(Handle &h, int precision);
void drop_graphs(Handle &h);
void drop_graph(Handle &h, int precision);
void matvec_dev(Handle &h, const double *x, double *y, int precision, bool tree);

// kNR-wide runtime (multi.hip)
// Y = A X for nrhs <= kNR columns (column stride ldx / ldy), device pointers; tree = tree order
void matvec_multi_dev(Handle &h, const double *X, int64_t ldx, double *Y, int64_t ldy, int nrhs, int precision,
                      bool tree);
// nrhs <= kNR GMRES solves / FP64 refinements on interleaved tree-order B and X (device)
void solve_tree_m(Handle &h, const double *B, double *X, int nrhs, int max_iter, double tol, int mode,
                  h2w_gmres_stats *st);

// Lockstep solves over S systems x kNR columns on a lockstep vector (n slots, the
// systems' interleaved tree-order vectors one after another).  Column (s, r) is
// right-hand side r of system s.
struct Lockstep {
    hipStream_t s = nullptr;
    int64_t n = 0;                    // slots
    int S = 1;                        // systems
    const int32_t *seg = nullptr;     // system of every slot (device; NULL: S == 1)
    const int2 *blk = nullptr;        // dot-product blocks (device): [first, one-past-last) slot
    int nblk = 0;
    const int32_t *sblk = nullptr;    // blocks of system s: sblk[s] .. sblk[s+1]-1 (device)
    std::vector<double> ir_floor;     // per system: 10^-p of its inner operator
    std::function<void(const double *, double *, int)> matvec;  // Y = A X, interleaved, a precision
    LockBufs *buf = nullptr;
    // deflation spaces shared by the columns of each system: row-major U and C (n x dk, every system's
    // slots holding its own space, zero-padded to dk), built for the inner precision dprec; every
    // correction of the refinement starts from them
    const double *dU = nullptr, *dC = nullptr;
    int dk = 0, dprec = -1;
};
// Split S consecutive slot ranges (system s: [start[s], start[s+1])) into dot-product blocks.
void lock_blocks(const std::vector<int64_t> &start, std::vector<int2> &blk, std::vector<int32_t> &sblk);
// nrhs <= kNR right-hand sides per system: B, X lockstep vectors (device); st: S x nrhs,
// entry (s, r) at s * nrhs + r.  Modes H2W_GMRES_FP64, _FP32_IR, _FP32_INNER_IR.
void lock_solve(Lockstep &L, const double *B, double *X, int nrhs, int max_iter, double tol, int mode,
                h2w_gmres_stats *st);

}  // namespace h2w
