// h2w_internal.hpp -- shared declarations of libh2w (host plan, device work
// descriptors, launch wrappers).  Not part of the public C ABI (include/h2w.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <array>
#include <vector>

#include "../../include/h2w.h"

namespace h2w {

// ------------------------------------------------------------------ errors
void set_error(const std::string &msg);
struct Error {
    int code;
    std::string msg;
};
#define H2W_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::h2w::Error{H2W_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

// ------------------------------------------------------------------ kernel parameters
// Matrix entry A(i,j) of the reference's userkernel (*/codes/kernel.hpp); i, j are
// tree-order point indices, i == j takes the diagonal rule.
struct KParams {
    int dim;
    int choice;
    double diag;
    const double *px, *py, *pz;  // tree-order coordinates (device)
};

// The same for the complex kernels (choice 2: 3D Helmholtz, 10: 2D Hankel); the diagonal
// value is complex (diag + i diag_im).
struct KParamsC {
    int dim;
    int choice;
    double diag, diag_im;
    const double *px, *py, *pz;
};

// Coordinates and point id of the point each vector-pool slot belongs to
// (x/y slots: the tree-order points; o/u slots: the skeleton points), so that
// near-field and M2L blocks can be evaluated on the fly from the work lists.
struct SlotPoints {
    const float *x, *y, *z;          // FP32 coordinates (rounded)
    const float *xl, *yl, *zl;       // FP32 remainders: x = x + xl to ~2^-48 relative
    const double *xd, *yd, *zd;      // FP64 coordinates
    const int32_t *pid;              // tree-order point index, -1 for slots without a point
    int dim, choice;
    double diag;
};

// ------------------------------------------------------------------ host plan
// Tree, tree-order permutation and interaction lists, built from integer box
// coordinates (the reference builds them from hand-written child tables,
// 3D/codes/HODLR3Dtree.hpp:155-1680; the lists and their order are the same).
struct Plan {
    int dim = 3, nchild = 8, nL = 1;
    int64_t N = 0;
    double L = 1.0;
    std::vector<int32_t> perm;                          // tree index -> input index
    std::vector<double> px, py, pz;                     // tree-order coordinates
    std::vector<int> nboxes;                            // per level
    std::vector<std::vector<int64_t>> bstart, bcount;   // [level][box] point range (tree order)
    std::vector<std::vector<std::vector<int32_t>>> il;  // [level][box] interaction list
    std::vector<std::vector<std::vector<uint8_t>>> ilv; // vertex flags of il entries
    std::vector<std::vector<int32_t>> near;             // [leaf box] near neighbours (slot order)
    std::vector<std::vector<std::array<int, 3>>> icoord; // [level][box] integer box coordinates
    std::vector<std::vector<std::array<float, 6>>> bbox; // [level][box] bounding box of its points (min, max);
                                                         // empty boxes: min > max

    void build(int dim, int64_t N, const double *pts_soa, double L, int nlevels);
    std::vector<int32_t> vert_of(int j, int b) const;
    std::vector<int32_t> far_of(int j, int b) const;
};

// ------------------------------------------------------------------ device descriptors
// One ACA problem.  Its U (m) and V (n) columns are carved out of a shared
// workspace pool by a device-side bump allocator, one column pair per step;
// the pool offsets of column pair l are kept at tab[2l] (U) and tab[2l+1] (V).
struct AcaJob {
    int64_t rows;  // offset into the index array (int32 entries)
    int64_t cols;
    int32_t m, n;
    int32_t tcap;  // capacity of the column table (max rank)
    int32_t pad;
    int64_t tab;   // offset (int64 entries) of the column table, 2*tcap entries
    int64_t wsF;   // offset (bytes) of the used-row/used-col flags (m + n bytes)
    int64_t piv;   // offset into rowpiv/colpiv (tcap entries each)
};

struct LrJob {     // copy L (unit lower) and R (upper) out of an ACA workspace
    int64_t tab, piv;
    int32_t m, n, rank, pad;
    int64_t out;   // offset (doubles): L at out, R at out + rank*rank, both column-major
};

// Stored blocks are column-major with the leading dimension padded to a multiple
// of 4 elements (zero rows), so that every column starts 16-byte aligned in both
// the FP64 pool and its FP32 shadow.
inline int64_t pad_ld(int64_t rows) { return (rows + 3) / 4 * 4; }

struct EvalJob {   // A = K(rows, cols), m x n column-major at a, leading dimension lda
    int64_t rows, cols, a;
    int32_t m, n, lda, pad;
};

struct XferJob {   // E = K(S, Ic) (L R)^{-1}, m x r at e (ld lde); E^T (r x m) at et (ld ldet)
    int64_t s, ic, lr, e, et;
    int32_t m, r, lde, ldet;
};

struct CurJob {    // A = (L R)^{-1} K(Ir, V), r x n column-major at a, leading dimension lda
    int64_t ir, v, lr, a;
    int32_t r, n, lda, pad;
};

struct CopyJob {   // B(0:m, 0:n) = A(0:m, 0:n): leading sub-block copy between pool regions
    int64_t src, dst;
    int32_t m, n, lds, ldd;
};

// Batched skinny GEMV work: y[row0 .. row0+rows) (+)= sum_t A_t x_t
struct GTerm {
    int64_t a;     // element offset of A (column-major, leading dimension = item.lda) at row 0
    int64_t x;     // element offset of x in the vector pool
    int64_t t;     // symmetric term: vector-pool offset receiving A^T x_item (cols entries), else -1
    int32_t cols;
    int32_t ld;    // transposed term (16-wide transposed launch): leading dimension of the stored block
};
struct GItem {
    int64_t y;     // element offset of the output segment (first row of this tile)
    int64_t xb;    // offset of this tile's rows in the input vector (symmetric terms' transposed part)
    int32_t rows;  // <= kRowTile
    int32_t lda;
    int32_t row0;  // first row of this tile inside the terms' matrices
    int32_t t0, t1;
    int32_t flags; // bit 0: atomic add (else exclusive); bit 1: operator in the second pool;
                   // bit 2: fused M2L + near-field item (atomic in the single-vector product);
                   // bits 8..: system of a batch
};

constexpr int kRowTile = 128;  // rows of one GEMV item
constexpr int kNR = 16;        // right-hand sides of one multi-vector pass

// ------------------------------------------------------------------ launch wrappers (kernels.hip)
// wide cross for a few very large blocks (state per block, partials per workgroup)
struct WideState {
    // status: 0 running / finished, 1 out of workspace (re-run), 2 parked at the end of its fixed
    // phase (kfix steps without the normS update), 3 ended inside its fixed phase (exhausted)
    int32_t rank, rp, cp, l, next_rp, done, status, found, dots_rank, kfix;
    double piv, vn2, un, vn, normS;
    int64_t off;
};
struct WidePart {
    double ss, a;
    int32_t idx, pad;
    double sv;
};
void launch_wide_init(const AcaJob *jobs, int nb, int maxmn, WideState *state, uint8_t *flags, int64_t *tab,
                      unsigned long long *ws_top, int64_t ws_cap, hipStream_t s);
// one pivot step of every block; maxrank bounds the ranks reached so far (+1)
void launch_wide_step(const AcaJob *jobs, int nb, int maxm, int maxn, int maxrank, const int32_t *idx, double *ws,
                      int64_t *tab, uint8_t *flags, WideState *state, WidePart *part, int pstride, double *dots,
                      int dstride, int32_t *rowpiv, int32_t *colpiv, double *hist, unsigned long long *ws_top,
                      int64_t ws_cap, KParams kp, double tol, hipStream_t s);
void launch_wide_finish(const WideState *state, int nb, int32_t *rank_out, int32_t *status_out, int32_t *ndone,
                        hipStream_t s);
// fixed-rank phase (AcaJob.pad = kfix > 1 steps): one pivot step without the normS update
void launch_wide_fixed_step(const AcaJob *jobs, int nb, int maxm, int maxn, const int32_t *idx, double *ws,
                            int64_t *tab, uint8_t *flags, WideState *state, WidePart *part, int pstride,
                            int32_t *rowpiv, int32_t *colpiv, double *hist, unsigned long long *ws_top, int64_t ws_cap,
                            KParams kp, double tol, hipStream_t s);
// Gram products of parked blocks: for tile job (block, qi, ki) of 64 x 64 term pairs, prod[off + k(k+1)/2 + q]
// = |<U_q, U_k>| |<V_q, V_k>| for q <= k < rank
struct GramJob {
    int32_t blk, qi, ki, pad;
    int64_t off;
};
void launch_wide_gram(const GramJob *gj, int njobs, const AcaJob *jobs, const WideState *state, const double *ws,
                      const int64_t *tab, double *prod, hipStream_t s);
// the judged blocks: final_rank[b] >= 0 ends block b at that rank; -1 resumes the adaptive cross with
// normS = normS_in[b]; -2 leaves the block as it is
void launch_wide_judge(const AcaJob *jobs, int nb, WideState *state, const int32_t *final_rank, const double *normS_in,
                       int64_t *tab, unsigned long long *ws_top, int64_t ws_cap, hipStream_t s);
// sampled-entry check of finished blocks: err2[b], ref2[b] over ns entries (i, j) drawn from a hash of
// (b, sample): |A(i,j) - sum_q U_q[i] V_q[j]|^2 and |A(i,j)|^2
void launch_aca_sample(const AcaJob *jobs, int nb, const int32_t *rank, const int32_t *idx, const double *ws,
                       const int64_t *tab, KParams kp, int ns, double *err2, double *ref2, hipStream_t s);
void launch_aca(const AcaJob *jobs, int njobs, const int32_t *idx, double *ws, int64_t ws_cap,
                unsigned long long *ws_top, int64_t *tab, uint8_t *flags, int32_t *rank, int32_t *status,
                int32_t *rowpiv, int32_t *colpiv, KParams kp, double tol, double *hist, int big, hipStream_t s);
void launch_lr_extract(const LrJob *jobs, int njobs, const double *ws, const int64_t *tab,
                       const int32_t *rowpiv, const int32_t *colpiv, double *out, hipStream_t s);
void launch_eval(const EvalJob *jobs, int njobs, const int32_t *idx, double *A, KParams kp, hipStream_t s);
void launch_xfer(const XferJob *jobs, int njobs, const int32_t *idx, const double *lr, double *A, KParams kp,
                 hipStream_t s);
void launch_cur(const CurJob *jobs, int njobs, const int32_t *idx, const double *lr, double *A, KParams kp,
                hipStream_t s);
void launch_copy32(const CopyJob *jobs, int njobs, const float *src, float *dst, hipStream_t s);
// operator pool A0, or A1 for items whose flags have bit 1 set; unr in {2, 4, 8}.  Role 1:
// symmetric terms (t >= 0) also add A^T x_item to vec[t ..] atomically.
void launch_bgemv64(const GItem *items, int nitems, const GTerm *terms, const double *A0, const double *A1,
                    double *vec, int role, int unr, hipStream_t s);
void launch_bgemv32(const GItem *items, int nitems, const GTerm *terms, const float *A0, const float *A1,
                    float *vec, int role, int unr, hipStream_t s);
void launch_otf32(const GItem *items, int nitems, const GTerm *terms, SlotPoints sp, float *vec, hipStream_t s);
void launch_otf64(const GItem *items, int nitems, const GTerm *terms, SlotPoints sp, double *vec, hipStream_t s);
void launch_to_fp32(const double *a, float *b, int64_t n, hipStream_t s);
void launch_to_fp64(const float *a, double *b, int64_t n, hipStream_t s);
// read-only stream of bytes (whole chunks of stream_chunk_bytes(); the rest is not read) by grid
// workgroups of 256; out: grid words, written only in the rare case the loaded words fold to a fixed value
int64_t stream_chunk_bytes();
void launch_stream_read(const void *p, int64_t bytes, uint32_t *out, int grid, hipStream_t s);
void launch_fill(double *a, int64_t n, double v, hipStream_t s);
void launch_fill32(float *a, int64_t n, float v, hipStream_t s);
// tree-order gather/scatter with optional precision change, nrhs vectors of leading dimension ld
void launch_permute_in(const double *x, int64_t ldx, double *xt, int64_t ldt, const int32_t *perm, int64_t N,
                       int nrhs, hipStream_t s);
void launch_permute_in32(const double *x, int64_t ldx, float *xt, int64_t ldt, const int32_t *perm, int64_t N,
                         int nrhs, hipStream_t s);
void launch_permute_out(const double *yt, int64_t ldt, double *y, int64_t ldy, const int32_t *perm, int64_t N,
                        int nrhs, hipStream_t s);
void launch_permute_out32(const float *yt, int64_t ldt, double *y, int64_t ldy, const int32_t *perm, int64_t N,
                          int nrhs, hipStream_t s);
void launch_dense_rows(const int32_t *rows, int64_t nrows, const double *xt, double *y, int64_t N, KParams kp,
                       hipStream_t s);
// BLAS-1 for the Arnoldi process (all FP64)
void launch_mdot(const double *V, int64_t ldv, int k, const double *w, int64_t n, double *partial, int nblk,
                 hipStream_t s);
void launch_mdot_reduce(const double *partial, int nblk, int k, double *out, hipStream_t s);
void launch_mgemv_sub(const double *V, int64_t ldv, int k, const double *h, double *w, int64_t n, hipStream_t s);
void launch_axpby(double a, const double *x, double b, double *y, int64_t n, hipStream_t s);
void launch_scal(double a, double *y, int64_t n, hipStream_t s);  // y *= a
void launch_vec_update(const double *V, int64_t ldv, int k, const double *c, double *x, int64_t n, hipStream_t s);

// kNR right-hand sides, the fused M2L + near-field launch: every stored block applied directly and, for
// symmetric terms, transposed (K^T X_b into the partner's rows, atomics) in one read of the block
void launch_bgemm_sym64(const GItem *items, int nitems, const GTerm *terms, const double *A0, const double *A1,
                        const double *const *tab, double *vec, hipStream_t s);
void launch_bgemm_sym32(const GItem *items, int nitems, const GTerm *terms, const float *A0, const float *A1,
                        const float *const *tab, float *vec, hipStream_t s);
// kNR right-hand sides, transposed products of the stored symmetric blocks (role 3 lists)
// (narrow: slabs of 4 VEC inner indices, 16 columns per load, four slabs per step -- for
// blocks of few rows; else slabs of 16 VEC inner indices, 4 columns per load)
void launch_bgemm_t64(const GItem *items, int nitems, const GTerm *terms, const double *A0, const double *A1,
                      const double *const *tab, double *vec, int narrow, hipStream_t s);
void launch_bgemm_t32(const GItem *items, int nitems, const GTerm *terms, const float *A0, const float *A1,
                      const float *const *tab, float *vec, int narrow, hipStream_t s);
// kNR right-hand sides: vector pools and work vectors hold element (slot, r) at slot * kNR + r;
// per_wave: one item per wave instead of per workgroup (launches of many small items)
void launch_bgemm64(const GItem *items, int nitems, const GTerm *terms, const double *A0, const double *A1,
                    double *vec, int role, bool per_wave, hipStream_t s);
void launch_bgemm32(const GItem *items, int nitems, const GTerm *terms, const float *A0, const float *A1, float *vec,
                    int role, bool per_wave, hipStream_t s);
// dst(t, r) = X(perm[t] + r ldx) for r < nrhs, 0 beyond (perm == NULL: identity)
void launch_gather_m64(const double *X, int64_t ldx, int nrhs, const int32_t *perm, int64_t N, double *dst,
                       hipStream_t s);
void launch_gather_m32(const double *X, int64_t ldx, int nrhs, const int32_t *perm, int64_t N, float *dst,
                       hipStream_t s);
void launch_scatter_m64(const double *src, int64_t N, double *Y, int64_t ldy, int nrhs, const int32_t *perm,
                        hipStream_t s);
void launch_scatter_m32(const float *src, int64_t N, double *Y, int64_t ldy, int nrhs, const int32_t *perm,
                        hipStream_t s);
// batched systems: the operator pools of system f >> 8 of an item come from tab[2 f + 0/1]
void launch_bgemm64_tab(const GItem *items, int nitems, const GTerm *terms, const double *const *tab, double *vec,
                        int role, bool per_wave, hipStream_t s);
void launch_bgemm32_tab(const GItem *items, int nitems, const GTerm *terms, const float *const *tab, float *vec,
                        int role, bool per_wave, hipStream_t s);
void launch_map_in64(const double *X, const int32_t *map, int64_t n, double *pool, hipStream_t s);
void launch_map_in32(const double *X, const int32_t *map, int64_t n, float *pool, hipStream_t s);
void launch_map_out64(const double *pool, const int32_t *map, int64_t n, double *Y, hipStream_t s);
void launch_map_out32(const float *pool, const int32_t *map, int64_t n, double *Y, hipStream_t s);
void launch_gather_b(const double *X, const int64_t *src, const int32_t *ld, int nrhs, int64_t n, double *dst,
                     hipStream_t s);
void launch_scatter_b(const double *v, const int64_t *src, const int32_t *ld, int nrhs, int64_t n, double *Y,
                      hipStream_t s);
// rank predictor: out[i] = log2(1 + predicted rank) of block i (16 features each, BF16 parameters)
void launch_rank_mlp(const float *feat, int n, const uint16_t *w, float *out, hipStream_t s);
// lockstep solves over S systems x kNR columns (segmented interleaved vectors, seg == NULL: S == 1)
// a space of k vectors shared by all kNR columns of a system (row-major, n x k; several systems: each
// system's slots hold its own space): per-block partial dot products with a lockstep vector (reduced per
// system by launch_bdot16_reduce), w += sign Cs h (h: k x S x kNR), and the row-major copy
void launch_sdot16(const double *Cs, int k, const double *w, const int2 *blk, int nblk, double *partial, hipStream_t s);
void launch_supd16(const double *Cs, int k, const double *h, int S, double sign, double *w, int64_t n,
                   const int32_t *seg, hipStream_t s);
void launch_to_rowmajor(const double *src, int64_t n, int k, double *dst, hipStream_t s);
void launch_bdot16(const double *V, int64_t ldv, int k, const double *w, const int2 *blk, int nblk, double *partial,
                   hipStream_t s);
void launch_bdot16_reduce(const double *partial, int nblk, int k, const int32_t *sblk, int S, double *out,
                          hipStream_t s);
void launch_bgemv16(const double *V, int64_t ldv, int k, const double *h, int S, double sign, double *w, int64_t n,
                    const int32_t *seg, hipStream_t s);
void launch_bscale16(const double *src, const double *sc, double *dst, int64_t n, const int32_t *seg, hipStream_t s);
int mdot_blocks(int64_t n);

// complex kernels (complex.hip): operator and vector pools of double2 (re, im), element offsets as
// in the real path
void launch_aca_c(const AcaJob *jobs, int njobs, const int32_t *idx, double2 *ws, int64_t ws_cap,
                  unsigned long long *ws_top, int64_t *tab, uint8_t *flags, int32_t *rank, int32_t *status,
                  int32_t *rowpiv, int32_t *colpiv, KParamsC kp, double tol, double *hist, int big, hipStream_t s);
void launch_lr_extract_c(const LrJob *jobs, int njobs, const double2 *ws, const int64_t *tab, const int32_t *rowpiv,
                         const int32_t *colpiv, double2 *out, hipStream_t s);
void launch_eval_c(const EvalJob *jobs, int njobs, const int32_t *idx, double2 *A, KParamsC kp, hipStream_t s);
void launch_xfer_c(const XferJob *jobs, int njobs, const int32_t *idx, const double2 *lr, double2 *A, KParamsC kp,
                   hipStream_t s);
void launch_cur_c(const CurJob *jobs, int njobs, const int32_t *idx, const double2 *lr, double2 *A, KParamsC kp,
                  hipStream_t s);
void launch_cgemv(const GItem *items, int nitems, const GTerm *terms, const double2 *A, double2 *vec, hipStream_t s);
void launch_cfill(double2 *a, int64_t n, hipStream_t s);
void launch_cpermute_in(const double2 *x, double2 *xt, const int32_t *perm, int64_t N, hipStream_t s);
void launch_cpermute_out(const double2 *yt, double2 *y, const int32_t *perm, int64_t N, hipStream_t s);
void launch_cdense_rows(const int32_t *rows, int64_t nrows, const double2 *xt, double2 *y, int64_t N, KParamsC kp,
                        hipStream_t s);
int cmdot_blocks(int64_t n);
void launch_cmdot(const double2 *V, int64_t ldv, int k, const double2 *w, int64_t n, double2 *partial, int nblk,
                  hipStream_t s);
void launch_cmdot_reduce(const double2 *partial, int nblk, int k, double2 *out, hipStream_t s);
void launch_cmgemv(const double2 *V, int64_t ldv, int k, const double2 *h, double2 *w, int64_t n, double sign,
                   hipStream_t s);
void launch_caxpby(double2 a, const double2 *x, double2 b, double2 *y, int64_t n, hipStream_t s);

}  // namespace h2w
