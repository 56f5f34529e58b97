/*
 * h2w.h -- C ABI of the MI355X-native H2* / (H2+H)* fast matrix-vector
 * product (libh2w.so).  Plain pointers and sizes only.
 *
 * It replaces the reference's operator facade, the class `initialisation`
 * (1D/codes/matvec.hpp:6-113, 2D/codes/matvec.hpp:6-172,
 * 3D/codes/matvec.hpp:6-178), and the GMRES driver that calls it
 * ({1,2,3}D/codes/gmres.hpp:91-200, {1,2,3}D/codes/main.cpp).
 *
 * Conventions
 *   - All functions return an H2W_* status; they never exit the process
 *     (the reference exits on a bad `choice`, 3D/codes/kernel.hpp:175-179,
 *     and on nLevels < 1, 3D/codes/main.cpp:12-16).  h2w_last_error() gives
 *     a message for the last failure on the calling thread.
 *   - Vectors are FP64 and in INPUT order unless the call says otherwise.
 *   - `where` = H2W_HOST (host pointers, copied through pinned staging) or
 *     H2W_DEVICE (device pointers on the handle's HIP device).
 *   - Calls on one handle are serialized by the caller (the reference's
 *     matvec is not re-entrant either); distinct handles are independent.
 */
#ifndef H2W_H
#define H2W_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define H2W_OK 0
#define H2W_EINVAL 1   /* invalid argument (dim, choice, N, n_max, nLevels < 1, ...) */
#define H2W_ESTATE 2   /* wrong call order, e.g. matvec before assemble            */
#define H2W_EHIP 3     /* a HIP runtime call failed                                 */
#define H2W_ENOMEM 4   /* device or host allocation failed                          */
#define H2W_ENOCONV 5  /* GMRES stopped at max_iter without reaching tol            */
#define H2W_ENOLIB 6   /* no usable GPU (the library has no CPU fallback)           */

/* hierarchical format (compile-time -DUSE_nHODLRdD / -DUSE_snHODLRdD in the reference) */
#define H2W_NESTED 0      /* H2*    : nHODLRdD, nested far + nested vertex-sharing blocks */
#define H2W_SEMI_NESTED 1 /* (H2+H)*: s-nHODLRdD, nested far + per-pair CUR vertex blocks */
#define H2W_NON_NESTED 2  /* H*     : HODLRdD (-DUSE_HODLRdD), per-pair CUR on every interaction-list
                                       pair, 2D/3D only (2D/codes/HODLR2Dtree.hpp:971-1073,
                                       3D/codes/HODLR3Dtree.hpp:2239-2373)                        */

/* kernel choice, as argv[5] of the reference's main */
#define H2W_CHOICE_IE 0   /* integral equation kernel with its diagonal rule */
#define H2W_CHOICE_RBF 1  /* the repo's RBF kernel                            */
/* complex kernels (the reference's compile-time -DUSE_Helm / -DUSE_Hankel builds, with its ker_choice
   numbers).  A handle with one of them is complex: every vector it takes or returns holds N complex
   values as interleaved (re, im) doubles, operators are complex FP64 (precision H2W_FP64 and GMRES mode
   H2W_GMRES_FP64 only), and entry points without a complex form refuse it with H2W_EINVAL. */
#define H2W_CHOICE_HELMHOLTZ 2 /* 3D: e^{iR} / (4 pi), diagonal 500 sqrt(N) (3D/codes/kernel.hpp:132-144) */
#define H2W_CHOICE_HANKEL 10   /* 2D: (i/4) H0^(1)(R), diagonal (1 + 15 log N + Kii) / h^2
                                  (2D/codes/kernel.hpp:126-145, 2D/codes/integral2d.hpp:16-57)     */

#define H2W_HOST 0
#define H2W_DEVICE 1

#define H2W_FP64 0      /* stored FP64 operators                                                  */
#define H2W_FP32 1      /* stored FP32 copies of every operator, FP32 accumulation                */
#define H2W_FP32_OTF 2  /* FP32: near-field and M2L entries evaluated on the fly, transfers stored */
#define H2W_FP64_OTF 3  /* FP64: near-field and M2L entries evaluated on the fly, transfers stored */
#define H2W_FP32_INNER 4 /* FP32 inner operator of the refinement: every box keeps the leading
                            pivots at which its cross would have stopped at 10^-ir_tol_pow       */

#define H2W_ORDER_INPUT 0
#define H2W_ORDER_TREE 1

typedef struct h2w_handle h2w_handle;

typedef struct {
    int compat;              /* 1 (default): reproduce the reference's quirks, i.e. the 1D RBF
                                diagonal sqrt(2N) of the doubled particle vector
                                (1D/codes/matvec.hpp:35,38-39; 1D/codes/kernel.hpp:180)          */
    int nlevels;             /* <= 0 (default): the reference main's formula
                                ceil(log_{2^d}(N / n_max)) ({1,2,3}D/codes/main.cpp:11-13)             */
    int fp32_shadow;         /* 1 (default): also keep an FP32 copy of every stored operator     */
    int device;              /* HIP device ordinal (default 0)                                   */
    int rank_cap;            /* first-pass rank estimate (default 256) used to admit ACA blocks
                                of the top vertex level and of the semi-nested pairs to a shared
                                workspace batch; a block that outgrows the batch's pool is re-run,
                                so results do not depend on it                                   */
    int64_t workspace_bytes; /* ACA workspace budget in bytes (0 = automatic)                    */
    int plan_only;           /* 1: build the tree and lists only, touch no device (assembly and
                                products then return H2W_ESTATE); for host-side inspection        */
    int ir_tol_pow;          /* tolerance exponent of the refinement's truncated inner operator
                                (default 6; 0 or >= tol_pow: not built)                            */
} h2w_opts;

typedef struct {
    int dim, variant, choice, nlevels;
    int64_t N;
    double memory;           /* stored-element count as the reference's findMemory_n /
                                findMemory_sn counts it (3D/codes/HODLR3Dtree.hpp:2505-2628)     */
    double diag;             /* value of the diagonal entries                                    */
    double assembly_seconds; /* wall time of h2w_assemble                                        */
    int max_rank_far, max_rank_vertex;
    double mean_rank_far, mean_rank_vertex;
    int64_t stored_bytes_fp64;   /* bytes one FP64 matvec streams from the operator pools     */
    int64_t stored_bytes_fp32;   /* same for the FP32 shadow (0 if not built)                 */
    int n_launches_matvec;   /* kernel launches per matvec                                       */
    int aca_reruns;          /* blocks re-run because they outgrew their workspace batch         */
    int64_t stored_bytes_inner_fp32; /* bytes one H2W_FP32_INNER matvec streams                      */
    int aca_batches;         /* ACA launch batches of the assembly                                 */
    int64_t aca_pool_bytes;  /* largest ACA workspace of the assembly                              */
    int rank_pred_blocks;    /* blocks whose workspace came from the rank model (0: no model)      */
    int rank_pred_under;     /* of those, blocks whose rank exceeded the estimate (re-run)         */
    double rank_pred_mae_log2; /* mean |log2(1 + predicted) - log2(1 + rank)| over those blocks    */
    double inner_tol_exp;    /* tolerance exponent p (10^-p) of the inner operator, 0 if none          */
    int rank_fixed_blocks;   /* wide-cross blocks compressed with a fixed-rank phase from the rank model
                                (kfix = 1.03 x predicted + 2 steps, then the reference's stopping rule
                                applied to all of them at once)                                       */
    int rank_fixed_resumed;  /* of those, blocks whose rank the model under-predicted: the stopping rule
                                held at every fixed step and the adaptive cross continued (fallback)  */
    int64_t rank_fixed_extra_steps; /* pivot steps run beyond where the stopping rule ended a block    */
    double rank_fixed_sampled_err;  /* largest relative error of those blocks' cross approximation on
                                       256 sampled entries each                                       */
    double diag_im;          /* imaginary part of the diagonal value (complex kernels), else 0        */
} h2w_stats;

typedef struct {
    int iterations;          /* Arnoldi steps (FP64 mode) / total inner steps (FP32-IR mode),
                                counted like the reference's `max_iter` output                   */
    int outer_iterations;    /* refinement steps (FP32-IR mode), 0 for FP64 mode                 */
    int matvecs_fp64, matvecs_fp32;
    double resid;            /* final |s_{i+1}|/||b|| (FP64) or true ||b - A x||/||b|| (IR)      */
    double seconds;          /* wall time, first residual matvec to convergence                  */
    int converged;
} h2w_gmres_stats;

#define H2W_GMRES_FP64 0         /* unrestarted GMRES, stored FP64 operators                          */
#define H2W_GMRES_FP32_IR 1      /* FP64 iterative refinement (stored FP64 residuals) around GMRES on
                                    the stored FP32 operators; the first correction starts from the
                                    FP32 operator's residual of x0 (convergence is decided on FP64
                                    residuals only)                                                   */
#define H2W_GMRES_FP32_OTF_IR 2  /* as H2W_GMRES_FP32_IR with the inner operator's near-field and M2L
                                    blocks evaluated on the fly in FP32                               */
#define H2W_GMRES_FP32_INNER_IR 3 /* as H2W_GMRES_FP32_IR with the truncated FP32 inner operator     */

void h2w_default_opts(h2w_opts *opts);

/* Constructor.  Replaces `initialisation(N, nLevels, n_max, L, TOL_POW, choice)`
 * (1D/codes/matvec.hpp:18; 2D/codes/matvec.hpp:20 and 3D/codes/matvec.hpp:20
 * take sqrt(N)/cbrt(N) and the per-axis leaf size instead).
 *   points: dim x N, SoA (x[0..N), y[0..N), z[0..N)), host memory; NULL selects
 *           the reference's uniform grid (1D/codes/HODLR1Dtree.hpp:59-65,
 *           2D/codes/HODLR2Dtree.hpp:75-91, 3D/codes/HODLR3Dtree.hpp:82-102),
 *           which needs N to be a square (2D) or cube (3D).
 *   L: half-width of the root box [-L, L]^dim; tol_pow: ACA/NCA tolerance 10^-tol_pow. */
int h2w_create(int dim, int variant, int64_t N, const double *points, int n_max, double L, int tol_pow,
               int choice, const h2w_opts *opts, h2w_handle **out);

/* Assembly.  Replaces `initialisation::tree_initialisation()`
 * (3D/codes/matvec.hpp:33-61 nested, :86-113 semi-nested; 1D/codes/matvec.hpp:28-49,67-87). */
int h2w_assemble(h2w_handle *h);
/* No counterpart in the reference (one system per process): h2w_assemble on n handles from up to
 * nthreads host threads, each system's launches on its own stream, so that the latency-bound level
 * loops of small systems overlap on the device.  The operators are those of h2w_assemble one by one.
 * Every handle is assembled or the first failure is returned (with the system's index in
 * h2w_last_error()). */
int h2w_assemble_many(h2w_handle *const *hs, int n, int nthreads);

/* y = A x (overwrites y).  Replaces `initialisation::mat_vec_prod(Vec& charge, Vec& output)`
 * (3D/codes/matvec.hpp:63-83, :115-136; 1D/codes/matvec.hpp:51-65,89-103). */
int h2w_matvec(h2w_handle *h, const double *x, double *y, int where);

/* General form: precision H2W_FP64, H2W_FP32 (FP32 operators, FP32 accumulation), H2W_FP32_INNER,
 * H2W_FP32_OTF or H2W_FP64_OTF (x and y are FP64 either way), order H2W_ORDER_INPUT or H2W_ORDER_TREE,
 * nrhs right-hand sides stored one after another with leading dimension ld (>= N).  With nrhs > 1 and a
 * stored-operator precision the product runs 16 right-hand sides per pass on the matrix cores
 * (each operator block read once per pass); the on-the-fly precisions loop over the columns. */
int h2w_matvec_ex(h2w_handle *h, const double *x, double *y, int nrhs, int64_t ld, int precision, int order,
                  int where);

/* GMRES driver.  Replaces `iterative_solver(max_iter, tol_pow).GMRES(A, x, b)`
 * (1D/codes/gmres.hpp:22-28,91-200): unrestarted (restart = max_iter); x holds x0 on entry
 * (the reference's main passes ones, 1D/codes/main.cpp:34) and the solution on return.
 * Returns H2W_ENOCONV (with x updated) when max_iter is reached first. */
int h2w_gmres(h2w_handle *h, const double *b, double *x, int max_iter, double tol, int mode, int where,
              h2w_gmres_stats *stats);

/* nrhs independent GMRES solves A x_c = b_c (columns of B and X, leading dimensions ldb, ldx >= N),
 * 16 at a time in lockstep: every Arnoldi step is one 16-wide product on the matrix cores, each
 * column keeps its own Krylov basis, Hessenberg matrix, Givens rotations and stopping test (those of
 * h2w_gmres), and stops at its own iteration.  Modes H2W_GMRES_FP64, H2W_GMRES_FP32_IR and
 * H2W_GMRES_FP32_INNER_IR.  stats receives nrhs entries.  No counterpart in the reference, whose
 * driver solves one right-hand side (1D/codes/gmres.hpp:91-200). */
int h2w_gmres_multi(h2w_handle *h, const double *B, int64_t ldb, double *X, int64_t ldx, int nrhs, int max_iter,
                    double tol, int mode, int where, h2w_gmres_stats *stats);

/* Dense FP64 product of selected rows of A with x, entries evaluated on the fly on the GPU.
 * rows == NULL means all N rows.  Replaces `getTruePoten` (3D/codes/HODLR3Dtree.hpp:2630-2643). */
int h2w_dense_matvec(h2w_handle *h, const double *x, double *y, const int64_t *rows, int64_t nrows, int where);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The product of h2w_dense_matvec with
 * device pointers only (x: N doubles and y: nrows doubles in input order, rows: nrows int64 input-order
 * row indices or NULL for all N rows), without any allocation: its staging (x in tree order, the rows in
 * tree order, a count of bad row indices) lives in one fixed device array per GPU of 12 MiB, set up when
 * the first handle on that GPU is created and held until the process ends.  Limits: N <= 2^20 points and
 * nrows <= 2^20 rows per call (larger: H2W_EINVAL, use h2w_dense_matvec).  A row index outside [0, N)
 * gives H2W_EINVAL after the call (y is then unspecified).  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other host threads, other handles' streams) race. */
int h2w_dense_matvec_noalloc(h2w_handle *h, const double *x, double *y, const int64_t *rows, int64_t nrows);

int h2w_get_stats(const h2w_handle *h, h2w_stats *stats);

/* Timing of the matvec on device-resident data (no counterpart in the reference, whose
 * transcripts time whole GMRES solves with omp_get_wtime, 1D/codes/main.cpp:37-40).
 * Runs nrep tree-order matvecs back to back on the handle's stream, bracketed by HIP
 * events; also records each launch of the sequence (average over nrep) and the operator
 * bytes that launch streams. */
#define H2W_MAX_LAUNCHES 64
typedef struct {
    double ms_per_matvec;
    double launch_ms[H2W_MAX_LAUNCHES];
    int n_launch;
    int64_t launch_bytes[H2W_MAX_LAUNCHES]; /* operator bytes the launch streams; for the on-the-fly
                                               precisions' M2L + near-field launch: entries it evaluates
                                               (every directed pair) times the element size        */
    int launch_role[H2W_MAX_LAUNCHES]; /* -1 zero fill, 0 transfer pass, 1 M2L + near field (one
                                          stored block per unordered pair, applied both ways),
                                          2 semi-nested column factors                       */
} h2w_bench;
int h2w_bench_matvec(h2w_handle *h, int nrep, int precision, h2w_bench *out);
/* The same for the 16-wide product (matrix cores, 16 right-hand sides per pass): ms_per_matvec is the
 * time of one pass over all 16 columns.  Entry 0 is the whole pass (launch_role[0] = -2,
 * launch_bytes[0] the operator bytes one pass multiplies: the symmetric blocks count twice, for
 * their direct and transposed products); entries 1.. are its launches one by one (role 4: the
 * transposed products of the symmetric blocks, a launch of their own on this path). */
int h2w_bench_matvec_multi(h2w_handle *h, int nrep, int precision, h2w_bench *out);
/* No counterpart in the reference: a measurement helper.  Read bandwidth of a plain read-only
 * stream over `bytes` (>= 1 MiB, rounded down to a multiple of 32 KiB) of device memory on `device`, averaged over
 * nrep launches after one warm-up (HIP events; allocates and frees its own buffer): the achievable
 * rate the streaming products are compared with, beside the published 8 TB/s. */
int h2w_bench_stream_read(int device, int64_t bytes, int nrep, double *gbps);

/* Rebuild the truncated FP32 inner operator of H2W_GMRES_FP32_INNER_IR for tolerance
 * 10^-ir_tol_pow (0 < ir_tol_pow < tol_pow, fractional exponents allowed as the autotuner's half
 * decades; 0 removes it).  No counterpart in the reference
 * (which has one FP64 operator); the stored FP64 / FP32 operators are untouched. */
int h2w_set_inner_tolerance(h2w_handle *h, double ir_tol_pow);

/* Deflation space of the refinement (no counterpart in the reference, whose GMRES is
 * unpreconditioned): m Arnoldi steps (2 <= m <= 512, m < N) of the inner operator `precision`
 * (H2W_FP32, H2W_FP32_INNER or H2W_FP32_OTF) from a fixed pseudo-random vector.  k = 0 keeps that
 * whole Krylov space, 0 < k < m the k Ritz vectors whose Ritz values lie farthest from the
 * diagonal value (the operator is that value times I plus a kernel matrix; its spectrum clusters
 * there).  Kept as U and C = A U with orthonormal C (and a row-major copy of both for the
 * 16-column solves: 4 k N doubles of device memory).  Every correction of a later h2w_gmres,
 * h2w_gmres_multi (one copy shared by the columns) or h2w_batch_gmres (when every system of the
 * batch has one) in the matching refinement mode starts from x = U C^T r and orthogonalises its
 * Arnoldi vectors against C as well (GCRO), which removes the operator's outlying eigendirections
 * from every solve.  It depends on the operator only, never on a right-hand side; m = 0 drops it,
 * and rebuilding the inner operator (h2w_set_inner_tolerance, h2w_autotune) invalidates it for
 * H2W_FP32_INNER.  Plain FP64 GMRES does not use it. */
int h2w_set_deflation(h2w_handle *h, int m, int k, int precision);

/* Autotuner (no counterpart in the reference).  With H2W_TUNE_UNROLL it times every launch of
 * the FP64, FP32 and FP32-inner matvecs with 2, 4 and 8 column loads in flight per lane and
 * keeps the fastest per launch.  With H2W_TUNE_INNER it rebuilds the inner operator for
 * tolerances 10^-p, p = 4, 4.5, ..., 8 below tol_pow, solves A x = b with
 * H2W_GMRES_FP32_INNER_IR from x0 = 0 and from x0 = 1, scores each p by the slower solve
 * (raised by a fifth when a final residual exceeds tol / 2) and keeps the best.  With
 * H2W_TUNE_DEFLATION it then scores deflation spaces (h2w_set_deflation on the inner operator;
 * (m, k) = (0, 0) i.e. none, (64, 32), (128, 48), (256, 64), (256, 96), (384, 128)) the same way,
 * keeps the best, and with H2W_TUNE_INNER also re-scores the inner tolerances half a decade either
 * side of the chosen one under that space.
 * b has N entries in input order, on the host or the device (`where`); choose a b other than the
 * right-hand sides to be timed. */
#define H2W_TUNE_UNROLL 1
#define H2W_TUNE_INNER 2
#define H2W_TUNE_DEFLATION 4
#define H2W_TUNE_MAX_CAND 12
typedef struct {
    double ir_tol_exp;                       /* chosen inner exponent p (0: no inner operator) */
    int n_cand;
    double cand_ir_tol_exp[H2W_TUNE_MAX_CAND];
    double cand_seconds[H2W_TUNE_MAX_CAND];  /* the slower of the two solves */
    int cand_iterations[H2W_TUNE_MAX_CAND];  /* inner (FP32) GMRES iterations of that solve */
    int cand_outer[H2W_TUNE_MAX_CAND];       /* refinement steps of that solve */
    int cand_converged[H2W_TUNE_MAX_CAND];
    int64_t cand_inner_bytes[H2W_TUNE_MAX_CAND]; /* FP32 bytes one inner matvec streams */
    double matvec_ms_before[3];              /* FP64, FP32, FP32 inner matvec before the unroll tuning */
    double matvec_ms_after[3];
    int defl_m, defl_k;                      /* chosen deflation space (0, 0: none) */
    double defl_build_seconds;               /* time to build it */
    int n_defl_cand;
    int cand_defl_m[H2W_TUNE_MAX_CAND], cand_defl_k[H2W_TUNE_MAX_CAND];
    double cand_defl_ir_tol_exp[H2W_TUNE_MAX_CAND];
    double cand_defl_seconds[H2W_TUNE_MAX_CAND];  /* the slower of the two solves */
    int cand_defl_iterations[H2W_TUNE_MAX_CAND];
    int cand_defl_outer[H2W_TUNE_MAX_CAND];
    int cand_defl_converged[H2W_TUNE_MAX_CAND];
} h2w_tune;
int h2w_autotune(h2w_handle *h, const double *b, int where, double tol, int max_iter, int flags, h2w_tune *out);

/* ---- saving and loading an assembled operator (no counterpart in the reference, which re-assembles on
 * every run).  The file holds the skeletons (per box the rank, row and column pivots, L and R, and the
 * cross's per-step history; per semi-nested / non-nested pair the same), the input points and the build
 * parameters -- not the operator pools.  h2w_load rebuilds the plan and re-evaluates every operator block
 * with the same kernels: the loaded operator equals the saved one bit for bit, and no cross approximation
 * runs.  The file ends with an FNV-1a hash; truncated or altered files are refused (H2W_EINVAL).  The hash
 * only detects accidents: every count, rank, box, pair and pivot index in the file is also checked against
 * the plan rebuilt from its points before any device work, so a crafted file is refused as well. */
int h2w_save(const h2w_handle *h, const char *path);
/* opts: device, fp32_shadow, ir_tol_pow, ... of the new handle (nlevels and compat come from the file).
 * With opts->plan_only = 1 the file is read and checked only, no device is touched, and the handle
 * answers plan queries (h2w_tree_order, h2w_interaction_list, ...). */
int h2w_load(const char *path, const h2w_opts *opts, h2w_handle **out);

/* ---- learned rank predictor (no counterpart in the reference).  A 16-64-64-64-1 perceptron (GELU,
 * BF16 parameters, FP32 accumulation, evaluated on the GPU) predicts log2(1 + rank) of every ACA block
 * from 16 geometry features (family, level, row/column and point counts, interaction-set size and
 * separations, box fill, box centre).  With a model set, each block reserves workspace for 1.2 x its
 * predicted rank + 4 and blocks are batched in descending predicted rank; the adaptive cross still decides
 * pivots and ranks (a block that outgrows its estimate is re-run), so the assembled operator does not
 * depend on the model.  Parameter layout: W1 (16 x 64, input-major), b1, W2 (64 x 64), b2, W3, b3,
 * w4 (64), b4. */
#define H2W_RANK_FEATURES 16
#define H2W_RANK_PARAMS 9473
/* params == NULL removes the model; takes effect at the next h2w_assemble. */
int h2w_set_rank_model(h2w_handle *h, const uint16_t *params, int64_t count);
/* (features, rank) of every ACA block of the last assembly (the training harvest); copies up to cap
 * samples (feat: cap x H2W_RANK_FEATURES) and returns the number available (or -H2W_*). */
int64_t h2w_rank_samples(const h2w_handle *h, float *feat, int32_t *rank, int64_t cap);
/* The model's output log2(1 + rank) for n feature rows (host pointers), by the same GPU kernel. */
int h2w_rank_predict(h2w_handle *h, const float *feat, int64_t n, float *out);

/* ---- batches of independent systems (no counterpart in the reference, which solves one system per run).
 * A batch refers to assembled handles on one device (they must outlive it and not be re-assembled; an
 * inner operator rebuilt after h2w_batch_create makes H2W_FP32_INNER calls fail with H2W_ESTATE).  Every
 * product is one launch per phase for all systems (their work lists merged by phase, up passes aligned at
 * the leaves, down passes ending there), 16 right-hand sides per pass on the matrix cores.  Vectors: for
 * each system in order an N_s x nrhs column-major block (leading dimension N_s), blocks concatenated. */
typedef struct h2w_batch h2w_batch;
int h2w_batch_create(h2w_handle *const *systems, int nsys, h2w_batch **out);
/* Y = A X for every system (precision H2W_FP64, H2W_FP32 or H2W_FP32_INNER). */
int h2w_batch_matvec(h2w_batch *b, const double *X, double *Y, int nrhs, int precision, int where);
/* nsys x nrhs independent solves in lockstep (the per-column algorithm of h2w_gmres_multi); stats has
 * nsys * nrhs entries, system-major. */
int h2w_batch_gmres(h2w_batch *b, const double *B, double *X, int nrhs, int max_iter, double tol, int mode,
                    int where, h2w_gmres_stats *stats);
/* HIP-event time of one batched 16-wide product; launch_bytes / launch_role per merged launch. */
int h2w_batch_bench_matvec(h2w_batch *b, int nrep, int precision, h2w_bench *out);
void h2w_batch_destroy(h2w_batch *b);

/* Tree-order permutation: perm[t] = input index of the t-th point in tree order
 * (leaf boxes' chargeLocations concatenated, 3D/codes/HODLR3Dtree.hpp:2489-2503). */
int h2w_tree_order(const h2w_handle *h, int32_t *perm);

/* Skeleton of one box after assembly.  family 0 = far field (incoming_checkPoints /
 * incoming_chargePoints, 3D/codes/HODLR3Dtree.hpp:1915-1924), 1 = vertex-sharing nested
 * (…_ver, :2083-2092).  ir/ic receive input-order point indices (capacity cap each).
 * Returns the rank (>= 0) or -H2W_* on error. */
int h2w_box_skeleton(const h2w_handle *h, int family, int level, int box, int32_t *ir, int32_t *ic, int cap);

/* Interaction list of a box in the reference's order (3D/codes/HODLR3Dtree.hpp:155-1680) with
 * the vertex flag of each entry (:1744-1767); near neighbours (leaf level) in slot order.
 * Returns the list length or -H2W_* on error. */
int h2w_interaction_list(const h2w_handle *h, int level, int box, int32_t *il, int32_t *is_vertex, int cap);
int h2w_near_list(const h2w_handle *h, int level, int box, int32_t *nb, int cap);

void h2w_destroy(h2w_handle *h);

const char *h2w_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* H2W_H */
