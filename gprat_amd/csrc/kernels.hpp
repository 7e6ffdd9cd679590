// Host-side launchers for the gfx950 kernels. Every matrix is row-major
// double (or float where stated) with an explicit leading dimension; every
// launcher checks the shapes it relies on before launching.
#pragma once

#include "common.hpp"

namespace gpamd::k
{

// ---- covariance family (k_covariance.hip) ---------------------------------

// Lower band of the padded training covariance into A (n_pad x n_pad, lda):
// all 64x64 blocks on or below the diagonal. Rows/columns >= n form an
// identity block, so the Cholesky factor of the padded matrix is
// blockdiag(L, I).
void assemble_lower(hipStream_t s, const double *x, int r, SEK sek, double *A, long lda, int n, int n_pad);
// Strict upper triangle of the same matrix: A[i][j] = K(i, j) for i < j
// (zero where i or j >= n); the lower triangle and diagonal are not touched.
void assemble_upper(hipStream_t s, const double *x, int r, SEK sek, double *A, long lda, int n, int n_pad);

// B = K*^T (n_pad x m_pad, ldb): B[i][j] = k(test j, train i); zero outside n x m.
// Rows of K* (m x n, row-major at S with row stride lds): entry (j, i) for
// test point j and training point i; zero for i >= n or j >= m up to
// n_pad x m_pad.
void assemble_cross_rows(hipStream_t s, const double *x, const double *z, int r, SEK sek, double *S, long lds, int n, int n_pad, int m,
                         int m_pad);
// out[j] = sum_{i < cols} S[j][i]^2 for j < rows, in a fixed order.
void row_sumsq(hipStream_t s, const double *S, long lds, int rows, int cols, double *out);
void assemble_cross_t(hipStream_t s,
                      const double *x,
                      const double *z,
                      int r,
                      SEK sek,
                      double *B,
                      long ldb,
                      int n,
                      int n_pad,
                      int m,
                      int m_pad);

// The same in FP32 (values computed in FP64, rounded once).
void assemble_cross_t_f32(hipStream_t s,
                          const double *x,
                          const double *z,
                          int r,
                          SEK sek,
                          float *B,
                          long ldb,
                          int n,
                          int n_pad,
                          int m,
                          int m_pad);

// mean[j] = sum_i k(test j, train i) alpha[i], K* generated on the fly.
// `partial` holds splits * m doubles; returns nothing until the reduction
// kernel (same stream) has written mean.
int mean_splits(int n, int m);
void predict_mean(hipStream_t s,
                  const double *x,
                  const double *z,
                  const double *alpha,
                  int r,
                  SEK sek,
                  int n,
                  int m,
                  double *partial,
                  int splits,
                  double *mean);

// Reference tile layout of cholesky() (lower tiles; K in the strict upper
// triangle of diagonal tiles).
void gather_tiles(hipStream_t s, const double *L, long lda, const double *x, int r, SEK sek, int T, int b, double *out);

// Sums for the hyperparameter gradient over the lower triangle of K^-1
// (i, j < n); dK is regenerated on the fly. `part` holds grad_tiles(n) * 5
// doubles. out[0..4] = tr(K^-1 dK_l), tr(K^-1 dK_v), a^T dK_l a, a^T dK_v a,
// tr(K^-1), where dK_l, dK_v include the softplus factors sig_l, sig_v.
long grad_tiles(int n);
void gradient_sums(hipStream_t s,
                   const double *Kinv,
                   long lda,
                   const double *x,
                   const double *alpha,
                   int r,
                   SEK sek,
                   double sig_l,
                   double sig_v,
                   int n,
                   double *part,
                   double *out);

// ---- dense products on MFMA (k_gemm.hip) ----------------------------------

// Which algorithm a product belongs to; each role is its own kernel symbol
// so profilers can attribute time (same code otherwise).
// The role only separates kernel names in profiles (it is the last template
// argument of k_gemm_f64); the values are kept stable for that reason.
enum class GemmRole
{
    Inverse = 3,    // W^T W (A transposed, triangular K range)
    TrailingT = 5,  // Cholesky trailing update from a k-major panel copy
    SolveT = 6,     // TRSM update from a k-major panel copy
    LeafT = 7,      // TRSM diagonal step: k-major block inverse times the right-hand side
    Residual = 8,   // FP64 product with the strictly lower covariance (variance refinement)
};

struct GemmArgs
{
    // C[i][j] (op)= sum_k A(i,k) B(k,j), i < M, j < N, k < K (K % 16 == 0).
    // A(i,k) = a_trans ? A[k*lda+i] : A[i*lda+k]
    // B(k,j) = b_trans ? B[j*ldb+k] : B[k*ldb+j]
    const double *A;
    long lda;
    const double *B;
    long ldb;
    double *C;
    long ldc;
    int M, N, K;
    bool a_trans;
    bool b_trans;
    // Only elements with i + diag_off >= j are touched (lower triangle of a
    // matrix whose (0,0) sits diag_off rows below its diagonal).
    bool lower;
    long diag_off;
    // Entries of A/B with k < max(row tile start, col tile start) + kstart_off
    // are known zero and skipped (triangular operands, e.g. W^T W).
    bool k_from_tile;
    long kstart_off;
    // A(i, k) is known zero for k > i (A lower triangular): each row tile
    // stops K at its last row.
    bool k_to_row;
    // K ends at each row tile's first row (Residual role only): with A(i, k)
    // read from the strict upper triangle of a symmetric matrix (a_trans),
    // each tile takes the part of the strictly lower triangle left of its
    // 128-row diagonal block.
    bool k_before_tile;
    // true: C = acc; false: C -= acc.
    bool set;
    GemmRole role;
    // > 0: at most this many workgroups, each looping over tiles, so the
    // launch leaves workgroup slots free for kernels on other streams.
    int max_wgs;
    // With max_wgs: a device int holding 0 at launch; the workgroups then
    // take tiles in order from this counter (a workgroup that starts late,
    // or shares its CU, takes fewer) instead of a fixed share each.
    int *queue = nullptr;
};
void gemm_f64(hipStream_t s, const GemmArgs &g);

// ---- FP32 matrix-core trailing update, k-major panel copies (k_gemm32.hip) --

// P[k][r] = (float) A[(r0 + r) * lda + p0 + k], r < R, k < W (k-major copy).
void panel_to_f32t(hipStream_t s, const double *A, long lda, long r0, long p0, int R, int W, float *P, long ldp, double *back = nullptr,
                   long ldback = 0);
// The same copy in FP64: the A^T operand layout of gemm_f64 (a_trans).
void panel_to_f64t(hipStream_t s, const double *A, long lda, long r0, long p0, int R, int W, double *P, long ldp, double *back = nullptr,
                   long ldback = 0);
// C[i][j] -= sum_k P[k][a_off + i] P[k][b_off + j] (FP32 products and sums),
// for i < M, j < N, i + diag_off >= j; K % 16 == 0.
// max_wgs > 0 caps the launch (workgroups loop over tiles; with a queue, a
// device int holding 0 at launch, they take tiles in order from it).
void gemm_f32_trailing(hipStream_t s, const float *P, long ldp, long a_off, long b_off, double *C, long ldc, int M, int N, int K, long diag_off,
                       int max_wgs = 0, int *queue = nullptr);
// C (= or -=) A B over FP32 operands, A(i,k) = A[k*lda + i], B(k,j) = B[k*ldb + j],
// FP32 C; K % 16 == 0, M, N, lda, ldb multiples of 4, A and B 16-byte aligned.
// a_lower: A(i, k) is known zero for k > i (each row tile stops K there)
void gemm_f32_tn(hipStream_t s, const float *A, long lda, const float *B, long ldb, float *C, long ldc, int M, int N, int K, bool set,
                 int max_wgs = 0, bool a_lower = false);

// ---- batched fit of many GPs (k_batch.hip) --------------------------------

// One GP of a batch: device addresses of its series (n + n_reg - 1),
// targets (n_pad, zero padded), factor (n_pad x n_pad), diagonal-block
// inverses (n_pad x 64) and alpha (n_pad); n_pad = n rounded up to 64.
struct BatchGP
{
    const double *x, *y;
    double *L, *W, *alpha;
    double *work;  // n_pad doubles, mixed precision only (refinement residual and correction)
    int n, n_pad, n_reg, pad_;
    double lengthscale, vertical, noise;
    double *W128;  // 128-wide level path only: ceil(n_pad / 128) row-major 128x128 block-column inverses
    double *pre;   // 64-wide level path: FP64, n_pad >= 192: BATCH_CARRY_DOUBLES, the next diagonal block's partial sums;
                   // mixed precision: BATCH_REFINE_STATE_DOUBLES, the state of alpha's refinement between launches
};
// Doubles of a BatchGP's `pre` (one thread's 16 accumulator entries and one
// forward-solve partial per thread of a 256-thread workgroup).
constexpr int BATCH_CARRY_DOUBLES = 17 * 256;
constexpr int BATCH_REFINE_STATE_DOUBLES = 4;
// Most refinement steps of alpha in a mixed-precision batched fit.
constexpr int BATCH_REFINE_STEPS = 10;
// Persistent fit of `count` GPs (visited in `order`, largest first) by
// `workgroups` workgroups; losses[g] gets GP g's loss, info[g] its first
// non-positive pivot column (INT_MAX if none). mixed: the block products
// of the factorisation on FP32 matrix cores (FP64 diagonal blocks, FP64
// storage) and alpha refined in FP64 against K regenerated from the
// series; refine_steps[g] (when given) gets GP g's refinement steps.
// info must be reset (batch_info_reset) before the first launch of a batch.
void batch_fit(hipStream_t s, const BatchGP *gps, const int *order, int count, int *counter, double *losses, int *info,
               int workgroups, bool mixed = false, int *refine_steps = nullptr);
// info[g] = INT_MAX for g < count (a kernel, so it may be captured in a graph).
void batch_info_reset(hipStream_t s, int *info, int count);
// Level-synchronous path for large GPs: one launch per step over all of
// them, one workgroup per item (int2 {GP index, block row}); the same block
// functions as the persistent path, so results are bit-identical. Order:
// Assemble (items: every 64-row block of every GP), then for jb = 0, 1, ...:
// Diagonal (items: GPs with more than jb blocks) and OffDiagonal (items:
// blocks ib > jb of those GPs), then Finish (items: every GP; alpha and
// the loss) or, for mixed precision, the refinement steps below.
// FP64 OffDiagonal launches also take items {GP, -1}: the first jb sums of
// the GP's diagonal block jb + 1 (for jb >= 1), which the Diagonal launch of
// column jb + 1 then continues (the GP needs `pre`); the rounding is that
// of the whole sum in one launch.
// Diagonal128 / OffDiagonal128: the same steps on 128-wide block columns
// (FP64 only; each GP needs W128; its rounding differs from the 64-wide
// steps').
// Mixed precision ends with the refinement of alpha instead of Finish:
// RefineStart (items: every GP; alpha = K^-1 y from the factor), then for
// it = 0 .. BATCH_REFINE_STEPS - 1 RefineResidual (items: every 64-row
// block of every GP) and RefineCorrect (items: every GP, jb = it), then
// RefineFinish (items: every GP; the loss and the step count). Each GP
// needs `pre` (BATCH_REFINE_STATE_DOUBLES) and `work`.
enum class BatchLevelStep
{
    Assemble,
    Diagonal,
    OffDiagonal,
    Diagonal128,
    OffDiagonal128,
    Finish,
    Finish128,  // Finish after 128-wide steps (their diagonal steps leave alpha's forward solve to it)
    RefineStart,
    RefineResidual,
    RefineCorrect,
    RefineFinish
};
void batch_level_step(hipStream_t s, BatchLevelStep step, const BatchGP *gps, const int2 *items, int n_items, int jb, bool mixed,
                      double *losses, int *info, int *refine_steps);

// ---- triangular family (k_triangular.hip) ---------------------------------

// One launch of a diagonal block's task schedule (panel_plan.hpp): the
// nb x nb block diagonal block of A at (p0, p0) is factored in place (lower
// triangle), the row-major inverses of its 64x64 diagonal blocks go to
// Winv + (p0 / 64 + k) * 4096 and its whole inverse, k-major (W(r, c) at
// Wb[c * ldw + r], zero above the diagonal), to Wb; tmp holds the plan's
// tmp_doubles. info (device int) takes the smallest global column whose
// pivot is not positive. Head tasks (build_head_plan) read the block's rows
// of the previous panel (columns hp0 .., nbh blocks wide) and that panel's
// k-major inverse Wh (ld ldwh), and use Xh (ld ldx, nb*64 x nbh*64) as
// scratch. tasks (4 ints each) and wg_off (groups + 1 offsets) are device
// copies of the plan's arrays for this launch.
struct PanelTaskArgs
{
    double *A;
    long lda;
    long p0;
    int nb;
    double *Winv;
    double *Wb;
    long ldw;
    double *tmp;
    int *info;
    long hp0;
    int nbh;
    const double *Wh;
    long ldwh;
    double *Xh;
    long ldx;
};
void panel_tasks(hipStream_t s, const PanelTaskArgs &a, const int *tasks, const int *wg_off, int groups);

// Rows below a factored diagonal block: X = A W^T for A[r0 : r0 + rows,
// p0 : p0 + width] with W the block's k-major inverse (W(r, c) at
// Wb[c * ldw + r], lower triangular), into X (row-major, ld ldx). rows and
// width are multiples of 64.
void rows_solve(hipStream_t s, const double *A, long lda, long r0, long p0, int rows, int width, const double *Wb, long ldw, double *X, long ldx);

// Diagonal step of the blocked TRSM: V = W B for W lower triangular,
// k-major (W(r, c) at Wk[c * ldw + r]), rows x rows; B, V row-major, rows
// x cols. rows and cols are multiples of 64.
void leaf_solve(hipStream_t s, const double *Wk, long ldw, int rows, const double *B, long ldb, int cols, double *V, long ldv);

// Block steps of the vector solves with the saved diagonal-block inverses
// Winv (block J at Winv + J*64*64). Forward: out[J] = W_JJ rhs[J],
// rhs[later] -= L[later, J] out[J]; run J = 0 .. n_pad/64-1 in order.
// Backward (L^T x = rhs): out[J] = W_JJ^T rhs[J], rhs[earlier] -= L[J, earlier]^T out[J];
// run J descending.
void solve_lower_step(hipStream_t s, const double *L, long lda, const double *Winv, double *rhs, double *out, int n_pad, int j);
void solve_upper_step(hipStream_t s, const double *L, long lda, const double *Winv, double *rhs, double *out, int n_pad, int j);

// Dense FP64 matrix-vector products for the wide-block solves (row-major M):
// coldot: out[c] = (or -=) sum_{r<K} M[r][c] v[r], c < N;
// rowdot: out[i] = (or -=) sum_{c<K} M[i][c] v[c], i < N, K % 64 == 0.
// v must not overlap out.
void coldot(hipStream_t s, const double *M, long ld, int K, int N, const double *v, double *out, bool sub);
void rowdot(hipStream_t s, const double *M, long ld, int K, int N, const double *v, double *out, bool sub);

// Inverses of the 256-wide diagonal blocks of L covering rows
// [row0, row0 + rows) (the last block may be narrower), built from the
// 64-wide inverses Winv by W_ij = -W_ii sum_{t=j}^{i-1} L_it W_tj over
// 64-blocks. Block b is stored k-major at Wt + b * block_stride with leading
// dimension ldw (W(r, c) at [c * ldw + r]), zero above the diagonal and past
// the block's last row within its 256 x 256 square.
// zero_above (only with block_stride = 256 * ldw + 256, i.e. one inverse of
// the whole range stored with its blocks along the diagonal): each block
// also zeroes the rows above it in its columns, the inverse's upper part.
void invert_blocks256(hipStream_t s, const double *L, long lda, const double *Winv, long row0, long rows, double *Wt, long ldw,
                      long block_stride, bool zero_above = false);

// colsq[j] += sum_{i<rows} V[i][j]^2 for j < m (fixed summation order;
// FP64 sums for FP32 V too).
void colsq_accumulate(hipStream_t s, const double *V, long ldv, int rows, int m, double *colsq);
void colsq_accumulate_f32(hipStream_t s, const float *V, long ldv, int rows, int m, double *colsq);
// dst[i] = (float) src[i]
void to_f32(hipStream_t s, const double *src, float *dst, long count);

// colsq[j] = sum_{i<n} V[i][j]^2 for j < m (deterministic two-pass);
// `part` holds colsq_splits(n, m) * m doubles.
int colsq_splits(int n, int m);
void column_sumsq(hipStream_t s, const double *V, long ldv, int n, int m, double *part, int splits, double *colsq);

// var[j] = v - colsq[j]
void variance_finish(hipStream_t s, const double *colsq, double v, int m, double *var);

// out[0] = y^T alpha, out[1] = sum_i log(L_ii^2), out[2] = alpha^T alpha, over i < n.
void loss_terms(hipStream_t s, const double *L, long lda, const double *y, const double *alpha, int n, double *out);

// r[i] = y[i] - s[i] - noise * alpha[i] (i < n), the residual of K alpha = y
// given s = K_noiseless alpha.
void residual_finish(hipStream_t s, const double *y, const double *kx, const double *alpha, double noise, int n, double *r);
// y[i] += x[i], i < n.
void add_vec(hipStream_t s, double *y, const double *x, long n);

// Fill a device vector.
void fill(hipStream_t s, double *p, long count, double value);
// The same with at most 8 workgroups, for streams that run beside a
// saturating kernel (a wide launch would wait for free slots).
void fill_narrow(hipStream_t s, double *p, long count, double value);
// A[i][i] = value for i < n.
void set_diagonal(hipStream_t s, double *A, long lda, int n, double value);
// Store one int (used to reset the pivot-failure flag without a memset node).
void set_int(hipStream_t s, int *p, int value);

// ---- FP64 refinement of the variance solves (k_refine.hip) ---------------

// WT[b][k][i] = (float) W[b][i][k] for `count` square nb x nb blocks stored
// one after another (nb a multiple of 64).
void transpose_blocks_f32(hipStream_t s, const double *W, float *WT, int nb, int count);
// R[k][i] = (float) A[(r0 + k) * lda + i] for k < rows, i < cols: a row block
// of A as the k-major A operand of gemm_f32_tn.
void rows_to_f32(hipStream_t s, const double *A, long lda, long r0, int rows, int cols, float *R, long ldr);
// dst[i] = (double) src[i]
void to_f64(hipStream_t s, const float *src, double *dst, long count);
// Z[i][c] += sum_{k = t0}^{i-1} U[k * ldu + i] X[k][c], t0 = 128 floor(i / 128),
// for i < rows, c < cols (cols % 64 == 0): the part of K_sl X inside the
// 128-row diagonal blocks, K_sl read from the strict upper triangle of U.
void diag_strict_product(hipStream_t s, const double *U, long ldu, const double *X, long ldx, double *Z, long ldz, int rows, int cols);
// Refined quadratic forms, column c < mt: s1 = sum_i B[i][c] X[i][c],
// s2 = sum_i X[i][c] Z[i][c], s3 = sum_i X[i][c]^2 over i < rows (fixed
// order, two passes through `part`: quad_form_splits(rows, m) * 3 * m
// doubles); q = 2 s1 - (2 s2 + d s3); var[c] = v - q and, when corr is
// given, corr[c] = q - q0[c].
int quad_form_splits(int rows, int m);
void quad_form(hipStream_t s, const double *B, const double *X, const double *Z, long ld, int rows, int m, double d, double v, const double *q0,
               int mt, double *part, int splits, double *var, double *corr);

// ---- fixed scratch of the no-allocation entry point (k_noalloc.hip) ------

// Bytes of the per-GPU scratch array.
std::size_t noalloc_scratch_bytes();
// Looks up the array's device address on `device` (once per device; done
// when a context is created there).
void noalloc_scratch_init(int device);
// The array's device address on `device` (after noalloc_scratch_init).
void *noalloc_scratch(int device);

// ---- calibration (k_calibrate.hip) ---------------------------------------

// Back-to-back MFMAs from registers on `workgroups` x 256 threads (FP64
// 16x16x4 or FP32 16x16x4); returns the flops of the launch. out receives one
// double per thread.
long mfma_loop(hipStream_t s, bool fp64, double *out, int workgroups, int iters);
// dst[i] = src[i], i < count (count even, 16-byte aligned pointers).
void copy_f64(hipStream_t s, const double *src, double *dst, long count, int workgroups);

}  // namespace gpamd::k
