// FP64 refinement of the predictive-variance solves of mixed-precision
// models on gfx950.
//
// The latent variance at test point c is v - q_c with q_c = k_c^T K^-1 k_c.
// From an approximate solution x~ = K~^-1 k_c (FP32 forward and backward
// solves with the mixed-precision factor), one FP64 residual r = k - K x~
// gives the corrected form q1 = k^T x~ + x~^T r = 2 k^T x~ - x~^T K x~, whose
// error is e^T K e for e = K^-1 k - x~: the first-order error of the
// solves cancels. K x~ is needed only through x~^T K x~ =
// 2 x~^T K_sl x~ + d |x~|^2 (K_sl the strictly lower part, d = v + noise the
// diagonal), so the FP64 product is the triangular Z = K_sl X~ (half the
// flops of a full one). K_sl is read from the strict upper triangle of the
// factor's storage, where the fit writes the covariance after the
// factorisation: A(i, k) = U[k][i] for k < i is the k-major operand layout
// of the FP64 matrix-core product. The product skips each 128-row tile's
// own diagonal block (the lower triangle there holds the factor);
// `k_diag_strict_product` adds those blocks.
//
// This file holds the helpers of that scheme: FP32 operand copies for the
// backward solve, the diagonal-block part of Z, and the column reductions.
#include "kernels.hpp"

namespace gpamd::k
{
namespace
{

constexpr int THREADS = 256;

// WT[b][k][i] = (float) W[b][i][k] through a 64x64 LDS tile (coalesced
// reads and writes).
__global__ __launch_bounds__(THREADS) void k_transpose_blocks_f32(const double *__restrict__ W, float *__restrict__ WT, int nb)
{
    __shared__ float t[64][65];
    const long base = static_cast<long>(blockIdx.z) * nb * nb;
    const int i0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
    const int c = threadIdx.x & 63, q = threadIdx.x >> 6;
    for (int rr = q; rr < 64; rr += 4)
        t[rr][c] = static_cast<float>(W[base + static_cast<long>(i0 + rr) * nb + k0 + c]);  // W[i0 + rr][k0 + c]
    __syncthreads();
    for (int kk = q; kk < 64; kk += 4)
        WT[base + static_cast<long>(k0 + kk) * nb + i0 + c] = t[c][kk];  // WT[k0 + kk][i0 + c] = W[i0 + c][k0 + kk]
}

// R[k][i] = (float) A[(r0 + k) * lda + i], k < rows, i < cols; four
// columns per thread (two 16-byte loads, one 16-byte store).
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(THREADS) void k_rows_to_f32(const double *__restrict__ A, long lda, long r0, int rows, int cols,
                                                          float *__restrict__ R, long ldr)
{
    const int k = blockIdx.y;
    const double *src = A + (r0 + k) * lda;
    float *dst = R + k * ldr;
    for (long i = (static_cast<long>(blockIdx.x) * THREADS + threadIdx.x) * 4; i < cols; i += static_cast<long>(gridDim.x) * THREADS * 4)
    {
        const d2v a = *reinterpret_cast<const d2v *>(src + i);
        const d2v b = *reinterpret_cast<const d2v *>(src + i + 2);
        *reinterpret_cast<f4v *>(dst + i) = f4v{ static_cast<float>(a.x), static_cast<float>(a.y), static_cast<float>(b.x), static_cast<float>(b.y) };
    }
}

__global__ __launch_bounds__(THREADS) void k_to_f64(const float *__restrict__ src, double *__restrict__ dst, long count)
{
    for (long i = static_cast<long>(blockIdx.x) * THREADS + threadIdx.x; i < count; i += static_cast<long>(gridDim.x) * THREADS)
        dst[i] = static_cast<double>(src[i]);
}

// Z[i][c] += sum_{k = r0}^{i-1} U[k * ldu + i] X[k][c] for the 128-row tile
// [r0, r0 + 128) of workgroup row blockIdx.y and the 64 columns of
// blockIdx.x, on the FP64 matrix cores. K (its strictly lower part, zeros
// elsewhere) and X are staged 16 k-rows at a time through LDS, k-major;
// wavefront w accumulates rows 32 w .. 32 w + 31 of the tile over all 64
// columns (two 16-row by four 16-column v_mfma_f64_16x16x4f64 tiles) and
// skips the k-rows at or below its last row, whose K entries are zero.
__global__ __launch_bounds__(THREADS) void k_diag_strict_product(const double *__restrict__ U, long ldu, const double *__restrict__ X,
                                                                  long ldx, double *__restrict__ Z, long ldz, int rows)
{
    typedef double v4d __attribute__((ext_vector_type(4)));
    __shared__ double Us[16][128];
    __shared__ double Xs[16][64];
    const long r0 = static_cast<long>(blockIdx.y) * 128, c0 = static_cast<long>(blockIdx.x) * 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const int span = static_cast<int>(min(static_cast<long>(128), rows - r0));
    const int wend = 32 * w + 32;  // k-rows of this wavefront's products: k < wend
    v4d acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
            acc[i][j] = v4d{ 0.0, 0.0, 0.0, 0.0 };
    for (int kk0 = 0; kk0 < span; kk0 += 16)
    {
        // U rows r0 + kk0 .. +16, columns r0 .. r0 + 128: K(i, k) where k < i
        for (int e = threadIdx.x; e < 16 * 128; e += THREADS)
        {
            const int kr = e >> 7, ic = e & 127, k = kk0 + kr;
            Us[kr][ic] = (k < ic && ic < span) ? U[(r0 + k) * ldu + r0 + ic] : 0.0;
        }
        for (int e = threadIdx.x; e < 16 * 64; e += THREADS)
        {
            const int kr = e >> 6, cc = e & 63;
            Xs[kr][cc] = (kk0 + kr < span) ? X[(r0 + kk0 + kr) * ldx + c0 + cc] : 0.0;
        }
        __syncthreads();
        if (kk0 < wend)
        {
#pragma unroll
            for (int k4 = 0; k4 < 16; k4 += 4)
            {
                double a[2], b[4];
#pragma unroll
                for (int i = 0; i < 2; i++)
                    a[i] = Us[k4 + lk][32 * w + 16 * i + lm];
#pragma unroll
                for (int j = 0; j < 4; j++)
                    b[j] = Xs[k4 + lk][16 * j + lm];
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // accumulator layout: lane holds column lane % 16, rows lane / 16 + 4 e
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int e = 0; e < 4; e++)
        {
            const int r = 32 * w + 16 * i + lk + 4 * e;
            if (r < span)
            {
                double *zr = Z + (r0 + r) * ldz + c0 + lm;
                double cur[4];
#pragma unroll
                for (int j = 0; j < 4; j++)
                    cur[j] = zr[16 * j];
#pragma unroll
                for (int j = 0; j < 4; j++)
                    zr[16 * j] = cur[j] + acc[i][j][e];
            }
        }
}

// Partial column sums over the row range of split blockIdx.y:
// part[(split * 3 + t) * m + c], t = 0: sum B X, 1: sum X Z, 2: sum X^2.
// 64 columns per workgroup; its four wavefronts take interleaved rows and
// their sums are combined in a fixed order.
__global__ __launch_bounds__(THREADS) void k_quad_form_partial(const double *__restrict__ B, const double *__restrict__ X,
                                                                const double *__restrict__ Z, long ld, int rows, int m, int per_split,
                                                                double *__restrict__ part)
{
    __shared__ double red[3][4][64];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const long j = static_cast<long>(blockIdx.x) * 64 + c;
    const int split = blockIdx.y;
    const long i0 = static_cast<long>(split) * per_split, i1 = min(static_cast<long>(rows), i0 + per_split);
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (j < m)
        for (long i = i0 + g; i < i1; i += 4)
        {
            const double x = X[i * ld + j];
            s1 = fma(B[i * ld + j], x, s1);
            s2 = fma(Z[i * ld + j], x, s2);
            s3 = fma(x, x, s3);
        }
    red[0][g][c] = s1;
    red[1][g][c] = s2;
    red[2][g][c] = s3;
    __syncthreads();
    if (g < 3 && j < m)
    {
        const double t = ((red[g][0][c] + red[g][1][c]) + red[g][2][c]) + red[g][3][c];
        part[(static_cast<long>(split) * 3 + g) * m + j] = t;
    }
}

__global__ __launch_bounds__(THREADS) void k_quad_form_finish(const double *__restrict__ part, int splits, int m, int mt, double d, double v,
                                                               const double *__restrict__ q0, double *__restrict__ var,
                                                               double *__restrict__ corr)
{
    const int j = blockIdx.x * THREADS + threadIdx.x;
    if (j >= mt)
        return;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int sp = 0; sp < splits; sp++)
    {
        s1 += part[(static_cast<long>(sp) * 3 + 0) * m + j];
        s2 += part[(static_cast<long>(sp) * 3 + 1) * m + j];
        s3 += part[(static_cast<long>(sp) * 3 + 2) * m + j];
    }
    const double q = 2.0 * s1 - (2.0 * s2 + d * s3);
    var[j] = v - q;
    if (corr)
        corr[j] = q - q0[j];
}

}  // namespace

void transpose_blocks_f32(hipStream_t s, const double *W, float *WT, int nb, int count)
{
    require(nb > 0 && nb % 64 == 0 && count >= 1 && count <= 65535, "transpose_blocks_f32: bad shape");
    dim3 grid(nb / 64, nb / 64, count);
    hipLaunchKernelGGL(k_transpose_blocks_f32, grid, dim3(THREADS), 0, s, W, WT, nb);
    GP_CHECK_LAUNCH();
}

void rows_to_f32(hipStream_t s, const double *A, long lda, long r0, int rows, int cols, float *R, long ldr)
{
    require(rows >= 0 && cols >= 0 && ldr >= cols && rows <= 65535, "rows_to_f32: bad shape");
    require(cols % 4 == 0 && lda % 2 == 0 && ldr % 4 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0
                && reinterpret_cast<uintptr_t>(R) % 16 == 0,
            "rows_to_f32: rows must come in aligned groups of four");
    if (rows == 0 || cols == 0)
        return;
    dim3 grid(static_cast<unsigned>(std::min<long>(ceil_div(cols, 4 * THREADS), 64)), static_cast<unsigned>(rows));
    hipLaunchKernelGGL(k_rows_to_f32, grid, dim3(THREADS), 0, s, A, lda, r0, rows, cols, R, ldr);
    GP_CHECK_LAUNCH();
}

void to_f64(hipStream_t s, const float *src, double *dst, long count)
{
    if (count <= 0)
        return;
    const unsigned blocks = static_cast<unsigned>(std::min<long>(ceil_div(count, THREADS), 8192));
    hipLaunchKernelGGL(k_to_f64, dim3(blocks), dim3(THREADS), 0, s, src, dst, count);
    GP_CHECK_LAUNCH();
}

void diag_strict_product(hipStream_t s, const double *U, long ldu, const double *X, long ldx, double *Z, long ldz, int rows, int cols)
{
    require(rows >= 0 && cols >= 0 && cols % 64 == 0 && ldx >= cols && ldz >= cols && ldu >= rows, "diag_strict_product: bad shape");
    if (rows == 0 || cols == 0)
        return;
    dim3 grid(static_cast<unsigned>(cols / 64), static_cast<unsigned>(ceil_div(rows, 128)));
    require(grid.y <= 65535, "diag_strict_product: too many rows");
    hipLaunchKernelGGL(k_diag_strict_product, grid, dim3(THREADS), 0, s, U, ldu, X, ldx, Z, ldz, rows);
    GP_CHECK_LAUNCH();
}

int quad_form_splits(int rows, int m)
{
    const long col_blocks = ceil_div(std::max(m, 1), 64);
    long splits = ceil_div(2048, col_blocks);
    splits = std::min<long>(splits, ceil_div(std::max(rows, 1), 256));
    return static_cast<int>(std::max<long>(std::min<long>(splits, 65535), 1));
}

void quad_form(hipStream_t s, const double *B, const double *X, const double *Z, long ld, int rows, int m, double d, double v, const double *q0,
               int mt, double *part, int splits, double *var, double *corr)
{
    require(rows >= 0 && m >= 0 && mt <= m && ld >= m && splits >= 1 && (!corr || q0), "quad_form: bad shape");
    if (mt == 0)
        return;
    const int per_split = static_cast<int>(ceil_div(std::max(rows, 1), splits));
    dim3 grid(static_cast<unsigned>(ceil_div(m, 64)), static_cast<unsigned>(splits));
    hipLaunchKernelGGL(k_quad_form_partial, grid, dim3(THREADS), 0, s, B, X, Z, ld, rows, m, per_split, part);
    GP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_quad_form_finish, dim3(static_cast<unsigned>(ceil_div(mt, THREADS))), dim3(THREADS), 0, s, part, splits, m, mt, d, v,
                       q0, var, corr);
    GP_CHECK_LAUNCH();
}

}  // namespace gpamd::k
