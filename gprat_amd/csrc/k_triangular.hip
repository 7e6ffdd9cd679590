// Triangular kernels on gfx950: the 64-wide Cholesky panel step, the
// blocked vector solves, the multi-right-hand-side TRSM base case and the
// loss reduction. All reductions run in a fixed order, so results do not
// depend on scheduling.
#include "kernels.hpp"

namespace gpamd::k
{
namespace
{

constexpr int NB = 64;        // diagonal block edge
constexpr int THREADS = 256;

// Lower triangle of the NB x NB block at (r0, r0) into LDS (strict upper zeroed).
__device__ __forceinline__ void load_lower_block(const double *__restrict__ A, long lda, long r0, double (*D)[NB + 1])
{
    for (int e = threadIdx.x; e < NB * NB; e += THREADS)
    {
        const int r = e / NB, c = e % NB;
        D[r][c] = (c <= r) ? A[(r0 + r) * lda + r0 + c] : 0.0;
    }
}

// grid.x = 1 + ceil((row_end - c0 - NB) / THREADS). Every workgroup factors the
// diagonal block in LDS (identical arithmetic, so identical results);
// workgroup 0 stores L_cc, the others each solve THREADS panel rows
// x L_cc^T = a, one row per thread held in registers.
__global__ __launch_bounds__(THREADS) void k_potrf_trsm64(double *__restrict__ A, long lda, int c0, int row_end, int *__restrict__ info)
{
    __shared__ double D[NB][NB + 1];
    const int tid = threadIdx.x;
    load_lower_block(A, lda, c0, D);
    __syncthreads();
    for (int j = 0; j < NB; j++)
    {
        if (tid == 0)
        {
            const double d = D[j][j];
            if (!(d > 0.0) && blockIdx.x == 0)
                atomicMin(info, c0 + j);
            D[j][j] = sqrt(d);
        }
        __syncthreads();
        if (tid > j && tid < NB)
            D[tid][j] = D[tid][j] / D[j][j];
        __syncthreads();
        const int w = NB - 1 - j;
        for (int e = tid; e < w * w; e += THREADS)
        {
            const int rr = j + 1 + e / w, cc = j + 1 + e % w;
            if (cc <= rr)
                D[rr][cc] -= D[rr][j] * D[cc][j];
        }
        __syncthreads();
    }
    if (blockIdx.x == 0)
    {
        for (int e = tid; e < NB * NB; e += THREADS)
        {
            const int r = e / NB, c = e % NB;
            if (c <= r)
                A[(static_cast<long>(c0) + r) * lda + c0 + c] = D[r][c];
        }
        return;
    }
    const long row = static_cast<long>(c0) + NB + static_cast<long>(blockIdx.x - 1) * THREADS + tid;
    if (row >= row_end)
        return;
    double *prow = A + row * lda + c0;
    double xr[NB];
#pragma unroll
    for (int q = 0; q < NB / 2; q++)
    {
        const double2 t = reinterpret_cast<const double2 *>(prow)[q];
        xr[2 * q] = t.x;
        xr[2 * q + 1] = t.y;
    }
#pragma unroll
    for (int j = 0; j < NB; j++)
    {
        double s = xr[j];
#pragma unroll
        for (int p = 0; p < j; p++)
            s -= xr[p] * D[j][p];
        xr[j] = s / D[j][j];
    }
#pragma unroll
    for (int q = 0; q < NB / 2; q++)
        reinterpret_cast<double2 *>(prow)[q] = make_double2(xr[2 * q], xr[2 * q + 1]);
}

// Forward block step J: every workgroup solves L_JJ x_J = rhs_J redundantly
// (wavefront 0, lane i owns row i); workgroup 0 stores x_J, workgroup w >= 1
// updates rhs_{J+w} -= L_{J+w,J} x_J.
__global__ __launch_bounds__(THREADS) void k_solve_lower_step(const double *__restrict__ L,
                                                               long lda,
                                                               double *__restrict__ rhs,
                                                               double *__restrict__ out,
                                                               int J)
{
    __shared__ double D[NB][NB + 1];
    __shared__ double xs[NB];
    __shared__ double part[4][NB];
    const int tid = threadIdx.x;
    const long base = static_cast<long>(J) * NB;
    load_lower_block(L, lda, base, D);
    if (tid < NB)
        xs[tid] = rhs[base + tid];
    __syncthreads();
    if (tid < NB)
    {
        double r = xs[tid];
        double mine = 0.0;
        for (int p = 0; p < NB; p++)
        {
            const double xp = __shfl(r, p) / D[p][p];
            if (tid == p)
                mine = xp;
            if (tid > p)
                r -= D[tid][p] * xp;
        }
        xs[tid] = mine;
    }
    __syncthreads();
    if (blockIdx.x == 0)
    {
        if (tid < NB)
            out[base + tid] = xs[tid];
        return;
    }
    const long I = static_cast<long>(J) + blockIdx.x;
    const int rr = tid >> 2, q = tid & 3;
    const double *prow = L + (I * NB + rr) * lda + base + q * 16;
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 16; c++)
        s += prow[c] * xs[q * 16 + c];
    part[q][rr] = s;
    __syncthreads();
    if (tid < NB)
        rhs[I * NB + tid] -= (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
}

// Backward block step J: x_J = L_JJ^-T rhs_J, then rhs_{J-w} -= L_{J,J-w}^T x_J.
__global__ __launch_bounds__(THREADS) void k_solve_upper_step(const double *__restrict__ L,
                                                               long lda,
                                                               double *__restrict__ rhs,
                                                               double *__restrict__ out,
                                                               int J)
{
    __shared__ double D[NB][NB + 1];
    __shared__ double xs[NB];
    __shared__ double part[4][NB];
    const int tid = threadIdx.x;
    const long base = static_cast<long>(J) * NB;
    load_lower_block(L, lda, base, D);
    if (tid < NB)
        xs[tid] = rhs[base + tid];
    __syncthreads();
    if (tid < NB)
    {
        double r = xs[tid];
        double mine = 0.0;
        for (int p = NB - 1; p >= 0; p--)
        {
            const double xp = __shfl(r, p) / D[p][p];
            if (tid == p)
                mine = xp;
            if (tid < p)
                r -= D[p][tid] * xp;
        }
        xs[tid] = mine;
    }
    __syncthreads();
    if (blockIdx.x == 0)
    {
        if (tid < NB)
            out[base + tid] = xs[tid];
        return;
    }
    const long I = static_cast<long>(J) - blockIdx.x;
    const int c = tid & 63, q = tid >> 6;
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < 16; r++)
        s += L[(base + q * 16 + r) * lda + I * NB + c] * xs[q * 16 + r];
    part[q][c] = s;
    __syncthreads();
    if (tid < NB)
        rhs[I * NB + tid] -= (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
}

// Solve L_rr X = B[r0:r0+64, :] in place, one right-hand-side column per
// thread held in registers; optionally accumulate the column sums of X^2.
__global__ __launch_bounds__(THREADS) void k_trsm_left64(const double *__restrict__ L,
                                                          long lda,
                                                          double *__restrict__ B,
                                                          long ldb,
                                                          int r0,
                                                          int ncols,
                                                          double *__restrict__ colsq)
{
    __shared__ double D[NB][NB + 1];
    load_lower_block(L, lda, r0, D);
    __syncthreads();
    const long c = static_cast<long>(blockIdx.x) * THREADS + threadIdx.x;
    if (c >= ncols)
        return;
    double xv[NB];
#pragma unroll
    for (int i = 0; i < NB; i++)
        xv[i] = B[(static_cast<long>(r0) + i) * ldb + c];
#pragma unroll
    for (int i = 0; i < NB; i++)
    {
        double s = xv[i];
#pragma unroll
        for (int p = 0; p < i; p++)
            s -= D[i][p] * xv[p];
        xv[i] = s / D[i][i];
    }
#pragma unroll
    for (int i = 0; i < NB; i++)
        B[(static_cast<long>(r0) + i) * ldb + c] = xv[i];
    if (colsq)
    {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < NB; i++)
            acc += xv[i] * xv[i];
        colsq[c] += acc;
    }
}

__global__ void k_variance_finish(const double *__restrict__ colsq, double v, int m, double *__restrict__ var)
{
    const long j = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < m)
        var[j] = v - colsq[j];
}

// Single workgroup, fixed-order tree reduction:
// out[0] = y^T alpha, out[1] = sum log(L_ii^2), out[2] = alpha^T alpha.
__global__ __launch_bounds__(THREADS) void k_loss_terms(const double *__restrict__ L,
                                                         long lda,
                                                         const double *__restrict__ y,
                                                         const double *__restrict__ alpha,
                                                         int n,
                                                         double *__restrict__ out)
{
    __shared__ double s1[THREADS], s2[THREADS], s3[THREADS];
    double a = 0.0, b = 0.0, c = 0.0;
    for (long i = threadIdx.x; i < n; i += THREADS)
    {
        const double al = alpha[i];
        a += y[i] * al;
        c += al * al;
        const double d = L[i * lda + i];
        b += log(d * d);
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    s3[threadIdx.x] = c;
    __syncthreads();
    for (int w = THREADS / 2; w > 0; w >>= 1)
    {
        if (threadIdx.x < w)
        {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
            s3[threadIdx.x] += s3[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        out[0] = s1[0];
        out[1] = s2[0];
        out[2] = s3[0];
    }
}

__global__ void k_set_diagonal(double *__restrict__ A, long lda, int n, double value)
{
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        A[i * lda + i] = value;
}

__global__ void k_fill(double *__restrict__ p, long count, double value)
{
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count)
        p[i] = value;
}

__global__ void k_set_int(int *p, int value) { *p = value; }

}  // namespace

void potrf_trsm64(hipStream_t s, double *A, long lda, int c0, int row_end, int *info)
{
    require(c0 >= 0 && c0 + NB <= row_end && lda % 2 == 0, "potrf_trsm64: bad shape");
    const long below = row_end - c0 - NB;
    const unsigned blocks = static_cast<unsigned>(1 + ceil_div(below, THREADS));
    hipLaunchKernelGGL(k_potrf_trsm64, dim3(blocks), dim3(THREADS), 0, s, A, lda, c0, row_end, info);
    GP_CHECK_LAUNCH();
}

void solve_lower_step(hipStream_t s, const double *L, long lda, double *rhs, double *out, int n_pad, int j)
{
    const int nb = n_pad / NB;
    require(n_pad % NB == 0 && j >= 0 && j < nb, "solve_lower_step: bad shape");
    hipLaunchKernelGGL(k_solve_lower_step, dim3(nb - j), dim3(THREADS), 0, s, L, lda, rhs, out, j);
    GP_CHECK_LAUNCH();
}

void solve_upper_step(hipStream_t s, const double *L, long lda, double *rhs, double *out, int n_pad, int j)
{
    const int nb = n_pad / NB;
    require(n_pad % NB == 0 && j >= 0 && j < nb, "solve_upper_step: bad shape");
    hipLaunchKernelGGL(k_solve_upper_step, dim3(j + 1), dim3(THREADS), 0, s, L, lda, rhs, out, j);
    GP_CHECK_LAUNCH();
}

void trsm_left64(hipStream_t s, const double *L, long lda, double *B, long ldb, int r0, int ncols, double *colsq)
{
    require(r0 >= 0 && ncols >= 0, "trsm_left64: bad shape");
    if (ncols == 0)
        return;
    hipLaunchKernelGGL(k_trsm_left64, dim3(ceil_div(ncols, THREADS)), dim3(THREADS), 0, s, L, lda, B, ldb, r0, ncols, colsq);
    GP_CHECK_LAUNCH();
}

void variance_finish(hipStream_t s, const double *colsq, double v, int m, double *var)
{
    if (m <= 0)
        return;
    hipLaunchKernelGGL(k_variance_finish, dim3(ceil_div(m, 256)), dim3(256), 0, s, colsq, v, m, var);
    GP_CHECK_LAUNCH();
}

void loss_terms(hipStream_t s, const double *L, long lda, const double *y, const double *alpha, int n, double *out)
{
    hipLaunchKernelGGL(k_loss_terms, dim3(1), dim3(THREADS), 0, s, L, lda, y, alpha, n, out);
    GP_CHECK_LAUNCH();
}

void set_diagonal(hipStream_t s, double *A, long lda, int n, double value)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_set_diagonal, dim3(ceil_div(n, 256)), dim3(256), 0, s, A, lda, n, value);
    GP_CHECK_LAUNCH();
}

void fill(hipStream_t s, double *p, long count, double value)
{
    if (count <= 0)
        return;
    hipLaunchKernelGGL(k_fill, dim3(ceil_div(count, 256)), dim3(256), 0, s, p, count, value);
    GP_CHECK_LAUNCH();
}

void set_int(hipStream_t s, int *p, int value)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, p, value);
    GP_CHECK_LAUNCH();
}

}  // namespace gpamd::k
