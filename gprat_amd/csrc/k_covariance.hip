// SE covariance generation on gfx950: training covariance (lower band),
// transposed cross-covariance, the fused predictive mean and the tile gather
// for cholesky(). All are HBM-write-bound or exp-bound streaming kernels;
// 64-lane wavefronts own 64 consecutive columns so every store is a fully
// coalesced 512-byte row segment.
#include "kernels.hpp"

namespace gpamd::k
{
namespace
{

constexpr int TB = 64;       // tile edge
constexpr int THREADS = 256;  // 4 wavefronts, each covering 16 rows of the tile

// SE entry as core/src/cpu/gp_algorithms.cpp:11-28 evaluates it.
__device__ __forceinline__ double se(double scale, double v, double d) { return v * exp(scale * d); }

__global__ __launch_bounds__(THREADS) void k_assemble_lower(const double *__restrict__ x,
                                                             int r,
                                                             double scale,
                                                             double v,
                                                             double noise,
                                                             double *__restrict__ A,
                                                             long lda,
                                                             int n)
{
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (tj > ti)
        return;
    const int c = threadIdx.x % TB;
    const long j = static_cast<long>(tj) * TB + c;
    for (int rr = threadIdx.x / TB; rr < TB; rr += THREADS / TB)
    {
        const long i = static_cast<long>(ti) * TB + rr;
        double val;
        if (i < n && j < n)
        {
            val = se(scale, v, window_sqdist(x, i, x, j, r));
            if (i == j)
                val += noise;
        }
        else
        {
            val = (i == j) ? 1.0 : 0.0;
        }
        A[i * lda + j] = val;
    }
}

// The same for window lengths R <= 16: a lane keeps its column's window in
// registers, and the row's window (one row per wavefront) is read through
// the scalar cache. Same operations in the same order as window_sqdist.
// UPPER: the 64x64 blocks on or above the diagonal instead, and only the
// elements strictly above it (K(i, j), i < j; zero in the padding).
template <int R, bool UPPER = false>
__global__ __launch_bounds__(THREADS) void k_assemble_lower_w(const double *__restrict__ x,
                                                               double scale,
                                                               double v,
                                                               double noise,
                                                               double *__restrict__ A,
                                                               long lda,
                                                               int n)
{
#pragma clang fp contract(off)
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (UPPER ? tj < ti : tj > ti)
        return;
    const int c = threadIdx.x % TB;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / TB);
    const int j = tj * TB + c;
    const int jc = j < n ? j : n - 1;  // lanes past the end read a valid window and store the padding value
    double xj[R];
#pragma unroll
    for (int k = 0; k < R; k++)
        xj[k] = x[jc + k];
    for (int rr = w; rr < TB; rr += THREADS / TB)
    {
        const int i = ti * TB + rr;
        double val;
        if (i < n && j < n)
        {
            double d = 0.0;
#pragma unroll
            for (int k = 0; k < R; k++)
            {
                const double t = x[i + k] - xj[k];
                d += t * t;
            }
            val = se(scale, v, d);
            if (i == j)
                val += noise;
        }
        else
            val = (i == j) ? 1.0 : 0.0;
        if (!UPPER || i < j)
            A[static_cast<long>(i) * lda + j] = val;
    }
}

template <class T, int R>
__global__ __launch_bounds__(THREADS) void k_assemble_cross_t_w(const double *__restrict__ x,
                                                                 const double *__restrict__ z,
                                                                 double scale,
                                                                 double v,
                                                                 T *__restrict__ B,
                                                                 long ldb,
                                                                 int n,
                                                                 int m)
{
#pragma clang fp contract(off)
    const int ti = blockIdx.y, tj = blockIdx.x;
    const int c = threadIdx.x % TB;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / TB);
    const int j = tj * TB + c;  // test index
    const int jc = j < m ? j : m - 1;
    double zj[R];
#pragma unroll
    for (int k = 0; k < R; k++)
        zj[k] = z[jc + k];
    for (int rr = w; rr < TB; rr += THREADS / TB)
    {
        const int i = ti * TB + rr;  // train index
        double val = 0.0;
        if (i < n && j < m)
        {
            double d = 0.0;
#pragma unroll
            for (int k = 0; k < R; k++)
            {
                const double t = zj[k] - x[i + k];  // (test - train), as gp_algorithms.cpp:113-140
                d += t * t;
            }
            val = se(scale, v, d);
        }
        B[static_cast<long>(i) * ldb + j] = static_cast<T>(val);
    }
}

template <class T>
__global__ __launch_bounds__(THREADS) void k_assemble_cross_t(const double *__restrict__ x,
                                                               const double *__restrict__ z,
                                                               int r,
                                                               double scale,
                                                               double v,
                                                               T *__restrict__ B,
                                                               long ldb,
                                                               int n,
                                                               int m)
{
    const int ti = blockIdx.y, tj = blockIdx.x;
    const int c = threadIdx.x % TB;
    const long j = static_cast<long>(tj) * TB + c;  // test index
    for (int rr = threadIdx.x / TB; rr < TB; rr += THREADS / TB)
    {
        const long i = static_cast<long>(ti) * TB + rr;  // train index
        double val = 0.0;
        if (i < n && j < m)
            val = se(scale, v, window_sqdist(z, j, x, i, r));  // (test - train), as gp_algorithms.cpp:113-140
        B[i * ldb + j] = static_cast<T>(val);
    }
}

// Grid (ceil(m/64), splits). Thread (lane = test point, wave = train stride
// class) sums k(z_j, x_i) alpha_i over its share of split s; the four wave
// partials are combined in a fixed order, so the result is deterministic.
__global__ __launch_bounds__(THREADS) void k_mean_partial(const double *__restrict__ x,
                                                           const double *__restrict__ z,
                                                           const double *__restrict__ alpha,
                                                           int r,
                                                           double scale,
                                                           double v,
                                                           int n,
                                                           int m,
                                                           int chunk,
                                                           double *__restrict__ partial)
{
    __shared__ double red[THREADS];
    const int lane = threadIdx.x % TB, w = threadIdx.x / TB;
    const long j = static_cast<long>(blockIdx.x) * TB + lane;
    const long i0 = static_cast<long>(blockIdx.y) * chunk;
    const long i1 = min(static_cast<long>(n), i0 + chunk);
    double acc = 0.0;
    if (j < m)
        for (long i = i0 + w; i < i1; i += THREADS / TB)
            acc += se(scale, v, window_sqdist(z, j, x, i, r)) * alpha[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    if (w == 0 && j < m)
    {
        double s = red[lane];
        for (int q = 1; q < THREADS / TB; q++)
            s += red[q * TB + lane];
        partial[static_cast<long>(blockIdx.y) * m + j] = s;
    }
}

// Window lengths up to 16 known at compile time: the lane's test window is
// held in registers for the whole chunk, the train index is wavefront-
// uniform (so the train window and alpha arrive through scalar loads and
// feed the FP64 VALU as scalar operands), and each wavefront walks its own
// contiguous quarter of the chunk, two train points per step. The four
// wavefront partials are combined in a fixed order, as above.
template <int R>
__global__ __launch_bounds__(THREADS) void k_mean_partial_w(const double *__restrict__ x,
                                                             const double *__restrict__ z,
                                                             const double *__restrict__ alpha,
                                                             double scale,
                                                             double v,
                                                             int n,
                                                             int m,
                                                             int chunk,
                                                             double *__restrict__ partial)
{
    __shared__ double red[THREADS];
    const int lane = threadIdx.x % TB;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / TB);
    const int j = blockIdx.x * TB + lane;
    const int jc = j < m ? j : m - 1;  // lanes past the end read a valid window and discard the sum
    double zw[R];
#pragma unroll
    for (int k = 0; k < R; k++)
        zw[k] = z[jc + k];
    const int c0 = blockIdx.y * chunk;
    const int c1 = min(n, c0 + chunk);
    const int share = (chunk + 3) / 4;
    const int i0 = min(c1, c0 + w * share);
    const int i1 = min(c1, i0 + share);
    double acc0 = 0.0, acc1 = 0.0;
    int i = i0;
    for (; i + 1 < i1; i += 2)
    {
        double d0 = 0.0, d1 = 0.0;
#pragma unroll
        for (int k = 0; k < R; k++)
        {
            const double t0 = zw[k] - x[i + k];
            const double t1 = zw[k] - x[i + 1 + k];
            d0 = fma(t0, t0, d0);
            d1 = fma(t1, t1, d1);
        }
        acc0 = fma(se(scale, v, d0), alpha[i], acc0);
        acc1 = fma(se(scale, v, d1), alpha[i + 1], acc1);
    }
    if (i < i1)
    {
        double d0 = 0.0;
#pragma unroll
        for (int k = 0; k < R; k++)
        {
            const double t0 = zw[k] - x[i + k];
            d0 = fma(t0, t0, d0);
        }
        acc0 = fma(se(scale, v, d0), alpha[i], acc0);
    }
    red[threadIdx.x] = acc0 + acc1;
    __syncthreads();
    if (w == 0 && j < m)
    {
        double s = red[lane];
        for (int q = 1; q < THREADS / TB; q++)
            s += red[q * TB + lane];
        partial[static_cast<long>(blockIdx.y) * m + j] = s;
    }
}

template <int R>
void launch_mean_w(hipStream_t s, dim3 grid, const double *x, const double *z, const double *alpha, double scale, double v, int n, int m,
                   int chunk, double *partial)
{
    hipLaunchKernelGGL(k_mean_partial_w<R>, grid, dim3(THREADS), 0, s, x, z, alpha, scale, v, n, m, chunk, partial);
}

using MeanLaunch = void (*)(hipStream_t, dim3, const double *, const double *, const double *, double, double, int, int, int, double *);
constexpr MeanLaunch mean_w_table[16] = { launch_mean_w<1>,  launch_mean_w<2>,  launch_mean_w<3>,  launch_mean_w<4>,
                                          launch_mean_w<5>,  launch_mean_w<6>,  launch_mean_w<7>,  launch_mean_w<8>,
                                          launch_mean_w<9>,  launch_mean_w<10>, launch_mean_w<11>, launch_mean_w<12>,
                                          launch_mean_w<13>, launch_mean_w<14>, launch_mean_w<15>, launch_mean_w<16> };

__global__ __launch_bounds__(THREADS) void k_mean_reduce(const double *__restrict__ partial, int splits, int m, double *__restrict__ mean)
{
    const long j = static_cast<long>(blockIdx.x) * THREADS + threadIdx.x;
    if (j >= m)
        return;
    double s = 0.0;
    for (int q = 0; q < splits; q++)
        s += partial[static_cast<long>(q) * m + j];
    mean[j] = s;
}

// grid.x = lower tile index t = i(i+1)/2 + j, grid.y = row chunk of the tile.
__global__ __launch_bounds__(THREADS) void k_gather_tiles(const double *__restrict__ L,
                                                           long lda,
                                                           const double *__restrict__ x,
                                                           int r,
                                                           double scale,
                                                           double v,
                                                           int b,
                                                           int rows_per_block,
                                                           double *__restrict__ out)
{
    const long t = blockIdx.x;
    // invert t = i(i+1)/2 + j
    long i = static_cast<long>((sqrt(8.0 * static_cast<double>(t) + 1.0) - 1.0) / 2.0);
    while (i * (i + 1) / 2 > t)
        i--;
    while ((i + 1) * (i + 2) / 2 <= t)
        i++;
    const long j = t - i * (i + 1) / 2;
    const long rbeg = static_cast<long>(blockIdx.y) * rows_per_block;
    const long rend = min(static_cast<long>(b), rbeg + rows_per_block);
    double *tile = out + t * static_cast<long>(b) * b;
    for (long rr = rbeg; rr < rend; rr++)
        for (long cc = threadIdx.x; cc < b; cc += THREADS)
        {
            const long gi = i * b + rr, gj = j * b + cc;
            double val;
            if (i == j && cc > rr)
                val = se(scale, v, window_sqdist(x, gi, x, gj, r));  // strict upper: original K, no noise
            else
                val = L[gi * lda + gj];
            tile[rr * b + cc] = val;
        }
}

// Gradient sums over the lower triangle of K^-1 (one 64x64 tile per
// workgroup), with dK regenerated from the series instead of stored:
//   D = -0.5/(l*l) d2, e = exp(D), dK_l = (-2 v / l) D e s_l, dK_v = e s_v
// part[t][0..4] = sum w Kinv dK_l, sum w Kinv dK_v, sum w a_i a_j dK_l,
//                 sum w a_i a_j dK_v, sum_diag Kinv   (w = 2 off the diagonal).
__global__ __launch_bounds__(THREADS) void k_grad_partial(const double *__restrict__ Kinv,
                                                           long lda,
                                                           const double *__restrict__ x,
                                                           const double *__restrict__ alpha,
                                                           int r,
                                                           double scale,
                                                           double fl,
                                                           double sv,
                                                           int n,
                                                           int tiles,
                                                           double *__restrict__ part)
{
    __shared__ double red[5][THREADS];
    const long t = blockIdx.x;
    long ti = static_cast<long>((sqrt(8.0 * static_cast<double>(t) + 1.0) - 1.0) / 2.0);
    while (ti * (ti + 1) / 2 > t)
        ti--;
    while ((ti + 1) * (ti + 2) / 2 <= t)
        ti++;
    const long tj = t - ti * (ti + 1) / 2;
    const int c = threadIdx.x % TB;
    const long j = tj * TB + c;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
    if (j < n)
    {
        const double aj = alpha[j];
        for (int rr = threadIdx.x / TB; rr < TB; rr += THREADS / TB)
        {
            const long i = ti * TB + rr;
            if (i >= n || i < j)
                continue;
            const double kij = Kinv[i * lda + j];
            const double D = scale * window_sqdist(x, i, x, j, r);
            const double e = exp(D);
            const double dkl = fl * D * e;
            const double dkv = e * sv;
            const double w = (i == j) ? 1.0 : 2.0;
            const double aa = alpha[i] * aj;
            s0 += w * kij * dkl;
            s1 += w * kij * dkv;
            s2 += w * aa * dkl;
            s3 += w * aa * dkv;
            if (i == j)
                s4 += kij;
        }
    }
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    red[2][threadIdx.x] = s2;
    red[3][threadIdx.x] = s3;
    red[4][threadIdx.x] = s4;
    __syncthreads();
    for (int w = THREADS / 2; w > 0; w >>= 1)
    {
        if (threadIdx.x < w)
            for (int q = 0; q < 5; q++)
                red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 5)
        part[t * 5 + threadIdx.x] = red[threadIdx.x][0];
}

// out[q] = sum_t part[t][q], fixed order (single workgroup).
__global__ __launch_bounds__(THREADS) void k_grad_reduce(const double *__restrict__ part, int tiles, double *__restrict__ out)
{
    __shared__ double red[5][THREADS];
    double s[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
    for (long t = threadIdx.x; t < tiles; t += THREADS)
        for (int q = 0; q < 5; q++)
            s[q] += part[t * 5 + q];
    for (int q = 0; q < 5; q++)
        red[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int w = THREADS / 2; w > 0; w >>= 1)
    {
        if (threadIdx.x < w)
            for (int q = 0; q < 5; q++)
                red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 5)
        out[threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace

long grad_tiles(int n) { return ceil_div(n, TB) * (ceil_div(n, TB) + 1) / 2; }

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
                   double *out)
{
    const long tiles = grad_tiles(n);
    require(tiles >= 1 && tiles < (1L << 31), "gradient_sums: bad shape");
    const double fl = -2.0 * sek.vertical / sek.lengthscale * sig_l;
    hipLaunchKernelGGL(k_grad_partial, dim3(static_cast<unsigned>(tiles)), dim3(THREADS), 0, s, Kinv, lda, x, alpha, r, sek.scale(),
                       fl, sig_v, n, static_cast<int>(tiles), part);
    GP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_grad_reduce, dim3(1), dim3(THREADS), 0, s, part, static_cast<int>(tiles), out);
    GP_CHECK_LAUNCH();
}

template <int R, bool UPPER = false>
void launch_lower_w(hipStream_t s, dim3 grid, const double *x, double scale, double v, double noise, double *A, long lda, int n)
{
    hipLaunchKernelGGL((k_assemble_lower_w<R, UPPER>), grid, dim3(THREADS), 0, s, x, scale, v, noise, A, lda, n);
}
using LowerLaunch = void (*)(hipStream_t, dim3, const double *, double, double, double, double *, long, int);
constexpr LowerLaunch lower_w_table[16] = { launch_lower_w<1>,  launch_lower_w<2>,  launch_lower_w<3>,  launch_lower_w<4>,
                                            launch_lower_w<5>,  launch_lower_w<6>,  launch_lower_w<7>,  launch_lower_w<8>,
                                            launch_lower_w<9>,  launch_lower_w<10>, launch_lower_w<11>, launch_lower_w<12>,
                                            launch_lower_w<13>, launch_lower_w<14>, launch_lower_w<15>, launch_lower_w<16> };
constexpr LowerLaunch upper_w_table[16] = { launch_lower_w<1, true>,  launch_lower_w<2, true>,  launch_lower_w<3, true>,
                                            launch_lower_w<4, true>,  launch_lower_w<5, true>,  launch_lower_w<6, true>,
                                            launch_lower_w<7, true>,  launch_lower_w<8, true>,  launch_lower_w<9, true>,
                                            launch_lower_w<10, true>, launch_lower_w<11, true>, launch_lower_w<12, true>,
                                            launch_lower_w<13, true>, launch_lower_w<14, true>, launch_lower_w<15, true>,
                                            launch_lower_w<16, true> };

// Generic window length: the strict upper triangle, as the windowed kernel above.
__global__ __launch_bounds__(THREADS) void k_assemble_upper(const double *__restrict__ x, int r, double scale, double v,
                                                             double *__restrict__ A, long lda, int n)
{
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (tj < ti)
        return;
    const int c = threadIdx.x % TB;
    const long j = static_cast<long>(tj) * TB + c;
    for (int rr = threadIdx.x / TB; rr < TB; rr += THREADS / TB)
    {
        const long i = static_cast<long>(ti) * TB + rr;
        if (i >= j)
            continue;
        A[i * lda + j] = (i < n && j < n) ? se(scale, v, window_sqdist(x, i, x, j, r)) : 0.0;
    }
}

// Rows of K* (test point j by training point i) stored row-major, as the
// rows below the covariance that the one-call prediction factors with it:
// lane = training column (its window in registers), the test row
// wavefront-uniform (its window through scalar loads). Same operations in
// the same order as the transposed assembly ((test - train), no
// contraction), so the entries equal it bit for bit.
template <int R>
__global__ __launch_bounds__(THREADS) void k_assemble_cross_rows_w(const double *__restrict__ x,
                                                                    const double *__restrict__ z,
                                                                    double scale,
                                                                    double v,
                                                                    double *__restrict__ S,
                                                                    long lds,
                                                                    int n,
                                                                    int m)
{
#pragma clang fp contract(off)
    const int ti = blockIdx.x, tj = blockIdx.y;
    const int c = threadIdx.x % TB;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / TB);
    const int i = ti * TB + c;  // training index
    const int ic = i < n ? i : n - 1;
    double xi[R];
#pragma unroll
    for (int k = 0; k < R; k++)
        xi[k] = x[ic + k];
    for (int rr = w; rr < TB; rr += THREADS / TB)
    {
        const int j = tj * TB + rr;  // test index
        double val = 0.0;
        if (i < n && j < m)
        {
            double d = 0.0;
#pragma unroll
            for (int k = 0; k < R; k++)
            {
                const double t = z[j + k] - xi[k];
                d += t * t;
            }
            val = se(scale, v, d);
        }
        S[static_cast<long>(j) * lds + i] = val;
    }
}

__global__ __launch_bounds__(THREADS) void k_assemble_cross_rows(const double *__restrict__ x,
                                                                  const double *__restrict__ z,
                                                                  int r,
                                                                  double scale,
                                                                  double v,
                                                                  double *__restrict__ S,
                                                                  long lds,
                                                                  int n,
                                                                  int m)
{
    const int ti = blockIdx.x, tj = blockIdx.y;
    const long i = static_cast<long>(ti) * TB + threadIdx.x % TB;
    for (int rr = threadIdx.x / TB; rr < TB; rr += THREADS / TB)
    {
        const long j = static_cast<long>(tj) * TB + rr;
        double val = 0.0;
        if (i < n && j < m)
            val = se(scale, v, window_sqdist(z, j, x, i, r));
        S[j * lds + i] = val;
    }
}

// out[j] = sum over i < cols of S[j][i]^2, one workgroup per row, in a fixed order.
__global__ __launch_bounds__(THREADS) void k_row_sumsq(const double *__restrict__ S, long lds, int cols, double *__restrict__ out)
{
    __shared__ double red[THREADS];
    const double *row = S + static_cast<long>(blockIdx.x) * lds;
    const int tid = threadIdx.x;
    double s = 0.0;
    for (int i = 2 * tid; i < cols; i += 2 * THREADS)
    {
        const double a = row[i], b = row[i + 1];
        s = fma(a, a, s);
        s = fma(b, b, s);
    }
    red[tid] = s;
    __syncthreads();
    for (int h = THREADS / 2; h > 0; h >>= 1)
    {
        if (tid < h)
            red[tid] += red[tid + h];
        __syncthreads();
    }
    if (tid == 0)
        out[blockIdx.x] = red[0];
}

template <int R>
void launch_cross_rows_w(hipStream_t s, dim3 grid, const double *x, const double *z, double scale, double v, double *S, long lds, int n, int m)
{
    hipLaunchKernelGGL((k_assemble_cross_rows_w<R>), grid, dim3(THREADS), 0, s, x, z, scale, v, S, lds, n, m);
}
using CrossRowsLaunch = void (*)(hipStream_t, dim3, const double *, const double *, double, double, double *, long, int, int);
constexpr CrossRowsLaunch cross_rows_w_table[16] = {
    launch_cross_rows_w<1>,  launch_cross_rows_w<2>,  launch_cross_rows_w<3>,  launch_cross_rows_w<4>,
    launch_cross_rows_w<5>,  launch_cross_rows_w<6>,  launch_cross_rows_w<7>,  launch_cross_rows_w<8>,
    launch_cross_rows_w<9>,  launch_cross_rows_w<10>, launch_cross_rows_w<11>, launch_cross_rows_w<12>,
    launch_cross_rows_w<13>, launch_cross_rows_w<14>, launch_cross_rows_w<15>, launch_cross_rows_w<16>,
};

template <class T, int R>
void launch_cross_w(hipStream_t s, dim3 grid, const double *x, const double *z, double scale, double v, T *B, long ldb, int n, int m)
{
    hipLaunchKernelGGL((k_assemble_cross_t_w<T, R>), grid, dim3(THREADS), 0, s, x, z, scale, v, B, ldb, n, m);
}
template <class T>
using CrossLaunch = void (*)(hipStream_t, dim3, const double *, const double *, double, double, T *, long, int, int);
template <class T>
constexpr CrossLaunch<T> cross_w_table[16] = { launch_cross_w<T, 1>,  launch_cross_w<T, 2>,  launch_cross_w<T, 3>,  launch_cross_w<T, 4>,
                                               launch_cross_w<T, 5>,  launch_cross_w<T, 6>,  launch_cross_w<T, 7>,  launch_cross_w<T, 8>,
                                               launch_cross_w<T, 9>,  launch_cross_w<T, 10>, launch_cross_w<T, 11>, launch_cross_w<T, 12>,
                                               launch_cross_w<T, 13>, launch_cross_w<T, 14>, launch_cross_w<T, 15>, launch_cross_w<T, 16> };

void assemble_lower(hipStream_t s, const double *x, int r, SEK sek, double *A, long lda, int n, int n_pad)
{
    require(n_pad % TB == 0 && n <= n_pad && lda >= n_pad && r >= 1, "assemble_lower: bad shape");
    dim3 grid(n_pad / TB, n_pad / TB);
    if (r <= 16 && n >= 1 && n_pad < (1 << 30) - 16)
        lower_w_table[r - 1](s, grid, x, sek.scale(), sek.vertical, sek.noise, A, lda, n);
    else
        hipLaunchKernelGGL(k_assemble_lower, grid, dim3(THREADS), 0, s, x, r, sek.scale(), sek.vertical, sek.noise, A, lda, n);
    GP_CHECK_LAUNCH();
}

void assemble_upper(hipStream_t s, const double *x, int r, SEK sek, double *A, long lda, int n, int n_pad)
{
    require(n_pad % TB == 0 && n <= n_pad && lda >= n_pad && r >= 1, "assemble_upper: bad shape");
    dim3 grid(n_pad / TB, n_pad / TB);
    if (r <= 16 && n >= 1 && n_pad < (1 << 30) - 16)
        upper_w_table[r - 1](s, grid, x, sek.scale(), sek.vertical, sek.noise, A, lda, n);
    else
        hipLaunchKernelGGL(k_assemble_upper, grid, dim3(THREADS), 0, s, x, r, sek.scale(), sek.vertical, A, lda, n);
    GP_CHECK_LAUNCH();
}

template <class T>
static void assemble_cross_t_impl(hipStream_t s, const double *x, const double *z, int r, SEK sek, T *B, long ldb, int n, int n_pad, int m, int m_pad)
{
    require(n_pad % TB == 0 && m_pad % TB == 0 && n <= n_pad && m <= m_pad && ldb >= m_pad, "assemble_cross_t: bad shape");
    dim3 grid(m_pad / TB, n_pad / TB);
    if (r <= 16 && m >= 1 && n_pad < (1 << 30) - 16 && m_pad < (1 << 30) - 16)
        cross_w_table<T>[r - 1](s, grid, x, z, sek.scale(), sek.vertical, B, ldb, n, m);
    else
        hipLaunchKernelGGL(k_assemble_cross_t<T>, grid, dim3(THREADS), 0, s, x, z, r, sek.scale(), sek.vertical, B, ldb, n, m);
    GP_CHECK_LAUNCH();
}

void assemble_cross_t(hipStream_t s, const double *x, const double *z, int r, SEK sek, double *B, long ldb, int n, int n_pad, int m, int m_pad)
{
    assemble_cross_t_impl<double>(s, x, z, r, sek, B, ldb, n, n_pad, m, m_pad);
}

void assemble_cross_t_f32(hipStream_t s, const double *x, const double *z, int r, SEK sek, float *B, long ldb, int n, int n_pad, int m, int m_pad)
{
    assemble_cross_t_impl<float>(s, x, z, r, sek, B, ldb, n, n_pad, m, m_pad);
}

void assemble_cross_rows(hipStream_t s, const double *x, const double *z, int r, SEK sek, double *S, long lds, int n, int n_pad, int m,
                         int m_pad)
{
    require(n_pad % TB == 0 && m_pad % TB == 0 && n <= n_pad && m <= m_pad && lds >= n_pad && r >= 1, "assemble_cross_rows: bad shape");
    dim3 grid(n_pad / TB, m_pad / TB);
    if (r <= 16 && n >= 1 && m >= 1 && n_pad < (1 << 30) - 16 && m_pad < (1 << 30) - 16)
        cross_rows_w_table[r - 1](s, grid, x, z, sek.scale(), sek.vertical, S, lds, n, m);
    else
        hipLaunchKernelGGL(k_assemble_cross_rows, grid, dim3(THREADS), 0, s, x, z, r, sek.scale(), sek.vertical, S, lds, n, m);
    GP_CHECK_LAUNCH();
}

void row_sumsq(hipStream_t s, const double *S, long lds, int rows, int cols, double *out)
{
    require(rows >= 0 && cols >= 0 && cols % 2 == 0 && lds % 2 == 0, "row_sumsq: bad shape");
    if (rows == 0)
        return;
    hipLaunchKernelGGL(k_row_sumsq, dim3(rows), dim3(THREADS), 0, s, S, lds, cols, out);
    GP_CHECK_LAUNCH();
}

int mean_splits(int n, int m)
{
    const long col_blocks = ceil_div(m, TB);
    long splits = ceil_div(2048, col_blocks);
    splits = std::min<long>(splits, ceil_div(n, 512));
    return static_cast<int>(std::max<long>(splits, 1));
}

void predict_mean(hipStream_t s, const double *x, const double *z, const double *alpha, int r, SEK sek, int n, int m, double *partial, int splits, double *mean)
{
    require(n >= 1 && m >= 1 && splits >= 1 && r >= 1, "predict_mean: bad shape");
    const int chunk = static_cast<int>(ceil_div(n, splits));
    require(n + r < (1 << 30) && m + r < (1 << 30), "predict_mean: sizes beyond 32-bit indexing");
    dim3 grid(ceil_div(m, TB), splits);
    if (r <= 16)
        mean_w_table[r - 1](s, grid, x, z, alpha, sek.scale(), sek.vertical, n, m, chunk, partial);
    else
        hipLaunchKernelGGL(k_mean_partial, grid, dim3(THREADS), 0, s, x, z, alpha, r, sek.scale(), sek.vertical, n, m, chunk, partial);
    GP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_mean_reduce, dim3(ceil_div(m, THREADS)), dim3(THREADS), 0, s, partial, splits, m, mean);
    GP_CHECK_LAUNCH();
}

void gather_tiles(hipStream_t s, const double *L, long lda, const double *x, int r, SEK sek, int T, int b, double *out)
{
    require(T >= 1 && b >= 1, "gather_tiles: bad shape");
    const int rows_per_block = 16;
    dim3 grid(static_cast<unsigned>(static_cast<long>(T) * (T + 1) / 2), ceil_div(b, rows_per_block));
    hipLaunchKernelGGL(k_gather_tiles, grid, dim3(THREADS), 0, s, L, lda, x, r, sek.scale(), sek.vertical, b, rows_per_block, out);
    GP_CHECK_LAUNCH();
}

}  // namespace gpamd::k
