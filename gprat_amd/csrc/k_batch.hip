// Batched fit of many independent GPs of uneven size in one persistent
// launch. Each workgroup repeatedly takes the next GP from a device counter
// (GPs are ordered largest first) and, without leaving the workgroup,
// assembles its covariance, factors it with 64-wide blocks (diagonal block
// factor + inverse, then 64x64x64 FP64 MFMA block products for the panel
// solve and the trailing update), solves alpha = K^-1 y and writes the loss.
// Every workgroup exits as soon as the counter passes the batch size.
#include "block64.hpp"
#include "kernels.hpp"

#include <climits>

namespace gpamd::k
{

namespace
{

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int NB = 64;
constexpr int THREADS = 256;
constexpr int XS = NB + 2;  // LDS row stride of staged 64x64 blocks (bank-spread MFMA reads)

struct Stage
{
    double a[NB * XS];
    double b[NB * XS];
};
struct Factor
{
    double D[NB][NB + 1];
    double W[NB][NB + 1];
    double S[3][16][17];
};
union Shared
{
    Stage st;
    Factor fa;
};

// Copy a 64x64 row-major block (leading dimension ld) into LDS stride XS.
__device__ __forceinline__ void stage_block(double *dst, const double *__restrict__ src, long ld)
{
    for (int e = threadIdx.x; e < NB * NB / 2; e += THREADS)
    {
        const int r = e >> 5, c = (e & 31) * 2;
        const double2 v = *reinterpret_cast<const double2 *>(src + static_cast<long>(r) * ld + c);
        dst[r * XS + c] = v.x;
        dst[r * XS + c + 1] = v.y;
    }
}

// acc (this wavefront's 32x32 quarter of a 64x64 block) = sum_k A[r][k] B[c][k]
// with A and B staged in LDS. Wavefront w owns rows 32*(w>>1).., columns 32*(w&1)..
__device__ __forceinline__ void mfma_nt64(const double *A, const double *B, v4d (&acc)[2][2])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const int r0 = (w >> 1) * 32, c0 = (w & 1) * 32;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
            acc[i][j] = v4d{ 0.0, 0.0, 0.0, 0.0 };
#pragma unroll 4
    for (int k0 = 0; k0 < NB; k0 += 4)
    {
        double a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++)
            a[i] = A[(r0 + i * 16 + lm) * XS + k0 + lk];
#pragma unroll
        for (int j = 0; j < 2; j++)
            b[j] = B[(c0 + j * 16 + lm) * XS + k0 + lk];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
}

// C (global, 64x64 block, ld) = acc  or  C -= acc, optionally only on or
// below the block's diagonal. f64 16x16x4 accumulator layout: lane holds
// column lane%16, rows lane/16 + 4e.
__device__ __forceinline__ void store_acc(double *C, long ld, const v4d (&acc)[2][2], bool set, bool lower)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const int r0 = (w >> 1) * 32, c0 = (w & 1) * 32;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 4; e++)
            {
                const int r = r0 + i * 16 + lk + 4 * e, c = c0 + j * 16 + lm;
                if (lower && c > r)
                    continue;
                double *dst = C + static_cast<long>(r) * ld + c;
                if (set)
                    *dst = acc[i][j][e];
                else
                    *dst -= acc[i][j][e];
            }
}

__global__ __launch_bounds__(THREADS) void k_batch_fit(const BatchGP *__restrict__ gps,
                                                        const int *__restrict__ order,
                                                        int count,
                                                        int *__restrict__ counter,
                                                        double *__restrict__ arena,
                                                        double *__restrict__ losses,
                                                        int *__restrict__ info)
{
    __shared__ Shared sh;
    __shared__ int next;
    __shared__ double red[2][THREADS];
    const int tid = threadIdx.x;
    for (;;)
    {
        if (tid == 0)
            next = atomicAdd(counter, 1);
        __syncthreads();
        const int slot = next;
        __syncthreads();
        if (slot >= count)
            break;
        const int gi = order[slot];
        const BatchGP g = gps[gi];
        const double *x = arena + g.x_off;
        const double *y = arena + g.y_off;
        double *L = arena + g.l_off;
        double *Wg = arena + g.w_off;
        double *alpha = arena + g.a_off;
        const long np = g.n_pad;
        const int nblk = static_cast<int>(np / NB);
        const double scale = -0.5 / (g.lengthscale * g.lengthscale);

        // 1. lower band of K (whole 64-blocks on and below the diagonal),
        // identity padding; wavefronts take rows, lanes take columns
        const int lane = tid & 63, wv = tid >> 6;
        for (long i = wv; i < np; i += THREADS / 64)
            for (long j = lane; j < ((i >> 6) + 1) * 64; j += 64)
            {
            double val;
            if (i < g.n && j < g.n)
            {
                val = g.vertical * exp(scale * window_sqdist(x, i, x, j, g.n_reg));
                if (i == j)
                    val += g.noise;
            }
            else
            {
                val = (i == j) ? 1.0 : 0.0;
            }
            L[i * np + j] = val;
            }
        __syncthreads();

        // 2. blocked right-looking Cholesky
        for (int kb = 0; kb < nblk; kb++)
        {
            double *Lkk = L + static_cast<long>(kb) * NB * np + kb * NB;
            for (int e = tid; e < NB * NB; e += THREADS)
            {
                const int r = e / NB, c = e % NB;
                sh.fa.D[r][c] = (c <= r) ? Lkk[static_cast<long>(r) * np + c] : 0.0;
            }
            __syncthreads();
            factor_block64(sh.fa.D, sh.fa.W, sh.fa.S, info + gi, kb * NB);
            double *Wkk = Wg + static_cast<long>(kb) * NB * NB;
            for (int e = tid; e < NB * NB; e += THREADS)
            {
                const int r = e / NB, c = e % NB;
                if (c <= r)
                    Lkk[static_cast<long>(r) * np + c] = sh.fa.D[r][c];
                Wkk[e] = (c <= r) ? sh.fa.W[r][c] : 0.0;
            }
            __syncthreads();
            // panel solve: X_ib = A_ib W_kk^T (in place)
            for (int ib = kb + 1; ib < nblk; ib++)
            {
                double *Aik = L + static_cast<long>(ib) * NB * np + kb * NB;
                stage_block(sh.st.a, Aik, np);
                stage_block(sh.st.b, Wkk, NB);
                __syncthreads();
                v4d acc[2][2];
                mfma_nt64(sh.st.a, sh.st.b, acc);
                __syncthreads();
                store_acc(Aik, np, acc, true, false);
                __syncthreads();
            }
            // trailing update: C_ij -= X_i X_j^T for kb < j <= i
            for (int ib = kb + 1; ib < nblk; ib++)
            {
                stage_block(sh.st.a, L + static_cast<long>(ib) * NB * np + kb * NB, np);
                for (int jb = kb + 1; jb <= ib; jb++)
                {
                    stage_block(sh.st.b, L + static_cast<long>(jb) * NB * np + kb * NB, np);
                    __syncthreads();
                    v4d acc[2][2];
                    mfma_nt64(sh.st.a, sh.st.b, acc);
                    store_acc(L + static_cast<long>(ib) * NB * np + jb * NB, np, acc, false, jb == ib);
                    __syncthreads();
                }
            }
        }

        // 3. alpha: forward then backward block substitution with W
        double *xs = red[0];
        for (long i = tid; i < np; i += THREADS)
            alpha[i] = (i < g.n) ? y[i] : 0.0;
        __syncthreads();
        for (int J = 0; J < nblk; J++)
        {
            const double *W = Wg + static_cast<long>(J) * NB * NB;
            const long base = static_cast<long>(J) * NB;
            if (tid < NB)
            {
                double s = 0.0;
                for (int k = 0; k < NB; k++)
                    s += W[tid * NB + k] * alpha[base + k];
                xs[tid] = s;
            }
            __syncthreads();
            if (tid < NB)
                alpha[base + tid] = xs[tid];
            for (long i = base + NB + tid; i < np; i += THREADS)
            {
                const double *row = L + i * np + base;
                double s = 0.0;
                for (int k = 0; k < NB; k++)
                    s += row[k] * xs[k];
                alpha[i] -= s;
            }
            __syncthreads();
        }
        for (int J = nblk - 1; J >= 0; J--)
        {
            const double *W = Wg + static_cast<long>(J) * NB * NB;
            const long base = static_cast<long>(J) * NB;
            if (tid < NB)
            {
                double s = 0.0;
                for (int k = 0; k < NB; k++)
                    s += W[k * NB + tid] * alpha[base + k];
                xs[tid] = s;
            }
            __syncthreads();
            if (tid < NB)
                alpha[base + tid] = xs[tid];
            for (long jcol = tid; jcol < base; jcol += THREADS)
            {
                double s = 0.0;
                for (int k = 0; k < NB; k++)
                    s += L[(base + k) * np + jcol] * xs[k];
                alpha[jcol] -= s;
            }
            __syncthreads();
        }

        // 4. loss = 0.5 (y^T alpha + sum log L_ii^2 + n log 2 pi) / n
        double s1 = 0.0, s2 = 0.0;
        for (long i = tid; i < g.n; i += THREADS)
        {
            s1 += y[i] * alpha[i];
            const double d = L[i * np + i];
            s2 += log(d * d);
        }
        red[0][tid] = s1;
        red[1][tid] = s2;
        __syncthreads();
        for (int w = THREADS / 2; w > 0; w >>= 1)
        {
            if (tid < w)
            {
                red[0][tid] += red[0][tid + w];
                red[1][tid] += red[1][tid + w];
            }
            __syncthreads();
        }
        if (tid == 0)
        {
            const double Nn = static_cast<double>(g.n);
            double l = red[0][0] + red[1][0];
            l += Nn * log(2.0 * 3.14159265358979323846);
            losses[gi] = 0.5 * l / Nn;
        }
        __syncthreads();
    }
}

__global__ void k_set_ints(int *p, int count, int value)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        p[i] = value;
}

}  // namespace

void batch_fit(hipStream_t s, const BatchGP *gps, const int *order, int count, int *counter, double *arena, double *losses,
               int *info, int workgroups)
{
    require(count >= 1 && workgroups >= 1, "batch_fit: bad shape");
    hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(1), 0, s, counter, 1, 0);
    GP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_set_ints, dim3(ceil_div(count, 256)), dim3(256), 0, s, info, count, INT_MAX);
    GP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_batch_fit, dim3(workgroups), dim3(THREADS), 0, s, gps, order, count, counter, arena, losses, info);
    GP_CHECK_LAUNCH();
}

}  // namespace gpamd::k
