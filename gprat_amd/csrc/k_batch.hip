// Batched fit of many independent GPs of uneven size in one persistent
// launch. Each workgroup repeatedly takes the next GP from a device counter
// (GPs are ordered largest first) and, without leaving the workgroup,
// assembles its covariance, factors it, solves alpha = K^-1 y and writes
// the loss. Every workgroup exits as soon as the counter passes the batch
// size.
//
// The factorisation is left-looking over 64-wide block columns: block
// (i, j) of column j is generated from the series (no assembly pass: each
// block of K is needed once), accumulates sum_{k<j} L_ik L_jk^T on FP64 MFMA in
// registers, with the next pair of blocks already on its way from memory
// (register prefetch) while the current pair is multiplied; the sum is then
// subtracted from K_ij once and the block is either factored (i == j:
// Cholesky factor and inverse of the 64x64 block in LDS) or multiplied by
// the diagonal block's inverse, which stays in LDS for the whole column.
#include "block64.hpp"
#include "block_mfma.hpp"
#include "gemm_tile.hpp"
#include "kernels.hpp"

#include <climits>

namespace gpamd::k
{

namespace
{

using blk::NB;
using blk::THREADS;
using blk::XS;
using blk::v4d;

typedef double d2 __attribute__((ext_vector_type(2)));

struct Stage
{
    double a[NB * XS];
    double b[NB * XS];
};
struct Factor
{
    double D[NB][NB + 1];
    double W[NB][NB + 1];
    double R[NB];
    double T[NB];  // the diagonal step's L_jb,0:jb z_0:jb (alpha's forward solve)
};
// FP32 copies of two 64x64 blocks for the mixed-precision products; row
// stride XF floats (the 16 rows x 4 k-slots of an MFMA operand read fall in
// distinct banks)
constexpr int XF = NB + 4;
struct Stage32
{
    float a[NB * XF];
    float b[NB * XF];
};
union Shared
{
    Stage st;
    Stage32 st32;
    Factor fa;
};

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));

// One thread's share of a 64x64 row-major block: 8 pairs of doubles, row
// e >> 5 and columns (e & 31) * 2 for e = tid + 256 s.
struct BlockRegs
{
    d2 v[8];

    __device__ __forceinline__ void load(const double *__restrict__ src, long ld)
    {
#pragma unroll
        for (int s = 0; s < 8; s++)
        {
            const int e = threadIdx.x + THREADS * s;
            v[s] = *reinterpret_cast<const d2 *>(src + static_cast<long>(e >> 5) * ld + (e & 31) * 2);
        }
    }

    __device__ __forceinline__ void store(double *dst) const
    {
#pragma unroll
        for (int s = 0; s < 8; s++)
        {
            const int e = threadIdx.x + THREADS * s;
            *reinterpret_cast<d2 *>(dst + (e >> 5) * XS + (e & 31) * 2) = v[s];
        }
    }

    // rounded to FP32, row stride XF
    __device__ __forceinline__ void store32(float *dst) const
    {
#pragma unroll
        for (int s = 0; s < 8; s++)
        {
            const int e = threadIdx.x + THREADS * s;
            *reinterpret_cast<f2 *>(dst + (e >> 5) * XF + (e & 31) * 2) = f2{ static_cast<float>(v[s].x), static_cast<float>(v[s].y) };
        }
    }
};

// Position (row, column within the 64x64 block) of this lane's accumulator
// entry q = (i * 2 + j) * 4 + e. f64 16x16x4: column lane % 16, rows
// lane / 16 + 4 e; f32 16x16x4: column lane % 16, rows 4 (lane / 16) + e.
template <bool F32>
__device__ __forceinline__ void acc_position(int q, int &r, int &c)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const int i = q >> 3, j = (q >> 2) & 1, e = q & 3;
    r = (w >> 1) * 32 + i * 16 + (F32 ? 4 * lk + e : lk + 4 * e);
    c = (w & 1) * 32 + j * 16 + lm;
}

// acc += sum_k A[r][k] B[c][k] over a 64-deep FP32 block pair in LDS, FP32
// accumulation; wavefront w owns rows 32 (w >> 1).., columns 32 (w & 1)..
__device__ __forceinline__ void mfma_nt_acc32(const float *A, const float *B, v4f (&acc)[2][2])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const int r0 = (w >> 1) * 32, c0 = (w & 1) * 32;
#pragma unroll 4
    for (int k0 = 0; k0 < NB; k0 += 4)
    {
        float a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++)
            a[i] = A[(r0 + i * 16 + lm) * XF + k0 + lk];
#pragma unroll
        for (int j = 0; j < 2; j++)
            b[j] = B[(c0 + j * 16 + lm) * XF + k0 + lk];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
}

// Carried state of a partial accumulate_row_products (a diagonal block's
// sum over its first k-blocks, computed one launch early): each thread's 16
// accumulator entries, then its forward-solve partial, at
// carry[q * THREADS + tid] (q < 16) and carry[16 * THREADS + tid].
static_assert(BATCH_CARRY_DOUBLES == 17 * THREADS, "carry layout");

__device__ __forceinline__ void save_carry(double *carry, const v4d (&acc)[2][2], double tp)
{
#pragma unroll
    for (int q = 0; q < 16; q++)
        carry[q * THREADS + threadIdx.x] = acc[q >> 3][(q >> 2) & 1][q & 3];
    carry[16 * THREADS + threadIdx.x] = tp;
}

__device__ __forceinline__ void load_carry(const double *carry, v4d (&acc)[2][2], double &tp)
{
#pragma unroll
    for (int q = 0; q < 16; q++)
        acc[q >> 3][(q >> 2) & 1][q & 3] = carry[q * THREADS + threadIdx.x];
    tp = carry[16 * THREADS + threadIdx.x];
}

// acc = sum_{k < kn} L[i, k] L[j, k]^T over 64-blocks of block rows i and j
// (block (r, k) at L + r*64*np + k*64), streamed through the LDS stage with
// the next pair prefetched into registers. With zv, also tout = L[i, 0:kn]
// zv[0:kn] (64 values, fixed order: each block's 64 columns in four
// 16-column parts per row, the parts summed in a fixed order at the end),
// from the blocks of row i while they are in LDS. With carry, the sums over
// k < k0 come from there (save_carry of the same sums) and the loop starts
// at block k0: every accumulator then sees the same operations in the same
// order as without.
__device__ __forceinline__ void accumulate_row_products(const double *L, long np, int ib, int jb, int kn, Stage &st, v4d (&acc)[2][2],
                                                        const double *zv = nullptr, double *tout = nullptr, int k0 = 0,
                                                        const double *carry = nullptr)
{
    const int trow = threadIdx.x & 63, tq = threadIdx.x >> 6;
    double tp = 0.0;
    if (carry)
        load_carry(carry, acc, tp);
    else
    {
        blk::zero(acc);
        k0 = 0;
    }
    if (kn == 0)
    {
        if (tout && threadIdx.x < NB)
            tout[threadIdx.x] = 0.0;
        return;
    }
    const double *ai = L + static_cast<long>(ib) * NB * np;
    const double *bj = L + static_cast<long>(jb) * NB * np;
    BlockRegs ra, rb;
    ra.load(ai + k0 * NB, np);
    rb.load(bj + k0 * NB, np);
    for (int kb = k0; kb < kn; kb++)
    {
        __syncthreads();  // every wavefront is done with the previous pair
        ra.store(st.a);
        rb.store(st.b);
        __syncthreads();
        if (kb + 1 < kn)
        {
            ra.load(ai + (kb + 1) * NB, np);
            rb.load(bj + (kb + 1) * NB, np);
        }
        // keep the prefetch ahead of the products
        __builtin_amdgcn_sched_barrier(0);
        blk::mfma_nt_acc(st.a, st.b, acc);
        if (zv)
        {
            const double *zb = zv + kb * NB + tq * 16;
#pragma unroll
            for (int c = 0; c < 16; c++)
                tp = fma(st.a[trow * XS + tq * 16 + c], zb[c], tp);
        }
    }
    __syncthreads();
    if (tout)
    {
        st.b[tq * NB + trow] = tp;
        __syncthreads();
        if (threadIdx.x < NB)
            tout[threadIdx.x] = (st.b[threadIdx.x] + st.b[NB + threadIdx.x]) + (st.b[2 * NB + threadIdx.x] + st.b[3 * NB + threadIdx.x]);
        __syncthreads();
    }
}

// The same sum with FP32 operands and FP32 accumulation.
__device__ __forceinline__ void accumulate_row_products32(const double *L, long np, int ib, int jb, int kn, Stage32 &st, v4f (&acc)[2][2])
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
            acc[i][j] = v4f{ 0.0f, 0.0f, 0.0f, 0.0f };
    if (kn == 0)
        return;
    const double *ai = L + static_cast<long>(ib) * NB * np;
    const double *bj = L + static_cast<long>(jb) * NB * np;
    BlockRegs ra, rb;
    ra.load(ai, np);
    rb.load(bj, np);
    for (int kb = 0; kb < kn; kb++)
    {
        __syncthreads();  // every wavefront is done with the previous pair
        ra.store32(st.a);
        rb.store32(st.b);
        __syncthreads();
        if (kb + 1 < kn)
        {
            ra.load(ai + (kb + 1) * NB, np);
            rb.load(bj + (kb + 1) * NB, np);
        }
        __builtin_amdgcn_sched_barrier(0);
        mfma_nt_acc32(st.a, st.b, acc);
    }
    __syncthreads();
}

// vec <- L^-1 vec for one GP from its factor: forward block substitution
// with the saved 64x64 inverses (W), by a workgroup of NT >= 256 threads.
// xs and part are two 256-double LDS buffers. The first 256 threads form
// the products with W_JJ (four fixed-order partial sums per row) and every
// row's update sums its 64 products in k order, so the result does not
// depend on NT.
template <int NT>
__device__ __forceinline__ void forward_solve(const double *L, const double *Wg, long np, int nblk, double *vec, double *xs, double *part)
{
    const int tid = threadIdx.x;
    const int row = tid & 63, quarter = tid >> 6;
    for (int J = 0; J < nblk; J++)
    {
        const double *W = Wg + static_cast<long>(J) * NB * NB;
        const long base = static_cast<long>(J) * NB;
        if (tid < THREADS)
        {
            // x_J = W_JJ vec_J: four partial sums per row, fixed order
            double w[16], a[16];
#pragma unroll
            for (int k = 0; k < 16; k++)
            {
                w[k] = W[row * NB + quarter * 16 + k];
                a[k] = vec[base + quarter * 16 + k];
            }
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 16; k++)
                s += w[k] * a[k];
            part[quarter * 64 + row] = s;
        }
        __syncthreads();
        if (tid < NB)
        {
            xs[tid] = (part[tid] + part[64 + tid]) + (part[128 + tid] + part[192 + tid]);
            vec[base + tid] = xs[tid];
        }
        __syncthreads();
        for (long i = base + NB + tid; i < np; i += NT)
        {
            const double *lr = L + i * np + base;
            // the row's 64 values all in flight before its products, which
            // run in k order
            d2 v[NB / 2];
#pragma unroll
            for (int q = 0; q < NB / 2; q++)
                v[q] = *reinterpret_cast<const d2 *>(lr + 2 * q);
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < NB / 2; q++)
                s += v[q].x * xs[2 * q] + v[q].y * xs[2 * q + 1];
            vec[i] -= s;
        }
        __syncthreads();
    }
}

// vec <- L^-T vec: the backward block substitution (NT as in forward_solve).
template <int NT>
__device__ __forceinline__ void backward_solve(const double *L, const double *Wg, long np, int nblk, double *vec, double *xs, double *part)
{
    const int tid = threadIdx.x;
    const int row = tid & 63, quarter = tid >> 6;
    for (int J = nblk - 1; J >= 0; J--)
    {
        const double *W = Wg + static_cast<long>(J) * NB * NB;
        const long base = static_cast<long>(J) * NB;
        if (tid < THREADS)
        {
            // x_J = W_JJ^T vec_J
            double w[16], a[16];
#pragma unroll
            for (int k = 0; k < 16; k++)
            {
                w[k] = W[(quarter * 16 + k) * NB + row];
                a[k] = vec[base + quarter * 16 + k];
            }
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 16; k++)
                s += w[k] * a[k];
            part[quarter * 64 + row] = s;
        }
        __syncthreads();
        if (tid < NB)
        {
            xs[tid] = (part[tid] + part[64 + tid]) + (part[128 + tid] + part[192 + tid]);
            vec[base + tid] = xs[tid];
        }
        __syncthreads();
        // a column's values in flight before their products (a step is a
        // chain of memory latencies otherwise): all 64 with 256 threads,
        // two halves of 32 with more (the register budget halves); the
        // products in k order either way
        constexpr int HALF = NT > THREADS ? NB / 2 : NB;
        for (long jcol = tid; jcol < base; jcol += NT)
        {
            double s = 0.0;
#pragma unroll
            for (int h = 0; h < NB; h += HALF)
            {
                double v[HALF];
#pragma unroll
                for (int q = 0; q < HALF; q++)
                    v[q] = L[(base + h + q) * np + jcol];
#pragma unroll
                for (int q = 0; q < HALF; q++)
                    s += v[q] * xs[h + q];
            }
            vec[jcol] -= s;
        }
        __syncthreads();
    }
}

// vec <- K^-1 vec: forward, then backward.
template <int NT>
__device__ __forceinline__ void block_solves(const double *L, const double *Wg, long np, int nblk, double *vec, double *xs, double *part)
{
    forward_solve<NT>(L, Wg, np, nblk, vec, xs, part);
    backward_solve<NT>(L, Wg, np, nblk, vec, xs, part);
}

// max over the workgroup of NT threads of v (all threads get it); red holds
// NT doubles
template <int NT>
__device__ __forceinline__ double block_max(double v, double *red)
{
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1)
    {
        if (tid < w)
            red[tid] = fmax(red[tid], red[tid + w]);
        __syncthreads();
    }
    const double m = red[0];
    __syncthreads();
    return m;
}

// r_i = y_i - (K alpha)_i for row i < n of one GP, K regenerated from the
// series in FP64. The row's window is in registers; the column index is
// uniform across the wavefront, so the column's window and alpha_j arrive
// through scalar loads. R: the window length (compile time).
template <int R>
__device__ __forceinline__ double residual_row_w(const BatchGP &g, double scale, const double *alpha, long i)
{
    const double *x = g.x;
    double xi[R];
#pragma unroll
    for (int k = 0; k < R; k++)
        xi[k] = x[i + k];
    double s = 0.0;
    for (long j = 0; j < g.n; j++)
    {
        double d = 0.0;
#pragma unroll
        for (int k = 0; k < R; k++)
        {
            const double t = xi[k] - x[j + k];
            d = fma(t, t, d);
        }
        s = fma(g.vertical * exp(scale * d), alpha[j], s);
    }
    return g.y[i] - (s + g.noise * alpha[i]);
}

__device__ __forceinline__ double residual_row(const BatchGP &g, double scale, const double *alpha, long i)
{
    if (g.n_reg == 4)
        return residual_row_w<4>(g, scale, alpha, i);
    if (g.n_reg == 8)
        return residual_row_w<8>(g, scale, alpha, i);
    if (g.n_reg == 16)
        return residual_row_w<16>(g, scale, alpha, i);
    double s = 0.0;
    for (long j = 0; j < g.n; j++)
        s += g.vertical * exp(scale * window_sqdist(g.x, i, g.x, j, g.n_reg)) * alpha[j];
    s += g.noise * alpha[i];
    return g.y[i] - s;
}

// r = y - K alpha for one GP by the whole workgroup (r[i] = 0 for padding
// rows).
__device__ __forceinline__ void residual_rows(const BatchGP &g, double scale, const double *alpha, double *r)
{
    for (long i = threadIdx.x; i < g.n_pad; i += THREADS)
        r[i] = i < g.n ? residual_row(g, scale, alpha, i) : 0.0;
}

// Rows [r0, r1) (whole 64-row blocks) of one GP's lower band: K on and
// below the diagonal blocks, identity padding; wavefronts take rows, lanes
// take columns. For window lengths R = 4, 8, 16 a lane keeps its column's
// window in registers and the row's window (the row is wavefront-uniform)
// arrives through scalar loads; the same operations in the same order as
// window_sqdist (no contraction), so every entry matches the generic loop.
template <int R>
__device__ __forceinline__ void assemble_rows_w(const BatchGP &g, long r0, long r1)
{
#pragma clang fp contract(off)
    const double scale = -0.5 / (g.lengthscale * g.lengthscale);
    const long np = g.n_pad;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const double *x = g.x;
    for (long j0 = 0; j0 < r1; j0 += 64)
    {
        const long j = j0 + lane;
        const long jc = j < g.n ? j : g.n - 1;  // lanes past the end read a valid window and store the padding value
        double xj[R];
#pragma unroll
        for (int k = 0; k < R; k++)
            xj[k] = x[jc + k];
        // rows of this 64-block column that lie in [r0, r1) and on or below its diagonal block
        const long rs = r0 > j0 ? r0 : j0;
        for (long i = rs + wv; i < r1; i += THREADS / 64)
        {
            double val;
            if (i < g.n && j < g.n)
            {
                double d = 0.0;
#pragma unroll
                for (int k = 0; k < R; k++)
                {
                    const double t = x[i + k] - xj[k];
                    d += t * t;
                }
                val = g.vertical * exp(scale * d);
                if (i == j)
                    val += g.noise;
            }
            else
                val = (i == j) ? 1.0 : 0.0;
            g.L[i * np + j] = val;
        }
    }
}

__device__ __forceinline__ void assemble_rows(const BatchGP &g, long r0, long r1)
{
    if (g.n_reg == 4)
        return assemble_rows_w<4>(g, r0, r1);
    if (g.n_reg == 8)
        return assemble_rows_w<8>(g, r0, r1);
    if (g.n_reg == 16)
        return assemble_rows_w<16>(g, r0, r1);
    const double scale = -0.5 / (g.lengthscale * g.lengthscale);
    const long np = g.n_pad;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (long i = r0 + wv; i < r1; i += THREADS / 64)
        for (long j = lane; j < ((i >> 6) + 1) * 64; j += 64)
        {
            double val;
            if (i < g.n && j < g.n)
            {
                val = g.vertical * exp(scale * window_sqdist(g.x, i, g.x, j, g.n_reg));
                if (i == j)
                    val += g.noise;
            }
            else
            {
                val = (i == j) ? 1.0 : 0.0;
            }
            g.L[i * np + j] = val;
        }
}

// Block (ib, jb) of one GP's covariance, generated from its series into D
// (64 rows, stride XS): the operations and their order of assemble_rows_w
// (lane = column, its window in registers; the row wavefront-uniform, its
// window through scalar loads), so every entry equals the assembled one.
template <int R>
__device__ __forceinline__ void gen_block_w(const BatchGP &g, int ib, int jb, double *D)
{
#pragma clang fp contract(off)
    const double scale = -0.5 / (g.lengthscale * g.lengthscale);
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const double *x = g.x;
    const long j = static_cast<long>(jb) * NB + lane;
    const long jc = j < g.n ? j : g.n - 1;  // lanes past the end read a valid window and store the padding value
    double xj[R];
#pragma unroll
    for (int k = 0; k < R; k++)
        xj[k] = x[jc + k];
    for (int rr = wv; rr < NB; rr += THREADS / 64)
    {
        const long i = static_cast<long>(ib) * NB + rr;
        double val;
        if (i < g.n && j < g.n)
        {
            double d = 0.0;
#pragma unroll
            for (int k = 0; k < R; k++)
            {
                const double t = x[i + k] - xj[k];
                d += t * t;
            }
            val = g.vertical * exp(scale * d);
            if (i == j)
                val += g.noise;
        }
        else
            val = (i == j) ? 1.0 : 0.0;
        D[rr * XS + lane] = val;
    }
}

__device__ __forceinline__ void gen_block(const BatchGP &g, int ib, int jb, double *D)
{
    if (g.n_reg == 4)
        return gen_block_w<4>(g, ib, jb, D);
    if (g.n_reg == 8)
        return gen_block_w<8>(g, ib, jb, D);
    if (g.n_reg == 16)
        return gen_block_w<16>(g, ib, jb, D);
    const double scale = -0.5 / (g.lengthscale * g.lengthscale);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long j = static_cast<long>(jb) * NB + lane;
    for (int rr = wv; rr < NB; rr += THREADS / 64)
    {
        const long i = static_cast<long>(ib) * NB + rr;
        double val;
        if (i < g.n && j < g.n)
        {
            val = g.vertical * exp(scale * window_sqdist(g.x, i, g.x, j, g.n_reg));
            if (i == j)
                val += g.noise;
        }
        else
            val = (i == j) ? 1.0 : 0.0;
        D[rr * XS + lane] = val;
    }
}

// This lane's 16 entries of K_ij - sum_{k<jb} L_ik L_jk^T at the
// accumulator positions of the product's precision (left-looking step of
// block (ib, jb)). K_ij is generated from the series here (each block of K
// is reduced exactly once, so no pass assembles K first), through the
// stage's first block.
template <bool MIXED>
__device__ __forceinline__ void reduced_block(const BatchGP &g, int ib, int jb, Shared &sh, double (&kij)[16], bool forward = false,
                                              const double *carry = nullptr)
{
    const long np = g.n_pad;
    gen_block(g, ib, jb, sh.st.a);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; q++)
    {
        int r, c;
        acc_position<MIXED>(q, r, c);
        kij[q] = sh.st.a[r * XS + c];
    }
    __syncthreads();  // every entry is read before the stage is reused
    if constexpr (MIXED)
    {
        v4f acc[2][2];
        accumulate_row_products32(g.L, np, ib, jb, jb, sh.st32, acc);
#pragma unroll
        for (int q = 0; q < 16; q++)
            kij[q] -= static_cast<double>(acc[q >> 3][(q >> 2) & 1][q & 3]);
    }
    else
    {
        v4d acc[2][2];
        accumulate_row_products(g.L, np, ib, jb, jb, sh.st, acc, forward ? g.alpha : nullptr, forward ? sh.fa.T : nullptr, jb - 1, carry);
#pragma unroll
        for (int q = 0; q < 16; q++)
            kij[q] -= acc[q >> 3][(q >> 2) & 1][q & 3];
    }
}

// Diagonal block jb: reduced, then factored with its inverse in LDS; L_jj
// goes back to the factor, W_jj (row-major, zero above the diagonal) to
// the GP's inverses and, when wjj is given, to LDS (stride XS). In FP64 the
// step also takes alpha's forward solve one block further: the products of
// row block jb with z_0:jb come from the blocks the reduction streams
// through LDS anyway, and z_jb = W_jj (y_jb - those) goes to alpha.
// carry: the block's sums over its first jb - 1 k-blocks (FP64 only), from
// the previous level launch (diag_prefix_halves).
template <bool MIXED>
__device__ __forceinline__ void diag_block(const BatchGP &g, int gi, int jb, Shared &sh, double *wjj, int *info, const double *carry = nullptr)
{
    const long np = g.n_pad;
    double kij[16];
    reduced_block<MIXED>(g, jb, jb, sh, kij, !MIXED, carry);
#pragma unroll
    for (int q = 0; q < 16; q++)
    {
        int r, c;
        acc_position<MIXED>(q, r, c);
        sh.fa.D[r][c] = (c <= r) ? kij[q] : 0.0;
    }
    __syncthreads();
    factor_block64(sh.fa.D, sh.fa.W, sh.fa.R, info + gi, jb * NB);
    double *Cjj = g.L + static_cast<long>(jb) * NB * np + jb * NB;
    double *Wjj = g.W + static_cast<long>(jb) * NB * NB;
    for (int e = threadIdx.x; e < NB * NB; e += THREADS)
    {
        const int r = e / NB, c = e % NB;
        const double w = (c <= r) ? sh.fa.W[r][c] : 0.0;
        if (c <= r)
            Cjj[static_cast<long>(r) * np + c] = sh.fa.D[r][c];
        Wjj[e] = w;
        if (wjj)
            wjj[r * XS + c] = w;
    }
    if constexpr (!MIXED)
    {
        if (threadIdx.x < NB)
        {
            const int r = threadIdx.x;
            const double *yb = g.y + static_cast<long>(jb) * NB;
            double s = 0.0;
            for (int c = 0; c <= r; c++)
                s = fma(sh.fa.W[r][c], yb[c] - sh.fa.T[c], s);
            g.alpha[static_cast<long>(jb) * NB + r] = s;
        }
    }
    __syncthreads();
}

// Off-diagonal block (ib, jb), ib > jb: reduced, then L_ij = (K_ij - acc)
// W_jj^T with W_jj from LDS (wjj) or, when wjj is null, staged from the
// GP's inverses into the second half of the stage.
template <bool MIXED>
__device__ __forceinline__ void off_block(const BatchGP &g, int ib, int jb, Shared &sh, const double *wjj)
{
    const long np = g.n_pad;
    double kij[16];
    reduced_block<MIXED>(g, ib, jb, sh, kij);
#pragma unroll
    for (int q = 0; q < 16; q++)
    {
        int r, c;
        acc_position<MIXED>(q, r, c);
        sh.st.a[r * XS + c] = kij[q];
    }
    if (!wjj)
        blk::stage_block(sh.st.b, g.W + static_cast<long>(jb) * NB * NB, NB);
    __syncthreads();
    v4d xb[2][2];
    blk::mfma_nt(sh.st.a, wjj ? wjj : sh.st.b, xb);
    blk::store_acc(g.L + static_cast<long>(ib) * NB * np + jb * NB, np, xb, true, false);
    __syncthreads();
}

// loss = 0.5 (y^T alpha + sum log L_ii^2 + n log 2 pi) / n of one factored
// GP, summed by the workgroup's first THREADS threads in a fixed order (the
// same bits for any workgroup size); red: [2][THREADS].
__device__ __forceinline__ void gp_loss(const BatchGP &g, int gi, double (*red)[THREADS], double *losses)
{
    const int tid = threadIdx.x;
    const long np = g.n_pad;
    if (tid < THREADS)
    {
        double s1 = 0.0, s2 = 0.0;
        for (long i = tid; i < g.n; i += THREADS)
        {
            s1 += g.y[i] * g.alpha[i];
            const double d = g.L[i * np + i];
            s2 += log(d * d);
        }
        red[0][tid] = s1;
        red[1][tid] = s2;
    }
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

// alpha's FP64 refinement (mixed precision): r = y - K alpha with K
// regenerated from the series, delta = K^-1 r from the factor, alpha +=
// delta, for at most BATCH_REFINE_STEPS steps, until the correction (or the
// next one, as the contraction so far predicts it) is at most
// 1e-12 max|alpha|, or stops shrinking.

// Step `it`'s decision from its correction (dmax = max|delta|, amax =
// max|alpha|) and the previous step's dmax: whether delta is applied, and
// whether another step follows.
struct RefineDecision
{
    bool apply, more;
};
__device__ __forceinline__ RefineDecision refine_decision(int it, double dmax, double amax, double prev)
{
    // a correction that did not shrink is not applied
    if (it > 0 && dmax >= prev)
        return { false, false };
    const bool converged = dmax <= 1e-12 * amax || (it > 0 && dmax * (dmax / prev) <= 1e-12 * amax);
    return { true, !(converged || !(dmax < 0.5 * prev)) && it + 1 < BATCH_REFINE_STEPS };
}

// alpha (refined in FP64 for MIXED) and the loss of one factored GP, by a
// workgroup of NT threads (THREADS when MIXED). With forward_done, alpha
// already holds L^-1 y (the 64-wide FP64 diagonal steps compute it) and
// only the backward substitution is left.
template <bool MIXED, int NT>
__device__ __forceinline__ void finish_gp(const BatchGP &g, int gi, double (*red)[THREADS], double *losses, int *refine_steps,
                                          bool forward_done = false)
{
    static_assert(NT >= THREADS && (!MIXED || NT == THREADS), "finish_gp: workgroup size");
    const int tid = threadIdx.x;
    const double *y = g.y;
    const double *L = g.L;
    const double *Wg = g.W;
    double *alpha = g.alpha;
    const long np = g.n_pad;
    const int nblk = static_cast<int>(np / NB);

    // alpha: forward then backward block substitution with W
    double *xs = red[0];
    double *part = red[1];
    if (!forward_done)
    {
        for (long i = tid; i < np; i += NT)
            alpha[i] = (i < g.n) ? y[i] : 0.0;
        __syncthreads();
        forward_solve<NT>(L, Wg, np, nblk, alpha, xs, part);
    }
    backward_solve<NT>(L, Wg, np, nblk, alpha, xs, part);
    if constexpr (MIXED)
    {
        const double scale = -0.5 / (g.lengthscale * g.lengthscale);
        double *r = g.work;
        double prev = 1e300;
        int steps = 0;
        for (int it = 0; it < BATCH_REFINE_STEPS; it++)
        {
            residual_rows(g, scale, alpha, r);
            __syncthreads();
            block_solves<NT>(L, Wg, np, nblk, r, xs, part);
            double dmax = 0.0, amax = 0.0;
            for (long i = tid; i < np; i += NT)
            {
                dmax = fmax(dmax, fabs(r[i]));
                amax = fmax(amax, fabs(alpha[i]));
            }
            dmax = block_max<NT>(dmax, part);
            amax = block_max<NT>(amax, part);
            const RefineDecision d = refine_decision(it, dmax, amax, prev);
            if (!d.apply)
                break;
            for (long i = tid; i < np; i += NT)
                alpha[i] += r[i];
            __syncthreads();
            steps = it + 1;
            if (!d.more)
                break;
            prev = dmax;
        }
        if (refine_steps && tid == 0)
            refine_steps[gi] = steps;
    }
    gp_loss(g, gi, red, losses);
}

// Persistent path: each workgroup takes whole GPs from a device counter
// (largest first) and runs assembly, the left-looking factorisation and
// finish_gp without leaving the workgroup; W_jj stays in LDS for its column.
template <bool MIXED>
__global__ __launch_bounds__(THREADS) void k_batch_fit(const BatchGP *__restrict__ gps,
                                                        const int *__restrict__ order,
                                                        int count,
                                                        int *__restrict__ counter,
                                                        double *__restrict__ losses,
                                                        int *__restrict__ info,
                                                        int *__restrict__ refine_steps)
{
    __shared__ Shared sh;
    __shared__ double wjj[NB * XS];  // inverse of the current diagonal block, rows = B^T rows
    __shared__ int next;
    __shared__ double red[2][THREADS];
    const int tid = threadIdx.x;
    for (;;)
    {
        if (tid == 0)
            next = atomicAdd(counter, 1);
        __syncthreads();
        const int slot = __builtin_amdgcn_readfirstlane(next);  // uniform: the GP descriptor then comes through scalar loads
        __syncthreads();
        if (slot >= count)
            break;
        const int gi = order[slot];
        const BatchGP g = gps[gi];
        const int nblk = static_cast<int>(g.n_pad / NB);
        for (int jb = 0; jb < nblk; jb++)
        {
            diag_block<MIXED>(g, gi, jb, sh, wjj, info);
            for (int ib = jb + 1; ib < nblk; ib++)
                off_block<MIXED>(g, ib, jb, sh, wjj);
        }
        __syncthreads();  // the diagonal steps' alpha entries are visible to every thread
        finish_gp<MIXED, THREADS>(g, gi, red, losses, refine_steps, !MIXED);
    }
}

// ---- 64-wide level path, FP64 off-diagonal blocks on 32-deep stages -----
// The persistent kernel keeps whole 64x64 blocks in LDS (two staged blocks,
// the diagonal inverse, the factor's workspace: two workgroups per CU). An
// off-diagonal level step needs less: staging 32-deep halves of the block
// pairs takes 35 KB, so four workgroups share a CU and hide each other's
// staging. Each accumulator sees the same MFMAs in the same order as with
// whole blocks (k ascending), so the results are bit-identical.
using blk::XH;
union HalfShared
{
    struct
    {
        double a[NB * XH];
        double b[NB * XH];
    } h;
    double kb[NB * XS];  // the generated block of K (stride XS)
};

// off_block<false> with wjj == nullptr, on half stages.
__device__ __forceinline__ void off_block_halves(const BatchGP &g, int ib, int jb, HalfShared &sh)
{
    const long np = g.n_pad;
    double kij[16];
    gen_block(g, ib, jb, sh.kb);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; q++)
    {
        int r, c;
        acc_position<false>(q, r, c);
        kij[q] = sh.kb[r * XS + c];
    }
    __syncthreads();
    v4d acc[2][2];
    blk::zero(acc);
    const int steps = 2 * jb;
    if (steps > 0)
    {
        const double *ai = g.L + static_cast<long>(ib) * NB * np;
        const double *bj = g.L + static_cast<long>(jb) * NB * np;
        d2 ra[4], rb[4];
        blk::load_half(ai, np, ra);
        blk::load_half(bj, np, rb);
        for (int t = 0; t < steps; t++)
        {
            __syncthreads();  // every wavefront is done with the previous pair
            blk::store_half(sh.h.a, ra);
            blk::store_half(sh.h.b, rb);
            __syncthreads();
            if (t + 1 < steps)
            {
                blk::load_half(ai + (t + 1) * 32, np, ra);
                blk::load_half(bj + (t + 1) * 32, np, rb);
            }
            __builtin_amdgcn_sched_barrier(0);
            blk::mfma_nt_acc_half(sh.h.a, sh.h.b, acc);
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 16; q++)
        kij[q] -= acc[q >> 3][(q >> 2) & 1][q & 3];
    // L_ij = (K_ij - acc) W_jj^T, the 64-deep product as two halves of C's
    // and W_jj's columns
    v4d xb[2][2];
    blk::zero(xb);
    const double *W = g.W + static_cast<long>(jb) * NB * NB;
    for (int h = 0; h < 2; h++)
    {
        d2 rw[4];
        blk::load_half(W + h * 32, NB, rw);
#pragma unroll
        for (int q = 0; q < 16; q++)
        {
            int r, c;
            acc_position<false>(q, r, c);
            if ((c >> 5) == h)
                sh.h.a[r * XH + (c & 31)] = kij[q];
        }
        blk::store_half(sh.h.b, rw);
        __syncthreads();
        blk::mfma_nt_acc_half(sh.h.a, sh.h.b, xb);
        __syncthreads();
    }
    blk::store_acc(g.L + static_cast<long>(ib) * NB * np + jb * NB, np, xb, true, false);
}

// The mixed-precision level path's off-diagonal block (ib, jb) on 35 KB of
// LDS instead of 66: the generated block of K, then the FP32 block pairs,
// then the FP64 product with W_jj in 32-deep halves, each in turn in the
// same space (three workgroups per CU instead of two). The operations and
// their order are off_block<true>'s, so the results are the same bits.
union MixedHalfShared
{
    struct
    {
        double a[NB * XH];
        double b[NB * XH];
    } h;
    double kb[NB * XS];  // the generated block of K (stride XS)
    Stage32 st32;
};

__device__ __forceinline__ void off_block_mixed_halves(const BatchGP &g, int ib, int jb, MixedHalfShared &sh)
{
    const long np = g.n_pad;
    double kij[16];
    gen_block(g, ib, jb, sh.kb);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; q++)
    {
        int r, c;
        acc_position<true>(q, r, c);
        kij[q] = sh.kb[r * XS + c];
    }
    __syncthreads();  // every entry is read before the space is reused
    v4f acc[2][2];
    accumulate_row_products32(g.L, np, ib, jb, jb, sh.st32, acc);  // ends with a barrier
#pragma unroll
    for (int q = 0; q < 16; q++)
        kij[q] -= static_cast<double>(acc[q >> 3][(q >> 2) & 1][q & 3]);
    // L_ij = (K_ij - acc) W_jj^T, the 64-deep product as two halves of C's
    // and W_jj's columns
    v4d xb[2][2];
    blk::zero(xb);
    const double *W = g.W + static_cast<long>(jb) * NB * NB;
    for (int h = 0; h < 2; h++)
    {
        d2 rw[4];
        blk::load_half(W + h * 32, NB, rw);
#pragma unroll
        for (int q = 0; q < 16; q++)
        {
            int r, c;
            acc_position<true>(q, r, c);
            if ((c >> 5) == h)
                sh.h.a[r * XH + (c & 31)] = kij[q];
        }
        blk::store_half(sh.h.b, rw);
        __syncthreads();
        blk::mfma_nt_acc_half(sh.h.a, sh.h.b, xb);
        __syncthreads();
    }
    blk::store_acc(g.L + static_cast<long>(ib) * NB * np + jb * NB, np, xb, true, false);
}

// The first jn - 1 k-blocks of diagonal block jn's reduction (FP64), run in
// the off-diagonal launch of column jn - 1 beside that column's blocks, so
// the diagonal launch of column jn only adds block jn - 1 before it factors.
// The same MFMAs in the same order as accumulate_row_products (row block
// jn is both operands, staged once per half), and the same per-thread
// forward-solve sums: thread (row, part) takes its 16 columns of a block
// from the half that holds them. The state goes to g.pre (save_carry).
__device__ __forceinline__ void diag_prefix_halves(const BatchGP &g, int jn, HalfShared &sh)
{
    const long np = g.n_pad;
    const int trow = threadIdx.x & 63, tq = threadIdx.x >> 6;
    v4d acc[2][2];
    blk::zero(acc);
    double tp = 0.0;
    const int steps = 2 * (jn - 1);
    const double *aj = g.L + static_cast<long>(jn) * NB * np;
    d2 ra[4];
    blk::load_half(aj, np, ra);
    for (int t = 0; t < steps; t++)
    {
        __syncthreads();  // every wavefront is done with the previous half
        blk::store_half(sh.h.a, ra);
        __syncthreads();
        if (t + 1 < steps)
            blk::load_half(aj + (t + 1) * 32, np, ra);
        __builtin_amdgcn_sched_barrier(0);
        blk::mfma_nt_acc_half(sh.h.a, sh.h.a, acc);
        if ((tq >> 1) == (t & 1))
        {
            const double *zb = g.alpha + (t >> 1) * NB + tq * 16;
            const double *arow = sh.h.a + trow * XH + (tq & 1) * 16;
#pragma unroll 4
            for (int c = 0; c < 16; c++)
                tp = fma(arow[c], zb[c], tp);
        }
    }
    save_carry(g.pre, acc, tp);
}

// Level-synchronous path for large GPs, one launch per step over all of
// them: workgroup w takes item w of its launch's list (GP index and block
// row). The same block functions run in the same order per block as in the
// persistent path, so a GP gets bit-identical results on either path.
__global__ __launch_bounds__(THREADS) void k_batch_assemble_level(const BatchGP *__restrict__ gps, const int2 *__restrict__ items)
{
    const int2 it = items[blockIdx.x];
    const BatchGP g = gps[it.x];
    assemble_rows(g, static_cast<long>(it.y) * NB, static_cast<long>(it.y + 1) * NB);
}

template <bool MIXED>
__global__ __launch_bounds__(THREADS) void k_batch_diag_level(const BatchGP *__restrict__ gps, const int2 *__restrict__ items, int jb,
                                                               int *__restrict__ info)
{
    __shared__ Shared sh;
    const int2 it = items[blockIdx.x];
    const BatchGP g = gps[it.x];
    // FP64 steps from column 2 on continue the sums the previous launch began
    diag_block<MIXED>(g, it.x, jb, sh, nullptr, info, !MIXED && jb >= 2 ? g.pre : nullptr);
}

template <bool MIXED>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(3))) void k_batch_off_level(const BatchGP *__restrict__ gps, const int2 *__restrict__ items, int jb)
{
    const int2 it = items[blockIdx.x];
    const BatchGP g = gps[it.x];
    if constexpr (MIXED)
    {
        __shared__ MixedHalfShared sh;
        off_block_mixed_halves(g, it.y, jb, sh);
    }
    else
    {
        __shared__ HalfShared sh;
        if (it.y < 0)
            diag_prefix_halves(g, jb + 1, sh);  // item {GP, -1}: the next diagonal block's first sums
        else
            off_block_halves(g, it.y, jb, sh);
    }
}

// FP64 finish of the level path: alpha's backward substitution (and the
// forward one unless the diagonal steps did it) and the loss, by a
// workgroup of SOLVE_THREADS threads: the substitutions stream each GP's
// factor through one CU, so more wavefronts keep more of it in flight.
constexpr int SOLVE_THREADS = 512;

__global__ __launch_bounds__(SOLVE_THREADS) void k_batch_finish_level(const BatchGP *__restrict__ gps, const int2 *__restrict__ items,
                                                                       double *__restrict__ losses, int forward_done)
{
    __shared__ double red[2][THREADS];
    const int2 it = items[blockIdx.x];
    const BatchGP g = gps[it.x];
    finish_gp<false, SOLVE_THREADS>(g, it.x, red, losses, nullptr, forward_done != 0);
}

// ---- mixed precision: alpha's refinement over several launches ----------
// finish_gp's refinement loop spread over launches, so that its costly part,
// the residual (K regenerated in FP64, n^2 exponentials per step), runs on
// one wavefront per 64-row block instead of one workgroup per GP: a start
// launch solves alpha = K^-1 y from the factor; each of BATCH_REFINE_STEPS
// pairs of launches forms r = y - K alpha (per block), then solves for the
// correction and decides (per GP); a finish launch writes the loss and the
// step count. A GP's state (previous correction, whether it goes on, steps)
// lives in g.pre; a GP that stopped returns at once from the remaining
// launches. Per GP the operations and their order are finish_gp's, so the
// persistent and level paths give the same bits.
__global__ __launch_bounds__(SOLVE_THREADS) void k_batch_refine_start(const BatchGP *__restrict__ gps, const int2 *__restrict__ items)
{
    __shared__ double red[2][THREADS];
    const BatchGP g = gps[items[blockIdx.x].x];
    const long np = g.n_pad;
    for (long i = threadIdx.x; i < np; i += SOLVE_THREADS)
        g.alpha[i] = (i < g.n) ? g.y[i] : 0.0;
    __syncthreads();
    block_solves<SOLVE_THREADS>(g.L, g.W, np, static_cast<int>(np / NB), g.alpha, red[0], red[1]);
    if (threadIdx.x == 0)
    {
        g.pre[0] = 1e300;  // previous correction
        g.pre[1] = 1.0;    // refining
        g.pre[2] = 0.0;    // steps applied
    }
}

// Items {GP, block row}: r over the block's 64 rows, one row per lane.
__global__ __launch_bounds__(NB) void k_batch_refine_residual(const BatchGP *__restrict__ gps, const int2 *__restrict__ items)
{
    const int2 it = items[blockIdx.x];
    const BatchGP g = gps[it.x];
    if (g.pre[1] == 0.0)
        return;
    const double scale = -0.5 / (g.lengthscale * g.lengthscale);
    const long i = static_cast<long>(it.y) * NB + threadIdx.x;
    g.work[i] = i < g.n ? residual_row(g, scale, g.alpha, i) : 0.0;
}

__global__ __launch_bounds__(SOLVE_THREADS) void k_batch_refine_correct(const BatchGP *__restrict__ gps, const int2 *__restrict__ items, int it)
{
    __shared__ double xs[NB];
    __shared__ double red[SOLVE_THREADS];
    const int tid = threadIdx.x;
    const BatchGP g = gps[items[blockIdx.x].x];
    if (g.pre[1] == 0.0)
        return;
    const double prev = g.pre[0];
    const long np = g.n_pad;
    double *r = g.work;
    block_solves<SOLVE_THREADS>(g.L, g.W, np, static_cast<int>(np / NB), r, xs, red);
    double dmax = 0.0, amax = 0.0;
    for (long i = tid; i < np; i += SOLVE_THREADS)
    {
        dmax = fmax(dmax, fabs(r[i]));
        amax = fmax(amax, fabs(g.alpha[i]));
    }
    dmax = block_max<SOLVE_THREADS>(dmax, red);
    amax = block_max<SOLVE_THREADS>(amax, red);
    const RefineDecision d = refine_decision(it, dmax, amax, prev);
    if (d.apply)
        for (long i = tid; i < np; i += SOLVE_THREADS)
            g.alpha[i] += r[i];
    if (tid == 0)
    {
        if (d.apply)
        {
            g.pre[0] = dmax;
            g.pre[2] = it + 1;
        }
        g.pre[1] = d.more ? 1.0 : 0.0;
    }
}

__global__ __launch_bounds__(THREADS) void k_batch_refine_finish(const BatchGP *__restrict__ gps, const int2 *__restrict__ items,
                                                                  double *__restrict__ losses, int *__restrict__ refine_steps)
{
    __shared__ double red[2][THREADS];
    const int gi = items[blockIdx.x].x;
    const BatchGP g = gps[gi];
    if (refine_steps && threadIdx.x == 0)
        refine_steps[gi] = static_cast<int>(g.pre[2]);
    gp_loss(g, gi, red, losses);
}

// ---------------------------------------------------------------------------
// Level-synchronous path on 128-wide block columns: the accumulation of a
// 128x128 block, C -= L_i,0:j0 L_j,0:j0^T, is one 128x128 matrix-core tile of
// k_gemm.hip (K = j0 in 16-deep slabs), where the 64-wide path runs 64x64
// block pairs. The diagonal step factors the 128x128 block as two 64x64
// blocks (L00, L10 = D10 W00^T, D11 -= L10 L10^T, L11) and forms its inverse
// [[W00, 0], [-W11 L10 W00, W11]] into the GP's W128 (row-major 128x128 per
// block column); the 64x64 inverses also go to W, where the solves and
// predictions read them. The off-diagonal step then computes L_ij = C_ij
// W128^T in place (a second tile product with K = the block's width).

union Shared128
{
    struct
    {
        double As[2][gt::BK * gt::LDS_STRIDE];
        double Bs[2][gt::BK * gt::LDS_STRIDE];
    } g;
    Factor fa;
    Stage st;
};

// C(rows x cols block at (i0, j0) of the GP's factor) -= sum over the first
// j0 columns of L_i L_j^T (lower only when i0 == j0).
__device__ __forceinline__ void tile_update(const BatchGP &g, long i0, long j0, int rows, int cols, Shared128 &sh)
{
    if (j0 == 0)
        return;
    const long np = g.n_pad;
    gt::Params p{};
    p.A = g.L + i0 * np;  // A(i, k) = L[i0 + i][k]
    p.lda = np;
    p.B = g.L + j0 * np;  // B(k, j) = L[j0 + j][k]
    p.ldb = np;
    p.C = g.L + i0 * np + j0;
    p.ldc = np;
    p.M = rows;
    p.N = cols;
    p.K = static_cast<int>(j0);
    p.lower = i0 == j0 ? 1 : 0;
    p.diag_off = 0;
    p.set = 0;
    gt::gemm_tile<false, true>(p, 0, 0, sh.g.As, sh.g.Bs);
    __syncthreads();
}

// Diagonal step of 128-wide block column jb of one GP (the block may be 64
// wide at the end of the matrix).
__device__ __forceinline__ void diag_block128(const BatchGP &g, int gi, int jb, Shared128 &sh, int *info)
{
    const long np = g.n_pad;
    const long j0 = static_cast<long>(jb) * 128;
    const int w = static_cast<int>(min(static_cast<long>(128), np - j0));
    tile_update(g, j0, j0, w, w, sh);
    double *C = g.L + j0 * np + j0;
    double *W00 = g.W + (j0 / NB) * NB * NB;
    double *Wt = g.W128 + static_cast<long>(jb) * 128 * 128;
    const int tid = threadIdx.x;
    // D00 -> L00, W00
    for (int e = tid; e < NB * NB; e += THREADS)
    {
        const int r = e / NB, c = e % NB;
        sh.fa.D[r][c] = (c <= r) ? C[static_cast<long>(r) * np + c] : 0.0;
    }
    __syncthreads();
    factor_block64(sh.fa.D, sh.fa.W, sh.fa.R, info + gi, static_cast<int>(j0));
    for (int e = tid; e < NB * NB; e += THREADS)
    {
        const int r = e / NB, c = e % NB;
        const double wv = (c <= r) ? sh.fa.W[r][c] : 0.0;
        if (c <= r)
            C[static_cast<long>(r) * np + c] = sh.fa.D[r][c];
        W00[e] = wv;
        Wt[r * 128 + c] = wv;
        if (w == 128)
            Wt[r * 128 + NB + c] = 0.0;
    }
    __syncthreads();
    if (w < 128)
        return;
    double *C10 = C + NB * np, *C11 = C + NB * np + NB;
    double *W11 = W00 + NB * NB;
    v4d acc[2][2];
    // L10 = D10 W00^T
    blk::stage_block(sh.st.a, C10, np);
    blk::stage_block(sh.st.b, W00, NB);
    __syncthreads();
    blk::mfma_nt(sh.st.a, sh.st.b, acc);
    __syncthreads();
    blk::store_acc(C10, np, acc, true, false);
    __syncthreads();
    // D11 = C11 - L10 L10^T -> L11, W11
    blk::stage_block(sh.st.a, C10, np);
    blk::stage_block(sh.st.b, C10, np);
    __syncthreads();
    blk::mfma_nt(sh.st.a, sh.st.b, acc);
    double c11[16];
    blk::load_at_acc(C11, np, c11);
    __syncthreads();
    int q = 0;
    blk::for_each(acc,
                  [&](int r, int c, double v)
                  {
                      sh.fa.D[r][c] = (c <= r) ? c11[q] - v : 0.0;
                      q++;
                  });
    __syncthreads();
    factor_block64(sh.fa.D, sh.fa.W, sh.fa.R, info + gi, static_cast<int>(j0 + NB));
    for (int e = tid; e < NB * NB; e += THREADS)
    {
        const int r = e / NB, c = e % NB;
        const double wv = (c <= r) ? sh.fa.W[r][c] : 0.0;
        if (c <= r)
            C11[static_cast<long>(r) * np + c] = sh.fa.D[r][c];
        W11[e] = wv;
        Wt[(NB + r) * 128 + NB + c] = wv;
    }
    __syncthreads();
    // W10 = -W11 (L10 W00): T = L10 W00 first (B = W00^T staged transposed)
    blk::stage_block(sh.st.a, C10, np);
    blk::stage_block_t(sh.st.b, W00, NB);
    __syncthreads();
    blk::mfma_nt(sh.st.a, sh.st.b, acc);
    __syncthreads();
    // T^T into the second stage half, W11 into the first
    blk::for_each(acc, [&](int r, int c, double v) { sh.st.b[c * XS + r] = v; });
    blk::stage_block(sh.st.a, W11, NB);
    __syncthreads();
    blk::mfma_nt(sh.st.a, sh.st.b, acc);
    blk::for_each(acc, [&](int r, int c, double v) { Wt[(NB + r) * 128 + c] = -v; });
    __syncthreads();
}

// Off-diagonal step: block row i of block column jb, L_ij = (C_ij - acc) W128_jj^T.
__device__ __forceinline__ void off_block128(const BatchGP &g, int ib, int jb, Shared128 &sh)
{
    const long np = g.n_pad;
    const long i0 = static_cast<long>(ib) * 128, j0 = static_cast<long>(jb) * 128;
    const int rows = static_cast<int>(min(static_cast<long>(128), np - i0));
    const int w = static_cast<int>(min(static_cast<long>(128), np - j0));
    tile_update(g, i0, j0, rows, w, sh);
    gt::Params p{};
    p.A = g.L + i0 * np + j0;  // the updated block, read whole before the tile is written
    p.lda = np;
    p.B = g.W128 + static_cast<long>(jb) * 128 * 128;  // B(k, c) = W(c, k)
    p.ldb = 128;
    p.C = g.L + i0 * np + j0;
    p.ldc = np;
    p.M = rows;
    p.N = w;
    p.K = w;
    p.set = 1;
    gt::gemm_tile<false, true>(p, 0, 0, sh.g.As, sh.g.Bs);
}

__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_batch_diag_level128(const BatchGP *__restrict__ gps,
                                                                                                           const int2 *__restrict__ items,
                                                                                                           int jb, int *__restrict__ info)
{
    __shared__ Shared128 sh;
    const int2 it = items[blockIdx.x];
    const BatchGP g = gps[it.x];
    diag_block128(g, it.x, jb, sh, info);
}

__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_batch_off_level128(const BatchGP *__restrict__ gps,
                                                                                                          const int2 *__restrict__ items,
                                                                                                          int jb)
{
    __shared__ Shared128 sh;
    const int2 it = items[blockIdx.x];
    const BatchGP g = gps[it.x];
    off_block128(g, it.y, jb, sh);
}

__global__ void k_set_ints(int *p, int count, int value)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        p[i] = value;
}

}  // namespace

void batch_info_reset(hipStream_t s, int *info, int count)
{
    require(count >= 1, "batch: bad count");
    hipLaunchKernelGGL(k_set_ints, dim3(ceil_div(count, 256)), dim3(256), 0, s, info, count, INT_MAX);
    GP_CHECK_LAUNCH();
}

void batch_fit(hipStream_t s, const BatchGP *gps, const int *order, int count, int *counter, double *losses, int *info,
               int workgroups, bool mixed, int *refine_steps)
{
    require(count >= 1 && workgroups >= 1, "batch_fit: bad shape");
    hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(1), 0, s, counter, 1, 0);
    GP_CHECK_LAUNCH();
    if (mixed)
        hipLaunchKernelGGL(k_batch_fit<true>, dim3(workgroups), dim3(THREADS), 0, s, gps, order, count, counter, losses, info, refine_steps);
    else
        hipLaunchKernelGGL(k_batch_fit<false>, dim3(workgroups), dim3(THREADS), 0, s, gps, order, count, counter, losses, info,
                           refine_steps);
    GP_CHECK_LAUNCH();
}

void batch_level_step(hipStream_t s, BatchLevelStep step, const BatchGP *gps, const int2 *items, int n_items, int jb, bool mixed,
                      double *losses, int *info, int *refine_steps)
{
    if (n_items <= 0)
        return;
    const dim3 grid(static_cast<unsigned>(n_items));
    switch (step)
    {
    case BatchLevelStep::Assemble:
        hipLaunchKernelGGL(k_batch_assemble_level, grid, dim3(THREADS), 0, s, gps, items);
        break;
    case BatchLevelStep::Diagonal:
        if (mixed)
            hipLaunchKernelGGL(k_batch_diag_level<true>, grid, dim3(THREADS), 0, s, gps, items, jb, info);
        else
            hipLaunchKernelGGL(k_batch_diag_level<false>, grid, dim3(THREADS), 0, s, gps, items, jb, info);
        break;
    case BatchLevelStep::OffDiagonal:
        if (mixed)
            hipLaunchKernelGGL(k_batch_off_level<true>, grid, dim3(THREADS), 0, s, gps, items, jb);
        else
            hipLaunchKernelGGL(k_batch_off_level<false>, grid, dim3(THREADS), 0, s, gps, items, jb);
        break;
    case BatchLevelStep::Diagonal128:
        require(!mixed, "batch: the 128-wide level path is FP64");
        hipLaunchKernelGGL(k_batch_diag_level128, grid, dim3(THREADS), 0, s, gps, items, jb, info);
        break;
    case BatchLevelStep::OffDiagonal128:
        require(!mixed, "batch: the 128-wide level path is FP64");
        hipLaunchKernelGGL(k_batch_off_level128, grid, dim3(THREADS), 0, s, gps, items, jb);
        break;
    case BatchLevelStep::Finish:
        // the 64-wide FP64 diagonal steps did alpha's forward solve
        require(!mixed, "batch: the mixed-precision level path ends with the refinement steps");
        hipLaunchKernelGGL(k_batch_finish_level, grid, dim3(SOLVE_THREADS), 0, s, gps, items, losses, 1);
        break;
    case BatchLevelStep::Finish128:
        require(!mixed, "batch: the 128-wide level path is FP64");
        hipLaunchKernelGGL(k_batch_finish_level, grid, dim3(SOLVE_THREADS), 0, s, gps, items, losses, 0);
        break;
    case BatchLevelStep::RefineStart:
        require(mixed, "batch: refinement steps are mixed precision");
        hipLaunchKernelGGL(k_batch_refine_start, grid, dim3(SOLVE_THREADS), 0, s, gps, items);
        break;
    case BatchLevelStep::RefineResidual:
        require(mixed, "batch: refinement steps are mixed precision");
        hipLaunchKernelGGL(k_batch_refine_residual, grid, dim3(NB), 0, s, gps, items);
        break;
    case BatchLevelStep::RefineCorrect:
        require(mixed && jb >= 0 && jb < BATCH_REFINE_STEPS, "batch: bad refinement step");
        hipLaunchKernelGGL(k_batch_refine_correct, grid, dim3(SOLVE_THREADS), 0, s, gps, items, jb);
        break;
    case BatchLevelStep::RefineFinish:
        require(mixed, "batch: refinement steps are mixed precision");
        hipLaunchKernelGGL(k_batch_refine_finish, grid, dim3(THREADS), 0, s, gps, items, losses, refine_steps);
        break;
    }
    GP_CHECK_LAUNCH();
}

}  // namespace gpamd::k
