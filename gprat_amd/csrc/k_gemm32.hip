// FP32 matrix-core trailing update for the mixed-precision factorisation.
//
// The factored FP64 panel is first copied to FP32 in k-major order
// (P[k][row], `panel_to_f32t`), so both operands of C -= A B^T are read as
// contiguous 16-byte row segments and land in LDS without a transpose.
// One workgroup (4 wavefronts) owns a 128x128 tile of the FP64 trailing
// matrix; each wavefront accumulates a 64x64 quarter in FP32 on
// v_mfma_f32_16x16x4f32 over 32-deep LDS slabs (double buffered), and the
// epilogue subtracts the FP32 product from the FP64 matrix.
#include "kernels.hpp"

namespace gpamd::k
{
namespace
{

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int BM = 128;
constexpr int BK = 32;
constexpr int KS = BM + 16;  // floats per LDS k-row: the 4 k-slots of a read hit disjoint banks
constexpr int THREADS = 256;

struct Params32
{
    const float *P;  // k-major panel copy
    long ldp;
    long a_off, b_off;  // row offsets of the A and B operands inside P
    double *C;
    long ldc;
    int M, N, K;
    long diag_off;
};

// Element offsets (within one slab) of this thread's four float4 loads,
// clamped to the operand's last row pair; rows past the edge only feed
// accumulators the epilogue never stores.
__device__ __forceinline__ void slab_offsets(long ldp, int rows, int tid, unsigned (&off)[4], int (&lds)[4])
{
#pragma unroll
    for (int s = 0; s < 4; s++)
    {
        const int idx = tid + THREADS * s;
        const int kk = idx >> 5;         // 0..31
        const int col = (idx & 31) * 4;  // 0..124
        off[s] = static_cast<unsigned>(kk * ldp + min(col, rows - 4));
        lds[s] = kk * KS + col;
    }
}

__device__ __forceinline__ void mfma_slab32(const float *As, const float *Bs, int wr, int wc, int lm, int lk, v4f (&acc)[4][4])
{
#pragma unroll
    for (int kq = 0; kq < BK / 4; kq++)
    {
        float a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; i++)
            a[i] = As[(kq * 4 + lk) * KS + wr * 64 + i * 16 + lm];
#pragma unroll
        for (int j = 0; j < 4; j++)
            b[j] = Bs[(kq * 4 + lk) * KS + wc * 64 + j * 16 + lm];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
}

__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(2))) void k_gemm_f32_trailing(Params32 p)
{
    __shared__ float As[2][BK * KS];
    __shared__ float Bs[2][BK * KS];

    const long r0 = static_cast<long>(blockIdx.y) * BM;
    const long c0 = static_cast<long>(blockIdx.x) * BM;
    if (r0 + (BM - 1) + p.diag_off < c0)
        return;  // tile entirely above the diagonal
    const int rows_a = static_cast<int>(min(static_cast<long>(BM), p.M - r0));
    const int rows_b = static_cast<int>(min(static_cast<long>(BM), p.N - c0));

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int lm = lane & 15, lk = lane >> 4;

    v4f acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
            acc[i][j] = v4f{ 0.0f, 0.0f, 0.0f, 0.0f };

    unsigned oa[4], ob[4];
    int la[4], lb[4];
    slab_offsets(p.ldp, rows_a, tid, oa, la);
    slab_offsets(p.ldp, rows_b, tid, ob, lb);
    const float *ak = p.P + p.a_off + r0;
    const float *bk = p.P + p.b_off + c0;
    const long step = BK * p.ldp;

    float4 ra[4], rb[4];
#pragma unroll
    for (int s = 0; s < 4; s++)
    {
        ra[s] = *reinterpret_cast<const float4 *>(ak + oa[s]);
        rb[s] = *reinterpret_cast<const float4 *>(bk + ob[s]);
    }
#pragma unroll
    for (int s = 0; s < 4; s++)
    {
        *reinterpret_cast<float4 *>(&As[0][la[s]]) = ra[s];
        *reinterpret_cast<float4 *>(&Bs[0][lb[s]]) = rb[s];
    }
    __syncthreads();

    int buf = 0;
    const int steps = p.K / BK;
    for (int t = 1; t < steps; t++)
    {
        ak += step;
        bk += step;
#pragma unroll
        for (int s = 0; s < 4; s++)
        {
            ra[s] = *reinterpret_cast<const float4 *>(ak + oa[s]);
            rb[s] = *reinterpret_cast<const float4 *>(bk + ob[s]);
        }
        mfma_slab32(As[buf], Bs[buf], wr, wc, lm, lk, acc);
#pragma unroll
        for (int s = 0; s < 4; s++)
        {
            *reinterpret_cast<float4 *>(&As[buf ^ 1][la[s]]) = ra[s];
            *reinterpret_cast<float4 *>(&Bs[buf ^ 1][lb[s]]) = rb[s];
        }
        __syncthreads();
        buf ^= 1;
    }
    mfma_slab32(As[buf], Bs[buf], wr, wc, lm, lk, acc);

    // Accumulator layout of the 16x16x4 f32 MFMA on gfx950: lane holds
    // column (lane % 16) and rows 4 (lane / 16) + e, e = 0..3.
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int e = 0; e < 4; e++)
            {
                const int rr = wr * 64 + i * 16 + lk * 4 + e;
                const int cc = wc * 64 + j * 16 + lm;
                if (rr >= rows_a || cc >= rows_b)
                    continue;
                const long gi = r0 + rr, gj = c0 + cc;
                if (gi + p.diag_off < gj)
                    continue;
                p.C[gi * p.ldc + gj] -= static_cast<double>(acc[i][j][e]);
            }
}

// P[k][r] = (float) A[(r0 + r) * lda + p0 + k] for r < R, k < W, through a
// 64x64 LDS tile so both the FP64 reads and the FP32 writes are coalesced.
__global__ __launch_bounds__(THREADS) void k_panel_to_f32t(const double *__restrict__ A, long lda, long r0, long p0, int R, int W,
                                                            float *__restrict__ P, long ldp)
{
    __shared__ float t[64][65];
    const long rb = static_cast<long>(blockIdx.x) * 64, kb = static_cast<long>(blockIdx.y) * 64;
    const int c = threadIdx.x & 63, q = threadIdx.x >> 6;
    for (int rr = q; rr < 64; rr += 4)
    {
        const long r = rb + rr, k = kb + c;
        t[rr][c] = (r < R && k < W) ? static_cast<float>(A[(r0 + r) * lda + p0 + k]) : 0.0f;
    }
    __syncthreads();
    for (int kk = q; kk < 64; kk += 4)
    {
        const long k = kb + kk, r = rb + c;
        if (k < W && r < R)
            P[k * ldp + r] = t[c][kk];
    }
}

}  // namespace

void panel_to_f32t(hipStream_t s, const double *A, long lda, long r0, long p0, int R, int W, float *P, long ldp)
{
    require(R >= 0 && W >= 0 && ldp >= R, "panel_to_f32t: bad shape");
    if (R == 0 || W == 0)
        return;
    dim3 grid(static_cast<unsigned>(ceil_div(R, 64)), static_cast<unsigned>(ceil_div(W, 64)));
    hipLaunchKernelGGL(k_panel_to_f32t, grid, dim3(THREADS), 0, s, A, lda, r0, p0, R, W, P, ldp);
    GP_CHECK_LAUNCH();
}

void gemm_f32_trailing(hipStream_t s, const float *P, long ldp, long a_off, long b_off, double *C, long ldc, int M, int N, int K, long diag_off)
{
    if (M <= 0 || N <= 0)
        return;
    require(K % BK == 0, "gemm_f32_trailing: K must be a multiple of 32");
    require(M % 4 == 0 && N % 4 == 0 && ldp % 4 == 0 && a_off % 4 == 0 && b_off % 4 == 0,
            "gemm_f32_trailing: rows must come in float4 groups");
    require(static_cast<long>(BK) * ldp + BM < (1L << 31), "gemm_f32_trailing: panel too tall");
    Params32 p{ P, ldp, a_off, b_off, C, ldc, M, N, K, diag_off };
    dim3 grid(static_cast<unsigned>(ceil_div(N, BM)), static_cast<unsigned>(ceil_div(M, BM)));
    require(grid.y <= 65535, "gemm_f32_trailing: M too large for one launch");
    hipLaunchKernelGGL(k_gemm_f32_trailing, grid, dim3(THREADS), 0, s, p);
    GP_CHECK_LAUNCH();
}

}  // namespace gpamd::k
