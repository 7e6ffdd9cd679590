// 64x64 FP64 block products on the matrix cores for one 256-thread
// workgroup, shared by the kernels that keep whole 64x64 blocks in LDS (the
// batched fit and the 256-block inverses). Header-only so each kernel file
// inlines them.
#pragma once

#include <hip/hip_runtime.h>

namespace gpamd::k::blk
{

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int NB = 64;
constexpr int THREADS = 256;
constexpr int XS = NB + 2;  // LDS row stride of staged blocks (bank-spread MFMA reads)

typedef double d2 __attribute__((ext_vector_type(2)));

// Copy a 64x64 row-major block (leading dimension ld) into LDS stride XS:
// all eight 16-byte loads of a thread are issued before the first LDS
// store, so the block costs one memory latency, not eight.
__device__ __forceinline__ void stage_block(double *dst, const double *__restrict__ src, long ld)
{
    d2 v[8];
#pragma unroll
    for (int s = 0; s < 8; s++)
    {
        const int e = threadIdx.x + THREADS * s;
        v[s] = *reinterpret_cast<const d2 *>(src + static_cast<long>(e >> 5) * ld + (e & 31) * 2);
    }
#pragma unroll
    for (int s = 0; s < 8; s++)
    {
        const int e = threadIdx.x + THREADS * s;
        *reinterpret_cast<d2 *>(dst + (e >> 5) * XS + (e & 31) * 2) = v[s];
    }
}

// Copy a 64x64 row-major block into LDS transposed: dst[c][r] = src[r][c]
// (stride XS); loads as in stage_block, then scattered LDS stores.
__device__ __forceinline__ void stage_block_t(double *dst, const double *__restrict__ src, long ld)
{
    d2 v[8];
#pragma unroll
    for (int s = 0; s < 8; s++)
    {
        const int e = threadIdx.x + THREADS * s;
        v[s] = *reinterpret_cast<const d2 *>(src + static_cast<long>(e >> 5) * ld + (e & 31) * 2);
    }
#pragma unroll
    for (int s = 0; s < 8; s++)
    {
        const int e = threadIdx.x + THREADS * s;
        const int r = e >> 5, c = (e & 31) * 2;
        dst[c * XS + r] = v[s][0];
        dst[(c + 1) * XS + r] = v[s][1];
    }
}

// stage_block / stage_block_t in two halves, so the loads of the next
// block can be in flight while the current one is multiplied: load_block
// issues this thread's eight 16-byte loads, store_block / store_block_t
// write them to LDS (stride XS; transposed: dst[c][r] = src[r][c]).
__device__ __forceinline__ void load_block(const double *__restrict__ src, long ld, d2 (&v)[8])
{
#pragma unroll
    for (int s = 0; s < 8; s++)
    {
        const int e = threadIdx.x + THREADS * s;
        v[s] = *reinterpret_cast<const d2 *>(src + static_cast<long>(e >> 5) * ld + (e & 31) * 2);
    }
}

__device__ __forceinline__ void store_block(double *dst, const d2 (&v)[8])
{
#pragma unroll
    for (int s = 0; s < 8; s++)
    {
        const int e = threadIdx.x + THREADS * s;
        *reinterpret_cast<d2 *>(dst + (e >> 5) * XS + (e & 31) * 2) = v[s];
    }
}

__device__ __forceinline__ void store_block_t(double *dst, const d2 (&v)[8])
{
#pragma unroll
    for (int s = 0; s < 8; s++)
    {
        const int e = threadIdx.x + THREADS * s;
        const int r = e >> 5, c = (e & 31) * 2;
        dst[c * XS + r] = v[s][0];
        dst[(c + 1) * XS + r] = v[s][1];
    }
}

__device__ __forceinline__ void zero(v4d (&acc)[2][2])
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
            acc[i][j] = v4d{ 0.0, 0.0, 0.0, 0.0 };
}

// acc (this wavefront's 32x32 quarter of a 64x64 block) += sum_k A[r][k] B[c][k]
// with A and B staged in LDS (stride XS). Wavefront w owns rows
// 32*(w>>1).., columns 32*(w&1)..
__device__ __forceinline__ void mfma_nt_acc(const double *A, const double *B, v4d (&acc)[2][2])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const int r0 = (w >> 1) * 32, c0 = (w & 1) * 32;
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

// ---- 32-deep half stages -------------------------------------------------
// A product of two 64x64 blocks staged as two 64x32 halves per operand (LDS
// stride XH) needs half the LDS of whole blocks, so twice the workgroups fit
// a CU; mfma_nt_acc_half runs the MFMAs of one half in mfma_nt_acc's order,
// so the two halves give bit-identical results to one whole-block product.
constexpr int XH = 32 + 2;

// this thread's share of a 64x32 part of a row-major matrix (rows e >> 4,
// columns (e & 15) * 2, e = tid + 256 s)
__device__ __forceinline__ void load_half(const double *__restrict__ src, long ld, d2 (&v)[4])
{
#pragma unroll
    for (int s = 0; s < 4; s++)
    {
        const int e = threadIdx.x + THREADS * s;
        v[s] = *reinterpret_cast<const d2 *>(src + static_cast<long>(e >> 4) * ld + (e & 15) * 2);
    }
}

__device__ __forceinline__ void store_half(double *dst, const d2 (&v)[4])
{
#pragma unroll
    for (int s = 0; s < 4; s++)
    {
        const int e = threadIdx.x + THREADS * s;
        *reinterpret_cast<d2 *>(dst + (e >> 4) * XH + (e & 15) * 2) = v[s];
    }
}

// this thread's share of a 32x64 part (rows e >> 5, columns (e & 31) * 2),
// to be stored transposed: dst[c][r] = src[r][c] (stride XH)
__device__ __forceinline__ void load_half_t(const double *__restrict__ src, long ld, d2 (&v)[4])
{
#pragma unroll
    for (int s = 0; s < 4; s++)
    {
        const int e = threadIdx.x + THREADS * s;
        v[s] = *reinterpret_cast<const d2 *>(src + static_cast<long>(e >> 5) * ld + (e & 31) * 2);
    }
}

__device__ __forceinline__ void store_half_t(double *dst, const d2 (&v)[4])
{
#pragma unroll
    for (int s = 0; s < 4; s++)
    {
        const int e = threadIdx.x + THREADS * s;
        const int r = e >> 5, c = (e & 31) * 2;
        dst[c * XH + r] = v[s][0];
        dst[(c + 1) * XH + r] = v[s][1];
    }
}

__device__ __forceinline__ void mfma_nt_acc_half(const double *A, const double *B, v4d (&acc)[2][2])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const int r0 = (w >> 1) * 32, c0 = (w & 1) * 32;
#pragma unroll 4
    for (int k0 = 0; k0 < 32; k0 += 4)
    {
        double a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++)
            a[i] = A[(r0 + i * 16 + lm) * XH + k0 + lk];
#pragma unroll
        for (int j = 0; j < 2; j++)
            b[j] = B[(c0 + j * 16 + lm) * XH + k0 + lk];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
}

__device__ __forceinline__ void mfma_nt(const double *A, const double *B, v4d (&acc)[2][2])
{
    zero(acc);
    mfma_nt_acc(A, B, acc);
}

// Visit this lane's accumulator entries as f(row, col, value) within the
// 64x64 block. f64 16x16x4 accumulator layout (measured on gfx950): lane
// holds column lane%16, rows lane/16 + 4e.
template <class F>
__device__ __forceinline__ void for_each(const v4d (&acc)[2][2], F f)
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
                f(r0 + i * 16 + lk + 4 * e, c0 + j * 16 + lm, acc[i][j][e]);
}

// The 16 entries of a 64x64 global block (leading dimension ld) at this
// lane's accumulator positions, in for_each order; all loads are issued
// before any is used.
__device__ __forceinline__ void load_at_acc(const double *__restrict__ C, long ld, double (&v)[16])
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
                v[(i * 2 + j) * 4 + e] = C[static_cast<long>(r0 + i * 16 + lk + 4 * e) * ld + c0 + j * 16 + lm];
}

// C (global 64x64 block, ld) = acc or C -= acc, optionally only on or below
// the block's diagonal.
__device__ __forceinline__ void store_acc(double *C, long ld, const v4d (&acc)[2][2], bool set, bool lower)
{
    for_each(acc,
             [&](int r, int c, double v)
             {
                 if (lower && c > r)
                     return;
                 double *dst = C + static_cast<long>(r) * ld + c;
                 if (set)
                     *dst = v;
                 else
                     *dst -= v;
             });
}

// C (global 64x64 block, ld) -= acc, optionally only on or below the
// block's diagonal. C is read and written through one pointer, so every
// load is issued before the first store (left to itself the compiler keeps
// each element's load, wait and store in order).
__device__ __forceinline__ void sub_acc(double *C, long ld, const v4d (&acc)[2][2], bool lower)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const int r0 = (w >> 1) * 32, c0 = (w & 1) * 32;
    double v[16];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 4; e++)
            {
                const int r = r0 + i * 16 + lk + 4 * e, c = c0 + j * 16 + lm;
                v[(i * 2 + j) * 4 + e] = (lower && c > r) ? 0.0 : C[static_cast<long>(r) * ld + c];
            }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 4; e++)
            {
                const int r = r0 + i * 16 + lk + 4 * e, c = c0 + j * 16 + lm;
                if (!(lower && c > r))
                    C[static_cast<long>(r) * ld + c] = v[(i * 2 + j) * 4 + e] - acc[i][j][e];
            }
}

}  // namespace gpamd::k::blk
