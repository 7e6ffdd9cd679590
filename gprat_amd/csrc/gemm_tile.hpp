// The 128x128 FP64 matrix-core tile of k_gemm.hip as device code shared with
// other kernels (the batched fit's level-synchronous path): one 256-thread
// workgroup computes C(r0.., c0..) (op)= A B over a K range in 16-deep slabs
// through a double-buffered k-major LDS stage (see k_gemm.hip).
#pragma once

#include "kernels.hpp"

namespace gpamd::k::gt
{

typedef double v4d __attribute__((ext_vector_type(4)));
// staging registers of the global->LDS copy: a native vector type, so they
// stay in registers (the HIP double2 struct ends up in scratch here)
typedef double d2 __attribute__((ext_vector_type(2)));

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int BK = 16;
constexpr int LDS_STRIDE = BM + 16;
constexpr int THREADS = 256;

struct Params
{
    const double *A;
    long lda;
    const double *B;
    long ldb;
    double *C;
    long ldc;
    int M, N, K;
    int lower;
    long diag_off;
    int k_from_tile;
    long kstart_off;
    int set;
    int k_to_row;
    int k_before_tile;  // K ends at the row tile's first row; row tiles launched last first
    int paired;  // k_gemm_f64 of the TRSM diagonal steps: two row tiles per workgroup
    int *queue;  // k_gemm_f64_capped: tile counter (0 at launch) or null for fixed shares
    int tiles;   // k_gemm_f64_capped: tiles holding part of the update
};

// Per-thread geometry of one BK-deep slab of a 128-wide operand, fixed for
// the whole K loop. `trans == false`: element (row, k) is P[row*ld + k]
// (rows contiguous in k); `trans == true`: element (row, k) is P[k*ld + row].
// Rows past the operand edge are clamped to its last row: they only feed
// accumulator rows/columns that the epilogue never stores, so no load in the
// loop is predicated.
template <bool TRANS>
struct Slab
{
    unsigned off[4];  // element offsets from the slab origin (fit in 32 bits)
    int lds[4];       // LDS element offsets of the first value

    __device__ __forceinline__ void init(long ld, int rows, int tid)
    {
#pragma unroll
        for (int s = 0; s < 4; s++)
        {
            const int idx = tid + THREADS * s;
            if (!TRANS)
            {
                const int row = min(idx >> 3, rows - 1);
                const int kk = (idx & 7) * 2;
                off[s] = static_cast<unsigned>(row * ld + kk);
                lds[s] = kk * LDS_STRIDE + (idx >> 3);
            }
            else
            {
                const int kk = idx >> 6;
                const int col = min((idx & 63) * 2, rows - 2);
                off[s] = static_cast<unsigned>(kk * ld + col);
                lds[s] = kk * LDS_STRIDE + (idx & 63) * 2;
            }
        }
    }

    __device__ __forceinline__ void load(const double *__restrict__ origin, d2 (&reg)[4]) const
    {
#pragma unroll
        for (int s = 0; s < 4; s++)
            reg[s] = *reinterpret_cast<const d2 *>(origin + off[s]);
    }

    __device__ __forceinline__ void store(double *S, const d2 (&reg)[4]) const
    {
#pragma unroll
        for (int s = 0; s < 4; s++)
        {
            if (!TRANS)
            {
                S[lds[s]] = reg[s].x;
                S[lds[s] + LDS_STRIDE] = reg[s].y;
            }
            else
            {
                *reinterpret_cast<d2 *>(S + lds[s]) = reg[s];
            }
        }
    }
};

// 16 MFMAs per 4-deep k step, four k steps per slab.
__device__ __forceinline__ void mfma_slab(const double *As, const double *Bs, int wr, int wc, int lm, int lk, v4d (&acc)[4][4])
{
#pragma unroll
    for (int kq = 0; kq < BK / 4; kq++)
    {
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; i++)
            a[i] = As[(kq * 4 + lk) * LDS_STRIDE + wr * 64 + i * 16 + lm];
#pragma unroll
        for (int j = 0; j < 4; j++)
            b[j] = Bs[(kq * 4 + lk) * LDS_STRIDE + wc * 64 + j * 16 + lm];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
}

// One 128x128 tile of C at (r0, c0).
// C (op)= acc over one wavefront's 64x64 quarter, element (i, j, e) at
// cw + (16 i + 4 e) ldc + 16 j. C is read and written through the same
// pointer, so the compiler keeps every element's load, wait and store in
// program order; the loads of a 16-row strip are therefore all issued before
// its first store, and the strip pays one memory latency instead of sixteen.
// FULL: an interior tile (no edge, wholly inside the lower band), no
// predicates.
template <bool FULL>
__device__ __forceinline__ void store_quarter(const Params &p, const v4d (&acc)[4][4], double *cw, int rb, int cb, int rows_a, int rows_b,
                                              long r0, long c0)
{
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        bool ok[16];
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int e = 0; e < 4; e++)
            {
                const int rr = rb + i * 16 + 4 * e, cc = cb + j * 16;
                ok[j * 4 + e] = FULL || ((rr < rows_a) & (cc < rows_b) & !(p.lower && r0 + rr + p.diag_off < c0 + cc));
            }
        double cv[16];
        if (!p.set)
        {
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int e = 0; e < 4; e++)
                    cv[j * 4 + e] = ok[j * 4 + e] ? cw[static_cast<long>(i * 16 + 4 * e) * p.ldc + j * 16] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (ok[j * 4 + e])
                    cw[static_cast<long>(i * 16 + 4 * e) * p.ldc + j * 16] = p.set ? acc[i][j][e] : cv[j * 4 + e] - acc[i][j][e];
    }
}

// part 0: the tile's whole K range; parts 1 and 2: its first and second
// half (in slabs), for C -= updates only.
template <bool AT, bool BT>
__device__ __forceinline__ void gemm_tile(const Params &p, long r0, long c0, double (*As)[BK * LDS_STRIDE], double (*Bs)[BK * LDS_STRIDE],
                                          int part = 0)
{
    long kbeg = 0;
    if (p.k_from_tile)
    {
        kbeg = (r0 > c0 ? r0 : c0) + p.kstart_off;
        kbeg = kbeg < 0 ? 0 : kbeg;
        kbeg = (kbeg / BK) * BK;
    }
    long kend = p.K;
    if (p.k_to_row)
        kend = min(kend, r0 + BM);
    if (p.k_before_tile)
        kend = min(kend, r0);
    if (part != 0)
    {
        const long mid = kbeg < kend ? kbeg + ((kend - kbeg) / BK / 2) * BK : kend;
        if (part == 1)
            kend = mid;
        else
            kbeg = mid;
    }
    const int rows_a = static_cast<int>(min(static_cast<long>(BM), p.M - r0));
    const int rows_b = static_cast<int>(min(static_cast<long>(BN), p.N - c0));

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int lm = lane & 15, lk = lane >> 4;

    v4d acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
            acc[i][j] = v4d{ 0.0, 0.0, 0.0, 0.0 };

    if (kbeg < kend)
    {
        Slab<AT> sa;
        Slab<!BT> sb;
        sa.init(p.lda, rows_a, tid);
        sb.init(p.ldb, rows_b, tid);
        // slab origins: A(r0, k0) and B(k0, c0) in the operands' own layouts
        const double *a0 = AT ? p.A + r0 : p.A + r0 * p.lda;
        const double *b0 = BT ? p.B + c0 * p.ldb : p.B + c0;
        const long a_step = AT ? BK * p.lda : BK;
        const long b_step = BT ? BK : BK * p.ldb;
        const double *ak = a0 + (AT ? kbeg * p.lda : kbeg);
        const double *bk = b0 + (BT ? kbeg : kbeg * p.ldb);

        d2 ra[4], rb[4];
        sa.load(ak, ra);
        sb.load(bk, rb);
        sa.store(As[0], ra);
        sb.store(Bs[0], rb);
        __syncthreads();

        int buf = 0;
        const long steps = (kend - kbeg) / BK;
        for (long t = 1; t < steps; t++)
        {
            ak += a_step;
            bk += b_step;
            sa.load(ak, ra);
            sb.load(bk, rb);
            // keep the next slab's loads ahead of this slab's products: left
            // to itself the scheduler sinks them to the LDS stores and every
            // slab then waits out a full memory latency
            __builtin_amdgcn_sched_barrier(0);
            mfma_slab(As[buf], Bs[buf], wr, wc, lm, lk, acc);
            __builtin_amdgcn_sched_barrier(0);
            sa.store(As[buf ^ 1], ra);
            sb.store(Bs[buf ^ 1], rb);
            __syncthreads();
            buf ^= 1;
        }
        mfma_slab(As[buf], Bs[buf], wr, wc, lm, lk, acc);
    }

    // Accumulator layout of the 16x16x4 f64 MFMA (measured on gfx950): lane
    // holds column (lane % 16) and rows (lane / 16) + 4 e, e = 0..3.
    // C is read and written through the same pointer, so the compiler keeps
    // every element's load, wait and store in program order; the loads of a
    // 16-row strip are therefore all issued before its first store, and the
    // strip pays one memory latency instead of sixteen. Interior tiles (no
    // edge, wholly on or below the diagonal band) skip the predicates.
    const bool full = rows_a == BM && rows_b == BN && (!p.lower || r0 + p.diag_off >= c0 + BN - 1);
    const int rb = wr * 64 + lk, cb = wc * 64 + lm;
    double *cw = p.C + (r0 + rb) * p.ldc + c0 + cb;
    if (full)
        store_quarter<true>(p, acc, cw, rb, cb, rows_a, rows_b, r0, c0);
    else
        store_quarter<false>(p, acc, cw, rb, cb, rows_a, rows_b, r0, c0);
}

}  // namespace gpamd::k::gt
