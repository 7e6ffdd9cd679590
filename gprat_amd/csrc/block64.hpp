// Device routine shared by the panel kernel and the batched fit: Cholesky
// factor and inverse of one 64x64 block held in LDS by one 256-thread
// workgroup. Header-only so each kernel file inlines it.
#pragma once

#include <hip/hip_runtime.h>

namespace gpamd::k
{
namespace b64
{

typedef double v4d __attribute__((ext_vector_type(4)));

// Lane l's value of v on every lane (l is a compile-time constant once the
// callers' loops are unrolled): two v_readlane_b32 into scalar registers.
__device__ __forceinline__ double lane_value(double v, int l)
{
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffull), l));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), l));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// 1/sqrt(d): the hardware estimate refined by two Newton steps
// y <- y (3/2 - d/2 y^2) (each at least doubles the correct bits). A
// non-positive or NaN d gives NaN or infinity.
__device__ __forceinline__ double rsqrt_refined(double d)
{
    double y = __builtin_amdgcn_rsq(d);
    const double h = 0.5 * d;
    y = fma(y, fma(-h * y, y, 0.5), y);
    y = fma(y, fma(-h * y, y, 0.5), y);
    return y;
}

}  // namespace b64

// Cholesky factor and inverse of one 64x64 block held in LDS, by one
// workgroup of 256 threads, in 16-wide steps:
//  * factor, per step: wavefront 0 factors the 16x16 diagonal block in
//    registers (lane i owns row i; each pivot's column reaches the other
//    lanes through v_readlane, and the pivot's reciprocal square root is a
//    refined hardware estimate), one thread per row below solves against it
//    column by column, and the rank-16 update of the trailing block runs on
//    the matrix cores (one 16x16 output tile per wavefront at a time);
//  * inverse: the four 16x16 diagonal inverses are formed in parallel (one
//    wavefront each, one column per lane), then the off-diagonal blocks by
//    distance, W_ij = -W_ii * sum_{t=j}^{i-1} L_it W_tj, as two chained
//    matrix-core products whose intermediate never leaves registers.
// On exit D holds L in its lower triangle (the strict upper triangle of the
// 16x16 diagonal blocks holds scratch values), W = L^-1 (lower, zero above)
// and R the reciprocals of L's diagonal. A non-positive pivot stores the
// smallest failing global column in *info.
template <int NB = 64, int THREADS = 256>
__device__ __forceinline__ void factor_block64(double (*D)[NB + 1], double (*W)[NB + 1], double *R, int *info, int col0)
{
    using b64::v4d;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lm = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
    {
        const int o = kb * 16;
        if (wave == 0)
        {
            double r[16];
#pragma unroll
            for (int t = 0; t < 16; t++)
                r[t] = (lane < 16) ? D[o + lane][o + t] : 0.0;
            double piv = 1.0, rinv = 0.0;
#pragma unroll
            for (int j = 0; j < 16; j++)
            {
                const double djj = b64::lane_value(r[j], j);
                const double inv = b64::rsqrt_refined(djj);
                if (lane == j)
                {
                    piv = djj;
                    rinv = inv;
                }
                const double lij = (lane > j) ? r[j] * inv : (lane == j ? djj * inv : 0.0);
                r[j] = lij;
#pragma unroll
                for (int c = j + 1; c < 16; c++)
                    r[c] = fma(-lij, b64::lane_value(lij, c), r[c]);
            }
            if (lane < 16)
            {
                if (!(piv > 0.0))
                    atomicMin(info, col0 + o + lane);
                R[o + lane] = rinv;
#pragma unroll
                for (int t = 0; t < 16; t++)
                    if (t <= lane)
                        D[o + lane][o + t] = r[t];
            }
        }
        __syncthreads();
        const int below = NB - 16 - o;  // rows o+16 .. 63
        if (tid < below)
        {
            // x L_oo^T = a, column by column: x_j = x_j / l_jj, then
            // x_c -= x_j l_cj for the later columns
            const int rr = o + 16 + tid;
            double x[16];
#pragma unroll
            for (int t = 0; t < 16; t++)
                x[t] = D[rr][o + t];
#pragma unroll
            for (int j = 0; j < 16; j++)
            {
                x[j] *= R[o + j];
#pragma unroll
                for (int c = j + 1; c < 16; c++)
                    x[c] = fma(-x[j], D[o + c][o + j], x[c]);
            }
#pragma unroll
            for (int t = 0; t < 16; t++)
                D[rr][o + t] = x[t];
        }
        __syncthreads();
        if (below > 0)
        {
            // trailing block -= A A^T, A = D[o+16.., o..o+15], lower tiles only
            const int nt = below / 16, base = o + 16;
            for (int t = wave; t < nt * (nt + 1) / 2; t += THREADS / 64)
            {
                int tr = 0, tc = t;
                while (tc > tr)
                {
                    tc -= tr + 1;
                    tr++;
                }
                v4d acc = { 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
                for (int k0 = 0; k0 < 16; k0 += 4)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(D[base + tr * 16 + lm][o + k0 + lk], D[base + tc * 16 + lm][o + k0 + lk], acc,
                                                               0, 0, 0);
                // accumulator: lane holds column lm, rows lk + 4 e
#pragma unroll
                for (int e = 0; e < 4; e++)
                {
                    const int rw = tr * 16 + lk + 4 * e, cl = tc * 16 + lm;
                    if (cl <= rw)
                        D[base + rw][base + cl] -= acc[e];
                }
            }
        }
        __syncthreads();
    }

    // diagonal 16x16 inverses, one wavefront per block, one column per lane:
    // L w = e_c column by column (w_j /= l_jj, then w_i -= l_ij w_j)
    {
        const int o = wave * 16;
        if (lane < 16)
        {
            double w[16];
#pragma unroll
            for (int i = 0; i < 16; i++)
                w[i] = (i == lane) ? 1.0 : 0.0;
#pragma unroll
            for (int j = 0; j < 16; j++)
            {
                w[j] *= R[o + j];
#pragma unroll
                for (int i = j + 1; i < 16; i++)
                    w[i] = fma(-w[j], D[o + i][o + j], w[i]);
            }
#pragma unroll
            for (int i = 0; i < 16; i++)
                W[o + i][o + lane] = w[i];
        }
        // zero the strict upper blocks of W
        for (int e = tid; e < NB * NB; e += THREADS)
        {
            const int r = e / NB, c = e % NB;
            if ((c >> 4) > (r >> 4))
                W[r][c] = 0.0;
        }
    }
    __syncthreads();
#pragma unroll
    for (int d = 1; d < 4; d++)
    {
        // wavefront p forms block pair (i, j) = (p + d, p)
        if (wave < 4 - d)
        {
            const int bj = wave, bi = wave + d;
            v4d s = { 0.0, 0.0, 0.0, 0.0 };
            for (int bt = bj; bt < bi; bt++)
            {
#pragma unroll
                for (int k0 = 0; k0 < 16; k0 += 4)
                    s = __builtin_amdgcn_mfma_f64_16x16x4f64(D[bi * 16 + lm][bt * 16 + k0 + lk], W[bt * 16 + k0 + lk][bj * 16 + lm], s, 0, 0,
                                                             0);
            }
            // the B operand of k-step 4e (row 4e + lk, column lm) is exactly
            // this lane's accumulator entry e
            v4d acc = { 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
            for (int e = 0; e < 4; e++)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(W[bi * 16 + lm][bi * 16 + 4 * e + lk], s[e], acc, 0, 0, 0);
#pragma unroll
            for (int e = 0; e < 4; e++)
                W[bi * 16 + lk + 4 * e][bj * 16 + lm] = -acc[e];
        }
        __syncthreads();
    }
}

}  // namespace gpamd::k
