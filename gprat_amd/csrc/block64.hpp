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
//  * factor, per step: wavefront 0 holds the step's rows o.. 63 one per lane
//    (16 columns each) and eliminates the 16 pivots in registers: each
//    pivot's column reaches the other lanes through v_readlane and its
//    reciprocal square root is a refined hardware estimate, so the lanes
//    below the diagonal block finish the panel solve in the same pass.
//    Wavefront 1 runs the same elimination on the diagonal rows with the
//    identity rows e_s in lanes 16..31: what those lanes end with is
//    e_s L^-T, column s of the diagonal block's inverse. The rank-16 update
//    of the trailing block then runs on the matrix cores (one 16x16 output
//    tile per wavefront at a time);
//  * inverse: the off-diagonal blocks by distance,
//    W_ij = -W_ii * sum_{t=j}^{i-1} L_it W_tj, as two chained matrix-core
//    products whose intermediate never leaves registers.
// On exit D holds L in its lower triangle (the strict upper triangle of the
// 16x16 diagonal blocks holds scratch values), W holds L^-1 on and below
// the diagonal (exact zeros above it inside the 16x16 diagonal blocks, the
// 16x16 blocks above those untouched) and R the reciprocals of L's
// diagonal. A non-positive pivot stores the
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
        const int rows = NB - o;  // rows o .. 63 of the step's columns
        double r[16];
        if (wave < 2)
        {
#pragma unroll
            for (int t = 0; t < 16; t++)
            {
                if (wave == 0)
                    r[t] = (lane < rows) ? D[o + lane][o + t] : 0.0;
                else
                    r[t] = (lane < 16) ? D[o + lane][o + t] : ((lane - 16 == t) ? 1.0 : 0.0);
            }
        }
        __syncthreads();  // wavefront 1 has read the diagonal rows before wavefront 0 overwrites them
        if (wave < 2)
        {
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
            if (wave == 0)
            {
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
                else if (lane < rows)
                {
#pragma unroll
                    for (int t = 0; t < 16; t++)
                        D[o + lane][o + t] = r[t];
                }
            }
            else if (lane >= 16 && lane < 32)
            {
                // zeros above the diagonal come out exact: e_s has none
                // before column s
#pragma unroll
                for (int t = 0; t < 16; t++)
                    W[o + t][o + lane - 16] = r[t];
            }
        }
        __syncthreads();
        const int below = NB - 16 - o;  // rows o+16 .. 63
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
            __syncthreads();
        }
    }

    // the d = 1 step needs the diagonal inverses written by wavefront 1 in
    // the last step (the barrier after the kb = 3 pivots covers it); the
    // blocks above the diagonal of W are never read or written here
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
