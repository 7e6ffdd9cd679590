// Device routine shared by the panel kernel and the batched fit: Cholesky
// factor and inverse of one 64x64 block held in LDS by one 256-thread
// workgroup. Header-only so each kernel file inlines it.
#pragma once

#include <hip/hip_runtime.h>

namespace gpamd::k
{

// Cholesky factor and inverse of one 64x64 block held in LDS, by one
// workgroup of 256 threads, blocked in 16-wide steps so only ~20 workgroup
// barriers are needed:
//  * factor: per 16-column step, wavefront 0 factors the 16x16 diagonal
//    block in registers (lane i owns row i, columns broadcast by shuffles),
//    one thread per row below solves against it, and all threads apply the
//    rank-16 update to the rest of the block;
//  * inverse: the four 16x16 diagonal inverses are formed in parallel (one
//    wavefront each, one column per lane), then the off-diagonal blocks by
//    distance: W_ij = -W_ii * sum_{t=j}^{i-1} L_it W_tj.
// On exit D holds L in its lower triangle and W = L^-1 (lower, zero above),
// R the reciprocals of L's diagonal (divisions past the pivots are products
// with them). A non-positive pivot stores the smallest failing global column
// in *info.
template <int NB = 64, int THREADS = 256>
__device__ __forceinline__ void factor_block64(double (*D)[NB + 1], double (*W)[NB + 1], double (*S)[16][17], double *R, int *info, int col0)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
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
#pragma unroll
            for (int j = 0; j < 16; j++)
            {
                const double djj = __shfl(r[j], j);
                if (lane == 0 && !(djj > 0.0))
                    atomicMin(info, col0 + o + j);
                const double ljj = sqrt(djj);
                const double inv = 1.0 / ljj;
                if (lane == 0)
                    R[o + j] = inv;
                const double lij = (lane > j) ? r[j] * inv : (lane == j ? ljj : 0.0);
                r[j] = lij;
#pragma unroll
                for (int c = j + 1; c < 16; c++)
                    r[c] -= lij * __shfl(lij, c);
            }
            if (lane < 16)
            {
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
            const int rr = o + 16 + tid;
            double x[16];
#pragma unroll
            for (int t = 0; t < 16; t++)
                x[t] = D[rr][o + t];
#pragma unroll
            for (int j = 0; j < 16; j++)
            {
                double sum = x[j];
#pragma unroll
                for (int q = 0; q < j; q++)
                    sum -= x[q] * D[o + j][o + q];
                x[j] = sum * R[o + j];
            }
#pragma unroll
            for (int t = 0; t < 16; t++)
                D[rr][o + t] = x[t];
        }
        __syncthreads();
        if (below > 0)
        {
            for (int e = tid; e < below * below; e += THREADS)
            {
                const int rr = e / below, cc = e % below;  // below is a compile-time constant per unrolled kb
                if (cc <= rr)
                {
                    double sum = 0.0;
#pragma unroll
                    for (int t = 0; t < 16; t++)
                        sum += D[o + 16 + rr][o + t] * D[o + 16 + cc][o + t];
                    D[o + 16 + rr][o + 16 + cc] -= sum;
                }
            }
        }
        __syncthreads();
    }

    // diagonal 16x16 inverses, one wavefront per block, one column per lane
    {
        const int o = wave * 16;
        if (lane < 16)
        {
            double wc[16];
#pragma unroll
            for (int i = 0; i < 16; i++)
            {
                double sum = (i == lane) ? 1.0 : 0.0;
#pragma unroll
                for (int t = 0; t < i; t++)
                    sum -= D[o + i][o + t] * wc[t];
                wc[i] = sum * R[o + i];
            }
#pragma unroll
            for (int i = 0; i < 16; i++)
                W[o + i][o + lane] = wc[i];
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
        const int pairs = 4 - d;
        // S_p = sum_{t=j}^{i-1} L_it W_tj for block pair p: i = p + d, j = p
        for (int e = tid; e < pairs * 256; e += THREADS)
        {
            const int pr = e >> 8, rr = (e >> 4) & 15, cc = e & 15;
            const int bi = pr + d, bj = pr;
            double sum = 0.0;
            for (int bt = bj; bt < bi; bt++)
            {
#pragma unroll
                for (int t = 0; t < 16; t++)
                    sum += D[bi * 16 + rr][bt * 16 + t] * W[bt * 16 + t][bj * 16 + cc];
            }
            S[pr][rr][cc] = sum;
        }
        __syncthreads();
        for (int e = tid; e < pairs * 256; e += THREADS)
        {
            const int pr = e >> 8, rr = (e >> 4) & 15, cc = e & 15;
            const int bi = pr + d, bj = pr;
            double sum = 0.0;
#pragma unroll
            for (int t = 0; t < 16; t++)
                sum += W[bi * 16 + rr][bi * 16 + t] * S[pr][t][cc];
            W[bi * 16 + rr][bj * 16 + cc] = -sum;
        }
        __syncthreads();
    }
}


}  // namespace gpamd::k
