// FP32 matrix-core products for the mixed-precision path: the trailing
// update of the factorisation and the variance solve.
//
// The factored FP64 panel is first copied to FP32 in k-major order
// (P[k][row], `panel_to_f32t`), so both operands of C -= A B^T are read as
// contiguous 16-byte row segments and land in LDS without a transpose.
// One workgroup (4 wavefronts) owns a 128x128 tile of the FP64 trailing
// matrix; each wavefront accumulates a 64x64 quarter in FP32 on
// v_mfma_f32_16x16x4f32 over 16-deep LDS slabs (double buffered), and the
// epilogue subtracts the FP32 product from the FP64 matrix. The shallow
// slabs (36 KiB of LDS per workgroup) and a 168-register budget let a CU
// hold three workgroups; on MI355X that ran the variance solve's updates at
// 124 TF/s against 118 with 32-deep slabs and two workgroups.
#include "kernels.hpp"

namespace gpamd::k
{
namespace
{

typedef float v4f __attribute__((ext_vector_type(4)));
// staging registers of the global->LDS copy: a native vector type, so they
// stay in registers (the HIP float4 struct ends up in scratch here)
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128;
constexpr int BK = 16;
constexpr int LPT = BK * 128 / (256 * 4);  // 16-byte loads per thread per operand slab
constexpr int KS = BM + 16;  // floats per LDS k-row: the 4 k-slots of a read hit disjoint banks
constexpr int THREADS = 256;

// C (op)= A B with A(i, k) = A[k * lda + i] and B(k, j) = B[k * ldb + j]
// (both operands k-major, as the panel copies are), FP32 products and sums.
template <class TC>
struct Params32
{
    const float *A;
    long lda;
    const float *B;
    long ldb;
    TC *C;
    long ldc;
    int M, N, K;
    int lower;  // only i + diag_off >= j
    long diag_off;
    int set;  // C = acc instead of C -= acc
    int max_wgs;  // > 0: at most this many workgroups, looping over tiles
    int k_to_row;  // A(i, k) known zero for k > i: each row tile stops K at its last row
    int *queue;    // capped: tile counter (0 at launch) or null for fixed shares
    int tiles;     // capped: tiles holding part of the update
};

// Element offsets (within one slab) of this thread's LPT 16-byte loads,
// clamped to the operand's last group of four; rows past the edge only feed
// accumulators the epilogue never stores.
__device__ __forceinline__ void slab_offsets(long ld, int rows, int tid, unsigned (&off)[LPT], int (&lds)[LPT])
{
#pragma unroll
    for (int s = 0; s < LPT; s++)
    {
        const int idx = tid + THREADS * s;
        const int kk = idx >> 5;         // 0..BK-1
        const int col = (idx & 31) * 4;  // 0..124
        off[s] = static_cast<unsigned>(kk * ld + min(col, rows - 4));
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

// One 128x128 tile of C at (r0, c0).
// C (op)= acc over one wavefront's 64x64 quarter, element (i, j, e) at
// cw + (16 i + e) ldc + 16 j. C is read and written through the same
// pointer, so the compiler keeps every element's load, wait and store in
// program order; the loads of a 16-row strip are therefore all issued before
// its first store (one memory latency per strip, not sixteen). FULL: an
// interior tile (no edge, wholly inside the lower band), no predicates.
template <bool FULL, class TC>
__device__ __forceinline__ void store_quarter32(const Params32<TC> &p, const v4f (&acc)[4][4], TC *cw, int rb, int cb, int rows_a, int rows_b,
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
                const int rr = rb + i * 16 + e, cc = cb + j * 16;
                ok[j * 4 + e] = FULL || ((rr < rows_a) & (cc < rows_b) & !(p.lower && r0 + rr + p.diag_off < c0 + cc));
            }
        TC cv[16];
        if (!p.set)
        {
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int e = 0; e < 4; e++)
                    cv[j * 4 + e] = ok[j * 4 + e] ? cw[static_cast<long>(i * 16 + e) * p.ldc + j * 16] : TC(0);
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (ok[j * 4 + e])
                    cw[static_cast<long>(i * 16 + e) * p.ldc + j * 16]
                        = p.set ? static_cast<TC>(acc[i][j][e]) : cv[j * 4 + e] - static_cast<TC>(acc[i][j][e]);
    }
}

// part 0: the tile's whole K range; parts 1 and 2: its first and second
// half (in slabs), for C -= updates only.
template <class TC>
__device__ __forceinline__ void gemm_tile32(const Params32<TC> &p, long r0, long c0, float (*As)[BK * KS], float (*Bs)[BK * KS], int part = 0)
{
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

    unsigned oa[LPT], ob[LPT];
    int la[LPT], lb[LPT];
    slab_offsets(p.lda, rows_a, tid, oa, la);
    slab_offsets(p.ldb, rows_b, tid, ob, lb);
    const int kmax = p.k_to_row ? static_cast<int>(min(static_cast<long>(p.K), r0 + BM)) : p.K;
    const int half = (kmax / BK / 2) * BK;
    const int k0 = part == 2 ? half : 0, k1 = part == 1 ? half : kmax;
    const float *ak = p.A + r0 + static_cast<long>(k0) * p.lda;
    const float *bk = p.B + c0 + static_cast<long>(k0) * p.ldb;
    const long astep = BK * p.lda, bstep = BK * p.ldb;

    if (k1 > k0)
    {
    f4 ra[LPT], rb[LPT];
#pragma unroll
    for (int s = 0; s < LPT; s++)
    {
        ra[s] = *reinterpret_cast<const f4 *>(ak + oa[s]);
        rb[s] = *reinterpret_cast<const f4 *>(bk + ob[s]);
    }
#pragma unroll
    for (int s = 0; s < LPT; s++)
    {
        *reinterpret_cast<f4 *>(&As[0][la[s]]) = ra[s];
        *reinterpret_cast<f4 *>(&Bs[0][lb[s]]) = rb[s];
    }
    __syncthreads();

    int buf = 0;
    const int steps = (k1 - k0) / BK;
    for (int t = 1; t < steps; t++)
    {
        ak += astep;
        bk += bstep;
#pragma unroll
        for (int s = 0; s < LPT; s++)
        {
            ra[s] = *reinterpret_cast<const f4 *>(ak + oa[s]);
            rb[s] = *reinterpret_cast<const f4 *>(bk + ob[s]);
        }
        // the next slab's loads stay ahead of this slab's products
        __builtin_amdgcn_sched_barrier(0);
        mfma_slab32(As[buf], Bs[buf], wr, wc, lm, lk, acc);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int s = 0; s < LPT; s++)
        {
            *reinterpret_cast<f4 *>(&As[buf ^ 1][la[s]]) = ra[s];
            *reinterpret_cast<f4 *>(&Bs[buf ^ 1][lb[s]]) = rb[s];
        }
        __syncthreads();
        buf ^= 1;
    }
    mfma_slab32(As[buf], Bs[buf], wr, wc, lm, lk, acc);
    }

    // Accumulator layout of the 16x16x4 f32 MFMA on gfx950: lane holds
    // column (lane % 16) and rows 4 (lane / 16) + e, e = 0..3.
    const bool full = rows_a == BM && rows_b == BM && (!p.lower || r0 + p.diag_off >= c0 + BM - 1);
    const int er = wr * 64 + lk * 4, ec = wc * 64 + lm;
    TC *cw = p.C + (r0 + er) * p.ldc + c0 + ec;
    if (full)
        store_quarter32<true>(p, acc, cw, er, ec, rows_a, rows_b, r0, c0);
    else
        store_quarter32<false>(p, acc, cw, er, ec, rows_a, rows_b, r0, c0);
}

// ROLE only separates profiler names: 0 trailing update (FP64 C), 1 TRSM
// diagonal step, 2 TRSM update. Tiles are numbered row-major; a launch with
// fewer workgroups than tiles loops over them (tile w, w + G, ...).
template <class TC, int ROLE>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(3))) void k_gemm_f32_tn(Params32<TC> p)
{
    __shared__ float As[2][BK * KS];
    __shared__ float Bs[2][BK * KS];
    const int gx = (p.N + BM - 1) / BM;
    const int gy = (p.M + BM - 1) / BM;
    if (p.max_wgs <= 0)
    {
        // one workgroup per tile, row-major; tiles above the band do nothing
        const long r0 = static_cast<long>(blockIdx.x / gx) * BM, c0 = static_cast<long>(blockIdx.x % gx) * BM;
        if (p.lower && r0 + (BM - 1) + p.diag_off < c0)
            return;
        gemm_tile32<TC>(p, r0, c0, As, Bs, 0);
        return;
    }
    // Capped: the tiles holding part of the update (for a lower update, those
    // touching the band) are numbered row by row, neighbours in flight
    // together. With a queue, each workgroup takes the next tile number from
    // a device counter when it is ready (one that starts late or shares its
    // CU takes fewer); without one, workgroup w takes w, w + G, .... For C -=
    // updates the tiles numbered G/2 .. G-1 run as two K halves, the second
    // when the workgroup has nothing else left, so the two workgroups of a
    // CU do not reach their epilogues together; the choice depends only on
    // tile numbers, so rounding does not depend on scheduling.
    __shared__ int next;
    const int G = gridDim.x;
    long first_r = -1, first_c = 0;  // this workgroup's deferred second half
    auto row_tiles = [&](int tr)
    {
        if (!p.lower)
            return gx;
        const long last = static_cast<long>(tr) * BM + (BM - 1) + p.diag_off;
        return last < 0 ? 0 : static_cast<int>(min(static_cast<long>(gx), last / BM + 1));
    };
    auto run = [&](int t, long r0, long c0)
    {
        if (p.set || t < G / 2 || t >= G)
            gemm_tile32<TC>(p, r0, c0, As, Bs, 0);
        else if (first_r < 0)
        {
            first_r = r0;
            first_c = c0;
            gemm_tile32<TC>(p, r0, c0, As, Bs, 1);
        }
        else
        {
            gemm_tile32<TC>(p, r0, c0, As, Bs, 1);
            __syncthreads();
            gemm_tile32<TC>(p, r0, c0, As, Bs, 2);
        }
    };
    if (p.queue)
    {
        int tr = 0, row_first = 0, nc = gy > 0 ? row_tiles(0) : 0;
        for (;;)
        {
            __syncthreads();  // the previous tile's LDS reads and `next` reads are done
            if (threadIdx.x == 0)
                next = atomicAdd(p.queue, 1);
            __syncthreads();
            const int t = next;
            if (t >= p.tiles)
                break;
            while (t >= row_first + nc && tr + 1 < gy)
            {
                row_first += nc;
                nc = row_tiles(++tr);
            }
            run(t, static_cast<long>(tr) * BM, static_cast<long>(t - row_first) * BM);
        }
    }
    else
    {
        int rem = blockIdx.x;  // this workgroup's next tile, counted from the current row's first
        int row_first = 0;
        for (int tr = 0; tr < gy; tr++)
        {
            const int nc = row_tiles(tr);
            for (; rem < nc; rem += G)
            {
                __syncthreads();  // the previous tile's last slab has been read by every wavefront
                run(row_first + rem, static_cast<long>(tr) * BM, static_cast<long>(rem) * BM);
            }
            rem -= nc;
            row_first += nc;
        }
    }
    if (first_r >= 0)
    {
        __syncthreads();
        gemm_tile32<TC>(p, first_r, first_c, As, Bs, 2);
    }
}

// P[k][r] = (T) A[(r0 + r) * lda + p0 + k] for r < R, k < W, through a
// 64x64 LDS tile so both the FP64 reads and the k-major writes are coalesced;
// with `back`, the same FP64 values are also copied to back[r * ldback + k].
template <class T>
__global__ __launch_bounds__(THREADS) void k_panel_to_t(const double *__restrict__ A, long lda, long r0, long p0, int R, int W,
                                                         T *__restrict__ P, long ldp, double *__restrict__ back, long ldback)
{
    __shared__ T t[64][65];
    const long rb = static_cast<long>(blockIdx.x) * 64, kb = static_cast<long>(blockIdx.y) * 64;
    const int c = threadIdx.x & 63, q = threadIdx.x >> 6;
    for (int rr = q; rr < 64; rr += 4)
    {
        const long r = rb + rr, k = kb + c;
        const double v = (r < R && k < W) ? A[(r0 + r) * lda + p0 + k] : 0.0;
        t[rr][c] = static_cast<T>(v);
        if (back && r < R && k < W)
            back[r * ldback + k] = v;
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

template <class T>
static void panel_to_t(hipStream_t s, const double *A, long lda, long r0, long p0, int R, int W, T *P, long ldp, double *back, long ldback)
{
    require(R >= 0 && W >= 0 && ldp >= R, "panel copy: bad shape");
    if (R == 0 || W == 0)
        return;
    dim3 grid(static_cast<unsigned>(ceil_div(R, 64)), static_cast<unsigned>(ceil_div(W, 64)));
    require(grid.y <= 65535, "panel copy: panel too wide");
    require(!back || ldback >= W, "panel copy: bad copy-back layout");
    hipLaunchKernelGGL(k_panel_to_t<T>, grid, dim3(THREADS), 0, s, A, lda, r0, p0, R, W, P, ldp, back, ldback);
    GP_CHECK_LAUNCH();
}

void panel_to_f32t(hipStream_t s, const double *A, long lda, long r0, long p0, int R, int W, float *P, long ldp, double *back, long ldback)
{
    panel_to_t<float>(s, A, lda, r0, p0, R, W, P, ldp, back, ldback);
}

void panel_to_f64t(hipStream_t s, const double *A, long lda, long r0, long p0, int R, int W, double *P, long ldp, double *back, long ldback)
{
    panel_to_t<double>(s, A, lda, r0, p0, R, W, P, ldp, back, ldback);
}

template <class TC, int ROLE>
static void launch_f32_tn(hipStream_t s, const Params32<TC> &p)
{
    if (p.M <= 0 || p.N <= 0)
        return;
    require(p.K % BK == 0, "FP32 product: K must be a multiple of 16");
    require(p.M % 4 == 0 && p.N % 4 == 0 && p.lda % 4 == 0 && p.ldb % 4 == 0
                && reinterpret_cast<uintptr_t>(p.A) % 16 == 0 && reinterpret_cast<uintptr_t>(p.B) % 16 == 0,
            "FP32 product: operand rows must come in aligned groups of four");
    require(static_cast<long>(BK) * std::max(p.lda, p.ldb) + BM < (1L << 31), "FP32 product: operand too tall");
    const long gx = ceil_div(static_cast<long>(p.N), static_cast<long>(BM)), gy = ceil_div(static_cast<long>(p.M), static_cast<long>(BM));
    require(gx * gy < (1L << 31), "FP32 product: too many tiles for one launch");
    // tiles holding part of the update, counted as the capped kernel numbers them
    long active = 0;
    for (long tr = 0; tr < gy; tr++)
    {
        const long last = tr * BM + (BM - 1) + p.diag_off;
        active += p.lower ? (last < 0 ? 0 : std::min(gx, last / BM + 1)) : gx;
    }
    if (active == 0)
        return;
    Params32<TC> q = p;
    q.tiles = static_cast<int>(std::min(active, (1L << 30)));
    if (q.max_wgs >= active)
    {
        q.max_wgs = 0;  // the cap does not bind: one workgroup per tile
        q.queue = nullptr;
    }
    dim3 grid(static_cast<unsigned>(q.max_wgs > 0 ? q.max_wgs : gx * gy));
    hipLaunchKernelGGL((k_gemm_f32_tn<TC, ROLE>), grid, dim3(THREADS), 0, s, q);
    GP_CHECK_LAUNCH();
}

void gemm_f32_trailing(hipStream_t s, const float *P, long ldp, long a_off, long b_off, double *C, long ldc, int M, int N, int K, long diag_off,
                       int max_wgs, int *queue)
{
    Params32<double> p{ P + a_off, ldp, P + b_off, ldp, C, ldc, M, N, K, 1, diag_off, 0, max_wgs, 0, queue, 0 };
    launch_f32_tn<double, 0>(s, p);
}

void gemm_f32_tn(hipStream_t s, const float *A, long lda, const float *B, long ldb, float *C, long ldc, int M, int N, int K, bool set,
                 int max_wgs, bool a_lower)
{
    Params32<float> p{ A, lda, B, ldb, C, ldc, M, N, K, 0, 0, set ? 1 : 0, max_wgs, a_lower ? 1 : 0, nullptr, 0 };
    if (set)
        launch_f32_tn<float, 1>(s, p);
    else
        launch_f32_tn<float, 2>(s, p);
}

}  // namespace gpamd::k
