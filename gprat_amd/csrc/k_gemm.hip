// FP64 dense update on gfx950 matrix cores (v_mfma_f64_16x16x4_f64).
//
// One workgroup = 256 threads = 4 wavefronts computes a 128x128 tile of
// C (op)= A * B; each wavefront owns a 64x64 quarter as a 4x4 grid of 16x16
// MFMA accumulators (64 doubles per lane). K advances in chunks of 16
// through a double-buffered LDS stage stored k-major ([k][row]), so each MFMA
// operand is one ds_read_b64 per lane: lanes 0-15 read 16 consecutive rows
// of k-slot 0, lanes 16-31 of k-slot 1, and so on. The LDS row stride of 144
// doubles puts the four k-slots of a read in disjoint bank halves. The
// steady-state loop has no branch and no predicated load, so the
// accumulators stay in place and two workgroups share a CU.
//
// The same kernel serves the Cholesky trailing update (both operands the
// k-major panel copy, lower triangle only), the blocked TRSM's diagonal
// steps and updates, and W^T W (triangular K range).
#include "kernels.hpp"

namespace gpamd::k
{
namespace
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

// One workgroup per tile (grid x = column tiles, y = row tiles). The TRSM
// diagonal steps (K ends at each row tile's last row) give row tile t of n
// a K range of (t + 1) tiles: when the grid exceeds the device's workgroup
// slots (p.paired), a workgroup takes row tiles t and n - 1 - t, so every
// workgroup does the same work (grid halved along the rows); a grid that
// fits in one round keeps one tile per workgroup (shorter critical tile).
template <bool AT, bool BT, int ROLE>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_gemm_f64(Params p)
{
    __shared__ double As[2][BK * LDS_STRIDE];
    __shared__ double Bs[2][BK * LDS_STRIDE];
    long r0 = static_cast<long>(blockIdx.y) * BM;
    const long c0 = static_cast<long>(blockIdx.x) * BN;
    if (p.k_before_tile)
        r0 = static_cast<long>(gridDim.y - 1 - blockIdx.y) * BM;  // longest K ranges first
    if constexpr (ROLE == static_cast<int>(GemmRole::LeafT))
    {
        if (p.paired)
        {
            const int n = (p.M + BM - 1) / BM;
            const int a = blockIdx.y, b = n - 1 - a;
            if (!(p.lower && r0 + (BM - 1) + p.diag_off < c0))
                gemm_tile<AT, BT>(p, r0, c0, As, Bs);
            if (b <= a)
                return;
            r0 = static_cast<long>(b) * BM;
            if (p.lower && r0 + (BM - 1) + p.diag_off < c0)
                return;
            __syncthreads();  // the first tile's LDS reads are done
            gemm_tile<AT, BT>(p, r0, c0, As, Bs);
            return;
        }
    }
    if (p.lower && r0 + (BM - 1) + p.diag_off < c0)
        return;  // tile entirely above the diagonal
    gemm_tile<AT, BT>(p, r0, c0, As, Bs);
}

// Capped launch: G workgroups share the tiles that hold part of the update
// (for a lower update, only tiles touching the lower band), numbered row by
// row, so the tiles in flight together are neighbours. The launch leaves
// workgroup slots free for kernels on other streams. With a queue, each
// workgroup takes the next tile number from a device counter when it is
// ready, so one that starts late (its slot was still held by another
// stream's kernel) or shares its CU takes fewer tiles; without one,
// workgroup w takes tiles w, w + G, .... Two workgroups share a CU and their
// tiles take equal time, so left alone they would reach their epilogues
// (memory-bound, no matrix-core work) together. For C -= updates the tiles
// numbered G/2 .. G-1 (the first tiles of the upper half of the grid) are
// therefore run as two halves of their K range, the first at once and the
// second when the workgroup has no other tile left: those epilogues fall
// half a tile after the others'. Which tiles are split depends only on
// their numbers, so the rounding of every element does not depend on which
// workgroup ran it.
template <bool AT, bool BT, int ROLE>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_gemm_f64_capped(Params p)
{
    __shared__ double As[2][BK * LDS_STRIDE];
    __shared__ double Bs[2][BK * LDS_STRIDE];
    __shared__ int next;
    const int gx = (p.N + BN - 1) / BN;
    const int gy = (p.M + BM - 1) / BM;
    const int G = gridDim.x;
    long first_r = -1, first_c = 0;  // this workgroup's deferred second half
    // column tiles of tile row tr that touch the update
    auto row_tiles = [&](int tr)
    {
        if (!p.lower)
            return gx;
        const long last = static_cast<long>(tr) * BM + (BM - 1) + p.diag_off;  // last column the band reaches
        return last < 0 ? 0 : static_cast<int>(min(static_cast<long>(gx), last / BN + 1));
    };
    auto run = [&](int t, long r0, long c0)
    {
        if (p.set || t < G / 2 || t >= G)
            gemm_tile<AT, BT>(p, r0, c0, As, Bs, 0);
        else if (first_r < 0)
        {
            first_r = r0;
            first_c = c0;
            gemm_tile<AT, BT>(p, r0, c0, As, Bs, 1);
        }
        else
        {
            // a second split tile for this workgroup: both halves now
            gemm_tile<AT, BT>(p, r0, c0, As, Bs, 1);
            __syncthreads();
            gemm_tile<AT, BT>(p, r0, c0, As, Bs, 2);
        }
    };
    if (p.queue)
    {
        // tile numbers only grow, so the row walk is incremental
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
            run(t, static_cast<long>(tr) * BM, static_cast<long>(t - row_first) * BN);
        }
    }
    else
    {
        int rem = blockIdx.x;  // this workgroup's next tile, counted from the current row's first
        int row_first = 0;     // number of the current row's first tile
        for (int tr = 0; tr < gy; tr++)
        {
            const int nc = row_tiles(tr);
            for (; rem < nc; rem += G)
            {
                __syncthreads();  // the previous tile's last slab has been read by every wavefront
                run(row_first + rem, static_cast<long>(tr) * BM, static_cast<long>(rem) * BN);
            }
            rem -= nc;
            row_first += nc;
        }
    }
    if (first_r >= 0)
    {
        __syncthreads();
        gemm_tile<AT, BT>(p, first_r, first_c, As, Bs, 2);
    }
}

// Compute units of the current device (queried once per device).
int device_cus()
{
    static int cus[64] = {};
    int dev = 0;
    GP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64)
        return 256;
    if (cus[dev] == 0)
    {
        int v = 0;
        GP_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
        cus[dev] = v > 0 ? v : 256;
    }
    return cus[dev];
}

}  // namespace

void gemm_f64(hipStream_t s, const GemmArgs &g)
{
    if (g.M <= 0 || g.N <= 0)
        return;
    require(g.K >= 0 && g.K % BK == 0, "gemm_f64: K must be a multiple of 16");
    require(g.M % 2 == 0 && g.N % 2 == 0, "gemm_f64: M and N must be even");
    // per-lane slab offsets are 32-bit element offsets
    require(128L * std::max(g.lda, g.ldb) + 128 < (1L << 31), "gemm_f64: leading dimension too large");
    require(g.lda % 2 == 0 && g.ldb % 2 == 0, "gemm_f64: leading dimensions must be even");
    require(reinterpret_cast<uintptr_t>(g.A) % 16 == 0 && reinterpret_cast<uintptr_t>(g.B) % 16 == 0,
            "gemm_f64: operands must be 16-byte aligned");
    Params p{ g.A, g.lda, g.B, g.ldb, g.C, g.ldc, g.M, g.N, g.K, g.lower ? 1 : 0, g.diag_off, g.k_from_tile ? 1 : 0,
              g.kstart_off, g.set ? 1 : 0, g.k_to_row ? 1 : 0, g.k_before_tile ? 1 : 0, 0, nullptr, 0 };
    require(!g.k_before_tile || (g.max_wgs <= 0 && g.role == GemmRole::Residual), "gemm_f64: K before the row tile is for the residual product");
    dim3 grid(static_cast<unsigned>(ceil_div(g.N, BN)), static_cast<unsigned>(ceil_div(g.M, BM)));
    require(grid.y <= 65535, "gemm_f64: M too large for one launch");
    // tiles holding part of the update, counted as the capped kernel numbers them
    long tiles = 0;
    for (long tr = 0; tr < grid.y; tr++)
    {
        const long last = tr * BM + (BM - 1) + g.diag_off;
        tiles += g.lower ? (last < 0 ? 0 : std::min<long>(grid.x, last / BN + 1)) : grid.x;
    }
    if (g.max_wgs > 0 && g.max_wgs < tiles)
    {
        // only the trailing update runs capped
        require(g.a_trans && !g.b_trans && g.role == GemmRole::TrailingT, "gemm_f64: capped launch for this role");
        require(tiles < (1L << 30), "gemm_f64: too many tiles");
        p.queue = g.queue;
        p.tiles = static_cast<int>(tiles);
        hipLaunchKernelGGL((k_gemm_f64_capped<true, false, 5>), dim3(static_cast<unsigned>(g.max_wgs)), dim3(THREADS), 0, s, p);
        GP_CHECK_LAUNCH();
        return;
    }
    // paired row tiles (see k_gemm_f64) when the grid needs more than one
    // round of the device's workgroup slots (two per CU)
    if (g.role == GemmRole::LeafT && tiles > 2L * device_cus())
    {
        require(g.k_to_row, "gemm_f64: TRSM diagonal steps pair row tiles by their K range");
        grid.y = (grid.y + 1) / 2;
        p.paired = 1;
    }
    // Instantiated combinations: (T,N) with both operands k-major for the
    // updates, the TRSM steps and W^T W.
    if (g.a_trans && !g.b_trans && g.role == GemmRole::TrailingT)
        hipLaunchKernelGGL((k_gemm_f64<true, false, 5>), grid, dim3(THREADS), 0, s, p);
    else if (g.a_trans && !g.b_trans && g.role == GemmRole::SolveT)
        hipLaunchKernelGGL((k_gemm_f64<true, false, 6>), grid, dim3(THREADS), 0, s, p);
    else if (g.a_trans && !g.b_trans && g.role == GemmRole::LeafT)
        hipLaunchKernelGGL((k_gemm_f64<true, false, 7>), grid, dim3(THREADS), 0, s, p);
    else if (g.a_trans && !g.b_trans && g.role == GemmRole::Inverse)
        hipLaunchKernelGGL((k_gemm_f64<true, false, 3>), grid, dim3(THREADS), 0, s, p);
    else if (g.a_trans && !g.b_trans && g.role == GemmRole::Residual)
        hipLaunchKernelGGL((k_gemm_f64<true, false, 8>), grid, dim3(THREADS), 0, s, p);
    else
        throw Error(GPRAT_ERR_INTERNAL, "gemm_f64: operand layout and role not instantiated");
    GP_CHECK_LAUNCH();
}

}  // namespace gpamd::k
