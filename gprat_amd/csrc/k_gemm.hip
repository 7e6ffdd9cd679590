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
#include "gemm_tile.hpp"
#include "kernels.hpp"

namespace gpamd::k
{
namespace
{

using namespace gt;

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
