// Triangular kernels on gfx950: the 64-wide Cholesky panel step, the
// blocked vector solves, the multi-right-hand-side TRSM base case and the
// loss reduction. All reductions run in a fixed order, so results do not
// depend on scheduling.
#include "block64.hpp"
#include "block_mfma.hpp"
#include "kernels.hpp"
#include "panel_plan.hpp"

namespace gpamd::k
{
namespace
{

constexpr int NB = 64;        // diagonal block edge
constexpr int THREADS = 256;

// Lower triangle of the NB x NB block at (r0, r0) into LDS (strict upper
// zeroed); every load of a thread is issued before the first LDS store.
__device__ __forceinline__ void load_lower_block(const double *__restrict__ A, long lda, long r0, double (*D)[NB + 1])
{
    double v[NB * NB / THREADS];
#pragma unroll
    for (int s = 0; s < NB * NB / THREADS; s++)
    {
        const int e = threadIdx.x + THREADS * s, r = e / NB, c = e % NB;
        v[s] = (c <= r) ? A[(r0 + r) * lda + r0 + c] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < NB * NB / THREADS; s++)
    {
        const int e = threadIdx.x + THREADS * s;
        D[e / NB][e % NB] = v[s];
    }
}

// Task interpreter of the diagonal-block factorisation (panel_plan.hpp):
// workgroup g of a launch runs the tasks [wg_off[g], wg_off[g + 1]) of
// the host-built schedule, each on 64x64 blocks of the diagonal block at
// (p0, p0). Block (i, j) of A is at A + (p0 + 64 i) lda + p0 + 64 j;
// Winv_k (row-major) at Winv + (p0 / 64 + k) 4096; the k-major inverse has
// W(64 i + r, 64 j + c) at Wb[(64 j + c) ldw + 64 i + r]. Every task reads
// only blocks that earlier launches finished and writes blocks no other
// task of its launch touches, so the tasks of a launch need no ordering.
struct PanelArgs
{
    PanelTaskArgs h;  // A, p0, Winv, Wb, tmp, info, head operands
    const int4 *tasks;
    const int *wg_off;
};

__device__ __forceinline__ double *blockA(const PanelArgs &a, int i, int j)
{
    return a.h.A + (a.h.p0 + static_cast<long>(NB) * i) * a.h.lda + a.h.p0 + NB * j;
}
__device__ __forceinline__ double *blockWinv(const PanelArgs &a, int k) { return a.h.Winv + (a.h.p0 / NB + k) * static_cast<long>(NB * NB); }
// k-major W block (i, j): element (r, c) at [c * ldw + r]
__device__ __forceinline__ double *blockW(const PanelArgs &a, int i, int j) { return a.h.Wb + static_cast<long>(NB) * j * a.h.ldw + NB * i; }

// FACTOR / UFACTOR: block (k, k) (minus L_kq L_kq^T when q >= 0) factored
// and inverted in LDS; L to A (lower triangle), the inverse to Winv_k
// (row-major, zero upper).
__device__ __noinline__ void task_factor(const PanelArgs &a, int k, int q, double *buf0, double *buf1, double *R)
{
    double(*D)[NB + 1] = reinterpret_cast<double(*)[NB + 1]>(buf0);
    double(*Wl)[NB + 1] = reinterpret_cast<double(*)[NB + 1]>(buf1);
    double *blk = blockA(a, k, k);
    load_lower_block(blk, a.h.lda, 0, D);
    if (q >= 0)
    {
        blk::stage_block(buf1, blockA(a, k, q), a.h.lda);
        __syncthreads();
        blk::v4d acc[2][2];
        blk::mfma_nt(buf1, buf1, acc);
        blk::for_each(acc,
                      [&](int r, int c, double v)
                      {
                          if (c <= r)
                              D[r][c] -= v;
                      });
    }
    __syncthreads();  // D final; every read of buf1 done before W lands there
    factor_block64(D, Wl, R, a.h.info, static_cast<int>(a.h.p0) + NB * k);
    double *w = blockWinv(a, k);
    for (int e = threadIdx.x; e < NB * NB; e += THREADS)
    {
        const int r = e / NB, c = e % NB;
        if (c <= r)
            blk[static_cast<long>(r) * a.h.lda + c] = D[r][c];
        w[e] = (c <= r) ? Wl[r][c] : 0.0;
    }
}

// L_ik = A_ik Winv_k^T in place.
__device__ __noinline__ void task_solve(const PanelArgs &a, int i, int k, double *buf0, double *buf1)
{
    double *p = blockA(a, i, k);
    blk::stage_block(buf0, p, a.h.lda);
    blk::stage_block(buf1, blockWinv(a, k), NB);
    __syncthreads();
    blk::v4d acc[2][2];
    blk::mfma_nt(buf0, buf1, acc);
    blk::store_acc(p, a.h.lda, acc, true, false);
}

// A_ij -= L_ik L_jk^T (lower half of a diagonal block).
__device__ __noinline__ void task_update(const PanelArgs &a, int i, int j, int k, double *buf0, double *buf1)
{
    blk::stage_block(buf0, blockA(a, i, k), a.h.lda);
    if (i != j)
        blk::stage_block(buf1, blockA(a, j, k), a.h.lda);
    __syncthreads();
    blk::v4d acc[2][2];
    blk::mfma_nt(buf0, i != j ? buf1 : buf0, acc);
    blk::sub_acc(blockA(a, i, j), a.h.lda, acc, i == j);
}

// W_kk = Winv_k (transposed through LDS: both sides coalesced), W_kj = 0
// right of it.
__device__ __noinline__ void task_place(const PanelArgs &a, int k, double *buf0)
{
    blk::stage_block(buf0, blockWinv(a, k), NB);
    __syncthreads();
    double *out = blockW(a, k, k);
    for (int e = threadIdx.x; e < NB * NB; e += THREADS)
    {
        const int c = e >> 6, r = e & 63;
        out[static_cast<long>(c) * a.h.ldw + r] = buf0[r * blk::XS + c];
    }
    for (int j = k + 1; j < a.h.nb; j++)
    {
        double *z = blockW(a, k, j);
        for (int e = threadIdx.x; e < NB * NB; e += THREADS)
            z[static_cast<long>(e >> 6) * a.h.ldw + (e & 63)] = 0.0;
    }
}

// W_ba = -Winv_b (L_ba Winv_a), b = a + 1; the intermediate stays on chip.
__device__ __noinline__ void task_joinf(const PanelArgs &a, int ka, double *buf0, double *buf1)
{
    const int kb = ka + 1;
    blk::stage_block(buf0, blockA(a, kb, ka), a.h.lda);  // [r][k] = L_ba(r, k)
    blk::stage_block_t(buf1, blockWinv(a, ka), NB);    // [c][k] = W_a(k, c)
    __syncthreads();
    blk::v4d acc[2][2];
    blk::mfma_nt(buf0, buf1, acc);
    __syncthreads();
    blk::for_each(acc, [&](int r, int c, double v) { buf1[c * blk::XS + r] = v; });  // [c][k] = T(k, c)
    blk::stage_block(buf0, blockWinv(a, kb), NB);                                     // [r][k] = W_b(r, k)
    __syncthreads();
    blk::mfma_nt(buf0, buf1, acc);
    double *out = blockW(a, kb, ka);
    blk::for_each(acc, [&](int r, int c, double v) { out[static_cast<long>(c) * a.h.ldw + r] = -v; });
}

// Doubling join at level h, pair p: a = blocks [2hp, 2hp + h), b = the
// next h; the pair's T (64h x 64h) at tmp + w * 4096, T(r, c) at
// [c * 64h + r].
//   phase 0: T(rb, cb) = sum_{s >= cb} L_{b0+rb, a0+s} W_{a0+s, a0+cb} (W_a lower);
//   phase 1: W_{b0+rb, a0+cb} = -sum_{s <= rb} W_{b0+rb, b0+s} T(s, cb) (W_b lower).
__device__ __noinline__ void task_join(const PanelArgs &a, int phase, int h, int p, int rb, int cb, int w, double *buf0, double *buf1)
{
    const int a0 = 2 * h * p, b0 = a0 + h;
    const long size = static_cast<long>(NB) * h;
    double *tp = a.h.tmp + static_cast<long>(w) * NB * NB;
    blk::v4d acc[2][2];
    blk::zero(acc);
    if (phase == 0)
    {
        for (int s = cb; s < h; s++)
        {
            __syncthreads();
            blk::stage_block(buf0, blockA(a, b0 + rb, a0 + s), a.h.lda);  // [r][k] = L(r, k)
            blk::stage_block(buf1, blockW(a, a0 + s, a0 + cb), a.h.ldw);  // [c][k] = W(k, c)
            __syncthreads();
            blk::mfma_nt_acc(buf0, buf1, acc);
        }
        blk::for_each(acc, [&](int r, int c, double v) { tp[(NB * cb + c) * size + NB * rb + r] = v; });
    }
    else
    {
        for (int s = 0; s <= rb; s++)
        {
            __syncthreads();
            blk::stage_block_t(buf0, blockW(a, b0 + rb, b0 + s), a.h.ldw);  // [r][k] = W(r, k)
            blk::stage_block(buf1, tp + (NB * cb) * size + NB * s, size);  // [c][k] = T(k, c)
            __syncthreads();
            blk::mfma_nt_acc(buf0, buf1, acc);
        }
        double *out = blockW(a, b0 + rb, a0 + cb);
        blk::for_each(acc, [&](int r, int c, double v) { out[static_cast<long>(c) * a.h.ldw + r] = -v; });
    }
}

// Head of the diagonal block at p0 in the previous panel (columns hp0 ..,
// nbh blocks): X_ic = sum_{t <= c} A(p0 + 64 i, hp0 + 64 t) W_P(c, t)^T,
// W_P the panel's k-major inverse (lower), into Xh (row-major, ld ldx).
__device__ __noinline__ void task_headx(const PanelArgs &a, int i, int c, double *buf0, double *buf1)
{
    const PanelTaskArgs &h = a.h;
    const double *row = h.A + (h.p0 + static_cast<long>(NB) * i) * h.lda + h.hp0;
    blk::v4d acc[2][2];
    blk::zero(acc);
    for (int t = 0; t <= c; t++)
    {
        __syncthreads();
        blk::stage_block(buf0, row + NB * t, h.lda);                                           // [r][k] = A(r, k)
        blk::stage_block_t(buf1, h.Wh + static_cast<long>(NB) * t * h.ldwh + NB * c, h.ldwh);  // [cc][k] = W_P(c, t)(cc, k)
        __syncthreads();
        blk::mfma_nt_acc(buf0, buf1, acc);
    }
    double *out = h.Xh + static_cast<long>(NB) * i * h.ldx + NB * c;
    blk::store_acc(out, h.ldx, acc, true, false);
}

// A(p0 + 64 i, p0 + 64 j) -= sum_{t < nbh} X_it X_jt^T (lower half if i == j).
__device__ __noinline__ void task_heads(const PanelArgs &a, int i, int j, double *buf0, double *buf1)
{
    const PanelTaskArgs &h = a.h;
    blk::v4d acc[2][2];
    blk::zero(acc);
    for (int t = 0; t < h.nbh; t++)
    {
        __syncthreads();
        blk::stage_block(buf0, h.Xh + static_cast<long>(NB) * i * h.ldx + NB * t, h.ldx);
        if (i != j)
            blk::stage_block(buf1, h.Xh + static_cast<long>(NB) * j * h.ldx + NB * t, h.ldx);
        __syncthreads();
        blk::mfma_nt_acc(buf0, i != j ? buf1 : buf0, acc);
    }
    blk::sub_acc(blockA(a, i, j), h.lda, acc, i == j);
}

__global__ __launch_bounds__(THREADS) void k_panel_tasks(PanelArgs a)
{
    __shared__ double buf0[NB * blk::XS];
    __shared__ double buf1[NB * blk::XS];
    __shared__ double R[NB];
    const int first = a.wg_off[blockIdx.x], last = a.wg_off[blockIdx.x + 1];
    for (int t = first; t < last; t++)
    {
        const int4 d = a.tasks[t];
        const int type = d.x & 0xff;
        __syncthreads();  // the previous task's LDS reads are done
        if (type == PT_FACTOR || type == PT_UFACTOR)
            task_factor(a, d.y, type == PT_UFACTOR ? d.z : -1, buf0, buf1, R);
        else if (type == PT_SOLVE)
            task_solve(a, d.y, d.z, buf0, buf1);
        else if (type == PT_UPDATE)
            task_update(a, d.y, d.z, d.w, buf0, buf1);
        else if (type == PT_PLACE)
            task_place(a, d.y, buf0);
        else if (type == PT_JOINF)
            task_joinf(a, d.y, buf0, buf1);
        else if (type == PT_HEADX)
            task_headx(a, d.y, d.z, buf0, buf1);
        else if (type == PT_HEADS)
            task_heads(a, d.y, d.z, buf0, buf1);
        else
            task_join(a, type == PT_JOIN0 ? 0 : 1, d.x >> 8, d.y, d.z & 0xffff, d.z >> 16, d.w, buf0, buf1);
    }
}

// Panel rows below a factored diagonal block: X = A W^T for the R rows at
// r0 of the panel's columns [p0, p0 + 64 nbw), W the block's k-major
// inverse (lower: X(i, c) = sum_{t <= c} A(i, t) W(c, t)^T over 64x64
// blocks), into X (row-major, ld ldx). Column block c costs c + 1 block
// products, so workgroup (p, i) computes column blocks p and nbw - 1 - p of
// row block i: every workgroup does nbw + 1 products (one extra for the
// middle block of an odd nbw). Each product is staged as two 32-deep
// halves (35 KB of LDS; with the register budget held to four wavefronts
// per SIMD, a few values spill, four workgroups share a CU and hide each
// other's staging: on MI355X whole blocks ran two per CU, halves with the
// compiler's budget three, this four, the n = 32768 factorisation 204 ->
// 202.5 -> 200.7 ms; the MFMAs run in the whole product's order, so the
// result is the same), and the next half is loaded into registers while
// the current one is multiplied.
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_rows_solve(const double *__restrict__ A,
                                                         long lda,
                                                         long r0,
                                                         long p0,
                                                         int nbw,
                                                         const double *__restrict__ Wb,
                                                         long ldw,
                                                         double *__restrict__ X,
                                                         long ldx)
{
    __shared__ double a[NB * blk::XH];
    __shared__ double b[NB * blk::XH];
    const int pair = blockIdx.x, i = blockIdx.y;
    const double *row = A + (r0 + static_cast<long>(NB) * i) * lda + p0;
    for (int side = 0; side < 2; side++)
    {
        const int c = side == 0 ? pair : nbw - 1 - pair;
        if (side == 1 && c <= pair)
            break;
        blk::v4d acc[2][2];
        blk::zero(acc);
        // half-step s covers k = 32 s .. 32 s + 31 of the product's sum:
        // A's columns there, and W(c, .)'s k-major rows
        const double *wc = Wb + static_cast<long>(NB) * c;
        const int steps = 2 * (c + 1);
        blk::d2 va[4], vb[4];
        blk::load_half(row, lda, va);
        blk::load_half_t(wc, ldw, vb);
        for (int st = 0; st < steps; st++)
        {
            __syncthreads();  // the previous half's LDS reads are done
            blk::store_half(a, va);    // [r][k] = A(r, k)
            blk::store_half_t(b, vb);  // [cc][k] = W(c, t)(cc, k)
            __syncthreads();
            if (st + 1 < steps)
            {
                blk::load_half(row + 32 * (st + 1), lda, va);
                blk::load_half_t(wc + static_cast<long>(32) * (st + 1) * ldw, ldw, vb);
            }
            __builtin_amdgcn_sched_barrier(0);
            blk::mfma_nt_acc_half(a, b, acc);
        }
        blk::store_acc(X + static_cast<long>(NB) * i * ldx + NB * c, ldx, acc, true, false);
    }
}

// Diagonal step of the blocked TRSM: V = W B with W lower triangular
// (k-major: W(r, c) at Wk[c * ldw + r], nbr 64-blocks), B and V row-major.
// V(i, j) = sum_{t <= i} W(i, t) B(t, j) over 64x64 blocks: row block i
// costs i + 1 block products, so workgroup (j, p) computes row blocks p and
// nbr - 1 - p of column block j. Products on 32-deep half stages, the next
// half loaded while the current one is multiplied (as k_rows_solve).
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_leaf_solve(const double *__restrict__ Wk,
                                                         long ldw,
                                                         int nbr,
                                                         const double *__restrict__ B,
                                                         long ldb,
                                                         double *__restrict__ V,
                                                         long ldv)
{
    __shared__ double a[NB * blk::XH];
    __shared__ double b[NB * blk::XH];
    const int j = blockIdx.x, pair = blockIdx.y;
    for (int side = 0; side < 2; side++)
    {
        const int i = side == 0 ? pair : nbr - 1 - pair;
        if (side == 1 && i <= pair)
            break;
        blk::v4d acc[2][2];
        blk::zero(acc);
        // half-step s covers k = 32 s .. 32 s + 31: those k-major rows of
        // W(i, .) and rows of B(., j)
        const double *wi = Wk + static_cast<long>(NB) * i;
        const double *bj = B + static_cast<long>(NB) * j;
        const int steps = 2 * (i + 1);
        blk::d2 va[4], vb[4];
        blk::load_half_t(wi, ldw, va);
        blk::load_half_t(bj, ldb, vb);
        for (int st = 0; st < steps; st++)
        {
            __syncthreads();            // the previous half's LDS reads are done
            blk::store_half_t(a, va);   // [r][k] = W(i, t)(r, k)
            blk::store_half_t(b, vb);   // [cc][k] = B(t, j)(k, cc)
            __syncthreads();
            if (st + 1 < steps)
            {
                blk::load_half_t(wi + static_cast<long>(32) * (st + 1) * ldw, ldw, va);
                blk::load_half_t(bj + static_cast<long>(32) * (st + 1) * ldb, ldb, vb);
            }
            __builtin_amdgcn_sched_barrier(0);
            blk::mfma_nt_acc_half(a, b, acc);
        }
        blk::store_acc(V + static_cast<long>(NB) * i * ldv + NB * j, ldv, acc, true, false);
    }
}

// x_J = W_JJ rhs_J (or W_JJ^T rhs_J), four column quarters per row combined
// in a fixed order; every workgroup of a solve step computes it.
template <bool TRANSPOSED>
__device__ __forceinline__ void block_matvec(const double *__restrict__ W, const double *__restrict__ rhs, double *xs, double (*part)[NB])
{
    __shared__ double rs[NB];
    const int tid = threadIdx.x;
    if (tid < NB)
        rs[tid] = rhs[tid];
    __syncthreads();
    const int i = tid & (NB - 1), q = tid >> 6;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 16; k++)
    {
        const int kk = q * 16 + k;
        s += (TRANSPOSED ? W[kk * NB + i] : W[i * NB + kk]) * rs[kk];
    }
    part[q][i] = s;
    __syncthreads();
    if (tid < NB)
        xs[tid] = (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
    __syncthreads();
}

// Forward block step J of L x = rhs: x_J = W_JJ rhs_J (workgroup 0 stores it),
// workgroup w >= 1 updates rhs_{J+w} -= L_{J+w,J} x_J.
__global__ __launch_bounds__(THREADS) void k_solve_lower_step(const double *__restrict__ L,
                                                               long lda,
                                                               const double *__restrict__ Winv,
                                                               double *__restrict__ rhs,
                                                               double *__restrict__ out,
                                                               int J)
{
    __shared__ double xs[NB];
    __shared__ double part[4][NB];
    const int tid = threadIdx.x;
    const long base = static_cast<long>(J) * NB;
    block_matvec<false>(Winv + base * NB, rhs + base, xs, part);
    if (blockIdx.x == 0)
    {
        if (tid < NB)
            out[base + tid] = xs[tid];
        return;
    }
    const long I = static_cast<long>(J) + blockIdx.x;
    const int rr = tid >> 2, q = tid & 3;
    const double *prow = L + (I * NB + rr) * lda + base + q * 16;
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 16; c++)
        s += prow[c] * xs[q * 16 + c];
    part[q][rr] = s;
    __syncthreads();
    if (tid < NB)
        rhs[I * NB + tid] -= (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
}

// Backward block step J of L^T x = rhs: x_J = W_JJ^T rhs_J, then
// rhs_{J-w} -= L_{J,J-w}^T x_J.
__global__ __launch_bounds__(THREADS) void k_solve_upper_step(const double *__restrict__ L,
                                                               long lda,
                                                               const double *__restrict__ Winv,
                                                               double *__restrict__ rhs,
                                                               double *__restrict__ out,
                                                               int J)
{
    __shared__ double xs[NB];
    __shared__ double part[4][NB];
    const int tid = threadIdx.x;
    const long base = static_cast<long>(J) * NB;
    block_matvec<true>(Winv + base * NB, rhs + base, xs, part);
    if (blockIdx.x == 0)
    {
        if (tid < NB)
            out[base + tid] = xs[tid];
        return;
    }
    const long I = static_cast<long>(J) - blockIdx.x;
    const int c = tid & 63, q = tid >> 6;
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < 16; r++)
        s += L[(base + q * 16 + r) * lda + I * NB + c] * xs[q * 16 + r];
    part[q][c] = s;
    __syncthreads();
    if (tid < NB)
        rhs[I * NB + tid] -= (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
}

// out[c] = (or -=) sum_{r < K} M[r * ld + c] v[r] for c < N. Lane = column
// (each row read is a coalesced 512-byte segment), the sixteen wavefronts
// of a 1024-thread workgroup take contiguous sixteenths of the rows with 16
// loads in flight per lane (a 512-row block is then two batches deep), v
// comes in through scalar loads, and the partial sums are combined in a
// fixed order.
constexpr int COLDOT_WAVES = 16;
__global__ __launch_bounds__(64 * COLDOT_WAVES) void k_coldot(const double *__restrict__ M,
                                                               long ld,
                                                               int K,
                                                               int N,
                                                               const double *__restrict__ v,
                                                               double *__restrict__ out,
                                                               bool sub)
{
    __shared__ double part[COLDOT_WAVES][64];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = blockIdx.x * 64 + lane;
    const int q = (K + COLDOT_WAVES - 1) / COLDOT_WAVES;
    const int r0 = min(K, w * q), r1 = min(K, r0 + q);
    double s = 0.0;
    if (c < N)
    {
        int r = r0;
        for (; r + 16 <= r1; r += 16)
        {
            double m[16];
#pragma unroll
            for (int t = 0; t < 16; t++)
                m[t] = M[static_cast<long>(r + t) * ld + c];
#pragma unroll
            for (int t = 0; t < 16; t++)
                s = fma(m[t], v[r + t], s);
        }
        for (; r < r1; r++)
            s = fma(M[static_cast<long>(r) * ld + c], v[r], s);
    }
    part[w][lane] = s;
    __syncthreads();
    if (w == 0 && c < N)
    {
        double t = 0.0;
#pragma unroll
        for (int q2 = 0; q2 < COLDOT_WAVES; q2++)
            t += part[q2][lane];
        out[c] = sub ? out[c] - t : t;
    }
}

// out[i] = (or -=) sum_{c < K} M[i * ld + c] v[c] for i < N, K a multiple
// of 64. Each wavefront owns four consecutive rows: lane l reads columns
// l, l + 64, ... of each (coalesced 512-byte segments), then every row sum
// is reduced across the wavefront by a fixed butterfly.
__global__ __launch_bounds__(THREADS) void k_rowdot(const double *__restrict__ M,
                                                     long ld,
                                                     int K,
                                                     int N,
                                                     const double *__restrict__ v,
                                                     double *__restrict__ out,
                                                     bool sub)
{
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i0 = (blockIdx.x * 4 + w) * 4;
    if (i0 >= N)
        return;
    const int rows = min(4, N - i0);
    double s[4] = { 0.0, 0.0, 0.0, 0.0 };
    for (int c0 = 0; c0 < K; c0 += 256)
    {
        double m[4][4], x[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
        {
            const int c = c0 + u * 64 + lane;
            x[u] = c < K ? v[c] : 0.0;
#pragma unroll
            for (int rr = 0; rr < 4; rr++)
                m[rr][u] = (c < K && rr < rows) ? M[static_cast<long>(i0 + rr) * ld + c] : 0.0;
        }
#pragma unroll
        for (int rr = 0; rr < 4; rr++)
#pragma unroll
            for (int u = 0; u < 4; u++)
                s[rr] = fma(m[rr][u], x[u], s[rr]);
    }
#pragma unroll
    for (int rr = 0; rr < 4; rr++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            s[rr] += __shfl_xor(s[rr], off);
    if (lane < rows)
    {
        const double t = lane == 0 ? s[0] : lane == 1 ? s[1] : lane == 2 ? s[2] : s[3];
        out[i0 + lane] = sub ? out[i0 + lane] - t : t;
    }
}

// colsq[j] over rows [0, n) of V (row-major, ldv): grid (ceil(m/64), splits),
// lane = column, the four waves stride over rows; partials combined in a
// fixed order per split, then across splits by k_colsq_reduce.
__global__ __launch_bounds__(THREADS) void k_colsq_partial(const double *__restrict__ V, long ldv, int n, int m, int chunk, double *__restrict__ part)
{
    __shared__ double red[THREADS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long j = static_cast<long>(blockIdx.x) * 64 + lane;
    const long i0 = static_cast<long>(blockIdx.y) * chunk;
    const long i1 = min(static_cast<long>(n), i0 + chunk);
    double s = 0.0;
    if (j < m)
        for (long i = i0 + w; i < i1; i += 4)
        {
            const double v = V[i * ldv + j];
            s += v * v;
        }
    red[threadIdx.x] = s;
    __syncthreads();
    if (w == 0 && j < m)
        part[static_cast<long>(blockIdx.y) * m + j] = (red[lane] + red[64 + lane]) + (red[128 + lane] + red[192 + lane]);
}

__global__ __launch_bounds__(THREADS) void k_colsq_reduce(const double *__restrict__ part, int splits, int m, double *__restrict__ colsq)
{
    const long j = static_cast<long>(blockIdx.x) * THREADS + threadIdx.x;
    if (j >= m)
        return;
    double s = 0.0;
    for (int q = 0; q < splits; q++)
        s += part[static_cast<long>(q) * m + j];
    colsq[j] = s;
}

__global__ void k_variance_finish(const double *__restrict__ colsq, double v, int m, double *__restrict__ var)
{
    const long j = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < m)
        var[j] = v - colsq[j];
}

// Single workgroup, fixed-order tree reduction:
// out[0] = y^T alpha, out[1] = sum log(L_ii^2), out[2] = alpha^T alpha.
__global__ __launch_bounds__(THREADS) void k_loss_terms(const double *__restrict__ L,
                                                         long lda,
                                                         const double *__restrict__ y,
                                                         const double *__restrict__ alpha,
                                                         int n,
                                                         double *__restrict__ out)
{
    __shared__ double s1[THREADS], s2[THREADS], s3[THREADS];
    double a = 0.0, b = 0.0, c = 0.0;
    for (long i = threadIdx.x; i < n; i += THREADS)
    {
        const double al = alpha[i];
        a += y[i] * al;
        c += al * al;
        const double d = L[i * lda + i];
        b += log(d * d);
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    s3[threadIdx.x] = c;
    __syncthreads();
    for (int w = THREADS / 2; w > 0; w >>= 1)
    {
        if (threadIdx.x < w)
        {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
            s3[threadIdx.x] += s3[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        out[0] = s1[0];
        out[1] = s2[0];
        out[2] = s3[0];
    }
}

__global__ void k_set_diagonal(double *__restrict__ A, long lda, int n, double value)
{
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        A[i * lda + i] = value;
}

__global__ void k_residual_finish(const double *__restrict__ y, const double *__restrict__ kx, const double *__restrict__ alpha, double noise, int n, double *__restrict__ r)
{
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        r[i] = y[i] - kx[i] - noise * alpha[i];
}

__global__ void k_add_vec(double *__restrict__ y, const double *__restrict__ x, long n)
{
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        y[i] += x[i];
}

__global__ void k_fill(double *__restrict__ p, long count, double value)
{
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count)
        p[i] = value;
}

__global__ void k_set_int(int *p, int value) { *p = value; }

// One workgroup per 256-wide diagonal block: the 64-wide diagonal
// inverses are placed first, then the off-diagonal 64-blocks by increasing
// distance d = i - j, each from blocks of smaller distance:
//   S = sum_{t=j}^{i-1} L_it W_tj,   W_ij = -W_ii S.
// Block b covers rows row0 + 256 b .. (at most 256, up to row0 + rows) and is
// written k-major (W(r, c) at out[c * ldw + r], out = Wt + b * block_stride),
// zero above its diagonal and past its last row; a 64-block of W read as
// B^T rows is then a plain strided 64x64 read.
__global__ __launch_bounds__(THREADS) void k_invert_blocks256(const double *__restrict__ L, long lda, const double *__restrict__ Winv,
                                                               long row0, long rows, double *Wt, long ldw, long block_stride,
                                                               int zero_above)
{
    __shared__ double a[blk::NB * blk::XS];
    __shared__ double b[blk::NB * blk::XS];
    const int tid = threadIdx.x;
    const long j0 = row0 + static_cast<long>(blockIdx.x) * 256;
    const int q = static_cast<int>(min(4L, (row0 + rows - j0) / NB));
    const long jb = j0 / NB;  // first 64-block of this 256-block
    double *out = Wt + static_cast<long>(blockIdx.x) * block_stride;
    // zero the off-diagonal 64-blocks (stores only) ...
    for (int e = tid; e < 256 * 256; e += THREADS)
    {
        const int c = e >> 8, r = e & 255;
        if ((r >> 6) != (c >> 6) || (r >> 6) >= q)
            out[c * ldw + r] = 0.0;
    }
    // ... and place the 64-wide inverses, transposed into the k-major
    // layout through LDS, with every load of a block issued up front
    for (int bi = 0; bi < q; bi++)
    {
        blk::stage_block(a, Winv + (jb + bi) * NB * NB, NB);
        __syncthreads();
        for (int e = tid; e < NB * NB; e += THREADS)
        {
            const int c = e >> 6, r = e & 63;
            out[static_cast<long>(NB * bi + c) * ldw + NB * bi + r] = a[r * blk::XS + c];
        }
        __syncthreads();
    }
    if (zero_above)
    {
        // one inverse of the whole range (blocks along its diagonal): the
        // rows above this block, in its columns, are the inverse's zero part
        const long above = j0 - row0;
        for (long e = tid; e < above * 256; e += THREADS)
        {
            const long c = e / above, r = e % above;
            out[c * ldw + r - above] = 0.0;
        }
    }
    __syncthreads();
    for (int d = 1; d < q; d++)
        for (int j = 0; j + d < q; j++)
        {
            const int i = j + d;
            blk::v4d acc[2][2];
            blk::zero(acc);
            for (int t = j; t < i; t++)
            {
                blk::stage_block(a, L + (j0 + NB * i) * lda + j0 + NB * t, lda);
                // rows c of the staged block: W(64t + k, 64j + c) over k
                blk::stage_block(b, out + static_cast<long>(NB * j) * ldw + NB * t, ldw);
                __syncthreads();
                blk::mfma_nt_acc(a, b, acc);
                __syncthreads();
            }
            blk::stage_block(a, Winv + (jb + i) * NB * NB, NB);
            blk::for_each(acc, [&](int r, int c, double v) { b[c * blk::XS + r] = v; });
            __syncthreads();
            blk::v4d w[2][2];
            blk::mfma_nt(a, b, w);
            blk::for_each(w, [&](int r, int c, double v) { out[static_cast<long>(NB * j + c) * ldw + NB * i + r] = -v; });
            __syncthreads();
        }
}

// colsq[j] += sum_{i<rows} V[i][j]^2: 64 columns per workgroup, the
// sixteen wavefronts of a 1024-thread workgroup take rows g, g + 16, ...
// with eight loads in flight per lane; the sixteen partial sums are
// combined in a fixed order.
constexpr int CQ_WAVES = 16;
template <class T>
__global__ __launch_bounds__(64 * CQ_WAVES) void k_colsq_accumulate(const T *__restrict__ V, long ldv, int rows, int m,
                                                                    double *__restrict__ colsq)
{
    __shared__ double part[CQ_WAVES][64];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const long j = static_cast<long>(blockIdx.x) * 64 + c;
    double s = 0.0;
    if (j < m)
    {
        int i = g;
        for (; i + 7 * CQ_WAVES < rows; i += 8 * CQ_WAVES)
        {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; u++)
                v[u] = static_cast<double>(V[static_cast<long>(i + u * CQ_WAVES) * ldv + j]);
#pragma unroll
            for (int u = 0; u < 8; u++)
                s = fma(v[u], v[u], s);
        }
        for (; i < rows; i += CQ_WAVES)
        {
            const double v = static_cast<double>(V[static_cast<long>(i) * ldv + j]);
            s = fma(v, v, s);
        }
    }
    part[g][c] = s;
    __syncthreads();
    if (g == 0 && j < m)
    {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < CQ_WAVES; q++)
            t += part[q][c];
        colsq[j] += t;
    }
}

__global__ void k_fill_strided(double *__restrict__ p, long count, double value)
{
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += static_cast<long>(gridDim.x) * blockDim.x)
        p[i] = value;
}

__global__ void k_to_f32(const double *__restrict__ src, float *__restrict__ dst, long count)
{
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count)
        dst[i] = static_cast<float>(src[i]);
}

}  // namespace

void invert_blocks256(hipStream_t s, const double *L, long lda, const double *Winv, long row0, long rows, double *Wt, long ldw,
                      long block_stride, bool zero_above)
{
    require(row0 % NB == 0 && rows % NB == 0 && rows > 0 && lda >= row0 + rows && lda % 2 == 0 && ldw % 2 == 0 && ldw >= 256,
            "invert_blocks256: bad shape");
    require(reinterpret_cast<uintptr_t>(L) % 16 == 0 && reinterpret_cast<uintptr_t>(Wt) % 16 == 0,
            "invert_blocks256: operands must be 16-byte aligned");
    require(!zero_above || block_stride == 256 * ldw + 256, "invert_blocks256: zero_above needs the diagonal layout");
    hipLaunchKernelGGL(k_invert_blocks256, dim3(static_cast<unsigned>(ceil_div(rows, 256L))), dim3(THREADS), 0, s, L, lda, Winv, row0,
                       rows, Wt, ldw, block_stride, zero_above ? 1 : 0);
    GP_CHECK_LAUNCH();
}

template <class T>
static void colsq_accumulate_impl(hipStream_t s, const T *V, long ldv, int rows, int m, double *colsq)
{
    require(rows >= 0 && m >= 0 && ldv >= m, "colsq_accumulate: bad shape");
    if (rows == 0 || m == 0)
        return;
    hipLaunchKernelGGL(k_colsq_accumulate<T>, dim3(static_cast<unsigned>(ceil_div(m, 64))), dim3(64 * CQ_WAVES), 0, s, V, ldv, rows, m,
                       colsq);
    GP_CHECK_LAUNCH();
}

void colsq_accumulate(hipStream_t s, const double *V, long ldv, int rows, int m, double *colsq)
{
    colsq_accumulate_impl<double>(s, V, ldv, rows, m, colsq);
}

void colsq_accumulate_f32(hipStream_t s, const float *V, long ldv, int rows, int m, double *colsq)
{
    colsq_accumulate_impl<float>(s, V, ldv, rows, m, colsq);
}

void to_f32(hipStream_t s, const double *src, float *dst, long count)
{
    require(count >= 0, "to_f32: negative count");
    if (count == 0)
        return;
    hipLaunchKernelGGL(k_to_f32, dim3(static_cast<unsigned>(ceil_div(count, 256L))), dim3(THREADS), 0, s, src, dst, count);
    GP_CHECK_LAUNCH();
}

void panel_tasks(hipStream_t s, const PanelTaskArgs &h, const int *tasks, const int *wg_off, int groups)
{
    require(h.p0 >= 0 && h.p0 % NB == 0 && h.nb >= 1 && h.lda % 2 == 0 && groups >= 1, "panel_tasks: bad shape");
    require(!h.Wb || (h.ldw % 2 == 0 && h.ldw >= static_cast<long>(h.nb) * NB), "panel_tasks: bad inverse layout");
    require(!h.Xh || (h.hp0 >= 0 && h.hp0 % NB == 0 && h.nbh >= 1 && h.ldx % 2 == 0 && h.ldx >= static_cast<long>(h.nbh) * NB
                      && h.ldwh % 2 == 0 && h.ldwh >= static_cast<long>(h.nbh) * NB && h.Wh),
            "panel_tasks: bad head layout");
    PanelArgs a{ h, reinterpret_cast<const int4 *>(tasks), wg_off };
    hipLaunchKernelGGL(k_panel_tasks, dim3(groups), dim3(THREADS), 0, s, a);
    GP_CHECK_LAUNCH();
}

void rows_solve(hipStream_t s, const double *A, long lda, long r0, long p0, int rows, int width, const double *Wb, long ldw, double *X, long ldx)
{
    require(rows > 0 && rows % NB == 0 && width > 0 && width % NB == 0 && lda % 2 == 0 && ldw % 2 == 0 && ldx % 2 == 0 && ldw >= width
                && ldx >= width && p0 % 2 == 0,
            "rows_solve: bad shape");
    const int nbw = width / NB;
    require(rows / NB <= 65535, "rows_solve: too many rows for one launch");
    hipLaunchKernelGGL(k_rows_solve, dim3((nbw + 1) / 2, rows / NB), dim3(THREADS), 0, s, A, lda, r0, p0, nbw, Wb, ldw, X, ldx);
    GP_CHECK_LAUNCH();
}

void leaf_solve(hipStream_t s, const double *Wk, long ldw, int rows, const double *B, long ldb, int cols, double *V, long ldv)
{
    require(rows > 0 && rows % NB == 0 && cols > 0 && cols % NB == 0 && ldw % 2 == 0 && ldb % 2 == 0 && ldv % 2 == 0 && ldw >= rows
                && ldb >= cols && ldv >= cols,
            "leaf_solve: bad shape");
    const int nbr = rows / NB;
    require((nbr + 1) / 2 <= 65535, "leaf_solve: too many rows for one launch");
    hipLaunchKernelGGL(k_leaf_solve, dim3(cols / NB, (nbr + 1) / 2), dim3(THREADS), 0, s, Wk, ldw, nbr, B, ldb, V, ldv);
    GP_CHECK_LAUNCH();
}

void solve_lower_step(hipStream_t s, const double *L, long lda, const double *Winv, double *rhs, double *out, int n_pad, int j)
{
    const int nb = n_pad / NB;
    require(n_pad % NB == 0 && j >= 0 && j < nb, "solve_lower_step: bad shape");
    hipLaunchKernelGGL(k_solve_lower_step, dim3(nb - j), dim3(THREADS), 0, s, L, lda, Winv, rhs, out, j);
    GP_CHECK_LAUNCH();
}

void solve_upper_step(hipStream_t s, const double *L, long lda, const double *Winv, double *rhs, double *out, int n_pad, int j)
{
    const int nb = n_pad / NB;
    require(n_pad % NB == 0 && j >= 0 && j < nb, "solve_upper_step: bad shape");
    hipLaunchKernelGGL(k_solve_upper_step, dim3(j + 1), dim3(THREADS), 0, s, L, lda, Winv, rhs, out, j);
    GP_CHECK_LAUNCH();
}

void coldot(hipStream_t s, const double *M, long ld, int K, int N, const double *v, double *out, bool sub)
{
    require(K >= 1 && N >= 1 && ld >= N, "coldot: bad shape");
    hipLaunchKernelGGL(k_coldot, dim3(ceil_div(N, 64)), dim3(64 * COLDOT_WAVES), 0, s, M, ld, K, N, v, out, sub);
    GP_CHECK_LAUNCH();
}

void rowdot(hipStream_t s, const double *M, long ld, int K, int N, const double *v, double *out, bool sub)
{
    require(K >= 64 && K % 64 == 0 && N >= 1 && ld >= K, "rowdot: bad shape");
    hipLaunchKernelGGL(k_rowdot, dim3(ceil_div(N, 16)), dim3(THREADS), 0, s, M, ld, K, N, v, out, sub);
    GP_CHECK_LAUNCH();
}

int colsq_splits(int n, int m)
{
    const long col_blocks = ceil_div(m, 64);
    long splits = ceil_div(1024, col_blocks);
    splits = std::min<long>(splits, ceil_div(n, 256));
    return static_cast<int>(std::max<long>(splits, 1));
}

void column_sumsq(hipStream_t s, const double *V, long ldv, int n, int m, double *part, int splits, double *colsq)
{
    if (m <= 0)
        return;
    require(n >= 1 && splits >= 1, "column_sumsq: bad shape");
    const int chunk = static_cast<int>(ceil_div(n, splits));
    hipLaunchKernelGGL(k_colsq_partial, dim3(ceil_div(m, 64), splits), dim3(THREADS), 0, s, V, ldv, n, m, chunk, part);
    GP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_colsq_reduce, dim3(ceil_div(m, THREADS)), dim3(THREADS), 0, s, part, splits, m, colsq);
    GP_CHECK_LAUNCH();
}

void variance_finish(hipStream_t s, const double *colsq, double v, int m, double *var)
{
    if (m <= 0)
        return;
    hipLaunchKernelGGL(k_variance_finish, dim3(ceil_div(m, 256)), dim3(256), 0, s, colsq, v, m, var);
    GP_CHECK_LAUNCH();
}

void loss_terms(hipStream_t s, const double *L, long lda, const double *y, const double *alpha, int n, double *out)
{
    hipLaunchKernelGGL(k_loss_terms, dim3(1), dim3(THREADS), 0, s, L, lda, y, alpha, n, out);
    GP_CHECK_LAUNCH();
}

void set_diagonal(hipStream_t s, double *A, long lda, int n, double value)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_set_diagonal, dim3(ceil_div(n, 256)), dim3(256), 0, s, A, lda, n, value);
    GP_CHECK_LAUNCH();
}

void residual_finish(hipStream_t s, const double *y, const double *kx, const double *alpha, double noise, int n, double *r)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_residual_finish, dim3(ceil_div(n, 256)), dim3(256), 0, s, y, kx, alpha, noise, n, r);
    GP_CHECK_LAUNCH();
}

void add_vec(hipStream_t s, double *y, const double *x, long n)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_add_vec, dim3(ceil_div(n, 256)), dim3(256), 0, s, y, x, n);
    GP_CHECK_LAUNCH();
}

void fill_narrow(hipStream_t s, double *p, long count, double value)
{
    require(count >= 0, "fill_narrow: negative count");
    if (count == 0)
        return;
    hipLaunchKernelGGL(k_fill_strided, dim3(static_cast<unsigned>(std::min<long>(ceil_div(count, 256L), 8))), dim3(THREADS), 0, s, p, count,
                       value);
    GP_CHECK_LAUNCH();
}

void fill(hipStream_t s, double *p, long count, double value)
{
    if (count <= 0)
        return;
    hipLaunchKernelGGL(k_fill, dim3(ceil_div(count, 256)), dim3(256), 0, s, p, count, value);
    GP_CHECK_LAUNCH();
}

void set_int(hipStream_t s, int *p, int value)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, p, value);
    GP_CHECK_LAUNCH();
}

}  // namespace gpamd::k
