// Task schedule of one diagonal block of the blocked Cholesky: its factor
// and its inverse as 64x64 block tasks grouped into launches of one
// kernel (k_panel_tasks). Plain C++ (no HIP), so the schedule can be built
// and checked on any host.
#pragma once

#include <vector>

namespace gpamd
{

// Block (i, j) below is the 64x64 block at rows 64 i, columns 64 j of the
// diagonal block; L its factor, W = L^-1, Winv_k the row-major inverse of
// L's diagonal block k.
enum PanelTaskType : int
{
    PT_FACTOR = 0,   // (k): factor block (k, k) in place; Winv_k = its inverse
    PT_UFACTOR = 1,  // (k, q): block (k, k) -= L_kq L_kq^T, then as PT_FACTOR
    PT_SOLVE = 2,    // (i, k): L_ik = A_ik Winv_k^T in place
    PT_UPDATE = 3,   // (i, j, k): A_ij -= L_ik L_jk^T (lower half if i == j)
    PT_PLACE = 4,    // (k): W_kk = Winv_k; W_kj = 0 for j > k
    PT_JOINF = 5,    // (a): W_ba = -Winv_b L_ba Winv_a, b = a + 1
    PT_JOIN0 = 6,    // doubling level h, pair p, block (rb, cb): T = L_ba W_a (into tmp)
    PT_JOIN1 = 7,    // the same: W_ba = -W_b T
    // head of the next diagonal block (its rows in the previous panel P,
    // whose diagonal block has the k-major inverse W_P):
    PT_HEADX = 8,  // (i, c): X_ic = sum_{t <= c} A_{i, P t} W_P(c, t)^T into scratch
    PT_HEADS = 9,  // (i, j): A_ij -= sum_t X_it X_jt^T (lower half if i == j)
};

// Four ints per task (an int4 on the device):
//   x = type | h << 8, y, z, w as listed per type:
//   FACTOR y = k; UFACTOR y = k, z = q; SOLVE y = i, z = k;
//   UPDATE y = i, z = j, w = k; PLACE y = k; JOINF y = a;
//   JOIN0 / JOIN1 y = p (a = 2 h p, b = a + h), z = rb | cb << 16,
//   w = offset of the pair's T in tmp, in 64x64 blocks (T(r, c) of the
//   64h x 64h pair matrix at tmp[w * 4096 + c * 64 h + r]);
//   HEADX y = i, z = c; HEADS y = i, z = j.
struct PanelTask
{
    int x, y, z, w;
};

// Launch l runs `groups` workgroups; workgroup g executes the tasks
// [wg_off[first + g], wg_off[first + g + 1]) in order. Tasks of one launch
// never touch a block that another task of the same launch writes.
struct PanelLaunch
{
    int first;
    int groups;
};

struct PanelPlan
{
    int nb = 0;
    std::vector<PanelTask> tasks;
    std::vector<int> wg_off;
    std::vector<PanelLaunch> launches;
    long tmp_doubles = 0;  // scratch the JOIN tasks need
};

// Schedule of an nb x nb block diagonal block (nb 64-wide blocks):
// launch 0 factors block 0; then per step k a launch of the solves of
// column k and a launch of the rank-64 updates, whose workgroup 0 also
// factors block k + 1 (the chain). The inverse (PLACE and the doubling
// joins) is slotted into the chain's launches as soon as its inputs are
// final and it does not lengthen the launch (estimated cost), the rest
// into launches after the chain. Chain launches use at most chain_wgs
// workgroups, the later ones at most tail_wgs. Throws std::runtime_error
// if the result violates its own dependency check.
PanelPlan build_panel_plan(int nb, int chain_wgs, int tail_wgs);

// The previous panel's (nbh blocks wide) contribution to the next
// diagonal block (nb blocks): launch 0 the HEADX tasks (the block's rows of
// the panel, solved), launch 1 the HEADS tasks (the rank-64 nbh update of
// the block's lower half), each on at most wgs workgroups.
PanelPlan build_head_plan(int nbh, int nb, int wgs);

}  // namespace gpamd
