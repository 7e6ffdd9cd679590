// Task schedule of one diagonal block (panel_plan.hpp).
//
// Dependencies are tracked on 64x64 blocks: the final factor block L_ij
// (written last by SOLVE(i, j), or by the factor of block j when i == j),
// Winv_k, W_ij and the joins' intermediate T blocks. A task may run in a
// launch after every launch that produces one of its inputs; tasks of one
// launch must not write what another of them reads or writes.
#include "panel_plan.hpp"

#include <algorithm>
#include <map>
#include <set>
#include <stdexcept>
#include <string>

namespace gpamd
{
namespace
{

// Estimated cost of a task on one workgroup (microseconds beside the
// trailing update, roughly): a staged 64x64 block ~1.2, a 64x64x64
// product on the matrix cores ~1.7, the 64x64 factor + inverse ~20.
double cost_of(int stages, int products, double extra = 0.0) { return 1.0 + 1.2 * stages + 1.7 * products + extra; }

enum Kind : long
{
    K_L = 1,     // final factor block
    K_WINV = 2,  // row-major inverse of a diagonal block
    K_W = 3,     // block of the k-major inverse
    K_T = 4,     // join intermediate (tmp)
    K_A = 5,     // any version of a factor block (same-launch conflicts)
};

long key(long kind, long a, long b) { return (kind << 48) | (a << 24) | b; }

struct Task
{
    PanelTask t;
    double cost;
    std::vector<long> reads;   // K_L / K_WINV / K_W / K_T inputs (final versions)
    std::vector<long> writes;  // everything written (K_A for factor blocks)
    std::vector<long> touch;   // factor blocks read in any version (chain tasks)
    std::vector<long> makes;   // final versions this task produces
};

struct Launch
{
    std::vector<Task> tasks;
    std::vector<double> loads;  // estimated per-workgroup work (placement only)
    int cap = 1;

    double makespan() const { return loads.empty() ? 0.0 : *std::max_element(loads.begin(), loads.end()); }
    // where the next task would go: an unused workgroup while the cap allows
    double least() const
    {
        if (static_cast<int>(loads.size()) < cap)
            return 0.0;
        return *std::min_element(loads.begin(), loads.end());
    }
    void add(Task t)
    {
        if (static_cast<int>(loads.size()) < cap)
            loads.push_back(t.cost);
        else
            *std::min_element(loads.begin(), loads.end()) += t.cost;
        tasks.push_back(std::move(t));
    }
};

PanelTask pt(int type, int y, int z = 0, int w = 0, int h = 0) { return PanelTask{ type | (h << 8), y, z, w }; }

// Longest task first onto the least loaded workgroup; appends the launch
// to the plan (nothing for an empty launch).
void append_launch(PanelPlan &plan, Launch &l)
{
    if (l.tasks.empty())
        return;
    std::stable_sort(l.tasks.begin(), l.tasks.end(), [](const Task &a, const Task &b) { return a.cost > b.cost; });
    const int g = std::min(static_cast<int>(l.tasks.size()), l.cap);
    std::vector<double> load(g, 0.0);
    std::vector<std::vector<PanelTask>> per(g);
    for (const auto &t : l.tasks)
    {
        const int w = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
        load[w] += t.cost;
        per[w].push_back(t.t);
    }
    PanelLaunch pl{ static_cast<int>(plan.wg_off.size()), g };
    for (int w = 0; w < g; w++)
    {
        plan.wg_off.push_back(static_cast<int>(plan.tasks.size()));
        plan.tasks.insert(plan.tasks.end(), per[w].begin(), per[w].end());
    }
    plan.wg_off.push_back(static_cast<int>(plan.tasks.size()));
    plan.launches.push_back(pl);
}

}  // namespace

PanelPlan build_panel_plan(int nb, int chain_wgs, int tail_wgs)
{
    if (nb < 1 || nb > 4096 || chain_wgs < 1 || tail_wgs < 1)
        throw std::runtime_error("build_panel_plan: bad arguments");
    std::vector<Launch> L;
    auto new_launch = [&](bool chain)
    {
        Launch l;
        l.cap = chain ? chain_wgs : tail_wgs;
        L.push_back(std::move(l));
        return static_cast<int>(L.size()) - 1;
    };
    std::map<long, int> made_in;  // final version -> launch index

    // ---- the chain: factor, solves, updates (+ the next factor) ----------
    auto chain_task = [&](int li, PanelTask t, double cost, std::vector<long> touch, std::vector<long> writes, std::vector<long> makes,
                          std::vector<long> reads)
    {
        Task k{ t, cost, std::move(reads), std::move(writes), std::move(touch), std::move(makes) };
        for (long m : k.makes)
            made_in[m] = li;
        L[li].add(std::move(k));
    };
    {
        const int l0 = new_launch(true);
        chain_task(l0, pt(PT_FACTOR, 0), cost_of(1, 0, 19.0), { key(K_A, 0, 0) }, { key(K_A, 0, 0), key(K_WINV, 0, 0) },
                   { key(K_L, 0, 0), key(K_WINV, 0, 0) }, {});
    }
    for (int k = 0; k + 1 < nb; k++)
    {
        const int ls = new_launch(true);
        for (int i = k + 1; i < nb; i++)
            chain_task(ls, pt(PT_SOLVE, i, k), cost_of(2, 1), { key(K_A, i, k) }, { key(K_A, i, k) }, { key(K_L, i, k) }, { key(K_WINV, k, 0) });
        const int lu = new_launch(true);
        chain_task(lu, pt(PT_UFACTOR, k + 1, k), cost_of(2, 1, 19.0), { key(K_A, k + 1, k + 1), key(K_A, k + 1, k) },
                   { key(K_A, k + 1, k + 1), key(K_WINV, k + 1, 0) }, { key(K_L, k + 1, k + 1), key(K_WINV, k + 1, 0) }, {});
        for (int i = k + 1; i < nb; i++)
            for (int j = k + 1; j <= i; j++)
                if (!(i == k + 1 && j == k + 1))
                    chain_task(lu, pt(PT_UPDATE, i, j, k), cost_of(3, 1), { key(K_A, i, j), key(K_A, i, k), key(K_A, j, k) }, { key(K_A, i, j) }, {},
                               {});
    }
    const int chain_end = static_cast<int>(L.size());

    // ---- the inverse, in dependency order --------------------------------
    std::vector<Task> inv;
    for (int k = 0; k < nb; k++)
    {
        Task t{ pt(PT_PLACE, k), cost_of(1, 0, 1.0), { key(K_WINV, k, 0) }, {}, {}, {} };
        for (int j = k; j < nb; j++)
            t.writes.push_back(key(K_W, k, j));
        t.makes.push_back(key(K_W, k, k));
        inv.push_back(std::move(t));
    }
    for (int a = 0; a + 1 < nb; a += 2)
    {
        const int b = a + 1;
        inv.push_back(Task{ pt(PT_JOINF, a), cost_of(3, 2),
                            { key(K_WINV, a, 0), key(K_WINV, b, 0), key(K_L, b, a) },
                            { key(K_W, b, a) },
                            {},
                            { key(K_W, b, a) } });
    }
    long tmp_blocks = 0;
    for (int h = 2; h < nb; h *= 2)
    {
        for (int p = 0; 2 * h * p + h < nb; p++)
        {
            const int a0 = 2 * h * p, b0 = a0 + h, hb = std::min(h, nb - b0);
            const long off = tmp_blocks;
            tmp_blocks += static_cast<long>(h) * h;
            auto tkey = [&](int r, int c) { return key(K_T, off + static_cast<long>(c) * h + r, 0); };
            for (int cb = 0; cb < h; cb++)
                for (int rb = 0; rb < hb; rb++)
                {
                    Task t{ pt(PT_JOIN0, p, rb | (cb << 16), static_cast<int>(off), h), cost_of(2 * (h - cb), h - cb), {}, { tkey(rb, cb) }, {}, { tkey(rb, cb) } };
                    for (int s = cb; s < h; s++)
                    {
                        t.reads.push_back(key(K_L, b0 + rb, a0 + s));
                        t.reads.push_back(key(K_W, a0 + s, a0 + cb));
                    }
                    inv.push_back(std::move(t));
                }
            for (int cb = 0; cb < h; cb++)
                for (int rb = 0; rb < hb; rb++)
                {
                    Task t{ pt(PT_JOIN1, p, rb | (cb << 16), static_cast<int>(off), h), cost_of(2 * (rb + 1), rb + 1), {}, { key(K_W, b0 + rb, a0 + cb) }, {},
                            { key(K_W, b0 + rb, a0 + cb) } };
                    for (int s = 0; s <= rb; s++)
                    {
                        t.reads.push_back(key(K_W, b0 + rb, b0 + s));
                        t.reads.push_back(tkey(s, cb));
                    }
                    inv.push_back(std::move(t));
                }
        }
    }

    auto ready = [&](const Task &t)
    {
        int r = 0;
        for (long k : t.reads)
        {
            auto it = made_in.find(k);
            if (it == made_in.end())
                throw std::runtime_error("build_panel_plan: input produced by no earlier task");
            r = std::max(r, it->second + 1);
        }
        return r;
    };
    for (auto &t : inv)
    {
        const int r = ready(t);
        int at = -1;
        for (int li = r; li < chain_end && at < 0; li++)
            if (L[li].least() + t.cost <= L[li].makespan())
                at = li;
        // otherwise the chain launch it lengthens least, if that costs less
        // than running it after the chain
        double best = t.cost;
        for (int li = r; li < chain_end && at < 0; li++)
            best = std::min(best, L[li].least() + t.cost - L[li].makespan());
        for (int li = r; li < chain_end && at < 0 && best < t.cost; li++)
            if (L[li].least() + t.cost - L[li].makespan() <= best)
                at = li;
        if (at < 0)
        {
            at = std::max(r, chain_end);
            while (static_cast<int>(L.size()) <= at)
                new_launch(false);
        }
        for (long m : t.makes)
            made_in[m] = at;
        L[at].add(std::move(t));
    }

    // ---- self-check ------------------------------------------------------
    for (std::size_t li = 0; li < L.size(); li++)
    {
        std::set<long> written;
        for (const auto &t : L[li].tasks)
        {
            for (long w : t.writes)
                if (!written.insert(w).second)
                    throw std::runtime_error("build_panel_plan: two tasks of launch " + std::to_string(li) + " write one block");
            for (long r : t.reads)
                if (made_in.at(r) >= static_cast<int>(li))
                    throw std::runtime_error("build_panel_plan: task reads a block its own launch or a later one produces");
        }
        for (const auto &t : L[li].tasks)
        {
            auto others_write = [&](long k)
            {
                bool mine = std::find(t.writes.begin(), t.writes.end(), k) != t.writes.end();
                return !mine && written.count(k);
            };
            for (long r : t.touch)
                if (others_write(r))
                    throw std::runtime_error("build_panel_plan: a task reads a block another task of its launch writes");
            for (long r : t.reads)
            {
                // final factor blocks are K_A blocks nobody writes any more
                const long kind = r >> 48;
                const long alias = kind == K_L ? key(K_A, (r >> 24) & 0xffffff, r & 0xffffff) : r;
                if (others_write(alias))
                    throw std::runtime_error("build_panel_plan: a task reads a block another task of its launch writes");
            }
        }
    }

    PanelPlan plan;
    plan.nb = nb;
    plan.tmp_doubles = tmp_blocks * 4096;
    for (auto &l : L)
        append_launch(plan, l);
    return plan;
}

PanelPlan build_head_plan(int nbh, int nb, int wgs)
{
    if (nbh < 1 || nb < 1 || nbh > 4096 || nb > 4096 || wgs < 1)
        throw std::runtime_error("build_head_plan: bad arguments");
    PanelPlan plan;
    plan.nb = nb;
    Launch x, u;
    x.cap = u.cap = wgs;
    for (int i = 0; i < nb; i++)
        for (int c = 0; c < nbh; c++)
            x.tasks.push_back(Task{ pt(PT_HEADX, i, c), cost_of(2 * (c + 1), c + 1), {}, {}, {}, {} });
    for (int i = 0; i < nb; i++)
        for (int j = 0; j <= i; j++)
            u.tasks.push_back(Task{ pt(PT_HEADS, i, j), cost_of(2 * nbh, nbh, 1.5), {}, {}, {}, {} });
    append_launch(plan, x);
    append_launch(plan, u);
    return plan;
}

}  // namespace gpamd
