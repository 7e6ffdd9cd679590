// Batched fit of many GPs: routing between the persistent kernel (one
// workgroup runs a whole GP) and the level-synchronous path (one launch per
// factorisation step over all large GPs, one workgroup per 64x64 block),
// per size class from the autotuned table, and the captured launch
// sequences (HIP graphs) of the level path.
#include <dlfcn.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <sstream>

#include "engine.hpp"

namespace gpamd
{

namespace
{

// One row of the tuning table: GPs with n_pad <= n_pad_max (and above the
// previous row's) take `level` (level-synchronous path) or not, and mixed
// precision when it was measured faster than FP64 for the class.
struct TuneRow
{
    int n_pad_max;
    int path;  // 0 persistent, 1 level (64-wide), 2 level128
    bool mixed_faster;
};

int path_code(const std::string &p) { return p == "level128" ? 2 : (p == "level" ? 1 : 0); }

struct Tuning
{
    std::vector<TuneRow> rows;
    std::string source = "built-in defaults";
};

std::string library_dir()
{
    Dl_info info{};
    if (dladdr(reinterpret_cast<void *>(&library_dir), &info) && info.dli_fname)
    {
        std::string p(info.dli_fname);
        const auto slash = p.rfind('/');
        if (slash != std::string::npos)
            return p.substr(0, slash);
    }
    return ".";
}

// The table for the current device's architecture: GPRAT_BATCH_TUNING, or
// tuning/batch_<arch>.txt next to the library. Lines "n_pad_max path
// mixed_faster ..." (path "level" or "persistent"); '#' starts a comment.
Tuning load_tuning(int device)
{
    Tuning t;
    std::string path;
    if (const char *e = std::getenv("GPRAT_BATCH_TUNING"))
        path = e;
    else
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        {
            (void)hipGetLastError();
            return t;
        }
        std::string arch(prop.gcnArchName);
        arch = arch.substr(0, arch.find(':'));
        path = library_dir() + "/tuning/batch_" + arch + ".txt";
    }
    std::ifstream in(path);
    if (!in)
        return t;
    std::string line;
    while (std::getline(in, line))
    {
        const auto hash = line.find('#');
        if (hash != std::string::npos)
            line.resize(hash);
        std::istringstream ls(line);
        TuneRow r{};
        std::string p;
        int mf = 0;
        if (!(ls >> r.n_pad_max >> p >> mf))
            continue;
        r.path = path_code(p);
        r.mixed_faster = mf != 0;
        t.rows.push_back(r);
    }
    std::sort(t.rows.begin(), t.rows.end(), [](const TuneRow &a, const TuneRow &b) { return a.n_pad_max < b.n_pad_max; });
    if (!t.rows.empty())
        t.source = path;
    return t;
}

const Tuning &tuning(int device)
{
    static std::mutex mu;
    static std::map<int, Tuning> cache;
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(device);
    if (it == cache.end())
        it = cache.emplace(device, load_tuning(device)).first;
    return it->second;
}

// Path and precision of one GP. GPRAT_BATCH_PATH (persistent | level |
// level128) and GPRAT_BATCH_MIXED (requested: mixed precision for every GP
// of a mixed batch) override the table, for measurement. The 128-wide level
// path is FP64: a mixed GP routed there takes the 64-wide one.
struct Route
{
    int path;
    bool mixed;
};

Route route(const Tuning &t, int n_pad, bool mixed_requested)
{
    Route r{ n_pad > 1024 ? 1 : 0, mixed_requested };
    const TuneRow *row = nullptr;
    for (const auto &x : t.rows)
        if (n_pad <= x.n_pad_max)
        {
            row = &x;
            break;
        }
    if (!row && !t.rows.empty())
        row = &t.rows.back();
    if (row)
    {
        r.path = row->path;
        r.mixed = mixed_requested && row->mixed_faster;
    }
    if (const char *e = std::getenv("GPRAT_BATCH_PATH"))
        r.path = path_code(e);
    if (const char *e = std::getenv("GPRAT_BATCH_MIXED"))
        if (std::string(e) == "requested")
            r.mixed = mixed_requested;
    if (r.mixed && r.path == 2)
        r.path = 1;
    return r;
}

// Items of the level path's launches for the GPs `ids` (64-blocks per GP
// in nb): Assemble (every block row), per column jb Diagonal (GPs with more
// than jb blocks) and OffDiagonal (their blocks below jb, and with prefix,
// for jb >= 1, {GP, -1} for the GPs with more than jb + 1 blocks: their
// next diagonal block's first sums), then Finish (every GP) and, with
// refine, every 64-row block for the refinement's residual.
struct LevelPlan
{
    std::vector<int2> items;
    long assemble_off = 0, assemble_n = 0, finish_off = 0, finish_n = 0, rows_off = 0, rows_n = 0;
    std::vector<long> diag_off, diag_n, off_off, off_n;
};

// nb: 64-row blocks per GP; width: the column width of the factorisation
// steps (64 or 128; blocks per step column = ceil(nb * 64 / width)).
LevelPlan level_plan(const std::vector<int> &ids, const std::vector<int> &nb64, int width, bool prefix, bool refine)
{
    LevelPlan p;
    std::vector<int> nb(nb64.size());
    for (std::size_t g = 0; g < nb64.size(); g++)
        nb[g] = static_cast<int>(ceil_div(static_cast<long>(nb64[g]) * 64, width));
    int maxb = 0;
    for (int g : ids)
        maxb = std::max(maxb, nb[g]);
    // the 64-wide steps generate each block of K when they reduce it; the
    // 128-wide ones update the factor in place, so K is assembled first
    p.assemble_off = 0;
    if (width == 128)
        for (int g : ids)
            for (int i = 0; i < nb64[g]; i++)
                p.items.push_back(make_int2(g, i));
    p.assemble_n = static_cast<long>(p.items.size());
    for (int jb = 0; jb < maxb; jb++)
    {
        p.diag_off.push_back(static_cast<long>(p.items.size()));
        for (int g : ids)
            if (nb[g] > jb)
                p.items.push_back(make_int2(g, jb));
        p.diag_n.push_back(static_cast<long>(p.items.size()) - p.diag_off.back());
        p.off_off.push_back(static_cast<long>(p.items.size()));
        // the next diagonal blocks' sums are as long as this column's blocks
        if (prefix && jb >= 1)
            for (int g : ids)
                if (nb[g] > jb + 1)
                    p.items.push_back(make_int2(g, -1));
        // rows first, so the workgroups of one GP's column run side by side
        for (int g : ids)
            for (int i = jb + 1; i < nb[g]; i++)
                p.items.push_back(make_int2(g, i));
        p.off_n.push_back(static_cast<long>(p.items.size()) - p.off_off.back());
    }
    p.finish_off = static_cast<long>(p.items.size());
    for (int g : ids)
        p.items.push_back(make_int2(g, 0));
    p.finish_n = static_cast<long>(p.items.size()) - p.finish_off;
    // the mixed-precision refinement's residual: every 64-row block
    p.rows_off = static_cast<long>(p.items.size());
    if (refine)
        for (int g : ids)
            for (int i = 0; i < nb64[g]; i++)
                p.items.push_back(make_int2(g, i));
    p.rows_n = static_cast<long>(p.items.size()) - p.rows_off;
    return p;
}

void launch_level(hipStream_t s, const LevelPlan &p, const int2 *items, const k::BatchGP *gps, bool mixed, bool wide, double *losses,
                  int *info, int *steps)
{
    k::batch_level_step(s, k::BatchLevelStep::Assemble, gps, items + p.assemble_off, static_cast<int>(p.assemble_n), 0, mixed, losses,
                        info, steps);
    const auto diag = wide ? k::BatchLevelStep::Diagonal128 : k::BatchLevelStep::Diagonal;
    const auto off = wide ? k::BatchLevelStep::OffDiagonal128 : k::BatchLevelStep::OffDiagonal;
    for (std::size_t jb = 0; jb < p.diag_off.size(); jb++)
    {
        k::batch_level_step(s, diag, gps, items + p.diag_off[jb], static_cast<int>(p.diag_n[jb]), static_cast<int>(jb), mixed, losses, info,
                            steps);
        k::batch_level_step(s, off, gps, items + p.off_off[jb], static_cast<int>(p.off_n[jb]), static_cast<int>(jb), mixed, losses, info,
                            steps);
    }
    const int2 *fin = items + p.finish_off;
    const int nfin = static_cast<int>(p.finish_n);
    if (!mixed)
    {
        k::batch_level_step(s, wide ? k::BatchLevelStep::Finish128 : k::BatchLevelStep::Finish, gps, fin, nfin, 0, false, losses, info,
                            steps);
        return;
    }
    k::batch_level_step(s, k::BatchLevelStep::RefineStart, gps, fin, nfin, 0, true, losses, info, steps);
    for (int it = 0; it < k::BATCH_REFINE_STEPS; it++)
    {
        k::batch_level_step(s, k::BatchLevelStep::RefineResidual, gps, items + p.rows_off, static_cast<int>(p.rows_n), 0, true, losses,
                            info, steps);
        k::batch_level_step(s, k::BatchLevelStep::RefineCorrect, gps, fin, nfin, it, true, losses, info, steps);
    }
    k::batch_level_step(s, k::BatchLevelStep::RefineFinish, gps, fin, nfin, 0, true, losses, info, steps);
}

}  // namespace

// A captured level-path launch sequence with the buffers its kernels read
// and write (the descriptor table is refreshed before every replay).
struct BatchGraph
{
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf<double> desc;
    DevBuf<int2> items;
    DevBuf<int> info, steps;
    DevBuf<double> losses;
    unsigned long long last_use = 0;
    ~BatchGraph()
    {
        if (exec)
            hipGraphExecDestroy(exec);
        if (graph)
            hipGraphDestroy(graph);
    }
};

std::string batch_tuning_source(int device) { return tuning(device).source; }

// Fits `gps` (device buffers filled): GPs routed to the persistent kernel on
// the main stream, the level-synchronous ones on the panel stream at the
// same time. A level sequence whose batch composition (sizes, order,
// precisions) was seen before on this context is replayed from a captured
// HIP graph (captured on its second occurrence). Returns each GP's first
// non-positive pivot column (INT_MAX if none).
std::vector<int> batch_run(Context &c, const std::vector<k::BatchGP> &gps, double *losses_host, bool mixed, int *refine_host,
                           std::vector<char> *mixed_out)
{
    const int count = static_cast<int>(gps.size());
    require(count >= 1, "batch: need at least one GP");
    c.make_current();
    const Tuning &tune = tuning(c.device);
    std::vector<int> nb(count);
    std::vector<Route> routes(count);
    for (int g = 0; g < count; g++)
    {
        require(gps[g].n_pad % 64 == 0 && gps[g].n_pad >= gps[g].n, "batch: bad padding");
        nb[g] = gps[g].n_pad / 64;
        routes[g] = route(tune, gps[g].n_pad, mixed);
        if (routes[g].mixed)
            require(gps[g].work != nullptr, "batch: mixed precision needs a work vector per GP");
    }
    // largest first within each group
    std::vector<int> order(count);
    for (int g = 0; g < count; g++)
        order[g] = g;
    std::stable_sort(order.begin(), order.end(), [&](int a, int z) { return gps[a].n > gps[z].n; });
    // persistent kernel: [mixed]; level path: FP64 64-wide, mixed 64-wide, FP64 128-wide
    std::vector<int> pers[2];
    struct LevelGroup
    {
        std::vector<int> ids;
        bool mixed, wide;
        LevelPlan plan;
        long base = 0;
    };
    LevelGroup groups[3] = { { {}, false, false, {}, 0 }, { {}, true, false, {}, 0 }, { {}, false, true, {}, 0 } };
    for (int g : order)
    {
        if (routes[g].path == 0)
            pers[routes[g].mixed ? 1 : 0].push_back(g);
        else
            groups[routes[g].path == 2 ? 2 : (routes[g].mixed ? 1 : 0)].ids.push_back(g);
    }
    // the 128-wide steps keep each block column's 128x128 inverse
    std::vector<k::BatchGP> dgps = gps;
    long w128_total = 0;
    for (int g : groups[2].ids)
        w128_total += ceil_div(gps[g].n_pad, 128) * 128 * 128;
    DevBuf<double> w128;
    if (w128_total > 0)
    {
        w128 = DevBuf<double>(c, static_cast<std::size_t>(w128_total));
        long off = 0;
        for (int g : groups[2].ids)
        {
            dgps[g].W128 = w128.get() + off;
            off += ceil_div(gps[g].n_pad, 128) * 128 * 128;
        }
    }

    // the FP64 64-wide level steps carry each next diagonal block's first
    // sums from one launch to the next; the mixed ones, the refinement's
    // state
    long pre_total = 0;
    for (int g : groups[0].ids)
        pre_total += nb[g] >= 3 ? k::BATCH_CARRY_DOUBLES : 0;
    pre_total += static_cast<long>(groups[1].ids.size()) * k::BATCH_REFINE_STATE_DOUBLES;
    DevBuf<double> pre;
    if (pre_total > 0)
    {
        pre = DevBuf<double>(c, static_cast<std::size_t>(pre_total));
        long off = 0;
        for (int g : groups[0].ids)
            if (nb[g] >= 3)
            {
                dgps[g].pre = pre.get() + off;
                off += k::BATCH_CARRY_DOUBLES;
            }
        for (int g : groups[1].ids)
        {
            dgps[g].pre = pre.get() + off;
            off += k::BATCH_REFINE_STATE_DOUBLES;
        }
    }

    DevBuf<double> desc(c, (sizeof(k::BatchGP) * count + 7) / 8);
    DevBuf<int> dorder(c, count), info(c, count), counter(c, 2);
    DevBuf<double> losses(c, count);
    DevBuf<int> steps(c, count);
    GP_HIP(hipMemcpyAsync(desc.get(), dgps.data(), sizeof(k::BatchGP) * count, hipMemcpyDefault, c.main));
    std::vector<int> porder;
    for (int m = 0; m < 2; m++)
        porder.insert(porder.end(), pers[m].begin(), pers[m].end());
    if (!porder.empty())
        GP_HIP(hipMemcpyAsync(dorder.get(), porder.data(), sizeof(int) * porder.size(), hipMemcpyDefault, c.main));
    k::batch_info_reset(c.main, info.get(), count);
    auto *gdesc = reinterpret_cast<const k::BatchGP *>(desc.get());

    // the level path runs beside the persistent kernel on the panel stream
    auto whole = std::make_unique<Timed>(c, c.main, "batch");
    hipStream_t ls = c.panel;
    hipEvent_t ready = c.event(), done = c.event();
    GP_HIP(hipEventRecord(ready, c.main));
    GP_HIP(hipStreamWaitEvent(ls, ready, 0));
    std::vector<std::pair<std::shared_ptr<BatchGraph>, std::vector<int>>> replayed;  // graph and its GPs
    DevBuf<int2> items;
    long item_total = 0;
    for (auto &gr : groups)
        if (!gr.ids.empty())
        {
            gr.plan = level_plan(gr.ids, nb, gr.wide ? 128 : 64, !gr.mixed && !gr.wide, gr.mixed);
            gr.base = item_total;
            item_total += static_cast<long>(gr.plan.items.size());
        }
    {
        Timed t(c, ls, "batch_fit_level");
        for (int q = 0; q < 3; q++)
        {
            LevelGroup &gr = groups[q];
            if (gr.ids.empty())
                continue;
            // the composition: step kind, then each GP's size in item order
            std::string key = std::to_string(q) + ":" + std::to_string(count);
            for (int g : gr.ids)
                key += "," + std::to_string(g) + "/" + std::to_string(nb[g]);
            if (c.batch_seen.size() > 64)
                c.batch_seen.clear();
            const int seen = ++c.batch_seen[key];
            auto git = c.batch_graphs.find(key);
            if (git == c.batch_graphs.end() && seen >= 2)
            {
                // capture this composition's sequence into a graph of its own buffers
                auto bg = std::make_shared<BatchGraph>();
                bg->desc = DevBuf<double>(c, (sizeof(k::BatchGP) * count + 7) / 8);
                bg->items = DevBuf<int2>(c, gr.plan.items.size());
                bg->info = DevBuf<int>(c, count);
                bg->steps = DevBuf<int>(c, count);
                bg->losses = DevBuf<double>(c, count);
                GP_HIP(hipMemcpyAsync(bg->items.get(), gr.plan.items.data(), sizeof(int2) * gr.plan.items.size(), hipMemcpyDefault, ls));
                GP_HIP(hipStreamSynchronize(ls));
                GP_HIP(hipStreamBeginCapture(ls, hipStreamCaptureModeThreadLocal));
                k::batch_info_reset(ls, bg->info.get(), count);
                launch_level(ls, gr.plan, bg->items.get(), reinterpret_cast<const k::BatchGP *>(bg->desc.get()), gr.mixed, gr.wide,
                             bg->losses.get(), bg->info.get(), bg->steps.get());
                GP_HIP(hipStreamEndCapture(ls, &bg->graph));
                GP_HIP(hipGraphInstantiate(&bg->exec, bg->graph, nullptr, nullptr, 0));
                // at most eight graphs per context, least recently used out
                if (c.batch_graphs.size() >= 8)
                {
                    auto victim = c.batch_graphs.begin();
                    for (auto it = c.batch_graphs.begin(); it != c.batch_graphs.end(); ++it)
                        if (it->second->last_use < victim->second->last_use)
                            victim = it;
                    GP_HIP(hipStreamSynchronize(ls));
                    c.batch_graphs.erase(victim);
                }
                git = c.batch_graphs.emplace(key, bg).first;
            }
            if (git != c.batch_graphs.end())
            {
                BatchGraph &bg = *git->second;
                bg.last_use = ++c.batch_clock;
                GP_HIP(hipMemcpyAsync(bg.desc.get(), dgps.data(), sizeof(k::BatchGP) * count, hipMemcpyDefault, ls));
                GP_HIP(hipGraphLaunch(bg.exec, ls));
                replayed.emplace_back(git->second, gr.ids);
            }
            else
            {
                if (!items.get())
                    items = DevBuf<int2>(c, static_cast<std::size_t>(item_total));
                GP_HIP(hipMemcpyAsync(items.get() + gr.base, gr.plan.items.data(), sizeof(int2) * gr.plan.items.size(), hipMemcpyDefault,
                                      ls));
                launch_level(ls, gr.plan, items.get() + gr.base, gdesc, gr.mixed, gr.wide, losses.get(), info.get(), steps.get());
            }
        }
    }
    GP_HIP(hipEventRecord(done, ls));
    {
        Timed t(c, c.main, "batch_fit");
        hipDeviceProp_t prop;
        GP_HIP(hipGetDeviceProperties(&prop, c.device));
        long off = 0;
        for (int m = 0; m < 2; m++)
        {
            const int cnt = static_cast<int>(pers[m].size());
            if (cnt == 0)
                continue;
            const int workgroups = std::max(1, std::min(cnt, prop.multiProcessorCount));
            k::batch_fit(c.main, gdesc, dorder.get() + off, cnt, counter.get() + m, losses.get(), info.get(), workgroups, m == 1,
                         steps.get());
            off += cnt;
        }
    }
    GP_HIP(hipStreamWaitEvent(c.main, done, 0));
    whole.reset();
    std::vector<int> hinfo(count, INT_MAX), hsteps(count, 0);
    std::vector<double> hloss(count, 0.0);
    GP_HIP(hipMemcpyAsync(hinfo.data(), info.get(), sizeof(int) * count, hipMemcpyDefault, c.main));
    GP_HIP(hipMemcpyAsync(hloss.data(), losses.get(), sizeof(double) * count, hipMemcpyDefault, c.main));
    GP_HIP(hipMemcpyAsync(hsteps.data(), steps.get(), sizeof(int) * count, hipMemcpyDefault, c.main));
    std::vector<std::vector<int>> ginfo, gsteps;
    std::vector<std::vector<double>> gloss;
    for (auto &r : replayed)
    {
        ginfo.emplace_back(count);
        gsteps.emplace_back(count);
        gloss.emplace_back(count);
        GP_HIP(hipMemcpyAsync(ginfo.back().data(), r.first->info.get(), sizeof(int) * count, hipMemcpyDefault, c.main));
        GP_HIP(hipMemcpyAsync(gloss.back().data(), r.first->losses.get(), sizeof(double) * count, hipMemcpyDefault, c.main));
        GP_HIP(hipMemcpyAsync(gsteps.back().data(), r.first->steps.get(), sizeof(int) * count, hipMemcpyDefault, c.main));
    }
    c.sync();
    c.recycle(ready);
    c.recycle(done);
    for (std::size_t q = 0; q < replayed.size(); q++)
        for (int g : replayed[q].second)
        {
            hinfo[g] = ginfo[q][g];
            hloss[g] = gloss[q][g];
            hsteps[g] = gsteps[q][g];
        }
    for (int g = 0; g < count; g++)
    {
        losses_host[g] = hloss[g];
        if (refine_host)
            refine_host[g] = routes[g].mixed ? hsteps[g] : 0;
    }
    if (mixed_out)
    {
        mixed_out->resize(count);
        for (int g = 0; g < count; g++)
            (*mixed_out)[g] = routes[g].mixed ? 1 : 0;
    }
    return hinfo;
}

}  // namespace gpamd
