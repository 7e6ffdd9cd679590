// Host check of the diagonal-block task schedule (gprat_amd/csrc/panel_plan.cpp).
//
// Runs every plan on the CPU with blocks of B x B doubles instead of 64 x 64,
// with the tasks' block arithmetic restated in plain loops, and launch
// semantics made strict: all tasks of a launch read the state left by the
// previous launch, and any element written by one task and read or written
// by another task of the same launch is reported. The result must be the
// Cholesky factor (against an unblocked factorisation of the same matrix),
// the row-major inverses of its diagonal blocks and the whole inverse with
// zeros above the diagonal, every element of it written.
//
// usage: panel_plan_check  (exit status 0 when every case passes)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "panel_plan.hpp"

using namespace gpamd;

namespace
{

constexpr int B = 4;  // block edge of the check (64 on the device)

struct State
{
    std::vector<double> buf[4];  // 0: A (row-major n x n), 1: Winv, 2: Wb (k-major), 3: tmp
};

struct Exec
{
    int n, nb;
    State old, cur;
    std::map<long, int> writer;  // (buffer, index) -> task id within the launch
    std::vector<std::vector<long>> reads;
    int task = 0;

    static long id(int b, long i) { return (static_cast<long>(b) << 40) | i; }
    double rd(int b, long i)
    {
        reads[task].push_back(id(b, i));
        return old.buf[b].at(i);
    }
    void wr(int b, long i, double v)
    {
        auto it = writer.find(id(b, i));
        if (it != writer.end() && it->second != task)
            throw std::runtime_error("two tasks of one launch write the same element");
        writer[id(b, i)] = task;
        cur.buf[b].at(i) = v;
    }
    // block accessors (block coordinates in units of B)
    double a(int bi, int bj, int r, int c) { return rd(0, static_cast<long>(bi * B + r) * n + bj * B + c); }
    void set_a(int bi, int bj, int r, int c, double v) { wr(0, static_cast<long>(bi * B + r) * n + bj * B + c, v); }
    double winv(int k, int r, int c) { return rd(1, static_cast<long>(k) * B * B + r * B + c); }
    void set_winv(int k, int r, int c, double v) { wr(1, static_cast<long>(k) * B * B + r * B + c, v); }
    double w(int bi, int bj, int r, int c) { return rd(2, static_cast<long>(bj * B + c) * n + bi * B + r); }
    void set_w(int bi, int bj, int r, int c, double v) { wr(2, static_cast<long>(bj * B + c) * n + bi * B + r, v); }

    void factor(int k, const std::vector<double> &C)
    {
        std::vector<double> L(B * B, 0.0), W(B * B, 0.0);
        for (int j = 0; j < B; j++)
        {
            double d = C[j * B + j];
            for (int t = 0; t < j; t++)
                d -= L[j * B + t] * L[j * B + t];
            if (!(d > 0))
                throw std::runtime_error("non-positive pivot");
            L[j * B + j] = std::sqrt(d);
            for (int i = j + 1; i < B; i++)
            {
                double s = C[i * B + j];
                for (int t = 0; t < j; t++)
                    s -= L[i * B + t] * L[j * B + t];
                L[i * B + j] = s / L[j * B + j];
            }
        }
        for (int c = 0; c < B; c++)
            for (int r = c; r < B; r++)
            {
                double s = (r == c) ? 1.0 : 0.0;
                for (int t = c; t < r; t++)
                    s -= L[r * B + t] * W[t * B + c];
                W[r * B + c] = s / L[r * B + r];
            }
        for (int r = 0; r < B; r++)
            for (int c = 0; c < B; c++)
            {
                if (c <= r)
                    set_a(k, k, r, c, L[r * B + c]);
                set_winv(k, r, c, c <= r ? W[r * B + c] : 0.0);
            }
    }

    void run(const PanelTask &t, long tmp_stride_unused)
    {
        (void)tmp_stride_unused;
        const int type = t.x & 0xff, h = t.x >> 8;
        switch (type)
        {
        case PT_FACTOR:
        case PT_UFACTOR:
        {
            const int k = t.y;
            std::vector<double> C(B * B, 0.0);
            for (int r = 0; r < B; r++)
                for (int c = 0; c <= r; c++)
                {
                    double v = a(k, k, r, c);
                    if (type == PT_UFACTOR)
                        for (int s = 0; s < B; s++)
                            v -= a(k, t.z, r, s) * a(k, t.z, c, s);
                    C[r * B + c] = v;
                    C[c * B + r] = v;
                }
            factor(k, C);
            break;
        }
        case PT_SOLVE:
        {
            double X[B][B];
            for (int r = 0; r < B; r++)
                for (int c = 0; c < B; c++)
                {
                    double s = 0;
                    for (int q = 0; q < B; q++)
                        s += a(t.y, t.z, r, q) * winv(t.z, c, q);
                    X[r][c] = s;
                }
            for (int r = 0; r < B; r++)
                for (int c = 0; c < B; c++)
                    set_a(t.y, t.z, r, c, X[r][c]);
            break;
        }
        case PT_UPDATE:
        {
            const int i = t.y, j = t.z, k = t.w;
            for (int r = 0; r < B; r++)
                for (int c = 0; c < B; c++)
                {
                    if (i == j && c > r)
                        continue;
                    double s = a(i, j, r, c);
                    for (int q = 0; q < B; q++)
                        s -= a(i, k, r, q) * a(j, k, c, q);
                    set_a(i, j, r, c, s);
                }
            break;
        }
        case PT_PLACE:
        {
            const int k = t.y;
            for (int r = 0; r < B; r++)
                for (int c = 0; c < B; c++)
                    set_w(k, k, r, c, winv(k, r, c));
            for (int j = k + 1; j < nb; j++)
                for (int r = 0; r < B; r++)
                    for (int c = 0; c < B; c++)
                        set_w(k, j, r, c, 0.0);
            break;
        }
        case PT_JOINF:
        {
            const int aa = t.y, bb = aa + 1;
            double T[B][B];
            for (int r = 0; r < B; r++)
                for (int c = 0; c < B; c++)
                {
                    double s = 0;
                    for (int q = 0; q < B; q++)
                        s += a(bb, aa, r, q) * winv(aa, q, c);
                    T[r][c] = s;
                }
            for (int r = 0; r < B; r++)
                for (int c = 0; c < B; c++)
                {
                    double s = 0;
                    for (int q = 0; q < B; q++)
                        s += winv(bb, r, q) * T[q][c];
                    set_w(bb, aa, r, c, -s);
                }
            break;
        }
        case PT_JOIN0:
        case PT_JOIN1:
        {
            const int p = t.y, rb = t.z & 0xffff, cb = t.z >> 16;
            const int a0 = 2 * h * p, b0 = a0 + h;
            const long size = static_cast<long>(B) * h;
            const long base = static_cast<long>(t.w) * B * B;
            auto tix = [&](int r, int c) { return base + static_cast<long>(c) * size + r; };  // pair-matrix coordinates
            if (type == PT_JOIN0)
            {
                for (int r = 0; r < B; r++)
                    for (int c = 0; c < B; c++)
                    {
                        double s = 0;
                        for (int sb = cb; sb < h; sb++)
                            for (int q = 0; q < B; q++)
                                s += a(b0 + rb, a0 + sb, r, q) * w(a0 + sb, a0 + cb, q, c);
                        wr(3, tix(rb * B + r, cb * B + c), s);
                    }
            }
            else
            {
                for (int r = 0; r < B; r++)
                    for (int c = 0; c < B; c++)
                    {
                        double s = 0;
                        for (int sb = 0; sb <= rb; sb++)
                            for (int q = 0; q < B; q++)
                                s += w(b0 + rb, b0 + sb, r, q) * rd(3, tix(sb * B + q, cb * B + c));
                        set_w(b0 + rb, a0 + cb, r, c, -s);
                    }
            }
            break;
        }
        default:
            throw std::runtime_error("unknown task type");
        }
    }
};

unsigned long long rng = 88172645463325252ull;
double uniform()
{
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (rng >> 11) * (1.0 / 9007199254740992.0);
}

void check(int nb, int chain_wgs, int tail_wgs)
{
    const PanelPlan plan = build_panel_plan(nb, chain_wgs, tail_wgs);
    const int n = nb * B;
    // SPD matrix: M M^T / n + I
    std::vector<double> M(static_cast<std::size_t>(n) * n), A0(static_cast<std::size_t>(n) * n);
    for (auto &v : M)
        v = uniform() - 0.5;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
        {
            double s = (i == j) ? 1.0 : 0.0;
            for (int t = 0; t < n; t++)
                s += M[i * n + t] * M[j * n + t] / n;
            A0[i * n + j] = s;
        }
    Exec e{ n, nb, {}, {}, {}, {}, 0 };
    e.old.buf[0] = A0;
    e.old.buf[1].assign(static_cast<std::size_t>(nb) * B * B, NAN);
    e.old.buf[2].assign(static_cast<std::size_t>(n) * n, NAN);
    e.old.buf[3].assign(static_cast<std::size_t>(plan.tmp_doubles / 4096 * B * B) + 1, NAN);
    int tasks = 0;
    for (const auto &l : plan.launches)
    {
        if (l.groups < 1 || l.groups > std::max(chain_wgs, tail_wgs))
            throw std::runtime_error("launch with a bad workgroup count");
        e.cur = e.old;
        e.writer.clear();
        std::vector<const PanelTask *> list;
        for (int g = 0; g < l.groups; g++)
            for (int t = plan.wg_off[l.first + g]; t < plan.wg_off[l.first + g + 1]; t++)
                list.push_back(&plan.tasks[t]);
        e.reads.assign(list.size(), {});
        for (std::size_t t = 0; t < list.size(); t++)
        {
            e.task = static_cast<int>(t);
            e.run(*list[t], 0);
        }
        for (std::size_t t = 0; t < list.size(); t++)
            for (long r : e.reads[t])
            {
                auto it = e.writer.find(r);
                if (it != e.writer.end() && it->second != static_cast<int>(t))
                    throw std::runtime_error("a task reads an element another task of its launch writes");
            }
        e.old = e.cur;
        tasks += static_cast<int>(list.size());
    }
    // reference: unblocked Cholesky
    std::vector<double> L(static_cast<std::size_t>(n) * n, 0.0);
    for (int j = 0; j < n; j++)
    {
        double d = A0[j * n + j];
        for (int t = 0; t < j; t++)
            d -= L[j * n + t] * L[j * n + t];
        L[j * n + j] = std::sqrt(d);
        for (int i = j + 1; i < n; i++)
        {
            double s = A0[i * n + j];
            for (int t = 0; t < j; t++)
                s -= L[i * n + t] * L[j * n + t];
            L[i * n + j] = s / L[j * n + j];
        }
    }
    const auto &A = e.old.buf[0], &Wb = e.old.buf[2], &Wi = e.old.buf[1];
    double err_l = 0, err_w = 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j <= i; j++)
            err_l = std::max(err_l, std::fabs(A[i * n + j] - L[i * n + j]));
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++)
        {
            const double wv = Wb[static_cast<long>(c) * n + r];
            if (std::isnan(wv))
                throw std::runtime_error("inverse element never written");
            if (c > r && wv != 0.0)
                throw std::runtime_error("inverse not zero above the diagonal");
            double s = 0;  // (L W)(r, c)
            for (int t = 0; t < n; t++)
                s += L[r * n + t] * Wb[static_cast<long>(c) * n + t];
            err_w = std::max(err_w, std::fabs(s - (r == c ? 1.0 : 0.0)));
        }
    for (int k = 0; k < nb; k++)
        for (int r = 0; r < B; r++)
            for (int c = 0; c < B; c++)
            {
                const double v = Wi[static_cast<long>(k) * B * B + r * B + c];
                const double ref = Wb[static_cast<long>(k * B + c) * n + k * B + r];
                if (std::isnan(v) || std::fabs(v - ref) > 1e-12)
                    throw std::runtime_error("diagonal-block inverse differs from the whole inverse");
            }
    if (err_l > 1e-12 || err_w > 1e-10)
        throw std::runtime_error("wrong result: factor err " + std::to_string(err_l) + ", inverse err " + std::to_string(err_w));
    std::printf("ok nb=%d chain_wgs=%d tail_wgs=%d launches=%zu tasks=%d tmp=%ld\n", nb, chain_wgs, tail_wgs, plan.launches.size(), tasks,
                plan.tmp_doubles);
}

}  // namespace

int main()
{
    const int caps[][2] = { { 1, 1 }, { 2, 3 }, { 16, 64 }, { 4096, 4096 } };
    try
    {
        for (int nb = 1; nb <= 16; nb++)
            for (const auto &c : caps)
                check(nb, c[0], c[1]);
        check(24, 16, 64);
        check(32, 16, 64);
    }
    catch (const std::exception &ex)
    {
        std::printf("FAILED: %s\n", ex.what());
        return 1;
    }
    return 0;
}
