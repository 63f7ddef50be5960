// Seeded synthetic graph generators for the benchmark workloads: 2D grid,
// 2D random geometric graph and R-MAT.  All randomness comes from xoshiro256**
// seeded by splitmix64, so every graph is a pure function of its parameters.
// Compiled with -ffp-contract=off so the RGG distance test is the same IEEE
// expression everywhere.
#include "host_graph.hpp"

#include <algorithm>
#include <cmath>

namespace heipa {
namespace {

// Builds a symmetric CSR from per-row neighbour lists already containing both
// directions; rows are sorted ascending and deduplicated.
void finish_rows(HostGraph &g, std::vector<u32> &deg_or_off, std::vector<u32> &nbr) {
    // deg_or_off holds exclusive offsets (n+1), nbr holds the entries
    g.xadj.assign(g.n + 1, 0);
    u32 out = 0;
    for (u32 u = 0; u < g.n; ++u) {
        u32 b = deg_or_off[u], e = deg_or_off[u + 1];
        std::sort(nbr.begin() + b, nbr.begin() + e);
        u32 start = out;
        for (u32 j = b; j < e; ++j) {
            if (j > b && nbr[j] == nbr[j - 1]) continue;
            nbr[out++] = nbr[j];
        }
        (void) start;
        g.xadj[u + 1] = out;
    }
    nbr.resize(out);
    g.adjncy = std::move(nbr);
    g.m = out;
}

void add_weights(HostGraph &g, u64 seed, u32 wmod) {
    if (wmod == 0) {
        g.uniform_edge_weights = true;
        g.adjwgt.clear();
        return;
    }
    g.uniform_edge_weights = false;
    g.adjwgt.resize(g.m);
    for (u32 u = 0; u < g.n; ++u)
        for (u32 j = g.xadj[u]; j < g.xadj[u + 1]; ++j)
            g.adjwgt[j] = edge_weight_hash(seed, u, g.adjncy[j], wmod);
}

} // namespace

HostGraph gen_grid(u32 rows, u32 cols) {
    HostGraph g;
    g.n = rows * cols;
    g.xadj.assign(g.n + 1, 0);
    g.adjncy.reserve((size_t) 4 * g.n);
    for (u32 r = 0; r < rows; ++r) {
        for (u32 c = 0; c < cols; ++c) {
            const u32 id = r * cols + c;
            if (r > 0) g.adjncy.push_back(id - cols);
            if (c > 0) g.adjncy.push_back(id - 1);
            if (c + 1 < cols) g.adjncy.push_back(id + 1);
            if (r + 1 < rows) g.adjncy.push_back(id + cols);
            g.xadj[id + 1] = (u32) g.adjncy.size();
        }
    }
    g.m = (u32) g.adjncy.size();
    g.uniform_edge_weights = true;
    g.uniform_vertex_weights = true;
    g.g_weight = g.n;
    return g;
}

HostGraph gen_rgg(u32 n, double avg_deg, u64 seed, u32 wmod) {
    HostGraph g;
    g.n = n;
    std::vector<double> x(n), y(n);
    Xoshiro256 rng(seed);
    for (u32 i = 0; i < n; ++i) {
        x[i] = rng.uniform();
        y[i] = rng.uniform();
    }
    const double r = std::sqrt(avg_deg / (M_PI * (double) n));
    const double r2 = r * r;
    u32 cells = (u32) std::max(1.0, std::floor(1.0 / r));
    if (cells > 65536) cells = 65536;
    auto cell_of = [&](double v) {
        u32 c = (u32) (v * (double) cells);
        return c >= cells ? cells - 1 : c;
    };
    // counting sort of points by cell
    const u64 ncell = (u64) cells * cells;
    std::vector<u32> cstart(ncell + 1, 0), order(n);
    std::vector<u32> pc(n);
    for (u32 i = 0; i < n; ++i) {
        pc[i] = cell_of(y[i]) * cells + cell_of(x[i]);
        cstart[pc[i] + 1]++;
    }
    for (u64 c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
    {
        std::vector<u32> fill(cstart.begin(), cstart.end() - 1);
        for (u32 i = 0; i < n; ++i) order[fill[pc[i]]++] = i;
    }
    // two passes: count then fill
    std::vector<u32> deg(n + 1, 0);
    std::vector<u32> nbr;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            for (u32 i = 0; i < n; ++i) deg[i + 1] += deg[i];
            nbr.resize(deg[n]);
        }
        std::vector<u32> pos;
        if (pass == 1) pos.assign(deg.begin(), deg.end() - 1);
        for (u32 i = 0; i < n; ++i) {
            const u32 cx = cell_of(x[i]), cy = cell_of(y[i]);
            for (int dy = -1; dy <= 1; ++dy) {
                const s64 yy = (s64) cy + dy;
                if (yy < 0 || yy >= (s64) cells) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    const s64 xx = (s64) cx + dx;
                    if (xx < 0 || xx >= (s64) cells) continue;
                    const u64 c = (u64) yy * cells + (u64) xx;
                    for (u32 t = cstart[c]; t < cstart[c + 1]; ++t) {
                        const u32 j = order[t];
                        if (j == i) continue;
                        const double ddx = x[i] - x[j], ddy = y[i] - y[j];
                        if (ddx * ddx + ddy * ddy <= r2) {
                            if (pass == 0) deg[i + 1]++;
                            else nbr[pos[i]++] = j;
                        }
                    }
                }
            }
        }
    }
    finish_rows(g, deg, nbr);
    add_weights(g, seed, wmod);
    g.uniform_vertex_weights = true;
    g.g_weight = n;
    return g;
}

HostGraph gen_rmat(u32 scale, u32 n, double edge_factor, double a, double b, double c,
                   u64 seed, u32 wmod, bool explicit_vw) {
    HostGraph g;
    const u64 full = (u64) 1 << scale;
    if (n == 0 || n > full) n = (u32) full;
    g.n = n;
    const u64 samples = (u64) (edge_factor * (double) full);
    Xoshiro256 rng(seed);
    const double ab = a + b, abc = a + b + c;
    std::vector<u32> src, dst;
    src.reserve(samples);
    dst.reserve(samples);
    for (u64 s = 0; s < samples; ++s) {
        u64 u = 0, v = 0;
        for (u32 l = 0; l < scale; ++l) {
            const double r = rng.uniform();
            u64 bu, bv;
            if (r < a) { bu = 0; bv = 0; }
            else if (r < ab) { bu = 0; bv = 1; }
            else if (r < abc) { bu = 1; bv = 0; }
            else { bu = 1; bv = 1; }
            u = 2 * u + bu;
            v = 2 * v + bv;
        }
        if (u == v || u >= n || v >= n) continue;
        src.push_back((u32) u);
        dst.push_back((u32) v);
    }
    std::vector<u32> off(n + 1, 0);
    for (size_t i = 0; i < src.size(); ++i) {
        off[src[i] + 1]++;
        off[dst[i] + 1]++;
    }
    for (u32 i = 0; i < n; ++i) off[i + 1] += off[i];
    std::vector<u32> nbr(off[n]);
    {
        std::vector<u32> pos(off.begin(), off.end() - 1);
        for (size_t i = 0; i < src.size(); ++i) {
            nbr[pos[src[i]]++] = dst[i];
            nbr[pos[dst[i]]++] = src[i];
        }
    }
    src.clear();
    src.shrink_to_fit();
    dst.clear();
    dst.shrink_to_fit();
    finish_rows(g, off, nbr);
    add_weights(g, seed, wmod);
    if (explicit_vw) {
        g.uniform_vertex_weights = false;
        g.vwgt.assign(n, 1);
    } else {
        g.uniform_vertex_weights = true;
    }
    g.g_weight = n;
    return g;
}

} // namespace heipa
