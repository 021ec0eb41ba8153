// planner.cpp -- host-side planning code for the bucket-elimination engine:
// elimination ordering and the plan compiler's interval allocator.
//
// Min-degree greedy ordering of an interaction (line) graph, the rule of
// qtensor/utils.py:12-44 (get_neighbours_peo_vars) used by
// qtensor/optimisation/Optimizer.py:46-88: repeatedly eliminate the node with
// the fewest distinct neighbours -- the first such node in graph order -- and
// join its neighbours into a clique.  The width of a step is the neighbour
// count at elimination.  The same elimination with a fixed order gives the
// widths of that order.  Plain host code: no device work.

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/qtensor_amd.h"
#include "internal.h"

namespace {

struct Graph {
    std::vector<std::vector<int32_t>> nb;   // distinct neighbours, no self loops
    std::vector<uint32_t> mark;
    uint32_t stamp = 0;

    bool load(int32_t n, const int32_t* ptr, const int32_t* idx) {
        nb.assign(n, {});
        mark.assign(n, 0u);
        for (int32_t v = 0; v < n; ++v) {
            if (ptr[v + 1] < ptr[v]) return false;
            ++stamp;
            mark[v] = stamp;      // drop self loops
            for (int32_t k = ptr[v]; k < ptr[v + 1]; ++k) {
                const int32_t u = idx[k];
                if (u < 0 || u >= n) return false;
                if (mark[u] != stamp) {
                    mark[u] = stamp;
                    nb[v].push_back(u);
                }
            }
        }
        // symmetric closure, as an undirected graph
        for (int32_t v = 0; v < n; ++v)
            for (int32_t u : nb[v]) {
                bool found = false;
                for (int32_t w : nb[u])
                    if (w == v) { found = true; break; }
                if (!found) nb[u].push_back(v);
            }
        return true;
    }

    // Remove v: its neighbours lose v and become a clique.  Returns its width.
    int32_t eliminate(int32_t v) {
        std::vector<int32_t> nv;
        return eliminate(v, nv);
    }

    // The same, leaving v's former neighbours in `nv`.
    int32_t eliminate(int32_t v, std::vector<int32_t>& nv) {
        nv.clear();
        nv.swap(nb[v]);
        for (int32_t u : nv) {
            std::vector<int32_t>& s = nb[u];
            ++stamp;
            size_t w = 0;
            for (size_t k = 0; k < s.size(); ++k) {
                if (s[k] == v) continue;
                mark[s[k]] = stamp;
                s[w++] = s[k];
            }
            s.resize(w);
            mark[u] = stamp;
            for (int32_t x : nv)
                if (mark[x] != stamp) {
                    mark[x] = stamp;
                    s.push_back(x);
                }
        }
        return (int32_t)nv.size();
    }
};

// Live nodes in node order (a doubly linked list) with their neighbour counts,
// for the elimination loops: each step scans only the nodes still present.
struct LiveList {
    std::vector<int32_t> next, prev, deg;
    int32_t head = -1;

    void reset(const Graph& g) {
        const int32_t n = (int32_t)g.nb.size();
        next.resize(n);
        prev.resize(n);
        deg.resize(n);
        for (int32_t v = 0; v < n; ++v) {
            next[v] = v + 1 < n ? v + 1 : -1;
            prev[v] = v - 1;
            deg[v] = (int32_t)g.nb[v].size();
        }
        head = n > 0 ? 0 : -1;
    }
    void remove(int32_t v) {
        if (prev[v] >= 0) next[prev[v]] = next[v]; else head = next[v];
        if (next[v] >= 0) prev[next[v]] = prev[v];
    }
    // first live node with the fewest neighbours
    int32_t first_min(int32_t& lo) const {
        int32_t best = -1;
        lo = INT32_MAX;
        for (int32_t v = head; v >= 0; v = next[v])
            if (deg[v] < lo) { lo = deg[v]; best = v; }
        return best;
    }
    // eliminate v from g and the list; returns its width
    int32_t eliminate(Graph& g, int32_t v, std::vector<int32_t>& nv) {
        remove(v);
        const int32_t w = g.eliminate(v, nv);
        for (int32_t u : nv) deg[u] = (int32_t)g.nb[u].size();
        return w;
    }
};

}  // namespace

extern "C" {

int qb_greedy_order(int32_t n_nodes, const int32_t* adj_ptr, const int32_t* adj_idx, int32_t* order,
                    int32_t* widths) {
    if (n_nodes < 0 || (n_nodes > 0 && (!adj_ptr || !adj_idx || !order || !widths)))
        return qb_internal::fail(QB_ERR_INVALID, "qb_greedy_order: bad node count or null pointer");
    Graph g;
    if (!g.load(n_nodes, adj_ptr, adj_idx))
        return qb_internal::fail(QB_ERR_INVALID, "qb_greedy_order: malformed adjacency");
    LiveList live;
    live.reset(g);
    std::vector<int32_t> nv;
    for (int32_t step = 0; step < n_nodes; ++step) {
        int32_t lo;
        const int32_t best = live.first_min(lo);
        order[step] = best;
        widths[step] = live.eliminate(g, best, nv);
    }
    return QB_OK;
}

// Randomized greedy orderings, the rule of qtensor/optimisation/RGreedy.py:15-77
// (RGreedyOptimizer): at every step pick a remaining node with probability
// proportional to exp(-(neighbours - fewest neighbours) / temp).  Sample 0 is
// the min-degree greedy order itself; samples 1..repeats draw from a
// splitmix64 stream seeded by `seed`.  The kept sample is the one with the
// smallest (largest width, sum of 2^width), the earliest on ties: the
// reference keeps the smallest largest width; the sum breaks its ties by the
// total size of the step results.
int qb_rgreedy_order(int32_t n_nodes, const int32_t* adj_ptr, const int32_t* adj_idx, double temp,
                     uint64_t seed, int32_t repeats, int32_t* order, int32_t* widths) {
    if (n_nodes < 0 || repeats < 0 || !(temp > 0.0) || (n_nodes > 0 && (!adj_ptr || !adj_idx || !order || !widths)))
        return qb_internal::fail(QB_ERR_INVALID, "qb_rgreedy_order: bad node count, repeats, temp or null pointer");
    Graph g0;
    if (!g0.load(n_nodes, adj_ptr, adj_idx))
        return qb_internal::fail(QB_ERR_INVALID, "qb_rgreedy_order: malformed adjacency");
    uint64_t state = seed;
    auto uniform = [&state]() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    std::vector<int32_t> cur_order(n_nodes), cur_width(n_nodes);
    // exp(-d / temp) for every neighbour-count gap d (a node never has more
    // than n_nodes - 1 neighbours): the same values, computed once
    std::vector<double> boltzmann((size_t)n_nodes + 1);
    for (int32_t d = 0; d <= n_nodes; ++d) boltzmann[d] = std::exp(-(double)d / temp);
    int32_t best_max = -1;
    double best_sum = 0.0;
    Graph g = g0;
    LiveList live;
    std::vector<int32_t> nv;
    for (int32_t sample = 0; sample <= repeats; ++sample) {
        if (sample > 0)
            for (int32_t v = 0; v < n_nodes; ++v) g.nb[v].assign(g0.nb[v].begin(), g0.nb[v].end());
        live.reset(g);
        int32_t w_max = 0;
        double w_sum = 0.0;
        for (int32_t step = 0; step < n_nodes; ++step) {
            int32_t lo;
            int32_t pick = live.first_min(lo);
            if (sample > 0) {
                // the Boltzmann weights of the live nodes, in node order (removed
                // nodes weigh 0 and change no partial sum)
                double total = 0.0;
                for (int32_t v = live.head; v >= 0; v = live.next[v]) total += boltzmann[live.deg[v] - lo];
                const double r = uniform() * total;
                double acc = 0.0;
                for (int32_t v = live.head; v >= 0; v = live.next[v]) {
                    acc += boltzmann[live.deg[v] - lo];
                    pick = v;
                    if (acc > r) break;
                }
            }
            cur_order[step] = pick;
            cur_width[step] = live.eliminate(g, pick, nv);
            w_max = std::max(w_max, cur_width[step]);
            w_sum += std::ldexp(1.0, cur_width[step]);
        }
        if (best_max < 0 || w_max < best_max || (w_max == best_max && w_sum < best_sum)) {
            best_max = w_max;
            best_sum = w_sum;
            std::copy(cur_order.begin(), cur_order.end(), order);
            std::copy(cur_width.begin(), cur_width.end(), widths);
        }
    }
    return QB_OK;
}

// Offline interval allocation of a plan's buffers (the plan compiler's own
// layout step, no counterpart in the reference): buffer i lives over ops
// [starts[i], ends[i]] and needs sizes[i] elements; buffers are placed in
// `order`, each at the lowest offset >= base that is free over its whole
// interval (first fit over the placed buffers whose intervals overlap, sorted
// by offset, ties in placement order).
int qb_place_intervals(int32_t m, const int64_t* sizes, const int64_t* starts, const int64_t* ends,
                       const int64_t* order, int64_t base, int64_t* offsets, int64_t* peak) {
    if (m < 0 || (m > 0 && (!sizes || !starts || !ends || !order || !offsets)) || !peak)
        return qb_internal::fail(QB_ERR_INVALID, "qb_place_intervals: bad count or null pointer");
    std::vector<char> seen(m, 0);
    for (int32_t k = 0; k < m; ++k) {
        const int64_t i = order[k];
        if (i < 0 || i >= m || seen[i] || sizes[i] < 0 || ends[i] < starts[i])
            return qb_internal::fail(QB_ERR_INVALID, "qb_place_intervals: bad order, size or interval");
        seen[i] = 1;
    }
    struct Placed { int64_t off, end, s, e; };
    std::vector<Placed> placed;
    placed.reserve(m);
    std::vector<int32_t> live;
    int64_t top = base;
    for (int32_t k = 0; k < m; ++k) {
        const int64_t i = order[k];
        const int64_t n = sizes[i], s = starts[i], e = ends[i];
        live.clear();
        for (int32_t q = 0; q < k; ++q)
            if (placed[q].s <= e && s <= placed[q].e) live.push_back(q);
        std::stable_sort(live.begin(), live.end(),
                         [&](int32_t a, int32_t b) { return placed[a].off < placed[b].off; });
        int64_t off = base;
        for (int32_t q : live) {
            if (off + n <= placed[q].off) break;
            if (placed[q].end > off) off = placed[q].end;
        }
        offsets[i] = off;
        placed.push_back({off, off + n, s, e});
        top = std::max(top, off + n);
    }
    *peak = top;
    return QB_OK;
}

int qb_elimination_widths(int32_t n_nodes, const int32_t* adj_ptr, const int32_t* adj_idx, int32_t n_order,
                          const int32_t* order, int32_t* widths) {
    if (n_nodes < 0 || n_order < 0 || n_order > n_nodes || (n_nodes > 0 && (!adj_ptr || !adj_idx)) ||
        (n_order > 0 && (!order || !widths)))
        return qb_internal::fail(QB_ERR_INVALID, "qb_elimination_widths: bad size or null pointer");
    Graph g;
    if (!g.load(n_nodes, adj_ptr, adj_idx))
        return qb_internal::fail(QB_ERR_INVALID, "qb_elimination_widths: malformed adjacency");
    std::vector<char> alive(n_nodes, 1);
    for (int32_t k = 0; k < n_order; ++k) {
        const int32_t v = order[k];
        if (v < 0 || v >= n_nodes || !alive[v])
            return qb_internal::fail(QB_ERR_INVALID, "qb_elimination_widths: order repeats or leaves the graph");
        alive[v] = 0;
        widths[k] = g.eliminate(v);
    }
    return QB_OK;
}

}  // extern "C"
