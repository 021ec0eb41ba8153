// lightcone_planner.cpp -- the batched engine's host planner in native code.
//
// One call turns a graph's edges into canonical lightcone classes; one call per
// class compiles its executor plan.  The steps, each written to give exactly
// the result of the engine's Python planner (qtensor_amd/lightcone.py,
// composer.py, circuit.py, ordering.py, plan.py), which the tests compare word
// for word:
//
//   canonical lightcone  radius-p ball around the edge, colour refinement seeded
//                        with (endpoint, distance), individualisation of the
//                        smallest ambiguous class's first node
//                        (the lightcone of qtensor/utils.py:176-189)
//   circuit              QAOA cone ansatz U, Z_i Z_j, U^dagger on the canonical
//                        graph (qtensor/CircuitComposer.py:82-96,131-169): cost as
//                        cX.ZPhase.cX, as ZZ, or ZZ with commutation-based
//                        cancellation (qtensor/simplify_circuit/)
//   network              circ2buckets: bra wires closed by M tensors, gates in
//                        reverse time order, a new variable per changed qubit
//                        (qtensor/optimisation/TensorNet.py:60-68)
//   order and slices     line graph without bra/ket variables; min-degree greedy
//                        or randomized greedy (qtensor/utils.py:12-44,
//                        RGreedy.py:15-77); slices by GreedyParvars
//                        (Optimizer.py:90-159, Greedy.py:78-92) or tree trimming
//                        (Optimizer.py:178-228)
//   plan                 bucket elimination run symbolically
//                        (qtensor/Simulate.py:84-94, ProcessingFrameworks.py:13-37),
//                        chains of simple sums as reduction ops, list scheduling
//                        into gather and step ops, interval placement in LDS and
//                        scratch, and the executor's program encoding.
//
// Host code only: no device work.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/qtensor_amd.h"
#include "internal.h"

namespace {

// ---------------------------------------------------------------------------
// Canonical lightcones
// ---------------------------------------------------------------------------

struct Csr {
    int32_t n = 0;
    const int32_t* ptr = nullptr;
    const int32_t* idx = nullptr;
};

struct Lightcone {
    int32_t n = 0, i = 0, j = 0;
    std::vector<std::pair<int32_t, int32_t>> edges;   // canonical labels, sorted
};

// Scratch buffers of one thread's canonical labelling (reused across edges).
struct RefineScratch {
    std::vector<int32_t> sig, off, order, out, uniq;
};

// Colour refinement: new colour = rank of (colour, sorted neighbour colours)
// in lexicographic order; stops when the number of colours no longer grows and
// returns that ranking.  Signatures live in one flat buffer (node v's at
// off[v] .. off[v+1]).
std::vector<int32_t> refine(const std::vector<std::vector<int32_t>>& adj, std::vector<int32_t> color,
                            RefineScratch& rs) {
    const size_t n = color.size();
    rs.off.resize(n + 1);
    rs.off[0] = 0;
    for (size_t v = 0; v < n; ++v) rs.off[v + 1] = rs.off[v] + 1 + (int32_t)adj[v].size();
    rs.sig.resize(rs.off[n]);
    rs.order.resize(n);
    rs.out.resize(n);
    const int32_t* sig = nullptr;
    const int32_t* off = rs.off.data();
    auto less = [&](int32_t a, int32_t b) {
        return std::lexicographical_compare(sig + off[a], sig + off[a + 1], sig + off[b], sig + off[b + 1]);
    };
    auto equal = [&](int32_t a, int32_t b) {
        return off[a + 1] - off[a] == off[b + 1] - off[b] &&
               std::equal(sig + off[a], sig + off[a + 1], sig + off[b]);
    };
    for (;;) {
        for (size_t v = 0; v < n; ++v) {
            int32_t* s = rs.sig.data() + rs.off[v];
            s[0] = color[v];
            int32_t k = 1;
            for (int32_t w : adj[v]) s[k++] = color[w];
            std::sort(s + 1, s + k);
        }
        sig = rs.sig.data();
        for (size_t v = 0; v < n; ++v) rs.order[v] = (int32_t)v;
        std::sort(rs.order.begin(), rs.order.end(), less);
        int32_t rank = -1;
        for (size_t k = 0; k < n; ++k) {
            if (k == 0 || !equal(rs.order[k], rs.order[k - 1])) ++rank;
            rs.out[rs.order[k]] = rank;
        }
        rs.uniq.assign(color.begin(), color.end());
        std::sort(rs.uniq.begin(), rs.uniq.end());
        const size_t before = std::unique(rs.uniq.begin(), rs.uniq.end()) - rs.uniq.begin();
        if ((size_t)(rank + 1) == before) return rs.out;
        color.assign(rs.out.begin(), rs.out.end());
    }
}

std::vector<int32_t> rank_values(const std::vector<int64_t>& vals) {
    std::vector<int64_t> u(vals);
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    std::vector<int32_t> out(vals.size());
    for (size_t k = 0; k < vals.size(); ++k)
        out[k] = (int32_t)(std::lower_bound(u.begin(), u.end(), vals[k]) - u.begin());
    return out;
}

// Nodes within `dist` hops of {u, v}, ascending.
void ball(const Csr& g, int32_t u, int32_t v, int32_t dist, std::vector<int32_t>& stamp, int32_t& tick,
          std::vector<int32_t>& out) {
    ++tick;
    out.clear();
    std::vector<int32_t> frontier;
    for (int32_t s : {u, v})
        if (stamp[s] != tick) {
            stamp[s] = tick;
            out.push_back(s);
            frontier.push_back(s);
        }
    for (int32_t d = 0; d < dist; ++d) {
        std::vector<int32_t> nxt;
        for (int32_t x : frontier)
            for (int32_t k = g.ptr[x]; k < g.ptr[x + 1]; ++k) {
                const int32_t y = g.idx[k];
                if (stamp[y] != tick) {
                    stamp[y] = tick;
                    out.push_back(y);
                    nxt.push_back(y);
                }
            }
        frontier.swap(nxt);
    }
    std::sort(out.begin(), out.end());
}

Lightcone canonical_lightcone(const Csr& g, int32_t u, int32_t v, int32_t p, std::vector<int32_t>& stamp,
                              int32_t& tick, std::vector<int32_t>& local, RefineScratch& rs) {
    std::vector<int32_t> nodes;
    ball(g, u, v, p, stamp, tick, nodes);
    const int32_t n = (int32_t)nodes.size();
    for (int32_t k = 0; k < n; ++k) local[nodes[k]] = k;
    const int32_t in_tick = tick;   // stamp[x] == in_tick marks ball membership
    std::vector<std::vector<int32_t>> adj(n);
    for (int32_t k = 0; k < n; ++k) {
        const int32_t x = nodes[k];
        for (int32_t q = g.ptr[x]; q < g.ptr[x + 1]; ++q) {
            const int32_t y = g.idx[q];
            if (y != x && stamp[y] == in_tick) adj[k].push_back(local[y]);
        }
    }
    const int32_t lu = local[u], lv = local[v];
    std::vector<int32_t> dist(n, -1), dq;
    for (int32_t s : {lu, lv})
        if (dist[s] < 0) {
            dist[s] = 0;
            dq.push_back(s);
        }
    for (size_t h = 0; h < dq.size(); ++h) {
        const int32_t x = dq[h];
        for (int32_t y : adj[x])
            if (dist[y] < 0) {
                dist[y] = dist[x] + 1;
                dq.push_back(y);
            }
    }
    std::vector<int64_t> c0(n);
    for (int32_t k = 0; k < n; ++k) c0[k] = ((k == lu || k == lv) ? 0 : 1) * (int64_t)(1 << 20) + dist[k];
    std::vector<int32_t> color = refine(adj, rank_values(c0), rs);
    for (;;) {
        int32_t n_col = *std::max_element(color.begin(), color.end()) + 1;
        if (n_col >= n) break;
        std::vector<int32_t> size(n_col, 0), first(n_col, -1);
        for (int32_t k = 0; k < n; ++k) {
            if (first[color[k]] < 0) first[color[k]] = k;
            ++size[color[k]];
        }
        int32_t amb = -1;
        for (int32_t c = 0; c < n_col; ++c)
            if (size[c] > 1 && (amb < 0 || size[c] < size[amb])) amb = c;
        const int32_t pick = first[amb];
        std::vector<int64_t> c2(n);
        for (int32_t k = 0; k < n; ++k) c2[k] = 2 * (int64_t)color[k] + (k == pick ? 0 : 1);
        color = refine(adj, rank_values(c2), rs);
    }
    Lightcone lc;
    lc.n = n;
    for (int32_t a = 0; a < n; ++a)
        for (int32_t b : adj[a])
            if (color[a] < color[b]) lc.edges.emplace_back(color[a], color[b]);
    std::sort(lc.edges.begin(), lc.edges.end());
    lc.i = std::min(color[lu], color[lv]);
    lc.j = std::max(color[lu], color[lv]);
    return lc;
}

// ---------------------------------------------------------------------------
// Circuits
// ---------------------------------------------------------------------------

enum GateType : int32_t { G_H, G_Z, G_CX, G_M, G_ZPHASE, G_XPHASE, G_YPHASE, G_ZZ };
// device gate kinds (qtensor_amd.hip GateKind, gates.py / plan.py codes)
int32_t device_kind(int32_t t) {
    switch (t) {
    case G_ZPHASE: return 1;
    case G_XPHASE: return 2;
    case G_ZZ: return 3;
    case G_YPHASE: return 4;
    case G_H: return 5;
    case G_CX: return 6;
    case G_Z: return 7;
    default: return 8;   // M
    }
}
bool parametric(int32_t t) { return t == G_ZPHASE || t == G_XPHASE || t == G_YPHASE || t == G_ZZ; }

struct Gate {
    int32_t type = G_H;
    int32_t q[2] = {0, 0};
    int32_t nq = 1;
    std::vector<std::pair<int32_t, double>> terms;   // alpha = sum coeff * params[j]
};

// cone_ansatz on the canonical graph: H on every qubit, then for layer l the
// cost on the radius-(p-l) ball's edges and the mixer on its nodes.
void cone_ansatz(const Lightcone& lc, int32_t p, int32_t composer, std::vector<Gate>& out) {
    const int32_t n = lc.n;
    std::vector<int32_t> ptr(n + 1, 0), idx;
    std::vector<std::vector<int32_t>> adj(n);
    for (auto& e : lc.edges) {
        adj[e.first].push_back(e.second);
        adj[e.second].push_back(e.first);
    }
    for (auto& a : adj) std::sort(a.begin(), a.end());
    for (int32_t k = 0; k < n; ++k) {
        ptr[k + 1] = ptr[k] + (int32_t)adj[k].size();
        idx.insert(idx.end(), adj[k].begin(), adj[k].end());
    }
    Csr g{n, ptr.data(), idx.data()};
    for (int32_t q = 0; q < n; ++q) {
        Gate h;
        h.type = G_H;
        h.q[0] = q;
        out.push_back(h);
    }
    std::vector<int32_t> stamp(n, 0), nodes;
    int32_t tick = 0;
    std::vector<char> in(n);
    for (int32_t l = 0; l < p; ++l) {
        ball(g, lc.i, lc.j, p - l, stamp, tick, nodes);
        std::fill(in.begin(), in.end(), 0);
        for (int32_t x : nodes) in[x] = 1;
        for (int32_t a : nodes)
            for (int32_t b : adj[a]) {
                if (!in[b] || b <= a) continue;
                if (composer == 0) {
                    Gate c;
                    c.type = G_CX;
                    c.nq = 2;
                    c.q[0] = a;
                    c.q[1] = b;
                    out.push_back(c);
                    Gate z;
                    z.type = G_ZPHASE;
                    z.q[0] = b;
                    z.terms = {{l, 2.0}};
                    out.push_back(z);
                    out.push_back(c);
                } else {
                    Gate z;
                    z.type = G_ZZ;
                    z.nq = 2;
                    z.q[0] = a;
                    z.q[1] = b;
                    z.terms = {{l, 2.0}};
                    out.push_back(z);
                }
            }
        for (int32_t x : nodes) {
            Gate m;
            m.type = G_XPHASE;
            m.q[0] = x;
            m.terms = {{p + l, 2.0}};
            out.push_back(m);
        }
    }
}

// Commutation-based cancellation (composer.simplify_circuit): same-type,
// same-site gate pairs, nearest first, that commute with every live gate between
// them; Z and cX pairs cancel, parametric pairs cancel when their summed angle
// terms vanish and merge otherwise; H pairs stay.
enum Basis : int32_t { B_Z, B_X, B_Y, B_XY, B_CA, B_CB };
int32_t basis_of(const Gate& g, int32_t pos) {
    switch (g.type) {
    case G_ZZ: case G_ZPHASE: case G_Z: return B_Z;
    case G_XPHASE: return B_X;
    case G_YPHASE: return B_Y;
    case G_H: return B_XY;
    case G_CX: return pos == 0 ? B_CA : B_CB;
    default: return -1;
    }
}

bool commute(const Gate& a, const Gate& b) {
    for (int32_t x = 0; x < a.nq; ++x)
        for (int32_t y = 0; y < b.nq; ++y)
            if (a.q[x] == b.q[y] && basis_of(a, x) != basis_of(b, y)) return false;
    return true;
}

std::vector<std::pair<int32_t, double>> merge_terms(const std::vector<std::pair<int32_t, double>>& a,
                                                     const std::vector<std::pair<int32_t, double>>& b) {
    std::map<int32_t, double> acc;
    for (auto& t : a) acc[t.first] += t.second;
    for (auto& t : b) acc[t.first] += t.second;
    std::vector<std::pair<int32_t, double>> out;
    for (auto& kv : acc)
        if (kv.second != 0.0) out.push_back(kv);
    return out;
}

std::vector<Gate> simplify(std::vector<Gate> gates) {
    const int32_t n = (int32_t)gates.size();
    // groups by (type, site); ZZ sites are sorted qubit pairs
    std::map<std::vector<int32_t>, std::vector<int32_t>> groups;
    std::vector<std::vector<int32_t>> group_order;
    for (int32_t k = 0; k < n; ++k) {
        const Gate& g = gates[k];
        std::vector<int32_t> key{g.type, g.q[0], g.nq == 2 ? g.q[1] : -1};
        if (g.type == G_ZZ && key[1] > key[2]) std::swap(key[1], key[2]);
        groups[key].push_back(k);
    }
    struct Pair { int32_t d, i, j; };
    std::vector<Pair> pairs;
    for (auto& kv : groups) {
        const auto& m = kv.second;
        for (size_t a = 0; a < m.size(); ++a)
            for (size_t b = a + 1; b < m.size(); ++b) pairs.push_back({m[b] - m[a], m[a], m[b]});
    }
    std::sort(pairs.begin(), pairs.end(), [](const Pair& x, const Pair& y) {
        if (x.d != y.d) return x.d < y.d;
        if (x.i != y.i) return x.i < y.i;
        return x.j < y.j;
    });
    std::vector<char> alive(n, 1);
    for (const Pair& pr : pairs) {
        const int32_t i = pr.i, j = pr.j;
        if (!alive[i] || !alive[j]) continue;
        const Gate& gi = gates[i];
        if (!parametric(gi.type) && gi.type != G_Z && gi.type != G_CX) continue;
        bool ok = true;
        for (int32_t k = i + 1; k < j && ok; ++k)
            if (alive[k] && !commute(gi, gates[k])) ok = false;
        if (!ok) continue;
        if (!parametric(gi.type)) {
            alive[i] = alive[j] = 0;
            continue;
        }
        auto terms = merge_terms(gi.terms, gates[j].terms);
        if (terms.empty()) {
            alive[i] = alive[j] = 0;
            continue;
        }
        gates[j].terms = terms;      // the merged gate takes the first gate's qubits
        gates[j].q[0] = gi.q[0];
        gates[j].q[1] = gi.q[1];
        alive[i] = 0;
    }
    std::vector<Gate> out;
    for (int32_t k = 0; k < n; ++k)
        if (alive[k]) out.push_back(gates[k]);
    return out;
}

std::vector<Gate> lightcone_circuit(const Lightcone& lc, int32_t p, int32_t composer) {
    std::vector<Gate> first, second;
    cone_ansatz(lc, p, composer == 0 ? 0 : 1, first);
    for (int32_t q : {lc.i, lc.j}) {
        Gate z;
        z.type = G_Z;
        z.q[0] = q;
        first.push_back(z);
    }
    cone_ansatz(lc, p, composer == 0 ? 0 : 1, second);
    for (auto it = second.rbegin(); it != second.rend(); ++it) {
        Gate g = *it;
        for (auto& t : g.terms) t.second = -t.second;
        first.push_back(g);
    }
    if (composer == 2) return simplify(first);
    return first;
}

// ---------------------------------------------------------------------------
// Bucket network (circ2buckets) and its line graph
// ---------------------------------------------------------------------------

struct Leaf {
    int32_t type;
    std::vector<std::pair<int32_t, double>> terms;
    std::vector<int32_t> idx;   // variable ids in data-axis order
};

struct Network {
    int32_t n_vars = 0;
    std::vector<std::vector<int32_t>> buckets;   // leaf indices
    std::vector<Leaf> leaves;
    std::vector<int32_t> bra, ket;
};

Network circ2buckets(int32_t n_qubits, const std::vector<Gate>& circuit) {
    Network net;
    std::vector<int32_t> layer(n_qubits);
    for (int32_t q = 0; q < n_qubits; ++q) layer[q] = q;
    net.bra = layer;
    net.buckets.assign(n_qubits, {});
    net.n_vars = n_qubits;
    auto place = [&](int32_t type, const int32_t* qs, int32_t nq,
                     const std::vector<std::pair<int32_t, double>>& terms) {
        Leaf lf;
        lf.type = type;
        lf.terms = terms;
        for (int32_t pos = 0; pos < nq; ++pos) {
            const int32_t q = qs[pos];
            const bool changes = (type == G_H || type == G_M || type == G_XPHASE || type == G_YPHASE) ? pos == 0
                                 : (type == G_CX) ? pos == 1 : false;
            if (changes) {
                const int32_t nv = net.n_vars++;
                net.buckets.emplace_back();
                lf.idx.push_back(layer[q]);
                lf.idx.push_back(nv);
                layer[q] = nv;
            } else {
                lf.idx.push_back(layer[q]);
            }
        }
        const int32_t b = *std::min_element(lf.idx.begin(), lf.idx.end());
        net.buckets[b].push_back((int32_t)net.leaves.size());
        net.leaves.push_back(lf);
    };
    for (int32_t q = 0; q < n_qubits; ++q) place(G_M, &q, 1, {});
    for (auto it = circuit.rbegin(); it != circuit.rend(); ++it) place(it->type, it->q, it->nq, it->terms);
    for (int32_t q = 0; q < n_qubits; ++q) {
        place(G_M, &q, 1, {});
        net.ket.push_back(layer[q]);
    }
    return net;
}

// Line graph without the bra/ket variables: nodes in first-appearance order
// (buckets in order, tensors in order, indices in order), a clique per tensor.
struct LineGraph {
    std::vector<int32_t> nodes;                  // variable ids in node order
    std::vector<std::vector<int32_t>> adj;       // by node index, distinct
};

LineGraph line_graph(const Network& net) {
    std::vector<char> ignore(net.n_vars, 0);
    for (int32_t v : net.bra) ignore[v] = 1;
    for (int32_t v : net.ket) ignore[v] = 1;
    std::vector<int32_t> pos(net.n_vars, -1);
    LineGraph lg;
    std::vector<std::vector<int32_t>> nb;
    for (const auto& bucket : net.buckets)
        for (int32_t li : bucket) {
            std::vector<int32_t> vs;
            for (int32_t v : net.leaves[li].idx)
                if (!ignore[v]) vs.push_back(v);
            for (int32_t v : vs)
                if (pos[v] < 0) {
                    pos[v] = (int32_t)lg.nodes.size();
                    lg.nodes.push_back(v);
                    nb.emplace_back();
                }
            for (size_t a = 0; a < vs.size(); ++a)
                for (size_t b = a + 1; b < vs.size(); ++b) {
                    const int32_t x = pos[vs[a]], y = pos[vs[b]];
                    if (x == y) continue;
                    if (std::find(nb[x].begin(), nb[x].end(), y) == nb[x].end()) {
                        nb[x].push_back(y);
                        nb[y].push_back(x);
                    }
                }
        }
    lg.adj = std::move(nb);
    return lg;
}

// ---------------------------------------------------------------------------
// Orders and slices on a node subset of the line graph
// ---------------------------------------------------------------------------

struct OrderOpts {
    int32_t ordering = 0;      // 0 greedy, 1 randomized greedy (best of samples)
    double temp = 0.002;
    int32_t repeats = 16;
    uint64_t seed = 0;
};

// Order of the live nodes (graph-node order preserved), as node indices, and widths.
void order_nodes(const LineGraph& lg, const std::vector<char>& live, const OrderOpts& o, std::vector<int32_t>& order,
                 std::vector<int32_t>& widths) {
    std::vector<int32_t> ids, local(lg.nodes.size(), -1);
    for (size_t k = 0; k < lg.nodes.size(); ++k)
        if (live[k]) {
            local[k] = (int32_t)ids.size();
            ids.push_back((int32_t)k);
        }
    const int32_t m = (int32_t)ids.size();
    std::vector<int32_t> ptr(m + 1, 0), idx;
    for (int32_t a = 0; a < m; ++a) {
        for (int32_t b : lg.adj[ids[a]])
            if (live[b]) idx.push_back(local[b]);
        ptr[a + 1] = (int32_t)idx.size();
    }
    order.assign(m, 0);
    widths.assign(m, 0);
    if (m == 0) return;
    const int32_t* ip = idx.empty() ? ptr.data() : idx.data();
    if (o.ordering == 0)
        qb_greedy_order(m, ptr.data(), ip, order.data(), widths.data());
    else
        qb_rgreedy_order(m, ptr.data(), ip, o.temp, o.seed, o.repeats, order.data(), widths.data());
    for (auto& v : order) v = ids[v];
}

void widths_of(const LineGraph& lg, const std::vector<char>& live, const std::vector<int32_t>& order,
               std::vector<int32_t>& widths) {
    std::vector<int32_t> ids, local(lg.nodes.size(), -1);
    for (size_t k = 0; k < lg.nodes.size(); ++k)
        if (live[k]) {
            local[k] = (int32_t)ids.size();
            ids.push_back((int32_t)k);
        }
    const int32_t m = (int32_t)ids.size();
    std::vector<int32_t> ptr(m + 1, 0), idx, ord;
    for (int32_t a = 0; a < m; ++a) {
        for (int32_t b : lg.adj[ids[a]])
            if (live[b]) idx.push_back(local[b]);
        ptr[a + 1] = (int32_t)idx.size();
    }
    for (int32_t v : order) ord.push_back(local[v]);
    widths.assign(ord.size(), 0);
    if (ord.empty()) return;
    qb_elimination_widths(m, ptr.data(), idx.empty() ? ptr.data() : idx.data(), (int32_t)ord.size(), ord.data(),
                          widths.data());
}

int32_t max_of(const std::vector<int32_t>& w) { return w.empty() ? 0 : *std::max_element(w.begin(), w.end()); }

// Tree-trimming cost of each live node for an order (ordering._trim_costs).
void trim_costs(const LineGraph& lg, const std::vector<char>& live, const std::vector<int32_t>& order,
                std::vector<unsigned __int128>& cost) {
    const size_t N = lg.nodes.size();
    std::vector<std::vector<int32_t>> adj(N);
    for (size_t k = 0; k < N; ++k)
        if (live[k])
            for (int32_t b : lg.adj[k])
                if (live[b]) adj[k].push_back(b);
    cost.assign(N, 0);
    std::vector<char> gone(N, 0);
    std::vector<int32_t> mark(N, 0);
    int32_t tick = 0;
    for (int32_t v : order) {
        std::vector<int32_t> nb;
        nb.swap(adj[v]);
        gone[v] = 1;
        const unsigned __int128 w = (unsigned __int128)1 << nb.size();
        cost[v] += w;
        for (int32_t u : nb) {
            cost[u] += w;
            std::vector<int32_t>& s = adj[u];
            ++tick;
            size_t keep = 0;
            for (size_t q = 0; q < s.size(); ++q) {
                if (s[q] == v) continue;
                mark[s[q]] = tick;
                s[keep++] = s[q];
            }
            s.resize(keep);
            mark[u] = tick;
            for (int32_t x : nb)
                if (mark[x] != tick) {
                    mark[x] = tick;
                    s.push_back(x);
                }
        }
    }
}

struct Ordered {
    std::vector<int32_t> order;     // node indices, eliminated first to last
    std::vector<int32_t> par;       // sliced node indices, in selection order
    int32_t treewidth = 0;
};

// slicer: 0 none, 1 GreedyParvars (SlicesOptimizer), 2 tree trimming.
int optimize(const LineGraph& lg, const OrderOpts& o, int32_t slicer, int32_t max_tw, int32_t tw_bias, Ordered& res) {
    const size_t N = lg.nodes.size();
    std::vector<char> live(N, 1);
    std::vector<int32_t> widths;
    order_nodes(lg, live, o, res.order, widths);
    res.treewidth = max_of(widths);
    if (slicer == 0) return QB_OK;
    const int32_t target = max_tw - tw_bias;
    if (slicer == 1) {
        std::vector<int32_t> peo = res.order;
        int32_t tw = res.treewidth;
        for (;;) {
            if (tw < target) break;
            // highest degree among live nodes, first in node order
            int32_t best = -1, best_deg = -1;
            for (size_t k = 0; k < N; ++k) {
                if (!live[k]) continue;
                int32_t d = 0;
                for (int32_t b : lg.adj[k]) d += live[b] ? 1 : 0;
                if (d > best_deg) {
                    best_deg = d;
                    best = (int32_t)k;
                }
            }
            if (best < 0) return qb_internal::fail(QB_ERR_UNSUPPORTED, "Estimated OOM");
            res.par.push_back(best);
            live[best] = 0;
            order_nodes(lg, live, o, peo, widths);
            tw = max_of(widths);
        }
        res.order = peo;
        res.treewidth = tw;
        return QB_OK;
    }
    std::vector<int32_t> peo = res.order;
    widths_of(lg, live, peo, widths);
    int32_t tw = max_of(widths);
    std::vector<unsigned __int128> cost;
    while (tw > target) {
        const int32_t delta = tw - target;
        const int32_t want = (int32_t)(delta * 0.2) + 1;
        trim_costs(lg, live, peo, cost);
        std::vector<int32_t> pos(N, -1), cand;
        for (size_t k = 0; k < peo.size(); ++k) pos[peo[k]] = (int32_t)k;
        for (size_t k = 0; k < N; ++k)
            if (live[k]) cand.push_back((int32_t)k);
        std::stable_sort(cand.begin(), cand.end(), [&](int32_t a, int32_t b) {
            if (cost[a] != cost[b]) return cost[a] > cost[b];
            return pos[a] < pos[b];
        });
        if (cand.empty()) return qb_internal::fail(QB_ERR_UNSUPPORTED, "Estimated OOM");
        for (int32_t k = 0; k < want && k < (int32_t)cand.size(); ++k) {
            live[cand[k]] = 0;
            res.par.push_back(cand[k]);
        }
        order_nodes(lg, live, o, peo, widths);
        tw = max_of(widths);
    }
    res.order = peo;
    res.treewidth = tw;
    return QB_OK;
}

// ---------------------------------------------------------------------------
// Plan compiler (item plans: plan.PlanCompiler with gate_space "lds")
// ---------------------------------------------------------------------------

constexpr int32_t kPlanHdr = 12, kGateWords = 16, kMaxIn = 4, kGatherHdr = 4, kReduceHdr = 4;
constexpr int32_t kGatherMaxSteps = 16, kReduceMaxOut = 64, kGateBase = 4;
constexpr uint32_t kGlobal = 0x80000000u, kOpGather = 1u << 15, kOpReduce = 1u << 13;
constexpr uint32_t kLocOne = 0, kLocZero = 1;

struct Buf {
    int64_t n = 0;
    int64_t loc = -1;
    bool leaf = false;
    int32_t last_op = -1;
    bool global() const { return loc >= 0 && (loc & kGlobal); }
};

struct Node {
    std::vector<std::pair<int32_t, int32_t>> idx;   // (var id, size), ascending id
    std::vector<int32_t> dev;                      // ids of size-2 variables
    int32_t buf = -1;
    int32_t step = -1;
    int64_t n() const { return (int64_t)1 << dev.size(); }
};

struct Step {
    std::vector<int32_t> inputs;   // node indices
    int32_t out = -1;
    int32_t rank_out = 0, sum = 0;
    std::vector<uint32_t> lows, tops;
    // reduction op (a chain of simple sums): k > 0
    int32_t k = 0;
};

uint32_t pext(uint32_t x, uint32_t m) {
    uint32_t r = 0, bit = 1;
    for (; m; m &= m - 1, bit <<= 1)
        if (x & m & (0u - m)) r |= bit;
    return r;
}

uint32_t runs_word(uint32_t mask) {
    uint32_t src[2] = {0, 0}, w[2] = {0, 0}, dst[2] = {0, 0};
    int n = 0;
    uint32_t d = 0;
    for (uint32_t b = 0; b < 32;) {
        if (!((mask >> b) & 1u)) { ++b; continue; }
        const uint32_t s = b;
        while (b < 32 && ((mask >> b) & 1u)) ++b;
        if (n < 2) { src[n] = s; w[n] = b - s; dst[n] = d; }
        d += b - s;
        ++n;
    }
    if (n > 2) return 3u << 30;
    uint32_t r = (uint32_t)n << 30;
    for (int k = 0; k < n; ++k) r |= (src[k] | (w[k] << 5) | (dst[k] << 10)) << (15 * k);
    return r;
}

int32_t f32_bits(double x) {
    const float f = (float)x;
    int32_t r;
    std::memcpy(&r, &f, 4);
    return r;
}

struct CompileOpts {
    int32_t lanes = 64;
    int64_t lds_tensor_limit = 2048;
    int64_t lds_capacity = -1;
};

struct PlanOut {
    std::vector<int32_t> words;
    qb_lightcone_plan_info info{};
};

struct Compiler {
    std::vector<Buf> bufs;
    std::vector<Node> nodes;
    std::vector<Step> steps;

    int32_t new_buf(int64_t n, int64_t loc = -1, bool leaf = false) {
        Buf b;
        b.n = n;
        b.loc = loc;
        b.leaf = leaf;
        bufs.push_back(b);
        return (int32_t)bufs.size() - 1;
    }

    int emit(const std::vector<int32_t>& inputs, bool no_sum, int32_t& out) {
        if (inputs.empty() || (int32_t)inputs.size() > kMaxIn)
            return qb_internal::fail(QB_ERR_INVALID, "step with more inputs than the kernel takes");
        std::vector<std::pair<int32_t, int32_t>> U;
        for (int32_t nd : inputs)
            for (auto& v : nodes[nd].idx) {
                bool seen = false;
                for (auto& u : U)
                    if (u.first == v.first) { seen = true; break; }
                if (!seen) U.push_back(v);
            }
        std::stable_sort(U.begin(), U.end(), [](const std::pair<int32_t, int32_t>& a,
                                                const std::pair<int32_t, int32_t>& b) { return a.first < b.first; });
        std::vector<int32_t> U2;
        for (auto& v : U)
            if (v.second == 2) U2.push_back(v.first);
        const bool do_sum = !no_sum && !U.empty();
        const int32_t sum_flag = (do_sum && U[0].second == 2) ? 1 : 0;
        Node o;
        o.idx.assign(U.begin() + (do_sum ? 1 : 0), U.end());
        o.dev.assign(U2.begin() + (sum_flag ? 1 : 0), U2.end());
        const int32_t rank_u = (int32_t)U2.size();
        if (rank_u > 31) return qb_internal::fail(QB_ERR_INVALID, "step rank exceeds 31");
        const uint32_t top_bit = sum_flag ? (1u << (rank_u - 1)) : 0u;
        Step st;
        for (int32_t nd : inputs) {
            uint32_t m = 0;
            for (int32_t vid : nodes[nd].dev) {
                const int32_t a = (int32_t)(std::find(U2.begin(), U2.end(), vid) - U2.begin());
                m |= 1u << (rank_u - 1 - a);
            }
            const uint32_t low = m & ~top_bit;
            st.lows.push_back(low);
            st.tops.push_back((m & top_bit) ? (1u << __builtin_popcount(low)) : 0u);
        }
        o.buf = new_buf((int64_t)1 << o.dev.size());
        o.step = (int32_t)steps.size();
        nodes.push_back(o);
        out = (int32_t)nodes.size() - 1;
        st.inputs = inputs;
        st.out = out;
        st.rank_out = rank_u - sum_flag;
        st.sum = sum_flag;
        steps.push_back(st);
        return QB_OK;
    }

    int process(std::vector<int32_t> in, bool no_sum, int32_t& out) {
        while ((int32_t)in.size() > kMaxIn) {
            int32_t head;
            std::vector<int32_t> first(in.begin(), in.begin() + kMaxIn);
            int rc = emit(first, true, head);
            if (rc) return rc;
            std::vector<int32_t> rest{head};
            rest.insert(rest.end(), in.begin() + kMaxIn, in.end());
            in.swap(rest);
        }
        if (in.size() == 1 && !nodes[in[0]].idx.empty() && nodes[in[0]].idx[0].second == 1 && !no_sum) {
            Node pass;
            pass.idx.assign(nodes[in[0]].idx.begin() + 1, nodes[in[0]].idx.end());
            pass.dev = nodes[in[0]].dev;
            pass.buf = nodes[in[0]].buf;
            pass.step = nodes[in[0]].step;
            nodes.push_back(pass);
            out = (int32_t)nodes.size() - 1;
            return QB_OK;
        }
        return emit(in, no_sum, out);
    }
};

bool simple_sum(const Step& st) {
    if (st.k || st.inputs.size() != 1 || !st.sum) return false;
    const uint32_t n_out = 1u << st.rank_out;
    return st.lows[0] == n_out - 1 && st.tops[0] == n_out;
}

// Chains of >= 2 simple sums whose intermediates have no other reader become
// reduction steps (plan.reduce_chains).
std::vector<Step> reduce_chains(Compiler& c, int32_t result_step, int64_t lds_limit) {
    std::vector<Step>& steps = c.steps;
    const int32_t n = (int32_t)steps.size();
    std::vector<int32_t> readers(n, 0), nxt(n, -1);
    std::vector<char> has_prev(n, 0);
    for (auto& st : steps)
        for (int32_t nd : st.inputs)
            if (c.nodes[nd].step >= 0) ++readers[c.nodes[nd].step];
    for (int32_t i = 0; i < n; ++i)
        if (simple_sum(steps[i])) {
            const int32_t j = c.nodes[steps[i].inputs[0]].step;
            if (j >= 0 && readers[j] == 1 && j != result_step && simple_sum(steps[j])) nxt[j] = i;
        }
    for (int32_t j = 0; j < n; ++j)
        if (nxt[j] >= 0) has_prev[nxt[j]] = 1;
    std::vector<int32_t> repl_kind(n, 0);   // 0 keep, 1 replaced by a reduction, 2 dropped
    std::vector<Step> reduction(n);
    bool any = false;
    for (int32_t i = 0; i < n; ++i) {
        if (has_prev[i] || nxt[i] < 0 || !simple_sum(steps[i])) continue;
        std::vector<int32_t> chain{i};
        while (nxt[chain.back()] >= 0) chain.push_back(nxt[chain.back()]);
        const Step& first = steps[chain.front()];
        const Step& last = steps[chain.back()];
        if (((int64_t)1 << last.rank_out) > kReduceMaxOut || c.nodes[first.inputs[0]].n() > lds_limit) continue;
        Step r;
        r.inputs = {first.inputs[0]};
        r.out = last.out;
        r.k = (int32_t)chain.size();
        r.rank_out = last.rank_out;
        reduction[chain.back()] = r;
        repl_kind[chain.back()] = 1;
        for (size_t q = 0; q + 1 < chain.size(); ++q) repl_kind[chain[q]] = 2;
        any = true;
    }
    if (!any) return steps;
    std::vector<Step> out;
    std::vector<int32_t> index(n, -1);
    for (int32_t i = 0; i < n; ++i) {
        if (repl_kind[i] == 2) continue;
        index[i] = (int32_t)out.size();
        out.push_back(repl_kind[i] == 1 ? reduction[i] : steps[i]);
    }
    // renumber producers of every node the new steps touch
    std::vector<char> seen(c.nodes.size(), 0);
    for (auto& st : out) {
        std::vector<int32_t> touch(st.inputs);
        touch.push_back(st.out);
        for (int32_t nd : touch)
            if (!seen[nd] && c.nodes[nd].step >= 0) {
                seen[nd] = 1;
                c.nodes[nd].step = index[c.nodes[nd].step];
            }
    }
    return out;
}

int compile_plan(const Network& net, const Ordered& ord, const LineGraph& lg, const CompileOpts& co, PlanOut& po) {
    // new variable ids: ket, bra, then the order, then the sliced variables
    std::vector<int32_t> perm(net.n_vars, -1);
    int32_t nid = 0;
    for (int32_t v : net.ket) perm[v] = nid++;
    for (int32_t v : net.bra) perm[v] = nid++;
    for (int32_t k : ord.order) perm[lg.nodes[k]] = nid++;
    for (int32_t k : ord.par) perm[lg.nodes[k]] = nid++;
    if (nid != net.n_vars) return qb_internal::fail(QB_ERR_INVALID, "order does not cover every variable");
    const int32_t n_fixed = 2 * (int32_t)net.ket.size();
    const int32_t n_par = (int32_t)ord.par.size();
    if (n_par > 24) return qb_internal::fail(QB_ERR_INVALID, "more than 24 sliced variables");
    const int32_t first_par = net.n_vars - n_par;
    // reorder_buckets: each tensor to the bucket of its smallest new id, in
    // the old buckets' order
    std::vector<std::vector<int32_t>> buckets(net.n_vars);
    for (const auto& b : net.buckets)
        for (int32_t li : b) {
            int32_t m = INT32_MAX;
            for (int32_t v : net.leaves[li].idx) m = std::min(m, perm[v]);
            buckets[m].push_back(li);
        }
    Compiler c;
    // leaves: distinct gate tensors in the LDS gate area
    std::vector<std::vector<int32_t>> gates;
    std::map<std::vector<int32_t>, int32_t> gate_buf;
    int64_t gate_area = kGateBase;
    std::vector<std::vector<int32_t>> sym(net.n_vars);
    for (int32_t b = 0; b < net.n_vars; ++b)
        for (int32_t li : buckets[b]) {
            const Leaf& lf = net.leaves[li];
            const int32_t r = (int32_t)lf.idx.size();
            std::vector<int32_t> order(r);
            for (int32_t a = 0; a < r; ++a) order[a] = a;
            std::stable_sort(order.begin(), order.end(),
                             [&](int32_t x, int32_t y) { return perm[lf.idx[x]] < perm[lf.idx[y]]; });
            int64_t base_off = 0;
            std::vector<int64_t> dev_axes;
            std::vector<std::pair<int32_t, int64_t>> slice_axes;
            Node nd;
            for (int32_t a : order) {
                const int32_t vid = perm[lf.idx[a]];
                const int64_t stride = (int64_t)1 << (r - 1 - a);
                if (vid >= first_par) {
                    slice_axes.emplace_back(vid - first_par, stride);
                    continue;
                }
                if (vid < n_fixed) {     // bra / ket wires, sliced to state 0
                    nd.idx.emplace_back(vid, 1);
                    continue;
                }
                dev_axes.push_back(stride);
                nd.idx.emplace_back(vid, 2);
                nd.dev.push_back(vid);
            }
            const int32_t n = 1 << dev_axes.size();
            uint32_t src_map = 0;
            for (int32_t e = 0; e < n; ++e) {
                int64_t off = base_off;
                for (size_t k = 0; k < dev_axes.size(); ++k)
                    if ((e >> (dev_axes.size() - 1 - k)) & 1) off += dev_axes[k];
                src_map |= (uint32_t)(off & 15) << (4 * e);
            }
            if (lf.terms.size() > 2)
                return qb_internal::fail(QB_ERR_INVALID, "a gate angle depends on more than 2 parameters");
            if (slice_axes.size() > 2) return qb_internal::fail(QB_ERR_INVALID, "a gate touches more than 2 sliced variables");
            const int32_t sl = (int32_t)slice_axes.size();
            const int32_t slw = sl | ((sl > 0 ? slice_axes[0].first : 0) << 8) | ((sl > 1 ? slice_axes[1].first : 0) << 16);
            int32_t j0 = -1, j1 = -1;
            double c0 = 0.0, c1 = 0.0;
            if (!lf.terms.empty()) {
                j0 = lf.terms[0].first;
                c0 = lf.terms[0].second;
                if (lf.terms.size() > 1) {
                    j1 = lf.terms[1].first;
                    c1 = lf.terms[1].second;
                }
            }
            std::vector<int32_t> entry{device_kind(lf.type), 0, n, (int32_t)src_map, j0, f32_bits(c0), j1, f32_bits(c1),
                                       0, 0, 0, slw,
                                       sl > 0 ? (int32_t)slice_axes[0].second : 0,
                                       sl > 1 ? (int32_t)slice_axes[1].second : 0, 0, 0};
            std::vector<int32_t> sig(entry);
            sig.erase(sig.begin() + 1);
            auto f = gate_buf.find(sig);
            if (f == gate_buf.end()) {
                entry[1] = (int32_t)gate_area;
                const int32_t bi = c.new_buf(n, gate_area, true);
                gate_buf.emplace(sig, bi);
                gates.push_back(entry);
                gate_area += n;
                nd.buf = bi;
            } else {
                nd.buf = f->second;
            }
            c.nodes.push_back(nd);
            sym[b].push_back((int32_t)c.nodes.size() - 1);
        }
    // phase 1: symbolic elimination
    const int32_t n_contract = net.n_vars - n_par;
    std::vector<int32_t> scalars;
    for (int32_t b = 0; b < n_contract; ++b) {
        if (sym[b].empty()) continue;
        int32_t res;
        int rc = c.process(sym[b], false, res);
        if (rc) return rc;
        if (!c.nodes[res].idx.empty())
            sym[c.nodes[res].idx[0].first].push_back(res);
        else
            scalars.push_back(res);
    }
    std::vector<int32_t> rest;
    for (int32_t b = n_contract; b < net.n_vars; ++b) rest.insert(rest.end(), sym[b].begin(), sym[b].end());
    if (!rest.empty()) {
        int32_t res;
        int rc = c.process(rest, true, res);
        if (rc) return rc;
        for (auto& v : c.nodes[res].idx)
            if (v.second == 2) return qb_internal::fail(QB_ERR_INVALID, "plans with free (batch) variables are not supported");
        scalars.push_back(res);
    }
    if (scalars.empty()) return qb_internal::fail(QB_ERR_INVALID, "network has no scalar result");
    int32_t result = scalars[0];
    if (scalars.size() > 1) {
        int rc = c.process(scalars, true, result);
        if (rc) return rc;
    }
    if (!c.nodes[result].dev.empty()) return qb_internal::fail(QB_ERR_INVALID, "network result is not a scalar");
    // step-model figures of the bucket steps
    int64_t total_bytes = 0, total_flops = 0;
    int32_t max_rank = 0;
    for (auto& st : c.steps) {
        const int64_t n_out = c.nodes[st.out].n();
        const int32_t rank_u = st.rank_out + st.sum;
        max_rank = std::max(max_rank, rank_u);
        int64_t in_sum = 0;
        for (int32_t nd : st.inputs) in_sum += c.nodes[nd].n();
        total_bytes += 8 * (in_sum + n_out);
        const int64_t n_union = (int64_t)1 << rank_u;
        total_flops += 6 * ((int64_t)st.inputs.size() - 1) * n_union + 2 * (n_union - n_out);
    }
    const int32_t n_steps = (int32_t)c.steps.size();
    // phase 2: reduction chains and list scheduling
    const int64_t lim = co.lds_tensor_limit;
    std::vector<Step> ex = reduce_chains(c, c.nodes[result].step, lim);
    const int32_t n = (int32_t)ex.size();
    const int32_t lanes = co.lanes;
    std::vector<char> gok(n, 0);
    for (int32_t i = 0; i < n; ++i) {
        const Step& st = ex[i];
        if (st.k) continue;
        const int64_t n_out = c.nodes[st.out].n();
        bool ok = n_out <= lim && n_out <= lanes && (int32_t)st.inputs.size() <= kMaxIn;
        for (int32_t nd : st.inputs)
            if (c.nodes[nd].n() > lim && !c.bufs[c.nodes[nd].buf].leaf) ok = false;
        gok[i] = ok;
    }
    std::vector<std::vector<int32_t>> prods(n), cons(n);
    for (int32_t i = 0; i < n; ++i) {
        for (int32_t nd : ex[i].inputs)
            if (c.nodes[nd].step >= 0) prods[i].push_back(c.nodes[nd].step);
        std::vector<int32_t> u(prods[i]);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        for (int32_t j : u) cons[j].push_back(i);
        prods[i] = u;
    }
    std::vector<int32_t> height(n, 0), waiting(n);
    for (int32_t i = n - 1; i >= 0; --i) {
        int32_t h = 0;
        for (int32_t cc : cons[i]) h = std::max(h, height[cc]);
        height[i] = (gok[i] ? 1 : 2) + h;
    }
    std::vector<int32_t> ready;
    for (int32_t i = 0; i < n; ++i) {
        waiting[i] = (int32_t)prods[i].size();
        if (!waiting[i]) ready.push_back(i);
    }
    std::vector<std::vector<int32_t>> ops;
    std::vector<char> op_gather;
    int32_t done = 0;
    while (done < n) {
        std::sort(ready.begin(), ready.end(), [&](int32_t a, int32_t b) {
            if (height[a] != height[b]) return height[a] > height[b];
            return a < b;
        });
        std::vector<int32_t> chosen;
        bool kind = false;
        if (!gok[ready[0]]) {
            chosen.push_back(ready[0]);
        } else {
            int64_t fill = 0;
            for (int32_t i : ready) {
                const int64_t m = c.nodes[ex[i].out].n();
                if (gok[i] && fill + m <= lanes && (int32_t)chosen.size() < kGatherMaxSteps) {
                    chosen.push_back(i);
                    fill += m;
                }
            }
            kind = true;
        }
        std::sort(chosen.begin(), chosen.end());
        ops.push_back(chosen);
        op_gather.push_back(kind);
        std::vector<int32_t> keep;
        for (int32_t i : ready)
            if (!std::binary_search(chosen.begin(), chosen.end(), i)) keep.push_back(i);
        ready.swap(keep);
        done += (int32_t)chosen.size();
        for (int32_t i : chosen)
            for (int32_t cc : cons[i])
                if (--waiting[cc] == 0) ready.push_back(cc);
    }
    std::vector<int32_t> op_of(n, 0);
    for (size_t k = 0; k < ops.size(); ++k)
        for (int32_t i : ops[k]) op_of[i] = (int32_t)k;
    // phase 3: interval placement
    for (int32_t i = 0; i < n; ++i)
        for (int32_t nd : ex[i].inputs) {
            Buf& b = c.bufs[c.nodes[nd].buf];
            b.last_op = std::max(b.last_op, op_of[i]);
        }
    struct Item { int32_t buf; int64_t s, e; };
    std::vector<Item> lds_items, glb_items;
    const int64_t end_all = (int64_t)ops.size();
    for (size_t k = 0; k < ops.size(); ++k)
        for (int32_t i : ops[k]) {
            const int32_t bi = c.nodes[ex[i].out].buf;
            const Buf& b = c.bufs[bi];
            const int64_t e = b.last_op >= 0 ? b.last_op : end_all;
            (b.n <= lim ? lds_items : glb_items).push_back({bi, (int64_t)k, e});
        }
    auto place = [&](std::vector<Item>& items, int64_t base, int64_t align, int64_t& peak) -> int {
        peak = base;
        if (items.empty()) return QB_OK;
        const int32_t m = (int32_t)items.size();
        std::vector<int64_t> sizes(m), starts(m), ends(m), offs(m), best_offs;
        for (int32_t k = 0; k < m; ++k) {
            sizes[k] = (c.bufs[items[k].buf].n + align - 1) / align * align;
            starts[k] = items[k].s;
            ends[k] = items[k].e;
        }
        int64_t best_peak = -1;
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<int64_t> order(m);
            for (int32_t k = 0; k < m; ++k) order[k] = k;
            std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
                const int64_t ka = pass == 0 ? sizes[a] : sizes[a] * (ends[a] - starts[a] + 1);
                const int64_t kb = pass == 0 ? sizes[b] : sizes[b] * (ends[b] - starts[b] + 1);
                if (ka != kb) return ka > kb;
                return starts[a] < starts[b];
            });
            int64_t pk = 0;
            int rc = qb_place_intervals(m, sizes.data(), starts.data(), ends.data(), order.data(), base, offs.data(), &pk);
            if (rc) return rc;
            if (best_peak < 0 || pk < best_peak) {
                best_peak = pk;
                best_offs = offs;
            }
        }
        for (int32_t k = 0; k < m; ++k) c.bufs[items[k].buf].loc = best_offs[k];
        peak = best_peak;
        return QB_OK;
    };
    int64_t lds_peak = 0, glb_peak = 0;
    int rc = place(lds_items, gate_area, 2, lds_peak);
    if (rc) return rc;
    if (co.lds_capacity >= 0 && lds_peak > co.lds_capacity)
        return qb_internal::fail(QB_ERR_INVALID, "LDS capacity exhausted; use a larger capacity, a smaller "
                                                 "LDS tensor limit or more slicing");
    rc = place(glb_items, 0, 64, glb_peak);
    if (rc) return rc;
    for (auto& it : glb_items) c.bufs[it.buf].loc |= kGlobal;
    // phase 4: encoding
    auto loc_of = [&](int32_t nd) { return (int32_t)(uint32_t)c.bufs[c.nodes[nd].buf].loc; };
    auto glb_of = [&](int32_t nd) { return c.bufs[c.nodes[nd].buf].global(); };
    std::vector<std::vector<int32_t>> op_words;
    bool glb_written = false;
    int32_t barriers = 0, n_gather = 0, n_reduce = 0;
    for (size_t k = 0; k < ops.size(); ++k) {
        bool reads_glb = false, writes_glb = false;
        for (int32_t i : ops[k]) {
            for (int32_t nd : ex[i].inputs) reads_glb = reads_glb || glb_of(nd);
            writes_glb = writes_glb || glb_of(ex[i].out);
        }
        int32_t sync = 0;
        if ((reads_glb || writes_glb) && glb_written) {
            sync = 1;
            ++barriers;
            glb_written = false;
        }
        if (writes_glb) glb_written = true;
        std::vector<int32_t> w;
        if (op_gather[k]) {
            int32_t ni = 0;
            for (int32_t i : ops[k]) ni = std::max(ni, (int32_t)ex[i].inputs.size());
            std::vector<std::vector<uint32_t>> cols(ni + 1);
            for (int32_t i : ops[k]) {
                const Step& st = ex[i];
                const int64_t nn = c.nodes[st.out].n();
                const uint32_t base = (uint32_t)loc_of(st.out);
                for (int64_t e = 0; e < nn; ++e) cols[0].push_back(base + (uint32_t)e);
                for (int32_t j = 0; j < ni; ++j) {
                    if (j < (int32_t)st.inputs.size()) {
                        const uint32_t loc = (uint32_t)loc_of(st.inputs[j]);
                        const uint32_t top = st.tops[j];
                        const uint32_t last = pext((uint32_t)(nn - 1), st.lows[j]);
                        if ((uint64_t)loc + last + (st.sum ? top : 0) >= 65536)
                            return qb_internal::fail(QB_ERR_INVALID, "gather op input beyond 16-bit LDS offsets");
                        for (int64_t e = 0; e < nn; ++e) {
                            const uint32_t t = pext((uint32_t)e, st.lows[j]);
                            if (st.sum)
                                cols[1 + j].push_back((loc + t) | ((loc + top + t) << 16));
                            else
                                cols[1 + j].push_back((loc + t) + ((j == 0 ? kLocZero : kLocOne) << 16));
                        }
                    } else {
                        for (int64_t e = 0; e < nn; ++e) cols[1 + j].push_back(kLocOne | (kLocOne << 16));
                    }
                }
            }
            const int32_t n_act = (int32_t)cols[0].size();
            const int32_t length = kGatherHdr + (ni + 1) * n_act;
            w = {(int32_t)(kOpGather | (uint32_t)ops[k].size() | ((uint32_t)length << 16)), n_act, kGatherHdr, ni};
            for (int32_t l = 0; l < n_act; ++l)
                for (int32_t j = 0; j <= ni; ++j) w.push_back((int32_t)cols[j][l]);
            ++n_gather;
        } else {
            const Step& st = ex[ops[k][0]];
            if (st.k) {
                const uint32_t w0 = kOpReduce | ((uint32_t)st.rank_out << 4) | ((uint32_t)sync << 11) |
                                    ((uint32_t)kReduceHdr << 16);
                w = {(int32_t)w0, st.k, loc_of(st.out), loc_of(st.inputs[0])};
                ++n_reduce;
            } else {
                const int32_t rec_len = 3 + 4 * (int32_t)st.inputs.size();
                bool all_lds = !glb_of(st.out);
                for (int32_t nd : st.inputs) all_lds = all_lds && !glb_of(nd);
                const uint32_t w0 = (uint32_t)st.inputs.size() | ((uint32_t)st.rank_out << 4) |
                                    ((uint32_t)st.sum << 9) | ((uint32_t)sync << 11) |
                                    ((uint32_t)(all_lds ? 1 : 0) << 12) | ((uint32_t)rec_len << 16);
                w = {(int32_t)w0, 0, loc_of(st.out)};
                for (size_t j = 0; j < st.inputs.size(); ++j) {
                    w.push_back(loc_of(st.inputs[j]));
                    w.push_back((int32_t)st.tops[j]);
                    w.push_back((int32_t)st.lows[j]);
                    w.push_back((int32_t)runs_word(st.lows[j]));
                }
            }
        }
        op_words.push_back(std::move(w));
    }
    int32_t steps_words = 0;
    for (auto& w : op_words) steps_words += (int32_t)w.size();
    const int32_t n_ops = (int32_t)op_words.size();
    std::vector<int32_t>& words = po.words;
    words = {(int32_t)gates.size(), n_ops, kPlanHdr, kPlanHdr + kGateWords * (int32_t)gates.size(), loc_of(result),
             (int32_t)lds_peak, (int32_t)(uint32_t)(glb_peak & 0xFFFFFFFF), n_par, steps_words,
             (int32_t)(glb_peak >> 32), 0, 0};
    for (auto& g : gates) words.insert(words.end(), g.begin(), g.end());
    for (auto& w : op_words) words.insert(words.end(), w.begin(), w.end());
    qb_lightcone_plan_info& in = po.info;
    in.n_steps = n_steps;
    in.n_ops = n_ops;
    in.n_gates = (int32_t)gates.size();
    in.treewidth = ord.treewidth;
    in.max_rank = max_rank;
    in.lds_need = lds_peak;
    in.glb_need = glb_peak;
    in.n_sliced = n_par;
    in.bytes_step_model = total_bytes;
    in.flops = total_flops;
    in.n_barriers = barriers;
    in.n_gather = n_gather;
    in.n_reduce = n_reduce;
    in.steps_words = steps_words;
    return QB_OK;
}

}  // namespace

extern "C" {

int qb_lightcone_classes(int32_t n_nodes, const int32_t* adj_ptr, const int32_t* adj_idx, int32_t n_edges,
                         const int32_t* edges, int32_t p, int32_t* edge_class, int32_t* n_classes,
                         int32_t* class_words, int64_t capacity, int64_t* used) {
    if (n_nodes < 0 || n_edges < 0 || p < 0 || !edge_class || !n_classes || !used ||
        (n_nodes > 0 && (!adj_ptr || !adj_idx)) || (n_edges > 0 && !edges) || capacity < 0 ||
        (capacity > 0 && !class_words))
        return qb_internal::fail(QB_ERR_INVALID, "qb_lightcone_classes: bad size or null pointer");
    for (int32_t v = 0; v < n_nodes; ++v)
        if (adj_ptr[v + 1] < adj_ptr[v]) return qb_internal::fail(QB_ERR_INVALID, "qb_lightcone_classes: malformed adjacency");
    for (int32_t k = 0; k < adj_ptr[n_nodes]; ++k)
        if (adj_idx[k] < 0 || adj_idx[k] >= n_nodes)
            return qb_internal::fail(QB_ERR_INVALID, "qb_lightcone_classes: neighbour outside the graph");
    Csr g{n_nodes, adj_ptr, adj_idx};
    for (int32_t e = 0; e < n_edges; ++e) {
        const int32_t u = edges[2 * e], v = edges[2 * e + 1];
        if (u < 0 || u >= n_nodes || v < 0 || v >= n_nodes || u == v)
            return qb_internal::fail(QB_ERR_INVALID, "qb_lightcone_classes: edge endpoint outside the graph");
    }
    // every edge's canonical record, computed on up to 16 threads (each with
    // its own scratch), then numbered into classes in edge order
    std::vector<std::vector<int32_t>> keys(n_edges);
    const int32_t hw = (int32_t)std::max(1u, std::thread::hardware_concurrency());
    const int32_t n_thr = std::max(1, std::min({16, hw, n_edges / 64}));
    auto work = [&](int32_t t) {
        std::vector<int32_t> stamp(n_nodes, 0), local(n_nodes, 0);
        int32_t tick = 0;
        RefineScratch rs;
        for (int32_t e = t; e < n_edges; e += n_thr) {
            Lightcone lc = canonical_lightcone(g, edges[2 * e], edges[2 * e + 1], p, stamp, tick, local, rs);
            std::vector<int32_t>& key = keys[e];
            key = {lc.n, (int32_t)lc.edges.size(), lc.i, lc.j};
            for (auto& x : lc.edges) {
                key.push_back(x.first);
                key.push_back(x.second);
            }
        }
    };
    if (n_thr == 1) {
        work(0);
    } else {
        std::vector<std::thread> threads;
        for (int32_t t = 0; t < n_thr; ++t) threads.emplace_back(work, t);
        for (auto& th : threads) th.join();
    }
    std::map<std::vector<int32_t>, int32_t> classes;
    int64_t pos = 0;
    for (int32_t e = 0; e < n_edges; ++e) {
        const std::vector<int32_t>& key = keys[e];
        auto f = classes.find(key);
        if (f == classes.end()) {
            const int32_t id = (int32_t)classes.size();
            classes.emplace(key, id);
            if (pos + (int64_t)key.size() <= capacity) std::copy(key.begin(), key.end(), class_words + pos);
            pos += (int64_t)key.size();
            edge_class[e] = id;
        } else {
            edge_class[e] = f->second;
        }
    }
    *n_classes = (int32_t)classes.size();
    *used = pos;
    if (pos > capacity)
        return qb_internal::fail(QB_ERR_UNSUPPORTED, "qb_lightcone_classes: class_words too small (see *used)");
    return QB_OK;
}

int qb_compile_lightcone(const int32_t* lightcone, int32_t p, const qb_lightcone_plan_opts* o, int32_t* words,
                         int64_t capacity, int64_t* used, qb_lightcone_plan_info* info) {
    if (!lightcone || !o || !used || !info || p < 0 || capacity < 0 || (capacity > 0 && !words))
        return qb_internal::fail(QB_ERR_INVALID, "qb_compile_lightcone: null pointer or bad size");
    if (o->composer < 0 || o->composer > 2 || o->ordering < 0 || o->ordering > 1 || o->slicer < 0 || o->slicer > 2 ||
        (o->lanes != 64 && o->lanes != 128 && o->lanes != 256 && o->lanes != 512 && o->lanes != 1024) ||
        o->lds_tensor_limit < 0)
        return qb_internal::fail(QB_ERR_INVALID, "qb_compile_lightcone: bad options");
    Lightcone lc;
    lc.n = lightcone[0];
    const int32_t ne = lightcone[1];
    lc.i = lightcone[2];
    lc.j = lightcone[3];
    if (lc.n < 2 || ne < 1 || lc.i < 0 || lc.j >= lc.n || lc.i >= lc.j)
        return qb_internal::fail(QB_ERR_INVALID, "qb_compile_lightcone: bad lightcone header");
    for (int32_t k = 0; k < ne; ++k) {
        const int32_t a = lightcone[4 + 2 * k], b = lightcone[5 + 2 * k];
        if (a < 0 || b < 0 || a >= lc.n || b >= lc.n || a == b)
            return qb_internal::fail(QB_ERR_INVALID, "qb_compile_lightcone: edge outside the lightcone");
        lc.edges.emplace_back(std::min(a, b), std::max(a, b));
    }
    std::sort(lc.edges.begin(), lc.edges.end());
    lc.edges.erase(std::unique(lc.edges.begin(), lc.edges.end()), lc.edges.end());
    const std::vector<Gate> circuit = lightcone_circuit(lc, p, o->composer);
    const Network net = circ2buckets(lc.n, circuit);
    const LineGraph lg = line_graph(net);
    OrderOpts oo;
    oo.ordering = o->ordering;
    oo.temp = o->rg_temp;
    oo.repeats = o->rg_repeats;
    oo.seed = o->rg_seed;
    Ordered ord;
    int rc = optimize(lg, oo, o->slicer, o->max_tw, o->tw_bias, ord);
    if (rc) return rc;
    CompileOpts co;
    co.lanes = o->lanes;
    co.lds_tensor_limit = o->lds_tensor_limit;
    co.lds_capacity = o->lds_capacity;
    PlanOut po;
    rc = compile_plan(net, ord, lg, co, po);
    if (rc) return rc;
    *info = po.info;
    *used = (int64_t)po.words.size();
    if (*used > capacity)
        return qb_internal::fail(QB_ERR_UNSUPPORTED, "qb_compile_lightcone: words too small (see *used)");
    std::copy(po.words.begin(), po.words.end(), words);
    return QB_OK;
}

}  // extern "C"
