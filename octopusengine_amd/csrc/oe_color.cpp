// Graph colouring, bit-exact with the reference (include/Tools/Graph.h).
//
// The reference stores adjacency as std::vector<std::multiset<int>>: a pair of
// vertices sharing k elements appears k times, "degree" is the multiset size,
// and every traversal visits neighbours in ascending id with repeats.  Here the
// same multiset is a CSR whose rows are sorted with repeats kept, which gives
// identical traversal order without a node allocation per entry.
//
// DSAT (Graph.h:321-371) keeps the uncoloured vertices in a std::set ordered by
// (saturation, remaining degree, id), largest first, re-keying a neighbour on
// every visit.  Here the same order is served by an indexed max-heap whose
// entries are re-sifted in place, and saturation sets are bitsets.
#include "oe_host.hpp"

#include <algorithm>
#include <queue>
#include <stdexcept>
#include <tuple>

namespace oe {

namespace {

MultiGraph finish_csr(int n, std::vector<int64_t>& counts, std::vector<int32_t>&& adj_unsorted) {
    MultiGraph g;
    g.n = n;
    g.off.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) g.off[i + 1] = g.off[i] + counts[i];
    g.adj = std::move(adj_unsorted);
    for (int i = 0; i < n; ++i) std::sort(g.adj.begin() + g.off[i], g.adj.begin() + g.off[i + 1]);
    return g;
}

}  // namespace

MultiGraph build_primal_graph(Element e, const std::vector<int32_t>& topo, int nb_vertices) {
    const int S = elem_nb_vertices(e);
    // The reference sizes the graph by the largest vertex id in the topology.
    int n = 0;
    for (int32_t v : topo) n = std::max(n, v + 1);
    if (nb_vertices > 0 && n > nb_vertices) throw std::invalid_argument("topology references unknown vertex");
    std::vector<int64_t> cnt(n, 0);
    for (size_t i = 0; i < topo.size(); i += S)
        for (int j = 0; j < S; ++j) cnt[topo[i + j]] += S - 1;
    std::vector<int64_t> pos(n + 1, 0);
    for (int i = 0; i < n; ++i) pos[i + 1] = pos[i] + cnt[i];
    std::vector<int32_t> adj(pos[n]);
    std::vector<int64_t> cur(pos.begin(), pos.end() - 1);
    for (size_t i = 0; i < topo.size(); i += S)
        for (int j = 0; j < S; ++j) {
            const int32_t a = topo[i + j];
            for (int k = 0; k < S; ++k)
                if (k != j) adj[cur[a]++] = topo[i + k];
        }
    return finish_csr(n, cnt, std::move(adj));
}

MultiGraph build_dual_vertex_graph(Element e, const std::vector<int32_t>& topo, int nb_vertices) {
    const int S = elem_nb_vertices(e);
    const int E = static_cast<int>(topo.size() / S);
    int nv = 0;
    for (int32_t v : topo) nv = std::max(nv, v + 1);
    (void)nb_vertices;
    // owners of every vertex in ascending element id (Graph.h:56-85)
    std::vector<int64_t> ocount(nv + 1, 0);
    for (int32_t v : topo) ocount[v + 1]++;
    for (int i = 0; i < nv; ++i) ocount[i + 1] += ocount[i];
    std::vector<int32_t> owners(topo.size());
    std::vector<int64_t> fill(ocount.begin(), ocount.end() - 1);
    for (int eid = 0; eid < E; ++eid)
        for (int i = 0; i < S; ++i) owners[fill[topo[eid * S + i]]++] = eid;
    std::vector<int64_t> cnt(E, 0);
    for (int v = 0; v < nv; ++v) {
        const int64_t d = ocount[v + 1] - ocount[v];
        for (int64_t j = ocount[v]; j < ocount[v + 1]; ++j) cnt[owners[j]] += d - 1;
    }
    std::vector<int64_t> pos(E + 1, 0);
    for (int i = 0; i < E; ++i) pos[i + 1] = pos[i] + cnt[i];
    std::vector<int32_t> adj(pos[E]);
    std::vector<int64_t> cur(pos.begin(), pos.end() - 1);
    for (int v = 0; v < nv; ++v)
        for (int64_t j = ocount[v]; j < ocount[v + 1]; ++j)
            for (int64_t k = j + 1; k < ocount[v + 1]; ++k) {
                const int32_t a = owners[j], b = owners[k];
                adj[cur[a]++] = b;
                adj[cur[b]++] = a;
            }
    return finish_csr(E, cnt, std::move(adj));
}

namespace {

// Max-heap of vertices keyed by (saturation, remaining degree, id) with
// in-place key updates (a vertex's key changes every time a neighbour is
// coloured; the heap position is tracked so the entry is re-sifted instead of
// re-inserted).
class DsatHeap {
public:
    DsatHeap(int n, const std::vector<int32_t>& sat, const std::vector<int32_t>& deg)
        : sat_(sat), deg_(deg), heap_(n), pos_(n) {
        for (int v = 0; v < n; ++v) heap_[v] = pos_[v] = v;
        for (int i = n / 2 - 1; i >= 0; --i) down(i);
    }
    bool empty() const { return heap_.empty(); }
    int pop() {
        const int top = heap_[0];
        swap_at(0, static_cast<int>(heap_.size()) - 1);
        heap_.pop_back();
        pos_[top] = -1;
        if (!heap_.empty()) down(0);
        return top;
    }
    // key of v only grows in saturation or shrinks in degree: sift both ways
    void update(int v) {
        const int i = pos_[v];
        if (i < 0) return;
        up(i);
        down(pos_[v]);
    }

private:
    bool above(int a, int b) const {
        if (sat_[a] != sat_[b]) return sat_[a] > sat_[b];
        if (deg_[a] != deg_[b]) return deg_[a] > deg_[b];
        return a > b;
    }
    void swap_at(int i, int j) {
        std::swap(heap_[i], heap_[j]);
        pos_[heap_[i]] = i;
        pos_[heap_[j]] = j;
    }
    void up(int i) {
        while (i > 0) {
            const int p = (i - 1) / 2;
            if (!above(heap_[i], heap_[p])) break;
            swap_at(i, p);
            i = p;
        }
    }
    void down(int i) {
        const int n = static_cast<int>(heap_.size());
        while (true) {
            const int l = 2 * i + 1, r = l + 1;
            int b = i;
            if (l < n && above(heap_[l], heap_[b])) b = l;
            if (r < n && above(heap_[r], heap_[b])) b = r;
            if (b == i) return;
            swap_at(i, b);
            i = b;
        }
    }
    const std::vector<int32_t>& sat_;
    const std::vector<int32_t>& deg_;
    std::vector<int> heap_, pos_;
};

}  // namespace

Coloration color_dsat(const MultiGraph& g) {
    const int n = g.n;
    constexpr int W = 4;   // saturation bitset: up to 256 colours
    std::vector<int32_t> colors(n, -1), deg(n), sat(n, 0);
    std::vector<uint64_t> bits(static_cast<size_t>(n) * W, 0);
    for (int v = 0; v < n; ++v) deg[v] = g.degree(v);
    DsatHeap heap(n, sat, deg);
    std::vector<char> used(n + 1, 0);
    int max_color = -1;
    while (!heap.empty()) {
        const int v = heap.pop();
        for (int64_t i = g.off[v]; i < g.off[v + 1]; ++i) {
            const int32_t c = colors[g.adj[i]];
            if (c != -1) used[c] = 1;
        }
        int c = 0;
        while (used[c]) ++c;
        for (int64_t i = g.off[v]; i < g.off[v + 1]; ++i) {
            const int32_t cc = colors[g.adj[i]];
            if (cc != -1) used[cc] = 0;
        }
        if (c >= 64 * W) throw std::runtime_error("DSAT: more than 256 colours");
        colors[v] = c;
        max_color = std::max(max_color, c);
        for (int64_t i = g.off[v]; i < g.off[v + 1]; ++i) {
            const int32_t u = g.adj[i];
            if (colors[u] != -1) continue;
            uint64_t& word = bits[static_cast<size_t>(u) * W + (c >> 6)];
            const uint64_t bit = 1ull << (c & 63);
            if (!(word & bit)) {
                word |= bit;
                sat[u]++;
            }
            deg[u]--;
            heap.update(u);
        }
    }
    Coloration out;
    out.nb_color = max_color + 1;
    out.colors = std::move(colors);
    return out;
}

void color_reduction(const MultiGraph& g, Coloration& col, int target) {
    if (target <= 0) throw std::invalid_argument("ColorReduction needs target_color > 0");
    std::vector<int32_t>& colors = col.colors;
    col.nb_color = target;
    std::vector<int32_t> occ(target);
    for (int v = 0; v < g.n; ++v) {
        if (colors[v] < target) continue;
        std::fill(occ.begin(), occ.end(), 0);
        for (int64_t i = g.off[v]; i < g.off[v + 1]; ++i) {
            const int32_t c = colors[g.adj[i]];
            if (c >= target) continue;
            occ[c]++;
        }
        // least frequent colour, first one on ties (std::min_element)
        colors[v] = static_cast<int32_t>(std::min_element(occ.begin(), occ.end()) - occ.begin());
    }
}

Coloration color_greedy_slf(const MultiGraph& g) {
    const int n = g.n;
    using P = std::pair<int32_t, int32_t>;  // (degree, id), smallest first
    std::priority_queue<P, std::vector<P>, std::greater<P>> heap;
    std::vector<int32_t> deg(n);
    for (int v = 0; v < n; ++v) {
        deg[v] = g.degree(v);
        heap.emplace(deg[v], v);
    }
    std::vector<int32_t> order(n);
    std::vector<char> removed(n, 0);
    for (int i = n - 1; i >= 0; --i) {
        while (!heap.empty() && removed[heap.top().second]) heap.pop();
        const int32_t v = heap.top().second;
        heap.pop();
        removed[v] = 1;
        order[i] = v;
        for (int64_t k = g.off[v]; k < g.off[v + 1]; ++k) {
            const int32_t u = g.adj[k];
            if (!removed[u]) {
                deg[u]--;
                heap.emplace(deg[u], u);
            }
        }
    }
    std::vector<int32_t> colors(n, -1);
    int nb_color = 0;
    std::vector<char> used;
    for (int32_t v : order) {
        used.assign(nb_color + 1, 0);
        for (int64_t k = g.off[v]; k < g.off[v + 1]; ++k) {
            const int32_t c = colors[g.adj[k]];
            if (c != -1) used[c] = 1;
        }
        int c = 0;
        while (c < nb_color && used[c]) ++c;
        colors[v] = c;
        if (c == nb_color) ++nb_color;
    }
    Coloration out;
    out.nb_color = nb_color;
    out.colors = std::move(colors);
    return out;
}

int64_t count_conflicts(const MultiGraph& g, const Coloration& c) {
    int64_t k = 0;
    for (int v = 0; v < g.n; ++v)
        for (int64_t i = g.off[v]; i < g.off[v + 1]; ++i)
            if (c.colors[g.adj[i]] == c.colors[v]) k += 2;
    return k;
}

}  // namespace oe
