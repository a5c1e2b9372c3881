// Host symbolic analysis: CSR pattern, atom-graph ordering, tile structure.
//
// build_pattern restates BlochPattern.from_bundle (src/aospectrum/sparse.py:63-140)
// without a global sort: blocks are grouped per (row atom, column atom) pair,
// and because every block covers its whole atom pair, the CSR row of orbital r
// of atom i is the concatenation of the column ranges of atom i's pairs in
// column-atom order.  Duplicate keys (same atom pair, several lattice shifts)
// keep the block order of the input, as the reference's stable argsort does.
#include "symbolic.h"

#include <algorithm>
#include <climits>
#include <array>
#include <cstring>
#include <functional>
#include <map>
#include <numeric>

namespace aosx {

namespace {

struct Bitset {
    int64_t words = 0;
    std::vector<uint64_t> bits;
    void resize(int64_t rows, int64_t cols) {
        words = (cols + 63) / 64;
        bits.assign((size_t)(rows * words), 0ull);
    }
    uint64_t *row(int64_t r) { return bits.data() + r * words; }
    const uint64_t *row(int64_t r) const { return bits.data() + r * words; }
    void set(int64_t r, int64_t c) { row(r)[c >> 6] |= 1ull << (c & 63); }
    void clear(int64_t r, int64_t c) { row(r)[c >> 6] &= ~(1ull << (c & 63)); }
    bool test(int64_t r, int64_t c) const { return (row(r)[c >> 6] >> (c & 63)) & 1ull; }
};

template <typename F>
void for_each_bit(const uint64_t *w, int64_t words, int64_t from, F f) {
    for (int64_t q = from >> 6; q < words; ++q) {
        uint64_t v = w[q];
        if (q == (from >> 6)) v &= ~0ull << (from & 63);
        while (v) {
            int b = __builtin_ctzll(v);
            f(q * 64 + b);
            v &= v - 1;
        }
    }
}

}  // namespace

std::string build_pattern(const BlockTable &bt, Pattern &out, int &code) {
    const int64_t n = bt.n_orbitals, na = bt.n_atoms, nb = bt.n_blocks;
    code = 50;
    if (n <= 0 || na <= 0 || nb <= 0) return "empty bundle";
    if (n > INT32_MAX) { code = 11; return "orbital count exceeds the int32 CSR ABI"; }
    if (bt.atom_offsets[0] != 0 || bt.atom_offsets[na] != n) return "atom offsets do not span the orbitals";
    for (int64_t a = 0; a < na; ++a)
        if (bt.atom_offsets[a + 1] <= bt.atom_offsets[a]) return "each atom must own at least one AO";
    if (bt.block_offsets[0] != 0) return "block offsets must start at zero";
    for (int64_t b = 0; b < nb; ++b) {
        const int32_t i = bt.row_atom[b], j = bt.col_atom[b];
        if (i < 0 || j < 0 || i >= na || j >= na) return "operator block references an unknown atom";
        const int64_t ni = bt.atom_offsets[i + 1] - bt.atom_offsets[i];
        const int64_t nj = bt.atom_offsets[j + 1] - bt.atom_offsets[j];
        if (bt.block_offsets[b + 1] - bt.block_offsets[b] != ni * nj)
            return "operator block offsets do not match the AO layout";
    }

    // unique lattice shifts, ascending lexicographic (np.unique(axis=0) order)
    std::vector<std::array<int32_t, 3>> sh(nb);
    for (int64_t b = 0; b < nb; ++b)
        sh[b] = {bt.lattice_shift[3 * b], bt.lattice_shift[3 * b + 1], bt.lattice_shift[3 * b + 2]};
    std::vector<std::array<int32_t, 3>> uniq(sh);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    std::vector<int32_t> bgroup(nb);
    for (int64_t b = 0; b < nb; ++b)
        bgroup[b] = (int32_t)(std::lower_bound(uniq.begin(), uniq.end(), sh[b]) - uniq.begin());
    out.shifts.resize(uniq.size() * 3);
    for (size_t g = 0; g < uniq.size(); ++g)
        for (int d = 0; d < 3; ++d) out.shifts[3 * g + d] = uniq[g][d];

    // atom pairs, sorted by (row atom, col atom); block lists in input order
    std::vector<int64_t> border(nb);
    std::iota(border.begin(), border.end(), 0);
    std::stable_sort(border.begin(), border.end(), [&](int64_t x, int64_t y) {
        if (bt.row_atom[x] != bt.row_atom[y]) return bt.row_atom[x] < bt.row_atom[y];
        return bt.col_atom[x] < bt.col_atom[y];
    });
    std::vector<int64_t> pair_ptr{0};
    std::vector<int32_t> pair_i, pair_j;
    for (int64_t q = 0; q < nb; ++q) {
        const int64_t b = border[q];
        if (q == 0 || bt.row_atom[b] != pair_i.back() || bt.col_atom[b] != pair_j.back()) {
            if (q) pair_ptr.push_back(q);
            pair_i.push_back(bt.row_atom[b]);
            pair_j.push_back(bt.col_atom[b]);
        }
    }
    pair_ptr.push_back(nb);
    const int64_t np_ = (int64_t)pair_i.size();
    std::vector<int64_t> row_pair_ptr(na + 1, 0);
    for (int64_t p = 0; p < np_; ++p) row_pair_ptr[pair_i[p] + 1]++;
    for (int64_t a = 0; a < na; ++a) row_pair_ptr[a + 1] += row_pair_ptr[a];
    std::vector<int64_t> colpos(np_), rowlen(na, 0);
    for (int64_t a = 0; a < na; ++a)
        for (int64_t p = row_pair_ptr[a]; p < row_pair_ptr[a + 1]; ++p) {
            colpos[p] = rowlen[a];
            rowlen[a] += bt.atom_offsets[pair_j[p] + 1] - bt.atom_offsets[pair_j[p]];
        }
    // Hermitian topology: pair (j, i) must exist for every pair (i, j)
    auto find_pair = [&](int32_t i, int32_t j) -> int64_t {
        auto lo = pair_j.begin() + row_pair_ptr[i], hi = pair_j.begin() + row_pair_ptr[i + 1];
        auto it = std::lower_bound(lo, hi, j);
        return (it != hi && *it == j) ? (int64_t)(it - pair_j.begin()) : -1;
    };
    std::vector<int64_t> pmirror(np_);
    for (int64_t p = 0; p < np_; ++p) {
        pmirror[p] = find_pair(pair_j[p], pair_i[p]);
        if (pmirror[p] < 0) { code = 51; return "operator topology is not Hermitian"; }
    }

    out.n = n;
    out.indptr.assign(n + 1, 0);
    for (int64_t a = 0; a < na; ++a)
        for (int64_t r = bt.atom_offsets[a]; r < bt.atom_offsets[a + 1]; ++r) out.indptr[r + 1] = (int32_t)rowlen[a];
    int64_t total = 0;
    for (int64_t r = 0; r < n; ++r) {
        total += out.indptr[r + 1];
        if (total > INT32_MAX) { code = 11; return "pattern nnz exceeds the int32 CSR ABI"; }
        out.indptr[r + 1] = (int32_t)total;
    }
    out.indices.resize(total);
    out.rows.resize(total);
    out.mirror.resize(total);
    out.starts.resize(total + 1);
    int64_t n_source = 0;
    for (int64_t p = 0; p < np_; ++p) {
        const int64_t ni = bt.atom_offsets[pair_i[p] + 1] - bt.atom_offsets[pair_i[p]];
        const int64_t nj = bt.atom_offsets[pair_j[p] + 1] - bt.atom_offsets[pair_j[p]];
        n_source += ni * nj * (pair_ptr[p + 1] - pair_ptr[p]);
    }
    out.source.resize(n_source);
    out.group.resize(n_source);
    // entries are written row by row so that `starts` is monotone
    int64_t s = 0;
    for (int64_t a = 0; a < na; ++a) {
        const int64_t oi = bt.atom_offsets[a], ni = bt.atom_offsets[a + 1] - oi;
        for (int64_t r = 0; r < ni; ++r) {
            int64_t pos = out.indptr[oi + r];
            for (int64_t p = row_pair_ptr[a]; p < row_pair_ptr[a + 1]; ++p) {
                const int32_t j = pair_j[p];
                const int64_t oj = bt.atom_offsets[j], nj = bt.atom_offsets[j + 1] - oj;
                const int64_t mp = pmirror[p];
                for (int64_t c = 0; c < nj; ++c, ++pos) {
                    out.indices[pos] = (int32_t)(oj + c);
                    out.rows[pos] = (int32_t)(oi + r);
                    out.mirror[pos] = (int32_t)(out.indptr[oj + c] + colpos[mp] + r);
                    out.starts[pos] = s;
                    for (int64_t q = pair_ptr[p]; q < pair_ptr[p + 1]; ++q) {
                        const int64_t b = border[q];
                        out.source[s] = bt.block_offsets[b] + r * nj + c;
                        out.group[s] = bgroup[b];
                        ++s;
                    }
                }
            }
        }
    }
    out.starts[total] = s;
    for (int64_t r = 0; r < n; ++r) {
        bool diag = false;
        for (int32_t p = out.indptr[r]; p < out.indptr[r + 1]; ++p) diag |= out.indices[p] == r;
        if (!diag) { code = 33; return "pattern is missing a diagonal entry"; }
    }
    code = 0;
    return "";
}

// Minimum external degree on the atom quotient graph (degrees weighted by AO
// counts), with explicit fill kept in bitsets.  Ties go to the lowest atom.
std::string order_atoms(const BlockTable &bt, int ordering, std::vector<int32_t> &order,
                        int64_t &atom_fill_entries) {
    const int64_t na = bt.n_atoms;
    std::vector<int64_t> w(na);
    for (int64_t a = 0; a < na; ++a) w[a] = bt.atom_offsets[a + 1] - bt.atom_offsets[a];
    Bitset adj;
    adj.resize(na, na);
    for (int64_t b = 0; b < bt.n_blocks; ++b) {
        const int32_t i = bt.row_atom[b], j = bt.col_atom[b];
        if (i != j) { adj.set(i, j); adj.set(j, i); }
    }
    // weight classes for popcount-based degrees
    std::map<int64_t, int> cls_of;
    for (int64_t a = 0; a < na; ++a) cls_of.emplace(w[a], (int)cls_of.size());
    std::vector<int64_t> cls_w(cls_of.size());
    for (auto &kv : cls_of) cls_w[kv.second] = kv.first;
    Bitset cls;
    cls.resize((int64_t)cls_w.size(), na);
    for (int64_t a = 0; a < na; ++a) cls.set(cls_of[w[a]], a);
    auto degree = [&](int64_t u) {
        int64_t d = 0;
        const uint64_t *r = adj.row(u);
        for (size_t c = 0; c < cls_w.size(); ++c) {
            const uint64_t *m = cls.row((int64_t)c);
            int64_t cnt = 0;
            for (int64_t q = 0; q < adj.words; ++q) cnt += __builtin_popcountll(r[q] & m[q]);
            d += cnt * cls_w[c];
        }
        return d;
    };
    std::vector<int64_t> deg(na);
    for (int64_t a = 0; a < na; ++a) deg[a] = degree(a);
    std::vector<char> done(na, 0);
    order.clear();
    atom_fill_entries = 0;
    std::vector<int64_t> nbr;
    for (int64_t step = 0; step < na; ++step) {
        int64_t v = -1;
        if (ordering == 1) {
            v = step;
        } else {
            for (int64_t a = 0; a < na; ++a)
                if (!done[a] && (v < 0 || deg[a] < deg[v])) v = a;
        }
        order.push_back((int32_t)v);
        done[v] = 1;
        atom_fill_entries += w[v] * w[v] + 2 * w[v] * deg[v];
        nbr.clear();
        for_each_bit(adj.row(v), adj.words, 0, [&](int64_t u) { nbr.push_back(u); });
        const uint64_t *nv = adj.row(v);
        for (int64_t u : nbr) {
            uint64_t *ru = adj.row(u);
            for (int64_t q = 0; q < adj.words; ++q) ru[q] |= nv[q];
            adj.clear(u, u);
            adj.clear(u, v);
        }
        for (int64_t u : nbr) deg[u] = degree(u);
        std::fill(adj.row(v), adj.row(v) + adj.words, 0ull);
    }
    return "";
}

// Nested dissection on the atom graph.  A region is split at the median
// coordinate of its widest axis; the separator is a vertex cover of the graph
// edges that cross the split (greedy, highest remaining cut degree first), so
// periodic-image edges -- which the block table carries -- are separated too.
// Regions are ordered before their separator; regions of at most `leaf` atoms
// (or whose separator would take most of them) become leaves.  Inside every
// node (leaf or separator) the atoms follow recursive coordinate bisection
// (median of the widest axis, down to pairs), so that the consecutive atoms a
// tile holds are spatial neighbours: a separator plane is cut into compact
// patches instead of strips along its normal, and the tiles of a node couple
// to few tiles of the nodes around it.  A separator's atoms are then ordered
// (stably) by the node of their earliest already-ordered neighbour, so that
// the atoms a lower node couples to sit in few consecutive tiles of the
// separator (tile-level factorisation flops: bulk3d 3.361e12 -> 3.295e12,
// 52,000-orbital slab 3.327e13 -> 3.293e13).
std::string order_atoms_nd(const BlockTable &bt, const double *pos, std::vector<int32_t> &order,
                           std::vector<int32_t> &groups) {
    const int64_t na = bt.n_atoms;
    std::vector<std::vector<int32_t>> adj(na);
    for (int64_t b = 0; b < bt.n_blocks; ++b) {
        const int32_t i = bt.row_atom[b], j = bt.col_atom[b];
        if (i != j) adj[i].push_back(j);
    }
    for (auto &a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    std::vector<int32_t> side(na, -1);  // region-local marks: 0 A, 1 B, -1 outside
    std::vector<int32_t> cutdeg(na, 0);
    std::vector<char> covered(na, 0);
    const size_t leaf = 32;
    order.clear();
    order.reserve(na);
    groups.assign(1, 0);
    std::function<void(int32_t *, int32_t *)> bisect = [&](int32_t *b, int32_t *e) {
        if (e - b <= 2) return;
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (const int32_t *p = b; p < e; ++p)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], pos[3 * *p + d]);
                hi[d] = std::max(hi[d], pos[3 * *p + d]);
            }
        int ax = 0;
        for (int d = 1; d < 3; ++d)
            if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
        int32_t *mid = b + (e - b) / 2;
        std::nth_element(b, mid, e, [&](int32_t x, int32_t y) {
            return pos[3 * x + ax] < pos[3 * y + ax] || (pos[3 * x + ax] == pos[3 * y + ax] && x < y);
        });
        bisect(b, mid);
        bisect(mid, e);
    };
    std::vector<int32_t> where(na, -1);  // elimination position of the atoms ordered so far
    auto emit = [&](std::vector<int32_t> &node, bool separator) {
        bisect(node.data(), node.data() + node.size());
        if (separator) {
            std::vector<std::pair<int64_t, int32_t>> kv(node.size());
            for (size_t q = 0; q < node.size(); ++q) {
                int32_t first = -1;
                for (int32_t nb : adj[node[q]])
                    if (where[nb] >= 0 && (first < 0 || where[nb] < first)) first = where[nb];
                const int64_t g = first < 0 ? INT64_MAX
                                            : std::upper_bound(groups.begin(), groups.end(), first) - groups.begin();
                kv[q] = {g, node[q]};
            }
            std::stable_sort(kv.begin(), kv.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
            for (size_t q = 0; q < node.size(); ++q) node[q] = kv[q].second;
        }
        for (size_t q = 0; q < node.size(); ++q) where[node[q]] = (int32_t)(order.size() + q);
        order.insert(order.end(), node.begin(), node.end());
        groups.push_back((int32_t)order.size());
    };
    struct Frame {
        std::vector<int32_t> set;
    };
    // explicit recursion: emit(region) = emit(A'), emit(B'), separator
    std::function<void(std::vector<int32_t> &)> nd = [&](std::vector<int32_t> &set) {
        if (set.size() <= leaf) {
            emit(set, false);
            return;
        }
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int32_t a : set)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], pos[3 * a + d]);
                hi[d] = std::max(hi[d], pos[3 * a + d]);
            }
        int axis = 0;
        for (int d = 1; d < 3; ++d)
            if (hi[d] - lo[d] > hi[axis] - lo[axis]) axis = d;
        std::stable_sort(set.begin(), set.end(),
                         [&](int32_t x, int32_t y) { return pos[3 * x + axis] < pos[3 * y + axis]; });
        const size_t half = set.size() / 2;
        for (size_t q = 0; q < set.size(); ++q) side[set[q]] = q < half ? 0 : 1;
        // cut graph and greedy vertex cover
        std::vector<int32_t> cutv;
        for (int32_t a : set) {
            int c = 0;
            for (int32_t nb : adj[a])
                if (side[nb] >= 0 && side[nb] != side[a]) ++c;
            cutdeg[a] = c;
            if (c) cutv.push_back(a);
        }
        std::vector<int32_t> sep;
        while (true) {
            int32_t best = -1;
            for (int32_t v : cutv)
                if (!covered[v] && cutdeg[v] > 0 && (best < 0 || cutdeg[v] > cutdeg[best])) best = v;
            if (best < 0) break;
            covered[best] = 1;
            sep.push_back(best);
            for (int32_t nb : adj[best])
                if (side[nb] >= 0 && side[nb] != side[best] && !covered[nb]) cutdeg[nb]--;
            cutdeg[best] = 0;
        }
        std::vector<int32_t> A, B;
        for (int32_t a : set)
            if (!covered[a]) (side[a] == 0 ? A : B).push_back(a);
        for (int32_t a : set) {
            side[a] = -1;
            cutdeg[a] = 0;
        }
        for (int32_t a : sep) covered[a] = 0;
        if (A.empty() || B.empty() || sep.size() * 10 > set.size() * 9) {
            emit(set, false);
            return;
        }
        nd(A);
        nd(B);
        emit(sep, true);
    };
    std::vector<int32_t> all(na);
    std::iota(all.begin(), all.end(), 0);
    nd(all);
    if ((int64_t)order.size() != na) return "nested dissection lost atoms";
    return "";
}

// nnz(L+U) of the atom-level symbolic factor for a given elimination order.
int64_t atom_fill(const BlockTable &bt, const std::vector<int32_t> &order) {
    const int64_t na = bt.n_atoms;
    std::vector<int32_t> rank(na);
    for (int64_t q = 0; q < na; ++q) rank[order[q]] = (int32_t)q;
    Bitset g;
    g.resize(na, na);
    for (int64_t b = 0; b < bt.n_blocks; ++b) {
        const int32_t i = rank[bt.row_atom[b]], j = rank[bt.col_atom[b]];
        g.set(i, j);
        g.set(j, i);
    }
    int64_t fill = 0;
    std::vector<int64_t> nbr;
    for (int64_t k = 0; k < na; ++k) {
        const int64_t wk = bt.atom_offsets[order[k] + 1] - bt.atom_offsets[order[k]];
        nbr.clear();
        for_each_bit(g.row(k), g.words, k + 1, [&](int64_t i) { nbr.push_back(i); });
        int64_t deg = 0;
        for (int64_t i : nbr) deg += bt.atom_offsets[order[i] + 1] - bt.atom_offsets[order[i]];
        fill += wk * wk + 2 * wk * deg;
        const uint64_t *rk = g.row(k);
        for (int64_t i : nbr) {
            uint64_t *ri = g.row(i);
            for (int64_t q = (k + 1) >> 6; q < g.words; ++q) ri[q] |= rk[q];
        }
    }
    return fill;
}

std::string build_tiles(const BlockTable &bt, const Pattern &pat, const std::vector<int32_t> &atom_order,
                        const std::vector<int32_t> &groups, int tile, TileStructure &ts) {
    const int64_t n = bt.n_orbitals, na = bt.n_atoms;
    ts.tile = tile;
    ts.n = n;
    ts.atom_order = atom_order;
    ts.perm.clear();
    ts.perm.reserve(n);
    std::vector<int64_t> atom_start(na);
    std::vector<char> first(atom_order.size() + 1, 0);  // a group starts at this position
    for (size_t g = 1; g + 1 < groups.size(); ++g) first[(size_t)groups[g]] = 1;
    for (size_t q = 0; q < atom_order.size(); ++q) {
        if (first[q])
            while (ts.perm.size() % tile) ts.perm.push_back(-1);
        const int32_t a = atom_order[q];
        atom_start[a] = (int64_t)ts.perm.size();
        for (int64_t r = bt.atom_offsets[a]; r < bt.atom_offsets[a + 1]; ++r) ts.perm.push_back((int32_t)r);
    }
    while (ts.perm.size() % tile) ts.perm.push_back(-1);
    ts.n_rows = (int64_t)ts.perm.size();
    if (ts.n_rows > INT32_MAX) return "factor rows exceed int32";
    ts.iperm.assign(n, 0);
    ts.padding_rows.clear();
    for (int64_t q = 0; q < ts.n_rows; ++q) {
        if (ts.perm[q] >= 0) ts.iperm[ts.perm[q]] = (int32_t)q;
        else ts.padding_rows.push_back((int32_t)q);
    }
    const int nt = (int)(ts.n_rows / tile);
    ts.n_diag = nt;
    Bitset g;
    g.resize(nt, nt);
    for (int64_t b = 0; b < bt.n_blocks; ++b) {
        const int32_t i = bt.row_atom[b], j = bt.col_atom[b];
        const int64_t r0 = atom_start[i], r1 = r0 + (bt.atom_offsets[i + 1] - bt.atom_offsets[i]) - 1;
        const int64_t c0 = atom_start[j], c1 = c0 + (bt.atom_offsets[j + 1] - bt.atom_offsets[j]) - 1;
        for (int64_t I = r0 / tile; I <= r1 / tile; ++I)
            for (int64_t J = c0 / tile; J <= c1 / tile; ++J) { g.set(I, J); g.set(J, I); }
    }
    for (int k = 0; k < nt; ++k) g.set(k, k);
    // symbolic right-looking elimination on the tile graph
    ts.next_ptr.assign(nt + 1, 0);
    ts.next_idx.clear();
    std::vector<int64_t> sk;
    for (int k = 0; k < nt; ++k) {
        sk.clear();
        for_each_bit(g.row(k), g.words, k + 1, [&](int64_t i) { sk.push_back(i); });
        const uint64_t *rk = g.row(k);
        for (int64_t i : sk) {
            uint64_t *ri = g.row(i);
            for (int64_t q = (k + 1) >> 6; q < g.words; ++q) ri[q] |= rk[q];
        }
        for (int64_t i : sk) ts.next_idx.push_back((int32_t)i);
        ts.next_ptr[k + 1] = (int32_t)ts.next_idx.size();
    }
    // predecessors (transpose of next)
    ts.prev_ptr.assign(nt + 1, 0);
    for (int32_t i : ts.next_idx) ts.prev_ptr[i + 1]++;
    for (int k = 0; k < nt; ++k) ts.prev_ptr[k + 1] += ts.prev_ptr[k];
    ts.prev_idx.assign(ts.next_idx.size(), 0);
    {
        std::vector<int32_t> fillp(ts.prev_ptr.begin(), ts.prev_ptr.end() - 1);
        for (int k = 0; k < nt; ++k)
            for (int32_t q = ts.next_ptr[k]; q < ts.next_ptr[k + 1]; ++q) ts.prev_idx[fillp[ts.next_idx[q]]++] = k;
    }
    // tile ids: per step k the diagonal, then L(i,k), then U(k,i)
    ts.tile_id.assign((size_t)nt * nt, -1);
    int64_t id = 0;
    const double t3 = (double)tile * tile * tile;
    ts.factor_flops = 0.0;
    ts.schur_flops = 0.0;
    for (int k = 0; k < nt; ++k) {
        ts.tile_id[(size_t)k * nt + k] = (int32_t)id++;
        for (int32_t q = ts.next_ptr[k]; q < ts.next_ptr[k + 1]; ++q)
            ts.tile_id[(size_t)ts.next_idx[q] * nt + k] = (int32_t)id++;
        for (int32_t q = ts.next_ptr[k]; q < ts.next_ptr[k + 1]; ++q)
            ts.tile_id[(size_t)k * nt + ts.next_idx[q]] = (int32_t)id++;
        const double s = ts.next_ptr[k + 1] - ts.next_ptr[k];
        ts.factor_flops += 8.0 * t3 * (2.0 / 3.0 + 2.0 * s + s * s);
        ts.schur_flops += 8.0 * t3 * s * s;
    }
    ts.n_tiles = id;
    if (id > INT32_MAX) return "tile count exceeds int32";
    (void)pat;
    return "";
}

}  // namespace aosx
