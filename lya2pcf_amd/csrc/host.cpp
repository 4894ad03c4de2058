// Host-side runtime of the lya2pcf library: forest-pair list and work
// decomposition.  Built with -ffp-contract=off so that the neighbour test
// evaluates (x1*x2 + y1*y2) + z1*z2 exactly like the FP64 dot product of
// reference forest_class.py:70-76 (IEEE double, left to right, no FMA).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <vector>
#include <omp.h>

#include "../../include/lya2pcf.h"
#include "errors.h"

namespace {

// Cubic grid over the unit-vector cube [-1,1]^3 with cell edge >= the chord of
// the cut angle, so every neighbour of a forest sits in one of the 27 cells
// around it.
struct Grid {
    double h;
    int64_t m;
    std::unordered_map<int64_t, std::pair<int32_t, int32_t>> cells;  // key -> [begin,end) in order
    std::vector<int32_t> order;

    int64_t coord(double u) const {
        int64_t c = (int64_t)std::floor((u + 1.0) / h);
        return std::min<int64_t>(std::max<int64_t>(c, 0), m - 1);
    }
    int64_t key(int64_t a, int64_t b, int64_t c) const { return (a * m + b) * m + c; }
};

}  // namespace

extern "C" int lya_build_pairs(int32_t n, const double* ux, const double* uy, const double* uz,
                               const double* ra, const int64_t* name, double cos_max, int64_t* nb_off,
                               int32_t* nb_idx, int64_t cap, int64_t* n_pairs) {
    if (n < 0 || !ux || !uy || !uz || !ra || !name || !nb_off || !n_pairs)
        return lya_fail(LYA_ERR_ARG, "lya_build_pairs: null argument or n < 0");
    nb_off[0] = 0;
    if (n == 0) {
        *n_pairs = 0;
        return LYA_OK;
    }
    // chord length for the cut angle, with a margin for rounding
    double cm = std::max(-1.0, std::min(1.0, cos_max));
    double chord = std::sqrt(std::max(0.0, 2.0 - 2.0 * cm));
    Grid g;
    // cells no smaller than 2/4096: a coarser grid only adds candidates
    g.h = std::max(chord * 1.0001 + 1e-9, 2.0 / 4096.0);
    g.m = std::max<int64_t>(1, (int64_t)std::ceil(2.0 / g.h) + 1);

    std::vector<int64_t> keys(n);
    for (int32_t i = 0; i < n; ++i) keys[i] = g.key(g.coord(ux[i]), g.coord(uy[i]), g.coord(uz[i]));
    g.order.resize(n);
    std::iota(g.order.begin(), g.order.end(), 0);
    std::stable_sort(g.order.begin(), g.order.end(), [&](int32_t a, int32_t b) { return keys[a] < keys[b]; });
    for (int32_t s = 0; s < n;) {
        int32_t e = s;
        while (e < n && keys[g.order[e]] == keys[g.order[s]]) ++e;
        g.cells[keys[g.order[s]]] = {s, e};
        s = e;
    }

    auto visit = [&](int32_t i, auto&& emit) {
        int64_t a = g.coord(ux[i]), b = g.coord(uy[i]), c = g.coord(uz[i]);
        for (int64_t da = -1; da <= 1; ++da)
            for (int64_t db = -1; db <= 1; ++db)
                for (int64_t dc = -1; dc <= 1; ++dc) {
                    int64_t A = a + da, B = b + db, C = c + dc;
                    if (A < 0 || B < 0 || C < 0 || A >= g.m || B >= g.m || C >= g.m) continue;
                    auto it = g.cells.find(g.key(A, B, C));
                    if (it == g.cells.end()) continue;
                    for (int32_t q = it->second.first; q < it->second.second; ++q) {
                        int32_t j = g.order[q];
                        double mu = ux[i] * ux[j] + uy[i] * uy[j];
                        mu = mu + uz[i] * uz[j];
                        if (mu > cos_max && name[i] != name[j] && ra[i] < ra[j]) emit(j);
                    }
                }
    };

    std::vector<int64_t> counts(n, 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t i = 0; i < n; ++i) {
        int64_t k = 0;
        visit(i, [&](int32_t) { ++k; });
        counts[i] = k;
    }
    for (int32_t i = 0; i < n; ++i) nb_off[i + 1] = nb_off[i] + counts[i];
    *n_pairs = nb_off[n];
    if (!nb_idx) return LYA_OK;
    if (cap < nb_off[n]) return lya_fail(LYA_ERR_CAPACITY, "lya_build_pairs: capacity %lld < %lld pairs",
                                         (long long)cap, (long long)nb_off[n]);
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t i = 0; i < n; ++i) {
        int32_t* out = nb_idx + nb_off[i];
        int64_t k = 0;
        visit(i, [&](int32_t j) { out[k++] = j; });
        std::sort(out, out + k);
    }
    return LYA_OK;
}

namespace {

// Tiles: runs of consecutive items of one slot, cost <= tile_cost (a tile holds at
// least one item), ordered by decreasing cost (stable).
void make_tiles(int64_t n_items, const int32_t* items, const int64_t* cost, int64_t tile_cost,
                std::vector<std::pair<int32_t, int32_t>>& out) {
    struct Tile { int32_t b, e; int64_t cost; };
    std::vector<Tile> tl;
    for (int64_t s = 0; s < n_items;) {
        int64_t e = s;
        int64_t acc = 0;
        while (e < n_items && items[4 * e + 3] == items[4 * s + 3] && (e == s || acc + cost[e] <= tile_cost)) {
            acc += cost[e];
            ++e;
        }
        tl.push_back({(int32_t)s, (int32_t)e, acc});
        s = e;
    }
    std::stable_sort(tl.begin(), tl.end(), [](const Tile& a, const Tile& b) { return a.cost > b.cost; });
    out.resize(tl.size());
    for (size_t k = 0; k < tl.size(); ++k) out[k] = {tl[k].b, tl[k].e};
}

}  // namespace

extern "C" int lya_build_work(int32_t n_f1, const int32_t* f1_list, const int32_t* f1_slot,
                              const int64_t* pix_off, const int64_t* nb_off, const int32_t* nb_idx,
                              int64_t item_cost, int64_t tile_cost, int32_t* items, int32_t* tiles,
                              int32_t* n_items_out, int32_t* n_tiles_out, double* total_cost) {
    if (n_f1 < 0 || !f1_list || !f1_slot || !pix_off || !nb_off || !nb_idx || !n_items_out || !n_tiles_out)
        return lya_fail(LYA_ERR_ARG, "lya_build_work: null argument");
    if (item_cost < 1 || tile_cost < 1) return lya_fail(LYA_ERR_ARG, "lya_build_work: costs must be >= 1");
    for (int32_t k = 1; k < n_f1; ++k)
        if (f1_slot[k] < f1_slot[k - 1])
            return lya_fail(LYA_ERR_ARG, "lya_build_work: f1_list must be grouped by ascending slot");

    // Items: the neighbours of each forest 1 cut greedily into runs of cost <=
    // item_cost (at least one neighbour each).  Forests are independent, so the
    // cut runs in parallel: count, prefix sum, fill.
    auto cut = [&](int32_t k, auto&& emit) {
        const int32_t f = f1_list[k];
        const int64_t n1 = pix_off[f + 1] - pix_off[f];
        const int64_t b = nb_off[f], e = nb_off[f + 1];
        int64_t s = b, acc = 0, tot = 0;
        for (int64_t q = b; q < e; ++q) {
            const int32_t f2 = nb_idx[q];
            const int64_t c = n1 * (pix_off[f2 + 1] - pix_off[f2]);
            tot += c;
            if (q > s && acc + c > item_cost) {
                emit(f, s, q, acc);
                s = q;
                acc = 0;
            }
            acc += c;
        }
        if (e > s) emit(f, s, e, acc);
        return tot;
    };
    std::vector<int64_t> first(n_f1 + 1, 0);
    double total = 0.0;  // integer costs, exact in double below 2^53: any summation order
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total)
    for (int32_t k = 0; k < n_f1; ++k) {
        int64_t ni = 0;
        total += (double)cut(k, [&](int32_t, int64_t, int64_t, int64_t) { ++ni; });
        first[k + 1] = ni;
    }
    for (int32_t k = 0; k < n_f1; ++k) first[k + 1] += first[k];
    const int64_t ni = first[n_f1];
    if (ni > (int64_t)INT32_MAX) return lya_fail(LYA_ERR_CAPACITY, "lya_build_work: too many items");
    std::vector<int32_t> it(4 * (size_t)ni);
    std::vector<int64_t> cost(ni);
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t k = 0; k < n_f1; ++k) {
        int64_t o = first[k];
        cut(k, [&](int32_t f, int64_t s, int64_t e, int64_t acc) {
            it[4 * o + 0] = f;
            it[4 * o + 1] = (int32_t)s;
            it[4 * o + 2] = (int32_t)e;
            it[4 * o + 3] = f1_slot[k];
            cost[o] = acc;
            ++o;
        });
    }
    std::vector<std::pair<int32_t, int32_t>> tl;
    make_tiles(ni, it.data(), cost.data(), tile_cost, tl);

    *n_items_out = (int32_t)ni;
    *n_tiles_out = (int32_t)tl.size();
    if (total_cost) *total_cost = total;
    if (!items || !tiles) return LYA_OK;
    std::copy(it.begin(), it.end(), items);
    for (size_t k = 0; k < tl.size(); ++k) {
        tiles[2 * k + 0] = tl[k].first;
        tiles[2 * k + 1] = tl[k].second;
    }
    return LYA_OK;
}

extern "C" int lya_build_tiles(int64_t n_items, const int32_t* items, const int64_t* item_cost,
                               int64_t tile_cost, int32_t* tiles, int32_t* n_tiles) {
    if (n_items < 0 || !n_tiles || (n_items > 0 && (!items || !item_cost)))
        return lya_fail(LYA_ERR_ARG, "lya_build_tiles: null argument or n_items < 0");
    if (tile_cost < 1) return lya_fail(LYA_ERR_ARG, "lya_build_tiles: tile_cost must be >= 1");
    if (n_items > (int64_t)INT32_MAX) return lya_fail(LYA_ERR_CAPACITY, "lya_build_tiles: too many items");
    std::vector<std::pair<int32_t, int32_t>> tl;
    make_tiles(n_items, items, item_cost, tile_cost, tl);
    *n_tiles = (int32_t)tl.size();
    if (!tiles) return LYA_OK;
    for (size_t k = 0; k < tl.size(); ++k) {
        tiles[2 * k + 0] = tl[k].first;
        tiles[2 * k + 1] = tl[k].second;
    }
    return LYA_OK;
}
