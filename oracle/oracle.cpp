// CPU oracle for the Vamana / FreshVamana hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: a deterministic,
// sequential C++17 restatement of the reference's algorithms, used by tests/,
// by the package's smoke check and as the timed CPU baseline in bench.py. The
// product path (the HIP library) never links, loads or calls it.
//
// Every function cites the reference file:line it restates. Two families:
//
//  * parth mode  -- parth/src/build_graph.cu, parth/src/engine.cu,
//                   parth/src/common.h (CandidateList), parth/src/device_funcs.cu.
//                   Pinned bit-exactly by the reference's shipped graphs
//                   parth/data/{sift,custom_sift,toy,vis}_initial_graph.bin.
//  * fresh mode  -- vamana/include/{search,kernels,bloom_filter}.cuh (G2).
//                   The reference's GPU order of equal-distance items is
//                   atomic-order dependent; this restatement fixes the order
//                   a sequential loop produces (adjacency order / (dist, id)).
//                   Parity of fresh mode is pinned by invariants and by the
//                   closed-form bloom hashes, not by a reference fixture.
//
// Distances: squared L2 summed over dimensions in index order. `dist_mode`
// selects how each term is accumulated: 0 = fused multiply-add chain
// acc = fma(diff, diff, acc) (what nvcc emits for `dist += diff * diff` in
// device code with its default --fmad=true), 1 = separately rounded multiply
// and add (host code, no contraction). On integer-valued data below 2^24 the
// two are identical and exact.
//
// Build: g++ -O2 -std=c++17 -fopenmp -mfma -ffp-contract=off -shared -fPIC
// (see oracle/Makefile). -ffp-contract=off keeps every a*b+c that is not an
// explicit fmaf() separately rounded.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// ---------------------------------------------------------------------------
// Planted mistakes, for the tests that check that the fresh-mode checks can
// fail (0 = none; never set outside those tests). Each value breaks one rule:
//  1 prune keeps p' on alpha * d(p*, p') == d(node, p') (`<` for `<=`);
//  2 search merge puts new neighbours before worklist entries on ties;
//  3 search does not skip the query's own node id;
//  4 reverse candidates capped at M + 1; 5 at M - 1;
//  6 out-prune candidates put the current neighbours first on ties;
//  7 reverse-prune candidates put the current neighbours first on ties.
// ---------------------------------------------------------------------------
static int g_mut = 0;
void orc_set_mutation(int m) { g_mut = m; }

// ---------------------------------------------------------------------------
// Distances
// ---------------------------------------------------------------------------

// parth/src/device_funcs.cu:17-49 (device `distance<D>`), vamana/include/
// kernels.cuh:45-51 (partials), utils.cuh:150-158 (`l2_distance_sq`).
static inline float l2sq(const float* a, const float* b, int d, int dist_mode) {
    float acc = 0.0f;
    if (dist_mode == 0) {
        for (int i = 0; i < d; ++i) {
            float diff = a[i] - b[i];
            acc = fmaf(diff, diff, acc);
        }
    } else if (dist_mode == 1) {
        for (int i = 0; i < d; ++i) {
            float diff = a[i] - b[i];
            float sq = diff * diff;
            acc = acc + sq;
        }
    } else {
        // dist_mode 2: the D=2 specialisation written as d0*d0 + d1*d1
        // (parth/src/device_funcs.cu:43-48) with the first product fused
        // into the add: fma(d0, d0, d1*d1). For d != 2 this falls back to
        // the fused chain.
        if (d == 2) {
            float d0 = a[0] - b[0], d1 = a[1] - b[1];
            float p1 = d1 * d1;
            return fmaf(d0, d0, p1);
        }
        return l2sq(a, b, d, 0);
    }
    return acc;
}

// dist_mode 3 for d == 2: fma(d1, d1, d0*d0).
static inline float l2sq_m(const float* a, const float* b, int d, int dist_mode) {
    if (dist_mode == 3) {
        if (d == 2) {
            float d0 = a[0] - b[0], d1 = a[1] - b[1];
            float p0 = d0 * d0;
            return fmaf(d1, d1, p0);
        }
        return l2sq(a, b, d, 0);
    }
    return l2sq(a, b, d, dist_mode);
}

float orc_l2sq(const float* a, const float* b, int d, int dist_mode) {
    return l2sq_m(a, b, d, dist_mode);
}

// ---------------------------------------------------------------------------
// parth mode
// ---------------------------------------------------------------------------

// parth/src/build_graph.cu:28-56 (`randomInitKernel`): per node a 64-bit LCG
// seeded with seed + node id; rejection of self and duplicates; row padded
// with -1 to `row_width` (the reference's MAX_DEGREE).
void orc_random_init(int64_t n, int R, int row_width, uint64_t seed, int32_t* adj) {
#pragma omp parallel for schedule(static)
    for (int64_t tid = 0; tid < n; ++tid) {
        uint64_t state = seed + (uint64_t)tid;
        int32_t* row = adj + tid * row_width;
        for (int i = 0; i < row_width; ++i) row[i] = -1;
        int count = 0;
        while (count < R) {
            state = state * 6364136223846793005ULL + 1442695040888963407ULL;
            int32_t target = (int32_t)((state >> 32) % (uint64_t)n);
            if (target != (int32_t)tid) {
                bool exists = false;
                for (int i = 0; i < count; ++i)
                    if (row[i] == target) { exists = true; break; }
                if (!exists) row[count++] = target;
            }
        }
    }
}

// A caller's initial graph [n x init_width] as rows of width `row_width`:
// entries up to the first -1 in order, repeated ids and the node itself
// dropped, at most R kept, -1 padded (the rule of vmn_build_from).
void orc_init_rows(const int32_t* init, int64_t n, int init_width, int R, int row_width,
                   int32_t* adj) {
    for (int64_t i = 0; i < n; ++i) {
        int32_t* row = adj + i * row_width;
        for (int j = 0; j < row_width; ++j) row[j] = -1;
        int w = 0;
        for (int j = 0; j < init_width && w < R; ++j) {
            const int32_t v = init[i * init_width + j];
            if (v == -1) break;
            if (v == (int32_t)i) continue;
            bool dup = false;
            for (int t = 0; t < w; ++t) dup |= row[t] == v;
            if (!dup) row[w++] = v;
        }
    }
}

// parth/src/build_graph.cu:326-343: float centroid accumulated node by node,
// divided by n, then the node nearest to it (separately rounded mul/add, strict
// `<` so the lowest id wins ties).
int64_t orc_medoid(const float* vecs, int64_t n, int d) {
    std::vector<float> c(d, 0.0f);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < d; ++k) c[k] += vecs[i * d + k];
    for (int k = 0; k < d; ++k) c[k] /= (float)n;
    int64_t medoid = 0;
    float best = 1e30f;
    for (int64_t i = 0; i < n; ++i) {
        float s = 0.0f;
        for (int k = 0; k < d; ++k) {
            float diff = vecs[i * d + k] - c[k];
            float sq = diff * diff;
            s = s + sq;
        }
        if (s < best) { best = s; medoid = i; }
    }
    return medoid;
}

// parth/src/common.h:89-155 (`CandidateList<K>`), sized by `cap` at run time.
struct CandList {
    std::vector<int32_t> ids;
    std::vector<float> dists;
    int size = 0;
    explicit CandList(int cap) : ids(cap, -1), dists(cap, 1e30f) {}
    // common.h:117-143
    void insert(int32_t id, float dist, int limit) {
        for (int i = 0; i < size; ++i)
            if (ids[i] == id) return;
        if (size < limit) {
            ids[size] = id; dists[size] = dist; ++size;
        } else if (dist < dists[size - 1]) {
            ids[size - 1] = id; dists[size - 1] = dist;
        } else {
            return;
        }
        for (int i = size - 1; i > 0; --i) {
            if (dists[i] < dists[i - 1]) {
                std::swap(dists[i], dists[i - 1]);
                std::swap(ids[i], ids[i - 1]);
            } else {
                break;
            }
        }
    }
    bool contains(int32_t id) const {
        for (int i = 0; i < size; ++i)
            if (ids[i] == id) return true;
        return false;
    }
};

// parth/src/engine.cu:29-87 (`greedy_search`) and build_graph.cu:68-119
// (`greedy_search_build`). The two differ only in the delete mask (the build
// passes none) and an explicit `contains` test that `insert` repeats anyway.
// Expands the closest unvisited candidate until none is left, the visited list
// (cap `visited_cap`, reference VISITED_LIST_SIZE=128) is full, or `max_hops`
// iterations (reference 200) have run. Writes the first `out_n` candidates
// (ids, dists; -1 / 1e30 padded) and the expansion count.
static void parth_search_one(const float* vecs, int d, const int32_t* adj, int row_width,
                             const uint8_t* del, const float* q, int32_t start, int L,
                             int visited_cap, int max_hops, int dist_mode, int out_n,
                             int32_t* out_ids, float* out_d, int32_t* out_hops,
                             int32_t* out_ndist) {
    CandList cand(std::max(L, 1));
    std::vector<int32_t> visited(visited_cap > 0 ? visited_cap : 1);
    int visited_count = 0;
    int ndist = 1;
    float sd = l2sq_m(q, vecs + (int64_t)start * d, d, dist_mode);
    if (del && del[start]) sd = 1e30f;
    cand.insert(start, sd, L);
    int iter = 0;
    while (iter < max_hops) {
        ++iter;
        int32_t next = -1;
        for (int i = 0; i < cand.size; ++i) {
            int32_t id = cand.ids[i];
            bool vis = false;
            for (int k = 0; k < visited_count; ++k)
                if (visited[k] == id) { vis = true; break; }
            if (!vis) { next = id; break; }
        }
        if (next == -1) break;
        if (visited_count < visited_cap) visited[visited_count++] = next;
        else break;
        const int32_t* nb = adj + (int64_t)next * row_width;
        for (int i = 0; i < row_width; ++i) {
            int32_t nbr = nb[i];
            if (nbr == -1) break;
            if (del && del[nbr]) continue;
            if (cand.contains(nbr)) continue;
            float dd = l2sq_m(q, vecs + (int64_t)nbr * d, d, dist_mode);
            ++ndist;
            cand.insert(nbr, dd, L);
        }
    }
    for (int i = 0; i < out_n; ++i) {
        out_ids[i] = i < cand.size ? cand.ids[i] : -1;
        if (out_d) out_d[i] = i < cand.size ? cand.dists[i] : 1e30f;
    }
    if (out_hops) *out_hops = visited_count;
    if (out_ndist) *out_ndist = ndist;
}

void orc_parth_search(const float* vecs, int64_t n, int d, const int32_t* adj, int row_width,
                      const uint8_t* del, const float* queries, int64_t nq, int32_t start,
                      int L, int visited_cap, int max_hops, int dist_mode, int out_n,
                      int32_t* out_ids, float* out_dists, int32_t* out_hops,
                      int32_t* out_ndist) {
    (void)n;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t qi = 0; qi < nq; ++qi) {
        parth_search_one(vecs, d, adj, row_width, del, queries + qi * d, start, L, visited_cap,
                         max_hops, dist_mode, out_n, out_ids + qi * out_n,
                         out_dists ? out_dists + qi * out_n : nullptr,
                         out_hops ? out_hops + qi : nullptr,
                         out_ndist ? out_ndist + qi : nullptr);
    }
}

// parth/src/build_graph.cu:173-264 (`robustPruneKernel`) for nodes [0, n):
// candidates = search list (minus -1 and self) then current neighbours not
// already present (fresh distances), stable bubble sort by distance, then keep
// p* in order while fewer than R are kept, dropping p' when
// alpha * d(p*, p') < d(node, p') (one float multiply, strict <).
void orc_parth_prune(const float* vecs, int64_t n, int d, int32_t* adj, int row_width, int R,
                     float alpha, const int32_t* cand_ids, const float* cand_dists, int L,
                     int dist_mode) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t tid = 0; tid < n; ++tid) {
        std::vector<int32_t> ci;
        std::vector<float> cd;
        ci.reserve(L + row_width);
        cd.reserve(L + row_width);
        const float* my = vecs + tid * d;
        for (int i = 0; i < L; ++i) {
            int32_t id = cand_ids[tid * L + i];
            if (id != -1 && id != (int32_t)tid) {
                ci.push_back(id);
                cd.push_back(cand_dists[tid * L + i]);
            }
        }
        int32_t* row = adj + tid * row_width;
        for (int i = 0; i < row_width; ++i) {
            int32_t nbr = row[i];
            if (nbr != -1 && nbr != (int32_t)tid) {
                bool dup = false;
                for (size_t j = 0; j < ci.size(); ++j)
                    if (ci[j] == nbr) { dup = true; break; }
                if (!dup) {
                    ci.push_back(nbr);
                    cd.push_back(l2sq_m(my, vecs + (int64_t)nbr * d, d, dist_mode));
                }
            }
        }
        const int cnt = (int)ci.size();
        for (int i = 0; i < cnt - 1; ++i)
            for (int j = 0; j < cnt - i - 1; ++j)
                if (cd[j] > cd[j + 1]) {
                    std::swap(cd[j], cd[j + 1]);
                    std::swap(ci[j], ci[j + 1]);
                }
        std::vector<int32_t> nadj(row_width, -1);
        std::vector<char> kept(cnt, 1);
        int nc = 0;
        for (int i = 0; i < cnt && nc < R; ++i) {
            if (!kept[i]) continue;
            int32_t ps = ci[i];
            nadj[nc++] = ps;
            for (int j = i + 1; j < cnt; ++j) {
                if (!kept[j]) continue;
                float dps = l2sq_m(vecs + (int64_t)ps * d, vecs + (int64_t)ci[j] * d, d, dist_mode);
                float lhs = alpha * dps;
                if (lhs < cd[j]) kept[j] = 0;
            }
        }
        for (int i = 0; i < row_width; ++i) row[i] = i < nc ? nadj[i] : -1;
    }
}

// parth/src/build_graph.cu:282-378 (`build_graph`): random init, medoid, then
// `passes` x (search every node from the medoid with beam L, prune every node).
// alphas[p] is the pass-p alpha (the reference uses one alpha for both passes).
// The search of a pass completes for every node before any row is pruned.
int64_t orc_parth_build(const float* vecs, int64_t n, int d, int R, int row_width, int L,
                        const float* alphas, int passes, uint64_t seed, int visited_cap,
                        int max_hops, int dist_mode, int32_t* adj, const int32_t* init,
                        int init_width) {
    if (init)
        orc_init_rows(init, n, init_width, R, row_width, adj);
    else
        orc_random_init(n, R, row_width, seed, adj);
    int64_t medoid = orc_medoid(vecs, n, d);
    std::vector<int32_t> ci((size_t)n * L);
    std::vector<float> cd((size_t)n * L);
    for (int p = 0; p < passes; ++p) {
        orc_parth_search(vecs, n, d, adj, row_width, nullptr, vecs, n, (int32_t)medoid, L,
                         visited_cap, max_hops, dist_mode, L, ci.data(), cd.data(), nullptr,
                         nullptr);
        orc_parth_prune(vecs, n, d, adj, row_width, R, alphas[p], ci.data(), cd.data(), L,
                        dist_mode);
    }
    return medoid;
}

// ---------------------------------------------------------------------------
// fresh mode (G2: vamana/include/search.cuh, kernels.cuh, bloom_filter.cuh)
// ---------------------------------------------------------------------------

// vamana/include/bloom_filter.cuh:11-30
uint32_t orc_bf_hash1(uint32_t x) {
    uint64_t h = 0xcbf29ce4ull;
    h = (h ^ (x & 0xffu)) * 0x01000193ull;
    h = (h ^ ((x >> 8) & 0xffu)) * 0x01000193ull;
    h = (h ^ ((x >> 16) & 0xffu)) * 0x01000193ull;
    h = (h ^ ((x >> 24) & 0xffu)) * 0x01000193ull;
    return (uint32_t)(h % 39988u);
}
uint32_t orc_bf_hash2(uint32_t x) {
    uint64_t h = 0x84222325ull;
    h = (h ^ (x & 0xffu)) * 0x1B3ull;
    h = (h ^ ((x >> 8) & 0xffu)) * 0x1B3ull;
    h = (h ^ ((x >> 16) & 0xffu)) * 0x1B3ull;
    h = (h ^ ((x >> 24) & 0xffu)) * 0x1B3ull;
    return (uint32_t)(h % 39988u);
}

struct Item {
    int32_t id;
    float d;
    bool vis;
};

// Greedy search of vamana/include/search.cuh:226-367 for one query, in the
// canonical order: the worklist (cap L) starts with the start node; each
// iteration expands the closest unexpanded entry (search.cuh:205-221), records
// it in the visited set (cap P, :104-111), takes its neighbours in adjacency
// order, drops tombstoned ones (:122), the query's own node (:126; only when
// the query is a graph node) and those whose two bloom bits are both set in
// the filter state from before this expansion (:130), sets the bits of the
// rest, sorts them stably by distance (sortByDistance) and merges them into
// the worklist with worklist entries first on ties (mergeIntoWorklist
// :140-223), truncating to L. The start node is entered into the filter.
static void fresh_search_one(const float* vecs, int d, const int32_t* adj, int W,
                             const uint8_t* del, const float* q, int32_t self, int32_t start,
                             int L, int max_hops, int P, int dist_mode, int out_n,
                             int32_t* out_ids, float* out_d, int32_t* vis_ids, float* vis_d,
                             int32_t* vis_cnt, int32_t* out_hops, int32_t* out_ndist) {
    std::vector<uint8_t> bloom(39988, 0);
    bloom[orc_bf_hash1((uint32_t)start)] = 1;
    bloom[orc_bf_hash2((uint32_t)start)] = 1;
    std::vector<Item> wl;
    float sd = l2sq_m(q, vecs + (int64_t)start * d, d, dist_mode);
    if (del && del[start]) sd = 1e30f;  // as parth/src/engine.cu:45
    wl.push_back({start, sd, false});
    int nexp = 0, hops = 0, ndist = 1;
    std::vector<int32_t> news;
    while (hops < max_hops) {
        ++hops;
        int p = -1;
        for (size_t i = 0; i < wl.size(); ++i)
            if (!wl[i].vis) { p = (int)i; break; }
        if (p < 0) break;
        wl[p].vis = true;
        const int32_t node = wl[p].id;
        ++nexp;
        if (vis_ids && nexp <= P) {
            vis_ids[nexp - 1] = node;
            vis_d[nexp - 1] = wl[p].d;
        }
        news.clear();
        const int32_t* nb = adj + (int64_t)node * W;
        for (int j = 0; j < W; ++j) {
            const int32_t v = nb[j];
            if (v == -1) break;
            if (del && del[v]) continue;
            if (v == self && g_mut != 3) continue;
            if (bloom[orc_bf_hash1((uint32_t)v)] && bloom[orc_bf_hash2((uint32_t)v)]) continue;
            news.push_back(v);
        }
        std::vector<Item> add;
        for (int32_t v : news) {
            bloom[orc_bf_hash1((uint32_t)v)] = 1;
            bloom[orc_bf_hash2((uint32_t)v)] = 1;
            add.push_back({v, l2sq_m(q, vecs + (int64_t)v * d, d, dist_mode), false});
            ++ndist;
        }
        std::stable_sort(add.begin(), add.end(),
                         [](const Item& a, const Item& b) { return a.d < b.d; });
        std::vector<Item> merged;
        merged.reserve(wl.size() + add.size());
        size_t a = 0, b = 0;
        while ((a < wl.size() || b < add.size()) && (int)merged.size() < L) {
            if (b >= add.size() ||
                (a < wl.size() && (g_mut == 2 ? wl[a].d < add[b].d : wl[a].d <= add[b].d)))
                merged.push_back(wl[a++]);
            else merged.push_back(add[b++]);
        }
        wl.swap(merged);
    }
    for (int i = 0; i < out_n; ++i) {
        out_ids[i] = i < (int)wl.size() ? wl[i].id : -1;
        if (out_d) out_d[i] = i < (int)wl.size() ? wl[i].d : 1e30f;
    }
    if (vis_cnt) *vis_cnt = std::min(nexp, P);
    if (out_hops) *out_hops = nexp;
    if (out_ndist) *out_ndist = ndist;
}

void orc_fresh_search(const float* vecs, int64_t n, int d, const int32_t* adj, int W,
                      const uint8_t* del, const float* queries, int64_t nq, int queries_are_nodes,
                      int64_t node_base, int32_t start, int L, int max_hops, int P, int dist_mode, int out_n,
                      int32_t* out_ids, float* out_dists, int32_t* vis_ids, float* vis_d,
                      int32_t* vis_cnt, int32_t* out_hops, int32_t* out_ndist) {
    (void)n;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t qi = 0; qi < nq; ++qi) {
        fresh_search_one(vecs, d, adj, W, del, queries + qi * d,
                         queries_are_nodes ? (int32_t)(node_base + qi) : -1, start, L, max_hops, P, dist_mode,
                         out_n, out_ids ? out_ids + qi * out_n : nullptr,
                         out_dists ? out_dists + qi * out_n : nullptr,
                         vis_ids ? vis_ids + qi * P : nullptr, vis_d ? vis_d + qi * P : nullptr,
                         vis_cnt ? vis_cnt + qi : nullptr, out_hops ? out_hops + qi : nullptr,
                         out_ndist ? out_ndist + qi : nullptr);
    }
}

// Robust prune with `<=` (vamana/include/kernels.cuh:231-356, 626-743) over a
// sorted candidate list: the first still-INIT candidate becomes p* (while
// fewer than R are kept), then every INIT candidate p' with
// d(p*, p') * alpha <= d(node, p') is PRUNED. Returns kept ids and dists.
static void prune_le(const float* vecs, int d, const std::vector<Item>& c, float alpha, int R,
                     int dist_mode, std::vector<int32_t>& out, std::vector<float>& outd) {
    out.clear();
    outd.clear();
    std::vector<uint8_t> st(c.size(), 0);  // 0 INIT, 1 PRUNED, 2 NEIGHBOR
    while ((int)out.size() < R) {
        int p = -1;
        for (size_t i = 0; i < c.size(); ++i)
            if (st[i] == 0) { p = (int)i; break; }
        if (p < 0) break;
        st[p] = 2;
        out.push_back(c[p].id);
        outd.push_back(c[p].d);
        const float* pv = vecs + (int64_t)c[p].id * d;
        for (size_t j = 0; j < c.size(); ++j) {
            if (st[j] != 0) continue;
            float dp = l2sq_m(pv, vecs + (int64_t)c[j].id * d, d, dist_mode);
            float lhs = dp * alpha;
            if (g_mut == 1 ? lhs < c[j].d : lhs <= c[j].d) st[j] = 1;
        }
    }
}

// Out-neighbour prune of vamana/include/kernels.cuh:358-500 for nodes
// [base, base+cnt): candidates = the visited set stably sorted by distance,
// merged (visited entries first on ties, mergeIntoVisitedSet :171-229) with
// the current neighbours stably sorted by distance, capped at P; the node
// itself and tombstoned nodes are left out. Writes the new rows and, per kept p*, the reverse pair
// (p* <- node, d(node, p*)) into pair_dst/pair_d [cnt x R] (-1 unused).
void orc_fresh_out_prune(const float* vecs, int d, int32_t* adj, int W, int R, float alpha,
                         int64_t base, int64_t cnt, const int32_t* vis_ids, const float* vis_d,
                         const int32_t* vis_cnt, int P, int dist_mode, int32_t* pair_dst,
                         float* pair_d, const uint8_t* del) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < cnt; ++i) {
        const int64_t node = base + i;
        std::vector<Item> A, B, C;
        for (int j = 0; j < vis_cnt[i]; ++j) {
            int32_t v = vis_ids[i * P + j];
            if (v != (int32_t)node && !(del && del[v])) A.push_back({v, vis_d[i * P + j], false});
        }
        std::stable_sort(A.begin(), A.end(), [](const Item& a, const Item& b) { return a.d < b.d; });
        int32_t* row = adj + node * W;
        const float* nv = vecs + node * d;
        for (int j = 0; j < W; ++j) {
            int32_t v = row[j];
            if (v == -1 || v == (int32_t)node || (del && del[v])) continue;
            B.push_back({v, l2sq_m(nv, vecs + (int64_t)v * d, d, dist_mode), false});
        }
        std::stable_sort(B.begin(), B.end(), [](const Item& a, const Item& b) { return a.d < b.d; });
        size_t a = 0, b = 0;
        while ((a < A.size() || b < B.size()) && (int)C.size() < P) {
            if (b >= B.size() ||
                (a < A.size() && (g_mut == 6 ? A[a].d < B[b].d : A[a].d <= B[b].d)))
                C.push_back(A[a++]);
            else C.push_back(B[b++]);
        }
        std::vector<int32_t> out;
        std::vector<float> outd;
        prune_le(vecs, d, C, alpha, R, dist_mode, out, outd);
        for (int j = 0; j < W; ++j) row[j] = j < (int)out.size() ? out[j] : -1;
        for (int j = 0; j < R; ++j) {
            pair_dst[i * R + j] = j < (int)out.size() ? out[j] : -1;
            pair_d[i * R + j] = j < (int)out.size() ? outd[j] : 0.0f;
        }
    }
}

// Reverse-edge pass of vamana/include/kernels.cuh:745-903 for every node that
// received at least one reverse pair: reverse candidates ordered by
// (distance, source id) and capped at M, merged with the node's current
// neighbours (stable by distance; reverse candidates first on ties,
// mergeIntoReverseEdges :566-624), capped at M, pruned with <= into R.
void orc_fresh_reverse(const float* vecs, int64_t n, int d, int32_t* adj, int W, int R,
                       float alpha, int64_t base, int64_t cnt, const int32_t* pair_dst,
                       const float* pair_d, int M, int dist_mode) {
    std::vector<std::vector<Item>> rev(n);
    for (int64_t i = 0; i < cnt; ++i)
        for (int j = 0; j < R; ++j) {
            int32_t dst = pair_dst[i * R + j];
            if (dst >= 0) rev[dst].push_back({(int32_t)(base + i), pair_d[i * R + j], false});
        }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t node = 0; node < n; ++node) {
        std::vector<Item>& A = rev[node];
        if (A.empty()) continue;
        std::sort(A.begin(), A.end(), [](const Item& a, const Item& b) {
            return a.d < b.d || (a.d == b.d && a.id < b.id);
        });
        const int Mc = g_mut == 4 ? M + 1 : g_mut == 5 ? M - 1 : M;
        if ((int)A.size() > Mc) A.resize(Mc);
        std::vector<Item> B, C;
        int32_t* row = adj + node * W;
        const float* nv = vecs + node * d;
        for (int j = 0; j < W; ++j) {
            int32_t v = row[j];
            if (v == -1 || v == (int32_t)node) continue;
            B.push_back({v, l2sq_m(nv, vecs + (int64_t)v * d, d, dist_mode), false});
        }
        std::stable_sort(B.begin(), B.end(), [](const Item& a, const Item& b) { return a.d < b.d; });
        size_t a = 0, b = 0;
        while ((a < A.size() || b < B.size()) && (int)C.size() < Mc) {
            if (b >= B.size() ||
                (a < A.size() && (g_mut == 7 ? A[a].d < B[b].d : A[a].d <= B[b].d)))
                C.push_back(A[a++]);
            else C.push_back(B[b++]);
        }
        std::vector<int32_t> out;
        std::vector<float> outd;
        prune_le(vecs, d, C, alpha, R, dist_mode, out, outd);
        for (int j = 0; j < W; ++j) row[j] = j < (int)out.size() ? out[j] : -1;
    }
}

// FreshVamana build (vamana_impl.cuh:289-360 runVamana) from the parth random
// initial graph and the computed medoid, `passes` times: search every node,
// out-prune every node, reverse-edge prune the nodes that received pairs.
int64_t orc_fresh_build(const float* vecs, int64_t n, int d, int R, int L, const float* alphas,
                        int passes, uint64_t seed, int P, int M, int max_hops, int dist_mode,
                        int32_t* adj, const int32_t* init, int init_width) {
    if (init)
        orc_init_rows(init, n, init_width, R, R, adj);
    else
        orc_random_init(n, R, R, seed, adj);
    int64_t medoid = orc_medoid(vecs, n, d);
    std::vector<int32_t> vis((size_t)n * P), vcnt(n), pdst((size_t)n * R);
    std::vector<float> visd((size_t)n * P), pd((size_t)n * R);
    for (int p = 0; p < passes; ++p) {
        orc_fresh_search(vecs, n, d, adj, R, nullptr, vecs, n, 1, 0, (int32_t)medoid, L, max_hops, P,
                         dist_mode, 0, nullptr, nullptr, vis.data(), visd.data(), vcnt.data(),
                         nullptr, nullptr);
        orc_fresh_out_prune(vecs, d, adj, R, R, alphas[p], 0, n, vis.data(), visd.data(),
                            vcnt.data(), P, dist_mode, pdst.data(), pd.data(), nullptr);
        orc_fresh_reverse(vecs, n, d, adj, R, R, alphas[p], 0, n, pdst.data(), pd.data(), M,
                          dist_mode);
    }
    return medoid;
}

// Batched FreshVamana insert (vamana_impl.cuh:182-265 `patchGraph` applied to
// one batch): nodes [n_old, n_old + n_new) already hold their vectors and
// empty rows; each is searched from `start` over the graph (tombstones
// skipped, its own id skipped), out-pruned from its visited set, and the
// reverse pairs are merged into their destinations. All searches of the batch
// see the graph as it was before the batch.
void orc_fresh_insert(const float* vecs, int64_t n_old, int64_t n_new, int d, int32_t* adj, int R,
                      const uint8_t* del, int32_t start, int L, float alpha, int P, int M,
                      int max_hops, int dist_mode) {
    const int64_t N = n_old + n_new;
    std::vector<int32_t> vis((size_t)n_new * P), vcnt(n_new), pdst((size_t)n_new * R);
    std::vector<float> visd((size_t)n_new * P), pd((size_t)n_new * R);
    orc_fresh_search(vecs, N, d, adj, R, del, vecs + n_old * d, n_new, 1, n_old, start, L,
                     max_hops, P, dist_mode, 0, nullptr, nullptr, vis.data(), visd.data(),
                     vcnt.data(), nullptr, nullptr);
    orc_fresh_out_prune(vecs, d, adj, R, R, alpha, n_old, n_new, vis.data(), visd.data(),
                        vcnt.data(), P, dist_mode, pdst.data(), pd.data(), del);
    orc_fresh_reverse(vecs, N, d, adj, R, R, alpha, n_old, n_new, pdst.data(), pd.data(), M,
                      dist_mode);
}

// Delete consolidation (the reference has none; after FreshDiskANN, Singh et
// al. 2021, Algorithm 4, with the live part of each row kept). For every live
// node p with a tombstoned out-neighbour: K = p's live out-neighbours in row
// order, all kept; candidates = the live out-neighbours of p's tombstoned
// out-neighbours (rows in adjacency order), without p and K, ordered by
// (d(p, c), id), the first M kept; each candidate in order joins the row
// unless a row member r occludes it (alpha * d(r, c) <= d(p, c)), until the
// row holds R. Then the rows of tombstoned nodes are emptied. Returns the
// number of repaired nodes.
int64_t orc_consolidate(const float* vecs, int64_t n, int d, int32_t* adj, int R,
                        const uint8_t* del, float alpha, int M, int dist_mode) {
    std::vector<int64_t> todo;
    for (int64_t p = 0; p < n; ++p) {
        if (del[p]) continue;
        for (int j = 0; j < R; ++j) {
            int32_t v = adj[p * R + j];
            if (v >= 0 && del[v]) { todo.push_back(p); break; }
        }
    }
    std::vector<std::vector<int32_t>> rows(todo.size());
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < (int64_t)todo.size(); ++t) {
        const int64_t p = todo[t];
        std::vector<int32_t> row;
        for (int j = 0; j < R; ++j) {
            int32_t w = adj[p * R + j];
            if (w >= 0 && !del[w]) row.push_back(w);
        }
        std::vector<int32_t> ids;
        for (int j = 0; j < R; ++j) {
            int32_t v = adj[p * R + j];
            if (v < 0 || !del[v]) continue;
            for (int k = 0; k < R; ++k) {
                int32_t w = adj[(int64_t)v * R + k];
                if (w < 0 || del[w] || w == (int32_t)p) continue;
                bool seen = false;
                for (int32_t x : row) seen |= x == w;
                for (int32_t x : ids) seen |= x == w;
                if (!seen) ids.push_back(w);
            }
        }
        std::vector<Item> C;
        for (int32_t w : ids)
            C.push_back({w, l2sq_m(vecs + p * d, vecs + (int64_t)w * d, d, dist_mode), false});
        std::sort(C.begin(), C.end(), [](const Item& a, const Item& b) {
            return a.d < b.d || (a.d == b.d && a.id < b.id);
        });
        if ((int)C.size() > M) C.resize(M);
        for (const Item& c : C) {
            if ((int)row.size() >= R) break;
            bool occ = false;
            for (int32_t r : row) {
                float dr = l2sq_m(vecs + (int64_t)r * d, vecs + (int64_t)c.id * d, d, dist_mode);
                if (dr * alpha <= c.d) { occ = true; break; }
            }
            if (!occ) row.push_back(c.id);
        }
        rows[t] = row;
    }
    for (size_t t = 0; t < todo.size(); ++t)
        for (int j = 0; j < R; ++j) adj[todo[t] * R + j] = j < (int)rows[t].size() ? rows[t][j] : -1;
    for (int64_t p = 0; p < n; ++p)
        if (del[p])
            for (int j = 0; j < R; ++j) adj[p * R + j] = -1;
    return (int64_t)todo.size();
}

// Entry point over live nodes: the medoid rule of parth/src/build_graph.cu:
// 326-343 (sequential float centroid, nearest node, lowest id on ties)
// restricted to nodes whose tombstone is clear; divides by the live count.
int64_t orc_medoid_live(const float* vecs, int64_t n, int d, const uint8_t* del) {
    std::vector<float> c(d, 0.0f);
    int64_t live = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (del && del[i]) continue;
        ++live;
        for (int k = 0; k < d; ++k) c[k] += vecs[i * d + k];
    }
    if (live == 0) return -1;
    for (int k = 0; k < d; ++k) c[k] /= (float)live;
    int64_t medoid = -1;
    float best = 1e30f;
    for (int64_t i = 0; i < n; ++i) {
        if (del && del[i]) continue;
        float s = 0.0f;
        for (int k = 0; k < d; ++k) {
            float diff = vecs[i * d + k] - c[k];
            float sq = diff * diff;
            s = s + sq;
        }
        if (s < best) { best = s; medoid = i; }
    }
    return medoid;
}

// ---------------------------------------------------------------------------
// Exact k-NN ground truth in FP64 (replaces parth/src/engine.cu:253-279
// `bruteForceKernel` (FP32) and the external DiskANN compute_groundtruth called
// at vamana.cu:1644). Each distance is sum_k (double(q_k) - double(x_k))^2 in
// index order; ties ordered by id. Rows whose `del` flag is set are skipped.
// ---------------------------------------------------------------------------
void orc_bruteforce_fp64(const float* vecs, int64_t n, int d, const uint8_t* del,
                         const float* queries, int64_t nq, int k, int32_t* out_ids,
                         double* out_dists) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t qi = 0; qi < nq; ++qi) {
        const float* q = queries + qi * d;
        std::vector<std::pair<double, int32_t>> best;
        best.reserve(k + 1);
        for (int64_t i = 0; i < n; ++i) {
            if (del && del[i]) continue;
            const float* x = vecs + i * d;
            double s = 0.0;
            for (int t = 0; t < d; ++t) {
                double df = (double)q[t] - (double)x[t];
                s += df * df;
            }
            std::pair<double, int32_t> e(s, (int32_t)i);
            if ((int)best.size() < k) {
                best.insert(std::upper_bound(best.begin(), best.end(), e), e);
            } else if (e < best.back()) {
                best.pop_back();
                best.insert(std::upper_bound(best.begin(), best.end(), e), e);
            }
        }
        for (int j = 0; j < k; ++j) {
            out_ids[qi * k + j] = j < (int)best.size() ? best[j].second : -1;
            if (out_dists) out_dists[qi * k + j] = j < (int)best.size() ? best[j].first : INFINITY;
        }
    }
}

// parth/src/engine.cu:407-568 (`VamanaEngine::process_workload`) with the
// kernels it launches, run in a fixed sequential order. The reference runs
// each kernel's threads concurrently; this restatement fixes the order a
// single converged wavefront would give:
//  * a batch is ops [b, b + batch): its deletes, then its inserts, then its
//    searches, each in op order; a type other than 0 / 1 is a search;
//  * deleteKernel (:131-147): an id in [0, cap) is marked deleted and stored
//    at free-list slot c1++ (a slot outside [0, cap) is dropped: the
//    reference writes outside its list there);
//  * insertPrepareKernel (:161-199): every insert of the batch reads the
//    count c1 of the batch start; when it is positive each insert decrements
//    c1 and, if the old value was positive, takes free-list slot old - 1 (and
//    unmarks it); otherwise it takes id c0++. Then each writes its vector and
//    a -1 row (op order: a repeated id keeps the last vector);
//  * insertLinkKernel (:213-239): the nearest node of each new node's stored
//    vector by greedy_search (beam link_L) on the graph before any link of the
//    batch; slot 0 of every new row is set to it; then, in op order, the new id
//    goes into the first -1 slot of that node's row;
//  * searchKernel (:101-120): greedy_search top 10 with beam L from `start`;
//  * bruteForceKernel (:254-279): top 10 live ids below c0, CandidateList order.
// vecs [cap x d], adj [cap x W] and del [cap] are updated in place, *n_io is c0
// and *c1_io the free-list counter (start 0), freelist [cap]. found / gt
// [searches x 10], assigned [inserts], *search_s the wall time of the search
// loops (ground truth excluded). Returns the number of searches, or -1 when an
// insert would pass cap.
int64_t orc_engine_run(float* vecs, int32_t* adj, uint8_t* del, int64_t cap, int d, int W,
                       int64_t* n_io, int64_t* c1_io, int32_t* freelist, const int32_t* types,
                       const int32_t* ids, const float* opvecs, int64_t n_ops, int batch, int L,
                       int link_L, int visited_cap, int max_hops, int dist_mode, int32_t start,
                       int32_t* found, int32_t* gt, int32_t* assigned, double* search_s) {
    int64_t c0 = *n_io, c1 = *c1_io, n_search = 0, n_assigned = 0;
    double t_search = 0.0;
    for (int64_t b = 0; b < n_ops; b += batch) {
        const int64_t e = std::min<int64_t>(b + batch, n_ops);
        std::vector<int64_t> dels, ins, srch;
        for (int64_t i = b; i < e; ++i) {
            if (types[i] == 0) dels.push_back(i);
            else if (types[i] == 1) ins.push_back(i);
            else srch.push_back(i);
        }
        for (int64_t i : dels) {
            const int32_t t = ids[i];
            if (t < 0 || t >= cap) continue;
            del[t] = 1;
            const int64_t slot = c1++;
            if (slot >= 0 && slot < cap) freelist[slot] = t;
        }
        if (!ins.empty()) {
            const int64_t m = (int64_t)ins.size();
            const int64_t fc = c1;
            const int64_t popped = fc > 0 ? std::min<int64_t>(m, fc) : 0;
            if (c0 + (m - popped) > cap) return -1;
            std::vector<int32_t> my(m);
            for (int64_t j = 0; j < m; ++j) {
                int32_t id = -1;
                if (fc > 0) {
                    const int64_t old = c1--;
                    if (old > 0) {
                        id = freelist[old - 1];
                        del[id] = 0;
                    }
                }
                if (id == -1) id = (int32_t)c0++;
                my[j] = id;
                if (assigned) assigned[n_assigned] = id;
                ++n_assigned;
                std::memcpy(vecs + (int64_t)id * d, opvecs + ins[j] * d, d * sizeof(float));
                for (int k = 0; k < W; ++k) adj[(int64_t)id * W + k] = -1;
            }
            std::vector<int32_t> best(m);
#pragma omp parallel for schedule(dynamic, 4)
            for (int64_t j = 0; j < m; ++j)
                parth_search_one(vecs, d, adj, W, del, vecs + (int64_t)my[j] * d, start, link_L,
                                 visited_cap, max_hops, dist_mode, 1, &best[j], nullptr, nullptr,
                                 nullptr);
            for (int64_t j = 0; j < m; ++j)
                if (best[j] != -1) adj[(int64_t)my[j] * W] = best[j];
            for (int64_t j = 0; j < m; ++j) {
                if (best[j] == -1) continue;
                int32_t* row = adj + (int64_t)best[j] * W;
                for (int k = 0; k < W; ++k)
                    if (row[k] == -1) {
                        row[k] = my[j];
                        break;
                    }
            }
        }
        const int64_t ns = (int64_t)srch.size();
        const auto ts = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t j = 0; j < ns; ++j)
            parth_search_one(vecs, d, adj, W, del, opvecs + srch[j] * d, start, L, visited_cap,
                             max_hops, dist_mode, 10, found + (n_search + j) * 10, nullptr,
                             nullptr, nullptr);
        t_search += std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t j = 0; j < ns; ++j) {
            const float* q = opvecs + srch[j] * d;
            CandList top(10);
            for (int64_t i = 0; i < c0; ++i) {
                if (del[i]) continue;
                top.insert((int32_t)i, l2sq_m(q, vecs + i * d, d, dist_mode), 10);
            }
            for (int k = 0; k < 10; ++k) gt[(n_search + j) * 10 + k] = k < top.size ? top.ids[k] : -1;
        }
        n_search += ns;
    }
    *n_io = c0;
    *c1_io = c1;
    if (search_s) *search_s = t_search;
    return n_search;
}

// ---------------------------------------------------------------------------
// BANG exact search (bang_exact/src/parANN.cu, the schedule its host loop
// :500-668 drives). Per query and round:
//  * neighbor_filtering_new (:996-1052): candidates are, in round 1, the entry
//    point (:1017-1030), then the parent's neighbours in adjacency order whose
//    single bloom cell hashFn1_d(x) (:1055-1064, FNV-1a over the four bytes,
//    modulo 399,887) is clear; their cells are set. The reference tests and
//    sets cells from concurrent threads, so two neighbours of one parent with
//    the same cell may both pass there; here the first in adjacency order does
//    and the later one is dropped. Each candidate is one distance computation
//    (the `bang_counter` total, :519-526, :596-599).
//  * compute_neighborDist_par (:1118-1158): the FP32 distance of each
//    candidate, summed by 8 threads (thread s adds (x_i - q_i)^2 for
//    i = s, s + 8, ... as one fused multiply-add chain) and combined by
//    cub::WarpReduce<float, 8> as ((p0+p1)+(p2+p3))+((p4+p5)+(p6+p7)).
//  * compute_BestLSets_par_sort_msort_new (:1475-1669): a stable merge sort of
//    the candidates by distance (ties in arrival order), then round 1 takes the
//    first min(count, L) as the list with only the entry point visited; later
//    rounds merge the first nb sorted candidates into the list of B entries,
//    nb = max(#leading candidates closer than the list's last entry (among
//    the first min(L, count)), min(L - B, count)), new entries before list
//    entries of equal distance, the list cut at min(B + nb, L); the first
//    unvisited list entry is marked visited and becomes the next parent.
//  * A round that yields no candidate ends the query: the reference skips the
//    merge then and leaves the list bound of its parent scan unset
//    (:1588-1590 vs :1641); its printed totals for SIFT10K (641,853
//    computations, 47 rounds) are those of this reading, while scanning the
//    whole list gives 642,920.
//  * The host loop stops after round 4L + 19 (MAX_PARENTS_PERQUERY - 1,
//    :49-50, :665-668); a query stops when it has no parent.
// Results: the first k list entries (parANN.cu:1252-1254 copies d_BestLSets),
// -1 / +inf past the list's end. out_stats [nq x 2]: computations, rounds run.
// Rows whose `del` flag is set are never candidates (the reference has no
// tombstones). Mutations: 8 scans the whole list in a round without
// candidates, 9 puts list entries before new ones on ties, 10 drops the
// min(L - B, count) term of nb.
// ---------------------------------------------------------------------------
uint32_t orc_bang_hash(uint32_t x) {
    uint64_t h = 0xcbf29ce4ull;
    h = (h ^ (x & 0xffu)) * 0x01000193ull;
    h = (h ^ ((x >> 8) & 0xffu)) * 0x01000193ull;
    h = (h ^ ((x >> 16) & 0xffu)) * 0x01000193ull;
    h = (h ^ ((x >> 24) & 0xffu)) * 0x01000193ull;
    return (uint32_t)(h % 399887u);
}

float orc_bang_l2sq(const float* q, const float* x, int d) {
    float p[8];
    for (int s = 0; s < 8; ++s) {
        float acc = 0.0f;
        for (int i = s; i < d; i += 8) {
            float diff = x[i] - q[i];
            acc = fmaf(diff, diff, acc);
        }
        p[s] = acc;
    }
    return ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
}

void orc_bang_search(const float* vecs, int64_t n, int d, const int32_t* adj, int R,
                     const uint8_t* del, int64_t entry, const float* queries, int64_t nq, int L,
                     int k, int32_t* out_ids, float* out_dists, int32_t* out_stats) {
    (void)n;
    const int max_rounds = 4 * L + 19;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t qi = 0; qi < nq; ++qi) {
        const float* q = queries + qi * d;
        std::vector<uint8_t> bloom(399887, 0);
        std::vector<Item> list;  // sorted by distance; vis = expanded
        std::vector<int32_t> cand;
        std::vector<std::pair<float, int32_t>> sc;  // (distance, arrival)
        int64_t parent = entry;
        int comps = 0, rounds = 0;
        for (int it = 1; it <= max_rounds; ++it) {
            rounds = it;
            cand.clear();
            if (it == 1) {
                bloom[orc_bang_hash((uint32_t)entry)] = 1;
                cand.push_back((int32_t)entry);
            }
            const int32_t* row = adj + parent * R;
            for (int j = 0; j < R && row[j] >= 0; ++j) {
                const int32_t x = row[j];
                if (del && del[x]) continue;
                const uint32_t h = orc_bang_hash((uint32_t)x);
                if (bloom[h]) continue;
                bloom[h] = 1;
                cand.push_back(x);
            }
            const int nc = (int)cand.size();
            comps += nc;
            int bound = 0;
            if (nc > 0) {
                sc.clear();
                for (int c = 0; c < nc; ++c)
                    sc.emplace_back(orc_bang_l2sq(q, vecs + (int64_t)cand[c] * d, d), c);
                std::stable_sort(sc.begin(), sc.end(),
                                 [](const std::pair<float, int32_t>& a,
                                    const std::pair<float, int32_t>& b) { return a.first < b.first; });
                if (it == 1) {
                    list.clear();
                    for (int c = 0; c < std::min(nc, L); ++c) {
                        const int32_t id = cand[sc[c].second];
                        list.push_back({id, sc[c].first, id == entry});
                    }
                } else {
                    const int B = (int)list.size();
                    const float mx = list[B - 1].d;
                    const int T = std::min(L, nc);
                    int nb = 0;
                    while (nb < T && sc[nb].first < mx) ++nb;
                    if (g_mut != 10) nb = std::max(nb, std::min(L - B, nc));
                    const int nb2 = std::min(B + nb, L);
                    std::vector<Item> merged;
                    merged.reserve(B + nb);
                    int a = 0, b = 0;
                    while ((int)merged.size() < nb2) {
                        // new entry first unless the list entry is strictly closer
                        const bool take_list =
                            b >= nb ||
                            (a < B && (g_mut == 9 ? list[a].d <= sc[b].first
                                                  : list[a].d < sc[b].first));
                        if (take_list) {
                            merged.push_back(list[a++]);
                        } else {
                            merged.push_back({cand[sc[b].second], sc[b].first, false});
                            ++b;
                        }
                    }
                    list.swap(merged);
                }
                bound = (int)list.size();
            } else if (g_mut == 8) {
                bound = (int)list.size();
            }
            int next = -1;
            for (int j = 0; j < bound; ++j) {
                if (!list[j].vis) {
                    list[j].vis = true;
                    next = list[j].id;
                    break;
                }
            }
            if (next < 0) break;
            parent = next;
        }
        for (int j = 0; j < k; ++j) {
            const bool ok = j < (int)list.size();
            out_ids[qi * k + j] = ok ? list[j].id : -1;
            if (out_dists) out_dists[qi * k + j] = ok ? list[j].d : INFINITY;
        }
        if (out_stats) {
            out_stats[qi * 2] = comps;
            out_stats[qi * 2 + 1] = rounds;
        }
    }
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int t) {
#ifdef _OPENMP
    omp_set_num_threads(t);
#else
    (void)t;
#endif
}

}  // extern "C"
