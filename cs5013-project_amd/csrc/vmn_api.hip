// Host runtime behind the C ABI (include/vamana_mi355x.h): the index handle,
// device memory, the build / search orchestration and error reporting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vmn_internal.hpp"
#include "vmn_kernels.hpp"
#include "vmn_update.hpp"

namespace vmn_rt {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int set_device(const vmn_index* idx) {
    HIPCHK(hipSetDevice(idx->p.device));
    return VMN_OK;
}

// Places `n` rows of dimension D (host or device pointer) as [n x DP] rows in
// dst (device). Returns the device pointer holding padded rows (may be src
// itself when it is a device pointer with D == DP).
int stage_rows(vmn_index* idx, const float* src, int64_t n, int32_t mem, DevBuf& scratch,
               const float** out) {
    const int D = idx->p.dim, DP = idx->DP;
    if (mem == VMN_MEM_DEVICE && D == DP) {
        *out = src;
        return VMN_OK;
    }
    HIPCHK(scratch.reserve((size_t)n * DP * sizeof(float)));
    if (mem == VMN_MEM_DEVICE) {
        HIPCHK(vmn::launch_pad_rows(src, n, D, scratch.as<float>(), DP, idx->stream));
    } else if (D == DP) {
        HIPCHK(hipMemcpyAsync(scratch.p, src, (size_t)n * D * sizeof(float), hipMemcpyHostToDevice,
                              idx->stream));
    } else {
        std::vector<float> tmp((size_t)n * DP, 0.0f);
        for (int64_t i = 0; i < n; ++i) std::memcpy(&tmp[i * DP], src + i * D, D * sizeof(float));
        HIPCHK(hipMemcpyAsync(scratch.p, tmp.data(), tmp.size() * sizeof(float),
                              hipMemcpyHostToDevice, idx->stream));
        HIPCHK(hipStreamSynchronize(idx->stream));
    }
    *out = scratch.as<float>();
    return VMN_OK;
}

int zero_counter(vmn_index* idx) {
    HIPCHK(idx->counter.reserve(64));
    HIPCHK(vmn::launch_fill_u64(idx->counter.as<unsigned long long>(), 1, 0ull, idx->stream));
    return VMN_OK;
}

int compute_medoid(vmn_index* idx) {
    HIPCHK(idx->centroid.reserve((size_t)idx->DP * sizeof(float)));
    HIPCHK(idx->medoid_key.reserve(64));
    HIPCHK(vmn::launch_medoid(idx->vecs.as<float>(), idx->n, idx->p.dim, idx->DP,
                              idx->n_deleted ? idx->del_bits.as<uint32_t>() : nullptr,
                              idx->centroid.as<float>(),
                              idx->medoid_key.as<unsigned long long>(), idx->stream));
    unsigned long long key = 0;
    HIPCHK(hipMemcpyAsync(&key, idx->medoid_key.p, sizeof key, hipMemcpyDeviceToHost, idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    idx->medoid = key == ~0ull ? 0 : (int64_t)(uint32_t)(key & 0xffffffffull);
    return VMN_OK;
}

int alloc_graph(vmn_index* idx, int64_t n) {
    ++idx->epoch;
    const int64_t cap = std::max<int64_t>(idx->p.capacity, n);
    idx->cap = cap;
    idx->n_deleted = 0;
    idx->n_pend = 0;
    idx->freelist.clear();
    idx->free_count = 0;
    HIPCHK(idx->vecs.reserve((size_t)cap * idx->DP * sizeof(float)));
    HIPCHK(idx->adj.reserve((size_t)cap * idx->p.R * sizeof(int32_t)));
    HIPCHK(idx->del_bits.reserve((size_t)((cap + 31) / 32) * sizeof(uint32_t)));
    HIPCHK(vmn::launch_fill_i32(idx->del_bits.as<int32_t>(), (cap + 31) / 32, 0, idx->stream));
    if (idx->p.low_precision & VMN_STORE_FP8)
        HIPCHK(idx->adj_fp8n.reserve((size_t)cap * idx->p.R * sizeof(float)));
    if (idx->p.low_precision & VMN_STORE_FP8)
        HIPCHK(idx->adj_hash.reserve((size_t)cap * idx->p.R * sizeof(uint32_t)));
    return VMN_OK;
}

// (Re)computes the low-precision copies of rows [from, to). The FP8 centre
// and scales are fixed from the rows present at the first call (later rows
// are coded with them and clamped).
int refresh_lowp(vmn_index* idx, int64_t from, int64_t to) {
    const int DP = idx->DP;
    const int64_t cap = idx->cap, cnt = to - from;
    if (cnt <= 0) return VMN_OK;
    const float* src = idx->vecs.as<float>() + from * DP;
    if ((idx->p.low_precision & VMN_STORE_BF16) || idx->exact_bf16) {
        HIPCHK(idx->bf16_rows.reserve((size_t)cap * DP * 2));
        HIPCHK(idx->bf16_norms.reserve((size_t)cap * sizeof(float)));
        HIPCHK(vmn::launch_to_bf16(src, cnt, DP, static_cast<char*>(idx->bf16_rows.p) + from * DP * 2,
                                   idx->bf16_norms.as<float>() + from, idx->stream));
    }
    if (idx->p.low_precision & VMN_STORE_FP8) {
        HIPCHK(idx->fp8_rows.reserve((size_t)cap * DP));
        HIPCHK(idx->fp8_norms.reserve((size_t)cap * sizeof(float)));
        HIPCHK(idx->fp8_scales.reserve((size_t)DP * sizeof(float)));
        HIPCHK(idx->fp8_centre.reserve((size_t)DP * sizeof(float)));
        HIPCHK(idx->fp8_absmax.reserve((size_t)DP * sizeof(unsigned int)));
        if (!idx->fp8_scaled) {
            HIPCHK(idx->fp8_temp.reserve(vmn::fp8_scales_temp_bytes(to, DP)));
            HIPCHK(vmn::launch_fp8_scales(idx->vecs.as<float>(), to, idx->p.dim, DP,
                                          idx->fp8_absmax.as<unsigned int>(),
                                          idx->fp8_temp.as<double>(), idx->fp8_centre.as<float>(),
                                          idx->fp8_scales.as<float>(), idx->stream));
            idx->fp8_scaled = true;
            from = 0;
            src = idx->vecs.as<float>();
        }
        HIPCHK(vmn::launch_to_fp8(src, to - from, DP, idx->fp8_scales.as<float>(),
                                  idx->fp8_centre.as<float>(),
                                  static_cast<char*>(idx->fp8_rows.p) + from * DP,
                                  idx->fp8_norms.as<float>() + from, idx->stream));
    }
    return VMN_OK;
}

int update_exact(vmn_index* idx, int64_t from, int64_t to) {
    if (from == 0) idx->exact_bf16 = true;
    return check_exact(idx, from, to);
}

int check_exact(vmn_index* idx, int64_t from, int64_t to) {
    if (!idx->exact_bf16 || to <= from) return VMN_OK;
    HIPCHK(idx->exact_flag.reserve(64));
    HIPCHK(vmn::launch_fill_i32(idx->exact_flag.as<int32_t>(), 1, 1, idx->stream));
    HIPCHK(vmn::launch_bf16_exact(idx->vecs.as<float>(), from, to, idx->DP,
                                  idx->exact_flag.as<unsigned int>(), idx->stream));
    unsigned int ok = 0;
    HIPCHK(hipMemcpyAsync(&ok, idx->exact_flag.p, sizeof ok, hipMemcpyDeviceToHost, idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    idx->exact_bf16 = ok != 0;
    return VMN_OK;
}

// Traversal precision of the build and insert searches: BF16 rows when every
// row makes it exact (the same graph as FP32, half the row bytes).
int build_prec(const vmn_index* idx) { return idx->exact_bf16 ? VMN_PREC_BF16 : VMN_PREC_FP32; }

// Traversal precision for nq padded query rows q (device) that are not index
// rows: BF16 only when the index rows and the queries all qualify, since a
// query rounded to BF16 would give other distances than the FP32 ones.
int query_prec(vmn_index* idx, const float* q, int64_t nq, int* prec) {
    *prec = VMN_PREC_FP32;
    if (!idx->exact_bf16 || nq <= 0) return VMN_OK;
    HIPCHK(idx->exact_flag.reserve(64));
    unsigned int* flag = idx->exact_flag.as<unsigned int>() + 1;
    HIPCHK(vmn::launch_fill_i32(reinterpret_cast<int32_t*>(flag), 1, 1, idx->stream));
    HIPCHK(vmn::launch_bf16_exact(q, 0, nq, idx->DP, flag, idx->stream));
    unsigned int ok = 0;
    HIPCHK(hipMemcpyAsync(&ok, flag, sizeof ok, hipMemcpyDeviceToHost, idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    if (ok) *prec = VMN_PREC_BF16;
    return VMN_OK;
}

// The prunes read the same BF16 rows under the same condition.
void set_prune_rows(const vmn_index* idx, vmn::PruneArgs& a) {
    a.bf16_rows = idx->exact_bf16 ? idx->bf16_rows.p : nullptr;
    a.bf16_norms = idx->exact_bf16 ? idx->bf16_norms.as<float>() : nullptr;
}

void search_args(const vmn_index* idx, vmn::SearchArgs& a, const float* q, int64_t nq, int L,
                 int out_n, int32_t* out_ids, float* out_d, bool queries_are_nodes, int mode,
                 int64_t node_base, int prec, int krr) {
    a = vmn::SearchArgs{};
    a.mode = mode;
    a.DP = idx->DP;
    a.vecs = idx->vecs.as<float>();
    a.adj = idx->adj.as<int32_t>();
    a.del_bits = idx->n_deleted ? idx->del_bits.as<uint32_t>() : nullptr;
    a.W = idx->p.R;
    a.queries = q;
    a.nq = nq;
    a.query_is_node = queries_are_nodes ? 1 : 0;
    a.node_base = node_base;
    a.start = (int32_t)idx->medoid;
    a.L = L;
    a.visited_cap = idx->p.visited_cap;
    a.max_hops = idx->p.max_hops;
    a.out_n = out_n;
    a.out_ids = out_ids;
    a.out_d = out_d;
    a.P = idx->p.P;
    a.prec = prec;
    a.krr = krr;
    if (prec == VMN_PREC_BF16) {
        a.low_rows = idx->bf16_rows.p;
        a.low_norms = idx->bf16_norms.as<float>();
    } else if (prec == VMN_PREC_FP8) {
        a.low_rows = idx->fp8_rows.p;
        a.low_norms = idx->fp8_norms.as<float>();
        a.adj_norms = idx->adj_fp8n.as<float>();
        a.fp8_scales = idx->fp8_scales.as<float>();
        a.fp8_centre = idx->fp8_centre.as<float>();
    }
    if (mode == VMN_MODE_FRESH && prec == VMN_PREC_FP8) a.adj_hash = idx->adj_hash.as<uint32_t>();
    a.lcap = idx->list_cap;
    // the tile shape applies where its kernels exist (FP32 gathers 16 rows a round)
    const int need = L > idx->list_cap ? L : idx->list_cap;
    a.tiles = prec != VMN_PREC_FP32 && idx->tiles == 4 && vmn::search_tiles4(idx->DP, need, mode) ? 4 : 2;
    // a batch that leaves most of the GPU idle is latency-bound: there the
    // BF16 row norms are read from the table (an extra 4-byte load per
    // candidate) instead of being summed from the row registers
    a.table_norms = nq <= (int64_t)idx->cus * 4;
}

int prepare_search(vmn_index* idx, int prec, int mode, bool sync) {
    const bool norms = prec == VMN_PREC_FP8 && idx->adj_fp8n_epoch != idx->epoch;
    // the FP8 FRESH search reads the neighbours' bloom cells
    const bool hashes = mode == VMN_MODE_FRESH && prec == VMN_PREC_FP8 &&
                        idx->adj_hash_epoch != idx->epoch;
    if (!norms && !hashes) return VMN_OK;
    // both allocated with the graph (alloc_graph); the norms for indexes that
    // keep FP8 rows
    if (norms && idx->adj_fp8n.bytes < (size_t)idx->cap * idx->p.R * sizeof(float))
        return fail(VMN_ERR_STATE, "FP8 norms array missing (index keeps no FP8 rows)");
    if (hashes && idx->adj_hash.bytes < (size_t)idx->cap * idx->p.R * sizeof(uint32_t))
        return fail(VMN_ERR_STATE, "adjacency hash array missing");
    if (norms)
        HIPCHK(vmn::launch_adj_norms(idx->adj.as<int32_t>(), idx->n * idx->p.R,
                                     idx->fp8_norms.as<float>(), idx->adj_fp8n.as<float>(),
                                     idx->stream));
    if (hashes)
        HIPCHK(vmn::launch_adj_hash(idx->adj.as<int32_t>(), idx->n * idx->p.R,
                                    idx->adj_hash.as<uint32_t>(), idx->stream));
    // searches on other streams (contexts) read them next
    if (sync) HIPCHK(hipStreamSynchronize(idx->stream));
    if (norms) idx->adj_fp8n_epoch = idx->epoch;
    if (hashes) idx->adj_hash_epoch = idx->epoch;
    return VMN_OK;
}

int search_grid(const vmn_index* idx, int64_t nq) {
    return (int)std::max<int64_t>(std::min<int64_t>(nq, (int64_t)idx->cus * idx->waves_per_cu), 1);
}

int run_search(vmn_index* idx, const float* q, int64_t nq, int L, int out_n, int32_t* out_ids,
               float* out_d, bool queries_are_nodes, int mode, bool want_vis, int64_t node_base,
               int prec, int krr, vmn_ctx* ctx) {
    hipStream_t stream = ctx ? ctx->stream : idx->stream;
    DevBuf& counter = ctx ? ctx->counter : idx->counter;
    DevBuf& stat_hops = ctx ? ctx->stat_hops : idx->stat_hops;
    DevBuf& stat_ndist = ctx ? ctx->stat_ndist : idx->stat_ndist;
    int rc = prepare_search(idx, prec, mode, true);
    if (rc) return rc;
    const int ng = groups_wanted(idx, nq) ? idx->query_groups : 0;
    if (ng) {
        bool fresh = false;
        rc = ensure_centroids(idx, &fresh);
        if (rc) return rc;
        // computed on the index stream; a context's stream reads them next
        if (fresh && ctx) HIPCHK(hipStreamSynchronize(idx->stream));
    }
    const vmn::GroupLayout lay = ng ? group_layout(idx, nq) : vmn::GroupLayout{false, 0, 0};
    HIPCHK(counter.reserve(ng ? (size_t)lay.nlists * 256 : 64));
    HIPCHK(vmn::launch_fill_u64(counter.as<unsigned long long>(), ng ? 32 * lay.nlists : 1, 0ull, stream));
    DevBuf& grp_order = ctx ? ctx->grp_order : idx->grp_order;
    if (ng) {
        rc = group_queries(idx, q, nq, grp_order, counter.as<unsigned long long>(), stream);
        if (rc) return rc;
    }
    HIPCHK(stat_hops.reserve((size_t)nq * sizeof(int32_t)));
    HIPCHK(stat_ndist.reserve((size_t)nq * sizeof(int32_t)));
    vmn::SearchArgs a;
    search_args(idx, a, q, nq, L, out_n, out_ids, out_d, queries_are_nodes, mode, node_base, prec,
                krr);
    if (want_vis) {
        HIPCHK(idx->vis_ids.reserve((size_t)nq * idx->p.P * sizeof(int32_t)));
        HIPCHK(idx->vis_d.reserve((size_t)nq * idx->p.P * sizeof(float)));
        HIPCHK(idx->vis_cnt.reserve((size_t)nq * sizeof(int32_t)));
        a.vis_ids = idx->vis_ids.as<int32_t>();
        a.vis_d = idx->vis_d.as<float>();
        a.vis_cnt = idx->vis_cnt.as<int32_t>();
    }
    a.stat_hops = stat_hops.as<int32_t>();
    a.stat_ndist = stat_ndist.as<int32_t>();
    a.counter = counter.as<unsigned long long>();
    if (ng) {
        a.order = grp_order.as<int32_t>();
        a.gstride = lay.stride;
        a.ngroups = ng;
        a.nlists = lay.nlists;
    }
    if (idx->build_count && !ctx) a.acc = idx->build_acc.as<unsigned long long>();
    HIPCHK(vmn::launch_search(a, search_grid(idx, nq), stream));
    if (ctx) ctx->last_nq = nq;
    else idx->last_nq = nq;
    return VMN_OK;
}

int prune_grid(const vmn_index* idx, int64_t n) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)idx->cus * 16));
}

int build_parth(vmn_index* idx, bool random_init) {
    const int64_t n = idx->n;
    const int L = idx->p.L_build;
    if (random_init)
        HIPCHK(vmn::launch_random_init(idx->adj.as<int32_t>(), n, idx->p.R, idx->p.R, idx->p.seed,
                                       idx->stream));
    HIPCHK(idx->cand_ids.reserve((size_t)n * L * sizeof(int32_t)));
    HIPCHK(idx->cand_d.reserve((size_t)n * L * sizeof(float)));
    for (int pass = 0; pass < idx->p.passes; ++pass) {
        int rc = run_search(idx, idx->vecs.as<float>(), n, L, L, idx->cand_ids.as<int32_t>(),
                            idx->cand_d.as<float>(), true, VMN_MODE_PARTH, false, 0,
                            build_prec(idx));
        if (rc) return rc;
        rc = zero_counter(idx);
        if (rc) return rc;
        vmn::PruneArgs a{};
        a.mode = VMN_MODE_PARTH;
        a.DP = idx->DP;
        a.vecs = idx->vecs.as<float>();
        set_prune_rows(idx, a);
        a.adj = idx->adj.as<int32_t>();
        a.W = idx->p.R;
        a.R = idx->p.R;
        a.alpha = idx->p.alpha[pass];
        a.n = n;
        a.node_base = 0;
        a.cand_ids = idx->cand_ids.as<int32_t>();
        a.cand_d = idx->cand_d.as<float>();
        a.Lc = L;
        a.counter = idx->counter.as<unsigned long long>();
        if (idx->build_count) a.acc = idx->build_acc.as<unsigned long long>();
        HIPCHK(vmn::launch_parth_prune(a, prune_grid(idx, n), idx->stream));
    }
    return VMN_OK;
}

// A FRESH prune over `nitems` items (rev: the reverse prune). The out-prune
// runs in two tiers when its full candidate list is longer than 256: a
// 256-entry launch (twice the resident waves) lists the items whose
// candidates could exceed it, and a full-list launch takes those. The reverse
// prune stays one launch: its large items are the hubs with hundreds of
// reverse candidates, and run on their own they form a tail longer than the
// time a first tier saves (2 x 28 ms -> 2 x 31.5 ms per 1M build). They are
// handed out first instead, so the light items fill in behind them.
// `a` is the caller's, by value.
int tiered_prune(vmn_index* idx, vmn::PruneArgs a, int64_t nitems, bool rev) {
    if (nitems <= 0) return VMN_OK;
    auto launch = [&](const vmn::PruneArgs& x, int64_t n) {
        return rev ? vmn::launch_fresh_rev_prune(x, prune_grid(idx, n), idx->stream)
                   : vmn::launch_fresh_out_prune(x, prune_grid(idx, n), idx->stream);
    };
    a.n = nitems;
    a.items = nullptr;
    a.overflow = nullptr;
    a.overflow_n = nullptr;
    const int full = rev ? a.M : a.P;
    int rc = zero_counter(idx);
    if (rc) return rc;
    if (rev) {
        // one launch, hubs first
        HIPCHK(idx->prune_ovf.reserve((size_t)nitems * sizeof(int64_t)));
        unsigned long long* cnts = idx->counter.as<unsigned long long>();
        HIPCHK(vmn::launch_fill_u64(cnts + 1, 2, 0ull, idx->stream));
        HIPCHK(vmn::launch_heavy_first(a.node_list, a.list_off, a.rev_off, nitems, a.W, 256,
                                       idx->prune_ovf.as<int64_t>(), cnts + 1, idx->stream));
        a.items = idx->prune_ovf.as<int64_t>();
        HIPCHK(launch(a, nitems));
        return VMN_OK;
    }
    // a batch too small to fill the GPU gains nothing from the first tier's
    // occupancy and would pay its host round trip (small inserts)
    if (full <= 256 || a.W >= 256 || nitems < (int64_t)idx->cus * 16) {
        HIPCHK(launch(a, nitems));
        return VMN_OK;
    }
    HIPCHK(idx->prune_ovf.reserve((size_t)nitems * sizeof(int64_t)));
    unsigned long long* cnts = idx->counter.as<unsigned long long>();
    HIPCHK(vmn::launch_fill_u64(cnts, 2, 0ull, idx->stream));
    a.overflow = idx->prune_ovf.as<int64_t>();
    a.overflow_n = cnts + 1;
    HIPCHK(launch(a, nitems));
    unsigned long long novf = 0;
    HIPCHK(hipMemcpyAsync(&novf, cnts + 1, sizeof novf, hipMemcpyDeviceToHost, idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    if (!novf) return VMN_OK;
    rc = zero_counter(idx);
    if (rc) return rc;
    a.overflow = nullptr;
    a.overflow_n = nullptr;
    a.items = idx->prune_ovf.as<int64_t>();
    a.n = (int64_t)novf;
    HIPCHK(launch(a, (int64_t)novf));
    return VMN_OK;
}

// One FRESH pass over nodes [base, base + cnt): search, out-prune emitting
// reverse pairs, group pairs by destination, reverse prune. The search and
// out-prune run in chunks when the visited sets of the whole range would not
// fit the build's scratch budget: a chunk's rows go to a shadow adjacency, and
// the shadow replaces the range's rows after the last chunk, so every search
// of the pass sees the graph from before it (one batch, as the reference's
// pass) and the result is the one-chunk result.
constexpr int64_t kVisitedBudget = (int64_t)24 << 30;  // bytes of visited sets per chunk

int64_t pass_chunk(const vmn_index* idx, int64_t cnt) {
    if (idx->p.build_chunk > 0) return std::min<int64_t>(idx->p.build_chunk, cnt);
    const int64_t per = (int64_t)idx->p.P * (int64_t)(sizeof(int32_t) + sizeof(float));
    return std::min<int64_t>(cnt, std::max<int64_t>(1 << 20, kVisitedBudget / per));
}

int fresh_pass(vmn_index* idx, int64_t base, int64_t cnt, float alpha) {
    const int R = idx->p.R;
    const int64_t N = idx->n;
    const int64_t chunk = pass_chunk(idx, cnt);
    const bool shadow = chunk < cnt;
    HIPCHK(idx->pair_dst.reserve((size_t)cnt * R * sizeof(int32_t)));
    HIPCHK(idx->pair_d.reserve((size_t)cnt * R * sizeof(float)));
    if (shadow) HIPCHK(idx->adj_shadow.reserve((size_t)cnt * R * sizeof(int32_t)));
    vmn::PruneArgs a{};
    a.mode = VMN_MODE_FRESH;
    a.DP = idx->DP;
    a.vecs = idx->vecs.as<float>();
    set_prune_rows(idx, a);
    a.adj = idx->adj.as<int32_t>();
    a.W = R;
    a.R = R;
    a.alpha = alpha;
    a.P = idx->p.P;
    a.del_bits = idx->n_deleted ? idx->del_bits.as<uint32_t>() : nullptr;
    a.M = idx->p.M;
    if (idx->build_count) a.acc = idx->build_acc.as<unsigned long long>();
    int rc = VMN_OK;
    for (int64_t c0 = 0; c0 < cnt; c0 += chunk) {
        const int64_t cs = std::min(chunk, cnt - c0);
        rc = run_search(idx, idx->vecs.as<float>() + (base + c0) * idx->DP, cs, idx->p.L_build, 0,
                        nullptr, nullptr, true, VMN_MODE_FRESH, true, base + c0, build_prec(idx));
        if (rc) return rc;
        vmn::PruneArgs c = a;
        // the work counter exists once a search has run (run_search allocates it)
        c.counter = idx->counter.as<unsigned long long>();
        c.node_base = base + c0;
        c.vis_ids = idx->vis_ids.as<int32_t>();
        c.vis_d = idx->vis_d.as<float>();
        c.vis_cnt = idx->vis_cnt.as<int32_t>();
        c.pair_dst = idx->pair_dst.as<int32_t>() + c0 * R;
        c.pair_d = idx->pair_d.as<float>() + c0 * R;
        c.adj_out = shadow ? idx->adj_shadow.as<int32_t>() + c0 * R : nullptr;
        rc = tiered_prune(idx, c, cs, false);
        if (rc) return rc;
    }
    if (shadow)
        HIPCHK(hipMemcpyAsync(idx->adj.as<int32_t>() + base * R, idx->adj_shadow.p,
                              (size_t)cnt * R * sizeof(int32_t), hipMemcpyDeviceToDevice,
                              idx->stream));
    a.n = cnt;
    a.node_base = base;
    a.pair_dst = idx->pair_dst.as<int32_t>();
    a.pair_d = idx->pair_d.as<float>();
    a.counter = idx->counter.as<unsigned long long>();

    const int64_t npairs = cnt * R;
    unsigned long long nlist = 0;
    if (npairs * 8 < N) {
        // a batch much smaller than the index: group the pairs by sorting them
        const size_t tb = vmn::reverse_group_temp_bytes(npairs);
        HIPCHK(idx->gtemp.reserve(tb));
        HIPCHK(idx->rsrc.reserve((size_t)npairs * sizeof(int32_t)));
        HIPCHK(idx->rd.reserve((size_t)npairs * sizeof(float)));
        HIPCHK(idx->rlist.reserve((size_t)npairs * sizeof(int64_t)));
        HIPCHK(idx->roff.reserve((size_t)(npairs + 1) * sizeof(int64_t)));
        HIPCHK(vmn::launch_reverse_group(idx->pair_dst.as<int32_t>(), idx->pair_d.as<float>(), npairs,
                                         R, base, idx->gtemp.p, idx->gtemp.bytes,
                                         idx->rsrc.as<int32_t>(), idx->rd.as<float>(),
                                         idx->rlist.as<int64_t>(), idx->roff.as<int64_t>(),
                                         idx->counter.as<unsigned long long>(), idx->stream));
        a.list_off = idx->roff.as<int64_t>();
    } else {
        // reverse pairs -> CSR by destination; per-node arrays are sized for
        // the index capacity so that growing the index does not reallocate them
        const int64_t NC = std::max<int64_t>(idx->cap, N);
        HIPCHK(idx->rcount.reserve((size_t)NC * sizeof(int32_t)));
        HIPCHK(idx->rfill.reserve((size_t)NC * sizeof(int32_t)));
        HIPCHK(idx->roff.reserve((size_t)(NC + 1) * sizeof(int64_t)));
        HIPCHK(idx->rscan_tmp.reserve((size_t)((NC + 1023) / 1024 + 2) * sizeof(int64_t)));
        HIPCHK(idx->rsrc.reserve((size_t)npairs * sizeof(int32_t)));
        HIPCHK(idx->rd.reserve((size_t)npairs * sizeof(float)));
        HIPCHK(idx->rlist.reserve((size_t)NC * sizeof(int64_t)));
        HIPCHK(vmn::launch_fill_i32(idx->rcount.as<int32_t>(), N, 0, idx->stream));
        HIPCHK(vmn::launch_fill_i32(idx->rfill.as<int32_t>(), N, 0, idx->stream));
        HIPCHK(vmn::launch_reverse_count(idx->pair_dst.as<int32_t>(), npairs,
                                         idx->rcount.as<int32_t>(), idx->stream));
        HIPCHK(vmn::launch_exclusive_scan(idx->rcount.as<int32_t>(), N, idx->roff.as<int64_t>(),
                                          idx->rscan_tmp.as<int64_t>(), idx->stream));
        HIPCHK(vmn::launch_reverse_scatter(idx->pair_dst.as<int32_t>(), idx->pair_d.as<float>(),
                                           npairs, R, base, idx->roff.as<int64_t>(),
                                           idx->rfill.as<int32_t>(), idx->rsrc.as<int32_t>(),
                                           idx->rd.as<float>(), idx->stream));
        rc = zero_counter(idx);
        if (rc) return rc;
        HIPCHK(vmn::launch_compact_nonzero(idx->rcount.as<int32_t>(), N, idx->rlist.as<int64_t>(),
                                           idx->counter.as<unsigned long long>(), idx->stream));
        a.rev_off = idx->roff.as<int64_t>();
        a.list_off = nullptr;
    }
    HIPCHK(hipMemcpyAsync(&nlist, idx->counter.p, sizeof nlist, hipMemcpyDeviceToHost, idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    rc = zero_counter(idx);
    if (rc) return rc;
    a.node_list = idx->rlist.as<int64_t>();
    a.rev_src = idx->rsrc.as<int32_t>();
    a.rev_d = idx->rd.as<float>();
    rc = tiered_prune(idx, a, (int64_t)nlist, true);
    // the graph changed: the next search lays out its adjacency-side arrays again
    ++idx->epoch;
    return rc;
}

int build_fresh(vmn_index* idx, bool random_init) {
    if (random_init)
        HIPCHK(vmn::launch_random_init(idx->adj.as<int32_t>(), idx->n, idx->p.R, idx->p.R,
                                       idx->p.seed, idx->stream));
    for (int pass = 0; pass < idx->p.passes; ++pass) {
        int rc = fresh_pass(idx, 0, idx->n, idx->p.alpha[pass]);
        if (rc) return rc;
    }
    return VMN_OK;
}

int set_tombstones(vmn_index* idx, const int32_t* dev_ids, int64_t m) {
    ++idx->epoch;
    HIPCHK(idx->bad.reserve(64));
    HIPCHK(vmn::launch_fill_u64(idx->bad.as<unsigned long long>(), 1, 0ull, idx->stream));
    HIPCHK(vmn::launch_tombstone(dev_ids, m, idx->n, idx->del_bits.as<uint32_t>(),
                                 idx->bad.as<unsigned long long>(), idx->stream));
    unsigned long long bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, idx->bad.p, sizeof bad, hipMemcpyDeviceToHost, idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    idx->n_deleted += m - (int64_t)bad;
    if (bad) return fail(VMN_ERR_ARG, "%llu ids outside [0, %lld) were ignored", bad, (long long)idx->n);
    return VMN_OK;
}

int normalize_rows(const int32_t* src, int64_t n, int row_width, int R, bool drop_self,
                   int32_t mem, std::vector<int32_t>& rows) {
    std::vector<int32_t> tmp;
    if (mem == VMN_MEM_DEVICE) {
        tmp.resize((size_t)n * row_width);
        HIPCHK(hipMemcpy(tmp.data(), src, tmp.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        src = tmp.data();
    }
    rows.assign((size_t)n * R, -1);
    for (int64_t i = 0; i < n; ++i) {
        int w = 0;
        for (int j = 0; j < row_width && w < R; ++j) {
            const int32_t v = src[i * row_width + j];
            if (v == -1) break;
            if (v < 0 || v >= n) return fail(VMN_ERR_ARG, "adjacency id %d out of range", v);
            if (drop_self && v == i) continue;
            bool dup = false;
            for (int t = 0; t < w; ++t) dup |= rows[i * R + t] == v;
            if (!dup) rows[i * R + w++] = v;
        }
    }
    return VMN_OK;
}

int check_params(const vmn_params* p) {
    if (!p) return fail(VMN_ERR_ARG, "params is NULL");
    if (p->dim < 1 || p->dim > 256) return fail(VMN_ERR_ARG, "dim %d not in [1, 256]", p->dim);
    if (p->R < 1 || p->R > 64) return fail(VMN_ERR_ARG, "R %d not in [1, 64]", p->R);
    if (p->L_build < 1 || p->L_build > 512) return fail(VMN_ERR_ARG, "L_build %d not in [1, 512]", p->L_build);
    if (p->passes < 0 || p->passes > 4) return fail(VMN_ERR_ARG, "passes %d not in [0, 4]", p->passes);
    if (p->mode != VMN_MODE_PARTH && p->mode != VMN_MODE_FRESH) return fail(VMN_ERR_ARG, "bad mode %d", p->mode);
    if (p->visited_cap < 1 || p->max_hops < 1) return fail(VMN_ERR_ARG, "visited_cap/max_hops must be >= 1");
    if (p->P < 1 || p->P > 640) return fail(VMN_ERR_ARG, "P %d not in [1, 640]", p->P);
    if (p->M < 1 || p->M > 512) return fail(VMN_ERR_ARG, "M %d not in [1, 512]", p->M);
    if (p->mode == VMN_MODE_PARTH && p->L_build + p->R > 576)
        return fail(VMN_ERR_ARG, "L_build + R must be <= 576");
    if (p->build_chunk < 0) return fail(VMN_ERR_ARG, "build_chunk must be >= 0");
    return VMN_OK;
}

}  // namespace vmn_rt

using namespace vmn_rt;

extern "C" {

void vmn_default_params(vmn_params* p, int32_t mode) {
    std::memset(p, 0, sizeof *p);
    p->dim = 128;
    p->mode = mode;
    p->seed = 1234;
    p->visited_cap = 128;
    p->P = 600;
    p->M = 500;
    p->device = 0;
    if (mode == VMN_MODE_FRESH) {
        p->R = 64;
        p->L_build = 128;
        p->passes = 1;
        p->alpha[0] = p->alpha[1] = p->alpha[2] = p->alpha[3] = 1.5f;
        p->max_hops = 1 << 20;
    } else {
        p->R = 32;
        p->L_build = 64;
        p->passes = 2;
        p->alpha[0] = p->alpha[1] = p->alpha[2] = p->alpha[3] = 1.2f;
        p->max_hops = 200;
    }
}

int vmn_create(const vmn_params* p, vmn_index** out) {
    if (!out) return fail(VMN_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_params(p);
    if (rc) return rc;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (p->device < 0 || p->device >= ndev) return fail(VMN_ERR_ARG, "device %d of %d", p->device, ndev);
    HIPCHK(hipSetDevice(p->device));
    auto* idx = new vmn_index();
    idx->p = *p;
    idx->DP = pad_dim(p->dim);
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, p->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&idx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete idx;
        return fail(VMN_ERR_HIP, "device setup: %s", hipGetErrorString(e));
    }
    idx->cus = prop.multiProcessorCount;
    *out = idx;
    return VMN_OK;
}

int vmn_destroy(vmn_index* idx) {
    if (!idx) return VMN_OK;
    (void)hipSetDevice(idx->p.device);
    (void)hipStreamSynchronize(idx->stream);
    (void)hipStreamDestroy(idx->stream);
    delete idx;  // every device buffer is freed by its destructor
    return VMN_OK;
}

namespace vmn_rt {

// The build's algorithmic bytes from its work counters (the byte model of
// vmn_build_stats in the header).
static void set_build_stats(vmn_index* idx, const unsigned long long* acc, bool random_init) {
    vmn_build_stats& b = idx->build_stats;
    b = vmn_build_stats{};
    const double n = (double)idx->n, DP = idx->DP, R = idx->p.R;
    const bool bf = idx->exact_bf16;
    b.searches = (int64_t)acc[VMN_ACC_SEARCHES];
    b.expansions = (int64_t)acc[VMN_ACC_EXPANSIONS];
    b.dist_evals = (int64_t)acc[VMN_ACC_DISTS];
    b.prune_items = (int64_t)acc[VMN_ACC_ITEMS];
    b.prune_rows = (int64_t)acc[VMN_ACC_ROWS];
    b.reverse_pairs = (int64_t)acc[VMN_ACC_PAIRS];
    b.list_entries = (int64_t)acc[VMN_ACC_LISTS];
    b.reverse_entries = (int64_t)acc[VMN_ACC_REVS];
    b.row_bytes = (int32_t)(bf ? 2 * DP : 4 * DP);
    b.passes = idx->p.passes;
    double by = 8.0 * n * DP;                        // medoid: centroid and distance passes
    if (random_init) by += 4.0 * n * R;              // random initial graph
    if ((idx->p.low_precision & VMN_STORE_BF16) || bf) by += n * (4 * DP + 2 * DP + 4);
    if (idx->p.low_precision & VMN_STORE_FP8) by += n * (3 * 4 * DP + DP + 4);
    by += (double)b.searches * 4 * DP;               // the searched nodes' own rows
    by += (double)b.expansions * (4 * R + 8);        // adjacency rows, visited entries written
    by += (double)b.dist_evals * b.row_bytes;        // candidate rows
    by += (double)b.prune_rows * (b.row_bytes + (bf ? 4 : 0));  // prune rows (+ BF16 norms)
    by += (double)b.list_entries * 8;                // candidate lists the prunes read
    by += (double)b.prune_items * 8 * R;             // rows read and written by the prunes
    by += (double)b.reverse_pairs * 24;              // pairs written, grouped, read
    by += (double)b.reverse_entries * 8;             // reverse candidates the prunes read
    b.algorithmic_bytes = by;
}

// vmn_build / vmn_build_from: `init` (host rows, -1 padded to R) replaces the
// seeded random initial graph when not NULL.
int build_common(vmn_index* idx, const float* vecs, int64_t n, int32_t mem,
                 const std::vector<int32_t>* init) {
    int rc = set_device(idx);
    if (rc) return rc;
    idx->built = false;
    rc = alloc_graph(idx, n);
    if (rc) return rc;
    idx->n = n;
    const float* staged = nullptr;
    rc = stage_rows(idx, vecs, n, mem, idx->vecs, &staged);
    if (rc) return rc;
    if (staged != idx->vecs.as<float>())
        HIPCHK(hipMemcpyAsync(idx->vecs.p, staged, (size_t)n * idx->DP * sizeof(float),
                              hipMemcpyDeviceToDevice, idx->stream));
    if (init)
        HIPCHK(hipMemcpyAsync(idx->adj.p, init->data(), init->size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, idx->stream));
    rc = compute_medoid(idx);
    if (rc) return rc;
    rc = update_exact(idx, 0, n);
    if (rc) return rc;
    idx->fp8_scaled = false;
    rc = refresh_lowp(idx, 0, n);
    if (rc) return rc;
    HIPCHK(idx->build_acc.reserve(VMN_ACC_SLOTS * VMN_ACC_N * sizeof(unsigned long long)));
    HIPCHK(vmn::launch_fill_u64(idx->build_acc.as<unsigned long long>(),
                                VMN_ACC_SLOTS * VMN_ACC_N, 0ull, idx->stream));
    idx->build_count = true;
    rc = idx->p.mode == VMN_MODE_PARTH ? build_parth(idx, !init) : build_fresh(idx, !init);
    idx->build_count = false;
    if (rc) return rc;
    unsigned long long slots[VMN_ACC_SLOTS * VMN_ACC_N] = {}, acc[VMN_ACC_N] = {};
    HIPCHK(hipMemcpyAsync(slots, idx->build_acc.p, sizeof slots, hipMemcpyDeviceToHost,
                          idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    for (int s = 0; s < VMN_ACC_SLOTS; ++s)
        for (int k = 0; k < VMN_ACC_N; ++k) acc[k] += slots[s * VMN_ACC_N + k];
    set_build_stats(idx, acc, !init);
    idx->built = true;
    return VMN_OK;
}


int check_build_args(const vmn_index* idx, const float* vecs, int64_t n) {
    if (!idx || !vecs) return fail(VMN_ERR_ARG, "NULL argument");
    if (n <= idx->p.R) return fail(VMN_ERR_ARG, "n = %lld must exceed R = %d", (long long)n, idx->p.R);
    if (n >= (int64_t)1 << 31) return fail(VMN_ERR_ARG, "n must be < 2^31");
    if (idx->p.capacity && n > idx->p.capacity) return fail(VMN_ERR_CAPACITY, "n exceeds capacity");
    return VMN_OK;
}

}  // namespace vmn_rt

int vmn_build(vmn_index* idx, const float* vecs, int64_t n, int32_t mem) {
    int rc = check_build_args(idx, vecs, n);
    if (rc) return rc;
    return build_common(idx, vecs, n, mem, nullptr);
}

int vmn_build_from(vmn_index* idx, const float* vecs, int64_t n, const int32_t* init_adj,
                   int32_t row_width, int32_t mem) {
    int rc = check_build_args(idx, vecs, n);
    if (rc) return rc;
    if (!init_adj) return build_common(idx, vecs, n, mem, nullptr);
    if (row_width < 1) return fail(VMN_ERR_ARG, "row_width %d < 1", row_width);
    rc = set_device(idx);
    if (rc) return rc;
    std::vector<int32_t> rows;
    rc = normalize_rows(init_adj, n, row_width, idx->p.R, true, mem, rows);
    if (rc) return rc;
    return build_common(idx, vecs, n, mem, &rows);
}

int vmn_get_vectors(const vmn_index* idx, float* out, int32_t mem) {
    if (!idx || !out) return fail(VMN_ERR_ARG, "NULL argument");
    if (!idx->built) return fail(VMN_ERR_STATE, "index not built");
    int rc = set_device(idx);
    if (rc) return rc;
    const int D = idx->p.dim, DP = idx->DP;
    HIPCHK(hipStreamSynchronize(idx->stream));
    HIPCHK(hipMemcpy2D(out, (size_t)D * sizeof(float), idx->vecs.p, (size_t)DP * sizeof(float),
                       (size_t)D * sizeof(float), (size_t)idx->n,
                       mem == VMN_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost));
    return VMN_OK;
}

int vmn_get_params(const vmn_index* idx, vmn_params* out) {
    if (!idx || !out) return fail(VMN_ERR_ARG, "NULL argument");
    *out = idx->p;
    return VMN_OK;
}

int vmn_load_graph(vmn_index* idx, const float* vecs, int64_t n, const int32_t* adj,
                   int32_t row_width, int64_t medoid, int32_t mem) {
    if (!idx || !vecs || !adj) return fail(VMN_ERR_ARG, "NULL argument");
    if (n < 1 || n >= (int64_t)1 << 31) return fail(VMN_ERR_ARG, "bad n");
    if (row_width < 1 || row_width > idx->p.R) return fail(VMN_ERR_ARG, "row_width %d > R %d", row_width, idx->p.R);
    if (medoid >= n) return fail(VMN_ERR_ARG, "medoid out of range");
    int rc = set_device(idx);
    if (rc) return rc;
    rc = alloc_graph(idx, n);
    if (rc) return rc;
    idx->n = n;
    const float* staged = nullptr;
    rc = stage_rows(idx, vecs, n, mem, idx->vecs, &staged);
    if (rc) return rc;
    if (staged != idx->vecs.as<float>())
        HIPCHK(hipMemcpyAsync(idx->vecs.p, staged, (size_t)n * idx->DP * sizeof(float),
                              hipMemcpyDeviceToDevice, idx->stream));
    // rows are widened to R with -1 on the host side of the copy
    std::vector<int32_t> rows;
    rc = normalize_rows(adj, n, row_width, idx->p.R, false, mem, rows);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(idx->adj.p, rows.data(), rows.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, idx->stream));
    if (medoid < 0) {
        rc = compute_medoid(idx);
        if (rc) return rc;
    } else {
        idx->medoid = medoid;
    }
    rc = update_exact(idx, 0, n);
    if (rc) return rc;
    idx->fp8_scaled = false;
    rc = refresh_lowp(idx, 0, n);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(idx->stream));
    idx->built = true;
    return VMN_OK;
}

int vmn_search_ex(vmn_index* idx, const float* queries, int64_t nq, int32_t k, int32_t L,
                  int32_t prec, int32_t k_rerank, int32_t* ids_out, float* dists_out,
                  int32_t mem) {
    if (!idx || !queries || !ids_out) return fail(VMN_ERR_ARG, "NULL argument");
    if (!idx->built) return fail(VMN_ERR_STATE, "index has no graph");
    if (nq < 0 || k < 1 || k > L || L < 1 || L > 512)
        return fail(VMN_ERR_ARG, "need 1 <= k <= L <= 512 (k=%d, L=%d)", k, L);
    if (prec == VMN_PREC_BF16 && !(idx->p.low_precision & VMN_STORE_BF16))
        return fail(VMN_ERR_STATE, "index keeps no BF16 rows (low_precision)");
    if (prec == VMN_PREC_FP8 && !(idx->p.low_precision & VMN_STORE_FP8))
        return fail(VMN_ERR_STATE, "index keeps no FP8 rows (low_precision)");
    if (prec < VMN_PREC_FP32 || prec > VMN_PREC_FP8) return fail(VMN_ERR_ARG, "bad precision %d", prec);
    if (k_rerank != 0 && (k_rerank < k || k_rerank > 64 || k_rerank > L))
        return fail(VMN_ERR_ARG, "need k <= k_rerank <= min(64, L) (k_rerank=%d)", k_rerank);
    if (nq == 0) return VMN_OK;
    int rc = set_device(idx);
    if (rc) return rc;
    const float* q = nullptr;
    rc = stage_rows(idx, queries, nq, mem, idx->qpad, &q);
    if (rc) return rc;
    int32_t* oi = ids_out;
    float* od = dists_out;
    if (mem == VMN_MEM_HOST) {
        HIPCHK(idx->out_ids.reserve((size_t)nq * k * sizeof(int32_t)));
        HIPCHK(idx->out_d.reserve((size_t)nq * k * sizeof(float)));
        oi = idx->out_ids.as<int32_t>();
        od = idx->out_d.as<float>();
    }
    rc = search_into(idx, q, nq, k, L, prec, prec == VMN_PREC_FP32 ? 0 : k_rerank, oi, od);
    if (rc) return rc;
    if (mem == VMN_MEM_HOST) {
        HIPCHK(hipMemcpyAsync(ids_out, oi, (size_t)nq * k * sizeof(int32_t), hipMemcpyDeviceToHost,
                              idx->stream));
        if (dists_out)
            HIPCHK(hipMemcpyAsync(dists_out, od, (size_t)nq * k * sizeof(float),
                                  hipMemcpyDeviceToHost, idx->stream));
        HIPCHK(hipStreamSynchronize(idx->stream));
    }
    return VMN_OK;
}

int vmn_search(vmn_index* idx, const float* queries, int64_t nq, int32_t k, int32_t L,
               int32_t* ids_out, float* dists_out, int32_t mem) {
    return vmn_search_ex(idx, queries, nq, k, L, VMN_PREC_FP32, 0, ids_out, dists_out, mem);
}

int vmn_search_vectors(vmn_index* idx, const float* queries, int64_t nq, int32_t k, int32_t L,
                       float* vecs_out, int32_t mem) {
    if (!idx || !queries || !vecs_out) return fail(VMN_ERR_ARG, "NULL argument");
    if (!idx->built) return fail(VMN_ERR_STATE, "index has no graph");
    if (nq < 0 || k < 1 || k > L || L < 1 || L > 512)
        return fail(VMN_ERR_ARG, "need 1 <= k <= L <= 512 (k=%d, L=%d)", k, L);
    if (nq == 0) return VMN_OK;
    int rc = set_device(idx);
    if (rc) return rc;
    const float* q = nullptr;
    rc = stage_rows(idx, queries, nq, mem, idx->qpad, &q);
    if (rc) return rc;
    HIPCHK(idx->out_ids.reserve((size_t)nq * k * sizeof(int32_t)));
    rc = search_into(idx, q, nq, k, L, VMN_PREC_FP32, 0, idx->out_ids.as<int32_t>(), nullptr);
    if (rc) return rc;
    const int D = idx->p.dim;
    float* dst = vecs_out;
    if (mem == VMN_MEM_HOST) {
        HIPCHK(idx->staging.reserve((size_t)nq * k * D * sizeof(float)));
        dst = idx->staging.as<float>();
    }
    HIPCHK(vmn::launch_rows_by_id(idx->out_ids.as<int32_t>(), nq * k, idx->vecs.as<float>(),
                                  idx->DP, D, idx->pend.as<float>(), idx->n, dst, idx->stream));
    if (mem == VMN_MEM_HOST) {
        HIPCHK(hipMemcpyAsync(vecs_out, dst, (size_t)nq * k * D * sizeof(float),
                              hipMemcpyDeviceToHost, idx->stream));
        HIPCHK(hipStreamSynchronize(idx->stream));
    }
    return VMN_OK;
}

int vmn_get_lowp_rows(const vmn_index* idx, int32_t prec, float* out) {
    if (!idx || !out) return fail(VMN_ERR_ARG, "NULL argument");
    const int DP = idx->DP, D = idx->p.dim;
    const int64_t n = idx->n;
    int rc = set_device(idx);
    if (rc) return rc;
    std::vector<float> rows((size_t)n * DP);
    if (prec == VMN_PREC_BF16) {
        if (!(idx->p.low_precision & VMN_STORE_BF16)) return fail(VMN_ERR_STATE, "no BF16 rows");
        std::vector<uint16_t> h((size_t)n * DP);
        HIPCHK(hipMemcpy(h.data(), idx->bf16_rows.p, h.size() * 2, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < h.size(); ++i) {
            uint32_t w = (uint32_t)h[i] << 16;
            std::memcpy(&rows[i], &w, 4);
        }
    } else if (prec == VMN_PREC_FP8) {
        if (!(idx->p.low_precision & VMN_STORE_FP8)) return fail(VMN_ERR_STATE, "no FP8 rows");
        DevBuf tmp;
        HIPCHK(tmp.reserve(rows.size() * sizeof(float)));
        HIPCHK(vmn::launch_fp8_dequant(idx->fp8_rows.p, n, DP, idx->fp8_scales.as<float>(),
                                       idx->fp8_centre.as<float>(), tmp.as<float>(), idx->stream));
        HIPCHK(hipMemcpyAsync(rows.data(), tmp.p, rows.size() * sizeof(float), hipMemcpyDeviceToHost,
                              idx->stream));
        HIPCHK(hipStreamSynchronize(idx->stream));
        tmp.release();
    } else {
        return fail(VMN_ERR_ARG, "prec must be BF16 or FP8");
    }
    for (int64_t i = 0; i < n; ++i) std::memcpy(out + i * D, &rows[i * DP], D * sizeof(float));
    return VMN_OK;
}

int vmn_get_adjacency(const vmn_index* idx, int32_t* out, int32_t mem) {
    if (!idx || !out) return fail(VMN_ERR_ARG, "NULL argument");
    if (!idx->built) return fail(VMN_ERR_STATE, "index has no graph");
    int rc = set_device(idx);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, idx->adj.p, (size_t)idx->n * idx->p.R * sizeof(int32_t),
                          mem == VMN_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                          idx->stream));
    if (mem == VMN_MEM_HOST) HIPCHK(hipStreamSynchronize(idx->stream));
    return VMN_OK;
}

int64_t vmn_size(const vmn_index* idx) { return idx ? idx->n : -1; }
int64_t vmn_medoid(const vmn_index* idx) { return idx ? idx->medoid : -1; }
int vmn_get_build_stats(const vmn_index* idx, vmn_build_stats* out) {
    if (!idx || !out) return fail(VMN_ERR_ARG, "NULL argument");
    if (!idx->built) return fail(VMN_ERR_STATE, "index not built");
    *out = idx->build_stats;
    return VMN_OK;
}

int32_t vmn_build_precision(const vmn_index* idx) {
    return idx ? build_prec(idx) : VMN_PREC_FP32;
}
void* vmn_stream(const vmn_index* idx) { return idx ? (void*)idx->stream : nullptr; }

int vmn_search_stats(const vmn_index* idx, int64_t* expansions, int64_t* dist_evals) {
    if (!idx) return fail(VMN_ERR_ARG, "NULL index");
    int rc = set_device(idx);
    if (rc) return rc;
    const int64_t nq = idx->last_nq;
    std::vector<int32_t> h(nq), d(nq);
    if (nq) {
        HIPCHK(hipMemcpyAsync(h.data(), idx->stat_hops.p, nq * sizeof(int32_t), hipMemcpyDeviceToHost,
                              idx->stream));
        HIPCHK(hipMemcpyAsync(d.data(), idx->stat_ndist.p, nq * sizeof(int32_t),
                              hipMemcpyDeviceToHost, idx->stream));
        HIPCHK(hipStreamSynchronize(idx->stream));
    }
    int64_t sh = 0, sd = 0;
    for (int64_t i = 0; i < nq; ++i) {
        sh += h[i];
        sd += d[i];
    }
    if (expansions) *expansions = sh;
    if (dist_evals) *dist_evals = sd;
    return VMN_OK;
}

// FP64 exact k nearest of nq padded queries among n padded rows (device), by
// (distance, id), into device out_i / out_d. Widths up to 128 and k <= 64
// take the MFMA-filtered path (same results; the queries its error bound
// cannot settle go through the direct kernel); the rest the direct kernel.
int exact_knn(vmn_index* idx, const float* base, int64_t n, const float* q, int64_t nq, int k,
              int32_t* out_i, double* out_d) {
    const int DP = idx->DP;
    auto direct = [&](const float* qq, int64_t m, int32_t* oi, double* od) -> int {
        int splits = (int)std::max<int64_t>(1, std::min<int64_t>(256, (int64_t)idx->cus * 256 / std::max<int64_t>(m, 1)));
        splits = (int)std::min<int64_t>(splits, n);
        const int KP = k <= 16 ? 16 : 128;
        DevBuf pd, pi;
        HIPCHK(pd.reserve((size_t)m * splits * KP * sizeof(double)));
        HIPCHK(pi.reserve((size_t)m * splits * KP * sizeof(int32_t)));
        HIPCHK(vmn::launch_gt(base, n, DP, nullptr, qq, m, k, splits, pd.as<double>(), pi.as<int32_t>(),
                              oi, od, idx->stream));
        HIPCHK(hipStreamSynchronize(idx->stream));
        return VMN_OK;
    };
    idx->gt_fallback = 0;
    if (DP > 128 || k > 64 || n < 4096) return direct(q, nq, out_i, out_d);
    const int S = vmn::gt_fast_splits(n, nq, idx->cus);
    DevBuf scratch, misc;
    HIPCHK(scratch.reserve(vmn::gt_fast_scratch_bytes(n, nq, S)));
    HIPCHK(misc.reserve(64));
    int32_t* fb = nullptr;
    HIPCHK(vmn::launch_gt_fast(base, n, DP, nullptr, q, nq, k, S, scratch.p, out_i, out_d,
                               misc.as<unsigned int>(), &fb, idx->stream));
    unsigned int nfb = 0;
    HIPCHK(hipMemcpyAsync(&nfb, misc.as<unsigned int>() + 1, sizeof nfb, hipMemcpyDeviceToHost,
                          idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    idx->gt_fallback = nfb;
    if (!nfb) return VMN_OK;
    DevBuf qf, fi, fd;
    HIPCHK(qf.reserve((size_t)nfb * DP * sizeof(float)));
    HIPCHK(fi.reserve((size_t)nfb * k * sizeof(int32_t)));
    HIPCHK(fd.reserve((size_t)nfb * k * sizeof(double)));
    HIPCHK(vmn::launch_gt_gather(q, fb, nfb, DP, qf.as<float>(), idx->stream));
    int rc = direct(qf.as<float>(), nfb, fi.as<int32_t>(), fd.as<double>());
    if (rc) return rc;
    HIPCHK(vmn::launch_gt_scatter(fb, nfb, k, fi.as<int32_t>(), fd.as<double>(), out_i, out_d, idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    return VMN_OK;
}

thread_local int64_t g_gt_fallbacks = 0;

int64_t vmn_groundtruth_fallbacks(void) { return g_gt_fallbacks; }

int vmn_groundtruth(const float* base, int64_t n, int32_t dim, const float* queries, int64_t nq,
                    int32_t k, int32_t* ids_out, double* dists_out, int32_t mem, int32_t device) {
    if (!base || !queries || !ids_out) return fail(VMN_ERR_ARG, "NULL argument");
    if (dim < 1 || dim > 256 || k < 1 || k > 128 || n < 1 || nq < 0)
        return fail(VMN_ERR_ARG, "bad shape");
    if (nq == 0) return VMN_OK;
    HIPCHK(hipSetDevice(device));
    vmn_params p;
    vmn_default_params(&p, VMN_MODE_PARTH);
    p.dim = dim;
    p.device = device;
    vmn_index* idx = nullptr;
    int rc = vmn_create(&p, &idx);
    if (rc) return rc;
    struct Guard {
        vmn_index* i;
        ~Guard() { vmn_destroy(i); }
    } guard{idx};
    const float* b = nullptr;
    const float* q = nullptr;
    rc = stage_rows(idx, base, n, mem, idx->vecs, &b);
    if (rc) return rc;
    rc = stage_rows(idx, queries, nq, mem, idx->qpad, &q);
    if (rc) return rc;
    DevBuf oi, od;
    int32_t* dst_i = ids_out;
    double* dst_d = dists_out;
    if (mem == VMN_MEM_HOST) {
        HIPCHK(oi.reserve((size_t)nq * k * sizeof(int32_t)));
        HIPCHK(od.reserve((size_t)nq * k * sizeof(double)));
        dst_i = oi.as<int32_t>();
        dst_d = od.as<double>();
    }
    rc = exact_knn(idx, b, n, q, nq, k, dst_i, dst_d);
    g_gt_fallbacks = idx->gt_fallback;
    if (rc) return rc;
    if (mem == VMN_MEM_HOST) {
        HIPCHK(hipMemcpyAsync(ids_out, dst_i, (size_t)nq * k * sizeof(int32_t), hipMemcpyDeviceToHost,
                              idx->stream));
        if (dists_out)
            HIPCHK(hipMemcpyAsync(dists_out, dst_d, (size_t)nq * k * sizeof(double),
                                  hipMemcpyDeviceToHost, idx->stream));
    }
    HIPCHK(hipStreamSynchronize(idx->stream));
    oi.release();
    od.release();
    return VMN_OK;
}

double vmn_recall(const int32_t* found, int32_t k_found, const int32_t* gt, const float* gt_dists,
                  int32_t k_gt, int64_t nq, int32_t recall_at) {
    if (!found || !gt || nq <= 0 || recall_at < 1 || k_found < recall_at || k_gt < recall_at)
        return fail(VMN_ERR_ARG, "bad recall arguments"), -1.0;
    double total = 0.0;
    std::vector<int32_t> res;
    for (int64_t i = 0; i < nq; ++i) {
        const int32_t* g = gt + i * k_gt;
        int64_t m = recall_at;
        if (gt_dists) {
            const float* gd = gt_dists + i * k_gt;
            m = recall_at - 1;
            while (m < k_gt && gd[m] == gd[recall_at - 1]) ++m;
        }
        res.assign(found + i * k_found, found + i * k_found + recall_at);
        std::sort(res.begin(), res.end());
        res.erase(std::unique(res.begin(), res.end()), res.end());
        std::vector<int32_t> gs(g, g + m);
        std::sort(gs.begin(), gs.end());
        gs.erase(std::unique(gs.begin(), gs.end()), gs.end());
        int hits = 0;
        for (int32_t v : gs)
            if (v >= 0 && std::binary_search(res.begin(), res.end(), v)) ++hits;
        total += hits;
    }
    return total / ((double)nq * recall_at);
}

}  // extern "C"

namespace vmn_rt {

// FreshVamana insert of n rows: host or device rows of width dim (`mem`), or,
// with padded, device rows of stride DP.
int insert_rows(vmn_index* idx, const float* vecs, int64_t n, int32_t* ids_out, int32_t mem,
                bool padded) {
    int rc = set_device(idx);
    if (rc) return rc;
    ++idx->epoch;
    const int64_t base = idx->n;
    const int DP = idx->DP;
    float* dst = idx->vecs.as<float>() + base * DP;
    if (padded) {
        HIPCHK(hipMemcpyAsync(dst, vecs, (size_t)n * DP * sizeof(float), hipMemcpyDeviceToDevice,
                              idx->stream));
    } else if (mem == VMN_MEM_DEVICE) {
        HIPCHK(vmn::launch_pad_rows(vecs, n, idx->p.dim, dst, DP, idx->stream));
    } else {
        const float* staged = nullptr;
        rc = stage_rows(idx, vecs, n, mem, idx->staging, &staged);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(dst, staged, (size_t)n * DP * sizeof(float), hipMemcpyDeviceToDevice,
                              idx->stream));
    }
    HIPCHK(vmn::launch_fill_i32(idx->adj.as<int32_t>() + base * idx->p.R, n * idx->p.R, -1,
                                idx->stream));
    idx->n = base + n;
    rc = update_exact(idx, base, base + n);
    if (rc) return rc;
    rc = refresh_lowp(idx, base, base + n);
    if (rc) return rc;
    const int last = idx->p.passes > 0 ? idx->p.passes - 1 : 0;
    rc = fresh_pass(idx, base, n, idx->p.alpha[last]);
    if (rc) return rc;
    std::vector<int32_t> ids(n);
    for (int64_t i = 0; i < n; ++i) ids[i] = (int32_t)(base + i);
    if (ids_out && mem == VMN_MEM_HOST) std::memcpy(ids_out, ids.data(), n * sizeof(int32_t));
    if (ids_out && mem == VMN_MEM_DEVICE)
        HIPCHK(hipMemcpyAsync(ids_out, ids.data(), n * sizeof(int32_t), hipMemcpyHostToDevice,
                              idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    return VMN_OK;
}

int apply_staged(vmn_index* idx) {
    const int64_t m = idx->n_pend;
    if (!m) return VMN_OK;
    idx->n_pend = 0;
    return insert_rows(idx, idx->pend.as<float>(), m, nullptr, VMN_MEM_DEVICE, true);
}

// Grows `b` to `want` bytes keeping its first `used` bytes.
int grow_keep(DevBuf& b, size_t want, size_t used, hipStream_t s) {
    if (want <= b.bytes) return VMN_OK;
    DevBuf nb;
    HIPCHK(nb.reserve(std::max(want, b.bytes * 2)));
    if (used) HIPCHK(hipMemcpyAsync(nb.p, b.p, used, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    b.release();
    b.p = nb.p;
    b.bytes = nb.bytes;
    nb.p = nullptr;
    nb.bytes = 0;
    return VMN_OK;
}

// Search of nq padded query rows into device buffers ri / rd (rd may be
// NULL); with staged rows present, their exact k nearest (FP64 brute force)
// are merged into the graph's k by (distance, id), as merge_and_rerank_kernel
// (utils.cuh:199-285) merges the insert list into the worklist.
int search_into(vmn_index* idx, const float* q, int64_t nq, int k, int L, int prec, int krr,
                int32_t* ri, float* rd) {
    if (!idx->n_pend)
        return run_search(idx, q, nq, L, k, ri, rd, false, idx->p.mode, false, 0, prec, krr);
    if (k > 128) return fail(VMN_ERR_ARG, "searches over staged inserts take k <= 128");
    HIPCHK(idx->pend_i.reserve((size_t)nq * k * sizeof(int32_t)));
    HIPCHK(idx->pend_d.reserve((size_t)nq * k * sizeof(float)));
    int rc = run_search(idx, q, nq, L, k, idx->pend_i.as<int32_t>(), idx->pend_d.as<float>(), false,
                        idx->p.mode, false, 0, prec, krr);
    if (rc) return rc;
    const int64_t m = idx->n_pend;
    int splits = (int)std::max<int64_t>(
        1, std::min<int64_t>(256, (int64_t)idx->cus * 256 / std::max<int64_t>(nq, 1)));
    splits = (int)std::min<int64_t>(splits, m);
    const int KP = k <= 16 ? 16 : 128;
    HIPCHK(idx->pend_part_d.reserve((size_t)nq * splits * KP * sizeof(double)));
    HIPCHK(idx->pend_part_i.reserve((size_t)nq * splits * KP * sizeof(int32_t)));
    HIPCHK(idx->pend_gi.reserve((size_t)nq * k * sizeof(int32_t)));
    HIPCHK(idx->pend_gd.reserve((size_t)nq * k * sizeof(double)));
    HIPCHK(vmn::launch_gt(idx->pend.as<float>(), m, idx->DP, nullptr, q, nq, k, splits,
                          idx->pend_part_d.as<double>(), idx->pend_part_i.as<int32_t>(),
                          idx->pend_gi.as<int32_t>(), idx->pend_gd.as<double>(), idx->stream));
    HIPCHK(vmn::launch_merge_staged(idx->pend_i.as<int32_t>(), idx->pend_d.as<float>(),
                                    idx->pend_gi.as<int32_t>(), idx->pend_gd.as<double>(), nq, k,
                                    idx->n, ri, rd, idx->stream));
    return VMN_OK;
}

}  // namespace vmn_rt

extern "C" {

int vmn_insert(vmn_index* idx, const float* vecs, int64_t n, int32_t* ids_out, int32_t mem) {
    if (!idx || (!vecs && n)) return fail(VMN_ERR_ARG, "NULL argument");
    if (!idx->built) return fail(VMN_ERR_STATE, "index has no graph");
    if (n < 0) return fail(VMN_ERR_ARG, "negative count");
    if (n == 0) return VMN_OK;
    if (idx->n + idx->n_pend + n > idx->cap)
        return fail(VMN_ERR_CAPACITY, "insert of %lld exceeds capacity %lld (size %lld, staged %lld)",
                    (long long)n, (long long)idx->cap, (long long)idx->n, (long long)idx->n_pend);
    // staged rows were promised the next ids: they go in first
    int rc = apply_staged(idx);
    if (rc) return rc;
    return insert_rows(idx, vecs, n, ids_out, mem, false);
}

int vmn_insert_staged(vmn_index* idx, const float* vecs, int64_t n, int32_t* ids_out, int32_t mem) {
    if (!idx || (!vecs && n)) return fail(VMN_ERR_ARG, "NULL argument");
    if (!idx->built) return fail(VMN_ERR_STATE, "index has no graph");
    if (n < 0) return fail(VMN_ERR_ARG, "negative count");
    if (n == 0) return VMN_OK;
    if (idx->n + idx->n_pend + n > idx->cap)
        return fail(VMN_ERR_CAPACITY, "insert of %lld exceeds capacity %lld (size %lld, staged %lld)",
                    (long long)n, (long long)idx->cap, (long long)idx->n, (long long)idx->n_pend);
    int rc = set_device(idx);
    if (rc) return rc;
    const int DP = idx->DP;
    const float* staged = nullptr;
    rc = stage_rows(idx, vecs, n, mem, idx->staging, &staged);
    if (rc) return rc;
    const size_t row = (size_t)DP * sizeof(float);
    rc = grow_keep(idx->pend, (size_t)(idx->n_pend + n) * row, (size_t)idx->n_pend * row,
                   idx->stream);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(static_cast<char*>(idx->pend.p) + (size_t)idx->n_pend * row, staged,
                          (size_t)n * row, hipMemcpyDeviceToDevice, idx->stream));
    std::vector<int32_t> ids(n);
    for (int64_t i = 0; i < n; ++i) ids[i] = (int32_t)(idx->n + idx->n_pend + i);
    idx->n_pend += n;
    if (ids_out && mem == VMN_MEM_HOST) std::memcpy(ids_out, ids.data(), n * sizeof(int32_t));
    if (ids_out && mem == VMN_MEM_DEVICE)
        HIPCHK(hipMemcpyAsync(ids_out, ids.data(), n * sizeof(int32_t), hipMemcpyHostToDevice,
                              idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    return VMN_OK;
}

int64_t vmn_staged_count(const vmn_index* idx) { return idx ? idx->n_pend : -1; }

int vmn_delete_ids(vmn_index* idx, const int32_t* ids, int64_t n, int32_t mem) {
    if (!idx || (!ids && n)) return fail(VMN_ERR_ARG, "NULL argument");
    if (!idx->built) return fail(VMN_ERR_STATE, "index has no graph");
    if (n <= 0) return n == 0 ? VMN_OK : fail(VMN_ERR_ARG, "negative count");
    int rc = set_device(idx);
    if (rc) return rc;
    const int32_t* d = ids;
    if (mem == VMN_MEM_HOST) {
        HIPCHK(idx->staging.reserve((size_t)n * sizeof(int32_t)));
        HIPCHK(hipMemcpyAsync(idx->staging.p, ids, n * sizeof(int32_t), hipMemcpyHostToDevice,
                              idx->stream));
        d = idx->staging.as<int32_t>();
    }
    return set_tombstones(idx, d, n);
}

int vmn_delete_vectors(vmn_index* idx, const float* vecs, int64_t n, int32_t* ids_out, int32_t mem) {
    if (!idx || (!vecs && n)) return fail(VMN_ERR_ARG, "NULL argument");
    if (!idx->built) return fail(VMN_ERR_STATE, "index has no graph");
    if (n <= 0) return n == 0 ? VMN_OK : fail(VMN_ERR_ARG, "negative count");
    if (n > 65535) return fail(VMN_ERR_ARG, "at most 65535 vectors per call");
    int rc = set_device(idx);
    if (rc) return rc;
    const float* q = nullptr;
    rc = stage_rows(idx, vecs, n, mem, idx->qpad, &q);
    if (rc) return rc;
    HIPCHK(idx->find_out.reserve((size_t)n * sizeof(int32_t)));
    HIPCHK(vmn::launch_fill_i32(idx->find_out.as<int32_t>(), n, -1, idx->stream));
    HIPCHK(vmn::launch_find_vectors(idx->vecs.as<float>(), idx->n, idx->p.dim, idx->DP,
                                    idx->del_bits.as<uint32_t>(), q, n,
                                    idx->find_out.as<uint32_t>(), idx->stream));
    std::vector<int32_t> found(n);
    HIPCHK(hipMemcpyAsync(found.data(), idx->find_out.p, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                          idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    // not-found vectors are reported as -1 and not tombstoned (the reference
    // appends the -1 to its delete list, vamana_impl.cuh:49-50)
    std::vector<int32_t> hits;
    for (int32_t v : found)
        if (v >= 0) hits.push_back(v);
    if (!hits.empty()) {
        HIPCHK(idx->staging.reserve(hits.size() * sizeof(int32_t)));
        HIPCHK(hipMemcpyAsync(idx->staging.p, hits.data(), hits.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, idx->stream));
        rc = set_tombstones(idx, idx->staging.as<int32_t>(), (int64_t)hits.size());
        if (rc) return rc;
    }
    if (ids_out && mem == VMN_MEM_HOST) std::memcpy(ids_out, found.data(), n * sizeof(int32_t));
    if (ids_out && mem == VMN_MEM_DEVICE) {
        HIPCHK(hipMemcpyAsync(ids_out, found.data(), n * sizeof(int32_t), hipMemcpyHostToDevice,
                              idx->stream));
        HIPCHK(hipStreamSynchronize(idx->stream));
    }
    return VMN_OK;
}

int vmn_consolidate(vmn_index* idx) {
    if (!idx) return fail(VMN_ERR_ARG, "NULL index");
    if (!idx->built) return fail(VMN_ERR_STATE, "index has no graph");
    // patchGraph: the staged inserts join the graph first
    int rc = apply_staged(idx);
    if (rc) return rc;
    if (idx->n_deleted == 0) return VMN_OK;
    rc = set_device(idx);
    if (rc) return rc;
    ++idx->epoch;
    const int64_t N = idx->n;
    const int R = idx->p.R;
    HIPCHK(idx->rcount.reserve((size_t)N * sizeof(int32_t)));
    HIPCHK(idx->rlist.reserve((size_t)N * sizeof(int64_t)));
    // the node list's buffer holds the tombstone summary until the list is made
    HIPCHK(vmn::launch_cons_flags(idx->adj.as<int32_t>(), N, R, idx->del_bits.as<uint32_t>(),
                                  idx->rcount.as<int32_t>(), idx->rlist.as<uint32_t>(), idx->cus,
                                  idx->stream));
    rc = zero_counter(idx);
    if (rc) return rc;
    HIPCHK(vmn::launch_compact_nonzero(idx->rcount.as<int32_t>(), N, idx->rlist.as<int64_t>(),
                                       idx->counter.as<unsigned long long>(), idx->stream));
    unsigned long long nlist = 0;
    HIPCHK(hipMemcpyAsync(&nlist, idx->counter.p, sizeof nlist, hipMemcpyDeviceToHost, idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    if (nlist) {
        rc = zero_counter(idx);
        if (rc) return rc;
        vmn::ConsArgs a{};
        a.DP = idx->DP;
        a.vecs = idx->vecs.as<float>();
        a.bf16_rows = idx->exact_bf16 ? idx->bf16_rows.p : nullptr;
        a.bf16_norms = idx->exact_bf16 ? idx->bf16_norms.as<float>() : nullptr;
        a.adj = idx->adj.as<int32_t>();
        a.del_bits = idx->del_bits.as<uint32_t>();
        a.W = R;
        a.R = R;
        a.M = idx->p.M;
        a.alpha = idx->p.alpha[idx->p.passes > 0 ? idx->p.passes - 1 : 0];
        a.list = idx->rlist.as<int64_t>();
        a.nlist = (int64_t)nlist;
        a.counter = idx->counter.as<unsigned long long>();
        if (R >= 256) {
            HIPCHK(vmn::launch_consolidate(a, prune_grid(idx, (int64_t)nlist), idx->stream));
        } else {
            // two tiers: pools of at most 256 first, the listed rest with the full pool
            HIPCHK(idx->prune_ovf.reserve((size_t)nlist * sizeof(int64_t)));
            unsigned long long* cnts = idx->counter.as<unsigned long long>();
            HIPCHK(vmn::launch_fill_u64(cnts, 2, 0ull, idx->stream));
            a.overflow = idx->prune_ovf.as<int64_t>();
            a.overflow_n = cnts + 1;
            HIPCHK(vmn::launch_consolidate(a, prune_grid(idx, (int64_t)nlist), idx->stream));
            unsigned long long novf = 0;
            HIPCHK(hipMemcpyAsync(&novf, cnts + 1, sizeof novf, hipMemcpyDeviceToHost, idx->stream));
            HIPCHK(hipStreamSynchronize(idx->stream));
            if (novf) {
                rc = zero_counter(idx);
                if (rc) return rc;
                a.overflow = nullptr;
                a.overflow_n = nullptr;
                a.items = idx->prune_ovf.as<int64_t>();
                a.nlist = (int64_t)novf;
                HIPCHK(vmn::launch_consolidate(a, prune_grid(idx, (int64_t)novf), idx->stream));
            }
        }
    }
    HIPCHK(vmn::launch_clear_deleted_rows(idx->adj.as<int32_t>(), N, R, idx->del_bits.as<uint32_t>(),
                                          idx->stream));
    // move the entry point off a tombstoned node
    uint32_t word = 0;
    HIPCHK(hipMemcpyAsync(&word, idx->del_bits.as<uint32_t>() + (idx->medoid >> 5), sizeof word,
                          hipMemcpyDeviceToHost, idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    if ((word >> (idx->medoid & 31)) & 1u) {
        rc = compute_medoid(idx);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(idx->stream));
    return VMN_OK;
}

int vmn_get_deleted(const vmn_index* idx, uint8_t* mask_out) {
    if (!idx || !mask_out) return fail(VMN_ERR_ARG, "NULL argument");
    int rc = set_device(idx);
    if (rc) return rc;
    const int64_t words = (idx->n + 31) / 32;
    std::vector<uint32_t> bits(words);
    if (words) {
        HIPCHK(hipMemcpyAsync(bits.data(), idx->del_bits.p, words * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, idx->stream));
        HIPCHK(hipStreamSynchronize(idx->stream));
    }
    for (int64_t i = 0; i < idx->n; ++i) mask_out[i] = (bits[i >> 5] >> (i & 31)) & 1u;
    return VMN_OK;
}

int vmn_ctx_create(vmn_index* idx, vmn_ctx** out) {
    if (!idx || !out) return fail(VMN_ERR_ARG, "NULL argument");
    *out = nullptr;
    int rc = set_device(idx);
    if (rc) return rc;
    auto* c = new vmn_ctx();
    c->idx = idx;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(VMN_ERR_HIP, "stream: %s", hipGetErrorString(e));
    }
    *out = c;
    return VMN_OK;
}

int vmn_ctx_destroy(vmn_ctx* c) {
    if (!c) return VMN_OK;
    (void)hipSetDevice(c->idx->p.device);
    (void)hipStreamSynchronize(c->stream);
    if (c->exec) (void)hipGraphExecDestroy(c->exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    (void)hipStreamDestroy(c->stream);
    delete c;  // frees the context's device buffers
    return VMN_OK;
}

void* vmn_ctx_stream(const vmn_ctx* c) { return c ? (void*)c->stream : nullptr; }

int vmn_ctx_search(vmn_ctx* c, const float* queries, int64_t nq, int32_t k, int32_t L,
                   int32_t prec, int32_t k_rerank, int32_t* ids_out, float* dists_out) {
    if (!c || !queries || !ids_out) return fail(VMN_ERR_ARG, "NULL argument");
    vmn_index* idx = c->idx;
    if (!idx->built) return fail(VMN_ERR_STATE, "index has no graph");
    if (idx->p.dim != idx->DP) return fail(VMN_ERR_ARG, "context search needs dim % 32 == 0");
    if (nq < 0 || k < 1 || k > L || L < 1 || L > 512)
        return fail(VMN_ERR_ARG, "need 1 <= k <= L <= 512 (k=%d, L=%d)", k, L);
    if (prec == VMN_PREC_BF16 && !(idx->p.low_precision & VMN_STORE_BF16))
        return fail(VMN_ERR_STATE, "index keeps no BF16 rows (low_precision)");
    if (prec == VMN_PREC_FP8 && !(idx->p.low_precision & VMN_STORE_FP8))
        return fail(VMN_ERR_STATE, "index keeps no FP8 rows (low_precision)");
    if (prec < VMN_PREC_FP32 || prec > VMN_PREC_FP8) return fail(VMN_ERR_ARG, "bad precision %d", prec);
    if (k_rerank != 0 && (k_rerank < k || k_rerank > 64 || k_rerank > L))
        return fail(VMN_ERR_ARG, "need k <= k_rerank <= min(64, L) (k_rerank=%d)", k_rerank);
    if (nq == 0) return VMN_OK;
    int rc = set_device(idx);
    if (rc) return rc;
    return run_search(idx, queries, nq, L, k, ids_out, dists_out, false, idx->p.mode, false, 0, prec,
                      prec == VMN_PREC_FP32 ? 0 : k_rerank, c);
}

int vmn_ctx_capture(vmn_ctx* c, const float* queries, int64_t nq, int32_t k, int32_t L,
                    int32_t prec, int32_t k_rerank, int32_t* ids_out, float* dists_out) {
    if (!c) return fail(VMN_ERR_ARG, "NULL context");
    // one eager call sizes every scratch buffer, so nothing allocates during capture
    int rc = vmn_ctx_search(c, queries, nq, k, L, prec, k_rerank, ids_out, dists_out);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->exec) {
        HIPCHK(hipGraphExecDestroy(c->exec));
        c->exec = nullptr;
    }
    if (c->graph) {
        HIPCHK(hipGraphDestroy(c->graph));
        c->graph = nullptr;
    }
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = vmn_ctx_search(c, queries, nq, k, L, prec, k_rerank, ids_out, dists_out);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(VMN_ERR_HIP, "capture: %s", hipGetErrorString(e));
    c->graph = g;
    HIPCHK(hipGraphInstantiate(&c->exec, g, nullptr, nullptr, 0));
    c->exec_epoch = c->idx->epoch;
    return VMN_OK;
}

int vmn_ctx_replay(vmn_ctx* c) {
    if (!c || !c->exec) return fail(VMN_ERR_STATE, "no captured search on this context");
    if (c->exec_epoch != c->idx->epoch)
        return fail(VMN_ERR_STATE,
                    "the index changed since the capture (build, insert, delete or "
                    "consolidation): capture the search again");
    int rc = set_device(c->idx);
    if (rc) return rc;
    HIPCHK(hipGraphLaunch(c->exec, c->stream));
    return VMN_OK;
}

int vmn_set_tiles(vmn_index* idx, int32_t tiles) {
    if (!idx) return fail(VMN_ERR_ARG, "NULL index");
    if (tiles != 0 && tiles != 2 && tiles != 4) return fail(VMN_ERR_ARG, "tiles must be 0, 2 or 4 (%d)", tiles);
    idx->tiles = tiles == 0 ? 2 : tiles;
    return VMN_OK;
}

int vmn_set_launch(vmn_index* idx, int32_t waves_per_cu, int32_t list_cap) {
    if (!idx) return fail(VMN_ERR_ARG, "NULL index");
    if (waves_per_cu < 1 || waves_per_cu > 64) return fail(VMN_ERR_ARG, "waves_per_cu in [1, 64]");
    if (list_cap < 0 || list_cap > 512) return fail(VMN_ERR_ARG, "list_cap in [0, 512]");
    idx->waves_per_cu = waves_per_cu;
    idx->list_cap = list_cap;
    return VMN_OK;
}

const char* vmn_last_error(void) { return g_err.c_str(); }

}  // extern "C"
