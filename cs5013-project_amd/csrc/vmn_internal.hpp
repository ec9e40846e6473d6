// Internal view of the index handle shared by the runtime sources (not part
// of the C ABI): device buffers, the vmn_index / vmn_ctx structs and the
// runtime helpers the C entry points are built from.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vamana_mi355x.h"
#include "vmn_kernels.hpp"

namespace vmn_rt {

// Records the message for vmn_last_error() and returns `code`.
int fail(int code, const char* fmt, ...);

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(VMN_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                \
    } while (0)

// Device allocation that grows on demand and is kept across calls; freed
// with its owner (the owner's device must be current then).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    hipError_t reserve(size_t want) {
        if (want <= bytes) return hipSuccess;
        const size_t old = bytes;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            bytes = 0;
        }
        // grow by at least half the old size: buffers sized by batch or node
        // count are not freed and reallocated (a device-wide sync) on every
        // slightly larger call
        size_t b = std::max<size_t>(std::max<size_t>(want, 256), old + old / 2);
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Row stride in floats: the dimension rounded up to one of the instantiated
// kernel widths {32, 64, 96, 128, 160, 192, 256}.
inline int pad_dim(int d) {
    const int p = (d + 31) / 32 * 32;
    return p > 192 ? 256 : p;
}

}  // namespace vmn_rt

struct vmn_index {
    vmn_params p{};
    int DP = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    int64_t n = 0, cap = 0;
    int64_t medoid = -1;
    bool built = false;
    vmn_rt::DevBuf vecs, adj, del_bits;
    // scratch
    vmn_rt::DevBuf qpad, out_ids, out_d, stat_hops, stat_ndist, counter;
    vmn_rt::DevBuf cand_ids, cand_d, vis_ids, vis_d, vis_cnt;
    vmn_rt::DevBuf pair_dst, pair_d, rcount, rfill, roff, rscan_tmp, rsrc, rd, rlist, gtemp;
    vmn_rt::DevBuf prune_ovf;  // items left to the full-capacity prune launch
    vmn_rt::DevBuf adj_shadow;  // out-pruned rows of a chunked FRESH pass
    vmn_rt::DevBuf centroid, medoid_key;
    int64_t last_nq = 0;
    // query groups of large batches (one per XCD; 0 or 1: one queue), the
    // k-means centroids they are formed by (ncent of them, computed at size
    // cent_n) and the index stream's group lists
    int query_groups = 8;
    int ncent = 0;
    int64_t cent_n = 0;
    vmn_rt::DevBuf cent, cent_norm, grp_sample, grp_keys, grp_order;
    int64_t n_deleted = 0;  // tombstones set since the build (upper bound)
    vmn_rt::DevBuf staging, find_out, bad;
    vmn_rt::DevBuf bf16_rows, bf16_norms, fp8_rows, fp8_norms, fp8_scales, fp8_centre, fp8_absmax,
        fp8_temp;
    bool fp8_scaled = false;
    // every row so far makes BF16 traversal exact: the build and insert
    // searches then read BF16 rows (half the bytes, bit-identical result)
    bool exact_bf16 = false;
    vmn_rt::DevBuf exact_flag;
    // parth engine updates: free list of deleted ids and its counter (signed:
    // the engine's counter can run below zero), reset by alloc_graph
    std::vector<int32_t> freelist;
    int64_t free_count = 0;
    vmn_rt::DevBuf eng_ids, eng_src, eng_q, eng_best, eng_marks, eng_found, eng_gt_i, eng_gt_d,
        eng_part_d, eng_part_i;
    // staged inserts (insertPoints before patchGraph): [n_pend x DP] rows that
    // take ids n .. n + n_pend - 1 when vmn_consolidate adds them
    vmn_rt::DevBuf pend, pend_part_d, pend_part_i, pend_i, pend_d, pend_gi, pend_gd;
    int64_t n_pend = 0;
    int waves_per_cu = 32;  // search grid = CUs x waves_per_cu (persistent work queue)
    int list_cap = 0;       // minimum beam-list capacity of the search kernel
    int tiles = 2;          // low-precision search: MFMA tiles gathered per round (2 or 4)
    // queries of the last exact k-NN that its MFMA filter left to the direct kernel
    int64_t gt_fallback = 0;
    // work counters of the last build (VMN_ACC_*), counted while build_count
    vmn_rt::DevBuf build_acc;
    bool build_count = false;
    vmn_build_stats build_stats{};
    // FP8 searches: row norms beside the adjacency entries ([cap x R], the
    // norm of adj[i] at i), valid for index epoch adj_fp8n_epoch
    vmn_rt::DevBuf adj_fp8n;
    uint64_t adj_fp8n_epoch = ~0ull;
    // FRESH searches: both bloom cells of every adjacency entry ([cap x R],
    // bf_hash_pair(adj[i]) at i), valid for index epoch adj_hash_epoch
    vmn_rt::DevBuf adj_hash;
    uint64_t adj_hash_epoch = ~0ull;
    // per-query (distance computations, rounds) of a BANG search into host memory
    vmn_rt::DevBuf bang_stats;
    // bumped by every change of the graph, tombstones or rows: a search
    // captured as a HIP graph holds the entry point, the tombstone pointer
    // and the buffers of its capture, so it may only be replayed while this
    // is unchanged
    uint64_t epoch = 0;
};

// A search context: its own stream and per-launch scratch, so several search
// batches can run on the device at once (the index is read-only meanwhile).
struct vmn_ctx {
    vmn_index* idx = nullptr;
    hipStream_t stream = nullptr;
    vmn_rt::DevBuf counter, stat_hops, stat_ndist, grp_order;
    int64_t last_nq = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t exec_epoch = 0;  // index epoch the captured search belongs to
};

namespace vmn_rt {

int set_device(const vmn_index* idx);
// Query groups (one per XCD) for a batch of nq queries: wanted, the centroids
// they come from (computed on the index stream, once), the batch's list
// layout, and its lists (written on `stream`; heads = 32 x nlists zeroed
// counters).
bool groups_wanted(const vmn_index* idx, int64_t nq);
int ensure_centroids(vmn_index* idx, bool* computed);
vmn::GroupLayout group_layout(const vmn_index* idx, int64_t nq);
int group_queries(vmn_index* idx, const float* q, int64_t nq, DevBuf& order,
                  unsigned long long* heads, hipStream_t stream);
// Rows of dimension D (host or device) placed as [n x DP] rows on the device.
int stage_rows(vmn_index* idx, const float* src, int64_t n, int32_t mem, DevBuf& scratch,
               const float** out);
int compute_medoid(vmn_index* idx);
// Allocates rows, adjacency and tombstones for max(capacity, n) nodes.
int alloc_graph(vmn_index* idx, int64_t n);
// (Re)computes the BF16 / FP8 copies of rows [from, to).
int refresh_lowp(vmn_index* idx, int64_t from, int64_t to);
// Keeps idx->exact_bf16 current after rows [from, to) were written (from == 0
// starts over).
int update_exact(vmn_index* idx, int64_t from, int64_t to);
// Clears idx->exact_bf16 unless rows [from, to) qualify too (never sets it).
int check_exact(vmn_index* idx, int64_t from, int64_t to);
// Traversal precision of the build, insert and engine searches: BF16 rows
// while every row makes it exact, else FP32.
int build_prec(const vmn_index* idx);
// Traversal precision for query rows that are not index rows (engine
// searches): BF16 when the index rows and these rows all make it exact.
int query_prec(vmn_index* idx, const float* q, int64_t nq, int* prec);
// Host copy of an adjacency given by a caller ([n x row_width], -1 padded):
// each row keeps its entries up to the first -1, in order, without repeated
// ids (and without i itself when drop_self), -1 padded to width R.
int normalize_rows(const int32_t* src, int64_t n, int row_width, int R, bool drop_self,
                   int32_t mem, std::vector<int32_t>& rows);
int check_params(const vmn_params* p);
// Adds the staged rows to the graph as one insert batch.
int apply_staged(vmn_index* idx);
// Search of nq padded query rows into device ids / dists (dists may be NULL),
// with the staged rows merged in.
int search_into(vmn_index* idx, const float* q, int64_t nq, int k, int L, int prec, int krr,
                int32_t* ri, float* rd);
// Search kernel arguments for `nq` padded query rows [nq x DP] from
// idx->medoid, without the work counter, statistics and visited-set buffers
// (left null), and the persistent grid for nq queries.
void search_args(const vmn_index* idx, vmn::SearchArgs& a, const float* q, int64_t nq, int L,
                 int out_n, int32_t* out_ids, float* out_d, bool queries_are_nodes, int mode,
                 int64_t node_base, int prec, int krr);
int search_grid(const vmn_index* idx, int64_t nq);
// Makes the buffers a search of precision `prec` reads current (FP8: the
// adjacency-aligned row norms, rebuilt on the index stream when the index
// changed; sync waits for it, for searches on other streams).
int prepare_search(vmn_index* idx, int prec, int mode, bool sync);
// Search kernel over `nq` padded query rows [nq x DP] from idx->medoid.
int run_search(vmn_index* idx, const float* q, int64_t nq, int L, int out_n, int32_t* out_ids,
               float* out_d, bool queries_are_nodes, int mode, bool want_vis,
               int64_t node_base = 0, int prec = VMN_PREC_FP32, int krr = 0,
               vmn_ctx* ctx = nullptr);

}  // namespace vmn_rt
