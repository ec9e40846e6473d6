// Kernel argument blocks and launchers (internal to the library).
#pragma once
#include "vmn_device.hpp"

#define VMN_MODE_PARTH 0
#define VMN_MODE_FRESH 1

namespace vmn {

struct SearchArgs {
    int mode;                 // VMN_MODE_PARTH / VMN_MODE_FRESH
    int DP;                   // padded dimension
    const float* vecs;        // [cap x DP]
    const int32_t* adj;       // [cap x W]
    const uint32_t* del_bits; // tombstone bitmap or null
    int W;
    const float* queries;     // [nq x DP]
    int64_t nq;
    int query_is_node;        // queries are graph nodes node_base + qi (skip self in FRESH)
    int64_t node_base;
    int32_t start;
    int L, visited_cap, max_hops;
    int out_n;                // entries of the final list written per query
    int32_t* out_ids;         // [nq x out_n] or null
    float* out_d;             // [nq x out_n] or null
    int32_t* vis_ids;         // FRESH: expanded nodes [nq x P] or null
    float* vis_d;
    int32_t* vis_cnt;         // [nq]
    int P;
    int prec;                 // VMN_PREC_FP32 / BF16 / FP8 distances during traversal
    const void* low_rows;     // [cap x DP] BF16 or FP8 rows
    const float* low_norms;   // [cap] ||x_hat||^2 of the low-precision rows
    const float* adj_norms;   // FP8: [cap x W] ||x_hat||^2 of adj[i] beside entry i (0 past -1)
    const uint32_t* adj_hash; // FRESH: [cap x W] bf_hash_pair(adj[i]) beside entry i (0 past -1)
    const float* fp8_scales;  // [DP] per-dimension FP8 scales
    const float* fp8_centre;  // [DP] per-dimension centre the FP8 rows are coded about
    int krr;                  // > 0: exact FP32 re-rank of the first krr list entries (<= 64)
    int lcap;                 // minimum list capacity template (0: smallest >= L)
    int table_norms;          // BF16, FRESH: row norms from low_norms (latency-bound batches)
    int tiles;                // BF16 / FP8: MFMA tiles of 16 candidates gathered per round (2 or 4)
    int32_t* stat_hops;       // [nq] expansions
    int32_t* stat_ndist;      // [nq] distance evaluations
    unsigned long long* counter;  // work queue head(s), zero before launch
    // Query groups: with ngroups > 0 the batch's queries sit in nlists lists
    // ordered by group (a group's lists consecutive), list L at
    // order[L * gstride ...] with counter[32 L + 1] entries and its queue head
    // at counter[32 L]. Workgroup b walks the lists cyclically from the first
    // of group b % ngroups.
    const int32_t* order;
    int64_t gstride;
    int ngroups;
    int nlists;
    unsigned long long* acc;      // build work counters (VMN_ACC_*) or null
};

// Work counters the build kernels add to (u64 each), from which the build's
// algorithmic bytes are summed.
#define VMN_ACC_SEARCHES 0    // searches run
#define VMN_ACC_EXPANSIONS 1  // nodes expanded (adjacency rows read)
#define VMN_ACC_DISTS 2       // distance evaluations (candidate rows read)
#define VMN_ACC_ITEMS 3       // nodes pruned
#define VMN_ACC_ROWS 4        // rows the prunes read for distances
#define VMN_ACC_PAIRS 5       // reverse pairs emitted by out-prunes
#define VMN_ACC_LISTS 6       // candidate-list entries the prunes read (visited / search lists)
#define VMN_ACC_REVS 7        // reverse candidates the reverse prunes read
#define VMN_ACC_N 8
// The counters are kept in VMN_ACC_SLOTS copies of VMN_ACC_N, one per
// 64-byte line, each workgroup adding to the copy of its index modulo
// VMN_ACC_SLOTS (so the atomics do not queue on one line); the host sums them.
#define VMN_ACC_SLOTS 64
__device__ __forceinline__ unsigned long long* acc_slot(unsigned long long* acc) {
    return acc + (size_t)(blockIdx.x % VMN_ACC_SLOTS) * VMN_ACC_N;
}

hipError_t launch_search(const SearchArgs& a, int grid, hipStream_t s);
// Whether search kernels with 4 gather tiles exist for padded width DP, beam
// capacity need (max of L and the list-capacity minimum) and mode.
bool search_tiles4(int DP, int need, int mode);
// out[i] = norms[adj[i]] (0 where adj[i] is -1) for the n x W entries.
hipError_t launch_adj_norms(const int32_t* adj, int64_t entries, const float* norms, float* out,
                            hipStream_t s);
// out[i] = bf_hash_pair(adj[i]) (0 where adj[i] is -1) for the n x W entries.
hipError_t launch_adj_hash(const int32_t* adj, int64_t entries, uint32_t* out, hipStream_t s);
// Clears *flag unless every row of [from, to) makes BF16 traversal exact.
hipError_t launch_bf16_exact(const float* vecs, int64_t from, int64_t to, int DP, unsigned int* flag,
                             hipStream_t s);
hipError_t launch_to_bf16(const float* vecs, int64_t n, int DP, void* rows, float* norms,
                          hipStream_t s);
// Per-dimension centre (mean, FP64 in fixed chunks) and scales (absmax of the
// centred values / 448) of rows [0, n); temp holds fp8_scales_temp_bytes(n, DP).
size_t fp8_scales_temp_bytes(int64_t n, int DP);
hipError_t launch_fp8_scales(const float* vecs, int64_t n, int D, int DP, unsigned int* absmax_bits,
                             double* temp, float* centre, float* scales, hipStream_t s);
hipError_t launch_to_fp8(const float* vecs, int64_t n, int DP, const float* scales,
                         const float* centre, void* rows, float* norms, hipStream_t s);
hipError_t launch_fp8_dequant(const void* rows, int64_t n, int DP, const float* scales,
                              const float* centre, float* out, hipStream_t s);

struct PruneArgs {
    int mode;
    int DP;
    const float* vecs;
    const void* bf16_rows;    // BF16 copies of vecs when every row is exact in BF16, else null
    const float* bf16_norms;
    int32_t* adj;             // rows rewritten in place
    int32_t* adj_out;         // FRESH out-prune: rows written here instead ([n x W], by item)
    const uint32_t* del_bits; // FRESH out-prune: tombstoned ids are not candidates (or null)
    int W, R;
    float alpha;
    int64_t n;                // nodes handled: node = node_base + i, i < n
    int64_t node_base;
    const int64_t* node_list; // optional explicit node list (overrides node_base)
    // PARTH: search lists [n x Lc]
    const int32_t* cand_ids;
    const float* cand_d;
    int Lc;
    // FRESH out-prune: visited sets [n x P]
    const int32_t* vis_ids;
    const float* vis_d;
    const int32_t* vis_cnt;
    int P;
    // FRESH out-prune: emitted reverse pairs [n x R] (dst = -1 when unused)
    int32_t* pair_dst;
    float* pair_d;
    // FRESH reverse prune: CSR of reverse candidates per node
    const int64_t* rev_off;   // [N+1], indexed by node
    const int64_t* list_off;  // or [n+1], indexed by node_list position (when not null)
    const int32_t* rev_src;
    const float* rev_d;
    int M;                    // cap on merged candidates (reference 500)
    unsigned long long* counter;
    // FRESH prunes, two tiers (the host uses them for the out-prune): with
    // `overflow` set the launch uses a 256-entry candidate list (twice the
    // resident waves of the full one) and lists the items whose candidates
    // could exceed it in overflow[0..*overflow_n); a second launch with
    // `items` = that list and a.n = its length handles them with the full
    // list. Each item's prune touches only its own row, so the split does not
    // change the graph.
    int64_t* overflow;
    unsigned long long* overflow_n;
    const int64_t* items;     // optional: work item k is item items[k]
    unsigned long long* acc;  // build work counters (VMN_ACC_*) or null
};

hipError_t launch_parth_prune(const PruneArgs& a, int grid, hipStream_t s);
hipError_t launch_fresh_out_prune(const PruneArgs& a, int grid, hipStream_t s);
hipError_t launch_fresh_rev_prune(const PruneArgs& a, int grid, hipStream_t s);

hipError_t launch_random_init(int32_t* adj, int64_t n, int R, int W, uint64_t seed, hipStream_t s);
hipError_t launch_medoid(const float* vecs, int64_t n, int D, int DP, const uint32_t* del_bits,
                         float* centroid, unsigned long long* best, hipStream_t s);
hipError_t launch_pad_rows(const float* src, int64_t n, int D, float* dst, int DP, hipStream_t s);
hipError_t launch_rows_by_id(const int32_t* ids, int64_t m, const float* vecs, int DP, int D,
                             const float* pend, int64_t id_base, float* dst, hipStream_t s);
hipError_t launch_merge_staged(const int32_t* gi, const float* gd, const int32_t* pi,
                               const double* pd, int64_t nq, int k, int64_t id_base, int32_t* ri,
                               float* rd, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s);
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
// Sub-lists per query group in the bounded layout (separate counters keep
// the atomics of the lists' appends and takes apart).
#define VMN_GROUP_SUBLISTS 8
// Lists of a grouped batch: one per centroid (by_centroid), or
// groups x VMN_GROUP_SUBLISTS bounded ones; stride entries each.
struct GroupLayout {
    bool by_centroid;
    int nlists;
    int64_t stride;
};
GroupLayout list_layout(int64_t nrows, int C, int G, int S);
// Nearest of C BF16 centroids per row: its index to keys_out, or the row
// appended to its list of list_layout(nrows, C, G, S) (entries at
// order[L * stride ...], count at heads[32 L + 1]).
hipError_t launch_assign(const float* rows, int64_t nrows, int DP, const void* cent,
                         const float* cnorm, int C, int32_t* keys_out, int32_t* order,
                         unsigned long long* heads, int G, int S, hipStream_t s);
hipError_t launch_reverse_count(const int32_t* pair_dst, int64_t npairs, int32_t* counts,
                                hipStream_t s);
hipError_t launch_exclusive_scan(const int32_t* counts, int64_t n, int64_t* off, int64_t* tmp,
                                 hipStream_t s);
hipError_t launch_reverse_scatter(const int32_t* pair_dst, const float* pair_d, int64_t npairs,
                                  int R, int64_t node_base, const int64_t* off, int32_t* fill,
                                  int32_t* rev_src, float* rev_d, hipStream_t s);
// Reverse pairs grouped by destination with a stable radix sort (work in the
// number of pairs, for batches much smaller than the node count): writes
// rev_src / rev_d in destination order, node_list[k] / list_off[k] per
// destination and *nseg (device). temp is sized by reverse_group_temp_bytes.
size_t reverse_group_temp_bytes(int64_t npairs);
hipError_t launch_reverse_group(const int32_t* pair_dst, const float* pair_d, int64_t npairs, int R,
                                int64_t node_base, void* temp, size_t temp_bytes,
                                int32_t* rev_src, float* rev_d, int64_t* node_list,
                                int64_t* list_off, unsigned long long* nseg, hipStream_t s);
// Reverse-prune work order with the items of more than `heavy` candidates
// first (ctr: two zeroed counters).
hipError_t launch_heavy_first(const int64_t* node_list, const int64_t* list_off,
                              const int64_t* rev_off, int64_t n, int W, int heavy, int64_t* order,
                              unsigned long long* ctr, hipStream_t s);
hipError_t launch_compact_nonzero(const int32_t* counts, int64_t n, int64_t* list,
                                  unsigned long long* nlist, hipStream_t s);

// FP64 exact k-NN: partial top-k per (query, base split), then a merge.
// f32: the parth engine's FP32 ground truth instead (sequential FP32 fused
// multiply-add distances, first k by (distance, id)).
hipError_t launch_gt(const float* base, int64_t n, int DP, const uint32_t* del_bits,
                     const float* queries, int64_t nq, int k, int splits, double* part_d,
                     int32_t* part_i, int32_t* out_ids, double* out_d, hipStream_t s,
                     bool f32 = false);
// The FP64 ground truth of launch_gt for DP <= 128 and k <= 64 through an FP32
// MFMA filter with a per-query error bound (same results): S row splits
// (gt_fast_splits), scratch of gt_fast_scratch_bytes, misc = 2 device words.
// Queries the bound cannot settle are listed in *fb_list_out (count in
// misc[1] once the stream reaches it) for launch_gt.
int gt_fast_splits(int64_t n, int64_t nq, int cus);
size_t gt_fast_scratch_bytes(int64_t n, int64_t nq, int S);
hipError_t launch_gt_fast(const float* base, int64_t n, int DP, const uint32_t* del_bits,
                          const float* queries, int64_t nq, int k, int S, void* scratch,
                          int32_t* out_ids, double* out_d, unsigned int* misc,
                          int32_t** fb_list_out, hipStream_t s);
// dst[i] = src row rows[i] (m rows of DP floats); out row rows[i] = result row i.
hipError_t launch_gt_gather(const float* src, const int32_t* rows, int64_t m, int DP, float* dst,
                            hipStream_t s);
hipError_t launch_gt_scatter(const int32_t* rows, int64_t m, int k, const int32_t* ids, const double* d,
                             int32_t* out_ids, double* out_d, hipStream_t s);

}  // namespace vmn
