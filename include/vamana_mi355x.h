/*
 * vamana_mi355x.h -- C ABI of the MI355X-native Vamana / FreshVamana index.
 *
 * Drop-in boundary for the reference's index API (VishankSingh/CS5013-Project):
 * the C++ class `Vamana<T>` of vamana/include/vamana.cuh:36-146 (G2, mirrored
 * at vamana.cu:1245-1527) and the deterministic build/search engine of
 * parth/src/build_graph.cu and parth/src/engine.cu. Plain pointers and sizes,
 * no C++ or torch types. Every call returns VMN_OK (0) or a negative status;
 * vmn_last_error() describes the last failure on the calling thread. Nothing
 * exits the process (the reference calls exit() through gpuErrchk,
 * vamana/include/utils.cuh:68-78).
 *
 * Memory: arguments tagged `mem` are host pointers (VMN_MEM_HOST; the call is
 * synchronous) or device pointers on the index's device (VMN_MEM_DEVICE; the
 * call is asynchronous on the index's stream, vmn_stream(), and the caller
 * synchronises before reading outputs). Outputs are caller-owned (the
 * reference returns callee-allocated device buffers the caller must free,
 * vamana/include/vamana.cuh:96,127).
 *
 * Threading: all state lives in the handle (the reference keeps process-global
 * state, vamana/include/globals.cuh:4-14); calls on one handle must be
 * serialised by the caller; separate handles are independent.
 */
#ifndef VAMANA_MI355X_H
#define VAMANA_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VMN_OK 0
#define VMN_ERR_ARG (-1)      /* invalid argument or unsupported shape */
#define VMN_ERR_HIP (-2)      /* HIP runtime or kernel launch failure */
#define VMN_ERR_STATE (-3)    /* call not valid in the index's state (e.g. search before build) */
#define VMN_ERR_CAPACITY (-4) /* capacity exceeded */
#define VMN_ERR_IO (-5)       /* file open / short read / write failure */

/* Traversal and prune semantics. */
#define VMN_MODE_PARTH 0 /* parth/src/build_graph.cu + engine.cu (deterministic; bit-exact pinned) */
#define VMN_MODE_FRESH 1 /* vamana/include/{search,kernels}.cuh (G2: bloom filter, reverse edges) */

#define VMN_MEM_HOST 0
#define VMN_MEM_DEVICE 1

/* Distance precision during traversal (vmn_search_ex). */
#define VMN_PREC_FP32 0 /* FP32 rows; PARTH mode sums each distance as one fused multiply-add
                          chain in index order (parth/src/device_funcs.cu:17-49), FRESH mode
                          in a lane tree (the reference's order there is atomic arrival,
                          vamana/include/kernels.cuh:40-52) */
#define VMN_PREC_BF16 1 /* BF16 rows on MFMA, FP32 accumulation */
#define VMN_PREC_FP8 2  /* FP8 E4M3 rows, centred on the per-dimension mean and scaled by the
                          centred absmax, on MFMA; their squared norms are kept beside the
                          adjacency ([capacity x R] floats, read with each adjacency row and
                          laid out again on the index stream after an update) */

/* vmn_params.low_precision bits: which low-precision copies the index keeps. */
#define VMN_STORE_BF16 1
#define VMN_STORE_FP8 2

typedef struct vmn_index vmn_index;
typedef struct vmn_ctx vmn_ctx;       /* search context: own stream + scratch */
typedef struct vmn_server vmn_server; /* continuous-batching request server */

typedef struct vmn_params {
    int32_t dim;         /* vector dimension D (any D <= 256) */
    int32_t R;           /* max out-degree = adjacency row width (<= 64) */
    int32_t L_build;     /* beam width of the build searches */
    int32_t passes;      /* build passes (parth: 2, build_graph.cu:345; G2: 1) */
    float alpha[4];      /* alpha of each pass (reference: one alpha, 1.2 parth / 1.5 G2) */
    uint64_t seed;       /* random initial graph seed (build_graph.cu:317 uses 1234) */
    int32_t mode;        /* VMN_MODE_PARTH / VMN_MODE_FRESH */
    int32_t visited_cap; /* PARTH: expansions per search (common.h:16 VISITED_LIST_SIZE = 128) */
    int32_t max_hops;    /* PARTH: iterations per search (engine.cu:50, 200); FRESH: safety bound */
    int32_t P;           /* FRESH: visited-set cap (constants.cuh:13, 600) */
    int32_t M;           /* FRESH: reverse candidate cap (constants.cuh:20, 500) */
    int64_t capacity;    /* node capacity for inserts (0: the build size) */
    int32_t device;      /* HIP device ordinal */
    int32_t low_precision; /* VMN_STORE_* bits: BF16 / FP8 row copies kept beside FP32 */
    int64_t build_chunk; /* FRESH passes: nodes searched and out-pruned per chunk (their rows go
                            to a shadow adjacency until the pass's last chunk, so every search
                            still sees the graph from before the pass); 0: as many as 24 GiB
                            of visited sets hold */
} vmn_params;

/* Reference defaults for `mode` (parth: R=32, L=64, alpha 1.2 x 2 passes). */
void vmn_default_params(vmn_params* p, int32_t mode);

/* Create an empty index on p->device. Replaces the constructor half of
 * `Vamana<T>::Vamana(std::unique_ptr<GraphT>)` (vamana_impl.cuh:10-20). */
int vmn_create(const vmn_params* p, vmn_index** out);
int vmn_destroy(vmn_index* idx);

/* Build the graph over n vectors [n x dim]. Replaces `runVamana`
 * (vamana_impl.cuh:289-360) and parth `build_graph` (build_graph.cu:282-378):
 * random initial graph, medoid, then `passes` x (search all, prune all). */
int vmn_build(vmn_index* idx, const float* vecs, int64_t n, int32_t mem);

/* vmn_build starting from a given initial graph instead of the seeded random
 * one: `init_adj` [n x row_width] (same `mem` as vecs), each row read up to its
 * first -1, repeated ids and the node itself dropped, at most R kept. The
 * reference's constructor builds from the random-graph file it is handed
 * (`Vamana<T>(std::unique_ptr<GraphT>)`, vamana_impl.cuh:10-20, with
 * initGraph, utils.cuh:113-148). init_adj == NULL is vmn_build. */
int vmn_build_from(vmn_index* idx, const float* vecs, int64_t n, const int32_t* init_adj,
                   int32_t row_width, int32_t mem);

/* Install an existing graph ([n x dim] vectors, [n x row_width] adjacency,
 * -1 padded, row_width <= R) with entry point `medoid` (-1: compute it).
 * Replaces `initGraph` (utils.cuh:113-148) / parth load_data (engine.cu:338-357). */
int vmn_load_graph(vmn_index* idx, const float* vecs, int64_t n, const int32_t* adj,
                   int32_t row_width, int64_t medoid, int32_t mem);

/* Index file: parameters, vectors, adjacency, tombstones, entry point and the
 * FP8 scales and centre, so the loaded index searches and updates exactly as the saved
 * one. vmn_load creates the index on `device`. Replaces `saveGraph`
 * (vamana.cu:1346-1362), which writes only vectors and adjacency. */
int vmn_save(const vmn_index* idx, const char* path);
int vmn_load(const char* path, int32_t device, vmn_index** out);

/* The reference's FreshVamana graph file: size() records of { float vec[dim];
 * uint32 degree; uint32 nbr[R] } (772 bytes at dim 128, R 64), no header
 * (constants.cuh:17-18; written by saveGraph, vamana.cu:1346-1362; read by
 * initGraph, utils.cuh:113-148). Saving writes 0xffffffff past each degree.
 * Loading takes dim and R (the record layout) and the index parameters from
 * *p, the node count from the file size; build != 0 builds from the file's
 * graph as the initial graph (the reference's constructor, vmn_build_from),
 * build == 0 installs it with entry point `medoid` (-1: compute it; the
 * reference uses the constant 5000, constants.cuh:12). */
int vmn_save_fresh_graph(const vmn_index* idx, const char* path);
int vmn_load_fresh_graph(const char* path, const vmn_params* p, int32_t build, int64_t medoid,
                         vmn_index** out);

/* The parth engine's graph file: int32 n, dim, max_degree; float
 * vectors[n x dim]; int32 adjacency[n x max_degree], -1 padded (written by
 * build_graph.cu:361-369, read by engine.cu:338-357). Saving writes R as
 * max_degree. Loading takes dim and R from the file, the other parameters
 * from *p, and installs the graph with entry point `medoid` (-1: compute it;
 * the engine's searches start at node 0, engine.cu:476-477). */
int vmn_save_parth_graph(const vmn_index* idx, const char* path);
int vmn_load_parth_graph(const char* path, const vmn_params* p, int64_t medoid, vmn_index** out);

/* Vector files. Element type per format: float (FVECS, BIN), int32 (IVECS),
 * uint8 (BVECS).
 *   FVECS / IVECS / BVECS: per vector int32 dim, then dim elements
 *     (the siftsmall .fvecs / .ivecs files; FVecs::load_from_fvecs_file,
 *     vamana.cu:108-153);
 *   BIN: uint32 n, uint32 dim, then n x dim floats (FVecs::save_to_binary /
 *     load_from_binary, vamana.cu:156-195; scripts/vecs_to_binary.py:21-39;
 *     data/sift10k/siftsmall_query.bin).
 * Reading: with out == NULL only *n and *dim are set; otherwise at most
 * `capacity` vectors are copied. A BIN header that claims more vectors than
 * the file holds (the reference's siftsmall_query.bin says 10,000 for its 100)
 * reads the vectors present. */
#define VMN_FILE_FVECS 0
#define VMN_FILE_IVECS 1
#define VMN_FILE_BVECS 2
#define VMN_FILE_BIN 3
int vmn_read_vectors(const char* path, int32_t format, void* out, int64_t capacity, int64_t* n,
                     int32_t* dim);
int vmn_write_vectors(const char* path, int32_t format, const void* vecs, int64_t n, int32_t dim);

/* DiskANN ground-truth ("truthset") file: uint32 n, uint32 k, uint32 ids[n x
 * k], then float dists[n x k] when the file size has room for them (the
 * reference's data/sift10k/sift10k_groundtruth.bin has them,
 * siftsmall_groundtruth.bin does not; read by load_truthset,
 * bang_exact/src/parANN.cuh:371; written by DiskANN's compute_groundtruth,
 * vamana.cu:1644). *has_dists (may be NULL) says which. dists == NULL on
 * writing leaves them out. */
int vmn_read_truthset(const char* path, int32_t* ids, float* dists, int64_t capacity, int64_t* n,
                      int32_t* k, int32_t* has_dists);
int vmn_write_truthset(const char* path, const int32_t* ids, const float* dists, int64_t n,
                       int32_t k);

/* BANG / DiskANN disk index (scripts/bang-preprocess.py:37-77;
 * data/sift10k/sift10k_index_disk.bin + _metadata.bin): n records { float
 * vec[dim]; uint32 degree; uint32 nbr[R] } with the neighbours in ascending id
 * order (0xffffffff past the degree when saving), and a 32-byte metadata file
 * { uint64 medoid; uint64 record bytes; uint32 data type (2 = float); uint32
 * dim; uint32 R; uint32 n }. Loading takes dim, R, n and the entry point from
 * the metadata and the other parameters from *p. */
int vmn_save_bang_index(const vmn_index* idx, const char* index_path, const char* metadata_path);
int vmn_load_bang_index(const char* index_path, const char* metadata_path, const vmn_params* p,
                        vmn_index** out);

/* BANG's exact search schedule over the index's graph: replaces the search of
 * `parANN` (bang_exact/src/parANN.cu:96-978: its host round loop :500-668 and
 * the kernels neighbor_filtering_new :996-1052, compute_neighborDist_par
 * :1118-1158, compute_BestLSets_par_sort_msort_new :1475-1669,
 * compute_NearestNeighbours :1191-1255), with L (compile-time L, parANN.cuh:41),
 * the entry point (MEDOID, parANN.cuh:125; entry < 0 takes vmn_medoid()) and
 * recall_at (k) as arguments. Per query: a 399,887-cell bloom filter on one
 * hash; each round's candidates are the entry point (round 1) and the parent's
 * unseen neighbours; FP32 distances summed as the reference does; a stable
 * sort and BANG's best-L merge; the first unvisited list entry is the next
 * parent; a query ends at a round without candidates or parent, or after
 * round 4L + 19. Writes the first k list entries per query (ids -1 and dists
 * +inf past the list's end) and, in stats_out [nq x 2] (may be NULL), the
 * distance computations and rounds of each query: their sum and maximum are
 * BANG's "total computations" and "iterations". Tombstoned rows are never
 * candidates; staged rows (vmn_insert_staged) are not searched. Needs
 * 1 <= k <= L <= 256. `mem` as for vmn_search. */
int vmn_bang_search(vmn_index* idx, const float* queries, int64_t nq, int32_t k, int32_t L,
                    int64_t entry, int32_t* ids_out, float* dists_out, int32_t* stats_out,
                    int32_t mem);

/* Batched greedy search of nq queries [nq x dim] from the medoid with beam L;
 * writes the k best ids (-1 padded) and squared L2 distances (may be NULL) per
 * query. Replaces `searchPoints` (vamana.cuh:96, vamana_impl.cuh:55-96) and
 * parth `searchKernel` (engine.cu:100-120). Tombstoned nodes are skipped. */
int vmn_search(vmn_index* idx, const float* queries, int64_t nq, int32_t k, int32_t L,
               int32_t* ids_out, float* dists_out, int32_t mem);

/* vmn_search with the traversal distance in `prec` (VMN_PREC_*; BF16/FP8 need
 * the matching VMN_STORE_* copy) and an exact FP32 re-rank of the first
 * `k_rerank` (k <= k_rerank <= 64; 0 = none) list entries before the k
 * results are taken (the reference re-ranks its worklist in
 * merge_and_rerank_kernel, utils.cuh:199-285). */
int vmn_search_ex(vmn_index* idx, const float* queries, int64_t nq, int32_t k, int32_t L,
                  int32_t prec, int32_t k_rerank, int32_t* ids_out, float* dists_out,
                  int32_t mem);

/* vmn_search_ex on device buffers that allocates nothing: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * queries [nq x dim], ids_out [nq x k] and dists_out [nq x k] (may be NULL)
 * are device pointers on the index's device; the call is asynchronous on
 * vmn_stream(). Same results as vmn_search_ex. Its scratch (the work-queue
 * counter and, when dim is not a multiple of 32, the zero-padded queries) is
 * one static 4 MiB device array per GPU, held until the process ends: with
 * padding, nq * padded_dim * 4 must fit in 4 MiB - 256 bytes (8,191 queries
 * at dim 100), else VMN_ERR_CAPACITY. Every call on one GPU shares that
 * array, so calls whose kernels overlap on one GPU (two streams, two indexes
 * on one device) race: serialise them. Refused (VMN_ERR_STATE) while staged
 * inserts are pending. Search statistics (vmn_search_stats) are not kept. */
int vmn_search_noalloc(vmn_index* idx, const float* queries, int64_t nq, int32_t k, int32_t L,
                       int32_t prec, int32_t k_rerank, int32_t* ids_out, float* dists_out);

/* vmn_search returning the k nearest rows themselves, vecs_out [nq x k x dim]
 * (zeros past the results found), like `searchPoints2` (vamana.cuh:127,
 * vamana_impl.cuh:99-180), which returns the worklist's vectors. */
int vmn_search_vectors(vmn_index* idx, const float* queries, int64_t nq, int32_t k, int32_t L,
                       float* vecs_out, int32_t mem);

/* The rows as the low-precision traversal sees them, decoded to float
 * ([size x dim], host): BF16 values, or FP8 values times their scales plus the centre. */
int vmn_get_lowp_rows(const vmn_index* idx, int32_t prec, float* out);

/* Batched insert of n vectors [n x dim]: they become nodes size .. size+n-1
 * (written to ids_out when not NULL). Each is searched from the entry point
 * over the current graph (tombstones skipped), out-pruned from its visited
 * set, and its reverse edges are merged into its out-neighbours, as the
 * reference's patchGraph does for its staged insert list (vamana_impl.cuh:
 * 182-265 with insertPoints, vamana.cuh:42,130). Needs capacity. */
int vmn_insert(vmn_index* idx, const float* vecs, int64_t n, int32_t* ids_out, int32_t mem);

/* Staged insert, the reference's insertPoints (vamana_impl.cuh:22-33, InsertList
 * insert_list.cuh:12-87): the vectors are held aside and are promised ids
 * size() + staged + i (ids_out). vmn_search, vmn_search_ex and
 * vmn_search_vectors merge their exact k nearest (FP64 brute force over the
 * staged rows) into the graph's results by (distance, id), as searchPoints2's
 * merge_and_rerank_kernel (utils.cuh:199-285) does; k <= 128 then. Search
 * contexts and the server see the graph only. vmn_consolidate (patchGraph,
 * vamana_impl.cuh:182-265) adds them to the graph as one batch, and so does
 * vmn_insert before its own rows. The savers refuse while rows are staged. */
int vmn_insert_staged(vmn_index* idx, const float* vecs, int64_t n, int32_t* ids_out,
                      int32_t mem);
int64_t vmn_staged_count(const vmn_index* idx);

/* Tombstone ids: searches no longer expand or return them. Ids outside
 * [0, size) give VMN_ERR_ARG (the valid ones are still applied). Replaces the
 * delete list (delete_list.cuh:9-115) and parth deleteKernel (engine.cu:131-147). */
int vmn_delete_ids(vmn_index* idx, const int32_t* ids, int64_t n, int32_t mem);

/* Delete by exact vector match among live nodes (deletePoints, vamana.cuh:65,
 * findPointsKernel kernels.cuh:905-933): ids_out[i] = the lowest matching id,
 * or -1 when none matches (then nothing is tombstoned for it). n <= 65535. */
int vmn_delete_vectors(vmn_index* idx, const float* vecs, int64_t n, int32_t* ids_out,
                       int32_t mem);

/* patchGraph (vamana_impl.cuh:182-265) plus delete consolidation. First the
 * staged inserts join the graph (one insert batch). Then (FreshDiskANN
 * Algorithm 4 in spirit; the reference consolidates nothing) every live node
 * with a tombstoned out-neighbour keeps its live out-neighbours in order and
 * fills the freed slots from the live out-neighbours of its tombstoned ones,
 * nearest first, skipping a candidate that a kept neighbour r occludes
 * (alpha * d(r, c) <= d(p, c)); rows of tombstoned nodes are emptied; a
 * tombstoned entry point moves to the medoid of the live nodes. Ids stay
 * tombstoned (not reused). */
int vmn_consolidate(vmn_index* idx);

/* Tombstone mask [size] (1 = deleted) into host memory. */
int vmn_get_deleted(const vmn_index* idx, uint8_t* mask_out);

/* Copy the adjacency [size x R] (int32, -1 padded). */
int vmn_get_adjacency(const vmn_index* idx, int32_t* out, int32_t mem);

/* Copy the vectors [size x dim] (float32, as given to build / insert). The
 * reference's graph records hold them (graphT.cuh:56-59). */
int vmn_get_vectors(const vmn_index* idx, float* out, int32_t mem);

/* The parameters the index was created with. */
int vmn_get_params(const vmn_index* idx, vmn_params* out);

int64_t vmn_size(const vmn_index* idx);
int64_t vmn_medoid(const vmn_index* idx);

/* Row precision the build and insert searches traverse: VMN_PREC_BF16 while
 * every row is an integer vector with |v| <= 256 and squared norm <= 2^22
 * (BF16 distances then equal FP32 ones bit for bit, so the graph is the same
 * as an FP32 traversal's at half the row bytes), else VMN_PREC_FP32. The
 * reference's build and insert searches always read its FP32 rows
 * (greedySearch calls, vamana_impl.cuh:235, :325). */
int32_t vmn_build_precision(const vmn_index* idx);

/* Work of the last vmn_build / vmn_build_from, counted by its kernels, and
 * the build's algorithmic bytes from them (the measure of build bandwidth;
 * no counterpart in the reference). row_bytes is the bytes of one row the
 * build read: 2 x padded dim while every row is exact in BF16, else 4 x. The
 * byte sum is: 8 n DP (medoid: centroid and distance passes over the FP32
 * rows) + 4 n R (random initial graph, when used) + the low-precision copies
 * (BF16: n (6 DP + 4); FP8: n (13 DP + 4)) + per search its node's FP32 row
 * (4 DP) + per expansion its adjacency row and visited entry (4 R + 8) + per
 * distance evaluation a row (row_bytes) + per row a prune read for distances
 * row_bytes (+ 4 for its BF16 norm) + 8 per candidate-list entry the prunes
 * read + 8 R per pruned node (its row read and written) + 24 per reverse pair
 * (written, grouped, read) + 8 per reverse candidate a reverse prune read. */
typedef struct vmn_build_stats {
    int64_t searches, expansions, dist_evals; /* the build's searches */
    int64_t prune_items, prune_rows;          /* nodes pruned, rows read for their distances */
    int64_t reverse_pairs;                    /* (p* <- node) pairs the out-prunes emitted */
    int64_t list_entries, reverse_entries;    /* candidate-list / reverse-candidate entries read */
    int32_t row_bytes, passes;
    double algorithmic_bytes;
} vmn_build_stats;
int vmn_get_build_stats(const vmn_index* idx, vmn_build_stats* out);

/* Sum over the queries of the last vmn_search: expansions (hops) and distance
 * evaluations. Synchronises the index's stream. */
int vmn_search_stats(const vmn_index* idx, int64_t* expansions, int64_t* dist_evals);

/* The hipStream_t all of the index's work is queued on. */
void* vmn_stream(const vmn_index* idx);

/* Exact k nearest neighbours (k <= 128) of nq queries over n base vectors in
 * FP64, ids ordered by (distance, id); dists_out may be NULL. Replaces
 * parth `bruteForceKernel` (engine.cu:253-279) and the external DiskANN
 * compute_groundtruth (vamana.cu:1644). Runs on `device`; synchronous. */
int vmn_groundtruth(const float* base, int64_t n, int32_t dim, const float* queries, int64_t nq,
                    int32_t k, int32_t* ids_out, double* dists_out, int32_t mem, int32_t device);
/* Queries of the calling thread's last vmn_groundtruth whose candidates the
 * MFMA filter (FP32 q.x with a per-query error bound, dimensions up to 128,
 * k <= 64, n >= 4096) could not settle and the direct FP64 kernel computed;
 * either way the results are the direct kernel's. No counterpart in the
 * reference. */
int64_t vmn_groundtruth_fallbacks(void);

/* Mean recall@recall_at of nq result rows (found [nq x k_found], -1 padded)
 * against ground truth rows gt [nq x k_gt] (k_gt >= recall_at): the share of
 * the first recall_at ground-truth ids found among the first recall_at
 * results. With gt_dists [nq x k_gt] (may be NULL) the ground-truth set also
 * takes every further id whose distance equals the recall_at-th one (ties),
 * as calculate_recall (bang_exact/src/main.cu:87-137) does; the result is a
 * fraction, not a percentage. Host memory. */
double vmn_recall(const int32_t* found, int32_t k_found, const int32_t* gt, const float* gt_dists,
                  int32_t k_gt, int64_t nq, int32_t recall_at);

/* Search contexts. Each has its own HIP stream and scratch, so searches on
 * different contexts run concurrently on the device. Searches only read the
 * index: the caller must not run an update (build, insert, delete,
 * consolidate) on the index while a context search is in flight.
 * vmn_ctx_search takes device pointers (dim a multiple of 32) and is
 * asynchronous on vmn_ctx_stream(ctx); otherwise as vmn_search_ex. No
 * counterpart in the reference (its search is synchronous on the default
 * stream, search.cuh:344). */
int vmn_ctx_create(vmn_index* idx, vmn_ctx** out);
int vmn_ctx_destroy(vmn_ctx* ctx);
void* vmn_ctx_stream(const vmn_ctx* ctx);
int vmn_ctx_search(vmn_ctx* ctx, const float* queries, int64_t nq, int32_t k, int32_t L,
                   int32_t prec, int32_t k_rerank, int32_t* ids_out, float* dists_out);

/* Continuous-batching server (no counterpart in the reference, whose request
 * loops sort fixed batches into delete/insert/search phases,
 * parth/src/engine.cu:407-568, vamana.cu:1782-1895). Requests are queued in
 * arrival order; consecutive searches are packed into the batch of a free
 * slot (a search context, up to max_batch queries) and launched while other
 * slots' batches are still on the device; an insert or delete runs when it
 * reaches the head of the queue, after the launched searches finish, so each
 * search sees exactly the updates that arrived before it and returns the same
 * result as when run alone. Submission fails with VMN_ERR_CAPACITY once the
 * queued payload exceeds max_queued_bytes (admission by memory). Data is
 * copied at submit; dim must be a multiple of 32. While a server runs, the
 * index must only be used through it. */
#define VMN_REQ_SEARCH 0 /* data: float [n x dim] queries; results [n x k] ids, dists */
#define VMN_REQ_INSERT 1 /* data: float [n x dim] vectors; results: n assigned ids */
#define VMN_REQ_DELETE 2 /* data: int32 [n] ids */

typedef struct vmn_server_params {
    int32_t k, L;              /* results per query, beam width */
    int32_t prec, k_rerank;    /* traversal precision (VMN_PREC_*), FP32 re-rank count */
    int32_t max_batch;         /* queries per search launch (per slot) */
    int32_t n_slots;           /* search launches in flight at once (1..16) */
    int64_t max_queued_bytes;  /* admission limit on queued request payload */
    int64_t consolidate_every; /* consolidate after this many deletes (0: never) */
    int32_t device;
} vmn_server_params;

void vmn_default_server_params(vmn_server_params* p);
int vmn_server_create(vmn_index* idx, const vmn_server_params* p, int32_t dim, vmn_server** out);
/* Seconds since the server was created (the clock t_arrival is measured on). */
double vmn_server_now(const vmn_server* srv);
int vmn_server_submit(vmn_server* srv, int32_t kind, const void* data, int64_t n,
                      double t_arrival, int64_t* ticket);
/* 1: done (results copied, ticket released), 0: pending, < 0: the request's
 * error status (ticket released). */
int vmn_server_poll(vmn_server* srv, int64_t ticket, int32_t* ids_out, float* dists_out,
                    double* t_arrival, double* t_done);
int64_t vmn_server_pending(const vmn_server* srv);

/* Engine counters since the server was created: where the engine's time goes
 * (no counterpart in the reference's request loop, engine.cu:407-568). */
typedef struct vmn_server_stats {
    int64_t batches, queries, max_batch_queries; /* search launches and their sizes */
    int64_t inserts, deletes, consolidations;    /* updates applied */
    double insert_s, delete_s;                   /* engine time applying them (incl. consolidation) */
    double max_update_s;                         /* longest single update */
    double drain_s;                              /* time updates waited for launched searches */
    int64_t failed_batches;                      /* search launches whose requests failed with a
                                                    runtime error (VMN_ERR_HIP) */
} vmn_server_stats;
int vmn_server_get_stats(const vmn_server* srv, vmn_server_stats* out);
/* Zeroes the counters (e.g. after a warm-up), maxima included. */
int vmn_server_reset_stats(vmn_server* srv);

/* Finishes the queued requests, then stops the engine. */
int vmn_server_destroy(vmn_server* srv);

/* Capture one context search (fixed pointers and sizes) as a HIP graph of
 * kernel nodes (the work-queue counter is cleared by a kernel, not a memset
 * node), and launch it again on the context's stream. Replays re-read the
 * query buffer, so new queries written there are searched. The graph holds
 * the index state of its capture (entry point, tombstones, buffers): after
 * any build, load, insert, delete or consolidation of the index,
 * vmn_ctx_replay returns VMN_ERR_STATE until the search is captured again. */
int vmn_ctx_capture(vmn_ctx* ctx, const float* queries, int64_t nq, int32_t k, int32_t L,
                    int32_t prec, int32_t k_rerank, int32_t* ids_out, float* dists_out);
int vmn_ctx_replay(vmn_ctx* ctx);

/* Search launch shape chosen by the autotuner: persistent grid of
 * waves_per_cu x CUs workgroups (one wavefront each) and the minimum
 * beam-list capacity template (0: the smallest of 64/128/256/512 >= L). */
int vmn_set_launch(vmn_index* idx, int32_t waves_per_cu, int32_t list_cap);
/* Tile shape of the BF16 / FP8 searches, also chosen by the autotuner (no
 * counterpart in the reference): MFMA tiles of 16 candidates whose rows are
 * gathered per round, 2 (the default; 0 selects it) or 4 (a hop's new
 * candidates in one round, at 4 wavefronts per SIMD; FRESH mode, dimensions
 * 33 to 128, beam capacity up to 256, else 2 is used). Results do not depend
 * on it. */
int vmn_set_tiles(vmn_index* idx, int32_t tiles);

/* Query groups of large batches (no counterpart in the reference; default 8,
 * one per XCD of an MI355X in its default single-partition mode; 0 or 1: one
 * queue). A batch of at least 8 queries per CU is split by the nearest of 64
 * k-means centroids of the index's rows (computed on the index stream by the
 * first such search, and again once the index has doubled), one group per
 * XCD, and each workgroup searches its own XCD's group first, so the rows one
 * L2 holds serve the queries that XCD runs. The grouping changes only the
 * order in which queries are searched, never a result. */
int vmn_set_query_groups(vmn_index* idx, int32_t groups);
/* The group (0 .. groups-1) each of nq queries ([nq x dim]) would be searched
 * in, one byte per query into gids_out (host or device, per mem). Synchronous. */
int vmn_query_group_ids(vmn_index* idx, const float* queries, int64_t nq, int32_t mem,
                        uint8_t* gids_out);

/* ---- parth engine: workload files, free-list updates, workload replay ----
 *
 * The parth engine (parth/src/engine.cu:304-568, VamanaEngine) replays a
 * workload file of operations in batches: per batch the deletes, then the
 * inserts, then the searches (from node 0, top 10), each search scored
 * against brute force over the live nodes. Its updates differ from the
 * FreshVamana ones: a delete also pushes the id on a free list, and an
 * insert takes an id from that list (else the next new id), links the new
 * node to its nearest node both ways and prunes nothing. Load the graph with
 * vmn_load_parth_graph(path, &p, 0, &idx) (medoid 0 = the engine's start
 * node) and p.capacity = the engine's MAX_N (20000 for dim 128, engine.cu:721).
 * The free list starts empty when the index is created or loaded. */
#define VMN_OP_DELETE 0 /* parth/src/common.h:51-55 OpType */
#define VMN_OP_INSERT 1
#define VMN_OP_SEARCH 2

/* Reads a workload file: records { int32 type; int32 id; float vec[dim] }
 * (common.h:63-67; written by parth/tools/gen_sift_workload.py:80-86, read by
 * engine.cu:359-372) up to the first incomplete record. With types == NULL
 * only *count is set; otherwise at most `capacity` records are copied into
 * types [n], ids [n] and vecs [n x dim]. */
int vmn_read_workload(const char* path, int32_t dim, int32_t* types, int32_t* ids, float* vecs,
                      int64_t capacity, int64_t* count);

/* deleteKernel (engine.cu:131-147): each id in [0, capacity) is marked
 * deleted and pushed on the free list at the list's counter (a push at a
 * negative counter is dropped, as the reference's write lands outside its
 * list); other ids are ignored. Ids in order, host memory. */
int vmn_engine_delete(vmn_index* idx, const int32_t* ids, int64_t n);

/* insertPrepareKernel + insertLinkKernel (engine.cu:161-239) as one batch, in
 * this order: ids assigned (every insert sees the free-list count of the
 * batch start; when it is positive each insert decrements it and takes the
 * slot below the old value if that is positive, otherwise the next new id);
 * vectors written (a reused id is unmarked), rows reset to -1; the nearest
 * node of every new vector searched on the graph before any new link (beam
 * link_L, the engine's BEAM_WIDTH = 128); slot 0 of every new row set to it;
 * then, in input order, the new id written into the first -1 slot of its
 * nearest node's row. ids_out [n] gets the assigned ids. vecs [n x dim] in
 * `mem`; ids_out in host memory (or NULL). */
int vmn_engine_insert(vmn_index* idx, const float* vecs, int64_t n, int32_t link_L,
                      int32_t* ids_out, int32_t mem);

typedef struct vmn_engine_result {
    int64_t searches;      /* search operations run */
    int64_t inserts, deletes;
    int64_t batches;
    double search_ms;      /* sum of the search launches' device time */
    double qps;            /* searches / search time (the reference's `qps`) */
    double recall;         /* mean recall@10 (the reference's `recall_10`) */
    double wall_s;         /* host wall time of the whole replay */
    int64_t expansions;    /* searches' node expansions (hops), summed */
    int64_t dist_evals;    /* searches' distance evaluations, summed */
} vmn_engine_result;

/* VamanaEngine::process_workload (engine.cu:407-568): ops [n_ops] (types,
 * ids, vecs [n_ops x dim], host memory) in batches of batch_size; searches
 * of beam L from the entry point return 10 ids (-1 padded); ground truth is
 * the first 10 live ids below the node count by (FP64 distance, id)
 * (bruteForceKernel, engine.cu:254-279); recall of a search = found ids in
 * its ground truth / 10. found_out and gt_out ([searches x 10], host, in op
 * order) may be NULL. */
int vmn_engine_run(vmn_index* idx, const int32_t* types, const int32_t* ids, const float* vecs,
                   int64_t n_ops, int32_t batch_size, int32_t L, int32_t link_L,
                   vmn_engine_result* result, int32_t* found_out, int32_t* gt_out);

/* Measurement helper, no counterpart in the reference: reads rows ids[0 ..
 * n_reads) of `row_bytes` (128, 256 or 512) bytes each from `table` (device,
 * table_rows rows; every id must be below table_rows) the way the search
 * kernels gather candidate rows -- each row in one load instruction
 * (split = 0), or 64 bytes of 16 rows per instruction (split = 1) -- and
 * writes one word per lane to `sink` (device, sink_words >= 256). Synchronous
 * on the null stream; *ms_out gets the kernel time. With distinct ids the
 * bytes it must fetch are known (n_reads x row_bytes + 4 n_reads), which
 * calibrates the HBM traffic counters for random row gathers. */
int vmn_gather_probe(const void* table, int64_t table_rows, int32_t row_bytes, int32_t split,
                     const int32_t* ids, int64_t n_reads, uint32_t* sink, int64_t sink_words,
                     int32_t device, double* ms_out);

/* Message for the last failing call on this thread ("" if none). */
const char* vmn_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* VAMANA_MI355X_H */
