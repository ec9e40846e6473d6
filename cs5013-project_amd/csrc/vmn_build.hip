// Graph construction kernels: random initial graph, medoid, alpha robust
// prune (PARTH and FRESH forms) and reverse-edge insertion.
#include "vmn_prune.hpp"

namespace vmn {

// ---------------------------------------------------------------------------
// Small utility kernels
// ---------------------------------------------------------------------------
__global__ void fill_u64_kernel(unsigned long long* p, int64_t n, unsigned long long v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
static int fill_grid(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (int)(g < 4096 ? (g < 1 ? 1 : g) : 4096);
}
hipError_t launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s) {
    hipLaunchKernelGGL(fill_u64_kernel, dim3(fill_grid(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3(fill_grid(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// Copies [n x D] rows into [n x DP] rows, zero padding columns D..DP.
__global__ void pad_rows_kernel(const float* src, int64_t n, int D, float* dst, int DP) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = n * DP;
    for (; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / DP;
        const int c = (int)(i - r * DP);
        dst[i] = c < D ? src[r * D + c] : 0.0f;
    }
}
hipError_t launch_pad_rows(const float* src, int64_t n, int D, float* dst, int DP, hipStream_t s) {
    hipLaunchKernelGGL(pad_rows_kernel, dim3(fill_grid(n * DP)), dim3(256), 0, s, src, n, D, dst, DP);
    return hipGetLastError();
}

// dst[i] = row ids[i] cut to D columns, zeros for id -1: ids below id_base
// from vecs, the rest from the staged rows pend (both [.. x DP]).
__global__ void rows_by_id_kernel(const int32_t* ids, int64_t m, const float* vecs, int DP, int D,
                                  const float* pend, int64_t id_base, float* dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = m * D;
    for (; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / D;
        const int c = (int)(i - r * D);
        const int64_t id = ids[r];
        dst[i] = id < 0 ? 0.0f
                        : (id < id_base ? vecs[id * DP + c] : pend[(id - id_base) * DP + c]);
    }
}
hipError_t launch_rows_by_id(const int32_t* ids, int64_t m, const float* vecs, int DP, int D,
                             const float* pend, int64_t id_base, float* dst, hipStream_t s) {
    hipLaunchKernelGGL(rows_by_id_kernel, dim3(fill_grid(m * D)), dim3(256), 0, s, ids, m, vecs, DP,
                       D, pend, id_base, dst);
    return hipGetLastError();
}

// Per query, the first k of two lists sorted by (distance, id): the graph's
// (gi, gd) and the staged rows' exact nearest (pi as staged index, FP64 pd),
// which take ids id_base + index. Equal distances keep the lower id first.
__global__ void merge_staged_kernel(const int32_t* gi, const float* gd, const int32_t* pi,
                                    const double* pd, int64_t nq, int k, int64_t id_base,
                                    int32_t* ri, float* rd) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int64_t o = q * k;
    int a = 0, b = 0;
    for (int j = 0; j < k; ++j) {
        const bool ha = a < k && gi[o + a] >= 0;
        const bool hb = b < k && pi[o + b] >= 0;
        int32_t id = -1;
        float d = 1e30f;
        if (ha && hb) {
            const float db = (float)pd[o + b];
            const int64_t ib = id_base + pi[o + b];
            if (gd[o + a] < db || (gd[o + a] == db && gi[o + a] < ib)) {
                id = gi[o + a];
                d = gd[o + a];
                ++a;
            } else {
                id = (int32_t)ib;
                d = db;
                ++b;
            }
        } else if (ha) {
            id = gi[o + a];
            d = gd[o + a];
            ++a;
        } else if (hb) {
            id = (int32_t)(id_base + pi[o + b]);
            d = (float)pd[o + b];
            ++b;
        }
        ri[o + j] = id;
        if (rd) rd[o + j] = d;
    }
}
hipError_t launch_merge_staged(const int32_t* gi, const float* gd, const int32_t* pi,
                               const double* pd, int64_t nq, int k, int64_t id_base, int32_t* ri,
                               float* rd, hipStream_t s) {
    hipLaunchKernelGGL(merge_staged_kernel, dim3((unsigned)((nq + 63) / 64)), dim3(64), 0, s, gi, gd,
                       pi, pd, nq, k, id_base, ri, rd);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Random initial graph: parth/src/build_graph.cu:28-56. Per node a 64-bit LCG
// (multiplier 6364136223846793005, increment 1442695040888963407) seeded with
// seed + node id draws targets (state >> 32) % n, rejecting self and repeats,
// until R distinct neighbours are found; the rest of the row is -1.
// ---------------------------------------------------------------------------
__global__ void random_init_kernel(int32_t* adj, int64_t n, int R, int W, uint64_t seed) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n) return;
    int32_t* row = adj + tid * W;
    int32_t loc[64];
    uint64_t state = seed + (uint64_t)tid;
    int count = 0;
    // n > R is checked on the host, so R distinct non-self targets exist; the
    // draw count is still bounded so a degenerate generator cannot spin forever.
    for (int64_t draws = 0; count < R && draws < ((int64_t)1 << 24); ++draws) {
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        const int32_t target = (int32_t)((state >> 32) % (uint64_t)n);
        if (target == (int32_t)tid) continue;
        bool exists = false;
        for (int i = 0; i < count; ++i) exists |= loc[i] == target;
        if (!exists) loc[count++] = target;
    }
    for (int i = 0; i < W; ++i) row[i] = i < count ? loc[i] : -1;
}
hipError_t launch_random_init(int32_t* adj, int64_t n, int R, int W, uint64_t seed, hipStream_t s) {
    hipLaunchKernelGGL(random_init_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, adj,
                       n, R, W, seed);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Medoid: parth/src/build_graph.cu:326-343. The reference sums the centroid on
// the host node by node in float; float addition is not associative, so the
// same sequential order is kept here: lane c of wave 0 adds column c of every
// row in row order. The chain is fed from LDS. A block owns CCOLS columns;
// waves 1-3 stage tiles of CROWS rows (double-buffered), issuing all of a
// tile's loads before storing any, so about a tile of rows is in flight per
// column. Tombstoned rows and rows past n are staged as +0.0: a sum that
// starts at +0.0 never becomes -0.0, and adding +0.0 to it is exact, so this
// equals skipping them. Then every node's distance to the centroid
// (separately rounded multiply and add, index order) and the lowest id with
// the smallest distance.
// ---------------------------------------------------------------------------
constexpr int CCOLS = 16;   // columns per block (one 64-byte segment of a row)
constexpr int CROWS = 1024; // rows per staged tile
constexpr int CSTEP = 48;   // rows one staging round covers (3 waves x 16 rows)
constexpr int CLOADS = (CROWS + CSTEP - 1) / CSTEP;

__global__ __launch_bounds__(256) void centroid_kernel(const float* __restrict__ vecs, int64_t n,
                                                       int D, int DP, const uint32_t* del_bits,
                                                       float* centroid) {
#pragma clang fp contract(off)
    __shared__ __align__(16) float tile[2][CROWS][CCOLS];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int c0 = blockIdx.x * CCOLS;
    const int64_t ntiles = (n + CROWS - 1) / CROWS;
    // staging lane (waves 1-3): row (lane / 4) of its wave's 16, columns
    // c0 + 4 * (lane % 4) .. + 3 (one float4; DP is a multiple of 32, so the
    // columns past D that the float4 reads are zero padding)
    const int sr = (w - 1) * 16 + (l >> 2), sc = 4 * (l & 3);
    auto stage = [&](int b, int64_t tt) {
        float4 v[CLOADS];
#pragma unroll
        for (int k = 0; k < CLOADS; ++k) {
            const int r = sr + CSTEP * k;
            const int64_t i = tt * CROWS + r;
            v[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (r < CROWS && i < n && !(del_bits && bit_test(del_bits, (uint32_t)i)))
                v[k] = *reinterpret_cast<const float4*>(vecs + i * DP + c0 + sc);
        }
#pragma unroll
        for (int k = 0; k < CLOADS; ++k) {
            const int r = sr + CSTEP * k;
            if (r < CROWS) *reinterpret_cast<float4*>(&tile[b][r][sc]) = v[k];
        }
    };
    if (w > 0) stage(0, 0);
    __syncthreads();
    float acc = 0.0f;
    int64_t live = 0;
    for (int64_t tt = 0; tt < ntiles; ++tt) {
        const int b = (int)(tt & 1);
        if (w > 0) {
            if (tt + 1 < ntiles) stage(b ^ 1, tt + 1);
        } else {
            if (l < CCOLS) {
#pragma unroll 32
                for (int r = 0; r < CROWS; ++r) acc = acc + tile[b][r][l];
            }
            // live rows of this tile: lane l looks at rows l, l + 64, ...
            int cnt = 0;
            for (int r = l; r < CROWS; r += 64) {
                const int64_t i = tt * CROWS + r;
                cnt += i < n && !(del_bits && bit_test(del_bits, (uint32_t)i));
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
            live += cnt;
        }
        __syncthreads();
    }
    const int d = c0 + l;
    if (w == 0 && l < CCOLS && d < D) centroid[d] = live ? acc / (float)live : 0.0f;
}

__global__ void centroid_dist_kernel(const float* __restrict__ vecs, int64_t n, int D, int DP,
                                     const uint32_t* del_bits, const float* __restrict__ centroid,
                                     unsigned long long* best) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long key = ~0ull;
    if (i < n && !(del_bits && bit_test(del_bits, (uint32_t)i))) {
        float s = 0.0f;
        for (int k = 0; k < D; ++k) {
            const float diff = vecs[i * DP + k] - centroid[k];
            const float sq = diff * diff;
            s = s + sq;
        }
        // reference starts from min_dist = 1e30f with strict <
        if (s < 1e30f) key = ((unsigned long long)__float_as_uint(s) << 32) | (uint64_t)(uint32_t)i;
    }
    // wave min then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(key, off);
        key = o < key ? o : key;
    }
    if (lane() == 0 && key != ~0ull) atomicMin(best, key);
}

hipError_t launch_medoid(const float* vecs, int64_t n, int D, int DP, const uint32_t* del_bits,
                         float* centroid, unsigned long long* best, hipStream_t s) {
    hipLaunchKernelGGL(centroid_kernel, dim3((D + CCOLS - 1) / CCOLS), dim3(256), 0, s, vecs, n,
                       D, DP, del_bits, centroid);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_fill_u64(best, 1, ~0ull, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(centroid_dist_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       vecs, n, D, DP, del_bits, centroid, best);
    return hipGetLastError();
}

// Build work counters of one pruned item (lane 0 adds them).
__device__ __forceinline__ void count_item(unsigned long long* acc, int nread, int pairs,
                                           int64_t lists, int64_t revs) {
    if (!acc || lane() != 0) return;
    unsigned long long* c = acc_slot(acc);
    atomicAdd(&c[VMN_ACC_ITEMS], 1ull);
    atomicAdd(&c[VMN_ACC_ROWS], (unsigned long long)nread);
    if (pairs) atomicAdd(&c[VMN_ACC_PAIRS], (unsigned long long)pairs);
    if (lists) atomicAdd(&c[VMN_ACC_LISTS], (unsigned long long)lists);
    if (revs) atomicAdd(&c[VMN_ACC_REVS], (unsigned long long)revs);
}

// ---------------------------------------------------------------------------
// PARTH prune: parth/src/build_graph.cu:173-264 (`robustPruneKernel`).
// Candidates = the node's search list (minus -1 and the node itself, already
// in distance order) followed by its current neighbours that are not in the
// list (fresh distances); a stable sort by distance (the reference's bubble
// sort) equals a stable merge of the two with list entries first on ties.
// ---------------------------------------------------------------------------
template <int DP, int CCAP, bool BF>
__global__ __launch_bounds__(64) void parth_prune_kernel(PruneArgs a) {
    __shared__ __align__(16) uint32_t c_id[2][CCAP];
    __shared__ __align__(16) float c_d[2][CCAP];
    __shared__ int32_t e_id[64];
    __shared__ float e_d[64];
    __shared__ uint8_t kept[CCAP];
    __shared__ int32_t act_id[CCAP];
    __shared__ int32_t act_pos[CCAP];
    __shared__ float act_d[CCAP];
    __shared__ int32_t newrow[64];
    __shared__ int64_t s_q;
    const int l = threadIdx.x;
    const int W = a.W, Lc = a.Lc;
    const PruneRows rows{a.vecs, a.bf16_rows, a.bf16_norms};
    for (;;) {
        const int64_t i = next_item(a.counter, &s_q);
        if (i >= a.n) return;
        const int64_t node = a.node_base + i;
        int m = 0;
        for (int b = 0; b < Lc; b += 64) {
            const int j = b + l;
            const int32_t id = j < Lc ? a.cand_ids[i * Lc + j] : -1;
            const float d = j < Lc ? a.cand_d[i * Lc + j] : 0.0f;
            const bool f = id != -1 && id != (int32_t)node;
            const uint64_t mk = ballot(f);
            if (f) {
                const int k = m + __popcll((unsigned long long)(mk & lanemask_lt()));
                c_id[0][k] = (uint32_t)id;
                c_d[0][k] = d;
            }
            m += __popcll((unsigned long long)mk);
        }
        const int32_t nb = l < W ? a.adj[node * W + l] : -1;
        __syncthreads();
        bool pass = nb != -1 && nb != (int32_t)node;
        {
            bool dup = false;
            for (int t = 0; t < m; ++t) dup |= c_id[0][t] == (uint32_t)nb;
            pass = pass && !dup;
        }
        // FP32 rows: distances in the reference's sequential order
        int nread = 0;
        PruneQ<DP, BF, !BF> q;
        q.load(rows, node);
        const int ne = row_extras<DP, BF, !BF>(rows, q, pass, nb, e_id, e_d, &nread);
        const uint32_t my_id = l < ne ? (uint32_t)e_id[l] : 0u;
        const float my_d = l < ne ? e_d[l] : 0.0f;
        const int cnt = merge_sorted<CCAP, false>(c_id[0], c_d[0], m, my_id, my_d, ne, c_id[1],
                                                  c_d[1], CCAP);
        __syncthreads();
        const int nk = prune_list<DP, true, BF, CCAP, !BF>(rows, c_id[1], c_d[1], cnt, a.alpha,
                                                           a.R, kept, act_id, act_pos, act_d,
                                                           newrow, nullptr, &nread);
        for (int j = l; j < W; j += 64) a.adj[node * W + j] = j < nk ? newrow[j] : -1;
        count_item(a.acc, nread, 0, Lc, 0);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// FRESH out-neighbour prune: vamana/include/kernels.cuh:358-500
// (`computeOutNeighbors`) + `pruneOutNeighbors` :231-356. Candidates = the
// visited set stably sorted by distance (expansion order on ties) merged with
// the current neighbours stably sorted by distance (adjacency order on ties),
// visited-set entries first on ties, capped at P; the node itself is left out
// (the reference can select the medoid as its own p*), and so are tombstoned
// nodes (a tombstoned start node can be expanded). Prune with <=. Each
// kept p* emits the reverse pair (p* <- node) with d(node, p*) (:292-301).
// ---------------------------------------------------------------------------
template <int DP, int CCAP, bool BF>
__global__ __launch_bounds__(64) void fresh_out_prune_kernel(PruneArgs a) {
    __shared__ __align__(16) uint32_t c_id[2][CCAP];
    __shared__ __align__(16) float c_d[2][CCAP];
    __shared__ int32_t e_id[64];
    __shared__ float e_d[64];
    __shared__ uint8_t kept[CCAP];
    __shared__ int32_t act_id[CCAP];
    __shared__ int32_t act_pos[CCAP];
    __shared__ float act_d[CCAP];
    __shared__ int32_t newrow[64];
    __shared__ float newd[64];
    __shared__ int64_t s_q;
    const int l = threadIdx.x;
    const int W = a.W, P = a.P;
    const PruneRows rows{a.vecs, a.bf16_rows, a.bf16_norms};
    for (;;) {
        const int64_t it = next_item(a.counter, &s_q);
        if (it >= a.n) return;
        const int64_t i = a.items ? a.items[it] : it;
        const int64_t node = a.node_base + i;
        const int vc = a.vis_cnt[i];
        // at most vc + W candidates: beyond this list's capacity, leave the
        // node to the full-capacity launch
        if (a.overflow && vc + W > CCAP) {
            if (l == 0) a.overflow[atomicAdd(a.overflow_n, 1ull)] = i;
            continue;
        }
        int cur = 0, cnt = 0;
        for (int b = 0; b < vc; b += 64) {
            const int j = b + l;
            const int32_t id = j < vc ? a.vis_ids[i * P + j] : -1;
            const float d = j < vc ? a.vis_d[i * P + j] : 0.0f;
            const bool f = j < vc && id != (int32_t)node &&
                           !(a.del_bits && bit_test(a.del_bits, (uint32_t)id));
            const uint64_t mk = ballot(f);
            const int k = __popcll((unsigned long long)(mk & lanemask_lt()));
            const int nn = __popcll((unsigned long long)mk);
            if (f) {
                e_id[k] = id;
                e_d[k] = d;
            }
            __syncthreads();
            const uint32_t my_id = l < nn ? (uint32_t)e_id[l] : 0u;
            const float my_d = l < nn ? e_d[l] : 0.0f;
            __syncthreads();
            cnt = merge_sorted<CCAP, false>(c_id[cur], c_d[cur], cnt, my_id, my_d, nn,
                                            c_id[cur ^ 1], c_d[cur ^ 1], CCAP);
            cur ^= 1;
            __syncthreads();
        }
        const int32_t nb = l < W ? a.adj[node * W + l] : -1;
        const bool pass = nb != -1 && nb != (int32_t)node &&
                          !(a.del_bits && bit_test(a.del_bits, (uint32_t)nb));
        int nread = 0;
        PruneQ<DP, BF> q;
        q.load(rows, node);
        const int ne = row_extras<DP, BF>(rows, q, pass, nb, e_id, e_d, &nread);
        const uint32_t my_id = l < ne ? (uint32_t)e_id[l] : 0u;
        const float my_d = l < ne ? e_d[l] : 0.0f;
        cnt = merge_sorted<CCAP, false>(c_id[cur], c_d[cur], cnt, my_id, my_d, ne, c_id[cur ^ 1],
                                        c_d[cur ^ 1], P);
        cur ^= 1;
        __syncthreads();
        const int nk = prune_list<DP, false, BF, CCAP>(rows, c_id[cur], c_d[cur], cnt, a.alpha, a.R,
                                               kept, act_id, act_pos, act_d, newrow, newd,
                                               &nread);
        count_item(a.acc, nread, nk, vc, 0);
        int32_t* row = a.adj_out ? a.adj_out + i * W : a.adj + node * W;
        for (int j = l; j < W; j += 64) row[j] = j < nk ? newrow[j] : -1;
        for (int j = l; j < a.R; j += 64) {
            a.pair_dst[i * a.R + j] = j < nk ? newrow[j] : -1;
            a.pair_d[i * a.R + j] = j < nk ? newd[j] : 0.0f;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// FRESH reverse-edge prune: vamana/include/kernels.cuh:745-903
// (`computeReverseEdges`) + `mergeIntoReverseEdges` :566-624 +
// `pruneReverseEdges` :626-743. For a node with reverse candidates: those
// candidates ordered by (distance, source id) (the reference's order is
// atomic-arrival dependent), first M kept, merged with the node's current
// neighbours (stable by distance, reverse candidates first on ties), capped at
// M, then pruned with <= into at most R neighbours. Only nodes that received
// reverse candidates are visited (the compacted list the reference builds at
// :539-564); re-pruning an unchanged row with the alpha it was pruned with
// returns it unchanged.
// ---------------------------------------------------------------------------
template <int DP, int CCAP, bool BF>
__global__ __launch_bounds__(64) void fresh_rev_prune_kernel(PruneArgs a) {
    __shared__ __align__(16) uint32_t c_id[2][CCAP];
    __shared__ __align__(16) float c_d[2][CCAP];
    __shared__ int32_t e_id[64];
    __shared__ float e_d[64];
    __shared__ uint8_t kept[CCAP];
    __shared__ int32_t act_id[CCAP];
    __shared__ int32_t act_pos[CCAP];
    __shared__ float act_d[CCAP];
    __shared__ int32_t newrow[64];
    __shared__ int64_t s_q;
    const int l = threadIdx.x;
    const int W = a.W, M = a.M;
    const PruneRows rows{a.vecs, a.bf16_rows, a.bf16_norms};
    for (;;) {
        const int64_t it = next_item(a.counter, &s_q);
        if (it >= a.n) return;
        const int64_t i = a.items ? a.items[it] : it;
        const int64_t node = a.node_list[i];
        const int64_t o0 = a.list_off ? a.list_off[i] : a.rev_off[node];
        const int64_t o1 = a.list_off ? a.list_off[i + 1] : a.rev_off[node + 1];
        // at most (o1 - o0) + W candidates: see the out-prune
        if (a.overflow && (o1 - o0) + W > CCAP) {
            if (l == 0) a.overflow[atomicAdd(a.overflow_n, 1ull)] = i;
            continue;
        }
        int cur = 0, cnt = 0;
        for (int64_t b = o0; b < o1; b += 64) {
            const int64_t j = b + l;
            const bool f = j < o1;
            const uint32_t my_id = f ? (uint32_t)a.rev_src[j] : 0u;
            const float my_d = f ? a.rev_d[j] : 0.0f;
            const int nn = (int)((o1 - b) < 64 ? (o1 - b) : 64);
            cnt = merge_sorted<CCAP, true>(c_id[cur], c_d[cur], cnt, my_id, my_d, nn,
                                           c_id[cur ^ 1], c_d[cur ^ 1], M);
            cur ^= 1;
            __syncthreads();
        }
        const int32_t nb = l < W ? a.adj[node * W + l] : -1;
        const bool pass = nb != -1 && nb != (int32_t)node;
        int nread = 0;
        PruneQ<DP, BF> q;
        q.load(rows, node);
        const int ne = row_extras<DP, BF>(rows, q, pass, nb, e_id, e_d, &nread);
        const uint32_t my_id = l < ne ? (uint32_t)e_id[l] : 0u;
        const float my_d = l < ne ? e_d[l] : 0.0f;
        cnt = merge_sorted<CCAP, false>(c_id[cur], c_d[cur], cnt, my_id, my_d, ne, c_id[cur ^ 1],
                                        c_d[cur ^ 1], M);
        cur ^= 1;
        __syncthreads();
        const int nk = prune_list<DP, false, BF, CCAP>(rows, c_id[cur], c_d[cur], cnt, a.alpha, a.R,
                                               kept, act_id, act_pos, act_d, newrow, nullptr,
                                               &nread);
        count_item(a.acc, nread, 0, 0, o1 - o0);
        for (int j = l; j < W; j += 64) a.adj[node * W + j] = j < nk ? newrow[j] : -1;
        __syncthreads();
    }
}

#define VMN_DISPATCH_DP(DPV, CALL)          \
    switch (DPV) {                           \
        case 32: { constexpr int DPc = 32; CALL; } break;   \
        case 64: { constexpr int DPc = 64; CALL; } break;   \
        case 96: { constexpr int DPc = 96; CALL; } break;   \
        case 128: { constexpr int DPc = 128; CALL; } break; \
        case 160: { constexpr int DPc = 160; CALL; } break; \
        case 192: { constexpr int DPc = 192; CALL; } break; \
        case 256: { constexpr int DPc = 256; CALL; } break; \
        default: return hipErrorInvalidValue;             \
    }

// BF16 rows when the index supplies them (exact distances), else FP32 rows
#define VMN_PRUNE_LAUNCH(KERNEL, CAP)                                                        \
    do {                                                                                    \
        if (a.bf16_rows) {                                                                  \
            VMN_DISPATCH_DP(a.DP, hipLaunchKernelGGL((KERNEL<DPc, CAP, true>), dim3(grid),   \
                                                     dim3(64), 0, s, a));                   \
        } else {                                                                            \
            VMN_DISPATCH_DP(a.DP, hipLaunchKernelGGL((KERNEL<DPc, CAP, false>), dim3(grid),  \
                                                     dim3(64), 0, s, a));                   \
        }                                                                                   \
    } while (0)

hipError_t launch_parth_prune(const PruneArgs& a, int grid, hipStream_t s) {
    const int need = a.Lc + a.W;
    if (need <= 128) {
        VMN_PRUNE_LAUNCH(parth_prune_kernel, 128);
    } else if (need <= 256) {
        VMN_PRUNE_LAUNCH(parth_prune_kernel, 256);
    } else if (need <= 576) {
        VMN_PRUNE_LAUNCH(parth_prune_kernel, 576);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fresh_out_prune(const PruneArgs& a, int grid, hipStream_t s) {
    // list capacity must hold min(P, visited + W) entries (first tier: 256,
    // the larger ones are listed)
    const int need = a.overflow ? 256 : a.P;
    if (need <= 256) {
        VMN_PRUNE_LAUNCH(fresh_out_prune_kernel, 256);
    } else if (need <= 640) {
        VMN_PRUNE_LAUNCH(fresh_out_prune_kernel, 640);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fresh_rev_prune(const PruneArgs& a, int grid, hipStream_t s) {
    const int need = a.overflow ? 256 : a.M;
    if (need <= 256) {
        VMN_PRUNE_LAUNCH(fresh_rev_prune_kernel, 256);
    } else if (need <= 512) {
        VMN_PRUNE_LAUNCH(fresh_rev_prune_kernel, 512);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Reverse-edge grouping: count pairs per destination (atomics), exclusive scan
// to CSR offsets, scatter (src, d) into each destination's segment (slot from
// an atomic per destination; the order inside a segment is then fixed by the
// (distance, id) keyed merge of the prune kernel), and the compacted list of
// destinations with at least one reverse candidate.
// ---------------------------------------------------------------------------
__global__ void reverse_count_kernel(const int32_t* pair_dst, int64_t npairs, int32_t* counts) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < npairs; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t d = pair_dst[i];
        if (d >= 0) atomicAdd(&counts[d], 1);
    }
}
hipError_t launch_reverse_count(const int32_t* pair_dst, int64_t npairs, int32_t* counts, hipStream_t s) {
    hipLaunchKernelGGL(reverse_count_kernel, dim3(fill_grid(npairs)), dim3(256), 0, s, pair_dst, npairs, counts);
    return hipGetLastError();
}

// Exclusive scan in three passes: per-1024-block sums, a single-block scan of
// those sums, then each block's local scan plus its offset. off has n+1 entries.
__global__ void scan_block_sums(const int32_t* in, int64_t n, int64_t* bsum) {
    __shared__ int64_t red[256];
    const int64_t base = (int64_t)blockIdx.x * 1024;
    int64_t s = 0;
    for (int k = threadIdx.x; k < 1024; k += 256) {
        const int64_t i = base + k;
        if (i < n) s += in[i];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}
__global__ void scan_sums_single(int64_t* bsum, int64_t nb) {
    if (threadIdx.x != 0) return;
    int64_t run = 0;
    for (int64_t i = 0; i < nb; ++i) {
        const int64_t v = bsum[i];
        bsum[i] = run;
        run += v;
    }
    bsum[nb] = run;
}
__global__ void scan_local(const int32_t* in, int64_t n, const int64_t* bsum, int64_t* out) {
    __shared__ int64_t buf[1024];
    const int64_t base = (int64_t)blockIdx.x * 1024;
    for (int k = threadIdx.x; k < 1024; k += 256) {
        const int64_t i = base + k;
        buf[k] = i < n ? in[i] : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = bsum[blockIdx.x];
        for (int k = 0; k < 1024; ++k) {
            const int64_t v = buf[k];
            buf[k] = run;
            run += v;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 1024; k += 256) {
        const int64_t i = base + k;
        if (i < n) out[i] = buf[k];
    }
}
__global__ void scan_tail(int64_t* out, int64_t n, const int64_t* bsum, int64_t nb) {
    out[n] = bsum[nb];
}
hipError_t launch_exclusive_scan(const int32_t* counts, int64_t n, int64_t* off, int64_t* tmp,
                                 hipStream_t s) {
    const int64_t nb = (n + 1023) / 1024;
    hipLaunchKernelGGL(scan_block_sums, dim3((unsigned)nb), dim3(256), 0, s, counts, n, tmp);
    hipLaunchKernelGGL(scan_sums_single, dim3(1), dim3(64), 0, s, tmp, nb);
    hipLaunchKernelGGL(scan_local, dim3((unsigned)nb), dim3(256), 0, s, counts, n, tmp, off);
    hipLaunchKernelGGL(scan_tail, dim3(1), dim3(1), 0, s, off, n, tmp, nb);
    return hipGetLastError();
}

__global__ void reverse_scatter_kernel(const int32_t* pair_dst, const float* pair_d, int64_t npairs,
                                       int R, int64_t node_base, const int64_t* off, int32_t* fill,
                                       int32_t* rev_src, float* rev_d) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < npairs; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t d = pair_dst[i];
        if (d < 0) continue;
        const int64_t pos = off[d] + atomicAdd(&fill[d], 1);
        rev_src[pos] = (int32_t)(node_base + i / R);
        rev_d[pos] = pair_d[i];
    }
}
hipError_t launch_reverse_scatter(const int32_t* pair_dst, const float* pair_d, int64_t npairs,
                                  int R, int64_t node_base, const int64_t* off, int32_t* fill,
                                  int32_t* rev_src, float* rev_d, hipStream_t s) {
    hipLaunchKernelGGL(reverse_scatter_kernel, dim3(fill_grid(npairs)), dim3(256), 0, s, pair_dst,
                       pair_d, npairs, R, node_base, off, fill, rev_src, rev_d);
    return hipGetLastError();
}

// getPrunableQueryIDs (vamana/include/kernels.cuh:539-564): wave-ballot
// compaction of nodes with a non-zero count.
// list[0..nlist) = the indices i with counts[i] != 0, in no particular order
// (nlist zero on entry). A workgroup of 256 takes a tile of 4,096 entries:
// it counts them, reserves its range with one atomic, then writes them in
// tile order. One atomic per wavefront queued them on the counter's L2 line
// when most wavefronts hold a nonzero entry (10M entries: 1.2 ms).
#define VMN_COMPACT_CHUNKS 16
__global__ __launch_bounds__(256) void compact_nonzero_kernel(const int32_t* counts, int64_t n,
                                                              int64_t* list, unsigned long long* nlist) {
    __shared__ int s_w[4];
    __shared__ unsigned long long s_base;
    const int t = threadIdx.x, w = t >> 6;
    const int64_t tile = (int64_t)blockIdx.x * (256 * VMN_COMPACT_CHUNKS);
    int cnt = 0;
#pragma unroll
    for (int c = 0; c < VMN_COMPACT_CHUNKS; ++c) {
        const int64_t i = tile + 256 * c + t;
        cnt += (i < n && counts[i] != 0) ? 1 : 0;
    }
    // block total: wave sums, then thread 0
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if (lane() == 0) s_w[w] = cnt;
    __syncthreads();
    if (t == 0) {
        const int tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        s_base = tot ? atomicAdd(nlist, (unsigned long long)tot) : 0ull;
    }
    __syncthreads();
    unsigned long long base = s_base;
    for (int c = 0; c < VMN_COMPACT_CHUNKS; ++c) {
        const int64_t i = tile + 256 * c + t;
        const bool f = i < n && counts[i] != 0;
        const uint64_t m = ballot(f);
        __syncthreads();  // s_w of the previous chunk has been read
        if (lane() == 0) s_w[w] = __popcll((unsigned long long)m);
        __syncthreads();
        unsigned long long off = base;
        for (int k = 0; k < w; ++k) off += (unsigned long long)s_w[k];
        if (f) list[off + __popcll((unsigned long long)(m & lanemask_lt()))] = i;
        base += (unsigned long long)(s_w[0] + s_w[1] + s_w[2] + s_w[3]);
    }
}
hipError_t launch_compact_nonzero(const int32_t* counts, int64_t n, int64_t* list,
                                  unsigned long long* nlist, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t per = 256 * VMN_COMPACT_CHUNKS;
    hipLaunchKernelGGL(compact_nonzero_kernel, dim3((unsigned)((n + per - 1) / per)), dim3(256), 0, s,
                       counts, n, list, nlist);
    return hipGetLastError();
}

// Work order of the reverse prune: items whose candidates (reverse ones plus
// the row) exceed `heavy` go first, the rest after them, so the hubs start at
// the beginning of the launch instead of forming its tail. The order inside
// each class follows atomic arrival, which does not matter: each item's prune
// reads and writes only its own row.
__global__ void heavy_first_kernel(const int64_t* node_list, const int64_t* list_off,
                                   const int64_t* rev_off, int64_t n, int W, int heavy,
                                   int64_t* order, unsigned long long* ctr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t node = node_list[i];
    const int64_t c = list_off ? list_off[i + 1] - list_off[i] : rev_off[node + 1] - rev_off[node];
    if (c + W > heavy) order[atomicAdd(&ctr[0], 1ull)] = i;
    else order[n - 1 - (int64_t)atomicAdd(&ctr[1], 1ull)] = i;
}
hipError_t launch_heavy_first(const int64_t* node_list, const int64_t* list_off,
                              const int64_t* rev_off, int64_t n, int W, int heavy, int64_t* order,
                              unsigned long long* ctr, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(heavy_first_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       node_list, list_off, rev_off, n, W, heavy, order, ctr);
    return hipGetLastError();
}

}  // namespace vmn
