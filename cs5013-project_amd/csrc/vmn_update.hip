// FreshVamana updates: tombstones, delete-by-vector, delete consolidation and
// the entry-point move after deletes.
#include "vmn_prune.hpp"
#include "vmn_update.hpp"

namespace vmn {

// Tombstone a list of ids (vamana/include/delete_list.cuh keeps an append-only
// id list scanned per neighbour; a bitmap gives the same membership answer in
// one load). Ids outside [0, n) are counted in *bad and ignored.
__global__ void tombstone_kernel(const int32_t* ids, int64_t m, int64_t n, uint32_t* del_bits,
                                 unsigned long long* bad) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t id = ids[i];
        if (id < 0 || id >= n) {
            atomicAdd(bad, 1ull);
            continue;
        }
        atomicOr(&del_bits[id >> 5], 1u << (id & 31));
    }
}
hipError_t launch_tombstone(const int32_t* ids, int64_t m, int64_t n, uint32_t* del_bits,
                            unsigned long long* bad, hipStream_t s) {
    const int64_t g = (m + 255) / 256;
    hipLaunchKernelGGL(tombstone_kernel, dim3((unsigned)(g < 1 ? 1 : (g > 4096 ? 4096 : g))),
                       dim3(256), 0, s, ids, m, n, del_bits, bad);
    return hipGetLastError();
}

// Exact-match lookup of query vectors among live nodes
// (vamana/include/kernels.cuh:905-933 `findPointsKernel`): out[q] = the lowest
// matching live id, or -1. out must hold 0xffffffff before the launch.
__global__ void find_vectors_kernel(const float* __restrict__ vecs, int64_t n, int D, int DP,
                                    const uint32_t* del_bits, const float* __restrict__ qv,
                                    int64_t nq, uint32_t* out) {
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t q = blockIdx.y;
    if (node >= n || q >= nq) return;
    if (bit_test(del_bits, (uint32_t)node)) return;
    const float* a = vecs + node * DP;
    const float* b = qv + q * DP;
    for (int k = 0; k < D; ++k)
        if (a[k] != b[k]) return;
    atomicMin(&out[q], (uint32_t)node);
}
hipError_t launch_find_vectors(const float* vecs, int64_t n, int D, int DP, const uint32_t* del_bits,
                               const float* qv, int64_t nq, uint32_t* out, hipStream_t s) {
    dim3 g((unsigned)((n + 255) / 256), (unsigned)nq);
    hipLaunchKernelGGL(find_vectors_kernel, g, dim3(256), 0, s, vecs, n, D, DP, del_bits, qv, nq, out);
    return hipGetLastError();
}

// flags[p] = 1 for live nodes with at least one tombstoned out-neighbour;
// flags must be zero on entry. One lane per adjacency entry, so a wavefront
// reads 256 contiguous bytes of the rows (one lane walking its own row made
// each load instruction touch 64 rows: 10M x 64 rows took 11 ms). Lanes that
// find a tombstone write 1 (every writer stores the same value). The
// tombstone test first reads a summary bit per 64 nodes held in LDS (set when
// any of the 64 is tombstoned; 19.5 KB at 10M nodes), so with few tombstones
// almost no entry reads the full bitmap; `summary` null: every entry reads it.
__global__ void del_summary_kernel(const uint32_t* del_bits, int64_t nwords, uint32_t* summary) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // 64-node block
    if (2 * b >= nwords) return;
    const uint32_t w = del_bits[2 * b] | (2 * b + 1 < nwords ? del_bits[2 * b + 1] : 0u);
    if (w) atomicOr(&summary[b >> 5], 1u << (b & 31));
}
__global__ __launch_bounds__(256) void cons_flags_kernel(const int32_t* adj, int64_t n, int W,
                                                         const uint32_t* del_bits,
                                                         const uint32_t* summary, int nsum,
                                                         int32_t* flags) {
    extern __shared__ uint32_t s_sum[];
    for (int i = threadIdx.x; i < nsum; i += blockDim.x) s_sum[i] = summary[i];
    __syncthreads();
    // four entries per lane and step, their loads issued together
    const int64_t ne = n * W, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < ne; i0 += 4 * stride) {
        int32_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t i = i0 + u * stride;
            v[u] = i < ne ? adj[i] : -1;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (v[u] < 0) continue;
            const uint32_t blk = (uint32_t)v[u] >> 6;
            if (summary && !((s_sum[blk >> 5] >> (blk & 31)) & 1u)) continue;
            if (bit_test(del_bits, (uint32_t)v[u])) {
                const int64_t p = (i0 + u * stride) / W;
                if (!bit_test(del_bits, (uint32_t)p)) flags[p] = 1;
            }
        }
    }
}
__global__ void zero_i32_kernel(int32_t* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
// summary: scratch of (n + 2047) / 2048 words, or null
hipError_t launch_cons_flags(const int32_t* adj, int64_t n, int W, const uint32_t* del_bits,
                             int32_t* flags, uint32_t* summary, int cus, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, flags, n);
    // the summary is read into 48 KB of LDS at most (25M nodes)
    const int64_t nblk = (n + 63) / 64, nsum64 = (nblk + 31) / 32;
    const int nsum = summary && nsum64 <= 12288 ? (int)nsum64 : 0;
    if (nsum) {
        hipLaunchKernelGGL(zero_i32_kernel, dim3((unsigned)((nsum + 255) / 256)), dim3(256), 0, s,
                           reinterpret_cast<int32_t*>(summary), (int64_t)nsum);
        const int64_t nwords = (n + 31) / 32;
        hipLaunchKernelGGL(del_summary_kernel, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, s,
                           del_bits, nwords, summary);
    }
    const int64_t ne = n * W;
    const int64_t want = (ne + 255) / 256, cap = (int64_t)(cus > 0 ? cus : 256) * 8;
    const unsigned grid = (unsigned)(want < cap ? want : cap);
    hipLaunchKernelGGL(cons_flags_kernel, dim3(grid), dim3(256), (size_t)nsum * 4, s, adj, n, W,
                       del_bits, nsum ? summary : nullptr, nsum, flags);
    return hipGetLastError();
}

// Delete consolidation, one wavefront per listed live node p with a tombstoned
// out-neighbour (the reference has none; after FreshDiskANN, Singh et al.
// 2021, Algorithm 4, with the live part of the row kept). The row's live
// neighbours K all stay, in row order. Candidates for the slots the
// tombstones free are the live out-neighbours of p's tombstoned neighbours
// (rows in adjacency order; not p, not in K, not already pooled), kept as the
// first M by (distance, id). Each K member drops the candidates c it occludes
// (alpha * d(k, c) <= d(p, c)); the rest are taken in order as robust prune
// does (each taken candidate drops the later ones it occludes) until the row
// holds R. Re-pruning the whole pool instead lets the 2-hop candidates
// occlude long edges the build chose from a full search, and recall decays
// with every round of deletes.
template <int DP, int CCAP, bool BF>
__global__ __launch_bounds__(64) void consolidate_kernel(ConsArgs a) {
    __shared__ __align__(16) uint32_t c_id[2][CCAP];
    __shared__ __align__(16) float c_d[2][CCAP];
    __shared__ int32_t e_id[64];
    __shared__ float e_d[64];
    __shared__ uint8_t kept[CCAP];
    __shared__ int32_t act_id[CCAP];
    __shared__ int32_t act_pos[CCAP];
    __shared__ float act_d[CCAP];
    __shared__ __align__(16) int32_t k_id[64];
    __shared__ int32_t newrow[64];
    __shared__ int64_t s_q;
    const int l = threadIdx.x;
    const int W = a.W;
    for (;;) {
        const int64_t it = next_item(a.counter, &s_q);
        if (it >= a.nlist) return;
        const int64_t i = a.items ? a.items[it] : it;
        const int64_t p = a.list[i];
        const PruneRows rows{a.vecs, a.bf16_rows, a.bf16_norms};
        PruneQ<DP, BF> q;
        q.load(rows, p);
        const int32_t own = l < W ? a.adj[p * W + l] : -1;
        const bool own_dead = own >= 0 && bit_test(a.del_bits, (uint32_t)own);
        const uint64_t dmask = ballot(own_dead);
        const uint64_t kmask = ballot(own >= 0 && !own_dead);
        const int nK = __popcll((unsigned long long)kmask);
        // the pool takes at most W entries per tombstoned neighbour
        if (a.overflow && __popcll((unsigned long long)dmask) * W > CCAP) {
            if (l == 0) a.overflow[atomicAdd(a.overflow_n, 1ull)] = i;
            continue;
        }
        if (own >= 0 && !own_dead) k_id[__popcll((unsigned long long)(kmask & lanemask_lt()))] = own;
        __syncthreads();
        int cur = 0, cnt = 0;
        for (uint64_t rem = dmask; rem; rem &= rem - 1) {
            const int32_t src = readlane_i(own, __ffsll((unsigned long long)rem) - 1);
            const int32_t v = l < W ? a.adj[(int64_t)src * W + l] : -1;
            bool pass = v >= 0 && v != (int32_t)p && !bit_test(a.del_bits, (uint32_t)v);
            if (pass) {
                bool dup = false;
                for (int t = 0; t < nK; ++t) dup |= k_id[t] == v;
                for (int t = 0; t < cnt; ++t) dup |= c_id[cur][t] == (uint32_t)v;
                pass = !dup;
            }
            const int ne = row_extras<DP, BF>(rows, q, pass, v, e_id, e_d);
            const uint32_t my_id = l < ne ? (uint32_t)e_id[l] : 0u;
            const float my_d = l < ne ? e_d[l] : 0.0f;
            cnt = merge_sorted<CCAP, true>(c_id[cur], c_d[cur], cnt, my_id, my_d, ne,
                                           c_id[cur ^ 1], c_d[cur ^ 1], a.M);
            cur ^= 1;
            __syncthreads();
        }
        // K stays; each member drops the candidates it occludes
        for (int t = l; t < cnt; t += 64) kept[t] = 1;
        if (l < nK) newrow[l] = k_id[l];
        __syncthreads();
        int nk = nK;
        if constexpr (BF) {
            occlude_mfma<DP>(rows, k_id, nK, c_id[cur], c_d[cur], cnt, a.alpha, kept, act_id,
                             act_pos, reinterpret_cast<uint32_t*>(act_d));
            // then the remaining candidates in order, as robust prune takes them
            nk = robust_prune_mfma<DP, false, CCAP>(rows, c_id[cur], c_d[cur], cnt, a.alpha, a.R,
                                                    kept, act_id, act_pos,
                                                    reinterpret_cast<uint32_t*>(act_d), newrow,
                                                    nullptr, nK, false);
        } else {
            occlude_fp32<DP>(rows, k_id, nK, c_id[cur], c_d[cur], cnt, a.alpha, kept, act_id,
                             act_pos);
            // then the remaining candidates in order, as robust prune takes them
            nk = robust_prune_batch<DP, false, CCAP>(rows, c_id[cur], c_d[cur], cnt, a.alpha, a.R,
                                                     kept, act_id, act_pos,
                                                     reinterpret_cast<uint32_t*>(act_d), newrow,
                                                     nullptr, nK, false);
        }
        __syncthreads();
        for (int j = l; j < W; j += 64) a.adj[p * W + j] = j < nk ? newrow[j] : -1;
        __syncthreads();
    }
}

hipError_t launch_consolidate(const ConsArgs& a, int grid, hipStream_t s) {
    if (a.M > 512) return hipErrorInvalidValue;
    switch (a.DP) {
#define VMN_CONS_CASE(DPV)                                                                        \
    case DPV:                                                                                     \
        if (a.overflow && a.bf16_rows)                                                            \
            hipLaunchKernelGGL((consolidate_kernel<DPV, 256, true>), dim3(grid), dim3(64), 0, s, a); \
        else if (a.overflow)                                                                      \
            hipLaunchKernelGGL((consolidate_kernel<DPV, 256, false>), dim3(grid), dim3(64), 0, s, a); \
        else if (a.bf16_rows)                                                                     \
            hipLaunchKernelGGL((consolidate_kernel<DPV, 512, true>), dim3(grid), dim3(64), 0, s, a); \
        else                                                                                      \
            hipLaunchKernelGGL((consolidate_kernel<DPV, 512, false>), dim3(grid), dim3(64), 0, s, a); \
        break;
        VMN_CONS_CASE(32)
        VMN_CONS_CASE(64)
        VMN_CONS_CASE(96)
        VMN_CONS_CASE(128)
        VMN_CONS_CASE(160)
        VMN_CONS_CASE(192)
        VMN_CONS_CASE(256)
#undef VMN_CONS_CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Rows of tombstoned nodes are emptied once consolidation no longer reads them.
// Rows of tombstoned nodes become empty: one thread per 32-node word of the
// tombstone bitmap writes -1 over the rows of its set bits (only those rows
// are written; one thread per adjacency entry took 0.7 ms at 10M x 64).
__global__ void clear_deleted_rows_kernel(int32_t* adj, int64_t n, int W, const uint32_t* del_bits) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w * 32 >= n) return;
    uint32_t bits = del_bits[w];
    while (bits) {
        const int b = __ffs(bits) - 1;
        bits &= bits - 1;
        const int64_t node = w * 32 + b;
        if (node >= n) break;
        int32_t* row = adj + node * W;
        if ((W & 3) == 0) {
            for (int j = 0; j < W; j += 4) *reinterpret_cast<int4*>(row + j) = make_int4(-1, -1, -1, -1);
        } else {
            for (int j = 0; j < W; ++j) row[j] = -1;
        }
    }
}
hipError_t launch_clear_deleted_rows(int32_t* adj, int64_t n, int W, const uint32_t* del_bits,
                                     hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t nwords = (n + 31) / 32;
    hipLaunchKernelGGL(clear_deleted_rows_kernel, dim3((unsigned)((nwords + 255) / 256)), dim3(256),
                       0, s, adj, n, W, del_bits);
    return hipGetLastError();
}

}  // namespace vmn
