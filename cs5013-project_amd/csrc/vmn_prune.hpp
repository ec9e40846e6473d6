// Robust prune and candidate helpers shared by the build and update kernels.
#pragma once
#include "vmn_kernels.hpp"

namespace vmn {

// Where a prune reads node rows: the FP32 rows, and the BF16 copies with
// their squared norms (null unless every row is exact in BF16).
struct PruneRows {
    const float* f32;
    const void* bf16;
    const float* bf16_norms;
};

// One node's row as the "query" of a prune's distances. BF = false: FP32 rows,
// 8 lanes per row. BF = true: BF16 rows on MFMA, 16 rows per tile. BF is only
// taken when every row of the index holds integers of magnitude <= 256 with
// squared norm <= 2^22: BF16 then holds the values exactly and every partial
// sum of ||a||^2 + ||b||^2 - 2 a.b is an integer below 2^24, so the distance
// equals the FP32 sum of squared differences bit for bit, from half the bytes.
template <int DP, bool BF>
struct PruneQ;

template <int DP>
struct PruneQ<DP, false> {
    QReg<DP> q;
    __device__ __forceinline__ void load(const PruneRows& r, int64_t id) {
        q.load(r.f32 + id * DP, lane() & 7);
    }
    __device__ __forceinline__ void dists(const PruneRows& r, const int32_t* ids, float* out,
                                          int cnt) const {
        gather_dists<DP>(q, r.f32, ids, out, cnt);
    }
};

template <int DP>
struct PruneQ<DP, true> {
    QLow<DP> q;
    __device__ __forceinline__ void load(const PruneRows& r, int64_t id) {
        prep_query<DP, VMN_PREC_BF16>(q, r.f32 + id * DP, nullptr);
    }
    __device__ __forceinline__ void dists(const PruneRows& r, const int32_t* ids, float* out,
                                          int cnt) const {
        gather_dists_low<DP, VMN_PREC_BF16, VMN_GATHER_TILES>(q, r.bf16, r.bf16_norms, ids, out,
                                                              cnt);
    }
};

// ---------------------------------------------------------------------------
// Robust prune of a sorted candidate list held in LDS.
//
// Candidates are taken in list order; a candidate that is still kept becomes
// p* and is added to the output, and every later kept candidate p' is dropped
// when alpha * d(p*, p') < d(node, p') (STRICT, parth/src/build_graph.cu:
// 241-258) or <= (vamana/include/kernels.cuh:329-353). The product alpha*d is
// one float rounding, compared without fusion. Stops once R are kept. The
// d(p*, p') of one round are computed by the whole wave (8 rows at a time).
// ---------------------------------------------------------------------------
template <int DP, bool STRICT, bool BF = false>
__device__ int robust_prune(const PruneRows& rows, const uint32_t* c_id,
                            const float* c_d, int cnt, float alpha, int R, uint8_t* s_kept,
                            int32_t* s_act_id, int32_t* s_act_pos, float* s_act_d,
                            int32_t* out_ids, float* out_d) {
    const int l = lane();
    for (int i = l; i < cnt; i += 64) s_kept[i] = 1;
    __syncthreads();
    int nk = 0, from = 0;
    while (nk < R) {
        int p = -1;
        for (int b = from; b < cnt; b += 64) {
            const int j = b + l;
            const uint64_t m = ballot(j < cnt && s_kept[j]);
            if (m) {
                p = b + __ffsll((unsigned long long)m) - 1;
                break;
            }
        }
        if (p < 0) break;
        const int32_t ps = (int32_t)(c_id[p] & VMN_ID_MASK);
        if (l == 0) {
            out_ids[nk] = ps;
            if (out_d) out_d[nk] = c_d[p];
        }
        ++nk;
        if (nk >= R) break;
        PruneQ<DP, BF> pq;
        pq.load(rows, ps);
        int nact = 0;
        for (int b = p + 1; b < cnt; b += 64) {
            const int j = b + l;
            const bool f = j < cnt && s_kept[j];
            const uint64_t m = ballot(f);
            const int k = nact + __popcll((unsigned long long)(m & lanemask_lt()));
            if (f) {
                s_act_id[k] = (int32_t)(c_id[j] & VMN_ID_MASK);
                s_act_pos[k] = j;
            }
            nact += __popcll((unsigned long long)m);
        }
        __syncthreads();
        pq.dists(rows, s_act_id, s_act_d, nact);
        __syncthreads();
        for (int k = l; k < nact; k += 64) {
            const int j = s_act_pos[k];
            const float lhs = alpha * s_act_d[k];
            const bool pr = STRICT ? (lhs < c_d[j]) : (lhs <= c_d[j]);
            if (pr) s_kept[j] = 0;
        }
        __syncthreads();
        from = p + 1;
    }
    __syncthreads();
    return nk;
}

// Distances from `node` to its current row entries that pass `keep`, appended
// (compacted, adjacency order) to e_id/e_d. Returns the count.
template <int DP, bool BF>
__device__ __forceinline__ int row_extras(const PruneRows& rows, const PruneQ<DP, BF>& q,
                                          bool pass, int32_t nb, int32_t* e_id, float* e_d) {
    const uint64_t m = ballot(pass);
    const int ne = __popcll((unsigned long long)m);
    if (pass) e_id[__popcll((unsigned long long)(m & lanemask_lt()))] = nb;
    __syncthreads();
    q.dists(rows, e_id, e_d, ne);
    __syncthreads();
    return ne;
}

__device__ __forceinline__ int64_t next_item(unsigned long long* counter, int64_t* s_q) {
    if (lane() == 0) *s_q = (int64_t)atomicAdd(counter, 1ull);
    __syncthreads();
    const int64_t v = *s_q;
    __syncthreads();
    return v;
}

}  // namespace vmn
