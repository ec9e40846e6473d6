// BANG's exact search schedule (bang_exact/src/parANN.cu:96-978) over the
// index's graph, as one kernel: a wavefront owns one query and runs all of its
// rounds, with the query's bloom filter in LDS and its best-L list in
// registers (one entry per lane).
//
// The reference drives the rounds from the host: per round a memset, three
// kernels (neighbor_filtering_new :996-1052, compute_neighborDist_par
// :1118-1158, compute_BestLSets_par_sort_msort_new :1475-1669), a blocking copy
// of the per-query candidate counts (:594-599) and one of the continue flag
// (:652-654), over a 399,888-byte global bloom filter per query (:346). Here
// nothing leaves the device between rounds and a query ends as soon as it has
// no parent; its bloom filter is 399,887 bits (12,500 words, 50 KB) in LDS,
// cleared at the query's start, so three queries share a CU. The cells of ids
// 0 .. 32,791 are all distinct (the first two ids with one cell are 16,573 and
// 32,792), so on a graph of at most 32,792 nodes the filter is exactly the set
// of ids seen: there it is a bitmap over ids, n bits (1.25 KB for SIFT10K),
// and the CU holds as many queries as registers allow.
//
// Per round (semantics of the reference, order fixed where it is racy there):
//  * candidates: in round 1 the entry point, then the parent's neighbours in
//    adjacency order (up to the first -1) whose bloom cell hashFn1_d
//    (:1055-1064) is clear; their cells are set. Neighbours of one parent that
//    share a cell: the first in adjacency order is kept (the reference's
//    threads race there). Tombstoned rows are never candidates.
//  * distance: FP32, as compute_neighborDist_par sums it: eight fused
//    multiply-add chains over dimensions s, s + 8, ... (s = 0..7) combined as
//    ((p0+p1)+(p2+p3))+((p4+p5)+(p6+p7)) (cub::WarpReduce<float, 8>): 8 lanes
//    per candidate, lane s loading and summing chain s, the tree in three DPP
//    adds, several candidates' loads in flight per lane group.
//  * stable sort by distance, ties in arrival order; round 1 takes the first
//    min(count, L) as the list (entry visited); later rounds merge the first
//    nb candidates, nb = max(#leading candidates below the list's last
//    distance, min(L - B, count)), new entries first on ties, cut at L;
//  * the first unvisited list entry is the next parent (known before the
//    merge lays the list out, so its adjacency row loads meanwhile); a round
//    without candidates or parent ends the query, and so does round 4L + 19
//    (MAX_PARENTS_PERQUERY - 1, :49-50, :665-668).
// Outputs: the first k list entries (-1 / +inf past the list), and per query
// the number of distance computations and of rounds run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "vmn_device.hpp"
#include "vmn_internal.hpp"
#include "vmn_kernels.hpp"

using namespace vmn_rt;

namespace {

constexpr uint32_t kCells = 399887u;  // BF_ENTRIES, parANN.cu:36
constexpr int kWords = 12500;         // >= ceil(kCells / 32), a multiple of 4
constexpr int64_t kInjective = 32792;  // ids below this have distinct cells

struct BangArgs {
    const float* vecs;       // [n x DP]
    const int32_t* adj;      // [n x R], -1 padded
    const uint32_t* del;     // tombstone bits or nullptr
    const float* q;          // [nq x DP]
    int64_t nq;
    int R, L, k, max_rounds;
    int32_t entry;
    int words;               // filter words in LDS (a multiple of 4)
    int by_id;               // filter indexed by id (n <= kInjective)
    int32_t* out_ids;        // [nq x k]
    float* out_d;            // [nq x k] or nullptr
    int32_t* out_stats;      // [nq x 2] or nullptr
};

__device__ __forceinline__ uint32_t bang_cell(uint32_t x) {
    uint64_t h = 0xcbf29ce4ull;
    h = (h ^ (x & 0xffu)) * 0x01000193ull;
    h = (h ^ ((x >> 8) & 0xffu)) * 0x01000193ull;
    h = (h ^ ((x >> 16) & 0xffu)) * 0x01000193ull;
    h = (h ^ ((x >> 24) & 0xffu)) * 0x01000193ull;
    return (uint32_t)(h % kCells);
}

// The filter bit of id x: its cell, or x itself where cells are distinct.
__device__ __forceinline__ uint32_t filter_bit(uint32_t x, int by_id) {
    return by_id ? x : bang_cell(x);
}

// Squared L2 of one row against the query in LDS, summed in the reference's
// order (8 strided chains, then the pairwise tree). Padded dimensions are
// zero in both and leave every chain unchanged.
template <int DP>
__device__ __forceinline__ float bang_dist(const float* __restrict__ row, const float* qs) {
    const float4* r = reinterpret_cast<const float4*>(row);
    const float4* qv = reinterpret_cast<const float4*>(qs);
    float p0 = 0.f, p1 = 0.f, p2 = 0.f, p3 = 0.f, p4 = 0.f, p5 = 0.f, p6 = 0.f, p7 = 0.f;
#pragma unroll
    for (int j = 0; j < DP / 8; ++j) {
        const float4 a = r[2 * j], b = r[2 * j + 1];
        const float4 qa = qv[2 * j], qb = qv[2 * j + 1];
        float t;
        t = a.x - qa.x; p0 = __fmaf_rn(t, t, p0);
        t = a.y - qa.y; p1 = __fmaf_rn(t, t, p1);
        t = a.z - qa.z; p2 = __fmaf_rn(t, t, p2);
        t = a.w - qa.w; p3 = __fmaf_rn(t, t, p3);
        t = b.x - qb.x; p4 = __fmaf_rn(t, t, p4);
        t = b.y - qb.y; p5 = __fmaf_rn(t, t, p5);
        t = b.z - qb.z; p6 = __fmaf_rn(t, t, p6);
        t = b.w - qb.w; p7 = __fmaf_rn(t, t, p7);
    }
    return ((p0 + p1) + (p2 + p3)) + ((p4 + p5) + (p6 + p7));
}

// Element i of a list held one entry per lane (entry i in lane i % 64, slot
// i / 64); i is uniform.
template <int PER, class T>
__device__ __forceinline__ T list_at(const T (&v)[PER], int i) {
    T r{};
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (k == i / 64) {
            if constexpr (sizeof(T) == 4 && !__is_same(T, float))
                r = (T)vmn::readlane_i((int)v[k], i % 64);
            else
                r = vmn::readlane_f(v[k], i % 64);
        }
    }
    return r;
}

// A round's candidates are read in arrival order, 8 lanes each: lane s loads
// dimensions s, s + 8, ... of its candidate's row and sums chain s, so a load
// instruction covers 8 rows in 32-byte runs (one lane loading a whole row
// touches 64 rows per instruction and was a third slower once many waves
// gather at once). NP candidates' loads per lane are in flight: 8 (all of a
// round's at most 64 candidates, one memory round trip) with up to 256
// registers at 2 waves per SIMD, or 4 with up to 168 at 3 waves per SIMD for
// batches that fill the GPU several times over.
template <int DP, int LCAP, int NP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, NP == 8 ? 2 : 3))) void
bang_kernel(BangArgs a) {
    extern __shared__ __align__(16) uint32_t bloom[];  // a.words words
    __shared__ __align__(16) float qs[DP];
    // staging: the sorted candidates (at most R + 1 = 65), then the merged list
    constexpr int SZ = LCAP > 72 ? LCAP : 72;
    __shared__ int32_t sid[SZ];
    __shared__ float sd[SZ];
    __shared__ uint8_t sv[SZ];
    constexpr int PER = LCAP / 64;
    // the best-L list in registers: entry i in lane i % 64, slot i / 64
    int32_t lid[PER];
    float ld[PER];
    bool lv[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        lid[k] = -1;
        ld[k] = INFINITY;
        lv[k] = true;
    }

    const int l = vmn::lane();
    const int64_t qi = blockIdx.x;
    for (int i = l; i < a.words / 4; i += 64)
        reinterpret_cast<uint4*>(bloom)[i] = make_uint4(0u, 0u, 0u, 0u);
    for (int i = l; i < DP; i += 64) qs[i] = a.q[qi * DP + i];
    __syncthreads();

    // lane s = l & 7 of each 8-lane group sums chain s: dims s, s + 8, ...
    const int gs = l & 7, gg = l >> 3;
    float qv[DP / 8];
#pragma unroll
    for (int j = 0; j < DP / 8; ++j) qv[j] = qs[8 * j + gs];
    __shared__ int32_t cid[64];  // the round's candidates in arrival order
    __shared__ float cdist[64];
    const int32_t entry = a.entry;
    int B = 0, comps = 0, rounds = 0;
    // the parent's adjacency row: loaded as soon as the parent is known,
    // before the merge that lays out the new list
    int32_t row = l < a.R ? a.adj[(size_t)entry * a.R + l] : -1;
    for (int it = 1; it <= a.max_rounds; ++it) {
        rounds = it;
        const bool first = it == 1;
        float ed = 0.f;
        if (first) {
            const uint32_t c = filter_bit((uint32_t)entry, a.by_id);
            if (l == 0) bloom[c >> 5] |= 1u << (c & 31);
            ed = bang_dist<DP>(a.vecs + (size_t)entry * DP, qs);
            __syncthreads();
        }
        // neighbour filtering (neighbor_filtering_new)
        int32_t nbr = row;
        const uint64_t ends = vmn::ballot(nbr < 0);
        const int deg = ends ? __ffsll((unsigned long long)ends) - 1 : 64;
        bool want = l < deg;
        if (want && a.del && vmn::bit_test(a.del, (uint32_t)nbr)) want = false;
        const uint32_t c = want ? filter_bit((uint32_t)nbr, a.by_id) : 0u;
        const uint32_t w = c >> 5, bit = 1u << (c & 31);
        if (want) want = !(bloom[w] & bit);
        uint32_t old = 0u;
        if (want) old = atomicOr(&bloom[w], bit);
        if (vmn::ballot(want && (old & bit))) {
            // two neighbours with one cell: keep the first in adjacency order
            bool later = false;
            for (uint64_t m = vmn::ballot(want); m; m &= m - 1) {
                const int t = __ffsll((unsigned long long)m) - 1;
                later |= t < l && (uint32_t)vmn::readlane_i((int)c, t) == c;
            }
            want = want && !later;
        }
        uint64_t km = vmn::ballot(want);
        const int nn = __popcll((unsigned long long)km);
        const int nc = nn + (first ? 1 : 0);
        comps += nc;
        if (nc == 0) break;

        // distances (compute_neighborDist_par), 8 lanes per candidate
        if (want) cid[__popcll((unsigned long long)(km & vmn::lanemask_lt()))] = nbr;
        __syncthreads();
        for (int p0 = 0; p0 < nn; p0 += 8 * NP) {
            float x[NP][DP / 8];
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                if (p0 + 8 * p < nn) {
                    const int cc = p0 + 8 * p + gg;
                    const float* r = a.vecs + (size_t)cid[cc < nn ? cc : 0] * DP + gs;
#pragma unroll
                    for (int j = 0; j < DP / 8; ++j) x[p][j] = r[8 * j];
                }
            }
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                if (p0 + 8 * p < nn) {
                    float acc = 0.f;
#pragma unroll
                    for (int j = 0; j < DP / 8; ++j) {
                        const float t = x[p][j] - qv[j];
                        acc = __fmaf_rn(t, t, acc);
                    }
                    // ((p0+p1)+(p2+p3))+((p4+p5)+(p6+p7)) in every lane of the group
                    acc = vmn::reduce8(acc);
                    const int cc = p0 + 8 * p + gg;
                    if (gs == 0 && cc < nn) cdist[cc] = acc;
                }
            }
        }
        __syncthreads();
        // one candidate per lane again, in arrival order
        want = l < nn;
        nbr = want ? cid[l] : -1;
        const float d = want ? cdist[l] : 0.f;
        km = vmn::ballot(want);

        // stable sort by (distance, arrival): arrival is the entry, then lane
        // order. After round 1 only the candidates that enter the list need a
        // rank: while the list is full (nb = #candidates below its last
        // distance, at most min(L, count)) those are the ones below that
        // distance, whose ranks among themselves are their ranks among all.
        int nb = 0;
        uint64_t rm = km;
        if (!first) {
            const float mx = list_at<PER>(ld, B - 1);
            const uint64_t below = vmn::ballot(want && d < mx);
            const int nb0 = min(__popcll((unsigned long long)below), min(a.L, nc));
            nb = max(nb0, min(a.L - B, nc));
            if (nb == nb0) rm = below;
        }
        int rank = first && ed <= d ? 1 : 0;
        for (uint64_t m = rm; m; m &= m - 1) {
            const int t = __ffsll((unsigned long long)m) - 1;
            const float dt = vmn::readlane_f(d, t);
            rank += dt < d || (dt == d && t < l);
        }
        if (((rm >> l) & 1) && (first || rank < nb)) {
            sid[rank] = nbr;
            sd[rank] = d;
        }
        if (first) {
            int erank = 0;
            for (uint64_t m = km; m; m &= m - 1)
                erank += vmn::readlane_f(d, __ffsll((unsigned long long)m) - 1) < ed;
            if (l == 0) {
                sid[erank] = entry;
                sd[erank] = ed;
            }
        }
        __syncthreads();

        int32_t next = -1;
        if (first) {
            // round 1: the first min(count, L) sorted candidates, the entry visited
            B = min(nc, a.L);
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int i = l + 64 * k;
                if (i < B) {
                    lid[k] = sid[i];
                    ld[k] = sd[i];
                    lv[k] = sid[i] == entry;
                }
            }
            // next parent: the first unvisited entry
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const uint64_t m = vmn::ballot(l + 64 * k < B && !lv[k]);
                if (m && next < 0) {
                    const int t = __ffsll((unsigned long long)m) - 1;
                    next = vmn::readlane_i(lid[k], t);
                    if (l == t) lv[k] = true;
                }
            }
            if (next >= 0) row = l < a.R ? a.adj[(size_t)next * a.R + l] : -1;
        } else {
            // best-L merge (compute_BestLSets_par_sort_msort_new): the nb sorted
            // candidates, new entries before list entries on ties, cut at L
            float xs = 0.f;
            int32_t xid = 0;
            if (l < nb) {
                xs = sd[l];
                xid = sid[l];
            }
            // The next parent is the merged list's first unvisited entry: the
            // closest new entry if it is at or below the first unvisited list
            // entry u (new entries come first on ties), else u. Its adjacency
            // row is loaded now, so the load overlaps the merge.
            int u = -1;
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const uint64_t m = vmn::ballot(l + 64 * k < B && !lv[k]);
                if (m && u < 0) u = 64 * k + __ffsll((unsigned long long)m) - 1;
            }
            const bool take_new =
                nb > 0 && (u < 0 || vmn::readlane_f(xs, 0) <= list_at<PER>(ld, u));
            next = take_new ? vmn::readlane_i(xid, 0) : (u >= 0 ? list_at<PER>(lid, u) : -1);
            if (next >= 0) row = l < a.R ? a.adj[(size_t)next * a.R + l] : -1;

            // positions: new entry t at #(list entries below it) + t, list
            // entry j at j + #(new entries at or below it)
            const int nB = min(B + nb, a.L);
            int np = l;
            int cle[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) cle[k] = 0;
            for (int t = 0; t < nb; ++t) {
                const float xt = vmn::readlane_f(xs, t);
                int less = 0;
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    less += __popcll((unsigned long long)vmn::ballot(l + 64 * k < B && ld[k] < xt));
                    cle[k] += xt <= ld[k];
                }
                np = vmn::writelane_i(less + t, t, np);
            }
            __syncthreads();
            if (l < nb && np < nB) {
                sid[np] = xid;
                sd[np] = xs;
                sv[np] = take_new && l == 0;
            }
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int j = l + 64 * k;
                const int p = j + cle[k];
                if (j < B && p < nB) {
                    sid[p] = lid[k];
                    sd[p] = ld[k];
                    sv[p] = lv[k] || (!take_new && j == u);
                }
            }
            __syncthreads();
            B = nB;
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int i = l + 64 * k;
                if (i < B) {
                    lid[k] = sid[i];
                    ld[k] = sd[i];
                    lv[k] = sv[i];
                }
            }
        }
        __syncthreads();
        if (next < 0) break;
    }

#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int j = l + 64 * k;
        if (j < a.k) {
            a.out_ids[qi * a.k + j] = j < B ? lid[k] : -1;
            if (a.out_d) a.out_d[qi * a.k + j] = j < B ? ld[k] : INFINITY;
        }
    }
    if (a.out_stats && l == 0) {
        a.out_stats[qi * 2] = comps;
        a.out_stats[qi * 2 + 1] = rounds;
    }
}

template <int DP, int NP>
hipError_t launch_dp(const BangArgs& a, hipStream_t s) {
    const dim3 grid((unsigned)a.nq), block(64);
    const size_t lds = (size_t)a.words * sizeof(uint32_t);
    if (a.L <= 64)
        hipLaunchKernelGGL((bang_kernel<DP, 64, NP>), grid, block, lds, s, a);
    else
        hipLaunchKernelGGL((bang_kernel<DP, 256, NP>), grid, block, lds, s, a);
    return hipGetLastError();
}

template <int NP>
hipError_t launch_np(const BangArgs& a, int DP, hipStream_t s) {
    switch (DP) {
        case 32: return launch_dp<32, NP>(a, s);
        case 64: return launch_dp<64, NP>(a, s);
        case 96: return launch_dp<96, NP>(a, s);
        case 128: return launch_dp<128, NP>(a, s);
        case 160: return launch_dp<160, NP>(a, s);
        case 192: return launch_dp<192, NP>(a, s);
        case 256: return launch_dp<256, NP>(a, s);
        default: return hipErrorInvalidValue;
    }
}

// All of a round's loads in flight up to 16 queries per CU (4,096 on 256 CUs:
// 0.273 against 0.288 ms), 4 candidates' loads at 3 waves per SIMD beyond
// (10,000 queries: 0.560 against 0.634 ms)
hipError_t launch_bang(const BangArgs& a, int DP, int cus, hipStream_t s) {
    return a.nq <= 16 * (int64_t)cus ? launch_np<8>(a, DP, s) : launch_np<4>(a, DP, s);
}

}  // namespace

extern "C" int vmn_bang_search(vmn_index* idx, const float* queries, int64_t nq, int32_t k,
                               int32_t L, int64_t entry, int32_t* ids_out, float* dists_out,
                               int32_t* stats_out, int32_t mem) {
    if (!idx || !queries || !ids_out) return fail(VMN_ERR_ARG, "NULL argument");
    if (!idx->built) return fail(VMN_ERR_STATE, "index has no graph");
    if (nq < 0 || nq > 0x7fffffff) return fail(VMN_ERR_ARG, "nq = %lld out of range", (long long)nq);
    if (k < 1 || k > L || L > 256)
        return fail(VMN_ERR_ARG, "need 1 <= k <= L <= 256 (k=%d, L=%d)", k, L);
    if (mem != VMN_MEM_HOST && mem != VMN_MEM_DEVICE) return fail(VMN_ERR_ARG, "bad mem %d", mem);
    if (entry < 0) entry = idx->medoid;
    if (entry < 0 || entry >= idx->n)
        return fail(VMN_ERR_ARG, "entry point %lld not in [0, %lld)", (long long)entry,
                    (long long)idx->n);
    if (nq == 0) return VMN_OK;
    int rc = set_device(idx);
    if (rc) return rc;
    const float* q = nullptr;
    rc = stage_rows(idx, queries, nq, mem, idx->qpad, &q);
    if (rc) return rc;
    int32_t* oi = ids_out;
    float* od = dists_out;
    int32_t* os = stats_out;
    if (mem == VMN_MEM_HOST) {
        HIPCHK(idx->out_ids.reserve((size_t)nq * k * sizeof(int32_t)));
        HIPCHK(idx->out_d.reserve((size_t)nq * k * sizeof(float)));
        HIPCHK(idx->bang_stats.reserve((size_t)nq * 2 * sizeof(int32_t)));
        oi = idx->out_ids.as<int32_t>();
        od = dists_out ? idx->out_d.as<float>() : nullptr;
        os = stats_out ? idx->bang_stats.as<int32_t>() : nullptr;
    }
    BangArgs a{};
    a.vecs = idx->vecs.as<float>();
    a.adj = idx->adj.as<int32_t>();
    a.del = idx->n_deleted ? idx->del_bits.as<uint32_t>() : nullptr;
    a.q = q;
    a.nq = nq;
    a.R = idx->p.R;
    a.L = L;
    a.k = k;
    a.max_rounds = 4 * L + 19;
    a.entry = (int32_t)entry;
    a.by_id = idx->n <= kInjective;
    a.words = a.by_id ? (int)((idx->n + 127) / 128 * 4) : kWords;
    if (a.by_id) {
        // A batch smaller than the CUs can hold at once: pad each query's LDS
        // so that a CU takes at most ceil(nq / CUs) of them (160 KB of LDS per
        // CU, about 4 KB of it the kernel's own arrays) and the queries spread
        // over all CUs instead of sharing SIMDs on a few.
        const int64_t per_cu = std::max<int64_t>(1, (nq + idx->cus - 1) / std::max(1, idx->cus));
        if (per_cu < 12) {
            const int pad = (int)std::min<int64_t>(kWords, ((163840 / per_cu) - 4096) / 4 / 4 * 4);
            a.words = std::max(a.words, pad);
        }
    }
    a.out_ids = oi;
    a.out_d = od;
    a.out_stats = os;
    // one workgroup per query; a dispatch's grid counts work-items in 32 bits
    constexpr int64_t kChunk = (int64_t)1 << 24;
    for (int64_t q0 = 0; q0 < nq; q0 += kChunk) {
        BangArgs c = a;
        c.nq = std::min(kChunk, nq - q0);
        c.q = q + q0 * idx->DP;
        c.out_ids = oi + q0 * k;
        c.out_d = od ? od + q0 * k : nullptr;
        c.out_stats = os ? os + q0 * 2 : nullptr;
        HIPCHK(launch_bang(c, idx->DP, idx->cus, idx->stream));
    }
    if (mem == VMN_MEM_HOST) {
        HIPCHK(hipMemcpyAsync(ids_out, oi, (size_t)nq * k * sizeof(int32_t), hipMemcpyDeviceToHost,
                              idx->stream));
        if (dists_out)
            HIPCHK(hipMemcpyAsync(dists_out, od, (size_t)nq * k * sizeof(float),
                                  hipMemcpyDeviceToHost, idx->stream));
        if (stats_out)
            HIPCHK(hipMemcpyAsync(stats_out, os, (size_t)nq * 2 * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, idx->stream));
        HIPCHK(hipStreamSynchronize(idx->stream));
    }
    return VMN_OK;
}
