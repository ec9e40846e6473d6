// Batched greedy beam search, one wavefront per query, persistent launch.
//
// Two traversal semantics share one kernel body:
//
// PARTH (parth/src/engine.cu:29-87 `greedy_search`, build_graph.cu:68-119
//   `greedy_search_build`, CandidateList common.h:89-155). The reference inserts
//   each neighbour of the expanded node one at a time into a sorted list of at
//   most L entries, skipping ids already in the list, with equal distances
//   placed after existing entries and rejected when the list is full. That
//   sequence equals: keep the L first entries of every id ever offered,
//   ordered by (distance, time first offered). An id that left the list can
//   never re-enter it (the list's last distance only falls), so a per-entry
//   "expanded" flag replaces the reference's visited list, and one hop is:
//   drop neighbours already in the list, compute the rest, and stable-merge
//   them (list entries first on ties, new ones in adjacency order). Search
//   stops when no unexpanded entry is left, after `visited_cap` expansions
//   (reference VISITED_LIST_SIZE = 128) or after `max_hops` iterations
//   (reference 200).
//
// FRESH (vamana/include/search.cuh:20-223, G2). New neighbours are those not
//   yet in the per-query bloom filter (bloom_filter.cuh hashes, one bit per
//   cell in LDS); they are sorted by distance and merged into the worklist
//   (worklist entries first on ties), truncated to L; the closest unexpanded
//   entry is expanded next, until none is left. Decisions where the reference
//   is racy or self-inconsistent are fixed as follows: every neighbour of one
//   expanded node is tested against the filter state from before that node
//   (what one warp of the reference's filterNeighbors does), the start node
//   is entered into the filter (the reference omits it and can list the
//   medoid twice), and the query's own node id is skipped only when the query
//   is a graph node (the reference compares against the batch index,
//   search.cuh:126). Expanded nodes (the "visited set", search.cuh:104-111,
//   capped at P) are written out with their distances for the prune.
//
// Both modes skip tombstoned nodes while expanding (search.cuh:122,
// engine.cu:76) and give a tombstoned start node distance 1e30 (engine.cu:45).
//
// PREC selects the distance used while traversing: FP32 rows, or BF16 / FP8
// rows on MFMA (vmn_device.hpp). FP32 distances are sequential fused
// multiply-add chains over the dimensions in PARTH mode (the reference's
// distance<D>, parth/src/device_funcs.cu:17-49, so real-valued rows give the
// reference's results) and lane-tree sums in FRESH mode (the reference sums
// 8 strided partials with float atomics, kernels.cuh:40-52, in no fixed
// order). With krr > 0 the first krr entries of the final list are re-ranked
// by exact FP32 distance (stable on ties) before the k results are written;
// this is the re-rank step the reference runs on the worklist (utils.cuh:
// 199-285, merge_and_rerank_kernel), done here on ids inside the search
// kernel.
#include "vmn_kernels.hpp"

namespace vmn {

__global__ void fill_zero_u32(unsigned int* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// PREC_ARG is a VMN_PREC_* value, or VMN_PREC_BF16_TABLE_NORMS: BF16 rows with
// the row norms read from the norms table (small batches, FRESH mode).
// FRESH mode at row widths up to 128 with beam lists up to 256 entries, and
// PARTH-mode FP8 at width 128 with lists up to 128, are held to 5 waves per
// SIMD (96 VGPRs; the compiler's own choice takes 4): large batches are
// bandwidth-bound and run faster with the extra wave (FP8 10k-query batch
// 7%), single queries about 3% slower. PARTH FP32 / BF16, and PARTH FP8 at
// widths 64 and 96, would spill there.
// NG is the low-precision forms' tile shape: MFMA tiles of 16 candidates
// gathered per round. 2 (32 candidates) is every form's default; 4 (64
// candidates, held to 4 waves per SIMD, 128 VGPRs) fetches a whole hop's new
// candidates in one round (half of all hops need two rounds of 32), which
// shortens the latency of small batches; instantiated for FRESH BF16 / FP8
// at row widths 64 to 128 with lists up to 256 (the autotuner's choice).
template <int DP, int LCAP, int MODE, int PREC_ARG, int NG>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(
    NG == 4 ? 4
            : (DP <= 128 && ((MODE == VMN_MODE_FRESH && LCAP <= 256) ||
                             (PREC_ARG == VMN_PREC_FP8 && LCAP <= 128 && DP == 128))) ? 5 : 1, 8))) void search_kernel(SearchArgs a) {
    constexpr int PREC = PREC_ARG == VMN_PREC_BF16_TABLE_NORMS ? VMN_PREC_BF16 : PREC_ARG;
    constexpr bool TN = PREC_ARG == VMN_PREC_BF16_TABLE_NORMS;
    // FP32 distances in the reference's summation order (PARTH)
    constexpr bool SEQ = MODE == VMN_MODE_PARTH;
    // FRESH bloom cells of the neighbours: read from the table beside the
    // adjacency by the FP8 form (its hops are short of ALU time), hashed in
    // registers by the FP32 and BF16 forms (short of HBM bandwidth: the
    // table's 4 bytes per neighbour cost them more than the hashing)
    constexpr bool HTAB = MODE == VMN_MODE_FRESH && PREC == VMN_PREC_FP8;
    // FRESH lists of up to 128 entries live in registers, entry i in lane
    // i % 64, slot i / 64: the select reads them with ballots, the merge
    // scatters through LDS once (longer lists, and the 4-tile table-norm
    // form at 128, would spill)
    constexpr bool RL = MODE == VMN_MODE_FRESH &&
                        (LCAP == 64 || (LCAP == 128 && !(TN && NG == 4)));
    constexpr int RPER = RL ? LCAP / 64 : 1;
    __shared__ __align__(16) uint32_t s_id[2][LCAP];
    __shared__ __align__(16) float s_d[2][LCAP];
    __shared__ __align__(16) int32_t s_sid[64];
    __shared__ float s_sd[64];
    // FP8: the candidates' row norms, from the adjacency-aligned norms
    __shared__ float s_sn[PREC_ARG == VMN_PREC_FP8 ? 64 : 1];
    __shared__ uint32_t s_bloom[MODE == VMN_MODE_FRESH ? VMN_BF_WORDS : 1];
    __shared__ int64_t s_q;

    const int l = threadIdx.x, sub = l & 7;
    const float* __restrict__ vecs = a.vecs;
    const int W = a.W;

    // query groups: this workgroup's cyclic walk over the lists (lane 0), from
    // its group's first; lists only empty during a launch, so the walk only
    // moves forward
    const int nlists = a.ngroups > 0 ? a.nlists : 0;
    const int fstart = a.ngroups > 0
                           ? (int)(((int64_t)(blockIdx.x % (unsigned)a.ngroups) * nlists + a.ngroups - 1) / a.ngroups)
                           : 0;
    int ft = 0;
    unsigned long long fcnt = 0;  // entries of the list at step ft, once read
    bool fknown = false;

    for (;;) {
        if (l == 0) {
            int64_t qsel = a.nq;
            if (a.ngroups > 0) {
                for (; ft < nlists; ++ft, fknown = false) {
                    const int L = fstart + ft < nlists ? fstart + ft : fstart + ft - nlists;
                    unsigned long long* h = a.counter + 32 * L;
                    if (!fknown) {
                        // first visit: an emptied list is passed over without
                        // a read-modify-write
                        fcnt = __hip_atomic_load(h + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (__hip_atomic_load(h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= fcnt)
                            continue;
                        fknown = true;
                    }
                    const unsigned long long c = atomicAdd(h, 1ull);
                    if (c < fcnt) {
                        qsel = a.order[(int64_t)L * a.gstride + (int64_t)c];
                        break;
                    }
                }
            } else {
                qsel = (int64_t)atomicAdd(a.counter, 1ull);
            }
            s_q = qsel;
        }
        __syncthreads();
        const int64_t qi = s_q;
        __syncthreads();
        if (qi >= a.nq) return;

        // The FP32 query lives in registers through the walk only for FP32
        // lane-tree traversal; the low-precision forms load it for the
        // re-rank. PARTH-mode FP32 distances are sequential chains read from
        // the query row itself.
        const float* qrow = a.queries + qi * DP;
        QReg<DP> q;
        if (PREC == VMN_PREC_FP32 && !SEQ) q.load(qrow, sub);
        QLow<DP> ql;
        if (PREC != VMN_PREC_FP32) prep_query<DP, PREC>(ql, a.queries + qi * DP, a.fp8_scales, a.fp8_centre,
                                 reinterpret_cast<uint8_t*>(s_sid));
        const int32_t self = a.query_is_node ? (int32_t)(a.node_base + qi) : -1;

        if (MODE == VMN_MODE_FRESH) {
            for (int w = l; w < (int)VMN_BF_WORDS; w += 64) s_bloom[w] = 0u;
        }
        const int32_t st = a.start;
        float sd;
        if (PREC == VMN_PREC_FP32 && !SEQ) {
            sd = reduce8(q.partial(vecs + (int64_t)st * DP, sub));
        } else if (PREC == VMN_PREC_FP32) {
            if (l == 0) s_sid[0] = st;
            __syncthreads();
            gather_dists_seq<DP>(qrow, vecs, s_sid, s_sd, 1);
            __syncthreads();
            sd = s_sd[0];
        } else {
            if (l == 0) s_sid[0] = st;
            __syncthreads();
            gather_dists_low<DP, PREC, 2, TN>(ql, a.low_rows, a.low_norms, s_sid, s_sd, 1);
            __syncthreads();
            sd = s_sd[0];
        }
        if (a.del_bits && bit_test(a.del_bits, (uint32_t)st)) sd = 1e30f;
        __syncthreads();
        if (l == 0) {
            s_id[0][0] = (uint32_t)st;
            s_d[0][0] = sd;
            if (MODE == VMN_MODE_FRESH) {
                uint32_t h1 = bf_hash1((uint32_t)st), h2 = bf_hash2((uint32_t)st);
                s_bloom[h1 >> 5] |= 1u << (h1 & 31);
                s_bloom[h2 >> 5] |= 1u << (h2 & 31);
            }
        }
        __syncthreads();

        uint32_t r_id[RPER];
        float r_d[RPER];
#pragma unroll
        for (int k = 0; k < RPER; ++k) {
            r_id[k] = k == 0 && l == 0 ? (uint32_t)st : VMN_VIS_BIT;
            r_d[k] = k == 0 && l == 0 ? sd : 0.0f;
        }
        int cur = 0, cnt = 1, nexp = 0, hops = 0, ndist = 1;
        // adjacency row of the node predicted to be expanded next, loaded
        // before the merge that decides it (pf: its id, -1 for none)
        int32_t pf = -1, pf_nb = -1;
        uint32_t pf_nh = 0u;
        float pf_nbn = 0.0f;
        while (hops < a.max_hops) {
            ++hops;
            // closest unexpanded entry (its id and distance taken from the
            // lane that read it), and the one after it (p2), which is the
            // closest once this one is expanded unless a new candidate beats it
            int p = -1, p2 = -1;
            uint32_t e = 0u, e2 = 0u;
            float nd = 0.0f, nd2 = 0.0f;
            auto scan = [&](int b, uint32_t v, float dv) {
                uint64_t m = ballot(!(v & VMN_VIS_BIT));
                if (m && p < 0) {
                    const int t = __ffsll((unsigned long long)m) - 1;
                    p = b + t;
                    e = (uint32_t)readlane_i((int)v, t);
                    nd = readlane_f(dv, t);
                    m &= m - 1;
                }
                if (m) {
                    const int t = __ffsll((unsigned long long)m) - 1;
                    p2 = b + t;
                    e2 = (uint32_t)readlane_i((int)v, t);
                    nd2 = readlane_f(dv, t);
                }
            };
            if constexpr (RL) {
#pragma unroll
                for (int k = 0; k < RPER; ++k) {
                    if (64 * k < cnt && p2 < 0) {
                        const bool in = 64 * k + l < cnt;
                        scan(64 * k, in ? r_id[k] : VMN_VIS_BIT, in ? r_d[k] : 0.0f);
                    }
                }
            } else {
                for (int b = 0; b < cnt && p2 < 0; b += 64) {
                    const int i = b + l;
                    scan(b, i < cnt ? s_id[cur][i] : VMN_VIS_BIT, i < cnt ? s_d[cur][i] : 0.0f);
                }
            }
            if (p < 0) break;
            if (MODE == VMN_MODE_PARTH && nexp >= a.visited_cap) break;
            ++nexp;
            const int32_t node = (int32_t)(e & VMN_ID_MASK);
            __syncthreads();
            if constexpr (RL) {
#pragma unroll
                for (int k = 0; k < RPER; ++k)
                    if (64 * k + l == p) r_id[k] |= VMN_VIS_BIT;
            }
            if (l == 0) {
                if (!RL) s_id[cur][p] = e | VMN_VIS_BIT;
                if (MODE == VMN_MODE_FRESH && a.vis_ids && nexp <= a.P) {
                    a.vis_ids[qi * a.P + nexp - 1] = node;
                    a.vis_d[qi * a.P + nexp - 1] = nd;
                }
            }
            // PARTH: the list's ids, LCAP / 64 per lane, for the membership test
            // below (read before the adjacency row arrives)
            constexpr int EPL = MODE == VMN_MODE_PARTH ? LCAP / 64 : 1;
            uint32_t ent[EPL];
            if (MODE == VMN_MODE_PARTH) {
#pragma unroll
                for (int k = 0; k < EPL; ++k) {
                    const int i = l + 64 * k;
                    ent[k] = i < cnt ? (s_id[cur][i] & VMN_ID_MASK) : 0xffffffffu;
                }
            }
            // neighbours up to the first -1 (FRESH: with their bloom cells;
            // FP8: with their row norms)
            int32_t nb;
            uint32_t nh = 0u;
            float nbn = 0.0f;
            if (node == pf) {
                nb = pf_nb;
                nh = pf_nh;
                nbn = pf_nbn;
            } else {
                nb = l < W ? a.adj[(int64_t)node * W + l] : -1;
                if constexpr (HTAB) nh = l < W ? a.adj_hash[(int64_t)node * W + l] : 0u;
                if constexpr (PREC == VMN_PREC_FP8)
                    nbn = l < W ? a.adj_norms[(int64_t)node * W + l] : 0.0f;
            }
            const uint64_t endm = ballot(l < W && nb == -1);
            const int deg = endm ? __ffsll((unsigned long long)endm) - 1 : W;
            bool valid = l < deg;
            // the tombstone word is consumed after the membership test, so its
            // load overlaps that work
            const bool gone = valid && a.del_bits && bit_test(a.del_bits, (uint32_t)nb);
            if (MODE == VMN_MODE_FRESH && valid && nb == self) valid = false;
            if (MODE == VMN_MODE_PARTH) {
                // skip ids already in the list. A short list is scanned from
                // LDS by every lane; a long one is matched the other way round:
                // neighbour j is broadcast and compared with every lane's list
                // entries in one ballot (deg steps instead of cnt / 4 loads).
                bool dup = false;
                if (cnt <= 32) {
                    const uint32_t key = (uint32_t)nb;
                    int i = 0;
                    for (; i + 4 <= cnt; i += 4) {
                        const uint4 v = *reinterpret_cast<const uint4*>(&s_id[cur][i]);
                        dup |= ((v.x & VMN_ID_MASK) == key) | ((v.y & VMN_ID_MASK) == key) |
                               ((v.z & VMN_ID_MASK) == key) | ((v.w & VMN_ID_MASK) == key);
                    }
                    for (; i < cnt; ++i) dup |= (s_id[cur][i] & VMN_ID_MASK) == key;
                } else {
                    uint64_t dupm = 0;
                    for (int j = 0; j < deg; ++j) {
                        const uint32_t c = (uint32_t)__builtin_amdgcn_readlane(nb, j);
                        bool hit = false;
#pragma unroll
                        for (int k = 0; k < EPL; ++k) hit |= ent[k] == c;
                        if (ballot(hit)) dupm |= 1ull << j;
                    }
                    dup = (dupm >> l) & 1ull;
                }
                valid = valid && !dup && !gone;
                __syncthreads();
            } else {
                valid = valid && !gone;
                __syncthreads();
                if constexpr (!HTAB) nh = valid ? bf_hash_pair((uint32_t)nb) : 0u;
                const uint32_t h1 = nh & 0xffffu, h2 = nh >> 16;
                bool seen = true;
                // both cells read at once (no short circuit)
                if (valid) seen = ((int)bit_test(s_bloom, h1) & (int)bit_test(s_bloom, h2)) != 0;
                __syncthreads();
                valid = valid && !seen;
                if (valid) {
                    atomicOr(&s_bloom[h1 >> 5], 1u << (h1 & 31));
                    atomicOr(&s_bloom[h2 >> 5], 1u << (h2 & 31));
                }
            }
            const uint64_t vm = ballot(valid);
            const int nnew = __popcll((unsigned long long)vm);
            const int slot = __popcll((unsigned long long)(vm & lanemask_lt()));
            if (valid) {
                s_sid[slot] = nb;
                if constexpr (PREC == VMN_PREC_FP8) s_sn[slot] = nbn;
            }
            __syncthreads();
            if (PREC == VMN_PREC_FP32 && SEQ)
                gather_dists_seq<DP>(qrow, vecs, s_sid, s_sd, nnew);
            else if (PREC == VMN_PREC_FP32)
                gather_dists<DP>(q, vecs, s_sid, s_sd, nnew);
            else
                gather_dists_low<DP, PREC, NG, TN>(
                    ql, a.low_rows, a.low_norms, s_sid, s_sd, nnew,
                    PREC == VMN_PREC_FP8 ? (const float*)s_sn : nullptr);
            __syncthreads();
            ndist += nnew;
            // A full list keeps its entries on ties, so a candidate at or beyond
            // its last distance can never enter: drop those before ranking.
            const float thr =
                cnt == a.L ? (RL ? reg_at<RPER>(r_d, cnt - 1) : s_d[cur][cnt - 1]) : INFINITY;
            const uint32_t my_id = l < nnew ? (uint32_t)s_sid[l] : 0u;
            const float my_d = l < nnew ? s_sd[l] : 0.0f;
            const int my_k = fkey(my_d);
            const uint64_t km = ballot(l < nnew && my_d < thr);
            // The merged list's closest unexpanded entry is the closest kept
            // candidate (lowest slot on ties) if it is below p2's distance
            // (list entries come first on ties), else p2: start loading its
            // adjacency row now, so the load overlaps the merge.
            {
                const uint64_t bm = km & ballot(p2 < 0 || my_d < nd2);
                pf = -1;
                if (bm) {
                    const int mn = wave_min_key(((bm >> l) & 1ull) ? my_k : 0x7fffffff);
                    const uint64_t mm = bm & ballot(my_k == mn);
                    if (mm) pf = readlane_i((int)my_id, __ffsll((unsigned long long)mm) - 1);
                } else if (p2 >= 0) {
                    pf = (int32_t)(e2 & VMN_ID_MASK);
                }
                if (pf >= 0) {
                    pf_nb = l < W ? a.adj[(int64_t)pf * W + l] : -1;
                    if constexpr (HTAB) pf_nh = l < W ? a.adj_hash[(int64_t)pf * W + l] : 0u;
                    if constexpr (PREC == VMN_PREC_FP8)
                        pf_nbn = l < W ? a.adj_norms[(int64_t)pf * W + l] : 0.0f;
                }
            }
            if constexpr (RL) {
                cnt = merge_regs<RPER>(r_id, r_d, cnt, my_id, my_d, my_k, km, s_id[0], s_d[0], a.L);
            } else {
                cnt = merge_masked<LCAP>(s_id[cur], s_d[cur], cnt, my_id, my_d, my_k, km,
                                         s_id[cur ^ 1], s_d[cur ^ 1], a.L);
                cur ^= 1;
            }
            __syncthreads();
        }
        if constexpr (RL) {
            // the list back to LDS for the re-rank and the output
#pragma unroll
            for (int k = 0; k < RPER; ++k) {
                if (64 * k + l < cnt) {
                    s_id[0][64 * k + l] = r_id[k];
                    s_d[0][64 * k + l] = r_d[k];
                }
            }
            __syncthreads();
        }

        if (PREC != VMN_PREC_FP32 && a.krr > 0) {
            // exact FP32 re-rank of the first krr entries, stable on ties
            const int nr = cnt < a.krr ? cnt : a.krr;
            if (l < nr) s_sid[l] = (int32_t)(s_id[cur][l] & VMN_ID_MASK);
            if (!SEQ) q.load(qrow, sub);
            __syncthreads();
            if (SEQ) gather_dists_seq<DP>(qrow, vecs, s_sid, s_sd, nr);
            else gather_dists<DP>(q, vecs, s_sid, s_sd, nr);
            __syncthreads();
            const uint32_t rid = l < nr ? (uint32_t)s_sid[l] : 0u;
            const float rd = l < nr ? s_sd[l] : 0.0f;
            merge_sorted<LCAP, false>(s_id[cur], s_d[cur], 0, rid, rd, nr, s_id[cur ^ 1],
                                      s_d[cur ^ 1], nr);
            __syncthreads();
            if (l < nr) {
                s_id[cur][l] = s_id[cur ^ 1][l];
                s_d[cur][l] = s_d[cur ^ 1][l];
            }
            __syncthreads();
            if (cnt > nr) cnt = nr;  // entries past krr carry low-precision distances
        }
        if (a.out_ids) {
            for (int j = l; j < a.out_n; j += 64) {
                a.out_ids[qi * a.out_n + j] = j < cnt ? (int32_t)(s_id[cur][j] & VMN_ID_MASK) : -1;
                if (a.out_d) a.out_d[qi * a.out_n + j] = j < cnt ? s_d[cur][j] : 1e30f;
            }
        }
        if (l == 0) {
            if (a.vis_cnt) a.vis_cnt[qi] = nexp < a.P ? nexp : a.P;
            if (a.stat_hops) a.stat_hops[qi] = nexp;
            if (a.stat_ndist) a.stat_ndist[qi] = ndist;
            if (a.acc) {
                unsigned long long* c = acc_slot(a.acc);
                atomicAdd(&c[VMN_ACC_SEARCHES], 1ull);
                atomicAdd(&c[VMN_ACC_EXPANSIONS], (unsigned long long)nexp);
                atomicAdd(&c[VMN_ACC_DISTS], (unsigned long long)ndist);
            }
        }
        __syncthreads();
    }
}

// Kernel instances with 4 gather tiles exist for these shapes.
template <int DP, int LCAP, int MODE>
constexpr bool tiles4_shape() {
    return MODE == VMN_MODE_FRESH && DP >= 64 && DP <= 128 && LCAP <= 256;
}

bool search_tiles4(int DP, int need, int mode) {
    return mode == VMN_MODE_FRESH && DP >= 64 && DP <= 128 && need <= 256;
}

template <int DP, int LCAP, int MODE, int PREC_ARG>
static void launch_lowp(const SearchArgs& a, int grid, hipStream_t s) {
    if constexpr (tiles4_shape<DP, LCAP, MODE>()) {
        if (a.tiles == 4) {
            hipLaunchKernelGGL((search_kernel<DP, LCAP, MODE, PREC_ARG, 4>), dim3(grid), dim3(64), 0, s, a);
            return;
        }
    }
    hipLaunchKernelGGL((search_kernel<DP, LCAP, MODE, PREC_ARG, 2>), dim3(grid), dim3(64), 0, s, a);
}

template <int DP, int LCAP, int MODE>
static hipError_t launch_search_p(const SearchArgs& a, int grid, hipStream_t s) {
    switch (a.prec) {
        case VMN_PREC_FP32:
            hipLaunchKernelGGL((search_kernel<DP, LCAP, MODE, VMN_PREC_FP32, 2>), dim3(grid), dim3(64), 0, s, a);
            break;
        case VMN_PREC_BF16:
            if (MODE == VMN_MODE_FRESH && a.table_norms)
                launch_lowp<DP, LCAP, MODE, MODE == VMN_MODE_FRESH ? VMN_PREC_BF16_TABLE_NORMS : VMN_PREC_BF16>(a, grid, s);
            else
                launch_lowp<DP, LCAP, MODE, VMN_PREC_BF16>(a, grid, s);
            break;
        case VMN_PREC_FP8:
            launch_lowp<DP, LCAP, MODE, VMN_PREC_FP8>(a, grid, s);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int DP, int LCAP>
static hipError_t launch_search_l(const SearchArgs& a, int grid, hipStream_t s) {
    if (a.mode == VMN_MODE_PARTH) return launch_search_p<DP, LCAP, VMN_MODE_PARTH>(a, grid, s);
    return launch_search_p<DP, LCAP, VMN_MODE_FRESH>(a, grid, s);
}

template <int DP>
static hipError_t launch_search_d(const SearchArgs& a, int grid, hipStream_t s) {
    const int need = a.L > a.lcap ? a.L : a.lcap;
    if (need <= 64) return launch_search_l<DP, 64>(a, grid, s);
    if (need <= 128) return launch_search_l<DP, 128>(a, grid, s);
    if (need <= 256) return launch_search_l<DP, 256>(a, grid, s);
    return launch_search_l<DP, 512>(a, grid, s);
}

__global__ void adj_norms_kernel(const int32_t* __restrict__ adj, int64_t entries,
                                 const float* __restrict__ norms, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= entries) return;
    const int32_t v = adj[i];
    out[i] = v >= 0 ? norms[v] : 0.0f;
}
hipError_t launch_adj_norms(const int32_t* adj, int64_t entries, const float* norms, float* out,
                            hipStream_t s) {
    if (entries <= 0) return hipSuccess;
    hipLaunchKernelGGL(adj_norms_kernel, dim3((unsigned)((entries + 255) / 256)), dim3(256), 0, s,
                       adj, entries, norms, out);
    return hipGetLastError();
}

__global__ void adj_hash_kernel(const int32_t* __restrict__ adj, int64_t entries, uint32_t* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= entries) return;
    const int32_t v = adj[i];
    out[i] = v >= 0 ? bf_hash_pair((uint32_t)v) : 0u;
}
hipError_t launch_adj_hash(const int32_t* adj, int64_t entries, uint32_t* out, hipStream_t s) {
    if (entries <= 0) return hipSuccess;
    hipLaunchKernelGGL(adj_hash_kernel, dim3((unsigned)((entries + 255) / 256)), dim3(256), 0, s,
                       adj, entries, out);
    return hipGetLastError();
}

hipError_t launch_search(const SearchArgs& a, int grid, hipStream_t s) {
    if (a.prec == VMN_PREC_FP8 && !a.adj_norms) return hipErrorInvalidValue;
    if (a.mode == VMN_MODE_FRESH && a.prec == VMN_PREC_FP8 && !a.adj_hash)
        return hipErrorInvalidValue;
    switch (a.DP) {
        case 32: return launch_search_d<32>(a, grid, s);
        case 64: return launch_search_d<64>(a, grid, s);
        case 96: return launch_search_d<96>(a, grid, s);
        case 128: return launch_search_d<128>(a, grid, s);
        case 160: return launch_search_d<160>(a, grid, s);
        case 192: return launch_search_d<192>(a, grid, s);
        case 256: return launch_search_d<256>(a, grid, s);
        default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// Low-precision copies of the rows.
// ---------------------------------------------------------------------------

// Rows on which BF16 traversal equals FP32 traversal bit for bit: every value
// an integer with |v| <= 256 (exact in BF16) and ||x||^2 <= 2^22. Between two
// such rows every product, dot product and norm, ||q||^2 + ||x||^2 - 2 q.x,
// and every partial sum of sum (x - q)^2 (at most (||x|| + ||q||)^2 <= 2^24)
// is an integer an FP32 holds exactly, in any summation order. Clears *flag
// when a row of [from, to) is not such a row. One wavefront per row.
__global__ void bf16_exact_kernel(const float* __restrict__ vecs, int64_t from, int64_t to, int DP,
                                  unsigned int* flag) {
    const int64_t row = from + (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int l = threadIdx.x & 63;
    if (row >= to) return;
    bool ok = true;
    double s = 0.0;
    for (int k = l; k < DP; k += 64) {
        const float v = vecs[row * DP + k];
        ok = ok && v == rintf(v) && fabsf(v) <= 256.0f;
        s += (double)v * (double)v;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    const uint64_t bad = ballot(!ok);
    if (l == 0 && (bad || s > 4194304.0)) atomicAnd(flag, 0u);
}
hipError_t launch_bf16_exact(const float* vecs, int64_t from, int64_t to, int DP, unsigned int* flag,
                             hipStream_t s) {
    if (to <= from) return hipSuccess;
    const int64_t rows = to - from;
    hipLaunchKernelGGL(bf16_exact_kernel, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, s, vecs,
                       from, to, DP, flag);
    return hipGetLastError();
}

// BF16 rows and their squared norms. The norm is summed in the order the
// search kernel sums it from the row registers (gather_dists_low): four
// chains, one per 8-element slice g of each 32-element k-block, each a
// v_dot2_f32_bf16 chain over the element pairs (k-blocks ascending, pairs
// ascending), then (c0 + c1) + (c2 + c3). So a distance read with the table's
// norm (small batches) equals the one summed in registers (large batches)
// bit for bit, and a query's result does not depend on its batch's size.
__global__ void to_bf16_kernel(const float* __restrict__ vecs, int64_t n, int DP, __bf16* rows,
                               float* norms) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int kb = 0; kb < DP / 32; ++kb) {
        for (int g = 0; g < 4; ++g) {
            for (int e = 0; e < 4; ++e) {
                const int k = 32 * kb + 8 * g + 2 * e;
                const __bf16 h0 = to_bf16(vecs[i * DP + k]);
                const __bf16 h1 = to_bf16(vecs[i * DP + k + 1]);
                rows[i * DP + k] = h0;
                rows[i * DP + k + 1] = h1;
                const bf16x2_t pr = {h0, h1};
                c[g] = __builtin_amdgcn_fdot2_f32_bf16(pr, pr, c[g], false);
            }
        }
    }
    norms[i] = (c[0] + c[1]) + (c[2] + c[3]);
}
hipError_t launch_to_bf16(const float* vecs, int64_t n, int DP, void* rows, float* norms,
                          hipStream_t s) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, vecs, n,
                       DP, reinterpret_cast<__bf16*>(rows), norms);
    return hipGetLastError();
}

// FP8 rows are coded about the per-dimension mean c of the build's rows.
// Squared L2 is unchanged by the shift, the centred values are smaller, and
// e4m3's rounding error is relative to the value coded: on 100k rows of the
// 256-component 128-d integer mixture the relative distance error over each
// query's 1000 nearest rows falls from 3.5% to 1.0% rms. The mean is summed in FP64 over fixed row
// chunks (colsum_kernel) and the chunks in order (centre_kernel), so it does
// not depend on scheduling.
__global__ void colsum_kernel(const float* __restrict__ vecs, int64_t n, int DP, int64_t chunk,
                              double* part) {
    const int d = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const int64_t r1 = r0 + chunk < n ? r0 + chunk : n;
    double s = 0.0;
    for (int64_t i = r0; i < r1; ++i) s += (double)vecs[i * DP + d];
    part[(int64_t)blockIdx.x * DP + d] = s;
}
__global__ void centre_kernel(const double* part, int nb, int64_t n, int DP, float* centre) {
    const int d = threadIdx.x;
    if (d >= DP) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[(int64_t)b * DP + d];
    centre[d] = (float)(s / (double)n);
}
// Per-dimension absmax of the centred rows (non-negative float bits order as uints).
__global__ void absmax_kernel(const float* __restrict__ vecs, int64_t n, int DP,
                              const float* __restrict__ centre, unsigned int* absmax_bits) {
    const int d = threadIdx.x;
    const float c = centre[d];
    unsigned int m = 0;
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        const unsigned int b = __float_as_uint(fabsf(vecs[i * DP + d] - c));
        m = b > m ? b : m;
    }
    atomicMax(&absmax_bits[d], m);
}
__global__ void scales_kernel(const unsigned int* absmax_bits, int DP, float* scales) {
    const int d = threadIdx.x;
    if (d >= DP) return;
    const float m = __uint_as_float(absmax_bits[d]);
    scales[d] = m > 0.0f ? m / 448.0f : 1.0f;
}
size_t fp8_scales_temp_bytes(int64_t n, int DP) {
    const int64_t nb = n < 1024 ? (n < 1 ? 1 : n) : 1024;
    return (size_t)nb * DP * sizeof(double);
}
hipError_t launch_fp8_scales(const float* vecs, int64_t n, int D, int DP, unsigned int* absmax_bits,
                             double* temp, float* centre, float* scales, hipStream_t s) {
    (void)D;
    if (n < 1) return hipErrorInvalidValue;
    int64_t nb = n < 1024 ? n : 1024;
    const int64_t chunk = (n + nb - 1) / nb;
    nb = (n + chunk - 1) / chunk;
    hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)nb), dim3(DP), 0, s, vecs, n, DP, chunk, temp);
    hipLaunchKernelGGL(centre_kernel, dim3(1), dim3(256), 0, s, temp, (int)nb, n, DP, centre);
    hipLaunchKernelGGL(fill_zero_u32, dim3(1), dim3(256), 0, s, absmax_bits, DP);
    const int64_t g = n < 2048 ? n : 2048;
    hipLaunchKernelGGL(absmax_kernel, dim3((unsigned)g), dim3(DP), 0, s, vecs, n, DP, centre,
                       absmax_bits);
    hipLaunchKernelGGL(scales_kernel, dim3(1), dim3(256), 0, s, absmax_bits, DP, scales);
    return hipGetLastError();
}

// x~ = e4m3(clamp((x - c) / s)), norms = ||s * x~||^2 with the hardware decode.
__global__ void to_fp8_kernel(const float* __restrict__ vecs, int64_t n, int DP,
                              const float* __restrict__ scales, const float* __restrict__ centre,
                              uint8_t* rows, float* norms) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int k = 0; k < DP; k += 2) {
        const float v0 = fminf(fmaxf((vecs[i * DP + k] - centre[k]) / scales[k], -448.0f), 448.0f);
        const float v1 =
            fminf(fmaxf((vecs[i * DP + k + 1] - centre[k + 1]) / scales[k + 1], -448.0f), 448.0f);
        const uint32_t w = pk_fp8(v0, v1);
        rows[i * DP + k] = (uint8_t)(w & 0xff);
        rows[i * DP + k + 1] = (uint8_t)(w >> 8);
        const float d0 = fp8_to_f32(w, 0) * scales[k];
        const float d1 = fp8_to_f32(w, 1) * scales[k + 1];
        s += d0 * d0;
        s += d1 * d1;
    }
    norms[i] = s;
}
hipError_t launch_to_fp8(const float* vecs, int64_t n, int DP, const float* scales,
                         const float* centre, void* rows, float* norms, hipStream_t s) {
    hipLaunchKernelGGL(to_fp8_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, vecs, n,
                       DP, scales, centre, reinterpret_cast<uint8_t*>(rows), norms);
    return hipGetLastError();
}

// Dequantised FP8 rows (s * x~ + c), for error reports and tests.
__global__ void fp8_dequant_kernel(const uint8_t* rows, int64_t n, int DP, const float* scales,
                                   const float* centre, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * DP) return;
    out[i] = fp8_to_f32(rows[i], 0) * scales[i % DP] + centre[i % DP];
}
hipError_t launch_fp8_dequant(const void* rows, int64_t n, int DP, const float* scales,
                              const float* centre, float* out, hipStream_t s) {
    hipLaunchKernelGGL(fp8_dequant_kernel, dim3((unsigned)((n * DP + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const uint8_t*>(rows), n, DP, scales, centre, out);
    return hipGetLastError();
}

}  // namespace vmn
