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
// SEQ (FP32 rows only): sequential fused multiply-add chains over the
// dimensions, one lane per candidate (gather_dists_seq), the summation order
// of the reference's parth distance<D> (parth/src/device_funcs.cu:17-49).
template <int DP, bool BF, bool SEQ = false>
struct PruneQ;

template <int DP>
struct PruneQ<DP, false, false> {
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
struct PruneQ<DP, false, true> {
    const float* row;
    __device__ __forceinline__ void load(const PruneRows& r, int64_t id) { row = r.f32 + id * DP; }
    __device__ __forceinline__ void dists(const PruneRows& r, const int32_t* ids, float* out,
                                          int cnt) const {
        gather_dists_seq<DP>(row, r.f32, ids, out, cnt);
    }
};

template <int DP>
struct PruneQ<DP, true, false> {
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
template <int DP, bool STRICT, bool SEQ = false>
__device__ int robust_prune(const PruneRows& rows, const uint32_t* c_id,
                            const float* c_d, int cnt, float alpha, int R, uint8_t* s_kept,
                            int32_t* s_act_id, int32_t* s_act_pos, float* s_act_d,
                            int32_t* out_ids, float* out_d, int* nread = nullptr) {
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
        PruneQ<DP, false, SEQ> pq;
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
        if (nread) *nread += nact;
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

// ---------------------------------------------------------------------------
// Occlusion masks on MFMA for rows exact in BF16. Up to 16 "row" nodes (this
// lane's row is `sid`, the node of A row l % 16; rows at or past `nrows` are
// ignored) against the active candidates s_act_id[0..nact): s_pm[k] gets bit
// j when row j occludes candidate k, alpha * d(row j, k) < d(node, k)
// (STRICT) or <= (c_d[s_act_pos[k]] is d(node, k)); ORDERED keeps only
// candidates after row j (k >= j, row j >= 1 being active entry j - 1).
// Distances are ||a||^2 + ||b||^2 - 2 a.b from the BF16 rows and their norms:
// exact integers, equal to the FP32 sums of squared differences. Lane l loads
// 16 bytes of tile row l/4 per k-block (coalesced); ds_bpermute hands each
// dword to the lane that holds it as the B operand. Two tiles of 16
// candidates per round trip.
// ---------------------------------------------------------------------------
template <int DP, bool STRICT, bool ORDERED>
__device__ __forceinline__ void gram_masks(const PruneRows& rows, int32_t sid, int nrows,
                                           int nact, const int32_t* s_act_id,
                                           const int32_t* s_act_pos, const float* c_d,
                                           float alpha, uint32_t* s_pm, int* nread = nullptr) {
    constexpr int KB = DP / 32;
    // rows read: the 16 A rows and the active candidates' rows
    if (nread) *nread += 16 + nact;
    const int l = lane(), g = l >> 4, rr = l >> 2, sub = l & 3, col = l & 15;
    const __bf16* bro = reinterpret_cast<const __bf16*>(rows.bf16);
    const float* nrm = rows.bf16_norms;
    bf16x8_t A[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
        A[kb] = *reinterpret_cast<const bf16x8_t*>(bro + (int64_t)sid * DP + 32 * kb + 8 * g);
    // squared norms of rows 4g .. 4g+3 (lane r < 16 holds row r's node)
    float an[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) an[i] = nrm[__builtin_amdgcn_ds_bpermute((4 * g + i) << 2, sid)];
    const int src = (4 * col + g) << 2;
    for (int base = 0; base < nact; base += 32) {
        int32_t cid[2];
        uint4 v[2][KB];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int c = base + 16 * t + rr;
            cid[t] = c < nact ? s_act_id[c] : s_act_id[0];
            if (t == 0 || base + 16 * t < nact) {
                const uint4* r = reinterpret_cast<const uint4*>(bro + (int64_t)cid[t] * DP);
#pragma unroll
                for (int kb = 0; kb < KB; ++kb) v[t][kb] = r[4 * kb + sub];
            }
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            if (t == 0 || base + 16 * t < nact) {
                f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kb = 0; kb < KB; ++kb) {
                    const uint4 u = {bperm(src, v[t][kb].x), bperm(src, v[t][kb].y),
                                     bperm(src, v[t][kb].z), bperm(src, v[t][kb].w)};
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[kb], __builtin_bit_cast(bf16x8_t, u), acc, 0, 0, 0);
                }
                // this lane: candidate k (column col) against rows 4g .. 4g+3
                const int k = base + 16 * t + col;
                uint32_t bits = 0;
                if (k < nact) {
                    const float xn = nrm[s_act_id[k]];
                    const float dn = c_d[s_act_pos[k]];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int j = 4 * g + i;
                        const float lhs = alpha * ((an[i] + xn) - 2.0f * acc[i]);
                        const bool pr = STRICT ? (lhs < dn) : (lhs <= dn);
                        bits |= (uint32_t)(pr && j < nrows && (!ORDERED || k >= j)) << j;
                    }
                }
                bits |= __shfl_xor(bits, 16);
                bits |= __shfl_xor(bits, 32);
                if (g == 0 && k < nact) s_pm[k] = bits;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// The same prune for rows exact in BF16, with the p*-to-candidate distances
// as MFMA tiles. The next 16 kept candidates (p0, then the first 15 of the
// active list behind it) are the 16 rows of A; one pass over the active
// candidates' BF16 rows gives the distances from all 16 of them to every
// active candidate, reduced at once to a 16-bit mask per candidate
// (gram_masks, ORDERED). The 16 are then taken in list order as the
// sequential loop takes them: one still kept at its turn is a p*, is output,
// and drops the later candidates its mask bit marks; one an earlier p*
// dropped is skipped. Every candidate between two of the 16 was already
// dropped, so the next scan, from just after the 16th, finds the next p*, and
// the output equals robust_prune's (the distances are exact integers either
// way). One gather of the active rows serves up to 16 rounds.
// nk0 entries of out_ids are already taken (output continues after them);
// with init_kept false, s_kept[0..cnt) holds the caller's kept flags.
// ---------------------------------------------------------------------------
template <int DP, bool STRICT, int CCAP>
__device__ int robust_prune_mfma(const PruneRows& rows, const uint32_t* c_id, const float* c_d,
                                 int cnt, float alpha, int R, uint8_t* s_kept,
                                 int32_t* s_act_id, int32_t* s_act_pos, uint32_t* s_pm,
                                 int32_t* out_ids, float* out_d, int nk0 = 0,
                                 bool init_kept = true, int* nread = nullptr) {
    constexpr int NCH = (CCAP + 63) / 64;
    const int l = lane(), col = l & 15;
    if (init_kept) {
        for (int i = l; i < cnt; i += 64) s_kept[i] = 1;
    }
    __syncthreads();
    int nk = nk0, from = 0;
    while (nk < R) {
        // kept candidates from `from` on: the first is p0, the others the active list
        int nall = 0, p0 = -1;
        for (int b = from; b < cnt; b += 64) {
            const int j = b + l;
            const bool f = j < cnt && s_kept[j];
            const uint64_t m = ballot(f);
            const int r = nall + __popcll((unsigned long long)(m & lanemask_lt()));
            if (p0 < 0 && m) p0 = b + __ffsll((unsigned long long)m) - 1;
            if (f && r > 0) {
                s_act_id[r - 1] = (int32_t)(c_id[j] & VMN_ID_MASK);
                s_act_pos[r - 1] = j;
            }
            nall += __popcll((unsigned long long)m);
        }
        if (nall == 0) break;
        const int nact = nall - 1;
        const int ns = nall < 16 ? nall : 16;
        const int32_t id0 = (int32_t)(c_id[p0] & VMN_ID_MASK);
        __syncthreads();
        if (nact > 0) {
            // A row `col`: p0 (row 0) or active entry col - 1 (rows past the
            // list repeat p0 and are ignored)
            const int32_t sid = (col > 0 && col - 1 < nact) ? s_act_id[col - 1] : id0;
            gram_masks<DP, STRICT, true>(rows, sid, ns, nact, s_act_id, s_act_pos, c_d, alpha,
                                         s_pm, nread);
        }
        __syncthreads();
        uint32_t pm[NCH];
        uint64_t keep[NCH];
#pragma unroll
        for (int m = 0; m < NCH; ++m) {
            const int k = 64 * m + l;
            pm[m] = k < nact ? s_pm[k] : 0u;
            keep[m] = ballot(k < nact);
        }
        bool done = false;
        for (int j = 0; j < ns; ++j) {
            // row j >= 1 is active entry j - 1: still kept?
            if (j > 0 && !((keep[0] >> (j - 1)) & 1ull)) continue;
            if (l == 0) {
                out_ids[nk] = j == 0 ? id0 : s_act_id[j - 1];
                if (out_d) out_d[nk] = c_d[j == 0 ? p0 : s_act_pos[j - 1]];
            }
            ++nk;
            if (nk >= R) {
                done = true;
                break;
            }
#pragma unroll
            for (int m = 0; m < NCH; ++m) keep[m] &= ~ballot((pm[m] >> j) & 1u);
        }
        if (done) break;
        // drops back to the list; the next scan starts after the last of the 16
#pragma unroll
        for (int m = 0; m < NCH; ++m) {
            const int k = 64 * m + l;
            if (k < nact && !((keep[m] >> l) & 1ull)) s_kept[s_act_pos[k]] = 0;
        }
        from = (ns == 1 ? p0 : s_act_pos[ns - 2]) + 1;
        __syncthreads();
    }
    __syncthreads();
    return nk;
}

// Every node of members[0..nm) drops the kept candidates it occludes
// (alpha * d(member, c) <= d(node, c)), 16 members per pass over the kept
// candidates' BF16 rows. The members all apply, in any order, so a candidate
// stays exactly when no member occludes it.
template <int DP>
__device__ void occlude_mfma(const PruneRows& rows, const int32_t* members, int nm,
                             const uint32_t* c_id, const float* c_d, int cnt, float alpha,
                             uint8_t* s_kept, int32_t* s_act_id, int32_t* s_act_pos,
                             uint32_t* s_pm, int* nread = nullptr) {
    const int l = lane(), col = l & 15;
    for (int g0 = 0; g0 < nm; g0 += 16) {
        int nact = 0;
        for (int b = 0; b < cnt; b += 64) {
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
        if (nact == 0) break;
        __syncthreads();
        const int nrows = nm - g0 < 16 ? nm - g0 : 16;
        const int32_t sid = members[g0 + (col < nrows ? col : 0)];
        gram_masks<DP, false, false>(rows, sid, nrows, nact, s_act_id, s_act_pos, c_d, alpha,
                                     s_pm, nread);
        __syncthreads();
        for (int k = l; k < nact; k += 64)
            if (s_pm[k]) s_kept[s_act_pos[k]] = 0;
        __syncthreads();
    }
}

// robust_prune over FP32 rows with the next NB kept candidates taken as
// pivots per pass over the active rows (the FP32 counterpart of
// robust_prune_mfma's 16-row Gram). Each row group holds the NB pivot rows in
// registers (QReg layout) and loads two active rows per round trip; every
// pivot-row distance is gather_dists's sum, in its order (8 lane partials,
// reduce8), so each occlusion decision is the one-pivot-at-a-time loop's.
// The decisions become an NB-bit mask per active candidate (bit j: pivot j,
// which is p0 or active entry j - 1, occludes it and comes before it), and
// the pivots are then taken in list order as robust_prune takes them: one
// still kept at its turn is a p*, is output, and drops what its bit marks;
// one an earlier p* dropped is skipped. The next pass scans from after the
// last pivot. One gather of the active rows serves up to NB rounds.
template <int DP, bool STRICT, int CCAP>
__device__ int robust_prune_batch(const PruneRows& rows, const uint32_t* c_id, const float* c_d,
                                  int cnt, float alpha, int R, uint8_t* s_kept,
                                  int32_t* s_act_id, int32_t* s_act_pos, uint32_t* s_pm,
                                  int32_t* out_ids, float* out_d, int nk0 = 0,
                                  bool init_kept = true, int* nread = nullptr) {
    constexpr int NB = DP <= 128 ? 4 : 2;
    constexpr int NCH = (CCAP + 63) / 64;
    constexpr int T = DP / 32;
    const int l = lane(), sub = l & 7, grp = l >> 3;
    if (init_kept) {
        for (int i = l; i < cnt; i += 64) s_kept[i] = 1;
    }
    __syncthreads();
    int nk = nk0, from = 0;
    while (nk < R) {
        // kept candidates from `from` on: the first is p0, the others the active list
        int nall = 0, p0 = -1;
        for (int b = from; b < cnt; b += 64) {
            const int j = b + l;
            const bool f = j < cnt && s_kept[j];
            const uint64_t m = ballot(f);
            const int r = nall + __popcll((unsigned long long)(m & lanemask_lt()));
            if (p0 < 0 && m) p0 = b + __ffsll((unsigned long long)m) - 1;
            if (f && r > 0) {
                s_act_id[r - 1] = (int32_t)(c_id[j] & VMN_ID_MASK);
                s_act_pos[r - 1] = j;
            }
            nall += __popcll((unsigned long long)m);
        }
        if (nall == 0) break;
        const int nact = nall - 1;
        const int ns = nall < NB ? nall : NB;
        const int32_t id0 = (int32_t)(c_id[p0] & VMN_ID_MASK);
        __syncthreads();
        if (nact > 0) {
            if (nread) *nread += ns + nact;
            QReg<DP> pv[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int32_t pid = j == 0 ? id0 : (j - 1 < nact ? s_act_id[j - 1] : id0);
                if (j < ns) pv[j].load(rows.f32 + (int64_t)pid * DP, sub);
            }
            for (int base = 0; base < nact; base += 16) {
                const int r0 = base + grp, r1 = base + 8 + grp;
                float4 x0[T], x1[T];
                const float4* row0 = reinterpret_cast<const float4*>(rows.f32 + (int64_t)s_act_id[r0 < nact ? r0 : 0] * DP);
                const float4* row1 = reinterpret_cast<const float4*>(rows.f32 + (int64_t)s_act_id[r1 < nact ? r1 : 0] * DP);
                const bool two = base + 8 < nact;
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    x0[t] = row0[sub + 8 * t];
                    if (two) x1[t] = row1[sub + 8 * t];
                }
                uint32_t b0 = 0u, b1 = 0u;
                const float dn0 = r0 < nact ? c_d[s_act_pos[r0]] : 0.0f;
                const float dn1 = r1 < nact ? c_d[s_act_pos[r1]] : 0.0f;
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    if (j >= ns) continue;
                    float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
                    for (int t = 0; t < T; ++t) {
                        float a = x0[t].x - pv[j].v[t].x, b = x0[t].y - pv[j].v[t].y;
                        float c = x0[t].z - pv[j].v[t].z, d = x0[t].w - pv[j].v[t].w;
                        s0 += a * a; s0 += b * b; s0 += c * c; s0 += d * d;
                        if (two) {
                            a = x1[t].x - pv[j].v[t].x; b = x1[t].y - pv[j].v[t].y;
                            c = x1[t].z - pv[j].v[t].z; d = x1[t].w - pv[j].v[t].w;
                            s1 += a * a; s1 += b * b; s1 += c * c; s1 += d * d;
                        }
                    }
                    s0 = reduce8(s0);
                    s1 = reduce8(s1);
                    const float l0 = alpha * s0, l1 = alpha * s1;
                    const bool o0 = STRICT ? (l0 < dn0) : (l0 <= dn0);
                    const bool o1 = STRICT ? (l1 < dn1) : (l1 <= dn1);
                    // pivot j >= 1 is active entry j - 1: it drops only entries after it
                    b0 |= (uint32_t)(o0 && r0 >= j) << j;
                    b1 |= (uint32_t)(o1 && r1 >= j) << j;
                }
                if (sub == 0 && r0 < nact) s_pm[r0] = b0;
                if (sub == 0 && two && r1 < nact) s_pm[r1] = b1;
            }
        }
        __syncthreads();
        uint32_t pm[NCH];
        uint64_t keep[NCH];
#pragma unroll
        for (int m = 0; m < NCH; ++m) {
            const int k = 64 * m + l;
            pm[m] = k < nact ? s_pm[k] : 0u;
            keep[m] = ballot(k < nact);
        }
        bool done = false;
        for (int j = 0; j < ns; ++j) {
            // pivot j >= 1 is active entry j - 1: still kept?
            if (j > 0 && !((keep[0] >> (j - 1)) & 1ull)) continue;
            if (l == 0) {
                out_ids[nk] = j == 0 ? id0 : s_act_id[j - 1];
                if (out_d) out_d[nk] = c_d[j == 0 ? p0 : s_act_pos[j - 1]];
            }
            ++nk;
            if (nk >= R) {
                done = true;
                break;
            }
#pragma unroll
            for (int m = 0; m < NCH; ++m) keep[m] &= ~ballot((pm[m] >> j) & 1u);
        }
        if (done) break;
        // drops back to the list; the next scan starts after the last pivot
#pragma unroll
        for (int m = 0; m < NCH; ++m) {
            const int k = 64 * m + l;
            if (k < nact && !((keep[m] >> l) & 1ull)) s_kept[s_act_pos[k]] = 0;
        }
        from = (ns == 1 ? p0 : s_act_pos[ns - 2]) + 1;
        __syncthreads();
    }
    __syncthreads();
    return nk;
}

// occlude_mfma over FP32 rows: every node of members[0..nm) drops the kept
// candidates it occludes (alpha * d(member, c) <= d(node, c)), NB members per
// pass over the kept candidates' rows (two rows per row group and round
// trip), each distance gather_dists's sum in its order. The members all
// apply, in any order, so a candidate stays exactly when no member occludes
// it.
template <int DP>
__device__ void occlude_fp32(const PruneRows& rows, const int32_t* members, int nm,
                             const uint32_t* c_id, const float* c_d, int cnt, float alpha,
                             uint8_t* s_kept, int32_t* s_act_id, int32_t* s_act_pos,
                             int* nread = nullptr) {
    constexpr int NB = DP <= 128 ? 4 : 2;
    constexpr int T = DP / 32;
    const int l = lane(), sub = l & 7, grp = l >> 3;
    for (int g0 = 0; g0 < nm; g0 += NB) {
        int nact = 0;
        for (int b = 0; b < cnt; b += 64) {
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
        if (nact == 0) break;
        __syncthreads();
        const int nrows = nm - g0 < NB ? nm - g0 : NB;
        if (nread) *nread += nrows + nact;
        QReg<DP> pv[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j)
            if (j < nrows) pv[j].load(rows.f32 + (int64_t)members[g0 + j] * DP, sub);
        for (int base = 0; base < nact; base += 16) {
            const int r0 = base + grp, r1 = base + 8 + grp;
            float4 x0[T], x1[T];
            const float4* row0 = reinterpret_cast<const float4*>(rows.f32 + (int64_t)s_act_id[r0 < nact ? r0 : 0] * DP);
            const float4* row1 = reinterpret_cast<const float4*>(rows.f32 + (int64_t)s_act_id[r1 < nact ? r1 : 0] * DP);
            const bool two = base + 8 < nact;
#pragma unroll
            for (int t = 0; t < T; ++t) {
                x0[t] = row0[sub + 8 * t];
                if (two) x1[t] = row1[sub + 8 * t];
            }
            const float dn0 = r0 < nact ? c_d[s_act_pos[r0]] : 0.0f;
            const float dn1 = r1 < nact ? c_d[s_act_pos[r1]] : 0.0f;
            bool o0 = false, o1 = false;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                if (j >= nrows) continue;
                float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    float a = x0[t].x - pv[j].v[t].x, b = x0[t].y - pv[j].v[t].y;
                    float c = x0[t].z - pv[j].v[t].z, d = x0[t].w - pv[j].v[t].w;
                    s0 += a * a; s0 += b * b; s0 += c * c; s0 += d * d;
                    if (two) {
                        a = x1[t].x - pv[j].v[t].x; b = x1[t].y - pv[j].v[t].y;
                        c = x1[t].z - pv[j].v[t].z; d = x1[t].w - pv[j].v[t].w;
                        s1 += a * a; s1 += b * b; s1 += c * c; s1 += d * d;
                    }
                }
                s0 = reduce8(s0);
                s1 = reduce8(s1);
                o0 |= alpha * s0 <= dn0;
                o1 |= alpha * s1 <= dn1;
            }
            if (sub == 0 && r0 < nact && o0) s_kept[s_act_pos[r0]] = 0;
            if (sub == 0 && two && r1 < nact && o1) s_kept[s_act_pos[r1]] = 0;
        }
        __syncthreads();
    }
}

// robust_prune_batch over FP32 rows, or robust_prune_mfma when the rows are
// exact in BF16 (same output). SEQ: FP32 distances as sequential chains, one
// p* per round (robust_prune with PruneQ).
template <int DP, bool STRICT, bool BF, int CCAP, bool SEQ = false>
__device__ __forceinline__ int prune_list(const PruneRows& rows, const uint32_t* c_id,
                                          const float* c_d, int cnt, float alpha, int R,
                                          uint8_t* s_kept, int32_t* s_act_id, int32_t* s_act_pos,
                                          float* s_act_d, int32_t* out_ids, float* out_d,
                                          int* nread = nullptr) {
    if constexpr (BF)
        return robust_prune_mfma<DP, STRICT, CCAP>(rows, c_id, c_d, cnt, alpha, R, s_kept,
                                                   s_act_id, s_act_pos,
                                                   reinterpret_cast<uint32_t*>(s_act_d), out_ids,
                                                   out_d, 0, true, nread);
    else if constexpr (!SEQ)
        return robust_prune_batch<DP, STRICT, CCAP>(rows, c_id, c_d, cnt, alpha, R, s_kept,
                                                    s_act_id, s_act_pos,
                                                    reinterpret_cast<uint32_t*>(s_act_d), out_ids,
                                                    out_d, 0, true, nread);
    else
        return robust_prune<DP, STRICT, SEQ>(rows, c_id, c_d, cnt, alpha, R, s_kept, s_act_id,
                                             s_act_pos, s_act_d, out_ids, out_d, nread);
}

// Distances from `node` to its current row entries that pass `keep`, appended
// (compacted, adjacency order) to e_id/e_d. Returns the count.
template <int DP, bool BF, bool SEQ = false>
__device__ __forceinline__ int row_extras(const PruneRows& rows, const PruneQ<DP, BF, SEQ>& q,
                                          bool pass, int32_t nb, int32_t* e_id, float* e_d,
                                          int* nread = nullptr) {
    const uint64_t m = ballot(pass);
    const int ne = __popcll((unsigned long long)m);
    if (nread) *nread += ne;
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
