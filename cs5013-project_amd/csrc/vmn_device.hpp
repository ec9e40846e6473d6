// Device-side building blocks shared by the Vamana kernels (gfx950, wave64).
//
// Layout conventions used everywhere:
//  * vectors: row-major [capacity x DP] float, DP = dimension rounded up to a
//    multiple of 32 and zero padded (padding adds (0-0)^2 = 0 to every
//    distance, so distances are unchanged).
//  * adjacency: row-major [capacity x W] int32, W = row width (max degree),
//    -1 terminated/padded, no duplicate ids in a row.
//  * one wavefront (64 lanes) owns one query / node; lanes are split into 8
//    row groups of 8 lanes; a row group reads one candidate vector as 8 lanes x
//    DP/32 float4 loads (each load instruction of a group covers 128
//    contiguous bytes), so 8 candidate rows are in flight per instruction.
//
// Distances are sums of squared differences reduced with a lane tree. On
// integer-valued data below 2^24 (SIFT, the Gaussian-mixture benchmark sets)
// every partial sum is an exact float, so the result is bit-identical to the
// reference's sequential sum regardless of order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define VMN_WAVE 64
#define VMN_VIS_BIT 0x80000000u
#define VMN_ID_MASK 0x7fffffffu

namespace vmn {

__device__ __forceinline__ int lane() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    int l = lane();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ int readlane_i(int v, int l) {
    return __builtin_amdgcn_readlane(v, l);
}

// Sum over the 8 lanes of a row group (lanes 8g..8g+7).
__device__ __forceinline__ float reduce8(float v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    return v;
}

// Query (or pivot) vector held in registers: lane sub (0..7) holds float4
// indices sub, sub+8, sub+16, ... of the padded row.
template <int DP>
struct QReg {
    static constexpr int T = DP / 32;
    float4 v[T];
    __device__ __forceinline__ void load(const float* row, int sub) {
        const float4* r = reinterpret_cast<const float4*>(row);
#pragma unroll
        for (int t = 0; t < T; ++t) v[t] = r[sub + 8 * t];
    }
    // This lane's share of ||row - q||^2.
    __device__ __forceinline__ float partial(const float* row, int sub) const {
        const float4* r = reinterpret_cast<const float4*>(row);
        float acc = 0.0f;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            float4 x = r[sub + 8 * t];
            float a = x.x - v[t].x, b = x.y - v[t].y, c = x.z - v[t].z, d = x.w - v[t].w;
            acc += a * a;
            acc += b * b;
            acc += c * c;
            acc += d * d;
        }
        return acc;
    }
};

// Distances from q to `cnt` rows whose ids sit in LDS ids[0..cnt); lane s of
// the result slot dist_out[s] (LDS) receives the distance of ids[s].
// Must be called by the whole wave.
template <int DP>
__device__ __forceinline__ void gather_dists(const QReg<DP>& q, const float* __restrict__ vecs,
                                             const int32_t* ids, float* dist_out, int cnt) {
    const int l = lane(), sub = l & 7, grp = l >> 3;
    for (int base = 0; base < cnt; base += 8) {
        const int r = base + grp;
        float acc = 0.0f;
        if (r < cnt) {
            const int32_t id = ids[r];
            acc = q.partial(vecs + (int64_t)id * DP, sub);
        }
        acc = reduce8(acc);
        if (sub == 0 && r < cnt) dist_out[r] = acc;
    }
}

// FreshVamana bloom filter hashes (vamana/include/bloom_filter.cuh:11-30):
// FNV-style over the 4 bytes of the id with a 64-bit accumulator, reduced
// modulo the filter size (39,988 cells per query in the reference; here one
// bit per cell).
#define VMN_BF_CELLS 39988u
#define VMN_BF_WORDS ((VMN_BF_CELLS + 31u) / 32u)

__host__ __device__ __forceinline__ uint32_t bf_hash1(uint32_t x) {
    uint64_t h = 0xcbf29ce4ull;
    h = (h ^ (x & 0xffu)) * 0x01000193ull;
    h = (h ^ ((x >> 8) & 0xffu)) * 0x01000193ull;
    h = (h ^ ((x >> 16) & 0xffu)) * 0x01000193ull;
    h = (h ^ ((x >> 24) & 0xffu)) * 0x01000193ull;
    return (uint32_t)(h % VMN_BF_CELLS);
}

__host__ __device__ __forceinline__ uint32_t bf_hash2(uint32_t x) {
    uint64_t h = 0x84222325ull;
    h = (h ^ (x & 0xffu)) * 0x1B3ull;
    h = (h ^ ((x >> 8) & 0xffu)) * 0x1B3ull;
    h = (h ^ ((x >> 16) & 0xffu)) * 0x1B3ull;
    h = (h ^ ((x >> 24) & 0xffu)) * 0x1B3ull;
    return (uint32_t)(h % VMN_BF_CELLS);
}

__device__ __forceinline__ bool bit_test(const uint32_t* bits, uint32_t i) {
    return (bits[i >> 5] >> (i & 31)) & 1u;
}

// Merge `nnew` new candidates (one per lane: my_id/my_d valid for lane < nnew,
// arbitrary order) into the sorted list (src_id/src_d, cnt entries), writing
// the first `limit` entries of the merged order to dst_id/dst_d. Returns the
// new count. CAP is a compile-time bound on cnt.
//
// BY_ID = false (stable merge): order by distance; equal distances keep list
//   entries before new ones and new ones in lane order. This is the order a
//   stable sort of [list..., new...] gives (and the order the reference's
//   sequential insert / stable merge sorts produce).
// BY_ID = true (keyed merge): order by (distance, id & VMN_ID_MASK); used where
//   the reference's order is atomic-arrival dependent and a canonical order is
//   chosen instead.
template <int CAP, bool BY_ID>
__device__ __forceinline__ int merge_sorted(const uint32_t* src_id, const float* src_d, int cnt,
                                            uint32_t my_id, float my_d, int nnew,
                                            uint32_t* dst_id, float* dst_d, int limit) {
    constexpr int PER = (CAP + 63) / 64;
    const int l = lane();
    const uint32_t my_key = my_id & VMN_ID_MASK;
    float bdv[PER];
    uint32_t bidv[PER];
    int less[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        int bi = l + 64 * j;
        bdv[j] = bi < cnt ? src_d[bi] : 0.0f;
        bidv[j] = bi < cnt ? src_id[bi] : 0u;
        less[j] = 0;
    }
    int rank = 0;
    for (int t = 0; t < nnew; ++t) {
        float dt = readlane_f(my_d, t);
        if (BY_ID) {
            uint32_t kt = (uint32_t)readlane_i((int)my_key, t);
            rank += (dt < my_d) || (dt == my_d && kt < my_key);
#pragma unroll
            for (int j = 0; j < PER; ++j)
                less[j] += (dt < bdv[j]) || (dt == bdv[j] && kt < (bidv[j] & VMN_ID_MASK));
        } else {
            rank += (dt < my_d) || (dt == my_d && t < l);
#pragma unroll
            for (int j = 0; j < PER; ++j) less[j] += dt < bdv[j];
        }
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        int bi = l + 64 * j;
        if (bi < cnt) {
            int pos = bi + less[j];
            if (pos < limit) {
                dst_id[pos] = bidv[j];
                dst_d[pos] = bdv[j];
            }
        }
    }
    if (l < nnew) {
        // list entries ordered before the new element
        int lo = 0, hi = cnt;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            float dm = src_d[mid];
            bool before = BY_ID ? (dm < my_d || (dm == my_d && (src_id[mid] & VMN_ID_MASK) < my_key))
                                : (dm <= my_d);
            if (before) lo = mid + 1;
            else hi = mid;
        }
        int pos = lo + rank;
        if (pos < limit) {
            dst_id[pos] = my_id;
            dst_d[pos] = my_d;
        }
    }
    int nc = cnt + nnew;
    return nc < limit ? nc : limit;
}

}  // namespace vmn
