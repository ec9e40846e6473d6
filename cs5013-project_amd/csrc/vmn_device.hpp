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
// MFMA tiles of 16 candidates whose rows are in flight together (the prunes'
// and the search's default; the search autotuner may take 4)
#define VMN_GATHER_TILES 2
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

// v_writelane_b32 (the compiler's intrinsic; it has no builtin here): `old`
// with lane l replaced by the uniform value v.
__device__ int writelane_i(int v, int l, int old) __asm("llvm.amdgcn.writelane.i32");

// DPP lane moves within rows of 16 lanes: quad_perm [1,0,3,2] and [2,3,0,1]
// read lanes l^1 and l^2; row_half_mirror reads lane 7 - l of the 8-lane
// half, row_mirror lane 15 - l of the row. Every source lane is in the row,
// so the moves fold into the instruction that consumes them.
#define VMN_DPP_XOR1 0xB1
#define VMN_DPP_XOR2 0x4E
#define VMN_DPP_HALF_MIRROR 0x141
#define VMN_DPP_MIRROR 0x140

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}

// Order-preserving integer key of a distance: x < y iff fkey(x) < fkey(y)
// (signed), equal values have equal keys (-0 is folded into +0 first).
__device__ __forceinline__ int fkey(float x) {
    const int b = __float_as_int(x + 0.0f);
    return b ^ ((b >> 31) & 0x7fffffff);
}

// Minimum of k over the wave, returned uniform: within each row by DPP, then
// over the four rows' results in scalar registers.
__device__ __forceinline__ int wave_min_key(int k) {
    k = min(k, dpp_i<VMN_DPP_XOR1>(k));
    k = min(k, dpp_i<VMN_DPP_XOR2>(k));
    k = min(k, dpp_i<VMN_DPP_HALF_MIRROR>(k));
    k = min(k, dpp_i<VMN_DPP_MIRROR>(k));
    return min(min(readlane_i(k, 0), readlane_i(k, 16)), min(readlane_i(k, 32), readlane_i(k, 48)));
}

// Sum over the 8 lanes of a row group (lanes 8g..8g+7), every lane of the
// group receiving it: pairs l, l^1, then l, l^2, then the two quads (the
// quads are uniform by then, so the mirror reads the l^4 quad's value).
__device__ __forceinline__ float reduce8(float v) {
    v += dpp_f<VMN_DPP_XOR1>(v);
    v += dpp_f<VMN_DPP_XOR2>(v);
    v += dpp_f<VMN_DPP_HALF_MIRROR>(v);
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
    __device__ __forceinline__ void partial2(const float* row0, const float* row1, int sub,
                                             float& s0, float& s1) const {
        const float4* r0 = reinterpret_cast<const float4*>(row0);
        const float4* r1 = reinterpret_cast<const float4*>(row1);
        float4 x0[T], x1[T];
#pragma unroll
        for (int t = 0; t < T; ++t) { x0[t] = r0[sub + 8 * t]; x1[t] = r1[sub + 8 * t]; }
        s0 = 0.0f; s1 = 0.0f;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            float a = x0[t].x - v[t].x, b = x0[t].y - v[t].y, c = x0[t].z - v[t].z, d = x0[t].w - v[t].w;
            s0 += a * a; s0 += b * b; s0 += c * c; s0 += d * d;
            a = x1[t].x - v[t].x; b = x1[t].y - v[t].y; c = x1[t].z - v[t].z; d = x1[t].w - v[t].w;
            s1 += a * a; s1 += b * b; s1 += c * c; s1 += d * d;
        }
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
    for (int base = 0; base < cnt; base += 16) {
        const int r0 = base + grp, r1 = base + 8 + grp;
        float a0 = 0.0f, a1 = 0.0f;
        const int32_t i0 = r0 < cnt ? ids[r0] : ids[0];
        const int32_t i1 = r1 < cnt ? ids[r1] : ids[0];
        if (base + 8 < cnt) {
            q.partial2(vecs + (int64_t)i0 * DP, vecs + (int64_t)i1 * DP, sub, a0, a1);
        } else {
            a0 = q.partial(vecs + (int64_t)i0 * DP, sub);
        }
        a0 = reduce8(a0);
        a1 = reduce8(a1);
        if (sub == 0 && r0 < cnt) dist_out[r0] = a0;
        if (sub == 0 && r1 < cnt) dist_out[r1] = a1;
    }
}

// Distances from the row q (a padded row in global memory, the same for every
// lane) to `cnt` rows whose ids sit in LDS ids[0..cnt), one lane per row,
// summed in index order as one fused multiply-add chain acc = fma(d_i, d_i,
// acc), d_i = q_i - x_i: the order of the reference's device distance<D>
// (parth/src/device_funcs.cu:17-49, `dist += diff * diff` contracted by nvcc's
// default --fmad=true). The padded tail adds fma(0, 0, acc) = acc. Equal to
// the lane-tree sums bit for bit on integer data; on real-valued rows this is
// the form whose results match the reference's. Must be called by the whole
// wave; dist_out[s] (LDS) receives the distance of ids[s].
template <int DP>
__device__ __forceinline__ void gather_dists_seq(const float* __restrict__ q,
                                                 const float* __restrict__ vecs,
                                                 const int32_t* ids, float* dist_out, int cnt) {
    constexpr int T = DP / 4;
    constexpr int CH = 8;  // float4 loads in flight per lane
    const float4* qq = reinterpret_cast<const float4*>(q);
    for (int base = 0; base < cnt; base += 64) {
        const int c = base + lane();
        if (c < cnt) {
            const float4* x = reinterpret_cast<const float4*>(vecs + (int64_t)ids[c] * DP);
            float acc = 0.0f;
#pragma unroll
            for (int t0 = 0; t0 < T; t0 += CH) {
                float4 xv[CH];
#pragma unroll
                for (int t = 0; t < CH; ++t)
                    if (t0 + t < T) xv[t] = x[t0 + t];
#pragma unroll
                for (int t = 0; t < CH; ++t) {
                    if (t0 + t < T) {
                        const float4 qv = qq[t0 + t];
                        const float a = qv.x - xv[t].x, b = qv.y - xv[t].y;
                        const float e = qv.z - xv[t].z, f = qv.w - xv[t].w;
                        acc = __builtin_fmaf(a, a, acc);
                        acc = __builtin_fmaf(b, b, acc);
                        acc = __builtin_fmaf(e, e, acc);
                        acc = __builtin_fmaf(f, f, acc);
                    }
                }
            }
            dist_out[c] = acc;
        }
    }
}

// ---------------------------------------------------------------------------
// Low-precision candidate rows on MFMA.
//
// A wave computes q . x for 16 candidates per v_mfma_f32_16x16x32_{bf16,fp8}
// straight from coalesced row loads: lane l holds 8 elements of candidate
// l / 4, chunk c = l % 4 (elements 8c .. 8c + 7 of the 32-element k-block;
// FP8: of a 16-byte load, one MFMA per half). The MFMA reads lane l's 8
// elements as column l % 16, k-slice l / 16 of B, so column j's k-slices
// come from four different candidates. The query operand A is therefore
// block-diagonal: row i = 4a + c of A holds the query's chunk c in k-slice a
// and zeros elsewhere (lane l carries it when l / 16 == (l / 4) % 4). Then
// C[4a + c][4m + c] = (query chunk c) . (chunk c of candidate m + 4a), which
// the MFMA leaves in register c of lane 16a + 4m + c: lane l holds the chunk
// l % 4 partial of its own candidate l / 4 in register l % 4, accumulated
// over the k-blocks, and the four lanes of a quad sum to q . x. No lane
// permutes are needed between the loads and the MFMA.
//
// FP8 rows narrower than a multiple of 64 bytes load in operand layout
// instead: lane l supplies, for k-block kk, elements 32*kk + 8*(l/16) .. +7
// of the query (A, every row) and of candidate (l % 16) (B), and column n of
// the result holds q . x_n.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

#define VMN_PREC_FP32 0
#define VMN_PREC_BF16 1
#define VMN_PREC_FP8 2
// kernel-internal: BF16 traversal with row norms read from the norms table
#define VMN_PREC_BF16_TABLE_NORMS 3

// Block-diagonal A operand (lane l of A non-zero): see above.
__device__ __forceinline__ bool diag_lane(int l) { return (l >> 4) == ((l >> 2) & 3); }

template <int DP>
struct QLow {
    static constexpr int KB = DP / 32;  // k-blocks of 32
    // FP8 rows load coalesced (block-diagonal A) at multiples of 64 bytes
    static constexpr bool DIAG8 = DP % 64 == 0;
    bf16x8_t a[KB];                     // BF16 query operand (block-diagonal)
    long a8[KB];                        // FP8 query operand (8 x e4m3 per lane)
    float qn;                           // ||q_hat||^2 (BF16) or ||q||^2 (FP8)
    float t;                            // FP8: query scale
};

__device__ __forceinline__ __bf16 to_bf16(float f) { return (__bf16)f; }

// Two floats -> two FP8 (hardware E4M3 encoding) packed in the low 16 bits.
__device__ __forceinline__ uint32_t pk_fp8(float a, float b) {
    return (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false) & 0xffffu;
}

__device__ __forceinline__ float fp8_to_f32(uint32_t word, int byte) {
    switch (byte) {
        case 0: return __builtin_amdgcn_cvt_f32_fp8((int)word, 0);
        case 1: return __builtin_amdgcn_cvt_f32_fp8((int)word, 1);
        case 2: return __builtin_amdgcn_cvt_f32_fp8((int)word, 2);
        default: return __builtin_amdgcn_cvt_f32_fp8((int)word, 3);
    }
}

// Query preparation (whole wave). qf: padded FP32 query row. BF16: q_hat =
// bf16(q), qn = ||q_hat||^2. FP8 (rows stored as e4m3((x_j - c_j) / s_j)
// about the per-dimension centre c): q'_j = (q_j - c_j) * s_j,
// t = max|q'| / 448, A = e4m3(q' / t), qn = ||q - c||^2. The FP8 codes are
// made one element per lane per 64 (j = l + 64 r) and handed to the operand
// layout (8 consecutive codes per lane and MFMA) through `stage`, DP bytes of
// LDS, so only a few values per lane are live at once. qn is summed over the
// natural k-slices (8 elements per lane and k-block) in either layout.
template <int DP, int PREC>
__device__ __forceinline__ void prep_query(QLow<DP>& Q, const float* qf, const float* scales,
                                           const float* centre = nullptr,
                                           uint8_t* stage = nullptr) {
    const int l = lane(), g = l >> 4;
    const bool dg = diag_lane(l);
    if (PREC == VMN_PREC_BF16) {
        // every lane converts chunk l % 4 of each k-block (the operand keeps
        // it in the block-diagonal lanes); a quad's four chunk sums give qn
        float part = 0.0f;
#pragma unroll
        for (int kb = 0; kb < QLow<DP>::KB; ++kb) {
            bf16x8_t h8;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const __bf16 h = to_bf16(qf[32 * kb + 8 * (l & 3) + e]);
                h8[e] = h;
                const float hf = (float)h;
                part += hf * hf;
            }
            Q.a[kb] = dg ? h8 : bf16x8_t{};
        }
        // (c0 + c1) + (c2 + c3) over the quad's chunks
        part += dpp_f<VMN_DPP_XOR1>(part);
        part += dpp_f<VMN_DPP_XOR2>(part);
        Q.qn = part;
        Q.t = 1.0f;
    } else {
        constexpr int NR = (DP + 63) / 64;
        float v[NR];
        float mx = 0.0f, part = 0.0f;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int j = l + 64 * r;
            v[r] = 0.0f;
            if (j < DP) {
                const float u = qf[j] - centre[j];
                v[r] = u * scales[j];
                mx = fmaxf(mx, fabsf(v[r]));
                part += u * u;
            }
        }
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            mx = fmaxf(mx, __shfl_xor(mx, o));
            part += __shfl_xor(part, o);
        }
        const float t = mx > 0.0f ? mx / 448.0f : 1.0f;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int j = l + 64 * r;
            if (j < DP) stage[j] = (uint8_t)(pk_fp8(fminf(fmaxf(v[r] / t, -448.0f), 448.0f), 0.0f) & 0xff);
        }
        __syncthreads();
#pragma unroll
        for (int kb = 0; kb < QLow<DP>::KB; ++kb) {
            if constexpr (QLow<DP>::DIAG8)
                Q.a8[kb] = dg ? *reinterpret_cast<const long*>(stage + 64 * (kb >> 1) + 16 * (l & 3) + 8 * (kb & 1)) : 0l;
            else
                Q.a8[kb] = *reinterpret_cast<const long*>(stage + 32 * kb + 8 * g);
        }
        __syncthreads();
        Q.qn = part;
        Q.t = t;
    }
}

// Moves a dword to the lane that uses it as an MFMA operand: lane L receives
// the value lane src_lane(L) holds.
__device__ __forceinline__ uint32_t bperm(int src_byte, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src_byte, (int)v);
}

// Distances from the prepared query to `cnt` candidates whose ids sit in LDS
// ids[0..cnt); dist_out[s] (LDS) receives the distance of ids[s]. Two tiles
// of 16 candidates are loaded before either is multiplied, so one memory round
// trip serves 32 candidates. FP8 row norms (per-dimension scales) come from
// `cand_norms[s]` (LDS, the norm of ids[s], loaded with the adjacency row the
// ids came from) when given, else from the table `norms` (a cache line per
// candidate). BF16 row norms come from the loaded registers (v_dot2_f32_bf16),
// or, with TN, from `norms` as well: one 4-byte load per candidate issued
// with the row loads. That saves the dot products and two cross-lane sums on a
// latency-bound batch, but costs a cache line per candidate, which a
// bandwidth-bound batch cannot afford.
//
// Loads are coalesced per row: in load step s, lane l reads 16 bytes of tile
// row l/4 at byte 64*s + 16*(l%4), so each 4-lane quad reads 64 contiguous
// bytes of one row (one L1 access, against four when lanes are laid out as
// the MFMA operand wants), and the block-diagonal query operand multiplies
// them where they are (see the section comment). A quad's lanes hold its
// candidate's four chunk partials of the dot product (register l % 4) and of
// the row norm; both are summed (c0 + c1) + (c2 + c3) across the quad, the
// norm in the norms table's order.
template <int DP, int PREC, int NG = 2, bool TN = false>
__device__ __forceinline__ void gather_dists_low(const QLow<DP>& Q, const void* __restrict__ rows,
                                                 const float* __restrict__ norms,
                                                 const int32_t* ids, float* dist_out, int cnt,
                                                 const float* cand_norms = nullptr) {
    const int l = lane(), rr = l >> 2, sub = l & 3;
    constexpr int KB = QLow<DP>::KB;
    // the quad's partial sums: (c0 + c1) + (c2 + c3) in every lane of the quad
    auto quad_sum = [](float v) {
        v += dpp_f<VMN_DPP_XOR1>(v);
        return v + dpp_f<VMN_DPP_XOR2>(v);
    };
    auto own = [&](const f32x4_t& c) {
        return sub == 0 ? c[0] : sub == 1 ? c[1] : sub == 2 ? c[2] : c[3];
    };
    for (int base = 0; base < cnt; base += 16 * NG) {
        int32_t id[NG];
#pragma unroll
        for (int t = 0; t < NG; ++t) {
            const int c = base + 16 * t + rr;
            id[t] = c < cnt ? ids[c] : ids[0];
        }
        f32x4_t acc[NG];
        float x[NG];
#pragma unroll
        for (int t = 0; t < NG; ++t) { acc[t] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f}; x[t] = 0.0f; }
        if (PREC == VMN_PREC_BF16) {
            // step kb holds k-block kb (64 bytes of every row)
            uint4 v[NG][KB];
#pragma unroll
            for (int t = 0; t < NG; ++t) {
                if (t == 0 || base + 16 * t < cnt) {
                    const uint4* r = reinterpret_cast<const uint4*>(reinterpret_cast<const __bf16*>(rows) + (int64_t)id[t] * DP);
#pragma unroll
                    for (int kb = 0; kb < KB; ++kb) v[t][kb] = r[4 * kb + sub];
                    if constexpr (TN) x[t] = norms[id[t]];
                }
            }
#pragma unroll
            for (int t = 0; t < NG; ++t) {
                if (t == 0 || base + 16 * t < cnt) {
#pragma unroll
                    for (int kb = 0; kb < KB; ++kb) {
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Q.a[kb], __builtin_bit_cast(bf16x8_t, v[t][kb]), acc[t], 0, 0, 0);
                        if constexpr (!TN) {
                            const uint32_t w[4] = {v[t][kb].x, v[t][kb].y, v[t][kb].z, v[t][kb].w};
#pragma unroll
                            for (int e = 0; e < 4; ++e) { const bf16x2_t p = __builtin_bit_cast(bf16x2_t, w[e]); x[t] = __builtin_amdgcn_fdot2_f32_bf16(p, p, x[t], false); }
                        }
                    }
                    if constexpr (!TN) x[t] = quad_sum(x[t]);
                }
            }
        } else if constexpr (QLow<DP>::DIAG8) {
            // FP8: step s holds k-blocks 2s, 2s+1 (64 bytes of every row);
            // MFMA kb multiplies the low (kb even) or high half of each
            // lane's 16 bytes
            constexpr int NS = DP / 64;
            uint4 v[NG][NS];
#pragma unroll
            for (int t = 0; t < NG; ++t) {
                if (t == 0 || base + 16 * t < cnt) {
                    const uint4* r = reinterpret_cast<const uint4*>(reinterpret_cast<const uint8_t*>(rows) + (int64_t)id[t] * DP);
#pragma unroll
                    for (int st = 0; st < NS; ++st) v[t][st] = r[4 * st + sub];
                    x[t] = cand_norms ? cand_norms[base + 16 * t + rr] : norms[id[t]];
                }
            }
#pragma unroll
            for (int t = 0; t < NG; ++t) {
                if (t == 0 || base + 16 * t < cnt) {
#pragma unroll
                    for (int kb = 0; kb < KB; ++kb) {
                        const uint4 u = v[t][kb >> 1];
                        const unsigned long op = (kb & 1) ? (((unsigned long)u.w << 32) | u.z)
                                                          : (((unsigned long)u.y << 32) | u.x);
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(Q.a8[kb], (long)op, acc[t], 0, 0, 0);
                    }
                }
            }
        } else {
            // FP8 rows shorter than a multiple of 64 bytes: operand-shaped loads
            const int col = l & 15, g = l >> 4;
            unsigned long v[NG][KB];
            int32_t oid[NG];
#pragma unroll
            for (int t = 0; t < NG; ++t) {
                oid[t] = base + 16 * t + col < cnt ? ids[base + 16 * t + col] : ids[0];
                if (t == 0 || base + 16 * t < cnt) {
                    const unsigned long* r = reinterpret_cast<const unsigned long*>(reinterpret_cast<const uint8_t*>(rows) + (int64_t)oid[t] * DP);
#pragma unroll
                    for (int kb = 0; kb < KB; ++kb) v[t][kb] = r[4 * kb + g];
                    x[t] = cand_norms ? cand_norms[base + 16 * t + col] : norms[oid[t]];
                }
            }
#pragma unroll
            for (int t = 0; t < NG; ++t) {
                if (t == 0 || base + 16 * t < cnt) {
#pragma unroll
                    for (int kb = 0; kb < KB; ++kb)
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(Q.a8[kb], (long)v[t][kb], acc[t], 0, 0, 0);
                }
            }
            // column l % 16 of the result, row 0, in lanes 0..15
            if (g == 0) {
#pragma unroll
                for (int t = 0; t < NG; ++t) {
                    const int c = base + 16 * t + l;
                    if (c < cnt) dist_out[c] = (Q.qn + x[t]) - 2.0f * (acc[t][0] * Q.t);
                }
            }
            continue;
        }
        const float ts = PREC == VMN_PREC_FP8 ? Q.t : 1.0f;
#pragma unroll
        for (int t = 0; t < NG; ++t) {
            if (t == 0 || base + 16 * t < cnt) {
                const float dot = quad_sum(own(acc[t]));
                const int c = base + 16 * t + rr;
                if (sub == 0 && c < cnt) dist_out[c] = (Q.qn + x[t]) - 2.0f * (dot * ts);
            }
        }
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

// Both cells of id x, h1 | h2 << 16 (each below VMN_BF_CELLS < 2^16). The
// FRESH search reads these from a table laid out beside the adjacency
// (adj_hash_kernel) instead of hashing every neighbour on every hop.
__host__ __device__ __forceinline__ uint32_t bf_hash_pair(uint32_t x) {
    return bf_hash1(x) | (bf_hash2(x) << 16);
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
    int rank = 0, ub = 0;
    for (int t = 0; t < nnew; ++t) {
        float dt = readlane_f(my_d, t);
        if (!BY_ID) {
            int c = 0;
#pragma unroll
            for (int j = 0; j < PER; ++j) c += __popcll((unsigned long long)ballot((l + 64 * j) < cnt && bdv[j] <= dt));
            if (l == t) ub = c;
        }
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
    if (!BY_ID && l < nnew) {
        const int pos = ub + rank;
        if (pos < limit) {
            dst_id[pos] = my_id;
            dst_d[pos] = my_d;
        }
    }
    if (BY_ID && l < nnew) {
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

// merge_sorted<CAP, false> for new candidates left in place: the lanes set
// in `km` hold them (my_id / my_d, my_k = fkey(my_d)), in lane order for
// ties. Same merged order as compacting them to lanes 0.. and calling
// merge_sorted, without the compaction's LDS round trip. The candidates are
// visited in lane order t, comparing order-preserving integer keys:
//   * list entries at or below candidate t's key: the list entries placed
//     before it (written into lane t);
//   * list entry j counts the candidates at or above its key (placed after
//     it); the others are placed before it;
//   * candidate c counts the candidates before it: keys below c's, and equal
//     keys from lower lanes. Lane c compares against thr = my_k + 1 (at or
//     below) until t reaches c, when thr drops to my_k (strictly below).
template <int CAP>
__device__ __forceinline__ int merge_masked(const uint32_t* src_id, const float* src_d, int cnt,
                                            uint32_t my_id, float my_d, int my_k, uint64_t km,
                                            uint32_t* dst_id, float* dst_d, int limit) {
    constexpr int PER = (CAP + 63) / 64;
    const int l = lane();
    float bdv[PER];
    uint32_t bidv[PER];
    int bk[PER], ab[PER];
    uint64_t live[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int bi = l + 64 * j;
        bdv[j] = bi < cnt ? src_d[bi] : 0.0f;
        bidv[j] = bi < cnt ? src_id[bi] : 0u;
        bk[j] = fkey(bdv[j]);
        ab[j] = 0;
        const int nl = cnt - 64 * j;
        live[j] = nl >= 64 ? ~0ull : (nl > 0 ? (1ull << nl) - 1ull : 0ull);
    }
    int rank = 0, ub = 0, thr = my_k + 1;
    for (uint64_t m = km; m; m &= m - 1) {
        const int t = __ffsll((unsigned long long)m) - 1;
        const int kt = readlane_i(my_k, t);
        int c = 0;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const bool at_or_below = bk[j] <= kt;
            c += __popcll((unsigned long long)(ballot(at_or_below) & live[j]));
            ab[j] += at_or_below ? 1 : 0;
        }
        ub = writelane_i(c, t, ub);
        thr = writelane_i(kt, t, thr);
        rank += kt < thr ? 1 : 0;
    }
    const int nk = __popcll((unsigned long long)km);
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int bi = l + 64 * j;
        if (bi < cnt) {
            // candidates with keys below entry j's: all but those at or above it
            const int pos = bi + (nk - ab[j]);
            if (pos < limit) {
                dst_id[pos] = bidv[j];
                dst_d[pos] = bdv[j];
            }
        }
    }
    if ((km >> l) & 1ull) {
        const int pos = ub + rank;
        if (pos < limit) {
            dst_id[pos] = my_id;
            dst_d[pos] = my_d;
        }
    }
    const int nc = cnt + nk;
    return nc < limit ? nc : limit;
}

// Element i (uniform) of a list held in registers, entry i in lane i % 64,
// slot i / 64.
template <int PER>
__device__ __forceinline__ float reg_at(const float (&v)[PER], int i) {
    float r = 0.0f;
#pragma unroll
    for (int k = 0; k < PER; ++k)
        if (k == i / 64) r = readlane_f(v[k], i % 64);
    return r;
}

// merge_masked for a list held in registers, entry i in lane i % 64, slot
// i / 64 (entries at or past cnt hold VMN_VIS_BIT): the same positions, the
// merged list scattered to dst (LDS, 64 PER entries) and read back into the
// registers. Returns the new count.
template <int PER>
__device__ __forceinline__ int merge_regs(uint32_t (&r_id)[PER], float (&r_d)[PER], int cnt,
                                          uint32_t my_id, float my_d, int my_k, uint64_t km,
                                          uint32_t* dst_id, float* dst_d, int limit) {
    const int l = lane();
    int bk[PER], ab[PER];
    uint64_t live[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        bk[j] = fkey(r_d[j]);
        ab[j] = 0;
        const int nl = cnt - 64 * j;
        live[j] = nl >= 64 ? ~0ull : (nl > 0 ? (1ull << nl) - 1ull : 0ull);
    }
    int rank = 0, ub = 0, thr = my_k + 1;
    for (uint64_t m = km; m; m &= m - 1) {
        const int t = __ffsll((unsigned long long)m) - 1;
        const int kt = readlane_i(my_k, t);
        int c = 0;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const bool at_or_below = bk[j] <= kt;
            c += __popcll((unsigned long long)(ballot(at_or_below) & live[j]));
            ab[j] += at_or_below ? 1 : 0;
        }
        ub = writelane_i(c, t, ub);
        thr = writelane_i(kt, t, thr);
        rank += kt < thr ? 1 : 0;
    }
    const int nk = __popcll((unsigned long long)km);
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int bi = l + 64 * j;
        if (bi < cnt) {
            // candidates with keys below this entry's: all but those at or above it
            const int pos = bi + (nk - ab[j]);
            if (pos < limit) {
                dst_id[pos] = r_id[j];
                dst_d[pos] = r_d[j];
            }
        }
    }
    if ((km >> l) & 1ull) {
        const int pos = ub + rank;
        if (pos < limit) {
            dst_id[pos] = my_id;
            dst_d[pos] = my_d;
        }
    }
    const int nc = cnt + nk < limit ? cnt + nk : limit;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = l + 64 * j;
        r_id[j] = i < nc ? dst_id[i] : VMN_VIS_BIT;
        r_d[j] = i < nc ? dst_d[i] : 0.0f;
    }
    return nc;
}

}  // namespace vmn
