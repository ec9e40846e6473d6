// Exact k-nearest neighbours in FP64 (the ground truth recall is scored
// against). Replaces the external DiskANN compute_groundtruth the reference
// shells out to (vamana.cu:1644). Each distance is sum_k (double(q_k) -
// double(x_k))^2 in index order, every square rounded before it is added (no
// fused multiply-add, as the oracle's FP64 brute force); equal distances are
// ordered by id.
//
// F32 form: the parth engine's own ground truth (parth/src/engine.cu:253-279
// `bruteForceKernel`), a CandidateList<10> over the live ids in increasing
// order with FP32 distances summed as one fused multiply-add chain in index
// order (distance<D>, parth/src/device_funcs.cu:17-49). Ties after existing
// entries and replacement only on a strictly smaller distance make its list
// the first k by (FP32 distance, id), which is the order kept here; the FP32
// value is stored exactly as a double.
//
// Pass 1: one lane per (query, base split) keeps a sorted top-KP list in
// registers while base rows stream through LDS 32 at a time. Pass 2: one
// wavefront per query merges its splits' lists.
#include <algorithm>

#include "vmn_kernels.hpp"

namespace vmn {

template <int DP, int KP, bool F32>
__global__ __launch_bounds__(64) void gt_partial_kernel(const float* __restrict__ base, int64_t n,
                                                        const uint32_t* del_bits,
                                                        const float* __restrict__ queries,
                                                        int64_t nq, int splits, double* part_d,
                                                        int32_t* part_i) {
    constexpr int TB = 32;
    __shared__ __align__(16) float tile[TB * DP];
    const int l = threadIdx.x;
    const int64_t qi = (int64_t)blockIdx.x * 64 + l;
    const int split = blockIdx.y;
    const int64_t per = (n + splits - 1) / splits;
    const int64_t r0 = split * per, r1 = r0 + per < n ? r0 + per : n;
    float4 qv[DP / 4];
    const bool active = qi < nq;
#pragma unroll
    for (int t = 0; t < DP / 4; ++t)
        qv[t] = active ? reinterpret_cast<const float4*>(queries + qi * DP)[t] : make_float4(0, 0, 0, 0);
    double bd[KP];
    int32_t bi[KP];
#pragma unroll
    for (int j = 0; j < KP; ++j) {
        bd[j] = INFINITY;
        bi[j] = -1;
    }
    for (int64_t t0 = r0; t0 < r1; t0 += TB) {
        const int rows = (int)(r1 - t0 < TB ? r1 - t0 : TB);
        __syncthreads();
        for (int e = l; e < rows * DP / 4; e += 64)
            reinterpret_cast<float4*>(tile)[e] = reinterpret_cast<const float4*>(base + t0 * DP)[e];
        __syncthreads();
        for (int r = 0; r < rows; ++r) {
            const int64_t id = t0 + r;
            if (del_bits && bit_test(del_bits, (uint32_t)id)) continue;
            double acc = 0.0;
            const float4* x = reinterpret_cast<const float4*>(tile + r * DP);
            if constexpr (F32) {
                float f = 0.0f;
#pragma unroll
                for (int t = 0; t < DP / 4; ++t) {
                    const float4 xv = x[t];
                    const float a = qv[t].x - xv.x, b = qv[t].y - xv.y;
                    const float c = qv[t].z - xv.z, d = qv[t].w - xv.w;
                    f = __builtin_fmaf(a, a, f);
                    f = __builtin_fmaf(b, b, f);
                    f = __builtin_fmaf(c, c, f);
                    f = __builtin_fmaf(d, d, f);
                }
                acc = (double)f;
            } else {
                // every square rounded before it is added
#pragma clang fp contract(off)
#pragma unroll
                for (int t = 0; t < DP / 4; ++t) {
                    const float4 xv = x[t];
                    double a = (double)qv[t].x - (double)xv.x;
                    double b = (double)qv[t].y - (double)xv.y;
                    double c = (double)qv[t].z - (double)xv.z;
                    double d = (double)qv[t].w - (double)xv.w;
                    acc += a * a;
                    acc += b * b;
                    acc += c * c;
                    acc += d * d;
                }
            }
            if (acc < bd[KP - 1]) {
                // insert keeping (dist, id) order; ids arrive increasing
                double cd = acc;
                int32_t ci = (int32_t)id;
#pragma unroll
                for (int j = 0; j < KP; ++j) {
                    if (cd < bd[j]) {
                        double td = bd[j];
                        int32_t ti = bi[j];
                        bd[j] = cd;
                        bi[j] = ci;
                        cd = td;
                        ci = ti;
                    }
                }
            }
        }
    }
    if (active) {
        const int64_t o = (qi * splits + split) * KP;
#pragma unroll
        for (int j = 0; j < KP; ++j) {
            part_d[o + j] = bd[j];
            part_i[o + j] = bi[j];
        }
    }
}

// One wavefront per query: k rounds, each taking the smallest (dist, id)
// after the previous pick; every lane scans a strided share of the query's
// split lists and a butterfly over the wave keeps the minimum.
template <int KP>
__global__ __launch_bounds__(64) void gt_merge_kernel(const double* part_d, const int32_t* part_i,
                                                      int64_t nq, int splits, int k,
                                                      int32_t* out_ids, double* out_d) {
    const int64_t qi = blockIdx.x;
    if (qi >= nq) return;
    const int l = threadIdx.x;
    const int64_t o = qi * splits * KP;
    const int total = splits * KP;
    double last_d = -INFINITY;
    int32_t last_i = -1;
    for (int j = 0; j < k; ++j) {
        double bd = INFINITY;
        int32_t bi = -1;
        for (int c = l; c < total; c += 64) {
            const double d = part_d[o + c];
            const int32_t id = part_i[o + c];
            if (id < 0) continue;
            const bool after_last = d > last_d || (d == last_d && id > last_i);
            const bool better = bi < 0 || d < bd || (d == bd && id < bi);
            if (after_last && better) {
                bd = d;
                bi = id;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double od = __shfl_xor(bd, off);
            const int32_t oi = __shfl_xor(bi, off);
            if (oi >= 0 && (bi < 0 || od < bd || (od == bd && oi < bi))) {
                bd = od;
                bi = oi;
            }
        }
        if (l == 0) {
            out_ids[qi * k + j] = bi;
            if (out_d) out_d[qi * k + j] = bd;
        }
        if (bi >= 0) {
            last_d = bd;
            last_i = bi;
        }
    }
}

template <int DP, int KP>
static void gt_launch(const float* base, int64_t n, const uint32_t* del_bits, const float* queries,
                      int64_t nq, int k, int splits, double* part_d, int32_t* part_i,
                      int32_t* out_ids, double* out_d, bool f32, hipStream_t s) {
    dim3 g((unsigned)((nq + 63) / 64), (unsigned)splits);
    if (f32)
        hipLaunchKernelGGL((gt_partial_kernel<DP, KP, true>), g, dim3(64), 0, s, base, n, del_bits,
                           queries, nq, splits, part_d, part_i);
    else
        hipLaunchKernelGGL((gt_partial_kernel<DP, KP, false>), g, dim3(64), 0, s, base, n, del_bits,
                           queries, nq, splits, part_d, part_i);
    hipLaunchKernelGGL((gt_merge_kernel<KP>), dim3((unsigned)nq), dim3(64), 0, s, part_d, part_i,
                       nq, splits, k, out_ids, out_d);
}

hipError_t launch_gt(const float* base, int64_t n, int DP, const uint32_t* del_bits,
                     const float* queries, int64_t nq, int k, int splits, double* part_d,
                     int32_t* part_i, int32_t* out_ids, double* out_d, hipStream_t s, bool f32) {
#define VMN_GT_CASE(DPV)                                                                            \
    case DPV:                                                                                       \
        if (k <= 16) gt_launch<DPV, 16>(base, n, del_bits, queries, nq, k, splits, part_d, part_i, \
                                        out_ids, out_d, f32, s);                                    \
        else gt_launch<DPV, 128>(base, n, del_bits, queries, nq, k, splits, part_d, part_i,        \
                                 out_ids, out_d, f32, s);                                           \
        break;
    if (k > 128) return hipErrorInvalidValue;
    switch (DP) {
        VMN_GT_CASE(32)
        VMN_GT_CASE(64)
        VMN_GT_CASE(96)
        VMN_GT_CASE(128)
        VMN_GT_CASE(160)
        VMN_GT_CASE(192)
        VMN_GT_CASE(256)
        default: return hipErrorInvalidValue;
    }
#undef VMN_GT_CASE
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The same FP64 ground truth, filtered on MFMA.
//
// Pass 1 (gt_filter_kernel) scores every (query, row) pair with the FP32
// matrix form d~ = ||q||^2 + ||x||^2 - 2 q.x, q.x from v_mfma_f32_16x16x4_f32,
// and keeps, per query and per (row split, lane column), the 4 rows of least
// d~. |d~ - d| <= E_q = 4 DP 2^-24 (||q|| + M)^2 for the exact distance d,
// M = max_x ||x|| (the dot product's summation error is at most DP 2^-24
// ||q|| ||x|| in any order, doubled by the factor 2, and the norms' and the
// final additions' roundings are below the rest). Pass 2 (gt_select_kernel)
// takes T = the k-th least d~ + E_q over the kept rows: at least k rows have
// d <= T. A row left out of every list has d~ at least the least 4th entry t
// of the lists, so when t - E_q > T' (T' = T widened by the relative rounding
// of FP64 distances) it cannot be among the first k by FP64 distance. The
// kept rows with d~ - E_q <= T' then get the exact FP64 distance of
// gt_partial_kernel (same operations, same order) and the first k by
// (distance, id) are the result, identical to launch_gt's. A query whose
// check fails (near ties in bulk, non-finite values, more than 64 rows to
// recompute) is listed and run through launch_gt instead.
// ---------------------------------------------------------------------------

#define VMN_GTF_QPW 16    // queries per wavefront (one MFMA A tile)
#define VMN_GTF_WAVES 8   // wavefronts per workgroup: 128 queries share each row tile
#define VMN_GTF_ROWS 64   // rows per LDS tile
#define VMN_GTF_KEEP 4    // rows kept per (query, split, lane column)

__global__ void gt_row_norms_kernel(const float* __restrict__ base, int64_t n, int DP,
                                    const uint32_t* del_bits, float* xn, unsigned int* max_norm_bits) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double s = 0.0;
    const float4* x = reinterpret_cast<const float4*>(base + r * DP);
    for (int t = 0; t < DP / 4; ++t) {
        const float4 v = x[t];
        s += (double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z + (double)v.w * v.w;
    }
    xn[r] = (float)s;
    if (del_bits && bit_test(del_bits, (uint32_t)r)) return;
    // an upper bound of ||x||, as non-negative float bits (ordered like the values)
    atomicMax(max_norm_bits, __float_as_uint(__double2float_ru(sqrt(s))));
}

template <int DP>
__global__ __launch_bounds__(64 * VMN_GTF_WAVES) void gt_filter_kernel(
    const float* __restrict__ base, int64_t n, const uint32_t* del_bits, const float* __restrict__ xn,
    const float* __restrict__ queries, int64_t nq, float* cand_d, int32_t* cand_i) {
    constexpr int KS = DP / 4;   // MFMA k-steps of 4 dimensions
    constexpr int RS = DP + 4;   // LDS row stride in floats (16 consecutive rows hit distinct banks)
    constexpr int TR = VMN_GTF_ROWS;
    constexpr int KEEP = VMN_GTF_KEEP;
    __shared__ __align__(16) float tile[TR * RS];
    __shared__ float tn[TR];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, m = l & 15, kk = l >> 4;
    const int S = gridDim.y, split = blockIdx.y;
    const int64_t q0 = ((int64_t)blockIdx.x * VMN_GTF_WAVES + w) * VMN_GTF_QPW;
    // A operand of step s: query q0 + m, dimension KS kk + s (the B operand
    // pairs the same dimension of row m of the row tile)
    float a[KS];
    {
        const int64_t qa = q0 + m;
        const float4* qp = reinterpret_cast<const float4*>(queries + (qa < nq ? qa : 0) * DP + KS * kk);
#pragma unroll
        for (int s4 = 0; s4 < KS / 4; ++s4) {
            const float4 v = qa < nq ? qp[s4] : make_float4(0.f, 0.f, 0.f, 0.f);
            a[4 * s4] = v.x; a[4 * s4 + 1] = v.y; a[4 * s4 + 2] = v.z; a[4 * s4 + 3] = v.w;
        }
    }
    // the C rows this lane holds: queries q0 + 4 kk + i
    float qn[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t qi = q0 + 4 * kk + i;
        double sum = 0.0;
        if (qi < nq)
            for (int d = 0; d < DP; ++d) {
                const double v = queries[qi * DP + d];
                sum += v * v;
            }
        qn[i] = (float)sum;
    }
    float bd[4][KEEP];
    int32_t bi[4][KEEP];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < KEEP; ++j) {
            bd[i][j] = INFINITY;
            bi[i][j] = -1;
        }
    const int64_t per = (n + S - 1) / S;
    const int64_t r0 = (int64_t)split * per, r1 = r0 + per < n ? r0 + per : n;
    for (int64_t t0 = r0; t0 < r1; t0 += TR) {
        const int rows = (int)(r1 - t0 < TR ? r1 - t0 : TR);
        __syncthreads();
        for (int e = threadIdx.x; e < TR * (DP / 4); e += 64 * VMN_GTF_WAVES) {
            const int r = e / (DP / 4), c4 = e % (DP / 4);
            const float4 v = r < rows ? reinterpret_cast<const float4*>(base + (t0 + r) * DP)[c4]
                                      : make_float4(0.f, 0.f, 0.f, 0.f);
            *reinterpret_cast<float4*>(tile + r * RS + 4 * c4) = v;
        }
        for (int e = threadIdx.x; e < TR; e += 64 * VMN_GTF_WAVES)
            tn[e] = e < rows && !(del_bits && bit_test(del_bits, (uint32_t)(t0 + e))) ? xn[t0 + e]
                                                                                     : INFINITY;
        __syncthreads();
        for (int sub = 0; sub < TR / 16 && 16 * sub < rows; ++sub) {
            f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
            const float* brow = tile + (16 * sub + m) * RS + KS * kk;
#pragma unroll
            for (int s4 = 0; s4 < KS / 4; ++s4) {
                const float4 b = *reinterpret_cast<const float4*>(brow + 4 * s4);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[4 * s4], b.x, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[4 * s4 + 1], b.y, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[4 * s4 + 2], b.z, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[4 * s4 + 3], b.w, acc, 0, 0, 0);
            }
            // lane l holds query q0 + 4 kk + i against row t0 + 16 sub + m
            const int rr = 16 * sub + m;
            const float xr = tn[rr];
            const int32_t id = (int32_t)(t0 + rr);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float cd = (qn[i] + xr) - 2.0f * acc[i];
                if (cd < bd[i][KEEP - 1]) {
                    int32_t ci = id;
#pragma unroll
                    for (int j = 0; j < KEEP; ++j) {
                        if (cd < bd[i][j]) {
                            const float td = bd[i][j];
                            const int32_t ti = bi[i][j];
                            bd[i][j] = cd;
                            bi[i][j] = ci;
                            cd = td;
                            ci = ti;
                        }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t qi = q0 + 4 * kk + i;
        if (qi >= nq) continue;
        const int64_t o = ((qi * S + split) * 16 + m) * KEEP;
#pragma unroll
        for (int j = 0; j < KEEP; ++j) {
            cand_d[o + j] = bd[i][j];
            cand_i[o + j] = bi[i][j];
        }
    }
}

// Exact FP64 distance of gt_partial_kernel (same operations in the same order).
template <int DP>
__device__ __forceinline__ double gt_exact_dist(const float* q, const float* x) {
#pragma clang fp contract(off)
    double acc = 0.0;
    const float4* qq = reinterpret_cast<const float4*>(q);
    const float4* xx = reinterpret_cast<const float4*>(x);
#pragma unroll 8
    for (int t = 0; t < DP / 4; ++t) {
        const float4 qv = qq[t], xv = xx[t];
        const double a = (double)qv.x - (double)xv.x, b = (double)qv.y - (double)xv.y;
        const double c = (double)qv.z - (double)xv.z, d = (double)qv.w - (double)xv.w;
        acc += a * a;
        acc += b * b;
        acc += c * c;
        acc += d * d;
    }
    return acc;
}

// One wavefront per query: the check and the exact first k of pass 2.
template <int DP>
__global__ __launch_bounds__(64) void gt_select_kernel(
    const float* __restrict__ base, const float* __restrict__ queries, int64_t nq, int S,
    const float* cand_d, const int32_t* cand_i, const unsigned int* max_norm_bits, int k,
    int32_t* out_ids, double* out_d, int32_t* fb_list, unsigned int* fb_count) {
    constexpr int KEEP = VMN_GTF_KEEP;
    __shared__ int32_t s_id[64];
    __shared__ double s_d[64];
    const int64_t qi = blockIdx.x;
    if (qi >= nq) return;
    const int l = threadIdx.x;
    const float* q = queries + qi * DP;
    double qq = 0.0;
    for (int d = l; d < DP; d += 64) qq += (double)q[d] * q[d];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) qq += __shfl_xor(qq, o);
    const double M = (double)__uint_as_float(*max_norm_bits);
    const double r = sqrt(qq) + M;
    const double E = 4.0 * DP * 0x1p-24 * r * r * (1.0 + 0x1p-20);
    const int total = S * 16 * KEEP;
    const float* cd = cand_d + qi * total;
    const int32_t* ci = cand_i + qi * total;
    // t: the least 4th entry over the lists (a list that is not full keeps all its rows)
    float tmin = INFINITY;
    for (int c = l; c < S * 16; c += 64) tmin = fminf(tmin, cd[c * KEEP + KEEP - 1]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) tmin = fminf(tmin, __shfl_xor(tmin, o));
    // the k-th least d~ (k rounds of the wave minimum over (d~, slot))
    float last = -INFINITY;
    int last_c = -1;
    bool ok = true;
    for (int j = 0; j < k; ++j) {
        float bdv = INFINITY;
        int bc = 0x7fffffff;
        for (int c = l; c < total; c += 64) {
            const float v = cd[c];
            if (ci[c] < 0) continue;
            const bool after = v > last || (v == last && c > last_c);
            if (after && (v < bdv || (v == bdv && c < bc))) {
                bdv = v;
                bc = c;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float ov = __shfl_xor(bdv, o);
            const int oc = __shfl_xor(bc, o);
            if (ov < bdv || (ov == bdv && oc < bc)) {
                bdv = ov;
                bc = oc;
            }
        }
        if (bc == 0x7fffffff) {
            ok = false;  // fewer than k rows
            break;
        }
        last = bdv;
        last_c = bc;
    }
    const double T = (double)last + E;
    const double Tw = T + fabs(T) * 0x1p-40 + 0x1p-1000;
    ok = ok && ((double)tmin - E > Tw);
    // the rows to recompute: d~ - E <= T'
    int cnt = 0;
    if (ok) {
        for (int c0 = 0; c0 < total; c0 += 64) {
            const int c = c0 + l;
            const bool take = c < total && ci[c] >= 0 && (double)cd[c] - E <= Tw;
            const uint64_t mk = __ballot(take);
            const int slot = cnt + __popcll(mk & ((1ull << l) - 1ull));
            if (take && slot < 64) s_id[slot] = ci[c];
            cnt += __popcll(mk);
        }
        ok = cnt >= k && cnt <= 64;
    }
    if (!ok) {
        if (l == 0) fb_list[atomicAdd(fb_count, 1u)] = (int32_t)qi;
        return;
    }
    __syncthreads();
    const int32_t my_id = l < cnt ? s_id[l] : -1;
    const double my_d = l < cnt ? gt_exact_dist<DP>(q, base + (int64_t)my_id * DP) : INFINITY;
    double last_d = -INFINITY;
    int32_t last_i = -1;
    for (int j = 0; j < k; ++j) {
        double bdv = INFINITY;
        int32_t bidv = -1;
        const bool after = my_id >= 0 && (my_d > last_d || (my_d == last_d && my_id > last_i));
        if (after) {
            bdv = my_d;
            bidv = my_id;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double od = __shfl_xor(bdv, o);
            const int32_t oi = __shfl_xor(bidv, o);
            if (oi >= 0 && (bidv < 0 || od < bdv || (od == bdv && oi < bidv))) {
                bdv = od;
                bidv = oi;
            }
        }
        if (l == 0) {
            out_ids[qi * k + j] = bidv;
            if (out_d) out_d[qi * k + j] = bdv;
        }
        last_d = bdv;
        last_i = bidv;
    }
    (void)s_d;
}

__global__ void gt_gather_rows_kernel(const float* __restrict__ src, const int32_t* rows, int64_t m,
                                      int DP, float* dst) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * DP) return;
    dst[e] = src[(int64_t)rows[e / DP] * DP + e % DP];
}

__global__ void gt_scatter_kernel(const int32_t* rows, int64_t m, int k, const int32_t* ids,
                                  const double* d, int32_t* out_ids, double* out_d) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * k) return;
    const int64_t q = rows[e / k];
    out_ids[q * k + e % k] = ids[e];
    if (out_d) out_d[q * k + e % k] = d[e];
}

int gt_fast_splits(int64_t n, int64_t nq, int cus) {
    const int64_t blocks = (nq + VMN_GTF_QPW * VMN_GTF_WAVES - 1) / (VMN_GTF_QPW * VMN_GTF_WAVES);
    int64_t S = (4 * (int64_t)cus + blocks - 1) / blocks;
    S = std::max<int64_t>(1, std::min<int64_t>(S, 64));
    return (int)std::min<int64_t>(S, std::max<int64_t>(1, n / VMN_GTF_ROWS));
}

size_t gt_fast_scratch_bytes(int64_t n, int64_t nq, int S) {
    const size_t cand = (size_t)nq * S * 16 * VMN_GTF_KEEP;
    return cand * (sizeof(float) + sizeof(int32_t)) + (size_t)n * sizeof(float) +
           (size_t)nq * sizeof(int32_t) + 256;
}

template <int DP>
static hipError_t gt_fast_launch(const float* base, int64_t n, const uint32_t* del_bits,
                                 const float* queries, int64_t nq, int k, int S, float* cand_d,
                                 int32_t* cand_i, float* xn, unsigned int* misc, int32_t* fb_list,
                                 int32_t* out_ids, double* out_d, hipStream_t s) {
    hipLaunchKernelGGL(gt_row_norms_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, base, n,
                       DP, del_bits, xn, misc);
    const unsigned qb = (unsigned)((nq + VMN_GTF_QPW * VMN_GTF_WAVES - 1) / (VMN_GTF_QPW * VMN_GTF_WAVES));
    hipLaunchKernelGGL((gt_filter_kernel<DP>), dim3(qb, (unsigned)S), dim3(64 * VMN_GTF_WAVES), 0, s,
                       base, n, del_bits, xn, queries, nq, cand_d, cand_i);
    hipLaunchKernelGGL((gt_select_kernel<DP>), dim3((unsigned)nq), dim3(64), 0, s, base, queries, nq,
                       S, cand_d, cand_i, misc, k, out_ids, out_d, fb_list, misc + 1);
    return hipGetLastError();
}

hipError_t launch_gt_fast(const float* base, int64_t n, int DP, const uint32_t* del_bits,
                          const float* queries, int64_t nq, int k, int S, void* scratch,
                          int32_t* out_ids, double* out_d, unsigned int* fb_count_dev,
                          int32_t** fb_list_out, hipStream_t s) {
    if (k < 1 || k > 64 || nq < 1 || n < 1) return hipErrorInvalidValue;
    const size_t cand = (size_t)nq * S * 16 * VMN_GTF_KEEP;
    char* p = static_cast<char*>(scratch);
    float* cand_d = reinterpret_cast<float*>(p);
    int32_t* cand_i = reinterpret_cast<int32_t*>(p + cand * sizeof(float));
    float* xn = reinterpret_cast<float*>(p + cand * (sizeof(float) + sizeof(int32_t)));
    int32_t* fb_list = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(xn) + (size_t)n * sizeof(float));
    *fb_list_out = fb_list;
    // misc[0]: max-norm bits, misc[1]: fallback count (cleared by a kernel)
    unsigned int* misc = fb_count_dev;
    const hipError_t e = launch_fill_i32(reinterpret_cast<int32_t*>(misc), 2, 0, s);
    if (e != hipSuccess) return e;
    switch (DP) {
        case 32: return gt_fast_launch<32>(base, n, del_bits, queries, nq, k, S, cand_d, cand_i, xn, misc, fb_list, out_ids, out_d, s);
        case 64: return gt_fast_launch<64>(base, n, del_bits, queries, nq, k, S, cand_d, cand_i, xn, misc, fb_list, out_ids, out_d, s);
        case 96: return gt_fast_launch<96>(base, n, del_bits, queries, nq, k, S, cand_d, cand_i, xn, misc, fb_list, out_ids, out_d, s);
        case 128: return gt_fast_launch<128>(base, n, del_bits, queries, nq, k, S, cand_d, cand_i, xn, misc, fb_list, out_ids, out_d, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_gt_gather(const float* src, const int32_t* rows, int64_t m, int DP, float* dst,
                            hipStream_t s) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(gt_gather_rows_kernel, dim3((unsigned)((m * DP + 255) / 256)), dim3(256), 0, s,
                       src, rows, m, DP, dst);
    return hipGetLastError();
}

hipError_t launch_gt_scatter(const int32_t* rows, int64_t m, int k, const int32_t* ids, const double* d,
                             int32_t* out_ids, double* out_d, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(gt_scatter_kernel, dim3((unsigned)((m * k + 255) / 256)), dim3(256), 0, s, rows,
                       m, k, ids, d, out_ids, out_d);
    return hipGetLastError();
}

}  // namespace vmn
