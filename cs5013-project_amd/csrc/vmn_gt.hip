// Exact k-nearest neighbours in FP64 (the ground truth recall is scored
// against). Replaces the external DiskANN compute_groundtruth the reference
// shells out to (vamana.cu:1644). Each distance is sum_k (double(q_k) -
// double(x_k))^2; equal distances are ordered by id.
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
#pragma unroll
                for (int t = 0; t < DP / 4; ++t) {
                    const float4 xv = x[t];
                    double a = (double)qv[t].x - (double)xv.x;
                    double b = (double)qv[t].y - (double)xv.y;
                    double c = (double)qv[t].z - (double)xv.z;
                    double d = (double)qv[t].w - (double)xv.w;
                    acc = fma(a, a, acc);
                    acc = fma(b, b, acc);
                    acc = fma(c, c, acc);
                    acc = fma(d, d, acc);
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

}  // namespace vmn
