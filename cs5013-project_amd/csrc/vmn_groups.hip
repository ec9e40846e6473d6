// Query groups for the search kernel: locality per XCD.
//
// MI355X has 8 XCDs, each with its own L2, and workgroups are dispatched to
// them round-robin (workgroup b runs on XCD b % 8). A large batch's queries
// are split into one group per XCD by the region of the space they fall in,
// so the rows and adjacency lists one L2 holds serve the queries that XCD
// runs: each group's searches walk fewer distinct parts of the graph. The
// region is the nearest of C centroids (k-means over a sample of the index's
// rows, computed once per index), and centroid c belongs to group c * G / C.
// This file computes the centroids and lists the batch's queries: one list per
// centroid, in centroid order, so a group's lists are consecutive. The search
// kernel's workgroups of one XCD walk the lists cyclically from their group's
// first, together, so at any moment they search few regions; they go on to the
// next groups' lists once their own are empty, so the groups need not be
// balanced. (Batches of more than 65,536 queries, a build's, keep per-group
// lists in a bounded layout instead: see list_layout.) The grouping decides
// only the order in which queries are searched: each query's search, and so
// every result, is the same in any order.
#include <hip/hip_runtime.h>

#include <vector>

#include "vmn_internal.hpp"
#include "vmn_device.hpp"
#include "vmn_kernels.hpp"

using namespace vmn_rt;

namespace vmn {

namespace {

// Nearest centroid of each of `nrows` rows ([nrows x DP] FP32, KB = DP / 32)
// on MFMA: score ||c||^2 - 2 r.c with BF16 rows and centroids (a scheduling
// choice needs no more), lowest centroid index on ties. Each wavefront takes
// 16 rows. Lane l holds row l % 16's k-slice l / 16 of every k-block (the B
// operand of v_mfma_f32_16x16x32_bf16) and, per tile of 16 centroids, loads
// centroid l % 16's (the A operand); the product tile D[m][n] (centroid m,
// row n) sits in lane 16 (m / 4) + n, element m % 4, so every lane follows one
// row and the four lanes of a row meet in two exchanges. Four centroid tiles
// are accumulated at once, k-blocks outermost. With keys_out the centroid
// index is written per row. Otherwise the row is appended to list L of the
// batch (entries order[L * stride + ...], count at heads[32 L + 1]): L is the
// centroid (by_centroid), or, in the bounded layout, L = g S + b % S for group
// g = index * G / C and workgroup b, S sub-lists per group. A one-wavefront
// workgroup (WAVES = 1; it starts beside a persistent search still holding
// the GPU) adds each list's rows with one atomic, the lists' atomics in flight
// together; WAVES = 16 (bounded layout) counts its 256 rows per group in LDS
// first, for batches so large that even those would queue.
template <int KB, int WAVES>
__device__ __forceinline__ void assign_body(const float* __restrict__ rows, int64_t nrows,
                                            const __bf16* __restrict__ cent,
                                            const float* __restrict__ cnorm, int C,
                                            int32_t* keys_out, int32_t* order,
                                            unsigned long long* heads, int G, int S,
                                            int64_t gstride, bool by_centroid) {
    constexpr int DP = 32 * KB;
    __shared__ int s_cnt[WAVES > 1 ? 32 : 1];
    __shared__ unsigned long long s_base[WAVES > 1 ? 32 : 1];
    const int w = WAVES > 1 ? (int)(threadIdx.x >> 6) : 0;
    const int l = threadIdx.x & 63, g = l >> 4, col = l & 15;
    if (WAVES > 1 && threadIdx.x < 32) s_cnt[threadIdx.x] = 0;
    const int64_t r0 = ((int64_t)blockIdx.x * WAVES + w) * 16;
    const int64_t brow = r0 + col;
    float best = INFINITY;
    int bid = 0x7fffffff;
    // centroid tiles of 4 at a time, k-blocks outermost: the row's k-slice is
    // loaded once per k-block and four accumulators stay live
    for (int c0 = 0; c0 < C; c0 += 64) {
        f32x4_t acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            float4 f0 = make_float4(0.f, 0.f, 0.f, 0.f), f1 = f0;
            if (brow < nrows) {
                const float4* p = reinterpret_cast<const float4*>(rows + brow * DP + 32 * kb + 8 * g);
                f0 = p[0];
                f1 = p[1];
            }
            const bf16x8_t b = bf16x8_t{to_bf16(f0.x), to_bf16(f0.y), to_bf16(f0.z), to_bf16(f0.w),
                                        to_bf16(f1.x), to_bf16(f1.y), to_bf16(f1.z), to_bf16(f1.w)};
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int ca = c0 + 16 * t + col;  // the A row this lane loads
                const bf16x8_t a = ca < C ? *reinterpret_cast<const bf16x8_t*>(cent + (int64_t)ca * DP + 32 * kb + 8 * g)
                                          : bf16x8_t{};
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int c = c0 + 16 * t + 4 * g + i;
                if (c < C) {
                    const float sc = __builtin_fmaf(-2.0f, acc[t][i], cnorm[c]);
                    if (sc < best || (sc == best && c < bid)) {
                        best = sc;
                        bid = c;
                    }
                }
            }
    }
    // lanes n, n + 16, n + 32, n + 48 hold row n against centroids 4 g + i
#pragma unroll
    for (int o = 16; o < 64; o <<= 1) {
        const float ob = __shfl_xor(best, o);
        const int oi = __shfl_xor(bid, o);
        if (ob < best || (ob == best && oi < bid)) {
            best = ob;
            bid = oi;
        }
    }
    // a row with no finite score (NaN components) goes to centroid 0
    const int key = bid >= 0 && bid < C ? bid : 0;
    const int64_t r = r0 + l;
    const bool valid = l < 16 && r < nrows;
    if (keys_out) {
        if (valid) keys_out[r] = key;
        return;
    }
    const int gr = (int)((int64_t)key * G / C);
    const int sub = (int)(blockIdx.x % (unsigned)S);
    if constexpr (WAVES == 1) {
        // rows of one list meet in a mask; its lowest lane adds them with one
        // atomic, the lists' atomics in flight together
        const int Lr = !valid ? -1 : by_centroid ? key : gr * S + sub;
        uint64_t mine = 0;
        for (int j = 0; j < 16; ++j) {
            const int Lj = __shfl(Lr, j);
            const uint64_t m = __ballot(valid && Lr == Lj);
            if (l == j) mine = m;
        }
        const int lead = valid ? __ffsll((unsigned long long)mine) - 1 : 0;
        unsigned long long base = 0;
        if (valid && lead == l) base = atomicAdd(heads + 32 * Lr + 1, (unsigned long long)__popcll(mine));
        base = __shfl(base, lead);
        if (valid)
            order[(int64_t)Lr * gstride + (int64_t)base + __popcll(mine & ((1ull << l) - 1ull))] = (int32_t)r;
    } else {
        __syncthreads();
        const int local = valid ? atomicAdd(&s_cnt[gr], 1) : 0;
        __syncthreads();
        if (threadIdx.x < G && s_cnt[threadIdx.x] > 0)
            s_base[threadIdx.x] = atomicAdd(heads + 32 * (threadIdx.x * S + sub) + 1,
                                            (unsigned long long)s_cnt[threadIdx.x]);
        __syncthreads();
        if (valid) order[(int64_t)(gr * S + sub) * gstride + (int64_t)s_base[gr] + local] = (int32_t)r;
    }
}

#define VMN_ASSIGN_ARGS                                                                            \
    const float *__restrict__ rows, int64_t nrows, const __bf16 *__restrict__ cent,                \
        const float *__restrict__ cnorm, int C, int32_t *keys_out, int32_t *order,                 \
        unsigned long long *heads, int G, int S, int64_t gstride, bool by_centroid
#define VMN_ASSIGN_PASS rows, nrows, cent, cnorm, C, keys_out, order, heads, G, S, gstride, by_centroid
template <int KB>
__global__ __launch_bounds__(64) void assign_kernel_one(VMN_ASSIGN_ARGS) {
    assign_body<KB, 1>(VMN_ASSIGN_PASS);
}
template <int KB>
__global__ __launch_bounds__(1024) void assign_kernel_wide(VMN_ASSIGN_ARGS) {
    assign_body<KB, 16>(VMN_ASSIGN_PASS);
}
#undef VMN_ASSIGN_ARGS
#undef VMN_ASSIGN_PASS

// Sample row i = row (i * n) / S of `vecs`.
__global__ void sample_rows_kernel(const float* __restrict__ vecs, int64_t n, int DP, int S,
                                   float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)S * DP) return;
    const int64_t s = i / DP, d = i % DP;
    out[i] = vecs[(s * n / S) * DP + d];
}

// Centroid c = mean of the sample rows with key c, summed in row order (the
// result does not depend on the launch), stored in BF16 with the squared norm
// of the stored values; a centroid with no rows keeps its place. One
// workgroup per centroid: its threads find the members of contiguous key
// ranges, a scan orders them, and thread d sums dimension d.
__global__ __launch_bounds__(256) void kmeans_update_kernel(const float* __restrict__ smp, int S, int DP,
                                                            const int32_t* __restrict__ keys,
                                                            __bf16* cent, float* cnorm) {
    __shared__ int32_t s_mem[4096];
    __shared__ int s_cnt[256];
    __shared__ float s_red[256];
    const int c = blockIdx.x, t = threadIdx.x;
    const int per = (S + 255) / 256;
    const int lo = t * per, hi = min(S, lo + per);
    int m = 0;
    for (int i = lo; i < hi; ++i) m += keys[i] == c;
    s_cnt[t] = m;
    __syncthreads();
    if (t == 0) {
        int run = 0;
        for (int j = 0; j < 256; ++j) {
            const int v = s_cnt[j];
            s_cnt[j] = run;
            run += v;
        }
        s_red[0] = (float)run;
    }
    __syncthreads();
    const int total = (int)s_red[0];
    const int nm = total < 4096 ? total : 4096;  // members used: the first 4096 in row order
    int pos = s_cnt[t];
    for (int i = lo; i < hi && pos < 4096; ++i)
        if (keys[i] == c) s_mem[pos++] = i;
    __syncthreads();
    float part = 0.0f;
    for (int d = t; d < DP; d += 256) {
        __bf16 v = cent[(int64_t)c * DP + d];
        if (nm > 0) {
            float sum = 0.0f;
            for (int j = 0; j < nm; ++j) sum += smp[(int64_t)s_mem[j] * DP + d];
            v = to_bf16(sum / (float)nm);
            cent[(int64_t)c * DP + d] = v;
        }
        const float vf = (float)v;
        part = __builtin_fmaf(vf, vf, part);
    }
    __syncthreads();
    s_red[t] = part;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) s_red[t] += s_red[t + o];
        __syncthreads();
    }
    if (t == 0) cnorm[c] = s_red[0];
}

__global__ void cent_init_kernel(const float* __restrict__ smp, int S, int DP, int C, __bf16* cent) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)C * DP) return;
    const int64_t c = i / DP, d = i % DP;
    cent[i] = to_bf16(smp[(c * S / C) * DP + d]);
}

}  // namespace

// Lists of a batch of nrows queries: one per centroid (C lists of nrows
// entries each), or past 65,536 queries G x S sub-lists bounded by the rows of
// the workgroups (of 256 rows) that append to each.
constexpr int64_t kCentroidListsMax = 65536;

GroupLayout list_layout(int64_t nrows, int C, int G, int S) {
    GroupLayout y;
    y.by_centroid = nrows <= kCentroidListsMax;
    if (y.by_centroid) {
        y.nlists = C;
        y.stride = nrows;
    } else {
        const int64_t blocks = (nrows + 255) / 256;
        y.nlists = G * S;
        y.stride = (blocks + S - 1) / S * 256;
    }
    return y;
}

hipError_t launch_assign(const float* rows, int64_t nrows, int DP, const void* cent,
                         const float* cnorm, int C, int32_t* keys_out, int32_t* order,
                         unsigned long long* heads, int G, int S, hipStream_t s) {
    if (nrows <= 0) return hipSuccess;
    if (!keys_out && (G < 1 || G > 32 || S < 1 || G * S > 256 || C > 256)) return hipErrorInvalidValue;
    if (DP % 32 || DP < 32 || DP > 256) return hipErrorInvalidValue;
    const GroupLayout y = keys_out ? GroupLayout{true, 0, 0} : list_layout(nrows, C, G, S);
    const bool wide = !y.by_centroid;
    const int per = wide ? 256 : 16;
    const unsigned blocks = (unsigned)((nrows + per - 1) / per);
    const __bf16* cb = static_cast<const __bf16*>(cent);
#define VMN_ASSIGN(KB)                                                                              \
    do {                                                                                            \
        if (wide)                                                                                   \
            hipLaunchKernelGGL((assign_kernel_wide<KB>), dim3(blocks), dim3(1024), 0, s, rows,      \
                               nrows, cb, cnorm, C, keys_out, order, heads, G, S, y.stride, false); \
        else                                                                                        \
            hipLaunchKernelGGL((assign_kernel_one<KB>), dim3(blocks), dim3(64), 0, s, rows, nrows,  \
                               cb, cnorm, C, keys_out, order, heads, G, S, y.stride, true);         \
    } while (0)
    switch (DP / 32) {
        case 1: VMN_ASSIGN(1); break;
        case 2: VMN_ASSIGN(2); break;
        case 3: VMN_ASSIGN(3); break;
        case 4: VMN_ASSIGN(4); break;
        case 5: VMN_ASSIGN(5); break;
        case 6: VMN_ASSIGN(6); break;
        case 7: VMN_ASSIGN(7); break;
        default: VMN_ASSIGN(8); break;
    }
#undef VMN_ASSIGN
    return hipGetLastError();
}

}  // namespace vmn

namespace vmn_rt {

// Sample size and centroid count of the grouping's k-means, and its rounds.
// 64 centroids: on the 1M x 128 mixture 32 leave FP32 batches 3% slower and
// 128 / 256 leave BF16 batches 5% / 12% slower (more, smaller lists to walk).
constexpr int kSample = 16384;
constexpr int kCentroids = 64;
constexpr int kRounds = 6;
// Batches of at least this many queries per CU are grouped.
constexpr int kMinQueriesPerCu = 8;

bool groups_wanted(const vmn_index* idx, int64_t nq) {
    return idx->query_groups > 1 && idx->n >= 4 * kCentroids && idx->DP <= 256 &&
           nq >= (int64_t)idx->cus * kMinQueriesPerCu;
}

int ensure_centroids(vmn_index* idx, bool* computed) {
    // computed once, and again once the index has doubled since
    *computed = false;
    if (idx->ncent > 0 && idx->n < 2 * idx->cent_n) return VMN_OK;
    *computed = true;
    const int DP = idx->DP;
    const int S = (int)std::min<int64_t>(idx->n, kSample);
    const int C = std::min(kCentroids, S / 4);
    HIPCHK(idx->grp_sample.reserve((size_t)S * DP * sizeof(float)));
    HIPCHK(idx->grp_keys.reserve((size_t)S * sizeof(int32_t)));
    HIPCHK(idx->cent.reserve((size_t)C * DP * sizeof(__bf16)));
    HIPCHK(idx->cent_norm.reserve((size_t)C * sizeof(float)));
    float* smp = idx->grp_sample.as<float>();
    const int64_t se = (int64_t)S * DP, ce = (int64_t)C * DP;
    hipLaunchKernelGGL(vmn::sample_rows_kernel, dim3((unsigned)((se + 255) / 256)), dim3(256), 0,
                       idx->stream, idx->vecs.as<float>(), idx->n, DP, S, smp);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(vmn::cent_init_kernel, dim3((unsigned)((ce + 255) / 256)), dim3(256), 0,
                       idx->stream, smp, S, DP, C, idx->cent.as<__bf16>());
    HIPCHK(hipGetLastError());
    // the norms of the initial centroids: an update with no members keeps them
    HIPCHK(vmn::launch_fill_i32(idx->grp_keys.as<int32_t>(), S, -1, idx->stream));
    hipLaunchKernelGGL(vmn::kmeans_update_kernel, dim3(C), dim3(256), 0, idx->stream, smp, S, DP,
                       idx->grp_keys.as<int32_t>(), idx->cent.as<__bf16>(), idx->cent_norm.as<float>());
    HIPCHK(hipGetLastError());
    for (int r = 0; r < kRounds; ++r) {
        HIPCHK(vmn::launch_assign(smp, S, DP, idx->cent.p, idx->cent_norm.as<float>(), C,
                                  idx->grp_keys.as<int32_t>(), nullptr, nullptr, 0, 1, idx->stream));
        hipLaunchKernelGGL(vmn::kmeans_update_kernel, dim3(C), dim3(256), 0, idx->stream, smp, S, DP,
                           idx->grp_keys.as<int32_t>(), idx->cent.as<__bf16>(),
                           idx->cent_norm.as<float>());
        HIPCHK(hipGetLastError());
    }
    idx->ncent = C;
    idx->cent_n = idx->n;
    return VMN_OK;
}

vmn::GroupLayout group_layout(const vmn_index* idx, int64_t nq) {
    return vmn::list_layout(nq, idx->ncent, idx->query_groups, VMN_GROUP_SUBLISTS);
}

int group_queries(vmn_index* idx, const float* q, int64_t nq, DevBuf& order,
                  unsigned long long* heads, hipStream_t stream) {
    const vmn::GroupLayout y = group_layout(idx, nq);
    HIPCHK(order.reserve((size_t)y.nlists * y.stride * sizeof(int32_t)));
    HIPCHK(vmn::launch_assign(q, nq, idx->DP, idx->cent.p, idx->cent_norm.as<float>(), idx->ncent,
                              nullptr, order.as<int32_t>(), heads, idx->query_groups,
                              VMN_GROUP_SUBLISTS, stream));
    return VMN_OK;
}

}  // namespace vmn_rt

extern "C" int vmn_set_query_groups(vmn_index* idx, int32_t groups) {
    if (!idx) return fail(VMN_ERR_ARG, "NULL index");
    if (groups < 0 || groups > 32) return fail(VMN_ERR_ARG, "query groups must be in [0, 32] (%d)", groups);
    idx->query_groups = groups;
    return VMN_OK;
}

extern "C" int vmn_query_group_ids(vmn_index* idx, const float* queries, int64_t nq, int32_t mem,
                                   uint8_t* gids_out) {
    if (!idx || !queries || !gids_out || nq < 0) return fail(VMN_ERR_ARG, "NULL argument");
    if (!idx->built) return fail(VMN_ERR_STATE, "index has no graph");
    if (idx->query_groups < 2) return fail(VMN_ERR_STATE, "query groups are off");
    if (idx->n < 256 || idx->DP > 256) return fail(VMN_ERR_STATE, "index too small or too wide for query groups");
    if (nq == 0) return VMN_OK;
    int rc = set_device(idx);
    if (rc) return rc;
    bool fresh = false;
    rc = ensure_centroids(idx, &fresh);
    if (rc) return rc;
    const float* q = nullptr;
    rc = stage_rows(idx, queries, nq, mem, idx->qpad, &q);
    if (rc) return rc;
    const int G = idx->query_groups;
    const vmn::GroupLayout y = group_layout(idx, nq);
    HIPCHK(idx->counter.reserve((size_t)y.nlists * 256));
    unsigned long long* heads = idx->counter.as<unsigned long long>();
    HIPCHK(vmn::launch_fill_u64(heads, 32 * y.nlists, 0ull, idx->stream));
    rc = group_queries(idx, q, nq, idx->grp_order, heads, idx->stream);
    if (rc) return rc;
    std::vector<unsigned long long> h((size_t)32 * y.nlists);
    std::vector<int32_t> ord((size_t)y.nlists * y.stride);
    HIPCHK(hipMemcpyAsync(h.data(), heads, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          idx->stream));
    HIPCHK(hipMemcpyAsync(ord.data(), idx->grp_order.p, ord.size() * sizeof(int32_t),
                          hipMemcpyDeviceToHost, idx->stream));
    HIPCHK(hipStreamSynchronize(idx->stream));
    std::vector<uint8_t> gid((size_t)nq, 0xff);
    for (int L = 0; L < y.nlists; ++L)
        for (unsigned long long i = 0; i < h[32 * L + 1] && i < (unsigned long long)y.stride; ++i) {
            const int g = y.by_centroid ? (int)((int64_t)L * G / idx->ncent) : L / VMN_GROUP_SUBLISTS;
            const int32_t r = ord[(size_t)L * y.stride + i];
            if (r < 0 || r >= nq || gid[r] != 0xff)
                return fail(VMN_ERR_STATE, "query group lists are not a partition of the batch");
            gid[r] = (uint8_t)g;
        }
    for (int64_t r = 0; r < nq; ++r)
        if (gid[r] == 0xff) return fail(VMN_ERR_STATE, "query %lld in no group", (long long)r);
    HIPCHK(hipMemcpy(gids_out, gid.data(), (size_t)nq,
                     mem == VMN_MEM_DEVICE ? hipMemcpyHostToDevice : hipMemcpyHostToHost));
    return VMN_OK;
}
