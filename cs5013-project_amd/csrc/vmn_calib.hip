// Random-row gather probe: reads `n_reads` rows of `row_bytes` bytes at the
// row ids given (each row once), the access pattern of the search kernels'
// candidate gathers, and writes one word per lane. Its byte count is known
// exactly, so the HBM traffic counters (FETCH_SIZE) read during it give the
// scale that turns those counters into bytes for this kind of access. A
// measurement helper, not part of the index.
#include <hip/hip_runtime.h>

#include "vmn_internal.hpp"

using namespace vmn_rt;

namespace {

// Lanes per row per load = LPR (each lane loads 16 bytes): LPR = row_bytes / 16
// reads every row in one load instruction; LPR = 4 (SPLIT) reads 64 bytes of
// 16 rows per instruction and the rest of those rows in the next ones (the
// BF16 search gather's layout). A wave keeps UNR load steps in flight.
template <int ROW_BYTES, int LPR>
__global__ __launch_bounds__(256) void gather_probe_kernel(const uint4* __restrict__ table,
                                                           const int32_t* __restrict__ ids,
                                                           int64_t n_reads, uint32_t* sink) {
    constexpr int CPR = ROW_BYTES / 16;  // 16-byte chunks per row
    constexpr int RPW = 64 / LPR;        // rows per wave per step
    constexpr int NS = CPR / LPR;        // load instructions per row
    constexpr int UNR = NS >= 4 ? 1 : 4 / NS;
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int r_in = lane / LPR, part = lane % LPR;
    uint32_t acc = 0;
    for (int64_t base = wave * RPW * UNR; base < n_reads; base += waves * RPW * UNR) {
        uint4 v[UNR][NS];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int64_t r = base + (int64_t)u * RPW + r_in;
            const int64_t row = r < n_reads ? (int64_t)ids[r] : -1;
#pragma unroll
            for (int s = 0; s < NS; ++s)
                v[u][s] = row >= 0 ? table[row * CPR + s * LPR + part] : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u)
#pragma unroll
            for (int s = 0; s < NS; ++s) acc ^= v[u][s].x ^ v[u][s].y ^ v[u][s].z ^ v[u][s].w;
    }
    sink[wave * 64 + lane] = acc;
}

}  // namespace

extern "C" int vmn_gather_probe(const void* table, int64_t table_rows, int32_t row_bytes,
                                int32_t split, const int32_t* ids, int64_t n_reads, uint32_t* sink,
                                int64_t sink_words, int32_t device, double* ms_out) {
    if (!table || !ids || !sink || n_reads < 1 || table_rows < 1)
        return fail(VMN_ERR_ARG, "gather probe: NULL buffer or empty read list");
    if (row_bytes != 128 && row_bytes != 256 && row_bytes != 512)
        return fail(VMN_ERR_ARG, "gather probe: row_bytes must be 128, 256 or 512 (%d)", row_bytes);
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    // 8 workgroups of 4 waves per CU; the sink holds one word per lane
    int64_t blocks = (int64_t)prop.multiProcessorCount * 8;
    const int64_t need = blocks * 256;
    if (sink_words < need) blocks = sink_words / 256;
    if (blocks < 1) return fail(VMN_ERR_ARG, "gather probe: sink needs >= 256 words");
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, nullptr));
    const uint4* t = static_cast<const uint4*>(table);
#define VMN_PROBE(RB, LPR)                                                                    \
    hipLaunchKernelGGL((gather_probe_kernel<RB, LPR>), dim3((unsigned)blocks), dim3(256), 0, \
                       nullptr, t, ids, n_reads, sink)
    if (row_bytes == 128) {
        if (split) VMN_PROBE(128, 4); else VMN_PROBE(128, 8);
    } else if (row_bytes == 256) {
        if (split) VMN_PROBE(256, 4); else VMN_PROBE(256, 16);
    } else {
        if (split) VMN_PROBE(512, 4); else VMN_PROBE(512, 32);
    }
#undef VMN_PROBE
    hipError_t le = hipGetLastError();
    HIPCHK(hipEventRecord(e1, nullptr));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(le);
    if (ms_out) *ms_out = ms;
    return VMN_OK;
}
