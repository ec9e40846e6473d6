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

// Lanes per row = row_bytes / 16 (each lane loads 16 bytes); a wave covers
// 64 / lanes-per-row rows per load step and keeps UNR steps in flight.
template <int ROW_BYTES>
__global__ __launch_bounds__(256) void gather_probe_kernel(const uint4* __restrict__ table,
                                                           const int32_t* __restrict__ ids,
                                                           int64_t n_reads, uint32_t* sink) {
    constexpr int LPR = ROW_BYTES / 16;
    constexpr int RPW = 64 / LPR;  // rows per wave per step
    constexpr int UNR = 4;
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int r_in = lane / LPR, part = lane % LPR;
    uint32_t acc = 0;
    for (int64_t base = wave * RPW * UNR; base < n_reads; base += waves * RPW * UNR) {
        uint4 v[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int64_t r = base + (int64_t)u * RPW + r_in;
            v[u] = r < n_reads ? table[(int64_t)ids[r] * LPR + part] : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    sink[wave * 64 + lane] = acc;
}

}  // namespace

extern "C" int vmn_gather_probe(const void* table, int64_t table_rows, int32_t row_bytes,
                                const int32_t* ids, int64_t n_reads, uint32_t* sink,
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
    if (row_bytes == 128)
        hipLaunchKernelGGL(gather_probe_kernel<128>, dim3((unsigned)blocks), dim3(256), 0, nullptr, t,
                           ids, n_reads, sink);
    else if (row_bytes == 256)
        hipLaunchKernelGGL(gather_probe_kernel<256>, dim3((unsigned)blocks), dim3(256), 0, nullptr, t,
                           ids, n_reads, sink);
    else
        hipLaunchKernelGGL(gather_probe_kernel<512>, dim3((unsigned)blocks), dim3(256), 0, nullptr, t,
                           ids, n_reads, sink);
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
