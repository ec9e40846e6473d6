// Reverse pairs grouped by destination with a stable device radix sort.
//
// An insert batch of c points emits c x R (destination, source) pairs. The
// build groups them with a count / scan / scatter over every node (a CSR),
// which is the right shape when the pairs outnumber the nodes. A small batch
// into a large index would pay for passes over all n nodes instead; here the
// pairs are sorted by destination (unused pairs, dst = -1, sort last as
// 0xffffffff), segment heads are flagged and scanned, and every distinct
// destination gets a list entry with its segment's offsets: all work is in
// the number of pairs. The reverse prune orders each node's candidates by
// (distance, id) itself, so the order inside a segment does not matter.
#include <hipcub/hipcub.hpp>

#include "vmn_kernels.hpp"

namespace vmn {

namespace {

constexpr uint32_t kUnused = 0xffffffffu;

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Layout {
    size_t keys_in, keys_out, vals_in, vals_out, flags, scan, lib, total, lib_bytes;
};

Layout layout(int64_t npairs) {
    size_t sort_b = 0, scan_b = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_b, (const uint32_t*)nullptr,
                                             (uint32_t*)nullptr, (const int32_t*)nullptr,
                                             (int32_t*)nullptr, (int)npairs, 0, 32);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan_b, (const int32_t*)nullptr,
                                           (int32_t*)nullptr, (int)npairs);
    Layout L{};
    const size_t k = up256((size_t)npairs * 4);
    L.keys_in = 0;
    L.keys_out = L.keys_in + k;
    L.vals_in = L.keys_out + k;
    L.vals_out = L.vals_in + k;
    L.flags = L.vals_out + k;
    L.scan = L.flags + k;
    L.lib = L.scan + k;
    L.lib_bytes = up256(std::max(sort_b, scan_b));
    L.total = L.lib + L.lib_bytes;
    return L;
}

__global__ void group_keys_kernel(const int32_t* pair_dst, int64_t npairs, uint32_t* keys,
                                  int32_t* vals, unsigned long long* nseg) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *nseg = 0ull;
    if (i >= npairs) return;
    keys[i] = (uint32_t)pair_dst[i];
    vals[i] = (int32_t)i;
}

__global__ void group_heads_kernel(const uint32_t* keys, int64_t npairs, int32_t* flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npairs) return;
    const uint32_t k = keys[i];
    flags[i] = (k != kUnused && (i == 0 || keys[i - 1] != k)) ? 1 : 0;
}

__global__ void group_write_kernel(const uint32_t* keys, const int32_t* vals, const int32_t* flags,
                                   const int32_t* scan, int64_t npairs, const float* pair_d, int R,
                                   int64_t node_base, int32_t* rev_src, float* rev_d,
                                   int64_t* node_list, int64_t* list_off,
                                   unsigned long long* nseg) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npairs) return;
    const uint32_t k = keys[i];
    if (k == kUnused) return;
    const int32_t v = vals[i];
    rev_src[i] = (int32_t)(node_base + v / R);
    rev_d[i] = pair_d[v];
    if (flags[i]) {
        node_list[scan[i]] = (int64_t)k;
        list_off[scan[i]] = i;
    }
    if (i + 1 == npairs || keys[i + 1] == kUnused) {  // last used pair
        const int64_t total = (int64_t)scan[i] + flags[i];
        list_off[total] = i + 1;
        *nseg = (unsigned long long)total;
    }
}

unsigned grid_for(int64_t n) {
    const int64_t g = (n + 255) / 256;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace

size_t reverse_group_temp_bytes(int64_t npairs) { return layout(npairs).total; }

hipError_t launch_reverse_group(const int32_t* pair_dst, const float* pair_d, int64_t npairs, int R,
                                int64_t node_base, void* temp, size_t temp_bytes,
                                int32_t* rev_src, float* rev_d, int64_t* node_list,
                                int64_t* list_off, unsigned long long* nseg, hipStream_t s) {
    if (npairs < 1 || npairs > (int64_t)0x7fffffff) return hipErrorInvalidValue;
    const Layout L = layout(npairs);
    if (temp_bytes < L.total) return hipErrorInvalidValue;
    char* t = static_cast<char*>(temp);
    uint32_t* keys_in = reinterpret_cast<uint32_t*>(t + L.keys_in);
    uint32_t* keys_out = reinterpret_cast<uint32_t*>(t + L.keys_out);
    int32_t* vals_in = reinterpret_cast<int32_t*>(t + L.vals_in);
    int32_t* vals_out = reinterpret_cast<int32_t*>(t + L.vals_out);
    int32_t* flags = reinterpret_cast<int32_t*>(t + L.flags);
    int32_t* scan = reinterpret_cast<int32_t*>(t + L.scan);
    const unsigned g = grid_for(npairs);
    hipLaunchKernelGGL(group_keys_kernel, dim3(g), dim3(256), 0, s, pair_dst, npairs, keys_in,
                       vals_in, nseg);
    size_t lb = L.lib_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(t + L.lib, lb, keys_in, keys_out, vals_in,
                                                      vals_out, (int)npairs, 0, 32, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(group_heads_kernel, dim3(g), dim3(256), 0, s, keys_out, npairs, flags);
    lb = L.lib_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(t + L.lib, lb, flags, scan, (int)npairs, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(group_write_kernel, dim3(g), dim3(256), 0, s, keys_out, vals_out, flags, scan,
                       npairs, pair_d, R, node_base, rev_src, rev_d, node_list, list_off, nseg);
    return hipGetLastError();
}

}  // namespace vmn
