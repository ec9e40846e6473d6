// Update-path kernel arguments and launchers (internal to the library).
#pragma once
#include "vmn_kernels.hpp"

namespace vmn {

struct ConsArgs {
    int DP;
    const float* vecs;
    const void* bf16_rows;    // BF16 copies of vecs when every row is exact in BF16, else null
    const float* bf16_norms;
    int32_t* adj;
    const uint32_t* del_bits;
    int W, R, M;
    float alpha;
    const int64_t* list;   // nodes to re-prune
    int64_t nlist;
    unsigned long long* counter;
    // two tiers, as the FRESH prunes: with `overflow` set the launch uses a
    // 256-entry pool and lists the list positions whose pool could exceed it
    // (tombstoned neighbours x W > 256); `items` then names the positions a
    // full-pool launch takes (nlist = their number)
    int64_t* overflow;
    unsigned long long* overflow_n;
    const int64_t* items;
};

hipError_t launch_tombstone(const int32_t* ids, int64_t m, int64_t n, uint32_t* del_bits,
                            unsigned long long* bad, hipStream_t s);
hipError_t launch_find_vectors(const float* vecs, int64_t n, int D, int DP, const uint32_t* del_bits,
                               const float* qv, int64_t nq, uint32_t* out, hipStream_t s);
hipError_t launch_cons_flags(const int32_t* adj, int64_t n, int W, const uint32_t* del_bits,
                             int32_t* flags, uint32_t* summary, int cus, hipStream_t s);
hipError_t launch_consolidate(const ConsArgs& a, int grid, hipStream_t s);
hipError_t launch_clear_deleted_rows(int32_t* adj, int64_t n, int W, const uint32_t* del_bits,
                                     hipStream_t s);

}  // namespace vmn
