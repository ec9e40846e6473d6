// vmn_search_noalloc: the batched search of vmn_search_ex on device buffers,
// for callers that may not allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc). Its scratch -- the
// persistent launch's work-queue counter and, when the dimension is not one of
// the kernel widths, the zero-padded query rows -- lives in one file-scope
// device array per GPU instead of buffers the library grows on demand.
//
// The array's host-side name is not a device address: the address of each
// device's copy is taken once, with that device current, by
// hipGetSymbolAddress, and only that address is handed to kernels.
#include <hip/hip_runtime.h>

#include <mutex>

#include "vmn_internal.hpp"
#include "vmn_kernels.hpp"

using namespace vmn_rt;

namespace {

// Work counter (first 256 bytes) + padded query rows.
constexpr size_t kScratchBytes = (size_t)4 << 20;
constexpr size_t kCounterBytes = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (the device must be current).
int scratch_ptr(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(VMN_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(VMN_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                        hipGetErrorString(e));
        g_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_ptr[dev]);
    return VMN_OK;
}

}  // namespace

extern "C" int vmn_search_noalloc(vmn_index* idx, const float* queries, int64_t nq, int32_t k,
                                  int32_t L, int32_t prec, int32_t k_rerank, int32_t* ids_out,
                                  float* dists_out) {
    if (!idx || !queries || !ids_out) return fail(VMN_ERR_ARG, "NULL argument");
    if (!idx->built) return fail(VMN_ERR_STATE, "index has no graph");
    if (idx->n_pend)
        return fail(VMN_ERR_STATE,
                    "staged inserts pending: their merge needs scratch; vmn_consolidate first");
    if (nq < 0 || k < 1 || k > L || L < 1 || L > 512)
        return fail(VMN_ERR_ARG, "need 1 <= k <= L <= 512 (k=%d, L=%d)", k, L);
    if (prec < VMN_PREC_FP32 || prec > VMN_PREC_FP8) return fail(VMN_ERR_ARG, "bad precision %d", prec);
    if (prec == VMN_PREC_BF16 && !(idx->p.low_precision & VMN_STORE_BF16))
        return fail(VMN_ERR_STATE, "index keeps no BF16 rows (low_precision)");
    if (prec == VMN_PREC_FP8 && !(idx->p.low_precision & VMN_STORE_FP8))
        return fail(VMN_ERR_STATE, "index keeps no FP8 rows (low_precision)");
    if (k_rerank != 0 && (k_rerank < k || k_rerank > 64 || k_rerank > L))
        return fail(VMN_ERR_ARG, "need k <= k_rerank <= min(64, L) (k_rerank=%d)", k_rerank);
    const int D = idx->p.dim, DP = idx->DP;
    const bool pad = D != DP;
    const size_t qbytes = pad ? (size_t)nq * DP * sizeof(float) : 0;
    if (qbytes > kScratchBytes - kCounterBytes)
        return fail(VMN_ERR_CAPACITY,
                    "vmn_search_noalloc: %lld queries of dimension %d need %zu bytes of padded "
                    "rows, the scratch holds %zu (at most %lld queries; dimensions that are a "
                    "multiple of 32 need no padding and have no limit)",
                    (long long)nq, D, qbytes, kScratchBytes - kCounterBytes,
                    (long long)((kScratchBytes - kCounterBytes) / ((size_t)DP * sizeof(float))));
    if (nq == 0) return VMN_OK;
    int rc = set_device(idx);
    if (rc) return rc;
    unsigned char* base = nullptr;
    rc = scratch_ptr(idx->p.device, &base);
    if (rc) return rc;
    // FP8 reads row norms laid out beside the adjacency (allocated with the
    // graph): after an index update they are refreshed on the index stream,
    // ahead of this search
    rc = prepare_search(idx, prec, idx->p.mode, false);
    if (rc) return rc;
    unsigned long long* counter = reinterpret_cast<unsigned long long*>(base);
    const float* q = queries;
    if (pad) {
        float* qp = reinterpret_cast<float*>(base + kCounterBytes);
        HIPCHK(vmn::launch_pad_rows(queries, nq, D, qp, DP, idx->stream));
        q = qp;
    }
    // the counter keeps the last call's value: cleared by a kernel (no
    // memset node when a caller captures this call)
    HIPCHK(vmn::launch_fill_u64(counter, 1, 0ull, idx->stream));
    vmn::SearchArgs a;
    search_args(idx, a, q, nq, L, k, ids_out, dists_out, false, idx->p.mode, 0, prec,
                prec == VMN_PREC_FP32 ? 0 : k_rerank);
    a.counter = counter;
    HIPCHK(vmn::launch_search(a, search_grid(idx, nq), idx->stream));
    return VMN_OK;
}
