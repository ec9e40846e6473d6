// The parth engine's operation stream behind the C ABI: workload files, the
// free-list delete and the link-to-nearest insert, and the batched replay of
// VamanaEngine::process_workload (parth/src/engine.cu:304-568).
//
// The reference runs each of its update kernels with one thread per operation
// and lets atomics decide the order. Here the id assignment and the back links
// run in operation order (what one converged wavefront does), the vector
// writes and the nearest-node searches run in parallel, and every search of a
// batch sees the graph from before the batch's links. Searches and the
// 1-nearest searches use the PARTH search kernel (FP32 distances summed in the
// reference's order, or BF16 rows where those give the same values bit for
// bit); the ground truth is the reference's FP32 brute force (first 10 live
// ids below the node count by FP32 distance, then id).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "vmn_internal.hpp"
#include "vmn_kernels.hpp"

using namespace vmn_rt;

namespace {

// Sets (set != 0) or clears the tombstone bit of each id (ids checked by the
// host against the capacity).
__global__ void engine_mark_kernel(const int32_t* __restrict__ ids, int64_t n, uint32_t* del_bits,
                                   int set) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = (uint32_t)ids[i];
    if (set) atomicOr(&del_bits[id >> 5], 1u << (id & 31));
    else atomicAnd(&del_bits[id >> 5], ~(1u << (id & 31)));
}

// insertPrepareKernel's writes (engine.cu:161-199), one workgroup per insert:
// the batch row src[i] (the last insert given id ids[i]) is copied to the
// node's row, to the nearest-node query row i, and the node's adjacency row
// is reset to -1.
__global__ void engine_place_kernel(const float* __restrict__ staged, const int32_t* __restrict__ ids,
                                    const int32_t* __restrict__ src, int64_t n, int DP, float* vecs,
                                    int32_t* adj, int W, float* queries) {
    const int64_t i = blockIdx.x;
    if (i >= n) return;
    const int64_t id = ids[i];
    const int64_t s = src[i];
    for (int k = threadIdx.x; k < DP; k += blockDim.x) {
        const float v = staged[s * DP + k];
        queries[i * DP + k] = v;
        if (s == i) vecs[id * DP + k] = v;
    }
    for (int k = threadIdx.x; k < W; k += blockDim.x) adj[id * W + k] = -1;
}

// insertLinkKernel's writes (engine.cu:227-238): slot 0 of every new row is
// its nearest node; then, in batch order, the new id goes into the first -1
// slot of the nearest node's row (a full row takes no back link). One lane:
// the result depends on this order.
__global__ void engine_link_kernel(int32_t* adj, int W, const int32_t* __restrict__ ids,
                                   const int32_t* __restrict__ best, int64_t n) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int64_t i = 0; i < n; ++i)
        if (best[i] != -1) adj[(int64_t)ids[i] * W] = best[i];
    for (int64_t i = 0; i < n; ++i) {
        const int32_t b = best[i];
        if (b == -1) continue;
        int32_t* row = adj + (int64_t)b * W;
        for (int k = 0; k < W; ++k)
            if (row[k] == -1) {
                row[k] = ids[i];
                break;
            }
    }
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

int check_engine(const vmn_index* idx) {
    if (!idx) return fail(VMN_ERR_ARG, "NULL index");
    if (!idx->built) return fail(VMN_ERR_STATE, "index has no graph");
    if (idx->p.mode != VMN_MODE_PARTH)
        return fail(VMN_ERR_STATE, "parth engine operations need a PARTH-mode index");
    if (idx->n_pend)
        return fail(VMN_ERR_STATE, "staged inserts pending: vmn_consolidate adds them first");
    return VMN_OK;
}

int mark(vmn_index* idx, const std::vector<int32_t>& ids, int set) {
    if (ids.empty()) return VMN_OK;
    ++idx->epoch;
    HIPCHK(idx->eng_marks.reserve(ids.size() * sizeof(int32_t)));
    HIPCHK(hipMemcpyAsync(idx->eng_marks.p, ids.data(), ids.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, idx->stream));
    hipLaunchKernelGGL(engine_mark_kernel, dim3(blocks_for((int64_t)ids.size())), dim3(256), 0,
                       idx->stream, idx->eng_marks.as<int32_t>(), (int64_t)ids.size(),
                       idx->del_bits.as<uint32_t>(), set);
    HIPCHK(hipGetLastError());
    // the host vector must outlive the copy
    HIPCHK(hipStreamSynchronize(idx->stream));
    return VMN_OK;
}

}  // namespace

int vmn_read_workload(const char* path, int32_t dim, int32_t* types, int32_t* ids, float* vecs,
                      int64_t capacity, int64_t* count) {
    if (!path || !count) return fail(VMN_ERR_ARG, "NULL argument");
    if (dim < 1) return fail(VMN_ERR_ARG, "dim %d", dim);
    if (types && (!ids || !vecs)) return fail(VMN_ERR_ARG, "types given without ids / vecs");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(VMN_ERR_IO, "cannot open %s", path);
    if (std::fseek(f, 0, SEEK_END) != 0) {
        std::fclose(f);
        return fail(VMN_ERR_IO, "cannot seek %s", path);
    }
    const long bytes = std::ftell(f);
    std::rewind(f);
    const int64_t rec = 8 + 4 * (int64_t)dim;
    const int64_t total = bytes > 0 ? bytes / rec : 0;
    *count = total;
    if (!types) {
        std::fclose(f);
        return VMN_OK;
    }
    const int64_t m = total < capacity ? total : capacity;
    std::vector<unsigned char> buf((size_t)rec);
    for (int64_t i = 0; i < m; ++i) {
        if (std::fread(buf.data(), 1, (size_t)rec, f) != (size_t)rec) {
            std::fclose(f);
            return fail(VMN_ERR_IO, "short read in %s at record %lld", path, (long long)i);
        }
        std::memcpy(&types[i], buf.data(), 4);
        std::memcpy(&ids[i], buf.data() + 4, 4);
        std::memcpy(vecs + i * dim, buf.data() + 8, 4 * (size_t)dim);
    }
    std::fclose(f);
    *count = m;
    return VMN_OK;
}

int vmn_engine_delete(vmn_index* idx, const int32_t* ids, int64_t n) {
    int rc = check_engine(idx);
    if (rc) return rc;
    if (n < 0 || (!ids && n)) return fail(VMN_ERR_ARG, "bad id list");
    if (n == 0) return VMN_OK;
    rc = set_device(idx);
    if (rc) return rc;
    if ((int64_t)idx->freelist.size() < idx->cap) idx->freelist.resize((size_t)idx->cap, -1);
    std::vector<int32_t> valid;
    valid.reserve((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t t = ids[i];
        if (t < 0 || t >= idx->cap) continue;
        valid.push_back(t);
        const int64_t slot = idx->free_count++;
        if (slot >= 0 && slot < idx->cap) idx->freelist[(size_t)slot] = t;
    }
    idx->n_deleted += (int64_t)valid.size();
    return mark(idx, valid, 1);
}

int vmn_engine_insert(vmn_index* idx, const float* vecs, int64_t n, int32_t link_L,
                      int32_t* ids_out, int32_t mem) {
    int rc = check_engine(idx);
    if (rc) return rc;
    if (n < 0 || (!vecs && n)) return fail(VMN_ERR_ARG, "bad vector list");
    if (link_L < 1 || link_L > 512) return fail(VMN_ERR_ARG, "need 1 <= link_L <= 512 (%d)", link_L);
    if (n == 0) return VMN_OK;
    rc = set_device(idx);
    if (rc) return rc;
    ++idx->epoch;
    // ids: every insert sees the count of the batch start
    const int64_t fc = idx->free_count;
    const int64_t popped = fc > 0 ? std::min<int64_t>(n, fc) : 0;
    if (idx->n + (n - popped) > idx->cap)
        return fail(VMN_ERR_CAPACITY, "insert of %lld exceeds capacity %lld (size %lld)",
                    (long long)n, (long long)idx->cap, (long long)idx->n);
    std::vector<int32_t> ids((size_t)n), src((size_t)n), reused;
    int64_t c0 = idx->n;
    int32_t lo = INT32_MAX, hi = -1;
    for (int64_t i = 0; i < n; ++i) {
        int32_t id = -1;
        if (fc > 0) {
            const int64_t old = idx->free_count--;
            if (old > 0) {
                id = idx->freelist[(size_t)(old - 1)];
                reused.push_back(id);
            }
        }
        if (id == -1) id = (int32_t)c0++;
        ids[(size_t)i] = id;
        lo = std::min(lo, id);
        hi = std::max(hi, id);
    }
    {
        std::unordered_map<int32_t, int32_t> last;
        for (int64_t i = 0; i < n; ++i) last[ids[(size_t)i]] = (int32_t)i;
        for (int64_t i = 0; i < n; ++i) src[(size_t)i] = last[ids[(size_t)i]];
    }
    const int DP = idx->DP, W = idx->p.R;
    const float* staged = nullptr;
    rc = stage_rows(idx, vecs, n, mem, idx->staging, &staged);
    if (rc) return rc;
    HIPCHK(idx->eng_ids.reserve((size_t)n * sizeof(int32_t)));
    HIPCHK(idx->eng_src.reserve((size_t)n * sizeof(int32_t)));
    HIPCHK(idx->eng_best.reserve((size_t)n * sizeof(int32_t)));
    HIPCHK(idx->eng_q.reserve((size_t)n * DP * sizeof(float)));
    HIPCHK(hipMemcpyAsync(idx->eng_ids.p, ids.data(), (size_t)n * sizeof(int32_t),
                          hipMemcpyHostToDevice, idx->stream));
    HIPCHK(hipMemcpyAsync(idx->eng_src.p, src.data(), (size_t)n * sizeof(int32_t),
                          hipMemcpyHostToDevice, idx->stream));
    hipLaunchKernelGGL(engine_place_kernel, dim3((unsigned)n), dim3(64), 0, idx->stream, staged,
                       idx->eng_ids.as<int32_t>(), idx->eng_src.as<int32_t>(), n, DP,
                       idx->vecs.as<float>(), idx->adj.as<int32_t>(), W, idx->eng_q.as<float>());
    HIPCHK(hipGetLastError());
    rc = mark(idx, reused, 0);
    if (rc) return rc;
    idx->n = c0;
    // the traversal stays on exact BF16 rows only if the new rows qualify too
    // (rows [lo, hi] hold them; the others there were checked before)
    rc = check_exact(idx, lo, (int64_t)hi + 1);
    if (rc) return rc;
    rc = refresh_lowp(idx, lo, (int64_t)hi + 1);
    if (rc) return rc;
    rc = run_search(idx, idx->eng_q.as<float>(), n, link_L, 1, idx->eng_best.as<int32_t>(),
                    nullptr, false, VMN_MODE_PARTH, false, 0, build_prec(idx));
    if (rc) return rc;
    hipLaunchKernelGGL(engine_link_kernel, dim3(1), dim3(64), 0, idx->stream,
                       idx->adj.as<int32_t>(), W, idx->eng_ids.as<int32_t>(),
                       idx->eng_best.as<int32_t>(), n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(idx->stream));
    if (ids_out) std::memcpy(ids_out, ids.data(), (size_t)n * sizeof(int32_t));
    return VMN_OK;
}

int vmn_engine_run(vmn_index* idx, const int32_t* types, const int32_t* ids, const float* vecs,
                   int64_t n_ops, int32_t batch_size, int32_t L, int32_t link_L,
                   vmn_engine_result* result, int32_t* found_out, int32_t* gt_out) {
    int rc = check_engine(idx);
    if (rc) return rc;
    if (!result || n_ops < 0 || (n_ops && (!types || !ids || !vecs)))
        return fail(VMN_ERR_ARG, "NULL argument");
    if (batch_size < 1) return fail(VMN_ERR_ARG, "batch_size %d", batch_size);
    if (L < 1 || L > 512) return fail(VMN_ERR_ARG, "need 1 <= L <= 512 (%d)", L);
    rc = set_device(idx);
    if (rc) return rc;
    std::memset(result, 0, sizeof *result);
    const auto t0 = std::chrono::steady_clock::now();
    const int D = idx->p.dim, DP = idx->DP;
    constexpr int K = 10;  // searchKernel / bruteForceKernel keep 10 (engine.cu:112, :267)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    struct Events {
        hipEvent_t a, b;
        ~Events() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } guard{ev0, ev1};
    double recall_sum = 0.0;
    std::vector<int32_t> del_ids, h_found, h_gt, h_hops, h_nd;
    std::vector<float> ins, qs;
    for (int64_t b = 0; b < n_ops; b += batch_size) {
        const int64_t e = std::min<int64_t>(b + batch_size, n_ops);
        del_ids.clear();
        ins.clear();
        qs.clear();
        for (int64_t i = b; i < e; ++i) {
            const float* v = vecs + i * D;
            if (types[i] == VMN_OP_DELETE) del_ids.push_back(ids[i]);
            else if (types[i] == VMN_OP_INSERT) ins.insert(ins.end(), v, v + D);
            else qs.insert(qs.end(), v, v + D);
        }
        if (!del_ids.empty()) {
            rc = vmn_engine_delete(idx, del_ids.data(), (int64_t)del_ids.size());
            if (rc) return rc;
            result->deletes += (int64_t)del_ids.size();
        }
        if (!ins.empty()) {
            const int64_t m = (int64_t)ins.size() / D;
            rc = vmn_engine_insert(idx, ins.data(), m, link_L, nullptr, VMN_MEM_HOST);
            if (rc) return rc;
            result->inserts += m;
        }
        const int64_t nq = (int64_t)qs.size() / D;
        if (nq) {
            const float* q = nullptr;
            rc = stage_rows(idx, qs.data(), nq, VMN_MEM_HOST, idx->qpad, &q);
            if (rc) return rc;
            HIPCHK(idx->eng_found.reserve((size_t)nq * K * sizeof(int32_t)));
            // BF16 rows only when the queries are exact in BF16 too
            int prec = VMN_PREC_FP32;
            rc = query_prec(idx, q, nq, &prec);
            if (rc) return rc;
            HIPCHK(hipEventRecord(ev0, idx->stream));
            rc = run_search(idx, q, nq, L, K, idx->eng_found.as<int32_t>(), nullptr, false,
                            VMN_MODE_PARTH, false, 0, prec);
            if (rc) return rc;
            HIPCHK(hipEventRecord(ev1, idx->stream));
            const int64_t n = idx->n;
            int splits = (int)std::max<int64_t>(
                1, std::min<int64_t>(256, (int64_t)idx->cus * 256 / std::max<int64_t>(nq, 1)));
            splits = (int)std::min<int64_t>(splits, n);
            HIPCHK(idx->eng_part_d.reserve((size_t)nq * splits * 16 * sizeof(double)));
            HIPCHK(idx->eng_part_i.reserve((size_t)nq * splits * 16 * sizeof(int32_t)));
            HIPCHK(idx->eng_gt_i.reserve((size_t)nq * K * sizeof(int32_t)));
            HIPCHK(idx->eng_gt_d.reserve((size_t)nq * K * sizeof(double)));
            HIPCHK(vmn::launch_gt(idx->vecs.as<float>(), n, DP, idx->del_bits.as<uint32_t>(), q, nq,
                                  K, splits, idx->eng_part_d.as<double>(),
                                  idx->eng_part_i.as<int32_t>(), idx->eng_gt_i.as<int32_t>(),
                                  idx->eng_gt_d.as<double>(), idx->stream, true));
            h_found.resize((size_t)nq * K);
            h_gt.resize((size_t)nq * K);
            h_hops.resize((size_t)nq);
            h_nd.resize((size_t)nq);
            HIPCHK(hipMemcpyAsync(h_hops.data(), idx->stat_hops.p, (size_t)nq * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, idx->stream));
            HIPCHK(hipMemcpyAsync(h_nd.data(), idx->stat_ndist.p, (size_t)nq * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, idx->stream));
            HIPCHK(hipMemcpyAsync(h_found.data(), idx->eng_found.p, h_found.size() * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, idx->stream));
            HIPCHK(hipMemcpyAsync(h_gt.data(), idx->eng_gt_i.p, h_gt.size() * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, idx->stream));
            HIPCHK(hipStreamSynchronize(idx->stream));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
            result->search_ms += ms;
            for (int64_t j = 0; j < nq; ++j) {
                result->expansions += h_hops[(size_t)j];
                result->dist_evals += h_nd[(size_t)j];
            }
            for (int64_t j = 0; j < nq; ++j) {
                int matches = 0;
                for (int a = 0; a < K; ++a) {
                    const int32_t f = h_found[(size_t)(j * K + a)];
                    if (f == -1) continue;
                    for (int g = 0; g < K; ++g)
                        if (h_gt[(size_t)(j * K + g)] == f) {
                            ++matches;
                            break;
                        }
                }
                recall_sum += matches / (double)K;
            }
            if (found_out)
                std::memcpy(found_out + result->searches * K, h_found.data(),
                            h_found.size() * sizeof(int32_t));
            if (gt_out)
                std::memcpy(gt_out + result->searches * K, h_gt.data(), h_gt.size() * sizeof(int32_t));
            result->searches += nq;
        }
        ++result->batches;
    }
    if (result->searches) {
        result->recall = recall_sum / (double)result->searches;
        result->qps = result->search_ms > 0 ? result->searches / (result->search_ms * 1e-3) : 0.0;
    }
    result->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return VMN_OK;
}
