// Index persistence and the reference's graph files (host-side I/O behind the
// C ABI of include/vamana_mi355x.h).
//
// * Index file (vmn_save / vmn_load): everything a later search or update
//   needs, so a loaded index behaves exactly like the saved one: parameters,
//   vectors, adjacency, tombstones, entry point and the FP8 per-dimension
//   scales and centre (fixed at the build; re-deriving them from all rows
//   after inserts would change the FP8 rows). Little-endian:
//     char magic[8] = "VMNIDX01"; u32 version = 2; u32 DP;
//     i32 dim, R, L_build, passes; f32 alpha[4]; u64 seed;
//     i32 mode, visited_cap, max_hops, P, M; i64 capacity;
//     i32 low_precision, fp8_scaled; i64 n, medoid, n_deleted;
//     f32 vectors[n x dim]; i32 adjacency[n x R] (-1 padded);
//     u32 tombstone words[(n + 31) / 32];
//     if fp8_scaled: f32 fp8_scales[DP], f32 fp8_centre[DP].
//   Version-1 files (no centre: their FP8 rows were coded about 0) load with
//   a zero centre.
// * FreshVamana record file (saveGraph vamana.cu:1346-1362, read by initGraph
//   utils.cuh:113-148): n records of { f32 vec[dim]; u32 degree; u32
//   nbr[R] }, no header (772 bytes at dim 128, R 64).
// * parth graph file (written by build_graph.cu:361-369, read by
//   engine.cu:338-357): i32 n, dim, max_degree; f32 vectors[n x dim];
//   i32 adjacency[n x max_degree] (-1 padded).
// * vector files (fvecs / ivecs / bvecs, and the .bin form with a u32 n, dim
//   header), DiskANN ground-truth files (u32 n, k; ids; optional distances)
//   and the BANG / DiskANN disk index (772-byte records with ascending
//   neighbour ids plus a 32-byte metadata file).
#include <cstdio>
#include <cstring>
#include <vector>

#include "vmn_internal.hpp"

using namespace vmn_rt;

namespace {

struct File {
    FILE* f = nullptr;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
    bool put(const void* p, size_t n) { return std::fwrite(p, 1, n, f) == n; }
    bool get(void* p, size_t n) { return std::fread(p, 1, n, f) == n; }
    template <class T>
    bool put(const T& v) { return put(&v, sizeof v); }
    template <class T>
    bool get(T& v) { return get(&v, sizeof v); }
};

const char kMagic[8] = {'V', 'M', 'N', 'I', 'D', 'X', '0', '1'};

// Vectors [n x dim], adjacency [n x R] and tombstone words of a built index,
// copied to the host.
int fetch(const vmn_index* idx, std::vector<float>& vecs, std::vector<int32_t>& adj,
          std::vector<uint32_t>* del) {
    if (!idx->built) return fail(VMN_ERR_STATE, "index not built");
    if (idx->n_pend)
        return fail(VMN_ERR_STATE, "%lld staged inserts pending: vmn_consolidate adds them first",
                    (long long)idx->n_pend);
    int rc = set_device(idx);
    if (rc) return rc;
    const int64_t n = idx->n;
    vecs.resize((size_t)n * idx->p.dim);
    adj.resize((size_t)n * idx->p.R);
    HIPCHK(hipStreamSynchronize(idx->stream));
    HIPCHK(hipMemcpy2D(vecs.data(), (size_t)idx->p.dim * sizeof(float), idx->vecs.p,
                       (size_t)idx->DP * sizeof(float), (size_t)idx->p.dim * sizeof(float),
                       (size_t)n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(adj.data(), idx->adj.p, adj.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (del) {
        del->resize((size_t)(n + 31) / 32);
        HIPCHK(hipMemcpy(del->data(), idx->del_bits.p, del->size() * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
        if (n % 32) del->back() &= (1u << (n % 32)) - 1u;
    }
    return VMN_OK;
}

long file_size(FILE* f) {
    if (std::fseek(f, 0, SEEK_END) != 0) return -1;
    const long s = std::ftell(f);
    std::rewind(f);
    return s;
}

}  // namespace

int vmn_save(const vmn_index* idx, const char* path) {
    if (!idx || !path) return fail(VMN_ERR_ARG, "NULL argument");
    std::vector<float> vecs;
    std::vector<int32_t> adj;
    std::vector<uint32_t> del;
    int rc = fetch(idx, vecs, adj, &del);
    if (rc) return rc;
    const bool fp8 = (idx->p.low_precision & VMN_STORE_FP8) && idx->fp8_scaled;
    std::vector<float> scales, centre;
    if (fp8) {
        scales.resize(idx->DP);
        centre.resize(idx->DP);
        HIPCHK(hipMemcpy(scales.data(), idx->fp8_scales.p, scales.size() * sizeof(float),
                         hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(centre.data(), idx->fp8_centre.p, centre.size() * sizeof(float),
                         hipMemcpyDeviceToHost));
    }
    File f(path, "wb");
    if (!f.f) return fail(VMN_ERR_IO, "cannot open %s for writing", path);
    const vmn_params& p = idx->p;
    bool ok = f.put(kMagic, 8) && f.put((uint32_t)2) && f.put((uint32_t)idx->DP) &&
              f.put(p.dim) && f.put(p.R) && f.put(p.L_build) && f.put(p.passes) &&
              f.put(p.alpha, sizeof p.alpha) && f.put(p.seed) && f.put(p.mode) &&
              f.put(p.visited_cap) && f.put(p.max_hops) && f.put(p.P) && f.put(p.M) &&
              f.put(p.capacity) && f.put(p.low_precision) && f.put((int32_t)fp8) &&
              f.put(idx->n) && f.put(idx->medoid) && f.put(idx->n_deleted) &&
              f.put(vecs.data(), vecs.size() * sizeof(float)) &&
              f.put(adj.data(), adj.size() * sizeof(int32_t)) &&
              f.put(del.data(), del.size() * sizeof(uint32_t)) &&
              (!fp8 || (f.put(scales.data(), scales.size() * sizeof(float)) &&
                        f.put(centre.data(), centre.size() * sizeof(float))));
    if (!ok) return fail(VMN_ERR_IO, "short write to %s", path);
    return VMN_OK;
}

int vmn_load(const char* path, int32_t device, vmn_index** out) {
    if (!path || !out) return fail(VMN_ERR_ARG, "NULL argument");
    *out = nullptr;
    File f(path, "rb");
    if (!f.f) return fail(VMN_ERR_IO, "cannot open %s", path);
    char magic[8];
    uint32_t version = 0, DP = 0;
    if (!f.get(magic, 8) || std::memcmp(magic, kMagic, 8) != 0 || !f.get(version) ||
        (version != 1 && version != 2) || !f.get(DP))
        return fail(VMN_ERR_IO, "%s is not a version-1 or version-2 index file", path);
    vmn_params p{};
    int32_t fp8 = 0;
    int64_t n = 0, medoid = 0, n_deleted = 0;
    bool ok = f.get(p.dim) && f.get(p.R) && f.get(p.L_build) && f.get(p.passes) &&
              f.get(p.alpha, sizeof p.alpha) && f.get(p.seed) && f.get(p.mode) &&
              f.get(p.visited_cap) && f.get(p.max_hops) && f.get(p.P) && f.get(p.M) &&
              f.get(p.capacity) && f.get(p.low_precision) && f.get(fp8) && f.get(n) &&
              f.get(medoid) && f.get(n_deleted);
    if (!ok) return fail(VMN_ERR_IO, "truncated header in %s", path);
    p.device = device;
    int rc = check_params(&p);
    if (rc) return rc;
    if (n < 1 || n >= (int64_t)1 << 31 || medoid < 0 || medoid >= n || (p.capacity && n > p.capacity))
        return fail(VMN_ERR_IO, "bad sizes in %s", path);
    std::vector<float> vecs((size_t)n * p.dim);
    std::vector<int32_t> adj((size_t)n * p.R);
    std::vector<uint32_t> del((size_t)(n + 31) / 32);
    std::vector<float> scales(fp8 ? DP : 0), centre(fp8 ? DP : 0, 0.0f);
    ok = f.get(vecs.data(), vecs.size() * sizeof(float)) &&
         f.get(adj.data(), adj.size() * sizeof(int32_t)) &&
         f.get(del.data(), del.size() * sizeof(uint32_t)) &&
         (!fp8 || f.get(scales.data(), scales.size() * sizeof(float))) &&
         (!fp8 || version < 2 || f.get(centre.data(), centre.size() * sizeof(float)));
    if (!ok) return fail(VMN_ERR_IO, "truncated body in %s", path);
    for (int32_t v : adj)
        if (v < -1 || v >= n) return fail(VMN_ERR_IO, "adjacency id %d out of range in %s", v, path);

    vmn_index* idx = nullptr;
    rc = vmn_create(&p, &idx);
    if (rc) return rc;
    if ((uint32_t)idx->DP != DP) {
        vmn_destroy(idx);
        return fail(VMN_ERR_IO, "%s was written with row stride %u, this build uses %d", path, DP,
                    idx->DP);
    }
    struct Guard {
        vmn_index* i;
        ~Guard() {
            if (i) vmn_destroy(i);
        }
    } guard{idx};
    rc = alloc_graph(idx, n);
    if (rc) return rc;
    idx->n = n;
    const float* staged = nullptr;
    rc = stage_rows(idx, vecs.data(), n, VMN_MEM_HOST, idx->vecs, &staged);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(idx->adj.p, adj.data(), adj.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, idx->stream));
    HIPCHK(hipMemcpyAsync(idx->del_bits.p, del.data(), del.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice, idx->stream));
    idx->medoid = medoid;
    idx->n_deleted = n_deleted;
    rc = update_exact(idx, 0, n);
    if (rc) return rc;
    if (fp8) {
        HIPCHK(idx->fp8_scales.reserve((size_t)DP * sizeof(float)));
        HIPCHK(idx->fp8_centre.reserve((size_t)DP * sizeof(float)));
        HIPCHK(hipMemcpyAsync(idx->fp8_scales.p, scales.data(), scales.size() * sizeof(float),
                              hipMemcpyHostToDevice, idx->stream));
        HIPCHK(hipMemcpyAsync(idx->fp8_centre.p, centre.data(), centre.size() * sizeof(float),
                              hipMemcpyHostToDevice, idx->stream));
        idx->fp8_scaled = true;
    } else {
        idx->fp8_scaled = false;
    }
    rc = refresh_lowp(idx, 0, n);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(idx->stream));
    idx->built = true;
    guard.i = nullptr;
    *out = idx;
    return VMN_OK;
}

int vmn_save_fresh_graph(const vmn_index* idx, const char* path) {
    if (!idx || !path) return fail(VMN_ERR_ARG, "NULL argument");
    std::vector<float> vecs;
    std::vector<int32_t> adj;
    int rc = fetch(idx, vecs, adj, nullptr);
    if (rc) return rc;
    const int D = idx->p.dim, R = idx->p.R;
    File f(path, "wb");
    if (!f.f) return fail(VMN_ERR_IO, "cannot open %s for writing", path);
    std::vector<uint32_t> nbr(R);
    for (int64_t i = 0; i < idx->n; ++i) {
        uint32_t deg = 0;
        for (int j = 0; j < R; ++j) {
            const int32_t v = adj[(size_t)i * R + j];
            nbr[j] = (uint32_t)v;
            if (v >= 0) ++deg;
        }
        if (!f.put(&vecs[(size_t)i * D], (size_t)D * sizeof(float)) || !f.put(deg) ||
            !f.put(nbr.data(), (size_t)R * sizeof(uint32_t)))
            return fail(VMN_ERR_IO, "short write to %s", path);
    }
    return VMN_OK;
}

int vmn_load_fresh_graph(const char* path, const vmn_params* p, int32_t build, int64_t medoid,
                         vmn_index** out) {
    if (!path || !p || !out) return fail(VMN_ERR_ARG, "NULL argument");
    *out = nullptr;
    int rc = check_params(p);
    if (rc) return rc;
    const int D = p->dim, R = p->R;
    File f(path, "rb");
    if (!f.f) return fail(VMN_ERR_IO, "cannot open %s", path);
    const long size = file_size(f.f);
    const long rec = (long)D * 4 + 4 + (long)R * 4;
    if (size <= 0 || size % rec != 0)
        return fail(VMN_ERR_IO, "%s: %ld bytes is not a whole number of %ld-byte records (dim %d, R %d)",
                    path, size, rec, D, R);
    const int64_t n = size / rec;
    std::vector<float> vecs((size_t)n * D);
    std::vector<int32_t> adj((size_t)n * R, -1);
    std::vector<uint32_t> nbr(R);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t deg = 0;
        if (!f.get(&vecs[(size_t)i * D], (size_t)D * sizeof(float)) || !f.get(deg) ||
            !f.get(nbr.data(), (size_t)R * sizeof(uint32_t)))
            return fail(VMN_ERR_IO, "short read from %s", path);
        if (deg > (uint32_t)R) return fail(VMN_ERR_IO, "%s: record %lld has degree %u > R", path,
                                           (long long)i, deg);
        for (uint32_t j = 0; j < deg; ++j) {
            if (nbr[j] >= (uint64_t)n)
                return fail(VMN_ERR_IO, "%s: record %lld names node %u", path, (long long)i, nbr[j]);
            adj[(size_t)i * R + j] = (int32_t)nbr[j];
        }
    }
    vmn_index* idx = nullptr;
    rc = vmn_create(p, &idx);
    if (rc) return rc;
    rc = build ? vmn_build_from(idx, vecs.data(), n, adj.data(), R, VMN_MEM_HOST)
               : vmn_load_graph(idx, vecs.data(), n, adj.data(), R, medoid, VMN_MEM_HOST);
    if (rc) {
        vmn_destroy(idx);
        return rc;
    }
    *out = idx;
    return VMN_OK;
}

int vmn_save_parth_graph(const vmn_index* idx, const char* path) {
    if (!idx || !path) return fail(VMN_ERR_ARG, "NULL argument");
    std::vector<float> vecs;
    std::vector<int32_t> adj;
    int rc = fetch(idx, vecs, adj, nullptr);
    if (rc) return rc;
    File f(path, "wb");
    if (!f.f) return fail(VMN_ERR_IO, "cannot open %s for writing", path);
    const int32_t hdr[3] = {(int32_t)idx->n, idx->p.dim, idx->p.R};
    if (!f.put(hdr, sizeof hdr) || !f.put(vecs.data(), vecs.size() * sizeof(float)) ||
        !f.put(adj.data(), adj.size() * sizeof(int32_t)))
        return fail(VMN_ERR_IO, "short write to %s", path);
    return VMN_OK;
}

int vmn_load_parth_graph(const char* path, const vmn_params* p, int64_t medoid, vmn_index** out) {
    if (!path || !p || !out) return fail(VMN_ERR_ARG, "NULL argument");
    *out = nullptr;
    File f(path, "rb");
    if (!f.f) return fail(VMN_ERR_IO, "cannot open %s", path);
    int32_t hdr[3];
    if (!f.get(hdr, sizeof hdr)) return fail(VMN_ERR_IO, "short read from %s", path);
    const int64_t n = hdr[0];
    vmn_params q = *p;
    q.dim = hdr[1];
    q.R = hdr[2];
    int rc = check_params(&q);
    if (rc) return rc;
    const long want = 12 + (long)n * q.dim * 4 + (long)n * q.R * 4;
    if (n < 1 || file_size(f.f) != want) return fail(VMN_ERR_IO, "%s: size does not match its header", path);
    if (!f.get(hdr, sizeof hdr)) return fail(VMN_ERR_IO, "short read from %s", path);
    std::vector<float> vecs((size_t)n * q.dim);
    std::vector<int32_t> adj((size_t)n * q.R);
    if (!f.get(vecs.data(), vecs.size() * sizeof(float)) ||
        !f.get(adj.data(), adj.size() * sizeof(int32_t)))
        return fail(VMN_ERR_IO, "short read from %s", path);
    vmn_index* idx = nullptr;
    rc = vmn_create(&q, &idx);
    if (rc) return rc;
    rc = vmn_load_graph(idx, vecs.data(), n, adj.data(), q.R, medoid, VMN_MEM_HOST);
    if (rc) {
        vmn_destroy(idx);
        return rc;
    }
    *out = idx;
    return VMN_OK;
}

// ---------------------------------------------------------------------------
// Vector files, ground-truth files and the BANG / DiskANN disk index.
// ---------------------------------------------------------------------------

namespace {

int elem_bytes(int32_t format) {
    switch (format) {
        case VMN_FILE_FVECS: return 4;
        case VMN_FILE_IVECS: return 4;
        case VMN_FILE_BVECS: return 1;
        case VMN_FILE_BIN: return 4;
        default: return 0;
    }
}

}  // namespace

int vmn_read_vectors(const char* path, int32_t format, void* out, int64_t capacity, int64_t* n,
                     int32_t* dim) {
    if (!path || !n || !dim) return fail(VMN_ERR_ARG, "NULL argument");
    const int eb = elem_bytes(format);
    if (!eb) return fail(VMN_ERR_ARG, "unknown vector file format %d", format);
    File f(path, "rb");
    if (!f.f) return fail(VMN_ERR_IO, "cannot open %s", path);
    const long size = file_size(f.f);
    int64_t count = 0;
    int32_t d = 0;
    if (format == VMN_FILE_BIN) {
        uint32_t hdr[2];
        if (!f.get(hdr, sizeof hdr)) return fail(VMN_ERR_IO, "%s: no header", path);
        count = hdr[0];
        d = (int32_t)hdr[1];
        // a header may claim more vectors than the file holds (the reference's
        // data/sift10k/siftsmall_query.bin says 10,000 for its 100): the
        // vectors present are read
        const int64_t held = d >= 1 ? (size - 8) / ((int64_t)d * 4) : 0;
        if (d < 1 || (size - 8) % ((int64_t)d * 4) != 0 || held > count)
            return fail(VMN_ERR_IO, "%s: %ld bytes do not hold whole %d-float vectors (header %lld)",
                        path, size, d, (long long)count);
        count = held;
    } else {
        if (!f.get(d) || d < 1) return fail(VMN_ERR_IO, "%s: bad first record", path);
        const int64_t rec = 4 + (int64_t)d * eb;
        if (size % rec != 0)
            return fail(VMN_ERR_IO, "%s: %ld bytes is not a whole number of %lld-byte records", path,
                        size, (long long)rec);
        count = size / rec;
        std::rewind(f.f);
    }
    *n = count;
    *dim = d;
    if (!out) return VMN_OK;
    const int64_t m = count < capacity ? count : capacity;
    char* o = static_cast<char*>(out);
    if (format == VMN_FILE_BIN) {
        if (!f.get(o, (size_t)(m * d) * 4)) return fail(VMN_ERR_IO, "short read from %s", path);
    } else {
        for (int64_t i = 0; i < m; ++i) {
            int32_t di = 0;
            if (!f.get(di) || !f.get(o + (size_t)(i * d) * eb, (size_t)d * eb))
                return fail(VMN_ERR_IO, "short read from %s", path);
            if (di != d)
                return fail(VMN_ERR_IO, "%s: record %lld has dimension %d, not %d", path,
                            (long long)i, di, d);
        }
    }
    *n = m;
    return VMN_OK;
}

int vmn_write_vectors(const char* path, int32_t format, const void* vecs, int64_t n, int32_t dim) {
    if (!path || (!vecs && n)) return fail(VMN_ERR_ARG, "NULL argument");
    const int eb = elem_bytes(format);
    if (!eb) return fail(VMN_ERR_ARG, "unknown vector file format %d", format);
    if (n < 0 || dim < 1 || (format == VMN_FILE_BIN && n > (int64_t)UINT32_MAX))
        return fail(VMN_ERR_ARG, "bad shape");
    File f(path, "wb");
    if (!f.f) return fail(VMN_ERR_IO, "cannot open %s for writing", path);
    const char* v = static_cast<const char*>(vecs);
    bool ok = true;
    if (format == VMN_FILE_BIN) {
        const uint32_t hdr[2] = {(uint32_t)n, (uint32_t)dim};
        ok = f.put(hdr, sizeof hdr) && f.put(v, (size_t)(n * dim) * eb);
    } else {
        for (int64_t i = 0; ok && i < n; ++i)
            ok = f.put(dim) && f.put(v + (size_t)(i * dim) * eb, (size_t)dim * eb);
    }
    if (!ok) return fail(VMN_ERR_IO, "short write to %s", path);
    return VMN_OK;
}

int vmn_read_truthset(const char* path, int32_t* ids, float* dists, int64_t capacity, int64_t* n,
                      int32_t* k, int32_t* has_dists) {
    if (!path || !n || !k) return fail(VMN_ERR_ARG, "NULL argument");
    File f(path, "rb");
    if (!f.f) return fail(VMN_ERR_IO, "cannot open %s", path);
    const long size = file_size(f.f);
    uint32_t hdr[2];
    if (!f.get(hdr, sizeof hdr)) return fail(VMN_ERR_IO, "%s: no header", path);
    const int64_t cnt = hdr[0], kk = hdr[1];
    const int64_t cell = cnt * kk * 4;
    int with = 0;
    if (size == 8 + 2 * cell) with = 1;
    else if (size != 8 + cell)
        return fail(VMN_ERR_IO, "%s: %ld bytes fit neither %lld x %lld ids nor ids + distances",
                    path, size, (long long)cnt, (long long)kk);
    *n = cnt;
    *k = (int32_t)kk;
    if (has_dists) *has_dists = with;
    if (!ids) return VMN_OK;
    const int64_t m = cnt < capacity ? cnt : capacity;
    if (!f.get(ids, (size_t)(m * kk) * 4)) return fail(VMN_ERR_IO, "short read from %s", path);
    if (dists && with) {
        if (std::fseek(f.f, (long)(8 + cell), SEEK_SET) != 0 ||
            !f.get(dists, (size_t)(m * kk) * 4))
            return fail(VMN_ERR_IO, "short read from %s", path);
    }
    *n = m;
    return VMN_OK;
}

int vmn_write_truthset(const char* path, const int32_t* ids, const float* dists, int64_t n,
                       int32_t k) {
    if (!path || (!ids && n)) return fail(VMN_ERR_ARG, "NULL argument");
    if (n < 0 || k < 1 || n > (int64_t)UINT32_MAX) return fail(VMN_ERR_ARG, "bad shape");
    File f(path, "wb");
    if (!f.f) return fail(VMN_ERR_IO, "cannot open %s for writing", path);
    const uint32_t hdr[2] = {(uint32_t)n, (uint32_t)k};
    bool ok = f.put(hdr, sizeof hdr) && f.put(ids, (size_t)(n * k) * 4);
    if (ok && dists) ok = f.put(dists, (size_t)(n * k) * 4);
    if (!ok) return fail(VMN_ERR_IO, "short write to %s", path);
    return VMN_OK;
}

int vmn_save_bang_index(const vmn_index* idx, const char* index_path, const char* metadata_path) {
    if (!idx || !index_path || !metadata_path) return fail(VMN_ERR_ARG, "NULL argument");
    std::vector<float> vecs;
    std::vector<int32_t> adj;
    int rc = fetch(idx, vecs, adj, nullptr);
    if (rc) return rc;
    const int D = idx->p.dim, R = idx->p.R;
    {
        File f(index_path, "wb");
        if (!f.f) return fail(VMN_ERR_IO, "cannot open %s for writing", index_path);
        std::vector<uint32_t> nbr(R);
        for (int64_t i = 0; i < idx->n; ++i) {
            uint32_t deg = 0;
            for (int j = 0; j < R; ++j)
                if (adj[(size_t)i * R + j] >= 0) nbr[deg++] = (uint32_t)adj[(size_t)i * R + j];
            std::sort(nbr.begin(), nbr.begin() + deg);
            for (int j = (int)deg; j < R; ++j) nbr[j] = 0xffffffffu;
            if (!f.put(&vecs[(size_t)i * D], (size_t)D * sizeof(float)) || !f.put(deg) ||
                !f.put(nbr.data(), (size_t)R * sizeof(uint32_t)))
                return fail(VMN_ERR_IO, "short write to %s", index_path);
        }
    }
    File m(metadata_path, "wb");
    if (!m.f) return fail(VMN_ERR_IO, "cannot open %s for writing", metadata_path);
    const uint64_t med = (uint64_t)idx->medoid, rec = (uint64_t)D * 4 + 4 + (uint64_t)R * 4;
    const uint32_t tail[4] = {2u, (uint32_t)D, (uint32_t)R, (uint32_t)idx->n};
    if (!m.put(med) || !m.put(rec) || !m.put(tail, sizeof tail))
        return fail(VMN_ERR_IO, "short write to %s", metadata_path);
    return VMN_OK;
}

int vmn_load_bang_index(const char* index_path, const char* metadata_path, const vmn_params* p,
                        vmn_index** out) {
    if (!index_path || !metadata_path || !p || !out) return fail(VMN_ERR_ARG, "NULL argument");
    *out = nullptr;
    uint64_t med = 0, rec = 0;
    uint32_t tail[4];
    {
        File m(metadata_path, "rb");
        if (!m.f) return fail(VMN_ERR_IO, "cannot open %s", metadata_path);
        if (file_size(m.f) != 32 || !m.get(med) || !m.get(rec) || !m.get(tail, sizeof tail))
            return fail(VMN_ERR_IO, "%s is not a 32-byte index metadata file", metadata_path);
    }
    if (tail[0] != 2u)
        return fail(VMN_ERR_IO, "%s: data type %u (only 2 = float is read)", metadata_path, tail[0]);
    vmn_params q = *p;
    q.dim = (int32_t)tail[1];
    q.R = (int32_t)tail[2];
    const int64_t n = tail[3];
    if (rec != (uint64_t)q.dim * 4 + 4 + (uint64_t)q.R * 4 || med >= (uint64_t)n)
        return fail(VMN_ERR_IO, "%s: record size or medoid does not fit its shape", metadata_path);
    {
        File f(index_path, "rb");
        if (!f.f) return fail(VMN_ERR_IO, "cannot open %s", index_path);
        if (file_size(f.f) != (long)(n * (int64_t)rec))
            return fail(VMN_ERR_IO, "%s: size is not %lld records of %llu bytes", index_path,
                        (long long)n, (unsigned long long)rec);
    }
    return vmn_load_fresh_graph(index_path, &q, 0, (int64_t)med, out);
}
