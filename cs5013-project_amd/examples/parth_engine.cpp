// The parth engine's command-line driver in C++ over the C ABI of
// include/vamana_mi355x.h, in place of the reference's `main`
// (parth/src/engine.cu:709-729) and `run_automated_benchmark` (:578-600):
// read the graph and workload files, then for each beam of the sweep reload
// the graph (reset_gpu_state, :382-390), replay the workload in batches and
// print the row; <output_dir>/metrics.json keeps the last row, as the
// reference writes it.
//
// Build and run from the repository root:
//   g++ -O2 -std=c++17 -I include cs5013-project_amd/examples/parth_engine.cpp
//       -L cs5013-project_amd -lvmn -Wl,-rpath,$PWD/cs5013-project_amd -o parth_engine
//   ./parth_engine sift_initial_graph.bin sift_workload.bin 50 out_dir
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "vamana_mi355x.h"

static int fail(const char* what) {
    std::fprintf(stderr, "%s: %s\n", what, vmn_last_error());
    return 1;
}

int main(int argc, char** argv) {
    if (argc < 5) {
        std::printf("Usage: %s <initial_graph> <workload> <batch_size> <output_dir>\n", argv[0]);
        return 1;
    }
    const char* graph = argv[1];
    const char* workload = argv[2];
    const int batch = std::atoi(argv[3]);
    const char* out_dir = argv[4];
    // the graph header (int32 n, dim, max_degree) gives the record width
    int32_t hdr[3] = {0, 0, 0};
    FILE* f = std::fopen(graph, "rb");
    if (!f) {
        std::perror("fopen graph");
        return 1;
    }
    const bool got = std::fread(hdr, sizeof(int32_t), 3, f) == 3;
    std::fclose(f);
    if (!got) {
        std::fprintf(stderr, "short graph header\n");
        return 1;
    }
    const int dim = hdr[1];
    int64_t n_ops = 0;
    if (vmn_read_workload(workload, dim, nullptr, nullptr, nullptr, 0, &n_ops)) return fail("workload");
    std::vector<int32_t> types((size_t)n_ops), ids((size_t)n_ops);
    std::vector<float> vecs((size_t)n_ops * dim);
    if (vmn_read_workload(workload, dim, types.data(), ids.data(), vecs.data(), n_ops, &n_ops))
        return fail("workload");
    std::printf("Loaded graph: %d nodes\nLoaded %lld operations\n", hdr[0], (long long)n_ops);

    vmn_params p;
    vmn_default_params(&p, VMN_MODE_PARTH);
    p.capacity = dim == 128 ? 20000 : 1000;  // VamanaEngine<128, 20000> / <2, 1000>
    std::printf("Running Automated Benchmark...\n\n%-8s %-8s %-16s %-8s\n", "L", "Time", "QPS",
                "10-r@10");
    std::printf("%-8s %-8s %-16s %-8s\n", "---", "----", "---", "-------");
    const int sweep[] = {10, 20, 32, 64, 100};
    for (int L : sweep) {
        vmn_index* idx = nullptr;
        if (vmn_load_parth_graph(graph, &p, 0, &idx)) return fail("graph");
        vmn_engine_result r{};
        const int rc = vmn_engine_run(idx, types.data(), ids.data(), vecs.data(), n_ops, batch, L,
                                      128, &r, nullptr, nullptr);
        vmn_destroy(idx);
        if (rc) return fail("replay");
        std::printf("%-8d %-8.2f %-16.2f %-8.2f\n", L, r.qps > 0 ? n_ops / r.qps : 0.0, r.qps,
                    r.recall);
        char path[4096];
        std::snprintf(path, sizeof path, "%s/metrics.json", out_dir);
        if (FILE* m = std::fopen(path, "w")) {
            std::fprintf(m, "{\"qps\": %.2f, \"recall_10\": %.4f}\n", r.qps, r.recall);
            std::fclose(m);
        }
    }
    return 0;
}
