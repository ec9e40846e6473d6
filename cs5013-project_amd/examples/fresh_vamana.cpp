// The FreshVamana driver of the reference (vamana/src/main.cu:151-269) in C++
// over the C ABI of include/vamana_mi355x.h:
//   * build the index from a random-graph file of 772-byte records (initGraph,
//     utils.cuh:113-148, then the Vamana<T> constructor, vamana_impl.cuh:10-20);
//   * search for the first base vector and print its 5 nearest rows
//     (searchPoints2 + printFinalVectors);
//   * stage that vector with 1 added to its first coordinate (insertPoints),
//     search for the shifted vector (the staged row is merged in by exact
//     distance), patch (patchGraph adds it to the graph) and search again;
//   * write the graph as records (saveGraph, vamana.cu:1346-1362).
//
// Build and run from the repository root:
//   g++ -O2 -std=c++17 -I include cs5013-project_amd/examples/fresh_vamana.cpp
//       -L cs5013-project_amd -lvmn -Wl,-rpath,$PWD/cs5013-project_amd -o fresh_vamana
//   ./fresh_vamana sift10k_randomgraph.bin unused out_graph.bin
#include <cstdint>
#include <cstdio>
#include <vector>

#include "vamana_mi355x.h"

static int fail(const char* what) {
    std::fprintf(stderr, "%s: %s\n", what, vmn_last_error());
    return 1;
}

// ids and distances of the 5 nearest, then the first 4 coordinates of each row
static int show(vmn_index* idx, const float* q, int dim) {
    const int k = 5, L = 64;
    int32_t ids[k];
    float d[k];
    std::vector<float> rows((size_t)k * dim);
    if (vmn_search(idx, q, 1, k, L, ids, d, VMN_MEM_HOST)) return fail("search");
    if (vmn_search_vectors(idx, q, 1, k, L, rows.data(), VMN_MEM_HOST)) return fail("search");
    for (int j = 0; j < k; ++j)
        std::printf("  id %d dist %.1f vec [%.1f %.1f %.1f %.1f ...]\n", ids[j], d[j],
                    rows[(size_t)j * dim], rows[(size_t)j * dim + 1], rows[(size_t)j * dim + 2],
                    rows[(size_t)j * dim + 3]);
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 4) {
        std::printf("Usage: %s <graph> <basepoints> <output>\n", argv[0]);
        return 1;
    }
    vmn_params p;
    vmn_default_params(&p, VMN_MODE_FRESH);  // G2 semantics: alpha 1.5, R 64, L 128
    p.dim = 128;
    p.L_build = 150;
    // room to insert past the file's records
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) {
        std::perror("fopen graph");
        return 1;
    }
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fclose(f);
    p.capacity = bytes / (p.dim * 4 + 4 + p.R * 4) + 16;
    vmn_index* idx = nullptr;
    if (vmn_load_fresh_graph(argv[1], &p, 1, -1, &idx)) return fail("build");
    std::printf("built %lld nodes, medoid %lld\n", (long long)vmn_size(idx),
                (long long)vmn_medoid(idx));

    const int dim = p.dim;
    std::vector<float> all((size_t)vmn_size(idx) * dim);
    if (vmn_get_vectors(idx, all.data(), VMN_MEM_HOST)) return fail("vectors");
    std::vector<float> q(all.begin(), all.begin() + dim);
    std::printf("search node 0:\n");
    if (show(idx, q.data(), dim)) return 1;

    q[0] += 1.0f;
    int32_t new_id = -1;
    if (vmn_insert_staged(idx, q.data(), 1, &new_id, VMN_MEM_HOST)) return fail("insert");
    std::printf("inserted id %d\nsearch the inserted vector:\n", new_id);
    if (show(idx, q.data(), dim)) return 1;
    if (vmn_consolidate(idx)) return fail("patch");
    std::printf("after patch:\n");
    if (show(idx, q.data(), dim)) return 1;

    if (vmn_save_fresh_graph(idx, argv[3])) return fail("save");
    std::printf("saved %lld records to %s\n", (long long)vmn_size(idx), argv[3]);
    vmn_destroy(idx);
    return 0;
}
