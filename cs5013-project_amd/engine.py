"""Host-side mirror of the parth engine's `VamanaEngine<DIM_SZ, MAX_N>`
(parth/src/engine.cu:304-700) over the C ABI.

`load_data`, `reset_gpu_state`, `process_workload` and
`run_automated_benchmark` keep the reference's names, arguments and outputs
(the sweep table and `metrics.json` with `qps` and `recall_10`). The graph
lives in a PARTH-mode index; every delete, insert, search and ground-truth
scan runs in libvmn.so (vmn_engine_run). Command line, as the reference's
`main` (engine.cu:709-729):

    python -m cs5013-project_amd.engine <initial_graph> <workload> <batch_size> <output_dir>
"""
import os
import sys
from collections import namedtuple

from .vamana import Vamana, read_workload

BenchmarkResult = namedtuple("BenchmarkResult", ["qps", "recall"])

# engine.cu:584, the beams of the automated sweep
SWEEP_L = (10, 20, 32, 64, 100)


class VamanaEngine:
    """dim and max_n are the reference's template arguments (128 / 20000 for
    SIFT, 2 / 1000 for the 2-D data, engine.cu:720-724); link_L is the beam of
    the insert's nearest-node search (BEAM_WIDTH, common.h:17)."""

    def __init__(self, dim=128, max_n=20000, device=0, link_L=128):
        self.dim, self.max_n, self.device, self.link_L = dim, max_n, device, link_L
        self.index = None
        self.graph_file = None
        self.workload = None
        self.last = None

    def load_data(self, graph_file, workload_file):
        """engine.cu:338-374: the parth graph file and the workload file."""
        self.graph_file = graph_file
        self.workload = read_workload(workload_file, self.dim)
        self.reset_gpu_state()

    def reset_gpu_state(self):
        """engine.cu:382-390: the loaded graph again, nothing deleted, an empty
        free list, searches from node 0."""
        if self.index is not None:
            self.index.close()
        self.index = Vamana.load_parth_graph(self.graph_file, medoid=0, capacity=self.max_n,
                                             device=self.device)
        if self.index.params.dim != self.dim:
            raise ValueError("Dim mismatch")

    def process_workload(self, batch_size, output_dir=None, beam_width=100):
        """engine.cu:407-568; writes output_dir/metrics.json like the reference."""
        t, i, v = self.workload
        r = self.index.run_workload(t, i, v, batch_size, beam_width, link_L=self.link_L)
        if output_dir:
            with open(os.path.join(output_dir, "metrics.json"), "w") as f:
                f.write('{"qps": %.2f, "recall_10": %.4f}\n' % (r["qps"], r["recall"]))
        self.last = r
        return BenchmarkResult(r["qps"], r["recall"])

    def run_automated_benchmark(self, graph_file, workload_file, batch_size, output_dir,
                                L_values=SWEEP_L):
        """engine.cu:578-600: reset, replay and print one row per beam."""
        self.load_data(graph_file, workload_file)
        print("Running Automated Benchmark...")
        print("\n%-8s %-8s %-16s %-8s" % ("L", "Time", "QPS", "10-r@10"))
        print("%-8s %-8s %-16s %-8s" % ("---", "----", "---", "-------"))
        rows = []
        for L in L_values:
            self.reset_gpu_state()
            res = self.process_workload(batch_size, output_dir, L)
            time_sec = len(self.workload[0]) / res.qps if res.qps else 0.0
            print("%-8d %-8.2f %-16.2f %-8.2f" % (L, time_sec, res.qps, res.recall))
            rows.append({"L": L, "qps": res.qps, "recall": res.recall,
                         "search_ms": self.last["search_ms"]})
        return rows

    def close(self):
        if self.index is not None:
            self.index.close()
            self.index = None


def main(argv):
    if len(argv) < 5:
        print("Usage: %s <initial_graph> <workload> <batch_size> <output_dir>" % argv[0])
        return 1
    with open(argv[1], "rb") as f:
        dim = int.from_bytes(f.read(12)[4:8], "little", signed=True)
    eng = VamanaEngine(dim=dim, max_n=20000 if dim == 128 else 1000)
    eng.run_automated_benchmark(argv[1], argv[2], int(argv[3]), argv[4])
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
