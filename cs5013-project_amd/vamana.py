"""Python mirror of the reference's index API over the C ABI (libvmn.so).

`Vamana` follows the reference's `Vamana<T>` (vamana/include/vamana.cuh:36-146):
the constructor builds the graph, `searchPoints` runs a batched search, and the
graph can be read back. Inputs may be numpy arrays (host) or torch tensors on
the GPU (device pointers, no copy). The HIP library is required: nothing here
falls back to a CPU path; a missing or unloadable libvmn.so raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvmn.so")

MODE_PARTH = 0
MODE_FRESH = 1
MEM_HOST = 0
MEM_DEVICE = 1
PREC = {"fp32": 0, "bf16": 1, "fp8": 2}
STORE_BF16 = 1
STORE_FP8 = 2

STATUS = {0: "VMN_OK", -1: "VMN_ERR_ARG", -2: "VMN_ERR_HIP", -3: "VMN_ERR_STATE",
          -4: "VMN_ERR_CAPACITY", -5: "VMN_ERR_IO"}


class VamanaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    """Mirror of `vmn_params` (include/vamana_mi355x.h)."""
    _fields_ = [("dim", ctypes.c_int32), ("R", ctypes.c_int32), ("L_build", ctypes.c_int32),
                ("passes", ctypes.c_int32), ("alpha", ctypes.c_float * 4),
                ("seed", ctypes.c_uint64), ("mode", ctypes.c_int32),
                ("visited_cap", ctypes.c_int32), ("max_hops", ctypes.c_int32),
                ("P", ctypes.c_int32), ("M", ctypes.c_int32), ("capacity", ctypes.c_int64),
                ("device", ctypes.c_int32), ("low_precision", ctypes.c_int32),
                ("build_chunk", ctypes.c_int64)]


class ServerStats(ctypes.Structure):
    """Mirror of `vmn_server_stats`."""
    _fields_ = [("batches", ctypes.c_int64), ("queries", ctypes.c_int64),
                ("max_batch_queries", ctypes.c_int64), ("inserts", ctypes.c_int64),
                ("deletes", ctypes.c_int64), ("consolidations", ctypes.c_int64),
                ("insert_s", ctypes.c_double), ("delete_s", ctypes.c_double),
                ("max_update_s", ctypes.c_double), ("drain_s", ctypes.c_double),
                ("failed_batches", ctypes.c_int64)]


class BuildStats(ctypes.Structure):
    """Mirror of `vmn_build_stats`."""
    _fields_ = [("searches", ctypes.c_int64), ("expansions", ctypes.c_int64),
                ("dist_evals", ctypes.c_int64), ("prune_items", ctypes.c_int64),
                ("prune_rows", ctypes.c_int64), ("reverse_pairs", ctypes.c_int64),
                ("list_entries", ctypes.c_int64), ("reverse_entries", ctypes.c_int64),
                ("row_bytes", ctypes.c_int32), ("passes", ctypes.c_int32),
                ("algorithmic_bytes", ctypes.c_double)]


class EngineResult(ctypes.Structure):
    _fields_ = [("searches", ctypes.c_int64), ("inserts", ctypes.c_int64),
                ("deletes", ctypes.c_int64), ("batches", ctypes.c_int64),
                ("search_ms", ctypes.c_double), ("qps", ctypes.c_double),
                ("recall", ctypes.c_double), ("wall_s", ctypes.c_double),
                ("expansions", ctypes.c_int64), ("dist_evals", ctypes.c_int64)]


class ServerParams(ctypes.Structure):
    """Mirror of `vmn_server_params`."""
    _fields_ = [("k", ctypes.c_int32), ("L", ctypes.c_int32), ("prec", ctypes.c_int32),
                ("k_rerank", ctypes.c_int32), ("max_batch", ctypes.c_int32),
                ("n_slots", ctypes.c_int32), ("max_queued_bytes", ctypes.c_int64),
                ("consolidate_every", ctypes.c_int64), ("device", ctypes.c_int32)]


REQ_SEARCH, REQ_INSERT, REQ_DELETE = 0, 1, 2

_lib = None


def lib():
    """Load libvmn.so (built by build_lib.py). Raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.vmn_default_params.argtypes = [ctypes.POINTER(Params), i32]
        L.vmn_default_params.restype = None
        L.vmn_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(P)]
        L.vmn_destroy.argtypes = [P]
        L.vmn_build.argtypes = [P, P, i64, i32]
        L.vmn_build_from.argtypes = [P, P, i64, P, i32, i32]
        L.vmn_load_graph.argtypes = [P, P, i64, P, i32, i64, i32]
        cs = ctypes.c_char_p
        L.vmn_save.argtypes = [P, cs]
        L.vmn_load.argtypes = [cs, i32, ctypes.POINTER(P)]
        L.vmn_save_fresh_graph.argtypes = [P, cs]
        L.vmn_load_fresh_graph.argtypes = [cs, ctypes.POINTER(Params), i32, i64, ctypes.POINTER(P)]
        L.vmn_save_parth_graph.argtypes = [P, cs]
        L.vmn_load_parth_graph.argtypes = [cs, ctypes.POINTER(Params), i64, ctypes.POINTER(P)]
        L.vmn_get_vectors.argtypes = [P, P, i32]
        L.vmn_get_params.argtypes = [P, ctypes.POINTER(Params)]
        L.vmn_search.argtypes = [P, P, i64, i32, i32, P, P, i32]
        L.vmn_search_ex.argtypes = [P, P, i64, i32, i32, i32, i32, P, P, i32]
        L.vmn_search_vectors.argtypes = [P, P, i64, i32, i32, P, i32]
        L.vmn_search_noalloc.argtypes = [P, P, i64, i32, i32, i32, i32, P, P]
        L.vmn_bang_search.argtypes = [P, P, i64, i32, i32, i64, P, P, P, i32]
        L.vmn_get_build_stats.argtypes = [P, ctypes.POINTER(BuildStats)]
        L.vmn_gather_probe.argtypes = [P, i64, i32, i32, P, i64, P, i64, i32, P]
        L.vmn_get_lowp_rows.argtypes = [P, i32, P]
        L.vmn_get_adjacency.argtypes = [P, P, i32]
        L.vmn_size.argtypes = [P]
        L.vmn_size.restype = i64
        L.vmn_medoid.argtypes = [P]
        L.vmn_medoid.restype = i64
        L.vmn_build_precision.argtypes = [P]
        L.vmn_build_precision.restype = i32
        L.vmn_search_stats.argtypes = [P, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.vmn_stream.argtypes = [P]
        L.vmn_stream.restype = P
        L.vmn_groundtruth.argtypes = [P, i64, i32, P, i64, i32, P, P, i32, i32]
        L.vmn_insert.argtypes = [P, P, i64, P, i32]
        L.vmn_delete_ids.argtypes = [P, P, i64, i32]
        L.vmn_delete_vectors.argtypes = [P, P, i64, P, i32]
        L.vmn_consolidate.argtypes = [P]
        L.vmn_get_deleted.argtypes = [P, P]
        L.vmn_ctx_create.argtypes = [P, ctypes.POINTER(P)]
        L.vmn_ctx_destroy.argtypes = [P]
        L.vmn_ctx_stream.argtypes = [P]
        L.vmn_ctx_stream.restype = P
        L.vmn_ctx_search.argtypes = [P, P, i64, i32, i32, i32, i32, P, P]
        L.vmn_ctx_capture.argtypes = [P, P, i64, i32, i32, i32, i32, P, P]
        L.vmn_ctx_replay.argtypes = [P]
        L.vmn_set_launch.argtypes = [P, i32, i32]
        L.vmn_set_tiles.argtypes = [P, i32]
        L.vmn_groundtruth_fallbacks.argtypes = []
        L.vmn_groundtruth_fallbacks.restype = ctypes.c_int64
        L.vmn_set_query_groups.argtypes = [P, i32]
        L.vmn_query_group_ids.argtypes = [P, ctypes.c_void_p, i64, i32, ctypes.c_void_p]
        L.vmn_default_server_params.argtypes = [ctypes.POINTER(ServerParams)]
        L.vmn_default_server_params.restype = None
        L.vmn_server_create.argtypes = [P, ctypes.POINTER(ServerParams), i32, ctypes.POINTER(P)]
        L.vmn_server_now.argtypes = [P]
        L.vmn_server_now.restype = ctypes.c_double
        L.vmn_server_submit.argtypes = [P, i32, P, i64, ctypes.c_double, ctypes.POINTER(i64)]
        L.vmn_server_poll.argtypes = [P, i64, P, P, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]
        L.vmn_server_pending.argtypes = [P]
        L.vmn_server_get_stats.argtypes = [P, ctypes.POINTER(ServerStats)]
        L.vmn_server_get_stats.restype = ctypes.c_int
        L.vmn_server_reset_stats.argtypes = [P]
        L.vmn_server_reset_stats.restype = ctypes.c_int
        L.vmn_server_pending.restype = i64
        L.vmn_server_destroy.argtypes = [P]
        L.vmn_recall.argtypes = [P, i32, P, P, i32, i64, i32]
        L.vmn_recall.restype = ctypes.c_double
        L.vmn_read_workload.argtypes = [ctypes.c_char_p, i32, P, P, P, i64, ctypes.POINTER(i64)]
        L.vmn_read_vectors.argtypes = [ctypes.c_char_p, i32, P, i64, ctypes.POINTER(i64),
                                       ctypes.POINTER(i32)]
        L.vmn_write_vectors.argtypes = [ctypes.c_char_p, i32, P, i64, i32]
        L.vmn_read_truthset.argtypes = [ctypes.c_char_p, P, P, i64, ctypes.POINTER(i64),
                                        ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.vmn_write_truthset.argtypes = [ctypes.c_char_p, P, P, i64, i32]
        L.vmn_save_bang_index.argtypes = [P, ctypes.c_char_p, ctypes.c_char_p]
        L.vmn_load_bang_index.argtypes = [ctypes.c_char_p, ctypes.c_char_p,
                                          ctypes.POINTER(Params), ctypes.POINTER(P)]
        L.vmn_insert_staged.argtypes = [P, P, i64, P, i32]
        L.vmn_staged_count.argtypes = [P]
        L.vmn_staged_count.restype = i64
        L.vmn_engine_delete.argtypes = [P, P, i64]
        L.vmn_engine_insert.argtypes = [P, P, i64, i32, P, i32]
        L.vmn_engine_run.argtypes = [P, P, P, P, i64, i32, i32, i32, ctypes.POINTER(EngineResult),
                                     P, P]
        L.vmn_last_error.argtypes = []
        L.vmn_last_error.restype = ctypes.c_char_p
        for name in ("vmn_create", "vmn_destroy", "vmn_build", "vmn_build_from", "vmn_load_graph",
                     "vmn_save", "vmn_load", "vmn_save_fresh_graph", "vmn_load_fresh_graph",
                     "vmn_save_parth_graph", "vmn_load_parth_graph", "vmn_get_vectors",
                     "vmn_get_params", "vmn_search",
                     "vmn_get_adjacency", "vmn_search_stats", "vmn_groundtruth", "vmn_insert",
                     "vmn_delete_ids", "vmn_delete_vectors", "vmn_consolidate",
                     "vmn_get_deleted", "vmn_search_ex", "vmn_get_lowp_rows", "vmn_ctx_create",
                     "vmn_ctx_destroy", "vmn_ctx_search", "vmn_ctx_capture", "vmn_ctx_replay",
                     "vmn_set_launch", "vmn_set_tiles", "vmn_set_query_groups",
                     "vmn_query_group_ids",
                     "vmn_server_create",
                     "vmn_server_submit",
                     "vmn_server_poll", "vmn_server_destroy", "vmn_read_workload",
                     "vmn_engine_delete", "vmn_engine_insert", "vmn_engine_run",
                     "vmn_read_vectors", "vmn_write_vectors", "vmn_read_truthset",
                     "vmn_write_truthset", "vmn_save_bang_index", "vmn_load_bang_index",
                     "vmn_search_vectors", "vmn_insert_staged", "vmn_bang_search"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise VamanaError(rc, lib().vmn_last_error().decode())


def default_params(mode=MODE_PARTH):
    p = Params()
    lib().vmn_default_params(ctypes.byref(p), mode)
    return p


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _torch_ready(x):
    """The library reads device tensors on its own streams: wait for the work
    torch has queued on its current stream (e.g. the kernel that produced the
    tensor) before handing the pointer over."""
    if _is_torch(x) and x.is_cuda:
        import torch
        torch.cuda.current_stream(x.device).synchronize()


def _device_tensor(x, what, dtype, device, min_numel=None, width=None):
    """Checks a tensor handed to the library as a raw device pointer: a
    contiguous CUDA tensor of `dtype` on `device`, holding at least
    `min_numel` elements (rows of `width` when given). Raises VamanaError
    (VMN_ERR_ARG) before any library call otherwise: a wrong pointer would be
    read or written out of bounds on the device."""
    import torch
    tdt = {np.float32: torch.float32, np.int32: torch.int32}[dtype]
    if not _is_torch(x):
        raise VamanaError(-1, f"{what} must be a torch tensor on cuda:{device}")
    if not x.is_cuda or x.device.index != device:
        raise VamanaError(-1, f"{what} is on {x.device}, the index is on cuda:{device}")
    if x.dtype != tdt:
        raise VamanaError(-1, f"{what} has dtype {x.dtype}, expected {tdt}")
    if not x.is_contiguous():
        raise VamanaError(-1, f"{what} must be contiguous")
    if width is not None and (x.dim() != 2 or int(x.shape[1]) != width):
        raise VamanaError(-1, f"{what} has shape {tuple(x.shape)}, expected [n, {width}]")
    if min_numel is not None and x.numel() < min_numel:
        raise VamanaError(-1, f"{what} holds {x.numel()} elements, needs {min_numel}")


def _arg(x, dtype):
    """(pointer, mem kind, keepalive) for a numpy array or a torch tensor."""
    if _is_torch(x):
        import torch
        tdt = {np.float32: torch.float32, np.int32: torch.int32}[dtype]
        assert x.dtype == tdt and x.is_contiguous(), "tensor must be contiguous " + str(tdt)
        _torch_ready(x)
        mem = MEM_DEVICE if x.is_cuda else MEM_HOST
        return ctypes.c_void_p(x.data_ptr()), mem, x
    a = np.ascontiguousarray(x, dtype=dtype)
    return a.ctypes.data_as(ctypes.c_void_p), MEM_HOST, a


class Vamana:
    """Vamana / FreshVamana index resident on one GPU.

    Vamana(vectors, mode="parth"|"fresh", **params) builds the graph (as the
    reference's constructor does), from the seeded random initial graph or from
    `init_adj` (an [n, w] -1-padded adjacency, e.g. the reference's random-graph
    file). Vamana.from_graph(vectors, adj, medoid) installs an existing graph
    instead; Vamana.load / load_fresh_graph / load_parth_graph read files.
    """

    def __init__(self, vectors=None, mode="parth", *, dim=None, R=None, L_build=None,
                 alphas=None, passes=None, seed=None, visited_cap=None, max_hops=None, P=None,
                 M=None, capacity=0, device=0, low_precision=0, init_adj=None, build_chunk=None,
                 _build=True):
        m = {"parth": MODE_PARTH, "fresh": MODE_FRESH}[mode] if isinstance(mode, str) else mode
        p = default_params(m)
        if vectors is not None:
            p.dim = int(vectors.shape[1])
        if dim is not None:
            p.dim = dim
        for name, v in (("R", R), ("L_build", L_build), ("passes", passes), ("seed", seed),
                        ("visited_cap", visited_cap), ("max_hops", max_hops), ("P", P), ("M", M),
                        ("build_chunk", build_chunk)):
            if v is not None:
                setattr(p, name, v)
        if alphas is not None:
            al = list(alphas)
            if passes is None:
                p.passes = len(al)
            for i in range(4):
                p.alpha[i] = al[min(i, len(al) - 1)]
        p.capacity = capacity
        p.device = device
        if isinstance(low_precision, str):
            low_precision = {"": 0, "bf16": STORE_BF16, "fp8": STORE_FP8,
                             "both": STORE_BF16 | STORE_FP8}[low_precision]
        p.low_precision = low_precision
        self.params = p
        self._h = ctypes.c_void_p()
        _check(lib().vmn_create(ctypes.byref(p), ctypes.byref(self._h)))
        if vectors is not None and _build:
            ptr, mem, keep = _arg(vectors, np.float32)
            if init_adj is None:
                _check(lib().vmn_build(self._h, ptr, int(vectors.shape[0]), mem))
            else:
                ap, am, ak = _arg(init_adj, np.int32)
                assert am == mem, "vectors and init_adj must both be host or both be device"
                _check(lib().vmn_build_from(self._h, ptr, int(vectors.shape[0]), ap,
                                            int(init_adj.shape[1]), mem))

    @classmethod
    def _wrap(cls, handle):
        self = cls.__new__(cls)
        self._h = handle
        self.params = Params()
        _check(lib().vmn_get_params(handle, ctypes.byref(self.params)))
        return self

    @classmethod
    def load(cls, path, device=0):
        """An index written by save() (vmn_load)."""
        h = ctypes.c_void_p()
        _check(lib().vmn_load(os.fsencode(path), device, ctypes.byref(h)))
        return cls._wrap(h)

    def save(self, path):
        """Parameters, vectors, graph, tombstones, entry point, FP8 scales (vmn_save)."""
        _check(lib().vmn_save(self._h, os.fsencode(path)))

    @classmethod
    def load_fresh_graph(cls, path, dim, R, build=False, medoid=-1, mode="fresh", **kw):
        """The reference's FreshVamana record file (n x {float vec[dim]; uint32
        degree; uint32 nbr[R]}): installed as the graph, or with build=True used
        as the initial graph of a build (vmn_load_fresh_graph)."""
        p = cls._params(mode, dim=dim, R=R, **kw)
        h = ctypes.c_void_p()
        _check(lib().vmn_load_fresh_graph(os.fsencode(path), ctypes.byref(p), int(bool(build)),
                                          int(medoid), ctypes.byref(h)))
        return cls._wrap(h)

    def save_fresh_graph(self, path):
        _check(lib().vmn_save_fresh_graph(self._h, os.fsencode(path)))

    @classmethod
    def load_parth_graph(cls, path, medoid=-1, mode="parth", **kw):
        """The parth engine's graph file (header n, dim, max_degree; vectors;
        -1-padded adjacency), installed with entry point `medoid` (-1: computed)."""
        p = cls._params(mode, **kw)
        h = ctypes.c_void_p()
        _check(lib().vmn_load_parth_graph(os.fsencode(path), ctypes.byref(p), int(medoid),
                                          ctypes.byref(h)))
        return cls._wrap(h)

    def save_parth_graph(self, path):
        _check(lib().vmn_save_parth_graph(self._h, os.fsencode(path)))

    @classmethod
    def load_bang_index(cls, index_path, metadata_path, mode="fresh", **kw):
        """BANG / DiskANN disk index (records + 32-byte metadata with the medoid)."""
        p = cls._params(mode, **kw)
        h = ctypes.c_void_p()
        _check(lib().vmn_load_bang_index(os.fsencode(index_path), os.fsencode(metadata_path),
                                         ctypes.byref(p), ctypes.byref(h)))
        return cls._wrap(h)

    def save_bang_index(self, index_path, metadata_path):
        _check(lib().vmn_save_bang_index(self._h, os.fsencode(index_path),
                                         os.fsencode(metadata_path)))

    @staticmethod
    def _params(mode, dim=None, R=None, L_build=None, alphas=None, capacity=0, device=0,
                low_precision=0):
        m = {"parth": MODE_PARTH, "fresh": MODE_FRESH}[mode] if isinstance(mode, str) else mode
        p = default_params(m)
        if dim is not None:
            p.dim = dim
        if R is not None:
            p.R = R
        if L_build is not None:
            p.L_build = L_build
        if alphas is not None:
            al = list(alphas)
            p.passes = len(al)
            for i in range(4):
                p.alpha[i] = al[min(i, len(al) - 1)]
        p.capacity, p.device = capacity, device
        if isinstance(low_precision, str):
            low_precision = {"": 0, "bf16": STORE_BF16, "fp8": STORE_FP8,
                             "both": STORE_BF16 | STORE_FP8}[low_precision]
        p.low_precision = low_precision
        return p

    @classmethod
    def from_graph(cls, vectors, adj, medoid=-1, mode="parth", R=None, **kw):
        R = R if R is not None else int(adj.shape[1])
        self = cls(vectors, mode=mode, R=R, _build=False, **kw)
        vp, vm, vk = _arg(vectors, np.float32)
        ap, am, ak = _arg(adj, np.int32)
        assert vm == am
        _check(lib().vmn_load_graph(self._h, vp, int(vectors.shape[0]), ap, int(adj.shape[1]),
                                    int(medoid), vm))
        return self

    def close(self):
        if self._h:
            lib().vmn_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def size(self):
        return int(lib().vmn_size(self._h))

    @property
    def medoid(self):
        return int(lib().vmn_medoid(self._h))

    @property
    def stream(self):
        return lib().vmn_stream(self._h)

    @property
    def build_precision(self):
        """Rows the build and insert searches read: "bf16" while every row is
        exact in BF16 (integers, |v| <= 256, squared norm <= 2^22), else "fp32".
        Both give the same graph."""
        return {PREC["bf16"]: "bf16", PREC["fp32"]: "fp32"}[lib().vmn_build_precision(self._h)]

    def searchPoints(self, queries, k=10, L=None, ids_out=None, dists_out=None, sync=True,
                     prec="fp32", rerank=0):
        """Search a batch; returns (ids [nq,k] int32, dists [nq,k] float32).

        prec: traversal distances "fp32" | "bf16" | "fp8"; rerank: exact FP32
        re-rank of that many list entries (k <= rerank <= 64, 0 = none).
        numpy in -> numpy out. CUDA tensor in -> CUDA tensors out (written on
        the index's stream; sync=False leaves the call asynchronous)."""
        L = k if L is None else L
        pc = PREC[prec] if isinstance(prec, str) else prec
        nq = int(queries.shape[0])
        qp, mem, keep = _arg(queries, np.float32)
        if mem == MEM_DEVICE:
            import torch
            dev = int(self.params.device)
            _device_tensor(queries, "queries", np.float32, dev, width=int(self.params.dim))
            ids = ids_out if ids_out is not None else torch.empty((nq, k), dtype=torch.int32,
                                                                  device=queries.device)
            dists = dists_out if dists_out is not None else torch.empty(
                (nq, k), dtype=torch.float32, device=queries.device)
            _device_tensor(ids, "ids_out", np.int32, dev, min_numel=nq * k)
            _device_tensor(dists, "dists_out", np.float32, dev, min_numel=nq * k)
            _check(lib().vmn_search_ex(self._h, qp, nq, k, L, pc, rerank,
                                       ctypes.c_void_p(ids.data_ptr()),
                                       ctypes.c_void_p(dists.data_ptr()), mem))
            if sync:
                self.synchronize()
            return ids, dists
        ids = np.empty((nq, k), dtype=np.int32)
        dists = np.empty((nq, k), dtype=np.float32)
        _check(lib().vmn_search_ex(self._h, qp, nq, k, L, pc, rerank,
                                   ids.ctypes.data_as(ctypes.c_void_p),
                                   dists.ctypes.data_as(ctypes.c_void_p), mem))
        return ids, dists

    search = searchPoints

    def bang_search(self, queries, k=10, L=40, entry=-1):
        """BANG's exact search schedule (bang_exact/src/parANN.cu:96-978) over
        this graph from `entry` (-1: the index's medoid) with list size L.
        Returns (ids [nq,k] int32, dists [nq,k] float32, stats [nq,2] int32 =
        per-query distance computations and rounds run). numpy in -> numpy out;
        CUDA tensor in -> CUDA tensors out (synchronised)."""
        nq = int(queries.shape[0])
        qp, mem, keep = _arg(queries, np.float32)
        if mem == MEM_DEVICE:
            import torch
            dev = int(self.params.device)
            _device_tensor(queries, "queries", np.float32, dev, width=int(self.params.dim))
            ids = torch.empty((nq, k), dtype=torch.int32, device=queries.device)
            dists = torch.empty((nq, k), dtype=torch.float32, device=queries.device)
            stats = torch.empty((nq, 2), dtype=torch.int32, device=queries.device)
            _check(lib().vmn_bang_search(self._h, qp, nq, k, L, entry,
                                         ctypes.c_void_p(ids.data_ptr()),
                                         ctypes.c_void_p(dists.data_ptr()),
                                         ctypes.c_void_p(stats.data_ptr()), mem))
            self.synchronize()
            return ids, dists, stats
        ids = np.empty((nq, k), dtype=np.int32)
        dists = np.empty((nq, k), dtype=np.float32)
        stats = np.empty((nq, 2), dtype=np.int32)
        _check(lib().vmn_bang_search(self._h, qp, nq, k, L, entry,
                                     ids.ctypes.data_as(ctypes.c_void_p),
                                     dists.ctypes.data_as(ctypes.c_void_p),
                                     stats.ctypes.data_as(ctypes.c_void_p), mem))
        return ids, dists, stats

    def searchPoints2(self, queries, k=10, L=None):
        """The k nearest rows of each query, [nq, k, dim] (vamana.cuh:127)."""
        L = k if L is None else L
        nq = int(queries.shape[0])
        qp, mem, keep = _arg(queries, np.float32)
        if mem == MEM_DEVICE:
            import torch
            out = torch.empty((nq, k, self.params.dim), dtype=torch.float32, device=queries.device)
            _check(lib().vmn_search_vectors(self._h, qp, nq, k, L,
                                            ctypes.c_void_p(out.data_ptr()), mem))
            self.synchronize()
            return out
        out = np.empty((nq, k, self.params.dim), np.float32)
        _check(lib().vmn_search_vectors(self._h, qp, nq, k, L,
                                        out.ctypes.data_as(ctypes.c_void_p), mem))
        return out

    def synchronize(self):
        import torch
        torch.cuda.ExternalStream(self.stream).synchronize()

    def insertPoints(self, vectors):
        """Batched insert; returns the new ids (numpy int32)."""
        n = int(vectors.shape[0])
        ids = np.empty(n, dtype=np.int32)
        vp, mem, keep = _arg(vectors, np.float32)
        if mem == MEM_DEVICE:
            import torch
            dids = torch.empty(n, dtype=torch.int32, device=vectors.device)
            _check(lib().vmn_insert(self._h, vp, n, ctypes.c_void_p(dids.data_ptr()), mem))
            return dids.cpu().numpy()
        _check(lib().vmn_insert(self._h, vp, n, ids.ctypes.data_as(ctypes.c_void_p), mem))
        return ids

    def deleteIds(self, ids):
        ip, mem, keep = _arg(ids, np.int32)
        _check(lib().vmn_delete_ids(self._h, ip, int(len(ids)), mem))

    def deletePoints(self, vectors):
        """Delete by exact vector match; returns the matched ids (-1: not found)."""
        n = int(vectors.shape[0])
        vp, mem, keep = _arg(vectors, np.float32)
        out = np.empty(n, dtype=np.int32)
        if mem == MEM_DEVICE:
            import torch
            d = torch.empty(n, dtype=torch.int32, device=vectors.device)
            _check(lib().vmn_delete_vectors(self._h, vp, n, ctypes.c_void_p(d.data_ptr()), mem))
            return d.cpu().numpy()
        _check(lib().vmn_delete_vectors(self._h, vp, n, out.ctypes.data_as(ctypes.c_void_p), mem))
        return out

    def consolidate(self):
        _check(lib().vmn_consolidate(self._h))

    patchGraph = consolidate

    @property
    def deleted(self):
        out = np.empty(self.size, dtype=np.uint8)
        _check(lib().vmn_get_deleted(self._h, out.ctypes.data_as(ctypes.c_void_p)))
        return out.astype(bool)

    @property
    def vectors(self):
        out = np.empty((self.size, self.params.dim), dtype=np.float32)
        _check(lib().vmn_get_vectors(self._h, out.ctypes.data_as(ctypes.c_void_p), MEM_HOST))
        return out

    @property
    def adjacency(self):
        out = np.empty((self.size, self.params.R), dtype=np.int32)
        _check(lib().vmn_get_adjacency(self._h, out.ctypes.data_as(ctypes.c_void_p), MEM_HOST))
        return out

    def insert_staged(self, vectors):
        """The reference's insertPoints (vamana_impl.cuh:22-33): held aside, merged
        into searches by exact distance, added to the graph by patchGraph();
        returns the ids they will take."""
        vp, mem, keep = _arg(vectors, np.float32)
        n = int(vectors.shape[0])
        ids = np.empty(max(n, 1), np.int32)
        if mem == MEM_DEVICE:
            import torch
            t = torch.empty(max(n, 1), dtype=torch.int32, device=vectors.device)
            _check(lib().vmn_insert_staged(self._h, vp, n, ctypes.c_void_p(t.data_ptr()), mem))
            self.synchronize()
            return t[:n].cpu().numpy()
        _check(lib().vmn_insert_staged(self._h, vp, n, ids.ctypes.data_as(ctypes.c_void_p), mem))
        return ids[:n]

    @property
    def staged_count(self):
        return int(lib().vmn_staged_count(self._h))

    def patchGraph(self):
        """vamana_impl.cuh:182-265: staged inserts join the graph (then
        tombstones are consolidated)."""
        self.consolidate()

    def engine_delete(self, ids):
        """parth engine delete (engine.cu:131-147): mark and push on the free list."""
        a = np.ascontiguousarray(ids, dtype=np.int32)
        _check(lib().vmn_engine_delete(self._h, a.ctypes.data_as(ctypes.c_void_p), len(a)))

    def engine_insert(self, vectors, link_L=128):
        """parth engine insert (engine.cu:161-239): free-list or new ids, each new
        node linked to its nearest node both ways; returns the ids."""
        vp, mem, keep = _arg(vectors, np.float32)
        n = int(vectors.shape[0])
        ids = np.empty(max(n, 1), np.int32)
        _check(lib().vmn_engine_insert(self._h, vp, n, int(link_L),
                                       ids.ctypes.data_as(ctypes.c_void_p), mem))
        return ids[:n]

    def run_workload(self, types, ids, vectors, batch_size, L, link_L=128, results=False):
        """VamanaEngine::process_workload (engine.cu:407-568): returns a dict with
        qps, recall, counts and times; with results=True also the found and
        ground-truth ids [searches, 10]."""
        t = np.ascontiguousarray(types, dtype=np.int32)
        i = np.ascontiguousarray(ids, dtype=np.int32)
        v = np.ascontiguousarray(vectors, dtype=np.float32)
        ns = int(np.count_nonzero((t != 0) & (t != 1)))
        found = np.full((max(ns, 1), 10), -1, np.int32) if results else None
        gt = np.full((max(ns, 1), 10), -1, np.int32) if results else None
        r = EngineResult()
        _check(lib().vmn_engine_run(self._h, t.ctypes.data_as(ctypes.c_void_p),
                                    i.ctypes.data_as(ctypes.c_void_p),
                                    v.ctypes.data_as(ctypes.c_void_p), len(t), int(batch_size),
                                    int(L), int(link_L), ctypes.byref(r),
                                    found.ctypes.data_as(ctypes.c_void_p) if results else None,
                                    gt.ctypes.data_as(ctypes.c_void_p) if results else None))
        out = {f: getattr(r, f) for f, _ in EngineResult._fields_}
        if results:
            out["found"], out["gt"] = found[:ns], gt[:ns]
        return out

    def set_launch(self, waves_per_cu=32, list_cap=0):
        _check(lib().vmn_set_launch(self._h, waves_per_cu, list_cap))

    def set_tiles(self, tiles=2):
        """Tile shape of the BF16 / FP8 searches: MFMA tiles of 16 candidates
        gathered per round (2 or 4). Changes the speed, never a result."""
        _check(lib().vmn_set_tiles(self._h, tiles))

    def set_query_groups(self, groups=8):
        """Query groups of large batches (one per XCD; 0 or 1: one queue).
        Changes the order queries are searched in, never a result."""
        _check(lib().vmn_set_query_groups(self._h, groups))

    def query_group_ids(self, queries):
        """Group of each query (uint8 numpy array), as a large batch uses."""
        nq = int(queries.shape[0])
        qp, mem, keep = _arg(queries, np.float32)
        if mem == MEM_DEVICE:
            import torch
            out = torch.empty(nq, dtype=torch.uint8, device=queries.device)
            _check(lib().vmn_query_group_ids(self._h, qp, nq, mem, ctypes.c_void_p(out.data_ptr())))
            return out.cpu().numpy()
        out = np.empty(nq, dtype=np.uint8)
        _check(lib().vmn_query_group_ids(self._h, qp, nq, mem, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def search_kernel_name(self, k, L, prec="fp32", list_cap=0, tiles=2):
        """Name of the search kernel instance a search with these arguments
        launches, as a profiler prints it (template arguments: padded
        dimension, beam-list capacity, mode, precision, gather tiles)."""
        d = (int(self.params.dim) + 31) // 32 * 32
        dp = 256 if d > 192 else d
        need = max(int(L), int(list_cap))
        cap = next(c for c in (64, 128, 256, 512) if need <= c)
        pc = PREC[prec] if isinstance(prec, str) else int(prec)
        mode = int(self.params.mode)
        ng = 4 if (tiles == 4 and pc != 0 and mode == MODE_FRESH and 64 <= dp <= 128
                   and need <= 256) else 2
        return f"void vmn::search_kernel<{dp}, {cap}, {mode}, {pc}, {ng}>(vmn::SearchArgs)"

    def lowp_rows(self, prec):
        """Rows as the BF16 / FP8 traversal sees them, decoded to float32."""
        out = np.empty((self.size, self.params.dim), dtype=np.float32)
        pc = PREC[prec] if isinstance(prec, str) else prec
        _check(lib().vmn_get_lowp_rows(self._h, pc, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def build_stats(self):
        """Work counted by the last build's kernels and its algorithmic bytes
        (vmn_build_stats)."""
        st = BuildStats()
        _check(lib().vmn_get_build_stats(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in BuildStats._fields_}

    def search_stats(self):
        e, d = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().vmn_search_stats(self._h, ctypes.byref(e), ctypes.byref(d)))
        return int(e.value), int(d.value)


VEC_FORMATS = {"fvecs": (0, np.float32), "ivecs": (1, np.int32), "bvecs": (2, np.uint8),
               "bin": (3, np.float32)}


def read_vectors(path, fmt):
    """fvecs / ivecs / bvecs / bin (u32 n, dim header + floats) -> [n, dim]."""
    code, dt = VEC_FORMATS[fmt]
    n, d = ctypes.c_int64(), ctypes.c_int32()
    _check(lib().vmn_read_vectors(os.fsencode(path), code, None, 0, ctypes.byref(n),
                                  ctypes.byref(d)))
    out = np.empty((max(n.value, 1), d.value), dt)
    _check(lib().vmn_read_vectors(os.fsencode(path), code, out.ctypes.data_as(ctypes.c_void_p),
                                  n.value, ctypes.byref(n), ctypes.byref(d)))
    return out[:n.value]


def write_vectors(path, vecs, fmt):
    code, dt = VEC_FORMATS[fmt]
    a = np.ascontiguousarray(vecs, dtype=dt)
    _check(lib().vmn_write_vectors(os.fsencode(path), code, a.ctypes.data_as(ctypes.c_void_p),
                                   a.shape[0], a.shape[1]))


def read_truthset(path):
    """DiskANN ground-truth file -> (ids [n, k] int32, dists [n, k] float32 or None)."""
    n, k, hd = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
    _check(lib().vmn_read_truthset(os.fsencode(path), None, None, 0, ctypes.byref(n),
                                   ctypes.byref(k), ctypes.byref(hd)))
    ids = np.empty((max(n.value, 1), k.value), np.int32)
    dists = np.empty((max(n.value, 1), k.value), np.float32) if hd.value else None
    _check(lib().vmn_read_truthset(os.fsencode(path), ids.ctypes.data_as(ctypes.c_void_p),
                                   dists.ctypes.data_as(ctypes.c_void_p) if hd.value else None,
                                   n.value, ctypes.byref(n), ctypes.byref(k), ctypes.byref(hd)))
    return ids[:n.value], (dists[:n.value] if hd.value else None)


def write_truthset(path, ids, dists=None):
    i = np.ascontiguousarray(ids, dtype=np.int32)
    d = None if dists is None else np.ascontiguousarray(dists, dtype=np.float32)
    _check(lib().vmn_write_truthset(os.fsencode(path), i.ctypes.data_as(ctypes.c_void_p),
                                    None if d is None else d.ctypes.data_as(ctypes.c_void_p),
                                    i.shape[0], i.shape[1]))


def read_workload(path, dim):
    """The parth engine's workload file (records {int32 type; int32 id; float
    vec[dim]}, parth/src/common.h:63-67) -> (types, ids, vecs [n, dim])."""
    n = ctypes.c_int64()
    _check(lib().vmn_read_workload(os.fsencode(path), int(dim), None, None, None, 0,
                                   ctypes.byref(n)))
    t = np.empty(max(n.value, 1), np.int32)
    i = np.empty(max(n.value, 1), np.int32)
    v = np.empty((max(n.value, 1), dim), np.float32)
    _check(lib().vmn_read_workload(os.fsencode(path), int(dim), t.ctypes.data_as(ctypes.c_void_p),
                                   i.ctypes.data_as(ctypes.c_void_p),
                                   v.ctypes.data_as(ctypes.c_void_p), n.value, ctypes.byref(n)))
    return t[:n.value], i[:n.value], v[:n.value]


def groundtruth(base, queries, k, device=0):
    """Exact FP64 k-NN on the GPU: (ids [nq,k] int32, dists [nq,k] float64)."""
    bp, bm, bk = _arg(base, np.float32)
    qp, qm, qk = _arg(queries, np.float32)
    assert bm == qm
    nq = int(queries.shape[0])
    if bm == MEM_DEVICE:
        import torch
        ids = torch.empty((nq, k), dtype=torch.int32, device=queries.device)
        d = torch.empty((nq, k), dtype=torch.float64, device=queries.device)
        _check(lib().vmn_groundtruth(bp, int(base.shape[0]), int(base.shape[1]), qp, nq, k,
                                     ctypes.c_void_p(ids.data_ptr()),
                                     ctypes.c_void_p(d.data_ptr()), bm, device))
        return ids, d
    ids = np.empty((nq, k), dtype=np.int32)
    d = np.empty((nq, k), dtype=np.float64)
    _check(lib().vmn_groundtruth(bp, int(base.shape[0]), int(base.shape[1]), qp, nq, k,
                                 ids.ctypes.data_as(ctypes.c_void_p),
                                 d.ctypes.data_as(ctypes.c_void_p), bm, device))
    return ids, d


def groundtruth_fallbacks():
    """Queries of this thread's last groundtruth() call that the MFMA filter's
    error bound left to the direct FP64 kernel (results are the same)."""
    return int(lib().vmn_groundtruth_fallbacks())


def recall_at_k(found, gt_ids, k=10, gt_dists=None):
    """Mean |found[:, :k] ∩ gt[:, :k]| / k (engine.cu:506-528 counts the same);
    with gt_dists the ground-truth set also takes the ids tied with the k-th
    (calculate_recall, bang_exact/src/main.cu:87-137). Computed by vmn_recall."""
    if _is_torch(found):
        found = found.cpu().numpy()
    if _is_torch(gt_ids):
        gt_ids = gt_ids.cpu().numpy()
    f = np.ascontiguousarray(found, dtype=np.int32)
    g = np.ascontiguousarray(gt_ids, dtype=np.int32)
    d = None if gt_dists is None else np.ascontiguousarray(
        gt_dists.cpu().numpy() if _is_torch(gt_dists) else gt_dists, dtype=np.float32)
    r = lib().vmn_recall(f.ctypes.data_as(ctypes.c_void_p), f.shape[1],
                         g.ctypes.data_as(ctypes.c_void_p),
                         None if d is None else d.ctypes.data_as(ctypes.c_void_p), g.shape[1],
                         f.shape[0], int(k))
    if r < 0:
        raise VamanaError(-1, lib().vmn_last_error().decode())
    return r


class SearchContext:
    """A search context (own stream + scratch); searches on different contexts
    run concurrently. capture() records one search as a HIP graph that
    replay() launches again (device tensors only)."""

    def __init__(self, index):
        self.index = index
        self._h = ctypes.c_void_p()
        _check(lib().vmn_ctx_create(index._h, ctypes.byref(self._h)))

    @property
    def stream(self):
        return lib().vmn_ctx_stream(self._h)

    def _args(self, q, k, L, prec, rerank, ids, dists):
        dev, dim = int(self.index.params.device), int(self.index.params.dim)
        _device_tensor(q, "queries", np.float32, dev, width=dim)
        nq = int(q.shape[0])
        _device_tensor(ids, "ids", np.int32, dev, min_numel=nq * int(k))
        _device_tensor(dists, "dists", np.float32, dev, min_numel=nq * int(k))
        _torch_ready(q)
        pc = PREC[prec] if isinstance(prec, str) else prec
        return (self._h, ctypes.c_void_p(q.data_ptr()), int(q.shape[0]), k, L, pc, rerank,
                ctypes.c_void_p(ids.data_ptr()), ctypes.c_void_p(dists.data_ptr()))

    def search(self, q, ids, dists, k=10, L=64, prec="fp32", rerank=0):
        _check(lib().vmn_ctx_search(*self._args(q, k, L, prec, rerank, ids, dists)))

    def capture(self, q, ids, dists, k=10, L=64, prec="fp32", rerank=0):
        _check(lib().vmn_ctx_capture(*self._args(q, k, L, prec, rerank, ids, dists)))

    def replay(self):
        _check(lib().vmn_ctx_replay(self._h))

    def close(self):
        if self._h:
            lib().vmn_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request server over an index (vmn_server_*).

    submit() queues a request with its arrival time (server clock, now());
    poll() returns None while pending, else a dict with the results and the
    arrival / completion times. While the server runs, use the index only
    through it."""

    def __init__(self, index, k=10, L=64, prec="fp32", rerank=0, max_batch=4096, n_slots=3,
                 max_queued_bytes=1 << 30, consolidate_every=0):
        p = ServerParams()
        lib().vmn_default_server_params(ctypes.byref(p))
        p.k, p.L = k, L
        p.prec = PREC[prec] if isinstance(prec, str) else prec
        p.k_rerank, p.max_batch, p.n_slots = rerank, max_batch, n_slots
        p.max_queued_bytes, p.consolidate_every = max_queued_bytes, consolidate_every
        p.device = index.params.device
        self.index, self.k, self.dim = index, k, int(index.params.dim)
        self._h = ctypes.c_void_p()
        rc = lib().vmn_server_create(index._h, ctypes.byref(p), self.dim, ctypes.byref(self._h))
        if rc != 0:
            raise VamanaError(rc, "vmn_server_create failed")
        self._meta = {}

    def now(self):
        return float(lib().vmn_server_now(self._h))

    def submit(self, kind, data, t_arrival=None):
        t = self.now() if t_arrival is None else float(t_arrival)
        dt = np.int32 if kind == REQ_DELETE else np.float32
        a = np.ascontiguousarray(data, dtype=dt)
        n = int(a.shape[0])
        tk = ctypes.c_int64()
        rc = lib().vmn_server_submit(self._h, kind, a.ctypes.data_as(ctypes.c_void_p), n, t,
                                     ctypes.byref(tk))
        if rc != 0:
            raise VamanaError(rc, "submit rejected")
        self._meta[tk.value] = (kind, n)
        return tk.value

    def poll(self, ticket):
        kind, n = self._meta[ticket]
        ids = np.empty((n, self.k) if kind == REQ_SEARCH else (n,), dtype=np.int32)
        d = np.empty((n, self.k), dtype=np.float32) if kind == REQ_SEARCH else None
        ta, td = ctypes.c_double(), ctypes.c_double()
        rc = lib().vmn_server_poll(self._h, ticket, ids.ctypes.data_as(ctypes.c_void_p),
                                   None if d is None else d.ctypes.data_as(ctypes.c_void_p),
                                   ctypes.byref(ta), ctypes.byref(td))
        if rc == 0:
            return None
        del self._meta[ticket]
        if rc < 0:
            raise VamanaError(rc, f"request {ticket} failed")
        return {"kind": kind, "ids": ids if kind != REQ_DELETE else None, "dists": d,
                "t_arrival": ta.value, "t_done": td.value}

    def pending(self):
        return int(lib().vmn_server_pending(self._h))

    def stats(self):
        """Engine counters (vmn_server_stats) as a dict."""
        st = ServerStats()
        _check(lib().vmn_server_get_stats(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in ServerStats._fields_}

    def reset_stats(self):
        """Zero the engine counters (after a warm-up)."""
        _check(lib().vmn_server_reset_stats(self._h))
    def close(self):
        if self._h:
            lib().vmn_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
