"""MI355X-native Vamana / FreshVamana index (gfx950 HIP kernels behind a C ABI).

Import with importlib (the directory name is not an identifier):

    vmn = importlib.import_module("cs5013-project_amd")
    index = vmn.Vamana(base, mode="parth", R=32, L_build=64)
    ids, dists = index.searchPoints(queries, k=10, L=40)
"""
from .vamana import (MODE_FRESH, MODE_PARTH, REQ_DELETE, REQ_INSERT, REQ_SEARCH, Params,
                     SearchContext, Server, Vamana, VamanaError, default_params, groundtruth,
                     groundtruth_fallbacks, lib,
                     read_truthset, read_vectors, read_workload, recall_at_k,
                     write_truthset, write_vectors)
from .engine import VamanaEngine
from . import datasets, tuning

__all__ = ["Vamana", "VamanaEngine", "VamanaError", "Params", "SearchContext", "Server", "tuning",
           "read_workload", "read_vectors", "write_vectors", "read_truthset", "write_truthset",
           "default_params", "groundtruth", "groundtruth_fallbacks", "lib",
           "recall_at_k", "datasets", "MODE_PARTH", "MODE_FRESH", "REQ_SEARCH", "REQ_INSERT",
           "REQ_DELETE"]
