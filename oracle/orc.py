"""ctypes wrapper over the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the package smoke check and by
bench.py's cpu_baseline leg, always as the checker or the CPU baseline, never
as the thing measured or shipped.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liborc.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32, i64, u64, f32 = ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float
        _lib.orc_l2sq.argtypes = [P, P, i32, i32]
        _lib.orc_l2sq.restype = f32
        _lib.orc_random_init.argtypes = [i64, i32, i32, u64, P]
        _lib.orc_medoid.argtypes = [P, i64, i32]
        _lib.orc_medoid.restype = i64
        _lib.orc_parth_search.argtypes = [P, i64, i32, P, i32, P, P, i64, i32, i32, i32, i32,
                                          i32, i32, P, P, P, P]
        _lib.orc_parth_prune.argtypes = [P, i64, i32, P, i32, i32, f32, P, P, i32, i32]
        _lib.orc_parth_build.argtypes = [P, i64, i32, i32, i32, i32, P, i32, u64, i32, i32,
                                         i32, P, P, i32]
        _lib.orc_parth_build.restype = i64
        _lib.orc_bruteforce_fp64.argtypes = [P, i64, i32, P, P, i64, i32, P, P]
        _lib.orc_bf_hash1.argtypes = [ctypes.c_uint32]
        _lib.orc_bf_hash1.restype = ctypes.c_uint32
        _lib.orc_bf_hash2.argtypes = [ctypes.c_uint32]
        _lib.orc_bf_hash2.restype = ctypes.c_uint32
        _lib.orc_fresh_search.argtypes = [P, i64, i32, P, i32, P, P, i64, i32, i64, i32, i32,
                                          i32, i32, i32, i32, P, P, P, P, P, P, P]
        _lib.orc_fresh_insert.argtypes = [P, i64, i64, i32, P, i32, P, i32, i32, f32, i32, i32,
                                          i32, i32]
        _lib.orc_consolidate.argtypes = [P, i64, i32, P, i32, P, f32, i32, i32]
        _lib.orc_consolidate.restype = i64
        _lib.orc_init_rows.argtypes = [P, i64, i32, i32, i32, P]
        _lib.orc_engine_run.argtypes = [P, P, P, i64, i32, i32, P, P, P, P, P, P, i64, i32, i32,
                                        i32, i32, i32, i32, i32, P, P, P, P]
        _lib.orc_engine_run.restype = i64
        _lib.orc_medoid_live.argtypes = [P, i64, i32, P]
        _lib.orc_medoid_live.restype = i64
        _lib.orc_fresh_out_prune.argtypes = [P, i32, P, i32, i32, f32, i64, i64, P, P, P, i32,
                                             i32, P, P, P]
        _lib.orc_fresh_reverse.argtypes = [P, i64, i32, P, i32, i32, f32, i64, i64, P, P, i32,
                                           i32]
        _lib.orc_fresh_build.argtypes = [P, i64, i32, i32, i32, P, i32, u64, i32, i32, i32, i32,
                                         P, P, i32]
        _lib.orc_fresh_build.restype = i64
        _lib.orc_set_mutation.argtypes = [i32]
        _lib.orc_bang_hash.argtypes = [ctypes.c_uint32]
        _lib.orc_bang_hash.restype = ctypes.c_uint32
        _lib.orc_bang_l2sq.argtypes = [P, P, i32]
        _lib.orc_bang_l2sq.restype = f32
        _lib.orc_bang_search.argtypes = [P, i64, i32, P, i32, P, i64, P, i64, i32, i32, P, P, P]
        _lib.orc_num_threads.restype = i32
        _lib.orc_set_num_threads.argtypes = [i32]
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def l2sq(a, b, dist_mode=0):
    a, b = _f32(a), _f32(b)
    return lib().orc_l2sq(_p(a), _p(b), a.shape[0], dist_mode)


def random_init(n, R, row_width, seed=1234):
    adj = np.empty((n, row_width), dtype=np.int32)
    lib().orc_random_init(n, R, row_width, seed, _p(adj))
    return adj


def medoid(vecs):
    vecs = _f32(vecs)
    return int(lib().orc_medoid(_p(vecs), vecs.shape[0], vecs.shape[1]))


def parth_search(vecs, adj, queries, start, L, out_n=10, visited_cap=128, max_hops=200,
                 dist_mode=0, deleted=None):
    """Returns (ids [nq,out_n] int32, dists [nq,out_n] f32, hops [nq], ndist [nq])."""
    vecs, queries = _f32(vecs), _f32(queries)
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    nq = queries.shape[0]
    ids = np.empty((nq, out_n), dtype=np.int32)
    dists = np.empty((nq, out_n), dtype=np.float32)
    hops = np.empty(nq, dtype=np.int32)
    nd = np.empty(nq, dtype=np.int32)
    dm = None if deleted is None else np.ascontiguousarray(deleted, dtype=np.uint8)
    lib().orc_parth_search(_p(vecs), vecs.shape[0], vecs.shape[1], _p(adj), adj.shape[1],
                           _p(dm), _p(queries), nq, int(start), L, visited_cap, max_hops,
                           dist_mode, out_n, _p(ids), _p(dists), _p(hops), _p(nd))
    return ids, dists, hops, nd


def parth_prune(vecs, adj, R, alpha, cand_ids, cand_dists, dist_mode=0):
    """Prunes every row of `adj` in place from the [n, L] candidate lists."""
    vecs = _f32(vecs)
    assert adj.dtype == np.int32 and adj.flags.c_contiguous
    cand_ids = np.ascontiguousarray(cand_ids, dtype=np.int32)
    cand_dists = _f32(cand_dists)
    lib().orc_parth_prune(_p(vecs), vecs.shape[0], vecs.shape[1], _p(adj), adj.shape[1], R,
                          float(alpha), _p(cand_ids), _p(cand_dists), cand_ids.shape[1],
                          dist_mode)
    return adj


def parth_build(vecs, R, L, alphas=(1.2, 1.2), row_width=None, seed=1234, visited_cap=128,
                max_hops=200, dist_mode=0, init=None):
    """Returns (adj [n,row_width] int32, medoid). `init` ([n, w] int32, -1
    padded) replaces the seeded random initial graph."""
    vecs = _f32(vecs)
    n, d = vecs.shape
    rw = R if row_width is None else row_width
    adj = np.empty((n, rw), dtype=np.int32)
    al = np.asarray(alphas, dtype=np.float32)
    ini = None if init is None else np.ascontiguousarray(init, dtype=np.int32)
    med = lib().orc_parth_build(_p(vecs), n, d, R, rw, L, _p(al), len(al), seed, visited_cap,
                                max_hops, dist_mode, _p(adj), _p(ini),
                                0 if ini is None else ini.shape[1])
    return adj, int(med)


def bf_hash(x):
    return int(lib().orc_bf_hash1(x)), int(lib().orc_bf_hash2(x))


def fresh_search(vecs, adj, queries, start, L, out_n=10, P=600, max_hops=1 << 20, dist_mode=0,
                 deleted=None, queries_are_nodes=False, want_visited=False, node_base=0):
    """G2 greedy search. Returns (ids, dists, hops, ndist[, vis_ids, vis_d, vis_cnt])."""
    vecs, queries = _f32(vecs), _f32(queries)
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    nq = queries.shape[0]
    ids = np.empty((nq, out_n), dtype=np.int32)
    dists = np.empty((nq, out_n), dtype=np.float32)
    hops = np.empty(nq, dtype=np.int32)
    nd = np.empty(nq, dtype=np.int32)
    vi = np.full((nq, P), -1, dtype=np.int32) if want_visited else None
    vd = np.zeros((nq, P), dtype=np.float32) if want_visited else None
    vc = np.zeros(nq, dtype=np.int32) if want_visited else None
    dm = None if deleted is None else np.ascontiguousarray(deleted, dtype=np.uint8)
    lib().orc_fresh_search(_p(vecs), vecs.shape[0], vecs.shape[1], _p(adj), adj.shape[1], _p(dm),
                           _p(queries), nq, int(queries_are_nodes), int(node_base), int(start),
                           L, max_hops, P,
                           dist_mode, out_n, _p(ids), _p(dists), _p(vi), _p(vd), _p(vc),
                           _p(hops), _p(nd))
    if want_visited:
        return ids, dists, hops, nd, vi, vd, vc
    return ids, dists, hops, nd


def fresh_build(vecs, R=64, L=128, alphas=(1.5,), seed=1234, P=600, M=500, max_hops=1 << 20,
                dist_mode=0, init=None):
    """Returns (adj [n,R] int32, medoid). `init` ([n, w] int32, -1 padded)
    replaces the seeded random initial graph."""
    vecs = _f32(vecs)
    n, d = vecs.shape
    adj = np.empty((n, R), dtype=np.int32)
    al = np.asarray(alphas, dtype=np.float32)
    ini = None if init is None else np.ascontiguousarray(init, dtype=np.int32)
    med = lib().orc_fresh_build(_p(vecs), n, d, R, L, _p(al), len(al), seed, P, M, max_hops,
                                dist_mode, _p(adj), _p(ini), 0 if ini is None else ini.shape[1])
    return adj, int(med)


def fresh_out_prune(vecs, adj, R, alpha, vis_ids, vis_d, vis_cnt, base=0, deleted=None,
                    dist_mode=0):
    """Out-prune of nodes base .. base + len(vis_cnt) - 1 (rows of `adj`
    rewritten in place) from their visited sets [cnt, P]. Returns the reverse
    pairs (pair_dst [cnt, R], -1 unused; pair_d [cnt, R])."""
    vecs = _f32(vecs)
    assert adj.dtype == np.int32 and adj.flags.c_contiguous
    vi = np.ascontiguousarray(vis_ids, dtype=np.int32)
    vd = _f32(vis_d)
    vc = np.ascontiguousarray(vis_cnt, dtype=np.int32)
    cnt, P = vi.shape
    pdst = np.full((cnt, R), -1, np.int32)
    pd = np.zeros((cnt, R), np.float32)
    dm = None if deleted is None else np.ascontiguousarray(deleted, dtype=np.uint8)
    lib().orc_fresh_out_prune(_p(vecs), vecs.shape[1], _p(adj), adj.shape[1], R, float(alpha),
                              base, cnt, _p(vi), _p(vd), _p(vc), P, dist_mode, _p(pdst), _p(pd),
                              _p(dm))
    return pdst, pd


def fresh_reverse(vecs, adj, R, alpha, pair_dst, pair_d, base=0, M=500, dist_mode=0):
    """Reverse-edge prune of every node that received a pair (rows of `adj`
    rewritten in place); pair row i belongs to source node base + i."""
    vecs = _f32(vecs)
    assert adj.dtype == np.int32 and adj.flags.c_contiguous
    pdst = np.ascontiguousarray(pair_dst, dtype=np.int32)
    pd = _f32(pair_d)
    lib().orc_fresh_reverse(_p(vecs), vecs.shape[0], vecs.shape[1], _p(adj), adj.shape[1], R,
                            float(alpha), base, pdst.shape[0], _p(pdst), _p(pd), M, dist_mode)
    return adj


def set_mutation(m):
    """Plants mistake `m` in the fresh-mode restatement (0 = none); used only
    by the tests that show the fresh-mode checks can fail."""
    lib().orc_set_mutation(int(m))


def fresh_insert(vecs, n_old, adj, start, L, alpha, deleted=None, P=600, M=500, max_hops=1 << 20,
                 dist_mode=0):
    """vecs [N,d] holds old and new nodes; adj [N,R] with -1 rows for the new
    nodes is updated in place."""
    vecs = _f32(vecs)
    assert adj.dtype == np.int32 and adj.flags.c_contiguous
    N, d = vecs.shape
    dm = None if deleted is None else np.ascontiguousarray(deleted, dtype=np.uint8)
    lib().orc_fresh_insert(_p(vecs), n_old, N - n_old, d, _p(adj), adj.shape[1], _p(dm),
                           int(start), L, float(alpha), P, M, max_hops, dist_mode)
    return adj


def consolidate(vecs, adj, deleted, alpha, M=500, dist_mode=0):
    vecs = _f32(vecs)
    assert adj.dtype == np.int32 and adj.flags.c_contiguous
    dm = np.ascontiguousarray(deleted, dtype=np.uint8)
    return int(lib().orc_consolidate(_p(vecs), vecs.shape[0], vecs.shape[1], _p(adj),
                                     adj.shape[1], _p(dm), float(alpha), M, dist_mode))


def medoid_live(vecs, deleted):
    vecs = _f32(vecs)
    dm = np.ascontiguousarray(deleted, dtype=np.uint8)
    return int(lib().orc_medoid_live(_p(vecs), vecs.shape[0], vecs.shape[1], _p(dm)))


def bruteforce_fp64(vecs, queries, k, deleted=None):
    vecs, queries = _f32(vecs), _f32(queries)
    nq = queries.shape[0]
    ids = np.empty((nq, k), dtype=np.int32)
    dists = np.empty((nq, k), dtype=np.float64)
    dm = None if deleted is None else np.ascontiguousarray(deleted, dtype=np.uint8)
    lib().orc_bruteforce_fp64(_p(vecs), vecs.shape[0], vecs.shape[1], _p(dm), _p(queries), nq,
                              k, _p(ids), _p(dists))
    return ids, dists


def bang_hash(x):
    return int(lib().orc_bang_hash(x & 0xFFFFFFFF))


def bang_l2sq(q, x):
    q, x = _f32(q), _f32(x)
    return float(lib().orc_bang_l2sq(_p(q), _p(x), q.shape[0]))


def bang_search(vecs, adj, queries, entry, L, k=10, deleted=None):
    """BANG exact search (oracle.cpp orc_bang_search): ids [nq x k], dists
    [nq x k] and per-query (distance computations, rounds run) [nq x 2]."""
    vecs, queries = _f32(vecs), _f32(queries)
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    nq = queries.shape[0]
    ids = np.empty((nq, k), dtype=np.int32)
    dists = np.empty((nq, k), dtype=np.float32)
    stats = np.empty((nq, 2), dtype=np.int32)
    dm = None if deleted is None else np.ascontiguousarray(deleted, dtype=np.uint8)
    lib().orc_bang_search(_p(vecs), vecs.shape[0], vecs.shape[1], _p(adj), adj.shape[1], _p(dm),
                          int(entry), _p(queries), nq, L, k, _p(ids), _p(dists), _p(stats))
    return ids, dists, stats


def engine_run(vecs, adj, n, cap, types, ids, opvecs, batch, L, link_L=128, start=0,
               visited_cap=128, max_hops=200, dist_mode=0, stats=None):
    """parth engine workload replay (engine.cu:407-568). vecs [n, d] and adj
    [n, W] are the initial graph; returns (found, gt, vecs [cap, d], adj [cap, W],
    deleted [cap], c0, assigned) with found / gt [searches, 10]. A dict `stats`
    receives search_s, the wall time of the search loops."""
    vecs = _f32(vecs)
    d = vecs.shape[1]
    W = adj.shape[1]
    V = np.zeros((cap, d), np.float32)
    V[:n] = vecs[:n]
    A = np.full((cap, W), -1, np.int32)
    A[:n] = adj[:n]
    dm = np.zeros(cap, np.uint8)
    fl = np.zeros(cap, np.int32)
    types = np.ascontiguousarray(types, dtype=np.int32)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    opvecs = _f32(opvecs)
    n_ops = len(types)
    nsr = int((types != 0).sum() - (types == 1).sum())
    found = np.full((max(nsr, 1), 10), -1, np.int32)
    gt = np.full((max(nsr, 1), 10), -1, np.int32)
    assigned = np.full(max(int((types == 1).sum()), 1), -1, np.int32)
    n_io = ctypes.c_int64(n)
    c1 = ctypes.c_int64(0)
    ts = ctypes.c_double(0.0)
    r = lib().orc_engine_run(_p(V), _p(A), _p(dm), cap, d, W, ctypes.byref(n_io), ctypes.byref(c1),
                             _p(fl), _p(types), _p(ids), _p(opvecs), n_ops, batch, L, link_L,
                             visited_cap, max_hops, dist_mode, start, _p(found), _p(gt),
                             _p(assigned), ctypes.byref(ts))
    if stats is not None:
        stats["search_s"] = ts.value
    if r < 0:
        raise RuntimeError("engine replay exceeds capacity")
    return (found[:r], gt[:r], V, A, dm, int(n_io.value),
            assigned[:int((types == 1).sum())])


def num_threads():
    return int(lib().orc_num_threads())


def set_num_threads(t):
    lib().orc_set_num_threads(int(t))
