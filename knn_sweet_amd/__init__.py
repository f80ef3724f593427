"""knn_sweet_amd -- MI355X-native exact kNN join (Python binding of libknnj.so).

The library is the C ABI declared in include/knnj.h; this module mirrors it
with ctypes.  It has no CPU fallback: if libknnj.so is missing or no GPU is
visible, the calls raise.

    import knn_sweet_amd as kj
    ctx = kj.Context()
    idx, dist, stats = ctx.join(R, Q, k)     # numpy (host) or torch CUDA tensors

``R`` and ``Q`` are 2-D float32 arrays (row-major).  For torch tensors on the
GPU the outputs are torch tensors on the same device; for numpy inputs they
are numpy arrays.  ``idx[q]`` lists the k reference rows nearest to ``Q[q]``
in ascending (squared distance, row index) order; ``dist`` is the Euclidean
distance (or the squared one with ``out_squared=True``).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libknnj.so")

KNNJ_OK, KNNJ_EINVAL, KNNJ_ENOMEM, KNNJ_EDEVICE, KNNJ_EUNSUPPORTED, KNNJ_EOVERFLOW = range(6)
PREC_AUTO, PREC_FP32, PREC_BF16, PREC_FP8, PREC_BF16X1 = range(5)
KERNEL_AUTO, KERNEL_VALU, KERNEL_MFMA = range(3)
_STATUS_NAMES = {0: "OK", 1: "EINVAL", 2: "ENOMEM", 3: "EDEVICE", 4: "EUNSUPPORTED", 5: "EOVERFLOW"}


class KnnjError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"knnj {_STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Matrix(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("n", ctypes.c_int64), ("d", ctypes.c_int32),
                ("ld", ctypes.c_int64), ("on_device", ctypes.c_int32)]


class Opts(ctypes.Structure):
    _fields_ = [("m_q", ctypes.c_int32), ("m_r", ctypes.c_int32), ("trials", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("seed", ctypes.c_uint64),
                ("out_squared", ctypes.c_int32), ("out_aos", ctypes.c_int32),
                ("brute_force", ctypes.c_int32), ("tile_q", ctypes.c_int32),
                ("kernel", ctypes.c_int32), ("refine_iters", ctypes.c_int32),
                ("xcd_tiles", ctypes.c_int32), ("tile_waves", ctypes.c_int32),
                ("tile_split", ctypes.c_int32), ("split_rounds", ctypes.c_int32),
                ("autotune", ctypes.c_int32), ("graph", ctypes.c_int32), ("kcap", ctypes.c_int32),
                ("kcap_mfma_max", ctypes.c_int32), ("dedup", ctypes.c_int32),
                ("filter_stats", ctypes.c_int32), ("work_queue", ctypes.c_int32),
                ("rep_scale", ctypes.c_int32), ("reserved", ctypes.c_int32 * 8)]


class Stats(ctypes.Structure):
    _fields_ = [("t_total_ms", ctypes.c_double), ("t_index_ms", ctypes.c_double),
                ("t_bounds_ms", ctypes.c_double), ("t_join_ms", ctypes.c_double),
                ("t_fallback_ms", ctypes.c_double), ("n_dist_evals", ctypes.c_int64),
                ("pruning_ratio", ctypes.c_double), ("n_cluster_pairs", ctypes.c_int64),
                ("n_overflow", ctypes.c_int64), ("exact_mode", ctypes.c_int32),
                ("precision_used", ctypes.c_int32), ("m_q", ctypes.c_int32), ("m_r", ctypes.c_int32),
                ("tile_q", ctypes.c_int32), ("kcap", ctypes.c_int32), ("n_tiles", ctypes.c_int64),
                ("filter_max_abs_err", ctypes.c_double), ("kernel_used", ctypes.c_int32),
                ("n_launches", ctypes.c_int32), ("n_merges", ctypes.c_int64),
                ("n_appended", ctypes.c_int64), ("tuned", ctypes.c_int32), ("graph", ctypes.c_int32),
                ("rerun_stages", ctypes.c_int32), ("queue_launches", ctypes.c_int32),
                ("filter_err_max", ctypes.c_double), ("filter_bound_max", ctypes.c_double),
                ("filter_err_ratio", ctypes.c_double), ("n_distinct_queries", ctypes.c_int64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_
                if not name.startswith("reserved")}


class Instance(ctypes.Structure):
    _fields_ = [("R", Matrix), ("Q", Matrix), ("k", ctypes.c_int32),
                ("out_idx", ctypes.c_void_p), ("out_dist", ctypes.c_void_p)]


class IndexDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("d", ctypes.c_int32), ("m_r", ctypes.c_int32),
                ("kernel", ctypes.c_int32), ("precision", ctypes.c_int32), ("device", ctypes.c_int32),
                ("reserved0", ctypes.c_int32), ("bytes", ctypes.c_int64), ("t_build_ms", ctypes.c_double)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int32), ("slots", ctypes.c_int32), ("max_batch_rows", ctypes.c_int64),
                ("slot_budget", ctypes.c_int64), ("opts", Opts)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("batch_rows", ctypes.c_int32),
                ("batch_requests", ctypes.c_int32), ("slot", ctypes.c_int32),
                ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("t_submit_ms", ctypes.c_double), ("t_start_ms", ctypes.c_double),
                ("t_end_ms", ctypes.c_double), ("message", ctypes.c_char * 160)]


class PlanDesc(ctypes.Structure):
    _fields_ = [("n_r", ctypes.c_int64), ("n_q", ctypes.c_int64), ("d", ctypes.c_int32),
                ("dpad", ctypes.c_int32), ("m_r", ctypes.c_int32), ("m_q", ctypes.c_int32),
                ("k", ctypes.c_int32), ("self_join", ctypes.c_int32), ("valid", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 5)]


# parts of the last join's plan (knnj_plan_part): name -> (id, dtype, shape from the desc)
PLAN_PARTS = {
    "r_reps": (0, np.float32, lambda p: (p["m_r"], p["dpad"])),
    "r_radius": (1, np.float64, lambda p: (p["m_r"],)),
    "r_offsets": (2, np.int32, lambda p: (p["m_r"] + 1,)),
    "r_perm": (3, np.int32, lambda p: (p["n_r"],)),
    "r_pdist": (4, np.float32, lambda p: (p["n_r"],)),
    "q_reps": (5, np.float32, lambda p: (p["m_q"], p["dpad"])),
    "q_radius": (6, np.float64, lambda p: (p["m_q"],)),
    "q_offsets": (7, np.int32, lambda p: (p["m_q"] + 1,)),
    "q_perm": (8, np.int32, lambda p: (p["n_q"],)),
    "q_pdist": (9, np.float32, lambda p: (p["n_q"],)),
    "theta": (10, np.float64, lambda p: (p["m_q"],)),
    "pair_count": (11, np.int32, lambda p: (p["m_q"],)),
    "pair_list": (12, np.int32, lambda p: (p["m_q"], p["m_r"])),
    "pair_lb": (13, np.float32, lambda p: (p["m_q"], p["m_r"])),
}


EXPORTS = ("knnj_version", "knnj_last_error", "knnj_ctx_create", "knnj_ctx_destroy",
           "knnj_ctx_set_budget", "knnj_join", "knnj_join_batch", "knnj_index_build",
           "knnj_index_destroy", "knnj_index_info", "knnj_index_query", "knnj_server_create",
           "knnj_server_submit", "knnj_server_wait", "knnj_server_counters", "knnj_server_destroy",
           "knnj_ctx_last_plan", "knnj_ctx_copy_plan_part", "knnj_index_query_noalloc")

# bytes of the per-device scratch array of knnj_index_query_noalloc
NOALLOC_SCRATCH_BYTES = 64 << 20

_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libknnj.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m knn_sweet_amd.build`")
    lib = ctypes.CDLL(path)
    lib.knnj_version.restype = ctypes.c_char_p
    lib.knnj_last_error.restype = ctypes.c_char_p
    lib.knnj_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.knnj_ctx_create.restype = ctypes.c_int
    lib.knnj_ctx_destroy.argtypes = [ctypes.c_void_p]
    lib.knnj_ctx_destroy.restype = None
    lib.knnj_ctx_set_budget.argtypes = [ctypes.c_void_p, ctypes.c_int64]
    lib.knnj_ctx_set_budget.restype = ctypes.c_int
    lib.knnj_join.argtypes = [ctypes.c_void_p, Matrix, Matrix, ctypes.c_int32, ctypes.POINTER(Opts),
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Stats)]
    lib.knnj_join.restype = ctypes.c_int
    lib.knnj_join_batch.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Instance),
                                    ctypes.POINTER(Opts), ctypes.POINTER(Stats)]
    lib.knnj_join_batch.restype = ctypes.c_int
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    lib.knnj_index_build.argtypes = [vp, Matrix, ctypes.POINTER(Opts), ctypes.POINTER(vp)]
    lib.knnj_index_build.restype = ctypes.c_int
    lib.knnj_index_destroy.argtypes = [vp]
    lib.knnj_index_destroy.restype = None
    lib.knnj_index_info.argtypes = [vp, ctypes.POINTER(IndexDesc)]
    lib.knnj_index_info.restype = ctypes.c_int
    lib.knnj_index_query.argtypes = [vp, vp, Matrix, i32, ctypes.POINTER(Opts), vp, vp, ctypes.POINTER(Stats)]
    lib.knnj_index_query.restype = ctypes.c_int
    lib.knnj_index_query_noalloc.argtypes = [vp, vp, Matrix, i32, ctypes.POINTER(Opts), vp, vp, ctypes.POINTER(Stats)]
    lib.knnj_index_query_noalloc.restype = ctypes.c_int
    lib.knnj_server_create.argtypes = [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
    lib.knnj_server_create.restype = ctypes.c_int
    lib.knnj_server_submit.argtypes = [vp, vp, i64, i64, vp, vp, ctypes.POINTER(ctypes.c_uint64)]
    lib.knnj_server_submit.restype = ctypes.c_int
    lib.knnj_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(RequestInfo)]
    lib.knnj_server_wait.restype = ctypes.c_int
    lib.knnj_server_counters.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
    lib.knnj_server_counters.restype = ctypes.c_int
    lib.knnj_server_destroy.argtypes = [vp]
    lib.knnj_server_destroy.restype = None
    lib.knnj_ctx_last_plan.argtypes = [vp, ctypes.POINTER(PlanDesc)]
    lib.knnj_ctx_last_plan.restype = ctypes.c_int
    lib.knnj_ctx_copy_plan_part.argtypes = [vp, i32, vp, i64]
    lib.knnj_ctx_copy_plan_part.restype = ctypes.c_int
    _lib = lib
    return lib


def _check(status: int):
    if status != KNNJ_OK:
        raise KnnjError(status, _lib.knnj_last_error().decode())


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _matrix(x):
    """(Matrix, keepalive, on_device) for a numpy array or a torch tensor."""
    if _is_torch(x):
        import torch
        if x.dtype != torch.float32 or x.dim() != 2:
            raise ValueError("expected a 2-D float32 tensor")
        if x.stride(1) != 1:
            x = x.contiguous()
        dev = 1 if x.is_cuda else 0
        if not dev:
            x = x.contiguous()
        return Matrix(x.data_ptr(), x.shape[0], x.shape[1], x.stride(0), dev), x, bool(dev)
    a = np.asarray(x)
    if a.ndim != 2:
        raise ValueError("expected a 2-D array")
    a = np.ascontiguousarray(a, dtype=np.float32)
    return Matrix(a.ctypes.data, a.shape[0], a.shape[1], a.shape[1], 0), a, False


def _outputs(nq: int, k: int, aos: bool, qdev):
    """Output arrays (idx, dist) and their pointers: torch on qdev's device, else numpy."""
    if qdev is not None:
        import torch
        idx = torch.empty((nq, k, 2) if aos else (nq, k), dtype=torch.int32, device=qdev.device)
        dist = None if aos else torch.empty((nq, k), dtype=torch.float32, device=qdev.device)
        return idx, dist, idx.data_ptr(), (dist.data_ptr() if dist is not None else None)
    idx = np.empty((nq, k, 2) if aos else (nq, k), np.int32)
    dist = None if aos else np.empty((nq, k), np.float32)
    return idx, dist, idx.ctypes.data, (dist.ctypes.data if dist is not None else None)


def make_opts(m_q=0, m_r=0, trials=0, precision=PREC_AUTO, seed=0, out_squared=False, out_aos=False,
              brute_force=False, tile_q=0, kernel=KERNEL_AUTO, refine_iters=0, xcd_tiles=False,
              tile_waves=0, tile_split=0, split_rounds=0, autotune=False, graph=False, kcap=0,
              kcap_mfma_max=0, dedup=0, filter_stats=False, work_queue=False, rep_scale=0) -> Opts:
    return Opts(m_q, m_r, trials, precision, seed, int(out_squared), int(out_aos), int(brute_force),
                tile_q, kernel, refine_iters, int(xcd_tiles), tile_waves, tile_split, split_rounds,
                int(autotune), int(graph), kcap, kcap_mfma_max, dedup, int(filter_stats), int(work_queue),
                rep_scale)


class Context:
    """A device context (stream, events, workspace) on one GPU."""

    def __init__(self, device: int = 0):
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.knnj_ctx_create(device, ctypes.byref(h)))
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            _lib.knnj_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_budget(self, nbytes: int):
        _check(_lib.knnj_ctx_set_budget(self._h, nbytes))

    def join(self, R, Q, k: int, **opts):
        """Exact kNN join; returns (idx, dist, stats-dict)."""
        mR, keepR, _ = _matrix(R)
        mQ, keepQ, qdev = _matrix(Q)
        if R is Q:
            mQ = mR
        o = make_opts(**opts)
        idx, dist, pi, pd = _outputs(mQ.n, k, bool(o.out_aos), keepQ if qdev else None)
        st = Stats()
        _check(_lib.knnj_join(self._h, mR, mQ, k, ctypes.byref(o), pi, pd, ctypes.byref(st)))
        del keepR, keepQ
        return idx, dist, st.as_dict()

    def last_plan(self) -> dict:
        """Sizes of the last single join's cluster index and bounds (knnj_ctx_last_plan)."""
        v = PlanDesc()
        _check(_lib.knnj_ctx_last_plan(self._h, ctypes.byref(v)))
        return {name: getattr(v, name) for name, _ in v._fields_ if name != "reserved"}

    def plan_part(self, name: str) -> np.ndarray:
        """One part of the last single join's plan (see PLAN_PARTS) as a host array."""
        pid, dt, shape = PLAN_PARTS[name]
        out = np.empty(shape(self.last_plan()), dtype=dt)
        _check(_lib.knnj_ctx_copy_plan_part(self._h, pid, out.ctypes.data, out.nbytes))
        return out

    def build_index(self, R, **opts) -> "Index":
        """A resident index of R on this context's device (see Index)."""
        return Index(self, R, **opts)

    def join_batch(self, instances, outputs=None, **opts):
        """Many independent joins in one call.

        instances: sequence of (R, Q, k).  Returns ([(idx, dist), ...], stats-dict)
        with each result equal to ``join(R, Q, k)`` of that instance alone.
        outputs: optional [(idx, dist), ...] from an earlier call to write into
        (with ``graph=True`` a plan whose pointers do not change replays its
        captured HIP graph).
        """
        o = make_opts(**opts)
        arr = (Instance * max(len(instances), 1))()
        keep, results = [], []
        for i, (R, Q, k) in enumerate(instances):
            mR, keepR, _ = _matrix(R)
            mQ, keepQ, qdev = _matrix(Q)
            if R is Q:
                mQ = mR
            if outputs is not None:
                idx, dist = outputs[i]
                pi = idx.data_ptr() if _is_torch(idx) else idx.ctypes.data
                pd = None if dist is None else (dist.data_ptr() if _is_torch(dist) else dist.ctypes.data)
            else:
                idx, dist, pi, pd = _outputs(mQ.n, k, bool(o.out_aos), keepQ if qdev else None)
            arr[i] = Instance(mR, mQ, k, pi, pd)
            keep += [keepR, keepQ]
            results.append((idx, dist))
        st = Stats()
        _check(_lib.knnj_join_batch(self._h, len(instances), arr, ctypes.byref(o), ctypes.byref(st)))
        del keep
        return results, st.as_dict()


class Index:
    """A resident reference-side index (knnj_index): R's clusters, permuted
    rows and centred operand rows kept in HBM for repeated queries."""

    def __init__(self, ctx: "Context", R, **opts):
        mR, keep, _ = _matrix(R)
        o = make_opts(**opts)
        h = ctypes.c_void_p()
        _check(_lib.knnj_index_build(ctx._h, mR, ctypes.byref(o), ctypes.byref(h)))
        del keep
        self._h = h
        self.ctx = ctx

    def info(self) -> dict:
        v = IndexDesc()
        _check(_lib.knnj_index_info(self._h, ctypes.byref(v)))
        return {name: getattr(v, name) for name, _ in v._fields_ if not name.startswith("reserved")}

    def query(self, Q, k: int, ctx: "Context" = None, no_alloc: bool = False, out=None, **opts):
        """Exact kNN of Q against the indexed R; returns (idx, dist, stats-dict).

        no_alloc: run knnj_index_query_noalloc (no device allocation in the
        library; its buffers come from a fixed per-device scratch array).
        out: optional (idx, dist) to write into instead of new arrays."""
        ctx = ctx or self.ctx
        mQ, keepQ, qdev = _matrix(Q)
        o = make_opts(**opts)
        if out is not None:
            idx, dist = out
            pi = idx.data_ptr() if _is_torch(idx) else idx.ctypes.data
            pd = None if dist is None else (dist.data_ptr() if _is_torch(dist) else dist.ctypes.data)
        else:
            idx, dist, pi, pd = _outputs(mQ.n, k, bool(o.out_aos), keepQ if qdev else None)
        st = Stats()
        fn = _lib.knnj_index_query_noalloc if no_alloc else _lib.knnj_index_query
        _check(fn(ctx._h, self._h, mQ, k, ctypes.byref(o), pi, pd, ctypes.byref(st)))
        del keepQ
        return idx, dist, st.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            _lib.knnj_index_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop over an Index (knnj_server).

    submit() queues host query rows and returns a ticket at once; wait()
    blocks until that request's results are written and returns
    (idx, dist, info).  Thread-safe; ctypes releases the GIL in both calls.
    """

    def __init__(self, index: Index, k: int, slots: int = 2, max_batch_rows: int = 4096,
                 slot_budget: int = 0, **opts):
        cfg = ServerConfig(k, slots, max_batch_rows, slot_budget, make_opts(**opts))
        h = ctypes.c_void_p()
        st = _lib.knnj_server_create(index._h, ctypes.byref(cfg), ctypes.byref(h))
        if st != KNNJ_OK:
            raise KnnjError(st, "server creation failed")
        self._h, self.index, self.k = h, index, k
        self._pending = {}

    def submit(self, Q) -> int:
        a = np.ascontiguousarray(Q, dtype=np.float32)
        if a.ndim != 2:
            raise ValueError("expected a 2-D array")
        idx = np.empty((a.shape[0], self.k), np.int32)
        dist = np.empty((a.shape[0], self.k), np.float32)
        t = ctypes.c_uint64()
        st = _lib.knnj_server_submit(self._h, a.ctypes.data, a.shape[0], a.shape[1], idx.ctypes.data,
                                     dist.ctypes.data, ctypes.byref(t))
        if st != KNNJ_OK:
            raise KnnjError(st, "submit rejected")
        self._pending[t.value] = (a, idx, dist)
        return t.value

    def wait(self, ticket: int):
        info = RequestInfo()
        st = _lib.knnj_server_wait(self._h, ticket, ctypes.byref(info))
        _, idx, dist = self._pending.pop(ticket)
        if st != KNNJ_OK:
            raise KnnjError(st, info.message.decode())
        d = {name: getattr(info, name) for name, _ in info._fields_ if name != "message"}
        return idx, dist, d

    def counters(self):
        b, r = ctypes.c_int64(), ctypes.c_int64()
        _check(_lib.knnj_server_counters(self._h, ctypes.byref(b), ctypes.byref(r)))
        return b.value, r.value

    def close(self):
        if getattr(self, "_h", None):
            _lib.knnj_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = {}


def join(R, Q, k: int, device: int = 0, **opts):
    """Exact kNN join on a per-process default context of `device`."""
    ctx = _default_ctx.get(device)
    if ctx is None:
        ctx = _default_ctx[device] = Context(device)
    return ctx.join(R, Q, k, **opts)


def version() -> str:
    return load_library().knnj_version().decode()
