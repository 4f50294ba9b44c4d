"""Python mirror of the manifolds-rs `adam_gpu` fitting path over the C ABI.

Loads ``libmrs_amd.so`` (built in-tree next to this file) with ctypes.  The
library is the product: there is no fallback, and every call fails loudly if
the library or a HIP device is missing.  Names follow the reference crate
(src/lib.rs, src/training/umap_optimiser_gpu.rs); buffers are numpy arrays
(host) or torch tensors (device: used in place).
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libmrs_amd.so"
HEADER_PATH = _HERE.parent / "include" / "mrs_amd.h"

MRS_OK = 0
MRS_NO_DATA = 1
MRS_NO_GRAPH_EDGES = 2
MRS_DEGENERATE_LOCAL_RADII = 3
MRS_DEVICE_ERROR = 4
MRS_INVALID_ARGUMENT = 5

EPOCH_EDGE_ATOMIC = 0
EPOCH_NODE_EXACT = 1

KNN_FP32 = 0
KNN_BF16_RERANK = 1

_STATUS_NAMES = {
    1: "NoData",
    2: "NoGraphEdges",
    3: "DegenerateLocalRadii",
    4: "DeviceError",
    5: "InvalidArgument",
}


class ManifoldsError(RuntimeError):
    """A non-zero status from the C ABI (ManifoldsError in the reference)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class UmapOptimParams(C.Structure):
    """UmapOptimParams<f32> (src/training/umap_optimisers.rs:22-44)."""

    _fields_ = [
        ("a", C.c_float), ("b", C.c_float), ("lr", C.c_float), ("gamma", C.c_float),
        ("n_epochs", C.c_int32), ("neg_sample_rate", C.c_int32),
        ("min_dist", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
    ]

    @classmethod
    def default_2d(cls) -> "UmapOptimParams":
        # UmapOptimParams::default_2d, src/training/umap_optimisers.rs:56-69
        return cls(a=1.5, b=0.9, lr=1.0, gamma=1.0, n_epochs=500, neg_sample_rate=5,
                   min_dist=0.1, beta1=0.5, beta2=0.9, eps=1e-7)

    @classmethod
    def from_min_dist_spread(cls, min_dist: float, spread: float, lr=None, gamma=None,
                             n_epochs=None, neg_sample_rate=None, beta1=None, beta2=None,
                             eps=None) -> "UmapOptimParams":
        # UmapOptimParams::from_min_dist_spread, src/training/umap_optimisers.rs:89-122:
        # a, b fitted to (min_dist, spread) by the library; None = the reference's default
        nan = float("nan")
        f = lambda v: nan if v is None else float(v)
        i = lambda v: -1 if v is None else int(v)
        out = cls()
        _check(lib().mrs_umap_optim_params_from_min_dist_spread(
            min_dist, spread, f(lr), f(gamma), i(n_epochs), i(neg_sample_rate), f(beta1),
            f(beta2), f(eps), C.byref(out)))
        return out


class DensParams(C.Structure):
    """DensParams<f32> (src/utils/density.rs:69-79); defaults :38-55."""

    _fields_ = [("lambda_", C.c_float), ("frac", C.c_float), ("var_shift", C.c_float)]

    @classmethod
    def densmap_default(cls) -> "DensParams":
        return cls(lambda_=2.0, frac=0.3, var_shift=0.1)


class DensState(C.Structure):
    """DensState<f32> (src/utils/density.rs:140-160)."""

    _fields_ = [("params", DensParams), ("r", C.c_void_p), ("mu_sum", C.c_void_p),
                ("mu_tot", C.c_float)]


class UmapGraphParams(C.Structure):
    """UmapGraphParams<f32> (src/data/graph.rs:30-61)."""

    _fields_ = [("bandwidth", C.c_float), ("local_connectivity", C.c_float),
                ("mix_weight", C.c_float)]

    @classmethod
    def default(cls) -> "UmapGraphParams":
        return cls(bandwidth=1e-5, local_connectivity=1.0, mix_weight=1.0)


class UmapParamsGpu(C.Structure):
    """UmapParamsGpu<f32> (src/lib.rs:3218-3241) + kNN/epoch knobs of this build."""

    _fields_ = [
        ("n_dim", C.c_int32), ("k", C.c_int32), ("optimiser", C.c_char_p),
        ("ann_type", C.c_char_p), ("initialisation", C.c_char_p), ("init_range", C.c_float),
        ("dist_metric", C.c_char_p), ("knn_precision", C.c_int32), ("knn_shortlist", C.c_int32),
        ("umap_graph_params", UmapGraphParams), ("optim_params", UmapOptimParams),
        ("randomised", C.c_int32), ("epoch_strategy", C.c_int32),
    ]

    @classmethod
    def default(cls) -> "UmapParamsGpu":
        p = cls()
        lib().mrs_umap_params_gpu_default(C.byref(p))
        return p

    @classmethod
    def new_default_2d(cls, min_dist=None, spread=None) -> "UmapParamsGpu":
        # UmapParamsGpu::new_default_2d, src/lib.rs:3338-3346
        p = cls()
        nan = float("nan")
        lib().mrs_umap_params_gpu_default_2d(nan if min_dist is None else float(min_dist),
                                             nan if spread is None else float(spread), C.byref(p))
        return p


_lib = None


def _declared_functions() -> list[str]:
    """Names of every function declared in include/mrs_amd.h."""
    import re

    text = HEADER_PATH.read_text()
    return re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(mrs_\w+)\s*\(", text, re.M)


def lib() -> C.CDLL:
    """Load libmrs_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    # One HIP runtime per process.  PyTorch's ROCm wheels carry their own
    # libamdhip64 / libhsa-runtime64 under the same sonames: with torch loaded
    # first this library binds to those; loaded first itself, it would bring
    # up /opt/rocm's runtime beside torch's, and torch would then find no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    i32, i64, u32, u64, f32p = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_void_p
    vp = C.c_void_p
    sig = {
        "mrs_last_error": (C.c_char_p, []),
        "mrs_ctx_create": (C.c_int, [i32, C.c_size_t, C.POINTER(vp)]),
        "mrs_ctx_destroy": (C.c_int, [vp]),
        "mrs_ctx_synchronize": (C.c_int, [vp]),
        "mrs_ctx_stream": (vp, [vp]),
        "mrs_ctx_set_knn_pruning": (C.c_int, [vp, i32]),
        "mrs_ctx_set_knn_work_chunk": (C.c_int, [vp, i32]),
        "mrs_umap_optim_params_from_min_dist_spread": (C.c_int, [
            C.c_float, C.c_float, C.c_float, C.c_float, i32, i32, C.c_float, C.c_float,
            C.c_float, C.POINTER(UmapOptimParams)]),
        "mrs_umap_params_gpu_default_2d": (None, [C.c_float, C.c_float,
                                                  C.POINTER(UmapParamsGpu)]),
        "mrs_probe_random_gather": (C.c_int, [vp, i64, i64, i32, C.POINTER(C.c_double)]),
        "mrs_probe_random_atomic": (C.c_int, [vp, i64, i64, i32, C.POINTER(C.c_double)]),
        "mrs_ctx_cached_bytes": (C.c_size_t, [vp]),
        "mrs_ctx_release_cache": (C.c_int, [vp]),
        "mrs_parse_umap_optimiser_gpu": (i32, [C.c_char_p]),
        "mrs_parse_ann_search_gpu": (i32, [C.c_char_p]),
        "mrs_hash_neg": (u32, [u32, u32, u32, u32, u32, u32]),
        "mrs_umap_state_create": (C.c_int, [vp, vp, i64, i32, vp, vp, vp,
                                            C.POINTER(UmapOptimParams), C.POINTER(DensState),
                                            C.POINTER(vp)]),
        "mrs_umap_state_destroy": (C.c_int, [vp]),
        "mrs_umap_state_info": (C.c_int, [vp, C.POINTER(i64), C.POINTER(i32), C.POINTER(i64)]),
        "mrs_umap_state_read": (C.c_int, [vp, i32, vp]),
        "mrs_umap_state_write": (C.c_int, [vp, i32, vp]),
        "mrs_launch_grad_accum": (C.c_int, [vp, i32, u32, C.POINTER(C.c_float)]),
        "mrs_launch_dens_update": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "mrs_launch_adam_update": (C.c_int, [vp, i32]),
        "mrs_launch_edge_schedule_update": (C.c_int, [vp, i32]),
        "mrs_launch_grad_norm": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "mrs_run_epochs": (C.c_int, [vp, i32, i32, u64, i32]),
        "mrs_launch_epoch_edges": (C.c_int, [vp, i32, u32, C.POINTER(C.c_float)]),
        "mrs_launch_adam_clear": (C.c_int, [vp, i32]),
        "mrs_umap_state_set_profiling": (C.c_int, [vp, i32]),
        "mrs_umap_state_reset": (C.c_int, [vp, vp]),
        "mrs_schedule_active_counts": (C.c_int, [vp, i64, i32, vp]),
        "mrs_umap_state_kernel_times": (C.c_int, [vp, C.POINTER(C.c_double),
                                                  C.POINTER(C.c_int64)]),
        "mrs_knn_exhaustive": (C.c_int, [vp, vp, i64, i32, i32, i32, i32, vp, vp]),
        "mrs_knn_exhaustive_noalloc": (C.c_int, [vp, vp, i64, i32, i32, i32, i32, vp, vp]),
        "mrs_knn_exhaustive_timed": (C.c_int, [vp, vp, i64, i32, i32, i32, i32, vp, vp,
                                               C.POINTER(C.c_double)]),
        "mrs_fuzzy_graph": (C.c_int, [vp, vp, vp, i64, i32, C.POINTER(UmapGraphParams), i32,
                                      C.POINTER(vp)]),
        "mrs_graph_from_coo": (C.c_int, [vp, i64, vp, vp, vp, i64, i32, C.POINTER(vp)]),
        "mrs_graph_destroy": (C.c_int, [vp]),
        "mrs_graph_info": (C.c_int, [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]),
        "mrs_graph_export_coo": (C.c_int, [vp, vp, vp, vp]),
        "mrs_graph_sigma_rho": (C.c_int, [vp, vp, vp]),
        "mrs_dens_state_from_graph": (C.c_int, [vp, vp, vp, vp, i32, vp, vp,
                                                C.POINTER(C.c_float)]),
        "mrs_umap_state_from_graph": (C.c_int, [vp, vp, vp, i32, C.POINTER(UmapOptimParams),
                                                C.POINTER(DensState), C.POINTER(vp)]),
        "mrs_umap_params_gpu_default": (None, [C.POINTER(UmapParamsGpu)]),
        "mrs_spectral_layout": (C.c_int, [vp, vp, i32, u64, C.c_float, vp, i32, vp]),
        "mrs_umap_gpu": (C.c_int, [vp, vp, i64, i32, vp, vp, i32, C.POINTER(UmapParamsGpu), u64,
                                   i32, vp]),
        "mrs_densmap_gpu": (C.c_int, [vp, vp, i64, i32, vp, vp, i32, C.POINTER(UmapParamsGpu),
                                      C.POINTER(DensParams), u64, i32, vp]),
        "mrs_last_fit_timings": (C.c_int, [vp]),
        "mrs_fit_knn_copy_to": (C.c_int, [vp, vp, i64]),
        "mrs_fit_bytes_estimate": (C.c_size_t, [i64, i32, i32, i32]),
        "mrs_tridiag_eigen": (C.c_int, [vp, vp, i32, vp, vp]),
        "mrs_server_create": (C.c_int, [i32, C.c_size_t, i32, i32, C.POINTER(vp)]),
        "mrs_server_submit": (C.c_int, [vp, vp, i64, i32, C.POINTER(UmapParamsGpu), u64,
                                        C.POINTER(i64)]),
        "mrs_server_poll": (C.c_int, [vp, i64, C.POINTER(i32)]),
        "mrs_server_wait": (C.c_int, [vp, i64, vp, vp]),
        "mrs_server_stats": (C.c_int, [vp, vp]),
        "mrs_server_destroy": (C.c_int, [vp]),
        "mrs_pca_layout": (C.c_int, [vp, vp, i64, i32, i32, i32, C.c_float, u64, vp]),
        "mrs_autotune_set_mode": (C.c_int, [i32]),
        "mrs_autotune_force": (C.c_int, [i32, i32]),
        "mrs_autotune_clear": (C.c_int, []),
        "mrs_autotune_dump": (i64, [C.c_char_p, i64]),
        "mrs_autotune_save": (C.c_int, [C.c_char_p]),
        "mrs_autotune_load": (C.c_int, [C.c_char_p]),
        "mrs_optimise_embedding_adam_gpu": (C.c_int, [vp, vp, i64, i32, vp, vp, vp,
                                                      C.POINTER(UmapOptimParams),
                                                      C.POINTER(DensState), u64, i32, i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != MRS_OK:
        msg = lib().mrs_last_error().decode(errors="replace")
        raise ManifoldsError(rc, msg)


def _ptr(x) -> int | None:
    """Address of a numpy array (host) or torch tensor (host or device)."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        if not x.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous")
        return x.ctypes.data
    if hasattr(x, "data_ptr"):
        if not x.is_contiguous():
            raise ValueError("tensors must be contiguous")
        return x.data_ptr()
    raise TypeError(f"unsupported buffer type {type(x)}")


class Context:
    """mrs_ctx: one HIP stream on one device.  Graphs and optimiser states made
    on a context hold its stream: closing the context closes them first."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._deps = weakref.WeakSet()
        h = C.c_void_p()
        _check(lib().mrs_ctx_create(device, hbm_budget, C.byref(h)))
        self.handle = h

    def _adopt(self, obj) -> None:
        self._deps.add(obj)

    def synchronize(self) -> None:
        _check(lib().mrs_ctx_synchronize(self.handle))

    @property
    def stream(self) -> int:
        return lib().mrs_ctx_stream(self.handle)

    def cached_bytes(self) -> int:
        """Device memory the context keeps for reuse by later calls."""
        return lib().mrs_ctx_cached_bytes(self.handle)

    def release_cache(self) -> None:
        """Wait for the device and return the context's cached memory to it."""
        _check(lib().mrs_ctx_release_cache(self.handle))

    def set_knn_pruning(self, on) -> None:
        """Exhaustive kNN skips database stages that provably hold no candidate
        (True or 1, the default), or sweeps every stage: in cell order (False or
        0) or in row order (-1).  All give the same results."""
        mode = on if isinstance(on, int) and not isinstance(on, bool) else (1 if on else 0)
        _check(lib().mrs_ctx_set_knn_pruning(self.handle, mode))

    def set_knn_work_chunk(self, stages: int) -> None:
        """Second-pass stage lists longer than `stages` are cut into work items
        merged afterwards (0: the size-based default); results do not change."""
        _check(lib().mrs_ctx_set_knn_work_chunk(self.handle, int(stages)))

    def probe_random_gather(self, rows: int, gathers: int = 1 << 24, reps: int = 10) -> float:
        """Uniform random 8-byte gathers per second from a `rows`-row float2
        table (the epoch kernel's request pattern), measured on this context."""
        out = C.c_double(0.0)
        _check(lib().mrs_probe_random_gather(self.handle, int(rows), int(gathers), int(reps),
                                             C.byref(out)))
        return out.value

    def probe_random_atomic(self, rows: int, ops: int = 1 << 24, reps: int = 10) -> float:
        """FP32 atomic adds (no return) per second to uniform random rows of a
        `rows`-row float2 table (the epoch kernel's gradient scatter)."""
        out = C.c_double(0.0)
        _check(lib().mrs_probe_random_atomic(self.handle, int(rows), int(ops), int(reps),
                                             C.byref(out)))
        return out.value

    def close(self) -> None:
        if self.handle:
            for obj in list(self._deps):
                obj.close()
            lib().mrs_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def adjacency_to_csr(graph: list[list[tuple[int, float]]]):
    """Vec<Vec<(usize, T)>> adjacency list -> (offsets int64, cols int32, vals f32)."""
    off = np.zeros(len(graph) + 1, dtype=np.int64)
    for i, row in enumerate(graph):
        off[i + 1] = off[i] + len(row)
    cols = np.array([j for row in graph for j, _ in row], dtype=np.int32)
    vals = np.array([w for row in graph for _, w in row], dtype=np.float32)
    return off, cols, vals


# state buffer ids of mrs_umap_state_read/write
BUF_EMBD, BUF_M, BUF_V, BUF_GRAD, BUF_HAS_UPDATE = 0, 1, 2, 3, 4
BUF_OFFSETS, BUF_EDGE_IDX, BUF_OTHER, BUF_PER, BUF_NEXT, BUF_MU = 5, 6, 7, 8, 9, 10
BUF_DENS_RE, BUF_DENS_PHI, BUF_DENS_W, BUF_DENS_R, BUF_EDGE_IJ = 11, 12, 13, 14, 15
BUF_EDGE_SLOTS = 16


def knn_exhaustive(ctx: Context, x, k: int, precision: int = KNN_FP32, shortlist: int = 0,
                   timed: bool = False):
    """run_ann_search_gpu("exhaustive_gpu"): (idx [n,k] int32, dist [n,k] f32 squared).

    x may be a numpy array (results returned as numpy) or a device tensor
    (results returned as device tensors)."""
    if isinstance(x, np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        n, d = x.shape
        idx = np.empty((n, k), np.int32)
        dist = np.empty((n, k), np.float32)
    else:
        import torch

        n, d = x.shape
        idx = torch.empty((n, k), dtype=torch.int32, device=x.device)
        dist = torch.empty((n, k), dtype=torch.float32, device=x.device)
    if timed:
        ms = (C.c_double * 10)()
        _check(lib().mrs_knn_exhaustive_timed(ctx.handle, _ptr(x), n, d, k, precision, shortlist,
                                              _ptr(idx), _ptr(dist), ms))
        return idx, dist, dict(pack_ms=ms[0], shortlist_ms=ms[1], rerank_ms=ms[2],
                               splits=int(ms[3]), shortlist=int(ms[4]), swept=ms[5],
                               cells=int(ms[6]), list_max=int(ms[7]), list_mean=ms[8],
                               aux_items=int(ms[9]))
    _check(lib().mrs_knn_exhaustive(ctx.handle, _ptr(x), n, d, k, precision, shortlist,
                                    _ptr(idx), _ptr(dist)))
    return idx, dist


def knn_exhaustive_noalloc(ctx: Context, x, idx, dist, k: int, precision: int = KNN_FP32,
                           shortlist: int = 0) -> None:
    """The exhaustive kNN into caller-provided device tensors, allocating
    nothing (n <= 16384, d <= 128, shortlist <= 64); queued on the context's
    stream, not waited for."""
    n, d = x.shape
    _check(lib().mrs_knn_exhaustive_noalloc(ctx.handle, _ptr(x), n, d, k, precision, shortlist,
                                            _ptr(idx), _ptr(dist)))


class FuzzyGraph:
    """The device graph of construct_umap_graph_gpu (src/lib.rs:3376-3441)."""

    def __init__(self, ctx: Context, knn_idx, knn_dist, params: UmapGraphParams | None = None,
                 n_epochs: int = 500):
        params = params or UmapGraphParams.default()
        if isinstance(knn_idx, np.ndarray):
            knn_idx = np.ascontiguousarray(knn_idx, dtype=np.int32)
            knn_dist = np.ascontiguousarray(knn_dist, dtype=np.float32)
        n, k = knn_idx.shape
        self.ctx, self.k = ctx, k
        h = C.c_void_p()
        _check(lib().mrs_fuzzy_graph(ctx.handle, _ptr(knn_idx), _ptr(knn_dist), n, k,
                                     C.byref(params), n_epochs, C.byref(h)))
        self.handle = h
        ctx._adopt(self)
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().mrs_graph_info(h, C.byref(a), C.byref(b), C.byref(c)))
        self.n, self.nnz, self.n_edges = a.value, b.value, c.value

    @classmethod
    def from_coo(cls, ctx: Context, n: int, rows, cols, vals, n_epochs: int = 500) -> "FuzzyGraph":
        """A graph from a CoordinateList (symmetric, any entry order): the
        import counterpart of coo()."""
        if isinstance(rows, np.ndarray):
            rows = np.ascontiguousarray(rows, dtype=np.int32)
            cols = np.ascontiguousarray(cols, dtype=np.int32)
            vals = np.ascontiguousarray(vals, dtype=np.float32)
        self = cls.__new__(cls)
        self.ctx, self.k = ctx, 0
        h = C.c_void_p()
        _check(lib().mrs_graph_from_coo(ctx.handle, n, _ptr(rows), _ptr(cols), _ptr(vals),
                                        len(rows), n_epochs, C.byref(h)))
        self.handle = h
        ctx._adopt(self)
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().mrs_graph_info(h, C.byref(a), C.byref(b), C.byref(c)))
        self.n, self.nnz, self.n_edges = a.value, b.value, c.value
        return self

    def coo(self):
        """(rows, cols, vals) of the filtered CoordinateList."""
        r = np.empty(self.nnz, np.int32)
        c = np.empty(self.nnz, np.int32)
        v = np.empty(self.nnz, np.float32)
        _check(lib().mrs_graph_export_coo(self.handle, _ptr(r), _ptr(c), _ptr(v)))
        return r, c, v

    def sigma_rho(self):
        s = np.empty(self.n, np.float32)
        r = np.empty(self.n, np.float32)
        _check(lib().mrs_graph_sigma_rho(self.handle, _ptr(s), _ptr(r)))
        return s, r

    def dens_state(self, knn_idx, knn_dist, distances_squared: bool = True):
        """DensState::new on the device: (r, mu_sum, mu_tot)."""
        if isinstance(knn_idx, np.ndarray):
            knn_idx = np.ascontiguousarray(knn_idx, dtype=np.int32)
            knn_dist = np.ascontiguousarray(knn_dist, dtype=np.float32)
        r = np.empty(self.n, np.float32)
        ms = np.empty(self.n, np.float32)
        tot = C.c_float()
        _check(lib().mrs_dens_state_from_graph(self.ctx.handle, self.handle, _ptr(knn_idx),
                                               _ptr(knn_dist), 1 if distances_squared else 0,
                                               _ptr(r), _ptr(ms), C.byref(tot)))
        return r, ms, tot.value

    def spectral_layout(self, n_dim: int = 2, seed: int = 42, range_: float = float("nan"),
                        data=None) -> np.ndarray:
        """spectral_layout (src/data/init.rs:986-1004) of this graph: [n, n_dim]."""
        out = np.empty((self.n, n_dim), np.float32)
        d = 0
        if data is not None:
            if isinstance(data, np.ndarray):
                data = np.ascontiguousarray(data, dtype=np.float32)
            d = data.shape[1]
        _check(lib().mrs_spectral_layout(self.ctx.handle, self.handle, n_dim, seed, range_,
                                         _ptr(data), d, _ptr(out)))
        return out

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().mrs_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class UmapGpuState:
    """UmapGpuState (src/training/umap_optimiser_gpu.rs:278-452) on the device."""

    def __init__(self, ctx: Context, embd: np.ndarray, adj_off=None, adj_cols=None,
                 adj_vals=None, params: UmapOptimParams | None = None,
                 dens: DensState | None = None, graph: FuzzyGraph | None = None):
        embd = np.ascontiguousarray(embd, dtype=np.float32) if isinstance(embd, np.ndarray) else embd
        n, dim = embd.shape
        self.ctx = ctx
        self._keep = (embd, adj_off, adj_cols, adj_vals, dens, graph)
        h = C.c_void_p()
        if graph is not None:
            _check(lib().mrs_umap_state_from_graph(
                ctx.handle, graph.handle, _ptr(embd), dim, C.byref(params),
                C.byref(dens) if dens is not None else None, C.byref(h)))
        else:
            _check(lib().mrs_umap_state_create(
                ctx.handle, _ptr(embd), n, dim, _ptr(adj_off), _ptr(adj_cols), _ptr(adj_vals),
                C.byref(params), C.byref(dens) if dens is not None else None, C.byref(h)))
        self.handle = h
        ctx._adopt(self)
        n_, d_, e_ = C.c_int64(), C.c_int32(), C.c_int64()
        _check(lib().mrs_umap_state_info(h, C.byref(n_), C.byref(d_), C.byref(e_)))
        self.n, self.n_dim, self.n_edges = n_.value, d_.value, e_.value

    def _shape(self, which: int):
        n, d, e = self.n, self.n_dim, self.n_edges
        table = {0: (n * d, np.float32), 1: (n * d, np.float32), 2: (n * d, np.float32),
                 3: (n * d, np.float32), 4: (n, np.uint32), 5: (n + 1, np.uint32),
                 6: (2 * e, np.uint32), 7: (2 * e, np.uint32), 8: (e, np.float32),
                 9: (e, np.float32), 10: (e, np.float32), 15: (2 * e, np.uint32),
                 16: (2 * e, np.uint32)}
        if which in (11, 12, 13, 14):
            return (n, np.float32)
        return table[which]

    def read(self, which: int) -> np.ndarray:
        size, dt = self._shape(which)
        out = np.empty(size, dtype=dt)
        _check(lib().mrs_umap_state_read(self.handle, which, _ptr(out)))
        return out

    def write(self, which: int, data: np.ndarray) -> None:
        size, dt = self._shape(which)
        arr = np.ascontiguousarray(data, dtype=dt).reshape(-1)
        if arr.size != size:
            raise ValueError(f"buffer {which} expects {size} elements, got {arr.size}")
        _check(lib().mrs_umap_state_write(self.handle, which, _ptr(arr)))

    def launch_grad_accum(self, epoch: int, seed: int, dens_coeff: float | None = None) -> None:
        c = C.c_float(dens_coeff) if dens_coeff is not None else None
        _check(lib().mrs_launch_grad_accum(self.handle, epoch, seed,
                                           C.byref(c) if c is not None else None))

    def launch_dens_update(self) -> float:
        c = C.c_float()
        _check(lib().mrs_launch_dens_update(self.handle, C.byref(c)))
        return c.value

    def launch_adam_update(self, epoch: int) -> None:
        _check(lib().mrs_launch_adam_update(self.handle, epoch))

    def launch_edge_schedule_update(self, epoch: int) -> None:
        _check(lib().mrs_launch_edge_schedule_update(self.handle, epoch))

    def launch_epoch_edges(self, epoch: int, seed: int, dens_coeff: float | None = None) -> None:
        c = C.c_float(dens_coeff) if dens_coeff is not None else None
        _check(lib().mrs_launch_epoch_edges(self.handle, epoch, seed,
                                            C.byref(c) if c is not None else None))

    def launch_adam_clear(self, epoch: int) -> None:
        _check(lib().mrs_launch_adam_clear(self.handle, epoch))

    def launch_grad_norm(self) -> float:
        c = C.c_float()
        _check(lib().mrs_launch_grad_norm(self.handle, C.byref(c)))
        return c.value

    def run_epochs(self, first: int, count: int, seed: int, strategy: int = EPOCH_EDGE_ATOMIC):
        _check(lib().mrs_run_epochs(self.handle, first, count, seed, strategy))

    def reset(self, embd) -> None:
        """Back to the start of a fit with the given initial embedding."""
        if isinstance(embd, np.ndarray):
            embd = np.ascontiguousarray(embd, dtype=np.float32)
        _check(lib().mrs_umap_state_reset(self.handle, _ptr(embd)))

    def set_profiling(self, on: bool) -> None:
        _check(lib().mrs_umap_state_set_profiling(self.handle, 1 if on else 0))

    KERNEL_KINDS = ("epoch_edges", "adam_clear", "grad_accum_node", "adam_elem", "schedule",
                    "dens_update", "fire_bits")

    def kernel_times(self) -> dict:
        """{kind: (total_ms, launches)} of the launches recorded while profiling."""
        ms = (C.c_double * 8)()
        cnt = (C.c_int64 * 8)()
        _check(lib().mrs_umap_state_kernel_times(self.handle, ms, cnt))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.KERNEL_KINDS) if cnt[i]}

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().mrs_umap_state_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def optimise_embedding_adam_gpu(ctx: Context, embd: np.ndarray, adj_off, adj_cols, adj_vals,
                                params: UmapOptimParams, dens: DensState | None = None,
                                seed: int = 42, verbose: int = 0,
                                strategy: int = EPOCH_EDGE_ATOMIC) -> np.ndarray:
    """optimise_embedding_adam_gpu (umap_optimiser_gpu.rs:1302-1369); returns a new array."""
    out = np.ascontiguousarray(embd, dtype=np.float32).copy()
    n, dim = out.shape
    _check(lib().mrs_optimise_embedding_adam_gpu(
        ctx.handle, _ptr(out), n, dim, _ptr(adj_off), _ptr(adj_cols), _ptr(adj_vals),
        C.byref(params), C.byref(dens) if dens is not None else None, seed, verbose, strategy))
    return out


def schedule_active_counts(epochs_per_sample, n_epochs: int) -> np.ndarray:
    """Edges ticking per epoch (host replay of the f32 schedule); sum = edge-updates."""
    per = np.ascontiguousarray(epochs_per_sample, dtype=np.float32)
    out = np.zeros(n_epochs, np.int64)
    _check(lib().mrs_schedule_active_counts(_ptr(per), per.size, n_epochs, _ptr(out)))
    return out


def umap_gpu(ctx: Context, data, params: UmapParamsGpu | None = None, precomputed_knn=None,
             seed: int = 42, verbose: int = 0, dens: DensParams | None = None):
    """umap_gpu / densmap_gpu (src/lib.rs:3465, :3829): returns [n_dim, n] (numpy for host
    input, device tensor for device input)."""
    params = params or UmapParamsGpu.default()
    on_host = data is None or isinstance(data, np.ndarray)
    if data is not None and on_host:
        data = np.ascontiguousarray(data, dtype=np.float32)
    if data is not None:
        n, d = data.shape
    else:
        n, d = precomputed_knn[0].shape[0], 0
    kidx = kdist = None
    kk = 0
    if precomputed_knn is not None:
        kidx, kdist = precomputed_knn
        if isinstance(kidx, np.ndarray):
            kidx = np.ascontiguousarray(kidx, dtype=np.int32)
            kdist = np.ascontiguousarray(kdist, dtype=np.float32)
        kk = kidx.shape[1]
    if on_host:
        out = np.empty((params.n_dim, n), np.float32)
    else:
        import torch

        out = torch.empty((params.n_dim, n), dtype=torch.float32, device=data.device)
    if dens is None:
        _check(lib().mrs_umap_gpu(ctx.handle, _ptr(data), n, d, _ptr(kidx), _ptr(kdist), kk,
                                  C.byref(params), seed, verbose, _ptr(out)))
    else:
        _check(lib().mrs_densmap_gpu(ctx.handle, _ptr(data), n, d, _ptr(kidx), _ptr(kdist), kk,
                                     C.byref(params), C.byref(dens), seed, verbose, _ptr(out)))
    return out


def tridiag_eigen(diag, offdiag):
    """Eigenpairs of a symmetric tridiagonal matrix (the spectral layout's host
    solver): (eigenvalues ascending, eigenvectors as columns)."""
    d = np.ascontiguousarray(diag, dtype=np.float64)
    e = np.ascontiguousarray(offdiag, dtype=np.float64)
    m = d.size
    ev = np.empty(m, np.float64)
    vv = np.empty((m, m), np.float64)
    _check(lib().mrs_tridiag_eigen(_ptr(d), _ptr(e) if m > 1 else None, m, _ptr(ev), _ptr(vv)))
    return ev, vv


SERVE_QUEUED, SERVE_PREPARING, SERVE_EPOCHS, SERVE_DONE, SERVE_FAILED = range(5)


def fit_bytes_estimate(n: int, d: int, k: int = 15, n_dim: int = 2) -> int:
    return int(lib().mrs_fit_bytes_estimate(n, d, k, n_dim))


class Server:
    """Continuous-batching fit server (mrs_server_*): submit() queues one
    umap_gpu fit and returns a ticket; wait(ticket) returns its [n_dim, n]
    embedding (numpy for host data, a device tensor for device data) and the
    request's {queued, preparing, latency} milliseconds.  Submitted arrays are
    kept alive here until their request is collected.  prep_workers 1..8."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, prep_workers: int = 4,
                 max_active: int = 256):
        h = C.c_void_p()
        _check(lib().mrs_server_create(device, hbm_budget, prep_workers, max_active, C.byref(h)))
        self.handle = h
        self._live = {}

    def submit(self, data, params: UmapParamsGpu | None = None, seed: int = 42) -> int:
        params = params or UmapParamsGpu.default()
        if isinstance(data, np.ndarray):
            data = np.ascontiguousarray(data, dtype=np.float32)
        n, d = data.shape
        t = C.c_int64()
        _check(lib().mrs_server_submit(self.handle, _ptr(data), n, d, C.byref(params), seed,
                                       C.byref(t)))
        self._live[t.value] = (data, params.n_dim, n)
        return t.value

    def poll(self, ticket: int) -> int:
        st = C.c_int32()
        _check(lib().mrs_server_poll(self.handle, ticket, C.byref(st)))
        return st.value

    def wait(self, ticket: int):
        data, n_dim, n = self._live.pop(ticket)
        if isinstance(data, np.ndarray):
            out = np.empty((n_dim, n), np.float32)
        else:
            import torch

            out = torch.empty((n_dim, n), dtype=torch.float32, device=data.device)
        times = (C.c_double * 3)()
        _check(lib().mrs_server_wait(self.handle, ticket, _ptr(out), times))
        return out, dict(queued_ms=times[0], prepare_ms=times[1], latency_ms=times[2])

    def stats(self) -> dict:
        v = (C.c_int64 * 7)()
        _check(lib().mrs_server_stats(self.handle, v))
        keys = ("submitted", "completed", "failed", "ticks", "peak_resident", "peak_bytes",
                "budget_bytes")
        return dict(zip(keys, list(v)))

    def close(self) -> None:
        if self.handle:
            _check(lib().mrs_server_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fit_knn_copy_to(idx=None, dist=None) -> None:
    """Later fits on this thread copy the kNN lists they compute into these
    device tensors ([n, k] each, n*k entries at most); None switches it off."""
    cap = 0 if idx is None else int(idx.numel())
    _check(lib().mrs_fit_knn_copy_to(_ptr(idx), _ptr(dist), cap))


def pca_layout(ctx: "Context", x, n_comp: int = 2, randomised: bool = True, range_=None,
               seed: int = 42):
    """pca_layout (src/data/init.rs:1065-1140): [n, n_comp] float32 principal
    component scores, each column centred and scaled to std `range_` (None: 1)."""
    xa = np.ascontiguousarray(x, dtype=np.float32)
    n, d = xa.shape
    out = np.zeros((n, n_comp), np.float32)
    rng = float("nan") if range_ is None else float(range_)
    _check(lib().mrs_pca_layout(ctx.handle, xa.ctypes.data, n, d, n_comp, int(randomised), rng,
                                seed, out.ctypes.data))
    return out


TUNE_OFF, TUNE_ON_FIRST_USE = 0, 1
TUNE_EPOCH_SPAN, TUNE_KNN_SPLITS = 1, 2


def autotune_set_mode(mode: int) -> None:
    """TUNE_ON_FIRST_USE (default): undecided launch shapes are timed on first
    use; TUNE_OFF: built-in choices."""
    _check(lib().mrs_autotune_set_mode(mode))


def autotune_force(family: int, variant: int) -> None:
    """Force a family's variant (TUNE_EPOCH_SPAN: 64/128/256 edges per
    wavefront; TUNE_KNN_SPLITS: 1/2/4/8); -1 releases it."""
    _check(lib().mrs_autotune_force(family, variant))


def autotune_clear() -> None:
    _check(lib().mrs_autotune_clear())


def autotune_records() -> list[dict]:
    """The recorded decisions: family, shape class, chosen variant and the
    measured milliseconds of every candidate."""
    need = lib().mrs_autotune_dump(None, 0)
    buf = C.create_string_buffer(int(need))
    lib().mrs_autotune_dump(buf, need)
    out = []
    for line in buf.value.decode().splitlines():
        f = line.split()
        out.append(dict(family=int(f[0]), shape=int(f[1]), variant=int(f[2]),
                        ms={int(a): float(b) for a, b in (t.split(":") for t in f[3:])}))
    return out


def autotune_save(path: str) -> None:
    _check(lib().mrs_autotune_save(str(path).encode()))


def autotune_load(path: str) -> None:
    _check(lib().mrs_autotune_load(str(path).encode()))


def last_fit_timings() -> dict:
    out = (C.c_double * 8)()
    _check(lib().mrs_last_fit_timings(out))
    keys = ("knn_ms", "graph_ms", "dens_ms", "init_ms", "epochs_ms", "total_ms", "n_edges",
            "knn_shortlist_ms")
    return dict(zip(keys, list(out)))
