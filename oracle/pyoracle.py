"""ctypes wrapper over oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU restatement of the reference path (see oracle.c).  Imported only by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, always as the
checker or the timed CPU baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "liboracle.so"

_lib = None


class OptimParams(C.Structure):
    _fields_ = [
        ("a", C.c_float), ("b", C.c_float), ("lr", C.c_float), ("gamma", C.c_float),
        ("n_epochs", C.c_int32), ("neg_sample_rate", C.c_int32),
        ("min_dist", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
    ]


def params_from(p) -> OptimParams:
    return OptimParams(*(getattr(p, f) for f, _ in OptimParams._fields_))


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: build the oracle (make -C oracle)")
        _lib = C.CDLL(str(LIB_PATH))
        _lib.orc_hash_neg.restype = C.c_uint32
        _lib.orc_hash_neg.argtypes = [C.c_uint32] * 6
        _lib.orc_dens_is_active.restype = C.c_int
        _lib.orc_dens_is_active.argtypes = [C.c_float, C.c_float, C.c_int32, C.c_int32]
        _lib.orc_dens_weights.restype = C.c_float
        for name in ("orc_knn_to_coo", "orc_symmetrise", "orc_filter_weak_edges"):
            getattr(_lib, name).restype = C.c_int64
    return _lib


class _Ptr(C.c_void_p):
    """c_void_p that keeps the array it points into alive for the call."""


def P(a):
    if a is None:
        return None
    p = _Ptr(a.ctypes.data)
    p.keep = a
    return p


def f32(x):
    return np.ascontiguousarray(x, dtype=np.float32)


def i32(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def hash_neg(seed, node, epoch, edge_local, neg, n) -> int:
    return lib().orc_hash_neg(seed, node, epoch, edge_local, neg, n)


def csr_from_graph(adj_off, adj_cols, adj_vals):
    n = len(adj_off) - 1
    ne = C.c_int64()
    adj_off = np.ascontiguousarray(adj_off, dtype=np.int64)
    adj_cols, adj_vals = i32(adj_cols), f32(adj_vals)
    st = lib().orc_csr_from_graph(C.c_int64(n), P(adj_off), P(adj_cols), P(adj_vals),
                                  C.byref(ne), None, None, None, None, None, None, None)
    if st:
        return st, None
    E = ne.value
    out = dict(off=np.zeros(n + 1, np.uint32), edge_idx=np.zeros(2 * E, np.uint32),
               other=np.zeros(2 * E, np.uint32), per=np.zeros(E, np.float32),
               mu=np.zeros(E, np.float32), edge_i=np.zeros(E, np.int32),
               edge_j=np.zeros(E, np.int32))
    lib().orc_csr_from_graph(C.c_int64(n), P(adj_off), P(adj_cols), P(adj_vals), C.byref(ne),
                             P(out["off"]), P(out["edge_idx"]), P(out["other"]), P(out["per"]),
                             P(out["mu"]), P(out["edge_i"]), P(out["edge_j"]))
    out["n_edges"] = E
    return 0, out


def grad_accum(embd, csr, next_sample, epoch, seed, neg_rate, params, dens=None,
               with_abs=False, exact=False):
    """dens = dict(re, phi_sum, weight, coeff) or None; returns (grad, has_update),
    plus, with with_abs, the per-coordinate sum of |term| of each gradient.
    exact: the same f32 terms summed in float64 (rounded once at the end)."""
    embd = f32(embd)
    n, dim = embd.shape
    g = np.zeros((n, dim), np.float32)
    hu = np.zeros(n, np.uint32)
    ab = np.zeros((n, dim), np.float32) if with_abs else None
    p = params_from(params)
    if dens is None:
        re = phi = w = None
        coeff = 0.0
    else:
        re, phi, w = f32(dens["re"]), f32(dens["phi_sum"]), f32(dens["weight"])
        coeff = float(dens["coeff"])
    fn = lib().orc_grad_accum_exact if exact else lib().orc_grad_accum
    fn(P(embd), C.c_int64(n), C.c_int32(dim), P(csr["off"]),
                         P(csr["edge_idx"]), P(csr["other"]), P(f32(next_sample)),
                         C.c_int32(epoch), C.c_uint32(seed), C.c_int32(neg_rate), C.byref(p),
                         P(csr["mu"]), P(re), P(phi), P(w), C.c_float(coeff), P(g), P(hu),
                         P(ab))
    return (g, hu, ab) if with_abs else (g, hu)


def adam_update(grad, has_update, m, v, embd, params, epoch, n_epochs):
    grad, m, v, embd = f32(grad).copy(), f32(m).copy(), f32(v).copy(), f32(embd).copy()
    hu = np.ascontiguousarray(has_update, dtype=np.uint32)
    n = hu.size
    dim = embd.size // n
    p = params_from(params)
    lib().orc_adam_update(P(grad), P(hu), P(m), P(v), P(embd), C.c_int64(n), C.c_int32(dim),
                          C.byref(p), C.c_int32(epoch), C.c_int32(n_epochs))
    return embd, m, v


def schedule_update(per, nxt, epoch):
    nxt = f32(nxt).copy()
    per = f32(per)
    lib().orc_schedule_update(P(per), P(nxt), C.c_int64(per.size), C.c_int32(epoch))
    return nxt


def dens_radii(embd, csr, a, b):
    embd = f32(embd)
    n, dim = embd.shape
    re = np.zeros(n, np.float32)
    phi = np.zeros(n, np.float32)
    lib().orc_dens_radii(P(embd), C.c_int64(n), C.c_int32(dim), P(csr["off"]), P(csr["other"]),
                         C.c_float(a), C.c_float(b), P(re), P(phi))
    return re, phi


def correlation_stats(re, r, var_shift):
    out = np.zeros(3, np.float64)
    re, r = f32(re), f32(r)
    lib().orc_correlation_stats(P(re), P(r), C.c_int64(re.size), C.c_float(var_shift), P(out))
    return out


def dens_weights(r, re, var_shift, lam, mu_tot):
    r, re = f32(r), f32(re)
    w = np.zeros_like(r)
    coeff = lib().orc_dens_weights(P(r), P(re), C.c_int64(r.size), C.c_float(var_shift),
                                   C.c_float(lam), C.c_float(mu_tot), P(w))
    return w, coeff


def dens_is_active(lam, frac, epoch, n_epochs) -> bool:
    return bool(lib().orc_dens_is_active(lam, frac, epoch, n_epochs))


def optimise_adam_gpu(embd, adj_off, adj_cols, adj_vals, params, seed, dens_r=None,
                      dens_params=None, mu_tot=0.0):
    out = f32(embd).copy()
    n, dim = out.shape
    p = params_from(params)
    adj_off = np.ascontiguousarray(adj_off, dtype=np.int64)
    lam, frac, vs = (dens_params if dens_params is not None else (0.0, 0.0, 0.0))
    st = lib().orc_optimise_adam_gpu(P(out), C.c_int64(n), C.c_int32(dim), P(adj_off),
                                     P(i32(adj_cols)), P(f32(adj_vals)), C.byref(p),
                                     P(f32(dens_r)) if dens_r is not None else None,
                                     C.c_float(lam), C.c_float(frac), C.c_float(vs),
                                     C.c_float(mu_tot), C.c_uint64(seed))
    if st:
        raise RuntimeError(f"oracle status {st}")
    return out


def smooth_knn_dist(dist, k, local_connectivity=1.0, bandwidth=1e-5, n_iter=64):
    dist = f32(dist)
    n = dist.shape[0]
    sigma = np.zeros(n, np.float32)
    rho = np.zeros(n, np.float32)
    lib().orc_smooth_knn_dist(P(dist), C.c_int64(n), C.c_int32(k), C.c_float(local_connectivity),
                              C.c_float(bandwidth), C.c_int32(n_iter), P(sigma), P(rho))
    return sigma, rho


def knn_to_coo(idx, dist, sigma, rho):
    idx, dist = i32(idx), f32(dist)
    n, k = idx.shape
    rows = np.zeros(n * k, np.int32)
    cols = np.zeros(n * k, np.int32)
    vals = np.zeros(n * k, np.float32)
    m = lib().orc_knn_to_coo(P(idx), P(dist), C.c_int64(n), C.c_int32(k), P(f32(sigma)),
                             P(f32(rho)), P(rows), P(cols), P(vals))
    return rows[:m], cols[:m], vals[:m]


def symmetrise(n, rows, cols, vals, mix_weight=1.0):
    rows, cols, vals = i32(rows), i32(cols), f32(vals)
    nnz = rows.size
    orow = np.zeros(2 * nnz, np.int32)
    ocol = np.zeros(2 * nnz, np.int32)
    oval = np.zeros(2 * nnz, np.float32)
    m = lib().orc_symmetrise(C.c_int64(n), P(rows), P(cols), P(vals), C.c_int64(nnz),
                             C.c_float(mix_weight), P(orow), P(ocol), P(oval))
    return orow[:m], ocol[:m], oval[:m]


def filter_weak_edges(rows, cols, vals, n_epochs):
    rows, cols, vals = i32(rows).copy(), i32(cols).copy(), f32(vals).copy()
    m = lib().orc_filter_weak_edges(P(rows), P(cols), P(vals), C.c_int64(rows.size),
                                    C.c_int32(n_epochs))
    return rows[:m], cols[:m], vals[:m]


def knn_exact(x, k, q0=0, q1=None, f64=False):
    x = f32(x)
    n, d = x.shape
    q1 = n if q1 is None else q1
    idx = np.zeros((q1 - q0, k), np.int32)
    if f64:
        dist = np.zeros((q1 - q0, k), np.float64)
        lib().orc_knn_exact_f64(P(x), C.c_int64(n), C.c_int32(d), C.c_int32(k), C.c_int64(q0),
                                C.c_int64(q1), P(idx), P(dist))
    else:
        dist = np.zeros((q1 - q0, k), np.float32)
        lib().orc_knn_exact_f32(P(x), C.c_int64(n), C.c_int32(d), C.c_int32(k), C.c_int64(q0),
                                C.c_int64(q1), P(idx), P(dist))
    return idx, dist


def knn_exact_rows_f64(x, k, rows):
    """FP64 brute-force kNN (self excluded, ties by index) for the given query rows."""
    x = f32(x)
    n, d = x.shape
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    idx = np.zeros((rows.size, k), np.int32)
    dist = np.zeros((rows.size, k), np.float64)
    lib().orc_knn_exact_f64_rows(P(x), C.c_int64(n), C.c_int32(d), C.c_int32(k), P(rows),
                                 C.c_int64(rows.size), P(idx), P(dist))
    return idx, dist


def dens_state(n, rows, cols, vals, knn_idx, knn_dist, distances_squared=True):
    rows, cols, vals = i32(rows), i32(cols), f32(vals)
    knn_idx, knn_dist = i32(knn_idx), f32(knn_dist)
    k = knn_idx.shape[1]
    r = np.zeros(n, np.float32)
    ms = np.zeros(n, np.float32)
    tot = C.c_float()
    st = lib().orc_dens_state(C.c_int64(n), P(rows), P(cols), P(vals), C.c_int64(rows.size),
                              P(knn_idx), P(knn_dist), C.c_int32(k),
                              C.c_int32(1 if distances_squared else 0), P(r), P(ms), C.byref(tot))
    return st, r, ms, tot.value


def fit_params(min_dist, spread, n_iter=300):
    a, b = C.c_float(), C.c_float()
    lib().orc_fit_params(C.c_float(min_dist), C.c_float(spread), C.c_int32(n_iter), C.byref(a),
                         C.byref(b))
    return a.value, b.value


def fit_params_grad(min_dist, spread, a, b):
    """One iteration's normalised (grad_a, grad_b) of fit_params at (a, b)."""
    ga, gb = C.c_float(), C.c_float()
    lib().orc_fit_params_grad(C.c_float(min_dist), C.c_float(spread), C.c_float(a), C.c_float(b),
                              C.byref(ga), C.byref(gb))
    return ga.value, gb.value


def lut_pow(b, x):
    """The parallel optimiser's FastPowLut(b, 25, 65 536).get over x."""
    x = f32(x).ravel()
    out = np.zeros_like(x)
    lib().orc_lut_pow(C.c_float(b), P(x), C.c_int64(x.size), P(out))
    return out


def density_edge_coeff(a, b, dist_sq, dist_sq_b, denom, phi_i, phi_o, re_i, re_o, w_i, w_o,
                       coeff, mu_edge):
    fn = lib().orc_density_edge_coeff
    fn.restype = C.c_float
    fn.argtypes = [C.c_float] * 13
    return fn(a, b, dist_sq, dist_sq_b, denom, phi_i, phi_o, re_i, re_o, w_i, w_o, coeff, mu_edge)


def optimise_adam_parallel(embd, adj_off, adj_cols, adj_vals, params, seed, use_lut=True,
                           n_threads=0, dens_r=None, dens_params=None, mu_tot=0.0):
    out = f32(embd).copy()
    n, dim = out.shape
    p = params_from(params)
    adj_off = np.ascontiguousarray(adj_off, dtype=np.int64)
    lam, frac, vs = (dens_params if dens_params is not None else (0.0, 0.0, 0.0))
    st = lib().orc_optimise_adam_parallel_dens(
        P(out), C.c_int64(n), C.c_int32(dim), P(adj_off), P(i32(adj_cols)), P(f32(adj_vals)),
        C.byref(p), C.c_uint64(seed), C.c_int32(1 if use_lut else 0), C.c_int32(n_threads),
        P(f32(dens_r)) if dens_r is not None else None, C.c_float(lam), C.c_float(frac),
        C.c_float(vs), C.c_float(mu_tot))
    if st:
        raise RuntimeError(f"oracle status {st}")
    return out


def optimise_sgd(embd, adj_off, adj_cols, adj_vals, params, seed):
    out = f32(embd).copy()
    n, dim = out.shape
    p = params_from(params)
    adj_off = np.ascontiguousarray(adj_off, dtype=np.int64)
    st = lib().orc_optimise_sgd(P(out), C.c_int64(n), C.c_int32(dim), P(adj_off),
                                P(i32(adj_cols)), P(f32(adj_vals)), C.byref(p), C.c_uint64(seed))
    if st:
        raise RuntimeError(f"oracle status {st}")
    return out


def spectral_layout(n, rows, cols, vals, n_comp=2, seed=42, range_=10.0, data=None, align=None):
    """spectral_layout of a symmetric COO graph: [n, n_comp] float32.  align
    ([n, n_comp], optional): another layout of the graph whose eigenvector
    signs the oracle adopts (each component's and the meta layout's)."""
    rows, cols, vals = i32(rows), i32(cols), f32(vals)
    out = np.zeros((n, n_comp), np.float32)
    d = 0
    if data is not None:
        data = f32(data)
        d = data.shape[1]
    if align is not None:
        align = f32(align)
        assert align.shape == (n, n_comp)
    st = lib().orc_spectral_layout_aligned(C.c_int64(n), P(rows), P(cols), P(vals),
                                           C.c_int64(rows.size), C.c_int32(n_comp),
                                           C.c_uint64(seed), C.c_float(range_), P(data),
                                           C.c_int32(d), P(align), P(out))
    if st:
        raise RuntimeError(f"oracle status {st}")
    return out


def pca_layout(x, n_comp=2, range_=1.0):
    """pca_layout of x [n, d]: [n, n_comp] float32 (exact leading components)."""
    x = f32(x)
    n, d = x.shape
    out = np.zeros((n, n_comp), np.float32)
    st = lib().orc_pca_layout(P(x), C.c_int64(n), C.c_int32(d), C.c_int32(n_comp),
                              C.c_float(range_), P(out))
    if st:
        raise RuntimeError(f"oracle status {st}")
    return out
