"""CPU oracle for the fixed-effects estimation core -- TEST INFRASTRUCTURE ONLY.

This package is the checker the HIP product path is compared against.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import it; the product package ``fe_gpu_amd`` never does.

The C library ``libfe_oracle.so`` (built by ``make -C oracle``) restates the
reference's algorithms (see ``fe_oracle.c`` for file:line citations).  The
reference's own self-contained intersection-id builder
(``src/core/cluster_id_builder.c``) is compiled from the reference tree into
``oracle/_ref/libref_cluster_ids.so`` when that tree is present; it is used
only to pin the restatement.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

I64 = C.c_int64
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int32)


class FoOptions(C.Structure):
    _fields_ = [("conv", I64), ("ap_kind", I64), ("max_iter", I64), ("tol", C.c_double)]


class FoResult(C.Structure):
    _fields_ = [
        ("converged", I64), ("fe_iterations", I64), ("solver_info", I64), ("rank", I64),
        ("dof_model", I64), ("dof_fes", I64), ("dof_resid", I64), ("is_iv", I64),
        ("n_instruments", I64), ("n_cluster_dims", I64), ("cluster_dims", I64 * 3),
        ("n_clusters", I64 * 7), ("min_single_cluster", I64), ("psd_fix_applied", I64),
        ("rss", C.c_double), ("tss_total", C.c_double), ("tss_within", C.c_double),
        ("r2", C.c_double), ("r2_adj", C.c_double), ("r2_within", C.c_double),
        ("f_stat", C.c_double), ("max_update", C.c_double),
    ]


def lib():
    """Load (building on first use if the sources are newer) the oracle library."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libfe_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = C.CDLL(path)
        L.fo_group_sizes.argtypes = [IP, I64, I64, IP]
        L.fo_build_cluster_ids.argtypes = [IP, C.c_int, C.c_longlong, IP, IP, C.c_int, IP, IP]
        L.fo_cluster_ids_dimmajor.argtypes = [IP, I64, IP, IP, C.c_int, IP, IP]
        L.fo_demean_forward.argtypes = [DP, DP, I64, I64, IP, I64, IP, C.c_double, I64,
                                        C.POINTER(I64), DP]
        L.fo_demean_forward.restype = I64
        L.fo_demean_relaxed.argtypes = [DP, DP, DP, I64, I64, I64, IP, I64, IP, C.c_double, I64,
                                        C.POINTER(I64), DP]
        L.fo_demean_relaxed.restype = I64
        L.fo_chol_solve_and_invert.argtypes = [DP, DP, I64, DP, DP]
        L.fo_sym_eig.argtypes = [DP, I64, DP, DP]
        L.fo_estimate.argtypes = [I64, I64, I64, I64, DP, DP, DP, IP, IP, C.POINTER(FoOptions),
                                  IP, I64, IP, I64, IP, I64, DP, DP, DP, C.POINTER(FoResult)]
        L.fo_num_threads.restype = C.c_int
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own cluster_id_builder.c compiled from source, or None if not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_cluster_ids.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.fe_build_cluster_ids.argtypes = [IP, C.c_int, C.c_longlong, IP, IP, C.c_int, IP, IP]
        _REF = L
    return _REF


def _dp(a):
    return a.ctypes.data_as(DP)


def _ip(a):
    return a.ctypes.data_as(IP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _fortran2d(a, n):
    """n x K column-major copy (K may be 0)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(n, -1)
    return np.asfortranarray(a.reshape(n, -1)).copy(order="F")


def group_sizes(fe):
    """fe: (n_fe, n) int32.  Returns (status, G)."""
    fe = np.ascontiguousarray(fe, dtype=np.int32)
    G = np.zeros(fe.shape[0], dtype=np.int32)
    st = lib().fo_group_sizes(_ip(fe), fe.shape[0], fe.shape[1], _ip(G))
    return st, G


def cluster_ids(fe, subset_dims, subset_sizes):
    """Intersection ids for 1-based ``subset_dims`` of dim-major ``fe``; returns (status, ids, n)."""
    fe = np.ascontiguousarray(fe, dtype=np.int32)
    sd = np.ascontiguousarray(subset_dims, dtype=np.int32)
    ss = np.ascontiguousarray(subset_sizes, dtype=np.int32)
    out = np.zeros(fe.shape[1], dtype=np.int32)
    nc = np.zeros(1, dtype=np.int32)
    st = lib().fo_cluster_ids_dimmajor(_ip(fe), fe.shape[1], _ip(sd), _ip(ss), len(sd), _ip(out), _ip(nc))
    return st, out, int(nc[0])


def cluster_ids_interleaved(fe_il, subset_dims, subset_sizes, use_ref=False):
    """Same numbering through the reference signature (fe interleaved (n_fe, n) Fortran order)."""
    fe_il = np.ascontiguousarray(fe_il, dtype=np.int32)  # shape (n, ld_fe): row i holds its ids
    n, ld = fe_il.shape
    sd = np.ascontiguousarray(subset_dims, dtype=np.int32)
    ss = np.ascontiguousarray(subset_sizes, dtype=np.int32)
    out = np.zeros(n, dtype=np.int32)
    nc = np.zeros(1, dtype=np.int32)
    if use_ref:
        L = ref_lib()
        st = L.fe_build_cluster_ids(_ip(fe_il), ld, n, _ip(sd), _ip(ss), len(sd), _ip(out), _ip(nc))
    else:
        st = lib().fo_build_cluster_ids(_ip(fe_il), ld, n, _ip(sd), _ip(ss), len(sd), _ip(out), _ip(nc))
    return st, out, int(nc[0])


def demean_forward(y, W, fe, G, tol=1e-8, max_iter=500):
    n = len(y)
    y = _f64(y).copy()
    Wf = _fortran2d(W, n)
    fe = np.ascontiguousarray(fe, dtype=np.int32)
    G = np.ascontiguousarray(G, dtype=np.int32)
    conv = I64(0)
    mu = C.c_double(0)
    it = lib().fo_demean_forward(_dp(y), Wf.ctypes.data_as(DP), n, Wf.shape[1], _ip(fe), fe.shape[0], _ip(G),
                                 tol, max_iter, C.byref(conv), C.byref(mu))
    return y, Wf, int(it), bool(conv.value), mu.value


def demean_relaxed(y, W, Z, fe, G, tol=1e-8, max_iter=500):
    n = len(y)
    y = _f64(y).copy()
    Wf = _fortran2d(W, n)
    Zf = _fortran2d(Z if Z is not None else np.zeros((n, 0)), n)
    fe = np.ascontiguousarray(fe, dtype=np.int32)
    G = np.ascontiguousarray(G, dtype=np.int32)
    conv = I64(0)
    mu = C.c_double(0)
    it = lib().fo_demean_relaxed(_dp(y), Wf.ctypes.data_as(DP), Zf.ctypes.data_as(DP), n, Wf.shape[1],
                                 Zf.shape[1], _ip(fe), fe.shape[0], _ip(G), tol, max_iter, C.byref(conv),
                                 C.byref(mu))
    return y, Wf, Zf, int(it), bool(conv.value), mu.value


def chol_solve_and_invert(Q, b):
    Q = np.asfortranarray(Q, dtype=np.float64)
    b = _f64(b)
    n = len(b)
    beta = np.zeros(n)
    Qi = np.zeros((n, n), order="F")
    info = lib().fo_chol_solve_and_invert(Q.ctypes.data_as(DP), _dp(b), n, _dp(beta), Qi.ctypes.data_as(DP))
    return info, beta, Qi


def sym_eig(A):
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.zeros(n)
    V = np.zeros((n, n), order="F")
    lib().fo_sym_eig(A.ctypes.data_as(DP), n, _dp(w), V.ctypes.data_as(DP))
    return w, V


@dataclass
class OracleFit:
    beta: np.ndarray
    se: np.ndarray
    vcov: np.ndarray
    res: dict
    y_tilde: np.ndarray = field(repr=False)
    W_tilde: np.ndarray = field(repr=False)
    Z_tilde: np.ndarray = field(repr=False)


def estimate(y, W, fe, G=None, Z=None, cluster=(), iv_reg=(), iv_inst=(), tol=1e-8, max_iter=500,
             conv="gpu", ap="relaxed"):
    """Full estimation.  conv='gpu' restates fe_gpu_estimate, conv='cpu' fe_cpu_estimate."""
    n = len(y)
    y = _f64(y).copy()
    Wf = _fortran2d(W, n)
    Zf = _fortran2d(Z if Z is not None else np.zeros((n, 0)), n)
    fe = np.ascontiguousarray(fe, dtype=np.int32).reshape(-1, n)
    if G is None:
        st, G = group_sizes(fe)
        if st:
            raise ValueError("FE indices must be positive")
    G = np.ascontiguousarray(G, dtype=np.int32)
    K, L = Wf.shape[1], Zf.shape[1]
    opt = FoOptions(0 if conv == "gpu" else 1, 0 if ap == "relaxed" else 1, int(max_iter), float(tol))
    cl = np.ascontiguousarray(cluster, dtype=np.int32)
    ivr = np.ascontiguousarray(iv_reg, dtype=np.int32)
    ivi = np.ascontiguousarray(iv_inst, dtype=np.int32)
    beta = np.zeros(K)
    se = np.zeros(K)
    vcov = np.zeros((K, K), order="F")
    r = FoResult()
    st = lib().fo_estimate(n, K, L, fe.shape[0], _dp(y), Wf.ctypes.data_as(DP), Zf.ctypes.data_as(DP), _ip(fe),
                           _ip(G), C.byref(opt), _ip(cl), len(cl), _ip(ivr), len(ivr), _ip(ivi), len(ivi),
                           _dp(beta), _dp(se), vcov.ctypes.data_as(DP), C.byref(r))
    if st:
        raise RuntimeError(f"oracle estimate failed: {st}")
    d = {f: getattr(r, f) for f, _ in FoResult._fields_}
    d["cluster_dims"] = list(r.cluster_dims)[: r.n_cluster_dims]
    d["n_clusters"] = list(r.n_clusters)
    return OracleFit(beta, se, np.array(vcov), d, y, Wf, Zf)


def reconstruct_fe(y, W, beta, fe, tol=1e-8, max_iter=10000):
    """reconstruct_and_save_fixed_effects (src/core/fe_pipeline.f90:1619-1715), restated in numpy:
    r = y - W beta (:1653-1659); level counts over every row (:1666-1673); then forward alternating
    projection over the dims: group means of the current residual (:1680-1689), centred by their
    row-weighted mean (:1691-1704), added to the dim's values and subtracted from the residual
    (:1706-1717), until the largest update is below tol (:1718).  Returns [(values, counts)] per dim
    (level id - 1 indexes both; levels without rows have value 0) and the iteration count."""
    fe = np.asarray(fe)
    r = np.asarray(y, dtype=np.float64).copy()
    if W is not None and len(beta):
        r -= np.asarray(W, dtype=np.float64) @ np.asarray(beta, dtype=np.float64)
    m = fe.shape[0]
    G = [int(fe[d].max()) for d in range(m)]
    idx = [fe[d].astype(np.int64) - 1 for d in range(m)]
    cnt = [np.bincount(idx[d], minlength=G[d]) for d in range(m)]
    val = [np.zeros(G[d]) for d in range(m)]
    it = 0
    for it in range(1, max_iter + 1):
        mx = 0.0
        for d in range(m):
            s = np.bincount(idx[d], weights=r, minlength=G[d])
            has = cnt[d] > 0
            mean = np.where(has, s / np.maximum(cnt[d], 1), 0.0)
            tot = cnt[d].sum()
            center = (mean * cnt[d]).sum() / tot if tot > 0 else 0.0
            mean = np.where(has, mean - center, 0.0)
            val[d] += mean
            if has.any():
                mx = max(mx, float(np.abs(mean[has]).max()))
            r -= mean[idx[d]]
        if mx < tol:
            break
    return [(val[d], cnt[d]) for d in range(m)], it
