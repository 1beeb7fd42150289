"""Host mirror of the reference's GPU estimation pipeline, driven entirely through the
fine-grained drop-in C-ABI (include/fe_gpu_runtime.h).

It reproduces the call sequence of the Fortran host so the ABI can be validated the way the
reference uses it:
  * fe_gpu_dataset_upload            src/gpu/fe_gpu_data.f90:41-120
  * fe_gpu_within_transform          src/gpu/fe_gpu_demean.f90:63-311 (relaxed symmetric AP)
  * fe_gpu_compute_cross_products    src/gpu/fe_gpu_linalg.f90:106-127
  * build_iv_normal_equations        src/core/fe_pipeline.f90:947-1032
  * solve_with_column_filter         src/core/fe_pipeline.f90:1034-1107 (LAPACK dpotrf/dpotrs/dpotri)
  * compute_standard_errors          src/core/fe_pipeline.f90:1109-1420
  * finalize_regression_stats        src/core/fe_pipeline.f90 (end of module)
Cluster intersections use the device builder (the reference's --fast route), whose numbering is
the CPU builder's, so the result is that of the reference's standard mode.  Host K x K algebra
uses LAPACK through scipy, as the reference uses LAPACK.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
from scipy.linalg import lapack

from ._lib import check, load
from .device import DeviceBuffer

F8 = 8
I4 = 4


@dataclass
class FeDim:
    ids: DeviceBuffer
    n_groups: int
    mean_y: DeviceBuffer
    mean_W: DeviceBuffer
    mean_Z: DeviceBuffer
    counts: DeviceBuffer


@dataclass
class GpuDataset:
    n_obs: int
    n_reg: int
    n_inst: int
    d_y: DeviceBuffer
    d_W: DeviceBuffer
    d_Z: DeviceBuffer
    dims: list = field(default_factory=list)
    d_proj_W: DeviceBuffer = None


def dataset_upload(y, W, Z, fe, G) -> GpuDataset:
    """fe_gpu_dataset_upload (src/gpu/fe_gpu_data.f90:41-120)."""
    n = len(y)
    K = W.shape[1]
    L = Z.shape[1] if Z is not None else 0
    ds = GpuDataset(n, K, L, DeviceBuffer.from_array(np.ascontiguousarray(y, dtype=np.float64)),
                    DeviceBuffer.from_array(np.asfortranarray(W, dtype=np.float64)),
                    DeviceBuffer.from_array(np.asfortranarray(Z, dtype=np.float64)) if L else DeviceBuffer(0))
    for d in range(fe.shape[0]):
        g = int(G[d])
        dim = FeDim(DeviceBuffer.from_array(np.ascontiguousarray(fe[d], dtype=np.int32)), g, DeviceBuffer(g * F8),
                    DeviceBuffer(g * K * F8), DeviceBuffer(g * L * F8) if L else DeviceBuffer(0),
                    DeviceBuffer(g * I4))
        for b in (dim.mean_y, dim.mean_W, dim.mean_Z, dim.counts):
            b.zero()
        ds.dims.append(dim)
    return ds


def within_transform(ds: GpuDataset, tol: float, max_iter: int):
    """fe_gpu_within_transform (src/gpu/fe_gpu_demean.f90:63-311): relaxed symmetric AP."""
    lib = load()
    n_fe = len(ds.dims)
    K, L, n = ds.n_reg, ds.n_inst, ds.n_obs
    relax = [1.5] * n_fe
    prev_change = [float(np.finfo(np.float64).max)] * n_fe  # huge(1.0d0); 1.05x of it is +inf, as in Fortran
    order = list(range(n_fe))
    for a in range(n_fe - 1):                       # exchange sort by ascending group count
        for b in range(a + 1, n_fe):
            if ds.dims[order[b]].n_groups < ds.dims[order[a]].n_groups:
                order[a], order[b] = order[b], order[a]
    prev = []
    for dim in ds.dims:
        py = DeviceBuffer(dim.n_groups * F8)
        pw = DeviceBuffer(dim.n_groups * K * F8)
        pz = DeviceBuffer(dim.n_groups * L * F8)
        for b in (py, pw, pz):
            b.zero()
        prev.append((py, pw, pz))
    if n_fe == 0 or n == 0:
        return True, 0, 0.0
    out = C.c_double()

    def absmax(buf, length):
        if length <= 0:
            return 0.0
        check(lib.fe_gpu_absmax(buf.ptr, length, C.byref(out)), "absmax")
        return out.value

    converged, iterations, max_update = False, 0, 0.0
    for it in range(1, max_iter + 1):
        max_update = 0.0
        for _sweep in range(2):
            for d in order + order[::-1]:
                dim = ds.dims[d]
                G = dim.n_groups
                for b in (dim.mean_y, dim.mean_W, dim.counts) + ((dim.mean_Z,) if L else ()):
                    b.zero()
                check(lib.fe_gpu_fe_accumulate(ds.d_y.ptr, ds.d_W.ptr, ds.d_Z.ptr, dim.ids.ptr, n, G, K, L, n,
                                               dim.mean_y.ptr, dim.mean_W.ptr, dim.mean_Z.ptr, dim.counts.ptr),
                      "accumulating FE statistics")
                check(lib.fe_gpu_fe_compute_means(dim.mean_y.ptr, dim.mean_W.ptr, dim.mean_Z.ptr, dim.counts.ptr, G, K,
                                                  L), "computing FE group means")
                py, pw, pz = prev[d]
                check(lib.fe_gpu_mix_means(dim.mean_y.ptr, py.ptr, G, relax[d]), "relaxing FE means (y)")
                if K > 0:
                    check(lib.fe_gpu_mix_means(dim.mean_W.ptr, pw.ptr, G * K, relax[d]), "relaxing FE means (W)")
                if L > 0:
                    check(lib.fe_gpu_mix_means(dim.mean_Z.ptr, pz.ptr, G * L, relax[d]), "relaxing FE means (Z)")
                check(lib.fe_gpu_fe_subtract(ds.d_y.ptr, ds.d_W.ptr, ds.d_Z.ptr, dim.ids.ptr, n, K, L, n,
                                             dim.mean_y.ptr, dim.mean_W.ptr, dim.mean_Z.ptr, relax[d]),
                      "subtracting FE means")
                change_dim = absmax(dim.mean_y, G)
                max_update = max(max_update, change_dim)
                if K > 0:
                    c = absmax(dim.mean_W, G * K)
                    change_dim = max(change_dim, c)
                    max_update = max(max_update, c)
                if L > 0:
                    c = absmax(dim.mean_Z, G * L)
                    change_dim = max(change_dim, c)
                    max_update = max(max_update, c)
                if change_dim > prev_change[d] * 1.05:
                    relax[d] = max(0.4, relax[d] * 0.5)
                elif change_dim < prev_change[d] * 0.8:
                    relax[d] = min(2.0, relax[d] * 1.10)
                prev_change[d] = change_dim
        iterations = it
        if max_update < tol:
            converged = True
            break
    return converged, iterations, max_update


def _chol_solve_invert(Q, b):
    """chol_solve_and_invert (src/core/fe_solver.f90:41-74): dpotrf/dpotrs/dpotri, upper."""
    c, info = lapack.dpotrf(np.asfortranarray(Q), lower=0)
    if info != 0:
        return info, None, None
    x, info = lapack.dpotrs(c, np.asarray(b, dtype=np.float64), lower=0)
    if info != 0:
        return info, None, None
    inv, info = lapack.dpotri(c, lower=0)
    if info != 0:
        return info, None, None
    inv = np.triu(inv) + np.triu(inv, 1).T
    return 0, x, inv


def fe_gpu_estimate(y, W, fe, G=None, Z=None, cluster=(), iv_reg=(), iv_inst=(), tol=1e-8, max_iter=500):
    """Estimation through the fine-grained ABI; returns (beta, se, vcov, result dict)."""
    lib = load()
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = len(y)
    W = np.asarray(W, dtype=np.float64).reshape(n, -1)
    K = W.shape[1]
    Z = None if Z is None else np.asarray(Z, dtype=np.float64).reshape(n, -1)
    L = 0 if Z is None else Z.shape[1]
    fe = np.ascontiguousarray(fe, dtype=np.int32).reshape(-1, n)
    if G is None:
        if fe.size and fe.min() < 1:
            raise ValueError("FE indices must be positive and contiguous")
        G = fe.max(axis=1).astype(np.int32)
    n_fe = fe.shape[0]
    res = {}
    cdims = []
    for d in cluster:                                   # initialize_cluster_dims
        if d < 1 or d > n_fe or d in cdims:
            continue
        if len(cdims) == 3:
            break
        cdims.append(int(d))
    yc = y - y.mean()
    res["tss_total"] = float(yc @ yc)
    dof_fes = int(sum(max(0, int(g) - 1) for g in G))
    dof_nested = int(sum(max(0, int(G[d - 1]) - 1) for d in cdims))
    res["dof_fes"] = dof_fes
    ds = dataset_upload(y, W, Z, fe, G)
    check(lib.fe_gpu_linalg_init(), "linalg init")
    conv, iters, mu = within_transform(ds, tol, max_iter)
    res.update(converged=conv, fe_iterations=iters, max_update=mu)
    beta = np.zeros(K)
    se = np.zeros(K)
    vcov = np.zeros((K, K))
    if K <= 0:
        return beta, se, vcov, res
    dot = C.c_double()
    check(lib.fe_gpu_dot(n, ds.d_y.ptr, ds.d_y.ptr, C.byref(dot)), "dot")
    res["tss_within"] = dot.value

    def cross_products(d_yv, d_X, k):
        dQ, db = DeviceBuffer(k * k * F8), DeviceBuffer(k * F8)
        check(lib.fe_gpu_syrk(n, k, 1.0, d_X.ptr if hasattr(d_X, "ptr") else d_X, n, 0.0, dQ.ptr), "syrk")
        check(lib.fe_gpu_gemv(n, k, 1.0, d_X.ptr if hasattr(d_X, "ptr") else d_X, n, d_yv.ptr, 0.0, db.ptr), "gemv")
        Qh = dQ.to_array(np.float64, (k, k), order="F")
        Qh = np.triu(Qh) + np.triu(Qh, 1).T                 # symmetrize_upper
        return Qh, db.to_array(np.float64, (k,))

    WW, Wy = cross_products(ds.d_y, ds.d_W, K)
    vdiag = np.diag(WW).copy()
    is_endog = np.zeros(K, bool)
    for v in iv_reg:
        if 1 <= v <= K:
            is_endog[v - 1] = True
    n_endog = int(is_endog.sum())
    exog = [k for k in range(K) if not is_endog[k]]
    inst = []
    if L > 0:
        inst = [v for v in iv_inst if 1 <= v <= L] or list(range(1, L + 1))
    use_iv = n_endog > 0 and len(inst) >= n_endog
    n_aug = len(exog) + len(inst)
    res.update(is_iv=int(use_iv), n_instruments=n_aug if use_iv else 0)
    if use_iv:
        d_Zaug = DeviceBuffer(n * n_aug * F8)
        if exog:
            idx = np.array([k + 1 for k in exog], dtype=np.int32)
            check(lib.fe_gpu_copy_columns(ds.d_W.ptr, n, idx.ctypes.data, len(idx), n, d_Zaug.ptr, n, 0), "copy")
        if inst:
            idx = np.array(inst, dtype=np.int32)
            check(lib.fe_gpu_copy_columns(ds.d_Z.ptr, n, idx.ctypes.data, len(idx), n, d_Zaug.ptr, n, len(exog)),
                  "copy")
        Qzz, Qzy = cross_products(ds.d_y, d_Zaug, n_aug)
        d_cross = DeviceBuffer(n_aug * K * F8)
        check(lib.fe_gpu_gemm(b"T", b"N", n_aug, K, n, 1.0, d_Zaug.ptr, n, ds.d_W.ptr, n, 0.0, d_cross.ptr, n_aug),
              "gemm")
        Qzx = d_cross.to_array(np.float64, (n_aug, K), order="F")
        info, rhs, Qzinv = _chol_solve_invert(Qzz, Qzy)
        if info != 0:
            res["solver_info"] = info
            return beta, se, vcov, res
        proj = Qzinv @ Qzx
        Q = Qzx.T @ proj
        Q = 0.5 * (Q + Q.T)
        bvec = Qzx.T @ rhs
        d_proj = DeviceBuffer.from_array(np.asfortranarray(proj))
        ds.d_proj_W = DeviceBuffer(n * K * F8)
        check(lib.fe_gpu_gemm(b"N", b"N", n, K, n_aug, 1.0, d_Zaug.ptr, n, d_proj.ptr, n_aug, 0.0, ds.d_proj_W.ptr, n),
              "gemm")
    else:
        Q, bvec = WW, Wy
    # solve_with_column_filter
    dmax = max(float(np.max(np.abs(vdiag))) if K else 0.0, 1.0)
    keep = [k for k in range(K) if not (abs(vdiag[k]) <= 1e-10 * dmax)]
    res["dof_model"] = len(keep)
    if not keep:
        res["solver_info"] = -1
        return beta, se, vcov, res
    info, bsel, Qinv = _chol_solve_invert(Q[np.ix_(keep, keep)], bvec[keep])
    res["solver_info"] = info
    if info != 0:
        return beta, se, vcov, res
    beta[keep] = bsel
    res["rank"] = len(keep)
    # compute_standard_errors
    d_beta = DeviceBuffer.from_array(beta)
    d_res = DeviceBuffer(n * F8)
    check(lib.fe_gpu_residual(n, K, ds.d_W.ptr, n, d_beta.ptr, ds.d_y.ptr, d_res.ptr), "residual")
    check(lib.fe_gpu_dot(n, d_res.ptr, d_res.ptr, C.byref(dot)), "dot")
    rss = dot.value
    df = max(1, n - (len(keep) + dof_fes + 1))
    res.update(rss=rss, dof_resid=df)
    sigma2 = rss / df
    kept = len(keep)
    if not cdims:
        cov = sigma2 * Qinv
    else:
        reg = ds.d_proj_W if use_iv else ds.d_W
        cov = np.zeros((kept, kept))
        d_ids = DeviceBuffer(n * I4)
        min_single = None
        ncl_list = [0] * 7
        for mask in range(1, 1 << len(cdims)):
            sub = [cdims[j] for j in range(len(cdims)) if mask >> j & 1]
            w = 1.0 if len(sub) % 2 == 1 else -1.0
            if len(sub) == 1:
                ncl = int(G[sub[0] - 1])
                min_single = ncl if min_single is None else min(min_single, ncl)
                ids_ptr = ds.dims[sub[0] - 1].ids.ptr
            else:
                strides = np.cumprod([1] + [int(G[d - 1]) for d in sub[:-1]]).astype(np.uint64)
                ptrs = (C.c_void_p * len(sub))(*[ds.dims[d - 1].ids.ptr for d in sub])
                nc = C.c_int()
                check(lib.fe_gpu_build_multi_cluster_ids(ptrs, strides.ctypes.data, len(sub), n, d_ids.ptr,
                                                         C.byref(nc)), "cluster ids")
                ncl = nc.value
                ids_ptr = d_ids.ptr
            ncl_list[mask - 1] = ncl
            d_scores = DeviceBuffer(ncl * K * F8)
            d_scores.zero()
            check(lib.fe_gpu_cluster_scores(d_res.ptr, reg.ptr, ids_ptr, n, K, n, ncl, d_scores.ptr), "scores")
            d_meat = DeviceBuffer(K * K * F8)
            check(lib.fe_gpu_cluster_meat(ncl, K, d_scores.ptr, ncl, d_meat.ptr), "meat")
            M = d_meat.to_array(np.float64, (K, K), order="F")
            M = np.triu(M) + np.triu(M, 1).T
            Mk = M[np.ix_(keep, keep)]
            cov += w * (Qinv @ Mk @ Qinv)
        nested_adj = 1 if dof_nested > 0 else 0
        scalar = max(1, n - 1) / max(1, n - (kept + nested_adj))
        if min_single > 1:
            scalar *= min_single / (min_single - 1)
        cov = cov * scalar
        cov = np.triu(cov) + np.triu(cov, 1).T
        lam, V = np.linalg.eigh(cov, UPLO="U")
        res["psd_fix_applied"] = int(np.any(lam < 0))
        lam = np.clip(lam, 0, None)
        cov = (V * lam) @ V.T
        res.update(dof_resid=max(1, min_single - 1), min_single_cluster=min_single, n_clusters=ncl_list)
    se[keep] = np.sqrt(np.clip(np.diag(cov), 0, None))
    vcov[np.ix_(keep, keep)] = cov
    _finalize_stats(res, n)
    return beta, se, vcov, res


def _finalize_stats(r, n_obs):
    EPS = 1e-12
    n = float(max(1, n_obs))
    dfm = float(max(0, r.get("dof_model", 0)))
    dfr = float(max(1, r.get("dof_resid", 1)))
    dev, dev0 = r["rss"], max(EPS, r["tss_total"])
    r["r2"] = max(0.0, min(1.0, 1 - dev / dev0)) if dev0 > EPS else 0.0
    r["r2_within"] = max(0.0, min(1.0, 1 - dev / r["tss_within"])) if r["tss_within"] > EPS else r["r2"]
    k_total = r.get("dof_model", 0) + r["dof_fes"] + 1
    denom = max(n - k_total, 1.0)
    r["r2_adj"] = max(0.0, min(1.0, 1 - (dev * (n - 1)) / (dev0 * denom))) if n > 1 and dev0 > EPS else r["r2"]
    ssr = max(0.0, dev0 - dev)
    r["f_stat"] = (ssr / dfm) / (dev / dfr) if dfm > 0 and dfr > 0 and dev > EPS else 0.0
