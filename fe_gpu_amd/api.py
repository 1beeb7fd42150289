"""Python front end of the coarse C-ABI (include/fehip.h): device-resident panels and
whole-regression calls.  Host arrays are numpy; device memory comes from the library."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _coarse_sigs as S
from ._lib import check, load
from .device import DeviceBuffer


class Panel:
    """fehip_panel: FE ids resident on the device, rows ordered by the largest FE dim."""

    def __init__(self, fe_dev_ptrs, n_obs):
        self.lib = load()
        arr = (C.c_void_p * len(fe_dev_ptrs))(*fe_dev_ptrs)
        h = C.c_void_p()
        check(self.lib.fehip_panel_create(arr, len(fe_dev_ptrs), n_obs, C.byref(h)), "fehip_panel_create")
        self.handle = h.value
        self.n_obs = n_obs

    @classmethod
    def from_host(cls, fe: np.ndarray) -> "Panel":
        fe = np.ascontiguousarray(fe, dtype=np.int32)
        bufs = [DeviceBuffer.from_array(fe[d]) for d in range(fe.shape[0])]
        p = cls([b.ptr for b in bufs], fe.shape[1])
        for b in bufs:
            b.free()
        return p

    def info(self) -> dict:
        inf = S.PanelInfo()
        check(self.lib.fehip_panel_get_info(self.handle, C.byref(inf)), "panel info")
        m = inf.n_fe
        return dict(n_obs=inf.n_obs, n_fe=m, primary_dim=inf.primary_dim, sorted_input=inf.sorted_input,
                    max_segment=inf.max_segment, group_sizes=list(inf.group_sizes)[:m],
                    n_runs=list(inf.n_runs)[:m], device_bytes=inf.device_bytes)

    def close(self):
        if self.handle:
            check(self.lib.fehip_panel_destroy(self.handle), "fehip_panel_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def demean_opts(tol=1e-8, max_iter=500, fe_mask=0, method=0, precision=0, check_every=4):
    return S.DemeanOpts(tol, max_iter, method, fe_mask, precision, check_every, 0)


def demean(panel: Panel, cols_ptr: int, n_cols: int, ld: int, weights_ptr: int = 0, **kw):
    st = S.DemeanStats()
    o = demean_opts(**kw)
    check(load().fehip_demean(panel.handle, cols_ptr, n_cols, ld, weights_ptr or None, C.byref(o), C.byref(st)),
          "fehip_demean")
    return dict(iterations=st.iterations, converged=st.converged, max_update=st.max_update, seconds=st.seconds,
                bytes_moved=st.bytes_moved)


@dataclass
class Fit:
    beta: np.ndarray
    se: np.ndarray
    vcov: np.ndarray
    res: dict


def estimate_device(panel: Panel, y_ptr: int, W_ptr: int, K: int, Z_ptr: int = 0, L: int = 0, ld: int = None,
                    weights_ptr: int = 0, cluster=(), iv_reg=(), iv_inst=(), tol=1e-8, max_iter=500, fe_mask=0,
                    method=0, precision=0, check_every=4) -> Fit:
    lib = load()
    ld = panel.n_obs if ld is None else ld
    opts = S.Opts()
    opts.demean = demean_opts(tol, max_iter, fe_mask, method, precision, check_every)
    cl = list(cluster)[:4]
    opts.n_cluster = len(cl)
    for i, d in enumerate(cl):
        opts.cluster_dims[i] = int(d)
    ivr = (S.I32 * max(1, len(iv_reg)))(*[int(v) for v in iv_reg])
    ivi = (S.I32 * max(1, len(iv_inst)))(*[int(v) for v in iv_inst])
    opts.n_iv_reg = len(iv_reg)
    opts.iv_reg = C.cast(ivr, C.POINTER(S.I32))
    opts.n_iv_inst = len(iv_inst)
    opts.iv_inst = C.cast(ivi, C.POINTER(S.I32))
    r = S.Result()
    beta = np.zeros(max(K, 1))
    se = np.zeros(max(K, 1))
    vcov = np.zeros((max(K, 1), max(K, 1)), order="F")
    dp = lambda a: a.ctypes.data_as(S.DP)  # noqa: E731
    check(lib.fehip_estimate(panel.handle, y_ptr, W_ptr or None, K, Z_ptr or None, L, ld, weights_ptr or None,
                             C.byref(opts), C.byref(r), dp(beta), dp(se), dp(vcov)), "fehip_estimate")
    res = {f: getattr(r, f) for f, _ in S.Result._fields_}
    res["cluster_dims"] = list(r.cluster_dims)[: r.n_cluster_dims]
    res["n_clusters"] = list(r.n_clusters)
    return Fit(beta[:K], se[:K], np.array(vcov)[:K, :K], res)


def estimate(y, W, fe, Z=None, weights=None, panel: Panel = None, **kw) -> Fit:
    """Host-array convenience: uploads y, W, Z (and weights), builds the panel if not given."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = len(y)
    W = np.asfortranarray(np.asarray(W, dtype=np.float64).reshape(n, -1))
    K = W.shape[1]
    pan = panel or Panel.from_host(np.asarray(fe, dtype=np.int32).reshape(-1, n))
    dy = DeviceBuffer.from_array(y)
    dW = DeviceBuffer.from_array(W)
    dZ, L = None, 0
    if Z is not None:
        Zf = np.asfortranarray(np.asarray(Z, dtype=np.float64).reshape(n, -1))
        L = Zf.shape[1]
        dZ = DeviceBuffer.from_array(Zf)
    dw = DeviceBuffer.from_array(np.ascontiguousarray(weights, dtype=np.uint8)) if weights is not None else None
    fit = estimate_device(pan, dy.ptr, dW.ptr, K, dZ.ptr if dZ else 0, L, n, dw.ptr if dw else 0, **kw)
    if panel is None:
        pan.close()
    return fit
