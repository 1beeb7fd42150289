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

    def register_columns(self, cols_ptr: int, n_cols: int, ld: int = None):
        """fehip_panel_register_columns: keep FP32 copies of n_cols FP64 device columns (caller row order) for
        the FP32-storage requests that read them; the columns must not change while registered."""
        ld = self.n_obs if ld is None else ld
        check(self.lib.fehip_panel_register_columns(self.handle, cols_ptr, n_cols, ld), "fehip_panel_register_columns")

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


class _Spec:
    """Host-side storage of one request's options and outputs (kept alive across the C call)."""

    def __init__(self, K, cluster=(), iv_reg=(), iv_inst=(), tol=1e-8, max_iter=500, fe_mask=0, method=0,
                 precision=0, check_every=4):
        self.K = K
        self.opts = S.Opts()
        self.opts.demean = demean_opts(tol, max_iter, fe_mask, method, precision, check_every)
        cl = list(cluster)[:4]
        self.opts.n_cluster = len(cl)
        for i, d in enumerate(cl):
            self.opts.cluster_dims[i] = int(d)
        self._ivr = (S.I32 * max(1, len(iv_reg)))(*[int(v) for v in iv_reg])
        self._ivi = (S.I32 * max(1, len(iv_inst)))(*[int(v) for v in iv_inst])
        self.opts.n_iv_reg = len(iv_reg)
        self.opts.iv_reg = C.cast(self._ivr, C.POINTER(S.I32))
        self.opts.n_iv_inst = len(iv_inst)
        self.opts.iv_inst = C.cast(self._ivi, C.POINTER(S.I32))
        self.r = S.Result()
        k1 = max(K, 1)
        self.beta = np.zeros(k1)
        self.se = np.zeros(k1)
        self.vcov = np.zeros((k1, k1), order="F")

    def fit(self) -> "Fit":
        r, K = self.r, self.K
        res = {f: getattr(r, f) for f, _ in S.Result._fields_}
        res["cluster_dims"] = list(r.cluster_dims)[: r.n_cluster_dims]
        res["n_clusters"] = list(r.n_clusters)
        return Fit(self.beta[:K].copy(), self.se[:K].copy(), np.array(self.vcov)[:K, :K], res)


def _dp(a):
    return a.ctypes.data_as(S.DP)


def estimate_device(panel: Panel, y_ptr: int, W_ptr: int, K: int, Z_ptr: int = 0, L: int = 0, ld: int = None,
                    weights_ptr: int = 0, **kw) -> Fit:
    lib = load()
    ld = panel.n_obs if ld is None else ld
    sp = _Spec(K, **kw)
    check(lib.fehip_estimate(panel.handle, y_ptr, W_ptr or None, K, Z_ptr or None, L, ld, weights_ptr or None,
                             C.byref(sp.opts), C.byref(sp.r), _dp(sp.beta), _dp(sp.se), _dp(sp.vcov)),
          "fehip_estimate")
    return sp.fit()


@dataclass
class Request:
    """One specification of a batch (device pointers; options as for estimate_device)."""
    y_ptr: int
    W_ptr: int
    K: int
    Z_ptr: int = 0
    L: int = 0
    ld: int = None
    weights_ptr: int = 0
    options: dict = None


def estimate_batch(panel: Panel, requests) -> list:
    """fehip_estimate_batch: all requests in one call; returns one Fit per request."""
    lib = load()
    specs = []
    arr = (S.Request * max(1, len(requests)))()
    for i, q in enumerate(requests):
        sp = _Spec(q.K, **(q.options or {}))
        specs.append(sp)
        _fill_request(arr[i], panel, q, sp)
    check(lib.fehip_estimate_batch(panel.handle, arr, len(requests)), "fehip_estimate_batch")
    return [sp.fit() for sp in specs]


def _fill_request(a, panel, q, sp):
    a.d_y, a.d_W, a.d_Z = q.y_ptr, q.W_ptr or None, q.Z_ptr or None
    a.K, a.L = q.K, q.L
    a.ld = panel.n_obs if q.ld is None else q.ld
    a.d_weights = q.weights_ptr or None
    a.opts = sp.opts
    a.result = C.pointer(sp.r)
    a.beta, a.se, a.vcov = _dp(sp.beta), _dp(sp.se), _dp(sp.vcov)


class Server:
    """fehip_server: a worker thread that batches queued requests on a device-resident panel."""

    def __init__(self, panel: Panel, hbm_budget: int = 0, max_batch: int = 0, workers: int = 0, window_us: int = 0):
        self.lib = load()
        self.panel = panel
        o = S.ServerOpts(int(hbm_budget), int(max_batch), int(workers), int(window_us), 0)
        h = C.c_void_p()
        check(self.lib.fehip_server_create(panel.handle, C.byref(o), C.byref(h)), "fehip_server_create")
        self.handle = h.value
        self._pending = {}

    def submit(self, q: Request) -> int:
        sp = _Spec(q.K, **(q.options or {}))
        req = S.Request()
        _fill_request(req, self.panel, q, sp)
        t = S.I64()
        check(self.lib.fehip_server_submit(self.handle, C.byref(req), C.byref(t)), "fehip_server_submit")
        self._pending[t.value] = sp   # keeps the host outputs alive until wait()
        return t.value

    def wait(self, ticket: int):
        lat = C.c_double()
        sp = self._pending.pop(ticket)
        check(self.lib.fehip_server_wait(self.handle, ticket, C.byref(lat)), "fehip_server_wait")
        return sp.fit(), lat.value

    def stats(self) -> dict:
        st = S.ServerStats()
        check(self.lib.fehip_server_get_stats(self.handle, C.byref(st)), "fehip_server_get_stats")
        return {f: getattr(st, f) for f, _ in S.ServerStats._fields_}

    def reset_stats(self):
        check(self.lib.fehip_server_reset_stats(self.handle), "fehip_server_reset_stats")

    def close(self):
        if self.handle:
            check(self.lib.fehip_server_destroy(self.handle), "fehip_server_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


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


def fe_values(panel: Panel, y_ptr: int, W_ptr: int, K: int, beta, ld: int = None, **kw):
    """fehip_fe_values: the fixed effects of the fit y ~ W beta, per absorbed dim (ascending), as the
    reference's --save-fe reconstruction defines them (src/core/fe_pipeline.f90:1619-1715).  Returns
    [(values, counts)] per dim (level id - 1 indexes both), and the solve's statistics."""
    lib = load()
    info = panel.info()
    fm = kw.get("fe_mask", 0) or ((1 << info["n_fe"]) - 1)
    dims = [d for d in range(info["n_fe"]) if fm >> d & 1]
    Gs = [info["group_sizes"][d] for d in dims]
    tot = sum(Gs)
    vals, cnts = DeviceBuffer(8 * max(1, tot)), DeviceBuffer(4 * max(1, tot))
    o = demean_opts(**kw)
    st = S.DemeanStats()
    b = np.ascontiguousarray(beta, dtype=np.float64) if K else np.zeros(1)
    check(lib.fehip_fe_values(panel.handle, y_ptr, W_ptr or None, K, ld or panel.n_obs, _dp(b), C.byref(o), vals.ptr,
                              cnts.ptr, C.byref(st)), "fehip_fe_values")
    v = vals.to_array(np.float64, tot)
    c = cnts.to_array(np.int32, tot)
    vals.free(), cnts.free()
    out, off = [], 0
    for g in Gs:
        out.append((v[off:off + g], c[off:off + g]))
        off += g
    return out, dict(iterations=st.iterations, converged=st.converged, max_update=st.max_update)


# ---- `.bin` dataset files (src/io/fe_data_io.f90) straight into device memory ----

def bin_header(path) -> dict:
    """fehip_bin_read_header: the header of a reference `.bin` file (layout and precision decided from
    the file size as load_dataset_from_file does, src/io/fe_data_io.f90:78-202)."""
    h = S.BinHeader()
    check(load().fehip_bin_read_header(str(path).encode(), C.byref(h)), "fehip_bin_read_header")
    return {name: getattr(h, name) for name, _ in S.BinHeader._fields_}


@dataclass
class BinData:
    """The blocks of a `.bin` file in device memory (values in FP64, ids int32; None for absent blocks)."""
    header: dict
    y: DeviceBuffer
    W: DeviceBuffer | None
    Z: DeviceBuffer | None
    fe: DeviceBuffer | None
    cluster: DeviceBuffer | None
    weights: DeviceBuffer | None

    def fe_ptrs(self) -> list:
        n = self.header["n_obs"]
        return [self.fe.ptr + 4 * n * d for d in range(self.header["n_fe"])]

    def panel(self) -> Panel:
        return Panel(self.fe_ptrs(), self.header["n_obs"])

    def free(self):
        for b in (self.y, self.W, self.Z, self.fe, self.cluster, self.weights):
            if b is not None:
                b.free()


def load_bin(path) -> BinData:
    """fehip_bin_load: stream every block of a `.bin` file into device buffers of its own."""
    h = S.BinHeader()
    lib = load()
    check(lib.fehip_bin_read_header(str(path).encode(), C.byref(h)), "fehip_bin_read_header")
    n = h.n_obs
    y = DeviceBuffer(8 * n)
    W = DeviceBuffer(8 * n * h.n_regressors) if h.n_regressors > 0 else None
    Z = DeviceBuffer(8 * n * h.n_instruments) if h.n_instruments > 0 else None
    fe = DeviceBuffer(4 * n * h.n_fe) if h.n_fe > 0 else None
    cl = DeviceBuffer(4 * n) if h.has_cluster else None
    w = DeviceBuffer(8 * n) if h.has_weights else None
    ptr = lambda b: b.ptr if b is not None else None  # noqa: E731
    check(lib.fehip_bin_load(str(path).encode(), C.byref(h), y.ptr, ptr(W), ptr(Z), ptr(fe), ptr(cl), ptr(w)),
          "fehip_bin_load")
    return BinData({name: getattr(h, name) for name, _ in S.BinHeader._fields_}, y, W, Z, fe, cl, w)
