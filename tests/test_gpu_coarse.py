"""GPU parity of the coarse path (fehip_* : device-resident panel, FE-space CG within-transform,
fused normal equations, CGM sandwich, 2SLS) against the oracle.

Tolerances: integer outputs (group sizes, dof, cluster counts, drop decisions) exact; beta and SE
within 1e-8 relative of the oracle with both sides converged to a tight tolerance (the two AP
schemes stop at different points at the production tolerance 1e-8).
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    from fe_gpu_amd import device as d
    d.init(0)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))) if a.size else 0.0


def test_panel_structures():
    from fe_gpu_amd import api, synth
    note("panel structures")
    p = synth.random_fe_problem(1, n=5000, K=1, G=(40, 1500, 7))
    pan = api.Panel.from_host(p.fe)
    inf = pan.info()
    assert inf["group_sizes"] == [40, 1500, 7]
    assert inf["primary_dim"] == 1 and inf["sorted_input"] == 0
    # runs of dim 0 inside primary segments (rows stably sorted by dim 1)
    order = np.argsort(p.fe[1], kind="stable")
    f0, f1 = p.fe[0][order], p.fe[1][order]
    runs = 1 + int(np.sum((f0[1:] != f0[:-1]) | (f1[1:] != f1[:-1])))
    assert inf["n_runs"][0] == 0            # 40 groups -> LDS dim, no runs
    pan.close()
    q = synth.random_fe_problem(2, n=5000, K=1, G=(3000, 1500))
    pan = api.Panel.from_host(q.fe)
    inf = pan.info()
    order = np.argsort(q.fe[0], kind="stable")
    f0, f1 = q.fe[0][order], q.fe[1][order]
    assert inf["primary_dim"] == 0
    assert inf["n_runs"][1] == 1 + int(np.sum((f0[1:] != f0[:-1]) | (f1[1:] != f1[:-1])))
    pan.close()


@pytest.mark.parametrize("G", [(9,), (25, 8), (3000, 1200, 6), (40, 2000, 1500, 9)])
def test_demean_matches_dense_projection(G):
    from fe_gpu_amd import api, synth
    from fe_gpu_amd.device import DeviceBuffer
    from tests import bruteforce as bf
    note(f"demean G={G}")
    p = synth.random_fe_problem(sum(G), n=6000, K=5, G=G)
    X = np.asfortranarray(np.column_stack([p.y, p.W]))
    exact = bf.annihilate(X, p.fe)
    pan = api.Panel.from_host(p.fe)
    d = DeviceBuffer.from_array(X)
    st = api.demean(pan, d.ptr, X.shape[1], X.shape[0], tol=1e-13, max_iter=20000)
    got = d.to_array(np.float64, X.shape, order="F")
    assert st["converged"] == 1
    assert np.max(np.abs(got - exact)) < 1e-9 * max(1.0, np.max(np.abs(X)))
    pan.close()


def test_demean_weights_leave_masked_rows_untouched():
    from fe_gpu_amd import api, synth
    from fe_gpu_amd.device import DeviceBuffer
    from tests import bruteforce as bf
    note("demean weights")
    p = synth.random_fe_problem(5, n=4000, K=2, G=(300, 1100, 5))
    rng = np.random.default_rng(0)
    w = (rng.random(p.n) < 0.6).astype(np.uint8)
    X = np.asfortranarray(np.column_stack([p.y, p.W]))
    pan = api.Panel.from_host(p.fe)
    d = DeviceBuffer.from_array(X)
    dw = DeviceBuffer.from_array(w)
    api.demean(pan, d.ptr, 3, p.n, dw.ptr, tol=1e-13, max_iter=20000)
    got = d.to_array(np.float64, X.shape, order="F")
    m = w == 1
    assert np.array_equal(got[~m], X[~m])
    exact = bf.annihilate(X[m], p.fe[:, m])
    assert np.max(np.abs(got[m] - exact)) < 1e-9 * np.max(np.abs(X))
    pan.close()


@pytest.mark.parametrize("case", ["small3", "large2", "unsorted4", "one"])
@pytest.mark.parametrize("cluster", [(), (1,), (1, 2), (1, 2, 3)])
def test_estimate_matches_oracle(case, cluster):
    import oracle
    from fe_gpu_amd import api, synth
    G = {"small3": (50, 12, 6), "large2": (2500, 1300), "unsorted4": (700, 1600, 30, 4), "one": (90,)}[case]
    cl = tuple(c for c in cluster if c <= len(G))
    note(f"estimate case={case} cluster={cl}")
    p = synth.random_fe_problem(len(G) * 7 + len(cl), n=8000, K=3, G=G)
    fit = api.estimate(p.y, p.W, p.fe, cluster=cl, tol=1e-13, max_iter=20000)
    ref = oracle.estimate(p.y, p.W, p.fe, cluster=cl, tol=1e-13, max_iter=200000)
    assert fit.res["converged"] == 1
    assert rel(fit.beta, ref.beta) < 1e-8
    assert rel(fit.se, ref.se) < 1e-8
    for k in ("dof_fes", "dof_resid", "dof_model", "rank"):
        assert fit.res[k] == ref.res[k], k
    if cl:
        assert fit.res["n_clusters"] == ref.res["n_clusters"]
        assert fit.res["psd_fix_applied"] == ref.res["psd_fix_applied"]
    for k in ("rss", "tss_total", "tss_within", "r2", "r2_within", "r2_adj", "f_stat"):
        assert abs(fit.res[k] - ref.res[k]) <= 1e-8 * max(1.0, abs(ref.res[k])), k


def test_estimate_nlsw_like_homoskedastic_and_3way():
    import oracle
    from fe_gpu_amd import api, synth
    note("estimate nlsw-like")
    p = synth.nlsw_like()
    for cl in [(), (1, 2, 3)]:
        fit = api.estimate(p.y, p.W, p.fe, cluster=cl, tol=1e-12, max_iter=20000)
        ref = oracle.estimate(p.y, p.W, p.fe, cluster=cl, tol=1e-12, max_iter=20000)
        assert rel(fit.beta, ref.beta) < 1e-8 and rel(fit.se, ref.se) < 1e-8
        assert fit.res["dof_resid"] == ref.res["dof_resid"]


def test_estimate_iv_matches_oracle():
    import oracle
    from fe_gpu_amd import api, synth
    note("estimate IV")
    p = synth.random_fe_problem(91, n=6000, K=4, L=3, G=(1500, 1100, 8))
    for cl in [(), (1,), (2, 3)]:
        fit = api.estimate(p.y, p.W, p.fe, Z=p.Z, iv_reg=[1, 3], cluster=cl, tol=1e-13, max_iter=20000)
        ref = oracle.estimate(p.y, p.W, p.fe, Z=p.Z, iv_reg=[1, 3], cluster=cl, tol=1e-13, max_iter=200000)
        assert fit.res["is_iv"] == 1 and fit.res["n_instruments"] == ref.res["n_instruments"]
        assert rel(fit.beta, ref.beta) < 1e-8 and rel(fit.se, ref.se) < 1e-8


def test_estimate_column_filter():
    import oracle
    from fe_gpu_amd import api, synth
    note("column filter")
    p = synth.random_fe_problem(93, n=5000, K=3, G=(1500, 20))
    W = p.W.copy()
    W[:, 2] = np.arange(1, 1501)[p.fe[0] - 1] * 0.1
    fit = api.estimate(p.y, W, p.fe, tol=1e-13, max_iter=20000)
    ref = oracle.estimate(p.y, W, p.fe, tol=1e-13, max_iter=200000)
    assert fit.res["dof_model"] == ref.res["dof_model"] == 2
    assert fit.beta[2] == 0.0 and fit.se[2] == 0.0
    assert rel(fit.beta[:2], ref.beta[:2]) < 1e-8


def _redense(ids):
    u, inv = np.unique(ids, return_inverse=True)
    return (inv + 1).astype(np.int32)


@pytest.mark.parametrize("fe_mask,cluster", [(0b011, (1,)), (0b110, (2, 3)), (0b111, (1, 2, 3))])
def test_masked_request_equals_oracle_on_subset(fe_mask, cluster):
    """Serving semantics: a request with a row mask and an FE subset equals the oracle run on
    the masked rows with re-densified ids (G_d = non-empty groups)."""
    import oracle
    from fe_gpu_amd import api, synth
    note(f"masked request fe_mask={fe_mask:b} cluster={cluster}")
    p = synth.random_fe_problem(55, n=9000, K=3, G=(2000, 1200, 9))
    rng = np.random.default_rng(1)
    w = (rng.random(p.n) < 0.45).astype(np.uint8)
    dims = [d for d in range(3) if fe_mask >> d & 1]
    fit = api.estimate(p.y, p.W, p.fe, weights=w, cluster=cluster, fe_mask=fe_mask, tol=1e-13, max_iter=20000)
    m = w == 1
    fe_sub = np.vstack([_redense(p.fe[d][m]) for d in dims])
    cl_sub = tuple(dims.index(c - 1) + 1 for c in cluster)
    ref = oracle.estimate(p.y[m], p.W[m], fe_sub, cluster=cl_sub, tol=1e-13, max_iter=200000)
    assert fit.res["n_obs"] == int(m.sum())
    assert fit.res["dof_fes"] == ref.res["dof_fes"] and fit.res["dof_resid"] == ref.res["dof_resid"]
    assert rel(fit.beta, ref.beta) < 1e-8 and rel(fit.se, ref.se) < 1e-8


def test_synth_generator_is_deterministic_and_shaped():
    from fe_gpu_amd import _lib
    from fe_gpu_amd.device import DeviceBuffer
    note("synth generator")
    L = _lib.load()
    nw, T, nf = 20000, 30, 4000
    n = nw * T
    outs = []
    for _ in range(2):
        fe, y, W = DeviceBuffer(3 * n * 4), DeviceBuffer(n * 8), DeviceBuffer(8 * n * 8)
        assert L.fehip_synth_worker_firm(nw, T, nf, 0.1, 202, fe.ptr, y.ptr, W.ptr) == 0
        outs.append((fe.to_array(np.int32, (3, n)), y.to_array(np.float64, (n,)),
                     W.to_array(np.float64, (n, 8), order="F")))
    assert all(np.array_equal(a, b) for a, b in zip(outs[0], outs[1]))
    fe, y, W = outs[0]
    assert fe[0].min() == 1 and fe[0].max() == nw and np.all(np.diff(fe[0]) >= 0)
    assert fe[2].min() == 1 and fe[2].max() == T and fe[1].max() <= nf
    firm = fe[1].reshape(nw, T)
    switch = np.mean(firm[:, 1:] != firm[:, :-1])
    assert 0.09 < switch < 0.11
    assert abs(W[:, 1].mean() - 0.05) < 0.005                 # sick_shock ~ Bernoulli(0.05)


@pytest.mark.parametrize("K,cluster", [(3, (2, 3)), (12, (3,))])
def test_sorted_subsets_with_clusters_across_waves(K, cluster):
    """Cluster subsets routed through the sorted partition whose clusters hold hundreds of rows, so
    they cross many 64-position waves (pieces merged at the boundaries)."""
    import oracle
    from fe_gpu_amd import api, synth
    note(f"sorted subsets K={K} cluster={cluster}")
    p = synth.random_fe_problem(77 + K, n=40000, K=K, G=(4000, 40, 30))
    fit = api.estimate(p.y, p.W, p.fe, cluster=cluster, tol=1e-13, max_iter=20000)
    ref = oracle.estimate(p.y, p.W, p.fe, cluster=cluster, tol=1e-13, max_iter=200000)
    assert fit.res["n_clusters"] == ref.res["n_clusters"]
    assert rel(fit.beta, ref.beta) < 1e-8 and rel(fit.se, ref.se) < 1e-8


@pytest.mark.parametrize("K,cluster,weighted,nw,nf", [(3, (2, 3), False, 2000, 1500), (8, (1, 2, 3), False, 2000, 1500),
                                                       (8, (2, 3), True, 2000, 1500), (12, (1, 2, 3), True, 2000, 1500),
                                                       (8, (2, 3), False, 6000, 1100)])
def test_large_by_small_cluster_subset(K, cluster, weighted, nw, nf):
    """A cluster subset of one large dim (firms, more groups than the LDS tables take) and a small one
    (periods) on a sticky worker-firm panel: the score rows are written in the firm partition, whole runs
    at a time, with each row's period cell, and the cells of a firm are summed by one wave (the last case
    has firms of more than 64 rows, which the wave takes in several passes)."""
    import oracle
    from fe_gpu_amd import api, synth
    note(f"large x small subset K={K} cluster={cluster} weighted={weighted} workers={nw} firms={nf}")
    p = synth.worker_firm_panel(seed=300 + K, n_workers=nw, n_periods=20, n_firms=nf)
    rng = np.random.default_rng(K)
    W = p.W[:, :K] if K <= 8 else np.column_stack([p.W, rng.normal(size=(p.W.shape[0], K - 8))])
    n = W.shape[0]
    fe = np.vstack([_redense(p.fe[d]) for d in range(3)])  # no unused firm ids (G_d = max id)
    w = (rng.random(n) < 0.7).astype(np.uint8) if weighted else None
    fit = api.estimate(p.y, W, fe, weights=w, cluster=cluster, tol=1e-13, max_iter=20000)
    m = np.ones(n, bool) if w is None else w == 1
    fe_sub = np.vstack([_redense(fe[d][m]) for d in range(3)])
    ref = oracle.estimate(p.y[m], W[m], fe_sub, cluster=cluster, tol=1e-13, max_iter=200000)
    if nw > 2000:
        assert np.bincount(fe[1]).max() > 64  # firms longer than one wave pass
    assert fit.res["n_clusters"] == ref.res["n_clusters"]
    assert fit.res["dof_resid"] == ref.res["dof_resid"]
    assert rel(fit.beta, ref.beta) < 1e-8 and rel(fit.se, ref.se) < 1e-8


def test_wide_iv_design_matches_oracle():
    """A wide 2SLS design (40 regressors, 8 endogenous, 12 instruments): the normal equations, the
    first stage and X^ run through the wide GEMM paths."""
    import oracle
    from fe_gpu_amd import api, synth
    note("wide IV design")
    p = synth.random_fe_problem(808, n=20000, K=40, L=12, G=(1200, 300, 15))
    Z = p.Z + 0.5 * p.W[:, :12]
    iv = list(range(1, 9))
    fit = api.estimate(p.y, p.W, p.fe, Z=Z, iv_reg=iv, cluster=(1, 2), tol=1e-13, max_iter=20000)
    ref = oracle.estimate(p.y, p.W, p.fe, Z=Z, iv_reg=iv, cluster=(1, 2), tol=1e-13, max_iter=200000)
    assert fit.res["is_iv"] == 1 and fit.res["n_instruments"] == ref.res["n_instruments"]
    assert fit.res["dof_model"] == ref.res["dof_model"]
    assert rel(fit.beta, ref.beta) < 1e-8 and rel(fit.se, ref.se) < 1e-8


def test_iv_reverts_to_ols_and_skips_bad_indices():
    """The reference's selection (src/core/fe_pipeline.f90:664-753): out-of-range endogenous and
    instrument indices are ignored, no valid instrument index means every instrument column, and fewer
    instruments than endogenous regressors reverts to OLS (is_iv 0, n_instruments 0)."""
    import oracle
    from fe_gpu_amd import api, synth
    note("IV selection and reversion")
    p = synth.random_fe_problem(92, n=5000, K=4, L=2, G=(900, 700, 6))
    cases = [
        dict(iv_reg=[1, 2, 3], iv_inst=[]),      # 3 endogenous, 2 instruments: OLS
        dict(iv_reg=[2, 9], iv_inst=[]),         # index 9 ignored: 1 endogenous, IV
        dict(iv_reg=[1], iv_inst=[5, 0]),        # no valid instrument index: every instrument column
        dict(iv_reg=[1, 3], iv_inst=[2, 7]),     # one valid instrument for two endogenous: OLS
    ]
    for c in cases:
        fit = api.estimate(p.y, p.W, p.fe, Z=p.Z, cluster=(1,), tol=1e-13, max_iter=20000, **c)
        ref = oracle.estimate(p.y, p.W, p.fe, Z=p.Z, cluster=(1,), tol=1e-13, max_iter=200000, **c)
        assert fit.res["is_iv"] == ref.res["is_iv"], c
        assert fit.res["n_instruments"] == ref.res["n_instruments"], c
        assert rel(fit.beta, ref.beta) < 1e-8 and rel(fit.se, ref.se) < 1e-8, c
    ols = oracle.estimate(p.y, p.W, p.fe, cluster=(1,), tol=1e-13, max_iter=200000)
    fit = api.estimate(p.y, p.W, p.fe, Z=p.Z, iv_reg=[1, 2, 3], cluster=(1,), tol=1e-13, max_iter=20000)
    assert fit.res["is_iv"] == 0 and fit.res["n_instruments"] == 0
    assert rel(fit.beta, ols.beta) < 1e-8 and rel(fit.se, ols.se) < 1e-8


def test_homoskedastic_se_of_a_nearly_exact_fit():
    """Noise of sd 1e-5 leaves rss about 1e-10 of y~'y~: rss is then taken from the residual, not from
    the cross products (whose cancellation would cost about 1e-6 relative), so the SEs keep 1e-8."""
    import oracle
    from fe_gpu_amd import api, synth
    note("nearly exact fit")
    p = synth.random_fe_problem(94, n=8000, K=3, G=(1500, 900, 7))
    rng = np.random.default_rng(5)
    y = p.W @ np.array([0.7, -1.2, 0.4]) + 3.0 * rng.normal(size=1500)[p.fe[0] - 1] + 1e-5 * rng.normal(size=p.n)
    fit = api.estimate(y, p.W, p.fe, tol=1e-14, max_iter=20000)
    ref = oracle.estimate(y, p.W, p.fe, tol=1e-14, max_iter=200000)
    assert ref.res["rss"] < 1e-8 * ref.res["tss_within"]
    assert rel(fit.beta, ref.beta) < 1e-8
    assert rel(fit.se, ref.se) < 1e-8
    assert abs(fit.res["rss"] - ref.res["rss"]) <= 1e-8 * ref.res["rss"]


@pytest.mark.parametrize("max_iter,check_every", [(7, 4), (5, 3), (8, 4)])
def test_cg_iteration_cap(max_iter, check_every):
    """A solve stopped by max_iter before converging: the captured graph of check_every steps plus the
    steps past the last whole graph give the same result as launching every step directly (the
    per-kernel timing path), and report converged 0 with max_iter iterations."""
    from fe_gpu_amd import _lib, api, synth
    from fe_gpu_amd.device import DeviceBuffer
    note(f"cg cap max_iter={max_iter} check_every={check_every}")
    p = synth.worker_firm_panel(seed=11, n_workers=3000, n_periods=12, n_firms=600)
    X = np.asfortranarray(np.column_stack([p.y, p.W[:, :4]]))
    pan = api.Panel.from_host(p.fe)
    L = _lib.load()
    outs = []
    for timed in (False, True):
        d = DeviceBuffer.from_array(X)
        L.fehip_kstats_enable(1 if timed else 0)
        try:
            st = api.demean(pan, d.ptr, X.shape[1], X.shape[0], tol=1e-15, max_iter=max_iter,
                            check_every=check_every)
        finally:
            L.fehip_kstats_enable(0)
        assert st["converged"] == 0 and st["iterations"] == max_iter, st
        outs.append(d.to_array(np.float64, X.shape, order="F"))
    assert np.max(np.abs(outs[0] - outs[1])) <= 1e-14 * np.max(np.abs(outs[1]))
    pan.close()


def test_server_ticket_takes_one_waiter():
    """Two threads waiting on one ticket: exactly one receives the result; the other is refused (the
    ticket is freed when its wait returns)."""
    import ctypes as C
    import threading
    from fe_gpu_amd import _lib, api, synth
    from fe_gpu_amd.device import DeviceBuffer
    note("server ticket waiters")
    p = synth.random_fe_problem(95, n=20000, K=2, G=(3000, 800, 9))
    pan = api.Panel.from_host(p.fe)
    dy, dW = DeviceBuffer.from_array(p.y), DeviceBuffer.from_array(np.asfortranarray(p.W))
    srv = api.Server(pan, workers=1)
    t = srv.submit(api.Request(dy.ptr, dW.ptr, 2, options=dict(cluster=(1, 2), tol=1e-10)))
    L = _lib.load()
    status = []
    th = threading.Thread(target=lambda: status.append(L.fehip_server_wait(srv.handle, t, None)))
    th.start()
    status.append(L.fehip_server_wait(srv.handle, t, C.byref(C.c_double())))
    th.join()
    assert sorted(s == 0 for s in status) == [False, True], status
    L.fe_gpu_runtime_clear_error()
    srv._pending.clear()
    srv.close()
    pan.close()


def test_config3_wide_2sls_matches_oracle():
    """The wide 2SLS workload at a size the oracle finishes in seconds: the device generator's panel
    (1,500 individuals x 20 periods, 400 firms, 10% switching) with K = 252 design columns (xend and the
    7 cat8 dummies endogenous, x1..x4, the 29 + 7 + 203 dummies of i.ind&&i.msp, x1&x2), 21 excluded
    instruments (z1, z2, 19 zcat20 dummies), so L_aug = 244 + 21 = 265; fe(id firm time), cluster(id
    firm).  It runs the normal equations of both stages on the matrix cores
    (src/core/fe_pipeline.f90:947-1032, selection :664-753)."""
    import oracle
    from fe_gpu_amd import _lib, api
    from fe_gpu_amd._lib import check
    from fe_gpu_amd.device import DeviceBuffer
    ni, T, nf, K, L = 1500, 20, 400, 252, 21
    n = ni * T
    note(f"config-3 shape 2SLS n={n} K={K} L={L}")
    lib = _lib.load()
    d_fe, d_y = DeviceBuffer(3 * n * 4), DeviceBuffer(n * 8)
    d_W, d_Z = DeviceBuffer(K * n * 8), DeviceBuffer(L * n * 8)
    check(lib.fehip_synth_iv_wide(ni, T, nf, 0.10, 303, d_fe.ptr, d_y.ptr, d_W.ptr, d_Z.ptr), "iv wide synth")
    fe = d_fe.to_array(np.int32, (3, n))
    y = d_y.to_array(np.float64, n)
    W = d_W.to_array(np.float64, (n, K), order="F")
    Z = d_Z.to_array(np.float64, (n, L), order="F")
    G = fe.max(axis=1)
    kw = dict(cluster=(1, 2), iv_reg=list(range(1, 9)))
    pan = api.Panel([d_fe.ptr, d_fe.ptr + n * 4, d_fe.ptr + 2 * n * 4], n)
    fit = api.estimate_device(pan, d_y.ptr, d_W.ptr, K, d_Z.ptr, L, n, tol=1e-13, max_iter=20000, **kw)
    pan.close()
    note("oracle")
    ref = oracle.estimate(y, W, fe, G=G, Z=Z, tol=1e-13, max_iter=200000, **kw)
    assert ref.res["is_iv"] == 1 and ref.res["n_instruments"] == 265 and ref.res["dof_model"] == 252
    for k in ("is_iv", "n_instruments", "dof_model", "dof_fes", "dof_resid", "rank"):
        assert fit.res[k] == ref.res[k], k
    assert fit.res["n_clusters"] == ref.res["n_clusters"]
    assert fit.res["psd_fix_applied"] == ref.res["psd_fix_applied"]
    assert rel(fit.beta, ref.beta) < 1e-8
    assert rel(fit.se, ref.se) < 1e-8
    for k in ("rss", "tss_within", "r2", "f_stat"):
        assert abs(fit.res[k] - ref.res[k]) <= 1e-8 * max(1.0, abs(ref.res[k])), k


def _returning_panel(seed, n_workers, n_periods, n_firms, K):
    """Workers moving between a home firm and a second one, so a worker's rows of one firm form several
    runs (a return), with the rows shuffled (the panel sorts them)."""
    rng = np.random.default_rng(seed)
    home = rng.integers(1, n_firms + 1, n_workers)
    other = rng.integers(1, n_firms + 1, n_workers)
    at_home = rng.random((n_workers, n_periods)) < 0.6
    firm = np.where(at_home, home[:, None], other[:, None]).astype(np.int32)
    worker = np.repeat(np.arange(1, n_workers + 1, dtype=np.int32), n_periods)
    period = np.tile(np.arange(1, n_periods + 1, dtype=np.int32), n_workers)
    fe = np.vstack([worker, firm.reshape(-1), period])
    n = fe.shape[1]
    W = rng.normal(size=(n, K)) + 0.3 * rng.normal(size=n_workers)[worker - 1, None]
    y = W @ rng.normal(size=K) + rng.normal(size=n_workers)[worker - 1] + rng.normal(size=n_firms + 1)[fe[1]] \
        + rng.normal(size=n)
    perm = rng.permutation(n)
    fe = np.vstack([_redense(f) for f in fe[:, perm]])
    return y[perm], np.asfortranarray(W[perm]), fe


@pytest.mark.parametrize("cluster,weighted,precision", [((1, 2), False, 0), ((2,), False, 0), ((1,), False, 0),
                                                        ((1, 2, 3), False, 0), ((1, 2), True, 0), ((1, 2), False, 1)])
def test_wide_scores_cluster_sums(cluster, weighted, precision):
    """Scores wider than the fused pass (K = 20): the clusters of the primary dim (workers), of a large
    dim (firms, 1,500 > the LDS tables) and of their cells come from one pass of segment and run sums.
    Workers return to an earlier firm, so a cell holds several runs (merged in group order); the 3-way
    case mixes these subsets with the period subsets, which take the other methods; the weighted case
    counts only the cells holding a weighted row.  FP32 storage (precision 1) is held to the FP32
    bounds: beta 1e-6, SE 1e-5 relative."""
    import oracle
    from fe_gpu_amd import api
    note(f"wide scores cluster={cluster} weighted={weighted} precision={precision}")
    y, W, fe = _returning_panel(41, 2500, 12, 1500, 20)
    n = len(y)
    order = np.lexsort((fe[2], fe[0]))
    f0, f1 = fe[0][order], fe[1][order]
    runs = 1 + int(np.sum((f0[1:] != f0[:-1]) | (f1[1:] != f1[:-1])))
    cells = len(np.unique(f0.astype(np.int64) * 4096 + f1))
    assert runs > cells + 1000  # many returns
    rng = np.random.default_rng(9)
    w = (rng.random(n) < 0.7).astype(np.uint8) if weighted else None
    fit = api.estimate(y, W, fe, weights=w, cluster=cluster, tol=1e-13, max_iter=20000, precision=precision)
    m = np.ones(n, bool) if w is None else w == 1
    fe_sub = np.vstack([_redense(fe[d][m]) for d in range(3)])
    ref = oracle.estimate(y[m], W[m], fe_sub, cluster=cluster, tol=1e-13, max_iter=200000)
    assert fit.res["n_clusters"] == ref.res["n_clusters"]
    assert fit.res["dof_resid"] == ref.res["dof_resid"]
    bt, st = (1e-8, 1e-8) if precision == 0 else (1e-6, 1e-5)
    assert rel(fit.beta, ref.beta) < bt and rel(fit.se, ref.se) < st


@pytest.mark.parametrize("K", [3, 20])
def test_pairwise_subset_keys_match_oracle(K, monkeypatch):
    """Cluster subsets whose mixed-radix key would pass 64 bits are partitioned pairwise (dense rank of the
    dims so far, then the next dim).  The test hook FEHIP_PARTITION_PAIRWISE=1 sends every multi-dim subset
    there; 3-way clustered fits (the sorted-partition route at K = 3, the partition sums at K = 20) then
    match the oracle, whose builder takes its hash path only past 2^62 (cluster_id_builder.c:259-269).
    At K = 3 the two large non-primary dims put {2,3} on a two-dim sorted partition; at K = 20 the
    subsets with the 9-group dim ({1,3}, {2,3}, {1,2,3}) take two- and three-dim partitions."""
    import oracle
    from fe_gpu_amd import api, synth
    note(f"pairwise subset keys K={K}")
    monkeypatch.setenv("FEHIP_PARTITION_PAIRWISE", "1")
    G = (3000, 1500, 1100) if K == 3 else (2000, 1200, 9)
    p = synth.random_fe_problem(91 + K, n=12000, K=K, G=G)
    fit = api.estimate(p.y, p.W, p.fe, cluster=(1, 2, 3), tol=1e-13, max_iter=20000)
    monkeypatch.delenv("FEHIP_PARTITION_PAIRWISE")
    ref = oracle.estimate(p.y, p.W, p.fe, cluster=(1, 2, 3), tol=1e-13, max_iter=200000)
    assert fit.res["n_clusters"] == ref.res["n_clusters"]
    assert rel(fit.beta, ref.beta) < 1e-8 and rel(fit.se, ref.se) < 1e-8


@pytest.mark.parametrize("case", ["worker_firm_time", "with_region", "worker_firm", "mixed_precision", "returns"])
def test_run_operator_matches_row_operator_and_oracle(case, monkeypatch):
    """The run-level operator (an unweighted solve absorbing the primary dim, one large dim and small dims:
    the large dim's block from its runs, the small dims through the couplings B = D_D'M_P D_s and
    C = D_s'M_P D_s built once) gives the fit of the row kernels (FEHIP_RUN_OPERATOR=0) to 1e-10 and the
    oracle's to 1e-8 (FP64) or to the FP32 bounds (beta 1e-6, SE 1e-5) for the mixed-precision solve.
    Cases: worker x firm x period; plus a 50-group region dim (two small dims); worker x firm only (no
    couplings); the FP32 operator rounds; workers returning to an earlier firm (several runs of one group
    in a segment)."""
    import oracle
    from fe_gpu_amd import api, synth
    note(f"run operator {case}")
    if case == "returns":
        y, W, fe = _returning_panel(43, 3000, 12, 1500, 6)
        cl = (1, 2)
    else:
        # the FP32 operator rounds need a table of more than 4,096 entries: 6,000 firms there
        p = synth.worker_firm_panel(seed=61, n_workers=6000, n_periods=15,
                                    n_firms=6000 if case == "mixed_precision" else 2500)
        y, W = p.y, p.W
        fe = np.vstack([_redense(p.fe[d]) for d in range(3)])
        cl = (1, 2, 3)
        if case == "with_region":
            region = (np.arange(fe[1].max() + 1) % 50 + 1).astype(np.int32)[fe[1]]
            fe = np.vstack([fe, region])
            cl = (1, 2)
        if case == "worker_firm":
            fe = fe[:2]
            cl = (1, 2)
    prec = 1 if case == "mixed_precision" else 0
    kw = dict(cluster=cl, tol=1e-13 if prec == 0 else 1e-12, max_iter=20000, precision=prec)
    fit = api.estimate(y, W, fe, **kw)
    monkeypatch.setenv("FEHIP_RUN_OPERATOR", "0")
    row = api.estimate(y, W, fe, **kw)
    monkeypatch.delenv("FEHIP_RUN_OPERATOR")
    ref = oracle.estimate(y, W, fe, cluster=cl, tol=1e-13, max_iter=200000)
    assert fit.res["n_clusters"] == ref.res["n_clusters"] and fit.res["dof_resid"] == ref.res["dof_resid"]
    if prec == 0:
        assert rel(fit.beta, row.beta) < 1e-10 and rel(fit.se, row.se) < 1e-10
        assert rel(fit.beta, ref.beta) < 1e-8 and rel(fit.se, ref.se) < 1e-8
    else:
        assert rel(fit.beta, ref.beta) < 1e-6 and rel(fit.se, ref.se) < 1e-5
        assert fit.res["fe_refinements"] >= 1


@pytest.mark.parametrize("fe_mask", [0b0110, 0b0010, 0b0101, 0b1100, 0b1110, 0b1101])
def test_table_operator_matches_row_operator_and_oracle(fe_mask, monkeypatch):
    """Unweighted solves without the run structure take table-level products: with the primary dim not
    absorbed, one large dim D and small dims s give S p = (n_D p_D + B p_s, B' p_D + C p_s) with the counts
    B = D_D'D_s, C = D_s'D_s ({firm, period}, {firm}, {firm, period, region}); with no large dim S = C
    ({worker, period}: C = D_s'M_P D_s; {period, region}; {worker, period, region}).  Each fit equals the
    row kernels' (FEHIP_RUN_OPERATOR=0) to 1e-10 and the oracle's on the same FE dims to 1e-8, with the
    cluster dims inside the FE set."""
    import oracle
    from fe_gpu_amd import api, synth
    note(f"table operator fe_mask={fe_mask:04b}")
    p = synth.worker_firm_panel(seed=73, n_workers=3000, n_periods=12, n_firms=1500)
    fe3 = np.vstack([_redense(p.fe[d]) for d in range(3)])
    region = (np.arange(fe3[1].max() + 1) % 40 + 1).astype(np.int32)[fe3[1] - 1]
    fe = np.vstack([fe3, region])
    dims = [d for d in range(4) if fe_mask >> d & 1]
    cl = tuple(d + 1 for d in dims[:2])
    kw = dict(cluster=cl, fe_mask=fe_mask, tol=1e-13, max_iter=20000)
    fit = api.estimate(p.y, p.W, fe, **kw)
    monkeypatch.setenv("FEHIP_RUN_OPERATOR", "0")
    row = api.estimate(p.y, p.W, fe, **kw)
    monkeypatch.delenv("FEHIP_RUN_OPERATOR")
    ref = oracle.estimate(p.y, p.W, fe[dims], cluster=tuple(dims.index(c - 1) + 1 for c in cl), tol=1e-13,
                          max_iter=200000)
    assert fit.res["dof_fes"] == ref.res["dof_fes"] and fit.res["dof_resid"] == ref.res["dof_resid"]
    assert fit.res["n_clusters"][:len(cl)] == ref.res["n_clusters"][:len(cl)]
    assert rel(fit.beta, row.beta) < 1e-10 and rel(fit.se, row.se) < 1e-10
    assert rel(fit.beta, ref.beta) < 1e-8 and rel(fit.se, ref.se) < 1e-8


@pytest.mark.parametrize("fe_mask", [0b0111, 0b0011, 0b1011, 0b0110, 0b0101, 0b1100])
def test_subsample_table_operator_matches_row_operator_and_oracle(fe_mask, monkeypatch):
    """A solve over a row subsample (one weights mask) with the primary dim absorbed takes the run-level
    products with the runs' lengths in the primary segments counted over the subsample's rows and the
    couplings B, C built from the masked row operator ({worker, firm}, {worker, firm, period},
    {worker, firm, region}); the other FE sets keep the row kernels.  The mask drops every row of each 7th
    worker (primary segments without rows) and of each 11th firm.  Each fit equals the row kernels'
    (FEHIP_RUN_OPERATOR=0) to 1e-10 and the oracle's on the masked rows to 1e-8; a mixed-precision fit (FP32
    operator rounds) is within 1e-6 (beta) and 1e-5 (SE) of the FP64 fit."""
    import oracle
    from fe_gpu_amd import _lib, api, synth
    note(f"subsample table operator fe_mask={fe_mask:04b}")
    lib = _lib.load()
    p = synth.worker_firm_panel(seed=79, n_workers=3000, n_periods=12, n_firms=1500)
    fe3 = np.vstack([_redense(p.fe[d]) for d in range(3)])
    region = (np.arange(fe3[1].max() + 1) % 40 + 1).astype(np.int32)[fe3[1] - 1]
    fe = np.vstack([fe3, region])
    rng = np.random.default_rng(7)
    w = (rng.random(p.n) < 0.6) & (fe[0] % 7 != 0) & (fe[1] % 11 != 0)
    dims = [d for d in range(4) if fe_mask >> d & 1]
    cl = tuple(d + 1 for d in dims[:2])
    kw = dict(cluster=cl, fe_mask=fe_mask, tol=1e-13, max_iter=20000)
    pan = api.Panel.from_host(fe)
    lib.fehip_kstats_reset()
    lib.fehip_kstats_enable(1)
    try:
        fit = api.estimate(p.y, p.W, fe, weights=w.astype(np.uint8), panel=pan, **kw)
    finally:
        lib.fehip_kstats_enable(0)
    name = C.create_string_buffer(32)
    launches = C.c_int64()
    dd = C.c_double()
    lib.fehip_kstats_get(8, name, 32, C.byref(launches), C.byref(dd), C.byref(dd), C.byref(dd))
    if fe_mask in (0b0111, 0b1011):
        assert launches.value > 0, "the couplings of the run-level operator were not built"
    else:
        assert launches.value == 0
    fit32 = api.estimate(p.y, p.W, fe, weights=w.astype(np.uint8), panel=pan, precision=1, **kw)
    monkeypatch.setenv("FEHIP_RUN_OPERATOR", "0")
    row = api.estimate(p.y, p.W, fe, weights=w.astype(np.uint8), panel=pan, **kw)
    monkeypatch.delenv("FEHIP_RUN_OPERATOR")
    m = w
    fe_sub = np.vstack([_redense(fe[d][m]) for d in dims])
    ref = oracle.estimate(p.y[m], p.W[m], fe_sub, cluster=tuple(dims.index(c - 1) + 1 for c in cl), tol=1e-13,
                          max_iter=200000)
    assert fit.res["n_obs"] == int(m.sum())
    assert fit.res["dof_fes"] == ref.res["dof_fes"] and fit.res["dof_resid"] == ref.res["dof_resid"]
    assert rel(fit.beta, row.beta) < 1e-10 and rel(fit.se, row.se) < 1e-10
    assert rel(fit.beta, ref.beta) < 1e-8 and rel(fit.se, ref.se) < 1e-8
    assert rel(fit32.beta, fit.beta) < 1e-6 and rel(fit32.se, fit.se) < 1e-5
    pan.close()
