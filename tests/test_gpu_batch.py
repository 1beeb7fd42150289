"""GPU parity of fehip_estimate_batch: every request of a mixed batch (different FE sets, row
subsamples, regressor sets, cluster sets, an IV request) gives the same results as the same request
solved alone, and as the oracle on its own rows with re-densified ids.

Tolerances: integer outputs exact; beta/SE within 1e-10 relative between batched and alone (same
per-column stopping rule; only the order of LDS additions differs); 1e-8 against the oracle with
both sides converged to 1e-13.
"""
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


def _redense(ids):
    _, inv = np.unique(ids, return_inverse=True)
    return (inv + 1).astype(np.int32)


INT_KEYS = ("n_obs", "dof_fes", "dof_resid", "dof_model", "rank", "is_iv", "n_instruments", "n_clusters",
            "psd_fix_applied", "converged")


def _specs(rng, n):
    """(fe_mask, first W column, K, cluster dims, weights or None, iv)"""
    out = []
    masks = [0b111, 0b011, 0b110, 0b111, 0b101, 0b111]
    for i, fm in enumerate(masks * 2):
        dims = [d + 1 for d in range(3) if fm >> d & 1]
        k = int(rng.integers(1, 5))
        j0 = int(rng.integers(0, 6 - k + 1))
        ncl = int(rng.integers(0, len(dims) + 1))
        cl = tuple(sorted(rng.choice(dims, size=ncl, replace=False).tolist())) if ncl else ()
        w = (rng.random(n) < rng.uniform(0.3, 1.0)).astype(np.uint8) if i % 3 != 0 else None
        out.append((fm, j0, k, cl, w, False))
    out.append((0b111, 0, 4, (1,), None, True))  # IV: W cols 1 and 3 endogenous, Z instruments
    return out


@pytest.mark.parametrize("multi_slot", [False, True])
def test_batch_equals_alone_and_oracle(multi_slot, monkeypatch):
    """Requests with row subsamples take one chunk per subsample, or with FEHIP_MULTI_SLOT=1 share
    multi-slot sweeps (per-row mask words, per-column slot bits); either way each equals the same request
    alone and the oracle."""
    import oracle
    if multi_slot:
        monkeypatch.setenv("FEHIP_MULTI_SLOT", "1")
    from fe_gpu_amd import api, synth
    from fe_gpu_amd.device import DeviceBuffer
    note("batch of mixed requests")
    p = synth.random_fe_problem(404, n=7000, K=6, L=3, G=(1800, 900, 7))
    n = p.n
    rng = np.random.default_rng(7)
    specs = _specs(rng, n)
    pan = api.Panel.from_host(p.fe)
    dy = DeviceBuffer.from_array(p.y)
    dW = DeviceBuffer.from_array(np.asfortranarray(p.W))
    dZ = DeviceBuffer.from_array(np.asfortranarray(p.Z))
    dws = [DeviceBuffer.from_array(w) if w is not None else None for (_, _, _, _, w, _) in specs]
    reqs = []
    for (fm, j0, k, cl, w, iv), dw in zip(specs, dws):
        opt = dict(cluster=cl, fe_mask=fm, tol=1e-13, max_iter=20000)
        if iv:
            opt["iv_reg"] = [1, 3]
        reqs.append(api.Request(dy.ptr, dW.ptr + 8 * n * j0, k, dZ.ptr if iv else 0, 3 if iv else 0, n,
                                dw.ptr if dw else 0, opt))
    fits = api.estimate_batch(pan, reqs)
    assert len(fits) == len(reqs)
    for i, ((fm, j0, k, cl, w, iv), q, fb) in enumerate(zip(specs, reqs, fits)):
        note(f"request {i}: fe_mask={fm:b} K={k} cluster={cl} weighted={w is not None} iv={iv}")
        fa = api.estimate_device(pan, q.y_ptr, q.W_ptr, q.K, q.Z_ptr, q.L, n, q.weights_ptr, **q.options)
        for key in INT_KEYS:
            assert fb.res[key] == fa.res[key], (i, key)
        assert rel(fb.beta, fa.beta) < 1e-10 and rel(fb.se, fa.se) < 1e-10, i
        # the oracle on the request's rows: a full-sample request keeps the ids (G_d = max id, as the
        # reference counts groups); a subsample is re-densified over the absorbed dims, as if its
        # rows had been written to their own panel file
        m = np.ones(n, bool) if w is None else w == 1
        dims = [d for d in range(3) if fm >> d & 1]
        fe_sub = np.vstack([p.fe[d] if w is None else _redense(p.fe[d][m]) for d in dims])
        cl_sub = tuple(dims.index(c - 1) + 1 for c in cl)
        W = p.W[m][:, j0:j0 + k]
        if iv:
            ref = oracle.estimate(p.y[m], W, fe_sub, Z=p.Z[m], iv_reg=[1, 3], cluster=cl_sub, tol=1e-13,
                                  max_iter=200000)
        else:
            ref = oracle.estimate(p.y[m], W, fe_sub, cluster=cl_sub, tol=1e-13, max_iter=200000)
        assert fb.res["n_obs"] == int(m.sum())
        assert fb.res["dof_fes"] == ref.res["dof_fes"] and fb.res["dof_resid"] == ref.res["dof_resid"], i
        assert rel(fb.beta, ref.beta) < 1e-8 and rel(fb.se, ref.se) < 1e-8, i
    pan.close()


def test_batch_beyond_32_slots_and_duplicates(monkeypatch):
    """40 weighted requests with one FE set, 20 subsamples each used twice: shared multi-slot sweeps
    (FEHIP_MULTI_SLOT=1); duplicates give identical outputs."""
    from fe_gpu_amd import api, synth
    monkeypatch.setenv("FEHIP_MULTI_SLOT", "1")
    from fe_gpu_amd.device import DeviceBuffer
    note("batch of 40 weighted requests")
    p = synth.random_fe_problem(405, n=5000, K=3, G=(1200, 700, 5))
    n = p.n
    rng = np.random.default_rng(9)
    pan = api.Panel.from_host(p.fe)
    dy = DeviceBuffer.from_array(p.y)
    dW = DeviceBuffer.from_array(np.asfortranarray(p.W))
    ws = [(rng.random(n) < 0.6).astype(np.uint8) for _ in range(20)]
    dws = [DeviceBuffer.from_array(w) for w in ws]
    reqs = [api.Request(dy.ptr, dW.ptr, 3, weights_ptr=dws[i % 20].ptr, options=dict(cluster=(1, 2), tol=1e-12,
                                                                                       max_iter=20000))
            for i in range(40)]
    fits = api.estimate_batch(pan, reqs)
    for i in range(20):
        a, b = fits[i], fits[i + 20]
        assert a.res["n_obs"] == b.res["n_obs"] == int(ws[i].sum())
        assert rel(a.beta, b.beta) < 1e-12 and rel(a.se, b.se) < 1e-12
    alone = api.estimate_device(pan, dy.ptr, dW.ptr, 3, weights_ptr=dws[5].ptr, cluster=(1, 2), tol=1e-12,
                                max_iter=20000)
    assert rel(fits[25].beta, alone.beta) < 1e-10 and rel(fits[25].se, alone.se) < 1e-10
    pan.close()


@pytest.mark.parametrize("workers", [1, 3])
def test_server_results_same_alone_and_in_the_mix(workers):
    """Requests submitted back to back are batched by the server (fewer batches than requests),
    with one worker or several whose batches overlap on the device; each result equals the same
    request solved alone."""
    import threading
    from fe_gpu_amd import api, synth
    from fe_gpu_amd.device import DeviceBuffer
    note(f"server mix workers={workers}")
    p = synth.random_fe_problem(406, n=6000, K=6, L=3, G=(1500, 800, 6))
    n = p.n
    rng = np.random.default_rng(11)
    specs = _specs(rng, n)
    pan = api.Panel.from_host(p.fe)
    dy = DeviceBuffer.from_array(p.y)
    dW = DeviceBuffer.from_array(np.asfortranarray(p.W))
    dZ = DeviceBuffer.from_array(np.asfortranarray(p.Z))
    dws = [DeviceBuffer.from_array(w) if w is not None else None for (_, _, _, _, w, _) in specs]
    reqs = []
    for (fm, j0, k, cl, w, iv), dw in zip(specs, dws):
        opt = dict(cluster=cl, fe_mask=fm, tol=1e-12, max_iter=20000)
        if iv:
            opt["iv_reg"] = [1, 3]
        reqs.append(api.Request(dy.ptr, dW.ptr + 8 * n * j0, k, dZ.ptr if iv else 0, 3 if iv else 0, n,
                                dw.ptr if dw else 0, opt))
    srv = api.Server(pan, max_batch=8, workers=workers)
    tickets = [None] * len(reqs)

    def client(idx):
        for i in idx:
            tickets[i] = srv.submit(reqs[i])

    th = [threading.Thread(target=client, args=(list(range(j, len(reqs), 2)),)) for j in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    outs = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["batches"] <= len(reqs) and st["p99_s"] >= st["p50_s"] > 0.0
    if workers == 1:
        assert st["batches"] < len(reqs)
    for i, (q, (fit, lat)) in enumerate(zip(reqs, outs)):
        alone = api.estimate_device(pan, q.y_ptr, q.W_ptr, q.K, q.Z_ptr, q.L, n, q.weights_ptr, **q.options)
        for key in INT_KEYS:
            assert fit.res[key] == alone.res[key], (i, key)
        assert rel(fit.beta, alone.beta) < 1e-10 and rel(fit.se, alone.se) < 1e-10, i
        assert lat > 0.0
    pan.close()


def test_server_rejects_bad_requests_and_drains_on_close():
    from fe_gpu_amd import api, synth
    from fe_gpu_amd._lib import FeGpuError
    from fe_gpu_amd.device import DeviceBuffer
    note("server errors")
    p = synth.random_fe_problem(407, n=3000, K=2, G=(300, 40))
    pan = api.Panel.from_host(p.fe)
    dy = DeviceBuffer.from_array(p.y)
    dW = DeviceBuffer.from_array(np.asfortranarray(p.W))
    srv = api.Server(pan, hbm_budget=1)          # budget below any request: each runs alone
    with pytest.raises(FeGpuError):
        srv.submit(api.Request(dy.ptr, dW.ptr, 2, ld=10))   # ld < n_obs
    ts = [srv.submit(api.Request(dy.ptr, dW.ptr, 2, options=dict(tol=1e-12))) for _ in range(4)]
    fits = [srv.wait(t)[0] for t in ts]
    st = srv.stats()
    assert st["batches"] == 4 and st["completed"] == 4
    assert all(np.array_equal(f.beta, fits[0].beta) or rel(f.beta, fits[0].beta) < 1e-12 for f in fits)
    srv.close()
    pan.close()


def test_chunks_sharing_right_hand_sides():
    """Unweighted requests over every FE subset of a 3-dim panel: the chunks that absorb the primary dim
    share one right-hand-side pass over the union of their dims and columns, and so do the chunks that do
    not.  Each request, FP64 or FP32 storage, equals the same request solved alone (1e-10; FP32 storage
    1e-6 / 1e-5) and the oracle on its FE dims (1e-8; FP32 storage 1e-6 / 1e-5); a weighted request in the
    same batch keeps its own chunk; two requests repeated at the end get their originals' results."""
    import oracle
    from fe_gpu_amd import api, synth
    from fe_gpu_amd.device import DeviceBuffer
    note("batch of chunks sharing right-hand sides")
    p = synth.random_fe_problem(406, n=8000, K=5, G=(1500, 800, 9))
    n = p.n
    rng = np.random.default_rng(11)
    pan = api.Panel.from_host(p.fe)
    dy = DeviceBuffer.from_array(p.y)
    dW = DeviceBuffer.from_array(np.asfortranarray(p.W))
    w = (rng.random(n) < 0.5).astype(np.uint8)
    dw = DeviceBuffer.from_array(w)
    specs = []
    for fm in range(1, 8):
        dims = [d + 1 for d in range(3) if fm >> d & 1]
        for prec in (0, 1):
            k = int(rng.integers(1, 6))
            j0 = int(rng.integers(0, 5 - k + 1))
            ncl = int(rng.integers(0, len(dims) + 1))
            cl = tuple(sorted(rng.choice(dims, size=ncl, replace=False).tolist())) if ncl else ()
            specs.append((fm, j0, k, cl, prec, False))
    specs.append((0b011, 0, 3, (1,), 0, True))
    specs += [specs[3], specs[10]]  # requests equal to earlier ones are served from their computation
    reqs = [api.Request(dy.ptr, dW.ptr + 8 * n * j0, k, 0, 0, n, dw.ptr if wt else 0,
                        dict(cluster=cl, fe_mask=fm, tol=1e-13, max_iter=20000, precision=prec))
            for (fm, j0, k, cl, prec, wt) in specs]
    fits = api.estimate_batch(pan, reqs)
    for i, ((fm, j0, k, cl, prec, wt), q, fb) in enumerate(zip(specs, reqs, fits)):
        note(f"request {i}: fe_mask={fm:03b} K={k} cluster={cl} precision={prec} weighted={wt}")
        fa = api.estimate_device(pan, q.y_ptr, q.W_ptr, q.K, 0, 0, n, q.weights_ptr, **q.options)
        for key in INT_KEYS:
            assert fb.res[key] == fa.res[key], (i, key)
        bt, st = (1e-10, 1e-10) if prec == 0 else (1e-6, 1e-5)
        assert rel(fb.beta, fa.beta) < bt and rel(fb.se, fa.se) < st, i
        m = np.ones(n, bool) if not wt else w == 1
        dims = [d for d in range(3) if fm >> d & 1]
        fe_sub = np.vstack([p.fe[d] if not wt else _redense(p.fe[d][m]) for d in dims])
        cl_sub = tuple(dims.index(c - 1) + 1 for c in cl)
        ref = oracle.estimate(p.y[m], p.W[m][:, j0:j0 + k], fe_sub, cluster=cl_sub, tol=1e-13, max_iter=200000)
        assert fb.res["dof_fes"] == ref.res["dof_fes"] and fb.res["dof_resid"] == ref.res["dof_resid"], i
        bt, st = (1e-8, 1e-8) if prec == 0 else (1e-6, 1e-5)
        assert rel(fb.beta, ref.beta) < bt and rel(fb.se, ref.se) < st, i
    for dup, orig in ((len(specs) - 2, 3), (len(specs) - 1, 10)):
        assert np.array_equal(fits[dup].beta, fits[orig].beta) and np.array_equal(fits[dup].se, fits[orig].se)
        assert np.array_equal(fits[dup].vcov, fits[orig].vcov)
    pan.close()
