"""GPU parity of the mixed-precision form (precision=1): the within-transformed columns are stored in
FP32.  Their cross products are accumulated in FP64 by the transform's final pass before the rounding,
so beta is the FP64 solution; where a request's columns span several column tiles, beta is instead
refined in FP64 on the residual M_F(y - W beta), formed from the caller's FP64 columns and the solved
FE tables.  FP64 is kept for the Cholesky solve, the 2SLS projection and the sandwich.

Bar (the stated accuracy target): beta within 1e-6 and SE within 1e-5 relative of the FP64 path
(and of the oracle); integer outputs identical to the FP64 path.
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


INT_KEYS = ("n_obs", "dof_fes", "dof_resid", "dof_model", "rank", "is_iv", "n_instruments", "n_clusters",
            "psd_fix_applied", "converged")


@pytest.mark.parametrize("case", ["ols3", "ols_cl2", "iv", "weighted"])
def test_mixed_precision_matches_fp64_and_oracle(case):
    import oracle
    from fe_gpu_amd import api, synth
    note(f"mixed precision case={case}")
    p = synth.random_fe_problem(606, n=20000, K=5, L=3, G=(3000, 800, 12))
    # scale the columns apart so FP32 storage is exercised away from unit magnitude
    W = p.W * np.array([1.0, 250.0, 0.01, 7.0, 1.0])
    y = p.y + 3.0 * W[:, 1] / 250.0
    kw = dict(tol=1e-12, max_iter=20000)
    w = None
    if case == "ols3":
        kw["cluster"] = (1, 2, 3)
    elif case == "ols_cl2":
        kw["cluster"] = (2,)
    elif case == "iv":
        kw.update(cluster=(1,), iv_reg=[2, 4])
    else:
        w = (np.random.default_rng(3).random(p.n) < 0.4).astype(np.uint8)
        kw.update(cluster=(1, 2), fe_mask=0b011)
    Z = p.Z if case == "iv" else None
    f64 = api.estimate(y, W, p.fe, Z=Z, weights=w, precision=0, **kw)
    f32 = api.estimate(y, W, p.fe, Z=Z, weights=w, precision=1, **kw)
    for k in INT_KEYS:
        assert f32.res[k] == f64.res[k], k
    eb, es = rel(f32.beta, f64.beta), rel(f32.se, f64.se)
    note(f"  beta rel err {eb:.2e}, se rel err {es:.2e}")
    assert eb <= 1e-6 and es <= 1e-5
    assert abs(f32.res["rss"] - f64.res["rss"]) <= 1e-9 * f64.res["rss"]
    if case != "weighted":
        ref = oracle.estimate(y, W, p.fe, Z=Z, cluster=kw.get("cluster", ()), iv_reg=kw.get("iv_reg", ()),
                              tol=1e-13, max_iter=200000)
        assert rel(f32.beta, ref.beta) <= 1e-6 and rel(f32.se, ref.se) <= 1e-5


def test_mixed_precision_in_a_batch_equals_alone():
    from fe_gpu_amd import api, synth
    from fe_gpu_amd.device import DeviceBuffer
    note("mixed precision in a batch")
    p = synth.random_fe_problem(607, n=8000, K=4, G=(1500, 600, 9))
    n = p.n
    pan = api.Panel.from_host(p.fe)
    dy = DeviceBuffer.from_array(p.y)
    dW = DeviceBuffer.from_array(np.asfortranarray(p.W))
    reqs = [api.Request(dy.ptr, dW.ptr, 4, options=dict(cluster=(1, 3), tol=1e-11, precision=1)),
            api.Request(dy.ptr, dW.ptr + 8 * n, 3, options=dict(cluster=(2,), tol=1e-11, precision=0)),
            api.Request(dy.ptr, dW.ptr, 2, options=dict(tol=1e-11, precision=1, fe_mask=0b101))]
    fits = api.estimate_batch(pan, reqs)
    for q, fb in zip(reqs, fits):
        fa = api.estimate_device(pan, q.y_ptr, q.W_ptr, q.K, **q.options)
        assert rel(fb.beta, fa.beta) < 1e-10 and rel(fb.se, fa.se) < 1e-10
    pan.close()


@pytest.mark.parametrize("tol", [1e-8, 1e-12])
def test_fp32_operator_solve_meets_the_fp64_criterion(tol):
    """On a worker-firm-period panel (every primary segment in wave tiles) the mixed-precision solve runs
    its CG on the FP32 operator, with reliable updates from FP64 products of the true residual between
    rounds.  Its stopping rule is checked on that FP64 residual, so it ends where the FP64 solve ends:
    converged, the FE-space criterion below tol, and beta / SE of the FP32-storage fit within the stated
    bars of the FP64 fit (and of the oracle at the tight tolerance)."""
    import oracle
    from fe_gpu_amd import api, synth
    note(f"FP32 operator tol={tol}")
    # 5,000 firms: a table space large enough for the FP32 operator (small ones keep the FP64 operator)
    p = synth.worker_firm_panel(seed=41, n_workers=6000, n_periods=15, n_firms=5000)
    kw = dict(cluster=(1, 2, 3), tol=tol, max_iter=20000)
    f64 = api.estimate(p.y, p.W, p.fe, precision=0, **kw)
    f32 = api.estimate(p.y, p.W, p.fe, precision=1, **kw)
    assert f64.res["fe_refinements"] == 0
    assert f32.res["fe_refinements"] >= 1          # the FP32-operator path ran
    assert f32.res["converged"] == 1 and f32.res["max_update"] < tol
    for k in INT_KEYS:
        assert f32.res[k] == f64.res[k], k
    eb, es = rel(f32.beta, f64.beta), rel(f32.se, f64.se)
    note(f"  refinements {f32.res['fe_refinements']}, steps {f32.res['fe_iterations']} vs {f64.res['fe_iterations']}, "
         f"beta rel err {eb:.2e}, se rel err {es:.2e}")
    assert eb <= 1e-6 and es <= 1e-5
    if tol < 1e-10:
        ref = oracle.estimate(p.y, p.W, p.fe, cluster=(1, 2, 3), tol=1e-13, max_iter=200000)
        assert rel(f32.beta, ref.beta) <= 1e-8 and rel(f32.se, ref.se) <= 1e-5


def test_small_or_diagonal_systems_keep_the_fp64_operator():
    """A mixed-precision request whose FE-space system is diagonal (one absorbed dim besides the
    unabsorbed primary) or small takes no FP32 rounds (no FP64 refinement products), and equals the
    FP64 path."""
    from fe_gpu_amd import api, synth
    note("small / diagonal systems")
    p = synth.worker_firm_panel(seed=42, n_workers=4000, n_periods=10, n_firms=3000)
    for fe_mask in (0b010, 0b101):   # {firm} without the primary; {worker, period}: 10 table entries
        kw = dict(cluster=(2,) if fe_mask == 0b010 else (1,), fe_mask=fe_mask, tol=1e-10, max_iter=20000)
        f64 = api.estimate(p.y, p.W, p.fe, precision=0, **kw)
        f32 = api.estimate(p.y, p.W, p.fe, precision=1, **kw)
        assert f32.res["fe_refinements"] == 0
        assert rel(f32.beta, f64.beta) <= 1e-6 and rel(f32.se, f64.se) <= 1e-5


def test_registered_fp32_columns():
    """FP32-storage requests read the FP32 copies of registered columns in their right-hand-side and final
    passes (half the bytes): alone and in a batch they stay within the FP32 bars of the FP64 fits (beta 1e-6,
    SE 1e-5) and of the oracle, differ from the unregistered FP32 fits only by the inputs' FP32 rounding, and a
    column changed and registered again gives the fit of the new data."""
    import oracle
    from fe_gpu_amd import api, synth
    from fe_gpu_amd.device import DeviceBuffer
    note("registered FP32 columns")
    p = synth.worker_firm_panel(seed=43, n_workers=5000, n_periods=12, n_firms=4000)
    n = p.n
    K = p.W.shape[1]
    dy = DeviceBuffer.from_array(p.y)
    dW = DeviceBuffer.from_array(np.asfortranarray(p.W))
    plain = api.Panel.from_host(p.fe)
    pan = api.Panel.from_host(p.fe)
    pan.register_columns(dy.ptr, 1)
    pan.register_columns(dW.ptr, K)
    specs = [(0b111, 0, K, (1, 2, 3)), (0b011, 1, 4, (2,)), (0b110, 0, 3, (2, 3)), (0b101, 2, 5, ()), (0b010, 0, 2, (2,))]
    reqs = [api.Request(dy.ptr, dW.ptr + 8 * n * j0, k, options=dict(cluster=cl, fe_mask=fm, tol=1e-12, max_iter=20000,
                                                                     precision=1))
            for (fm, j0, k, cl) in specs]
    batch = api.estimate_batch(pan, reqs)
    moved = 0.0
    for (fm, j0, k, cl), q, fb in zip(specs, reqs, batch):
        fa = api.estimate_device(pan, q.y_ptr, q.W_ptr, q.K, **q.options)
        f64 = api.estimate_device(plain, q.y_ptr, q.W_ptr, q.K, **dict(q.options, precision=0))
        fu = api.estimate_device(plain, q.y_ptr, q.W_ptr, q.K, **q.options)
        assert rel(fb.beta, fa.beta) < 1e-10 and rel(fb.se, fa.se) < 1e-10
        assert rel(fa.beta, f64.beta) <= 1e-6 and rel(fa.se, f64.se) <= 1e-5
        assert rel(fa.beta, fu.beta) <= 1e-6 and rel(fa.se, fu.se) <= 1e-5
        moved = max(moved, rel(fa.beta, fu.beta))
        dims = [d for d in range(3) if fm >> d & 1]
        ref = oracle.estimate(p.y, p.W[:, j0:j0 + k], p.fe[dims], cluster=tuple(dims.index(c - 1) + 1 for c in cl),
                              tol=1e-12, max_iter=200000)
        assert rel(fa.beta, ref.beta) <= 1e-6 and rel(fa.se, ref.se) <= 1e-5
    assert moved > 1e-12  # the registered copies were read (their rounding moves beta)
    # change the first regressor on the device, register again: the fit follows the new data
    W2 = p.W.copy()
    W2[:, 0] *= 2.0
    dW.copy_from(np.asfortranarray(W2))
    pan.register_columns(dW.ptr, K)
    fa = api.estimate_device(pan, dy.ptr, dW.ptr, K, cluster=(1,), tol=1e-12, max_iter=20000, precision=1)
    f64 = api.estimate_device(plain, dy.ptr, dW.ptr, K, cluster=(1,), tol=1e-12, max_iter=20000, precision=0)
    assert rel(fa.beta, f64.beta) <= 1e-6 and rel(fa.se, f64.se) <= 1e-5
    pan.close()
    plain.close()
