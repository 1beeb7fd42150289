"""GPU parity of the cross products formed in the final pass of the within-transform.

The normal equations (and, with FP32 storage, beta itself) come from the FP64 Gram matrix that the
final pass accumulates on the matrix cores while it writes the transformed columns.  The pass runs in
three kernels by primary-segment length (wave tiles up to 64 rows, block tiles up to 256 rows, one
thread per longer segment), so the panels below put rows in each, and the column counts cover tile
widths from 2 to 16.  Bar: FP64 beta/SE within 1e-8 of the oracle, tss_within and rss within 1e-10;
FP32 storage: beta within 1e-6 and SE within 1e-5 of the oracle.
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


def panel_with_segments(seed, shape, n=24000, K=5):
    """Random problem whose primary (largest) FE dim has segments of the given kind."""
    from fe_gpu_amd import synth
    rng = np.random.default_rng(seed)
    if shape == "short":
        G = (3000, 700, 9)
        p = synth.random_fe_problem(seed, n=n, K=K, G=G)
        return p
    if shape == "medium":      # ~150 rows per primary group: block tiles
        sizes = np.full(160, n // 160)
    elif shape == "long":      # ~800 rows per primary group: one thread per segment
        sizes = np.full(30, n // 30)
    else:                      # every kind in one panel
        sizes = np.concatenate([np.full(6, 700), np.full(40, 150), np.full(1000, 12)])
    sizes[-1] += n - sizes.sum()
    g0 = np.repeat(np.arange(1, len(sizes) + 1), sizes)
    rng.shuffle(g0)
    g1 = rng.integers(1, 21, n)
    g1[0] = 20
    g2 = rng.integers(1, 8, n)
    g2[1] = 7
    fe = np.vstack([g0, g1, g2]).astype(np.int32)
    X = rng.normal(size=(n, K))
    y = X @ rng.normal(size=K) + rng.normal(size=len(sizes))[g0 - 1] + rng.normal(size=20)[g1 - 1] \
        + 0.3 * rng.normal(size=n)
    return synth.Panel(y, X, fe, [f"x{j}" for j in range(K)], ["a", "b", "c"], None, None)


@pytest.mark.parametrize("shape", ["short", "medium", "long", "mixed"])
@pytest.mark.parametrize("precision", [0, 1])
def test_final_pass_cross_products_match_oracle(shape, precision):
    import oracle
    from fe_gpu_amd import api
    note(f"final-pass cross products shape={shape} precision={precision}")
    p = panel_with_segments(700 + len(shape), shape)
    fit = api.estimate(p.y, p.W, p.fe, cluster=(1, 2), tol=1e-12, max_iter=20000, precision=precision)
    ref = oracle.estimate(p.y, p.W, p.fe, cluster=(1, 2), tol=1e-13, max_iter=200000)
    eb, es = rel(fit.beta, ref.beta), rel(fit.se, ref.se)
    note(f"  beta rel err {eb:.2e}, se rel err {es:.2e}")
    if precision == 0:
        assert eb <= 1e-8 and es <= 1e-8
        assert abs(fit.res["tss_within"] - ref.res["tss_within"]) <= 1e-10 * ref.res["tss_within"]
    else:
        assert eb <= 1e-6 and es <= 1e-5
    assert abs(fit.res["rss"] - ref.res["rss"]) <= 1e-10 * ref.res["rss"]


@pytest.mark.parametrize("K", [1, 7, 11, 15])
def test_final_pass_cross_products_tile_widths(K):
    """1 + K columns in one tile of width 2, 8, 12 and 16."""
    import oracle
    from fe_gpu_amd import api, synth
    note(f"final-pass cross products K={K}")
    p = synth.random_fe_problem(720 + K, n=12000, K=K, G=(2500, 300, 11))
    fit = api.estimate(p.y, p.W, p.fe, cluster=(1,), tol=1e-12, max_iter=20000)
    ref = oracle.estimate(p.y, p.W, p.fe, cluster=(1,), tol=1e-13, max_iter=200000)
    assert rel(fit.beta, ref.beta) <= 1e-8 and rel(fit.se, ref.se) <= 1e-8
    assert abs(fit.res["tss_within"] - ref.res["tss_within"]) <= 1e-10 * ref.res["tss_within"]


def test_batch_cross_products_with_row_masks_equal_alone():
    """Requests with their own row subsamples share one tile; each takes its block of the cross
    products (rows outside a column's subsample count as zero)."""
    from fe_gpu_amd import api
    from fe_gpu_amd.device import DeviceBuffer
    note("batch cross products with row masks")
    p = panel_with_segments(731, "mixed", K=6)
    n = p.n
    rng = np.random.default_rng(5)
    pan = api.Panel.from_host(p.fe)
    dy = DeviceBuffer.from_array(p.y)
    dW = DeviceBuffer.from_array(np.asfortranarray(p.W))
    ws = [DeviceBuffer.from_array((rng.random(n) < f).astype(np.uint8)) for f in (0.5, 0.8)]
    reqs = [api.Request(dy.ptr, dW.ptr, 3, weights_ptr=ws[0].ptr, options=dict(cluster=(1,), tol=1e-12)),
            api.Request(dy.ptr, dW.ptr + 8 * n, 4, weights_ptr=ws[1].ptr, options=dict(cluster=(2,), tol=1e-12)),
            api.Request(dy.ptr, dW.ptr + 8 * n * 2, 4, options=dict(tol=1e-12, precision=1))]
    fits = api.estimate_batch(pan, reqs)
    for q, fb in zip(reqs, fits):
        fa = api.estimate_device(pan, q.y_ptr, q.W_ptr, q.K, weights_ptr=q.weights_ptr, **q.options)
        assert rel(fb.beta, fa.beta) < 1e-10 and rel(fb.se, fa.se) < 1e-10
        assert abs(fb.res["tss_within"] - fa.res["tss_within"]) <= 1e-12 * fa.res["tss_within"]
    pan.close()


def test_wide_requests_in_a_batch_take_the_matrix_core_gram():
    """Two requests of 13 and 16 columns with their own row subsamples: their 29 distinct columns span
    two column tiles, so each request's cross products come from the 16-column matrix-core Gram of its
    own columns; they must equal the same requests alone (cross products from the final pass)."""
    from fe_gpu_amd import api, synth
    from fe_gpu_amd.device import DeviceBuffer
    note("wide requests in a batch")
    p = synth.random_fe_problem(741, n=9000, K=15, G=(1800, 200, 9))
    n = p.n
    rng = np.random.default_rng(8)
    pan = api.Panel.from_host(p.fe)
    dy = DeviceBuffer.from_array(p.y)
    dW = DeviceBuffer.from_array(np.asfortranarray(p.W))
    ws = [DeviceBuffer.from_array((rng.random(n) < f).astype(np.uint8)) for f in (0.7, 0.9)]
    reqs = [api.Request(dy.ptr, dW.ptr, 12, weights_ptr=ws[0].ptr, options=dict(cluster=(1,), tol=1e-12)),
            api.Request(dy.ptr, dW.ptr, 15, weights_ptr=ws[1].ptr, options=dict(cluster=(2,), tol=1e-12))]
    fits = api.estimate_batch(pan, reqs)
    for q, fb in zip(reqs, fits):
        fa = api.estimate_device(pan, q.y_ptr, q.W_ptr, q.K, weights_ptr=q.weights_ptr, **q.options)
        assert rel(fb.beta, fa.beta) < 1e-10 and rel(fb.se, fa.se) < 1e-10
        assert abs(fb.res["tss_within"] - fa.res["tss_within"]) <= 1e-12 * fa.res["tss_within"]
    pan.close()
