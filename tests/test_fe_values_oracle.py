"""Pins of the oracle's FE reconstruction (reconstruct_and_save_fixed_effects restated): a single dim
has the closed form 'group means of r minus their row-weighted mean'; on a small connected design the
fitted part equals the least-squares projection of r on the FE dummies minus the mean of r (dense
brute force), and each dim's values have row-weighted mean zero."""
import numpy as np

import oracle


def test_single_dim_closed_form():
    rng = np.random.default_rng(1)
    n = 500
    fe = rng.integers(1, 30, (1, n)).astype(np.int32)
    fe[0, 0] = 29
    W = rng.normal(size=(n, 2))
    beta = np.array([0.5, -1.0])
    y = rng.normal(size=n)
    (v, c), = oracle.reconstruct_fe(y, W, beta, fe, tol=1e-14)[0]
    r = y - W @ beta
    g = fe[0] - 1
    means = np.bincount(g, weights=r, minlength=29) / np.maximum(np.bincount(g, minlength=29), 1)
    cnt = np.bincount(g, minlength=29)
    expect = np.where(cnt > 0, means - (means * cnt).sum() / cnt.sum(), 0.0)
    assert np.array_equal(c, cnt)
    assert np.max(np.abs(v - expect)) < 1e-13


def test_fitted_part_is_the_centred_projection():
    rng = np.random.default_rng(2)
    n = 400
    fe = np.vstack([rng.integers(1, 41, n), rng.integers(1, 13, n), rng.integers(1, 6, n)]).astype(np.int32)
    for d, g in enumerate((40, 12, 5)):
        fe[d, d] = g
    y = rng.normal(size=n) + rng.normal(size=40)[fe[0] - 1]
    out, it = oracle.reconstruct_fe(y, None, [], fe, tol=1e-13, max_iter=100000)
    fitted = sum(out[d][0][fe[d] - 1] for d in range(3))
    D = np.hstack([np.eye(g)[fe[d] - 1] for d, g in enumerate((40, 12, 5))])
    proj = D @ np.linalg.lstsq(D, y, rcond=None)[0]
    assert np.max(np.abs(fitted - (proj - y.mean()))) < 1e-9
    for v, c in out:
        assert abs((v * c).sum()) < 1e-9 * max(1.0, np.abs(v).max() * c.sum())
