"""Pins the CPU oracle (no GPU).

Checks come from (i) the reference's own tests (tests/test_gpu_regression.f90,
tests/test_gpu_demean.f90), (ii) the reference's own C intersection-id builder
compiled from source into oracle/_ref, and (iii) what the mathematics fixes:
dense dummy-variable least squares, np.unique partitions, LAPACK through numpy,
and structural identities of the CGM estimator.
"""
import numpy as np
import pytest

import oracle
from fe_gpu_amd import synth
from tests import bruteforce as bf


def rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))) if a.size else 0.0


# --- reference known answers -------------------------------------------------

def _ref_regression_case():
    # tests/test_gpu_regression.f90:49-83
    W = np.array([[1.0, -2.0], [1.5, 0.5], [-0.5, 1.0], [-1.0, -1.5], [0.25, 0.75], [2.0, 1.25]])
    fe = np.array([[1, 1, 2, 2, 3, 3], [1, 2, 1, 2, 1, 2]], dtype=np.int32)
    fe1 = np.array([0.5, -0.25, 0.8])
    fe2 = np.array([-0.4, 0.2])
    beta = np.array([1.5, -0.75])
    y = W @ beta + fe1[fe[0] - 1] + fe2[fe[1] - 1]
    return y, W, fe, beta


@pytest.mark.parametrize("conv,ap", [("gpu", "relaxed"), ("gpu", "forward"), ("cpu", "forward")])
def test_known_answer_regression(conv, ap):
    """tests/test_gpu_regression.f90:17,30,40: beta = (1.5, -0.75) within 1e-8 at fe_tol 1e-10."""
    y, W, fe, beta = _ref_regression_case()
    fit = oracle.estimate(y, W, fe, tol=1e-10, max_iter=500, conv=conv, ap=ap)
    assert fit.res["solver_info"] == 0
    assert np.max(np.abs(fit.beta - beta)) < 1e-8


def test_known_answer_demean_invariant():
    """tests/test_gpu_demean.f90:30,43-84: every group mean of y and W below 1e-6 at tol 1e-8/200 its."""
    y = np.array([1.0, 2.0, 3.0, -2.0, -1.0, 0.5])
    W = np.array([[1.0, 0.0], [1.5, -2.0], [-0.5, 1.0], [-1.0, 0.5], [2.0, 1.5], [-0.5, -1.0]])
    fe = np.array([[1, 1, 2, 2, 3, 3], [1, 2, 1, 2, 1, 2]], dtype=np.int32)
    G = np.array([3, 2], dtype=np.int32)
    yt, Wt, Zt, it, conv, mu = oracle.demean_relaxed(y, W, None, fe, G, tol=1e-8, max_iter=200)
    assert conv
    for d in range(2):
        for g in range(1, G[d] + 1):
            m = fe[d] == g
            assert abs(yt[m].mean()) < 1e-6
            assert np.all(np.abs(Wt[m].mean(axis=0)) < 1e-6)


# --- demeaning: fixed point is the exact within-transform ---------------------

@pytest.mark.parametrize("seed,G", [(1, (7, 5)), (2, (11, 4, 3)), (3, (30, 9, 6, 2))])
def test_demean_fixed_point_matches_dense_projection(seed, G):
    p = synth.random_fe_problem(seed, n=300, K=3, L=2, G=G)
    exact_y = bf.annihilate(p.y, p.fe)
    exact_W = bf.annihilate(p.W, p.fe)
    exact_Z = bf.annihilate(p.Z, p.fe)
    yt, Wt, it, conv, mu = oracle.demean_forward(p.y, p.W, p.fe, p.G, tol=1e-13, max_iter=20000)
    assert conv
    assert np.max(np.abs(yt - exact_y)) < 1e-9
    assert np.max(np.abs(Wt - exact_W)) < 1e-9
    yr, Wr, Zr, it2, conv2, mu2 = oracle.demean_relaxed(p.y, p.W, p.Z, p.fe, p.G, tol=1e-13, max_iter=20000)
    assert conv2
    assert np.max(np.abs(yr - exact_y)) < 1e-9
    assert np.max(np.abs(Wr - exact_W)) < 1e-9
    assert np.max(np.abs(Zr - exact_Z)) < 1e-9


def test_forward_iterations_follow_fortran_do_semantics():
    p = synth.random_fe_problem(4, n=200, K=2, G=(20, 15))
    _, _, it, conv, _ = oracle.demean_forward(p.y, p.W, p.fe, p.G, tol=1e-300, max_iter=3)
    assert not conv and it == 4     # DO variable after normal completion = max_iter + 1


def test_one_dim_demean_is_group_centering():
    rng = np.random.default_rng(5)
    fe = rng.integers(1, 9, size=(1, 100)).astype(np.int32)
    fe[0, 0] = 8
    y = rng.normal(size=100)
    yt, _, it, conv, _ = oracle.demean_forward(y, np.zeros((100, 0)), fe, [8], tol=1e-12)
    means = np.bincount(fe[0] - 1, weights=y) / np.bincount(fe[0] - 1)
    assert np.allclose(yt, y - means[fe[0] - 1], atol=1e-14, rtol=0)


# --- group sizes ----------------------------------------------------------------

def test_group_sizes():
    fe = np.array([[3, 1, 2], [5, 5, 1]], dtype=np.int32)
    st, G = oracle.group_sizes(fe)
    assert st == 0 and list(G) == [3, 5]
    st, _ = oracle.group_sizes(np.array([[1, 0, 2]], dtype=np.int32))
    assert st == 1


# --- intersection ids ------------------------------------------------------------

def _keys_rank(fe, dims, sizes):
    mult = np.cumprod([1] + list(sizes[:-1])).astype(np.uint64)
    key = np.zeros(fe.shape[1], dtype=np.uint64)
    for j, d in enumerate(dims):
        key += mult[j] * (fe[d - 1].astype(np.uint64) - np.uint64(1))
    _, inv = np.unique(key, return_inverse=True)
    return inv.astype(np.int32) + 1


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_cluster_ids_sort_path_is_dense_key_rank(seed):
    rng = np.random.default_rng(seed)
    n = 5000
    G = [50, 7, 300]
    fe = np.vstack([rng.integers(1, g + 1, n) for g in G]).astype(np.int32)
    for dims in ([1, 2], [2, 3], [1, 3], [1, 2, 3], [3, 1]):
        sizes = [G[d - 1] for d in dims]
        st, ids, nc = oracle.cluster_ids(fe, dims, sizes)
        assert st == 0
        expect = _keys_rank(fe, dims, sizes)
        assert np.array_equal(ids, expect)
        assert nc == expect.max()


def _first_appearance(fe, dims):
    seen = {}
    out = np.empty(fe.shape[1], dtype=np.int32)
    for i in range(fe.shape[1]):
        k = tuple(int(fe[d - 1, i]) for d in dims)
        if k not in seen:
            seen[k] = len(seen) + 1
        out[i] = seen[k]
    return out


def test_cluster_ids_hash_path_first_appearance():
    rng = np.random.default_rng(7)
    n = 3000
    big = 2 ** 21 + 5           # 3 dims of this size: span > 2^62 -> hash path
    fe = np.vstack([rng.integers(1, 40, n) for _ in range(3)]).astype(np.int32)
    for d in range(3):
        fe[d, rng.integers(0, n)] = big
    sizes = [big, big, big]
    st, ids, nc = oracle.cluster_ids(fe, [1, 2, 3], sizes)
    assert st == 0
    expect = _first_appearance(fe, [1, 2, 3])
    assert np.array_equal(ids, expect) and nc == expect.max()


@pytest.mark.parametrize("seed,big", [(0, False), (1, False), (2, True)])
def test_cluster_ids_match_reference_builder(seed, big):
    """Bit-exact against src/core/cluster_id_builder.c compiled from the reference sources.

    The compiled builder is looked up inside the test (not at collection), so a run that only
    collects this module, such as a GPU-marked selection, never maps it.
    """
    if oracle.ref_lib() is None:
        pytest.skip("reference tree absent: oracle/_ref not built")
    rng = np.random.default_rng(seed)
    n = 70000 if not big else 4000          # > 2^16 exercises the reference's threaded radix path
    G = [900, 13, 12] if not big else [2 ** 21 + 3] * 3
    fe = np.vstack([rng.integers(1, min(g, 500) + 1, n) for g in G]).astype(np.int32)
    for d, g in enumerate(G):
        fe[d, rng.integers(0, n)] = g
    fe_il = np.ascontiguousarray(fe.T)      # (n, n_fe): the host's interleaved (n_fe, N) Fortran layout
    for dims in ([1, 2], [1, 3], [2, 3], [1, 2, 3]):
        sizes = [G[d - 1] for d in dims]
        a = oracle.cluster_ids_interleaved(fe_il, dims, sizes, use_ref=True)
        b = oracle.cluster_ids_interleaved(fe_il, dims, sizes, use_ref=False)
        c = oracle.cluster_ids(fe, dims, sizes)
        assert a[0] == b[0] == c[0] == 0
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[1], c[1])
        assert a[2] == b[2] == c[2]


def test_cluster_ids_status_codes_match_reference_contract():
    fe = np.ones((1, 4), dtype=np.int32)
    st, _, _ = oracle.cluster_ids_interleaved(np.ascontiguousarray(fe.T), [1, 1, 1, 1, 1], [1] * 5)
    assert st == 3


# --- dense solver ---------------------------------------------------------------

def test_cholesky_solve_and_inverse_against_lapack():
    rng = np.random.default_rng(11)
    A = rng.normal(size=(9, 9))
    Q = A @ A.T + 0.5 * np.eye(9)
    b = rng.normal(size=9)
    info, beta, Qi = oracle.chol_solve_and_invert(Q, b)
    assert info == 0
    assert rel(beta, np.linalg.solve(Q, b)) < 1e-11
    assert np.max(np.abs(Qi - np.linalg.inv(Q))) < 1e-11 * np.max(np.abs(np.linalg.inv(Q)))


def test_cholesky_reports_first_nonpositive_pivot():
    Q = np.array([[1.0, 2.0, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    info, _, _ = oracle.chol_solve_and_invert(Q, np.ones(3))
    assert info == 2        # dpotrf reports the order of the leading minor that is not positive


def test_sym_eig_against_eigh():
    rng = np.random.default_rng(12)
    A = rng.normal(size=(8, 8))
    A = A + A.T
    w, V = oracle.sym_eig(A)
    assert np.allclose(w, np.linalg.eigvalsh(A), atol=1e-12, rtol=0)
    assert np.allclose(V @ np.diag(w) @ V.T, A, atol=1e-11)


# --- full estimator ----------------------------------------------------------------

@pytest.mark.parametrize("cluster", [(), (1,), (2,), (1, 2), (1, 2, 3), (3, 1)])
def test_estimate_matches_brute_force(cluster):
    p = synth.random_fe_problem(21, n=400, K=3, G=(25, 8, 6))
    fit = oracle.estimate(p.y, p.W, p.fe, cluster=cluster, tol=1e-13, max_iter=20000)
    ref = bf.fit(p.y, p.W, p.fe, cluster=cluster)
    assert rel(fit.beta, ref["beta"]) < 1e-9
    assert rel(fit.se, ref["se"]) < 1e-8
    assert fit.res["dof_fes"] == ref["dof_fes"] == 25 + 8 + 6 - 3
    assert fit.res["dof_resid"] == ref["dof_resid"]
    assert abs(fit.res["rss"] - ref["rss"]) < 1e-9 * ref["rss"]


def test_iv_matches_closed_form_2sls():
    rng = np.random.default_rng(31)
    n = 600
    fe = np.vstack([rng.integers(1, 21, n), rng.integers(1, 6, n)]).astype(np.int32)
    fe[0, 0], fe[1, 0] = 20, 5
    z = rng.normal(size=(n, 3))
    u = rng.normal(size=n)
    xend = z @ np.array([0.6, -0.4, 0.3]) + 0.5 * u + rng.normal(size=n)
    xexo = rng.normal(size=(n, 2))
    X = np.column_stack([xend, xexo])
    y = X @ np.array([1.0, -0.5, 0.25]) + u + np.arange(21)[fe[0]] * 0.1
    for cluster in [(), (1,), (1, 2)]:
        fit = oracle.estimate(y, X, fe, Z=z, iv_reg=[1], cluster=cluster, tol=1e-13, max_iter=20000)
        ref = bf.fit(y, X, fe, cluster=cluster, Z=z, endog=[1])
        assert fit.res["is_iv"] == 1 and fit.res["n_instruments"] == 2 + 3
        assert rel(fit.beta, ref["beta"]) < 1e-9
        assert rel(fit.se, ref["se"]) < 1e-8


def test_iv_reverts_to_ols_with_too_few_instruments():
    p = synth.random_fe_problem(32, n=300, K=3, L=1, G=(9, 4))
    fit = oracle.estimate(p.y, p.W, p.fe, Z=p.Z, iv_reg=[1, 2], tol=1e-12)
    ols = oracle.estimate(p.y, p.W, p.fe, tol=1e-12)
    assert fit.res["is_iv"] == 0
    assert np.allclose(fit.beta, ols.beta, rtol=1e-12, atol=0)


def test_column_filter_drops_within_constant_regressor():
    p = synth.random_fe_problem(33, n=300, K=3, G=(12, 5))
    W = p.W.copy()
    W[:, 1] = np.arange(1, 13)[p.fe[0] - 1] * 0.7      # constant within dim-1 groups
    fit = oracle.estimate(p.y, W, p.fe, tol=1e-13, max_iter=20000)
    assert fit.res["dof_model"] == 2 and fit.res["rank"] == 2
    assert fit.beta[1] == 0.0 and fit.se[1] == 0.0
    ref = bf.fit(p.y, W[:, [0, 2]], p.fe)
    assert rel(fit.beta[[0, 2]], ref["beta"]) < 1e-9


def test_cgm_structural_identities():
    """Inclusion-exclusion fixes the subset weights: a dim nested in another makes the
    two-way VCE equal the coarser one-way VCE (up to the scalar's G_min factor)."""
    rng = np.random.default_rng(41)
    n = 800
    fine = rng.integers(1, 81, n).astype(np.int32)
    fine[0] = 80
    coarse = ((fine - 1) // 10 + 1).astype(np.int32)       # fine nested in coarse (8 groups)
    other = rng.integers(1, 5, n).astype(np.int32)
    other[0] = 4
    fe = np.vstack([fine, coarse, other])
    X = rng.normal(size=(n, 2))
    y = X @ np.array([0.3, -0.2]) + rng.normal(size=n)
    two = oracle.estimate(y, X, fe, cluster=(1, 2), tol=1e-13, max_iter=20000)
    one = oracle.estimate(y, X, fe, cluster=(2,), tol=1e-13, max_iter=20000)
    # fine cap coarse = fine  =>  V_fine + V_coarse - V_fine = V_coarse; G_min = 8 in both
    assert rel(two.se, one.se) < 1e-10
    assert two.res["n_clusters"][:3] == [80, 8, 80]
    dup = oracle.estimate(y, X, fe, cluster=(2, 2, 2), tol=1e-13, max_iter=20000)
    assert dup.res["n_cluster_dims"] == 1 and rel(dup.se, one.se) < 1e-14


def test_cpu_fallback_bridge_to_gpu_conventions():
    """SE_cpu / SE_gpu = sqrt((N-K-dof_fes-1)/(N-K)) homoskedastic; sqrt((N-K-nested_adj)/(N-K))
    clustered (dense ids, no PSD clipping).  Relates fe_pipeline.f90:185-189,467-479 to
    :1171-1176,1351-1358."""
    p = synth.random_fe_problem(51, n=500, K=3, G=(30, 6, 4))
    N, K = p.n, 3
    dof_fes = int(np.sum(p.G - 1))
    for cluster, factor in [((), (N - K - dof_fes - 1) / (N - K)), ((1,), (N - K - 1) / (N - K)),
                            ((3,), (N - K - 1) / (N - K))]:
        g = oracle.estimate(p.y, p.W, p.fe, cluster=cluster, tol=1e-13, max_iter=20000, conv="gpu", ap="forward")
        c = oracle.estimate(p.y, p.W, p.fe, cluster=cluster, tol=1e-13, max_iter=20000, conv="cpu")
        assert rel(c.beta, g.beta) < 1e-10
        if cluster:
            assert g.res["psd_fix_applied"] == 0
        assert rel(c.se / g.se, np.full(K, np.sqrt(factor))) < 1e-10


def test_nlsw_like_panel_runs_homoskedastic_and_3way():
    p = synth.nlsw_like()
    assert 24000 < p.n < 28000
    for cluster in [(), (1, 2, 3)]:
        fit = oracle.estimate(p.y, p.W, p.fe, cluster=cluster, tol=1e-10, max_iter=5000)
        assert fit.res["converged"] == 1
        assert np.all(np.abs(fit.beta - p.beta_true) < 8 * fit.se)
    assert fit.res["dof_resid"] == 12          # G_min = 13 occupations


def test_stats_closed_forms():
    p = synth.random_fe_problem(61, n=300, K=2, G=(10, 5))
    fit = oracle.estimate(p.y, p.W, p.fe, tol=1e-13, max_iter=20000)
    ref = bf.fit(p.y, p.W, p.fe)
    tss = float(np.sum((p.y - p.y.mean()) ** 2))
    assert abs(fit.res["tss_total"] - tss) < 1e-10 * tss
    assert abs(fit.res["r2"] - (1 - ref["rss"] / tss)) < 1e-10
    tw = float(ref["yt"] @ ref["yt"])
    assert abs(fit.res["r2_within"] - (1 - ref["rss"] / tw)) < 1e-10


def test_f_stat_is_the_wald_statistic_over_k():
    """With one FE dim of a single group (an intercept), the reference's F
    (src/core/fe_pipeline.f90:2000-2005) is the textbook overall-regression F, which equals the Wald
    statistic beta' V^-1 beta / K of the homoskedastic covariance V; for K = 1 it is t^2."""
    for K, seed in [(1, 3), (3, 4)]:
        p = synth.random_fe_problem(seed, n=400, K=K, G=(5,))
        fe = np.ones((1, p.n), dtype=np.int32)
        fit = oracle.estimate(p.y, p.W, fe, tol=1e-13, max_iter=20000)
        wald = float(fit.beta @ np.linalg.solve(fit.vcov, fit.beta)) / K
        assert fit.res["dof_fes"] == 0 and fit.res["dof_resid"] == p.n - K - 1
        assert abs(fit.res["f_stat"] - wald) < 1e-9 * wald
        if K == 1:
            assert abs(fit.res["f_stat"] - (fit.beta[0] / fit.se[0]) ** 2) < 1e-9 * wald


def test_adjusted_r2_from_the_dense_design_rank():
    """Adjusted R^2 (src/core/fe_pipeline.f90:1990-1997) is 1 - (rss / (n - p)) / (tss / (n - 1)) with
    p the rank of the dense design [1 | dummies | W] of a connected panel: p counted by LAPACK here,
    not by the DoF rules."""
    p = synth.random_fe_problem(62, n=500, K=2, G=(30, 8))
    fit = oracle.estimate(p.y, p.W, p.fe, tol=1e-13, max_iter=20000)
    X = np.hstack([np.ones((p.n, 1)), bf.dummies(p.fe), p.W])
    rank = np.linalg.matrix_rank(X)
    coef, *_ = np.linalg.lstsq(X, p.y, rcond=None)
    rss = float(np.sum((p.y - X @ coef) ** 2))
    tss = float(np.sum((p.y - p.y.mean()) ** 2))
    assert abs(fit.res["rss"] - rss) < 1e-9 * rss
    assert abs(fit.res["r2_adj"] - (1 - (rss / (p.n - rank)) / (tss / (p.n - 1)))) < 1e-10
