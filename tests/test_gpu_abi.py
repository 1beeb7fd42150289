"""GPU parity of the drop-in C-ABI (fine-grained entry points) against numpy and the oracle.

Each call goes through libfe_gpu_hip.so exactly as the Fortran host would call it.  Integer
outputs (counts, intersection ids) must be bit-exact; FP64 kernels within 1e-12 relative of a
numpy FP64 evaluation; the full estimation mirror within 1e-8 relative of the oracle (beta, SE).
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    from fe_gpu_amd import device
    from fe_gpu_amd._lib import load
    info = device.init(0)
    note(f"device {info.name.decode()} CUs={info.multiprocessor_count} mem={info.total_global_mem >> 30} GiB")
    return load()


def dev(a):
    from fe_gpu_amd.device import DeviceBuffer
    return DeviceBuffer.from_array(a)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = max(np.max(np.abs(b)), 1e-300) if b.size else 1.0
    return float(np.max(np.abs(a - b)) / scale) if a.size else 0.0


@pytest.mark.parametrize("G,sorted_ids", [(13, False), (5000, False), (5000, True)])
def test_accumulate_means_subtract(lib, G, sorted_ids):
    note(f"accumulate G={G} sorted={sorted_ids}")
    rng = np.random.default_rng(G)
    n, K, L = 20000, 3, 2
    ids = rng.integers(1, G + 1, n).astype(np.int32)
    if sorted_ids:
        ids.sort()
    y = rng.normal(size=n)
    W = np.asfortranarray(rng.normal(size=(n, K)))
    Z = np.asfortranarray(rng.normal(size=(n, L)))
    dy, dW, dZ, di = dev(y), dev(W), dev(Z), dev(ids)
    from fe_gpu_amd.device import DeviceBuffer
    sy, sW, sZ, cnt = DeviceBuffer(G * 8), DeviceBuffer(G * K * 8), DeviceBuffer(G * L * 8), DeviceBuffer(G * 4)
    for b in (sy, sW, sZ, cnt):
        b.zero()
    assert lib.fe_gpu_fe_accumulate(dy.ptr, dW.ptr, dZ.ptr, di.ptr, n, G, K, L, n, sy.ptr, sW.ptr, sZ.ptr,
                                    cnt.ptr) == 0
    ey = np.zeros(G); np.add.at(ey, ids - 1, y)
    eW = np.zeros((G, K)); np.add.at(eW, ids - 1, W)
    eZ = np.zeros((G, L)); np.add.at(eZ, ids - 1, Z)
    ec = np.bincount(ids - 1, minlength=G)
    assert np.array_equal(cnt.to_array(np.int32, (G,)), ec)
    assert rel(sy.to_array(np.float64, (G,)), ey) < 1e-12
    assert rel(sW.to_array(np.float64, (G, K)), eW) < 1e-12      # group-major [g*K + k]
    assert rel(sZ.to_array(np.float64, (G, L)), eZ) < 1e-12
    assert lib.fe_gpu_fe_compute_means(sy.ptr, sW.ptr, sZ.ptr, cnt.ptr, G, K, L) == 0
    inv = np.where(ec > 0, 1.0 / np.maximum(ec, 1), 0.0)
    my = sy.to_array(np.float64, (G,))
    assert rel(my, ey * inv) < 1e-12
    assert lib.fe_gpu_fe_subtract(dy.ptr, dW.ptr, dZ.ptr, di.ptr, n, K, L, n, sy.ptr, sW.ptr, sZ.ptr, 1.7) == 0
    assert rel(dy.to_array(np.float64, (n,)), y - (ey * inv)[ids - 1]) < 1e-12
    Wt = dW.to_array(np.float64, (n, K), order="F")
    assert rel(Wt, W - (eW * inv[:, None])[ids - 1]) < 1e-12


def test_mix_absmax_axpy_copy(lib):
    note("mix/absmax/axpy/copy_columns")
    rng = np.random.default_rng(3)
    m = rng.normal(size=1000)
    p = rng.normal(size=1000)
    dm, dp_ = dev(m), dev(p)
    assert lib.fe_gpu_mix_means(dm.ptr, dp_.ptr, 1000, 1.3) == 0
    mixed = p + 1.3 * (m - p)
    assert np.allclose(dm.to_array(np.float64, (1000,)), mixed, rtol=0, atol=1e-15)
    assert np.allclose(dp_.to_array(np.float64, (1000,)), mixed, rtol=0, atol=1e-15)
    out = C.c_double()
    assert lib.fe_gpu_absmax(dm.ptr, 1000, C.byref(out)) == 0 and out.value == np.max(np.abs(mixed))
    assert lib.fe_gpu_absmax(dm.ptr, 0, C.byref(out)) != 0 and out.value == 0.0
    x = rng.normal(size=777)
    yv = rng.normal(size=777)
    dx, dyv = dev(x), dev(yv)
    assert lib.fe_gpu_axpy(777, -0.5, dx.ptr, dyv.ptr) == 0
    assert np.allclose(dyv.to_array(np.float64, (777,)), yv - 0.5 * x, rtol=0, atol=1e-15)
    src = np.asfortranarray(rng.normal(size=(500, 6)))
    ds_ = dev(src)
    from fe_gpu_amd.device import DeviceBuffer
    dst = DeviceBuffer(500 * 5 * 8)
    dst.zero()
    idx = np.array([6, 2, 4], dtype=np.int32)
    assert lib.fe_gpu_copy_columns(ds_.ptr, 500, idx.ctypes.data, 3, 500, dst.ptr, 500, 1) == 0
    got = dst.to_array(np.float64, (500, 5), order="F")
    assert np.array_equal(got[:, 1:4], src[:, [5, 1, 3]]) and np.all(got[:, 0] == 0) and np.all(got[:, 4] == 0)
    bad = np.array([0], dtype=np.int32)
    assert lib.fe_gpu_copy_columns(ds_.ptr, 500, bad.ctypes.data, 1, 500, dst.ptr, 500, 0) != 0


@pytest.mark.parametrize("tA,tB", [("T", "N"), ("N", "N"), ("N", "T"), ("T", "T")])
def test_gemm_all_transposes(lib, tA, tB):
    note(f"gemm {tA}{tB}")
    rng = np.random.default_rng(5)
    m, n, k = 37, 29, 50001 if tA == "T" and tB == "N" else 301
    A = rng.normal(size=(k, m) if tA == "T" else (m, k))
    B = rng.normal(size=(n, k) if tB == "T" else (k, n))
    Cm = rng.normal(size=(m + 3, n))
    dA, dB, dC = dev(np.asfortranarray(A)), dev(np.asfortranarray(B)), dev(np.asfortranarray(Cm))
    assert lib.fe_gpu_gemm(tA.encode(), tB.encode(), m, n, k, 0.7, dA.ptr, A.shape[0], dB.ptr, B.shape[0], -0.3,
                           dC.ptr, m + 3) == 0
    opA = A.T if tA == "T" else A
    opB = B.T if tB == "T" else B
    exp = Cm.copy()
    exp[:m] = 0.7 * opA @ opB - 0.3 * Cm[:m]
    assert rel(dC.to_array(np.float64, (m + 3, n), order="F"), exp) < 1e-12


@pytest.mark.parametrize("m,n,k", [(265, 252, 20011), (64, 64, 4096), (100, 17, 9001)])
@pytest.mark.parametrize("shape", ["TN", "NT"])
def test_gemm_wide_cross_products_on_matrix_cores(lib, m, n, k, shape):
    """A^T B and A B^T with a long reduction run on the FP64 matrix cores (tile padding, split-K)."""
    note(f"gemm mfma {shape} {m}x{n}x{k}")
    rng = np.random.default_rng(m + n)
    if shape == "TN":
        A, B = rng.normal(size=(k, m)), rng.normal(size=(k, n))
        exp = A.T @ B
    else:
        A, B = rng.normal(size=(m, k)), rng.normal(size=(n, k))
        exp = A @ B.T
    dA, dB = dev(np.asfortranarray(A)), dev(np.asfortranarray(B))
    dC = dev(np.zeros((m, n), order="F"))
    assert lib.fe_gpu_gemm(shape[0].encode(), shape[1].encode(), m, n, k, 1.0, dA.ptr, A.shape[0], dB.ptr,
                           B.shape[0], 0.0, dC.ptr, m) == 0
    assert rel(dC.to_array(np.float64, (m, n), order="F"), exp) < 1e-12


@pytest.mark.parametrize("case", ["TN_odd_rows", "TN_same", "NT_even", "NT_same", "NT_odd"])
def test_gemm_pair_loads_and_symmetric_tiles(lib, case):
    """The matrix-core GEMM's pair loads (even leading dims): an odd row count on the A^T B shape leaves its
    last row to the finish pass; A = B shares one operand on the diagonal tiles; the A B^T shape takes
    pairs along even m and n, and single loads for odd ones."""
    note(f"gemm pair loads {case}")
    rng = np.random.default_rng(len(case))
    k = 20011
    if case.startswith("TN"):
        m, n = (130, 130) if case == "TN_same" else (130, 120)
        ld = k + 1                                   # even leading dim, odd row count
        Ab = np.asfortranarray(rng.normal(size=(ld, m)))
        Bb = Ab if case == "TN_same" else np.asfortranarray(rng.normal(size=(ld, n)))
        exp = Ab[:k].T @ Bb[:k]
        dA = dev(Ab)
        dB = dA if case == "TN_same" else dev(Bb)
        args = (b"T", b"N", m, n, k, 1.0, dA.ptr, ld, dB.ptr, ld)
    else:
        m, n = {"NT_even": (264, 120), "NT_same": (264, 264), "NT_odd": (265, 121)}[case]
        ld = 272                                     # row-major view: element (p, i) at p * ld + i
        Ab = np.asfortranarray(rng.normal(size=(ld, k)))
        Bb = Ab if case == "NT_same" else np.asfortranarray(rng.normal(size=(ld, k)))
        exp = Ab[:m] @ Bb[:n].T
        dA = dev(Ab)
        dB = dA if case == "NT_same" else dev(Bb)
        args = (b"N", b"T", m, n, k, 1.0, dA.ptr, ld, dB.ptr, ld)
    dC = dev(np.zeros((m, n), order="F"))
    assert lib.fe_gpu_gemm(*args, 0.0, dC.ptr, m) == 0
    assert rel(dC.to_array(np.float64, (m, n), order="F"), exp) < 1e-12


def test_syrk_pair_loads_odd_rows_upper(lib):
    """SYRK on the matrix cores with an even leading dim and an odd row count: pair loads, the last row
    from the finish pass, one operand and no below-diagonal blocks on the diagonal tiles."""
    note("syrk pair loads odd rows")
    rng = np.random.default_rng(12)
    n, K, ld = 30011, 140, 30012
    Wb = np.asfortranarray(rng.normal(size=(ld, K)))
    dW = dev(Wb)
    dQ = dev(np.full((K, K), 7.0, order="F"))
    assert lib.fe_gpu_syrk(n, K, 1.0, dW.ptr, ld, 0.0, dQ.ptr) == 0
    Q = dQ.to_array(np.float64, (K, K), order="F")
    iu = np.triu_indices(K)
    assert rel(Q[iu], (Wb[:n].T @ Wb[:n])[iu]) < 1e-12
    assert np.all(Q[np.tril_indices(K, -1)] == 7.0)


def test_syrk_wide_upper_on_matrix_cores(lib):
    note("syrk mfma 140 cols")
    rng = np.random.default_rng(11)
    n, K = 30011, 140
    W = np.asfortranarray(rng.normal(size=(n, K)))
    dW = dev(W)
    dQ = dev(np.full((K, K), 7.0, order="F"))
    assert lib.fe_gpu_syrk(n, K, 1.0, dW.ptr, n, 0.0, dQ.ptr) == 0
    Q = dQ.to_array(np.float64, (K, K), order="F")
    iu = np.triu_indices(K)
    assert rel(Q[iu], (W.T @ W)[iu]) < 1e-12
    assert np.all(Q[np.tril_indices(K, -1)] == 7.0)


def test_syrk_gemv_residual_dot(lib):
    note("syrk/gemv/residual/dot")
    rng = np.random.default_rng(6)
    n, K = 123457, 8
    W = np.asfortranarray(rng.normal(size=(n, K)))
    y = rng.normal(size=n)
    dW, dy = dev(W), dev(y)
    from fe_gpu_amd.device import DeviceBuffer
    dQ = dev(np.full((K, K), 7.0, order="F"))
    assert lib.fe_gpu_syrk(n, K, 1.0, dW.ptr, n, 0.0, dQ.ptr) == 0
    Q = dQ.to_array(np.float64, (K, K), order="F")
    exp = W.T @ W
    iu = np.triu_indices(K)
    assert rel(Q[iu], exp[iu]) < 1e-12
    assert np.all(Q[np.tril_indices(K, -1)] == 7.0)          # lower triangle untouched (UPLO = U)
    db = DeviceBuffer(K * 8)
    assert lib.fe_gpu_gemv(n, K, 1.0, dW.ptr, n, dy.ptr, 0.0, db.ptr) == 0
    assert rel(db.to_array(np.float64, (K,)), W.T @ y) < 1e-12
    beta = rng.normal(size=K)
    dbeta, dr = dev(beta), DeviceBuffer(n * 8)
    assert lib.fe_gpu_residual(n, K, dW.ptr, n, dbeta.ptr, dy.ptr, dr.ptr) == 0
    r = dr.to_array(np.float64, (n,))
    assert rel(r, y - W @ beta) < 1e-12
    out = C.c_double()
    assert lib.fe_gpu_dot(n, dr.ptr, dr.ptr, C.byref(out)) == 0
    assert abs(out.value - r @ r) < 1e-11 * (r @ r)


@pytest.mark.parametrize("sorted_ids", [False, True])
def test_cluster_scores_and_meat(lib, sorted_ids):
    note(f"cluster scores sorted={sorted_ids}")
    rng = np.random.default_rng(7)
    n, K, G = 100003, 5, 3001
    ids = rng.integers(1, G + 1, n).astype(np.int32)
    if sorted_ids:
        ids.sort()
    ids[5] = 0                                          # skipped (cluster < 0)
    r = rng.normal(size=n)
    W = np.asfortranarray(rng.normal(size=(n, K)))
    dr, dW, di = dev(r), dev(W), dev(ids)
    from fe_gpu_amd.device import DeviceBuffer
    S = DeviceBuffer(G * K * 8)
    S.zero()
    assert lib.fe_gpu_cluster_scores(dr.ptr, dW.ptr, di.ptr, n, K, n, G, S.ptr) == 0
    m = ids >= 1
    exp = np.zeros((G, K))
    np.add.at(exp, ids[m] - 1, r[m, None] * W[m])
    got = S.to_array(np.float64, (G, K), order="F")              # col-major G x K, ld G
    assert rel(got, exp) < 1e-12
    M = DeviceBuffer(K * K * 8)
    assert lib.fe_gpu_cluster_meat(G, K, S.ptr, G, M.ptr) == 0
    Mh = M.to_array(np.float64, (K, K), order="F")
    iu = np.triu_indices(K)
    assert rel(Mh[iu], (exp.T @ exp)[iu]) < 1e-12


@pytest.mark.parametrize("dims", [[1, 2], [2, 3], [1, 2, 3], [3, 1]])
def test_multi_cluster_ids_bit_exact(lib, dims):
    """Numbering equals the CPU builder's (src/core/cluster_id_builder.c) -> the oracle's."""
    import oracle
    note(f"multi cluster ids dims={dims}")
    rng = np.random.default_rng(sum(dims))
    n = 300001
    G = [40000, 900, 30]
    fe = np.vstack([rng.integers(1, g + 1, n) for g in G]).astype(np.int32)
    for d, g in enumerate(G):
        fe[d, 0] = g
    bufs = [dev(fe[d]) for d in range(3)]
    sizes = [G[d - 1] for d in dims]
    strides = np.cumprod([1] + sizes[:-1]).astype(np.uint64)
    ptrs = (C.c_void_p * len(dims))(*[bufs[d - 1].ptr for d in dims])
    from fe_gpu_amd.device import DeviceBuffer
    out = DeviceBuffer(n * 4)
    nc = C.c_int()
    assert lib.fe_gpu_build_multi_cluster_ids(ptrs, strides.ctypes.data, len(dims), n, out.ptr, C.byref(nc)) == 0
    st, exp, nexp = oracle.cluster_ids(fe, dims, sizes)
    assert st == 0 and nc.value == nexp
    assert np.array_equal(out.to_array(np.int32, (n,)), exp)


def _fits_close(beta, se, fit, tb=1e-8, ts=1e-8):
    assert rel(beta, fit.beta) < tb, (beta, fit.beta)
    assert rel(se, fit.se) < ts, (se, fit.se)


def test_pipeline_known_answer(lib):
    """tests/test_gpu_regression.f90:17,40: beta = (1.5, -0.75) within 1e-8, through the ABI."""
    from fe_gpu_amd import pipeline
    note("pipeline known answer")
    W = np.array([[1.0, -2.0], [1.5, 0.5], [-0.5, 1.0], [-1.0, -1.5], [0.25, 0.75], [2.0, 1.25]])
    fe = np.array([[1, 1, 2, 2, 3, 3], [1, 2, 1, 2, 1, 2]], dtype=np.int32)
    y = W @ np.array([1.5, -0.75]) + np.array([0.5, -0.25, 0.8])[fe[0] - 1] + np.array([-0.4, 0.2])[fe[1] - 1]
    beta, se, vcov, res = pipeline.fe_gpu_estimate(y, W, fe, tol=1e-10, max_iter=500)
    assert res["converged"] and np.max(np.abs(beta - [1.5, -0.75])) < 1e-8


def test_pipeline_demean_invariant(lib):
    """tests/test_gpu_demean.f90:43-84 through the ABI AP loop."""
    from fe_gpu_amd import pipeline
    note("pipeline demean invariant")
    y = np.array([1.0, 2.0, 3.0, -2.0, -1.0, 0.5])
    W = np.array([[1.0, 0.0], [1.5, -2.0], [-0.5, 1.0], [-1.0, 0.5], [2.0, 1.5], [-0.5, -1.0]])
    fe = np.array([[1, 1, 2, 2, 3, 3], [1, 2, 1, 2, 1, 2]], dtype=np.int32)
    ds = pipeline.dataset_upload(y, W, None, fe, [3, 2])
    conv, it, mu = pipeline.within_transform(ds, 1e-8, 200)
    assert conv
    yt = ds.d_y.to_array(np.float64, (6,))
    Wt = ds.d_W.to_array(np.float64, (6, 2), order="F")
    for d in range(2):
        for g in (1, 2, 3)[: [3, 2][d]]:
            msk = fe[d] == g
            assert abs(yt[msk].mean()) < 1e-6 and np.all(np.abs(Wt[msk].mean(0)) < 1e-6)


@pytest.mark.parametrize("cluster", [(), (1,), (1, 2, 3)])
def test_pipeline_matches_oracle_nlsw_like(lib, cluster):
    import oracle
    from fe_gpu_amd import pipeline, synth
    note(f"pipeline vs oracle nlsw-like cluster={cluster}")
    p = synth.nlsw_like()
    beta, se, vcov, res = pipeline.fe_gpu_estimate(p.y, p.W, p.fe, cluster=cluster, tol=1e-12, max_iter=5000)
    fit = oracle.estimate(p.y, p.W, p.fe, cluster=cluster, tol=1e-12, max_iter=5000)
    _fits_close(beta, se, fit)
    assert res["dof_fes"] == fit.res["dof_fes"] and res["dof_resid"] == fit.res["dof_resid"]
    assert res["dof_model"] == fit.res["dof_model"]
    if cluster:
        assert res["n_clusters"] == fit.res["n_clusters"]


def test_pipeline_iv_matches_oracle(lib):
    import oracle
    from fe_gpu_amd import pipeline, synth
    note("pipeline IV vs oracle")
    p = synth.random_fe_problem(77, n=3000, K=3, L=3, G=(60, 12, 5))
    for cluster in [(), (1, 2)]:
        beta, se, vcov, res = pipeline.fe_gpu_estimate(p.y, p.W, p.fe, Z=p.Z, iv_reg=[1], cluster=cluster, tol=1e-12,
                                                       max_iter=20000)
        fit = oracle.estimate(p.y, p.W, p.fe, Z=p.Z, iv_reg=[1], cluster=cluster, tol=1e-12, max_iter=20000)
        assert res["is_iv"] == 1 and fit.res["is_iv"] == 1
        _fits_close(beta, se, fit)


@pytest.mark.parametrize("G", [13, 5000])
def test_accumulate_skips_ids_outside_the_groups(lib, G):
    """Ids below 1 or above n_groups are skipped by both accumulate paths (LDS tables for small G,
    merged runs otherwise), as the reference's small-G kernel does."""
    from fe_gpu_amd.device import DeviceBuffer
    note(f"accumulate out-of-range ids G={G}")
    rng = np.random.default_rng(G + 1)
    n, K = 30000, 2
    ids = rng.integers(-2, G + 4, n).astype(np.int32)
    y = rng.normal(size=n)
    W = np.asfortranarray(rng.normal(size=(n, K)))
    dy, dW, di = dev(y), dev(W), dev(ids)
    sy, sW, cnt = DeviceBuffer(G * 8), DeviceBuffer(G * K * 8), DeviceBuffer(G * 4)
    for b in (sy, sW, cnt):
        b.zero()
    assert lib.fe_gpu_fe_accumulate(dy.ptr, dW.ptr, None, di.ptr, n, G, K, 0, n, sy.ptr, sW.ptr, None, cnt.ptr) == 0
    ok = (ids >= 1) & (ids <= G)
    ey = np.zeros(G); np.add.at(ey, ids[ok] - 1, y[ok])
    eW = np.zeros((G, K)); np.add.at(eW, ids[ok] - 1, W[ok])
    ec = np.bincount(ids[ok] - 1, minlength=G)
    assert np.array_equal(cnt.to_array(np.int32, G), ec)
    assert rel(sy.to_array(np.float64, G), ey) < 1e-12
    assert rel(sW.to_array(np.float64, (G, K)), eW) < 1e-12
