"""GPU parity of the `.bin` ingest (fehip_bin_load): the device blocks equal the reference reader's
restatement (oracle/bin_format.py) bit for bit, FP32 files widened on the device; a file larger than
the 64 MiB staging chunk exercises the double-buffered upload; a regression on the loaded panel
equals the one on the same host arrays."""
import os
import sys

import numpy as np
import pytest

from oracle import bin_format as bf

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    from fe_gpu_amd import device as d
    d.init(0)


def _check_blocks(path):
    from fe_gpu_amd import api
    h, ref = bf.read(path)
    d = api.load_bin(path)
    n, K, L, m = h["n_obs"], h["n_regressors"], h["n_instruments"], h["n_fe"]
    assert all(d.header[k] == h[k] for k in h)
    assert np.array_equal(d.y.to_array(np.float64, n), ref["y"])
    if K:
        assert np.array_equal(d.W.to_array(np.float64, (n, K), order="F"), ref["W"])
    if L:
        assert np.array_equal(d.Z.to_array(np.float64, (n, L), order="F"), ref["Z"])
    assert np.array_equal(d.fe.to_array(np.int32, (m, n)), ref["fe"])
    if h["has_cluster"]:
        assert np.array_equal(d.cluster.to_array(np.int32, n), ref["cluster"])
    if h["has_weights"]:
        assert np.array_equal(d.weights.to_array(np.float64, n), ref["weights"])
    return d


@pytest.mark.parametrize("precision,flag,old", [("float64", True, False), ("float32", True, False),
                                                ("float64", False, True), ("float32", False, True)])
def test_blocks_equal_reference_reader(precision, flag, old, tmp_path):
    note(f"bin ingest {precision} flag={flag} old={old}")
    rng = np.random.default_rng(21)
    n, K = 5003, 4
    y = rng.normal(size=n)
    W = rng.normal(size=(n, K))
    Z = None if old else rng.normal(size=(n, 2))
    fe = np.vstack([rng.integers(1, 400, n), rng.integers(1, 30, n)]).astype(np.int32)
    path = os.path.join(tmp_path, "d.bin")
    bf.write(path, y, W, fe, Z=Z, cluster=rng.integers(1, 9, n).astype(np.int32), weights=rng.random(n),
             precision=precision, flag=flag, old=old, meta={"y": "y"} if flag and not old else None)
    _check_blocks(path).free()


def test_multi_chunk_file_and_regression_on_loaded_panel(tmp_path):
    from fe_gpu_amd import api, synth
    note("bin ingest across staging chunks + regression")
    p = synth.random_fe_problem(77, n=3_000_000, K=3, G=(60000, 900, 25))
    path = os.path.join(tmp_path, "big.bin")
    bf.write(path, p.y, p.W, p.fe)
    assert os.path.getsize(path) > 64 << 20
    d = _check_blocks(path)
    pan = d.panel()
    fit = api.estimate_device(pan, d.y.ptr, d.W.ptr, 3, cluster=(1, 2), tol=1e-10, max_iter=5000)
    ref = api.estimate(p.y, p.W, p.fe, cluster=(1, 2), tol=1e-10, max_iter=5000)
    assert np.max(np.abs(fit.beta - ref.beta) / np.abs(ref.beta)) < 1e-12
    assert np.max(np.abs(fit.se - ref.se) / np.abs(ref.se)) < 1e-12
    pan.close()
    d.free()


@pytest.mark.parametrize("name", ["loader_case_fp64_noflag.bin", "loader_case_fp32_flag.bin"])
def test_reference_loader_fixtures_on_device(name):
    """The reference loader test's two files (tests/test_data_loader.f90:17-132): the device ingest equals
    the reader bit for bit and the values that test expects (FP32 widened on the device)."""
    import json
    note(f"bin ingest fixture {name}")
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    exp = json.load(open(os.path.join(golden, "loader_cases.json")))[name]
    d = _check_blocks(os.path.join(golden, name))
    n = exp["header"]["n_obs"]
    for k, v in exp["header"].items():
        assert d.header[k] == v, k
    assert np.max(np.abs(d.y.to_array(np.float64, n) - exp["y"])) <= exp["tol"]
    assert np.array_equal(d.fe.to_array(np.int32, (len(exp["fe"]), n)), np.array(exp["fe"], dtype=np.int32))
    d.free()
