"""The `.bin` dataset format (src/io/fe_data_io.f90): the numpy restatement in oracle/bin_format.py
pinned by hand-counted sizes and round trips, and the library's header parser (fehip_bin_read_header,
host code, no GPU) equal to it on every layout the reference's size rule separates."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import bin_format as bf


def test_size_formula_hand_counted():
    # N=10, K=2, L=1, two FE dims, cluster and weights, FP64: y 80 + W 160 + Z 80 + w 80 + ids 80 + cl 40
    assert bf.data_bytes(10, 2, 1, 2, True, True, 8) == 80 + 160 + 80 + 80 + 80 + 40
    # FP32 halves the value blocks only
    assert bf.data_bytes(10, 2, 1, 2, True, True, 4) == 40 + 80 + 40 + 40 + 80 + 40
    assert bf.data_bytes(7, 0, 0, 1, False, False, 8) == 56 + 28


def _data(rng, n=37, K=3, L=2, m=3):
    y = rng.normal(size=n)
    W = rng.normal(size=(n, K))
    Z = rng.normal(size=(n, L)) if L else None
    fe = np.vstack([rng.integers(1, 6 + 3 * d, n) for d in range(m)]).astype(np.int32)
    return y, W, Z, fe


VARIANTS = [dict(precision=p, flag=fl, old=o, cl=c, w=w, meta=mt)
            for p in ("float64", "float32") for fl in (True, False) for o in (False, True)
            for (c, w, mt) in ((False, False, None), (True, True, None), (True, False, {"y": "wage", "x": ["a", "b", "c"],
                                                                                       "fe": ["id", "firm", "t"]}))
            # the reference writes META only after the new header with its flag word; an old-layout file with
            # trailing bytes matches no size and the reader takes it for the new layout
            if not (mt and (not fl or o))]


@pytest.mark.parametrize("v", VARIANTS, ids=lambda v: "-".join(f"{k}={v[k]}" for k in ("precision", "flag", "old", "cl",
                                                                                        "w")) + ("-meta" if v["meta"] else ""))
def test_oracle_round_trip_and_library_header(v, tmp_path):
    rng = np.random.default_rng(11)
    y, W, Z, fe = _data(rng, L=0 if v["old"] else 2)
    n = len(y)
    cl = rng.integers(1, 5, n).astype(np.int32) if v["cl"] else None
    w = rng.random(n) if v["w"] else None
    path = os.path.join(tmp_path, "d.bin")
    bf.write(path, y, W, fe, Z=Z, cluster=cl, weights=w, precision=v["precision"], flag=v["flag"], old=v["old"],
             meta=v["meta"])
    h, blocks = bf.read(path)
    f32 = v["precision"] == "float32"
    cast = (lambda a: a.astype(np.float32).astype(np.float64)) if f32 else (lambda a: a)
    assert h["precision"] == int(f32) and h["new_format"] == int(not v["old"])
    assert h["has_precision_flag"] == int(v["flag"])
    assert np.array_equal(blocks["y"], cast(y)) and np.array_equal(blocks["W"], cast(W))
    if Z is not None:
        assert np.array_equal(blocks["Z"], cast(Z))
    assert np.array_equal(blocks["fe"], fe)
    if cl is not None:
        assert np.array_equal(blocks["cluster"], cl)
    if w is not None:
        assert np.array_equal(blocks["weights"], cast(w))
    if v["meta"]:
        assert h["metadata_bytes"] > 0
    # the library's parser (host code) decides the same layout
    from fe_gpu_amd import _coarse_sigs as S
    from fe_gpu_amd._lib import load
    lib = load()
    hh = S.BinHeader()
    assert lib.fehip_bin_read_header(path.encode(), C.byref(hh)) == 0
    got = {name: getattr(hh, name) for name, _ in S.BinHeader._fields_}
    assert got == h


def test_library_header_errors(tmp_path):
    from fe_gpu_amd import _coarse_sigs as S
    from fe_gpu_amd._lib import load
    lib = load()
    hh = S.BinHeader()
    assert lib.fehip_bin_read_header(os.path.join(str(tmp_path), "missing.bin").encode(), C.byref(hh)) != 0
    short = os.path.join(str(tmp_path), "short.bin")
    open(short, "wb").write(b"\0" * 10)
    assert lib.fehip_bin_read_header(short.encode(), C.byref(hh)) != 0
    # a flag word that is neither 0 nor 1, with trailing bytes so the size rule reads it
    rng = np.random.default_rng(3)
    y, W, Z, fe = _data(rng, L=0)
    bad = os.path.join(str(tmp_path), "bad.bin")
    bf.write(bad, y, W, fe, meta={"y": "v"})
    raw = bytearray(open(bad, "rb").read())
    raw[bf.HEADER_NEW:bf.HEADER_NEW + 4] = (7).to_bytes(4, "little")
    open(bad, "wb").write(bytes(raw))
    assert lib.fehip_bin_read_header(bad.encode(), C.byref(hh)) != 0
    with pytest.raises(ValueError):
        bf.read_header(bad)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["loader_case_fp64_noflag.bin", "loader_case_fp32_flag.bin"])
def test_reference_loader_fixtures(name):
    """The two files of the reference's loader test (tests/test_data_loader.f90:17-132, written by
    tests/golden/make_loader_fixtures.py): the numpy reader and the library's header parser decide the
    layout and precision the reference test asserts, and the blocks equal its expected values to its
    tolerances (1e-12 FP64, 1e-6 FP32)."""
    import json
    exp = json.load(open(os.path.join(GOLDEN, "loader_cases.json")))[name]
    path = os.path.join(GOLDEN, name)
    h, b = bf.read(path)
    for k, v in exp["header"].items():
        assert h[k] == v, k
    tol = exp["tol"]
    assert np.max(np.abs(b["y"] - exp["y"])) <= tol
    assert np.max(np.abs(b["W"] - np.array(exp["W"]).T)) <= tol
    assert np.array_equal(b["fe"], np.array(exp["fe"], dtype=np.int32))
    if exp["cluster"] is None:
        assert b["cluster"] is None
    else:
        assert np.array_equal(b["cluster"], exp["cluster"])
    if exp["weights"] is None:
        assert b["weights"] is None
    else:
        assert np.max(np.abs(b["weights"] - exp["weights"])) <= tol
    from fe_gpu_amd import _coarse_sigs as S
    from fe_gpu_amd._lib import load
    hh = S.BinHeader()
    assert load().fehip_bin_read_header(path.encode(), C.byref(hh)) == 0
    assert {name_: getattr(hh, name_) for name_, _ in S.BinHeader._fields_} == h
