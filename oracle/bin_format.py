"""The reference's `.bin` dataset format in numpy -- TEST INFRASTRUCTURE ONLY.

A restatement of load_dataset_from_file (src/io/fe_data_io.f90:23-66) for checking the device
ingest (fehip_bin_read_header / fehip_bin_load): only tests/ may import it.  ``write`` follows the
reference's writer (tools/dataframe_to_fe_binary.py:137-180; the old 24-byte header layout without L
and files without the precision flag are written here too, since the reader accepts them).

Parity pins: the size formula against hand-counted byte totals, every header variant the reader's
size rule separates (new/old layout x flag/no flag x FP64/FP32), a META block of known length, and
round trips of the blocks (the reader returns exactly what the writer wrote; FP32 values widened).
"""
from __future__ import annotations

import struct

import numpy as np

HEADER_NEW = 8 + 5 * 4  # fe_data_io.f90:7
HEADER_OLD = 8 + 4 * 4  # fe_data_io.f90:8


def data_bytes(n, K, L, n_fe, has_cluster, has_weights, vb):
    """compute_data_bytes (fe_data_io.f90:342-358)."""
    t = n * (1 + K) * vb
    if L > 0:
        t += n * L * vb
    if has_weights:
        t += n * vb
    t += n_fe * n * 4
    if has_cluster:
        t += n * 4
    return t


def _meta_bytes(meta):
    if not meta:
        return b""
    out = [b"META", struct.pack("<i", 1)]

    def s(x):
        d = x.encode("utf-8")
        out.append(struct.pack("<i", len(d)) + d)

    def lst(xs):
        out.append(struct.pack("<i", len(xs)))
        for x in xs:
            s(x)

    s(meta.get("y", ""))
    lst(meta.get("x", []))
    lst(meta.get("iv", []))
    lst(meta.get("fe", []))
    s(meta.get("cluster", ""))
    s(meta.get("weights", ""))
    return b"".join(out)


def write(path, y, W, fe, Z=None, cluster=None, weights=None, precision="float64", flag=True, old=False,
          meta=None):
    """A `.bin` file: header, optional precision flag, y, W, Z, FE ids (dim-major), cluster, weights, META."""
    n = len(y)
    W = np.zeros((n, 0)) if W is None else np.asarray(W)
    K, L = W.shape[1], (0 if Z is None else Z.shape[1])
    fe = np.asarray(fe, dtype="<i4")
    if old and L:
        raise ValueError("the old layout has no instruments")
    ft = np.dtype("<f4") if precision == "float32" else np.dtype("<f8")
    with open(path, "wb") as f:
        if old:
            f.write(struct.pack("<qiiii", n, K, fe.shape[0], int(cluster is not None), int(weights is not None)))
        else:
            f.write(struct.pack("<qiiiii", n, K, L, fe.shape[0], int(cluster is not None), int(weights is not None)))
        if flag:
            f.write(struct.pack("<i", 1 if precision == "float32" else 0))
        np.asarray(y, dtype=ft).tofile(f)
        np.asarray(W, dtype=ft).T.tofile(f)
        if L:
            np.asarray(Z, dtype=ft).T.tofile(f)
        for d in range(fe.shape[0]):
            fe[d].tofile(f)
        if cluster is not None:
            np.asarray(cluster, dtype="<i4").tofile(f)
        if weights is not None:
            np.asarray(weights, dtype=ft).tofile(f)
        f.write(_meta_bytes(meta))


def read_header(path):
    """read_header (fe_data_io.f90:78-202): layout and precision from the file size."""
    import os
    size = os.path.getsize(path)
    if size < HEADER_OLD:
        raise ValueError("File too small to contain required header")
    with open(path, "rb") as f:
        hb = f.read(HEADER_NEW + 4).ljust(HEADER_NEW + 4, b"\0")
    n, K = struct.unpack_from("<qi", hb, 0)
    v = struct.unpack_from("<5i", hb, 12)
    if n <= 0 or K < 0:
        raise ValueError("bad N or K")

    def matches(base, L, nfe, cl, w):
        t64 = base + data_bytes(n, K, L, nfe, cl, w, 8)
        t32 = base + data_bytes(n, K, L, nfe, cl, w, 4)
        return size in (t64, t32, t64 + 4, t32 + 4)

    mnew = matches(HEADER_NEW, v[0], v[1], v[2] != 0, v[3] != 0)
    mold = matches(HEADER_OLD, 0, v[0], v[1] != 0, v[2] != 0)
    new = mnew or not mold  # the reader falls back to the new layout (fe_data_io.f90:122-124)
    if new:
        h = dict(n_obs=n, n_regressors=K, n_instruments=v[0], n_fe=v[1], has_cluster=int(v[2] != 0),
                 has_weights=int(v[3] != 0), new_format=1)
        base = HEADER_NEW
    else:
        h = dict(n_obs=n, n_regressors=K, n_instruments=0, n_fe=v[0], has_cluster=int(v[1] != 0),
                 has_weights=int(v[2] != 0), new_format=0)
        base = HEADER_OLD
    args = (n, K, h["n_instruments"], h["n_fe"], h["has_cluster"], h["has_weights"])
    d64, d32 = data_bytes(*args, 8), data_bytes(*args, 4)
    h.update(has_precision_flag=0, metadata_bytes=0)
    if size == base + d64:
        h.update(precision=0, data_offset=base)
        return h
    if size == base + d32:
        h.update(precision=1, data_offset=base)
        return h
    flag = struct.unpack_from("<i", hb, base)[0]
    if flag not in (0, 1):
        raise ValueError(f"Unsupported precision flag value: {flag}")
    expected = base + 4 + (d64 if flag == 0 else d32)
    if size < expected:
        raise ValueError("File size does not match header metadata for declared precision")
    h.update(precision=flag, has_precision_flag=1, data_offset=base + 4, metadata_bytes=size - expected)
    return h


def read(path):
    """The blocks as the reference loads them (FP32 values widened to FP64, fe_data_io.f90:243-299)."""
    h = read_header(path)
    n, K, L, m = h["n_obs"], h["n_regressors"], h["n_instruments"], h["n_fe"]
    ft = np.dtype("<f4") if h["precision"] == 1 else np.dtype("<f8")
    with open(path, "rb") as f:
        f.seek(h["data_offset"])
        y = np.fromfile(f, ft, n).astype(np.float64)
        W = np.fromfile(f, ft, n * K).astype(np.float64).reshape(K, n).T
        Z = np.fromfile(f, ft, n * L).astype(np.float64).reshape(L, n).T
        fe = np.fromfile(f, "<i4", n * m).reshape(m, n)
        cl = np.fromfile(f, "<i4", n) if h["has_cluster"] else None
        w = np.fromfile(f, ft, n).astype(np.float64) if h["has_weights"] else None
    return h, dict(y=y, W=W, Z=Z, fe=fe, cluster=cl, weights=w)
