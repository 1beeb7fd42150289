"""Writes the two `.bin` files of the reference's loader test (tests/test_data_loader.f90:17-132) byte for
byte from the layout that test writes, and the values its assertions expect, into this folder:

  loader_case_fp64_noflag.bin  24-byte header (N=3, K=2, n_fe=1, cluster, no weights), no precision
                               flag, FP64 y and W, one FE column, a cluster column (:81-104)
  loader_case_fp32_flag.bin    24-byte header (N=2, K=1, n_fe=2, no cluster, weights), the precision
                               flag PRECISION_FLOAT32 = 1 (src/core/fe_types.f90:7), FP32 y, W, two FE
                               columns and FP32 weights (:106-132)
  loader_cases.json            the expected header fields and blocks (:17-74), with the tolerances the
                               reference test checks them to (1e-12 for FP64 values, 1e-6 for FP32)

Run from anywhere: python tests/golden/make_loader_fixtures.py
"""
import json
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    # case 1: int64 N, int32 K, n_fe, has_cluster, has_weights; y, W (column-major), fe, cluster
    with open(os.path.join(HERE, "loader_case_fp64_noflag.bin"), "wb") as f:
        f.write(struct.pack("<qiiii", 3, 2, 1, 1, 0))
        f.write(np.array([1.0, 2.0, -1.0], "<f8").tobytes())
        f.write(np.array([1.0, 0.0, 1.0, 0.5, -0.5, 2.0], "<f8").tobytes())
        f.write(np.array([1, 1, 2], "<i4").tobytes())
        f.write(np.array([10, 10, 99], "<i4").tobytes())
    # case 2: the same header fields, then the precision flag; y, W, fe1, fe2, weights in FP32
    with open(os.path.join(HERE, "loader_case_fp32_flag.bin"), "wb") as f:
        f.write(struct.pack("<qiiii", 2, 1, 2, 0, 1))
        f.write(struct.pack("<i", 1))
        f.write(np.array([0.25, 0.75], "<f4").tobytes())
        f.write(np.array([-1.5, 2.5], "<f4").tobytes())
        f.write(np.array([3, 4], "<i4").tobytes())
        f.write(np.array([1, 1], "<i4").tobytes())
        f.write(np.array([1.5, 2.5], "<f4").tobytes())
    cases = {
        "loader_case_fp64_noflag.bin": {
            "header": {"n_obs": 3, "n_regressors": 2, "n_fe": 1, "has_cluster": 1, "has_weights": 0, "precision": 0},
            "tol": 1e-12, "y": [1.0, 2.0, -1.0], "W": [[1.0, 0.0, 1.0], [0.5, -0.5, 2.0]],
            "fe": [[1, 1, 2]], "cluster": [10, 10, 99], "weights": None},
        "loader_case_fp32_flag.bin": {
            "header": {"n_obs": 2, "n_regressors": 1, "n_fe": 2, "has_cluster": 0, "has_weights": 1, "precision": 1},
            "tol": 1e-6, "y": [0.25, 0.75], "W": [[-1.5, 2.5]],
            "fe": [[3, 4], [1, 1]], "cluster": None, "weights": [1.5, 2.5]},
    }
    with open(os.path.join(HERE, "loader_cases.json"), "w") as f:
        json.dump(cases, f, indent=1)


if __name__ == "__main__":
    main()
