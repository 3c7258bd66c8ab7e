"""Writes reference_kats.json: known-answer vectors for the OFRR path.

The numbers are data transcribed from the reference's own GoogleTest suites
(paths relative to the reference repository root, cited per entry), from the
Random123 distribution's published Philox4x32-10 known-answer vectors, and
from closed forms (Dirichlet Laplacian spectra).  Nothing here runs reference
code.  Re-run to regenerate: python tests/golden/make_kats.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def lap_eigs(dim, N, count):
    q = np.arange(1, N + 1)
    l1 = 4.0 * np.sin(q * np.pi / (2.0 * (N + 1))) ** 2
    if dim == 2:
        L = (l1[:, None] + l1[None, :]).ravel()
    else:
        L = (l1[:, None, None] + l1[None, :, None] + l1[None, None, :]).ravel()
    return [float(v) for v in np.sort(L)[::-1][:count]]


kats = {
    "philox4x32_10": {
        "source": "Random123 kat_vectors (Salmon et al., SC'11), philox4x32_10",
        "cases": [
            {"ctr": [0, 0, 0, 0], "key": [0, 0],
             "out": [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]},
            {"ctr": [0xffffffff] * 4, "key": [0xffffffff] * 2,
             "out": [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]},
            {"ctr": [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
             "key": [0xa4093822, 0x299f31d0],
             "out": [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]},
        ],
    },
    "syev_tridiag10": {
        "source": "CPP/tests/mvops_test.cpp:2292-2447 (tridiag(-1,2,-1), n=10, ascending)",
        "n": 10,
        "eigenvalues": [0.0810140527710052, 0.317492934337638, 0.69027853210943,
                        1.16916997399623, 1.71537032345343, 2.28462967654657,
                        2.83083002600377, 3.30972146789057, 3.68250706566236,
                        3.91898594722899],
        "tol_double": 1e-10,
    },
    "sygv_tridiag10_diag": {
        "source": "CPP/tests/mvops_test.cpp:2450-2626 (A = tridiag(-1,2,-1), B = diag(1..10))",
        "n": 10,
        "eigenvalues": [0.0141309446511971, 0.0594567850513288, 0.131202907506177,
                        0.220848003169762, 0.316795717268112, 0.410267468002017,
                        0.52158439013008, 0.705660103820319, 1.08997474266915,
                        2.38801544566837],
        "tol_double": 1e-10,
    },
    "gemm_3x2_2x4": {
        "source": "CPP/tests/mvops_test.cpp:455-721",
        "A": [[1.0, 3.0], [2.0, 4.0], [5.0, 6.0]],
        "B": [[1.0, 2.0, 3.0, 4.0], [5.0, 6.0, 7.0, 8.0]],
        "C": [[16.0, 20.0, 24.0, 28.0], [22.0, 28.0, 34.0, 40.0], [35.0, 46.0, 57.0, 68.0]],
        "fp16_tol": 0.2,
    },
    "iamax": {
        "source": "CPP/tests/mvops_test.cpp:774-1150 (x[i] = (i - 65536)/131072, one planted "
                  "entry; result is the 1-based index and the signed value)",
        "n": 131072,
        "cases": [
            {"dtype": "f64", "idx0": 72345, "val": -8.0},
            {"dtype": "f32", "idx0": 77654, "val": 10.0},
            {"dtype": "f64", "idx0": 76666, "val": -9.0},
            {"dtype": "f32", "idx0": 88888, "val": 12.0},
            {"dtype": "f16", "idx0": 73333, "val": -21.0},
            {"dtype": "f16", "idx0": 73333, "val": 33.0},
            {"dtype": "f16", "idx0": 75678, "val": 15.0},
        ],
        "alternating": {"source": "mvops_test.cpp:938-962 (x[i] = +-i, answer n)", "n": 131072},
    },
    "matrix_iamax": {
        "source": "CPP/tests/mvops_test.cpp:1151-1497 (rows 10000, cols 1000, "
                  "A(r,c) = (r-5000)/10000, planted A(rand()%rows, c) = +-(c+1) after srand(42))",
        "rows": 10000, "cols": 1000, "srand": 42,
    },
    "hessenberg": {
        "source": "CPP/tests/hessenberg_test.cpp:29-358",
        "shapes": {"f64": [2000, 200], "f32": [2000, 200], "f16": [2000, 100]},
        "qr_tol": {"f64": 1e-10, "f32": 1e-5, "f16": 5e-2},
        "dup_select_tol": {"f64": 1e-12, "f32": 1e-6, "f16": 1e-2},
        "dup_columns": {"copied_from": [0, 1, 2], "copied_to": [3, 4, 5]},
    },
    "qr": {
        "source": "CPP/tests/qr_test.cpp:30-651 (A uniform [-1, 1), Q^T Q = I and A = Q R "
                  "within tol, then columns 3..5 copied from 0..2 and dropped)",
        "shape": [2000, 200],
        "tol": {"mgs": {"f64": 1e-10, "f32": 2e-6, "f16": 2e-2},
                "cgs2": {"f64": 1e-10, "f32": 2e-6, "f16": 2e-2},
                "mgs_v2": {"f64": 1e-10}},
        "dup_select_tol": {"mgs": {"f64": 1e-12, "f32": 1e-6, "f16": 1e-2},
                           "cgs2": {"f64": 1e-12, "f32": 5e-6, "f16": 5e-2},
                           "mgs_v2": {"f64": 1e-12}},
        "dup_orth_tol": {"mgs": {"f64": 1e-12, "f32": 1e-6, "f16": 1e-2}},
        "reorth_tol": 0.7071067811865476,
        "dup_columns": {"copied_from": [0, 1, 2], "copied_to": [3, 4, 5]},
    },
    "laplacian": {
        "source": "closed form lambda = sum_d 4 sin^2(p_d pi / (2 (N+1)))",
        "lap2d_N64": {"n": 4096, "nnz": 20224, "top": lap_eigs(2, 64, 24)},
        "lap3d_N128": {"n": 2097152, "nnz": 14581760, "top": lap_eigs(3, 128, 70)},
        "lap3d_N256": {"n": 16777216, "nnz": 117047296},
    },
}

if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)
    print("wrote reference_kats.json")
