"""Write the SuiteSparse fixtures of the sparse-eigensolver experiment.

Input: the three MAT v5 matrices the reference's experiment loads
(MATLAB/data/{bcsstk01,bcsstk03,1138_bus}.mat, read with scipy.io.loadmat,
which parses the MAT format and executes nothing).

Output under tests/golden/suitesparse/:
* <name>_row_ptr.bin (int32, rows + 1), <name>_col_ind.bin (int32),
  <name>_values.bin (float64): the raw matrix in the binary CSR format that
  MatrixReader::read_csr_matrix reads (CPP/io/matrix_reader.cpp:8-45);
* test4.json: per matrix the preprocessing of MATLAB/+tests/+test4/test.m:31-49
  (K = A / scale; K = (K + K^T) / 2; when min eig(K) <= 0, K += 1.01 |min eig| I),
  the (k, kdim, maxiter) of test_configs.m:25-52, and the 40 largest eigenvalues
  of the preprocessed K from LAPACK's dense symmetric eigensolver (numpy.linalg.eigvalsh,
  FP64) as golden values.

Run in a container that holds the reference checkout:
    python tests/golden/make_suitesparse.py /path/to/reference
"""
import json
import os
import sys

import numpy as np
import scipy.io as sio
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "suitesparse")

# name: (scale, k, kdim, maxiter) -- test4/test.m:31-39, test4/test_configs.m:25-52
CASES = {
    "bcsstk01": (1e8, 5, 20, 20),
    "bcsstk03": (1e9, 10, 50, 200),
    "1138_bus": (1e2, 20, 100, 400),
}


def main(ref):
    os.makedirs(OUT, exist_ok=True)
    meta = {}
    for name, (scale, k, kdim, maxiter) in CASES.items():
        d = sio.loadmat(os.path.join(ref, "MATLAB", "data", name + ".mat"))
        A = sp.csr_matrix(d["Problem"]["A"][0, 0])
        A.sort_indices()
        A.indptr.astype(np.int32).tofile(os.path.join(OUT, name + "_row_ptr.bin"))
        A.indices.astype(np.int32).tofile(os.path.join(OUT, name + "_col_ind.bin"))
        A.data.astype(np.float64).tofile(os.path.join(OUT, name + "_values.bin"))
        K = A.toarray() / scale
        K = (K + K.T) / 2
        w = np.linalg.eigvalsh(K)
        shift = 0.0
        if w[0] <= 0:
            shift = abs(w[0]) * 1.01
            w = np.linalg.eigvalsh(K + shift * np.eye(K.shape[0]))
        meta[name] = {"rows": int(A.shape[0]), "cols": int(A.shape[1]), "nnz": int(A.nnz),
                      "scale": scale, "shift": shift, "k": k, "kdim": kdim, "maxiter": maxiter,
                      "top_eigenvalues": [float(x) for x in w[::-1][:40]],
                      "min_eigenvalue": float(w[0])}
        print(name, A.shape, A.nnz, "shift", shift)
    with open(os.path.join(OUT, "test4.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
