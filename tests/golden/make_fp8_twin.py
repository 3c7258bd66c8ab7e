"""Golden eigenvalues of the scaled twin of the large FP8 solve: the 3-D
7-point variable-coefficient diffusion operator (face coefficients
10^(2u-1), u ~ U[0,1) from the Philox stream of coefficient seed 5, the
generator of ofrr_stencil_csr restated by the oracle) on a 32^3 grid,
its 110 largest eigenvalues by scipy's FP64 Lanczos (ARPACK eigsh), checked
against a second run from another start vector.  Writes fp8_twin.json."""
import json
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

N, SEED, NEV = 32, 5, 110
A = O.stencil(3, N, seed=SEED)
M = sp.csr_matrix((A.vals, A.col_idx, A.row_ptr), shape=(A.n, A.n))
assert abs(M - M.T).max() == 0.0
w1 = np.sort(sla.eigsh(M, k=NEV, which="LA", tol=1e-14, ncv=300,
                       v0=np.ones(A.n))[0])[::-1]
w2 = np.sort(sla.eigsh(M, k=NEV, which="LA", tol=1e-14, ncv=300,
                       v0=np.random.default_rng(1).standard_normal(A.n))[0])[::-1]
assert np.max(np.abs(w1 - w2) / w1[0]) < 1e-12, np.max(np.abs(w1 - w2))
out = {"dim": 3, "N": N, "coef_seed": SEED, "n": int(A.n), "nnz": int(A.nnz),
       "top_eigenvalues": [float(x) for x in w1],
       "method": "scipy.sparse.linalg.eigsh (ARPACK), which='LA', tol=1e-14, two start vectors"}
with open(os.path.join(HERE, "fp8_twin.json"), "w") as f:
    json.dump(out, f, indent=1)
print("lambda_1", w1[0], "lambda_100", w1[99], "lambda_101", w1[100])
