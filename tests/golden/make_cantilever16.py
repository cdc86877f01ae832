"""Writes cantilever16.json and cantilever16_U.npy: frozen fixtures of the 16^3 cantilever (the
CPU-runnable workload: 17^3 nodes, x = 0 clamped, 13,872 free dofs, E = 1, nu = 0.3).

* K_ff (FemCantileverBase.cs:46-104 with FemCantilever3D.cs:32-120; sorted-column CSR of the
  free-dof submatrix, DokRowMajor.cs:286-324): nnz and SHA-256 of rowptr (int32), col (int32)
  and val (float64, little endian) -- the assembly is bit-exact, so the hashes pin it;
* U_f (FemCantileverBase.cs:53-62, the manufactured Euler-Bernoulli displacements of the free
  dofs) as float64 in cantilever16_U.npy, and b = K_ff U_f summed in the reference's DOK order
  (FemCantileverBase.cs:70-82) by its SHA-256;
* the reference restriction / interpolation pair 16^3 -> 8^3 (IGeometricMultigridModel.cs:76-194,
  cone weights, rows normalised, free-dof submatrix): shapes, nnz and SHA-256 of each array.

Produced by the oracle (oracle/), whose assembly and transfers are pinned independently (Gauss
quadrature of Ke, brute-force scatter, closed-form cone weights; tests/test_oracle_fem.py);
freezing its output catches any later change of these bytes.  Run from the repository root:
python tests/golden/make_cantilever16.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402


def sha(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a, dtype).astype(np.dtype(dtype).newbyteorder("<")).tobytes()).hexdigest()


def csr_digest(M):
    return {"nrows": int(M.nrows), "ncols": int(M.ncols), "nnz": int(len(M.col)),
            "rowptr_sha256": sha(M.rowptr, np.int32), "col_sha256": sha(M.col, np.int32),
            "val_sha256": sha(M.val, np.float64)}


def main():
    n = 16
    A = O.cantilever(n, n, n)
    U = O.manufactured(n, n, n)
    b = O.multiply_dok_order(n, n, n, A, U)
    R, P = O.rp_reference(n, n, n)
    out = {
        "source": "oracle restatement of FemCantileverBase.cs:46-104 / FemCantilever3D.cs:32-120 (K_ff), "
                  "FemCantileverBase.cs:53-82 (U_f, b), IGeometricMultigridModel.cs:76-194 (R, P)",
        "mesh": [n, n, n], "E": 1.0, "nu": 0.3,
        "K_ff": csr_digest(A),
        "U_f": {"file": "cantilever16_U.npy", "n": int(U.size), "sha256": sha(U, np.float64)},
        "b_dok_order_sha256": sha(b, np.float64),
        "R": csr_digest(R), "P": csr_digest(P),
    }
    np.save(os.path.join(HERE, "cantilever16_U.npy"), np.ascontiguousarray(U, "<f8"), allow_pickle=False)
    with open(os.path.join(HERE, "cantilever16.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
