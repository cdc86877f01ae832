"""Writes duvi_kat.json: the DuVi known-answer vectors of the reference's unit test
CompressedSparseMatrixTests.CheckDenseAndSparseToCompressedConversion
(tests/MGroup.LinearAlgebra.Tests/Matrices/CompressedSparseMatrixTests.cs:15-39).

The input matrix is the one that test builds (2 x 1024; row 0: cols 5, 6, 8, 1000, 1023 with
values 1, 1, 1, 2, 2; row 1: cols 50, 1020 with values 3, 3).  The expected encodings were
derived by hand from the encoder source (DuViCompressedSparseMatrix.cs:13-88,285-400 for the
unit stream; OpenCLDuViGeometricMultigridSolver.cs:753-851 for the GPU block layout):
  unit 1: header 0x01 (new row, 1-byte deltas), 2 deltas, next = 5, deltas 1, 2
  unit 2: header 0x00, 1 delta, next = 992 = 128 + 0x360 -> varint 0x83 0x60, delta 23
  unit 3: header 0x03 (new row, 2-byte deltas), 1 delta, next = 50, delta 970 = 0x03CA (LE)
and y = A x with x = row 0 of the dense matrix is [11, 0] (the test's assertion).
No reference code is run.
"""
import json
import os

KAT = {
    "source": "CompressedSparseMatrixTests.cs:15-39 (matrix and y = [11, 0]); encodings derived from "
              "DuViCompressedSparseMatrix.cs:13-88,285-400 and OpenCLDuViGeometricMultigridSolver.cs:753-851",
    "nrows": 2, "ncols": 1024,
    "rowptr": [0, 5, 7], "col": [5, 6, 8, 1000, 1023, 50, 1020], "val": [1, 1, 1, 2, 2, 3, 3],
    "x_nonzeros": {"5": 1, "6": 1, "8": 1, "1000": 2, "1023": 2},
    "y": [11, 0],
    "units": {
        "bytes": [0x01, 0x02, 0x05, 0x01, 0x02, 0x00, 0x01, 0x83, 0x60, 0x17, 0x03, 0x01, 0x32, 0xCA, 0x03],
        "unit_off": [0, 5, 10, 15],
        "value_indices": [0, 0, 0, 1, 1, 2, 2],
        "unique_values": [1.0, 2.0, 3.0],
    },
    "block": {
        "row_blocks": [0, 2, 4], "row_dist": [0, 5, 7], "block_col": [5, 1000, 50, 1020],
        "dist": [2, 0, 1, 1, 22, 0, 0], "vi_bytes": 1, "vi": [0, 0, 0, 1, 1, 2, 2], "values": [1.0, 2.0, 3.0],
    },
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "duvi_kat.json")
    with open(out, "w") as f:
        json.dump(KAT, f, indent=1)
    print(out)
