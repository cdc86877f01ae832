"""The algorithmic byte counts the rooflines are computed from: the published per-SpMV figures of
the workloads (128^3: 509,597,550 nonzeros, 6,390,144 free dofs; 64^3: 63,695,790 and 811,200),
and the colour-split block count against the oracle's node colouring."""
import numpy as np
import pytest

from compressiontests_amd import traffic as T
from oracle import oracle as O


def test_csr_spmv_bytes_of_the_workloads():
    assert round(T.csr_spmv_bytes(63_695_790, 811_200, 811_200) / 1e6, 1) == 780.6
    assert round(T.csr_spmv_bytes(509_597_550, 6_390_144, 6_390_144) / 1e6) == 6243
    assert abs(T.csr_spmv_bytes(509_597_550, 6_390_144, 6_390_144, s=4) / 1e6 - 4154) < 1   # 4,153.5
    assert T.csr_jacobi_bytes(100, 10) == T.csr_spmv_bytes(100, 10, 10) + 160
    assert T.csr_residual_bytes(100, 10) == T.csr_spmv_bytes(100, 10, 10) + 80


def test_gs_sweep_counts_x_once():
    """128^3, FP32, 32 columns: blocks (2,281.9 MB with the node map) + X read + B read + X
    written (817.9 MB each)."""
    nnz, n = 509_597_550, 6_390_144
    m = T.nodeblock_matrix_bytes(nnz // 9, n // 3, 4, node_map=True)
    assert round(m / 1e6, 1) == 2281.9
    assert round(4 * 32 * n / 1e6, 1) == 817.9
    assert T.nodeblock_gs_sweep_bytes(nnz, n, 32, 4) == m + 3 * 4 * 32 * n


def test_duvi_bytes_raw_and_indexed():
    # raw values: 9 bytes per nonzero and no table; u8 indices: 2 bytes per nonzero and the table
    assert T.duvi_spmv_bytes(10, 2, 5, 3, 4, 0) == 90 + 12 + 24 + 56
    assert T.duvi_spmv_bytes(10, 2, 5, 3, 4, 1) == 20 + 12 + 24 + 32 + 56


@pytest.mark.parametrize("n", [4, 6])
def test_half_blocks_matches_the_colouring(n):
    """Own block plus the blocks of earlier-coloured neighbours, summed over all nodes, equals
    half_blocks for the oracle's greedy node colouring of the stiffness graph."""
    A = O.cantilever(n, n, n)
    ncol, col = O.block_coloring(A)
    rp, ci = A.rowptr, A.col
    npos = A.nrows // 3
    lower = upper = nblk = 0
    for q in range(npos):
        nodes = np.unique(ci[rp[3 * q]:rp[3 * q + 1]] // 3)
        nblk += len(nodes)
        lower += 1 + int(np.sum(col[nodes] < col[q]))
        upper += 1 + int(np.sum(col[nodes] > col[q]))
    assert lower == upper == T.half_blocks(nblk, npos)
    assert nblk * 9 == A.val.size


def test_block_solve_bytes_add_up():
    """The solve total is linear in the iteration count, with one V-cycle, SpMM and vector
    updates per iteration."""
    lv = [dict(n=300, nnz=9 * 300 * 20, r_nnz=3 * 120, p_nnz=3 * 120), dict(n=54)]
    a = T.block_mixed_solve_bytes(lv, 4, 10, 2)
    b = T.block_mixed_solve_bytes(lv, 4, 11, 2)
    per_it = b - a
    assert per_it == (T.nodeblock_spmm_bytes(lv[0]["nnz"], 300, 4, 4) + T.block_vcycle_bytes(lv, 4, 4)
                      + 10 * 4 * 4 * 300)
