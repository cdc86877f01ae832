"""Algorithmic HBM bytes of the MG-PCG kernels and of a whole batched solve.

Every figure counts each operand the operation needs once: the matrix in its stored format,
x (or X) read once, b (or B) read once, y (or Y) written once.  Re-reads of x that a kernel
makes because a cache turned over are not algorithmic bytes; the PMC traffic of a profile,
divided by these figures, shows them.  `s` is the byte size of a value (8 FP64, 4 FP32), K the
number of right-hand sides a block kernel carries.
"""


def csr_spmv_bytes(nnz, nrows, ncols, s=8):
    """CSR SpMV: values + int32 columns, nrows + 1 offsets, x read, y written."""
    return (s + 4) * nnz + 4 * (nrows + 1) + s * (ncols + nrows)


def csr_jacobi_bytes(nnz, n, s=8):
    """Damped Jacobi sweep y = x + pre (b - A x): the SpMV plus b and pre read."""
    return csr_spmv_bytes(nnz, n, n, s) + 2 * s * n


def csr_residual_bytes(nnz, n, s=8):
    """Residual r = b - A x: the SpMV plus b read."""
    return csr_spmv_bytes(nnz, n, n, s) + s * n


def duvi_spmv_bytes(nnz, nrows, ncols, nblocks, nuniq, vi_bytes):
    """DuVi SpMV: a distance byte and a value index (vi_bytes) per nonzero, or the raw FP64
    value per nonzero (vi_bytes = 0, no table), a column per block, row offsets into the blocks
    and into the distances (two int32 arrays of nrows + 1), the unique-value table, x read, y
    written."""
    per_nnz = 1 + (vi_bytes if vi_bytes else 8)
    table = 8 * nuniq if vi_bytes else 0
    return per_nnz * nnz + 4 * nblocks + 8 * (nrows + 1) + table + 8 * (ncols + nrows)


def nodeblock_matrix_bytes(nblocks, npos, s, node_map=False):
    """A 3x3 node-block copy: 9 values and one int32 neighbour index per block, npos + 1 int32
    offsets, and (colour-ordered copies) an int32 node per position."""
    return 9 * s * nblocks + 4 * nblocks + 4 * (npos + 1) + (4 * npos if node_map else 0)


def nodeblock_spmm_bytes(nnz, n, k, s):
    """Node-block SpMM Y = A X with k columns: the blocks once, X read, Y written."""
    return nodeblock_matrix_bytes(nnz // 9, n // 3, s) + 2 * s * k * n


def nodeblock_gs_sweep_bytes(nnz, n, k, s):
    """One general multicolour Gauss-Seidel sweep on node blocks with k columns: the blocks once
    (colour order, with the position-to-node map), X read once, B read once, X written once
    (each colour writes its own rows).  The node's own 3x3 block replaces a separate D^-1."""
    return nodeblock_matrix_bytes(nnz // 9, n // 3, s, node_map=True) + 3 * s * k * n


def half_blocks(nblocks, npos):
    """Blocks of [own, earlier colours] (or [own, later colours]) over all nodes: every pair of
    coupled nodes has different colours, so each off-diagonal block is earlier for exactly one
    of its two nodes."""
    return npos + (nblocks - npos) // 2


def node_transfer_bytes(pairs, n_out, n_in, k, s, add):
    """Node transfer Y (+)= M X: one weight and one int32 input node per node pair, n_out / 3 + 1
    offsets, X read, Y written (and read when adding)."""
    return (4 + s) * pairs + 4 * (n_out // 3 + 1) + s * k * n_in + (2 if add else 1) * s * k * n_out


def block_vcycle_bytes(levels, k, s):
    """One V-cycle of the batching form (Gauss-Seidel, one sweep each way) applied to k columns.
    levels: finest first, dicts with n, nnz (level matrix) and, above the coarsest, r_nnz and
    p_nnz (dof entries of R and P, three per node pair); the coarsest is a dense FP64 inverse.
    Per level: the forward sweep from x = 0 (own and earlier-coloured blocks, B read, X
    written), the residual after it (own and later-coloured blocks, X read, R written), the
    restriction, the coarser cycle, the prolongation added to X, and the backward sweep."""
    total = 0
    for l, lv in enumerate(levels):
        n = lv["n"]
        if l == len(levels) - 1:
            total += 8 * n * n + 2 * s * k * n
            break
        npos, nblk = n // 3, lv["nnz"] // 9
        half = half_blocks(nblk, npos)
        nc = levels[l + 1]["n"]
        total += nodeblock_matrix_bytes(half, npos, s, node_map=True) + 4 * npos + 2 * s * k * n   # zero-start sweep
        total += nodeblock_matrix_bytes(half, npos, s) + 2 * s * k * n                             # residual
        total += node_transfer_bytes(lv["r_nnz"] // 3, nc, n, k, s, add=False)
        total += node_transfer_bytes(lv["p_nnz"] // 3, n, nc, k, s, add=True)
        total += nodeblock_gs_sweep_bytes(lv["nnz"], n, k, s)                                     # backward sweep
    return total


def block_mixed_solve_bytes(levels, k, iterations, outer_steps):
    """One solve of a block of k load cases in the precision + batching form: `outer_steps` FP64
    refinement steps holding `iterations` FP32 block-PCG iterations between them (each iteration:
    SpMM with the p.q partials, r update, one V-cycle, r.z, x and p update), an FP64 true residual
    before every step and after the last (b itself before the first, x being 0), and the
    conversions between the two precisions."""
    f, d = 4, 8
    n, nnz = levels[0]["n"], levels[0]["nnz"]
    kn = k * n
    spmm32 = nodeblock_spmm_bytes(nnz, n, k, f)
    resid64 = nodeblock_matrix_bytes(nnz // 9, n // 3, d) + 3 * d * kn
    vcyc = block_vcycle_bytes(levels, k, f)
    per_iteration = spmm32 + 3 * f * kn + vcyc + 2 * f * kn + 5 * f * kn
    per_step = (d + f) * kn + f * kn + f * kn + 2 * f * kn + (f + 2 * d) * kn   # convert, clear, norms, p = z, x += d
    total = iterations * per_iteration + outer_steps * per_step
    total += 2 * d * kn + outer_steps * resid64 + (outer_steps + 1) * 2 * d * kn   # true residuals and their norms
    return total
