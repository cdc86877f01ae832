"""The identities behind the zero-start pre-smoothing of the batching form's V-cycle (CPU, numpy):
a forward multicolour Gauss-Seidel sweep from x = 0, node by node with each node's dofs relaxed
in order, equals the sweep that reads only the blocks of earlier colours with the node's own
x taken as 0; and its residual b - A x equals minus the pull of the later colours plus the
strictly upper part of each node's own 3x3 block.  Pinned against a direct dense sweep and
b - A x on the 4^3 cantilever with heterogeneous moduli."""
import numpy as np
import scipy.sparse as sp

from oracle import oracle as O


def _colouring(A3, nn):
    adj = [set() for _ in range(nn)]
    for i in range(nn):
        for j in range(nn):
            if i != j and np.any(A3[3 * i:3 * i + 3, 3 * j:3 * j + 3] != 0):
                adj[i].add(j)
    col = [-1] * nn
    for i in range(nn):
        used = {col[j] for j in adj[i] if col[j] >= 0}
        c = 0
        while c in used:
            c += 1
        col[i] = c
    return adj, col


def test_zero_start_lower_sweep_and_upper_residual():
    n = 4
    A = O.cantilever(n, n, n, dens=np.exp(O.rng_normal(5, n ** 3)))
    S = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.nrows, A.ncols)).toarray()
    nn = A.nrows // 3
    adj, col = _colouring(S, nn)
    nc = max(col) + 1
    assert nc == 8
    b = O.rng_normal(9, A.nrows)
    # direct forward sweep from zero, colour by colour, dofs in order
    x = np.zeros(A.nrows)
    for c in range(nc):
        for i in (i for i in range(nn) if col[i] == c):
            for a in range(3):
                r = 3 * i + a
                x[r] = (b[r] - S[r] @ x + S[r, r] * x[r]) / S[r, r]
    # the lower-colour sweep: earlier colours only, own x from 0; x never cleared (NaN elsewhere)
    xl = np.full(A.nrows, np.nan)
    for c in range(nc):
        for i in (i for i in range(nn) if col[i] == c):
            y = np.zeros(3)
            for j in adj[i]:
                if col[j] < c:
                    y += S[3 * i:3 * i + 3, 3 * j:3 * j + 3] @ xl[3 * j:3 * j + 3]
            d = S[3 * i:3 * i + 3, 3 * i:3 * i + 3]
            x0 = (b[3 * i] - y[0]) / d[0, 0]
            x1 = (b[3 * i + 1] - y[1] - d[1, 0] * x0) / d[1, 1]
            x2 = (b[3 * i + 2] - y[2] - d[2, 0] * x0 - d[2, 1] * x1) / d[2, 2]
            xl[3 * i:3 * i + 3] = (x0, x1, x2)
    assert not np.any(np.isnan(xl))
    assert np.max(np.abs(x - xl)) <= 1e-13 * np.max(np.abs(x))
    # residual as the pull of later colours + strictly upper own block
    r_direct = b - S @ x
    r_up = np.zeros(A.nrows)
    for i in range(nn):
        y = np.zeros(3)
        for j in adj[i]:
            if col[j] > col[i]:
                y += S[3 * i:3 * i + 3, 3 * j:3 * j + 3] @ xl[3 * j:3 * j + 3]
        d = S[3 * i:3 * i + 3, 3 * i:3 * i + 3]
        y[0] += d[0, 1] * xl[3 * i + 1] + d[0, 2] * xl[3 * i + 2]
        y[1] += d[1, 2] * xl[3 * i + 2]
        r_up[3 * i:3 * i + 3] = -y
    assert np.max(np.abs(r_direct - r_up)) <= 1e-12 * np.max(np.abs(r_direct))
    # and a mistake in the identity (the own block's upper part left out) is visible
    r_bad = r_up.copy()
    for i in range(nn):
        d = S[3 * i:3 * i + 3, 3 * i:3 * i + 3]
        r_bad[3 * i] += d[0, 1] * xl[3 * i + 1] + d[0, 2] * xl[3 * i + 2]
    assert np.max(np.abs(r_direct - r_bad)) > 1e-6 * np.max(np.abs(r_direct))
