"""Pins the oracle's sparse kernels, smoothers, colouring, dense solve and Galerkin product
against scipy / numpy, closed forms and the reference's own KAT
(GMCantileverOpenCLTests.cs:16-52)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from oracle import oracle as O


def rand_csr(n, m, density, seed, empty_rows=()):
    rng = np.random.default_rng(seed)
    A = sp.random(n, m, density=density, random_state=rng, format="csr")
    A = A.tolil()
    for r in empty_rows:
        A.rows[r] = []
        A.data[r] = []
    A = A.tocsr()
    A.sort_indices()
    return O.Csr(A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64), n, m), A


def test_opencl_diagonal_kat():
    """GMCantileverOpenCLTests.cs:16-52: 16x16 diagonal, x = values => y_i = v_i^2 exactly."""
    idx = np.arange(17, dtype=np.int32)
    vals = np.arange(16, 0, -1).astype(np.float64)
    A = O.Csr(idx, idx[:16].copy(), vals, 16, 16)
    y = O.spmv(A, vals)
    assert np.array_equal(y, vals * vals)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_spmv_residual_jacobi_against_scipy(seed):
    A, S = rand_csr(200, 200, 0.05, seed, empty_rows=(3, 77))
    rng = np.random.default_rng(seed + 10)
    x, b, pre = rng.normal(size=200), rng.normal(size=200), rng.uniform(0.5, 1.5, 200)
    ref = S @ x
    assert np.allclose(O.spmv(A, x), ref, rtol=1e-14, atol=1e-14)
    y0 = rng.normal(size=200)
    assert np.allclose(O.spmv(A, x, y0.copy(), clear=False), y0 + ref, rtol=1e-13, atol=1e-13)
    assert np.allclose(O.residual(A, b, x), b - ref, rtol=1e-13, atol=1e-13)
    assert np.allclose(O.jacobi(A, pre, b, x), x + pre * (b - ref), rtol=1e-13, atol=1e-13)
    assert np.array_equal(O.jacobi_initial(pre, b), pre * b)
    assert np.isclose(O.lib().orc_dot(200, O._p(x), O._p(b)), x @ b, rtol=1e-13)


def test_residual_check_flag_semantics():
    """CsrGeometricMultigrid.cl:138-140: 1 = converged, bit 1 = NaN/overflow, 0 = not yet."""
    idx = np.arange(5, dtype=np.int32)
    A = O.Csr(idx, idx[:4].copy(), np.ones(4), 4, 4)
    x = np.zeros(4)
    assert O.residual_check(A, np.full(4, 1e-7), x, 1e-6)[1] == 1
    assert O.residual_check(A, np.array([1e-7, 1e-3, 0, 0]), x, 1e-6)[1] == 0
    assert O.residual_check(A, np.array([np.nan, 0, 0, 0]), x, 1e-6)[1] & 2
    assert O.residual_check(A, np.array([1e51, 0, 0, 0]), x, 1e-6)[1] & 2
    assert O.residual_check(A, np.array([1e-3, 1e51, 0, 0]), x, 1e-6)[1] == 2


def spd_csr(n, seed):
    A, S = rand_csr(n, n, 0.08, seed)
    S = S + S.T + sp.eye(n) * (abs(S).sum(axis=1).max() + 2)
    S = S.tocsr(); S.sort_indices()
    return O.Csr(S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data, n, n), S


def test_gauss_seidel_closed_form():
    """One forward sweep = (D + L)^-1 (b - U x)."""
    A, S = spd_csr(90, 4)
    rng = np.random.default_rng(1)
    b, x = rng.normal(size=90), rng.normal(size=90)
    Lw = sp.tril(S, 0).tocsr()
    U = sp.triu(S, 1).tocsr()
    ref = spl.spsolve_triangular(Lw, b - U @ x, lower=True)
    assert np.allclose(O.gs_forward(A, b, x), ref, rtol=1e-12, atol=1e-12)


def test_multicolor_gs_is_gs_in_colour_order():
    n = 6
    A = O.cantilever(n, n, n)
    ncol, color = O.block_coloring(A)
    # a valid colouring: no two coupled nodes share a colour
    for r in range(A.nrows):
        cols = A.col[A.rowptr[r]:A.rowptr[r + 1]]
        nb = np.unique(cols // 3)
        nb = nb[nb != r // 3]
        assert np.all(color[nb] != color[r // 3])
    assert ncol == 8
    # forward multicolour sweep == plain forward GS on the colour-ordered permutation
    order = np.concatenate([np.nonzero(color == c)[0] for c in range(ncol)])
    perm = (3 * order[:, None] + np.arange(3)[None, :]).reshape(-1)
    S = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.nrows, A.nrows))
    Sp = S[perm][:, perm].tocsr(); Sp.sort_indices()
    Ap = O.Csr(Sp.indptr.astype(np.int32), Sp.indices.astype(np.int32), Sp.data, A.nrows, A.nrows)
    rng = np.random.default_rng(3)
    b, x = rng.normal(size=A.nrows), rng.normal(size=A.nrows)
    y1 = O.gs_multicolor(A, ncol, color, b, x)
    y2 = O.gs_forward(Ap, b[perm], x[perm])
    assert np.allclose(y1[perm], y2, rtol=1e-12, atol=1e-12)
    # backward sweep == forward sweep of the reversed ordering
    rev = perm[::-1]
    Sr = S[rev][:, rev].tocsr(); Sr.sort_indices()
    Ar = O.Csr(Sr.indptr.astype(np.int32), Sr.indices.astype(np.int32), Sr.data, A.nrows, A.nrows)
    y3 = O.gs_multicolor(A, ncol, color, b, x, backward=True)
    y4 = O.gs_forward(Ar, b[rev], x[rev])
    assert np.allclose(y3[rev], y4, rtol=1e-12, atol=1e-12)


def test_ordered_gs_is_gs_in_that_order():
    """orc_gs_ordered: the forward sweep in any node order equals the textbook forward GS on the
    matrix permuted to that order (a closed-form lower-triangular solve), the backward sweep the
    forward GS of the reversed order; with the colour-list order it is the multicolour sweep."""
    n = 5
    A = O.cantilever(n, n, n, dens=10 ** (2 * O.rng_uniform(4, n ** 3) - 1))
    S = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.nrows, A.nrows))
    rng = np.random.default_rng(8)
    b, x = rng.normal(size=A.nrows), rng.normal(size=A.nrows)
    order = rng.permutation(A.nrows // 3).astype(np.int32)
    perm = (3 * order[:, None] + np.arange(3)[None, :]).reshape(-1)
    Sp = S[perm][:, perm].tocsr()
    ref = spl.spsolve_triangular(sp.tril(Sp, 0).tocsr(), b[perm] - sp.triu(Sp, 1).tocsr() @ x[perm], lower=True)
    assert np.allclose(O.gs_ordered(A, order, b, x)[perm], ref, rtol=1e-12, atol=1e-12)
    rev = perm[::-1]
    Sr = S[rev][:, rev].tocsr()
    ref = spl.spsolve_triangular(sp.tril(Sr, 0).tocsr(), b[rev] - sp.triu(Sr, 1).tocsr() @ x[rev], lower=True)
    assert np.allclose(O.gs_ordered(A, order, b, x, backward=True)[rev], ref, rtol=1e-12, atol=1e-12)
    ncol, color = O.block_coloring(A)
    corder = np.concatenate([np.nonzero(color == c)[0] for c in range(ncol)]).astype(np.int32)
    assert np.array_equal(O.gs_ordered(A, corder, b, x), O.gs_multicolor(A, ncol, color, b, x))
    assert np.array_equal(O.gs_ordered(A, corder, b, x, True), O.gs_multicolor(A, ncol, color, b, x, True))


def test_cholesky_solve_against_numpy():
    A, S = spd_csr(120, 9)
    b = np.random.default_rng(2).normal(size=120)
    assert np.allclose(O.cholesky_solve_csr(A, b), np.linalg.solve(S.toarray(), b), rtol=1e-12, atol=1e-12)


def test_eigen_upper_bound_closed_form():
    """max_i sum_j |a_ij| / max_j |a_ij| (GeometricMultigridSolver.cs:144-151)."""
    A, S = spd_csr(50, 5)
    D = abs(S).toarray()
    ref = max(D[i].sum() / D[i].max() for i in range(50))
    assert np.isclose(O.eigen_upper_bound(A), ref, rtol=1e-14)


def test_galerkin_against_dense():
    n = 4
    A = O.cantilever(n, n, n, dens=10 ** np.random.default_rng(0).uniform(-1, 1, n ** 3))
    P = O.p_trilinear(n, n, n)
    C = O.galerkin(A, P)
    Sd = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.nrows, A.nrows)).toarray()
    Pd = sp.csr_matrix((P.val, P.col, P.rowptr), shape=(P.nrows, P.ncols)).toarray()
    ref = Pd.T @ Sd @ Pd
    Cd = sp.csr_matrix((C.val, C.col, C.rowptr), shape=(C.nrows, C.ncols)).toarray()
    assert np.allclose(Cd, ref, rtol=1e-12, atol=1e-12 * abs(ref).max())
    # the Galerkin operator keeps the 27-node stencil pattern of the coarse mesh
    A2 = O.cantilever(n // 2, n // 2, n // 2)
    assert np.array_equal(C.rowptr, A2.rowptr) and np.array_equal(C.col, A2.col)


def test_transpose_roundtrip():
    A, S = rand_csr(40, 70, 0.1, 3)
    T = O.transpose(A)
    ref = S.T.tocsr(); ref.sort_indices()
    assert np.array_equal(T.rowptr, ref.indptr) and np.array_equal(T.col, ref.indices)
    assert np.array_equal(T.val, ref.data)
