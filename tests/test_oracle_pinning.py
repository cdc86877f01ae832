"""Pins two oracle functions that the solver tests lean on, against mathematics rather than a
second copy of their code:

* orc_power_bound -- EigenValueUpperBound, GeometricMultigridSolver.cs:80-103 (power method on
  D^-1 A from x = 1/sqrt(n), stop when successive norms agree within 5 %, then extrapolate),
  and the 1.2 safety factor of RelaxedJacobiPreconditioner (:110-121) as the oracle V-cycle
  applies it.  Closed forms: matrices whose start vector is an eigenvector of D^-1 A (one or
  two eigenvalues), where every step of the iteration is known exactly; numpy's spectrum of
  D^-1 A on heterogeneous stiffness matrices.

* orc_gmg_solve -- the GPU stationary V-cycle of OpenCLCsrGeometricMultigridSolver.cs:334-536
  at depth 1 (levelDown = [T, F]): level-0 smoothing with the zero-start rules (:390-392,
  :422-439), the Jacobi ping-pong of sweeps in pairs (:441-452, odd counts rounded up), the
  residual check before restriction (:458-489), the coarse solve and the prolongation added to
  x (:357-382).  The check is a dense numpy two-grid iteration written from that definition:
  each smoother is a dense matrix formula (Jacobi x + W(b - A x); Gauss-Seidel
  x + (D + L_pi)^-1 (b - A x) in the node order pi of the colouring), the coarse correction
  x + P A_c^-1 R (b - A x).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O


def dense(M):
    return sp.csr_matrix((M.val, M.col, M.rowptr), shape=(M.nrows, M.ncols)).toarray()


def csr_from_dense(D):
    S = sp.csr_matrix(D)
    S.sort_indices()
    return O.Csr(S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.astype(np.float64),
                 D.shape[0], D.shape[1])


def circulant(n, diag, off):
    """Symmetric circulant tridiagonal matrix: every row sums to diag + 2 off, so the constant
    vector is an eigenvector of D^-1 A with eigenvalue (diag + 2 off) / diag."""
    D = np.zeros((n, n))
    for i in range(n):
        D[i, i] = diag
        D[i, (i + 1) % n] += off
        D[i, (i - 1) % n] += off
    return D


# ---------------------------------------------------------------- power method
def test_power_bound_diagonal_matrix_is_one():
    """D^-1 A = I: the first step gives norm 1 and ratio 1, inside the 5 % band below the
    extrapolating branch, so l = (9 * 1 + 1) / 10 = 1 exactly."""
    d = np.array([1.0, 2.0, 5.0, 11.0, 0.5, 3.0])
    A = csr_from_dense(np.diag(d))
    assert O.power_bound(A, 1 / d) == 1.0


@pytest.mark.parametrize("diag,off,expect", [
    # eigenvalue 0.5, far outside the band: l = 0.5 after step 1, step 2 has ratio 1 -> l = 0.5
    (4.0, -1.0, lambda lam: lam),
    # eigenvalue 1.02 inside the band at step 1 (ratio 1.02 > 1): l = 1.02 * 1.02^(1/sqrt(0.05))
    (1.0, 0.01, lambda lam: lam * lam ** (1 / np.sqrt(0.05))),
    # eigenvalue 0.98 inside the band at step 1 (ratio 0.98 < 1): l = (9 * 0.98 + 1) / 10
    (1.0, -0.01, lambda lam: (9 * lam + 1) / 10),
    # eigenvalue 3, ratio 3 at step 1, then ratio 1: l = (9 * 3 + 3) / 10 = 3
    (2.0, 2.0, lambda lam: lam),
])
def test_power_bound_start_vector_eigenvector(diag, off, expect):
    """The start vector 1/sqrt(n) is an eigenvector of D^-1 A with eigenvalue (diag + 2 off) /
    diag; every branch of the stop rule has a closed-form result.  Scaling the rows by a
    diagonal also pins the D^-1 multiply: the row sums of A alone are diag + 2 off times the
    row scale, not the eigenvalue."""
    n = 30
    scale = np.linspace(1.0, 7.0, n)
    D = circulant(n, diag, off) * scale[:, None]          # D^-1 A is the unscaled circulant / diag
    A = csr_from_dense(D)
    invd = 1 / np.diag(D)
    lam = (diag + 2 * off) / diag
    got = O.power_bound(A, invd)
    assert got == pytest.approx(expect(lam), rel=1e-13, abs=0)


def test_power_bound_two_eigenvalues_closed_form():
    """2 x 2 blocks [[1, c], [c, 1]] with c = 0.9 (first half) and c = 0.2 (second half): the
    start vector is the (1, 1) eigenvector of every block, with eigenvalues m1 = 1.9 and
    m2 = 1.2 on the two halves, so after k steps the iterate is (m1^k, ..., m2^k, ...) up to
    scale and the k-th norm is sqrt((m1^2k + m2^2k) / (m1^2(k-1) + m2^2(k-1))).  The reference's
    stop rule on that sequence gives the expected bound and step count."""
    c1, c2, nb = 0.9, 0.2, 8
    blocks = [np.array([[1, c], [c, 1]]) for c in [c1] * nb + [c2] * nb]
    D = np.zeros((4 * nb, 4 * nb))
    for i, B in enumerate(blocks):
        D[2 * i:2 * i + 2, 2 * i:2 * i + 2] = B
    A = csr_from_dense(D)
    m1, m2 = 1 + c1, 1 + c2
    l, k = 1.0, 0
    while True:
        k += 1
        l2 = np.sqrt((m1 ** (2 * k) + m2 ** (2 * k)) / (m1 ** (2 * k - 2) + m2 ** (2 * k - 2)))
        ratio = l2 / l
        if 0.95 < ratio < 1.05:
            expect = l2 * ratio ** (1 / np.sqrt(0.05)) if ratio > 1 else (9 * l2 + l) / 10
            break
        l = l2
    assert k >= 3                                          # the test exercises several steps
    assert O.power_bound(A, np.ones(4 * nb)) == pytest.approx(expect, rel=1e-12)


@pytest.mark.parametrize("n", [4, 8])
def test_power_bound_brackets_spectral_radius(n):
    """On heterogeneous cantilever stiffness matrices (moduli log-uniform over four decades)
    the power bound lies within 10 % above the largest eigenvalue of D^-1 K (numpy on the
    symmetric similar matrix D^-1/2 K D^-1/2): the extrapolation overshoots, never undershoots,
    so 2 / (1.2 l) is a convergent Jacobi weight."""
    dens = 10 ** (4 * O.rng_uniform(7 + n, n ** 3) - 2)
    A = O.cantilever(n, n, n, dens=dens)
    K = dense(A)
    h = 1 / np.sqrt(np.diag(K))
    lam_max = np.linalg.eigvalsh(h[:, None] * K * h[None, :]).max()
    l = O.power_bound(A, O.jacobi_preconditioner(A))
    assert lam_max <= l <= 1.1 * lam_max, (l, lam_max)


def test_power_weight_carries_the_safety_factor():
    """omega <= 0 makes the oracle V-cycle weight its Jacobi smoother by 2 / (1.2 l) with l the
    power bound (RelaxedJacobiPreconditioner, GeometricMultigridSolver.cs:110-121).  On a level
    matrix whose D^-1 A has the start vector as eigenvector (eigenvalue 0.5, so l = 0.5 exactly)
    the automatic weight must reproduce the explicit omega = 2 / 0.6 and differ from 2 / 0.5."""
    n = 12                                                 # 4 nodes x 3 dofs
    K0 = circulant(n, 4.0, -1.0)
    P = np.zeros((n, 3))
    for i in range(n):
        P[i, i % 3] = 1.0                                  # one coarse node, its 3 dofs
    K1 = P.T @ K0 @ P
    lv = [dict(A=csr_from_dense(K0), R=csr_from_dense(P.T), P=csr_from_dense(P)), dict(A=csr_from_dense(K1))]
    r = np.random.default_rng(5).normal(size=n)
    auto = O.vcycle_apply(lv, r, smoother=0, omega=0.0, nu=2)
    explicit = O.vcycle_apply(lv, r, smoother=0, omega=2 / (1.2 * 0.5), nu=2)
    no_factor = O.vcycle_apply(lv, r, smoother=0, omega=2 / 0.5, nu=2)
    assert np.allclose(auto, explicit, rtol=1e-12, atol=1e-14 * np.abs(explicit).max())
    assert np.abs(auto - no_factor).max() > 1e-3 * np.abs(auto).max()


# ---------------------------------------------------------------- stationary V-cycle
def two_grid_dense(K0, R, P, K1, b, smoother, sweeps, cycles, W=None, order=None, x0=None):
    """The GPU driver's depth-1 cycle written as dense matrix algebra.  Returns the level-0
    iterate at the residual check of cycles 0 .. cycles-1 (the state the solver returns when
    it stops at that check)."""
    n = K0.shape[0]
    x = np.zeros(n) if x0 is None else np.array(x0, float)
    given = x0 is not None
    if smoother == "gs":
        perm = np.concatenate([np.arange(3 * q, 3 * q + 3) for q in order])
        Kp = K0[np.ix_(perm, perm)]
        Lp = np.tril(Kp)                                   # D + L in the sweep's order
    out = []
    for _ in range(cycles):
        if smoother == "jacobi":
            if not given:
                x = np.zeros(n)
                count = 2 + 2 * int(np.ceil(max(sweeps - 2, 0) / 2))
            else:
                count = 2 * int(np.ceil(sweeps / 2))
            for _ in range(count):
                x = x + W * (b - K0 @ x)
        else:
            if not given:
                x = np.zeros(n)
            for _ in range(sweeps):
                d = np.zeros(n)
                d[perm] = np.linalg.solve(Lp, (b - K0 @ x)[perm])
                x = x + d
        given = True
        out.append(x.copy())
        x = x + P @ np.linalg.solve(K1, R @ (b - K0 @ x))
    return out


@pytest.fixture(scope="module")
def two_grid_4():
    n = 4
    dens = 10 ** (2 * O.rng_uniform(20261204, n ** 3) - 1)
    A0 = O.cantilever(n, n, n, dens=dens)
    R, P = O.rp_reference(n, n, n)
    A1 = O.cantilever(n // 2, n // 2, n // 2)
    b = O.tip_load(n, n, n) + 0.1 * O.rng_normal(11, A0.nrows)
    lv = [dict(A=A0, R=R, P=P), dict(A=A1)]
    return lv, dense(A0), dense(R), dense(P), dense(A1), b


@pytest.mark.parametrize("sweeps", [2, 3, 4])
@pytest.mark.parametrize("with_x0", [False, True])
def test_gmg_jacobi_matches_dense_two_grid(two_grid_4, sweeps, with_x0):
    lv, K0, R, P, K1, b = two_grid_4
    W = 0.6 / np.diag(K0)
    x0 = 0.01 * O.rng_normal(12, K0.shape[0]) if with_x0 else None
    ref = two_grid_dense(K0, R, P, K1, b, "jacobi", sweeps, 4, W=W, x0=x0)
    for stop_at in range(4):
        # tol = 0 never converges; the solver returns at the check of cycle max_cycles + 1
        x, st = O.gmg_solve(lv, b, pre=[W], smoother=0, level_iterations=sweeps, max_cycles=stop_at - 1,
                            tol=0.0, x0=x0)
        assert st.iterations == stop_at and st.converged == 0
        e = np.abs(x - ref[stop_at]).max() / np.abs(ref[stop_at]).max()
        assert e <= 1e-13, (sweeps, with_x0, stop_at, e)


@pytest.mark.parametrize("sweeps", [1, 2])
def test_gmg_gauss_seidel_matches_dense_two_grid(two_grid_4, sweeps):
    lv, K0, R, P, K1, b = two_grid_4
    ncol, col = O.block_coloring(lv[0]["A"])
    # a valid colouring: no two coupled nodes share a colour
    nodes = K0.shape[0] // 3
    coupled = np.abs(K0.reshape(nodes, 3, nodes, 3)).sum(axis=(1, 3)) > 0
    same = col[:, None] == col[None, :]
    assert not np.any(coupled & same & ~np.eye(nodes, dtype=bool))
    order = np.lexsort((np.arange(nodes), col))
    ref = two_grid_dense(K0, R, P, K1, b, "gs", sweeps, 3, order=order)
    for stop_at in range(3):
        x, st = O.gmg_solve(lv, b, smoother=1, level_iterations=sweeps, max_cycles=stop_at - 1, tol=0.0)
        assert st.iterations == stop_at
        e = np.abs(x - ref[stop_at]).max() / np.abs(ref[stop_at]).max()
        assert e <= 1e-13, (sweeps, stop_at, e)


def test_gmg_stops_at_the_first_passing_check(two_grid_4):
    """The solver returns at the first residual check whose max-norm is within tol, with the
    iteration count of that cycle (OpenCLCsrGeometricMultigridSolver.cs:473-486)."""
    lv, K0, R, P, K1, b = two_grid_4
    W = 0.6 / np.diag(K0)
    ref = two_grid_dense(K0, R, P, K1, b, "jacobi", 2, 60, W=W)
    res = [np.abs(b - K0 @ x).max() for x in ref]
    tol = 0.5 * (res[9] + res[10])                         # passes first at cycle 10 if monotone there
    first = next(i for i, r in enumerate(res) if r <= tol)
    x, st = O.gmg_solve(lv, b, pre=[W], smoother=0, level_iterations=2, tol=tol)
    assert st.converged == 1 and st.iterations == first
    assert np.abs(x - ref[first]).max() <= 1e-13 * np.abs(ref[first]).max()
