"""Pins the oracle's solver drivers on the reference's manufactured exact solution
(FemCantileverBase.cs:53-62: b = K_ff U_f, so the discrete solution is U_f), on scipy's
direct solve, and on structural properties (PCG iteration bound, preconditioner symmetry)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from oracle import oracle as O


def S(M):
    return sp.csr_matrix((M.val, M.col, M.rowptr), shape=(M.nrows, M.ncols))


def energy_err(A, x, xs):
    e = x - xs
    return np.sqrt(e @ (S(A) @ e)) / np.sqrt(xs @ (S(A) @ xs))


def reference_levels(n, depth):
    levels, m, A = [], n, O.cantilever(n, n, n)
    for _ in range(depth):
        R, P = O.rp_reference(m, m, m)
        levels.append(dict(A=A, R=R, P=P))
        m //= 2
        A = O.cantilever(m, m, m)
    levels.append(dict(A=A))
    return levels


def galerkin_levels(n, nlev, dens=None):
    levels, m, A = [], n, O.cantilever(n, n, n, dens=dens)
    for _ in range(nlev - 1):
        P = O.p_trilinear(m, m, m)
        levels.append(dict(A=A, R=O.transpose(P), P=P))
        A = O.galerkin(A, P)
        m //= 2
    levels.append(dict(A=A))
    return levels


@pytest.fixture(scope="module")
def sys16():
    n = 16
    A = O.cantilever(n, n, n)
    U = O.manufactured(n, n, n)
    b = O.multiply_dok_order(n, n, n, A, U)
    return n, A, U, b


def test_gmg_reference_jacobi_and_gs_converge(sys16):
    n, A, U, b = sys16
    lv = reference_levels(n, 2)
    lam = O.eigen_upper_bound(lv[-1]["A"])
    pre = [O.jacobi_preconditioner(l["A"]) * (2 / lam) for l in lv[:-1]]
    x, st = O.gmg_solve(lv, b, pre=pre, smoother=0, level_iterations=4, tol=1e-7)
    assert st.converged == 1
    r = b - O.spmv(A, x)
    assert np.abs(r).max() <= 1e-7
    x2, st2 = O.gmg_solve(lv, b, smoother=1, level_iterations=4, tol=1e-7)
    assert st2.converged == 1 and np.abs(b - O.spmv(A, x2)).max() <= 1e-7
    # x converges to the manufactured solution as the tolerance tightens
    assert energy_err(A, x, U) < 1e-4


def test_gmg_failure_and_cap(sys16):
    n, A, U, b = sys16
    lv = reference_levels(n, 1)
    lam = O.eigen_upper_bound(lv[-1]["A"])
    pre = [O.jacobi_preconditioner(lv[0]["A"]) * (2 / lam)]
    bb = b.copy(); bb[5] = np.nan
    x, st = O.gmg_solve(lv, bb, pre=pre, level_iterations=2, tol=1e-6)
    assert st.converged == -1
    x, st = O.gmg_solve(lv, b, pre=pre, level_iterations=2, tol=0.0, max_cycles=3)
    assert st.converged == 0 and st.iterations == 4


def test_pcg_jacobi_matches_manufactured(sys16):
    n, A, U, b = sys16
    x, st = O.pcg_jacobi(A, b, tol=1e-9, max_it=5000)
    assert st.converged == 1
    assert energy_err(A, x, U) < 1e-6


def test_pcg_finite_termination():
    """CG on a diagonal system with k distinct eigenvalues converges in k iterations."""
    d = np.repeat([1.0, 2.0, 5.0, 11.0], 10)
    n = len(d)
    A = O.Csr(np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d, n, n)
    b = np.random.default_rng(0).normal(size=n)
    # with M = D^-1 the preconditioned operator is the identity: one iteration
    x, st = O.pcg_jacobi(A, b, tol=1e-12)
    assert st.converged == 1 and st.iterations == 0
    assert np.allclose(x, b / d, rtol=1e-14)


@pytest.mark.parametrize("smoother", [0, 1])
def test_mgpcg_manufactured(sys16, smoother):
    n, A, U, b = sys16
    lv = galerkin_levels(n, 3)
    x, st = O.mgpcg(lv, b, smoother=smoother, omega=0.6, nu=2, tol=1e-10)
    assert st.converged == 1 and st.iterations <= 20
    assert energy_err(A, x, U) < 1e-8


def test_mgpcg_tip_load_against_spsolve():
    n = 8
    dens = 10 ** np.random.default_rng(3).uniform(-2, 2, n ** 3)
    A = O.cantilever(n, n, n, dens=dens)
    f = O.tip_load(n, n, n)
    xs = spl.spsolve(S(A).tocsc(), f)
    lv = galerkin_levels(n, 3, dens)
    for sm in (0, 1):
        x, st = O.mgpcg(lv, f, smoother=sm, omega=0.6, nu=2, tol=1e-11)
        assert st.converged == 1
        assert energy_err(A, x, xs) < 1e-9


def test_vcycle_is_symmetric():
    """v^T M u == u^T M v and u^T M u > 0 for the MG-PCG preconditioner (R = P^T,
    symmetric smoothing), the property PCG needs."""
    n = 8
    lv = galerkin_levels(n, 3, dens=10 ** np.random.default_rng(1).uniform(-1, 1, n ** 3))
    rng = np.random.default_rng(4)
    u, v = rng.normal(size=lv[0]["A"].nrows), rng.normal(size=lv[0]["A"].nrows)
    for sm in (0, 1):
        Mu = O.vcycle_apply(lv, u, smoother=sm)
        Mv = O.vcycle_apply(lv, v, smoother=sm)
        lhs, rhs = v @ Mu, u @ Mv
        assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), abs(rhs))
        assert u @ Mu > 0


def test_openmp_build_is_bit_identical():
    """liboracle_omp.so (the all-core CPU baseline) parallelises only row-independent loops and
    keeps the dot products serial, so it must return the serial build's bits; the timed variant
    runs the same iterations as the untimed one."""
    n = 8
    dens = 10 ** np.random.default_rng(6).uniform(-2, 2, n ** 3)
    lv = galerkin_levels(n, 3, dens)
    f = O.tip_load(n, n, n)
    assert O.num_threads(omp=False) == 1
    for sm in (0, 1):
        x1, s1 = O.mgpcg(lv, f, smoother=sm, omega=0.6, nu=2, tol=1e-10)
        x2, s2 = O.mgpcg(lv, f, smoother=sm, omega=0.6, nu=2, tol=1e-10, omp=True)
        assert s1.iterations == s2.iterations and x1.tobytes() == x2.tobytes()
        x3, s3 = O.mgpcg(lv, f, smoother=sm, omega=0.6, nu=2, tol=0.0, max_it=5)
        x4, s4, t = O.mgpcg_timed(lv, f, 5, smoother=sm, omega=0.6, nu=2, omp=True)
        assert s4.iterations == 5 and x3.tobytes() == x4.tobytes()
        assert np.all(t > 0)
