"""Solver handle (C ABI) against the oracle drivers: the stationary V-cycle (CSR and DuVi,
Jacobi and multicolour GS), Jacobi PCG, and MG-PCG.  Parity = same iteration count (+-1,
the stop test sits on a tolerance boundary and sums round differently) and solutions that
agree to the solve tolerance; MG-PCG additionally reproduces the manufactured exact solution."""
import numpy as np
import pytest
import scipy.sparse as sp

from compressiontests_amd import _lib as L
from compressiontests_amd.model import FemCantilever3D
from compressiontests_amd.solvers import (HipCsrGeometricMultigridSolver, HipCsrPcgSolver,
                                          HipDuViGeometricMultigridSolver, HipMgPcgSolver)
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def S(M):
    return sp.csr_matrix((M.val, M.col, M.rowptr), shape=(M.nrows, M.ncols))


def energy_err(A, x, xs):
    e = x - xs
    return np.sqrt(e @ (S(A) @ e)) / np.sqrt(xs @ (S(A) @ xs))


def oracle_reference_levels(n, depth):
    lv, m, A = [], n, O.cantilever(n, n, n)
    for _ in range(depth):
        R, P = O.rp_reference(m, m, m)
        lv.append(dict(A=A, R=R, P=P))
        m //= 2
        A = O.cantilever(m, m, m)
    lv.append(dict(A=A))
    return lv


def oracle_galerkin_levels(n, nlev, dens=None):
    lv, m, A = [], n, O.cantilever(n, n, n, dens=dens)
    for _ in range(nlev - 1):
        P = O.p_trilinear(m, m, m)
        lv.append(dict(A=A, R=O.transpose(P), P=P))
        A = O.galerkin(A, P)
        m //= 2
    lv.append(dict(A=A))
    return lv


@pytest.fixture(scope="module")
def m16(gpu):
    m = FemCantilever3D((16, 16, 16))
    A, U, b = m.CreateAndSolveLinearSystem()
    return m, A, U, b


@pytest.mark.parametrize("gs", [False, True])
@pytest.mark.parametrize("cls", [HipCsrGeometricMultigridSolver, HipDuViGeometricMultigridSolver])
def test_gmg_parity(m16, gs, cls):
    m, A, U, b = m16
    depth, it, tol = 2, 4, 1e-5
    solver = cls.CreateDeepV(m, GaussSeidel=gs, maxCircleIterations=20000, coarseRelaxation=True,
                             convergenceTolerance=tol, depth=depth, levelIterations=it)
    x, st, times = solver.Solve(None)
    solver.ReleaseOpenCLResources()
    lv = oracle_reference_levels(16, depth)
    lam = O.eigen_upper_bound(lv[-1]["A"])
    pre = [O.jacobi_preconditioner(l["A"]) * (2 / lam) for l in lv[:-1]]
    xo, so = O.gmg_solve(lv, b, pre=pre, smoother=1 if gs else 0, level_iterations=it, max_cycles=20000, tol=tol)
    assert st.HasConverged and so.converged == 1
    assert abs(st.NumIterationsRequired - so.iterations) <= 1
    assert np.abs(b - O.spmv(A, x)).max() <= tol
    assert energy_err(A, x, xo) < 1e-5
    assert len(times) == depth + 1


@pytest.mark.parametrize("cls", [HipCsrGeometricMultigridSolver, HipDuViGeometricMultigridSolver])
def test_gmg_device_stop_state(m16, cls):
    """The V-cycle's residual_with_check verdict stays on the device and the host reads it every
    checkEvery cycles (the reference reads it every cycle, OpenCLCsrGeometricMultigridSolver.cs:
    458-489); the cycles queued past the stopping one return at once.  Reading every cycle and
    every 5 cycles gives bit-identical solutions and the same cycle counts: converged, at the
    cycle cap, and failed on a NaN initial guess."""
    m, A, U, b = m16
    x0 = np.zeros(A.nrows)
    x0[7] = np.nan
    out = []
    for ce in (1, 5):
        s = cls.CreateDeepV(m, GaussSeidel=True, maxCircleIterations=20000, convergenceTolerance=1e-5, depth=2,
                            levelIterations=4, checkEvery=ce)
        x, st, _ = s.Solve(None)
        xn, stn, _ = s.Solve(x0)
        s.ReleaseOpenCLResources()
        c = cls.CreateDeepV(m, GaussSeidel=True, maxCircleIterations=6, convergenceTolerance=1e-12, depth=2,
                            levelIterations=4, checkEvery=ce)
        xc, stc, _ = c.Solve(None)
        c.ReleaseOpenCLResources()
        assert st.HasConverged and not stn.HasConverged and stn.Failed
        assert not stc.HasConverged and not stc.Failed
        out.append((x, st.NumIterationsRequired, stn.NumIterationsRequired, xc, stc.NumIterationsRequired))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][3], out[1][3])
    assert out[0][1:3] == out[1][1:3] and out[0][4] == out[1][4] == 7
    assert out[0][2] == 0                      # the first check sees the NaN


def test_pcg_jacobi_parity(m16):
    m, A, U, b = m16
    s = HipCsrPcgSolver(20000, 1e-5)
    s.Initialize(m)
    x, st = s.Solve(None)
    xo, so = O.pcg_jacobi(A, b, max_it=20000, tol=1e-5)
    assert st.HasConverged and so.converged == 1
    assert abs(st.NumIterationsRequired - so.iterations) <= 1
    assert np.abs(b - O.spmv(A, x)).max() <= 1e-5
    # both stop on max|r| <= 1e-5: they agree to that level, and both approach U
    assert energy_err(A, x, xo) < 1e-5
    assert abs(energy_err(A, x, U) - energy_err(A, xo, U)) < 1e-5


@pytest.mark.parametrize("smoother", ["jacobi", "gs"])
def test_mgpcg_manufactured_and_parity(m16, smoother):
    m, A, U, b = m16
    s = HipMgPcgSolver(m, levels=3, smoother=smoother, omega=0.6, nu=2, tolerance=1e-10)
    s.Initialize(b)
    x, st = s.Solve()
    xo, so = O.mgpcg(oracle_galerkin_levels(16, 3), b, smoother=1 if smoother == "gs" else 0, omega=0.6, nu=2,
                     tol=1e-10)
    assert st.HasConverged and so.converged == 1
    assert abs(st.NumIterationsRequired - so.iterations) <= max(1, so.iterations // 10)
    assert st.RelativeResidual <= 1e-10 * 1.01
    assert energy_err(A, x, U) < 1e-8
    assert energy_err(A, x, xo) < 1e-9


def test_mgpcg_heterogeneous_tip_load(gpu):
    n = 16
    dens = 10 ** (4 * O.rng_uniform(20261205, n ** 3) - 2)
    m = FemCantilever3D((n, n, n), ElementDensities=dens, rhs="tip")
    f = m.rhs()
    A = O.cantilever(n, n, n, dens=dens)
    xo, so = O.mgpcg(oracle_galerkin_levels(n, 3, dens), f, smoother=1, nu=2, tol=1e-12)
    # omega = 0: per-level weight 2 / (1.2 lambda) (power method); a fixed 0.6 diverges here
    # because lambda_max(D^-1 K) = 5.1 at 1e4 modulus contrast
    for smoother, sm in (("jacobi", 0), ("gs", 1)):
        s = HipMgPcgSolver(m, levels=3, smoother=smoother, omega=0.0, nu=2, tolerance=1e-10)
        s.Initialize(f)
        x, st = s.Solve()
        assert st.HasConverged
        assert energy_err(A, x, xo) < 1e-8
        xs, ss = O.mgpcg(oracle_galerkin_levels(n, 3, dens), f, smoother=sm, omega=0.0, nu=2, tol=1e-10)
        assert ss.converged == 1
        assert abs(st.NumIterationsRequired - ss.iterations) <= max(1, ss.iterations // 10)


def test_warm_start_and_caps(m16):
    m, A, U, b = m16
    s = HipMgPcgSolver(m, levels=3, smoother="gs", nu=1, tolerance=1e-8)
    s.Initialize(b)
    x, st = s.Solve()
    assert st.HasConverged
    x2, st2 = s.Solve(x)          # warm start from the solution: stop test passes at once
    assert st2.HasConverged and st2.NumIterationsRequired == 0
    assert np.array_equal(x2, x)
    s2 = HipMgPcgSolver(m, levels=3, smoother="gs", nu=1, tolerance=0.0, max_iterations=3)
    s2.Initialize(b)
    x3, st3 = s2.Solve()
    assert not st3.HasConverged and not st3.Failed and st3.NumIterationsRequired == 3


def test_nan_rhs_reports_failure(m16):
    m, A, U, b = m16
    bb = b.copy(); bb[100] = np.nan
    s = HipMgPcgSolver(m, levels=3, smoother="jacobi", nu=2, tolerance=1e-8)
    s.Initialize(bb)
    x, st = s.Solve()
    assert st.Failed and not st.HasConverged
    g = HipCsrGeometricMultigridSolver(100, 1e-5, [True, False], [2], GaussSeidel=False)
    with pytest.raises(L.MgxInvalidOperation):
        g.Solve(None)
    with pytest.raises(L.MgxNotImplemented):
        HipCsrGeometricMultigridSolver(100, 1e-5, [False, True], [2]).Initialize(m)


@pytest.mark.parametrize("smoother", ["jacobi", "gs"])
def test_duvi_mgpcg(m16, smoother):
    """FP64 MG-PCG on the DuVi layout (the level-0 products, the V-cycle's residual, smoother and
    transfers all decode DuVi): the manufactured solution and the oracle's iteration count."""
    m, A, U, b = m16
    s = HipMgPcgSolver(m, levels=3, smoother=smoother, omega=0.6, nu=2, tolerance=1e-10, format="duvi")
    s.Initialize(b)
    x, st = s.Solve()
    xo, so = O.mgpcg(oracle_galerkin_levels(16, 3), b, smoother=1 if smoother == "gs" else 0, omega=0.6, nu=2,
                     tol=1e-10)
    assert st.HasConverged
    assert abs(st.NumIterationsRequired - so.iterations) <= max(1, so.iterations // 10)
    assert energy_err(A, x, U) < 1e-8


def test_graph_replays_match_eager(gpu):
    """The captured hipGraph of PCG iterations, replayed over several solves with eager work and
    synchronisations in between, gives the eager result bit for bit (deterministic reductions)."""
    n = 8
    dens = 10 ** (4 * O.rng_uniform(20261205, n ** 3) - 2)
    m = FemCantilever3D((n, n, n), ElementDensities=dens, rhs="tip")
    f = m.rhs()
    eager = HipMgPcgSolver(m, levels=3, smoother="gs", nu=1, tolerance=1e-10, use_graph=False)
    eager.Initialize(f)
    xe, se = eager.Solve()
    g = HipMgPcgSolver(m, levels=3, smoother="gs", nu=1, tolerance=1e-10, use_graph=True, check_every=3)
    g.Initialize(f)
    for rep in range(4):
        xg, sg = g.Solve()
        assert sg.NumIterationsRequired == se.NumIterationsRequired
        assert np.array_equal(xg, xe), rep
        eager.Solve()                   # eager work on another solver between replays
    gj = HipMgPcgSolver(m, levels=3, smoother="jacobi", omega=0.0, nu=2, tolerance=1e-10, use_graph=True)
    gj.Initialize(f)
    ej = HipMgPcgSolver(m, levels=3, smoother="jacobi", omega=0.0, nu=2, tolerance=1e-10, use_graph=False)
    ej.Initialize(f)
    assert np.array_equal(gj.Solve()[0], ej.Solve()[0])
    assert np.array_equal(gj.Solve()[0], ej.Solve()[0])


@pytest.mark.slow
def test_mgpcg_64_heterogeneous(gpu):
    """Config 2 moduli at 64^3 (811,200 dofs): the solve reaches 1e-8 relative residual and
    agrees with a 1e-12 solve of the same system to 1e-8 in the energy norm."""
    n = 64
    dens = 10 ** (4 * O.rng_uniform(20261205, n ** 3) - 2)
    m = FemCantilever3D((n, n, n), ElementDensities=dens, rhs="tip")
    f = m.rhs()
    s = HipMgPcgSolver(m, levels=4, smoother="gs", nu=1, tolerance=1e-8)
    s.Initialize(f)
    x, st = s.Solve()
    assert st.HasConverged and st.RelativeResidual <= 1.01e-8
    t = HipMgPcgSolver(m, levels=4, smoother="gs", nu=1, tolerance=1e-12)
    t.Initialize(f)
    xt, stt = t.Solve()
    A = O.cantilever(n, n, n, dens=dens)
    assert energy_err(A, x, xt) < 1e-8


@pytest.mark.parametrize("smoother", ["gs", "jacobi"])
def test_mixed_precision_matches_fp64(gpu, smoother):
    """Precision form: FP32 SpMV / V-cycle / vectors inside FP64 iterative refinement (FP64 true
    residual, FP64 dots and alpha/beta, FP64 coarse inverse).  The refined solution meets the
    FP64 stop test on the true residual and agrees with the oracle's FP64 MG-PCG solution of the
    same system to 1e-8 in the energy norm (16^3, 1e4 modulus contrast, tip load)."""
    n = 16
    dens = 10 ** (4 * O.rng_uniform(20261205, n ** 3) - 2)
    m = FemCantilever3D((n, n, n), ElementDensities=dens, rhs="tip")
    f = m.rhs()
    A = O.cantilever(n, n, n, dens=dens)
    xo, so = O.mgpcg(oracle_galerkin_levels(n, 3, dens), f, smoother=1, nu=2, tol=1e-12)
    assert so.converged == 1
    s = HipMgPcgSolver(m, levels=3, smoother=smoother, omega=0.0, nu=2, tolerance=1e-10, precision="mixed",
                       inner_tolerance=1e-4)
    s.Initialize(f)
    for _ in range(2):                       # the second solve reuses every buffer
        x, st = s.Solve()
        assert st.HasConverged and not st.Failed
        assert st.OuterIterations >= 2       # one FP32 inner solve cannot reach 1e-10
        assert st.RelativeResidual <= 1.01e-10
        r = f - S(A) @ x
        assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(f)
        assert energy_err(A, x, xo) < 1e-8


def test_mixed_precision_maxnorm_and_caps(m16):
    m, A, U, b = m16
    s = HipMgPcgSolver(m, levels=3, smoother="gs", nu=1, tolerance=1e-9, stop="maxnorm", precision="mixed")
    s.Initialize(b)
    x, st = s.Solve()
    assert st.HasConverged
    assert np.max(np.abs(b - S(A) @ x)) <= 1e-9
    assert energy_err(A, x, U) < 1e-8
    s2 = HipMgPcgSolver(m, levels=3, smoother="gs", nu=1, tolerance=0.0, max_iterations=2, precision="mixed")
    s2.Initialize(b)
    x2, st2 = s2.Solve()
    assert not st2.HasConverged and not st2.Failed and st2.OuterIterations == 2


@pytest.mark.parametrize("smoother", ["gs", "jacobi"])
def test_duvi_mixed_precision(gpu, smoother):
    """Precision form on the DuVi layout: FP32 value table and vectors in the V-cycle and the inner
    SpMV (DuVi residual, Jacobi, Gauss-Seidel and transfer kernels), FP64 DuVi true residual.
    Against the oracle's FP64 MG-PCG on the CSR matrices (the DuVi values merge near-duplicates
    within 1e-10), 16^3 with a 1e4 modulus contrast."""
    n = 16
    dens = 10 ** (4 * O.rng_uniform(20261205, n ** 3) - 2)
    m = FemCantilever3D((n, n, n), ElementDensities=dens, rhs="tip")
    f = m.rhs()
    A = O.cantilever(n, n, n, dens=dens)
    xo, so = O.mgpcg(oracle_galerkin_levels(n, 3, dens), f, smoother=1, nu=2, tol=1e-12)
    s = HipMgPcgSolver(m, levels=3, smoother=smoother, omega=0.0, nu=2, tolerance=1e-10, precision="mixed",
                       format="duvi")
    s.Initialize(f)
    x, st = s.Solve()
    assert st.HasConverged and not st.Failed and st.OuterIterations >= 2
    assert st.RelativeResidual <= 1.01e-10
    assert np.linalg.norm(f - S(A) @ x) <= 1.02e-10 * np.linalg.norm(f)
    assert energy_err(A, x, xo) < 1e-8


def test_mixed_precision_stagnation_stops(m16):
    """A tolerance below the FP64 true-residual floor: the refinement stops after three steps
    that fail to halve the residual instead of running to the iteration cap."""
    m, A, U, b = m16
    s = HipMgPcgSolver(m, levels=3, smoother="gs", nu=1, tolerance=1e-30, max_iterations=500, precision="mixed")
    s.Initialize(b)
    x, st = s.Solve()
    assert not st.HasConverged and not st.Failed
    assert st.OuterIterations < 30
    assert st.RelativeResidual < 1e-13
    assert energy_err(A, x, U) < 1e-10


@pytest.mark.slow
def test_mixed_precision_64_heterogeneous(gpu):
    """64^3 with 1e4 modulus contrast: the precision form reaches 1e-9 on the FP64 true residual
    and agrees with a 1e-12 FP64 solve to 1e-8 in the energy norm."""
    n = 64
    dens = 10 ** (4 * O.rng_uniform(20261205, n ** 3) - 2)
    m = FemCantilever3D((n, n, n), ElementDensities=dens, rhs="tip")
    f = m.rhs()
    s = HipMgPcgSolver(m, levels=4, smoother="gs", nu=1, tolerance=1e-9, precision="mixed")
    s.Initialize(f)
    x, st = s.Solve()
    assert st.HasConverged and st.RelativeResidual <= 1.01e-9
    t = HipMgPcgSolver(m, levels=4, smoother="gs", nu=1, tolerance=1e-12)
    t.Initialize(f)
    xt, stt = t.Solve()
    A = O.cantilever(n, n, n, dens=dens)
    assert energy_err(A, x, xt) < 1e-8
