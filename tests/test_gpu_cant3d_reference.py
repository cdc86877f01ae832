"""The reference's own 3D cantilever test shapes (GMCantileverTests.cs:105-109: 256 x 16 x 16 and
512 x 16 x 32 elements on a 20 x 1 x 1 beam) on the GPU: the stationary V-cycle (CSR and DuVi,
multicolour Gauss-Seidel and Jacobi, depth 3) and Jacobi PCG on 256 x 16 x 16 against the oracle
(its OpenMP build, bit-identical to the serial one), and MG-PCG on 512 x 16 x 32 against the
manufactured solution."""
import numpy as np
import pytest
import scipy.sparse as sp

from compressiontests_amd.model import FemCantilever3D
from compressiontests_amd.solvers import (HipCsrGeometricMultigridSolver, HipCsrPcgSolver,
                                          HipDuViGeometricMultigridSolver, HipMgPcgSolver)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LEN = (20.0, 1.0, 1.0)
DIMS = (256, 16, 16)


def S(M):
    return sp.csr_matrix((M.val, M.col, M.rowptr), shape=(M.nrows, M.ncols))


def energy_err(A, x, xs):
    e = x - xs
    return np.sqrt(e @ (S(A) @ e)) / np.sqrt(xs @ (S(A) @ xs))


@pytest.fixture(scope="module")
def beam(gpu):
    m = FemCantilever3D(DIMS, LEN)
    A, U, b = m.CreateAndSolveLinearSystem()
    lv, d = [], DIMS
    for _ in range(3):
        R, P = O.rp_reference(*d)
        lv.append(dict(A=O.cantilever(*d), R=R, P=P))
        d = tuple(k >> 1 if k > 1 else 1 for k in d)
    lv.append(dict(A=O.cantilever(*d)))
    return m, A, U, b, lv, {}


@pytest.mark.parametrize("cls", [HipCsrGeometricMultigridSolver, HipDuViGeometricMultigridSolver])
@pytest.mark.parametrize("gs", [True, False])
def test_beam_gmg_parity(beam, cls, gs):
    """CheckCantileverSolutionDeepVWithOpenCL on 256 x 16 x 16, depth 3, 4 sweeps, tolerance 1e-5,
    coarse relaxation: the oracle's cycle count within one (the same Gauss-Seidel colouring)."""
    m, A, U, b, lv, cache = beam
    tol = 1e-5
    s = cls.CreateDeepV(m, GaussSeidel=gs, maxCircleIterations=20000, coarseRelaxation=True,
                        convergenceTolerance=tol, depth=3, levelIterations=4)
    x, st, _ = s.Solve(None)
    s.ReleaseOpenCLResources()
    if gs not in cache:
        pre = None
        if not gs:
            lam = O.eigen_upper_bound(lv[-1]["A"])
            pre = [O.jacobi_preconditioner(l["A"]) * (2 / lam) for l in lv[:-1]]
        cache[gs] = O.gmg_solve(lv, b, pre=pre, smoother=1 if gs else 0, level_iterations=4, max_cycles=20000,
                                tol=tol, omp=True)
    xo, so = cache[gs]
    print(f"beam 256x16x16 {cls.__name__} gs={gs}: cycles {st.NumIterationsRequired} (oracle {so.iterations})",
          flush=True)
    assert so.converged == 1 and st.HasConverged
    assert abs(st.NumIterationsRequired - so.iterations) <= 1
    assert np.abs(b - O.spmv(A, x)).max() <= tol
    assert energy_err(A, x, xo) < 1e-6


def test_beam_pcg_jacobi_parity(beam):
    """CheckCantileverSolutionCGWithOpenCL on 256 x 16 x 16."""
    m, A, U, b, lv, _ = beam
    s = HipCsrPcgSolver(20000, 1e-5)
    s.Initialize(m)
    x, st = s.Solve(None)
    s.ReleaseOpenCLResources()
    xo, so = O.pcg_jacobi(A, b, max_it=20000, tol=1e-5, omp=True)
    print(f"beam 256x16x16 PCG: iterations {st.NumIterationsRequired} (oracle {so.iterations})", flush=True)
    assert st.HasConverged and so.converged == 1
    assert abs(st.NumIterationsRequired - so.iterations) <= so.iterations // 20
    assert np.abs(b - O.spmv(A, x)).max() <= 1e-5


def test_beam_512_mgpcg_manufactured(gpu):
    """512 x 16 x 32 (GMCantileverTests.cs:107, 0.86 M dofs): MG-PCG with Gauss-Seidel over the
    Galerkin hierarchy (4 levels) reproduces the manufactured U_f to 1e-8 in the energy norm."""
    m = FemCantilever3D((512, 16, 32), LEN)
    b = m.rhs()
    U = m.CalcFreeDisplacements()
    s = HipMgPcgSolver(m, levels=4, smoother="gs", nu=2, tolerance=1e-10)
    s.Initialize(b)
    x, st = s.Solve()
    K = m.device_matrix().to_host()
    e = energy_err(K, x, U)
    print(f"beam 512x16x32 MG-PCG: iterations {st.NumIterationsRequired}, {st.SolveMs:.1f} ms, energy-norm error {e:.2e}",
          flush=True)
    assert st.HasConverged and e <= 1e-8
