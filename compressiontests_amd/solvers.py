"""Python mirror of the reference's solver interfaces over the C ABI of libmgx.

  IGeometricMultigridSolver.Solve(xInitialGuess) -> (x, IterativeStatistics, levelTimes)
      (src/MGroup.Solvers/Multigrid/IGeometricMultigridSolver.cs:6-9)
  HipCsrGeometricMultigridSolver   <- OpenCLCsrGeometricMultigridSolver (CreateSimpleV/CreateDeepV,
                                      Initialize, Solve, ReleaseOpenCLResources)
  HipDuViGeometricMultigridSolver  <- OpenCLDuViGeometricMultigridSolver
  HipCsrPcgSolver                  <- OpenCLCsrPcgSolver (OpenCLCsrPcgSolver.cs:40-331)
  HipMgPcgSolver                   new: PCG preconditioned by one symmetric V-cycle

Same argument names and meaning, same error behaviour: Solve before Initialize raises
MgxInvalidOperation (InvalidOperationException), a first level other than 0 raises
MgxNotImplemented, non-convergence is reported in the statistics, not raised.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .device import DeviceArray, DeviceCsr, galerkin_hierarchy, transfer_trilinear
from .model import CreateCoarserModelAndSmoothenerMatrices

MAX_CIRCLE_ITERATIONS = 10000


@dataclass
class IterativeStatistics:
    NumIterationsRequired: int = 0
    HasConverged: bool = False
    Failed: bool = False
    ConvergenceCriterion: tuple = ("", 0.0)
    ResidualNorm: float = 0.0
    RelativeResidual: float = 0.0
    SolveMs: float = 0.0
    SpmvMs: float = 0.0
    SpmvLaunches: int = 0
    OuterIterations: int = 0
    BlockWidth: int = 1
    GsMs: float = 0.0
    GsLaunches: int = 0

    @classmethod
    def from_c(cls, s, criterion_name, criterion_value):
        return cls(NumIterationsRequired=s.iterations, HasConverged=s.converged == 1, Failed=s.converged == -1,
                   ConvergenceCriterion=(criterion_name, criterion_value), ResidualNorm=s.residual_norm,
                   RelativeResidual=s.rel_residual, SolveMs=s.solve_ms, SpmvMs=s.spmv_ms,
                   SpmvLaunches=s.spmv_launches, OuterIterations=s.outer_iterations, BlockWidth=s.block_width,
                   GsMs=s.gs_ms, GsLaunches=s.gs_launches)


def make_path(level_down):
    """GeometricMultigridSolver.MakePath (GeometricMultigridSolver.cs:178-201)."""
    level_down = list(level_down)
    current, total, first = 0, 1, 0
    for d in level_down:
        if d:
            current += 1
            if current == total:
                total += 1
        elif current == 0:
            first += 1
            total += 1
        else:
            current -= 1
    if current > first:
        level_down += [False] * (current - first)
    elif current < first:
        level_down += [True] * (first - current)
    return total, level_down, first


class _Handle:
    """Owns one mgx_solver."""

    def __init__(self, opt):
        h = C.c_void_p()
        L.call("mgx_solver_create", C.byref(opt), C.byref(h))
        self.h = h.value
        self.opt = opt
        self.keep = []
        self.n = 0

    def set_level_device(self, level, A, R=None, P=None):
        self.keep += [A, R, P]
        L.call("mgx_solver_set_level_device", C.c_void_p(self.h), level, C.c_void_p(A.handle),
               C.c_void_p(R.handle) if R is not None else None, C.c_void_p(P.handle) if P is not None else None)
        if level == 0:
            self.n = A.nrows

    def set_level_host(self, level, A, R=None, P=None):
        def host(M):
            if M is None:
                return None, None
            arrs = (np.ascontiguousarray(M.rowptr, np.int32), np.ascontiguousarray(M.col, np.int32),
                    np.ascontiguousarray(M.val, np.float64))
            h = L.CsrHost(M.nrows, M.ncols, len(arrs[1]), *(a.ctypes.data for a in arrs))
            return h, arrs
        hA, kA = host(A)
        hR, kR = host(R)
        hP, kP = host(P)
        L.call("mgx_solver_set_level", C.c_void_p(self.h), level, C.byref(hA), C.byref(hR) if hR else None,
               C.byref(hP) if hP else None)
        if level == 0:
            self.n = A.nrows

    def set_rhs(self, b, nrhs=1):
        b = np.ascontiguousarray(b, np.float64).reshape(-1)
        n = b.size // nrhs
        L.call("mgx_solver_set_rhs", C.c_void_p(self.h), b.ctypes.data_as(C.c_void_p), n, nrhs)

    def initialize(self):
        L.call("mgx_solver_initialize", C.c_void_p(self.h))

    def solve(self, x0=None, nrhs=1):
        n = self.n
        x = np.zeros(n * nrhs) if x0 is None else np.array(x0, np.float64).reshape(-1).copy()
        st = (L.Stats * nrhs)()
        L.call("mgx_solver_solve", C.c_void_p(self.h), x.ctypes.data_as(C.c_void_p), int(x0 is not None), st)
        return x, list(st)

    def solve_device(self, b_dev, x_dev, has_x0=False, nrhs=1):
        st = (L.Stats * nrhs)()
        L.call("mgx_solver_solve_device", C.c_void_p(self.h), b_dev.vp, x_dev.vp, int(has_x0), st)
        return list(st)

    def block_width(self):
        w = C.c_int32()
        ms = (C.c_double * 4)()
        L.call("mgx_solver_block_width", C.c_void_p(self.h), C.byref(w), ms)
        return w.value, list(ms)

    def level_layout(self, level):
        v = C.c_int32()
        L.call("mgx_solver_level_layout", C.c_void_p(self.h), level, C.byref(v))
        return v.value

    def transfer_layout(self, level):
        v = C.c_int32()
        L.call("mgx_solver_transfer_layout", C.c_void_p(self.h), level, C.byref(v))
        return v.value

    def strip_window(self):
        b = C.c_double()
        ms = (C.c_double * 4)()
        L.call("mgx_solver_strip_window", C.c_void_p(self.h), C.byref(b), ms)
        return b.value, list(ms)

    def gs_order(self, level, nnodes):
        o = np.empty(nnodes, np.int32)
        L.call("mgx_solver_gs_order", C.c_void_p(self.h), level, o.ctypes.data_as(C.c_void_p))
        return o

    def level_groups(self, level):
        """The level's Gauss-Seidel row groups: None for row triples, else the group starts."""
        v = C.c_int32()
        L.call("mgx_solver_level_groups", C.c_void_p(self.h), level, C.byref(v), None)
        if v.value == 0:
            return None
        g = np.empty(v.value + 1, np.int32)
        L.call("mgx_solver_level_groups", C.c_void_p(self.h), level, C.byref(v), g.ctypes.data_as(C.c_void_p))
        return g

    def level_colors(self, level):
        v = C.c_int32()
        L.call("mgx_solver_level_colors", C.c_void_p(self.h), level, C.byref(v))
        return v.value

    def gs_pairs(self, level):
        v = C.c_int32()
        L.call("mgx_solver_gs_pairs", C.c_void_p(self.h), level, C.byref(v))
        return v.value

    def release(self):
        if self.h:
            L.call("mgx_solver_release", C.c_void_p(self.h))
            self.h = None
        self.keep = []

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class HipCsrGeometricMultigridSolver:
    """Stationary V-cycle multigrid (OpenCLCsrGeometricMultigridSolver.cs:14-557)."""

    FORMAT = L.FORMAT_CSR

    def __init__(self, maxCircleIterations, convergenceTolerance, levelDown, levelIterations, GaussSeidel=True,
                 checkEvery=None):
        """checkEvery (no reference counterpart): the host reads the device's stop state every
        checkEvery cycles (None: the library default, 4); the result is the same for every value."""
        self.MaxIterations = maxCircleIterations
        self.ConvergenceTolerance = convergenceTolerance
        self.LevelDown = list(levelDown)
        self.LevelIterations = list(levelIterations) if levelIterations is not None else [4]
        self.GaussSeidel = GaussSeidel
        self.CheckEvery = checkEvery
        self._h = None
        self.LevelDoFs = None

    @classmethod
    def CreateSimpleV(cls, model, GaussSeidel=True, maxCircleIterations=MAX_CIRCLE_ITERATIONS, coarseRelaxation=True,
                      convergenceTolerance=1e-6, levelIterations=4, checkEvery=None):
        a = cls(maxCircleIterations, convergenceTolerance, [True, False], [levelIterations], GaussSeidel, checkEvery)
        a.Initialize(model, coarseRelaxation)
        return a

    @classmethod
    def CreateDeepV(cls, model, GaussSeidel=True, maxCircleIterations=MAX_CIRCLE_ITERATIONS, coarseRelaxation=True,
                    convergenceTolerance=1e-6, depth=2, levelIterations=4, checkEvery=None):
        a = cls(maxCircleIterations, convergenceTolerance, [True] * depth + [False] * depth, [levelIterations],
                GaussSeidel, checkEvery)
        a.Initialize(model, coarseRelaxation)
        return a

    def Initialize(self, model, coarseRelaxation=True):
        total, level_down, first = make_path(self.LevelDown)
        self.LevelDown = level_down
        depth = total - 1
        if first != 0:
            raise L.MgxNotImplemented("firstLevel != 0 is not implementent yet")
        if level_down != [True] * depth + [False] * depth:
            raise L.MgxNotImplemented("only V paths [down x depth, up x depth] are supported")
        if len(set(self.LevelIterations)) != 1:
            raise L.MgxNotImplemented("one smoother iteration count for every level")
        opt = L.default_options(method=L.METHOD_GMG, format=self.FORMAT,
                                smoother=L.SMOOTHER_GS if self.GaussSeidel else L.SMOOTHER_JACOBI,
                                coarse_relaxation=int(bool(coarseRelaxation)), omega=0.0,
                                max_iterations=self.MaxIterations, tolerance=self.ConvergenceTolerance,
                                stop=L.STOP_MAXNORM, depth=depth, level_iterations=self.LevelIterations[0])
        if self.CheckEvery is not None:
            opt.check_every = int(self.CheckEvery)
        h = _Handle(opt)
        A = model.device_matrix()
        b = model.rhs()
        self.LevelDoFs = [A.nrows]
        m = model
        for level in range(depth):
            coarser, R, P = CreateCoarserModelAndSmoothenerMatrices(m)
            h.set_level_device(level, m.device_matrix(), R, P)
            self.LevelDoFs.append(R.nrows)
            m = coarser
        h.set_level_device(depth, m.device_matrix())
        h.set_rhs(b)
        h.initialize()
        self._h = h

    def Solve(self, xInitialGuess=None):
        if self._h is None:
            raise L.MgxInvalidOperation("You must call Initialize(model) first")
        x, st = self._h.solve(xInitialGuess)
        stats = IterativeStatistics.from_c(st[0], "dumb text", self.ConvergenceTolerance)
        return x, stats, np.zeros(len(self.LevelDoFs))

    def ReleaseOpenCLResources(self):
        if self._h is not None:
            self._h.release()
            self._h = None

    ReleaseResources = ReleaseOpenCLResources


class HipDuViGeometricMultigridSolver(HipCsrGeometricMultigridSolver):
    """Same V-cycle over the DuVi block layout (OpenCLDuViGeometricMultigridSolver.cs:14-861)."""

    FORMAT = L.FORMAT_DUVI


class HipCsrPcgSolver:
    """Jacobi-preconditioned CG (OpenCLCsrPcgSolver.cs:40-331): stop when every |r_i| <= tol,
    iteration count is the 0-based index of the last iteration, as the reference reports."""

    def __init__(self, maxIterations, convergenceTolerance):
        self.MaxIterations = maxIterations
        self.ConvergenceTolerance = convergenceTolerance
        self._h = None

    def Initialize(self, model):
        opt = L.default_options(method=L.METHOD_PCG_JACOBI, max_iterations=self.MaxIterations,
                                tolerance=self.ConvergenceTolerance, stop=L.STOP_MAXNORM)
        h = _Handle(opt)
        h.set_level_device(0, model.device_matrix())
        h.set_rhs(model.rhs())
        h.initialize()
        self._h = h

    def Solve(self, xInitialGuess=None):
        if self._h is None:
            raise L.MgxInvalidOperation("You must call Initialize(model) first")
        x, st = self._h.solve(xInitialGuess)
        return x, IterativeStatistics.from_c(st[0], "dumb text", self.ConvergenceTolerance)

    def ReleaseOpenCLResources(self):
        if self._h is not None:
            self._h.release()
            self._h = None

    ReleaseResources = ReleaseOpenCLResources


class HipMgPcgSolver:
    """PCG with one V-cycle as preconditioner (the new composition of OpenCLCsrPcgSolver and the
    V-cycle of OpenCLCsrGeometricMultigridSolver).  hierarchy:
      "galerkin":  K_{l+1} = P^T K_l P with exact trilinear P and R = P^T (symmetric, handles
                   heterogeneous moduli);
      "reference": the reference's rediscretised coarse models with R = P^T of the
                   reference interpolation (its cone-weight R is not symmetric to P).
    smoother "jacobi" (nu pre and nu post sweeps; omega > 0 fixed, omega <= 0 per level
    2 / (1.2 lambda_max(D^-1 K)) by the reference's power method) or "gs" (nu forward / nu
    backward multicolour Gauss-Seidel sweeps).

    nrhs > 1: b is (nrhs, n), one load case per row, and Solve returns x of the same shape and
    a list of statistics.  block_width: right-hand sides advanced together by one batched PCG
    (4, 8, 16, 32), 1 = one at a time, 0 = chosen at Initialize by timing each width.
    strip_window: X window (bytes) of the strip order of the block kernels on structured levels,
    0 = whole planes, < 0 = chosen at Initialize by timing the candidates.
    gs_pairs: block Gauss-Seidel relaxes two colours per launch, line by line, where the colouring
    allows it (the same sweep); False = one launch per colour.
    precision "mixed": FP32 SpMV / V-cycle inside FP64 iterative refinement.
    duvi_mode (format "duvi"): 0 = the narrowest value index, 4 = 32-bit indices, 8 = raw values."""

    def __init__(self, model, levels=3, smoother="jacobi", omega=0.0, nu=2, tolerance=1e-8, stop="rel",
                 max_iterations=1000, hierarchy="galerkin", format="csr", profile=False, check_every=4, nrhs=1,
                 use_graph=True, precision="fp64", inner_tolerance=1e-4, inner_iterations=200, block_width=0,
                 strip_window=-1.0, gs_pairs=True, duvi_mode=0):
        self.model = model
        self.levels = levels
        opt = L.default_options(method=L.METHOD_MGPCG,
                                format=L.FORMAT_DUVI if format == "duvi" else L.FORMAT_CSR,
                                smoother=L.SMOOTHER_GS if smoother == "gs" else L.SMOOTHER_JACOBI,
                                omega=omega, max_iterations=max_iterations, tolerance=tolerance,
                                stop=L.STOP_REL_L2 if stop == "rel" else L.STOP_MAXNORM, level_iterations=nu,
                                profile=int(profile), check_every=check_every, nrhs=nrhs,
                                use_graph=int(bool(use_graph)),
                                precision=L.PRECISION_MIXED if precision == "mixed" else L.PRECISION_FP64,
                                inner_tolerance=inner_tolerance, inner_iterations=inner_iterations,
                                block_width=block_width, strip_window=float(strip_window),
                                gs_pairs=int(bool(gs_pairs)))
        opt.duvi_mode = duvi_mode
        self.opt = opt
        self.nrhs = nrhs
        h = _Handle(opt)
        self.hierarchy = []             # (A, R, P) device matrices per level, finest first
        if isinstance(hierarchy, (list, tuple)):
            self.hierarchy = list(hierarchy)      # shared with another solver of the same model
        elif hierarchy == "galerkin":
            if len(model.NumElements) != 3:
                raise L.MgxNotImplemented("the Galerkin hierarchy is built for the 3D cantilever; "
                                          "use hierarchy='reference' for the plate")
            nx, ny, nz = model.NumElements
            Ks = galerkin_hierarchy(nx, ny, nz, levels, model.ElasticityModulus, model.PoissonRatio,
                                    model.ElementDensities)
            model._K = Ks[0]
            dims = (nx, ny, nz)
            for l in range(levels - 1):
                R, P = transfer_trilinear(*dims)
                self.hierarchy.append((Ks[l], R, P))
                dims = tuple(d // 2 for d in dims)
            self.hierarchy.append((Ks[-1], None, None))
        elif hierarchy == "reference":
            m = model
            for l in range(levels - 1):
                coarser, _, P = CreateCoarserModelAndSmoothenerMatrices(m)
                Ph = P.to_host()
                import scipy.sparse as sp   # host transpose of the reference P (setup only)
                PT = sp.csr_matrix((Ph.val, Ph.col, Ph.rowptr), shape=(Ph.nrows, Ph.ncols)).T.tocsr()
                PT.sort_indices()
                from .device import CsrMatrix
                R = DeviceCsr.from_host(CsrMatrix(PT.indptr.astype(np.int32), PT.indices.astype(np.int32),
                                                  PT.data, PT.shape[0], PT.shape[1]))
                self.hierarchy.append((m.device_matrix(), R, P))
                m = coarser
            self.hierarchy.append((m.device_matrix(), None, None))
        else:
            raise ValueError("hierarchy must be 'galerkin', 'reference' or a list of (A, R, P)")
        if len(self.hierarchy) != levels:
            raise ValueError("hierarchy depth differs from levels")
        for l, (A, R, P) in enumerate(self.hierarchy):
            h.set_level_device(l, A, R, P)
        self._h = h

    def Initialize(self, b=None):
        if b is None:
            b = self.model.rhs()
        self._h.set_rhs(b, self.nrhs)
        self._h.initialize()

    def Solve(self, xInitialGuess=None):
        x, st = self._h.solve(xInitialGuess, self.nrhs)
        stats = [IterativeStatistics.from_c(s, "relative residual", self.opt.tolerance) for s in st]
        if self.nrhs == 1:
            return x, stats[0]
        return x.reshape(self.nrhs, -1), stats

    @property
    def BlockWidth(self):
        """(block width in use, ms per block PCG iteration timed for K = 4, 8, 16, 32)"""
        return self._h.block_width()

    def LevelLayout(self, level=0):
        """Layout the block kernels read on a level: "csr", "node-triple", "node-block" or "duvi"."""
        return {0: "csr", 1: "node-triple", 2: "node-block", 4: "duvi"}[self._h.level_layout(level)]

    def TransferLayout(self, level=0):
        """Layout of the level's R and P in the block V-cycle: "csr", "node-transfer" or "duvi"."""
        return {0: "csr", 3: "node-transfer", 4: "duvi"}[self._h.transfer_layout(level)]

    @property
    def StripWindow(self):
        """(X window in bytes of the block kernels' strip order, 0 = whole planes; ms per block PCG
        iteration the autotuner timed for the windows 6e6, 3e6, 1.5e6 and 0)"""
        return self._h.strip_window()

    def GsOrder(self, level=0):
        """Node order of the level's forward Gauss-Seidel sweep as the solver runs it (row-group
        numbers where the level's nodes are not all row triples: see GsGroups)."""
        g = self._h.level_groups(level)
        A = self.hierarchy[level][0]
        return self._h.gs_order(level, A.nrows // 3 if g is None else len(g) - 1)

    def GsGroups(self, level=0):
        """None when the level's Gauss-Seidel nodes are the row triples, else the starts of its
        row groups (runs of up to 3 consecutive rows with one column list)."""
        return self._h.level_groups(level)

    def LevelColors(self, level=0):
        """Colours of the level's multicolour Gauss-Seidel (0 without that smoother)."""
        return self._h.level_colors(level)

    def GsPairs(self, level=0):
        """Colour pairs of the level relaxed two colours per launch, line by line (0: one launch
        per colour)."""
        return self._h.gs_pairs(level)

    def solve_device(self, b_dev, x_dev):
        return self._h.solve_device(b_dev, x_dev, False, self.nrhs)

    def ReleaseResources(self):
        self._h.release()
