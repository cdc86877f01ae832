// HipCsrGeometricMultigridSolver.cs -- drop-in for OpenCLCsrGeometricMultigridSolver
// (src/MGroup.Solvers/Multigrid/OpenCLCsrGeometricMultigridSolver.cs:14-536) on libmgx.so.
//
// Same constructor arguments, factories (CreateSimpleV / CreateDeepV, :55-83; the OpenCL Device and
// context arguments are dropped), Initialize(model, coarseRelaxation) (:145-285), Solve contract
// (:334-536: the solution is written into xInitialGuess when given, else into a new vector;
// non-convergence is HasConverged = false, not an exception) and ReleaseOpenCLResources
// (:106-138).  The levels are built exactly as the reference builds them (the model's DOK
// matrices, CreateCoarserModelAndSmoothenerMatrices, BuildCsrArrays(true)) and copied to the
// device by the library; the smoothing, residual check, restriction, coarse solve and
// prolongation run in libmgx.so.  Gauss-Seidel is the library's multicolour sweep (the
// reference's GPU kernel updates x in place with racy reads).
using System;
using System.Linq;
using MGroup.LinearAlgebra.Iterative;
using MGroup.LinearAlgebra.Vectors;
using MGroup.Solvers.LinearAlgebraExtensions.Matrices.Builders;

namespace Compression.src.MGroup.Solvers.Multigrid
{
    public class HipCsrGeometricMultigridSolver : IOpenCLGeometricMultigridSolver, IDisposable
    {
        private const int MaxCircleIterations = 10000;

        private bool[] LevelDown;
        private readonly int[] LevelIterations;
        private readonly bool GaussSeidel;
        private int firstLevel, totalLevels;
        private int[]? LevelDoFs;
        private IntPtr solver = IntPtr.Zero;

        public int MaxIterations { get; set; }
        public double ConvergenceTolerance { get; set; }

        /// <summary>Matrix layout the device kernels decode (CSR here, DuVi in the subclass).</summary>
        protected virtual int Format => Mgx.FORMAT_CSR;

        public static HipCsrGeometricMultigridSolver CreateSimpleV(IGeometricMultigridModel model, bool GaussSeidel = true,
            int maxCircleIterations = MaxCircleIterations, bool coarseRelaxation = true, double convergenceTolerance = 1e-6,
            int levelIterations = 4)
        {
            var a = new HipCsrGeometricMultigridSolver(maxCircleIterations, convergenceTolerance, new bool[] { true, false },
                                                       new int[] { levelIterations }, GaussSeidel);
            a.Initialize(model, coarseRelaxation);
            return a;
        }

        public static HipCsrGeometricMultigridSolver CreateDeepV(IGeometricMultigridModel model, bool GaussSeidel = true,
            int maxCircleIterations = MaxCircleIterations, bool coarseRelaxation = true, double convergenceTolerance = 1e-6,
            int depth = 2, int levelIterations = 4)
        {
            bool[] levelDown = Enumerable.Repeat(true, depth).Concat(Enumerable.Repeat(false, depth)).ToArray();
            var a = new HipCsrGeometricMultigridSolver(maxCircleIterations, convergenceTolerance, levelDown,
                                                       new int[] { levelIterations }, GaussSeidel);
            a.Initialize(model, coarseRelaxation);
            return a;
        }

        public HipCsrGeometricMultigridSolver(int maxCircleIterations, double convergenceTolerance, bool[] levelDown,
                                              int[] levelIterations, bool GaussSeidel = true)
        {
            this.GaussSeidel = GaussSeidel;
            LevelDown = levelDown;
            LevelIterations = levelIterations;
            MaxIterations = maxCircleIterations;
            ConvergenceTolerance = convergenceTolerance;
        }

        public unsafe void Initialize(IGeometricMultigridModel model, bool coarseRelaxation)
        {
            GeometricMultigridSolver.MakePath(ref totalLevels, ref LevelDown, ref firstLevel);
            int depth = totalLevels - 1;
            // the library runs the symmetric path [down x depth, up x depth] with one sweep count
            bool symmetric = LevelDown.Length == 2 * depth &&
                             LevelDown.Take(depth).All(d => d) && LevelDown.Skip(depth).All(d => !d);
            int sweeps = LevelIterations is { Length: > 0 } ? LevelIterations[0] : 4;
            if (!symmetric || (LevelIterations != null && LevelIterations.Any(i => i != sweeps)))
                throw new NotImplementedException("libmgx runs the V path [down x depth, up x depth] with one sweep count");

            ReleaseOpenCLResources();
            var o = new Mgx.mgx_options();
            Mgx.mgx_options_default(ref o);
            o.method = Mgx.METHOD_GMG;
            o.format = Format;
            o.smoother = GaussSeidel ? Mgx.SMOOTHER_GS : Mgx.SMOOTHER_JACOBI;
            o.coarse_relaxation = coarseRelaxation ? 1 : 0;
            o.omega = 0;                                   // the reference's weight rules
            o.max_iterations = MaxIterations;
            o.tolerance = ConvergenceTolerance;
            o.stop = Mgx.STOP_MAXNORM;                     // residual_with_check, CsrGeometricMultigrid.cl:121-141
            o.depth = depth;
            o.level_iterations = sweeps;
            o.nrhs = 1;
            Mgx.Check(Mgx.mgx_solver_create(ref o, out solver));

            LevelDoFs = new int[totalLevels];
            (DokRowMajor A, Vector b) = model.CreateLinearSystem();
            Vector rhs = b;
            for (int level = 0; level < totalLevels; ++level)
            {
                LevelDoFs[level] = A.NumRows;
                var a = A.BuildCsrArrays(true);
                if (level == totalLevels - 1)
                {
                    Mgx.SetLevel(solver, level, a, A.NumRows, null, null);
                    break;
                }
                (model, DokRowMajor R, DokRowMajor P) = IGeometricMultigridModel.CreateCoarserModelAndSmoothenerMatrices(model);
                var r = R.BuildCsrArrays(true);
                var p = P.BuildCsrArrays(true);
                Mgx.SetLevel(solver, level, a, A.NumRows, (r.values, r.colIndices, r.rowOffsets, R.NumRows, R.NumColumns),
                             (p.values, p.colIndices, p.rowOffsets, P.NumRows, P.NumColumns));
                (A, _) = model.CreateLinearSystem();
            }
            fixed (double* pb = rhs.RawData) Mgx.Check(Mgx.mgx_solver_set_rhs(solver, pb, rhs.Length, 1));
            Mgx.Check(Mgx.mgx_solver_initialize(solver));
        }

        public unsafe (Vector, IterativeStatistics, double[]) Solve(Vector? xInitialGuess)
        {
            if (LevelDoFs == null || solver == IntPtr.Zero)
                throw new InvalidOperationException("You must call Initialize(model) first");
            if (firstLevel > 0) throw new NotImplementedException("firstLevel != 0 is not implementent yet");
            bool given = xInitialGuess != null;
            Vector x = xInitialGuess ?? Vector.CreateZero(LevelDoFs[0]);
            Mgx.mgx_stats st;
            fixed (double* px = x.RawData) Mgx.Check(Mgx.mgx_solver_solve(solver, px, given ? 1 : 0, &st));
            // NaN / overflow (converged = -1) and the cycle cap (0) both report HasConverged = false
            var stats = new IterativeStatistics
            {
                NumIterationsRequired = st.iterations,
                ConvergenceCriterion = ("dumb text", ConvergenceTolerance),
                HasConverged = st.converged == 1,
            };
            return (x, stats, new double[totalLevels]);
        }

        public void ReleaseOpenCLResources()
        {
            if (solver != IntPtr.Zero) { Mgx.mgx_solver_release(solver); solver = IntPtr.Zero; }
        }

        public void Dispose()
        {
            ReleaseOpenCLResources();
            GC.SuppressFinalize(this);
        }
    }
}
