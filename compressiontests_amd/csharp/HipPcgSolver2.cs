// HipPcgSolver2.cs -- drop-in for PcgSolver2 (src/MGroup.Solvers/Iterative/PcgSolver2.cs:20-155), the
// MSolve ISolver that runs Jacobi-preconditioned CG on a single-subdomain model, with the CG
// iterations in libmgx.so.
//
// Kept from the reference: the base class and its overrides (HandleMatrixWillBeSet, Initialize,
// PreventFromOverwrittingSystemMatrices, Solve, InverseSystemMatrixTimesOtherMatrix), the solution
// cleared to zero before every solve (:62-66), the matrix handed over again only after
// HandleMatrixWillBeSet (:40-43, the reference's preconditioner update, :68-77), the
// IterativeSolverNotConvergedException when a Solve does not converge (:84-89), one zero-start
// solve per column in InverseSystemMatrixTimesOtherMatrix without a convergence check (:92-136),
// and a nested Factory with DofOrderer, MatrixPatternWillNotBeModified, BuildSolver and
// BuildAlgebraicModel (:138-153).
//
// Changed: the algebraic model assembles a DokRowMajor (DokRowMajorMatrixAssembler, the
// assembler FemPlate.cs:59-61 uses) instead of the DuVi matrix, whose entries the library would
// have to decode again; its CSR arrays (BuildCsrArrays(true)) are copied to the device once per
// matrix.  The Factory's PcgAlgorithm and Preconditioner objects (MSolve CPU algorithms) become
// MaxIterations and ResidualTolerance of the device solver; the preconditioner is always Jacobi,
// the reference's default (:147).  The stop test is ||b - A x||_2 <= ResidualTolerance ||b||_2
// on the device's recurrence residual.  All columns of InverseSystemMatrixTimesOtherMatrix are
// solved in one library call.
using System;
using System.Diagnostics;
using Compression.src.MGroup.Solvers.Multigrid;
using MGroup.LinearAlgebra.Iterative;
using MGroup.LinearAlgebra.Matrices;
using MGroup.LinearAlgebra.Vectors;
using MGroup.MSolve.Discretization.Entities;
using MGroup.MSolve.Solution.LinearSystem;
using MGroup.Solvers.Assemblers;
using MGroup.Solvers.DofOrdering;
using MGroup.Solvers.DofOrdering.Reordering;
using MGroup.Solvers.LinearAlgebraExtensions.Matrices.Builders;
using MGroup.Solvers.LinearSystem;

namespace MGroup.Solvers.Iterative
{
    /// <summary>
    /// Iterative solver for models with only 1 subdomain: Jacobi-preconditioned CG on the GPU.
    /// </summary>
    public class HipPcgSolver2 : SingleSubdomainSolverBase<DokRowMajor>
    {
        private readonly int maxIterations;
        private readonly double residualTolerance;
        private readonly bool matrixPatternWillNotBeModified;   // kept for the Factory; the arrays are rebuilt either way

        private bool mustUpdateMatrix = true;
        private (double[] v, int[] c, int[] r) csr;
        private int systemSize;

        private HipPcgSolver2(GlobalAlgebraicModel<DokRowMajor> model, int maxIterations, double residualTolerance,
            bool matrixPatternWillNotBeModified)
            : base(model, "HipPcgSolver2")
        {
            this.maxIterations = maxIterations;
            this.residualTolerance = residualTolerance;
            this.matrixPatternWillNotBeModified = matrixPatternWillNotBeModified;
        }

        public override void HandleMatrixWillBeSet()
        {
            mustUpdateMatrix = true;
        }

        public override void Initialize() { }

        public override void PreventFromOverwrittingSystemMatrices()
        {
            // No factorization is done.
        }

        /// <summary>
        /// Solves the linear system with PCG on the GPU, from x = 0.  The matrix is copied to the device
        /// again only if it has been set since the last solve.
        /// </summary>
        public override unsafe void Solve()
        {
            var watch = new Stopwatch();
            IMatrix matrix = LinearSystem.Matrix;
            int n = matrix.NumRows;
            if (LinearSystem.Solution == null)
            {
                ((IGlobalLinearSystem) LinearSystem).Solution = Vector.CreateZero(n);
            }
            else LinearSystem.Solution.Clear();
            UpdateMatrix(matrix, watch);

            watch.Start();
            Vector b = AsVector(LinearSystem.RhsVector);
            Vector x = AsVector(LinearSystem.Solution);
            Mgx.mgx_stats st = SolveColumns(b.RawData, x.RawData, 1)[0];
            if (st.converged != 1)
            {
                throw new IterativeSolverNotConvergedException(Name + " did not converge to a solution. PCG algorithm run for"
                    + $" {st.iterations} iterations and the residual norm ratio was {st.rel_residual}");
            }
            watch.Stop();
            Logger.LogTaskDuration("Iterative algorithm", watch.ElapsedMilliseconds);
            Logger.LogIterativeAlgorithm(st.iterations, st.rel_residual);
            Logger.IncrementAnalysisStep();
        }

        /// <summary>
        /// A^-1 times every column of otherMatrix, each from x = 0, in one library call (the columns
        /// are solved one after another on the device).
        /// </summary>
        protected override Matrix InverseSystemMatrixTimesOtherMatrix(IReadOnlyMatrix otherMatrix)
        {
            var watch = new Stopwatch();
            IMatrix matrix = LinearSystem.Matrix;
            UpdateMatrix(matrix, watch);

            watch.Start();
            int n = systemSize, numRhs = otherMatrix.NumColumns;
            var solutionVectors = Matrix.CreateZero(n, numRhs);
            if (numRhs > 0)
            {
                var rhs = new double[(long)n * numRhs];
                var sol = new double[(long)n * numRhs];
                for (int j = 0; j < numRhs; ++j)
                    Array.Copy(otherMatrix.GetColumn(j).RawData, 0, rhs, (long)j * n, n);
                SolveColumns(rhs, sol, numRhs);
                var column = Vector.CreateZero(n);
                for (int j = 0; j < numRhs; ++j)
                {
                    Array.Copy(sol, (long)j * n, column.RawData, 0, n);
                    solutionVectors.SetSubcolumn(j, column);
                }
            }
            watch.Stop();
            Logger.LogTaskDuration("Iterative algorithm", watch.ElapsedMilliseconds);
            Logger.IncrementAnalysisStep();
            return solutionVectors;
        }

        // the CSR arrays of the assembled matrix, rebuilt after HandleMatrixWillBeSet
        private void UpdateMatrix(IMatrix matrix, Stopwatch watch)
        {
            if (!mustUpdateMatrix) return;
            watch.Start();
            var dok = (DokRowMajor) matrix;
            csr = dok.BuildCsrArrays(true);
            systemSize = dok.NumRows;
            watch.Stop();
            Logger.LogTaskDuration("Calculating preconditioner", watch.ElapsedMilliseconds);
            watch.Reset();
            mustUpdateMatrix = false;
        }

        // one solver handle per call: the matrix, the Jacobi weights and nrhs columns on the device
        private unsafe Mgx.mgx_stats[] SolveColumns(double[] rhs, double[] sol, int nrhs)
        {
            var o = new Mgx.mgx_options();
            Mgx.mgx_options_default(ref o);
            o.method = Mgx.METHOD_PCG_JACOBI;
            o.format = Mgx.FORMAT_CSR;
            o.max_iterations = maxIterations > 0 ? maxIterations : systemSize;
            o.tolerance = residualTolerance;
            o.stop = Mgx.STOP_REL_L2;
            o.nrhs = nrhs;
            Mgx.Check(Mgx.mgx_solver_create(ref o, out IntPtr solver));
            try
            {
                Mgx.SetLevel(solver, 0, csr, systemSize, null, null);
                fixed (double* pb = rhs) Mgx.Check(Mgx.mgx_solver_set_rhs(solver, pb, systemSize, nrhs));
                Mgx.Check(Mgx.mgx_solver_initialize(solver));
                var st = new Mgx.mgx_stats[nrhs];
                fixed (double* px = sol) fixed (Mgx.mgx_stats* ps = st)
                    Mgx.Check(Mgx.mgx_solver_solve(solver, px, 0, ps));
                return st;
            }
            finally
            {
                Mgx.mgx_solver_release(solver);
            }
        }

        // the single-subdomain vector behind a global one (FemPlate.cs:56 reads the rhs this way)
        private static Vector AsVector(object v) => v switch
        {
            Vector x => x,
            GlobalVector g => g.SingleVector,
            _ => throw new NotSupportedException("HipPcgSolver2 needs the single-subdomain vectors of a GlobalAlgebraicModel"),
        };

        public class Factory
        {
            public IDofOrderer DofOrderer { get; set; }
                = new DofOrderer(new NodeMajorDofOrderingStrategy(), new NullReordering());

            public bool MatrixPatternWillNotBeModified { get; set; } = false;

            /// <summary>CG iterations per solve; 0 = the system size.</summary>
            public int MaxIterations { get; set; } = 0;

            /// <summary>Stop when ||b - A x||_2 <= ResidualTolerance ||b||_2.</summary>
            public double ResidualTolerance { get; set; } = 1E-10;

            public HipPcgSolver2 BuildSolver(GlobalAlgebraicModel<DokRowMajor> model)
                => new HipPcgSolver2(model, MaxIterations, ResidualTolerance, MatrixPatternWillNotBeModified);

            public GlobalAlgebraicModel<DokRowMajor> BuildAlgebraicModel(IModel model)
                => new GlobalAlgebraicModel<DokRowMajor>(model, DofOrderer, new DokRowMajorMatrixAssembler(true));
        }
    }
}
