// HipCsrPcgSolver.cs -- drop-in for OpenCLCsrPcgSolver (src/MGroup.Solvers/Multigrid/OpenCLCsrPcgSolver.cs:12-331)
// on libmgx.so: Jacobi-preconditioned CG on the model's free-DOF stiffness matrix, stopping when every
// |r_i| <= tolerance (update_x_r_z_check_r, CsrCG.cl:257-277), the iteration count reported as the
// reference reports it (the 0-based index of the last iteration, :305-316).  The constructor,
// Initialize(model) (:84-218; the OpenCL Device argument is dropped), Solve (:250-331) and
// ReleaseOpenCLResources (:47-78) keep the reference's contract.
using System;
using MGroup.LinearAlgebra.Iterative;
using MGroup.LinearAlgebra.Vectors;
using MGroup.Solvers.LinearAlgebraExtensions.Matrices.Builders;

namespace Compression.src.MGroup.Solvers.Multigrid
{
    public class HipCsrPcgSolver : IDisposable
    {
        private IntPtr solver = IntPtr.Zero;
        private int DoFs;

        public int MaxIterations { get; set; }
        public double ConvergenceTolerance { get; set; }

        public HipCsrPcgSolver(int maxIterations, double convergenceTolerance)
        {
            MaxIterations = maxIterations;
            ConvergenceTolerance = convergenceTolerance;
        }

        public unsafe void Initialize(IGeometricMultigridModel model)
        {
            ReleaseOpenCLResources();
            var o = new Mgx.mgx_options();
            Mgx.mgx_options_default(ref o);
            o.method = Mgx.METHOD_PCG_JACOBI;
            o.format = Mgx.FORMAT_CSR;
            o.max_iterations = MaxIterations;
            o.tolerance = ConvergenceTolerance;
            o.stop = Mgx.STOP_MAXNORM;
            o.nrhs = 1;
            Mgx.Check(Mgx.mgx_solver_create(ref o, out solver));
            (DokRowMajor A, Vector b) = model.CreateLinearSystem();
            DoFs = A.NumRows;
            Mgx.SetLevel(solver, 0, A.BuildCsrArrays(true), DoFs, null, null);
            fixed (double* pb = b.RawData) Mgx.Check(Mgx.mgx_solver_set_rhs(solver, pb, DoFs, 1));
            Mgx.Check(Mgx.mgx_solver_initialize(solver));
        }

        public unsafe (Vector, IterativeStatistics) Solve(Vector? xInitialGuess)
        {
            if (DoFs == 0 || solver == IntPtr.Zero) throw new InvalidOperationException("You must call Initialize(model) first");
            bool given = xInitialGuess != null;
            Vector x = xInitialGuess ?? Vector.CreateZero(DoFs);
            Mgx.mgx_stats st;
            fixed (double* px = x.RawData) Mgx.Check(Mgx.mgx_solver_solve(solver, px, given ? 1 : 0, &st));
            return (x, new IterativeStatistics
            {
                NumIterationsRequired = st.iterations,
                ConvergenceCriterion = ("dumb text", ConvergenceTolerance),
                HasConverged = st.converged == 1,
            });
        }

        public void ReleaseOpenCLResources()
        {
            if (solver != IntPtr.Zero) { Mgx.mgx_solver_release(solver); solver = IntPtr.Zero; }
            DoFs = 0;
        }

        public void Dispose()
        {
            ReleaseOpenCLResources();
            GC.SuppressFinalize(this);
        }
    }
}
