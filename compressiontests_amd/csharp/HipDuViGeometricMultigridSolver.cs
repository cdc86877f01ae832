// HipDuViGeometricMultigridSolver.cs -- drop-in for OpenCLDuViGeometricMultigridSolver
// (src/MGroup.Solvers/Multigrid/OpenCLDuViGeometricMultigridSolver.cs:33-751) on libmgx.so.
//
// The same solver as HipCsrGeometricMultigridSolver with every level matrix (A, R, P) encoded
// in the DuVi GPU block layout (FromDokRowMajor, :753-851, encoded by the library with the
// reference's value deduplication) and decoded by the device kernels: SpMV, residual,
// residual_with_check, Jacobi and multicolour Gauss-Seidel (DuViGeometricMultigrid.cl).
// More than 65,536 distinct values are stored as raw values (the intent of the reference's
// fallback, :814) instead of failing.
using System.Linq;

namespace Compression.src.MGroup.Solvers.Multigrid
{
    public class HipDuViGeometricMultigridSolver : HipCsrGeometricMultigridSolver
    {
        private const int MaxCircleIterations = 10000;

        protected override int Format => Mgx.FORMAT_DUVI;

        public new static HipDuViGeometricMultigridSolver CreateSimpleV(IGeometricMultigridModel model, bool GaussSeidel = true,
            int maxCircleIterations = MaxCircleIterations, bool coarseRelaxation = true, double convergenceTolerance = 1e-6,
            int levelIterations = 4)
        {
            var a = new HipDuViGeometricMultigridSolver(maxCircleIterations, convergenceTolerance, new bool[] { true, false },
                                                        new int[] { levelIterations }, GaussSeidel);
            a.Initialize(model, coarseRelaxation);
            return a;
        }

        public new static HipDuViGeometricMultigridSolver CreateDeepV(IGeometricMultigridModel model, bool GaussSeidel = true,
            int maxCircleIterations = MaxCircleIterations, bool coarseRelaxation = true, double convergenceTolerance = 1e-6,
            int depth = 2, int levelIterations = 4)
        {
            bool[] levelDown = Enumerable.Repeat(true, depth).Concat(Enumerable.Repeat(false, depth)).ToArray();
            var a = new HipDuViGeometricMultigridSolver(maxCircleIterations, convergenceTolerance, levelDown,
                                                        new int[] { levelIterations }, GaussSeidel);
            a.Initialize(model, coarseRelaxation);
            return a;
        }

        public HipDuViGeometricMultigridSolver(int maxCircleIterations, double convergenceTolerance, bool[] levelDown,
                                               int[] levelIterations, bool GaussSeidel = true)
            : base(maxCircleIterations, convergenceTolerance, levelDown, levelIterations, GaussSeidel)
        {
        }
    }
}
