// Mgx.cs -- P/Invoke binding of libmgx.so (include/mgx.h) for the reference's C# solvers.
//
// The structs mirror the C structs field by field (LayoutKind.Sequential gives the same natural
// alignment as the C compiler: a double after an int32 starts at the next 8-byte boundary on
// both sides); tests/test_csharp_binding.py checks names, order and types against include/mgx.h.
// Status codes are mapped back to the exceptions the reference throws.
using System;
using System.Runtime.InteropServices;

namespace Compression.src.MGroup.Solvers.Multigrid
{
    internal static class Mgx
    {
        private const string Lib = "mgx";   // libmgx.so next to the application binaries

        public const int OK = 0;
        public const int ERR_INVALID_ARGUMENT = 1;
        public const int ERR_INVALID_OPERATION = 2;
        public const int ERR_NOT_IMPLEMENTED = 3;
        public const int ERR_DEVICE = 4;
        public const int ERR_OUT_OF_MEMORY = 5;
        public const int ERR_NOT_SPD = 6;

        public const int METHOD_GMG = 0, METHOD_PCG_JACOBI = 1, METHOD_MGPCG = 2;
        public const int FORMAT_CSR = 0, FORMAT_DUVI = 1;
        public const int SMOOTHER_JACOBI = 0, SMOOTHER_GS = 1;
        public const int STOP_MAXNORM = 0, STOP_REL_L2 = 1;
        public const int PRECISION_FP64 = 0, PRECISION_MIXED = 1;

        [StructLayout(LayoutKind.Sequential)]
        public struct mgx_options
        {
            public int method;
            public int format;
            public int smoother;
            public int coarse_relaxation;
            public double omega;
            public int max_iterations;
            public double tolerance;
            public int stop;
            public int depth;
            public int level_iterations;
            public int precision;
            public int nrhs;
            public double duvi_tolerance;
            public int duvi_mode;
            public int inner_iterations;
            public double inner_tolerance;
            public int profile;
            public int check_every;
            public int use_graph;
            public int block_width;
            public double strip_window;
            public int gs_pairs;
        }

        [StructLayout(LayoutKind.Sequential)]
        public struct mgx_stats
        {
            public int iterations;
            public int converged;
            public double residual_norm;
            public double rel_residual;
            public double solve_ms;
            public int outer_iterations;
            public int spmv_launches;
            public double spmv_ms;
            public int block_width;
            public int gs_launches;
            public double gs_ms;
        }

        [StructLayout(LayoutKind.Sequential)]
        public unsafe struct mgx_csr_host
        {
            public int nrows;
            public int ncols;
            public long nnz;
            public int* rowptr;
            public int* col;
            public double* val;
        }

        [DllImport(Lib)] public static extern IntPtr mgx_last_error();
        [DllImport(Lib)] public static extern int mgx_version();
        [DllImport(Lib)] public static extern int mgx_device_count();
        [DllImport(Lib)] public static extern void mgx_options_default(ref mgx_options opt);
        [DllImport(Lib)] public static extern int mgx_solver_create(ref mgx_options opt, out IntPtr solver);
        [DllImport(Lib)] public static extern unsafe int mgx_solver_set_level(IntPtr solver, int level, mgx_csr_host* A,
                                                                              mgx_csr_host* R, mgx_csr_host* P);
        [DllImport(Lib)] public static extern unsafe int mgx_solver_set_rhs(IntPtr solver, double* b, int n, int nrhs);
        [DllImport(Lib)] public static extern int mgx_solver_initialize(IntPtr solver);
        [DllImport(Lib)] public static extern unsafe int mgx_solver_solve(IntPtr solver, double* x, int has_x0,
                                                                          mgx_stats* stats);
        [DllImport(Lib)] public static extern int mgx_solver_release(IntPtr solver);

        /// <summary>Throws the exception the reference throws for a failed status.</summary>
        public static void Check(int rc)
        {
            if (rc == OK) return;
            string msg = Marshal.PtrToStringAnsi(mgx_last_error()) ?? "";
            throw rc switch
            {
                ERR_INVALID_OPERATION => new InvalidOperationException(msg),
                ERR_NOT_IMPLEMENTED => new NotImplementedException(msg),
                ERR_INVALID_ARGUMENT or ERR_NOT_SPD => new ArgumentException(msg),
                ERR_OUT_OF_MEMORY => new OutOfMemoryException(msg),
                _ => new InvalidOperationException("libmgx device failure: " + msg),
            };
        }

        /// <summary>Hands one level (A, and R / P towards the next coarser level) to the solver; the
        /// library copies the arrays, as the reference's blocking WriteBuffer does.</summary>
        public static unsafe void SetLevel(IntPtr solver, int level, (double[] v, int[] c, int[] r) A, int n,
                                           (double[] v, int[] c, int[] r, int rows, int cols)? R,
                                           (double[] v, int[] c, int[] r, int rows, int cols)? P)
        {
            fixed (double* av = A.v) fixed (int* ac = A.c) fixed (int* ar = A.r)
            {
                var a = new mgx_csr_host { nrows = n, ncols = n, nnz = A.v.Length, rowptr = ar, col = ac, val = av };
                if (R == null || P == null)
                {
                    Check(mgx_solver_set_level(solver, level, &a, null, null));
                    return;
                }
                var (rv, rc, rr, rrows, rcols) = R.Value;
                var (pv, pc, pr, prows, pcols) = P.Value;
                fixed (double* rvp = rv) fixed (int* rcp = rc) fixed (int* rrp = rr)
                fixed (double* pvp = pv) fixed (int* pcp = pc) fixed (int* prp = pr)
                {
                    var hr = new mgx_csr_host { nrows = rrows, ncols = rcols, nnz = rv.Length, rowptr = rrp, col = rcp, val = rvp };
                    var hp = new mgx_csr_host { nrows = prows, ncols = pcols, nnz = pv.Length, rowptr = prp, col = pcp, val = pvp };
                    Check(mgx_solver_set_level(solver, level, &a, &hr, &hp));
                }
            }
        }
    }
}
