/*
 * mgx.h -- C ABI of the MI355X (gfx950) MG-PCG / SpMV solver backend (libmgx.so).
 *
 * Plain C: pointers, sizes, status codes.  No torch or HIP types in any signature;
 * a stream is passed as `void*` (a hipStream_t, NULL = the library's own stream).
 * Every function returns MGX_OK (0) or an error status; mgx_last_error() returns the
 * message of the calling thread's last failure.
 *
 * Two groups of entry points:
 *   1. Device-level kernels (device pointers in and out).  One per OpenCL kernel of the
 *      reference, same argument meaning; the reference kernel each replaces is cited.
 *   2. The solver handle, replacing the reference's OpenCL solver classes
 *      (Initialize / Solve / ReleaseOpenCLResources) with host arrays in and out.
 *
 * Reference paths are relative to nickreserved/CompressionTests.
 */
#ifndef MGX_H
#define MGX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status codes
 * The reference throws; a C# binding maps these back to its exceptions:
 *   MGX_ERR_INVALID_OPERATION  <- InvalidOperationException ("You must call Initialize(model) first",
 *                                 OpenCLCsrGeometricMultigridSolver.cs:336, OpenCLCsrPcgSolver.cs:252)
 *   MGX_ERR_NOT_IMPLEMENTED    <- NotImplementedException (firstLevel != 0,
 *                                 OpenCLCsrGeometricMultigridSolver.cs:345; varint range,
 *                                 DuViCompressedSparseMatrix.cs:22)
 *   MGX_ERR_DEVICE             <- OpenCLException (OpenCLDuViGeometricMultigridSolver.cs:859)
 *   MGX_ERR_INVALID_ARGUMENT   <- ArgumentException / Debug.Assert on sizes
 * Non-convergence is not an error: it is reported in mgx_stats.converged. */
#define MGX_OK                     0
#define MGX_ERR_INVALID_ARGUMENT   1
#define MGX_ERR_INVALID_OPERATION  2
#define MGX_ERR_NOT_IMPLEMENTED    3
#define MGX_ERR_DEVICE             4
#define MGX_ERR_OUT_OF_MEMORY      5
#define MGX_ERR_NOT_SPD            6

const char *mgx_last_error(void);
/* library version, e.g. 0x000100 */
int32_t mgx_version(void);
/* number of visible devices (0 on a host without a GPU; never an error) */
int32_t mgx_device_count(void);

/* ================================================================== 1. device-level kernels */

/* y = A x (clear = 1) or y += A x (clear = 0); CSR with int32 row offsets / columns.
 * Replaces matrix_vector_product, src/MGroup.OCL/CsrGeometricMultigrid.cl:49-62 (= CsrCG.cl:137-150). */
int mgx_csr_spmv(int32_t nrows, const int32_t *rowptr, const int32_t *col, const double *val,
                 const double *x, double *y, int32_t clear, void *stream);

/* FP32 values and vectors, int32 indices (precision form; no reference counterpart). */
int mgx_csr_spmv_f32(int32_t nrows, const int32_t *rowptr, const int32_t *col, const float *val,
                     const float *x, float *y, int32_t clear, void *stream);

/* Y = A X (clear = 1) or Y += A X (clear = 0) for k right-hand sides at once: X (ncols x k) and
 * Y (nrows x k) are row-major blocks (element (i, c) at i*k + c), k in {4, 8, 16, 32}; the
 * matrix is read once for all k columns.  Batching form (no reference counterpart; the
 * reference solves one right-hand side at a time, PcgSolver2.cs:92-136). */
int mgx_csr_spmm(int32_t nrows, const int32_t *rowptr, const int32_t *col, const double *val, int32_t k,
                 const double *X, double *Y, int32_t clear, void *stream);
int mgx_csr_spmm_f32(int32_t nrows, const int32_t *rowptr, const int32_t *col, const float *val, int32_t k,
                     const float *X, float *Y, int32_t clear, void *stream);

/* r = b - A x.  Replaces residual, CsrGeometricMultigrid.cl:96-110. */
int mgx_csr_residual(int32_t nrows, const int32_t *rowptr, const int32_t *col, const double *val,
                     const double *b, const double *x, double *r, void *stream);

/* r = b - A x and the convergence flag of residual_with_check, CsrGeometricMultigrid.cl:121-141:
 * *flag must hold 1 on entry; bit 1 is set if any r_i is NaN or |r_i| > 1e50, bit 0 is cleared
 * if any other |r_i| > tol.  flag is a device pointer. */
int mgx_csr_residual_check(int32_t nrows, const int32_t *rowptr, const int32_t *col, const double *val,
                           const double *b, const double *x, double *r, double tol,
                           int32_t *flag, void *stream);

/* x = pre .* b.  Replaces jacobi_initial_iteration, CsrGeometricMultigrid.cl:153-164. */
int mgx_jacobi_initial(int32_t n, const double *pre, const double *b, double *x, void *stream);

/* y = x + pre .* (b - A x).  Replaces jacobi_iteration, CsrGeometricMultigrid.cl:176-194. */
int mgx_csr_jacobi(int32_t n, const double *pre, const int32_t *rowptr, const int32_t *col,
                   const double *val, const double *b, const double *x, double *y, void *stream);

/* One colour of a multicolour Gauss-Seidel sweep over node blocks of 3 rows: for each node
 * k in nodes[0..nnodes), rows 3k..3k+2 are relaxed in order (reverse order when
 * backward != 0) using the current x.  Nodes of one colour must not be coupled.
 * Replaces gauss_seidel_iteration, CsrGeometricMultigrid.cl:207-225, whose in-place update
 * races; the multicolour order is the documented deviation. */
int mgx_csr_gs_color(int32_t nnodes, const int32_t *nodes, const int32_t *rowptr, const int32_t *col,
                     const double *val, const double *b, double *x, int32_t backward, void *stream);

/* *result = a . b (device pointer result).  Replaces dot_partial + dot_finalize, CsrCG.cl:43-85. */
int mgx_dot(int32_t n, const double *a, const double *b, double *result, void *stream);

/* *result = a . b, the same kernels and the same result bits as mgx_dot, but the call allocates
 * nothing: no hipMalloc / hipFree, no allocating library routine.  No counterpart in the reference;
 * it is for callers that cannot allocate device memory (code under stream capture, tight serving
 * loops, hosts that forbid allocation).  The per-block partial sums live in a 16 KiB array of
 * device memory inside the library (2048 slots; dot sizes needing more partials are rejected with
 * MGX_ERR_INVALID_ARGUMENT -- the grid is capped at 2048 blocks, so every int32 n fits), one copy
 * per GPU, held until the process ends.  Asynchronous on `stream` (mgx_dot synchronizes): *result
 * is ready when the stream reaches this point.  Every call on one GPU shares that GPU's array, so
 * calls that overlap on one GPU (two streams, two host threads) race and must be serialized by
 * the caller. */
int mgx_dot_noalloc(int32_t n, const double *a, const double *b, double *result, void *stream);

/* DuVi GPU block layout (OpenCLDuViGeometricMultigridSolver.cs:16-31) on the device.
 * vi_bytes: 1 (u8), 2 (u16), 4 (u32) value indices into values[nuniq], or 0 = "DU + raw
 * values" (values[nnz] in CSR order, no index stream). */
typedef struct {
    int32_t nrows;
    int32_t vi_bytes;
    int64_t nnz, nblocks, nuniq;
    const int32_t *row_blocks;   /* [nrows+1] RowToColumnIndices */
    const int32_t *row_dist;     /* [nrows+1] RowToDistances     */
    const int32_t *block_col;    /* [nblocks] ColumnIndices       */
    const uint8_t *dist;         /* [nnz]     Distances           */
    const void    *vi;           /* [nnz]     ValueIndices or NULL */
    const double  *values;
} mgx_duvi;

/* y = A x / y += A x decoding the DuVi block layout.
 * Replaces matrix_vector_product, src/MGroup.OCL/DuViGeometricMultigrid.cl:122-159. */
int mgx_duvi_spmv(const mgx_duvi *A, const double *x, double *y, int32_t clear, void *stream);

/* r = b - A x on the DuVi layout, one fused pass.  Replaces residual,
 * DuViGeometricMultigrid.cl:220-254. */
int mgx_duvi_residual(const mgx_duvi *A, const double *b, const double *x, double *r, void *stream);

/* r = b - A x and the convergence flag of residual_with_check (DuViGeometricMultigrid.cl:275-317),
 * the same protocol as mgx_csr_residual_check: *flag (device) holds 1 on entry; bit 1 is set if
 * any r_i is NaN or |r_i| > 1e50, bit 0 is cleared if any other |r_i| > tol. */
int mgx_duvi_residual_check(const mgx_duvi *A, const double *b, const double *x, double *r, double tol,
                            int32_t *flag, void *stream);

/* y = x + pre .* (b - A x) on the DuVi layout, one fused pass.  Replaces jacobi_iteration,
 * DuViGeometricMultigrid.cl:362-402. */
int mgx_duvi_jacobi(const mgx_duvi *A, const double *pre, const double *b, const double *x, double *y,
                    void *stream);

/* One colour of a multicolour Gauss-Seidel sweep on the DuVi layout, rows in node triples: for
 * each node k of nodes[0..nnodes), rows 3k..3k+2 are relaxed in order (reverse order when
 * backward != 0) with the current x; nodes of one colour must not be coupled.  Replaces
 * gauss_seidel_iteration, DuViGeometricMultigrid.cl:422-462, whose in-place update races (and
 * whose 8/16-bit value-index types are swapped, :428-432); the multicolour order is the
 * documented deviation. */
int mgx_duvi_gs_color(const mgx_duvi *A, int32_t nnodes, const int32_t *nodes, const double *b, double *x,
                      int32_t backward, void *stream);

/* y = A x / y += A x with FP32 values and vectors: values32 is an FP32 copy (device) of A's value
 * table (nuniq entries, or nnz raw values), which replaces A->values; the index streams are
 * shared.  Precision form (no reference counterpart). */
int mgx_duvi_spmv_f32(const mgx_duvi *A, const float *values32, const float *x, float *y, int32_t clear,
                      void *stream);

/* Host-side DuVi encoder: CSR -> block layout.  Restates FromDokRowMajor,
 * OpenCLDuViGeometricMultigridSolver.cs:753-851 (tolerance dedup of ToleranceComparer,
 * DuViCompressedSparseMatrix.cs:168-174).  mode 0: u8 if <= 256 unique values, u16 if
 * <= 65536, else DU + raw values (the intent of the reference's defective 16-bit switch
 * :818-825 and its >65536 fallback :814); mode 4: u32 indices; mode 8: raw values.
 * Output arrays are allocated by the library (host memory) and freed by mgx_duvi_host_free. */
typedef struct {
    int32_t nrows, vi_bytes;
    int64_t nnz, nblocks, nuniq;
    int32_t *row_blocks, *row_dist, *block_col;
    uint8_t *dist;
    void *vi;
    double *values;
} mgx_duvi_host;
int mgx_duvi_encode(int32_t nrows, const int32_t *rowptr, const int32_t *col, const double *val,
                    double tol, int32_t mode, mgx_duvi_host *out);
void mgx_duvi_host_free(mgx_duvi_host *m);

/* Node-block (3x3 block CSR) copy of a matrix whose rows come in node triples (3q, 3q+1, 3q+2)
 * sharing one column list made of whole nodes (3j, 3j+1, 3j+2) -- every level matrix of the
 * 3-dof hexahedral hierarchy.  This is the layout the batching form's SpMM and Gauss-Seidel read
 * (one column index per 9 values).  Built on the device from a device CSR with FP64 values;
 * fp32 = 1 stores FP32 blocks.  MGX_ERR_INVALID_ARGUMENT if the matrix is not of that form or a
 * node has more than 27 neighbour nodes.  No reference counterpart (batching form). */
typedef struct mgx_nodeblock_s *mgx_nodeblock;
int mgx_nodeblock_from_csr(int32_t nrows, const int32_t *rowptr, const int32_t *col, const double *val,
                           int32_t fp32, mgx_nodeblock *out);
/* Y = A X (clear = 1) or Y += A X (clear = 0); X, Y: n x k row-major device blocks of the copy's
 * type (float for an FP32 copy, double otherwise), k in {4, 8, 16, 32}. */
int mgx_nodeblock_spmm(mgx_nodeblock A, int32_t k, const void *X, void *Y, int32_t clear, void *stream);
int mgx_nodeblock_destroy(mgx_nodeblock A);

/* Node transfer: a prolongation or restriction between two levels whose dofs come in node triples,
 * acting on the three dofs of a node alike (entry (3i+a, 3j+b) = w_ij when a == b, else 0), stored
 * as one weight per (output node, input node) pair.  It replaces the CSR R / P products of the
 * V-cycle in the batching form (the reference applies R and P as CSR SpMVs:
 * CsrGeometricMultigrid.cl, driven from OpenCLCsrGeometricMultigridSolver.cs:334-536).  Built
 * on the device from a device CSR with FP64 values (nrows x ncols, both multiples of 3); fp32 = 1
 * stores FP32 weights.  MGX_ERR_INVALID_ARGUMENT if the matrix is not of that form. */
typedef struct mgx_nodetransfer_s *mgx_nodetransfer;
int mgx_nodetransfer_from_csr(int32_t nrows, int32_t ncols, const int32_t *rowptr, const int32_t *col,
                              const double *val, int32_t fp32, mgx_nodetransfer *out);
/* Y = M X (add = 0) or Y += M X (add = 1); X: ncols x k and Y: nrows x k row-major device blocks of
 * the copy's type, k in {4, 8, 16, 32}. */
int mgx_nodetransfer_apply(mgx_nodetransfer M, int32_t k, const void *X, void *Y, int32_t add, void *stream);
int mgx_nodetransfer_destroy(mgx_nodetransfer M);

/* SIMP compliance sensitivity, one thread per element (no reference counterpart; the hook
 * is ElementDensities x Ke, FemCantileverBase.cs:99-100):
 *   grad_e = -penal * rho_e^(penal-1) * (E0 - Emin) * u_e^T Ke0 u_e
 * u is the free-dof displacement (device), rho and grad per element (device), on the
 * structured cantilever mesh nx x ny x nz with x = 0 clamped. */
int mgx_simp_gradient(int32_t nx, int32_t ny, int32_t nz, double nu, const double *rho, double penal,
                      double Emin, double E0, const double *u, double *grad, void *stream);

/* ================================================================== structured generator
 * The reference builds K_ff, R and P on the host from DokRowMajor dictionaries
 * (FemCantileverBase.cs:87-104, IGeometricMultigridModel.cs:76-194), which cannot reach
 * 128^3.  These build them on the device, in the same sorted-CSR layout and the same
 * element-order summation (bit-identical K_ff).  Matrices are owned by the library. */
typedef struct mgx_matrix_s *mgx_matrix;

/* device view of a matrix owned by the library */
typedef struct {
    int32_t nrows, ncols;
    int64_t nnz;
    int32_t *rowptr, *col;
    double *val;
} mgx_csr_view;

/* K_ff of the x = 0 clamped cantilever on nx*ny*nz unit-cube-Ke hexahedra
 * (FemCantilever3D.cs:32-120); dens (host, nx*ny*nz, may be NULL = 1) scales each element. */
int mgx_matrix_cantilever(int32_t nx, int32_t ny, int32_t nz, double E, double nu, const double *dens,
                          mgx_matrix *out);
/* Reference restriction / interpolation pair of the level nx*ny*nz -> coarser
 * (IGeometricMultigridModel.cs:76-194, cone weights, rows normalised, free-dof submatrix). */
int mgx_matrix_transfer_reference(int32_t nx, int32_t ny, int32_t nz, mgx_matrix *R, mgx_matrix *P);
/* Exact trilinear interpolation P (fine free x coarse free) and R = P^T. */
int mgx_matrix_transfer_trilinear(int32_t nx, int32_t ny, int32_t nz, mgx_matrix *R, mgx_matrix *P);

/* The Kirchhoff plate workload (FemPlate.cs): nx*ny rectangular plate elements (Adini-Clough-Melosh,
 * PlateElementRectangle4Nodes3Dofs.cs:134-222) on Lx*Ly, thickness t; dofs (w, rx, ry) per node, node
 * by node, w fixed on the lines x = 0, x = Lx/2, y = 0, y = Ly/2, y = Ly (FemPlate.cs:181-196).
 * Free dofs (FindFreeDofs, FemPlate.cs:72-101): *count of them, and when dofs is not NULL their
 * all-dof indices 3 node + d (host array of *count entries). */
int mgx_plate_free_dofs(int32_t nx, int32_t ny, double Lx, double Ly, int32_t *count, int32_t *dofs);
/* K_ff (sorted columns, element sums in ascending element order as the reference's DOK assembly) on
 * the device, and, when b is not NULL, the load vector of a uniform load q (host, count entries;
 * PlateElementVectorAssembler.cs:14-30).  Replaces FemPlate.CreateLinearSystem (FemPlate.cs:41-70). */
int mgx_matrix_plate(int32_t nx, int32_t ny, double Lx, double Ly, double t, double E, double nu, double q,
                     mgx_matrix *A, double *b);
/* Restriction / interpolation between the plate and its coarser model (element counts halved):
 * IGeometricMultigridModel.cs:76-194 on two axes, free-dof submatrices. */
int mgx_matrix_transfer_plate(int32_t nx, int32_t ny, double Lx, double Ly, mgx_matrix *R, mgx_matrix *P);
/* The plane-stress cantilever (FemCantilever2D.cs with FemCantileverBase.cs): nx*ny 4-node elements of
 * the 99-line topology optimization code (FemCantilever2D.cs:29-49; the matrix does not depend on
 * the element size), dofs (ux, uy) per node, node by node, every node with ix == 0 clamped
 * (FemCantileverBase.cs:137-158); dens (host, nx*ny, e = ex + nx ey, may be NULL = 1) scales each
 * element.  K_ff on the device with sorted columns, each entry summed in ascending element order as
 * the reference's DOK assembly (FemCantileverBase.cs:87-104). */
int mgx_matrix_cantilever2d(int32_t nx, int32_t ny, double E, double nu, const double *dens, mgx_matrix *A);
/* Restriction / interpolation between the 2D cantilever and its coarser model: IGeometricMultigridModel.cs:
 * 76-194 on two axes with two dofs per node, free-dof submatrices. */
int mgx_matrix_transfer_cantilever2d(int32_t nx, int32_t ny, mgx_matrix *R, mgx_matrix *P);
/* Galerkin hierarchy: K_0 from dens, K_{l+1} = P_l^T K_l P_l for trilinear P_l, computed
 * element by element on the device; levels[l] receives K_l (nlevels entries). */
int mgx_matrix_galerkin_hierarchy(int32_t nx, int32_t ny, int32_t nz, double E, double nu,
                                  const double *dens, int32_t nlevels, mgx_matrix *levels);
/* Upload a host CSR into a library-owned device matrix. */
int mgx_matrix_from_host(int32_t nrows, int32_t ncols, const int32_t *rowptr, const int32_t *col,
                         const double *val, mgx_matrix *out);
int mgx_matrix_view(mgx_matrix m, mgx_csr_view *view);
/* Structured grid of a matrix (elements per axis; zeros when unknown): set by the generators, read
 * by the solver for its structured-mesh node orders; settable on a matrix uploaded from host CSR
 * (its rows must be 3 x the free nodes nx (ny+1) (nz+1) of that grid). */
int mgx_matrix_grid(mgx_matrix m, int32_t *nxyz);
int mgx_matrix_set_grid(mgx_matrix m, const int32_t *nxyz);
/* copy to host arrays sized from the view (rowptr[nrows+1], col[nnz], val[nnz]) */
int mgx_matrix_download(mgx_matrix m, int32_t *rowptr, int32_t *col, double *val);
int mgx_matrix_destroy(mgx_matrix m);

/* ================================================================== 2. solver handle
 * Replaces OpenCLCsrGeometricMultigridSolver / OpenCLDuViGeometricMultigridSolver
 * (IGeometricMultigridSolver.cs:6-9, IOpenCLGeometricMultigridSolver.cs:3-6) and
 * OpenCLCsrPcgSolver (OpenCLCsrPcgSolver.cs:40-331), plus the new MG-PCG composition. */

#define MGX_METHOD_GMG        0   /* stationary V-cycle solver (OpenCL*GeometricMultigridSolver.Solve) */
#define MGX_METHOD_PCG_JACOBI 1   /* OpenCLCsrPcgSolver.Solve */
#define MGX_METHOD_MGPCG      2   /* PCG preconditioned by one symmetric V-cycle (new) */

#define MGX_FORMAT_CSR   0
#define MGX_FORMAT_DUVI  1

#define MGX_SMOOTHER_JACOBI      0   /* damped Jacobi, omega * D^-1 */
#define MGX_SMOOTHER_GS          1   /* multicolour Gauss-Seidel (symmetric in MG-PCG) */

#define MGX_STOP_MAXNORM 0   /* reference: every |r_i| <= tol (CsrGeometricMultigrid.cl:138-140) */
#define MGX_STOP_REL_L2  1   /* ||r||_2 <= tol * ||b||_2 */

#define MGX_PRECISION_FP64  0
#define MGX_PRECISION_MIXED 1   /* FP32 V-cycle and SpMV inside FP64 iterative refinement */

typedef struct {
    int32_t method;
    int32_t format;
    int32_t smoother;
    int32_t coarse_relaxation;  /* reference flag: Jacobi weight 2/lambda from the coarsest matrix
                                   (GeometricMultigridSolver.cs:144-151,209-214); used when omega <= 0 */
    double  omega;              /* Jacobi damping; > 0 overrides the reference rule */
    int32_t max_iterations;     /* cycles (GMG) or iterations (PCG) */
    double  tolerance;
    int32_t stop;
    int32_t depth;              /* GMG: levelDown = [T x depth, F x depth] (CreateDeepV) */
    int32_t level_iterations;   /* smoother sweeps per step (GMG) / pre and post sweeps (MG-PCG) */
    int32_t precision;
    int32_t nrhs;               /* right-hand sides (load cases), columns of b and x */
    double  duvi_tolerance;     /* value dedup tolerance (1e-10 in the reference) */
    int32_t duvi_mode;          /* see mgx_duvi_encode */
    int32_t inner_iterations;   /* mixed precision: max inner PCG iterations per refinement step */
    double  inner_tolerance;    /* mixed precision: inner relative reduction */
    int32_t profile;            /* 1: time every level-0 A*p SpMV with HIP events (mgx_stats.spmv_ms) */
    int32_t check_every;        /* PCG iterations (GMG: V-cycles) queued per host check of the device stop
                                   flag (default 4); the result does not depend on it.  The reference's GMG
                                   reads its flag back every cycle (OpenCLCsrGeometricMultigridSolver.cs:458-489) */
    int32_t use_graph;          /* 1: replay the queued PCG iterations as one captured hipGraph (default) */
    int32_t block_width;        /* MG-PCG on CSR with nrhs > 1: right-hand sides advanced together by one
                                   batched PCG (its SpMM reads each matrix once per block); 4, 8, 16 or 32,
                                   1 = one right-hand side at a time, 0 = chosen at initialize by timing
                                   PCG iterations of each width on the first right-hand sides (default) */
    double  strip_window;       /* batching form on structured levels: X window (bytes) of the strip order in
                                   which the block kernels visit the nodes, 0 = whole planes, < 0 = chosen at
                                   initialize by timing 6e6, 3e6, 1.5e6 and 0 (default) */
    int32_t gs_pairs;           /* batching form, multicolour Gauss-Seidel on structured levels: 1 = relax two
                                   colours per launch, line by line, where every edge between the two colours
                                   lies on one mesh line (default; the same sweep), 0 = one launch per colour */
} mgx_options;

typedef struct {
    int32_t iterations;      /* cycles (GMG), 0-based iteration index (PCG_JACOBI, as the
                                reference), or A*p products (MG-PCG) */
    int32_t converged;       /* 1 converged, 0 iteration cap, -1 failed (NaN or > 1e50) */
    double  residual_norm;   /* ||b - A x||_2 at exit */
    double  rel_residual;    /* ||b - A x||_2 / ||b||_2 */
    double  solve_ms;        /* device time of the solve */
    int32_t outer_iterations;/* mixed precision: refinement steps */
    int32_t spmv_launches;   /* level-0 A*p SpMV launches timed (options.profile) */
    double  spmv_ms;         /* their summed device time (HIP events on the solver stream) */
    int32_t block_width;     /* right-hand sides solved together with this one (1: alone); with
                                block_width > 1, solve_ms is the device time of that whole block */
    int32_t gs_launches;     /* batching form with options.profile: launches of the level-0 general Gauss-Seidel
                                sweeps timed (one per colour, or per colour pair with options.gs_pairs; the
                                sweeps from x = 0 are not timed) */
    double  gs_ms;           /* their summed device time (HIP events on the solver stream around each sweep) */
} mgx_stats;

typedef struct mgx_solver_s *mgx_solver;

void mgx_options_default(mgx_options *opt);
int mgx_solver_create(const mgx_options *opt, mgx_solver *out);

/* Level data, host arrays (copied at initialize, like the reference's blocking WriteBuffer,
 * OpenCLCsrGeometricMultigridSolver.cs:194-220).  Level 0 is the finest.  R maps level ->
 * level+1 and P maps level+1 -> level; both NULL on the coarsest level.  For MG-PCG, pass
 * R = P^T.  For PCG_JACOBI set level 0 only. */
typedef struct {
    int32_t nrows, ncols;
    int64_t nnz;
    const int32_t *rowptr, *col;
    const double *val;
} mgx_csr_host;
int mgx_solver_set_level(mgx_solver s, int32_t level, const mgx_csr_host *A,
                         const mgx_csr_host *R, const mgx_csr_host *P);
/* Same, from library-owned device matrices (no copy; they must outlive the solver). */
int mgx_solver_set_level_device(mgx_solver s, int32_t level, mgx_matrix A, mgx_matrix R, mgx_matrix P);
/* b: n x nrhs, column-major (one right-hand side after the other), host memory. */
int mgx_solver_set_rhs(mgx_solver s, const double *b, int32_t n, int32_t nrhs);
/* Builds D^-1, the Jacobi weights, the coarse factor, colourings, DuVi encodings and
 * FP32 copies; uploads everything.  Replaces Initialize (OpenCLCsrGeometricMultigridSolver.cs:145-285). */
int mgx_solver_initialize(mgx_solver s);
/* x: n x nrhs host array; in: initial guess when has_x0, out: solution (the reference writes
 * into xInitialGuess.RawData, OpenCLCsrGeometricMultigridSolver.cs:480).  stats: one per rhs. */
int mgx_solver_solve(mgx_solver s, double *x, int32_t has_x0, mgx_stats *stats);
/* Same with device-resident b and x (n x nrhs each, column-major); no host transfer. */
int mgx_solver_solve_device(mgx_solver s, const double *b_dev, double *x_dev, int32_t has_x0, mgx_stats *stats);
/* Block width chosen at initialize (1 = one right-hand side at a time) and, when the options
 * asked for the automatic choice, the timed ms per block PCG iteration for K = 4, 8, 16, 32
 * (0 where not timed); tune_ms may be NULL.  No reference counterpart (batching form). */
int mgx_solver_block_width(mgx_solver s, int32_t *width, double *tune_ms);
/* Matrix layout the block kernels read on a level (batching form; MGX_LAYOUT_CSR whenever the
 * right-hand sides are solved one at a time).  No reference counterpart. */
#define MGX_LAYOUT_CSR          0   /* CSR (row kernels / generic SpMM) */
#define MGX_LAYOUT_NODE_TRIPLE  1   /* CSR read by node triples (rows 3q..3q+2 share their columns) */
#define MGX_LAYOUT_NODEBLOCK    2   /* 3x3 node blocks (mgx_nodeblock_from_csr) */
int mgx_solver_level_layout(mgx_solver s, int32_t level, int32_t *layout);
/* Layout of the level's R and P in the batching form's V-cycle: MGX_LAYOUT_NODE_TRANSFER when both
 * are node transfers (mgx_nodetransfer_from_csr), else MGX_LAYOUT_CSR.  No reference counterpart. */
#define MGX_LAYOUT_NODE_TRANSFER 3
/* The DuVi block layout (options.format = MGX_FORMAT_DUVI): the batching form's SpMM, Gauss-Seidel
 * and transfers decode it, K columns per pass (level and transfer layouts). */
#define MGX_LAYOUT_DUVI          4
int mgx_solver_transfer_layout(mgx_solver s, int32_t level, int32_t *layout);
/* X window (bytes) of the strip order in which the block kernels visit the nodes of structured
 * levels (0: whole planes) and, when the autotuner chose it, the ms per block PCG iteration it timed
 * for the windows 6e6, 3e6, 1.5e6 and 0 (tune_ms: 4 entries, may be NULL).  No reference
 * counterpart (batching form). */
int mgx_solver_strip_window(mgx_solver s, double *bytes, double *tune_ms);
/* Colours of the multicolour Gauss-Seidel node colouring of a level (0 without that smoother). */
int mgx_solver_level_colors(mgx_solver s, int32_t level, int32_t *ncolors);
/* The node order of the level's forward Gauss-Seidel sweep as the solver runs it: the sweep
 * relaxes node order[0], then order[1], ... (rows 3q, 3q+1, 3q+2 of node q in turn), the
 * backward sweep the exact reverse; nodes that the kernels relax together (one colour) are
 * uncoupled, so the parallel sweep equals this sequential one: colour 0's nodes, then colour
 * 1's, ...  A level whose rows do not all come in triples with one shared column list (the
 * plate's supported nodes keep one or two free dofs) has row groups instead: maximal runs of up
 * to 3 consecutive rows with identical column lists, numbered in row order, and order[] holds
 * group numbers (mgx_solver_level_groups gives their count).  n/3 entries for triples, else one
 * per group.  The reference's GPU "Gauss-Seidel" (CsrGeometricMultigrid.cl:207-225) updates in
 * place with racy reads; this ordering is the documented deviation. */
int mgx_solver_gs_order(mgx_solver s, int32_t level, int32_t *order);
/* Row groups of a level's Gauss-Seidel nodes: 0 when the nodes are the row triples, else the
 * group count and, if gptr is not null, the group starts gptr[0 .. count] (group q is rows
 * gptr[q] .. gptr[q+1]-1).  No counterpart in the reference, whose kernels take rows one by one. */
int mgx_solver_level_groups(mgx_solver s, int32_t level, int32_t *count, int32_t *gptr);
/* Colour pairs (2k, 2k+1) of the level that the batching form's Gauss-Seidel relaxes in one launch
 * (options.gs_pairs): block i of the launch relaxes the pair's nodes on mesh line i, the first
 * colour's, then after a block barrier the second colour's.  A pair qualifies when every edge
 * between its two colours joins two nodes of one line, so the sweep is the one gs_order gives.
 * 0 when the sweeps run one launch per colour.  No reference counterpart (batching form). */
int mgx_solver_gs_pairs(mgx_solver s, int32_t level, int32_t *pairs);
/* Replaces ReleaseOpenCLResources (OpenCLCsrGeometricMultigridSolver.cs:106-138). */
int mgx_solver_release(mgx_solver s);

/* ================================================================== on-disk hierarchy
 * A multigrid hierarchy stored once and loaded without assembly, Galerkin products or DuVi
 * encoding.  No reference counterpart: the reference rebuilds every level from the model's DOK
 * matrices at each Initialize and keeps its DuVi encodings (the DuViMat types,
 * OpenCLDuViGeometricMultigridSolver.cs:16-31, filled by FromDokRowMajor, :753-851) in memory
 * only.  Versioned little-endian binary: a header, one section per matrix (CSR A of every level, R
 * and P towards the next coarser one, optionally the DuVi block layout of each, and the level's
 * structured grid), each with a 64-bit checksum of its payload, and a trailer checksumming the
 * header and the section heads.  Every reader verifies what it reads and fails with
 * MGX_ERR_INVALID_ARGUMENT on a bad magic, checksum or truncation, MGX_ERR_NOT_IMPLEMENTED on
 * another format version. */
#define MGX_HIERARCHY_VERSION 1
/* Host levels, finest first: A[nlevels], R[nlevels-1] (level l -> l+1), P[nlevels-1]; grid:
 * nlevels x 3 element counts (nx, ny, nz; zeros when unstructured) or NULL; duvi_mode: -1 stores
 * no DuVi sections, else the mode of mgx_duvi_encode (0, 4, 8) with tolerance duvi_tol.  The
 * file is written under path + ".part" and renamed when complete. */
int mgx_hierarchy_write(const char *path, int32_t nlevels, const mgx_csr_host *A, const mgx_csr_host *R,
                        const mgx_csr_host *P, const int32_t *grid, int32_t duvi_mode, double duvi_tol);
typedef struct {
    int32_t version, nlevels, has_grid, has_duvi;
    int64_t payload_bytes;
} mgx_hierarchy_summary;
/* Reads and verifies the whole file (every checksum). */
int mgx_hierarchy_info(const char *path, mgx_hierarchy_summary *info);
/* CSR of one level's A (which = 0), R (1) or P (2), in host arrays allocated by the library. */
typedef struct {
    int32_t nrows, ncols;
    int64_t nnz;
    int32_t *rowptr, *col;
    double *val;
} mgx_csr_owned;
int mgx_hierarchy_read_csr(const char *path, int32_t level, int32_t which, mgx_csr_owned *out);
void mgx_csr_owned_free(mgx_csr_owned *m);
/* The plate matrices of mgx_matrix_plate / mgx_matrix_transfer_plate in host arrays (freed by
 * mgx_csr_owned_free); b (NULL or the free-dof count) receives the load vector. */
int mgx_plate_system_host(int32_t nx, int32_t ny, double Lx, double Ly, double t, double E, double nu, double q,
                          mgx_csr_owned *K, double *b);
int mgx_plate_transfer_host(int32_t nx, int32_t ny, double Lx, double Ly, mgx_csr_owned *R, mgx_csr_owned *P);
/* The 2D cantilever matrices of mgx_matrix_cantilever2d / mgx_matrix_transfer_cantilever2d in host arrays
 * (freed by mgx_csr_owned_free). */
int mgx_cantilever2d_system_host(int32_t nx, int32_t ny, double E, double nu, const double *dens, mgx_csr_owned *K);
int mgx_cantilever2d_transfer_host(int32_t nx, int32_t ny, mgx_csr_owned *R, mgx_csr_owned *P);
/* DuVi block layout of one level's A / R / P (freed by mgx_duvi_host_free). */
int mgx_hierarchy_read_duvi(const char *path, int32_t level, int32_t which, mgx_duvi_host *out);
/* The level's structured grid (3 element counts). */
int mgx_hierarchy_read_grid(const char *path, int32_t level, int32_t *nxyz);
/* Library-owned device matrices in and out (the grid of each A travels with it). */
int mgx_hierarchy_save(const char *path, int32_t nlevels, const mgx_matrix *A, const mgx_matrix *R,
                       const mgx_matrix *P, int32_t duvi_mode, double duvi_tol);
int mgx_hierarchy_load(const char *path, int32_t max_levels, int32_t *nlevels, mgx_matrix *A, mgx_matrix *R,
                       mgx_matrix *P);

/* ================================================================== plumbing
 * Device memory and timing helpers so that a host program (the C# binding, the Python test
 * harness) needs no other GPU runtime binding.  Plain hipMalloc / hipMemcpy wrappers. */
int mgx_device_malloc(void **ptr, uint64_t bytes);
int mgx_device_free(void *ptr);
int mgx_memcpy_htod(void *dst, const void *src, uint64_t bytes);
int mgx_memcpy_dtoh(void *dst, const void *src, uint64_t bytes);
int mgx_memcpy_dtod(void *dst, const void *src, uint64_t bytes);
int mgx_memset(void *dst, int32_t value, uint64_t bytes);
int mgx_device_synchronize(void);
/* Average device time (HIP events on one stream, after `warmup` untimed launches) of
 * `repeats` back-to-back launches of a CSR row kernel on device arrays:
 * kind 0: y = A x; 1: y = x + pre .* (b - A x); 2: y = b - A x; 3: the same as 0 with FP32
 * values and vectors (val, x, y point to float); 4: y = A x with the x.y partials and their
 * reduction, the PCG step q = A p, p.q (square matrices). */
int mgx_time_csr_kernel(int32_t kind, int32_t nrows, const int32_t *rowptr, const int32_t *col, const void *val,
                        const void *x, const void *b, const void *pre, void *y, int32_t warmup, int32_t repeats,
                        double *avg_ms);
/* Same for the block SpMM Y = A X with k right-hand sides.  flags bit 0: FP32 values and vectors
 * (mgx_csr_spmm_f32) instead of FP64; bit 1: the node-triple kernel of the block solver (rows
 * 3q..3q+2 share one column list; triples of unequal length are rejected); bit 2: the block
 * PCG's form that also produces the per-column partials of X.Y (Y = A X itself unchanged); bit 3: the
 * node-block kernel (mgx_nodeblock_spmm; the copy is built from the CSR before timing, FP32 blocks
 * with bit 0; takes bit 2, not bit 1). */
int mgx_time_spmm(int32_t flags, int32_t k, int32_t nrows, const int32_t *rowptr, const int32_t *col, const void *val,
                  const void *X, void *Y, int32_t warmup, int32_t repeats, double *avg_ms);
/* Coarse direct solve applied to k columns at once: Y = M X with M the dense FP64 inverse of the
 * coarsest matrix (nc x nc, row-major, device) and X, Y nc x k row-major device blocks (FP32 with
 * fp32 = 1, else FP64); the batching form's coarse step (the reference solves the coarsest system
 * directly for one right-hand side, OpenCLCsrGeometricMultigridSolver.cs:364-367).  Synchronous. */
int mgx_dense_block_apply(int32_t nc, const double *M, int32_t k, const void *X, void *Y, int32_t fp32, void *stream);
/* Average device time of a plain copy kernel (16-byte vector loads and stores) moving `bytes`
 * (a multiple of 16) from src to dst: the calibration of HBM byte counters. */
int mgx_time_copy_kernel(const void *src, void *dst, uint64_t bytes, int32_t warmup, int32_t repeats, double *avg_ms);
/* Average device time of a streaming read of `bytes` (a multiple of 16) from src (16-byte
 * non-temporal loads, nothing written; sink: 16 bytes of device memory the kernel may write): the
 * practical HBM read ceiling that the gather kernels are compared with.  No reference counterpart. */
int mgx_time_read_kernel(const void *src, uint64_t bytes, void *sink, int32_t warmup, int32_t repeats, double *avg_ms);
/* Same for the DuVi SpMV y = A x.  form 0: the kernel the library chooses (the stream form for
 * matrices of 16 or more slots per row, else the wave form); 1: the wave form for any matrix. */
int mgx_time_duvi_spmv(const mgx_duvi *A, const double *x, double *y, int32_t form, int32_t warmup, int32_t repeats,
                       double *avg_ms);

/* ================================================================== workload helpers
 * SplitMix64-seeded xoshiro256** (U[0,1) from the top 53 bits; N(0,1) by Box-Muller,
 * cosine branch), used to draw element moduli / load cases reproducibly. */
int mgx_rng_uniform(uint64_t seed, int64_t n, double *out);
int mgx_rng_normal(uint64_t seed, int64_t n, double *out);

#ifdef __cplusplus
}
#endif
#endif
