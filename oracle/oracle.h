/*
 * oracle.h -- CPU restatement of the reference algorithms on the MG-PCG / SpMV path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * there only as the checker (or the timed CPU baseline), never as the thing
 * measured or shipped.  The product (compressiontests_amd/libmgx.so) never
 * links, loads or calls it.
 *
 * Every function names the reference file:line it restates.  Paths are
 * relative to the reference repository (nickreserved/CompressionTests).
 * Built with -O3 -ffp-contract=off so that products and sums round exactly as
 * the reference's managed (.NET, no FMA contraction) code does.  A second build
 * (liboracle_omp.so, -fopenmp) runs the row loops of the SpMV, residual, Jacobi
 * and within-colour Gauss-Seidel in parallel: each row is computed by one thread
 * exactly as in the serial build and the dot products stay serial, so both builds
 * return bit-identical results; it serves as the all-core CPU baseline.
 *
 * Pinning (what checks each function against something other than itself) is
 * listed in the header comment of each source file.
 */
#ifndef MGX_ORACLE_H
#define MGX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* row-independent loops: parallel in the OpenMP build, plain loops otherwise */
#ifdef _OPENMP
#define ORC_PARALLEL_FOR _Pragma("omp parallel for schedule(static)")
#else
#define ORC_PARALLEL_FOR
#endif
/* threads the library runs its parallel loops on (1 in the serial build) */
int orc_num_threads(void);

/* ---------------- RNG (workload definition; SplitMix64-seeded xoshiro256**) */
void orc_rng_uniform(uint64_t seed, int64_t n, double *out);  /* U[0,1), top 53 bits */
void orc_rng_normal(uint64_t seed, int64_t n, double *out);   /* N(0,1), Box-Muller (cos branch) */
uint64_t orc_splitmix64_next(uint64_t *state);

/* ---------------- FE cantilever (FemCantileverBase.cs, FemCantilever3D.cs) */
void orc_ke_hex8(double E, double nu, double *Ke /*24x24 row-major*/);
int64_t orc_cantilever_nnz(int nx, int ny, int nz);
/* K_ff in sorted CSR; element contributions summed in element order like
   DokRowMajor.AddSubmatrix.  dens may be NULL (all 1.0). */
void orc_cantilever_assemble(int nx, int ny, int nz, double E, double nu, const double *dens,
                             int32_t *rowptr, int32_t *col, double *val);
/* generic: per-element 24x24 matrices (row-major, element order e = ex + nx*(ey + ny*ez)) */
void orc_assemble_elements(int nx, int ny, int nz, const double *Kel,
                           int32_t *rowptr, int32_t *col, double *val);
void orc_cantilever_manufactured(int nx, int ny, int nz, double Lx, double Ly, double Lz,
                                 double E, double q, double *Uf);
void orc_cantilever_tip_load(int nx, int ny, int nz, double total, double *f);
void orc_csr_multiply_dok_order(int nx, int ny, int nz, const int32_t *rowptr, const int32_t *col,
                                const double *val, const double *x, double *y);

/* ---------------- reference restriction / interpolation (IGeometricMultigridModel.cs:76-194) */
void orc_rp_reference_nnz(int nx, int ny, int nz, int64_t *nnzR, int64_t *nnzP);
void orc_rp_reference(int nx, int ny, int nz,
                      int32_t *Rrowptr, int32_t *Rcol, double *Rval,
                      int32_t *Prowptr, int32_t *Pcol, double *Pval);
/* exact trilinear prolongation P (fine free x coarse free), and R = P^T */
void orc_p_trilinear_nnz(int nx, int ny, int nz, int64_t *nnzP);
void orc_p_trilinear(int nx, int ny, int nz, int32_t *Prowptr, int32_t *Pcol, double *Pval);

/* ---------------- generic CSR helpers */
void orc_csr_transpose(int32_t nrows, int32_t ncols, const int32_t *rp, const int32_t *ci, const double *v,
                       int32_t *rpT, int32_t *ciT, double *vT);
/* Galerkin triple product C = P^T A P (textbook SpGEMM, sorted columns) */
int64_t orc_galerkin_nnz(int32_t nf, int32_t nc, const int32_t *Arp, const int32_t *Aci,
                         const int32_t *Prp, const int32_t *Pci);
void orc_galerkin(int32_t nf, int32_t nc, const int32_t *Arp, const int32_t *Aci, const double *Av,
                  const int32_t *Prp, const int32_t *Pci, const double *Pv,
                  int32_t *Crp, int32_t *Cci, double *Cv);

/* ---------------- kernels (CsrGeometricMultigrid.cl, CsrCG.cl) */
void orc_csr_spmv(int32_t nrows, const int32_t *rp, const int32_t *ci, const double *v,
                  const double *x, double *y, int clear);
void orc_csr_residual(int32_t nrows, const int32_t *rp, const int32_t *ci, const double *v,
                      const double *b, const double *x, double *r);
int orc_csr_residual_check(int32_t nrows, const int32_t *rp, const int32_t *ci, const double *v,
                           const double *b, const double *x, double *r, double tol);
void orc_jacobi_initial(int32_t n, const double *pre, const double *b, double *x);
void orc_csr_jacobi(int32_t n, const double *pre, const int32_t *rp, const int32_t *ci, const double *v,
                    const double *b, const double *x, double *y);
void orc_jacobi_preconditioner(int32_t n, const int32_t *rp, const int32_t *ci, const double *v, double *d);
double orc_eigen_upper_bound(int32_t n, const int32_t *rp, const double *v);
double orc_dot(int32_t n, const double *a, const double *b);
double orc_power_bound(int32_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *invd);
void orc_gs_forward(int32_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *b, double *x);
/* node-block greedy colouring (block = 3 dofs); returns number of colours */
int orc_block_coloring(int32_t nblocks, int bs, const int32_t *rp, const int32_t *ci, int32_t *color);
void orc_gs_multicolor(int32_t n, int bs, const int32_t *rp, const int32_t *ci, const double *v,
                       int ncolors, const int32_t *color, const double *b, double *x, int backward);
/* Gauss-Seidel sweep over node blocks of bs rows in a given node order: forward = nodes
   order[0 .. nn), rows in order; backward = the exact reverse (nodes and rows). */
void orc_gs_ordered(int32_t n, int bs, const int32_t *rp, const int32_t *ci, const double *v,
                    const int32_t *order, const double *b, double *x, int backward);

/* ---------------- dense direct solve (LDL^T / Cholesky of the coarsest level) */
int orc_dense_cholesky(int32_t n, double *A /*in: SPD, out: L lower*/);
void orc_dense_cholesky_solve(int32_t n, const double *L, const double *b, double *x);
void orc_csr_to_dense(int32_t n, const int32_t *rp, const int32_t *ci, const double *v, double *A);

/* ---------------- DuVi GPU block layout (OpenCLDuViGeometricMultigridSolver.cs:753-851) */
typedef struct {
    int32_t nrows;
    int64_t nnz, nblocks, nuniq;
    int32_t vi_bytes;        /* 1, 2, 4, or 0 = raw values (DU + raw) */
    int32_t *row_blocks;     /* [nrows+1] RowToColumnIndices */
    int32_t *row_dist;       /* [nrows+1] RowToDistances   */
    int32_t *block_col;      /* [nblocks] ColumnIndices     */
    uint8_t *dist;           /* [nnz]     Distances         */
    void    *vi;             /* [nnz]     ValueIndices (u8/u16/u32) or NULL */
    double  *values;         /* [nuniq] unique values, or [nnz] raw values */
} orc_duvi_block;
/* mode: 0 = reference selection (u8 if <=256 unique, else u16 if <=65536, else raw),
         4 = force u32 indices, 8 = force raw values.  tol <= 0: exact dedup. */
int orc_duvi_block_encode(int32_t nrows, const int32_t *rp, const int32_t *ci, const double *v,
                          double tol, int mode, orc_duvi_block *out);
void orc_duvi_block_free(orc_duvi_block *m);
void orc_duvi_block_spmv(const orc_duvi_block *m, const double *x, double *y, int clear);
/* decode to CSR (rp[nrows+1], ci[nnz], v[nnz]), DuVi2Mat (OpenCLDuViGeometricMultigridSolver.cs:394-415) */
void orc_duvi_block_to_csr(const orc_duvi_block *m, int32_t *rp, int32_t *ci, double *v);
/* .NET Dictionary<double,int>(ToleranceComparer) emulation: value -> index, first-seen order */
int64_t orc_dedup(int64_t n, const double *v, double tol, int32_t *idx, double *uniq);

/* ---------------- DuVi CPU Unit stream (DuViCompressedSparseMatrix.cs:11-441) */
typedef struct {
    int32_t nrows, ncols;
    int64_t nunits, nbytes, nvals, nuniq;
    int64_t *unit_off;       /* [nunits+1] byte offsets into bytes */
    uint8_t *bytes;
    int32_t *vi;             /* [nvals] */
    double  *uniq;           /* [nuniq] */
} orc_duvi_units;
int orc_duvi_units_from_csr(int32_t nrows, int32_t ncols, const int32_t *rp, const int32_t *ci,
                            const double *v, double tol, orc_duvi_units *out);
int orc_duvi_units_from_dense(int32_t nrows, int32_t ncols, const double *dense, double tol,
                              orc_duvi_units *out);
void orc_duvi_units_free(orc_duvi_units *m);
void orc_duvi_units_spmv(const orc_duvi_units *m, const double *x, double *y);
void orc_duvi_units_gs(const orc_duvi_units *m, const double *b, double *x);

/* ---------------- solvers */
typedef struct {
    int32_t n;                      /* rows of this level */
    const int32_t *Arp, *Aci; const double *Av;   /* level matrix */
    const int32_t *Rrp, *Rci; const double *Rv;   /* restriction to level+1 (NULL on coarsest) */
    const int32_t *Prp, *Pci; const double *Pv;   /* interpolation from level+1 (NULL on coarsest) */
    const int32_t *gs_order;        /* Gauss-Seidel node order of the forward sweep (n/3 nodes, or the
                                       ngroups groups; NULL: multicolour order of orc_block_coloring /
                                       orc_group_coloring) */
    const int32_t *gptr;            /* node groups of the Gauss-Seidel sweep (orc_row_groups); NULL: triples */
    int32_t ngroups;
} orc_level;

typedef struct {
    int32_t iterations;
    int32_t converged;     /* 1 converged, 0 not, -1 failed (NaN / overflow) */
    double  residual_norm; /* ||b - A x||_2 at exit */
    double  rel_residual;  /* ||b - A x||_2 / ||b||_2 */
} orc_stats;

/* GPU-semantics stationary V-cycle (OpenCLCsrGeometricMultigridSolver.cs:334-536).
   smoother: 0 = damped Jacobi (ping-pong), 1 = multicolour Gauss-Seidel (bs = 3) */
int orc_gmg_solve(int nlevels, const orc_level *lv, const double *const *pre, int smoother,
                  int depth, int level_iterations, int max_cycles, double tol,
                  const double *b, double *x, int has_x0, orc_stats *st);
/* Jacobi PCG (OpenCLCsrPcgSolver.cs:250-331); stop on max-norm of r <= tol */
int orc_pcg_jacobi(int32_t n, const int32_t *rp, const int32_t *ci, const double *v,
                   const double *b, double *x, int has_x0, int max_it, double tol, orc_stats *st);
/* MG-preconditioned CG (new composition): symmetric V-cycle, R = P^T,
   nu pre/post sweeps; smoother 0 = Jacobi(omega; omega <= 0: per level 2 / (1.2 lambda) with
   lambda from orc_power_bound), 1 = symmetric multicolour GS.
   stop: 0 = max-norm (reference), 1 = ||r||/||b|| <= tol. */
int orc_vcycle_apply(int nlevels, const orc_level *lv, int smoother, double omega, int nu,
                     const double *r, double *z);
int orc_mgpcg(int nlevels, const orc_level *lv, int smoother, double omega, int nu,
              int max_it, double tol, int stop, const double *b, double *x, orc_stats *st);
/* orc_mgpcg that also records the wall time (monotonic clock, seconds) of every PCG iteration
   in iter_seconds[0 .. max_it) (set-up and the first V-cycle are not part of any entry): the
   timed CPU baseline. */
int orc_mgpcg_timed(int nlevels, const orc_level *lv, int smoother, double omega, int nu,
                    int max_it, double tol, int stop, const double *b, double *x, orc_stats *st,
                    double *iter_seconds);

/* ---------------- Kirchhoff plate workload (oracle_fem2d.c) */
void orc_plate_ke(double a, double b, double E, double nu, double t, double *Ke);
void orc_plate_load(double a, double b, double q, double *fe);
int32_t orc_plate_nfree(int nx, int ny, double Lx, double Ly);
void orc_plate_free_dofs(int nx, int ny, double Lx, double Ly, int32_t *dofs);
int64_t orc_plate_nnz(int nx, int ny, double Lx, double Ly);
void orc_plate_assemble(int nx, int ny, double Lx, double Ly, double t, double E, double nu, double q,
                        int32_t *rowptr, int32_t *col, double *val, double *b);
void orc_plate_transfer_nnz(int nx, int ny, double Lx, double Ly, int64_t *nnzR, int64_t *nnzP);
void orc_plate_transfer(int nx, int ny, double Lx, double Ly, int32_t *Rrowptr, int32_t *Rcol, double *Rval,
                        int32_t *Prowptr, int32_t *Pcol, double *Pval);
/* ---------------- plane-stress cantilever, FemCantilever2D (oracle_fem2d.c) */
void orc_ke_q4(double E, double nu, double *Ke);
int32_t orc_cant2d_nfree(int nx, int ny);
int64_t orc_cant2d_nnz(int nx, int ny);
void orc_cant2d_assemble(int nx, int ny, double E, double nu, const double *dens, int32_t *rowptr, int32_t *col,
                         double *val);
void orc_cant2d_manufactured(int nx, int ny, double Lx, double Ly, double Lz, double E, double q, double *Uf);
void orc_cant2d_multiply_dok_order(int nx, int ny, const int32_t *rowptr, const int32_t *col, const double *val,
                                   const double *x, double *y);
void orc_cant2d_transfer_nnz(int nx, int ny, int64_t *nnzR, int64_t *nnzP);
void orc_cant2d_transfer(int nx, int ny, int32_t *Rrowptr, int32_t *Rcol, double *Rval, int32_t *Prowptr,
                         int32_t *Pcol, double *Pval);
/* node groups: rows gptr[g] .. gptr[g+1]-1 (the free dofs of one mesh node) */
int32_t orc_row_groups(int32_t n, const int32_t *rp, const int32_t *ci, int32_t *gptr);
int orc_group_coloring(int32_t ngroups, const int32_t *gptr, const int32_t *rp, const int32_t *ci, int32_t *color);
void orc_gs_groups(int32_t ngroups, const int32_t *gptr, const int32_t *rp, const int32_t *ci, const double *v,
                   const int32_t *order, int ncolors, const int32_t *color, const double *b, double *x,
                   int backward);

/* ---------------- SIMP compliance adjoint */
void orc_simp_gradient(int nx, int ny, int nz, const double *Ke0, const double *rho, double penal,
                       double Emin, double E0, const double *u_free, double *grad);

#ifdef __cplusplus
}
#endif
#endif
