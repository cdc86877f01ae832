// mgx_internal.h -- declarations shared by the HIP kernels, the C-ABI layer and the
// host-side setup code of libmgx.so.  Not part of the public ABI (include/mgx.h).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mgx.h"

namespace mgx {

// ------------------------------------------------------------------ errors
void set_error(const std::string &msg);
struct Error {
    int code;
    std::string msg;
};
#define MGX_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::mgx::Error{e_ == hipErrorOutOfMemory ? MGX_ERR_OUT_OF_MEMORY : MGX_ERR_DEVICE, \
                               std::string(#call) + ": " + hipGetErrorString(e_)};            \
    } while (0)
#define MGX_REQUIRE(cond, code, msg)                                                         \
    do { if (!(cond)) throw ::mgx::Error{(code), (msg)}; } while (0)

// ------------------------------------------------------------------ device control block
// Scalars of one solve that stay on the device (no host round trip per iteration).
struct Ctl {
    double rz;        // r . z (current)
    double pq;        // p . A p
    double alpha, beta;
    double rnorm2;    // ||r||^2
    double rmax;      // max |r_i| (NaN-propagating)
    double bnorm2;    // ||b||^2
    double scratch[9];
    int32_t iter;     // A*p products done
    int32_t done;     // 1 once the stop test passed or failed; kernels of later iterations return
    int32_t status;   // 1 converged, -1 failed
    int32_t stop;     // MGX_STOP_*
    double tol;
    int32_t max_it;
    uint32_t counter[16];   // last-block tickets of grid reductions
};

constexpr int kBlock = 256;

// ------------------------------------------------------------------ kernel launchers (mgx_kernels.hip)
// lanes-per-row chosen from the mean row length
int lanes_for(int64_t nnz, int32_t nrows);

void launch_spmv(int32_t nrows, const int32_t *rp, const int32_t *ci, const double *v,
                 const double *x, double *y, bool clear, int lanes, const int32_t *done, hipStream_t s);
void launch_spmv_f32(int32_t nrows, const int32_t *rp, const int32_t *ci, const float *v,
                     const float *x, float *y, bool clear, int lanes, const int32_t *done, hipStream_t s);
void launch_residual(int32_t nrows, const int32_t *rp, const int32_t *ci, const double *v,
                     const double *b, const double *x, double *r, int lanes, const int32_t *done, hipStream_t s);
void launch_residual_f32(int32_t nrows, const int32_t *rp, const int32_t *ci, const float *v,
                         const float *b, const float *x, float *r, int lanes, const int32_t *done, hipStream_t s);
void launch_residual_flag(int32_t nrows, const int32_t *rp, const int32_t *ci, const double *v,
                          const double *b, const double *x, double *r, double tol, int32_t *flag,
                          int lanes, hipStream_t s, const int32_t *done = nullptr);
void launch_jacobi(int32_t nrows, const double *pre, const int32_t *rp, const int32_t *ci, const double *v,
                   const double *b, const double *x, double *y, int lanes, const int32_t *done, hipStream_t s);
void launch_jacobi_f32(int32_t nrows, const float *pre, const int32_t *rp, const int32_t *ci, const float *v,
                       const float *b, const float *x, float *y, int lanes, const int32_t *done, hipStream_t s);
void launch_jacobi_initial(int32_t n, const double *pre, const double *b, double *x, const int32_t *done, hipStream_t s);
void launch_jacobi_initial_f32(int32_t n, const float *pre, const float *b, float *x, const int32_t *done, hipStream_t s);
// one colour of a multicolour Gauss-Seidel sweep over nodes: node q = rows 3q .. 3q+2, or, with gptr,
// rows gptr[q] .. gptr[q+1]-1 (1 to 3 rows: the plate's supported nodes have 2 free dofs)
void launch_gs_color(int32_t nnodes, const int32_t *nodes, const int32_t *rp, const int32_t *ci, const double *v,
                     const double *b, double *x, bool backward, const int32_t *done, hipStream_t s,
                     const int32_t *gptr = nullptr);
void launch_gs_color_f32(int32_t nnodes, const int32_t *nodes, const int32_t *rp, const int32_t *ci, const float *v,
                         const float *b, float *x, bool backward, const int32_t *done, hipStream_t s,
                         const int32_t *gptr = nullptr);
void launch_dot(int32_t n, const double *a, const double *b, double *result, double *partials,
                uint32_t *counter, hipStream_t s);
int dot_partial_count(int64_t n);   // partial sums launch_dot writes (its first kernel's grid)
void launch_dense_gemv(int32_t n, const double *M, const double *x, double *y, bool accumulate_into_y,
                       const int32_t *done, hipStream_t s);
void launch_duvi_spmv(const mgx_duvi &A, const double *x, double *y, bool clear, const int32_t *done, hipStream_t s);
// DuVi row kernels with a fused epilogue (values: A's table, or an FP32 copy of it for T = float)
enum DuviEpi : int { DEPI_SPMV = 0, DEPI_ADD = 1, DEPI_RESID = 2, DEPI_JACOBI = 3, DEPI_CHECK = 4 };
template <typename T>
void launch_duvi(const mgx_duvi &A, const T *values, int epi, const T *x, const T *b, const T *pre, T *y,
                 int32_t *flag, double tol, const int32_t *done, hipStream_t s);
void set_duvi_wave_only(bool on);   // timing comparisons: the wave form for long rows too
// batching form on the DuVi layout: Y (op)= A X with K columns (BlockEpi epilogues; BEPI_DOT publishes
// per-column X.Y partials of duvi_spmm_grid(K, nrows) blocks), and one colour of a multicolour
// Gauss-Seidel sweep for K columns (node triples, or the row groups gptr)
template <typename T>
void launch_duvi_spmm(int K, int epi, const mgx_duvi &A, const T *values, const T *X, const T *B, const T *pre, T *Y,
                      const int32_t *done, double *partials, int stride, hipStream_t s);
int duvi_spmm_grid(int K, int32_t nrows);
template <typename T>
void launch_duvi_gs_block(int K, int32_t nnodes, const int32_t *nodes, const int32_t *gptr, const mgx_duvi &A,
                          const T *values, const T *B, T *X, bool backward, const int32_t *done, hipStream_t s);
// one colour of a multicolour Gauss-Seidel sweep on the DuVi layout (nodes of 3 rows)
template <typename T>
void launch_duvi_gs(const mgx_duvi &A, const T *values, int32_t nnodes, const int32_t *nodes, const T *b, T *x,
                    bool backward, const int32_t *done, hipStream_t s, const int32_t *gptr = nullptr);
void launch_simp_gradient(int32_t nx, int32_t ny, int32_t nz, const double *Ke0, const double *rho, double penal,
                          double Emin, double E0, const double *u, double *grad, hipStream_t s);

// PCG pieces working on Ctl (all honour ctl->done)
// two-phase reductions: producers store per-block partials, k_finalize (one block) combines them
enum FinMode : int { FIN_ALPHA = 0, FIN_UPDATE = 1, FIN_BETA = 2, FIN_BETA_FIRST = 3, FIN_NORMS = 4, FIN_DOT = 5 };
constexpr int kPartialStride = 4096;   // slot stride of the 3-slot reductions (grid_stream <= 2048 blocks)
void launch_finalize(int nparts, const double *partials, int stride, Ctl *ctl, int mode, double *result, hipStream_t s);
void launch_pcg_norms(int32_t n, const double *b, const double *r, Ctl *ctl, double *partials, int stride, hipStream_t s);
void launch_spmv_dot_alpha(int32_t n, const int32_t *rp, const int32_t *ci, const double *v,
                           const double *p, double *q, int lanes, Ctl *ctl, double *partials, hipStream_t s);
int reduce_slots_for_rows(int64_t nrows, int lanes);
void launch_pcg_update(int32_t n, const double *p, const double *q, double *x, double *r,
                       Ctl *ctl, double *partials, hipStream_t s);
void launch_dot_beta(int32_t n, const double *r, const double *z, Ctl *ctl, double *partials, bool first, hipStream_t s);
void launch_update_p(int32_t n, const double *z, double *p, const Ctl *ctl, bool first, hipStream_t s);
// precision form: FP32 vectors, FP64 scalars / accumulation
void launch_pcg_norms_f32(int32_t n, const float *b, const float *r, Ctl *ctl, double *partials, int stride, hipStream_t s);
void launch_spmv_dot_alpha_f32(int32_t n, const int32_t *rp, const int32_t *ci, const float *v, const float *p,
                               float *q, int lanes, Ctl *ctl, double *partials, hipStream_t s);
template <typename T>
void launch_dot_alpha(int32_t n, const T *p, const T *q, Ctl *ctl, double *partials, hipStream_t s);
void launch_pcg_update_f32(int32_t n, const float *p, const float *q, float *x, float *r, Ctl *ctl,
                           double *partials, hipStream_t s);
void launch_dot_beta_f32(int32_t n, const float *r, const float *z, Ctl *ctl, double *partials, bool first, hipStream_t s);
void launch_update_p_f32(int32_t n, const float *z, float *p, const Ctl *ctl, bool first, hipStream_t s);
void launch_dense_gemv_f32io(int32_t n, const double *M, const float *x, float *y, const int32_t *done, hipStream_t s);
// precision helpers
void launch_convert_f64_f32(int64_t n, const double *a, float *b, const int32_t *done, hipStream_t s);
void launch_convert_f32_f64(int64_t n, const float *a, double *b, const int32_t *done, hipStream_t s);
void launch_fill_zero_f32(int64_t n, float *a, const int32_t *done, hipStream_t s);
void launch_fill_zero(int64_t n, double *a, const int32_t *done, hipStream_t s);
void launch_axpy_f32_into_f64(int64_t n, const float *a, double *x, const int32_t *done, hipStream_t s);

// batching form: K right-hand sides as n x K row-major block vectors (mgx_block.hip)
constexpr int kMaxBlock = 32;
// per-column PCG scalars of a block solve; columns that are done are frozen (alpha = beta = 0)
struct BCtl {
    double rz[kMaxBlock], pq[kMaxBlock], alpha[kMaxBlock], beta[kMaxBlock];
    double rnorm2[kMaxBlock], rmax[kMaxBlock], bnorm2[kMaxBlock];
    int32_t iter[kMaxBlock], cdone[kMaxBlock], status[kMaxBlock];
    int32_t frozen[kMaxBlock];   // set by the host: the column takes no part in this solve
    int32_t done;                // every column done: kernels of later iterations return
    int32_t xflushed;            // no x += alpha p pending (set at the start and once done, cleared by each alpha)
    int32_t k;                   // block width
    int32_t stop, max_it;
    double tol;
    double ctol[kMaxBlock];      // per-column tolerance of the stop test (tol unless the host sets it)
};
enum BlockEpi : int { BEPI_SPMM = 0, BEPI_ADD = 1, BEPI_RESID = 2, BEPI_JACOBI = 3, BEPI_DOT = 4,
                      BEPI_UPRESID = 5 };   // node-block only, on an own-and-later-colours copy: residual after a
                           // zero-start forward sweep
// partials of the block reductions: 3 slots x kMaxBlock columns x kPartialStride
constexpr size_t kBlockPartials = (size_t)3 * kMaxBlock * kPartialStride;
template <typename TV, typename TX>
void launch_spmm(int K, int epi, int32_t nrows, const int32_t *rp, const int32_t *ci, const TV *v, const TX *X,
                 const TX *B, const TV *pre, TX *Y, const int32_t *done, double *partials, int stride, hipStream_t s);
// node-triple forms (rows 3q..3q+2 share their column list): nnodes = nrows / 3
template <typename TV, typename TX>
void launch_spmm_node(int K, int epi, int32_t nnodes, const int32_t *rp, const int32_t *ci, const TV *v, const TX *X,
                      const TX *B, const TV *pre, TX *Y, const int32_t *done, double *partials, int stride,
                      hipStream_t s, const int32_t *order = nullptr);
int spmm_node_grid(int K, int32_t nnodes);
int spmm_node_tile(int K);   // nodes per tile of the node-triple SpMM
template <typename T>
void launch_gs_node(int K, int32_t nnodes, const int32_t *nodes, const int32_t *rp, const int32_t *ci, const T *v,
                    const T *B, T *X, bool backward, const int32_t *done, hipStream_t s);
template <typename T>
void launch_gs_block(int K, int32_t nnodes, const int32_t *nodes, const int32_t *rp, const int32_t *ci, const T *v,
                     const T *B, T *X, bool backward, const int32_t *done, hipStream_t s,
                     const int32_t *gptr = nullptr);
template <typename T>
void launch_bnorms(int K, int32_t n, const T *B, const T *R, BCtl *ctl, double *partials, hipStream_t s);
template <typename T>
void launch_bdot_beta(int K, int32_t n, const T *R, const T *Z, BCtl *ctl, double *partials, bool first, hipStream_t s);
template <typename T>
void launch_bupdate(int K, int32_t n, const T *P, const T *Q, T *X, T *R, BCtl *ctl, double *partials, hipStream_t s);
// the iteration's updates split around the preconditioner: r -= alpha q (with the norms and the
// stop test) right after alpha; x += alpha p fused with p = z + beta p after beta
template <typename T>
void launch_bupdate_r(int K, int32_t n, const T *Q, T *R, BCtl *ctl, double *partials, hipStream_t s);
template <typename T>
void launch_bupdate_xp(int K, int32_t n, const T *Z, T *P, T *X, const BCtl *ctl, hipStream_t s);
template <typename T>
void launch_bupdate_p(int K, int32_t n, const T *Z, T *P, const BCtl *ctl, bool first, hipStream_t s);
template <typename T>
void launch_bjacobi_initial(int K, int32_t n, const T *pre, const T *B, T *X, const int32_t *done, hipStream_t s);
// one-block combine of per-column partials (FinMode) and the partial count of a k_spmm launch
void launch_bfinalize(int nparts, const double *partials, BCtl *ctl, int mode, hipStream_t s);
int spmm_grid(int K, int32_t nrows);
template <typename TV, typename TX>
void launch_spmm_dot_alpha(int K, int32_t n, const int32_t *rp, const int32_t *ci, const TV *v, const TX *P, TX *Q,
                           BCtl *ctl, double *partials, hipStream_t s);
// coarse solve Y = M X for K columns; scratch: dense_block_scratch(nc, K) doubles
size_t dense_block_scratch(int32_t nc, int K);
template <typename TX>
void launch_dense_block(int K, int32_t nc, const double *M, const TX *X, TX *Y, double *scratch, const int32_t *done,
                        hipStream_t s);
template <typename T>
void launch_bcopy(int64_t n, const T *src, T *dst, const int32_t *done, hipStream_t s);
void launch_pack_cols(int32_t n, int K, int kc, const double *src, int64_t ld, double *dst, hipStream_t s);
void launch_unpack_cols(int32_t n, int K, int kc, const double *src, double *dst, int64_t ld, hipStream_t s);

// node-block (3x3) layout of node-triple level matrices (mgx_nodeblock.hip); the kernels take
// nodes of at most 27 blocks (nb_count flags larger ones)
struct NbView {
    int32_t npos = 0;          // positions (nodes)
    int64_t nblk = 0;          // 3x3 blocks
    const int32_t *nrp = nullptr, *ncol = nullptr;
    const int32_t *node = nullptr;   // node at each position (nullptr: the identity)
    const int32_t *split = nullptr;  // colour-sorted copies: first block of a later-coloured neighbour
    const void *val = nullptr;       // float or double [9 * nblk] + 16 bytes of slack
};
template <typename T>
int nb_tile(int K);                  // positions per wave tile
template <typename T>
int nb_spmm_grid(int K, int32_t npos);       // bounded grid (waves walk several tiles)
template <typename T>
int nb_spmm_full_grid(int K, int32_t npos);  // one tile per wave

template <typename T>
void launch_nb_spmm(int K, int epi, const NbView &A, const T *X, const T *B, const T *pre, T *Y, const int32_t *done,
                    double *partials, int stride, hipStream_t s, const int32_t *order = nullptr);
// start: NB_GS_GENERAL, NB_GS_ZERO (X is 0 on entry and this is the sweep's first colour: only the
// node's own block is read), NB_GS_LOWER (a zero-start forward sweep past its first colour, on a
// colour-sorted copy: only the blocks of earlier colours are read, and the node's own X is taken
// as 0, so X need not be cleared first)
enum NbGsStart : int { NB_GS_GENERAL = 0, NB_GS_ZERO = 1, NB_GS_LOWER = 2 };
template <typename T>
void launch_nb_gs(int K, const NbView &A, int32_t p0, int32_t p1, const T *B, T *X, bool backward, bool reverse,
                  const int32_t *done, hipStream_t s, int start = NB_GS_GENERAL);
// two colours a, b of a sweep in one launch over nlines mesh lines: rng[4i .. 4i+3] = the
// positions [a0, a1) of colour a and [b0, b1) of colour b on line i; colour b first when
// second_first; lower: the NB_GS_LOWER form for both colours; line_nodes: nodes per mesh line (0:
// unknown), which picks the block size
template <typename T>
void launch_nb_gs_pair(int K, const NbView &A, const int32_t *rng, int32_t nlines, bool second_first, const T *B,
                       T *X, bool backward, bool reverse, bool lower, const int32_t *done, hipStream_t s,
                       int32_t line_nodes = 0);
// colour + upper (with diag_first): count / fill only the node's own block (slot 0) and the blocks
// of later-coloured neighbours, in column order (the copy the zero-start sweep's residual reads)
void nb_count(int32_t npos, const int32_t *node, const int32_t *rp, const int32_t *ci, bool diag_first, int32_t *cnt,
              int32_t *bad, hipStream_t s, const int32_t *colour = nullptr, bool upper = false);
// colour (per node, may be nullptr): sort each position's other blocks by the colour of their
// neighbour node and write split (per position)
template <typename TI, typename TV>
void nb_fill(int32_t npos, const int32_t *node, const int32_t *rp, const int32_t *ci, const TI *v, const int32_t *nrp,
             bool diag_first, int32_t *ncol, TV *val, hipStream_t s, const int32_t *colour = nullptr,
             int32_t *split = nullptr, bool upper = false);

// node transfers: output node i, input node j, one weight for the three dofs of the node
struct NtView {
    int32_t nout = 0, nin = 0;       // output / input nodes
    int64_t nnz = 0;
    const int32_t *rp = nullptr, *ci = nullptr;
    const void *w = nullptr;         // float or double [nnz]
};
template <typename T>
void launch_nt_apply(int K, bool add, const NtView &M, const T *X, T *Y, const int32_t *done, hipStream_t s);
void nt_count(int32_t nout, const int32_t *rp, const int32_t *ci, const double *v, int32_t *cnt, int32_t *bad,
              hipStream_t s);
template <typename TV>
void nt_fill(int32_t nout, const int32_t *rp, const int32_t *ci, const double *v, const int32_t *nrp, int32_t *ncol,
             TV *w, hipStream_t s);

// structured assembly
void launch_assemble_count(int nx, int ny, int nz, int32_t *rowlen, hipStream_t s);
void launch_assemble_fill(int nx, int ny, int nz, const double *Ke, const double *dens, const double *Kel,
                          const int32_t *rowptr, int32_t *col, double *val, hipStream_t s);
void launch_galerkin_elements(int ncx, int ncy, int ncz, const double *Ke0, const double *dens,
                              const double *Kfine, double *Kcoarse, hipStream_t s);
void exclusive_scan_i32(const int32_t *in, int32_t *out, int32_t n, hipStream_t s);

// ------------------------------------------------------------------ host setup (mgx_host.cpp)
void ke_hex8(double E, double nu, double *Ke);
int64_t cantilever_nnz(int nx, int ny, int nz);
struct HostCsr {
    int32_t nrows = 0, ncols = 0;
    std::vector<int32_t> rowptr, col;
    std::vector<double> val;
};
void transfer_reference(int nx, int ny, int nz, HostCsr &R, HostCsr &P);
void transfer_trilinear(int nx, int ny, int nz, HostCsr &R, HostCsr &P);
void csr_transpose(const HostCsr &A, HostCsr &T);
// the 2D workloads (mgx_fem2d.cpp): the Kirchhoff plate and the plane-stress cantilever
void plate_element(double a, double b, double E, double nu, double t, double Ke[144]);
int32_t plate_free_dofs(int nx, int ny, double Lx, double Ly, std::vector<int32_t> *dofs);
void plate_system(int nx, int ny, double Lx, double Ly, double t, double E, double nu, double q, HostCsr &K,
                  std::vector<double> *rhs);
void plate_transfer(int nx, int ny, double Lx, double Ly, HostCsr &R, HostCsr &P);
void cant2d_element(double E, double nu, double Ke[64]);
void cant2d_system(int nx, int ny, double E, double nu, const double *dens, HostCsr &K);
void cant2d_transfer(int nx, int ny, HostCsr &R, HostCsr &P);
int block_coloring(int32_t nblocks, int bs, const int32_t *rp, const int32_t *ci, std::vector<int32_t> &color);
int32_t row_groups(int32_t n, const int32_t *rp, const int32_t *ci, std::vector<int32_t> &gptr);
int group_coloring(int32_t ngroups, const int32_t *gptr, const int32_t *rp, const int32_t *ci,
                   std::vector<int32_t> &color);
// dense SPD inverse via Cholesky; returns false if not SPD
bool dense_spd_inverse(int32_t n, std::vector<double> &A);
int duvi_encode(int32_t nrows, const int32_t *rp, const int32_t *ci, const double *v, double tol, int mode,
                mgx_duvi_host *out);
void rng_uniform(uint64_t seed, int64_t n, double *out);
void rng_normal(uint64_t seed, int64_t n, double *out);

}  // namespace mgx
