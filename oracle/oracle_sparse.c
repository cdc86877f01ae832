/*
 * oracle_sparse.c -- CPU restatement of the reference's sparse kernels, smoothers,
 * Jacobi weights, dense coarse solve and the Galerkin triple product.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates:
 *   CsrGeometricMultigrid.cl:16-62     matrix_vector_product (y[row] accumulated in CSR order)
 *   CsrGeometricMultigrid.cl:75-141    residual, residual_with_check (flag bits)
 *   CsrGeometricMultigrid.cl:153-194   jacobi_initial_iteration, jacobi_iteration
 *   CsrCG.cl:9-85                      dot products (sequential here)
 *   GeometricMultigridSolver.cs:128-151 JacobiPreconditioner, EigenValueUpperBound
 *   GaussSeidelIterationCsrDuVi.cs:22-50 forward Gauss-Seidel (textbook order)
 *   GeometricMultigridSolver.cs:255,337 LDL^T coarse solve (Cholesky here: same factorisation
 *                                        of an SPD matrix up to rounding)
 *
 * Pinned by (tests/test_oracle_sparse.py): spmv/residual/jacobi against scipy.sparse on
 * random CSR; the OpenCL diagonal KAT y_i = v_i^2 (GMCantileverOpenCLTests.cs:16-52);
 * the flag semantics on planted NaN / overflow / tolerance cases; GS against the closed-form
 * lower-triangular solve x = (D+L)^-1 (b - U x); multicolour GS against GS on the
 * symmetrically permuted matrix; colouring validity; Cholesky against numpy.linalg.solve;
 * Galerkin against dense P^T A P.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc_csr_spmv(int32_t nrows, const int32_t *rp, const int32_t *ci, const double *v,
                  const double *x, double *y, int clear)
{
    ORC_PARALLEL_FOR
    for (int32_t r = 0; r < nrows; ++r) {
        double acc = clear ? 0.0 : y[r];
        for (int32_t k = rp[r]; k < rp[r + 1]; ++k) acc += v[k] * x[ci[k]];
        y[r] = acc;
    }
}

void orc_csr_residual(int32_t nrows, const int32_t *rp, const int32_t *ci, const double *v,
                      const double *b, const double *x, double *r)
{
    ORC_PARALLEL_FOR
    for (int32_t i = 0; i < nrows; ++i) {
        double acc = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) acc += v[k] * x[ci[k]];
        r[i] = b[i] - acc;
    }
}

/* returns the reference flag: starts at 1; |= 2 on NaN or > 1e50; &= 2 when > tol. */
int orc_csr_residual_check(int32_t nrows, const int32_t *rp, const int32_t *ci, const double *v,
                           const double *b, const double *x, double *r, double tol)
{
    int flag = 1;
    orc_csr_residual(nrows, rp, ci, v, b, x, r);
    for (int32_t i = 0; i < nrows; ++i) {
        double c = fabs(r[i]);
        if (isnan(r[i]) || c > 1e50) flag |= 2;
        else if (c > tol) flag &= 2;
    }
    return flag;
}

void orc_jacobi_initial(int32_t n, const double *pre, const double *b, double *x)
{
    ORC_PARALLEL_FOR
    for (int32_t i = 0; i < n; ++i) x[i] = pre[i] * b[i];
}

void orc_csr_jacobi(int32_t n, const double *pre, const int32_t *rp, const int32_t *ci, const double *v,
                    const double *b, const double *x, double *y)
{
    ORC_PARALLEL_FOR
    for (int32_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) acc += v[k] * x[ci[k]];
        double t = b[i] - acc;
        y[i] = x[i] + pre[i] * t;
    }
}

void orc_jacobi_preconditioner(int32_t n, const int32_t *rp, const int32_t *ci, const double *v, double *d)
{
    for (int32_t i = 0; i < n; ++i) {
        double diag = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) if (ci[k] == i) diag = v[k];
        d[i] = 1 / diag;
    }
}

double orc_eigen_upper_bound(int32_t n, const int32_t *rp, const double *v)
{
    double l = 0;
    for (int32_t i = 0; i < n; ++i) {
        double s = 0, m = 0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) { double a = fabs(v[k]); s += a; if (a > m) m = a; }
        double q = s / m;
        if (q > l) l = q;
    }
    return l;
}

/* EigenValueUpperBound (power method), GeometricMultigridSolver.cs:80-103:
   x = 1/sqrt(n); repeat y = D^-1 A x, l2 = ||y||, x = y / l2, ratio = l2 / l until
   0.95 < ratio < 1.05, then l = ratio > 1 ? l2 ratio^(1/sqrt(0.05)) : (9 l2 + l) / 10. */
double orc_power_bound(int32_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *invd)
{
    const double percent = 0.05;
    double *x = (double *)malloc(sizeof(double) * n), *y = (double *)malloc(sizeof(double) * n);
    for (int32_t i = 0; i < n; ++i) x[i] = 1 / sqrt((double)n);
    double l = 1;
    for (int it = 0; it < 100000; ++it) {
        for (int32_t i = 0; i < n; ++i) {
            double s = 0;
            for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += v[k] * x[ci[k]];
            y[i] = s * invd[i];
        }
        double l2 = 0;
        for (int32_t i = 0; i < n; ++i) l2 += y[i] * y[i];
        l2 = sqrt(l2);
        for (int32_t i = 0; i < n; ++i) x[i] = y[i] * (1 / l2);
        double ratio = l2 / l;
        if (ratio > 1 - percent && ratio < 1 + percent) {
            l = ratio > 1 ? l2 * pow(ratio, 1 / sqrt(percent)) : (l2 * 9 + l) / 10;
            break;
        }
        l = l2;
    }
    free(x); free(y);
    return l;
}

double orc_dot(int32_t n, const double *a, const double *b)
{
    double s = 0.0;
    for (int32_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static inline void gs_row(int32_t i, const int32_t *rp, const int32_t *ci, const double *v,
                          const double *b, double *x)
{
    double xi = b[i], div = 0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        if (ci[k] != i) xi -= v[k] * x[ci[k]];
        else div = v[k];
    }
    x[i] = xi / div;
}

void orc_gs_forward(int32_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *b, double *x)
{
    for (int32_t i = 0; i < n; ++i) gs_row(i, rp, ci, v, b, x);
}

int orc_block_coloring(int32_t nblocks, int bs, const int32_t *rp, const int32_t *ci, int32_t *color)
{
    int ncol = 0;
    uint64_t used[4];
    for (int32_t k = 0; k < nblocks; ++k) color[k] = -1;
    for (int32_t k = 0; k < nblocks; ++k) {
        memset(used, 0, sizeof used);
        for (int a = 0; a < bs; ++a) {
            int32_t r = k * bs + a;
            for (int32_t q = rp[r]; q < rp[r + 1]; ++q) {
                int32_t nb = ci[q] / bs;
                if (nb != k && color[nb] >= 0 && color[nb] < 256) used[color[nb] >> 6] |= 1ULL << (color[nb] & 63);
            }
        }
        int c = 0;
        while (c < 256 && (used[c >> 6] >> (c & 63) & 1)) ++c;
        color[k] = c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    return ncol;
}

/* Any fixed node order gives a Gauss-Seidel sweep (each node relaxed once with the latest values
   of its neighbours); the multicolour sweep above is the order colour 0, colour 1, ...  A
   parallel sweep that relaxes pairwise uncoupled nodes together equals the sequential sweep in
   any order that keeps its groups in sequence. */
void orc_gs_ordered(int32_t n, int bs, const int32_t *rp, const int32_t *ci, const double *v,
                    const int32_t *order, const double *b, double *x, int backward)
{
    int32_t nb = n / bs;
    if (!backward) {
        for (int32_t i = 0; i < nb; ++i)
            for (int a = 0; a < bs; ++a) gs_row(order[i] * bs + a, rp, ci, v, b, x);
    } else {
        for (int32_t i = nb - 1; i >= 0; --i)
            for (int a = bs - 1; a >= 0; --a) gs_row(order[i] * bs + a, rp, ci, v, b, x);
    }
}

void orc_gs_multicolor(int32_t n, int bs, const int32_t *rp, const int32_t *ci, const double *v,
                       int ncolors, const int32_t *color, const double *b, double *x, int backward)
{
    int32_t nb = n / bs;
    if (!backward) {
        for (int c = 0; c < ncolors; ++c) {
            ORC_PARALLEL_FOR
            for (int32_t k = 0; k < nb; ++k)
                if (color[k] == c)
                    for (int a = 0; a < bs; ++a) gs_row(k * bs + a, rp, ci, v, b, x);
        }
    } else {
        for (int c = ncolors - 1; c >= 0; --c) {
            ORC_PARALLEL_FOR
            for (int32_t k = nb - 1; k >= 0; --k)
                if (color[k] == c)
                    for (int a = bs - 1; a >= 0; --a) gs_row(k * bs + a, rp, ci, v, b, x);
        }
    }
}

/* ---------------------------------------------------------------- node groups
 * Multicolour Gauss-Seidel by mesh node when nodes carry different numbers of free dofs (the
 * plate: 2 at supported nodes, 3 elsewhere).  A group is a maximal run of at most 3 consecutive
 * rows with identical column lists -- the free dofs of one node, which couple to the same
 * neighbours; on the cantilever these are the node triples of orc_block_coloring. */
int32_t orc_row_groups(int32_t n, const int32_t *rp, const int32_t *ci, int32_t *gptr)
{
    int32_t ng = 0, r = 0;
    while (r < n) {
        gptr[ng++] = r;
        int32_t e = r + 1;
        while (e < n && e - r < 3 && rp[e + 1] - rp[e] == rp[r + 1] - rp[r] &&
               memcmp(ci + rp[e], ci + rp[r], sizeof(int32_t) * (size_t)(rp[r + 1] - rp[r])) == 0)
            ++e;
        r = e;
    }
    gptr[ng] = n;
    return ng;
}

/* greedy distance-1 colouring of the groups in natural order (orc_block_coloring with groups) */
int orc_group_coloring(int32_t ngroups, const int32_t *gptr, const int32_t *rp, const int32_t *ci, int32_t *color)
{
    const int32_t n = gptr[ngroups];
    int32_t *grp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int32_t g = 0; g < ngroups; ++g)
        for (int32_t r = gptr[g]; r < gptr[g + 1]; ++r) grp[r] = g;
    int ncol = 0;
    uint64_t used[4];
    for (int32_t k = 0; k < ngroups; ++k) color[k] = -1;
    for (int32_t k = 0; k < ngroups; ++k) {
        memset(used, 0, sizeof used);
        for (int32_t r = gptr[k]; r < gptr[k + 1]; ++r)
            for (int32_t q = rp[r]; q < rp[r + 1]; ++q) {
                const int32_t nb = grp[ci[q]];
                if (nb != k && color[nb] >= 0 && color[nb] < 256) used[color[nb] >> 6] |= 1ULL << (color[nb] & 63);
            }
        int c = 0;
        while (c < 256 && (used[c >> 6] >> (c & 63) & 1)) ++c;
        color[k] = c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    free(grp);
    return ncol;
}

/* Gauss-Seidel sweep over the groups: in `order` (group ids) when given, else colour by colour
 * (groups ascending within a colour); a group's rows in order, all reversed when backward */
void orc_gs_groups(int32_t ngroups, const int32_t *gptr, const int32_t *rp, const int32_t *ci, const double *v,
                   const int32_t *order, int ncolors, const int32_t *color, const double *b, double *x,
                   int backward)
{
    if (order) {
        if (!backward) {
            for (int32_t i = 0; i < ngroups; ++i)
                for (int32_t r = gptr[order[i]]; r < gptr[order[i] + 1]; ++r) gs_row(r, rp, ci, v, b, x);
        } else {
            for (int32_t i = ngroups - 1; i >= 0; --i)
                for (int32_t r = gptr[order[i] + 1] - 1; r >= gptr[order[i]]; --r) gs_row(r, rp, ci, v, b, x);
        }
        return;
    }
    if (!backward) {
        for (int c = 0; c < ncolors; ++c)
            for (int32_t k = 0; k < ngroups; ++k)
                if (color[k] == c)
                    for (int32_t r = gptr[k]; r < gptr[k + 1]; ++r) gs_row(r, rp, ci, v, b, x);
    } else {
        for (int c = ncolors - 1; c >= 0; --c)
            for (int32_t k = ngroups - 1; k >= 0; --k)
                if (color[k] == c)
                    for (int32_t r = gptr[k + 1] - 1; r >= gptr[k]; --r) gs_row(r, rp, ci, v, b, x);
    }
}

/* ---------------------------------------------------------------- dense coarse solve */
void orc_csr_to_dense(int32_t n, const int32_t *rp, const int32_t *ci, const double *v, double *A)
{
    memset(A, 0, sizeof(double) * (size_t)n * n);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) A[(size_t)i * n + ci[k]] = v[k];
}

int orc_dense_cholesky(int32_t n, double *A)
{
    for (int32_t j = 0; j < n; ++j) {
        double d = A[(size_t)j * n + j];
        for (int32_t k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(d > 0)) return -1;
        d = sqrt(d);
        A[(size_t)j * n + j] = d;
        for (int32_t i = j + 1; i < n; ++i) {
            double s = A[(size_t)i * n + j];
            for (int32_t k = 0; k < j; ++k) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = s / d;
        }
        for (int32_t i = 0; i < j; ++i) A[(size_t)i * n + j] = 0;
    }
    return 0;
}

void orc_dense_cholesky_solve(int32_t n, const double *L, const double *b, double *x)
{
    double *y = (double *)malloc(sizeof(double) * n);
    for (int32_t i = 0; i < n; ++i) {
        double s = b[i];
        for (int32_t k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * y[k];
        y[i] = s / L[(size_t)i * n + i];
    }
    for (int32_t i = n - 1; i >= 0; --i) {
        double s = y[i];
        for (int32_t k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * x[k];
        x[i] = s / L[(size_t)i * n + i];
    }
    free(y);
}

/* ---------------------------------------------------------------- CSR transpose / Galerkin */
void orc_csr_transpose(int32_t nrows, int32_t ncols, const int32_t *rp, const int32_t *ci, const double *v,
                       int32_t *rpT, int32_t *ciT, double *vT)
{
    memset(rpT, 0, sizeof(int32_t) * (ncols + 1));
    for (int32_t k = 0; k < rp[nrows]; ++k) rpT[ci[k] + 1]++;
    for (int32_t c = 0; c < ncols; ++c) rpT[c + 1] += rpT[c];
    int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (ncols + 1));
    memcpy(fill, rpT, sizeof(int32_t) * (ncols + 1));
    for (int32_t r = 0; r < nrows; ++r)
        for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
            int32_t p = fill[ci[k]]++;
            ciT[p] = r; vT[p] = v[k];
        }
    free(fill);
}

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* C = P^T (A P): AP row by row with a dense accumulator, then P^T * AP. */
typedef struct { int32_t *rp, *ci; double *v; } csr_t;

static void spgemm(int32_t nr, int32_t nc, const int32_t *Arp, const int32_t *Aci, const double *Av,
                   const int32_t *Brp, const int32_t *Bci, const double *Bv, csr_t *C, int values)
{
    int32_t *mark = (int32_t *)malloc(sizeof(int32_t) * nc);
    double *acc = (double *)calloc(nc, sizeof(double));
    int32_t *list = (int32_t *)malloc(sizeof(int32_t) * nc);
    for (int32_t j = 0; j < nc; ++j) mark[j] = -1;
    int64_t nnz = 0;
    C->rp[0] = 0;
    for (int32_t i = 0; i < nr; ++i) {
        int32_t cnt = 0;
        for (int32_t k = Arp[i]; k < Arp[i + 1]; ++k) {
            int32_t a = Aci[k];
            for (int32_t q = Brp[a]; q < Brp[a + 1]; ++q) {
                int32_t j = Bci[q];
                if (mark[j] != i) { mark[j] = i; list[cnt++] = j; acc[j] = 0.0; }
                if (values) acc[j] += Av[k] * Bv[q];
            }
        }
        qsort(list, cnt, sizeof(int32_t), cmp_i32);
        if (values)
            for (int32_t t = 0; t < cnt; ++t) { C->ci[nnz + t] = list[t]; C->v[nnz + t] = acc[list[t]]; }
        nnz += cnt;
        C->rp[i + 1] = (int32_t)nnz;
    }
    free(mark); free(acc); free(list);
}

static void galerkin_impl(int32_t nf, int32_t nc, const int32_t *Arp, const int32_t *Aci, const double *Av,
                          const int32_t *Prp, const int32_t *Pci, const double *Pv,
                          int32_t *Crp, int32_t *Cci, double *Cv, int values)
{
    /* AP */
    csr_t AP; AP.rp = (int32_t *)malloc(sizeof(int32_t) * (nf + 1));
    AP.ci = NULL; AP.v = NULL;
    spgemm(nf, nc, Arp, Aci, NULL, Prp, Pci, NULL, &AP, 0);
    int64_t nap = AP.rp[nf];
    AP.ci = (int32_t *)malloc(sizeof(int32_t) * nap);
    AP.v = (double *)malloc(sizeof(double) * nap);
    spgemm(nf, nc, Arp, Aci, Av ? Av : NULL, Prp, Pci, Pv, &AP, 1);
    /* P^T */
    int64_t np = Prp[nf];
    int32_t *PTrp = (int32_t *)malloc(sizeof(int32_t) * (nc + 1));
    int32_t *PTci = (int32_t *)malloc(sizeof(int32_t) * np);
    double *PTv = (double *)malloc(sizeof(double) * np);
    orc_csr_transpose(nf, nc, Prp, Pci, Pv, PTrp, PTci, PTv);
    csr_t C = {Crp, Cci, Cv};
    spgemm(nc, nc, PTrp, PTci, PTv, AP.rp, AP.ci, AP.v, &C, values);
    free(AP.rp); free(AP.ci); free(AP.v); free(PTrp); free(PTci); free(PTv);
}

int64_t orc_galerkin_nnz(int32_t nf, int32_t nc, const int32_t *Arp, const int32_t *Aci,
                         const int32_t *Prp, const int32_t *Pci)
{
    int64_t np = Prp[nf], na = Arp[nf];
    double *ones_a = (double *)malloc(sizeof(double) * na);
    double *ones_p = (double *)malloc(sizeof(double) * np);
    for (int64_t i = 0; i < na; ++i) ones_a[i] = 1;
    for (int64_t i = 0; i < np; ++i) ones_p[i] = 1;
    int32_t *Crp = (int32_t *)malloc(sizeof(int32_t) * (nc + 1));
    galerkin_impl(nf, nc, Arp, Aci, ones_a, Prp, Pci, ones_p, Crp, NULL, NULL, 0);
    int64_t nnz = Crp[nc];
    free(Crp); free(ones_a); free(ones_p);
    return nnz;
}

void orc_galerkin(int32_t nf, int32_t nc, const int32_t *Arp, const int32_t *Aci, const double *Av,
                  const int32_t *Prp, const int32_t *Pci, const double *Pv,
                  int32_t *Crp, int32_t *Cci, double *Cv)
{
    galerkin_impl(nf, nc, Arp, Aci, Av, Prp, Pci, Pv, Crp, Cci, Cv, 1);
}
