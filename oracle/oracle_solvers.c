/*
 * oracle_solvers.c -- CPU restatement of the solver drivers.  TEST INFRASTRUCTURE ONLY.
 *
 * Restates:
 *   GeometricMultigridSolver.cs:56-64,178-201   CreateDeepV path: levelDown = [T x depth, F x depth]
 *   OpenCLCsrGeometricMultigridSolver.cs:334-536 GPU V-cycle semantics (zero-init rules :390-392,
 *     Jacobi ping-pong through one shared buffer :421-452, residual_with_check + flag :458-489,
 *     restriction clear=1 :508-515, coarse solve then prolongation clear=0 :357-382, :519-526)
 *   OpenCLCsrPcgSolver.cs:250-331               Jacobi PCG (iteration count = 0-based index)
 *   The multicolour Gauss-Seidel smoother replaces the reference's racy in-place weighted
 *   Jacobi (CsrGeometricMultigrid.cl:207-225) -- the documented ordering deviation.
 *   orc_mgpcg is a new composition (PCG of OpenCLCsrPcgSolver with M = one symmetric V-cycle).
 *
 * Pinned by (tests/test_oracle_solvers.py): the manufactured exact discrete solution
 * (FemCantileverBase.cs:53-62) -- every converged solve must reproduce it to the stated
 * energy-norm tolerance; scipy spsolve on the 16^3 system; iteration-count invariants
 * (PCG on an n-distinct-eigenvalue diagonal system converges in n steps); the V-cycle as a
 * preconditioner must be symmetric (u^T M v == v^T M u) for the MG-PCG configuration.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static double now_seconds(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static double *dalloc(int64_t n) { return (double *)calloc(n > 0 ? n : 1, sizeof(double)); }

static double *coarse_factor(const orc_level *c)
{
    double *L = dalloc((int64_t)c->n * c->n);
    orc_csr_to_dense(c->n, c->Arp, c->Aci, c->Av, L);
    if (orc_dense_cholesky(c->n, L) != 0) { free(L); return NULL; }
    return L;
}

/* Gauss-Seidel node sets of a level: its groups when given (nodes of 2 or 3 free dofs), else
   the dof triples */
static int level_nodes(const orc_level *q) { return q->gptr ? q->ngroups : q->n / 3; }

static int level_coloring(const orc_level *q, int32_t *color)
{
    if (q->gptr) return orc_group_coloring(q->ngroups, q->gptr, q->Arp, q->Aci, color);
    return orc_block_coloring(q->n / 3, 3, q->Arp, q->Aci, color);
}

static void level_gs(const orc_level *q, int ncol, const int32_t *color, const double *b, double *x, int backward)
{
    if (q->gptr) orc_gs_groups(q->ngroups, q->gptr, q->Arp, q->Aci, q->Av, q->gs_order, ncol, color, b, x, backward);
    else if (q->gs_order) orc_gs_ordered(q->n, 3, q->Arp, q->Aci, q->Av, q->gs_order, b, x, backward);
    else orc_gs_multicolor(q->n, 3, q->Arp, q->Aci, q->Av, ncol, color, b, x, backward);
}

int orc_gmg_solve(int nlevels, const orc_level *lv, const double *const *pre, int smoother,
                  int depth, int level_iterations, int max_cycles, double tol,
                  const double *b, double *x0, int has_x0, orc_stats *st)
{
    const int L = nlevels;
    if (L != depth + 1 || L < 2) return -1;
    const int nsteps = 2 * depth;
    double *L_coarse = coarse_factor(&lv[L - 1]);
    if (!L_coarse) return -2;
    double **X = (double **)calloc(L, sizeof(double *));
    double **B = (double **)calloc(L, sizeof(double *));
    int32_t **color = (int32_t **)calloc(L, sizeof(int32_t *));
    int *ncol = (int *)calloc(L, sizeof(int));
    X[0] = x0; B[0] = (double *)b;
    for (int l = 1; l < L; ++l) { B[l] = dalloc(lv[l].n); if (l < L - 1) X[l] = dalloc(lv[l].n); }
    if (smoother == 1)
        for (int l = 0; l < L - 1; ++l) {
            color[l] = (int32_t *)malloc(sizeof(int32_t) * (size_t)(level_nodes(&lv[l]) + 1));
            ncol[l] = level_coloring(&lv[l], color[l]);
        }
    double *R = dalloc(lv[0].n);
    double *tmp = dalloc(lv[L - 1].n);
    int x_given = has_x0;
    int rc = 0;
    for (int cycle = 0;; ++cycle) {
        int cur = 0;
        for (int step = 0; step < nsteps; ++step) {
            int down = step < depth;
            if (cur == L - 1) {
                orc_dense_cholesky_solve(lv[cur].n, L_coarse, B[cur], tmp);
                memcpy(B[cur], tmp, sizeof(double) * lv[cur].n);
                const orc_level *f = &lv[cur - 1];
                orc_csr_spmv(f->n, f->Prp, f->Pci, f->Pv, B[cur], X[cur - 1], 0);
            } else {
                const orc_level *q = &lv[cur];
                int ld = step == 0 ? 0 /* levelDown.Last() is false */ : (step - 1 < depth);
                int reset = (cur == 0 && !x_given) || (cur != 0 && ld);
                int it = level_iterations;
                if (smoother == 1) {
                    if (reset) memset(X[cur], 0, sizeof(double) * q->n);
                    for (int i = 0; i < it; ++i) level_gs(q, ncol[cur], color[cur], B[cur], X[cur], 0);
                } else {
                    int i;
                    if (reset) {
                        orc_jacobi_initial(q->n, pre[cur], B[cur], R);
                        orc_csr_jacobi(q->n, pre[cur], q->Arp, q->Aci, q->Av, B[cur], R, X[cur]);
                        i = 2;
                    } else i = 0;
                    for (; i < it; i += 2) {
                        orc_csr_jacobi(q->n, pre[cur], q->Arp, q->Aci, q->Av, B[cur], X[cur], R);
                        orc_csr_jacobi(q->n, pre[cur], q->Arp, q->Aci, q->Av, B[cur], R, X[cur]);
                    }
                }
                if (down) {
                    if (cur == 0) {
                        int flag = orc_csr_residual_check(q->n, q->Arp, q->Aci, q->Av, B[0], X[0], R, tol);
                        x_given = 1;
                        int converged = flag == 1, failed = (flag & 2) == 2;
                        if (converged || failed || cycle > max_cycles) {
                            double rn = sqrt(orc_dot(q->n, R, R)), bn = sqrt(orc_dot(q->n, B[0], B[0]));
                            st->iterations = cycle;
                            st->converged = failed ? -1 : converged;
                            st->residual_norm = rn;
                            st->rel_residual = bn > 0 ? rn / bn : rn;
                            goto done;
                        }
                    } else orc_csr_residual(q->n, q->Arp, q->Aci, q->Av, B[cur], X[cur], R);
                    orc_csr_spmv(lv[cur + 1].n, q->Rrp, q->Rci, q->Rv, R, B[cur + 1], 1);
                } else {
                    const orc_level *f = &lv[cur - 1];
                    orc_csr_spmv(f->n, f->Prp, f->Pci, f->Pv, X[cur], X[cur - 1], 0);
                }
            }
            cur += down ? 1 : -1;
        }
    }
done:
    for (int l = 1; l < L; ++l) { free(B[l]); free(X[l]); }
    for (int l = 0; l < L; ++l) free(color[l]);
    free(X); free(B); free(color); free(ncol); free(R); free(tmp); free(L_coarse);
    return rc;
}

int orc_pcg_jacobi(int32_t n, const int32_t *rp, const int32_t *ci, const double *v,
                   const double *b, double *x, int has_x0, int max_it, double tol, orc_stats *st)
{
    double *M = dalloc(n), *r = dalloc(n), *p = dalloc(n), *z = dalloc(n), *Ap = dalloc(n);
    orc_jacobi_preconditioner(n, rp, ci, v, M);
    memcpy(r, b, sizeof(double) * n);
    if (has_x0) {                                  /* CsrCG.cl:191-211 */
        for (int32_t i = 0; i < n; ++i) {
            double acc = -r[i];
            for (int32_t k = rp[i]; k < rp[i + 1]; ++k) acc += v[k] * x[ci[k]];
            r[i] = -acc;
        }
    } else memset(x, 0, sizeof(double) * n);
    for (int32_t i = 0; i < n; ++i) p[i] = z[i] = M[i] * r[i];
    double rz = orc_dot(n, r, z);
    for (int it = 0;; ++it) {
        orc_csr_spmv(n, rp, ci, v, p, Ap, 1);
        double pAp = orc_dot(n, p, Ap);
        double a = rz / pAp;
        int flag = 1;
        for (int32_t i = 0; i < n; ++i) {
            x[i] += a * p[i];
            r[i] -= a * Ap[i];
            z[i] = M[i] * r[i];
            double c = fabs(r[i]);
            if (isnan(r[i]) || c > 1e50) flag |= 2;
            else if (c > tol) flag &= 2;
        }
        int converged = flag == 1, failed = (flag & 2) == 2;
        if (converged || failed || it > max_it) {
            double rn = sqrt(orc_dot(n, r, r)), bn = sqrt(orc_dot(n, b, b));
            st->iterations = it; st->converged = failed ? -1 : converged;
            st->residual_norm = rn; st->rel_residual = bn > 0 ? rn / bn : rn;
            break;
        }
        double rz2 = orc_dot(n, r, z);
        double beta = rz2 / rz;
        for (int32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        rz = rz2;
    }
    free(M); free(r); free(p); free(z); free(Ap);
    return 0;
}

/* ---------------------------------------------------------------- MG-PCG */
typedef struct {
    int L, smoother, nu;
    const orc_level *lv;
    double **pre;          /* omega * D^-1 per level (Jacobi) */
    int32_t **color; int *ncol;
    double *Lc;            /* coarse Cholesky factor */
    double **x, **b, **r;  /* per-level work */
} vcyc;

static void vcycle(vcyc *V, int l)
{
    const orc_level *q = &V->lv[l];
    double *x = V->x[l], *b = V->b[l], *r = V->r[l];
    if (l == V->L - 1) { orc_dense_cholesky_solve(q->n, V->Lc, b, x); return; }
    memset(x, 0, sizeof(double) * q->n);
    for (int s = 0; s < V->nu; ++s) {
        if (V->smoother == 1) level_gs(q, V->ncol[l], V->color[l], b, x, 0);
        else if (s == 0) orc_jacobi_initial(q->n, V->pre[l], b, x);
        else { orc_csr_jacobi(q->n, V->pre[l], q->Arp, q->Aci, q->Av, b, x, r); memcpy(x, r, sizeof(double) * q->n); }
    }
    orc_csr_residual(q->n, q->Arp, q->Aci, q->Av, b, x, r);
    orc_csr_spmv(V->lv[l + 1].n, q->Rrp, q->Rci, q->Rv, r, V->b[l + 1], 1);
    vcycle(V, l + 1);
    orc_csr_spmv(q->n, q->Prp, q->Pci, q->Pv, V->x[l + 1], x, 0);
    for (int s = 0; s < V->nu; ++s) {
        if (V->smoother == 1) level_gs(q, V->ncol[l], V->color[l], b, x, 1);
        else { orc_csr_jacobi(q->n, V->pre[l], q->Arp, q->Aci, q->Av, b, x, r); memcpy(x, r, sizeof(double) * q->n); }
    }
}

static int vcyc_setup(vcyc *V, int nlevels, const orc_level *lv, int smoother, double omega, int nu);
static void vcyc_free(vcyc *V);

/* z = M r for the MG-PCG preconditioner (one symmetric V-cycle). */
int orc_vcycle_apply(int nlevels, const orc_level *lv, int smoother, double omega, int nu,
                     const double *r, double *z)
{
    vcyc V;
    if (vcyc_setup(&V, nlevels, lv, smoother, omega, nu) != 0) return -2;
    memcpy(V.b[0], r, sizeof(double) * lv[0].n);
    vcycle(&V, 0);
    memcpy(z, V.x[0], sizeof(double) * lv[0].n);
    vcyc_free(&V);
    return 0;
}

static int vcyc_setup(vcyc *Vp, int nlevels, const orc_level *lv, int smoother, double omega, int nu)
{
    vcyc V; memset(&V, 0, sizeof V);
    V.L = nlevels; V.smoother = smoother; V.nu = nu; V.lv = lv;
    V.pre = (double **)calloc(nlevels, sizeof(double *));
    V.color = (int32_t **)calloc(nlevels, sizeof(int32_t *));
    V.ncol = (int *)calloc(nlevels, sizeof(int));
    V.x = (double **)calloc(nlevels, sizeof(double *));
    V.b = (double **)calloc(nlevels, sizeof(double *));
    V.r = (double **)calloc(nlevels, sizeof(double *));
    for (int l = 0; l < nlevels; ++l) {
        V.x[l] = dalloc(lv[l].n); V.b[l] = dalloc(lv[l].n); V.r[l] = dalloc(lv[l].n);
        if (l < nlevels - 1) {
            if (smoother == 1) {
                V.color[l] = (int32_t *)malloc(sizeof(int32_t) * (size_t)(level_nodes(&lv[l]) + 1));
                V.ncol[l] = level_coloring(&lv[l], V.color[l]);
            } else {
                V.pre[l] = dalloc(lv[l].n);
                orc_jacobi_preconditioner(lv[l].n, lv[l].Arp, lv[l].Aci, lv[l].Av, V.pre[l]);
                double w = omega;
                if (omega <= 0) w = 2 / (1.2 * orc_power_bound(lv[l].n, lv[l].Arp, lv[l].Aci, lv[l].Av, V.pre[l]));
                for (int32_t i = 0; i < lv[l].n; ++i) V.pre[l][i] *= w;
            }
        }
    }
    V.Lc = coarse_factor(&lv[nlevels - 1]);
    *Vp = V;
    return V.Lc ? 0 : -2;
}

static void vcyc_free(vcyc *Vp)
{
    vcyc V = *Vp;
    for (int l = 0; l < V.L; ++l) { free(V.x[l]); free(V.b[l]); free(V.r[l]); free(V.pre[l]); free(V.color[l]); }
    free(V.x); free(V.b); free(V.r); free(V.pre); free(V.color); free(V.ncol); free(V.Lc);
}

int orc_mgpcg(int nlevels, const orc_level *lv, int smoother, double omega, int nu,
              int max_it, double tol, int stop, const double *bvec, double *x, orc_stats *st)
{
    return orc_mgpcg_timed(nlevels, lv, smoother, omega, nu, max_it, tol, stop, bvec, x, st, NULL);
}

int orc_mgpcg_timed(int nlevels, const orc_level *lv, int smoother, double omega, int nu,
                    int max_it, double tol, int stop, const double *bvec, double *x, orc_stats *st,
                    double *iter_seconds)
{
    vcyc V;
    if (vcyc_setup(&V, nlevels, lv, smoother, omega, nu) != 0) { vcyc_free(&V); return -2; }
    const int32_t n = lv[0].n;
    const int32_t *rp = lv[0].Arp, *ci = lv[0].Aci; const double *v = lv[0].Av;
    double *r = dalloc(n), *p = dalloc(n), *q = dalloc(n);
    memcpy(r, bvec, sizeof(double) * n);
    memset(x, 0, sizeof(double) * n);
    double bn = sqrt(orc_dot(n, bvec, bvec));
    int it = 0, conv = 0;
    double rn = bn;
    if (bn == 0) conv = 1;
    if (!conv) {
        memcpy(V.b[0], r, sizeof(double) * n);
        vcycle(&V, 0);
        memcpy(p, V.x[0], sizeof(double) * n);
        double rz = orc_dot(n, r, V.x[0]);
        for (it = 1; it <= max_it; ++it) {
            const double t0 = iter_seconds ? now_seconds() : 0.0;
            orc_csr_spmv(n, rp, ci, v, p, q, 1);
            double pq = orc_dot(n, p, q);
            double alpha = rz / pq;
            double mx = 0; int bad = 0;
            /* NaN never wins a max comparison, so the maximum is the same in any order */
            #ifdef _OPENMP
            #pragma omp parallel for schedule(static) reduction(max : mx) reduction(| : bad)
            #endif
            for (int32_t i = 0; i < n; ++i) {
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
                double c = fabs(r[i]);
                if (isnan(r[i]) || c > 1e50) bad = 1;
                if (c > mx) mx = c;
            }
            rn = sqrt(orc_dot(n, r, r));
            if (bad) { conv = -1; break; }
            if ((stop == 1 && rn <= tol * bn) || (stop == 0 && mx <= tol)) {
                conv = 1;
                if (iter_seconds) iter_seconds[it - 1] = now_seconds() - t0;
                break;
            }
            memcpy(V.b[0], r, sizeof(double) * n);
            vcycle(&V, 0);
            double rz2 = orc_dot(n, r, V.x[0]);
            double beta = rz2 / rz;
            ORC_PARALLEL_FOR
            for (int32_t i = 0; i < n; ++i) p[i] = V.x[0][i] + beta * p[i];
            rz = rz2;
            if (iter_seconds) iter_seconds[it - 1] = now_seconds() - t0;
        }
        if (it > max_it) it = max_it;
    }
    st->iterations = it; st->converged = conv; st->residual_norm = rn;
    st->rel_residual = bn > 0 ? rn / bn : rn;
    vcyc_free(&V);
    free(r); free(p); free(q);
    return 0;
}
