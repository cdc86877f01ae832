/* oracle_fem2d.c -- CPU restatement of the reference's two 2D workloads, the Kirchhoff plate and
 * the plane-stress cantilever (test infrastructure: checker only, never the measured or shipped
 * path).
 *
 *   FemPlate              tests/MGroup.LinearAlgebra.Tests/TestData/SparseLinearSystems/FemPlate.cs
 *   FemCantilever2D       tests/MGroup.LinearAlgebra.Tests/TestData/SparseLinearSystems/FemCantilever2D.cs
 *                         with FemCantileverBase.cs (assembly, free dofs, manufactured solution)
 *   plate element         tests/MGroup.FEM.Structural.Tests/FemExtentions/PlateElementRectangle4Nodes3Dofs.cs
 *   load vector           src/MGroup.Solvers/Assemblers/PlateElementVectorAssembler.cs
 *   matrix assembly       src/MGroup.Solvers/Assemblers/DokRowMajorMatrixAssembler.cs:30-57 with
 *                         src/MGroup.LinearAlgebra/Matrices/Builders/DokRowMajor.cs:244-270
 *   transfers             src/MGroup.Solvers/Multigrid/IGeometricMultigridModel.cs:76-194 (2 axes)
 *
 * Mesh conventions (the reference takes them from the unvendored UniformCartesianMesh2D): node id
 * = ix + (nx+1) iy at (ix Lx/nx, iy Ly/ny); element e = ex + nx ey with corners (ex,ey),
 * (ex+1,ey), (ex+1,ey+1), (ex,ey+1) -- the counter-clockwise order the element's rectangle check
 * demands (PlateElementRectangle4Nodes3Dofs.cs:33-39); dofs per node (w, rx, ry) in that order,
 * numbered node by node, constrained dofs left out (FemPlate.cs:72-101).  The cantilever has the
 * same mesh with dofs (ux, uy) per node and every node with ix == 0 clamped
 * (FemCantileverBase.cs:137-158).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* Element stiffness, PlateElementRectangle4Nodes3Dofs.cs:134-222 (Adini-Clough-Melosh rectangle of
 * half-sides a, b): the lower triangle of the closed form, mirrored, times D / (60 a b) with the
 * bending stiffness D = E t^3 / (12 (1 - nu^2)) (:42).  Evaluated in the reference's operation
 * order. */
void orc_plate_ke(double a, double b, double E, double nu, double t, double *Ke)
{
    const double D = E * pow(t, 3) / (12 * (1 - pow(nu, 2)));
    const double p = pow(a / b, 2), v = nu;
    const double k33 = a * a * (80 / p + 16 - 16 * v), k43 = a * (-60 / p - 6 + 6 * v), k53 = 0;
    const double k63 = a * a * (40 / p - 4 + 4 * v), k73 = a * (-30 / p + 6 - 6 * v), k83 = 0;
    const double k93 = a * a * (20 / p + 4 - 4 * v), k10_3 = a * (30 / p - 6 - 24 * v), k11_3 = 0;
    const double k12_3 = a * a * (40 / p - 16 + 16 * v);
    const double k22 = b * b * (80 * p + 16 - 16 * v), k32 = 60 * v * a * b, k42 = b * (30 * p - 6 - 24 * v);
    const double k52 = b * b * (40 * p - 16 + 16 * v), k62 = -k53, k72 = b * (-30 * p + 6 - 6 * v);
    const double k82 = b * b * (20 * p + 4 - 4 * v), k92 = k83, k10_2 = b * (-60 * p - 6 + 6 * v);
    const double k11_2 = b * b * (40 * p - 4 + 4 * v), k12_2 = -k11_3;
    const double k11 = 60 * p + 60 / p + 42 - 12 * v, k21 = b * (60 * p + 6 + 24 * v);
    const double k31 = a * (60 / p + 6 + 24 * v), k41 = 30 * p - 60 / p - 42 + 12 * v, k51 = k42, k61 = -k43;
    const double k71 = -30 * p - 30 / p + 42 - 12 * v, k81 = -k72, k91 = -k73;
    const double k10_1 = -60 * p + 30 / p - 42 + 12 * v, k11_1 = -k10_2, k12_1 = k10_3;
    const double L[12][12] = {
        {k11},
        {k21, k22},
        {k31, k32, k33},
        {k41, k42, k43, k11},
        {k51, k52, k53, k21, k22},
        {k61, k62, k63, -k31, -k32, k33},
        {k71, k72, k73, k10_1, k10_2, -k10_3, k11},
        {k81, k82, k83, k11_1, k11_2, -k11_3, -k21, k22},
        {k91, k92, k93, -k12_1, -k12_2, k12_3, -k31, k32, k33},
        {k10_1, k10_2, k10_3, k71, k72, -k73, k41, -k42, -k43, k11},
        {k11_1, k11_2, k11_3, k81, k82, -k83, -k51, k52, k53, -k21, k22},
        {k12_1, k12_2, k12_3, -k91, -k92, k93, -k61, k62, k63, k31, -k32, k33},
    };
    const double s = D / (60 * a * b);
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) Ke[i * 12 + j] = (j <= i ? L[i][j] : L[j][i]) * s;
}

/* Nodal forces of a uniform load q on one element, ConvertBodyForcesToNodal (:104-116). */
void orc_plate_load(double a, double b, double q, double *fe)
{
    const double r[12] = {1, b / 3, a / 3, 1, b / 3, -a / 3, 1, -b / 3, -a / 3, 1, -b / 3, a / 3};
    const double s = q * a * b;
    for (int i = 0; i < 12; ++i) fe[i] = r[i] * s;
}

/* Nodes whose w is fixed, FemPlate.cs:181-196: every node within 0.1 min(Lx/(nx-1), Ly/(ny-1))
 * of the lines x = 0, x = Lx/2, y = 0, y = Ly/2, y = Ly. */
static int plate_fixed(int ix, int iy, int nx, int ny, double Lx, double Ly)
{
    const double tol = 0.1 * fmin(Lx / (nx - 1), Ly / (ny - 1));
    const double x = ix * (Lx / nx), y = iy * (Ly / ny);
    return fabs(x) < tol || fabs(x - Lx / 2) < tol || fabs(y) < tol || fabs(y - Ly / 2) < tol || fabs(y - Ly) < tol;
}

/* free dof index of dof d of node (ix, iy), or -1 (node-major numbering of FemPlate.cs:72-101 and
 * FemCantileverBase.cs:137-158); dpn dofs per node, first[] = free index of each node's first free
 * dof, fixed[] = bit d set when dof d of the node is constrained (plate: bit 0 at supported nodes;
 * cantilever: both bits at ix == 0) */
typedef struct { int nx, ny, dpn; int32_t *first; uint8_t *fixed; int32_t nfree; } plate_dofs;

static void dofs_alloc(int nx, int ny, int dpn, plate_dofs *P)
{
    const int64_t nn = (int64_t)(nx + 1) * (ny + 1);
    P->nx = nx; P->ny = ny; P->dpn = dpn;
    P->first = (int32_t *)malloc(sizeof(int32_t) * nn);
    P->fixed = (uint8_t *)malloc(nn);
}

static void dofs_number(plate_dofs *P)
{
    const int64_t nn = (int64_t)(P->nx + 1) * (P->ny + 1);
    int32_t f = 0;
    for (int64_t id = 0; id < nn; ++id) {
        P->first[id] = f;
        for (int d = 0; d < P->dpn; ++d) f += !((P->fixed[id] >> d) & 1);
    }
    P->nfree = f;
}

static void plate_dofs_make(int nx, int ny, double Lx, double Ly, plate_dofs *P)
{
    dofs_alloc(nx, ny, 3, P);
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix)
            P->fixed[ix + (int64_t)(nx + 1) * iy] = (uint8_t)plate_fixed(ix, iy, nx, ny, Lx, Ly);
    dofs_number(P);
}

static void cant2d_dofs_make(int nx, int ny, plate_dofs *P)
{
    dofs_alloc(nx, ny, 2, P);
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) P->fixed[ix + (int64_t)(nx + 1) * iy] = ix == 0 ? 3 : 0;
    dofs_number(P);
}

static void plate_dofs_free(plate_dofs *P) { free(P->first); free(P->fixed); }

static inline int32_t plate_dof(const plate_dofs *P, int64_t id, int d)
{
    const unsigned m = P->fixed[id];
    if ((m >> d) & 1) return -1;
    int32_t k = P->first[id];
    for (int j = 0; j < d; ++j) k += !((m >> j) & 1);
    return k;
}

int32_t orc_plate_nfree(int nx, int ny, double Lx, double Ly)
{
    plate_dofs P;
    plate_dofs_make(nx, ny, Lx, Ly, &P);
    const int32_t n = P.nfree;
    plate_dofs_free(&P);
    return n;
}

/* FindFreeDofs (FemPlate.cs:72-101): the all-dof index 3 id + d of every free dof */
void orc_plate_free_dofs(int nx, int ny, double Lx, double Ly, int32_t *dofs)
{
    plate_dofs P;
    plate_dofs_make(nx, ny, Lx, Ly, &P);
    int32_t k = 0;
    const int64_t nn = (int64_t)(nx + 1) * (ny + 1);
    for (int64_t id = 0; id < nn; ++id)
        for (int d = P.fixed[id] ? 1 : 0; d < 3; ++d) dofs[k++] = (int32_t)(3 * id + d);
    plate_dofs_free(&P);
}

/* nnz of K_ff: every free dof row couples the free dofs of the (up to 9) nodes sharing an
 * element with its node */
int64_t orc_plate_nnz(int nx, int ny, double Lx, double Ly)
{
    plate_dofs P;
    plate_dofs_make(nx, ny, Lx, Ly, &P);
    int64_t nnz = 0;
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) {
            const int64_t id = ix + (int64_t)(nx + 1) * iy;
            int cols = 0;
            for (int jy = iy - 1; jy <= iy + 1; ++jy)
                for (int jx = ix - 1; jx <= ix + 1; ++jx)
                    if (jx >= 0 && jx <= nx && jy >= 0 && jy <= ny)
                        cols += P.fixed[jx + (int64_t)(nx + 1) * jy] ? 2 : 3;
            nnz += (int64_t)cols * (P.fixed[id] ? 2 : 3);
        }
    plate_dofs_free(&P);
    return nnz;
}

/* K_ff (sorted columns) and the load vector b of FemPlate.CreateLinearSystem (FemPlate.cs:41-70).
 * Entry (r, c) is the sum, over the elements holding both dofs in ascending element order, of the
 * element matrix entry -- what AddSubmatrixSymmetric accumulates element after element (each pair
 * of global dofs of an element is added once, to the upper entry, and mirrored); b sums the
 * element load vectors in element order. */
void orc_plate_assemble(int nx, int ny, double Lx, double Ly, double t, double E, double nu, double q,
                        int32_t *rowptr, int32_t *col, double *val, double *b)
{
    plate_dofs P;
    plate_dofs_make(nx, ny, Lx, Ly, &P);
    const double a = 0.5 * (Lx / nx), bb = 0.5 * (Ly / ny);
    double Ke[144], fe[12];
    orc_plate_ke(a, bb, E, nu, t, Ke);
    orc_plate_load(a, bb, q, fe);
    /* corner c of element (ex, ey): offsets (0,0), (1,0), (1,1), (0,1) */
    static const int cx[4] = {0, 1, 1, 0}, cy[4] = {0, 0, 1, 1};
    int64_t pos = 0, row = 0;
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) {
            const int64_t id = ix + (int64_t)(nx + 1) * iy;
            for (int d = P.fixed[id] ? 1 : 0; d < 3; ++d) {
                rowptr[row++] = (int32_t)pos;
                /* columns: neighbour nodes in id order, their free dofs in order */
                for (int jy = iy - 1; jy <= iy + 1; ++jy)
                    for (int jx = ix - 1; jx <= ix + 1; ++jx) {
                        if (jx < 0 || jx > nx || jy < 0 || jy > ny) continue;
                        const int64_t jd = jx + (int64_t)(nx + 1) * jy;
                        for (int e2 = P.fixed[jd] ? 1 : 0; e2 < 3; ++e2) {
                            double s = 0;
                            /* elements holding both nodes, ascending e = ex + nx ey */
                            for (int ey = (iy > jy ? iy : jy) - 1; ey <= (iy < jy ? iy : jy); ++ey)
                                for (int ex = (ix > jx ? ix : jx) - 1; ex <= (ix < jx ? ix : jx); ++ex) {
                                    if (ex < 0 || ex >= nx || ey < 0 || ey >= ny) continue;
                                    int li = -1, lj = -1;
                                    for (int c = 0; c < 4; ++c) {
                                        if (ex + cx[c] == ix && ey + cy[c] == iy) li = c;
                                        if (ex + cx[c] == jx && ey + cy[c] == jy) lj = c;
                                    }
                                    s += Ke[(3 * li + d) * 12 + 3 * lj + e2];
                                }
                            col[pos] = plate_dof(&P, jd, e2);
                            val[pos] = s;
                            ++pos;
                        }
                    }
            }
        }
    rowptr[row] = (int32_t)pos;
    if (b) {
        memset(b, 0, sizeof(double) * (size_t)P.nfree);
        for (int ey = 0; ey < ny; ++ey)
            for (int ex = 0; ex < nx; ++ex)
                for (int c = 0; c < 4; ++c) {
                    const int64_t id = ex + cx[c] + (int64_t)(nx + 1) * (ey + cy[c]);
                    for (int d = 0; d < 3; ++d) {
                        const int32_t g = plate_dof(&P, id, d);
                        if (g >= 0) b[g] += fe[3 * c + d];
                    }
                }
    }
    plate_dofs_free(&P);
}

/* ---------------------------------------------------------------- transfers
 * IGeometricMultigridModel.cs:76-194 on two axes: cone weights 1 - |w - cursor| / distance of the
 * fine nodes in each coarse node's window, restriction rows normalised over all fine nodes of
 * the window, interpolation rows over all coarse nodes reaching the fine node, the same weight for
 * each of the three dofs, then the free-dof submatrices (the fixed dofs count in the sums). */
typedef struct { int nf[2], nc[2]; double nid[2]; } prp_dims;

static void prp_setup(int nx, int ny, prp_dims *d)
{
    const int n[2] = {nx, ny};
    for (int i = 0; i < 2; ++i) {
        const int c = n[i] > 1 ? n[i] >> 1 : n[i];                     /* :54-58 */
        d->nid[i] = (double)n[i] / c;                                   /* :92 */
        d->nf[i] = n[i] + 1;
        d->nc[i] = c + 1;
    }
}

static void prp_window(const prp_dims *d, int i, int a, int *s, int *e, double *cur)
{
    const double c = a * d->nid[i];                                     /* :123 */
    int lo = (int)ceil(nextafter(c - d->nid[i], INFINITY));             /* :124 */
    if (lo < 0) lo = 0;
    int hi = (int)floor(nextafter(c + d->nid[i], -INFINITY)) + 1;       /* :126 */
    if (hi > d->nf[i]) hi = d->nf[i];
    *s = lo; *e = hi; *cur = c;
}

static double prp_influence(const prp_dims *d, const int *w, const double *cur)
{
    double s = 0;                                                       /* :145-151 */
    for (int i = 0; i < 2; ++i) { const double d0 = (w[i] - cur[i]) / d->nid[i]; s += d0 * d0; }
    return 1 - sqrt(s);
}

/* the transfers between the dof maps F (fine) and Cd (its coarser model) */
static void transfer_nnz(const plate_dofs *F_, const plate_dofs *Cd_, int64_t *nnzR, int64_t *nnzP)
{
    prp_dims d; prp_setup(F_->nx, F_->ny, &d);
    const plate_dofs F = *F_, Cd = *Cd_;
    int64_t nr = 0;
    for (int a1 = 0; a1 < d.nc[1]; ++a1)
        for (int a0 = 0; a0 < d.nc[0]; ++a0) {
            int s[2], e[2]; double cur[2];
            prp_window(&d, 0, a0, &s[0], &e[0], &cur[0]);
            prp_window(&d, 1, a1, &s[1], &e[1], &cur[1]);
            const int64_t cid = a0 + (int64_t)d.nc[0] * a1;
            for (int dd = 0; dd < F.dpn; ++dd) {
                if (plate_dof(&Cd, cid, dd) < 0) continue;
                for (int w1 = s[1]; w1 < e[1]; ++w1)
                    for (int w0 = s[0]; w0 < e[0]; ++w0)
                        nr += plate_dof(&F, w0 + (int64_t)d.nf[0] * w1, dd) >= 0;
            }
        }
    *nnzR = nr;
    *nnzP = nr;                                                         /* the same pairs, transposed */
}

static void transfer_fill(const plate_dofs *F_, const plate_dofs *Cd_, int32_t *Rrowptr, int32_t *Rcol,
                          double *Rval, int32_t *Prowptr, int32_t *Pcol, double *Pval)
{
    prp_dims d; prp_setup(F_->nx, F_->ny, &d);
    const plate_dofs F = *F_, Cd = *Cd_;
    const int64_t nfn = (int64_t)d.nf[0] * d.nf[1], ncn = (int64_t)d.nc[0] * d.nc[1];
    double *sumR = (double *)calloc(ncn, sizeof(double)), *sumP = (double *)calloc(nfn, sizeof(double));
    int s[2], e[2], w[2], a[2]; double cur[2];
    /* the reference's loop order (a0 outer, a1 inner; w0 outer, w1 inner) for the row sums */
    for (a[0] = 0; a[0] < d.nc[0]; ++a[0])
        for (a[1] = 0; a[1] < d.nc[1]; ++a[1]) {
            for (int i = 0; i < 2; ++i) prp_window(&d, i, a[i], &s[i], &e[i], &cur[i]);
            double sr = 0;
            for (w[0] = s[0]; w[0] < e[0]; ++w[0])
                for (w[1] = s[1]; w[1] < e[1]; ++w[1]) {
                    const double infl = prp_influence(&d, w, cur);
                    sr += infl;
                    sumP[w[0] + (int64_t)d.nf[0] * w[1]] += infl;
                }
            sumR[a[0] + (int64_t)d.nc[0] * a[1]] = sr;
        }
    /* R: coarse free dofs in order; columns: fine free dofs of the same dof index, ascending */
    int64_t pos = 0, row = 0;
    for (a[1] = 0; a[1] < d.nc[1]; ++a[1])
        for (a[0] = 0; a[0] < d.nc[0]; ++a[0]) {
            const int64_t cid = a[0] + (int64_t)d.nc[0] * a[1];
            for (int i = 0; i < 2; ++i) prp_window(&d, i, a[i], &s[i], &e[i], &cur[i]);
            for (int dd = 0; dd < F.dpn; ++dd) {
                if (plate_dof(&Cd, cid, dd) < 0) continue;
                Rrowptr[row++] = (int32_t)pos;
                for (w[1] = s[1]; w[1] < e[1]; ++w[1])
                    for (w[0] = s[0]; w[0] < e[0]; ++w[0]) {
                        const int32_t c = plate_dof(&F, w[0] + (int64_t)d.nf[0] * w[1], dd);
                        if (c < 0) continue;
                        Rcol[pos] = c;
                        Rval[pos] = prp_influence(&d, w, cur) / sumR[cid];            /* :165 */
                        ++pos;
                    }
            }
        }
    Rrowptr[row] = (int32_t)pos;
    /* P: fine free dofs in order; columns: coarse free dofs of the same dof index whose window
       holds the fine node, ascending */
    pos = 0; row = 0;
    for (w[1] = 0; w[1] < d.nf[1]; ++w[1])
        for (w[0] = 0; w[0] < d.nf[0]; ++w[0]) {
            const int64_t fid = w[0] + (int64_t)d.nf[0] * w[1];
            int lo[2], hi[2];
            for (int i = 0; i < 2; ++i) {
                lo[i] = d.nc[i]; hi[i] = -1;
                for (int aa = 0; aa < d.nc[i]; ++aa) {
                    int ss, ee; double cc;
                    prp_window(&d, i, aa, &ss, &ee, &cc);
                    if (w[i] >= ss && w[i] < ee) { if (aa < lo[i]) lo[i] = aa; if (aa > hi[i]) hi[i] = aa; }
                }
            }
            for (int dd = 0; dd < F.dpn; ++dd) {
                if (plate_dof(&F, fid, dd) < 0) continue;
                Prowptr[row++] = (int32_t)pos;
                for (a[1] = lo[1]; a[1] <= hi[1]; ++a[1])
                    for (a[0] = lo[0]; a[0] <= hi[0]; ++a[0]) {
                        const int32_t c = plate_dof(&Cd, a[0] + (int64_t)d.nc[0] * a[1], dd);
                        if (c < 0) continue;
                        for (int i = 0; i < 2; ++i) prp_window(&d, i, a[i], &s[i], &e[i], &cur[i]);
                        Pcol[pos] = c;
                        Pval[pos] = prp_influence(&d, w, cur) / sumP[fid];            /* :180 */
                        ++pos;
                    }
            }
        }
    Prowptr[row] = (int32_t)pos;
    free(sumR); free(sumP);
}

void orc_plate_transfer_nnz(int nx, int ny, double Lx, double Ly, int64_t *nnzR, int64_t *nnzP)
{
    prp_dims d; prp_setup(nx, ny, &d);
    plate_dofs F, Cd;
    plate_dofs_make(nx, ny, Lx, Ly, &F);
    plate_dofs_make(d.nc[0] - 1, d.nc[1] - 1, Lx, Ly, &Cd);
    transfer_nnz(&F, &Cd, nnzR, nnzP);
    plate_dofs_free(&F); plate_dofs_free(&Cd);
}

void orc_plate_transfer(int nx, int ny, double Lx, double Ly, int32_t *Rrowptr, int32_t *Rcol, double *Rval,
                        int32_t *Prowptr, int32_t *Pcol, double *Pval)
{
    prp_dims d; prp_setup(nx, ny, &d);
    plate_dofs F, Cd;
    plate_dofs_make(nx, ny, Lx, Ly, &F);
    plate_dofs_make(d.nc[0] - 1, d.nc[1] - 1, Lx, Ly, &Cd);
    transfer_fill(&F, &Cd, Rrowptr, Rcol, Rval, Prowptr, Pcol, Pval);
    plate_dofs_free(&F); plate_dofs_free(&Cd);
}

/* ================================================================ plane-stress cantilever */

/* ElementStiffness, FemCantilever2D.cs:29-49 (the 4-node plane-stress element of "A 99 line
 * topology optimization code", Sigmund 2001): eight distinct entries k[0..7] in the reference's
 * operation order, placed by the reference's 8x8 index table, then every entry times
 * E / (1 - nu^2) (ScaleIntoThis). */
void orc_ke_q4(double E, double nu, double *Ke)
{
    const double v = nu;
    const double k[8] = {0.5 - v / 6.0, 0.125 + v / 8.0, -0.25 - v / 12.0, -0.125 + 3 * v / 8.0,
                         -0.25 + v / 12.0, -0.125 - v / 8.0, v / 6.0, 0.125 - 3 * v / 8.0};
    static const int T[8][8] = {{0, 1, 2, 3, 4, 5, 6, 7}, {1, 0, 7, 6, 5, 4, 3, 2}, {2, 7, 0, 5, 6, 3, 4, 1},
                                {3, 6, 5, 0, 7, 2, 1, 4}, {4, 5, 6, 7, 0, 1, 2, 3}, {5, 4, 3, 2, 1, 0, 7, 6},
                                {6, 3, 4, 1, 2, 7, 0, 5}, {7, 2, 1, 4, 3, 6, 5, 0}};
    const double s = E / (1 - v * v);
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) Ke[i * 8 + j] = k[T[i][j]] * s;
}

int32_t orc_cant2d_nfree(int nx, int ny) { return 2 * nx * (ny + 1); }

/* nnz of K_ff: the free dofs of the (up to 9) nodes sharing an element with the row's node */
int64_t orc_cant2d_nnz(int nx, int ny)
{
    int64_t nnz = 0;
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 1; ix <= nx; ++ix) {
            int cx = 0, cy = 0;
            for (int dd = -1; dd <= 1; ++dd) {
                cx += ix + dd >= 1 && ix + dd <= nx;
                cy += iy + dd >= 0 && iy + dd <= ny;
            }
            nnz += 4LL * cx * cy;
        }
    return nnz;
}

/* K_ff of AssembleGlobalMatrix + GetSubmatrix (FemCantileverBase.cs:87-104,137-158): entry (r, c)
 * sums density_e * Ke over the elements holding both dofs in ascending element order
 * e = ex + nx ey, each added to the running entry as sub + old (DokRowMajor.cs:182-207); rows and
 * the sorted columns (BuildCsrArrays(true)) in free-dof order.  dens = NULL: every density 1. */
void orc_cant2d_assemble(int nx, int ny, double E, double nu, const double *dens, int32_t *rowptr, int32_t *col,
                         double *val)
{
    double Ke[64];
    orc_ke_q4(E, nu, Ke);
    plate_dofs P;
    cant2d_dofs_make(nx, ny, &P);
    static const int cx[4] = {0, 1, 1, 0}, cy[4] = {0, 0, 1, 1};
    int64_t pos = 0, row = 0;
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 1; ix <= nx; ++ix)
            for (int a = 0; a < 2; ++a) {
                rowptr[row++] = (int32_t)pos;
                for (int jy = iy - 1; jy <= iy + 1; ++jy)
                    for (int jx = ix - 1; jx <= ix + 1; ++jx) {
                        if (jx < 1 || jx > nx || jy < 0 || jy > ny) continue;
                        const int64_t jd = jx + (int64_t)(nx + 1) * jy;
                        for (int b = 0; b < 2; ++b) {
                            double acc = 0;
                            for (int ey = (iy > jy ? iy : jy) - 1; ey <= (iy < jy ? iy : jy); ++ey)
                                for (int ex = (ix > jx ? ix : jx) - 1; ex <= (ix < jx ? ix : jx); ++ex) {
                                    if (ex < 0 || ex >= nx || ey < 0 || ey >= ny) continue;
                                    int li = -1, lj = -1;
                                    for (int c = 0; c < 4; ++c) {
                                        if (ex + cx[c] == ix && ey + cy[c] == iy) li = c;
                                        if (ex + cx[c] == jx && ey + cy[c] == jy) lj = c;
                                    }
                                    const double de = dens ? dens[ex + (int64_t)nx * ey] : 1.0;
                                    acc = de * Ke[(2 * li + a) * 8 + 2 * lj + b] + acc;
                                }
                            col[pos] = plate_dof(&P, jd, b);
                            val[pos] = acc;
                            ++pos;
                        }
                    }
            }
    rowptr[row] = (int32_t)pos;
    plate_dofs_free(&P);
}

/* U_f of CalcFreeDisplacements (FemCantileverBase.cs:160-182) with FemCantilever2D.cs:51-57: at
 * node (ix Lx/nx, iy Ly/ny), x the first coordinate and z = y - Ly/2, (u, w) of the Euler-Bernoulli
 * beam (:64-77) with I = Ly Lz / 12 Ly Ly (FemCantilever2D.cs:24, Lz the depth). */
void orc_cant2d_manufactured(int nx, int ny, double Lx, double Ly, double Lz, double E, double q, double *Uf)
{
    const double I = Ly * Lz / 12 * Ly * Ly, L = Lx;
    const double hx = Lx / nx, hy = Ly / ny;
    int64_t pos = 0;
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 1; ix <= nx; ++ix) {
            const double x = ix * hx;
            const double z = iy * hy - (Ly - 0) / 2;
            const double w = -q / (24 * E * I) * (pow(x, 4) - 4 * L * pow(x, 3) + 6 * L * L * x * x);
            const double rot = -q / (6 * E * I) * (pow(x, 3) - 3 * L * x * x + 3 * L * L * x);
            const double u = -rot * z;
            Uf[pos++] = u;
            Uf[pos++] = w;
        }
}

/* y = K_ff x with each row summed in the DOK insertion order (DokRowMajor.cs:660-674): a row's
 * columns in the order the element loop first touched them (elements ascending, then the
 * element's local dofs), FemCantileverBase.cs:95-101. */
void orc_cant2d_multiply_dok_order(int nx, int ny, const int32_t *rowptr, const int32_t *col, const double *val,
                                   const double *x, double *y)
{
    static const int cx[4] = {0, 1, 1, 0}, cy[4] = {0, 0, 1, 1};
    int64_t row = 0;
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 1; ix <= nx; ++ix)
            for (int a = 0; a < 2; ++a, ++row) {
                const int32_t s0 = rowptr[row], t0 = rowptr[row + 1];
                int32_t seen[18];
                int ns = 0;
                double dot = 0;
                for (int ey = iy - 1; ey <= iy; ++ey) {
                    if (ey < 0 || ey >= ny) continue;
                    for (int ex = ix - 1; ex <= ix; ++ex) {
                        if (ex < 0 || ex >= nx) continue;
                        for (int c = 0; c < 4; ++c) {
                            const int jx = ex + cx[c], jy = ey + cy[c];
                            if (jx == 0) continue;
                            for (int b = 0; b < 2; ++b) {
                                const int32_t cc = (int32_t)(2 * ((int64_t)jy * nx + (jx - 1)) + b);
                                int dup = 0;
                                for (int k = 0; k < ns; ++k) dup |= seen[k] == cc;
                                if (dup) continue;
                                seen[ns++] = cc;
                                int32_t lo = s0, hi = t0 - 1;
                                while (lo < hi) { const int32_t mid = (lo + hi) / 2; if (col[mid] < cc) lo = mid + 1; else hi = mid; }
                                dot += val[lo] * x[cc];
                            }
                        }
                    }
                }
                y[row] = dot;
            }
}

void orc_cant2d_transfer_nnz(int nx, int ny, int64_t *nnzR, int64_t *nnzP)
{
    prp_dims d; prp_setup(nx, ny, &d);
    plate_dofs F, Cd;
    cant2d_dofs_make(nx, ny, &F);
    cant2d_dofs_make(d.nc[0] - 1, d.nc[1] - 1, &Cd);
    transfer_nnz(&F, &Cd, nnzR, nnzP);
    plate_dofs_free(&F); plate_dofs_free(&Cd);
}

/* CreateRestrictionAndInterpolationMatrix (IGeometricMultigridModel.cs:76-194) on two axes with
 * two dofs per node, then the free-dof submatrices of the clamped ix == 0 nodes */
void orc_cant2d_transfer(int nx, int ny, int32_t *Rrowptr, int32_t *Rcol, double *Rval, int32_t *Prowptr,
                         int32_t *Pcol, double *Pval)
{
    prp_dims d; prp_setup(nx, ny, &d);
    plate_dofs F, Cd;
    cant2d_dofs_make(nx, ny, &F);
    cant2d_dofs_make(d.nc[0] - 1, d.nc[1] - 1, &Cd);
    transfer_fill(&F, &Cd, Rrowptr, Rcol, Rval, Prowptr, Pcol, Pval);
    plate_dofs_free(&F); plate_dofs_free(&Cd);
}
