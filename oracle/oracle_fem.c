/*
 * oracle_fem.c -- CPU restatement of the FE cantilever generator and the
 * reference multigrid transfer operators.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates:
 *   FemCantilever3D.cs:32-120      ElementStiffness (top3d closed form)          -> orc_ke_hex8
 *   FemCantileverBase.cs:87-104    AssembleGlobalMatrix (element-order sums)     -> orc_cantilever_assemble
 *   FemCantileverBase.cs:137-158   FindFreeDofs (nodes with ix == 0 clamped)     -> free numbering
 *   FemCantileverBase.cs:53-82,160-182 + FemCantilever3D.cs:122-128 manufactured solution
 *   DokRowMajor.cs:182-207         AddSubmatrix (new = sub + old)
 *   DokRowMajor.cs:286-324         BuildCsrArrays(sorted)
 *   DokRowMajor.cs:471-501,660-674 GetSubmatrix + MultiplyRight (insertion-order sums)
 *   IGeometricMultigridModel.cs:54-58,76-194  coarsening + restriction/interpolation
 *
 * Mesh conventions (the unvendored UniformCartesianMesh3D is not in the
 * reference; these are assumptions, the only ones consistent with the top3d Ke):
 *   node id n = ix + (nx+1)*(iy + (ny+1)*iz); element e = ex + nx*(ey + ny*ez);
 *   element corner order (0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1);
 *   dofs (ux,uy,uz) per node; coordinates = index * (L / N).
 *
 * Pinned by (tests/test_oracle_fem.py): Ke against textbook 2x2x2 Gauss quadrature of the
 * trilinear hexahedron (numpy) and against its invariants (symmetry, 6 rigid-body null
 * vectors for this corner order); the stencil nnz closed form; K * (rigid motion) = 0 on the
 * unconstrained assembly; R/P against their closed-form interior weights and row sums;
 * exact trilinear P against the reference-weight P; manufactured b against the element loop.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- RNG */
/* SplitMix64 (Steele, Lea, Flood 2014) seeds xoshiro256** (Blackman & Vigna 2018). */
uint64_t orc_splitmix64_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

typedef struct { uint64_t s[4]; } xo_t;

static void xo_seed(xo_t *g, uint64_t seed)
{
    uint64_t sm = seed;
    for (int i = 0; i < 4; ++i) g->s[i] = orc_splitmix64_next(&sm);
}

static inline uint64_t xo_next(xo_t *g)
{
    uint64_t *s = g->s;
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

static inline double xo_uniform(xo_t *g) { return (double)(xo_next(g) >> 11) * 0x1.0p-53; }

void orc_rng_uniform(uint64_t seed, int64_t n, double *out)
{
    xo_t g; xo_seed(&g, seed);
    for (int64_t i = 0; i < n; ++i) out[i] = xo_uniform(&g);
}

void orc_rng_normal(uint64_t seed, int64_t n, double *out)
{
    xo_t g; xo_seed(&g, seed);
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < n; ++i) {
        double u1 = xo_uniform(&g), u2 = xo_uniform(&g);
        out[i] = sqrt(-2.0 * log(1.0 - u1)) * cos(two_pi * u2);
    }
}

/* ---------------------------------------------------------------- Ke */
/* FemCantilever3D.cs:32-120: k = A^T [1, nu] / 72, blocks K1..K6, scaled by E/((nu+1)(1-2nu)). */
void orc_ke_hex8(double E, double nu, double *Ke)
{
    static const double A0[14] = {32, 6, -8, 6, -6, 4, 3, -6, -10, 3, -3, -3, -4, -8};
    static const double A1[14] = {-48, 0, 0, -24, 24, 0, 0, 0, 12, -12, 0, 12, 12, 12};
    double k[15];
    const double s72 = 1.0 / 72.0;
    k[0] = 0.0;
    for (int i = 0; i < 14; ++i) {
        double t = 0.0;
        t = t + A0[i] * 1.0;
        t = t + A1[i] * nu;
        k[i + 1] = t * s72;
    }
#define K6(M, a,b,c,d,e,f, g,h,i_,j,l,m, n,o,p,q,r,s, t,u,v,w,x,y, z1,z2,z3,z4,z5,z6, z7,z8,z9,z10,z11,z12) \
    const int M[36] = {a,b,c,d,e,f, g,h,i_,j,l,m, n,o,p,q,r,s, t,u,v,w,x,y, z1,z2,z3,z4,z5,z6, z7,z8,z9,z10,z11,z12}
    K6(K1, 1,2,2,3,5,5, 2,1,2,4,6,7, 2,2,1,4,7,6, 3,4,4,1,8,8, 5,6,7,8,1,2, 5,7,6,8,2,1);
    K6(K2, 9,8,12,6,4,7, 8,9,12,5,3,5, 10,10,13,7,4,6, 6,5,11,9,2,10, 4,3,5,2,9,12, 11,4,6,12,10,13);
    K6(K3, 6,7,4,9,12,8, 7,6,4,10,13,10, 5,5,3,8,12,9, 9,10,2,6,11,5, 12,13,10,11,6,4, 2,12,9,4,5,3);
    K6(K4, 14,11,11,13,10,10, 11,14,11,12,9,8, 11,11,14,12,8,9, 13,12,12,14,7,7, 10,9,8,7,14,11, 10,8,9,7,11,14);
    K6(K5, 1,2,8,3,5,4, 2,1,8,4,6,11, 8,8,1,5,11,6, 3,4,5,1,8,2, 5,6,11,8,1,8, 4,11,6,2,8,1);
    K6(K6_, 14,11,7,13,10,12, 11,14,7,12,9,2, 7,7,14,10,2,9, 13,12,10,14,7,11, 10,9,2,7,14,7, 12,2,9,11,7,14);
#undef K6
    /* block layout: {K1,K2,K3,K4},{K2',K5,K6,K3'},{K3',K6,K5',K2'},{K4,K3,K2,K1'} */
    const int *blk[4][4] = {{K1, K2, K3, K4}, {K2, K5, K6_, K3}, {K3, K6_, K5, K2}, {K4, K3, K2, K1}};
    const int tr[4][4] = {{0, 0, 0, 0}, {1, 0, 0, 1}, {1, 0, 1, 1}, {0, 0, 0, 1}};
    const double scale = E / ((nu + 1) * (1 - 2 * nu));
    for (int bi = 0; bi < 4; ++bi)
        for (int bj = 0; bj < 4; ++bj)
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j) {
                    int idx = tr[bi][bj] ? blk[bi][bj][j * 6 + i] : blk[bi][bj][i * 6 + j];
                    Ke[(bi * 6 + i) * 24 + bj * 6 + j] = k[idx] * scale;
                }
}

/* ---------------------------------------------------------------- mesh helpers */
/* local corner index from offsets (ox,oy,oz) in {0,1}^3 */
static const int LOC[2][2][2] = {{{0, 4}, {3, 7}}, {{1, 5}, {2, 6}}}; /* LOC[ox][oy][oz] */

static inline int64_t free_node(int ix, int iy, int iz, int nx, int ny)
{
    return ((int64_t)iz * (ny + 1) + iy) * nx + (ix - 1);
}

int64_t orc_cantilever_nnz(int nx, int ny, int nz)
{
    int64_t nnz = 0;
    for (int iz = 0; iz <= nz; ++iz)
        for (int iy = 0; iy <= ny; ++iy)
            for (int ix = 1; ix <= nx; ++ix) {
                int cx = 0, cy = 0, cz = 0;
                for (int d = -1; d <= 1; ++d) {
                    if (ix + d >= 1 && ix + d <= nx) cx++;
                    if (iy + d >= 0 && iy + d <= ny) cy++;
                    if (iz + d >= 0 && iz + d <= nz) cz++;
                }
                nnz += 9LL * cx * cy * cz;
            }
    return nnz;
}

/* element ranges shared by nodes i and j = i + d along one axis with n elements */
static inline int elem_range(int i, int d, int n, int *e0)
{
    if (d == 0) {
        int lo = i - 1 < 0 ? 0 : i - 1, hi = i > n - 1 ? n - 1 : i;
        *e0 = lo; return hi - lo + 1;
    }
    *e0 = d > 0 ? i : i - 1;
    return 1;
}

/* Shared core: per-element matrices given either as rho*Ke (Kel == NULL) or per element. */
static void assemble_core(int nx, int ny, int nz, const double *Ke, const double *dens,
                          const double *Kel, int32_t *rowptr, int32_t *col, double *val)
{
    int64_t pos = 0;
    int64_t row = 0;
    for (int iz = 0; iz <= nz; ++iz)
        for (int iy = 0; iy <= ny; ++iy)
            for (int ix = 1; ix <= nx; ++ix)
                for (int a = 0; a < 3; ++a) {
                    rowptr[row++] = (int32_t)pos;
                    for (int dz = -1; dz <= 1; ++dz) {
                        int jz = iz + dz; if (jz < 0 || jz > nz) continue;
                        for (int dy = -1; dy <= 1; ++dy) {
                            int jy = iy + dy; if (jy < 0 || jy > ny) continue;
                            for (int dx = -1; dx <= 1; ++dx) {
                                int jx = ix + dx; if (jx < 1 || jx > nx) continue;
                                int64_t fj = free_node(jx, jy, jz, nx, ny);
                                int ex0, ey0, ez0;
                                int cx = elem_range(ix, dx, nx, &ex0);
                                int cy = elem_range(iy, dy, ny, &ey0);
                                int cz = elem_range(iz, dz, nz, &ez0);
                                for (int b = 0; b < 3; ++b) {
                                    double acc = 0.0;
                                    for (int ez = ez0; ez < ez0 + cz; ++ez)
                                        for (int ey = ey0; ey < ey0 + cy; ++ey)
                                            for (int ex = ex0; ex < ex0 + cx; ++ex) {
                                                int64_t e = ex + (int64_t)nx * (ey + (int64_t)ny * ez);
                                                int li = LOC[ix - ex][iy - ey][iz - ez];
                                                int lj = LOC[jx - ex][jy - ey][jz - ez];
                                                double c;
                                                if (Kel) c = Kel[e * 576 + (3 * li + a) * 24 + 3 * lj + b];
                                                else c = (dens ? dens[e] : 1.0) * Ke[(3 * li + a) * 24 + 3 * lj + b];
                                                acc = c + acc;   /* DokRowMajor.cs:203: sub + old */
                                            }
                                    col[pos] = (int32_t)(3 * fj + b);
                                    val[pos] = acc;
                                    ++pos;
                                }
                            }
                        }
                    }
                }
    rowptr[row] = (int32_t)pos;
}

void orc_cantilever_assemble(int nx, int ny, int nz, double E, double nu, const double *dens,
                             int32_t *rowptr, int32_t *col, double *val)
{
    double Ke[576];
    orc_ke_hex8(E, nu, Ke);
    assemble_core(nx, ny, nz, Ke, dens, NULL, rowptr, col, val);
}

void orc_assemble_elements(int nx, int ny, int nz, const double *Kel,
                           int32_t *rowptr, int32_t *col, double *val)
{
    assemble_core(nx, ny, nz, NULL, NULL, Kel, rowptr, col, val);
}

/* FemCantileverBase.cs:70-82,160-182; FemCantilever3D.cs:122-128 (q = DistributedLoad, E = ElasticityModulus) */
void orc_cantilever_manufactured(int nx, int ny, int nz, double Lx, double Ly, double Lz,
                                 double E, double q, double *Uf)
{
    const double I = Ly * Lz / 12 * Lz * Lz;                 /* FemCantilever3D.cs:28 */
    const double L = Lx;
    const double hx = Lx / nx, hz = Lz / nz;
    int64_t pos = 0;
    for (int iz = 0; iz <= nz; ++iz)
        for (int iy = 0; iy <= ny; ++iy)
            for (int ix = 1; ix <= nx; ++ix) {
                double x = 0 + ix * hx;
                double z = (0 + iz * hz) - (Lz - 0) / 2;
                double w = -q / (24 * E * I) * (pow(x, 4) - 4 * L * pow(x, 3) + 6 * L * L * x * x);
                double rot = -q / (6 * E * I) * (pow(x, 3) - 3 * L * x * x + 3 * L * L * x);
                double u = -rot * z;
                Uf[pos++] = u; Uf[pos++] = 0; Uf[pos++] = w;
            }
}

/* Tip load of total magnitude `total` along z, split equally over the x = Lx face nodes
   (the reference's per-node split, CantileverBeam.cs:150-156). */
void orc_cantilever_tip_load(int nx, int ny, int nz, double total, double *f)
{
    int64_t nfree = 3LL * nx * (ny + 1) * (nz + 1);
    memset(f, 0, sizeof(double) * nfree);
    double per = total / ((double)(ny + 1) * (nz + 1));
    for (int iz = 0; iz <= nz; ++iz)
        for (int iy = 0; iy <= ny; ++iy)
            f[3 * free_node(nx, iy, iz, nx, ny) + 2] = per;
}

/* y = K_ff x with each row summed in DOK insertion order (DokRowMajor.cs:660-674 builds an
   unsorted CSR from the dictionaries; insertion order = first touch in element order, then
   local column order, FemCantileverBase.cs:95-101 + DokRowMajor.cs:182-207). */
void orc_csr_multiply_dok_order(int nx, int ny, int nz, const int32_t *rowptr, const int32_t *col,
                                const double *val, const double *x, double *y)
{
    static const int OFF[8][3] = {{0,0,0},{1,0,0},{1,1,0},{0,1,0},{0,0,1},{1,0,1},{1,1,1},{0,1,1}};
    int64_t row = 0;
    for (int iz = 0; iz <= nz; ++iz)
        for (int iy = 0; iy <= ny; ++iy)
            for (int ix = 1; ix <= nx; ++ix)
                for (int a = 0; a < 3; ++a, ++row) {
                    int32_t s = rowptr[row], t = rowptr[row + 1];
                    /* sorted columns -> position lookup via binary search */
                    double dot = 0.0;
                    int32_t seen[81]; int ns = 0;
                    for (int ez = iz - 1; ez <= iz; ++ez) {
                        if (ez < 0 || ez >= nz) continue;
                        for (int ey = iy - 1; ey <= iy; ++ey) {
                            if (ey < 0 || ey >= ny) continue;
                            for (int ex = ix - 1; ex <= ix; ++ex) {
                                if (ex < 0 || ex >= nx) continue;
                                for (int lj = 0; lj < 8; ++lj) {
                                    int jx = ex + OFF[lj][0], jy = ey + OFF[lj][1], jz = ez + OFF[lj][2];
                                    if (jx == 0) continue;
                                    for (int b = 0; b < 3; ++b) {
                                        int32_t c = (int32_t)(3 * free_node(jx, jy, jz, nx, ny) + b);
                                        int dup = 0;
                                        for (int k = 0; k < ns; ++k) if (seen[k] == c) { dup = 1; break; }
                                        if (dup) continue;
                                        seen[ns++] = c;
                                        int32_t lo = s, hi = t - 1;
                                        while (lo < hi) { int32_t mid = (lo + hi) / 2; if (col[mid] < c) lo = mid + 1; else hi = mid; }
                                        dot += val[lo] * x[c];
                                    }
                                }
                            }
                        }
                    }
                    y[row] = dot;
                }
}

/* ---------------------------------------------------------------- reference R / P */
/* IGeometricMultigridModel.cs:54-58 */
static inline int coarser(int n) { return n > 1 ? n >> 1 : n; }

typedef struct { int nf[3], nc[3]; double nid[3]; } rp_dims;

static void rp_setup(int nx, int ny, int nz, rp_dims *d)
{
    int nf[3] = {nx, ny, nz};
    for (int i = 0; i < 3; ++i) {
        int c = coarser(nf[i]);
        d->nid[i] = (double)nf[i] / c;               /* :92 */
        d->nf[i] = nf[i] + 1; d->nc[i] = c + 1;       /* node counts */
    }
}

static inline void rp_window(const rp_dims *d, int i, int a, int *start, int *end, double *cursor)
{
    double c = a * d->nid[i];                                         /* :123 */
    int s = (int)ceil(nextafter(c - d->nid[i], INFINITY));             /* :124 */
    if (s < 0) s = 0;
    int e = (int)floor(nextafter(c + d->nid[i], -INFINITY)) + 1;       /* :126 */
    if (e > d->nf[i]) e = d->nf[i];
    *start = s; *end = e; *cursor = c;
}

static inline double rp_influence(const rp_dims *d, const int *w, const double *cursor)
{
    double s = 0;                                                      /* :145-151 */
    for (int i = 0; i < 3; ++i) { double d0 = (w[i] - cursor[i]) / d->nid[i]; s += d0 * d0; }
    return 1 - sqrt(s);
}

void orc_rp_reference_nnz(int nx, int ny, int nz, int64_t *nnzR, int64_t *nnzP)
{
    rp_dims d; rp_setup(nx, ny, nz, &d);
    int64_t nr = 0, np = 0;
    /* R rows: coarse free nodes; entries: free fine nodes in window */
    for (int a2 = 0; a2 < d.nc[2]; ++a2)
        for (int a1 = 0; a1 < d.nc[1]; ++a1)
            for (int a0 = 1; a0 < d.nc[0]; ++a0) {
                int a[3] = {a0, a1, a2}, s[3], e[3]; double c[3];
                for (int i = 0; i < 3; ++i) rp_window(&d, i, a[i], &s[i], &e[i], &c[i]);
                int sx = s[0] < 1 ? 1 : s[0];
                int cnt = (e[0] > sx ? e[0] - sx : 0) * (e[1] - s[1]) * (e[2] - s[2]);
                nr += 3LL * cnt;
            }
    /* P rows: fine free nodes; entries: free coarse nodes whose window contains it */
    for (int a2 = 0; a2 < d.nc[2]; ++a2)
        for (int a1 = 0; a1 < d.nc[1]; ++a1)
            for (int a0 = 1; a0 < d.nc[0]; ++a0) {
                int a[3] = {a0, a1, a2}, s[3], e[3]; double c[3];
                for (int i = 0; i < 3; ++i) rp_window(&d, i, a[i], &s[i], &e[i], &c[i]);
                int sx = s[0] < 1 ? 1 : s[0];
                int cnt = (e[0] > sx ? e[0] - sx : 0) * (e[1] - s[1]) * (e[2] - s[2]);
                np += 3LL * cnt;
            }
    *nnzR = nr; *nnzP = np;
}

void orc_rp_reference(int nx, int ny, int nz,
                      int32_t *Rrowptr, int32_t *Rcol, double *Rval,
                      int32_t *Prowptr, int32_t *Pcol, double *Pval)
{
    rp_dims d; rp_setup(nx, ny, nz, &d);
    const int fx = nx, fy = ny; const int cxn = coarser(nx), cyn = coarser(ny);
    int64_t nfnodes = (int64_t)d.nf[0] * d.nf[1] * d.nf[2];
    double *sumP = (double *)calloc(nfnodes, sizeof(double));
    /* Pass 1 (reference loop order a0, a2, a1 and w0, w2, w1): row sums. */
    int64_t ncnodes = (int64_t)d.nc[0] * d.nc[1] * d.nc[2];
    double *sumR = (double *)calloc(ncnodes, sizeof(double));
    int a[3], w[3], s[3], e[3]; double cur[3];
    for (a[0] = 0; a[0] < d.nc[0]; ++a[0])
        for (a[2] = 0; a[2] < d.nc[2]; ++a[2])
            for (a[1] = 0; a[1] < d.nc[1]; ++a[1]) {
                for (int i = 0; i < 3; ++i) rp_window(&d, i, a[i], &s[i], &e[i], &cur[i]);
                int64_t cn = a[0] + (int64_t)d.nc[0] * (a[1] + (int64_t)d.nc[1] * a[2]);
                double sr = 0;
                for (w[0] = s[0]; w[0] < e[0]; ++w[0])
                    for (w[2] = s[2]; w[2] < e[2]; ++w[2])
                        for (w[1] = s[1]; w[1] < e[1]; ++w[1]) {
                            double infl = rp_influence(&d, w, cur);
                            int64_t fnode = w[0] + (int64_t)d.nf[0] * (w[1] + (int64_t)d.nf[1] * w[2]);
                            sr += infl;
                            sumP[fnode] += infl;
                        }
                sumR[cn] = sr;
            }
    /* R: rows coarse free dofs in free order; columns fine free dofs ascending. */
    int64_t pos = 0, row = 0;
    for (a[2] = 0; a[2] < d.nc[2]; ++a[2])
        for (a[1] = 0; a[1] < d.nc[1]; ++a[1])
            for (a[0] = 1; a[0] < d.nc[0]; ++a[0]) {
                for (int i = 0; i < 3; ++i) rp_window(&d, i, a[i], &s[i], &e[i], &cur[i]);
                int64_t cn = a[0] + (int64_t)d.nc[0] * (a[1] + (int64_t)d.nc[1] * a[2]);
                for (int dof = 0; dof < 3; ++dof) {
                    Rrowptr[row++] = (int32_t)pos;
                    for (w[2] = s[2]; w[2] < e[2]; ++w[2])
                        for (w[1] = s[1]; w[1] < e[1]; ++w[1])
                            for (w[0] = s[0] < 1 ? 1 : s[0]; w[0] < e[0]; ++w[0]) {
                                double infl = rp_influence(&d, w, cur);
                                Rcol[pos] = (int32_t)(3 * free_node(w[0], w[1], w[2], fx, fy) + dof);
                                Rval[pos] = infl / sumR[cn];                    /* :165 */
                                ++pos;
                            }
                }
            }
    Rrowptr[row] = (int32_t)pos;
    /* P: rows fine free dofs; columns coarse free dofs ascending. */
    pos = 0; row = 0;
    for (w[2] = 0; w[2] < d.nf[2]; ++w[2])
        for (w[1] = 0; w[1] < d.nf[1]; ++w[1])
            for (w[0] = 1; w[0] < d.nf[0]; ++w[0]) {
                int64_t fnode = w[0] + (int64_t)d.nf[0] * (w[1] + (int64_t)d.nf[1] * w[2]);
                /* candidate coarse indices per axis: a with start(a) <= w < end(a) */
                int lo[3], hi[3];
                for (int i = 0; i < 3; ++i) {
                    lo[i] = d.nc[i]; hi[i] = -1;
                    for (int aa = 0; aa < d.nc[i]; ++aa) {
                        int ss, ee; double cc; rp_window(&d, i, aa, &ss, &ee, &cc);
                        if (w[i] >= ss && w[i] < ee) { if (aa < lo[i]) lo[i] = aa; if (aa > hi[i]) hi[i] = aa; }
                    }
                }
                for (int dof = 0; dof < 3; ++dof) {
                    Prowptr[row++] = (int32_t)pos;
                    for (a[2] = lo[2]; a[2] <= hi[2]; ++a[2])
                        for (a[1] = lo[1]; a[1] <= hi[1]; ++a[1])
                            for (a[0] = lo[0] < 1 ? 1 : lo[0]; a[0] <= hi[0]; ++a[0]) {
                                for (int i = 0; i < 3; ++i) rp_window(&d, i, a[i], &s[i], &e[i], &cur[i]);
                                double infl = rp_influence(&d, w, cur);
                                Pcol[pos] = (int32_t)(3 * free_node(a[0], a[1], a[2], cxn, cyn) + dof);
                                Pval[pos] = infl / sumP[fnode];                /* :180 */
                                ++pos;
                            }
                }
            }
    Prowptr[row] = (int32_t)pos;
    free(sumP); free(sumR);
}

/* ---------------------------------------------------------------- exact trilinear P */
void orc_p_trilinear_nnz(int nx, int ny, int nz, int64_t *nnzP)
{
    int64_t n = 0;
    for (int iz = 0; iz <= nz; ++iz)
        for (int iy = 0; iy <= ny; ++iy)
            for (int ix = 1; ix <= nx; ++ix) {
                int cx = (ix & 1) ? 2 : 1, cy = (iy & 1) ? 2 : 1, cz = (iz & 1) ? 2 : 1;
                if ((ix & 1) && ix == 1) cx = 1;   /* coarse node 0 along x is clamped */
                n += 3LL * cx * cy * cz;
            }
    *nnzP = n;
}

void orc_p_trilinear(int nx, int ny, int nz, int32_t *Prowptr, int32_t *Pcol, double *Pval)
{
    const int cnx = nx / 2, cny = ny / 2;
    int64_t pos = 0, row = 0;
    for (int iz = 0; iz <= nz; ++iz)
        for (int iy = 0; iy <= ny; ++iy)
            for (int ix = 1; ix <= nx; ++ix) {
                int c0[3], cn[3]; double wt[3];
                int f[3] = {ix, iy, iz};
                for (int i = 0; i < 3; ++i) {
                    if (f[i] & 1) { c0[i] = (f[i] - 1) / 2; cn[i] = 2; wt[i] = 0.5; }
                    else { c0[i] = f[i] / 2; cn[i] = 1; wt[i] = 1.0; }
                }
                for (int dof = 0; dof < 3; ++dof) {
                    Prowptr[row++] = (int32_t)pos;
                    for (int cz = c0[2]; cz < c0[2] + cn[2]; ++cz)
                        for (int cy = c0[1]; cy < c0[1] + cn[1]; ++cy)
                            for (int cx = c0[0]; cx < c0[0] + cn[0]; ++cx) {
                                if (cx == 0) continue;
                                Pcol[pos] = (int32_t)(3 * free_node(cx, cy, cz, cnx, cny) + dof);
                                Pval[pos] = wt[0] * wt[1] * wt[2];
                                ++pos;
                            }
                }
            }
    Prowptr[row] = (int32_t)pos;
}

/* ---------------------------------------------------------------- SIMP adjoint */
/* dC/drho_e = -p rho_e^(p-1) (E0 - Emin) u_e^T Ke0 u_e  (modified SIMP, E_e = Emin + rho^p (E0-Emin)) */
void orc_simp_gradient(int nx, int ny, int nz, const double *Ke0, const double *rho, double penal,
                       double Emin, double E0, const double *u, double *grad)
{
    static const int OFF[8][3] = {{0,0,0},{1,0,0},{1,1,0},{0,1,0},{0,0,1},{1,0,1},{1,1,1},{0,1,1}};
    for (int ez = 0; ez < nz; ++ez)
        for (int ey = 0; ey < ny; ++ey)
            for (int ex = 0; ex < nx; ++ex) {
                int64_t e = ex + (int64_t)nx * (ey + (int64_t)ny * ez);
                double ue[24];
                for (int l = 0; l < 8; ++l) {
                    int jx = ex + OFF[l][0], jy = ey + OFF[l][1], jz = ez + OFF[l][2];
                    for (int b = 0; b < 3; ++b)
                        ue[3 * l + b] = jx == 0 ? 0.0 : u[3 * free_node(jx, jy, jz, nx, ny) + b];
                }
                double ce = 0;
                for (int i = 0; i < 24; ++i) {
                    double s = 0;
                    for (int j = 0; j < 24; ++j) s += Ke0[i * 24 + j] * ue[j];
                    ce += ue[i] * s;
                }
                grad[e] = -penal * pow(rho[e], penal - 1) * (E0 - Emin) * ce;
            }
}
