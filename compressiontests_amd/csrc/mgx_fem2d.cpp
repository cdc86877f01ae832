// mgx_fem2d.cpp -- host generators of the reference's two 2D workloads (in
// tests/MGroup.LinearAlgebra.Tests/TestData/SparseLinearSystems/): the Kirchhoff plate (FemPlate.cs)
// and the plane-stress cantilever (FemCantilever2D.cs with FemCantileverBase.cs).  For each: the
// element matrix, the free-dof stiffness K_ff (and the plate's load vector), the free-dof list,
// and the restriction / interpolation between a model and its coarser one.  Host code (no
// kernels); the matrices go to the device through the C ABI like the 3D cantilever's.
//
// Mesh (the reference takes it from the unvendored UniformCartesianMesh2D): (nx+1) x (ny+1)
// nodes, node id = ix + (nx+1) iy at (ix Lx/nx, iy Ly/ny); element e = ex + nx ey has the corners
// (ex,ey), (ex+1,ey), (ex+1,ey+1), (ex,ey+1), counter-clockwise as the plate element's rectangle
// check requires (PlateElementRectangle4Nodes3Dofs.cs:33-39) and the cantilever's Q4 matrix
// assumes.  Dofs are numbered node by node with the constrained ones left out: the plate has
// (w, rx, ry) per node and w fixed at supported nodes (FemPlate.cs:72-101), the cantilever
// (ux, uy) and both fixed at ix == 0 (FemCantileverBase.cs:137-158).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "mgx_internal.h"

namespace mgx {

namespace {

// per node: index of its first free dof and the mask of its constrained dofs (bit d: dof d)
struct GridDofs {
    int nx = 0, ny = 0, dpn = 0;
    std::vector<int32_t> first;
    std::vector<uint8_t> fixed;
    int32_t nfree = 0;
    int64_t id(int ix, int iy) const { return ix + (int64_t)(nx + 1) * iy; }
    int ndofs(int64_t node) const
    {
        int k = 0;
        for (int d = 0; d < dpn; ++d) k += !((fixed[node] >> d) & 1);
        return k;
    }
    // free index of dof d of a node, -1 when constrained
    int32_t dof(int64_t node, int d) const
    {
        if ((fixed[node] >> d) & 1) return -1;
        int32_t k = first[node];
        for (int j = 0; j < d; ++j) k += !((fixed[node] >> j) & 1);
        return k;
    }
    void number()
    {
        nfree = 0;
        for (size_t n = 0; n < first.size(); ++n) { first[n] = nfree; nfree += ndofs((int64_t)n); }
    }
};

// Supports of FemPlate.cs:181-196: w fixed on every node within 0.1 min(Lx/(nx-1), Ly/(ny-1))
// of x = 0, x = Lx/2, y = 0, y = Ly/2 or y = Ly (FindNodesWithX/Y of the test NodeLocator: a
// strict |coordinate - target| < tolerance)
GridDofs plate_dofs(int nx, int ny, double Lx, double Ly)
{
    GridDofs P;
    P.nx = nx; P.ny = ny; P.dpn = 3;
    const int64_t nn = (int64_t)(nx + 1) * (ny + 1);
    P.first.resize(nn);
    P.fixed.resize(nn);
    const double tol = 0.1 * std::fmin(Lx / (nx - 1), Ly / (ny - 1));
    const double hx = Lx / nx, hy = Ly / ny;
    for (int iy = 0; iy <= ny; ++iy) {
        const double y = iy * hy;
        const bool yline = std::fabs(y) < tol || std::fabs(y - Ly / 2) < tol || std::fabs(y - Ly) < tol;
        for (int ix = 0; ix <= nx; ++ix) {
            const double x = ix * hx;
            P.fixed[P.id(ix, iy)] = yline || std::fabs(x) < tol || std::fabs(x - Lx / 2) < tol;
        }
    }
    P.number();
    return P;
}

// FemCantileverBase.cs:137-158 on a 2D mesh: (ux, uy) per node, every node with ix == 0 clamped
GridDofs cant2d_dofs(int nx, int ny)
{
    GridDofs P;
    P.nx = nx; P.ny = ny; P.dpn = 2;
    const int64_t nn = (int64_t)(nx + 1) * (ny + 1);
    P.first.resize(nn);
    P.fixed.resize(nn);
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) P.fixed[P.id(ix, iy)] = ix == 0 ? 3 : 0;
    P.number();
    return P;
}

// The pattern of K_ff on the 9-node stencil: a row couples the free dofs of the up to 9 nodes
// around its node, in node id order (sorted columns).  nboff[9 n + slot] = the column offset of
// neighbour slot (dy + 1) 3 + (dx + 1) within the rows of node n (-1: outside the mesh).
void stencil_pattern(const GridDofs &P, HostCsr &K, std::vector<int32_t> &nboff)
{
    const int nx = P.nx, ny = P.ny;
    const int64_t nn = (int64_t)(nx + 1) * (ny + 1);
    std::vector<int32_t> rowlen(nn, 0);
    nboff.assign((size_t)nn * 9, -1);
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) {
            const int64_t n = P.id(ix, iy);
            int32_t off = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int jx = ix + dx, jy = iy + dy;
                    if (jx < 0 || jx > nx || jy < 0 || jy > ny) continue;
                    nboff[(size_t)n * 9 + (dy + 1) * 3 + (dx + 1)] = off;
                    off += P.ndofs(P.id(jx, jy));
                }
            rowlen[n] = off;
        }
    K.nrows = K.ncols = P.nfree;
    K.rowptr.assign((size_t)P.nfree + 1, 0);
    for (int64_t n = 0; n < nn; ++n)
        for (int d = 0; d < P.ndofs(n); ++d) K.rowptr[P.first[n] + d + 1] = rowlen[n];
    for (int32_t r = 0; r < P.nfree; ++r) K.rowptr[r + 1] += K.rowptr[r];
    K.col.assign(K.rowptr[P.nfree], 0);
    K.val.assign(K.rowptr[P.nfree], 0.0);
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) {
            const int64_t n = P.id(ix, iy);
            for (int d = 0; d < P.ndofs(n); ++d) {
                int32_t k = K.rowptr[P.first[n] + d];
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int jx = ix + dx, jy = iy + dy;
                        if (jx < 0 || jx > nx || jy < 0 || jy > ny) continue;
                        const int64_t m = P.id(jx, jy);
                        for (int e = 0; e < P.ndofs(m); ++e) K.col[k++] = P.first[m] + e;
                    }
            }
        }
}

// Adds every element's matrix (times its density, when given) into its slots in ascending element
// order e = ex + nx ey -- the order in which the reference's element loop accumulates each entry
// into its dictionary-of-keys matrix.  Ke is (4 dpn) x (4 dpn) over the corners (0,0), (1,0),
// (1,1), (0,1), dofs of a corner in order.
void scatter_elements(const GridDofs &P, const std::vector<int32_t> &nboff, const double *Ke, const double *dens,
                      HostCsr &K)
{
    static const int kx[4] = {0, 1, 1, 0}, ky[4] = {0, 0, 1, 1};
    const int dpn = P.dpn, ne = 4 * dpn;
    for (int ey = 0; ey < P.ny; ++ey)
        for (int ex = 0; ex < P.nx; ++ex) {
            const double de = dens ? dens[ex + (int64_t)P.nx * ey] : 1.0;
            int64_t node[4];
            int32_t g[12];
            for (int c = 0; c < 4; ++c) {
                node[c] = P.id(ex + kx[c], ey + ky[c]);
                for (int d = 0; d < dpn; ++d) g[dpn * c + d] = P.dof(node[c], d);
            }
            for (int i = 0; i < ne; ++i) {
                if (g[i] < 0) continue;
                const int ci = i / dpn;
                for (int j = 0; j < ne; ++j) {
                    if (g[j] < 0) continue;
                    const int cj = j / dpn;
                    const int slot = (ky[cj] - ky[ci] + 1) * 3 + (kx[cj] - kx[ci] + 1);
                    const int32_t col0 = nboff[(size_t)node[ci] * 9 + slot];
                    const int32_t k = K.rowptr[g[i]] + col0 + (g[j] - P.first[node[cj]]);
                    K.val[k] += dens ? de * Ke[i * ne + j] : Ke[i * ne + j];
                }
            }
        }
}

inline int coarser_count(int n) { return n > 1 ? n >> 1 : n; }     // IGeometricMultigridModel.cs:54-58

}  // namespace

// Element stiffness of PlateElementRectangle4Nodes3Dofs (:134-222): the Adini-Clough-Melosh
// rectangle of half-sides a, b in closed form, the reference's coefficient table evaluated in its
// own operation order (so the bits agree), times D / (60 a b), D = E t^3 / (12 (1 - nu^2)) (:42).
void plate_element(double a, double b, double E, double nu, double t, double Ke[144])
{
    const double D = E * std::pow(t, 3) / (12 * (1 - std::pow(nu, 2)));
    const double p = std::pow(a / b, 2), v = nu;
    // coefficients of the first three columns of the lower triangle
    const double c11 = 60 * p + 60 / p + 42 - 12 * v, c21 = b * (60 * p + 6 + 24 * v);
    const double c31 = a * (60 / p + 6 + 24 * v), c41 = 30 * p - 60 / p - 42 + 12 * v;
    const double c71 = -30 * p - 30 / p + 42 - 12 * v, c101 = -60 * p + 30 / p - 42 + 12 * v;
    const double c22 = b * b * (80 * p + 16 - 16 * v), c32 = 60 * v * a * b, c42 = b * (30 * p - 6 - 24 * v);
    const double c52 = b * b * (40 * p - 16 + 16 * v), c72 = b * (-30 * p + 6 - 6 * v);
    const double c82 = b * b * (20 * p + 4 - 4 * v), c102 = b * (-60 * p - 6 + 6 * v);
    const double c112 = b * b * (40 * p - 4 + 4 * v);
    const double c33 = a * a * (80 / p + 16 - 16 * v), c43 = a * (-60 / p - 6 + 6 * v);
    const double c63 = a * a * (40 / p - 4 + 4 * v), c73 = a * (-30 / p + 6 - 6 * v);
    const double c93 = a * a * (20 / p + 4 - 4 * v), c103 = a * (30 / p - 6 - 24 * v);
    const double c123 = a * a * (40 / p - 16 + 16 * v);
    // the three zero coefficients of the table, and the ones it defines through others
    const double z = 0, c53 = z, c83 = z, c113 = z;
    const double c51 = c42, c61 = -c43, c81 = -c72, c91 = -c73, c111 = -c102, c121 = c103;
    const double c62 = -c53, c92 = c83, c122 = -c113;
    const double Lw[12][12] = {
        {c11},
        {c21, c22},
        {c31, c32, c33},
        {c41, c42, c43, c11},
        {c51, c52, c53, c21, c22},
        {c61, c62, c63, -c31, -c32, c33},
        {c71, c72, c73, c101, c102, -c103, c11},
        {c81, c82, c83, c111, c112, -c113, -c21, c22},
        {c91, c92, c93, -c121, -c122, c123, -c31, c32, c33},
        {c101, c102, c103, c71, c72, -c73, c41, -c42, -c43, c11},
        {c111, c112, c113, c81, c82, -c83, -c51, c52, c53, -c21, c22},
        {c121, c122, c123, -c91, -c92, c93, -c61, c62, c63, c31, -c32, c33},
    };
    const double scale = D / (60 * a * b);
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j <= i; ++j) Ke[i * 12 + j] = Ke[j * 12 + i] = Lw[i][j] * scale;
}

int32_t plate_free_dofs(int nx, int ny, double Lx, double Ly, std::vector<int32_t> *dofs)
{
    const GridDofs P = plate_dofs(nx, ny, Lx, Ly);
    if (dofs) {
        dofs->clear();
        dofs->reserve(P.nfree);
        for (int64_t n = 0; n < (int64_t)P.first.size(); ++n)
            for (int d = 0; d < 3; ++d)
                if (P.dof(n, d) >= 0) dofs->push_back((int32_t)(3 * n + d));
    }
    return P.nfree;
}

// K_ff with sorted columns and the load vector of FemPlate.CreateLinearSystem (FemPlate.cs:41-70).
// The reference assembles element by element into a dictionary-of-keys matrix
// (DokRowMajorMatrixAssembler.cs:30-57, AddSubmatrixSymmetric DokRowMajor.cs:244-270: each pair
// of an element's free dofs added once), and the load vector element by element
// (PlateElementVectorAssembler.cs:14-30, the first element's nodal forces for all): so every
// entry is the sum of its element contributions in ascending element order.  Here the pattern
// is laid out first (a row couples the free dofs of the up to 9 nodes around its node, in node
// id order), then the elements are added in that order straight into their slots.
void plate_system(int nx, int ny, double Lx, double Ly, double t, double E, double nu, double q, HostCsr &K,
                  std::vector<double> *rhs)
{
    const GridDofs P = plate_dofs(nx, ny, Lx, Ly);
    std::vector<int32_t> nboff;
    stencil_pattern(P, K, nboff);
    const double a = 0.5 * (Lx / nx), b = 0.5 * (Ly / ny);
    double Ke[144];
    plate_element(a, b, E, nu, t, Ke);
    scatter_elements(P, nboff, Ke, nullptr, K);
    if (!rhs) return;
    // nodal forces of a uniform load q (ConvertBodyForcesToNodal, PlateElementRectangle4Nodes3Dofs.cs:104-116),
    // summed element by element
    const double fr[12] = {1, b / 3, a / 3, 1, b / 3, -a / 3, 1, -b / 3, -a / 3, 1, -b / 3, a / 3};
    const double qs = q * a * b;
    rhs->assign(P.nfree, 0.0);
    static const int kx[4] = {0, 1, 1, 0}, ky[4] = {0, 0, 1, 1};
    for (int ey = 0; ey < ny; ++ey)
        for (int ex = 0; ex < nx; ++ex)
            for (int c = 0; c < 4; ++c)
                for (int d = 0; d < 3; ++d) {
                    const int32_t g = P.dof(P.id(ex + kx[c], ey + ky[c]), d);
                    if (g >= 0) (*rhs)[g] += fr[3 * c + d] * qs;
                }
}

namespace {

// IGeometricMultigridModel.cs:76-194 on two axes: the cone weight 1 - |w - c| / s of each fine
// node in the window of each coarse node (s = fine / coarse element ratio per axis), restriction
// rows normalised by the window sum, interpolation rows by the sum of the weights a fine node
// receives (accumulated in the reference's loop order: coarse x outer, y inner; fine x outer, y
// inner), the same weight for each dof of a node, then the free-dof submatrices -- the
// constrained dofs count in both sums.  F is the model's dof map, Cd its coarser model's.
void transfer2d(const GridDofs &F, const GridDofs &Cd, HostCsr &R, HostCsr &P)
{
    const int ne[2] = {F.nx, F.ny};
    int nf[2], nc[2];
    double s[2];
    for (int i = 0; i < 2; ++i) {
        const int c = coarser_count(ne[i]);
        s[i] = (double)ne[i] / c;
        nf[i] = ne[i] + 1;
        nc[i] = c + 1;
    }
    auto window = [&](int i, int a, int &lo, int &hi) {
        const double c = a * s[i];
        lo = std::max(0, (int)std::ceil(std::nextafter(c - s[i], INFINITY)));
        hi = std::min(nf[i], (int)std::floor(std::nextafter(c + s[i], -INFINITY)) + 1);
    };
    auto weight = [&](int w0, int w1, int a0, int a1) {
        const double t0 = (w0 - a0 * s[0]) / s[0], t1 = (w1 - a1 * s[1]) / s[1];
        double r = 0;
        r += t0 * t0;
        r += t1 * t1;
        return 1 - std::sqrt(r);
    };
    const int64_t nfn = (int64_t)nf[0] * nf[1], ncn = (int64_t)nc[0] * nc[1];
    std::vector<double> sumF(nfn, 0.0), sumC(ncn, 0.0);
    for (int a0 = 0; a0 < nc[0]; ++a0)
        for (int a1 = 0; a1 < nc[1]; ++a1) {
            int l0, h0, l1, h1;
            window(0, a0, l0, h0);
            window(1, a1, l1, h1);
            double sr = 0;
            for (int w0 = l0; w0 < h0; ++w0)
                for (int w1 = l1; w1 < h1; ++w1) {
                    const double f = weight(w0, w1, a0, a1);
                    sr += f;
                    sumF[w0 + (int64_t)nf[0] * w1] += f;
                }
            sumC[a0 + (int64_t)nc[0] * a1] = sr;
        }
    // the coarse nodes reaching fine node index f along each axis form one contiguous range
    std::vector<int> alo[2], ahi[2];
    for (int i = 0; i < 2; ++i) {
        alo[i].assign(nf[i], nc[i]);
        ahi[i].assign(nf[i], -1);
        for (int a = 0; a < nc[i]; ++a) {
            int l, h;
            window(i, a, l, h);
            for (int f = l; f < h; ++f) { alo[i][f] = std::min(alo[i][f], a); ahi[i][f] = std::max(ahi[i][f], a); }
        }
    }
    R.nrows = Cd.nfree; R.ncols = F.nfree;
    R.rowptr.assign((size_t)Cd.nfree + 1, 0); R.col.clear(); R.val.clear();
    int32_t row = 0;
    for (int a1 = 0; a1 < nc[1]; ++a1)
        for (int a0 = 0; a0 < nc[0]; ++a0) {
            const int64_t cn = a0 + (int64_t)nc[0] * a1;
            int l0, h0, l1, h1;
            window(0, a0, l0, h0);
            window(1, a1, l1, h1);
            for (int d = 0; d < F.dpn; ++d) {
                if (Cd.dof(cn, d) < 0) continue;
                R.rowptr[row++] = (int32_t)R.col.size();
                for (int w1 = l1; w1 < h1; ++w1)
                    for (int w0 = l0; w0 < h0; ++w0) {
                        const int32_t c = F.dof(w0 + (int64_t)nf[0] * w1, d);
                        if (c < 0) continue;
                        R.col.push_back(c);
                        R.val.push_back(weight(w0, w1, a0, a1) / sumC[cn]);
                    }
            }
        }
    R.rowptr[row] = (int32_t)R.col.size();
    P.nrows = F.nfree; P.ncols = Cd.nfree;
    P.rowptr.assign((size_t)F.nfree + 1, 0); P.col.clear(); P.val.clear();
    row = 0;
    for (int w1 = 0; w1 < nf[1]; ++w1)
        for (int w0 = 0; w0 < nf[0]; ++w0) {
            const int64_t fn = w0 + (int64_t)nf[0] * w1;
            for (int d = 0; d < F.dpn; ++d) {
                if (F.dof(fn, d) < 0) continue;
                P.rowptr[row++] = (int32_t)P.col.size();
                for (int a1 = alo[1][w1]; a1 <= ahi[1][w1]; ++a1)
                    for (int a0 = alo[0][w0]; a0 <= ahi[0][w0]; ++a0) {
                        const int32_t c = Cd.dof(a0 + (int64_t)nc[0] * a1, d);
                        if (c < 0) continue;
                        P.col.push_back(c);
                        P.val.push_back(weight(w0, w1, a0, a1) / sumF[fn]);
                    }
            }
        }
    P.rowptr[row] = (int32_t)P.col.size();
}

}  // namespace

void plate_transfer(int nx, int ny, double Lx, double Ly, HostCsr &R, HostCsr &P)
{
    const GridDofs F = plate_dofs(nx, ny, Lx, Ly);
    const GridDofs Cd = plate_dofs(coarser_count(nx), coarser_count(ny), Lx, Ly);
    transfer2d(F, Cd, R, P);
}

// ---------------------------------------------------------------- plane-stress cantilever
// Element stiffness of FemCantilever2D (FemCantilever2D.cs:29-49, the 4-node element of Sigmund's
// 99-line topology optimization code): eight distinct entries placed by the reference's index
// table, each times E / (1 - nu^2); dofs (ux, uy) of the corners (0,0), (1,0), (1,1), (0,1).
void cant2d_element(double E, double nu, double Ke[64])
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

// K_ff of FemCantileverBase.CreateLinearSystem (:87-104 assembly with the element densities,
// :137-158 free dofs) for the 2D mesh, sorted columns; dens = nullptr: every density 1
void cant2d_system(int nx, int ny, double E, double nu, const double *dens, HostCsr &K)
{
    const GridDofs P = cant2d_dofs(nx, ny);
    std::vector<int32_t> nboff;
    stencil_pattern(P, K, nboff);
    double Ke[64];
    cant2d_element(E, nu, Ke);
    scatter_elements(P, nboff, Ke, dens, K);
}

void cant2d_transfer(int nx, int ny, HostCsr &R, HostCsr &P)
{
    transfer2d(cant2d_dofs(nx, ny), cant2d_dofs(coarser_count(nx), coarser_count(ny)), R, P);
}

}  // namespace mgx
