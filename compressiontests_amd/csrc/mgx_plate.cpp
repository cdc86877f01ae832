// mgx_plate.cpp -- host generator of the reference's second workload, the Kirchhoff plate
// (FemPlate, tests/MGroup.LinearAlgebra.Tests/TestData/SparseLinearSystems/FemPlate.cs): its
// element matrix, the free-dof stiffness K_ff and load vector, the free-dof list, and the
// restriction / interpolation between a plate and its coarser model.  Host code (no kernels);
// the matrices go to the device through the C ABI like the cantilever's.
//
// Mesh (the reference takes it from the unvendored UniformCartesianMesh2D): (nx+1) x (ny+1)
// nodes, node id = ix + (nx+1) iy at (ix Lx/nx, iy Ly/ny); element (ex, ey) has the corners
// (ex,ey), (ex+1,ey), (ex+1,ey+1), (ex,ey+1), counter-clockwise as the element's rectangle
// check requires (PlateElementRectangle4Nodes3Dofs.cs:33-39); three dofs per node (w, rx, ry),
// numbered node by node with the constrained w of supported nodes left out (FemPlate.cs:72-101).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "mgx_internal.h"

namespace mgx {

namespace {

// per node: index of its first free dof and whether its w is constrained
struct PlateDofs {
    int nx = 0, ny = 0;
    std::vector<int32_t> first;
    std::vector<uint8_t> supported;
    int32_t nfree = 0;
    int64_t id(int ix, int iy) const { return ix + (int64_t)(nx + 1) * iy; }
    int ndofs(int64_t node) const { return supported[node] ? 2 : 3; }
    // free index of dof d (0 = w, 1 = rx, 2 = ry) of a node, -1 when constrained
    int32_t dof(int64_t node, int d) const
    {
        if (supported[node]) return d == 0 ? -1 : first[node] + d - 1;
        return first[node] + d;
    }
};

// Supports of FemPlate.cs:181-196: w fixed on every node within 0.1 min(Lx/(nx-1), Ly/(ny-1))
// of x = 0, x = Lx/2, y = 0, y = Ly/2 or y = Ly (FindNodesWithX/Y of the test NodeLocator: a
// strict |coordinate - target| < tolerance)
PlateDofs plate_dofs(int nx, int ny, double Lx, double Ly)
{
    PlateDofs P;
    P.nx = nx; P.ny = ny;
    const int64_t nn = (int64_t)(nx + 1) * (ny + 1);
    P.first.resize(nn);
    P.supported.resize(nn);
    const double tol = 0.1 * std::fmin(Lx / (nx - 1), Ly / (ny - 1));
    const double hx = Lx / nx, hy = Ly / ny;
    int32_t next = 0;
    for (int iy = 0; iy <= ny; ++iy) {
        const double y = iy * hy;
        const bool yline = std::fabs(y) < tol || std::fabs(y - Ly / 2) < tol || std::fabs(y - Ly) < tol;
        for (int ix = 0; ix <= nx; ++ix) {
            const double x = ix * hx;
            const int64_t n = P.id(ix, iy);
            P.supported[n] = yline || std::fabs(x) < tol || std::fabs(x - Lx / 2) < tol;
            P.first[n] = next;
            next += P.ndofs(n);
        }
    }
    P.nfree = next;
    return P;
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
    const PlateDofs P = plate_dofs(nx, ny, Lx, Ly);
    if (dofs) {
        dofs->clear();
        dofs->reserve(P.nfree);
        for (int64_t n = 0; n < (int64_t)P.first.size(); ++n)
            for (int d = P.supported[n] ? 1 : 0; d < 3; ++d) dofs->push_back((int32_t)(3 * n + d));
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
    const PlateDofs P = plate_dofs(nx, ny, Lx, Ly);
    const int64_t nn = (int64_t)(nx + 1) * (ny + 1);
    // per node: the column offset of each of its 3 x 3 neighbour nodes (-1: outside the mesh)
    std::vector<int32_t> nboff((size_t)nn * 9, -1), rowlen(nn, 0);
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
    const double a = 0.5 * (Lx / nx), b = 0.5 * (Ly / ny);
    double Ke[144];
    plate_element(a, b, E, nu, t, Ke);
    // nodal forces of a uniform load q (ConvertBodyForcesToNodal, PlateElementRectangle4Nodes3Dofs.cs:104-116)
    const double fr[12] = {1, b / 3, a / 3, 1, b / 3, -a / 3, 1, -b / 3, -a / 3, 1, -b / 3, a / 3};
    const double qs = q * a * b;
    if (rhs) rhs->assign(P.nfree, 0.0);
    static const int kx[4] = {0, 1, 1, 0}, ky[4] = {0, 0, 1, 1};
    for (int ey = 0; ey < ny; ++ey)
        for (int ex = 0; ex < nx; ++ex) {
            int64_t node[4];
            int32_t g[12];
            for (int c = 0; c < 4; ++c) {
                node[c] = P.id(ex + kx[c], ey + ky[c]);
                for (int d = 0; d < 3; ++d) g[3 * c + d] = P.dof(node[c], d);
            }
            for (int i = 0; i < 12; ++i) {
                if (g[i] < 0) continue;
                const int ci = i / 3;
                for (int j = 0; j < 12; ++j) {
                    if (g[j] < 0) continue;
                    const int cj = j / 3;
                    const int slot = (ky[cj] - ky[ci] + 1) * 3 + (kx[cj] - kx[ci] + 1);
                    const int32_t col0 = nboff[(size_t)node[ci] * 9 + slot];
                    const int32_t k = K.rowptr[g[i]] + col0 + (g[j] - P.first[node[cj]]);
                    K.val[k] += Ke[i * 12 + j];
                }
                if (rhs) (*rhs)[g[i]] += fr[i] * qs;
            }
        }
}

// IGeometricMultigridModel.cs:76-194 for a plate (two axes): the cone weight 1 - |w - c| / s of
// each fine node in the window of each coarse node (s = fine / coarse element ratio per axis),
// restriction rows normalised by the window sum, interpolation rows by the sum of the weights a
// fine node receives (accumulated in the reference's loop order: coarse x outer, y inner; fine x
// outer, y inner), the same weight for each of the three dofs, then the free-dof submatrices --
// the constrained dofs count in both sums.
void plate_transfer(int nx, int ny, double Lx, double Ly, HostCsr &R, HostCsr &P)
{
    const int ne[2] = {nx, ny};
    int nf[2], nc[2];
    double s[2];
    for (int i = 0; i < 2; ++i) {
        const int c = coarser_count(ne[i]);
        s[i] = (double)ne[i] / c;
        nf[i] = ne[i] + 1;
        nc[i] = c + 1;
    }
    const PlateDofs F = plate_dofs(nx, ny, Lx, Ly), Cd = plate_dofs(nc[0] - 1, nc[1] - 1, Lx, Ly);
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
            for (int d = 0; d < 3; ++d) {
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
            for (int d = 0; d < 3; ++d) {
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

}  // namespace mgx
