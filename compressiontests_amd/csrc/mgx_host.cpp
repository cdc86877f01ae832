// mgx_host.cpp -- host-side setup of libmgx.so: element stiffness, transfer operators,
// colouring, coarse-level inverse, DuVi encoder, workload RNG.  Runs once per solver
// (the reference does the same work in C# at Initialize); nothing here is on the solve path.
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <algorithm>

#include "mgx_internal.h"

namespace mgx {

// ------------------------------------------------------------------ element stiffness
// FemCantilever3D.cs:32-120 (top3d closed form).  Same operation order as the reference:
// k = (A0 * 1 + A1 * nu) * (1/72), then the blocks scaled by E / ((nu + 1)(1 - 2 nu)).
void ke_hex8(double E, double nu, double *Ke)
{
    static const double A0[14] = {32, 6, -8, 6, -6, 4, 3, -6, -10, 3, -3, -3, -4, -8};
    static const double A1[14] = {-48, 0, 0, -24, 24, 0, 0, 0, 12, -12, 0, 12, 12, 12};
    double k[15];
    const double s72 = 1.0 / 72.0;
    k[0] = 0;
    for (int i = 0; i < 14; ++i) {
        double t = 0.0;   // two roundings, no contraction (file built with -ffp-contract=off)
        t = t + A0[i] * 1.0;
        t = t + A1[i] * nu;
        k[i + 1] = t * s72;
    }
    static const int K1[36] = {1, 2, 2, 3, 5, 5, 2, 1, 2, 4, 6, 7, 2, 2, 1, 4, 7, 6, 3, 4, 4, 1, 8, 8, 5, 6, 7, 8, 1, 2, 5, 7, 6, 8, 2, 1};
    static const int K2[36] = {9, 8, 12, 6, 4, 7, 8, 9, 12, 5, 3, 5, 10, 10, 13, 7, 4, 6, 6, 5, 11, 9, 2, 10, 4, 3, 5, 2, 9, 12, 11, 4, 6, 12, 10, 13};
    static const int K3[36] = {6, 7, 4, 9, 12, 8, 7, 6, 4, 10, 13, 10, 5, 5, 3, 8, 12, 9, 9, 10, 2, 6, 11, 5, 12, 13, 10, 11, 6, 4, 2, 12, 9, 4, 5, 3};
    static const int K4[36] = {14, 11, 11, 13, 10, 10, 11, 14, 11, 12, 9, 8, 11, 11, 14, 12, 8, 9, 13, 12, 12, 14, 7, 7, 10, 9, 8, 7, 14, 11, 10, 8, 9, 7, 11, 14};
    static const int K5[36] = {1, 2, 8, 3, 5, 4, 2, 1, 8, 4, 6, 11, 8, 8, 1, 5, 11, 6, 3, 4, 5, 1, 8, 2, 5, 6, 11, 8, 1, 8, 4, 11, 6, 2, 8, 1};
    static const int K6[36] = {14, 11, 7, 13, 10, 12, 11, 14, 7, 12, 9, 2, 7, 7, 14, 10, 2, 9, 13, 12, 10, 14, 7, 11, 10, 9, 2, 7, 14, 7, 12, 2, 9, 11, 7, 14};
    // Ke = [[K1, K2, K3, K4], [K2', K5, K6, K3'], [K3', K6, K5', K2'], [K4, K3, K2, K1']]
    const int *blk[4][4] = {{K1, K2, K3, K4}, {K2, K5, K6, K3}, {K3, K6, K5, K2}, {K4, K3, K2, K1}};
    const bool tr[4][4] = {{false, false, false, false}, {true, false, false, true}, {true, false, true, true}, {false, false, false, true}};
    const double scale = E / ((nu + 1) * (1 - 2 * nu));
    for (int bi = 0; bi < 4; ++bi)
        for (int bj = 0; bj < 4; ++bj)
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j) {
                    const int idx = tr[bi][bj] ? blk[bi][bj][j * 6 + i] : blk[bi][bj][i * 6 + j];
                    Ke[(bi * 6 + i) * 24 + bj * 6 + j] = k[idx] * scale;
                }
}

int64_t cantilever_nnz(int nx, int ny, int nz)
{
    auto cnt = [](int i, int lo, int hi) { int c = 0; for (int d = -1; d <= 1; ++d) c += (i + d >= lo && i + d <= hi); return c; };
    int64_t n = 0;
    for (int iz = 0; iz <= nz; ++iz)
        for (int iy = 0; iy <= ny; ++iy) {
            const int64_t yz = (int64_t)cnt(iy, 0, ny) * cnt(iz, 0, nz);
            for (int ix = 1; ix <= nx; ++ix) n += 9 * cnt(ix, 1, nx) * yz;
        }
    return n;
}

static inline int64_t free_node(int ix, int iy, int iz, int nx, int ny)
{
    return ((int64_t)iz * (ny + 1) + iy) * nx + (ix - 1);
}

// ------------------------------------------------------------------ reference R / P
// IGeometricMultigridModel.cs:76-194.  Window per axis [ceil(nextup(c - s)), floor(nextdown(c + s)) + 1)
// around the coarse node at c = a * s (s = fine/coarse element ratio); weight
// 1 - sqrt(sum ((w - c)/s)^2); R rows normalised by the window sum (summed in the loop order
// x, z, y), P rows by the per-fine-node sum accumulated in the coarse loop order x, z, y.
namespace {
struct RPDims {
    int nf[3], nc[3];
    double s[3];
};
inline int coarser(int n) { return n > 1 ? n >> 1 : n; }
inline void window(const RPDims &d, int i, int a, int &lo, int &hi, double &c)
{
    c = a * d.s[i];
    lo = (int)std::ceil(std::nextafter(c - d.s[i], INFINITY));
    if (lo < 0) lo = 0;
    hi = (int)std::floor(std::nextafter(c + d.s[i], -INFINITY)) + 1;
    if (hi > d.nf[i]) hi = d.nf[i];
}
inline double influence(const RPDims &d, const int *w, const double *c)
{
    double s = 0;
    for (int i = 0; i < 3; ++i) { const double t = (w[i] - c[i]) / d.s[i]; s += t * t; }
    return 1 - std::sqrt(s);
}
}  // namespace

void transfer_reference(int nx, int ny, int nz, HostCsr &R, HostCsr &P)
{
    RPDims d;
    const int nfe[3] = {nx, ny, nz};
    for (int i = 0; i < 3; ++i) {
        const int c = coarser(nfe[i]);
        d.s[i] = (double)nfe[i] / c;
        d.nf[i] = nfe[i] + 1;
        d.nc[i] = c + 1;
    }
    const int cnx = d.nc[0] - 1, cny = d.nc[1] - 1;
    const int64_t nfn = (int64_t)d.nf[0] * d.nf[1] * d.nf[2], ncn = (int64_t)d.nc[0] * d.nc[1] * d.nc[2];
    std::vector<double> sumP(nfn, 0.0), sumR(ncn, 0.0);
    int a[3], w[3], lo[3], hi[3];
    double c[3];
    for (a[0] = 0; a[0] < d.nc[0]; ++a[0])
        for (a[2] = 0; a[2] < d.nc[2]; ++a[2])
            for (a[1] = 0; a[1] < d.nc[1]; ++a[1]) {
                for (int i = 0; i < 3; ++i) window(d, i, a[i], lo[i], hi[i], c[i]);
                double sr = 0;
                for (w[0] = lo[0]; w[0] < hi[0]; ++w[0])
                    for (w[2] = lo[2]; w[2] < hi[2]; ++w[2])
                        for (w[1] = lo[1]; w[1] < hi[1]; ++w[1]) {
                            const double f = influence(d, w, c);
                            sr += f;
                            sumP[w[0] + (int64_t)d.nf[0] * (w[1] + (int64_t)d.nf[1] * w[2])] += f;
                        }
                sumR[a[0] + (int64_t)d.nc[0] * (a[1] + (int64_t)d.nc[1] * a[2])] = sr;
            }
    const int64_t nfree_c = 3LL * cnx * d.nc[1] * d.nc[2], nfree_f = 3LL * nx * d.nf[1] * d.nf[2];
    R.nrows = (int32_t)nfree_c; R.ncols = (int32_t)nfree_f;
    R.rowptr.assign(nfree_c + 1, 0); R.col.clear(); R.val.clear();
    int64_t row = 0;
    for (a[2] = 0; a[2] < d.nc[2]; ++a[2])
        for (a[1] = 0; a[1] < d.nc[1]; ++a[1])
            for (a[0] = 1; a[0] < d.nc[0]; ++a[0]) {
                for (int i = 0; i < 3; ++i) window(d, i, a[i], lo[i], hi[i], c[i]);
                const double sr = sumR[a[0] + (int64_t)d.nc[0] * (a[1] + (int64_t)d.nc[1] * a[2])];
                for (int dof = 0; dof < 3; ++dof) {
                    R.rowptr[row++] = (int32_t)R.col.size();
                    for (w[2] = lo[2]; w[2] < hi[2]; ++w[2])
                        for (w[1] = lo[1]; w[1] < hi[1]; ++w[1])
                            for (w[0] = std::max(lo[0], 1); w[0] < hi[0]; ++w[0]) {
                                R.col.push_back((int32_t)(3 * free_node(w[0], w[1], w[2], nx, ny) + dof));
                                R.val.push_back(influence(d, w, c) / sr);
                            }
                }
            }
    R.rowptr[row] = (int32_t)R.col.size();
    // P: the coarse nodes whose window holds fine node w, per axis a contiguous range
    std::vector<int> alo[3], ahi[3];
    for (int i = 0; i < 3; ++i) {
        alo[i].assign(d.nf[i], d.nc[i]); ahi[i].assign(d.nf[i], -1);
        for (int aa = 0; aa < d.nc[i]; ++aa) {
            int l, h; double cc;
            window(d, i, aa, l, h, cc);
            for (int f = l; f < h; ++f) { alo[i][f] = std::min(alo[i][f], aa); ahi[i][f] = std::max(ahi[i][f], aa); }
        }
    }
    P.nrows = (int32_t)nfree_f; P.ncols = (int32_t)nfree_c;
    P.rowptr.assign(nfree_f + 1, 0); P.col.clear(); P.val.clear();
    row = 0;
    for (w[2] = 0; w[2] < d.nf[2]; ++w[2])
        for (w[1] = 0; w[1] < d.nf[1]; ++w[1])
            for (w[0] = 1; w[0] < d.nf[0]; ++w[0]) {
                const double sp = sumP[w[0] + (int64_t)d.nf[0] * (w[1] + (int64_t)d.nf[1] * w[2])];
                for (int dof = 0; dof < 3; ++dof) {
                    P.rowptr[row++] = (int32_t)P.col.size();
                    for (a[2] = alo[2][w[2]]; a[2] <= ahi[2][w[2]]; ++a[2])
                        for (a[1] = alo[1][w[1]]; a[1] <= ahi[1][w[1]]; ++a[1])
                            for (a[0] = std::max(alo[0][w[0]], 1); a[0] <= ahi[0][w[0]]; ++a[0]) {
                                for (int i = 0; i < 3; ++i) c[i] = a[i] * d.s[i];
                                P.col.push_back((int32_t)(3 * free_node(a[0], a[1], a[2], cnx, cny) + dof));
                                P.val.push_back(influence(d, w, c) / sp);
                            }
                }
            }
    P.rowptr[row] = (int32_t)P.col.size();
}

void csr_transpose(const HostCsr &A, HostCsr &T)
{
    T.nrows = A.ncols; T.ncols = A.nrows;
    T.rowptr.assign(A.ncols + 1, 0);
    const int64_t nnz = A.rowptr[A.nrows];
    for (int64_t k = 0; k < nnz; ++k) T.rowptr[A.col[k] + 1]++;
    for (int32_t c = 0; c < A.ncols; ++c) T.rowptr[c + 1] += T.rowptr[c];
    T.col.resize(nnz); T.val.resize(nnz);
    std::vector<int32_t> fill(T.rowptr.begin(), T.rowptr.end() - 1);
    for (int32_t r = 0; r < A.nrows; ++r)
        for (int32_t k = A.rowptr[r]; k < A.rowptr[r + 1]; ++k) {
            const int32_t p = fill[A.col[k]]++;
            T.col[p] = r; T.val[p] = A.val[k];
        }
}

// Exact trilinear interpolation (weights 1, 1/2, 1/4, 1/8) and R = P^T.
void transfer_trilinear(int nx, int ny, int nz, HostCsr &R, HostCsr &P)
{
    const int cnx = nx / 2, cny = ny / 2, cnz = nz / 2;
    P.nrows = 3 * nx * (ny + 1) * (nz + 1);
    P.ncols = 3 * cnx * (cny + 1) * (cnz + 1);
    P.rowptr.assign(P.nrows + 1, 0); P.col.clear(); P.val.clear();
    P.col.reserve((size_t)P.nrows * 8); P.val.reserve((size_t)P.nrows * 8);
    int64_t row = 0;
    for (int iz = 0; iz <= nz; ++iz)
        for (int iy = 0; iy <= ny; ++iy)
            for (int ix = 1; ix <= nx; ++ix) {
                const int f[3] = {ix, iy, iz};
                int c0[3], cn[3]; double wt[3];
                for (int i = 0; i < 3; ++i) {
                    if (f[i] & 1) { c0[i] = (f[i] - 1) / 2; cn[i] = 2; wt[i] = 0.5; }
                    else { c0[i] = f[i] / 2; cn[i] = 1; wt[i] = 1.0; }
                }
                for (int dof = 0; dof < 3; ++dof) {
                    P.rowptr[row++] = (int32_t)P.col.size();
                    for (int cz = c0[2]; cz < c0[2] + cn[2]; ++cz)
                        for (int cy = c0[1]; cy < c0[1] + cn[1]; ++cy)
                            for (int cx = c0[0]; cx < c0[0] + cn[0]; ++cx) {
                                if (cx == 0) continue;
                                P.col.push_back((int32_t)(3 * free_node(cx, cy, cz, cnx, cny) + dof));
                                P.val.push_back(wt[0] * wt[1] * wt[2]);
                            }
                }
            }
    P.rowptr[row] = (int32_t)P.col.size();
    csr_transpose(P, R);
}

// ------------------------------------------------------------------ colouring
// Greedy distance-1 colouring of node blocks (bs rows each), natural order.
int block_coloring(int32_t nblocks, int bs, const int32_t *rp, const int32_t *ci, std::vector<int32_t> &color)
{
    color.assign(nblocks, -1);
    int ncol = 0;
    for (int32_t k = 0; k < nblocks; ++k) {
        uint64_t used[4] = {0, 0, 0, 0};
        for (int a = 0; a < bs; ++a) {
            const int32_t r = k * bs + a;
            for (int32_t q = rp[r]; q < rp[r + 1]; ++q) {
                const int32_t nb = ci[q] / bs;
                if (nb != k && color[nb] >= 0 && color[nb] < 256) used[color[nb] >> 6] |= 1ULL << (color[nb] & 63);
            }
        }
        int c = 0;
        while (c < 256 && ((used[c >> 6] >> (c & 63)) & 1)) ++c;
        color[k] = c;
        ncol = std::max(ncol, c + 1);
    }
    return ncol;
}

// Node groups for Gauss-Seidel when a level's nodes carry different numbers of free dofs (the
// plate: w is constrained on its support lines): a group is a maximal run of at most 3
// consecutive rows with identical column lists (the free dofs of one node couple to the same
// neighbours).  gptr[g] .. gptr[g+1]-1 are the rows of group g.
int32_t row_groups(int32_t n, const int32_t *rp, const int32_t *ci, std::vector<int32_t> &gptr)
{
    gptr.clear();
    for (int32_t r = 0; r < n;) {
        gptr.push_back(r);
        const int32_t len = rp[r + 1] - rp[r];
        int32_t e = r + 1;
        while (e < n && e - r < 3 && rp[e + 1] - rp[e] == len && std::equal(ci + rp[e], ci + rp[e + 1], ci + rp[r]))
            ++e;
        r = e;
    }
    gptr.push_back(n);
    return (int32_t)gptr.size() - 1;
}

// block_coloring over the groups (greedy, natural order)
int group_coloring(int32_t ngroups, const int32_t *gptr, const int32_t *rp, const int32_t *ci,
                   std::vector<int32_t> &color)
{
    std::vector<int32_t> group_of(gptr[ngroups]);
    for (int32_t g = 0; g < ngroups; ++g) std::fill(group_of.begin() + gptr[g], group_of.begin() + gptr[g + 1], g);
    color.assign(ngroups, -1);
    int ncol = 0;
    for (int32_t k = 0; k < ngroups; ++k) {
        uint64_t used[4] = {0, 0, 0, 0};
        for (int32_t r = gptr[k]; r < gptr[k + 1]; ++r)
            for (int32_t q = rp[r]; q < rp[r + 1]; ++q) {
                const int32_t nb = group_of[ci[q]];
                if (nb != k && color[nb] >= 0 && color[nb] < 256) used[color[nb] >> 6] |= 1ULL << (color[nb] & 63);
            }
        int c = 0;
        while (c < 256 && ((used[c >> 6] >> (c & 63)) & 1)) ++c;
        color[k] = c;
        ncol = std::max(ncol, c + 1);
    }
    return ncol;
}

// ------------------------------------------------------------------ coarse inverse
// A (n x n, row-major, SPD) -> A^-1 by Cholesky; columns of the inverse solved in parallel.
bool dense_spd_inverse(int32_t n, std::vector<double> &A)
{
    std::vector<double> L(A);
    for (int32_t j = 0; j < n; ++j) {
        double d = L[(size_t)j * n + j];
        for (int32_t k = 0; k < j; ++k) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        if (!(d > 0)) return false;
        d = std::sqrt(d);
        L[(size_t)j * n + j] = d;
        const double *Lj = &L[(size_t)j * n];
        for (int32_t i = j + 1; i < n; ++i) {
            double *Li = &L[(size_t)i * n];
            double s = Li[j];
            for (int32_t k = 0; k < j; ++k) s -= Li[k] * Lj[k];
            Li[j] = s / d;
        }
    }
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            std::vector<double> y(n), x(n);
            for (int32_t col = (int32_t)t; col < n; col += (int32_t)nt) {
                for (int32_t i = 0; i < n; ++i) {
                    double s = (i == col) ? 1.0 : 0.0;
                    const double *Li = &L[(size_t)i * n];
                    for (int32_t k = col; k < i; ++k) s -= Li[k] * y[k];
                    y[i] = (i < col) ? 0.0 : s / Li[i];
                }
                for (int32_t i = n - 1; i >= 0; --i) {
                    double s = y[i];
                    for (int32_t k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * x[k];
                    x[i] = s / L[(size_t)i * n + i];
                }
                for (int32_t i = 0; i < n; ++i) A[(size_t)i * n + col] = x[i];
            }
        });
    for (auto &t : th) t.join();
    return true;
}

// ------------------------------------------------------------------ DuVi encoder
// .NET Dictionary<double,int>(ToleranceComparer) lookup: candidates with the same 32-bit
// hash of Math.Round(v / (2 tol)), newest first; first with |key - v| < tol wins.
namespace {
inline int32_t net_double_hash(double d)
{
    uint64_t bits;
    std::memcpy(&bits, &d, 8);
    if (((bits - 1) & 0x7FFFFFFFFFFFFFFFULL) >= 0x7FF0000000000000ULL) bits &= 0x7FF0000000000000ULL;
    return (int32_t)((uint32_t)bits ^ (uint32_t)(bits >> 32));
}
// Entries of one hash value form a chain through `next`, newest first, from the slot of that
// hash in an open-addressing table (linear probing, kept at most half full): the same candidate
// order as the .NET Dictionary scan.
struct Dedup {
    double tol;
    std::vector<double> keys;
    std::vector<int32_t> next;       // previous entry with the same hash, -1 at the chain's end
    std::vector<int32_t> slot_hash;  // hash of the slot's chain
    std::vector<int32_t> slot_head;  // newest entry of the chain, -1 for a free slot
    uint64_t mask = 0;
    explicit Dedup(double t) : tol(t) { rehash(1 << 12); }
    static uint64_t mix(int32_t h) { return (uint64_t)(uint32_t)h * 0x9E3779B97F4A7C15ULL >> 17; }
    void rehash(uint64_t cap)
    {
        std::vector<int32_t> oh, ohead;
        oh.swap(slot_hash);
        ohead.swap(slot_head);
        slot_hash.assign(cap, 0);
        slot_head.assign(cap, -1);
        mask = cap - 1;
        for (size_t i = 0; i < ohead.size(); ++i) {
            if (ohead[i] < 0) continue;
            int32_t *hd = find(oh[i]);          // a free slot: every hash has one chain
            slot_hash[hd - slot_head.data()] = oh[i];
            *hd = ohead[i];
        }
    }
    // the head of the chain of hash h, or the free slot where it would go
    int32_t *find(int32_t h)
    {
        uint64_t i = mix(h) & mask;
        while (slot_head[i] >= 0 && slot_hash[i] != h) i = (i + 1) & mask;
        return &slot_head[i];
    }
    int32_t get(double v)
    {
        const int32_t h = tol > 0 ? net_double_hash(std::nearbyint(v / (2 * tol))) : net_double_hash(v);
        int32_t *head = find(h);
        for (int32_t e = *head; e >= 0; e = next[e]) {
            const double k = keys[e];
            if (tol > 0 ? std::fabs(k - v) < tol : (k == v || (k != k && v != v))) return e;
        }
        const int32_t idx = (int32_t)keys.size();
        keys.push_back(v);
        if (*head < 0) slot_hash[head - slot_head.data()] = h;
        next.push_back(*head);
        *head = idx;
        if (2 * (uint64_t)keys.size() > mask) rehash(2 * (mask + 1));
        return idx;
    }
};
}  // namespace

int duvi_encode(int32_t nrows, const int32_t *rp, const int32_t *ci, const double *v, double tol, int mode,
                mgx_duvi_host *out)
{
    std::memset(out, 0, sizeof *out);
    const int64_t nnz = rp[nrows];
    out->nrows = nrows;
    out->nnz = nnz;
    out->row_blocks = (int32_t *)std::malloc(sizeof(int32_t) * (nrows + 1));
    out->row_dist = (int32_t *)std::malloc(sizeof(int32_t) * (nrows + 1));
    out->dist = (uint8_t *)std::malloc(nnz > 0 ? nnz : 1);
    std::vector<int32_t> blocks;
    blocks.reserve(nnz / 8 + 16);
    // DU: OpenCLDuViGeometricMultigridSolver.cs:771-794
    int64_t vi = 0, vbeg = 0;
    for (int32_t i = 0;; ++i) {
        out->row_blocks[i] = (int32_t)blocks.size();
        out->row_dist[i] = (int32_t)vi;
        if (i == nrows) break;
        const int32_t from = rp[i], to = rp[i + 1];
        for (int32_t j = from; j < to; ++j) {
            int count = 0;
            blocks.push_back(ci[j]);
            for (; j + 1 < to && count < 255; ++j) {
                const int32_t diff = ci[j + 1] - ci[j] - 1;
                if (diff < 256) { ++count; ++vi; out->dist[vi] = (uint8_t)diff; }
                else break;
            }
            out->dist[vbeg] = (uint8_t)count;
            ++vi;
            vbeg = vi;
        }
    }
    out->nblocks = (int64_t)blocks.size();
    out->block_col = (int32_t *)std::malloc(sizeof(int32_t) * (blocks.empty() ? 1 : blocks.size()));
    std::memcpy(out->block_col, blocks.data(), sizeof(int32_t) * blocks.size());
    // VI: :798-829, with the intended 16-bit switch and raw-value fallback
    // mode 8 needs no table; mode 0 falls back to raw values once the table passes 65,536 entries,
    // so the deduplication can stop there
    Dedup dd(tol);
    std::vector<int32_t> idx(mode == 8 ? 0 : nnz);
    bool raw = mode == 8;
    for (int64_t k = 0; k < nnz && !raw; ++k) {
        idx[k] = dd.get(v[k]);
        if (mode == 0 && dd.keys.size() > 65536) raw = true;
    }
    const int64_t nu = (int64_t)dd.keys.size();
    int vib = raw ? 0 : (mode == 4 ? 4 : (nu <= 256 ? 1 : (nu <= 65536 ? 2 : 0)));
    out->vi_bytes = vib;
    if (vib == 0) {
        out->nuniq = nnz;
        out->values = (double *)std::malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
        std::memcpy(out->values, v, sizeof(double) * nnz);
        out->vi = nullptr;
    } else {
        out->nuniq = nu;
        out->values = (double *)std::malloc(sizeof(double) * (nu > 0 ? nu : 1));
        std::memcpy(out->values, dd.keys.data(), sizeof(double) * nu);
        out->vi = std::malloc((size_t)vib * (nnz > 0 ? nnz : 1));
        for (int64_t k = 0; k < nnz; ++k) {
            if (vib == 1) ((uint8_t *)out->vi)[k] = (uint8_t)idx[k];
            else if (vib == 2) ((uint16_t *)out->vi)[k] = (uint16_t)idx[k];
            else ((uint32_t *)out->vi)[k] = (uint32_t)idx[k];
        }
    }
    return MGX_OK;
}

// ------------------------------------------------------------------ RNG
// SplitMix64 seeding of xoshiro256** (Blackman & Vigna), U[0,1) = (next >> 11) * 2^-53.
namespace {
inline uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) { uint64_t sm = seed; for (auto &w : s) w = splitmix(sm); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};
}  // namespace

void rng_uniform(uint64_t seed, int64_t n, double *out)
{
    Xoshiro g(seed);
    for (int64_t i = 0; i < n; ++i) out[i] = g.uniform();
}

void rng_normal(uint64_t seed, int64_t n, double *out)
{
    Xoshiro g(seed);
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < n; ++i) {
        const double u1 = g.uniform(), u2 = g.uniform();
        out[i] = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(two_pi * u2);
    }
}

}  // namespace mgx
