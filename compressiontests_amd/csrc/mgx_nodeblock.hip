// mgx_nodeblock.hip -- node-block (3x3) layout of the level matrices of the batching form, and
// its SpMM and multicolour Gauss-Seidel kernels.
//
// Every level matrix of the 3-dof hexahedral hierarchy has its rows in node triples (3q, 3q+1,
// 3q+2) sharing one column list made of whole neighbour nodes (columns 3j, 3j+1, 3j+2).  Stored
// as 3x3 blocks the matrix needs one column index per 9 values instead of one per value (FP32:
// 4.44 bytes per nonzero instead of 8; FP64: 8.44 instead of 12), and the lane that gathers a
// neighbour node's 3 x W values of X applies the whole block to them.
//
// Layout of one copy: positions p = 0 .. npos-1 (node[p] = the node at position p; no node
// array = the identity), blocks nrp[p] .. nrp[p+1]-1 of position p, ncol[b] = the neighbour node
// of block b, val[9b + 3a + c] = entry (row 3q+a, column 3j+c).  The SpMM copy is in node order
// with the blocks of a node in column order; the Gauss-Seidel copy lists the nodes colour by
// colour with each node's blocks as [own, earlier colours, later colours] and split[p] between
// the last two; the residual copy of a zero-start sweep is in node order with only [own, later
// colours].  val carries 16 bytes of slack at the end (the staging loads below round a byte range
// out to 16-byte boundaries).
//
// Kernels: a wavefront owns a tile of NW consecutive positions.  It copies the tile's blocks
// (one contiguous range of ncol and of val; node by node where only part of each node's blocks
// is read) into its own LDS region with coalesced 16-byte non-temporal loads, then G lanes handle
// one node: lane (s, v) takes the W columns v*W .. v*W+W-1 (one 16-byte vector of X per dof row)
// and every S-th block of the node.  Waves proceed independently (no block-wide barrier, except
// the final partials of the DOT form and the one between the two colours of a colour-pair
// launch); LDS operations of one wavefront execute in order, so a compiler barrier plus a wait
// for the wave's own LDS writes separates staging from use.
//
// Also here: the node transfers (R, P acting on the three dofs of a node alike).
#include <type_traits>

#include "mgx_device.h"

namespace mgx {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <typename T> struct NbVec;
template <> struct NbVec<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct NbVec<double> { typedef double type __attribute__((ext_vector_type(2))); };

constexpr int kNbMaxBlocks = 27;   // neighbour nodes of the 27-node hexahedral stencil

template <typename T, int K> struct NbShape {
    static constexpr int W = 16 / (int)sizeof(T);           // columns per lane
    static constexpr int V = K / W;                         // lanes covering the K columns of one row
    static constexpr int SMIN = 2;                          // at least 2 lanes split a node's blocks
    static constexpr int GMIN = 2 * (int)sizeof(T);         // and at least 8 (FP32) / 16 (FP64) lanes per node
    static constexpr int G = SMIN * V > GMIN ? SMIN * V : GMIN;   // lanes per node
    static constexpr int S = G / V;                         // lanes splitting a node's blocks
    static constexpr int NW = 64 / G;                       // nodes per wave tile
    static constexpr int MB = kNbMaxBlocks;                 // blocks per node staged
    // 16-byte LDS slots per wave for the values and the columns of a tile (+ alignment), rounded
    // up to whole 64-slot rounds of stage16
    static constexpr int VQ = ((NW * MB * 9 * (int)sizeof(T)) / 16 + 2 + 63) / 64 * 64;
    static constexpr int CQ = ((NW * MB * 4) / 16 + 2 + 63) / 64 * 64;
    static constexpr int U = 3;                             // blocks in flight per lane
    // per-node staging (kernels that read part of each node's blocks): a region of RQ value slots
    // and RC column slots per node, whole rounds of the node's G lanes
    static constexpr int RQ = ((MB * 9 * (int)sizeof(T)) / 16 + 2 + G - 1) / G * G;
    static constexpr int RC = ((MB * 4) / 16 + 2 + G - 1) / G * G;
    static_assert(NW * RQ <= VQ && NW * RC <= CQ, "per-node regions fit the wave's staging area");
};
static_assert(NbShape<float, 4>::G == 8 && NbShape<float, 16>::S == 2 && NbShape<float, 32>::G == 16 &&
                  NbShape<double, 32>::NW == 2,
              "shape");

constexpr int kNbWaves = kBlock / 64;

// wait for this wave's LDS writes and keep the compiler from moving LDS accesses across
__device__ __forceinline__ void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0), vmcnt and expcnt unconstrained
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// copy bytes [lo, hi) of src into dst[0 ..) shifted by lo & 15, with 16-byte loads that are all
// issued before the first LDS store.  Every lane loads and stores NI = NQ/64 slots without
// predication: slots past the range reload its last 16 bytes (same cache line) into dst slots
// < 64 * NI that the tile does not use.  The loads reach up to 15 bytes past hi.
template <int NQ>
__device__ __forceinline__ void stage16(const char *__restrict__ src, size_t lo, size_t hi, u32x4 *dst, int lane)
{
    constexpr int NI = NQ / 64;
    const size_t a0 = lo & ~(size_t)15;
    const int last = max((int)((hi - a0 + 15) >> 4), 1) - 1;
    const u32x4 *s = (const u32x4 *)(src + a0);
    u32x4 t[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) t[i] = __builtin_nontemporal_load(s + min(lane + 64 * i, last));
#pragma unroll
    for (int i = 0; i < NI; ++i) dst[lane + 64 * i] = t[i];
}

// stage16 for the G lanes of one node group (l = lane in the group): bytes [lo, hi) into
// dst[0 .. NQ), NQ a multiple of G; the same clamping (loads reach up to 15 bytes past hi)
template <int NQ, int G>
__device__ __forceinline__ void stage16_group(const char *__restrict__ src, size_t lo, size_t hi, u32x4 *dst, int l)
{
    constexpr int NI = NQ / G;
    const size_t a0 = lo & ~(size_t)15;
    const int last = max((int)((hi - a0 + 15) >> 4), 1) - 1;
    const u32x4 *s = (const u32x4 *)(src + a0);
    u32x4 t[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) t[i] = __builtin_nontemporal_load(s + min(l + G * i, last));
#pragma unroll
    for (int i = 0; i < NI; ++i) dst[l + G * i] = t[i];
}

// Blocks [m0, m1) of one node into its group's regions of the wave's staging area; returns the
// LDS addresses of block m0's values and column (alignment shift applied).  Call between two
// wave_lds_fence()s.
template <typename T, int K>
__device__ __forceinline__ void nb_stage_node(const T *__restrict__ val, const int32_t *__restrict__ ncol, int m0,
                                              int m1, u32x4 *sv, u32x4 *sc, int g, int l, const T *&lv,
                                              const int32_t *&lc)
{
    using Sh = NbShape<T, K>;
    const size_t vlo = (size_t)m0 * 9 * sizeof(T), clo = (size_t)m0 * 4;
    stage16_group<Sh::RQ, Sh::G>((const char *)val, vlo, (size_t)m1 * 9 * sizeof(T), sv + g * Sh::RQ, l);
    stage16_group<Sh::RC, Sh::G>((const char *)ncol, clo, (size_t)m1 * 4, sc + g * Sh::RC, l);
    lv = (const T *)((const char *)(sv + g * Sh::RQ) + (vlo & 15));
    lc = (const int32_t *)((const char *)(sc + g * Sh::RC) + (clo & 15));
}

// Sum over the blocks m = m0, m0 + S, ... < m1 of (block m) x (X rows of its neighbour node), for
// columns c0 .. c0+W-1: y[a] += sum_c A_m[a][c] X[3 j_m + c].  cols/vals: the wave's staged tile
// in LDS, block indices relative to the tile.  Groups of U blocks are loaded before use; the
// remainder goes one block at a time (no predicated loads).
template <typename T, int K>
__device__ __forceinline__ void nb_block(const T *a, const typename NbVec<T>::type (&x)[3], typename NbVec<T>::type (&y)[3])
{
    // one FMA per entry and column, accumulated straight into y
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) y[r] += a[3 * r + c] * x[c];
}

template <typename T, int K>
__device__ __forceinline__ void nb_accumulate(const int32_t *cols, const T *vals, int m0, int m1, const T *__restrict__ X,
                                              int c0, typename NbVec<T>::type (&y)[3])
{
    using Sh = NbShape<T, K>;
    using VT = typename NbVec<T>::type;
    int m = m0;
    for (; m + (Sh::U - 1) * Sh::S < m1; m += Sh::U * Sh::S) {
        VT x[Sh::U][3];
#pragma unroll
        for (int u = 0; u < Sh::U; ++u) {
            const T *xr = X + (size_t)3 * cols[m + u * Sh::S] * K + c0;
#pragma unroll
            for (int c = 0; c < 3; ++c) x[u][c] = *(const VT *)(xr + c * K);
        }
#pragma unroll
        for (int u = 0; u < Sh::U; ++u) nb_block<T, K>(vals + (size_t)(m + u * Sh::S) * 9, x[u], y);
    }
    for (; m < m1; m += Sh::S) {
        VT x[3];
        const T *xr = X + (size_t)3 * cols[m] * K + c0;
#pragma unroll
        for (int c = 0; c < 3; ++c) x[c] = *(const VT *)(xr + c * K);
        nb_block<T, K>(vals + (size_t)m * 9, x, y);
    }
}

// reduce the S partial sums of a node (lanes v, v + V, ... of its G-lane group)
template <typename T, int K>
__device__ __forceinline__ void nb_reduce(typename NbVec<T>::type (&y)[3])
{
    using Sh = NbShape<T, K>;
#pragma unroll
    for (int o = Sh::V; o < Sh::G; o <<= 1)
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int w = 0; w < Sh::W; ++w) y[r][w] += __shfl_xor(y[r][w], o, Sh::G);
}

// K column partials of a block: thread t holds columns (t % V) * W .. + W - 1
// (red: W * kBlock doubles of LDS, here the staging area, free once every wave is past its tiles)
template <typename T, int K>
__device__ __forceinline__ void nb_publish(const double *dot, double *red, double *__restrict__ partials, int stride)
{
    using Sh = NbShape<T, K>;
    const int t = threadIdx.x;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < Sh::W; ++w) red[w * kBlock + t] = dot[w];
    __syncthreads();
    if (t < K) {
        const int v = t / Sh::W, w = t % Sh::W;
        double a = 0;
        for (int i = v; i < kBlock; i += Sh::V) a += red[w * kBlock + i];
        partials[(size_t)t * stride + blockIdx.x] = a;
    }
}

// ------------------------------------------------------------------ SpMM  Y (op)= A X
template <typename T, int K, int EPI>
__global__ __launch_bounds__(kBlock) void k_nb_spmm(int32_t npos, const int32_t *__restrict__ nrp,
                                                    const int32_t *__restrict__ ncol, const int32_t *__restrict__ node,
                                                    const T *__restrict__ val, const T *__restrict__ X,
                                                    const T *__restrict__ B, const T *__restrict__ pre,
                                                    T *__restrict__ Y, const int32_t *__restrict__ done,
                                                    double *__restrict__ partials, int stride,
                                                    const int32_t *__restrict__ order)
{
    using Sh = NbShape<T, K>;
    using VT = typename NbVec<T>::type;
    constexpr int W = Sh::W;
    __shared__ u32x4 s_val[kNbWaves][Sh::VQ];
    __shared__ u32x4 s_col[kNbWaves][Sh::CQ];
    if (done && *done) return;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane / Sh::G, l = lane % Sh::G, s = l / Sh::V, c0 = (l % Sh::V) * W;
    const int ntiles = (npos + Sh::NW - 1) / Sh::NW;
    // XCD x (blocks b with b % 8 == x) walks the x-th eighth of the tiles
    const int xcd = blockIdx.x & 7;
    const int per = (gridDim.x >> 3) * kNbWaves;
    const int hi = (int)((int64_t)ntiles * (xcd + 1) / 8);
    double dot[W];
#pragma unroll
    for (int w = 0; w < W; ++w) dot[w] = 0;
    for (int pos = (int)((int64_t)ntiles * xcd / 8) + (int)(blockIdx.x >> 3) * kNbWaves + wv; pos < hi; pos += per) {
        const int tile = order ? order[pos] : pos;
        const int p0 = tile * Sh::NW, p1 = min(p0 + Sh::NW, npos);
        const int b0 = nrp[p0], b1 = nrp[p1];
        const int p = p0 + g;
        // the DOT form's own X rows, loaded with the tile's first loads so that they do not add a
        // memory round trip after the sums (its loop state already holds the form to 4 waves per
        // SIMD; in the one-tile forms the extra registers would halve the waves)
        const int qe = p < p1 ? (node ? node[p] : p) : 0;
        VT own[3] = {VT(0), VT(0), VT(0)};
        if (EPI == BEPI_DOT) {
#pragma unroll
            for (int a = 0; a < 3; ++a) own[a] = *(const VT *)(X + ((size_t)3 * qe + a) * K + c0);
        }
        VT y[3] = {VT(0), VT(0), VT(0)};
        wave_lds_fence();                                        // the previous tile's LDS reads are issued
        stage16<Sh::VQ>((const char *)val, (size_t)b0 * 9 * sizeof(T), (size_t)b1 * 9 * sizeof(T), s_val[wv], lane);
        stage16<Sh::CQ>((const char *)ncol, (size_t)b0 * 4, (size_t)b1 * 4, s_col[wv], lane);
        wave_lds_fence();
        const T *lv = (const T *)((const char *)s_val[wv] + (((size_t)b0 * 9 * sizeof(T)) & 15));
        const int32_t *lc = (const int32_t *)((const char *)s_col[wv] + (((size_t)b0 * 4) & 15));
        // UPRESID (own-and-later-colours copy): slot 0 is the node's own block, not part of the sum
        if (p < p1)
            nb_accumulate<T, K>(lc, lv, nrp[p] - b0 + (EPI == BEPI_UPRESID ? 1 : 0) + s, nrp[p + 1] - b0, X, c0, y);
        nb_reduce<T, K>(y);
        if (p < p1 && s == 0) {
            const int q = qe;
            if (EPI == BEPI_UPRESID) {
                // after a forward sweep from X = 0 (colours in order, each node's dofs 0, 1, 2 in
                // order) rows are left with the pull of what was relaxed after them: the later
                // colours and, inside the node, the strictly upper part of its own block (slot 0)
                const T *d = lv + (size_t)(nrp[p] - b0) * 9;
                const size_t i0 = (size_t)3 * q * K + c0;
                const VT x1 = *(const VT *)(X + i0 + K), x2 = *(const VT *)(X + i0 + 2 * K);
                y[0] += d[1] * x1;
                y[0] += d[2] * x2;
                y[1] += d[5] * x2;
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const size_t i = ((size_t)3 * q + a) * K + c0;
                VT *yp = (VT *)(Y + i);
                if (EPI == BEPI_SPMM) *yp = y[a];
                else if (EPI == BEPI_UPRESID) *yp = -y[a];
                else if (EPI == BEPI_ADD) *yp = *yp + y[a];
                else if (EPI == BEPI_RESID) *yp = *(const VT *)(B + i) - y[a];
                else if (EPI == BEPI_JACOBI) *yp = *(const VT *)(X + i) + pre[3 * q + a] * (*(const VT *)(B + i) - y[a]);
                else {
                    *yp = y[a];
                    const VT xp = own[a];
#pragma unroll
                    for (int w = 0; w < W; ++w) dot[w] += (double)xp[w] * (double)y[a][w];
                }
            }
        }
        // the other forms run the full grid (launch_nb_spmm): every wave has at most one tile, and
        // leaving here lets the compiler drop the loop-carried state (60 instead of 109 VGPRs)
        if (EPI != BEPI_DOT) break;
    }
    if (EPI == BEPI_DOT) nb_publish<T, K>(dot, (double *)&s_val[0][0], partials, stride);
}

template <typename T>
int nb_tile(int K)
{
    switch (K) {
    case 4: return NbShape<T, 4>::NW;
    case 8: return NbShape<T, 8>::NW;
    case 16: return NbShape<T, 16>::NW;
    default: return NbShape<T, 32>::NW;
    }
}
template int nb_tile<float>(int);
template int nb_tile<double>(int);

// waves of the bounded SpMM grid: at most 1024 blocks (4 waves each), a multiple of 8
template <typename T>
int nb_spmm_grid(int K, int32_t npos)
{
    const int64_t ntiles = (npos + nb_tile<T>(K) - 1) / nb_tile<T>(K);
    int64_t nb = (ntiles + kNbWaves - 1) / kNbWaves;
    nb = std::min<int64_t>(std::max<int64_t>(nb, 1), 1024);
    return (int)((nb + 7) & ~int64_t(7));
}
template int nb_spmm_grid<float>(int, int32_t);
template int nb_spmm_grid<double>(int, int32_t);

// one tile per wave, blocks dispatched in tile order
template <typename T>
int nb_spmm_full_grid(int K, int32_t npos)
{
    const int64_t ntiles = (npos + nb_tile<T>(K) - 1) / nb_tile<T>(K);
    return grid_rows(ntiles, kNbWaves);
}
template int nb_spmm_full_grid<float>(int, int32_t);
template int nb_spmm_full_grid<double>(int, int32_t);

// One tile per wave (blocks dispatched in tile order; the kernel relies on it and processes at most
// one tile per wave for these forms: 4m waves per XCD for ntiles <= 32m cover the XCD's
// ceil(ntiles / 8) tiles), except the DOT form: it runs the bounded
// grid of nb_spmm_grid (waves walk several tiles), whose block count is the number of partials
// the caller's finalize reads.  (One tile per wave was faster for the other forms, 0.75 against
// 0.91 ms for the 128^3 residual, and slower for DOT, whose per-block publication then
// dominates: 1.40 against 1.31 ms with a second reduction stage.)
template <typename T>
void launch_nb_spmm(int K, int epi, const NbView &A, const T *X, const T *B, const T *pre, T *Y, const int32_t *done,
                    double *partials, int stride, hipStream_t s, const int32_t *order)
{
    if (A.npos <= 0) return;
    const int g = epi == BEPI_DOT ? nb_spmm_grid<T>(K, A.npos) : nb_spmm_full_grid<T>(K, A.npos);
#define MGX_NBS(KK, EE)                                                                                          \
    hipLaunchKernelGGL((k_nb_spmm<T, KK, EE>), dim3(g), dim3(kBlock), 0, s, A.npos, A.nrp, A.ncol, A.node,     \
                       (const T *)A.val, X, B, pre, Y, done, partials, stride, order)
#define MGX_NBS_K(KK)                                       \
    switch (epi) {                                          \
    case BEPI_SPMM: MGX_NBS(KK, BEPI_SPMM); break;          \
    case BEPI_ADD: MGX_NBS(KK, BEPI_ADD); break;            \
    case BEPI_RESID: MGX_NBS(KK, BEPI_RESID); break;        \
    case BEPI_JACOBI: MGX_NBS(KK, BEPI_JACOBI); break;      \
    case BEPI_UPRESID: MGX_NBS(KK, BEPI_UPRESID); break;    \
    default: MGX_NBS(KK, BEPI_DOT); break;                  \
    }
    switch (K) {
    case 4: MGX_NBS_K(4); break;
    case 8: MGX_NBS_K(8); break;
    case 16: MGX_NBS_K(16); break;
    case 32: MGX_NBS_K(32); break;
    default: throw Error{MGX_ERR_INVALID_ARGUMENT, "block width must be 4, 8, 16 or 32"};
    }
#undef MGX_NBS_K
#undef MGX_NBS
    MGX_HIP(hipGetLastError());
}
template void launch_nb_spmm<float>(int, int, const NbView &, const float *, const float *, const float *, float *,
                                    const int32_t *, double *, int, hipStream_t, const int32_t *);
template void launch_nb_spmm<double>(int, int, const NbView &, const double *, const double *, const double *, double *,
                                     const int32_t *, double *, int, hipStream_t, const int32_t *);

// ------------------------------------------------------------------ multicolour Gauss-Seidel
// Positions [p0, p1) of the colour-ordered copy (one colour: no two of its nodes are coupled),
// tiles visited in ascending or (reverse) descending order: the nodes of one colour are
// independent, so the order changes no result, and alternating it between the colour launches
// of a sweep lets each launch start on the X rows the previous one touched last (still in the
// memory-side cache).
// Each node: a[r] = sum over its off-diagonal blocks (slots 1 ..) of A X, then the three dofs are
// relaxed in order with the node's own block D (slot 0): x_0 = (b_0 - a_0 - D01 x_1 - D02 x_2) /
// D00, then x_1 with the new x_0, then x_2 (reverse order when backward).
// One wave tile [pa, pb) of the colour-ordered copy (pb - pa <= NW): stage its blocks, sum the
// off-diagonal blocks, relax each node.  sv / sc: the wave's LDS staging region.
// the node's 3x3 solve: x_0 from row 0 with the current x_1, x_2, then x_1, then x_2 (reverse
// order when BW).  The direction is a template parameter: with a run-time flag the compiler merges
// the two orders into stores through a selected address, a stack copy of x whose reload waits
// for every memory load in flight.
template <typename T, bool BW>
__device__ __forceinline__ void nb_relax(const T *d, const typename NbVec<T>::type &r0,
                                         const typename NbVec<T>::type &r1, const typename NbVec<T>::type &r2,
                                         typename NbVec<T>::type &x0, typename NbVec<T>::type &x1,
                                         typename NbVec<T>::type &x2)
{
    if constexpr (!BW) {
        x0 = (r0 - d[1] * x1 - d[2] * x2) / d[0];
        x1 = (r1 - d[3] * x0 - d[5] * x2) / d[4];
        x2 = (r2 - d[6] * x0 - d[7] * x1) / d[8];
    } else {
        x2 = (r2 - d[6] * x0 - d[7] * x1) / d[8];
        x1 = (r1 - d[3] * x0 - d[5] * x2) / d[4];
        x0 = (r0 - d[1] * x1 - d[2] * x2) / d[0];
    }
}

template <typename T, int K, bool LOWER, bool BW>
__device__ __forceinline__ void nb_gs_tile(int pa, int pb, const int32_t *__restrict__ nrp,
                                           const int32_t *__restrict__ ncol, const int32_t *__restrict__ node,
                                           const int32_t *__restrict__ split, const T *__restrict__ val,
                                           const T *__restrict__ B, T *__restrict__ X, u32x4 *sv,
                                           u32x4 *sc, int lane)
{
    using Sh = NbShape<T, K>;
    using VT = typename NbVec<T>::type;
    const int g = lane / Sh::G, l = lane % Sh::G, s = l / Sh::V, c0 = (l % Sh::V) * Sh::W;
    const int b0 = nrp[pa], b1 = nrp[pb];
    const int p = pa + g;
    // the node's index is loaded with the tile's first loads, so the X and B rows of the relaxation
    // wait for one memory round trip after the sums, not two
    const int q = p < pb ? (node ? node[p] : p) : 0;
    VT y[3] = {VT(0), VT(0), VT(0)};
    const T *d;                                                  // the node's own block (slot 0) in LDS
    if (LOWER) {
        // the node's own block and the blocks of earlier colours only (the later ones multiply
        // zeros), staged node by node
        const int m0 = p < pb ? nrp[p] : b1, m1 = p < pb ? split[p] : b1;
        const T *lv;
        const int32_t *lc;
        nb_stage_node<T, K>(val, ncol, m0, m1, sv, sc, g, l, lv, lc);
        wave_lds_fence();
        if (p < pb) nb_accumulate<T, K>(lc, lv, 1 + s, m1 - m0, X, c0, y);
        d = lv;
    } else {
        stage16<Sh::VQ>((const char *)val, (size_t)b0 * 9 * sizeof(T), (size_t)b1 * 9 * sizeof(T), sv, lane);
        stage16<Sh::CQ>((const char *)ncol, (size_t)b0 * 4, (size_t)b1 * 4, sc, lane);
        wave_lds_fence();
        const T *lv = (const T *)((const char *)sv + (((size_t)b0 * 9 * sizeof(T)) & 15));
        const int32_t *lc = (const int32_t *)((const char *)sc + (((size_t)b0 * 4) & 15));
        if (p < pb) nb_accumulate<T, K>(lc, lv, nrp[p] - b0 + 1 + s, nrp[p + 1] - b0, X, c0, y);
        d = lv + (size_t)(p < pb ? nrp[p] - b0 : 0) * 9;
    }
    nb_reduce<T, K>(y);
    if (p >= pb || s != 0) return;
    const size_t i0 = (size_t)3 * q * K + c0;
    VT x0 = VT(0), x1 = VT(0), x2 = VT(0);
    if (!LOWER) x0 = *(const VT *)(X + i0), x1 = *(const VT *)(X + i0 + K), x2 = *(const VT *)(X + i0 + 2 * K);
    const VT r0 = *(const VT *)(B + i0) - y[0], r1 = *(const VT *)(B + i0 + K) - y[1],
             r2 = *(const VT *)(B + i0 + 2 * K) - y[2];
    nb_relax<T, BW>(d, r0, r1, r2, x0, x1, x2);
    *(VT *)(X + i0) = x0;
    *(VT *)(X + i0 + K) = x1;
    *(VT *)(X + i0 + 2 * K) = x2;
}

// The first colour launch of a sweep that starts from X = 0: every neighbour value is zero, so a
// node's relaxation reads only its own block and its B rows (the off-diagonal sums are 0 and the
// node's own X values 0, exactly what the general kernel would read).  Thread (p, v) handles the
// W columns v*W .. v*W+W-1 of position p.
template <typename T, int K, bool BW>
__global__ __launch_bounds__(kBlock) void k_nb_gs_zero(int32_t p0c, int32_t p1c, const int32_t *__restrict__ nrp,
                                                       const int32_t *__restrict__ node, const T *__restrict__ val,
                                                       const T *__restrict__ B, T *__restrict__ X,
                                                       const int32_t *__restrict__ done)
{
    using VT = typename NbVec<T>::type;
    constexpr int W = 16 / (int)sizeof(T);
    constexpr int V = K / W;
    if (done && *done) return;
    const int64_t t = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const int64_t pp = t / V;
    if (pp >= p1c - p0c) return;
    const int p = p0c + (int)pp, c0 = (int)(t - pp * V) * W;
    const int q = node ? node[p] : p;
    T d[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) d[e] = val[(size_t)nrp[p] * 9 + e];
    const size_t i0 = (size_t)3 * q * K + c0;
    VT x0 = VT(0), x1 = VT(0), x2 = VT(0);
    const VT r0 = *(const VT *)(B + i0), r1 = *(const VT *)(B + i0 + K), r2 = *(const VT *)(B + i0 + 2 * K);
    nb_relax<T, BW>(d, r0, r1, r2, x0, x1, x2);
    *(VT *)(X + i0) = x0;
    *(VT *)(X + i0 + K) = x1;
    *(VT *)(X + i0 + 2 * K) = x2;
}

template <typename T, int K, bool LOWER, bool BW>
__global__ __launch_bounds__(kBlock) void k_nb_gs(int32_t p0c, int32_t p1c, const int32_t *__restrict__ nrp,
                                                  const int32_t *__restrict__ ncol, const int32_t *__restrict__ node,
                                                  const int32_t *__restrict__ split, const T *__restrict__ val,
                                                  const T *__restrict__ B, T *__restrict__ X, int reverse,
                                                  const int32_t *__restrict__ done)
{
    using Sh = NbShape<T, K>;
    __shared__ u32x4 s_val[kNbWaves][Sh::VQ];
    __shared__ u32x4 s_col[kNbWaves][Sh::CQ];
    if (done && *done) return;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ntiles = (p1c - p0c + Sh::NW - 1) / Sh::NW;
    const int t0 = xcd_remap(blockIdx.x, gridDim.x) * kNbWaves + wv;
    if (t0 >= ntiles) return;                                    // the whole wave leaves together
    const int tile = reverse ? ntiles - 1 - t0 : t0;
    const int pa = p0c + tile * Sh::NW;
    nb_gs_tile<T, K, LOWER, BW>(pa, min(pa + Sh::NW, p1c), nrp, ncol, node, split, val, B, X, s_val[wv],
                                s_col[wv], lane);
}

// Two colours of a sweep in one launch.  Block b takes mesh line b: rng[4b] .. rng[4b+1]-1 are
// its positions of colour a, rng[4b+2] .. rng[4b+3]-1 those of colour b; the block relaxes the
// first colour's positions (a, or b when second_first), waits at a block barrier, then relaxes
// the other colour's.  This is the sweep's order whenever every edge between nodes of the two
// colours joins two nodes of one line (the host checks the matrix for it): a node of the first
// colour reads only old values of the second colour (none is written before the barrier in its
// line, and no other line's is a neighbour), and a node of the second colour reads the new values
// of the first colour of its own line, written by this block before the barrier.  The second half
// reads the same X neighbourhood as the first, from cache.
template <typename T, int K, bool LOWER, bool BW, int NB>
__global__ __launch_bounds__(NB) void k_nb_gs_pair(const int32_t *__restrict__ rng, int32_t nlines,
                                                           int second_first, const int32_t *__restrict__ nrp,
                                                           const int32_t *__restrict__ ncol,
                                                           const int32_t *__restrict__ node,
                                                           const int32_t *__restrict__ split, const T *__restrict__ val,
                                                           const T *__restrict__ B, T *__restrict__ X, int reverse,
                                                           const int32_t *__restrict__ done)
{
    using Sh = NbShape<T, K>;
    constexpr int NWB = NB / 64;
    __shared__ u32x4 s_val[NWB][Sh::VQ];
    __shared__ u32x4 s_col[NWB][Sh::CQ];
    if (done && *done) return;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    if (b >= nlines) return;                                     // the whole block leaves together
    const int line = reverse ? nlines - 1 - b : b;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;   // wave-uniform
    // wave w takes tiles w, w + NWB, ... of each colour; the second colour's in reverse, so a wave
    // starts it on the stretch of the line its last first-colour tile has just gathered
    for (int h = 0; h < 2; ++h) {
        const int o = 4 * line + 2 * (h ^ second_first);
        const int p0 = rng[o], p1 = rng[o + 1];
        const int nt = (p1 - p0 + Sh::NW - 1) / Sh::NW;
        if (wv < nt)
            for (int i = 0, last = wv + (nt - 1 - wv) / NWB * NWB; wv + i * NWB <= last; ++i) {
                const int pa = p0 + (h == 0 ? wv + i * NWB : last - i * NWB) * Sh::NW;
                nb_gs_tile<T, K, LOWER, BW>(pa, min(pa + Sh::NW, p1), nrp, ncol, node, split, val, B, X,
                                            s_val[wv], s_col[wv], lane);
            }
        if (h == 0) __syncthreads();
    }
}

template <typename T>
void launch_nb_gs(int K, const NbView &A, int32_t p0, int32_t p1, const T *B, T *X, bool backward, bool reverse,
                  const int32_t *done, hipStream_t s, int start)
{
    if (p1 <= p0) return;
    if (start == NB_GS_LOWER && (!A.split || backward))
        throw Error{MGX_ERR_INVALID_OPERATION, "lower-colour sweep needs a colour-sorted copy and a forward sweep"};
    if (start == NB_GS_ZERO) {
        const int g = grid_rows((int64_t)(p1 - p0) * (K / (16 / (int)sizeof(T))), kBlock);
#define MGX_NBZ(KK)                                                                                              \
    if (backward)                                                                                                \
        hipLaunchKernelGGL((k_nb_gs_zero<T, KK, true>), dim3(g), dim3(kBlock), 0, s, p0, p1, A.nrp, A.node,    \
                           (const T *)A.val, B, X, done);                                                        \
    else                                                                                                         \
        hipLaunchKernelGGL((k_nb_gs_zero<T, KK, false>), dim3(g), dim3(kBlock), 0, s, p0, p1, A.nrp, A.node,   \
                           (const T *)A.val, B, X, done)
        switch (K) {
        case 4: MGX_NBZ(4); break;
        case 8: MGX_NBZ(8); break;
        case 16: MGX_NBZ(16); break;
        case 32: MGX_NBZ(32); break;
        default: throw Error{MGX_ERR_INVALID_ARGUMENT, "block width must be 4, 8, 16 or 32"};
        }
#undef MGX_NBZ
        MGX_HIP(hipGetLastError());
        return;
    }
    const int nw = nb_tile<T>(K);
    const int g = grid_rows((p1 - p0 + nw - 1) / nw, kNbWaves);
#define MGX_NBG1(KK, LO, BWD)                                                                                    \
    hipLaunchKernelGGL((k_nb_gs<T, KK, LO, BWD>), dim3(g), dim3(kBlock), 0, s, p0, p1, A.nrp, A.ncol, A.node,   \
                       A.split, (const T *)A.val, B, X, reverse ? 1 : 0, done)
#define MGX_NBG(KK)                                                                                              \
    if (start == NB_GS_LOWER) MGX_NBG1(KK, true, false);                                                         \
    else if (backward) MGX_NBG1(KK, false, true);                                                                \
    else MGX_NBG1(KK, false, false)
    switch (K) {
    case 4: MGX_NBG(4); break;
    case 8: MGX_NBG(8); break;
    case 16: MGX_NBG(16); break;
    case 32: MGX_NBG(32); break;
    default: throw Error{MGX_ERR_INVALID_ARGUMENT, "block width must be 4, 8, 16 or 32"};
    }
#undef MGX_NBG
#undef MGX_NBG1
    MGX_HIP(hipGetLastError());
}
template <typename T>
void launch_nb_gs_pair(int K, const NbView &A, const int32_t *rng, int32_t nlines, bool second_first, const T *B,
                       T *X, bool backward, bool reverse, bool lower, const int32_t *done, hipStream_t s,
                       int32_t line_nodes)
{
    if (nlines <= 0) return;
    if (lower && (!A.split || backward))
        throw Error{MGX_ERR_INVALID_OPERATION, "lower-colour sweep needs a colour-sorted copy and a forward sweep"};
    const int g = grid_rows(nlines, 1);
    // The FP32 general form at 32 columns on lines of 128 nodes or more (each of 16 waves then has
    // a tile of each colour) runs 1024-thread blocks: 0.467 against 0.492 ms per launch at 128^3;
    // with them the zero-start (lower) form made the batch slower.
    const bool wide = sizeof(T) == 4 && K == 32 && !lower && line_nodes / 2 >= 16 * nb_tile<T>(K);
#define MGX_NBP1(KK, LO, BWD, NB)                                                                                \
    hipLaunchKernelGGL((k_nb_gs_pair<T, KK, LO, BWD, NB>), dim3(g), dim3(NB), 0, s, rng, nlines,                 \
                       second_first ? 1 : 0, A.nrp, A.ncol, A.node, A.split, (const T *)A.val, B, X, reverse ? 1 : 0, \
                       done)
#define MGX_NBP(KK, NB)                                                                                          \
    if (lower) MGX_NBP1(KK, true, false, NB);                                                                    \
    else if (backward) MGX_NBP1(KK, false, true, NB);                                                            \
    else MGX_NBP1(KK, false, false, NB)
    switch (K) {
    case 4: MGX_NBP(4, 512); break;
    case 8: MGX_NBP(8, 512); break;
    case 16: MGX_NBP(16, 512); break;
    case 32:
        if (wide) MGX_NBP(32, 1024);
        else MGX_NBP(32, 512);
        break;
    default: throw Error{MGX_ERR_INVALID_ARGUMENT, "block width must be 4, 8, 16 or 32"};
    }
#undef MGX_NBP
#undef MGX_NBP1
    MGX_HIP(hipGetLastError());
}

template void launch_nb_gs<float>(int, const NbView &, int32_t, int32_t, const float *, float *, bool, bool,
                                  const int32_t *, hipStream_t, int);
template void launch_nb_gs_pair<float>(int, const NbView &, const int32_t *, int32_t, bool, const float *, float *,
                                       bool, bool, bool, const int32_t *, hipStream_t, int32_t);
template void launch_nb_gs_pair<double>(int, const NbView &, const int32_t *, int32_t, bool, const double *,
                                        double *, bool, bool, bool, const int32_t *, hipStream_t, int32_t);
template void launch_nb_gs<double>(int, const NbView &, int32_t, int32_t, const double *, double *, bool, bool,
                                   const int32_t *, hipStream_t, int);

// ------------------------------------------------------------------ node transfers
// A transfer matrix (prolongation P or restriction R) between two levels whose dofs come in node
// triples, when it acts on the three dofs of a node alike: entry (3i+a, 3j+b) = w_ij if a == b,
// else 0.  Stored per node: rows i of the output level, columns j of the input level, one weight
// each (a third of the CSR's entries, and no dof column indices).  Y[3i+a][:] (+)= sum_j w_ij
// X[3j+a][:].  Thread (i, v) owns the 16-byte column vector v (columns v*W .. v*W+W-1) of the three
// rows of output node i and reads the same vector of the three rows of each input node, 4 entries
// in flight: V = K / W threads per node, so a wave covers 64 / V consecutive output nodes and each
// row segment it reads or writes is V * 16 contiguous bytes.
template <typename T, int K, bool ADD>
__global__ __launch_bounds__(kBlock) void k_nt_apply(int32_t nout, const int32_t *__restrict__ rp,
                                                     const int32_t *__restrict__ ci, const T *__restrict__ w,
                                                     const T *__restrict__ X, T *__restrict__ Y,
                                                     const int32_t *__restrict__ done)
{
    using VT = typename NbVec<T>::type;
    constexpr int W = 16 / (int)sizeof(T);
    constexpr int V = K / W;                                     // threads per node
    if (done && *done) return;
    const int64_t t = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const int64_t i = t / V;
    if (i >= nout) return;
    const int c0 = (int)(t - i * V) * W;
    VT *yp = (VT *)(Y + (size_t)3 * i * K + c0);
    VT yv[3] = {VT(0), VT(0), VT(0)};
    if (ADD) {                                                   // issued with the first loads
#pragma unroll
        for (int a = 0; a < 3; ++a) yv[a] = yp[a * V];
    }
    const int b1 = rp[i + 1];
    int b = rp[i];
    VT acc[3] = {VT(0), VT(0), VT(0)};
    for (; b + 3 < b1; b += 4) {
        int j[4];
        T wv[4];
        VT x[4][3];
#pragma unroll
        for (int u = 0; u < 4; ++u) { j[u] = ci[b + u]; wv[u] = w[b + u]; }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int a = 0; a < 3; ++a) x[u][a] = *(const VT *)(X + ((size_t)3 * j[u] + a) * K + c0);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int a = 0; a < 3; ++a) acc[a] += wv[u] * x[u][a];
    }
    for (; b < b1; ++b) {
        const int j = ci[b];
        const T wv = w[b];
#pragma unroll
        for (int a = 0; a < 3; ++a) acc[a] += wv * *(const VT *)(X + ((size_t)3 * j + a) * K + c0);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) yp[a * V] = ADD ? yv[a] + acc[a] : acc[a];
}

template <typename T>
void launch_nt_apply(int K, bool add, const NtView &M, const T *X, T *Y, const int32_t *done, hipStream_t s)
{
    if (M.nout <= 0) return;
    const int W = 16 / (int)sizeof(T);
    const int g = grid_rows((int64_t)M.nout * (K / W), kBlock);
#define MGX_NTA(KK, AA)                                                                                          \
    hipLaunchKernelGGL((k_nt_apply<T, KK, AA>), dim3(g), dim3(kBlock), 0, s, M.nout, M.rp, M.ci, (const T *)M.w, \
                       X, Y, done)
#define MGX_NTA_K(KK)                    \
    if (add) MGX_NTA(KK, true);          \
    else MGX_NTA(KK, false)
    switch (K) {
    case 4: MGX_NTA_K(4); break;
    case 8: MGX_NTA_K(8); break;
    case 16: MGX_NTA_K(16); break;
    case 32: MGX_NTA_K(32); break;
    default: throw Error{MGX_ERR_INVALID_ARGUMENT, "block width must be 4, 8, 16 or 32"};
    }
#undef MGX_NTA_K
#undef MGX_NTA
    MGX_HIP(hipGetLastError());
}
template void launch_nt_apply<float>(int, bool, const NtView &, const float *, float *, const int32_t *, hipStream_t);
template void launch_nt_apply<double>(int, bool, const NtView &, const double *, double *, const int32_t *,
                                      hipStream_t);

// Entries per output node, and the check that the CSR is a node transfer: the three rows of node
// i have one length, row 3i+a holds only columns 3j+a, and the three rows list the same input
// nodes j in the same order with bit-identical values.  *bad = 1 on any violation.
__global__ __launch_bounds__(kBlock) void k_nt_count(int32_t nout, const int32_t *__restrict__ rp,
                                                     const int32_t *__restrict__ ci, const double *__restrict__ v,
                                                     int32_t *__restrict__ cnt, int32_t *__restrict__ bad)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nout) return;
    const int r0 = rp[3 * i], len = rp[3 * i + 1] - r0;
    bool ok = rp[3 * i + 2] - rp[3 * i + 1] == len && rp[3 * i + 3] - rp[3 * i + 2] == len;
    for (int a = 0; ok && a < 3; ++a) {
        const int ra = r0 + a * len;
        for (int m = 0; ok && m < len; ++m) {
            const int c = ci[ra + m];
            ok = c % 3 == a && c / 3 == ci[r0 + m] / 3 && v[ra + m] == v[r0 + m];
        }
    }
    if (!ok) atomicOr(bad, 1);
    cnt[i] = ok ? len : 0;
}

template <typename TV>
__global__ __launch_bounds__(kBlock) void k_nt_fill(int32_t nout, const int32_t *__restrict__ rp,
                                                    const int32_t *__restrict__ ci, const double *__restrict__ v,
                                                    const int32_t *__restrict__ nrp, int32_t *__restrict__ ncol,
                                                    TV *__restrict__ w)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nout) return;
    const int r0 = rp[3 * i], o = nrp[i], len = nrp[i + 1] - o;
    for (int m = 0; m < len; ++m) {
        ncol[o + m] = ci[r0 + m] / 3;
        w[o + m] = (TV)v[r0 + m];
    }
}

void nt_count(int32_t nout, const int32_t *rp, const int32_t *ci, const double *v, int32_t *cnt, int32_t *bad,
              hipStream_t s)
{
    if (nout <= 0) return;
    hipLaunchKernelGGL(k_nt_count, dim3((nout + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nout, rp, ci, v, cnt, bad);
    MGX_HIP(hipGetLastError());
}

template <typename TV>
void nt_fill(int32_t nout, const int32_t *rp, const int32_t *ci, const double *v, const int32_t *nrp, int32_t *ncol,
             TV *w, hipStream_t s)
{
    if (nout <= 0) return;
    hipLaunchKernelGGL((k_nt_fill<TV>), dim3((nout + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nout, rp, ci, v, nrp,
                       ncol, w);
    MGX_HIP(hipGetLastError());
}
template void nt_fill<float>(int32_t, const int32_t *, const int32_t *, const double *, const int32_t *, int32_t *,
                             float *, hipStream_t);
template void nt_fill<double>(int32_t, const int32_t *, const int32_t *, const double *, const int32_t *, int32_t *,
                              double *, hipStream_t);

// ------------------------------------------------------------------ building a copy from CSR
// Blocks per position, and the checks that make the CSR a node-block matrix: the three rows of
// the node have one column list, its length a multiple of 3, and every column triple is a whole
// node (3j, 3j+1, 3j+2).  Bit 0 of *bad
// on any violation (with diag_first, also when a node has no block of its own); bit 1 when a node
// has more blocks than the kernels stage per node (kNbMaxBlocks).
// upper (with diag_first and colour): count only the node's own block and the blocks of
// neighbours coloured after it.
__global__ __launch_bounds__(kBlock) void k_nb_count(int32_t npos, const int32_t *__restrict__ node,
                                                     const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                     int diag_first, int32_t *__restrict__ cnt, int32_t *__restrict__ bad,
                                                     const int32_t *__restrict__ colour, int upper)
{
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npos) return;
    const int q = node ? node[p] : p;
    const int s0 = rp[3 * q], len = rp[3 * q + 1] - s0;
    bool ok = len % 3 == 0 && rp[3 * q + 2] - rp[3 * q + 1] == len && rp[3 * q + 3] - rp[3 * q + 2] == len;
    bool own = false;
    for (int m = 0; ok && m < len / 3; ++m) {
        const int c = ci[s0 + 3 * m];
        ok = c % 3 == 0 && ci[s0 + 3 * m + 1] == c + 1 && ci[s0 + 3 * m + 2] == c + 2;
        own = own || c == 3 * q;
    }
    for (int k = 0; ok && k < len; ++k) ok = ci[s0 + len + k] == ci[s0 + k] && ci[s0 + 2 * len + k] == ci[s0 + k];
    if (!ok || (diag_first && !own)) atomicOr(bad, 1);
    if (len / 3 > kNbMaxBlocks) atomicOr(bad, 2);
    int n = ok ? len / 3 : 0;
    if (upper && ok) {
        n = 1;
        for (int m = 0; m < len / 3; ++m) n += colour[ci[s0 + 3 * m] / 3] > colour[q] ? 1 : 0;
    }
    cnt[p] = n;
}

// Slots of a position's blocks: with diag_first its own block first; with colour (diag_first
// only) the other blocks ordered by the colour of their neighbour node, ties in column order, and
// split[p] = the first slot of a neighbour coloured after the node itself (neighbours never share
// the node's colour).  Without colour the column order is kept.
template <typename TI, typename TV>
__global__ __launch_bounds__(kBlock) void k_nb_fill(int32_t npos, const int32_t *__restrict__ node,
                                                    const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                    const TI *__restrict__ v, const int32_t *__restrict__ nrp,
                                                    int diag_first, int32_t *__restrict__ ncol, TV *__restrict__ val,
                                                    const int32_t *__restrict__ colour, int32_t *__restrict__ split,
                                                    int upper)
{
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npos) return;
    const int q = node ? node[p] : p;
    const int nb = nrp[p + 1] - nrp[p];
    if (nb == 0) {
        if (split) split[p] = nrp[p];
        return;
    }
    const int r0 = rp[3 * q];
    int own = -1;
    if (diag_first)
        for (int m = 0; m < (upper ? (rp[3 * q + 1] - r0) / 3 : nb); ++m) if (ci[r0 + 3 * m] == 3 * q) own = m;
    if (upper) {
        // the own block, then the later-coloured neighbours in column order
        const int len = (rp[3 * q + 1] - r0) / 3, cq = colour[q];
        int slot = nrp[p];
        for (int m = 0; m < len; ++m) {
            if (m != own && colour[ci[r0 + 3 * m] / 3] <= cq) continue;
            const int sl = m == own ? nrp[p] : ++slot;
            ncol[sl] = ci[r0 + 3 * m] / 3;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int c = 0; c < 3; ++c) val[(size_t)sl * 9 + 3 * a + c] = (TV)v[rp[3 * q + a] + 3 * m + c];
        }
        return;
    }
    const bool bycol = colour && own >= 0;
    if (bycol) {
        const int cq = colour[q];
        int before = 0;
        for (int m = 0; m < nb; ++m)
            if (m != own && colour[ci[r0 + 3 * m] / 3] < cq) ++before;
        split[p] = nrp[p] + 1 + before;
    } else if (split) {
        split[p] = nrp[p + 1];
    }
    for (int m = 0; m < nb; ++m) {
        int slot;
        if (bycol) {
            // rank among the other blocks by (neighbour colour, position in the column list)
            const int cm = colour[ci[r0 + 3 * m] / 3];
            int rank = 0;
            for (int o = 0; o < nb; ++o) {
                if (o == own || o == m) continue;
                const int co = colour[ci[r0 + 3 * o] / 3];
                rank += (co < cm || (co == cm && o < m)) ? 1 : 0;
            }
            slot = nrp[p] + (m == own ? 0 : 1 + rank);
        } else {
            slot = nrp[p] + (own < 0 ? m : (m == own ? 0 : (m < own ? m + 1 : m)));
        }
        ncol[slot] = ci[r0 + 3 * m] / 3;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) val[(size_t)slot * 9 + 3 * a + c] = (TV)v[rp[3 * q + a] + 3 * m + c];
    }
}

void nb_count(int32_t npos, const int32_t *node, const int32_t *rp, const int32_t *ci, bool diag_first, int32_t *cnt,
              int32_t *bad, hipStream_t s, const int32_t *colour, bool upper)
{
    if (npos <= 0) return;
    hipLaunchKernelGGL(k_nb_count, dim3((npos + kBlock - 1) / kBlock), dim3(kBlock), 0, s, npos, node, rp, ci,
                       diag_first ? 1 : 0, cnt, bad, colour, upper ? 1 : 0);
    MGX_HIP(hipGetLastError());
}

template <typename TI, typename TV>
void nb_fill(int32_t npos, const int32_t *node, const int32_t *rp, const int32_t *ci, const TI *v, const int32_t *nrp,
             bool diag_first, int32_t *ncol, TV *val, hipStream_t s, const int32_t *colour, int32_t *split,
             bool upper)
{
    if (npos <= 0) return;
    hipLaunchKernelGGL((k_nb_fill<TI, TV>), dim3((npos + kBlock - 1) / kBlock), dim3(kBlock), 0, s, npos, node, rp, ci,
                       v, nrp, diag_first ? 1 : 0, ncol, val, colour, split, upper ? 1 : 0);
    MGX_HIP(hipGetLastError());
}
template void nb_fill<double, double>(int32_t, const int32_t *, const int32_t *, const int32_t *, const double *,
                                      const int32_t *, bool, int32_t *, double *, hipStream_t, const int32_t *,
                                      int32_t *, bool);
template void nb_fill<double, float>(int32_t, const int32_t *, const int32_t *, const int32_t *, const double *,
                                     const int32_t *, bool, int32_t *, float *, hipStream_t, const int32_t *,
                                     int32_t *, bool);
template void nb_fill<float, float>(int32_t, const int32_t *, const int32_t *, const int32_t *, const float *,
                                    const int32_t *, bool, int32_t *, float *, hipStream_t, const int32_t *,
                                    int32_t *, bool);

}  // namespace mgx
