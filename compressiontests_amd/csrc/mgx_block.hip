// mgx_block.hip -- kernels of the batching form: K right-hand sides (K in {4, 8, 16, 32})
// advance together through one PCG, so every pass over a matrix serves K columns.
//
// Block vectors are n x K, row-major (element (i, c) at i*K + c): the K values a matrix entry
// multiplies are one contiguous 8K- (FP64) or 4K-byte (FP32) segment, read by K adjacent lanes.
//
//  * k_spmm: a 256-thread block takes a tile of consecutive rows, stages the tile's contiguous
//    (col, val) range into LDS with coalesced non-temporal loads, and then G = K*S lanes work
//    on one row: lane (c, s) handles column c and every S-th entry of the row, reading the
//    entry from LDS (broadcast) and gathering X[col*K + c].  S = 16/K for K < 16 so that a row
//    is still split over 16 lanes.  The grid is bounded (<= 2048 blocks): each XCD walks a
//    contiguous eighth of the tiles, so its L2 sees one window of X.  Epilogues: Y = AX,
//    Y += AX, Y = B - AX, Jacobi Y = X + pre (B - AX), and Y = AX with per-column X.Y partials.
//  * k_gs_block: multicolour Gauss-Seidel over 3-row node blocks for K columns at once; the
//    node's 3x3 diagonal block is common to all columns.
//  * Vector kernels (norms, fused x/r update, r.z, p update) give every thread a fixed column
//    (block size and grid stride are multiples of K); per-column partial sums go to
//    partials[(slot*K + c)*stride + block] and k_bfinalize (one block) combines them in a
//    fixed order and updates the per-column PCG scalars of a BCtl.  Converged columns are
//    frozen (alpha = beta = 0) while the others continue; every kernel returns at once when
//    all columns are done.
#include <type_traits>

#include "mgx_device.h"

namespace mgx {

// ------------------------------------------------------------------ SpMM
template <typename TV, typename TX, int K, int EPI>
__global__ __launch_bounds__(kBlock) void k_spmm(int32_t nrows, const int32_t *__restrict__ rp,
                                                 const int32_t *__restrict__ ci, const TV *__restrict__ v,
                                                 const TX *__restrict__ X, const TX *__restrict__ B,
                                                 const TV *__restrict__ pre, TX *__restrict__ Y,
                                                 const int32_t *__restrict__ done, double *__restrict__ partials,
                                                 int stride)
{
    using Sh = BShape<K>;
    __shared__ int32_t s_ci[Sh::CAP];
    __shared__ TV s_v[Sh::CAP];
    if (done && *done) return;
    const int t = threadIdx.x, g = t / Sh::G, l = t % Sh::G, c = l % K, sl = l / K;
    const int ntiles = (nrows + Sh::RPB - 1) / Sh::RPB;
    const int per = gridDim.x >> 3, xcd = blockIdx.x & 7;
    const int hi = (int)((int64_t)ntiles * (xcd + 1) / 8);
    double dot = 0;
    for (int tile = (int)((int64_t)ntiles * xcd / 8) + (int)(blockIdx.x >> 3); tile < hi; tile += per) {
        const int r0 = tile * Sh::RPB, r1 = min(r0 + Sh::RPB, nrows);
        const int e0 = rp[r0], e1 = rp[r1];
        const bool staged = e1 - e0 <= Sh::CAP;     // the same for the whole block
        if (staged) {
            __syncthreads();                        // the previous tile's readers are done
            for (int k = t; k < e1 - e0; k += kBlock) { s_ci[k] = ntload(ci + e0 + k); s_v[k] = ntload(v + e0 + k); }
            __syncthreads();
        }
        const int row = r0 + g;
        if (row < r1) {
            const int s = rp[row], e = rp[row + 1];
            TV acc = staged ? row_cols<K, TV, TX>(s_ci, s_v, s - e0, e - e0, X, c, sl)
                            : row_cols<K, TV, TX>(ci, v, s, e, X, c, sl);
#pragma unroll
            for (int o = K; o < Sh::G; o <<= 1) acc += __shfl_xor(acc, o, Sh::G);
            if (sl == 0) {
                const size_t i = (size_t)row * K + c;
                if (EPI == BEPI_SPMM) Y[i] = (TX)acc;
                else if (EPI == BEPI_ADD) Y[i] = (TX)((TV)Y[i] + acc);
                else if (EPI == BEPI_RESID) Y[i] = (TX)((TV)B[i] - acc);
                else if (EPI == BEPI_JACOBI) Y[i] = (TX)((TV)X[i] + pre[row] * ((TV)B[i] - acc));
                else {
                    const TX q = (TX)acc;
                    Y[i] = q;
                    dot += (double)X[i] * (double)q;
                }
            }
        }
    }
    if (EPI == BEPI_DOT) publish_cols<K, 1, -1>(dot, 0, 0, partials, stride);
}

template <typename TV, typename TX>
void launch_spmm(int K, int epi, int32_t nrows, const int32_t *rp, const int32_t *ci, const TV *v, const TX *X,
                 const TX *B, const TV *pre, TX *Y, const int32_t *done, double *partials, int stride, hipStream_t s)
{
    if (nrows <= 0) return;
#define MGX_SPMM(KK, EE)                                                                                       \
    hipLaunchKernelGGL((k_spmm<TV, TX, KK, EE>), dim3(grid_tiles((nrows + BShape<KK>::RPB - 1) / BShape<KK>::RPB)), \
                       dim3(kBlock), 0, s, nrows, rp, ci, v, X, B, pre, Y, done, partials, stride)
#define MGX_SPMM_K(KK)                                        \
    switch (epi) {                                            \
    case BEPI_SPMM: MGX_SPMM(KK, BEPI_SPMM); break;           \
    case BEPI_ADD: MGX_SPMM(KK, BEPI_ADD); break;             \
    case BEPI_RESID: MGX_SPMM(KK, BEPI_RESID); break;         \
    case BEPI_JACOBI: MGX_SPMM(KK, BEPI_JACOBI); break;       \
    default: MGX_SPMM(KK, BEPI_DOT); break;                   \
    }
    switch (K) {
    case 4: MGX_SPMM_K(4); break;
    case 8: MGX_SPMM_K(8); break;
    case 16: MGX_SPMM_K(16); break;
    case 32: MGX_SPMM_K(32); break;
    default: throw Error{MGX_ERR_INVALID_ARGUMENT, "block width must be 4, 8, 16 or 32"};
    }
#undef MGX_SPMM_K
#undef MGX_SPMM
    MGX_HIP(hipGetLastError());
}

template void launch_spmm<double, double>(int, int, int32_t, const int32_t *, const int32_t *, const double *,
                                          const double *, const double *, const double *, double *, const int32_t *,
                                          double *, int, hipStream_t);
template void launch_spmm<float, float>(int, int, int32_t, const int32_t *, const int32_t *, const float *,
                                        const float *, const float *, const float *, float *, const int32_t *,
                                        double *, int, hipStream_t);
template void launch_spmm<double, float>(int, int, int32_t, const int32_t *, const int32_t *, const double *,
                                         const float *, const float *, const double *, float *, const int32_t *,
                                         double *, int, hipStream_t);


// ------------------------------------------------------------------ node-triple kernels
// For matrices whose rows come in node triples (rows 3q, 3q+1, 3q+2 with identical column lists,
// as every level matrix of the 3-dof hexahedral hierarchy), each gathered X row serves all three
// rows.  A lane holds W = 16 bytes of columns (4 FP32 or 2 FP64), so the K values of one X row
// come in with V = K / W lanes of one 16-byte load each; 32 lanes per node split the node's
// entries S = 32 / V ways.  (With one column per lane the kernels were bound by the latency of
// many 4-byte gathers, not by bandwidth.)
template <typename T> struct Vec16;
template <> struct Vec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr int W = 4;
};
template <> struct Vec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
    static constexpr int W = 2;
};

template <typename T, int K> struct NShape {
    static constexpr int W = Vec16<T>::W;     // columns per lane
    static constexpr int V = K / W;           // lanes covering one X row
    static constexpr int G = 32;              // lanes per node
    static constexpr int S = G / V;           // lanes splitting a node's entries
    static constexpr int NPB = kBlock / G;    // nodes per tile
    static constexpr int CAP = NPB * 3 * 84;  // staged entries per tile
    static constexpr int U = 4;               // entries in flight per lane
};

int spmm_node_tile(int) { return kBlock / 32; }

// K column partials of a block: thread t holds columns c0 .. c0 + W - 1 (c0 = (t % V) W)
template <typename T, int K>
__device__ __forceinline__ void publish_vcols(const double *dot, double *__restrict__ partials, int stride)
{
    using Sh = NShape<T, K>;
    __shared__ double red[Sh::W][kBlock];
    const int t = threadIdx.x;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < Sh::W; ++w) red[w][t] = dot[w];
    __syncthreads();
    if (t < K) {
        const int v = t / Sh::W, w = t % Sh::W;
        double a = 0;
        for (int j = v; j < kBlock; j += Sh::V) a += red[w][j];
        partials[(size_t)t * stride + blockIdx.x] = a;
    }
}

template <typename T, int K, typename PI, typename PV>
__device__ __forceinline__ void node_cols(PI cis, PV vs, int o0, int o1, int o2, int len, const T *__restrict__ X,
                                          int c0, int sl, typename Vec16<T>::type &y0, typename Vec16<T>::type &y1,
                                          typename Vec16<T>::type &y2)
{
    using Sh = NShape<T, K>;
    using VT = typename Vec16<T>::type;
    for (int j = sl; j < len; j += Sh::U * Sh::S) {
        int cc[Sh::U];
        T w0[Sh::U], w1[Sh::U], w2[Sh::U];
#pragma unroll
        for (int u = 0; u < Sh::U; ++u) {
            const int jj = j + u * Sh::S;
            if (jj < len) { cc[u] = cis[o0 + jj]; w0[u] = vs[o0 + jj]; w1[u] = vs[o1 + jj]; w2[u] = vs[o2 + jj]; }
            else { cc[u] = -1; w0[u] = w1[u] = w2[u] = 0; }
        }
        VT xv[Sh::U];
#pragma unroll
        for (int u = 0; u < Sh::U; ++u)
            xv[u] = cc[u] >= 0 ? *(const VT *)(X + (size_t)cc[u] * K + c0) : VT(0);
#pragma unroll
        for (int u = 0; u < Sh::U; ++u) { y0 += w0[u] * xv[u]; y1 += w1[u] * xv[u]; y2 += w2[u] * xv[u]; }
    }
}

template <typename T, int K, int EPI>
__global__ __launch_bounds__(kBlock) void k_spmm_node(int32_t nnodes, const int32_t *__restrict__ rp,
                                                      const int32_t *__restrict__ ci, const T *__restrict__ v,
                                                      const T *__restrict__ X, const T *__restrict__ B,
                                                      const T *__restrict__ pre, T *__restrict__ Y,
                                                      const int32_t *__restrict__ done, double *__restrict__ partials,
                                                      int stride, const int32_t *__restrict__ order)
{
    using Sh = NShape<T, K>;
    using VT = typename Vec16<T>::type;
    constexpr int W = Sh::W;
    __shared__ int32_t s_ci[Sh::CAP];
    __shared__ T s_v[Sh::CAP];
    if (done && *done) return;
    const int t = threadIdx.x, g = t / Sh::G, l = t % Sh::G, c0 = (l % Sh::V) * W, sl = l / Sh::V;
    const int ntiles = (nnodes + Sh::NPB - 1) / Sh::NPB;
    const int per = gridDim.x >> 3, xcd = blockIdx.x & 7;
    const int hi = (int)((int64_t)ntiles * (xcd + 1) / 8);
    double dot[W];
#pragma unroll
    for (int w = 0; w < W; ++w) dot[w] = 0;
    for (int pos = (int)((int64_t)ntiles * xcd / 8) + (int)(blockIdx.x >> 3); pos < hi; pos += per) {
        // order: the tiles in y-strips through z, so that an XCD's window of X fits its L2
        const int tile = order ? order[pos] : pos;
        const int q0 = tile * Sh::NPB, q1 = min(q0 + Sh::NPB, nnodes);
        const int e0 = rp[3 * q0], e1 = rp[3 * q1];
        const bool staged = e1 - e0 <= Sh::CAP;     // the same for the whole block
        if (staged) {
            __syncthreads();
            for (int k = t; k < e1 - e0; k += kBlock) { s_ci[k] = ntload(ci + e0 + k); s_v[k] = ntload(v + e0 + k); }
            __syncthreads();
        }
        const int q = q0 + g;
        if (q < q1) {
            const int r = 3 * q;
            const int s0 = rp[r], s1 = rp[r + 1], s2 = rp[r + 2];
            VT y[3] = {VT(0), VT(0), VT(0)};
            if (staged) node_cols<T, K>(s_ci, s_v, s0 - e0, s1 - e0, s2 - e0, s1 - s0, X, c0, sl, y[0], y[1], y[2]);
            else node_cols<T, K>(ci, v, s0, s1, s2, s1 - s0, X, c0, sl, y[0], y[1], y[2]);
#pragma unroll
            for (int o = Sh::V; o < Sh::G; o <<= 1)
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int w = 0; w < W; ++w) y[a][w] += __shfl_xor(y[a][w], o, Sh::G);
            if (sl == 0) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const size_t i = (size_t)(r + a) * K + c0;
                    VT *yp = (VT *)(Y + i);
                    if (EPI == BEPI_SPMM) *yp = y[a];
                    else if (EPI == BEPI_ADD) *yp = *yp + y[a];
                    else if (EPI == BEPI_RESID) *yp = *(const VT *)(B + i) - y[a];
                    else if (EPI == BEPI_JACOBI)
                        *yp = *(const VT *)(X + i) + pre[r + a] * (*(const VT *)(B + i) - y[a]);
                    else {
                        *yp = y[a];
                        const VT xp = *(const VT *)(X + i);
#pragma unroll
                        for (int w = 0; w < W; ++w) dot[w] += (double)xp[w] * (double)y[a][w];
                    }
                }
            }
        }
    }
    if (EPI == BEPI_DOT) publish_vcols<T, K>(dot, partials, stride);
}

int spmm_node_grid(int K, int32_t nnodes)
{
    const int npb = spmm_node_tile(K);
    return grid_tiles((nnodes + npb - 1) / npb);
}

template <typename TV, typename TX>
void launch_spmm_node(int K, int epi, int32_t nnodes, const int32_t *rp, const int32_t *ci, const TV *v, const TX *X,
                      const TX *B, const TV *pre, TX *Y, const int32_t *done, double *partials, int stride,
                      hipStream_t s, const int32_t *order)
{
    static_assert(std::is_same<TV, TX>::value, "node-triple kernels take matrix and vectors of one type");
    if (nnodes <= 0) return;
    const int g = spmm_node_grid(K, nnodes);
#define MGX_SPMMN(KK, EE)                                                                                      \
    hipLaunchKernelGGL((k_spmm_node<TX, KK, EE>), dim3(g), dim3(kBlock), 0, s, nnodes, rp, ci, v, X, B, pre, Y, \
                       done, partials, stride, order)
#define MGX_SPMMN_K(KK)                                       \
    switch (epi) {                                            \
    case BEPI_SPMM: MGX_SPMMN(KK, BEPI_SPMM); break;          \
    case BEPI_ADD: MGX_SPMMN(KK, BEPI_ADD); break;            \
    case BEPI_RESID: MGX_SPMMN(KK, BEPI_RESID); break;        \
    case BEPI_JACOBI: MGX_SPMMN(KK, BEPI_JACOBI); break;      \
    default: MGX_SPMMN(KK, BEPI_DOT); break;                  \
    }
    switch (K) {
    case 4: MGX_SPMMN_K(4); break;
    case 8: MGX_SPMMN_K(8); break;
    case 16: MGX_SPMMN_K(16); break;
    case 32: MGX_SPMMN_K(32); break;
    default: throw Error{MGX_ERR_INVALID_ARGUMENT, "block width must be 4, 8, 16 or 32"};
    }
#undef MGX_SPMMN_K
#undef MGX_SPMMN
    MGX_HIP(hipGetLastError());
}
template void launch_spmm_node<double, double>(int, int, int32_t, const int32_t *, const int32_t *, const double *,
                                               const double *, const double *, const double *, double *,
                                               const int32_t *, double *, int, hipStream_t, const int32_t *);
template void launch_spmm_node<float, float>(int, int, int32_t, const int32_t *, const int32_t *, const float *,
                                             const float *, const float *, const float *, float *, const int32_t *,
                                             double *, int, hipStream_t, const int32_t *);

// Multicolour GS step on node triples: thread (node, s, v) walks every S-th entry of the node's
// three rows for columns c0 .. c0 + W - 1, gathering X once (16 bytes) for the three rows; the
// node's 3x3 diagonal block comes from the same pass.  No LDS and no block barrier (staging the
// node rows through LDS was measured slower: the kernel waits on the X gathers).
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_gs_node(int32_t nnodes, const int32_t *__restrict__ nodes,
                                                    const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                    const T *__restrict__ v, const T *__restrict__ B,
                                                    T *__restrict__ X, int backward, const int32_t *__restrict__ done)
{
    using Sh = NShape<T, K>;
    using VT = typename Vec16<T>::type;
    constexpr int G = Sh::G, N = kBlock / G, S = Sh::S, U = Sh::U, W = Sh::W;
    if (done && *done) return;
    const int t = threadIdx.x, l = t % G, c0 = (l % Sh::V) * W, sl = l / Sh::V;
    const int nd = xcd_remap(blockIdx.x, gridDim.x) * N + t / G;
    if (nd >= nnodes) return;                      // the whole G-lane group leaves together
    const int q3 = 3 * nodes[nd];
    const int s0 = rp[q3], s1 = rp[q3 + 1], s2 = rp[q3 + 2], len = s1 - s0;
    const int o1 = s1 - s0, o2 = s2 - s0;
    const int32_t *cis = ci + s0;
    const T *vs = v + s0;
    VT a0 = VT(0), a1 = VT(0), a2 = VT(0);
    T d[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = sl; j < len; j += U * S) {
        int cc[U];
        T w0[U], w1[U], w2[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int jj = j + u * S;
            if (jj < len) { cc[u] = cis[jj]; w0[u] = vs[jj]; w1[u] = vs[o1 + jj]; w2[u] = vs[o2 + jj]; }
            else { cc[u] = -1; w0[u] = w1[u] = w2[u] = 0; }
        }
        VT xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const unsigned off = (unsigned)(cc[u] - q3);
            xv[u] = (cc[u] >= 0 && off >= 3u) ? *(const VT *)(X + (size_t)cc[u] * K + c0) : VT(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (cc[u] < 0) continue;
            const unsigned off = (unsigned)(cc[u] - q3);
            if (off >= 3u) { a0 += w0[u] * xv[u]; a1 += w1[u] * xv[u]; a2 += w2[u] * xv[u]; }
            else if (off == 0) { d[0] += w0[u]; d[3] += w1[u]; d[6] += w2[u]; }
            else if (off == 1) { d[1] += w0[u]; d[4] += w1[u]; d[7] += w2[u]; }
            else { d[2] += w0[u]; d[5] += w1[u]; d[8] += w2[u]; }
        }
    }
#pragma unroll
    for (int o = Sh::V; o < G; o <<= 1) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            a0[w] += __shfl_xor(a0[w], o, G); a1[w] += __shfl_xor(a1[w], o, G); a2[w] += __shfl_xor(a2[w], o, G);
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) d[k] += __shfl_xor(d[k], o, G);
    }
    if (sl != 0) return;
    const size_t i0 = (size_t)q3 * K + c0;
    VT x0 = *(const VT *)(X + i0), x1 = *(const VT *)(X + i0 + K), x2 = *(const VT *)(X + i0 + 2 * K);
    const VT b0 = *(const VT *)(B + i0) - a0, b1 = *(const VT *)(B + i0 + K) - a1, b2 = *(const VT *)(B + i0 + 2 * K) - a2;
    if (!backward) {
        x0 = (b0 - d[1] * x1 - d[2] * x2) / d[0];
        x1 = (b1 - d[3] * x0 - d[5] * x2) / d[4];
        x2 = (b2 - d[6] * x0 - d[7] * x1) / d[8];
    } else {
        x2 = (b2 - d[6] * x0 - d[7] * x1) / d[8];
        x1 = (b1 - d[3] * x0 - d[5] * x2) / d[4];
        x0 = (b0 - d[1] * x1 - d[2] * x2) / d[0];
    }
    *(VT *)(X + i0) = x0;
    *(VT *)(X + i0 + K) = x1;
    *(VT *)(X + i0 + 2 * K) = x2;
}

template <typename T>
void launch_gs_node(int K, int32_t nnodes, const int32_t *nodes, const int32_t *rp, const int32_t *ci, const T *v,
                    const T *B, T *X, bool backward, const int32_t *done, hipStream_t s)
{
    if (nnodes <= 0) return;
#define MGX_GSN(KK)                                                                                                \
    hipLaunchKernelGGL((k_gs_node<T, KK>), dim3(grid_rows(nnodes, kBlock / NShape<T, KK>::G)), dim3(kBlock), 0, s, \
                       nnodes, nodes, rp, ci, v, B, X, backward ? 1 : 0, done)
    switch (K) {
    case 4: MGX_GSN(4); break;
    case 8: MGX_GSN(8); break;
    case 16: MGX_GSN(16); break;
    case 32: MGX_GSN(32); break;
    default: throw Error{MGX_ERR_INVALID_ARGUMENT, "block width must be 4, 8, 16 or 32"};
    }
#undef MGX_GSN
    MGX_HIP(hipGetLastError());
}
template void launch_gs_node<double>(int, int32_t, const int32_t *, const int32_t *, const int32_t *, const double *,
                                     const double *, double *, bool, const int32_t *, hipStream_t);
template void launch_gs_node<float>(int, int32_t, const int32_t *, const int32_t *, const int32_t *, const float *,
                                    const float *, float *, bool, const int32_t *, hipStream_t);

// ------------------------------------------------------------------ multicolour GS, K columns
// A block of 192 threads relaxes N nodes: thread (m, a, s, c) reduces row a of node m for
// column c over the entries s, s + S, ...: sa = sum over columns outside the node, and the
// node's 3x3 diagonal block (the same for every column).  Then one thread per (node, column)
// relaxes the three dofs in order (reverse order when backward).  With gptr, node q is the group
// of one to three rows gptr[q] .. gptr[q+1] (threads of the missing rows idle).
constexpr int kGsbBlock = 192;

template <typename T, int K>
__global__ __launch_bounds__(kGsbBlock) void k_gs_block(int32_t nnodes, const int32_t *__restrict__ nodes,
                                                        const int32_t *__restrict__ gptr,
                                                        const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                        const T *__restrict__ v, const T *__restrict__ B,
                                                        T *__restrict__ X, int backward, const int32_t *__restrict__ done)
{
    constexpr int S = BShape<K>::S, G = BShape<K>::G, N = kGsbBlock / (3 * G), U = S == 1 ? 8 : 4;
    __shared__ T sh[N][3][K];
    __shared__ T dg[N][3][3];
    if (done && *done) return;
    const int base = xcd_remap(blockIdx.x, gridDim.x) * N;
    const int t = threadIdx.x, l = t % G, c = l % K, sl = l / K, a = (t / G) % 3, m = t / (3 * G);
    const int nd = base + m;
    T sa = 0, d0 = 0, d1 = 0, d2 = 0;
    int r0 = 0, nr = 0;
    if (nd < nnodes) {
        const int q = nodes[nd];
        r0 = gptr ? gptr[q] : 3 * q;
        nr = gptr ? gptr[q + 1] - r0 : 3;
    }
    if (a < nr) {
        const int s = rp[r0 + a], e = rp[r0 + a + 1];
        for (int k = s + sl; k < e; k += U * S) {
            int cc[U];
            T w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kk = k + u * S;
                if (kk < e) { cc[u] = ci[kk]; w[u] = v[kk]; } else { cc[u] = -1; w[u] = 0; }
            }
            T xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const unsigned off = (unsigned)(cc[u] - r0);
                xv[u] = (cc[u] >= 0 && off >= (unsigned)nr) ? X[(size_t)cc[u] * K + c] : T(0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (cc[u] < 0) continue;
                const unsigned off = (unsigned)(cc[u] - r0);
                if (off < (unsigned)nr) { if (off == 0) d0 += w[u]; else if (off == 1) d1 += w[u]; else d2 += w[u]; }
                else sa += w[u] * xv[u];
            }
        }
    }
#pragma unroll
    for (int o = K; o < G; o <<= 1) {
        sa += __shfl_xor(sa, o, G);
        d0 += __shfl_xor(d0, o, G);
        d1 += __shfl_xor(d1, o, G);
        d2 += __shfl_xor(d2, o, G);
    }
    if (nd < nnodes && sl == 0) {
        sh[m][a][c] = sa;
        if (c == 0) { dg[m][a][0] = d0; dg[m][a][1] = d1; dg[m][a][2] = d2; }
    }
    __syncthreads();
    if (t < N * K) {
        const int m2 = t / K, c2 = t % K, nd2 = base + m2;
        if (nd2 < nnodes) {
            const int q = nodes[nd2], g0 = gptr ? gptr[q] : 3 * q, n2 = gptr ? gptr[q + 1] - g0 : 3;
            const size_t q0 = (size_t)g0 * K + c2;
            const T(*D)[3] = dg[m2];
            T xv[3], bv[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                xv[r] = r < n2 ? X[q0 + r * K] : T(0);
                bv[r] = r < n2 ? B[q0 + r * K] - sh[m2][r][c2] : T(0);
            }
#pragma unroll
            for (int t2 = 0; t2 < 3; ++t2) {
                const int r = backward ? 2 - t2 : t2;
                if (r >= n2) continue;
                T y = bv[r];
#pragma unroll
                for (int cc2 = 0; cc2 < 3; ++cc2)
                    if (cc2 != r) y -= D[r][cc2] * xv[cc2];     // zero past the group's rows
                xv[r] = y / D[r][r];
            }
#pragma unroll
            for (int r = 0; r < 3; ++r)
                if (r < n2) X[q0 + r * K] = xv[r];
        }
    }
}

template <typename T>
void launch_gs_block(int K, int32_t nnodes, const int32_t *nodes, const int32_t *rp, const int32_t *ci, const T *v,
                     const T *B, T *X, bool backward, const int32_t *done, hipStream_t s, const int32_t *gptr)
{
    if (nnodes <= 0) return;
#define MGX_GSB(KK)                                                                                          \
    hipLaunchKernelGGL((k_gs_block<T, KK>), dim3(grid_rows(nnodes, kGsbBlock / (3 * BShape<KK>::G))),       \
                       dim3(kGsbBlock), 0, s, nnodes, nodes, gptr, rp, ci, v, B, X, backward ? 1 : 0, done)
    switch (K) {
    case 4: MGX_GSB(4); break;
    case 8: MGX_GSB(8); break;
    case 16: MGX_GSB(16); break;
    case 32: MGX_GSB(32); break;
    default: throw Error{MGX_ERR_INVALID_ARGUMENT, "block width must be 4, 8, 16 or 32"};
    }
#undef MGX_GSB
    MGX_HIP(hipGetLastError());
}
template void launch_gs_block<double>(int, int32_t, const int32_t *, const int32_t *, const int32_t *, const double *,
                                      const double *, double *, bool, const int32_t *, hipStream_t, const int32_t *);
template void launch_gs_block<float>(int, int32_t, const int32_t *, const int32_t *, const int32_t *, const float *,
                                     const float *, float *, bool, const int32_t *, hipStream_t, const int32_t *);

// ------------------------------------------------------------------ block vector kernels
// Every thread keeps one column: blockDim (256) and the grid stride are multiples of K.
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_bnorms(int64_t nk, const T *__restrict__ B, const T *__restrict__ R,
                                                   double *__restrict__ partials, int stride)
{
    double bb = 0, rr = 0, m = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nk; i += (int64_t)gridDim.x * kBlock) {
        const double b = B[i], r = R[i];
        bb += b * b;
        rr += r * r;
        m = nanmax(m, fabs(r));
    }
    publish_cols<K, 3, 2>(bb, rr, m, partials, stride);
}

template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_bdot(int64_t nk, const T *__restrict__ A, const T *__restrict__ Bv,
                                                 const int32_t *__restrict__ done, double *__restrict__ partials,
                                                 int stride)
{
    if (*done) return;
    double d = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nk; i += (int64_t)gridDim.x * kBlock)
        d += (double)A[i] * (double)Bv[i];
    publish_cols<K, 1, -1>(d, 0, 0, partials, stride);
}

// x += alpha_c p, r -= alpha_c q, with per-column ||r||^2 and max |r|
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_bupdate(int64_t nk, const T *__restrict__ P, const T *__restrict__ Q,
                                                    T *__restrict__ X, T *__restrict__ R, const BCtl *__restrict__ ctl,
                                                    double *__restrict__ partials, int stride)
{
    if (ctl->done) return;
    const double a = ctl->alpha[threadIdx.x % K];
    double rr = 0, m = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nk; i += (int64_t)gridDim.x * kBlock) {
        const T r = (T)((double)R[i] - a * (double)Q[i]);
        X[i] = (T)((double)X[i] + a * (double)P[i]);
        R[i] = r;
        rr += (double)r * (double)r;
        m = nanmax(m, fabs((double)r));
    }
    publish_cols<K, 2, 1>(rr, m, 0, partials, stride);
}

// r -= alpha_c q, with per-column ||r||^2 and max |r|
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_bupdate_r(int64_t nk, const T *__restrict__ Q, T *__restrict__ R,
                                                      const BCtl *__restrict__ ctl, double *__restrict__ partials,
                                                      int stride)
{
    if (ctl->done) return;
    const double a = ctl->alpha[threadIdx.x % K];
    double rr = 0, m = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nk; i += (int64_t)gridDim.x * kBlock) {
        const T r = (T)((double)R[i] - a * (double)Q[i]);
        R[i] = r;
        rr += (double)r * (double)r;
        m = nanmax(m, fabs((double)r));
    }
    publish_cols<K, 2, 1>(rr, m, 0, partials, stride);
}

// x += alpha_c p with the p that gave alpha, then p = z + beta_c p; once every column is done
// only the pending x update is applied (xflushed guards against applying it twice)
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_bupdate_xp(int64_t nk, const T *__restrict__ Z, T *__restrict__ P,
                                                       T *__restrict__ X, const BCtl *__restrict__ ctl)
{
    if (ctl->xflushed) return;
    const bool more = !ctl->done;
    const double a = ctl->alpha[threadIdx.x % K], b = ctl->beta[threadIdx.x % K];
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nk; i += (int64_t)gridDim.x * kBlock) {
        const double p = (double)P[i];
        X[i] = (T)((double)X[i] + a * p);
        if (more) P[i] = (T)((double)Z[i] + b * p);
    }
}

// p = z + beta_c p (p = z on the first iteration)
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_bupdate_p(int64_t nk, const T *__restrict__ Z, T *__restrict__ P,
                                                      const BCtl *__restrict__ ctl, int first)
{
    if (ctl->done) return;
    const double b = first ? 0.0 : ctl->beta[threadIdx.x % K];
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nk; i += (int64_t)gridDim.x * kBlock)
        P[i] = first ? Z[i] : (T)((double)Z[i] + b * (double)P[i]);
}

// x = pre_row * b
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_bjacobi_initial(int64_t nk, const T *__restrict__ pre,
                                                            const T *__restrict__ B, T *__restrict__ X,
                                                            const int32_t *__restrict__ done)
{
    if (done && *done) return;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nk; i += (int64_t)gridDim.x * kBlock)
        X[i] = pre[i / K] * B[i];
}

// Per-column scalars from the partials of the producing kernel (one block, 16 waves; wave w
// handles columns w, w + 16).
__global__ __launch_bounds__(1024) void k_bfinalize(int nparts, const double *__restrict__ partials, int stride,
                                                    BCtl *ctl, int mode)
{
    __shared__ int s_done[kMaxBlock];
    if (mode != FIN_NORMS && ctl->done) {
        if (mode == FIN_ALPHA && threadIdx.x == 0) ctl->xflushed = 1;   // the last x update is applied
        return;
    }
    if (threadIdx.x == 0) {
        if (mode == FIN_ALPHA) ctl->xflushed = 0;      // x += alpha p now pending
        if (mode == FIN_NORMS) ctl->xflushed = 1;
    }
    const int K = ctl->k;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int c = w; c < K; c += 16) {
        double a = 0, b = 0, m = 0;
        for (int i = lane; i < nparts; i += 64) {
            a += partials[(size_t)c * stride + i];
            if (mode == FIN_UPDATE) m = nanmax(m, partials[(size_t)(K + c) * stride + i]);
            if (mode == FIN_NORMS) {
                b += partials[(size_t)(K + c) * stride + i];
                m = nanmax(m, partials[(size_t)(2 * K + c) * stride + i]);
            }
        }
        a = wave_sum(a); b = wave_sum(b); m = wave_max(m);
        if (lane == 0) {
            int cd = ctl->cdone[c];
            switch (mode) {
            case FIN_ALPHA: ctl->pq[c] = a; ctl->alpha[c] = cd ? 0.0 : ctl->rz[c] / a; break;
            case FIN_BETA_FIRST: ctl->rz[c] = a; break;
            case FIN_BETA: ctl->beta[c] = cd ? 0.0 : a / ctl->rz[c]; ctl->rz[c] = a; break;
            case FIN_UPDATE:
                if (!cd) {
                    // a = ||r||^2, m = max |r|
                    ctl->rnorm2[c] = a;
                    ctl->rmax[c] = m;
                    const int it = ctl->iter[c] + 1;
                    ctl->iter[c] = it;
                    if (!(m == m) || m > 1e50 || !(a == a)) { cd = 1; ctl->status[c] = -1; }
                    else if (ctl->stop == MGX_STOP_REL_L2 ? (a <= ctl->ctol[c] * ctl->ctol[c] * ctl->bnorm2[c])
                                                          : (m <= ctl->ctol[c])) {
                        cd = 1; ctl->status[c] = 1;
                    } else if (it >= ctl->max_it) cd = 1;
                }
                break;
            case FIN_NORMS: {
                // a = ||b||^2, b = ||r||^2, m = max |r|
                ctl->bnorm2[c] = a;
                ctl->rnorm2[c] = b;
                ctl->rmax[c] = m;
                ctl->iter[c] = 0;
                int st = 0;
                cd = 0;
                if (!(m == m) || m > 1e50) { cd = 1; st = -1; }
                else if (ctl->stop == MGX_STOP_REL_L2 ? (b <= ctl->ctol[c] * ctl->ctol[c] * a) : (m <= ctl->ctol[c])) {
                    cd = 1;
                    st = 1;
                }
                else if (ctl->frozen[c] || ctl->max_it <= 0) cd = 1;
                ctl->status[c] = st;
                break;
            }
            default: break;
            }
            ctl->cdone[c] = cd;
            s_done[c] = cd;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int all = 1;
        for (int c = 0; c < K; ++c) all &= s_done[c];
        ctl->done = all;
    }
}

void launch_bfinalize(int nparts, const double *partials, BCtl *ctl, int mode, hipStream_t s)
{
    hipLaunchKernelGGL(k_bfinalize, dim3(1), dim3(1024), 0, s, nparts, partials, kPartialStride, ctl, mode);
    MGX_HIP(hipGetLastError());
}

static inline int grid_cols(int64_t nk) { return grid_stream(nk); }

#define MGX_K_SWITCH(K, LAUNCH)                                                          \
    switch (K) {                                                                         \
    case 4: LAUNCH(4); break;                                                            \
    case 8: LAUNCH(8); break;                                                            \
    case 16: LAUNCH(16); break;                                                          \
    case 32: LAUNCH(32); break;                                                          \
    default: throw Error{MGX_ERR_INVALID_ARGUMENT, "block width must be 4, 8, 16 or 32"}; \
    }

template <typename T>
void launch_bnorms(int K, int32_t n, const T *B, const T *R, BCtl *ctl, double *partials, hipStream_t s)
{
    const int64_t nk = (int64_t)n * K;
    const int g = grid_cols(nk);
#define L_(KK) hipLaunchKernelGGL((k_bnorms<T, KK>), dim3(g), dim3(kBlock), 0, s, nk, B, R, partials, kPartialStride)
    MGX_K_SWITCH(K, L_)
#undef L_
    MGX_HIP(hipGetLastError());
    launch_bfinalize(g, partials, ctl, FIN_NORMS, s);
}

template <typename T>
void launch_bdot_beta(int K, int32_t n, const T *R, const T *Z, BCtl *ctl, double *partials, bool first, hipStream_t s)
{
    const int64_t nk = (int64_t)n * K;
    const int g = grid_cols(nk);
#define L_(KK) hipLaunchKernelGGL((k_bdot<T, KK>), dim3(g), dim3(kBlock), 0, s, nk, R, Z, &ctl->done, partials, kPartialStride)
    MGX_K_SWITCH(K, L_)
#undef L_
    MGX_HIP(hipGetLastError());
    launch_bfinalize(g, partials, ctl, first ? FIN_BETA_FIRST : FIN_BETA, s);
}

template <typename T>
void launch_bupdate(int K, int32_t n, const T *P, const T *Q, T *X, T *R, BCtl *ctl, double *partials, hipStream_t s)
{
    const int64_t nk = (int64_t)n * K;
    const int g = grid_cols(nk);
#define L_(KK) hipLaunchKernelGGL((k_bupdate<T, KK>), dim3(g), dim3(kBlock), 0, s, nk, P, Q, X, R, ctl, partials, kPartialStride)
    MGX_K_SWITCH(K, L_)
#undef L_
    MGX_HIP(hipGetLastError());
    launch_bfinalize(g, partials, ctl, FIN_UPDATE, s);
}

template <typename T>
void launch_bupdate_r(int K, int32_t n, const T *Q, T *R, BCtl *ctl, double *partials, hipStream_t s)
{
    const int64_t nk = (int64_t)n * K;
    const int g = grid_cols(nk);
#define L_(KK) hipLaunchKernelGGL((k_bupdate_r<T, KK>), dim3(g), dim3(kBlock), 0, s, nk, Q, R, ctl, partials, kPartialStride)
    MGX_K_SWITCH(K, L_)
#undef L_
    MGX_HIP(hipGetLastError());
    launch_bfinalize(g, partials, ctl, FIN_UPDATE, s);
}

template <typename T>
void launch_bupdate_xp(int K, int32_t n, const T *Z, T *P, T *X, const BCtl *ctl, hipStream_t s)
{
    const int64_t nk = (int64_t)n * K;
    const int g = grid_cols(nk);
#define L_(KK) hipLaunchKernelGGL((k_bupdate_xp<T, KK>), dim3(g), dim3(kBlock), 0, s, nk, Z, P, X, ctl)
    MGX_K_SWITCH(K, L_)
#undef L_
    MGX_HIP(hipGetLastError());
}

template <typename T>
void launch_bupdate_p(int K, int32_t n, const T *Z, T *P, const BCtl *ctl, bool first, hipStream_t s)
{
    const int64_t nk = (int64_t)n * K;
    const int g = grid_cols(nk);
#define L_(KK) hipLaunchKernelGGL((k_bupdate_p<T, KK>), dim3(g), dim3(kBlock), 0, s, nk, Z, P, ctl, first ? 1 : 0)
    MGX_K_SWITCH(K, L_)
#undef L_
    MGX_HIP(hipGetLastError());
}

template <typename T>
void launch_bjacobi_initial(int K, int32_t n, const T *pre, const T *B, T *X, const int32_t *done, hipStream_t s)
{
    const int64_t nk = (int64_t)n * K;
    const int g = grid_cols(nk);
#define L_(KK) hipLaunchKernelGGL((k_bjacobi_initial<T, KK>), dim3(g), dim3(kBlock), 0, s, nk, pre, B, X, done)
    MGX_K_SWITCH(K, L_)
#undef L_
    MGX_HIP(hipGetLastError());
}

int spmm_grid(int K, int32_t nrows)
{
    const int rpb = K == 32 ? BShape<32>::RPB : BShape<16>::RPB;   // 8 or 16 rows per tile
    return grid_tiles((nrows + rpb - 1) / rpb);
}

template <typename TV, typename TX>
void launch_spmm_dot_alpha(int K, int32_t n, const int32_t *rp, const int32_t *ci, const TV *v, const TX *P, TX *Q,
                           BCtl *ctl, double *partials, hipStream_t s)
{
    launch_spmm<TV, TX>(K, BEPI_DOT, n, rp, ci, v, P, nullptr, nullptr, Q, &ctl->done, partials, kPartialStride, s);
    launch_bfinalize(spmm_grid(K, n), partials, ctl, FIN_ALPHA, s);
}

#define MGX_INST_T(T)                                                                                             \
    template void launch_bnorms<T>(int, int32_t, const T *, const T *, BCtl *, double *, hipStream_t);             \
    template void launch_bdot_beta<T>(int, int32_t, const T *, const T *, BCtl *, double *, bool, hipStream_t);    \
    template void launch_bupdate<T>(int, int32_t, const T *, const T *, T *, T *, BCtl *, double *, hipStream_t);  \
    template void launch_bupdate_r<T>(int, int32_t, const T *, T *, BCtl *, double *, hipStream_t);                 \
    template void launch_bupdate_xp<T>(int, int32_t, const T *, T *, T *, const BCtl *, hipStream_t);               \
    template void launch_bupdate_p<T>(int, int32_t, const T *, T *, const BCtl *, bool, hipStream_t);              \
    template void launch_bjacobi_initial<T>(int, int32_t, const T *, const T *, T *, const int32_t *, hipStream_t); \
    template void launch_spmm_dot_alpha<T, T>(int, int32_t, const int32_t *, const int32_t *, const T *, const T *, \
                                              T *, BCtl *, double *, hipStream_t);
MGX_INST_T(double)
MGX_INST_T(float)
#undef MGX_INST_T

// ------------------------------------------------------------------ dense coarse solve, K columns
// Y = M X with M the dense FP64 inverse of the coarsest matrix (nc x nc, row-major) and X, Y
// nc x K blocks, split over the inner index: block (row tile, split) takes kDenseRows rows and
// the inner indices of one of kDenseSplit equal slices; thread (slice, c) accumulates column c
// over j = slice, slice + J, ... of that range for all of the tile's rows (one X load feeds
// kDenseRows FMAs); the J slices are summed in LDS and the kDenseSplit split sums are added in a
// fixed order by a second kernel.  Deterministic.
constexpr int kDenseRows = 8, kDenseSplit = 8;

template <typename TX, int K>
__global__ __launch_bounds__(kBlock) void k_dense_split(int32_t nc, const double *__restrict__ M,
                                                        const TX *__restrict__ X, double *__restrict__ part,
                                                        const int32_t *__restrict__ done)
{
    constexpr int J = kBlock / K;
    __shared__ double red[kDenseRows][J][K];
    if (done && *done) return;
    const int r0 = blockIdx.x * kDenseRows, sp = blockIdx.y, t = threadIdx.x, c = t % K, sl = t / K;
    const int nr = min(kDenseRows, nc - r0);
    const int chunk = (nc + kDenseSplit - 1) / kDenseSplit;
    const int j0 = sp * chunk, j1 = min(nc, j0 + chunk);
    double acc[kDenseRows];
#pragma unroll
    for (int r = 0; r < kDenseRows; ++r) acc[r] = 0;
    for (int j = j0 + sl; j < j1; j += J) {
        const double xv = (double)X[(size_t)j * K + c];
#pragma unroll
        for (int r = 0; r < kDenseRows; ++r)
            if (r < nr) acc[r] += M[(size_t)(r0 + r) * nc + j] * xv;
    }
#pragma unroll
    for (int r = 0; r < kDenseRows; ++r) red[r][sl][c] = acc[r];
    __syncthreads();
    if (t < kDenseRows * K) {
        const int r = t / K, cc = t % K;
        if (r < nr) {
            double a = 0;
            for (int q = 0; q < J; ++q) a += red[r][q][cc];
            part[((size_t)sp * nc + r0 + r) * K + cc] = a;
        }
    }
}

template <typename TX>
__global__ __launch_bounds__(kBlock) void k_dense_sum(int64_t nk, const double *__restrict__ part,
                                                      TX *__restrict__ Y, const int32_t *__restrict__ done)
{
    if (done && *done) return;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nk; i += (int64_t)gridDim.x * kBlock) {
        double a = 0;
#pragma unroll
        for (int sp = 0; sp < kDenseSplit; ++sp) a += part[sp * nk + i];
        Y[i] = (TX)a;
    }
}

// The same product on the FP64 matrix cores: v_mfma_f64_16x16x4f64 (A 16 x 4, B 4 x 16, C 16 x 16;
// lane l holds A[l % 16][l / 16], B[l / 16][l % 16] and C[l / 16 + 4 i][l % 16], i < 4).  One
// wave per 16 x 16 output tile (rows of M x columns of X) and inner slice of kDenseSplit; every
// load is guarded against the matrix edge (zeros past it), and the split sums go to the same
// partial layout k_dense_sum adds.  X is widened to FP64 as it is loaded.
typedef double dbl4 __attribute__((ext_vector_type(4)));
template <typename TX, int K>
__global__ __launch_bounds__(64) void k_dense_mfma(int32_t nc, const double *__restrict__ M,
                                                   const TX *__restrict__ X, double *__restrict__ part,
                                                   const int32_t *__restrict__ done)
{
    if (done && *done) return;
    const int l = threadIdx.x;
    const int r0 = blockIdx.x * 16, c0 = blockIdx.y * 16, sp = blockIdx.z;
    const int chunk = ((nc + kDenseSplit - 1) / kDenseSplit + 3) & ~3;
    const int j0 = sp * chunk, j1 = min(nc, j0 + chunk);
    const int ra = r0 + (l & 15), ka = l >> 4, cb = c0 + (l & 15);
    const bool rowok = ra < nc;
    const double *mrow = M + (size_t)(rowok ? ra : 0) * nc;
    dbl4 acc = {0, 0, 0, 0};
    for (int j = j0; j < j1; j += 4) {
        const int k = j + ka;
        const double a = (rowok && k < j1) ? mrow[k] : 0.0;
        const double b = (k < j1) ? (double)X[(size_t)k * K + cb] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = r0 + (l >> 4) + 4 * i;
        if (r < nc) part[((size_t)sp * nc + r) * K + cb] = acc[i];
    }
}

size_t dense_block_scratch(int32_t nc, int K) { return (size_t)kDenseSplit * nc * K; }

template <typename TX>
void launch_dense_block(int K, int32_t nc, const double *M, const TX *X, TX *Y, double *scratch, const int32_t *done,
                        hipStream_t s)
{
    if (K >= 16) {
        const dim3 grid((nc + 15) / 16, K / 16, kDenseSplit);
#define L_(KK) hipLaunchKernelGGL((k_dense_mfma<TX, KK>), grid, dim3(64), 0, s, nc, M, X, scratch, done)
        if (K == 16) L_(16);
        else L_(32);
#undef L_
    } else {
        const dim3 grid((nc + kDenseRows - 1) / kDenseRows, kDenseSplit);
#define L_(KK) hipLaunchKernelGGL((k_dense_split<TX, KK>), grid, dim3(kBlock), 0, s, nc, M, X, scratch, done)
        MGX_K_SWITCH(K, L_)
#undef L_
    }
    MGX_HIP(hipGetLastError());
    const int64_t nk = (int64_t)nc * K;
    hipLaunchKernelGGL((k_dense_sum<TX>), dim3(grid_stream(nk)), dim3(kBlock), 0, s, nk, scratch, Y, done);
    MGX_HIP(hipGetLastError());
}
template void launch_dense_block<double>(int, int32_t, const double *, const double *, double *, double *,
                                         const int32_t *, hipStream_t);
template void launch_dense_block<float>(int, int32_t, const double *, const float *, float *, double *,
                                        const int32_t *, hipStream_t);

// ------------------------------------------------------------------ column-major <-> block layout
// dst[i*K + c] = src[c*ld + i] for c < kc, 0 for kc <= c < K.  A 64-row tile goes through LDS so
// that both the column reads and the row-major writes are coalesced.
template <typename TS, typename TD>
__global__ __launch_bounds__(kBlock) void k_pack(int32_t n, int K, int kc, const TS *__restrict__ src, int64_t ld,
                                                 TD *__restrict__ dst)
{
    __shared__ TD tile[kMaxBlock][65];
    const int i0 = blockIdx.x * 64, t = threadIdx.x;
    for (int c = t / 64; c < K; c += kBlock / 64) {
        const int i = i0 + (t & 63);
        tile[c][t & 63] = (c < kc && i < n) ? (TD)src[(size_t)c * ld + i] : TD(0);
    }
    __syncthreads();
    for (int e = t; e < 64 * K; e += kBlock) {
        const int r = e / K, c = e % K;
        if (i0 + r < n) dst[(size_t)(i0 + r) * K + c] = tile[c][r];
    }
}

// dst[c*ld + i] = src[i*K + c] for c < kc
template <typename TS, typename TD>
__global__ __launch_bounds__(kBlock) void k_unpack(int32_t n, int K, int kc, const TS *__restrict__ src,
                                                   TD *__restrict__ dst, int64_t ld)
{
    __shared__ TS tile[kMaxBlock][65];
    const int i0 = blockIdx.x * 64, t = threadIdx.x;
    for (int e = t; e < 64 * K; e += kBlock) {
        const int r = e / K, c = e % K;
        tile[c][r] = (i0 + r < n) ? src[(size_t)(i0 + r) * K + c] : TS(0);
    }
    __syncthreads();
    for (int c = t / 64; c < kc; c += kBlock / 64) {
        const int i = i0 + (t & 63);
        if (i < n) dst[(size_t)c * ld + i] = (TD)tile[c][t & 63];
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bcopy(int64_t n, const T *__restrict__ src, T *__restrict__ dst,
                                                  const int32_t *__restrict__ done)
{
    if (done && *done) return;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) dst[i] = src[i];
}
template <typename T>
void launch_bcopy(int64_t n, const T *src, T *dst, const int32_t *done, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL((k_bcopy<T>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, src, dst, done);
    MGX_HIP(hipGetLastError());
}
template void launch_bcopy<double>(int64_t, const double *, double *, const int32_t *, hipStream_t);
template void launch_bcopy<float>(int64_t, const float *, float *, const int32_t *, hipStream_t);

void launch_pack_cols(int32_t n, int K, int kc, const double *src, int64_t ld, double *dst, hipStream_t s)
{
    hipLaunchKernelGGL((k_pack<double, double>), dim3((n + 63) / 64), dim3(kBlock), 0, s, n, K, kc, src, ld, dst);
    MGX_HIP(hipGetLastError());
}
void launch_unpack_cols(int32_t n, int K, int kc, const double *src, double *dst, int64_t ld, hipStream_t s)
{
    hipLaunchKernelGGL((k_unpack<double, double>), dim3((n + 63) / 64), dim3(kBlock), 0, s, n, K, kc, src, dst, ld);
    MGX_HIP(hipGetLastError());
}

}  // namespace mgx
