// mgx_kernels.hip -- hand-written gfx950 kernels of the MG-PCG / SpMV path.
//
// All kernels are HBM-bound integer/FP64 streaming work (no MFMA).  Design points:
//  * CSR rows are processed by a group of G lanes of a 64-wide wavefront (G = 4..32 from
//    the mean row length), so the col/val streams of a row are read as contiguous
//    G-element segments; the matrix streams use non-temporal loads so that L2 keeps the
//    gathered vector.
//  * blockIdx is remapped so that each of the 8 XCDs walks one contiguous slab of rows:
//    consecutive blocks are dispatched round-robin over the XCDs, and without the remap every
//    XCD's L2 would fetch the same x window.
//  * Grid reductions are deterministic and fence-free: every block stores its partial with
//    a plain store, and a one-block finalize kernel (the next launch on the stream) sums the
//    partials in a fixed order.
//  * Kernels of the PCG loop take `done` (device flag): once the stop test has fired they
//    return at once, so the host can queue several iterations per synchronisation.
#include "mgx_device.h"

namespace mgx {

// Two-phase grid reductions.  Phase 1: every block of the producing kernel reduces its
// contributions and thread 0 stores them with plain stores into partials[slot * stride + block].
// Phase 2: k_finalize (one block, next launch on the same stream, so the stores are visible)
// sums the partials in a fixed order and updates the solver scalars.  No per-block atomics or
// cache-wide release fences on the hot path.
__device__ __forceinline__ void block_publish(double a, double b, double m, double *partials, int stride, int nslots)
{
    block_reduce3(a, b, m);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = a;
        if (nslots > 1) partials[stride + blockIdx.x] = b;
        if (nslots > 2) partials[2 * stride + blockIdx.x] = m;
    }
}

// Partial sum of one CSR row held by one lane of a G-lane group: the lane takes entries
// s + lane, s + lane + G, ...  Every trip issues U predicated (col, val) loads and then U
// gathers before any FMA, so a K_ff row (<= 81 entries, G = 16, U = 6) costs one round trip.
template <int G, int U, typename T>
__device__ __forceinline__ T row_partial(const int32_t *__restrict__ ci, const T *__restrict__ v,
                                         const T *__restrict__ x, int s, int e, int lane)
{
    T a0 = 0, a1 = 0;
    for (int k = s + lane; k < e; k += U * G) {
        int c[U];
        T w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = k + u * G;
            if (kk < e) { c[u] = ntload(ci + kk); w[u] = ntload(v + kk); }
            else { c[u] = -1; w[u] = 0; }
        }
        T xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = c[u] >= 0 ? x[c[u]] : T(0);
#pragma unroll
        for (int u = 0; u < U; ++u) { if (u & 1) a1 += w[u] * xv[u]; else a0 += w[u] * xv[u]; }
    }
    return a0 + a1;
}

template <int G> struct Unroll { static constexpr int value = G >= 16 ? 6 : 4; };

// Sum of one CSR row by a group of G lanes; every lane of the group gets the total.
template <int G, typename T>
__device__ __forceinline__ T row_dot(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                     const T *__restrict__ v, const T *__restrict__ x, int row, int lane)
{
    T acc = row_partial<G, Unroll<G>::value, T>(ci, v, x, rp[row], rp[row + 1], lane);
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, G);
    return acc;
}

// ------------------------------------------------------------------ SpMV family
enum : int { EPI_SPMV = 0, EPI_SPMV_ADD = 1, EPI_RESID = 2, EPI_JACOBI = 3 };

template <int G, int EPI, typename T>
__global__ __launch_bounds__(kBlock) void k_csr_rows(int32_t nrows, const int32_t *__restrict__ rp,
                                                     const int32_t *__restrict__ ci, const T *__restrict__ v,
                                                     const T *__restrict__ x, const T *__restrict__ b,
                                                     const T *__restrict__ pre, T *__restrict__ y,
                                                     const int32_t *__restrict__ done)
{
    if (done && *done) return;
    constexpr int RPB = kBlock / G;
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int row = blk * RPB + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    if (row >= nrows) return;   // whole group leaves together
    const T s = row_dot<G, T>(rp, ci, v, x, row, lane);
    if (lane == 0) {
        if (EPI == EPI_SPMV) y[row] = s;
        else if (EPI == EPI_SPMV_ADD) y[row] = y[row] + s;
        else if (EPI == EPI_RESID) y[row] = b[row] - s;
        else y[row] = x[row] + pre[row] * (b[row] - s);
    }
}


// CSR-stream SpMV for the stiffness matrices (mean row length >= 16 lanes' worth): a 256-thread
// block takes a tile of RB consecutive rows whose (col, val) range is contiguous; every thread
// loads E entries at stride 256 (each wave load is one contiguous, fully used segment), gathers
// x, and stores the products in LDS; then G = 256 / RB lanes sum each row's products.  Tiles
// longer than the LDS chunk are processed in chunks, so any row length is handled.  Each XCD
// walks a contiguous eighth of the tiles; with a capped grid (the DOT form, whose per-block
// partials feed k_finalize) blocks loop over the tiles of their eighth.
template <typename T> struct StreamShape {
    static constexpr int RB = sizeof(T) == 8 ? 64 : 32;                 // rows per tile
    static constexpr int G = kBlock / RB;                                // lanes per row in the sum
    static constexpr int E = (RB * 84 + kBlock - 1) / kBlock;           // entries per thread per chunk
    static constexpr int C = E * kBlock;                                 // entries per chunk
};

template <typename T, int EPI, bool DOT>
__global__ __launch_bounds__(kBlock) void k_csr_stream(int32_t nrows, const int32_t *__restrict__ rp,
                                                       const int32_t *__restrict__ ci, const T *__restrict__ v,
                                                       const T *__restrict__ x, const T *__restrict__ b,
                                                       const T *__restrict__ pre, T *__restrict__ y,
                                                       const int32_t *__restrict__ done, double *__restrict__ partials)
{
    using Sh = StreamShape<T>;
    constexpr int RB = Sh::RB, G = Sh::G, E = Sh::E, C = Sh::C;
    __shared__ T sp[C];
    if (done && *done) return;
    const int t = threadIdx.x, lane = t % G;
    const int ntiles = (nrows + RB - 1) / RB;
    const int per = gridDim.x >> 3, xcd = blockIdx.x & 7;
    const int tend = (int)((int64_t)ntiles * (xcd + 1) / 8);
    double dot = 0;
    bool first = true;
    for (int tile = (int)((int64_t)ntiles * xcd / 8) + (int)(blockIdx.x >> 3); tile < tend; tile += per) {
        const int r0 = tile * RB, r1 = min(r0 + RB, nrows);
        const int e0 = rp[r0], cnt = rp[r1] - e0;
        const int row = r0 + t / G;
        int s = 0, e = 0;
        if (row < r1) { s = rp[row] - e0; e = rp[row + 1] - e0; }
        T a = 0;
        if (!first) __syncthreads();                  // the previous tile's readers are done
        first = false;
        if (cnt <= C) {
            // the common case: the whole tile in one chunk, all loads issued before any use
            int c[E];
            T w[E];
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const int k = t + kBlock * j;
                if (k < cnt) { c[j] = ntload(ci + e0 + k); w[j] = ntload(v + e0 + k); } else { c[j] = 0; w[j] = 0; }
            }
            T xv[E];
#pragma unroll
            for (int j = 0; j < E; ++j) xv[j] = x[c[j]];
#pragma unroll
            for (int j = 0; j < E; ++j) sp[t + kBlock * j] = w[j] * xv[j];
            __syncthreads();
            for (int k = s + lane; k < e; k += G) a += sp[k];
        } else {
            for (int base = 0; base < cnt; base += C) {
                if (base > 0) __syncthreads();
                for (int j = 0; j < E; ++j) {
                    const int k = base + t + kBlock * j;
                    sp[t + kBlock * j] = k < cnt ? ntload(v + e0 + k) * x[ntload(ci + e0 + k)] : T(0);
                }
                __syncthreads();
                const int lo = max(s, base), hi = min(e, base + C);
                for (int k = lo + lane; k < hi; k += G) a += sp[k - base];
            }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) a += __shfl_xor(a, o, G);
        if (lane == 0 && row < r1) {
            if (DOT) { y[row] = a; dot += (double)x[row] * (double)a; }
            else if (EPI == EPI_SPMV) y[row] = a;
            else if (EPI == EPI_SPMV_ADD) y[row] = y[row] + a;
            else if (EPI == EPI_RESID) y[row] = b[row] - a;
            else y[row] = x[row] + pre[row] * (b[row] - a);
        }
    }
    if (DOT) block_publish(dot, 0, 0, partials, 0, 1);
}

static inline int stream_tiles(int64_t nrows, int rb) { return (int)((nrows + rb - 1) / rb); }

template <int EPI, typename T>
static void dispatch_rows(int32_t nrows, const int32_t *rp, const int32_t *ci, const T *v, const T *x,
                          const T *b, const T *pre, T *y, int lanes, const int32_t *done, hipStream_t s)
{
    if (nrows <= 0) return;
    if (lanes >= 16) {
        const int nb = grid_rows(stream_tiles(nrows, StreamShape<T>::RB), 1);
        hipLaunchKernelGGL((k_csr_stream<T, EPI, false>), dim3(nb), dim3(kBlock), 0, s, nrows, rp, ci, v, x, b, pre, y,
                           done, nullptr);
        MGX_HIP(hipGetLastError());
        return;
    }
#define MGX_ROWS(G)                                                                                      \
    hipLaunchKernelGGL((k_csr_rows<G, EPI, T>), dim3(grid_rows(nrows, kBlock / G)), dim3(kBlock), 0, s, \
                       nrows, rp, ci, v, x, b, pre, y, done)
    switch (lanes) {
    case 2: MGX_ROWS(2); break;
    case 4: MGX_ROWS(4); break;
    case 8: MGX_ROWS(8); break;
    case 16: MGX_ROWS(16); break;
    case 32: MGX_ROWS(32); break;
    default: MGX_ROWS(64); break;
    }
#undef MGX_ROWS
    MGX_HIP(hipGetLastError());
}

int lanes_for(int64_t nnz, int32_t nrows)
{
    double avg = nrows > 0 ? (double)nnz / nrows : 0;
    if (avg <= 3) return 2;
    if (avg <= 6) return 4;
    if (avg <= 14) return 8;
    if (avg <= 120) return 16;
    if (avg <= 240) return 32;
    return 64;
}

void launch_spmv(int32_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *x, double *y,
                 bool clear, int lanes, const int32_t *done, hipStream_t s)
{
    if (clear) dispatch_rows<EPI_SPMV, double>(n, rp, ci, v, x, nullptr, nullptr, y, lanes, done, s);
    else dispatch_rows<EPI_SPMV_ADD, double>(n, rp, ci, v, x, nullptr, nullptr, y, lanes, done, s);
}
void launch_spmv_f32(int32_t n, const int32_t *rp, const int32_t *ci, const float *v, const float *x, float *y,
                     bool clear, int lanes, const int32_t *done, hipStream_t s)
{
    if (clear) dispatch_rows<EPI_SPMV, float>(n, rp, ci, v, x, nullptr, nullptr, y, lanes, done, s);
    else dispatch_rows<EPI_SPMV_ADD, float>(n, rp, ci, v, x, nullptr, nullptr, y, lanes, done, s);
}
void launch_residual(int32_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *b,
                     const double *x, double *r, int lanes, const int32_t *done, hipStream_t s)
{
    dispatch_rows<EPI_RESID, double>(n, rp, ci, v, x, b, nullptr, r, lanes, done, s);
}
void launch_residual_f32(int32_t n, const int32_t *rp, const int32_t *ci, const float *v, const float *b,
                         const float *x, float *r, int lanes, const int32_t *done, hipStream_t s)
{
    dispatch_rows<EPI_RESID, float>(n, rp, ci, v, x, b, nullptr, r, lanes, done, s);
}
void launch_jacobi(int32_t n, const double *pre, const int32_t *rp, const int32_t *ci, const double *v,
                   const double *b, const double *x, double *y, int lanes, const int32_t *done, hipStream_t s)
{
    dispatch_rows<EPI_JACOBI, double>(n, rp, ci, v, x, b, pre, y, lanes, done, s);
}
void launch_jacobi_f32(int32_t n, const float *pre, const int32_t *rp, const int32_t *ci, const float *v,
                       const float *b, const float *x, float *y, int lanes, const int32_t *done, hipStream_t s)
{
    dispatch_rows<EPI_JACOBI, float>(n, rp, ci, v, x, b, pre, y, lanes, done, s);
}

// residual_with_check: same flag protocol as CsrGeometricMultigrid.cl:121-141
template <int G>
__global__ __launch_bounds__(kBlock) void k_residual_flag(int32_t nrows, const int32_t *__restrict__ rp,
                                                          const int32_t *__restrict__ ci, const double *__restrict__ v,
                                                          const double *__restrict__ b, const double *__restrict__ x,
                                                          double *__restrict__ r, double tol, int32_t *flag,
                                                          const int32_t *__restrict__ done)
{
    constexpr int RPB = kBlock / G;
    if (done && *done) return;
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int row = blk * RPB + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    if (row >= nrows) return;
    const double s = row_dot<G, double>(rp, ci, v, x, row, lane);
    if (lane == 0) {
        const double ri = b[row] - s;
        r[row] = ri;
        const double c = fabs(ri);
        if (ri != ri || c > 1e50) atomicOr(flag, 2);
        else if (c > tol) atomicAnd(flag, 2);
    }
}

void launch_residual_flag(int32_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *b,
                          const double *x, double *r, double tol, int32_t *flag, int lanes, hipStream_t s,
                          const int32_t *done)
{
    if (n <= 0) return;
#define MGX_RF(G) hipLaunchKernelGGL((k_residual_flag<G>), dim3(grid_rows(n, kBlock / G)), dim3(kBlock), 0, s, \
                                     n, rp, ci, v, b, x, r, tol, flag, done)
    switch (lanes) {
    case 2: MGX_RF(2); break;
    case 4: MGX_RF(4); break;
    case 8: MGX_RF(8); break;
    case 16: MGX_RF(16); break;
    case 32: MGX_RF(32); break;
    default: MGX_RF(64); break;
    }
#undef MGX_RF
    MGX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ streaming vector kernels

template <typename T>
__global__ __launch_bounds__(kBlock) void k_jacobi_initial(int32_t n, const T *__restrict__ pre, const T *__restrict__ b,
                                                           T *__restrict__ x, const int32_t *__restrict__ done)
{
    if (done && *done) return;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        x[i] = pre[i] * b[i];
}

void launch_jacobi_initial(int32_t n, const double *pre, const double *b, double *x, const int32_t *done, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_jacobi_initial<double>, dim3(grid_stream(n)), dim3(kBlock), 0, s, n, pre, b, x, done);
    MGX_HIP(hipGetLastError());
}
void launch_jacobi_initial_f32(int32_t n, const float *pre, const float *b, float *x, const int32_t *done, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_jacobi_initial<float>, dim3(grid_stream(n)), dim3(kBlock), 0, s, n, pre, b, x, done);
    MGX_HIP(hipGetLastError());
}

// phase 1 of a dot product (FP64 accumulation for FP64 and FP32 vectors)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot_partial(int32_t n, const T *__restrict__ a,
                                                        const T *__restrict__ b, const int32_t *__restrict__ done,
                                                        double *partials)
{
    if (done && *done) return;
    double s = 0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        s += (double)a[i] * (double)b[i];
    block_publish(s, 0, 0, partials, 0, 1);
}

// phase 2: one block sums the partials in a fixed order and applies `mode`
__global__ __launch_bounds__(1024) void k_finalize(int nparts, const double *__restrict__ partials, int stride,
                                                   Ctl *ctl, int mode, double *result)
{
    if (mode != FIN_NORMS && mode != FIN_DOT && ctl->done) return;
    const int nslots = (mode == FIN_UPDATE || mode == FIN_NORMS) ? 3 : 1;
    double a = 0, b = 0, m = 0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        a += partials[i];
        if (nslots > 1) { b += partials[stride + i]; m = nanmax(m, partials[2 * stride + i]); }
    }
    block_reduce3(a, b, m);
    if (threadIdx.x != 0) return;
    switch (mode) {
    case FIN_DOT: *result = a; break;
    case FIN_ALPHA: ctl->pq = a; ctl->alpha = ctl->rz / a; break;
    case FIN_BETA_FIRST: ctl->rz = a; break;
    case FIN_BETA: ctl->beta = a / ctl->rz; ctl->rz = a; break;
    case FIN_UPDATE: {
        // a = ||r||^2 (slot 0 holds it for FIN_UPDATE), m = max |r|
        ctl->rnorm2 = a;
        ctl->rmax = m;
        const int it = ctl->iter + 1;
        ctl->iter = it;
        int done = 0, status = 0;
        if (!(m == m) || m > 1e50 || !(a == a)) { done = 1; status = -1; }
        else if (ctl->stop == MGX_STOP_REL_L2 ? (a <= ctl->tol * ctl->tol * ctl->bnorm2) : (m <= ctl->tol)) {
            done = 1; status = 1;
        } else if (it >= ctl->max_it) done = 1;
        ctl->status = status;
        ctl->done = done;
        break;
    }
    case FIN_NORMS: {
        // a = ||b||^2, b = ||r||^2, m = max |r|
        ctl->bnorm2 = a;
        ctl->rnorm2 = b;
        ctl->rmax = m;
        ctl->iter = 0;
        int done = 0, status = 0;
        if (!(m == m) || m > 1e50) { done = 1; status = -1; }
        else if (ctl->stop == MGX_STOP_REL_L2 ? (b <= ctl->tol * ctl->tol * a) : (m <= ctl->tol)) { done = 1; status = 1; }
        else if (ctl->max_it <= 0) done = 1;
        ctl->status = status;
        ctl->done = done;
        break;
    }
    }
}

// First stage for long one-slot partial lists (the per-tile partials of the streaming SpMV + dot):
// block b sums partials[b*2048 .. b*2048 + 2048) in a fixed order into out[b].
__global__ __launch_bounds__(kBlock) void k_partials_reduce(int nparts, const double *__restrict__ partials,
                                                            double *__restrict__ out, const Ctl *ctl)
{
    if (ctl->done) return;
    double a = 0, b = 0, m = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int i = blockIdx.x * 2048 + j * kBlock + threadIdx.x;
        if (i < nparts) a += partials[i];
    }
    block_reduce3(a, b, m);
    if (threadIdx.x == 0) out[blockIdx.x] = a;
}

void launch_finalize(int nparts, const double *partials, int stride, Ctl *ctl, int mode, double *result, hipStream_t s)
{
    const bool one_slot = mode != FIN_UPDATE && mode != FIN_NORMS;
    if (one_slot && nparts > 4096 && ctl) {
        // the stage-1 sums go right after the partials (the buffer holds at least 3x as many)
        double *out = const_cast<double *>(partials) + ((nparts + 255) & ~255);
        const int g1 = (nparts + 2047) / 2048;
        hipLaunchKernelGGL(k_partials_reduce, dim3(g1), dim3(kBlock), 0, s, nparts, partials, out, ctl);
        MGX_HIP(hipGetLastError());
        partials = out;
        nparts = g1;
    }
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, s, nparts, partials, stride, ctl, mode, result);
    MGX_HIP(hipGetLastError());
}

int dot_partial_count(int64_t n) { return grid_stream(n); }

void launch_dot(int32_t n, const double *a, const double *b, double *result, double *partials, uint32_t *counter,
                hipStream_t s)
{
    (void)counter;
    const int g = grid_stream(n);
    hipLaunchKernelGGL(k_dot_partial<double>, dim3(g), dim3(kBlock), 0, s, n, a, b, nullptr, partials);
    MGX_HIP(hipGetLastError());
    launch_finalize(g, partials, 0, nullptr, FIN_DOT, result, s);
}

// ||b||^2, ||r||^2, max|r| before the first iteration (r = b - A x0 or b)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_norms(int32_t n, const T *__restrict__ b,
                                                      const T *__restrict__ r, double *partials, int stride)
{
    double sb = 0, sr = 0, m = 0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double bi = b[i], ri = r[i];
        sb += bi * bi;
        sr += ri * ri;
        m = nanmax(m, fabs(ri));
    }
    block_publish(sb, sr, m, partials, stride, 3);
}

void launch_pcg_norms(int32_t n, const double *b, const double *r, Ctl *ctl, double *partials, int stride, hipStream_t s)
{
    const int g = grid_stream(n);
    hipLaunchKernelGGL(k_pcg_norms<double>, dim3(g), dim3(kBlock), 0, s, n, b, r, partials, stride);
    MGX_HIP(hipGetLastError());
    launch_finalize(g, partials, stride, ctl, FIN_NORMS, nullptr, s);
}

void launch_pcg_norms_f32(int32_t n, const float *b, const float *r, Ctl *ctl, double *partials, int stride, hipStream_t s)
{
    const int g = grid_stream(n);
    hipLaunchKernelGGL(k_pcg_norms<float>, dim3(g), dim3(kBlock), 0, s, n, b, r, partials, stride);
    MGX_HIP(hipGetLastError());
    launch_finalize(g, partials, stride, ctl, FIN_NORMS, nullptr, s);
}

// q = A p and the per-block partials of p.q (phase 1 of alpha = rz / pq).  Same row mapping as
// k_csr_rows (one G-lane group per row, XCD-remapped blocks).
template <int G, typename T>
__global__ __launch_bounds__(kBlock) void k_spmv_dot(int32_t nrows, const int32_t *__restrict__ rp,
                                                     const int32_t *__restrict__ ci, const T *__restrict__ v,
                                                     const T *__restrict__ p, T *__restrict__ q,
                                                     const int32_t *__restrict__ done, double *partials)
{
    if (*done) return;
    constexpr int RPB = kBlock / G;
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int row = blk * RPB + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    double contrib = 0;
    if (row < nrows) {
        const T s = row_dot<G, T>(rp, ci, v, p, row, lane);
        if (lane == 0) { q[row] = s; contrib = (double)p[row] * (double)s; }
    }
    block_publish(contrib, 0, 0, partials, 0, 1);
}

template <typename T>
int launch_spmv_dot_t(int32_t n, const int32_t *rp, const int32_t *ci, const T *v, const T *p, T *q,
                      int lanes, const int32_t *done, double *partials, hipStream_t s)
{
    int grid = 0;
    if (lanes >= 16) {
        grid = grid_rows(stream_tiles(n, StreamShape<T>::RB), 1);
        hipLaunchKernelGGL((k_csr_stream<T, EPI_SPMV, true>), dim3(grid), dim3(kBlock), 0, s, n, rp, ci, v, p, nullptr,
                           nullptr, q, done, partials);
        MGX_HIP(hipGetLastError());
        return grid;
    }
#define MGX_SDA(G) grid = grid_rows(n, kBlock / G); \
    hipLaunchKernelGGL((k_spmv_dot<G, T>), dim3(grid), dim3(kBlock), 0, s, n, rp, ci, v, p, q, done, partials)
    switch (lanes) {
    case 2: MGX_SDA(2); break;
    case 4: MGX_SDA(4); break;
    case 8: MGX_SDA(8); break;
    case 16: MGX_SDA(16); break;
    case 32: MGX_SDA(32); break;
    default: MGX_SDA(64); break;
    }
#undef MGX_SDA
    MGX_HIP(hipGetLastError());
    return grid;
}

int reduce_slots_for_rows(int64_t nrows, int lanes) { return grid_rows(nrows, kBlock / lanes); }

void launch_spmv_dot_alpha(int32_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *p,
                           double *q, int lanes, Ctl *ctl, double *partials, hipStream_t s)
{
    const int g = launch_spmv_dot_t<double>(n, rp, ci, v, p, q, lanes, &ctl->done, partials, s);
    launch_finalize(g, partials, 0, ctl, FIN_ALPHA, nullptr, s);
}

void launch_spmv_dot_alpha_f32(int32_t n, const int32_t *rp, const int32_t *ci, const float *v, const float *p,
                               float *q, int lanes, Ctl *ctl, double *partials, hipStream_t s)
{
    const int g = launch_spmv_dot_t<float>(n, rp, ci, v, p, q, lanes, &ctl->done, partials, s);
    launch_finalize(g, partials, 0, ctl, FIN_ALPHA, nullptr, s);
}

// alpha = rz / p.q from a q = A p computed by another kernel (the DuVi SpMV)
template <typename T>
void launch_dot_alpha(int32_t n, const T *p, const T *q, Ctl *ctl, double *partials, hipStream_t s)
{
    const int g = grid_stream(n);
    hipLaunchKernelGGL(k_dot_partial<T>, dim3(g), dim3(kBlock), 0, s, n, p, q, &ctl->done, partials);
    MGX_HIP(hipGetLastError());
    launch_finalize(g, partials, 0, ctl, FIN_ALPHA, nullptr, s);
}
template void launch_dot_alpha<double>(int32_t, const double *, const double *, Ctl *, double *, hipStream_t);
template void launch_dot_alpha<float>(int32_t, const float *, const float *, Ctl *, double *, hipStream_t);

// x += alpha p; r -= alpha q; partials of ||r||^2 and max|r| (phase 1 of the stop test)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int32_t n, const T *__restrict__ p,
                                                       const T *__restrict__ q, T *__restrict__ x,
                                                       T *__restrict__ r, const Ctl *ctl, double *partials, int stride)
{
    if (ctl->done) return;
    const double alpha = ctl->alpha;
    double s = 0, m = 0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        x[i] = (T)((double)x[i] + alpha * (double)p[i]);
        const T ri = (T)((double)r[i] - alpha * (double)q[i]);
        r[i] = ri;
        s += (double)ri * (double)ri;
        m = nanmax(m, fabs((double)ri));
    }
    block_publish(s, 0, m, partials, stride, 3);
}

void launch_pcg_update(int32_t n, const double *p, const double *q, double *x, double *r, Ctl *ctl,
                       double *partials, hipStream_t s)
{
    const int g = grid_stream(n);
    const int stride = kPartialStride;
    hipLaunchKernelGGL(k_pcg_update<double>, dim3(g), dim3(kBlock), 0, s, n, p, q, x, r, ctl, partials, stride);
    MGX_HIP(hipGetLastError());
    launch_finalize(g, partials, stride, ctl, FIN_UPDATE, nullptr, s);
}

void launch_pcg_update_f32(int32_t n, const float *p, const float *q, float *x, float *r, Ctl *ctl,
                           double *partials, hipStream_t s)
{
    const int g = grid_stream(n);
    const int stride = kPartialStride;
    hipLaunchKernelGGL(k_pcg_update<float>, dim3(g), dim3(kBlock), 0, s, n, p, q, x, r, ctl, partials, stride);
    MGX_HIP(hipGetLastError());
    launch_finalize(g, partials, stride, ctl, FIN_UPDATE, nullptr, s);
}

void launch_dot_beta(int32_t n, const double *r, const double *z, Ctl *ctl, double *partials, bool first, hipStream_t s)
{
    const int g = grid_stream(n);
    hipLaunchKernelGGL(k_dot_partial<double>, dim3(g), dim3(kBlock), 0, s, n, r, z, &ctl->done, partials);
    MGX_HIP(hipGetLastError());
    launch_finalize(g, partials, 0, ctl, first ? FIN_BETA_FIRST : FIN_BETA, nullptr, s);
}

void launch_dot_beta_f32(int32_t n, const float *r, const float *z, Ctl *ctl, double *partials, bool first, hipStream_t s)
{
    const int g = grid_stream(n);
    hipLaunchKernelGGL(k_dot_partial<float>, dim3(g), dim3(kBlock), 0, s, n, r, z, &ctl->done, partials);
    MGX_HIP(hipGetLastError());
    launch_finalize(g, partials, 0, ctl, first ? FIN_BETA_FIRST : FIN_BETA, nullptr, s);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_update_p(int32_t n, const T *__restrict__ z, T *__restrict__ p,
                                                     const Ctl *ctl, int first)
{
    if (ctl->done) return;
    const double beta = first ? 0.0 : ctl->beta;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        p[i] = first ? z[i] : (T)((double)z[i] + beta * (double)p[i]);
}

void launch_update_p(int32_t n, const double *z, double *p, const Ctl *ctl, bool first, hipStream_t s)
{
    hipLaunchKernelGGL(k_update_p<double>, dim3(grid_stream(n)), dim3(kBlock), 0, s, n, z, p, ctl, first ? 1 : 0);
    MGX_HIP(hipGetLastError());
}

void launch_update_p_f32(int32_t n, const float *z, float *p, const Ctl *ctl, bool first, hipStream_t s)
{
    hipLaunchKernelGGL(k_update_p<float>, dim3(grid_stream(n)), dim3(kBlock), 0, s, n, z, p, ctl, first ? 1 : 0);
    MGX_HIP(hipGetLastError());
}

// coarse solve of the precision form: y (float) = M (FP64) x (float), FP64 accumulation
__global__ __launch_bounds__(kBlock) void k_dense_gemv_f32io(int32_t n, const double *__restrict__ M,
                                                             const float *__restrict__ x, float *__restrict__ y,
                                                             const int32_t *__restrict__ done)
{
    if (done && *done) return;
    const int row = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double *Mr = M + (int64_t)row * n;
    double s = 0;
    for (int j = lane; j < n; j += 64) s += Mr[j] * (double)x[j];
    s = wave_sum(s);
    if (lane == 0) y[row] = (float)s;
}

void launch_dense_gemv_f32io(int32_t n, const double *M, const float *x, float *y, const int32_t *done, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_dense_gemv_f32io, dim3((n + 3) / 4), dim3(kBlock), 0, s, n, M, x, y, done);
    MGX_HIP(hipGetLastError());
}

template <typename A, typename B>
__global__ __launch_bounds__(kBlock) void k_convert(int64_t n, const A *__restrict__ a, B *__restrict__ b,
                                                    const int32_t *__restrict__ done)
{
    if (done && *done) return;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) b[i] = (B)a[i];
}
void launch_convert_f64_f32(int64_t n, const double *a, float *b, const int32_t *done, hipStream_t s)
{
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, a, b, done);
    MGX_HIP(hipGetLastError());
}
void launch_convert_f32_f64(int64_t n, const float *a, double *b, const int32_t *done, hipStream_t s)
{
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, a, b, done);
    MGX_HIP(hipGetLastError());
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_zero(int64_t n, T *__restrict__ a, const int32_t *__restrict__ done)
{
    if (done && *done) return;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) a[i] = T(0);
}
void launch_fill_zero_f32(int64_t n, float *a, const int32_t *done, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_zero<float>, dim3(grid_stream(n)), dim3(kBlock), 0, s, n, a, done);
    MGX_HIP(hipGetLastError());
}
void launch_fill_zero(int64_t n, double *a, const int32_t *done, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_zero<double>, dim3(grid_stream(n)), dim3(kBlock), 0, s, n, a, done);
    MGX_HIP(hipGetLastError());
}

__global__ __launch_bounds__(kBlock) void k_axpy_f32_into_f64(int64_t n, const float *__restrict__ a,
                                                              double *__restrict__ x, const int32_t *__restrict__ done)
{
    if (done && *done) return;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        x[i] = x[i] + (double)a[i];
}
void launch_axpy_f32_into_f64(int64_t n, const float *a, double *x, const int32_t *done, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpy_f32_into_f64, dim3(grid_stream(n)), dim3(kBlock), 0, s, n, a, x, done);
    MGX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ multicolour Gauss-Seidel
// A block of 192 threads relaxes 4 nodes (12 row slots): each row is reduced by 16 lanes into
// s_a = sum over columns outside the node and d_ab = the node's own block, with all loads of the
// row issued up front; the per-row results go through LDS and one thread per node relaxes its
// dofs in order (reverse order when backward).  A node is 3 rows, or 1 to 3 with a group pointer
// (nodes whose free dofs differ in number; slots past a node's rows stay idle).  Nodes of one
// colour are uncoupled, so the x values read here are not written by this launch except the
// node's own.
constexpr int kGsBlock = 192, kGsNodes = 4;

template <typename T>
__global__ __launch_bounds__(kGsBlock) void k_gs_color(int32_t nnodes, const int32_t *__restrict__ nodes,
                                                       const int32_t *__restrict__ gptr,
                                                       const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                       const T *__restrict__ v, const T *__restrict__ b, T *__restrict__ x,
                                                       int backward, const int32_t *__restrict__ done)
{
    constexpr int G = 16, U = 6;
    __shared__ T sh[kGsNodes][3][4];
    if (done && *done) return;
    const int base = xcd_remap(blockIdx.x, gridDim.x) * kGsNodes;
    const int rowslot = threadIdx.x / G;           // 0..11
    const int ln = rowslot / 3, a = rowslot % 3;   // node within block, row within node
    const int lane = threadIdx.x % G;
    const int grp = base + ln;
    T sa = 0, d0 = 0, d1 = 0, d2 = 0;
    int r0 = 0, gs = 3;
    if (grp < nnodes) {
        const int g = nodes[grp];
        r0 = gptr ? gptr[g] : 3 * g;
        if (gptr) gs = gptr[g + 1] - r0;
    }
    if (grp < nnodes && a < gs) {
        const int s = rp[r0 + a], e = rp[r0 + a + 1];
        for (int k = s + lane; k < e; k += U * G) {
            int c[U];
            T w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kk = k + u * G;
                if (kk < e) { c[u] = ci[kk]; w[u] = v[kk]; } else { c[u] = -1; w[u] = 0; }
            }
            T xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const unsigned off = (unsigned)(c[u] - r0);
                xv[u] = (c[u] >= 0 && off >= (unsigned)gs) ? x[c[u]] : T(0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const unsigned off = (unsigned)(c[u] - r0);
                if (c[u] < 0) continue;
                if (off < (unsigned)gs) { if (off == 0) d0 += w[u]; else if (off == 1) d1 += w[u]; else d2 += w[u]; }
                else sa += w[u] * xv[u];
            }
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        sa += __shfl_xor(sa, o, G);
        d0 += __shfl_xor(d0, o, G);
        d1 += __shfl_xor(d1, o, G);
        d2 += __shfl_xor(d2, o, G);
    }
    if (lane == 0) { sh[ln][a][0] = sa; sh[ln][a][1] = d0; sh[ln][a][2] = d1; sh[ln][a][3] = d2; }
    __syncthreads();
    if (threadIdx.x < kGsNodes && base + (int)threadIdx.x < nnodes) {
        // the node's rows in order (reverse when backward): x_a = (b_a - s_a - sum_{c != a} d_ac x_c) / d_aa
        const int m = threadIdx.x;
        const int g = nodes[base + m];
        const int q0 = gptr ? gptr[g] : 3 * g, n = gptr ? gptr[g + 1] - q0 : 3;
        const T(*R)[4] = sh[m];
        T xv[3], bv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xv[c] = c < n ? x[q0 + c] : T(0);
            bv[c] = c < n ? b[q0 + c] - R[c][0] : T(0);
        }
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int r = backward ? 2 - t : t;
            if (r >= n) continue;
            T acc = bv[r];
#pragma unroll
            for (int c = 0; c < 3; ++c)
                if (c != r && c < n) acc -= R[r][c + 1] * xv[c];
            xv[r] = acc / R[r][r + 1];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (c < n) x[q0 + c] = xv[c];
    }
}

void launch_gs_color(int32_t nnodes, const int32_t *nodes, const int32_t *rp, const int32_t *ci, const double *v,
                     const double *b, double *x, bool backward, const int32_t *done, hipStream_t s,
                     const int32_t *gptr)
{
    if (nnodes <= 0) return;
    const int nb = grid_rows(nnodes, kGsNodes);
    hipLaunchKernelGGL(k_gs_color<double>, dim3(nb), dim3(kGsBlock), 0, s, nnodes, nodes, gptr, rp, ci, v, b, x,
                       backward ? 1 : 0, done);
    MGX_HIP(hipGetLastError());
}
void launch_gs_color_f32(int32_t nnodes, const int32_t *nodes, const int32_t *rp, const int32_t *ci, const float *v,
                         const float *b, float *x, bool backward, const int32_t *done, hipStream_t s,
                         const int32_t *gptr)
{
    if (nnodes <= 0) return;
    const int nb = grid_rows(nnodes, kGsNodes);
    hipLaunchKernelGGL(k_gs_color<float>, dim3(nb), dim3(kGsBlock), 0, s, nnodes, nodes, gptr, rp, ci, v, b, x,
                       backward ? 1 : 0, done);
    MGX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ dense coarse solve: y (+)= M x
__global__ __launch_bounds__(kBlock) void k_dense_gemv(int32_t n, const double *__restrict__ M,
                                                       const double *__restrict__ x, double *__restrict__ y,
                                                       int accumulate, const int32_t *__restrict__ done)
{
    if (done && *done) return;
    const int row = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double *Mr = M + (int64_t)row * n;
    double s = 0;
    for (int j = lane; j < n; j += 64) s += Mr[j] * x[j];
    s = wave_sum(s);
    if (lane == 0) y[row] = accumulate ? y[row] + s : s;
}

void launch_dense_gemv(int32_t n, const double *M, const double *x, double *y, bool acc, const int32_t *done,
                       hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_dense_gemv, dim3((n + 3) / 4), dim3(kBlock), 0, s, n, M, x, y, acc ? 1 : 0, done);
    MGX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ SIMP adjoint
__constant__ double c_Ke0[576];

__global__ __launch_bounds__(kBlock) void k_simp_gradient(int nx, int ny, int nz, const double *__restrict__ rho,
                                                          double penal, double Emin, double E0,
                                                          const double *__restrict__ u, double *__restrict__ grad)
{
    const int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    const int64_t ne = (int64_t)nx * ny * nz;
    if (e >= ne) return;
    const int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((int64_t)nx * ny));
    const int ox[8] = {0, 1, 1, 0, 0, 1, 1, 0}, oy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, oz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    double ue[24];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        const int jx = ex + ox[l], jy = ey + oy[l], jz = ez + oz[l];
        const int64_t fn = ((int64_t)jz * (ny + 1) + jy) * nx + (jx - 1);
#pragma unroll
        for (int b = 0; b < 3; ++b) ue[3 * l + b] = jx == 0 ? 0.0 : u[3 * fn + b];
    }
    double ce = 0;
    for (int i = 0; i < 24; ++i) {
        double s = 0;
#pragma unroll
        for (int j = 0; j < 24; ++j) s += c_Ke0[i * 24 + j] * ue[j];
        ce += ue[i] * s;
    }
    grad[e] = -penal * pow(rho[e], penal - 1) * (E0 - Emin) * ce;
}

void launch_simp_gradient(int32_t nx, int32_t ny, int32_t nz, const double *Ke0, const double *rho, double penal,
                          double Emin, double E0, const double *u, double *grad, hipStream_t s)
{
    MGX_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_Ke0), Ke0, sizeof(double) * 576, 0, hipMemcpyHostToDevice, s));
    const int64_t ne = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_simp_gradient, dim3((ne + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nx, ny, nz, rho, penal,
                       Emin, E0, u, grad);
    MGX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ structured assembly
// Row (node (ix,iy,iz) with ix >= 1, dof a) of K_ff in sorted CSR; value of column (node j,
// dof b) = sum over the elements holding both nodes, in increasing element index, of
// rho_e * Ke[3 li + a][3 lj + b] (or of Kel_e[...]); sums start at 0.0 and add each
// contribution as `c + acc`, as DokRowMajor.AddSubmatrix does (FemCantileverBase.cs:95-101).
__constant__ double c_Ke[576];

__device__ __forceinline__ int loc_index(int ox, int oy, int oz)
{
    // (0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1)
    return oz * 4 + (oy ? (ox ? 2 : 3) : (ox ? 1 : 0));
}

__device__ __forceinline__ int axis_count(int i, int lo, int hi)
{
    int c = 0;
    for (int d = -1; d <= 1; ++d) c += (i + d >= lo && i + d <= hi);
    return c;
}

__global__ __launch_bounds__(kBlock) void k_assemble_count(int nx, int ny, int nz, int32_t *rowlen)
{
    const int64_t row = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    const int64_t nrows = 3LL * nx * (ny + 1) * (nz + 1);
    if (row >= nrows) return;
    const int64_t fn = row / 3;
    const int ix = (int)(fn % nx) + 1;
    const int iy = (int)((fn / nx) % (ny + 1));
    const int iz = (int)(fn / ((int64_t)nx * (ny + 1)));
    rowlen[row] = 3 * axis_count(ix, 1, nx) * axis_count(iy, 0, ny) * axis_count(iz, 0, nz);
}

__global__ __launch_bounds__(kBlock) void k_assemble_fill(int nx, int ny, int nz, const double *__restrict__ dens,
                                                          const double *__restrict__ Kel,
                                                          const int32_t *__restrict__ rowptr, int32_t *__restrict__ col,
                                                          double *__restrict__ val)
{
#pragma clang fp contract(off)
    const int64_t row = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    const int64_t nrows = 3LL * nx * (ny + 1) * (nz + 1);
    if (row >= nrows) return;
    const int a = (int)(row % 3);
    const int64_t fn = row / 3;
    const int ix = (int)(fn % nx) + 1;
    const int iy = (int)((fn / nx) % (ny + 1));
    const int iz = (int)(fn / ((int64_t)nx * (ny + 1)));
    int64_t pos = rowptr[row];
    for (int dz = -1; dz <= 1; ++dz) {
        const int jz = iz + dz;
        if (jz < 0 || jz > nz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int jy = iy + dy;
            if (jy < 0 || jy > ny) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int jx = ix + dx;
                if (jx < 1 || jx > nx) continue;
                const int64_t fj = ((int64_t)jz * (ny + 1) + jy) * nx + (jx - 1);
                int ex0, ey0, ez0, cx, cy, cz;
                if (dx == 0) { ex0 = ix - 1 < 0 ? 0 : ix - 1; cx = (ix > nx - 1 ? nx - 1 : ix) - ex0 + 1; } else { ex0 = dx > 0 ? ix : ix - 1; cx = 1; }
                if (dy == 0) { ey0 = iy - 1 < 0 ? 0 : iy - 1; cy = (iy > ny - 1 ? ny - 1 : iy) - ey0 + 1; } else { ey0 = dy > 0 ? iy : iy - 1; cy = 1; }
                if (dz == 0) { ez0 = iz - 1 < 0 ? 0 : iz - 1; cz = (iz > nz - 1 ? nz - 1 : iz) - ez0 + 1; } else { ez0 = dz > 0 ? iz : iz - 1; cz = 1; }
                for (int b = 0; b < 3; ++b) {
                    double acc = 0.0;
                    for (int ez = ez0; ez < ez0 + cz; ++ez)
                        for (int ey = ey0; ey < ey0 + cy; ++ey)
                            for (int ex = ex0; ex < ex0 + cx; ++ex) {
                                const int64_t e = ex + (int64_t)nx * (ey + (int64_t)ny * ez);
                                const int li = loc_index(ix - ex, iy - ey, iz - ez);
                                const int lj = loc_index(jx - ex, jy - ey, jz - ez);
                                const int kk = (3 * li + a) * 24 + 3 * lj + b;
                                double c;
                                if (Kel) c = Kel[e * 576 + kk];
                                else c = (dens ? dens[e] : 1.0) * c_Ke[kk];
                                acc = c + acc;
                            }
                    col[pos] = (int32_t)(3 * fj + b);
                    val[pos] = acc;
                    ++pos;
                }
            }
        }
    }
}

void launch_assemble_count(int nx, int ny, int nz, int32_t *rowlen, hipStream_t s)
{
    const int64_t nrows = 3LL * nx * (ny + 1) * (nz + 1);
    hipLaunchKernelGGL(k_assemble_count, dim3((nrows + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nx, ny, nz, rowlen);
    MGX_HIP(hipGetLastError());
}

void launch_assemble_fill(int nx, int ny, int nz, const double *Ke, const double *dens, const double *Kel,
                          const int32_t *rowptr, int32_t *col, double *val, hipStream_t s)
{
    if (Ke) MGX_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_Ke), Ke, sizeof(double) * 576, 0, hipMemcpyHostToDevice, s));
    const int64_t nrows = 3LL * nx * (ny + 1) * (nz + 1);
    hipLaunchKernelGGL(k_assemble_fill, dim3((nrows + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nx, ny, nz, dens, Kel,
                       rowptr, col, val);
    MGX_HIP(hipGetLastError());
}

// exclusive scan of int32 (single-block chunked scan: chunk sums, then offsets)
__global__ __launch_bounds__(1024) void k_scan_chunks(const int32_t *__restrict__ in, int32_t *__restrict__ out,
                                                      int32_t n, int32_t chunk, int64_t *__restrict__ sums)
{
    // each thread scans one chunk sequentially
    const int t = blockIdx.x * 1024 + threadIdx.x;
    const int64_t s = (int64_t)t * chunk;
    if (s >= n) return;
    int64_t e = s + chunk; if (e > n) e = n;
    int64_t acc = 0;
    for (int64_t i = s; i < e; ++i) { out[i] = (int32_t)acc; acc += in[i]; }
    sums[t] = acc;
}
__global__ void k_scan_sums(int64_t *sums, int nchunks, int32_t *total_out, int32_t n)
{
    if (threadIdx.x || blockIdx.x) return;
    int64_t acc = 0;
    for (int i = 0; i < nchunks; ++i) { int64_t v = sums[i]; sums[i] = acc; acc += v; }
    total_out[n] = (int32_t)acc;
}
__global__ __launch_bounds__(kBlock) void k_scan_add(int32_t *out, int32_t n, int32_t chunk, const int64_t *sums)
{
    const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (i >= n) return;
    out[i] = (int32_t)(out[i] + sums[i / chunk]);
}

void exclusive_scan_i32(const int32_t *in, int32_t *out, int32_t n, hipStream_t s)
{
    // out has n+1 entries; out[n] = total
    const int32_t chunk = 256;
    const int nchunks = (n + chunk - 1) / chunk;
    int64_t *sums = nullptr;
    MGX_HIP(hipMalloc((void **)&sums, sizeof(int64_t) * (nchunks > 0 ? nchunks : 1)));
    if (n > 0) {
        hipLaunchKernelGGL(k_scan_chunks, dim3((nchunks + 1023) / 1024), dim3(1024), 0, s, in, out, n, chunk, sums);
        MGX_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(64), 0, s, sums, nchunks, out, n);
    MGX_HIP(hipGetLastError());
    if (n > 0) {
        hipLaunchKernelGGL(k_scan_add, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, out, n, chunk, sums);
        MGX_HIP(hipGetLastError());
    }
    MGX_HIP(hipStreamSynchronize(s));
    MGX_HIP(hipFree(sums));
}

// ------------------------------------------------------------------ Galerkin element coarsening
// Coarse element E (ncx x ncy x ncz) holds children c = (cx,cy,cz) in {0,1}^3, fine element
// (2X+cx, 2Y+cy, 2Z+cz).  Fine corner l of child c sits at (c + off(l)) / 2 in the coarse
// element; its trilinear weight from coarse corner L is t_c[l][L] = prod_axis (1 - |p - L|).
// Kc = sum_c (t_c x I3)^T K_c (t_c x I3), K_c = rho_c Ke0 (level 0) or Kfine[c].
__global__ __launch_bounds__(kBlock) void k_galerkin_elements(int ncx, int ncy, int ncz,
                                                              const double *__restrict__ dens,
                                                              const double *__restrict__ Kfine,
                                                              double *__restrict__ Kcoarse)
{
    __shared__ double Kc[576];
    __shared__ double W[576];
    __shared__ double acc[576];
    const int64_t E = blockIdx.x;
    const int X = (int)(E % ncx), Y = (int)((E / ncx) % ncy), Z = (int)(E / ((int64_t)ncx * ncy));
    const int nfx = 2 * ncx, nfy = 2 * ncy;
    const int ox[8] = {0, 1, 1, 0, 0, 1, 1, 0}, oy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, oz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (int i = threadIdx.x; i < 576; i += kBlock) acc[i] = 0;
    for (int c = 0; c < 8; ++c) {
        const int fx = 2 * X + ox[c], fy = 2 * Y + oy[c], fz = 2 * Z + oz[c];
        const int64_t fe = fx + (int64_t)nfx * (fy + (int64_t)nfy * fz);
        __syncthreads();
        for (int i = threadIdx.x; i < 576; i += kBlock)
            Kc[i] = Kfine ? Kfine[fe * 576 + i] : dens[fe] * c_Ke[i];
        __syncthreads();
        // W[3l+a][3M+b] = sum_m K[3l+a][3m+b] t[m][M]
        for (int i = threadIdx.x; i < 576; i += kBlock) {
            const int r = i / 24, q = i % 24, M = q / 3, bb = q % 3;
            double s = 0;
            for (int m = 0; m < 8; ++m) {
                const double px = 0.5 * (ox[c] + ox[m]), py = 0.5 * (oy[c] + oy[m]), pz = 0.5 * (oz[c] + oz[m]);
                const double t = (1 - fabs(px - ox[M])) * (1 - fabs(py - oy[M])) * (1 - fabs(pz - oz[M]));
                if (t != 0) s += Kc[r * 24 + 3 * m + bb] * t;
            }
            W[i] = s;
        }
        __syncthreads();
        // acc[3L+a][3M+b] += sum_l t[l][L] W[3l+a][3M+b]
        for (int i = threadIdx.x; i < 576; i += kBlock) {
            const int r = i / 24, q = i % 24, L = r / 3, aa = r % 3;
            double s = 0;
            for (int l = 0; l < 8; ++l) {
                const double px = 0.5 * (ox[c] + ox[l]), py = 0.5 * (oy[c] + oy[l]), pz = 0.5 * (oz[c] + oz[l]);
                const double t = (1 - fabs(px - ox[L])) * (1 - fabs(py - oy[L])) * (1 - fabs(pz - oz[L]));
                if (t != 0) s += t * W[(3 * l + aa) * 24 + q];
            }
            acc[i] += s;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 576; i += kBlock) Kcoarse[E * 576 + i] = acc[i];
}

void launch_galerkin_elements(int ncx, int ncy, int ncz, const double *Ke0, const double *dens,
                              const double *Kfine, double *Kcoarse, hipStream_t s)
{
    if (Ke0) MGX_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_Ke), Ke0, sizeof(double) * 576, 0, hipMemcpyHostToDevice, s));
    const int64_t ne = (int64_t)ncx * ncy * ncz;
    hipLaunchKernelGGL(k_galerkin_elements, dim3(ne), dim3(kBlock), 0, s, ncx, ncy, ncz, dens, Kfine, Kcoarse);
    MGX_HIP(hipGetLastError());
}

}  // namespace mgx
