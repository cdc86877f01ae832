// mgx_device.h -- device-side helpers shared by the kernel files of libmgx.so.
#pragma once
#include "mgx_internal.h"

namespace mgx {

__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    // nb is a multiple of 8 for every launch that uses this (see grid_rows)
    const int per = nb >> 3;
    return (b & 7) * per + (b >> 3);
}

static inline int grid_rows(int64_t nrows, int rows_per_block)
{
    int64_t nb = (nrows + rows_per_block - 1) / rows_per_block;
    if (nb < 1) nb = 1;
    nb = (nb + 7) & ~int64_t(7);
    return (int)nb;
}

template <typename T>
__device__ __forceinline__ T ntload(const T *p) { return __builtin_nontemporal_load(p); }

__device__ __forceinline__ double nanmax(double a, double b)
{
    return (a != a) ? a : ((b != b) ? b : (a > b ? a : b));
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Block reduction of (sum a, sum b, max m) for blocks of up to 1024 threads; totals valid in
// thread 0.  Fixed tree order: deterministic.
__device__ __forceinline__ void block_reduce3(double &a, double &b, double &m)
{
    __shared__ double sh[3][16];
    a = wave_sum(a); b = wave_sum(b); m = wave_max(m);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) { sh[0][w] = a; sh[1][w] = b; sh[2][w] = m; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        double sa = 0, sb = 0, sm = 0;
        for (int i = 0; i < nw; ++i) { sa += sh[0][i]; sb += sh[1][i]; sm = nanmax(sm, sh[2][i]); }
        a = sa; b = sb; m = sm;
    }
}

static inline int grid_stream(int64_t n)
{
    int64_t nb = (n + kBlock * 4 - 1) / (kBlock * 4);
    if (nb < 1) nb = 1;
    if (nb > 2048) nb = 2048;
    return (int)nb;
}

// ------------------------------------------------------------------ batching-form shapes and helpers
template <int K> struct BShape {
    static constexpr int S = K >= 16 ? 1 : 16 / K;   // lanes sharing one row for one column
    static constexpr int G = K * S;                  // lanes per row
    static constexpr int RPB = kBlock / G;           // rows per tile
    static constexpr int CAP = RPB * 96;             // staged entries per tile
    static constexpr int U = S == 1 ? 8 : 4;         // entries in flight per lane
};

static inline int grid_tiles(int64_t ntiles)
{
    int64_t nb = std::min<int64_t>(std::max<int64_t>(ntiles, 1), 2048);
    nb = (nb + 7) & ~int64_t(7);
    return (int)nb;
}

// Per-column block reduction.  Thread t holds values of column t % K (blockDim = 256); the
// K column totals of slot j go to partials[(j*K + c)*stride + blockIdx.x].  Slot MAXSLOT is a
// NaN-propagating max, the others sums.  Fixed order: deterministic.
template <int K, int NS, int MAXSLOT>
__device__ __forceinline__ void publish_cols(double v0, double v1, double v2, double *__restrict__ partials, int stride)
{
    __shared__ double red[NS][kBlock];
    const int t = threadIdx.x;
    __syncthreads();
    red[0][t] = v0;
    if constexpr (NS > 1) red[1][t] = v1;
    if constexpr (NS > 2) red[2][t] = v2;
    __syncthreads();
    if (t < K * NS) {
        const int slot = t / K, c = t % K;
        double a = 0;
        for (int j = c; j < kBlock; j += K) {
            const double x = red[slot][j];
            a = (slot == MAXSLOT) ? nanmax(a, x) : a + x;
        }
        partials[(size_t)(slot * K + c) * stride + blockIdx.x] = a;
    }
}

// One lane's share of a row of a block SpMM, entries [s, e) of cis / vs, for column c: entries
// s + sl, s + sl + S, ...
template <int K, typename TV, typename TX, typename PI, typename PV>
__device__ __forceinline__ TV row_cols(PI cis, PV vs, int s, int e, const TX *__restrict__ X, int c, int sl)
{
    using Sh = BShape<K>;
    TV a0 = 0, a1 = 0;
    for (int k = s + sl; k < e; k += Sh::U * Sh::S) {
        int cc[Sh::U];
        TV w[Sh::U];
#pragma unroll
        for (int u = 0; u < Sh::U; ++u) {
            const int kk = k + u * Sh::S;
            if (kk < e) { cc[u] = cis[kk]; w[u] = vs[kk]; } else { cc[u] = -1; w[u] = 0; }
        }
        TX xv[Sh::U];
#pragma unroll
        for (int u = 0; u < Sh::U; ++u) xv[u] = cc[u] >= 0 ? X[(size_t)cc[u] * K + c] : TX(0);
#pragma unroll
        for (int u = 0; u < Sh::U; ++u) { if (u & 1) a1 += w[u] * (TV)xv[u]; else a0 += w[u] * (TV)xv[u]; }
    }
    return a0 + a1;
}

}  // namespace mgx
