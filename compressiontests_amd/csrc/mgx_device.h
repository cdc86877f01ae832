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

}  // namespace mgx
