// mgx_duvi.hip -- SpMV over the DuVi GPU block layout (OpenCLDuViGeometricMultigridSolver.cs:16-31).
//
// Decode rule of DuViGeometricMultigrid.cl:57-99: a row's blocks are consecutive in
// block_col; a block's entries are consecutive slots of dist/vi starting at the row's offset
// row_dist[r]; at a block's first slot the distance byte is the count of the block's further
// entries and the column is block_col[b]; at a later slot the column is the previous one plus
// the distance byte plus 1.  The value of slot d is values[vi[d]] (u8/u16/u32 indices) or
// values[d] ("DU + raw values").
//
// A wavefront owns a tile of 8 consecutive rows (8 lanes per row) and needs no block-wide
// barrier: (1) every lane issues the loads of the raw values of its first slots (raw mode); the
// tile's distance bytes and, with value indices, its index stream are one contiguous range each
// and go to the wave's LDS region with coalesced 16-byte loads; (2) each row's lanes walk the
// row's block chain together (one step per block: the count byte) and mark each block's first
// slot with its column; (3) in rounds of 96 slots per row, lane l takes the 12 slots 12l ..
// 12l+11 of the round: it folds its slots into (a block starts among them, column or column
// offset), a segmented scan over the row's 8 lanes gives every lane the column before its first
// slot, and the lane decodes its slots again, gathering x and accumulating the products; (4) the
// 8 lanes sum the row.  A tile of more than kDwCap slots falls back to one lane per row decoding
// straight from global memory.
//
// The row result goes through one of the epilogues of the reference's DuVi kernels, fused into
// the same pass (no intermediate A x vector): y = A x, y += A x (matrix_vector_product,
// DuViGeometricMultigrid.cl:122-159), r = b - A x (residual, :220-254), the same with the
// convergence flag (residual_with_check, :275-317), y = x + pre (b - A x) (jacobi_iteration,
// :362-402).  Values and vectors are FP64, or FP32 for the precision form (an FP32 copy of the
// value table; indices unchanged).
//
// k_duvi_gs is one colour of a multicolour Gauss-Seidel sweep on the same layout (the
// reference's gauss_seidel_iteration, :422-462, updates x in place with racy reads; the
// multicolour order is the documented deviation).
#include <type_traits>

#include "mgx_device.h"

namespace mgx {

typedef unsigned int dw_u32x4 __attribute__((ext_vector_type(4)));

template <int VIB, typename T>
__device__ __forceinline__ T duvi_value(const void *vi, const T *values, int64_t d)
{
    if (VIB == 1) return values[((const uint8_t *)vi)[d]];
    if (VIB == 2) return values[((const uint16_t *)vi)[d]];
    if (VIB == 4) return values[((const uint32_t *)vi)[d]];
    return values[d];
}

constexpr int kDwRows = 8;                     // rows per wave tile
constexpr int kDwLanes = 8;                    // lanes per row
constexpr int kDwSeg = 12;                     // slots per lane per round (96 per row and round)
constexpr int kDwCap = 768;                    // slots staged per wave tile
constexpr int kDwDistNI = (kDwCap + 32 + 1023) / 1024;   // 64-slot rounds of 16 bytes for the distance bytes
constexpr int kDwWaves = kBlock / 64;

__device__ __forceinline__ void dw_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);        // lgkmcnt(0): this wave's LDS writes are done
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Copy bytes [lo, hi) of src (tot bytes long) into dst[0 ..) shifted by lo & 15.  The 16-byte
// slots wholly inside src are loaded as vectors, NI rounds of 64 slots issued before the first
// store (slots past the range reload the last whole slot into dst slots < 64 NI that the tile
// does not use); slots reaching past the end of src are copied byte by byte.
template <int NI>
__device__ __forceinline__ void dw_stage(const char *__restrict__ src, int64_t lo, int64_t hi, int64_t tot,
                                         dw_u32x4 *dst, int lane)
{
    const int64_t a0 = lo & ~(int64_t)15;
    const int nq = (int)((hi - a0 + 15) >> 4);
    const int nin = (int)min<int64_t>((int64_t)nq, (tot - a0) >> 4);
    if (nin > 0) {
        const dw_u32x4 *s = (const dw_u32x4 *)(src + a0);
        dw_u32x4 t[NI];
#pragma unroll
        for (int i = 0; i < NI; ++i) t[i] = __builtin_nontemporal_load(s + min(lane + 64 * i, nin - 1));
#pragma unroll
        for (int i = 0; i < NI; ++i) dst[lane + 64 * i] = t[i];
    }
    for (int q = max(nin, 0) + lane; q < nq; q += 64) {
        uint8_t *o = (uint8_t *)&dst[q];
        for (int i = 0; i < 16; ++i) o[i] = a0 + 16 * q + i < tot ? (uint8_t)src[a0 + 16 * q + i] : 0;
    }
}

// staged index stream: 16-byte rounds of 64 slots per wave (none in raw mode)
template <int VIB> struct DwIdx { static constexpr int NI = VIB == 0 ? 1 : (kDwCap * VIB / 16 + 2 + 63) / 64; };

template <int VIB, typename T>
__device__ __forceinline__ T dw_value(const uint8_t *si, const T *__restrict__ values, int64_t d0, int k)
{
    if (VIB == 0) return values[d0 + k];
    if (VIB == 1) return values[si[k]];
    if (VIB == 2) return values[((const uint16_t *)si)[k]];
    return values[((const uint32_t *)si)[k]];
}

template <int VIB> __device__ __forceinline__ uint32_t dw_index(const uint8_t *si, int64_t k)
{
    if (VIB == 1) return __builtin_nontemporal_load(si + k);
    if (VIB == 2) return __builtin_nontemporal_load((const uint16_t *)si + k);
    return __builtin_nontemporal_load((const uint32_t *)si + k);
}

// operands of the fused epilogues
template <typename T> struct DuviOp {
    const T *values;          // value table (or the raw values)
    const T *x, *b, *pre;     // b: RESID, JACOBI, CHECK; pre: JACOBI
    T *y;
    int32_t *flag;            // CHECK
    double tol;               // CHECK
};

// result of one row through the epilogue; the CHECK form also returns the flag bits of the row
// (2: NaN or |r| > 1e50, 1: |r| > tol)
template <int EPI, typename T>
__device__ __forceinline__ int duvi_store(const DuviOp<T> &op, int row, T acc)
{
    if (EPI == DEPI_SPMV) op.y[row] = acc;
    else if (EPI == DEPI_ADD) op.y[row] = op.y[row] + acc;
    else if (EPI == DEPI_RESID) op.y[row] = op.b[row] - acc;
    else if (EPI == DEPI_JACOBI) op.y[row] = op.x[row] + op.pre[row] * (op.b[row] - acc);
    else {
        const T ri = op.b[row] - acc;
        op.y[row] = ri;
        const double c = fabs((double)ri);
        if (ri != ri || c > 1e50) return 2;
        if (c > op.tol) return 1;
    }
    return 0;
}

// the flag protocol of residual_with_check (bit 1 set on NaN / overflow, bit 0 cleared when a row
// exceeds tol), one atomic per wavefront and kind; the final flag does not depend on the order
template <int EPI, typename T>
__device__ __forceinline__ void duvi_flag(const DuviOp<T> &op, int bits)
{
    if (EPI != DEPI_CHECK) return;
    const bool bad = __any(bits & 2), over = __any(bits & 1);
    if ((threadIdx.x & 63) == 0) {
        if (bad) atomicOr(op.flag, 2);
        if (over) atomicAnd(op.flag, 2);
    }
}

template <int VIB, typename T, int EPI>
__global__ __launch_bounds__(kBlock) void k_duvi_wave(int32_t nrows, int64_t nnz, const int32_t *__restrict__ row_blocks,
                                                      const int32_t *__restrict__ row_dist,
                                                      const int32_t *__restrict__ block_col,
                                                      const uint8_t *__restrict__ dist, const void *__restrict__ vi,
                                                      const DuviOp<T> op, const int32_t *__restrict__ done)
{
    __shared__ dw_u32x4 s_iq[kDwWaves][64 * DwIdx<VIB>::NI];  // value indices (not used in raw mode)
    __shared__ dw_u32x4 s_dq[kDwWaves][64 * kDwDistNI];       // distance bytes
    __shared__ int32_t s_h[kDwWaves][kDwCap];                 // column + 1 at a block's first slot, else 0
    static_assert(64 * 16 * kDwDistNI >= kDwCap + 32, "distance staging");
    if (done && *done) return;
    const T *__restrict__ values = op.values;
    const T *__restrict__ x = op.x;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane / kDwLanes, l = lane % kDwLanes;
    const int ntiles = (nrows + kDwRows - 1) / kDwRows;
    const int tile = xcd_remap(blockIdx.x, gridDim.x) * kDwWaves + wv;
    if (tile >= ntiles) return;                                // the whole wave leaves together
    const int r0 = tile * kDwRows, r1 = min(r0 + kDwRows, nrows);
    const int row = r0 + r;
    // every offset the tile needs, loaded together (one global round trip)
    const int64_t d0 = row_dist[r0];
    const int64_t dr1 = row_dist[r1];
    int64_t ds = d0, de = d0;                                  // an empty range past the last row
    int rb0 = 0, rb1 = 0;
    if (row < r1) { ds = row_dist[row]; de = row_dist[row + 1]; rb0 = row_blocks[row]; rb1 = row_blocks[row + 1]; }
    const int cnt = (int)(dr1 - d0);
    if (cnt > kDwCap) {
        // long tile: one lane per row, decoding straight from global memory
        int bits = 0;
        if (l == 0 && row < r1) {
            T acc = 0;
            int64_t d = row_dist[row];
            for (int b = row_blocks[row]; b < row_blocks[row + 1]; ++b) {
                int col = block_col[b];
                const int64_t dend = d + dist[d] + 1;
                for (;;) {
                    acc += duvi_value<VIB, T>(vi, values, d) * x[col];
                    if (++d == dend) break;
                    col += dist[d] + 1;
                }
            }
            bits = duvi_store<EPI, T>(op, row, acc);
        }
        duvi_flag<EPI, T>(op, bits);
        return;
    }
    const int s = (int)(ds - d0), e = (int)(de - d0);
    // the first two rounds of block columns (up to 2 * kDwLanes blocks per row), loaded before
    // the staging so that the chain walk below starts without a global round trip
    const int bcA = rb0 + l < rb1 ? block_col[rb0 + l] : 0;
    const int bcB = rb0 + kDwLanes + l < rb1 ? block_col[rb0 + kDwLanes + l] : 0;
    // (1) raw values of this lane's slots in round 0 (they do not depend on the decode); the
    //     index stream and the distance bytes to LDS; cleared start marks
    T w0[kDwSeg];
    if (VIB == 0) {
        const int a = s + l * kDwSeg;
#pragma unroll
        for (int j = 0; j < kDwSeg; ++j) w0[j] = a + j < e ? values[d0 + a + j] : T(0);
    }
    if (VIB != 0) dw_stage<DwIdx<VIB>::NI>((const char *)vi, d0 * VIB, (d0 + cnt) * VIB, nnz * VIB, s_iq[wv], lane);
    dw_stage<kDwDistNI>((const char *)dist, d0, d0 + cnt, nnz, s_dq[wv], lane);
    for (int k = lane; k < cnt; k += 64) s_h[wv][k] = 0;
    const uint8_t *sd = (const uint8_t *)s_dq[wv] + (d0 & 15);               // sd[k]: distance byte of slot d0 + k
    const uint8_t *si = (const uint8_t *)s_iq[wv] + ((d0 * (VIB ? VIB : 1)) & 15);
    dw_fence();
    // (2) the block chain of each row, walked by its 8 lanes together (the same d in every
    //     lane); the block columns come 8 at a time, one per lane
    {
        int d = s;
        for (int bw = rb0, it = 0; bw < rb1; bw += kDwLanes, ++it) {
            const int bc = it == 0 ? bcA : it == 1 ? bcB : (bw + l < rb1 ? block_col[bw + l] : 0);
            const int nb = min(kDwLanes, rb1 - bw);
            for (int j = 0; j < nb; ++j) {
                const int c = __shfl(bc, j, kDwLanes);
                if (l == 0 && d < e) s_h[wv][d] = c + 1;
                d += sd[d] + 1;
            }
        }
    }
    dw_fence();
    // (3) rounds of 96 slots per row
    T acc = 0;
    int carry = 0;                                             // column before the round's first slot
    for (int base = s, round = 0; base < e; base += kDwLanes * kDwSeg, ++round) {
        const int a = base + l * kDwSeg, bnd = min(a + kDwSeg, e);
        // fold this lane's slots: h = a block starts among them; v = column of the last slot
        // (h) or its offset from the column before the first slot (!h)
        int h = 0, v = 0;
#pragma unroll
        for (int j = 0; j < kDwSeg; ++j) {
            const int k = a + j;
            if (k < bnd) {
                const int hc = s_h[wv][k];
                if (hc) { h = 1; v = hc - 1; } else v += sd[k] + 1;
            }
        }
        // inclusive segmented scan over the row's 8 lanes
#pragma unroll
        for (int o = 1; o < kDwLanes; o <<= 1) {
            const int ph = __shfl_up(h, o, kDwLanes), pv = __shfl_up(v, o, kDwLanes);
            if (l >= o && !h) { v += pv; h = ph; }
        }
        // column before this lane's first slot
        const int ph = __shfl_up(h, 1, kDwLanes), pv = __shfl_up(v, 1, kDwLanes);
        int col = (l == 0) ? carry : (ph ? pv : carry + pv);
        // the round's last column, for the next round
        const int th = __shfl(h, kDwLanes - 1, kDwLanes), tv = __shfl(v, kDwLanes - 1, kDwLanes);
        carry = th ? tv : carry + tv;
        // decode again, gathering x
        T xv[kDwSeg], w[kDwSeg];
#pragma unroll
        for (int j = 0; j < kDwSeg; ++j) {
            const int k = a + j;
            int c = 0;
            if (k < bnd) {
                const int hc = s_h[wv][k];
                col = hc ? hc - 1 : col + sd[k] + 1;
                c = col;
            }
            xv[j] = k < bnd ? x[c] : T(0);
            if (VIB == 0 && round == 0) w[j] = w0[j];
            else w[j] = k < bnd ? dw_value<VIB, T>(si, values, d0, k) : T(0);
        }
#pragma unroll
        for (int j = 0; j < kDwSeg; ++j) acc += w[j] * xv[j];
    }
    // (4) row sum over the 8 lanes
#pragma unroll
    for (int o = kDwLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kDwLanes);
    int bits = 0;
    if (l == 0 && row < r1) bits = duvi_store<EPI, T>(op, row, acc);
    duvi_flag<EPI, T>(op, bits);
}

// ------------------------------------------------------------------ stream form (long rows)
// For matrices with long rows (the stiffness levels: 9 blocks of up to 9 slots per row) a 256-thread
// block takes a tile of RB = 32 consecutive rows, whose slots [d0, d1) and blocks [b0, b1) are
// contiguous, the layout the CSR-stream kernel reads (64-row tiles take 43 KB of LDS, 3 blocks per
// CU, and ran 1.2-1.8x slower at 128^3; 16-row tiles ran 1.4-1.5x slower):
//   (1) each thread loads its share of the tile's row offsets, distance bytes (16-byte loads) and
//       block columns, a fixed number of loads each; then the E slots t, t + 256, ... of the value
//       stream: raw values, or value indices, whose table entries are loaded in step (3) beside x.
//       A one-byte index table (256 values at most) is copied to LDS: looked up in global memory,
//       every wave of the chip gathers from the same two or three cache lines;
//   (2) the staged offsets, distance bytes and block columns go to LDS; the G = 256 / RB lanes of a
//       row walk its block chain together (the count byte of each block) and lane g decodes the
//       blocks g, g + G, ...: column of the block's first slot from block_col, then + distance + 1
//       per slot, into an int32 column per slot;
//   (3) every thread reads the columns of its E slots, gathers x and the table values of its value
//       indices, and stores the products;
//   (4) the G lanes of a row sum its products in the CSR-stream kernel's order.
// The column, distance and block-column buffers alias the product buffer.  A tile with more
// slots than the buffer or more blocks than fit (not the case for the stiffness matrices) is
// decoded one row per thread from global memory.
template <typename T> struct DsShape {
    static constexpr int RB = 32;                                         // rows per tile
    static constexpr int G = kBlock / RB;                                 // lanes per row
    static constexpr int E = (RB * 84 + kBlock - 1) / kBlock;            // slots per thread
    static constexpr int C = E * kBlock;                                  // slots per tile
    static constexpr int NQ = (C + 32 + 15) / 16;                         // 16-byte rounds of distance bytes
    static constexpr int NQT = (NQ + kBlock - 1) / kBlock;                // ... per thread
    static constexpr int DIST_OFF = 4 * C;                                // after the int32 columns
    static constexpr int BC_OFF = DIST_OFF + 16 * NQ;                     // block columns
    static constexpr int NBT = 2;                                         // block-column loads per thread
    static constexpr int BCAP = NBT * kBlock;                             // block columns per tile
    static constexpr int BYTES_DECODE = BC_OFF + 4 * BCAP;
    static constexpr int BYTES = BYTES_DECODE > (int)sizeof(T) * C ? BYTES_DECODE : (int)sizeof(T) * C;
};

template <int VIB, typename T, int EPI>
__global__ __launch_bounds__(kBlock) void k_duvi_stream(int32_t nrows, int64_t nnz, int64_t nblocks, int32_t nvals,
                                                        const int32_t *__restrict__ row_blocks,
                                                        const int32_t *__restrict__ row_dist,
                                                        const int32_t *__restrict__ block_col,
                                                        const uint8_t *__restrict__ dist, const void *__restrict__ vi,
                                                        const DuviOp<T> op, const int32_t *__restrict__ done)
{
    using Sh = DsShape<T>;
    constexpr int RB = Sh::RB, G = Sh::G, E = Sh::E, C = Sh::C;
    __shared__ __align__(16) unsigned char buf[Sh::BYTES];
    __shared__ int32_t s_rd[RB + 1], s_rb[RB + 1];
    __shared__ T s_tab[VIB == 1 ? 256 : 1];           // one-byte indices: the whole value table
    if (done && *done) return;
    const T *__restrict__ values = op.values;
    const T *__restrict__ x = op.x;
    const int t = threadIdx.x, lane = t % G, r = t / G;
    const int ntiles = (nrows + RB - 1) / RB;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;                                    // the whole block leaves together
    const int r0 = tile * RB, r1 = min(r0 + RB, nrows), nr = r1 - r0;
    const int64_t d0 = row_dist[r0], d1 = row_dist[r1];
    const int b0 = row_blocks[r0], b1 = row_blocks[r1];
    const int cnt = (int)(d1 - d0), nb = b1 - b0;
    const int row = r0 + r;
    if (cnt > C || nb > Sh::BCAP) {
        // long or fragmented tile: one thread per row, decoding straight from global memory
        int bits = 0;
        if (lane == 0 && row < r1) {
            T acc = 0;
            int64_t d = row_dist[row];
            for (int bb = row_blocks[row]; bb < row_blocks[row + 1]; ++bb) {
                int col = block_col[bb];
                const int64_t dend = d + dist[d] + 1;
                for (;;) {
                    acc += duvi_value<VIB, T>(vi, values, d) * x[col];
                    if (++d == dend) break;
                    col += dist[d] + 1;
                }
            }
            bits = duvi_store<EPI, T>(op, row, acc);
        }
        duvi_flag<EPI, T>(op, bits);
        return;
    }
    // (1) staged inputs of the decode: a fixed number of loads per thread, then the value stream
    int off = 0;
    if (t <= nr) off = row_dist[r0 + t];
    else if (t >= 128 && t - 128 <= nr) off = row_blocks[r0 + t - 128];
    const int64_t a0 = d0 & ~(int64_t)15;
    const int nq = (int)((d1 - a0 + 15) >> 4);
    const int nin = (int)min<int64_t>((int64_t)nq, (nnz - a0) >> 4);   // 16-byte rounds wholly inside dist
    dw_u32x4 q[Sh::NQT];
    if (nin > 0) {                                   // rounds past nin reload round nin - 1 (unused)
#pragma unroll
        for (int i = 0; i < Sh::NQT; ++i)
            q[i] = __builtin_nontemporal_load((const dw_u32x4 *)(dist + a0) + min(t + kBlock * i, nin - 1));
    } else {
#pragma unroll
        for (int i = 0; i < Sh::NQT; ++i) q[i] = dw_u32x4{0, 0, 0, 0};
    }
    int bc[Sh::NBT];
#pragma unroll
    for (int i = 0; i < Sh::NBT; ++i) bc[i] = block_col[min(b0 + t + kBlock * i, (int)nblocks - 1)];
    // raw values stream in now; value indices too, their table lookups wait for the gather of x
    // (every thread loads all E slots, those past the tile clamped to its last slot, and drops them:
    // no branch per load, no bytes outside the tile)
    T w[E];
    uint32_t vx[E];
#pragma unroll
    for (int j = 0; j < E; ++j) {
        const int k = t + kBlock * j;
        const int64_t ks = min<int64_t>(d0 + min(k, max(cnt - 1, 0)), nnz - 1);
        if constexpr (VIB == 0) w[j] = __builtin_nontemporal_load(values + ks);
        else vx[j] = dw_index<VIB>((const uint8_t *)vi, ks);
    }
#pragma unroll
    for (int j = 0; j < E; ++j) {
        if (t + kBlock * j >= cnt) {
            if constexpr (VIB == 0) w[j] = T(0);
            else vx[j] = 0u;
        }
    }
    // (2) to LDS, then the decode
    int32_t *scol = (int32_t *)buf;
    dw_u32x4 *sq = (dw_u32x4 *)(buf + Sh::DIST_OFF);
    int32_t *sbc = (int32_t *)(buf + Sh::BC_OFF);
    if constexpr (VIB == 1)
        if (t < nvals) s_tab[t] = values[t];
    if (t <= nr) s_rd[t] = (int)(off - d0);
    else if (t >= 128 && t - 128 <= nr) s_rb[t - 128] = off - b0;
#pragma unroll
    for (int i = 0; i < Sh::NQT; ++i)
        if (t + kBlock * i < nin) sq[t + kBlock * i] = q[i];
    for (int i = max(nin, 0) + t; i < nq; i += kBlock) {               // the last bytes of dist, one by one
        uint8_t *o = (uint8_t *)&sq[i];
        for (int k = 0; k < 16; ++k) o[k] = a0 + 16 * i + k < nnz ? dist[a0 + 16 * i + k] : 0;
    }
#pragma unroll
    for (int i = 0; i < Sh::NBT; ++i)
        if (t + kBlock * i < nb) sbc[t + kBlock * i] = bc[i];
    __syncthreads();
    const uint8_t *sd = (const uint8_t *)sq + (d0 & 15);              // sd[k]: distance byte of slot d0 + k
    if (row < r1) {
        // the row's lanes walk its block chain together (the count byte of each block) and lane g
        // notes the first slots of its blocks g and g + G; then every lane decodes its blocks at
        // once (a block past 2 G per row, none in the stiffness matrices, is decoded in the walk)
        // (the distance bytes of a block's further slots are read four at a time, as 32-bit words
        // of the staging area)
        const uint32_t *sw = (const uint32_t *)buf;
        const int boff = Sh::DIST_OFF + (int)(d0 & 15);              // byte offset of slot 0 in buf
        auto decode = [&](int j, int s0) {
            int col = sbc[j];
            const int c = sd[s0];
            scol[s0] = col;
            int bo = boff + s0 + 1, wi = bo >> 2, left = 4 - (bo & 3);
            uint32_t wv = sw[wi] >> (8 * (bo & 3));
            for (int i = 1; i <= c; ++i) {
                if (left == 0) { wv = sw[++wi]; left = 4; }
                col += (int)(wv & 0xffu) + 1;
                wv >>= 8;
                --left;
                scol[s0 + i] = col;
            }
        };
        int sl = s_rd[r], own0 = -1, own1 = -1;
        const int jb = s_rb[r], be = s_rb[r + 1];
        for (int j = jb; j < be; ++j) {
            const int k = j - jb;
            if (k == lane) own0 = sl;
            else if (k == lane + G) own1 = sl;
            sl += sd[sl] + 1;
        }
        if (own0 >= 0) decode(jb + lane, own0);
        if (own1 >= 0) decode(jb + lane + G, own1);
        if (be - jb > 2 * G) {                                        // blocks past 2 G: a second walk
            sl = s_rd[r];
            for (int j = jb; j < be; ++j) {
                const int k = j - jb;
                if (k >= 2 * G && k % G == lane) decode(j, sl);
                sl += sd[sl] + 1;
            }
        }
    }
    __syncthreads();
    // (3) columns of this thread's slots, x, products
    int c[E];
#pragma unroll
    for (int j = 0; j < E; ++j) {
        const int k = t + kBlock * j;
        c[j] = scol[k];                                            // k < C: inside the buffer
        if (k >= cnt) c[j] = 0;
    }
    __syncthreads();                                                  // the products overwrite the columns
    T xv[E];
#pragma unroll
    for (int j = 0; j < E; ++j) xv[j] = x[c[j]];
    if constexpr (VIB == 1) {
#pragma unroll
        for (int j = 0; j < E; ++j) {
            w[j] = s_tab[vx[j]];                                   // index 0 past the tile
            if (t + kBlock * j >= cnt) w[j] = T(0);
        }
    } else if constexpr (VIB != 0) {
#pragma unroll
        for (int j = 0; j < E; ++j) {
            w[j] = values[vx[j]];                                  // index 0 past the tile
            if (t + kBlock * j >= cnt) w[j] = T(0);
        }
    }
    T *sp = (T *)buf;
#pragma unroll
    for (int j = 0; j < E; ++j) sp[t + kBlock * j] = w[j] * xv[j];
    __syncthreads();
    // (4) row sums
    T a = 0;
    if (row < r1)
        for (int k = s_rd[r] + lane; k < s_rd[r + 1]; k += G) a += sp[k];
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) a += __shfl_xor(a, o, G);
    int bits = 0;
    if (lane == 0 && row < r1) bits = duvi_store<EPI, T>(op, row, a);
    duvi_flag<EPI, T>(op, bits);
}

// One colour of a multicolour Gauss-Seidel sweep: the nodes nodes[0 .. nnodes) are pairwise
// uncoupled; node q's rows 3q, 3q+1, 3q+2 (with gptr: the one to three rows gptr[q] ..
// gptr[q+1]) are relaxed in that order (the reverse when backward) with the current x.  A wavefront takes 2 nodes, 8 lanes per row (lanes 48-63 idle).  A row's
// lanes walk its block chain together (one distance byte per block, the same address in all 8
// lanes); lane l keeps the first slot of blocks l, l + 8, ... of each round of 8 and decodes
// them one slot at a time, gathering x and recording the entries of the node's own three
// columns.  The row sums (the node's own x still old) and the own 3x3 block meet in the node's
// first lane, which relaxes the three rows in turn, correcting each row sum for the node's
// values already updated: x_a = x_a + (b_a - sum_j A_aj x_j) / A_aa with x current.
template <int VIB, typename T>
__global__ __launch_bounds__(kBlock) void k_duvi_gs(int32_t nnodes, const int32_t *__restrict__ nodes,
                                                    const int32_t *__restrict__ gptr,
                                                    const int32_t *__restrict__ row_blocks,
                                                    const int32_t *__restrict__ row_dist,
                                                    const int32_t *__restrict__ block_col,
                                                    const uint8_t *__restrict__ dist, const void *__restrict__ vi,
                                                    const T *__restrict__ values, const T *__restrict__ b,
                                                    T *__restrict__ x, int backward, const int32_t *__restrict__ done)
{
    if (done && *done) return;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, grp = lane >> 3, l = lane & 7;
    const int ns = grp / 3, ra = grp - 3 * ns;                 // node slot (2: idle lanes), row of the node
    const int64_t k = ((int64_t)xcd_remap(blockIdx.x, gridDim.x) * kDwWaves + wv) * 2 + ns;
    if (k - ns >= nnodes) return;                              // the whole wave leaves together
    const bool valid = ns < 2 && k < nnodes;
    const int q = valid ? nodes[k] : 0;
    const int r0 = gptr ? gptr[q] : 3 * q, nr = gptr ? (valid ? gptr[q + 1] - r0 : 0) : 3;
    const bool act = valid && ra < nr;
    T acc = 0, own[3] = {0, 0, 0};
    if (act) {
        const int row = r0 + ra;
        const int rb1 = row_blocks[row + 1];
        int64_t d = row_dist[row];
        for (int bj = row_blocks[row]; bj < rb1; bj += kDwLanes) {
            const int nb = min(kDwLanes, rb1 - bj);
            int64_t my = -1;
            for (int j = 0; j < nb; ++j) {
                if (l == j) my = d;
                d += dist[d] + 1;
            }
            if (my < 0) continue;
            int col = block_col[bj + l];
            int64_t dd = my;
            const int64_t dend = dd + dist[dd] + 1;
            for (;;) {
                const T a = duvi_value<VIB, T>(vi, values, dd);
                acc += a * x[col];
                const int oc = col - r0;
                if (oc >= 0 && oc < nr) own[oc] = a;
                if (++dd == dend) break;
                col += dist[dd] + 1;
            }
        }
    }
#pragma unroll
    for (int o = kDwLanes / 2; o > 0; o >>= 1) {
        acc += __shfl_xor(acc, o, kDwLanes);
#pragma unroll
        for (int c = 0; c < 3; ++c) own[c] += __shfl_xor(own[c], o, kDwLanes);
    }
    // the node's three row sums and its own block, in the node's first lane
    const int base = 24 * min(ns, 1);
    T S[3], A[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        S[a] = __shfl(acc, base + 8 * a);
#pragma unroll
        for (int c = 0; c < 3; ++c) A[a][c] = __shfl(own[c], base + 8 * a);
    }
    if (!act || ra != 0 || l != 0) return;
    T xo[3], xn[3];
    const size_t i0 = (size_t)r0;
#pragma unroll
    for (int a = 0; a < 3; ++a) { xo[a] = a < nr ? x[i0 + a] : T(0); xn[a] = xo[a]; }
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int a = backward ? 2 - t : t;
        if (a >= nr) continue;
        T r = b[i0 + a] - S[a];
#pragma unroll
        for (int c = 0; c < 3; ++c) r -= A[a][c] * (xn[c] - xo[c]);   // zero for rows not yet relaxed
        xn[a] = xo[a] + r / A[a][a];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (a < nr) x[i0 + a] = xn[a];
}

// the wave form for every matrix (timing comparisons of the two forms)
// ------------------------------------------------------------------ batching form on the DuVi layout
// K right-hand sides at once (block vectors n x K, row-major), the matrix read once per block of
// columns.  A row's G lanes (the CSR block kernels' shape) turn the row into (column, value) pairs
// in LDS: the row's distance bytes and values are staged, the lanes walk its block chain together
// (the count byte of each block) and lane l decodes its own blocks l and l + G (further blocks in a
// second walk); then the per-row arithmetic of the CSR block kernels runs on the staged pairs.
// A tile or row longer than the staging area is decoded by every lane straight from global memory.

// columns of a row's slots [s, s + len) into scol (tile-relative slots), sd[k] = distance byte of
// slot k; jb .. je the row's blocks; lane l of the row's G lanes
__device__ __forceinline__ void duvi_row_columns(const uint8_t *sd, int32_t *scol, int s, int jb, int je,
                                                 const int32_t *__restrict__ block_col, int l, int G)
{
    auto decode = [&](int j, int p0) {
        int col = block_col[j];
        const int n = sd[p0];
        scol[p0] = col;
        for (int i = 1; i <= n; ++i) {
            col += sd[p0 + i] + 1;
            scol[p0 + i] = col;
        }
    };
    int pos = s, own0 = -1, own1 = -1;
    for (int j = jb; j < je; ++j) {
        const int k = j - jb;
        if (k == l) own0 = pos;
        else if (k == l + G) own1 = pos;
        pos += sd[pos] + 1;
    }
    if (own0 >= 0) decode(jb + l, own0);
    if (own1 >= 0) decode(jb + l + G, own1);
    if (je - jb > 2 * G) {
        pos = s;
        for (int j = jb; j < je; ++j) {
            const int k = j - jb;
            if (k >= 2 * G && k % G == l) decode(j, pos);
            pos += sd[pos] + 1;
        }
    }
}

// one lane's share (slots s + sl, s + sl + S, ...) of row `row` for column c, decoded from global
// memory (rows longer than the staging area)
template <int VIB, int K, typename TV, typename TX>
__device__ TV duvi_row_cols_global(int row, const int32_t *__restrict__ row_blocks, const int32_t *__restrict__ row_dist,
                                   const int32_t *__restrict__ block_col, const uint8_t *__restrict__ dist,
                                   const void *__restrict__ vi, const TV *__restrict__ values,
                                   const TX *__restrict__ X, int c, int sl)
{
    constexpr int S = BShape<K>::S;
    TV a = 0;
    int64_t d = row_dist[row];
    const int64_t ds = d;
    for (int j = row_blocks[row]; j < row_blocks[row + 1]; ++j) {
        int col = block_col[j];
        const int64_t dend = d + dist[d] + 1;
        for (;;) {
            if ((int)(d - ds) % S == sl) a += duvi_value<VIB, TV>(vi, values, d) * (TV)X[(size_t)col * K + c];
            if (++d == dend) break;
            col += dist[d] + 1;
        }
    }
    return a;
}

template <int VIB, typename TV, typename TX, int K, int EPI>
__global__ __launch_bounds__(kBlock) void k_duvi_spmm(int32_t nrows, const int32_t *__restrict__ row_blocks,
                                                      const int32_t *__restrict__ row_dist,
                                                      const int32_t *__restrict__ block_col,
                                                      const uint8_t *__restrict__ dist, const void *__restrict__ vi,
                                                      const TV *__restrict__ values, const TX *__restrict__ X,
                                                      const TX *__restrict__ B, const TV *__restrict__ pre,
                                                      TX *__restrict__ Y, const int32_t *__restrict__ done,
                                                      double *__restrict__ partials, int stride)
{
    using Sh = BShape<K>;
    __shared__ int32_t s_ci[Sh::CAP];
    __shared__ TV s_v[Sh::CAP];
    __shared__ uint8_t s_d[Sh::CAP];
    if (done && *done) return;
    const int t = threadIdx.x, g = t / Sh::G, l = t % Sh::G, c = l % K, sl = l / K;
    const int ntiles = (nrows + Sh::RPB - 1) / Sh::RPB;
    const int per = gridDim.x >> 3, xcd = blockIdx.x & 7;
    const int hi = (int)((int64_t)ntiles * (xcd + 1) / 8);
    double dot = 0;
    for (int tile = (int)((int64_t)ntiles * xcd / 8) + (int)(blockIdx.x >> 3); tile < hi; tile += per) {
        const int r0 = tile * Sh::RPB, r1 = min(r0 + Sh::RPB, nrows);
        const int64_t d0 = row_dist[r0], d1 = row_dist[r1];
        const bool staged = d1 - d0 <= Sh::CAP;     // the same for the whole block
        const int row = r0 + g;
        if (staged) {
            const int cnt = (int)(d1 - d0);
            __syncthreads();                        // the previous tile's readers are done
            for (int k = t; k < cnt; k += kBlock) {
                s_d[k] = dist[d0 + k];
                s_v[k] = duvi_value<VIB, TV>(vi, values, d0 + k);
            }
            __syncthreads();
            if (row < r1)
                duvi_row_columns(s_d, s_ci, (int)(row_dist[row] - d0), row_blocks[row], row_blocks[row + 1], block_col,
                                 l, Sh::G);
            __syncthreads();
        }
        if (row < r1) {
            const int64_t s = row_dist[row], e = row_dist[row + 1];
            TV acc = staged ? row_cols<K, TV, TX>(s_ci, s_v, (int)(s - d0), (int)(e - d0), X, c, sl)
                            : duvi_row_cols_global<VIB, K, TV, TX>(row, row_blocks, row_dist, block_col, dist, vi,
                                                                   values, X, c, sl);
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

int duvi_spmm_grid(int K, int32_t nrows)
{
    const int rpb = kBlock / (K >= 16 ? K : 16);
    return grid_tiles((nrows + rpb - 1) / rpb);
}

template <typename T>
void launch_duvi_spmm(int K, int epi, const mgx_duvi &A, const T *values, const T *X, const T *B, const T *pre, T *Y,
                      const int32_t *done, double *partials, int stride, hipStream_t s)
{
    if (A.nrows <= 0) return;
    const int grid = duvi_spmm_grid(K, A.nrows);
#define MGX_DSM(V, KK, EE)                                                                                      \
    hipLaunchKernelGGL((k_duvi_spmm<V, T, T, KK, EE>), dim3(grid), dim3(kBlock), 0, s, A.nrows, A.row_blocks,   \
                       A.row_dist, A.block_col, A.dist, A.vi, values, X, B, pre, Y, done, partials, stride)
#define MGX_DSM_E(V, KK)                                         \
    switch (epi) {                                               \
    case BEPI_SPMM: MGX_DSM(V, KK, BEPI_SPMM); break;            \
    case BEPI_ADD: MGX_DSM(V, KK, BEPI_ADD); break;              \
    case BEPI_RESID: MGX_DSM(V, KK, BEPI_RESID); break;          \
    case BEPI_JACOBI: MGX_DSM(V, KK, BEPI_JACOBI); break;        \
    default: MGX_DSM(V, KK, BEPI_DOT); break;                    \
    }
#define MGX_DSM_K(V)                                                                           \
    switch (K) {                                                                               \
    case 4: MGX_DSM_E(V, 4); break;                                                            \
    case 8: MGX_DSM_E(V, 8); break;                                                            \
    case 16: MGX_DSM_E(V, 16); break;                                                          \
    case 32: MGX_DSM_E(V, 32); break;                                                          \
    default: throw Error{MGX_ERR_INVALID_ARGUMENT, "block width must be 4, 8, 16 or 32"};      \
    }
    switch (A.vi_bytes) {
    case 1: MGX_DSM_K(1); break;
    case 2: MGX_DSM_K(2); break;
    case 4: MGX_DSM_K(4); break;
    default: MGX_DSM_K(0); break;
    }
#undef MGX_DSM_K
#undef MGX_DSM_E
#undef MGX_DSM
    MGX_HIP(hipGetLastError());
}
template void launch_duvi_spmm<double>(int, int, const mgx_duvi &, const double *, const double *, const double *,
                                       const double *, double *, const int32_t *, double *, int, hipStream_t);
template void launch_duvi_spmm<float>(int, int, const mgx_duvi &, const float *, const float *, const float *,
                                      const float *, float *, const int32_t *, double *, int, hipStream_t);

// Multicolour Gauss-Seidel for K columns on the DuVi layout: the node-group arithmetic of the CSR
// kernel (a block of kDgbBlock threads relaxes N nodes; thread (m, a, s, c) sums row a of node m
// for column c over the slots s, s + S, ... outside the node, and collects the node's diagonal
// block; one thread per (node, column) then relaxes the node's rows in order, the reverse order
// when backward), with each row's (column, value) pairs decoded into LDS first by its G lanes (all
// in one wavefront).  With gptr, node q is the group of one to three rows gptr[q] .. gptr[q+1].
constexpr int kDgbBlock = 192;
constexpr int kDgbRow = 96;                          // staged slots per row

template <int VIB, typename T, int K>
__global__ __launch_bounds__(kDgbBlock) void k_duvi_gs_block(int32_t nnodes, const int32_t *__restrict__ nodes,
                                                             const int32_t *__restrict__ gptr,
                                                             const int32_t *__restrict__ row_blocks,
                                                             const int32_t *__restrict__ row_dist,
                                                             const int32_t *__restrict__ block_col,
                                                             const uint8_t *__restrict__ dist,
                                                             const void *__restrict__ vi, const T *__restrict__ values,
                                                             const T *__restrict__ B, T *__restrict__ X, int backward,
                                                             const int32_t *__restrict__ done)
{
    constexpr int S = BShape<K>::S, G = BShape<K>::G, N = kDgbBlock / (3 * G), U = S == 1 ? 8 : 4;
    static_assert(64 % G == 0, "a row's lanes lie in one wavefront");
    __shared__ T sh[N][3][K];
    __shared__ T dg[N][3][3];
    __shared__ int32_t s_ci[N][3][kDgbRow];
    __shared__ T s_v[N][3][kDgbRow];
    __shared__ uint8_t s_d[N][3][kDgbRow];
    if (done && *done) return;
    const int base = xcd_remap(blockIdx.x, gridDim.x) * N;
    const int t = threadIdx.x, l = t % G, c = l % K, sl = l / K, a = (t / G) % 3, m = t / (3 * G);
    const int nd = base + m;
    T sa = 0, e0 = 0, e1 = 0, e2 = 0;
    int r0 = 0, nr = 0;
    if (nd < nnodes) {
        const int q = nodes[nd];
        r0 = gptr ? gptr[q] : 3 * q;
        nr = gptr ? gptr[q + 1] - r0 : 3;
    }
    if (a < nr) {
        const int row = r0 + a;
        const int64_t ds = row_dist[row], de = row_dist[row + 1];
        const int len = (int)(de - ds);
        if (len <= kDgbRow) {
            uint8_t *sd = s_d[m][a];
            int32_t *sc = s_ci[m][a];
            T *sv = s_v[m][a];
            for (int k = l; k < len; k += G) {
                sd[k] = dist[ds + k];
                sv[k] = duvi_value<VIB, T>(vi, values, ds + k);
            }
            dw_fence();                                          // the row's lanes: one wavefront
            duvi_row_columns(sd, sc, 0, row_blocks[row], row_blocks[row + 1], block_col, l, G);
            dw_fence();
            for (int k = sl; k < len; k += U * S) {
                int cc[U];
                T w[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int kk = k + u * S;
                    cc[u] = kk < len ? sc[kk] : -1;
                    w[u] = kk < len ? sv[kk] : T(0);
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
                    if (off < (unsigned)nr) { if (off == 0) e0 += w[u]; else if (off == 1) e1 += w[u]; else e2 += w[u]; }
                    else sa += w[u] * xv[u];
                }
            }
        } else {
            int64_t d = ds;
            for (int j = row_blocks[row]; j < row_blocks[row + 1]; ++j) {
                int col = block_col[j];
                const int64_t dend = d + dist[d] + 1;
                for (;;) {
                    if ((int)(d - ds) % S == sl) {
                        const T w = duvi_value<VIB, T>(vi, values, d);
                        const unsigned off = (unsigned)(col - r0);
                        if (off < (unsigned)nr) { if (off == 0) e0 += w; else if (off == 1) e1 += w; else e2 += w; }
                        else sa += w * X[(size_t)col * K + c];
                    }
                    if (++d == dend) break;
                    col += dist[d] + 1;
                }
            }
        }
    }
#pragma unroll
    for (int o = K; o < G; o <<= 1) {
        sa += __shfl_xor(sa, o, G);
        e0 += __shfl_xor(e0, o, G);
        e1 += __shfl_xor(e1, o, G);
        e2 += __shfl_xor(e2, o, G);
    }
    if (nd < nnodes && sl == 0) {
        sh[m][a][c] = sa;
        if (c == 0) { dg[m][a][0] = e0; dg[m][a][1] = e1; dg[m][a][2] = e2; }
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
void launch_duvi_gs_block(int K, int32_t nnodes, const int32_t *nodes, const int32_t *gptr, const mgx_duvi &A,
                          const T *values, const T *B, T *X, bool backward, const int32_t *done, hipStream_t s)
{
    if (nnodes <= 0) return;
#define MGX_DGB(V, KK)                                                                                          \
    hipLaunchKernelGGL((k_duvi_gs_block<V, T, KK>), dim3(grid_rows(nnodes, kDgbBlock / (3 * BShape<KK>::G))),  \
                       dim3(kDgbBlock), 0, s, nnodes, nodes, gptr, A.row_blocks, A.row_dist, A.block_col, A.dist, \
                       A.vi, values, B, X, backward ? 1 : 0, done)
#define MGX_DGB_K(V)                                                                           \
    switch (K) {                                                                               \
    case 4: MGX_DGB(V, 4); break;                                                              \
    case 8: MGX_DGB(V, 8); break;                                                              \
    case 16: MGX_DGB(V, 16); break;                                                            \
    case 32: MGX_DGB(V, 32); break;                                                            \
    default: throw Error{MGX_ERR_INVALID_ARGUMENT, "block width must be 4, 8, 16 or 32"};      \
    }
    switch (A.vi_bytes) {
    case 1: MGX_DGB_K(1); break;
    case 2: MGX_DGB_K(2); break;
    case 4: MGX_DGB_K(4); break;
    default: MGX_DGB_K(0); break;
    }
#undef MGX_DGB_K
#undef MGX_DGB
    MGX_HIP(hipGetLastError());
}
template void launch_duvi_gs_block<double>(int, int32_t, const int32_t *, const int32_t *, const mgx_duvi &,
                                           const double *, const double *, double *, bool, const int32_t *, hipStream_t);
template void launch_duvi_gs_block<float>(int, int32_t, const int32_t *, const int32_t *, const mgx_duvi &,
                                          const float *, const float *, float *, bool, const int32_t *, hipStream_t);

static bool g_duvi_wave_only = false;
void set_duvi_wave_only(bool on) { g_duvi_wave_only = on; }

template <typename T>
void launch_duvi(const mgx_duvi &A, const T *values, int epi, const T *x, const T *b, const T *pre, T *y,
                 int32_t *flag, double tol, const int32_t *done, hipStream_t s)
{
    if (A.nrows <= 0) return;
    if (epi == DEPI_CHECK && !std::is_same<T, double>::value)
        throw Error{MGX_ERR_INVALID_ARGUMENT, "the residual check runs in FP64"};
    const DuviOp<T> op{values, x, b, pre, y, flag, tol};
    // long rows (the stiffness levels) take the stream form, short rows (transfers) the wave form
    const bool stream = A.nnz >= 16 * (int64_t)A.nrows && A.nblocks > 0 && !g_duvi_wave_only;
    const dim3 g(stream ? grid_rows((A.nrows + DsShape<T>::RB - 1) / DsShape<T>::RB, 1)
                        : grid_rows((A.nrows + kDwRows - 1) / kDwRows, kDwWaves)),
        bl(kBlock);
#define MGX_DV(V, E)                                                                                             \
    if (stream)                                                                                                  \
        hipLaunchKernelGGL((k_duvi_stream<V, T, E>), g, bl, 0, s, A.nrows, A.nnz, A.nblocks,                     \
                           (int32_t)std::min<int64_t>(A.nuniq, 256), A.row_blocks, A.row_dist, A.block_col,        \
                           A.dist, A.vi, op, done);                                                               \
    else                                                                                                         \
        hipLaunchKernelGGL((k_duvi_wave<V, T, E>), g, bl, 0, s, A.nrows, A.nnz, A.row_blocks, A.row_dist,        \
                           A.block_col, A.dist, A.vi, op, done)
#define MGX_DVE(V)                                                                                               \
    switch (epi) {                                                                                               \
    case DEPI_SPMV: MGX_DV(V, DEPI_SPMV); break;                                                                 \
    case DEPI_ADD: MGX_DV(V, DEPI_ADD); break;                                                                   \
    case DEPI_RESID: MGX_DV(V, DEPI_RESID); break;                                                               \
    case DEPI_JACOBI: MGX_DV(V, DEPI_JACOBI); break;                                                             \
    default:                                                                                                     \
        if constexpr (std::is_same<T, double>::value) MGX_DV(V, DEPI_CHECK);                                     \
        break;                                                                                                   \
    }
    switch (A.vi_bytes) {
    case 1: MGX_DVE(1); break;
    case 2: MGX_DVE(2); break;
    case 4: MGX_DVE(4); break;
    default: MGX_DVE(0); break;
    }
#undef MGX_DVE
#undef MGX_DV
    MGX_HIP(hipGetLastError());
}

template <typename T>
void launch_duvi_gs(const mgx_duvi &A, const T *values, int32_t nnodes, const int32_t *nodes, const T *b, T *x,
                    bool backward, const int32_t *done, hipStream_t s, const int32_t *gptr)
{
    if (nnodes <= 0) return;
    const dim3 g(grid_rows((nnodes + 1) / 2, kDwWaves)), bl(kBlock);
#define MGX_DG(V) hipLaunchKernelGGL((k_duvi_gs<V, T>), g, bl, 0, s, nnodes, nodes, gptr, A.row_blocks, A.row_dist, \
                                     A.block_col, A.dist, A.vi, values, b, x, backward ? 1 : 0, done)
    switch (A.vi_bytes) {
    case 1: MGX_DG(1); break;
    case 2: MGX_DG(2); break;
    case 4: MGX_DG(4); break;
    default: MGX_DG(0); break;
    }
#undef MGX_DG
    MGX_HIP(hipGetLastError());
}

template void launch_duvi<double>(const mgx_duvi &, const double *, int, const double *, const double *,
                                  const double *, double *, int32_t *, double, const int32_t *, hipStream_t);
template void launch_duvi<float>(const mgx_duvi &, const float *, int, const float *, const float *, const float *,
                                 float *, int32_t *, double, const int32_t *, hipStream_t);
template void launch_duvi_gs<double>(const mgx_duvi &, const double *, int32_t, const int32_t *, const double *,
                                     double *, bool, const int32_t *, hipStream_t, const int32_t *);
template void launch_duvi_gs<float>(const mgx_duvi &, const float *, int32_t, const int32_t *, const float *, float *,
                                    bool, const int32_t *, hipStream_t, const int32_t *);

void launch_duvi_spmv(const mgx_duvi &A, const double *x, double *y, bool clear, const int32_t *done, hipStream_t s)
{
    launch_duvi<double>(A, A.values, clear ? DEPI_SPMV : DEPI_ADD, x, nullptr, nullptr, y, nullptr, 0.0, done, s);
}

}  // namespace mgx
