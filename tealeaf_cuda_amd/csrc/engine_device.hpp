// engine_device.hpp -- device-side building blocks of the TeaLeaf engine (gfx950).
//
// Work decomposition: a workgroup is 4 wavefronts of 64 lanes.  A wavefront owns
// a 64-column x 4-row tile; each lane owns one column of it ("strip"), so the
// 4-cell vertical tridiagonal of the block-Jacobi preconditioner
// (tea_block_jacobi.cuknl:11-46) is solved in registers by one lane, and a
// lane's vertical stencil neighbours are reused from registers.  The 4 waves
// of a workgroup stack vertically: a workgroup covers 64 x 16 cells.
//
// Workgroups are dispatched round-robin over the 8 XCDs; tile_of_block() maps
// them so that each XCD walks one contiguous band of tiles, keeping the rows
// shared by vertically adjacent tiles in the same L2.
//
// Reductions use compensated (double-double) accumulation: products are split
// exactly with fma, sums carried with TwoSum.  The rounded result is the
// correctly rounded dot product for all practical purposes, so it does not
// depend on the order in which lanes, waves and workgroups are combined.
#pragma once
#include <hip/hip_runtime.h>

namespace tea {

constexpr int WAVE = 64;
constexpr int TILE_X = 64;
constexpr int STRIP = 4;          // block-Jacobi strip height (JACOBI_BLOCK_SIZE)
constexpr int WAVES = 4;
constexpr int TILE_Y = STRIP * WAVES;
constexpr int BLOCK = WAVE * WAVES;
constexpr int NUM_XCD = 8;
constexpr int MAX_RED = 54;       // values reduced by one launch at most

enum { PREC_NONE = 1, PREC_DIAG = 2, PREC_BLOCK = 3 };

// Inclusive cell extent in reference coordinates (column, row; halo included).
struct Ext {
    int x0, x1, y0, y1;
};

struct DD {
    double s, c;
};

__device__ __forceinline__ void dd_add_prod(DD &a, double x, double y)
{
    const double p = x * y;
    const double pe = __builtin_fma(x, y, -p);
    const double s = a.s + p;
    const double bb = s - a.s;
    const double se = (a.s - (s - bb)) + (p - bb);
    a.s = s;
    a.c += se + pe;
}

__device__ __forceinline__ void dd_merge(DD &a, const DD b)
{
    const double s = a.s + b.s;
    const double bb = s - a.s;
    const double se = (a.s - (s - bb)) + (b.s - bb);
    a.s = s;
    a.c += (se + b.c);
}

__device__ __forceinline__ DD wave_reduce(DD v)
{
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        DD o;
        o.s = __shfl_xor(v.s, off, WAVE);
        o.c = __shfl_xor(v.c, off, WAVE);
        dd_merge(v, o);
    }
    return v;
}

// Reduction output: one compensated partial per workgroup and value in device
// memory; a one-workgroup finishing launch (k_finish) combines them in a fixed
// order and writes the rounded results to pinned host memory.  No atomics:
// every workgroup owns its slot.
struct Red {
    DD *partials;        // MAX_RED * max_blocks
    int max_blocks;
};

// Combines NV per-lane accumulators over the workgroup and publishes the
// workgroup's partials.
template <int NV>
__device__ __forceinline__ void block_reduce(DD (&v)[NV], const Red red)
{
    __shared__ DD lds[NV][WAVES];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
#pragma unroll
    for (int j = 0; j < NV; j++) v[j] = wave_reduce(v[j]);
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++) lds[j][wave] = v[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++) {
            DD t = lds[j][0];
            for (int w = 1; w < WAVES; w++) dd_merge(t, lds[j][w]);
            red.partials[j * red.max_blocks + blockIdx.x] = t;
        }
    }
}

// Combines the nblocks partials of each of nv values (lane t takes partials
// t, t+1024, ...; then the wave and workgroup trees) and writes out[j] =
// rounded sum.  Launched with one workgroup of FINISH_THREADS threads.
constexpr int FINISH_THREADS = 1024;

__global__ __launch_bounds__(FINISH_THREADS) void k_finish(const DD *__restrict__ partials, int max_blocks,
                                                          int nblocks, int nv, double *__restrict__ out,
                                                          double *__restrict__ dev_out)
{
    __shared__ DD lds[FINISH_THREADS / WAVE];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
    for (int j = 0; j < nv; j++) {
        DD acc{0.0, 0.0};
        const DD *p = partials + (size_t)j * max_blocks;
        // lane t merges partials t, t+1024, ... in that order; eight loads are
        // issued before the first merge waits
        for (int b0 = threadIdx.x; b0 < nblocks; b0 += 8 * FINISH_THREADS) {
            DD x[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int b = b0 + u * FINISH_THREADS;
                x[u] = b < nblocks ? p[b] : DD{0.0, 0.0};
            }
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (b0 + u * FINISH_THREADS < nblocks) dd_merge(acc, x[u]);
        }
        acc = wave_reduce(acc);
        if (lane == 0) lds[wave] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            DD t = lds[0];
            for (int w = 1; w < FINISH_THREADS / WAVE; w++) dd_merge(t, lds[w]);
            out[j] = t.s + t.c;
            if (dev_out) dev_out[j] = t.s + t.c;   // for a following launch on the device
        }
        __syncthreads();
    }
}

// Column and first row of a lane's strip; len = rows in the strip (0 if none).
struct Strip {
    int c, r0, len;
};

// Walks the tiles of extent e assigned to this workgroup.  Workgroup b runs on
// XCD b % 8; the tiles (row-major, 64 x 16 cells) are split into 8 contiguous
// bands, one per XCD, and the workgroups of an XCD take the tiles of its band
// round-robin, so vertically adjacent tiles share an L2.
// TX: tile width in cells (TILE_X, or 2*TILE_X for the FP32 kernels whose
// lanes hold two columns).
template <int TX = TILE_X>
struct TileLoopT {
    int nbx, count, start, k, nk;
    int t;
    Ext e;
    __device__ __forceinline__ explicit TileLoopT(const Ext ext) : e(ext)
    {
        nbx = (e.x1 - e.x0 + TX) / TX;
        const int nby = (e.y1 - e.y0 + TILE_Y) / TILE_Y;
        const int ntiles = (e.x1 < e.x0 || e.y1 < e.y0) ? 0 : nbx * nby;
        const int b = (int)blockIdx.x, g = (int)gridDim.x;
        const int xcd = b % NUM_XCD;
        const int q = ntiles / NUM_XCD, rem = ntiles % NUM_XCD;
        count = q + (xcd < rem ? 1 : 0);
        start = xcd * q + (xcd < rem ? xcd : rem);
        k = b / NUM_XCD;
        nk = g / NUM_XCD + (xcd < g % NUM_XCD ? 1 : 0);
        t = k;
    }
    __device__ __forceinline__ bool more() const { return t < count; }
    __device__ __forceinline__ void advance() { t += nk; }
    __device__ __forceinline__ Strip strip() const
    {
        const int tile = start + t;
        const int ty = tile / nbx;
        const int tx = tile - ty * nbx;
        const int lane = threadIdx.x & (WAVE - 1);
        const int wave = threadIdx.x / WAVE;
        Strip s;
        s.c = e.x0 + tx * TX + (TX / TILE_X) * lane;
        s.r0 = e.y0 + ty * TILE_Y + wave * STRIP;
        int len = e.y1 - s.r0 + 1;
        len = len < STRIP ? len : STRIP;
        if (s.c > e.x1 || len < 0) len = 0;
        s.len = len;
        return s;
    }
};
using TileLoop = TileLoopT<TILE_X>;


// 5-point operator (I - dt div K grad) x at flat index i, in the expression
// order of tea_leaf_cg.cuknl:53-57, evaluated in the field type T (FP64, or
// FP32 for the mixed-precision inner solver).
template <typename T>
__device__ __forceinline__ T apply_A(const T *__restrict__ x, const T *__restrict__ Kx, const T *__restrict__ Ky,
                                     long i, long pitch)
{
    return (T(1.0) + (Ky[i + pitch] + Ky[i]) + (Kx[i + 1] + Kx[i])) * x[i]
           - (Ky[i + pitch] * x[i + pitch] + Ky[i] * x[i - pitch])
           - (Kx[i + 1] * x[i + 1] + Kx[i] * x[i - 1]);
}

// Diagonal of the operator at i (COEF_B of tea_block_jacobi.cuknl:6).
template <typename T>
__device__ __forceinline__ T diag_A(const T *__restrict__ Kx, const T *__restrict__ Ky, long i, long pitch)
{
    return T(1.0) + (Ky[i + pitch] + Ky[i]) + (Kx[i + 1] + Kx[i]);
}

// z = M_block^-1 rhs on one strip of len (1..4) rows starting at flat index i0
// (Thomas algorithm, tea_block_jacobi.cuknl:11-46), with the factors cp, bfp
// recomputed in registers from Kx, Ky.  The factor expressions are those of
// block_init (tea_block_jacobi.cuknl:83-111), so every value equals the stored
// one bit for bit, and the pass reads 16 bytes per cell less.
template <typename T>
__device__ __forceinline__ void block_solve_strip(const T (&rhs)[STRIP], T (&z)[STRIP], int len,
                                                  const T *__restrict__ Kx, const T *__restrict__ Ky, long i0,
                                                  long pitch)
{
    T cp[STRIP], bfp[STRIP], dp[STRIP];
    const T b0 = diag_A(Kx, Ky, i0, pitch);
    cp[0] = (-Ky[i0 + pitch]) / b0;
    dp[0] = rhs[0] / b0;
#pragma unroll
    for (int k = 1; k < STRIP; k++) {
        if (k < len) {
            const long i = i0 + k * pitch;
            bfp[k] = T(1.0) / (diag_A(Kx, Ky, i, pitch) - (-Ky[i]) * cp[k - 1]);
            cp[k] = (-Ky[i + pitch]) * bfp[k];
            dp[k] = (rhs[k] - (-Ky[i]) * dp[k - 1]) * bfp[k];
        }
    }
#pragma unroll
    for (int k = STRIP - 1; k >= 0; k--) {
        if (k == len - 1) z[k] = dp[k];
        else if (k < len - 1) z[k] = dp[k] - cp[k] * z[k + 1];
    }
}

}  // namespace tea
