// warmstart.hip -- the warm-start network's kernels (gfx950) and its trainer.
//
// Convolutions with 32 input channels run on the matrix cores as implicit
// GEMMs with v_mfma_f32_16x16x32_bf16: M = 16 output channels, N = 16
// pixels of one image row, K = the 32 input channels of one tap, accumulated
// over the 9 taps in FP32.  The backward pass of such a layer with respect to
// its input is the same kernel with the taps flipped and the channel roles
// swapped; its weight gradient is a GEMM with K = pixels (M = input channels,
// N = output channels).  The 32 -> 1 output layer is VALU work.
//
// Operand layout of the 16x16x32 BF16 MFMA (64 lanes): lane l holds row
// l % 16 of A and column l % 16 of B, each with 8 K-values at positions
// 8*(l/16) .. 8*(l/16)+7; the result lane l holds column l % 16, rows
// 4*(l/16) .. 4*(l/16)+3.  A and B use the same K positions, so any
// permutation of K inside the hardware leaves the product unchanged.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/tea_warmstart.h"
#include "ws_net.hpp"

#define EXPORT extern "C" __attribute__((visibility("default")))

namespace ws {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int BLOCK = 256;
constexpr int NPLANES = 6;   // training sample: f0, f1, f2, target, Kx, Ky

__device__ __forceinline__ floatx4 mfma(bf16x8 a, bf16x8 b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float bf16_round(float x) { return (float)(__bf16)x; }

// Flat index of pixel (y, x) of image b in a padded image batch (y, x interior
// coordinates; -1 and H / W address the zero border).
__device__ __forceinline__ long padded(int b, int y, int x, int H, int W)
{
    return ((long)b * (H + 2) + (y + 1)) * (W + 2) + (x + 1);
}

// ---------------------------------------------------------------- parameters

__global__ void k_pack(const float *__restrict__ p, __bf16 *__restrict__ packed, float *__restrict__ w6)
{
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e < TAPS * CH) w6[e] = bf16_round(p[OFF_W6 + e]);
    if (e >= PK_TOTAL) return;
    float v;
    if (e >= PK_OUT) {   // layer 6: W6 in row 0 (lanes 0, 16, 32, 48), zeros elsewhere
        const int r = e - PK_OUT, t = r / FRAG, lane = (r / 8) % 64, i = r % 8;
        v = lane % 16 == 0 ? p[OFF_W6 + t * CH + (lane / 16) * 8 + i] : 0.0f;
    } else if (e < PK_HID) {
        const int nb = e / FRAG, lane = (e / 8) % 64, i = e % 8;
        const int k = (lane / 16) * 8 + i, co = nb * 16 + lane % 16;
        v = k < K1 ? p[OFF_W1 + k * CH + co] : 0.0f;
    } else {
        const int e2 = e - PK_HID;
        const int per_layer = TAPS * 2 * FRAG;
        const bool bwd = e2 >= HIDDEN * per_layer;
        const int r0 = e2 % (HIDDEN * per_layer);
        const int l = r0 / per_layer + 2, r = r0 % per_layer;
        const int t = r / (2 * FRAG), nb = (r / FRAG) % 2, lane = (r / 8) % 64, i = r % 8;
        const int k = (lane / 16) * 8 + i, n = nb * 16 + lane % 16;
        const float *w = p + off_w(l);
        // forward: K = ci, N = co; backward: K = co, N = ci, tap mirrored
        v = bwd ? w[((TAPS - 1 - t) * CH + n) * CH + k] : w[(t * CH + k) * CH + n];
    }
    packed[e] = (__bf16)v;
}

// ---------------------------------------------------------------- forward

// planes[b][f][y][x] (npl planes per image, the first three are the features)
// -> x1[b][y][x][k], k = tap*3 + f, zero for k >= 27 and outside the image.
__global__ void k_im2col(const float *__restrict__ planes, int npl, __bf16 *__restrict__ x1, int B, int H, int W)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= (long)B * H * W) return;
    const int x = (int)(p % W), y = (int)((p / W) % H), b = (int)(p / ((long)W * H));
    const float *pl = planes + (long)b * npl * H * W;
    bf16x8 v[4];
#pragma unroll
    for (int k = 0; k < CH; k++) {
        float f = 0.0f;
        if (k < K1) {
            const int t = k / 3, c = k % 3;
            const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
            if (yy >= 0 && yy < H && xx >= 0 && xx < W) f = pl[(long)c * H * W + (long)yy * W + xx];
        }
        v[k / 8][k % 8] = (__bf16)f;
    }
    bf16x8 *dst = (bf16x8 *)(x1 + padded(b, y, x, H, W) * CH);
#pragma unroll
    for (int q = 0; q < 4; q++) dst[q] = v[q];
}

// Convolution kernels: a workgroup of CONV_ROWS waves computes rows
// y0 .. y0+CONV_ROWS-1, 64 pixels from x0, one row per wave (4 blocks of 16
// pixels) and all 32 output channels (2 blocks of 16).  It first stages its
// input window (rows y0-1 .. y0+CONV_ROWS, pixels x0-1 .. x0+64) in LDS with
// 16-byte loads; the taps then read the window from LDS.  The MFMA's A operand
// is the weight fragment (rows = output channels), B the activations
// (columns = pixels), so each lane's result is 4 consecutive channels of one
// pixel, stored as one 8-byte vector.  Grid (W/64, H/ROWS, B); ROWS = 8
// where H allows it (a smaller share of re-read halo rows), else 4.
constexpr int WIN_X = 64 + 2;
constexpr int LDS_PIX = CH + 8;   // bf16 per staged pixel: an 80-byte stride spreads 16-byte reads over the banks
inline int conv_rows(int H) { return H % 8 == 0 ? 8 : 4; }

// Stages rows y0-1 .. y0+ROWS of a padded NHWC32 image into win.
template <int ROWS>
__device__ __forceinline__ void stage_nhwc32(const __bf16 *__restrict__ in, __bf16 *win, int b, int y0, int x0,
                                             int H, int W)
{
    constexpr int CHUNKS = (ROWS + 2) * WIN_X * 4;   // 16-byte pieces, 4 per pixel
    for (int c = threadIdx.x; c < CHUNKS; c += ROWS * 64) {
        const int r = c / (WIN_X * 4), rem = c - r * (WIN_X * 4), px = rem >> 2, q = rem & 3;
        const bf16x8 v = *(const bf16x8 *)(in + padded(b, y0 - 1 + r, x0 - 1 + px, H, W) * CH + q * 8);
        *(bf16x8 *)(win + (r * WIN_X + px) * LDS_PIX + q * 8) = v;
    }
    __syncthreads();
}

// acc (4 pixel blocks x 2 channel blocks) -> out at row y, pixels x0..x0+63.
// Forward: bf16(relu(acc + bias)); backward (BWD): bf16(acc * (mask > 0)).
template <bool BWD>
__device__ __forceinline__ void store_nhwc32(const floatx4 (&acc)[4][2], __bf16 *__restrict__ out,
                                             const float *__restrict__ bias, const __bf16 *__restrict__ mask, int b,
                                             int y, int x0, int H, int W)
{
    const int lane = threadIdx.x & 63, m = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
        for (int nb = 0; nb < 2; nb++) {
            const int co = nb * 16 + kg * 4;
            const long idx = padded(b, y, x0 + mb * 16 + m, H, W) * CH + co;
            bf16x4 o;
            bf16x4 mk;
            if (BWD) mk = *(const bf16x4 *)(mask + idx);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                float v = acc[mb][nb][i];
                if (BWD) v = (float)mk[i] > 0.0f ? v : 0.0f;
                else v = fmaxf(v + bias[co + i], 0.0f);
                o[i] = (__bf16)v;
            }
            *(bf16x4 *)(out + idx) = o;
        }
}

// 32 -> 32 channels, 3x3 taps.
// Forward:  out = bf16(relu(conv(in) + bias)).
// Backward (BWD): out = bf16(conv(in) * (mask > 0)), mask = the layer's forward input.
template <int ROWS, bool BWD>
__global__ __launch_bounds__(ROWS * 64) void k_conv32(const __bf16 *__restrict__ in, __bf16 *__restrict__ out,
                                                      const bf16x8 *__restrict__ frag, const float *__restrict__ bias,
                                                      const __bf16 *__restrict__ mask, int H, int W)
{
    __shared__ __attribute__((aligned(16))) __bf16 win[(ROWS + 2) * WIN_X * LDS_PIX];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kg = lane >> 4;
    const int y0 = blockIdx.y * ROWS, x0 = blockIdx.x * 64, b = blockIdx.z;
    stage_nhwc32<ROWS>(in, win, b, y0, x0, H, W);
    floatx4 acc[4][2];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) acc[mb][0] = acc[mb][1] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < TAPS; t++) {
        const int dy = t / 3 - 1, dx = t % 3 - 1;
        const bf16x8 w0 = frag[(t * 2 + 0) * 64 + lane];
        const bf16x8 w1 = frag[(t * 2 + 1) * 64 + lane];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const bf16x8 a = *(const bf16x8 *)(win + ((wave + 1 + dy) * WIN_X + 1 + mb * 16 + m + dx) * LDS_PIX + kg * 8);
            acc[mb][0] = mfma(w0, a, acc[mb][0]);
            acc[mb][1] = mfma(w1, a, acc[mb][1]);
        }
    }
    store_nhwc32<BWD>(acc, out, bias, mask, b, y0 + wave, x0, H, W);
}

// Layer 1 straight from the feature planes (planes[b][f][y][x], npl planes
// per image, features f = 0..2): the im2col value k = tap*3 + f (zero for
// k >= 27 and outside the image) of each pixel is formed from the staged
// planes, then one MFMA K-block does the whole 3x3x3 -> 32 convolution.
template <int ROWS>
__global__ __launch_bounds__(ROWS * 64) void k_conv_first(const float *__restrict__ planes, int npl,
                                                          __bf16 *__restrict__ out, const bf16x8 *__restrict__ frag,
                                                          const float *__restrict__ bias, int H, int W)
{
    constexpr int WIN_Y = ROWS + 2;
    __shared__ float feat[NFEAT][WIN_Y][WIN_X];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kg = lane >> 4;
    const int y0 = blockIdx.y * ROWS, x0 = blockIdx.x * 64, b = blockIdx.z;
    const float *pl = planes + (long)b * npl * H * W;
    for (int c = threadIdx.x; c < NFEAT * WIN_Y * WIN_X; c += ROWS * 64) {
        const int f = c / (WIN_Y * WIN_X), rem = c - f * (WIN_Y * WIN_X), r = rem / WIN_X, px = rem - r * WIN_X;
        const int y = y0 - 1 + r, x = x0 - 1 + px;
        feat[f][r][px] = (y >= 0 && y < H && x >= 0 && x < W) ? pl[(long)f * H * W + (long)y * W + x] : 0.0f;
    }
    __syncthreads();
    // this lane's 8 K-values: offsets into feat relative to the pixel
    int off[8];
    bool real[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int k = kg * 8 + i, t = k / 3, f = k - 3 * t;
        real[i] = k < K1;
        off[i] = real[i] ? (f * WIN_Y + (t / 3 - 1)) * WIN_X + (t % 3 - 1) : 0;
    }
    const float *base = &feat[0][wave + 1][1];
    const bf16x8 w0 = frag[lane], w1 = frag[64 + lane];
    floatx4 acc[4][2];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        bf16x8 a;
#pragma unroll
        for (int i = 0; i < 8; i++) a[i] = (__bf16)(real[i] ? base[mb * 16 + m + off[i]] : 0.0f);
        const floatx4 z = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        acc[mb][0] = mfma(w0, a, z);
        acc[mb][1] = mfma(w1, a, z);
    }
    store_nhwc32<false>(acc, out, bias, (const __bf16 *)nullptr, b, y0 + wave, x0, H, W);
}

// Layer 6 (32 -> 1) on the matrix cores: the A fragment (packed at PK_OUT)
// holds W6 in its first row (output channel 0) and zeros elsewhere;
// out[b][y][x] = b6 + acc.
template <int ROWS>
__global__ __launch_bounds__(ROWS * 64) void k_conv_out(const __bf16 *__restrict__ a5, const bf16x8 *__restrict__ frag,
                                                        const float *__restrict__ b6, float *__restrict__ out, int H,
                                                        int W)
{
    __shared__ __attribute__((aligned(16))) __bf16 win[(ROWS + 2) * WIN_X * LDS_PIX];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kg = lane >> 4;
    const int y0 = blockIdx.y * ROWS, x0 = blockIdx.x * 64, b = blockIdx.z;
    stage_nhwc32<ROWS>(a5, win, b, y0, x0, H, W);
    floatx4 acc[4];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) acc[mb] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < TAPS; t++) {
        const int dy = t / 3 - 1, dx = t % 3 - 1;
        const bf16x8 w = frag[t * 64 + lane];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const bf16x8 a = *(const bf16x8 *)(win + ((wave + 1 + dy) * WIN_X + 1 + mb * 16 + m + dx) * LDS_PIX + kg * 8);
            acc[mb] = mfma(w, a, acc[mb]);
        }
    }
    if (kg == 0) {
        const float bv = b6[0];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) out[((long)b * H + y0 + wave) * W + x0 + mb * 16 + m] = acc[mb][0] + bv;
    }
}

// The 32-channel-input layers (k_conv32 forward, k_conv_out) as a march down a
// run of run_rows rows of one 64-pixel band (inference): the workgroup keeps
// the ROWS + 2 input rows of the current block of ROWS output rows in an LDS
// ring (row y in slot (y + 1) % (ROWS + 2)) and, while the matrix cores work
// on that block, has the next block's ROWS new rows in flight into registers;
// they replace the rows no later block needs once every wave is done.  So
// each input row is read once per run (a window per block re-reads 2 of every
// ROWS + 2), and the loads of a block overlap the products of the one before.
// The operands and their order are those of the windowed kernels, so the
// values are the same.  Grid (W/64, H/run_rows, B); run_rows a multiple of ROWS.
template <int ROWS> struct MarchRing {
    static constexpr int NR = ROWS + 2;
    static constexpr int PIECES = WIN_X * 4;                    // 16-byte pieces per staged row
    static constexpr int NEW = ROWS * PIECES;                   // pieces per block of new rows
    static constexpr int PER = (NEW + ROWS * 64 - 1) / (ROWS * 64);
    __bf16 *win;
    const __bf16 *in;
    int b, x0, H, W;
    __device__ __forceinline__ static int slot(int y) { return (y + 1) % NR; }
    __device__ __forceinline__ const bf16x8 *src(int y, int px, int q) const
    {
        return (const bf16x8 *)(in + padded(b, y, x0 - 1 + px, H, W) * CH + q * 8);
    }
    __device__ __forceinline__ bf16x8 *dst(int y, int px, int q) const
    {
        return (bf16x8 *)(win + (slot(y) * WIN_X + px) * LDS_PIX + q * 8);
    }
    // rows y_first .. y_first + n - 1 straight into the ring
    __device__ __forceinline__ void stage(int y_first, int n) const
    {
        for (int c = threadIdx.x; c < n * PIECES; c += ROWS * 64) {
            const int r = c / PIECES, rem = c - r * PIECES;
            *dst(y_first + r, rem >> 2, rem & 3) = *src(y_first + r, rem >> 2, rem & 3);
        }
    }
    __device__ __forceinline__ void prefetch(bf16x8 (&pf)[PER], int y_first) const
    {
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int c = threadIdx.x + j * ROWS * 64;
            if (c < NEW) {
                const int r = c / PIECES, rem = c - r * PIECES;
                pf[j] = *src(y_first + r, rem >> 2, rem & 3);
            }
        }
    }
    __device__ __forceinline__ void commit(const bf16x8 (&pf)[PER], int y_first) const
    {
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int c = threadIdx.x + j * ROWS * 64;
            if (c < NEW) {
                const int r = c / PIECES, rem = c - r * PIECES;
                *dst(y_first + r, rem >> 2, rem & 3) = pf[j];
            }
        }
    }
    // input row y, pixel x0 + px (px = -1 .. 64), this lane's 8 channels
    __device__ __forceinline__ bf16x8 at(int y, int px, int kg) const
    {
        return *(const bf16x8 *)(win + (slot(y) * WIN_X + 1 + px) * LDS_PIX + kg * 8);
    }
};

// body(ring, y0): the products and stores of the block of output rows y0 .. y0+ROWS-1
template <int ROWS, typename Body>
__device__ __forceinline__ void conv_march(const __bf16 *__restrict__ in, __bf16 *win, int H, int W, int run_rows,
                                           Body body)
{
    using Ring = MarchRing<ROWS>;
    const Ring ring{win, in, (int)blockIdx.z, (int)blockIdx.x * 64, H, W};
    const int ya = blockIdx.y * run_rows, yb = ya + run_rows < H ? ya + run_rows : H;
    ring.stage(ya - 1, ROWS + 2);
    bf16x8 pf[Ring::PER];
    for (int y0 = ya; y0 < yb; y0 += ROWS) {
        const bool more = y0 + ROWS < yb;
        if (more) ring.prefetch(pf, y0 + ROWS + 1);   // the next block's new rows
        __syncthreads();                              // rows y0-1 .. y0+ROWS in the ring
        body(ring, y0);
        __syncthreads();                              // rows y0-1 .. y0+ROWS-2 no longer read
        if (more) ring.commit(pf, y0 + ROWS + 1);
    }
}

// The weight fragments and biases are read once per workgroup, into LDS and
// registers: a global read inside the march would wait for the next block's
// rows in flight before it (the memory counter completes in order).
template <int ROWS>
__global__ __launch_bounds__(ROWS * 64) void k_conv32_march(const __bf16 *__restrict__ in, __bf16 *__restrict__ out,
                                                            const bf16x8 *__restrict__ frag,
                                                            const float *__restrict__ bias, int H, int W, int run_rows)
{
    __shared__ __attribute__((aligned(16))) __bf16 win[(ROWS + 2) * WIN_X * LDS_PIX];
    __shared__ bf16x8 wsh[TAPS * 2 * 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kg = lane >> 4;
    for (int i = threadIdx.x; i < TAPS * 2 * 64; i += ROWS * 64) wsh[i] = frag[i];
    float bv[2][4];
#pragma unroll
    for (int nb = 0; nb < 2; nb++)
#pragma unroll
        for (int i = 0; i < 4; i++) bv[nb][i] = bias[nb * 16 + kg * 4 + i];
    conv_march<ROWS>(in, win, H, W, run_rows, [&](const MarchRing<ROWS> &ring, int y0) {
        floatx4 acc[4][2];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) acc[mb][0] = acc[mb][1] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < TAPS; t++) {
            const int dy = t / 3 - 1, dx = t % 3 - 1;
            const bf16x8 w0 = wsh[(t * 2 + 0) * 64 + lane];
            const bf16x8 w1 = wsh[(t * 2 + 1) * 64 + lane];
#pragma unroll
            for (int mb = 0; mb < 4; mb++) {
                const bf16x8 a = ring.at(y0 + wave + dy, mb * 16 + m + dx, kg);
                acc[mb][0] = mfma(w0, a, acc[mb][0]);
                acc[mb][1] = mfma(w1, a, acc[mb][1]);
            }
        }
        // store_nhwc32's forward expression with the biases held in registers
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int nb = 0; nb < 2; nb++) {
                const long idx = padded(ring.b, y0 + wave, ring.x0 + mb * 16 + m, H, W) * CH + nb * 16 + kg * 4;
                bf16x4 o;
#pragma unroll
                for (int i = 0; i < 4; i++) o[i] = (__bf16)fmaxf(acc[mb][nb][i] + bv[nb][i], 0.0f);
                *(bf16x4 *)(out + idx) = o;
            }
    });
}

// Layers 1 and 2 in one march (inference, one image): the rows of layer 1's
// output that layer 2's window needs are formed in the LDS ring from the
// feature planes as the march reaches them (k_conv_first's operands and
// expressions, zero outside the image as its padded output holds), instead
// of being stored and read back: 12 bytes per pixel read instead of
// 12 + 64 written + 64 read.  Layer 1 covers 80 pixels (five 16-pixel
// blocks from x0 - 8) of which the ring keeps x0 - 1 .. x0 + 64.  The next
// block's feature rows are in flight into registers while layer 2 works.
// The same values as k_conv_first followed by k_conv32.
template <int ROWS> struct FirstRows {
    static constexpr int FX = 82;                       // feature columns x0 - 9 .. x0 + 72
    static constexpr int FY = ROWS + 2;                 // feature rows of ROWS layer-1 rows
    static constexpr int N = NFEAT * FY * FX;
    static constexpr int PER = (N + ROWS * 64 - 1) / (ROWS * 64);
    const float *pl;
    int H, W, x0;
    // feature f at row y, column x0 - 9 + px (zero outside the image)
    __device__ __forceinline__ float load(int c, int y_first) const
    {
        const int f = c / (FY * FX), rem = c - f * (FY * FX), r = rem / FX, px = rem - r * FX;
        const int y = y_first - 1 + r, x = x0 - 9 + px;
        return (y >= 0 && y < H && x >= 0 && x < W) ? pl[(long)f * H * W + (long)y * W + x] : 0.0f;
    }
    __device__ __forceinline__ void prefetch(float (&pf)[PER], int y_first) const
    {
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int c = threadIdx.x + j * ROWS * 64;
            if (c < N) pf[j] = load(c, y_first);
        }
    }
    __device__ __forceinline__ void commit(float *fb, const float (&pf)[PER]) const
    {
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int c = threadIdx.x + j * ROWS * 64;
            if (c < N) fb[c] = pf[j];
        }
    }
};

template <int ROWS>
__global__ __launch_bounds__(ROWS * 64) void k_first_conv32_march(const float *__restrict__ planes, int npl,
                                                                  __bf16 *__restrict__ out,
                                                                  const bf16x8 *__restrict__ frag1,
                                                                  const float *__restrict__ bias1,
                                                                  const bf16x8 *__restrict__ frag2,
                                                                  const float *__restrict__ bias2, int H, int W,
                                                                  int run_rows)
{
    using FR = FirstRows<ROWS>;
    __shared__ __attribute__((aligned(16))) __bf16 win[(ROWS + 2) * WIN_X * LDS_PIX];
    __shared__ float fb[FR::N];
    __shared__ bf16x8 wsh1[2 * 64], wsh2[TAPS * 2 * 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kg = lane >> 4;
    const int x0 = blockIdx.x * 64, b = blockIdx.z;
    const FR fr{planes + (long)b * npl * H * W, H, W, x0};
    const MarchRing<ROWS> ring{win, nullptr, b, x0, H, W};
    for (int i = threadIdx.x; i < TAPS * 2 * 64; i += ROWS * 64) wsh2[i] = frag2[i];
    for (int i = threadIdx.x; i < 2 * 64; i += ROWS * 64) wsh1[i] = frag1[i];
    float bv1[2][4], bv2[2][4];
#pragma unroll
    for (int nb = 0; nb < 2; nb++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            bv1[nb][i] = bias1[nb * 16 + kg * 4 + i];
            bv2[nb][i] = bias2[nb * 16 + kg * 4 + i];
        }
    // this lane's 8 im2col K-values: offsets into fb relative to the pixel (k_conv_first's)
    int off[8];
    bool real[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int k = kg * 8 + i, t = k / 3, f = k - 3 * t;
        real[i] = k < K1;
        off[i] = real[i] ? (f * FR::FY + (t / 3 - 1)) * FR::FX + (t % 3 - 1) : 0;
    }
    // layer-1 rows y_first .. y_first + n - 1 (wave w: row y_first + w) into the ring from the
    // features in fb (rows y_first - 1 .. y_first + n)
    auto form = [&](int y_first, int n) {
        if (wave >= n) return;
        const int y = y_first + wave;
        const bool row_in = y >= 0 && y < H;
        const float *base = fb + (wave + 1) * FR::FX + 9;   // pixel x0 + 0 of row y
        const bf16x8 w0 = wsh1[lane], w1 = wsh1[64 + lane];
#pragma unroll
        for (int mb = 0; mb < 5; mb++) {
            const int px = mb * 16 + m - 8;   // relative to x0
            bf16x8 a;
#pragma unroll
            for (int i = 0; i < 8; i++) a[i] = (__bf16)(real[i] ? base[px + off[i]] : 0.0f);
            const floatx4 z = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
            const floatx4 acc[2] = {mfma(w0, a, z), mfma(w1, a, z)};
            if (px >= -1 && px <= 64) {
                const bool in = row_in && x0 + px >= 0 && x0 + px < W;
#pragma unroll
                for (int nb = 0; nb < 2; nb++) {
                    bf16x4 o;
#pragma unroll
                    for (int i = 0; i < 4; i++) o[i] = in ? (__bf16)fmaxf(acc[nb][i] + bv1[nb][i], 0.0f) : (__bf16)0.0f;
                    *(bf16x4 *)(win + (MarchRing<ROWS>::slot(y) * WIN_X + 1 + px) * LDS_PIX + nb * 16 + kg * 4) = o;
                }
            }
        }
    };
    const int ya = blockIdx.y * run_rows, yb = ya + run_rows < H ? ya + run_rows : H;
    float pf[FR::PER];
    // prologue: layer-1 rows ya - 1 .. ya + ROWS
    fr.prefetch(pf, ya - 1);
    fr.commit(fb, pf);
    __syncthreads();
    form(ya - 1, ROWS);
    __syncthreads();
    fr.prefetch(pf, ya - 1 + ROWS);
    fr.commit(fb, pf);
    __syncthreads();
    form(ya - 1 + ROWS, 2);
    for (int y0 = ya; y0 < yb; y0 += ROWS) {
        const bool more = y0 + ROWS < yb;
        if (more) fr.prefetch(pf, y0 + ROWS + 1);   // features of the next block's layer-1 rows
        __syncthreads();                            // layer-1 rows y0-1 .. y0+ROWS in the ring
        floatx4 acc[4][2];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) acc[mb][0] = acc[mb][1] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < TAPS; t++) {
            const int dy = t / 3 - 1, dx = t % 3 - 1;
            const bf16x8 w0 = wsh2[(t * 2 + 0) * 64 + lane];
            const bf16x8 w1 = wsh2[(t * 2 + 1) * 64 + lane];
#pragma unroll
            for (int mb = 0; mb < 4; mb++) {
                const bf16x8 a = ring.at(y0 + wave + dy, mb * 16 + m + dx, kg);
                acc[mb][0] = mfma(w0, a, acc[mb][0]);
                acc[mb][1] = mfma(w1, a, acc[mb][1]);
            }
        }
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int nb = 0; nb < 2; nb++) {
                const long idx = padded(b, y0 + wave, x0 + mb * 16 + m, H, W) * CH + nb * 16 + kg * 4;
                bf16x4 o;
#pragma unroll
                for (int i = 0; i < 4; i++) o[i] = (__bf16)fmaxf(acc[mb][nb][i] + bv2[nb][i], 0.0f);
                *(bf16x4 *)(out + idx) = o;
            }
        __syncthreads();   // layer-1 rows y0-1 .. y0+ROWS-2 and the features no longer read
        if (more) {
            fr.commit(fb, pf);
            __syncthreads();
            form(y0 + ROWS + 1, ROWS);
        }
    }
}

template <int ROWS>
__global__ __launch_bounds__(ROWS * 64) void k_conv_out_march(const __bf16 *__restrict__ a5,
                                                              const bf16x8 *__restrict__ frag,
                                                              const float *__restrict__ b6, float *__restrict__ out,
                                                              int H, int W, int run_rows)
{
    __shared__ __attribute__((aligned(16))) __bf16 win[(ROWS + 2) * WIN_X * LDS_PIX];
    __shared__ bf16x8 wsh[TAPS * 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kg = lane >> 4;
    for (int i = threadIdx.x; i < TAPS * 64; i += ROWS * 64) wsh[i] = frag[i];
    const float bv = b6[0];
    conv_march<ROWS>(a5, win, H, W, run_rows, [&](const MarchRing<ROWS> &ring, int y0) {
        floatx4 acc[4];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) acc[mb] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < TAPS; t++) {
            const int dy = t / 3 - 1, dx = t % 3 - 1;
            const bf16x8 w = wsh[t * 64 + lane];
#pragma unroll
            for (int mb = 0; mb < 4; mb++) acc[mb] = mfma(w, ring.at(y0 + wave + dy, mb * 16 + m + dx, kg), acc[mb]);
        }
        if (kg == 0) {
#pragma unroll
            for (int mb = 0; mb < 4; mb++)
                out[((long)ring.b * H + y0 + wave) * W + ring.x0 + mb * 16 + m] = acc[mb][0] + bv;
        }
    });
}

// ---------------------------------------------------------------- loss and backward

// Loss over the cells M at least `margin` from the crop edge, e = out - target:
//   MSE:    sum_M e^2 / |M|,            g6 = 2 e / |M|
//   energy: sum_M e (A_M e) / |M|,      g6 = 2 (A_M e) / |M|
// where A_M e = e + sum over the faces between two cells of M of
// k_face (e_i - e_j), k_face = Kx (left face, plane 4) or Ky (bottom face,
// plane 5): the error's squared norm in the timestep operator I - dt div K grad,
// restricted to M.  One FP64 partial sum per workgroup.
__device__ __forceinline__ float err_at(const float *out, const float *tgt, long p, long q) { return out[p] - tgt[q]; }

__global__ __launch_bounds__(BLOCK) void k_loss(const float *__restrict__ out, const float *__restrict__ samples,
                                                float *__restrict__ g6, double *__restrict__ part, int B, int H,
                                                int W, int margin, float scale, int energy)
{
    __shared__ double lds[BLOCK / 64];
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    double s = 0.0;
    if (p < (long)B * H * W) {
        const int x = (int)(p % W), y = (int)((p / W) % H), b = (int)(p / ((long)W * H));
        const bool in = x >= margin && x < W - margin && y >= margin && y < H - margin;
        const long hw = (long)H * W, q = (long)y * W + x;
        const float *smp = samples + (long)b * NPLANES * hw;
        const float *tgt = smp + 3 * hw;
        const float *pout = out + (long)b * hw;
        const float e = pout[q] - tgt[q];
        float ae = e;
        if (energy && in) {
            const float *kx = smp + 4 * hw, *ky = smp + 5 * hw;
            if (x - 1 >= margin) ae += kx[q] * (e - err_at(pout, tgt, q - 1, q - 1));
            if (x + 1 < W - margin) ae += kx[q + 1] * (e - err_at(pout, tgt, q + 1, q + 1));
            if (y - 1 >= margin) ae += ky[q] * (e - err_at(pout, tgt, q - W, q - W));
            if (y + 1 < H - margin) ae += ky[q + W] * (e - err_at(pout, tgt, q + W, q + W));
        }
        g6[padded(b, y, x, H, W)] = in ? scale * ae : 0.0f;
        if (in) s = (double)e * ae;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}

// Layer 6 backward to its input: g5 = bf16(relu'(a5) * sum_tap g6[p - tap] W6[tap]).
__global__ void k_out_bwd(const float *__restrict__ g6, const float *__restrict__ w6, const __bf16 *__restrict__ a5,
                          __bf16 *__restrict__ g5, int B, int H, int W)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= (long)B * H * W) return;
    const int x = (int)(p % W), y = (int)((p / W) % H), b = (int)(p / ((long)W * H));
    float s[CH];
#pragma unroll
    for (int c = 0; c < CH; c++) s[c] = 0.0f;
    for (int t = 0; t < TAPS; t++) {
        const float g = g6[padded(b, y - (t / 3 - 1), x - (t % 3 - 1), H, W)];
#pragma unroll
        for (int c = 0; c < CH; c++) s[c] += g * w6[t * CH + c];
    }
    const long base = padded(b, y, x, H, W) * CH;
    const bf16x8 *am = (const bf16x8 *)(a5 + base);
    bf16x8 *dst = (bf16x8 *)(g5 + base);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const bf16x8 a = am[q];
        bf16x8 v;
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = (__bf16)((float)a[i] > 0.0f ? s[q * 8 + i] : 0.0f);
        dst[q] = v;
    }
}

// Layer 6 weight and bias gradients: thread j < 288 owns W6[tap][ci], thread
// 288 owns b6; workgroup c sums pixels [c*n/chunks, (c+1)*n/chunks).
constexpr int OUT_WG_THREADS = 320;
__global__ __launch_bounds__(OUT_WG_THREADS) void k_out_wgrad(const __bf16 *__restrict__ a5,
                                                               const float *__restrict__ g6, float *__restrict__ part,
                                                               int B, int H, int W, int chunks)
{
    const int j = threadIdx.x;
    if (j > TAPS * CH) return;
    const long n = (long)B * H * W;
    const long p0 = n * blockIdx.x / chunks, p1 = n * (blockIdx.x + 1) / chunks;
    const int t = j / CH, ci = j % CH;
    float acc = 0.0f;
    for (long p = p0; p < p1; p++) {
        const int x = (int)(p % W), y = (int)((p / W) % H), b = (int)(p / ((long)W * H));
        const float g = g6[padded(b, y, x, H, W)];
        if (j == TAPS * CH) acc += g;
        else acc += (float)a5[padded(b, y + t / 3 - 1, x + t % 3 - 1, H, W) * CH + ci] * g;
    }
    part[(long)blockIdx.x * (TAPS * CH + 1) + j] = acc;
}

// Weight gradient of a 32-channel layer: dW[tap][ci][co] = sum_p in[p + tap][ci] * g[p][co],
// pixels in groups of 32 along a row (K of one MFMA).  Grid (NTAP, chunks),
// 4 waves per workgroup taking every 4th group of the workgroup's range; the
// workgroup's 32x32 sum goes to part[chunk][tap][ci][co].
template <int NTAP>
__global__ __launch_bounds__(BLOCK) void k_wgrad32(const __bf16 *__restrict__ in, const __bf16 *__restrict__ g,
                                                   float *__restrict__ part, int B, int H, int W, int chunks)
{
    __shared__ float lds[4][CH * CH];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kg = lane >> 4;
    const int t = blockIdx.x, c = blockIdx.y;
    const int dy = NTAP == 1 ? 0 : t / 3 - 1, dx = NTAP == 1 ? 0 : t % 3 - 1;
    const long groups = (long)B * H * W / 32;
    const long g0 = groups * c / chunks, g1 = groups * (c + 1) / chunks;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++) acc[i][0] = acc[i][1] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    for (long gi = g0 + wave; gi < g1; gi += 4) {
        const long p0 = gi * 32;
        const int b = (int)(p0 / ((long)W * H));
        const int rem = (int)(p0 % ((long)W * H));
        const int y = rem / W, xg = rem % W;
        bf16x8 av[2], gv[2];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int xk = xg + kg * 8 + i;
            const long pa = padded(b, y + dy, xk + dx, H, W) * CH;
            const long pg = padded(b, y, xk, H, W) * CH;
            av[0][i] = in[pa + m];
            av[1][i] = in[pa + 16 + m];
            gv[0][i] = g[pg + m];
            gv[1][i] = g[pg + 16 + m];
        }
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
            acc[cb][0] = mfma(av[cb], gv[0], acc[cb][0]);
            acc[cb][1] = mfma(av[cb], gv[1], acc[cb][1]);
        }
    }
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
        for (int nb = 0; nb < 2; nb++)
#pragma unroll
            for (int i = 0; i < 4; i++) lds[wave][(cb * 16 + kg * 4 + i) * CH + nb * 16 + m] = acc[cb][nb][i];
    __syncthreads();
    float *dst = part + ((long)c * NTAP + t) * CH * CH;
    for (int e = threadIdx.x; e < CH * CH; e += BLOCK) dst[e] = lds[0][e] + lds[1][e] + lds[2][e] + lds[3][e];
}

// Bias gradient of a 32-channel layer: per-workgroup channel sums of g.
__global__ __launch_bounds__(BLOCK) void k_chansum(const __bf16 *__restrict__ g, float *__restrict__ part, int B,
                                                   int H, int W, int chunks)
{
    __shared__ float lds[BLOCK];
    const int co = threadIdx.x % CH, sub = threadIdx.x / CH;
    const long n = (long)B * H * W;
    const long p0 = n * blockIdx.x / chunks, p1 = n * (blockIdx.x + 1) / chunks;
    float acc = 0.0f;
    for (long p = p0 + sub; p < p1; p += BLOCK / CH) {
        const int x = (int)(p % W), y = (int)((p / W) % H), b = (int)(p / ((long)W * H));
        acc += (float)g[padded(b, y, x, H, W) * CH + co];
    }
    lds[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < CH) {
        float s = 0.0f;
        for (int k = 0; k < BLOCK / CH; k++) s += lds[k * CH + threadIdx.x];
        part[(long)blockIdx.x * CH + threadIdx.x] = s;
    }
}

// dst[e] = sum_c part[c*stride + e], e < n (fixed order).
__global__ void k_sum_parts(const float *__restrict__ part, long stride, int chunks, int n, float *__restrict__ dst)
{
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= n) return;
    float s = 0.0f;
    for (int c = 0; c < chunks; c++) s += part[(long)c * stride + e];
    dst[e] = s;
}

// Adam (Kingma and Ba, 2015) on the FP32 master parameters.
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= n) return;
    const float ge = g[e];
    const float me = b1 * m[e] + (1.0f - b1) * ge;
    const float ve = b2 * v[e] + (1.0f - b2) * ge * ge;
    m[e] = me;
    v[e] = ve;
    p[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
}

// Training samples from dense FP64 interior fields (row-major ny x nx):
// sample k = crop of size `crop` at origins[2k], origins[2k+1] (x, y), planes
//   0: (log10 rho - 0.5) / 1.5
//   1: (u0 - u_prev) / s
//   2: u0 / s_u
//   3: target ((u_next - u0) - c0 (u0 - u_prev) - c1 (u_prev - u_prev2)) / s
//   4, 5: Kx (left face), Ky (bottom face) of the timestep operator
__global__ void k_make_samples(const double *__restrict__ rho, const double *__restrict__ u_prev2,
                               const double *__restrict__ u_prev, const double *__restrict__ u0,
                               const double *__restrict__ u_next, const double *__restrict__ kx,
                               const double *__restrict__ ky, int nx, int ny, double c0, double c1, double inv_s,
                               double inv_su, const int *__restrict__ origins, int n, int crop,
                               float *__restrict__ out)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    const long per = (long)crop * crop;
    if (p >= per * n) return;
    const int k = (int)(p / per);
    const int cy = (int)((p % per) / crop), cx = (int)(p % crop);
    const int ox = origins[2 * k], oy = origins[2 * k + 1];
    const int x = ox + cx, y = oy + cy;
    float *o = out + (long)k * NPLANES * per + (long)cy * crop + cx;
    if (x < 0 || x >= nx || y < 0 || y >= ny) {
        for (int q = 0; q < NPLANES; q++) o[q * per] = 0.0f;
        return;
    }
    const long i = (long)y * nx + x;
    const double dprev = u0[i] - u_prev[i];
    o[0] = (float)((log10(rho[i]) - 0.5) / 1.5);
    o[per] = (float)(dprev * inv_s);
    o[2 * per] = (float)(u0[i] * inv_su);
    o[3 * per] = (float)(((u_next[i] - u0[i]) - c0 * dprev - c1 * (u_prev[i] - u_prev2[i])) * inv_s);
    o[4 * per] = (float)kx[i];
    o[5 * per] = (float)ky[i];
}

// sums of (u0 - u_prev)^2 and u0^2 over dense fields, one FP64 pair per workgroup
__global__ __launch_bounds__(BLOCK) void k_scales(const double *__restrict__ u_prev, const double *__restrict__ u0,
                                                  long n, double *__restrict__ part)
{
    __shared__ double lds[2][BLOCK / 64];
    double a = 0.0, c = 0.0;
    for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK) {
        const double d = u0[i] - u_prev[i];
        a += d * d;
        c += u0[i] * u0[i];
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        lds[0][threadIdx.x >> 6] = a;
        lds[1][threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = lds[0][0] + lds[0][1] + lds[0][2] + lds[0][3];
        part[2 * blockIdx.x + 1] = lds[1][0] + lds[1][1] + lds[1][2] + lds[1][3];
    }
}

// ---------------------------------------------------------------- host side

static unsigned nblocks(long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

hipError_t net_alloc(Net &net)
{
    hipError_t e;
    if ((e = hipMalloc(&net.params, NPARAM * sizeof(float))) != hipSuccess) return e;
    if ((e = hipMalloc(&net.packed, PK_TOTAL * sizeof(__bf16))) != hipSuccess) return e;
    return hipMalloc(&net.w6, TAPS * CH * sizeof(float));
}

void net_free(Net &net)
{
    if (net.params) (void)hipFree(net.params);
    if (net.packed) (void)hipFree(net.packed);
    if (net.w6) (void)hipFree(net.w6);
    net = Net();
}

void net_pack(const Net &net, hipStream_t stream)
{
    hipLaunchKernelGGL(k_pack, dim3(nblocks(PK_TOTAL)), dim3(BLOCK), 0, stream, net.params, net.packed, net.w6);
}

hipError_t net_upload(Net &net, const float *host_params, hipStream_t stream)
{
    hipError_t e = hipMemcpyAsync(net.params, host_params, NPARAM * sizeof(float), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return e;
    net_pack(net, stream);
    return hipStreamSynchronize(stream);
}

hipError_t acts_alloc(Acts &a, int B, int H, int W, bool training)
{
    a.B = B;
    a.H = H;
    a.W = W;
    const size_t bytes = act_bytes(B, H, W);
    const int distinct = training ? 5 : 2;
    hipError_t e;
    if (training) {   // the im2col image is only needed for layer 1's weight gradient
        if ((e = hipMalloc(&a.x1, bytes)) != hipSuccess) return e;
        if ((e = hipMemset(a.x1, 0, bytes)) != hipSuccess) return e;
    }
    for (int l = 1; l <= distinct; l++) {
        if ((e = hipMalloc(&a.a[l], bytes)) != hipSuccess) return e;
        if ((e = hipMemset(a.a[l], 0, bytes)) != hipSuccess) return e;
    }
    for (int l = distinct + 1; l <= 5; l++) a.a[l] = a.a[l - 2];  // inference: ping-pong
    return hipMalloc(&a.out, (size_t)B * H * W * sizeof(float));
}

void acts_free(Acts &a)
{
    if (a.x1) (void)hipFree(a.x1);
    for (int l = 1; l <= 5; l++) {
        bool alias = false;
        for (int k = 1; k < l; k++) alias |= a.a[k] == a.a[l];
        if (a.a[l] && !alias) (void)hipFree(a.a[l]);
    }
    if (a.out) (void)hipFree(a.out);
    a = Acts();
}

void im2col(const Acts &a, const float *planes, int npl, hipStream_t stream)
{
    hipLaunchKernelGGL(k_im2col, dim3(nblocks((long)a.B * a.H * a.W)), dim3(BLOCK), 0, stream, planes, npl, a.x1,
                       a.B, a.H, a.W);
}

// Rows per run of the marching layers: about one workgroup (95 KB of LDS)
// per compute unit, a multiple of rows.
int march_run_rows(int H, int W, int B, int rows)
{
    static int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1)
            n = 256;
        return n;
    }();
    const long bands = (long)(W / 64) * B;
    const long runs = bands >= cus ? 1 : cus / bands;
    const long per = (H + runs - 1) / runs;
    return (int)((per + rows - 1) / rows * rows);
}

// The marching layers (k_first_conv32_march, k_conv32_march,
// k_conv_out_march) for one image where the height is a multiple of 16,
// unless TEA_WS_MARCH=0.
template <int ROWS> void forward_rows(const Net &net, const Acts &a, const float *planes, int npl, hipStream_t stream)
{
    const dim3 grid(a.W / 64, a.H / ROWS, a.B), block(ROWS * 64);
    const bf16x8 *pk = (const bf16x8 *)net.packed;
    const char *env = getenv("TEA_WS_MARCH");
    const bool march = ROWS == 16 && a.B == 1 && !(env && env[0] == '0');
    int run_rows = march_run_rows(a.H, a.W, a.B, ROWS);
    // TEA_WS_MARCH_ROWS=<rows> sets the run length (tests march several blocks on small images)
    if (const char *s = getenv("TEA_WS_MARCH_ROWS")) {
        const int v = atoi(s);
        if (v > 0) run_rows = (v + ROWS - 1) / ROWS * ROWS;
    }
    const dim3 mgrid(a.W / 64, (a.H + run_rows - 1) / run_rows, a.B);
    if (march)   // layers 1 and 2 in one march (layer 1's output stays in LDS)
        hipLaunchKernelGGL(k_first_conv32_march<ROWS>, mgrid, block, 0, stream, planes, npl, a.a[2], pk + PK_L1 / 8,
                           net.params + OFF_B1, pk + pk_fwd(2) / 8, net.params + off_b(2), a.H, a.W, run_rows);
    else
        hipLaunchKernelGGL(k_conv_first<ROWS>, grid, block, 0, stream, planes, npl, a.a[1], pk + PK_L1 / 8,
                           net.params + OFF_B1, a.H, a.W);
    for (int l = march ? 3 : 2; l <= 5; l++) {
        if (march)
            hipLaunchKernelGGL(k_conv32_march<ROWS>, mgrid, block, 0, stream, a.a[l - 1], a.a[l], pk + pk_fwd(l) / 8,
                               net.params + off_b(l), a.H, a.W, run_rows);
        else
            hipLaunchKernelGGL((k_conv32<ROWS, false>), grid, block, 0, stream, a.a[l - 1], a.a[l],
                               pk + pk_fwd(l) / 8, net.params + off_b(l), (const __bf16 *)nullptr, a.H, a.W);
    }
    if (march)
        hipLaunchKernelGGL(k_conv_out_march<ROWS>, mgrid, block, 0, stream, a.a[5], pk + PK_OUT / 8,
                           net.params + OFF_B6, a.out, a.H, a.W, run_rows);
    else
        hipLaunchKernelGGL(k_conv_out<ROWS>, grid, block, 0, stream, a.a[5], pk + PK_OUT / 8, net.params + OFF_B6,
                           a.out, a.H, a.W);
}

// Inference takes workgroups of 16 rows where the height allows it: a smaller
// share of re-read halo rows (18/16 against 10/8 per staged window).
void forward(const Net &net, const Acts &a, const float *planes, int npl, hipStream_t stream)
{
    if (a.H % 16 == 0) forward_rows<16>(net, a, planes, npl, stream);
    else if (conv_rows(a.H) == 8) forward_rows<8>(net, a, planes, npl, stream);
    else forward_rows<4>(net, a, planes, npl, stream);
}

}  // namespace ws

// =============================================================================
// Trainer C-ABI
// =============================================================================
namespace {

char g_ws_error[512] = "";

int ws_fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_ws_error, sizeof g_ws_error, fmt, ap);
    va_end(ap);
    return -1;
}

#define WS_OK(call)                                                                     \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) return ws_fail("%s: %s", #call, hipGetErrorString(e_));  \
    } while (0)

constexpr int WG_CHUNKS = 256;      // workgroups per weight-gradient reduction

struct Trainer {
    int B = 0, H = 0, W = 0, margin = 0, energy = 0;
    long step = 0;
    hipStream_t stream = nullptr;
    ws::Net net;
    ws::Acts acts;
    float *g6 = nullptr;            // padded FP32 gradient of the output layer
    __bf16 *g[2] = {};              // padded NHWC32 gradients (ping-pong)
    float *grad = nullptr, *m = nullptr, *v = nullptr;
    float *part = nullptr;          // weight-gradient partials
    double *loss_part = nullptr;
    int loss_blocks = 0;
};

void destroy(Trainer *t)
{
    if (!t) return;
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    ws::net_free(t->net);
    ws::acts_free(t->acts);
    for (void *p : {(void *)t->g6, (void *)t->g[0], (void *)t->g[1], (void *)t->grad, (void *)t->m, (void *)t->v,
                    (void *)t->part, (void *)t->loss_part})
        if (p) (void)hipFree(p);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
}

int forward_pass(Trainer *t, const float *samples)
{
    ws::im2col(t->acts, samples, ws::NPLANES, t->stream);
    ws::forward(t->net, t->acts, samples, ws::NPLANES, t->stream);
    WS_OK(hipGetLastError());
    return 0;
}

// loss, then every gradient into t->grad (same layout as the parameters)
int backward_pass(Trainer *t, const float *samples)
{
    using namespace ws;
    const int B = t->B, H = t->H, W = t->W;
    const long npix = (long)B * H * W;
    const long inner = (long)B * (H - 2 * t->margin) * (W - 2 * t->margin);
    const float scale = 2.0f / (float)inner;
    hipStream_t s = t->stream;
    hipLaunchKernelGGL(k_loss, dim3(t->loss_blocks), dim3(BLOCK), 0, s, t->acts.out, samples, t->g6, t->loss_part,
                       B, H, W, t->margin, scale, t->energy);
    // layer 6
    hipLaunchKernelGGL(k_out_wgrad, dim3(WG_CHUNKS), dim3(OUT_WG_THREADS), 0, s, t->acts.a[5], t->g6, t->part, B,
                       H, W, WG_CHUNKS);
    hipLaunchKernelGGL(k_sum_parts, dim3(nblocks(TAPS * CH + 1)), dim3(BLOCK), 0, s, t->part,
                       (long)(TAPS * CH + 1), WG_CHUNKS, TAPS * CH + 1, t->grad + OFF_W6);
    hipLaunchKernelGGL(k_out_bwd, dim3(nblocks(npix)), dim3(BLOCK), 0, s, t->g6, t->net.w6, t->acts.a[5], t->g[0],
                       B, H, W);
    // layers 5..2: weight and bias gradients from g(l), then g(l-1)
    const int rows = conv_rows(H);
    const dim3 grid(W / 64, H / rows, B), block(rows * 64);
    const bf16x8 *pk = (const bf16x8 *)t->net.packed;
    int cur = 0;
    for (int l = 5; l >= 2; l--) {
        hipLaunchKernelGGL((k_wgrad32<TAPS>), dim3(TAPS, WG_CHUNKS), dim3(BLOCK), 0, s, t->acts.a[l - 1], t->g[cur],
                           t->part, B, H, W, WG_CHUNKS);
        hipLaunchKernelGGL(k_sum_parts, dim3(nblocks(TAPS * CH * CH)), dim3(BLOCK), 0, s, t->part,
                           (long)TAPS * CH * CH, WG_CHUNKS, TAPS * CH * CH, t->grad + off_w(l));
        hipLaunchKernelGGL(k_chansum, dim3(WG_CHUNKS), dim3(BLOCK), 0, s, t->g[cur], t->part, B, H, W, WG_CHUNKS);
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(BLOCK), 0, s, t->part, (long)CH, WG_CHUNKS, CH,
                           t->grad + off_b(l));
        if (rows == 8)
            hipLaunchKernelGGL((k_conv32<8, true>), grid, block, 0, s, t->g[cur], t->g[1 - cur], pk + pk_bwd(l) / 8,
                               (const float *)nullptr, t->acts.a[l - 1], H, W);
        else
            hipLaunchKernelGGL((k_conv32<4, true>), grid, block, 0, s, t->g[cur], t->g[1 - cur], pk + pk_bwd(l) / 8,
                               (const float *)nullptr, t->acts.a[l - 1], H, W);
        cur = 1 - cur;
    }
    // layer 1 (im2col input, no input gradient)
    hipLaunchKernelGGL((k_wgrad32<1>), dim3(1, WG_CHUNKS), dim3(BLOCK), 0, s, t->acts.x1, t->g[cur], t->part, B, H,
                       W, WG_CHUNKS);
    hipLaunchKernelGGL(k_sum_parts, dim3(nblocks(K1 * CH)), dim3(BLOCK), 0, s, t->part, (long)CH * CH, WG_CHUNKS,
                       K1 * CH, t->grad + OFF_W1);
    hipLaunchKernelGGL(k_chansum, dim3(WG_CHUNKS), dim3(BLOCK), 0, s, t->g[cur], t->part, B, H, W, WG_CHUNKS);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(BLOCK), 0, s, t->part, (long)CH, WG_CHUNKS, CH, t->grad + OFF_B1);
    WS_OK(hipGetLastError());
    return 0;
}

int read_loss(Trainer *t, double *loss)
{
    if (!loss) return 0;
    std::vector<double> h(t->loss_blocks);
    WS_OK(hipMemcpyAsync(h.data(), t->loss_part, h.size() * sizeof(double), hipMemcpyDeviceToHost, t->stream));
    WS_OK(hipStreamSynchronize(t->stream));
    double s = 0.0;
    for (double x : h) s += x;
    const long inner = (long)t->B * (t->H - 2 * t->margin) * (t->W - 2 * t->margin);
    *loss = s / (double)inner;
    return 0;
}

}  // namespace

EXPORT const char *tea_ws_last_error(void) { return g_ws_error; }

EXPORT int tea_ws_param_count(void) { return ws::NPARAM; }

EXPORT void *tea_ws_trainer_create(int batch, int height, int width, int margin, int loss)
{
    if (loss != 0 && loss != 1) {
        ws_fail("tea_ws_trainer_create: loss %d is not 0 (MSE) or 1 (energy)", loss);
        return nullptr;
    }
    if (batch < 1 || !ws::shape_ok(height, width)) {
        ws_fail("tea_ws_trainer_create: need batch >= 1, height %% 4 == 0, width %% 64 == 0 (got %d, %d, %d)", batch,
                height, width);
        return nullptr;
    }
    if (margin < 0 || 2 * margin >= height || 2 * margin >= width) {
        ws_fail("tea_ws_trainer_create: margin %d leaves no pixels", margin);
        return nullptr;
    }
    Trainer *t = new Trainer();
    t->B = batch;
    t->H = height;
    t->W = width;
    t->margin = margin;
    t->energy = loss;
    const long npix = (long)batch * height * width;
    t->loss_blocks = (int)ws::nblocks(npix);
    const size_t gbytes = ws::act_bytes(batch, height, width);
    const size_t g6bytes = (size_t)batch * (height + 2) * (width + 2) * sizeof(float);
    const size_t part_floats = (size_t)WG_CHUNKS * ws::TAPS * ws::CH * ws::CH;
    bool ok = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) == hipSuccess &&
              ws::net_alloc(t->net) == hipSuccess && ws::acts_alloc(t->acts, batch, height, width, true) == hipSuccess &&
              hipMalloc(&t->g6, g6bytes) == hipSuccess && hipMemset(t->g6, 0, g6bytes) == hipSuccess &&
              hipMalloc(&t->g[0], gbytes) == hipSuccess && hipMemset(t->g[0], 0, gbytes) == hipSuccess &&
              hipMalloc(&t->g[1], gbytes) == hipSuccess && hipMemset(t->g[1], 0, gbytes) == hipSuccess &&
              hipMalloc(&t->grad, ws::NPARAM * sizeof(float)) == hipSuccess &&
              hipMalloc(&t->m, ws::NPARAM * sizeof(float)) == hipSuccess &&
              hipMalloc(&t->v, ws::NPARAM * sizeof(float)) == hipSuccess &&
              hipMalloc(&t->part, part_floats * sizeof(float)) == hipSuccess &&
              hipMalloc(&t->loss_part, t->loss_blocks * sizeof(double)) == hipSuccess &&
              hipDeviceSynchronize() == hipSuccess;
    if (!ok) {
        ws_fail("tea_ws_trainer_create: device allocation failed (%s)", hipGetErrorString(hipGetLastError()));
        destroy(t);
        return nullptr;
    }
    return t;
}

EXPORT void tea_ws_trainer_destroy(void *h) { destroy((Trainer *)h); }

EXPORT int tea_ws_trainer_set_params(void *h, const float *host_params)
{
    Trainer *t = (Trainer *)h;
    if (!t || !host_params) return ws_fail("tea_ws_trainer_set_params: null argument");
    WS_OK(ws::net_upload(t->net, host_params, t->stream));
    WS_OK(hipMemsetAsync(t->m, 0, ws::NPARAM * sizeof(float), t->stream));
    WS_OK(hipMemsetAsync(t->v, 0, ws::NPARAM * sizeof(float), t->stream));
    WS_OK(hipStreamSynchronize(t->stream));
    t->step = 0;
    return 0;
}

EXPORT int tea_ws_trainer_get_params(void *h, float *host_params)
{
    Trainer *t = (Trainer *)h;
    if (!t || !host_params) return ws_fail("tea_ws_trainer_get_params: null argument");
    WS_OK(hipMemcpyAsync(host_params, t->net.params, ws::NPARAM * sizeof(float), hipMemcpyDeviceToHost, t->stream));
    WS_OK(hipStreamSynchronize(t->stream));
    return 0;
}

EXPORT int tea_ws_trainer_forward(void *h, const float *samples, float *out_host)
{
    Trainer *t = (Trainer *)h;
    if (!t || !samples || !out_host) return ws_fail("tea_ws_trainer_forward: null argument");
    if (forward_pass(t, samples)) return -1;
    WS_OK(hipMemcpyAsync(out_host, t->acts.out, (size_t)t->B * t->H * t->W * sizeof(float), hipMemcpyDeviceToHost,
                         t->stream));
    WS_OK(hipStreamSynchronize(t->stream));
    return 0;
}

EXPORT int tea_ws_trainer_gradient(void *h, const float *samples, float *grad_host, double *loss)
{
    Trainer *t = (Trainer *)h;
    if (!t || !samples || !grad_host) return ws_fail("tea_ws_trainer_gradient: null argument");
    if (forward_pass(t, samples) || backward_pass(t, samples)) return -1;
    WS_OK(hipMemcpyAsync(grad_host, t->grad, ws::NPARAM * sizeof(float), hipMemcpyDeviceToHost, t->stream));
    WS_OK(hipStreamSynchronize(t->stream));
    return read_loss(t, loss);
}

EXPORT int tea_ws_trainer_step(void *h, const float *samples, float lr, double *loss)
{
    Trainer *t = (Trainer *)h;
    if (!t || !samples) return ws_fail("tea_ws_trainer_step: null argument");
    if (forward_pass(t, samples) || backward_pass(t, samples)) return -1;
    if (loss && read_loss(t, loss)) return -1;
    t->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)t->step), c2 = 1.0f - powf(b2, (float)t->step);
    hipLaunchKernelGGL(ws::k_adam, dim3(ws::nblocks(ws::NPARAM)), dim3(ws::BLOCK), 0, t->stream, t->net.params,
                       t->grad, t->m, t->v, ws::NPARAM, lr, b1, b2, 1e-8f, c1, c2);
    ws::net_pack(t->net, t->stream);
    WS_OK(hipGetLastError());
    return 0;
}

EXPORT int tea_ws_trainer_sync(void *h)
{
    Trainer *t = (Trainer *)h;
    if (!t) return ws_fail("tea_ws_trainer_sync: null argument");
    WS_OK(hipStreamSynchronize(t->stream));
    return 0;
}

EXPORT int tea_ws_make_samples(const double *rho, const double *u_prev2, const double *u_prev, const double *u0,
                               const double *u_next, const double *kx, const double *ky, int nx, int ny, double c0,
                               double c1, const int *origins, int n, int crop, float *out, double *scales)
{
    if (!rho || !u_prev2 || !u_prev || !u0 || !u_next || !kx || !ky || !origins || !out || !scales || nx < 1 ||
        ny < 1 || n < 1 || crop < 1)
        return ws_fail("tea_ws_make_samples: bad argument");
    const long cells = (long)nx * ny;
    const int blocks = 1024;
    double *part = nullptr;
    WS_OK(hipMalloc(&part, 2 * blocks * sizeof(double)));
    hipLaunchKernelGGL(ws::k_scales, dim3(blocks), dim3(ws::BLOCK), 0, nullptr, u_prev, u0, cells, part);
    std::vector<double> h(2 * blocks);
    hipError_t e = hipMemcpy(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(part);
    WS_OK(e);
    double a = 0.0, c = 0.0;
    for (int b = 0; b < blocks; b++) {
        a += h[2 * b];
        c += h[2 * b + 1];
    }
    scales[0] = std::sqrt(a / (double)cells);
    scales[1] = std::sqrt(c / (double)cells);
    if (!(scales[0] > 0.0) || !(scales[1] > 0.0)) return ws_fail("tea_ws_make_samples: zero field scale");
    const long total = (long)n * crop * crop;
    hipLaunchKernelGGL(ws::k_make_samples, dim3(ws::nblocks(total)), dim3(ws::BLOCK), 0, nullptr, rho, u_prev2,
                       u_prev, u0, u_next, kx, ky, nx, ny, c0, c1, 1.0 / scales[0], 1.0 / scales[1], origins, n, crop,
                       out);
    WS_OK(hipGetLastError());
    WS_OK(hipDeviceSynchronize());
    return 0;
}
