// ws_net.hpp -- the warm-start network: layout constants and the host-side
// entry points that the engine (inference on its own fields) and the trainer
// (warmstart.hip) share.
//
// Network: six 3x3 convolutions with zero "same" padding,
//   3 -> 32 (ReLU), 4 x (32 -> 32, ReLU), 32 -> 1 (linear),
// BF16 weights and activations, FP32 accumulation and biases.
//
// Parameter vector (FP32, the trainer's master copy and the weight file's
// payload), in this order:
//   W1[27][32]      k = tap*3 + input channel, tap = (dy+1)*3 + (dx+1)
//   b1[32]
//   for l = 2..5:  Wl[9][32][32] (tap, ci, co), bl[32]
//   W6[9][32]      (tap, ci)
//   b6[1]
//
// Activations are NHWC with 32 channels (64 bytes per pixel) on images padded
// by one zero pixel on every side, so the 3x3 taps need no bounds tests.
// Layer 1 reads an im2col image X1 whose 32 channels are the 27 (tap,
// feature) products of the three input features, zero-filled above 27.
#ifndef TEA_WS_NET_HPP
#define TEA_WS_NET_HPP

#include <hip/hip_runtime.h>

#include <cstddef>

namespace ws {

constexpr int CH = 32;
constexpr int TAPS = 9;
constexpr int NFEAT = 3;
constexpr int K1 = TAPS * NFEAT;                 // 27 real im2col channels
constexpr int HIDDEN = 4;                        // layers 2..5

constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + K1 * CH;
constexpr int HID_SIZE = TAPS * CH * CH + CH;    // one hidden layer's W and b
constexpr int off_w(int l) { return OFF_B1 + CH + (l - 2) * HID_SIZE; }
constexpr int off_b(int l) { return off_w(l) + TAPS * CH * CH; }
constexpr int OFF_W6 = off_w(2) + HIDDEN * HID_SIZE;
constexpr int OFF_B6 = OFF_W6 + TAPS * CH;
constexpr int NPARAM = OFF_B6 + 1;               // 38177

// Packed BF16 MFMA weight fragments (8 values per lane, 64 lanes; lane l
// holds output channel l % 16 of its block and K-values 8*(l/16) .. +7, the
// same contents as an A operand over output channels or a B operand over
// them):
//   layer 1:            [2 column blocks][64][8]
//   layer l forward:    [9 taps][2][64][8]
//   layer l backward:   [9 taps][2][64][8]  (taps flipped, ci/co swapped)
//   layer 6:            [9 taps][64][8]     (W6 as output channel 0 of 16)
constexpr int FRAG = 64 * 8;
constexpr int PK_L1 = 0;
constexpr int PK_HID = 2 * FRAG;
constexpr int pk_fwd(int l) { return PK_HID + (l - 2) * TAPS * 2 * FRAG; }
constexpr int pk_bwd(int l) { return PK_HID + (HIDDEN + l - 2) * TAPS * 2 * FRAG; }
constexpr int PK_OUT = PK_HID + 2 * HIDDEN * TAPS * 2 * FRAG;
constexpr int PK_TOTAL = PK_OUT + TAPS * FRAG;  // 79360

// Device copy of one parameter set ready for inference and training.
struct Net {
    float *params = nullptr;     // NPARAM (FP32 master values)
    __bf16 *packed = nullptr;    // PK_TOTAL
    float *w6 = nullptr;         // TAPS*CH: W6 rounded to BF16, held as float
};

// Buffers of one batch shape: B images of H x W interior pixels.
struct Acts {
    int B = 0, H = 0, W = 0;
    __bf16 *x1 = nullptr;        // im2col input, padded NHWC32 (training only)
    __bf16 *a[6] = {};           // a[1..5]: layer outputs, padded NHWC32
    float *out = nullptr;        // B*H*W network output
};

// Bytes of one padded NHWC32 BF16 image batch.
inline size_t act_bytes(int B, int H, int W) { return (size_t)B * (H + 2) * (W + 2) * CH * sizeof(__bf16); }

// Shape rules of the kernels: W a multiple of 64, H a multiple of 4.
inline bool shape_ok(int H, int W) { return H > 0 && W > 0 && W % 64 == 0 && H % 4 == 0; }

// All launches go to `stream`; none synchronises.
hipError_t net_alloc(Net &net);
void net_free(Net &net);
hipError_t net_upload(Net &net, const float *host_params, hipStream_t stream);  // then packs
void net_pack(const Net &net, hipStream_t stream);
hipError_t acts_alloc(Acts &a, int B, int H, int W, bool training);
void acts_free(Acts &a);

// Training: the im2col image x1 of the feature planes (planes[b][f][y][x],
// npl FP32 planes per image, features f = 0..2), for layer 1's weight gradient.
void im2col(const Acts &a, const float *planes, int npl, hipStream_t stream);
// Layers 1..6 from the feature planes; writes a.a[1..5] and a.out.
void forward(const Net &net, const Acts &a, const float *planes, int npl, hipStream_t stream);

}  // namespace ws

#endif
