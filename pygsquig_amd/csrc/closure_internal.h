// closure_internal.h -- kernels and device state of the learned subgrid
// closure: a 6-layer periodic 3x3 CNN (3 -> 64 -> 64 -> 64 -> 64 -> 64 -> 1
// channels, GELU) mapping (theta, u, v) on the coarse grid to the subgrid
// Jacobian flux tau.
//
// Layouts:
//   activations: NHWC bf16 (uint16_t), C = 64, pixel (b, i, j) at
//     ((b * H + i) * W + j) * 64; i is axis 0 (x) of the solver's physical
//     fields, j axis 1 (y);
//   tau, dtau: float32 (B, H, W);
//   parameters: one float32 vector, per layer W[t][ci][co] then b[co]
//     (t = 3 (dy + 1) + (dx + 1)); layer 5: W5[t][ci], b5;
//   derived per step: Wf[l][t][co][ci] and Wb[l][8 - t][ci][co] (bf16) for
//     l = 1..4, W0r, W5r (float32 holding bf16-rounded values).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cl {

constexpr int C = 64;  // channels of the hidden layers
constexpr int NW0 = 9 * 3 * C, NWM = 9 * C * C, NW5 = 9 * C;
// flat parameter offsets
constexpr int OFF_W0 = 0;
constexpr int OFF_B0 = OFF_W0 + NW0;
__host__ __device__ constexpr int off_w(int l) { return OFF_B0 + C + (l - 1) * (NWM + C); }
__host__ __device__ constexpr int off_b(int l) { return off_w(l) + NWM; }
constexpr int OFF_W5 = OFF_B0 + C + 4 * (NWM + C);
constexpr int OFF_B5 = OFF_W5 + NW5;
constexpr int NPARAM = OFF_B5 + 1;
// partial-sum record sizes of the gradient kernels
constexpr int PART_MID = NWM + C;
constexpr int PART_IN = NW0 + C;
constexpr int PART_OUT = NW5 + 1;

struct Feat {  // three float64 planes per image (theta, u, v)
  const double* plane[3];
  long long img_stride;  // doubles between images of one plane
  const int* sel;        // image index per batch slot (nullptr: slot b is image b)
  double in_scale[3];
};

struct Net {  // device views of the derived weights
  const float* P;        // flat parameters (biases read from here)
  const uint16_t* Wf;    // [4][9][64][64]
  const uint16_t* Wb;    // [4][9][64][64]
  const float* W0r;      // [9][3][64]
  const float* W5r;      // [9][64]
};

// Forward pass of a batch of B images of H x W (powers of two, W >= 64,
// H >= 4).  Writes z (optional) and a for layers 0..4 (z[l], a[l] may be
// ping-pong buffers when z is not needed), tau = out_scale (conv + b5).
hipError_t launch_forward(const Net& net, const Feat& f, int B, int H, int W, uint16_t* const z[5],
                          uint16_t* const a[5], float* tau, float out_scale, hipStream_t st);

// Loss and dtau against float32(target * tgt_scale); per-block squared-error
// sums (double) into part[B * H * W / 256].
hipError_t launch_loss(const float* tau, const double* tgt, long long img_stride, const int* sel,
                       double tgt_scale, float* dtau, double* part, int B, int H, int W,
                       hipStream_t st);

// Backward pass: gradient of the loss in every parameter into grad[NPARAM].
// Scratch: dz0/dz1 (B*H*W*64 bf16 each), part (>= part_floats(B, H, W)).
hipError_t launch_backward(const Net& net, const Feat& f, int B, int H, int W,
                           uint16_t* const z[5], uint16_t* const a[5], const float* dtau,
                           uint16_t* dz0, uint16_t* dz1, float* part, float* grad,
                           hipStream_t st);
size_t part_floats(int B, int H, int W);

// Adam step on the flat parameters, then rebuild the derived weights.
hipError_t launch_adam(float* P, const float* grad, float* m, float* v, double lr, double b1,
                       double b2, double eps, int t, hipStream_t st);
hipError_t launch_prepare(const float* P, uint16_t* Wf, uint16_t* Wb, float* W0r, float* W5r,
                          hipStream_t st);

// out = mask * (scale * in restricted to the coarse wavenumbers + add).
hipError_t launch_spectral_filter(const double2* in, int nf, double2* out, int nc, double scale,
                                  const uint8_t* mask, const double2* add, hipStream_t st);

}  // namespace cl

// Device state of one network (C ABI handle gsqg_cnn): parameters, Adam
// moments, derived weights and the activations of a batch of up to B images
// of H x W.
struct gsqg_cnn {
  int H = 0, W = 0, B = 0;
  float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;
  uint16_t *Wf = nullptr, *Wb = nullptr;
  float *W0r = nullptr, *W5r = nullptr;
  uint16_t* z[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  uint16_t* a[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  uint16_t *dz0 = nullptr, *dz1 = nullptr;
  float *tau = nullptr, *dtau = nullptr, *part = nullptr;
  double* lpart = nullptr;
  int t = 0;  // Adam steps taken
  size_t bytes = 0;
};

inline cl::Net cnn_net(const gsqg_cnn* n) {
  cl::Net r;
  r.P = n->P;
  r.Wf = n->Wf;
  r.Wb = n->Wb;
  r.W0r = n->W0r;
  r.W5r = n->W5r;
  return r;
}
