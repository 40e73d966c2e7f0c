/*
 * gsqg_closure.h -- C ABI of the learned subgrid closure.
 *
 * pygSQuiG has no closure: these entry points have no counterpart in the
 * reference.  They serve the closure workload (a coarse gSQG run whose
 * Jacobian is corrected by a CNN trained on sharp-spectrally filtered
 * high-resolution runs of this solver).
 *
 * The network: six periodic 3x3 convolutions, 3 -> 64 -> 64 -> 64 -> 64 ->
 * 64 -> 1 channels, GELU after the first five, inputs (theta, u, v) scaled
 * per channel, output the subgrid flux tau of the coarse Jacobian.  Weights
 * and activations are bfloat16, accumulation float32, optimiser Adam on
 * float32 master parameters.  Parameter vector layout (float32): per layer
 * W[t][ci][co] then b[co] (t = 3 (dy + 1) + (dx + 1) for the input pixel
 * (i + dy, j + dx)), layer 5 W5[t][ci] then b5.
 *
 * Images are (h, w) float64 planes, axis 0 first (the solver's physical
 * field layout); h, w powers of two, w >= 64, h >= 4.  All pointers named
 * *_dev are device pointers.  Calls are asynchronous on `stream` unless
 * they return a host value.
 */
#ifndef GSQG_CLOSURE_H
#define GSQG_CLOSURE_H

#include <stddef.h>
#include <stdint.h>

#include "gsqg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gsqg_cnn gsqg_cnn;

/* A batch of images: plane c of image k at theta/u/v + k * img_stride
 * (doubles); sel_dev (int32, may be NULL) picks image sel[b] for batch slot
 * b.  target (tau planes, same stride and selection) is read by the
 * training entry points only. */
typedef struct gsqg_cnn_batch {
  const double* theta;
  const double* u;
  const double* v;
  long long img_stride;
  const int32_t* sel_dev;
  int batch;
  double in_scale[3];
  const double* target;
  double tgt_scale;
} gsqg_cnn_batch;

/* Number of float32 parameters (150081). */
int gsqg_cnn_num_params(void);

/* Network and work space for batches of up to max_batch images of h x w. */
int gsqg_cnn_create(int h, int w, int max_batch, gsqg_cnn** out);
int gsqg_cnn_destroy(gsqg_cnn* net);
int gsqg_cnn_device_bytes(const gsqg_cnn* net, size_t* bytes);

/* Copy parameters in (resets the Adam moments and step count) or out. */
int gsqg_cnn_set_params(gsqg_cnn* net, const float* params_dev, void* stream);
int gsqg_cnn_get_params(const gsqg_cnn* net, float* params_dev, void* stream);

/* tau_dev[b][i][j] = out_scale * net(batch slot b), float32. */
int gsqg_cnn_forward(gsqg_cnn* net, const gsqg_cnn_batch* batch, double out_scale,
                     float* tau_dev, void* stream);

/* Gradient of loss = mean((tau - float(target * tgt_scale))^2) in every
 * parameter into grad_dev (may be NULL); *loss (may be NULL) receives the
 * loss and synchronises `stream`. */
int gsqg_cnn_gradient(gsqg_cnn* net, const gsqg_cnn_batch* batch, float* grad_dev, double* loss,
                      void* stream);

/* gsqg_cnn_gradient followed by one Adam step (beta1 0.9, beta2 0.999,
 * eps 1e-8, bias-corrected) at learning rate lr. */
int gsqg_cnn_train_step(gsqg_cnn* net, const gsqg_cnn_batch* batch, double lr, double* loss,
                        void* stream);

/* Sharp spectral filter of a fine (nf x nf) full complex128 spectrum onto a
 * coarse grid (nc x nc, nc <= nf, both even): for every coarse mode (i, j)
 * with mask_dev[i * nc + j] != 0, out = scale * in(same wavenumber) +
 * add_dev[i * nc + j] (add_dev may be NULL); 0 elsewhere.  With scale =
 * (nc / nf)^2 it maps the unnormalised transform of a field on the fine
 * grid to that of the filtered field on the coarse grid. */
int gsqg_spectral_filter(const void* in_dev, int nf, void* out_dev, int nc, double scale,
                         const uint8_t* mask_dev, const void* add_dev, void* stream);

/* Close a plan's dynamics with the network: tau = out_scale * net(theta,
 * u, v) of the stage state (inputs scaled by in_scale[3]) is added to the
 * physical Jacobian before its dealiased transform, so the stage right-hand
 * side gains -mask * fft2(tau).  every == 0 evaluates the network in every
 * RK4 stage; every == M (1, 2, 4, 8 or 16) evaluates it at the first stage of
 * steps 0, 1, 1 + M, 1 + 2M, ... of each gsqg_advance call (gsqg_rhs:
 * always) and holds tau until the next evaluation.  The network must be n x n
 * (the plan's grid), the plan a single realization.  net == NULL detaches.
 * The network must outlive the attachment; training it further changes the
 * closure in place. */
int gsqg_plan_attach_closure(gsqg_plan* plan, const gsqg_cnn* net, const double* in_scale,
                             double out_scale, int every);

#ifdef __cplusplus
}
#endif

#endif
