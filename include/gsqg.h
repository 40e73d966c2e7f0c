/*
 * gsqg.h -- C ABI of the MI355X-native pseudo-spectral gSQG time step.
 *
 * Every entry point returns an int status (GSQG_OK on success) and never
 * throws.  Pointers named *_dev are device pointers (hipMalloc'd memory or
 * torch CUDA tensors); full spectra are interleaved complex128, C order,
 * shape (n, n), axis 0 = kx, axis 1 = ky -- the layout of the reference's
 * `state["theta_hat"]`.  Physical fields are float64 (n, n), axis 0 = x.
 * `stream` is a hipStream_t (NULL = the legacy default stream).  All calls
 * are asynchronous on `stream` unless stated otherwise.
 *
 * A plan is not thread-safe; distinct plans are independent.  A plan holds
 * one device-resident solver state (theta and its passive scalars).
 *
 * Citations refer to pygSQuiG (dfielding14/pygSQuiG) paths.
 */
#ifndef GSQG_H
#define GSQG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSQG_OK 0
#define GSQG_ERR_ARG 1         /* invalid argument (message via gsqg_last_error) */
#define GSQG_ERR_HIP 2         /* HIP runtime error */
#define GSQG_ERR_NONFINITE 3   /* a non-finite state was produced */
#define GSQG_ERR_UNSUPPORTED 4 /* valid request outside what this build supports */

#define GSQG_SCHEME_EXPLICIT 0 /* reference RK4, hyperviscosity inside each stage */
#define GSQG_SCHEME_IF 1       /* integrating-factor RK4 (new; exact linear decay) */

#define GSQG_COMBINE_STEP 0      /* y + dt*(k1+2k2+2k3+k4)/6  (core/time_integrator.py:36) */
#define GSQG_COMBINE_MULTISTEP 1 /* y + (dt/6)*(k1+2k2+2k3+k4) (core/solver_optimized.py:62) */

#define GSQG_KEY_EXPLICIT 0 /* one caller key for the step(s) */
#define GSQG_KEY_DERIVED 1  /* per-step key = step_key(seed, step index) */

typedef struct gsqg_plan gsqg_plan;

typedef struct gsqg_plan_desc {
  int n;           /* points per side; power of two in [16, 8192] */
  double L;        /* box length (core/grid.py:72) */
  double alpha;    /* gSQG exponent, [-2, 2] (core/solver.py:91-92) */
  double nu_p;     /* solver hyperviscosity (core/solver.py:75) */
  int p;           /* 2, 4 or 8 (core/solver.py:93-94) */
  int scheme;      /* GSQG_SCHEME_* */
  int combine;     /* GSQG_COMBINE_* */
  int hv_dc_zero;  /* 1: hyperviscosity zeroed at k=0 (core/operators.py:205) */
  /* CombinedDamping(mu, kf, nu_p, p) (forcing/damping.py:49-92) */
  int damping;
  double damp_mu;
  double damp_nu_p;
  int damp_p;
  /* RingForcing(kf, dk, epsilon, amplitude) (forcing/ring_forcing.py:31-62); tau_f = 0 */
  int forcing;
  double epsilon;
  double amplitude;
  /* passive scalars (scalars/passive_scalar.py:106-132); kappa[n_scalars] */
  int n_scalars;
  const double* kappa;
  /* Host arrays computed with the reference's operation order:
   *   k1d[n]: fftfreq(n, L/n) * 2 * pi  (core/grid.py:94)
   *   mode_flags[(n/2) * n]: per half-plane mode, row = ky index, column = kx
   *     index; bit0 dealias (grid.py:102-105), bit1 |k| < kf/2
   *     (damping.py:34), bit2 forcing ring (ring_forcing.py:74-75).  Row 0
   *     holds ky = 0 in the low nibble and ky = n/2 in the high nibble.
   *   ring_modes[2 * n_ring]: (kx index, ky index <= n/2) of every mode that
   *     carries the anti-Hermitian part of theta_hat (the forcing ring). */
  const double* k1d;
  const uint8_t* mode_flags;
  const int32_t* ring_modes;
  int n_ring;
  /* Ensemble: n_members >= 1 independent realizations of this grid size,
   * advanced together (no forcing, no scalars when n_members > 1);
   * member_alpha[n_members] gives each member's gSQG exponent (NULL: alpha
   * for all).  Replaces looping gSQGSolver over realizations. */
  int n_members;
  const double* member_alpha;
} gsqg_plan_desc;

/* Library version (major*10000 + minor*100 + patch). */
int gsqg_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* gsqg_last_error(void);

/* Allocate a plan and its device work space on the current HIP device.
 * Replaces the set-up done by gSQGSolver.__init__ (core/solver.py:75-94),
 * RingForcing/CombinedDamping construction and make_grid (core/grid.py:72-107). */
int gsqg_plan_create(const gsqg_plan_desc* desc, gsqg_plan** out);
int gsqg_plan_destroy(gsqg_plan* plan);

/* Bytes of device memory held by the plan. */
int gsqg_plan_device_bytes(const gsqg_plan* plan, size_t* bytes);

/* Load theta_hat (full (n, n) complex128) into the plan's state.  The
 * Hermitian part drives the nonlinear dynamics; the anti-Hermitian part is
 * kept on the ring modes.  *offring_sq (may be NULL) receives the squared
 * norm of the anti-Hermitian part off those modes, which the state cannot
 * carry (synchronises `stream` when non-NULL).
 * Input side of gSQGSolver.step (core/solver.py:204). */
int gsqg_load_theta(gsqg_plan* plan, const void* theta_hat_dev, double* offring_sq, void* stream);

/* Write the plan's theta_hat as a full (n, n) complex128 spectrum.
 * Output side of gSQGSolver.step (core/solver.py:214). */
int gsqg_store_theta(gsqg_plan* plan, void* theta_hat_dev, void* stream);

/* Same for ensemble member `idx` of a plan with n_members > 1 (member 0 is
 * what gsqg_load_theta / gsqg_store_theta address). */
int gsqg_load_member(gsqg_plan* plan, int idx, const void* theta_hat_dev, void* stream);
int gsqg_store_member(gsqg_plan* plan, int idx, void* theta_hat_dev, void* stream);

/* Same for passive scalar `idx` (scalars/state.py:59-118 MultiScalarState.scalars). */
int gsqg_load_scalar(gsqg_plan* plan, int idx, const void* scalar_hat_dev, double* offring_sq,
                     void* stream);
int gsqg_store_scalar(gsqg_plan* plan, int idx, void* scalar_hat_dev, void* stream);

/* Advance the plan's state by n_steps RK4 steps of size dt: theta (with the
 * plan's forcing and damping), then every passive scalar with the velocity
 * of the new theta frozen over its four stages.
 * Replaces gSQGSolver.step (core/solver.py:182-214), gSQGSolverOptimized
 * .multistep (core/solver_optimized.py:168-199) and gSQGSolverWithScalars
 * .step (core/solver_with_scalars.py:131-178).
 * key_mode GSQG_KEY_EXPLICIT: `key` drives the forcing phases of every step
 * (the reference passes one key per step); GSQG_KEY_DERIVED: step s uses
 * step_key(key, step0 + s).  use_graph != 0 captures the steps between the
 * first and the last in HIP graphs (16 steps per graph, plus a one-step
 * graph for the remainder) and replays them; the graphs are kept until dt,
 * the key mode or the key changes. */
int gsqg_advance(gsqg_plan* plan, double dt, int n_steps, int key_mode, uint64_t key,
                 int64_t step0, int use_graph, void* stream);

/* Capture the HIP graphs gsqg_advance replays for (dt, key_mode, key) --
 * the 16-step graph and the one-step graph -- without running them: the
 * plan's state is unchanged, and a following gsqg_advance with the same
 * dt, key mode and key only replays.  Counterpart of the jit/lax.scan
 * compilation of gSQGSolverOptimized.multistep (core/solver_optimized.py:
 * 168-199), which the reference pays on its first call. */
int gsqg_graph_capture(gsqg_plan* plan, double dt, int key_mode, uint64_t key);

/* One RK4 step of every passive scalar of the plan with the given frozen
 * physical velocity (u_dev, v_dev: (n, n) float64).
 * Replaces MultiSpeciesEvolver.step / PassiveScalarEvolver.step
 * (scalars/passive_scalar.py:212-258, 341-376) without source terms. */
int gsqg_advance_scalars(gsqg_plan* plan, const double* u_dev, const double* v_dev, double dt,
                         void* stream);

/* rhs_dev = compute_rhs(theta_hat) with the plan's forcing/damping
 * (core/solver.py:150-180); key as in gsqg_advance with one step. */
int gsqg_rhs(gsqg_plan* plan, const void* theta_hat_dev, void* rhs_dev, uint64_t key,
             void* stream);

/* rhs_dev = compute_passive_scalar_rhs(scalar_hat, u, v, grid, kappa) with no
 * source (scalars/passive_scalar.py:63-103): fft2(-jacobian(ifft2(s), u, v))
 * - kappa k^2 s_hat, with the kappa of the plan's scalar `idx` and the caller's
 * physical (u, v).  Overwrites that scalar's device state.  The per-stage
 * right-hand side of PassiveScalarEvolver.step (:201-232) when a source term
 * is added on the host. */
int gsqg_scalar_rhs(gsqg_plan* plan, int idx, const void* scalar_hat_dev, const double* u_dev,
                    const double* v_dev, void* rhs_dev, void* stream);

/* (u, v) physical from theta_hat (core/operators.py:158-181). */
int gsqg_velocity(gsqg_plan* plan, const void* theta_hat_dev, double* u_dev, double* v_dev,
                  void* stream);

/* fft2 of a real field to a full complex spectrum (core/grid.py:110-124) and
 * ifft2(...).real (core/grid.py:127-138). */
int gsqg_fft2(gsqg_plan* plan, const double* field_dev, void* spectrum_dev, void* stream);
int gsqg_ifft2(gsqg_plan* plan, const void* spectrum_dev, double* field_dev, void* stream);

/* RingForcing.__call__(theta_hat, key) (forcing/ring_forcing.py:119-170):
 * writes the rescaled forcing spectrum into the ring modes of out_dev, a full
 * (n, n) complex128 array that the caller has zero-filled. */
int gsqg_ring_forcing(gsqg_plan* plan, const void* theta_hat_dev, uint64_t key, void* out_dev,
                      void* stream);

/* Number of entries in the plan's ring-mode list: the size of the OU state
 * of gsqg_ring_forcing_ou is 2 * n complex128 values.  The reference keeps
 * its state as a full (N, N) array (forcing/ring_forcing.py:61); only the
 * ring modes of it are ever non-zero. */
int gsqg_plan_ring_count(gsqg_plan* plan, int* n);

/* RingForcing.__call__(theta_hat, key, dt) with tau_f > 0
 * (forcing/ring_forcing.py:94-117, 144-145): the white ring noise of `key`
 * goes through the Ornstein-Uhlenbeck filter
 *     state = first ? white : decay * state + vfac * white,
 * decay = exp(-dt / tau_f), vfac = sqrt(1 - decay^2), state updated in place
 * (2 * gsqg_plan_ring_count values), and the filtered forcing is then made
 * real at DC and the Nyquist corners and rescaled to <theta f> = epsilon as
 * gsqg_ring_forcing does.  out_dev as there. */
int gsqg_ring_forcing_ou(gsqg_plan* plan, const void* theta_hat_dev, uint64_t key, double decay,
                         double vfac, void* ou_state_dev, int first, void* out_dev, void* stream);

/* gsqg_advance without graphs, with a HIP event pair around every kernel:
 * class_ms[5] / class_count[5] receive the summed duration and the number
 * of launches of the theta column, theta row, scalar column, scalar row and
 * other kernels (synchronises stream). */
int gsqg_profile_advance(gsqg_plan* plan, double dt, int n_steps, int key_mode, uint64_t key,
                         int64_t step0, double* class_ms, int* class_count, void* stream);

/* Tile-shape and radix autotuner (no counterpart in the reference, whose
 * FFTs are XLA's).  The column and row kernels are compiled in several
 * shapes, indexed 0 .. n_variants-1 with 0 the default: radix-8 schedules
 * (8 elements per lane) in the default 128-thread tile, a tile twice as
 * tall and one half as tall, a radix-16 schedule (16 elements per lane,
 * half the lanes per transform, fewer exchange passes) and a radix-4
 * schedule (4 elements per lane, twice the lanes: two waves per SIMD where
 * radix 8 has one), and for 64 <= n <= 512 a radix-4 column tile with two
 * field groups (two sets of lanes run a row's epilogue side by side, then
 * transform different prologue fields; its row tile is the radix-4 one).
 * n_variants = 3 for n <= 32 (radix-8 tiles), 6 for 64 <= n <= 512 (the
 * three tiles, radix 16, radix 4, field groups), 4 for n = 1024
 * (default and taller radix-8 tile, radix 16, radix 4), 3 above (radix 8,
 * radix 16 and radix 4; at n = 8192 column shape 2 is the radix-8 tile, a
 * radix-4 row would need 2048 lanes); gsqg_plan_shape_radix reports the
 * radix of a shape index (the row kernel's where the two differ).  gsqg_autotune times
 * gsqg_advance over n_steps steps (graphs on, explicit key 0) for every
 * (column, row) shape pair, best of `repeats` runs each, keeps the fastest
 * pair in the plan and returns it in best[2]; ms_table (n_variants^2 host
 * doubles, or NULL) receives the time of pair (c, r) at c * n_variants + r.
 * The plan's state (theta, ring side state, scalars) is saved into a
 * temporary device buffer and restored afterwards; synchronises stream.
 * gsqg_plan_set_shape / gsqg_plan_shape set and read the pair directly;
 * changing the pair synchronises the device and drops captured graphs.
 * Pairs of one radix give bit-identical states except with ring forcing,
 * whose <theta f> partial sums are then added in another order; radix-16
 * and radix-4 transforms round differently from radix-8 ones.  Both
 * differences are at the rounding level. */
int gsqg_autotune(gsqg_plan* plan, double dt, int n_steps, int repeats, int* best,
                  double* ms_table, void* stream);
int gsqg_plan_set_shape(gsqg_plan* plan, int col_shape, int row_shape);

/* Batched uneven ensembles: the plans of several grid sizes (each plan one
 * size, its members the y-slices of its launches, each with its own alpha)
 * advanced with ONE column launch and ONE row launch per RK4 stage for all
 * of them, the steps replayed as HIP graphs.  Counterpart of the
 * reference's ensemble step, one jax.pmap'd rk4_step of _compute_core_rhs
 * per member (core/multi_gpu.py:118-137), which runs members of one grid
 * size only.
 *
 * Plans must be unforced, without scalars or closure, with 64 <= n <= 1024
 * (at most 6 plans).  The batch runs every plan with its default radix-8
 * tile shapes whatever shapes the plan itself is set to, so each member's
 * state is that plan's own gsqg_advance with shape pair (0, 0) and key mode
 * GSQG_KEY_EXPLICIT, to rounding (the same arithmetic compiled into another
 * kernel may contract other multiply-adds).  The batch uses the plans' state
 * buffers: load members with gsqg_load_member first, store them after; do
 * not use the plans elsewhere while a batch call runs on them. */
typedef struct gsqg_batch gsqg_batch;
int gsqg_batch_create(gsqg_plan* const* plans, int n_plans, gsqg_batch** out);
int gsqg_batch_destroy(gsqg_batch* batch);
/* n_steps RK4 steps of every member of every plan (use_graph: replay the
 * middle steps from a captured 16-step graph). */
int gsqg_batch_advance(gsqg_batch* batch, double dt, int n_steps, int use_graph, void* stream);
int gsqg_plan_shape(const gsqg_plan* plan, int* col_shape, int* row_shape, int* n_variants);
int gsqg_plan_shape_radix(const gsqg_plan* plan, int shape, int* radix);

/* Non-finite flag raised by the last steps (synchronises stream). */
int gsqg_nonfinite(gsqg_plan* plan, int* flag, int clear, void* stream);

/* Ring-forcing <theta f> rate and scale of stage 1..4 of the last step
 * (synchronises stream); rate4/scale4 are 4-element arrays. */
int gsqg_forcing_stats(gsqg_plan* plan, double* rate4, double* scale4, void* stream);

/* Spectral diagnostics of a full (n, n) complex128 spectrum X, the sums
 * behind compute_energy_spectrum (utils/diagnostics.py:21-84),
 * compute_enstrophy (:119-141), compute_total_energy (:170-190) and
 * compute_palinstrophy (:193-216).  With |k|^2 = kx^2 + ky^2 (each square
 * rounded) and X[0,0] excluded from every density:
 *   bin_sums[b]   = sum of 0.5 |k|^(alpha-2) |X|^2 over edges[b] <= |k| < edges[b+1],
 *                   b < n_bins (edges: n_bins + 1 host doubles, edges[0] = 0);
 *   shell_sums[s] = sum of 0.5 |k|^(alpha-2) |X|^2 / n^4 over rint(|k|/dk) == s,
 *                   s <= n/2 (n/2 + 1 doubles);
 *   sums[0] = sum 0.5 (k^2)^((alpha-2)/2) |X|^2, sums[1] = sum (|k|^alpha)^2 |X|^2
 *             (k = 0 contributes 0), sums[2] = sum 0.5 (k^2)^(1+alpha) |X|^2.
 * Output pointers are host memory; synchronises `stream`. */
int gsqg_spectral_diagnostics(gsqg_plan* plan, const void* spectrum_dev, double alpha,
                              const double* edges, int n_bins, double dk, double* bin_sums,
                              double* shell_sums, double* sums, void* stream);

/* Allocation-free variant of gsqg_spectral_diagnostics: same arguments,
 * kernels and results (bit for bit), but its scratch -- per-block partial
 * sums, reduced slots, the device copy of `edges` -- lives in a file-scope
 * device array of the library (17.9 MB per GPU, held until the process
 * ends), so the call makes no hipMalloc/hipFree at all, on its first call
 * included.  n_bins must be in [1, 256] (GSQG_ERR_ARG otherwise).  Calls
 * that overlap on one GPU share the array and race; serialise them.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (tight diagnostic loops, serving loops, harnesses that
 * forbid allocation after set-up). */
int gsqg_spectral_diagnostics_noalloc(gsqg_plan* plan, const void* spectrum_dev, double alpha,
                                      const double* edges, int n_bins, double dk,
                                      double* bin_sums, double* shell_sums, double* sums,
                                      void* stream);

/* get_diagnostics (core/solver.py:216-245) of a full spectrum X with the
 * plan's alpha: out[0] = 0.5 mean(u^2 + v^2), out[1] = 0.5 mean(theta^2),
 * out[2] = max |theta|, theta = ifft2(X).real, (u, v) its velocity.  out is
 * 3 host doubles; synchronises `stream`. */
int gsqg_physical_diagnostics(gsqg_plan* plan, const void* spectrum_dev, double* out,
                              void* stream);

/* CFL controller quantities of a full (n, n) complex128 spectrum X:
 * out[0] = max sqrt(u^2 + v^2) of its velocity (compute_max_velocity,
 * core/adaptive_timestep.py:57-70; compute_velocity_from_theta,
 * core/adaptive_solver.py:106), out[1] = max |X| (the growth check of
 * AdaptiveTimestepper.check_stability, core/adaptive_timestep.py:256-262).
 * A NaN anywhere makes the maximum NaN.  out is 2 host doubles;
 * synchronises `stream`. */
int gsqg_cfl_extrema(gsqg_plan* plan, const void* spectrum_dev, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GSQG_H */
