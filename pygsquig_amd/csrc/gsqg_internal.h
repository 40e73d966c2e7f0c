// Argument blocks shared by the host launchers and the kernels.
//
// Spectral layout ("half-packed, ky-major"): a field is N/2 rows of N complex
// doubles, row j = ky index (0..N/2-1), column i = kx index in FFT order.
// Row 0 packs the two self-conjugate columns of the reference's full (N, N)
// spectrum: H[0][i] = X(i, 0) + i * X(i, N/2).  Mixed fields (x-inverse
// transformed) use the same shape with column i = x; physical fields are
// (N, N) doubles, row x, column y (the reference's C order).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gsqg {

// Row pitch (complex elements) of the fields that pass between the column
// and the row kernel -- the mixed fields and the y-transformed Jacobian.  The
// row kernel reads and writes them across rows, 16 bytes per row where its
// tile is one x-line (N >= 2048), with rows a power-of-two multiple of 32 KB
// apart; a pad of 16 elements (256 bytes) per row spreads those accesses
// over more memory channels: the N = 8192 row kernels measured 9% (theta)
// and 7% (scalars) faster with it, the step 48.6 -> 45.2 ms.  At N = 1024
// the step did not gain, so smaller grids keep pitch N.  The spectral state
// arrays keep pitch N at every size; a half-packed field slice holds N/2
// rows of the larger pitch.
#ifndef GSQG_PITCH_PAD
#define GSQG_PITCH_PAD 16
#endif
__host__ __device__ constexpr int mixed_pitch(int N) {
  return N >= 2048 ? N + GSQG_PITCH_PAD : N;
}

// Smallest grid whose scalar row launch loops over the scalars in one block
// per x-line tile (loading the frozen velocity once) instead of giving each
// scalar its own y-slice of blocks.  At N = 1024 with two scalars the loop
// measured 6% slower per forced step (0.2555 against 0.2417 ms): the longer
// chain per block overlaps less with the theta kernels running beside it.
#ifndef GSQG_SCALAR_LOOP_N
#define GSQG_SCALAR_LOOP_N 2048
#endif

enum FlagBits : unsigned {
  FLAG_DEALIAS = 1u,  // |k| < (2/3) pi N / L            (core/grid.py:102-105)
  FLAG_LSDAMP = 2u,   // |k| < kf/2                      (forcing/damping.py:34)
  FLAG_RING = 4u,     // ||k| - kf| <= dk/2              (forcing/ring_forcing.py:74-75)
};
// Row 0 of the flag array: low nibble = mode (i, 0), high nibble = (i, N/2).

// Device control block, one per plan.
struct Ctrl {
  double scale[8];   // ring-forcing scale per stage slot (1..4)
  double rate[8];    // <theta f> per stage slot
  uint64_t key;      // explicit per-call forcing key
  uint64_t seed;     // seed of derived per-step keys
  int64_t step;      // index of the step being taken
  int key_mode;      // 0: explicit key, 1: step_key(seed, step)
  int pad_;
  double nonfinite;  // set to 1 by any kernel that writes a non-finite state
};

struct Phys {
  int N;
  double inv_n2;
  double alpha;
  int alpha_kind;   // 0 generic pow, 1 alpha==1, 2 alpha==0, 3 alpha==2
  double nu;        // solver hyperviscosity
  int p;
  int damping;      // CombinedDamping present
  double mu;
  double nu_d;
  int p_d;
  int forcing;      // RingForcing present
  double amp;       // forcing amplitude
  double eps;       // target <theta F>
  int scheme;       // 0 explicit (reference), 1 integrating factor
  int combine;      // 0 dt*(k1+2k2+2k3+k4)/6, 1 (dt/6)*(...)
  int hv_dc_zero;   // 1: hyperviscosity zeroed at k=0 (reference gSQGSolver)
  double kappa;     // scalar diffusivity
  double dxn;       // (L / N) * N, the divisor of the wavenumber table k1d
};

// One mode of the anti-Hermitian ring side state with its wavenumbers and
// flags gathered once at plan creation, so the column kernel reads it in a
// single load issued before any of its stores.
struct ATab {
  double kx, ky;
  int ii, jy;  // kx index, ky index (<= N/2)
  unsigned fl;
  int dc;
};

enum CKJob { CK_THETA = 0, CK_SCALAR = 1, CK_VEL = 2, CK_FWD = 3, CK_INV = 4 };
// RK_THETA_CL: RK_THETA with a physical subgrid flux tau added to the
// Jacobian before its forward transform (learned closure).
enum RKJob { RK_THETA = 0, RK_SCALAR = 1, RK_VEL = 2, RK_R2C = 3, RK_C2R = 4, RK_THETA_CL = 5 };

// Optional start/stop events recorded at the kernel's own boundaries
// (hipExtLaunchKernelGGL), for live per-kernel timing.
struct LaunchEv {
  hipEvent_t start, stop;
};

struct CKArgs {
  const double2* J;      // mixed input (forward x-transform in the epilogue)
  double2* th;
  double2* acc;
  double2* ys;
  double2* F[4];         // prologue outputs (mixed)
  const uint8_t* flags;
  const double* k1d;
  const double2* tw;
  Ctrl* ctrl;
  double* part;          // rate partials, [8][part_stride]
  int part_stride;
  int e;                 // epilogue stage (0 = none)
  int prologue;          // write prologue fields
  int rate_slot;         // slot of the prologue state's <theta f> partials (0 = none)
  int rate_next_key;     // prologue rate uses the key of step+1
  double dt;
  Phys ph;
  const double* member_alpha;  // per-field gSQG exponent (ensembles; null: ph.alpha)
  const ATab* a_tab;     // anti-Hermitian ring side state: mode of each entry
  double2* a_val;
  int n_a;
  long long field_stride; // elements between consecutive scalars (blockIdx.y)
  double2* ystage;       // prologue state y written here when non-null (closure features)
};

struct RKArgs {
  const double2* F[4];
  const double* uphys;
  const double* vphys;
  double* uout;
  double* vout;
  double2* Jout;
  const double* phys_in;
  double* phys_out;
  const double2* tw;
  Ctrl* ctrl;
  const double* part;
  int part_stride;
  int nblk_ck;           // number of partials to reduce
  int red_slot;          // stage slot to reduce (0 = none)
  int inc_step;          // block 0 increments ctrl->step
  double eps;
  long long field_stride;
  const float* tau;      // RK_THETA_CL: physical (N, N) flux added to J
  int n_loop;            // RK_SCALAR, N >= 2048: scalars one block transforms in turn
};

// Several plans' column or row launches of one RK4 stage as one launch
// (uneven ensembles: one plan per grid size).  Block b belongs to plan g with
// blk0[g] <= b < blk0[g + 1]; within it, field slice by = (b - blk0[g]) / gx[g]
// and block bx = (b - blk0[g]) % gx[g] of that plan's own launch.
constexpr int kBatchMaxPlans = 6;
struct CKBatch {
  int n;
  int N[kBatchMaxPlans];
  int gx[kBatchMaxPlans];
  int blk0[kBatchMaxPlans + 1];
  CKArgs a[kBatchMaxPlans];
};
struct RKBatch {
  int n;
  int N[kBatchMaxPlans];
  int gx[kBatchMaxPlans];
  int blk0[kBatchMaxPlans + 1];
  RKArgs a[kBatchMaxPlans];
};

// Host-side argument checks of the launchers (gsqg_kernels.hip).
bool check_ck_args(int job, const CKArgs& a, int ny);
bool check_rk_args(int job, const RKArgs& a, int ny);

}  // namespace gsqg
