// Host side of the C ABI declared in include/gsqg.h: plan set-up, device
// buffers and the launch sequence of one RK4 step.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gsqg.h"
#include "../../include/gsqg_closure.h"
#include "closure_internal.h"
#include "gsqg_internal.h"

namespace gsqg {
hipError_t launch_ck(int N, int job, const CKArgs& a, int ny, hipStream_t s, const LaunchEv* ev,
                     int var);
hipError_t launch_rk(int N, int job, const RKArgs& a, int ny, hipStream_t s, const LaunchEv* ev,
                     int var);
int ck_num_blocks(int N, int var);
int num_shape_variants(int N);
int shape_radix(int N, int var);
hipError_t launch_full_to_half(const double2* X, double2* H, int N, double* norms, double* norm0,
                               hipStream_t s);
int full_to_half_num_blocks(int N);
hipError_t launch_half_to_full(const double2* H, double2* X, int N, hipStream_t s);
hipError_t launch_a_gather(const double2* X, const int2* modes, double2* aval, int n, int N,
                           double* norm, hipStream_t s);
hipError_t launch_a_scatter(double2* X, const int2* modes, const double2* aval, int n, int N,
                            hipStream_t s);
int diag_blocks();
hipError_t launch_spec_diag(const double2* X, const double* k1d, const double* edges, int N, int nb,
                            int nshell, double dk, double alpha, double* part, int stride,
                            hipStream_t s);
hipError_t launch_phys_diag(const double* th, const double* u, const double* v, long long n,
                            double* part, hipStream_t s);
hipError_t launch_diag_reduce(const double* part, int nblk, int stride, int nsum, double* out,
                              hipStream_t s);
hipError_t launch_cfl(const double* u, const double* v, const double2* X, long long n,
                      long long nx, double* part, hipStream_t s);
hipError_t launch_a_tab(const int2* modes, const double* k1d, const uint8_t* flags, ATab* out, int n,
                        int N, hipStream_t s);
hipError_t launch_a_rhs(const Phys& ph, const double* k1d, const uint8_t* flags, const Ctrl* ctrl,
                        const int2* modes, const double2* aval, double2* out, int n, int N,
                        hipStream_t s);
hipError_t launch_set_ctrl(Ctrl* c, int64_t step, uint64_t key, uint64_t seed, int key_mode,
                           hipStream_t s);
hipError_t launch_ring_field(const double2* X, const int2* modes, int n, int N, double amp,
                             double eps, uint64_t key, double2* out, double* rate_out,
                             double2* ou, double decay, double vfac, int ou_first,
                             hipStream_t s);
hipError_t launch_ck_batch(int e, const CKBatch& m, hipStream_t s);
hipError_t launch_rk_batch(const RKBatch& m, hipStream_t s);
bool batch_supports(int N);
int batch_blocks(int N, bool ck);
}  // namespace gsqg

using namespace gsqg;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(GSQG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

static const int kMaxScalars = 8;
static const int kGraphSteps = 16;  // steps per replayed graph (amortises the launch)
static const int kMaxBins = 4096;  // reference energy-spectrum bins per call

struct gsqg_plan {
  int N = 0;
  int device = 0;
  size_t hsz = 0;  // complex elements of one half-packed field
  Phys ph{};
  int n_scalars = 0;
  double kappa[kMaxScalars] = {0};
  int n_mem = 1;              // ensemble members (theta fields, stride hsz)
  double* malpha = nullptr;   // per-member alpha (device), n_mem > 1 only
  // theta
  double2 *th = nullptr, *th2 = nullptr, *acc = nullptr, *ys = nullptr, *J = nullptr;
  double2* F[4] = {nullptr, nullptr, nullptr, nullptr};
  // scalars (contiguous, stride hsz)
  double2 *s_th = nullptr, *s_acc = nullptr, *s_ys = nullptr, *s_J = nullptr;
  double2 *s_F0 = nullptr, *s_F1 = nullptr;
  double *uphys = nullptr, *vphys = nullptr;
  // tables
  double* k1d = nullptr;
  uint8_t* flags = nullptr;
  double2* tw = nullptr;
  Ctrl* ctrl = nullptr;
  double* part = nullptr;
  int part_stride = 0;  // >= the column kernel's block count in every shape variant
  // tile-shape variants of the column and row kernels (gsqg_autotune)
  int ck_var = 0, rk_var = 0, n_var = 1;
  int2* ring = nullptr;
  ATab* a_tab = nullptr;
  double2* a_val = nullptr;
  double2* a_tmp = nullptr;
  int n_ring = 0;
  double* norms = nullptr;
  int n_norms = 0;
  // diagnostics scratch (allocated on first use): partials, reduced slots,
  // bin edges, a physical field
  double *d_part = nullptr, *d_out = nullptr, *d_edges = nullptr, *d_phys = nullptr;
  int d_stride = 0;
  size_t bytes = 0;
  // live kernel timing (gsqg_profile_advance): events around every launch
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_class;
  size_t ev_used = 0;
  // graph cache; the scalar step of a captured step body runs as a second
  // branch (cap2) beside theta's stages 2-4
  hipStream_t cap = nullptr;
  hipStream_t cap2 = nullptr;
  // eager steps run their scalar step on this stream beside theta's
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipGraphExec_t gexec = nullptr;     // one captured step
  hipGraphExec_t gexec_k = nullptr;   // kGraphSteps captured steps
  double g_dt = 0.0;
  int g_key_mode = -1;
  uint64_t g_key = 0;
  // learned closure (gsqg_plan_attach_closure): the network, its input and
  // output scales, the stage state, its theta (mixed and physical), u, v,
  // two activation buffers and the flux tau
  const gsqg_cnn* cl_net = nullptr;
  double cl_in[3] = {1.0, 1.0, 1.0};
  double cl_out = 1.0;
  int cl_every = 0;  // 0: tau in every stage; M: from the first stage of every M-th step
  double2 *cl_ys = nullptr, *cl_f = nullptr;
  double *cl_th = nullptr, *cl_u = nullptr, *cl_v = nullptr;
  uint16_t *cl_a0 = nullptr, *cl_a1 = nullptr;
  float* cl_tau = nullptr;
};

template <typename T>
static hipError_t dalloc(gsqg_plan* p, T** ptr, size_t count) {
  size_t b = count * sizeof(T);
  if (b == 0) b = 16;
  hipError_t e = hipMalloc((void**)ptr, b);
  if (e == hipSuccess) p->bytes += b;
  return e;
}

static void free_plan(gsqg_plan* p) {
  if (!p) return;
  void* ptrs[] = {p->th,   p->th2,   p->acc,  p->ys,   p->J,     p->F[0], p->F[1], p->F[2],
                  p->F[3], p->s_th,  p->s_acc, p->s_ys, p->s_J,  p->s_F0, p->s_F1, p->uphys,
                  p->vphys, p->k1d,  p->flags, p->tw,   p->ctrl, p->part, p->ring, p->a_tab, p->a_val,
                  p->a_tmp, p->norms, p->d_part, p->d_out, p->d_edges, p->d_phys, p->malpha,
                  p->cl_ys, p->cl_f, p->cl_th, p->cl_u, p->cl_v, p->cl_a0, p->cl_a1, p->cl_tau};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  for (hipEvent_t e : p->ev_pool) (void)hipEventDestroy(e);
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->gexec_k) (void)hipGraphExecDestroy(p->gexec_k);
  if (p->cap) (void)hipStreamDestroy(p->cap);
  if (p->cap2) (void)hipStreamDestroy(p->cap2);
  if (p->side) (void)hipStreamDestroy(p->side);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_join) (void)hipEventDestroy(p->ev_join);
  delete p;
}

extern "C" int gsqg_version(void) { return 100; }

extern "C" const char* gsqg_last_error(void) { return g_err.c_str(); }

static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

extern "C" int gsqg_plan_create(const gsqg_plan_desc* d, gsqg_plan** out) {
  if (!d || !out) return fail(GSQG_ERR_ARG, "null argument");
  *out = nullptr;
  const int N = d->n;
  if (N % 2 != 0) return fail(GSQG_ERR_ARG, "N must be even, got " + std::to_string(N));
  if (!is_pow2(N) || N < 16 || N > 8192)
    return fail(GSQG_ERR_UNSUPPORTED,
                "this build supports power-of-two N in [16, 8192], got " + std::to_string(N));
  if (!(d->alpha >= -2.0 && d->alpha <= 2.0))
    return fail(GSQG_ERR_ARG, "alpha must be in [-2, 2]");
  if (d->p != 2 && d->p != 4 && d->p != 8) return fail(GSQG_ERR_ARG, "p must be 2, 4, or 8");
  if (d->damping && d->damp_p != 2 && d->damp_p != 4 && d->damp_p != 8)
    return fail(GSQG_ERR_ARG, "damping p must be 2, 4, or 8");
  const int nmem = d->n_members > 0 ? d->n_members : 1;
  if (nmem > 1 && (d->forcing || d->n_scalars > 0))
    return fail(GSQG_ERR_UNSUPPORTED, "ensembles (n_members > 1) carry no forcing and no scalars");
  if (nmem > 1 && d->member_alpha) {
    for (int m = 0; m < nmem; ++m)
      if (!(d->member_alpha[m] >= -2.0 && d->member_alpha[m] <= 2.0))
        return fail(GSQG_ERR_ARG, "member alpha must be in [-2, 2], got " +
                                      std::to_string(d->member_alpha[m]));
  }
  if (d->n_scalars < 0 || d->n_scalars > kMaxScalars)
    return fail(GSQG_ERR_UNSUPPORTED, "at most 8 passive scalars");
  if (!d->k1d || !d->mode_flags) return fail(GSQG_ERR_ARG, "k1d and mode_flags are required");
  if (d->forcing && d->n_ring > 0 && !d->ring_modes)
    return fail(GSQG_ERR_ARG, "ring_modes required");

  gsqg_plan* p = new gsqg_plan();
  p->N = N;
  HIPCHK(hipGetDevice(&p->device));
  p->hsz = (size_t)(N / 2) * (size_t)mixed_pitch(N);
  Phys& ph = p->ph;
  ph.N = N;
  ph.inv_n2 = 1.0 / ((double)N * (double)N);
  ph.alpha = d->alpha;
  ph.alpha_kind = d->alpha == 1.0 ? 1 : d->alpha == 0.0 ? 2 : d->alpha == 2.0 ? 3 : 0;
  ph.nu = d->nu_p;
  ph.p = d->p;
  ph.damping = d->damping;
  ph.mu = d->damp_mu;
  ph.nu_d = d->damp_nu_p;
  ph.p_d = d->damping ? d->damp_p : 2;
  ph.forcing = d->forcing;
  ph.amp = d->amplitude;
  ph.eps = d->epsilon;
  ph.scheme = d->scheme;
  ph.combine = d->combine;
  ph.hv_dc_zero = d->hv_dc_zero;
  ph.kappa = 0.0;
  // The register-tight kernels recompute kx from its index; the table must
  // follow the same formula, fftfreq(n, L/n) * 2 pi as k / (dx n) * 2 * pi.
  ph.dxn = (d->L / N) * N;
  for (int i = 0; i < N; ++i) {
    const int ki = i < N / 2 ? i : i - N;
    if (((double)ki / ph.dxn) * 2.0 * 3.141592653589793 != d->k1d[i])
      return fail(GSQG_ERR_ARG, "k1d[" + std::to_string(i) +
                                    "] is not k / (L/n * n) * 2 * pi (core/grid.py:94)");
  }
  p->n_scalars = d->n_scalars;
  for (int s = 0; s < d->n_scalars; ++s) p->kappa[s] = d->kappa ? d->kappa[s] : 0.0;

  hipError_t e = hipSuccess;
#define ALLOC(ptr, cnt)                                        \
  do {                                                         \
    e = dalloc(p, &(ptr), (cnt));                              \
    if (e != hipSuccess) {                                     \
      free_plan(p);                                            \
      return fail(GSQG_ERR_HIP, std::string("hipMalloc: ") +   \
                                    hipGetErrorString(e));     \
    }                                                          \
  } while (0)
  const size_t h = p->hsz;
  p->n_mem = nmem;
  const size_t hm = h * (size_t)nmem;
  ALLOC(p->th, hm);
  ALLOC(p->th2, h);
  ALLOC(p->acc, hm);
  ALLOC(p->ys, hm);
  ALLOC(p->J, hm);
  for (int q = 0; q < 4; ++q) ALLOC(p->F[q], hm);
  if (nmem > 1) ALLOC(p->malpha, (size_t)nmem);
  if (p->n_scalars > 0) {
    const size_t hs = h * p->n_scalars;
    ALLOC(p->s_th, hs);
    ALLOC(p->s_acc, hs);
    ALLOC(p->s_ys, hs);
    ALLOC(p->s_J, hs);
    ALLOC(p->s_F0, hs);
    ALLOC(p->s_F1, hs);
  }
  ALLOC(p->uphys, (size_t)N * N);
  ALLOC(p->vphys, (size_t)N * N);
  ALLOC(p->k1d, (size_t)N);
  ALLOC(p->flags, h);
  ALLOC(p->tw, (size_t)2 * 8192);
  ALLOC(p->ctrl, 1);
  p->n_var = num_shape_variants(N);
  for (int v = 0; v < p->n_var; ++v)
    p->part_stride = std::max(p->part_stride, ck_num_blocks(N, v));
  ALLOC(p->part, (size_t)8 * p->part_stride);
  p->n_ring = d->forcing ? d->n_ring : 0;
  ALLOC(p->ring, (size_t)(p->n_ring > 0 ? p->n_ring : 1));
  ALLOC(p->a_tab, (size_t)(p->n_ring > 0 ? p->n_ring : 1));
  ALLOC(p->a_val, (size_t)(p->n_ring > 0 ? p->n_ring : 1));
  ALLOC(p->a_tmp, (size_t)(p->n_ring > 0 ? p->n_ring : 1));
  p->n_norms = full_to_half_num_blocks(N) + 2;
  ALLOC(p->norms, (size_t)p->n_norms);
#undef ALLOC

  // Twiddle tables tw[T + k] = exp(-2 pi i k / T), long double on the host.
  std::vector<double2> tw(2 * 8192);
  tw[0].x = tw[0].y = tw[1].x = tw[1].y = 0.0;
  for (int T = 1; T <= 8192; T *= 2) {
    for (int k = 0; k < T; ++k) {
      long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / T;
      tw[T + k].x = (double)cosl(ang);
      tw[T + k].y = (double)sinl(ang);
    }
  }
  HIPCHK(hipMemcpy(p->tw, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p->k1d, d->k1d, (size_t)N * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p->flags, d->mode_flags, (size_t)(N / 2) * (size_t)N, hipMemcpyHostToDevice));
  if (p->n_ring > 0) {
    HIPCHK(hipMemcpy(p->ring, d->ring_modes, (size_t)p->n_ring * sizeof(int2),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemset(p->a_val, 0, (size_t)p->n_ring * sizeof(double2)));
    HIPCHK(launch_a_tab(p->ring, p->k1d, p->flags, p->a_tab, p->n_ring, N, nullptr));
  }
  Ctrl c0{};
  for (int i = 0; i < 8; ++i) c0.scale[i] = 1.0;
  HIPCHK(hipMemcpy(p->ctrl, &c0, sizeof(Ctrl), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(p->th, 0, hm * sizeof(double2)));
  if (nmem > 1) {
    std::vector<double> al(nmem, d->alpha);
    if (d->member_alpha)
      for (int m = 0; m < nmem; ++m) al[m] = d->member_alpha[m];
    HIPCHK(hipMemcpy(p->malpha, al.data(), nmem * sizeof(double), hipMemcpyHostToDevice));
  }
  if (p->n_scalars > 0) HIPCHK(hipMemset(p->s_th, 0, h * p->n_scalars * sizeof(double2)));
  HIPCHK(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&p->cap2, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming));
  HIPCHK(hipDeviceSynchronize());
  *out = p;
  return GSQG_OK;
}

extern "C" int gsqg_plan_destroy(gsqg_plan* p) {
  if (!p) return GSQG_OK;
  (void)hipDeviceSynchronize();
  free_plan(p);
  return GSQG_OK;
}

extern "C" int gsqg_plan_device_bytes(const gsqg_plan* p, size_t* bytes) {
  if (!p || !bytes) return fail(GSQG_ERR_ARG, "null argument");
  *bytes = p->bytes;
  return GSQG_OK;
}

// ----------------------------------------------------------- launch wrappers
// Kernel classes for live timing.
enum { KC_THETA_COL = 0, KC_THETA_ROW = 1, KC_SCALAR_COL = 2, KC_SCALAR_ROW = 3, KC_OTHER = 4,
       KC_COUNT = 5 };

// Live timing: a pair of pool events recorded by the launch itself at the
// kernel's start and end.
static hipError_t prof_pair(gsqg_plan* p, int cls, LaunchEv* ev) {
  for (int k = 0; k < 2; ++k) {
    if (p->ev_used >= p->ev_pool.size()) {
      hipEvent_t e;
      hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      p->ev_pool.push_back(e);
      p->ev_class.push_back(-1);
    }
    p->ev_class[p->ev_used] = k == 0 ? cls : -1;
    (k == 0 ? ev->start : ev->stop) = p->ev_pool[p->ev_used++];
  }
  return hipSuccess;
}

static hipError_t ck(gsqg_plan* p, int job, const CKArgs& a, int ny, hipStream_t st) {
  const int cls = job == CK_THETA ? KC_THETA_COL : job == CK_SCALAR ? KC_SCALAR_COL : KC_OTHER;
  if (!p->prof) return launch_ck(p->N, job, a, ny, st, nullptr, p->ck_var);
  LaunchEv ev{};
  hipError_t r = prof_pair(p, cls, &ev);
  return r != hipSuccess ? r : launch_ck(p->N, job, a, ny, st, &ev, p->ck_var);
}

static hipError_t rk(gsqg_plan* p, int job, const RKArgs& a, int ny, hipStream_t st) {
  const int cls = (job == RK_THETA || job == RK_THETA_CL) ? KC_THETA_ROW
                  : job == RK_SCALAR                         ? KC_SCALAR_ROW
                                                             : KC_OTHER;
  if (!p->prof) return launch_rk(p->N, job, a, ny, st, nullptr, p->rk_var);
  LaunchEv ev{};
  hipError_t r = prof_pair(p, cls, &ev);
  return r != hipSuccess ? r : launch_rk(p->N, job, a, ny, st, &ev, p->rk_var);
}

// ----------------------------------------------------------- arg builders
// Column-kernel arguments for theta.  On the plan's state (th == p->th) an
// ensemble's members are the y-slices of one launch (ny = n_mem).
static CKArgs ck_theta(gsqg_plan* p, double2* th, int e, int prologue, int rate_slot,
                       int rate_next, double dt) {
  CKArgs a{};
  a.J = p->J;
  a.th = th;
  a.acc = p->acc;
  a.ys = p->ys;
  for (int q = 0; q < 4; ++q) a.F[q] = p->F[q];
  a.flags = p->flags;
  a.k1d = p->k1d;
  a.tw = p->tw;
  a.ctrl = p->ctrl;
  a.part = p->part;
  a.part_stride = p->part_stride;
  a.e = e;
  a.prologue = prologue;
  a.rate_slot = p->ph.forcing ? rate_slot : 0;
  a.rate_next_key = rate_next;
  a.dt = dt;
  a.ph = p->ph;
  a.a_tab = p->a_tab;
  a.a_val = p->a_val;
  a.n_a = p->n_ring;
  a.field_stride = 0;
  a.ystage = p->cl_net ? p->cl_ys : nullptr;
  if (th == p->th && p->n_mem > 1) {
    a.field_stride = (long long)p->hsz;
    a.member_alpha = p->malpha;
  }
  return a;
}

static RKArgs rk_theta(gsqg_plan* p, int red_slot, int inc_step, bool uv_out) {
  RKArgs a{};
  for (int q = 0; q < 4; ++q) a.F[q] = p->F[q];
  a.uout = uv_out ? p->uphys : nullptr;
  a.vout = uv_out ? p->vphys : nullptr;
  a.Jout = p->J;
  a.tw = p->tw;
  a.ctrl = p->ctrl;
  a.part = p->part;
  a.part_stride = p->part_stride;
  a.nblk_ck = ck_num_blocks(p->N, p->ck_var);
  a.red_slot = p->ph.forcing ? red_slot : 0;
  a.inc_step = inc_step;
  a.eps = p->ph.eps;
  a.field_stride = (long long)p->hsz;  // ensemble members; 0 .. n_mem - 1 as y-slices
  return a;
}

// Closure flux of the current stage: theta of the stage state (stored by the
// previous column prologue), u and v from the prologue's velocity fields,
// then the network.
static int closure_tau(gsqg_plan* p, hipStream_t st) {
  const int N = p->N;
  CKArgs a = ck_theta(p, p->cl_ys, 0, 0, 0, 0, 0.0);
  a.F[0] = p->cl_f;
  a.ystage = nullptr;
  HIPCHK(ck(p, CK_INV, a, 1, st));
  RKArgs r{};
  r.F[0] = p->cl_f;
  r.phys_out = p->cl_th;
  r.tw = p->tw;
  r.ctrl = p->ctrl;
  HIPCHK(rk(p, RK_C2R, r, 1, st));
  RKArgs v{};
  v.F[0] = p->F[2];
  v.F[1] = p->F[3];
  v.uout = p->cl_u;
  v.vout = p->cl_v;
  v.tw = p->tw;
  v.ctrl = p->ctrl;
  HIPCHK(rk(p, RK_VEL, v, 1, st));
  cl::Feat f{};
  f.plane[0] = p->cl_th;
  f.plane[1] = p->cl_u;
  f.plane[2] = p->cl_v;
  f.img_stride = 0;
  f.sel = nullptr;
  for (int c = 0; c < 3; ++c) f.in_scale[c] = p->cl_in[c];
  uint16_t* const z[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  uint16_t* const act[5] = {p->cl_a0, p->cl_a1, p->cl_a0, p->cl_a1, p->cl_a0};
  HIPCHK(cl::launch_forward(cnn_net(p->cl_net), f, 1, N, N, z, act, p->cl_tau, (float)p->cl_out,
                            st));
  return GSQG_OK;
}

// Row pass of a theta stage (with the closure flux when one is attached).
// `fresh` asks for a new tau at this stage: the first stage of a step that
// evaluates the network; every stage when tau is per stage.
static int theta_row(gsqg_plan* p, RKArgs a, int ny, hipStream_t st, bool fresh) {
  if (p->cl_net) {
    if (fresh || p->cl_every == 0) {
      int rc = closure_tau(p, st);
      if (rc) return rc;
    }
    a.tau = p->cl_tau;
    HIPCHK(rk(p, RK_THETA_CL, a, 1, st));
    return GSQG_OK;
  }
  HIPCHK(rk(p, RK_THETA, a, ny, st));
  return GSQG_OK;
}

// The scalar kernels read kappa from ph.kappa; scalars with different kappa
// are launched one y-slice at a time when needed.
static bool uniform_kappa(const gsqg_plan* p) {
  for (int s = 1; s < p->n_scalars; ++s)
    if (p->kappa[s] != p->kappa[0]) return false;
  return true;
}

static int launch_scalar_ck(gsqg_plan* p, int e, int prologue, double dt, hipStream_t st) {
  CKArgs a{};
  a.flags = p->flags;
  a.k1d = p->k1d;
  a.tw = p->tw;
  a.ctrl = p->ctrl;
  a.part = p->part;
  a.part_stride = p->part_stride;
  a.e = e;
  a.prologue = prologue;
  a.dt = dt;
  a.ph = p->ph;
  a.ph.forcing = 0;
  a.ph.damping = 0;
  a.field_stride = (long long)p->hsz;
  const bool uni = uniform_kappa(p);
  const int groups = uni ? 1 : p->n_scalars;
  for (int g = 0; g < groups; ++g) {
    const size_t off = uni ? 0 : (size_t)g * p->hsz;
    a.J = p->s_J + off;
    a.th = p->s_th + off;
    a.acc = p->s_acc + off;
    a.ys = p->s_ys + off;
    a.F[0] = p->s_F0 + off;
    a.F[1] = p->s_F1 + off;
    a.F[2] = a.F[3] = nullptr;
    a.ph.kappa = p->kappa[uni ? 0 : g];
    HIPCHK(ck(p, CK_SCALAR, a, uni ? p->n_scalars : 1, st));
  }
  return GSQG_OK;
}

static int launch_scalar_rk(gsqg_plan* p, const double* u, const double* v, hipStream_t st) {
  RKArgs a{};
  a.F[0] = p->s_F0;
  a.F[1] = p->s_F1;
  a.uphys = u;
  a.vphys = v;
  a.Jout = p->s_J;
  a.tw = p->tw;
  a.ctrl = p->ctrl;
  a.part = p->part;
  a.field_stride = (long long)p->hsz;
  // Large grids: one block per x-line tile transforms every scalar against
  // the velocity it loads once (grids of N >= 2048 have blocks to spare);
  // below, each scalar is a y-slice of the launch, for parallelism.
  const bool loop = p->N >= GSQG_SCALAR_LOOP_N && p->n_scalars > 1;
  a.n_loop = loop ? p->n_scalars : 1;
  HIPCHK(rk(p, RK_SCALAR, a, loop ? 1 : p->n_scalars, st));
  return GSQG_OK;
}

// One scalar RK4 step with the frozen (u, v) in uphys/vphys; the scalar
// prologue fields of the current state must already be in s_F0/s_F1.
static int scalar_step(gsqg_plan* p, double dt, bool next_prologue, hipStream_t st,
                       const double* u = nullptr, const double* v = nullptr) {
  int rc;
  if (!u) u = p->uphys;
  if (!v) v = p->vphys;
  for (int e = 1; e <= 4; ++e) {
    if ((rc = launch_scalar_rk(p, u, v, st))) return rc;
    if ((rc = launch_scalar_ck(p, e, (e < 4) ? 1 : (next_prologue ? 1 : 0), dt, st))) return rc;
  }
  return GSQG_OK;
}

// Theta stages 2..4 and the final combine of one step (stage 1's RK is
// issued by the caller so that the scalar step can slot in after it).
static int theta_rest(gsqg_plan* p, double dt, bool derived, bool last_prologue, int last_rate,
                      hipStream_t st) {
  const int N = p->N;
  for (int e = 1; e <= 3; ++e) {
    HIPCHK(ck(p, CK_THETA, ck_theta(p, p->th, e, 1, e + 1, 0, dt), p->n_mem, st));
    int rc = theta_row(p, rk_theta(p, e + 1, 0, false), p->n_mem, st, false);
    if (rc) return rc;
  }
  HIPCHK(ck(p, CK_THETA,
                   ck_theta(p, p->th, 4, last_prologue ? 1 : 0, last_rate, derived ? 1 : 0, dt),
                   p->n_mem, st));
  return GSQG_OK;
}

// Steady-state step body: theta step s (s >= 1) and, if scalars exist, the
// scalar step s-1 slotted after theta's stage-1 row pass.  With a `side`
// stream the scalar step forks there after that row pass (which writes the
// velocity it reads) and joins `st` at the end: the two branches share no
// buffer -- the scalar kernels read uphys/vphys, the scalar arrays and the
// constant tables; theta's stages 2-4 write neither the velocity nor any
// scalar array -- so they may run concurrently.  The next step's stage-1 row
// pass, which rewrites the velocity, is ordered after the join.
static int step_body(gsqg_plan* p, double dt, bool derived, bool scalars_pending,
                     bool last_prologue, int last_rate, bool scalar_next_prologue,
                     hipStream_t st, hipStream_t side = nullptr, bool cl_eval = true) {
  int rc = theta_row(p, rk_theta(p, 1, derived ? 1 : 0, scalars_pending), p->n_mem, st, cl_eval);
  if (rc) return rc;
  if (scalars_pending && side) {
    HIPCHK(hipEventRecord(p->ev_fork, st));
    HIPCHK(hipStreamWaitEvent(side, p->ev_fork, 0));
    if ((rc = scalar_step(p, dt, scalar_next_prologue, side))) return rc;
    HIPCHK(hipEventRecord(p->ev_join, side));
    if ((rc = theta_rest(p, dt, derived, last_prologue, last_rate, st))) return rc;
    HIPCHK(hipStreamWaitEvent(st, p->ev_join, 0));
    return GSQG_OK;
  }
  if (scalars_pending) {
    if ((rc = scalar_step(p, dt, scalar_next_prologue, st))) return rc;
  }
  return theta_rest(p, dt, derived, last_prologue, last_rate, st);
}

// Whether middle step j (j = step index - 1 within a gsqg_advance call)
// evaluates the closure network: every M-th step, M dividing kGraphSteps,
// so the captured multi-step graph repeats the pattern.
static bool cl_step_evals(const gsqg_plan* p, int j) {
  return p->cl_every <= 1 || (j % p->cl_every) == 0;
}

// Capture nsteps steady-state step bodies into one executable graph.
static int capture_steps(gsqg_plan* p, int nsteps, double dt, bool derived, bool sc,
                         hipGraphExec_t* out) {
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
  int rc = GSQG_OK;
  for (int s = 0; s < nsteps && rc == GSQG_OK; ++s)
    rc = step_body(p, dt, derived, sc, true, 1, true, p->cap, p->cap2, cl_step_evals(p, s));
  hipError_t ce = hipStreamEndCapture(p->cap, &g);
  if (rc) return rc;
  if (ce != hipSuccess) return fail(GSQG_ERR_HIP, std::string("capture: ") + hipGetErrorString(ce));
  hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess)
    return fail(GSQG_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
  // upload at capture rather than at the first launch
  HIPCHK(hipGraphUpload(*out, p->cap));
  HIPCHK(hipStreamSynchronize(p->cap));
  return GSQG_OK;
}

// The cached graphs replay steps of one (dt, key mode, key); other values
// drop them.  Captures the multi-step graph (need_k) and the one-step graph
// (need_1) where missing.  Capture records launches without running them, so
// the plan's state is untouched.
static int ensure_graphs(gsqg_plan* p, double dt, int key_mode, uint64_t key, bool need_k,
                         bool need_1) {
  if (p->g_dt != dt || p->g_key_mode != key_mode || p->g_key != key) {
    if (p->gexec) HIPCHK(hipGraphExecDestroy(p->gexec));
    if (p->gexec_k) HIPCHK(hipGraphExecDestroy(p->gexec_k));
    p->gexec = p->gexec_k = nullptr;
    p->g_dt = dt;
    p->g_key_mode = key_mode;
    p->g_key = key;
  }
  const bool derived = key_mode == GSQG_KEY_DERIVED;
  const bool sc = p->n_scalars > 0;
  int rc;
  if (need_k && !p->gexec_k)
    if ((rc = capture_steps(p, kGraphSteps, dt, derived, sc, &p->gexec_k))) return rc;
  if (need_1 && !p->gexec)
    if ((rc = capture_steps(p, 1, dt, derived, sc, &p->gexec))) return rc;
  return GSQG_OK;
}

extern "C" int gsqg_graph_capture(gsqg_plan* p, double dt, int key_mode, uint64_t key) {
  if (!p) return fail(GSQG_ERR_ARG, "null plan");
  if (key_mode != GSQG_KEY_EXPLICIT && key_mode != GSQG_KEY_DERIVED)
    return fail(GSQG_ERR_ARG, "key_mode must be GSQG_KEY_EXPLICIT or GSQG_KEY_DERIVED");
  const bool graph1 = !(p->cl_net && p->cl_every > 1);
  return ensure_graphs(p, dt, key_mode, key, true, graph1);
}

extern "C" int gsqg_advance(gsqg_plan* p, double dt, int n_steps, int key_mode, uint64_t key,
                            int64_t step0, int use_graph, void* stream) {
  if (!p) return fail(GSQG_ERR_ARG, "null plan");
  if (n_steps < 0) return fail(GSQG_ERR_ARG, "n_steps must be >= 0");
  if (n_steps == 0) return GSQG_OK;
  hipStream_t st = (hipStream_t)stream;
  const int N = p->N;
  const bool derived = key_mode == GSQG_KEY_DERIVED;
  const bool sc = p->n_scalars > 0;
  int rc;
  HIPCHK(launch_set_ctrl(p->ctrl, derived ? step0 - 1 : step0, key, key, derived ? 1 : 0, st));
  // Prologue of the first stage.
  HIPCHK(ck(p, CK_THETA, ck_theta(p, p->th, 0, 1, 1, derived ? 1 : 0, dt), p->n_mem, st));
  if (sc) {
    if ((rc = launch_scalar_ck(p, 0, 1, dt, st))) return rc;
  }
  // step 0
  {
    const bool last = (n_steps == 1);
    if ((rc = theta_row(p, rk_theta(p, 1, derived ? 1 : 0, false), p->n_mem, st, true))) return rc;
    if ((rc = theta_rest(p, dt, derived, !last || sc, last ? 0 : 1, st))) return rc;
  }
  // Eager steps fork their scalar step onto the plan's side stream, as the
  // captured steps do (not while timing every launch, which wants them serial).
  hipStream_t eager_side = p->prof ? nullptr : p->side;
  // middle steps 1 .. n-2
  const int n_mid = n_steps >= 2 ? n_steps - 2 : 0;
  if (n_mid > 0) {
    if (use_graph && n_mid >= 2) {
      const int nk = n_mid / kGraphSteps, n1 = n_mid % kGraphSteps;
      // a one-step graph repeats one evaluation pattern: with a closure held
      // over several steps the remainder runs eagerly
      const bool graph1 = !(p->cl_net && p->cl_every > 1);
      if ((rc = ensure_graphs(p, dt, key_mode, key, nk > 0, n1 > 0 && graph1))) return rc;
      for (int s = 0; s < nk; ++s) HIPCHK(hipGraphLaunch(p->gexec_k, st));
      for (int s = 0; s < n1; ++s) {
        if (graph1) {
          HIPCHK(hipGraphLaunch(p->gexec, st));
        } else if ((rc = step_body(p, dt, derived, sc, true, 1, true, st, eager_side,
                                   cl_step_evals(p, s)))) {
          return rc;
        }
      }
    } else {
      for (int s = 0; s < n_mid; ++s) {
        if ((rc = step_body(p, dt, derived, sc, true, 1, true, st, eager_side,
                            cl_step_evals(p, s))))
          return rc;
      }
    }
  }
  // last step (if n >= 2)
  if (n_steps >= 2) {
    if ((rc = step_body(p, dt, derived, sc, sc, 0, true, st, eager_side,
                        cl_step_evals(p, n_mid))))
      return rc;
  }
  if (sc) {
    // velocity of the final theta (prologue fields F2, F3 of the last CK)
    RKArgs a{};
    a.F[0] = p->F[2];
    a.F[1] = p->F[3];
    a.uout = p->uphys;
    a.vout = p->vphys;
    a.tw = p->tw;
    a.ctrl = p->ctrl;
    HIPCHK(rk(p, RK_VEL, a, 1, st));
    if ((rc = scalar_step(p, dt, false, st))) return rc;
  }
  return GSQG_OK;
}

extern "C" int gsqg_advance_scalars(gsqg_plan* p, const double* u, const double* v, double dt,
                                    void* stream) {
  if (!p || !u || !v) return fail(GSQG_ERR_ARG, "null argument");
  if (p->n_scalars == 0) return fail(GSQG_ERR_ARG, "plan has no passive scalars");
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if ((rc = launch_scalar_ck(p, 0, 1, dt, st))) return rc;
  return scalar_step(p, dt, false, st, u, v);
}

// ------------------------------------------------------------ state I/O
static int load_half(gsqg_plan* p, const double2* X, double2* H, double* offsq, bool ring,
                     hipStream_t st) {
  const int N = p->N;
  const int nb = full_to_half_num_blocks(N);
  HIPCHK(launch_full_to_half(X, H, N, p->norms, p->norms + nb, st));
  if (ring && p->n_ring > 0)
    HIPCHK(launch_a_gather(X, p->ring, p->a_val, p->n_ring, N, p->norms + nb + 1, st));
  if (offsq) {
    std::vector<double> h(p->n_norms, 0.0);
    HIPCHK(hipMemcpyAsync(h.data(), p->norms, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                          st));
    HIPCHK(hipStreamSynchronize(st));
    double tot = 0.0;
    for (int i = 0; i <= nb; ++i) tot += h[i];
    double ringsq = (ring && p->n_ring > 0) ? h[nb + 1] : 0.0;
    double off = tot - ringsq;
    *offsq = off > 0.0 ? off : 0.0;
  }
  return GSQG_OK;
}

extern "C" int gsqg_load_theta(gsqg_plan* p, const void* X, double* offring_sq, void* stream) {
  if (!p || !X) return fail(GSQG_ERR_ARG, "null argument");
  return load_half(p, (const double2*)X, p->th, offring_sq, true, (hipStream_t)stream);
}

extern "C" int gsqg_load_member(gsqg_plan* p, int idx, const void* X, void* stream) {
  if (!p || !X) return fail(GSQG_ERR_ARG, "null argument");
  if (idx < 0 || idx >= p->n_mem)
    return fail(GSQG_ERR_ARG, "member index " + std::to_string(idx) + " out of range [0, " +
                                  std::to_string(p->n_mem) + ")");
  return load_half(p, (const double2*)X, p->th + (size_t)idx * p->hsz, nullptr, idx == 0,
                   (hipStream_t)stream);
}
extern "C" int gsqg_store_member(gsqg_plan* p, int idx, void* X, void* stream) {
  if (!p || !X) return fail(GSQG_ERR_ARG, "null argument");
  if (idx < 0 || idx >= p->n_mem)
    return fail(GSQG_ERR_ARG, "member index " + std::to_string(idx) + " out of range [0, " +
                                  std::to_string(p->n_mem) + ")");
  HIPCHK(launch_half_to_full(p->th + (size_t)idx * p->hsz, (double2*)X, p->N, (hipStream_t)stream));
  if (idx == 0 && p->n_ring > 0)
    HIPCHK(launch_a_scatter((double2*)X, p->ring, p->a_val, p->n_ring, p->N, (hipStream_t)stream));
  return GSQG_OK;
}
extern "C" int gsqg_store_theta(gsqg_plan* p, void* X, void* stream) {
  if (!p || !X) return fail(GSQG_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(launch_half_to_full(p->th, (double2*)X, p->N, st));
  if (p->n_ring > 0) HIPCHK(launch_a_scatter((double2*)X, p->ring, p->a_val, p->n_ring, p->N, st));
  return GSQG_OK;
}

extern "C" int gsqg_load_scalar(gsqg_plan* p, int idx, const void* X, double* offsq,
                                void* stream) {
  if (!p || !X) return fail(GSQG_ERR_ARG, "null argument");
  if (idx < 0 || idx >= p->n_scalars) return fail(GSQG_ERR_ARG, "scalar index out of range");
  return load_half(p, (const double2*)X, p->s_th + (size_t)idx * p->hsz, offsq, false,
                   (hipStream_t)stream);
}

extern "C" int gsqg_store_scalar(gsqg_plan* p, int idx, void* X, void* stream) {
  if (!p || !X) return fail(GSQG_ERR_ARG, "null argument");
  if (idx < 0 || idx >= p->n_scalars) return fail(GSQG_ERR_ARG, "scalar index out of range");
  HIPCHK(launch_half_to_full(p->s_th + (size_t)idx * p->hsz, (double2*)X, p->N,
                             (hipStream_t)stream));
  return GSQG_OK;
}

// ------------------------------------------------------------- operators
extern "C" int gsqg_rhs(gsqg_plan* p, const void* X, void* out, uint64_t key, void* stream) {
  if (!p || !X || !out) return fail(GSQG_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const int N = p->N;
  const int nb = full_to_half_num_blocks(N);
  HIPCHK(launch_full_to_half((const double2*)X, p->th2, N, p->norms, p->norms + nb, st));
  if (p->n_ring > 0)
    HIPCHK(launch_a_gather((const double2*)X, p->ring, p->a_tmp, p->n_ring, N, nullptr, st));
  HIPCHK(launch_set_ctrl(p->ctrl, 0, key, key, 0, st));
  // compute_rhs is the explicit right-hand side whatever the plan's scheme.
  CKArgs c0 = ck_theta(p, p->th2, 0, 1, 1, 0, 0.0);
  c0.ph.scheme = 0;
  HIPCHK(ck(p, CK_THETA, c0, 1, st));
  int rc = theta_row(p, rk_theta(p, 1, 0, false), 1, st, true);
  if (rc) return rc;
  // Stage-1 epilogue writes k1 into acc (th2 is left untouched for e == 1).
  CKArgs c1 = ck_theta(p, p->th2, 1, 0, 0, 0, 0.0);
  c1.ph.scheme = 0;
  HIPCHK(ck(p, CK_THETA, c1, 1, st));
  HIPCHK(launch_half_to_full(p->acc, (double2*)out, N, st));
  if (p->n_ring > 0) {
    HIPCHK(launch_a_rhs(p->ph, p->k1d, p->flags, p->ctrl, p->ring, p->a_tmp, p->a_tmp, p->n_ring,
                        N, st));
    HIPCHK(launch_a_scatter((double2*)out, p->ring, p->a_tmp, p->n_ring, N, st));
  }
  return GSQG_OK;
}

extern "C" int gsqg_scalar_rhs(gsqg_plan* p, int idx, const void* X, const double* u,
                               const double* v, void* out, void* stream) {
  if (!p || !X || !u || !v || !out) return fail(GSQG_ERR_ARG, "null argument");
  if (idx < 0 || idx >= p->n_scalars) return fail(GSQG_ERR_ARG, "scalar index out of range");
  hipStream_t st = (hipStream_t)stream;
  const size_t off = (size_t)idx * p->hsz;
  int rc;
  if ((rc = load_half(p, (const double2*)X, p->s_th + off, nullptr, false, st))) return rc;
  // prologue (i kx s, i ky s), the row pass against (u, v), then the explicit
  // stage-1 epilogue, which leaves k1 = rhs(s) in the accumulator
  CKArgs a{};
  a.J = p->s_J + off;
  a.th = p->s_th + off;
  a.acc = p->s_acc + off;
  a.ys = p->s_ys + off;
  a.F[0] = p->s_F0 + off;
  a.F[1] = p->s_F1 + off;
  a.flags = p->flags;
  a.k1d = p->k1d;
  a.tw = p->tw;
  a.ctrl = p->ctrl;
  a.part = p->part;
  a.part_stride = p->part_stride;
  a.ph = p->ph;
  a.ph.forcing = 0;
  a.ph.damping = 0;
  a.ph.scheme = 0;
  a.ph.kappa = p->kappa[idx];
  a.e = 0;
  a.prologue = 1;
  HIPCHK(ck(p, CK_SCALAR, a, 1, st));
  RKArgs r{};
  r.F[0] = a.F[0];
  r.F[1] = a.F[1];
  r.uphys = u;
  r.vphys = v;
  r.Jout = p->s_J + off;
  r.tw = p->tw;
  r.ctrl = p->ctrl;
  r.part = p->part;
  r.n_loop = 1;
  HIPCHK(rk(p, RK_SCALAR, r, 1, st));
  a.e = 1;
  a.prologue = 0;
  HIPCHK(ck(p, CK_SCALAR, a, 1, st));
  HIPCHK(launch_half_to_full(a.acc, (double2*)out, p->N, st));
  return GSQG_OK;
}

extern "C" int gsqg_velocity(gsqg_plan* p, const void* X, double* u, double* v, void* stream) {
  if (!p || !X || !u || !v) return fail(GSQG_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const int N = p->N;
  const int nb = full_to_half_num_blocks(N);
  HIPCHK(launch_full_to_half((const double2*)X, p->th2, N, p->norms, p->norms + nb, st));
  CKArgs a = ck_theta(p, p->th2, 0, 1, 0, 0, 0.0);
  a.F[0] = p->F[2];
  a.F[1] = p->F[3];
  HIPCHK(ck(p, CK_VEL, a, 1, st));
  RKArgs r{};
  r.F[0] = p->F[2];
  r.F[1] = p->F[3];
  r.uout = u;
  r.vout = v;
  r.tw = p->tw;
  r.ctrl = p->ctrl;
  HIPCHK(rk(p, RK_VEL, r, 1, st));
  return GSQG_OK;
}

extern "C" int gsqg_fft2(gsqg_plan* p, const double* f, void* X, void* stream) {
  if (!p || !f || !X) return fail(GSQG_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const int N = p->N;
  RKArgs r{};
  r.phys_in = f;
  r.Jout = p->J;
  r.tw = p->tw;
  r.ctrl = p->ctrl;
  HIPCHK(rk(p, RK_R2C, r, 1, st));
  CKArgs a = ck_theta(p, p->th2, 0, 0, 0, 0, 0.0);
  a.J = p->J;
  a.th = p->th2;
  HIPCHK(ck(p, CK_FWD, a, 1, st));
  HIPCHK(launch_half_to_full(p->th2, (double2*)X, N, st));
  return GSQG_OK;
}

extern "C" int gsqg_ifft2(gsqg_plan* p, const void* X, double* f, void* stream) {
  if (!p || !X || !f) return fail(GSQG_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const int N = p->N;
  const int nb = full_to_half_num_blocks(N);
  HIPCHK(launch_full_to_half((const double2*)X, p->th2, N, p->norms, p->norms + nb, st));
  CKArgs a = ck_theta(p, p->th2, 0, 0, 0, 0, 0.0);
  a.th = p->th2;
  a.F[0] = p->F[0];
  HIPCHK(ck(p, CK_INV, a, 1, st));
  RKArgs r{};
  r.F[0] = p->F[0];
  r.phys_out = f;
  r.tw = p->tw;
  r.ctrl = p->ctrl;
  HIPCHK(rk(p, RK_C2R, r, 1, st));
  return GSQG_OK;
}

extern "C" int gsqg_ring_forcing(gsqg_plan* p, const void* X, uint64_t key, void* out,
                                 void* stream) {
  if (!p || !X || !out) return fail(GSQG_ERR_ARG, "null argument");
  if (!p->ph.forcing) return fail(GSQG_ERR_ARG, "plan has no ring forcing");
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(launch_ring_field((const double2*)X, p->ring, p->n_ring, p->N, p->ph.amp, p->ph.eps, key,
                           (double2*)out, p->norms, nullptr, 0.0, 0.0, 0, st));
  return GSQG_OK;
}

extern "C" int gsqg_plan_ring_count(gsqg_plan* p, int* n) {
  if (!p || !n) return fail(GSQG_ERR_ARG, "null argument");
  *n = p->n_ring;
  return GSQG_OK;
}

extern "C" int gsqg_ring_forcing_ou(gsqg_plan* p, const void* X, uint64_t key, double decay,
                                    double vfac, void* ou_state, int first, void* out,
                                    void* stream) {
  if (!p || !X || !out || !ou_state) return fail(GSQG_ERR_ARG, "null argument");
  if (!p->ph.forcing) return fail(GSQG_ERR_ARG, "plan has no ring forcing");
  if (!(decay >= 0.0 && decay <= 1.0) || !(vfac >= 0.0 && vfac <= 1.0))
    return fail(GSQG_ERR_ARG, "decay and vfac must lie in [0, 1]");
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(launch_ring_field((const double2*)X, p->ring, p->n_ring, p->N, p->ph.amp, p->ph.eps, key,
                           (double2*)out, p->norms, (double2*)ou_state, decay, vfac,
                           first ? 1 : 0, st));
  return GSQG_OK;
}

extern "C" int gsqg_profile_advance(gsqg_plan* p, double dt, int n_steps, int key_mode,
                                    uint64_t key, int64_t step0, double* class_ms,
                                    int* class_count, void* stream) {
  if (!p || !class_ms || !class_count) return fail(GSQG_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  p->prof = true;
  p->ev_used = 0;
  int rc = gsqg_advance(p, dt, n_steps, key_mode, key, step0, 0, stream);
  p->prof = false;
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(st));
  for (int c = 0; c < KC_COUNT; ++c) {
    class_ms[c] = 0.0;
    class_count[c] = 0;
  }
  for (size_t i = 0; i + 1 < p->ev_used; i += 2) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, p->ev_pool[i], p->ev_pool[i + 1]));
    const int c = p->ev_class[i];
    if (c >= 0 && c < KC_COUNT) {
      class_ms[c] += ms;
      class_count[c] += 1;
    }
  }
  return GSQG_OK;
}

// ------------------------------------------------------------- autotuner
static int drop_graphs(gsqg_plan* p) {
  if (p->gexec) HIPCHK(hipGraphExecDestroy(p->gexec));
  if (p->gexec_k) HIPCHK(hipGraphExecDestroy(p->gexec_k));
  p->gexec = p->gexec_k = nullptr;
  return GSQG_OK;
}

extern "C" int gsqg_plan_set_shape(gsqg_plan* p, int col_shape, int row_shape) {
  if (!p) return fail(GSQG_ERR_ARG, "null plan");
  if (col_shape < 0 || col_shape >= p->n_var || row_shape < 0 || row_shape >= p->n_var)
    return fail(GSQG_ERR_ARG, "shape variant out of range [0, " + std::to_string(p->n_var) +
                                  ") for N = " + std::to_string(p->N));
  if (col_shape == p->ck_var && row_shape == p->rk_var) return GSQG_OK;
  HIPCHK(hipDeviceSynchronize());  // a replayed graph may still use the old kernels
  int rc = drop_graphs(p);
  if (rc) return rc;
  p->ck_var = col_shape;
  p->rk_var = row_shape;
  return GSQG_OK;
}

extern "C" int gsqg_plan_shape(const gsqg_plan* p, int* col_shape, int* row_shape,
                               int* n_variants) {
  if (!p) return fail(GSQG_ERR_ARG, "null plan");
  if (col_shape) *col_shape = p->ck_var;
  if (row_shape) *row_shape = p->rk_var;
  if (n_variants) *n_variants = p->n_var;
  return GSQG_OK;
}

extern "C" int gsqg_plan_shape_radix(const gsqg_plan* p, int shape, int* radix) {
  if (!p || !radix) return fail(GSQG_ERR_ARG, "null argument");
  if (shape < 0 || shape >= p->n_var)
    return fail(GSQG_ERR_ARG, "shape variant out of range [0, " + std::to_string(p->n_var) +
                                  ") for N = " + std::to_string(p->N));
  *radix = shape_radix(p->N, shape);
  return GSQG_OK;
}

// Device copies of the evolving state: theta (all members), the ring side
// state, the scalars and the control block.
struct StateSave {
  double2* buf = nullptr;
  size_t n_th = 0, n_a = 0, n_s = 0;
};

static int state_copy(gsqg_plan* p, StateSave& sv, bool save, hipStream_t st) {
  const size_t b = sizeof(double2);
  double2* th = sv.buf;
  double2* av = th + sv.n_th;
  double2* sc = av + sv.n_a;
  Ctrl* ct = reinterpret_cast<Ctrl*>(sc + sv.n_s);
  const hipMemcpyKind k = hipMemcpyDeviceToDevice;
  if (save) {
    HIPCHK(hipMemcpyAsync(th, p->th, sv.n_th * b, k, st));
    if (sv.n_a) HIPCHK(hipMemcpyAsync(av, p->a_val, sv.n_a * b, k, st));
    if (sv.n_s) HIPCHK(hipMemcpyAsync(sc, p->s_th, sv.n_s * b, k, st));
    HIPCHK(hipMemcpyAsync(ct, p->ctrl, sizeof(Ctrl), k, st));
  } else {
    HIPCHK(hipMemcpyAsync(p->th, th, sv.n_th * b, k, st));
    if (sv.n_a) HIPCHK(hipMemcpyAsync(p->a_val, av, sv.n_a * b, k, st));
    if (sv.n_s) HIPCHK(hipMemcpyAsync(p->s_th, sc, sv.n_s * b, k, st));
    HIPCHK(hipMemcpyAsync(p->ctrl, ct, sizeof(Ctrl), k, st));
  }
  return GSQG_OK;
}

// Time every (column, row) shape pair on n_steps graphed steps and keep the
// fastest.  All pairs run the same arithmetic, so the choice changes speed
// only.
static int autotune_body(gsqg_plan* p, double dt, int n_steps, int repeats, int* best,
                         double* ms_table, hipStream_t st, StateSave& sv, hipEvent_t e0,
                         hipEvent_t e1) {
  int rc;
  const int nv = p->n_var;
  double best_ms = 1e300;
  int bc = p->ck_var, br = p->rk_var;
  for (int c = 0; c < nv; ++c) {
    for (int r = 0; r < nv; ++r) {
      if ((rc = gsqg_plan_set_shape(p, c, r))) return rc;
      double t = 1e300;
      for (int k = 0; k <= repeats; ++k) {  // k = 0 captures the graphs
        if ((rc = state_copy(p, sv, false, st))) return rc;
        HIPCHK(hipEventRecord(e0, st));
        if ((rc = gsqg_advance(p, dt, n_steps, GSQG_KEY_EXPLICIT, 0, 0, 1, st))) return rc;
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (k > 0) t = std::min(t, (double)ms);
      }
      if (ms_table) ms_table[c * nv + r] = t;
      if (t < best_ms) {
        best_ms = t;
        bc = c;
        br = r;
      }
    }
  }
  if ((rc = state_copy(p, sv, false, st))) return rc;
  if ((rc = gsqg_plan_set_shape(p, bc, br))) return rc;
  best[0] = bc;
  best[1] = br;
  return GSQG_OK;
}

extern "C" int gsqg_autotune(gsqg_plan* p, double dt, int n_steps, int repeats, int* best,
                             double* ms_table, void* stream) {
  if (!p || !best) return fail(GSQG_ERR_ARG, "null argument");
  if (n_steps < 1 || repeats < 1) return fail(GSQG_ERR_ARG, "n_steps and repeats must be >= 1");
  hipStream_t st = (hipStream_t)stream;
  StateSave sv;
  sv.n_th = p->hsz * (size_t)p->n_mem;
  sv.n_a = (size_t)p->n_ring;
  sv.n_s = p->hsz * (size_t)p->n_scalars;
  const size_t nb = (sv.n_th + sv.n_a + sv.n_s) * sizeof(double2) + sizeof(Ctrl);
  HIPCHK(hipMalloc((void**)&sv.buf, nb));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = GSQG_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = fail(GSQG_ERR_HIP, "hipEventCreate failed");
  if (!rc) rc = state_copy(p, sv, true, st);
  if (!rc) rc = autotune_body(p, dt, n_steps, repeats, best, ms_table, st, sv, e0, e1);
  (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(sv.buf);
  return rc;
}

extern "C" int gsqg_nonfinite(gsqg_plan* p, int* flag, int clear, void* stream) {
  if (!p || !flag) return fail(GSQG_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  Ctrl c;
  HIPCHK(hipMemcpyAsync(&c, p->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *flag = c.nonfinite != 0.0;
  if (clear && *flag) {
    double z = 0.0;
    HIPCHK(hipMemcpyAsync((char*)p->ctrl + offsetof(Ctrl, nonfinite), &z, sizeof(double),
                          hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return GSQG_OK;
}

extern "C" int gsqg_forcing_stats(gsqg_plan* p, double* rate4, double* scale4, void* stream) {
  if (!p) return fail(GSQG_ERR_ARG, "null plan");
  hipStream_t st = (hipStream_t)stream;
  Ctrl c;
  HIPCHK(hipMemcpyAsync(&c, p->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int i = 0; i < 4; ++i) {
    if (rate4) rate4[i] = c.rate[i + 1];
    if (scale4) scale4[i] = c.scale[i + 1];
  }
  return GSQG_OK;
}

// ------------------------------------------------------------ diagnostics
static int diag_scratch(gsqg_plan* p, int stride) {
  if (stride > p->d_stride) {
    if (p->d_part) (void)hipFree(p->d_part);
    if (p->d_out) (void)hipFree(p->d_out);
    p->d_part = p->d_out = nullptr;
    p->d_stride = 0;
    HIPCHK(dalloc(p, &p->d_part, (size_t)diag_blocks() * stride));
    HIPCHK(dalloc(p, &p->d_out, (size_t)stride));
    p->d_stride = stride;
  }
  if (!p->d_edges) HIPCHK(dalloc(p, &p->d_edges, (size_t)kMaxBins + 1));
  return GSQG_OK;
}

extern "C" int gsqg_spectral_diagnostics(gsqg_plan* p, const void* X, double alpha,
                                         const double* edges, int n_bins, double dk,
                                         double* bin_sums, double* shell_sums, double* sums,
                                         void* stream) {
  if (!p || !X || !edges || !bin_sums || !shell_sums || !sums)
    return fail(GSQG_ERR_ARG, "null argument");
  if (n_bins < 1 || n_bins > kMaxBins)
    return fail(GSQG_ERR_ARG, "n_bins must be in [1, " + std::to_string(kMaxBins) + "], got " +
                                  std::to_string(n_bins));
  if (!(edges[0] == 0.0) || !(edges[1] > 0.0) || !(dk > 0.0))
    return fail(GSQG_ERR_ARG, "edges must start at 0 and increase; dk must be positive");
  hipStream_t st = (hipStream_t)stream;
  const int N = p->N;
  const int nshell = N / 2 + 1;
  const int stride = n_bins + nshell + 3;
  int rc;
  if ((rc = diag_scratch(p, stride))) return rc;
  HIPCHK(hipMemcpyAsync(p->d_edges, edges, (size_t)(n_bins + 1) * sizeof(double),
                        hipMemcpyHostToDevice, st));
  HIPCHK(launch_spec_diag((const double2*)X, p->k1d, p->d_edges, N, n_bins, nshell, dk, alpha,
                          p->d_part, stride, st));
  HIPCHK(launch_diag_reduce(p->d_part, diag_blocks(), stride, stride, p->d_out, st));
  std::vector<double> h(stride);
  HIPCHK(hipMemcpyAsync(h.data(), p->d_out, (size_t)stride * sizeof(double), hipMemcpyDeviceToHost,
                        st));
  HIPCHK(hipStreamSynchronize(st));
  for (int b = 0; b < n_bins; ++b) bin_sums[b] = h[b];
  for (int s = 0; s < nshell; ++s) shell_sums[s] = h[n_bins + s];
  for (int s = 0; s < 3; ++s) sums[s] = h[n_bins + nshell + s];
  return GSQG_OK;
}

// ---------------------------------------------- allocation-free diagnostics
// gsqg_spectral_diagnostics_noalloc runs the kernels of
// gsqg_spectral_diagnostics but keeps its scratch -- the per-block partial
// sums, the reduced slots and the device copy of the bin edges -- in the
// file-scope device array `scratch` instead of plan memory allocated on first
// use.  The array is sized for n_bins <= kNaMaxBins at every supported N
// (n <= 8192, so n/2 + 1 <= 4097 shells): 512 * 4356 + 4356 + 257 doubles,
// 17.9 MB, held by each GPU the library runs on until the process ends.
// Host code never passes `scratch` itself (in host code its name is a
// stand-in, not a device address): it passes the address that
// hipGetSymbolAddress returns for the current device, kept per device.
static constexpr int kNaMaxBins = 256;
static constexpr int kNaMaxShells = 8192 / 2 + 1;
static constexpr int kNaStride = kNaMaxBins + kNaMaxShells + 3;
static constexpr int kNaBlocks = 512;  // diag_blocks(), checked at run time
static constexpr size_t kNaDoubles =
    (size_t)kNaBlocks * kNaStride + (size_t)kNaStride + (size_t)(kNaMaxBins + 1);
static __device__ __align__(256) double scratch[kNaDoubles];

static constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mu;
static double* g_na_ptr[kNaMaxDevices] = {};

// Device address of `scratch` on `dev`, looked up once per device with that
// device current (each GPU holds its own copy at its own address).
static int noalloc_scratch(int dev, double** out) {
  if (dev < 0 || dev >= kNaMaxDevices)
    return fail(GSQG_ERR_UNSUPPORTED, "device index " + std::to_string(dev) + " out of range");
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess || !ptr)
      return fail(GSQG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                    hipGetErrorString(e));
    g_na_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_na_ptr[dev];
  return GSQG_OK;
}

extern "C" int gsqg_spectral_diagnostics_noalloc(gsqg_plan* p, const void* X, double alpha,
                                                 const double* edges, int n_bins, double dk,
                                                 double* bin_sums, double* shell_sums,
                                                 double* sums, void* stream) {
  if (!p || !X || !edges || !bin_sums || !shell_sums || !sums)
    return fail(GSQG_ERR_ARG, "null argument");
  if (n_bins < 1 || n_bins > kNaMaxBins)
    return fail(GSQG_ERR_ARG, "n_bins must be in [1, " + std::to_string(kNaMaxBins) +
                                  "] for the allocation-free diagnostics, got " +
                                  std::to_string(n_bins));
  if (!(edges[0] == 0.0) || !(edges[1] > 0.0) || !(dk > 0.0))
    return fail(GSQG_ERR_ARG, "edges must start at 0 and increase; dk must be positive");
  const int N = p->N;
  const int nshell = N / 2 + 1;
  const int stride = n_bins + nshell + 3;
  if (nshell > kNaMaxShells || diag_blocks() != kNaBlocks)
    return fail(GSQG_ERR_UNSUPPORTED, "allocation-free diagnostics: n = " + std::to_string(N) +
                                          " exceeds the scratch array");
  double* base = nullptr;
  int rc;
  if ((rc = noalloc_scratch(p->device, &base))) return rc;
  // Every slot read below is written earlier in the same call (each block
  // writes all `stride` slots of its partials, the reduction all of `out`),
  // so nothing is cleared; calls that overlap on one GPU share the array and
  // race.
  double* part = base;
  double* dout = base + (size_t)kNaBlocks * kNaStride;
  double* dedges = dout + kNaStride;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(dedges, edges, (size_t)(n_bins + 1) * sizeof(double),
                        hipMemcpyHostToDevice, st));
  HIPCHK(launch_spec_diag((const double2*)X, p->k1d, dedges, N, n_bins, nshell, dk, alpha, part,
                          stride, st));
  HIPCHK(launch_diag_reduce(part, kNaBlocks, stride, stride, dout, st));
  HIPCHK(hipMemcpyAsync(bin_sums, dout, (size_t)n_bins * sizeof(double), hipMemcpyDeviceToHost,
                        st));
  HIPCHK(hipMemcpyAsync(shell_sums, dout + n_bins, (size_t)nshell * sizeof(double),
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(sums, dout + n_bins + nshell, 3 * sizeof(double), hipMemcpyDeviceToHost,
                        st));
  HIPCHK(hipStreamSynchronize(st));
  return GSQG_OK;
}

extern "C" int gsqg_physical_diagnostics(gsqg_plan* p, const void* X, double* out, void* stream) {
  if (!p || !X || !out) return fail(GSQG_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const int N = p->N;
  int rc;
  if ((rc = diag_scratch(p, 3))) return rc;
  if (!p->d_phys) HIPCHK(dalloc(p, &p->d_phys, (size_t)N * N));
  if ((rc = gsqg_ifft2(p, X, p->d_phys, stream))) return rc;
  if ((rc = gsqg_velocity(p, X, p->uphys, p->vphys, stream))) return rc;
  HIPCHK(launch_phys_diag(p->d_phys, p->uphys, p->vphys, (long long)N * N, p->d_part, st));
  HIPCHK(launch_diag_reduce(p->d_part, diag_blocks(), 3, 2, p->d_out, st));
  double h[3];
  HIPCHK(hipMemcpyAsync(h, p->d_out, sizeof(h), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const double inv = 1.0 / ((double)N * (double)N);
  out[0] = 0.5 * (h[1] * inv);
  out[1] = 0.5 * (h[0] * inv);
  out[2] = h[2];
  return GSQG_OK;
}

extern "C" int gsqg_cfl_extrema(gsqg_plan* p, const void* X, double* out, void* stream) {
  if (!p || !X || !out) return fail(GSQG_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const int N = p->N;
  int rc;
  if ((rc = diag_scratch(p, 2))) return rc;
  if ((rc = gsqg_velocity(p, X, p->uphys, p->vphys, stream))) return rc;
  const long long n = (long long)N * N;
  HIPCHK(launch_cfl(p->uphys, p->vphys, (const double2*)X, n, n, p->d_part, st));
  const int nb = diag_blocks();
  std::vector<double> h((size_t)nb * 2);
  HIPCHK(hipMemcpyAsync(h.data(), p->d_part, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double a = 0.0, b = 0.0;
  for (int i = 0; i < nb; ++i) {
    const double x = h[2 * (size_t)i], y = h[2 * (size_t)i + 1];
    a = (x != x || x > a) ? x : a;
    b = (y != y || y > b) ? y : b;
  }
  out[0] = a;
  out[1] = b;
  return GSQG_OK;
}

// ------------------------------------------------------------ closure
// Host side of include/gsqg_closure.h.

template <typename T>
static hipError_t cnn_alloc(gsqg_cnn* n, T** ptr, size_t count) {
  size_t b = count * sizeof(T);
  if (b == 0) b = 16;
  hipError_t e = hipMalloc((void**)ptr, b);
  if (e == hipSuccess) n->bytes += b;
  return e;
}

static void free_cnn(gsqg_cnn* n) {
  if (!n) return;
  void* ptrs[] = {n->P,    n->G,    n->M,    n->V,    n->Wf,   n->Wb,   n->W0r,  n->W5r,
                  n->z[0], n->z[1], n->z[2], n->z[3], n->z[4], n->a[0], n->a[1], n->a[2],
                  n->a[3], n->a[4], n->dz0,  n->dz1,  n->tau,  n->dtau, n->part, n->lpart};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete n;
}

extern "C" int gsqg_cnn_num_params(void) { return cl::NPARAM; }

extern "C" int gsqg_cnn_create(int h, int w, int max_batch, gsqg_cnn** out) {
  if (!out) return fail(GSQG_ERR_ARG, "null output handle");
  *out = nullptr;
  if (!is_pow2(h) || !is_pow2(w) || w < 64 || h < 4)
    return fail(GSQG_ERR_ARG, "image sides must be powers of two with w >= 64 and h >= 4, got " +
                                  std::to_string(h) + " x " + std::to_string(w));
  if (max_batch < 1) return fail(GSQG_ERR_ARG, "max_batch must be >= 1");
  gsqg_cnn* n = new gsqg_cnn();
  n->H = h;
  n->W = w;
  n->B = max_batch;
  const size_t px = (size_t)max_batch * h * w;
  hipError_t e = hipSuccess;
  auto A = [&](auto** p, size_t c) {
    if (e == hipSuccess) e = cnn_alloc(n, p, c);
  };
  A(&n->P, cl::NPARAM);
  A(&n->G, cl::NPARAM);
  A(&n->M, cl::NPARAM);
  A(&n->V, cl::NPARAM);
  A(&n->Wf, 4 * (size_t)cl::NWM);
  A(&n->Wb, 4 * (size_t)cl::NWM);
  A(&n->W0r, cl::NW0);
  A(&n->W5r, cl::NW5);
  for (int l = 0; l < 5; ++l) {
    A(&n->z[l], px * cl::C);
    A(&n->a[l], px * cl::C);
  }
  A(&n->dz0, px * cl::C);
  A(&n->dz1, px * cl::C);
  A(&n->tau, px);
  A(&n->dtau, px);
  A(&n->part, cl::part_floats(max_batch, h, w));
  A(&n->lpart, px / 256);
  if (e == hipSuccess) e = hipMemset(n->P, 0, cl::NPARAM * sizeof(float));
  if (e == hipSuccess) e = hipMemset(n->M, 0, cl::NPARAM * sizeof(float));
  if (e == hipSuccess) e = hipMemset(n->V, 0, cl::NPARAM * sizeof(float));
  if (e == hipSuccess) e = cl::launch_prepare(n->P, n->Wf, n->Wb, n->W0r, n->W5r, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    free_cnn(n);
    return fail(GSQG_ERR_HIP, std::string("network allocation: ") + hipGetErrorString(e));
  }
  *out = n;
  return GSQG_OK;
}

extern "C" int gsqg_cnn_destroy(gsqg_cnn* n) {
  free_cnn(n);
  return GSQG_OK;
}

extern "C" int gsqg_cnn_device_bytes(const gsqg_cnn* n, size_t* bytes) {
  if (!n || !bytes) return fail(GSQG_ERR_ARG, "null argument");
  *bytes = n->bytes;
  return GSQG_OK;
}

extern "C" int gsqg_cnn_set_params(gsqg_cnn* n, const float* params, void* stream) {
  if (!n || !params) return fail(GSQG_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(n->P, params, cl::NPARAM * sizeof(float), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemsetAsync(n->M, 0, cl::NPARAM * sizeof(float), st));
  HIPCHK(hipMemsetAsync(n->V, 0, cl::NPARAM * sizeof(float), st));
  n->t = 0;
  HIPCHK(cl::launch_prepare(n->P, n->Wf, n->Wb, n->W0r, n->W5r, st));
  return GSQG_OK;
}

extern "C" int gsqg_cnn_get_params(const gsqg_cnn* n, float* params, void* stream) {
  if (!n || !params) return fail(GSQG_ERR_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(params, n->P, cl::NPARAM * sizeof(float), hipMemcpyDeviceToDevice,
                        (hipStream_t)stream));
  return GSQG_OK;
}

static int cnn_feat(const gsqg_cnn* n, const gsqg_cnn_batch* b, bool need_target, cl::Feat* f) {
  if (!n || !b) return fail(GSQG_ERR_ARG, "null argument");
  if (!b->theta || !b->u || !b->v) return fail(GSQG_ERR_ARG, "null input plane");
  if (b->batch < 1 || b->batch > n->B)
    return fail(GSQG_ERR_ARG, "batch must be in [1, " + std::to_string(n->B) + "], got " +
                                  std::to_string(b->batch));
  if (need_target && !b->target) return fail(GSQG_ERR_ARG, "null target");
  f->plane[0] = b->theta;
  f->plane[1] = b->u;
  f->plane[2] = b->v;
  f->img_stride = b->img_stride;
  f->sel = b->sel_dev;
  for (int c = 0; c < 3; ++c) f->in_scale[c] = b->in_scale[c];
  return GSQG_OK;
}

extern "C" int gsqg_cnn_forward(gsqg_cnn* n, const gsqg_cnn_batch* b, double out_scale, float* tau,
                                void* stream) {
  cl::Feat f{};
  int rc = cnn_feat(n, b, false, &f);
  if (rc) return rc;
  if (!tau) return fail(GSQG_ERR_ARG, "null output");
  uint16_t* const z[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  HIPCHK(cl::launch_forward(cnn_net(n), f, b->batch, n->H, n->W, z, n->a, tau, (float)out_scale,
                            (hipStream_t)stream));
  return GSQG_OK;
}

// Forward (keeping z and a), loss, backward into n->G.
static int cnn_gradient(gsqg_cnn* n, const gsqg_cnn_batch* b, hipStream_t st) {
  cl::Feat f{};
  int rc = cnn_feat(n, b, true, &f);
  if (rc) return rc;
  const int B = b->batch;
  HIPCHK(cl::launch_forward(cnn_net(n), f, B, n->H, n->W, n->z, n->a, n->tau, 1.0f, st));
  HIPCHK(cl::launch_loss(n->tau, b->target, b->img_stride, b->sel_dev, b->tgt_scale, n->dtau,
                         n->lpart, B, n->H, n->W, st));
  HIPCHK(cl::launch_backward(cnn_net(n), f, B, n->H, n->W, n->z, n->a, n->dtau, n->dz0, n->dz1,
                             n->part, n->G, st));
  return GSQG_OK;
}

static int cnn_loss(const gsqg_cnn* n, int B, hipStream_t st, double* loss) {
  const size_t nb = (size_t)B * n->H * n->W / 256;
  std::vector<double> h(nb);
  HIPCHK(hipMemcpyAsync(h.data(), n->lpart, nb * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double s = 0.0;
  for (double x : h) s += x;
  *loss = s / ((double)B * n->H * n->W);
  return GSQG_OK;
}

extern "C" int gsqg_cnn_gradient(gsqg_cnn* n, const gsqg_cnn_batch* b, float* grad, double* loss,
                                 void* stream) {
  hipStream_t st = (hipStream_t)stream;
  int rc = cnn_gradient(n, b, st);
  if (rc) return rc;
  if (grad)
    HIPCHK(hipMemcpyAsync(grad, n->G, cl::NPARAM * sizeof(float), hipMemcpyDeviceToDevice, st));
  return loss ? cnn_loss(n, b->batch, st, loss) : GSQG_OK;
}

extern "C" int gsqg_cnn_train_step(gsqg_cnn* n, const gsqg_cnn_batch* b, double lr, double* loss,
                                   void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (!(lr > 0.0)) return fail(GSQG_ERR_ARG, "learning rate must be positive");
  int rc = cnn_gradient(n, b, st);
  if (rc) return rc;
  n->t += 1;
  HIPCHK(cl::launch_adam(n->P, n->G, n->M, n->V, lr, 0.9, 0.999, 1e-8, n->t, st));
  HIPCHK(cl::launch_prepare(n->P, n->Wf, n->Wb, n->W0r, n->W5r, st));
  return loss ? cnn_loss(n, b->batch, st, loss) : GSQG_OK;
}

extern "C" int gsqg_plan_attach_closure(gsqg_plan* p, const gsqg_cnn* net, const double* in_scale,
                                        double out_scale, int every) {
  if (!p) return fail(GSQG_ERR_ARG, "null plan");
  // captured graphs hold the previous schedule
  if (p->gexec) HIPCHK(hipGraphExecDestroy(p->gexec));
  if (p->gexec_k) HIPCHK(hipGraphExecDestroy(p->gexec_k));
  p->gexec = p->gexec_k = nullptr;
  p->g_key_mode = -1;
  if (!net) {
    p->cl_net = nullptr;
    return GSQG_OK;
  }
  if (!in_scale) return fail(GSQG_ERR_ARG, "null input scales");
  if (every < 0 || every > kGraphSteps || (every > 0 && kGraphSteps % every != 0))
    return fail(GSQG_ERR_ARG, "every must be 0 (per stage) or divide " +
                                  std::to_string(kGraphSteps) + ", got " + std::to_string(every));
  const int N = p->N;
  if (net->H != N || net->W != N)
    return fail(GSQG_ERR_ARG, "the network is " + std::to_string(net->H) + " x " +
                                  std::to_string(net->W) + ", the plan's grid " +
                                  std::to_string(N) + " x " + std::to_string(N));
  if (p->n_mem > 1) return fail(GSQG_ERR_UNSUPPORTED, "closure on an ensemble plan");
  if (!p->cl_ys) {
    const size_t n2 = (size_t)N * N;
    HIPCHK(dalloc(p, &p->cl_ys, p->hsz));
    HIPCHK(dalloc(p, &p->cl_f, p->hsz));
    HIPCHK(dalloc(p, &p->cl_th, n2));
    HIPCHK(dalloc(p, &p->cl_u, n2));
    HIPCHK(dalloc(p, &p->cl_v, n2));
    HIPCHK(dalloc(p, &p->cl_a0, n2 * cl::C));
    HIPCHK(dalloc(p, &p->cl_a1, n2 * cl::C));
    HIPCHK(dalloc(p, &p->cl_tau, n2));
  }
  p->cl_net = net;
  for (int c = 0; c < 3; ++c) p->cl_in[c] = in_scale[c];
  p->cl_out = out_scale;
  p->cl_every = every;
  return GSQG_OK;
}

extern "C" int gsqg_spectral_filter(const void* in, int nf, void* out, int nc, double scale,
                                    const uint8_t* mask, const void* add, void* stream) {
  if (!in || !out || !mask) return fail(GSQG_ERR_ARG, "null argument");
  if (nf < 2 || nc < 2 || (nf & 1) || (nc & 1) || nc > nf)
    return fail(GSQG_ERR_ARG, "need even sizes with nc <= nf, got nf=" + std::to_string(nf) +
                                  " nc=" + std::to_string(nc));
  HIPCHK(cl::launch_spectral_filter((const double2*)in, nf, (double2*)out, nc, scale, mask,
                                    (const double2*)add, (hipStream_t)stream));
  return GSQG_OK;
}

// ------------------------------------------------------------- batches
// Uneven ensembles in one launch per RK4 stage (gsqg_batch_*).  The
// schedule is gsqg_advance's for an unforced, scalar-free plan with the
// explicit key mode -- prologue, then per step four (row pass, column pass)
// pairs, the last column pass of the last step without a prologue -- with
// each launch covering every plan.
struct gsqg_batch {
  std::vector<gsqg_plan*> plans;
  hipStream_t cap = nullptr;
  hipGraphExec_t gexec_k = nullptr;  // kGraphSteps captured steps
  double g_dt = 0.0;
};

static int batch_ck(gsqg_batch* b, int e, int prologue, double dt, hipStream_t st) {
  CKBatch m{};
  m.n = (int)b->plans.size();
  int blk = 0;
  for (int g = 0; g < m.n; ++g) {
    gsqg_plan* p = b->plans[g];
    m.N[g] = p->N;
    m.gx[g] = batch_blocks(p->N, true);
    m.blk0[g] = blk;
    m.a[g] = ck_theta(p, p->th, e, prologue, 0, 0, dt);
    blk += m.gx[g] * p->n_mem;
  }
  m.blk0[m.n] = blk;
  HIPCHK(launch_ck_batch(e, m, st));
  return GSQG_OK;
}

static int batch_rk(gsqg_batch* b, hipStream_t st) {
  RKBatch m{};
  m.n = (int)b->plans.size();
  int blk = 0;
  for (int g = 0; g < m.n; ++g) {
    gsqg_plan* p = b->plans[g];
    m.N[g] = p->N;
    m.gx[g] = batch_blocks(p->N, false);
    m.blk0[g] = blk;
    m.a[g] = rk_theta(p, 0, 0, false);
    blk += m.gx[g] * p->n_mem;
  }
  m.blk0[m.n] = blk;
  HIPCHK(launch_rk_batch(m, st));
  return GSQG_OK;
}

// One step: four (row, column) stage pairs; the last column pass writes the
// next step's prologue unless this is the final step.
static int batch_step(gsqg_batch* b, double dt, bool next_prologue, hipStream_t st) {
  int rc;
  for (int e = 1; e <= 4; ++e) {
    if ((rc = batch_rk(b, st))) return rc;
    if ((rc = batch_ck(b, e, (e < 4 || next_prologue) ? 1 : 0, dt, st))) return rc;
  }
  return GSQG_OK;
}

static int batch_drop_graph(gsqg_batch* b) {
  if (b->gexec_k) HIPCHK(hipGraphExecDestroy(b->gexec_k));
  b->gexec_k = nullptr;
  return GSQG_OK;
}

extern "C" int gsqg_batch_create(gsqg_plan* const* plans, int n_plans, gsqg_batch** out) {
  if (!plans || !out) return fail(GSQG_ERR_ARG, "null argument");
  *out = nullptr;
  if (n_plans < 1 || n_plans > kBatchMaxPlans)
    return fail(GSQG_ERR_ARG, "a batch holds 1 to " + std::to_string(kBatchMaxPlans) +
                                  " plans, got " + std::to_string(n_plans));
  for (int g = 0; g < n_plans; ++g) {
    const gsqg_plan* p = plans[g];
    if (!p) return fail(GSQG_ERR_ARG, "null plan in batch");
    for (int h = 0; h < g; ++h)
      if (plans[h] == p) return fail(GSQG_ERR_ARG, "a plan appears twice in the batch");
    if (!batch_supports(p->N))
      return fail(GSQG_ERR_UNSUPPORTED, "batched grids need 64 <= N <= 1024 (power of two), got " +
                                            std::to_string(p->N));
    if (p->ph.forcing || p->n_scalars > 0 || p->cl_net)
      return fail(GSQG_ERR_UNSUPPORTED, "batched plans carry no forcing, scalars or closure");
    if (p->device != plans[0]->device)
      return fail(GSQG_ERR_ARG, "batched plans must live on one device");
  }
  gsqg_batch* b = new gsqg_batch();
  b->plans.assign(plans, plans + n_plans);
  hipError_t e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete b;
    return fail(GSQG_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  *out = b;
  return GSQG_OK;
}

extern "C" int gsqg_batch_destroy(gsqg_batch* b) {
  if (!b) return GSQG_OK;
  int rc = batch_drop_graph(b);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  delete b;
  return rc;
}

extern "C" int gsqg_batch_advance(gsqg_batch* b, double dt, int n_steps, int use_graph,
                                  void* stream) {
  if (!b) return fail(GSQG_ERR_ARG, "null batch");
  if (n_steps < 0) return fail(GSQG_ERR_ARG, "n_steps must be >= 0");
  if (n_steps == 0) return GSQG_OK;
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if ((rc = batch_ck(b, 0, 1, dt, st))) return rc;  // prologue of the first stage
  // steps 0 .. n-2 end with the next step's prologue: whole 16-step graphs
  // where they fit, the rest eagerly; the last step ends without one
  int n_a = n_steps - 1;
  if (use_graph && n_a >= kGraphSteps) {
    if (b->gexec_k && b->g_dt != dt && (rc = batch_drop_graph(b))) return rc;
    if (!b->gexec_k) {
      hipGraph_t g;
      HIPCHK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
      for (int s = 0; s < kGraphSteps && rc == GSQG_OK; ++s) rc = batch_step(b, dt, true, b->cap);
      hipError_t ce = hipStreamEndCapture(b->cap, &g);
      if (rc) return rc;
      if (ce != hipSuccess)
        return fail(GSQG_ERR_HIP, std::string("capture: ") + hipGetErrorString(ce));
      hipError_t ie = hipGraphInstantiate(&b->gexec_k, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess)
        return fail(GSQG_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
      b->g_dt = dt;
    }
    for (; n_a >= kGraphSteps; n_a -= kGraphSteps) HIPCHK(hipGraphLaunch(b->gexec_k, st));
  }
  for (; n_a > 0; --n_a)
    if ((rc = batch_step(b, dt, true, st))) return rc;
  return batch_step(b, dt, false, st);
}
