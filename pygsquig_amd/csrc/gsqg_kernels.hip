// Fused pseudo-spectral RK4 kernels for the gSQG step on gfx950.
//
// One RHS stage of the reference (pygsquig/core/solver.py:28-58: 9 c2c FFTs
// over full (N, N) complex arrays) becomes two launches over the half-packed
// ky-major layout described in gsqg_internal.h:
//
//   column kernel  (ck_kernel, x direction, one block per N/8-lane row group)
//     epilogue of stage e:  forward x-FFT of the y-transformed Jacobian,
//                           dealias, hyperviscosity, damping, ring forcing,
//                           RK4 (or integrating-factor) combination
//     prologue of stage e+1: Hermitianised multipliers i kx, i ky,
//                           -i ky |k|^-a, i kx |k|^-a (x 1/N^2), inverse
//                           x-FFT of the four fields; ring <theta f> partials
//   row kernel     (rk_kernel, y direction, one block per WX x-rows)
//     four C2R y-transforms (half-length complex FFT + pre-twist), the
//     product u*theta_x + v*theta_y in registers, one R2C y-transform.
//
// Both kernels keep each transform in LDS/registers: a field crosses HBM
// exactly once per read or write.
#include <hip/hip_ext.h>

#include "gsqg_common.h"
#include "gsqg_internal.h"

namespace gsqg {

// ------------------------------------------------------------------ shapes
// Shape variant V, chosen per plan by the autotuner (gsqg_autotune).
// Variants 0-2 run the radix-8 schedule (8 elements per lane): 0 is the
// default tile of 128 threads; 1 doubles and 2 halves the rows (column
// kernel) or x-rows (row kernel) of a block.  Variant 3 runs the radix-16
// schedule (16 elements per lane, half the lanes per transform): fewer LDS
// exchange passes and, at large N, blocks whose register budget holds a
// whole row without spilling.  Variant 4 runs the radix-4 schedule (4
// elements per lane, twice the lanes of radix 8): where one radix-8 wave
// per SIMD is all a launch has (N <= 1024), two waves per SIMD hide each
// other's memory, LDS and FP64 latencies.  Which variants each N compiles
// is listed in variant_id() below.
template <int N, int V = 0>
struct CKShape {
  // elements per lane; variant 4 keeps radix 8 where radix 4 would need more
  // than 1024 lanes per row (N = 8192), so that it names the row kernel's
  // radix-4 tile alone there
  static constexpr int R = V == 3 ? 16 : (V == 4 && N / 4 > 1024) ? 8 : (V == 4 || V == 5) ? 4 : 8;
  static constexpr int Q = N / R;                        // lanes per row
  static constexpr int R00 = (128 / Q) > 1 ? (128 / Q) : 1;
  // radix 4: one row per block down to one wave, then one wave per block
  static constexpr int R0 = V == 1   ? 2 * R00
                            : V == 2 ? (R00 > 1 ? R00 / 2 : R00)
                            : V == 4 ? ((64 / Q) > 1 ? (64 / Q) : 1)
                            : V == 5 ? 1
                                     : R00;
  static constexpr int RPW = R0 < (N / 2) ? R0 : (N / 2);  // rows per block
  // Field groups (variant 5): FG sets of Q lanes run one row's epilogue
  // side by side and then transform different prologue fields, so a row's
  // inverse transforms run in parallel instead of one group after another.
  static constexpr int FG = V == 5 ? 2 : 1;
  static constexpr int NT = RPW * Q * FG;
  static constexpr int PADL = Padded<N>::value;
  static constexpr int REGION = RPW * PADL;  // LDS elements per field
  // Prologue fields transformed together: two where they fit in 80 KB of
  // LDS.  Not four: the smaller LDS and register footprint lets theta and
  // scalar blocks share a CU when the two run as concurrent graph branches.
  static constexpr int FP = REGION * 16 * 2 <= 80 * 1024 ? 2 : 1;
};

template <int N, int V = 0>
struct RKShape {
  static constexpr int M = N / 2;
  // variant 5 pairs the field-group column tile with the radix-4 row tile
  static constexpr int R = V == 3 ? 16 : (V == 4 || V == 5) ? 4 : 8;  // elements per lane
  static constexpr int Q = M / R;                        // lanes per x-row
  static constexpr int W00 = (128 / Q) > 1 ? (128 / Q) : 1;
  // Radix-16 tiles at 256 lanes per x-line take two x-lines, so that each
  // row of a tile is 32 contiguous bytes of the ky-major fields; radix-4
  // tiles take two x-lines, or one wave's worth where x-lines are shorter,
  // or one x-line where that alone has 512 lanes or more (N >= 4096).
  static constexpr int W0 = V == 1   ? 2 * W00
                            : V == 2 ? (W00 > 1 ? W00 / 2 : W00)
                            : V == 3 ? (Q >= 256 ? 2 : W00)
                            : (V == 4 || V == 5) ? (Q >= 512 ? 1 : (64 / Q) > 2 ? (64 / Q) : 2)
                                     : W00;
  static constexpr int WX = W0 < N ? W0 : N;  // x-rows per block
  static constexpr int NT = WX * Q;
  static constexpr int PADM = Padded<M>::value;
  // Stride between the x-lines of a tile.  Where a wave holds one x-line
  // (Q >= 64) the transforms never mix lines, and 8 / WX extra elements
  // spread the transposing accesses (WX lines x 64 / WX rows per
  // instruction) over the LDS banks instead of stacking them PADM apart.
  static constexpr int RSTR = PADM + ((Q >= 64 && WX > 1 && WX <= 8) ? 8 / WX : 0);
  static constexpr int REGION = WX * RSTR;  // LDS elements per field tile
  // Field tiles transformed together (two where they fit in 80 KB of LDS;
  // see CKShape::FP).
  static constexpr int FP = REGION * 16 * 2 <= 80 * 1024 ? 2 : 1;
};

// Spectral fields a row-kernel job reads.
template <int JOB>
struct RKFields {
  static constexpr int value =
      (JOB == RK_THETA || JOB == RK_THETA_CL) ? 4 : ((JOB == RK_SCALAR || JOB == RK_VEL) ? 2 : 1);
};

// ------------------------------------------------------- per-mode physics
// |k|^2 as the reference forms it, KX**2 + KY**2 (core/grid.py:96): each
// square and the sum rounded separately.  Never contracted to an fma, so
// every kernel and stage that needs |k|^2 of a mode gets the same value.
__device__ __forceinline__ double k2_of(double kx, double ky) {
  return __dadd_rn(__dmul_rn(kx, kx), __dmul_rn(ky, ky));
}

__device__ __forceinline__ double hv_factor(double k2, int p) {
  if (p == 2) return k2;
  double s = k2 * k2;
  if (p == 4) return s;
  return s * s;
}

// |k|^-alpha, zero at k = 0 (core/operators.py:149-153).
__device__ __forceinline__ int alpha_kind_of(double alpha) {
  return alpha == 1.0 ? 1 : alpha == 0.0 ? 2 : alpha == 2.0 ? 3 : 0;
}
__device__ __forceinline__ double inv_frac(double k2, double alpha, int kind) {
  if (!(k2 > 0.0)) return 0.0;
  switch (kind) {
    case 1:
      return 1.0 / sqrt(k2);
    case 2:
      return 1.0;
    case 3:
      return 1.0 / k2;
    default:
      return pow(sqrt(k2), -alpha);
  }
}

// Real diagonal linear operator: solver hyperviscosity plus CombinedDamping.
__device__ __forceinline__ double lin_coeff(double k2, bool dc, unsigned fl, const Phys& ph) {
  double l = (dc && ph.hv_dc_zero) ? 0.0 : -ph.nu * hv_factor(k2, ph.p);
  if (ph.damping) {
    double ls = (fl & FLAG_LSDAMP) ? -ph.mu : 0.0;
    double dh = dc ? 0.0 : -ph.nu_d * hv_factor(k2, ph.p_d);
    l = l + (ls + dh);
  }
  return l;
}

// Linear part of the explicit RHS, in the reference's order:
// (adv + hv*y) + s*F + (ls*y + dhv*y)  (solver.py:170-178, damping.py:75-92).
__device__ __forceinline__ cd explicit_rhs(cd adv, cd y, double k2, bool dc, unsigned fl,
                                           const Phys& ph, bool forced, cd sF) {
  double hv = (dc && ph.hv_dc_zero) ? 0.0 : -ph.nu * hv_factor(k2, ph.p);
  cd k = cadd(adv, cscale(y, hv));
  if (forced) k = cadd(k, sF);
  if (ph.damping) {
    cd ls = (fl & FLAG_LSDAMP) ? cscale(y, -ph.mu) : cmk(0.0, 0.0);
    double dh = dc ? 0.0 : -ph.nu_d * hv_factor(k2, ph.p_d);
    k = cadd(k, cadd(ls, cscale(y, dh)));
  }
  return k;
}

// Hermitian (PF) and anti-Hermitian (QF) parts of the unscaled ring forcing
// F = amp * exp(i phi) at mode (i, jy) (ring_forcing.py:138-150: phases drawn
// over the full plane, DC and the Nyquist corners made real).
__device__ __forceinline__ void ring_parts(uint64_t key, int i, int jy, int N, double amp, cd& PF,
                                           cd& QF) {
  const int mi = (N - i) & (N - 1), mj = (N - jy) & (N - 1);
  double s1, c1, s2, c2;
  // exp(i phi) with phi = 2 pi u: sincospi(2u) needs no argument reduction
  sincospi(2.0 * phase_u(key, (uint64_t)i * (uint64_t)N + (uint64_t)jy), &s1, &c1);
  sincospi(2.0 * phase_u(key, (uint64_t)mi * (uint64_t)N + (uint64_t)mj), &s2, &c2);
  cd f1 = cmk(amp * c1, amp * s1);
  cd f2 = cmk(amp * c2, amp * s2);
  const bool selfc = (i == 0 || i == N / 2) && (jy == 0 || jy == N / 2);
  if (selfc) {
    f1.y = 0.0;
    f2 = f1;
  }
  PF = cscale(cadd(f1, cconj(f2)), 0.5);
  QF = cscale(csub(f1, cconj(f2)), 0.5);
}

__device__ __forceinline__ uint64_t ctrl_key(const Ctrl* c, int64_t step) {
  return c->key_mode ? step_key(c->seed, step) : c->key;
}

// Per-mode epilogue state.
struct ModeIn {
  cd J;    // forward-transformed Jacobian (not yet dealiased / negated)
  cd th;   // theta_hat_n (Hermitian part)
  cd acc;  // RK accumulator
  cd ys;   // stored stage input
};
struct ModeOut {
  cd th, acc, ys, y;  // y = input of the next stage
};

// Stage e epilogue at one mode.  Explicit scheme follows
// core/time_integrator.py:30-36 (and core/solver_optimized.py:62 for the
// multistep combine); the IF scheme is Lawson RK4 with the diagonal linear
// operator integrated exactly.
template <bool SCALAR>
__device__ __forceinline__ ModeOut mode_epilogue(int e, const ModeIn& in, double k2, bool dc,
                                                 unsigned fl, const Phys& ph, double dt,
                                                 bool forced, cd sF) {
  ModeOut o;
  o.th = in.th;
  o.acc = in.acc;
  o.ys = in.ys;
  const double hdt = 0.5 * dt;
  cd adv = (fl & FLAG_DEALIAS) ? cmk(-in.J.x, -in.J.y) : cmk(0.0, 0.0);
  if (ph.scheme == 0) {
    cd y;
    if (e == 1)
      y = in.th;
    else if (e == 2)
      y = cadd(in.th, cscale(in.acc, hdt));
    else
      y = in.ys;
    cd k;
    if (SCALAR) {
      k = cadd(adv, cscale(y, -ph.kappa * k2));
    } else {
      k = explicit_rhs(adv, y, k2, dc, fl, ph, forced, sF);
    }
    if (e == 1) {
      o.acc = k;
      o.y = cadd(in.th, cscale(k, hdt));
    } else if (e == 2) {
      o.acc = cadd(in.acc, cscale(k, 2.0));
      o.ys = cadd(in.th, cscale(k, hdt));
      o.y = o.ys;
    } else if (e == 3) {
      o.acc = cadd(in.acc, cscale(k, 2.0));
      o.ys = cadd(in.th, cscale(k, dt));
      o.y = o.ys;
    } else {
      cd s = cadd(in.acc, k);
      cd d;
      if (ph.combine == 0) {
        cd t = cscale(s, dt);
        d = cmk(t.x / 6.0, t.y / 6.0);
      } else {
        d = cscale(s, dt / 6.0);
      }
      o.th = cadd(in.th, d);
      o.y = o.th;
    }
  } else {
    double L = SCALAR ? -ph.kappa * k2 : lin_coeff(k2, dc, fl, ph);
    // exp(L dt) = exp(L dt/2)^2: one exponential per mode
    const double e1 = exp(L * hdt), e2 = e1 * e1;
    cd k = adv;
    if (!SCALAR && forced) k = cadd(k, sF);
    if (e == 1) {
      o.acc = cscale(k, e2);
      o.y = cscale(cadd(in.th, cscale(k, hdt)), e1);
    } else if (e == 2) {
      o.acc = cadd(in.acc, cscale(cscale(k, e1), 2.0));
      o.y = cadd(cscale(in.th, e1), cscale(k, hdt));
    } else if (e == 3) {
      cd ek = cscale(k, e1);
      o.acc = cadd(in.acc, cscale(ek, 2.0));
      o.y = cadd(cscale(in.th, e2), cscale(ek, dt));
    } else {
      o.th = cadd(cscale(in.th, e2), cscale(cadd(in.acc, k), dt / 6.0));
      o.y = o.th;
    }
  }
  return o;
}

// Prologue multipliers (Hermitianised, core/operators.py:27-28,90-91 with
// the .real of ifft2 folded in: i kx vanishes on the kx-Nyquist row and i ky
// on the ky-Nyquist column), including the 1/N^2 of the inverse transform.
// Field q of the job, given K = |k|^-alpha.
template <int JOB>
__device__ __forceinline__ cd prologue_field(int q, cd y, double kx, double ky, double K, bool kxn,
                                             bool kyn, const Phys& ph) {
  const double s = ph.inv_n2;
  const double cx = kxn ? 0.0 : kx * s;
  const double cy = kyn ? 0.0 : ky * s;
  if (JOB == CK_THETA) {
    if (q == 0) return cmul_i(cscale(y, cx));        // theta_x
    if (q == 1) return cmul_i(cscale(y, cy));        // theta_y
    if (q == 2) return cmul_mi(cscale(y, cy * K));   // u = -d_y psi
    return cmul_i(cscale(y, cx * K));                // v =  d_x psi
  } else if (JOB == CK_SCALAR) {
    if (q == 0) return cmul_i(cscale(y, cx));
    return cmul_i(cscale(y, cy));
  }
  // CK_VEL
  if (q == 0) return cmul_mi(cscale(y, cy * K));
  return cmul_i(cscale(y, cx * K));
}

template <int JOB>
struct NFields {
  static constexpr int value = JOB == CK_THETA ? 4 : 2;
};

__device__ __forceinline__ unsigned mode_flags(const uint8_t* flags, int N, int i, int jy) {
  if (jy == N / 2) return (unsigned)(flags[i] >> 4);
  return (unsigned)(flags[(size_t)jy * N + i] & 15u);
}

// ------------------------------------------------------------ column kernel
// Row 0 of every spectral state array (theta, accumulator, stage input) is
// kept in the "R layout": R[r] = X(r, 0) for 0 < r < N/2,
// R[N/2 + r] = X(r, N/2) for 0 < r < N/2, and the four self-conjugate modes
// as real pairs R[0] = X(0,0) + i X(0,N/2), R[N/2] = X(N/2,0) + i X(N/2,N/2).
// Every element is then one mode (two real modes at r = 0, N/2) and is
// updated in place by the lane that owns it.  The x transforms need the
// packed form Z(i) = X(i,0) + i X(i,N/2); the conversion goes through LDS.

// Mode (kx index, ky index) of R-layout element r (r != 0, N/2).
__device__ __forceinline__ int2 r_mode(int r, int N) {
  int2 m;
  if (r < N / 2) {
    m.x = r;
    m.y = 0;
  } else {
    m.x = r - N / 2;
    m.y = N / 2;
  }
  return m;
}

// Z (packed) from an R-layout row in LDS.
__device__ __forceinline__ cd z_from_r(const cd* buf, int i, int N) {
  if (i == 0 || i == N / 2) return buf[lpad(i)];
  if (i < N / 2) return cadd(buf[lpad(i)], cmul_i(buf[lpad(N / 2 + i)]));
  return cadd(cconj(buf[lpad(N - i)]), cmul_i(cconj(buf[lpad(N / 2 + N - i)])));
}

// R-layout element r from a packed Z row in LDS.
__device__ __forceinline__ cd r_from_z(const cd* buf, int r, int N) {
  if (r == 0 || r == N / 2) return buf[lpad(r)];
  if (r < N / 2) return cscale(cadd(buf[lpad(r)], cconj(buf[lpad(N - r)])), 0.5);
  const int q = r - N / 2;
  return cscale(cmul_mi(csub(buf[lpad(q)], cconj(buf[lpad(N - q)]))), 0.5);
}

// Per-lane wavenumbers and flags of the column kernel, loaded at kernel
// start: kx and the raw flag byte of elements t + m*Q of row j, the row's ky
// and k1d[N/2] (row 0 holds the ky = N/2 modes too).
//
// With KC (the register-tight shapes) kx is not kept per element but
// recomputed from its index the way the host builds the table,
// k / (dx N) * 2 * pi with dxn = dx N: the same three correctly rounded
// operations, so the values are bit-identical (plan creation checks the
// table against this formula).  The flag bytes are packed four to a word.
constexpr double kPi = 3.141592653589793;
template <int N, int R, bool KC>
struct RowCtx {
  double kxs[KC ? 1 : R];
  uint32_t fbw[(R + 3) / 4];
  double ky, kyn, dxn;
  int t;
  __device__ __forceinline__ double kx(int m) const {
    if constexpr (KC) {
      const int i = t + m * (N / R);
      const int ki = i < N / 2 ? i : i - N;
      return ((double)ki / dxn) * 2.0 * kPi;
    } else {
      return kxs[m];
    }
  }
  __device__ __forceinline__ unsigned fb(int m) const { return (fbw[m >> 2] >> (8 * (m & 3))) & 0xffu; }
};

// Mode of element m (not a row-0 self-conjugate element): wavenumbers,
// flags, indices and the Nyquist lines its multipliers vanish on.  Row 0 is
// in R layout: with R elements per lane, elements m < R/2 are (r, 0) and
// elements m >= R/2 are (r - N/2, N/2), so both use kx and the flag byte of
// element m mod R/2.
struct MInfo {
  double kx, ky;
  unsigned fl;
  int ii, jy;
  bool kxnyq, kynyq;
};
template <int N, int R, bool KC>
__device__ __forceinline__ MInfo minfo(const RowCtx<N, R, KC>& c, int m, int t, int j, bool row0) {
  constexpr int Q = N / R;
  MInfo r;
  if (!row0) {
    r.kx = c.kx(m);
    r.ky = c.ky;
    r.fl = c.fb(m) & 15u;
    r.ii = t + m * Q;
    r.jy = j;
    r.kxnyq = (r.ii == N / 2);
    r.kynyq = false;
  } else {
    const int m4 = m & (R / 2 - 1);
    r.kx = c.kx(m4);
    r.ii = t + m4 * Q;
    r.kxnyq = false;
    if (m < R / 2) {
      r.ky = c.ky;
      r.fl = c.fb(m4) & 15u;
      r.jy = 0;
      r.kynyq = false;
    } else {
      r.ky = c.kyn;
      r.fl = c.fb(m4) >> 4;
      r.jy = N / 2;
      r.kynyq = true;
    }
  }
  return r;
}

// Epilogue at a self-conjugate row-0 element r (0 or N/2): two real modes,
// (r, 0) in the real parts and (r, N/2) in the imaginary parts; kxr = k1d[r],
// fbyte = flag byte of column r of row 0.
template <int JOB, bool FORCED>
__device__ __forceinline__ ModeOut row0_selfc_epilogue(int E, int r, const ModeIn& in, double kxr,
                                                       unsigned fbyte, double ky0, double kyn,
                                                       const Phys& ph, double dt, uint64_t key,
                                                       double sc) {
  const int N = ph.N;
  ModeOut oo[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int jy = h ? N / 2 : 0;
    const double ky = h ? kyn : ky0;
    const unsigned fl = h ? (fbyte >> 4) : (fbyte & 15u);
    ModeIn u;
    u.J = cmk(h ? in.J.y : in.J.x, 0.0);
    u.th = cmk(h ? in.th.y : in.th.x, 0.0);
    u.acc = cmk(h ? in.acc.y : in.acc.x, 0.0);
    u.ys = cmk(h ? in.ys.y : in.ys.x, 0.0);
    cd sF = cmk(0.0, 0.0);
    if (FORCED && (fl & FLAG_RING)) {
      cd PF, QF;
      ring_parts(key, r, jy, N, ph.amp, PF, QF);
      sF = cscale(PF, sc);
    }
    oo[h] = mode_epilogue<JOB == CK_SCALAR>(E, u, k2_of(kxr, ky), (r == 0 && jy == 0), fl,
                                            ph, dt, FORCED, sF);
  }
  ModeOut o;
  o.th = cmk(oo[0].th.x, oo[1].th.x);
  o.acc = cmk(oo[0].acc.x, oo[1].acc.x);
  o.ys = cmk(oo[0].ys.x, oo[1].ys.x);
  o.y = cmk(oo[0].y.x, oo[1].y.x);
  return o;
}

// LDS elements of a column-kernel block: the field regions and the twiddle
// octant table (N <= 2048; one element otherwise).
template <int N, int JOB, int V>
struct CKLds {
  using S = CKShape<N, V>;
  static constexpr int NFL = (JOB == CK_FWD || JOB == CK_INV) ? 1 : NFields<JOB>::value;
  // field groups: one region per prologue field (each API transform group
  // its own region)
  static constexpr int FIELDS = S::FG > 1 ? (NFL > S::FG ? NFL : S::FG) * S::PADL
                                          : (S::FP < NFL ? S::FP : NFL) * S::REGION;
  static constexpr int TW = N <= 2048 ? OctTab<N>::SIZE : 1;
};

// One column-kernel block: ky rows bx * RPW .. of field slice by (gx blocks
// per slice), with its LDS at lds (CKLds::FIELDS elements) and twl
// (CKLds::TW).  ck_kernel runs it as block (blockIdx.x, blockIdx.y); the
// multi-plan kernel of the batched path runs blocks of several plans in one
// launch.
template <int N, int JOB, int E, bool FORCED, int V>
__device__ __forceinline__ void ck_body(const CKArgs& a, const int bx, const int by, const int gx,
                                        cd* lds, cd* twl) {
#include "ck_body.inc"
}

template <int N, int JOB, int E, bool FORCED, int V>
__global__ void __launch_bounds__((CKShape<N, V>::NT)) ck_kernel(CKArgs a) {
  using LS = CKLds<N, JOB, V>;
  __shared__ cd lds[LS::FIELDS];
  __shared__ cd twl[LS::TW];
  const unsigned bx = blockIdx.x, by = blockIdx.y, gx = gridDim.x;
#include "ck_body.inc"
}

// --------------------------------------------------------------- row kernel
template <int N, int V>
__device__ __forceinline__ int rk_tile(int b) {
  constexpr int T = N / RKShape<N, V>::WX;
  if (T % 8 != 0) return b;
  // Blocks are dispatched round-robin over the 8 XCDs: give each XCD a
  // contiguous run of x-tiles so that tiles sharing 128-byte lines of the
  // ky-major fields meet in the same L2.
  return (b & 7) * (T / 8) + (b >> 3);
}

// Issue the global loads of G mixed fields' tiles (WX x-columns by M rows,
// coalesced along x) into registers.  Staged as native 2-vectors: plain
// values that stay in registers (an array of the aggregate type is copied
// through scratch).
template <int N, int V, int G>
__device__ __forceinline__ void rk_issue(const cd* const (&src)[G], int x0, int tid,
                                         vd2 (&tmp)[G * RKShape<N, V>::R]) {
  using S = RKShape<N, V>;
  constexpr int WX = S::WX, NT = S::NT, R = S::R;
#pragma unroll
  for (int e = 0; e < G * R; ++e) {
    const int l = tid + (e % R) * NT;
    const int jr = l / WX, bb = l - jr * WX;
    tmp[e] = *reinterpret_cast<const vd2*>(src[e / R] + (size_t)jr * mixed_pitch(N) + x0 + bb);
  }
}

// Pre/post-twist factor e^{-2 pi i k/N}, k < N/2: from the LDS octant
// table (TWL) or from the global table.
template <int N, bool TWL>
struct PreTw {
  const cd* t;
  __device__ __forceinline__ cd operator()(int k) const {
    if (TWL) return OctTab<N>{t}.fwd(k);
    return t[N + k];
  }
};

// Write staged tiles to the LDS regions (region g at lds + g * REGION), then
// pre-twist for the half-length complex inverse: lane element k of field g
// gets Z'(k) = (X(k) + conj X(M-k)) + i e^{+2 pi i k/N} (X(k) - conj X(M-k)),
// with pre(k) = e^{-2 pi i k/N}.  Ends without a barrier: the next LDS write
// (inside the transform) is preceded by one.
template <int N, int V, int G, class PT>
__device__ __forceinline__ void rk_commit(const vd2 (&tmp)[G * RKShape<N, V>::R], cd* lds, int tid,
                                          int b, int t, const PT& pre,
                                          cd (&v)[G][RKShape<N, V>::R]) {
  using S = RKShape<N, V>;
  constexpr int M = S::M, WX = S::WX, NT = S::NT, Q = S::Q, RG = S::REGION, R = S::R;
  __syncthreads();
#pragma unroll
  for (int e = 0; e < G * R; ++e) {
    const int l = tid + (e % R) * NT;
    const int jr = l / WX, bb = l - jr * WX;
    *reinterpret_cast<vd2*>(lds + (e / R) * RG + bb * S::RSTR + lpad(jr)) = tmp[e];
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int k = t + m * Q;
    const cd w = cconj(pre(k));
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const cd* buf = lds + g * RG + b * S::RSTR;
      const cd X = buf[lpad(k)];
      if (k == 0) {
        v[g][m] = cmk(X.x + X.y, X.x - X.y);
      } else {
        const cd Xp = cconj(buf[lpad(M - k)]);
        v[g][m] = cadd(cadd(X, Xp), cmul_i(cmul(w, csub(X, Xp))));
      }
    }
  }
}

// Forward half-length FFT result in v -> post-twist -> store X(k), packed
// slot 0 = X(0) + i X(M), at dst[k*N + x0 + b] (coalesced along x).
template <int N, int V, int R, class PT>
__device__ __forceinline__ void rk_store_r2c(cd (&v)[R], cd* dst, int x0, cd* lds, int tid, int b,
                                             int t, const PT& pre) {
  using S = RKShape<N, V>;
  static_assert(R == S::R, "elements per lane of the shape");
  constexpr int M = S::M, WX = S::WX, NT = S::NT, Q = S::Q;
  cd* buf = lds + b * S::RSTR;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < R; ++m) buf[lpad(t + m * Q)] = v[m];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int l = tid + r * NT;
    const int jr = l / WX, bb = l - jr * WX;
    const cd* bf = lds + bb * S::RSTR;
    const cd Z = bf[lpad(jr)];
    cd X;
    if (jr == 0) {
      X = cmk(Z.x + Z.y, Z.x - Z.y);
    } else {
      const cd Zp = cconj(bf[lpad(M - jr)]);
      const cd E = cscale(cadd(Z, Zp), 0.5);
      const cd O = cscale(cmul_mi(csub(Z, Zp)), 0.5);
      X = cadd(E, cmul(pre(jr), O));
    }
    dst[(size_t)jr * mixed_pitch(N) + x0 + bb] = X;
  }
}

// Row-kernel twiddles and pre/post-twist factors from an LDS octant table
// of W_N (W_M = W_N^2): for every job up to N = 2048, and above that for the
// time-step jobs where the table fits beside the field tiles in 160 KB.
// There the table (18 KB at N = 8192) costs no occupancy, since registers
// hold those blocks to one per CU, and without it each FFT pass's twiddle
// reads are global loads, which complete in issue order and so wait for the
// next field's tile loads issued ahead of the transform.
template <int N, int JOB, int V>
struct RKLds {
  using S = RKShape<N, V>;
  static constexpr int NF = RKFields<JOB>::value;
  static constexpr int FIELDS = (S::FP < NF ? S::FP : NF) * S::REGION;
  static constexpr bool STEP_JOB =
      JOB == RK_THETA || JOB == RK_THETA_CL || JOB == RK_SCALAR || JOB == RK_VEL;
  static constexpr bool TWL =
      N <= 2048 || (STEP_JOB && (FIELDS + OctTab<N>::SIZE) * (int)sizeof(cd) <= 160 * 1024);
  static constexpr int TW = TWL ? OctTab<N>::SIZE : 1;
};

// One row-kernel block: x-tile bx of field slice by (see ck_body).
template <int N, int JOB, int V>
__device__ __forceinline__ void rk_body(const RKArgs& a, const int bx, const int by,
                                        cd* lds, cd* twl) {
#include "rk_body.inc"
}

template <int N, int JOB, int V>
__global__ void __launch_bounds__((RKShape<N, V>::NT)) rk_kernel(RKArgs a) {
  using LS = RKLds<N, JOB, V>;
  __shared__ cd lds[LS::FIELDS];
  __shared__ cd twl[LS::TW];
  const unsigned bx = blockIdx.x, by = blockIdx.y;
#include "rk_body.inc"
}

#if defined(GSQG_INSTANTIATE_COMMON)
// ------------------------------------------------------- layout converters
// full (N, N) [kx][ky] complex -> half-packed ky-major P(X), rows 1..N/2-1.
// Also accumulates sum |Q(X)|^2 (weight 2: the mirror mode is not visited)
// per block into norms[blockIdx].
template <int TILE>
__global__ void full_to_half_kernel(const cd* __restrict__ X, cd* __restrict__ H, int N,
                                    double* norms) {
  __shared__ cd a[TILE][TILE + 1];
  __shared__ cd mrr[TILE][TILE + 1];
  __shared__ double red[256];
  const int i0 = blockIdx.x * TILE;  // kx
  const int j0 = blockIdx.y * TILE;  // ky (only 0..N/2)
  const int tx = threadIdx.x % TILE, ty = threadIdx.x / TILE;
  const int rows_per = 256 / TILE;
  for (int r = ty; r < TILE; r += rows_per) {
    const int i = i0 + r;
    const int jy = j0 + tx;
    if (jy < N / 2 && jy > 0) {
      a[r][tx] = X[(size_t)i * N + jy];
      const int mi = (N - i) & (N - 1), mj = N - jy;
      mrr[r][tx] = X[(size_t)mi * N + mj];
    }
  }
  __syncthreads();
  double nrm = 0.0;
  for (int r = ty; r < TILE; r += rows_per) {
    const int jy = j0 + r;
    const int i = i0 + tx;
    if (jy < N / 2 && jy > 0) {
      const cd z = a[tx][r], zm = cconj(mrr[tx][r]);
      H[(size_t)jy * N + i] = cscale(cadd(z, zm), 0.5);
      const cd q = cscale(csub(z, zm), 0.5);
      nrm += 2.0 * (q.x * q.x + q.y * q.y);
    }
  }
  red[threadIdx.x] = nrm;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0 && norms) norms[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// Row 0 of the spectral layout (ky = 0 and ky = N/2 columns) in R layout:
// R[r] = P(X)(r, 0), R[N/2 + r] = P(X)(r, N/2), self-conjugate modes as
// real pairs at r = 0 and r = N/2.
__global__ void full_to_half_row0_kernel(const cd* __restrict__ X, cd* __restrict__ H, int N,
                                         double* norm0) {
  __shared__ double red[256];
  double nrm = 0.0;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < N; r += gridDim.x * blockDim.x) {
    const int i = (r <= N / 2) ? r : r - N / 2;   // kx index
    const int jy = (r < N / 2 || r == 0) ? 0 : (r == N / 2 ? 0 : N / 2);
    const int mi = (N - i) & (N - 1);
    if (r == 0 || r == N / 2) {
      const cd a0 = X[(size_t)i * N], b0 = cconj(X[(size_t)mi * N]);
      const cd aN = X[(size_t)i * N + N / 2], bN = cconj(X[(size_t)mi * N + N / 2]);
      const cd p0 = cscale(cadd(a0, b0), 0.5), pN = cscale(cadd(aN, bN), 0.5);
      H[r] = cmk(p0.x, pN.x);
      const cd q0 = cscale(csub(a0, b0), 0.5), qN = cscale(csub(aN, bN), 0.5);
      nrm += q0.x * q0.x + q0.y * q0.y + qN.x * qN.x + qN.y * qN.y;
    } else {
      const cd a = X[(size_t)i * N + jy], b = cconj(X[(size_t)mi * N + jy]);
      H[r] = cscale(cadd(a, b), 0.5);
      const cd q = cscale(csub(a, b), 0.5);
      nrm += 2.0 * (q.x * q.x + q.y * q.y);  // the mirror mode -i is not visited
    }
  }
  red[threadIdx.x] = nrm;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0 && norm0) norm0[blockIdx.x] = red[0];
}

// half-packed ky-major -> full (N, N) [kx][ky] with Hermitian extension.
template <int TILE>
__global__ void half_to_full_kernel(const cd* __restrict__ H, cd* __restrict__ X, int N) {
  __shared__ cd a[TILE][TILE + 1];
  const int i0 = blockIdx.x * TILE;  // kx (output rows)
  const int c0 = blockIdx.y * TILE;  // ky (output columns)
  const int tx = threadIdx.x % TILE, ty = threadIdx.x / TILE;
  const int rows_per = 256 / TILE;
  // Gather: for output column jy = c0 + r, read along kx (contiguous in H).
  for (int r = ty; r < TILE; r += rows_per) {
    const int jy = c0 + r;
    const int i = i0 + tx;
    cd val = cmk(0.0, 0.0);
    if (jy > 0 && jy < N / 2) {
      val = H[(size_t)jy * N + i];
    } else if (jy > N / 2) {
      val = cconj(H[(size_t)(N - jy) * N + ((N - i) & (N - 1))]);
    } else {
      // row 0, R layout
      if (i == 0 || i == N / 2) {
        const cd R = H[i];
        val = cmk(jy == 0 ? R.x : R.y, 0.0);
      } else {
        const int base = (jy == 0) ? 0 : N / 2;
        val = (i < N / 2) ? H[base + i] : cconj(H[base + N - i]);
      }
    }
    a[tx][r] = val;
  }
  __syncthreads();
  for (int r = ty; r < TILE; r += rows_per) {
    const int i = i0 + r;
    X[(size_t)i * N + c0 + tx] = a[r][tx];
  }
}

// Anti-Hermitian side state <-> full spectra.
__global__ void a_gather_kernel(const cd* __restrict__ X, const int2* modes, cd* aval, int n,
                                int N, double* norm) {
  __shared__ double red[256];
  double s = 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int i = modes[e].x, jy = modes[e].y;
    const int mi = (N - i) & (N - 1), mj = (N - jy) & (N - 1);
    const cd z = X[(size_t)i * N + jy], zm = cconj(X[(size_t)mi * N + mj]);
    const cd q = cscale(csub(z, zm), 0.5);
    aval[e] = q;
    const double w = (jy == 0 || jy == N / 2) ? 1.0 : 2.0;
    s += w * (q.x * q.x + q.y * q.y);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0 && norm) norm[blockIdx.x] = red[0];
}

__global__ void a_scatter_kernel(cd* X, const int2* modes, const cd* aval, int n, int N) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int i = modes[e].x, jy = modes[e].y;
    const cd A = aval[e];
    const size_t o = (size_t)i * N + jy;
    X[o] = cadd(X[o], A);
    if (jy != 0 && jy != N / 2) {
      const int mi = (N - i) & (N - 1), mj = N - jy;
      const size_t om = (size_t)mi * N + mj;
      X[om] = csub(X[om], cconj(A));
    }
  }
}

// ------------------------------------------------------------ diagnostics
// Spectral diagnostics of a full (N, N) spectrum X (axis 0 = kx index i,
// axis 1 = ky index j), utils/diagnostics.py:21-84, 119-141, 170-216.  Per
// block: the reference's binned energy spectrum (density 0.5 |k|^(alpha-2)
// |X|^2, X[0,0] excluded, bin b when edges[b] <= |k| < edges[b+1]), the
// integer-shell spectrum (the same density / N^4 at shell rint(|k|/dk)), and
// the sums of the total-energy, enstrophy and palinstrophy densities.  |k|^2
// is formed as kx*kx + ky*ky with both products rounded (no contraction), as
// the reference's KX**2 + KY**2, so a mode sits in the reference's bin.
// Output per block at part[blk * stride]: nb bins, nshell shells, 3 sums.
__global__ void __launch_bounds__(256) spec_diag_kernel(const cd* __restrict__ X,
                                                        const double* __restrict__ k1d,
                                                        const double* __restrict__ edges, int N,
                                                        int nb, int nshell, double dk,
                                                        double alpha, double* part, int stride) {
  extern __shared__ double sh[];  // nb + nshell bins, then 3 * 256 sums
  double* ssum = sh + nb + nshell;
  const int tid = threadIdx.x;
  for (int b = tid; b < nb + nshell; b += 256) sh[b] = 0.0;
  __syncthreads();
  const double n4 = (double)N * (double)N * (double)N * (double)N;
  const double step = edges[1];
  const double kmax = edges[nb];
  double sE = 0.0, sQ = 0.0, sP = 0.0;
  const long long total = (long long)N * N;
  for (long long e = (long long)blockIdx.x * 256 + tid; e < total;
       e += (long long)gridDim.x * 256) {
    const int i = (int)(e / N), j = (int)(e - (long long)i * N);
    const cd x = X[e];
    const double kx = k1d[i], ky = k1d[j];
    const double k2 = k2_of(kx, ky);
    const double a2 = __dadd_rn(__dmul_rn(x.x, x.x), __dmul_rn(x.y, x.y));
    const bool dc = (i == 0 && j == 0);
    const double kmag = sqrt(k2);
    const double ksafe = kmag > 0.0 ? kmag : 1.0;
    const double ds = dc ? 0.0 : 0.5 * pow(ksafe, alpha - 2.0) * a2;
    if (kmag < kmax) {
      int g = (int)(kmag / step);
      g = g < 0 ? 0 : (g > nb - 1 ? nb - 1 : g);
      while (g > 0 && kmag < edges[g]) --g;
      while (g < nb - 1 && kmag >= edges[g + 1]) ++g;
      if (kmag >= edges[g] && kmag < edges[g + 1]) atomicAdd(&sh[g], ds);
    }
    const double sr = rint(kmag / dk);
    if (sr < (double)nshell) atomicAdd(&sh[nb + (int)sr], ds / n4);
    const double k2s = k2 > 0.0 ? k2 : 1.0;
    if (!dc) {
      sE += 0.5 * pow(k2s, 0.5 * (alpha - 2.0)) * a2;
      sP += 0.5 * pow(k2s, 1.0 + alpha) * a2;
    }
    if (k2 > 0.0) {
      const double f = pow(ksafe, alpha);
      sQ += (f * f) * a2;
    }
  }
  ssum[tid] = sE;
  ssum[256 + tid] = sQ;
  ssum[512 + tid] = sP;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) {
      ssum[tid] += ssum[tid + h];
      ssum[256 + tid] += ssum[256 + tid + h];
      ssum[512 + tid] += ssum[512 + tid + h];
    }
    __syncthreads();
  }
  double* out = part + (size_t)blockIdx.x * stride;
  for (int b = tid; b < nb + nshell; b += 256) out[b] = sh[b];
  if (tid == 0) {
    out[nb + nshell] = ssum[0];
    out[nb + nshell + 1] = ssum[256];
    out[nb + nshell + 2] = ssum[512];
  }
}

// Physical diagnostics (core/solver.py:216-245) of theta, u, v (n doubles
// each): per block sum theta^2, sum (u^2 + v^2), max |theta|.
__global__ void __launch_bounds__(256) phys_diag_kernel(const double* __restrict__ th,
                                                        const double* __restrict__ u,
                                                        const double* __restrict__ v,
                                                        long long n, double* part) {
  __shared__ double s0[256], s1[256], s2[256];
  const int tid = threadIdx.x;
  double a = 0.0, b = 0.0, m = 0.0;
  for (long long e = (long long)blockIdx.x * 256 + tid; e < n; e += (long long)gridDim.x * 256) {
    const double t = th[e];
    a += t * t;
    b += u[e] * u[e] + v[e] * v[e];
    m = fmax(m, fabs(t));
  }
  s0[tid] = a;
  s1[tid] = b;
  s2[tid] = m;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) {
      s0[tid] += s0[tid + h];
      s1[tid] += s1[tid + h];
      s2[tid] = fmax(s2[tid], s2[tid + h]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    part[(size_t)blockIdx.x * 3] = s0[0];
    part[(size_t)blockIdx.x * 3 + 1] = s1[0];
    part[(size_t)blockIdx.x * 3 + 2] = s2[0];
  }
}

// Per-block maxima for the CFL controller: part[2 b] = max sqrt(u^2 + v^2),
// part[2 b + 1] = max |X| over a complex array; a NaN anywhere makes the
// block's maximum NaN.
__device__ __forceinline__ double nan_max(double m, double x) { return (x != x || x > m) ? x : m; }

__global__ void __launch_bounds__(256) cfl_kernel(const double* __restrict__ u,
                                                  const double* __restrict__ v,
                                                  const double2* __restrict__ X, long long n,
                                                  long long nx, double* part) {
  __shared__ double s0[256], s1[256];
  const int tid = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (long long e = (long long)blockIdx.x * 256 + tid; e < n; e += (long long)gridDim.x * 256)
    a = nan_max(a, sqrt(u[e] * u[e] + v[e] * v[e]));
  for (long long e = (long long)blockIdx.x * 256 + tid; e < nx; e += (long long)gridDim.x * 256)
    b = nan_max(b, hypot(X[e].x, X[e].y));
  s0[tid] = a;
  s1[tid] = b;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) {
      s0[tid] = nan_max(s0[tid], s0[tid + h]);
      s1[tid] = nan_max(s1[tid], s1[tid + h]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    part[(size_t)blockIdx.x * 2] = s0[0];
    part[(size_t)blockIdx.x * 2 + 1] = s1[0];
  }
}

// Sum (or max, for slots >= nsum) the per-block partials slot by slot, in
// block order (deterministic).
__global__ void __launch_bounds__(256) diag_reduce_kernel(const double* part, int nblk, int stride,
                                                          int nsum, double* out) {
  for (int s = blockIdx.x * 256 + threadIdx.x; s < stride; s += gridDim.x * 256) {
    double acc = 0.0;
    for (int b = 0; b < nblk; ++b) {
      const double x = part[(size_t)b * stride + s];
      acc = s < nsum ? acc + x : fmax(acc, x);
    }
    out[s] = acc;
  }
}

// Side-state table: wavenumbers and flags of every ring mode, gathered once.
__global__ void a_tab_kernel(const int2* modes, const double* k1d, const uint8_t* flags, ATab* out,
                             int n, int N) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int i = modes[e].x, jy = modes[e].y;
    ATab t;
    t.kx = k1d[i];
    t.ky = k1d[jy];
    t.ii = i;
    t.jy = jy;
    t.fl = mode_flags(flags, N, i, jy);
    t.dc = (i == 0 && jy == 0) ? 1 : 0;
    out[e] = t;
  }
}

// Anti-Hermitian part of compute_rhs at the ring side-state modes.
__global__ void a_rhs_kernel(Phys ph, const double* k1d, const uint8_t* flags, const Ctrl* ctrl,
                             const int2* modes, const cd* aval, cd* out, int n, int N) {
  const uint64_t key = ctrl_key(ctrl, ctrl->step);
  const double s1 = ctrl->scale[1];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int i = modes[e].x, jy = modes[e].y;
    const double kx = k1d[i], ky = k1d[jy];
    const double k2 = k2_of(kx, ky);
    const unsigned fl = mode_flags(flags, N, i, jy);
    cd PF, QF;
    ring_parts(key, i, jy, N, ph.amp, PF, QF);
    const cd A = aval[e];
    out[e] = explicit_rhs(cmk(0.0, 0.0), A, k2, (i == 0 && jy == 0), fl, ph, ph.forcing != 0,
                          cscale(QF, s1));
  }
}

// White ring noise w = amp * exp(i phi) at mode (i, jy) and at its mirror
// (-i, -jy), before the DC / Nyquist realification (ring_forcing.py:138-141).
__device__ __forceinline__ void ring_white(uint64_t key, int i, int jy, int N, double amp, cd& f1,
                                           cd& f2) {
  const int mi = (N - i) & (N - 1), mj = (N - jy) & (N - 1);
  double s1, c1, s2, c2;
  sincospi(2.0 * phase_u(key, (uint64_t)i * (uint64_t)N + (uint64_t)jy), &s1, &c1);
  sincospi(2.0 * phase_u(key, (uint64_t)mi * (uint64_t)N + (uint64_t)mj), &s2, &c2);
  f1 = cmk(amp * c1, amp * s1);
  f2 = cmk(amp * c2, amp * s2);
}

// Ornstein-Uhlenbeck filter of one white value (ring_forcing.py:105-117):
// the state becomes the white value on the first call, else
// decay * state + vfac * white.  Returns the new state, which is also stored.
__device__ __forceinline__ cd ou_update(cd* st, cd w, double decay, double vfac, int first) {
  cd s;
  if (first) {
    s = w;
  } else {
    const cd o = *st;
    s = cmk(__dadd_rn(__dmul_rn(decay, o.x), __dmul_rn(vfac, w.x)),
            __dadd_rn(__dmul_rn(decay, o.y), __dmul_rn(vfac, w.y)));
  }
  *st = s;
  return s;
}

// The forcing of ring entry e: Hermitian and anti-Hermitian parts of
// F = amp * exp(i phi), or, with an OU state (two values per entry: the mode
// and its mirror), of the filtered state after this call's update; DC and
// the Nyquist corners made real afterwards (ring_forcing.py:138-150).
__device__ __forceinline__ void ring_entry_parts(uint64_t key, int i, int jy, int N, double amp,
                                                 cd* ou, int e, double decay, double vfac,
                                                 int first, cd& PF, cd& QF) {
  if (!ou) {
    ring_parts(key, i, jy, N, amp, PF, QF);
    return;
  }
  cd f1, f2;
  ring_white(key, i, jy, N, amp, f1, f2);
  f1 = ou_update(ou + 2 * (size_t)e, f1, decay, vfac, first);
  f2 = ou_update(ou + 2 * (size_t)e + 1, f2, decay, vfac, first);
  if ((i == 0 || i == N / 2) && (jy == 0 || jy == N / 2)) {
    f1.y = 0.0;
    f2 = f1;
  }
  PF = cscale(cadd(f1, cconj(f2)), 0.5);
  QF = cscale(csub(f1, cconj(f2)), 0.5);
}

// RingForcing.__call__ on a full spectrum: rate by Parseval over the ring
// list, then F = scale * (raw forcing) at every ring mode
// (ring_forcing.py:119-170).  With `ou` non-null the raw forcing is the OU
// state (2 n values), updated once here (ring_forcing.py:144-145).  One
// block; `out` must be zero-filled.  Each thread reads back only the OU
// entries it updated itself.
__global__ void ring_field_kernel(const cd* __restrict__ X, const int2* modes, int n, int N,
                                  double amp, double eps, uint64_t key, cd* out, double* rate_out,
                                  cd* ou, double decay, double vfac, int ou_first) {
  __shared__ double red[256];
  double s = 0.0;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const int i = modes[e].x, jy = modes[e].y;
    const int mi = (N - i) & (N - 1), mj = (N - jy) & (N - 1);
    const cd H = cscale(cadd(X[(size_t)i * N + jy], cconj(X[(size_t)mi * N + mj])), 0.5);
    cd PF, QF;
    ring_entry_parts(key, i, jy, N, amp, ou, e, decay, vfac, ou_first, PF, QF);
    const double w = (jy == 0 || jy == N / 2) ? 1.0 : 2.0;
    s += w * (H.x * PF.x + H.y * PF.y);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double inv = 1.0 / ((double)N * (double)N);
  const double rate = (red[0] * inv) * inv;
  const double sc = fabs(rate) > 1e-10 ? eps / rate : 1.0;
  if (threadIdx.x == 0 && rate_out) *rate_out = rate;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const int i = modes[e].x, jy = modes[e].y;
    const int mi = (N - i) & (N - 1), mj = (N - jy) & (N - 1);
    cd PF, QF;
    if (ou) {
      cd f1 = ou[2 * (size_t)e], f2 = ou[2 * (size_t)e + 1];
      if ((i == 0 || i == N / 2) && (jy == 0 || jy == N / 2)) {
        f1.y = 0.0;
        f2 = f1;
      }
      PF = cscale(cadd(f1, cconj(f2)), 0.5);
      QF = cscale(csub(f1, cconj(f2)), 0.5);
    } else {
      ring_parts(key, i, jy, N, amp, PF, QF);
    }
    // F(k) = PF + QF, F(-k) = conj(PF - QF)
    out[(size_t)i * N + jy] = cscale(cadd(PF, QF), sc);
    if (jy != 0 && jy != N / 2) out[(size_t)mi * N + mj] = cscale(cconj(csub(PF, QF)), sc);
  }
}

__global__ void set_ctrl_kernel(Ctrl* c, int64_t step, uint64_t key, uint64_t seed, int key_mode) {
  c->step = step;
  c->key = key;
  c->seed = seed;
  c->key_mode = key_mode;
}

#endif  // GSQG_INSTANTIATE_COMMON

// ------------------------------------------------------------- launchers
// The build compiles this file once per grid size (GSQG_INSTANTIATE_N) and
// once for the size-independent parts (GSQG_INSTANTIATE_COMMON).
template <int N, int JOB, int E, bool FORCED, int V>
static hipError_t launch_ck_n(const CKArgs& a, int nfields_y, hipStream_t s, const LaunchEv* ev) {
  using S = CKShape<N, V>;
  dim3 grid((N / 2) / S::RPW, nfields_y);
  if (ev)
    hipExtLaunchKernelGGL((ck_kernel<N, JOB, E, FORCED, V>), grid, dim3(S::NT), 0, s, ev->start,
                          ev->stop, 0, a);
  else
    hipLaunchKernelGGL((ck_kernel<N, JOB, E, FORCED, V>), grid, dim3(S::NT), 0, s, a);
  return hipGetLastError();
}

template <int N, int JOB, bool FORCED, int V>
static hipError_t launch_ck_e(const CKArgs& a, int ny, hipStream_t s, const LaunchEv* ev) {
  switch (a.e) {
    case 0: return launch_ck_n<N, JOB, 0, FORCED, V>(a, ny, s, ev);
    case 1: return launch_ck_n<N, JOB, 1, FORCED, V>(a, ny, s, ev);
    case 2: return launch_ck_n<N, JOB, 2, FORCED, V>(a, ny, s, ev);
    case 3: return launch_ck_n<N, JOB, 3, FORCED, V>(a, ny, s, ev);
    case 4: return launch_ck_n<N, JOB, 4, FORCED, V>(a, ny, s, ev);
  }
  return hipErrorInvalidValue;
}

template <int N, int JOB, int V>
static hipError_t launch_rk_n(const RKArgs& a, int nfields_y, hipStream_t s, const LaunchEv* ev) {
  using S = RKShape<N, V>;
  dim3 grid(N / S::WX, nfields_y);
  if (ev)
    hipExtLaunchKernelGGL((rk_kernel<N, JOB, V>), grid, dim3(S::NT), 0, s, ev->start, ev->stop, 0,
                          a);
  else
    hipLaunchKernelGGL((rk_kernel<N, JOB, V>), grid, dim3(S::NT), 0, s, a);
  return hipGetLastError();
}

template <int N, int V>
static hipError_t ck_by_job_v(int job, const CKArgs& a, int ny, hipStream_t s, const LaunchEv* ev) {
  switch (job) {
    case CK_THETA:
      return a.ph.forcing ? launch_ck_e<N, CK_THETA, true, V>(a, ny, s, ev)
                          : launch_ck_e<N, CK_THETA, false, V>(a, ny, s, ev);
    case CK_SCALAR: return launch_ck_e<N, CK_SCALAR, false, V>(a, ny, s, ev);
    case CK_VEL: return launch_ck_n<N, CK_VEL, 0, false, V>(a, ny, s, ev);
    case CK_FWD: return launch_ck_n<N, CK_FWD, 0, false, V>(a, ny, s, ev);
    case CK_INV: return launch_ck_n<N, CK_INV, 0, false, V>(a, ny, s, ev);
  }
  return hipErrorInvalidValue;
}

template <int N, int V>
static hipError_t rk_by_job_v(int job, const RKArgs& a, int ny, hipStream_t s, const LaunchEv* ev) {
  switch (job) {
    case RK_THETA: return launch_rk_n<N, RK_THETA, V>(a, ny, s, ev);
    case RK_SCALAR: return launch_rk_n<N, RK_SCALAR, V>(a, ny, s, ev);
    case RK_VEL: return launch_rk_n<N, RK_VEL, V>(a, ny, s, ev);
    case RK_R2C: return launch_rk_n<N, RK_R2C, V>(a, ny, s, ev);
    case RK_C2R: return launch_rk_n<N, RK_C2R, V>(a, ny, s, ev);
    case RK_THETA_CL: return launch_rk_n<N, RK_THETA_CL, V>(a, ny, s, ev);
  }
  return hipErrorInvalidValue;
}

// Shape variants compiled for grid size N, in the autotuner's index order;
// index 0 is the default.  Radix-8 tiles 0-2 where a block of 128 threads
// holds several rows (at N = 1024 the default column block is one row, so
// there is no smaller tile), and from N = 64 up the radix-16 variant 3 and
// the radix-4 variant 4 (at N = 8192 a radix-4 row tile beside the radix-8
// column tile), and up to N = 512 the field-group
// variant 5: at N = 1024 its 512-lane blocks need more than 128 registers
// per lane in the later stages, so only one fits per CU; measured 10.8 ms
// per 32 steps against 7.9 ms for the radix-4 pair).
template <int N>
constexpr int n_variants() {
  return N <= 32 ? 3 : N <= 512 ? 6 : N == 1024 ? 4 : 3;
}
template <int N>
constexpr int variant_id(int k) {
  if (N >= 2048) return k == 0 ? 0 : k == 1 ? 3 : 4;  // radix 8, 16, 4
  if (N == 1024) return k <= 1 ? k : k + 1;  // 0, 1, 3, 4
  return k;  // 0, 1, 2 (and 3, 4, 5 from N = 64)
}

template <int N, int K = 0>
hipError_t ck_by_job(int job, const CKArgs& a, int ny, hipStream_t s, const LaunchEv* ev, int var) {
  if constexpr (K < n_variants<N>()) {
    if (var == K || K + 1 == n_variants<N>()) return ck_by_job_v<N, variant_id<N>(K)>(job, a, ny, s, ev);
    return ck_by_job<N, K + 1>(job, a, ny, s, ev, var);
  }
  return hipErrorInvalidValue;
}

template <int N, int K = 0>
hipError_t rk_by_job(int job, const RKArgs& a, int ny, hipStream_t s, const LaunchEv* ev, int var) {
  if constexpr (K < n_variants<N>()) {
    if (var == K || K + 1 == n_variants<N>()) return rk_by_job_v<N, variant_id<N>(K)>(job, a, ny, s, ev);
    return rk_by_job<N, K + 1>(job, a, ny, s, ev, var);
  }
  return hipErrorInvalidValue;
}

template <int N, int K = 0>
int ck_blocks_n(int var) {
  if constexpr (K < n_variants<N>()) {
    if (var == K || K + 1 == n_variants<N>()) {
      using S = CKShape<N, variant_id<N>(K)>;
      return (N / 2) / S::RPW;
    }
    return ck_blocks_n<N, K + 1>(var);
  }
  return 0;
}

template <int N, int K = 0>
int radix_n(int var) {
  if constexpr (K < n_variants<N>()) {
    if (var == K || K + 1 == n_variants<N>()) return RKShape<N, variant_id<N>(K)>::R;
    return radix_n<N, K + 1>(var);
  }
  return 0;
}

#define GSQG_FOR_EACH_N(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192)

#if defined(GSQG_INSTANTIATE_N)
template hipError_t ck_by_job<GSQG_INSTANTIATE_N>(int, const CKArgs&, int, hipStream_t,
                                                  const LaunchEv*, int);
template hipError_t rk_by_job<GSQG_INSTANTIATE_N>(int, const RKArgs&, int, hipStream_t,
                                                  const LaunchEv*, int);
#endif

#if defined(GSQG_INSTANTIATE_COMMON)
#define GSQG_EXTERN(N_)                                                              \
  extern template hipError_t ck_by_job<N_>(int, const CKArgs&, int, hipStream_t, \
                                           const LaunchEv*, int);                \
  extern template hipError_t rk_by_job<N_>(int, const RKArgs&, int, hipStream_t, \
                                           const LaunchEv*, int);
GSQG_FOR_EACH_N(GSQG_EXTERN)
#undef GSQG_EXTERN

#define GSQG_CASE(N_, CALL) \
  case N_: return CALL;

// Host-side checks before a launch: every pointer the job and stage
// dereference is set, and the launch has at least one field slice.  A
// missing one is an error here rather than an invalid address on the GPU.
bool check_ck_args(int job, const CKArgs& a, int ny) {
  if (ny < 1 || !a.tw) return false;
  if (job == CK_FWD) return a.J && a.th;
  if (job == CK_INV) return a.th && a.F[0];
  if (job != CK_THETA && job != CK_SCALAR && job != CK_VEL) return false;
  if (!a.ctrl || !a.k1d || !a.flags || !a.th) return false;
  const int e = a.e;
  if (e < 0 || e > 4) return false;
  if (e > 0 && (!a.J || !a.acc)) return false;
  if (a.ph.scheme == 0 && e >= 2 && !a.ys) return false;
  const bool forced = job == CK_THETA && a.ph.forcing;
  if (forced && e == 4 && a.n_a > 0 && (!a.a_tab || !a.a_val)) return false;
  if (a.prologue) {
    const int nf = job == CK_THETA ? 4 : 2;
    for (int q = 0; q < nf; ++q)
      if (!a.F[q]) return false;
    if (forced && a.rate_slot > 0 && !a.part) return false;
  }
  return true;
}

bool check_rk_args(int job, const RKArgs& a, int ny) {
  if (ny < 1 || !a.tw) return false;
  switch (job) {
    case RK_R2C: return a.phys_in && a.Jout;
    case RK_C2R: return a.F[0] && a.phys_out;
    case RK_VEL: return a.F[0] && a.F[1] && a.uout && a.vout;
    case RK_SCALAR: return a.F[0] && a.F[1] && a.uphys && a.vphys && a.Jout;
    case RK_THETA:
    case RK_THETA_CL:
      if (!a.F[0] || !a.F[1] || !a.F[2] || !a.F[3] || !a.Jout) return false;
      if ((a.red_slot > 0 || a.inc_step) && !a.ctrl) return false;
      if (a.red_slot > 0 && !a.part) return false;
      return job == RK_THETA || a.tau;
  }
  return false;
}

hipError_t launch_ck(int N, int job, const CKArgs& a, int ny, hipStream_t s, const LaunchEv* ev,
                     int var) {
  if (!check_ck_args(job, a, ny)) return hipErrorInvalidValue;
  switch (N) {
#define X(N_) GSQG_CASE(N_, ck_by_job<N_>(job, a, ny, s, ev, var))
    GSQG_FOR_EACH_N(X)
#undef X
  }
  return hipErrorInvalidValue;
}
hipError_t launch_rk(int N, int job, const RKArgs& a, int ny, hipStream_t s, const LaunchEv* ev,
                     int var) {
  if (!check_rk_args(job, a, ny)) return hipErrorInvalidValue;
  switch (N) {
#define X(N_) GSQG_CASE(N_, rk_by_job<N_>(job, a, ny, s, ev, var))
    GSQG_FOR_EACH_N(X)
#undef X
  }
  return hipErrorInvalidValue;
}
int ck_num_blocks(int N, int var) {
  switch (N) {
#define X(N_) GSQG_CASE(N_, ck_blocks_n<N_>(var))
    GSQG_FOR_EACH_N(X)
#undef X
  }
  return 0;
}
int shape_radix(int N, int var) {
  switch (N) {
#define X(N_) GSQG_CASE(N_, radix_n<N_>(var))
    GSQG_FOR_EACH_N(X)
#undef X
  }
  return 0;
}
int num_shape_variants(int N) {
  switch (N) {
#define X(N_) GSQG_CASE(N_, n_variants<N_>())
    GSQG_FOR_EACH_N(X)
#undef X
  }
  return 0;
}

hipError_t launch_full_to_half(const cd* X, cd* H, int N, double* norms, double* norm0,
                               hipStream_t s) {
  constexpr int T = 16;
  if (N >= T) {
    dim3 grid(N / T, ((N / 2) + T - 1) / T);
    hipLaunchKernelGGL((full_to_half_kernel<T>), grid, dim3(256), 0, s, X, H, N, norms);
  }
  hipLaunchKernelGGL(full_to_half_row0_kernel, dim3(1), dim3(256), 0, s, X, H, N, norm0);
  return hipGetLastError();
}
int full_to_half_num_blocks(int N) { return (N / 16) * (((N / 2) + 15) / 16); }

hipError_t launch_half_to_full(const cd* H, cd* X, int N, hipStream_t s) {
  constexpr int T = 16;
  dim3 grid(N / T, N / T);
  hipLaunchKernelGGL((half_to_full_kernel<T>), grid, dim3(256), 0, s, H, X, N);
  return hipGetLastError();
}

hipError_t launch_a_gather(const cd* X, const int2* modes, cd* aval, int n, int N, double* norm,
                           hipStream_t s) {
  hipLaunchKernelGGL(a_gather_kernel, dim3(1), dim3(256), 0, s, X, modes, aval, n, N, norm);
  return hipGetLastError();
}
hipError_t launch_a_scatter(cd* X, const int2* modes, const cd* aval, int n, int N,
                            hipStream_t s) {
  int blocks = (n + 255) / 256;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(a_scatter_kernel, dim3(blocks), dim3(256), 0, s, X, modes, aval, n, N);
  return hipGetLastError();
}
int diag_blocks() { return 512; }
hipError_t launch_spec_diag(const cd* X, const double* k1d, const double* edges, int N, int nb,
                            int nshell, double dk, double alpha, double* part, int stride,
                            hipStream_t s) {
  const size_t shm = (size_t)(nb + nshell + 3 * 256) * sizeof(double);
  hipLaunchKernelGGL(spec_diag_kernel, dim3(diag_blocks()), dim3(256), shm, s, X, k1d, edges, N,
                     nb, nshell, dk, alpha, part, stride);
  return hipGetLastError();
}
hipError_t launch_phys_diag(const double* th, const double* u, const double* v, long long n,
                            double* part, hipStream_t s) {
  hipLaunchKernelGGL(phys_diag_kernel, dim3(diag_blocks()), dim3(256), 0, s, th, u, v, n, part);
  return hipGetLastError();
}
hipError_t launch_cfl(const double* u, const double* v, const double2* X, long long n,
                      long long nx, double* part, hipStream_t s) {
  hipLaunchKernelGGL(cfl_kernel, dim3(diag_blocks()), dim3(256), 0, s, u, v, X, n, nx, part);
  return hipGetLastError();
}
hipError_t launch_diag_reduce(const double* part, int nblk, int stride, int nsum, double* out,
                              hipStream_t s) {
  const int blocks = (stride + 255) / 256;
  hipLaunchKernelGGL(diag_reduce_kernel, dim3(blocks), dim3(256), 0, s, part, nblk, stride, nsum,
                     out);
  return hipGetLastError();
}
hipError_t launch_a_tab(const int2* modes, const double* k1d, const uint8_t* flags, ATab* out, int n,
                        int N, hipStream_t s) {
  int blocks = (n + 255) / 256;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(a_tab_kernel, dim3(blocks), dim3(256), 0, s, modes, k1d, flags, out, n, N);
  return hipGetLastError();
}
hipError_t launch_a_rhs(const Phys& ph, const double* k1d, const uint8_t* flags, const Ctrl* ctrl,
                        const int2* modes, const cd* aval, cd* out, int n, int N, hipStream_t s) {
  int blocks = (n + 255) / 256;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(a_rhs_kernel, dim3(blocks), dim3(256), 0, s, ph, k1d, flags, ctrl, modes,
                     aval, out, n, N);
  return hipGetLastError();
}
hipError_t launch_ring_field(const cd* X, const int2* modes, int n, int N, double amp, double eps,
                             uint64_t key, cd* out, double* rate_out, cd* ou, double decay,
                             double vfac, int ou_first, hipStream_t s) {
  hipLaunchKernelGGL(ring_field_kernel, dim3(1), dim3(256), 0, s, X, modes, n, N, amp, eps, key,
                     out, rate_out, ou, decay, vfac, ou_first);
  return hipGetLastError();
}
hipError_t launch_set_ctrl(Ctrl* c, int64_t step, uint64_t key, uint64_t seed, int key_mode,
                           hipStream_t s) {
  hipLaunchKernelGGL(set_ctrl_kernel, dim3(1), dim3(1), 0, s, c, step, key, seed, key_mode);
  return hipGetLastError();
}


#endif  // GSQG_INSTANTIATE_COMMON

#if defined(GSQG_INSTANTIATE_BATCH)
// ------------------------------------------------------ batched launches
// The theta column and row passes of several unforced, scalar-free plans
// (the groups of an uneven ensemble) in one launch per RK4 stage.  Every
// plan runs the default radix-8 tile (shape variant 0), whose blocks are 128
// threads for 64 <= N <= 1024, so one launch shape fits every group; a block
// finds its plan from the block table and runs that plan's body.
constexpr int kBatchNT = 128;
template <int N>
constexpr bool batch_n_ok() {
  return CKShape<N, 0>::NT == kBatchNT && RKShape<N, 0>::NT == kBatchNT;
}
static_assert(batch_n_ok<64>() && batch_n_ok<128>() && batch_n_ok<256>() && batch_n_ok<512>() &&
                  batch_n_ok<1024>(),
              "batched tiles are 128 threads at every supported N");
constexpr int cmax(int a, int b) { return a > b ? a : b; }
template <template <int, int, int> class LS, int JOB>
struct BatchLds {
  static constexpr int FIELDS =
      cmax(cmax(cmax(LS<64, JOB, 0>::FIELDS, LS<128, JOB, 0>::FIELDS),
                cmax(LS<256, JOB, 0>::FIELDS, LS<512, JOB, 0>::FIELDS)),
           LS<1024, JOB, 0>::FIELDS);
  static constexpr int TW = cmax(cmax(cmax(LS<64, JOB, 0>::TW, LS<128, JOB, 0>::TW),
                                      cmax(LS<256, JOB, 0>::TW, LS<512, JOB, 0>::TW)),
                                 LS<1024, JOB, 0>::TW);
};
static_assert(sizeof(CKBatch) <= 4096 && sizeof(RKBatch) <= 4096, "kernel argument size");

__device__ __forceinline__ int batch_plan(const int* blk0, int n, int b) {
  int g = 0;
  while (g + 1 < n && b >= blk0[g + 1]) ++g;
  return g;
}

// NMAX: the largest grid size the kernel holds a body for.  The register
// allocation is the largest body's, so launches without N = 1024 run the
// kernel without that body (two waves per SIMD instead of one).
template <int E, int NMAX>
__global__ void __launch_bounds__(kBatchNT) ck_batch_kernel(CKBatch m) {
  using LS = BatchLds<CKLds, CK_THETA>;
  __shared__ cd lds[LS::FIELDS];
  __shared__ cd twl[LS::TW];
  const int b = blockIdx.x;
  const int g = batch_plan(m.blk0, m.n, b);
  const int gx = m.gx[g];
  const int loc = b - m.blk0[g];
  const int by = loc / gx, bx = loc - by * gx;
  const CKArgs& a = m.a[g];
  switch (m.N[g]) {
    case 64: ck_body<64, CK_THETA, E, false, 0>(a, bx, by, gx, lds, twl); break;
    case 128: ck_body<128, CK_THETA, E, false, 0>(a, bx, by, gx, lds, twl); break;
    case 256: ck_body<256, CK_THETA, E, false, 0>(a, bx, by, gx, lds, twl); break;
    case 512: ck_body<512, CK_THETA, E, false, 0>(a, bx, by, gx, lds, twl); break;
    case 1024:
      if constexpr (NMAX >= 1024) ck_body<1024, CK_THETA, E, false, 0>(a, bx, by, gx, lds, twl);
      break;
  }
}

template <int NMAX>
__global__ void __launch_bounds__(kBatchNT) rk_batch_kernel(RKBatch m) {
  using LS = BatchLds<RKLds, RK_THETA>;
  __shared__ cd lds[LS::FIELDS];
  __shared__ cd twl[LS::TW];
  const int b = blockIdx.x;
  const int g = batch_plan(m.blk0, m.n, b);
  const int gx = m.gx[g];
  const int loc = b - m.blk0[g];
  const int by = loc / gx, bx = loc - by * gx;
  const RKArgs& a = m.a[g];
  switch (m.N[g]) {
    case 64: rk_body<64, RK_THETA, 0>(a, bx, by, lds, twl); break;
    case 128: rk_body<128, RK_THETA, 0>(a, bx, by, lds, twl); break;
    case 256: rk_body<256, RK_THETA, 0>(a, bx, by, lds, twl); break;
    case 512: rk_body<512, RK_THETA, 0>(a, bx, by, lds, twl); break;
    case 1024:
      if constexpr (NMAX >= 1024) rk_body<1024, RK_THETA, 0>(a, bx, by, lds, twl);
      break;
  }
}

bool batch_supports(int N) { return N == 64 || N == 128 || N == 256 || N == 512 || N == 1024; }

// Blocks per field slice of the column (ck) or row pass at grid size N.
int batch_blocks(int N, bool ck) {
  switch (N) {
    case 64: return ck ? 32 / CKShape<64, 0>::RPW : 64 / RKShape<64, 0>::WX;
    case 128: return ck ? 64 / CKShape<128, 0>::RPW : 128 / RKShape<128, 0>::WX;
    case 256: return ck ? 128 / CKShape<256, 0>::RPW : 256 / RKShape<256, 0>::WX;
    case 512: return ck ? 256 / CKShape<512, 0>::RPW : 512 / RKShape<512, 0>::WX;
    case 1024: return ck ? 512 / CKShape<1024, 0>::RPW : 1024 / RKShape<1024, 0>::WX;
  }
  return 0;
}

// Checks the block table against the plans before a launch: every block
// belongs to exactly one plan's launch, and the count of field slices it
// implies is what the plan's arguments are checked with.
static bool batch_slices(const int* N, const int* gx, const int* blk0, int n, bool ck, int g,
                         int* ny) {
  if (n < 1 || n > kBatchMaxPlans || blk0[0] != 0) return false;
  if (!batch_supports(N[g]) || gx[g] != batch_blocks(N[g], ck)) return false;
  const int span = blk0[g + 1] - blk0[g];
  if (span <= 0 || span % gx[g] != 0) return false;
  *ny = span / gx[g];
  return true;
}

template <class B>
static bool batch_has_1024(const B& m) {
  for (int g = 0; g < m.n; ++g)
    if (m.N[g] == 1024) return true;
  return false;
}

template <int NMAX>
static hipError_t launch_ck_batch_n(int e, const CKBatch& m, hipStream_t s) {
  const dim3 grid(m.blk0[m.n]);
  switch (e) {
    case 0: hipLaunchKernelGGL((ck_batch_kernel<0, NMAX>), grid, dim3(kBatchNT), 0, s, m); break;
    case 1: hipLaunchKernelGGL((ck_batch_kernel<1, NMAX>), grid, dim3(kBatchNT), 0, s, m); break;
    case 2: hipLaunchKernelGGL((ck_batch_kernel<2, NMAX>), grid, dim3(kBatchNT), 0, s, m); break;
    case 3: hipLaunchKernelGGL((ck_batch_kernel<3, NMAX>), grid, dim3(kBatchNT), 0, s, m); break;
    case 4: hipLaunchKernelGGL((ck_batch_kernel<4, NMAX>), grid, dim3(kBatchNT), 0, s, m); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_ck_batch(int e, const CKBatch& m, hipStream_t s) {
  if (m.n < 1 || m.n > kBatchMaxPlans) return hipErrorInvalidValue;
  for (int g = 0; g < m.n; ++g) {
    int ny = 0;
    if (!batch_slices(m.N, m.gx, m.blk0, m.n, true, g, &ny)) return hipErrorInvalidValue;
    const CKArgs& a = m.a[g];
    if (a.e != e || a.ph.forcing || !check_ck_args(CK_THETA, a, ny)) return hipErrorInvalidValue;
  }
  return batch_has_1024(m) ? launch_ck_batch_n<1024>(e, m, s) : launch_ck_batch_n<512>(e, m, s);
}

hipError_t launch_rk_batch(const RKBatch& m, hipStream_t s) {
  if (m.n < 1 || m.n > kBatchMaxPlans) return hipErrorInvalidValue;
  for (int g = 0; g < m.n; ++g) {
    int ny = 0;
    if (!batch_slices(m.N, m.gx, m.blk0, m.n, false, g, &ny)) return hipErrorInvalidValue;
    const RKArgs& a = m.a[g];
    if (a.red_slot != 0 || !check_rk_args(RK_THETA, a, ny)) return hipErrorInvalidValue;
  }
  if (batch_has_1024(m))
    hipLaunchKernelGGL(rk_batch_kernel<1024>, dim3(m.blk0[m.n]), dim3(kBatchNT), 0, s, m);
  else
    hipLaunchKernelGGL(rk_batch_kernel<512>, dim3(m.blk0[m.n]), dim3(kBatchNT), 0, s, m);
  return hipGetLastError();
}
#endif  // GSQG_INSTANTIATE_BATCH

}  // namespace gsqg
