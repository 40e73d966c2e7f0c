// Device-side building blocks shared by every gSQG kernel (gfx950, wave64).
//
// Complex FP64 helpers, the SplitMix64 phase hash used for ring forcing, and a
// Stockham radix-8 FFT that runs across L/8 lanes per transform with each
// lane holding 8 elements in registers and exchanging through LDS between
// passes.  Element ownership is the same on entry and exit: lane t of a
// transform holds elements t + m*L/8, m = 0..7, in natural order, so the first
// pass reads registers and the last pass writes registers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gsqg {

using cd = double2;
typedef double vd2 __attribute__((vector_size(16)));  // native 2-vector, same bytes as cd

__device__ __forceinline__ cd cmk(double a, double b) {
  cd r;
  r.x = a;
  r.y = b;
  return r;
}
__device__ __forceinline__ cd cadd(cd a, cd b) { return cmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cd csub(cd a, cd b) { return cmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cd cmul(cd a, cd b) {
  return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ cd cmulc(cd a, cd b) {
  return cmk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ cd cconj(cd a) { return cmk(a.x, -a.y); }
__device__ __forceinline__ cd cscale(cd a, double s) { return cmk(a.x * s, a.y * s); }
__device__ __forceinline__ cd cmul_i(cd a) { return cmk(-a.y, a.x); }    // i*a
__device__ __forceinline__ cd cmul_mi(cd a) { return cmk(a.y, -a.x); }   // -i*a

// Multiply by DIR*i (DIR = -1: forward transform sign, +1: inverse).
template <int DIR>
__device__ __forceinline__ cd rot90(cd a) {
  return DIR < 0 ? cmk(a.y, -a.x) : cmk(-a.y, a.x);
}

// ---------------------------------------------------------------- hashing
// SplitMix64 output function (Steele/Lea/Flood 2014, constants of Vigna's
// splitmix64.c).  Bit-identical to oracle/rng.py.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t phase_bits(uint64_t key, uint64_t idx) {
  return mix64(key + (idx + 1ull) * 0x9E3779B97F4A7C15ull);
}
__device__ __forceinline__ uint64_t step_key(uint64_t seed, int64_t n) {
  return mix64(seed + (uint64_t)(n + 1) * 0xD1B54A32D192ED03ull);
}
// u = (bits >> 11) * 2^-53 in [0, 1); the phase is phi = 2 pi u.
__device__ __forceinline__ double phase_u(uint64_t key, uint64_t idx) {
  return (double)(phase_bits(key, idx) >> 11) * 0x1.0p-53;
}

// ------------------------------------------------------- block reduction
// Sum of one double per thread over a block of NT threads, finished in
// wave 0: the levels whose partner lies in another wave go through LDS
// (red, NT doubles) with a barrier each; the last ones (partner offset < 64)
// are shuffles down the wave, with no barrier.  Every level adds
// red[i] + red[i + s] exactly as a tree reduction over LDS does, so the sum
// is the same bit for bit.  All threads of the block must call it; the
// result is valid in thread 0.  The caller has made red free for writing.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red, int tid) {
  if constexpr (NT > 64) {
    red[tid] = v;
    __syncthreads();
#pragma unroll
    for (int s = NT / 2; s >= 64; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    v = red[tid];  // wave 0's lanes
  }
  if (tid < 64) {
#pragma unroll
    for (int s = (NT < 64 ? NT : 64) / 2; s > 0; s >>= 1) v += __shfl_down(v, s);
  }
  return v;
}

// ---------------------------------------------------------------- FFT
// LDS index with one pad element per 8 to break the stride-8 bank pattern of
// the first Stockham pass.
__device__ __forceinline__ int lpad(int e) { return e + (e >> 3); }
template <int L>
struct Padded {
  static constexpr int value = L + L / 8;
};

template <int DIR>
__device__ __forceinline__ void dft2(cd& a, cd& b) {
  cd t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(cd& a0, cd& a1, cd& a2, cd& a3) {
  cd s02 = cadd(a0, a2), d02 = csub(a0, a2);
  cd s13 = cadd(a1, a3), d13 = rot90<DIR>(csub(a1, a3));
  a0 = cadd(s02, s13);
  a2 = csub(s02, s13);
  a1 = cadd(d02, d13);
  a3 = csub(d02, d13);
}

// In-register 8-point DFT, natural order in and out.
template <int DIR>
__device__ __forceinline__ void dft8(cd (&v)[8]) {
  const double r2 = 0.70710678118654752440;  // 1/sqrt(2)
  cd e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cd o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  // twiddles W8^k, W8 = exp(DIR * i*pi/4)
  cd t1, t3;
  if (DIR < 0) {
    t1 = cmk((o1.x + o1.y) * r2, (o1.y - o1.x) * r2);      // (1-i)/sqrt2
    t3 = cmk((-o3.x + o3.y) * r2, (-o3.y - o3.x) * r2);    // (-1-i)/sqrt2
  } else {
    t1 = cmk((o1.x - o1.y) * r2, (o1.y + o1.x) * r2);      // (1+i)/sqrt2
    t3 = cmk((-o3.x - o3.y) * r2, (o3.x - o3.y) * r2);     // (-1+i)/sqrt2
  }
  cd t2 = rot90<DIR>(o2);
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, t1);
  v[5] = csub(e1, t1);
  v[2] = cadd(e2, t2);
  v[6] = csub(e2, t2);
  v[3] = cadd(e3, t3);
  v[7] = csub(e3, t3);
}

// In-register 16-point DFT, natural order in and out: two 8-point DFTs of
// the even and odd elements, combined with W16^k = exp(DIR i pi k / 8).
template <int DIR>
__device__ __forceinline__ void dft16(cd (&v)[16]) {
  const double c1 = 0.92387953251128675613;  // cos(pi/8)
  const double s1 = 0.38268343236508977173;  // sin(pi/8)
  const double r2 = 0.70710678118654752440;  // 1/sqrt(2)
  cd e[8], o[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    e[k] = v[2 * k];
    o[k] = v[2 * k + 1];
  }
  dft8<DIR>(e);
  dft8<DIR>(o);
  // W16^k for k = 1..7 with DIR < 0 (the inverse uses the conjugates):
  // (c1,-s1) (r2,-r2) (s1,-c1) (0,-1) (-s1,-c1) (-r2,-r2) (-c1,-s1)
  const double sg = DIR < 0 ? -1.0 : 1.0;
  cd t[8];
  t[0] = o[0];
  t[1] = cmul(o[1], cmk(c1, sg * s1));
  t[2] = cmul(o[2], cmk(r2, sg * r2));
  t[3] = cmul(o[3], cmk(s1, sg * c1));
  t[4] = rot90<DIR>(o[4]);
  t[5] = cmul(o[5], cmk(-s1, sg * c1));
  t[6] = cmul(o[6], cmk(-r2, sg * r2));
  t[7] = cmul(o[7], cmk(-c1, sg * s1));
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = cadd(e[k], t[k]);
    v[k + 8] = csub(e[k], t[k]);
  }
}

// Twiddle tables: tw[T + k] = exp(-2 pi i k / T) for every power of two T.
template <int DIR>
__device__ __forceinline__ cd twiddle(const cd* __restrict__ tw, int T, int k) {
  cd w = tw[T + k];
  return DIR < 0 ? w : cconj(w);
}

// x[r] *= w^r, r = 1..7, from w, w^2, w^4 with the products of the twiddle
// sources' pass(): w^3 = w w^2, w^5 = w w^4, w^6 = w^2 w^4, w^7 = w^3 w^4.
__device__ __forceinline__ void apply_w124(cd w1, cd w2, cd w4, cd (&x)[8]) {
  x[1] = cmul(x[1], w1);
  x[2] = cmul(x[2], w2);
  x[4] = cmul(x[4], w4);
  const cd w3 = cmul(w1, w2);
  x[3] = cmul(x[3], w3);
  x[5] = cmul(x[5], cmul(w1, w4));
  x[6] = cmul(x[6], cmul(w2, w4));
  x[7] = cmul(x[7], cmul(w3, w4));
}

// x[r] *= w^r, r = 1..15, from w, w^2, w^4, w^8 with the products of
// pass16(): w^3 = w w^2, w^5 = w w^4, w^6 = w^2 w^4, w^7 = w^3 w^4 and
// w^(r+8) = w^r w^8.
__device__ __forceinline__ void apply_w1248(cd w1, cd w2, cd w4, cd w8, cd (&x)[16]) {
  x[8] = cmul(x[8], w8);
  x[1] = cmul(x[1], w1);
  x[9] = cmul(x[9], cmul(w1, w8));
  x[2] = cmul(x[2], w2);
  x[10] = cmul(x[10], cmul(w2, w8));
  x[4] = cmul(x[4], w4);
  x[12] = cmul(x[12], cmul(w4, w8));
  const cd w3 = cmul(w1, w2);
  x[3] = cmul(x[3], w3);
  x[11] = cmul(x[11], cmul(w3, w8));
  const cd w5 = cmul(w1, w4);
  x[5] = cmul(x[5], w5);
  x[13] = cmul(x[13], cmul(w5, w8));
  const cd w6 = cmul(w2, w4);
  x[6] = cmul(x[6], w6);
  x[14] = cmul(x[14], cmul(w6, w8));
  const cd w7 = cmul(w3, w4);
  x[7] = cmul(x[7], w7);
  x[15] = cmul(x[15], cmul(w7, w8));
}

// Twiddle sources.  Radix-8 pass p >= 1 (Ns = 8^p) needs w, w^2, w^4 with
// w = exp(DIR 2 pi i jm / (8 Ns)), jm = t mod Ns; the final radix-2 (radix-4)
// pass needs exp(DIR 2 pi i r j / L) for r = 1 (r = 1, 2, 3).  `base` is the
// offset of the pass in the compact layout (ignored by the global table).
struct TwGlobal {
  const cd* __restrict__ tw;  // tw[T + k] = exp(-2 pi i k / T)
  // w[r] = W^(r jm), r = 1..7: three table reads (w, w^2, w^4) and four
  // products, which keeps the live registers low.
  template <int DIR>
  __device__ __forceinline__ void pass(int Ns, int jm, cd (&w)[8]) const {
    w[1] = twiddle<DIR>(tw, Ns * 8, jm);
    w[2] = twiddle<DIR>(tw, Ns * 8, 2 * jm);
    w[4] = twiddle<DIR>(tw, Ns * 8, 4 * jm);
    w[3] = cmul(w[1], w[2]);
    w[5] = cmul(w[1], w[4]);
    w[6] = cmul(w[2], w[4]);
    w[7] = cmul(w[3], w[4]);
  }
  // x[r] *= W^(r jm) with the same products as pass(), formed where each is
  // used: three twiddles live instead of seven (one transform at a time).
  template <int DIR>
  __device__ __forceinline__ void apply(int Ns, int jm, cd (&x)[8]) const {
    const cd w1 = twiddle<DIR>(tw, Ns * 8, jm);
    const cd w2 = twiddle<DIR>(tw, Ns * 8, 2 * jm);
    const cd w4 = twiddle<DIR>(tw, Ns * 8, 4 * jm);
    apply_w124(w1, w2, w4, x);
  }
  // Radix-16 pass: w[r] = W^(r jm), W = exp(DIR 2 pi i / (16 Ns)), from
  // four table reads (w, w^2, w^4, w^8) and eleven products.
  template <int DIR>
  __device__ __forceinline__ void pass16(int Ns, int jm, cd (&w)[16]) const {
    w[1] = twiddle<DIR>(tw, Ns * 16, jm);
    w[2] = twiddle<DIR>(tw, Ns * 16, 2 * jm);
    w[4] = twiddle<DIR>(tw, Ns * 16, 4 * jm);
    w[8] = twiddle<DIR>(tw, Ns * 16, 8 * jm);
    w[3] = cmul(w[1], w[2]);
    w[5] = cmul(w[1], w[4]);
    w[6] = cmul(w[2], w[4]);
    w[7] = cmul(w[3], w[4]);
#pragma unroll
    for (int r = 1; r < 8; ++r) w[r + 8] = cmul(w[r], w[8]);
  }
  template <int DIR>
  __device__ __forceinline__ void apply16(int Ns, int jm, cd (&x)[16]) const {
    apply_w1248(twiddle<DIR>(tw, Ns * 16, jm), twiddle<DIR>(tw, Ns * 16, 2 * jm),
                twiddle<DIR>(tw, Ns * 16, 4 * jm), twiddle<DIR>(tw, Ns * 16, 8 * jm), x);
  }
  // Radix-4 pass: W^(r jm), r = 1..3, W = exp(DIR 2 pi i / (4 Ns)).
  template <int DIR>
  __device__ __forceinline__ void pass4(int Ns, int jm, cd& w1, cd& w2, cd& w3) const {
    w1 = twiddle<DIR>(tw, Ns * 4, jm);
    w2 = twiddle<DIR>(tw, Ns * 4, 2 * jm);
    w3 = twiddle<DIR>(tw, Ns * 4, 3 * jm);
  }
  template <int DIR>
  __device__ __forceinline__ cd fin2(int, int L, int j) const {
    return twiddle<DIR>(tw, L, j);
  }
  template <int DIR>
  __device__ __forceinline__ void fin4(int, int L, int j, cd& w1, cd& w2, cd& w3) const {
    w1 = twiddle<DIR>(tw, L, j);
    w2 = twiddle<DIR>(tw, L, 2 * j);
    w3 = twiddle<DIR>(tw, L, 3 * j);
  }
};

// Exact twiddles from an octant table oct[r] = W^r, r = 0 .. LT/8, with
// W = exp(-2 pi i / LT), resident in LDS (reads wait on lgkmcnt only, so they
// never wait behind the block's outstanding global stores).  W^k for any
// 0 <= k < LT follows exactly from the quarter turn W^(LT/4) = -i and
// W^(LT/4 - r) = -i conj(W^r): every value equals the full table's entry.
//
// Entry r is stored at r + r/8: the radix-8 passes read entries in strides
// of 2^n, which without the pad fall into one LDS bank (and serialise); the
// pad spreads every such set of eight over distinct banks.
template <int LT>
struct OctTab {
  static constexpr int ENTRIES = LT / 8 + 1;
  static constexpr int SIZE = ENTRIES + ENTRIES / 8 + 1;  // LDS elements
  const cd* t;
  __device__ __forceinline__ static int slot(int r) { return r + (r >> 3); }
  __device__ __forceinline__ cd fwd(int k) const {
    constexpr int QL = LT / 4;
    const int q = k / QL;
    const int r = k - q * QL;
    const bool lo = r <= LT / 8;
    const cd c = t[slot(lo ? r : QL - r)];
    const cd w = lo ? c : cmk(-c.y, -c.x);
    if (q == 0) return w;
    if (q == 1) return cmk(w.y, -w.x);
    if (q == 2) return cmk(-w.x, -w.y);
    return cmk(-w.y, w.x);
  }
  __device__ __forceinline__ static void fill(cd* dst, const cd* __restrict__ tw, int tid, int nt) {
    for (int e = tid; e < ENTRIES; e += nt) dst[slot(e)] = tw[LT + e];
  }
};

// Twiddles of a length-L transform from the octant table of LT (L | LT):
// W_L^k = W_LT^(k LT/L).
template <int L, int LT>
struct TwOct {
  OctTab<LT> o;
  template <int DIR>
  __device__ __forceinline__ static cd dir(cd w) {
    return DIR < 0 ? w : cconj(w);
  }
  // w[r] = W^(r jm), r = 1..7: three exact table values (w, w^2, w^4) and
  // four products, which keeps the live registers low.
  template <int DIR>
  __device__ __forceinline__ void pass(int Ns, int jm, cd (&w)[8]) const {
    const int s = LT / (8 * Ns);
    w[1] = dir<DIR>(o.fwd(jm * s));
    w[2] = dir<DIR>(o.fwd(2 * jm * s));
    w[4] = dir<DIR>(o.fwd(4 * jm * s));
    w[3] = cmul(w[1], w[2]);
    w[5] = cmul(w[1], w[4]);
    w[6] = cmul(w[2], w[4]);
    w[7] = cmul(w[3], w[4]);
  }
  template <int DIR>
  __device__ __forceinline__ void apply(int Ns, int jm, cd (&x)[8]) const {
    const int s = LT / (8 * Ns);
    apply_w124(dir<DIR>(o.fwd(jm * s)), dir<DIR>(o.fwd(2 * jm * s)), dir<DIR>(o.fwd(4 * jm * s)),
               x);
  }
  template <int DIR>
  __device__ __forceinline__ void pass16(int Ns, int jm, cd (&w)[16]) const {
    const int s = LT / (16 * Ns);
    w[1] = dir<DIR>(o.fwd(jm * s));
    w[2] = dir<DIR>(o.fwd(2 * jm * s));
    w[4] = dir<DIR>(o.fwd(4 * jm * s));
    w[8] = dir<DIR>(o.fwd(8 * jm * s));
    w[3] = cmul(w[1], w[2]);
    w[5] = cmul(w[1], w[4]);
    w[6] = cmul(w[2], w[4]);
    w[7] = cmul(w[3], w[4]);
#pragma unroll
    for (int r = 1; r < 8; ++r) w[r + 8] = cmul(w[r], w[8]);
  }
  template <int DIR>
  __device__ __forceinline__ void apply16(int Ns, int jm, cd (&x)[16]) const {
    const int s = LT / (16 * Ns);
    apply_w1248(dir<DIR>(o.fwd(jm * s)), dir<DIR>(o.fwd(2 * jm * s)), dir<DIR>(o.fwd(4 * jm * s)),
                dir<DIR>(o.fwd(8 * jm * s)), x);
  }
  // Radix-4 pass: one table value w = W^jm, W = exp(DIR 2 pi i / (4 Ns)),
  // and its powers w^2 = w w, w^3 = w w^2 (two products instead of two more
  // octant lookups, whose index arithmetic costs more than the products).
  template <int DIR>
  __device__ __forceinline__ void pass4(int Ns, int jm, cd& w1, cd& w2, cd& w3) const {
    w1 = dir<DIR>(o.fwd(jm * (LT / (4 * Ns))));
    w2 = cmul(w1, w1);
    w3 = cmul(w1, w2);
  }
  template <int DIR>
  __device__ __forceinline__ cd fin2(int, int, int j) const {
    return dir<DIR>(o.fwd(j * (LT / L)));
  }
  template <int DIR>
  __device__ __forceinline__ void fin4(int, int, int j, cd& w1, cd& w2, cd& w3) const {
    constexpr int s = LT / L;
    w1 = dir<DIR>(o.fwd(j * s));
    w2 = dir<DIR>(o.fwd(2 * j * s));
    w3 = dir<DIR>(o.fwd(3 * j * s));
  }
};

// G independent length-L transforms over L/8 lanes each (lane index t in
// [0, L/8)), advanced in lockstep: transform g keeps its elements in v[g]
// (lane t holds elements t + m*L/8, m = 0..7, natural order on entry and
// exit) and exchanges through the LDS region buf + g * gs.  The G transforms
// share one set of barriers and one set of twiddle reads.  All lanes of the
// block must call it together (it contains block barriers).
template <int L, int DIR, int G, class TW>
__device__ __forceinline__ void fft_core8(cd (&v)[G][8], cd* buf, int gs, int t, const TW& tws) {
  static_assert(L >= 8 && (L & (L - 1)) == 0, "power of two >= 8");
  constexpr int Q = L / 8;
  constexpr int LOG = __builtin_ctz(L);
  constexpr int P8 = LOG / 3;
  constexpr int REM = LOG % 3;
  int Ns = 1;
#pragma unroll
  for (int p = 0; p < P8; ++p) {
    if (p > 0) {
#pragma unroll
      for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int m = 0; m < 8; ++m) v[g][m] = buf[g * gs + lpad(t + m * Q)];
      }
    }
    if (Ns > 1) {
      const int jm = t % Ns;
      if constexpr (G == 1) {
        tws.template apply<DIR>(Ns, jm, v[0]);
      } else {
        cd w[8];
        tws.template pass<DIR>(Ns, jm, w);
#pragma unroll
        for (int g = 0; g < G; ++g) {
#pragma unroll
          for (int r = 1; r < 8; ++r) v[g][r] = cmul(v[g][r], w[r]);
        }
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) dft8<DIR>(v[g]);
    const bool last = (p == P8 - 1) && (REM == 0);
    if (!last) {
      __syncthreads();
      const int base = (t / Ns) * Ns * 8 + (t % Ns);
#pragma unroll
      for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int r = 0; r < 8; ++r) buf[g * gs + lpad(base + r * Ns)] = v[g][r];
      }
      __syncthreads();
    }
    Ns *= 8;
  }
  if (REM != 0 && P8 > 0) {
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
      for (int m = 0; m < 8; ++m) v[g][m] = buf[g * gs + lpad(t + m * Q)];
    }
  }
  if (REM == 1) {  // radix-2 final pass, Ns = L/2, 4 butterflies per lane
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const cd w = tws.template fin2<DIR>(0, L, t + q * Q);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        cd a = v[g][q], b = cmul(v[g][q + 4], w);
        dft2<DIR>(a, b);
        v[g][q] = a;
        v[g][q + 4] = b;
      }
    }
  } else if (REM == 2) {  // radix-4 final pass, Ns = L/4, 2 butterflies per lane
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      cd w1, w2, w3;
      tws.template fin4<DIR>(0, L, t + q * Q, w1, w2, w3);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        cd a0 = v[g][q], a1 = cmul(v[g][q + 2], w1), a2 = cmul(v[g][q + 4], w2),
           a3 = cmul(v[g][q + 6], w3);
        dft4<DIR>(a0, a1, a2, a3);
        v[g][q] = a0;
        v[g][q + 2] = a1;
        v[g][q + 4] = a2;
        v[g][q + 6] = a3;
      }
    }
  }
}

// The same over L/16 lanes with 16 elements per lane (lane t holds elements
// t + m*L/16, m = 0..15): radix-16 Stockham passes, then a radix-2, -4 or -8
// pass for the remaining factor of L, whose butterflies each lane does in
// place on registers q + s*16/F.
template <int L, int DIR, int G, class TW>
__device__ __forceinline__ void fft_core16(cd (&v)[G][16], cd* buf, int gs, int t, const TW& tws) {
  static_assert(L >= 16 && (L & (L - 1)) == 0, "power of two >= 16");
  constexpr int Q = L / 16;
  constexpr int LOG = __builtin_ctz(L);
  constexpr int P16 = LOG / 4;
  constexpr int REM = LOG % 4;
  int Ns = 1;
#pragma unroll
  for (int p = 0; p < P16; ++p) {
    if (p > 0) {
#pragma unroll
      for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int m = 0; m < 16; ++m) v[g][m] = buf[g * gs + lpad(t + m * Q)];
      }
    }
    if (Ns > 1) {
      const int jm = t % Ns;
      if constexpr (G == 1) {
        tws.template apply16<DIR>(Ns, jm, v[0]);
      } else {
        cd w[16];
        tws.template pass16<DIR>(Ns, jm, w);
#pragma unroll
        for (int g = 0; g < G; ++g) {
#pragma unroll
          for (int r = 1; r < 16; ++r) v[g][r] = cmul(v[g][r], w[r]);
        }
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) dft16<DIR>(v[g]);
    const bool last = (p == P16 - 1) && (REM == 0);
    if (!last) {
      __syncthreads();
      const int base = (t / Ns) * Ns * 16 + (t % Ns);
#pragma unroll
      for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int r = 0; r < 16; ++r) buf[g * gs + lpad(base + r * Ns)] = v[g][r];
      }
      __syncthreads();
    }
    Ns *= 16;
  }
  if (REM != 0 && P16 > 0) {
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
      for (int m = 0; m < 16; ++m) v[g][m] = buf[g * gs + lpad(t + m * Q)];
    }
  }
  if (REM == 1) {  // radix 2, Ns = L/2: pairs (q, q + 8)
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const cd w = tws.template fin2<DIR>(0, L, t + q * Q);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        cd a = v[g][q], b = cmul(v[g][q + 8], w);
        dft2<DIR>(a, b);
        v[g][q] = a;
        v[g][q + 8] = b;
      }
    }
  } else if (REM == 2) {  // radix 4, Ns = L/4: (q, q + 4, q + 8, q + 12)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      cd w1, w2, w3;
      tws.template fin4<DIR>(0, L, t + q * Q, w1, w2, w3);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        cd a0 = v[g][q], a1 = cmul(v[g][q + 4], w1), a2 = cmul(v[g][q + 8], w2),
           a3 = cmul(v[g][q + 12], w3);
        dft4<DIR>(a0, a1, a2, a3);
        v[g][q] = a0;
        v[g][q + 4] = a1;
        v[g][q + 8] = a2;
        v[g][q + 12] = a3;
      }
    }
  } else if (REM == 3) {  // radix 8, Ns = L/8: (q, q + 2, ..., q + 14)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      cd w[8];
      tws.template pass<DIR>(L / 8, t + q * Q, w);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        cd a[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) a[s] = v[g][q + 2 * s];
        if constexpr (G == 1)
          tws.template apply<DIR>(L / 8, t + q * Q, a);
        else
          for (int s = 1; s < 8; ++s) a[s] = cmul(a[s], w[s]);
        dft8<DIR>(a);
#pragma unroll
        for (int s = 0; s < 8; ++s) v[g][q + 2 * s] = a[s];
      }
    }
  }
}

// 4 x 4 transpose across the four 16-lane rows of a wave: afterwards lane
// (row r, l) holds in x[h] what lane (row h, l) held in x[r].  Two rounds of
// the gfx950 row swaps, one 32-bit word at a time: v_permlane32_swap trades
// rows 2-3 of its first operand for rows 0-1 of its second, and
// v_permlane16_swap the odd rows of its first for the even rows of its
// second.  Registers only: no LDS, no barrier, no arithmetic (the values
// move bit for bit).  Every lane of the wave must be active.
__device__ __forceinline__ unsigned cd_word(const cd& z, int c) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(c < 2 ? z.x : z.y);
  return (c & 1) ? (unsigned)(u >> 32) : (unsigned)u;
}
__device__ __forceinline__ void xpose_rows4(cd (&x)[4]) {
  unsigned w[4][4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const auto a = __builtin_amdgcn_permlane32_swap(cd_word(x[0], c), cd_word(x[2], c), false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(cd_word(x[1], c), cd_word(x[3], c), false, false);
    const auto lo = __builtin_amdgcn_permlane16_swap(a[0], b[0], false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(a[1], b[1], false, false);
    w[0][c] = lo[0];
    w[1][c] = lo[1];
    w[2][c] = hi[0];
    w[3][c] = hi[1];
  }
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    x[h].x = __longlong_as_double((long long)(((unsigned long long)w[h][1] << 32) | w[h][0]));
    x[h].y = __longlong_as_double((long long)(((unsigned long long)w[h][3] << 32) | w[h][2]));
  }
}

// The same over L/4 lanes with 4 elements per lane (lane t holds elements
// t + m*L/4, m = 0..3): radix-4 Stockham passes, then a radix-2 pass where
// log2 L is odd (pairs q, q + 2).  Twice the lanes of the radix-8 layout for
// the same transform: at grid sizes where one radix-8 wave per SIMD is all a
// launch has, this gives each SIMD two waves to switch between.
//
// L = 256 (Q = 64): one transform per wave, lane t of the transform being
// lane t of its wave (every caller lays lanes out as tid = b * Q + t).  The
// exchange before the last pass (Ns = 16 -> 64) sends element r of lane
// (row h, l) to element h of lane (row r, l): a transpose across the wave's
// 16-lane rows, done in registers with xpose_rows4 instead of through LDS.
template <int L, int DIR, int G, class TW>
__device__ __forceinline__ void fft_core4(cd (&v)[G][4], cd* buf, int gs, int t, const TW& tws) {
  static_assert(L >= 4 && (L & (L - 1)) == 0, "power of two >= 4");
  constexpr int Q = L / 4;
  constexpr int LOG = __builtin_ctz(L);
  constexpr int P4 = LOG / 2;
  constexpr int REM = LOG % 2;
  constexpr bool XROWS = (Q == 64) && (REM == 0);  // last exchange across rows
  int Ns = 1;
#pragma unroll
  for (int p = 0; p < P4; ++p) {
    if (p > 0 && !(XROWS && p == P4 - 1)) {
#pragma unroll
      for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int m = 0; m < 4; ++m) v[g][m] = buf[g * gs + lpad(t + m * Q)];
      }
    }
    if (Ns > 1) {
      cd w1, w2, w3;
      tws.template pass4<DIR>(Ns, t % Ns, w1, w2, w3);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        v[g][1] = cmul(v[g][1], w1);
        v[g][2] = cmul(v[g][2], w2);
        v[g][3] = cmul(v[g][3], w3);
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) dft4<DIR>(v[g][0], v[g][1], v[g][2], v[g][3]);
    const bool last = (p == P4 - 1) && (REM == 0);
    if (XROWS && p == P4 - 2) {
#pragma unroll
      for (int g = 0; g < G; ++g) xpose_rows4(v[g]);
    } else if (!last) {
      __syncthreads();
      const int base = (t / Ns) * Ns * 4 + (t % Ns);
#pragma unroll
      for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int r = 0; r < 4; ++r) buf[g * gs + lpad(base + r * Ns)] = v[g][r];
      }
      __syncthreads();
    }
    Ns *= 4;
  }
  if (REM == 1) {  // radix 2, Ns = L/2: pairs (q, q + 2)
    if (P4 > 0) {
#pragma unroll
      for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int m = 0; m < 4; ++m) v[g][m] = buf[g * gs + lpad(t + m * Q)];
      }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const cd w = tws.template fin2<DIR>(0, L, t + q * Q);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        cd a = v[g][q], b = cmul(v[g][q + 2], w);
        dft2<DIR>(a, b);
        v[g][q] = a;
        v[g][q + 2] = b;
      }
    }
  }
}

// Dispatch on the elements per lane (4, 8 or 16), deduced from v.
template <int L, int DIR, int G, class TW>
__device__ __forceinline__ void fft_core(cd (&v)[G][4], cd* buf, int gs, int t, const TW& tws) {
  fft_core4<L, DIR, G>(v, buf, gs, t, tws);
}
template <int L, int DIR, int G, class TW>
__device__ __forceinline__ void fft_core(cd (&v)[G][8], cd* buf, int gs, int t, const TW& tws) {
  fft_core8<L, DIR, G>(v, buf, gs, t, tws);
}
template <int L, int DIR, int G, class TW>
__device__ __forceinline__ void fft_core(cd (&v)[G][16], cd* buf, int gs, int t, const TW& tws) {
  fft_core16<L, DIR, G>(v, buf, gs, t, tws);
}

// One transform with twiddles from the global table; `buf` is its LDS
// region of Padded<L> elements.
template <int L, int DIR, int R>
__device__ __forceinline__ void fft_lanes(cd (&v)[R], cd* buf, int t, const cd* __restrict__ tw) {
  fft_core<L, DIR, 1>(*reinterpret_cast<cd(*)[1][R]>(&v), buf, 0, t, TwGlobal{tw});
}

}  // namespace gsqg
