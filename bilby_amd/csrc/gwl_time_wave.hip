// Time-marginalised TaylorF2 likelihood with one wavefront per parameter row
// (gfx950).
//
// The transform is the one k_time computes: d_inner_h_array = 4/T FFT_N of
// h conj(d) / S at the time-prior window (bilby/gw/likelihood/base.py:325-330),
// as a decimated two-stage DFT (element i of sub-sequence n2 is bin
// n = Q i + n2; 512-point transforms of the Q sub-sequences, then
// X[j] = sum_n2 W_N^(n2 j) Y_n2[j mod 512]).  It is organised for sources
// whose support [kb, ke] is a small part of the band (TaylorF2 ends at the
// ISCO frequency):
//
//  k_rowconst : per row, the source constants (phase coefficients, detector
//               amplitudes and delays, support) and <h|h> from per-event prefix
//               sums of f^-7/3 sigma^2 / S (gw/utils.py:118-138), one thread
//               per row.
//  k_time_w   : one wavefront per (row, group of sub-sequences).  For each
//               sub-sequence the lanes holding in-support elements evaluate
//               the integrand; the 512-point transform runs as three radix-8
//               passes with the data in registers (lane l holds elements
//               l + 64 s) and two LDS transposes private to the wavefront, so
//               the sub-sequence loop has no workgroup barrier; each lane then
//               accumulates the window bins j = l + 64 s (mod 512) in float64.
//               With one group per row the marginalisation epilogue
//               (base.py:775-820) runs in the same wavefront.
#include "gwl_internal.h"

#include <algorithm>

namespace gwl {

constexpr int kWP = 512;      // transform length
constexpr int kWaves = 4;     // wavefronts (rows) per workgroup
constexpr int kXch = 576;     // per-wave exchange buffer, padded (float2)
constexpr int kTw3 = 7 * 64;  // pass-3 twiddles W_512^(s l), [s - 1][l]
constexpr int kTw2 = 7 * 8;   // pass-2 twiddles W_64^(s k), [s - 1][k]

// LDS tables of the transform: tw[m] = W_512^m and tw3[s - 1][l] =
// W_512^(s l mod 512) (so that lane l reads consecutive entries), both
// from float64 sincospi.
__device__ inline void init_wave_tables(float2* tw, float2* tw3) {
  init_twiddles<kWP, 64 * kWaves>(tw);
  for (int k = threadIdx.x; k < kTw3; k += 64 * kWaves) {
    const int s = k / 64 + 1, l = k & 63;
    const int e = (s * l) & (kWP - 1);
    double sn, cs;
    sincospi(-2.0 * (double)e / (double)kWP, &sn, &cs);
    tw3[k] = make_float2((float)cs, (float)sn);
  }
}

// LDS tables of the row transform (k_time_w, k_time_w_multi): tw2[s - 1][k]
// = W_64^(s k) for pass 2, where lane l uses k = l mod 8 (eight consecutive
// entries per s: no bank conflicts); tw3 as above; wt[l + 64 s] = W_N^j in
// float64 for the window bin j = l + 64 s (mod 512) of lane l
// (j = jlo + ((l + 64 s - jlo) mod 512)), the Horner step of the second
// transform stage.
__device__ inline void init_row_tables(float2* tw2, float2* tw3, double2* wt, int N, int jlo) {
  for (int k = threadIdx.x; k < kTw2; k += blockDim.x) {
    const int s = k / 8 + 1, q = k & 7;
    double sn, cs;
    sincospi(-2.0 * (double)((s * q) & 63) / 64.0, &sn, &cs);
    tw2[k] = make_float2((float)cs, (float)sn);
  }
  for (int k = threadIdx.x; k < kTw3; k += blockDim.x) {
    const int s = k / 64 + 1, l = k & 63;
    const int e = (s * l) & (kWP - 1);
    double sn, cs;
    sincospi(-2.0 * (double)e / (double)kWP, &sn, &cs);
    tw3[k] = make_float2((float)cs, (float)sn);
  }
  for (int k = threadIdx.x; k < kWP; k += blockDim.x) {
    const long long j = (long long)jlo + ((k - jlo) & (kWP - 1));
    const long long e = j & (long long)(N - 1);
    double sn, cs;
    sincospi(-2.0 * (double)e / (double)N, &sn, &cs);
    wt[k] = make_double2(cs, sn);
  }
}

__device__ __forceinline__ int xpad(int e) { return e + (e >> 3); }

// Orders the LDS accesses of one wavefront (the wave executes them in order;
// this keeps the compiler from moving them across).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct RowArgs {
  EvDev ev;
  const double* params;
  int ld;
  long long n;
  RowW* rows;
  double2* c64 = nullptr;  // [n][GWL_MAX_DET] float64 detector amplitudes, or null
};

__device__ inline void rowconst_one(const EvDev& ev, const double* row, RowW* out,
                                    const Tf2Part* pre = nullptr, double2* c64 = nullptr) {
  SampleConst sc;
  compute_sample_const(ev, row, sc, pre);
  if (c64)
#pragma unroll
    for (int d = 0; d < GWL_MAX_DET; ++d)
      c64[d] = d < ev.nd ? make_double2(sc.cre[d], sc.cim[d]) : make_double2(0.0, 0.0);
  RowW w;
#pragma unroll
  for (int k = 0; k < 9; ++k) w.c[k] = sc.c[k];
  double hh = 0.0;
#pragma unroll
  for (int d = 0; d < GWL_MAX_DET; ++d) {
    const bool on = d < ev.nd;
    w.dt[d] = on ? sc.dt[d] : 0.0;
    w.cre[d] = on ? (float)sc.cre[d] : 0.f;
    w.cim[d] = on ? (float)sc.cim[d] : 0.f;
    w.cabs2[d] = on ? sc.cabs2[d] : 0.0;
    if (on && sc.kb <= sc.ke) {
      const double* ps = ev.hps + (size_t)d * (ev.N + 2);
      hh += sc.cabs2[d] * (ps[sc.ke + 1] - ps[sc.kb]);
    }
  }
  w.hh = hh;
  w.hscale = sc.hscale;
  w.jitter = sc.jitter;
  w.kb = sc.kb;
  w.ke = sc.ke;
  w.valid = sc.valid;
  {
    int lo = ev.mlo[0], hi = ev.mhi[0];
    for (int d = 1; d < ev.nd; ++d) {
      lo = min(lo, ev.mlo[d]);
      hi = max(hi, ev.mhi[d]);
    }
    w.nomask = (sc.kb <= lo && sc.ke >= hi) ? 1 : 0;
    w.pad = 0;
  }
  // reference detector: the largest |C_d|, so that |rho_k| <= 1
  int ref = 0;
  for (int d = 1; d < ev.nd && d < GWL_MAX_DET; ++d)
    if (sc.cabs2[d] > sc.cabs2[ref]) ref = d;
  const double cr = sc.cre[ref], ci = sc.cim[ref];
  const double a2 = cr * cr + ci * ci;
  w.ref = ref;
  w.amp = sqrt(a2);
  w.c4r = sc.c[4] - (a2 > 0.0 ? atan2(ci, cr) / kTwoPi : 0.0);
#pragma unroll
  for (int k = 0; k < GWL_MAX_DET - 1; ++k) {
    const int d = k < ref ? k : k + 1;
    double re = 0.0, im = 0.0;
    if (d < ev.nd && a2 > 0.0) {  // C_d conj(C_ref) / |C_ref|^2
      re = (sc.cre[d] * cr + sc.cim[d] * ci) / a2;
      im = (sc.cim[d] * cr - sc.cre[d] * ci) / a2;
    }
    w.rre[k] = (float)re;
    w.rim[k] = (float)im;
  }
  *out = w;
}

// Two lanes per row: the odd lane computes the TaylorF2 part of the row's
// constants while the even lane computes the sky part (GMST, antenna
// patterns, delays), then the even lane assembles them; the two float64
// chains run side by side instead of one after the other.
constexpr int kRowconstRows = 128;  // rows per 256-thread block

__global__ __launch_bounds__(256) void k_rowconst(RowArgs a) {
  __shared__ Tf2Part t[kRowconstRows];
  const int k = threadIdx.x >> 1;
  const long long r = (long long)blockIdx.x * kRowconstRows + k;
  const bool live = r < a.n;
  if (live && (threadIdx.x & 1)) tf2_part(a.ev, a.params + r * a.ld, t[k]);
  __syncthreads();
  if (live && !(threadIdx.x & 1))
    rowconst_one(a.ev, a.params + r * a.ld, a.rows + r, a.ev.approx == GWL_TAYLORF2 ? &t[k] : nullptr,
                 a.c64 ? a.c64 + r * GWL_MAX_DET : nullptr);
}

// The event fields the row loop reads.
struct WaveEv {
  const double* tf2b;
  const float2* dwa32;
  double df;
  int N, Q, jlo, L, nd, pad;
};

__device__ __forceinline__ WaveEv wave_ev(const EvDev& ev) {
  WaveEv w;
  w.tf2b = ev.tf2b;
  w.dwa32 = ev.dwa32;
  w.df = ev.df;
  w.N = ev.N;
  w.Q = ev.Q;
  w.jlo = ev.jlo;
  w.L = ev.L;
  w.nd = ev.nd;
  w.pad = 0;
  return w;
}

struct TimeWArgs {
  EvDev ev;
  const RowW* rows;
  long long n;
  int G;             // groups of sub-sequences per row
  int fused;         // G == 1: epilogue here
  double2* part;     // [n][G][L] window partials (G > 1)
  double* hhpart;    // [n][G]
  double* out;
  const int* only;   // if set: rows r with only[r] == 0 are skipped
};

// Psi(f) / 2pi + f dt_ref - arg(C_ref) / 2pi (cycles) of one row at a tabled
// basis point (u = f^1/3, v = f^-1/3, ln f): the nine-term sum of
// tf2_phase_cycles in Horner form with the reference detector's delay as the
// u^3 = f term and its amplitude phase in the constant, float64.
__device__ __forceinline__ double tf2_cycles_ref(const RowW& rw, double dtr, const double* tb) {
  const double u = tb[0], v = tb[1], lf = tb[2];
  // c0 v^5 + c1 v^3 + c2 v^2 + c3 v + c4r
  double a = fma(rw.c[0], v * v, rw.c[1]);
  a = fma(a, v, rw.c[2]);
  a = fma(a, v, rw.c[3]);
  double s = fma(a, v, rw.c4r);
  // + c5 u + c6 u^2 + dt_ref u^3 + (c7 + c8 u) ln f
  s = fma(fma(fma(dtr, u, rw.c[6]), u, rw.c[5]), u, s);
  return fma(fma(rw.c[8], u, rw.c[7]), lf, s);
}

// x - floor(x) in one instruction (v_fract_f64)
__device__ __forceinline__ double frac1(double x) { return __builtin_amdgcn_fract(x); }

// The transform of one row over sub-sequences [n2b, n2e): lane l returns
// the float64 window sums of bins j = l + 64 s (mod 512) in accr[s],
// acci[s].  tw2 / tw3 / wt: the LDS tables of init_row_tables; xb: this
// wavefront's exchange buffer.  nd <= ND detectors.
//
// The integrand h conj(d) / S = |C_r| exp(-2 pi i phi(f)) [conj(dw_r) +
// sum_k rho_k exp(-2 pi i f (dt_k - dt_r)) conj(dw_k)] (dw: the tabled
// dw f^-7/6), where r is the row's reference detector and phi(f) =
// Psi(f) / 2pi + f dt_r - arg(C_r) / 2pi is evaluated in float64 and reduced
// to [0, 1) cycles.  The other detectors' rotations start from an exact
// float64 phase at the lane's first element of the sub-sequence,
// f0 = (Q l + n2) df, and advance by w_k = exp(-2 pi i 64 Q df (dt_k - dt_r))
// (float64, rounded to float32) to the lane's elements l + 64 s.
//
// Second stage: X[j] = sum_n2 W_N^(n2 j) Y_n2[j mod 512] by Horner's rule
// over descending n2 (acc = acc W_N^j + Y_n2, float64, the step from the
// float64 table wt), then times W_N^(n2b j) and |C_r|.
// One sub-sequence n2 of wave_row: the integrand (elements outside the
// support zeroed when MASK), the 512-point transform and the Horner step.
template <int ND, bool MASK>
__device__ __forceinline__ void wave_step(const WaveEv& ev, const RowW& rw, int nd, int n2, int kb,
                                          int ke, double dtr, const double* dd, const float2* wd,
                                          const float2* const* dwo, const float2* dwr,
                                          const float2* tw2l, const float2* tw3, const double2* wt,
                                          float2* xb, int lane, int base2, double* accr,
                                          double* acci) {
  constexpr int NO = ND - 1;
  const int Q = ev.Q;
  // in-support elements of this sub-sequence: kb <= Q i + n2 <= ke
  const int ilo = (kb - n2 + Q - 1) / Q;
  const int ihi = ke >= n2 ? (ke - n2) / Q : -1;
  if (ilo <= ihi) {
    float2 x[8];
    const double f0 = (double)(Q * lane + n2) * ev.df;
    float2 r[NO > 0 ? NO : 1];
#pragma unroll
    for (int k = 0; k < NO; ++k)
      r[k] = cmul(make_float2(rw.rre[k], rw.rim[k]), expm2pi((float)frac1(f0 * dd[k])));
    const size_t e0 = (size_t)n2 * kWP + lane;
    const double* tb = ev.tf2b + 3 * e0;
    // every element is evaluated (the tables cover the whole sub-sequence);
    // with MASK those outside the support are zeroed, without divergence
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const double cyc = tf2_cycles_ref(rw, dtr, tb + 3 * 64 * s);
      const float2 d0 = dwr[e0 + 64 * s];
      float2 y = make_float2(d0.x, -d0.y);
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        if (k >= nd - 1) break;  // the event holds nd detectors' tables
        const float2 dw = dwo[k][e0 + 64 * s];
        y.x += r[k].x * dw.x + r[k].y * dw.y;
        y.y += r[k].y * dw.x - r[k].x * dw.y;
      }
      const float2 v = cmul(expm2pi((float)frac1(cyc)), y);
      if constexpr (MASK) {
        const int i = lane + 64 * s;
        const bool in = i >= ilo && i <= ihi;
        x[s] = make_float2(in ? v.x : 0.f, in ? v.y : 0.f);
      } else {
        x[s] = v;
      }
#pragma unroll
      for (int k = 0; k < NO; ++k) r[k] = cmul(r[k], wd[k]);
    }
    // pass 1 (span 1): inputs l + 64 s, outputs to 8 l + s.  When the
    // support ends below element 128 (uniform) only x[0], x[1] can be
    // nonzero and the butterfly is x0 + W_8^s x1 (the same arithmetic dft8
    // performs on those inputs).
    if (!MASK || ihi >= 128) {
      dft8(x);
    } else {
      const float2 x0 = x[0], a = x[1];
      const float c = 0.70710678118654752440f;
      const float p = (a.x + a.y) * c, q = (a.y - a.x) * c;
      x[0] = make_float2(x0.x + a.x, x0.y + a.y);
      x[4] = make_float2(x0.x - a.x, x0.y - a.y);
      x[1] = make_float2(x0.x + p, x0.y + q);
      x[5] = make_float2(x0.x - p, x0.y - q);
      x[2] = make_float2(x0.x + a.y, x0.y - a.x);
      x[6] = make_float2(x0.x - a.y, x0.y + a.x);
      x[3] = make_float2(x0.x + q, x0.y - p);
      x[7] = make_float2(x0.x - q, x0.y + p);
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) xb[xpad(8 * lane + s)] = x[s];
    wave_sync();
    // pass 2 (span 8): inputs l + 64 s, outputs to 64 (l / 8) + l mod 8 + 8 s
#pragma unroll
    for (int s = 0; s < 8; ++s) x[s] = xb[xpad(lane + 64 * s)];
#pragma unroll
    for (int s = 1; s < 8; ++s) x[s] = cmul(x[s], tw2l[(s - 1) * 8]);
    dft8(x);
    wave_sync();
#pragma unroll
    for (int s = 0; s < 8; ++s) xb[xpad(base2 + 8 * s)] = x[s];
    wave_sync();
    // pass 3 (span 64): inputs l + 64 s, outputs Y[l + 64 s] stay in x[s]
#pragma unroll
    for (int s = 0; s < 8; ++s) x[s] = xb[xpad(lane + 64 * s)];
#pragma unroll
    for (int s = 1; s < 8; ++s) x[s] = cmul(x[s], tw3[(s - 1) * 64 + lane]);
    dft8(x);
    wave_sync();
    // Horner step of the window bins j = l + 64 s (mod 512)
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const double2 w = wt[lane + 64 * s];
      const double ar = accr[s], ai = acci[s];
      accr[s] = fma(ar, w.x, fma(-ai, w.y, (double)x[s].x));
      acci[s] = fma(ar, w.y, fma(ai, w.x, (double)x[s].y));
    }
  } else {
    // nothing of the support in this sub-sequence (uniform): Y_n2 = 0, only
    // the Horner step
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const double2 w = wt[lane + 64 * s];
      const double ar = accr[s], ai = acci[s];
      accr[s] = fma(ar, w.x, -ai * w.y);
      acci[s] = fma(ar, w.y, ai * w.x);
    }
  }
}

// The sub-sequences [n2b, n2e) in descending order.
template <int ND, bool MASK>
__device__ __forceinline__ void wave_loop(const WaveEv& ev, const RowW& rw, int nd, int n2b,
                                          int n2e, double dtr, const double* dd, const float2* wd,
                                          const float2* const* dwo, const float2* dwr,
                                          const float2* tw2l, const float2* tw3, const double2* wt,
                                          float2* xb, int lane, double* accr, double* acci) {
  const int kb = rw.kb, ke = rw.ke;
  const int base2 = 64 * (lane >> 3) + (lane & 7);
  for (int n2 = n2e - 1; n2 >= n2b; --n2)
    wave_step<ND, MASK>(ev, rw, nd, n2, kb, ke, dtr, dd, wd, dwo, dwr, tw2l, tw3, wt, xb, lane,
                        base2, accr, acci);
}

// The transform of one row over sub-sequences [n2b, n2e): lane l returns
// the float64 window sums of bins j = l + 64 s (mod 512) in accr[s],
// acci[s].  tw2 / tw3 / wt: the LDS tables of init_row_tables; xb: this
// wavefront's exchange buffer.  nd <= ND detectors.  FAST: rows whose
// support covers every detector's mask run without element masking.
//
// The integrand h conj(d) / S = |C_r| exp(-2 pi i phi(f)) [conj(dw_r) +
// sum_k rho_k exp(-2 pi i f (dt_k - dt_r)) conj(dw_k)] (dw: the tabled
// dw f^-7/6), where r is the row's reference detector and phi(f) =
// Psi(f) / 2pi + f dt_r - arg(C_r) / 2pi is evaluated in float64 and reduced
// to [0, 1) cycles.  The other detectors' rotations start from an exact
// float64 phase at the lane's first element of the sub-sequence,
// f0 = (Q l + n2) df, and advance by w_k = exp(-2 pi i 64 Q df (dt_k - dt_r))
// (float64, rounded to float32) to the lane's elements l + 64 s.
//
// Second stage: X[j] = sum_n2 W_N^(n2 j) Y_n2[j mod 512] by Horner's rule
// over descending n2 (acc = acc W_N^j + Y_n2, float64, the step from the
// float64 table wt), then times W_N^(n2b j) and |C_r|.
template <int ND, bool FAST = true>
__device__ __forceinline__ void wave_row(const WaveEv& ev, const RowW& rw, int nd, int n2b, int n2e,
                                         const float2* tw2, const float2* tw3, const double2* wt,
                                         float2* xb, int lane, double* accr, double* acci) {
  constexpr int NO = ND - 1;  // detectors other than the reference
  const int Q = ev.Q;
  const float2* tw2l = tw2 + (lane & 7);
#pragma unroll
  for (int s = 0; s < 8; ++s) accr[s] = acci[s] = 0.0;
  const int ref = rw.ref;
  // (static indices only: rw may be a register copy)
  double dtr = rw.dt[0];
#pragma unroll
  for (int d = 1; d < ND; ++d)
    if (d == ref) dtr = rw.dt[d];
  // the others' delays relative to the reference and their steps (float64
  // once per row, rounded once to float32)
  double dd[NO > 0 ? NO : 1];
  float2 wd[NO > 0 ? NO : 1];
  const float2* dwo[NO > 0 ? NO : 1];
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    const int d = k < ref ? k : k + 1;
    dd[k] = (k < ref ? rw.dt[k] : rw.dt[k + 1]) - dtr;
    const double2 w = expm2pi(frac1((64.0 * (double)Q * ev.df) * dd[k]));
    wd[k] = make_float2((float)w.x, (float)w.y);
    dwo[k] = ev.dwa32 + (size_t)d * ev.N;
  }
  const float2* dwr = ev.dwa32 + (size_t)ref * ev.N;
  if (FAST && rw.nomask)
    wave_loop<ND, false>(ev, rw, nd, n2b, n2e, dtr, dd, wd, dwo, dwr, tw2l, tw3, wt, xb, lane,
                         accr, acci);
  else
    wave_loop<ND, true>(ev, rw, nd, n2b, n2e, dtr, dd, wd, dwo, dwr, tw2l, tw3, wt, xb, lane,
                        accr, acci);
  // the group's first sub-sequence n2b: times W_N^(n2b j); and |C_r|
  const double amp = rw.amp;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    double tr = amp, ti = 0.0;
    if (n2b > 0) {
      const long long j = (long long)ev.jlo + ((lane + 64 * s - ev.jlo) & (kWP - 1));
      const long long e = ((long long)n2b * j) & (long long)(ev.N - 1);
      const double2 t = expm2pi((double)e / (double)ev.N);
      tr = amp * t.x;
      ti = amp * t.y;
    }
    const double ar = accr[s], ai = acci[s];
    accr[s] = ar * tr - ai * ti;
    acci[s] = ar * ti + ai * tr;
  }
}

// Window partials of row r, group g (layout [row][G][L], <h|h> in group 0).
__device__ __forceinline__ void wave_store_partials(const WaveEv& ev, const RowW& rw, long long r,
                                                    int G, int g, int lane, const double* accr,
                                                    const double* acci, double2* part,
                                                    double* hhpart) {
  double2* dst = part + ((size_t)r * G + g) * ev.L;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int jj = (lane + 64 * s - ev.jlo) & (kWP - 1);
    if (jj < ev.L) dst[jj] = make_double2(accr[s], acci[s]);
  }
  if (lane == 0) hhpart[(size_t)r * G + g] = (g == 0) ? rw.hh : 0.0;
}

template <int ND>
__global__ __launch_bounds__(64 * kWaves, 4) void k_time_w(TimeWArgs a) {
  extern __shared__ double lds_d[];
  const EvDev& ev = a.ev;
  double2* wt = reinterpret_cast<double2*>(lds_d);                 // [512]
  float2* tw2 = reinterpret_cast<float2*>(wt + kWP);               // [7][8]
  float2* tw3 = tw2 + kTw2;                                        // [7][64]
  float2* xch = tw3 + kTw3;                                        // [kWaves][kXch]
  double* knots = reinterpret_cast<double*>(xch + kWaves * kXch);  // tx, ty
  const bool dist = (ev.flags & GWL_DISTANCE_MARG) != 0;
  init_row_tables(tw2, tw3, wt, ev.N, ev.jlo);
  if (a.fused && dist) stage_knots(ev, knots);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long item = (long long)blockIdx.x * kWaves + wid;
  const long long r = item / a.G;
  const int g = (int)(item - r * a.G);
  if (r >= a.n) return;
  if (a.only && !a.only[r]) return;
  const RowW& rw = a.rows[r];
  const int Q = ev.Q;
  const int n2b = (int)(((long long)g * Q) / a.G);
  const int n2e = (int)(((long long)(g + 1) * Q) / a.G);
  const WaveEv wv = wave_ev(ev);
  double accr[8], acci[8];
  wave_row<ND>(wv, rw, ND, n2b, n2e, tw2, tw3, wt, xch + wid * kXch, lane, accr, acci);

  if (!a.fused) {
    wave_store_partials(wv, rw, r, a.G, g, lane, accr, acci, a.part, a.hhpart);
    return;
  }
  // epilogue: logsumexp over the prior window (base.py:794-820)
  if (!rw.valid) {
    if (lane == 0) a.out[r] = row_valid_sentinel();
    return;
  }
  const double norm = 4.0 / ev.T;
  const double hh = rw.hh * norm, hs = rw.hscale, jit = rw.jitter;
  SplineY yy;
  if (dist) yy = spline_y(ev, knots + ev.sp.nx + 4, hh * hs * hs);
  double m = -INFINITY, sum = 0.0;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int jj = (lane + 64 * s - ev.jlo) & (kWP - 1);
    if (jj < ev.L) {
      const int j = ev.jlo + jj;
      const double t = (ev.start + (double)(j + 1) * ev.tstep) + jit;
      if (t >= ev.tmin && t <= ev.tmax)
        lse_push(m, sum, epi_value(ev, knots, yy, hh, hs, norm * accr[s], norm * acci[s]) +
                             time_log_weight(ev, t));
    }
  }
  lse_wave(m, sum);
  if (lane == 0) {
    const double b = time_lse_scale(ev, ev.tdelta);
    a.out[r] = (m > -INFINITY) ? log(b * sum) + m : -INFINITY;
  }
}

// ------------------------------------------------------------------ calibration
// Calibration and time marginalisation together (base.py:305-324, 862-877):
// for every response curve c the transform at the window of
//   sum_d cal_{c,d}(n) h_d(n) conj(d_d(n)) / S_d(n)
// and <h|h>_c = sum_d |C_d|^2 sum_n f^-7/3 sigma^2 / S_d |cal_{c,d}|^2 over
// the support (per-curve prefix sums), the time-marginalised lnL per curve,
// then the logsumexp over curves minus ln C.  One wavefront per (row, block
// of kCalCurves curves): the per-detector integrand of a sub-sequence is
// computed once and combined with each curve of the block before that
// curve's transform; the grid runs the curve blocks slowest, so a block's
// curve values stay in L2 while many rows use them.
constexpr int kCalCurves = 2;

// Per-detector h conj(d) / S at decimated element i (bin nbin), zero for
// detectors past nd (the float64 phase and float32 products of
// integrand_tf2, kept apart per detector).
template <int ND>
__device__ __forceinline__ void integrand_tf2_det(const WaveEv& ev, const RowW& rw, int i,
                                                  int nbin, float2* out) {
  const double f = (double)nbin * ev.df;
  const double* tb = ev.tf2b + 3 * (size_t)i;
  const double u = tb[0], v = tb[1], lf = tb[2];
  const double v2 = v * v, v3 = v2 * v;
  double s = rw.c[4];
  s = fma(rw.c[0], v3 * v2, s);
  s = fma(rw.c[1], v3, s);
  s = fma(rw.c[2], v2, s);
  s = fma(rw.c[3], v, s);
  s = fma(rw.c[5], u, s);
  s = fma(rw.c[6], u * u, s);
  s = fma(rw.c[7], lf, s);
  s = fma(rw.c[8], u * lf, s);
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const double cyc = fma(f, rw.dt[d], s);
    const float2 e = expm2pi((float)(cyc - floor(cyc)));
    const float2 dw = ev.dwa32[(size_t)d * ev.N + i];
    const float hr = rw.cre[d] * e.x - rw.cim[d] * e.y;
    const float hi = rw.cre[d] * e.y + rw.cim[d] * e.x;
    out[d] = make_float2(hr * dw.x + hi * dw.y, hi * dw.x - hr * dw.y);
  }
}

// The three radix-8 passes of one 512-point sub-sequence transform (inputs
// x[s] = element l + 64 s of lane l) and the window accumulation of
// sub-sequence n2 into accr / acci (bins j = l + 64 s mod 512), as in
// wave_row.
__device__ __forceinline__ void wave_transform_accumulate(const WaveEv& ev, float2* x, int n2,
                                                          const float2* tw, const float2* tw3,
                                                          float2* xb, int lane, double* accr,
                                                          double* acci) {
  const int k2 = 8 * (lane & 7);
  const int base2 = 64 * (lane >> 3) + (lane & 7);
  dft8(x);
#pragma unroll
  for (int s = 0; s < 8; ++s) xb[xpad(8 * lane + s)] = x[s];
  wave_sync();
#pragma unroll
  for (int s = 0; s < 8; ++s) x[s] = xb[xpad(lane + 64 * s)];
#pragma unroll
  for (int s = 1; s < 8; ++s) x[s] = cmul(x[s], tw[(s * k2) & (kWP - 1)]);
  dft8(x);
  wave_sync();
#pragma unroll
  for (int s = 0; s < 8; ++s) xb[xpad(base2 + 8 * s)] = x[s];
  wave_sync();
#pragma unroll
  for (int s = 0; s < 8; ++s) x[s] = xb[xpad(lane + 64 * s)];
#pragma unroll
  for (int s = 1; s < 8; ++s) x[s] = cmul(x[s], tw3[(s - 1) * 64 + lane]);
  dft8(x);
  wave_sync();
  const unsigned nmask = (unsigned)ev.N - 1u;
  const float invN = 1.0f / (float)ev.N;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int jj = (lane + 64 * s - ev.jlo) & (kWP - 1);
    if (jj < ev.L) {
      const unsigned j = (unsigned)(ev.jlo + jj);
      const unsigned idx = ((unsigned)n2 * j) & nmask;
      const float2 t = expm2pi((float)idx * invN);
      const float2 z = cmul(t, x[s]);
      accr[s] += (double)z.x;
      acci[s] += (double)z.y;
    }
  }
}

struct CalTimeArgs {
  EvDev ev;
  const RowW* rows;
  long long n;
  long long rowblocks;  // ceil(n / kWaves)
  int nc;
  const float2* cal;    // [nc][nd][N] curve values, decimated order
  const double* calps;  // [nc][nd][N + 2] prefix sums of f^-7/3 sigma^2 / S |cal|^2
  double* lnl;          // [n][nc] time-marginalised lnL per curve
  // float64 form: the curves' float32 remainders (cal = hi + lo to 48 bits),
  // the rows' float64 detector amplitudes, W_N^m (m < N)
  const float2* cal_lo;
  const double2* c64;
  const double2* wN;
};

// The float64 integrand of each detector at decimated element i (bin nbin):
// C_d exp(-2 pi i (Psi(f) / 2 pi + f dt_d)) f^-7/6 conj(d_d / S_d) sigma.
template <int ND>
__device__ __forceinline__ void integrand_tf2_det64(const WaveEv& ev, const RowW& rw,
                                                    const double2* c64, const double2* dw64,
                                                    int i, int nbin, double2* out) {
  const double f = (double)nbin * ev.df;
  const double* tb = ev.tf2b + 3 * (size_t)i;
  const double u = tb[0], v = tb[1], lf = tb[2];
  const double v2 = v * v, v3 = v2 * v;
  double s = rw.c[4];
  s = fma(rw.c[0], v3 * v2, s);
  s = fma(rw.c[1], v3, s);
  s = fma(rw.c[2], v2, s);
  s = fma(rw.c[3], v, s);
  s = fma(rw.c[5], u, s);
  s = fma(rw.c[6], u * u, s);
  s = fma(rw.c[7], lf, s);
  s = fma(rw.c[8], u * lf, s);
  const double amp = v3 * sqrt(v);  // f^-7/6 (v = f^-1/3)
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const double cyc = fma(f, rw.dt[d], s);
    const double2 h = cmul(c64[d], expm2pi(cyc - floor(cyc)));
    const double2 w = dw64[(size_t)d * ev.N + i];
    const double wr = w.x * amp, wi = w.y * amp;
    out[d] = make_double2(h.x * wr + h.y * wi, h.y * wr - h.x * wi);
  }
}

// float64 form of wave_transform_accumulate: the three radix-8 passes in
// float64 through the wave's exchange buffer, the window accumulation with
// W_N^(n2 j mod N) from the event's float64 table.
__device__ __forceinline__ void wave_transform_accumulate64(const WaveEv& ev, double2* x, int n2,
                                                            const double2* tw, const double2* tw3,
                                                            double2* xb, int lane,
                                                            const double2* wN, double* accr,
                                                            double* acci) {
  const int k2 = 8 * (lane & 7);
  const int base2 = 64 * (lane >> 3) + (lane & 7);
  dft8(x);
#pragma unroll
  for (int s = 0; s < 8; ++s) xb[xpad(8 * lane + s)] = x[s];
  wave_sync();
#pragma unroll
  for (int s = 0; s < 8; ++s) x[s] = xb[xpad(lane + 64 * s)];
#pragma unroll
  for (int s = 1; s < 8; ++s) x[s] = cmul(x[s], tw[(s * k2) & (kWP - 1)]);
  dft8(x);
  wave_sync();
#pragma unroll
  for (int s = 0; s < 8; ++s) xb[xpad(base2 + 8 * s)] = x[s];
  wave_sync();
#pragma unroll
  for (int s = 0; s < 8; ++s) x[s] = xb[xpad(lane + 64 * s)];
#pragma unroll
  for (int s = 1; s < 8; ++s) x[s] = cmul(x[s], tw3[(s - 1) * 64 + lane]);
  dft8(x);
  wave_sync();
  const unsigned nmask = (unsigned)ev.N - 1u;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int jj = (lane + 64 * s - ev.jlo) & (kWP - 1);
    if (jj < ev.L) {
      const unsigned j = (unsigned)(ev.jlo + jj);
      const double2 z = cmul(wN[((unsigned)n2 * j) & nmask], x[s]);
      accr[s] += z.x;
      acci[s] += z.y;
    }
  }
}

template <int ND>
__global__ __launch_bounds__(64 * kWaves, 2) void k_cal_time_w(CalTimeArgs a) {
  extern __shared__ double lds_d[];
  const EvDev& ev = a.ev;
  float2* tw = reinterpret_cast<float2*>(lds_d);
  float2* tw3 = tw + kWP;
  float2* xch = tw3 + kTw3;
  double* knots = reinterpret_cast<double*>(xch + kWaves * kXch);
  const bool dist = (ev.flags & GWL_DISTANCE_MARG) != 0;
  init_wave_tables(tw, tw3);
  if (dist) stage_knots(ev, knots);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long cb = blockIdx.x / a.rowblocks;
  const long long r = (blockIdx.x - cb * a.rowblocks) * kWaves + wid;
  if (r >= a.n) return;
  const int c0 = (int)cb * kCalCurves;
  const int ncb = min(kCalCurves, a.nc - c0);
  const RowW& rw = a.rows[r];
  if (!rw.valid) {
    if (lane < ncb) a.lnl[(size_t)r * a.nc + c0 + lane] = row_valid_sentinel();
    return;
  }
  const WaveEv wv = wave_ev(ev);
  const int Q = ev.Q;
  const int kb = rw.kb, ke = rw.ke;
  float2* xb = xch + wid * kXch;
  double accr[kCalCurves][8], acci[kCalCurves][8];
#pragma unroll
  for (int cc = 0; cc < kCalCurves; ++cc)
#pragma unroll
    for (int s = 0; s < 8; ++s) accr[cc][s] = acci[cc][s] = 0.0;
  for (int n2 = 0; n2 < Q; ++n2) {
    const int ilo = (kb - n2 + Q - 1) / Q;
    const int ihi = ke >= n2 ? (ke - n2) / Q : -1;
    if (ilo > ihi) continue;  // nothing of the support in this sub-sequence
    float2 xd[8][ND];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int i = lane + 64 * s;
      if (i >= ilo && i <= ihi) {
        integrand_tf2_det<ND>(wv, rw, n2 * kWP + i, Q * i + n2, xd[s]);
      } else {
#pragma unroll
        for (int d = 0; d < ND; ++d) xd[s][d] = make_float2(0.f, 0.f);
      }
    }
#pragma unroll
    for (int cc = 0; cc < kCalCurves; ++cc) {
      if (cc >= ncb) break;
      const float2* cal = a.cal + (size_t)(c0 + cc) * ND * ev.N + (size_t)n2 * kWP;
      float2 x[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        float xr = 0.f, xi = 0.f;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          const float2 k = cal[(size_t)d * ev.N + lane + 64 * s];
          xr += k.x * xd[s][d].x - k.y * xd[s][d].y;
          xi += k.x * xd[s][d].y + k.y * xd[s][d].x;
        }
        x[s] = make_float2(xr, xi);
      }
      wave_transform_accumulate(wv, x, n2, tw, tw3, xb, lane, accr[cc], acci[cc]);
    }
  }
  // epilogue per curve: time-marginalised lnL (base.py:794-820 per curve)
  const double norm = 4.0 / ev.T;
  const double hs = rw.hscale, jit = rw.jitter;
  const double b = time_lse_scale(ev, ev.tdelta);
#pragma unroll
  for (int cc = 0; cc < kCalCurves; ++cc) {
    if (cc >= ncb) break;
    const int c = c0 + cc;
    double hh = 0.0;
    if (kb <= ke) {
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        const double* ps = a.calps + ((size_t)c * ND + d) * (ev.N + 2);
        hh += rw.cabs2[d] * (ps[ke + 1] - ps[kb]);
      }
    }
    hh *= norm;
    SplineY yy;
    if (dist) yy = spline_y(ev, knots + ev.sp.nx + 4, hh * hs * hs);
    double m = -INFINITY, sum = 0.0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int jj = (lane + 64 * s - ev.jlo) & (kWP - 1);
      if (jj < ev.L) {
        const int j = ev.jlo + jj;
        const double t = (ev.start + (double)(j + 1) * ev.tstep) + jit;
        if (t >= ev.tmin && t <= ev.tmax)
          lse_push(m, sum, epi_value(ev, knots, yy, hh, hs, norm * accr[cc][s],
                                     norm * acci[cc][s]) +
                               time_log_weight(ev, t));
      }
    }
    lse_wave(m, sum);
    if (lane == 0) a.lnl[(size_t)r * a.nc + c] = (m > -INFINITY) ? log(b * sum) + m : -INFINITY;
  }
}

// The float64 form of k_cal_time_w (base.py:305-324 in float64): the
// integrand in float64 (float64 sin / cos of the phase, float64 amplitudes
// and data), curve values hi + lo, float64 transforms and twiddles.  Tables
// in LDS: W_512^m and the pass-3 twiddles in float64, the exchange buffers.
// The float64 integrand of a curve pair needs up to 256 registers per lane:
// one workgroup per CU is asked for so that nothing spills.
template <int ND>
__global__ __launch_bounds__(64 * kWaves, 1) void k_cal_time_w64(CalTimeArgs a) {
  extern __shared__ double lds_d[];
  const EvDev& ev = a.ev;
  double2* tw = reinterpret_cast<double2*>(lds_d);
  double2* tw3 = tw + kWP;
  double2* xch = tw3 + kTw3;
  double* knots = reinterpret_cast<double*>(xch + kWaves * kXch);
  const bool dist = (ev.flags & GWL_DISTANCE_MARG) != 0;
  for (int k = threadIdx.x; k < kWP; k += 64 * kWaves) {
    double sn, cs;
    sincospi(-2.0 * (double)k / (double)kWP, &sn, &cs);
    tw[k] = make_double2(cs, sn);
  }
  for (int k = threadIdx.x; k < kTw3; k += 64 * kWaves) {
    const int e = ((k / 64 + 1) * (k & 63)) & (kWP - 1);
    double sn, cs;
    sincospi(-2.0 * (double)e / (double)kWP, &sn, &cs);
    tw3[k] = make_double2(cs, sn);
  }
  if (dist) stage_knots(ev, knots);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long cb = blockIdx.x / a.rowblocks;
  const long long r = (blockIdx.x - cb * a.rowblocks) * kWaves + wid;
  if (r >= a.n) return;
  const int c0 = (int)cb * kCalCurves;
  const int ncb = min(kCalCurves, a.nc - c0);
  const RowW& rw = a.rows[r];
  if (!rw.valid) {
    if (lane < ncb) a.lnl[(size_t)r * a.nc + c0 + lane] = row_valid_sentinel();
    return;
  }
  double2 c64[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) c64[d] = a.c64[r * GWL_MAX_DET + d];
  const WaveEv wv = wave_ev(ev);
  const int Q = ev.Q;
  const int kb = rw.kb, ke = rw.ke;
  double2* xb = xch + wid * kXch;
  double accr[kCalCurves][8], acci[kCalCurves][8];
#pragma unroll
  for (int cc = 0; cc < kCalCurves; ++cc)
#pragma unroll
    for (int s = 0; s < 8; ++s) accr[cc][s] = acci[cc][s] = 0.0;
  for (int n2 = 0; n2 < Q; ++n2) {
    const int ilo = (kb - n2 + Q - 1) / Q;
    const int ihi = ke >= n2 ? (ke - n2) / Q : -1;
    if (ilo > ihi) continue;
    double2 xd[8][ND];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int i = lane + 64 * s;
      if (i >= ilo && i <= ihi) {
        integrand_tf2_det64<ND>(wv, rw, c64, ev.dw64, n2 * kWP + i, Q * i + n2, xd[s]);
      } else {
#pragma unroll
        for (int d = 0; d < ND; ++d) xd[s][d] = make_double2(0.0, 0.0);
      }
    }
#pragma unroll
    for (int cc = 0; cc < kCalCurves; ++cc) {
      if (cc >= ncb) break;
      const size_t off = (size_t)(c0 + cc) * ND * ev.N + (size_t)n2 * kWP;
      const float2* hi = a.cal + off;
      const float2* lo = a.cal_lo + off;
      double2 x[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        double xr = 0.0, xi = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          const size_t e = (size_t)d * ev.N + lane + 64 * s;
          const float2 kh = hi[e], kl = lo[e];
          const double kr = (double)kh.x + (double)kl.x, ki = (double)kh.y + (double)kl.y;
          xr += kr * xd[s][d].x - ki * xd[s][d].y;
          xi += kr * xd[s][d].y + ki * xd[s][d].x;
        }
        x[s] = make_double2(xr, xi);
      }
      wave_transform_accumulate64(wv, x, n2, tw, tw3, xb, lane, a.wN, accr[cc], acci[cc]);
    }
  }
  const double norm = 4.0 / ev.T;
  const double hs = rw.hscale, jit = rw.jitter;
  const double b = time_lse_scale(ev, ev.tdelta);
#pragma unroll
  for (int cc = 0; cc < kCalCurves; ++cc) {
    if (cc >= ncb) break;
    const int c = c0 + cc;
    double hh = 0.0;
    if (kb <= ke) {
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        const double* ps = a.calps + ((size_t)c * ND + d) * (ev.N + 2);
        hh += rw.cabs2[d] * (ps[ke + 1] - ps[kb]);
      }
    }
    hh *= norm;
    SplineY yy;
    if (dist) yy = spline_y(ev, knots + ev.sp.nx + 4, hh * hs * hs);
    double m = -INFINITY, sum = 0.0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int jj = (lane + 64 * s - ev.jlo) & (kWP - 1);
      if (jj < ev.L) {
        const int j = ev.jlo + jj;
        const double t = (ev.start + (double)(j + 1) * ev.tstep) + jit;
        if (t >= ev.tmin && t <= ev.tmax)
          lse_push(m, sum, epi_value(ev, knots, yy, hh, hs, norm * accr[cc][s],
                                     norm * acci[cc][s]) +
                               time_log_weight(ev, t));
      }
    }
    lse_wave(m, sum);
    if (lane == 0) a.lnl[(size_t)r * a.nc + c] = (m > -INFINITY) ? log(b * sum) + m : -INFINITY;
  }
}

// One wavefront per row, lanes over curves: logsumexp_c(lnL_c) - ln C
// (base.py:877).
__global__ __launch_bounds__(256) void k_cal_time_epi(const RowW* rows, long long n, int nc,
                                                      const double* lnl, double* out) {
  const long long r = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const int lane = threadIdx.x & 63;
  if (!rows[r].valid) {
    if (lane == 0) out[r] = row_valid_sentinel();
    return;
  }
  double m = -INFINITY, sum = 0.0;
  for (int c = lane; c < nc; c += 64) lse_push(m, sum, lnl[(size_t)r * nc + c]);
  lse_wave(m, sum);
  if (lane == 0) out[r] = (m > -INFINITY) ? log(sum) + m - log((double)nc) : -INFINITY;
}

// ------------------------------------------------------------------ multi-event
// Many TaylorF2 events in one launch: row constants of every event, then one
// workgroup per (event, group, 4-row block) item of the longest-first work
// list (the grid covers the largest list; workgroups past the current list
// exit at once), one wavefront per row; window partials go to the per-event
// buffers that the multi-event epilogue reads.  Workgroups are dispatched in
// list order, so the longest items start first, and kernels on other streams
// can interleave with this one.
struct MultiWArgs {
  const EvDev* evs;
  const double* const* params;
  const long long* rows;
  const int* groups;
  RowW* const* roww;
  double2* const* part;
  double* const* hhpart;
  const MultiItem* items;
  const int* n_items;
  int ld;
};

__global__ __launch_bounds__(256) void k_rowconst_multi(MultiWArgs a) {
  const int e = blockIdx.y;
  const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
  if (r >= a.rows[e]) return;
  rowconst_one(a.evs[e], a.params[e] + r * a.ld, a.roww[e] + r);
}

__global__ __launch_bounds__(64 * kWaves, 4) void k_time_w_multi(MultiWArgs a) {
  const int it = blockIdx.x;
  if (it >= *a.n_items) return;
  extern __shared__ double lds_d[];
  double2* wt = reinterpret_cast<double2*>(lds_d);    // [512]
  float2* tw2 = reinterpret_cast<float2*>(wt + kWP);  // [7][8]
  float2* tw3 = tw2 + kTw2;                            // [7][64]
  float2* xch = tw3 + kTw3;                            // [kWaves][kXch]
  const MultiItem m = a.items[it];
  init_row_tables(tw2, tw3, wt, a.evs[m.e].N, a.evs[m.e].jlo);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long r = m.row0 + wid;
  if (r >= a.rows[m.e]) return;
  const WaveEv ev = uniform_copy(wave_ev(a.evs[m.e]));
  const int G = __builtin_amdgcn_readfirstlane(a.groups[m.e]);
  const int Q = ev.Q;
  const int n2b = (int)(((long long)m.g * Q) / G);
  const int n2e = (int)(((long long)(m.g + 1) * Q) / G);
  const RowW& rw = a.roww[m.e][r];
  double accr[8], acci[8];
  // specialised on the detector count (wave-uniform)
  if (ev.nd == 1)
    wave_row<1, false>(ev, rw, 1, n2b, n2e, tw2, tw3, wt, xch + wid * kXch, lane, accr, acci);
  else if (ev.nd == 2)
    wave_row<2, false>(ev, rw, 2, n2b, n2e, tw2, tw3, wt, xch + wid * kXch, lane, accr, acci);
  else
    wave_row<3, false>(ev, rw, 3, n2b, n2e, tw2, tw3, wt, xch + wid * kXch, lane, accr, acci);
  wave_store_partials(ev, rw, r, G, m.g, lane, accr, acci, a.part[m.e], a.hhpart[m.e]);
}

// f^1/3, f^-1/3, ln f and dw32 f^-7/6 at every decimated element.
__global__ __launch_bounds__(256) void k_tf2_tables(EvDev ev, double* tb, float2* dwa) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ev.N) return;
  const int n = ev.Q * (i & (ev.P - 1)) + (i >> ev.logP);
  double u = 0.0, v = 0.0, lf = 0.0, amp = 0.0;
  if (n > 0) {
    double b[kNB];
    tf2_basis((double)n * ev.df, b);
    u = b[5];
    v = b[3];
    lf = b[7];
    amp = b[1] * sqrt(b[3]);
  }
  tb[3 * (size_t)i + 0] = u;
  tb[3 * (size_t)i + 1] = v;
  tb[3 * (size_t)i + 2] = lf;
  for (int d = 0; d < ev.nd; ++d) {
    const double2 w = ev.dw64[(size_t)d * ev.N + i];
    dwa[(size_t)d * ev.N + i] = make_float2((float)(w.x * amp), (float)(w.y * amp));
  }
}

bool time_wave_eligible(const gwl_event* e) {
  const EvDev& ev = e->dev;
  if (!(ev.flags & GWL_TIME_MARG) || ev.approx != GWL_TAYLORF2 || ev.P != kWP || !ev.tf2b)
    return false;
  if (e->transform == GWL_TRANSFORM_WAVE) return true;
  return e->transform == GWL_TRANSFORM_AUTO && e->tile_rows == 0;
}

int time_wave_groups(const gwl_event* e, long long n) {
  const int Q = e->dev.Q;
  const long long target = (long long)e->ctx->n_cu * (e->tile_wg_per_cu > 0 ? e->tile_wg_per_cu : 16);
  int G = 1;
  while (G < Q && G < 64 && n * G < target) G *= 2;
  return G;
}

size_t time_wave_ws_bytes(const gwl_event* e, int64_t rows) {
  const int wpc = e->tile_wg_per_cu > 0 ? e->tile_wg_per_cu : 16;
  const long long items = std::max<long long>(rows, 2LL * wpc * e->ctx->n_cu);
  return (size_t)items * (e->dev.L * sizeof(double2) + sizeof(double)) +
         (size_t)rows * sizeof(RowW) + 1024;
}

int init_time_wave_tables(const EvDev& ev, double* tb, float2* dwa) {
  hipLaunchKernelGGL(k_tf2_tables, dim3((unsigned)((ev.N + 255) / 256)), dim3(256), 0, 0, ev, tb, dwa);
  GWL_HIP_CHECK(hipGetLastError());
  GWL_HIP_CHECK(hipDeviceSynchronize());
  return GWL_OK;
}

size_t time_wave_lds_bytes(const EvDev& ev, bool fused) {
  size_t b = (size_t)kWP * sizeof(double2) + (size_t)(kTw2 + kTw3 + kWaves * kXch) * sizeof(float2);
  if (fused && (ev.flags & GWL_DISTANCE_MARG)) b += (size_t)(ev.sp.nx + ev.sp.ny + 8) * sizeof(double);
  return b;
}

// k_cal_time_w: the 512-entry pass-2 table, tw3, the exchange buffers, knots
// (float64 kernel: the same tables in float64)
static size_t cal_time_lds_bytes(const EvDev& ev, bool fp64) {
  size_t b = (size_t)(kWP + kTw3 + kWaves * kXch) * (fp64 ? sizeof(double2) : sizeof(float2));
  if (ev.flags & GWL_DISTANCE_MARG) b += (size_t)(ev.sp.nx + ev.sp.ny + 8) * sizeof(double);
  return b;
}

int launch_time_wave(gwl_event* e, const double* params, int64_t n, int ld, double* out,
                     hipStream_t st) {
  const EvDev& ev = e->dev;
  const int G = time_wave_groups(e, n);
  const size_t part_bytes = (size_t)n * G * (ev.L * sizeof(double2) + sizeof(double));
  char* ws = reinterpret_cast<char*>(e->ws);
  RowW* rows = reinterpret_cast<RowW*>(ws + (part_bytes + 255) / 256 * 256);
  if ((char*)(rows + n) > ws + e->ws_bytes) {
    set_error("workspace too small for the time transform");
    return GWL_E_INVALID;
  }
  RowArgs ra{ev, params, ld, n, rows};
  hipLaunchKernelGGL(k_rowconst, dim3((unsigned)((n + kRowconstRows - 1) / kRowconstRows)),
                     dim3(256), 0, st, ra);
  GWL_HIP_CHECK(hipGetLastError());
  TimeWArgs ta{};
  ta.ev = ev;
  ta.rows = rows;
  ta.n = n;
  ta.G = G;
  ta.fused = G == 1;
  ta.part = reinterpret_cast<double2*>(ws);
  ta.hhpart = reinterpret_cast<double*>(ws + (size_t)n * G * ev.L * sizeof(double2));
  ta.out = out;
  const size_t lds = time_wave_lds_bytes(ev, ta.fused);
  if (lds > e->ctx->lds_per_block) {
    set_error("time transform exceeds the LDS of a workgroup");
    return GWL_E_UNSUPPORTED;
  }
  const long long items = n * (long long)G;
  const dim3 grid((unsigned)((items + kWaves - 1) / kWaves));
  if (int rc = timing_begin(e, st)) return rc;
  switch (ev.nd) {
    case 1: hipLaunchKernelGGL(k_time_w<1>, grid, dim3(64 * kWaves), lds, st, ta); break;
    case 2: hipLaunchKernelGGL(k_time_w<2>, grid, dim3(64 * kWaves), lds, st, ta); break;
    case 3: hipLaunchKernelGGL(k_time_w<3>, grid, dim3(64 * kWaves), lds, st, ta); break;
    default: hipLaunchKernelGGL(k_time_w<4>, grid, dim3(64 * kWaves), lds, st, ta); break;
  }
  GWL_HIP_CHECK(hipGetLastError());
  if (int rc = timing_end(e, st)) return rc;
  if (!ta.fused) return launch_epi_time(e, params, n, ld, G, ta.part, ta.hhpart, out, st);
  return GWL_OK;
}

size_t roww_bytes(int64_t rows) { return (size_t)rows * sizeof(RowW); }

bool cal_time_eligible(const gwl_event* e) {
  const EvDev& ev = e->dev;
  return (ev.flags & GWL_TIME_MARG) && ev.approx == GWL_TAYLORF2 && ev.P == kWP && ev.tf2b &&
         ev.L <= kWP;
}

// rows: RowW, then the float64 detector amplitudes, then the per-curve lnL
static size_t cal_time_c64_offset(int64_t rows) {
  return ((size_t)rows * sizeof(RowW) + 255) / 256 * 256;
}
static size_t cal_time_lnl_offset(int64_t rows) {
  return cal_time_c64_offset(rows) + (size_t)rows * GWL_MAX_DET * sizeof(double2);
}

size_t cal_time_ws_bytes(int64_t rows, int nc) {
  return cal_time_lnl_offset(rows) + (size_t)rows * nc * sizeof(double);
}

int launch_cal_time(gwl_event* e, const double* params, int64_t n, int ld, int nc,
                    const float2* cal, const double* calps, void* ws, double* per_curve,
                    double* out, hipStream_t st, bool fp64, const float2* cal_lo,
                    const double2* wN) {
  const EvDev& ev = e->dev;
  if (fp64 && (!cal_lo || !wN)) {
    set_error("calibration time transform: float64 tables missing");
    return GWL_E_INVALID;
  }
  char* w = static_cast<char*>(ws);
  RowW* rows = reinterpret_cast<RowW*>(w);
  double2* c64 = reinterpret_cast<double2*>(w + cal_time_c64_offset(n));
  double* lnl = per_curve ? per_curve : reinterpret_cast<double*>(w + cal_time_lnl_offset(n));
  RowArgs ra{ev, params, ld, n, rows, fp64 ? c64 : nullptr};
  hipLaunchKernelGGL(k_rowconst, dim3((unsigned)((n + kRowconstRows - 1) / kRowconstRows)),
                     dim3(256), 0, st, ra);
  GWL_HIP_CHECK(hipGetLastError());
  CalTimeArgs ca{};
  ca.ev = ev;
  ca.rows = rows;
  ca.n = n;
  ca.rowblocks = (n + kWaves - 1) / kWaves;
  ca.nc = nc;
  ca.cal = cal;
  ca.calps = calps;
  ca.lnl = lnl;
  ca.cal_lo = cal_lo;
  ca.c64 = c64;
  ca.wN = wN;
  const size_t lds = cal_time_lds_bytes(ev, fp64);
  if (lds > e->ctx->lds_per_block) {
    set_error("calibration time transform exceeds the LDS of a workgroup");
    return GWL_E_UNSUPPORTED;
  }
  const long long blocks = ca.rowblocks * ((nc + kCalCurves - 1) / kCalCurves);
  if (blocks > 0x7fffffffLL) {
    set_error("gwl_calib_eval: too many rows x curves in one call");
    return GWL_E_INVALID;
  }
  const dim3 grid((unsigned)blocks);
  if (fp64) {
    switch (ev.nd) {
      case 1: hipLaunchKernelGGL(k_cal_time_w64<1>, grid, dim3(64 * kWaves), lds, st, ca); break;
      case 2: hipLaunchKernelGGL(k_cal_time_w64<2>, grid, dim3(64 * kWaves), lds, st, ca); break;
      case 3: hipLaunchKernelGGL(k_cal_time_w64<3>, grid, dim3(64 * kWaves), lds, st, ca); break;
      default: hipLaunchKernelGGL(k_cal_time_w64<4>, grid, dim3(64 * kWaves), lds, st, ca); break;
    }
  } else {
    switch (ev.nd) {
      case 1: hipLaunchKernelGGL(k_cal_time_w<1>, grid, dim3(64 * kWaves), lds, st, ca); break;
      case 2: hipLaunchKernelGGL(k_cal_time_w<2>, grid, dim3(64 * kWaves), lds, st, ca); break;
      case 3: hipLaunchKernelGGL(k_cal_time_w<3>, grid, dim3(64 * kWaves), lds, st, ca); break;
      default: hipLaunchKernelGGL(k_cal_time_w<4>, grid, dim3(64 * kWaves), lds, st, ca); break;
    }
  }
  GWL_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_cal_time_epi, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, rows,
                     (long long)n, nc, lnl, out);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

size_t multi_wave_lds_bytes() {
  return (size_t)kWP * sizeof(double2) + (size_t)(kTw2 + kTw3 + kWaves * kXch) * sizeof(float2);
}

int multi_launch_wave(gwl_multi* m, hipStream_t st) {
  MultiWArgs a;
  a.evs = m->d_evs;
  a.params = m->d_params;
  a.rows = m->d_rows;
  a.groups = m->d_groups;
  a.roww = reinterpret_cast<RowW* const*>(m->d_roww);
  a.part = reinterpret_cast<double2* const*>(m->d_parts);
  a.hhpart = m->d_hhs;
  a.items = m->d_items;
  a.n_items = m->d_n_items;
  a.ld = GWL_NPARAM;
  long long maxr = 1;
  for (int64_t r : m->max_rows) maxr = std::max<long long>(maxr, r);
  hipLaunchKernelGGL(k_rowconst_multi, dim3((unsigned)((maxr + 255) / 256), (unsigned)m->n_events),
                     dim3(256), 0, st, a);
  GWL_HIP_CHECK(hipGetLastError());
  // eager launches cover the current list; a captured graph, the largest one
  const size_t grid = m->capturing ? m->max_items : (size_t)std::max(1, m->n_items);
  hipLaunchKernelGGL(k_time_w_multi, dim3((unsigned)grid), dim3(64 * kWaves),
                     multi_wave_lds_bytes(), st, a);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

}  // namespace gwl
