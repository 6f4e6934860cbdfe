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
};

__device__ inline void rowconst_one(const EvDev& ev, const double* row, RowW* out) {
  SampleConst sc;
  compute_sample_const(ev, row, sc);
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
  w.pad = 0;
  *out = w;
}

__global__ __launch_bounds__(256) void k_rowconst(RowArgs a) {
  const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
  if (r >= a.n) return;
  rowconst_one(a.ev, a.params + r * a.ld, a.rows + r);
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

// h conj(d) / S at decimated element i (bin nbin) for one row, summed over
// detectors (interferometer.py:303-368, base.py:327-329): float64 phase,
// reduced to [0, 1) cycles, float32 rotation and products.
template <int ND>
__device__ __forceinline__ float2 integrand_tf2(const WaveEv& ev, const RowW& rw, int nd, int i,
                                                int nbin) {
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
  float ar = 0.f, ai = 0.f;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    if (d >= nd) break;
    const double cyc = fma(f, rw.dt[d], s);
    const float2 e = expm2pi((float)(cyc - floor(cyc)));
    const float2 dw = ev.dwa32[(size_t)d * ev.N + i];
    const float hr = rw.cre[d] * e.x - rw.cim[d] * e.y;
    const float hi = rw.cre[d] * e.y + rw.cim[d] * e.x;
    ar += hr * dw.x + hi * dw.y;
    ai += hi * dw.x - hr * dw.y;
  }
  return make_float2(ar, ai);
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

// Psi(f) / 2pi (cycles) of one row at a tabled basis point: the nine-term
// sum of tf2_phase_cycles in Horner form, float64.
__device__ __forceinline__ double tf2_cycles_horner(const RowW& rw, const double* tb) {
  const double u = tb[0], v = tb[1], lf = tb[2];
  // c0 v^5 + c1 v^3 + c2 v^2 + c3 v + c4
  double a = fma(rw.c[0], v * v, rw.c[1]);
  a = fma(a, v, rw.c[2]);
  a = fma(a, v, rw.c[3]);
  double s = fma(a, v, rw.c[4]);
  // + c5 u + c6 u^2 + (c7 + c8 u) ln f
  s = fma(fma(rw.c[6], u, rw.c[5]), u, s);
  return fma(fma(rw.c[8], u, rw.c[7]), lf, s);
}

// The transform of one row over sub-sequences [n2b, n2e): lane l returns
// the float64 window accumulators of bins j = l + 64 s (mod 512) in
// accr[s], acci[s].  tw: the 512-entry twiddle table in LDS; xb: this
// wavefront's exchange buffer.  nd <= ND detectors.
//
// The integrand h conj(d) / S = exp(-2 pi i Psi(f) / 2pi) sum_d C_d
// exp(-2 pi i f dt_d) conj(dw_d f^-7/6): the source phase (float64, reduced
// to [0, 1) cycles) is rotated once per element; each detector's delay
// rotation starts from an exact float64 phase at the lane's first element
// of the sub-sequence, f0 = (Q l + n2) df, and advances by the step
// w_d = exp(-2 pi i 64 Q df dt_d) (float64, rounded to float32) to the
// lane's elements l + 64 s, whose frequencies are f0 + 64 Q df s.
template <int ND>
__device__ __forceinline__ void wave_row(const WaveEv& ev, const RowW& rw, int nd, int n2b, int n2e,
                                         const float2* tw, const float2* tw3, float2* xb, int lane,
                                         double* accr, double* acci) {
  const int Q = ev.Q;
  // twiddles from LDS: pass 2 (span 8) W_64^(s k), k = l mod 8 (eight
  // distinct addresses per s); pass 3 (span 64) W_512^(s l) from tw3
  const int k2 = 8 * (lane & 7);
#pragma unroll
  for (int s = 0; s < 8; ++s) accr[s] = acci[s] = 0.0;
  const unsigned nmask = (unsigned)ev.N - 1u;
  const float invN = 1.0f / (float)ev.N;
  const int kb = rw.kb, ke = rw.ke;
  const int base2 = 64 * (lane >> 3) + (lane & 7);
  // the steps in float64 (once per row), rounded once to float32
  float2 wd[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const double st = (64.0 * (double)Q * ev.df) * rw.dt[d];
    const double2 w = expm2pi(st - floor(st));
    wd[d] = make_float2((float)w.x, (float)w.y);
  }

  for (int n2 = n2b; n2 < n2e; ++n2) {
    // in-support elements of this sub-sequence: kb <= Q i + n2 <= ke
    const int ilo = (kb - n2 + Q - 1) / Q;
    const int ihi = ke >= n2 ? (ke - n2) / Q : -1;
    float2 x[8];
    if (ilo <= ihi) {
      const double f0 = (double)(Q * lane + n2) * ev.df;
      float2 r[ND];
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        const double c0 = f0 * rw.dt[d];
        r[d] = cmul(make_float2(rw.cre[d], rw.cim[d]), expm2pi((float)(c0 - floor(c0))));
      }
      const float2* dwa = ev.dwa32 + (size_t)n2 * kWP + lane;
      const double* tb = ev.tf2b + 3 * ((size_t)n2 * kWP + lane);
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        x[s] = make_float2(0.f, 0.f);
        const int i = lane + 64 * s;
        if (i >= ilo && i <= ihi) {
          const double cyc = tf2_cycles_horner(rw, tb + 3 * 64 * s);
          float2 y = make_float2(0.f, 0.f);
#pragma unroll
          for (int d = 0; d < ND; ++d) {
            if (d >= nd) break;  // the event holds nd detectors' tables
            const float2 dw = dwa[(size_t)d * ev.N + 64 * s];
            y.x += r[d].x * dw.x + r[d].y * dw.y;
            y.y += r[d].y * dw.x - r[d].x * dw.y;
          }
          x[s] = cmul(expm2pi((float)(cyc - floor(cyc))), y);
        }
#pragma unroll
        for (int d = 0; d < ND; ++d) r[d] = cmul(r[d], wd[d]);
      }
    } else {
#pragma unroll
      for (int s = 0; s < 8; ++s) x[s] = make_float2(0.f, 0.f);
    }
    // pass 1 (span 1): inputs l + 64 s, outputs to 8 l + s.  When the
    // support ends below element 128 (uniform) only x[0], x[1] can be
    // nonzero and the butterfly is x0 + W_8^s x1 (the same arithmetic dft8
    // performs on those inputs).
    if (ilo <= ihi && ihi >= 128) {
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
    for (int s = 1; s < 8; ++s) x[s] = cmul(x[s], tw[(s * k2) & (kWP - 1)]);
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
    // window bins j = jlo + jj with j = l + 64 s (mod 512)
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
  float2* tw = reinterpret_cast<float2*>(lds_d);                  // [512]
  float2* tw3 = tw + kWP;                                          // [7][64]
  float2* xch = tw3 + kTw3;                                        // [kWaves][kXch]
  double* knots = reinterpret_cast<double*>(xch + kWaves * kXch);  // tx, ty
  const bool dist = (ev.flags & GWL_DISTANCE_MARG) != 0;
  init_wave_tables(tw, tw3);
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
  wave_row<ND>(wv, rw, ND, n2b, n2e, tw, tw3, xch + wid * kXch, lane, accr, acci);

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
};

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
  float2* tw = reinterpret_cast<float2*>(lds_d);  // [512]
  float2* tw3 = tw + kWP;                          // [7][64]
  float2* xch = tw3 + kTw3;                        // [kWaves][kXch]
  init_wave_tables(tw, tw3);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const MultiItem m = a.items[it];
  const long long r = m.row0 + wid;
  if (r >= a.rows[m.e]) return;
  const WaveEv ev = uniform_copy(wave_ev(a.evs[m.e]));
  const int G = __builtin_amdgcn_readfirstlane(a.groups[m.e]);
  const int Q = ev.Q;
  const int n2b = (int)(((long long)m.g * Q) / G);
  const int n2e = (int)(((long long)(m.g + 1) * Q) / G);
  const RowW rw = uniform_copy(a.roww[m.e][r]);
  double accr[8], acci[8];
  wave_row<3>(ev, rw, ev.nd, n2b, n2e, tw, tw3, xch + wid * kXch, lane, accr, acci);
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
  size_t b = (size_t)(kWP + kTw3 + kWaves * kXch) * sizeof(float2);
  if (fused && (ev.flags & GWL_DISTANCE_MARG)) b += (size_t)(ev.sp.nx + ev.sp.ny + 8) * sizeof(double);
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
  hipLaunchKernelGGL(k_rowconst, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ra);
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

size_t cal_time_ws_bytes(int64_t rows, int nc) {
  return ((size_t)rows * sizeof(RowW) + 255) / 256 * 256 + (size_t)rows * nc * sizeof(double);
}

int launch_cal_time(gwl_event* e, const double* params, int64_t n, int ld, int nc,
                    const float2* cal, const double* calps, void* ws, double* per_curve,
                    double* out, hipStream_t st) {
  const EvDev& ev = e->dev;
  char* w = static_cast<char*>(ws);
  RowW* rows = reinterpret_cast<RowW*>(w);
  double* lnl = per_curve ? per_curve
                          : reinterpret_cast<double*>(w + ((size_t)n * sizeof(RowW) + 255) / 256 * 256);
  RowArgs ra{ev, params, ld, n, rows};
  hipLaunchKernelGGL(k_rowconst, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ra);
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
  const size_t lds = time_wave_lds_bytes(ev, true);
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
  switch (ev.nd) {
    case 1: hipLaunchKernelGGL(k_cal_time_w<1>, grid, dim3(64 * kWaves), lds, st, ca); break;
    case 2: hipLaunchKernelGGL(k_cal_time_w<2>, grid, dim3(64 * kWaves), lds, st, ca); break;
    case 3: hipLaunchKernelGGL(k_cal_time_w<3>, grid, dim3(64 * kWaves), lds, st, ca); break;
    default: hipLaunchKernelGGL(k_cal_time_w<4>, grid, dim3(64 * kWaves), lds, st, ca); break;
  }
  GWL_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_cal_time_epi, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, rows,
                     (long long)n, nc, lnl, out);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

size_t multi_wave_lds_bytes() { return (size_t)(kWP + kTw3 + kWaves * kXch) * sizeof(float2); }

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
