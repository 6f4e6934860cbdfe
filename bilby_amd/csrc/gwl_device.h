// Device-side building blocks of the compact-binary likelihood kernels:
// complex arithmetic, the per-row FP64 prologue (sidereal time, antenna
// patterns, time delays, source constants), per-bin waveform evaluation,
// ln I0, the bicubic B-spline and the LDS radix FFT.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <cfloat>
#include "../../include/gwl.h"

namespace gwl {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kTwoPi = 6.283185307179586476925286766559005768;
// bilby/core/utils/constants.py:3-7
constexpr double kSpeedOfLight = 299792458.0;
constexpr double kParsec = 3.085677581491367e16;
constexpr double kSolarMass = 1.988409870698050731911960804878414216e30;
constexpr double kG = 6.6743e-11;
constexpr double kEulerGamma = 0.5772156649015329;

// ------------------------------------------------------------ complex
template <typename R> struct cplx_t;
template <> struct cplx_t<float> { using type = float2; };
template <> struct cplx_t<double> { using type = double2; };
template <typename R> using cplx = typename cplx_t<R>::type;

template <typename C> __device__ __forceinline__ C cmk(decltype(C::x) a, decltype(C::x) b) {
  C r; r.x = a; r.y = b; return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * (-i)
template <typename C> __device__ __forceinline__ C cmul_mi(C a) { return cmk<C>(a.y, -a.x); }

// exp(-2 pi i r) for r in [0, 1): (cos 2 pi r, -sin 2 pi r)
__device__ __forceinline__ float2 expm2pi(float r) {
  // v_sin_f32 / v_cos_f32 take their argument in revolutions.
  return make_float2(__builtin_amdgcn_cosf(r), -__builtin_amdgcn_sinf(r));
}
// float64: the nearest of the 64 points j / 64 of the circle from a table
// (cos, sin of 2 pi j / 64, correctly rounded), times the Taylor series of
// exp(-i theta) on |theta| <= pi / 64, whose truncation (theta^9 / 9! for
// the sine, theta^10 / 10! for the cosine) stays below 5e-18: the table
// method of P. T. P. Tang (ACM TOMS 15, 1989, 144-157) applied to the
// circle.  r - j / 64 is exact, so the error is that of the few final
// roundings (about 2 ulp).  Valid for |r| < 2^25.
__constant__ __attribute__((aligned(16))) static const double kCis64[64][2] = {
    {0x1.0000000000000p+0, 0x0.0p+0}, {0x1.fd88da3d12526p-1, 0x1.917a6bc29b42cp-4},
    {0x1.f6297cff75cb0p-1, 0x1.8f8b83c69a60bp-3}, {0x1.e9f4156c62ddap-1, 0x1.294062ed59f06p-2},
    {0x1.d906bcf328d46p-1, 0x1.87de2a6aea963p-2}, {0x1.c38b2f180bdb1p-1, 0x1.e2b5d3806f63bp-2},
    {0x1.a9b66290ea1a3p-1, 0x1.1c73b39ae68c8p-1}, {0x1.8bc806b151741p-1, 0x1.44cf325091dd6p-1},
    {0x1.6a09e667f3bcdp-1, 0x1.6a09e667f3bcdp-1}, {0x1.44cf325091dd6p-1, 0x1.8bc806b151741p-1},
    {0x1.1c73b39ae68c8p-1, 0x1.a9b66290ea1a3p-1}, {0x1.e2b5d3806f63bp-2, 0x1.c38b2f180bdb1p-1},
    {0x1.87de2a6aea963p-2, 0x1.d906bcf328d46p-1}, {0x1.294062ed59f06p-2, 0x1.e9f4156c62ddap-1},
    {0x1.8f8b83c69a60bp-3, 0x1.f6297cff75cb0p-1}, {0x1.917a6bc29b42cp-4, 0x1.fd88da3d12526p-1},
    {0x0.0p+0, 0x1.0000000000000p+0}, {-0x1.917a6bc29b42cp-4, 0x1.fd88da3d12526p-1},
    {-0x1.8f8b83c69a60bp-3, 0x1.f6297cff75cb0p-1}, {-0x1.294062ed59f06p-2, 0x1.e9f4156c62ddap-1},
    {-0x1.87de2a6aea963p-2, 0x1.d906bcf328d46p-1}, {-0x1.e2b5d3806f63bp-2, 0x1.c38b2f180bdb1p-1},
    {-0x1.1c73b39ae68c8p-1, 0x1.a9b66290ea1a3p-1}, {-0x1.44cf325091dd6p-1, 0x1.8bc806b151741p-1},
    {-0x1.6a09e667f3bcdp-1, 0x1.6a09e667f3bcdp-1}, {-0x1.8bc806b151741p-1, 0x1.44cf325091dd6p-1},
    {-0x1.a9b66290ea1a3p-1, 0x1.1c73b39ae68c8p-1}, {-0x1.c38b2f180bdb1p-1, 0x1.e2b5d3806f63bp-2},
    {-0x1.d906bcf328d46p-1, 0x1.87de2a6aea963p-2}, {-0x1.e9f4156c62ddap-1, 0x1.294062ed59f06p-2},
    {-0x1.f6297cff75cb0p-1, 0x1.8f8b83c69a60bp-3}, {-0x1.fd88da3d12526p-1, 0x1.917a6bc29b42cp-4},
    {-0x1.0000000000000p+0, 0x0.0p+0}, {-0x1.fd88da3d12526p-1, -0x1.917a6bc29b42cp-4},
    {-0x1.f6297cff75cb0p-1, -0x1.8f8b83c69a60bp-3}, {-0x1.e9f4156c62ddap-1, -0x1.294062ed59f06p-2},
    {-0x1.d906bcf328d46p-1, -0x1.87de2a6aea963p-2}, {-0x1.c38b2f180bdb1p-1, -0x1.e2b5d3806f63bp-2},
    {-0x1.a9b66290ea1a3p-1, -0x1.1c73b39ae68c8p-1}, {-0x1.8bc806b151741p-1, -0x1.44cf325091dd6p-1},
    {-0x1.6a09e667f3bcdp-1, -0x1.6a09e667f3bcdp-1}, {-0x1.44cf325091dd6p-1, -0x1.8bc806b151741p-1},
    {-0x1.1c73b39ae68c8p-1, -0x1.a9b66290ea1a3p-1}, {-0x1.e2b5d3806f63bp-2, -0x1.c38b2f180bdb1p-1},
    {-0x1.87de2a6aea963p-2, -0x1.d906bcf328d46p-1}, {-0x1.294062ed59f06p-2, -0x1.e9f4156c62ddap-1},
    {-0x1.8f8b83c69a60bp-3, -0x1.f6297cff75cb0p-1}, {-0x1.917a6bc29b42cp-4, -0x1.fd88da3d12526p-1},
    {0x0.0p+0, -0x1.0000000000000p+0}, {0x1.917a6bc29b42cp-4, -0x1.fd88da3d12526p-1},
    {0x1.8f8b83c69a60bp-3, -0x1.f6297cff75cb0p-1}, {0x1.294062ed59f06p-2, -0x1.e9f4156c62ddap-1},
    {0x1.87de2a6aea963p-2, -0x1.d906bcf328d46p-1}, {0x1.e2b5d3806f63bp-2, -0x1.c38b2f180bdb1p-1},
    {0x1.1c73b39ae68c8p-1, -0x1.a9b66290ea1a3p-1}, {0x1.44cf325091dd6p-1, -0x1.8bc806b151741p-1},
    {0x1.6a09e667f3bcdp-1, -0x1.6a09e667f3bcdp-1}, {0x1.8bc806b151741p-1, -0x1.44cf325091dd6p-1},
    {0x1.a9b66290ea1a3p-1, -0x1.1c73b39ae68c8p-1}, {0x1.c38b2f180bdb1p-1, -0x1.e2b5d3806f63bp-2},
    {0x1.d906bcf328d46p-1, -0x1.87de2a6aea963p-2}, {0x1.e9f4156c62ddap-1, -0x1.294062ed59f06p-2},
    {0x1.f6297cff75cb0p-1, -0x1.8f8b83c69a60bp-3}, {0x1.fd88da3d12526p-1, -0x1.917a6bc29b42cp-4}
};
__device__ __forceinline__ double2 expm2pi(double r) {
  const double j = rint(r * 64.0);
  const double t = fma(j, -0.015625, r);
  const double th = t * kTwoPi;
  const double t2 = th * th;
  const double sn = fma(th * t2, fma(t2, fma(t2, -1.0 / 5040.0, 1.0 / 120.0), -1.0 / 6.0), th);
  const double cs =
      fma(t2, fma(t2, fma(t2, fma(t2, 1.0 / 40320.0, -1.0 / 720.0), 1.0 / 24.0), -0.5), 1.0);
  const double2 w = *reinterpret_cast<const double2*>(kCis64[(int)j & 63]);
  // (w.x - i w.y)(cs - i sn)
  return make_double2(fma(w.x, cs, -w.y * sn), -fma(w.y, cs, w.x * sn));
}

// ------------------------------------------------------------ event (device view)
struct SplineDev {
  const double* tx;
  const double* ty;
  const double* c;
  int nx, ny;  // coefficient counts (knots: nx + 4, ny + 4)
  double xmin, xmax, ymin, ymax;
};

struct EvDev {
  int nd, nf, N, P, Q, logP;
  double T, fs, df, start;
  int mlo[GWL_MAX_DET], mhi[GWL_MAX_DET];
  double tensor[GWL_MAX_DET][9];
  double vertex[GWL_MAX_DET][3];
  int approx;
  double wf_fmin, wf_fmax, fref;
  unsigned flags;
  double tmin, tmax, dref, sigma;
  // time prior other than Uniform: its density on the grid tp_t0 + k / tp_ih,
  // k < tp_n (null: Uniform on [tmin, tmax])
  const double* tpd;
  double tp_t0, tp_ih;
  int tp_n;
  // data in decimated order: element i <-> bin n = Q*(i % P) + i / P, i < N.
  const float2* dw32;  // [nd][N]  (d/S) * sigma
  const float* w32;    // [nd][N]  sigma^2 / S
  const double2* dw64;
  const double* w64;
  // TaylorF2 tables for the wave-per-row time transform (null otherwise):
  const double* tf2b;   // [N][3] (f^1/3, f^-1/3, ln f), decimated order
  const float2* dwa32;  // [nd][N] dw32 f^-7/6, decimated order
  const double* hps;    // [nd][N + 2] prefix sums of f^-7/3 sigma^2 / S over bins
                        //   0..N in natural order, hps[d][n] = sum of bins < n
  // Nyquist bin (n = N), kept apart from the power-of-two FFT range
  double dwN[GWL_MAX_DET][2];
  double wN[GWL_MAX_DET];
  // time marginalisation window: j in [jlo, jlo + L)
  int jlo, L;
  double tstep;  // spacing of the time grid, T / N = 2 / fs
  double tdelta; // bilby's _delta_tc
  SplineDev sp;
  // GWL_HOST_POLARIZATIONS: the caller's [rows][2][nf] complex polarizations
  const double2* pol;
};

// ------------------------------------------------------------ sidereal time, geometry
// bilby/gw/time.py:74-93
__constant__ static const double kLeapSeconds[18] = {
    46828800, 78364801, 109900802, 173059203, 252028804, 315187205, 346723206,
    393984007, 425520008, 457056009, 504489610, 551750411, 599184012,
    820108813, 914803214, 1025136015, 1119744016, 1167264017};

// bilby/gw/time.py:114-192 (GPS -> UTC -> Julian day -> GMST in radians)
// The sidereal seconds reach ~7e8, whose last place is ~1e-11 rad: every
// operation is rounded on its own (no fused multiply-adds), as the reference
// computes it, so that the sidereal time, and with it each detector's delay,
// does not move with the compiler's contraction choices.
__device__ inline double gmst_rad(double gps) {
#pragma clang fp contract(off)
  double gps_int = floor(gps);
  int nleap = 0;
#pragma unroll
  for (int k = 0; k < 18; ++k) nleap += (gps_int > kLeapSeconds[k]) ? 1 : 0;
  double second = gps_int - (double)nleap;
  // datetime(1980, 1, 6) julian_day: 367 y - 7 (y + (m + 9) // 12) // 4 + 275 m // 9 + d
  double jd = (723231.0 + second / 86400.0) + 1721013.5;
  double t_hi = (jd - 2451545.0) / 36525.0;
  double t_lo = fmod(gps, 1.0) / (36525.0 * 86400.0);
  double t = t_hi + t_lo;
  double st = (0.0 + (-6.2e-6 * t + 0.093104) * (t * t)) + 67310.54841;
  st += 8640184.812866 * t_lo;
  st += 3155760000.0 * t_lo;
  st += 8640184.812866 * t_hi;
  st += 3155760000.0 * t_hi;
  return st * 2.0 * kPi / 86400.0;
}

// The sky's frame at one time: the polarisation tensors' vectors m, n
// (geometry.py:118-186, 261-279) and the propagation direction omega
// (geometry.py:282-343), shared by every detector of a row.
struct SkyBasis {
  double m[3], n[3], om[3];
};

__device__ inline SkyBasis sky_basis(double ra, double dec, double gps, double psi) {
  double g = gmst_rad(gps);
  // g mod 2 pi, exact as fmod is: the remainder is representable, and the
  // fused form rounds it once
  g = fma(-kTwoPi, floor(g / kTwoPi), g);
  double phi = ra - g;
  double theta = 0.5 * kPi - dec;
  double sphi, cphi, sth, cth, sps, cps;
  sincos(phi, &sphi, &cphi);
  sincos(theta, &sth, &cth);
  sincos(psi, &sps, &cps);
  double u[3] = {cphi * cth, cth * sphi, -sth};
  double v[3] = {-sphi, cphi, 0.0};
  SkyBasis b;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    b.m[i] = -u[i] * sps - v[i] * cps;
    b.n[i] = -u[i] * cps + v[i] * sps;
  }
  b.om[0] = sth * cphi;
  b.om[1] = sth * sphi;
  b.om[2] = cth;
  return b;
}

// Antenna patterns and geocentre time delay of one detector in a sky frame.
__device__ inline void detector_projection(const double* tensor, const double* vertex,
                                           const SkyBasis& b, double& fp, double& fc,
                                           double& delay) {
  double ap = 0.0, ac = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      ap += tensor[3 * i + j] * (b.m[i] * b.m[j] - b.n[i] * b.n[j]);
      ac += tensor[3 * i + j] * (b.m[i] * b.n[j] + b.n[i] * b.m[j]);
    }
  fp = ap;
  fc = ac;
  delay = (b.om[0] * (0.0 - vertex[0]) + b.om[1] * (0.0 - vertex[1]) +
           b.om[2] * (0.0 - vertex[2])) / kSpeedOfLight;
}

// Antenna patterns (bilby/gw/geometry.py:118-186, 261-279) and geocentre
// time delay (geometry.py:282-343) for one detector.
__device__ inline void antenna_and_delay(const double* tensor, const double* vertex, double ra,
                                         double dec, double gps, double psi, double& fp,
                                         double& fc, double& delay) {
  detector_projection(tensor, vertex, sky_basis(ra, dec, gps, psi), fp, fc, delay);
}

// ------------------------------------------------------------ per-row constants
// Phase basis of TaylorF2 (per bin, float64):
//   b = {f^-5/3, f^-1, f^-2/3, f^-1/3, 1, f^1/3, f^2/3, ln f, f^1/3 ln f}
constexpr int kNB = 9;

struct SampleConst {
  double c[10];                 // TF2: basis coefficients (cycles); SG: see below
  double dt[GWL_MAX_DET];       // total time shift of each detector [s]
  double cre[GWL_MAX_DET];      // complex detector amplitude (TF2) / F+ A+ (SG)
  double cim[GWL_MAX_DET];      //                                 / -Fx Ax (SG)
  double cabs2[GWL_MAX_DET];
  double hscale;                // d_L / d_ref (distance marginalisation) else 1
  double jitter;
  int kb, ke;                   // waveform support, bins [kb, ke]
  int valid;
  int row;                      // the row's index into caller data (host polarizations)
};

// TaylorF2 3.5PN phase coefficients (non-spinning: Buonanno et al. 2009,
// arXiv:0907.0700 Eq. 3.18; aligned spin: Khan et al. 2016,
// arXiv:1508.07253 App. B).  phil5, phil6 multiply ln v.
__device__ inline void tf2_phis(double m1, double m2, double chi1, double chi2, double* phi,
                                double& phil5, double& phil6, double& eta) {
  double M = m1 + m2;
  eta = m1 * m2 / (M * M);
  double delta = (m1 - m2) / M;
  double chis = 0.5 * (chi1 + chi2), chia = 0.5 * (chi1 - chi2);
  double pi = kPi, e2 = eta * eta, e3 = e2 * eta;
  phi[0] = 1.0;
  phi[1] = 0.0;
  phi[2] = 3715.0 / 756.0 + 55.0 * eta / 9.0;
  phi[3] = -16.0 * pi + 113.0 * delta * chia / 3.0 + (113.0 / 3.0 - 76.0 * eta / 3.0) * chis;
  phi[4] = 15293365.0 / 508032.0 + 27145.0 * eta / 504.0 + 3085.0 * e2 / 72.0 +
           (-405.0 / 8.0 + 200.0 * eta) * chia * chia - 405.0 / 4.0 * delta * chia * chis +
           (-405.0 / 8.0 + 5.0 * eta / 2.0) * chis * chis;
  double p5 = 38645.0 * pi / 756.0 - 65.0 * pi * eta / 9.0 +
              delta * (-732985.0 / 2268.0 - 140.0 * eta / 9.0) * chia +
              (-732985.0 / 2268.0 + 24260.0 * eta / 81.0 + 340.0 * e2 / 9.0) * chis;
  phi[5] = p5;
  phil5 = 3.0 * p5;
  phi[6] = 11583231236531.0 / 4694215680.0 - 6848.0 * kEulerGamma / 21.0 - 640.0 * pi * pi / 3.0 +
           (-15737765635.0 / 3048192.0 + 2255.0 * pi * pi / 12.0) * eta + 76055.0 * e2 / 1728.0 -
           127825.0 * e3 / 1296.0 - 6848.0 / 21.0 * log(4.0) + 2270.0 * pi * delta * chia / 3.0 +
           (2270.0 * pi / 3.0 - 520.0 * pi * eta) * chis;
  phil6 = -6848.0 / 21.0;
  phi[7] = 77096675.0 * pi / 254016.0 + 378515.0 * pi * eta / 1512.0 - 74045.0 * pi * e2 / 756.0 +
           delta * (-25150083775.0 / 3048192.0 + 26804935.0 * eta / 6048.0 - 1985.0 * e2 / 48.0) * chia +
           (-25150083775.0 / 3048192.0 + 10566655595.0 * eta / 762048.0 - 1042165.0 * e2 / 3024.0 +
            5345.0 * e3 / 36.0) * chis;
}

__device__ inline void tf2_basis(double f, double* b) {
  double f13 = cbrt(f);
  double u = 1.0 / f13;
  double u2 = u * u, u3 = u2 * u;
  double lnf = log(f);
  b[0] = u3 * u2;
  b[1] = u3;
  b[2] = u2;
  b[3] = u;
  b[4] = 1.0;
  b[5] = f13;
  b[6] = f13 * f13;
  b[7] = lnf;
  b[8] = f13 * lnf;
}

__device__ inline double tf2_phase_cycles(const double* c, const double* b) {
  double s = c[4];
  s = fma(c[0], b[0], s);
  s = fma(c[1], b[1], s);
  s = fma(c[2], b[2], s);
  s = fma(c[3], b[3], s);
  s = fma(c[5], b[5], s);
  s = fma(c[6], b[6], s);
  s = fma(c[7], b[7], s);
  s = fma(c[8], b[8], s);
  return s;
}

// First bin k in [lo, hi] with k*df >= f (returns hi + 1 if none).
__device__ inline int first_bin_ge(double f, double df, int lo, int hi) {
  double q = ceil(f / df);
  long long k = (long long)fmax(q, (double)lo - 1.0);
  if (k < lo) k = lo;
  while (k > lo && (double)(k - 1) * df >= f) --k;
  while (k <= hi && (double)k * df < f) ++k;
  return (int)k;
}
// Last bin k in [lo, hi] with k*df <= f (returns lo - 1 if none).
__device__ inline int last_bin_le(double f, double df, int lo, int hi) {
  double q = floor(f / df);
  long long k = (long long)fmin(q, (double)hi + 1.0);
  if (k > hi) k = hi;
  while (k < hi && (double)(k + 1) * df <= f) ++k;
  while (k >= lo && (double)k * df > f) --k;
  return (int)k;
}

// TaylorF2 per-row constants: phase-basis coefficients c[kNB] in cycles
// (Psi(f) / 2pi = sum_k c_k b_k(f), reference-frequency phase and t_c = 0
// folded in), the amplitude prefactor amp0 (divided by ``sigma``) of
// A(f) = amp0 f^-7/6, and the ISCO frequency.  amp0 is negative, as
// lalsimulation's TaylorF2 writes it (amp0 = -4 m1 m2 / r * MRSUN * MTSUN *
// sqrt(pi / 12) times the Newtonian sqrt(-dE/dv / F) v; bilby takes h+, hx
// from there, source.py:633-643), so h+ = amp0 f^-7/6 (1 + cos^2 i)/2
// exp(-i Psi).  Returns 0 for rows whose waveform is undefined
// (non-positive masses or distance, |chi| >= 1, non-finite values).
__device__ inline int tf2_row_constants(const double* row, double fref, double sigma, double* cout,
                                        double& amp0, double& fisco) {
  double m1 = row[GWL_P_MASS_1], m2 = row[GWL_P_MASS_2];
  double chi1 = row[GWL_P_CHI_1], chi2 = row[GWL_P_CHI_2];
  double dl = row[GWL_P_LUMINOSITY_DISTANCE];
  double iota = row[GWL_P_THETA_JN], phase = row[GWL_P_PHASE];
  int valid = 1;
  if (!(m1 > 0.0) || !(m2 > 0.0) || !(dl > 0.0) || !(fabs(chi1) < 1.0) || !(fabs(chi2) < 1.0) ||
      !isfinite(m1 + m2 + dl + iota + phase)) {
    valid = 0;
  }
  double M = m1 + m2;
  double mtsun = kG * kSolarMass / (kSpeedOfLight * kSpeedOfLight * kSpeedOfLight);
  double msec = M * mtsun;
  double phi[8], phil5, phil6, eta;
  tf2_phis(m1, m2, chi1, chi2, phi, phil5, phil6, eta);
  double piM = kPi * msec;
  double p13 = cbrt(piM), ip13 = 1.0 / p13;
  double lnpiM = log(piM);
  double pre = 3.0 / (128.0 * eta);
  double c[kNB];
  c[0] = pre * phi[0] * ip13 * ip13 * ip13 * ip13 * ip13;
  c[1] = pre * phi[2] * ip13 * ip13 * ip13;
  c[2] = pre * phi[3] * ip13 * ip13;
  c[3] = pre * phi[4] * ip13;
  c[4] = pre * (phi[5] + phil5 * lnpiM / 3.0);
  c[5] = pre * (phi[6] + phil6 * lnpiM / 3.0) * p13;
  c[6] = pre * phi[7] * p13 * p13;
  c[7] = pre * phil5 / 3.0;
  c[8] = pre * phil6 / 3.0 * p13;
  double bref[kNB];
  tf2_basis(fref, bref);
  double psi_ref = tf2_phase_cycles(c, bref);  // radians here (c not yet scaled)
  c[4] += -psi_ref - 2.0 * phase - 0.25 * kPi;
  for (int k = 0; k < kNB; ++k) cout[k] = c[k] / kTwoPi;
  cout[9] = 0.0;
  double mc = pow(eta, 0.6) * msec;
  amp0 = -sqrt(5.0 / 24.0) * pow(kPi, -2.0 / 3.0) * kSpeedOfLight / (dl * 1e6 * kParsec) *
         pow(mc, 5.0 / 6.0) / sigma;
  fisco = 1.0 / (pow(6.0, 1.5) * kPi * msec);
  return valid;
}

// The TaylorF2 source part of a row's constants (tf2_row_constants and
// cos iota), computable apart from the sky part.
struct Tf2Part {
  double c[10];
  double amp0, fisco, ci;
  int valid;
};

__device__ inline void tf2_part(const EvDev& ev, const double* row, Tf2Part& t) {
  t.valid = tf2_row_constants(row, ev.fref, ev.sigma, t.c, t.amp0, t.fisco);
  t.ci = cos(row[GWL_P_THETA_JN]);
}

// Per-row prologue: everything that depends on the parameter row only.
// pre: the TaylorF2 part computed elsewhere (another lane), or null.
__device__ inline void compute_sample_const(const EvDev& ev, const double* row, SampleConst& sc,
                                            const Tf2Part* pre = nullptr) {
  double ra = row[GWL_P_RA], dec = row[GWL_P_DEC], psi = row[GWL_P_PSI];
  double tc = row[GWL_P_GEOCENT_TIME], jit = row[GWL_P_TIME_JITTER];
  double dl = row[GWL_P_LUMINOSITY_DISTANCE];
  sc.row = 0;
  bool timem = (ev.flags & GWL_TIME_MARG) != 0;
  bool jitter = timem && (ev.flags & GWL_JITTER_TIME);
  // base.py:427-428: geocent_time += time_jitter before the detector loop
  double gt = jitter ? (tc + jit) : tc;
  sc.jitter = jitter ? jit : 0.0;
  sc.hscale = (ev.flags & GWL_DISTANCE_MARG) ? dl / ev.dref : 1.0;
  int lo = ev.mlo[0], hi = ev.mhi[0];
  for (int d = 1; d < ev.nd; ++d) {
    lo = min(lo, ev.mlo[d]);
    hi = max(hi, ev.mhi[d]);
  }
  double fp[GWL_MAX_DET], fc[GWL_MAX_DET];
  const SkyBasis sky = sky_basis(ra, dec, gt, psi);  // once per row, not per detector
  for (int d = 0; d < ev.nd; ++d) {
    double delay;
    detector_projection(ev.tensor[d], ev.vertex[d], sky, fp[d], fc[d], delay);
    double dtg = gt - ev.start;  // interferometer.py:355-358: subtract GPS times first
    sc.dt[d] = dtg + delay;
  }
  sc.valid = 1;
  if (ev.approx == GWL_TAYLORF2) {
    Tf2Part own;
    if (!pre) tf2_part(ev, row, own);
    const Tf2Part& t = pre ? *pre : own;
#pragma unroll
    for (int k = 0; k < 10; ++k) sc.c[k] = t.c[k];
    sc.valid = t.valid;
    const double amp0 = t.amp0, fisco = t.fisco, ci = t.ci;
    double ap = amp0 * (1.0 + ci * ci) / 2.0, ax = amp0 * ci;
    for (int d = 0; d < ev.nd; ++d) {
      sc.cre[d] = fp[d] * ap;
      sc.cim[d] = -fc[d] * ax;
      sc.cabs2[d] = sc.cre[d] * sc.cre[d] + sc.cim[d] * sc.cim[d];
    }
    // lalsimulation ends TaylorF2 at the fEnd bilby passes (maximum_frequency);
    // at the ISCO as well only when the event asks for it
    const double fmax = (ev.flags & GWL_TF2_END_ISCO) ? fmin(fisco, ev.wf_fmax) : ev.wf_fmax;
    int kb = ev.wf_fmin > 0.0 ? first_bin_ge(ev.wf_fmin, ev.df, lo, hi) : lo;
    int ke = last_bin_le(fmax, ev.df, lo, hi);
    sc.kb = max(kb, lo);
    sc.ke = min(ke, hi);
  } else if (ev.approx == GWL_HOST_POLARIZATIONS) {
    // the caller's polarizations: F+ and Fx (over sigma) per detector
    for (int k = 0; k < 10; ++k) sc.c[k] = 0.0;
    for (int d = 0; d < ev.nd; ++d) {
      sc.cre[d] = fp[d] / ev.sigma;
      sc.cim[d] = fc[d] / ev.sigma;
      sc.cabs2[d] = 0.0;
    }
    sc.kb = lo;
    sc.ke = hi;
  } else {  // sine-Gaussian, source.py:1153-1209
    double hrss = row[GWL_P_HRSS], Q = row[GWL_P_Q], f0 = row[GWL_P_FREQUENCY];
    if (!(Q > 0.0) || !(f0 > 0.0) || !isfinite(hrss)) sc.valid = 0;
    double tau = Q / (sqrt(2.0) * kPi * f0);
    double temp = Q / (4.0 * sqrt(kPi) * f0);
    double base = hrss * sqrt(kPi) * tau / 2.0;
    double ap = base / sqrt(temp * (1.0 + exp(-Q * Q))) / ev.sigma;
    double ax = base / sqrt(temp * (1.0 - exp(-Q * Q))) / ev.sigma;
    // exponent: exp(-pi^2 tau^2 (f -+ f0)^2) = exp2(-c0 (f -+ f0)^2)
    sc.c[0] = kPi * kPi * tau * tau * 1.4426950408889634074;
    sc.c[1] = f0;
    sc.c[2] = kPi * kPi * tau * tau;
    for (int k = 3; k < 10; ++k) sc.c[k] = 0.0;
    for (int d = 0; d < ev.nd; ++d) {
      sc.cre[d] = fp[d] * ap;
      sc.cim[d] = -fc[d] * ax;
      sc.cabs2[d] = 0.0;
    }
    sc.kb = lo;
    sc.ke = hi;
  }
  if (!sc.valid) {
    sc.kb = 1;
    sc.ke = 0;
  }
}

// ------------------------------------------------------------ special functions
// ln I0(x) (bilby/gw/utils.py:1006-1022 computes log(i0e(x)) + |x|).
// Power series I0(x) = sum_k c_k q^k, q = x^2/4, c_k = 1/(k!)^2, for
// x < 30, in Horner form over a compile-time table of c_k (all terms
// positive: no cancellation, relative error ~ K ulp) truncated at K terms
// chosen per range so that c_K q^K < 1e-17 I0(x): K = 24 for x <= 8,
// 40 for x <= 16, 60 below 30.  For x >= 30 the large-argument expansion
// I0(x) = e^x / sqrt(2 pi x) sum_k ((2k-1)!!)^2 / (k! (8x)^k)
// (Abramowitz & Stegun 9.7.1), summed until float64 round-off (its terms
// fall below 1e-17 long before they start to grow at k ~ 2x).
struct I0Series {
  double c[60];
  constexpr I0Series() : c() {
    double f = 1.0;
    c[0] = 1.0;
    for (int k = 1; k < 60; ++k) {
      f *= (double)k;
      c[k] = 1.0 / (f * f);
    }
  }
};
constexpr I0Series kI0Series{};

// The large-argument coefficients a_k = ((2k-1)!!)^2 / k!, k < 19: at
// x >= 30 (y = 1/(8x) <= 1/240) the term a_18 y^18 is below 1e-17 and the
// terms are still falling, so 19 terms give float64 round-off.
struct I0Asymptotic {
  double a[19];
  constexpr I0Asymptotic() : a() {
    a[0] = 1.0;
    for (int k = 1; k < 19; ++k) a[k] = a[k - 1] * (double)((2 * k - 1) * (2 * k - 1)) / (double)k;
  }
};
constexpr I0Asymptotic kI0Asymptotic{};

// sum_{k < K} c_k q^k in Horner form, unrolled (the coefficients become
// wave-uniform constants)
template <int K>
__host__ __device__ __forceinline__ double i0_series(double q) {
  double s = kI0Series.c[K - 1];
#pragma unroll
  for (int k = K - 2; k >= 0; --k) s = fma(s, q, kI0Series.c[k]);
  return s;
}

// Not inlined: inlined into a loop, its coefficients would be hoisted into
// vector registers of the caller.  The series length follows the range (the
// K of the comment above); lanes of one wavefront hold neighbouring window
// bins, so they mostly take the same branch.
__host__ __device__ inline __attribute__((noinline)) double ln_i0(double x) {
  x = fabs(x);
  if (x < 30.0) {
    const double q = 0.25 * x * x;
    if (x <= 8.0) return log(i0_series<24>(q));
    if (x <= 16.0) return log(i0_series<40>(q));
    return log(i0_series<60>(q));
  }
  const double y = 1.0 / (8.0 * x);
  double s = kI0Asymptotic.a[18];
#pragma unroll
  for (int k = 17; k >= 0; --k) s = fma(s, y, kI0Asymptotic.a[k]);
  return x - 0.5 * log(kTwoPi * x) + log(s);
}

// Largest l in [3, ncoef - 1] with t[l] <= x (FITPACK fpbisp interval rule).
__host__ __device__ inline int knot_interval(const double* t, int ncoef, double x) {
  int lo = 3, hi = ncoef - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (t[mid] <= x) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Cubic B-spline values at x in interval l (Cox-de Boor, as FITPACK fpbspl).
__host__ __device__ inline void bspline3(const double* t, int l, double x, double* h) {
  double hh[3];
  h[0] = 1.0;
#pragma unroll
  for (int j = 1; j <= 3; ++j) {
#pragma unroll
    for (int i = 0; i < j; ++i) hh[i] = h[i];
    h[0] = 0.0;
#pragma unroll
    for (int i = 1; i <= j; ++i) {
      int li = l + i, lj = li - j;
      double f = hh[i - 1] / (t[li] - t[lj]);
      h[i - 1] += f * (t[li] - x);
      h[i] = f * (x - t[lj]);
    }
  }
}

// BoundedRectBivariateSpline (bilby/core/utils/calculus.py:221-262):
// the bicubic interpolant, -inf outside the bounding box.
__host__ __device__ inline double spline2d(const SplineDev& sp, double x, double y) {
  if (!(x >= sp.xmin && x <= sp.xmax && y >= sp.ymin && y <= sp.ymax)) return -INFINITY;
  int lx = knot_interval(sp.tx, sp.nx, x);
  int ly = knot_interval(sp.ty, sp.ny, y);
  double bx[4], by[4];
  bspline3(sp.tx, lx, x, bx);
  bspline3(sp.ty, ly, y, by);
  double s = 0.0;
  const double* c = sp.c + (size_t)(lx - 3) * sp.ny + (ly - 3);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double r = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) r += by[j] * c[(size_t)i * sp.ny + j];
    s += bx[i] * r;
  }
  return s;
}

// ------------------------------------------------------------ LDS radix FFT
// Forward DFT X[k] = sum_n x[n] exp(-2 pi i n k / P) of B independent
// P-point sequences held in LDS (buf[b * P + n]), Stockham autosort with
// radix-8/4/2 passes; twiddles W_P^m from an LDS table tw[m], m < P.
template <typename C>
__device__ __forceinline__ void dft2(C& a, C& b) {
  C t = a;
  a = cadd(t, b);
  b = csub(t, b);
}
template <typename C>
__device__ __forceinline__ void dft4(C* x) {
  C a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
  C a2 = cadd(x[1], x[3]), a3 = cmul_mi(csub(x[1], x[3]));
  x[0] = cadd(a0, a2);
  x[2] = csub(a0, a2);
  x[1] = cadd(a1, a3);
  x[3] = csub(a1, a3);
}
template <typename C>
__device__ __forceinline__ void dft8(C* x) {
  using R = decltype(C::x);
  C e[4] = {x[0], x[2], x[4], x[6]};
  C o[4] = {x[1], x[3], x[5], x[7]};
  dft4(e);
  dft4(o);
  const R r = (R)0.70710678118654752440;
  // W8^1 = (1 - i)/sqrt2, W8^2 = -i, W8^3 = (-1 - i)/sqrt2
  C o1 = cmk<C>((o[1].x + o[1].y) * r, (o[1].y - o[1].x) * r);
  C o2 = cmul_mi(o[2]);
  C o3 = cmk<C>((-o[3].x + o[3].y) * r, (-o[3].y - o[3].x) * r);
  x[0] = cadd(e[0], o[0]);
  x[4] = csub(e[0], o[0]);
  x[1] = cadd(e[1], o1);
  x[5] = csub(e[1], o1);
  x[2] = cadd(e[2], o2);
  x[6] = csub(e[2], o2);
  x[3] = cadd(e[3], o3);
  x[7] = csub(e[3], o3);
}

template <int R, typename C>
__device__ __forceinline__ void dftR(C* x) {
  if constexpr (R == 8) dft8(x);
  else if constexpr (R == 4) dft4(x);
  else dft2(x[0], x[1]);
}

// Twiddle table W_P^m = exp(-2 pi i m / P), m < P.  Up to P = 1024 it is
// stored whole; beyond, as two levels W_P^m = H[m >> 6] L[m & 63] with
// H[q] = W_P^(64 q) and L[l] = W_P^l, so that LDS holds P/64 + 64 entries.
template <int P>
struct TwLayout {
  static constexpr bool kTwo = P > 1024;
  static constexpr int kHi = kTwo ? (P >> 6) : P;
  static constexpr int kEntries = kTwo ? kHi + 64 : P;
};

template <int P, typename C>
__device__ __forceinline__ C tw_get(const C* tw, int m) {
  if constexpr (TwLayout<P>::kTwo) {
    return cmul(tw[m >> 6], tw[TwLayout<P>::kHi + (m & 63)]);
  } else {
    return tw[m];
  }
}

template <int P, int NT, typename C>
__device__ inline void init_twiddles(C* tw) {
  using R = decltype(C::x);
  for (int m = threadIdx.x; m < TwLayout<P>::kEntries; m += NT) {
    int e;
    if constexpr (TwLayout<P>::kTwo) e = (m < TwLayout<P>::kHi) ? (m << 6) : (m - TwLayout<P>::kHi);
    else e = m;
    double s, c;
    sincospi(-2.0 * (double)e / (double)P, &s, &c);
    tw[m] = cmk<C>((R)c, (R)s);
  }
}

// One Stockham pass of radix RAD over B sequences of length P, span p.
// Each thread handles butterflies bf = tid, tid + NT, ...
template <int RAD, int NT, int MAXBF, int P, typename C>
__device__ __forceinline__ void stockham_pass(C* buf, const C* tw, int B, int p) {
  constexpr int nb = P / RAD;    // butterflies per sequence
  const int total = B * nb;
  const int tid = threadIdx.x;
  C x[MAXBF][RAD];
  int base[MAXBF];
#pragma unroll
  for (int q = 0; q < MAXBF; ++q) {
    int bf = tid + q * NT;
    base[q] = -1;
    if (bf < total) {
      int b = bf / nb, i = bf - b * nb;
      int k = i & (p - 1);
      const C* src = buf + (size_t)b * P;
#pragma unroll
      for (int r = 0; r < RAD; ++r) x[q][r] = src[i + r * nb];
      // twiddle W_{p RAD}^{r k} = W_P^{r k P/(p RAD)}
      int step = (P / (p * RAD)) * k;
#pragma unroll
      for (int r = 1; r < RAD; ++r) x[q][r] = cmul(x[q][r], tw_get<P>(tw, (r * step) & (P - 1)));
      dftR<RAD>(x[q]);
      base[q] = b * P + (i - k) * RAD + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXBF; ++q) {
    if (base[q] >= 0) {
#pragma unroll
      for (int r = 0; r < RAD; ++r) buf[base[q] + r * p] = x[q][r];
    }
  }
  __syncthreads();
}

// Full forward FFT of B sequences of length P (power of two, P >= 2).
// MAXBF8 bounds the radix-8 butterflies one thread handles per pass.
template <int NT, int MAXBF8, int P, typename C>
__device__ inline void fft_lds(C* buf, const C* tw, int B) {
  int p = 1;
  int rem = P;
  while (rem >= 8) {
    stockham_pass<8, NT, MAXBF8, P>(buf, tw, B, p);
    p *= 8;
    rem /= 8;
  }
  if (rem == 4) {
    stockham_pass<4, NT, 2 * MAXBF8, P>(buf, tw, B, p);
  } else if (rem == 2) {
    stockham_pass<2, NT, 4 * MAXBF8, P>(buf, tw, B, p);
  }
}

__device__ inline double row_valid_sentinel() { return -DBL_MAX; }  // base.py:424-425

// Knots of the distance spline staged in LDS: tx[nx + 4] then ty[ny + 4].
__device__ inline void stage_knots(const EvDev& ev, double* lds) {
  const int nx4 = ev.sp.nx + 4, ny4 = ev.sp.ny + 4;
  for (int i = threadIdx.x; i < nx4; i += blockDim.x) lds[i] = ev.sp.tx[i];
  for (int i = threadIdx.x; i < ny4; i += blockDim.x) lds[nx4 + i] = ev.sp.ty[i];
}

// y-part of the bicubic interpolant (fixed per row): interval and basis.
struct SplineY {
  int ly;
  double by[4];
  bool inside;
};

__device__ inline SplineY spline_y(const EvDev& ev, const double* ty, double y) {
  SplineY r;
  r.inside = (y >= ev.sp.ymin && y <= ev.sp.ymax);
  r.ly = 3;
  r.by[0] = r.by[1] = r.by[2] = r.by[3] = 0.0;
  if (r.inside) {
    r.ly = knot_interval(ty, ev.sp.ny, y);
    bspline3(ty, r.ly, y, r.by);
  }
  return r;
}

// BoundedRectBivariateSpline value at (x, y) with the y-part precomputed;
// -inf outside the bounding box (calculus.py:252-260).
__device__ inline double spline_x(const EvDev& ev, const double* tx, const SplineY& yy, double x) {
  if (!yy.inside || !(x >= ev.sp.xmin && x <= ev.sp.xmax)) return -INFINITY;
  const int lx = knot_interval(tx, ev.sp.nx, x);
  double bx[4];
  bspline3(tx, lx, x, bx);
  const double* c = ev.sp.c + (size_t)(lx - 3) * ev.sp.ny + (yy.ly - 3);
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double* ci = c + (size_t)i * ev.sp.ny;
    s += bx[i] * (yy.by[0] * ci[0] + yy.by[1] * ci[1] + yy.by[2] * ci[2] + yy.by[3] * ci[3]);
  }
  return s;
}

// Per-row constants of the TaylorF2 time transforms (k_rowconst).
struct RowW {
  double c[9];                 // phase coefficients (cycles), tf2_basis order
  double dt[GWL_MAX_DET];      // detector time shifts [s]
  float cre[GWL_MAX_DET];      // complex detector amplitude / sigma
  float cim[GWL_MAX_DET];
  double hh;                   // sum over the support of |h|^2 sigma^2 / S
  double hscale, jitter;
  double cabs2[GWL_MAX_DET];   // |complex detector amplitude / sigma|^2, float64
  // the wave-per-row transform's reference detector ref (the largest
  // amplitude C_ref): its delay and the phase of C_ref enter the float64
  // phase, |C_ref| scales the result, and the other detectors (in index
  // order, ref skipped) carry rho_k = C_k / C_ref and rotate relative to it
  float rre[GWL_MAX_DET - 1];
  float rim[GWL_MAX_DET - 1];
  double amp;                  // |C_ref|
  double c4r;                  // c[4] - arg(C_ref) / 2 pi
  int kb, ke, valid, ref;
  // the support [kb, ke] covers every detector's mask, outside which the
  // data are zero: no element needs masking
  int nomask, pad;
};

// Marginalised lnL at one time (base.py:775-792): distance spline (y-part
// precomputed per row), else ln I0 (phase) or the real part.
__device__ inline double epi_value(const EvDev& ev, const double* knots, const SplineY& yy,
                                   double hh, double hs, double xr, double xi) {
  if (ev.flags & GWL_DISTANCE_MARG) {
    const double d = (ev.flags & GWL_PHASE_MARG) ? hypot(xr, xi) * hs : xr * hs;
    return spline_x(ev, knots, yy, d);
  }
  if (ev.flags & GWL_PHASE_MARG) return ln_i0(hypot(xr, xi)) - hh / 2.0;
  return xr - hh / 2.0;
}

// Copy of a wave-uniform value held in scalar registers.
template <typename T>
__device__ __forceinline__ T uniform_copy(const T& v) {
  static_assert(sizeof(T) % 4 == 0, "word-sized type");
  T out;
  const int* src = reinterpret_cast<const int*>(&v);
  int* dst = reinterpret_cast<int*>(&out);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k) dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
  return out;
}

// ln of the geocent_time prior density at t (base.py:808-819 weighs each
// time by prior.prob(t) delta_tc): the event's table, linear between its
// points; -inf outside the table or where the density is zero.
__device__ inline double time_log_density(const double* tpd, double t0, double ih, int n,
                                          double t) {
  const double x = (t - t0) * ih;
  if (!(x >= 0.0) || !(x <= (double)(n - 1))) return -INFINITY;
  const int k = min((int)x, n - 2);
  const double w = x - (double)k;
  const double p = fma(w, tpd[k + 1] - tpd[k], tpd[k]);
  return p > 0.0 ? log(p) : -INFINITY;
}

// The time weight of one transform bin: with a Uniform prior nothing here
// (the constant delta_tc / (max - min) is applied after the logsumexp, see
// time_lse_scale), else ln prior.prob(t).
__device__ __forceinline__ double time_log_weight(const EvDev& ev, double t) {
  return ev.tpd ? time_log_density(ev.tpd, ev.tp_t0, ev.tp_ih, ev.tp_n, t) : 0.0;
}
__device__ __forceinline__ double time_lse_scale(const EvDev& ev, double delta_tc) {
  return ev.tpd ? delta_tc : delta_tc / (ev.tmax - ev.tmin);
}

// Running logsumexp state (max m, sum s of exp(v - m)).
__device__ __forceinline__ void lse_push(double& m, double& s, double v) {
  if (v > m) {
    s = (m > -INFINITY ? s * exp(m - v) : 0.0) + 1.0;
    m = v;
  } else if (v > -INFINITY) {
    s += exp(v - m);
  }
}

// Merge the logsumexp states of the 64 lanes of a wavefront.
__device__ __forceinline__ void lse_wave(double& m, double& s) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double mo = __shfl_xor(m, off, 64), so = __shfl_xor(s, off, 64);
    const double mn = fmax(m, mo);
    if (mn > -INFINITY) {
      s = (m > -INFINITY ? s * exp(m - mn) : 0.0) + (mo > -INFINITY ? so * exp(mo - mn) : 0.0);
    } else {
      s = 0.0;
    }
    m = mn;
  }
}

}  // namespace gwl
