// Calibration-marginalised likelihood (bilby/gw/likelihood/base.py:332-346
// calculate_snrs with calibration_marginalization, :860-877
// calibration_marginalized_likelihood).  With time marginalisation
// (base.py:305-324) every curve needs its own transform instead; that path
// (k_cal_time_w in gwl_time_wave.hip) uses the curves in the transform's
// decimated order and per-curve prefix sums for <h|h>, built here.
//
// For C response curves cal_c(f) on each detector's masked bins (the
// calibration lookup, base.py:1037-1051), per row
//   <d|h>_c = 4/T sum_{d,f} conj(d) h cal_c / S
//   <h|h>_c = 4/T sum_{d,f} |h|^2 |cal_c|^2 / S
//   ln L    = logsumexp_c(ln L_c) - ln C
// with ln L_c the plain / phase / distance marginalised value of each pair.
//
// Dense contraction on MFMA: k_cal_prep writes, per row, the per-bin
// integrands x = 4/T conj(d) h / S (real and imaginary halves) and
// y = 4/T |h|^2 / S as GEMM operands (float64, float32 or bfloat16 hi/lo)
// and their float64 totals D0, H0; two GEMMs against the curves minus one,
//   G1 = [Re x | Im x] . [Re dc | -Im dc ; Im dc | Re dc]^T,  dc = cal - 1
//   G2 = y . (|cal|^2 - 1)^T
// give <d|h>_c = D0 + G1 and <h|h>_c = H0 + G2, so the reduced-precision
// GEMMs only carry the calibration correction (|cal - 1| ~ 0.1), not the
// whole inner product; k_cal_epi forms the per-curve values and their
// logsumexp, one wavefront per row with the lanes over curves.
#include "gwl_internal.h"
#include "gwl_bins.h"
#include "gwl_window.h"

#include <cmath>
#include <vector>

using gwl::set_error;

namespace {

constexpr int kGM = 128;  // GEMM tile (M, N multiples), gwl_gemm
constexpr int kGK = 32;   // GEMM K multiple
// float32 GEMMs split K into at most this many chunks (about 2048 workgroups
// for the two long-K contractions, whose row x curve tiles alone are too
// few to fill 256 CUs); the epilogue adds the chunk partials in float64
constexpr int kCalSplit = 8;

size_t up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

struct gwl_calib;

namespace gwl {
int calib_create_time(gwl_event* ev, int nc, const double* curves, int K, gwl_calib** out);
}

struct gwl_calib {
  gwl_event* ev = nullptr;
  int nc = 0;           // curves
  int K = 0;            // masked bins over all detectors
  int K1p = 0, K2p = 0; // padded K of the two GEMMs (2K, K)
  int N1p = 0, N2p = 0; // padded N (2C, C)
  void* data = nullptr; // curve operands: fp64, fp32, bf16 hi/lo of B1 and B2
  size_t data_bytes = 0;
  double* b1_64 = nullptr;
  double* b2_64 = nullptr;
  float* b1_32 = nullptr;
  float* b2_32 = nullptr;
  __bf16* b1_hi = nullptr;
  __bf16* b1_lo = nullptr;
  __bf16* b2_hi = nullptr;
  __bf16* b2_lo = nullptr;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  int64_t ws_rows = 0;
  // time marginalisation: curves [nc][nd][N] float2 in decimated order and
  // prefix sums [nc][nd][N + 2] of f^-7/3 sigma^2 / S |cal|^2 (natural order)
  bool timem = false;
  float2* cal_t = nullptr;
  double* calps = nullptr;
  // float64 evaluations: the curves' float32 remainders (same layout as
  // cal_t, cal = hi + lo to 48 bits) and W_N^m = exp(-2 pi i m / N), m < N
  float2* cal_lo = nullptr;
  double2* wN = nullptr;
};

namespace gwl {
namespace {

enum { CAL_FP64 = 0, CAL_FP32 = 1, CAL_BF16X3 = 2 };

struct CalArgs {
  EvDev ev;
  const double* params;
  int ld;
  long long n;
  int Mp, K, K1p, K2p;
  int nc, N1p, N2p;
  int mode;
  void* a1;          // [Mp][K1p]
  void* a1_lo;       // bf16 low halves
  void* a2;          // [Mp][K2p]
  void* a2_lo;
  double* tot;       // [Mp][3]: Re D0, Im D0, H0
  int* valid;        // [Mp]
  const void* g1;    // [nz1][Mp][N1p] (double, nz1 = 1; or float chunk partials)
  const void* g2;    // [nz2][Mp][N2p]
  int nz1, nz2;      // K chunks of the float32 GEMMs, summed in float64 here
  double* out;
  double* per_curve; // [n][nc] or null
};

__device__ __forceinline__ void put(const CalArgs& a, void* base, void* lo, size_t i, double v) {
  if (a.mode == CAL_FP64) {
    static_cast<double*>(base)[i] = v;
  } else if (a.mode == CAL_FP32) {
    static_cast<float*>(base)[i] = (float)v;
  } else {
    const float f = (float)v;
    const __bf16 h = (__bf16)f;
    static_cast<__bf16*>(base)[i] = h;
    static_cast<__bf16*>(lo)[i] = (__bf16)(f - (float)h);
  }
}

// One workgroup per operand row: the integrands of every masked bin of every
// detector in detector order (interferometer.py:303-368, base.py:332-346),
// float64 phase argument and float64 arithmetic; rows past n and padding
// columns are written as zeros.
template <int WAVE>
__global__ __launch_bounds__(256) void k_cal_prep(CalArgs a) {
  const EvDev& ev = a.ev;
  const long long r = blockIdx.x;
  const bool live = r < a.n;
  __shared__ SampleConst sc[1];
  __shared__ SampleFast<double> sf[1];
  __shared__ double red[4][3];
  const int tid = threadIdx.x;
  if (tid == 0) {
    if (live) {
      compute_sample_const(ev, a.params + r * a.ld, sc[0]);
    } else {
      sc[0].valid = 0;
      sc[0].kb = 1;
      sc[0].ke = 0;
    }
    for (int d = 0; d < GWL_MAX_DET; ++d) {
      sf[0].cre[d] = d < ev.nd ? sc[0].cre[d] : 0.0;
      sf[0].cim[d] = d < ev.nd ? sc[0].cim[d] : 0.0;
      sf[0].cabs2[d] = d < ev.nd ? sc[0].cabs2[d] : 0.0;
    }
    sf[0].k2 = sc[0].c[0];
    a.valid[r] = live ? sc[0].valid : 0;
  }
  __syncthreads();
  const double norm = 4.0 / ev.T;
  double sr = 0.0, si = 0.0, sh = 0.0;
  int k0 = 0;
  for (int d = 0; d < ev.nd; ++d) {
    const int lo = ev.mlo[d], cnt = ev.mhi[d] - ev.mlo[d] + 1;
    SampleConst one = sc[0];
    SampleFast<double> onef = sf[0];
    one.dt[0] = sc[0].dt[d];
    onef.cre[0] = sf[0].cre[d];
    onef.cim[0] = sf[0].cim[d];
    onef.cabs2[0] = sf[0].cabs2[d];
    for (int j = tid; j < cnt; j += blockDim.x) {
      const int nbin = lo + j;
      double xr = 0.0, xi = 0.0, y = 0.0;
      if (live && nbin >= one.kb && nbin <= one.ke) {
        double2 dw[1];
        double w[1];
        if (nbin < ev.N) {
          const int i = (nbin % ev.Q) * ev.P + nbin / ev.Q;
          dw[0] = ev.dw64[(size_t)d * ev.N + i];
          w[0] = ev.w64[(size_t)d * ev.N + i];
        } else {
          dw[0] = make_double2(ev.dwN[d][0], ev.dwN[d][1]);
          w[0] = ev.wN[d];
        }
        double2 I[1];
        double h[1] = {0.0};
        bin_rows<double, WAVE, 1, 1>(nbin, (double)nbin * ev.df, dw, w, &one, &onef, I, h);
        xr = norm * I[0].x;
        xi = norm * I[0].y;
        y = norm * h[0];
      }
      const size_t k = (size_t)k0 + j;
      put(a, a.a1, a.a1_lo, (size_t)r * a.K1p + k, xr);
      put(a, a.a1, a.a1_lo, (size_t)r * a.K1p + a.K + k, xi);
      put(a, a.a2, a.a2_lo, (size_t)r * a.K2p + k, y);
      sr += xr;
      si += xi;
      sh += y;
    }
    k0 += cnt;
  }
  for (int k = 2 * a.K + tid; k < a.K1p; k += blockDim.x) put(a, a.a1, a.a1_lo, (size_t)r * a.K1p + k, 0.0);
  for (int k = a.K + tid; k < a.K2p; k += blockDim.x) put(a, a.a2, a.a2_lo, (size_t)r * a.K2p + k, 0.0);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    si += __shfl_xor(si, off, 64);
    sh += __shfl_xor(sh, off, 64);
  }
  if ((tid & 63) == 0) {
    red[tid >> 6][0] = sr;
    red[tid >> 6][1] = si;
    red[tid >> 6][2] = sh;
  }
  __syncthreads();
  if (tid == 0) {
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    for (int wv = 0; wv < 4; ++wv) {
      t0 += red[wv][0];
      t1 += red[wv][1];
      t2 += red[wv][2];
    }
    a.tot[3 * r + 0] = t0;
    a.tot[3 * r + 1] = t1;
    a.tot[3 * r + 2] = t2;
  }
}

__device__ __forceinline__ double gval(const CalArgs& a, const void* g, size_t i, int nz,
                                       size_t zs) {
  if (a.mode == CAL_FP64) return static_cast<const double*>(g)[i];
  const float* f = static_cast<const float*>(g);
  double s = 0.0;
  for (int z = 0; z < nz; ++z) s += (double)f[(size_t)z * zs + i];
  return s;
}

// One wavefront per row, lanes over curves: ln L_c and their logsumexp
// minus ln C (base.py:860-877; the value per curve as base.py:775-792).
__global__ __launch_bounds__(256) void k_cal_epi(CalArgs a) {
  extern __shared__ double knots[];
  const EvDev& ev = a.ev;
  const bool dist = (ev.flags & GWL_DISTANCE_MARG) != 0;
  if (dist) stage_knots(ev, knots);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long long r = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= a.n) return;
  if (!a.valid[r]) {
    if (lane == 0) a.out[r] = row_valid_sentinel();
    return;
  }
  const double d0r = a.tot[3 * r + 0], d0i = a.tot[3 * r + 1], h0 = a.tot[3 * r + 2];
  const double hs = dist ? a.params[r * a.ld + GWL_P_LUMINOSITY_DISTANCE] / ev.dref : 1.0;
  double m = -INFINITY, s = 0.0;
  for (int c = lane; c < a.nc; c += 64) {
    const size_t z1 = (size_t)a.Mp * a.N1p, z2 = (size_t)a.Mp * a.N2p;
    const double dr = d0r + gval(a, a.g1, (size_t)r * a.N1p + c, a.nz1, z1);
    const double di = d0i + gval(a, a.g1, (size_t)r * a.N1p + a.nc + c, a.nz1, z1);
    const double hh = h0 + gval(a, a.g2, (size_t)r * a.N2p + c, a.nz2, z2);
    double v;
    if (dist) {
      const double dd = (ev.flags & GWL_PHASE_MARG) ? hypot(dr, di) * hs : dr * hs;
      const SplineY yy = spline_y(ev, knots + ev.sp.nx + 4, hh * hs * hs);
      v = spline_x(ev, knots, yy, dd);
    } else if (ev.flags & GWL_PHASE_MARG) {
      v = ln_i0(hypot(dr, di)) - hh / 2.0;
    } else {
      v = dr - hh / 2.0;
    }
    if (a.per_curve) a.per_curve[(size_t)r * a.nc + c] = v;
    lse_push(m, s, v);
  }
  lse_wave(m, s);
  if (lane == 0) a.out[r] = (m > -INFINITY ? m + log(s) : -INFINITY) - log((double)a.nc);
}

size_t cal_ws_bytes(const gwl_calib* c, long long rows) {
  if (c->timem) return cal_time_ws_bytes(rows, c->nc);
  const size_t Mp = up((size_t)rows, kGM);
  size_t b = 0;
  b += up(Mp * c->K1p * 8, 256) + up(Mp * c->K1p * 4, 256);  // A1 (largest element) + bf16 lo
  b += up(Mp * c->K2p * 8, 256) + up(Mp * c->K2p * 4, 256);
  // G1, G2: float64, or up to kCalSplit float32 K-chunk partials
  b += up(Mp * c->N1p * 4 * kCalSplit, 256) + up(Mp * c->N2p * 4 * kCalSplit, 256);
  b += up(Mp * 3 * 8, 256) + up(Mp * 4, 256);                 // totals, valid
  return b;
}

int cal_reserve(gwl_calib* c, long long rows) {
  const size_t need = cal_ws_bytes(c, rows);
  if (need <= c->ws_bytes) return GWL_OK;
  GWL_HIP_CHECK(hipSetDevice(c->ev->ctx->device));
  if (c->ws) (void)hipFree(c->ws);
  c->ws = nullptr;
  c->ws_bytes = 0;
  if (hipMalloc(&c->ws, need) != hipSuccess) {
    set_error("hipMalloc failed for the calibration workspace");
    return GWL_E_NOMEM;
  }
  c->ws_bytes = need;
  c->ws_rows = rows;
  return GWL_OK;
}

}  // namespace

// Time-marginalised calibration tables (see gwl_calib): the curve values of
// each detector's masked bins n < N at decimated element (n % Q) P + n / Q
// (1 elsewhere; the integrand is zero there), and the prefix sums over bins
// 0..N of w_d(n) |cal(n)|^2 with w_d(n) = f^-7/3 sigma^2 / S_d the event's
// own prefix-sum increments (|cal|^2 = 1 outside the mask, where w_d = 0).
int calib_create_time(gwl_event* ev, int nc, const double* curves, int K, gwl_calib** out) {
  const EvDev& d = ev->dev;
  if (!cal_time_eligible(ev)) {
    set_error("gwl_calib_create: calibration with time marginalisation needs a TaylorF2 event on "
              "the wave-per-row transform with a time window of at most 512 bins");
    return GWL_E_UNSUPPORTED;
  }
  const int nd = d.nd, N = d.N;
  GWL_HIP_CHECK(hipSetDevice(ev->ctx->device));
  std::vector<double> hps((size_t)nd * (N + 2));
  GWL_HIP_CHECK(hipMemcpy(hps.data(), d.hps, hps.size() * 8, hipMemcpyDeviceToHost));
  std::vector<float2> tab, tab_lo;
  std::vector<double> ps;
  std::vector<double2> wn;
  try {
    tab.assign((size_t)nc * nd * N, make_float2(1.f, 0.f));
    tab_lo.assign((size_t)nc * nd * N, make_float2(0.f, 0.f));
    ps.assign((size_t)nc * nd * (N + 2), 0.0);
    wn.resize((size_t)N);
  } catch (...) {
    set_error("out of host memory for the calibration tables");
    return GWL_E_NOMEM;
  }
  for (int cv = 0; cv < nc; ++cv) {
    int k0 = 0;
    for (int k = 0; k < nd; ++k) {
      const int lo = d.mlo[k], hi = d.mhi[k];
      float2* t = tab.data() + ((size_t)cv * nd + k) * N;
      double* p = ps.data() + ((size_t)cv * nd + k) * (N + 2);
      const double* h = hps.data() + (size_t)k * (N + 2);
      double acc = 0.0;
      p[0] = 0.0;
      for (int n = 0; n <= N; ++n) {
        double c2 = 1.0;
        if (n >= lo && n <= hi) {
          const double* v = curves + ((size_t)cv * K + k0 + (n - lo)) * 2;
          c2 = v[0] * v[0] + v[1] * v[1];
          if (n < N) {
            const size_t e = (size_t)(n % d.Q) * d.P + n / d.Q;
            const float hr = (float)v[0], hi = (float)v[1];
            t[e] = make_float2(hr, hi);
            tab_lo[((size_t)cv * nd + k) * N + e] =
                make_float2((float)(v[0] - (double)hr), (float)(v[1] - (double)hi));
          }
        }
        acc += (h[n + 1] - h[n]) * c2;
        p[n + 1] = acc;
      }
      k0 += hi - lo + 1;
    }
  }
  gwl_calib* c = new (std::nothrow) gwl_calib;
  if (!c) { set_error("out of host memory"); return GWL_E_NOMEM; }
  c->ev = ev;
  c->nc = nc;
  c->K = K;
  c->timem = true;
  {
    const long double step = 2.0L * 3.141592653589793238462643383279502884L / (long double)N;
    for (int m = 0; m < N; ++m) {
      const long double a = step * (long double)m;
      wn[m] = make_double2((double)std::cos(a), (double)-std::sin(a));
    }
  }
  const size_t bt = up(tab.size() * sizeof(float2), 256), bp = up(ps.size() * 8, 256);
  const size_t bw = up(wn.size() * sizeof(double2), 256);
  c->data_bytes = 2 * bt + bp + bw;
  if (hipMalloc(&c->data, c->data_bytes) != hipSuccess) {
    delete c;
    set_error("hipMalloc failed for the calibration curves");
    return GWL_E_NOMEM;
  }
  c->cal_t = static_cast<float2*>(c->data);
  c->calps = reinterpret_cast<double*>(static_cast<char*>(c->data) + bt);
  c->cal_lo = reinterpret_cast<float2*>(static_cast<char*>(c->data) + bt + bp);
  c->wN = reinterpret_cast<double2*>(static_cast<char*>(c->data) + 2 * bt + bp);
  hipError_t err = hipMemcpy(c->cal_t, tab.data(), tab.size() * sizeof(float2), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(c->calps, ps.data(), ps.size() * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = hipMemcpy(c->cal_lo, tab_lo.data(), tab_lo.size() * sizeof(float2), hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = hipMemcpy(c->wN, wn.data(), wn.size() * sizeof(double2), hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    (void)hipFree(c->data);
    delete c;
    return hip_fail(err, "gwl_calib_create upload");
  }
  *out = c;
  return GWL_OK;
}

}  // namespace gwl

extern "C" {

int gwl_calib_create(gwl_event* ev, int32_t n_curves, const double* curves, gwl_calib** out) {
  if (!ev || n_curves < 1 || !curves || !out) {
    set_error("gwl_calib_create: invalid argument");
    return GWL_E_INVALID;
  }
  const gwl::EvDev& d = ev->dev;
  if (d.approx == GWL_HOST_POLARIZATIONS) {
    set_error("gwl_calib_create: calibration marginalisation needs a source model the library "
              "evaluates (sine-Gaussian or TaylorF2)");
    return GWL_E_UNSUPPORTED;
  }
  int K = 0;
  for (int k = 0; k < d.nd; ++k) K += d.mhi[k] - d.mlo[k] + 1;
  if (d.flags & GWL_TIME_MARG) return gwl::calib_create_time(ev, n_curves, curves, K, out);
  gwl_calib* c = new (std::nothrow) gwl_calib;
  if (!c) { set_error("out of host memory"); return GWL_E_NOMEM; }
  c->ev = ev;
  c->nc = n_curves;
  c->K = K;
  c->K1p = (int)up((size_t)2 * K, kGK);
  c->K2p = (int)up((size_t)K, kGK);
  c->N1p = (int)up((size_t)2 * n_curves, kGM);
  c->N2p = (int)up((size_t)n_curves, kGM);
  const size_t n1 = (size_t)c->N1p * c->K1p, n2 = (size_t)c->N2p * c->K2p;
  std::vector<double> b1(n1, 0.0), b2(n2, 0.0);
  for (int cv = 0; cv < n_curves; ++cv)
    for (int k = 0; k < K; ++k) {
      const double cr = curves[((size_t)cv * K + k) * 2 + 0];
      const double ci = curves[((size_t)cv * K + k) * 2 + 1];
      const double dr = cr - 1.0;
      b1[(size_t)cv * c->K1p + k] = dr;
      b1[(size_t)cv * c->K1p + K + k] = -ci;
      b1[(size_t)(n_curves + cv) * c->K1p + k] = ci;
      b1[(size_t)(n_curves + cv) * c->K1p + K + k] = dr;
      b2[(size_t)cv * c->K2p + k] = cr * cr + ci * ci - 1.0;
    }
  std::vector<float> f1(n1), f2(n2);
  std::vector<__bf16> h1(n1), l1(n1), h2(n2), l2(n2);
  for (size_t i = 0; i < n1; ++i) {
    f1[i] = (float)b1[i];
    h1[i] = (__bf16)f1[i];
    l1[i] = (__bf16)(f1[i] - (float)h1[i]);
  }
  for (size_t i = 0; i < n2; ++i) {
    f2[i] = (float)b2[i];
    h2[i] = (__bf16)f2[i];
    l2[i] = (__bf16)(f2[i] - (float)h2[i]);
  }
  const size_t s64_1 = up(n1 * 8, 256), s64_2 = up(n2 * 8, 256);
  const size_t s32_1 = up(n1 * 4, 256), s32_2 = up(n2 * 4, 256);
  const size_t s16_1 = up(n1 * 2, 256), s16_2 = up(n2 * 2, 256);
  c->data_bytes = s64_1 + s64_2 + s32_1 + s32_2 + 2 * s16_1 + 2 * s16_2;
  if (hipSetDevice(ev->ctx->device) != hipSuccess || hipMalloc(&c->data, c->data_bytes) != hipSuccess) {
    delete c;
    set_error("hipMalloc failed for the calibration curves");
    return GWL_E_NOMEM;
  }
  char* p = static_cast<char*>(c->data);
  c->b1_64 = reinterpret_cast<double*>(p); p += s64_1;
  c->b2_64 = reinterpret_cast<double*>(p); p += s64_2;
  c->b1_32 = reinterpret_cast<float*>(p); p += s32_1;
  c->b2_32 = reinterpret_cast<float*>(p); p += s32_2;
  c->b1_hi = reinterpret_cast<__bf16*>(p); p += s16_1;
  c->b1_lo = reinterpret_cast<__bf16*>(p); p += s16_1;
  c->b2_hi = reinterpret_cast<__bf16*>(p); p += s16_2;
  c->b2_lo = reinterpret_cast<__bf16*>(p);
  hipError_t err = hipMemcpy(c->b1_64, b1.data(), n1 * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(c->b2_64, b2.data(), n2 * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(c->b1_32, f1.data(), n1 * 4, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(c->b2_32, f2.data(), n2 * 4, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(c->b1_hi, h1.data(), n1 * 2, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(c->b1_lo, l1.data(), n1 * 2, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(c->b2_hi, h2.data(), n2 * 2, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(c->b2_lo, l2.data(), n2 * 2, hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    (void)hipFree(c->data);
    delete c;
    return gwl::hip_fail(err, "gwl_calib_create upload");
  }
  *out = c;
  return GWL_OK;
}

int gwl_calib_destroy(gwl_calib* c) {
  if (!c) return GWL_OK;
  (void)hipSetDevice(c->ev->ctx->device);
  if (c->data) (void)hipFree(c->data);
  if (c->ws) (void)hipFree(c->ws);
  delete c;
  return GWL_OK;
}

int gwl_calib_reserve(gwl_calib* c, int64_t rows) {
  if (!c || rows < 1) { set_error("gwl_calib_reserve: invalid argument"); return GWL_E_INVALID; }
  return gwl::cal_reserve(c, rows);
}

int gwl_calib_device_bytes(gwl_calib* c, int64_t* bytes) {
  if (!c || !bytes) { set_error("gwl_calib_device_bytes: invalid argument"); return GWL_E_INVALID; }
  *bytes = (int64_t)(c->data_bytes + c->ws_bytes);
  return GWL_OK;
}

int gwl_calib_eval(gwl_calib* c, const double* params, int64_t n, int32_t ld, int32_t mode,
                   double* out, double* per_curve, void* stream) {
  if (c && n == 0) return GWL_OK;  // an empty batch needs no buffers
  if (!c || !params || !out || n < 0 || ld < GWL_NPARAM || mode < 0 || mode > 2) {
    set_error("gwl_calib_eval: invalid argument");
    return GWL_E_INVALID;
  }
  if (n == 0) return GWL_OK;
  gwl_event* ev = c->ev;
  if ((ev->dev.flags & GWL_DISTANCE_MARG) && !ev->has_spline) {
    set_error("gwl_calib_eval: distance marginalisation requested but no spline attached");
    return GWL_E_INVALID;
  }
  if (n > (int64_t)1 << 30) {
    set_error("gwl_calib_eval: too many rows in one call");
    return GWL_E_INVALID;
  }
  if (c->timem && mode == gwl::CAL_BF16X3) {
    set_error("gwl_calib_eval: with time marginalisation the curves' transforms run in float32 "
              "or float64 (mode GWL_ROQ_FP32 or GWL_ROQ_FP64)");
    return GWL_E_UNSUPPORTED;
  }
  if (int rc = gwl::cal_reserve(c, n)) return rc;
  GWL_HIP_CHECK(hipSetDevice(ev->ctx->device));
  (void)hipGetLastError();
  if (c->timem) {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (int rc = gwl::timing_begin(ev, st)) return rc;
    if (int rc = gwl::launch_cal_time(ev, params, n, ld, c->nc, c->cal_t, c->calps, c->ws,
                                      per_curve, out, st, mode == gwl::CAL_FP64, c->cal_lo,
                                      c->wN))
      return rc;
    return gwl::timing_end(ev, st);
  }
  const int Mp = (int)up((size_t)n, kGM);
  gwl::CalArgs a;
  a.ev = ev->dev;
  a.params = params;
  a.ld = ld;
  a.n = n;
  a.Mp = Mp;
  a.K = c->K;
  a.K1p = c->K1p;
  a.K2p = c->K2p;
  a.nc = c->nc;
  a.N1p = c->N1p;
  a.N2p = c->N2p;
  a.mode = mode;
  const size_t el = mode == gwl::CAL_FP64 ? 8 : mode == gwl::CAL_FP32 ? 4 : 2;
  char* w = static_cast<char*>(c->ws);
  a.a1 = w;
  a.a1_lo = w + (size_t)Mp * c->K1p * 2;  // bf16: low halves follow the high halves
  w += up((size_t)Mp * c->K1p * 8, 256) + up((size_t)Mp * c->K1p * 4, 256);
  a.a2 = w;
  a.a2_lo = w + (size_t)Mp * c->K2p * 2;
  w += up((size_t)Mp * c->K2p * 8, 256) + up((size_t)Mp * c->K2p * 4, 256);
  void* g1 = w;
  w += up((size_t)Mp * c->N1p * 4 * kCalSplit, 256);
  void* g2 = w;
  w += up((size_t)Mp * c->N2p * 4 * kCalSplit, 256);
  a.tot = reinterpret_cast<double*>(w);
  w += up((size_t)Mp * 3 * 8, 256);
  a.valid = reinterpret_cast<int*>(w);
  a.g1 = g1;
  a.g2 = g2;
  a.out = out;
  a.per_curve = per_curve;
  (void)el;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (int rc = gwl::timing_begin(ev, st)) return rc;
  if (ev->dev.approx == GWL_TAYLORF2)
    hipLaunchKernelGGL(gwl::k_cal_prep<GWL_TAYLORF2>, dim3((unsigned)Mp), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(gwl::k_cal_prep<GWL_SINE_GAUSSIAN>, dim3((unsigned)Mp), dim3(256), 0, st, a);
  GWL_HIP_CHECK(hipGetLastError());
  const void* b1 = mode == gwl::CAL_FP64 ? (const void*)c->b1_64
                   : mode == gwl::CAL_FP32 ? (const void*)c->b1_32 : (const void*)c->b1_hi;
  const void* b2 = mode == gwl::CAL_FP64 ? (const void*)c->b2_64
                   : mode == gwl::CAL_FP32 ? (const void*)c->b2_32 : (const void*)c->b2_hi;
  const void* b1lo = mode == gwl::CAL_BF16X3 ? (const void*)c->b1_lo : nullptr;
  const void* b2lo = mode == gwl::CAL_BF16X3 ? (const void*)c->b2_lo : nullptr;
  a.nz1 = a.nz2 = 1;
  // split K: chunks (multiples of 32 columns) sized for ~2048 workgroups
  auto chunk = [&](int Np, int Kp, int* nz) {
      const long long tiles = (long long)(Np / kGM) * (Mp / kGM);
      const int want = (int)std::max<long long>(1, std::min<long long>(kCalSplit,
                                                                         (2048 + tiles - 1) / tiles));
      const int kc = (int)up((size_t)((Kp + want - 1) / want), kGK);
      *nz = (Kp + kc - 1) / kc;
      return kc;
  };
  if (mode == gwl::CAL_FP32) {
    const int kc1 = chunk(c->N1p, c->K1p, &a.nz1);
    const int kc2 = chunk(c->N2p, c->K2p, &a.nz2);
    if (int rc = gwl::win_gemm_f32(static_cast<const float*>(a.a1),
                                   static_cast<const float*>(b1), static_cast<float*>(g1), Mp,
                                   c->N1p, c->K1p, kc1, st))
      return rc;
    if (int rc = gwl::win_gemm_f32(static_cast<const float*>(a.a2),
                                   static_cast<const float*>(b2), static_cast<float*>(g2), Mp,
                                   c->N2p, c->K2p, kc2, st))
      return rc;
  } else if (mode == gwl::CAL_BF16X3) {
    const int kc1 = chunk(c->N1p, c->K1p, &a.nz1);
    const int kc2 = chunk(c->N2p, c->K2p, &a.nz2);
    if (int rc = gwl::gemm_bf16x3_split(
            static_cast<const __bf16*>(a.a1), static_cast<const __bf16*>(a.a1_lo),
            static_cast<const __bf16*>(b1), static_cast<const __bf16*>(b1lo),
            static_cast<float*>(g1), Mp, c->N1p, c->K1p, kc1, st))
      return rc;
    if (int rc = gwl::gemm_bf16x3_split(
            static_cast<const __bf16*>(a.a2), static_cast<const __bf16*>(a.a2_lo),
            static_cast<const __bf16*>(b2), static_cast<const __bf16*>(b2lo),
            static_cast<float*>(g2), Mp, c->N2p, c->K2p, kc2, st))
      return rc;
  } else {
    if (int rc = gwl_gemm(ev->ctx, mode, a.a1, nullptr, b1, nullptr, g1, Mp, c->N1p, c->K1p,
                          stream))
      return rc;
    if (int rc = gwl_gemm(ev->ctx, mode, a.a2, nullptr, b2, nullptr, g2, Mp, c->N2p, c->K2p,
                          stream))
      return rc;
  }
  const size_t lds = (ev->dev.flags & GWL_DISTANCE_MARG)
                         ? (size_t)(ev->dev.sp.nx + ev->dev.sp.ny + 8) * 8 : 0;
  hipLaunchKernelGGL(gwl::k_cal_epi, dim3((unsigned)((n + 3) / 4)), dim3(256), lds, st, a);
  GWL_HIP_CHECK(hipGetLastError());
  return gwl::timing_end(ev, st);
}

}  // extern "C"
