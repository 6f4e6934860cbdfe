// Time-marginalised TaylorF2 likelihood by a chirp-z transform (gfx950).
//
// The window of d_inner_h_array (bilby/gw/likelihood/base.py:325-330),
//   X[j] = sum_n a[n] W^(n j),  W = exp(-2 pi i / N),  j in [jlo, jlo + L),
// with a[n] = sum_d h_d(n) conj(d_d(n)) / S_d(n) nonzero only on the row's
// support n in [kb, kb + B) (TaylorF2 ends at the ISCO frequency), is a zoom
// DFT from B inputs to L outputs.  With n = kb + m, j = jlo + t and
// m t = (m^2 + t^2 - (t - m)^2) / 2 (Bluestein):
//   X[jlo + t] = W^(kb jlo + kb t + t^2/2) sum_m u[m] v[t - m],
//   u[m] = a[kb + m] W^(m jlo + m^2/2),  v[k] = W^(-k^2/2),
// a linear convolution of length B + L - 1, done as a circular one of size
// M = 1024, 2048 or 4096 >= B + L - 1: a forward mixed-radix
// decimation-in-frequency FFT of u in LDS (radices 8, 8, 8, M/512; the output
// in digit-reversed order), a pointwise product with the per-event spectrum
// of v (precomputed in float64 in the same order, scaled by 1/M), and the
// inverse as the decimation-in-time mirror of the forward stages (natural
// order out).  Phases are exact integers modulo 2N before the float32
// rotation (N a power of two).  One 512-thread workgroup per row; the marginalisation epilogue
// (base.py:775-820) runs in the same workgroup.  Rows whose support does not
// fit M = 4096 are flagged and go through the wave-per-row transform.
#include "gwl_internal.h"

#include <cmath>
#include <complex>
#include <vector>

namespace gwl {

constexpr int kCT = 512;                   // threads per row
constexpr int kCMax = 4096;                // largest transform
constexpr int kCBuf = kCMax + kCMax / 16;  // padded LDS buffer (float2)

__device__ __forceinline__ int cpad(int e) { return e + (e >> 4); }

// W_4096^e from the quarter table q[k] = W_4096^k, k < 1024:
// W^(e + 1024 s) = (-i)^s W^e.
__device__ __forceinline__ float2 twc(const float2* q, const float2* unused, int e) {
  (void)unused;
  const float2 w = q[e & 1023];
  switch ((e >> 10) & 3) {
    case 0: return w;
    case 1: return make_float2(w.y, -w.x);
    case 2: return make_float2(-w.x, -w.y);
    default: return make_float2(-w.y, w.x);
  }
}

// One decimation-in-frequency stage of radix R, span h = 2^lh, on M points:
// butterfly (b, j) on legs b R h + j + k h, outputs k times W_{R h}^(j k).
template <int R>
__device__ __forceinline__ void dif_stage(float2* buf, int M, int lh, const float2* hi,
                                          const float2* lo) {
  const int h = 1 << lh;
  const int nb = M / R;
  const int tstep = (kCMax / R) >> lh;  // W_{R h}^(j k) = W_4096^(tstep j k)
  for (int beta = threadIdx.x; beta < nb; beta += kCT) {
    const int b = beta >> lh, j = beta & (h - 1);
    const int base = b * R * h + j;
    float2 x[R];
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = buf[cpad(base + k * h)];
    dftR<R>(x);
    const int e1 = tstep * j;
#pragma unroll
    for (int k = 1; k < R; ++k) x[k] = cmul(x[k], twc(hi, lo, (e1 * k) & (kCMax - 1)));
#pragma unroll
    for (int k = 0; k < R; ++k) buf[cpad(base + k * h)] = x[k];
  }
}

// The inverse of dif_stage (times R): inputs k times conj(W_{R h}^(j k)),
// then the inverse radix-R DFT.
template <int R>
__device__ __forceinline__ void dit_stage_inv(float2* buf, int M, int lh, const float2* hi,
                                              const float2* lo) {
  const int h = 1 << lh;
  const int nb = M / R;
  const int tstep = (kCMax / R) >> lh;
  for (int beta = threadIdx.x; beta < nb; beta += kCT) {
    const int b = beta >> lh, j = beta & (h - 1);
    const int base = b * R * h + j;
    float2 x[R];
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = buf[cpad(base + k * h)];
    const int e1 = tstep * j;
#pragma unroll
    for (int k = 1; k < R; ++k) {
      const float2 w = twc(hi, lo, (e1 * k) & (kCMax - 1));
      x[k] = cmul(x[k], make_float2(w.x, -w.y));
    }
    // inverse DFT = conj(DFT(conj(x)))
#pragma unroll
    for (int k = 0; k < R; ++k) x[k].y = -x[k].y;
    dftR<R>(x);
#pragma unroll
    for (int k = 0; k < R; ++k) buf[cpad(base + k * h)] = make_float2(x[k].x, -x[k].y);
  }
}

__device__ __forceinline__ void forward_fft(float2* buf, int M, const float2* hi, const float2* lo) {
  const int lm = 31 - __builtin_clz(M);
  dif_stage<8>(buf, M, lm - 3, hi, lo);
  __syncthreads();
  dif_stage<8>(buf, M, lm - 6, hi, lo);
  __syncthreads();
  dif_stage<8>(buf, M, lm - 9, hi, lo);
  __syncthreads();
  if (M == 4096) dif_stage<8>(buf, M, 0, hi, lo);
  else if (M == 2048) dif_stage<4>(buf, M, 0, hi, lo);
  else dif_stage<2>(buf, M, 0, hi, lo);
  __syncthreads();
}

__device__ __forceinline__ void inverse_fft(float2* buf, int M, const float2* hi, const float2* lo) {
  const int lm = 31 - __builtin_clz(M);
  if (M == 4096) dit_stage_inv<8>(buf, M, 0, hi, lo);
  else if (M == 2048) dit_stage_inv<4>(buf, M, 0, hi, lo);
  else dit_stage_inv<2>(buf, M, 0, hi, lo);
  __syncthreads();
  dit_stage_inv<8>(buf, M, lm - 9, hi, lo);
  __syncthreads();
  dit_stage_inv<8>(buf, M, lm - 6, hi, lo);
  __syncthreads();
  dit_stage_inv<8>(buf, M, lm - 3, hi, lo);
  __syncthreads();
}

// a[n] (see integrand_tf2) from the natural-order tables.
template <int ND>
__device__ __forceinline__ float2 integrand_natural(const EvDev& ev, const RowW& rw, int n) {
  const double f = (double)n * ev.df;
  const double* tb = ev.tf2n + 3 * (size_t)n;
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
    const double cyc = fma(f, rw.dt[d], s);
    const float2 e = expm2pi((float)(cyc - floor(cyc)));
    const float2 dw = ev.dwan[(size_t)d * ev.N + n];
    const float hr = rw.cre[d] * e.x - rw.cim[d] * e.y;
    const float hi = rw.cre[d] * e.y + rw.cim[d] * e.x;
    ar += hr * dw.x + hi * dw.y;
    ai += hi * dw.x - hr * dw.y;
  }
  return make_float2(ar, ai);
}

struct ChirpArgs {
  EvDev ev;
  const RowW* rows;
  long long n;
  double* out;
  int* fallback;  // [n]: 1 where the support needs a transform longer than 4096
};

template <int ND>
__global__ __launch_bounds__(kCT) void k_time_c(ChirpArgs a) {
  extern __shared__ double lds_d[];
  float2* buf = reinterpret_cast<float2*>(lds_d);            // [kCBuf]
  float2* thi = buf + kCBuf;                                  // [1024] W_4096^k
  float2* tlo = thi;
  double* red = reinterpret_cast<double*>(thi + 1024);        // [2 kCT / 64]
  double* knots = red + 2 * kCT / 64;                          // tx, ty (distance)
  const EvDev& ev = a.ev;
  const long long r = blockIdx.x;
  const int tid = threadIdx.x;
  const bool dist = (ev.flags & GWL_DISTANCE_MARG) != 0;
  if (dist) stage_knots(ev, knots);
  for (int k = tid; k < 1024; k += kCT) {
    double sn, cs;
    sincospi(-2.0 * (double)k / (double)kCMax, &sn, &cs);
    thi[k] = make_float2((float)cs, (float)sn);
  }
  const RowW& rw = a.rows[r];
  const int N = ev.N, L = ev.L, jlo = ev.jlo;
  const int kb = rw.kb;
  const int ke = min(rw.ke, N - 1);  // the transform leaves out the Nyquist bin
  const int B = (rw.valid && ke >= kb) ? ke - kb + 1 : 0;
  const int need = B + L - 1;
  const int M = need <= 1024 ? 1024 : need <= 2048 ? 2048 : need <= kCMax ? kCMax : 0;
  if (tid == 0) a.fallback[r] = (rw.valid && M == 0) ? 1 : 0;
  if (!rw.valid) {
    if (tid == 0) a.out[r] = row_valid_sentinel();
    return;
  }
  if (M == 0) return;
  // exponents of W^(x/2) modulo 2N (a power of two): 32-bit unsigned
  // arithmetic wraps modulo 2^32, a multiple of 2N, so the masked value is exact
  const unsigned mask2n = 2u * (unsigned)N - 1u;
  const float inv2n = 1.0f / (2.0f * (float)N);
  // u[m] = a[kb + m] W^(m jlo + m^2/2), zero for m >= B
  for (int m = tid; m < M; m += kCT) {
    float2 u = make_float2(0.f, 0.f);
    if (m < B) {
      const float2 av = integrand_natural<ND>(ev, rw, kb + m);
      const unsigned um = (unsigned)m;
      const unsigned e2 = (2u * um * (unsigned)jlo + um * um) & mask2n;
      u = cmul(av, expm2pi((float)e2 * inv2n));
    }
    buf[cpad(m)] = u;
  }
  __syncthreads();
  forward_fft(buf, M, thi, tlo);
  const float2* V = ev.chirpv[M == 1024 ? 0 : M == 2048 ? 1 : 2];
  for (int k = tid; k < M; k += kCT) buf[cpad(k)] = cmul(buf[cpad(k)], V[k]);
  __syncthreads();
  inverse_fft(buf, M, thi, tlo);
  // window bins and the epilogue: logsumexp over the prior window
  const double norm = 4.0 / ev.T;
  const double hh = rw.hh * norm, hs = rw.hscale, jit = rw.jitter;
  SplineY yy;
  if (dist) yy = spline_y(ev, knots + ev.sp.nx + 4, hh * hs * hs);
  double m = -INFINITY, s = 0.0;
  for (int t = tid; t < L; t += kCT) {
    const unsigned ut = (unsigned)t, ukb = (unsigned)kb;
    const unsigned e2 = (2u * ukb * (unsigned)jlo + 2u * ukb * ut + ut * ut) & mask2n;
    const float2 x = cmul(buf[cpad(t)], expm2pi((float)e2 * inv2n));
    const int j = jlo + t;
    const double tt = (ev.start + (double)(j + 1) * ev.tstep) + jit;
    if (tt >= ev.tmin && tt <= ev.tmax)
      lse_push(m, s, epi_value(ev, knots, yy, hh, hs, norm * (double)x.x, norm * (double)x.y));
  }
  lse_wave(m, s);
  const int lane = tid & 63, wid = tid >> 6;
  if (lane == 0) {
    red[2 * wid] = m;
    red[2 * wid + 1] = s;
  }
  __syncthreads();
  if (tid == 0) {
    double mm = -INFINITY, ss = 0.0;
    for (int w = 0; w < kCT / 64; ++w) {
      const double mo = red[2 * w], so = red[2 * w + 1];
      const double mn = fmax(mm, mo);
      if (mn > -INFINITY)
        ss = (mm > -INFINITY ? ss * exp(mm - mn) : 0.0) + (mo > -INFINITY ? so * exp(mo - mn) : 0.0);
      mm = mn;
    }
    const double b = (1.0 / (ev.tmax - ev.tmin)) * ev.tdelta;
    a.out[r] = (mm > -INFINITY) ? log(b * ss) + mm : -INFINITY;
  }
}

// natural-order f^1/3, f^-1/3, ln f and dw f^-7/6
__global__ __launch_bounds__(256) void k_chirp_tables(EvDev ev, double* tb, float2* dwa) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= ev.N) return;
  double u = 0.0, v = 0.0, lf = 0.0, amp = 0.0;
  if (n > 0) {
    double b[kNB];
    tf2_basis((double)n * ev.df, b);
    u = b[5];
    v = b[3];
    lf = b[7];
    amp = b[1] * sqrt(b[3]);
  }
  tb[3 * (size_t)n + 0] = u;
  tb[3 * (size_t)n + 1] = v;
  tb[3 * (size_t)n + 2] = lf;
  const int i = (n % ev.Q) * ev.P + n / ev.Q;  // decimated position of bin n
  for (int d = 0; d < ev.nd; ++d) {
    const double2 w = ev.dw64[(size_t)d * ev.N + i];
    dwa[(size_t)d * ev.N + n] = make_float2((float)(w.x * amp), (float)(w.y * amp));
  }
}

// ------------------------------------------------------------------ host
// The device forward transform in float64: the same stages, legs and
// twiddles, so the output is in the same digit-reversed order.
static void dif_host(std::vector<std::complex<double>>& x, int M) {
  const double two_pi = 6.283185307179586476925286766559;
  auto stage = [&](int R, int h) {
    for (int beta = 0; beta < M / R; ++beta) {
      const int b = beta / h, j = beta % h;
      const int base = b * R * h + j;
      std::complex<double> in[8], out[8];
      for (int k = 0; k < R; ++k) in[k] = x[base + k * h];
      for (int k = 0; k < R; ++k) {
        std::complex<double> acc = 0.0;
        for (int q = 0; q < R; ++q) {
          const int e = (q * k) % R;
          acc += in[q] * std::polar(1.0, -two_pi * e / R);
        }
        const long long e = ((long long)j * k) % ((long long)R * h);
        out[k] = acc * std::polar(1.0, -two_pi * (double)e / (double)(R * h));
      }
      for (int k = 0; k < R; ++k) x[base + k * h] = out[k];
    }
  };
  stage(8, M / 8);
  stage(8, M / 64);
  stage(8, M / 512);
  stage(M / 512, 1);
}

void chirp_spectrum(int N, int L, int M, float2* out) {
  const double two_pi = 6.283185307179586476925286766559;
  std::vector<std::complex<double>> v(M);
  const long long two_n = 2LL * N;
  for (int i = 0; i < M; ++i) {
    const long long k = i < L ? i : i - M;
    const long long e = (k * k) % two_n;
    v[i] = std::polar(1.0, two_pi * (double)e / (double)two_n);  // W^(-k^2/2)
  }
  dif_host(v, M);
  for (int i = 0; i < M; ++i)
    out[i] = make_float2((float)(v[i].real() / M), (float)(v[i].imag() / M));
}

int init_chirp_tables(const EvDev& ev, double* tf2n, float2* dwan) {
  hipLaunchKernelGGL(k_chirp_tables, dim3((unsigned)((ev.N + 255) / 256)), dim3(256), 0, 0, ev, tf2n,
                     dwan);
  GWL_HIP_CHECK(hipGetLastError());
  GWL_HIP_CHECK(hipDeviceSynchronize());
  return GWL_OK;
}

size_t time_chirp_lds_bytes(const EvDev& ev) {
  size_t b = (size_t)(kCBuf + 1024) * sizeof(float2) + (size_t)(2 * kCT / 64) * sizeof(double);
  if (ev.flags & GWL_DISTANCE_MARG) b += (size_t)(ev.sp.nx + ev.sp.ny + 8) * sizeof(double);
  return b;
}

int launch_chirp_kernel(const EvDev& ev, const RowW* rows, long long n, double* out, int* fallback,
                        hipStream_t st) {
  ChirpArgs a{ev, rows, n, out, fallback};
  const dim3 grid((unsigned)n);
  const size_t lds = time_chirp_lds_bytes(ev);
  switch (ev.nd) {
    case 1: hipLaunchKernelGGL(k_time_c<1>, grid, dim3(kCT), lds, st, a); break;
    case 2: hipLaunchKernelGGL(k_time_c<2>, grid, dim3(kCT), lds, st, a); break;
    case 3: hipLaunchKernelGGL(k_time_c<3>, grid, dim3(kCT), lds, st, a); break;
    default: hipLaunchKernelGGL(k_time_c<4>, grid, dim3(kCT), lds, st, a); break;
  }
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

bool time_chirp_eligible(const gwl_event* e, long long n) {
  const EvDev& ev = e->dev;
  if (!(ev.flags & GWL_TIME_MARG) || ev.approx != GWL_TAYLORF2 || !ev.tf2n || ev.P != 512 ||
      (ev.N & (ev.N - 1)) != 0 || ev.N > (1 << 22))
    return false;
  (void)n;
  return e->transform == GWL_TRANSFORM_CHIRP;
}

}  // namespace gwl
