// C ABI of libgwl: context and event management, device residency of the
// event data, the distance-marginalisation table kernel, the host spline
// coefficient solve, and small operator entry points (batched LDS FFT,
// ln I0) used by the likelihood and exposed for checking.
#include "gwl_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace gwl {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
  return GWL_E_HIP;
}

static int ilog2(long long x) {
  int l = 0;
  while ((1LL << l) < x) ++l;
  return l;
}

// ------------------------------------------------------------------ distance table
// base.py:994-1018: table[ii][jj] = logsumexp_k(f(d_jj s_k) - h_ii s_k^2 / 2,
// b = prior_k * delta) - log_norm, s_k = d_ref / d_k, f = ln I0(|.|) with
// phase marginalisation, identity otherwise.  One workgroup per d-axis
// point; the per-distance terms of a chunk of k are staged in LDS and every
// thread (one h-axis point) streams them with a running-max logsumexp.
constexpr int kTabChunk = 2048;

__global__ __launch_bounds__(256) void k_distance_table(const double* dist, const double* prior,
                                                        int nk, double dref, int phase,
                                                        double delta, const double* daxis, int nd,
                                                        const double* haxis, int nh,
                                                        double lognorm, double* out) {
  __shared__ double dterm[kTabChunk];
  __shared__ double s2[kTabChunk];
  __shared__ double lb[kTabChunk];
  const int jj = blockIdx.x;
  const double x = daxis[jj];
  constexpr int HPT = 4;  // h-axis points per thread (nh <= 1024)
  double m[HPT], s[HPT], hv[HPT];
#pragma unroll
  for (int q = 0; q < HPT; ++q) {
    int ii = threadIdx.x + q * 256;
    hv[q] = ii < nh ? haxis[ii] : 0.0;
    m[q] = -INFINITY;
    s[q] = 0.0;
  }
  for (int k0 = 0; k0 < nk; k0 += kTabChunk) {
    const int kn = min(kTabChunk, nk - k0);
    __syncthreads();
    for (int k = threadIdx.x; k < kn; k += 256) {
      const double sc = dref / dist[k0 + k];
      const double dv = x * sc;
      dterm[k] = phase ? ln_i0(fabs(dv)) : dv;
      s2[k] = sc * sc;
      const double b = prior[k0 + k] * delta;
      lb[k] = b;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < HPT; ++q) {
      if (threadIdx.x + q * 256 >= nh) continue;
      double mm = m[q], ss = s[q];
      const double h = hv[q];
      for (int k = 0; k < kn; ++k) {
        const double b = lb[k];
        if (b == 0.0) continue;
        const double a = dterm[k] - h * s2[k] / 2.0;
        if (a > mm) {
          ss = ss * exp(mm - a) + b;
          mm = a;
        } else {
          ss += b * exp(a - mm);
        }
      }
      m[q] = mm;
      s[q] = ss;
    }
  }
#pragma unroll
  for (int q = 0; q < HPT; ++q) {
    int ii = threadIdx.x + q * 256;
    if (ii < nh) out[(size_t)ii * nd + jj] = (log(s[q]) + m[q]) - lognorm;
  }
}

// ------------------------------------------------------------------ batched FFT operator
template <typename R, int P>
struct FftTile {
  static constexpr int B = (P >= 4096 ? 1 : 4096 / P);
};

template <typename R, int P, int NT>
__global__ __launch_bounds__(NT) void k_fft_batch(const cplx<R>* in, cplx<R>* out, long long batch) {
  using C = cplx<R>;
  constexpr int B = FftTile<R, P>::B;
  constexpr int MAXBF8 = ((B * P) / 8 + NT - 1) / NT > 0 ? ((B * P) / 8 + NT - 1) / NT : 1;
  extern __shared__ __align__(16) unsigned char smem[];
  C* buf = reinterpret_cast<C*>(smem);
  C* tw = buf + B * P;
  init_twiddles<P, NT>(tw);
  const long long b0 = (long long)blockIdx.x * B;
  const int nb = (int)min((long long)B, batch - b0);
  for (int i = threadIdx.x; i < B * P; i += NT)
    buf[i] = (i < nb * P) ? in[b0 * P + i] : cmk<C>((R)0, (R)0);
  __syncthreads();
  fft_lds<NT, MAXBF8, P>(buf, tw, B);
  for (int i = threadIdx.x; i < nb * P; i += NT) out[b0 * P + i] = buf[i];
}

template <typename R, int P>
static int fft_launch(const gwl_ctx* ctx, const void* in, void* out, long long batch, hipStream_t st) {
  constexpr int NT = 256;
  constexpr int B = FftTile<R, P>::B;
  const size_t lds = (size_t)(B * P + TwLayout<P>::kEntries) * sizeof(cplx<R>);
  if (lds > ctx->lds_per_block) {
    set_error("gwl_fft_batch: transform does not fit the LDS of a workgroup");
    return GWL_E_UNSUPPORTED;
  }
  long long nblk = (batch + B - 1) / B;
  hipLaunchKernelGGL((k_fft_batch<R, P, NT>), dim3((unsigned)nblk), dim3(NT), lds, st,
                     reinterpret_cast<const cplx<R>*>(in), reinterpret_cast<cplx<R>*>(out), batch);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

// Transforms longer than one LDS transform: P = Rr P2 (four-step).  The
// radix-Rr step y[b][k1][n2] = W_P^(n2 k1) sum_n1 x[b][n2 + P2 n1] W_Rr^(n1 k1)
// (float64 twiddles from the exact integer phase), P2-point LDS transforms
// of the Rr sub-sequences, then X[b][k1 + Rr k2] = Y[b][k1][k2].
template <typename R>
__global__ __launch_bounds__(256) void k_fft_split_pre(const cplx<R>* in, cplx<R>* tmp,
                                                       long long batch, int P, int Rr) {
  using C = cplx<R>;
  const int P2 = P / Rr;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;  // (b, n2)
  if (i >= batch * P2) return;
  const long long b = i / P2;
  const int n2 = (int)(i - b * P2);
  const C* x = in + b * P;
  for (int k1 = 0; k1 < Rr; ++k1) {
    double ar = 0.0, ai = 0.0;
    for (int n1 = 0; n1 < Rr; ++n1) {
      double sn, cs;
      sincospi(-2.0 * (double)((n1 * k1) % Rr) / (double)Rr, &sn, &cs);
      const C v = x[n2 + (long long)P2 * n1];
      ar += (double)v.x * cs - (double)v.y * sn;
      ai += (double)v.x * sn + (double)v.y * cs;
    }
    double sn, cs;
    sincospi(-2.0 * (double)(((long long)n2 * k1) % P) / (double)P, &sn, &cs);
    tmp[(b * Rr + k1) * P2 + n2] = cmk<C>((R)(ar * cs - ai * sn), (R)(ar * sn + ai * cs));
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_fft_split_post(const cplx<R>* tmp, cplx<R>* out,
                                                        long long batch, int P, int Rr) {
  const int P2 = P / Rr;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;  // output index
  if (i >= batch * P) return;
  const long long b = i / P;
  const int k = (int)(i - b * P);
  const int k1 = k % Rr, k2 = k / Rr;
  out[i] = tmp[(b * Rr + k1) * P2 + k2];
}

template <typename R>
static int fft_dispatch(const gwl_ctx* ctx, const void* in, void* out, long long batch, int P,
                        hipStream_t st);

// the longest power-of-two transform whose workgroup fits the LDS
template <typename R>
static int fft_lds_max(const gwl_ctx* ctx) {
  int P = 16384;
  while (P > 2 && (size_t)(P + P / 64 + 64) * sizeof(cplx<R>) > ctx->lds_per_block) P /= 2;
  return P;
}

template <typename R>
static int fft_split(const gwl_ctx* ctx, const void* in, void* out, long long batch, int P,
                     hipStream_t st) {
  const int P2 = fft_lds_max<R>(ctx);
  const int Rr = P / P2;
  if (P % P2 || Rr < 2 || Rr > 16) {
    set_error("gwl_fft_batch: unsupported length");
    return GWL_E_UNSUPPORTED;
  }
  void* tmp = nullptr;
  const size_t bytes = (size_t)batch * P * sizeof(cplx<R>);
  GWL_HIP_CHECK(hipMallocAsync(&tmp, bytes, st));
  const unsigned g1 = (unsigned)((batch * P2 + 255) / 256), g2 = (unsigned)((batch * P + 255) / 256);
  hipLaunchKernelGGL(k_fft_split_pre<R>, dim3(g1), dim3(256), 0, st,
                     reinterpret_cast<const cplx<R>*>(in), reinterpret_cast<cplx<R>*>(tmp), batch,
                     P, Rr);
  hipError_t e = hipGetLastError();
  int rc = e == hipSuccess ? GWL_OK : hip_fail(e, "k_fft_split_pre");
  if (rc == GWL_OK) rc = fft_dispatch<R>(ctx, tmp, tmp, batch * Rr, P2, st);  // in place
  if (rc == GWL_OK) {
    hipLaunchKernelGGL(k_fft_split_post<R>, dim3(g2), dim3(256), 0, st,
                       reinterpret_cast<const cplx<R>*>(tmp), reinterpret_cast<cplx<R>*>(out),
                       batch, P, Rr);
    e = hipGetLastError();
    rc = e == hipSuccess ? GWL_OK : hip_fail(e, "k_fft_split_post");
  }
  (void)hipFreeAsync(tmp, st);
  return rc;
}

template <typename R>
static int fft_dispatch(const gwl_ctx* ctx, const void* in, void* out, long long batch, int P,
                        hipStream_t st) {
  if (P > fft_lds_max<R>(ctx)) return fft_split<R>(ctx, in, out, batch, P, st);
  switch (P) {
    case 2: return fft_launch<R, 2>(ctx, in, out, batch, st);
    case 4: return fft_launch<R, 4>(ctx, in, out, batch, st);
    case 8: return fft_launch<R, 8>(ctx, in, out, batch, st);
    case 16: return fft_launch<R, 16>(ctx, in, out, batch, st);
    case 32: return fft_launch<R, 32>(ctx, in, out, batch, st);
    case 64: return fft_launch<R, 64>(ctx, in, out, batch, st);
    case 128: return fft_launch<R, 128>(ctx, in, out, batch, st);
    case 256: return fft_launch<R, 256>(ctx, in, out, batch, st);
    case 512: return fft_launch<R, 512>(ctx, in, out, batch, st);
    case 1024: return fft_launch<R, 1024>(ctx, in, out, batch, st);
    case 2048: return fft_launch<R, 2048>(ctx, in, out, batch, st);
    case 4096: return fft_launch<R, 4096>(ctx, in, out, batch, st);
    case 8192: return fft_launch<R, 8192>(ctx, in, out, batch, st);
    case 16384: return fft_launch<R, 16384>(ctx, in, out, batch, st);
  }
  set_error("gwl_fft_batch: unsupported length");
  return GWL_E_UNSUPPORTED;
}

// -sum_det |4/T sum |d|^2/S| / 2 over the masked bins (base.py:402-413);
// |d|^2/S = |dw|^2 / w with the scaled arrays dw = d sigma / S, w = sigma^2 / S.
__global__ __launch_bounds__(1024) void k_noise_lnl(EvDev ev, double* out) {
  __shared__ double red[16];
  double tot = 0.0;
  for (int d = 0; d < ev.nd; ++d) {
    double s = 0.0;
    for (int i = threadIdx.x; i < ev.N; i += 1024) {
      double2 v = ev.dw64[(size_t)d * ev.N + i];
      double w = ev.w64[(size_t)d * ev.N + i];
      if (w > 0.0) s += (v.x * v.x + v.y * v.y) / w;
    }
    if (threadIdx.x == 0 && ev.wN[d] > 0.0)
      s += (ev.dwN[d][0] * ev.dwN[d][0] + ev.dwN[d][1] * ev.dwN[d][1]) / ev.wN[d];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double v = 0.0;
      for (int w = 0; w < 16; ++w) v += red[w];
      tot -= fabs(4.0 / ev.T * v / 2.0);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = tot;
}

__global__ void k_ln_i0(const double* in, double* out, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = ln_i0(in[i]);
}

// ------------------------------------------------------------------ host spline solve
// Not-a-knot cubic knots on the data abscissae (FITPACK's choice for s = 0):
// t = {x0 x4, x2 .. x_{n-3}, x_{n-1} x4}.
static void knots(const double* x, int n, double* t) {
  for (int i = 0; i < 4; ++i) t[i] = x[0];
  for (int i = 2; i <= n - 3; ++i) t[i + 2] = x[i];
  for (int i = 0; i < 4; ++i) t[n + i] = x[n - 1];
}

// LU with partial pivoting of the n x n collocation matrix A[i][m] = B_m(x_i).
struct LU {
  int n;
  std::vector<double> a;
  std::vector<int> piv;
  bool factor(const double* x, const double* t) {
    a.assign((size_t)n * n, 0.0);
    piv.resize(n);
    for (int i = 0; i < n; ++i) {
      int l = knot_interval(t, n, x[i]);
      double h[4];
      bspline3(t, l, x[i], h);
      for (int k = 0; k < 4; ++k) a[(size_t)i * n + (l - 3 + k)] = h[k];
    }
    for (int k = 0; k < n; ++k) {
      int p = k;
      double best = std::fabs(a[(size_t)k * n + k]);
      for (int i = k + 1; i < std::min(n, k + 8); ++i) {
        double v = std::fabs(a[(size_t)i * n + k]);
        if (v > best) { best = v; p = i; }
      }
      if (best == 0.0) return false;
      piv[k] = p;
      // interchange the not-yet-eliminated part only, so that the multipliers
      // stay where the interleaved forward substitution in solve() reads them
      if (p != k)
        for (int j = k; j < std::min(n, k + 12); ++j) std::swap(a[(size_t)k * n + j], a[(size_t)p * n + j]);
      const double inv = 1.0 / a[(size_t)k * n + k];
      for (int i = k + 1; i < std::min(n, k + 8); ++i) {
        double f = a[(size_t)i * n + k] * inv;
        if (f == 0.0) continue;
        a[(size_t)i * n + k] = f;
        for (int j = k + 1; j < std::min(n, k + 12); ++j) a[(size_t)i * n + j] -= f * a[(size_t)k * n + j];
      }
    }
    return true;
  }
  // solve A y = b in place (b has stride 'stride' between consecutive entries)
  void solve(double* b, size_t stride) const {
    for (int k = 0; k < n; ++k) {
      int p = piv[k];
      if (p != k) std::swap(b[k * stride], b[p * stride]);
      for (int i = k + 1; i < std::min(n, k + 8); ++i) b[i * stride] -= a[(size_t)i * n + k] * b[k * stride];
    }
    for (int k = n - 1; k >= 0; --k) {
      double s = b[k * stride];
      for (int j = k + 1; j < std::min(n, k + 12); ++j) s -= a[(size_t)k * n + j] * b[j * stride];
      b[k * stride] = s / a[(size_t)k * n + k];
    }
  }
};

// Per detector (one workgroup each): the sums over bins of conj(h F) d / S
// and |h F|^2 / S, h the detector response (physical units, [nd][nf][2]), F a
// per-bin complex factor ([nd][nf][2]), d / S and 1 / S from the event's
// float64 data (sigma-scaled, zero off the mask; the Nyquist bin from the
// argument block).  part[3 d + {0, 1, 2}] = Re, Im, |.|^2 sums (without 4/T).
__global__ __launch_bounds__(256) void k_weighted_inner(EvDev ev, const double* resp,
                                                         const double* fac, double* part) {
  __shared__ double red[4][3];
  const int d = blockIdx.x;
  double sr = 0.0, si = 0.0, sh = 0.0;
  for (int n = threadIdx.x; n < ev.nf; n += blockDim.x) {
    double dr, di, w;
    if (n == ev.N) {
      dr = ev.dwN[d][0];
      di = ev.dwN[d][1];
      w = ev.wN[d];
    } else {
      const size_t i = (size_t)d * ev.N + (size_t)(n % ev.Q) * ev.P + n / ev.Q;
      dr = ev.dw64[i].x;
      di = ev.dw64[i].y;
      w = ev.w64[i];
    }
    if (w == 0.0 && dr == 0.0 && di == 0.0) continue;
    const size_t o = ((size_t)d * ev.nf + n) * 2;
    const double hr = resp[o] * fac[o] - resp[o + 1] * fac[o + 1];
    const double hi = resp[o] * fac[o + 1] + resp[o + 1] * fac[o];
    // conj(h F) (d / S sigma) / sigma
    sr += (hr * dr + hi * di) / ev.sigma;
    si += (hr * di - hi * dr) / ev.sigma;
    sh += (hr * hr + hi * hi) * (w / (ev.sigma * ev.sigma));
  }
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    si += __shfl_xor(si, off, 64);
    sh += __shfl_xor(sh, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6][0] = sr;
    red[threadIdx.x >> 6][1] = si;
    red[threadIdx.x >> 6][2] = sh;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double v = 0.0;
    for (int w = 0; w < 4; ++w) v += red[w][threadIdx.x];
    part[3 * d + threadIdx.x] = v;
  }
}

// One thread per window bin j = j_first + jj (j may pass n_fft: the
// exponential wraps): sum over detectors and masked bins n of
// h_n conj(d_n) / S_n exp(-2 pi i n j / n_fft) (without 4/T), the transform
// of the zero-padded integrand of length n_fft; h the response
// ([nd][nf][2]), d / S sigma from the event's float64 data.  Every lane of a
// wave walks the same bins, so the loads are broadcasts.
__global__ __launch_bounds__(256) void k_window_dft(EvDev ev, const double* resp,
                                                     long long n_fft, long long j_first, int L,
                                                     double* out) {
  const int jj = blockIdx.x * blockDim.x + threadIdx.x;
  if (jj >= L) return;
  const unsigned long long j = (unsigned long long)(j_first + jj);
  double sr = 0.0, si = 0.0;
  for (int d = 0; d < ev.nd; ++d) {
    for (int n = ev.mlo[d]; n <= ev.mhi[d] && n < ev.nf; ++n) {
      double dr, di;
      if (n == ev.N) {
        dr = ev.dwN[d][0];
        di = ev.dwN[d][1];
      } else {
        const size_t i = (size_t)d * ev.N + (size_t)(n % ev.Q) * ev.P + n / ev.Q;
        dr = ev.dw64[i].x;
        di = ev.dw64[i].y;
      }
      const size_t o = ((size_t)d * ev.nf + n) * 2;
      const double hr = resp[o], hi = resp[o + 1];
      // h conj(d / S sigma) / sigma
      const double xr = (hr * dr + hi * di) / ev.sigma;
      const double xi = (hi * dr - hr * di) / ev.sigma;
      const unsigned long long r = ((unsigned long long)n * j) % (unsigned long long)n_fft;
      double sn, cs;
      sincospi(-2.0 * (double)r / (double)n_fft, &sn, &cs);
      sr += xr * cs - xi * sn;
      si += xr * sn + xi * cs;
    }
  }
  out[2 * jj] = sr;
  out[2 * jj + 1] = si;
}

}  // namespace gwl

using namespace gwl;

namespace gwl {
int upload_time_prior(int64_t n, double t0, double dt, const double* density, double** dst) {
  if (n < 2 || !density || !(dt > 0) || !std::isfinite(t0)) {
    set_error("time prior table: need n >= 2 points, a positive spacing and finite values");
    return GWL_E_INVALID;
  }
  for (int64_t k = 0; k < n; ++k) {
    if (!(density[k] >= 0) || !std::isfinite(density[k])) {
      set_error("time prior table: densities must be finite and non-negative");
      return GWL_E_INVALID;
    }
  }
  if (*dst) (void)hipFree(*dst);
  *dst = nullptr;
  GWL_HIP_CHECK(hipMalloc((void**)dst, (size_t)n * sizeof(double)));
  GWL_HIP_CHECK(hipMemcpy(*dst, density, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  return GWL_OK;
}
}  // namespace gwl

extern "C" {

const char* gwl_last_error(void) { return g_last_error.c_str(); }

int gwl_ctx_create(int32_t device, gwl_ctx** out) {
  if (!out) { set_error("gwl_ctx_create: out is NULL"); return GWL_E_INVALID; }
  int ndev = 0;
  GWL_HIP_CHECK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) { set_error("gwl_ctx_create: no such device"); return GWL_E_INVALID; }
  GWL_HIP_CHECK(hipSetDevice(device));
  hipDeviceProp_t prop;
  GWL_HIP_CHECK(hipGetDeviceProperties(&prop, device));
  gwl_ctx* c = new (std::nothrow) gwl_ctx;
  if (!c) { set_error("out of host memory"); return GWL_E_NOMEM; }
  c->device = device;
  c->n_cu = prop.multiProcessorCount;
  c->lds_per_block = prop.sharedMemPerBlock;
  if (int rc = noalloc_scratch_init(device, &c->scratch)) {
    delete c;
    return rc;
  }
  *out = c;
  return GWL_OK;
}

int gwl_ctx_destroy(gwl_ctx* ctx) {
  delete ctx;
  return GWL_OK;
}

int gwl_event_create(gwl_ctx* ctx, const gwl_event_desc* d, gwl_event** out, int64_t* hbm_bytes) {
  if (!ctx || !d || !out) { set_error("gwl_event_create: NULL argument"); return GWL_E_INVALID; }
  if (d->n_det < 1 || d->n_det > GWL_MAX_DET) { set_error("n_det must be 1..4"); return GWL_E_INVALID; }
  if (d->n_freq < 3) { set_error("n_freq too small"); return GWL_E_INVALID; }
  if (d->approximant != GWL_SINE_GAUSSIAN && d->approximant != GWL_TAYLORF2 &&
      d->approximant != GWL_HOST_POLARIZATIONS) {
    set_error("unknown approximant");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(ctx->device));
  const int nd = d->n_det, nf = d->n_freq, N = nf - 1;
  EvDev ev{};
  ev.nd = nd;
  ev.nf = nf;
  ev.N = N;
  ev.T = d->duration;
  ev.fs = d->sampling_frequency;
  ev.df = (d->sampling_frequency / 2.0) / (double)(nf - 1);  // linspace step, series.py:134
  ev.start = d->start_time;
  ev.approx = d->approximant;
  ev.wf_fmin = d->wf_minimum_frequency;
  ev.wf_fmax = d->wf_maximum_frequency > 0 ? d->wf_maximum_frequency : ev.df * (double)N;
  ev.fref = d->reference_frequency > 0 ? d->reference_frequency : 50.0;
  ev.flags = d->flags;
  ev.tmin = d->time_prior_minimum;
  ev.tmax = d->time_prior_maximum;
  ev.dref = d->reference_distance > 0 ? d->reference_distance : 1.0;
  for (int k = 0; k < nd; ++k) {
    ev.mlo[k] = std::max(0, d->mask_lo[k]);
    ev.mhi[k] = std::min(N, d->mask_hi[k]);
    if (ev.mlo[k] > ev.mhi[k]) { set_error("empty frequency mask"); return GWL_E_INVALID; }
    for (int i = 0; i < 9; ++i) ev.tensor[k][i] = d->detector_tensor[9 * k + i];
    for (int i = 0; i < 3; ++i) ev.vertex[k][i] = d->vertex[3 * k + i];
  }
  if ((ev.flags & GWL_DISTANCE_MARG) && !(d->reference_distance > 0)) {
    set_error("distance marginalisation needs reference_distance > 0");
    return GWL_E_INVALID;
  }
  // strain scale: sigma^2 = median PSD over detector 0's mask
  {
    std::vector<double> v;
    for (int n = ev.mlo[0]; n <= ev.mhi[0]; ++n) {
      double s = d->psd[n];
      if (std::isfinite(s) && s > 0) v.push_back(s);
    }
    if (v.empty()) { set_error("PSD not finite and positive inside the mask"); return GWL_E_INVALID; }
    std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
    ev.sigma = std::sqrt(v[v.size() / 2]);
  }
  // time marginalisation window and transform tile
  ev.tstep = d->duration / (double)N;
  ev.tdelta = 2.0 / d->sampling_frequency;
  if (ev.flags & GWL_TIME_MARG) {
    if (!(ev.tmax > ev.tmin)) { set_error("time prior must have maximum > minimum"); return GWL_E_INVALID; }
    long long jlo = (long long)std::floor((ev.tmin - ev.start - ev.tdelta) / ev.tstep) - 3;
    long long jhi = (long long)std::ceil((ev.tmax - ev.start + ev.tdelta) / ev.tstep) + 2;
    jlo = std::max(0LL, jlo);
    jhi = std::min((long long)N - 1, jhi);
    if (jhi < jlo) { set_error("time prior outside the segment"); return GWL_E_INVALID; }
    ev.jlo = (int)jlo;
    ev.L = (int)(jhi - jlo + 1);
    // transform length: the power of two covering the window, at most 8192
    // (float64 tiles of one row fit a workgroup's LDS up to that length);
    // a longer window is evaluated in slices of P bins, one transform pass
    // each (bilby's default prior, the whole segment, on long segments)
    long long P = 512;
    while (P < ev.L && P < 8192) P *= 2;
    if (P > N) P = N;
    if ((N % P) != 0 || (P & (P - 1)) != 0) {
      set_error("time marginalisation needs N = n_freq - 1 divisible by a power of two");
      return GWL_E_UNSUPPORTED;
    }
    ev.P = (int)P;
    ev.Q = (int)(N / P);
    ev.logP = ilog2(P);
  } else {
    ev.logP = ilog2(N);
    ev.P = 1 << ev.logP;
    ev.Q = 1;
    ev.jlo = 0;
    ev.L = 0;
  }
  // host staging in decimated order
  const size_t n_el = (size_t)nd * N;
  std::vector<float2> dw32(n_el);
  std::vector<float> w32(n_el);
  std::vector<double2> dw64(n_el);
  std::vector<double> w64(n_el);
  const double sg = ev.sigma, sg2 = ev.sigma * ev.sigma;
  for (int k = 0; k < nd; ++k) {
    for (int i = 0; i < N; ++i) {
      const int n = (ev.Q == 1) ? i : ev.Q * (i & (ev.P - 1)) + (i >> ev.logP);
      double re = 0, im = 0, w = 0;
      if (n >= ev.mlo[k] && n <= ev.mhi[k]) {
        const double s = d->psd[(size_t)k * nf + n];
        const double* h = d->strain + 2 * ((size_t)k * nf + n);
        re = h[0] / s * sg;
        im = h[1] / s * sg;
        w = sg2 / s;
        if (!std::isfinite(re) || !std::isfinite(im) || !std::isfinite(w)) {
          set_error("non-finite data or PSD inside the mask");
          return GWL_E_INVALID;
        }
      }
      const size_t o = (size_t)k * N + i;
      dw32[o] = make_float2((float)re, (float)im);
      w32[o] = (float)w;
      dw64[o] = make_double2(re, im);
      w64[o] = w;
    }
    const int n = N;
    if (n >= ev.mlo[k] && n <= ev.mhi[k]) {
      const double s = d->psd[(size_t)k * nf + n];
      const double* h = d->strain + 2 * ((size_t)k * nf + n);
      ev.dwN[k][0] = h[0] / s * sg;
      ev.dwN[k][1] = h[1] / s * sg;
      ev.wN[k] = sg2 / s;
    }
  }
  // TaylorF2 with a 512-point time transform: tables of the wave-per-row
  // kernel, including prefix sums of f^-7/3 sigma^2 / S over bins 0..N
  const bool wave_tables = ev.approx == GWL_TAYLORF2 && (ev.flags & GWL_TIME_MARG) && ev.P == 512;
  std::vector<double> hps;
  if (wave_tables) {
    hps.assign((size_t)nd * (N + 2), 0.0);
    for (int k = 0; k < nd; ++k) {
      double* ps = hps.data() + (size_t)k * (N + 2);
      for (int n = 0; n <= N; ++n) {
        double term = 0.0;
        if (n > 0 && n >= ev.mlo[k] && n <= ev.mhi[k]) {
          const double f = (double)n * ev.df;
          const double w = (n == N) ? ev.wN[k]
                                    : w64[(size_t)k * N + (size_t)(n % ev.Q) * ev.P + n / ev.Q];
          const double a = 1.0 / std::cbrt(f);  // f^-1/3
          const double a7 = a * a * a * a * a * a * a;
          term = a7 * w;
        }
        ps[n + 1] = ps[n] + term;
      }
    }
  }
  gwl_event* e = new (std::nothrow) gwl_event;
  if (!e) { set_error("out of host memory"); return GWL_E_NOMEM; }
  e->ctx = ctx;
  const size_t b32 = n_el * (sizeof(float2) + sizeof(float));
  const size_t b64 = n_el * (sizeof(double2) + sizeof(double));
  const size_t bw = wave_tables ? ((size_t)N * 3 * sizeof(double) + n_el * sizeof(float2)) +
                                      hps.size() * sizeof(double) + 4096
                                : 0;
  e->data_bytes = b32 + b64 + bw + 1024;
  if (hipMalloc(&e->data, e->data_bytes) != hipSuccess) {
    delete e;
    set_error("hipMalloc failed for event data");
    return GWL_E_NOMEM;
  }
  char* p = reinterpret_cast<char*>(e->data);
  ev.dw64 = reinterpret_cast<double2*>(p);
  p += n_el * sizeof(double2);
  ev.w64 = reinterpret_cast<double*>(p);
  p += n_el * sizeof(double);
  ev.dw32 = reinterpret_cast<float2*>(p);
  p += n_el * sizeof(float2);
  ev.w32 = reinterpret_cast<float*>(p);
  p += n_el * sizeof(float);
  if (wave_tables) {
    p = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(p) + 255) & ~(uintptr_t)255);
    ev.tf2b = reinterpret_cast<double*>(p);
    p += (size_t)N * 3 * sizeof(double);
    ev.dwa32 = reinterpret_cast<float2*>(p);
    p += n_el * sizeof(float2);
    ev.hps = reinterpret_cast<double*>(p);
    p += hps.size() * sizeof(double);
  }
  hipError_t err = hipSuccess;
  if (err == hipSuccess) err = hipMemcpy((void*)ev.dw64, dw64.data(), n_el * sizeof(double2), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy((void*)ev.w64, w64.data(), n_el * sizeof(double), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy((void*)ev.dw32, dw32.data(), n_el * sizeof(float2), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy((void*)ev.w32, w32.data(), n_el * sizeof(float), hipMemcpyHostToDevice);
  if (err == hipSuccess && wave_tables)
    err = hipMemcpy((void*)ev.hps, hps.data(), hps.size() * sizeof(double), hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    (void)hipFree(e->data);
    delete e;
    return hip_fail(err, "gwl_event_create upload");
  }
  if (wave_tables) {
    int rc = init_time_wave_tables(ev, const_cast<double*>(ev.tf2b), const_cast<float2*>(ev.dwa32));
    if (rc != GWL_OK) {
      (void)hipFree(e->data);
      delete e;
      return rc;
    }
  }
  e->dev = ev;
  *out = e;
  if (hbm_bytes) *hbm_bytes = (int64_t)e->data_bytes;
  return GWL_OK;
}

int gwl_event_set_tile(gwl_event* e, int32_t rows, int32_t wg_per_cu) {
  if (!e || rows < 0 || rows > 16 || wg_per_cu < 0 || wg_per_cu > 64) {
    set_error("gwl_event_set_tile: invalid argument");
    return GWL_E_INVALID;
  }
  e->tile_rows = rows;
  e->tile_wg_per_cu = wg_per_cu;
  return GWL_OK;
}

int gwl_event_set_escalation(gwl_event* e, double coeff, double tol) {
  if (!e || !(coeff >= 0.0) || !(tol > 0.0)) {
    set_error("gwl_event_set_escalation: invalid argument");
    return GWL_E_INVALID;
  }
  e->esc_coeff = coeff;
  e->esc_tol = tol;
  return GWL_OK;
}

int gwl_event_escalated_rows(gwl_event* e, int64_t* rows) {
  if (!e || !rows) {
    set_error("gwl_event_escalated_rows: invalid argument");
    return GWL_E_INVALID;
  }
  *rows = 0;
  if (!e->esc_count_dev) return GWL_OK;
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  GWL_HIP_CHECK(hipDeviceSynchronize());
  int c = 0;
  GWL_HIP_CHECK(hipMemcpy(&c, e->esc_count_dev, sizeof(int), hipMemcpyDeviceToHost));
  *rows = c;
  return GWL_OK;
}

int gwl_event_set_polarizations(gwl_event* e, const double* polarizations, int64_t n_rows) {
  if (!e || n_rows < 0 || (n_rows > 0 && !polarizations)) {
    set_error("gwl_event_set_polarizations: invalid argument");
    return GWL_E_INVALID;
  }
  if (e->dev.approx != GWL_HOST_POLARIZATIONS) {
    set_error("gwl_event_set_polarizations: the event's source model is evaluated by the "
              "library (create it with GWL_HOST_POLARIZATIONS)");
    return GWL_E_INVALID;
  }
  if (n_rows > 0x7fffffffLL) {
    set_error("gwl_event_set_polarizations: more than 2^31 - 1 rows");
    return GWL_E_INVALID;
  }
  e->dev.pol = reinterpret_cast<const double2*>(polarizations);
  e->pol_rows = n_rows;
  return GWL_OK;
}

int gwl_event_set_multi_steps(gwl_event* e, int32_t steps) {
  if (!e || steps < 0 || steps > 4096) {
    set_error("gwl_event_set_multi_steps: invalid argument");
    return GWL_E_INVALID;
  }
  e->multi_steps = steps;
  return GWL_OK;
}

int gwl_event_set_transform(gwl_event* e, int32_t transform) {
  if (!e || transform < GWL_TRANSFORM_AUTO || transform > GWL_TRANSFORM_WAVE) {
    set_error("gwl_event_set_transform: invalid argument");
    return GWL_E_INVALID;
  }
  e->transform = transform;
  return GWL_OK;
}

int gwl_event_timing(gwl_event* e, int32_t enable) {
  if (!e) { set_error("gwl_event_timing: NULL event"); return GWL_E_INVALID; }
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  if (enable && !e->t0) {
    GWL_HIP_CHECK(hipEventCreate(&e->t0));
    GWL_HIP_CHECK(hipEventCreate(&e->t1));
  }
  e->timing = enable != 0;
  return GWL_OK;
}

int gwl_event_timing_read(gwl_event* e, double* total_ms, int64_t* launches) {
  if (!e || !total_ms || !launches) { set_error("gwl_event_timing_read: invalid argument"); return GWL_E_INVALID; }
  *total_ms = e->timed_ms;
  *launches = e->timed_launches;
  e->timed_ms = 0.0;
  e->timed_launches = 0;
  return GWL_OK;
}

int gwl_event_destroy(gwl_event* e) {
  if (!e) return GWL_OK;
  (void)hipSetDevice(e->ctx->device);
  if (e->t0) (void)hipEventDestroy(e->t0);
  if (e->t1) (void)hipEventDestroy(e->t1);
  if (e->data) (void)hipFree(e->data);
  if (e->spline) (void)hipFree(e->spline);
  if (e->tprior) (void)hipFree(e->tprior);
  if (e->ws) (void)hipFree(e->ws);
  delete e;
  return GWL_OK;
}


int gwl_event_set_time_prior(gwl_event* e, int64_t n, double t0, double dt,
                             const double* density) {
  if (!e || !(e->dev.flags & GWL_TIME_MARG)) {
    set_error("gwl_event_set_time_prior: the event does not marginalise over time");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  if (int rc = upload_time_prior(n, t0, dt, density, &e->tprior)) return rc;
  e->dev.tpd = e->tprior;
  e->dev.tp_t0 = t0;
  e->dev.tp_ih = 1.0 / dt;
  e->dev.tp_n = (int)n;
  return GWL_OK;
}

int gwl_event_set_distance_spline(gwl_event* e, int32_t nx, int32_t ny, const double* tx,
                                  const double* ty, const double* coeffs, const double* bbox) {
  if (!e || !tx || !ty || !coeffs || !bbox || nx < 4 || ny < 4) {
    set_error("gwl_event_set_distance_spline: invalid argument");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  const size_t n = (size_t)(nx + 4) + (ny + 4) + (size_t)nx * ny;
  if (e->spline) {
    (void)hipFree(e->spline);
    e->spline = nullptr;
  }
  GWL_HIP_CHECK(hipMalloc(&e->spline, n * sizeof(double)));
  e->spline_bytes = n * sizeof(double);
  GWL_HIP_CHECK(hipMemcpy(e->spline, tx, (nx + 4) * sizeof(double), hipMemcpyHostToDevice));
  GWL_HIP_CHECK(hipMemcpy(e->spline + nx + 4, ty, (ny + 4) * sizeof(double), hipMemcpyHostToDevice));
  GWL_HIP_CHECK(hipMemcpy(e->spline + nx + 4 + ny + 4, coeffs, (size_t)nx * ny * sizeof(double),
                          hipMemcpyHostToDevice));
  SplineDev sp;
  sp.tx = e->spline;
  sp.ty = e->spline + nx + 4;
  sp.c = e->spline + nx + 4 + ny + 4;
  sp.nx = nx;
  sp.ny = ny;
  sp.xmin = bbox[0];
  sp.xmax = bbox[1];
  sp.ymin = bbox[2];
  sp.ymax = bbox[3];
  e->dev.sp = sp;
  e->has_spline = true;
  return GWL_OK;
}

int gwl_event_reserve(gwl_event* e, int64_t rows) {
  if (!e || rows < 1) { set_error("gwl_event_reserve: invalid argument"); return GWL_E_INVALID; }
  size_t need = workspace_bytes(e, rows);
  if (need <= e->ws_bytes) return GWL_OK;
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  if (e->ws) {
    (void)hipFree(e->ws);
    e->ws = nullptr;
    e->ws_bytes = 0;
    e->esc_count_dev = nullptr;  // lived in the old workspace
  }
  if (hipMalloc(&e->ws, need) != hipSuccess) {
    set_error("hipMalloc failed for workspace");
    return GWL_E_NOMEM;
  }
  e->ws_bytes = need;
  e->ws_rows = rows;
  return GWL_OK;
}

int gwl_event_device_bytes(gwl_event* e, int64_t* bytes) {
  if (!e || !bytes) { set_error("gwl_event_device_bytes: invalid argument"); return GWL_E_INVALID; }
  *bytes = (int64_t)(e->data_bytes + e->spline_bytes + e->ws_bytes);
  return GWL_OK;
}

int gwl_noise_log_likelihood(gwl_event* e, double* out) {
  if (!e || !out) { set_error("gwl_noise_log_likelihood: invalid argument"); return GWL_E_INVALID; }
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  double* d = nullptr;
  GWL_HIP_CHECK(hipMalloc(&d, sizeof(double)));
  hipLaunchKernelGGL(k_noise_lnl, dim3(1), dim3(1024), 0, nullptr, e->dev, d);
  hipError_t err = hipGetLastError();
  if (err == hipSuccess) err = hipMemcpy(out, d, sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (err != hipSuccess) return hip_fail(err, "gwl_noise_log_likelihood");
  return GWL_OK;
}

int gwl_eval(gwl_event* e, const double* params, int64_t n, int32_t ld, int32_t precision,
             double* out, double* aux, void* stream) {
  if (e && n == 0) return GWL_OK;  // an empty batch needs no buffers
  if (!e || !params || !out || n < 0 || ld < GWL_NPARAM) {
    set_error("gwl_eval: invalid argument");
    return GWL_E_INVALID;
  }
  if (n == 0) return GWL_OK;
  if (precision != GWL_FP32 && precision != GWL_FP64) {
    set_error("gwl_eval: unknown precision");
    return GWL_E_INVALID;
  }
  if ((e->dev.flags & GWL_DISTANCE_MARG) && !e->has_spline) {
    set_error("gwl_eval: distance marginalisation requested but no spline attached");
    return GWL_E_INVALID;
  }
  int rc = gwl_event_reserve(e, n);
  if (rc != GWL_OK) return rc;
  return launch_eval(e, params, n, ld, precision, out, aux, reinterpret_cast<hipStream_t>(stream));
}

int gwl_eval_time_series(gwl_event* e, const double* params, int64_t n, int32_t ld,
                         int32_t precision, double* d_inner_h, double* h_inner_h, void* stream) {
  if (e && n == 0) return GWL_OK;
  if (!e || !params || !d_inner_h || !h_inner_h || n < 0 || ld < GWL_NPARAM ||
      (precision != GWL_FP32 && precision != GWL_FP64)) {
    set_error("gwl_eval_time_series: invalid argument");
    return GWL_E_INVALID;
  }
  if (!(e->dev.flags & GWL_TIME_MARG)) {
    set_error("gwl_eval_time_series: the event does not marginalise over time");
    return GWL_E_UNSUPPORTED;
  }
  if (int rc = gwl_event_reserve(e, n)) return rc;
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  return launch_time_series(e, params, n, ld, precision, d_inner_h, h_inner_h,
                            reinterpret_cast<hipStream_t>(stream));
}

int gwl_event_time_window(gwl_event* e, int64_t* j_first, int32_t* n_window, double* t_step) {
  if (!e || !j_first || !n_window || !t_step) {
    set_error("gwl_event_time_window: invalid argument");
    return GWL_E_INVALID;
  }
  *j_first = e->dev.jlo;
  *n_window = e->dev.L;
  *t_step = e->dev.tstep;
  return GWL_OK;
}

int gwl_weighted_inner_products(gwl_event* e, const double* params, const double* factor,
                                double* out) {
  if (!e || !params || !factor || !out) {
    set_error("gwl_weighted_inner_products: invalid argument");
    return GWL_E_INVALID;
  }
  const int nd = e->dev.nd, nf = e->dev.nf;
  const size_t fb = (size_t)2 * nd * nf * sizeof(double);
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  char* mem = nullptr;
  const size_t pb = 256, pp = 256;
  if (hipMalloc((void**)&mem, pb + pp + 2 * fb) != hipSuccess) {
    set_error("hipMalloc failed");
    return GWL_E_NOMEM;
  }
  double* dp = reinterpret_cast<double*>(mem);
  double* part = reinterpret_cast<double*>(mem + pb);
  double* resp = reinterpret_cast<double*>(mem + pb + pp);
  double* fac = reinterpret_cast<double*>(mem + pb + pp + fb);
  int rc = GWL_OK;
  hipError_t err = hipMemcpy(dp, params, GWL_NPARAM * sizeof(double), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(fac, factor, fb, hipMemcpyHostToDevice);
  for (int d = 0; err == hipSuccess && rc == GWL_OK && d < nd; ++d)
    rc = launch_response(e, dp, d, GWL_FP64, resp + (size_t)2 * d * nf, nullptr);
  double h[3 * GWL_MAX_DET];
  if (err == hipSuccess && rc == GWL_OK) {
    hipLaunchKernelGGL(k_weighted_inner, dim3((unsigned)nd), dim3(256), 0, 0, e->dev, resp, fac,
                       part);
    err = hipGetLastError();
    if (err == hipSuccess) err = hipMemcpy(h, part, (size_t)3 * nd * 8, hipMemcpyDeviceToHost);
  }
  (void)hipFree(mem);
  if (err != hipSuccess) return hip_fail(err, "gwl_weighted_inner_products");
  if (rc != GWL_OK) return rc;
  const double norm = 4.0 / e->dev.T;
  out[0] = out[1] = out[2] = 0.0;
  for (int d = 0; d < nd; ++d)
    for (int c = 0; c < 3; ++c) out[c] += norm * h[3 * d + c];
  return GWL_OK;
}

int gwl_window_time_series(gwl_event* e, const double* params, int64_t n_fft, int64_t j_first,
                           int32_t n_window, double* d_inner_h, double* h_inner_h) {
  if (!e || !params || !d_inner_h || !h_inner_h || n_window < 1 || j_first < 0 ||
      n_fft < e->dev.nf || j_first + (int64_t)n_window > 2 * n_fft) {
    set_error("gwl_window_time_series: invalid argument");
    return GWL_E_INVALID;
  }
  const int nd = e->dev.nd, nf = e->dev.nf;
  const size_t fb = (size_t)2 * nd * nf * sizeof(double);
  const size_t ob = (size_t)2 * n_window * sizeof(double);
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  char* mem = nullptr;
  if (hipMalloc((void**)&mem, 256 + 256 + 2 * fb + ob) != hipSuccess) {
    set_error("hipMalloc failed");
    return GWL_E_NOMEM;
  }
  double* dp = reinterpret_cast<double*>(mem);
  double* part = reinterpret_cast<double*>(mem + 256);
  double* resp = reinterpret_cast<double*>(mem + 512);
  double* ones = reinterpret_cast<double*>(mem + 512 + fb);
  double* dout = reinterpret_cast<double*>(mem + 512 + 2 * fb);
  std::vector<double> one((size_t)2 * nd * nf, 0.0);
  for (size_t k = 0; k < one.size(); k += 2) one[k] = 1.0;
  int rc = GWL_OK;
  hipError_t err = hipMemcpy(dp, params, GWL_NPARAM * sizeof(double), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(ones, one.data(), fb, hipMemcpyHostToDevice);
  for (int d = 0; err == hipSuccess && rc == GWL_OK && d < nd; ++d)
    rc = launch_response(e, dp, d, GWL_FP64, resp + (size_t)2 * d * nf, nullptr);
  double h[3 * GWL_MAX_DET];
  if (err == hipSuccess && rc == GWL_OK) {
    hipLaunchKernelGGL(k_window_dft, dim3((unsigned)((n_window + 255) / 256)), dim3(256), 0, 0,
                       e->dev, resp, (long long)n_fft, (long long)j_first, (int)n_window, dout);
    err = hipGetLastError();
    if (err == hipSuccess)
      hipLaunchKernelGGL(k_weighted_inner, dim3((unsigned)nd), dim3(256), 0, 0, e->dev, resp,
                         ones, part);
    if (err == hipSuccess) err = hipGetLastError();
    if (err == hipSuccess) err = hipMemcpy(d_inner_h, dout, ob, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(h, part, (size_t)3 * nd * 8, hipMemcpyDeviceToHost);
  }
  (void)hipFree(mem);
  if (err != hipSuccess) return hip_fail(err, "gwl_window_time_series");
  if (rc != GWL_OK) return rc;
  const double norm = 4.0 / e->dev.T;
  for (int k = 0; k < 2 * n_window; ++k) d_inner_h[k] *= norm;
  double hh = 0.0;
  for (int d = 0; d < nd; ++d) hh += norm * h[3 * d + 2];
  *h_inner_h = hh;
  return GWL_OK;
}

int gwl_detector_response(gwl_event* e, const double* params, int32_t det, int32_t precision,
                          double* out) {
  if (!e || !params || !out || det < 0 || det >= e->dev.nd) {
    set_error("gwl_detector_response: invalid argument");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  double* dp = nullptr;
  double* dout = nullptr;
  GWL_HIP_CHECK(hipMalloc(&dp, GWL_NPARAM * sizeof(double)));
  if (hipMalloc(&dout, (size_t)2 * e->dev.nf * sizeof(double)) != hipSuccess) {
    (void)hipFree(dp);
    set_error("hipMalloc failed");
    return GWL_E_NOMEM;
  }
  int rc = GWL_OK;
  hipError_t err = hipMemcpy(dp, params, GWL_NPARAM * sizeof(double), hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    rc = launch_response(e, dp, det, precision, dout, nullptr);
    if (rc == GWL_OK) {
      err = hipMemcpy(out, dout, (size_t)2 * e->dev.nf * sizeof(double), hipMemcpyDeviceToHost);
    }
  }
  (void)hipFree(dp);
  (void)hipFree(dout);
  if (err != hipSuccess) return hip_fail(err, "gwl_detector_response");
  return rc;
}

int gwl_distance_table(gwl_ctx* ctx, const double* dist, const double* prior, int32_t nk,
                       double dref, int32_t phase, const double* daxis, int32_t nd,
                       const double* haxis, int32_t nh, double* out) {
  if (!ctx || !dist || !prior || !daxis || !haxis || !out || nk < 2 || nd < 1 || nh < 1 ||
      nh > 1024) {
    set_error("gwl_distance_table: invalid argument");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(ctx->device));
  const double delta = dist[1] - dist[0];  // base.py:887-888
  // log_norm = logsumexp(0 / d, b = prior * delta)  (base.py:1013-1016)
  double lognorm;
  {
    double s = 0.0;
    for (int k = 0; k < nk; ++k) s += prior[k] * delta;
    lognorm = std::log(s);
  }
  double *d_dist = nullptr, *d_prior = nullptr, *d_dax = nullptr, *d_hax = nullptr, *d_out = nullptr;
  size_t bytes = sizeof(double) * ((size_t)2 * nk + nd + nh + (size_t)nd * nh);
  char* base = nullptr;
  GWL_HIP_CHECK(hipMalloc(&base, bytes));
  d_dist = reinterpret_cast<double*>(base);
  d_prior = d_dist + nk;
  d_dax = d_prior + nk;
  d_hax = d_dax + nd;
  d_out = d_hax + nh;
  hipError_t err = hipMemcpy(d_dist, dist, nk * sizeof(double), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(d_prior, prior, nk * sizeof(double), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(d_dax, daxis, nd * sizeof(double), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(d_hax, haxis, nh * sizeof(double), hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    hipLaunchKernelGGL(k_distance_table, dim3(nd), dim3(256), 0, nullptr, d_dist, d_prior, nk, dref,
                       phase, delta, d_dax, nd, d_hax, nh, lognorm, d_out);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipMemcpy(out, d_out, (size_t)nd * nh * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(base);
  if (err != hipSuccess) return hip_fail(err, "gwl_distance_table");
  return GWL_OK;
}

int gwl_multi_create(gwl_ctx* ctx, int32_t n_events, gwl_event* const* events,
                     const double* const* params_dev, double* const* out_dev,
                     const int64_t* max_rows, gwl_multi** out) {
  if (!ctx || n_events < 1 || !events || !params_dev || !out_dev || !max_rows || !out) {
    set_error("gwl_multi_create: invalid argument");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(ctx->device));
  gwl_multi* m = new (std::nothrow) gwl_multi;
  if (!m) { set_error("out of host memory"); return GWL_E_NOMEM; }
  m->ctx = ctx;
  m->n_events = n_events;
  m->wave = events[0]->dev.approx;
  m->wave_path = true;
  for (int e = 0; e < n_events; ++e)
    if (!events[e] || events[e]->dev.approx != GWL_TAYLORF2 || !events[e]->dev.tf2b)
      m->wave_path = false;
  size_t items = 0, blocks = 0;
  for (int e = 0; e < n_events; ++e) {
    const gwl_event* ev = events[e];
    if (!ev || !(ev->dev.flags & GWL_TIME_MARG) || ev->dev.P != 512 || ev->dev.nd > 3 ||
        ev->dev.approx != m->wave || ev->dev.approx == GWL_HOST_POLARIZATIONS || max_rows[e] < 0 ||
        ((ev->dev.flags & GWL_DISTANCE_MARG) && !ev->has_spline)) {
      delete m;
      set_error("gwl_multi_create: every event must be time-marginalised with a 512-sample "
                "transform tile, <= 3 detectors, one source model (and a distance spline "
                "when distance-marginalised)");
      return GWL_E_UNSUPPORTED;
    }
    m->events.push_back(const_cast<gwl_event*>(ev));
    m->max_rows.push_back(max_rows[e]);
    m->rows.push_back(0);
    m->params.push_back(params_dev[e]);
    m->outs.push_back(out_dev[e]);
    const int Q = ev->dev.N / 512;
    // sub-sequence steps per item: the event's tuned choice, else ~32
    const int st = ev->multi_steps > 0 ? ev->multi_steps : 32;
    int G = (Q + st - 1) / st;
    G = std::max(1, std::min(G, std::min(Q, 16)));
    m->groups.push_back(G);
    items += (size_t)((max_rows[e] + 3) / 4) * std::min(Q, 16);
    blocks += (size_t)((max_rows[e] + 3) / 4);
  }
  m->max_items = std::max<size_t>(items, 1);
  m->max_blocks = std::max<size_t>(blocks, 1);
  // per-event partials for up to 16 groups
  for (int e = 0; e < n_events; ++e) {
    const size_t L = (size_t)m->events[e]->dev.L;
    void* part = nullptr;
    double* hh = nullptr;
    const size_t rows = (size_t)std::max<int64_t>(1, m->max_rows[e]);
    if (hipMalloc(&part, rows * 16 * L * sizeof(double2)) != hipSuccess ||
        hipMalloc((void**)&hh, rows * 16 * sizeof(double)) != hipSuccess) {
      if (part) (void)hipFree(part);
      m->parts.push_back(nullptr);
      m->hhs.push_back(nullptr);
      gwl_multi_destroy(m);
      set_error("hipMalloc failed for multi-event partials");
      return GWL_E_NOMEM;
    }
    m->parts.push_back(part);
    m->hhs.push_back(hh);
    if (m->wave_path) {
      void* rw = nullptr;
      if (hipMalloc(&rw, roww_bytes(rows)) != hipSuccess) {
        gwl_multi_destroy(m);
        set_error("hipMalloc failed for multi-event row constants");
        return GWL_E_NOMEM;
      }
      m->roww.push_back(rw);
    }
  }
  const size_t ne = (size_t)n_events;
  const size_t bytes = ne * sizeof(EvDev) + 6 * ne * sizeof(void*) + ne * sizeof(long long) +
                       ne * sizeof(int) + m->max_items * sizeof(MultiItem) +
                       m->max_blocks * (sizeof(int) + sizeof(long long)) + 64 * 8 + 16 * 256;
  if (hipMalloc(&m->dev, bytes) != hipSuccess) {
    gwl_multi_destroy(m);
    set_error("hipMalloc failed for multi-event tables");
    return GWL_E_NOMEM;
  }
  m->dev_bytes = bytes;
  char* p = reinterpret_cast<char*>(m->dev);
  auto take = [&](size_t nb) {
    char* q = p;
    p += (nb + 255) / 256 * 256;
    return q;
  };
  m->d_evs = reinterpret_cast<EvDev*>(take(ne * sizeof(EvDev)));
  m->d_params = reinterpret_cast<const double**>(take(ne * sizeof(void*)));
  m->d_outs = reinterpret_cast<double**>(take(ne * sizeof(void*)));
  m->d_parts = reinterpret_cast<void**>(take(ne * sizeof(void*)));
  m->d_hhs = reinterpret_cast<double**>(take(ne * sizeof(void*)));
  m->d_roww = reinterpret_cast<void**>(take(ne * sizeof(void*)));
  m->d_rows = reinterpret_cast<long long*>(take(ne * sizeof(long long)));
  m->d_groups = reinterpret_cast<int*>(take(ne * sizeof(int)));
  m->d_items = reinterpret_cast<MultiItem*>(take(m->max_items * sizeof(MultiItem)));
  m->d_block_event = reinterpret_cast<int*>(take(m->max_blocks * sizeof(int)));
  m->d_block_row0 = reinterpret_cast<long long*>(take(m->max_blocks * sizeof(long long)));
  m->d_n_items = reinterpret_cast<int*>(take(sizeof(int)));
  m->d_counter = reinterpret_cast<int*>(take(sizeof(int)));
  m->d_n_blocks = reinterpret_cast<int*>(take(sizeof(int)));
  std::vector<EvDev> evs;
  for (auto* ev : m->events) evs.push_back(ev->dev);
  hipError_t err = hipMemcpy(m->d_evs, evs.data(), ne * sizeof(EvDev), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(m->d_params, m->params.data(), ne * sizeof(void*), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(m->d_outs, m->outs.data(), ne * sizeof(void*), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(m->d_parts, m->parts.data(), ne * sizeof(void*), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(m->d_hhs, m->hhs.data(), ne * sizeof(void*), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(m->d_groups, m->groups.data(), ne * sizeof(int), hipMemcpyHostToDevice);
  if (err == hipSuccess && m->wave_path)
    err = hipMemcpy(m->d_roww, m->roww.data(), ne * sizeof(void*), hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    gwl_multi_destroy(m);
    return hip_fail(err, "gwl_multi_create upload");
  }
  std::vector<int64_t> zero(ne, 0);
  int rc = multi_set_rows(m, zero.data(), nullptr);
  if (rc != GWL_OK) {
    gwl_multi_destroy(m);
    return rc;
  }
  if ((m->wave_path ? multi_wave_lds_bytes() : multi_lds_bytes()) > ctx->lds_per_block) {
    gwl_multi_destroy(m);
    set_error("multi-event tile exceeds the LDS of a workgroup");
    return GWL_E_UNSUPPORTED;
  }
  *out = m;
  return GWL_OK;
}

int gwl_multi_set_tile(gwl_multi* m, const int32_t* steps_per_item) {
  if (!m || !steps_per_item) { set_error("gwl_multi_set_tile: invalid argument"); return GWL_E_INVALID; }
  for (int e = 0; e < m->n_events; ++e) {
    const int Q = m->events[e]->dev.N / 512;
    int st = steps_per_item[e];
    int G = st > 0 ? (Q + st - 1) / st : (Q + 31) / 32;
    m->groups[e] = std::max(1, std::min(G, std::min(Q, 16)));
  }
  GWL_HIP_CHECK(hipSetDevice(m->ctx->device));
  GWL_HIP_CHECK(hipMemcpy(m->d_groups, m->groups.data(), m->groups.size() * sizeof(int),
                          hipMemcpyHostToDevice));
  std::vector<int64_t> rows(m->rows.begin(), m->rows.end());
  return multi_set_rows(m, rows.data(), nullptr);
}

int gwl_multi_set_rows(gwl_multi* m, const int64_t* rows, void* stream) {
  if (!m || !rows) { set_error("gwl_multi_set_rows: invalid argument"); return GWL_E_INVALID; }
  GWL_HIP_CHECK(hipSetDevice(m->ctx->device));
  return multi_set_rows(m, rows, reinterpret_cast<hipStream_t>(stream));
}

int gwl_multi_launch(gwl_multi* m, void* stream) {
  if (!m) { set_error("gwl_multi_launch: NULL"); return GWL_E_INVALID; }
  GWL_HIP_CHECK(hipSetDevice(m->ctx->device));
  (void)hipGetLastError();
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (m->captured) {
    GWL_HIP_CHECK(hipGraphLaunch(m->exec, st));
    return GWL_OK;
  }
  return multi_launch_raw(m, st);
}

int gwl_multi_capture(gwl_multi* m, void* stream) {
  if (!m || !stream) {
    set_error("gwl_multi_capture: needs a non-default stream");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(m->ctx->device));
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (m->captured) {
    (void)hipGraphExecDestroy(m->exec);
    (void)hipGraphDestroy(m->graph);
    m->captured = false;
  }
  GWL_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  m->capturing = true;  // grids sized for the largest work list
  int rc = multi_launch_raw(m, st);
  m->capturing = false;
  hipGraph_t g = nullptr;
  hipError_t err = hipStreamEndCapture(st, &g);
  if (rc != GWL_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (err != hipSuccess) return hip_fail(err, "hipStreamEndCapture");
  err = hipGraphInstantiate(&m->exec, g, nullptr, nullptr, 0);
  if (err != hipSuccess) {
    (void)hipGraphDestroy(g);
    return hip_fail(err, "hipGraphInstantiate");
  }
  m->graph = g;
  m->captured = true;
  return GWL_OK;
}

int gwl_multi_destroy(gwl_multi* m) {
  if (!m) return GWL_OK;
  (void)hipSetDevice(m->ctx->device);
  if (m->captured) {
    (void)hipGraphExecDestroy(m->exec);
    (void)hipGraphDestroy(m->graph);
  }
  for (void* p : m->parts)
    if (p) (void)hipFree(p);
  for (double* p : m->hhs)
    if (p) (void)hipFree(p);
  for (void* p : m->roww)
    if (p) (void)hipFree(p);
  if (m->dev) (void)hipFree(m->dev);
  delete m;
  return GWL_OK;
}

int gwl_spline2d_coefficients(int32_t nx, int32_t ny, const double* x, const double* y,
                              const double* z, double* tx, double* ty, double* c) {
  if (nx < 4 || ny < 4 || !x || !y || !z || !tx || !ty || !c) {
    set_error("gwl_spline2d_coefficients: invalid argument");
    return GWL_E_INVALID;
  }
  for (int i = 1; i < nx; ++i)
    if (!(x[i] > x[i - 1])) { set_error("x must be strictly increasing"); return GWL_E_INVALID; }
  for (int i = 1; i < ny; ++i)
    if (!(y[i] > y[i - 1])) { set_error("y must be strictly increasing"); return GWL_E_INVALID; }
  knots(x, nx, tx);
  knots(y, ny, ty);
  LU lx{nx}, ly{ny};
  if (!lx.factor(x, tx) || !ly.factor(y, ty)) {
    set_error("singular collocation matrix");
    return GWL_E_INVALID;
  }
  std::memcpy(c, z, sizeof(double) * (size_t)nx * ny);
  // along x: columns of c (stride ny)
  for (int j = 0; j < ny; ++j) lx.solve(c + j, (size_t)ny);
  // along y: rows of c (stride 1)
  for (int i = 0; i < nx; ++i) ly.solve(c + (size_t)i * ny, 1);
  return GWL_OK;
}

int gwl_spline2d_eval_host(int32_t nx, int32_t ny, const double* tx, const double* ty,
                           const double* c, int64_t m, const double* x, const double* y,
                           double* out) {
  if (nx < 4 || ny < 4 || !tx || !ty || !c || m < 0 || (m && (!x || !y || !out))) {
    set_error("gwl_spline2d_eval_host: invalid argument");
    return GWL_E_INVALID;
  }
  SplineDev sp;
  sp.tx = tx;
  sp.ty = ty;
  sp.c = c;
  sp.nx = nx;
  sp.ny = ny;
  sp.xmin = tx[0];
  sp.xmax = tx[nx + 3];
  sp.ymin = ty[0];
  sp.ymax = ty[ny + 3];
  for (int64_t i = 0; i < m; ++i) out[i] = spline2d(sp, x[i], y[i]);
  return GWL_OK;
}

int gwl_fft_batch(gwl_ctx* ctx, const void* in, void* out, int64_t batch, int32_t P,
                  int32_t precision, void* stream) {
  if (!ctx || !in || !out || batch < 0) { set_error("gwl_fft_batch: invalid argument"); return GWL_E_INVALID; }
  if (batch == 0) return GWL_OK;
  GWL_HIP_CHECK(hipSetDevice(ctx->device));
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  (void)hipGetLastError();
  if (precision == GWL_FP32) return fft_dispatch<float>(ctx, in, out, batch, P, st);
  if (precision == GWL_FP64) return fft_dispatch<double>(ctx, in, out, batch, P, st);
  set_error("gwl_fft_batch: unknown precision");
  return GWL_E_INVALID;
}

int gwl_ln_i0(gwl_ctx* ctx, const double* in, double* out, int64_t n, void* stream) {
  if (!ctx || !in || !out || n < 0) { set_error("gwl_ln_i0: invalid argument"); return GWL_E_INVALID; }
  if (n == 0) return GWL_OK;
  GWL_HIP_CHECK(hipSetDevice(ctx->device));
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_ln_i0, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), in, out, (long long)n);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

}  // extern "C"
