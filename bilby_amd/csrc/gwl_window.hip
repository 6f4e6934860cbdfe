// Window transforms on MFMA for the time-marginalised binned likelihoods
// (relative binning and multibanding): basis construction, the operand
// GEMM and the window epilogue.  The layout and the algebra are described
// in gwl_window.h.
#include "gwl_window.h"

#include <algorithm>

namespace gwl {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

size_t al256(size_t x) { return (x + 255) / 256 * 256; }
long long up(long long x, long long a) { return (x + a - 1) / a * a; }

// ------------------------------------------------------------------ basis
struct RelbinBasisArgs {
  int nd, nb, L;
  long long N, j_first, n_x;
  const long long* edges;  // [nb + 1] grid indices
  const double* fe;        // [nb + 1] edge frequencies
  double df;
  const double2* x;        // [nd][n_x] from grid bin edges[0]
  double2* m;              // [K][L]
};

// One thread per (basis row k = d (nb + 1) + e, window bin jj).  Bin b holds
// grid bins [edges[b], edges[b + 1]), the last one also edges[nb]
// (relative.py:216-224, 384-391); inside bin b the rebuilt waveform is
// h0 (r_b (f_{b+1} - f) + r_{b+1} (f - f_b)) / w_b.
__global__ __launch_bounds__(256) void k_win_basis_relbin(RelbinBasisArgs a) {
  const int jj = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (jj >= a.L) return;
  const int d = k / (a.nb + 1), e = k % (a.nb + 1);
  const long long j = a.j_first + jj;
  const long long e0 = a.edges[0];
  const double2* xd = a.x + (size_t)d * a.n_x;
  double sr = 0.0, si = 0.0;
  for (int side = 0; side < 2; ++side) {
    const int b = side == 0 ? e - 1 : e;  // bin below / above the edge
    if (b < 0 || b >= a.nb) continue;
    const long long n0 = a.edges[b];
    const long long n1 = b == a.nb - 1 ? a.edges[b + 1] : a.edges[b + 1] - 1;
    const double flo = a.fe[b], fhi = a.fe[b + 1], w = fhi - flo;
    for (long long n = n0; n <= n1; ++n) {
      const double f = (double)n * a.df;
      const double wt = side == 0 ? (f - flo) / w : (fhi - f) / w;
      const double2 x = xd[n - e0];
      const long long r = (long long)(((unsigned long long)n * (unsigned long long)j) %
                                      (unsigned long long)a.N);
      double s, c;
      sincospi(-2.0 * (double)r / (double)a.N, &s, &c);
      const double xr = x.x * wt, xi = x.y * wt;
      sr += xr * c - xi * s;
      si += xr * s + xi * c;
    }
  }
  a.m[(size_t)k * a.L + jj] = make_double2(sr, si);
}

// One thread per (point p, window bin jj): exp(-2 pi i n_p j / N).
__global__ __launch_bounds__(256) void k_win_basis_points(const long long* idx, long long N,
                                                          long long j_first, int L, double2* m) {
  const int jj = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (jj >= L) return;
  const long long n = idx[p];
  double2 v = make_double2(0.0, 0.0);
  if (n >= 0) {
    const long long j = j_first + jj;
    const long long r = (long long)(((unsigned long long)n * (unsigned long long)j) %
                                    (unsigned long long)N);
    double s, c;
    sincospi(-2.0 * (double)r / (double)N, &s, &c);
    v = make_double2(c, s);
  }
  m[(size_t)p * L + jj] = v;
}

// Bt[Np][Kp] from M[K][L] (float64 and float32 copies).
__global__ __launch_bounds__(256) void k_win_pack(const double2* m, int K, int L, int Kp,
                                                  long long total, double* bt64, float* bt32) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const long long row = t / Kp;
  const int col = (int)(t % Kp);
  double v = 0.0;
  if (row < 2LL * L && col < 2 * K) {
    const int jj = (int)(row < L ? row : row - L);
    const int k = col < K ? col : col - K;
    const double2 z = m[(size_t)k * L + jj];
    if (row < L)
      v = col < K ? z.x : -z.y;
    else
      v = col < K ? z.y : z.x;
  }
  bt64[t] = v;
  bt32[t] = (float)v;
}

// ------------------------------------------------------------------ operand padding
__global__ __launch_bounds__(256) void k_win_clear(void* a, int fp64, long long n, long long Mp,
                                                   int K2, int Kp) {
  const long long pad = n * (long long)(Kp - K2);
  const long long total = pad + (Mp - n) * (long long)Kp;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    long long r, c;
    if (t < pad) {
      r = t / (Kp - K2);
      c = K2 + t % (Kp - K2);
    } else {
      r = n + (t - pad) / Kp;
      c = (t - pad) % Kp;
    }
    const size_t i = (size_t)r * Kp + c;
    if (fp64)
      static_cast<double*>(a)[i] = 0.0;
    else
      static_cast<float*>(a)[i] = 0.f;
  }
}

// ------------------------------------------------------------------ GEMM
// Split-K partials: C_z[M][N] = A[:, k in chunk z] . Bt[:, k in chunk z]^T
// (float32), chunk z = blockIdx.z of kWinKc columns; M % 128 == 0,
// N % 128 == 0, K % 32 == 0, row strides K.  Inside a chunk the MFMA
// accumulators restart every kFlush k-tiles (32 columns, 16 MFMA steps)
// after adding into a float32 chunk total; the epilogue adds the chunk
// totals in float64.  The chunks also multiply the workgroups of a short
// row tile.
// A workgroup of 4 waves computes a 128 x 128 tile, each wave 2 x 2 MFMA
// 32 x 32 blocks (output layout: lane l, accumulator v -> row
// 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32).
constexpr int kBM = 128, kBN = 128, kBK = 32, kFlush = 1;

__global__ __launch_bounds__(256) void k_win_gemm_f32(const float* __restrict__ A,
                                                      const float* __restrict__ Bt,
                                                      float* __restrict__ C, int M, int N, int K,
                                                      int kc) {
  __shared__ float As[kBM][kBK + 1];
  __shared__ float Bs[kBN][kBK + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
  const int kb = blockIdx.z * kc, ke = min(K, kb + kc);
  f32x16 acc[2][2];
  float tot[2][2][16];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        acc[i][j][v] = 0.f;
        tot[i][j][v] = 0.f;
      }
  int tiles = 0;
  for (int k0 = kb; k0 < ke; k0 += kBK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + 256 * q;
      const int row = idx >> 3, c4 = (idx & 7) * 4;
      const float4 va = *reinterpret_cast<const float4*>(A + (size_t)(m0 + row) * K + k0 + c4);
      const float4 vb = *reinterpret_cast<const float4*>(Bt + (size_t)(n0 + row) * K + k0 + c4);
      As[row][c4 + 0] = va.x; As[row][c4 + 1] = va.y; As[row][c4 + 2] = va.z; As[row][c4 + 3] = va.w;
      Bs[row][c4 + 0] = vb.x; Bs[row][c4 + 1] = vb.y; Bs[row][c4 + 2] = vb.z; Bs[row][c4 + 3] = vb.w;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 2) {
      float a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[wr + i * 32 + (lane & 31)][kk + (lane >> 5)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[wc + j * 32 + (lane & 31)][kk + (lane >> 5)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
    if (++tiles == kFlush || k0 + kBK >= ke) {
      tiles = 0;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            tot[i][j][v] += acc[i][j][v];
            acc[i][j][v] = 0.f;
          }
    }
  }
  float* Cz = C + (size_t)blockIdx.z * M * N;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = m0 + wr + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int col = n0 + wc + j * 32 + (lane & 31);
        Cz[(size_t)row * N + col] = tot[i][j][v];
      }
}

// ------------------------------------------------------------------ epilogue
struct WinEpiArgs {
  EvDev ev;
  const double* C64;  // [Mp][Np] (float64 GEMM), or null
  const float* C32;   // [nchunk][Mp][Np] split-K partials (float32 GEMM)
  int nchunk;
  long long Mp;
  int Np, L;
  long long n;
  const double* params;
  int ld;
  const double* hh;   // [n][3]; <h|h> at [2]
  const int* valid;
  int vstride;
  long long j_first;
  double start;
  int offset;
  double t_step, delta_tc;
  double* out;
};

constexpr int kWinEpiRows = 4;

// base.py:794-820 over the window: one wavefront per row, lanes over the
// window bins, running-max logsumexp with the uniform time prior weight
// delta_tc / (t_max - t_min).
__global__ __launch_bounds__(64 * kWinEpiRows) void k_win_epi(WinEpiArgs a) {
  extern __shared__ double knots[];
  const EvDev& ev = a.ev;
  const bool dist = (ev.flags & GWL_DISTANCE_MARG) != 0;
  if (dist) stage_knots(ev, knots);
  __syncthreads();
  const long long r = (long long)blockIdx.x * kWinEpiRows + (threadIdx.x >> 6);
  if (r >= a.n) return;
  const int lane = threadIdx.x & 63;
  if (!a.valid[r * a.vstride]) {
    if (lane == 0) a.out[r] = row_valid_sentinel();
    return;
  }
  const double* p = a.params + r * a.ld;
  const double hh = a.hh[3 * r + 2];
  const double hs = dist ? p[GWL_P_LUMINOSITY_DISTANCE] / ev.dref : 1.0;
  const double jit = (ev.flags & GWL_JITTER_TIME) ? p[GWL_P_TIME_JITTER] : 0.0;
  SplineY yy;
  if (dist) yy = spline_y(ev, knots + ev.sp.nx + 4, hh * hs * hs);
  double m = -INFINITY, s = 0.0;
  for (int jj = lane; jj < a.L; jj += 64) {
    const double t = (a.start + (double)(a.j_first + jj + a.offset) * a.t_step) + jit;
    if (!(t >= ev.tmin && t <= ev.tmax)) continue;
    double xr = 0.0, xi = 0.0;
    if (a.C64) {
      const double* crow = a.C64 + (size_t)r * a.Np;
      xr = crow[jj];
      xi = crow[a.L + jj];
    } else {
      for (int c = 0; c < a.nchunk; ++c) {
        const float* crow = a.C32 + ((size_t)c * a.Mp + r) * a.Np;
        xr += (double)crow[jj];
        xi += (double)crow[a.L + jj];
      }
    }
    lse_push(m, s, epi_value(ev, knots, yy, hh, hs, xr, xi) + time_log_weight(ev, t));
  }
  lse_wave(m, s);
  if (lane == 0) {
    const double b = time_lse_scale(ev, a.delta_tc);
    a.out[r] = (m > -INFINITY) ? log(b * s) + m : -INFINITY;
  }
}

// One thread per (row, window bin): the transform X_j itself (the float32
// chunk partials summed in float64) and the row's <h|h>; NaN for rows whose
// waveform failed.
__global__ __launch_bounds__(256) void k_win_series(WinEpiArgs a, double* dh, double* hho) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n * a.L) return;
  const long long r = t / a.L;
  const int jj = (int)(t - r * a.L);
  const bool ok = a.valid[r * a.vstride] != 0;
  double xr = NAN, xi = NAN;
  if (ok) {
    if (a.C64) {
      const double* crow = a.C64 + (size_t)r * a.Np;
      xr = crow[jj];
      xi = crow[a.L + jj];
    } else {
      xr = xi = 0.0;
      for (int c = 0; c < a.nchunk; ++c) {
        const float* crow = a.C32 + ((size_t)c * a.Mp + r) * a.Np;
        xr += (double)crow[jj];
        xi += (double)crow[a.L + jj];
      }
    }
  }
  dh[2 * t] = xr;
  dh[2 * t + 1] = xi;
  if (jj == 0) hho[r] = ok ? a.hh[3 * r + 2] : NAN;
}

int pack_basis(gwl_event* ev, const double2* m, int K, int L, WinBasis* wb) {
  wb->K = K;
  wb->L = L;
  wb->Kp = (int)up(2LL * K, 32);
  wb->Np = (int)up(2LL * L, 128);
  const size_t n = (size_t)wb->Np * wb->Kp;
  wb->bytes = al256(n * 8) + al256(n * 4);
  if (hipMalloc(&wb->data, wb->bytes) != hipSuccess) {
    wb->data = nullptr;
    set_error("hipMalloc failed for the time-window basis");
    return GWL_E_NOMEM;
  }
  wb->bt64 = static_cast<double*>(wb->data);
  wb->bt32 = reinterpret_cast<float*>(static_cast<char*>(wb->data) + al256(n * 8));
  hipLaunchKernelGGL(k_win_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, m, K, L,
                     wb->Kp, (long long)n, wb->bt64, wb->bt32);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

}  // namespace

int win_gemm_f32(const float* A, const float* Bt, float* C, int Mp, int Np, int Kp, int kc,
                 hipStream_t st) {
  if (Mp % kBM || Np % kBN || Kp % kBK || kc % kBK || kc <= 0) {
    set_error("win_gemm_f32: shapes must be multiples of the tile");
    return GWL_E_INVALID;
  }
  const int nz = (Kp + kc - 1) / kc;
  hipLaunchKernelGGL(k_win_gemm_f32, dim3((unsigned)(Np / kBN), (unsigned)(Mp / kBM), (unsigned)nz),
                     dim3(256), 0, st, A, Bt, C, Mp, Np, Kp, kc);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

void win_basis_free(WinBasis* wb) {
  if (wb->data) (void)hipFree(wb->data);
  *wb = WinBasis();
}

int win_basis_relbin(gwl_event* ev, int nd, int nb, const long long* edges, const double* fe,
                     const double2* x, long long n_x, WinBasis* wb) {
  const int K = nd * (nb + 1), L = wb->L;
  double2* m = nullptr;
  if (hipMalloc((void**)&m, (size_t)K * L * sizeof(double2)) != hipSuccess) {
    set_error("hipMalloc failed for the time-window basis");
    return GWL_E_NOMEM;
  }
  RelbinBasisArgs a;
  a.nd = nd;
  a.nb = nb;
  a.L = L;
  a.N = wb->n_fft > 0 ? wb->n_fft : ev->dev.N;
  a.j_first = wb->j_first;
  a.n_x = n_x;
  a.edges = edges;
  a.fe = fe;
  a.df = ev->dev.df;
  a.x = x;
  a.m = m;
  hipLaunchKernelGGL(k_win_basis_relbin, dim3((unsigned)((L + 255) / 256), (unsigned)K), dim3(256),
                     0, 0, a);
  hipError_t err = hipGetLastError();
  int rc = err == hipSuccess ? pack_basis(ev, m, K, L, wb) : hip_fail(err, "k_win_basis_relbin");
  if (rc == GWL_OK) {
    err = hipDeviceSynchronize();
    if (err != hipSuccess) rc = hip_fail(err, "time-window basis");
  }
  (void)hipFree(m);
  if (rc != GWL_OK) win_basis_free(wb);
  return rc;
}

int win_basis_points(gwl_event* ev, int n, const long long* idx, long long N, WinBasis* wb) {
  const int L = wb->L;
  double2* m = nullptr;
  if (hipMalloc((void**)&m, (size_t)n * L * sizeof(double2)) != hipSuccess) {
    set_error("hipMalloc failed for the time-window basis");
    return GWL_E_NOMEM;
  }
  hipLaunchKernelGGL(k_win_basis_points, dim3((unsigned)((L + 255) / 256), (unsigned)n), dim3(256),
                     0, 0, idx, N, wb->j_first, L, m);
  hipError_t err = hipGetLastError();
  int rc = err == hipSuccess ? pack_basis(ev, m, n, L, wb) : hip_fail(err, "k_win_basis_points");
  if (rc == GWL_OK) {
    err = hipDeviceSynchronize();
    if (err != hipSuccess) rc = hip_fail(err, "time-window basis");
  }
  (void)hipFree(m);
  if (rc != GWL_OK) win_basis_free(wb);
  return rc;
}

long long win_mp(long long rows) { return up(std::max(1LL, rows), 128); }

int win_chunks(const WinBasis& wb) { return (wb.Kp + kWinKc - 1) / kWinKc; }

size_t win_ws_bytes(const WinBasis& wb, long long rows) {
  const long long Mp = win_mp(std::min(rows, kWinTileRows));
  const size_t c = std::max<size_t>((size_t)Mp * wb.Np * 8, (size_t)win_chunks(wb) * Mp * wb.Np * 4);
  return al256((size_t)Mp * wb.Kp * 8) + al256(c);
}

void* win_a(const WinBasis&, void* ws) { return ws; }

void* win_c(const WinBasis& wb, void* ws, long long rows) {
  const long long Mp = win_mp(std::min(rows, kWinTileRows));
  return static_cast<char*>(ws) + al256((size_t)Mp * wb.Kp * 8);
}

int win_clear(const WinBasis& wb, void* a, bool fp64, long long n, hipStream_t st) {
  const long long Mp = win_mp(n);
  const long long total = n * (long long)(wb.Kp - 2 * wb.K) + (Mp - n) * (long long)wb.Kp;
  if (total <= 0) return GWL_OK;
  const unsigned blocks = (unsigned)std::min<long long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(k_win_clear, dim3(blocks), dim3(256), 0, st, a, fp64 ? 1 : 0, n, Mp,
                     2 * wb.K, wb.Kp);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

int win_finish(gwl_event* ev, const WinBasis& wb, const void* a, bool fp64, void* c,
               long long n, const double* params, int ld, const double* hh, const int* valid,
               int valid_stride, double* out, hipStream_t st, double* series_dh,
               double* series_hh) {
  const long long Mp = win_mp(n);
  const int nchunk = win_chunks(wb);
  if (fp64) {
    if (int rc = gwl_gemm(ev->ctx, GWL_ROQ_FP64, a, nullptr, wb.bt64, nullptr, c, (int32_t)Mp,
                          wb.Np, wb.Kp, st))
      return rc;
  } else {
    hipLaunchKernelGGL(k_win_gemm_f32,
                       dim3((unsigned)(wb.Np / kBN), (unsigned)(Mp / kBM), (unsigned)nchunk),
                       dim3(256), 0, st, static_cast<const float*>(a), wb.bt32,
                       static_cast<float*>(c), (int)Mp, wb.Np, wb.Kp, kWinKc);
    GWL_HIP_CHECK(hipGetLastError());
  }
  WinEpiArgs e;
  e.ev = ev->dev;
  e.C64 = fp64 ? static_cast<const double*>(c) : nullptr;
  e.C32 = fp64 ? nullptr : static_cast<const float*>(c);
  e.nchunk = nchunk;
  e.Mp = Mp;
  e.Np = wb.Np;
  e.L = wb.L;
  e.n = n;
  e.params = params;
  e.ld = ld;
  e.hh = hh;
  e.valid = valid;
  e.vstride = valid_stride;
  e.j_first = wb.j_first;
  e.start = wb.start;
  e.offset = wb.offset;
  e.t_step = wb.t_step;
  e.delta_tc = wb.delta_tc;
  e.out = out;
  if (series_dh) {
    const long long total = n * (long long)wb.L;
    hipLaunchKernelGGL(k_win_series, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, e,
                       series_dh, series_hh);
    GWL_HIP_CHECK(hipGetLastError());
    return GWL_OK;
  }
  const size_t lds = (ev->dev.flags & GWL_DISTANCE_MARG)
                         ? (size_t)(ev->dev.sp.nx + ev->dev.sp.ny + 8) * sizeof(double) : 0;
  hipLaunchKernelGGL(k_win_epi, dim3((unsigned)((n + kWinEpiRows - 1) / kWinEpiRows)),
                     dim3(64 * kWinEpiRows), lds, st, e);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

}  // namespace gwl
