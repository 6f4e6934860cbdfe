// Binned likelihoods: relative binning (below) and multibanding (gwl_mb,
// k_mb), sharing the per-row constants and the marginalisation epilogue.
//
// Relative-binning likelihood (bilby/gw/likelihood/relative.py,
// RelativeBinningGravitationalWaveTransient; Zackay, Dai & Venumadhav,
// arXiv:1806.08792).
//
// Setup (host, float64, in the caller): bin edges f_0 < ... < f_B on the
// frequency grid (relative.py:177-231), the fiducial detector responses
// h0_d(f_e) at the edges (:233-258) and the per-bin summary data
// a0, a1, b0, b1 (:314-360).  Per parameter row the kernel evaluates the
// TaylorF2 detector response only at the B + 1 edges, the ratio
// r_d(f_e) = h_d(f_e) / h0_d(f_e) (:362-374), and
//   <d|h> = sum_b a0 conj(r0) + a1 conj(r1)
//   <h|h> = sum_b b0 |r0|^2 + 2 b1 Re(r0 conj(r1))           (:393-407)
// with r0 = (r_{b+1} + r_b) / 2, r1 = (r_{b+1} - r_b) / (f_{b+1} - f_b),
// then the plain / phase / distance marginalised value of base.py:448-477.
//
// Device layout: the factors 1/2 and 1/width are folded into the summary
// at creation, so with S = r_{b+1} + r_b and D = r_{b+1} - r_b
//   <d|h> += A0 conj(S) + A1 conj(D),  A0 = a0 / 2,  A1 = a1 / w
//   <h|h> += B0 |S|^2 + B1 Re(S conj(D)),  B0 = b0 / 4,  B1 = b1 / w
// and g_d(e) = f_e^-7/6 / h0_d(f_e) is stored per edge, so that
// r_d(e) = C_d g_d(e) exp(-2 pi i (Psi(f_e) / 2 pi + f_e dt_d)), C_d being the
// complex detector amplitude of the row.
//
// Three launches: k_relbin_rows (one thread per row: antenna patterns,
// delays, TaylorF2 constants), k_relbin (one wavefront per row, lanes over
// edges: lane l evaluates edge 64 k + l and takes edge 64 k + l + 1 from
// lane l + 1, lane 63 from lane 0 of the next chunk, so every edge's
// waveform is computed once) and k_relbin_epi (one thread per row:
// marginalisation).  Summary data and edge tables are a few tens of KB per
// event and stay in L2.
#include "gwl_internal.h"
#include "gwl_window.h"

#include <cmath>
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <string>
#include <vector>

using gwl::set_error;

struct gwl_relbin {
  gwl_event* ev = nullptr;
  int nb = 0;
  void* data = nullptr;
  size_t data_bytes = 0;
  double* fe = nullptr;       // [nb + 1] edge frequencies
  double* basis = nullptr;    // [nb + 1][kNB] TaylorF2 phase basis at the edges
  double2* g64 = nullptr;     // [nd][nb + 1] f^-7/6 / h0
  float2* g32 = nullptr;
  double* s64 = nullptr;      // [nd][nb][6] A0 re, A0 im, A1 re, A1 im, B0, B1
  float* s32 = nullptr;
  void* ws = nullptr;         // row constants and per-row sums
  size_t ws_bytes = 0;
  gwl::WinBasis win;          // time marginalisation (gwl_relbin_set_time)
  bool has_win = false;
  void* tws = nullptr;        // GEMM operand and output of one row tile
  size_t tws_bytes = 0;
  // posterior reconstruction of the time on the 1/16384 s grid
  // (gwl_relbin_set_upsampled): the integrand, the <h|h> weights and the
  // bin of each grid bin on [edges[0], edges[nb]], and the basis of the
  // last window asked for
  void* up_data = nullptr;
  double2* up_x = nullptr;    // [nd][n_x] h0 conj(d) / S (zero off the mask)
  double* up_y = nullptr;     // [nd][n_x] 4/T |h0|^2 / S on the mask
  int* up_bin = nullptr;      // [n_x]
  long long* up_edges = nullptr;
  long long up_nx = 0, up_nfft = 0, up_edges_first = 0;
  gwl::WinBasis up;
  bool has_up = false;
};

namespace gwl {
// One band of the IFFT-FFT <h|h> (gwl_mb_set_fft), on the device: the
// band's M' = M_b / 2 point transforms split as M' = R P (P points in LDS).
struct MbFftBand {
  int mh, P, logP, R;     // M', LDS transform length, log2 P, M' / P
  int p0, L, ks;          // banded points [p0, p0 + L) at bins ks .. ks + L - 1
  const double2* tab;     // [4 M'] exp(-2 pi i k / (4 M'))
  const double* odd;      // [nd][M'] (4 / T^_b) Re Ibc_b[2 j + 1] / M'^2
  const double* sqw;      // [L] square root of the band's window at its points
};
}  // namespace gwl

// Multibanded likelihood (bilby/gw/likelihood/multiband.py:728-807): per
// banded point the frequency, TaylorF2 phase basis, f^-7/6, and per detector
// the linear (complex) and quadratic coefficients; for the IFFT-FFT <h|h>
// the quadratic coefficients hold the diagonal part and the bands' odd-bin
// transforms are added by k_mb_fft_lds / k_mb_fft_split.
struct gwl_mb {
  gwl_event* ev = nullptr;
  int n = 0;
  void* data = nullptr;
  size_t data_bytes = 0;
  double* fe = nullptr;       // [n]
  double* basis = nullptr;    // [n][kNB]
  double* amp = nullptr;      // [n] f^-7/6
  double2* lin64 = nullptr;   // [nd][n]
  float2* lin32 = nullptr;
  double* quad64 = nullptr;   // [nd][n]
  float* quad32 = nullptr;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  gwl::WinBasis win;          // time marginalisation (gwl_mb_set_time)
  bool has_win = false;
  void* tws = nullptr;
  size_t tws_bytes = 0;
  std::vector<gwl::MbFftBand> fft;   // IFFT-FFT bands (empty: linear interpolation)
  void* fft_data = nullptr;          // their tables
  void* fft_ws = nullptr;            // per-band row sums and split-transform scratch
  size_t fft_ws_bytes = 0;
};

namespace gwl {
namespace {

// Per-row constants (k_relbin_rows): phase coefficients in cycles, detector
// time shifts, complex detector amplitudes, the upper band edge.
struct RelRow {
  double c[kNB];
  double dt[GWL_MAX_DET];
  double cre[GWL_MAX_DET], cim[GWL_MAX_DET];
  double fhi;
  int valid, pad;
};

struct RelbinArgs {
  EvDev ev;
  int nb;
  const double* fe;
  const double* basis;
  const void* g;     // double2 or float2 [nd][nb + 1]
  const void* s;     // double or float [nd][nb][6]
  const double* params;
  int ld;
  long long n;
  RelRow* rows;      // [n]
  double* sums;      // [n][3]: Re <d|h>, Im <d|h>, <h|h>
  double* out;
  double* aux;
  double antenna_time;  // antenna patterns at this time (NaN: the row's time)
  void* win_a;          // time marginalisation: edge ratios as GEMM operand rows, or null
  int win_k, win_kp;
};

size_t relbin_ws_bytes(long long n) {
  return ((size_t)n * sizeof(RelRow) + 255) / 256 * 256 + (size_t)n * 3 * sizeof(double);
}

template <typename R> struct RelbinT;
template <> struct RelbinT<double> {
  using C = double2;
  __device__ static C load(const void* p, size_t i) { return static_cast<const double2*>(p)[i]; }
  __device__ static double sval(const void* p, size_t i) { return static_cast<const double*>(p)[i]; }
};
template <> struct RelbinT<float> {
  using C = float2;
  __device__ static C load(const void* p, size_t i) { return static_cast<const float2*>(p)[i]; }
  __device__ static float sval(const void* p, size_t i) { return static_cast<const float*>(p)[i]; }
};

__device__ __forceinline__ double2 shfl_c(double2 v, int src) {
  return make_double2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}
__device__ __forceinline__ float2 shfl_c(float2 v, int src) {
  return make_float2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

__device__ __forceinline__ double reduce_wave(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One thread per row: everything that depends on the parameter row only
// (interferometer.py:303-368 antenna and delays, the TaylorF2 constants).
// With time marginalisation and jitter the geocent time is shifted by the
// row's time_jitter first (base.py:427-428); the antenna patterns are taken
// at a.antenna_time when it is set (interferometer.py:336-339,
// reference_time).
__global__ __launch_bounds__(128) void k_relbin_rows(RelbinArgs a) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n) return;
  const EvDev& ev = a.ev;
  const double* p = a.params + r * a.ld;
  const double ra = p[GWL_P_RA], dec = p[GWL_P_DEC], psi = p[GWL_P_PSI];
  const bool jitter = (ev.flags & GWL_TIME_MARG) && (ev.flags & GWL_JITTER_TIME);
  const double tc = jitter ? p[GWL_P_GEOCENT_TIME] + p[GWL_P_TIME_JITTER] : p[GWL_P_GEOCENT_TIME];
  const double iota = p[GWL_P_THETA_JN];
  const bool fixed_antenna = !isnan(a.antenna_time);
  RelRow* o = a.rows + r;
  double c[10], amp0, fisco;
  o->valid = tf2_row_constants(p, ev.fref, 1.0, c, amp0, fisco);
  o->pad = 0;
#pragma unroll
  for (int k = 0; k < kNB; ++k) o->c[k] = c[k];
  const double ci = cos(iota);
  const double ap = amp0 * (1.0 + ci * ci) / 2.0, ax = amp0 * ci;
  // the sky frame once per row (and at the fixed antenna time when asked)
  const SkyBasis sky = sky_basis(ra, dec, tc, psi);
  const SkyBasis sky_ant = fixed_antenna ? sky_basis(ra, dec, a.antenna_time, psi) : sky;
  for (int d = 0; d < GWL_MAX_DET; ++d) {
    double dt = 0.0, cre = 0.0, cim = 0.0;
    if (d < ev.nd) {
      double fp, fc, delay;
      detector_projection(ev.tensor[d], ev.vertex[d], sky, fp, fc, delay);
      if (fixed_antenna) {
        double delay_ref;
        detector_projection(ev.tensor[d], ev.vertex[d], sky_ant, fp, fc, delay_ref);
      }
      dt = (tc - ev.start) + delay;  // interferometer.py:355-358
      cre = fp * ap;
      cim = -fc * ax;
    }
    o->dt[d] = dt;
    o->cre[d] = cre;
    o->cim[d] = cim;
  }
  o->fhi = (ev.flags & GWL_TF2_END_ISCO) ? fmin(fisco, ev.wf_fmax) : ev.wf_fmax;
}

// r_d at edge e for every detector (zero outside the waveform's band).
template <typename R, int ND>
__device__ __forceinline__ void edge_ratio(const RelbinArgs& a, const RelRow& rr, double flo, int e,
                                           typename RelbinT<R>::C* r) {
  using C = typename RelbinT<R>::C;
  const bool live = e <= a.nb;
  const int ee = live ? e : a.nb;
  const double f = a.fe[ee];
  double b[kNB];
#pragma unroll
  for (int k = 0; k < kNB; ++k) b[k] = a.basis[(size_t)ee * kNB + k];
  const double psi = tf2_phase_cycles(rr.c, b);
  const bool band = live && f >= flo && f <= rr.fhi;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    double ph = fma(f, rr.dt[d], psi);
    ph -= rint(ph);
    const C w = expm2pi((R)ph);
    const C g = RelbinT<R>::load(a.g, (size_t)d * (a.nb + 1) + ee);
    const C amp = cmk<C>((R)rr.cre[d], (R)rr.cim[d]);
    const C v = cmul(cmul(amp, g), w);
    r[d] = band ? v : cmk<C>((R)0, (R)0);
  }
}

// One wavefront per row: lane l evaluates edge 64 k + l, takes edge
// 64 k + l + 1 from lane l + 1 (lane 63: lane 0 of the next chunk) and
// accumulates bin 64 k + l; the wave sums go to a.sums.
template <typename R, int ND>
__global__ __launch_bounds__(256) void k_relbin(RelbinArgs a) {
  using C = typename RelbinT<R>::C;
  const int lane = threadIdx.x & 63;
  const long long row = __builtin_amdgcn_readfirstlane(
      (int)((long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  if (row >= a.n) return;
  const RelRow rr = uniform_copy(a.rows[row]);
  const double flo = a.ev.wf_fmin;
  R dhr = 0, dhi = 0, hh = 0;
  C cur[ND], nxt[ND], right[ND];
  R* wa = a.win_a ? static_cast<R*>(a.win_a) + (size_t)row * a.win_kp : nullptr;
  // operand column d (nb + 1) + e holds r_d(e) (real part; imaginary part at
  // column win_k + d (nb + 1) + e)
  auto put_edges = [&](int e, const C* v) {
    if (!wa || e > a.nb) return;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const int k = d * (a.nb + 1) + e;
      wa[k] = v[d].x;
      wa[a.win_k + k] = v[d].y;
    }
  };
  edge_ratio<R, ND>(a, rr, flo, lane, cur);
  for (int base = 0; base < a.nb; base += 64) {
    put_edges(base + lane, cur);
    if (base + 64 <= a.nb) {
      edge_ratio<R, ND>(a, rr, flo, base + 64 + lane, nxt);
    } else {
#pragma unroll
      for (int d = 0; d < ND; ++d) nxt[d] = cmk<C>((R)0, (R)0);
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const C up = shfl_c(cur[d], (lane + 1) & 63);
      const C wrap = shfl_c(nxt[d], 0);
      right[d] = lane == 63 ? wrap : up;
    }
    const int bin = base + lane;
    if (bin < a.nb) {
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        const size_t o = ((size_t)d * a.nb + bin) * 6;
        const R A0r = RelbinT<R>::sval(a.s, o + 0), A0i = RelbinT<R>::sval(a.s, o + 1);
        const R A1r = RelbinT<R>::sval(a.s, o + 2), A1i = RelbinT<R>::sval(a.s, o + 3);
        const R B0 = RelbinT<R>::sval(a.s, o + 4), B1 = RelbinT<R>::sval(a.s, o + 5);
        const R Sr = right[d].x + cur[d].x, Si = right[d].y + cur[d].y;
        const R Dr = right[d].x - cur[d].x, Di = right[d].y - cur[d].y;
        // A conj(S) = (Ar Sr + Ai Si) + i (Ai Sr - Ar Si)
        dhr += A0r * Sr + A0i * Si + A1r * Dr + A1i * Di;
        dhi += A0i * Sr - A0r * Si + A1i * Dr - A1r * Di;
        hh += B0 * (Sr * Sr + Si * Si) + B1 * (Sr * Dr + Si * Di);
      }
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) cur[d] = nxt[d];
  }
  put_edges((a.nb + 63) / 64 * 64 + lane, cur);
  const double dh_re = reduce_wave((double)dhr);
  const double dh_im = reduce_wave((double)dhi);
  const double h2 = reduce_wave((double)hh);
  if (lane == 0) {
    a.sums[3 * row + 0] = dh_re;
    a.sums[3 * row + 1] = dh_im;
    a.sums[3 * row + 2] = h2;
  }
}

struct MbArgs {
  int n;
  double flo;
  double inv_sigma;   // tables hold L sigma and Q sigma^2 (float32 range)
  const double* fe;
  const double* basis;
  const double* amp;
  const void* lin;    // double2 or float2 [nd][n]
  const void* quad;   // double or float [nd][n]
  long long rows_n;
  const RelRow* rows;
  double* sums;
  void* win_a;        // time marginalisation: sum_d h_d L_d per point as GEMM operand rows
  int win_kp;
};

// One wavefront per row, lanes over the banded points: the detector
// response at each point (interferometer.py:303-368 with the frequencies
// given), <d|h> = conj(sum_p h_p L_p) and <h|h> = sum_p |h_p|^2 Q_p
// (multiband.py:755-765), float64 per-lane sums.
template <typename R, int ND>
__global__ __launch_bounds__(256) void k_mb(MbArgs a) {
  using C = typename RelbinT<R>::C;
  const int lane = threadIdx.x & 63;
  const long long row = __builtin_amdgcn_readfirstlane(
      (int)((long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  if (row >= a.rows_n) return;
  const RelRow rr = uniform_copy(a.rows[row]);
  double cabs2[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const double cr = rr.cre[d] * a.inv_sigma, ci = rr.cim[d] * a.inv_sigma;
    cabs2[d] = cr * cr + ci * ci;
  }
  double dr = 0.0, di = 0.0, hh = 0.0;
  R* wa = a.win_a ? static_cast<R*>(a.win_a) + (size_t)row * a.win_kp : nullptr;
  for (int p = lane; p < a.n; p += 64) {
    const double f = a.fe[p];
    if (!(f >= a.flo && f <= rr.fhi)) {
      if (wa) {
        wa[p] = (R)0;
        wa[a.n + p] = (R)0;
      }
      continue;
    }
    double b[kNB];
#pragma unroll
    for (int k = 0; k < kNB; ++k) b[k] = a.basis[(size_t)p * kNB + k];
    const double psi = tf2_phase_cycles(rr.c, b);
    const double A = a.amp[p];
    const R Ar = (R)A;
    double yr = 0.0, yi = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      double ph = fma(f, rr.dt[d], psi);
      ph -= rint(ph);
      const C w = expm2pi((R)ph);
      const C h = cmul(cmk<C>((R)(rr.cre[d] * a.inv_sigma) * Ar, (R)(rr.cim[d] * a.inv_sigma) * Ar), w);
      const C L = RelbinT<R>::load(a.lin, (size_t)d * a.n + p);
      yr += (double)(h.x * L.x - h.y * L.y);
      yi += (double)(h.x * L.y + h.y * L.x);
      hh += cabs2[d] * A * A * (double)RelbinT<R>::sval(a.quad, (size_t)d * a.n + p);
    }
    dr += yr;
    di += yi;
    if (wa) {  // multiband.py:792-798: the point's term of the full d h array
      wa[p] = (R)yr;
      wa[a.n + p] = (R)yi;
    }
  }
  dr = reduce_wave(dr);
  di = reduce_wave(di);
  hh = reduce_wave(hh);
  if (lane == 0) {
    a.sums[3 * row + 0] = dr;
    a.sums[3 * row + 1] = -di;  // d_inner_h = conj(sum h L)
    a.sums[3 * row + 2] = hh;
  }
}

__global__ void k_mb_tables(const double* fe, int n, double* basis, double* amp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double f = fe[i];
  if (f > 0.0) {
    tf2_basis(f, basis + (size_t)i * kNB);
    amp[i] = basis[(size_t)i * kNB + 1] * sqrt(basis[(size_t)i * kNB + 3]);  // f^-1 f^-1/6
  } else {
    for (int k = 0; k < kNB; ++k) basis[(size_t)i * kNB + k] = 0.0;
    amp[i] = 0.0;
  }
}

template <typename R>
int launch_mb_nd(const MbArgs& a, int nd, hipStream_t st) {
  const dim3 grid((unsigned)((a.rows_n + 3) / 4)), block(256);
  switch (nd) {
    case 1: hipLaunchKernelGGL((k_mb<R, 1>), grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_mb<R, 2>), grid, block, 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_mb<R, 3>), grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL((k_mb<R, 4>), grid, block, 0, st, a); break;
  }
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

// ------------------------------------------------------------ IFFT-FFT <h|h>
// multiband.py:613-646, 766-783.  For band b >= 1 and detector d the
// reference places sqrt(w) h at bins K_s .. K_e of an M_b / 2 + 1 half
// spectrum, takes its inverse real transform x (M_b samples), zero-pads it in
// front to 2 M_b samples, transforms back and sums
//   (4 / T^_b) sum_{k <= M_b} Re Ibc_b[k] |thbc_k|^2.
// The even bins thbc_{2j} are the half spectrum itself, so their terms are
// per-point coefficients folded into the quadratic table on the host.  The
// odd bins are the spectrum at half-integer frequencies,
//   Y_j = sum_n x_n exp(-i pi (2 j + 1) n / M_b),  j < M' = M_b / 2,
// computed with two M'-point complex transforms per (row, detector, band):
//   1. Zc_k = (X_k + conj X_{M'-k}) / 2 + i e^{i pi k / M'} (X_k - conj X_{M'-k}) / 2
//      (X the half spectrum), z = inverse DFT(Zc): z_p = M' (x_{2p} + i x_{2p+1});
//   2. u_p = z_p e^{-i pi p / M'}, U = DFT(u);
//   3. E_j = (U_j + conj U_{M'-1-j}) / 2, O_j = (U_j - conj U_{M'-1-j}) / 2i,
//      Y_j = (E_j + e^{-i pi (2j+1) / M_b} O_j) / M'
// and sum_j odd_j |Y_j|^2 with the M'^-2 folded into the odd-bin weights.
// Transforms of P <= 8192 points run in LDS (Stockham passes of radix 16 and
// one of radix 2, 4 or 8, each thread holding 16 points); longer ones
// (M' = R P) add a radix-R step over global scratch (four-step split).
// Every twiddle is an entry of the band's table exp(-2 pi i k / (4 M')).

// LDS index with one pad slot per 16 points (the radix-16 pass reads and
// writes at strides of 16 points)
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }

__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// exp(-+2 pi i m / 16) for compile-time m (INV: the + sign)
template <bool INV>
__device__ __forceinline__ double2 w16(int m) {
  constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173;
  constexpr double h = 0.70710678118654752440;
  double c, s;
  switch (m & 15) {
    case 0: c = 1.0; s = 0.0; break;
    case 1: c = c1; s = s1; break;
    case 2: c = h; s = h; break;
    case 3: c = s1; s = c1; break;
    case 4: c = 0.0; s = 1.0; break;
    case 5: c = -s1; s = c1; break;
    case 6: c = -h; s = h; break;
    case 7: c = -c1; s = s1; break;
    case 8: c = -1.0; s = 0.0; break;
    case 9: c = -c1; s = -s1; break;
    case 10: c = -h; s = -h; break;
    case 11: c = -s1; s = -c1; break;
    case 12: c = 0.0; s = -1.0; break;
    case 13: c = s1; s = -c1; break;
    case 14: c = h; s = -h; break;
    default: c = c1; s = -s1; break;
  }
  return make_double2(c, INV ? s : -s);
}

// In-register DFT of R = 2, 4, 8, 16 points: a_k = sum_n a_n exp(-+2 pi i n k / R)
template <int R, bool INV>
__device__ __forceinline__ void dft_reg(double2* a) {
  if constexpr (R == 2) {
    const double2 t = a[1];
    a[1] = make_double2(a[0].x - t.x, a[0].y - t.y);
    a[0] = make_double2(a[0].x + t.x, a[0].y + t.y);
  } else if constexpr (R == 4) {
    const double2 s0 = make_double2(a[0].x + a[2].x, a[0].y + a[2].y);
    const double2 d0 = make_double2(a[0].x - a[2].x, a[0].y - a[2].y);
    const double2 s1 = make_double2(a[1].x + a[3].x, a[1].y + a[3].y);
    double2 d1 = make_double2(a[1].x - a[3].x, a[1].y - a[3].y);
    d1 = INV ? make_double2(-d1.y, d1.x) : make_double2(d1.y, -d1.x);  // * (+-i)
    a[0] = make_double2(s0.x + s1.x, s0.y + s1.y);
    a[2] = make_double2(s0.x - s1.x, s0.y - s1.y);
    a[1] = make_double2(d0.x + d1.x, d0.y + d1.y);
    a[3] = make_double2(d0.x - d1.x, d0.y - d1.y);
  } else {
    // R = A B with A = R / 4, B = 4: n = n1 + A n2, k = B k1 + k2;
    // a_k = sum_n1 W_A^{n1 k1} W_R^{n1 k2} sum_n2 a_{n1 + A n2} W_4^{n2 k2}
    constexpr int A = R / 4;
    double2 t[R];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) {
      double2 v[4] = {a[n1], a[n1 + A], a[n1 + 2 * A], a[n1 + 3 * A]};
      dft_reg<4, INV>(v);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) t[n1 * 4 + k2] = cmul(v[k2], w16<INV>(n1 * k2 * (16 / R)));
    }
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      double2 v[A];
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) v[n1] = t[n1 * 4 + k2];
      dft_reg<A, INV>(v);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) a[4 * k1 + k2] = v[k1];
    }
  }
}

// One Stockham pass of radix R over P points in LDS (E points per thread,
// P / E threads): butterfly j reads x[j + q P / R], is twiddled by
// W_{Ns R}^{(j mod Ns) q}, transformed, and written to
// x[(j - j mod Ns) R + j mod Ns + q Ns].  All reads finish before any write.
template <int R, bool INV, int E>
__device__ __forceinline__ void fft_pass(double2* x, int P, int Ns, const double2* tab, int tstep) {
  constexpr int G = E / R;
  const int T = P / E, PR = P / R;
  double2 v[G][R];
  int dst[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int j = threadIdx.x + g * T;
    const int k = j & (Ns - 1);
#pragma unroll
    for (int q = 0; q < R; ++q) v[g][q] = x[lpad(j + q * PR)];
    if (Ns > 1) {
      // W_{Ns R}^{k q} = w^q with w = tab[k (P / (Ns R)) tstep]: one table load
      // per butterfly, the powers by multiplication (a few units in the last
      // place after R - 1 products)
      const double2 t = tab[k * (P / (Ns * R)) * tstep];
      const double2 w = INV ? conj2(t) : t;
      double2 wq = w;
#pragma unroll
      for (int q = 1; q < R; ++q) {
        v[g][q] = cmul(v[g][q], wq);
        if (q + 1 < R) wq = cmul(wq, w);
      }
    }
    dft_reg<R, INV>(v[g]);
    dst[g] = (j - k) * R + k;
  }
  __syncthreads();
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int q = 0; q < R; ++q) x[lpad(dst[g] + q * Ns)] = v[g][q];
  __syncthreads();
}

// P-point transform in LDS, natural order in and out, unnormalised;
// tab[k tstep] = exp(-2 pi i k / P).  E = 16: radix-16 passes; E = 8:
// radix-8 passes (twice the threads for the same points); the remaining
// bits in one pass of radix 2, 4 or 8.
template <bool INV, int E = 16>
__device__ void lds_fft(double2* x, int P, int logP, const double2* tab, int tstep) {
  int Ns = 1, bits = logP;
  if constexpr (E == 16) {
    for (; bits >= 4; bits -= 4, Ns <<= 4) fft_pass<16, INV, 16>(x, P, Ns, tab, tstep);
  } else {
    for (; bits >= 3; bits -= 3, Ns <<= 3) fft_pass<8, INV, 8>(x, P, Ns, tab, tstep);
  }
  if (bits == 3) fft_pass<8, INV, E>(x, P, Ns, tab, tstep);
  else if (bits == 2) fft_pass<4, INV, E>(x, P, Ns, tab, tstep);
  else if (bits == 1) fft_pass<2, INV, E>(x, P, Ns, tab, tstep);
}

struct MbFftArgs {
  const RelRow* rows;
  long long n;
  int nd;
  double flo;
  const double* fe;
  const double* basis;
  const double* amp;
  MbFftBand band;
  double* part;        // [n][nd] this band's odd-bin sums
  double2* scratch;    // split transforms: [gridDim.x][2 M']
};

// sqrt(w_p) h_d(f_p) at banded point p (the float64 detector response of k_mb)
__device__ __forceinline__ double2 mb_point_strain(const MbFftArgs& a, const RelRow& rr, int d,
                                                   int i) {
  const int p = a.band.p0 + i;
  const double f = a.fe[p];
  if (!(f >= a.flo && f <= rr.fhi)) return make_double2(0.0, 0.0);
  double b[kNB];
#pragma unroll
  for (int k = 0; k < kNB; ++k) b[k] = a.basis[(size_t)p * kNB + k];
  double ph = fma(f, rr.dt[d], tf2_phase_cycles(rr.c, b));
  ph -= rint(ph);
  const double s = a.amp[p] * a.band.sqw[i];
  return cmul(make_double2(rr.cre[d] * s, rr.cim[d] * s), expm2pi(ph));
}

// Zc_k from X_k and X_{M'-k} (step 1 above); e^{i pi k / M'} = conj tab[2k]
__device__ __forceinline__ double2 mb_pack(double2 xk, double2 xm, double2 t2k) {
  const double2 s = make_double2(0.5 * (xk.x + xm.x), 0.5 * (xk.y - xm.y));
  const double2 dd = make_double2(0.5 * (xk.x - xm.x), 0.5 * (xk.y + xm.y));
  const double2 e = cmul(conj2(t2k), dd);
  return make_double2(s.x - e.y, s.y + e.x);  // s + i e
}

// odd_j |Y_j|^2 from U_j and U_{M'-1-j} (step 3 above); tab[2j+1] = e^{-i pi (2j+1) / M_b}
__device__ __forceinline__ double mb_odd_term(double2 u, double2 ur, double2 t, double wgt) {
  const double2 e = make_double2(0.5 * (u.x + ur.x), 0.5 * (u.y - ur.y));
  const double2 dd = make_double2(0.5 * (u.x - ur.x), 0.5 * (u.y + ur.y));
  const double2 o = make_double2(dd.y, -dd.x);  // dd / i
  const double2 y = cmul(t, o);
  const double yr = e.x + y.x, yi = e.y + y.y;
  return wgt * (yr * yr + yi * yi);
}

// sum over the workgroup, written by thread 0
__device__ __forceinline__ void mb_block_sum(double s, double* red, double* dst) {
  s = reduce_wave(s);
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < nw; ++k) t += red[k];
    *dst = t;
  }
}

// One workgroup of M' / E threads per (row, detector): the whole band in
// LDS (M' = P).  E = 8 points per thread (up to 1024 threads): four
// wavefronts per SIMD where one 139 KB workgroup fills a CU.
template <int E>
__global__ __launch_bounds__(64 * 16 * 8 / E) void k_mb_fft_lds(MbFftArgs a) {
  extern __shared__ double2 xs[];
  __shared__ double red[16];
  const MbFftBand& bd = a.band;
  const long long prob = blockIdx.x;
  const long long r = prob / a.nd;
  const int d = (int)(prob - r * a.nd);
  const RelRow& rr = a.rows[r];
  if (!rr.valid) {  // the epilogue discards the row
    if (threadIdx.x == 0) a.part[prob] = 0.0;
    return;
  }
  const int P = bd.P, T = P / E;
  for (int i = threadIdx.x; i < P; i += T) xs[lpad(i)] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int i = threadIdx.x; i < bd.L; i += T) xs[lpad(bd.ks + i)] = mb_point_strain(a, rr, d, i);
  __syncthreads();
  double2 z[E];
#pragma unroll
  for (int g = 0; g < E; ++g) {
    const int k = threadIdx.x + g * T;
    const double2 xm = k ? xs[lpad(P - k)] : make_double2(0.0, 0.0);  // X_{M'} = 0
    z[g] = mb_pack(xs[lpad(k)], xm, bd.tab[2 * k]);
  }
  __syncthreads();
#pragma unroll
  for (int g = 0; g < E; ++g) xs[lpad(threadIdx.x + g * T)] = z[g];
  __syncthreads();
  lds_fft<true, E>(xs, P, bd.logP, bd.tab, 4);
#pragma unroll
  for (int g = 0; g < E; ++g) {
    const int p = threadIdx.x + g * T;
    xs[lpad(p)] = cmul(xs[lpad(p)], bd.tab[2 * p]);
  }
  __syncthreads();
  lds_fft<false, E>(xs, P, bd.logP, bd.tab, 4);
  const double* odd = bd.odd + (size_t)d * bd.mh;
  double s = 0.0;
#pragma unroll 4
  for (int g = 0; g < E; ++g) {
    const int j = threadIdx.x + g * T;
    s += mb_odd_term(xs[lpad(j)], xs[lpad(P - 1 - j)], bd.tab[2 * j + 1], odd[j]);
  }
  mb_block_sum(s, red, a.part + prob);
}

// Bands with M' = R P > P: each workgroup takes (row, detector) problems in
// turn with 2 M' points of global scratch; both transforms are R LDS
// transforms of P points after a radix-R step:
//   X_{k1 + R k2} = DFT_P[ W_{M'}^{n2 k1} sum_n1 x_{n2 + P n1} W_R^{n1 k1} ](k2).
__global__ __launch_bounds__(256) void k_mb_fft_split(MbFftArgs a) {
  extern __shared__ double2 xs[];
  __shared__ double red[4];
  const MbFftBand& bd = a.band;
  const int P = bd.P, T = P >> 4, R = bd.R, M = bd.mh;
  const int rstep = 4 * M / R, tstep = 4 * R;  // tab[m rstep] = W_R^m, tab[k tstep] = W_P^k
  double2* A = a.scratch + (size_t)blockIdx.x * 2 * M;
  double2* B = A + M;
  const long long total = a.n * a.nd;
  for (long long prob = blockIdx.x; prob < total; prob += gridDim.x) {
    const long long r = prob / a.nd;
    const int d = (int)(prob - r * a.nd);
    const RelRow& rr = a.rows[r];
    if (!rr.valid) {
      if (threadIdx.x == 0) a.part[prob] = 0.0;
      continue;
    }
    for (int i = threadIdx.x; i < M; i += T) A[i] = make_double2(0.0, 0.0);
    __threadfence();
    __syncthreads();
    for (int i = threadIdx.x; i < bd.L; i += T) A[bd.ks + i] = mb_point_strain(a, rr, d, i);
    __threadfence();
    __syncthreads();
    // inverse transform of Zc; outputs modulated into B
    for (int k1 = 0; k1 < R; ++k1) {
#pragma unroll 4
      for (int g = 0; g < 16; ++g) {
        const int n2 = threadIdx.x + g * T;
        double2 acc = make_double2(0.0, 0.0);
        for (int n1 = 0; n1 < R; ++n1) {
          const int k = n2 + P * n1;
          const double2 xm = k ? A[M - k] : make_double2(0.0, 0.0);
          const double2 zc = mb_pack(A[k], xm, bd.tab[2 * k]);
          const double2 w = conj2(bd.tab[((n1 * k1) & (R - 1)) * rstep]);
          acc.x += zc.x * w.x - zc.y * w.y;
          acc.y += zc.x * w.y + zc.y * w.x;
        }
        xs[lpad(n2)] = cmul(acc, conj2(bd.tab[4 * n2 * k1]));
      }
      __syncthreads();
      lds_fft<true>(xs, P, bd.logP, bd.tab, tstep);
#pragma unroll 4
      for (int g = 0; g < 16; ++g) {
        const int k2 = threadIdx.x + g * T;
        const int p = k1 + R * k2;
        B[p] = cmul(xs[lpad(k2)], bd.tab[2 * p]);
      }
      __syncthreads();
    }
    __threadfence();
    __syncthreads();
    // forward transform of B into A
    for (int k1 = 0; k1 < R; ++k1) {
#pragma unroll 4
      for (int g = 0; g < 16; ++g) {
        const int n2 = threadIdx.x + g * T;
        double2 acc = make_double2(0.0, 0.0);
        for (int n1 = 0; n1 < R; ++n1) {
          const double2 u = B[n2 + P * n1];
          const double2 w = bd.tab[((n1 * k1) & (R - 1)) * rstep];
          acc.x += u.x * w.x - u.y * w.y;
          acc.y += u.x * w.y + u.y * w.x;
        }
        xs[lpad(n2)] = cmul(acc, bd.tab[4 * n2 * k1]);
      }
      __syncthreads();
      lds_fft<false>(xs, P, bd.logP, bd.tab, tstep);
#pragma unroll 4
      for (int g = 0; g < 16; ++g) {
        const int k2 = threadIdx.x + g * T;
        A[k1 + R * k2] = xs[lpad(k2)];
      }
      __syncthreads();
    }
    __threadfence();
    __syncthreads();
    const double* odd = bd.odd + (size_t)d * M;
    double s = 0.0;
    for (int j = threadIdx.x; j < M; j += T)
      s += mb_odd_term(A[j], A[M - 1 - j], bd.tab[2 * j + 1], odd[j]);
    mb_block_sum(s, red, a.part + prob);
    __syncthreads();  // red and the scratch are reused by the next problem
  }
}

// <h|h> += the bands' odd-bin sums, in band and detector order
__global__ __launch_bounds__(128) void k_mb_fft_sum(const double* part, long long n, int nd,
                                                    int nbands, double* sums) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double t = 0.0;
  for (int b = 0; b < nbands; ++b)
    for (int d = 0; d < nd; ++d) t += part[((size_t)b * n + r) * nd + d];
  sums[3 * r + 2] += t;
}

// One thread per row: base.py:448-477 on the summed inner products.
__global__ __launch_bounds__(128) void k_relbin_epi(RelbinArgs a) {
  extern __shared__ double knots[];
  const EvDev& ev = a.ev;
  const bool dist = (ev.flags & GWL_DISTANCE_MARG) != 0;
  if (dist) stage_knots(ev, knots);
  __syncthreads();
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n) return;
  const double dh_re = a.sums[3 * r + 0], dh_im = a.sums[3 * r + 1], h2 = a.sums[3 * r + 2];
  if (a.aux) {
    a.aux[3 * r + 0] = dh_re;
    a.aux[3 * r + 1] = dh_im;
    a.aux[3 * r + 2] = h2;
  }
  if (!a.rows[r].valid) {
    a.out[r] = row_valid_sentinel();
    return;
  }
  double ll;
  if (dist) {
    const double hs = a.params[r * a.ld + GWL_P_LUMINOSITY_DISTANCE] / ev.dref;
    const double dd = (ev.flags & GWL_PHASE_MARG) ? hypot(dh_re, dh_im) * hs : dh_re * hs;
    const SplineY yy = spline_y(ev, knots + ev.sp.nx + 4, h2 * hs * hs);
    ll = spline_x(ev, knots, yy, dd);
  } else if (ev.flags & GWL_PHASE_MARG) {
    ll = ln_i0(hypot(dh_re, dh_im)) - h2 / 2.0;
  } else {
    ll = dh_re - h2 / 2.0;
  }
  a.out[r] = ll;
}

__global__ void k_relbin_basis(const double* fe, int n, double* basis) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tf2_basis(fe[i], basis + (size_t)i * kNB);
}

template <typename R>
int launch_relbin_nd(const RelbinArgs& a, int nd, hipStream_t st) {
  const dim3 grid((unsigned)((a.n + 3) / 4)), block(256);
  switch (nd) {
    case 1: hipLaunchKernelGGL((k_relbin<R, 1>), grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_relbin<R, 2>), grid, block, 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_relbin<R, 3>), grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL((k_relbin<R, 4>), grid, block, 0, st, a); break;
  }
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

int grow_ws(gwl_event* ev, void** ws, size_t* ws_bytes, long long rows) {
  const size_t need = relbin_ws_bytes(rows);
  if (need <= *ws_bytes) return GWL_OK;
  GWL_HIP_CHECK(hipSetDevice(ev->ctx->device));
  if (*ws) (void)hipFree(*ws);
  *ws = nullptr;
  *ws_bytes = 0;
  if (hipMalloc(ws, need) != hipSuccess) {
    set_error("hipMalloc failed for the row workspace");
    return GWL_E_NOMEM;
  }
  *ws_bytes = need;
  return GWL_OK;
}

int relbin_reserve(gwl_relbin* rb, long long rows) {
  return grow_ws(rb->ev, &rb->ws, &rb->ws_bytes, rows);
}

int grow_buf(gwl_event* ev, void** buf, size_t* bytes, size_t need) {
  if (need <= *bytes) return GWL_OK;
  GWL_HIP_CHECK(hipSetDevice(ev->ctx->device));
  if (*buf) (void)hipFree(*buf);
  *buf = nullptr;
  *bytes = 0;
  if (hipMalloc(buf, need) != hipSuccess) {
    set_error("hipMalloc failed for the time-window workspace");
    return GWL_E_NOMEM;
  }
  *bytes = need;
  return GWL_OK;
}

static_assert(sizeof(RelRow) % sizeof(int) == 0, "RelRow holds whole ints");

// The IFFT-FFT bands' odd-bin sums of rows [0, n) (row constants rows,
// their <h|h> in sums[3 r + 2]), added in band and detector order.
constexpr long long kMbFftChunk = 1LL << 26;  // rows per launch of the LDS kernel
constexpr int kMbFftSplitP = 4096;            // LDS transform of the split bands
constexpr int kMbFftE = 8;                    // points per thread of the whole-band kernel

size_t mb_fft_lds_bytes(int P) { return (size_t)(P + P / 16) * sizeof(double2); }

int mb_fft_launch(gwl_mb* mb, const RelRow* rows, long long n, double* sums, hipStream_t st) {
  if (mb->fft.empty() || n < 1) return GWL_OK;
  gwl_event* ev = mb->ev;
  const int nd = ev->dev.nd, nb = (int)mb->fft.size();
  int split_m = 0;
  for (const MbFftBand& b : mb->fft)
    if (b.R > 1) split_m = std::max(split_m, b.mh);
  const long long slots = std::min<long long>(n * nd, 2LL * ev->ctx->n_cu);
  const size_t part_bytes = ((size_t)nb * n * nd * 8 + 255) / 256 * 256;
  const size_t need = part_bytes + (split_m ? (size_t)slots * 2 * split_m * sizeof(double2) : 0);
  if (int rc = grow_buf(ev, &mb->fft_ws, &mb->fft_ws_bytes, need)) return rc;
  double* part = static_cast<double*>(mb->fft_ws);
  MbFftArgs f;
  f.nd = nd;
  f.flo = ev->dev.wf_fmin;
  f.fe = mb->fe;
  f.basis = mb->basis;
  f.amp = mb->amp;
  f.scratch = reinterpret_cast<double2*>(static_cast<char*>(mb->fft_ws) + part_bytes);
  for (int b = 0; b < nb; ++b) {
    f.band = mb->fft[b];
    double* pb = part + (size_t)b * n * nd;
    const int P = f.band.P;
    if (f.band.R == 1) {
      for (long long r0 = 0; r0 < n; r0 += kMbFftChunk) {
        f.rows = rows + r0;
        f.n = std::min(kMbFftChunk, n - r0);
        f.part = pb + r0 * nd;
        hipLaunchKernelGGL(k_mb_fft_lds<kMbFftE>, dim3((unsigned)(f.n * nd)), dim3(P / kMbFftE),
                           mb_fft_lds_bytes(P), st, f);
        GWL_HIP_CHECK(hipGetLastError());
      }
    } else {
      f.rows = rows;
      f.n = n;
      f.part = pb;
      hipLaunchKernelGGL(k_mb_fft_split, dim3((unsigned)slots), dim3(P / 16), mb_fft_lds_bytes(P),
                         st, f);
      GWL_HIP_CHECK(hipGetLastError());
    }
  }
  hipLaunchKernelGGL(k_mb_fft_sum, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, part, n,
                     nd, nb, sums);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}
constexpr int kValidStride = (int)(sizeof(RelRow) / sizeof(int));

const int* valid_of(const RelRow* rows) {
  return reinterpret_cast<const int*>(reinterpret_cast<const char*>(rows) +
                                      offsetof(RelRow, valid));
}

// Row tiles of at most kWinTileRows: row constants, the binned kernel (which
// also writes the operand rows), the window GEMM and epilogue.  ``kernel``
// launches the binned kernel for the tile described by ``a``.
template <typename Launch>
int eval_time_tiles(gwl_event* ev, const WinBasis& win, void** ws, size_t* ws_bytes, void** tws,
                    size_t* tws_bytes, RelbinArgs a, const double* params, long long n, int ld,
                    bool fp64, double* out, double* aux, hipStream_t st, Launch kernel,
                    double* sdh = nullptr, double* shh = nullptr) {
  const long long tile = std::min<long long>(n, kWinTileRows);
  if (int rc = grow_ws(ev, ws, ws_bytes, tile)) return rc;
  if (int rc = grow_buf(ev, tws, tws_bytes, win_ws_bytes(win, tile))) return rc;
  char* w = static_cast<char*>(*ws);
  a.rows = reinterpret_cast<RelRow*>(w);
  a.sums = reinterpret_cast<double*>(w + ((size_t)tile * sizeof(RelRow) + 255) / 256 * 256);
  void* opa = win_a(win, *tws);
  void* c = win_c(win, *tws, tile);
  const size_t lds = (ev->dev.flags & GWL_DISTANCE_MARG)
                         ? (size_t)(ev->dev.sp.nx + ev->dev.sp.ny + 8) * 8 : 0;
  for (long long r0 = 0; r0 < n; r0 += tile) {
    const long long nt = std::min(tile, n - r0);
    a.params = params + r0 * ld;
    a.n = nt;
    a.out = out ? out + r0 : nullptr;
    a.aux = aux ? aux + 3 * r0 : nullptr;
    if (int rc = win_clear(win, opa, fp64, nt, st)) return rc;
    const unsigned g128 = (unsigned)((nt + 127) / 128);
    hipLaunchKernelGGL(k_relbin_rows, dim3(g128), dim3(128), 0, st, a);
    GWL_HIP_CHECK(hipGetLastError());
    if (int rc = kernel(a, opa)) return rc;
    if (a.aux) {  // the inner products at the row's own time
      hipLaunchKernelGGL(k_relbin_epi, dim3(g128), dim3(128), lds, st, a);
      GWL_HIP_CHECK(hipGetLastError());
    }
    if (int rc = win_finish(ev, win, opa, fp64, c, nt, a.params, ld, a.sums, valid_of(a.rows),
                            kValidStride, a.out, st, sdh ? sdh + 2 * r0 * win.L : nullptr,
                            sdh ? shh + r0 : nullptr))
      return rc;
  }
  return GWL_OK;
}

// <h|h> of the rebuilt full waveform (relative.py:380-393) on the mask,
// base.py:629 (optimal_snr_squared of the rebuilt signal), one workgroup
// per row: sum over detectors and grid bins of w |r(f)|^2 with r(f) the
// edge ratios of the row's operand (float64, column d (nb + 1) + e, the
// imaginary part K columns on) interpolated linearly across the bin.
__global__ __launch_bounds__(256) void k_relbin_full_hh(const double* A, int Kp, int K, int nb,
                                                         int nd, long long n_x, long long e0,
                                                         double df, const double* fe,
                                                         const int* bin_of, const double* y,
                                                         double* hh) {
  __shared__ double red[4];
  const double* ar = A + (size_t)blockIdx.x * Kp;
  double s = 0.0;
  for (long long t = threadIdx.x; t < (long long)nd * n_x; t += blockDim.x) {
    const int d = (int)(t / n_x);
    const long long i = t - (long long)d * n_x;
    const int b = bin_of[i];
    const double wy = y[t];
    if (b < 0 || wy == 0.0) continue;
    const double f = (double)(e0 + i) * df;
    const double w = fe[b + 1] - fe[b];
    const double lo = (fe[b + 1] - f) / w, hi = (f - fe[b]) / w;
    const int k = d * (nb + 1) + b;
    const double rr = ar[k] * lo + ar[k + 1] * hi;
    const double ri = ar[K + k] * lo + ar[K + k + 1] * hi;
    s += wy * (rr * rr + ri * ri);
  }
  s = reduce_wave(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) hh[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace
}  // namespace gwl

extern "C" {

int gwl_relbin_create(gwl_event* ev, const gwl_relbin_desc* desc, gwl_relbin** out) {
  if (!ev || !desc || !out || desc->n_bins < 1 || !desc->bin_frequencies ||
      !desc->fiducial_response || !desc->summary) {
    set_error("gwl_relbin_create: invalid argument");
    return GWL_E_INVALID;
  }
  const gwl::EvDev& d = ev->dev;
  if (d.approx != GWL_TAYLORF2) {
    set_error("gwl_relbin_create: the event must be TaylorF2");
    return GWL_E_UNSUPPORTED;
  }
  const int nb = desc->n_bins, ne = nb + 1, nd = d.nd;
  for (int e = 0; e < ne; ++e) {
    const double f = desc->bin_frequencies[e];
    if (!(f > 0.0) || !std::isfinite(f) || (e > 0 && !(f > desc->bin_frequencies[e - 1]))) {
      set_error("gwl_relbin_create: bin frequencies must be positive and increasing");
      return GWL_E_INVALID;
    }
  }
  // host tables: g = f^-7/6 / h0, folded summary
  std::vector<double2> g64((size_t)nd * ne);
  std::vector<float2> g32((size_t)nd * ne);
  for (int k = 0; k < nd; ++k)
    for (int e = 0; e < ne; ++e) {
      const double hr = desc->fiducial_response[((size_t)k * ne + e) * 2 + 0];
      const double hi = desc->fiducial_response[((size_t)k * ne + e) * 2 + 1];
      const double m2 = hr * hr + hi * hi;
      if (!(m2 > 0.0) || !std::isfinite(m2)) {
        set_error("gwl_relbin_create: the fiducial response is zero or not finite at a bin edge");
        return GWL_E_INVALID;
      }
      const double s = std::pow(desc->bin_frequencies[e], -7.0 / 6.0) / m2;
      g64[(size_t)k * ne + e] = make_double2(hr * s, -hi * s);
      g32[(size_t)k * ne + e] = make_float2((float)(hr * s), (float)(-hi * s));
    }
  std::vector<double> s64((size_t)nd * nb * 6);
  std::vector<float> s32((size_t)nd * nb * 6);
  for (int k = 0; k < nd; ++k)
    for (int b = 0; b < nb; ++b) {
      const double* src = desc->summary + ((size_t)k * nb + b) * 8;
      const double w = desc->bin_frequencies[b + 1] - desc->bin_frequencies[b];
      double* o = &s64[((size_t)k * nb + b) * 6];
      o[0] = src[0] / 2.0;
      o[1] = src[1] / 2.0;
      o[2] = src[2] / w;
      o[3] = src[3] / w;
      o[4] = src[4] / 4.0;   // b0 and b1 are real (|h0|^2 weights); their imaginary parts are dropped
      o[5] = src[6] / w;
      for (int j = 0; j < 6; ++j) s32[((size_t)k * nb + b) * 6 + j] = (float)o[j];
    }
  gwl_relbin* rb = new (std::nothrow) gwl_relbin;
  if (!rb) { set_error("out of host memory"); return GWL_E_NOMEM; }
  rb->ev = ev;
  rb->nb = nb;
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t b_fe = al((size_t)ne * 8), b_basis = al((size_t)ne * gwl::kNB * 8);
  const size_t b_g64 = al(g64.size() * sizeof(double2)), b_g32 = al(g32.size() * sizeof(float2));
  const size_t b_s64 = al(s64.size() * 8), b_s32 = al(s32.size() * 4);
  rb->data_bytes = b_fe + b_basis + b_g64 + b_g32 + b_s64 + b_s32;
  if (hipSetDevice(ev->ctx->device) != hipSuccess || hipMalloc(&rb->data, rb->data_bytes) != hipSuccess) {
    delete rb;
    set_error("hipMalloc failed for relative-binning tables");
    return GWL_E_NOMEM;
  }
  char* q = static_cast<char*>(rb->data);
  rb->fe = reinterpret_cast<double*>(q); q += b_fe;
  rb->basis = reinterpret_cast<double*>(q); q += b_basis;
  rb->g64 = reinterpret_cast<double2*>(q); q += b_g64;
  rb->g32 = reinterpret_cast<float2*>(q); q += b_g32;
  rb->s64 = reinterpret_cast<double*>(q); q += b_s64;
  rb->s32 = reinterpret_cast<float*>(q);
  hipError_t err = hipMemcpy(rb->fe, desc->bin_frequencies, (size_t)ne * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(rb->g64, g64.data(), g64.size() * sizeof(double2), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(rb->g32, g32.data(), g32.size() * sizeof(float2), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(rb->s64, s64.data(), s64.size() * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(rb->s32, s32.data(), s32.size() * 4, hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    hipLaunchKernelGGL(gwl::k_relbin_basis, dim3((unsigned)((ne + 127) / 128)), dim3(128), 0, 0,
                       rb->fe, ne, rb->basis);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipDeviceSynchronize();
  if (err != hipSuccess) {
    (void)hipFree(rb->data);
    delete rb;
    return gwl::hip_fail(err, "gwl_relbin_create upload");
  }
  *out = rb;
  return GWL_OK;
}

int gwl_relbin_destroy(gwl_relbin* rb) {
  if (!rb) return GWL_OK;
  (void)hipSetDevice(rb->ev->ctx->device);
  if (rb->data) (void)hipFree(rb->data);
  if (rb->ws) (void)hipFree(rb->ws);
  if (rb->tws) (void)hipFree(rb->tws);
  gwl::win_basis_free(&rb->win);
  gwl::win_basis_free(&rb->up);
  if (rb->up_data) (void)hipFree(rb->up_data);
  delete rb;
  return GWL_OK;
}

int gwl_relbin_reserve(gwl_relbin* rb, int64_t rows) {
  if (!rb || rows < 1) { set_error("gwl_relbin_reserve: invalid argument"); return GWL_E_INVALID; }
  return gwl::relbin_reserve(rb, rows);
}

int gwl_relbin_device_bytes(gwl_relbin* rb, int64_t* bytes) {
  if (!rb || !bytes) { set_error("gwl_relbin_device_bytes: invalid argument"); return GWL_E_INVALID; }
  *bytes = (int64_t)(rb->data_bytes + rb->ws_bytes + rb->win.bytes + rb->tws_bytes);
  return GWL_OK;
}

static int relbin_eval_impl(gwl_relbin* rb, const double* params, int64_t n, int32_t ld,
                          int32_t precision, double* out, double* aux, double* sdh, double* shh,
                          void* stream, gwl::WinBasis* basis = nullptr) {
  if (rb && n == 0) return GWL_OK;  // an empty batch needs no buffers
  if (!rb || !params || (!out && !sdh) || (sdh && !shh) || n < 0 || ld < GWL_NPARAM ||
      (precision != GWL_FP32 && precision != GWL_FP64)) {
    set_error("gwl_relbin_eval: invalid argument");
    return GWL_E_INVALID;
  }
  if (n == 0) return GWL_OK;
  gwl_event* ev = rb->ev;
  if ((ev->dev.flags & GWL_DISTANCE_MARG) && !ev->has_spline) {
    set_error("gwl_relbin_eval: distance marginalisation requested but no spline attached");
    return GWL_E_INVALID;
  }
  const bool timem = (ev->dev.flags & GWL_TIME_MARG) != 0;
  if (sdh && !timem) {
    set_error("gwl_relbin_time_series: the event does not marginalise over time");
    return GWL_E_UNSUPPORTED;
  }
  if (timem && !rb->has_win) {
    set_error("gwl_relbin_eval: the event marginalises over time; call gwl_relbin_set_time first");
    return GWL_E_INVALID;
  }
  if (!timem)
    if (int rc = gwl::relbin_reserve(rb, n)) return rc;
  GWL_HIP_CHECK(hipSetDevice(ev->ctx->device));
  (void)hipGetLastError();
  gwl::RelbinArgs a;
  std::memset(&a, 0, sizeof a);
  a.antenna_time = NAN;
  a.ev = ev->dev;
  a.nb = rb->nb;
  a.fe = rb->fe;
  a.basis = rb->basis;
  a.g = precision == GWL_FP64 ? (const void*)rb->g64 : (const void*)rb->g32;
  a.s = precision == GWL_FP64 ? (const void*)rb->s64 : (const void*)rb->s32;
  a.params = params;
  a.ld = ld;
  a.n = n;
  char* w = static_cast<char*>(rb->ws);
  a.rows = reinterpret_cast<gwl::RelRow*>(w);
  a.sums = reinterpret_cast<double*>(w + ((size_t)n * sizeof(gwl::RelRow) + 255) / 256 * 256);
  a.out = out;
  a.aux = aux;
  const size_t lds = (ev->dev.flags & GWL_DISTANCE_MARG)
                         ? (size_t)(ev->dev.sp.nx + ev->dev.sp.ny + 8) * 8 : 0;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (int rc = gwl::timing_begin(ev, st)) return rc;
  if (timem) {
    const bool fp64 = precision == GWL_FP64;
    gwl::WinBasis& win = basis ? *basis : rb->win;
    a.win_k = win.K;
    a.win_kp = win.Kp;
    const int nd = ev->dev.nd;
    const int rc = gwl::eval_time_tiles(
        ev, win, &rb->ws, &rb->ws_bytes, &rb->tws, &rb->tws_bytes, a, params, n, ld, fp64, out,
        aux, st, [&](gwl::RelbinArgs t, void* opa) {
          t.win_a = opa;
          return fp64 ? gwl::launch_relbin_nd<double>(t, nd, st)
                      : gwl::launch_relbin_nd<float>(t, nd, st);
        },
        sdh, shh);
    if (rc != GWL_OK) return rc;
    return gwl::timing_end(ev, st);
  }
  const unsigned g128 = (unsigned)((n + 127) / 128);
  hipLaunchKernelGGL(gwl::k_relbin_rows, dim3(g128), dim3(128), 0, st, a);
  GWL_HIP_CHECK(hipGetLastError());
  const int rc = precision == GWL_FP64 ? gwl::launch_relbin_nd<double>(a, ev->dev.nd, st)
                                       : gwl::launch_relbin_nd<float>(a, ev->dev.nd, st);
  if (rc != GWL_OK) return rc;
  hipLaunchKernelGGL(gwl::k_relbin_epi, dim3(g128), dim3(128), lds, st, a);
  GWL_HIP_CHECK(hipGetLastError());
  return gwl::timing_end(ev, st);
}

int gwl_relbin_eval(gwl_relbin* rb, const double* params, int64_t n, int32_t ld, int32_t precision,
                    double* out, double* aux, void* stream) {
  return relbin_eval_impl(rb, params, n, ld, precision, out, aux, nullptr, nullptr, stream);
}

int gwl_relbin_time_series(gwl_relbin* rb, const double* params, int64_t n, int32_t ld,
                           int32_t precision, double* d_inner_h, double* h_inner_h, void* stream) {
  if (rb && n > 0 && (!d_inner_h || !h_inner_h)) {
    set_error("gwl_relbin_time_series: invalid argument");
    return GWL_E_INVALID;
  }
  return relbin_eval_impl(rb, params, n, ld, precision, nullptr, nullptr, d_inner_h, h_inner_h,
                          stream);
}

int gwl_relbin_set_upsampled(gwl_relbin* rb, const int64_t* bin_index, const double* integrand,
                             const double* weight, int64_t n_fft) {
  if (!rb || !bin_index || !integrand || !weight) {
    set_error("gwl_relbin_set_upsampled: invalid argument");
    return GWL_E_INVALID;
  }
  gwl_event* ev = rb->ev;
  const gwl::EvDev& d = ev->dev;
  const int nb = rb->nb;
  for (int e = 0; e <= nb; ++e)
    if (bin_index[e] < 0 || bin_index[e] > d.N || (e > 0 && bin_index[e] <= bin_index[e - 1])) {
      set_error("gwl_relbin_set_upsampled: bin indices must increase within the grid");
      return GWL_E_INVALID;
    }
  if (n_fft <= d.N) {
    set_error("gwl_relbin_set_upsampled: n_fft must exceed the event's frequency bins");
    return GWL_E_INVALID;
  }
  const long long e0 = bin_index[0], n_x = bin_index[nb] - e0 + 1;
  std::vector<int> bin_of((size_t)n_x, -1);
  for (int b = 0; b < nb; ++b) {
    const long long hi = b == nb - 1 ? bin_index[b + 1] : bin_index[b + 1] - 1;
    for (long long n = bin_index[b]; n <= hi; ++n) bin_of[(size_t)(n - e0)] = b;
  }
  const size_t bx = (size_t)d.nd * n_x * sizeof(double2), by = (size_t)d.nd * n_x * 8;
  const size_t bb = (size_t)n_x * 4, be = (size_t)(nb + 1) * 8;
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  GWL_HIP_CHECK(hipSetDevice(ev->ctx->device));
  void* mem = nullptr;
  if (hipMalloc(&mem, al(bx) + al(by) + al(bb) + be) != hipSuccess) {
    set_error("hipMalloc failed for the upsampled time set-up");
    return GWL_E_NOMEM;
  }
  char* m = static_cast<char*>(mem);
  hipError_t err = hipMemcpy(m, integrand, bx, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(m + al(bx), weight, by, hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = hipMemcpy(m + al(bx) + al(by), bin_of.data(), bb, hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = hipMemcpy(m + al(bx) + al(by) + al(bb), bin_index, be, hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    (void)hipFree(mem);
    return gwl::hip_fail(err, "gwl_relbin_set_upsampled upload");
  }
  gwl::win_basis_free(&rb->up);
  if (rb->up_data) (void)hipFree(rb->up_data);
  rb->up_data = mem;
  rb->up_x = reinterpret_cast<double2*>(m);
  rb->up_y = reinterpret_cast<double*>(m + al(bx));
  rb->up_bin = reinterpret_cast<int*>(m + al(bx) + al(by));
  rb->up_edges = reinterpret_cast<long long*>(m + al(bx) + al(by) + al(bb));
  rb->up_nx = n_x;
  rb->up_nfft = n_fft;
  rb->up_edges_first = e0;
  rb->has_up = true;
  return GWL_OK;
}

int gwl_relbin_upsampled_series(gwl_relbin* rb, const double* params, int64_t n, int32_t ld,
                                int64_t j_first, int32_t n_window, double* d_inner_h,
                                double* h_inner_h, void* stream) {
  if (rb && n == 0) return GWL_OK;
  if (!rb || !params || !d_inner_h || !h_inner_h || n < 0 || ld < GWL_NPARAM || j_first < 0 ||
      n_window < 1) {
    set_error("gwl_relbin_upsampled_series: invalid argument");
    return GWL_E_INVALID;
  }
  if (!rb->has_up || !rb->has_win) {
    set_error("gwl_relbin_upsampled_series: call gwl_relbin_set_time and "
              "gwl_relbin_set_upsampled first");
    return GWL_E_INVALID;
  }
  if (n > gwl::kWinTileRows) {
    set_error("gwl_relbin_upsampled_series: at most 8192 rows per call");
    return GWL_E_INVALID;
  }
  if (j_first + (long long)n_window > 2 * rb->up_nfft) {
    set_error("gwl_relbin_upsampled_series: the window runs past one wrap of the transform");
    return GWL_E_INVALID;
  }
  gwl_event* ev = rb->ev;
  if (!rb->up.data || rb->up.j_first != j_first || rb->up.L != n_window) {
    gwl::win_basis_free(&rb->up);
    gwl::WinBasis wb;
    wb.L = n_window;
    wb.n_fft = rb->up_nfft;
    wb.j_first = j_first;
    wb.start = ev->dev.start;
    wb.offset = 0;
    wb.t_step = 1.0 / ((double)rb->up_nfft * ev->dev.df);
    wb.delta_tc = wb.t_step;
    GWL_HIP_CHECK(hipSetDevice(ev->ctx->device));
    if (int rc = gwl::win_basis_relbin(ev, ev->dev.nd, rb->nb, rb->up_edges, rb->fe, rb->up_x,
                                       rb->up_nx, &wb))
      return rc;
    rb->up = wb;
  }
  if (int rc = relbin_eval_impl(rb, params, n, ld, GWL_FP64, nullptr, nullptr, d_inner_h,
                                h_inner_h, stream, &rb->up))
    return rc;
  // the exact <h|h> of the rebuilt waveform replaces the binned one
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(gwl::k_relbin_full_hh, dim3((unsigned)n), dim3(256), 0, st,
                     static_cast<const double*>(gwl::win_a(rb->up, rb->tws)), rb->up.Kp,
                     rb->up.K, rb->nb, ev->dev.nd, rb->up_nx, rb->up_edges_first, ev->dev.df,
                     rb->fe, rb->up_bin, rb->up_y, h_inner_h);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

int gwl_relbin_set_time(gwl_relbin* rb, const int64_t* bin_index, const double* integrand) {
  if (!rb || !bin_index || !integrand) {
    set_error("gwl_relbin_set_time: invalid argument");
    return GWL_E_INVALID;
  }
  gwl_event* ev = rb->ev;
  const gwl::EvDev& d = ev->dev;
  if (!(d.flags & GWL_TIME_MARG)) {
    set_error("gwl_relbin_set_time: the event does not marginalise over time");
    return GWL_E_UNSUPPORTED;
  }
  const int nb = rb->nb;
  for (int e = 0; e <= nb; ++e)
    if (bin_index[e] < 0 || bin_index[e] > d.N || (e > 0 && bin_index[e] <= bin_index[e - 1])) {
      set_error("gwl_relbin_set_time: bin indices must increase within the grid");
      return GWL_E_INVALID;
    }
  const long long n_x = bin_index[nb] - bin_index[0] + 1;
  GWL_HIP_CHECK(hipSetDevice(ev->ctx->device));
  long long* de = nullptr;
  double2* dx = nullptr;
  if (hipMalloc((void**)&de, (size_t)(nb + 1) * 8) != hipSuccess ||
      hipMalloc((void**)&dx, (size_t)d.nd * n_x * sizeof(double2)) != hipSuccess) {
    if (de) (void)hipFree(de);
    set_error("hipMalloc failed for the time-window set-up");
    return GWL_E_NOMEM;
  }
  hipError_t err = hipMemcpy(de, bin_index, (size_t)(nb + 1) * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = hipMemcpy(dx, integrand, (size_t)d.nd * n_x * sizeof(double2), hipMemcpyHostToDevice);
  int rc = GWL_OK;
  gwl::WinBasis wb;
  if (err != hipSuccess) {
    rc = gwl::hip_fail(err, "gwl_relbin_set_time upload");
  } else {
    // the event's window and time grid (base.py:1027-1035)
    if (d.L > 16384) {
      (void)hipFree(de);
      (void)hipFree(dx);
      set_error("gwl_relbin_set_time: the binned window transform covers at most 16384 "
                "samples of time window");
      return GWL_E_UNSUPPORTED;
    }
    wb.L = d.L;
    wb.j_first = d.jlo;
    wb.start = d.start;
    wb.offset = 1;
    wb.t_step = d.tstep;
    wb.delta_tc = d.tdelta;
    rc = gwl::win_basis_relbin(ev, d.nd, nb, de, rb->fe, dx, n_x, &wb);
  }
  (void)hipFree(de);
  (void)hipFree(dx);
  if (rc != GWL_OK) return rc;
  gwl::win_basis_free(&rb->win);
  rb->win = wb;
  rb->has_win = true;
  return GWL_OK;
}

}  // extern "C"

extern "C" {

int gwl_mb_create(gwl_event* ev, int32_t n_points, const double* frequencies, const double* linear,
                  const double* quadratic, gwl_mb** out) {
  if (!ev || n_points < 1 || !frequencies || !linear || !quadratic || !out) {
    set_error("gwl_mb_create: invalid argument");
    return GWL_E_INVALID;
  }
  const gwl::EvDev& d = ev->dev;
  if (d.approx != GWL_TAYLORF2) {
    set_error("gwl_mb_create: the event must be TaylorF2");
    return GWL_E_UNSUPPORTED;
  }
  const int n = n_points, nd = d.nd;
  // scaled by the event's sigma (|h| / sigma ~ 1) so that float32 holds them
  const double sg = d.sigma;
  std::vector<double2> l64((size_t)nd * n);
  std::vector<float2> l32((size_t)nd * n);
  std::vector<double> q64((size_t)nd * n);
  std::vector<float> q32((size_t)nd * n);
  for (size_t i = 0; i < (size_t)nd * n; ++i) {
    l64[i] = make_double2(linear[2 * i] * sg, linear[2 * i + 1] * sg);
    l32[i] = make_float2((float)l64[i].x, (float)l64[i].y);
    q64[i] = quadratic[i] * sg * sg;
    q32[i] = (float)q64[i];
  }
  gwl_mb* mb = new (std::nothrow) gwl_mb;
  if (!mb) { set_error("out of host memory"); return GWL_E_NOMEM; }
  mb->ev = ev;
  mb->n = n;
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t b_fe = al((size_t)n * 8), b_basis = al((size_t)n * gwl::kNB * 8), b_amp = al((size_t)n * 8);
  const size_t b_l64 = al(l64.size() * 16), b_l32 = al(l32.size() * 8);
  const size_t b_q64 = al((size_t)nd * n * 8), b_q32 = al(q32.size() * 4);
  mb->data_bytes = b_fe + b_basis + b_amp + b_l64 + b_l32 + b_q64 + b_q32;
  if (hipSetDevice(ev->ctx->device) != hipSuccess || hipMalloc(&mb->data, mb->data_bytes) != hipSuccess) {
    delete mb;
    set_error("hipMalloc failed for multiband tables");
    return GWL_E_NOMEM;
  }
  char* q = static_cast<char*>(mb->data);
  mb->fe = reinterpret_cast<double*>(q); q += b_fe;
  mb->basis = reinterpret_cast<double*>(q); q += b_basis;
  mb->amp = reinterpret_cast<double*>(q); q += b_amp;
  mb->lin64 = reinterpret_cast<double2*>(q); q += b_l64;
  mb->lin32 = reinterpret_cast<float2*>(q); q += b_l32;
  mb->quad64 = reinterpret_cast<double*>(q); q += b_q64;
  mb->quad32 = reinterpret_cast<float*>(q);
  hipError_t err = hipMemcpy(mb->fe, frequencies, (size_t)n * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(mb->lin64, l64.data(), l64.size() * 16, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(mb->lin32, l32.data(), l32.size() * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(mb->quad64, q64.data(), q64.size() * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(mb->quad32, q32.data(), q32.size() * 4, hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    hipLaunchKernelGGL(gwl::k_mb_tables, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, 0,
                       mb->fe, n, mb->basis, mb->amp);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipDeviceSynchronize();
  if (err != hipSuccess) {
    (void)hipFree(mb->data);
    delete mb;
    return gwl::hip_fail(err, "gwl_mb_create upload");
  }
  *out = mb;
  return GWL_OK;
}

int gwl_mb_destroy(gwl_mb* mb) {
  if (!mb) return GWL_OK;
  (void)hipSetDevice(mb->ev->ctx->device);
  if (mb->data) (void)hipFree(mb->data);
  if (mb->ws) (void)hipFree(mb->ws);
  if (mb->tws) (void)hipFree(mb->tws);
  if (mb->fft_data) (void)hipFree(mb->fft_data);
  if (mb->fft_ws) (void)hipFree(mb->fft_ws);
  gwl::win_basis_free(&mb->win);
  delete mb;
  return GWL_OK;
}

int gwl_mb_reserve(gwl_mb* mb, int64_t rows) {
  if (!mb || rows < 1) { set_error("gwl_mb_reserve: invalid argument"); return GWL_E_INVALID; }
  return gwl::grow_ws(mb->ev, &mb->ws, &mb->ws_bytes, rows);
}

int gwl_mb_device_bytes(gwl_mb* mb, int64_t* bytes) {
  if (!mb || !bytes) { set_error("gwl_mb_device_bytes: invalid argument"); return GWL_E_INVALID; }
  size_t fft = mb->fft_ws_bytes;
  for (const gwl::MbFftBand& b : mb->fft)
    fft += (size_t)b.mh * (4 * 16 + mb->ev->dev.nd * 8) + (size_t)b.L * 8;
  *bytes = (int64_t)(mb->data_bytes + mb->ws_bytes + mb->win.bytes + mb->tws_bytes + fft);
  return GWL_OK;
}

static int mb_eval_impl(gwl_mb* mb, const double* params, int64_t n, int32_t ld,
                          int32_t precision, double* out, double* aux, double* sdh, double* shh,
                          void* stream) {
  if (mb && n == 0) return GWL_OK;  // an empty batch needs no buffers
  if (!mb || !params || (!out && !sdh) || (sdh && !shh) || n < 0 || ld < GWL_NPARAM ||
      (precision != GWL_FP32 && precision != GWL_FP64)) {
    set_error("gwl_mb_eval: invalid argument");
    return GWL_E_INVALID;
  }
  if (n == 0) return GWL_OK;
  gwl_event* ev = mb->ev;
  if ((ev->dev.flags & GWL_DISTANCE_MARG) && !ev->has_spline) {
    set_error("gwl_mb_eval: distance marginalisation requested but no spline attached");
    return GWL_E_INVALID;
  }
  if (n > (int64_t)1 << 30) {
    set_error("gwl_mb_eval: too many rows in one call");
    return GWL_E_INVALID;
  }
  const bool timem = (ev->dev.flags & GWL_TIME_MARG) != 0;
  if (sdh && !timem) {
    set_error("gwl_mb_time_series: the event does not marginalise over time");
    return GWL_E_UNSUPPORTED;
  }
  if (timem && !mb->has_win) {
    set_error("gwl_mb_eval: the event marginalises over time; call gwl_mb_set_time first");
    return GWL_E_INVALID;
  }
  if (!timem)
    if (int rc = gwl::grow_ws(ev, &mb->ws, &mb->ws_bytes, n)) return rc;
  GWL_HIP_CHECK(hipSetDevice(ev->ctx->device));
  (void)hipGetLastError();
  gwl::RelbinArgs a;
  std::memset(&a, 0, sizeof a);
  a.ev = ev->dev;
  a.params = params;
  a.ld = ld;
  a.n = n;
  char* w = static_cast<char*>(mb->ws);
  a.rows = reinterpret_cast<gwl::RelRow*>(w);
  a.sums = reinterpret_cast<double*>(w + ((size_t)n * sizeof(gwl::RelRow) + 255) / 256 * 256);
  a.out = out;
  a.aux = aux;
  a.antenna_time = NAN;
  gwl::MbArgs m;
  std::memset(&m, 0, sizeof m);
  m.n = mb->n;
  m.flo = ev->dev.wf_fmin;
  m.inv_sigma = 1.0 / ev->dev.sigma;
  m.fe = mb->fe;
  m.basis = mb->basis;
  m.amp = mb->amp;
  m.lin = precision == GWL_FP64 ? (const void*)mb->lin64 : (const void*)mb->lin32;
  m.quad = precision == GWL_FP64 ? (const void*)mb->quad64 : (const void*)mb->quad32;
  m.rows_n = n;
  m.rows = a.rows;
  m.sums = a.sums;
  const size_t lds = (ev->dev.flags & GWL_DISTANCE_MARG)
                         ? (size_t)(ev->dev.sp.nx + ev->dev.sp.ny + 8) * 8 : 0;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (int rc = gwl::timing_begin(ev, st)) return rc;
  if (timem) {
    const bool fp64 = precision == GWL_FP64;
    a.antenna_time = mb->win.antenna_time;
    const int nd = ev->dev.nd;
    const int rc = gwl::eval_time_tiles(
        ev, mb->win, &mb->ws, &mb->ws_bytes, &mb->tws, &mb->tws_bytes, a, params, n, ld, fp64, out,
        aux, st, [&](gwl::RelbinArgs t, void* opa) {
          gwl::MbArgs mt = m;
          mt.rows_n = t.n;
          mt.rows = t.rows;
          mt.sums = t.sums;
          mt.win_a = opa;
          mt.win_kp = mb->win.Kp;
          const int rc = fp64 ? gwl::launch_mb_nd<double>(mt, nd, st)
                              : gwl::launch_mb_nd<float>(mt, nd, st);
          return rc != GWL_OK ? rc : gwl::mb_fft_launch(mb, t.rows, t.n, t.sums, st);
        },
        sdh, shh);
    if (rc != GWL_OK) return rc;
    return gwl::timing_end(ev, st);
  }
  const unsigned g128 = (unsigned)((n + 127) / 128);
  hipLaunchKernelGGL(gwl::k_relbin_rows, dim3(g128), dim3(128), 0, st, a);
  GWL_HIP_CHECK(hipGetLastError());
  const int rc = precision == GWL_FP64 ? gwl::launch_mb_nd<double>(m, ev->dev.nd, st)
                                       : gwl::launch_mb_nd<float>(m, ev->dev.nd, st);
  if (rc != GWL_OK) return rc;
  if (int rf = gwl::mb_fft_launch(mb, a.rows, n, a.sums, st)) return rf;
  hipLaunchKernelGGL(gwl::k_relbin_epi, dim3(g128), dim3(128), lds, st, a);
  GWL_HIP_CHECK(hipGetLastError());
  return gwl::timing_end(ev, st);
}

int gwl_mb_eval(gwl_mb* mb, const double* params, int64_t n, int32_t ld, int32_t precision,
                double* out, double* aux, void* stream) {
  return mb_eval_impl(mb, params, n, ld, precision, out, aux, nullptr, nullptr, stream);
}

int gwl_mb_time_series(gwl_mb* mb, const double* params, int64_t n, int32_t ld, int32_t precision,
                       double* d_inner_h, double* h_inner_h, void* stream) {
  if (mb && n > 0 && (!d_inner_h || !h_inner_h)) {
    set_error("gwl_mb_time_series: invalid argument");
    return GWL_E_INVALID;
  }
  return mb_eval_impl(mb, params, n, ld, precision, nullptr, nullptr, d_inner_h, h_inner_h, stream);
}

int gwl_mb_set_time(gwl_mb* mb, const gwl_mb_time_desc* td) {
  if (!mb || !td || !td->grid_index || td->n_fft < 2 || td->j_first < 0 || td->n_window < 1 ||
      td->j_first + td->n_window > td->n_fft || !(td->t_step > 0.0)) {
    set_error("gwl_mb_set_time: invalid argument");
    return GWL_E_INVALID;
  }
  gwl_event* ev = mb->ev;
  if (!(ev->dev.flags & GWL_TIME_MARG)) {
    set_error("gwl_mb_set_time: the event does not marginalise over time");
    return GWL_E_UNSUPPORTED;
  }
  for (int p = 0; p < mb->n; ++p)
    if (td->grid_index[p] >= td->n_fft) {
      set_error("gwl_mb_set_time: a grid index lies beyond the transform length");
      return GWL_E_INVALID;
    }
  GWL_HIP_CHECK(hipSetDevice(ev->ctx->device));
  long long* di = nullptr;
  if (hipMalloc((void**)&di, (size_t)mb->n * 8) != hipSuccess) {
    set_error("hipMalloc failed for the time-window set-up");
    return GWL_E_NOMEM;
  }
  hipError_t err = hipMemcpy(di, td->grid_index, (size_t)mb->n * 8, hipMemcpyHostToDevice);
  int rc = GWL_OK;
  gwl::WinBasis wb;
  if (err != hipSuccess) {
    rc = gwl::hip_fail(err, "gwl_mb_set_time upload");
  } else {
    // multiband.py:714-726: times start + j T_0 / N, prior weight spacing T_0 / N
    wb.L = td->n_window;
    wb.j_first = td->j_first;
    wb.start = ev->dev.start;
    wb.offset = 0;
    wb.t_step = td->t_step;
    wb.delta_tc = td->t_step;
    wb.antenna_time = td->antenna_time;
    rc = gwl::win_basis_points(ev, mb->n, di, td->n_fft, &wb);
  }
  (void)hipFree(di);
  if (rc != GWL_OK) return rc;
  gwl::win_basis_free(&mb->win);
  mb->win = wb;
  mb->has_win = true;
  return GWL_OK;
}

int gwl_mb_set_fft(gwl_mb* mb, int32_t n_bands, const gwl_mb_fft_band* bands) {
  if (!mb || n_bands < 0 || (n_bands > 0 && !bands)) {
    set_error("gwl_mb_set_fft: invalid argument");
    return GWL_E_INVALID;
  }
  gwl_event* ev = mb->ev;
  const int nd = ev->dev.nd;
  for (int b = 0; b < n_bands; ++b) {
    const gwl_mb_fft_band& d = bands[b];
    const int m = d.half_length;
    if (m < 16 || m > (1 << 18) || (m & (m - 1)) || d.first_point < 0 || d.n_points < 1 ||
        d.first_point + (int64_t)d.n_points > mb->n || d.first_bin < 1 ||
        d.first_bin + (int64_t)d.n_points > m || !d.sqrt_window || !d.odd_weights) {
      set_error("gwl_mb_set_fft: invalid band " + std::to_string(b));
      return GWL_E_INVALID;
    }
  }
  GWL_HIP_CHECK(hipSetDevice(ev->ctx->device));
  // layout: per band the table, the odd-bin weights and the window roots
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  size_t bytes = 0;
  for (int b = 0; b < n_bands; ++b)
    bytes += al((size_t)4 * bands[b].half_length * 16) + al((size_t)nd * bands[b].half_length * 8) +
             al((size_t)bands[b].n_points * 8);
  void* data = nullptr;
  if (bytes && hipMalloc(&data, bytes) != hipSuccess) {
    set_error("hipMalloc failed for the IFFT-FFT tables");
    return GWL_E_NOMEM;
  }
  std::vector<gwl::MbFftBand> out;
  hipError_t err = hipSuccess;
  char* q = static_cast<char*>(data);
  const size_t lds_max = ev->ctx->lds_per_block;
  for (int b = 0; b < n_bands && err == hipSuccess; ++b) {
    const gwl_mb_fft_band& d = bands[b];
    const int m = d.half_length;
    gwl::MbFftBand fb;
    fb.mh = m;
    fb.P = m;
    if (m > 8192 || gwl::mb_fft_lds_bytes(m) > lds_max) {
      fb.P = gwl::kMbFftSplitP;
      while (gwl::mb_fft_lds_bytes(fb.P) > lds_max) fb.P /= 2;
      fb.P = std::min(fb.P, m);
    }
    fb.R = m / fb.P;
    fb.logP = 0;
    while ((1 << fb.logP) < fb.P) ++fb.logP;
    fb.p0 = d.first_point;
    fb.L = d.n_points;
    fb.ks = d.first_bin;
    // exp(-2 pi i k / (4 M')), in extended precision
    std::vector<double2> tab((size_t)4 * m);
    const long double step = 2.0L * 3.141592653589793238462643383279502884L / (4.0L * m);
    for (size_t k = 0; k < tab.size(); ++k) {
      const long double t = step * (long double)k;
      tab[k] = make_double2((double)std::cos(t), (double)-std::sin(t));
    }
    std::vector<double> odd((size_t)nd * m);
    const double inv = 1.0 / ((double)m * (double)m);
    for (size_t i = 0; i < odd.size(); ++i) odd[i] = d.odd_weights[i] * inv;
    double2* t_dev = reinterpret_cast<double2*>(q); q += al((size_t)4 * m * 16);
    double* o_dev = reinterpret_cast<double*>(q); q += al((size_t)nd * m * 8);
    double* w_dev = reinterpret_cast<double*>(q); q += al((size_t)d.n_points * 8);
    err = hipMemcpy(t_dev, tab.data(), tab.size() * 16, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(o_dev, odd.data(), odd.size() * 8, hipMemcpyHostToDevice);
    if (err == hipSuccess)
      err = hipMemcpy(w_dev, d.sqrt_window, (size_t)d.n_points * 8, hipMemcpyHostToDevice);
    fb.tab = t_dev;
    fb.odd = o_dev;
    fb.sqw = w_dev;
    out.push_back(fb);
  }
  if (err != hipSuccess) {
    (void)hipFree(data);
    return gwl::hip_fail(err, "gwl_mb_set_fft upload");
  }
  if (mb->fft_data) (void)hipFree(mb->fft_data);
  mb->fft_data = data;
  mb->fft = std::move(out);
  return GWL_OK;
}

}  // extern "C"
