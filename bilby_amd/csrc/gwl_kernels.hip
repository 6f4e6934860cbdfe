// Likelihood kernels for gfx950.
//
//  k_reduce : fused source model + detector projection + noise-weighted
//             inner products <d|h>, <h|h> (no time marginalisation).
//             bilby/gw/likelihood/base.py:260-354 (calculate_snrs) with
//             interferometer.py:303-368 and gw/utils.py:118-138.
//  k_time   : fused source model + projection + d_inner_h_array
//             = 4/T FFT_N(h conj(d) / S) at the time-prior window only
//             (base.py:325-330), as a decimated two-stage transform: P-point
//             LDS radix FFTs of the Q = N/P sub-sequences x[Q n1 + n2],
//             then X[j] = sum_n2 W_N^(n2 j) Y_n2[j mod P] for j in the window.
//  k_epi_*  : float64 marginalisation epilogues (base.py:448-477, 775-820).
//
// Data layout in HBM (per detector, decimated order): element i holds bin
// n = Q (i mod P) + i / P, so that one n2 step reads P consecutive elements.
// Work is tiled S rows x (bins) per workgroup: the per-bin float64 phase
// basis and the data are loaded once and reused by S parameter rows.
#include "gwl_internal.h"
#include "gwl_bins.h"
#include <algorithm>
#include <type_traits>
#include <cfloat>

namespace gwl {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ------------------------------------------------------------------ k_reduce
struct RedArgs {
  EvDev ev;
  const double* params;
  int ld;
  long long n;
  int chunk;    // elements per chunk
  int nchunk;
  double* part; // [n][nchunk][3]
  const SampleConst* rows;  // [n] row constants (k_sample_const)
  const int* count;         // if set: only rows r < *count exist (n bounds it)
};

__device__ __forceinline__ long long live_rows(long long n, const int* count) {
  return count ? min(n, (long long)*count) : n;
}

// One thread per row: the row constants every k_reduce tile of the row reads.
// zero (optional) is cleared by thread 0: the escalation counter the
// epilogue of the same evaluation fills.  scatter (optional): the rows'
// indices into the caller's data (the escalated rows of a float32 pass).
__global__ __launch_bounds__(128) void k_sample_const(EvDev ev, const double* params, int ld,
                                                      long long n, SampleConst* out,
                                                      const int* count, int* zero,
                                                      const int* scatter) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (zero && r == 0) *zero = 0;
  if (r >= live_rows(n, count)) return;
  SampleConst sc;
  compute_sample_const(ev, params + r * ld, sc);
  sc.row = scatter ? scatter[r] : (int)r;
  out[r] = sc;
}

// A workgroup evaluates S rows over the chunks blockIdx.x, blockIdx.x +
// gridDim.x, ... of chunk bins each; every chunk's sums go to its own
// partials, so the values do not depend on the grid.  In natural order
// (no time marginalisation) a chunk outside the union of the rows' supports
// [kb, ke] is skipped and its partials are not written: the epilogue sums
// only the chunks that meet a row's own support, which that row's tile
// always computes.
// The sums of one chunk for the S rows of a tile (see k_reduce).
template <typename R, int WAVE, int ND, int S, int NT>
__device__ __forceinline__ void reduce_chunk(const RedArgs& a, long long n, int chunk,
                                             long long row0, const SampleConst* sc,
                                             const SampleFast<R>* sf, const int* band,
                                             double (*red)[S][3]) {
  using C = cplx<R>;
  const EvDev& ev = a.ev;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int i0 = chunk * a.chunk;
  const int i1 = min(i0 + a.chunk, ev.N);
  if (ev.Q == 1) {
    const int b1 = chunk == a.nchunk - 1 ? ev.N : i1 - 1;  // the last also holds Nyquist
    if (band[1] < i0 || band[0] > b1) return;
  }
  R acc_r[S], acc_i[S], acc_h[S];
#pragma unroll
  for (int s = 0; s < S; ++s) acc_r[s] = acc_i[s] = acc_h[s] = 0;
  for (int i = i0 + tid; i < i1; i += NT) {
    const int n = ev.Q * (i & (ev.P - 1)) + (i >> ev.logP);
    if (n < band[0] || n > band[1]) continue;
    C dw[ND];
    R w[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      dw[d] = load_dw<R>(ev, d, i);
      w[d] = load_w<R>(ev, d, i);
    }
    C I[S];
    bin_rows<R, WAVE, ND, S>(n, (double)n * ev.df, dw, w, sc, sf, I, acc_h, ND, ev.pol, ev.nf);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      acc_r[s] += I[s].x;
      acc_i[s] += I[s].y;
    }
  }
  double dr[S], di[S], dh[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    dr[s] = acc_r[s];
    di[s] = acc_i[s];
    dh[s] = acc_h[s];
  }
  // Nyquist bin n = N (inside the inner-product mask, outside the FFT range)
  if (chunk == a.nchunk - 1 && tid == 0) {
    const int n = ev.N;
    if (n >= band[0] && n <= band[1]) {
      C dw[ND];
      R w[ND];
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        dw[d] = cmk<C>((R)ev.dwN[d][0], (R)ev.dwN[d][1]);
        w[d] = (R)ev.wN[d];
      }
      C I[S];
      R h[S];
#pragma unroll
      for (int s = 0; s < S; ++s) h[s] = 0;
      bin_rows<R, WAVE, ND, S>(n, (double)n * ev.df, dw, w, sc, sf, I, h, ND, ev.pol, ev.nf);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        dr[s] += I[s].x;
        di[s] += I[s].y;
        dh[s] += h[s];
      }
    }
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    double x = wave_sum(dr[s]), y = wave_sum(di[s]), z = wave_sum(dh[s]);
    if (lane == 0) {
      red[wid][s][0] = x;
      red[wid][s][1] = y;
      red[wid][s][2] = z;
    }
  }
  __syncthreads();
  if (tid < S * 3) {
    const int s = tid / 3, c = tid % 3;
    const long long r = row0 + s;
    if (r < n) {
      double v = 0.0;
      for (int w = 0; w < NT / 64; ++w) v += red[w][s][c];
      a.part[(r * a.nchunk + chunk) * 3 + c] = v;
    }
  }
}

// A workgroup evaluates S rows over one chunk (LOOP false: chunk blockIdx.x)
// or over the chunks blockIdx.x, blockIdx.x + gridDim.x, ... (LOOP true);
// every chunk's sums go to its own partials, so the values do not depend
// on the grid.  In natural order (no time marginalisation) a chunk outside
// the union of the rows' supports [kb, ke] is skipped and its partials are
// not written: the epilogue sums only the chunks that meet a row's own
// support, which that row's tile always computes.
template <typename R, int WAVE, int ND, int S, int NT, bool LOOP>
__global__ __launch_bounds__(NT) void k_reduce(RedArgs a) {
  __shared__ SampleConst sc[S];
  __shared__ SampleFast<R> sf[S];
  __shared__ int band[2];
  __shared__ double red[NT / 64][S][3];
  const long long row0 = (long long)blockIdx.y * S;
  const long long n = live_rows(a.n, a.count);
  if (row0 >= n) return;  // uniform over the workgroup
  tile_prologue_load<R, S>(a.ev, a.rows, n, row0, sc, sf, band);
  if constexpr (!LOOP) {
    reduce_chunk<R, WAVE, ND, S, NT>(a, n, blockIdx.x, row0, sc, sf, band, red);
  } else {
    for (int chunk = blockIdx.x; chunk < a.nchunk; chunk += gridDim.x) {
      reduce_chunk<R, WAVE, ND, S, NT>(a, n, chunk, row0, sc, sf, band, red);
      __syncthreads();  // red is reused by the next chunk
    }
  }
}

// ------------------------------------------------------------------ k_time
struct TimeArgs {
  EvDev ev;
  const double* params;
  int ld;
  long long n;
  int G;
  double2* part;    // [n][G][L]
  double* hhpart;   // [n][G]
  double* rinfo;    // [n][4]: hscale, jitter, valid of each row (written by group 0)
};

// One tile of the time-marginalised transform: rows [row0, row0 + S) of an
// event, decimated sub-sequences n2 in [n2b, n2e).  Writes the window
// partials part[r][g][:] and hhpart[r][g].  LDS: buf [S][P], tw (twiddles,
// initialised by the caller), sc/sf/band/red scratch.
template <typename R, int WAVE, int ND, int S, int P, int NT>
__device__ __forceinline__ void time_tile(const EvDev& ev, int nd, const double* params, int ld,
                                          long long n, long long row0, int n2b, int n2e,
                                          bool nyquist, int G, int g, double2* part,
                                          double* hhpart, cplx<R>* buf, const cplx<R>* tw,
                                          SampleConst* sc, SampleFast<R>* sf, int* band,
                                          double (*red)[S], double* rinfo = nullptr) {
  using C = cplx<R>;
  constexpr int MAXBF8 = (S * P / 8 + NT - 1) / NT;
  constexpr int NPT = (P + NT - 1) / NT;   // n1 values per thread per step
  constexpr int JT = (P + NT - 1) / NT;    // window slots per thread
  const int tid = threadIdx.x;
  tile_prologue<R, S>(ev, params, ld, n, row0, sc, sf, band);
  const int Q = ev.N / P;
  double acc[S][JT][2];
  double hacc[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    hacc[s] = 0.0;
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) acc[s][jt][0] = acc[s][jt][1] = 0.0;
  }
  const unsigned nmask = (unsigned)ev.N - 1u;
  const R invN = (R)1.0 / (R)ev.N;

  for (int n2 = n2b; n2 < n2e; ++n2) {
    // 1) integrand of this decimated sub-sequence, S rows, into LDS
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
      const int n1 = tid + q * NT;
      if (n1 < P) {
        const int nb = Q * n1 + n2;
        C I[S];
        R h[S];
#pragma unroll
        for (int s = 0; s < S; ++s) h[s] = 0;
        if (nb >= band[0] && nb <= band[1]) {
          const int i = n2 * P + n1;
          C dw[ND];
          R w[ND];
#pragma unroll
          for (int d = 0; d < ND; ++d) {
            if (d < nd) {
              dw[d] = load_dw<R>(ev, d, i);
              w[d] = load_w<R>(ev, d, i);
            } else {
              dw[d] = cmk<C>((R)0, (R)0);
              w[d] = (R)0;
            }
          }
          bin_rows<R, WAVE, ND, S>(nb, (double)nb * ev.df, dw, w, sc, sf, I, h, nd, ev.pol, ev.nf);
        } else {
#pragma unroll
          for (int s = 0; s < S; ++s) I[s] = cmk<C>((R)0, (R)0);
        }
#pragma unroll
        for (int s = 0; s < S; ++s) {
          buf[s * P + n1] = I[s];
          hacc[s] += (double)h[s];
        }
      }
    }
    __syncthreads();
    // 2) P-point transforms in LDS
    fft_lds<NT, MAXBF8, P>(buf, tw, S);
    // 3) twiddle + accumulate the window bins j: X[j] += W_N^(n2 j) Y[j mod P]
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) {
      const int jj = tid + jt * NT;
      if (jj < ev.L) {
        const unsigned j = (unsigned)(ev.jlo + jj);
        const unsigned idx = ((unsigned)n2 * j) & nmask;
        const C t = expm2pi((R)idx * invN);
        const int pj = (int)(j & (unsigned)(P - 1));
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const C y = buf[s * P + pj];
          const C z = cmul(t, y);
          acc[s][jt][0] += (double)z.x;
          acc[s][jt][1] += (double)z.y;
        }
      }
    }
    __syncthreads();
  }
  // Nyquist bin: enters <h|h> only (the FFT drops it, base.py:327-329)
  if (nyquist && tid == 0 && ev.N >= band[0] && ev.N <= band[1]) {
    C dw[ND];
    R w[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      dw[d] = cmk<C>((R)ev.dwN[d][0], (R)ev.dwN[d][1]);
      w[d] = (R)ev.wN[d];
    }
    C I[S];
    R h[S];
#pragma unroll
    for (int s = 0; s < S; ++s) h[s] = 0;
    bin_rows<R, WAVE, ND, S>(ev.N, (double)ev.N * ev.df, dw, w, sc, sf, I, h, nd, ev.pol, ev.nf);
#pragma unroll
    for (int s = 0; s < S; ++s) hacc[s] += (double)h[s];
  }
  // write window partials
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const long long r = row0 + s;
    if (r < n) {
      double2* dst = part + ((size_t)r * G + g) * ev.L;
#pragma unroll
      for (int jt = 0; jt < JT; ++jt) {
        const int jj = tid + jt * NT;
        if (jj < ev.L) dst[jj] = make_double2(acc[s][jt][0], acc[s][jt][1]);
      }
    }
  }
  const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    double v = wave_sum(hacc[s]);
    if (lane == 0) red[wid][s] = v;
  }
  __syncthreads();
  if (tid < S) {
    const long long r = row0 + tid;
    if (r < n) {
      double v = 0.0;
      for (int w = 0; w < NT / 64; ++w) v += red[w][tid];
      hhpart[(size_t)r * G + g] = v;
      // the row constants the epilogue needs, so that it does not recompute them
      if (rinfo && g == 0) {
        rinfo[4 * r + 0] = sc[tid].hscale;
        rinfo[4 * r + 1] = sc[tid].jitter;
        rinfo[4 * r + 2] = (double)sc[tid].valid;
        rinfo[4 * r + 3] = 0.0;
      }
    }
  }
}

template <typename R, int WAVE, int ND, int S, int P, int NT>
__global__ __launch_bounds__(NT) void k_time(TimeArgs a) {
  using C = cplx<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  C* buf = reinterpret_cast<C*>(smem);       // [S][P]
  C* tw = buf + S * P;                       // twiddles (TwLayout<P>)
  __shared__ SampleConst sc[S];
  __shared__ SampleFast<R> sf[S];
  __shared__ int band[2];
  __shared__ double red[NT / 64][S];
  const int g = blockIdx.x;
  init_twiddles<P, NT>(tw);
  const int Q = a.ev.N / P;
  const int n2b = (int)(((long long)g * Q) / a.G), n2e = (int)(((long long)(g + 1) * Q) / a.G);
  time_tile<R, WAVE, ND, S, P, NT>(a.ev, ND, a.params, a.ld, a.n, (long long)blockIdx.y * S, n2b,
                                   n2e, g == 0, a.G, g, a.part, a.hhpart, buf, tw, sc, sf, band,
                                   red, a.rinfo);
}

// ------------------------------------------------------------------ epilogues
struct EpiArgs {
  EvDev ev;
  const double* params;
  int ld;
  long long n;
  int nchunk;          // k_reduce chunks
  int chunk;           // bins per chunk
  const double* part3; // [n][nchunk][3]
  int G;
  const double2* part; // [n][G][L]
  const double* hhpart;
  double* out;
  double* aux;
  // a window evaluated in slices: running (max, scaled sum) logsumexp state
  // per row [n], first / last slice flags
  double2* lse;
  int first, last;
  const double* rinfo;      // [n][4] hscale, jitter, valid from the transform (or null)
  int gstride;              // groups per row in part / hhpart (0: G)
  const SampleConst* rows;  // [n] row constants, or null (then computed here)
  // live rows and output slots: with count set only rows r < *count exist
  // and row r's results go to out[scatter[r]] (the float64 pass over the
  // escalated rows of a float32 evaluation)
  const int* count;
  const int* scatter;
  // escalation (float32 pass): rows with esc_coeff * Q > esc_tol append
  // their index to esc_idx and their parameter row to esc_params
  // ([n][GWL_NPARAM]) through the counter esc_count
  int* esc_count;
  int* esc_idx;
  double* esc_params;
  double esc_coeff, esc_tol;
};

// The scale Q of a row's inner products that bounds the float32 error of
// its ln L as esc_coeff * Q: |<d|h>| + <h|h> for the plain and phase-
// marginalised forms; under distance marginalisation the table is read at
// dt = <d|h> d / d_ref, ht = <h|h> (d / d_ref)^2, and ln L moves by
// s dt - s^2 ht / 2 near its peak s = dt / ht, so Q = 2 dt^2 / ht.
__device__ inline double escalation_scale(unsigned flags, double dhr, double dhi, double hh,
                                          double hs) {
  if (flags & GWL_DISTANCE_MARG) {
    const double dt = ((flags & GWL_PHASE_MARG) ? hypot(dhr, dhi) : fabs(dhr)) * hs;
    const double ht = hh * hs * hs;
    return ht > 0.0 ? 2.0 * dt * dt / ht + dt + ht : dt;
  }
  return hypot(dhr, dhi) + hh;
}

// lnL from the summed inner products (base.py:448-477, 775-792); one thread
// per row, knots in LDS.
__global__ __launch_bounds__(128) void k_epi_plain(EpiArgs a) {
  extern __shared__ double knots[];
  const EvDev& ev = a.ev;
  const bool dist = (ev.flags & GWL_DISTANCE_MARG) != 0;
  if (dist) stage_knots(ev, knots);
  __syncthreads();
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= live_rows(a.n, a.count)) return;
  const long long o = a.scatter ? (long long)a.scatter[r] : r;
  SampleConst sc;
  if (a.rows)
    sc = a.rows[r];
  else
    compute_sample_const(ev, a.params + r * a.ld, sc);
  double dr = 0.0, di = 0.0, hh = 0.0;
  const double* p = a.part3 + (size_t)r * a.nchunk * 3;
  // the chunks that meet the row's support (natural order), else all
  int c_lo = 0, c_hi = a.nchunk - 1;
  if (ev.Q == 1) {
    c_lo = sc.kb <= sc.ke ? sc.kb / a.chunk : 1;
    c_hi = sc.kb <= sc.ke ? min(a.nchunk - 1, sc.ke / a.chunk) : 0;
  }
  for (int c = c_lo; c <= c_hi; ++c) {
    dr += p[3 * c + 0];
    di += p[3 * c + 1];
    hh += p[3 * c + 2];
  }
  const double norm = 4.0 / ev.T;
  // accumulated sum h conj(d)/S; <d|h> = sum conj(h) d / S is its conjugate
  const double dhr = norm * dr, dhi = -norm * di;
  hh *= norm;
  if (a.aux) {
    a.aux[3 * o + 0] = dhr;
    a.aux[3 * o + 1] = dhi;
    a.aux[3 * o + 2] = hh;
  }
  if (!sc.valid) {
    a.out[o] = row_valid_sentinel();
    return;
  }
  if (a.esc_count && a.esc_coeff * escalation_scale(ev.flags, dhr, dhi, hh, sc.hscale) > a.esc_tol) {
    const int slot = atomicAdd(a.esc_count, 1);
    a.esc_idx[slot] = (int)r;
    const double* src = a.params + r * a.ld;
    double* dst = a.esc_params + (size_t)slot * GWL_NPARAM;
    for (int k = 0; k < GWL_NPARAM; ++k) dst[k] = src[k];
  }
  double ll;
  if (dist) {
    const double hs = sc.hscale;
    const double d = (ev.flags & GWL_PHASE_MARG) ? hypot(dhr, dhi) * hs : dhr * hs;
    const SplineY yy = spline_y(ev, knots + ev.sp.nx + 4, hh * hs * hs);
    ll = spline_x(ev, knots, yy, d);
  } else if (ev.flags & GWL_PHASE_MARG) {
    ll = ln_i0(hypot(dhr, dhi)) - hh / 2.0;
  } else {
    ll = dhr - hh / 2.0;
  }
  a.out[o] = ll;
}

// Time-marginalised lnL: logsumexp over the prior window (base.py:794-820).
// One wavefront per row, kEpiRows rows per workgroup, knots in LDS.
constexpr int kEpiRows = 4;

// One wavefront finishes one row: sums the group partials, evaluates the
// marginalised lnL at every window bin and reduces them with a running-max
// logsumexp.  ``knots`` holds the spline knots in LDS (distance only).
// Merge two (running max, scaled sum) logsumexp states.
__device__ __forceinline__ void lse_merge(double& m, double& s, double mo, double so) {
  const double mn = fmax(m, mo);
  if (mn > -INFINITY)
    s = (m > -INFINITY ? s * exp(m - mn) : 0.0) + (mo > -INFINITY ? so * exp(mo - mn) : 0.0);
  else
    s = 0.0;
  m = mn;
}

// A row's epilogue inputs: <h|h> summed over its G groups (row stride Gs),
// jitter, distance scale and validity (from the transform's rinfo, or
// recomputed).
__device__ __forceinline__ void epi_row_inputs(const EvDev& ev, const double* params, int ld,
                                              long long r, int G, int Gs, const double* hhpart,
                                              const double* rinfo, double& hh, double& jit,
                                              double& hs, int& valid) {
  hh = 0.0;
  jit = 0.0;
  hs = 1.0;
  valid = 1;
  if (rinfo) {  // written by the transform's first group
    hs = rinfo[4 * r + 0];
    jit = rinfo[4 * r + 1];
    valid = rinfo[4 * r + 2] != 0.0;
  } else {
    SampleConst sc;
    compute_sample_const(ev, params + r * ld, sc);
    jit = sc.jitter;
    hs = sc.hscale;
    valid = sc.valid;
  }
#pragma unroll 8
  for (int g = 0; g < G; ++g) hh += hhpart[(size_t)r * Gs + g];
  hh *= 4.0 / ev.T;
}

// The time-conditioned lnL at window bin jj plus the log time-prior weight
// (-inf outside the prior).
__device__ __forceinline__ double epi_bin_value(const EvDev& ev, const double* knots,
                                                const SplineY& yy, double hh, double hs,
                                                double jit, const double2* prow, int G, int jj) {
  const int j = ev.jlo + jj;
  const double t = (ev.start + (double)(j + 1) * ev.tstep) + jit;
  if (!(t >= ev.tmin && t <= ev.tmax)) return -INFINITY;
  const double norm = 4.0 / ev.T;
  double xr = 0.0, xi = 0.0;
#pragma unroll 8
  for (int g = 0; g < G; ++g) {
    const double2 v = prow[(size_t)g * ev.L + jj];
    xr += v.x;
    xi += v.y;
  }
  xr *= norm;
  xi *= norm;
  double v;
  if (ev.flags & GWL_DISTANCE_MARG) {
    const double d = (ev.flags & GWL_PHASE_MARG) ? hypot(xr, xi) * hs : xr * hs;
    v = spline_x(ev, knots, yy, d);
  } else if (ev.flags & GWL_PHASE_MARG) {
    v = ln_i0(hypot(xr, xi)) - hh / 2.0;
  } else {
    v = xr - hh / 2.0;
  }
  return v + time_log_weight(ev, t);
}

// The row's logsumexp state merged with the earlier slices' (lse), then
// stored for the next slice or turned into the output.
__device__ __forceinline__ void epi_finish(const EvDev& ev, long long r, double m, double s,
                                           double* out, double2* lse, int first, int last) {
  if (!first) {  // merge the slices before this one
    const double2 st = lse[r];
    lse_merge(m, s, st.x, st.y);
  }
  if (!last) {
    lse[r] = make_double2(m, s);
  } else {
    const double b = time_lse_scale(ev, ev.tdelta);
    out[r] = (m > -INFINITY) ? log(b * s) + m : -INFINITY;
  }
}

// base.py:794-820 for one row, one wavefront: group sums, marginalised lnL
// per window bin, logsumexp.
__device__ inline void epi_time_row(const EvDev& ev, const double* params, int ld, long long r,
                                    int G, const double2* part, const double* hhpart, double* out,
                                    const double* knots, double2* lse = nullptr, int first = 1,
                                    int last = 1, const double* rinfo = nullptr, int Gs = 0) {
  if (Gs <= 0) Gs = G;  // groups per row in the buffers (G of them summed here)
  const int lane = threadIdx.x & 63;
  double hh = 0.0, jit = 0.0, hs = 1.0;
  int valid = 1;
  if (lane == 0) epi_row_inputs(ev, params, ld, r, G, Gs, hhpart, rinfo, hh, jit, hs, valid);
  hh = __shfl(hh, 0, 64);
  jit = __shfl(jit, 0, 64);
  hs = __shfl(hs, 0, 64);
  valid = __shfl(valid, 0, 64);
  if (!valid) {
    if (lane == 0) out[r] = row_valid_sentinel();
    return;
  }
  SplineY yy;
  if (ev.flags & GWL_DISTANCE_MARG) yy = spline_y(ev, knots + ev.sp.nx + 4, hh * hs * hs);
  // single-pass (running max, scaled sum) logsumexp per lane, then merged
  double m = -INFINITY, s = 0.0;
  const double2* prow = part + (size_t)r * Gs * ev.L;
  for (int jj = lane; jj < ev.L; jj += 64) lse_push(m, s, epi_bin_value(ev, knots, yy, hh, hs, jit, prow, G, jj));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lse_merge(m, s, __shfl_xor(m, off, 64), __shfl_xor(s, off, 64));
  if (lane == 0) epi_finish(ev, r, m, s, out, lse, first, last);
}

// The same for few rows: one workgroup per row, its threads over the window
// bins (the per-bin spline / ln I0 work of one row in parallel instead of
// serially in one wavefront), the waves' states merged through LDS.
constexpr int kEpiWideThreads = 512;

__global__ __launch_bounds__(kEpiWideThreads) void k_epi_time_wide(EpiArgs a) {
  extern __shared__ double knots[];
  __shared__ double s_in[4];
  __shared__ int s_valid;
  __shared__ double s_m[kEpiWideThreads / 64], s_s[kEpiWideThreads / 64];
  const EvDev& ev = a.ev;
  if (ev.flags & GWL_DISTANCE_MARG) stage_knots(ev, knots);
  const long long r = blockIdx.x;
  const int Gs = a.gstride > 0 ? a.gstride : a.G;
  if (threadIdx.x == 0) {
    double hh, jit, hs;
    int valid;
    epi_row_inputs(ev, a.params, a.ld, r, a.G, Gs, a.hhpart, a.rinfo, hh, jit, hs, valid);
    s_in[0] = hh;
    s_in[1] = jit;
    s_in[2] = hs;
    s_valid = valid;
  }
  __syncthreads();
  if (!s_valid) {
    if (threadIdx.x == 0) a.out[r] = row_valid_sentinel();
    return;
  }
  const double hh = s_in[0], jit = s_in[1], hs = s_in[2];
  SplineY yy;
  if (ev.flags & GWL_DISTANCE_MARG) yy = spline_y(ev, knots + ev.sp.nx + 4, hh * hs * hs);
  double m = -INFINITY, s = 0.0;
  const double2* prow = a.part + (size_t)r * Gs * ev.L;
  for (int jj = threadIdx.x; jj < ev.L; jj += kEpiWideThreads)
    lse_push(m, s, epi_bin_value(ev, knots, yy, hh, hs, jit, prow, a.G, jj));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lse_merge(m, s, __shfl_xor(m, off, 64), __shfl_xor(s, off, 64));
  if ((threadIdx.x & 63) == 0) {
    s_m[threadIdx.x >> 6] = m;
    s_s[threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kEpiWideThreads / 64; ++w) lse_merge(m, s, s_m[w], s_s[w]);
    epi_finish(ev, r, m, s, a.out, a.lse, a.first, a.last);
  }
}

__global__ __launch_bounds__(64 * kEpiRows) void k_epi_time(EpiArgs a) {
  extern __shared__ double knots[];
  if (a.ev.flags & GWL_DISTANCE_MARG) stage_knots(a.ev, knots);
  __syncthreads();
  const long long r = (long long)blockIdx.x * kEpiRows + (threadIdx.x >> 6);
  if (r >= a.n) return;
  epi_time_row(a.ev, a.params, a.ld, r, a.G, a.part, a.hhpart, a.out, knots, a.lse, a.first,
               a.last, a.rinfo, a.gstride);
}

// Few rows with many groups: the group sums of each window bin (and of
// <h|h>) by one thread per (row, bin), written over group 0, so that the
// one-wavefront-per-row epilogue then reads one group instead of summing G
// strided partials per bin serially.
__global__ __launch_bounds__(256) void k_epi_gsum(double2* part, double* hhpart, long long n,
                                                  int G, int L) {
  const long long r = blockIdx.y;
  const int jj = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  if (jj < L) {
    double2* p = part + (size_t)r * G * L + jj;
    double xr = 0.0, xi = 0.0;
#pragma unroll 8
    for (int g = 0; g < G; ++g) {
      const double2 v = p[(size_t)g * L];
      xr += v.x;
      xi += v.y;
    }
    p[0] = make_double2(xr, xi);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double hh = 0.0;
    for (int g = 0; g < G; ++g) hh += hhpart[(size_t)r * G + g];
    hhpart[(size_t)r * G] = hh;
  }
}

// ------------------------------------------------------------------ multi-event
// Uneven batches over many events in one persistent launch: a work list of
// (event, row tile, group) items in longest-first order is drained through
// an atomic counter; every workgroup exits once the counter passes the list.
struct MultiArgs {
  const EvDev* evs;
  const double* const* params;
  const long long* rows;
  const int* groups;
  double2* const* part;
  double* const* hhpart;
  const MultiItem* items;
  const int* n_items;
  int* counter;
  int ld;
};

constexpr int kMultiS = 8, kMultiP = 512, kMultiNT = 512;

template <int WAVE>
__global__ __launch_bounds__(kMultiNT) void k_time_multi(MultiArgs a) {
  using C = cplx<float>;
  extern __shared__ __align__(16) unsigned char smem[];
  C* buf = reinterpret_cast<C*>(smem);
  C* tw = buf + kMultiS * kMultiP;
  __shared__ SampleConst sc[kMultiS];
  __shared__ SampleFast<float> sf[kMultiS];
  __shared__ int band[2];
  __shared__ double red[kMultiNT / 64][kMultiS];
  __shared__ EvDev s_ev;
  __shared__ int s_item;
  init_twiddles<kMultiP, kMultiNT>(tw);
  const int n_items = *a.n_items;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(a.counter, 1);
    __syncthreads();
    const int it = s_item;
    if (it >= n_items) break;
    const MultiItem m = a.items[it];
    constexpr int nw = (int)(sizeof(EvDev) / sizeof(unsigned));
    const unsigned* src = reinterpret_cast<const unsigned*>(a.evs + m.e);
    unsigned* dst = reinterpret_cast<unsigned*>(&s_ev);
    for (int i = threadIdx.x; i < nw; i += kMultiNT) dst[i] = src[i];
    __syncthreads();
    const int G = a.groups[m.e];
    const int Q = s_ev.N / kMultiP;
    const int n2b = (int)(((long long)m.g * Q) / G), n2e = (int)(((long long)(m.g + 1) * Q) / G);
    time_tile<float, WAVE, 3, kMultiS, kMultiP, kMultiNT>(
        s_ev, s_ev.nd, a.params[m.e], a.ld, a.rows[m.e], m.row0, n2b, n2e, m.g == 0, G, m.g,
        a.part[m.e], a.hhpart[m.e], buf, tw, sc, sf, band, red);
  }
}

struct MultiEpiArgs {
  const EvDev* evs;
  const double* const* params;
  const long long* rows;
  const int* groups;
  double2* const* part;
  double* const* hhpart;
  double* const* out;
  const int* block_event;       // [blocks] event of each epilogue block
  const long long* block_row0;  // [blocks] first row of each block
  const int* n_blocks;
  int ld;
};

__global__ __launch_bounds__(64 * kEpiRows) void k_epi_time_multi(MultiEpiArgs a) {
  extern __shared__ double knots[];
  __shared__ EvDev s_ev;
  if ((int)blockIdx.x >= *a.n_blocks) return;
  const int e = a.block_event[blockIdx.x];
  constexpr int nw = (int)(sizeof(EvDev) / sizeof(unsigned));
  const unsigned* src = reinterpret_cast<const unsigned*>(a.evs + e);
  unsigned* dst = reinterpret_cast<unsigned*>(&s_ev);
  for (int i = threadIdx.x; i < nw; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
  if (s_ev.flags & GWL_DISTANCE_MARG) stage_knots(s_ev, knots);
  __syncthreads();
  const long long r = a.block_row0[blockIdx.x] + (threadIdx.x >> 6);
  if (r >= a.rows[e]) return;
  epi_time_row(s_ev, a.params[e], a.ld, r, a.groups[e], a.part[e], a.hhpart[e], a.out[e], knots);
}

__global__ void k_zero_int(int* p) { *p = 0; }

// ------------------------------------------------------------------ response
// get_detector_response (interferometer.py:303-368) of one row, all bins.
struct RespArgs {
  EvDev ev;
  const double* params;
  int det;
  double* out;  // [nf][2]
};

template <typename R, int WAVE>
__global__ void k_response(RespArgs a) {
  using C = cplx<R>;
  const EvDev& ev = a.ev;
  __shared__ SampleConst sc[1];
  __shared__ SampleFast<R> sf[1];
  __shared__ int band[2];
  tile_prologue<R, 1>(ev, a.params, GWL_NPARAM, 1, 0, sc, sf, band);
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ev.nf) return;
  // unit "data" isolates h: I = h * conj(1)
  double re = 0.0, im = 0.0;
  const int d = a.det;
  if (n >= ev.mlo[d] && n <= ev.mhi[d] && n >= sc[0].kb && n <= sc[0].ke) {
    // single-detector projection: evaluate with only detector d's amplitude
    SampleConst one = sc[0];
    SampleFast<R> onef = sf[0];
    one.dt[0] = sc[0].dt[d];
    one.cre[0] = sc[0].cre[d];
    one.cim[0] = sc[0].cim[d];
    onef.cre[0] = sf[0].cre[d];
    onef.cim[0] = sf[0].cim[d];
    onef.cabs2[0] = sf[0].cabs2[d];
    C dw[1] = {cmk<C>((R)1, (R)0)};
    R w[1] = {(R)0};
    C I[1];
    R h[1] = {(R)0};
    bin_rows<R, WAVE, 1, 1>(n, (double)n * ev.df, dw, w, &one, &onef, I, h, 1, ev.pol, ev.nf);
    re = (double)I[0].x * ev.sigma;
    im = (double)I[0].y * ev.sigma;
  }
  a.out[2 * n] = re;
  a.out[2 * n + 1] = im;
}

// ------------------------------------------------------------------ dispatch
static int pick_nd(int nd) { return nd; }

// rows per workgroup of the time-marginalised kernel (LDS: S P complex);
// instantiated: P = 512: 1, 2, 4, 8; P = 1024: 1, 2, 4; P = 2048: 1, 2.
static int time_tile_rows(int P, long long n, int want) {
  const int maxS = P == 512 ? 8 : P == 1024 ? 4 : P == 2048 ? 2 : 1;
  if (want > 0) {
    int S = 1;
    while (S * 2 <= want && S * 2 <= maxS) S *= 2;
    return S;
  }
  if (n < 64) return 1;
  return maxS;
}

// groups of sub-sequences per row: enough workgroups to fill the CUs
static int time_tile_groups(const EvDev& ev, long long n, int S, int n_cu, int wg_per_cu) {
  const int Q = ev.N / ev.P;
  const long long nblk = (n + S - 1) / S;
  const long long target = (long long)n_cu * (wg_per_cu > 0 ? wg_per_cu : 8);
  int G = 1;
  while (G < Q && G < 64 && nblk * G < target) G *= 2;
  return G;
}

// rows up to which the epilogue sums the groups bin-parallel first (k_epi_gsum)
constexpr int64_t kEpiGsumRows = 64;

int launch_epi_time(gwl_event* e, const double* params, int64_t n, int ld, int G,
                    const double2* part, const double* hhpart, double* out, hipStream_t st,
                    const EvDev* slice, double2* lse, int first, int last, const double* rinfo) {
  const EvDev& ev = slice ? *slice : e->dev;
  EpiArgs ea{};
  ea.ev = ev;
  ea.params = params;
  ea.ld = ld;
  ea.n = n;
  ea.G = G;
  ea.part = part;
  ea.hhpart = hhpart;
  ea.out = out;
  ea.lse = lse;
  ea.first = first;
  ea.last = last;
  ea.rinfo = rinfo;
  const size_t klds = (ev.flags & GWL_DISTANCE_MARG) ? (size_t)(ev.sp.nx + ev.sp.ny + 8) * sizeof(double) : 0;
  // few rows: the groups summed bin-parallel first, then one workgroup per row
  if (n <= kEpiGsumRows) {
    if (G > 1) {
      hipLaunchKernelGGL(k_epi_gsum, dim3((unsigned)((ev.L + 255) / 256), (unsigned)n), dim3(256),
                         0, st, const_cast<double2*>(part), const_cast<double*>(hhpart),
                         (long long)n, G, ev.L);
      GWL_HIP_CHECK(hipGetLastError());
      ea.G = 1;
      ea.gstride = G;
    }
    hipLaunchKernelGGL(k_epi_time_wide, dim3((unsigned)n), dim3(kEpiWideThreads), klds, st, ea);
    GWL_HIP_CHECK(hipGetLastError());
    return GWL_OK;
  }
  hipLaunchKernelGGL(k_epi_time, dim3((unsigned)((n + kEpiRows - 1) / kEpiRows)),
                     dim3(64 * kEpiRows), klds, st, ea);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

// Bytes of the k_reduce partials [n][nchunk][3] double (256-aligned); the
// row constants [n] SampleConst follow them in the event workspace.
static size_t reduce_part_bytes(const EvDev& ev, long long rows) {
  const long long nchunk = (ev.N + 1023) / 1024;
  return ((size_t)rows * nchunk * 3 * sizeof(double) + 255) / 256 * 256;
}

// Bytes in front of the tiled transform's partials in the event workspace:
// the per-row logsumexp state of a window evaluated in slices.
static size_t time_lse_bytes(long long rows) {
  return ((size_t)rows * sizeof(double2) + 255) / 256 * 256;
}

// The tiled time-marginalised transform: window partials [n][G][L] and
// <h|h> partials [n][G] in the event workspace after the logsumexp state
// (G returned).  slice: the event with jlo / L of one window slice (null:
// the whole window).
template <typename R, int WAVE, int ND>
static int launch_time_tiled(gwl_event* e, const double* params, int64_t n, int ld,
                             hipStream_t st, int* G_out, double2** part_out, double** hh_out,
                             const EvDev* slice = nullptr, double** rinfo_out = nullptr) {
  const EvDev& ev = slice ? *slice : e->dev;
  {
    // tile: S rows x P-point transforms, G groups of sub-sequences
    constexpr int NT = 512;
    const int P = ev.P;
    TimeArgs ta;
    ta.ev = ev;
    ta.params = params;
    ta.ld = ld;
    ta.n = n;
    const int S = time_tile_rows(P, n, e->tile_rows);
    const long long nblk = (n + S - 1) / S;
    const int G = time_tile_groups(ev, n, S, e->ctx->n_cu, e->tile_wg_per_cu);
    ta.G = G;
    char* base = reinterpret_cast<char*>(e->ws) + time_lse_bytes(n);
    ta.part = reinterpret_cast<double2*>(base);
    ta.hhpart = reinterpret_cast<double*>(base + (size_t)n * G * ev.L * sizeof(double2));
    ta.rinfo = reinterpret_cast<double*>(
        base + (((size_t)n * G * ev.L * sizeof(double2) + (size_t)n * G * sizeof(double)) + 255) /
                   256 * 256);
    dim3 grid(G, (unsigned)nblk);
#define GWL_LAUNCH_TIME(SS, PP)                                                                  \
  do {                                                                                           \
    const size_t lds = (size_t)(SS * PP + TwLayout<PP>::kEntries) * sizeof(cplx<R>);            \
    if (lds > e->ctx->lds_per_block) {                                                           \
      set_error("time-marginalisation tile exceeds the LDS of a workgroup");                     \
      return GWL_E_UNSUPPORTED;                                                                  \
    }                                                                                            \
    hipLaunchKernelGGL((k_time<R, WAVE, ND, SS, PP, NT>), grid, dim3(NT), lds, st, ta);          \
  } while (0)
    if (int rc = timing_begin(e, st)) return rc;
    if (P == 512 && S == 8) GWL_LAUNCH_TIME(8, 512);
    else if (P == 512 && S == 4) GWL_LAUNCH_TIME(4, 512);
    else if (P == 512 && S == 2) GWL_LAUNCH_TIME(2, 512);
    else if (P == 512) GWL_LAUNCH_TIME(1, 512);
    else if (P == 1024 && S == 4) GWL_LAUNCH_TIME(4, 1024);
    else if (P == 1024 && S == 2) GWL_LAUNCH_TIME(2, 1024);
    else if (P == 1024) GWL_LAUNCH_TIME(1, 1024);
    else if (P == 2048 && S == 2) GWL_LAUNCH_TIME(2, 2048);
    else if (P == 2048) GWL_LAUNCH_TIME(1, 2048);
    else if (P == 4096) GWL_LAUNCH_TIME(1, 4096);
    else if (P == 8192) GWL_LAUNCH_TIME(1, 8192);
    else if (P == 16384) GWL_LAUNCH_TIME(1, 16384);
    else {
      set_error("time-marginalisation transform size not supported");
      return GWL_E_UNSUPPORTED;
    }
#undef GWL_LAUNCH_TIME
    GWL_HIP_CHECK(hipGetLastError());
    if (int rc = timing_end(e, st)) return rc;
    *G_out = G;
    *part_out = ta.part;
    *hh_out = ta.hhpart;
    if (rinfo_out) *rinfo_out = ta.rinfo;
    return GWL_OK;
  }
}

// Workspace of the inner-product path: partials [n][nchunk][3], row
// constants [n], then the escalation list (parameter rows [n][GWL_NPARAM],
// row indices [n], the counter).
struct ReduceWs {
  double* part;
  SampleConst* rows;
  double* esc_params;
  int* esc_idx;
  int* esc_count;
};

static ReduceWs reduce_ws(gwl_event* e, long long n) {
  ReduceWs w;
  char* p = reinterpret_cast<char*>(e->ws);
  w.part = reinterpret_cast<double*>(p);
  p += reduce_part_bytes(e->dev, n);
  w.rows = reinterpret_cast<SampleConst*>(p);
  p += ((size_t)n * sizeof(SampleConst) + 255) / 256 * 256;
  w.esc_params = reinterpret_cast<double*>(p);
  p += ((size_t)n * GWL_NPARAM * sizeof(double) + 255) / 256 * 256;
  w.esc_idx = reinterpret_cast<int*>(p);
  p += ((size_t)n * sizeof(int) + 255) / 256 * 256;
  w.esc_count = reinterpret_cast<int*>(p);
  return w;
}

static size_t reduce_ws_bytes(const EvDev& ev, long long n) {
  return reduce_part_bytes(ev, n) + ((size_t)n * sizeof(SampleConst) + 255) / 256 * 256 +
         ((size_t)n * GWL_NPARAM * sizeof(double) + 255) / 256 * 256 +
         ((size_t)n * sizeof(int) + 255) / 256 * 256 + 256;
}

// One pass of the inner-product path (no time marginalisation): row
// constants, k_reduce, epilogue.  count / scatter: the float64 pass over the
// rows a float32 pass escalated (their number on the device, n bounds it);
// esc: the float32 pass that selects them.
template <typename R, int WAVE, int ND>
static int launch_reduce(gwl_event* e, const double* params, int64_t n, int ld, double* out,
                         double* aux, hipStream_t st, const int* count, const int* scatter,
                         bool esc) {
  const EvDev& ev = e->dev;
  constexpr int NT = 256;
  // float64 tiles of more than 4 rows hold too many registers for more than
  // one wavefront per SIMD; a row's sums do not depend on S
  const int S = std::is_same<R, double>::value
                    ? 4
                    : e->tile_rows >= 16 ? 16 : e->tile_rows >= 8 || e->tile_rows <= 0 ? 8 : 4;
  const ReduceWs ws = reduce_ws(e, n);
  RedArgs ra;
  ra.ev = ev;
  ra.params = params;
  ra.ld = ld;
  ra.n = n;
  // chunking fixed per event (independent of the batch) so that a row's
  // reduction order does not depend on what it is batched with
  ra.chunk = NT * 4;
  ra.nchunk = (ev.N + ra.chunk - 1) / ra.chunk;
  ra.part = ws.part;
  ra.rows = ws.rows;
  ra.count = count;
  const long long nblk = (n + S - 1) / S;
  // chunk groups per tile: every chunk in parallel when the rows' supports
  // span the band; with the ISCO termination TaylorF2 supports are short,
  // most chunks lie outside every row's support and fewer, longer
  // workgroups (one per tile once the tiles alone fill the CUs several
  // times) avoid launching workgroups that skip them all
  const long long target = 8LL * e->ctx->n_cu;
  const int G = (WAVE != GWL_TAYLORF2 || !(ev.flags & GWL_TF2_END_ISCO))
                    ? ra.nchunk
                    : (int)std::max<long long>(1, std::min<long long>(
                                                      ra.nchunk, (target + nblk - 1) / nblk));
  if (!count)
    if (int rc = timing_begin(e, st)) return rc;
  hipLaunchKernelGGL(k_sample_const, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, ev,
                     params, ld, (long long)n, ws.rows, count, esc ? ws.esc_count : nullptr,
                     scatter);
  const dim3 grid(G, (unsigned)nblk);
  if (G == ra.nchunk) {
    if (S == 16)
      hipLaunchKernelGGL((k_reduce<R, WAVE, ND, 16, NT, false>), grid, dim3(NT), 0, st, ra);
    else if (S == 4)
      hipLaunchKernelGGL((k_reduce<R, WAVE, ND, 4, NT, false>), grid, dim3(NT), 0, st, ra);
    else
      hipLaunchKernelGGL((k_reduce<R, WAVE, ND, 8, NT, false>), grid, dim3(NT), 0, st, ra);
  } else {
    if (S == 16)
      hipLaunchKernelGGL((k_reduce<R, WAVE, ND, 16, NT, true>), grid, dim3(NT), 0, st, ra);
    else if (S == 4)
      hipLaunchKernelGGL((k_reduce<R, WAVE, ND, 4, NT, true>), grid, dim3(NT), 0, st, ra);
    else
      hipLaunchKernelGGL((k_reduce<R, WAVE, ND, 8, NT, true>), grid, dim3(NT), 0, st, ra);
  }
  GWL_HIP_CHECK(hipGetLastError());
  if (!count)
    if (int rc = timing_end(e, st)) return rc;
  EpiArgs ea{};
  ea.ev = ev;
  ea.params = params;
  ea.ld = ld;
  ea.n = n;
  ea.nchunk = ra.nchunk;
  ea.chunk = ra.chunk;
  ea.part3 = ra.part;
  ea.out = out;
  ea.aux = aux;
  ea.rows = ra.rows;
  ea.count = count;
  ea.scatter = scatter;
  if (esc) {
    ea.esc_count = ws.esc_count;
    ea.esc_idx = ws.esc_idx;
    ea.esc_params = ws.esc_params;
    ea.esc_coeff = e->esc_coeff;
    ea.esc_tol = e->esc_tol;
  }
  const size_t klds = (ev.flags & GWL_DISTANCE_MARG) ? (size_t)(ev.sp.nx + ev.sp.ny + 8) * sizeof(double) : 0;
  hipLaunchKernelGGL(k_epi_plain, dim3((unsigned)((n + 127) / 128)), dim3(128), klds, st, ea);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

template <typename R, int WAVE, int ND>
static int eval_typed(gwl_event* e, const double* params, int64_t n, int ld, double* out,
                      double* aux, hipStream_t st) {
  const EvDev& ev = e->dev;
  if constexpr (std::is_same<R, float>::value && WAVE == GWL_TAYLORF2) {
    if (time_wave_eligible(e)) return launch_time_wave(e, params, n, ld, out, st);
  }
  if (ev.flags & GWL_TIME_MARG) {
    // a window longer than one transform: slices of P bins, one pass each,
    // the logsumexp carried between them
    const int passes = (ev.L + ev.P - 1) / ev.P;
    double2* lse = reinterpret_cast<double2*>(e->ws);
    for (int p = 0; p < passes; ++p) {
      EvDev slice = ev;
      slice.jlo = ev.jlo + p * ev.P;
      slice.L = std::min(ev.P, ev.L - p * ev.P);
      int G = 0;
      double2* part = nullptr;
      double* hh = nullptr;
      double* rinfo = nullptr;
      if (int rc = launch_time_tiled<R, WAVE, ND>(e, params, n, ld, st, &G, &part, &hh, &slice,
                                                  &rinfo))
        return rc;
      if (int rc = launch_epi_time(e, params, n, ld, G, part, hh, out, st, &slice, lse, p == 0,
                                   p == passes - 1, rinfo))
        return rc;
    }
    return GWL_OK;
  }
  const bool esc = std::is_same<R, float>::value && e->esc_coeff > 0.0;
  if (int rc = launch_reduce<R, WAVE, ND>(e, params, n, ld, out, aux, st, nullptr, nullptr, esc))
    return rc;
  if (!esc) return GWL_OK;
  // the escalated rows again in float64: same stream, no host round trip;
  // the grid is sized for n rows and the kernels read the live count
  const ReduceWs ws = reduce_ws(e, n);
  e->esc_count_dev = ws.esc_count;
  return launch_reduce<double, WAVE, ND>(e, ws.esc_params, n, GWL_NPARAM, out, aux, st,
                                         ws.esc_count, ws.esc_idx, false);
}

template <typename R, int WAVE>
static int eval_nd(gwl_event* e, const double* params, int64_t n, int ld, double* out, double* aux,
                   hipStream_t st) {
  switch (pick_nd(e->dev.nd)) {
    case 1: return eval_typed<R, WAVE, 1>(e, params, n, ld, out, aux, st);
    case 2: return eval_typed<R, WAVE, 2>(e, params, n, ld, out, aux, st);
    case 3: return eval_typed<R, WAVE, 3>(e, params, n, ld, out, aux, st);
    case 4: return eval_typed<R, WAVE, 4>(e, params, n, ld, out, aux, st);
  }
  set_error("unsupported detector count");
  return GWL_E_INVALID;
}

size_t workspace_bytes_for(const gwl_event* e, int64_t rows, int precision) {
  // a float32 evaluation on the wave-per-row transform uses that path's
  // workspace only
  if ((e->dev.flags & GWL_TIME_MARG) && precision == GWL_FP32 && e->dev.approx == GWL_TAYLORF2 &&
      time_wave_eligible(e))
    return time_wave_ws_bytes(e, rows);
  return workspace_bytes(e, rows);
}

size_t workspace_bytes(const gwl_event* e, int64_t rows) {
  const EvDev& ev = e->dev;
  if (ev.flags & GWL_TIME_MARG) {
    // partials [n][G][L] double2 + [n][G] double; n G <= max(n, 2 S 8 n_cu)
    // for every n <= rows (G halves as the row blocks double)
    const int S = time_tile_rows(ev.P, rows, e->tile_rows);
    const int wpc = e->tile_wg_per_cu > 0 ? e->tile_wg_per_cu : 8;
    const long long bound = std::max<long long>(
        (long long)rows * time_tile_groups(ev, rows, S, e->ctx->n_cu, e->tile_wg_per_cu),
        2LL * 8 * wpc * e->ctx->n_cu);
    size_t b = (size_t)bound * (std::min(ev.L, ev.P) * sizeof(double2) + sizeof(double)) +
               time_lse_bytes(rows) + (size_t)rows * 4 * sizeof(double) + 4096;
    if (ev.approx == GWL_TAYLORF2 && ev.tf2b) b = std::max(b, time_wave_ws_bytes(e, rows));
    return b;
  }
  return reduce_ws_bytes(ev, rows) + 4096;
}

// A host-polarization event needs the caller's polarizations for every row.
static int check_polarizations(const gwl_event* e, int64_t n, const char* who) {
  if (e->dev.approx != GWL_HOST_POLARIZATIONS) return GWL_OK;
  if (!e->dev.pol || n > e->pol_rows) {
    char msg[200];
    snprintf(msg, sizeof msg, "%s: the event's source model is evaluated by the caller; "
             "gwl_event_set_polarizations must hold at least %lld rows (it holds %lld)",
             who, (long long)n, (long long)e->pol_rows);
    set_error(msg);
    return GWL_E_INVALID;
  }
  return GWL_OK;
}

int launch_eval(gwl_event* e, const double* params, int64_t n, int ld, int precision, double* out,
                double* aux, hipStream_t st) {
  (void)hipGetLastError();  // clear a stale error left by an unrelated call
  if (int rc = check_polarizations(e, n, "gwl_eval")) return rc;
  const int wave = e->dev.approx;
  if (precision == GWL_FP32) {
    if (wave == GWL_TAYLORF2) return eval_nd<float, GWL_TAYLORF2>(e, params, n, ld, out, aux, st);
    if (wave == GWL_HOST_POLARIZATIONS)
      return eval_nd<float, GWL_HOST_POLARIZATIONS>(e, params, n, ld, out, aux, st);
    return eval_nd<float, GWL_SINE_GAUSSIAN>(e, params, n, ld, out, aux, st);
  }
  if (wave == GWL_TAYLORF2) return eval_nd<double, GWL_TAYLORF2>(e, params, n, ld, out, aux, st);
  if (wave == GWL_HOST_POLARIZATIONS)
    return eval_nd<double, GWL_HOST_POLARIZATIONS>(e, params, n, ld, out, aux, st);
  return eval_nd<double, GWL_SINE_GAUSSIAN>(e, params, n, ld, out, aux, st);
}

// d_inner_h_array at the window and <h|h> of every row (base.py:325-330
// before the time marginalisation): the group partials summed and scaled by
// 4/T; one thread per (row, window bin).
// A window slice of L bins goes to columns [off, off + L) of rows of ldo bins.
__global__ __launch_bounds__(256) void k_time_series_out(long long n, int G, int L, double norm,
                                                         const double2* part, const double* hhpart,
                                                         double* dh, double* hh, int ldo, int off) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * L) return;
  const long long r = t / L;
  const int jj = (int)(t - r * L);
  double xr = 0.0, xi = 0.0;
  for (int g = 0; g < G; ++g) {
    const double2 v = part[((size_t)r * G + g) * L + jj];
    xr += v.x;
    xi += v.y;
  }
  const size_t o = (size_t)r * ldo + off + jj;
  dh[2 * o + 0] = norm * xr;
  dh[2 * o + 1] = norm * xi;
  if (jj == 0) {
    double s = 0.0;
    for (int g = 0; g < G; ++g) s += hhpart[(size_t)r * G + g];
    hh[r] = norm * s;
  }
}

template <typename R, int WAVE, int ND>
static int series_typed(gwl_event* e, const double* params, int64_t n, int ld, double* dh,
                        double* hh, hipStream_t st) {
  const EvDev& ev = e->dev;
  // a window longer than one transform in slices of P bins
  for (int off = 0; off < ev.L; off += ev.P) {
    EvDev slice = ev;
    slice.jlo = ev.jlo + off;
    slice.L = std::min(ev.P, ev.L - off);
    int G = 0;
    double2* part = nullptr;
    double* hhp = nullptr;
    if (int rc = launch_time_tiled<R, WAVE, ND>(e, params, n, ld, st, &G, &part, &hhp, &slice))
      return rc;
    const long long total = (long long)n * slice.L;
    hipLaunchKernelGGL(k_time_series_out, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       st, (long long)n, G, slice.L, 4.0 / ev.T, part, hhp, dh, hh, ev.L, off);
    GWL_HIP_CHECK(hipGetLastError());
  }
  return GWL_OK;
}

template <typename R, int WAVE>
static int series_nd(gwl_event* e, const double* params, int64_t n, int ld, double* dh, double* hh,
                     hipStream_t st) {
  switch (pick_nd(e->dev.nd)) {
    case 1: return series_typed<R, WAVE, 1>(e, params, n, ld, dh, hh, st);
    case 2: return series_typed<R, WAVE, 2>(e, params, n, ld, dh, hh, st);
    case 3: return series_typed<R, WAVE, 3>(e, params, n, ld, dh, hh, st);
    case 4: return series_typed<R, WAVE, 4>(e, params, n, ld, dh, hh, st);
  }
  set_error("unsupported detector count");
  return GWL_E_INVALID;
}

int launch_time_series(gwl_event* e, const double* params, int64_t n, int ld, int precision,
                       double* dh, double* hh, hipStream_t st) {
  if (int rc = check_polarizations(e, n, "gwl_eval_time_series")) return rc;
  const int wave = e->dev.approx;
  if (precision == GWL_FP32) {
    if (wave == GWL_TAYLORF2) return series_nd<float, GWL_TAYLORF2>(e, params, n, ld, dh, hh, st);
    if (wave == GWL_HOST_POLARIZATIONS)
      return series_nd<float, GWL_HOST_POLARIZATIONS>(e, params, n, ld, dh, hh, st);
    return series_nd<float, GWL_SINE_GAUSSIAN>(e, params, n, ld, dh, hh, st);
  }
  if (wave == GWL_TAYLORF2) return series_nd<double, GWL_TAYLORF2>(e, params, n, ld, dh, hh, st);
  if (wave == GWL_HOST_POLARIZATIONS)
    return series_nd<double, GWL_HOST_POLARIZATIONS>(e, params, n, ld, dh, hh, st);
  return series_nd<double, GWL_SINE_GAUSSIAN>(e, params, n, ld, dh, hh, st);
}

size_t multi_lds_bytes() {
  return (size_t)(kMultiS * kMultiP + TwLayout<kMultiP>::kEntries) * sizeof(float2);
}

// Work list for the current row counts: items (event, group, row tile) in
// decreasing cost (sub-sequence steps) order, and the epilogue blocks.
int multi_set_rows(gwl_multi* m, const int64_t* rows, hipStream_t st) {
  std::vector<MultiItem> items;
  std::vector<int> bev;
  std::vector<long long> brow;
  std::vector<long long> cost;
  for (int e = 0; e < m->n_events; ++e) {
    if (rows[e] < 0 || rows[e] > m->max_rows[e]) {
      set_error("gwl_multi_set_rows: rows outside [0, max_rows]");
      return GWL_E_INVALID;
    }
    m->rows[e] = rows[e];
    const EvDev& ev = m->events[e]->dev;
    const int Q = ev.N / kMultiP;
    const int G = m->groups[e];
    const int tile = m->wave_path ? 4 : kMultiS;  // rows per item
    for (long long r0 = 0; r0 < rows[e]; r0 += tile) {
      for (int g = 0; g < G; ++g) {
        MultiItem it;
        it.e = e;
        it.g = g;
        it.row0 = r0;
        items.push_back(it);
        cost.push_back(((long long)(g + 1) * Q) / G - ((long long)g * Q) / G);
      }
    }
    for (long long r0 = 0; r0 < rows[e]; r0 += kEpiRows) {
      bev.push_back(e);
      brow.push_back(r0);
    }
  }
  if (items.size() > m->max_items || bev.size() > m->max_blocks) {
    set_error("gwl_multi_set_rows: work list larger than allocated");
    return GWL_E_INVALID;
  }
  std::vector<size_t> order(items.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
  std::vector<MultiItem> sorted(items.size());
  for (size_t i = 0; i < order.size(); ++i) sorted[i] = items[order[i]];
  const int n_items = (int)sorted.size();
  const int n_blocks = (int)bev.size();
  std::vector<long long> rr(m->rows.begin(), m->rows.end());
  GWL_HIP_CHECK(hipMemcpy(m->d_rows, rr.data(), rr.size() * sizeof(long long), hipMemcpyHostToDevice));
  if (n_items)
    GWL_HIP_CHECK(hipMemcpy(m->d_items, sorted.data(), n_items * sizeof(MultiItem), hipMemcpyHostToDevice));
  GWL_HIP_CHECK(hipMemcpy(m->d_n_items, &n_items, sizeof(int), hipMemcpyHostToDevice));
  if (n_blocks) {
    GWL_HIP_CHECK(hipMemcpy(m->d_block_event, bev.data(), n_blocks * sizeof(int), hipMemcpyHostToDevice));
    GWL_HIP_CHECK(hipMemcpy(m->d_block_row0, brow.data(), n_blocks * sizeof(long long), hipMemcpyHostToDevice));
  }
  GWL_HIP_CHECK(hipMemcpy(m->d_n_blocks, &n_blocks, sizeof(int), hipMemcpyHostToDevice));
  m->n_items = n_items;
  m->n_blocks = n_blocks;
  (void)st;
  return GWL_OK;
}

// zero the queue counter (a kernel, not a memset node), persistent pass,
// epilogue over every block the work list can hold.
int multi_launch_raw(gwl_multi* m, hipStream_t st) {
  if (m->wave_path) {
    if (int rc = multi_launch_wave(m, st)) return rc;
    return multi_launch_epi(m, st);
  }
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, m->d_counter);
  GWL_HIP_CHECK(hipGetLastError());
  MultiArgs a;
  a.evs = m->d_evs;
  a.params = m->d_params;
  a.rows = m->d_rows;
  a.groups = m->d_groups;
  a.part = reinterpret_cast<double2* const*>(m->d_parts);
  a.hhpart = m->d_hhs;
  a.items = m->d_items;
  a.n_items = m->d_n_items;
  a.counter = m->d_counter;
  a.ld = GWL_NPARAM;
  const dim3 grid((unsigned)(m->ctx->n_cu * 2));
  if (m->wave == GWL_TAYLORF2)
    hipLaunchKernelGGL(k_time_multi<GWL_TAYLORF2>, grid, dim3(kMultiNT), multi_lds_bytes(), st, a);
  else
    hipLaunchKernelGGL(k_time_multi<GWL_SINE_GAUSSIAN>, grid, dim3(kMultiNT), multi_lds_bytes(), st, a);
  GWL_HIP_CHECK(hipGetLastError());
  return multi_launch_epi(m, st);
}

// float64 epilogue of every row of every event (4 rows per block)
int multi_launch_epi(gwl_multi* m, hipStream_t st) {
  MultiEpiArgs ea;
  ea.evs = m->d_evs;
  ea.params = m->d_params;
  ea.rows = m->d_rows;
  ea.groups = m->d_groups;
  ea.part = reinterpret_cast<double2* const*>(m->d_parts);
  ea.hhpart = m->d_hhs;
  ea.out = m->d_outs;
  ea.block_event = m->d_block_event;
  ea.block_row0 = m->d_block_row0;
  ea.n_blocks = m->d_n_blocks;
  ea.ld = GWL_NPARAM;
  size_t klds = 0;
  for (auto* ev : m->events)
    if (ev->dev.flags & GWL_DISTANCE_MARG)
      klds = std::max(klds, (size_t)(ev->dev.sp.nx + ev->dev.sp.ny + 8) * sizeof(double));
  const size_t grid = m->capturing ? m->max_blocks : (size_t)std::max(1, m->n_blocks);
  hipLaunchKernelGGL(k_epi_time_multi, dim3((unsigned)grid), dim3(64 * kEpiRows), klds, st, ea);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

int timing_begin(gwl_event* e, hipStream_t st) {
  if (!e->timing) return GWL_OK;
  GWL_HIP_CHECK(hipEventRecord(e->t0, st));
  return GWL_OK;
}

int timing_end(gwl_event* e, hipStream_t st) {
  if (!e->timing) return GWL_OK;
  GWL_HIP_CHECK(hipEventRecord(e->t1, st));
  GWL_HIP_CHECK(hipEventSynchronize(e->t1));
  float ms = 0.f;
  GWL_HIP_CHECK(hipEventElapsedTime(&ms, e->t0, e->t1));
  e->timed_ms += ms;
  e->timed_launches += 1;
  return GWL_OK;
}

int launch_response(gwl_event* e, const double* params_dev, int det, int precision, double* out_dev,
                    hipStream_t st) {
  RespArgs ra;
  ra.ev = e->dev;
  ra.params = params_dev;
  ra.det = det;
  ra.out = out_dev;
  dim3 grid((e->dev.nf + 255) / 256);
  if (int rc = check_polarizations(e, 1, "gwl_detector_response")) return rc;
  const int wave = e->dev.approx;
  if (precision == GWL_FP32) {
    if (wave == GWL_TAYLORF2)
      hipLaunchKernelGGL((k_response<float, GWL_TAYLORF2>), grid, dim3(256), 0, st, ra);
    else if (wave == GWL_HOST_POLARIZATIONS)
      hipLaunchKernelGGL((k_response<float, GWL_HOST_POLARIZATIONS>), grid, dim3(256), 0, st, ra);
    else
      hipLaunchKernelGGL((k_response<float, GWL_SINE_GAUSSIAN>), grid, dim3(256), 0, st, ra);
  } else {
    if (wave == GWL_TAYLORF2)
      hipLaunchKernelGGL((k_response<double, GWL_TAYLORF2>), grid, dim3(256), 0, st, ra);
    else if (wave == GWL_HOST_POLARIZATIONS)
      hipLaunchKernelGGL((k_response<double, GWL_HOST_POLARIZATIONS>), grid, dim3(256), 0, st,
                         ra);
    else
      hipLaunchKernelGGL((k_response<double, GWL_SINE_GAUSSIAN>), grid, dim3(256), 0, st, ra);
  }
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

}  // namespace gwl
