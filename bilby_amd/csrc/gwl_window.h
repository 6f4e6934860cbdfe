// Time marginalisation for the binned likelihoods (relative binning,
// multibanding) as a window transform on MFMA.
//
// Both likelihoods rebuild an integrand that is linear in a short per-row
// vector a (the edge ratios of relative binning, the banded-point products
// of multibanding), so the transform at the time-prior window is a fixed
// linear map of a:
//   X_j = sum_k a_k M[k][j],  j = j_first + jj,  jj < L,
// evaluated for a tile of rows as one GEMM against the basis operand
//   Bt[Np][Kp]: row jj = [Re M | -Im M], row L + jj = [Im M | Re M]
// (columns k < K and K + k, zero padding), with the rows' operand
//   A[Mp][Kp] = [Re a | Im a | 0].
// float32 operands use an MFMA kernel split over chunks of kWinKc operand
// columns whose float32 partial products the epilogue adds in float64, so
// that sums over thousands of terms keep float32 rounding of short partial
// sums only; float64 operands use the float64 GEMM.  The epilogue is the
// time-marginalised lnL of base.py:794-820 over the window.
#pragma once
#include "gwl_internal.h"

#include <cmath>

namespace gwl {

struct WinBasis {
  int K = 0;                 // complex basis length
  int Kp = 0;                // padded real operand width (>= 2K, multiple of 32)
  int L = 0;                 // window bins
  int Np = 0;                // padded 2L (multiple of 128)
  long long n_fft = 0;       // transform length (0: the event's N)
  long long j_first = 0;     // first transform bin
  double start = 0.0;        // time of transform bin j: start + (j + offset) t_step
  int offset = 0;
  double t_step = 0.0;
  double delta_tc = 0.0;     // time prior weight spacing (bilby's _delta_tc)
  double antenna_time = NAN; // antenna patterns at this time; NaN: the row's time
  void* data = nullptr;      // device: bt64 [Np][Kp] double, bt32 [Np][Kp] float
  size_t bytes = 0;
  double* bt64 = nullptr;
  float* bt32 = nullptr;
};

// Rows per GEMM tile of one evaluation call (bounds the workspace), and
// operand columns per split-K chunk of the float32 GEMM.
constexpr long long kWinTileRows = 8192;
constexpr int kWinKc = 1024;

void win_basis_free(WinBasis* wb);

// C[z][Mp][Np] (float32 partials) = A[Mp][Kp] . Bt[Np][Kp]^T over the
// column chunks [z kc, (z + 1) kc), z < ceil(Kp / kc), each summed in float32
// with a second accumulation level every 256 columns; the caller adds the
// chunks in float64.  Mp, Np multiples of 128; Kp, kc multiples of 32.
int win_gemm_f32(const float* A, const float* Bt, float* C, int Mp, int Np, int Kp, int kc,
                 hipStream_t st);

// Relative-binning basis: for detector d and edge e (k = d (nb + 1) + e),
//   M[k][j] = sum_n x_d[n] phi_e(n) exp(-2 pi i n j / N)
// over the grid bins n of the two bins adjacent to edge e, phi_e the hat
// function of the edge (linear in frequency, 1 at f_e, 0 at the
// neighbouring edges); ``x`` (device) is [nd][n_x] double2 from grid bin
// edges[0], ``edges`` (device) [nb + 1] grid indices, ``fe`` (device) the
// edge frequencies.
int win_basis_relbin(gwl_event* ev, int nd, int nb, const long long* edges, const double* fe,
                     const double2* x, long long n_x, WinBasis* wb);

// Banded-point basis: M[p][j] = exp(-2 pi i n_p j / N) for n_p >= 0, zero
// for n_p < 0; ``idx`` (device) [n].
int win_basis_points(gwl_event* ev, int n, const long long* idx, long long N, WinBasis* wb);

// Device workspace of one tile: A [Mp][Kp] (double-sized) then C: [Mp][Np]
// double (float64 GEMM) or [chunks][Mp][Np] float partials.
size_t win_ws_bytes(const WinBasis& wb, long long rows);
void* win_a(const WinBasis& wb, void* ws);
void* win_c(const WinBasis& wb, void* ws, long long rows);
int win_chunks(const WinBasis& wb);
long long win_mp(long long rows);

// Zero the operand padding: columns [2K, Kp) of rows < n and every column
// of rows [n, Mp).
int win_clear(const WinBasis& wb, void* a, bool fp64, long long n, hipStream_t st);

// GEMM of the operand tile against the basis, then the window epilogue:
// out[r] for r < n from C, the row's <h|h> (hh[3 r + 2]), validity
// (valid[r * valid_stride]), luminosity distance and time jitter (params).
// With ``series_dh`` set, the epilogue instead writes the transform itself,
// series_dh[(r L + jj) 2 + {0, 1}] = X_{j_first + jj}, and series_hh[r] =
// <h|h> (NaN for rows whose waveform failed); ``out`` is not written.
int win_finish(gwl_event* ev, const WinBasis& wb, const void* a, bool fp64, void* c,
               long long n, const double* params, int ld, const double* hh, const int* valid,
               int valid_stride, double* out, hipStream_t st, double* series_dh = nullptr,
               double* series_hh = nullptr);

}  // namespace gwl
