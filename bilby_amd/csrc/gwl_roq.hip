// Reduced-order-quadrature likelihood (bilby/gw/likelihood/roq.py) on gfx950.
//
//  k_roq_prep : per row, TaylorF2 at the linear and quadratic frequency nodes
//               (source.py:802-898), the antenna pattern and time delay at the
//               beam-pattern reference time (roq.py:475-512), <h|h> =
//               sum_q |h_q|^2 w_q (roq.py:504-507), and the GEMM operand
//               X = conj(h_linear) split into real and imaginary halves.
//  GEMM       : Z = X . W^T for all ROQ time samples of all detectors, the
//               one dense contraction (roq.py:631-643 "weights_linear @
//               h_linear_conj"), on MFMA: float32 (v_mfma_f32_32x32x2f32) or
//               split-bfloat16 x3 (v_mfma_f32_32x32x16_bf16, A = Ah + Al,
//               B = Bh + Bl, Z ~ Ah Bh + Ah Bl + Al Bh), or a float64 FMA
//               reference.  The complex product is the real GEMM
//               [Xr Xi] [[Wr Wi] [-Wi Wr]].
//  k_roq_epi  : five-sample cubic interpolation in time (roq.py:576-651,
//               T2100224), detector sum, time / phase / distance
//               marginalisation (base.py:775-820), float64.
#include "gwl_internal.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <new>
#include <vector>

namespace gwl {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct RoqDev {
  int nd, L, Qn, NT, ntimes;
  int Kp, Np;  // GEMM K (2 L, multiple of 32) and N (2 nd NT padded to 128)
  double start, ts0, dts, tref, tcmin, dtc, tmin, tmax, sigma, fref, wf_fmin, dref;
  double idts;  // 1 / dts
  unsigned flags;
  double tensor[GWL_MAX_DET][9];
  double vertex[GWL_MAX_DET][3];
  const double* fl;   // [L] linear nodes
  const double* fq;   // [Qn] quadratic nodes
  const double* wq;   // [nd][Qn] quadratic weights (real part) * sigma^2
  const float* bt32;  // [Np][Kp] W^T real embedding, float32
  const __bf16* bth;  // [Np][Kp] high bfloat16 part
  const __bf16* btl;  // [Np][Kp] low bfloat16 part
  const double* bt64; // [Np][Kp] float64
  const double* ts;   // [NT] ROQ time samples (s from start)
  const double* nbl;  // [L][kNB + 1] TaylorF2 basis (tf2_basis order) and f^-7/6 at the linear nodes
  const double* nbq;  // [Qn][kNB + 1] the same at the quadratic nodes
  SplineDev sp;
  // time prior other than Uniform: density table (null: Uniform)
  const double* tpd;
  double tp_t0, tp_ih;
  int tp_n;
};

struct RoqRow {
  double cre[GWL_MAX_DET], cim[GWL_MAX_DET];  // conj(C_d)
  double dt[GWL_MAX_DET];                     // time delay at the reference time
  double dtg;                                 // geocent_time - start (incl. jitter)
  double hh;                                  // <h|h> (all detectors)
  double jitter;
  double hscale;
  int valid;
  int pad;
};

// TaylorF2 constants of a row for the node waveforms (k_roq_rows -> k_roq_prep)
struct RoqTf2 {
  double c[10];
  double fisco;
  int valid;
  int pad;
};

}  // namespace gwl

struct gwl_roq {
  gwl_ctx* ctx;
  gwl::RoqDev dev;
  void* data = nullptr;
  size_t data_bytes = 0;
  double* spline = nullptr;
  bool has_spline = false;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  double* tprior = nullptr;
};

namespace gwl {

enum { ROQ_FP64 = 0, ROQ_FP32 = 1, ROQ_BF16X3 = 2 };

// ------------------------------------------------------------------ prep
// One workgroup (128 threads) per row.  A layout [Mpad][Kp]: column l < L
// holds Re conj(h0(f_l)) = Re h0, column L + l holds Im conj(h0) = -Im h0,
// where the detector response is h_d = C_d h0, h0 = f^-7/6 e^{-i Psi(f)} and
// the amplitude prefactor lives in C_d.
// One thread per row: the row's TaylorF2 constants, antenna patterns and
// delays (what every node of the row shares), so that k_roq_prep's
// workgroups start from them instead of one lane computing them while the
// others wait.
__global__ __launch_bounds__(128) void k_roq_rows(RoqDev rd, const double* params, int ld,
                                                  long long n, RoqRow* rows, RoqTf2* tf2) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double* row = params + r * (long long)ld;
  RoqTf2 t;
  double amp0;
  const int valid = tf2_row_constants(row, rd.fref, rd.sigma, t.c, amp0, t.fisco);
  t.valid = valid;
  t.pad = 0;
  // antenna and delay at the reference time when time-marginalised
  // (roq.py:475-478), else at geocent_time
  const bool timem = (rd.flags & GWL_TIME_MARG) != 0;
  const bool jit = timem && (rd.flags & GWL_JITTER_TIME);
  const double jitter = jit ? row[GWL_P_TIME_JITTER] : 0.0;
  const double gt = row[GWL_P_GEOCENT_TIME] + jitter;  // base.py:427-428
  const double tref = timem ? rd.tref : gt;
  const double ci = cos(row[GWL_P_THETA_JN]);
  const double ap = amp0 * (1.0 + ci * ci) / 2.0, ax = amp0 * ci;
  RoqRow rr;
  const SkyBasis sky = sky_basis(row[GWL_P_RA], row[GWL_P_DEC], tref, row[GWL_P_PSI]);
  for (int d = 0; d < GWL_MAX_DET; ++d) rr.cre[d] = rr.cim[d] = rr.dt[d] = 0.0;
  for (int d = 0; d < rd.nd; ++d) {
    double fp, fc, delay;
    detector_projection(rd.tensor[d], rd.vertex[d], sky, fp, fc, delay);
    // C_d = F+ A+ - i Fx Ax; store conj(C_d)
    rr.cre[d] = fp * ap;
    rr.cim[d] = fc * ax;
    rr.dt[d] = delay;
  }
  rr.dtg = gt - rd.start;
  rr.jitter = jitter;
  rr.hscale = (rd.flags & GWL_DISTANCE_MARG) ? row[GWL_P_LUMINOSITY_DISTANCE] / rd.dref : 1.0;
  rr.valid = valid;
  rr.pad = 0;
  rr.hh = 0.0;
  rows[r] = rr;
  tf2[r] = t;
}

template <int MODE>
__global__ __launch_bounds__(128) void k_roq_prep(RoqDev rd, const RoqTf2* tf2, long long n,
                                                   void* aout, void* aout_lo, RoqRow* rows) {
  const long long r = blockIdx.x;
  __shared__ double c[10];
  __shared__ double s_fisco;
  __shared__ int s_valid;
  __shared__ double hpart[GWL_MAX_DET][2];
  const int tid = threadIdx.x;
  const bool live = r < n;
  if (tid < 10) c[tid] = live ? tf2[r].c[tid] : 0.0;
  if (tid == 0) {
    s_fisco = live ? tf2[r].fisco : 0.0;
    s_valid = live ? tf2[r].valid : 0;
    for (int d = 0; d < GWL_MAX_DET; ++d) hpart[d][0] = hpart[d][1] = 0.0;
  }
  __syncthreads();
  const bool valid = live && s_valid;
  const double fisco = s_fisco;
  // lalsimulation evaluates a frequency-sequence TaylorF2 at every node
  // (source.py:1125-1128); the waveform band and the ISCO cut the nodes only
  // when the likelihood asks for ISCO termination
  const bool cut = (rd.flags & GWL_TF2_END_ISCO) != 0;
  const size_t o = (size_t)r * rd.Kp;
  // padding columns [2L, Kp) of every row, and rows past n, are zero
  for (int k = 2 * rd.L + tid; k < rd.Kp; k += 128) {
    if (MODE == ROQ_FP64) reinterpret_cast<double*>(aout)[o + k] = 0.0;
    else if (MODE == ROQ_FP32) reinterpret_cast<float*>(aout)[o + k] = 0.f;
    else {
      reinterpret_cast<__bf16*>(aout)[o + k] = (__bf16)0.f;
      reinterpret_cast<__bf16*>(aout_lo)[o + k] = (__bf16)0.f;
    }
  }
  // linear nodes -> GEMM operand; the basis at each node is tabled
  for (int l = tid; l < rd.L; l += 128) {
    double xr = 0.0, xi = 0.0;
    if (valid) {
      const double f = rd.fl[l];
      if (!cut || (f >= rd.wf_fmin && f <= fisco)) {
        const double* b = rd.nbl + (size_t)l * (kNB + 1);
        const double cyc = tf2_phase_cycles(c, b);
        const double amp = b[kNB];  // f^-7/6; amp0 sits in C_d
        if (MODE == ROQ_FP64) {
          const double2 e = expm2pi(cyc - floor(cyc));
          xr = amp * e.x;   // Re conj(h0) = Re h0
          xi = -amp * e.y;  // Im conj(h0) = -Im h0 = +amp sin
        } else {
          // float32 operand: the rotation in float32 from the reduced phase
          const float2 e = expm2pi((float)(cyc - floor(cyc)));
          xr = amp * (double)e.x;
          xi = -amp * (double)e.y;
        }
      }
    }
    const int kl = l, ki = l + rd.L;
    if (MODE == ROQ_FP64) {
      double* a = reinterpret_cast<double*>(aout);
      a[o + kl] = xr;
      a[o + ki] = xi;
    } else if (MODE == ROQ_FP32) {
      float* a = reinterpret_cast<float*>(aout);
      a[o + kl] = (float)xr;
      a[o + ki] = (float)xi;
    } else {
      __bf16* ah = reinterpret_cast<__bf16*>(aout);
      __bf16* al = reinterpret_cast<__bf16*>(aout_lo);
      const float fr = (float)xr, fi = (float)xi;
      const __bf16 hr = (__bf16)fr, hi = (__bf16)fi;
      ah[o + kl] = hr;
      ah[o + ki] = hi;
      al[o + kl] = (__bf16)(fr - (float)hr);
      al[o + ki] = (__bf16)(fi - (float)hi);
    }
  }
  // quadratic nodes -> <h|h>_d = |C_d|^2 sum_q |h0_q|^2 w_dq
  if (valid) {
    double acc[GWL_MAX_DET] = {0.0, 0.0, 0.0, 0.0};
    for (int q = tid; q < rd.Qn; q += 128) {
      const double f = rd.fq[q];
      if (!cut || (f >= rd.wf_fmin && f <= fisco)) {
        const double amp = rd.nbq[(size_t)q * (kNB + 1) + kNB];
        for (int d = 0; d < rd.nd; ++d) acc[d] += amp * amp * rd.wq[(size_t)d * rd.Qn + q];
      }
    }
    for (int d = 0; d < rd.nd; ++d) {
      double v = acc[d];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if ((tid & 63) == 0) hpart[d][tid >> 6] = v;
    }
  }
  __syncthreads();
  if (tid == 0 && live) {
    double hh = 0.0;
    if (valid)
      for (int d = 0; d < rd.nd; ++d) {
        const double c2 = rows[r].cre[d] * rows[r].cre[d] + rows[r].cim[d] * rows[r].cim[d];
        hh += c2 * (hpart[d][0] + hpart[d][1]);
      }
    rows[r].hh = hh;
  }
}

// TaylorF2 basis (tf2_basis order) and f^-7/6 at n frequency nodes.
__global__ void k_roq_node_table(const double* f, int n, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double b[kNB];
  tf2_basis(f[i], b);
  for (int k = 0; k < kNB; ++k) out[(size_t)i * (kNB + 1) + k] = b[k];
  out[(size_t)i * (kNB + 1) + kNB] = b[1] * sqrt(b[3]);
}

// ------------------------------------------------------------------ GEMM kernels
// C[M][N] = A[M][K] . Bt[N][K]^T; M % 128 == 0, N % 128 == 0, K % 32 == 0.
// 256 threads = 4 waves; a workgroup computes a 128 x 128 tile, each wave a
// 64 x 64 block as 2 x 2 MFMA 32 x 32 outputs (16 accumulators per lane
// each).  32 x 32 output layout: lane l, accumulator v -> row
// 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32.
constexpr int kBM = 128, kBN = 128, kBK = 32;

__global__ __launch_bounds__(256) void k_gemm_f32(const float* __restrict__ A,
                                                  const float* __restrict__ Bt,
                                                  float* __restrict__ C, int M, int N, int K) {
  __shared__ float As[kBM][kBK + 1];
  __shared__ float Bs[kBN][kBK + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
  for (int k0 = 0; k0 < K; k0 += kBK) {
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
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = m0 + wr + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int col = n0 + wc + j * 32 + (lane & 31);
        C[(size_t)row * N + col] = acc[i][j][v];
      }
}

// Split-bfloat16 x3: per k16 step and 32 x 32 block, Ah Bh + Ah Bl + Al Bh.
// A fragment (lane l): row l % 32, k = 8 (l / 32) .. +7; B likewise.
constexpr int kPadH = 8;  // bf16 row padding in LDS

__global__ __launch_bounds__(256) void k_gemm_bf16x3(const __bf16* __restrict__ Ah,
                                                     const __bf16* __restrict__ Al,
                                                     const __bf16* __restrict__ Bh,
                                                     const __bf16* __restrict__ Bl,
                                                     float* __restrict__ C, int M, int N, int K,
                                                     int kcol) {
  __shared__ __attribute__((aligned(16))) __bf16 sAh[kBM][kBK + kPadH];
  __shared__ __attribute__((aligned(16))) __bf16 sAl[kBM][kBK + kPadH];
  __shared__ __attribute__((aligned(16))) __bf16 sBh[kBN][kBK + kPadH];
  __shared__ __attribute__((aligned(16))) __bf16 sBl[kBN][kBK + kPadH];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
  // K chunk blockIdx.z (columns [z kcol, (z + 1) kcol)) into partial plane z
  const int kb = blockIdx.z * kcol, ke = min(K, kb + kcol);
  C += (size_t)blockIdx.z * M * N;
  for (int k0 = kb; k0 < ke; k0 += kBK) {
    // 128 rows x 32 bf16 = 512 chunks of 8 per operand: 2 per thread
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int idx = tid + 256 * q;
      const int row = idx >> 2, c8 = (idx & 3) * 8;
      const size_t ga = (size_t)(m0 + row) * K + k0 + c8;
      const size_t gb = (size_t)(n0 + row) * K + k0 + c8;
      *reinterpret_cast<uint4*>(&sAh[row][c8]) = *reinterpret_cast<const uint4*>(Ah + ga);
      *reinterpret_cast<uint4*>(&sAl[row][c8]) = *reinterpret_cast<const uint4*>(Al + ga);
      *reinterpret_cast<uint4*>(&sBh[row][c8]) = *reinterpret_cast<const uint4*>(Bh + gb);
      *reinterpret_cast<uint4*>(&sBl[row][c8]) = *reinterpret_cast<const uint4*>(Bl + gb);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 16) {
      bf16x8 ah[2], al[2], bh[2], bl[2];
      const int kc = kk + 8 * (lane >> 5);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int rr = wr + i * 32 + (lane & 31);
        ah[i] = *reinterpret_cast<const bf16x8*>(&sAh[rr][kc]);
        al[i] = *reinterpret_cast<const bf16x8*>(&sAl[rr][kc]);
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int cc = wc + j * 32 + (lane & 31);
        bh[j] = *reinterpret_cast<const bf16x8*>(&sBh[cc][kc]);
        bl[j] = *reinterpret_cast<const bf16x8*>(&sBl[cc][kc]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = m0 + wr + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int col = n0 + wc + j * 32 + (lane & 31);
        C[(size_t)row * N + col] = acc[i][j][v];
      }
}

// The same products with 128 x 256 workgroup tiles: each wave computes a
// 64 x 128 block (2 x 4 MFMA 32 x 32 outputs), so a fragment read from LDS
// feeds four MFMAs instead of two and the LDS traffic per MFMA (operand
// writes + fragment reads) falls from 1.5 KB to 1.1 KB (measured: 10 % less
// time than the 128 x 128 tiles on the ROQ shapes).  N multiple of 256.
constexpr int kBNW = 256;

__global__ __launch_bounds__(256) void k_gemm_bf16x3_w(const __bf16* __restrict__ Ah,
                                                       const __bf16* __restrict__ Al,
                                                       const __bf16* __restrict__ Bh,
                                                       const __bf16* __restrict__ Bl,
                                                       float* __restrict__ C, int M, int N,
                                                       int K, int kcol) {
  __shared__ __attribute__((aligned(16))) __bf16 sAh[kBM][kBK + kPadH];
  __shared__ __attribute__((aligned(16))) __bf16 sAl[kBM][kBK + kPadH];
  __shared__ __attribute__((aligned(16))) __bf16 sBh[kBNW][kBK + kPadH];
  __shared__ __attribute__((aligned(16))) __bf16 sBl[kBNW][kBK + kPadH];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBNW;
  const int wr = (w >> 1) * 64, wc = (w & 1) * 128;
  f32x16 acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
  // K chunk blockIdx.z (columns [z kcol, (z + 1) kcol)) into partial plane z
  const int kb = blockIdx.z * kcol, ke = min(K, kb + kcol);
  C += (size_t)blockIdx.z * M * N;
  for (int k0 = kb; k0 < ke; k0 += kBK) {
    // A: 128 rows x 32 bf16 = 512 chunks of 8 per operand (2 per thread);
    // B: 256 rows (4 per thread)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int idx = tid + 256 * q;
      const int row = idx >> 2, c8 = (idx & 3) * 8;
      const size_t ga = (size_t)(m0 + row) * K + k0 + c8;
      *reinterpret_cast<uint4*>(&sAh[row][c8]) = *reinterpret_cast<const uint4*>(Ah + ga);
      *reinterpret_cast<uint4*>(&sAl[row][c8]) = *reinterpret_cast<const uint4*>(Al + ga);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + 256 * q;
      const int row = idx >> 2, c8 = (idx & 3) * 8;
      const size_t gb = (size_t)(n0 + row) * K + k0 + c8;
      *reinterpret_cast<uint4*>(&sBh[row][c8]) = *reinterpret_cast<const uint4*>(Bh + gb);
      *reinterpret_cast<uint4*>(&sBl[row][c8]) = *reinterpret_cast<const uint4*>(Bl + gb);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 16) {
      bf16x8 ah[2], al[2];
      const int kc = kk + 8 * (lane >> 5);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int rr = wr + i * 32 + (lane & 31);
        ah[i] = *reinterpret_cast<const bf16x8*>(&sAh[rr][kc]);
        al[i] = *reinterpret_cast<const bf16x8*>(&sAl[rr][kc]);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int cc = wc + j * 32 + (lane & 31);
        const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&sBh[cc][kc]);
        const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&sBl[cc][kc]);
        // per output the same order as k_gemm_bf16x3: lo.hi, hi.lo, hi.hi
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh, acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl, acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh, acc[i][j], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = m0 + wr + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int col = n0 + wc + j * 32 + (lane & 31);
        C[(size_t)row * N + col] = acc[i][j][v];
      }
}

// float64 reference GEMM: 64 x 64 tile, 256 threads x 4 x 4 outputs.
typedef double f64x4 __attribute__((ext_vector_type(4)));

// C[M][N] = A[M][K] . Bt[N][K]^T in float64 on the matrix cores
// (v_mfma_f64_16x16x4f64): 64 x 64 tiles, four wavefronts of 2 x 2 16 x 16
// blocks, K staged through LDS 32 columns at a time, each slab's loads all
// issued before its LDS writes (16 columns, or loads and writes paired
// element by element, left the MFMAs idle between slabs; this form has 1.5x
// the throughput).  Lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]
// of a 16 x 4 step and accumulator r holds D[4 r + l / 16][l % 16].  M, N
// multiples of 64, K of 32.
__global__ __launch_bounds__(256) void k_gemm_f64(const double* __restrict__ A,
                                                  const double* __restrict__ Bt,
                                                  double* __restrict__ C, int M, int N, int K) {
  __shared__ double As[64][32 + 1];
  __shared__ double Bs[64][32 + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  f64x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
  // a slab (64 rows x 32 k per operand, 8 elements per thread) is read into
  // registers, all loads in flight at once, then written to LDS
  double ra[8], rb[8];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int idx = tid + 256 * q;
      const int row = idx >> 5, kk = idx & 31;
      ra[q] = A[(size_t)(m0 + row) * K + k0 + kk];
      rb[q] = Bt[(size_t)(n0 + row) * K + k0 + kk];
    }
  };
  load(0);
  for (int k0 = 0; k0 < K; k0 += 32) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int idx = tid + 256 * q;
      const int row = idx >> 5, kk = idx & 31;
      As[row][kk] = ra[q];
      Bs[row][kk] = rb[q];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 32; kk += 4) {
      double a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[wr + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[wc + j * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
    if (k0 + 32 < K) load(k0 + 32);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        C[(size_t)(m0 + wr + i * 16 + 4 * r + (lane >> 4)) * N + n0 + wc + j * 16 + (lane & 15)] =
            acc[i][j][r];
}

// ------------------------------------------------------------------ epilogue
// d_inner_h at time t from the five nearest ROQ time samples
// (roq.py:576-602 / 636-651): a v0 + b v1 + c r1 + d r2, with the weight
// of the sample after t, b = 1 - a, given by the caller.
template <typename T>
__device__ inline void five_sample(const T* z, int base_c, double a, double& re, double& im) {
  // z: complex row (interleaved), base_c = complex index of sample closest - 2
  using C = typename cplx_t<T>::type;
  const C* zc = reinterpret_cast<const C*>(z) + base_c;
  double vr[5], vi[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const C v = zc[k];
    vr[k] = (double)v.x;
    vi[k] = (double)v.y;
  }
  constexpr double kSixth = 1.0 / 6.0;
  const double b = 1.0 - a;
  const double c = (a * a * a - a) * kSixth, d = (b * b * b - b) * kSixth;
  const double r1r = (-vr[0] + 8.0 * vr[1] - 14.0 * vr[2] + 8.0 * vr[3] - vr[4]) / 4.0;
  const double r1i = (-vi[0] + 8.0 * vi[1] - 14.0 * vi[2] + 8.0 * vi[3] - vi[4]) / 4.0;
  const double r2r = vr[2] - 2.0 * vr[3] + vr[4], r2i = vi[2] - 2.0 * vi[3] + vi[4];
  re = a * vr[2] + b * vr[3] + c * r1r + d * r2r;
  im = a * vi[2] + b * vi[3] + c * r1i + d * r2i;
}

constexpr int kRoqEpiRows = 4;

// Time-marginalisation grid point j (roq.py:320-331), rounded after each
// operation as numpy does.  The grid is in absolute GPS seconds (~1e9, one
// ulp = 2.4e-7 s) and the segment start is subtracted afterwards, so a fused
// multiply-add would move a point by up to one ulp, about 1e-3 of a ROQ time
// sample, and change the interpolated d_inner_h.
__device__ inline double roq_grid_time(double tcmin, double dtc, int j) {
#pragma clang fp contract(off)
  return (tcmin + dtc / 2.0) + (double)j * dtc;
}

// The interpolated, conj(C_d)-weighted detector sum of <d|h> at ifo time
// (toff + dt_d) + extra for each detector; false where the five samples
// leave the weights' time grid.  One time (roq.py:512-526):
// a = (t_{c+1} - t) / (t_{c-1} - t_{c-2}) on the samples themselves; the
// time-marginalisation array (roq.py:636-651): b = (t - t_0) / dt - c.
template <typename T>
__device__ __forceinline__ bool roq_dh_at(const RoqDev& rd, const T* zr, const RoqRow& rr,
                                          double toff, double extra, bool single, double& sr,
                                          double& si) {
  sr = si = 0.0;
#pragma unroll
  for (int d = 0; d < GWL_MAX_DET; ++d) {
    if (d >= rd.nd) break;
    const double ifo_t = (toff + rr.dt[d]) + extra;
    const double x = (ifo_t - rd.ts0) * rd.idts;
    const double cl = floor(x);
    const long long ci = (long long)cl;
    if (ci - 2 < 0 || ci + 2 >= rd.NT) return false;
    double a;
    if (single) {
      const int c = (int)ci;
      a = (rd.ts[c + 1] - ifo_t) / fmax(rd.ts[c - 1] - rd.ts[c - 2], 1e-12);
    } else {
      a = 1.0 - (x - cl);
    }
    double vr, vi;
    five_sample(zr, d * rd.NT + (int)ci - 2, a, vr, vi);
    // Y_d = conj(C_d) Z_d
    sr += rr.cre[d] * vr - rr.cim[d] * vi;
    si += rr.cre[d] * vi + rr.cim[d] * vr;
  }
  return true;
}

template <typename T>
__global__ __launch_bounds__(64 * kRoqEpiRows) void k_roq_epi(RoqDev rd, const T* Z,
                                                              const RoqRow* rows, long long n,
                                                              double* out) {
  extern __shared__ double knots[];
  const bool dist = (rd.flags & GWL_DISTANCE_MARG) != 0;
  if (dist) {
    const int nx4 = rd.sp.nx + 4, ny4 = rd.sp.ny + 4;
    for (int i = threadIdx.x; i < nx4; i += blockDim.x) knots[i] = rd.sp.tx[i];
    for (int i = threadIdx.x; i < ny4; i += blockDim.x) knots[nx4 + i] = rd.sp.ty[i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long long r = (long long)blockIdx.x * kRoqEpiRows + (threadIdx.x >> 6);
  if (r >= n) return;
  const RoqRow rr = rows[r];
  if (!rr.valid) {
    if (lane == 0) out[r] = -DBL_MAX;  // base.py:424-425
    return;
  }
  const T* zr = Z + (size_t)r * rd.Np;
  const double hs = rr.hscale;
  const double hh = rr.hh;
  double yy_by[4] = {0, 0, 0, 0};
  int yy_ly = 3;
  bool yy_in = false;
  if (dist) {
    const double y = hh * hs * hs;
    yy_in = (y >= rd.sp.ymin && y <= rd.sp.ymax);
    if (yy_in) {
      yy_ly = knot_interval(knots + rd.sp.nx + 4, rd.sp.ny, y);
      bspline3(knots + rd.sp.nx + 4, yy_ly, y, yy_by);
    }
  }
  auto lnl_of = [&](double xr, double xi) -> double {
    if (dist) {
      const double d = (rd.flags & GWL_PHASE_MARG) ? hypot(xr, xi) * hs : xr * hs;
      if (!yy_in || !(d >= rd.sp.xmin && d <= rd.sp.xmax)) return -INFINITY;
      const int lx = knot_interval(knots, rd.sp.nx, d);
      double bx[4];
      bspline3(knots, lx, d, bx);
      const double* c = rd.sp.c + (size_t)(lx - 3) * rd.sp.ny + (yy_ly - 3);
      double s = 0.0;
      for (int i = 0; i < 4; ++i) {
        const double* ci = c + (size_t)i * rd.sp.ny;
        s += bx[i] * (yy_by[0] * ci[0] + yy_by[1] * ci[1] + yy_by[2] * ci[2] + yy_by[3] * ci[3]);
      }
      return s;
    }
    if (rd.flags & GWL_PHASE_MARG) return ln_i0(hypot(xr, xi)) - hh / 2.0;
    return xr - hh / 2.0;
  };
  auto dh_at = [&](double toff, double extra, bool single, bool& ok, double& sr, double& si) {
    ok = roq_dh_at(rd, zr, rr, toff, extra, single, sr, si);
  };
  if (!(rd.flags & GWL_TIME_MARG)) {
    if (lane == 0) {
      bool ok;
      double sr, si;
      dh_at(rr.dtg, 0.0, true, ok, sr, si);
      // out-of-range time: log(in_bounds) = -inf (roq.py:529-532)
      out[r] = ok ? lnl_of(sr, si) : -INFINITY;
    }
    return;
  }
  double m = -INFINITY, s = 0.0;
  for (int j = lane; j < rd.ntimes; j += 64) {
    const double tj = roq_grid_time(rd.tcmin, rd.dtc, j);
    const double t = tj + rr.jitter;
    if (!(t >= rd.tmin && t <= rd.tmax)) continue;
    bool ok;
    double sr, si;
    // roq.py:537-540: ifo_times = (times - start) + dt + jitter
    dh_at(tj - rd.start, rr.jitter, false, ok, sr, si);
    double v = ok ? lnl_of(sr, si) : -INFINITY;
    if (rd.tpd) v += time_log_density(rd.tpd, rd.tp_t0, rd.tp_ih, rd.tp_n, t);
    if (v > m) {
      s = (m > -INFINITY ? s * exp(m - v) : 0.0) + 1.0;
      m = v;
    } else if (v > -INFINITY) {
      s += exp(v - m);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double mo = __shfl_xor(m, off, 64), so = __shfl_xor(s, off, 64);
    const double mn = fmax(m, mo);
    if (mn > -INFINITY)
      s = (m > -INFINITY ? s * exp(m - mn) : 0.0) + (mo > -INFINITY ? so * exp(mo - mn) : 0.0);
    else
      s = 0.0;
    m = mn;
  }
  if (lane == 0) {
    const double b = rd.tpd ? rd.dtc : rd.dtc / (rd.tmax - rd.tmin);
    out[r] = (m > -INFINITY) ? log(b * s) + m : -INFINITY;
  }
}

// One thread per (row, output): <d|h> at the row's own time (n_times = 0,
// roq.py:512-526) or on the time-marginalisation grid (roq.py:536-540,
// 604-651), NaN where the five samples leave the weights' grid or the
// waveform failed; and the row's <h|h>.
template <typename T>
__global__ __launch_bounds__(256) void k_roq_series(RoqDev rd, const T* Z, const RoqRow* rows,
                                                    long long n, int n_times, double* dh,
                                                    double* hh) {
  const int nout = n_times > 0 ? n_times : 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nout) return;
  const long long r = t / nout;
  const int j = (int)(t - r * nout);
  const RoqRow rr = rows[r];
  double sr = NAN, si = NAN;
  if (rr.valid) {
    const T* zr = Z + (size_t)r * rd.Np;
    const bool ok = n_times > 0
                        ? roq_dh_at(rd, zr, rr, roq_grid_time(rd.tcmin, rd.dtc, j) - rd.start,
                                    rr.jitter, false, sr, si)
                        : roq_dh_at(rd, zr, rr, rr.dtg, 0.0, true, sr, si);
    if (!ok) sr = si = NAN;
  }
  dh[2 * t] = sr;
  dh[2 * t + 1] = si;
  if (j == 0) hh[r] = rr.valid ? rr.hh : NAN;
}

}  // namespace gwl

using namespace gwl;

static size_t roq_ws_bytes(const RoqDev& rd, int64_t n, int mode) {
  const size_t Mp = (size_t)((n + kBM - 1) / kBM) * kBM;
  const size_t a = mode == ROQ_FP64 ? Mp * rd.Kp * 8 : Mp * rd.Kp * 4;  // bf16: two halves
  const size_t c = mode == ROQ_FP64 ? Mp * rd.Np * 8 : Mp * rd.Np * 4;
  return a + c + (size_t)n * sizeof(RoqRow) + (size_t)n * sizeof(RoqTf2) + 4 * 256;
}

// C[z][M][N] float32 partials of Ah/Al . Bh/Bl^T (split bfloat16 x3) over the
// column chunks [z kcol, (z + 1) kcol); M, N multiples of 128, K and kcol
// multiples of 32.
int gwl::gemm_bf16x3_split(const __bf16* Ah, const __bf16* Al, const __bf16* Bh,
                           const __bf16* Bl, float* C, int M, int N, int K, int kcol,
                           hipStream_t st) {
  if (M % kBM || N % kBN || K % kBK || kcol % kBK || kcol <= 0) {
    set_error("gemm_bf16x3_split: shapes must be multiples of the tile");
    return GWL_E_INVALID;
  }
  const int nz = (K + kcol - 1) / kcol;
  if (N % kBNW == 0)
    hipLaunchKernelGGL(k_gemm_bf16x3_w,
                       dim3((unsigned)(N / kBNW), (unsigned)(M / kBM), (unsigned)nz), dim3(256), 0,
                       st, Ah, Al, Bh, Bl, C, M, N, K, kcol);
  else
    hipLaunchKernelGGL(k_gemm_bf16x3,
                       dim3((unsigned)(N / kBN), (unsigned)(M / kBM), (unsigned)nz), dim3(256), 0,
                       st, Ah, Al, Bh, Bl, C, M, N, K, kcol);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

// Workspace, node waveforms (k_roq_prep) and the weight GEMM of one call:
// on return *C holds the products Z [Mp][Np] and *rows the row constants.
static int roq_prep_gemm(gwl_roq* e, const double* params, int64_t n, int32_t ld, int32_t mode,
                         hipStream_t st, const void** C, RoqRow** rows_out) {
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  const size_t need = roq_ws_bytes(e->dev, n, ROQ_FP64) > roq_ws_bytes(e->dev, n, mode)
                          ? roq_ws_bytes(e->dev, n, ROQ_FP64) : roq_ws_bytes(e->dev, n, mode);
  if (need > e->ws_bytes) {
    if (e->ws) (void)hipFree(e->ws);
    e->ws = nullptr;
    e->ws_bytes = 0;
    if (hipMalloc(&e->ws, need) != hipSuccess) {
      set_error("hipMalloc failed for ROQ workspace");
      return GWL_E_NOMEM;
    }
    e->ws_bytes = need;
  }
  (void)hipGetLastError();
  const RoqDev& rd = e->dev;
  const int Mp = (int)(((n + kBM - 1) / kBM) * kBM);
  char* p = reinterpret_cast<char*>(e->ws);
  const size_t abytes = (mode == ROQ_FP64 ? (size_t)Mp * rd.Kp * 8 : (size_t)Mp * rd.Kp * 4);
  void* A = p;
  void* Alo = p + (size_t)Mp * rd.Kp * 2;  // bf16 low half follows the high half
  char* Cp = p + (abytes + 255) / 256 * 256;
  const size_t cbytes = (mode == ROQ_FP64 ? (size_t)Mp * rd.Np * 8 : (size_t)Mp * rd.Np * 4);
  RoqRow* rows = reinterpret_cast<RoqRow*>(Cp + (cbytes + 255) / 256 * 256);
  RoqTf2* tf2 = reinterpret_cast<RoqTf2*>(reinterpret_cast<char*>(rows) +
                                          ((size_t)n * sizeof(RoqRow) + 255) / 256 * 256);
  hipLaunchKernelGGL(k_roq_rows, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, rd, params,
                     ld, (long long)n, rows, tf2);
  // one block per operand row: rows past n and padding columns are written as zeros
  if (mode == ROQ_FP64)
    hipLaunchKernelGGL(k_roq_prep<ROQ_FP64>, dim3((unsigned)Mp), dim3(128), 0, st, rd, tf2, n, A, nullptr, rows);
  else if (mode == ROQ_FP32)
    hipLaunchKernelGGL(k_roq_prep<ROQ_FP32>, dim3((unsigned)Mp), dim3(128), 0, st, rd, tf2, n, A, nullptr, rows);
  else
    hipLaunchKernelGGL(k_roq_prep<ROQ_BF16X3>, dim3((unsigned)Mp), dim3(128), 0, st, rd, tf2, n, A, Alo, rows);
  GWL_HIP_CHECK(hipGetLastError());
  if (mode == ROQ_FP64) {
    hipLaunchKernelGGL(k_gemm_f64, dim3(rd.Np / 64, Mp / 64), dim3(256), 0, st,
                       reinterpret_cast<const double*>(A), rd.bt64, reinterpret_cast<double*>(Cp),
                       Mp, rd.Np, rd.Kp);
  } else if (mode == ROQ_FP32) {
    hipLaunchKernelGGL(k_gemm_f32, dim3(rd.Np / kBN, Mp / kBM), dim3(256), 0, st,
                       reinterpret_cast<const float*>(A), rd.bt32, reinterpret_cast<float*>(Cp),
                       Mp, rd.Np, rd.Kp);
  } else if (rd.Np % kBNW == 0) {
    hipLaunchKernelGGL(k_gemm_bf16x3_w, dim3(rd.Np / kBNW, Mp / kBM), dim3(256), 0, st,
                       reinterpret_cast<const __bf16*>(A), reinterpret_cast<const __bf16*>(Alo),
                       rd.bth, rd.btl, reinterpret_cast<float*>(Cp), Mp, rd.Np, rd.Kp, rd.Kp);
  } else {
    hipLaunchKernelGGL(k_gemm_bf16x3, dim3(rd.Np / kBN, Mp / kBM), dim3(256), 0, st,
                       reinterpret_cast<const __bf16*>(A), reinterpret_cast<const __bf16*>(Alo),
                       rd.bth, rd.btl, reinterpret_cast<float*>(Cp), Mp, rd.Np, rd.Kp, rd.Kp);
  }
  GWL_HIP_CHECK(hipGetLastError());
  *C = Cp;
  *rows_out = rows;
  return GWL_OK;
}

extern "C" {

int gwl_roq_create(gwl_ctx* ctx, const gwl_roq_desc* d, gwl_roq** out, int64_t* hbm_bytes) {
  if (!ctx || !d || !out) { set_error("gwl_roq_create: NULL argument"); return GWL_E_INVALID; }
  if (d->n_det < 1 || d->n_det > GWL_MAX_DET || d->n_linear < 1 || d->n_quadratic < 1 ||
      d->n_time_samples < 5) {
    set_error("gwl_roq_create: invalid sizes");
    return GWL_E_INVALID;
  }
  if (!d->linear_nodes || !d->quadratic_nodes || !d->linear_weights || !d->quadratic_weights ||
      !d->detector_tensor || !d->vertex || !d->time_samples) {
    set_error("gwl_roq_create: NULL array in the descriptor");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(ctx->device));
  RoqDev rd{};
  rd.nd = d->n_det;
  rd.L = d->n_linear;
  rd.Qn = d->n_quadratic;
  rd.NT = d->n_time_samples;
  rd.Kp = ((2 * rd.L + kBK - 1) / kBK) * kBK;
  rd.Np = ((2 * rd.nd * rd.NT + kBN - 1) / kBN) * kBN;
  rd.start = d->start_time;
  rd.ts0 = d->time_samples_start;
  rd.dts = d->time_samples_step;
  rd.idts = 1.0 / rd.dts;
  rd.flags = d->flags;
  rd.fref = d->reference_frequency > 0 ? d->reference_frequency : 20.0;
  rd.wf_fmin = d->wf_minimum_frequency;
  rd.dref = d->reference_distance > 0 ? d->reference_distance : 1.0;
  rd.tmin = d->time_prior_minimum;
  rd.tmax = d->time_prior_maximum;
  if (rd.flags & GWL_TIME_MARG) {
    // roq.py:320-331
    if (!(rd.tmax > rd.tmin) || !(d->delta_tc > 0)) {
      set_error("gwl_roq_create: time marginalisation needs a prior range and delta_tc > 0");
      return GWL_E_INVALID;
    }
    const int nt = (int)std::ceil((rd.tmax - rd.tmin) / d->delta_tc);
    rd.dtc = (rd.tmax - rd.tmin) / nt;
    rd.ntimes = nt;
    rd.tcmin = rd.tmin;
    rd.tref = (rd.tmin + rd.tmax) / 2.0;
  }
  for (int k = 0; k < rd.nd; ++k) {
    for (int i = 0; i < 9; ++i) rd.tensor[k][i] = d->detector_tensor[9 * k + i];
    for (int i = 0; i < 3; ++i) rd.vertex[k][i] = d->vertex[3 * k + i];
  }
  // weight scale: sigma = 1 / median |W|
  const size_t nW = (size_t)rd.nd * rd.NT * rd.L;
  {
    std::vector<double> mags(nW);
    for (size_t i = 0; i < nW; ++i) mags[i] = std::hypot(d->linear_weights[2 * i], d->linear_weights[2 * i + 1]);
    std::nth_element(mags.begin(), mags.begin() + nW / 2, mags.end());
    const double med = mags[nW / 2];
    if (!(med > 0) || !std::isfinite(med)) { set_error("linear weights are zero or not finite"); return GWL_E_INVALID; }
    rd.sigma = 1.0 / med;
  }
  // host staging of W^T real embedding [Np][Kp]:
  //   column c = 2 (d NT + t) (real part):  k < L: Wr[t][k],  k >= L: -Wi[t][k-L]
  //   column c + 1 (imaginary part):        k < L: Wi[t][k],  k >= L:  Wr[t][k-L]
  const size_t nB = (size_t)rd.Np * rd.Kp;
  std::vector<double> b64(nB, 0.0);
  for (int dd = 0; dd < rd.nd; ++dd)
    for (int t = 0; t < rd.NT; ++t) {
      const size_t cr = (size_t)2 * (dd * rd.NT + t);
      for (int l = 0; l < rd.L; ++l) {
        const size_t wi = ((size_t)dd * rd.NT + t) * rd.L + l;
        const double wr = d->linear_weights[2 * wi] * rd.sigma, wim = d->linear_weights[2 * wi + 1] * rd.sigma;
        b64[cr * rd.Kp + l] = wr;
        b64[cr * rd.Kp + rd.L + l] = -wim;
        b64[(cr + 1) * rd.Kp + l] = wim;
        b64[(cr + 1) * rd.Kp + rd.L + l] = wr;
      }
    }
  std::vector<float> b32(nB);
  std::vector<__bf16> bh(nB), bl(nB);
  for (size_t i = 0; i < nB; ++i) {
    const float f = (float)b64[i];
    b32[i] = f;
    bh[i] = (__bf16)f;
    bl[i] = (__bf16)(f - (float)bh[i]);
  }
  std::vector<double> wq((size_t)rd.nd * rd.Qn);
  for (size_t i = 0; i < wq.size(); ++i) wq[i] = d->quadratic_weights[2 * i] * rd.sigma * rd.sigma;
  gwl_roq* e = new (std::nothrow) gwl_roq;
  if (!e) { set_error("out of host memory"); return GWL_E_NOMEM; }
  e->ctx = ctx;
  const size_t bytes =
      nB * (8 + 4 + 2 + 2) +
      sizeof(double) * ((size_t)rd.L + rd.Qn + wq.size() + rd.NT + ((size_t)rd.L + rd.Qn) * (kNB + 1)) +
      12 * 256;
  if (hipMalloc(&e->data, bytes) != hipSuccess) {
    delete e;
    set_error("hipMalloc failed for ROQ data");
    return GWL_E_NOMEM;
  }
  e->data_bytes = bytes;
  char* p = reinterpret_cast<char*>(e->data);
  auto take = [&](size_t nb) {
    char* q = p;
    p += (nb + 255) / 256 * 256;
    return q;
  };
  rd.bt64 = reinterpret_cast<double*>(take(nB * 8));
  rd.bt32 = reinterpret_cast<float*>(take(nB * 4));
  rd.bth = reinterpret_cast<__bf16*>(take(nB * 2));
  rd.btl = reinterpret_cast<__bf16*>(take(nB * 2));
  rd.fl = reinterpret_cast<double*>(take(rd.L * 8));
  rd.fq = reinterpret_cast<double*>(take(rd.Qn * 8));
  rd.wq = reinterpret_cast<double*>(take(wq.size() * 8));
  rd.ts = reinterpret_cast<double*>(take((size_t)rd.NT * 8));
  rd.nbl = reinterpret_cast<double*>(take((size_t)rd.L * (kNB + 1) * 8));
  rd.nbq = reinterpret_cast<double*>(take((size_t)rd.Qn * (kNB + 1) * 8));
  hipError_t err = hipMemcpy((void*)rd.bt64, b64.data(), nB * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy((void*)rd.bt32, b32.data(), nB * 4, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy((void*)rd.bth, bh.data(), nB * 2, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy((void*)rd.btl, bl.data(), nB * 2, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy((void*)rd.fl, d->linear_nodes, rd.L * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy((void*)rd.fq, d->quadratic_nodes, rd.Qn * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy((void*)rd.wq, wq.data(), wq.size() * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy((void*)rd.ts, d->time_samples, (size_t)rd.NT * 8, hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    hipLaunchKernelGGL(k_roq_node_table, dim3((unsigned)((rd.L + 127) / 128)), dim3(128), 0, 0, rd.fl,
                       rd.L, const_cast<double*>(rd.nbl));
    hipLaunchKernelGGL(k_roq_node_table, dim3((unsigned)((rd.Qn + 127) / 128)), dim3(128), 0, 0, rd.fq,
                       rd.Qn, const_cast<double*>(rd.nbq));
    err = hipGetLastError();
    if (err == hipSuccess) err = hipDeviceSynchronize();
  }
  if (err != hipSuccess) {
    (void)hipFree(e->data);
    delete e;
    return hip_fail(err, "gwl_roq_create upload");
  }
  e->dev = rd;
  *out = e;
  if (hbm_bytes) *hbm_bytes = (int64_t)bytes;
  return GWL_OK;
}

int gwl_roq_destroy(gwl_roq* e) {
  if (!e) return GWL_OK;
  (void)hipSetDevice(e->ctx->device);
  if (e->data) (void)hipFree(e->data);
  if (e->spline) (void)hipFree(e->spline);
  if (e->tprior) (void)hipFree(e->tprior);
  if (e->ws) (void)hipFree(e->ws);
  delete e;
  return GWL_OK;
}

int gwl_roq_set_time_prior(gwl_roq* e, int64_t n, double t0, double dt, const double* density) {
  if (!e || !(e->dev.flags & GWL_TIME_MARG)) {
    set_error("gwl_roq_set_time_prior: the likelihood does not marginalise over time");
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

int gwl_roq_set_distance_spline(gwl_roq* e, int32_t nx, int32_t ny, const double* tx,
                                const double* ty, const double* coeffs, const double* bbox) {
  if (!e || !tx || !ty || !coeffs || !bbox || nx < 4 || ny < 4) {
    set_error("gwl_roq_set_distance_spline: invalid argument");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  const size_t n = (size_t)(nx + 4) + (ny + 4) + (size_t)nx * ny;
  if (e->spline) (void)hipFree(e->spline);
  e->spline = nullptr;
  GWL_HIP_CHECK(hipMalloc(&e->spline, n * sizeof(double)));
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


int gwl_roq_reserve(gwl_roq* e, int64_t rows) {
  if (!e || rows < 1) {
    set_error("gwl_roq_reserve: invalid argument");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  size_t need = 0;
  for (int mode = ROQ_FP64; mode <= ROQ_BF16X3; ++mode)
    need = std::max(need, roq_ws_bytes(e->dev, rows, mode));
  if (need <= e->ws_bytes) return GWL_OK;
  if (e->ws) (void)hipFree(e->ws);
  e->ws = nullptr;
  e->ws_bytes = 0;
  if (hipMalloc(&e->ws, need) != hipSuccess) {
    set_error("hipMalloc failed for ROQ workspace");
    return GWL_E_NOMEM;
  }
  e->ws_bytes = need;
  return GWL_OK;
}

int gwl_roq_eval(gwl_roq* e, const double* params, int64_t n, int32_t ld, int32_t mode,
                 double* out, void* stream) {
  if (e && n == 0) return GWL_OK;  // an empty batch needs no buffers
  if (!e || !params || !out || n < 0 || ld < GWL_NPARAM || mode < 0 || mode > 2) {
    set_error("gwl_roq_eval: invalid argument");
    return GWL_E_INVALID;
  }
  if (n == 0) return GWL_OK;
  if ((e->dev.flags & GWL_DISTANCE_MARG) && !e->has_spline) {
    set_error("gwl_roq_eval: distance marginalisation requested but no spline attached");
    return GWL_E_INVALID;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const void* Cv = nullptr;
  RoqRow* rows = nullptr;
  if (int rc = roq_prep_gemm(e, params, n, ld, mode, st, &Cv, &rows)) return rc;
  const RoqDev& rd = e->dev;
  const char* Cp = static_cast<const char*>(Cv);
  const size_t klds = (rd.flags & GWL_DISTANCE_MARG) ? (size_t)(rd.sp.nx + rd.sp.ny + 8) * 8 : 0;
  const dim3 eg((unsigned)((n + kRoqEpiRows - 1) / kRoqEpiRows));
  if (mode == ROQ_FP64)
    hipLaunchKernelGGL(k_roq_epi<double>, eg, dim3(64 * kRoqEpiRows), klds, st, rd,
                       reinterpret_cast<const double*>(Cp), rows, (long long)n, out);
  else
    hipLaunchKernelGGL(k_roq_epi<float>, eg, dim3(64 * kRoqEpiRows), klds, st, rd,
                       reinterpret_cast<const float*>(Cp), rows, (long long)n, out);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

int gwl_roq_inner_products(gwl_roq* e, const double* params, int64_t n, int32_t ld,
                           int32_t mode, int32_t n_times, double* d_inner_h, double* h_inner_h,
                           void* stream) {
  if (e && n == 0) return GWL_OK;
  if (!e || !params || !d_inner_h || !h_inner_h || n < 0 || ld < GWL_NPARAM || mode < 0 ||
      mode > 2 || n_times < 0) {
    set_error("gwl_roq_inner_products: invalid argument");
    return GWL_E_INVALID;
  }
  if (n_times > 0 && (!(e->dev.flags & GWL_TIME_MARG) || n_times != e->dev.ntimes)) {
    set_error("gwl_roq_inner_products: n_times must be 0 or the length of the likelihood's "
              "time-marginalisation grid");
    return GWL_E_INVALID;
  }
  if (n > ((int64_t)1 << 31) / (n_times > 0 ? n_times : 1)) {
    set_error("gwl_roq_inner_products: too many outputs in one call");
    return GWL_E_INVALID;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const void* C = nullptr;
  RoqRow* rows = nullptr;
  if (int rc = roq_prep_gemm(e, params, n, ld, mode, st, &C, &rows)) return rc;
  const long long total = n * (long long)(n_times > 0 ? n_times : 1);
  const dim3 g((unsigned)((total + 255) / 256));
  if (mode == ROQ_FP64)
    hipLaunchKernelGGL(k_roq_series<double>, g, dim3(256), 0, st, e->dev,
                       static_cast<const double*>(C), rows, (long long)n, (int)n_times, d_inner_h,
                       h_inner_h);
  else
    hipLaunchKernelGGL(k_roq_series<float>, g, dim3(256), 0, st, e->dev,
                       static_cast<const float*>(C), rows, (long long)n, (int)n_times, d_inner_h,
                       h_inner_h);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

int gwl_gemm(gwl_ctx* ctx, int32_t mode, const void* A, const void* A_lo, const void* Bt,
             const void* Bt_lo, void* C, int32_t M, int32_t N, int32_t K, void* stream) {
  if (!ctx || !A || !Bt || !C || M % kBM || N % kBN || K % kBK || M <= 0 || N <= 0 || K <= 0 ||
      mode < 0 || mode > 2 || (mode == ROQ_BF16X3 && (!A_lo || !Bt_lo))) {
    set_error("gwl_gemm: invalid argument (M, N multiples of 128, K of 32)");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(ctx->device));
  (void)hipGetLastError();
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (mode == ROQ_FP64)
    hipLaunchKernelGGL(k_gemm_f64, dim3(N / 64, M / 64), dim3(256), 0, st, (const double*)A,
                       (const double*)Bt, (double*)C, M, N, K);
  else if (mode == ROQ_FP32)
    hipLaunchKernelGGL(k_gemm_f32, dim3(N / kBN, M / kBM), dim3(256), 0, st, (const float*)A,
                       (const float*)Bt, (float*)C, M, N, K);
  else if (N % kBNW == 0)
    hipLaunchKernelGGL(k_gemm_bf16x3_w, dim3(N / kBNW, M / kBM), dim3(256), 0, st,
                       (const __bf16*)A, (const __bf16*)A_lo, (const __bf16*)Bt,
                       (const __bf16*)Bt_lo, (float*)C, M, N, K, K);
  else
    hipLaunchKernelGGL(k_gemm_bf16x3, dim3(N / kBN, M / kBM), dim3(256), 0, st,
                       (const __bf16*)A, (const __bf16*)A_lo, (const __bf16*)Bt,
                       (const __bf16*)Bt_lo, (float*)C, M, N, K, K);
  GWL_HIP_CHECK(hipGetLastError());
  return GWL_OK;
}

}  // extern "C"
