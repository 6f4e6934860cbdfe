// Per-bin source model x detector projection x data shared by the
// likelihood kernels (gwl_kernels.hip) and the calibration prep kernel
// (gwl_calib.hip): row constants in registers / LDS, data loads in the
// decimated layout, and the per-bin integrand.
#pragma once
#include "gwl_internal.h"

namespace gwl {

template <typename R>
struct SampleFast {
  R cre[GWL_MAX_DET], cim[GWL_MAX_DET], cabs2[GWL_MAX_DET];
  R k2;  // sine-Gaussian exponent coefficient (log2 units)
};

template <typename R>
__device__ __forceinline__ R exp2r(R x);
template <>
__device__ __forceinline__ float exp2r<float>(float x) { return __builtin_amdgcn_exp2f(x); }
template <>
__device__ __forceinline__ double exp2r<double>(double x) { return exp2(x); }

template <typename R>
__device__ __forceinline__ cplx<R> load_dw(const EvDev& ev, int d, int i);
template <>
__device__ __forceinline__ float2 load_dw<float>(const EvDev& ev, int d, int i) {
  return ev.dw32[(size_t)d * ev.N + i];
}
template <>
__device__ __forceinline__ double2 load_dw<double>(const EvDev& ev, int d, int i) {
  return ev.dw64[(size_t)d * ev.N + i];
}
template <typename R>
__device__ __forceinline__ R load_w(const EvDev& ev, int d, int i);
template <>
__device__ __forceinline__ float load_w<float>(const EvDev& ev, int d, int i) {
  return ev.w32[(size_t)d * ev.N + i];
}
template <>
__device__ __forceinline__ double load_w<double>(const EvDev& ev, int d, int i) {
  return ev.w64[(size_t)d * ev.N + i];
}

// Row prologue for a tile of S rows (threads 0..S-1), into LDS.
template <typename R, int S>
__device__ inline void tile_prologue(const EvDev& ev, const double* params, int ld, long long n,
                                     long long row0, SampleConst* sc, SampleFast<R>* sf,
                                     int* band) {
  int t = threadIdx.x;
  if (t < S) {
    long long r = row0 + t;
    if (r < n) {
      compute_sample_const(ev, params + r * (long long)ld, sc[t]);
      sc[t].row = (int)r;
    } else {
      sc[t].valid = 0;
      sc[t].kb = 1;
      sc[t].ke = 0;
      sc[t].hscale = 1.0;
      sc[t].jitter = 0.0;
    }
#pragma unroll
    for (int d = 0; d < GWL_MAX_DET; ++d) {
      sf[t].cre[d] = (R)(d < ev.nd ? sc[t].cre[d] : 0.0);
      sf[t].cim[d] = (R)(d < ev.nd ? sc[t].cim[d] : 0.0);
      sf[t].cabs2[d] = (R)(d < ev.nd ? sc[t].cabs2[d] : 0.0);
    }
    sf[t].k2 = (R)sc[t].c[0];
  }
  __syncthreads();
  if (t == 0) {
    int lo = 1 << 30, hi = -1;
    for (int s = 0; s < S; ++s) {
      if (sc[s].kb <= sc[s].ke) {
        lo = min(lo, sc[s].kb);
        hi = max(hi, sc[s].ke);
      }
    }
    band[0] = lo;
    band[1] = hi;
  }
  __syncthreads();
}

// The same with the row constants read from a buffer filled by
// k_sample_const (one thread per row), so that a row's prologue runs once
// per evaluation rather than once per workgroup that touches the row.
template <typename R, int S>
__device__ inline void tile_prologue_load(const EvDev& ev, const SampleConst* rows, long long n,
                                          long long row0, SampleConst* sc, SampleFast<R>* sf,
                                          int* band) {
  int t = threadIdx.x;
  if (t < S) {
    long long r = row0 + t;
    if (r < n) {
      sc[t] = rows[r];
    } else {
      sc[t].valid = 0;
      sc[t].kb = 1;
      sc[t].ke = 0;
      sc[t].hscale = 1.0;
      sc[t].jitter = 0.0;
    }
#pragma unroll
    for (int d = 0; d < GWL_MAX_DET; ++d) {
      sf[t].cre[d] = (R)(d < ev.nd ? sc[t].cre[d] : 0.0);
      sf[t].cim[d] = (R)(d < ev.nd ? sc[t].cim[d] : 0.0);
      sf[t].cabs2[d] = (R)(d < ev.nd ? sc[t].cabs2[d] : 0.0);
    }
    sf[t].k2 = (R)sc[t].c[0];
  }
  __syncthreads();
  if (t == 0) {
    int lo = 1 << 30, hi = -1;
    for (int s = 0; s < S; ++s) {
      if (sc[s].kb <= sc[s].ke) {
        lo = min(lo, sc[s].kb);
        hi = max(hi, sc[s].ke);
      }
    }
    band[0] = lo;
    band[1] = hi;
  }
  __syncthreads();
}

// Source model x projection x data for one frequency bin n and S rows:
//   I[s]  = sum_d h_{s,d}(n) conj(dw_d(n))     (h conj(d) / S, base.py:327-329)
//   hh[s] += sum_d |h_{s,d}(n)|^2 w_d(n)        (|h|^2 / S, gw/utils.py:137)
// with h_{s,d} = H_{s,d}(f) exp(-2 pi i (Psi_s(f) / 2pi + f dt_{s,d})),
// the phase argument in float64 (interferometer.py:357-362).  pol / nf: the
// caller's polarizations (GWL_HOST_POLARIZATIONS) and their bins per row.
template <typename R, int WAVE, int ND, int S>
__device__ __forceinline__ void bin_rows(int n, double f, const cplx<R>* dw, const R* w,
                                         const SampleConst* sc, const SampleFast<R>* sf,
                                         cplx<R>* I, R* hh, int nd = ND,
                                         const double2* pol = nullptr, int nf = 0) {
  using C = cplx<R>;
  if constexpr (WAVE == GWL_TAYLORF2) {
    double b[kNB];
    tf2_basis(f, b);
    const R amp = (R)(b[1] * sqrt(b[3]));  // f^-7/6
    const R amp2 = amp * amp;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (n < sc[s].kb || n > sc[s].ke) {
        I[s] = cmk<C>((R)0, (R)0);
        continue;
      }
      const double psi = tf2_phase_cycles(sc[s].c, b);
      R ar = 0, ai = 0, hw = 0;
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        if (d >= nd) break;
        const double cyc = fma(f, sc[s].dt[d], psi);
        const C e = expm2pi((R)(cyc - floor(cyc)));
        const R cr = sf[s].cre[d] * amp, ci = sf[s].cim[d] * amp;
        const R hr = cr * e.x - ci * e.y, hi = cr * e.y + ci * e.x;
        ar += hr * dw[d].x + hi * dw[d].y;
        ai += hi * dw[d].x - hr * dw[d].y;
        hw += sf[s].cabs2[d] * w[d];
      }
      I[s] = cmk<C>(ar, ai);
      hh[s] += amp2 * hw;
    }
  } else if constexpr (WAVE == GWL_HOST_POLARIZATIONS) {
    // the caller's h+ and hx at bin n of row sc[s].row, projected with
    // F+/sigma and Fx/sigma (float64) and shifted by each detector's delay
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (n < sc[s].kb || n > sc[s].ke) {
        I[s] = cmk<C>((R)0, (R)0);
        continue;
      }
      const double2 hp = pol[(size_t)sc[s].row * 2 * nf + n];
      const double2 hx = pol[((size_t)sc[s].row * 2 + 1) * nf + n];
      R ar = 0, ai = 0, hw = 0;
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        if (d >= nd) break;
        const C p = cmk<C>((R)(sc[s].cre[d] * hp.x + sc[s].cim[d] * hx.x),
                           (R)(sc[s].cre[d] * hp.y + sc[s].cim[d] * hx.y));
        const double cyc = f * sc[s].dt[d];
        const C e = expm2pi((R)(cyc - floor(cyc)));
        const R hr = p.x * e.x - p.y * e.y, hi = p.x * e.y + p.y * e.x;
        ar += hr * dw[d].x + hi * dw[d].y;
        ai += hi * dw[d].x - hr * dw[d].y;
        hw += (hr * hr + hi * hi) * w[d];
      }
      I[s] = cmk<C>(ar, ai);
      hh[s] += hw;
    }
  } else {  // sine-Gaussian
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (n < sc[s].kb || n > sc[s].ke) {
        I[s] = cmk<C>((R)0, (R)0);
        continue;
      }
      const R xm = (R)(f - sc[s].c[1]), xp = (R)(f + sc[s].c[1]);
      const R k2 = sf[s].k2;
      const R em = exp2r<R>(-k2 * xm * xm), ep = exp2r<R>(-k2 * xp * xp);
      const R sp = em + ep, sm = em - ep;
      R ar = 0, ai = 0, hw = 0;
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        if (d >= nd) break;
        const double cyc = f * sc[s].dt[d];
        const C e = expm2pi((R)(cyc - floor(cyc)));
        const R pr = sf[s].cre[d] * sp, pim = sf[s].cim[d] * sm;
        const R hr = pr * e.x - pim * e.y, hi = pr * e.y + pim * e.x;
        ar += hr * dw[d].x + hi * dw[d].y;
        ai += hi * dw[d].x - hr * dw[d].y;
        hw += (pr * pr + pim * pim) * w[d];
      }
      I[s] = cmk<C>(ar, ai);
      hh[s] += hw;
    }
  }
}

}  // namespace gwl
