// ROQ linear weights from a reduced basis (bilby/gw/likelihood/roq.py:849-918,
// _set_weights_linear; LIGO-T2100125): for basis element i and ROQ time
// sample m = k0 + k (time m T / N, k < n_times)
//   W[k][i] = scale * sum_j x[i][j] exp(+2 pi i n_j m / N),
// x[i][j] = (d / S)(f_j) conj(B_i(f_j)) on the detector's masked bins n_j,
// scale = 4 / T.  bilby computes one length-N inverse FFT per basis element
// and keeps the window of times; here the window is evaluated directly as a
// float64 complex GEMM over chunks of frequencies,
//   [Re x | Im x] . [cos | -sin ; sin | cos]^T,
// with the exponentials generated on the device from the exact integer
// phase (n_j m mod N) / N, and the chunk products summed in float64.
#include "gwl_internal.h"

#include <cstring>
#include <vector>

using gwl::set_error;

namespace gwl {
namespace {

constexpr int kChunk = 8192;   // frequencies per GEMM (K = 2 kChunk)

size_t up128(size_t x) { return (x + 127) / 128 * 128; }

// Bt rows k < n_times: [cos | -sin], rows n_times + k: [sin | cos] of
// 2 pi (n_j m mod N) / N, m = k0 + k; zero padding elsewhere.
__global__ __launch_bounds__(256) void k_roq_twiddles(const long long* nidx, int nf, long long N,
                                                      long long k0, int n_times, int Np,
                                                      double* bt) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)Np * (2 * kChunk);
  if (t >= total) return;
  const int row = (int)(t / (2 * kChunk));
  const int col = (int)(t % (2 * kChunk));
  const int j = col < kChunk ? col : col - kChunk;
  double v = 0.0;
  const int k = row < n_times ? row : row - n_times;
  if (row < 2 * n_times && j < nf) {
    const long long m = k0 + k;
    const long long r = (long long)(((unsigned long long)nidx[j] * (unsigned long long)m) %
                                    (unsigned long long)N);
    double s, c;
    sincospi(2.0 * (double)r / (double)N, &s, &c);
    if (row < n_times)
      v = col < kChunk ? c : -s;
    else
      v = col < kChunk ? s : c;
  }
  bt[t] = v;
}

__global__ void k_accumulate(const double* c, double* acc, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) acc[i] += c[i];
}

}  // namespace
}  // namespace gwl

extern "C" {

int gwl_roq_weights_linear(gwl_ctx* ctx, int32_t n_basis, int64_t n_freq, const double* x,
                           const int64_t* bins, int64_t N, int64_t k0, int32_t n_times,
                           double scale, double* out) {
  if (!ctx || n_basis < 1 || n_freq < 1 || !x || !bins || N < 2 || k0 < 0 || n_times < 1 ||
      !out || k0 + n_times > N) {
    set_error("gwl_roq_weights_linear: invalid argument");
    return GWL_E_INVALID;
  }
  for (int64_t j = 0; j < n_freq; ++j)
    if (bins[j] < 0 || bins[j] >= N) {
      set_error("gwl_roq_weights_linear: frequency bin outside [0, N)");
      return GWL_E_INVALID;
    }
  GWL_HIP_CHECK(hipSetDevice(ctx->device));
  const int K = 2 * gwl::kChunk;
  const size_t Mp = gwl::up128((size_t)n_basis), Np = gwl::up128((size_t)2 * n_times);
  double *dA = nullptr, *dB = nullptr, *dC = nullptr, *dW = nullptr;
  long long* dn = nullptr;
  auto cleanup = [&]() {
    if (dA) (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    if (dC) (void)hipFree(dC);
    if (dW) (void)hipFree(dW);
    if (dn) (void)hipFree(dn);
  };
  if (hipMalloc((void**)&dA, Mp * K * 8) != hipSuccess ||
      hipMalloc((void**)&dB, Np * K * 8) != hipSuccess ||
      hipMalloc((void**)&dC, Mp * Np * 8) != hipSuccess ||
      hipMalloc((void**)&dW, Mp * Np * 8) != hipSuccess ||
      hipMalloc((void**)&dn, (size_t)gwl::kChunk * 8) != hipSuccess) {
    cleanup();
    set_error("gwl_roq_weights_linear: hipMalloc failed");
    return GWL_E_NOMEM;
  }
  hipError_t err = hipMemset(dW, 0, Mp * Np * 8);
  std::vector<double> a(Mp * K);
  std::vector<long long> nb(gwl::kChunk);
  for (int64_t j0 = 0; j0 < n_freq && err == hipSuccess; j0 += gwl::kChunk) {
    const int nf = (int)std::min<int64_t>(gwl::kChunk, n_freq - j0);
    std::fill(a.begin(), a.end(), 0.0);
    for (int i = 0; i < n_basis; ++i)
      for (int j = 0; j < nf; ++j) {
        const double* v = x + ((size_t)i * n_freq + j0 + j) * 2;
        a[(size_t)i * K + j] = v[0];
        a[(size_t)i * K + gwl::kChunk + j] = v[1];
      }
    for (int j = 0; j < gwl::kChunk; ++j) nb[j] = j < nf ? (long long)bins[j0 + j] : 0;
    err = hipMemcpy(dA, a.data(), a.size() * 8, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(dn, nb.data(), nb.size() * 8, hipMemcpyHostToDevice);
    if (err != hipSuccess) break;
    const long long total = (long long)Np * K;
    hipLaunchKernelGGL(gwl::k_roq_twiddles, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, 0,
                       dn, nf, (long long)N, (long long)k0, n_times, (int)Np, dB);
    err = hipGetLastError();
    if (err != hipSuccess) break;
    if (int rc = gwl_gemm(ctx, 0, dA, nullptr, dB, nullptr, dC, (int32_t)Mp, (int32_t)Np, K, nullptr)) {
      cleanup();
      return rc;
    }
    const long long nc = (long long)(Mp * Np);
    hipLaunchKernelGGL(gwl::k_accumulate, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, 0, dC,
                       dW, nc);
    err = hipGetLastError();
  }
  std::vector<double> w(Mp * Np);
  if (err == hipSuccess) err = hipMemcpy(w.data(), dW, w.size() * 8, hipMemcpyDeviceToHost);
  cleanup();
  if (err != hipSuccess) return gwl::hip_fail(err, "gwl_roq_weights_linear");
  for (int k = 0; k < n_times; ++k)
    for (int i = 0; i < n_basis; ++i) {
      out[((size_t)k * n_basis + i) * 2 + 0] = scale * w[(size_t)i * Np + k];
      out[((size_t)k * n_basis + i) * 2 + 1] = scale * w[(size_t)i * Np + n_times + k];
    }
  return GWL_OK;
}

}  // extern "C"
