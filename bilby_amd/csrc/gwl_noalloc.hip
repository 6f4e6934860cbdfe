// gwl_eval without allocation: the evaluation workspace (row constants,
// inner-product and window partials, the escalation list) lives in a
// file-scope device array instead of the event's hipMalloc'd workspace, for
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops).  Each device holds its own copy of the array, found with
// hipGetSymbolAddress once per device when a context is created; host code
// uses that address, never the array's name.
#include "gwl_internal.h"

#include <mutex>

namespace gwl {

constexpr size_t kNoAllocScratchBytes = GWL_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// The device address of ``scratch`` on ``device`` (which must be current).
int noalloc_scratch_init(int device, void** ptr) {
  *ptr = nullptr;
  if (device < 0 || device >= kMaxDevices) {
    set_error("gwl: device index out of range for the scratch table");
    return GWL_E_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !p) return hip_fail(err, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  return GWL_OK;
}

}  // namespace gwl

using namespace gwl;

extern "C" int gwl_eval_noalloc(gwl_event* e, const double* params, int64_t n, int32_t ld,
                                int32_t precision, double* out, double* aux, void* stream) {
  if (e && n == 0) return GWL_OK;
  if (!e || !params || !out || n < 0 || ld < GWL_NPARAM) {
    set_error("gwl_eval_noalloc: invalid argument");
    return GWL_E_INVALID;
  }
  if (precision != GWL_FP32 && precision != GWL_FP64) {
    set_error("gwl_eval_noalloc: unknown precision");
    return GWL_E_INVALID;
  }
  if ((e->dev.flags & GWL_DISTANCE_MARG) && !e->has_spline) {
    set_error("gwl_eval_noalloc: distance marginalisation requested but no spline attached");
    return GWL_E_INVALID;
  }
  void* ptr = e->ctx->scratch;
  if (!ptr) {
    set_error("gwl_eval_noalloc: the context has no scratch buffer");
    return GWL_E_INVALID;
  }
  const size_t need = workspace_bytes_for(e, n, precision);
  if (need > kNoAllocScratchBytes) {
    char msg[160];
    snprintf(msg, sizeof msg,
             "gwl_eval_noalloc: %lld rows need %zu bytes of scratch; the maximum is %zu",
             (long long)n, need, kNoAllocScratchBytes);
    set_error(msg);
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(e->ctx->device));
  // the launchers read the workspace from the event: point it at the scratch
  // for this call and restore it afterwards
  void* ws = e->ws;
  const size_t ws_bytes = e->ws_bytes;
  e->ws = ptr;
  e->ws_bytes = kNoAllocScratchBytes;
  const int rc = launch_eval(e, params, n, ld, precision, out, aux,
                             reinterpret_cast<hipStream_t>(stream));
  e->ws = ws;
  e->ws_bytes = ws_bytes;
  return rc;
}
