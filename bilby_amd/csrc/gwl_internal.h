// Host-side structures shared by the libgwl translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include "../../include/gwl.h"
#include "gwl_device.h"

struct gwl_ctx {
  int device;
  int n_cu;
  size_t lds_per_block;
  void* scratch = nullptr;  // this device's address of the gwl_eval_noalloc scratch array
};

// Tile of the time-marginalised kernel: S rows per workgroup, P-point LDS
// transforms, G groups of the Q = N / P decimated sub-sequences.
struct TimeTile {
  int S, P, G;
};

struct gwl_event {
  gwl_ctx* ctx;
  gwl::EvDev dev;
  void* data = nullptr;          // device: dw32, w32, dw64, w64
  size_t data_bytes = 0;
  double* spline = nullptr;      // device: tx, ty, coefficients
  size_t spline_bytes = 0;
  double* tprior = nullptr;      // device: time-prior density table
  void* ws = nullptr;            // device workspace (partials)
  size_t ws_bytes = 0;
  int64_t ws_rows = 0;
  bool has_spline = false;
  // tile overrides (0 = heuristic)
  int tile_rows = 0;
  int tile_wg_per_cu = 0;
  int transform = GWL_TRANSFORM_AUTO;
  int multi_steps = 0;  // sub-sequence steps per multi-event work item (0: default)
  int64_t pol_rows = 0;  // rows of the caller's polarizations (GWL_HOST_POLARIZATIONS)
  // float32 evaluations without time marginalisation re-evaluate in float64,
  // in the same call, every row whose float32 error bound esc_coeff * Q
  // exceeds esc_tol (Q: the row's inner-product scale); esc_coeff = 0: off
  double esc_coeff = 2e-7;
  double esc_tol = 8e-4;
  int* esc_count_dev = nullptr;  // counter of the last escalating evaluation
  // kernel timing (measurement only)
  bool timing = false;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  double timed_ms = 0.0;
  int64_t timed_launches = 0;
};

#include <vector>

namespace gwl {
// one item of the multi-event work list: rows [row0, row0 + 8) of event e,
// group g of its decimated sub-sequences
struct MultiItem {
  int e;
  int g;
  long long row0;
};
}

// Many events evaluated together: one persistent launch over a work list,
// optionally replayed from a captured hipGraph.
struct gwl_multi {
  gwl_ctx* ctx = nullptr;
  int n_events = 0;
  int wave = 0;
  std::vector<gwl_event*> events;
  std::vector<int64_t> max_rows, rows;
  std::vector<int> groups;              // G_e (groups of sub-sequences per row)
  std::vector<const double*> params;    // caller-owned device buffers
  std::vector<double*> outs;
  std::vector<void*> parts;             // [max_rows][16][L] double2 per event
  std::vector<double*> hhs;             // [max_rows][16] per event
  void* dev = nullptr;                  // device tables (events, pointers, items, blocks)
  size_t dev_bytes = 0;
  gwl::EvDev* d_evs = nullptr;
  const double** d_params = nullptr;
  double** d_outs = nullptr;
  long long* d_rows = nullptr;
  int* d_groups = nullptr;
  void** d_parts = nullptr;
  double** d_hhs = nullptr;
  gwl::MultiItem* d_items = nullptr;
  int* d_n_items = nullptr;
  int* d_counter = nullptr;
  int* d_block_event = nullptr;
  long long* d_block_row0 = nullptr;
  int* d_n_blocks = nullptr;
  size_t max_items = 0, max_blocks = 0;
  int n_items = 0;
  bool captured = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // TaylorF2 events with 512-point transforms: wave-per-row kernels, row
  // constants [max_rows] per event
  bool wave_path = false;
  bool capturing = false;   // a capture records grids for the largest work list
  int n_blocks = 0;         // epilogue blocks of the current rows
  std::vector<void*> roww;
  void** d_roww = nullptr;
};

namespace gwl {
int multi_launch_raw(gwl_multi* m, hipStream_t stream);
int multi_launch_epi(gwl_multi* m, hipStream_t stream);
int multi_set_rows(gwl_multi* m, const int64_t* rows, hipStream_t stream);
size_t multi_lds_bytes();
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);
size_t workspace_bytes(const gwl_event* ev, int64_t rows);
// the workspace of an evaluation at one precision (enum gwl_precision)
size_t workspace_bytes_for(const gwl_event* ev, int64_t rows, int precision);
// upload a time-prior density table (n values at t0 + k dt) into *dst (freed
// and replaced), checking the arguments
int upload_time_prior(int64_t n, double t0, double dt, const double* density, double** dst);
// this device's address of the gwl_eval_noalloc scratch array (gwl_noalloc.hip)
int noalloc_scratch_init(int device, void** ptr);
int launch_eval(gwl_event* ev, const double* params, int64_t n, int ld, int precision, double* out,
                double* aux, hipStream_t stream);
int timing_begin(gwl_event* ev, hipStream_t stream);
int timing_end(gwl_event* ev, hipStream_t stream);
// Split-bfloat16 x3 GEMM over K chunks into float32 partials C[z][M][N]
// (z < ceil(K / kcol)); M, N multiples of 128, K and kcol multiples of 32.
int gemm_bf16x3_split(const __bf16* Ah, const __bf16* Al, const __bf16* Bh, const __bf16* Bl,
                      float* C, int M, int N, int K, int kcol, hipStream_t st);
int launch_response(gwl_event* ev, const double* params_dev, int det, int precision, double* out_dev,
                    hipStream_t stream);
// float64 epilogue of the time-marginalised path over G window partials
// (slice: one window slice of the event, with the logsumexp state lse carried
// from the first slice to the last)
int launch_epi_time(gwl_event* ev, const double* params, int64_t n, int ld, int G,
                    const double2* part, const double* hhpart, double* out, hipStream_t stream,
                    const gwl::EvDev* slice = nullptr, double2* lse = nullptr, int first = 1,
                    int last = 1, const double* rinfo = nullptr);
// wave-per-row TaylorF2 time transform (gwl_time_wave.hip)
bool time_wave_eligible(const gwl_event* ev);
int time_wave_groups(const gwl_event* ev, long long n);
size_t time_wave_ws_bytes(const gwl_event* ev, int64_t rows);
size_t time_wave_lds_bytes(const EvDev& ev, bool fused);
int init_time_wave_tables(const EvDev& ev, double* tf2b, float2* dwa32);
int launch_time_wave(gwl_event* ev, const double* params, int64_t n, int ld, double* out,
                     hipStream_t stream);
size_t roww_bytes(int64_t rows);
int launch_time_series(gwl_event* ev, const double* params, int64_t n, int ld, int precision,
                       double* dh, double* hh, hipStream_t stream);
// Calibration with time marginalisation (gwl_time_wave.hip)
bool cal_time_eligible(const gwl_event* ev);
size_t cal_time_ws_bytes(int64_t rows, int nc);
// fp64: the float64 kernel, with the curves' float32 remainders cal_lo and
// the table wN[m] = W_N^m (m < N)
int launch_cal_time(gwl_event* ev, const double* params, int64_t n, int ld, int nc,
                    const float2* cal, const double* calps, void* ws, double* per_curve,
                    double* out, hipStream_t stream, bool fp64 = false,
                    const float2* cal_lo = nullptr, const double2* wN = nullptr);
size_t multi_wave_lds_bytes();
int multi_launch_wave(gwl_multi* m, hipStream_t stream);
}  // namespace gwl

#define GWL_HIP_CHECK(expr)                                   \
  do {                                                        \
    hipError_t _e = (expr);                                   \
    if (_e != hipSuccess) return gwl::hip_fail(_e, #expr);    \
  } while (0)
