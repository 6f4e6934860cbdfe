/*
 * gwl.h -- C ABI of the MI355X (gfx950) compact-binary likelihood library
 * (libgwl.so, built from bilby_amd/csrc/).
 *
 * The reference exposes this path only as a duck-typed Python object,
 * bilby.gw.likelihood.GravitationalWaveTransient
 * (bilby/gw/likelihood/base.py:150-229 constructor, :419-446
 * log_likelihood_ratio).  It has no FFI; these entry points are what a
 * Python (ctypes) binding of that interface calls.  Each declaration names
 * the reference code it replaces.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative GWL_E* code on
 *    failure; gwl_last_error() then holds a message (thread-local).
 *    No C++ exception crosses this boundary.
 *  - "device pointer" arguments live on the context's device; "host
 *    pointer" arguments are ordinary host memory.
 *  - Parameter rows are GWL_NPARAM float64 values in the column order of
 *    enum gwl_param.  Unused columns are ignored.
 *  - stream is a hipStream_t passed as void* (NULL = the legacy stream).
 */
#ifndef GWL_H
#define GWL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GWL_NPARAM 16
#define GWL_MAX_DET 4

enum gwl_status {
  GWL_OK = 0,
  GWL_E_INVALID = -1,   /* bad argument / shape (bilby: ValueError, TypeError) */
  GWL_E_HIP = -2,       /* HIP runtime error */
  GWL_E_NOMEM = -3,     /* allocation failed or HBM budget exceeded */
  GWL_E_UNSUPPORTED = -4
};

/* Column order of a parameter row (float64). */
enum gwl_param {
  GWL_P_RA = 0, GWL_P_DEC = 1, GWL_P_PSI = 2, GWL_P_GEOCENT_TIME = 3,
  GWL_P_TIME_JITTER = 4, GWL_P_LUMINOSITY_DISTANCE = 5, GWL_P_THETA_JN = 6,
  GWL_P_PHASE = 7, GWL_P_MASS_1 = 8, GWL_P_MASS_2 = 9, GWL_P_CHI_1 = 10,
  GWL_P_CHI_2 = 11, GWL_P_HRSS = 12, GWL_P_Q = 13, GWL_P_FREQUENCY = 14
};

/* Source models: bilby/gw/source.py:1153-1209 (sinegaussian) and the
 * TaylorF2 approximant reached through source.py:269-348, 552-690. */
/* GWL_HOST_POLARIZATIONS: the source model is evaluated by the caller (any
 * bilby frequency_domain_source_model), which hands its polarizations over
 * with gwl_event_set_polarizations before each evaluation. */
enum gwl_approximant { GWL_SINE_GAUSSIAN = 1, GWL_TAYLORF2 = 2, GWL_HOST_POLARIZATIONS = 3 };

/* Marginalisation flags: base.py:150-229 constructor arguments
 * time_marginalization, distance_marginalization, phase_marginalization,
 * jitter_time.
 *
 * GWL_TF2_END_ISCO has no counterpart in the reference: by default TaylorF2
 * runs from wf_minimum_frequency up to wf_maximum_frequency, which is what
 * lalsimulation's XLALSimInspiralTaylorF2 does for the non-zero fEnd bilby
 * always passes (maximum_frequency, source.py:338-343, 633-643; ROQ nodes:
 * every node, source.py:1125-1128).  With the flag the waveform also ends
 * at the Schwarzschild ISCO 1 / (6^3/2 pi M) (lalsimulation's rule for
 * fEnd = 0), for callers that want inspiral-only templates. */
enum gwl_flags {
  GWL_TIME_MARG = 1, GWL_DISTANCE_MARG = 2, GWL_PHASE_MARG = 4, GWL_JITTER_TIME = 8,
  GWL_TF2_END_ISCO = 16
};

/* Arithmetic of an evaluation: all-float64, or float32 bulk arithmetic with
 * float64 phase arguments and float64 final accumulation. */
enum gwl_precision { GWL_FP64 = 0, GWL_FP32 = 1 };

typedef struct gwl_ctx gwl_ctx;
typedef struct gwl_event gwl_event;

/* One analysis segment ("event"): the state bilby's likelihood holds in its
 * InterferometerList (bilby/gw/detector/networks.py:13-352) plus the
 * constructor options of GravitationalWaveTransient (base.py:150-229). */
typedef struct {
  int32_t n_det;                 /* 1..GWL_MAX_DET */
  int32_t n_freq;                /* N_f = round(T fs)/2 + 1 (series.py:115-134) */
  double duration;               /* T [s] */
  double sampling_frequency;     /* fs [Hz] */
  double start_time;             /* GPS start of the segment */
  const int32_t* mask_lo;        /* host [n_det]: first bin with f >= f_min (strain_data.py:142-164) */
  const int32_t* mask_hi;        /* host [n_det]: last bin with f <= min(f_max, fs/2) */
  const double* strain;          /* host [n_det][n_freq][2] frequency-domain data (complex128) */
  const double* psd;             /* host [n_det][n_freq] one-sided PSD on the grid */
  const double* detector_tensor; /* host [n_det][9] (detector/geometry.py:250-269) */
  const double* vertex;          /* host [n_det][3] metres (detector/geometry.py:200-214) */
  int32_t approximant;           /* enum gwl_approximant */
  double wf_minimum_frequency;   /* waveform band (source.py:618-619); <= 0: none */
  double wf_maximum_frequency;   /* <= 0: the grid's last frequency */
  double reference_frequency;    /* TaylorF2 f_ref (source.py:338-343 default 50 Hz) */
  uint32_t flags;                /* enum gwl_flags */
  double time_prior_minimum;     /* Uniform geocent_time prior (base.py:794-820) */
  double time_prior_maximum;
  double reference_distance;     /* d_ref = prior.rescale(0.5) (base.py:216) */
} gwl_event_desc;

/* Context on one HIP device. */
int gwl_ctx_create(int32_t device, gwl_ctx** out);
int gwl_ctx_destroy(gwl_ctx* ctx);
const char* gwl_last_error(void);

/* Upload one event and make it resident in HBM.  Replaces the per-call
 * re-reads of bilby's strain/PSD/mask accessors (strain_data.py:212-233,
 * interferometer.py:551-564, psd.py:246-251).  *hbm_bytes receives the
 * device bytes the event holds. */
int gwl_event_create(gwl_ctx* ctx, const gwl_event_desc* desc, gwl_event** out,
                     int64_t* hbm_bytes);
int gwl_event_destroy(gwl_event* ev);

/* Attach the distance-marginalisation interpolant: a bicubic tensor
 * B-spline (knots tx[nx+4], ty[ny+4], coefficients c[nx][ny]) evaluated
 * with value -inf outside bbox = {xmin, xmax, ymin, ymax}.  Replaces
 * BoundedRectBivariateSpline (bilby/core/utils/calculus.py:221-262) built
 * in base.py:916-934. */
int gwl_event_set_distance_spline(gwl_event* ev, int32_t nx, int32_t ny, const double* tx,
                                  const double* ty, const double* coeffs, const double* bbox);

/* Time prior other than Uniform (base.py:794-820 weighs each time by
 * priors['geocent_time'].prob(times) * delta_tc; base.py:362-369): its
 * density at n >= 2 points t0 + k dt, interpolated linearly at each
 * (jittered) time; the time_prior_minimum / maximum of the descriptor bound
 * the window as before.  Replaces the Uniform weight delta_tc / (max - min).
 * The event must marginalise over time. */
int gwl_event_set_time_prior(gwl_event* ev, int64_t n, double t0, double dt,
                             const double* density);

/* Pre-size the event's device workspace for batches up to max_rows (so
 * that gwl_eval allocates nothing afterwards). */
int gwl_event_reserve(gwl_event* ev, int64_t max_rows);

/* Noise log-likelihood of the event's data, float64 on the device:
 * -sum_det |4/T sum_mask |d|^2 / S| / 2 (base.py:402-413). */
int gwl_noise_log_likelihood(gwl_event* ev, double* out);

/* Device bytes the event currently holds (data, spline, workspace): the
 * figure a server admits against its HBM budget. */
int gwl_event_device_bytes(gwl_event* ev, int64_t* bytes);

/* Batched log-likelihood ratio: out[i] = lnL ratio of params[i*ld ..].
 * Replaces base.py:419-446 log_likelihood_ratio (one row per call there).
 * params, out: device pointers.  aux (device, may be NULL) receives per row
 * {Re<d|h>, Im<d|h>, <h|h>} summed over detectors (base.py:126-134) for
 * evaluations without time marginalisation.  Rows whose waveform is
 * undefined (non-physical masses, ...) get -1.7976931348623157e308 as
 * base.py:424-425 does. */
int gwl_eval(gwl_event* ev, const double* params, int64_t n_rows, int32_t ld,
             int32_t precision, double* out, double* aux, void* stream);

/* Polarizations of the next evaluations of a GWL_HOST_POLARIZATIONS event:
 * `polarizations` is device memory [n_rows][2][n_freq] of complex128 stored
 * as (re, im) float64 pairs, h_plus then h_cross of row r on the event's
 * whole frequency array (bins 0 .. n_freq - 1, as
 * WaveformGenerator.frequency_domain_strain returns them,
 * waveform_generator.py:113-141); row r of every later gwl_eval /
 * gwl_eval_time_series call (n_rows at most) reads row r here.  The caller
 * owns the buffer and keeps it until those evaluations have completed on
 * their stream.  Replaces the source-model call of calculate_snrs
 * (base.py:260-354) for source models the library does not evaluate; the
 * projection (interferometer.py:303-368), inner products and
 * marginalisations then run as for the built-in models. */
int gwl_event_set_polarizations(gwl_event* ev, const double* polarizations, int64_t n_rows);

/* gwl_eval that allocates nothing.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The evaluation's
 * workspace (row constants, partial sums, the float32 escalation list) is a
 * file-scope device array of GWL_NOALLOC_SCRATCH_BYTES per device, held
 * until the process ends, instead of the event's own workspace; a batch
 * whose workspace would exceed it is rejected with GWL_E_INVALID (no
 * partial evaluation).  The caller passes params, out and aux as for
 * gwl_eval, with the same results.  Every call on one device shares that
 * device's array: calls that overlap on one GPU (other streams or threads,
 * gwl_eval_noalloc on any event) race, and must be serialised by the
 * caller; the same event must not be evaluated concurrently by gwl_eval. */
#define GWL_NOALLOC_SCRATCH_BYTES (64u << 20)
int gwl_eval_noalloc(gwl_event* ev, const double* params, int64_t n_rows, int32_t ld,
                     int32_t precision, double* out, double* aux, void* stream);

/* The time-marginalisation arrays of n rows without the marginalisation:
 * d_inner_h_array = 4/T FFT_N(h conj(d) / S) at the event's time window
 * (base.py:325-330; bins j_first .. j_first + n_window - 1 of
 * gwl_event_time_window, transform bin j at shift j T / N from the row's
 * geocent time) and <h|h>.  For posterior reconstruction of the time
 * (base.py:580-640).  d_inner_h: device [n][n_window][2]; h_inner_h: device
 * [n].  The event must marginalise over time. */
int gwl_eval_time_series(gwl_event* ev, const double* params, int64_t n_rows, int32_t ld,
                         int32_t precision, double* d_inner_h, double* h_inner_h, void* stream);
/* The event's time window: first transform bin, number of bins, spacing T / N. */
int gwl_event_time_window(gwl_event* ev, int64_t* j_first, int32_t* n_window, double* t_step);

/* Tile selection for this event's kernels (0 = built-in heuristic):
 * rows_per_workgroup = parameter rows sharing one workgroup's bin loop
 * (1, 2, 4, 8 or 16), workgroups_per_cu = target resident workgroups per CU
 * used to split the time-marginalised transform into groups.  With
 * rows_per_workgroup = 0 a TaylorF2 event whose time window fits a 512-point
 * transform uses the wave-per-row transform, and workgroups_per_cu is the
 * target wavefronts per CU (default 16).  Float64 evaluations without time
 * marginalisation use 4 rows per workgroup whatever is set here.  Used by the
 * autotuner; no counterpart in the reference. */
int gwl_event_set_tile(gwl_event* ev, int32_t rows_per_workgroup, int32_t workgroups_per_cu);

/* This event's tile inside multi-event launches: decimated sub-sequence
 * steps per work item (0 = the default, ~32), read when a multi-event set
 * (gwl_multi_create, gwl_server_add_group) is created.  Set by the
 * autotuner; no counterpart in the reference. */
int gwl_event_set_multi_steps(gwl_event* ev, int32_t steps);

/* Transform of the time-marginalised TaylorF2 path: automatic (the
 * wave-per-row or tiled kernel as gwl_event_set_tile selects) or one of
 * them forced.  No counterpart in the reference (kernel choice). */
enum gwl_transform {
  GWL_TRANSFORM_AUTO = 0, GWL_TRANSFORM_TILED = 1, GWL_TRANSFORM_WAVE = 2
};
int gwl_event_set_transform(gwl_event* ev, int32_t transform);

/* Per-row precision escalation of float32 evaluations without time
 * marginalisation (no counterpart in the reference, which is float64
 * throughout): within the same gwl_eval call and stream, every row whose
 * float32 error bound coeff * Q exceeds tol is evaluated again in float64
 * and its float64 result returned, so each row meets |dlnL| <= tol against
 * float64.  Q is the row's inner-product scale, |<d|h>| + <h|h> (under
 * distance marginalisation 2 dt^2 / ht + dt + ht of the table arguments
 * dt = <d|h> d / d_ref, ht = <h|h> (d / d_ref)^2).  coeff = 0 turns it off;
 * defaults coeff = 2e-7 (the largest float32 |dlnL| / Q measured over 10^4
 * sine-Gaussian rows with |lnL| up to 5e5 is 1.4e-7), tol = 8e-4.  The workspace holds the selected rows
 * (gwl_event_reserve sizes it), so the call still allocates nothing. */
int gwl_event_set_escalation(gwl_event* ev, double coeff, double tol);
/* Rows escalated by the last escalating gwl_eval on this event (waits for
 * the device; for measurement). */
int gwl_event_escalated_rows(gwl_event* ev, int64_t* rows);

/* Kernel timing for measurement: while enabled, every gwl_eval on this event
 * brackets its dominant kernel (the fused inner-product or time-marginalised
 * kernel) with HIP events on the launch stream and waits for them; the
 * accumulated milliseconds and launch count are read (and reset) with
 * gwl_event_timing_read.  No counterpart in the reference (measurement). */
int gwl_event_timing(gwl_event* ev, int32_t enable);
int gwl_event_timing_read(gwl_event* ev, double* total_ms, int64_t* launches);

/* Inner products of one parameter row with a per-bin complex factor F on
 * the response (base.py:286-294: the signal times the drawn calibration
 * curve, signal[mask] *= calibration_draws[ifo][recalib_index]), summed over
 * detectors: out[0] + i out[1] = <d|h F> = 4/T sum conj(h F) d / S,
 * out[2] = <h F|h F>, over the masked bins (Nyquist included when masked).
 * params: host [GWL_NPARAM] (sky frame); factor: host [n_det][n_freq][2];
 * out: host [3].  Float64; allocates its own workspace (a reconstruction
 * call, not for the evaluation path). */
int gwl_weighted_inner_products(gwl_event* ev, const double* params, const double* factor,
                                double* out);

/* The transform of one row's zero-padded integrand at any window (the time
 * reconstruction of base.py:604-629 for a geocent_time whose window the
 * upsampled companion event does not hold):
 *   d_inner_h[jj] = 4/T sum_d sum_n h_n conj(d_n) / S_n exp(-2 pi i n j / n_fft),
 *   j = j_first + jj (j may pass n_fft, the exponential wraps), over each
 *   detector's masked bins; h_inner_h = <h|h>.
 * A direct float64 sum (one thread per window bin), for single rows.
 * params: host [GWL_NPARAM] (sky frame); d_inner_h: host [n_window][2];
 * h_inner_h: host [1].  n_fft >= n_freq; allocates its own workspace. */
int gwl_window_time_series(gwl_event* ev, const double* params, int64_t n_fft, int64_t j_first,
                           int32_t n_window, double* d_inner_h, double* h_inner_h);

/* Detector response of one parameter row: bilby
 * Interferometer.get_detector_response (interferometer.py:303-368) of the
 * event's source model.  params: host [GWL_NPARAM]; out: host
 * [n_freq][2] complex128. */
int gwl_detector_response(gwl_event* ev, const double* params, int32_t det,
                          int32_t precision, double* out);

/* Distance-marginalisation lookup table on the GPU: base.py:994-1018
 * (_create_lookup_table).  distance_array/prior_array: host [n_dist];
 * d_axis: host [n_d] <d|h> axis, h_axis: host [n_h] <h|h> axis (the
 * reference uses n_h = 400, n_d = 800 on the axes of base.py:890-914);
 * out: host [n_h][n_d] float64, normalised by the prior as base.py:1013-1017. */
int gwl_distance_table(gwl_ctx* ctx, const double* distance_array, const double* prior_array,
                       int32_t n_dist, double reference_distance, int32_t phase_marginalization,
                       const double* d_axis, int32_t n_d, const double* h_axis, int32_t n_h,
                       double* out);

/* Host-side coefficient solve of the interpolating (s = 0) bicubic spline on
 * a rectangular grid x[nx], y[ny], z[nx][ny] with not-a-knot knots, the
 * spline scipy's RectBivariateSpline(kx=3, ky=3, s=0) builds
 * (calculus.py:221-237).  Outputs tx[nx+4], ty[ny+4], c[nx][ny]. */
int gwl_spline2d_coefficients(int32_t nx, int32_t ny, const double* x, const double* y,
                              const double* z, double* tx, double* ty, double* c);

/* Host-side evaluation of such a spline at m points (for checks on hosts
 * without a GPU; the device evaluates the same formula). */
int gwl_spline2d_eval_host(int32_t nx, int32_t ny, const double* tx, const double* ty,
                           const double* c, int64_t m, const double* x, const double* y,
                           double* out);

/* Batched complex FFT of length P (power of two, 2..16 x the longest LDS
 * transform: 131072 in float64, 262144 in float32 on gfx950) with the LDS
 * radix kernel used by time marginalisation (numpy.fft.fft as called at
 * base.py:326); lengths past one workgroup's LDS run as a radix step over
 * global memory and LDS transforms of the parts (a temporary of batch x P
 * elements allocated on the stream).  in/out: device [batch][P][2];
 * precision selects float32 (GWL_FP32) or float64 data. */
int gwl_fft_batch(gwl_ctx* ctx, const void* in, void* out, int64_t batch, int32_t P,
                  int32_t precision, void* stream);

/* Natural log of the modified Bessel function I0 (bilby/gw/utils.py:1006-1022
 * ln_i0) evaluated on the device: in/out device float64 [n]. */
int gwl_ln_i0(gwl_ctx* ctx, const double* in, double* out, int64_t n, void* stream);

/* ------------------------------------------------------------------ many events */
/* Uneven batches over many events (segments of different lengths, sampling
 * rates and detector sets) in one persistent launch that drains a
 * longest-first work list, followed by one epilogue launch; the pair can be
 * captured once into a hipGraph and replayed.  The reference evaluates one
 * likelihood object per process (bilby/core/sampler/base_sampler.py:772-800);
 * this has no direct counterpart there.  Every event must be
 * time-marginalised with a 512-sample transform tile (time-prior windows up
 * to 512 samples), <= 3 detectors and one source model. */
typedef struct gwl_multi gwl_multi;
/* params_dev[e]: device [max_rows[e]][GWL_NPARAM]; out_dev[e]: device
 * [max_rows[e]]; both caller-owned and fixed for the object's lifetime. */
int gwl_multi_create(gwl_ctx* ctx, int32_t n_events, gwl_event* const* events,
                     const double* const* params_dev, double* const* out_dev,
                     const int64_t* max_rows, gwl_multi** out);
/* Rows to evaluate per event in the next launches (<= max_rows). */
int gwl_multi_set_rows(gwl_multi* m, const int64_t* rows, void* stream);
/* Work-item size per event in decimated sub-sequence steps (0 = default,
 * ~32); the autotuner sets it per bin-count bucket. */
int gwl_multi_set_tile(gwl_multi* m, const int32_t* steps_per_item);
/* Launch (replays the captured graph if gwl_multi_capture was called). */
int gwl_multi_launch(gwl_multi* m, void* stream);
/* Capture the launch sequence on a non-default stream into a hipGraph. */
int gwl_multi_capture(gwl_multi* m, void* stream);
int gwl_multi_destroy(gwl_multi* m);

/* ------------------------------------------------------------------ ROQ */
/* Reduced-order-quadrature likelihood, bilby/gw/likelihood/roq.py
 * (ROQGravitationalWaveTransient: calculate_snrs roq.py:467-549, time grid
 * roq.py:320-331, five-sample interpolation roq.py:576-651). */
typedef struct gwl_roq gwl_roq;

enum gwl_roq_mode {
  GWL_ROQ_FP64 = 0,    /* float64 FMA GEMM (reference) */
  GWL_ROQ_FP32 = 1,    /* float32 MFMA GEMM */
  GWL_ROQ_BF16X3 = 2   /* split-bfloat16 x3 MFMA GEMM, float32 accumulation */
};

typedef struct {
  int32_t n_det;
  int32_t n_linear;              /* L linear frequency nodes */
  int32_t n_quadratic;           /* quadratic frequency nodes */
  int32_t n_time_samples;        /* rows of each detector's linear weights */
  double start_time;
  double time_samples_start;     /* weights['time_samples'][0] (s from start_time) */
  double time_samples_step;      /* their spacing */
  const double* linear_nodes;    /* host [n_linear] Hz */
  const double* quadratic_nodes; /* host [n_quadratic] Hz */
  const double* linear_weights;  /* host [n_det][n_time_samples][n_linear][2] complex */
  const double* quadratic_weights; /* host [n_det][n_quadratic][2] complex (real part used) */
  const double* detector_tensor; /* host [n_det][9] */
  const double* vertex;          /* host [n_det][3] */
  double reference_frequency;    /* TaylorF2 f_ref */
  double wf_minimum_frequency;
  uint32_t flags;                /* enum gwl_flags */
  double time_prior_minimum, time_prior_maximum;
  double delta_tc;               /* requested time-grid spacing (roq.py:320-331) */
  double reference_distance;
  const double* time_samples;    /* host [n_time_samples]: weights['time_samples'] itself, read
                                    by the single-time interpolation (roq.py:576-602) */
} gwl_roq_desc;

/* ROQGravitationalWaveTransient.__init__ (roq.py:99-300): weights resident
 * in HBM in the layout the GEMM reads. */
int gwl_roq_create(gwl_ctx* ctx, const gwl_roq_desc* desc, gwl_roq** out, int64_t* hbm_bytes);
int gwl_roq_destroy(gwl_roq* roq);
/* Distance interpolant, as gwl_event_set_distance_spline. */
int gwl_roq_set_distance_spline(gwl_roq* roq, int32_t nx, int32_t ny, const double* tx,
                                const double* ty, const double* coeffs, const double* bbox);
/* Time prior other than Uniform, as gwl_event_set_time_prior (roq.py uses
 * the same time_marginalized_likelihood weights, base.py:794-820). */
int gwl_roq_set_time_prior(gwl_roq* roq, int64_t n, double t0, double dt, const double* density);
/* Workspace for up to max_rows rows in every mode (as gwl_event_reserve). */
int gwl_roq_reserve(gwl_roq* roq, int64_t max_rows);
/* Batched ROQ log-likelihood ratio (roq.py:463-465 -> base.py:419-446):
 * params, out device pointers; mode enum gwl_roq_mode. */
int gwl_roq_eval(gwl_roq* roq, const double* params, int64_t n_rows, int32_t ld, int32_t mode,
                 double* out, void* stream);
/* Inner products of the ROQ likelihood for posterior reconstruction
 * (base.py:642-742 via roq.py:467-549; roq.py:1238-1281):
 *   n_times = 0: d_inner_h [n_rows][2] = <d|h> at each row's own time (time
 *                jitter included; antenna patterns at the middle of the time
 *                prior when the likelihood marginalises over time);
 *   n_times > 0: the likelihood's time-marginalisation grid (n_times must be
 *                its length, tcmin + (k + 1/2) delta_tc): d_inner_h
 *                [n_rows][n_times][2] = d_inner_h_array shifted by the row's
 *                time jitter (roq.py:536-540, 604-651).
 * h_inner_h [n_rows] = <h|h>.  NaN where the five samples leave the
 * weights' time grid (the reference's log(in_bounds)) or the waveform
 * failed.  Device pointers; mode enum gwl_roq_mode. */
int gwl_roq_inner_products(gwl_roq* roq, const double* params, int64_t n_rows, int32_t ld,
                           int32_t mode, int32_t n_times, double* d_inner_h, double* h_inner_h,
                           void* stream);

/* ROQ linear weights from a reduced basis (roq.py:849-918,
 * _set_weights_linear) for one detector and ROQ time samples m = k0 .. k0 +
 * n_times - 1 (time m T / N):
 *   out[k][i] = scale * sum_j x[i][j] exp(2 pi i bins[j] (k0 + k) / N)
 * x: host [n_basis][n_freq][2] complex128 = (d / S) conj(B_i) on the masked
 * bins `bins` (host [n_freq], 0 <= bins < N); scale = 4 / T; out: host
 * [n_times][n_basis][2] complex128 (bilby's weights['<ifo>_linear'] layout).
 * Float64 GEMM on the device over chunks of frequencies; allocates its own
 * workspace (a set-up call, not for the evaluation path). */
int gwl_roq_weights_linear(gwl_ctx* ctx, int32_t n_basis, int64_t n_freq, const double* x,
                           const int64_t* bins, int64_t N, int64_t k0, int32_t n_times,
                           double scale, double* out);

/* The ROQ contraction as a plain GEMM, C[M][N] = A[M][K] Bt[N][K]^T
 * (M, N multiples of 128, K of 32), for checking the MFMA kernels:
 * mode GWL_ROQ_FP64 (double), GWL_ROQ_FP32 (float), GWL_ROQ_BF16X3
 * (A = A + A_lo, Bt = Bt + Bt_lo in bfloat16, C float).  Device pointers. */
int gwl_gemm(gwl_ctx* ctx, int32_t mode, const void* A, const void* A_lo, const void* Bt,
             const void* Bt_lo, void* C, int32_t M, int32_t N, int32_t K, void* stream);

/* ------------------------------------------------------------------ relative binning
 * RelativeBinningGravitationalWaveTransient (bilby/gw/likelihood/relative.py)
 * on a TaylorF2 event; the event supplies the detectors, the source model,
 * the time / phase / distance flags and the distance spline, and must
 * outlive the relative-binning object.  An event that marginalises over
 * time also needs gwl_relbin_set_time before the first evaluation.
 *   bin_frequencies:   [n_bins + 1] edge frequencies f_e, increasing
 *                      (relative.py:177-231, setup_bins)
 *   fiducial_response: [n_det][n_bins + 1][2] complex128 detector response
 *                      of the fiducial parameters at the edges
 *                      (relative.py:233-258; must be non-zero)
 *   summary:           [n_det][n_bins][8] float64: a0, a1, b0, b1 as complex
 *                      (re, im) pairs, relative.py:314-360 (b0 and b1 are
 *                      real; their imaginary parts are ignored) */
typedef struct gwl_relbin_desc {
  int32_t n_bins;
  const double* bin_frequencies;
  const double* fiducial_response;
  const double* summary;
} gwl_relbin_desc;
typedef struct gwl_relbin gwl_relbin;
int gwl_relbin_create(gwl_event* ev, const gwl_relbin_desc* desc, gwl_relbin** out);
int gwl_relbin_destroy(gwl_relbin* rb);
int gwl_relbin_device_bytes(gwl_relbin* rb, int64_t* bytes);
/* Pre-size the per-row workspace (grown on demand by gwl_relbin_eval). */
int gwl_relbin_reserve(gwl_relbin* rb, int64_t rows);
/* ln L ratio of n rows: calculate_snrs (relative.py:393-428) + the
 * marginalisation of base.py:448-477.  params: device [n][ld] float64;
 * out (and aux [n][3] = Re <d|h>, Im <d|h>, <h|h>, may be NULL): device.
 * precision GWL_FP64 or GWL_FP32 (float32 ratios and bin sums, float64
 * phase, float64 totals).  With time marginalisation the lnL is that of
 * base.py:794-820 over the event's time window, from the transform of the
 * rebuilt full waveform (relative.py:376-391, 410-420), and aux holds the
 * inner products at the row's own time.  Timed by gwl_event_timing of the
 * event. */
int gwl_relbin_eval(gwl_relbin* rb, const double* params, int64_t n, int32_t ld,
                    int32_t precision, double* out, double* aux, void* stream);
/* Time marginalisation for relative binning (relative.py:376-391, 410-420):
 *   bin_index: [n_bins + 1] grid indices of the edges (relative.py:216-224,
 *              bin_inds)
 *   integrand: [n_det][bin_index[n_bins] - bin_index[0] + 1][2] complex128,
 *              4/T h0 conj(d) / S on the grid from bin_index[0] (h0 the
 *              fiducial response), zero at the Nyquist bin
 * The transform of the rebuilt waveform is linear in the edge ratios, so
 * its values at the event's time window are built once here (one basis
 * vector per detector and edge, float64 on the device) and each evaluation
 * is one GEMM of the rows' edge ratios against them. */
int gwl_relbin_set_time(gwl_relbin* rb, const int64_t* bin_index, const double* integrand);
/* d_inner_h_array of relative binning over the event's time window
 * (relative.py:410-420: 4/T FFT of the rebuilt full waveform times
 * conj(d) / S), for time samples drawn on the likelihood's own grid:
 * d_inner_h [n][L][2] for transform bins j_first .. j_first + L - 1 of
 * gwl_event_time_window (time start + (j + 1) t_step + the row's jitter),
 * h_inner_h [n] the binned <h|h>; NaN for rows whose waveform failed.
 * Device pointers; needs gwl_relbin_set_time. */
int gwl_relbin_time_series(gwl_relbin* rb, const double* params, int64_t n, int32_t ld,
                           int32_t precision, double* d_inner_h, double* h_inner_h,
                           void* stream);
/* Posterior reconstruction of the time for relative binning (base.py:580-640
 * with the rebuilt waveform of relative.py:380-393 as the signal):
 *   gwl_relbin_set_upsampled: bin_index as for gwl_relbin_set_time;
 *     integrand [n_det][bin_index[n_bins] - bin_index[0] + 1][2] complex128
 *     h0 conj(d) / S (no 4/T, as base.py:617-622; zero off the mask, the
 *     Nyquist bin included), weight [n_det][same] 4/T |h0|^2 / S on the
 *     mask (zero off it), n_fft = 16384 T (the upsampled transform length);
 *   gwl_relbin_upsampled_series: for n <= 8192 rows, d_inner_h
 *     [n][n_window][2] = the transform of the rebuilt h conj(d) / S at bins
 *     j_first .. j_first + n_window - 1 of n_fft (time (geocent + jitter -
 *     start) + j / 16384 modulo the segment; j may run past n_fft, which
 *     wraps), h_inner_h [n] = <h|h> of the rebuilt waveform on the mask;
 *     NaN for rows whose waveform failed.  The basis of a window (one
 *     vector per detector and edge, float64, 12 (2 n_window) (2 K) bytes)
 *     is built on the first call for it and kept until another window is
 *     asked for.  Float64; device pointers. */
int gwl_relbin_set_upsampled(gwl_relbin* rb, const int64_t* bin_index, const double* integrand,
                             const double* weight, int64_t n_fft);
int gwl_relbin_upsampled_series(gwl_relbin* rb, const double* params, int64_t n, int32_t ld,
                                int64_t j_first, int32_t n_window, double* d_inner_h,
                                double* h_inner_h, void* stream);

/* ------------------------------------------------------------------ multibanding
 * MBGravitationalWaveTransient (bilby/gw/likelihood/multiband.py; <h|h> by
 * linear interpolation, or by the IFFT-FFT algorithm after gwl_mb_set_fft)
 * on a TaylorF2 event; the event supplies
 * detectors, source model, flags and the distance spline and must outlive
 * the object.  An event that marginalises over time also needs
 * gwl_mb_set_time before the first evaluation.
 *   frequencies: [n_points] banded frequency points (multiband.py:449-478)
 *   linear:      [n_det][n_points][2] complex linear coefficients
 *                (multiband.py:529-549)
 *   quadratic:   [n_det][n_points] quadratic coefficients
 *                (multiband.py:551-611; see gwl_mb_set_fft for IFFT-FFT) */
typedef struct gwl_mb gwl_mb;
int gwl_mb_create(gwl_event* ev, int32_t n_points, const double* frequencies,
                  const double* linear, const double* quadratic, gwl_mb** out);
int gwl_mb_destroy(gwl_mb* mb);
int gwl_mb_reserve(gwl_mb* mb, int64_t rows);
int gwl_mb_device_bytes(gwl_mb* mb, int64_t* bytes);
/* ln L ratio of n rows: multiband.py:728-807 + base.py:448-477; params,
 * out, aux (may be NULL, [n][3]) device pointers; precision GWL_FP64 or
 * GWL_FP32 (float32 response and coefficients, float64 phase and sums).
 * With time marginalisation: base.py:794-820 over the window given to
 * gwl_mb_set_time (aux: the inner products at the row's own time). */
int gwl_mb_eval(gwl_mb* mb, const double* params, int64_t n_rows, int32_t ld,
                int32_t precision, double* out, double* aux, void* stream);
/* Time marginalisation for multibanding (multiband.py:714-726, 792-800):
 * d_inner_h_array = FFT_N of the array that holds, at grid index
 * grid_index[p], the sum of the points' h L terms.
 *   grid_index:   [n_points] int(f_p T_0) (_full_to_multiband), or -1 for a
 *                 point whose term does not enter (a later point of the same
 *                 band writes the same index)
 *   n_fft:        N = Nbs[-1] / 2
 *   j_first, n_window: the transform bins evaluated (time start + j t_step)
 *   t_step:       T_0 / N (also the time prior weight spacing)
 *   antenna_time: GPS time of the antenna patterns (reference_time) */
typedef struct gwl_mb_time_desc {
  const int64_t* grid_index;
  int64_t n_fft;
  int64_t j_first;
  int32_t n_window;
  double t_step;
  double antenna_time;
} gwl_mb_time_desc;
int gwl_mb_set_time(gwl_mb* mb, const gwl_mb_time_desc* td);
/* The IFFT-FFT algorithm for <h|h> (linear_interpolation=False;
 * multiband.py:613-646, 766-783; replaces the linear-interpolation sum of
 * multiband.py:760-765).  The quadratic coefficients given to gwl_mb_create
 * then hold the diagonal part: 4/T mask w / S on the first band's points and
 * (4 / T^_b) Re Ibc_b[2 j] w_j on the points of band b >= 1 at bin j (the
 * even bins of that band's 2 M_b-point transform are its half spectrum).
 * Each band b >= 1 listed here adds the odd bins:
 *   half_length:  M' = M_b / 2 (a power of two, 16 .. 2^18)
 *   first_point, n_points: the band's banded points (start_end_idxs[b])
 *   first_bin:    K_s of the band (>= 1, K_s + n_points <= M')
 *   sqrt_window:  [n_points] square_root_windows of the band
 *   odd_weights:  [n_det][M'] (4 / T^_b) Re Ibc_b[2 j + 1], j < M'
 * n_bands = 0 returns to the linear-interpolation form.  Host pointers,
 * copied. */
typedef struct gwl_mb_fft_band {
  int32_t half_length;
  int32_t first_point;
  int32_t n_points;
  int32_t first_bin;
  const double* sqrt_window;
  const double* odd_weights;
} gwl_mb_fft_band;
int gwl_mb_set_fft(gwl_mb* mb, int32_t n_bands, const gwl_mb_fft_band* bands);
/* d_inner_h_array of multibanding over the window given to gwl_mb_set_time
 * (multiband.py:792-800), for the time reconstruction of
 * multiband.py:813-848: d_inner_h [n_rows][n_window][2] for transform bins
 * j_first .. j_first + n_window - 1 (time start + j t_step; the row's time
 * jitter enters the waveform's time shift), h_inner_h [n_rows] = <h|h>;
 * NaN for rows whose waveform failed.  Device pointers. */
int gwl_mb_time_series(gwl_mb* mb, const double* params, int64_t n_rows, int32_t ld,
                       int32_t precision, double* d_inner_h, double* h_inner_h, void* stream);

/* ------------------------------------------------------------------ calibration
 * Calibration marginalisation (base.py:332-346, 860-877, lookup
 * base.py:1037-1051); the event supplies data, source model, time / phase /
 * distance flags and the distance spline and must outlive the calibration
 * object.
 *   curves: [n_curves][K][2] complex128 response curves, K = the masked bins of
 *           every detector in detector order (bilby's calibration_draws[ifo]
 *           on frequency_array[frequency_mask], concatenated).
 * Device memory without time marginalisation: the curves in three
 * precisions, (2 n_curves x 2K + n_curves x K) x 16 bytes; the workspace per
 * row ~ 36 K bytes.  With time marginalisation (base.py:305-324; a TaylorF2
 * event on the wave-per-row transform, time window <= 512 bins): the curves
 * in the transform's order as float32 pairs hi + lo (the float64 value to
 * 48 bits) and per-curve prefix sums for <h|h>, n_curves x n_det x N x 24
 * bytes, and the N twiddles W_N^m in float64; per row (1 + n_curves) x 8
 * + 64 bytes. */
typedef struct gwl_calib gwl_calib;
int gwl_calib_create(gwl_event* ev, int32_t n_curves, const double* curves, gwl_calib** out);
int gwl_calib_destroy(gwl_calib* cal);
int gwl_calib_reserve(gwl_calib* cal, int64_t rows);
int gwl_calib_device_bytes(gwl_calib* cal, int64_t* bytes);
/* ln L ratio marginalised over the curves for n rows: the per-bin integrands
 * are contracted with the curves on MFMA in `mode` (enum gwl_roq_mode:
 * GWL_ROQ_FP64, GWL_ROQ_FP32, GWL_ROQ_BF16X3; the uncalibrated totals are
 * float64 in every mode).  With time marginalisation every curve's window
 * transform runs (GWL_ROQ_FP32: float32 integrand and transform, float64
 * window sums; GWL_ROQ_FP64: all float64, the curves to 48 bits) and each
 * curve's value is time-marginalised
 * (base.py:794-820) before the logsumexp over curves (:877).  per_curve (may
 * be NULL): device [n][n_curves], each curve's ln L
 * (get_calibration_log_likelihoods, base.py:822-858). */
int gwl_calib_eval(gwl_calib* cal, const double* params, int64_t n_rows, int32_t ld, int32_t mode,
                   double* out, double* per_curve, void* stream);

/* ------------------------------------------------------------------ serving
 * Continuous-batching request server (no counterpart in the reference, whose
 * parallelism is a pool of single-dict callers, base_sampler.py:772-800).
 * One scheduler thread keeps at most one batch in flight per unit (an event,
 * a ROQ likelihood, or a group of events evaluated by one multi-event
 * launch); requests carry host rows and a host output buffer that must stay
 * valid until gwl_server_wait returns.  Latency is measured from submission
 * to the copy of the results into the output buffer. */
typedef struct gwl_server gwl_server;
int gwl_server_create(gwl_ctx* ctx, int64_t max_rows, gwl_server** out);
int gwl_server_destroy(gwl_server* server);
/* Units; each returns the id (ids first_id .. first_id + n - 1 for a group)
 * that gwl_server_submit takes.  Add every unit before submitting. */
int gwl_server_add_event(gwl_server* server, gwl_event* ev, int32_t precision, int32_t* id);
int gwl_server_add_roq(gwl_server* server, gwl_roq* roq, int32_t mode, int32_t* id);
int gwl_server_add_group(gwl_server* server, int32_t n_events, gwl_event* const* events,
                         int32_t* first_id);
/* A relative-binning likelihood (gwl_relbin_create) as a unit; precision
 * GWL_FP32 or GWL_FP64. */
int gwl_server_add_relbin(gwl_server* server, gwl_relbin* rb, int32_t precision, int32_t* id);
/* A calibration-marginalised likelihood as a unit (gwl_calib_eval in `mode`,
 * enum gwl_roq_mode); no counterpart in the reference. */
int gwl_server_add_calib(gwl_server* server, gwl_calib* cal, int32_t mode, int32_t* id);
/* rows: host [n][16] float64, out: host [n] float64, 1 <= n <= max_rows. */
int gwl_server_submit(gwl_server* server, int32_t id, const double* rows, int64_t n_rows,
                      double* out, int64_t* ticket);
/* Blocks until the request is done; returns its status. */
int gwl_server_wait(gwl_server* server, int64_t ticket, double* latency_s);
int gwl_server_stats(gwl_server* server, int64_t* batches, int64_t* rows);
/* Device bytes of the server's own buffers (staging, group partials). */
int gwl_server_device_bytes(gwl_server* server, int64_t* bytes);
/* Serve requests from other processes on a Unix stream socket created at
 * `path` (mode 0600; the caller picks a private directory).  Each connection
 * sends requests and reads replies in order, native byte order:
 *   request {uint32 GWL_WIRE_REQUEST, int32 id, int64 n} + n x 16 float64 rows
 *   reply   {uint32 GWL_WIRE_REPLY, int32 status, int64 n} + n float64 lnL
 *           (status 0) or n bytes of error text (status != 0)
 * with 1 <= n <= max_rows; requests from all connections share launches with
 * gwl_server_submit.  gwl_server_destroy closes the socket.  Replaces the
 * pickled likelihood that bilby's pool workers hold (base_sampler.py:772-800). */
#define GWL_WIRE_REQUEST 0x51574C47u
#define GWL_WIRE_REPLY 0x52574C47u
int gwl_server_listen(gwl_server* server, const char* path);

#ifdef __cplusplus
}
#endif

#endif /* GWL_H */
