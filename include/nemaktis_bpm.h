/*
 * nemaktis_bpm.h - C ABI of the MI355X-native bpm-solver propagation step.
 *
 * Drop-in for the Python backend module `bpm_backend` of Nemaktis
 * (reference: BeamPropagationSolver/src/PythonBackend/bpm_backend.h:4-13,
 * bound by SWIG in bpm_backend.i:16-18 and called from
 * HighLevelPythonInterface/nemaktis/light_propagator.py:225-228).
 *
 * Array conventions (identical to the reference):
 *   lc_field_vals  (Nz, Ny, Nx, Nv) float64, C order, Nv = 3 (director) or
 *                  6 (Q-tensor xx, yy, zz, xy, xz, yz)
 *   mask_vals      (Nz, Ny, Nx) float64, >= 0 marks the LC domain
 *   E_field_vals   (N_wavelengths, N_wavevectors, 4, Ny, Nx) complex128 as
 *                  interleaved (re, im) float64 pairs; the 4 planes are
 *                  (Ex, Ey) for x-polarised then for y-polarised input.  It is
 *                  read on entry only with "Beam profile": "None", and
 *                  overwritten in place with the micrograph-plane fields.
 * Sizes are element counts (n_E_vals counts complex values).  Every entry
 * point returns 0 on success and a non-zero code on failure, with the
 * message copied into err_msg (if non-NULL); nothing is printed.  On success
 * err_msg holds the notice the reference prints to stdout, if any (a mask
 * formula that disagrees with the mask, UtilitiesLib/VectorField.cpp:107-111),
 * else the empty string.
 */
#ifndef NEMAKTIS_BPM_H
#define NEMAKTIS_BPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum nmk_bpm_status {
    NMK_BPM_OK = 0,
    NMK_BPM_EINPUT = 1,   /* invalid settings or array sizes (the reference's thrown strings) */
    NMK_BPM_EDEVICE = 2,  /* HIP runtime failure */
    NMK_BPM_ENOMEM = 3,   /* the run does not fit the device memory budget */
    NMK_BPM_EINTERNAL = 4
};

/* Replaces run_backend_without_mask (PythonBackend/bpm_backend.h:4-7,
 * bpm_backend.cpp:7-83): FP64 march, one wavelength at a time. */
int nmk_bpm_run_backend_without_mask(const char *json_str, const double *lc_field_vals, int64_t n_lc_vals,
                                     double *E_field_vals, int64_t n_E_vals, char *err_msg, size_t err_len);

/* Replaces run_backend_with_mask (PythonBackend/bpm_backend.h:9-13,
 * bpm_backend.cpp:85-164); mask_formula may be "" (no interface weights). */
int nmk_bpm_run_backend_with_mask(const char *json_str, const char *mask_formula, const double *lc_field_vals,
                                  int64_t n_lc_vals, const double *mask_vals, int64_t n_mask_vals,
                                  double *E_field_vals, int64_t n_E_vals, char *err_msg, size_t err_len);

/* ---- extensions beyond the reference interface ---------------------------- */

typedef struct {
    int32_t precision;   /* 64: complex128 march (default); 32: complex64 march
                            with FP64 permittivity / coefficient precompute */
    int32_t wl_batch;    /* wavelengths marched together; <= 0: all at once */
    int32_t use_graph;   /* 1: capture the z-march as a hipGraph and replay it
                            (contexts; one-shot nmk_bpm_run launches eagerly) */
    int32_t device;      /* HIP device ordinal, -1: current device */
    int32_t profile;     /* 1: time every kernel family (disables the graph) */
    int32_t x_chunks;    /* line-solver launch shapes; 0: built-in defaults      */
    int32_t x_rows;      /*   x lines: chunks per row, rows per block            */
    int32_t y_chunks;    /*   y lines: chunks per column, columns per block      */
    int32_t y_cols;
    int32_t item_order;  /*   0: by batch size; 1 + bits: bit 0 x components, bit 1
                              y components, bit 2 polarisation pairs adjacent
                              in the XCD work ranges, bit 3 the pairs of each
                              wavelength adjacent (bits 0-3 only)              */
    int32_t lean;        /*   1: register-lean solver variants                   */
    int32_t y_pair;      /*   1: y-line blocks hold both polarisations,
                              2: both field components of one polarisation     */
    int32_t x_both;      /*   x-line blocks hold both field components: bit 0 for
                              the first launch of a half-step, bit 1 for the
                              Neumann launches                                 */
    int32_t coef_serial; /*   1: slab coefficients on the march stream instead
                              of a second stream (no joins between streams);
                              the built-in shapes (x_chunks 0) always use 1    */
} nmk_bpm_options;

typedef struct {
    double setup_ms, march_ms, output_ms, total_ms; /* device time of the last propagate */
    double updates;                                  /* voxel-slice updates of one run */
    double updates_per_s;
    int64_t footprint_bytes;                         /* HBM held by the run */
    /* profile mode: device time per kernel family, indexed 0 material (now
       evaluated inside the coefficient kernel: always 0), 1 material and
       coefficients, 2 slab boundary, 3 reserved, 4 explicit half-step,
       5 x-line solves, 6 y-line solves, 7 output, 8 screen filter, 9 setup */
    double family_ms[10];
    int64_t family_launches[10];
} nmk_bpm_stats;

/* fills the defaults: FP64, one wavelength at a time, eager launches, built-in shapes */
void nmk_bpm_default_options(nmk_bpm_options *opts);

/* One-shot run with options (mask_vals == NULL selects the unmasked form). */
int nmk_bpm_run(const char *json_str, const char *mask_formula, const double *lc_field_vals, int64_t n_lc_vals,
                const double *mask_vals, int64_t n_mask_vals, double *E_field_vals, int64_t n_E_vals,
                const nmk_bpm_options *opts, nmk_bpm_stats *stats, char *err_msg, size_t err_len);

/* No counterpart in the reference.  Variant of nmk_bpm_run for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  The inputs and the output are the
 * caller's device arrays, in the layouts of nmk_bpm_run:
 *   lc_field_vals_dev  director / Q-tensor (n_lc_vals doubles)
 *   mask_vals_dev      float64 mask (>= 0 is LC), NULL: unmasked
 *   weights_dev        interface weights, 2*Nx*Ny*(Nz-1) doubles as
 *                      nmk_bpm_interface_weights computes them, NULL: all 0.5
 *   E_field_vals_dev   output fields (complex128 as (re, im) pairs); holds the
 *                      input fields when "Beam profile" is "None"
 * Every other buffer (slab coefficients, march state, line rows, FFT twiddles,
 * parameter blocks) is carved from one static device array of
 * NMK_BPM_NOALLOC_SCRATCH_BYTES on each GPU the library runs on, held until
 * the process ends.  The run allocates no device memory and creates no
 * streams or events: one wavelength at a time, eager launches on `stream`
 * (a hipStream_t; NULL: the per-thread default stream); it returns once the
 * output is written.  A run whose buffers exceed the scratch is refused with
 * NMK_BPM_ENOMEM before any device work; bulk output is refused with
 * NMK_BPM_EINPUT.  All calls on one GPU share its scratch array: calls that
 * overlap on one GPU race, so the caller serialises them. */
#define NMK_BPM_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int nmk_bpm_run_noalloc(const char *json_str, const double *lc_field_vals_dev, int64_t n_lc_vals,
                        const double *mask_vals_dev, int64_t n_mask_vals, const double *weights_dev,
                        int64_t n_weights, double *E_field_vals_dev, int64_t n_E_vals, int32_t precision,
                        void *stream, char *err_msg, size_t err_len);

/* Context API: inputs uploaded once and kept resident in HBM, so that repeated
 * propagations (benchmarks, parameter sweeps) time the device work alone. */
typedef struct nmk_bpm_ctx nmk_bpm_ctx;
nmk_bpm_ctx *nmk_bpm_create(const char *json_str, const char *mask_formula, const double *lc_field_vals,
                            int64_t n_lc_vals, const double *mask_vals, int64_t n_mask_vals, int64_t n_E_vals,
                            const nmk_bpm_options *opts, char *err_msg, size_t err_len);
/* input fields of "Beam profile": "None"; a propagation replaces them with
 * its output (as the reference writes E in place), so upload before each */
int nmk_bpm_upload_fields(nmk_bpm_ctx *ctx, const double *E_field_vals, int64_t n_E_vals, char *err_msg,
                          size_t err_len);
int nmk_bpm_propagate(nmk_bpm_ctx *ctx, nmk_bpm_stats *stats, char *err_msg, size_t err_len);
int nmk_bpm_download(nmk_bpm_ctx *ctx, double *E_field_vals, int64_t n_E_vals, char *err_msg, size_t err_len);
/* new line-solver launch shapes for the next propagations (the tuning fields
 * of opts; x_chunks = 0: built-in shapes); on refusal the context keeps its
 * shapes.  Used by the autotuner to time candidates on one resident job. */
int nmk_bpm_set_tuning(nmk_bpm_ctx *ctx, const nmk_bpm_options *opts, char *err_msg, size_t err_len);
void nmk_bpm_destroy(nmk_bpm_ctx *ctx);

/* HBM bytes a run of these settings would hold (serving admission). */
int nmk_bpm_estimate_footprint(const char *json_str, int64_t n_lc_vals, int has_mask, const nmk_bpm_options *opts,
                               int64_t *bytes, char *err_msg, size_t err_len);

/* Interface interpolation weights a masked run computes from its mask formula
 * (UtilitiesLib/VectorField.cpp:64-113, which keeps them in
 * VectorField::interp_weigths): weights[2*i] and weights[2*i+1] for voxel
 * i = ix + Nx*(iy + Ny*iz), iz < Nz-1, n_weights = 2*Nx*Ny*(Nz-1).  When the
 * formula's zero crossings disagree with the mask every weight is 0.5 and
 * err_msg holds the reference's notice (status 0).  Host-only: no device is
 * needed. */
int nmk_bpm_interface_weights(const char *json_str, const char *mask_formula, const double *mask_vals,
                              int64_t n_mask_vals, double *weights, int64_t n_weights, char *err_msg,
                              size_t err_len);

/* The same weights computed on the current HIP device, as a masked run
 * computes them (one thread per column and slab, the formula compiled to a postfix
 * program; arithmetic operation for operation the host version's, elementary
 * functions from the device math library).  No counterpart in the reference,
 * which evaluates the formula on the host; for callers that keep the weights
 * and for checking the device evaluation against the host one. */
int nmk_bpm_interface_weights_device(const char *json_str, const char *mask_formula, const double *mask_vals,
                                     int64_t n_mask_vals, double *weights, int64_t n_weights, char *err_msg,
                                     size_t err_len);

/* Evaluates an index expression (muParser syntax, variable `lambda` in um) as
 * the backend does for "no", "ne", ... (PhysicsLib/PhysicsCoefficients.cpp:27-103). */
int nmk_bpm_eval_expression(const char *expr, double lambda, double *value, char *err_msg, size_t err_len);

const char *nmk_bpm_version(void);

/* ---- serving loop (no counterpart in the reference, whose backend runs one
 * blocking propagation per call: PythonBackend/bpm_backend.cpp:7-164) --------
 *
 * A server owns one GPU and a pre-allocated HBM arena.  Submitted micrograph
 * jobs (the arguments of nmk_bpm_run) are validated at once, queued, and
 * admitted in arrival order whenever an execution slot and a contiguous arena
 * range of the job's footprint are free; a job that does not fit may be
 * overtaken by later jobs that do, at most max_bypass times.  Admitted jobs
 * run concurrently, each on its own stream with its state resident in its
 * arena range.  Submission only validates (settings, sizes, the mask
 * formula's syntax, footprint): the host work of size O(volume) -- the
 * per-voxel mask, the interface weights of a mask formula and the director
 * upload -- runs on the job's worker once it is admitted, so lc_field_vals,
 * mask_vals and E_field_vals are borrowed until the job has finished.
 * E_field_vals is read at admission ("Beam profile": "None") and written
 * when the job ends. */
typedef struct nmk_srv nmk_srv;

typedef struct {
    int32_t device;      /* HIP device, -1: current */
    int32_t max_active;  /* jobs on the device at once (default 4) */
    int64_t hbm_budget;  /* arena bytes; 0: 80 % of the free device memory */
    int32_t max_bypass;  /* default 8 */
    int32_t reserved[3];
} nmk_srv_config;

enum nmk_srv_job_state {
    NMK_SRV_QUEUED = 0, NMK_SRV_RUNNING = 1, NMK_SRV_DONE = 2, NMK_SRV_FAILED = 3, NMK_SRV_CANCELLED = 4
};

typedef struct {
    int32_t state;       /* nmk_srv_job_state */
    int32_t status;      /* nmk_bpm_status of a failed job */
    double t_submit_ms;  /* host clock, ms since the server was created */
    double t_admit_ms;   /* -1 until admitted */
    double t_end_ms;     /* -1 until finished */
    double device_ms;    /* device time of the propagation */
    double updates;      /* voxel-slice updates of the job */
    int64_t footprint_bytes;
    char error[256];
    double t_ready_ms;   /* -1 until the job's host preparation (mask, interface
                            weights) is done and its inputs are on the device */
    char notice[128];    /* what the reference prints on success, if anything */
} nmk_srv_job_info;

typedef struct {
    int64_t submitted, completed, failed, cancelled, running, queued;
    int64_t arena_bytes, arena_used_bytes, arena_peak_bytes;
    int32_t peak_active;
    int32_t reserved;
    double now_ms;
} nmk_srv_counters;

void nmk_srv_default_config(nmk_srv_config *cfg);
/* NULL on failure (no device, budget above the free memory, ...) */
nmk_srv *nmk_srv_create(const nmk_srv_config *cfg, char *err_msg, size_t err_len);
/* job id >= 0, or -(nmk_bpm_status) when the job is refused (invalid input,
 * or a footprint larger than the whole arena) */
int64_t nmk_srv_submit(nmk_srv *srv, const char *json_str, const char *mask_formula, const double *lc_field_vals,
                       int64_t n_lc_vals, const double *mask_vals, int64_t n_mask_vals, double *E_field_vals,
                       int64_t n_E_vals, const nmk_bpm_options *opts, char *err_msg, size_t err_len);
/* 0: finished (done, failed or cancelled), 1: timed out (timeout_ms < 0 waits
 * without limit), -1: unknown job */
int nmk_srv_wait(nmk_srv *srv, int64_t job, double timeout_ms, nmk_srv_job_info *info);
/* 0: cancelled, 1: already admitted or finished, -1: unknown job */
int nmk_srv_cancel(nmk_srv *srv, int64_t job);
int nmk_srv_stats(nmk_srv *srv, nmk_srv_counters *stats);
/* cancels queued jobs, waits for the running ones, frees the arena */
void nmk_srv_destroy(nmk_srv *srv);

#ifdef __cplusplus
}
#endif

#endif
