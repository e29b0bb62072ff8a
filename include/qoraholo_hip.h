/* C-ABI of the MI355X holotomography backend (Backend::hip of QoraHolo).
 *
 * Plain pointers, sizes and POD structs only. Every entry point returns a qh_status; on error
 * qh_hip_last_error() holds a message for the calling thread. Host-array entry points copy in,
 * compute on the device and copy out, as the reference's CUDA backend does per call; arguments
 * are validated here with the reference's rules (src/api.cpp:13-60,107-123), so the host layer
 * above may pass them through unchecked.
 *
 * Complex data are interleaved (re, im) float pairs: qh_c64. Arrays are row-major (z, y, x),
 * x fastest, centre-origin in real space, as the reference's Array3 (include/qoraholo/array.hpp).
 */
#ifndef QORAHOLO_HIP_H
#define QORAHOLO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The host layers rethrow them as the reference does: INVALID_ARGUMENT as
 * std::invalid_argument (src/api.cpp:13-15), RUNTIME and OUT_OF_MEMORY as std::runtime_error
 * (src/cuda/util.cuh:16-26), LOGIC as std::logic_error (src/pipeline.cpp:108). */
typedef enum {
  QH_OK = 0,
  QH_ERR_INVALID_ARGUMENT = 1,
  QH_ERR_RUNTIME = 2,
  QH_ERR_OUT_OF_MEMORY = 3,
  QH_ERR_LOGIC = 4
} qh_status;

typedef struct {
  float re;
  float im;
} qh_c64;

/* include/qoraholo/odt.hpp:13-18 (Geometry). Micrometres. */
typedef struct {
  float wavelength;
  float n_medium;
  float na;
  float pixel;
} qh_geometry;

/* include/qoraholo/odt.hpp:22-25 (Illumination), cycles per micrometre. */
typedef struct {
  float fy;
  float fx;
} qh_illumination;

/* src/backends.hpp:20-25 (CropPlan). */
typedef struct {
  int cy;
  int cx;
  float radius;
  int out;
} qh_crop_plan;

/* src/common.hpp:36-44 (OdtGrid). */
typedef struct {
  int n;
  int nz;
  float df;
  float dfz;
  float fm2;
  float fna2;
  float gain;
} qh_odt_grid;

/* include/qoraholo/odt.hpp:46-57 (OdtOptions without the backend). solver: 0 Gerchberg-
 * Papoulis, 1 FISTA-TV. gp_transform: 0 r2c, 1 c2c. mapping: 0 gather, 1 scatter; both run
 * the deterministic gather here, which odt.hpp:26-29 allows. */
typedef struct {
  int nz;
  int iterations;
  int solver;
  int gp_transform;
  float tv_weight;
  int tv_inner;
  int mapping;
} qh_odt_options;

/* src/backends.hpp:31-42 (OdtTimings), device event intervals in milliseconds, plus the summed
 * duration of the FGP dual-step kernels, the steps and the kernel passes they took (two steps
 * per pass where the row width allows; FISTA-TV only). */
typedef struct {
  size_t bytes;
  double setup;
  double upload;
  double fft2;
  double clear;
  double mapping;
  double normalize;
  double solve;
  double finish;
  double download;
  double fgp_ms;
  int fgp_steps;
  int fgp_launches;
} qh_odt_timings;

/* include/qoraholo/pipeline.hpp:31-43 (PipelineTimings), plus the FGP kernel totals. */
typedef struct {
  double upload;
  double fields;
  double rytov;
  double fft2;
  double clear;
  double mapping;
  double normalize;
  double solve;
  double finish;
  double download;
  double total;
  double fgp_ms;
  int fgp_steps;
  int fgp_launches;
} qh_pipeline_timings;

/* Message of the last failed call on this thread ("" when none). */
const char* qh_hip_last_error(void);

/* src/backends.hpp:109 (cuda::device_present) and include/qoraholo/backend.hpp:8
 * (cuda_available): 1 when a HIP device is visible. */
int qh_hip_device_present(void);

/* ---- host helpers the dispatch layer uses (no device work) ---- */

/* src/api.cpp:32-42 (make_grid). */
int qh_make_grid(int n, int nz, const qh_geometry* geometry, qh_odt_grid* out);

/* src/api.cpp:52-60 (crop_plan): integer centre by lround, default radius |c| / 3, checks. */
int qh_resolve_crop(float center_y, float center_x, float radius, int output_size, int n,
                 qh_crop_plan* out);

/* src/api.cpp:203-216 (circular_scan). */
int qh_circular_scan(int count, float na_illumination, const qh_geometry* geometry,
                     qh_illumination* out);

/* src/cpu/field.cpp:62-74 (field_ratio, host-only in the reference): out = sample / background
 * with std::complex<float> division, per pixel of `count` planes of `plane` values; bcount is 1
 * (one background for every plane) or count. Runs on the library's copy threads. */
int qh_field_ratio(const qh_c64* sample, int count, const qh_c64* background, int bcount,
                   int plane, qh_c64* out);

/* No counterpart in the reference (its bindings copy arrays with one thread,
 * python/bindings.cpp:27-32): dst = src for `bytes` bytes of host memory, spread over the
 * library's copy threads (half the CPUs the process may use, at most 8). The copies into and
 * out of pinned staging buffers of every host-array entry point use the same threads. */
int qh_host_copy(void* dst, const void* src, size_t bytes);

/* The noise stream of src/phantom.cpp:200-201: std::mt19937_64(seed) drawn through
 * std::normal_distribution<float>(0, sigma), `count` values in order. */
int qh_gaussian_noise(uint64_t seed, float sigma, size_t count, float* out);

/* ---- per-call entry points, host arrays (src/backends.hpp:112-118) ---- */

/* cuda::extract_fields: holograms (count, n, n) -> fields (count, plan.out, plan.out). */
int qh_hip_extract_fields(const float* holograms, int count, int n, const qh_crop_plan* plan,
                          qh_c64* fields);

/* cuda::unwrap: wrapped phase (nz, ny, nx) -> unwrapped phase. */
int qh_hip_unwrap(const float* wrapped, int nz, int ny, int nx, float* out);

/* cuda::rytov: field ratio (nz, ny, nx) -> Rytov field. */
int qh_hip_rytov(const qh_c64* ratio, int nz, int ny, int nx, qh_c64* out);

/* cuda::reconstruct: Rytov fields (count, n, n) -> index (nz, n, n) and the filled fraction.
 * timings may be NULL. */
int qh_hip_reconstruct(const qh_c64* fields, int count, int n, const qh_illumination* illum,
                       const qh_odt_grid* grid, const qh_geometry* geometry,
                       const qh_odt_options* options, float* ri, double* filled,
                       qh_odt_timings* timings);

/* The Ewald mapping alone (the first stage of cuda::reconstruct, src/cuda/odt.cu:243-262):
 * normalised k-space potential (nz, n, n) and per-voxel hit counts, for checks of the
 * intermediate. No counterpart entry in the reference. */
int qh_hip_map_kspace(const qh_c64* fields, int count, int n, const qh_illumination* illum,
                      const qh_odt_grid* grid, qh_c64* potential, uint16_t* counts,
                      double* filled);

/* Sideband of every hologram of a stack: find_sideband (src/cpu/field.cpp:35-60) on the device.
 * The reference runs it on the host for the first hologram only; here the spectra and the peak
 * search run on the device and the sub-bin refinement (field.cpp:24-33) on the host.
 * dc_exclusion <= 0 means n / 8. out: count x 5 floats (iy, ix, y, x, magnitude). */
int qh_hip_find_sidebands(const float* holograms, int count, int n, float dc_exclusion,
                          float* out);
/* Illumination frequencies estimated from background holograms as examples/readme_figure.py:39-46
 * does: the carrier c is the sideband of background 0 and f_a = (s_a - c) / (n pixel), with the
 * hologram pixel. carrier receives c as 5 floats (iy, ix, y, x, magnitude). The cold-start step
 * of a new geometry in the served loop. */
int qh_hip_estimate_illuminations(const float* backgrounds, int count, int n,
                                  float hologram_pixel, float dc_exclusion, qh_illumination* out,
                                  float* carrier);

/* ---- Pipeline (include/qoraholo/pipeline.hpp:50-85; src/backends.hpp:75-89,110-111) ---- */

typedef struct qh_pipeline qh_pipeline;

typedef struct {
  int count;          /* holograms per stack */
  int hologram_size;  /* side N of the square holograms */
  qh_crop_plan crop;  /* resolved from FieldOptions and the carrier */
  const qh_illumination* illuminations; /* count entries, copied */
  qh_geometry geometry;
  qh_odt_options odt;
  int upload_chunks;
} qh_pipeline_options;

int qh_hip_pipeline_create(const qh_pipeline_options* options, qh_pipeline** out);
void qh_hip_pipeline_destroy(qh_pipeline* p);
/* Background holograms (count, N, N), host memory; fields are extracted once and kept. */
int qh_hip_pipeline_set_background(qh_pipeline* p, const float* backgrounds);
/* Page-locked input (count, N, N) and output (nz, n, n) buffers owned by the pipeline. */
float* qh_hip_pipeline_input(qh_pipeline* p);
const float* qh_hip_pipeline_output(const qh_pipeline* p);
/* Reconstructs from input() into output(); QH_ERR_LOGIC without a background. */
int qh_hip_pipeline_process(qh_pipeline* p);
/* Pipeline::run (include/qoraholo/pipeline.hpp:72): the same from a caller's pageable
 * buffer of holograms (count, N, N) into a caller's volume (nz, n, n). The holograms pass
 * through input() in pieces whose host copies overlap their uploads, and the volume leaves in
 * parts whose host copies overlap the DMA of the next part; output() holds the volume too. */
int qh_hip_pipeline_run(qh_pipeline* p, const float* holograms, float* ri);
/* The same from holograms already in device memory into a device volume, on `stream`
 * (a hipStream_t, used as given: NULL is the null stream, which is also torch's default
 * stream). Returns once the work is queued; the caller synchronises the stream before reading
 * ri_device. Requests share the pipeline's device buffers, so the pipeline orders its own
 * work: a later process_device (on any stream), process, run or set_background starts after
 * this request's work; the caller need not synchronise between calls. */
int qh_hip_pipeline_process_device(qh_pipeline* p, const float* holograms_device,
                                   float* ri_device, void* stream);
double qh_hip_pipeline_filled(const qh_pipeline* p);
int qh_hip_pipeline_timings(const qh_pipeline* p, qh_pipeline_timings* out);
double qh_hip_pipeline_setup_ms(const qh_pipeline* p);
size_t qh_hip_pipeline_device_bytes(const qh_pipeline* p);
int qh_hip_pipeline_output_shape(const qh_pipeline* p, int* nz, int* ny, int* nx);
/* Records FGP kernel intervals in timings (a few events per outer iteration). */
int qh_hip_pipeline_time_fgp(qh_pipeline* p, int on);
/* No counterpart in the reference: 1 (default) replays the device work of a request (field
 * extraction through the index volume, stage events included) as one captured hipGraph, built
 * on the first request and again when the device input address or FGP timing changes; 0
 * launches it kernel by kernel. Results are the same bit for bit. */
int qh_hip_pipeline_use_graph(qh_pipeline* p, int on);

/* ---- Served loop (no counterpart in the reference, which serves one request per Pipeline
 * call): requests of several geometries run concurrently in device-resident slots admitted
 * against an HBM budget. ---- */

typedef struct qh_server qh_server;

typedef struct {
  int64_t submitted;
  int64_t completed;
  int64_t failed;
  int64_t queued;
  double p50_ms; /* latency from submit to result, completed requests */
  double p99_ms;
  double mean_ms;
  double recon_per_s; /* completed / (last completion - first submission) */
  size_t device_bytes;
  size_t budget_bytes;
  int slots;           /* slots resident on the device (idle ones keep their geometry's state) */
  int busy_slots;      /* slots running a request now */
  int peak_busy_slots; /* most slots running requests at once since the last reset */
  int graph_slots;     /* resident slots whose requests replay one captured hipGraph */
  int64_t batched_requests; /* requests that ran together with others of their geometry */
  int peak_batch;      /* most requests one slot ran at once since the last reset */
} qh_server_stats;

/* budget_bytes 0: 90% of the free device memory. max_slots: requests in flight at once; idle
 * slots stay resident until the budget needs their memory. */
int qh_hip_server_create(size_t budget_bytes, int max_slots, qh_server** out);
/* Finishes the requests on the device, drops the queued ones, frees everything. */
void qh_hip_server_destroy(qh_server* s);
/* A geometry (the options of a Pipeline) and its background holograms (count, N, N). */
int qh_hip_server_add_geometry(qh_server* s, const qh_pipeline_options* options,
                               const float* backgrounds, int* geometry);
/* The same with batched slots: a slot of the geometry takes up to max_batch (1 to 64) queued
 * requests of it at once, runs the front end of all their hologram stacks in one set of
 * launches and their solves in turn, as one captured graph per batch size; each result equals
 * the request's alone bit for bit. */
int qh_hip_server_add_geometry_batched(qh_server* s, const qh_pipeline_options* options,
                                       const float* backgrounds, int max_batch, int* geometry);
/* Queues holograms (count, N, N) of a geometry; the result (nz, n, n) goes to ri_out. Both
 * host buffers must stay valid until qh_hip_server_wait returns for the id. Thread-safe. */
int qh_hip_server_submit(qh_server* s, int geometry, const float* holograms, float* ri_out,
                         int64_t* id);
/* Blocks until the request is done and returns its status (QH_OK or the error it failed with,
 * message in qh_hip_last_error); one failed request does not affect the others. */
int qh_hip_server_wait(qh_server* s, int64_t id, double* filled, double* latency_ms);
int qh_hip_server_poll(qh_server* s, int64_t id, int* done);
int qh_hip_server_stats(qh_server* s, qh_server_stats* out);
int qh_hip_server_reset_stats(qh_server* s);

/* ---- beam propagation (replaces cuda::bpm_forward / bpm_reconstruct / bpm_vjp,
 *      reference src/backends.hpp:119-124, and the tables of src/bpm.cpp:10-60) ----
 * Volumes are the contrast x = ri - n_medium (nz, n, n), as the reference's backends receive
 * it (src/api.cpp:159-163); fields and residuals are (count, n, n). Lateral boundaries are
 * periodic; n must be even and of the form 2^a 3^b 5^c, at most 2048. */

/* Reference BpmOptions (include/qoraholo/bpm.hpp:18-25) plus the device layout. */
typedef struct {
  int obliquity;  /* weight the scattered part of each slice by its obliquity (reference: 1) */
  int chunk;      /* illuminations per device pass; 0: as many as keep the adjoint's buffers
                   * within half the free device memory */
  int checkpoint; /* slices per recomputed segment of the adjoint; 0: keep every envelope;
                   * -1: keep every envelope when one pass's illuminations then fit within half
                   * the free device memory (what the reference's CUDA backend stores), else 16 */
} qh_bpm_options;

/* Reference BpmSolveOptions (include/qoraholo/bpm.hpp:32-45). */
typedef struct {
  int iterations;
  int batch; /* illuminations per gradient step; 0 or count: all, with FISTA momentum */
  float step; /* multiple of 1 / L */
  float tv_weight; /* 0: n >= n_medium only; > 0: TV prox on the contrast (n <= 512) */
  int tv_inner;
} qh_bpm_solve_options;

/* A BPM inversion request kind of the served loop (no counterpart in the reference, whose
 * examples/bpm_figure.py:74-88 runs the same recipe call by call): Rytov fields of the field
 * ratios, a Gerchberg-Papoulis start of start_iterations on the (nz, n, n) grid, then
 * bpm_reconstruct with `solve` from that start. Results equal that chain of per-call functions
 * (qh_hip_rytov, qh_hip_reconstruct, qh_hip_bpm_reconstruct with the contrast conversions of
 * src/api.cpp:159-163) bit for bit. */
typedef struct {
  int n;
  int nz;
  int count;
  const qh_illumination* illuminations;
  qh_geometry geometry;
  qh_bpm_options bpm;
  int start_iterations;
  qh_bpm_solve_options solve;
} qh_bpm_request_options;

/* Registers a BPM inversion geometry with the server; its first slot is built (and its device
 * footprint measured) before the call returns. */
int qh_hip_server_add_bpm(qh_server* s, const qh_bpm_request_options* options, int* geometry);
/* Queues field ratios (count, n, n) of a BPM geometry; the refractive index (nz, n, n) goes to
 * ri_out and, when loss is not NULL, the data term of each iteration to loss (solve.iterations
 * values). Host buffers must stay valid until qh_hip_server_wait returns for the id. The same
 * wait, poll and stats calls serve both request kinds. Thread-safe. */
int qh_hip_server_submit_bpm(qh_server* s, int geometry, const qh_c64* fields, float* ri_out,
                             double* loss, int64_t* id);

/* The tables in float as bpm_tables (src/bpm.cpp:10-60) stores them: step, last (count, n, n)
 * complex, q (count, n, n), the screen phase k0 pixel and the Lipschitz constant. Host only. */
int qh_bpm_tables(int n, int nz, const qh_illumination* illum, int count,
                  const qh_geometry* geometry, int obliquity, qh_c64* step, qh_c64* last,
                  float* q, float* phase, float* lipschitz);
/* cuda::bpm_forward (src/backends.hpp:119): x (nz, n, n) -> field ratios (count, n, n) at z = 0. */
int qh_hip_bpm_forward(const float* x, int nz, int n, const qh_illumination* illum, int count,
                       const qh_geometry* geometry, const qh_bpm_options* options, qh_c64* out);
/* cuda::bpm_vjp (src/backends.hpp:124): grad (nz, n, n) = Re(J^H r) at x. */
int qh_hip_bpm_vjp(const float* x, const qh_c64* r, int nz, int n, const qh_illumination* illum,
                   int count, const qh_geometry* geometry, const qh_bpm_options* options,
                   float* grad);
/* cuda::bpm_reconstruct (src/backends.hpp:122-123): x holds the initial contrast (nz, n, n) and
 * receives the result's; loss receives solve->iterations data-term values. */
int qh_hip_bpm_reconstruct(const qh_c64* fields, int count, int n, const qh_illumination* illum,
                           const qh_geometry* geometry, float* x, int nz,
                           const qh_bpm_options* options, const qh_bpm_solve_options* solve,
                           double* loss);
/* bpm_jvp (src/cpu/bpm.cpp:141-176, CPU-only in the reference, src/backends.hpp:100-101): the
 * tangent fields J h (count, n, n) of the model linearised at x, for a direction h (nz, n, n). */
int qh_hip_bpm_jvp(const float* x, const float* h, int nz, int n, const qh_illumination* illum,
                   int count, const qh_geometry* geometry, const qh_bpm_options* options,
                   qh_c64* out);

/* The tables as src/backends.hpp:61-70 (BpmTables) holds them, borrowed for the call: step and
 * last (count, n, n) complex, q (count, n, n), lateral bins in FFT order, the screen phase
 * k0 pixel and the Lipschitz constant (what qh_bpm_tables returns). */
typedef struct {
  int n;
  int nz;
  int count;
  float phase;
  const qh_c64* step;
  const qh_c64* last;
  const float* q;
  float lipschitz;
} qh_bpm_table_view;
/* The same four calls on tables, with the arguments of cuda::bpm_forward(x, BpmTables),
 * cuda::bpm_vjp(x, r, BpmTables) and cuda::bpm_reconstruct(fields, x, BpmTables, options,
 * chunk) (src/backends.hpp:119-124) and cpu::bpm_jvp(x, h, BpmTables) (:100): volumes are
 * (tables->nz, n, n), fields (tables->count, n, n). options->obliquity is not read (the tables
 * carry it); chunk and checkpoint are the device layout as above. */
int qh_hip_bpm_forward_tables(const float* x, const qh_bpm_table_view* tables,
                              const qh_bpm_options* options, qh_c64* out);
int qh_hip_bpm_vjp_tables(const float* x, const qh_c64* r, const qh_bpm_table_view* tables,
                          const qh_bpm_options* options, float* grad);
int qh_hip_bpm_jvp_tables(const float* x, const float* h, const qh_bpm_table_view* tables,
                          const qh_bpm_options* options, qh_c64* out);
int qh_hip_bpm_reconstruct_tables(const qh_c64* fields, float* x,
                                  const qh_bpm_table_view* tables, const qh_bpm_options* options,
                                  const qh_bpm_solve_options* solve, double* loss);

/* Device-resident model for callers that keep x, fields and gradients in device memory (e.g.
 * a torch.autograd.Function): tables and work buffers stay on the device between calls. The
 * _dev calls take device pointers and a stream (a hipStream_t, used as given: NULL is the null
 * stream, torch's default), queue their work and return; the caller synchronises the stream.
 * A plan is used by one thread at a time. */
typedef struct qh_bpm_plan qh_bpm_plan;
int qh_hip_bpm_plan_create(int nz, int n, const qh_illumination* illum, int count,
                           const qh_geometry* geometry, const qh_bpm_options* options,
                           qh_bpm_plan** out);
void qh_hip_bpm_plan_destroy(qh_bpm_plan* plan);
/* Screen phase, Lipschitz constant, illuminations per pass and device bytes of the plan. */
int qh_hip_bpm_plan_info(const qh_bpm_plan* plan, float* phase, float* lipschitz, int* chunk,
                         size_t* device_bytes);
/* The plan's checkpoint interval as resolved (0: every envelope kept); no counterpart in the
 * reference, whose CUDA backend always keeps every envelope (src/cuda/bpm.cu:139-141). */
int qh_hip_bpm_plan_checkpoint(const qh_bpm_plan* plan, int* checkpoint);
int qh_hip_bpm_forward_dev(qh_bpm_plan* plan, const float* x, qh_c64* out, void* stream);
int qh_hip_bpm_vjp_dev(qh_bpm_plan* plan, const float* x, const qh_c64* r, float* grad,
                       void* stream);
int qh_hip_bpm_jvp_dev(qh_bpm_plan* plan, const float* x, const float* h, qh_c64* out,
                       void* stream);
/* gradient() of src/cpu/bpm.cpp:100-129: grad = Re(J^H (bpm(x) - fields)) / (last - first)
 * over illuminations [first, last); *loss = 0.5 / count times the sum of squared residuals.
 * Synchronises the stream to return the loss. */
int qh_hip_bpm_gradient_dev(qh_bpm_plan* plan, const float* x, const qh_c64* fields, int first,
                            int last, float* grad, double* loss, void* stream);

/* ---- transform and TV test hooks (no counterpart in the reference's ABI: it used cuFFT) ---- */

/* In-place complex FFT over the axes in `axes` (bit 0: z, 1: y, 2: x) of (d0, d1, d2). */
int qh_hip_fft_c2c(qh_c64* data, int d0, int d1, int d2, unsigned axes, int forward, float scale);
/* Real (nz, ny, nx) <-> half spectrum (nz, ny, nx/2+1). */
int qh_hip_rfft3(const float* x, int nz, int ny, int nx, qh_c64* out);
int qh_hip_irfft3(const qh_c64* in, int nz, int ny, int nx, float scale, float* x);
/* One TV prox call from a zero dual: argmin 0.5 |x - b|^2 + lambda TV(x), x >= 0
 * (src/cpu/tv.cpp:7-44). fused: 1 runs the FGP steps two per pass where the row width allows
 * (what the solvers do), 0 one per pass; the results are the same bits. */
int qh_hip_tv_prox(const float* b, int nz, int ny, int nx, float lambda, int inner, int fused,
                   float* x);
/* qh_hip_tv_prox of `batch` volumes at once (no counterpart in the reference, whose tv_prox
 * takes one): b and x hold the volumes one after another, every FGP pass and the final primal
 * run as one launch over all of them, and each volume gets the bits qh_hip_tv_prox gives it
 * alone; batch is 1 to 65535. The batched form the served loop's batched slots use. */
int qh_hip_tv_prox_batch(const float* b, int batch, int nz, int ny, int nx, float lambda,
                         int inner, int fused, float* x);

/* ---- allocation-free variant (no counterpart in the reference) ----
 * qh_hip_tv_prox for callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the same result bits, with b and x
 * device pointers of nz*ny*nx floats and the work queued on `stream` (NULL: the null stream);
 * the call returns without waiting for it. The FGP dual lives in a fixed device array of the
 * library, 96 MiB on each device it runs on, held until the process ends; volumes of more than
 * QH_TV_NOALLOC_MAX_VOXELS voxels are refused with QH_ERR_INVALID_ARGUMENT. The call makes no
 * hipMalloc or hipFree. Every call on one device shares that array (each call clears the dual
 * first): calls that overlap in time on one device, from other streams or host threads, race
 * and must be ordered by the caller. */
#define QH_TV_NOALLOC_MAX_VOXELS 2097152
int qh_hip_tv_prox_dev_noalloc(const float* b, int nz, int ny, int nx, float lambda, int inner,
                               int fused, float* x, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QORAHOLO_HIP_H */
