/*
 * ravesim.h -- C-ABI of the MI355X split-step wave-propagation library
 * (librsim.so).  Plain C types only: pointers, sizes, doubles, status codes.
 *
 * The reference has no C-ABI: fast-wave is a standalone executable run once
 * per source point (fast-wave/main.cpp:18-57) and big-wave drives its Python
 * path through multisim.run_single_simulation (big-wave/multisim.py:50-150).
 * Each entry point below names the reference interface it stands in for.
 *
 * Error behaviour: every call returns RS_OK (0) or a positive status code and
 * never exits or throws across the ABI (fast-wave exit(1)s on CUDA errors,
 * fast-wave/fwcuda/types.hpp:54-67, and throws std::runtime_error on config
 * errors, fast-wave/src/config_parsing.cpp).  rs_last_error() returns the
 * message of the calling thread's last failure.
 *
 * Threading: one host thread at a time per rs_ctx; distinct contexts may be
 * used from distinct threads.  All work of a context runs on its own HIP
 * streams.  A server (rs_server_create) reserves its context until
 * rs_server_destroy: meanwhile every other call that uses the context, its
 * fields included, returns RS_ERR_INVALID (rs_ctx_synchronize excepted).
 */
#ifndef RAVESIM_H
#define RAVESIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define RS_OK 0
#define RS_ERR_INVALID 1  /* bad argument or handle */
#define RS_ERR_CONFIG 2   /* malformed or inconsistent sim-dir configuration */
#define RS_ERR_IO 3       /* file could not be read or written */
#define RS_ERR_DEVICE 4   /* HIP runtime error */
#define RS_ERR_NOMEM 5    /* device memory exhausted / admission refused */
#define RS_ERR_UNSUPPORTED 6

/* field precision: complex64 ("c8") or complex128 ("c16"), big-wave/config.py:171-188 */
#define RS_C8 8
#define RS_C16 16

/* execution form */
#define RS_MODE_FUSED 0 /* fused passes: slice = column pass + row pass */
#define RS_MODE_PLAIN 1 /* one kernel per operation, fast-wave's launch structure */

typedef struct rs_ctx rs_ctx;
typedef struct rs_field rs_field;

/* Library version (major*10000 + minor*100 + patch).  The minor number
 * changes with the layout of any struct below; a binding checks it before
 * passing structs. */
#define RS_VERSION 300
int rs_version(void);

/* Message of this thread's last failed call ("" if none). */
const char *rs_last_error(void);

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */

/* Create a context on HIP device `device`. */
int rs_ctx_create(int device, rs_ctx **out);
int rs_ctx_destroy(rs_ctx *ctx);
/* Block until all work queued on the context's stream has finished. */
int rs_ctx_synchronize(rs_ctx *ctx);

/* ------------------------------------------------------------------ */
/* sim-dir entry point                                                 */
/* ------------------------------------------------------------------ */

typedef struct rs_run_options {
  int dtype;         /* 0: take `dtype` from config.yaml (default c16, big-wave/config.py:176-177); RS_C8 / RS_C16 force it */
  int mode;          /* RS_MODE_FUSED or RS_MODE_PLAIN */
  int write_output;  /* 1: write %08d/detected.npy (float64 (n_steps, n_px)) and u_%04d.npy if requested */
  int reserved;
  double *detected;  /* optional caller buffer for the detector rows (n_steps*n_px float64), may be NULL */
  int64_t detected_capacity; /* elements available in `detected` */
  int64_t *shape_out;        /* optional int64[2] receiving (n_steps, n_px) */
  /* > 0: record big-wave's history with this free-space z spacing
   * (big-wave/history.py:43-135, multisim.py:130-141): intensity rows every
   * history_dz between elements (each re-propagated from the segment start,
   * so the detector rows do not change), after every grating slice and
   * before every sample row, analytic rows for a point source; written as
   * %08d/history.npy (float64 (n_px, entries)), history_z.npy, history_x.npy.
   * 0: none.  fast-wave's `--history_dz` writes the transpose. */
  double history_dz;
  /* non-NULL: write the field before and after every element of the first
   * pass as <dir>/keypoint_%02d_%d.npy in the field precision
   * (big-wave/wavesim.py:35-58) */
  const char *keypoints_dir;
} rs_run_options;

/* Run source `source_idx` of the simulation directory `sim_dir`.
 * Replaces `fastwave <sim_dir> -s <idx>` (fast-wave/main.cpp:18-57 ->
 * simulation.cpp:375-549) and multisim.run_single_simulation
 * (big-wave/multisim.py:50-150).  Follows big-wave's semantics where the two
 * references differ.  `opts` may be NULL (defaults: config dtype, fused,
 * write_output=1). */
int rs_run_simdir(rs_ctx *ctx, const char *sim_dir, int source_idx, const rs_run_options *opts);

/* Parse and validate a sim dir without touching a device (host only).
 * info receives {N, dtype (8|16), n_steps, n_px, n_elements, n_slices}
 * where n_slices counts the reference's propagate calls of one phase step.
 * Same configuration errors as rs_run_simdir. */
int rs_validate_simdir(const char *sim_dir, int source_idx, int64_t *info, int64_t n_info);

/* ------------------------------------------------------------------ */
/* field-level operators (fast-wave/fwcuda/wrappers.hpp:10-40,         */
/* big-wave/propagation.py, optical_element.py, source.py)             */
/* ------------------------------------------------------------------ */

typedef struct rs_grid {
  int64_t N;         /* samples; a power of two, 64 <= N <= 2^30 */
  double dx;         /* sample spacing [m] */
  int64_t chunk_size;/* big-wave chunk length (grating x is linspace'd per chunk) */
} rs_grid;

typedef struct rs_grating {
  double pitch;
  double dc0, dc1;          /* duty cycle at the first and after the last slice */
  double z_start;           /* start of the grating on the optical axis [m] */
  double thickness;         /* without substrate */
  int32_t nr_steps;
  int32_t nr_x_positions;
  const double *x_positions;
  double substrate_thickness;
  double db_a_re, db_a_im;  /* delta + i beta of material A (0 = vacuum) */
  double db_b_re, db_b_im;
  double db_s_re, db_s_im;  /* substrate */
} rs_grating;

/* Double-pitch (DBA) grating: material A where both
 * mod(x/pitch0 + dc0/2, 1) < dc0 and mod(x/pitch1, 1) < dc1 hold, each duty
 * cycle interpolated linearly over the slices (big-wave/optical_element.py
 * :227-338, 359-375; fast-wave/include/optical_element.hpp:50-63). */
typedef struct rs_dba_grating {
  double pitch0, pitch1;
  double dc0[2];            /* first pitch's duty cycle at the first and after the last slice */
  double dc1[2];            /* second pitch's duty cycle, likewise */
  double z_start;
  double thickness;         /* without substrate */
  int32_t nr_steps;
  int32_t nr_x_positions;
  const double *x_positions;
  double substrate_thickness;
  double db_a_re, db_a_im;
  double db_b_re, db_b_im;
  double db_s_re, db_s_im;
} rs_dba_grating;

/* Sample element: a grid of material indices, row r (z) and column k (x),
 * 0 = vacuum, i = deltabeta[i - 1] (big-wave/optical_element.py:380-430). */
typedef struct rs_sample {
  double z_start;
  double pixel_size_x, pixel_size_z;
  int64_t rows, cols;
  const int8_t *grid;        /* rows x cols, row-major */
  int32_t nr_materials;
  int32_t nr_x_positions;
  const double *deltabeta;   /* nr_materials x (delta, beta) */
  const double *x_positions;
} rs_sample;

typedef struct rs_detector {
  double size;       /* detector_size [m] */
  double pixel_x;    /* detector_pixel_size_x */
  double pixel_y;    /* detector_pixel_size_y */
} rs_detector;

/* Allocate a field of N samples of precision `dtype` (RS_C8 / RS_C16),
 * initialised to zero.  Mode selects fused or plain kernels for operators. */
int rs_field_create(rs_ctx *ctx, const rs_grid *grid, int dtype, int mode, rs_field **out);
int rs_field_destroy(rs_field *f);
/* Copy a natural-order complex field from / to the host.  `host` holds N
 * interleaved complex values of the field's precision. */
int rs_field_upload(rs_field *f, const void *host);
int rs_field_download(rs_field *f, void *host);

/* dst <- src (same N and precision): big-wave Vector.copy_to
 * (big-wave/vector.py:281-282), one copy kernel on the context stream. */
int rs_field_copy(rs_field *dst, const rs_field *src);

/* u = exp(-2 pi i r / wl) / sqrt(r), r = sqrt((x - x_source)^2 + z^2):
 * big-wave/propagation.py:141-151, fast-wave/fwcuda/kernels.cu:6-28. */
int rs_point_source(rs_field *f, double wl, double x_source, double z);

/* u <- IFFT(FFT(u) * exp(-2 pi i dz/wl) exp(i pi wl dz f^2) * [|f| < f_c]):
 * big-wave/propagation.py:100-138, fast-wave/src/simulation.cpp:34-40.
 * The cutoff is applied exactly as big-wave applies it: bins
 * [ceil(N f_c dx), N - ceil(N f_c dx) + 1) are zeroed unless
 * ceil(N f_c dx) > N/2; pass INFINITY to keep every frequency. */
int rs_propagate(rs_field *f, double wl, double dz, double cutoff_freq);

/* Grating slices plus substrate (big-wave/optical_element.py:134-206,
 * fast-wave/src/simulation.cpp:131-150), using x_positions[phase_step]. */
int rs_apply_grating(rs_field *f, double wl, const rs_grating *g, int phase_step,
                     double cutoff_freq);

/* DBA grating slices plus substrate (big-wave/optical_element.py:247-316,
 * fast-wave/src/simulation.cpp:152-172 `apply_dbagrating`), using
 * x_positions[phase_step]. */
int rs_apply_dba_grating(rs_field *f, double wl, const rs_dba_grating *g, int phase_step,
                         double cutoff_freq);

/* Sample rows: per grid row, u *= exp(2 pi i pixel_size_z / wl * db(x)) with
 * db the np.interp of delta + i beta between the pixel positions
 * k px + x_positions[phase_step] - cols px / 2 (zero outside), then a
 * propagation over pixel_size_z (big-wave/optical_element.py:432-476;
 * fast-wave/fwcuda/kernels.cu:118-150 `apply_sample_factors`). */
int rs_apply_sample(rs_field *f, double wl, const rs_sample *s, int phase_step, double cutoff_freq);

/* u *= factor (fast-wave/fwcuda/kernels.cu:160-172). */
int rs_scale(rs_field *f, double re, double im);

/* Detector intensity at distance z into out[n_px], n_px = floor(size/pixel_x):
 * big-wave/propagation.py:279-320 (np.histogram binning, FP64 sums). */
int rs_square_and_downsample(rs_field *f, const rs_detector *det, double z, double *out,
                             int64_t n_out);

/* Allocation-free variant of rs_square_and_downsample, with the same
 * results bit for bit.  It has no counterpart in the reference; it is for
 * callers that cannot allocate device memory (serving loops and harnesses
 * that forbid hipMalloc): it calls no hipMalloc / hipFree and allocates no
 * host memory.  The pixel bounds and FP64 sums go to a fixed scratch array of
 * 1 MiB per GPU (held until the process ends, its address looked up when a
 * context is created), so n_px may be at most RS_NOALLOC_MAX_PIXELS (larger
 * detectors: RS_ERR_INVALID).  Every call on one GPU uses that GPU's array:
 * calls that overlap on one GPU, from different contexts or threads, race. */
#define RS_NOALLOC_MAX_PIXELS 65536
int rs_square_and_downsample_noalloc(rs_field *f, const rs_detector *det, double z, double *out,
                                     int64_t n_out);

/* Number of detector pixels for `det` (int(size // pixel_x)). */
int64_t rs_detector_pixels(const rs_detector *det);

/* Sample-index boundaries of the detector pixels (host-only helper, no
 * device work): bounds[k]..bounds[k+1] are the samples binned into pixel k.
 * bounds has n_px + 1 entries. */
int rs_detector_bounds(int64_t N, double dx, const rs_detector *det, int64_t *bounds,
                       int64_t n_bounds);

/* ------------------------------------------------------------------ */
/* batched runner (many energies / sources / sims in lockstep)        */
/* ------------------------------------------------------------------ */

/* One (energy, source) simulation of a batch.  All members of one batch call
 * share N, dtype and the element structure; they differ in wavelength,
 * source position and grating parameters. */
typedef struct rs_batch_member {
  double wl;               /* wavelength [m] */
  double weight;           /* weight of its detector rows in the accumulated sum */
  int32_t source_kind;     /* 0 = point source, 1 = plane wave (vector source of ones) */
  int32_t out_index;       /* which accumulator row receives this member */
  double source_x, source_z;
  const rs_grating *gratings; /* nr_gratings entries */
  const double *cutoff_angles; /* nr_gratings + 1 entries */
} rs_batch_member;

typedef struct rs_batch_desc {
  rs_grid grid;
  rs_detector det;
  double z_detector;
  int32_t dtype;
  int32_t mode;
  int32_t nr_gratings;
  int32_t nr_members;
  const rs_batch_member *members;
  int32_t nr_outputs;      /* accumulator rows */
  int32_t max_resident;    /* members resident at once, one launch per pass covering them; 0 = as many
                              as fit the 256 MiB last-level cache together (at least one), bounded by memory */
  int32_t use_graph;       /* capture the whole run, every lane, as one hipGraph (lanes fork from and
                              join the context stream through events), instantiate and launch it */
  int32_t lanes;           /* streams that groups alternate over (1..4); 0 = 2 with the default
                              grouping, 1 otherwise */
  int32_t spans;           /* three-group plans: run each propagation's middle, row and middle passes
                              over this many consecutive sub-ranges of the fields (a power of two, at
                              most the top group's length); 0 = the fewest whose share of a group's
                              fields fits the 256 MiB last-level cache */
} rs_batch_desc;

/* Run every member; out[nr_outputs * n_px] receives sum over members of
 * weight * detector row (FP64).  Only single-phase-step setups. */
int rs_batch_run(rs_ctx *ctx, const rs_batch_desc *desc, double *out, int64_t n_out);

/* One simulation of a mixed batch: its own grid, detector and gratings; its
 * members (energies, source points) add weight * detector row into its own
 * output row (their out_index is ignored). */
typedef struct rs_batch_sim {
  rs_grid grid;
  rs_detector det;
  double z_detector;
  int32_t nr_gratings;
  int32_t nr_members;
  const rs_batch_member *members;
  int64_t out_offset;      /* first element of its n_px values in `out` */
} rs_batch_sim;

typedef struct rs_mixed_desc {
  int32_t dtype;
  int32_t mode;
  int32_t nr_sims;
  int32_t use_graph;       /* as rs_batch_desc.use_graph, over the whole mixed run */
  const rs_batch_sim *sims;
} rs_mixed_desc;

/* Simulations of different grid sizes in one call (the reference runs one
 * fastwave process per source point, fast-wave/Readme.md:21-34): those that
 * share N, dx, chunk, detector and element count form one lockstep part, and
 * the parts' launches interleave over the context's two lanes between one
 * argument upload and one synchronisation; every simulation's row is what
 * rs_batch_run gives it.  Single-phase-step setups, one precision. */
int rs_batch_run_mixed(rs_ctx *ctx, const rs_mixed_desc *desc, double *out, int64_t n_out);

/* ------------------------------------------------------------------ */
/* serving (no counterpart in the reference, which runs one process per */
/* source point: fast-wave/Readme.md:21-34)                             */
/* ------------------------------------------------------------------ */

typedef struct rs_server rs_server;

/* One request: a setup at nr_energies energies (summed incoherently with
 * weights); gratings may carry several x_positions (phase stepping), giving
 * n_steps detector rows.  gratings holds nr_energies * nr_gratings entries
 * (delta/beta depend on the energy), energy-major. */
typedef struct rs_request {
  rs_grid grid;
  rs_detector det;
  double z_detector;
  int32_t dtype;            /* RS_C8 or RS_C16 */
  int32_t nr_gratings;
  int32_t nr_energies;
  int32_t source_kind;      /* 0 point source, 1 plane wave */
  const double *energies_ev;
  const double *weights;    /* NULL = all 1 */
  const rs_grating *gratings;
  const double *cutoff_angles;  /* nr_gratings + 1 */
  double source_x, source_z;
} rs_request;

/* Start a scheduler thread on the context's device; the context is reserved
 * for it until rs_server_destroy.  Jobs ((request, energy) pairs) are
 * admitted in arrival order while their buffers fit in the free device
 * memory, read at each admission, less reserve_bytes (<= 0: 4 GiB); a job
 * that cannot fit with nothing else resident fails its request (poll reports
 * state 3 and returns RS_ERR_DEVICE with an RS_ERR_NOMEM message).  max_jobs
 * bounds the number of jobs resident at once (<= 0: no bound).  Grid shapes
 * are spread over up to four streams; the stream of the smallest grid with
 * work goes first, and streams of larger grids get one tick a second while
 * it has work, so short requests are not held up by long ones. */
int rs_server_create(rs_ctx *ctx, double reserve_bytes, int max_jobs, rs_server **out);
/* Queue a request (all inputs copied); returns immediately. */
int rs_server_submit(rs_server *srv, const rs_request *req, int64_t *id);
/* state: 0 queued, 1 running, 2 done, 3 failed.  When done, out receives the
 * (n_steps, n_px) rows if out_cap suffices; shape2 = {n_steps, n_px};
 * times3 = {queue wait, service time, latency} in seconds. */
int rs_server_poll(rs_server *srv, int64_t id, int32_t *state, double *out, int64_t out_cap,
                   int64_t *shape2, double *times3);
/* Block until request id is done or timeout_s passes (1 = done, 0 = timeout). */
int rs_server_wait(rs_server *srv, int64_t id, double timeout_s, int32_t *done);
/* stats6 = {ticks, launches, jobs per launch (mean), jobs done, max resident jobs, busy s} */
int rs_server_stats(rs_server *srv, double *stats6);
/* Finish all queued work and stop the scheduler. */
int rs_server_destroy(rs_server *srv);

/* ------------------------------------------------------------------ */
/* measurement (no counterpart in the reference; fast-wave logs        */
/* per-element cudaEvent timings, fast-wave/src/simulation.cpp:13-32)  */
/* ------------------------------------------------------------------ */

/* Per kernel class: 0 row pass, 1 top column pass, 2 middle column pass,
 * 3 other (plain-form elementwise kernels). */
typedef struct rs_pass_stat {
  int64_t launches;
  double total_ms;   /* summed HIP-event durations on the context stream */
  double bytes;      /* summed algorithmic bytes (every field read + written once) */
} rs_pass_stat;

/* Autotuner (no counterpart in the reference): time every valid digit-group
 * split of an N-point field on a synthetic grating slice loop run as a batch
 * of `members` fields (<= 0: 1) runs in rs_batch_run by default (grouping and
 * lanes), by wall clock, and keep the fastest for this context and that batch
 * size; runs of other batch sizes use the entry tuned for the nearest one.
 * out4 (optional) = {groups, lg top, lg middle, lg row}; ms_per_slice
 * (optional) = measured time of one slice of one field with it (the
 * difference of a 4-slice and a 16-slice run, per slice and field, so that
 * what a run pays once is left out). */
int rs_autotune(rs_ctx *ctx, int64_t N, int dtype, int members, int32_t *out4, double *ms_per_slice);
/* Fix the digit-group split for (N, dtype) (lgR = log2 row length, lgA = log2
 * top column length, -1 = default); persisted tuning results are re-applied
 * this way. */
int rs_set_split(rs_ctx *ctx, int64_t N, int dtype, int lgR, int lgA);

/* One autotuner result: the digit-group split and the column tile variants
 * (0 = the default width, 1 = the alternative: twice as many columns per
 * workgroup for 128- and 512-long columns, half as many otherwise) chosen
 * for (N, dtype) at a batch size of `members` fields (0: every batch size). */
typedef struct rs_tune_entry {
  int64_t N;
  int32_t dtype;
  int32_t members;
  int32_t lgR, lgA;
  int32_t tileA, tileB;  /* top and middle column passes */
  double ms_per_slice;
} rs_tune_entry;

/* The entry runs of `members` fields would use (the nearest tuned batch size
 * in log2, or a fixed split); *found = 0 when there is none. */
int rs_tune_get(rs_ctx *ctx, int64_t N, int dtype, int members, rs_tune_entry *out, int32_t *found);
/* Persist the context's tuning entries as text (one line per entry:
 * device lgN c8|c16 members lgR lgA tileA tileB ms_per_slice, where device
 * is the architecture name and compute-unit count), or read such a file:
 * entries of this context's device are added (same key replaced), *n_loaded
 * receives their count.  A malformed or invalid entry is RS_ERR_CONFIG. */
int rs_tune_save(rs_ctx *ctx, const char *path);
int rs_tune_load(rs_ctx *ctx, const char *path, int32_t *n_loaded);

/* Per-launch event timing: 0 = off, 1 = every launch, k > 1 = every k-th
 * launch of each kernel class (a sample; each timed launch costs a few
 * microseconds).  Resets the counters. */
int rs_set_timing(rs_ctx *ctx, int enable);
/* Copy min(n, 4) class records. */
int rs_get_pass_stats(rs_ctx *ctx, rs_pass_stat *out, int n);

/* Streaming-copy bandwidth of this device (the achievable figure beside a
 * roofline): `reps` copies of `bytes` between two fresh device buffers with
 * the library's copy kernel, timed with HIP events on the context stream;
 * *gbs = 2 * bytes * reps / time (read + write) in GB/s. */
int rs_copy_bandwidth(rs_ctx *ctx, int64_t bytes, int reps, double *gbs);

#ifdef __cplusplus
}
#endif

#endif /* RAVESIM_H */
