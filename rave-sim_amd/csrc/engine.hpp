// Host engine: FFT plans, member programs and the lockstep executor.
//
// A *member* is one (source point, energy) simulation, the unit fast-wave runs
// per process (fast-wave/src/simulation.cpp:375-549).  Its work is a list of
// Ops mirroring big-wave's driver (big-wave/wavesim.py:84-298): generate or
// upload the source, x-space multiplies (grating slices, substrates),
// propagations, detector reads, snapshot / restore for phase stepping.
// Members whose op lists have the same shape run in lockstep: every fused
// pass is one launch over all of them (blockIdx.y = member).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rs_device.hpp"

namespace rs {

// MI355X last-level cache (AMD Infinity Cache / MALL) per device: default
// lockstep groups keep their fields within it
constexpr size_t kMallBytes = size_t(256) << 20;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what);

// ----------------------------------------------------------------------------
// exact host helpers (big-wave arithmetic)
// ----------------------------------------------------------------------------

// [rs, re) zeroed by the cutoff, big-wave/propagation.py:83-90; (0,0) if none
void cutoff_range(int64_t N, double fc, double dx, int64_t *rs, int64_t *re);

// Python's float floor division vx // wx
double py_floordiv(double vx, double wx);

// number of detector pixels, int(size // px) (big-wave/propagation.py:286)
int64_t detector_pixels(double size, double px);

// Sample boundaries per pixel reproducing np.histogram's uniform-bin
// assignment of x_j = (j - N/2) dx over [-(n/2) px, (n/2 + 1) px) with n+1
// bins (big-wave/propagation.py:293-308).  bounds has n_px + 1 entries.
std::vector<int64_t> detector_bounds(int64_t N, double dx, double size, double px);

// cos(atan(x/z)) and sqrt(x^2 + z^2) of pixel k, x = (k - n//2) px
void detector_pixel_scale(double size, double px, double z, int64_t k, double *cos_a, double *dist);

// final per-pixel factor cos(atan(x/z)) dx py / sqrt(x^2 + z^2),
// x = (k - n//2) px (big-wave/propagation.py:312-320, util.py:78-82)
std::vector<double> detector_scale(double size, double px, double py, double dx, double z);

// exp(2 pi i t / wl * db), big-wave/optical_element.py:36-39
void material_factor(double db_re, double db_im, double t, double wl, double out[2]);

double energy_to_wavelength(double e);

// ----------------------------------------------------------------------------
// FFT plan
// ----------------------------------------------------------------------------

struct Plan {
  int64_t N = 0;
  int lgN = 0;
  bool f64 = false;
  int ngroups = 0;
  int lgA = 0, lgB = 0, lgR = 0;  // top, middle, row (lgA = lgR when ngroups == 1)
  int tileA = 0, tileB = 0;       // column tile variants of the top and middle passes (col_fibers)
  // device tables (owned)
  void *fib[3] = {nullptr, nullptr, nullptr};  // per pass: W_L^j, then small-stage tables
  void *hi[2] = {nullptr, nullptr}, *lo[2] = {nullptr, nullptr};
  int sh[2] = {0, 0};
  void *cis = nullptr;  // complex64 plans: exp(2 pi i j / 256), j < 256
  Plan() = default;
  Plan(const Plan &) = delete;
  Plan &operator=(const Plan &) = delete;
  ~Plan();
  Geom geom_row() const;
  Geom geom_top() const;
  Geom geom_mid() const;
  std::string describe() const;
};

// split of lgN into digit groups; lgR_req / lgA_req >= 0 override the default
// (row as long as the kernels allow, remaining bits halved between columns)
void plan_split(int lgN, bool f64, int lgR_req, int lgA_req, int *ng, int *lgA, int *lgB, int *lgR);
std::unique_ptr<Plan> make_plan(int64_t N, bool f64, int lgR_req = -1, int lgA_req = -1, int tileA = 0,
                                int tileB = 0);
// whether the column tiles (tileA, tileB) fit the split lgR | lgA of lgN (the
// tile width divides the pass's stride)
bool plan_tiles_valid(int lgN, bool f64, int lgR, int lgA, int tileA, int tileB);
// every valid (lgR, lgA) split of lgN, the autotuner's candidates
std::vector<std::pair<int, int>> plan_candidates(int lgN, bool f64);

// ----------------------------------------------------------------------------
// member programs
// ----------------------------------------------------------------------------

// Snapshot / Restore: the phase-stepping restart point.  HSave / HRestore: a
// second saved state for history entries (big-wave/history.py:74-135
// re-propagates every entry from the segment's start).  SaveU / Keypoint:
// natural-order field downloads (u_%04d.npy, keypoint_%02d_%d.npy).
enum class OpKind { Source, Upload, Mask, Prop, Detect, Snapshot, Restore, SaveU, Flush, HSave, HRestore, Keypoint };

struct Op {
  OpKind kind;
  XOp x{};          // Source (POINT / ONES) and Mask (MASK / SCALAR / SAMPLE)
  double dz = 0;    // Prop
  double fc = 0;    // Prop cutoff frequency (big-wave semantics)
  int row = 0;      // Detect: output row; SaveU: phase step; Keypoint: 2 element + (0 before, 1 after)
  int sample = -1;  // SAMPLE masks: index into Member::samples and the grid row
  int sample_row = 0;
};

// Host arrays of one sample element at one phase step: pixel positions, and
// per grid row the pixels' delta + i beta and exp(i a db) (interleaved)
struct SampleRows {
  std::vector<double> rx;
  std::vector<double> db, f;  // rows x (2 len)
  int64_t len = 0;
};

struct Member {
  double wl = 0;
  double weight = 1.0;
  int out_base = 0;             // accumulator row of Detect(row = 0)
  std::vector<Op> ops;
  std::vector<SampleRows> samples;  // device copies are made per run
  // Upload op: a .npy vector (natural order) streamed from disk through the
  // context's pinned staging buffers (no host copy of the whole field)
  std::string upload_npy;       // path
  int64_t upload_offset = 0;    // byte offset of the data in the file
  bool upload_c16 = false;      // file elements are complex128 (else complex64)
  // .npy files the SaveU / Keypoint ops stream the field into (path of a
  // phase step / keypoint row), through the context's staging buffers; the
  // ops do nothing when unset
  std::function<std::string(int)> saveu_file, keypoint_file;
};

struct RunConfig {
  int64_t N = 0;
  double dx = 0;
  bool f64 = false;
  int mode = 0;              // 0 fused, 1 plain
  double det_size = 0, det_px = 0, det_py = 1, z_det = 0;
  int n_out = 1;             // accumulator rows
  std::vector<double> row_z; // detector distance of each row (empty: z_det for all)
  int max_resident = 0;      // members per group; 0 = as many as fit the last-level cache together
  // capture the whole run, every lane, as one hipGraph (lanes fork from and
  // join the context stream through events inside the capture)
  bool use_graph = false;
  bool whole_graph = false;  // set with use_graph by the runners
  int lanes = 0;             // streams groups alternate over; 0 = 2 with the default grouping
  int spans = 0;             // sub-ranges of the middle-row-middle passes (3 groups); 0 = by cache size
  int lgR = -1;              // row length override (autotuner), -1 = tuned or default
  int lgA = -1;              // top column length override
  int tileA = -1, tileB = -1;  // column tile variants with an override (-1 = default)
  void *const *fields = nullptr;  // caller-owned field buffers (one per member), else allocated
};

class Context;
struct GraphCache;

// Run all members; returns scaled detector rows, n_out x n_px (row-major).
std::vector<double> run_members(Context &ctx, const RunConfig &rc, std::vector<Member> &members);

// One lockstep part of a mixed run: members that share N, the detector and
// the element structure, with their run configuration.
struct MixedPart {
  RunConfig rc;
  std::vector<Member> *members = nullptr;
};
// Run parts of different N (one precision) together: one argument upload,
// the groups of every part interleaved over the context's lanes, one
// synchronisation; whole_graph captures all of it as one hipGraph.  Returns
// each part's scaled rows.
std::vector<std::vector<double>> run_mixed(Context &ctx, const std::vector<MixedPart> &parts, bool whole_graph);

// ----------------------------------------------------------------------------
// context: device + stream + plan cache
// ----------------------------------------------------------------------------

// per-kernel-class timing collected with HIP events on the context stream
struct PassStat {
  int64_t launches = 0;
  double ms = 0;      // summed kernel durations
  double bytes = 0;   // summed algorithmic bytes (fields read + written)
};
enum PassClass { PC_ROW = 0, PC_TOP = 1, PC_MID = 2, PC_OTHER = 3, PC_COUNT = 4 };

class Context {
 public:
  explicit Context(int device);
  ~Context();
  int device;
  hipStream_t stream = nullptr;   // all work of the context's calls is ordered on it
  // this device's scratch memory of the allocation-free detector read
  // (noalloc_scratch, looked up when the context is created)
  void *noalloc_scr = nullptr;
  // lane i of the lockstep runner (0 = `stream`; others created on first use
  // and joined into `stream` at the end of each run)
  hipStream_t lane_stream(int i);
  bool timing = false;
  int timing_stride = 1;              // time every timing_stride-th launch of a class
  int64_t timing_seen[PC_COUNT] = {};  // launches of each class since timing was set
  PassStat stats[PC_COUNT];
  // plan for (N, precision); explicit lgR/lgA (with tiles tileA/tileB, -1 =
  // default), else the split and tiles tuned for the nearest batch size
  // (fields per run), else the default split
  Plan &plan(int64_t N, bool f64, int lgR = -1, int lgA = -1, int members = 1, int tileA = -1, int tileB = -1);
  void sync();
  // autotuner results: (lgN, f64, fields per run) -> (lgR, lgA, ms per slice
  // and field); members == 0 marks a fixed split for every batch size
  struct Tuned {
    int lgN;
    bool f64;
    int lgR, lgA;
    double ms;
    int members;
    int tileA = 0, tileB = 0;
  };
  // identity of the device for persisted tuning tables: architecture name
  // and compute-unit count (e.g. "gfx950:sramecc+:xnack-/256")
  std::string device_key;
  std::vector<Tuned> tuned;
  const Tuned *find_tuned(int lgN, bool f64, int members = 1) const;
  // the instantiated graph of the last captured batch run and the device
  // buffers its launches address: a run of the same shape replays it with
  // new argument records instead of capturing again
  std::unique_ptr<GraphCache> graph_cache;
  // Copy N natural-order samples of a .npy vector (member m's upload_npy)
  // to the device field dst on stream st.  Up to kStageThreads host threads
  // each stream one contiguous part: chunks are read into the thread's two
  // pinned staging buffers in turn, each reused once its previous copy has
  // left it, so file reads overlap the transfers and each other.
  void stream_npy(const Member &m, bool f64, int64_t N, void *dst, hipStream_t st);
  // The reverse: N natural-order samples of the device field src, after the
  // work already on stream st, into a new .npy file at path (complex64 or
  // complex128 as f64), each thread downloading chunks of its part into its
  // two staging buffers in turn and writing one while the next arrives.
  // Returns when the file is complete.
  void stream_to_npy(const void *src, int64_t N, bool f64, const std::string &path, hipStream_t st);
  // Device-to-host copy of `bytes` after the work already on stream st,
  // through the staging buffers (pageable destinations included); returns
  // when host holds the data.
  void copy_to_host(void *host, const void *src, size_t bytes, hipStream_t st);
  // Host-to-device copy of `bytes` through the staging buffers, ordered on
  // stream st; host may be reused when it returns.
  void copy_from_host(void *dst, const void *host, size_t bytes, hipStream_t st);

 private:
  static constexpr size_t kStageBytes = size_t(4) << 20;
  static constexpr int kStageThreads = 4;
  struct Stage {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
  };
  Stage stage_[kStageThreads];
  // staging threads for a field of `bytes`, with their buffers allocated
  int stage_threads(int64_t bytes);
  // the download loop of stream_to_npy / copy_to_host: each thread takes a
  // contiguous part of the `bytes` at src, downloads it chunk by chunk into
  // its two staging buffers in turn and hands each chunk to sink(data,
  // byte offset, length) while the next one arrives
  void stream_out(const void *src, size_t bytes, hipStream_t st,
                  const std::function<void(const char *, size_t, size_t)> &sink);
  // the upload loop of stream_npy / copy_from_host: each thread takes a
  // contiguous part of the `count` elements of `esz` bytes, has fill(buffer,
  // first element, count) put each chunk into its staging buffers in turn and
  // copies it to dst while it fills the next
  void stream_in(void *dst, int64_t count, size_t esz, hipStream_t st,
                 const std::function<void(char *, int64_t, int64_t)> &fill);
  std::vector<hipStream_t> lanes_;
  std::vector<std::unique_ptr<Plan>> plans_;
  std::vector<std::array<int, 4>> plan_req_;  // lgR, lgA, tileA, tileB as requested
};

// Time every candidate split of N, with each column tile variant of its
// column passes, on a synthetic grating slice loop (mask + propagation) run
// as a batch of `members` fields runs by default (grouping and lanes of
// run_members), by wall clock, and record the fastest in ctx.tuned for that
// batch size.  The figure is the marginal cost of a slice: the difference of
// runs of reps + 1 and 4 (reps + 1) slices, per slice and field.  Returns it.
Context::Tuned autotune(Context &ctx, int64_t N, bool f64, int reps = 3, int members = 1);

// Tuning tables as text, one entry per line:
//   device lgN c8|c16 members lgR lgA tileA tileB ms_per_slice
// (members 0: a fixed split for every batch size).  save writes every entry
// of the context; load adds the entries recorded for this context's device
// (replacing entries with the same key) and returns how many it took.
void tune_save(const Context &ctx, const std::string &path);
int tune_load(Context &ctx, const std::string &path);

// device buffer RAII
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b);
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept;
  ~DevBuf();
};

// Device buffers of one lockstep run: pixel bounds, the per-lane FP64
// accumulators, and the field / snapshot / history buffers of one group per
// lane.
struct RunBuffers {
  DevBuf bounds;
  std::vector<DevBuf> acc;
  std::vector<std::vector<DevBuf>> field, snap, hist;
};

}  // namespace rs
