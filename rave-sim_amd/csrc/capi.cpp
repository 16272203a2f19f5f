// C-ABI of librsim.so (declarations and reference citations in include/ravesim.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ravesim.h"
#include "engine.hpp"
#include "npy.hpp"
#include "rs_kernels.hpp"
#include "simulation.hpp"
#include "serve.hpp"

static_assert(RS_NOALLOC_MAX_PIXELS == rs::kNoAllocMaxPixels, "allocation-free detector limit");

struct rs_ctx {
  std::unique_ptr<rs::Context> ctx;
  // servers attached to the context: while one lives, its scheduler thread
  // owns the context's streams and plan cache, and the other entry points
  // that use the context refuse to run
  std::atomic<int> servers{0};
};

struct rs_server {
  std::unique_ptr<rs::Server> srv;
  rs_ctx *owner = nullptr;
};

struct rs_field {
  rs_ctx *owner = nullptr;
  rs_grid grid{};
  bool f64 = false;
  int mode = 0;
  rs::DevBuf buf;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// the context is reserved by a live server (rs_server_create .. rs_server_destroy)
bool reserved(const rs_ctx *c) { return c && c->servers.load() > 0; }
int fail_reserved() {
  return fail(RS_ERR_INVALID, "the context is reserved by a server until rs_server_destroy");
}

template <typename F> int guarded(F &&f) {
  try {
    f();
    return RS_OK;
  } catch (const rs::Error &e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc &) {
    return fail(RS_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception &e) {
    return fail(RS_ERR_INVALID, e.what());
  }
}

rs::RunConfig field_rc(const rs_field *f) {
  rs::RunConfig rc;
  rc.N = f->grid.N;
  rc.dx = f->grid.dx;
  rc.f64 = f->f64;
  rc.mode = f->mode;
  rc.n_out = 1;
  return rc;
}

void run_on_field(rs_field *f, rs::Member &m, rs::RunConfig rc, double *out = nullptr, int64_t n_out = 0) {
  void *fields[1] = {f->buf.p};
  rc.fields = fields;
  std::vector<rs::Member> ms;
  ms.push_back(std::move(m));
  std::vector<double> det = rs::run_members(*f->owner->ctx, rc, ms);
  if (out) {
    if ((int64_t)det.size() > n_out) throw rs::Error(RS_ERR_INVALID, "output buffer too small");
    std::memcpy(out, det.data(), det.size() * sizeof(double));
  }
}

rs::SimConfig field_cfg(const rs_field *f, double wl) {
  rs::SimConfig c;
  c.N = f->grid.N;
  c.dx = f->grid.dx;
  c.chunk_size = f->grid.chunk_size > 0 ? f->grid.chunk_size : f->grid.N;
  c.f64 = f->f64;
  c.wl = wl;
  return c;
}

rs::GratingCfg grating_from(const rs_grating *g) {
  rs::GratingCfg c;
  c.pitch = g->pitch;
  c.dc0 = g->dc0;
  c.dc1 = g->dc1;
  c.z_start = g->z_start;
  c.thickness = g->thickness;
  c.nr_steps = g->nr_steps;
  c.x_positions.assign(g->x_positions, g->x_positions + g->nr_x_positions);
  c.substrate_thickness = g->substrate_thickness;
  c.db_a = {g->db_a_re, g->db_a_im};
  c.db_b = {g->db_b_re, g->db_b_im};
  c.db_s = {g->db_s_re, g->db_s_im};
  return c;
}

rs::GratingCfg dba_grating_from(const rs_dba_grating *g) {
  rs::GratingCfg c;
  c.dba = true;
  c.pitch = g->pitch0;
  c.pitch1 = g->pitch1;
  c.dc0 = g->dc0[0];
  c.dc1 = g->dc0[1];
  c.dcb0 = g->dc1[0];
  c.dcb1 = g->dc1[1];
  c.z_start = g->z_start;
  c.thickness = g->thickness;
  c.nr_steps = g->nr_steps;
  c.x_positions.assign(g->x_positions, g->x_positions + g->nr_x_positions);
  c.substrate_thickness = g->substrate_thickness;
  c.db_a = {g->db_a_re, g->db_a_im};
  c.db_b = {g->db_b_re, g->db_b_im};
  c.db_s = {g->db_s_re, g->db_s_im};
  return c;
}

}  // namespace

extern "C" {

int rs_version(void) { return RS_VERSION; }

const char *rs_last_error(void) { return g_err.c_str(); }

int rs_ctx_create(int device, rs_ctx **out) {
  if (!out) return fail(RS_ERR_INVALID, "out is NULL");
  return guarded([&] {
    int n = 0;
    rs::hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) throw rs::Error(RS_ERR_INVALID, "no HIP device " + std::to_string(device));
    auto c = std::make_unique<rs_ctx>();
    c->ctx = std::make_unique<rs::Context>(device);
    *out = c.release();
  });
}

int rs_ctx_destroy(rs_ctx *ctx) {
  if (!ctx) return fail(RS_ERR_INVALID, "ctx is NULL");
  if (reserved(ctx)) return fail_reserved();
  return guarded([&] { delete ctx; });
}

int rs_ctx_synchronize(rs_ctx *ctx) {
  if (!ctx) return fail(RS_ERR_INVALID, "ctx is NULL");
  return guarded([&] { ctx->ctx->sync(); });
}

int rs_run_simdir(rs_ctx *ctx, const char *sim_dir, int source_idx, const rs_run_options *opts) {
  if (!ctx || !sim_dir) return fail(RS_ERR_INVALID, "ctx or sim_dir is NULL");
  if (reserved(ctx)) return fail_reserved();
  rs_run_options o{};
  o.write_output = 1;
  if (opts) o = *opts;
  return guarded([&] {
    rs::SimConfig c = rs::parse_simdir(sim_dir, source_idx);
    if (o.dtype == RS_C8) c.f64 = false;
    else if (o.dtype == RS_C16) c.f64 = true;
    else if (o.dtype != 0) throw rs::Error(RS_ERR_INVALID, "dtype must be 0, RS_C8 or RS_C16");
    const std::string sub = std::string(sim_dir) + "/" + rs::subdir_name(source_idx);
    rs::Member m;
    const bool save_u = c.save_final_u_vectors && o.write_output;
    // u vectors and keypoints stream from the device into their files
    if (save_u)
      m.saveu_file = [&](int s) {
        char name[32];
        std::snprintf(name, sizeof name, "/u_%04d.npy", s);
        return sub + name;
      };
    if (c.source.kind == rs::SourceCfg::Vector) {
      // checked here, streamed to the device in the run (Context::stream_npy)
      const rs::NpyHeader h = rs::npy_header(c.source.input_path);
      if (h.shape.size() != 1 || h.shape[0] != c.N)
        throw rs::Error(RS_ERR_CONFIG, "source vector must have shape (N,)");
      if (h.descr != "<c16" && h.descr != "<c8")
        throw rs::Error(RS_ERR_CONFIG, "source vector must be complex64 or complex128");
      const int64_t esz = h.descr == "<c16" ? 16 : 8;
      if (h.file_bytes < h.data_offset + c.N * esz) throw rs::Error(3, c.source.input_path + ": truncated data");
      m.upload_npy = c.source.input_path;
      m.upload_offset = h.data_offset;
      m.upload_c16 = h.descr == "<c16";
    }
    if (!(o.history_dz >= 0)) throw rs::Error(RS_ERR_INVALID, "history_dz must be >= 0");
    rs::Recording rec;
    rec.history_dz = o.history_dz;
    rec.keypoints = o.keypoints_dir != nullptr;
    if (rec.keypoints)
      m.keypoint_file = [&](int k) {
        char name[48];
        std::snprintf(name, sizeof name, "/keypoint_%02d_%d.npy", k / 2, k % 2);
        return std::string(o.keypoints_dir) + name;
      };
    const int nsteps = rs::build_member(c, m, 0, save_u, &rec);
    rs::RunConfig rc;
    rc.N = c.N;
    rc.dx = c.dx;
    rc.f64 = c.f64;
    rc.mode = o.mode;
    rc.det_size = c.detector_size;
    rc.det_px = c.px;
    rc.det_py = c.py;
    rc.z_det = c.z_detector;
    rc.n_out = rec.n_rows;
    if (rec.n_rows > nsteps) {
      // history rows are scaled for their own z
      rc.row_z.assign((size_t)rec.n_rows, c.z_detector);
      for (size_t k = 0; k < rec.hist_z.size(); ++k)
        if (rec.hist_row[k] >= 0) rc.row_z[(size_t)rec.hist_row[k]] = rec.hist_z[k];
    }
    std::vector<rs::Member> ms;
    ms.push_back(std::move(m));
    std::vector<double> det = rs::run_members(*ctx->ctx, rc, ms);
    const int64_t npx = rs::detector_pixels(c.detector_size, c.px);
    const size_t ndet = (size_t)nsteps * (size_t)std::max<int64_t>(npx, 0);
    if (o.shape_out) {
      o.shape_out[0] = nsteps;
      o.shape_out[1] = npx;
    }
    if (o.detected) {
      if (o.detected_capacity < (int64_t)ndet) throw rs::Error(RS_ERR_INVALID, "detected buffer too small");
      std::memcpy(o.detected, det.data(), ndet * sizeof(double));
    }
    if (o.write_output) {
      rs::npy_save(sub + "/detected.npy", "<f8", {nsteps, npx}, det.data(), ndet * sizeof(double));
    }
    if (o.write_output && rec.history_dz > 0) {
      // np.stack(entries, axis=-1): (n_px, entries); np.empty((0, 0)) when none
      const int64_t H = (int64_t)rec.hist_z.size();
      std::vector<double> hm((size_t)(H * npx));
      for (int64_t k = 0; k < H; ++k) {
        const int r = rec.hist_row[(size_t)k];
        const double *src = r >= 0 ? det.data() + (size_t)r * npx : rec.analytic[(size_t)k].data();
        for (int64_t p = 0; p < npx; ++p) hm[(size_t)(p * H + k)] = src[p];
      }
      const std::vector<int64_t> shape = H ? std::vector<int64_t>{npx, H} : std::vector<int64_t>{0, 0};
      rs::npy_save(sub + "/history.npy", "<f8", shape, hm.data(), hm.size() * sizeof(double));
      rs::npy_save(sub + "/history_z.npy", "<f8", {H}, rec.hist_z.data(), rec.hist_z.size() * sizeof(double));
      const std::vector<double> hx = rs::detector_x(c.detector_size, c.px);
      rs::npy_save(sub + "/history_x.npy", "<f8", {(int64_t)hx.size()}, hx.data(), hx.size() * sizeof(double));
    }
  });
}

int rs_validate_simdir(const char *sim_dir, int source_idx, int64_t *info, int64_t n_info) {
  if (!sim_dir) return fail(RS_ERR_INVALID, "sim_dir is NULL");
  return guarded([&] {
    rs::SimConfig c = rs::parse_simdir(sim_dir, source_idx);
    rs::Member m;
    const int nsteps = rs::build_member(c, m, 0, false);
    int64_t props = 0;
    for (const auto &op : m.ops) {
      if (op.kind == rs::OpKind::Detect) break;
      props += op.kind == rs::OpKind::Prop;
    }
    const int64_t v[6] = {c.N, c.f64 ? 16 : 8, nsteps, rs::detector_pixels(c.detector_size, c.px),
                          (int64_t)c.elements.size(), props};
    if (info)
      for (int64_t i = 0; i < n_info && i < 6; ++i) info[i] = v[i];
  });
}

int rs_field_create(rs_ctx *ctx, const rs_grid *grid, int dtype, int mode, rs_field **out) {
  if (!ctx || !grid || !out) return fail(RS_ERR_INVALID, "NULL argument");
  if (reserved(ctx)) return fail_reserved();
  if (dtype != RS_C8 && dtype != RS_C16) return fail(RS_ERR_INVALID, "dtype must be RS_C8 or RS_C16");
  if (mode != RS_MODE_FUSED && mode != RS_MODE_PLAIN) return fail(RS_ERR_INVALID, "unknown mode");
  if (grid->N < 64 || (grid->N & (grid->N - 1)) != 0 || grid->N > (int64_t(1) << 30))
    return fail(RS_ERR_UNSUPPORTED, "N must be a power of two in [2^6, 2^30]");
  if (!(grid->dx > 0)) return fail(RS_ERR_INVALID, "dx must be positive");
  return guarded([&] {
    auto f = std::make_unique<rs_field>();
    f->owner = ctx;
    f->grid = *grid;
    if (f->grid.chunk_size <= 0) f->grid.chunk_size = grid->N;
    f->f64 = dtype == RS_C16;
    f->mode = mode;
    rs::hip_check(hipSetDevice(ctx->ctx->device), "hipSetDevice");
    (void)ctx->ctx->plan(grid->N, f->f64);
    f->buf = rs::DevBuf((size_t)grid->N * (f->f64 ? 16 : 8));
    rs::hip_check(hipMemsetAsync(f->buf.p, 0, f->buf.bytes, ctx->ctx->stream), "zero field");
    rs::hip_check(hipStreamSynchronize(ctx->ctx->stream), "sync");
    *out = f.release();
  });
}

int rs_field_destroy(rs_field *f) {
  if (!f) return fail(RS_ERR_INVALID, "field is NULL");
  return guarded([&] { delete f; });
}

int rs_field_upload(rs_field *f, const void *host) {
  if (!f || !host) return fail(RS_ERR_INVALID, "NULL argument");
  if (reserved(f->owner)) return fail_reserved();
  return guarded([&] {
    f->owner->ctx->copy_from_host(f->buf.p, host, f->buf.bytes, f->owner->ctx->stream);
    rs::hip_check(hipStreamSynchronize(f->owner->ctx->stream), "sync");
  });
}

int rs_field_download(rs_field *f, void *host) {
  if (!f || !host) return fail(RS_ERR_INVALID, "NULL argument");
  if (reserved(f->owner)) return fail_reserved();
  return guarded([&] {
    // through the context's pinned staging buffers: pageable destinations
    // get the transfer rate of pinned ones
    f->owner->ctx->copy_to_host(host, f->buf.p, f->buf.bytes, f->owner->ctx->stream);
  });
}

int rs_field_copy(rs_field *dst, const rs_field *src) {
  if (!dst || !src) return fail(RS_ERR_INVALID, "NULL argument");
  if (dst->grid.N != src->grid.N || dst->f64 != src->f64 || dst->owner != src->owner)
    return fail(RS_ERR_INVALID, "fields differ in size, precision or context");
  if (reserved(dst->owner)) return fail_reserved();
  return guarded([&] {
    hipStream_t st = dst->owner->ctx->stream;
    rs::hip_check(dst->f64 ? rs::launch_copy<double>(dst->buf.p, src->buf.p, dst->grid.N, st)
                           : rs::launch_copy<float>(dst->buf.p, src->buf.p, dst->grid.N, st),
                  "copy");
    rs::hip_check(hipStreamSynchronize(st), "sync");
  });
}

int rs_point_source(rs_field *f, double wl, double x_source, double z) {
  if (!f) return fail(RS_ERR_INVALID, "field is NULL");
  if (!(wl > 0)) return fail(RS_ERR_INVALID, "wavelength must be positive");
  if (reserved(f->owner)) return fail_reserved();
  return guarded([&] {
    rs::Member m;
    m.wl = wl;
    rs::Op s{rs::OpKind::Source};
    s.x.kind = rs::XOP_POINT;
    s.x.post[0] = 1.0;
    s.x.xs = x_source;
    s.x.z = z;
    s.x.dx = f->grid.dx;
    s.x.kphase = -2.0 * M_PI / wl;
    m.ops.push_back(s);
    m.ops.push_back(rs::Op{rs::OpKind::Flush});
    run_on_field(f, m, field_rc(f));
  });
}

int rs_propagate(rs_field *f, double wl, double dz, double cutoff_freq) {
  if (!f) return fail(RS_ERR_INVALID, "field is NULL");
  if (!(wl > 0)) return fail(RS_ERR_INVALID, "wavelength must be positive");
  if (reserved(f->owner)) return fail_reserved();
  return guarded([&] {
    rs::Member m;
    m.wl = wl;
    rs::Op p{rs::OpKind::Prop};
    p.dz = dz;
    p.fc = cutoff_freq;
    m.ops.push_back(p);
    m.ops.push_back(rs::Op{rs::OpKind::Flush});
    run_on_field(f, m, field_rc(f));
  });
}

int rs_apply_grating(rs_field *f, double wl, const rs_grating *g, int phase_step, double cutoff_freq) {
  if (!f || !g) return fail(RS_ERR_INVALID, "NULL argument");
  if (g->nr_steps < 1 || g->nr_x_positions < 1 || !g->x_positions)
    return fail(RS_ERR_INVALID, "grating needs nr_steps >= 1 and at least one x position");
  if (phase_step < 0 || phase_step >= g->nr_x_positions) return fail(RS_ERR_INVALID, "phase_step out of range");
  if (reserved(f->owner)) return fail_reserved();
  return guarded([&] {
    rs::SimConfig c = field_cfg(f, wl);
    const rs::GratingCfg gc = grating_from(g);
    rs::Member m;
    m.wl = wl;
    rs::grating_ops(c, gc, phase_step, cutoff_freq, m.ops);
    m.ops.push_back(rs::Op{rs::OpKind::Flush});
    run_on_field(f, m, field_rc(f));
  });
}

int rs_apply_dba_grating(rs_field *f, double wl, const rs_dba_grating *g, int phase_step, double cutoff_freq) {
  if (!f || !g) return fail(RS_ERR_INVALID, "NULL argument");
  if (g->nr_steps < 1 || g->nr_x_positions < 1 || !g->x_positions)
    return fail(RS_ERR_INVALID, "grating needs nr_steps >= 1 and at least one x position");
  if (phase_step < 0 || phase_step >= g->nr_x_positions) return fail(RS_ERR_INVALID, "phase_step out of range");
  if (reserved(f->owner)) return fail_reserved();
  return guarded([&] {
    rs::SimConfig c = field_cfg(f, wl);
    const rs::GratingCfg gc = dba_grating_from(g);
    rs::Member m;
    m.wl = wl;
    rs::grating_ops(c, gc, phase_step, cutoff_freq, m.ops);
    m.ops.push_back(rs::Op{rs::OpKind::Flush});
    run_on_field(f, m, field_rc(f));
  });
}

int rs_apply_sample(rs_field *f, double wl, const rs_sample *s, int phase_step, double cutoff_freq) {
  if (!f || !s || !s->grid || !s->x_positions || (s->nr_materials > 0 && !s->deltabeta))
    return fail(RS_ERR_INVALID, "NULL argument");
  if (s->rows < 1 || s->cols < 1) return fail(RS_ERR_INVALID, "sample grid must not be empty");
  if (!(s->pixel_size_x > 0) || !(s->pixel_size_z > 0))
    return fail(RS_ERR_INVALID, "sample pixel sizes must be positive");
  if (s->nr_materials < 0 || s->nr_x_positions < 1) return fail(RS_ERR_INVALID, "bad material or position count");
  if (phase_step < 0 || phase_step >= s->nr_x_positions) return fail(RS_ERR_INVALID, "phase_step out of range");
  if (reserved(f->owner)) return fail_reserved();
  return guarded([&] {
    rs::SimConfig c = field_cfg(f, wl);
    rs::GratingCfg g;
    g.sample = true;
    g.z_start = s->z_start;
    g.pixel_size_x = s->pixel_size_x;
    g.pixel_size_z = s->pixel_size_z;
    g.rows = s->rows;
    g.cols = s->cols;
    g.grid.assign(s->grid, s->grid + s->rows * s->cols);
    for (int8_t v : g.grid)
      if (v < 0 || v > s->nr_materials) throw rs::Error(RS_ERR_INVALID, "sample grid index out of range");
    g.db_list.assign(1, {0, 0});
    for (int i = 0; i < s->nr_materials; ++i) g.db_list.push_back({s->deltabeta[2 * i], s->deltabeta[2 * i + 1]});
    g.x_positions.assign(s->x_positions, s->x_positions + s->nr_x_positions);
    rs::Member m;
    m.wl = wl;
    rs::sample_ops(c, g, phase_step, cutoff_freq, m);
    m.ops.push_back(rs::Op{rs::OpKind::Flush});
    run_on_field(f, m, field_rc(f));
  });
}

int rs_scale(rs_field *f, double re, double im) {
  if (!f) return fail(RS_ERR_INVALID, "field is NULL");
  if (reserved(f->owner)) return fail_reserved();
  return guarded([&] {
    rs::Member m;
    m.wl = 1.0;
    rs::Op mk{rs::OpKind::Mask};
    mk.x.kind = rs::XOP_SCALAR;
    mk.x.fa[0] = re;
    mk.x.fa[1] = im;
    mk.x.post[0] = 1.0;
    m.ops.push_back(mk);
    m.ops.push_back(rs::Op{rs::OpKind::Flush});
    run_on_field(f, m, field_rc(f));
  });
}

int64_t rs_detector_pixels(const rs_detector *det) {
  if (!det) return -1;
  return rs::detector_pixels(det->size, det->pixel_x);
}

int rs_square_and_downsample(rs_field *f, const rs_detector *det, double z, double *out, int64_t n_out) {
  if (!f || !det || !out) return fail(RS_ERR_INVALID, "NULL argument");
  if (reserved(f->owner)) return fail_reserved();
  return guarded([&] {
    const int64_t npx = rs::detector_pixels(det->size, det->pixel_x);
    if (npx > f->grid.N) throw rs::Error(RS_ERR_INVALID, "detector has more pixels than samples");
    if (n_out < npx) throw rs::Error(RS_ERR_INVALID, "output buffer too small");
    rs::Member m;
    m.wl = 1.0;
    rs::Op d{rs::OpKind::Detect};
    d.row = 0;
    m.ops.push_back(d);
    rs::RunConfig rc = field_rc(f);
    rc.det_size = det->size;
    rc.det_px = det->pixel_x;
    rc.det_py = det->pixel_y;
    rc.z_det = z;
    run_on_field(f, m, rc, out, n_out);
  });
}

int rs_square_and_downsample_noalloc(rs_field *f, const rs_detector *det, double z, double *out, int64_t n_out) {
  if (!f || !det || !out) return fail(RS_ERR_INVALID, "NULL argument");
  if (reserved(f->owner)) return fail_reserved();
  const int64_t npx = rs::detector_pixels(det->size, det->pixel_x);
  if (npx > RS_NOALLOC_MAX_PIXELS)
    return fail(RS_ERR_INVALID, "detector has " + std::to_string(npx) + " pixels, more than the " +
                                    std::to_string(RS_NOALLOC_MAX_PIXELS) + " of the allocation-free read");
  if (npx > f->grid.N) return fail(RS_ERR_INVALID, "detector has more pixels than samples");
  if (n_out < npx) return fail(RS_ERR_INVALID, "output buffer too small");
  if (npx <= 0) return RS_OK;
  rs::Context &c = *f->owner->ctx;
  const hipError_t e =
      f->f64 ? rs::launch_detect_noalloc<double>(f->buf.p, f->grid.N, f->grid.dx, npx, det->pixel_x, c.noalloc_scr,
                                                 c.stream)
             : rs::launch_detect_noalloc<float>(f->buf.p, f->grid.N, f->grid.dx, npx, det->pixel_x, c.noalloc_scr,
                                                c.stream);
  if (e != hipSuccess) return fail(RS_ERR_DEVICE, std::string("detector kernels: ") + hipGetErrorString(e));
  // the FP64 pixel sums straight into the caller's buffer, then big-wave's
  // geometry factors in place: ((sum * cos) * dx * py) / dist
  hipError_t c1 = hipMemcpyAsync(out, c.noalloc_scr, (size_t)npx * sizeof(double), hipMemcpyDeviceToHost, c.stream);
  if (c1 == hipSuccess) c1 = hipStreamSynchronize(c.stream);
  if (c1 != hipSuccess) return fail(RS_ERR_DEVICE, std::string("detector read-back: ") + hipGetErrorString(c1));
  for (int64_t k = 0; k < npx; ++k) {
    double cos_a, dist;
    rs::detector_pixel_scale(det->size, det->pixel_x, z, k, &cos_a, &dist);
    out[k] = (((out[k] * cos_a) * f->grid.dx) * det->pixel_y) / dist;
  }
  return RS_OK;
}

int rs_detector_bounds(int64_t N, double dx, const rs_detector *det, int64_t *bounds, int64_t n_bounds) {
  if (!det || !bounds) return fail(RS_ERR_INVALID, "NULL argument");
  return guarded([&] {
    std::vector<int64_t> b = rs::detector_bounds(N, dx, det->size, det->pixel_x);
    if ((int64_t)b.size() > n_bounds) throw rs::Error(RS_ERR_INVALID, "bounds buffer too small");
    std::memcpy(bounds, b.data(), b.size() * sizeof(int64_t));
  });
}

int rs_autotune(rs_ctx *ctx, int64_t N, int dtype, int members, int32_t *out4, double *ms_per_slice) {
  if (!ctx) return fail(RS_ERR_INVALID, "ctx is NULL");
  if (reserved(ctx)) return fail_reserved();
  if (dtype != RS_C8 && dtype != RS_C16) return fail(RS_ERR_INVALID, "dtype must be RS_C8 or RS_C16");
  if (N < 64 || (N & (N - 1)) != 0 || N > (int64_t(1) << 30))
    return fail(RS_ERR_UNSUPPORTED, "N must be a power of two in [2^6, 2^30]");
  return guarded([&] {
    const rs::Context::Tuned t = rs::autotune(*ctx->ctx, N, dtype == RS_C16, 3, members);
    rs::Plan &p = ctx->ctx->plan(N, dtype == RS_C16, -1, -1, t.members);
    if (out4) {
      out4[0] = p.ngroups;
      out4[1] = p.lgA;
      out4[2] = p.lgB;
      out4[3] = p.lgR;
    }
    if (ms_per_slice) *ms_per_slice = t.ms;
  });
}

int rs_set_split(rs_ctx *ctx, int64_t N, int dtype, int lgR, int lgA) {
  if (!ctx) return fail(RS_ERR_INVALID, "ctx is NULL");
  if (reserved(ctx)) return fail_reserved();
  return guarded([&] {
    int lgN = 0;
    while ((int64_t(1) << lgN) < N) ++lgN;
    int ng, a, b, r;
    rs::plan_split(lgN, dtype == RS_C16, lgR, lgA, &ng, &a, &b, &r);  // validates
    auto &tv = ctx->ctx->tuned;
    tv.erase(std::remove_if(tv.begin(), tv.end(),
                            [&](const rs::Context::Tuned &t) { return t.lgN == lgN && t.f64 == (dtype == RS_C16); }),
             tv.end());
    tv.push_back(rs::Context::Tuned{lgN, dtype == RS_C16, lgR, lgA, 0.0, 0});
  });
}

int rs_tune_get(rs_ctx *ctx, int64_t N, int dtype, int members, rs_tune_entry *out, int32_t *found) {
  if (!ctx || !out || !found) return fail(RS_ERR_INVALID, "NULL argument");
  if (reserved(ctx)) return fail_reserved();
  if (dtype != RS_C8 && dtype != RS_C16) return fail(RS_ERR_INVALID, "dtype must be RS_C8 or RS_C16");
  return guarded([&] {
    int lgN = 0;
    while ((int64_t(1) << lgN) < N) ++lgN;
    const rs::Context::Tuned *t = ctx->ctx->find_tuned(lgN, dtype == RS_C16, members);
    *found = t ? 1 : 0;
    if (t) *out = rs_tune_entry{N, dtype, t->members, t->lgR, t->lgA, t->tileA, t->tileB, t->ms};
  });
}

int rs_tune_save(rs_ctx *ctx, const char *path) {
  if (!ctx || !path) return fail(RS_ERR_INVALID, "NULL argument");
  if (reserved(ctx)) return fail_reserved();
  return guarded([&] { rs::tune_save(*ctx->ctx, path); });
}

int rs_tune_load(rs_ctx *ctx, const char *path, int32_t *n_loaded) {
  if (!ctx || !path) return fail(RS_ERR_INVALID, "NULL argument");
  if (reserved(ctx)) return fail_reserved();
  return guarded([&] {
    const int n = rs::tune_load(*ctx->ctx, path);
    if (n_loaded) *n_loaded = n;
  });
}

int rs_server_create(rs_ctx *ctx, double reserve_bytes, int max_jobs, rs_server **out) {
  if (!ctx || !out) return fail(RS_ERR_INVALID, "NULL argument");
  if (reserved(ctx)) return fail_reserved();
  return guarded([&] {
    auto s = std::make_unique<rs_server>();
    s->owner = ctx;
    ctx->servers++;
    try {
      s->srv = std::make_unique<rs::Server>(*ctx->ctx, reserve_bytes, max_jobs);
    } catch (...) {
      ctx->servers--;
      throw;
    }
    *out = s.release();
  });
}

int rs_server_submit(rs_server *srv, const rs_request *q, int64_t *id) {
  if (!srv || !q || !id) return fail(RS_ERR_INVALID, "NULL argument");
  if (q->nr_energies < 1 || !q->energies_ev || (q->nr_gratings > 0 && !q->gratings) || !q->cutoff_angles)
    return fail(RS_ERR_INVALID, "request needs energies, gratings and cutoff angles");
  if (q->dtype != RS_C8 && q->dtype != RS_C16) return fail(RS_ERR_INVALID, "dtype must be RS_C8 or RS_C16");
  return guarded([&] {
    rs::ServeRequest r;
    for (int e = 0; e < q->nr_energies; ++e) {
      rs::SimConfig c;
      c.N = q->grid.N;
      c.dx = q->grid.dx;
      c.chunk_size = q->grid.chunk_size > 0 ? q->grid.chunk_size : q->grid.N;
      c.f64 = q->dtype == RS_C16;
      c.z_detector = q->z_detector;
      c.detector_size = q->det.size;
      c.px = q->det.pixel_x;
      c.py = q->det.pixel_y;
      c.energy = q->energies_ev[e];
      c.wl = rs::energy_to_wavelength(c.energy);
      c.source.kind = q->source_kind == 1 ? rs::SourceCfg::Plane : rs::SourceCfg::Point;
      c.source.x = q->source_x;
      c.source.z = q->source_z;
      for (int k = 0; k < q->nr_gratings; ++k)
        c.elements.push_back(grating_from(&q->gratings[(size_t)e * q->nr_gratings + k]));
      for (int k = 0; k <= q->nr_gratings; ++k) c.cutoff_angles.push_back(q->cutoff_angles[k]);
      r.members.push_back(std::move(c));
      r.weights.push_back(q->weights ? q->weights[e] : 1.0);
    }
    *id = srv->srv->submit(std::move(r));
  });
}

int rs_server_poll(rs_server *srv, int64_t id, int32_t *state, double *out, int64_t out_cap, int64_t *shape2,
                   double *times3) {
  if (!srv || !state) return fail(RS_ERR_INVALID, "NULL argument");
  return guarded([&] {
    const rs::ServeResult r = srv->srv->poll(id);
    *state = r.state;
    if (shape2) {
      shape2[0] = r.n_steps;
      shape2[1] = r.n_px;
    }
    if (r.state == 3) throw rs::Error(RS_ERR_DEVICE, "request failed: " + r.error);
    if (r.state == 2) {
      if (out && out_cap >= (int64_t)r.rows.size()) std::memcpy(out, r.rows.data(), r.rows.size() * sizeof(double));
      if (times3) {
        times3[0] = r.start_t - r.submit_t;
        times3[1] = r.done_t - r.start_t;
        times3[2] = r.done_t - r.submit_t;
      }
    }
  });
}

int rs_server_wait(rs_server *srv, int64_t id, double timeout_s, int32_t *done) {
  if (!srv) return fail(RS_ERR_INVALID, "server is NULL");
  return guarded([&] {
    const bool d = srv->srv->wait(id, timeout_s);
    if (done) *done = d ? 1 : 0;
  });
}

int rs_server_stats(rs_server *srv, double *o) {
  if (!srv || !o) return fail(RS_ERR_INVALID, "NULL argument");
  return guarded([&] {
    const rs::ServeStats s = srv->srv->stats();
    o[0] = (double)s.ticks;
    o[1] = (double)s.launches;
    o[2] = s.launches ? (double)s.merged_jobs / (double)s.launches : 0.0;
    o[3] = (double)s.jobs_done;
    o[4] = (double)s.max_active;
    o[5] = s.busy_s;
  });
}

int rs_server_destroy(rs_server *srv) {
  if (!srv) return fail(RS_ERR_INVALID, "server is NULL");
  return guarded([&] {
    srv->srv->drain();
    rs_ctx *owner = srv->owner;
    delete srv;
    owner->servers--;
  });
}

int rs_set_timing(rs_ctx *ctx, int enable) {
  if (!ctx) return fail(RS_ERR_INVALID, "ctx is NULL");
  if (reserved(ctx)) return fail_reserved();
  if (enable < 0) return fail(RS_ERR_INVALID, "timing stride must be >= 0");
  ctx->ctx->timing = enable != 0;
  ctx->ctx->timing_stride = enable > 0 ? enable : 1;
  for (auto &s : ctx->ctx->stats) s = rs::PassStat{};
  for (auto &n : ctx->ctx->timing_seen) n = 0;
  return RS_OK;
}

int rs_get_pass_stats(rs_ctx *ctx, rs_pass_stat *out, int n) {
  if (!ctx || !out) return fail(RS_ERR_INVALID, "NULL argument");
  for (int i = 0; i < n && i < rs::PC_COUNT; ++i) {
    out[i].launches = ctx->ctx->stats[i].launches;
    out[i].total_ms = ctx->ctx->stats[i].ms;
    out[i].bytes = ctx->ctx->stats[i].bytes;
  }
  return RS_OK;
}

namespace {
// member program of one (energy, source) batch member
void build_batch_member(const rs_grid &grid, const rs_detector &det, double z_detector, int dtype, int nr_gratings,
                        const rs_batch_member &bm, int out_index, rs::Member &m) {
  rs::SimConfig c;
  c.N = grid.N;
  c.dx = grid.dx;
  c.chunk_size = grid.chunk_size > 0 ? grid.chunk_size : grid.N;
  c.f64 = dtype == RS_C16;
  c.z_detector = z_detector;
  c.detector_size = det.size;
  c.px = det.pixel_x;
  c.py = det.pixel_y;
  c.wl = bm.wl;
  c.source.kind = bm.source_kind == 1 ? rs::SourceCfg::Plane : rs::SourceCfg::Point;
  c.source.x = bm.source_x;
  c.source.z = bm.source_z;
  if (nr_gratings > 0 && !bm.gratings) throw rs::Error(RS_ERR_INVALID, "member without gratings");
  if (!bm.cutoff_angles) throw rs::Error(RS_ERR_INVALID, "member without cutoff angles");
  for (int k = 0; k < nr_gratings; ++k) c.elements.push_back(grating_from(&bm.gratings[k]));
  for (int k = 0; k <= nr_gratings; ++k) c.cutoff_angles.push_back(bm.cutoff_angles[k]);
  const int nsteps = rs::build_member(c, m, out_index, false);
  if (nsteps != 1) throw rs::Error(RS_ERR_UNSUPPORTED, "batched members must not be phase stepped");
  m.weight = bm.weight;
}
}  // namespace

int rs_copy_bandwidth(rs_ctx *ctx, int64_t bytes, int reps, double *gbs) {
  if (!ctx || !gbs) return fail(RS_ERR_INVALID, "NULL argument");
  if (reserved(ctx)) return fail_reserved();
  if (bytes < 8 || bytes % 8 || reps < 1) return fail(RS_ERR_INVALID, "bytes must be a positive multiple of 8, reps >= 1");
  return guarded([&] {
    rs::Context &c = *ctx->ctx;
    rs::hip_check(hipSetDevice(c.device), "hipSetDevice");
    rs::DevBuf a((size_t)bytes), b((size_t)bytes);
    const int64_t n = bytes / 8;
    rs::hip_check(rs::launch_copy<float>(b.p, a.p, n, c.stream), "copy");  // warm
    hipEvent_t e0, e1;
    rs::hip_check(hipEventCreate(&e0), "event");
    rs::hip_check(hipEventCreate(&e1), "event");
    rs::hip_check(hipEventRecord(e0, c.stream), "event record");
    for (int r = 0; r < reps; ++r) rs::hip_check(rs::launch_copy<float>(b.p, a.p, n, c.stream), "copy");
    rs::hip_check(hipEventRecord(e1, c.stream), "event record");
    rs::hip_check(hipEventSynchronize(e1), "event sync");
    float ms = 0;
    rs::hip_check(hipEventElapsedTime(&ms, e0, e1), "event elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
  });
}

int rs_batch_run_mixed(rs_ctx *ctx, const rs_mixed_desc *d, double *out, int64_t n_out) {
  if (!ctx || !d || !out) return fail(RS_ERR_INVALID, "NULL argument");
  if (reserved(ctx)) return fail_reserved();
  if (d->nr_sims < 1 || !d->sims) return fail(RS_ERR_INVALID, "mixed batch has no simulations");
  if (d->dtype != RS_C8 && d->dtype != RS_C16) return fail(RS_ERR_INVALID, "dtype must be RS_C8 or RS_C16");
  if (d->mode != RS_MODE_FUSED && d->mode != RS_MODE_PLAIN) return fail(RS_ERR_INVALID, "unknown mode");
  return guarded([&] {
    // parts: simulations that share grid, detector and the operation
    // sequence of their member programs (slice counts, cutoffs, substrates)
    struct Part {
      const rs_batch_sim *first;
      std::string shape;
      std::vector<int> sims;
      std::vector<rs::Member> members;
    };
    auto shape_of = [](const rs::Member &m) {
      std::string sh;
      for (const auto &o : m.ops) sh.push_back((char)('A' + (int)o.kind));
      return sh;
    };
    std::vector<Part> parts;
    auto same = [](const rs_batch_sim &a, const rs_batch_sim &b) {
      return a.grid.N == b.grid.N && a.grid.dx == b.grid.dx && a.grid.chunk_size == b.grid.chunk_size &&
             a.det.size == b.det.size && a.det.pixel_x == b.det.pixel_x && a.det.pixel_y == b.det.pixel_y &&
             a.z_detector == b.z_detector && a.nr_gratings == b.nr_gratings;
    };
    for (int i = 0; i < d->nr_sims; ++i) {
      const rs_batch_sim &s = d->sims[i];
      if (s.nr_members < 1 || !s.members) throw rs::Error(RS_ERR_INVALID, "simulation without members");
      if (s.grid.N < 64 || (s.grid.N & (s.grid.N - 1)) != 0 || s.grid.N > (int64_t(1) << 30))
        throw rs::Error(RS_ERR_UNSUPPORTED, "N must be a power of two in [2^6, 2^30]");
      const int64_t npx = rs::detector_pixels(s.det.size, s.det.pixel_x);
      if (s.out_offset < 0 || s.out_offset + npx > n_out) throw rs::Error(RS_ERR_INVALID, "output buffer too small");
      std::vector<rs::Member> ms(s.nr_members);
      for (int k = 0; k < s.nr_members; ++k)
        build_batch_member(s.grid, s.det, s.z_detector, d->dtype, s.nr_gratings, s.members[k], 0, ms[k]);
      const std::string sh = shape_of(ms[0]);
      for (const auto &m : ms)
        if (shape_of(m) != sh) throw rs::Error(RS_ERR_INVALID, "the members of a simulation differ in structure");
      Part *p = nullptr;
      for (auto &q : parts)
        if (same(*q.first, s) && q.shape == sh) p = &q;
      if (!p) {
        parts.push_back(Part{&s, sh, {}, {}});
        p = &parts.back();
      }
      const int row = (int)p->sims.size();
      p->sims.push_back(i);
      for (auto &m : ms) {
        m.out_base = row;
        p->members.push_back(std::move(m));
      }
    }
    std::vector<rs::MixedPart> mp;
    for (auto &p : parts) {
      const rs_batch_sim &s = *p.first;
      rs::MixedPart x;
      x.rc.N = s.grid.N;
      x.rc.dx = s.grid.dx;
      x.rc.f64 = d->dtype == RS_C16;
      x.rc.mode = d->mode;
      x.rc.det_size = s.det.size;
      x.rc.det_px = s.det.pixel_x;
      x.rc.det_py = s.det.pixel_y;
      x.rc.z_det = s.z_detector;
      x.rc.n_out = (int)p.sims.size();
      x.members = &p.members;
      mp.push_back(x);
    }
    const std::vector<std::vector<double>> rows = rs::run_mixed(*ctx->ctx, mp, d->use_graph != 0);
    for (size_t q = 0; q < parts.size(); ++q) {
      const int64_t npx = rs::detector_pixels(parts[q].first->det.size, parts[q].first->det.pixel_x);
      for (size_t r = 0; r < parts[q].sims.size(); ++r)
        std::memcpy(out + d->sims[parts[q].sims[r]].out_offset, rows[q].data() + r * npx, (size_t)npx * sizeof(double));
    }
  });
}

int rs_batch_run(rs_ctx *ctx, const rs_batch_desc *d, double *out, int64_t n_out) {
  if (!ctx || !d || !out) return fail(RS_ERR_INVALID, "NULL argument");
  if (reserved(ctx)) return fail_reserved();
  if (d->nr_members < 1 || !d->members) return fail(RS_ERR_INVALID, "batch has no members");
  if (d->dtype != RS_C8 && d->dtype != RS_C16) return fail(RS_ERR_INVALID, "dtype must be RS_C8 or RS_C16");
  return guarded([&] {
    const int64_t npx = rs::detector_pixels(d->det.size, d->det.pixel_x);
    if (n_out < (int64_t)d->nr_outputs * npx) throw rs::Error(RS_ERR_INVALID, "output buffer too small");
    std::vector<rs::Member> ms(d->nr_members);
    for (int i = 0; i < d->nr_members; ++i) {
      const rs_batch_member &bm = d->members[i];
      if (bm.out_index < 0 || bm.out_index >= d->nr_outputs) throw rs::Error(RS_ERR_INVALID, "out_index out of range");
      build_batch_member(d->grid, d->det, d->z_detector, d->dtype, d->nr_gratings, bm, bm.out_index, ms[i]);
    }
    rs::RunConfig rc;
    rc.N = d->grid.N;
    rc.dx = d->grid.dx;
    rc.f64 = d->dtype == RS_C16;
    rc.mode = d->mode;
    rc.det_size = d->det.size;
    rc.det_px = d->det.pixel_x;
    rc.det_py = d->det.pixel_y;
    rc.z_det = d->z_detector;
    rc.n_out = d->nr_outputs;
    rc.max_resident = d->max_resident;
    rc.use_graph = rc.whole_graph = d->use_graph != 0;
    if (d->lanes < 0 || d->lanes > 4) throw rs::Error(RS_ERR_INVALID, "lanes must be 0..4");
    rc.lanes = d->lanes;
    if (d->spans < 0) throw rs::Error(RS_ERR_INVALID, "spans must be >= 0");
    rc.spans = d->spans;
    std::vector<double> det = rs::run_members(*ctx->ctx, rc, ms);
    std::memcpy(out, det.data(), det.size() * sizeof(double));
  });
}

}  // extern "C"
