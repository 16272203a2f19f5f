#include "simulation.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "npy.hpp"
#include "yaml_lite.hpp"

#pragma STDC FP_CONTRACT OFF

namespace rs {

std::string subdir_name(int source_idx) {
  char b[32];
  std::snprintf(b, sizeof b, "%08d", source_idx);
  return b;
}

namespace {

double get_f(const YNode &n, const std::string &k) { return n[k].as_double(k); }

struct DbEntry {
  std::string name;
  double density;
  std::complex<double> db;
};

// closest-density match within 1e-5 (big-wave/optical_element.py:42-63)
std::complex<double> lookup(const std::vector<DbEntry> &t, const YNode &mat, const std::string &what) {
  if (mat.is_null()) return {0, 0};
  if (mat.type != YNode::Seq || mat.size() < 2) throw Error(2, "material " + what + " must be [name, density]");
  const std::string name = mat[0].as_string(what);
  const double dens = mat[1].as_double(what + ".density");
  const DbEntry *best = nullptr;
  double bd = 0;
  for (auto &e : t) {
    if (e.name != name) continue;
    const double d = std::fabs(e.density - dens);
    if (!best || d < bd) {
      best = &e;
      bd = d;
    }
  }
  if (!best) throw Error(2, "Material " + name + " not found in deltabeta table");
  if (!(bd < 1e-5)) throw Error(2, "Material " + name + " found with density error of " + std::to_string(bd));
  return best->db;
}

}  // namespace

SimConfig parse_simdir(const std::string &sim_dir, int source_idx) {
  const YNode cfg = yaml_load_file(sim_dir + "/config.yaml");
  const YNode comp = yaml_load_file(sim_dir + "/computed.yaml");
  const YNode sub = yaml_load_file(sim_dir + "/" + subdir_name(source_idx) + "/subconfig.yaml");
  SimConfig c;
  const YNode &sp = cfg["sim_params"];
  c.N = sp["N"].as_int("N");
  c.dx = get_f(sp, "dx");
  c.z_detector = get_f(sp, "z_detector");
  c.detector_size = get_f(sp, "detector_size");
  c.px = get_f(sp, "detector_pixel_size_x");
  c.py = get_f(sp, "detector_pixel_size_y");
  c.chunk_size = sp["chunk_size"].as_int("chunk_size");
  if (cfg.has("dtype") && !cfg["dtype"].is_null()) {
    const std::string d = cfg["dtype"].as_string("dtype");
    if (d == "c8") c.f64 = false;
    else if (d == "c16") c.f64 = true;
    else throw Error(2, "Unknown dtype: " + d + ". Has to be either c8 or c16");
  }
  c.save_final_u_vectors = cfg["save_final_u_vectors"].as_bool("save_final_u_vectors");
  for (size_t i = 0; i < comp["cutoff_angles"].size(); ++i)
    c.cutoff_angles.push_back(comp["cutoff_angles"][i].as_double("cutoff_angles"));

  c.energy = sub["energy"].as_double("energy");
  c.wl = energy_to_wavelength(c.energy);
  std::vector<DbEntry> table;
  if (sub.has("deltabeta_table") && !sub["deltabeta_table"].is_null()) {
    const YNode &t = sub["deltabeta_table"];
    for (size_t i = 0; i < t.size(); ++i) {
      const YNode &e = t[i];
      DbEntry d;
      d.name = e[0][0].as_string("deltabeta_table material");
      d.density = e[0][1].as_double("deltabeta_table density");
      d.db = {e[1][0].as_double("delta"), e[1][1].as_double("beta")};
      // a non-finite delta/beta would turn every sample it touches into NaN
      if (!std::isfinite(d.db.real()) || !std::isfinite(d.db.imag()))
        throw Error(2, "deltabeta of material " + d.name + " is not finite");
      table.push_back(d);
    }
  }
  const YNode &src = sub["source"];
  const std::string st = src["type"].as_string("source.type");
  if (st == "point") {
    c.source.kind = SourceCfg::Point;
    c.source.x = get_f(src, "x");
    c.source.z = get_f(src, "z");
  } else if (st == "vector") {
    c.source.kind = SourceCfg::Vector;
    c.source.input_path = src["input_path"].as_string("input_path");
    c.source.z = get_f(src, "z");
  } else {
    throw Error(2, "Unknown source type: " + st + ".");
  }

  const YNode &els = cfg["elements"];
  for (size_t i = 0; i < els.size(); ++i) {
    const YNode &e = els[i];
    const std::string type = e["type"].as_string("type");
    if (type == "sample") {
      // big-wave/config.py:118-131: grid from an int8 .npy next to config.yaml
      GratingCfg g;
      g.sample = true;
      g.z_start = get_f(e, "z_start");
      g.pixel_size_x = get_f(e, "pixel_size_x");
      g.pixel_size_z = get_f(e, "pixel_size_z");
      g.x_positions.clear();
      const YNode &xp = e["x_positions"];
      for (size_t k = 0; k < xp.size(); ++k) g.x_positions.push_back(xp[k].as_double("x_positions"));
      const NpyArray a = npy_load(sim_dir + "/" + e["grid_path"].as_string("grid_path"));
      if (a.descr != "|i1" && a.descr != "<i1") throw Error(2, "sample grid must be int8");
      if (a.shape.size() != 2) throw Error(2, "sample grid must be two-dimensional");
      g.rows = a.shape[0];
      g.cols = a.shape[1];
      g.grid.assign((const int8_t *)a.data.data(), (const int8_t *)a.data.data() + a.count());
      const YNode &mats = e["materials"];
      g.db_list.assign(1, {0, 0});
      for (size_t k = 0; k < mats.size(); ++k) g.db_list.push_back(lookup(table, mats[k], "materials"));
      c.elements.push_back(g);
      continue;
    }
    if (type != "grating" && type != "dba_grating")
      throw Error(6, "optical element type '" + type +
                         "' is not supported by this build (grating, dba_grating, sample)");
    GratingCfg g;
    if (type == "dba_grating") {
      // fast-wave/src/config_parsing.cpp:107-137
      g.dba = true;
      g.pitch = get_f(e, "pitch0");
      g.pitch1 = get_f(e, "pitch1");
      g.dc0 = e["dc0"][0].as_double("dc0.0");
      g.dc1 = e["dc0"][1].as_double("dc0.1");
      g.dcb0 = e["dc1"][0].as_double("dc1.0");
      g.dcb1 = e["dc1"][1].as_double("dc1.1");
    } else {
      g.pitch = get_f(e, "pitch");
      g.dc0 = e["dc"][0].as_double("dc.0");
      g.dc1 = e["dc"][1].as_double("dc.1");
    }
    g.z_start = get_f(e, "z_start");
    g.thickness = get_f(e, "thickness");
    g.nr_steps = (int)e["nr_steps"].as_int("nr_steps");
    g.x_positions.clear();
    const YNode &xp = e["x_positions"];
    for (size_t k = 0; k < xp.size(); ++k) g.x_positions.push_back(xp[k].as_double("x_positions"));
    g.substrate_thickness = get_f(e, "substrate_thickness");
    auto mat = [&](const char *k) -> std::complex<double> {
      if (!e.has(k)) return {0, 0};
      return lookup(table, e[k], k);
    };
    g.db_a = mat("mat_a");
    g.db_b = mat("mat_b");
    g.db_s = mat("mat_substrate");
    c.elements.push_back(g);
  }
  return c;
}

void check_config(const SimConfig &c) {
  if (c.N < 64 || (c.N & (c.N - 1)) != 0) throw Error(6, "N must be a power of two >= 64");
  if (!(c.dx > 0)) throw Error(2, "dx must be positive");
  if (c.chunk_size <= 0) throw Error(2, "chunk_size must be positive");
  if (!(c.px > 0) || !(c.detector_size > 0)) throw Error(2, "Detector size and pixel size must be positive");
  if (detector_pixels(c.detector_size, c.px) > c.N) throw Error(2, "detector has more pixels than samples");
  // a sim dir set up with a save_and_exit element keeps the angles of the
  // elements cut after it (big-wave/multisim.py:166-192, 499-507)
  if (c.cutoff_angles.size() < c.elements.size() + 1)
    throw Error(2, "cutoff_angles must have at least len(elements) + 1 entries");
  const double tol = 1e-8;
  double cur = c.source.z;
  size_t nsteps = 0;
  for (const auto &g : c.elements) {
    if (g.sample) {
      // big-wave Sample.check_valid (optical_element.py:410-419)
      if (g.rows <= 0 || g.cols <= 0) throw Error(2, "sample grid must not be empty");
      if (!(g.z_start > 0)) throw Error(2, "sample z_start must be positive");
      if (!(g.pixel_size_x > 0) || !(g.pixel_size_z > 0)) throw Error(2, "sample pixel sizes must be positive");
      int maxidx = 0;
      for (int8_t v : g.grid) {
        if (v < 0) throw Error(2, "sample grid indices must be >= 0");
        maxidx = std::max(maxidx, (int)v);
      }
      if (maxidx >= (int)g.db_list.size()) throw Error(2, "sample grid indexes more materials than given");
    } else if (!(g.thickness > 0)) {
      throw Error(2, "grating thickness must be positive");
    }
    if (g.substrate_thickness < 0) throw Error(2, "substrate thickness must be >= 0");
    // big-wave DBAGrating.check_valid (optical_element.py:318-320)
    if (g.dba && !(g.thickness > 0)) throw Error(2, "DBA grating thickness must be positive");
    if (!g.sample && g.nr_steps < 1) throw Error(2, "nr_steps must be >= 1");
    if (g.x_positions.empty()) throw Error(2, "x_positions must not be empty");
    if (g.x_positions.size() > 1) {
      if (nsteps && nsteps != g.x_positions.size())
        throw Error(2, "If multiple elements have phase stepping enabled, then all elements must have the same number of steps.");
      nsteps = g.x_positions.size();
    }
    if (!(cur <= g.z_start + tol)) throw Error(2, "Overlapping optical elements are not supported");
    cur = g.z_start + g.total_thickness();
  }
  if (!(cur <= c.z_detector + tol)) throw Error(2, "Detector must be after the last optical element");
}

namespace {

XOp mask_op(const SimConfig &c, const GratingCfg &g, int slice, int step, const double fa[2], const double fb[2]) {
  XOp x{};
  x.kind = XOP_MASK;
  x.fa[0] = fa[0];
  x.fa[1] = fa[1];
  x.fb[0] = fb[0];
  x.fb[1] = fb[1];
  x.post[0] = 1.0;
  x.post[1] = 0.0;
  // big-wave/optical_element.py:156, 351: dc interpolated in FP64
  const double zf = (double)slice / (double)g.nr_steps;
  const double dc = (1.0 - zf) * g.dc0 + zf * g.dc1;
  x.pitch = g.pitch;
  x.dc = dc;
  x.half_dc = dc / 2.0;
  x.xpos = g.x_positions[step];
  x.dx = c.dx;
  x.chunk = c.chunk_size;
  x.single_chunk = c.N <= c.chunk_size;
  x.inv_pitch = 1.0 / g.pitch;
  if (g.dba) {
    // big-wave/optical_element.py:368-369: both duty cycles interpolated in FP64
    x.kind = XOP_DBA;
    x.pitch1 = g.pitch1;
    x.inv_pitch1 = 1.0 / g.pitch1;
    x.dc1 = (1.0 - zf) * g.dcb0 + zf * g.dcb1;
  }
  // np.linspace(x_start, x_start + N dx, N, endpoint=False) of the single chunk
  // (big-wave/optical_element.py:352-355), evaluated as numpy does
  const double x_start = (double)(-(c.N / 2)) * c.dx + x.xpos;
  const double x_stop = x_start + (double)c.N * c.dx;
  x.x0 = x_start;
  x.xstep = (x_stop - x_start) / (double)c.N;
  return x;
}

}  // namespace

void grating_ops(const SimConfig &c, const GratingCfg &g, int step, double fc, std::vector<Op> &ops,
                 const SliceHook &hook) {
  const double zs = g.thickness / (double)g.nr_steps;
  double fa[2], fb[2];
  material_factor(g.db_a.real(), g.db_a.imag(), zs, c.wl, fa);
  material_factor(g.db_b.real(), g.db_b.imag(), zs, c.wl, fb);
  for (int i = 0; i < g.nr_steps; ++i) {
    Op m{OpKind::Mask};
    m.x = mask_op(c, g, i, step, fa, fb);
    ops.push_back(m);
    Op p{OpKind::Prop};
    p.dz = zs;
    p.fc = fc;
    ops.push_back(p);
    if (hook) hook(g.z_start + (double)(i + 1) * zs);
  }
  if (g.substrate_thickness > 0) {
    double fs[2];
    material_factor(g.db_s.real(), g.db_s.imag(), g.substrate_thickness, c.wl, fs);
    Op m{OpKind::Mask};
    m.x.kind = XOP_SCALAR;
    m.x.fa[0] = fs[0];
    m.x.fa[1] = fs[1];
    m.x.post[0] = 1.0;
    m.x.post[1] = 0.0;
    ops.push_back(m);
    Op p{OpKind::Prop};
    p.dz = g.substrate_thickness;
    p.fc = fc;
    ops.push_back(p);
  }
}

void sample_ops(const SimConfig &c, const GratingCfg &g, int step, double fc, Member &m, const SliceHook &hook) {
  // row_x = np.arange(rowlen) * px + x_positions[step] - rowlen * px * 0.5
  SampleRows sr;
  sr.len = g.cols;
  sr.rx.resize((size_t)g.cols);
  const double xpos = g.x_positions[(size_t)step];
  const double half = (double)g.cols * g.pixel_size_x * 0.5;
  for (int64_t k = 0; k < g.cols; ++k) sr.rx[(size_t)k] = ((double)k * g.pixel_size_x + xpos) - half;
  sr.db.resize((size_t)(2 * g.rows * g.cols));
  sr.f.resize((size_t)(2 * g.rows * g.cols));
  for (int64_t r = 0; r < g.rows; ++r)
    for (int64_t k = 0; k < g.cols; ++k) {
      const std::complex<double> db = g.db_list[(size_t)g.grid[(size_t)(r * g.cols + k)]];
      const size_t o = (size_t)(2 * (r * g.cols + k));
      sr.db[o] = db.real();
      sr.db[o + 1] = db.imag();
      double f[2];
      material_factor(db.real(), db.imag(), g.pixel_size_z, c.wl, f);
      sr.f[o] = f[0];
      sr.f[o + 1] = f[1];
    }
  const int si = (int)m.samples.size();
  m.samples.push_back(std::move(sr));
  const double a = (2.0 * M_PI * g.pixel_size_z) / c.wl;  // as material_factor forms it
  for (int64_t r = 0; r < g.rows; ++r) {
    if (hook) hook(g.z_start + (double)r * g.pixel_size_z);
    Op mk{OpKind::Mask};
    mk.x.kind = XOP_SAMPLE;
    mk.x.post[0] = 1.0;
    mk.x.post[1] = 0.0;
    mk.x.dx = c.dx;
    mk.x.s_inv_px = 1.0 / g.pixel_size_x;
    mk.x.s_a = a;
    mk.sample = si;
    mk.sample_row = (int)r;
    m.ops.push_back(mk);
    Op p{OpKind::Prop};
    p.dz = g.pixel_size_z;
    p.fc = fc;
    m.ops.push_back(p);
  }
}

namespace {

void grid_density_check(double dz, double xs, int64_t N, double dx, double wl) {
  // big-wave/propagation.py:31-45
  auto dist = [&](double idx) {
    const double a = ((double)N / 2.0 - idx) * dx + std::fabs(xs);
    return std::sqrt(a * a + dz * dz);
  };
  if (!(std::fabs(dist(0) - dist(1)) <= wl / 2.0))
    throw Error(2, "Nyquist not satisfied, dx too large for the point source");
}

}  // namespace

std::vector<double> detector_x(double size, double px) {
  const double n = py_floordiv(size, px);
  const double h = py_floordiv(n, 2.0);
  std::vector<double> x((size_t)std::max<int64_t>((int64_t)n, 0));
  for (size_t k = 0; k < x.size(); ++k) x[k] = ((double)k - h) * px;
  return x;
}

namespace {

// Analytic history rows of a point source (big-wave/history.py:43-66) at
// z = i * hdz for 0 < i < ceil(z_out / hdz) (source.py:65-68).
void analytic_history(const SimConfig &c, double z_out, double hdz, double fc, Recording &rec) {
  const double gradient = fc * c.wl;
  const double n = py_floordiv(c.detector_size, c.px);
  const double h = py_floordiv(n, 2.0);
  const int64_t cnt = (int64_t)std::ceil(z_out / hdz);
  for (int64_t i = 1; i < cnt; ++i) {
    const double z = (double)i * hdz;
    std::vector<double> row((size_t)std::max<int64_t>((int64_t)n, 0));
    for (size_t k = 0; k < row.size(); ++k) {
      const double x = ((double)k - h) * c.px - c.source.x;
      const double ang = std::atan(x / z);
      const double r2 = x * x + z * z;
      row[k] = std::fabs(x) > gradient * z ? 0.0 : ((c.px * c.py) * std::cos(ang)) / r2;
    }
    rec.hist_z.push_back(z);
    rec.hist_row.push_back(-1);
    rec.analytic.push_back(std::move(row));
  }
}

}  // namespace

int build_member(const SimConfig &c, Member &m, int out_base, bool save_u, Recording *rec) {
  check_config(c);
  m.wl = c.wl;
  m.out_base = out_base;
  std::vector<Op> &ops = m.ops;
  ops.clear();
  m.samples.clear();
  std::vector<double> fcs;
  for (double a : c.cutoff_angles) fcs.push_back(std::sin(a) / c.wl);
  const double tol = 1e-8;

  size_t nsteps = 1;
  for (const auto &g : c.elements) nsteps = std::max(nsteps, g.x_positions.size());
  const bool hist = rec && rec->history_dz > 0;
  if (rec) {
    rec->hist_z.clear();
    rec->hist_row.clear();
    rec->analytic.clear();
    rec->n_rows = (int)nsteps;
  }
  // one history entry read from the device into the next free row
  auto hist_detect = [&](double z) {
    Op d{OpKind::Detect};
    d.row = rec->n_rows++;
    ops.push_back(d);
    rec->hist_z.push_back(z);
    rec->hist_row.push_back(d.row);
    rec->analytic.emplace_back();  // one slot per entry: device rows leave it empty
  };
  const SliceHook hook = hist ? SliceHook(hist_detect) : SliceHook();
  // big-wave/history.py:69-135 (propagate_with_history): entries at
  // zs = np.arange(0, dz, hdz) with the last one moved or appended to dz,
  // each propagated from the segment's start state
  auto propagate = [&](double dz, double z0, double fc_) {
    if (!hist) {
      Op p{OpKind::Prop};
      p.dz = dz;
      p.fc = fc_;
      ops.push_back(p);
      return;
    }
    if (dz == 0) return;
    const double hdz = rec->history_dz;
    const double cnt = std::ceil(dz / hdz);
    if (!(cnt <= 1e6)) throw Error(2, "history_dz gives more than 10^6 entries in one segment");
    std::vector<double> zs((size_t)cnt);
    for (size_t i = 0; i < zs.size(); ++i) zs[i] = (double)i * hdz;
    if (dz - zs.back() < hdz / 2.0 && zs.size() > 1) zs.back() = dz;
    else zs.push_back(dz);
    if (zs.size() > 2) ops.push_back(Op{OpKind::HSave});
    for (size_t i = 1; i < zs.size(); ++i) {
      Op p{OpKind::Prop};
      p.dz = zs[i] - zs[0];
      p.fc = fc_;
      ops.push_back(p);
      hist_detect(zs[i] + z0);
      if (i + 1 < zs.size()) ops.push_back(Op{OpKind::HRestore});
    }
  };

  double fc = fcs[0];
  double cur = c.elements.empty() ? c.z_detector : c.elements[0].z_start;
  // source (big-wave/source.py:46-102)
  if (c.source.kind == SourceCfg::Point) {
    const double dz = cur - c.source.z;
    grid_density_check(dz, c.source.x, c.N, c.dx, c.wl);
    Op s{OpKind::Source};
    s.x.kind = XOP_POINT;
    s.x.post[0] = 1.0;
    s.x.xs = c.source.x;
    s.x.z = dz;
    s.x.dx = c.dx;
    s.x.kphase = -2.0 * M_PI / c.wl;
    ops.push_back(s);
    Op p{OpKind::Prop};
    p.dz = 0.0;
    p.fc = fc;
    ops.push_back(p);
    if (hist) analytic_history(c, cur, rec->history_dz, fc, *rec);
  } else {
    if (c.source.kind == SourceCfg::Plane) {
      Op s{OpKind::Source};
      s.x.kind = XOP_ONES;
      s.x.post[0] = 1.0;
      ops.push_back(s);
    } else {
      ops.push_back(Op{OpKind::Upload});
    }
    const double dz = cur - c.source.z;
    if (dz != 0) propagate(dz, c.source.z, fc);
  }

  int snap_idx = -1;
  double snap_z = 0;
  for (size_t i = 0; i < c.elements.size(); ++i) {
    const GratingCfg &g = c.elements[i];
    if (cur < g.z_start) propagate(g.z_start - cur, cur, fc);
    cur = g.z_start;
    if (g.x_positions.size() > 1 && snap_idx < 0) {
      snap_idx = (int)i;
      snap_z = cur;
      ops.push_back(Op{OpKind::Snapshot});
    }
    fc = fcs[i + 1];
    if (rec && rec->keypoints) {
      Op k{OpKind::Keypoint};
      k.row = 2 * (int)i;
      ops.push_back(k);
    }
    if (g.sample) sample_ops(c, g, 0, fc, m, hook);
    else grating_ops(c, g, 0, fc, ops, hook);
    if (rec && rec->keypoints) {
      Op k{OpKind::Keypoint};
      k.row = 2 * (int)i + 1;
      ops.push_back(k);
    }
    cur = g.z_start + g.total_thickness();
  }
  auto final_hop = [&](bool first) {
    if (cur < c.z_detector - tol) {
      if (first) {
        propagate(c.z_detector - cur, cur, fc);
      } else {
        Op p{OpKind::Prop};
        p.dz = c.z_detector - cur;
        p.fc = fc;
        ops.push_back(p);
      }
    }
  };
  final_hop(true);
  {
    Op d{OpKind::Detect};
    d.row = 0;
    ops.push_back(d);
    if (save_u) {
      Op s{OpKind::SaveU};
      s.row = 0;
      ops.push_back(s);
    }
  }
  if (snap_idx >= 0) {
    for (size_t step = 1; step < nsteps; ++step) {
      ops.push_back(Op{OpKind::Restore});
      cur = snap_z;
      fc = fcs[snap_idx];
      for (size_t i = snap_idx; i < c.elements.size(); ++i) {
        const GratingCfg &g = c.elements[i];
        if (cur < g.z_start) {
          Op p{OpKind::Prop};
          p.dz = g.z_start - cur;
          p.fc = fc;
          ops.push_back(p);
        }
        cur = g.z_start;
        fc = fcs[i + 1];
        const int st = g.x_positions.size() > 1 ? (int)step : 0;
        if (g.sample) sample_ops(c, g, st, fc, m);
        else grating_ops(c, g, st, fc, ops);
        cur = g.z_start + g.total_thickness();
      }
      final_hop(false);
      Op d{OpKind::Detect};
      d.row = (int)step;
      ops.push_back(d);
      if (save_u) {
        Op s{OpKind::SaveU};
        s.row = (int)step;
        ops.push_back(s);
      }
    }
  }
  return (int)nsteps;
}

}  // namespace rs
