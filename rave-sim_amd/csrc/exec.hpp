// Internal: launch records and the member-program compiler / executor shared
// by the lockstep runner (engine.cpp) and the serving scheduler (serve.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "rs_kernels.hpp"

namespace rs {
namespace exec {

enum class LKind { Fiber, Xop, Kop, Detect, Copy, Fill, SaveU, Upload };

struct Launch {
  LKind kind;
  int lgL = 0;
  bool row = false;
  int xv = 0;  // fiber passes: XopVariant covering every field's x-space op
  Geom geom{};
  size_t arg_off = 0;  // into FieldArgs / DetArgs arrays
  int nf = 0;
  // copy / fill / saveU
  void *dst = nullptr;
  const void *src = nullptr;
  int64_t n = 0;
  int member = 0;
  int step = 0;
  bool keypoint = false;  // SaveU: a keypoint (the member's keypoint_file), not a u vector
};

// pending state of one lockstep group
struct PropAcc {
  bool on = false;
  double alpha = 0, beta = 0;
  int64_t rs = 0, re = 0;
  bool cut = false;
};

inline void prop_accumulate(PropAcc &a, const Op &op, double wl, int64_t N, double dx) {
  const double L = (double)N * dx;
  a.alpha += wl * op.dz / (2.0 * L * L);
  // global phase exactly as big-wave evaluates it in FP64,
  // np.exp(-1j * 2 * np.pi / wl * dz) (big-wave/propagation.py:126), then
  // reduced to revolutions in extended precision
  const double th = (-(2.0 * M_PI) / wl) * op.dz;
  const long double two_pi = 6.283185307179586476925286766559005768L;
  long double red = fmodl((long double)th, two_pi) / two_pi;
  double b = (double)(red - floorl(red));
  a.beta += b;
  a.beta -= std::floor(a.beta);
  int64_t rs_, re_;
  cutoff_range(N, op.fc, dx, &rs_, &re_);
  if (rs_ != re_) {
    if (!a.cut) {
      a.rs = rs_;
      a.re = re_;
      a.cut = true;
    } else {
      a.rs = std::min(a.rs, rs_);
      a.re = std::max(a.re, re_);
    }
  }
  a.on = true;
}

inline KOp make_kop(const PropAcc &a, int64_t N, bool fold_scale) {
  KOp k{};
  k.on = 1;
  k.alpha_rev = a.alpha;
  k.beta_rev = a.beta;
  k.rs = a.cut ? a.rs : 0;
  k.re = a.cut ? a.re : 0;
  k.scale[0] = fold_scale ? 1.0 / (double)N : 1.0;
  k.scale[1] = 0.0;
  const double w = 6.283185307179586 / (double)kCisTab;
  const double sc = k.scale[0];
  k.t_sc = (float)sc;
  k.t_s1 = (float)(sc * w);
  k.t_s3 = (float)(-sc * w * w * w / 6.0);
  k.t_c2 = (float)(-sc * w * w / 2.0);
  return k;
}

inline XOp scalar_xop(double re, double im) {
  XOp x{};
  x.kind = XOP_SCALAR;
  x.fa[0] = re;
  x.fa[1] = im;
  x.post[0] = 1.0;
  x.post[1] = 0.0;
  return x;
}

// x-space op composition: first `a`, then `b` (b is a MASK, DBA, SAMPLE or SCALAR)
inline XOp compose(const XOp &a, const XOp &b) {
  if (a.kind == XOP_NONE) return b;
  XOp r = a;
  double f[2];
  if (b.kind == XOP_SCALAR) {
    f[0] = b.fa[0] * b.post[0] - b.fa[1] * b.post[1];
    f[1] = b.fa[0] * b.post[1] + b.fa[1] * b.post[0];
    const double p0 = r.post[0] * f[0] - r.post[1] * f[1];
    const double p1 = r.post[0] * f[1] + r.post[1] * f[0];
    r.post[0] = p0;
    r.post[1] = p1;
    return r;
  }
  if ((a.kind == XOP_SCALAR || a.kind == XOP_ONES) &&
      (b.kind == XOP_MASK || b.kind == XOP_DBA || b.kind == XOP_SAMPLE)) {
    XOp m = b;
    // value of `a` at every sample: fa * post for a scalar, post for ones
    const double s0 = a.kind == XOP_ONES ? a.post[0] : a.fa[0] * a.post[0] - a.fa[1] * a.post[1];
    const double s1 = a.kind == XOP_ONES ? a.post[1] : a.fa[0] * a.post[1] + a.fa[1] * a.post[0];
    const double p0 = m.post[0] * s0 - m.post[1] * s1;
    const double p1 = m.post[0] * s1 + m.post[1] * s0;
    m.post[0] = p0;
    m.post[1] = p1;
    if (a.kind == XOP_ONES || a.from_ones) m.from_ones = 1;
    return m;
  }
  throw Error(6, "two x-space masks without a propagation in between");
}

// Prepare an x-space record for the fused passes: mask and scalar factors get
// the post factor folded in (the same FP64 products the generic kernel path
// forms per sample, so the cast factors are identical), and single-chunk
// masks get the constants of the fast test: t = fma(n, m_a, m_b) differs
// from big-wave's fl(fl(x)/pitch) + dc/2 by a few ulp of the largest term,
// (|x0| + 2 N |xstep|) / pitch, so m_guard (1e-13 of that plus 1) bounds it
// with a wide margin.
inline void finalize_xop(XOp &x, int64_t N) {
  if (x.kind != XOP_MASK && x.kind != XOP_DBA && x.kind != XOP_SCALAR) return;
  auto fold = [&](double *f) {
    const double r = f[0] * x.post[0] - f[1] * x.post[1];
    const double i = f[0] * x.post[1] + f[1] * x.post[0];
    f[0] = r;
    f[1] = i;
  };
  fold(x.fa);
  if (x.kind != XOP_SCALAR) fold(x.fb);
  x.post[0] = 1.0;
  x.post[1] = 0.0;
  if (x.kind == XOP_MASK && x.single_chunk) {
    x.m_a = x.xstep * x.inv_pitch;
    x.m_b = x.x0 * x.inv_pitch + x.half_dc;
    const double span = (std::fabs(x.x0) + 2.0 * (double)N * std::fabs(x.xstep)) * std::fabs(x.inv_pitch);
    x.m_guard = 1e-13 * (span + std::fabs(x.half_dc) + 1.0);
  }
}

template <typename S> struct Exec {
  Context &ctx;
  const RunConfig &rc;
  Plan &plan;
  std::vector<Member *> mem;          // this group
  std::vector<void *> field;          // per member
  std::vector<void *> snap;           // per member (phase stepping)
  std::vector<void *> hist;           // per member (history entries)
  std::vector<FieldArgs> fargs;
  std::vector<DetArgs> dargs;
  std::vector<Launch> launches;
  double *acc = nullptr;
  const int64_t *bounds = nullptr;
  int64_t npx = 0;
  hipStream_t st;  // stream the launches go to (the context stream by default)
  // three-group plans: the middle, row and middle passes of a propagation run
  // as `spans` consecutive sub-ranges of the fields, each small enough to stay
  // in the last-level cache from its first pass to its last
  int spans = 1;
  // fields larger than the last-level cache: the top column passes read them
  // with non-temporal loads (Geom::nt)
  bool stream_top = false;

  Exec(Context &c, const RunConfig &r, Plan &p) : ctx(c), rc(r), plan(p), st(c.stream) {}

  size_t push_args(const std::vector<FieldArgs> &v) {
    const size_t off = fargs.size();
    fargs.insert(fargs.end(), v.begin(), v.end());
    return off;
  }

  void fiber(int lgL, bool row, const Geom &g, std::vector<FieldArgs> a) {
    for (auto &r : a) finalize_xop(r.x, rc.N);
    Launch l{LKind::Fiber};
    l.lgL = lgL;
    l.row = row;
    l.xv = xop_variant(a.data(), (int)a.size(), g.top != 0);
    l.geom = g;
    l.geom.nt = stream_top && !row && g.top ? 1 : 0;
    l.nf = (int)a.size();
    l.arg_off = push_args(a);
    launches.push_back(l);
  }

  std::vector<FieldArgs> blank() const {
    std::vector<FieldArgs> a(mem.size());
    for (size_t i = 0; i < mem.size(); ++i) {
      std::memset(&a[i], 0, sizeof(FieldArgs));
      a[i].data = field[i];
    }
    return a;
  }

  // fused propagation from state X (transformed == false) or T
  void emit_prop_fused(bool transformed, const std::vector<XOp> &px, const std::vector<PropAcc> &pa) {
    const int K = (int)mem.size();
    if (plan.ngroups == 1) {
      auto a = blank();
      for (int i = 0; i < K; ++i) {
        a[i].x = px[i];
        a[i].do_fwd = 1;
        a[i].do_inv = 1;
        a[i].k = make_kop(pa[i], rc.N, true);
      }
      fiber(plan.lgR, true, plan.geom_row(), a);
      return;
    }
    {
      auto a = blank();
      for (int i = 0; i < K; ++i) {
        a[i].x = px[i];
        a[i].do_inv = transformed;
        a[i].do_fwd = 1;
      }
      fiber(plan.lgA, false, plan.geom_top(), a);
    }
    const int ns = plan.ngroups == 3 ? spans : 1;
    const int64_t len = rc.N / ns;
    for (int c = 0; c < ns; ++c) {
      Geom gm = plan.geom_mid(), gr = plan.geom_row();
      if (ns > 1) {
        gm.span_off = gr.span_off = c * len;
        gm.span_len = gr.span_len = len;
      }
      if (plan.ngroups == 3) {
        auto a = blank();
        for (int i = 0; i < K; ++i) a[i].do_fwd = 1;
        fiber(plan.lgB, false, gm, a);
      }
      {
        auto a = blank();
        for (int i = 0; i < K; ++i) {
          a[i].do_fwd = 1;
          a[i].do_inv = 1;
          a[i].k = make_kop(pa[i], rc.N, true);
        }
        fiber(plan.lgR, true, gr, a);
      }
      if (plan.ngroups == 3) {
        auto a = blank();
        for (int i = 0; i < K; ++i) a[i].do_inv = 1;
        fiber(plan.lgB, false, gm, a);
      }
    }
  }

  // one unfused FFT direction (plain form)
  void emit_fft_plain(bool inverse) {
    const int K = (int)mem.size();
    auto mk = [&](bool fwd, bool inv) {
      auto a = blank();
      for (int i = 0; i < K; ++i) { a[i].do_fwd = fwd; a[i].do_inv = inv; }
      return a;
    };
    Geom row = plan.geom_row();
    row.one_dir = 1;
    if (plan.ngroups == 1) {
      fiber(plan.lgR, true, row, mk(!inverse, inverse));
      return;
    }
    Geom top = plan.geom_top();
    top.top = 0;  // no x-space op inside an FFT pass of the plain form
    if (!inverse) {
      fiber(plan.lgA, false, top, mk(true, false));
      if (plan.ngroups == 3) fiber(plan.lgB, false, plan.geom_mid(), mk(true, false));
      fiber(plan.lgR, true, row, mk(true, false));
    } else {
      fiber(plan.lgR, true, row, mk(false, true));
      if (plan.ngroups == 3) fiber(plan.lgB, false, plan.geom_mid(), mk(false, true));
      fiber(plan.lgA, false, top, mk(false, true));
    }
  }

  void emit_xop(const std::vector<XOp> &px) {
    auto a = blank();
    for (size_t i = 0; i < mem.size(); ++i) a[i].x = px[i];
    Launch l{LKind::Xop};
    l.nf = (int)a.size();
    l.arg_off = push_args(a);
    launches.push_back(l);
  }

  void emit_kop(const std::vector<KOp> &ks) {
    auto a = blank();
    for (size_t i = 0; i < mem.size(); ++i) a[i].k = ks[i];
    Launch l{LKind::Kop};
    l.geom = plan.geom_row();
    l.nf = (int)a.size();
    l.arg_off = push_args(a);
    launches.push_back(l);
  }

  // flush a transformed field back to natural order (fused form)
  void emit_to_natural(const std::vector<XOp> &px) {
    auto a = blank();
    for (size_t i = 0; i < mem.size(); ++i) {
      a[i].x = px[i];
      a[i].do_inv = 1;
    }
    fiber(plan.lgA, false, plan.geom_top(), a);
  }

  void compile() {
    const int K = (int)mem.size();
    const size_t nops = mem[0]->ops.size();
    for (int i = 1; i < K; ++i)
      if (mem[i]->ops.size() != nops) throw Error(1, "lockstep members differ in length");
    std::vector<XOp> px(K);
    for (auto &x : px) std::memset(&x, 0, sizeof(XOp));
    std::vector<PropAcc> pa(K);
    bool transformed = false;
    const bool plain = rc.mode == 1;
    bool snap_transformed = false;
    bool hist_transformed = false;

    auto clear_px = [&]() {
      for (auto &x : px) std::memset(&x, 0, sizeof(XOp));
    };
    auto flush_natural = [&]() {
      bool any = false;
      for (auto &x : px) any |= x.kind != XOP_NONE;
      if (transformed) {
        emit_to_natural(px);
        transformed = false;
      } else if (any) {
        emit_xop(px);
      }
      clear_px();
    };

    for (size_t o = 0; o < nops; ++o) {
      const OpKind kind = mem[0]->ops[o].kind;
      for (int i = 1; i < K; ++i)
        if (mem[i]->ops[o].kind != kind) throw Error(1, "lockstep members differ in op kinds");
      switch (kind) {
        case OpKind::Source:
          for (int i = 0; i < K; ++i) px[i] = mem[i]->ops[o].x;
          transformed = false;
          break;
        case OpKind::Upload: {
          for (int i = 0; i < K; ++i) {
            Launch l{LKind::Upload};
            l.dst = field[i];
            l.n = rc.N;
            l.member = i;
            launches.push_back(l);
          }
          clear_px();
          transformed = false;
          break;
        }
        case OpKind::Mask:
          for (int i = 0; i < K; ++i) px[i] = compose(px[i], mem[i]->ops[o].x);
          break;
        case OpKind::Prop: {
          for (int i = 0; i < K; ++i) prop_accumulate(pa[i], mem[i]->ops[o], mem[i]->wl, rc.N, rc.dx);
          const bool next_prop = (o + 1 < nops) && mem[0]->ops[o + 1].kind == OpKind::Prop;
          if (next_prop && !plain) break;  // merge consecutive propagations
          if (plain) {
            bool any = false;
            for (auto &x : px) any |= x.kind != XOP_NONE;
            if (any) emit_xop(px);
            emit_fft_plain(false);
            std::vector<KOp> ks(K);
            for (int i = 0; i < K; ++i) ks[i] = make_kop(pa[i], rc.N, false);
            emit_kop(ks);
            emit_fft_plain(true);
            std::vector<XOp> sc(K, scalar_xop(1.0 / (double)rc.N, 0.0));
            emit_xop(sc);
          } else {
            emit_prop_fused(transformed, px, pa);
            transformed = plan.ngroups > 1;
          }
          clear_px();
          for (auto &a : pa) a = PropAcc{};
          break;
        }
        case OpKind::Detect: {
          flush_natural();
          const size_t off = dargs.size();
          for (int i = 0; i < K; ++i) {
            DetArgs d{};
            d.data = field[i];
            d.acc = acc + (size_t)(mem[i]->out_base + mem[i]->ops[o].row) * npx;
            d.weight = mem[i]->weight;
            d.accumulate = 1;
            dargs.push_back(d);
          }
          Launch l{LKind::Detect};
          l.nf = K;
          l.arg_off = off;
          launches.push_back(l);
          break;
        }
        case OpKind::Flush:
          flush_natural();
          break;
        case OpKind::Snapshot:
        case OpKind::HSave: {
          // the saved state is the in-flight one (natural or column-transformed)
          bool any = false;
          for (auto &x : px) any |= x.kind != XOP_NONE;
          if (any) flush_natural();
          const bool h = kind == OpKind::HSave;
          const std::vector<void *> &to = h ? hist : snap;
          if (h && to.empty()) throw Error(1, "history buffer missing");
          for (size_t i = 0; i < to.size(); ++i) {
            Launch l{LKind::Copy};
            l.dst = to[i];
            l.src = field[i];
            l.n = rc.N;
            launches.push_back(l);
          }
          (h ? hist_transformed : snap_transformed) = transformed;
          break;
        }
        case OpKind::Restore:
        case OpKind::HRestore: {
          const bool h = kind == OpKind::HRestore;
          for (int i = 0; i < K; ++i) {
            Launch l{LKind::Copy};
            l.dst = field[i];
            l.src = h ? hist[i] : snap[i];
            l.n = rc.N;
            launches.push_back(l);
          }
          clear_px();
          transformed = h ? hist_transformed : snap_transformed;
          break;
        }
        case OpKind::SaveU:
        case OpKind::Keypoint: {
          flush_natural();
          for (int i = 0; i < K; ++i) {
            Launch l{LKind::SaveU};
            l.src = field[i];
            l.n = rc.N;
            l.member = i;
            l.step = mem[i]->ops[o].row;
            l.keypoint = kind == OpKind::Keypoint;
            launches.push_back(l);
          }
          break;
        }
      }
    }
  }

  // per-launch event timing records of the enqueued launches
  std::vector<hipEvent_t> ev;
  std::vector<std::pair<int, double>> ev_meta;  // class, bytes

  // Enqueue the launches on the group's stream with argument records already
  // on the device (fa[i] = fargs[i], da[i] = dargs[i]); returns without
  // waiting unless a host download needs it.  Under a whole-run graph
  // capture (rc.whole_graph) only device work may be enqueued.
  void enqueue(const FieldArgs *fa, const DetArgs *da) {
    const size_t esz = 2 * sizeof(S);
    auto host_step = [&]() {
      if (rc.whole_graph) throw Error(6, "file uploads and downloads cannot run inside a captured graph");
    };

    // optional per-launch event timing (not combined with graph capture).
    // Timing only: the events carry no system-scope fence (cache write-back)
    // and are read after the stream has been synchronised.  Fiber passes
    // hand them to the launch, whose dispatch records the kernel's own start
    // and end; other kernels get event records around the launch.
    const bool timed = ctx.timing && !rc.use_graph;
    auto sampled = [&](int cls) { return timed && ctx.timing_seen[cls]++ % ctx.timing_stride == 0; };
    auto t_new = [&](int cls, double bytes) {
      hipEvent_t a, b;
      hip_check(hipEventCreateWithFlags(&a, hipEventDisableSystemFence), "event");
      hip_check(hipEventCreateWithFlags(&b, hipEventDisableSystemFence), "event");
      ev.push_back(a);
      ev.push_back(b);
      ev_meta.emplace_back(cls, bytes);
    };
    bool open = false;
    auto t_begin = [&](int cls, double bytes) {
      open = sampled(cls);
      if (!open) return;
      t_new(cls, bytes);
      hip_check(hipEventRecord(ev[ev.size() - 2], st), "event record");
    };
    auto t_end = [&]() {
      if (open) hip_check(hipEventRecord(ev.back(), st), "event record");
      open = false;
    };

    for (const Launch &l : launches) {
      switch (l.kind) {
        case LKind::Fiber: {
          const int cls = l.row ? PC_ROW : (l.geom.top ? PC_TOP : PC_MID);
          const double ns = (double)(l.geom.span_len ? l.geom.span_len : rc.N);
          hipEvent_t e0 = nullptr, e1 = nullptr;
          if (sampled(cls)) {
            t_new(cls, 2.0 * ns * (double)esz * l.nf);
            e0 = ev[ev.size() - 2];
            e1 = ev.back();
          }
          hip_check(launch_fiber_pass<S>(l.lgL, l.row, l.xv, fa + l.arg_off, l.nf, l.geom, st, e0, e1),
                    "fiber pass");
          break;
        }
        case LKind::Xop:
          t_begin(PC_OTHER, 2.0 * (double)rc.N * (double)esz * l.nf);
          hip_check(launch_xop<S>(fa + l.arg_off, l.nf, rc.N, st), "xop");
          t_end();
          break;
        case LKind::Kop:
          t_begin(PC_OTHER, 2.0 * (double)rc.N * (double)esz * l.nf);
          hip_check(launch_kop<S>(fa + l.arg_off, l.nf, l.geom, st), "kop");
          t_end();
          break;
        case LKind::Detect:
          hip_check(launch_detect<S>(da + l.arg_off, l.nf, bounds, npx, st), "detect");
          break;
        case LKind::Copy:
          hip_check(launch_copy<S>(l.dst, l.src, l.n, st), "copy");
          break;
        case LKind::Upload: {
          host_step();
          const Member &m = *mem[l.member];
          if (m.upload_npy.empty()) throw Error(1, "upload op without a source file");
          ctx.stream_npy(m, sizeof(S) == 8, l.n, l.dst, st);
          break;
        }
        case LKind::Fill:
          hip_check(launch_fill_f64((double *)l.dst, l.n, 0.0, st), "fill");
          break;
        case LKind::SaveU: {
          host_step();
          Member *m = mem[l.member];
          const auto &file = l.keypoint ? m->keypoint_file : m->saveu_file;
          if (file) ctx.stream_to_npy(l.src, l.n, sizeof(S) == 8, file(l.step), st);
          break;
        }
      }
    }
  }

  // after the stream has drained: add the timed launches to ctx.stats
  void collect_timing() {
    for (size_t i = 0; i < ev_meta.size(); ++i) {
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]), "event elapsed");
      PassStat &p = ctx.stats[ev_meta[i].first];
      p.launches += 1;
      p.ms += ms;
      p.bytes += ev_meta[i].second;
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    ev.clear();
    ev_meta.clear();
  }
};


// Grating masks shared by several top column passes of one run (the slices of
// a grating with constant duty cycle, the members of a batch): their
// decisions are computed once per run by mask_fill_kernel and read as bits by
// the passes.  Masks used by a single pass keep the inline test.
template <typename S> struct MaskCache {
  struct Entry {
    XOp key;
    int lgL;
    Geom g;
    std::vector<FieldArgs *> recs;
    DevBuf bits;
  };
  std::vector<Entry> entries;

  static bool same(const XOp &a, const XOp &b) {
    const double ka[8] = {a.m_a, a.m_b, a.m_guard, a.dc, a.pitch, a.half_dc, a.x0, a.xstep};
    const double kb[8] = {b.m_a, b.m_b, b.m_guard, b.dc, b.pitch, b.half_dc, b.x0, b.xstep};
    return std::memcmp(ka, kb, sizeof ka) == 0;
  }

  void collect(Exec<S> &ex) {
    for (const Launch &l : ex.launches) {
      if (l.kind != LKind::Fiber || l.row || l.xv != XV_MASK) continue;
      for (int i = 0; i < l.nf; ++i) {
        FieldArgs &a = ex.fargs[l.arg_off + i];
        if (a.x.kind != XOP_MASK || !a.x.single_chunk) continue;
        Entry *e = nullptr;
        for (auto &c : entries)
          if (c.lgL == l.lgL && same(c.key, a.x)) {
            e = &c;
            break;
          }
        if (!e) {
          entries.push_back(Entry{a.x, l.lgL, l.geom, {}, DevBuf()});
          e = &entries.back();
        }
        e->recs.push_back(&a);
      }
    }
  }

  // fill the masks used more than once on stream st and point their records
  // at them (before the records are uploaded)
  void fill(int64_t N, hipStream_t st) {
    for (auto &e : entries) {
      if (e.recs.size() < 2) continue;
      e.bits = DevBuf((size_t)N / 8);
      hip_check(launch_mask_fill<S>(e.lgL, e.key, e.g, e.bits.p, st), "mask fill");
      for (FieldArgs *r : e.recs) r->x.mbits = e.bits.p;
    }
  }
};

}  // namespace exec
}  // namespace rs
