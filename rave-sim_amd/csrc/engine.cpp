// Host engine: exact big-wave helper arithmetic, FFT plans, and the lockstep
// executor that compiles member programs into fused gfx950 launches.
#include "engine.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <sstream>

#include "rs_kernels.hpp"

#pragma STDC FP_CONTRACT OFF

namespace rs {

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    const int code = (e == hipErrorOutOfMemory) ? 5 : 4;
    throw Error(code, std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
  }
}

// ----------------------------------------------------------------------------
// exact helpers
// ----------------------------------------------------------------------------

// CPython's float floor division (Objects/floatobject.c float_floor_div)
static double py_floordiv(double vx, double wx) {
  double mod = std::fmod(vx, wx);
  double div = (vx - mod) / wx;
  if (mod != 0.0) {
    if ((wx < 0) != (mod < 0)) {
      mod += wx;
      div -= 1.0;
    }
  }
  double floordiv;
  if (div != 0.0) {
    floordiv = std::floor(div);
    if (div - floordiv > 0.5) floordiv += 1.0;
  } else {
    floordiv = std::copysign(0.0, vx / wx);
  }
  return floordiv;
}

void cutoff_range(int64_t N, double fc, double dx, int64_t *rs_, int64_t *re_) {
  // big-wave/propagation.py:83-90
  const double frac = fc * dx;
  const double start = std::ceil((double)N * frac);
  if (!(start <= (double)N / 2.0)) {  // also true for NaN / inf
    *rs_ = 0;
    *re_ = 0;
    return;
  }
  const int64_t s = (int64_t)start;
  *rs_ = s;
  *re_ = N - s + 1;
}

int64_t detector_pixels(double size, double px) {
  if (!(px > 0) || !(size > 0)) return 0;
  return (int64_t)py_floordiv(size, px);
}

std::vector<int64_t> detector_bounds(int64_t N, double dx, double size, double px) {
  const int64_t o = detector_pixels(size, px);
  std::vector<int64_t> b(o + 1, 0);
  if (o <= 0) return b;
  const double first = (-((double)o / 2.0)) * px;
  const double last = ((double)o / 2.0 + 1.0) * px;
  const int64_t nb = o + 1;
  // np.linspace(first, last, nb + 1, endpoint=True)
  std::vector<double> edges(nb + 1);
  const double delta = last - first;
  const double step = delta / (double)nb;
  for (int64_t i = 0; i <= nb; ++i) edges[i] = (double)i * step + first;
  edges[nb] = last;
  const double denom = last - first;
  auto binval = [&](int64_t j) -> int64_t {
    const double x = ((double)j - (double)N / 2.0) * dx;
    if (x < first) return -1;
    if (x > last) return nb + 1;
    const double f = ((x - first) / denom) * (double)nb;
    int64_t idx = (int64_t)f;
    if (idx == nb) idx -= 1;
    if (x < edges[idx]) idx -= 1;
    if (x >= edges[idx + 1] && idx != nb - 1) idx += 1;
    return idx;
  };
  for (int64_t k = 0; k <= o; ++k) {
    int64_t lo = 0, hi = N;  // first j in [0, N] with binval(j) >= k
    while (lo < hi) {
      const int64_t mid = lo + (hi - lo) / 2;
      if (binval(mid) >= k) hi = mid;
      else lo = mid + 1;
    }
    b[k] = lo;
  }
  return b;
}

std::vector<double> detector_scale(double size, double px, double py, double dx, double z) {
  // returned as pairs (cos(atan(x/z)), sqrt(x^2 + z^2)) so that the caller
  // applies them in big-wave's order: ((out * cos) * dx * py) / sqrt(...)
  const double n = py_floordiv(size, px);
  const double h = py_floordiv(n, 2.0);
  const int64_t o = (int64_t)n;
  std::vector<double> s(2 * std::max<int64_t>(o, 0));
  for (int64_t k = 0; k < o; ++k) {
    const double x = ((double)k - h) * px;
    const double ang = std::atan(x / z);
    s[2 * k] = std::cos(ang);
    s[2 * k + 1] = std::sqrt(x * x + z * z);
  }
  (void)dx;
  (void)py;
  return s;
}

void material_factor(double db_re, double db_im, double t, double wl, double out[2]) {
  // np.exp(2j * np.pi * t / wl * db), big-wave/optical_element.py:36-39
  const double a = (2.0 * M_PI * t) / wl;
  const double re = -(a * db_im) + 0.0 * db_re;
  const double im = a * db_re + 0.0 * db_im;
  const double e = std::exp(re);
  out[0] = e * std::cos(im);
  out[1] = e * std::sin(im);
}

double energy_to_wavelength(double e) {
  // big-wave/propagation.py:26-28, 69-71 (constants built as Python does)
  const double h = 6.62607004 * std::pow(10.0, -34.0);
  const double c0 = 299792458.0;
  const double ev = 1.602176634 * std::pow(10.0, -19.0);
  return h * c0 / (e * ev);
}

// ----------------------------------------------------------------------------
// device buffers / plans / context
// ----------------------------------------------------------------------------

DevBuf::DevBuf(size_t b) : bytes(b) {
  if (b) hip_check(hipMalloc(&p, b), "hipMalloc");
}
DevBuf &DevBuf::operator=(DevBuf &&o) noexcept {
  if (this != &o) {
    if (p) (void)hipFree(p);
    p = o.p;
    bytes = o.bytes;
    o.p = nullptr;
    o.bytes = 0;
  }
  return *this;
}
DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}

Plan::~Plan() {
  for (void *q : fib) if (q) (void)hipFree(q);
  for (void *q : hi) if (q) (void)hipFree(q);
  for (void *q : lo) if (q) (void)hipFree(q);
}

void plan_split(int lgN, bool f64, int lgR_req, int lgA_req, int *ng, int *lgA, int *lgB, int *lgR) {
  const int maxR = f64 ? kMaxRowLgF64 : kMaxRowLgF32;
  if (lgN < kMinRowLg || lgN > 30)
    throw Error(6, "N must be a power of two between 2^6 and 2^30");
  if (lgN <= maxR && (lgR_req < 0 || lgR_req >= lgN)) {
    *ng = 1; *lgA = *lgR = lgN; *lgB = 0;
    return;
  }
  int r = lgR_req >= 0 ? std::min(lgR_req, maxR) : maxR;
  r = std::max(r, kMinRowLg);
  if (lgN - r < kMinColLg) r = lgN - kMinColLg;
  if (r < kMinRowLg) throw Error(6, "no digit-group split for this N");
  const int rest = lgN - r;
  if (rest <= kMaxColLg && (lgA_req < 0 || lgA_req == rest)) {
    *ng = 2; *lgR = r; *lgA = rest; *lgB = 0;
    return;
  }
  const int maxW = f64 ? kMaxWideColLgF64 : kMaxWideColLgF32;
  if (lgA_req >= 0 && lgA_req == rest && rest <= maxW) {  // wide two-group split
    *ng = 2; *lgR = r; *lgA = rest; *lgB = 0;
    return;
  }
  int a = lgA_req >= 0 ? lgA_req : (rest + 1) / 2;
  int b = rest - a;
  if (a > maxW || a < kMinColLg || b < kMinColLg || b > maxW)
    throw Error(6, "no three-group split for this N");
  *ng = 3; *lgR = r; *lgA = a; *lgB = b;
}

std::vector<std::pair<int, int>> plan_candidates(int lgN, bool f64) {
  std::vector<std::pair<int, int>> out;
  const int maxR = f64 ? kMaxRowLgF64 : kMaxRowLgF32;
  if (lgN <= maxR) out.emplace_back(lgN, lgN);
  const int maxW = f64 ? kMaxWideColLgF64 : kMaxWideColLgF32;
  for (int r = kMinRowLg; r <= std::min(maxR, lgN - kMinColLg); ++r) {
    const int rest = lgN - r;
    if (rest <= maxW) out.emplace_back(r, rest);
    for (int a = kMinColLg; a <= kMaxColLg; ++a) {
      const int b = rest - a;
      if (b >= kMinColLg && b <= kMaxColLg) out.emplace_back(r, a);
    }
  }
  return out;
}

template <typename S>
static void *upload_table(const std::vector<long double> &re, const std::vector<long double> &im) {
  std::vector<S> h(2 * re.size());
  for (size_t i = 0; i < re.size(); ++i) {
    h[2 * i] = (S)re[i];
    h[2 * i + 1] = (S)im[i];
  }
  void *d = nullptr;
  hip_check(hipMalloc(&d, h.size() * sizeof(S)), "hipMalloc(twiddles)");
  hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(S), hipMemcpyHostToDevice), "twiddle upload");
  return d;
}

// W_M^(e * mult) for e < count, forward sign exp(-2 pi i e mult / M)
static void *make_twiddles(int lgM, int64_t count, int64_t mult, bool f64) {
  std::vector<long double> re(count), im(count);
  const long double M = std::ldexp(1.0L, lgM);
  const int64_t mask = (lgM >= 63) ? -1 : ((int64_t(1) << lgM) - 1);
  for (int64_t e = 0; e < count; ++e) {
    const int64_t ex = (e * mult) & mask;
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)ex / M;
    re[e] = cosl(ang);
    im[e] = sinl(ang);
  }
  return f64 ? upload_table<double>(re, im) : upload_table<float>(re, im);
}

std::unique_ptr<Plan> make_plan(int64_t N, bool f64, int lgR_req, int lgA_req) {
  int lgN = 0;
  while ((int64_t(1) << lgN) < N) ++lgN;
  if ((int64_t(1) << lgN) != N) throw Error(6, "N must be a power of two");
  auto p = std::make_unique<Plan>();
  p->N = N;
  p->lgN = lgN;
  p->f64 = f64;
  plan_split(lgN, f64, lgR_req, lgA_req, &p->ngroups, &p->lgA, &p->lgB, &p->lgR);
  // in-fiber tables: row, top, middle
  p->fib[0] = make_twiddles(p->lgR, int64_t(1) << p->lgR, 1, f64);
  if (p->ngroups >= 2) p->fib[1] = make_twiddles(p->lgA, int64_t(1) << p->lgA, 1, f64);
  if (p->ngroups == 3) p->fib[2] = make_twiddles(p->lgB, int64_t(1) << p->lgB, 1, f64);
  // inter-group tables: top column M = N, middle column M = 2^(lgB + lgR)
  for (int t = 0; t < p->ngroups - 1; ++t) {
    const int lgM = t == 0 ? lgN : p->lgB + p->lgR;
    const int s = (lgM + 1) / 2;
    p->sh[t] = s;
    p->lo[t] = make_twiddles(lgM, int64_t(1) << s, 1, f64);
    p->hi[t] = make_twiddles(lgM, int64_t(1) << (lgM - s), int64_t(1) << s, f64);
  }
  return p;
}

Geom Plan::geom_row() const {
  Geom g{};
  g.xcd_segs = 1;
  g.N = N;
  g.S = 1;
  g.top = ngroups == 1;
  g.ngroups = ngroups;
  g.lgA = lgA;
  g.lgB = lgB;
  g.lgR = lgR;
  g.tw_fib = fib[0];
  return g;
}
static int32_t line_sharing(int lgL, bool f64) {
  const int bytes = col_fibers(1 << lgL, f64) * (f64 ? 16 : 8);
  return bytes >= 128 ? 1 : 128 / bytes;
}

Geom Plan::geom_top() const {
  Geom g = geom_row();
  g.top = 1;
  g.xcd_segs = line_sharing(lgA, f64);
  g.S = int64_t(1) << (lgB + lgR);
  g.tw_fib = fib[1];
  g.tw_sh = sh[0];
  g.tw_hi = hi[0];
  g.tw_lo = lo[0];
  return g;
}
Geom Plan::geom_mid() const {
  Geom g = geom_row();
  g.top = 0;
  g.xcd_segs = line_sharing(lgB, f64);
  g.S = int64_t(1) << lgR;
  g.tw_fib = fib[2];
  g.tw_sh = sh[1];
  g.tw_hi = hi[1];
  g.tw_lo = lo[1];
  return g;
}
std::string Plan::describe() const {
  std::ostringstream s;
  s << "N=2^" << lgN << (f64 ? " c16" : " c8") << " groups=" << ngroups;
  if (ngroups == 1) s << " [row 2^" << lgR << "]";
  else if (ngroups == 2) s << " [col 2^" << lgA << " | row 2^" << lgR << "]";
  else s << " [col 2^" << lgA << " | col 2^" << lgB << " | row 2^" << lgR << "]";
  return s.str();
}

Context::Context(int dev) : device(dev) {
  hip_check(hipSetDevice(dev), "hipSetDevice");
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
}
Context::~Context() {
  plans_.clear();
  if (stream) (void)hipStreamDestroy(stream);
}
const Context::Tuned *Context::find_tuned(int lgN, bool f64) const {
  for (const auto &t : tuned)
    if (t.lgN == lgN && t.f64 == f64) return &t;
  return nullptr;
}

Plan &Context::plan(int64_t N, bool f64, int lgR, int lgA) {
  if (lgR < 0 && lgA < 0) {
    int lgN = 0;
    while ((int64_t(1) << lgN) < N) ++lgN;
    if (const Tuned *t = find_tuned(lgN, f64)) {
      lgR = t->lgR;
      lgA = t->lgA;
    }
  }
  for (size_t i = 0; i < plans_.size(); ++i)
    if (plans_[i]->N == N && plans_[i]->f64 == f64 && plan_req_[i] == std::make_pair(lgR, lgA))
      return *plans_[i];
  plans_.push_back(make_plan(N, f64, lgR, lgA));
  plan_req_.emplace_back(lgR, lgA);
  return *plans_.back();
}
void Context::sync() {
  hip_check(hipSetDevice(device), "hipSetDevice");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

// ----------------------------------------------------------------------------
// lockstep executor
// ----------------------------------------------------------------------------

namespace {

enum class LKind { Fiber, Xop, Kop, Detect, Copy, Fill, SaveU };

struct Launch {
  LKind kind;
  int lgL = 0;
  bool row = false;
  Geom geom{};
  size_t arg_off = 0;  // into FieldArgs / DetArgs arrays
  int nf = 0;
  // copy / fill / saveU
  void *dst = nullptr;
  const void *src = nullptr;
  int64_t n = 0;
  int member = 0;
  int step = 0;
};

// pending state of one lockstep group
struct PropAcc {
  bool on = false;
  double alpha = 0, beta = 0;
  int64_t rs = 0, re = 0;
  bool cut = false;
};

void prop_accumulate(PropAcc &a, const Op &op, double wl, int64_t N, double dx) {
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

KOp make_kop(const PropAcc &a, int64_t N, bool fold_scale) {
  KOp k{};
  k.on = 1;
  k.alpha_rev = a.alpha;
  k.beta_rev = a.beta;
  k.rs = a.cut ? a.rs : 0;
  k.re = a.cut ? a.re : 0;
  k.scale[0] = fold_scale ? 1.0 / (double)N : 1.0;
  k.scale[1] = 0.0;
  return k;
}

XOp scalar_xop(double re, double im) {
  XOp x{};
  x.kind = XOP_SCALAR;
  x.fa[0] = re;
  x.fa[1] = im;
  x.post[0] = 1.0;
  x.post[1] = 0.0;
  return x;
}

// x-space op composition: first `a`, then `b` (b is a MASK or SCALAR)
XOp compose(const XOp &a, const XOp &b) {
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
  if ((a.kind == XOP_SCALAR || a.kind == XOP_ONES) && b.kind == XOP_MASK) {
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

template <typename S> struct Exec {
  Context &ctx;
  const RunConfig &rc;
  Plan &plan;
  std::vector<Member *> mem;          // this group
  std::vector<void *> field;          // per member
  std::vector<void *> snap;           // per member (phase stepping)
  std::vector<FieldArgs> fargs;
  std::vector<DetArgs> dargs;
  std::vector<Launch> launches;
  double *acc = nullptr;
  const int64_t *bounds = nullptr;
  int64_t npx = 0;

  Exec(Context &c, const RunConfig &r, Plan &p) : ctx(c), rc(r), plan(p) {}

  size_t push_args(const std::vector<FieldArgs> &v) {
    const size_t off = fargs.size();
    fargs.insert(fargs.end(), v.begin(), v.end());
    return off;
  }

  void fiber(int lgL, bool row, const Geom &g, const std::vector<FieldArgs> &a) {
    Launch l{LKind::Fiber};
    l.lgL = lgL;
    l.row = row;
    l.geom = g;
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
    if (plan.ngroups == 3) {
      auto a = blank();
      for (int i = 0; i < K; ++i) a[i].do_fwd = 1;
      fiber(plan.lgB, false, plan.geom_mid(), a);
    }
    {
      auto a = blank();
      for (int i = 0; i < K; ++i) {
        a[i].do_fwd = 1;
        a[i].do_inv = 1;
        a[i].k = make_kop(pa[i], rc.N, true);
      }
      fiber(plan.lgR, true, plan.geom_row(), a);
    }
    if (plan.ngroups == 3) {
      auto a = blank();
      for (int i = 0; i < K; ++i) a[i].do_inv = 1;
      fiber(plan.lgB, false, plan.geom_mid(), a);
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
    if (plan.ngroups == 1) {
      fiber(plan.lgR, true, plan.geom_row(), mk(!inverse, inverse));
      return;
    }
    Geom top = plan.geom_top();
    top.top = 0;  // no x-space op inside an FFT pass of the plain form
    if (!inverse) {
      fiber(plan.lgA, false, top, mk(true, false));
      if (plan.ngroups == 3) fiber(plan.lgB, false, plan.geom_mid(), mk(true, false));
      fiber(plan.lgR, true, plan.geom_row(), mk(true, false));
    } else {
      fiber(plan.lgR, true, plan.geom_row(), mk(false, true));
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
            Launch l{LKind::Copy};
            l.dst = field[i];
            l.src = mem[i]->upload;
            l.n = rc.N;
            l.member = -1;  // host to device
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
        case OpKind::Snapshot: {
          bool any = false;
          for (auto &x : px) any |= x.kind != XOP_NONE;
          if (any) flush_natural();
          if (!snap.empty()) {
            for (int i = 0; i < K; ++i) {
              Launch l{LKind::Copy};
              l.dst = snap[i];
              l.src = field[i];
              l.n = rc.N;
              launches.push_back(l);
            }
          }
          snap_transformed = transformed;
          break;
        }
        case OpKind::Restore:
          for (int i = 0; i < K; ++i) {
            Launch l{LKind::Copy};
            l.dst = field[i];
            l.src = snap[i];
            l.n = rc.N;
            launches.push_back(l);
          }
          clear_px();
          transformed = snap_transformed;
          break;
        case OpKind::SaveU: {
          flush_natural();
          for (int i = 0; i < K; ++i) {
            Launch l{LKind::SaveU};
            l.src = field[i];
            l.n = rc.N;
            l.member = i;
            l.step = mem[i]->ops[o].row;
            launches.push_back(l);
          }
          break;
        }
      }
    }
  }

  void execute() {
    hipStream_t st = ctx.stream;
    DevBuf dfa(std::max<size_t>(fargs.size(), 1) * sizeof(FieldArgs));
    DevBuf dda(std::max<size_t>(dargs.size(), 1) * sizeof(DetArgs));
    if (!fargs.empty())
      hip_check(hipMemcpyAsync(dfa.p, fargs.data(), fargs.size() * sizeof(FieldArgs), hipMemcpyHostToDevice, st),
                "args upload");
    if (!dargs.empty())
      hip_check(hipMemcpyAsync(dda.p, dargs.data(), dargs.size() * sizeof(DetArgs), hipMemcpyHostToDevice, st),
                "det args upload");
    const FieldArgs *fa = (const FieldArgs *)dfa.p;
    const DetArgs *da = (const DetArgs *)dda.p;
    const size_t esz = 2 * sizeof(S);

    // graph capture covers the run of launches between host interactions
    bool capturing = false;
    hipGraph_t graph = nullptr;
    auto begin_capture = [&]() {
      if (rc.use_graph && !capturing) {
        hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
        capturing = true;
      }
    };
    auto end_capture = [&]() {
      if (capturing) {
        hip_check(hipStreamEndCapture(st, &graph), "end capture");
        capturing = false;
        hipGraphExec_t ex = nullptr;
        hip_check(hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0), "graph instantiate");
        hip_check(hipGraphLaunch(ex, st), "graph launch");
        hip_check(hipStreamSynchronize(st), "graph sync");
        (void)hipGraphExecDestroy(ex);
        (void)hipGraphDestroy(graph);
        graph = nullptr;
      }
    };

    // optional per-launch event timing (not combined with graph capture)
    const bool timed = ctx.timing && !rc.use_graph;
    std::vector<hipEvent_t> ev;
    std::vector<std::pair<int, double>> ev_meta;  // class, bytes
    auto t_begin = [&](int cls, double bytes) {
      if (!timed) return;
      hipEvent_t a, b;
      hip_check(hipEventCreate(&a), "event");
      hip_check(hipEventCreate(&b), "event");
      ev.push_back(a);
      ev.push_back(b);
      ev_meta.emplace_back(cls, bytes);
      hip_check(hipEventRecord(a, st), "event record");
    };
    auto t_end = [&]() {
      if (timed) hip_check(hipEventRecord(ev.back(), st), "event record");
    };

    for (const Launch &l : launches) {
      switch (l.kind) {
        case LKind::Fiber: {
          begin_capture();
          const int cls = l.row ? PC_ROW : (l.geom.top ? PC_TOP : PC_MID);
          t_begin(cls, 2.0 * (double)rc.N * (double)esz * l.nf);
          hip_check(launch_fiber_pass<S>(l.lgL, l.row, fa + l.arg_off, l.nf, l.geom, st), "fiber pass");
          t_end();
          break;
        }
        case LKind::Xop:
          begin_capture();
          t_begin(PC_OTHER, 2.0 * (double)rc.N * (double)esz * l.nf);
          hip_check(launch_xop<S>(fa + l.arg_off, l.nf, rc.N, st), "xop");
          t_end();
          break;
        case LKind::Kop:
          begin_capture();
          t_begin(PC_OTHER, 2.0 * (double)rc.N * (double)esz * l.nf);
          hip_check(launch_kop<S>(fa + l.arg_off, l.nf, l.geom, st), "kop");
          t_end();
          break;
        case LKind::Detect:
          begin_capture();
          hip_check(launch_detect<S>(da + l.arg_off, l.nf, bounds, npx, st), "detect");
          break;
        case LKind::Copy:
          if (l.member == -1) {
            end_capture();
            hip_check(hipMemcpyAsync(l.dst, l.src, l.n * esz, hipMemcpyHostToDevice, st), "upload");
          } else {
            begin_capture();
            hip_check(launch_copy<S>(l.dst, l.src, l.n, st), "copy");
          }
          break;
        case LKind::Fill:
          begin_capture();
          hip_check(launch_fill_f64((double *)l.dst, l.n, 0.0, st), "fill");
          break;
        case LKind::SaveU: {
          end_capture();
          Member *m = mem[l.member];
          if (m->saved_u) {
            std::vector<char> host(l.n * esz);
            hip_check(hipMemcpyAsync(host.data(), l.src, l.n * esz, hipMemcpyDeviceToHost, st), "download u");
            hip_check(hipStreamSynchronize(st), "sync");
            if ((int)m->saved_u->size() <= l.step) m->saved_u->resize(l.step + 1);
            (*m->saved_u)[l.step] = std::move(host);
          }
          break;
        }
      }
    }
    end_capture();
    hip_check(hipStreamSynchronize(st), "run sync");
    for (size_t i = 0; i < ev_meta.size(); ++i) {
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]), "event elapsed");
      PassStat &p = ctx.stats[ev_meta[i].first];
      p.launches += 1;
      p.ms += ms;
      p.bytes += ev_meta[i].second;
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
};

template <typename S>
std::vector<double> run_members_t(Context &ctx, const RunConfig &rc, std::vector<Member> &members) {
  hip_check(hipSetDevice(ctx.device), "hipSetDevice");
  Plan &plan = ctx.plan(rc.N, rc.f64, rc.lgR, rc.lgA);
  const int64_t npx = detector_pixels(rc.det_size, rc.det_px);
  const std::vector<int64_t> hb = detector_bounds(rc.N, rc.dx, rc.det_size, rc.det_px);
  DevBuf dbounds(hb.size() * sizeof(int64_t));
  hip_check(hipMemcpyAsync(dbounds.p, hb.data(), hb.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx.stream),
            "bounds upload");
  const size_t nacc = (size_t)rc.n_out * std::max<int64_t>(npx, 0);
  DevBuf dacc(std::max<size_t>(nacc, 1) * sizeof(double));
  hip_check(launch_fill_f64((double *)dacc.p, (int64_t)nacc, 0.0, ctx.stream), "acc fill");

  bool needs_snap = false;
  for (auto &m : members)
    for (auto &o : m.ops) needs_snap |= o.kind == OpKind::Restore;

  const size_t fbytes = (size_t)rc.N * 2 * sizeof(S);
  int group = rc.max_resident > 0 ? rc.max_resident : (int)members.size();
  if (rc.max_resident <= 0 && !rc.fields) {
    size_t freeb = 0, totalb = 0;
    hip_check(hipMemGetInfo(&freeb, &totalb), "hipMemGetInfo");
    const size_t per = fbytes * (needs_snap ? 2 : 1);
    const size_t budget = freeb > (size_t)(2ull << 30) ? freeb - (size_t)(2ull << 30) : freeb / 2;
    group = (int)std::max<size_t>(1, std::min<size_t>(members.size(), budget / per));
  }
  if (rc.fields) group = (int)members.size();
  group = std::max(1, std::min(group, 65535));

  for (size_t g0 = 0; g0 < members.size(); g0 += group) {
    const size_t g1 = std::min(members.size(), g0 + group);
    Exec<S> ex(ctx, rc, plan);
    std::vector<DevBuf> fbuf, sbuf;
    for (size_t i = g0; i < g1; ++i) {
      ex.mem.push_back(&members[i]);
      if (rc.fields) {
        ex.field.push_back(rc.fields[i]);
      } else {
        fbuf.emplace_back(fbytes);
        ex.field.push_back(fbuf.back().p);
      }
      if (needs_snap) {
        sbuf.emplace_back(fbytes);
        ex.snap.push_back(sbuf.back().p);
      }
    }
    ex.acc = (double *)dacc.p;
    ex.bounds = (const int64_t *)dbounds.p;
    ex.npx = npx;
    ex.compile();
    ex.execute();
  }

  std::vector<double> accs(nacc);
  if (nacc)
    hip_check(hipMemcpyAsync(accs.data(), dacc.p, nacc * sizeof(double), hipMemcpyDeviceToHost, ctx.stream),
              "acc download");
  hip_check(hipStreamSynchronize(ctx.stream), "sync");

  const std::vector<double> sc = detector_scale(rc.det_size, rc.det_px, rc.det_py, rc.dx, rc.z_det);
  std::vector<double> out(nacc);
  for (int r = 0; r < rc.n_out; ++r)
    for (int64_t k = 0; k < npx; ++k) {
      const double v = accs[(size_t)r * npx + k];
      out[(size_t)r * npx + k] = (((v * sc[2 * k]) * rc.dx) * rc.det_py) / sc[2 * k + 1];
    }
  return out;
}

}  // namespace

std::vector<double> run_members(Context &ctx, const RunConfig &rc, std::vector<Member> &members) {
  if (members.empty()) return {};
  return rc.f64 ? run_members_t<double>(ctx, rc, members) : run_members_t<float>(ctx, rc, members);
}

Context::Tuned autotune(Context &ctx, int64_t N, bool f64, int reps) {
  int lgN = 0;
  while ((int64_t(1) << lgN) < N) ++lgN;
  // synthetic slice loop: plane wave, `reps` grating slices of a 4.2 um
  // grating at 46 keV, 1 mm detector
  const double dx = 2e-9, wl = energy_to_wavelength(46000.0);
  Member m;
  m.wl = wl;
  Op src{OpKind::Source};
  src.x.kind = XOP_ONES;
  src.x.post[0] = 1.0;
  m.ops.push_back(src);
  for (int i = 0; i < reps + 1; ++i) {
    Op mk{OpKind::Mask};
    mk.x.kind = XOP_MASK;
    mk.x.fa[0] = 0.9;
    mk.x.fa[1] = 0.1;
    mk.x.fb[0] = 1.0;
    mk.x.post[0] = 1.0;
    mk.x.pitch = 4.2e-6;
    mk.x.inv_pitch = 1.0 / 4.2e-6;
    mk.x.dc = 0.5;
    mk.x.half_dc = 0.25;
    mk.x.dx = dx;
    mk.x.chunk = N;
    mk.x.single_chunk = 1;
    mk.x.x0 = (double)(-(N / 2)) * dx;
    mk.x.xstep = dx;
    m.ops.push_back(mk);
    Op p{OpKind::Prop};
    p.dz = 1e-6;
    p.fc = 1e30;
    m.ops.push_back(p);
  }
  m.ops.push_back(Op{OpKind::Flush});
  RunConfig rc;
  rc.N = N;
  rc.dx = dx;
  rc.f64 = f64;
  rc.n_out = 1;
  const bool was = ctx.timing;
  Context::Tuned best{lgN, f64, -1, -1, 1e300};
  for (auto cand : plan_candidates(lgN, f64)) {
    rc.lgR = cand.first;
    rc.lgA = cand.second;
    std::vector<Member> ms{m};
    ctx.timing = false;
    (void)run_members(ctx, rc, ms);  // warm: plan tables, code objects
    ms = {m};
    for (auto &st : ctx.stats) st = PassStat{};
    ctx.timing = true;
    (void)run_members(ctx, rc, ms);
    double tot = 0;
    for (auto &st : ctx.stats) tot += st.ms;
    const double per = tot / (reps + 1);
    if (per < best.ms) best = Context::Tuned{lgN, f64, cand.first, cand.second, per};
  }
  ctx.timing = was;
  for (auto &st : ctx.stats) st = PassStat{};
  bool replaced = false;
  for (auto &t : ctx.tuned)
    if (t.lgN == lgN && t.f64 == f64) {
      t = best;
      replaced = true;
    }
  if (!replaced) ctx.tuned.push_back(best);
  return best;
}

}  // namespace rs
