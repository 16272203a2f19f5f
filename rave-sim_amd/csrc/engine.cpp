// Host engine: exact big-wave helper arithmetic, FFT plans, and the lockstep
// executor that compiles member programs into fused gfx950 launches.
#include "engine.hpp"

#include <algorithm>
#include <memory>
#include <cstdio>
#include <exception>
#include <thread>

#include <fcntl.h>
#include <unistd.h>
#include <chrono>
#include <cstring>
#include <map>
#include <sstream>

#include "exec.hpp"
#include "npy.hpp"
#include "rs_detector.hpp"
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
double py_floordiv(double vx, double wx) {
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
  const DetGeom g = det_geom(N, dx, o, px);
  for (int64_t k = 0; k <= o; ++k) b[k] = det_bound(g, k);
  return b;
}

void detector_pixel_scale(double size, double px, double z, int64_t k, double *cos_a, double *dist) {
  // x = (k - n//2) px (util.py:78-82), cos(atan(x / z)) and sqrt(x^2 + z^2)
  const double n = py_floordiv(size, px);
  const double h = py_floordiv(n, 2.0);
  const double x = ((double)k - h) * px;
  *cos_a = std::cos(std::atan(x / z));
  *dist = std::sqrt(x * x + z * z);
}

std::vector<double> detector_scale(double size, double px, double py, double dx, double z) {
  // returned as pairs (cos(atan(x/z)), sqrt(x^2 + z^2)) so that the caller
  // applies them in big-wave's order: ((out * cos) * dx * py) / sqrt(...)
  const int64_t o = (int64_t)py_floordiv(size, px);
  std::vector<double> s(2 * std::max<int64_t>(o, 0));
  for (int64_t k = 0; k < o; ++k) detector_pixel_scale(size, px, z, k, &s[2 * k], &s[2 * k + 1]);
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
  if (cis) (void)hipFree(cis);
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
  int rest = lgN - r;
  if (rest <= kMaxColLg && (lgA_req < 0 || lgA_req == rest)) {
    *ng = 2; *lgR = r; *lgA = rest; *lgB = 0;
    return;
  }
  const int maxW = f64 ? kMaxWideColLgF64 : kMaxWideColLgF32;
  if (lgA_req >= 0 && lgA_req == rest && rest <= maxW) {  // wide two-group split
    *ng = 2; *lgR = r; *lgA = rest; *lgB = 0;
    return;
  }
  // three groups with no requested row length: rows one digit shorter than
  // the longest (4096 complex64, 2048 complex128), whose passes issue fewer
  // instructions per sample; 12|7|7 is among the fastest 2^26 complex64 splits
  if (lgR_req < 0) {
    r = maxR - 1;
    rest = lgN - r;
  }
  int a = lgA_req >= 0 ? lgA_req : (rest + 1) / 2;
  // untuned complex64 plans keep the top column at 256 or less, where its
  // tiles read 256 B row segments (2^28: 635 ms per Talbot-Lau simulation
  // with 12|8|8 against 689 with 12|9|7; 2^30: 3.1 s with 12|8|10 against
  // 3.3 s with 12|9|9)
  if (lgA_req < 0 && !f64) a = std::min(a, 8);
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

// In-fiber twiddles of a fiber of length L = 2^lgL: W_L^j for j < L, then
// the [k][r] tables of the small Stockham stages (small_tab_offset):
// entry k * kTwRow + r of stage NS is W_{16 NS}^{k r} = W_L^{k r L / (16 NS)}
// for r < 16, zero in the pad column.
static void *make_fiber_table(int lgL, bool f64) {
  const int L = 1 << lgL;
  const int n = fiber_tab_entries(L);
  std::vector<long double> re(n, 0.0L), im(n, 0.0L);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  auto w = [&](int64_t e, long double *c, long double *s) {
    const long double ang = -two_pi * (long double)(e & (L - 1)) / (long double)L;
    *c = cosl(ang);
    *s = sinl(ang);
  };
  for (int j = 0; j < L; ++j) w(j, &re[j], &im[j]);
  for (int ns = stage_r0(L); ns < L; ns *= 16) {
    if (ns <= 1 || ns > kSmallNS) continue;
    const int base = L + small_tab_offset(L, ns);
    const int64_t step = L / (16 * ns);
    for (int k = 0; k < ns; ++k)
      for (int r = 0; r < 16; ++r) w((int64_t)k * r * step, &re[base + k * kTwRow + r], &im[base + k * kTwRow + r]);
  }

  return f64 ? upload_table<double>(re, im) : upload_table<float>(re, im);
}

bool plan_tiles_valid(int lgN, bool f64, int lgR, int lgA, int tileA, int tileB) {
  int ng, a, b, r;
  plan_split(lgN, f64, lgR, lgA, &ng, &a, &b, &r);
  if (ng == 1) return tileA == 0 && tileB == 0;
  if (ng == 2 && tileB != 0) return false;
  if ((int64_t(1) << (b + r)) % col_fibers(1 << a, f64, tileA) != 0) return false;
  if (ng == 3 && (int64_t(1) << r) % col_fibers(1 << b, f64, tileB) != 0) return false;
  return true;
}

std::unique_ptr<Plan> make_plan(int64_t N, bool f64, int lgR_req, int lgA_req, int tileA, int tileB) {
  int lgN = 0;
  while ((int64_t(1) << lgN) < N) ++lgN;
  if ((int64_t(1) << lgN) != N) throw Error(6, "N must be a power of two");
  auto p = std::make_unique<Plan>();
  p->N = N;
  p->lgN = lgN;
  p->f64 = f64;
  plan_split(lgN, f64, lgR_req, lgA_req, &p->ngroups, &p->lgA, &p->lgB, &p->lgR);
  if (tileA < 0 || tileA >= kColTiles || tileB < 0 || tileB >= kColTiles) throw Error(1, "tile variant out of range");
  if (!plan_tiles_valid(lgN, f64, p->lgR, p->ngroups >= 2 ? p->lgA : p->lgR, p->ngroups >= 2 ? tileA : 0,
                        p->ngroups == 3 ? tileB : 0))
    throw Error(6, "column tile does not fit this split");
  p->tileA = p->ngroups >= 2 ? tileA : 0;
  p->tileB = p->ngroups == 3 ? tileB : 0;
  // in-fiber tables: row, top, middle
  p->fib[0] = make_fiber_table(p->lgR, f64);
  if (p->ngroups >= 2) p->fib[1] = make_fiber_table(p->lgA, f64);
  if (p->ngroups == 3) p->fib[2] = make_fiber_table(p->lgB, f64);
  // inter-group tables: top column M = N, middle column M = 2^(lgB + lgR)
  for (int t = 0; t < p->ngroups - 1; ++t) {
    const int lgM = t == 0 ? lgN : p->lgB + p->lgR;
    const int s = (lgM + 1) / 2;
    p->sh[t] = s;
    p->lo[t] = make_twiddles(lgM, int64_t(1) << s, 1, f64);
    p->hi[t] = make_twiddles(lgM, int64_t(1) << (lgM - s), int64_t(1) << s, f64);
  }
  // complex64 row-pass phase table exp(+2 pi i j / 256)
  if (!f64) {
    std::vector<long double> re(256), im(256);
    for (int j = 0; j < 256; ++j) {
      const long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)j / 256.0L;
      re[j] = cosl(ang);
      im[j] = sinl(ang);
    }
    p->cis = upload_table<float>(re, im);
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
  g.cis_tab = cis;
  return g;
}
static int32_t line_sharing(int lgL, bool f64, int tile) {
  const int bytes = col_fibers(1 << lgL, f64, tile) * (f64 ? 16 : 8);
  return bytes >= 128 ? 1 : 128 / bytes;
}

Geom Plan::geom_top() const {
  Geom g = geom_row();
  g.top = 1;
  g.xcd_segs = line_sharing(lgA, f64, tileA);
  g.tile = tileA;
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
  g.xcd_segs = line_sharing(lgB, f64, tileB);
  g.tile = tileB;
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
  if (tileA || tileB) s << " tiles " << tileA << tileB;
  return s.str();
}

Context::Context(int dev) : device(dev) {
  hip_check(hipSetDevice(dev), "hipSetDevice");
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  noalloc_scr = noalloc_scratch(dev);
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  device_key = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount);
  for (char &ch : device_key)
    if (ch == ' ') ch = '_';
}
Context::~Context() {
  graph_cache.reset();
  plans_.clear();
  for (Stage &sg : stage_)
    for (int b = 0; b < 2; ++b) {
      if (sg.ev[b]) {
        (void)hipEventSynchronize(sg.ev[b]);
        (void)hipEventDestroy(sg.ev[b]);
      }
      if (sg.buf[b]) (void)hipHostFree(sg.buf[b]);
    }
  for (hipStream_t s : lanes_) (void)hipStreamDestroy(s);
  if (stream) (void)hipStreamDestroy(stream);
}
void Context::stream_in(void *dst, int64_t count, size_t esz, hipStream_t st,
                        const std::function<void(char *, int64_t, int64_t)> &fill) {
  const int64_t per = (int64_t)(kStageBytes / esz);  // elements per chunk
  const int T = stage_threads(count * (int64_t)esz);
  std::vector<std::exception_ptr> err(T);
  auto part = [&](int t) {
    try {
      hip_check(hipSetDevice(device), "hipSetDevice");
      Stage &sg = stage_[t];
      const int64_t lo = count * t / T, hi = count * (t + 1) / T;
      int k = 0;
      for (int64_t off = lo; off < hi; off += per, ++k) {
        const int b = k & 1;
        const int64_t n = std::min(per, hi - off);
        if (sg.used[b]) hip_check(hipEventSynchronize(sg.ev[b]), "staging wait");
        fill((char *)sg.buf[b], off, n);
        hip_check(hipMemcpyAsync((char *)dst + (size_t)off * esz, sg.buf[b], (size_t)n * esz, hipMemcpyHostToDevice,
                                 st),
                  "upload");
        hip_check(hipEventRecord(sg.ev[b], st), "event record");
        sg.used[b] = true;
      }
    } catch (...) {
      err[t] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < T; ++t) th.emplace_back(part, t);
  part(0);
  for (auto &x : th) x.join();
  for (auto &e : err)
    if (e) std::rethrow_exception(e);
}

void Context::stream_npy(const Member &m, bool f64, int64_t N, void *dst, hipStream_t st) {
  const int fd = ::open(m.upload_npy.c_str(), O_RDONLY);
  if (fd < 0) throw Error(3, "cannot open " + m.upload_npy);
  struct Closer {
    int fd;
    ~Closer() { ::close(fd); }
  } closer{fd};
  const size_t desz = f64 ? 16 : 8, sesz = m.upload_c16 ? 16 : 8;
  auto read_at = [&](char *p, size_t want, int64_t at) {
    size_t got = 0;
    while (got < want) {
      const ssize_t r = ::pread(fd, p + got, want - got, (off_t)(at + (int64_t)got));
      if (r <= 0) throw Error(3, m.upload_npy + ": truncated data");
      got += (size_t)r;
    }
  };
  stream_in(dst, N, desz, st, [&](char *buf, int64_t first, int64_t n) {
    const int64_t at = m.upload_offset + first * (int64_t)sesz;
    if (sesz == desz) {
      read_at(buf, (size_t)n * sesz, at);
      return;
    }
    // precision conversion through a per-thread scratch chunk
    thread_local std::vector<char> raw;
    raw.resize((size_t)n * sesz);
    read_at(raw.data(), raw.size(), at);
    if (sesz == 16) {
      const double *s = (const double *)raw.data();
      float *d = (float *)buf;
      for (int64_t i = 0; i < 2 * n; ++i) d[i] = (float)s[i];
    } else {
      const float *s = (const float *)raw.data();
      double *d = (double *)buf;
      for (int64_t i = 0; i < 2 * n; ++i) d[i] = (double)s[i];
    }
  });
}

void Context::copy_from_host(void *dst, const void *host, size_t bytes, hipStream_t st) {
  stream_in(dst, (int64_t)bytes, 1, st,
            [&](char *buf, int64_t first, int64_t n) { std::memcpy(buf, (const char *)host + first, (size_t)n); });
}

int Context::stage_threads(int64_t bytes) {
  // one thread per 32 MiB of field, at most kStageThreads
  const int T = (int)std::max<int64_t>(1, std::min<int64_t>(kStageThreads, bytes >> 25));
  for (int t = 0; t < T; ++t)
    for (int b = 0; b < 2; ++b) {
      if (!stage_[t].buf[b])
        hip_check(hipHostMalloc(&stage_[t].buf[b], kStageBytes, hipHostMallocDefault), "hipHostMalloc");
      if (!stage_[t].ev[b]) hip_check(hipEventCreateWithFlags(&stage_[t].ev[b], hipEventDisableTiming), "event");
    }
  return T;
}

namespace {
void pwrite_full(int fd, const char *p, size_t n, int64_t at, const std::string &path) {
  while (n > 0) {
    const ssize_t w = ::pwrite(fd, p, n, (off_t)at);
    if (w <= 0) throw Error(3, "write failed: " + path);
    p += w;
    n -= (size_t)w;
    at += w;
  }
}
}  // namespace

void Context::stream_out(const void *src, size_t bytes, hipStream_t st,
                         const std::function<void(const char *, size_t, size_t)> &sink) {
  const int T = stage_threads((int64_t)bytes);
  std::vector<std::exception_ptr> err(T);
  auto part = [&](int t) {
    try {
      hip_check(hipSetDevice(device), "hipSetDevice");
      Stage &sg = stage_[t];
      const size_t lo = bytes * t / T, hi = bytes * (t + 1) / T;
      size_t prev_off = 0, prev_n = 0;
      int prev_b = 0, k = 0;
      bool have = false;
      auto flush = [&]() {
        hip_check(hipEventSynchronize(sg.ev[prev_b]), "staging wait");
        sink((const char *)sg.buf[prev_b], prev_off, prev_n);
      };
      for (size_t off = lo; off < hi; off += kStageBytes, ++k) {
        const int b = k & 1;
        const size_t n = std::min(kStageBytes, hi - off);
        if (sg.used[b]) hip_check(hipEventSynchronize(sg.ev[b]), "staging wait");
        hip_check(hipMemcpyAsync(sg.buf[b], (const char *)src + off, n, hipMemcpyDeviceToHost, st), "download");
        hip_check(hipEventRecord(sg.ev[b], st), "event record");
        sg.used[b] = true;
        if (have) flush();
        prev_off = off;
        prev_n = n;
        prev_b = b;
        have = true;
      }
      if (have) flush();
    } catch (...) {
      err[t] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < T; ++t) th.emplace_back(part, t);
  part(0);
  for (auto &x : th) x.join();
  for (auto &e : err)
    if (e) std::rethrow_exception(e);
}

void Context::stream_to_npy(const void *src, int64_t N, bool f64, const std::string &path, hipStream_t st) {
  const std::string hdr = npy_header_bytes(f64 ? "<c16" : "<c8", {N});
  const int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) throw Error(3, "cannot write " + path);
  struct Closer {
    int fd;
    ~Closer() { ::close(fd); }
  } closer{fd};
  pwrite_full(fd, hdr.data(), hdr.size(), 0, path);
  stream_out(src, (size_t)N * (f64 ? 16 : 8), st, [&](const char *p, size_t off, size_t n) {
    pwrite_full(fd, p, n, (int64_t)(hdr.size() + off), path);
  });
}

void Context::copy_to_host(void *host, const void *src, size_t bytes, hipStream_t st) {
  stream_out(src, bytes, st, [&](const char *p, size_t off, size_t n) { std::memcpy((char *)host + off, p, n); });
}

hipStream_t Context::lane_stream(int i) {
  if (i == 0) return stream;
  while ((int)lanes_.size() < i) {
    hipStream_t s = nullptr;
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    lanes_.push_back(s);
  }
  return lanes_[i - 1];
}
const Context::Tuned *Context::find_tuned(int lgN, bool f64, int members) const {
  const Tuned *best = nullptr;
  double bd = 1e300;
  for (const auto &t : tuned) {
    if (t.lgN != lgN || t.f64 != f64) continue;
    if (t.members == 0) return &t;  // fixed split
    const double d = std::fabs(std::log2((double)t.members) - std::log2((double)std::max(members, 1)));
    if (d < bd) {
      bd = d;
      best = &t;
    }
  }
  return best;
}

Plan &Context::plan(int64_t N, bool f64, int lgR, int lgA, int members, int tileA, int tileB) {
  if (lgR < 0 && lgA < 0) {
    int lgN = 0;
    while ((int64_t(1) << lgN) < N) ++lgN;
    if (const Tuned *t = find_tuned(lgN, f64, members)) {
      lgR = t->lgR;
      lgA = t->lgA;
      if (tileA < 0) tileA = t->tileA;
      if (tileB < 0) tileB = t->tileB;
    }
  }
  tileA = std::max(tileA, 0);
  tileB = std::max(tileB, 0);
  const std::array<int, 4> req{lgR, lgA, tileA, tileB};
  for (size_t i = 0; i < plans_.size(); ++i)
    if (plans_[i]->N == N && plans_[i]->f64 == f64 && plan_req_[i] == req) return *plans_[i];
  plans_.push_back(make_plan(N, f64, lgR, lgA, tileA, tileB));
  plan_req_.push_back(req);
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
using namespace exec;

// Default lockstep group: the members whose buffers fit the device's
// last-level cache (MALL) together, at least one, so that a group runs all
// of its passes out of the cache; bounded by free memory (minus 2 GiB).
int default_group(size_t bytes_per_member, int64_t n_members) {
  size_t freeb = 0, totalb = 0;
  hip_check(hipMemGetInfo(&freeb, &totalb), "hipMemGetInfo");
  const size_t budget = freeb > (size_t)(2ull << 30) ? freeb - (size_t)(2ull << 30) : freeb / 2;
  const size_t cached = std::max<size_t>(1, kMallBytes / bytes_per_member);
  const size_t g = std::min<size_t>({(size_t)n_members, budget / bytes_per_member, cached});
  return (int)std::max<size_t>(1, std::min<size_t>(g, 65535));
}

// One lockstep run of members that share N, precision, detector and
// element structure: its plan, groups, buffers and accumulators.  A call of
// run_members is one BatchRun; run_mixed holds one per distinct
// configuration and enqueues them all between one argument upload and one
// synchronisation.
template <typename S> struct BatchRun {
  Context &ctx;
  RunConfig rc;
  std::vector<Member> &members;
  int64_t npx = 0;
  size_t nacc = 0, fbytes = 0, per_member = 1, width = 0;
  bool needs_snap = false, needs_hist = false;
  int group = 1, lanes = 1, spans = 1;
  Plan *plan = nullptr;
  RunBuffers buf;
  std::vector<DevBuf> sample_bufs;
  std::vector<std::unique_ptr<Exec<S>>> groups;
  MaskCache<S> masks;
  std::vector<std::vector<double>> accs;

  BatchRun(Context &c, const RunConfig &r, std::vector<Member> &m) : ctx(c), rc(r), members(m) {}

  // the key under which this run's buffers and graph may be reused: what
  // decides the launch sequence before any buffer exists
  static std::string shape_key(const RunConfig &rc, const std::vector<Member> &members) {
    std::ostringstream k;
    k << rc.N << ' ' << rc.f64 << ' ' << rc.mode << ' ' << rc.n_out << ' ' << detector_pixels(rc.det_size, rc.det_px)
      << ' ' << rc.max_resident << ' ' << rc.lanes << ' ' << rc.spans << ' ' << rc.lgR << ' ' << rc.lgA << ' '
      << rc.tileA << ' ' << rc.tileB << ' ' << members.size() << ':';
    for (const auto &m : members) {
      for (const auto &o : m.ops) k << (char)('A' + (int)o.kind);
      k << '/' << m.out_base << ',';
    }
    return k.str();
  }

  // plan, buffers (taken from `reuse` when they fit), compiled groups and
  // the shared grating decisions, with the uploads on the context stream;
  // `lanes_max` bounds the lanes (0: no bound beyond the default)
  void setup(RunBuffers *reuse = nullptr, int lanes_max = 0) {
    npx = detector_pixels(rc.det_size, rc.det_px);
    const std::vector<int64_t> hb = detector_bounds(rc.N, rc.dx, rc.det_size, rc.det_px);
    nacc = (size_t)rc.n_out * std::max<int64_t>(npx, 0);
    for (auto &m : members)
      for (auto &o : m.ops) {
        needs_snap |= o.kind == OpKind::Restore;
        needs_hist |= o.kind == OpKind::HSave;
      }
    per_member = 1 + (needs_snap ? 1 : 0) + (needs_hist ? 1 : 0);  // buffers
    fbytes = (size_t)rc.N * 2 * sizeof(S);
    group = rc.max_resident > 0 ? rc.max_resident : (int)members.size();
    if (rc.max_resident <= 0 && !rc.fields) group = default_group(fbytes * per_member, (int64_t)members.size());
    if (rc.fields) group = (int)members.size();
    group = std::max(1, std::min(group, 65535));

    // Lanes: with the default grouping and more than one group, groups
    // alternate between the context's two streams, so that a VALU-bound pass
    // of one group runs beside a memory-bound pass of the other; the
    // resident fields are split between the lanes.  One lane when timing per
    // pass, with caller fields or an explicit group size.
    lanes = 1;
    if (!rc.fields && !ctx.timing && !(rc.use_graph && !rc.whole_graph) && members.size() >= 2) {
      if (rc.lanes > 0) lanes = rc.lanes;
      else if (rc.max_resident <= 0) lanes = 2;
    }
    lanes = std::max(1, std::min({lanes, 4, (int)members.size()}));
    if (lanes_max > 0) lanes = std::min(lanes, lanes_max);
    if (lanes > 1 && rc.max_resident <= 0) group = std::max(1, group / lanes);
    plan = &ctx.plan(rc.N, rc.f64, rc.lgR, rc.lgA, (int)std::min<size_t>(members.size(), 65535), rc.tileA,
                     rc.tileB);
    width = std::min(members.size(), (size_t)group);

    // Field (and snapshot) buffers of one group per lane, reused by every
    // group of the lane: each member program starts by generating or
    // uploading its field.  Per-lane accumulators (summed in lane order on
    // the host).  Buffers of a previous run of this shape are taken over
    // when they have the sizes this run needs.
    auto fits = [&](const std::vector<std::vector<DevBuf>> &v, bool want) {
      if ((int)v.size() != lanes) return false;
      for (const auto &lb : v) {
        if (lb.size() != (want ? width : 0)) return false;
        for (const auto &b : lb)
          if (b.bytes != fbytes) return false;
      }
      return true;
    };
    bool take = reuse && reuse->bounds.bytes == hb.size() * sizeof(int64_t) && (int)reuse->acc.size() == lanes &&
                fits(reuse->field, !rc.fields) && fits(reuse->snap, needs_snap) && fits(reuse->hist, needs_hist);
    if (take)
      for (const auto &a : reuse->acc) take &= a.bytes == std::max<size_t>(nacc, 1) * sizeof(double);
    if (take) {
      buf = std::move(*reuse);
    } else {
      buf = RunBuffers{};
      buf.bounds = DevBuf(hb.size() * sizeof(int64_t));
      buf.field.resize(lanes);
      buf.snap.resize(lanes);
      buf.hist.resize(lanes);
      for (int l = 0; l < lanes; ++l) {
        buf.acc.emplace_back(std::max<size_t>(nacc, 1) * sizeof(double));
        if (!rc.fields)
          for (size_t i = 0; i < width; ++i) buf.field[l].emplace_back(fbytes);
        if (needs_snap)
          for (size_t i = 0; i < width; ++i) buf.snap[l].emplace_back(fbytes);
        if (needs_hist)
          for (size_t i = 0; i < width; ++i) buf.hist[l].emplace_back(fbytes);
      }
    }
    hip_check(hipMemcpyAsync(buf.bounds.p, hb.data(), hb.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx.stream),
              "bounds upload");
    for (int l = 0; l < lanes; ++l)
      hip_check(launch_fill_f64((double *)buf.acc[l].p, (int64_t)nacc, 0.0, ctx.stream), "acc fill");
    // sample elements: device copies of the members' pixel tables for this
    // run, pointed at by their SAMPLE masks
    auto upload_d = [&](const std::vector<double> &h) -> const double * {
      sample_bufs.emplace_back(std::max<size_t>(h.size(), 1) * sizeof(double));
      hip_check(hipMemcpyAsync(sample_bufs.back().p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                               ctx.stream),
                "sample upload");
      return (const double *)sample_bufs.back().p;
    };
    for (auto &m : members) {
      if (m.samples.empty()) continue;
      std::vector<const double *> rx, db, f;
      for (auto &sr : m.samples) {
        rx.push_back(upload_d(sr.rx));
        db.push_back(upload_d(sr.db));
        f.push_back(upload_d(sr.f));
      }
      for (auto &op : m.ops) {
        if (op.sample < 0) continue;
        const SampleRows &sr = m.samples.at((size_t)op.sample);
        op.x.s_rx = rx[(size_t)op.sample];
        op.x.s_db = db[(size_t)op.sample] + (size_t)op.sample_row * 2 * sr.len;
        op.x.s_f = f[(size_t)op.sample] + (size_t)op.sample_row * 2 * sr.len;
        op.x.s_len = sr.len;
      }
    }
    // the member's host arrays need not outlive the copies
    if (!sample_bufs.empty()) hip_check(hipStreamSynchronize(ctx.stream), "sample upload sync");

    // spans of the middle-row-middle sequence (three-group plans): a group
    // whose fields overflow the last-level cache runs it over halves,
    // quarters, ... of the fields that fit
    spans = 1;
    if (rc.mode == 0 && plan->ngroups == 3 && plan->geom_mid().xcd_segs == 1) {
      if (rc.spans > 0) {
        if (rc.spans & (rc.spans - 1) || rc.spans > (1 << plan->lgA))
          throw Error(1, "spans must be a power of two no larger than the top group");
        spans = rc.spans;
      } else {
        const size_t gbytes = (size_t)lanes * width * fbytes;
        while (spans < (1 << plan->lgA) && gbytes / (size_t)spans > kMallBytes) spans *= 2;
      }
    }

    // resident fields larger than the last-level cache stream from HBM
    const size_t resident = (size_t)lanes * width * fbytes * per_member;
    const bool streaming = rc.mode == 0 && resident > kMallBytes;
    for (size_t g0 = 0, gi = 0; g0 < members.size(); g0 += group, ++gi) {
      const size_t g1 = std::min(members.size(), g0 + group);
      const int lane = (int)(gi % lanes);
      auto ex = std::make_unique<Exec<S>>(ctx, rc, *plan);
      ex->st = ctx.lane_stream(lane);
      ex->spans = spans;
      ex->stream_top = streaming;
      for (size_t i = g0; i < g1; ++i) {
        ex->mem.push_back(&members[i]);
        ex->field.push_back(rc.fields ? rc.fields[i] : buf.field[lane][i - g0].p);
        if (needs_snap) ex->snap.push_back(buf.snap[lane][i - g0].p);
        if (needs_hist) ex->hist.push_back(buf.hist[lane][i - g0].p);
      }
      ex->acc = (double *)buf.acc[lane].p;
      ex->bounds = (const int64_t *)buf.bounds.p;
      ex->npx = npx;
      ex->compile();
      groups.push_back(std::move(ex));
    }
    // Grating decisions shared by several passes: computed once, read as
    // bits (N/8 bytes per mask).  Only when the resident fields overflow the
    // last-level cache anyway: when they fit, the extra stream evicts them
    // and costs more than the decisions it saves.
    if (streaming)
      for (auto &ex : groups) masks.collect(*ex);
    masks.fill(rc.N, ctx.stream);
  }

  void append_args(std::vector<FieldArgs> &fa, std::vector<DetArgs> &da) const {
    for (auto &ex : groups) {
      fa.insert(fa.end(), ex->fargs.begin(), ex->fargs.end());
      da.insert(da.end(), ex->dargs.begin(), ex->dargs.end());
    }
  }

  // everything the captured launches depend on besides the argument
  // records' contents: launch records, streams, and the buffers and sizes
  // passed to kernels by value
  void launch_key(std::string &k) const {
    auto put = [&](const void *p, size_t n) { k.append((const char *)p, n); };
    for (auto &ex : groups) {
      put(&ex->st, sizeof ex->st);
      put(&ex->bounds, sizeof ex->bounds);
      put(&ex->npx, sizeof ex->npx);
      for (const auto &l : ex->launches) {
        put(&l.kind, sizeof l.kind);
        put(&l.lgL, sizeof l.lgL);
        put(&l.row, sizeof l.row);
        put(&l.xv, sizeof l.xv);
        put(&l.geom, sizeof l.geom);
        put(&l.arg_off, sizeof l.arg_off);
        put(&l.nf, sizeof l.nf);
        put(&l.dst, sizeof l.dst);
        put(&l.src, sizeof l.src);
        put(&l.n, sizeof l.n);
      }
    }
  }

  // every group's launches on its lane, argument records at fa / da
  void enqueue(const FieldArgs *fa, const DetArgs *da) {
    size_t ofa = 0, oda = 0;
    for (auto &ex : groups) {
      ex->enqueue(fa + ofa, da + oda);
      ofa += ex->fargs.size();
      oda += ex->dargs.size();
    }
  }

  // accumulator download on the context stream (after the lanes joined it)
  void download() {
    accs.assign(lanes, std::vector<double>(nacc));
    for (int l = 0; l < lanes && nacc; ++l)
      hip_check(
          hipMemcpyAsync(accs[l].data(), buf.acc[l].p, nacc * sizeof(double), hipMemcpyDeviceToHost, ctx.stream),
          "acc download");
  }

  // after the stream drained: pass timings and the scaled detector rows
  std::vector<double> finish() {
    for (auto &ex : groups) ex->collect_timing();
    if (!rc.row_z.empty() && (int)rc.row_z.size() != rc.n_out) throw Error(1, "row_z must have n_out entries");
    std::vector<double> sc = detector_scale(rc.det_size, rc.det_px, rc.det_py, rc.dx, rc.z_det);
    std::vector<double> out(nacc);
    for (int r = 0; r < rc.n_out; ++r) {
      if (!rc.row_z.empty() && (r == 0 || rc.row_z[(size_t)r] != rc.row_z[(size_t)r - 1]))
        sc = detector_scale(rc.det_size, rc.det_px, rc.det_py, rc.dx, rc.row_z[(size_t)r]);
      for (int64_t k = 0; k < npx; ++k) {
        double v = accs[0][(size_t)r * npx + k];
        for (int l = 1; l < lanes; ++l) v += accs[l][(size_t)r * npx + k];
        out[(size_t)r * npx + k] = (((v * sc[2 * k]) * rc.dx) * rc.det_py) / sc[2 * k + 1];
      }
    }
    return out;
  }
};

}  // namespace

// The last captured batch run of a context: its instantiated graph, the
// argument buffers and run buffers its launches address, and the keys that
// tell whether a new run would record the same launches.
struct GraphCache {
  std::string shape, launches;
  hipGraphExec_t exec = nullptr;
  DevBuf dfa, dda;
  std::vector<RunBuffers> bufs;
  ~GraphCache() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

namespace {

// Enqueue runs (already set up) on the context's lanes between one argument
// upload and one synchronisation.  whole_graph: the launches of every lane
// are captured as one hipGraph (the lanes fork from and join the context
// stream through events inside the capture) and instantiated; the graph and
// the buffers it addresses stay in ctx.graph_cache, and a later run whose
// launches are the same replays it with its own argument records.
template <typename S>
void execute_runs(Context &ctx, std::vector<std::unique_ptr<BatchRun<S>>> &runs, bool whole_graph,
                  const std::string &shape = std::string()) {
  std::vector<FieldArgs> hfa;
  std::vector<DetArgs> hda;
  int lanes = 1;
  for (auto &r : runs) {
    r->append_args(hfa, hda);
    lanes = std::max(lanes, r->lanes);
  }
  const size_t bfa = std::max<size_t>(hfa.size(), 1) * sizeof(FieldArgs);
  const size_t bda = std::max<size_t>(hda.size(), 1) * sizeof(DetArgs);
  GraphCache *gc = whole_graph ? ctx.graph_cache.get() : nullptr;
  DevBuf dfa, dda;
  if (gc && gc->dfa.bytes == bfa && gc->dda.bytes == bda) {
    dfa = std::move(gc->dfa);
    dda = std::move(gc->dda);
  } else {
    dfa = DevBuf(bfa);
    dda = DevBuf(bda);
  }
  if (!hfa.empty())
    hip_check(hipMemcpyAsync(dfa.p, hfa.data(), hfa.size() * sizeof(FieldArgs), hipMemcpyHostToDevice, ctx.stream),
              "args upload");
  if (!hda.empty())
    hip_check(hipMemcpyAsync(dda.p, hda.data(), hda.size() * sizeof(DetArgs), hipMemcpyHostToDevice, ctx.stream),
              "det args upload");
  std::string lkey;
  if (whole_graph) {
    lkey.append((const char *)&dfa.p, sizeof dfa.p);
    lkey.append((const char *)&dda.p, sizeof dda.p);
    for (auto &r : runs) r->launch_key(lkey);
  }
  const bool replay = gc && gc->exec && gc->launches == lkey;
  std::vector<hipEvent_t> evs;
  auto new_event = [&]() {
    hipEvent_t e = nullptr;
    hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
    evs.push_back(e);
    return e;
  };
  struct EventsFree {
    std::vector<hipEvent_t> &e;
    ~EventsFree() {
      for (hipEvent_t x : e) (void)hipEventDestroy(x);
    }
  } events_free{evs};
  hipGraphExec_t exec = replay ? gc->exec : nullptr;
  if (!replay) {
    hipGraph_t graph = nullptr;
    if (whole_graph) hip_check(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
      // lanes start after the uploads on the context stream and join it at the end
      if (lanes > 1) {
        hipEvent_t ready = new_event();
        hip_check(hipEventRecord(ready, ctx.stream), "event record");
        for (int l = 1; l < lanes; ++l) hip_check(hipStreamWaitEvent(ctx.lane_stream(l), ready, 0), "stream wait");
      }
      size_t ofa = 0, oda = 0;
      for (auto &r : runs) {
        r->enqueue((const FieldArgs *)dfa.p + ofa, (const DetArgs *)dda.p + oda);
        for (auto &ex : r->groups) {
          ofa += ex->fargs.size();
          oda += ex->dargs.size();
        }
      }
      for (int l = 1; l < lanes; ++l) {
        hipEvent_t done = new_event();
        hip_check(hipEventRecord(done, ctx.lane_stream(l)), "event record");
        hip_check(hipStreamWaitEvent(ctx.stream, done, 0), "stream wait");
      }
    } catch (...) {
      if (whole_graph) {
        (void)hipStreamEndCapture(ctx.stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
      }
      throw;
    }
    if (whole_graph) {
      hip_check(hipStreamEndCapture(ctx.stream, &graph), "end capture");
      const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      hip_check(ie, "graph instantiate");
    }
  }
  if (whole_graph) {
    // the cache owns the graph and buffers from here on, replayed or new
    if (!ctx.graph_cache) ctx.graph_cache = std::make_unique<GraphCache>();
    gc = ctx.graph_cache.get();
    if (!replay) {
      if (gc->exec) (void)hipGraphExecDestroy(gc->exec);
      gc->exec = exec;
      gc->launches = lkey;
    }
    hip_check(hipGraphLaunch(gc->exec, ctx.stream), "graph launch");
  }
  for (auto &r : runs) r->download();
  hip_check(hipStreamSynchronize(ctx.stream), "sync");
  if (whole_graph) {
    gc->dfa = std::move(dfa);
    gc->dda = std::move(dda);
    gc->bufs.clear();
    for (auto &r : runs) gc->bufs.push_back(std::move(r->buf));
    gc->shape = shape;
  }
}

// buffers of the cached graph run for the i-th of n runs of this shape
inline RunBuffers *cached_buffers(Context &ctx, const std::string &shape, size_t i, size_t n) {
  GraphCache *gc = ctx.graph_cache.get();
  if (!gc || gc->shape != shape || gc->bufs.size() != n) return nullptr;
  return &gc->bufs[i];
}

template <typename S>
std::vector<double> run_members_t(Context &ctx, const RunConfig &rc, std::vector<Member> &members) {
  hip_check(hipSetDevice(ctx.device), "hipSetDevice");
  const bool graph = rc.use_graph && rc.whole_graph && !rc.fields;
  const std::string shape = graph ? "1|" + BatchRun<S>::shape_key(rc, members) : std::string();
  std::vector<std::unique_ptr<BatchRun<S>>> runs;
  runs.push_back(std::make_unique<BatchRun<S>>(ctx, rc, members));
  runs[0]->setup(graph ? cached_buffers(ctx, shape, 0, 1) : nullptr);
  execute_runs<S>(ctx, runs, graph, shape);
  return runs[0]->finish();
}

template <typename S>
std::vector<std::vector<double>> run_mixed_t(Context &ctx, const std::vector<MixedPart> &parts, bool whole_graph) {
  hip_check(hipSetDevice(ctx.device), "hipSetDevice");
  std::string shape;
  if (whole_graph) {
    shape = std::to_string(parts.size()) + "|";
    for (const MixedPart &p : parts) shape += BatchRun<S>::shape_key(p.rc, *p.members) + ";";
  }
  std::vector<std::unique_ptr<BatchRun<S>>> runs;
  for (size_t i = 0; i < parts.size(); ++i) {
    RunConfig rc = parts[i].rc;
    rc.use_graph = whole_graph;
    rc.whole_graph = whole_graph;
    runs.push_back(std::make_unique<BatchRun<S>>(ctx, rc, *parts[i].members));
    runs.back()->setup(whole_graph ? cached_buffers(ctx, shape, i, parts.size()) : nullptr);
  }
  execute_runs<S>(ctx, runs, whole_graph, shape);
  std::vector<std::vector<double>> out;
  for (auto &r : runs) out.push_back(r->finish());
  return out;
}

}  // namespace

std::vector<double> run_members(Context &ctx, const RunConfig &rc, std::vector<Member> &members) {
  if (members.empty()) return {};
  return rc.f64 ? run_members_t<double>(ctx, rc, members) : run_members_t<float>(ctx, rc, members);
}

std::vector<std::vector<double>> run_mixed(Context &ctx, const std::vector<MixedPart> &parts, bool whole_graph) {
  if (parts.empty()) return {};
  const bool f64 = parts[0].rc.f64;
  for (const MixedPart &p : parts) {
    if (p.rc.f64 != f64) throw Error(1, "the parts of a mixed run share one precision");
    if (!p.members || p.members->empty()) throw Error(1, "a part of a mixed run has no members");
    if (p.rc.fields) throw Error(1, "mixed runs allocate their own fields");
  }
  return f64 ? run_mixed_t<double>(ctx, parts, whole_graph) : run_mixed_t<float>(ctx, parts, whole_graph);
}

Context::Tuned autotune(Context &ctx, int64_t N, bool f64, int reps, int members) {
  int lgN = 0;
  while ((int64_t(1) << lgN) < N) ++lgN;
  members = std::max(members, 1);
  {
    // all members resident at once: stay within half of the free memory
    size_t freeb = 0, totalb = 0;
    hip_check(hipSetDevice(ctx.device), "hipSetDevice");
    hip_check(hipMemGetInfo(&freeb, &totalb), "hipMemGetInfo");
    const size_t per = (size_t)N * (f64 ? 16 : 8);
    members = (int)std::max<size_t>(1, std::min<size_t>((size_t)members, freeb / 2 / per));
  }
  // synthetic slice loop: plane wave, `n` grating slices of a 4.2 um
  // grating at 46 keV, 1 mm detector
  const double dx = 2e-9, wl = energy_to_wavelength(46000.0);
  auto make_member = [&](int n) {
    Member m;
    m.wl = wl;
    Op src{OpKind::Source};
    src.x.kind = XOP_ONES;
    src.x.post[0] = 1.0;
    m.ops.push_back(src);
    for (int i = 0; i < n; ++i) {
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
    return m;
  };
  // Two runs of na and nb slices: their difference is the cost of nb - na
  // slices without what a run pays once (buffers, grating decisions, the
  // detector read), which would otherwise weigh on splits unevenly.
  const int na = reps + 1, nb = 4 * (reps + 1);
  const Member ma = make_member(na), mb = make_member(nb);
  RunConfig rc;
  rc.N = N;
  rc.dx = dx;
  rc.f64 = f64;
  rc.n_out = 1;
  // Time the runner as a batch of `members` fields runs by default (its
  // grouping and lanes), on up to four groups' worth of fields, by wall
  // clock around the run (lanes overlap, so kernel times would not add up).
  const int width = default_group((size_t)N * (f64 ? 16 : 8), members);
  const int key = members;
  const int run_n = std::min(members, 4 * width);
  rc.max_resident = 0;
  const bool was = ctx.timing;
  ctx.timing = false;
  auto time_run = [&](const Member &m) {
    double t_best = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
      std::vector<Member> ms(run_n, m);
      hip_check(hipStreamSynchronize(ctx.stream), "sync");
      const auto t0 = std::chrono::steady_clock::now();
      (void)run_members(ctx, rc, ms);
      t_best =
          std::min(t_best, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    return t_best;
  };
  Context::Tuned best{lgN, f64, -1, -1, 1e300, key};
  for (auto cand : plan_candidates(lgN, f64))
    for (int ta = 0; ta < kColTiles; ++ta)
      for (int tb = 0; tb < kColTiles; ++tb) {
        if (!plan_tiles_valid(lgN, f64, cand.first, cand.second, ta, tb)) continue;
        rc.lgR = cand.first;
        rc.lgA = cand.second;
        rc.tileA = ta;
        rc.tileB = tb;
        {
          std::vector<Member> ms(run_n, ma);
          (void)run_members(ctx, rc, ms);  // warm: plan tables, code objects
        }
        const double t_a = time_run(ma), t_b = time_run(mb);
        double per = (t_b - t_a) / (nb - na) / run_n;
        if (!(per > 0)) per = t_b / nb / run_n;
        if (per < best.ms) best = Context::Tuned{lgN, f64, cand.first, cand.second, per, key, ta, tb};
      }
  ctx.timing = was;
  bool replaced = false;
  for (auto &t : ctx.tuned)
    if (t.lgN == lgN && t.f64 == f64 && (t.members == key || t.members == 0)) {
      t = best;
      replaced = true;
      break;
    }
  if (!replaced) ctx.tuned.push_back(best);
  return best;
}

void tune_save(const Context &ctx, const std::string &path) {
  std::ostringstream o;
  o << "# ravesim tuning table v1: device lgN dtype members lgR lgA tileA tileB ms_per_slice\n";
  o.precision(6);
  for (const auto &t : ctx.tuned)
    o << ctx.device_key << ' ' << t.lgN << ' ' << (t.f64 ? "c16" : "c8") << ' ' << t.members << ' ' << t.lgR << ' '
      << t.lgA << ' ' << t.tileA << ' ' << t.tileB << ' ' << t.ms << '\n';
  const std::string txt = o.str();
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) throw Error(3, "cannot write " + path);
  const size_t n = std::fwrite(txt.data(), 1, txt.size(), f);
  const bool ok = n == txt.size() && std::fclose(f) == 0;
  if (!ok) throw Error(3, "write failed: " + path);
}

int tune_load(Context &ctx, const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) throw Error(3, "cannot read " + path);
  std::string txt;
  char buf[4096];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) txt.append(buf, n);
  std::fclose(f);
  std::istringstream in(txt);
  std::string line;
  int taken = 0, lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ls(line);
    std::string dev, dt;
    Context::Tuned t{};
    if (!(ls >> dev >> t.lgN >> dt >> t.members >> t.lgR >> t.lgA >> t.tileA >> t.tileB >> t.ms) ||
        (dt != "c8" && dt != "c16"))
      throw Error(2, path + ":" + std::to_string(lineno) + ": malformed tuning entry");
    if (dev != ctx.device_key) continue;
    t.f64 = dt == "c16";
    if (t.lgN < kMinRowLg || t.lgN > 30 || t.members < 0 || t.tileA < 0 || t.tileA >= kColTiles || t.tileB < 0 ||
        t.tileB >= kColTiles || !plan_tiles_valid(t.lgN, t.f64, t.lgR, t.lgA, t.tileA, t.tileB))
      throw Error(2, path + ":" + std::to_string(lineno) + ": invalid split or tiles");
    auto &tv = ctx.tuned;
    tv.erase(std::remove_if(tv.begin(), tv.end(),
                            [&](const Context::Tuned &u) {
                              return u.lgN == t.lgN && u.f64 == t.f64 && u.members == t.members;
                            }),
             tv.end());
    tv.push_back(t);
    ++taken;
  }
  return taken;
}

}  // namespace rs
