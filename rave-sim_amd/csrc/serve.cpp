// Serving scheduler (see serve.hpp for the model).
#include "serve.hpp"

#include <algorithm>
#include <chrono>

#include "exec.hpp"
#include "rs_kernels.hpp"

namespace rs {

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// fields per launch assumed when picking a tuned split for served jobs
constexpr int kServeBatchHint = 16;
// lanes (streams) the grid shapes are spread over, argument slots per lane
constexpr int kLanes = 4;
constexpr int kSlotsPerLane = 3;
// a job of N samples advances by kPaceSamples / N passes per tick (at least
// one, at most kMaxPassesPerTick): about the same device time per tick
constexpr int64_t kPaceSamples = int64_t(1) << 26;
constexpr int kMaxPassesPerTick = 64;
// lanes of larger grids wait while a lane of a smaller grid has passes to
// issue (shortest jobs first: a 2^26 request does not share the device with
// a 2^30 one), except for one tick per kDeferTickS seconds, so that they
// keep moving under a steady stream of small requests
constexpr double kDeferTickS = 1.0;

}  // namespace

// device / pinned buffers reused across jobs (hipFree would stall the stream)
struct Server::Pool {
  std::map<size_t, std::vector<void *>> dev, host;
  size_t dev_bytes = 0;   // allocated by the pool, in use or idle
  size_t idle_bytes = 0;  // cached in the free lists
  void *get_dev(size_t b) {
    auto &v = dev[b];
    if (!v.empty()) {
      void *p = v.back();
      v.pop_back();
      idle_bytes -= b;
      return p;
    }
    void *p = nullptr;
    hip_check(hipMalloc(&p, b), "hipMalloc(job)");
    dev_bytes += b;
    return p;
  }
  void put_dev(size_t b, void *p) {
    dev[b].push_back(p);
    idle_bytes += b;
  }
  void *get_host(size_t b) {
    auto &v = host[b];
    if (!v.empty()) {
      void *p = v.back();
      v.pop_back();
      return p;
    }
    void *p = nullptr;
    hip_check(hipHostMalloc(&p, b, hipHostMallocDefault), "hipHostMalloc");
    return p;
  }
  void put_host(size_t b, void *p) { host[b].push_back(p); }
  // bytes of the buffers `sizes` that idle buffers of exactly those sizes
  // would supply (no new allocation)
  size_t reusable(const std::vector<size_t> &sizes) const {
    std::map<size_t, size_t> want;
    for (size_t b : sizes) want[b]++;
    size_t r = 0;
    for (auto &kv : want) {
      auto it = dev.find(kv.first);
      const size_t have = it == dev.end() ? 0 : it->second.size();
      r += kv.first * std::min(have, kv.second);
    }
    return r;
  }
  // release idle cached device buffers, other than one of each of `keep`'s
  // sizes per entry, until at most `want` bytes are idle
  void trim(size_t want, const std::vector<size_t> &keep) {
    for (auto &kv : dev) {
      const size_t k = (size_t)std::count(keep.begin(), keep.end(), kv.first);
      while (kv.second.size() > k && idle_bytes > want) {
        (void)hipFree(kv.second.back());
        kv.second.pop_back();
        dev_bytes -= kv.first;
        idle_bytes -= kv.first;
      }
    }
  }
  ~Pool() {
    for (auto &kv : dev)
      for (void *p : kv.second) (void)hipFree(p);
    for (auto &kv : host)
      for (void *p : kv.second) (void)hipHostFree(p);
  }
};


struct Server::Req {
  ServeRequest r;
  ServeResult res;
  size_t admitted = 0, done = 0;
  std::vector<std::vector<double>> member_rows;
  std::vector<int64_t> bounds_host;
  void *bounds = nullptr;  // device
  int n_running = 0;
};

struct Server::Job {
  std::shared_ptr<Req> req;
  size_t member = 0;
  bool f64 = false;
  int64_t N = 0;
  int lane = 0;
  int pace = 1;  // passes per tick
  Member m;
  void *field = nullptr, *snap = nullptr;
  double *acc = nullptr;
  double *host_acc = nullptr;
  size_t field_bytes = 0, acc_bytes = 0;
  std::vector<exec::Launch> prog;
  std::vector<FieldArgs> fargs;
  std::vector<DetArgs> dargs;
  size_t cursor = 0;
  bool enqueued_all = false;
  hipEvent_t done_ev = nullptr;
  size_t bytes() const { return field_bytes * (snap ? 2 : 1) + acc_bytes; }
};

struct Server::Slot {
  void *host = nullptr, *dev = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool used = false;
};

struct Server::Lane {
  hipStream_t st = nullptr;
  std::vector<std::unique_ptr<Slot>> slots;
  size_t next = 0;
  size_t bytes = 0;          // resident bytes of its jobs
  double deferred_at = 0.0;  // when it last started waiting for a smaller grid's lane (0: not waiting)
};

Server::Server(Context &ctx, double reserve_bytes, int max_jobs)
    : ctx_(ctx), max_jobs_(max_jobs), pool_(std::make_unique<Pool>()) {
  hip_check(hipSetDevice(ctx_.device), "hipSetDevice");
  reserve_ = reserve_bytes > 0 ? (size_t)reserve_bytes : (size_t)(4ull << 30);
  if (max_jobs_ <= 0 || max_jobs_ > 4096) max_jobs_ = 4096;
  for (int l = 0; l < kLanes; ++l) {
    auto ln = std::make_unique<Lane>();
    ln->st = ctx_.lane_stream(l);
    for (int i = 0; i < kSlotsPerLane; ++i) {
      auto s = std::make_unique<Slot>();
      hip_check(hipEventCreateWithFlags(&s->ev, hipEventDisableTiming), "hipEventCreate");
      ln->slots.push_back(std::move(s));
    }
    lanes_.push_back(std::move(ln));
  }
  th_ = std::thread([this] { loop(); });
}

int Server::lane_of(int64_t N, bool f64) {
  const auto key = std::make_pair(N, f64);
  auto it = shape_lane_.find(key);
  if (it != shape_lane_.end()) return it->second;
  // a new shape takes the lane that holds the fewest shapes, then bytes
  std::vector<int> shapes(lanes_.size(), 0);
  for (auto &kv : shape_lane_) shapes[(size_t)kv.second]++;
  int best = 0;
  for (int l = 1; l < (int)lanes_.size(); ++l)
    if (shapes[l] < shapes[best] || (shapes[l] == shapes[best] && lanes_[l]->bytes < lanes_[best]->bytes)) best = l;
  shape_lane_[key] = best;
  return best;
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (th_.joinable()) th_.join();
  for (auto &ln : lanes_) (void)hipStreamSynchronize(ln->st);
  for (auto &j : active_) {
    if (j->done_ev) (void)hipEventDestroy(j->done_ev);
  }
  for (auto &ln : lanes_)
    for (auto &s : ln->slots) {
      if (s->ev) (void)hipEventDestroy(s->ev);
      if (s->host) (void)hipHostFree(s->host);
      if (s->dev) (void)hipFree(s->dev);
    }
  for (auto &kv : reqs_)
    if (kv.second->bounds) (void)hipFree(kv.second->bounds);
  pool_.reset();
}

int64_t Server::submit(ServeRequest req) {
  if (req.members.empty()) throw Error(1, "request has no members");
  const SimConfig &c0 = req.members[0];
  if (req.weights.empty()) req.weights.assign(req.members.size(), 1.0);
  if (req.weights.size() != req.members.size()) throw Error(1, "one weight per member");
  int nsteps0 = -1;
  for (const auto &c : req.members) {
    if (c.N != c0.N || c.f64 != c0.f64 || c.dx != c0.dx || c.detector_size != c0.detector_size ||
        c.px != c0.px || c.z_detector != c0.z_detector)
      throw Error(1, "all members of a request share N, precision, grid and detector");
    if (c.source.kind == SourceCfg::Vector) throw Error(6, "served requests use point or plane-wave sources");
    for (const auto &g : c.elements)
      if (g.sample) throw Error(6, "served requests hold gratings only");
    Member tmp;
    const int ns = build_member(c, tmp, 0, false);  // validates the configuration
    if (nsteps0 < 0) nsteps0 = ns;
    if (ns != nsteps0) throw Error(1, "members differ in phase steps");
  }
  auto r = std::make_shared<Req>();
  r->res.n_steps = nsteps0;
  r->res.n_px = detector_pixels(c0.detector_size, c0.px);
  r->res.submit_t = now_s();
  r->member_rows.resize(req.members.size());
  r->r = std::move(req);
  int64_t id;
  {
    std::lock_guard<std::mutex> lk(mu_);
    id = next_id_++;
    reqs_[id] = r;
    pending_.push_back(id);
  }
  cv_.notify_all();
  return id;
}

ServeResult Server::poll(int64_t id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) throw Error(1, "unknown request id");
  return it->second->res;
}

bool Server::wait(int64_t id, double timeout_s) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) throw Error(1, "unknown request id");
  auto r = it->second;
  return done_cv_.wait_for(lk, std::chrono::duration<double>(timeout_s),
                           [&] { return r->res.state >= 2; });
}

void Server::drain() {
  std::unique_lock<std::mutex> lk(mu_);
  done_cv_.wait(lk, [&] {
    if (!pending_.empty()) return false;
    for (auto &kv : reqs_)
      if (kv.second->res.state < 2) return false;
    return true;
  });
}

ServeStats Server::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  return stats_;
}

template <typename S> void Server::compile_job(Job &j) {
  const SimConfig &c = j.req->r.members[j.member];
  RunConfig rc;
  rc.N = c.N;
  rc.dx = c.dx;
  rc.f64 = c.f64;
  rc.mode = 0;
  rc.det_size = c.detector_size;
  rc.det_px = c.px;
  rc.det_py = c.py;
  rc.z_det = c.z_detector;
  rc.n_out = (int)j.req->res.n_steps;
  // jobs of one size are merged into shared launches: use the split tuned
  // for batched launches when there is one
  Plan &plan = ctx_.plan(c.N, c.f64, -1, -1, kServeBatchHint);
  exec::Exec<S> ex(ctx_, rc, plan);
  // a field larger than the last-level cache streams from HBM in its top passes
  ex.stream_top = (size_t)j.N * 2 * sizeof(S) > kMallBytes;
  ex.mem = {&j.m};
  ex.field = {j.field};
  if (j.snap) ex.snap = {j.snap};
  ex.acc = j.acc;
  ex.bounds = (const int64_t *)j.req->bounds;
  ex.npx = j.req->res.n_px;
  ex.compile();
  j.prog = std::move(ex.launches);
  j.fargs = std::move(ex.fargs);
  j.dargs = std::move(ex.dargs);
}

void Server::admit() {
  Pool &pool = *pool_;
  while (true) {
    std::shared_ptr<Req> r;
    int64_t rid = 0;
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (pending_.empty()) return;
      rid = pending_.front();
      r = reqs_[rid];
    }
    bool failed = false;
    if (!r->bounds) {
      const SimConfig &c = r->r.members[0];
      r->bounds_host = detector_bounds(c.N, c.dx, c.detector_size, c.px);
      hip_check(hipMalloc(&r->bounds, r->bounds_host.size() * sizeof(int64_t)), "hipMalloc(bounds)");
      hip_check(hipMemcpy(r->bounds, r->bounds_host.data(), r->bounds_host.size() * sizeof(int64_t),
                          hipMemcpyHostToDevice),
                "bounds upload");
    }
    while (r->admitted < r->r.members.size()) {
      if ((int)active_.size() >= max_jobs_) return;
      const SimConfig &c = r->r.members[r->admitted];
      auto j = std::make_unique<Job>();
      j->req = r;
      j->member = r->admitted;
      j->f64 = c.f64;
      j->N = c.N;
      const int ns = build_member(c, j->m, 0, false);
      bool stepped = false;
      for (auto &o : j->m.ops) stepped |= o.kind == OpKind::Restore;
      j->field_bytes = (size_t)c.N * (c.f64 ? 16 : 8);
      j->acc_bytes = (size_t)ns * std::max<int64_t>(r->res.n_px, 1) * sizeof(double);
      std::vector<size_t> sizes{j->field_bytes, j->acc_bytes};
      if (stepped) sizes.push_back(j->field_bytes);
      const size_t need = j->field_bytes * (stepped ? 2 : 1) + j->acc_bytes;
      // idle buffers of the right sizes are reused; the rest must come from
      // free device memory plus the other idle buffers (released on demand),
      // less the reserve
      const size_t reuse = pool.reusable(sizes);
      const size_t fresh = need - reuse;
      size_t freeb = 0, totalb = 0;
      hip_check(hipMemGetInfo(&freeb, &totalb), "hipMemGetInfo");
      const size_t other_idle = pool.idle_bytes - reuse;
      const size_t avail = freeb + other_idle > reserve_ ? freeb + other_idle - reserve_ : 0;
      if (fresh > avail) {
        if (!active_.empty()) return;  // wait for running jobs to finish
        // nothing else is resident: this job can never fit
        std::lock_guard<std::mutex> lk(mu_);
        r->res.state = 3;
        r->res.error = "request needs " + std::to_string(need >> 20) + " MiB per energy, " +
                       std::to_string((avail + reuse) >> 20) + " MiB can be allocated (RS_ERR_NOMEM)";
        r->res.done_t = now_s();
        r->member_rows.clear();
        done_cv_.notify_all();
        failed = true;
        break;
      }
      if (fresh > 0 && freeb < fresh + reserve_) {
        // release idle buffers of other sizes until the fresh part fits
        const size_t drop = fresh + reserve_ - freeb;
        pool.trim(pool.idle_bytes > drop ? pool.idle_bytes - drop : 0, sizes);
      }
      j->lane = lane_of(c.N, c.f64);
      j->pace = (int)std::max<int64_t>(1, std::min<int64_t>(kMaxPassesPerTick, kPaceSamples / c.N));
      hipStream_t st = lanes_[(size_t)j->lane]->st;
      j->field = pool.get_dev(j->field_bytes);
      if (stepped) j->snap = pool.get_dev(j->field_bytes);
      j->acc = (double *)pool.get_dev(j->acc_bytes);
      j->host_acc = (double *)pool.get_host(j->acc_bytes);
      hip_check(launch_fill_f64(j->acc, (int64_t)(j->acc_bytes / sizeof(double)), 0.0, st), "acc fill");
      hip_check(hipEventCreateWithFlags(&j->done_ev, hipEventDisableTiming), "hipEventCreate");
      if (c.f64) compile_job<double>(*j);
      else compile_job<float>(*j);
      r->admitted++;
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (r->res.state == 0) {
          r->res.state = 1;
          r->res.start_t = now_s();
        }
      }
      lanes_[(size_t)j->lane]->bytes += j->bytes();
      active_.push_back(std::move(j));
      stats_.max_active = std::max<int64_t>(stats_.max_active, (int64_t)active_.size());
    }
    std::lock_guard<std::mutex> lk(mu_);
    pending_.pop_front();
    (void)failed;
  }
}

bool Server::tick() {
  using exec::LKind;
  struct Key {
    int kind;
    bool f64;
    int64_t N;
    int lgL;
    bool row;
    int top;
    int xv;
    int64_t S;
    int64_t span_off, span_len;
    const void *tab;
    bool operator<(const Key &o) const {
      return std::tie(kind, f64, N, lgL, row, top, xv, S, span_off, span_len, tab) <
             std::tie(o.kind, o.f64, o.N, o.lgL, o.row, o.top, o.xv, o.S, o.span_off, o.span_len, o.tab);
    }
  };
  struct Group {
    std::vector<Job *> jobs;
    std::vector<size_t> at;  // launch index of each job
    size_t off = 0;
  };
  bool issued = false;
  // the largest grid with passes to issue in each lane, and the smallest of those
  std::vector<int64_t> lane_n(lanes_.size(), 0);
  for (auto &jp : active_)
    if (jp->cursor < jp->prog.size()) lane_n[(size_t)jp->lane] = std::max(lane_n[(size_t)jp->lane], jp->N);
  int64_t min_n = 0;
  for (int64_t n : lane_n)
    if (n > 0 && (min_n == 0 || n < min_n)) min_n = n;
  const double now = now_s();
  for (size_t li = 0; li < lanes_.size(); ++li) {
    Lane &ln = *lanes_[li];
    std::vector<Job *> js;
    int reps = 0;
    for (auto &jp : active_)
      if (jp->lane == (int)li && jp->cursor < jp->prog.size()) {
        js.push_back(jp.get());
        reps = std::max(reps, jp->pace);
      }
    if (js.empty()) {
      ln.deferred_at = 0.0;
      continue;
    }
    bool deferred_tick = false;
    if (lane_n[li] > min_n) {
      if (ln.deferred_at == 0.0) ln.deferred_at = now;
      if (now - ln.deferred_at < kDeferTickS) continue;
      deferred_tick = true;
    } else {
      ln.deferred_at = 0.0;
    }
    Slot &s = *ln.slots[ln.next];
    if (s.used) {
      if (hipEventQuery(s.ev) == hipErrorNotReady) continue;  // every slot of the lane is in flight
      hip_check(hipEventSynchronize(s.ev), "slot wait");
    }
    if (deferred_tick) ln.deferred_at = now;  // one tick now, the next after another kDeferTickS
    ln.next = (ln.next + 1) % ln.slots.size();
    hipStream_t st = ln.st;
    // sub-step r: the r-th next pass of every job that advances that far
    // this tick; copies and fills run on their own, the rest merge by kernel
    std::vector<std::map<Key, Group>> steps;
    std::vector<std::vector<std::pair<Job *, size_t>>> singles;
    for (int r = 0; r < reps; ++r) {
      std::map<Key, Group> groups;
      std::vector<std::pair<Job *, size_t>> one;
      for (Job *j : js) {
        const size_t c = j->cursor + (size_t)r;
        if (r >= j->pace || c >= j->prog.size()) continue;
        const exec::Launch &l = j->prog[c];
        if (l.kind == LKind::Copy || l.kind == LKind::Fill) {
          one.emplace_back(j, c);
          continue;
        }
        Key k{(int)l.kind, j->f64, j->N, l.lgL, l.row, l.geom.top, l.xv, l.geom.S, l.geom.span_off, l.geom.span_len,
              l.kind == LKind::Detect ? (const void *)j->req->bounds : l.geom.tw_fib};
        Group &g = groups[k];
        g.jobs.push_back(j);
        g.at.push_back(c);
      }
      if (groups.empty() && one.empty()) break;
      steps.push_back(std::move(groups));
      singles.push_back(std::move(one));
    }
    // stage the argument records of every merged launch in the lane's slot
    size_t need = 0;
    for (auto &groups : steps)
      for (auto &kv : groups)
        need += kv.second.jobs.size() * (kv.first.kind == (int)LKind::Detect ? sizeof(DetArgs) : sizeof(FieldArgs)) +
                64;
    if (need > s.cap) {
      if (s.host) (void)hipHostFree(s.host);
      if (s.dev) (void)hipFree(s.dev);
      s.host = s.dev = nullptr;
      s.cap = 0;
      const size_t cap = std::max(need * 2, (size_t)1 << 16);
      hip_check(hipHostMalloc(&s.host, cap, hipHostMallocDefault), "hipHostMalloc(args)");
      hip_check(hipMalloc(&s.dev, cap), "hipMalloc(args)");
      s.cap = cap;
    }
    size_t off = 0;
    for (auto &groups : steps)
      for (auto &kv : groups) {
        Group &g = kv.second;
        off = (off + 63) & ~(size_t)63;
        g.off = off;
        for (size_t i = 0; i < g.jobs.size(); ++i) {
          const exec::Launch &l = g.jobs[i]->prog[g.at[i]];
          if (kv.first.kind == (int)LKind::Detect) {
            std::memcpy((char *)s.host + off, &g.jobs[i]->dargs[l.arg_off], sizeof(DetArgs));
            off += sizeof(DetArgs);
          } else {
            std::memcpy((char *)s.host + off, &g.jobs[i]->fargs[l.arg_off], sizeof(FieldArgs));
            off += sizeof(FieldArgs);
          }
        }
      }
    if (off) hip_check(hipMemcpyAsync(s.dev, s.host, off, hipMemcpyHostToDevice, st), "args upload");
    for (size_t r = 0; r < steps.size(); ++r) {
      for (auto &jc : singles[r]) {
        const exec::Launch &l = jc.first->prog[jc.second];
        if (l.kind == LKind::Fill)
          hip_check(launch_fill_f64((double *)l.dst, l.n, 0.0, st), "fill");
        else
          hip_check(jc.first->f64 ? launch_copy<double>(l.dst, l.src, l.n, st)
                                  : launch_copy<float>(l.dst, l.src, l.n, st),
                    "copy");
        stats_.launches++;
        stats_.merged_jobs++;
      }
      for (auto &kv : steps[r]) {
        Group &g = kv.second;
        const int nf = (int)g.jobs.size();
        const char *dp = (const char *)s.dev + g.off;
        const exec::Launch &l = g.jobs[0]->prog[g.at[0]];
        const bool f64 = kv.first.f64;
        switch (l.kind) {
          case LKind::Fiber:
            hip_check(f64 ? launch_fiber_pass<double>(l.lgL, l.row, l.xv, (const FieldArgs *)dp, nf, l.geom, st)
                          : launch_fiber_pass<float>(l.lgL, l.row, l.xv, (const FieldArgs *)dp, nf, l.geom, st),
                      "fiber pass");
            break;
          case LKind::Xop:
            hip_check(f64 ? launch_xop<double>((const FieldArgs *)dp, nf, kv.first.N, st)
                          : launch_xop<float>((const FieldArgs *)dp, nf, kv.first.N, st),
                      "xop");
            break;
          case LKind::Kop:
            hip_check(f64 ? launch_kop<double>((const FieldArgs *)dp, nf, l.geom, st)
                          : launch_kop<float>((const FieldArgs *)dp, nf, l.geom, st),
                      "kop");
            break;
          case LKind::Detect: {
            Job *j0 = g.jobs[0];
            hip_check(f64 ? launch_detect<double>((const DetArgs *)dp, nf, (const int64_t *)j0->req->bounds,
                                                  j0->req->res.n_px, st)
                          : launch_detect<float>((const DetArgs *)dp, nf, (const int64_t *)j0->req->bounds,
                                                 j0->req->res.n_px, st),
                      "detect");
            break;
          }
          default:
            throw Error(6, "unsupported launch kind in serving");
        }
        stats_.launches++;
        stats_.merged_jobs += nf;
      }
    }
    hip_check(hipEventRecord(s.ev, st), "slot event");
    s.used = true;
    issued = true;
    // advance the lane's jobs; finished jobs read back their rows
    for (Job *j : js) {
      j->cursor = std::min(j->prog.size(), j->cursor + (size_t)j->pace);
      if (j->cursor == j->prog.size() && !j->enqueued_all) {
        hip_check(hipMemcpyAsync(j->host_acc, j->acc, j->acc_bytes, hipMemcpyDeviceToHost, st), "acc readback");
        hip_check(hipEventRecord(j->done_ev, st), "done event");
        j->enqueued_all = true;
      }
    }
    stats_.ticks++;
  }
  return issued;
}

void Server::reap() {
  Pool &pool = *pool_;
  for (size_t i = 0; i < active_.size();) {
    Job &j = *active_[i];
    if (!j.enqueued_all || hipEventQuery(j.done_ev) != hipSuccess) {
      ++i;
      continue;
    }
    auto r = j.req;
    const size_t n = j.acc_bytes / sizeof(double);
    if (r->res.state != 3) r->member_rows[j.member].assign(j.host_acc, j.host_acc + n);
    lanes_[(size_t)j.lane]->bytes -= j.bytes();
    pool.put_dev(j.field_bytes, j.field);
    if (j.snap) pool.put_dev(j.field_bytes, j.snap);
    pool.put_dev(j.acc_bytes, j.acc);
    pool.put_host(j.acc_bytes, j.host_acc);
    (void)hipEventDestroy(j.done_ev);
    r->done++;
    stats_.jobs_done++;
    if (r->done == r->r.members.size() && r->res.state != 3) {
      // weighted sum in member order, then the detector geometry factors
      const SimConfig &c = r->r.members[0];
      const int64_t npx = r->res.n_px, ns = r->res.n_steps;
      std::vector<double> sum((size_t)(ns * npx), 0.0);
      for (size_t m = 0; m < r->member_rows.size(); ++m)
        for (size_t k = 0; k < sum.size(); ++k) sum[k] += r->r.weights[m] * r->member_rows[m][k];
      const std::vector<double> sc = detector_scale(c.detector_size, c.px, c.py, c.dx, c.z_detector);
      for (int64_t s = 0; s < ns; ++s)
        for (int64_t k = 0; k < npx; ++k) {
          double &v = sum[(size_t)(s * npx + k)];
          v = (((v * sc[2 * k]) * c.dx) * c.py) / sc[2 * k + 1];
        }
      std::lock_guard<std::mutex> lk(mu_);
      r->res.rows = std::move(sum);
      r->res.state = 2;
      r->res.done_t = now_s();
      r->member_rows.clear();
      done_cv_.notify_all();
    }
    active_.erase(active_.begin() + (long)i);
  }
}

void Server::loop() {
  (void)hipSetDevice(ctx_.device);
  while (true) {
    {
      std::unique_lock<std::mutex> lk(mu_);
      if (stop_ && pending_.empty() && active_.empty()) break;
      if (pending_.empty() && active_.empty()) {
        cv_.wait_for(lk, std::chrono::milliseconds(50));
        continue;
      }
    }
    try {
      const double t0 = now_s();
      admit();
      const bool issued = tick();
      reap();
      if (!issued && !active_.empty()) {
        // every lane with work has all of its slots in flight, or every job
        // waits for its readback: block on the oldest of that work
        bool waited = false;
        for (size_t li = 0; li < lanes_.size() && !waited; ++li) {
          Slot &s = *lanes_[li]->slots[lanes_[li]->next];
          bool busy = false;
          for (auto &jp : active_) busy |= jp->lane == (int)li && jp->cursor < jp->prog.size();
          if (busy && s.used) {
            hip_check(hipEventSynchronize(s.ev), "slot wait");
            waited = true;
          }
        }
        if (!waited) {
          for (auto &jp : active_)
            if (jp->enqueued_all) {
              hip_check(hipEventSynchronize(jp->done_ev), "job wait");
              break;
            }
        }
        reap();
      }
      stats_.busy_s += now_s() - t0;
    } catch (const std::exception &e) {
      std::lock_guard<std::mutex> lk(mu_);
      for (auto &kv : reqs_)
        if (kv.second->res.state < 2) {
          kv.second->res.state = 3;
          kv.second->res.error = e.what();
        }
      pending_.clear();
      active_.clear();
      done_cv_.notify_all();
    }
  }
}

}  // namespace rs
