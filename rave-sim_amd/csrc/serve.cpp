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

}  // namespace

// device / pinned buffers reused across jobs (hipFree would stall the stream)
struct Server::Pool {
  std::map<size_t, std::vector<void *>> dev, host;
  size_t dev_bytes = 0;
  void *get_dev(size_t b) {
    auto &v = dev[b];
    if (!v.empty()) {
      void *p = v.back();
      v.pop_back();
      return p;
    }
    void *p = nullptr;
    hip_check(hipMalloc(&p, b), "hipMalloc(job)");
    dev_bytes += b;
    return p;
  }
  void put_dev(size_t b, void *p) { dev[b].push_back(p); }
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
  // release cached device buffers of sizes other than `keep` until `need` fits
  void trim(size_t budget, size_t need) {
    for (auto &kv : dev) {
      while (!kv.second.empty() && dev_bytes + need > budget) {
        (void)hipFree(kv.second.back());
        kv.second.pop_back();
        dev_bytes -= kv.first;
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

Server::Server(Context &ctx, double reserve_bytes, int max_jobs)
    : ctx_(ctx), max_jobs_(max_jobs), pool_(std::make_unique<Pool>()) {
  hip_check(hipSetDevice(ctx_.device), "hipSetDevice");
  size_t freeb = 0, totalb = 0;
  hip_check(hipMemGetInfo(&freeb, &totalb), "hipMemGetInfo");
  const size_t reserve = reserve_bytes > 0 ? (size_t)reserve_bytes : (size_t)(4ull << 30);
  budget_ = freeb > reserve ? freeb - reserve : freeb / 2;
  if (max_jobs_ <= 0 || max_jobs_ > 4096) max_jobs_ = 4096;
  for (int i = 0; i < 3; ++i) {
    auto s = std::make_unique<Slot>();
    hip_check(hipEventCreateWithFlags(&s->ev, hipEventDisableTiming), "hipEventCreate");
    slots_.push_back(std::move(s));
  }
  th_ = std::thread([this] { loop(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (th_.joinable()) th_.join();
  (void)hipStreamSynchronize(ctx_.stream);
  for (auto &j : active_) {
    if (j->done_ev) (void)hipEventDestroy(j->done_ev);
  }
  for (auto &s : slots_) {
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
      const size_t need = j->field_bytes * (stepped ? 2 : 1) + j->acc_bytes;
      size_t in_use = 0;
      for (auto &a : active_) in_use += a->bytes();
      if (in_use + need > budget_) return;  // wait for running jobs to finish
      if (pool.dev_bytes + need > budget_) pool.trim(budget_, need);
      j->field = pool.get_dev(j->field_bytes);
      if (stepped) j->snap = pool.get_dev(j->field_bytes);
      j->acc = (double *)pool.get_dev(j->acc_bytes);
      j->host_acc = (double *)pool.get_host(j->acc_bytes);
      hip_check(launch_fill_f64(j->acc, (int64_t)(j->acc_bytes / sizeof(double)), 0.0, ctx_.stream), "acc fill");
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
      active_.push_back(std::move(j));
      stats_.max_active = std::max<int64_t>(stats_.max_active, (int64_t)active_.size());
    }
    std::lock_guard<std::mutex> lk(mu_);
    pending_.pop_front();
  }
}

bool Server::tick() {
  using exec::LKind;
  hipStream_t st = ctx_.stream;
  struct Key {
    int kind;
    bool f64;
    int64_t N;
    int lgL;
    bool row;
    int top;
    int xv;
    int64_t S;
    const void *tab;
    bool operator<(const Key &o) const {
      return std::tie(kind, f64, N, lgL, row, top, xv, S, tab) <
             std::tie(o.kind, o.f64, o.N, o.lgL, o.row, o.top, o.xv, o.S, o.tab);
    }
  };
  struct Group {
    std::vector<Job *> jobs;
    const exec::Launch *l = nullptr;
    size_t off = 0;
  };
  std::map<Key, Group> groups;
  bool any = false;
  for (auto &jp : active_) {
    Job &j = *jp;
    if (j.cursor >= j.prog.size()) continue;
    any = true;
    const exec::Launch &l = j.prog[j.cursor];
    if (l.kind == LKind::Copy || l.kind == LKind::Fill) {
      hip_check(j.f64 ? launch_copy<double>(l.dst, l.src, l.n, st) : launch_copy<float>(l.dst, l.src, l.n, st),
                "copy");
      stats_.launches++;
      stats_.merged_jobs++;
      continue;
    }
    Key k{(int)l.kind, j.f64, j.N, l.lgL, l.row, l.geom.top, l.xv, l.geom.S,
          l.kind == LKind::Detect ? (const void *)j.req->bounds : l.geom.tw_fib};
    Group &g = groups[k];
    g.jobs.push_back(&j);
    g.l = &l;
  }
  if (!any) return false;
  // stage all argument records of this tick in one slot
  Slot &s = *slots_[slot_i_];
  slot_i_ = (slot_i_ + 1) % slots_.size();
  if (s.used) hip_check(hipEventSynchronize(s.ev), "slot wait");
  size_t need = 0;
  for (auto &kv : groups)
    need += kv.second.jobs.size() * (kv.first.kind == (int)LKind::Detect ? sizeof(DetArgs) : sizeof(FieldArgs)) + 64;
  if (need > s.cap) {
    if (s.host) (void)hipHostFree(s.host);
    if (s.dev) (void)hipFree(s.dev);
    s.cap = std::max(need * 2, (size_t)1 << 16);
    hip_check(hipHostMalloc(&s.host, s.cap, hipHostMallocDefault), "hipHostMalloc(args)");
    hip_check(hipMalloc(&s.dev, s.cap), "hipMalloc(args)");
  }
  size_t off = 0;
  for (auto &kv : groups) {
    Group &g = kv.second;
    off = (off + 63) & ~(size_t)63;
    g.off = off;
    for (Job *j : g.jobs) {
      const exec::Launch &l = j->prog[j->cursor];
      if (kv.first.kind == (int)LKind::Detect) {
        std::memcpy((char *)s.host + off, &j->dargs[l.arg_off], sizeof(DetArgs));
        off += sizeof(DetArgs);
      } else {
        std::memcpy((char *)s.host + off, &j->fargs[l.arg_off], sizeof(FieldArgs));
        off += sizeof(FieldArgs);
      }
    }
  }
  hip_check(hipMemcpyAsync(s.dev, s.host, off, hipMemcpyHostToDevice, st), "args upload");
  for (auto &kv : groups) {
    Group &g = kv.second;
    const int nf = (int)g.jobs.size();
    const char *dp = (const char *)s.dev + g.off;
    const exec::Launch &l = *g.l;
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
  hip_check(hipEventRecord(s.ev, st), "slot event");
  s.used = true;
  // advance every job by one pass; finished jobs read back their rows
  for (auto &jp : active_) {
    Job &j = *jp;
    if (j.cursor >= j.prog.size()) continue;
    j.cursor++;
    if (j.cursor == j.prog.size() && !j.enqueued_all) {
      hip_check(hipMemcpyAsync(j.host_acc, j.acc, j.acc_bytes, hipMemcpyDeviceToHost, st), "acc readback");
      hip_check(hipEventRecord(j.done_ev, st), "done event");
      j.enqueued_all = true;
    }
  }
  stats_.ticks++;
  return true;
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
    r->member_rows[j.member].assign(j.host_acc, j.host_acc + n);
    pool.put_dev(j.field_bytes, j.field);
    if (j.snap) pool.put_dev(j.field_bytes, j.snap);
    pool.put_dev(j.acc_bytes, j.acc);
    pool.put_host(j.acc_bytes, j.host_acc);
    (void)hipEventDestroy(j.done_ev);
    r->done++;
    stats_.jobs_done++;
    if (r->done == r->r.members.size()) {
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
      const bool worked = tick();
      reap();
      if (!worked && !active_.empty()) {
        // every admitted job is waiting for its readback
        (void)hipStreamSynchronize(ctx_.stream);
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
