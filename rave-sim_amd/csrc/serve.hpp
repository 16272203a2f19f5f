// Serving: continuous batching of simulation requests on one device.
//
// A request is a phase-stepping sweep (or a single exposure) of one setup at
// one or more energies.  Each (request, energy) pair is a *job*: one member
// program (simulation.cpp build_member) with its own field, snapshot and
// detector rows in device memory.  Jobs are admitted in arrival order while
// their memory fits: free device memory now (hipMemGetInfo, read at every
// admission) plus the scheduler's idle cached buffers, minus a reserve.  A
// job that cannot fit even with nothing else resident fails its request
// (RS_ERR_NOMEM) instead of blocking the queue.  Fields stay resident from
// admission to completion.
//
// Jobs run on lanes: each grid shape (N, precision) is given one of up to
// kLanes streams, so jobs of different sizes overlap on the device.  On every
// tick the scheduler gives each lane whose argument slots are free its next
// passes: a job of N samples advances by about 2^26 / N passes (1 to 64), so
// small fields are not paced by the launches of large ones, and the passes of
// all jobs of a lane that need the same kernel (same plan, pass kind, fiber
// length and span) merge into one launch (blockIdx.y = job).  Lanes of the
// smallest grid with passes to issue go first: a lane of a larger grid waits
// while one of a smaller grid has work, but for one tick a second, so short
// requests finish without sharing the device with long ones.  A job admitted
// while others are mid-way simply joins the next tick's launches.  Every job
// writes its own detector rows; a request's result is the weighted sum over
// its jobs taken on the host in energy order, so it does not depend on which
// other requests shared the launches.
#pragma once
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "simulation.hpp"

namespace rs {

struct ServeRequest {
  std::vector<SimConfig> members;  // one per energy (same N, precision, detector, phase steps)
  std::vector<double> weights;
};

struct ServeResult {
  int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
  std::vector<double> rows;  // n_steps x n_px
  int64_t n_steps = 0, n_px = 0;
  double submit_t = 0, start_t = 0, done_t = 0;
  std::string error;
};

struct ServeStats {
  int64_t ticks = 0, launches = 0, merged_jobs = 0;  // merged_jobs: sum over launches of jobs per launch
  int64_t jobs_done = 0, max_active = 0;
  double busy_s = 0;
};

class Server {
 public:
  Server(Context &ctx, double reserve_bytes, int max_jobs);
  ~Server();
  int64_t submit(ServeRequest req);
  ServeResult poll(int64_t id);
  bool wait(int64_t id, double timeout_s);
  void drain();
  ServeStats stats();

 private:
  struct Job;
  struct Req;
  void loop();
  void admit();
  bool tick();
  void reap();
  template <typename S> void compile_job(Job &j);

  // lane of a grid shape (assigned on first sight, fewest resident bytes)
  int lane_of(int64_t N, bool f64);

  Context &ctx_;
  int max_jobs_;
  struct Pool;
  std::unique_ptr<Pool> pool_;
  size_t reserve_ = 0;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<int64_t> pending_;
  std::map<int64_t, std::shared_ptr<Req>> reqs_;
  std::vector<std::unique_ptr<Job>> active_;
  int64_t next_id_ = 1;
  bool stop_ = false;
  ServeStats stats_;
  std::thread th_;
  // per lane: its stream and a ring of argument staging slots
  struct Slot;
  struct Lane;
  std::vector<std::unique_ptr<Lane>> lanes_;
  std::map<std::pair<int64_t, bool>, int> shape_lane_;
};

}  // namespace rs
