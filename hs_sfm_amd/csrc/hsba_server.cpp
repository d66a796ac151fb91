// hsba_server.cpp — serving loop: continuous batching of independent LM jobs.
//
// No counterpart in the reference (it solves one problem per optimizer call).
// Jobs are copied in at submit and queued in arrival order.  Worker threads,
// each with its own HIP stream, take the oldest queued job plus every later
// queued job of the same shape class (camera parameterisation, fix mask,
// solver choice), up to max_batch_jobs, while the jobs' estimated device bytes
// fit in what is left of the HBM budget and the admitted-job count stays
// within max_active_jobs (admission control).  A group runs as one batch: one
// set of launches per LM iteration covers all of its jobs, and each job's
// arithmetic is that of the job alone (every reduction, factorization and LM
// decision is per window).  Groups run concurrently on the workers' streams,
// so a job arriving while others run waits for a free worker, not for the
// running batch to end.  Device buffers come from one arena allocated at
// creation (the budget plus a quarter for fragmentation), so no hipMalloc /
// hipFree runs while jobs execute.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <string>
#include <thread>

#include "hsba_host.h"

namespace hsba {

DeviceArena::~DeviceArena() {
  if (base_) (void)hipFree(base_);
}

int DeviceArena::init(int64_t bytes) {
  bytes = (bytes + 255) & ~int64_t(255);
  if (bytes <= 0 || hipMalloc((void **)&base_, (size_t)bytes) != hipSuccess) {
    base_ = nullptr;
    return HSBA_ERR_OOM;
  }
  cap_ = bytes;
  used_ = 0;
  free_.clear();
  live_.clear();
  free_[0] = bytes;
  return HSBA_OK;
}

void *DeviceArena::alloc(int64_t bytes) {
  bytes = std::max<int64_t>(256, (bytes + 255) & ~int64_t(255));
  std::lock_guard<std::mutex> g(mu_);
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < bytes) continue;
    const int64_t off = it->first, sz = it->second;
    free_.erase(it);
    if (sz > bytes) free_[off + bytes] = sz - bytes;
    live_[off] = bytes;
    used_ += bytes;
    return base_ + off;
  }
  return nullptr;
}

void DeviceArena::release(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu_);
  auto it = live_.find((char *)p - base_);
  if (it == live_.end()) return;
  int64_t start = it->first, sz = it->second;
  used_ -= sz;
  live_.erase(it);
  auto nx = free_.lower_bound(start);
  if (nx != free_.end() && nx->first == start + sz) {
    sz += nx->second;
    nx = free_.erase(nx);
  }
  if (nx != free_.begin()) {
    auto pv = std::prev(nx);
    if (pv->first + pv->second == start) {
      start = pv->first;
      sz += pv->second;
      free_.erase(pv);
    }
  }
  free_[start] = sz;
}

int64_t DeviceArena::in_use() {
  std::lock_guard<std::mutex> g(mu_);
  return used_;
}

}  // namespace hsba

namespace {

using Clock = std::chrono::steady_clock;

template <class X>
void copy_in(std::vector<X> &v, const X *p, int64_t n) {
  if (p && n > 0) v.assign(p, p + n);
  else v.clear();
}

// One submitted job: owned copies of every input the caller passed.
struct Job {
  int64_t id = 0;
  std::string cls;
  int64_t bytes = 0;
  hsba_problem_desc desc;
  std::vector<int32_t> fm, icm;
  std::vector<double> intr, fixp, fixi, fixc;
  std::vector<hsba_constraint> pc, ic, cc;
  std::vector<double> x0, y, kb, cw, pb;
  hsba_ycov_inverse ycov;
  double *x_out = nullptr;
  hsba_job_result res;
};

}  // namespace

struct hsba_server {
  hsba_server_options opt;
  hsba::DeviceArena arena;
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::unique_ptr<Job>> queue;
  std::deque<std::unique_ptr<Job>> done;
  int64_t next_id = 1, outstanding = 0, resident = 0;
  int active = 0;
  bool stop = false;
  std::vector<std::thread> workers;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  // Admission (caller holds mu): the oldest job and the later queued jobs of
  // its class that fit the budget and the active-job limit, in arrival order.
  // Keys per batch: a bounded batch bounds the latency of its first job.
  static constexpr int64_t kMaxBatchKeys = int64_t(1) << 23;

  bool take_group(std::vector<std::unique_ptr<Job>> &g) {
    if (queue.empty()) return false;
    const Job &head = *queue.front();
    if (active + 1 > opt.max_active_jobs || resident + head.bytes > opt.hbm_budget_bytes) return false;
    int64_t bytes = 0, keys = 0;
    const double t = now();
    for (auto it = queue.begin(); it != queue.end() && (int)g.size() < opt.max_batch_jobs;) {
      Job &j = **it;
      if (j.cls == head.cls && active + (int)g.size() + 1 <= opt.max_active_jobs &&
          resident + bytes + j.bytes <= opt.hbm_budget_bytes &&
          (g.empty() || keys + j.desc.num_keys <= kMaxBatchKeys)) {
        bytes += j.bytes;
        keys += j.desc.num_keys;
        j.res.admit_time_s = t;
        g.push_back(std::move(*it));
        it = queue.erase(it);
      } else {
        ++it;
      }
    }
    active += (int)g.size();
    resident += bytes;
    return true;
  }

  void run_group(std::vector<std::unique_ptr<Job>> &g, hipStream_t st) {
    const int64_t n = (int64_t)g.size();
    std::vector<hsba_problem_desc> descs(n);
    std::vector<const double *> x0s(n), ys(n);
    std::vector<double *> xouts(n);
    std::vector<hsba_ycov_inverse> ycovs(n);
    std::vector<hsba_lm_stats> stats(n);
    for (int64_t i = 0; i < n; ++i) {
      descs[i] = g[i]->desc;
      x0s[i] = g[i]->x0.data();
      ys[i] = g[i]->y.data();
      xouts[i] = g[i]->x_out;
      ycovs[i] = g[i]->ycov;
    }
    hsba_options o;
    std::memset(&o, 0, sizeof(o));
    o.precision = opt.precision;
    o.solver = HSBA_SOLVER_AUTO;
    o.loss = opt.loss;
    o.loss_scale = opt.loss_scale;
    // rc is a failure of the group itself (allocation, device); a job's own
    // failure (an unobserved image, mu overflow, NaN) is its window's status
    // and leaves the other jobs of the group as if solved alone
    int rc = st ? HSBA_OK : HSBA_ERR_DEVICE;
    std::unique_ptr<hsba::EngineBase> eng;
    if (!rc) rc = hsba::create_engine(descs.data(), n, &o, st, &arena, eng);
    if (!rc) rc = eng->upload_x(x0s.data());
    if (!rc) rc = eng->upload_obs(ys.data(), ycovs.data());
    if (!rc) rc = eng->lm_run(&opt.lm, stats.data());
    if (!rc) rc = eng->download_x(xouts.data());
    eng.reset();  // waits for the stream, returns the buffers to the arena
    for (int64_t i = 0; i < n; ++i) {
      if (rc) {
        std::memset(&stats[i], 0, sizeof(stats[i]));
        stats[i].status = rc;
        stats[i].stop_reason = HSBA_STOP_FAILED;
      }
      g[i]->res.stats = stats[i];
      g[i]->res.status = stats[i].status;
    }
  }

  void worker_loop() {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      std::vector<std::unique_ptr<Job>> g;
      cv_work.wait(lk, [&] { return (stop && queue.empty()) || take_group(g); });
      if (g.empty()) break;
      lk.unlock();
      run_group(g, st);
      lk.lock();
      const double t = now();
      for (auto &j : g) {
        j->res.finish_time_s = t;
        resident -= j->bytes;
        done.push_back(std::move(j));
      }
      active -= (int)g.size();
      outstanding -= (int64_t)g.size();
      cv_done.notify_all();
      cv_work.notify_all();
    }
    lk.unlock();
    if (st) (void)hipStreamDestroy(st);
  }
};

extern "C" {

int hsba_server_create(const hsba_server_options *options, hsba_server **server) {
  return hsba::guarded([&]() -> int {
    if (!server) return HSBA_ERR_INVALID;
    *server = nullptr;
    std::unique_ptr<hsba_server> s(new hsba_server());
    std::memset(&s->opt, 0, sizeof(s->opt));
    if (options) s->opt = *options;
    hsba_server_options &o = s->opt;
    if (o.precision != HSBA_PRECISION_FP64 && o.precision != HSBA_PRECISION_MIXED) return HSBA_ERR_INVALID;
    if (o.loss != HSBA_LOSS_TRIVIAL && o.loss != HSBA_LOSS_HUBER) return HSBA_ERR_INVALID;
    if (o.hbm_budget_bytes <= 0) o.hbm_budget_bytes = int64_t(16) << 30;
    if (o.max_active_jobs <= 0) o.max_active_jobs = 4096;
    if (o.streams <= 0) o.streams = 4;
    if (o.max_batch_jobs <= 0) o.max_batch_jobs = 256;
    o.streams = std::min(o.streams, 16);
    int rc = s->arena.init(o.hbm_budget_bytes + o.hbm_budget_bytes / 4);
    if (rc) return rc;
    s->t0 = Clock::now();
    for (int i = 0; i < o.streams; ++i) s->workers.emplace_back([p = s.get()] { p->worker_loop(); });
    *server = s.release();
    return HSBA_OK;
  });
}

int hsba_server_destroy(hsba_server *s) {
  return hsba::guarded([&]() -> int {
    if (!s) return HSBA_OK;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
      s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    delete s;
    return HSBA_OK;
  });
}

int hsba_server_submit(hsba_server *s, const hsba_problem_desc *desc, const double *x0, const double *y_obs,
                       const hsba_ycov_inverse *ycov, double *x_out, int64_t *job_id) {
  return hsba::guarded([&]() -> int {
    if (!s || !desc || !x0 || !y_obs || !x_out) return HSBA_ERR_INVALID;
    int64_t xn = 0, yn = 0;
    int rc = hsba::problem_sizes(*desc, &xn, &yn);
    if (rc) return rc;
    // a malformed job is refused here, so it never reaches (and fails) a batch
    if ((rc = hsba::validate_desc(*desc, nullptr))) return rc;
    std::unique_ptr<Job> j(new Job());
    const hsba_problem_desc &d = *desc;
    const uint32_t fix = d.fix_mask & 7u;
    copy_in(j->fm, d.feature_maps, 2 * d.num_keys);
    copy_in(j->icm, d.image_camera_map, d.num_images);
    copy_in(j->intr, d.intrinsic_defaults, (int64_t)HSBA_NUM_INTRINSICS * d.num_cameras);
    copy_in(j->fixp, (fix & HSBA_FIX_POINTS) ? d.fix_points : nullptr, 3 * d.num_points);
    copy_in(j->fixi, (fix & HSBA_FIX_IMAGES) ? d.fix_images : nullptr, 6 * d.num_images);
    {
      int64_t xn_free = 0, dummy = 0;
      hsba_problem_desc all_free = d;
      all_free.fix_mask = 0;
      (void)hsba::problem_sizes(all_free, &xn_free, &dummy);
      const int64_t dcam = d.num_cameras > 0 ? (xn_free - 3 * d.num_points - 6 * d.num_images) / d.num_cameras : 0;
      copy_in(j->fixc, (fix & HSBA_FIX_CAMERAS) ? d.fix_cameras : nullptr, dcam * d.num_cameras);
    }
    copy_in(j->pc, d.point_constraints, d.num_point_constraints);
    copy_in(j->ic, d.image_constraints, d.num_image_constraints);
    copy_in(j->cc, d.camera_constraints, d.num_camera_constraints);
    copy_in(j->x0, x0, xn);
    copy_in(j->y, y_obs, yn);
    std::memset(&j->ycov, 0, sizeof(j->ycov));
    if (ycov) {
      copy_in(j->kb, ycov->key_blocks, 4 * d.num_keys);
      copy_in(j->cw, ycov->constraint_weights, yn - 2 * d.num_keys);
      j->ycov.key_blocks = j->kb.empty() ? nullptr : j->kb.data();
      j->ycov.constraint_weights = j->cw.empty() ? nullptr : j->cw.data();
      copy_in(j->pb, ycov->point_constraint_blocks, 9 * d.num_point_constraints);
      j->ycov.point_constraint_blocks = j->pb.empty() ? nullptr : j->pb.data();
    }
    j->desc = d;
    j->desc.feature_maps = j->fm.data();
    j->desc.image_camera_map = j->icm.empty() ? nullptr : j->icm.data();
    j->desc.intrinsic_defaults = j->intr.empty() ? nullptr : j->intr.data();
    j->desc.fix_points = j->fixp.empty() ? nullptr : j->fixp.data();
    j->desc.fix_images = j->fixi.empty() ? nullptr : j->fixi.data();
    j->desc.fix_cameras = j->fixc.empty() ? nullptr : j->fixc.data();
    j->desc.point_constraints = j->pc.empty() ? nullptr : j->pc.data();
    j->desc.image_constraints = j->ic.empty() ? nullptr : j->ic.data();
    j->desc.camera_constraints = j->cc.empty() ? nullptr : j->cc.data();
    j->x_out = x_out;
    const bool pcg = hsba::auto_pcg(d);
    j->bytes = hsba::estimate_device_bytes(d, s->opt.precision, pcg, !j->kb.empty());
    j->cls = std::to_string(fix) + "/" + std::to_string(d.intrinsic_computations_mask & 7u) + "/" +
             std::to_string(d.intrinsic_free_mask) + "/" + (pcg ? "pcg" : "dense");
    std::memset(&j->res, 0, sizeof(j->res));
    if (j->bytes > s->opt.hbm_budget_bytes) return HSBA_ERR_ADMISSION;
    {
      std::lock_guard<std::mutex> g(s->mu);
      j->id = s->next_id++;
      j->res.job_id = j->id;
      j->res.submit_time_s = s->now();
      if (job_id) *job_id = j->id;
      s->queue.push_back(std::move(j));
      ++s->outstanding;
    }
    s->cv_work.notify_one();
    return HSBA_OK;
  });
}

int hsba_server_poll(hsba_server *s, hsba_job_result *results, int32_t max_results, int32_t *count) {
  return hsba::guarded([&]() -> int {
    if (!s || !count || (max_results > 0 && !results)) return HSBA_ERR_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    int32_t n = 0;
    while (n < max_results && !s->done.empty()) {
      results[n++] = s->done.front()->res;
      s->done.pop_front();
    }
    *count = n;
    return HSBA_OK;
  });
}

int hsba_server_drain(hsba_server *s) {
  return hsba::guarded([&]() -> int {
    if (!s) return HSBA_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
    return HSBA_OK;
  });
}

double hsba_server_clock(const hsba_server *s) { return s ? s->now() : 0.0; }

}  // extern "C"
