// RVE job server: continuous batching of load increments on one GPU.
//
// A job is one RVE (geometry, material, numerics, load case) whose fields and
// constitutive state stay resident in HBM between its increments.  A request
// advances a job by one increment of its load case (the body of
// loop_over_loadcase, damask_grid.cpp:82-196).  Requests are admitted by HBM
// footprint: a job becomes resident (its buffers are allocated and it is
// initialised) only when the footprint fits the budget next to the jobs
// already resident; one-shot jobs release their memory when their last
// request is done.  Worker threads each drive one admitted job at a time on
// that job's own HIP stream, so the increments of different jobs overlap on
// the device and new requests join as soon as a worker and memory are free.
// Increments of one job run in submission order.  Every job's arithmetic is
// independent of the others (no shared device state, deterministic
// reductions), so a result is the same alone and in any mix.  When a waiting
// job does not fit, idle resident jobs (kept resident, no request pending)
// are evicted, oldest first: their state is checkpointed to host memory
// (Rve::save) and their device memory freed; their next request restores it
// (Rve::load) instead of initialising, so eviction changes no result.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.h"
#include "rve.h"

namespace sca {
namespace {

using clk = std::chrono::steady_clock;

struct Job {
  int64_t id = 0;
  sc_rve_desc desc{};
  std::vector<int> material_at;
  std::vector<double> orientations;
  std::vector<sc_loadstep> steps;
  bool keep_resident = false;
  int64_t footprint = 0;
  std::unique_ptr<Rve> rve;
  LoadCursor cur;
  bool busy = false, failed = false;
  std::deque<int64_t> pending;
  // evicted to host memory: the checkpoint the job resumes from
  std::vector<char> snapshot;
  bool evicted = false;
  uint64_t last_use = 0;  // order of the last request done (eviction picks the oldest idle job)
};

struct Request {
  sc_request_result res{};
  bool done = false;
  std::string err;  // message of a failed request (status 3), for the waiting client
};

}  // namespace

// worst-case device bytes of one context (see Rve::Rve): state, substep and
// N-GMRES planes per cell, spectra per wavevector, a full history of m pairs
int64_t rve_footprint(const sc_rve_desc& d) {
  const int64_t N = (int64_t)d.cells[0] * d.cells[1] * d.cells[2];
  const int64_t K = (int64_t)(d.cells[0] / 2 + 1) * d.cells[1] * d.cells[2];
  const int m = d.grid_num.ngmres_m > 0 ? d.grid_num.ngmres_m : 30;
  int64_t doubles = 288 * N + 18 * K + (int64_t)18 * m * N + 3 * (int64_t)(d.cells[0] + 1) * (d.cells[1] + 1) *
                                                               (d.cells[2] + 1);
  if (d.grid_num.derivative != 0) doubles += 6 * K;
  if (d.precision == SC_PRECISION_MIXED) doubles += 9 * K;
  return doubles * 8 + (int64_t)(64 << 20);  // + partials, tables and allocator slack
}

class Server {
 public:
  // batch_wait_ms >= 0: the residual evaluations of the running jobs are
  // computed in joint launches (Batcher), the oldest waiting at most that
  // long for the others
  Server(int64_t budget, int workers, double batch_wait_ms = -1.0) : budget_(budget), t0_(clk::now()) {
    if (workers < 1) workers = 1;
    if (batch_wait_ms >= 0.0) batcher_.reset(new Batcher(batch_wait_ms));
    for (int w = 0; w < workers; ++w) threads_.emplace_back([this] { worker(); });
  }
  ~Server() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& t : threads_) t.join();
    jobs_.clear();  // contexts go before the batcher they point to
  }

  int64_t add_job(const sc_rve_desc& d, const sc_loadstep* steps, int n_steps, bool keep) {
    validate_rve_desc(d);  // a job accepted here can be started
    if (n_steps < 1 || !steps) throw std::runtime_error("a job needs at least one load step");
    auto j = std::make_unique<Job>();
    const int64_t N = (int64_t)d.cells[0] * d.cells[1] * d.cells[2];
    j->desc = d;
    j->material_at.assign(d.material_at, d.material_at + N);
    j->orientations.assign(d.orientations, d.orientations + 4 * d.n_materials);
    j->desc.material_at = j->material_at.data();
    j->desc.orientations = j->orientations.data();
    j->steps.assign(steps, steps + n_steps);
    j->keep_resident = keep;
    j->footprint = rve_footprint(d);
    if (j->footprint > budget_) throw std::runtime_error("job does not fit the memory budget");
    std::lock_guard<std::mutex> lk(m_);
    j->id = ++next_job_;
    const int64_t id = j->id;
    jobs_[id] = std::move(j);
    return id;
  }

  int64_t submit(int64_t job) {
    std::lock_guard<std::mutex> lk(m_);
    auto it = jobs_.find(job);
    if (it == jobs_.end()) throw std::runtime_error("unknown job");
    const int64_t id = ++next_req_;
    Request& r = reqs_[id];
    r.res.request = id;
    r.res.job = job;
    r.res.t_submit = now();
    Job& j = *it->second;
    j.pending.push_back(id);
    if (!j.busy && j.pending.size() == 1) ready_.push_back(job);
    cv_work_.notify_one();
    return id;
  }

  // blocks until the request is done (timeout_s < 0: no limit); 1 on timeout
  int wait(int64_t request, sc_request_result* out, double timeout_s) {
    std::unique_lock<std::mutex> lk(m_);
    auto it = reqs_.find(request);
    if (it == reqs_.end()) throw std::runtime_error("unknown request");
    auto done = [&] { return reqs_[request].done; };
    if (timeout_s < 0) cv_done_.wait(lk, done);
    else if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) return 1;
    *out = reqs_[request].res;
    if (!reqs_[request].err.empty()) g_err = reqs_[request].err;  // for sc_last_error() of the client
    reqs_.erase(request);
    return 0;
  }

  void remove_job(int64_t job) {
    std::unique_ptr<Rve> drop;
    {
      std::lock_guard<std::mutex> lk(m_);
      auto it = jobs_.find(job);
      if (it == jobs_.end()) return;
      Job& j = *it->second;
      if (j.busy || !j.pending.empty()) {
        j.keep_resident = false;  // released once idle
        return;
      }
      if (j.rve) used_ -= j.footprint;
      drop = std::move(j.rve);
      jobs_.erase(it);
    }
    cv_work_.notify_all();
  }

  // a field of an idle resident job (the job is held busy meanwhile)
  void get_field(int64_t job, int field, double* out, int64_t n) {
    Job* j = nullptr;
    {
      std::lock_guard<std::mutex> lk(m_);
      auto it = jobs_.find(job);
      if (it == jobs_.end()) throw std::runtime_error("unknown job");
      j = it->second.get();
      if (j->busy || !j->pending.empty()) throw std::runtime_error("job has a request running or pending");
      if (!j->rve) throw std::runtime_error("job is not resident");
      j->busy = true;
    }
    int rc = 1;
    std::string err;
    try {
      rc = rve_get_field(*j->rve, field, out, n);
    } catch (std::exception& e) {
      err = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      j->busy = false;
      if (!j->pending.empty()) ready_.push_back(job);
    }
    cv_work_.notify_all();
    if (!err.empty()) throw std::runtime_error(err);
    if (rc != 0) throw std::runtime_error("field copy failed");
  }

  void stats(sc_server_stats* o) {
    std::lock_guard<std::mutex> lk(m_);
    o->budget_bytes = budget_;
    o->used_bytes = used_;
    o->peak_bytes = peak_;
    o->resident_jobs = 0;
    for (auto& kv : jobs_) o->resident_jobs += kv.second->rve ? 1 : 0;
    o->completed_requests = completed_;
    o->max_concurrent = max_concurrent_;
    o->evictions = evictions_;
    if (batcher_) {
      const BatchStats b = batcher_->stats();
      o->batches = b.batches;
      o->batched_evaluations = b.evaluations;
      o->max_batch = b.max_batch;
      o->batch_seconds = b.busy_seconds;
    }
  }

 private:
  double now() const { return std::chrono::duration<double>(clk::now() - t0_).count(); }

  // an idle resident job to evict when ready jobs wait for memory: oldest
  // last use first
  Job* pick_victim() {
    if (ready_.empty()) return nullptr;
    Job* v = nullptr;
    for (auto& kv : jobs_) {
      Job& j = *kv.second;
      if (j.rve && !j.busy && !j.failed && j.pending.empty() && (!v || j.last_use < v->last_use)) v = &j;
    }
    return v;
  }

  // first ready job (FIFO) that is resident or whose footprint fits
  bool pick(int64_t& job) {
    for (auto it = ready_.begin(); it != ready_.end(); ++it) {
      Job& j = *jobs_[*it];
      // a failed job is never started again: its requests are answered
      // without memory
      if (j.rve || j.failed || used_ + j.footprint <= budget_) {
        job = *it;
        ready_.erase(it);
        return true;
      }
    }
    return false;
  }

  void worker() {
    for (;;) {
      int64_t jid = 0, rid = 0;
      Job* j = nullptr;
      Job* victim = nullptr;
      bool admit = false;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_work_.wait(lk, [&] {
          victim = nullptr;
          return stop_ || pick(jid) || (victim = pick_victim()) != nullptr;
        });
        if (stop_) return;
        if (victim) {
          victim->busy = true;  // held while it is checkpointed
        } else {
          j = jobs_[jid].get();
          j->busy = true;
          rid = j->pending.front();
          j->pending.pop_front();
          if (!j->rve && !j->failed) {
            used_ += j->footprint;
            peak_ = std::max(peak_, used_);
            admit = true;
          }
          ++running_;
          max_concurrent_ = std::max(max_concurrent_, running_);
          reqs_[rid].res.t_start = now();
        }
      }
      if (victim) {
        evict(*victim);
        continue;
      }
      sc_request_result res{};
      res.request = rid;
      res.job = jid;
      std::string err;
      try {
        if (admit) {
          // jobs evaluate eagerly: the per-job streams keep the device busy,
          // and short jobs would not amortise graph captures
          j->rve.reset(new Rve(j->desc, false));
          j->rve->spin_wait = false;  // the workers share the host cores
          j->rve->batch = batcher_.get();
          if (j->evicted) {
            j->rve->load(j->snapshot);
            j->snapshot.clear();
            j->snapshot.shrink_to_fit();
            j->evicted = false;
          } else {
            j->rve->init();
          }
        }
        if (j->failed) {
          res.status = 2;
        } else {
          int iters = 0, evals = 0;
          BatchMembership member(batcher_.get());
          const int rc = j->rve->next_increment(j->steps.data(), (int)j->steps.size(), j->cur, &iters, &evals);
          res.status = rc;
          if (rc == 2) j->failed = true;
          res.increment = j->cur.total_inc;
          res.iterations = iters;
          res.evaluations = evals;
          for (int q = 0; q < 9; ++q) res.P_av[q] = j->rve->P_av[q];
          res.t = j->cur.t;
          res.material_point_updates = (int64_t)evals * j->rve->N;
        }
      } catch (std::exception& e) {
        res.status = 3;
        j->failed = true;
        err = e.what();
      }
      std::unique_ptr<Rve> drop;
      {
        std::lock_guard<std::mutex> lk(m_);
        --running_;
        if (j->failed && j->rve) {
          // a failed job holds no memory: its later requests are answered
          // as failed without starting it again
          used_ -= j->footprint;
          drop = std::move(j->rve);
        } else if (admit && !j->rve) {
          used_ -= j->footprint;  // the start threw before the context existed
        }
        Request& r = reqs_[rid];
        res.t_submit = r.res.t_submit;
        res.t_start = r.res.t_start;
        res.t_end = now();
        r.res = res;
        r.err = err;
        r.done = true;
        ++completed_;
        j->busy = false;
        j->last_use = ++use_clock_;
        if (!j->pending.empty()) {
          ready_.push_front(jid);  // keep this job's increments flowing
        } else if (!j->keep_resident) {
          if (j->rve) used_ -= j->footprint;
          if (j->rve) drop = std::move(j->rve);
          jobs_.erase(jid);
        }
      }
      drop.reset();  // release outside the lock
      trim_cache();
      cv_done_.notify_all();
      cv_work_.notify_all();
    }
  }

  // Released contexts leave their device blocks in the library's cache
  // (reused by the next jobs of the same sizes without hipMalloc); the cache
  // is held to the part of the budget the resident jobs leave free.
  void trim_cache() {
    int64_t free_budget;
    {
      std::lock_guard<std::mutex> lk(m_);
      free_budget = std::max<int64_t>(0, budget_ - used_);
    }
    if ((int64_t)pool_cached_bytes() > free_budget) pool_trim((size_t)free_budget);
  }

  // checkpoint an idle resident job to host memory and free its device
  // memory (the job was marked busy by the caller)
  void evict(Job& v) {
    std::vector<char> blob;
    bool ok = true;
    try {
      v.rve->save(blob);
    } catch (std::exception& e) {
      ok = false;
      g_err = e.what();
    }
    std::unique_ptr<Rve> drop;
    {
      std::lock_guard<std::mutex> lk(m_);
      if (ok) {
        v.snapshot = std::move(blob);
        v.evicted = true;
        ++evictions_;
      } else {
        v.failed = true;  // its state is lost: the job fails and gives its memory back
      }
      drop = std::move(v.rve);
      used_ -= v.footprint;
      v.busy = false;
      if (!v.pending.empty()) ready_.push_back(v.id);
    }
    drop.reset();
    trim_cache();
    cv_work_.notify_all();
  }

  std::mutex m_;
  std::condition_variable cv_work_, cv_done_;
  std::map<int64_t, std::unique_ptr<Job>> jobs_;
  std::map<int64_t, Request> reqs_;
  std::deque<int64_t> ready_;
  std::vector<std::thread> threads_;
  std::unique_ptr<Batcher> batcher_;
  int64_t budget_, used_ = 0, peak_ = 0, next_job_ = 0, next_req_ = 0, completed_ = 0, evictions_ = 0;
  uint64_t use_clock_ = 0;
  int running_ = 0, max_concurrent_ = 0;
  bool stop_ = false;
  clk::time_point t0_;
};

}  // namespace sca

struct sc_server {
  std::unique_ptr<sca::Server> s;
};

using sca::g_err;

extern "C" {

int64_t sc_rve_footprint(const sc_rve_desc* d) { return sca::rve_footprint(*d); }

int sc_server_create(int64_t hbm_budget_bytes, int workers, sc_server** out) {
  try {
    auto* s = new sc_server();
    s->s.reset(new sca::Server(hbm_budget_bytes, workers));
    *out = s;
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_server_create_batched(int64_t hbm_budget_bytes, int workers, double max_wait_ms, sc_server** out) {
  try {
    if (!(max_wait_ms >= 0.0)) throw std::runtime_error("max_wait_ms must be >= 0");
    auto* s = new sc_server();
    s->s.reset(new sca::Server(hbm_budget_bytes, workers, max_wait_ms));
    *out = s;
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

void sc_server_destroy(sc_server* s) { delete s; }

int sc_server_add_job(sc_server* s, const sc_rve_desc* d, const sc_loadstep* steps, int n_steps, int keep_resident,
                      int64_t* job) {
  try {
    *job = s->s->add_job(*d, steps, n_steps, keep_resident != 0);
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_server_submit(sc_server* s, int64_t job, int64_t* request) {
  try {
    *request = s->s->submit(job);
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_server_wait(sc_server* s, int64_t request, double timeout_s, sc_request_result* out) {
  try {
    return s->s->wait(request, out, timeout_s) ? 4 : 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_server_remove_job(sc_server* s, int64_t job) {
  try {
    s->s->remove_job(job);
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_server_get_field(sc_server* s, int64_t job, int field, double* out, int64_t n_values) {
  try {
    s->s->get_field(job, field, out, n_values);
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_server_get_stats(sc_server* s, sc_server_stats* out) {
  s->s->stats(out);
  return 0;
}

}  // extern "C"
