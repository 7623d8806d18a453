// Batched residual evaluations of many RVEs (batch.cpp).
#pragma once
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "rve.h"

namespace sca {

struct BatchStats {
  int64_t batches = 0, evaluations = 0, max_batch = 0;
  double busy_seconds = 0.0;  // wall time of the joint launches, their completion included
};

// Gathers the residual evaluations of the RVEs that are inside an increment
// (members) and computes them together on its own stream.
class Batcher {
 public:
  // max_wait_ms: longest time the oldest request waits for the other members
  explicit Batcher(double max_wait_ms);
  ~Batcher();
  Batcher(const Batcher&) = delete;
  Batcher& operator=(const Batcher&) = delete;
  void join();   // a context starts an increment: its evaluations are batched
  void leave();  // ... and ends it
  // r = form_residual(F) of rve, computed in a joint launch; F must be
  // complete (the caller synchronised its stream).  r holds floats while
  // rve is in its FP32 stage, doubles otherwise.  Returns once r, the
  // material-point state and rve->h_scal hold the results.
  void evaluate(Rve* rve, const double* F, void* r);
  BatchStats stats();

 private:
  struct Req {
    Rve* rve;
    const double* F;
    void* r;
    bool done = false;
    std::string err;
  };
  struct Group {
    bool init = false, fp32 = false;
    int cells[3] = {0, 0, 0};
    FftPlan<double> f64;
    FftPlan<float> f32;
    double* real = nullptr;  // 9 N doubles per RVE slot (P in; r out, as floats in an FP32 group)
    void* spec = nullptr;    // 9 K complex per RVE slot
    double* rowsum = nullptr;  // 9 Ny Nz x-pass row sums of P per RVE slot (pencil x lengths)
    int cap = 0;
    int tuned_for = 0;  // components / 9 the launch shapes were measured for
  };
  void run();
  void launch(const std::vector<Req*>& reqs);
  void ensure(Group& g, int cap);

  double max_wait_;
  std::mutex m_;
  std::condition_variable cv_req_, cv_done_;
  std::vector<Req*> waiting_;
  std::chrono::steady_clock::time_point first_;
  int members_ = 0;
  bool stop_ = false;
  BatchStats stats_;
  hipStream_t st_ = nullptr;
  TableArena ta_;
  std::map<std::tuple<int, int, int, int>, Group> groups_;
  double* d_scal_ = nullptr;
  double* h_scal_ = nullptr;
  int scal_cap_ = 0;
  std::thread th_;  // last: started after everything it uses
};

// membership of a context in a batcher for the duration of a scope
struct BatchMembership {
  Batcher* b;
  explicit BatchMembership(Batcher* b_) : b(b_) {
    if (b) b->join();
  }
  ~BatchMembership() {
    if (b) b->leave();
  }
};

}  // namespace sca
