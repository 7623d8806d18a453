// Batched residual evaluations of many RVEs: one launch per stage for all of
// them.
//
// The reference evaluates one RVE per process (src/grid/main.cpp:31-101,
// damask_grid.cpp:82-196), one form_residual at a time
// (mech_solver_basic.cpp:286-339).  Here every RVE keeps its own solver (its
// N-GMRES iteration on the host, its vectors on its own stream), but the
// residual evaluations that the running RVEs ask for are gathered and
// computed together:
//
//   material point   one launch over the cells of every RVE (a block table
//                    maps blocks to RVEs; launch_mp_response_batch), its
//                    pivoting pass and the P block sums likewise;
//   reductions       one launch for the P sums of all, one for the aims;
//   spectral stages  RVEs with the same grid and arithmetic form a group:
//                    their P planes are gathered into one contiguous arena,
//                    transformed as 9 x (group size) components by the
//                    pencil FFTs, convolved by one Gamma launch and
//                    transformed back; r is scattered to each RVE's buffer;
//   scalars          the response scalars of all RVEs in one copy.
//
// A batch is launched when every RVE that is inside an increment (a
// member) waits for an evaluation, or when the oldest request has waited
// max_wait.  Every kernel computes each RVE's values exactly as the RVE's
// own launches do (same cells per block, same reduction trees), so a
// result is bitwise the same in any batch and alone.
#include <chrono>
#include <map>
#include <tuple>

#include "batch.h"

namespace sca {

Batcher::Batcher(double max_wait_ms) : max_wait_(max_wait_ms) {
  st_ = stream_get();
  th_ = std::thread([this] { run(); });
}

Batcher::~Batcher() {
  {
    std::lock_guard<std::mutex> lk(m_);
    stop_ = true;
  }
  cv_req_.notify_all();
  th_.join();
  for (auto& kv : groups_) {
    Group& g = kv.second;
    g.f64.release();
    g.f32.release();
    dev_free(g.real);
    dev_free(g.spec);
    dev_free(g.rowsum);
  }
  dev_free(d_scal_);
  host_free(h_scal_);
  stream_put(st_);
}

void Batcher::join() {
  std::lock_guard<std::mutex> lk(m_);
  ++members_;
}

void Batcher::leave() {
  {
    std::lock_guard<std::mutex> lk(m_);
    --members_;
  }
  cv_req_.notify_all();  // the others may all be waiting now
}

void Batcher::evaluate(Rve* rve, const double* F, void* r) {
  Req q{rve, F, r, false, {}};
  std::unique_lock<std::mutex> lk(m_);
  if (waiting_.empty()) first_ = std::chrono::steady_clock::now();
  waiting_.push_back(&q);
  cv_req_.notify_all();
  cv_done_.wait(lk, [&] { return q.done; });
  if (!q.err.empty()) throw std::runtime_error(q.err);
}

BatchStats Batcher::stats() {
  std::lock_guard<std::mutex> lk(m_);
  return stats_;
}

void Batcher::run() {
  for (;;) {
    std::vector<Req*> take;
    {
      std::unique_lock<std::mutex> lk(m_);
      for (;;) {
        if (stop_ && waiting_.empty()) return;
        if (!waiting_.empty()) {
          if ((int)waiting_.size() >= members_ || stop_) break;
          const auto due = first_ + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                        std::chrono::duration<double, std::milli>(max_wait_));
          if (std::chrono::steady_clock::now() >= due) break;
          cv_req_.wait_until(lk, due);
        } else {
          cv_req_.wait(lk);
        }
      }
      take.swap(waiting_);
    }
    std::string err;
    const auto t0 = std::chrono::steady_clock::now();
    try {
      launch(take);
    } catch (std::exception& e) {
      err = e.what();
      (void)hipGetLastError();
    }
    const double busy = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    {
      std::lock_guard<std::mutex> lk(m_);
      stats_.busy_seconds += busy;
      for (Req* q : take) {
        q->err = err;
        q->done = true;
      }
      stats_.batches++;
      stats_.evaluations += (int64_t)take.size();
      stats_.max_batch = std::max<int64_t>(stats_.max_batch, (int64_t)take.size());
    }
    cv_done_.notify_all();
  }
}

// arena of a group of RVEs with the same grid and arithmetic, for up to cap
// of them at once
void Batcher::ensure(Group& g, int cap) {
  if (g.cap >= cap) return;
  const int64_t N = (int64_t)g.cells[0] * g.cells[1] * g.cells[2];
  const int64_t K = (int64_t)(g.cells[0] / 2 + 1) * g.cells[1] * g.cells[2];
  dev_free(g.real);  // the batch that used them has completed
  dev_free(g.spec);
  dev_free(g.rowsum);
  g.real = nullptr;
  g.spec = nullptr;
  g.rowsum = nullptr;
  g.cap = 0;
  g.real = static_cast<double*>(dev_alloc((size_t)9 * N * cap * sizeof(double)));
  g.spec = dev_alloc((size_t)2 * 9 * K * cap * (g.fp32 ? sizeof(float) : sizeof(double)));
  g.rowsum = static_cast<double*>(dev_alloc((size_t)9 * g.cells[1] * g.cells[2] * cap * sizeof(double)));
  g.cap = cap;
}

void Batcher::launch(const std::vector<Req*>& reqs) {
  const int n = (int)reqs.size();
  if (n == 0) return;
  ta_.reserve((size_t)n * batch_table_bytes_per_rve() + (64 << 10));
  // material point, P sums, aims.  A context whose x length has a pencil
  // kernel takes its P sums from the row sums of its forward transform (as
  // alone, Rve::launch_spectral), so its aim follows that transform; the
  // others sum P after the update, as alone.
  auto bc_item = [](Rve& v) {
    BcBatchItem b;
    b.sums = v.small + SM_SUMS;
    std::memcpy(b.S, v.S, sizeof b.S);
    double R[9];
    qu2om_host(v.rot_q, R);
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) b.R[3 * a + c] = R[a + 3 * c];
    b.wgt = v.wgt;
    b.rotate = 1;
    b.aim = v.small + SM_AIM;
    return b;
  };
  std::vector<MpBatchItem> mp(n);
  std::vector<BatchReduce> red;
  std::vector<BcBatchItem> bc;
  for (int i = 0; i < n; ++i) {
    Rve& v = *reqs[i]->rve;
    MpState s = v.mp;
    s.F = const_cast<double*>(reqs[i]->F);  // read only
    mp[i] = MpBatchItem{s, v.mc, v.delta_t, v.d_ill, v.sub, v.xsums ? nullptr : v.partial};
    if (!v.xsums) {
      red.push_back(BatchReduce{v.partial, mp_partial_blocks(v.N), 9, v.small + SM_SUMS});
      bc.push_back(bc_item(v));
    }
  }
  launch_mp_response_batch(mp.data(), n, ta_, st_);
  launch_final_reduce_batch(red.data(), (int)red.size(), ta_, st_);
  launch_bc_aim_batch(bc.data(), (int)bc.size(), ta_, st_);
  // spectral stages per group of equal grid and arithmetic
  std::map<std::tuple<int, int, int, int>, std::vector<int>> by_grid;
  for (int i = 0; i < n; ++i) {
    const Rve& v = *reqs[i]->rve;
    by_grid[{v.cells[0], v.cells[1], v.cells[2], v.use32 ? 1 : 0}].push_back(i);
  }
  for (auto& kv : by_grid) {
    const auto& key = kv.first;
    const std::vector<int>& idx = kv.second;
    Group& g = groups_[key];
    if (!g.init) {
      g.cells[0] = std::get<0>(key);
      g.cells[1] = std::get<1>(key);
      g.cells[2] = std::get<2>(key);
      g.fp32 = std::get<3>(key) != 0;
      if (g.fp32) g.f32.init(g.cells);
      else g.f64.init(g.cells);
      g.init = true;
    }
    const bool measure = reqs[idx[0]]->rve->desc.grid_num.fft_plan_mode == 1;
    const int64_t N = (int64_t)g.cells[0] * g.cells[1] * g.cells[2];
    const int64_t K = (int64_t)(g.cells[0] / 2 + 1) * g.cells[1] * g.cells[2];
    // the generic strided passes launch 9 x (RVEs) x (other length) block
    // rows, at most 65535
    const int ymax = std::max(g.cells[1], g.cells[2]);
    const int chunk = std::max(1, std::min((int)idx.size(), 65535 / (9 * ymax)));
    ensure(g, chunk);
    if (measure && g.tuned_for != chunk) {  // launch shapes for transforms of 9 x chunk components
      if (g.fp32) g.f32.tune(9 * chunk, st_);
      else g.f64.tune(9 * chunk, st_);
      g.tuned_for = chunk;
    }
    for (size_t c0 = 0; c0 < idx.size(); c0 += chunk) {
      const int m = (int)std::min<size_t>(chunk, idx.size() - c0);
      std::vector<BatchCopy> gather(m), scatter(m);
      std::vector<GammaBatchItem> gam(m);
      std::vector<BatchReduce> div(m);
      std::vector<RowSumItem> rsum;
      std::vector<BcBatchItem> bcx;
      const size_t spec_stride = (size_t)2 * 9 * K * (g.fp32 ? sizeof(float) : sizeof(double));
      for (int j = 0; j < m; ++j) {
        const Req& q = *reqs[idx[c0 + j]];
        Rve& v = *q.rve;
        double* slot = g.real + (size_t)9 * N * j;
        gather[j] = BatchCopy{v.mp.P, slot, 2 * 9 * N};
        // an FP32 group's residuals are floats (the FP32 stage stores them so)
        scatter[j] = g.fp32 ? BatchCopy{static_cast<float*>(static_cast<void*>(g.real)) + (size_t)9 * N * j, q.r, 9 * N}
                            : BatchCopy{slot, q.r, 2 * 9 * N};
        gam[j] = GammaBatchItem{static_cast<char*>(g.spec) + spec_stride * j, v.gamma_host(true), v.small + SM_AIM,
                                v.d_xi_tab, v.partial};
        div[j] = BatchReduce{v.partial, gamma_blocks(K), 1, v.small + SM_DIV};
        if (v.xsums) {
          const int64_t nrow = (int64_t)g.cells[1] * g.cells[2];
          rsum.push_back(RowSumItem{g.rowsum + (size_t)9 * nrow * j, nrow, v.small + SM_SUMS});
          bcx.push_back(bc_item(v));
        }
      }
      // the contexts of a group share the grid: all or none take row sums
      double* rowsum = rsum.empty() ? nullptr : g.rowsum;
      if (!rsum.empty() && (int)rsum.size() != m) throw std::runtime_error("batch: mixed P-sum paths in one group");
      launch_copy_batch(gather.data(), m, ta_, st_);
      // the x passes pair rows within each RVE's 9 planes (set_comp 9): the
      // transforms give the bits of the RVE transformed alone
      if (g.fp32) g.f32.forward(g.real, static_cast<float*>(g.spec), 9 * m, st_, 9, rowsum);
      else g.f64.forward(g.real, static_cast<double*>(g.spec), 9 * m, st_, 9, rowsum);
      if (rowsum) {
        launch_rowsum_reduce_batch(rsum.data(), m, ta_, st_);
        launch_bc_aim_batch(bcx.data(), m, ta_, st_);
      }
      if (g.fp32) launch_gamma_div_batch<float>(gam.data(), m, K, ta_, st_);
      else launch_gamma_div_batch<double>(gam.data(), m, K, ta_, st_);
      launch_final_reduce_batch(div.data(), m, ta_, st_);
      if (g.fp32) g.f32.backward(static_cast<float*>(g.spec), static_cast<float*>(static_cast<void*>(g.real)), 9 * m,
                                 1.0, st_, 9);
      else g.f64.backward(static_cast<double*>(g.spec), g.real, 9 * m, 1.0, st_, 9);
      launch_copy_batch(scatter.data(), m, ta_, st_);
    }
  }
  // response scalars (P sums, terminally-ill flag, divergence sum) of all
  constexpr int NS = 11;
  if (scal_cap_ < n) {
    const int cap = std::max(n, 2 * scal_cap_);
    dev_free(d_scal_);
    host_free(h_scal_);
    d_scal_ = nullptr;
    h_scal_ = nullptr;
    scal_cap_ = 0;
    d_scal_ = static_cast<double*>(dev_alloc((size_t)NS * cap * sizeof(double)));
    h_scal_ = static_cast<double*>(host_alloc((size_t)NS * cap * sizeof(double)));
    scal_cap_ = cap;
  }
  std::vector<BatchCopy> sc(n);
  for (int i = 0; i < n; ++i) sc[i] = BatchCopy{reqs[i]->rve->small + SM_SUMS, d_scal_ + (size_t)NS * i, 2 * NS};
  launch_copy_batch(sc.data(), n, ta_, st_);
  SCA_HIP(hipMemcpyAsync(h_scal_, d_scal_, (size_t)NS * n * sizeof(double), hipMemcpyDeviceToHost, st_));
  SCA_HIP(hipStreamSynchronize(st_));
  for (int i = 0; i < n; ++i) std::memcpy(reqs[i]->rve->h_scal, h_scal_ + (size_t)NS * i, NS * sizeof(double));
}

}  // namespace sca
