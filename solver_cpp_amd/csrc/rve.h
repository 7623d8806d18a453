// One representative volume element resident on the GPU (solver.cpp).
#pragma once
#include <algorithm>
#include <complex>
#include <functional>
#include <string>
#include <vector>

#include "common.h"
#include "spectral.h"

struct sc_rve;

// grids up to this many cells evaluate their residual as one HIP graph
constexpr int64_t RESIDUAL_GRAPH_MAX_CELLS = 64 * 64 * 64;

// mixed precision: the FP32 stage's relative divergence tolerance is never
// tighter than this (about 170 FP32 rounding units)
constexpr double MIXED_FP32_REL_DIV_MIN = 1.0e-5;

namespace sca {

class Batcher;
extern thread_local std::string g_err;

void masked_compliance_host(const double* C, const double q[4], const int mask[9], double* S_out);
// throws std::runtime_error naming the first problem of a description that
// a context cannot be built from (checked before anything is allocated)
void validate_rve_desc(const sc_rve_desc& d);
void sym_lstsq(int l, std::vector<double> H, std::vector<double>& b);

// offsets into the small device scratch (doubles); [0, MAX_RED_VEC) holds
// reduction results
enum {
  SM_SUMS = 100,  // 9: P sums of the last response
  SM_ILL = 109,   // int in a double slot: terminally-ill flag of the last response
  SM_DIV = 110,   // 1: divergence sum
  SM_AIM = 120,   // 9: DC of the Gamma convolution
  SM_AV = 130,    // 9: average rate
  SM_AIM2 = 140,  // 9: forward_field aim
  SM_FSUM = 150,  // 9: forward_field sums
  SM_FDIFF = 160, // 9: forward_field correction
  SM_ITERS = 111, // 2: Newton / state iterations of the last response (profiling)
  SM_G81 = 200    // 162: dPdF of the argmax / argmin cells
};

// HIP-event timing of kernel classes (sc_kernel_times), off by default
class KTimer {
 public:
  bool on = false;
  unsigned mask = ~0u;  // classes timed while on (bit SC_KT_*)
  sc_kernel_times acc{};
  void begin(int cls, hipStream_t st);
  void end(int cls, int64_t points, hipStream_t st);
  void flush();  // waits for the recorded events and accumulates them
  void reset() { acc = sc_kernel_times{}; }
  ~KTimer();

 private:
  struct Rec {
    int cls;
    int64_t pts;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  hipEvent_t open_a = nullptr;
  hipEvent_t get();
};

// PETSc SNESNGMRES (default options, see solver.cpp) on device vectors
class Ngmres {
 public:
  Ngmres(int64_t n, hipStream_t st, KTimer* kt = nullptr);
  ~Ngmres();
  // residual(x, f): f = F(x); conv(k): 1 converged, -1 diverged, 0 iterating.
  // TF: element type of the residuals (float for a residual computed in
  // FP32, which it holds exactly); iterates, inner products and the
  // combination are FP64.
  template <typename TF>
  int solve(double* X, int msize, const std::function<void(const double*, TF*)>& residual,
            const std::function<int(int)>& conv, const std::function<void(int)>& set_iter);
  template <typename T>
  std::vector<double> reduce(const std::vector<int>& ops, const std::vector<const T*>& a,
                             const std::vector<const T*>& b);
  // final sums of the first nout partial rows, to the host
  std::vector<double> collect(int nout, int64_t units);
  // called after every synchronisation of the reductions (the owner takes up
  // results it copied asynchronously)
  std::function<void()> after_sync;
  // waits for the stream (the owner's choice of waiting); reductions land
  // in h_out, pinned host memory the final reduction kernel writes
  std::function<void()> wait;
  // the point a solve returned was the point of its last evaluation
  bool solution_evaluated_last = false;
  int64_t n;
  hipStream_t st;
  double *partial = nullptr, *h_out = nullptr;
  KTimer* kt = nullptr;

 private:
  // The vectors of a solve are buffers of one pool, referenced by the
  // iterate (Xc, its residual Fv), the trial points and residuals of an
  // iteration and the history slots (a residual buffer holds n values of
  // the solve's residual type).  Accepting a point or recording it in
  // the history moves a reference instead of copying the vector; a buffer
  // is written only while nothing else references it.  The caller's X is
  // referenced until the first accepted point and receives the solution at
  // the end (one copy per solve).
  double *Fv = nullptr, *Xc = nullptr, *XM = nullptr, *FM = nullptr, *XA = nullptr, *FA = nullptr;
  std::vector<double*> Fh, Xh;
  std::vector<double*> bufs, pool;
  double* take();  // a pool buffer that nothing references
};

// position in a load case (the loop variables of loop_over_loadcase)
struct LoadCursor {
  int l = 0, inc = 0, total_inc = 0, cut_back_level = 0;
  double t = 0.0, t_0 = 0.0, dt = 1.0, dt_prev = 0.0;
  bool cut_back = false, guess = false, step_started = false;
};

class Rve {
 public:
  // residual_graphs: grids of at most RESIDUAL_GRAPH_MAX_CELLS cells
  // evaluate their residual as a HIP graph
  explicit Rve(const sc_rve_desc& d, bool residual_graphs = true);
  ~Rve();
  Rve(const Rve&) = delete;
  Rve& operator=(const Rve&) = delete;

  void init();
  void forward(bool cutBack, bool guess, double dt, double dt_old, double t_remaining, const sc_loadstep& stp);
  sc_solution calculate_solution();
  int run_loadcase(const sc_loadstep* steps, int n_steps, int* inc_iterations, int* inc_evaluations, int max_incs);
  int next_increment(const sc_loadstep* steps, int n_steps, LoadCursor& cur, int* iterations, int* evaluations);
  // r: floats during the FP32 stage of the mixed form (use32), doubles otherwise
  void form_residual(const double* Fdev, void* rdev);
  void launch_residual(const double* Fdev, void* rdev);
  void launch_spectral(void* rdev);
  bool ensure_residual_graph();  // false: evaluate eagerly this time
  void release_residual_graph();
  void constitutive_response(const double* Fdev, double dt);
  void fetch_response_scalars(bool rotate);
  void request_response_scalars(bool rotate, bool residual);
  void complete_response_scalars();  // no-op unless a request is pending
  void tangent();
  void update_coords();
  // mp.F <- the F of the last constitutive response (the response reads F
  // from the solver's vector; the copy is made once, when mp.F is used)
  void sync_response_F();
  // checkpoint of the state between increments (restart, eviction): the
  // material-point state and its converged copy, F, F_last_inc, F_dot, the
  // coordinates and the solver scalars; load() needs a context built from
  // the same descriptor
  void save(std::vector<char>& blob);
  void load(const std::vector<char>& blob);
  // called after a converged increment whose index in its load step is a
  // multiple of the step's f_out, with the total increment count and time
  // (damask_grid.cpp:165-168)
  std::function<void(int, double)> on_result;

  // residual evaluations are computed in joint launches with other contexts
  // (null: on this context's own stream)
  Batcher* batch = nullptr;
  GammaHost gamma_host(bool do_gamma) const;

  // descriptor and geometry
  sc_rve_desc desc;
  std::vector<int> material_at;
  std::vector<double> orientations;
  int cells[3];
  double size[3], scaled[3];
  int64_t N, K;
  int nr;
  double wgt;
  hipStream_t st = nullptr;
  // P sums from the forward transform's x-pass row sums (pencil x lengths)
  // instead of a pass over P after the material-point update
  bool xsums = false;
  double* d_rowsum = nullptr;  // 9 Ny Nz
  // after a forward step that exchanged the roles of the current and start
  // buffers of Fp, S, xi, gamma, until the next material-point update
  bool state_stale = false;
  void refresh_state();  // makes the current set readable (a copy, if stale)
  // polling waits (wait_stream); the job server's contexts wait blocking
  bool spin_wait = true;
  hipEvent_t wait_ev = nullptr;
  void wait_stream();
  MatConst mc;
  MpState mp;
  FftPlan<double> fft;
  // mixed precision: FP32 transforms and Gamma on an FP32 spectrum
  bool mixed = false;
  bool use32 = false;     // the current residual evaluations use the FP32 stages
  int iter_base = 0;      // total_iter offset of an FP64 finishing solve
  // the next evaluation takes the material point as last updated (at the
  // same F) instead of updating it again
  bool reuse_response = false;
  FftPlan<float> fft32;
  float* spec32 = nullptr;
  int64_t evals32 = 0, evals64 = 0;

  // device fields (SoA planes)
  double* X = nullptr;  // F_PETSc
  double *F_last_inc = nullptr, *F_dot = nullptr;
  Ngmres* ngm = nullptr;
  double* spec = nullptr;
  double* tpart = nullptr;  // the tangent's plane sums per 16 cells (81 x tangent_partials(N))
  double* sub = nullptr;
  double* partial = nullptr;
  double* small = nullptr;
  double* d_xi_tab = nullptr;
  int* d_ill = nullptr;
  long long* d_minmax_i = nullptr;
  double *d_u = nullptr, *d_xp = nullptr, *d_xn = nullptr, *d_fsum = nullptr;  // coordinates
  double* d_n2 = nullptr;  // per-cell |dPdF|^2
  double* d_scr_partial = nullptr;  // partial sums of the iteration counters
  const double* f_eval = nullptr;   // F of the last response, until sync_response_F
  double* h_scal = nullptr;  // pinned: response scalars in flight
  // residual evaluation as a HIP graph (grids of at most
  // RESIDUAL_GRAPH_MAX_CELLS cells): staging of F and r, the graph and the
  // arguments it was captured with
  double *g_in = nullptr, *g_out = nullptr;
  hipGraph_t g_graph = nullptr;
  hipGraphExec_t g_exec = nullptr;
  std::vector<char> g_sig, g_seen;
  int64_t stats_graph_captures = 0, stats_graph_updates = 0;
  bool scal_pending = false, scal_rotate = false, scal_residual = false, scal_reused = false;

  // MechBase state (mech_base.h:75-107), column-major
  double C_ref[81], C_volAvg[81], C_minMaxAvg[81], C_volAvgLastInc[81], C_minMaxAvgLastInc[81], S[81];
  double P_av[9], P_aim[9], F_aim_dot[9], F_aim[9], F_aim_last_inc[9];
  int stress_mask[9];
  double rot_q[4];
  double delta_t = 0.0, last_dt = 0.0, last_div_sum = 0.0;
  // Homogeneous fields, known on the host: X (F of the solver) after init,
  // F_last_inc forwarded from it and the rate of an increment without a rate
  // guess; their common 9 values.  forward() averages a homogeneous
  // forwarded field the way the reference's sum does (see forward).
  bool x_uniform = false, last_uniform = false, rate_uniform = false;
  double x_value[9] = {}, last_value[9] = {}, rate_value[9] = {};
  int total_iter = 0;
  double err_BC = 0.0, err_div = 0.0;
  bool terminally_ill = false;
  int snes_iter = 0, nfuncs = 0, solve_evals = 0;
  int solve_updates = 0;  // evaluations of this solve that updated the material point
  sc_stats stats;
  KTimer kt;

 private:
  std::vector<double*> allocs;
  void setup(const sc_rve_desc& d, bool residual_graphs);
  void release();
  double* dalloc(size_t n);
  int converged() const;
  int snes_ngmres();
  std::vector<double> reduce(const std::vector<int>& ops, const std::vector<const double*>& a,
                             const std::vector<const double*>& b, int64_t n);
};

// host copy of a field of a context, sc_rve_get_field's layouts (capi.cpp)
int rve_get_field(Rve& v, int field, double* out, int64_t n);

}  // namespace sca
