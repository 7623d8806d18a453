// Device-resident spectral basic solver for one RVE, host side.
//
// Mirrors src/grid/mech_solver_basic.cpp (init :19-108, calculate_solution
// :110-161, forward :163-220, converged :242-284, form_residual :286-339),
// src/grid/mech_base.cpp (forward_field :213-247, calculate_masked_compliance
// :249-300, calculate_rate :479-497), src/grid/spectral.cpp
// (constitutive_response :132-219) and src/grid/damask_grid.cpp
// (loop_over_loadcase :82-196), with PETSc's SNESNGMRES as the outer solver
// (PETSc >= 3.12, < 3.21 defaults: m = 30, restart_it = 2, gammaA = gammaC
// = 2, epsilonB = 0.1, deltaB = 0.9, difference selection and restart, basic
// line search).  Fields live on the GPU; per residual evaluation the host
// receives 11 scalars (P_av sums, divergence sum, terminally-ill flag) and
// per N-GMRES iteration a few dot products.
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "batch.h"
#include "rve.h"

namespace sca {

thread_local std::string g_err;

// ---- small host tensor helpers (column-major (3,3) and (3,3,3,3)) ------------
namespace h {
inline double& at(double* m, int i, int j) { return m[i + 3 * j]; }
inline double at(const double* m, int i, int j) { return m[i + 3 * j]; }
void mul(const double* A, const double* B, double* C) {
  double t[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += at(A, i, k) * at(B, k, j);
      t[i + 3 * j] = s;
    }
  std::memcpy(C, t, sizeof t);
}
void tr(const double* A, double* B) {
  double t[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[i + 3 * j] = at(A, j, i);
  std::memcpy(B, t, sizeof t);
}
// rotate_tensor2, rotations.f90:289-301 (passive R T R^T)
void rot2(const double q[4], const double* T, double* out) {
  double R[9], Rt[9], t[9];
  qu2om_host(q, R);
  tr(R, Rt);
  mul(R, T, t);
  mul(t, Rt, out);
}
// active R^T T R: load frame -> field frame
void rot2_active(const double q[4], const double* T, double* out) {
  double R[9], Rt[9], t[9];
  qu2om_host(q, R);
  tr(R, Rt);
  mul(Rt, T, t);
  mul(t, R, out);
}
// rotate_tensor4, rotations.f90:308-328 (passive)
void rot4(const double q[4], const double* T, double* out) {
  double R[9];
  qu2om_host(q, R);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          double s = 0.0;
          for (int m = 0; m < 3; ++m)
            for (int n = 0; n < 3; ++n)
              for (int o = 0; o < 3; ++o)
                for (int p = 0; p < 3; ++p)
                  s += at(R, i, m) * at(R, j, n) * at(R, k, o) * at(R, l, p) * T[m + 3 * n + 9 * o + 27 * p];
          out[i + 3 * j + 9 * k + 27 * l] = s;
        }
}
// math_mul3333xx33
void mul3333xx33(const double* A, const double* B, double* out) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int n = 0; n < 3; ++n)
        for (int m = 0; m < 3; ++m) s += A[i + 3 * j + 9 * m + 27 * n] * B[m + 3 * n];
      out[i + 3 * j] = s;
    }
}
// LU inverse with partial pivoting (row-major n x n); returns false on a zero pivot
bool invert(int n, const double* a, double* inv) {
  std::vector<double> lu(a, a + n * n);
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double mx = std::fabs(lu[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(lu[i * n + k]) > mx) {
        mx = std::fabs(lu[i * n + k]);
        p = i;
      }
    piv[k] = p;
    if (lu[p * n + k] == 0.0) return false;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(lu[k * n + j], lu[p * n + j]);
    const double d = 1.0 / lu[k * n + k];
    for (int i = k + 1; i < n; ++i) lu[i * n + k] *= d;
    for (int i = k + 1; i < n; ++i)
      for (int j = k + 1; j < n; ++j) lu[i * n + j] -= lu[i * n + k] * lu[k * n + j];
  }
  std::vector<double> col(n);
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) col[i] = (i == c) ? 1.0 : 0.0;
    for (int k = 0; k < n; ++k)
      if (piv[k] != k) std::swap(col[k], col[piv[k]]);
    for (int i = 1; i < n; ++i)
      for (int j = 0; j < i; ++j) col[i] -= lu[i * n + j] * col[j];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) col[i] -= lu[i * n + j] * col[j];
      col[i] /= lu[i * n + i];
    }
    for (int i = 0; i < n; ++i) inv[i * n + c] = col[i];
  }
  return true;
}
}  // namespace h

void masked_compliance_host(const double* C, const double q[4], const int mask[9], double* S_out) {
  bool m1d[9];
  int nred = 0;
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row) {
      m1d[col * 3 + row] = !mask[row + 3 * col];
      nred += m1d[col * 3 + row] ? 1 : 0;
    }
  double t99[81] = {0};  // plain map, row-major (p,q), p = 3i+j
  if (nred > 0) {
    double rot[81];
    h::rot4(q, C, rot);
    double c99[81];
    for (int p = 0; p < 9; ++p)
      for (int qq = 0; qq < 9; ++qq) {
        const int i = p / 3, j = p % 3, k = qq / 3, l = qq % 3;
        c99[p * 9 + qq] = rot[i + 3 * j + 9 * k + 27 * l];
      }
    std::vector<double> cr(nred * nred), sr(nred * nred);
    int idx = 0;
    for (int i = 0; i < 9; ++i)
      for (int j = 0; j < 9; ++j)
        if (m1d[i] && m1d[j]) {
          cr[(idx / nred) * nred + idx % nred] = c99[i * 9 + j];
          ++idx;
        }
    if (!h::invert(nred, cr.data(), sr.data())) throw std::runtime_error("Matrix inversion error");
    idx = 0;
    for (int i = 0; i < 9; ++i)
      for (int j = 0; j < 9; ++j)
        if (m1d[i] && m1d[j]) {
          t99[i * 9 + j] = sr[(idx / nred) * nred + idx % nred];
          ++idx;
        }
  }
  for (int p = 0; p < 9; ++p)
    for (int qq = 0; qq < 9; ++qq) {
      const int i = p / 3, j = p % 3, k = qq / 3, l = qq % 3;
      S_out[i + 3 * j + 9 * k + 27 * l] = t99[p * 9 + qq];
    }
}

// min-norm least squares of a symmetric l x l system via cyclic Jacobi,
// singular values <= eps*s_max dropped (LAPACK dgelss, rcond = -1)
void sym_lstsq(int l, std::vector<double> H, std::vector<double>& b) {
  std::vector<double> V(l * l, 0.0);
  for (int i = 0; i < l; ++i) V[i * l + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < l; ++p)
      for (int q = p + 1; q < l; ++q) off += H[p * l + q] * H[p * l + q];
    if (off < 1e-300) break;
    for (int p = 0; p < l; ++p)
      for (int q = p + 1; q < l; ++q) {
        const double apq = H[p * l + q];
        if (apq == 0.0) continue;
        const double theta = (H[q * l + q] - H[p * l + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < l; ++k) {
          const double a = H[k * l + p], bb = H[k * l + q];
          H[k * l + p] = c * a - s * bb;
          H[k * l + q] = s * a + c * bb;
        }
        for (int k = 0; k < l; ++k) {
          const double a = H[p * l + k], bb = H[q * l + k];
          H[p * l + k] = c * a - s * bb;
          H[q * l + k] = s * a + c * bb;
        }
        for (int k = 0; k < l; ++k) {
          const double a = V[k * l + p], bb = V[k * l + q];
          V[k * l + p] = c * a - s * bb;
          V[k * l + q] = s * a + c * bb;
        }
      }
  }
  double smax = 0.0;
  for (int i = 0; i < l; ++i) smax = std::max(smax, std::fabs(H[i * l + i]));
  const double thr = std::max(2.220446049250313e-16 * smax, 2.2250738585072014e-308);
  std::vector<double> x(l, 0.0);
  for (int i = 0; i < l; ++i) {
    const double lam = H[i * l + i];
    if (std::fabs(lam) <= thr) continue;
    double proj = 0.0;
    for (int k = 0; k < l; ++k) proj += V[k * l + i] * b[k];
    for (int k = 0; k < l; ++k) x[k] += V[k * l + i] * proj / lam;
  }
  b = x;
}

// ---------------------------------------------------------------------------
hipEvent_t KTimer::get() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e;
  SCA_HIP(hipEventCreate(&e));
  return e;
}
void KTimer::begin(int cls, hipStream_t st) {
  if (!on || !((mask >> cls) & 1u)) return;
  open_a = get();
  SCA_HIP(hipEventRecord(open_a, st));
}
void KTimer::end(int cls, int64_t points, hipStream_t st) {
  if (!on || !open_a) return;
  hipEvent_t b = get();
  SCA_HIP(hipEventRecord(b, st));
  pending.push_back({cls, points, open_a, b});
  open_a = nullptr;
}
void KTimer::flush() {
  if (pending.empty()) return;
  SCA_HIP(hipEventSynchronize(pending.back().b));
  for (const Rec& r : pending) {
    float ms = 0.0f;
    SCA_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    acc.ms[r.cls] += ms;
    acc.launches[r.cls] += 1;
    acc.points[r.cls] += r.pts;
    pool.push_back(r.a);
    pool.push_back(r.b);
  }
  pending.clear();
}
KTimer::~KTimer() {
  for (const Rec& r : pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  if (open_a) (void)hipEventDestroy(open_a);
}

// ---------------------------------------------------------------------------
double* Rve::dalloc(size_t n) {
  double* p = static_cast<double*>(dev_alloc(std::max<size_t>(n, 1) * sizeof(double)));
  allocs.push_back(p);
  return p;
}

// Every check a context's set-up makes, before anything is allocated: the
// job server runs it when a job is added, so that a job it accepts can be
// started.
void validate_rve_desc(const sc_rve_desc& d) {
  auto bad = [](const std::string& m) { throw std::runtime_error("invalid RVE description: " + m); };
  if (d.cells[0] < 2) bad("cells[0] must be larger than 1");
  for (int p = 0; p < 3; ++p) {
    if (d.cells[p] < 1) bad("cells must be positive");
    if (!(d.size[p] > 0.0)) bad("size must be positive");
    int m = d.cells[p];
    for (int f : {2, 3, 5, 7, 11, 13})
      while (m % f == 0) m /= f;
    if (m != 1) bad("grid length " + std::to_string(d.cells[p]) + " has a prime factor above 13");
  }
  const int64_t N = (int64_t)d.cells[0] * d.cells[1] * d.cells[2];
  if (N > ((int64_t)1 << 29)) bad("more than 2^29 cells");
  if (d.n_materials < 1) bad("no materials");
  if (!d.material_at || !d.orientations) bad("null material map or orientations");
  for (int64_t c = 0; c < N; ++c)
    if (d.material_at[c] < 0 || d.material_at[c] >= d.n_materials) bad("material_at out of range");
  if (d.precision != SC_PRECISION_FP64 && d.precision != SC_PRECISION_MIXED) bad("unknown precision");
  if (d.grid_num.derivative < 0 || d.grid_num.derivative > 2) bad("unknown derivative");
  if (d.grid_num.divergence_correction < 0 || d.grid_num.divergence_correction > 2)
    bad("unknown divergence correction");
}

Rve::Rve(const sc_rve_desc& d, bool residual_graphs) {
  validate_rve_desc(d);
  try {
    setup(d, residual_graphs);
  } catch (...) {
    release();  // a throwing constructor runs no destructor
    throw;
  }
}

void Rve::setup(const sc_rve_desc& d, bool residual_graphs) {
  desc = d;
  for (int p = 0; p < 3; ++p) {
    cells[p] = d.cells[p];
    size[p] = d.size[p];
  }
  N = (int64_t)cells[0] * cells[1] * cells[2];
  nr = cells[0] / 2 + 1;
  K = (int64_t)nr * cells[1] * cells[2];
  wgt = std::pow((double)(cells[0] * cells[1] * cells[2]), -1);
  material_at.assign(d.material_at, d.material_at + N);
  orientations.assign(d.orientations, d.orientations + 4 * d.n_materials);
  desc.material_at = material_at.data();
  desc.orientations = orientations.data();
  // scaled geometry, mech_base.cpp:11-33
  {
    int idx[3] = {0, 1, 2};
    const int dc = d.grid_num.divergence_correction;
    if (dc == 1) {
      std::sort(idx, idx + 3, [&](int a, int b) { return size[a] < size[b]; });
      for (int j = 0; j < 3; ++j) scaled[j] = size[j] / size[idx[1]];
    } else if (dc == 2) {
      double ng[3] = {size[0] / cells[0], size[1] / cells[1], size[2] / cells[2]};
      std::sort(idx, idx + 3, [&](int a, int b) { return ng[a] < ng[b]; });
      for (int j = 0; j < 3; ++j) scaled[j] = size[j] / size[idx[1]] * cells[idx[1]];
    } else {
      for (int j = 0; j < 3; ++j) scaled[j] = size[j];
    }
  }
  build_matconst(d.material, d.phase_num, mc);
  // non-blocking: contexts driven from several host threads never wait on
  // each other through the null stream
  st = stream_get();
  fft.init(cells);
  if (d.grid_num.fft_plan_mode == 1) fft.tune(9, st);
  SCA_HIP(hipEventCreateWithFlags(&wait_ev, hipEventDisableTiming));
  ngm = new Ngmres(9 * N, st, &kt);
  ngm->after_sync = [this] { complete_response_scalars(); };
  ngm->wait = [this] { wait_stream(); };
  h_scal = static_cast<double*>(host_alloc(16 * sizeof(double)));
  mixed = d.precision == SC_PRECISION_MIXED;
  if (mixed) {
    fft32.init(cells);
    if (d.grid_num.fft_plan_mode == 1) fft32.tune(9, st);
    spec32 = static_cast<float*>(dev_alloc(2 * 9 * K * sizeof(float)));
  }
  // material point planes
  mp.N = N;
  mp.F = dalloc(9 * N);
  mp.F0 = dalloc(9 * N);
  mp.Fp0 = dalloc(9 * N);
  mp.Lp0 = dalloc(9 * N);
  mp.S0 = dalloc(9 * N);
  mp.xi0 = dalloc(12 * N);
  mp.gamma0 = dalloc(12 * N);
  mp.Lp = dalloc(9 * N);
  mp.P = dalloc(9 * N);
  mp.Fp = dalloc(9 * N);
  mp.S = dalloc(9 * N);
  mp.Fe = dalloc(9 * N);
  mp.xi = dalloc(12 * N);
  mp.gamma = dalloc(12 * N);
  mp.iters = static_cast<int*>(dev_alloc(2 * N * sizeof(int)));
  mp.repivot = static_cast<int*>(dev_alloc((N + 1) * sizeof(int)));
  sub = dalloc((size_t)mp_sub_planes() * N);
  X = dalloc(9 * N);
  F_last_inc = dalloc(9 * N);
  F_dot = dalloc(9 * N);
  spec = dalloc(2 * 9 * K);
  tpart = dalloc((size_t)81 * tangent_partials(N));
  xsums = fft.row_sums_available();
  if (xsums) d_rowsum = dalloc((size_t)9 * cells[1] * cells[2]);
  partial = dalloc((size_t)std::max<int64_t>(red_blocks() * MAX_RED_VEC, 9 * mp_partial_blocks(N)) +
                   gamma_blocks(K));
  if (residual_graphs && N <= RESIDUAL_GRAPH_MAX_CELLS) {  // staging of the captured residual evaluation
    g_in = dalloc(9 * N);
    g_out = dalloc(9 * N);
  }
  small = dalloc(512);
  d_u = dalloc(3 * N);
  d_xp = dalloc(3 * N);
  d_xn = dalloc((size_t)3 * (cells[0] + 1) * (cells[1] + 1) * (cells[2] + 1));
  d_fsum = dalloc(9);
  d_n2 = dalloc(N);
  d_scr_partial = dalloc((size_t)2 * red_blocks());
  d_ill = reinterpret_cast<int*>(small + SM_ILL);
  d_minmax_i = static_cast<long long*>(dev_alloc(2 * red_blocks() * sizeof(long long)));
  if (d.grid_num.derivative != 0) {
    // xi table in the kernel layout: (3 complex) per k = i + nr*(y + Ny*z)
    std::vector<double> tab(6 * K);
    const int Ny = cells[1], Nz = cells[2];
    const double TAU = 2.0 * M_PI;
    for (int z = 0; z < Nz; ++z)
      for (int y = 0; y < Ny; ++y)
        for (int i = 0; i < nr; ++i) {
          const int ks[3] = {i, (y > Ny / 2) ? y - Ny : y, (z > Nz / 2) ? z - Nz : z};
          const int64_t kk = i + (int64_t)nr * (y + (int64_t)Ny * z);
          for (int p = 0; p < 3; ++p) {
            double re, im;
            if (d.grid_num.derivative == 1) {
              re = 0.0;
              im = std::sin(TAU * ks[p] / cells[p]) / (2.0 * size[p] / cells[p]);
            } else {
              const int p1 = (p + 1) % 3, p2 = (p + 2) % 3;
              std::complex<double> v = std::complex<double>(std::cos(TAU * ks[p] / cells[p]) - (p == 0 ? 1.0 : -1.0),
                                                            std::sin(TAU * ks[p] / cells[p])) *
                                       std::complex<double>(std::cos(TAU * ks[p1] / cells[p1]) + 1.0,
                                                            std::sin(TAU * ks[p1] / cells[p1])) *
                                       std::complex<double>(std::cos(TAU * ks[p2] / cells[p2]) + 1.0,
                                                            std::sin(TAU * ks[p2] / cells[p2])) /
                                       std::complex<double>(4.0 * size[p] / cells[p], 0.0);
              re = v.real();
              im = v.imag();
            }
            tab[6 * kk + 2 * p] = re;
            tab[6 * kk + 2 * p + 1] = im;
          }
        }
    d_xi_tab = dalloc(6 * K);
    SCA_HIP(hipMemcpyAsync(d_xi_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, st));
    SCA_HIP(hipStreamSynchronize(st));
  }
  // mechanical_init: F = I, Fp = O_0, Fe = Fp^T, Lp = S = 0, xi = xi_0, gamma = 0
  {
    std::vector<double> fp(9 * N);
    std::vector<double> om(9 * (size_t)d.n_materials);
    for (int m = 0; m < d.n_materials; ++m) qu2om_host(&orientations[4 * m], &om[9 * m]);
    for (int64_t c = 0; c < N; ++c) {
      const int ma = material_at[c];  // in range: validate_rve_desc
      for (int q = 0; q < 9; ++q) fp[q * N + c] = om[9 * ma + q];
    }
    double* d_fp = dalloc(9 * N);
    SCA_HIP(hipMemcpyAsync(d_fp, fp.data(), fp.size() * sizeof(double), hipMemcpyHostToDevice, st));
    launch_mp_init(mp, d_fp, d.material.xi_0_sl, st);
    SCA_HIP(hipStreamSynchronize(st));
    dev_free(d_fp);
    allocs.pop_back();
  }
  std::memset(C_ref, 0, sizeof C_ref);
  std::memset(C_volAvg, 0, sizeof C_volAvg);
  std::memset(C_minMaxAvg, 0, sizeof C_minMaxAvg);
  std::memset(C_volAvgLastInc, 0, sizeof C_volAvgLastInc);
  std::memset(C_minMaxAvgLastInc, 0, sizeof C_minMaxAvgLastInc);
  std::memset(S, 0, sizeof S);
  std::memset(P_av, 0, sizeof P_av);
  std::memset(P_aim, 0, sizeof P_aim);
  std::memset(F_aim_dot, 0, sizeof F_aim_dot);
  for (int q = 0; q < 9; ++q) F_aim[q] = F_aim_last_inc[q] = (q % 4 == 0) ? 1.0 : 0.0;
  for (int q = 0; q < 9; ++q) stress_mask[q] = 0;
  rot_q[0] = 1.0;
  rot_q[1] = rot_q[2] = rot_q[3] = 0.0;
  std::memset(&stats, 0, sizeof stats);
}

Rve::~Rve() { release(); }

// frees whatever set-up allocated (also after a set-up that threw part way)
void Rve::release() {
  // the blocks go back to the library's caches: nothing may still use them
  if (st) (void)hipStreamSynchronize(st);
  release_residual_graph();
  delete ngm;
  ngm = nullptr;
  if (wait_ev) (void)hipEventDestroy(wait_ev);
  wait_ev = nullptr;
  host_free(h_scal);
  h_scal = nullptr;
  dev_free(spec32);
  spec32 = nullptr;
  fft32.release();
  for (double* p : allocs) dev_free(p);
  allocs.clear();
  dev_free(mp.iters);
  dev_free(mp.repivot);
  mp.iters = nullptr;
  mp.repivot = nullptr;
  dev_free(d_minmax_i);
  d_minmax_i = nullptr;
  fft.release();
  stream_put(st);
  st = nullptr;
}

// deterministic device reduction of nvec vectors of length n -> host values
std::vector<double> Rve::reduce(const std::vector<int>& ops, const std::vector<const double*>& a,
                                const std::vector<const double*>& b, int64_t n) {
  const int nvec = (int)ops.size();
  std::vector<double> out(nvec);
  for (int v0 = 0; v0 < nvec; v0 += MAX_RED_VEC) {
    const int nv = std::min(MAX_RED_VEC, nvec - v0);
    launch_partial_reduce<double>(ops.data() + v0, n, nv, a.data() + v0, b.empty() ? nullptr : b.data() + v0,
                                  partial, st);
    launch_final_reduce(partial, nv, red_blocks(), small, st);
    SCA_HIP(hipMemcpyAsync(out.data() + v0, small, nv * sizeof(double), hipMemcpyDeviceToHost, st));
    SCA_HIP(hipStreamSynchronize(st));
  }
  return out;
}

// Spectral::constitutive_response (spectral.cpp:132-219) up to the P_av sums
// on the device; the tangent reductions are deferred to tangent().
// the current Fp, S, xi and gamma after a forward step that exchanged the
// buffers' roles: the start values (what the forward copy would have left)
void Rve::refresh_state() {
  if (!state_stale) return;
  launch_mp_copy(mp, 2, st);
  state_stale = false;
}

void Rve::constitutive_response(const double* Fdev, double dt) {
  state_stale = false;  // the update writes the current set of every cell
  MpState s = mp;
  s.F = const_cast<double*>(Fdev);  // read only
  f_eval = Fdev;
  kt.begin(SC_KT_RESPONSE, st);
  launch_mp_response(s, mc, dt, d_ill, sub, st);
  kt.end(SC_KT_RESPONSE, N, st);
  kt.begin(SC_KT_REDUCE, st);
  launch_pav_partial(mp.P, N, partial, st);
  launch_final_reduce(partial, 9, mp_partial_blocks(N), small + SM_SUMS, st);
  kt.end(SC_KT_REDUCE, N, st);
  if (kt.on) {  // iteration counts of this evaluation
    const int ops[2] = {2, 2};
    const int* a[2] = {mp.iters, mp.iters + N};
    launch_partial_reduce<int>(ops, N, 2, a, nullptr, d_scr_partial, st);
    launch_final_reduce(d_scr_partial, 2, red_blocks(), small + SM_ITERS, st);
  }
  last_dt = dt;
}

void Rve::sync_response_F() {
  if (f_eval && f_eval != mp.F)
    SCA_HIP(hipMemcpyAsync(mp.F, f_eval, 9 * N * sizeof(double), hipMemcpyDeviceToDevice, st));
  f_eval = mp.F;
}

// The response scalars (P sums, the terminally-ill flag, the divergence sum
// and, profiling, the iteration counts: adjacent slots) travel in one
// asynchronous copy to pinned memory; they are taken up at the next stream
// synchronisation the host makes anyway (the N-GMRES reductions), so an
// evaluation costs no round trip of its own.
void Rve::request_response_scalars(bool rotate, bool residual) {
  static_assert(SM_ILL == SM_SUMS + 9 && SM_DIV == SM_ILL + 1 && SM_ITERS == SM_DIV + 1,
                "response scalars must be adjacent");
  SCA_HIP(hipMemcpyAsync(h_scal, small + SM_SUMS, (kt.on ? 13 : 11) * sizeof(double), hipMemcpyDeviceToHost, st));
  scal_pending = true;
  scal_rotate = rotate;
  scal_residual = residual;
  scal_reused = false;
}

// Waits for the work issued on the stream.  A solve waits about twice per
// N-GMRES iteration, each time for a few scalars: a standalone context polls
// an event (bounded, then the blocking wait), which wakes up sooner than the
// runtime's yielding wait; jobs of the server, whose worker threads share the
// host cores, wait blocking.
void Rve::wait_stream() {
  if (!spin_wait) {
    SCA_HIP(hipStreamSynchronize(st));
    return;
  }
  SCA_HIP(hipEventRecord(wait_ev, st));
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t e = hipEventQuery(wait_ev);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) SCA_HIP(e);
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) break;
  }
  SCA_HIP(hipEventSynchronize(wait_ev));
}

void Rve::complete_response_scalars() {
  if (!scal_pending) return;
  wait_stream();
  scal_pending = false;
  int ill = 0;
  std::memcpy(&ill, &h_scal[9], sizeof ill);
  for (int q = 0; q < 9; ++q) P_av[q] = h_scal[q] * wgt;
  if (scal_rotate) h::rot2(rot_q, P_av, P_av);
  last_div_sum = h_scal[10];
  if (ill) terminally_ill = true;
  if (kt.on && !scal_reused) {  // a reused update's iterations are counted once
    stats.newton_iterations += (int64_t)h_scal[11];
    stats.state_iterations += (int64_t)h_scal[12];
  }
  kt.flush();
  if (!scal_residual) return;
  // the host part of form_residual (mech_solver_basic.cpp:286-339)
  err_div = std::sqrt(last_div_sum) * wgt;
  if (cells[0] == 1) err_div = err_div / 2;
  double dF[9];
  h::mul3333xx33(S, P_av, dF);
  for (int q = 0; q < 9; ++q) F_aim[q] = F_aim[q] - dF[q];
  double e = 0.0;
  for (int q = 0; q < 9; ++q)
    if (!stress_mask[q]) e = std::max(e, std::fabs(P_av[q] - P_aim[q]));
  err_BC = e;
}

void Rve::fetch_response_scalars(bool rotate) {
  request_response_scalars(rotate, false);
  complete_response_scalars();
}

// consistent tangent of the last evaluation: C_volAvg, C_minMaxAvg
// (spectral.cpp:189-217).  The tangent is reduced where it is formed (no
// dPdF field): per-cell norms and per-block plane sums; the two cells of
// the extreme norms are computed again for their dPdF.
void Rve::tangent() {
  complete_response_scalars();
  sync_response_F();
  kt.begin(SC_KT_TANGENT, st);
  launch_mp_tangent_reduce(mp, mc, last_dt, d_n2, tpart, st);
  kt.end(SC_KT_TANGENT, N, st);
  // C_volAvg: the blocks' plane sums in the fixed reduction trees
  const int rb = red_blocks();
  const int64_t nb = tangent_partials(N);
  int ops[81];
  const double* av[81];
  for (int q = 0; q < 81; ++q) {
    ops[q] = 2;
    av[q] = tpart + (size_t)q * nb;
  }
  launch_partial_reduce<double>(ops, nb, 81, av, nullptr, partial, st);
  launch_final_reduce(partial, 81, rb, small, st);
  double sums[81];
  SCA_HIP(hipMemcpyAsync(sums, small, sizeof sums, hipMemcpyDeviceToHost, st));
  // the cells of the extreme norms (the selection synchronises the stream)
  int64_t imx = -1, imn = -1;
  tangent_select_norms(d_n2, N, partial + (size_t)81 * rb, d_minmax_i, imx, imn, st);
  for (int q = 0; q < 81; ++q) C_volAvg[q] = sums[q] * wgt;
  double dmax[81], dmin[81];
  for (int q = 0; q < 81; ++q) {
    dmax[q] = 0.0;
    dmin[q] = std::numeric_limits<double>::max();
  }
  int64_t cl[2];
  int n = 0;
  if (imx >= 0) cl[n++] = imx;
  if (imn >= 0) cl[n++] = imn;
  for (int i = 0; i < n; ++i)
    if (cl[i] >= N) throw std::runtime_error("tangent: selected cell out of range");
  if (n > 0) {
    static_assert(sizeof(long long) == sizeof(int64_t), "cell list in the index scratch");
    SCA_HIP(hipMemcpyAsync(d_minmax_i, cl, n * sizeof(int64_t), hipMemcpyHostToDevice, st));
    launch_mp_tangent_cells(mp, mc, last_dt, reinterpret_cast<const int64_t*>(d_minmax_i), n, small + SM_G81, st);
    double h[162];
    SCA_HIP(hipMemcpyAsync(h, small + SM_G81, n * 81 * sizeof(double), hipMemcpyDeviceToHost, st));
    SCA_HIP(hipStreamSynchronize(st));
    int i = 0;
    if (imx >= 0) std::memcpy(dmax, h + 81 * i++, sizeof dmax);
    if (imn >= 0) std::memcpy(dmin, h + 81 * i, sizeof dmin);
  }
  for (int q = 0; q < 81; ++q) C_minMaxAvg[q] = 0.5 * (dmax[q] + dmin[q]);
  kt.flush();
}

GammaHost Rve::gamma_host(bool do_gamma) const {
  GammaHost g;
  for (int p = 0; p < 3; ++p) {
    g.cells[p] = cells[p];
    g.size[p] = size[p];
    g.rescale[p] = size[p] / scaled[p];
  }
  // C_ref (column-major i+3j+9k+27l) -> kernel layout 27l+9n+3m+o for C(l,n,m,o)
  for (int l = 0; l < 3; ++l)
    for (int n = 0; n < 3; ++n)
      for (int m = 0; m < 3; ++m)
        for (int o = 0; o < 3; ++o) g.C_ref[27 * l + 9 * n + 3 * m + o] = C_ref[l + 3 * n + 9 * m + 27 * o];
  g.derivative = desc.grid_num.derivative;
  g.do_gamma = do_gamma ? 1 : 0;
  return g;
}

// MechSolverBasic::form_residual, mech_solver_basic.cpp:286-339
// The device work of one residual evaluation on st, without host
// synchronisation: material-point update at F, P sums, the aim of the
// convolution, forward transform, Gamma convolution with the divergence,
// inverse transform into r, and the asynchronous copy of the response
// scalars.  Its launches depend on F, r, delta_t, the compliance S, the load
// rotation, C_ref and the precision phase only, so it can be captured.
void Rve::launch_residual(const double* Fdev, void* rdev) {
  state_stale = false;  // the update writes the current set of every cell
  MpState s = mp;
  s.F = const_cast<double*>(Fdev);  // read only
  kt.begin(SC_KT_RESPONSE, st);
  launch_mp_response(s, mc, delta_t, d_ill, sub, st);
  kt.end(SC_KT_RESPONSE, N, st);
  if (!xsums) {  // else the forward transform gives the P sums
    kt.begin(SC_KT_REDUCE, st);
    launch_pav_partial(mp.P, N, partial, st);
    launch_final_reduce(partial, 9, mp_partial_blocks(N), small + SM_SUMS, st);
    kt.end(SC_KT_REDUCE, N, st);
  }
  if (kt.on) {  // iteration counts of this evaluation
    const int ops[2] = {2, 2};
    const int* it[2] = {mp.iters, mp.iters + N};
    launch_partial_reduce<int>(ops, N, 2, it, nullptr, d_scr_partial, st);
    launch_final_reduce(d_scr_partial, 2, red_blocks(), small + SM_ITERS, st);
  }
  launch_spectral(rdev);
}

// The spectral part of an evaluation on the P of the last material-point
// update: the aim of the convolution, forward transform, Gamma convolution
// with the divergence, inverse transform into r, and the asynchronous copy of
// the response scalars.
void Rve::launch_spectral(void* rdev) {
  // DC of the convolution: rotated S : P_av, computed on the device once the
  // P sums are complete (with xsums, from the forward transform's row sums)
  double R[9], Rrm[9];
  qu2om_host(rot_q, R);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Rrm[3 * i + j] = R[i + 3 * j];
  auto sums_and_aim = [&] {
    if (xsums) launch_rowsum_reduce(RowSumItem{d_rowsum, (int64_t)cells[1] * cells[2], small + SM_SUMS}, st);
    launch_bc_aim(small + SM_SUMS, S, Rrm, wgt, 1, small + SM_AIM, st);
  };
  double* rowsum = xsums ? d_rowsum : nullptr;
  if (use32) {  // FP32 transforms and spectrum from the FP64 P; r out in FP32 (exactly the transform's values)
    kt.begin(SC_KT_FFT_FWD, st);
    fft32.forward(mp.P, spec32, 9, st, 0, rowsum);
    kt.end(SC_KT_FFT_FWD, N, st);
    sums_and_aim();
    kt.begin(SC_KT_GAMMA, st);
    launch_gamma_div<float>(spec32, K, gamma_host(true), small + SM_AIM, d_xi_tab, partial, st);
    launch_final_reduce(partial, 1, gamma_blocks(K), small + SM_DIV, st);
    kt.end(SC_KT_GAMMA, K, st);
    kt.begin(SC_KT_FFT_BWD, st);
    fft32.backward(spec32, static_cast<float*>(rdev), 9, 1.0, st);
    kt.end(SC_KT_FFT_BWD, N, st);
  } else {
    kt.begin(SC_KT_FFT_FWD, st);
    fft.forward(mp.P, spec, 9, st, 0, rowsum);
    kt.end(SC_KT_FFT_FWD, N, st);
    sums_and_aim();
    kt.begin(SC_KT_GAMMA, st);
    launch_gamma_div<double>(spec, K, gamma_host(true), small + SM_AIM, d_xi_tab, partial, st);
    launch_final_reduce(partial, 1, gamma_blocks(K), small + SM_DIV, st);
    kt.end(SC_KT_GAMMA, K, st);
    kt.begin(SC_KT_FFT_BWD, st);
    fft.backward(spec, static_cast<double*>(rdev), 9, 1.0, st);
    kt.end(SC_KT_FFT_BWD, N, st);
  }
  static_assert(SM_ILL == SM_SUMS + 9 && SM_DIV == SM_ILL + 1 && SM_ITERS == SM_DIV + 1,
                "response scalars must be adjacent");
  SCA_HIP(hipMemcpyAsync(h_scal, small + SM_SUMS, (kt.on ? 13 : 11) * sizeof(double), hipMemcpyDeviceToHost, st));
}

// Graph of launch_residual(g_in, g_out) for the current arguments.  A
// signature seen for the first time is evaluated eagerly (returns false): a
// capture pays off only when the arguments repeat.  A new capture updates the
// executable graph in place when only kernel arguments changed (the launch
// sequence is the same), instead of instantiating it again.
bool Rve::ensure_residual_graph() {
  struct Sig {
    double dt, S[81], rot_q[4], C_ref[81];
    int use32;
  } sig{};
  sig.dt = delta_t;
  std::memcpy(sig.S, S, sizeof sig.S);
  std::memcpy(sig.rot_q, rot_q, sizeof sig.rot_q);
  std::memcpy(sig.C_ref, C_ref, sizeof sig.C_ref);
  sig.use32 = use32 ? 1 : 0;
  const char* sb = reinterpret_cast<const char*>(&sig);
  auto same = [&](const std::vector<char>& v) { return v.size() == sizeof sig && std::memcmp(v.data(), sb, sizeof sig) == 0; };
  if (g_exec && same(g_sig)) return true;
  if (!same(g_seen)) {
    g_seen.assign(sb, sb + sizeof sig);
    return false;
  }
  SCA_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  try {
    launch_residual(g_in, g_out);
  } catch (...) {
    hipGraph_t broken = nullptr;
    (void)hipStreamEndCapture(st, &broken);
    if (broken) (void)hipGraphDestroy(broken);
    throw;
  }
  hipGraph_t graph = nullptr;
  SCA_HIP(hipStreamEndCapture(st, &graph));
  if (g_exec) {
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult res;
    if (hipGraphExecUpdate(g_exec, graph, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
      if (g_graph) (void)hipGraphDestroy(g_graph);
      g_graph = graph;
      g_sig.assign(sb, sb + sizeof sig);
      stats_graph_updates++;
      return true;
    }
    (void)hipGetLastError();  // a refused update leaves no error behind
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    SCA_HIP(e);
  }
  release_residual_graph();
  g_graph = graph;
  g_exec = exec;
  g_sig.assign(sb, sb + sizeof sig);
  stats_graph_captures++;
  return true;
}

void Rve::release_residual_graph() {
  if (g_exec) (void)hipGraphExecDestroy(g_exec);
  if (g_graph) (void)hipGraphDestroy(g_graph);
  g_exec = nullptr;
  g_graph = nullptr;
  g_sig.clear();
}

// MechSolverBasic::formResidual (mech_solver_basic.cpp:286-339).  Small grids
// are bound by launch latency: there the evaluation is one graph launch
// between copies of F in and r out (not while kernels are being timed).
//
// reuse_response (mixed precision, first evaluation of the FP64 finish): the
// material point was last updated, in FP64, at exactly this F (the point the
// FP32 stage returned was its last evaluation), so P and the material-point
// state are those of F already; the evaluation runs the FP64 spectral stages
// on them and counts no material-point update.
void Rve::form_residual(const double* Fdev, void* rdev) {
  if (nfuncs == 0 && snes_iter == 0) total_iter = iter_base - 1;
  if (total_iter - iter_base <= snes_iter) total_iter += 1;
  const bool reused = reuse_response;
  reuse_response = false;
  if (reused) {
    launch_spectral(rdev);
  } else if (batch) {
    SCA_HIP(hipStreamSynchronize(st));  // F and everything before it complete
    state_stale = false;                 // the joint update writes the current set
    batch->evaluate(this, Fdev, rdev);   // the response scalars land in h_scal
  } else if (g_in && !kt.on && ensure_residual_graph()) {
    SCA_HIP(hipMemcpyAsync(g_in, Fdev, 9 * N * sizeof(double), hipMemcpyDeviceToDevice, st));
    SCA_HIP(hipGraphLaunch(g_exec, st));
    SCA_HIP(hipMemcpyAsync(rdev, g_out, 9 * N * (use32 ? sizeof(float) : sizeof(double)), hipMemcpyDeviceToDevice,
                           st));
  } else {
    launch_residual(Fdev, rdev);
  }
  f_eval = Fdev;
  last_dt = delta_t;
  // P_av, err_div, F_aim and err_BC of this evaluation are completed at the
  // next synchronisation (complete_response_scalars)
  scal_pending = true;
  scal_rotate = true;
  scal_residual = true;
  scal_reused = reused;
  if (use32) {
    evals32++;
    stats.fp32_evaluations++;
  } else {
    evals64++;
  }
  nfuncs++;
  solve_evals++;
  stats.residual_evaluations++;
  if (!reused) {
    stats.material_point_updates += N;
    solve_updates++;
  }
}

// MechSolverBasic::converged, mech_solver_basic.cpp:242-284
// The FP32 stage of the mixed form stops at a divergence tolerance no
// tighter than MIXED_FP32_REL_DIV_MIN (its spectra carry FP32 rounding); the
// FP64 solve that follows holds the configured one.
int Rve::converged() const {
  const auto& gn = desc.grid_num;
  double maxP = -std::numeric_limits<double>::infinity();
  for (int q = 0; q < 9; ++q) maxP = std::max(maxP, P_av[q]);
  const double eps_div = use32 ? std::max(gn.eps_rel_div, MIXED_FP32_REL_DIV_MIN) : gn.eps_rel_div;
  const double divTol = std::max(maxP * eps_div, gn.eps_abs_div);
  const double BCTol = std::max(maxP * gn.eps_rel_P, gn.eps_abs_P);
  if ((total_iter >= gn.itmin && std::max(err_div / divTol, err_BC / BCTol) < 1.0) || terminally_ill) return 1;
  if (total_iter >= gn.itmax) return -1;
  return 0;
}

// ---------------------------------------------------------------------------
// SNESNGMRES on device vectors of length n
// ---------------------------------------------------------------------------
Ngmres::Ngmres(int64_t n_, hipStream_t st_, KTimer* kt_) : n(n_), st(st_), kt(kt_) {
  auto a = [&](double*& p, size_t m) {
    p = static_cast<double*>(dev_alloc(std::max<size_t>(m, 1) * sizeof(double)));
    bufs.push_back(p);
  };
  a(partial, (size_t)red_blocks() * MAX_RED_VEC);
  h_out = static_cast<double*>(host_alloc(MAX_RED_VEC * sizeof(double)));
  wait = [this] { SCA_HIP(hipStreamSynchronize(st)); };
}
Ngmres::~Ngmres() {
  for (double* p : bufs) dev_free(p);  // the owner synchronised the stream
  for (double* p : pool) dev_free(p);
  host_free(h_out);
}
double* Ngmres::take() {
  auto used = [&](const double* p) {
    if (p == Fv || p == Xc || p == XM || p == FM || p == XA || p == FA) return true;
    for (const double* q : Fh)
      if (q == p) return true;
    for (const double* q : Xh)
      if (q == p) return true;
    return false;
  };
  for (double* p : pool)
    if (!used(p)) return p;
  double* p = static_cast<double*>(dev_alloc(std::max<int64_t>(n, 1) * sizeof(double)));
  pool.push_back(p);
  return p;
}
template <typename T>
std::vector<double> Ngmres::reduce(const std::vector<int>& ops, const std::vector<const T*>& a,
                                   const std::vector<const T*>& b) {
  const int nvec = (int)ops.size();
  std::vector<double> out(nvec);
  for (int v0 = 0; v0 < nvec; v0 += MAX_RED_VEC) {
    const int nv = std::min(MAX_RED_VEC, nvec - v0);
    if (kt) kt->begin(SC_KT_REDUCE, st);
    launch_partial_reduce<T>(ops.data() + v0, n, nv, a.data() + v0, b.empty() ? nullptr : b.data() + v0, partial,
                             st);
    launch_final_reduce(partial, nv, red_blocks(), h_out, st);  // into pinned host memory
    if (kt) kt->end(SC_KT_REDUCE, n * nv, st);
    wait();
    std::memcpy(out.data() + v0, h_out, nv * sizeof(double));
  }
  if (after_sync) after_sync();
  if (kt) kt->flush();
  return out;
}

std::vector<double> Ngmres::collect(int nout, int64_t units) {
  std::vector<double> out(nout);
  launch_final_reduce(partial, nout, red_blocks(), h_out, st);  // into pinned host memory
  if (kt) kt->end(SC_KT_REDUCE, units, st);
  wait();
  std::memcpy(out.data(), h_out, nout * sizeof(double));
  if (after_sync) after_sync();
  if (kt) kt->flush();
  return out;
}

template <typename TF>
int Ngmres::solve(double* X, int msize, const std::function<void(const double*, TF*)>& residual,
                  const std::function<int(int)>& conv, const std::function<void(int)>& set_iter) {
  const int max_its = 50;
  const double gammaA = 2.0, gammaC = 2.0, epsilonB = 0.1, deltaB = 0.9;
  const int restart_it = 2;
  // residual buffers hold TF values
  auto tf = [](const double* p) { return reinterpret_cast<const TF*>(p); };
  auto Fdot = [&](int i) { return Fh[i]; };
  auto Xdot = [&](int i) { return Xh[i]; };
  std::vector<double> fnorms(msize, 0.0), Q(msize * msize, 0.0);
  Fv = XM = FM = XA = FA = nullptr;
  Fh.assign(msize, nullptr);
  Xh.assign(msize, nullptr);
  Xc = X;
  // the solution into the caller's X; every reference released
  // every evaluation goes through here: the point of the last one is kept
  const double* last_x = nullptr;
  auto eval = [&](const double* x, double* f) {
    last_x = x;
    residual(x, reinterpret_cast<TF*>(f));
  };
  auto done = [&](int r) {
    solution_evaluated_last = Xc == last_x;
    if (Xc != X) SCA_HIP(hipMemcpyAsync(X, Xc, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    Fv = Xc = XM = FM = XA = FA = nullptr;
    Fh.assign(msize, nullptr);
    Xh.assign(msize, nullptr);
    return r;
  };
  Fv = take();
  set_iter(0);
  eval(Xc, Fv);
  double fnorm = std::sqrt(reduce<TF>({0}, {tf(Fv)}, {tf(Fv)})[0]);
  if (std::isnan(fnorm) || std::isinf(fnorm)) return done(-4);  // SNES_DIVERGED_FNORM_NAN
  double fminnorm = fnorm;
  int reason = conv(0);
  if (reason) return done(reason > 0 ? 6 : -5);
  // SNESNGMRESUpdateSubspace_Private: the slot references the vectors
  auto update_subspace = [&](int iv, double* f, double fn, double* x) {
    Fh[iv] = f;
    Xh[iv] = x;
    fnorms[iv] = fn;
  };
  update_subspace(0, Fv, fnorm, Xc);
  int k_restart = 1, l = 1, ivec = 0, restart_count = 0;
  for (int k = 1; k < max_its + 1; ++k) {
    XM = nullptr;
    XM = take();
    if (kt) kt->begin(SC_KT_REDUCE, st);
    launch_axpby(1.0, Xc, -1.0, tf(Fv), XM, n, st);  // basic line search: XM = X - F
    if (kt) kt->end(SC_KT_REDUCE, n, st);
    FM = nullptr;
    FM = take();
    eval(XM, FM);
    // ||FM||^2, <FM, Fdot_i>, <Fdot_ivec, Fdot_i>: one pass over FM,
    // Fdot_ivec and the history (each vector read once)
    std::vector<double> dots(1 + 2 * l);
    if (2 * (l + 1) <= MAX_RED_VEC) {
      std::vector<const TF*> vs(l + 1);
      for (int i = 0; i < l; ++i) vs[i] = tf(Fdot(i));
      vs[l] = tf(FM);
      if (kt) kt->begin(SC_KT_REDUCE, st);
      launch_shared_reduce<TF>(0, n, tf(FM), tf(Fdot(ivec)), l + 1, vs.data(), partial, st);
      const std::vector<double> r = collect(2 * (l + 1), n * (l + 3));
      dots[0] = r[2 * l];
      for (int i = 0; i < l; ++i) {
        dots[1 + i] = r[2 * i];
        dots[1 + l + i] = r[2 * i + 1];
      }
    } else {
      std::vector<int> ops(1 + 2 * l, 0);
      std::vector<const TF*> va(1 + 2 * l), vb(1 + 2 * l);
      va[0] = tf(FM);
      vb[0] = tf(FM);
      for (int i = 0; i < l; ++i) {
        va[1 + i] = tf(FM);
        vb[1 + i] = tf(Fdot(i));
        va[1 + l + i] = tf(Fdot(ivec));
        vb[1 + l + i] = tf(Fdot(i));
      }
      dots = reduce<TF>(ops, va, vb);
    }
    const double fMnorm = std::sqrt(dots[0]);
    bool fused = false;
    if (std::isnan(fMnorm) || std::isinf(fMnorm)) return done(-6);  // SNES_DIVERGED_LINE_SEARCH
    {  // SNESNGMRESFormCombinedSolution_Private
      const double nu = fMnorm * fMnorm;
      std::vector<double> xi(l), beta(l);
      for (int i = 0; i < l; ++i) {
        xi[i] = dots[1 + i];
        beta[i] = dots[1 + l + i];
        Q[i * msize + ivec] = beta[i];
        Q[ivec * msize + i] = beta[i];
      }
      for (int i = 0; i < l; ++i) beta[i] = nu - xi[i];
      std::vector<double> H(l * l);
      for (int j = 0; j < l; ++j)
        for (int i = 0; i < l; ++i) H[i * l + j] = Q[i * msize + j] - xi[i] - xi[j] + nu;
      if (l == 1)
        beta[0] = (H[0] != 0.0) ? beta[0] / H[0] : 0.0;
      else
        sym_lstsq(l, H, beta);
      double alph_total = 0.0;
      for (int i = 0; i < l; ++i) alph_total += beta[i];
      std::vector<const double*> xs(l);
      for (int i = 0; i < l; ++i) xs[i] = Xdot(i);
      XA = nullptr;
      XA = take();
      // the distances of XA to XM and to the history (SNESNGMRESNorms_Private)
      // come with the combination when the history fits one pass
      if (kt) kt->begin(SC_KT_REDUCE, st);
      fused = launch_lincomb_norms(XM, 1.0 - alph_total, l, beta.data(), xs.data(), XA, n, partial + red_blocks(),
                                   st);
      if (!fused) launch_lincomb(XM, 1.0 - alph_total, l, beta.data(), xs.data(), XA, n, st);
      if (kt) kt->end(SC_KT_REDUCE, n * (l + 1), st);
      FA = nullptr;
      FA = take();
      eval(XA, FA);
    }
    if (fminnorm > fMnorm) fminnorm = fMnorm;
    // ||FA||^2, ||XA - XM||^2, ||Xdot_i - XA||^2 (SNESNGMRESNorms_Private):
    // FA alone, then XM and the history against XA in one pass
    std::vector<double> nrm;
    if (fused) {
      if (kt) kt->begin(SC_KT_REDUCE, st);
      const int op0 = 0;
      const TF* fa = tf(FA);
      launch_partial_reduce<TF>(&op0, n, 1, &fa, &fa, partial, st);
      nrm = collect(2 + l, n);
    } else if (2 + l <= MAX_RED_VEC) {
      if (kt) kt->begin(SC_KT_REDUCE, st);
      const int op0 = 0;
      const TF* fa = tf(FA);
      launch_partial_reduce<TF>(&op0, n, 1, &fa, &fa, partial, st);
      std::vector<const double*> vs(l + 1);
      vs[0] = XM;
      for (int i = 0; i < l; ++i) vs[1 + i] = Xdot(i);
      launch_shared_reduce<double>(1, n, XA, nullptr, l + 1, vs.data(), partial + red_blocks(), st);
      nrm = collect(2 + l, n * (l + 3));
    } else {
      // ||FA||^2 apart (another element type), then the distances; each
      // reduction is that vector's own, whatever it is launched with
      nrm = reduce<TF>({0}, {tf(FA)}, {tf(FA)});
      std::vector<int> ops2(1 + l, 1);
      std::vector<const double*> wa(1 + l), wb(1 + l);
      wa[0] = XA;
      wb[0] = XM;
      for (int i = 0; i < l; ++i) {
        wa[1 + i] = Xdot(i);
        wb[1 + i] = XA;
      }
      const std::vector<double> d = reduce<double>(ops2, wa, wb);
      nrm.insert(nrm.end(), d.begin(), d.end());
    }
    const double fAnorm = std::sqrt(nrm[0]);
    const double dnorm = std::sqrt(nrm[1]);
    double dminnorm = -1.0;
    for (int i = 0; i < l; ++i) {
      const double d = std::sqrt(nrm[2 + i]);
      if (d < dminnorm || dminnorm < 0.0) dminnorm = d;
    }
    if (std::isnan(fnorm) || std::isinf(fnorm)) return done(-4);
    // SNESNGMRESSelect_Private / SelectRestart_Private (difference)
    bool selectA = fAnorm < gammaA * fminnorm;
    if (!(epsilonB * dnorm < dminnorm || std::sqrt(fnorm) < deltaB * std::sqrt(fminnorm))) selectA = false;
    bool selectRestart = false;
    if (epsilonB * dnorm > dminnorm && fAnorm > deltaB * fminnorm && l > 0) selectRestart = true;
    if (fAnorm > gammaC * fminnorm) selectRestart = true;
    restart_count = selectRestart ? restart_count + 1 : 0;
    ivec = k_restart % msize;
    const bool do_restart = restart_count >= restart_it;
    if (do_restart) {
      update_subspace(0, FM, fMnorm, XM);
      for (int i = 1; i < msize; ++i) Fh[i] = Xh[i] = nullptr;
    }
    if (selectA) {
      fnorm = fAnorm;
      Fv = FA;
      Xc = XA;
    } else {
      fnorm = fMnorm;
      Fv = FM;
      Xc = XM;
    }
    if (do_restart) {
      restart_count = 0;
      k_restart = 1;
      l = 1;
      ivec = 0;
    } else {
      if (l < msize) l++;
      k_restart++;
      if (fminnorm > fnorm) fminnorm = fnorm;
      update_subspace(ivec, Fv, fnorm, Xc);
    }
    set_iter(k);
    reason = conv(k);
    if (reason) return done(reason > 0 ? 6 : -5);
  }
  return done(-5);  // SNES_DIVERGED_MAX_IT
}

template int Ngmres::solve<double>(double*, int, const std::function<void(const double*, double*)>&,
                                   const std::function<int(int)>&, const std::function<void(int)>&);
template int Ngmres::solve<float>(double*, int, const std::function<void(const double*, float*)>&,
                                  const std::function<int(int)>&, const std::function<void(int)>&);

int Rve::snes_ngmres() {
  const int msize = desc.grid_num.ngmres_m > 0 ? desc.grid_num.ngmres_m : 30;
  snes_iter = 0;
  nfuncs = 0;
  if (use32)
    return ngm->solve<float>(
        X, msize, [this](const double* x, float* f) { form_residual(x, f); }, [this](int) { return converged(); },
        [this](int k) { snes_iter = k; });
  return ngm->solve<double>(
      X, msize, [this](const double* x, double* f) { form_residual(x, f); }, [this](int) { return converged(); },
      [this](int k) { snes_iter = k; });
}

// MechSolverBasic::update_coords -> base_update_coords (mech_base.cpp:45-154):
// cell centres and nodes of the deformed configuration from the average F
// and the periodic displacement fluctuation u = FFT^-1(F_hat xi / |xi|^2)
void Rve::update_coords() {
  fft.forward(X, spec, 9, st);
  launch_coords_uhat(spec, K, gamma_host(false), d_xi_tab, d_fsum, st);
  fft.backward(spec, d_u, 3, wgt, st);
  launch_coords_nodes(d_u, d_fsum, cells, size, d_xn, st);
  launch_coords_points(d_u, d_fsum, cells, size, d_xp, st);
}

// Checkpoint layout: a header (magic, format version, N, cells, precision
// and a hash of the description's material map, orientations, material and
// numerics), the scalar state of the solver, then the device planes in a
// fixed order.
namespace {
constexpr char CKPT_MAGIC[8] = {'S', 'C', 'A', 'M', 'D', 'C', 'K', '1'};
constexpr int32_t CKPT_VERSION = 2;

// FNV-1a over the fields of a description that a checkpoint's state depends
// on (field by field: the structs have padding)
uint64_t desc_hash(const sc_rve_desc& d, int64_t N) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  };
  auto f64 = [&](double v) { mix(&v, sizeof v); };
  auto i32 = [&](int v) { mix(&v, sizeof v); };
  mix(d.material_at, (size_t)N * sizeof(int));
  i32(d.n_materials);
  mix(d.orientations, (size_t)4 * d.n_materials * sizeof(double));
  const sc_material& m = d.material;
  for (double v : {m.C_11, m.C_12, m.C_44, m.xi_0_sl, m.xi_inf_sl, m.h_0_sl_sl, m.n_sl, m.a_sl, m.dot_gamma_0_sl,
                   m.atol_xi, m.atol_gamma})
    f64(v);
  for (double v : m.h_sl_sl) f64(v);
  const sc_phase_numerics& p = d.phase_num;
  for (double v : {p.r_cutback_min, p.r_cutback, p.r_increase, p.eps_rel_state, p.r_linesearch_Lp, p.eps_rel_Lp,
                   p.eps_abs_Lp})
    f64(v);
  i32(p.N_iter_state_max);
  i32(p.N_iter_Lp_max);
  const sc_grid_numerics& g = d.grid_num;
  for (int v : {g.itmin, g.itmax, g.update_gamma, g.divergence_correction, g.derivative, g.stag_iter_max,
                g.max_cut_back, g.ngmres_m})
    i32(v);
  for (double v : {g.eps_abs_div, g.eps_rel_div, g.eps_abs_P, g.eps_rel_P}) f64(v);
  for (int q = 0; q < 3; ++q) f64(d.size[q]);
  return h;
}
struct CkptScalars {
  double C_ref[81], C_volAvg[81], C_minMaxAvg[81], C_volAvgLastInc[81], C_minMaxAvgLastInc[81], S[81];
  double P_av[9], P_aim[9], F_aim_dot[9], F_aim[9], F_aim_last_inc[9];
  int stress_mask[9];
  double rot_q[4];
  double delta_t, last_dt, last_div_sum, err_BC, err_div;
  int total_iter, terminally_ill, snes_iter, nfuncs, solve_evals, use32, iter_base;
  int64_t evals32, evals64;
  sc_stats stats;
};
struct CkptHeader {
  char magic[8];
  int32_t version;
  int32_t precision;
  int64_t N;
  int32_t cells[3];
  int32_t pad;
  uint64_t desc_hash;
};
}  // namespace

void Rve::save(std::vector<char>& blob) {
  complete_response_scalars();
  sync_response_F();
  refresh_state();
  const std::vector<std::pair<double*, int64_t>> planes = {
      {mp.F, 9 * N},       {mp.F0, 9 * N},    {mp.Fp0, 9 * N},   {mp.Lp0, 9 * N},   {mp.S0, 9 * N},
      {mp.xi0, 12 * N},    {mp.gamma0, 12 * N}, {mp.Lp, 9 * N},  {mp.P, 9 * N},     {mp.Fp, 9 * N},
      {mp.S, 9 * N},       {mp.Fe, 9 * N},    {mp.xi, 12 * N},   {mp.gamma, 12 * N}, {X, 9 * N},
      {F_last_inc, 9 * N}, {F_dot, 9 * N},    {d_u, 3 * N},      {d_xp, 3 * N},
      {d_xn, 3 * (int64_t)(cells[0] + 1) * (cells[1] + 1) * (cells[2] + 1)}};
  int64_t doubles = 0;
  for (auto& pl : planes) doubles += pl.second;
  CkptHeader h{};
  std::memcpy(h.magic, CKPT_MAGIC, 8);
  h.version = CKPT_VERSION;
  h.N = N;
  for (int p = 0; p < 3; ++p) h.cells[p] = cells[p];
  h.precision = desc.precision;
  h.desc_hash = desc_hash(desc, N);
  CkptScalars sc{};
  std::memcpy(sc.C_ref, C_ref, sizeof C_ref);
  std::memcpy(sc.C_volAvg, C_volAvg, sizeof C_volAvg);
  std::memcpy(sc.C_minMaxAvg, C_minMaxAvg, sizeof C_minMaxAvg);
  std::memcpy(sc.C_volAvgLastInc, C_volAvgLastInc, sizeof C_volAvgLastInc);
  std::memcpy(sc.C_minMaxAvgLastInc, C_minMaxAvgLastInc, sizeof C_minMaxAvgLastInc);
  std::memcpy(sc.S, S, sizeof S);
  std::memcpy(sc.P_av, P_av, sizeof P_av);
  std::memcpy(sc.P_aim, P_aim, sizeof P_aim);
  std::memcpy(sc.F_aim_dot, F_aim_dot, sizeof F_aim_dot);
  std::memcpy(sc.F_aim, F_aim, sizeof F_aim);
  std::memcpy(sc.F_aim_last_inc, F_aim_last_inc, sizeof F_aim_last_inc);
  std::memcpy(sc.stress_mask, stress_mask, sizeof stress_mask);
  std::memcpy(sc.rot_q, rot_q, sizeof rot_q);
  sc.delta_t = delta_t;
  sc.last_dt = last_dt;
  sc.last_div_sum = last_div_sum;
  sc.err_BC = err_BC;
  sc.err_div = err_div;
  sc.total_iter = total_iter;
  sc.terminally_ill = terminally_ill ? 1 : 0;
  sc.snes_iter = snes_iter;
  sc.nfuncs = nfuncs;
  sc.solve_evals = solve_evals;
  sc.use32 = use32 ? 1 : 0;
  sc.iter_base = iter_base;
  sc.evals32 = evals32;
  sc.evals64 = evals64;
  sc.stats = stats;
  blob.resize(sizeof h + sizeof sc + (size_t)doubles * sizeof(double));
  char* w = blob.data();
  std::memcpy(w, &h, sizeof h);
  std::memcpy(w + sizeof h, &sc, sizeof sc);
  size_t off = sizeof h + sizeof sc;
  for (auto& pl : planes) {
    SCA_HIP(hipMemcpyAsync(w + off, pl.first, pl.second * sizeof(double), hipMemcpyDeviceToHost, st));
    off += pl.second * sizeof(double);
  }
  SCA_HIP(hipStreamSynchronize(st));
}

void Rve::load(const std::vector<char>& blob) {
  CkptHeader h{};
  CkptScalars sc{};
  if (blob.size() < sizeof h + sizeof sc) throw std::runtime_error("checkpoint too short");
  std::memcpy(&h, blob.data(), sizeof h);
  if (std::memcmp(h.magic, CKPT_MAGIC, 8) != 0) throw std::runtime_error("not a solver checkpoint");
  if (h.version != CKPT_VERSION) throw std::runtime_error("checkpoint of another format version");
  if (h.N != N || h.cells[0] != cells[0] || h.cells[1] != cells[1] || h.cells[2] != cells[2] ||
      h.precision != desc.precision)
    throw std::runtime_error("checkpoint of another grid or precision");
  if (h.desc_hash != desc_hash(desc, N))
    throw std::runtime_error("checkpoint of another microstructure, material or numerics");
  // nothing of this context's last evaluation may land on the restored state
  complete_response_scalars();
  release_residual_graph();
  g_seen.clear();
  state_stale = false;  // both sets of the state are restored
  const std::vector<std::pair<double*, int64_t>> planes = {
      {mp.F, 9 * N},       {mp.F0, 9 * N},    {mp.Fp0, 9 * N},   {mp.Lp0, 9 * N},   {mp.S0, 9 * N},
      {mp.xi0, 12 * N},    {mp.gamma0, 12 * N}, {mp.Lp, 9 * N},  {mp.P, 9 * N},     {mp.Fp, 9 * N},
      {mp.S, 9 * N},       {mp.Fe, 9 * N},    {mp.xi, 12 * N},   {mp.gamma, 12 * N}, {X, 9 * N},
      {F_last_inc, 9 * N}, {F_dot, 9 * N},    {d_u, 3 * N},      {d_xp, 3 * N},
      {d_xn, 3 * (int64_t)(cells[0] + 1) * (cells[1] + 1) * (cells[2] + 1)}};
  int64_t doubles = 0;
  for (auto& pl : planes) doubles += pl.second;
  if (blob.size() != sizeof h + sizeof sc + (size_t)doubles * sizeof(double))
    throw std::runtime_error("checkpoint size does not match the grid");
  std::memcpy(&sc, blob.data() + sizeof h, sizeof sc);
  std::memcpy(C_ref, sc.C_ref, sizeof C_ref);
  std::memcpy(C_volAvg, sc.C_volAvg, sizeof C_volAvg);
  std::memcpy(C_minMaxAvg, sc.C_minMaxAvg, sizeof C_minMaxAvg);
  std::memcpy(C_volAvgLastInc, sc.C_volAvgLastInc, sizeof C_volAvgLastInc);
  std::memcpy(C_minMaxAvgLastInc, sc.C_minMaxAvgLastInc, sizeof C_minMaxAvgLastInc);
  std::memcpy(S, sc.S, sizeof S);
  std::memcpy(P_av, sc.P_av, sizeof P_av);
  std::memcpy(P_aim, sc.P_aim, sizeof P_aim);
  std::memcpy(F_aim_dot, sc.F_aim_dot, sizeof F_aim_dot);
  std::memcpy(F_aim, sc.F_aim, sizeof F_aim);
  std::memcpy(F_aim_last_inc, sc.F_aim_last_inc, sizeof F_aim_last_inc);
  std::memcpy(stress_mask, sc.stress_mask, sizeof stress_mask);
  std::memcpy(rot_q, sc.rot_q, sizeof rot_q);
  delta_t = sc.delta_t;
  last_dt = sc.last_dt;
  last_div_sum = sc.last_div_sum;
  err_BC = sc.err_BC;
  err_div = sc.err_div;
  total_iter = sc.total_iter;
  terminally_ill = sc.terminally_ill != 0;
  snes_iter = sc.snes_iter;
  nfuncs = sc.nfuncs;
  solve_evals = sc.solve_evals;
  use32 = sc.use32 != 0;
  iter_base = sc.iter_base;
  evals32 = sc.evals32;
  evals64 = sc.evals64;
  stats = sc.stats;
  size_t off = sizeof h + sizeof sc;
  for (auto& pl : planes) {
    SCA_HIP(hipMemcpyAsync(pl.first, blob.data() + off, pl.second * sizeof(double), hipMemcpyHostToDevice, st));
    off += pl.second * sizeof(double);
  }
  f_eval = nullptr;  // mp.F holds the restored F
  scal_pending = false;
  x_uniform = last_uniform = rate_uniform = false;
  SCA_HIP(hipStreamSynchronize(st));
}

void Rve::init() {
  std::vector<double> eye(9 * N);
  for (int q = 0; q < 9; ++q)
    for (int64_t c = 0; c < N; ++c) eye[q * N + c] = (q % 4 == 0) ? 1.0 : 0.0;
  SCA_HIP(hipMemcpyAsync(F_last_inc, eye.data(), 9 * N * sizeof(double), hipMemcpyHostToDevice, st));
  SCA_HIP(hipMemcpyAsync(X, eye.data(), 9 * N * sizeof(double), hipMemcpyHostToDevice, st));
  SCA_HIP(hipStreamSynchronize(st));
  x_uniform = last_uniform = true;
  rate_uniform = false;
  for (int q = 0; q < 9; ++q) x_value[q] = last_value[q] = (q % 4 == 0) ? 1.0 : 0.0;
  update_coords();
  constitutive_response(X, 0.0);
  fetch_response_scalars(false);
  tangent();
  for (int q = 0; q < 81; ++q) C_ref[q] = C_minMaxAvg[q] / wgt;
}

void Rve::forward(bool cutBack, bool guess, double dt, double dt_old, double t_remaining, const sc_loadstep& stp) {
  if (cutBack) {
    std::memcpy(C_volAvg, C_volAvgLastInc, sizeof C_volAvg);
    std::memcpy(C_minMaxAvg, C_minMaxAvgLastInc, sizeof C_minMaxAvg);
  } else {
    std::memcpy(C_volAvgLastInc, C_volAvg, sizeof C_volAvg);
    std::memcpy(C_minMaxAvgLastInc, C_minMaxAvg, sizeof C_minMaxAvg);
    double aim_diff[9];
    for (int q = 0; q < 9; ++q) aim_diff[q] = (F_aim[q] - F_aim_last_inc[q]) / dt_old;
    for (int q = 0; q < 9; ++q) F_aim_dot[q] = guess ? (stp.stress_mask[q] ? 0.0 : aim_diff[q]) : 0.0;
    std::memcpy(F_aim_last_inc, F_aim, sizeof F_aim);
    double step[9];
    for (int q = 0; q < 9; ++q)
      step[q] = (stp.deformation_type == SC_DEF_DOT_F) ? stp.deformation[q]
                                                        : (stp.deformation[q] - F_aim_last_inc[q]) / t_remaining;
    for (int q = 0; q < 9; ++q) F_aim_dot[q] += stp.deformation_mask[q] ? 0.0 : step[q];
    double rotated[9];
    h::rot2_active(stp.rot_q, F_aim_dot, rotated);  // load -> field frame
    SCA_HIP(hipMemcpyAsync(small + SM_AV, rotated, 9 * sizeof(double), hipMemcpyHostToDevice, st));
    launch_rate(guess ? 1 : 0, F_last_inc, X, dt_old, small + SM_AV, F_dot, N, st);
    // F_last_inc <- X by exchanging the buffers: X is written whole below
    // (launch_fwd_field, from F_last_inc), so its old contents are not needed
    std::swap(F_last_inc, X);
    rate_uniform = !guess;
    std::memcpy(rate_value, rotated, sizeof rate_value);
    last_uniform = x_uniform;
    std::memcpy(last_value, x_value, sizeof last_value);
    sync_response_F();
    // f_materialpoint_forward.  Fp, S, xi and gamma become the start values
    // by exchanging the buffers' roles (the next evaluation writes the
    // current set before anything reads it; refresh_state() copies it back
    // for a reader that comes first); F and Lp stay in use and are copied.
    // A captured residual graph holds the buffer addresses: with it the
    // whole state is copied.
    if (g_in) {
      launch_mp_forward(mp, st);
    } else {
      launch_mp_copy(mp, 1, st);
      std::swap(mp.Fp, mp.Fp0);
      std::swap(mp.S, mp.S0);
      std::swap(mp.xi, mp.xi0);
      std::swap(mp.gamma, mp.gamma0);
      state_stale = true;
    }
  }
  for (int q = 0; q < 9; ++q) F_aim[q] = F_aim_last_inc[q] + F_aim_dot[q] * dt;
  for (int q = 0; q < 9; ++q)
    if (stp.stress_mask[q]) {
      if (stp.stress_type == SC_STRESS_P)
        P_aim[q] += (stp.stress[q] - P_aim[q]) / t_remaining * dt;
      else
        P_aim[q] += stp.stress[q] * dt;
    }
  double rotated[9];
  h::rot2_active(stp.rot_q, F_aim, rotated);  // load -> field frame
  SCA_HIP(hipMemcpyAsync(small + SM_AIM2, rotated, 9 * sizeof(double), hipMemcpyHostToDevice, st));
  launch_fwd_field(F_last_inc, F_dot, dt, X, 9 * N, st);
  double diff[9];  // read by the copy below until the synchronisation
  if (last_uniform && rate_uniform) {
    // Every cell holds the same value v.  The reference averages the field
    // with one add chain per component over the cells in storage order
    // (mech_base.cpp:222, an Eigen sum over the cell dimensions); for N equal
    // terms that chain's rounding errors all have one sign and its average
    // is off by up to N/2 ulp (3.9e-12 for v = 1.0005 at 64^3).  The offset
    // stays in the average of F through the whole increment and moves <P>
    // of a first increment by about 1e-8 of itself, so here the sum is
    // formed as that chain does, on the host (N adds per component, only
    // for a homogeneous field: the first increment of a job).  Other
    // fields are summed by the device's blocked reduction, whose error is
    // of the order of log2(N) ulp.
    double v[9], sum[9] = {};
    for (int q = 0; q < 9; ++q) v[q] = last_value[q] + rate_value[q] * dt;
    for (int64_t c = 0; c < N; ++c)
      for (int q = 0; q < 9; ++q) sum[q] += v[q];
    for (int q = 0; q < 9; ++q) diff[q] = sum[q] * wgt - rotated[q];
    SCA_HIP(hipMemcpyAsync(small + SM_FDIFF, diff, sizeof diff, hipMemcpyHostToDevice, st));
  } else {
    std::vector<int> ops(9, 2);
    std::vector<const double*> a(9);
    for (int q = 0; q < 9; ++q) a[q] = X + q * N;
    launch_partial_reduce<double>(ops.data(), N, 9, a.data(), nullptr, partial, st);
    launch_final_reduce(partial, 9, red_blocks(), small + SM_FSUM, st);
    launch_fwd_diff(small + SM_FSUM, wgt, small + SM_AIM2, small + SM_FDIFF, st);
  }
  launch_sub_plane_const(X, small + SM_FDIFF, N, st);
  SCA_HIP(hipStreamSynchronize(st));
  x_uniform = false;  // X now holds the forwarded field, changed by the solve
  for (int q = 0; q < 9; ++q) stress_mask[q] = stp.stress_mask[q];
  for (int q = 0; q < 4; ++q) rot_q[q] = stp.rot_q[q];
  delta_t = dt;
}

sc_solution Rve::calculate_solution() {
  masked_compliance_host(C_volAvg, rot_q, stress_mask, S);
  if (desc.grid_num.update_gamma)
    for (int q = 0; q < 81; ++q) C_ref[q] = C_minMaxAvg[q] / wgt;
  solve_evals = 0;
  solve_updates = 0;
  iter_base = 0;
  int reason;
  if (mixed) {
    // FP32 spectral stages until the criteria hold (the divergence at no
    // tighter than MIXED_FP32_REL_DIV_MIN), then an FP64 solve continued from
    // that F (its iterations count on from the first solve's), which accepts
    // at its first evaluation when the FP64 residual already meets the
    // tolerances.  The FP64 solve also takes over when the FP32 stage stopped
    // without converging (iteration limit, a non-finite norm): only a
    // terminally ill point ends the increment there.  When the FP32 stage's
    // last evaluation was at the F it returned, the FP64 solve's first
    // evaluation reuses that FP64 material-point update.
    use32 = true;
    reason = snes_ngmres();
    use32 = false;
    if (!terminally_ill) {
      iter_base = total_iter + 1;
      reuse_response = ngm->solution_evaluated_last;
      reason = snes_ngmres();
      reuse_response = false;
    }
    iter_base = 0;
  } else {
    reason = snes_ngmres();
  }
  tangent();
  sc_solution out{};
  out.converged = reason > 0;
  out.stag_converged = 1;
  out.required_iterations = total_iter;
  out.terminally_ill = terminally_ill;
  out.residual_evaluations = solve_evals;
  out.err_div = err_div;
  out.err_BC = err_BC;
  terminally_ill = false;
  for (int q = 0; q < 9; ++q)
    if (stress_mask[q]) P_aim[q] = P_av[q];
  return out;
}

// One increment of DamaskGrid::loop_over_loadcase (damask_grid.cpp:82-196)
// at the cursor: time step, forward, solution and cut-backs.  Returns 0 when
// an increment was solved, 1 when the load case is exhausted, 2 when the
// increment could not be solved within max_cut_back.
int Rve::next_increment(const sc_loadstep* steps, int n_steps, LoadCursor& cur, int* iterations, int* evaluations) {
  const auto& gn = desc.grid_num;
  const int sub_step_factor = 2;
  if (cur.l >= n_steps) return 1;
  const sc_loadstep& stp = steps[cur.l];
  if (!cur.step_started) {
    cur.t_0 = cur.t;
    cur.guess = stp.estimate_rate && cur.l > 0;
    cur.step_started = true;
    cur.inc = 0;
  }
  const int inc = cur.inc;
  cur.total_inc++;
  cur.dt_prev = cur.dt;
  if (std::fabs(stp.r - 1.0) < 1.e-9)
    cur.dt = stp.t / stp.N;
  else
    cur.dt = stp.t * (std::pow(stp.r, inc) - std::pow(stp.r, inc + 1)) / (1.0 - std::pow(stp.r, stp.N));
  cur.dt *= std::pow((double)sub_step_factor, -(double)cur.cut_back_level);
  int step_fraction = 0, iters = 0, evals = 0;
  while (step_fraction < std::pow(sub_step_factor, cur.cut_back_level)) {
    const double t_remaining = stp.t + cur.t_0 - cur.t;
    cur.t += cur.dt;
    step_fraction++;
    forward(cur.cut_back, cur.guess, cur.dt, cur.dt_prev, t_remaining, stp);
    const sc_solution sol = calculate_solution();
    iters += sol.required_iterations;
    evals += sol.residual_evaluations;
    if (sol.converged && sol.stag_converged && !sol.terminally_ill) {
      update_coords();
      cur.dt_prev = cur.dt;
      cur.cut_back = false;
      cur.guess = true;
    } else if (cur.cut_back_level < gn.max_cut_back) {
      cur.cut_back = true;
      step_fraction = (step_fraction - 1) * sub_step_factor;
      cur.cut_back_level++;
      cur.t = cur.t - cur.dt;
      cur.dt = cur.dt / sub_step_factor;
      stats.cutbacks++;
    } else {
      g_err = "no solution found for increment " + std::to_string(cur.total_inc);
      return 2;
    }
  }
  cur.cut_back_level = std::max(0, (int)cur.cut_back - 1);  // damask_grid.cpp:153
  stats.increments++;
  if (on_result && stp.f_out > 0 && inc % stp.f_out == 0) on_result(cur.total_inc, cur.t);
  if (iterations) *iterations = iters;
  if (evaluations) *evaluations = evals;
  cur.inc++;
  if (cur.inc >= stp.N) {
    cur.l++;
    cur.step_started = false;
  }
  return 0;
}

int Rve::run_loadcase(const sc_loadstep* steps, int n_steps, int* inc_iterations, int* inc_evaluations,
                      int max_incs) {
  LoadCursor cur;
  for (;;) {
    int iters = 0, evals = 0;
    const int rc = next_increment(steps, n_steps, cur, &iters, &evals);
    if (rc == 1) return 0;
    if (rc != 0) return rc;
    if (inc_iterations && cur.total_inc <= max_incs) inc_iterations[cur.total_inc - 1] = iters;
    if (inc_evaluations && cur.total_inc <= max_incs) inc_evaluations[cur.total_inc - 1] = evals;
  }
}

}  // namespace sca
