// C-ABI of the HIP solver library (include/solver_cpp_amd.h).
//
// Every entry point catches exceptions and reports them through
// sc_last_error(); the compute always runs on the GPU (there is no host
// fallback: without a HIP device every call that computes fails).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "dadf5.h"
#include "result.h"
#include "rve.h"

using sca::g_err;

struct sc_rve {
  std::unique_ptr<sca::Rve> r;
  std::unique_ptr<sca::ResultFile> result;
};

struct sc_dadf5 {
  std::unique_ptr<sca::Dadf5Writer> w;
};

namespace {

using clk = std::chrono::steady_clock;

template <typename Fn>
int guarded(Fn&& fn) {
  try {
    return fn();
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// RAII device buffer on a stream
struct DBuf {
  double* p = nullptr;
  size_t n = 0;
  explicit DBuf(size_t n_) : n(n_) { SCA_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double))); }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  void put(const double* h, hipStream_t st) {
    SCA_HIP(hipMemcpyAsync(p, h, n * sizeof(double), hipMemcpyHostToDevice, st));
  }
  void get(double* h, hipStream_t st) const {
    SCA_HIP(hipMemcpyAsync(h, p, n * sizeof(double), hipMemcpyDeviceToHost, st));
  }
};

struct Stream {
  hipStream_t s = nullptr;
  Stream() { SCA_HIP(hipStreamCreate(&s)); }
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
  void sync() const { SCA_HIP(hipStreamSynchronize(s)); }
};

// ---- host tensor helpers used by the small C-ABI functions ---------------
inline double at(const double* m, int i, int j) { return m[i + 3 * j]; }
void mul33(const double* A, const double* B, double* C) {
  double t[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[i + 3 * j] = at(A, i, 0) * at(B, 0, j) + at(A, i, 1) * at(B, 1, j) + at(A, i, 2) * at(B, 2, j);
  std::memcpy(C, t, sizeof t);
}
void rotation_matrix(const double q[4], int active, double R[9]) {
  sca::qu2om_host(q, R);
  if (active == 1) {
    double t[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) t[i + 3 * j] = R[j + 3 * i];
    std::memcpy(R, t, sizeof t);
  }
}

// Spectral table (3, nr, Nz, Ny) interleaved (the reference's xi1st
// layout) -> kernel table (3 complex per k = i + nr*(y + Ny*z))
std::vector<double> xi_to_kernel_table(const int cells[3], const double* xi) {
  const int nr = cells[0] / 2 + 1, Ny = cells[1], Nz = cells[2];
  const int64_t K = (int64_t)nr * Ny * Nz;
  std::vector<double> tab(6 * K);
  for (int y = 0; y < Ny; ++y)
    for (int z = 0; z < Nz; ++z)
      for (int i = 0; i < nr; ++i) {
        const int64_t src = i + (int64_t)nr * (z + (int64_t)Nz * y);
        const int64_t dst = i + (int64_t)nr * (y + (int64_t)Ny * z);
        for (int p = 0; p < 6; ++p) tab[6 * dst + p] = xi[6 * src + p];
      }
  return tab;
}

// (howmany, N) interleaved host reals <-> howmany device planes
void interleaved_to_planes(const double* in, int howmany, int64_t N, std::vector<double>& planes) {
  planes.resize((size_t)howmany * N);
  for (int64_t n = 0; n < N; ++n)
    for (int c = 0; c < howmany; ++c) planes[c * N + n] = in[howmany * n + c];
}

// kernel spectrum planes [c][z][y][i] <-> (howmany, nr, Nz, Ny) interleaved complex
void spec_planes_to_ref(const int cells[3], int howmany, const std::vector<double>& sp, double* out) {
  const int nr = cells[0] / 2 + 1, Ny = cells[1], Nz = cells[2];
  const int64_t K = (int64_t)nr * Ny * Nz;
  for (int c = 0; c < howmany; ++c)
    for (int z = 0; z < Nz; ++z)
      for (int y = 0; y < Ny; ++y)
        for (int i = 0; i < nr; ++i) {
          const int64_t kd = i + (int64_t)nr * (y + (int64_t)Ny * z);
          const int64_t kr = i + (int64_t)nr * (z + (int64_t)Nz * y);
          out[2 * (howmany * kr + c)] = sp[2 * (c * K + kd)];
          out[2 * (howmany * kr + c) + 1] = sp[2 * (c * K + kd) + 1];
        }
}
void ref_to_spec_planes(const int cells[3], int howmany, const double* in, std::vector<double>& sp) {
  const int nr = cells[0] / 2 + 1, Ny = cells[1], Nz = cells[2];
  const int64_t K = (int64_t)nr * Ny * Nz;
  sp.resize((size_t)2 * howmany * K);
  for (int c = 0; c < howmany; ++c)
    for (int z = 0; z < Nz; ++z)
      for (int y = 0; y < Ny; ++y)
        for (int i = 0; i < nr; ++i) {
          const int64_t kd = i + (int64_t)nr * (y + (int64_t)Ny * z);
          const int64_t kr = i + (int64_t)nr * (z + (int64_t)Nz * y);
          sp[2 * (c * K + kd)] = in[2 * (howmany * kr + c)];
          sp[2 * (c * K + kd) + 1] = in[2 * (howmany * kr + c) + 1];
        }
}

sca::GammaHost gamma_host_table(const int cells[3], const double size[3], const double scaled[3],
                                const double* C_ref, int do_gamma) {
  sca::GammaHost g{};
  for (int p = 0; p < 3; ++p) {
    g.cells[p] = cells[p];
    g.size[p] = size ? size[p] : 1.0;
    g.rescale[p] = (size && scaled) ? size[p] / scaled[p] : 1.0;
  }
  if (C_ref)
    for (int l = 0; l < 3; ++l)
      for (int n = 0; n < 3; ++n)
        for (int m = 0; m < 3; ++m)
          for (int o = 0; o < 3; ++o) g.C_ref[27 * l + 9 * n + 3 * m + o] = C_ref[l + 3 * n + 9 * m + 27 * o];
  g.derivative = 1;  // table mode: the caller's xi1st
  g.do_gamma = do_gamma;
  return g;
}

// wall time of a solver call, accumulated in the context's statistics
template <typename Fn>
int timed(sc_rve* R, Fn&& fn) {
  return guarded([&] {
    const auto t0 = clk::now();
    const int rc = fn();
    SCA_HIP(hipStreamSynchronize(R->r->st));
    R->r->stats.seconds_total += std::chrono::duration<double>(clk::now() - t0).count();
    return rc;
  });
}

}  // namespace

extern "C" {

const char* sc_last_error(void) { return g_err.c_str(); }
const char* sc_backend_name(void) { return "hip-gfx950"; }

// examples/grid/material.yaml:7-24 (Aluminum phenopowerlaw)
void sc_default_material(sc_material* m) {
  m->C_11 = 106.75e9;
  m->C_12 = 60.41e9;
  m->C_44 = 28.34e9;
  m->xi_0_sl = 31.e6;
  m->xi_inf_sl = 63.e6;
  m->h_0_sl_sl = 75.e6;
  m->n_sl = 20;
  m->a_sl = 2.25;
  m->dot_gamma_0_sl = 0.001;
  const double h[7] = {1, 1, 1.4, 1.4, 1.4, 1.4, 1.4};
  for (int i = 0; i < 7; ++i) m->h_sl_sl[i] = h[i];
  m->atol_xi = 1.0;
  m->atol_gamma = 1e-6;
}
// phase_mechanical.f90:280-294
void sc_default_phase_numerics(sc_phase_numerics* p) {
  p->r_cutback_min = 1e-3;
  p->r_cutback = 0.25;
  p->r_increase = 1.5;
  p->eps_rel_state = 1e-6;
  p->N_iter_state_max = 20;
  p->N_iter_Lp_max = 40;
  p->r_linesearch_Lp = 0.5;
  p->eps_rel_Lp = 1e-6;
  p->eps_abs_Lp = 1e-8;
}
// src/grid/config.h:50-78
void sc_default_grid_numerics(sc_grid_numerics* g) {
  g->itmin = 1;
  g->itmax = 100;
  g->update_gamma = 0;
  g->eps_abs_div = 1e-4;
  g->eps_rel_div = 5e-4;
  g->eps_abs_P = 1e3;
  g->eps_rel_P = 1e-3;
  g->divergence_correction = 2;
  g->derivative = 0;
  g->stag_iter_max = 10;
  g->max_cut_back = 3;
  g->ngmres_m = 30;
  g->fft_plan_mode = 1;
}

// ---- RVE context -------------------------------------------------------------
int sc_rve_create(const sc_rve_desc* d, sc_rve** out) {
  return guarded([&] {
    if (d->precision != SC_PRECISION_FP64 && d->precision != SC_PRECISION_MIXED)
      throw std::runtime_error("unknown precision");
    auto* R = new sc_rve();
    try {
      R->r.reset(new sca::Rve(*d));
    } catch (...) {
      delete R;
      throw;
    }
    *out = R;
    return 0;
  });
}

void sc_rve_destroy(sc_rve* R) { delete R; }

int sc_rve_init(sc_rve* R) {
  return timed(R, [&] {
    R->r->init();
    return 0;
  });
}

int sc_rve_forward(sc_rve* R, int cut_back, int guess, double dt, double dt_old, double t_remaining,
                   const sc_loadstep* step) {
  return timed(R, [&] {
    R->r->forward(cut_back != 0, guess != 0, dt, dt_old, t_remaining, *step);
    return 0;
  });
}

int sc_rve_solution(sc_rve* R, sc_solution* out) {
  return timed(R, [&] {
    *out = R->r->calculate_solution();
    return 0;
  });
}

int sc_rve_update_coords(sc_rve* R) {
  return timed(R, [&] {
    R->r->update_coords();
    return 0;
  });
}

int sc_rve_checkpoint(sc_rve* R, const char* path) {
  return guarded([&] {
    std::vector<char> blob;
    R->r->save(blob);
    FILE* f = std::fopen(path, "wb");
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
    const size_t w = std::fwrite(blob.data(), 1, blob.size(), f);
    const int c = std::fclose(f);
    if (w != blob.size() || c != 0) throw std::runtime_error(std::string("short write to ") + path);
    return 0;
  });
}

int sc_rve_restore(sc_rve* R, const char* path) {
  return guarded([&] {
    FILE* f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot read ") + path);
    std::vector<char> blob;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) blob.insert(blob.end(), buf, buf + n);
    std::fclose(f);
    R->r->load(blob);
    return 0;
  });
}

// Displacements of the cell centres and of the first (Nx+1)(Ny+1)Nz nodes
// (the reference passes that count as sharedNodesBegin for a single process,
// discretization_grid.cpp:81), from the coordinates of the last
// update_coords, as discretization_result writes them.
static void rve_displacements(sca::Rve& v, std::vector<double>& u_p, std::vector<double>& u_n) {
  const int nx = v.cells[0], ny = v.cells[1], nz = v.cells[2];
  const int64_t N = v.N, nodes = (int64_t)(nx + 1) * (ny + 1) * nz;
  u_p.resize(3 * (size_t)N);
  u_n.resize(3 * (size_t)nodes);
  SCA_HIP(hipMemcpyAsync(u_p.data(), v.d_xp, 3 * N * sizeof(double), hipMemcpyDeviceToHost, v.st));
  SCA_HIP(hipMemcpyAsync(u_n.data(), v.d_xn, 3 * nodes * sizeof(double), hipMemcpyDeviceToHost, v.st));
  SCA_HIP(hipStreamSynchronize(v.st));
  const double step[3] = {v.size[0] / nx, v.size[1] / ny, v.size[2] / nz};
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int64_t c = i + (int64_t)nx * (j + (int64_t)ny * k);
        u_p[3 * c] -= (i + 0.5) * step[0];
        u_p[3 * c + 1] -= (j + 0.5) * step[1];
        u_p[3 * c + 2] -= (k + 0.5) * step[2];
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        const int64_t n = i + (int64_t)(nx + 1) * (j + (int64_t)(ny + 1) * k);
        u_n[3 * n] -= i * step[0];
        u_n[3 * n + 1] -= j * step[1];
        u_n[3 * n + 2] -= k * step[2];
      }
}

int sc_rve_result_open(sc_rve* R, const char* path, const sc_result_desc* d) {
  return guarded([&] {
    if (!path || !*path) throw std::runtime_error("no result file name");
    sca::ResultSetup s;
    if (d && d->material_yaml) s = sca::result_setup_from_material(d->material_yaml);
    if (d && d->numerics_yaml) s.numerics_yaml = d->numerics_yaml;
    s.compression = d ? d->compression != 0 : true;
    R->r->on_result = nullptr;
    R->result.reset();
    R->result.reset(new sca::ResultFile(path, s, R->r->N));
    R->r->on_result = [R](int inc, double t) {
      R->r->sync_response_F();
      std::vector<double> u_p, u_n;
      rve_displacements(*R->r, u_p, u_n);
      R->r->refresh_state();
      R->result->increment(inc, t, R->r->mp, R->r->st, u_p, u_n);
    };
    return 0;
  });
}

int sc_rve_result_write(sc_rve* R, int inc, double t) {
  return guarded([&] {
    if (!R->result) throw std::runtime_error("no result file open");
    R->r->on_result(inc, t);
    return 0;
  });
}

int sc_rve_result_close(sc_rve* R) {
  return guarded([&] {
    R->r->on_result = nullptr;
    R->result.reset();
    return 0;
  });
}

int sc_dadf5_create(const char* path, int compression, sc_dadf5** out) {
  return guarded([&] {
    auto* h = new sc_dadf5();
    try {
      h->w.reset(new sca::Dadf5Writer(path, compression != 0));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return 0;
  });
}

int sc_dadf5_dataset_f64(sc_dadf5* h, const char* name, int rank, const int64_t* dims, const double* data) {
  return guarded([&] {
    if (rank < 1 || rank > 8) throw std::runtime_error("dataset rank must be 1..8");
    std::vector<uint64_t> d;
    for (int k = 0; k < rank; ++k) {
      if (dims[k] < 0) throw std::runtime_error("negative dataset dimension");
      d.push_back((uint64_t)dims[k]);
    }
    h->w->dataset_f64(name, d, data);
    return 0;
  });
}

int sc_dadf5_dataset_str(sc_dadf5* h, const char* name, const char* text) {
  return guarded([&] {
    h->w->dataset_str(name, text);
    return 0;
  });
}

int sc_dadf5_attr_str(sc_dadf5* h, const char* obj, const char* name, const char* value) {
  return guarded([&] {
    h->w->attr_str(obj, name, value);
    return 0;
  });
}

int sc_dadf5_attr_f64(sc_dadf5* h, const char* obj, const char* name, double value) {
  return guarded([&] {
    h->w->attr_f64(obj, name, value);
    return 0;
  });
}

int sc_dadf5_close(sc_dadf5* h) {
  if (!h) return 0;
  const int rc = guarded([&] {
    h->w->flush();
    return 0;
  });
  delete h;
  return rc;
}

void sc_om2qu(const double om[9], double q[4]) { sca::om2qu_host(om, q); }

int sc_rve_run_loadcase(sc_rve* R, const sc_loadstep* steps, int n_steps, int* inc_iterations, int* inc_evaluations,
                        int max_incs) {
  return timed(R, [&] { return R->r->run_loadcase(steps, n_steps, inc_iterations, inc_evaluations, max_incs); });
}

int sc_rve_form_residual(sc_rve* R, const double* F, double* r) {
  return timed(R, [&] {
    sca::Rve& v = *R->r;
    const int64_t N = v.N;
    DBuf a(9 * N), soa(9 * N), res(9 * N);
    a.put(F, v.st);
    sca::launch_aos_to_soa(a.p, soa.p, N, 9, v.st);
    v.form_residual(soa.p, res.p);
    v.complete_response_scalars();
    v.sync_response_F();
    sca::launch_soa_to_aos(res.p, a.p, N, 9, v.st);
    a.get(r, v.st);
    SCA_HIP(hipStreamSynchronize(v.st));
    return 0;
  });
}

int sc_rve_get_field(sc_rve* R, int field, double* out, int64_t n) {
  return guarded([&] { return sca::rve_get_field(*R->r, field, out, n); });
}

}  // extern "C"

// host copy of a field in the reference's layout (sc_rve_get_field)
int sca::rve_get_field(sca::Rve& v, int field, double* out, int64_t n) {
  v.refresh_state();
  {
    const int64_t N = v.N;
    const double* src = nullptr;
    int ncomp = 9;
    switch (field) {
      case SC_FIELD_F: src = v.X; break;
      case SC_FIELD_P: src = v.mp.P; break;
      case SC_FIELD_FP: src = v.mp.Fp; break;
      case SC_FIELD_LP: src = v.mp.Lp; break;
      case SC_FIELD_S: src = v.mp.S; break;
      case SC_FIELD_FE: src = v.mp.Fe; break;
      case SC_FIELD_F_LAST_INC: src = v.F_last_inc; break;
      case SC_FIELD_XI: src = v.mp.xi; ncomp = 12; break;
      case SC_FIELD_GAMMA: src = v.mp.gamma; ncomp = 12; break;
      case SC_FIELD_ITERATIONS: {
        if (n < 2 * N) throw std::runtime_error("output buffer too small");
        std::vector<int> it(2 * N);
        SCA_HIP(hipMemcpyAsync(it.data(), v.mp.iters, 2 * N * sizeof(int), hipMemcpyDeviceToHost, v.st));
        SCA_HIP(hipStreamSynchronize(v.st));
        for (int64_t c = 0; c < N; ++c) {
          out[2 * c] = it[c];
          out[2 * c + 1] = it[N + c];
        }
        return 0;
      }
      case SC_FIELD_X_P:
      case SC_FIELD_X_N: {
        const int64_t m = field == SC_FIELD_X_P ? 3 * N
                                                 : (int64_t)3 * (v.cells[0] + 1) * (v.cells[1] + 1) * (v.cells[2] + 1);
        if (n < m) throw std::runtime_error("output buffer too small");
        SCA_HIP(hipMemcpyAsync(out, field == SC_FIELD_X_P ? v.d_xp : v.d_xn, m * sizeof(double),
                               hipMemcpyDeviceToHost, v.st));
        SCA_HIP(hipStreamSynchronize(v.st));
        return 0;
      }
      default: throw std::runtime_error("unknown field");
    }
    if (n < ncomp * N) throw std::runtime_error("output buffer too small");
    DBuf a((size_t)ncomp * N);
    sca::launch_soa_to_aos(src, a.p, N, ncomp, v.st);
    a.get(out, v.st);
    SCA_HIP(hipStreamSynchronize(v.st));
    return 0;
  }
}

extern "C" {

int sc_rve_set_field(sc_rve* R, int field, const double* in, int64_t n) {
  return guarded([&] {
    sca::Rve& v = *R->r;
    if (field != SC_FIELD_F) throw std::runtime_error("field not settable");
    if (n < 9 * v.N) throw std::runtime_error("input too small");
    DBuf a(9 * v.N);
    a.put(in, v.st);
    sca::launch_aos_to_soa(a.p, v.X, v.N, 9, v.st);
    SCA_HIP(hipStreamSynchronize(v.st));
    v.x_uniform = false;
    return 0;
  });
}

int sc_rve_get_scalars(sc_rve* R, sc_solver_scalars* o) {
  const sca::Rve& v = *R->r;
  std::memcpy(o->P_av, v.P_av, sizeof o->P_av);
  std::memcpy(o->P_aim, v.P_aim, sizeof o->P_aim);
  std::memcpy(o->F_aim, v.F_aim, sizeof o->F_aim);
  std::memcpy(o->F_aim_dot, v.F_aim_dot, sizeof o->F_aim_dot);
  std::memcpy(o->F_aim_last_inc, v.F_aim_last_inc, sizeof o->F_aim_last_inc);
  std::memcpy(o->C_volAvg, v.C_volAvg, sizeof o->C_volAvg);
  std::memcpy(o->C_minMaxAvg, v.C_minMaxAvg, sizeof o->C_minMaxAvg);
  std::memcpy(o->S, v.S, sizeof o->S);
  std::memcpy(o->C_ref, v.C_ref, sizeof o->C_ref);
  o->err_div = v.err_div;
  o->err_BC = v.err_BC;
  o->total_iter = v.total_iter;
  return 0;
}
int sc_rve_get_stats(sc_rve* R, sc_stats* o) {
  *o = R->r->stats;
  return 0;
}
int sc_rve_reset_stats(sc_rve* R) {
  std::memset(&R->r->stats, 0, sizeof R->r->stats);
  return 0;
}

int sc_rve_profile(sc_rve* R, int enable) { return sc_rve_profile_classes(R, enable ? ~0u : 0u); }
int sc_rve_profile_classes(sc_rve* R, unsigned mask) {
  return guarded([&] {
    R->r->kt.flush();
    R->r->kt.on = mask != 0;
    R->r->kt.mask = mask;
    if (mask) R->r->kt.reset();
    return 0;
  });
}
int sc_rve_get_kernel_times(sc_rve* R, sc_kernel_times* out) {
  return guarded([&] {
    R->r->kt.flush();
    *out = R->r->kt.acc;
    return 0;
  });
}

// ---- outer solver replay -------------------------------------------------------
int sc_ngmres_replay(int64_t n, int m, const double* x0, int n_evals, const double* r_rec, double* x_eval) {
  return guarded([&] {
    Stream s;
    sca::Ngmres ng(n, s.s);
    DBuf X(n);
    X.put(x0, s.s);
    int e = 0;
    const int reason = ng.solve<double>(
        X.p, m,
        [&](const double* x, double* f) {
          if (e >= n_evals) throw std::runtime_error("replay ran past the recorded evaluations");
          SCA_HIP(hipMemcpyAsync(x_eval + (int64_t)e * n, x, n * sizeof(double), hipMemcpyDeviceToHost, s.s));
          SCA_HIP(hipMemcpyAsync(f, r_rec + (int64_t)e * n, n * sizeof(double), hipMemcpyHostToDevice, s.s));
          s.sync();
          e++;
        },
        [&](int) { return e >= n_evals ? 1 : 0; }, [](int) {});
    (void)reason;
    s.sync();
    return e == n_evals ? 0 : 3;
  });
}

// ---- spectral building blocks ---------------------------------------------------
int sc_spectral_xi(const int cells[3], const double size[3], int derivative, double* xi1st, double* xi2nd) {
  return guarded([&] {
    const int Nx = cells[0], Ny = cells[1], Nz = cells[2], nr = Nx / 2 + 1;
    const double TAU = 2.0 * M_PI;
    for (int y = 0; y < Ny; ++y)
      for (int z = 0; z < Nz; ++z)
        for (int i = 0; i < nr; ++i) {
          const int ks[3] = {i, (y > Ny / 2) ? y - Ny : y, (z > Nz / 2) ? z - Nz : z};
          const int idx[3] = {i, y, z};
          const int64_t base = 3 * (i + (int64_t)nr * (z + (int64_t)Nz * y));
          for (int p = 0; p < 3; ++p) {
            std::complex<double> v;
            if (derivative == 0) {
              v = {0.0, TAU * ks[p] / size[p]};
            } else if (derivative == 1) {
              v = {0.0, std::sin(TAU * ks[p] / cells[p]) / (2.0 * size[p] / cells[p])};
            } else {
              const int p1 = (p + 1) % 3, p2 = (p + 2) % 3;
              v = std::complex<double>(std::cos(TAU * ks[p] / cells[p]) - (p == 0 ? 1.0 : -1.0),
                                       std::sin(TAU * ks[p] / cells[p])) *
                  std::complex<double>(std::cos(TAU * ks[p1] / cells[p1]) + 1.0, std::sin(TAU * ks[p1] / cells[p1])) *
                  std::complex<double>(std::cos(TAU * ks[p2] / cells[p2]) + 1.0, std::sin(TAU * ks[p2] / cells[p2])) /
                  std::complex<double>(4.0 * size[p] / cells[p], 0.0);
            }
            const bool nyq = derivative == 0 && cells[p] % 2 == 0 && idx[p] == cells[p] / 2;
            xi2nd[2 * (base + p)] = v.real();
            xi2nd[2 * (base + p) + 1] = v.imag();
            xi1st[2 * (base + p)] = nyq ? 0.0 : v.real();
            xi1st[2 * (base + p) + 1] = nyq ? 0.0 : v.imag();
          }
        }
    return 0;
  });
}

int sc_fft_r2c(const int cells[3], int howmany, const double* in, double* out) {
  return guarded([&] {
    Stream s;
    sca::FftPlan<double> plan;
    plan.init(cells);
    std::vector<double> planes, sp((size_t)2 * howmany * plan.K);
    interleaved_to_planes(in, howmany, plan.N, planes);
    DBuf re(planes.size()), spec(sp.size());
    re.put(planes.data(), s.s);
    plan.forward(re.p, spec.p, howmany, s.s);
    spec.get(sp.data(), s.s);
    s.sync();
    plan.release();
    spec_planes_to_ref(cells, howmany, sp, out);
    return 0;
  });
}

int sc_fft_c2r(const int cells[3], int howmany, const double* in, double wgt, double* out) {
  return guarded([&] {
    Stream s;
    sca::FftPlan<double> plan;
    plan.init(cells);
    std::vector<double> sp, planes((size_t)howmany * plan.N);
    ref_to_spec_planes(cells, howmany, in, sp);
    DBuf re(planes.size()), spec(sp.size());
    spec.put(sp.data(), s.s);
    plan.backward(spec.p, re.p, howmany, wgt, s.s);
    re.get(planes.data(), s.s);
    s.sync();
    plan.release();
    for (int64_t n = 0; n < plan.N; ++n)
      for (int c = 0; c < howmany; ++c) out[howmany * n + c] = planes[c * plan.N + n];
    return 0;
  });
}

int sc_divergence_rms(const int cells[3], const double size[3], const double scaled_size[3], const double* xi1st,
                      const double* field, double* rms) {
  return guarded([&] {
    Stream s;
    sca::FftPlan<double> plan;
    plan.init(cells);
    std::vector<double> planes;
    interleaved_to_planes(field, 9, plan.N, planes);
    const std::vector<double> tab = xi_to_kernel_table(cells, xi1st);
    DBuf re(planes.size()), spec((size_t)18 * plan.K), xt(tab.size()), part(sca::gamma_blocks(plan.K)), sum(1);
    re.put(planes.data(), s.s);
    xt.put(tab.data(), s.s);
    plan.forward(re.p, spec.p, 9, s.s);
    sca::launch_gamma_div<double>(spec.p, plan.K, gamma_host_table(cells, size, scaled_size, nullptr, 0), nullptr,
                                  xt.p, part.p, s.s);
    sca::launch_final_reduce(part.p, 1, sca::gamma_blocks(plan.K), sum.p, s.s);
    double h = 0.0;
    sum.get(&h, s.s);
    s.sync();
    plan.release();
    // an x extent of 1 has nr = 1: the reference counts that plane twice and halves
    if (plan.nr == 1) h *= 2.0;
    double r = std::sqrt(h) / (double)plan.N;
    if (cells[0] == 1) r /= 2;
    *rms = r;
    return 0;
  });
}

int sc_gamma_convolution(const int cells[3], const double* xi1st, const double* C_ref, const double field_aim[9],
                         double* field) {
  return guarded([&] {
    Stream s;
    sca::FftPlan<double> plan;
    plan.init(cells);
    std::vector<double> planes;
    interleaved_to_planes(field, 9, plan.N, planes);
    const std::vector<double> tab = xi_to_kernel_table(cells, xi1st);
    DBuf re(planes.size()), spec((size_t)18 * plan.K), xt(tab.size()), aim(9);
    re.put(planes.data(), s.s);
    xt.put(tab.data(), s.s);
    aim.put(field_aim, s.s);
    plan.forward(re.p, spec.p, 9, s.s);
    sca::launch_gamma_div<double>(spec.p, plan.K, gamma_host_table(cells, nullptr, nullptr, C_ref, 1), aim.p, xt.p,
                                  nullptr, s.s);
    plan.backward(spec.p, re.p, 9, 1.0, s.s);
    re.get(planes.data(), s.s);
    s.sync();
    plan.release();
    for (int64_t n = 0; n < plan.N; ++n)
      for (int c = 0; c < 9; ++c) field[9 * n + c] = planes[c * plan.N + n];
    return 0;
  });
}

int sc_masked_compliance(const double C[81], const double rot_q[4], const int stress_mask[9], double S_out[81]) {
  return guarded([&] {
    sca::masked_compliance_host(C, rot_q, stress_mask, S_out);
    return 0;
  });
}

int sc_scaled_geom_size(const int cells[3], const double size[3], int dc, double scaled[3]) {
  return guarded([&] {
    int idx[3] = {0, 1, 2};
    if (dc == 1) {
      std::sort(idx, idx + 3, [&](int a, int b) { return size[a] < size[b]; });
      for (int j = 0; j < 3; ++j) scaled[j] = size[j] / size[idx[1]];
    } else if (dc == 2) {
      const double ng[3] = {size[0] / cells[0], size[1] / cells[1], size[2] / cells[2]};
      std::sort(idx, idx + 3, [&](int a, int b) { return ng[a] < ng[b]; });
      for (int j = 0; j < 3; ++j) scaled[j] = size[j] / size[idx[1]] * cells[idx[1]];
    } else {
      for (int j = 0; j < 3; ++j) scaled[j] = size[j];
    }
    return 0;
  });
}

int sc_forward_field(int64_t n, double dt, const double* last, const double* rate, const double* aim, double* out) {
  return guarded([&] {
    Stream s;
    DBuf a(9 * n), b(9 * n), sa(9 * n), sb(9 * n), part((size_t)sca::red_blocks() * 9), sm(32);
    a.put(last, s.s);
    b.put(rate, s.s);
    sca::launch_aos_to_soa(a.p, sa.p, n, 9, s.s);
    sca::launch_aos_to_soa(b.p, sb.p, n, 9, s.s);
    sca::launch_fwd_field(sa.p, sb.p, dt, sa.p, 9 * n, s.s);
    if (aim) {
      std::vector<int> ops(9, 2);
      std::vector<const double*> v(9);
      for (int q = 0; q < 9; ++q) v[q] = sa.p + q * n;
      SCA_HIP(hipMemcpyAsync(sm.p + 16, aim, 9 * sizeof(double), hipMemcpyHostToDevice, s.s));
      sca::launch_partial_reduce<double>(ops.data(), n, 9, v.data(), nullptr, part.p, s.s);
      sca::launch_final_reduce(part.p, 9, sca::red_blocks(), sm.p, s.s);
      sca::launch_fwd_diff(sm.p, 1.0 / (double)n, sm.p + 16, sm.p + 9 + 16, s.s);
      sca::launch_sub_plane_const(sa.p, sm.p + 25, n, s.s);
    }
    sca::launch_soa_to_aos(sa.p, a.p, n, 9, s.s);
    a.get(out, s.s);
    s.sync();
    return 0;
  });
}

int sc_calculate_rate(int64_t n, int het, const double* f0, const double* f, double dt, const double avRate[9],
                      double* out) {
  return guarded([&] {
    Stream s;
    DBuf a(9 * n), b(9 * n), sa(9 * n), sb(9 * n), av(9), o(9 * n);
    a.put(f0, s.s);
    b.put(f, s.s);
    av.put(avRate, s.s);
    sca::launch_aos_to_soa(a.p, sa.p, n, 9, s.s);
    sca::launch_aos_to_soa(b.p, sb.p, n, 9, s.s);
    sca::launch_rate(het, sa.p, sb.p, dt, av.p, o.p, n, s.s);
    sca::launch_soa_to_aos(o.p, a.p, n, 9, s.s);
    a.get(out, s.s);
    s.sync();
    return 0;
  });
}

int sc_response_reductions(int64_t n, const double* P, const double* dPdF, const double* rot_q, double P_av[9],
                           double C_volAvg[81], double C_minMaxAvg[81]) {
  return guarded([&] {
    Stream s;
    const int nb = sca::red_blocks();
    DBuf a(9 * n), sp(9 * n), d(81 * n), sd(81 * n), part((size_t)std::max<int64_t>(nb * 81, 9 * sca::mp_partial_blocks(n))),
        sm(256);
    long long* d_i = nullptr;
    SCA_HIP(hipMalloc(&d_i, 2 * nb * sizeof(long long)));
    std::unique_ptr<long long, void (*)(long long*)> guard_i(d_i, [](long long* p) { (void)hipFree(p); });
    a.put(P, s.s);
    d.put(dPdF, s.s);
    sca::launch_aos_to_soa(a.p, sp.p, n, 9, s.s);
    sca::launch_aos_to_soa(d.p, sd.p, n, 81, s.s);
    sca::launch_pav_partial(sp.p, n, part.p, s.s);
    sca::launch_final_reduce(part.p, 9, sca::mp_partial_blocks(n), sm.p, s.s);
    double sums[9];
    SCA_HIP(hipMemcpyAsync(sums, sm.p, sizeof sums, hipMemcpyDeviceToHost, s.s));
    s.sync();
    const double wgt = 1.0 / (double)n;
    double pav[9];
    for (int q = 0; q < 9; ++q) pav[q] = sums[q] * wgt;
    if (rot_q) {
      double R[9], Rt[9], t[9];
      sca::qu2om_host(rot_q, R);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Rt[i + 3 * j] = R[j + 3 * i];
      mul33(R, pav, t);
      mul33(t, Rt, pav);
    }
    std::memcpy(P_av, pav, sizeof pav);
    std::vector<int> ops(81, 2);
    std::vector<const double*> v(81);
    for (int q = 0; q < 81; ++q) v[q] = sd.p + q * n;
    sca::launch_partial_reduce<double>(ops.data(), n, 81, v.data(), nullptr, part.p, s.s);
    sca::launch_final_reduce(part.p, 81, nb, sm.p, s.s);
    double vol[81];
    SCA_HIP(hipMemcpyAsync(vol, sm.p, sizeof vol, hipMemcpyDeviceToHost, s.s));
    s.sync();
    for (int q = 0; q < 81; ++q) C_volAvg[q] = vol[q] * wgt;
    DBuf n2(n);
    int64_t imx = -1, imn = -1;
    sca::tangent_minmax_select(sd.p, n, n2.p, part.p, d_i, imx, imn, s.s);
    for (int q = 0; q < 81; ++q) {
      const double x = imx >= 0 ? dPdF[81 * imx + q] : 0.0;
      const double y = imn >= 0 ? dPdF[81 * imn + q] : std::numeric_limits<double>::max();
      C_minMaxAvg[q] = 0.5 * (x + y);
    }
    return 0;
  });
}

// ---- material-point batch ---------------------------------------------------------
int sc_mp_update(int64_t n, double dt, const sc_material* m, const sc_phase_numerics* num, const double* F,
                 const double* F0, const double* Fp0, const double* Lp0, const double* S0, const double* xi0,
                 const double* gamma0, double* Lp, double* P, double* Fp, double* S, double* Fe, double* xi,
                 double* gamma, double* dPdF, int* broken, int* iterations) {
  return guarded([&] {
    if (n <= 0) return 0;
    Stream s;
    sca::MatConst mc;
    sca::build_matconst(*m, *num, mc);
    // SoA planes: 9 each for F F0 Fp0 Lp0 S0 Lp P Fp S Fe, 12 each for xi0 gamma0 xi gamma
    DBuf planes((size_t)(10 * 9 + 4 * 12) * n), stage((size_t)81 * n), sub((size_t)sca::mp_sub_planes() * n);
    double* b = planes.p;
    sca::MpState st;
    st.N = n;
    double** t9[10] = {&st.F, &st.F0, &st.Fp0, &st.Lp0, &st.S0, &st.Lp, &st.P, &st.Fp, &st.S, &st.Fe};
    for (int k = 0; k < 10; ++k) *t9[k] = b + (int64_t)9 * n * k;
    double** t12[4] = {&st.xi0, &st.gamma0, &st.xi, &st.gamma};
    for (int k = 0; k < 4; ++k) *t12[k] = b + (int64_t)90 * n + (int64_t)12 * n * k;
    int *d_ill = nullptr, *d_br = nullptr, *d_it = nullptr;
    SCA_HIP(hipMalloc(&d_ill, sizeof(int)));
    std::unique_ptr<int, void (*)(int*)> g1(d_ill, [](int* p) { (void)hipFree(p); });
    SCA_HIP(hipMalloc(&d_br, n * sizeof(int)));
    std::unique_ptr<int, void (*)(int*)> g2(d_br, [](int* p) { (void)hipFree(p); });
    SCA_HIP(hipMalloc(&d_it, (3 * n + 1) * sizeof(int)));
    std::unique_ptr<int, void (*)(int*)> g3(d_it, [](int* p) { (void)hipFree(p); });
    st.iters = d_it;
    st.repivot = d_it + 2 * n;
    auto up = [&](const double* h, double* dst, int nc) {
      SCA_HIP(hipMemcpyAsync(stage.p, h, nc * n * sizeof(double), hipMemcpyHostToDevice, s.s));
      sca::launch_aos_to_soa(stage.p, dst, n, nc, s.s);
    };
    auto down = [&](const double* src, double* h, int nc) {
      sca::launch_soa_to_aos(src, stage.p, n, nc, s.s);
      SCA_HIP(hipMemcpyAsync(h, stage.p, nc * n * sizeof(double), hipMemcpyDeviceToHost, s.s));
      s.sync();
    };
    up(F, st.F, 9);
    up(F0, st.F0, 9);
    up(Fp0, st.Fp0, 9);
    up(Lp0, st.Lp0, 9);
    up(S0, st.S0, 9);
    up(Lp, st.Lp, 9);
    up(xi0, st.xi0, 12);
    up(gamma0, st.gamma0, 12);
    sca::launch_mp_response(st, mc, dt, d_ill, sub.p, s.s, d_br);
    down(st.Lp, Lp, 9);
    down(st.P, P, 9);
    down(st.Fp, Fp, 9);
    down(st.S, S, 9);
    down(st.Fe, Fe, 9);
    down(st.xi, xi, 12);
    down(st.gamma, gamma, 12);
    if (dPdF) {
      DBuf d((size_t)81 * n);
      sca::launch_mp_tangent(st, mc, dt, d.p, s.s);
      sca::launch_soa_to_aos(d.p, stage.p, n, 81, s.s);
      SCA_HIP(hipMemcpyAsync(dPdF, stage.p, 81 * n * sizeof(double), hipMemcpyDeviceToHost, s.s));
    }
    if (broken) SCA_HIP(hipMemcpyAsync(broken, d_br, n * sizeof(int), hipMemcpyDeviceToHost, s.s));
    if (iterations) {
      std::vector<int> it(2 * n);
      SCA_HIP(hipMemcpyAsync(it.data(), d_it, 2 * n * sizeof(int), hipMemcpyDeviceToHost, s.s));
      s.sync();
      for (int64_t c = 0; c < n; ++c) {
        iterations[2 * c] = it[c];
        iterations[2 * c + 1] = it[n + c];
      }
    }
    s.sync();
    return 0;
  });
}

void sc_qu2om(const double q[4], double om[9]) { sca::qu2om_host(q, om); }
void sc_rotational_part(const double F[9], double R[9]) {
  try {
    sca::rotational_part_host(F, R);
  } catch (std::exception& e) {
    g_err = e.what();
    for (int k = 0; k < 9; ++k) R[k] = std::nan("");
  }
}

// ---- math / rotation helpers of utilities_fortran.h ---------------------------------
// math_invert (math.f90:599): LU with partial pivoting, column-major n x n
void f_math_invert(double* invA, int* err, double* A, int* n_) {
  const int n = *n_;
  std::vector<double> a((size_t)n * n), inv((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) a[(size_t)i * n + j] = A[i + (size_t)n * j];
  std::vector<int> piv(n);
  bool ok = true;
  for (int k = 0; k < n && ok; ++k) {
    int p = k;
    double mx = std::fabs(a[(size_t)k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(a[(size_t)i * n + k]) > mx) {
        mx = std::fabs(a[(size_t)i * n + k]);
        p = i;
      }
    piv[k] = p;
    if (a[(size_t)p * n + k] == 0.0) {
      ok = false;
      break;
    }
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(a[(size_t)k * n + j], a[(size_t)p * n + j]);
    for (int i = k + 1; i < n; ++i) {
      a[(size_t)i * n + k] /= a[(size_t)k * n + k];
      for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= a[(size_t)i * n + k] * a[(size_t)k * n + j];
    }
  }
  if (ok) {
    std::vector<double> col(n);
    for (int c = 0; c < n; ++c) {
      for (int i = 0; i < n; ++i) col[i] = (i == c) ? 1.0 : 0.0;
      for (int k = 0; k < n; ++k)
        if (piv[k] != k) std::swap(col[k], col[piv[k]]);
      for (int i = 1; i < n; ++i)
        for (int j = 0; j < i; ++j) col[i] -= a[(size_t)i * n + j] * col[j];
      for (int i = n - 1; i >= 0; --i) {
        for (int j = i + 1; j < n; ++j) col[i] -= a[(size_t)i * n + j] * col[j];
        col[i] /= a[(size_t)i * n + i];
      }
      for (int i = 0; i < n; ++i) invA[i + (size_t)n * c] = col[i];
    }
  } else {
    for (int k = 0; k < n * n; ++k) invA[k] = 0.0;
  }
  *err = ok ? 0 : 1;
}

// math_3333to99 / math_99to3333 (math.f90:819, :849): plain map, row 3(i-1)+j
void f_math_3333to99(double* m3333, double* m99) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) m99[(3 * i + j) + 9 * (3 * k + l)] = m3333[i + 3 * j + 9 * k + 27 * l];
}
void f_math_99to3333(double* m99, double* m3333) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) m3333[i + 3 * j + 9 * k + 27 * l] = m99[(3 * i + j) + 9 * (3 * k + l)];
}
// math_mul3333xx33 (math.f90:428)
void f_math_mul3333xx33(double* A, double* B, double* res) {
  double t[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int l = 0; l < 3; ++l)
        for (int k = 0; k < 3; ++k) s += A[i + 3 * j + 9 * k + 27 * l] * B[k + 3 * l];
      t[i + 3 * j] = s;
    }
  std::memcpy(res, t, sizeof t);
}
// rotate_tensor2 / rotate_tensor4 (rotations.f90:289-328): passive R T R^T
// unless active == 1 (R^T T R)
void f_rotate_tensor2(double* q, double* T, double* out, int* active) {
  double R[9], Rt[9], t[9];
  rotation_matrix(q, active ? *active : 0, R);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Rt[i + 3 * j] = R[j + 3 * i];
  mul33(R, T, t);
  mul33(t, Rt, out);
}
void f_rotate_tensor4(double* q, double* T, double* out, int* active) {
  double R[9];
  rotation_matrix(q, active ? *active : 0, R);
  std::vector<double> tmp(81);
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
          tmp[i + 3 * j + 9 * k + 27 * l] = s;
        }
  std::memcpy(out, tmp.data(), 81 * sizeof(double));
}

}  // extern "C"
