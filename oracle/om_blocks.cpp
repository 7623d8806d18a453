// ORACLE -- test infrastructure only (see oracle_internal.h).
// C-ABI building blocks: spectral utilities, material-point batch update and
// the f_math_* / f_rotate_* helpers (src/math.f90:428,599,819,849;
// src/rotations.f90:331-360).
#include <omp.h>

#include <cstring>
#include <stdexcept>

#include "oracle_internal.h"

namespace orc {
extern thread_local std::string g_err;
}
using namespace orc;

extern "C" {

int sc_spectral_xi(const int cells[3], const double size[3], int derivative, double* xi1st, double* xi2nd) {
  std::vector<cplx> a, b;
  spectral_xi(cells, size, derivative, a, b);
  std::memcpy(xi1st, a.data(), a.size() * sizeof(cplx));
  std::memcpy(xi2nd, b.data(), b.size() * sizeof(cplx));
  return 0;
}

int sc_fft_r2c(const int cells[3], int howmany, const double* in, double* out) {
  const int64_t N = (int64_t)cells[0] * cells[1] * cells[2];
  const int64_t K = (int64_t)(cells[0] / 2 + 1) * cells[1] * cells[2];
  std::vector<double> comp(N);
  std::vector<cplx> spec(K);
  cplx* o = reinterpret_cast<cplx*>(out);
  for (int c = 0; c < howmany; ++c) {
    for (int64_t n = 0; n < N; ++n) comp[n] = in[howmany * n + c];
    fft3d_r2c(cells, comp.data(), spec.data());
    for (int64_t k = 0; k < K; ++k) o[howmany * k + c] = spec[k];
  }
  return 0;
}

int sc_fft_c2r(const int cells[3], int howmany, const double* in, double wgt, double* out) {
  const int64_t N = (int64_t)cells[0] * cells[1] * cells[2];
  const int64_t K = (int64_t)(cells[0] / 2 + 1) * cells[1] * cells[2];
  std::vector<double> comp(N);
  std::vector<cplx> spec(K);
  const cplx* i_ = reinterpret_cast<const cplx*>(in);
  for (int c = 0; c < howmany; ++c) {
    for (int64_t k = 0; k < K; ++k) spec[k] = i_[howmany * k + c];
    fft3d_c2r(cells, spec.data(), comp.data());
    for (int64_t n = 0; n < N; ++n) out[howmany * n + c] = comp[n] * wgt;
  }
  return 0;
}

int sc_divergence_rms(const int cells[3], const double size[3], const double scaled_size[3], const double* xi1st,
                      const double* field, double* rms) {
  Grid g;
  grid_setup(g, cells, size, 0, 0);
  for (int p = 0; p < 3; ++p) g.scaled_size[p] = scaled_size[p];
  std::memcpy((void*)g.xi1st.data(), xi1st, g.xi1st.size() * sizeof(cplx));
  std::vector<cplx> hat;
  tensor_r2c(g, field, hat);
  *rms = divergence_rms(g, hat);
  return 0;
}

int sc_gamma_convolution(const int cells[3], const double* xi1st, const double* C_ref, const double field_aim[9],
                         double* field) {
  Grid g;
  const double size[3] = {1, 1, 1};
  grid_setup(g, cells, size, 0, 0);
  std::memcpy((void*)g.xi1st.data(), xi1st, g.xi1st.size() * sizeof(cplx));
  std::vector<cplx> hat;
  tensor_r2c(g, field, hat);
  gamma_apply(g, load3333(C_ref), hat, load33(field_aim));
  tensor_c2r(g, hat, 1.0, field);
  return 0;
}

int sc_masked_compliance(const double C[81], const double rot_q[4], const int stress_mask[9], double S_out[81]) {
  try {
    store3333(masked_compliance(load3333(C), rot_q, stress_mask), S_out);
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_scaled_geom_size(const int cells[3], const double size[3], int dc, double scaled[3]) {
  scaled_geom(cells, size, dc, scaled);
  return 0;
}

int sc_forward_field(int64_t n, double dt, const double* last, const double* rate, const double* aim, double* out) {
  double wgt = 1.0 / (double)n;
  for (int64_t i = 0; i < 9 * n; ++i) out[i] = last[i] + rate[i] * dt;
  if (aim) {
    double s[9] = {0};
    for (int64_t c = 0; c < n; ++c)
      for (int q = 0; q < 9; ++q) s[q] += out[9 * c + q];
    for (int64_t c = 0; c < n; ++c)
      for (int q = 0; q < 9; ++q) out[9 * c + q] -= s[q] * wgt - aim[q];
  }
  return 0;
}

int sc_calculate_rate(int64_t n, int het, const double* f0, const double* f, double dt, const double avRate[9],
                      double* out) {
  for (int64_t c = 0; c < n; ++c)
    for (int q = 0; q < 9; ++q) out[9 * c + q] = het ? (f[9 * c + q] - f0[9 * c + q]) / dt : avRate[q];
  return 0;
}

int sc_response_reductions(int64_t n, const double* P, const double* dPdF, const double* rot_q, double P_av[9],
                           double C_volAvg[81], double C_minMaxAvg[81]) {
  double wgt = 1.0 / (double)n;
  double pav[9] = {0};
  for (int64_t c = 0; c < n; ++c)
    for (int q = 0; q < 9; ++q) pav[q] += P[9 * c + q];
  for (int q = 0; q < 9; ++q) pav[q] *= wgt;
  M3 Pm = load33(pav);
  if (rot_q) Pm = rotTensor2(rot_q, Pm, false);
  store33(Pm, P_av);
  std::vector<double> vol(81, 0.0), n2(n);
  for (int64_t c = 0; c < n; ++c) {
    const double* d = dPdF + 81 * c;
    double s = 0.0;
    for (int q = 0; q < 81; ++q) s += d[q] * d[q];
    n2[c] = s;
    for (int q = 0; q < 81; ++q) vol[q] += d[q];
  }
  int64_t imax = -1, imin = -1;
  tangent_minmax_select(n2.data(), n, imax, imin);
  for (int q = 0; q < 81; ++q) {
    double a = imax >= 0 ? dPdF[81 * imax + q] : 0.0;
    double b = imin >= 0 ? dPdF[81 * imin + q] : std::numeric_limits<double>::max();
    C_minMaxAvg[q] = 0.5 * (a + b);
    C_volAvg[q] = vol[q] * wgt;
  }
  return 0;
}

int sc_mp_update(int64_t n, double dt, const sc_material* m, const sc_phase_numerics* num, const double* F,
                 const double* F0, const double* Fp0, const double* Lp0, const double* S0, const double* xi0,
                 const double* gamma0, double* Lp, double* P, double* Fp, double* S, double* Fe, double* xi,
                 double* gamma, double* dPdF, int* broken, int* iterations) {
  Phase ph;
  phase_setup(ph, *m, *num);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t c = 0; c < n; ++c) {
    PointState s;
    s.F = load33(F + 9 * c);
    s.F0 = load33(F0 + 9 * c);
    s.Fp0 = load33(Fp0 + 9 * c);
    s.Lp0 = load33(Lp0 + 9 * c);
    s.S0 = load33(S0 + 9 * c);
    s.Lp = load33(Lp + 9 * c);
    for (int i = 0; i < NSLIP; ++i) {
      s.xi0[i] = xi0[12 * c + i];
      s.gamma0[i] = gamma0[12 * c + i];
    }
    bool ok = mp_response(ph, s, dt);
    store33(s.Lp, Lp + 9 * c);
    store33(s.P, P + 9 * c);
    store33(s.Fp, Fp + 9 * c);
    store33(s.S, S + 9 * c);
    store33(s.Fe, Fe + 9 * c);
    for (int i = 0; i < NSLIP; ++i) {
      xi[12 * c + i] = s.xi[i];
      gamma[12 * c + i] = s.gamma[i];
    }
    if (dPdF) store3333(mp_dPdF(ph, s, dt), dPdF + 81 * c);
    if (broken) broken[c] = ok ? 0 : 1;
    if (iterations) {
      iterations[2 * c] = s.newton_iter;
      iterations[2 * c + 1] = s.state_iter;
    }
  }
  return 0;
}

void sc_qu2om(const double q[4], double om[9]) { store33(qu2om(q), om); }
void sc_rotational_part(const double F[9], double R[9]) { store33(rotationalPart(load33(F)), R); }

void f_math_invert(double* invA, int* err, double* A, int* n) {
  int N = *n;
  // column-major in and out: invert the transpose and transpose back == invert
  std::vector<double> a(N * N), inv(N * N);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) a[i * N + j] = A[i + N * j];
  bool e = invert(N, a.data(), inv.data());
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) invA[i + N * j] = inv[i * N + j];
  *err = e ? 1 : 0;
}
void f_math_3333to99(double* m3333, double* m99) {
  M99 r = t3333to99(load3333(m3333));
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) m99[i + 9 * j] = r.a[i][j];
}
void f_math_99to3333(double* m99, double* m3333) {
  M99 r;
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) r.a[i][j] = m99[i + 9 * j];
  store3333(t99to3333(r), m3333);
}
void f_math_mul3333xx33(double* A, double* B, double* res) { store33(mul3333xx33(load3333(A), load33(B)), res); }
void f_rotate_tensor2(double* q, double* T, double* out, int* active) {
  store33(rotTensor2(q, load33(T), *active == 1), out);
}
void f_rotate_tensor4(double* q, double* T, double* out, int* active) {
  store3333(rotTensor4(q, load3333(T), *active == 1), out);
}

}  // extern "C"
