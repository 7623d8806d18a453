// ORACLE -- test infrastructure only (see oracle_internal.h).
// The spectral basic solver on one RVE, restated from
//   src/grid/mech_solver_basic.cpp  init :19-108, calculate_solution :110-161,
//                                   forward :163-220, update_coords :222-240,
//                                   converged :242-284, form_residual :286-339
//   src/grid/mech_base.cpp          forward_field :213-247, calculate_rate :479-497,
//                                   base_update_coords :45-154
//   src/grid/spectral.cpp           constitutive_response :132-219
//   src/grid/damask_grid.cpp        forward_fields :19-34, solve_fields :36-67,
//                                   forward_time :69-80, loop_over_loadcase :82-196
// and PETSc's SNESNGMRES (PETSc >= 3.12, < 3.21; snesngmres.c / ngmresfunc.c,
// default options: m = 30, restart_it = 2, gammaA = gammaC = 2, epsilonB = 0.1,
// deltaB = 0.9, select and restart type "difference", basic line search with
// damping 1, max_it = 50), which is not part of the reference tree.
#include <omp.h>

#include <chrono>
#include <functional>
#include <cstring>
#include <stdexcept>

#include "oracle_internal.h"

namespace orc {
thread_local std::string g_err;
}

using namespace orc;

struct sc_rve {
  sc_rve_desc desc;
  std::vector<int> material_at;
  std::vector<double> orientations;
  Phase ph;
  Grid g;
  std::vector<PointState> pts;
  std::vector<double> homP, homdPdF;  // homogenization_P, homogenization_dPdF
  // MechBase / MechSolverBasic members (mech_base.h:75-107)
  std::vector<double> F, F_last_inc, F_dot;
  T4 C_ref, C_volAvg, C_minMaxAvg, C_volAvgLastInc, C_minMaxAvgLastInc, S;
  M3 P_av, P_aim, F_aim_dot, F_aim, F_aim_last_inc;
  int stress_mask[9];
  double rot_q[4];
  double delta_t = 0.0;
  int total_iter = 0;
  double err_BC = 0.0, err_div = 0.0;
  bool terminally_ill = false;
  // SNES bookkeeping
  int snes_iter = 0, nfuncs = 0;
  int solve_evals = 0;
  // coordinates
  std::vector<double> x_p, x_n;
  sc_stats stats;
};

namespace {

bool is_identity_q(const double q[4]) { return q[0] == 1.0 && q[1] == 0.0 && q[2] == 0.0 && q[3] == 0.0; }

M3 maskmerge0(const M3& A, const int mask[9]) {  // merge(0.0, A, mask)
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = mask[i + 3 * j] ? 0.0 : A.a[i][j];
  return r;
}

// Spectral::constitutive_response, spectral.cpp:132-219.  r receives P.
void constitutive_response(sc_rve* R, const double* F, double dt, const double* rot_q, double* r) {
  const int64_t N = R->g.N;
  int ill = 0;
  int64_t newton = 0, fpi = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : ill) reduction(+ : newton, fpi)
  for (int64_t c = 0; c < N; ++c) {
    PointState& s = R->pts[c];
    s.F = load33(F + 9 * c);
    bool ok = mp_response(R->ph, s, dt);
    newton += s.newton_iter;
    fpi += s.state_iter;
    if (!ok) ill = 1;
  }
  R->stats.newton_iterations += newton;
  R->stats.state_iterations += fpi;
  if (ill) R->terminally_ill = true;
  if (!R->terminally_ill) {
    // mechanical_response2 -> mechanical_homogenize (v = 1)
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < N; ++c) {
      PointState& s = R->pts[c];
      store33(s.P, &R->homP[9 * c]);
      T4 d = mp_dPdF(R->ph, s, dt);
      store3333(d, &R->homdPdF[81 * c]);
    }
  }
  for (int64_t i = 0; i < 9 * N; ++i) r[i] = R->homP[i];
  double pav[9] = {0};
  for (int64_t c = 0; c < N; ++c)
    for (int q = 0; q < 9; ++q) pav[q] += R->homP[9 * c + q];
  for (int q = 0; q < 9; ++q) pav[q] *= R->g.wgt;
  R->P_av = load33(pav);
  if (rot_q) R->P_av = rotTensor2(rot_q, R->P_av, false);
  // C_minMaxAvg / C_volAvg
  std::vector<double> n2(N);
  std::vector<double> vol(81, 0.0);
  for (int64_t c = 0; c < N; ++c) {
    const double* d = &R->homdPdF[81 * c];
    double s = 0.0;
    for (int q = 0; q < 81; ++q) s += d[q] * d[q];
    n2[c] = s;
    for (int q = 0; q < 81; ++q) vol[q] += d[q];
  }
  int64_t imax = -1, imin = -1;
  tangent_minmax_select(n2.data(), N, imax, imin);
  double mm[81];
  for (int q = 0; q < 81; ++q) {
    double a = imax >= 0 ? R->homdPdF[81 * imax + q] : 0.0;
    double b = imin >= 0 ? R->homdPdF[81 * imin + q] : std::numeric_limits<double>::max();
    mm[q] = 0.5 * (a + b);
    vol[q] = vol[q] * R->g.wgt;
  }
  R->C_minMaxAvg = load3333(mm);
  R->C_volAvg = load3333(vol.data());
}

// MechBase::forward_field, mech_base.cpp:213-247
void forward_field(int64_t N, double wgt, double dt, const double* last, const double* rate, const M3* aim,
                   double* out) {
  for (int64_t i = 0; i < 9 * N; ++i) out[i] = last[i] + rate[i] * dt;
  if (aim) {
    double s[9] = {0};
    for (int64_t c = 0; c < N; ++c)
      for (int q = 0; q < 9; ++q) s[q] += out[9 * c + q];
    double diff[9];
    for (int q = 0; q < 9; ++q) diff[q] = s[q] * wgt - aim->a[q % 3][q / 3];
    for (int64_t c = 0; c < N; ++c)
      for (int q = 0; q < 9; ++q) out[9 * c + q] -= diff[q];
  }
}

// MechBase::calculate_rate, mech_base.cpp:479-497
void calculate_rate(int64_t N, bool het, const double* f0, const double* f, double dt, const M3& av, double* out) {
  if (het) {
    for (int64_t i = 0; i < 9 * N; ++i) out[i] = (f[i] - f0[i]) / dt;
  } else {
    for (int64_t c = 0; c < N; ++c)
      for (int q = 0; q < 9; ++q) out[9 * c + q] = av.a[q % 3][q / 3];
  }
}

void gamma_convolution(sc_rve* R, double* field, const M3& aim) {
  std::vector<cplx> hat;
  tensor_r2c(R->g, field, hat);
  gamma_apply(R->g, R->C_ref, hat, aim);
  tensor_c2r(R->g, hat, 1.0, field);
}

double divergence(sc_rve* R, const double* field) {
  std::vector<cplx> hat;
  tensor_r2c(R->g, field, hat);
  return divergence_rms(R->g, hat);
}

// MechSolverBasic::form_residual, mech_solver_basic.cpp:286-339
void form_residual(sc_rve* R, const double* F, double* r) {
  if (R->nfuncs == 0 && R->snes_iter == 0) R->total_iter = -1;
  if (R->total_iter <= R->snes_iter) R->total_iter += 1;
  constitutive_response(R, F, R->delta_t, R->rot_q, r);
  R->err_div = divergence(R, r);
  M3 dFaim = mul3333xx33(R->S, R->P_av);
  R->F_aim = sub(R->F_aim, dFaim);
  M3 full = sub(R->P_av, R->P_aim);
  M3 m = maskmerge0(full, R->stress_mask);
  double e = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) e = std::max(e, std::fabs(m.a[i][j]));
  R->err_BC = e;
  // load frame -> field frame: the active rotation mech_solver_basic.cpp:333
  // asks for (its -1 flag reaches rotations.f90:359 as passive; see DESIGN)
  M3 rotated = rotTensor2(R->rot_q, dFaim, true);
  gamma_convolution(R, r, rotated);
  R->nfuncs++;
  R->solve_evals++;
  R->stats.residual_evaluations++;
  R->stats.material_point_updates += R->g.N;
}

// MechSolverBasic::converged, mech_solver_basic.cpp:242-284.
// 1 = converged (breakout), -1 = diverged (max it), 0 = iterating
int converged(sc_rve* R) {
  const auto& gn = R->desc.grid_num;
  double maxP = -std::numeric_limits<double>::infinity();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) maxP = std::max(maxP, R->P_av.a[i][j]);
  double divTol = std::max(maxP * gn.eps_rel_div, gn.eps_abs_div);
  double BCTol = std::max(maxP * gn.eps_rel_P, gn.eps_abs_P);
  if ((R->total_iter >= gn.itmin && std::max(R->err_div / divTol, R->err_BC / BCTol) < 1.0) || R->terminally_ill)
    return 1;
  if (R->total_iter >= gn.itmax) return -1;
  return 0;
}

double vnorm(const std::vector<double>& v) {
  double s = 0.0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}
double vdot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}
double vdiffnorm(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) {
    double d = a[i] - b[i];
    s += d * d;
  }
  return std::sqrt(s);
}

// Minimum-norm least squares of a symmetric l x l system through its
// eigen-decomposition (cyclic Jacobi), singular values below eps*s_max
// dropped -- LAPACK dgelss with rcond = -1 as PETSc calls it.
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
        double apq = H[p * l + q];
        if (apq == 0.0) continue;
        double app = H[p * l + p], aqq = H[q * l + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < l; ++k) {
          double hkp = H[k * l + p], hkq = H[k * l + q];
          H[k * l + p] = c * hkp - s * hkq;
          H[k * l + q] = s * hkp + c * hkq;
        }
        for (int k = 0; k < l; ++k) {
          double hpk = H[p * l + k], hqk = H[q * l + k];
          H[p * l + k] = c * hpk - s * hqk;
          H[q * l + k] = s * hpk + c * hqk;
        }
        for (int k = 0; k < l; ++k) {
          double vkp = V[k * l + p], vkq = V[k * l + q];
          V[k * l + p] = c * vkp - s * vkq;
          V[k * l + q] = s * vkp + c * vkq;
        }
      }
  }
  double smax = 0.0;
  for (int i = 0; i < l; ++i) smax = std::max(smax, std::fabs(H[i * l + i]));
  double thr = std::max(2.220446049250313e-16 * smax, 2.2250738585072014e-308);
  std::vector<double> x(l, 0.0);
  for (int i = 0; i < l; ++i) {
    double lam = H[i * l + i];
    if (std::fabs(lam) <= thr) continue;
    double proj = 0.0;
    for (int k = 0; k < l; ++k) proj += V[k * l + i] * b[k];
    for (int k = 0; k < l; ++k) x[k] += V[k * l + i] * proj / lam;
  }
  b = x;
}

// SNESSolve with SNESNGMRES.  `residual(x, f)` evaluates the function,
// `conv(k)` is the convergence test after iteration k (1 converged, -1
// diverged, 0 iterating); `set_iter(k)` mirrors snes->iter.  Returns the
// SNES converged reason (> 0 converged).
int ngmres_solve(int64_t n, int msize, std::vector<double>& X,
                 const std::function<void(const std::vector<double>&, std::vector<double>&)>& residual,
                 const std::function<int(int)>& conv, const std::function<void(int)>& set_iter) {
  const int max_its = 50;
  const double gammaA = 2.0, gammaC = 2.0, epsilonB = 0.1, deltaB = 0.9;
  const int restart_it = 2;
  std::vector<std::vector<double>> Fdot(msize), Xdot(msize);
  std::vector<double> fnorms(msize, 0.0), Q(msize * msize, 0.0);
  std::vector<double> Fv(n), XM(n), FM(n), XA(n), FA(n);
  set_iter(0);
  residual(X, Fv);
  double fnorm = vnorm(Fv);
  if (std::isnan(fnorm) || std::isinf(fnorm)) return -4;  // SNES_DIVERGED_FNORM_NAN
  double fminnorm = fnorm;
  int reason = conv(0);
  if (reason) return reason > 0 ? 6 : -5;
  auto update_subspace = [&](int ivec, const std::vector<double>& f, double fn, const std::vector<double>& x) {
    Fdot[ivec] = f;
    Xdot[ivec] = x;
    fnorms[ivec] = fn;
  };
  update_subspace(0, Fv, fnorm, X);
  int k_restart = 1, l = 1, ivec = 0, restart_count = 0;
  for (int k = 1; k < max_its + 1; ++k) {
    // x^M: basic line search with damping 1, XM = X - F
    for (int64_t i = 0; i < n; ++i) XM[i] = X[i] - 1.0 * Fv[i];
    residual(XM, FM);
    double fMnorm = vnorm(FM);
    if (std::isnan(fMnorm) || std::isinf(fMnorm)) return -6;  // SNES_DIVERGED_LINE_SEARCH
    // SNESNGMRESFormCombinedSolution_Private
    {
      double nu = fMnorm * fMnorm;
      std::vector<double> xi(l), beta(l);
      for (int i = 0; i < l; ++i) xi[i] = vdot(FM, Fdot[i]);
      for (int i = 0; i < l; ++i) beta[i] = vdot(Fdot[ivec], Fdot[i]);
      for (int i = 0; i < l; ++i) {
        Q[i * msize + ivec] = beta[i];
        Q[ivec * msize + i] = beta[i];
      }
      for (int i = 0; i < l; ++i) beta[i] = nu - xi[i];
      std::vector<double> H(l * l);
      for (int j = 0; j < l; ++j)
        for (int i = 0; i < l; ++i) H[i * l + j] = Q[i * msize + j] - xi[i] - xi[j] + nu;
      if (l == 1) {
        beta[0] = (H[0] != 0.0) ? beta[0] / H[0] : 0.0;
      } else {
        sym_lstsq(l, H, beta);
      }
      double alph_total = 0.0;
      for (int i = 0; i < l; ++i) alph_total += beta[i];
      for (int64_t t = 0; t < n; ++t) {
        double v = XM[t] * (1.0 - alph_total);
        for (int i = 0; i < l; ++i) v += beta[i] * Xdot[i][t];
        XA[t] = v;
      }
      residual(XA, FA);
    }
    if (fminnorm > fMnorm) fminnorm = fMnorm;
    double dnorm = vdiffnorm(XA, XM);
    double dminnorm = -1.0;
    for (int i = 0; i < l; ++i) {
      double d = vdiffnorm(Xdot[i], XA);
      if (d < dminnorm || dminnorm < 0.0) dminnorm = d;
    }
    double fAnorm = vnorm(FA);
    if (std::isnan(fnorm) || std::isinf(fnorm)) return -4;
    // SNESNGMRESSelect_Private, difference selection
    bool selectA = fAnorm < gammaA * fminnorm;
    if (!(epsilonB * dnorm < dminnorm || std::sqrt(fnorm) < deltaB * std::sqrt(fminnorm))) selectA = false;
    if (selectA) {
      fnorm = fAnorm;
      Fv = FA;
      X = XA;
    } else {
      fnorm = fMnorm;
      Fv = FM;
      X = XM;
    }
    // SNESNGMRESSelectRestart_Private, difference restart
    bool selectRestart = false;
    if (epsilonB * dnorm > dminnorm && fAnorm > deltaB * fminnorm && l > 0) selectRestart = true;
    if (fAnorm > gammaC * fminnorm) selectRestart = true;
    restart_count = selectRestart ? restart_count + 1 : 0;
    ivec = k_restart % msize;
    if (restart_count >= restart_it) {
      restart_count = 0;
      k_restart = 1;
      l = 1;
      ivec = 0;
      update_subspace(0, FM, fMnorm, XM);
    } else {
      if (l < msize) l++;
      k_restart++;
      if (fminnorm > fnorm) fminnorm = fnorm;
      update_subspace(ivec, Fv, fnorm, X);
    }
    set_iter(k);
    reason = conv(k);
    if (reason) return reason > 0 ? 6 : -5;
  }
  return -5;  // SNES_DIVERGED_MAX_IT
}

int snes_ngmres(sc_rve* R) {
  const int msize = R->desc.grid_num.ngmres_m > 0 ? R->desc.grid_num.ngmres_m : 30;
  R->snes_iter = 0;
  R->nfuncs = 0;
  R->solve_evals = 0;
  return ngmres_solve(
      9 * R->g.N, msize, R->F,
      [R](const std::vector<double>& x, std::vector<double>& f) { form_residual(R, x.data(), f.data()); },
      [R](int) { return converged(R); }, [R](int k) { R->snes_iter = k; });
}

void base_update_coords(sc_rve* R) {
  const Grid& g = R->g;
  const int Nx = g.cells[0], Ny = g.cells[1], Nz = g.cells[2], nr = g.nr;
  std::vector<cplx> hat;
  tensor_r2c(g, R->F.data(), hat);
  M3 Favg;
  for (int l = 0; l < 3; ++l)
    for (int m = 0; m < 3; ++m) Favg.a[l][m] = hat[l + 3 * m].real() * g.wgt;
  // displacement fluctuations: u_hat = F_hat xi / sum(conj(-xi) xi)  (DAMASK
  // spectral_utilities convention; the C++ port flips the denominator sign)
  std::vector<cplx> uhat(3 * g.K, cplx(0, 0));
  for (int j = 0; j < Ny; ++j)
    for (int k = 0; k < Nz; ++k)
      for (int i = 0; i < nr; ++i) {
        if (i == 0 && j == 0 && k == 0) continue;
        size_t kk = i + (size_t)nr * (k + (size_t)Nz * j);
        const cplx* xi = &g.xi2nd[3 * kk];
        cplx den(0, 0);
        for (int p = 0; p < 3; ++p) den += std::conj(-xi[p]) * xi[p];
        for (int l = 0; l < 3; ++l) {
          cplx s(0, 0);
          for (int m = 0; m < 3; ++m) s += hat[9 * kk + l + 3 * m] * xi[m];
          uhat[3 * kk + l] = s / den;
        }
      }
  std::vector<double> u(3 * g.N);
  std::vector<double> comp(g.N);
  std::vector<cplx> spec(g.K);
  for (int l = 0; l < 3; ++l) {
    for (int64_t kk = 0; kk < g.K; ++kk) spec[kk] = uhat[3 * kk + l];
    fft3d_c2r(g.cells, spec.data(), comp.data());
    for (int64_t c = 0; c < g.N; ++c) u[3 * c + l] = comp[c] * g.wgt;
  }
  auto mod = [](int x, int n) { return (x % n + n) % n; };
  double step[3] = {g.size[0] / Nx, g.size[1] / Ny, g.size[2] / Nz};
  R->x_n.assign((size_t)3 * (Nx + 1) * (Ny + 1) * (Nz + 1), 0.0);
  R->x_p.assign((size_t)3 * g.N, 0.0);
  static const int nb[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int k = 0; k <= Nz; ++k)
    for (int j = 0; j <= Ny; ++j)
      for (int i = 0; i <= Nx; ++i) {
        double pos[3] = {i * step[0], j * step[1], k * step[2]};
        size_t idx = 3 * (i + (size_t)(Nx + 1) * (j + (size_t)(Ny + 1) * k));
        for (int p = 0; p < 3; ++p) {
          double s = 0;
          for (int q = 0; q < 3; ++q) s += Favg.a[p][q] * pos[q];
          R->x_n[idx + p] = s;
        }
        for (int nn = 0; nn < 8; ++nn) {
          // padded z layer me[2] maps to the periodic neighbour cell
          int cx = mod(i + nb[nn][0] - 1, Nx), cy = mod(j + nb[nn][1] - 1, Ny), cz = mod(k + nb[nn][2] - 1, Nz);
          size_t c = cx + (size_t)Nx * (cy + (size_t)Ny * cz);
          for (int p = 0; p < 3; ++p) R->x_n[idx + p] += u[3 * c + p] * 0.125;
        }
      }
  for (int k = 0; k < Nz; ++k)
    for (int j = 0; j < Ny; ++j)
      for (int i = 0; i < Nx; ++i) {
        double pos[3] = {(i + 0.5) * step[0], (j + 0.5) * step[1], (k + 0.5) * step[2]};
        size_t c = i + (size_t)Nx * (j + (size_t)Ny * k);
        for (int p = 0; p < 3; ++p) {
          double s = 0;
          for (int q = 0; q < 3; ++q) s += Favg.a[p][q] * pos[q];
          R->x_p[3 * c + p] = u[3 * c + p] + s;
        }
      }
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* sc_last_error(void) { return orc::g_err.c_str(); }
const char* sc_backend_name(void) { return "cpu-oracle"; }

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
  g->fft_plan_mode = 1;  // no planner here: any mode computes the same
}

int sc_rve_create(const sc_rve_desc* d, sc_rve** out) {
  try {
    sc_rve* R = new sc_rve();
    R->desc = *d;
    int64_t N = (int64_t)d->cells[0] * d->cells[1] * d->cells[2];
    R->material_at.assign(d->material_at, d->material_at + N);
    R->orientations.assign(d->orientations, d->orientations + 4 * d->n_materials);
    R->desc.material_at = R->material_at.data();
    R->desc.orientations = R->orientations.data();
    phase_setup(R->ph, d->material, d->phase_num);
    grid_setup(R->g, d->cells, d->size, d->grid_num.derivative, d->grid_num.divergence_correction);
    R->pts.resize(N);
    // mechanical_init (phase_mechanical.f90:253-262), plastic_phenopowerlaw_init state
    for (int64_t c = 0; c < N; ++c) {
      int ma = R->material_at[c];
      if (ma < 0 || ma >= d->n_materials) throw std::runtime_error("material_at out of range");
      PointState& s = R->pts[c];
      M3 O = qu2om(&R->orientations[4 * ma]);
      s.F = eye3();
      s.Fp = O;
      s.Fe = tr(s.Fp);
      s.Lp = zero3();
      s.S = zero3();
      s.P = zero3();
      s.F0 = s.F;
      s.Fp0 = s.Fp;
      s.Lp0 = zero3();
      s.S0 = zero3();
      for (int i = 0; i < NSLIP; ++i) {
        s.xi[i] = s.xi0[i] = d->material.xi_0_sl;
        s.gamma[i] = s.gamma0[i] = 0.0;
      }
      s.newton_iter = s.state_iter = 0;
    }
    R->homP.assign(9 * N, 0.0);
    R->homdPdF.assign(81 * N, 0.0);
    R->F.assign(9 * N, 0.0);
    R->F_last_inc.assign(9 * N, 0.0);
    R->F_dot.assign(9 * N, 0.0);
    R->C_ref = R->C_volAvg = R->C_minMaxAvg = R->C_volAvgLastInc = R->C_minMaxAvgLastInc = R->S = zero4();
    R->P_av = R->P_aim = R->F_aim_dot = zero3();
    R->F_aim = R->F_aim_last_inc = eye3();
    for (int i = 0; i < 9; ++i) R->stress_mask[i] = 0;
    R->rot_q[0] = 1;
    R->rot_q[1] = R->rot_q[2] = R->rot_q[3] = 0;
    std::memset(&R->stats, 0, sizeof R->stats);
    *out = R;
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

void sc_rve_destroy(sc_rve* R) { delete R; }

int sc_rve_init(sc_rve* R) {
  try {
    auto t0 = std::chrono::steady_clock::now();
    const int64_t N = R->g.N;
    for (int64_t c = 0; c < N; ++c)
      for (int q = 0; q < 9; ++q) R->F_last_inc[9 * c + q] = (q % 4 == 0) ? 1.0 : 0.0;
    R->F = R->F_last_inc;
    base_update_coords(R);
    std::vector<double> P(9 * N);
    constitutive_response(R, R->F.data(), 0.0, nullptr, P.data());
    R->C_ref = R->C_minMaxAvg;  // update_gamma: C_ref = C / wgt
    for (int i = 0; i < 81; ++i) (&R->C_ref.a[0][0][0][0])[i] = (&R->C_minMaxAvg.a[0][0][0][0])[i] / R->g.wgt;
    R->stats.seconds_total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_rve_forward(sc_rve* R, int cut_back, int guess, double dt, double dt_old, double t_remaining,
                   const sc_loadstep* st) {
  try {
    auto t0 = std::chrono::steady_clock::now();
    const int64_t N = R->g.N;
    if (cut_back) {
      R->C_volAvg = R->C_volAvgLastInc;
      R->C_minMaxAvg = R->C_minMaxAvgLastInc;
    } else {
      R->C_volAvgLastInc = R->C_volAvg;
      R->C_minMaxAvgLastInc = R->C_minMaxAvg;
      M3 aim_diff = scal(sub(R->F_aim, R->F_aim_last_inc), 1.0 / dt_old);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) aim_diff.a[i][j] = (R->F_aim.a[i][j] - R->F_aim_last_inc.a[i][j]) / dt_old;
      M3 stress_components = maskmerge0(aim_diff, st->stress_mask);
      R->F_aim_dot = guess ? stress_components : zero3();
      R->F_aim_last_inc = R->F_aim;
      M3 vals = load33(st->deformation);
      M3 step;
      if (st->deformation_type == SC_DEF_DOT_F) {
        step = vals;
      } else {
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) step.a[i][j] = (vals.a[i][j] - R->F_aim_last_inc.a[i][j]) / t_remaining;
      }
      R->F_aim_dot = add(R->F_aim_dot, maskmerge0(step, st->deformation_mask));
      M3 rotated = rotTensor2(st->rot_q, R->F_aim_dot, true);  // load -> field frame, active
      std::vector<double> F0 = R->F;
      calculate_rate(N, guess != 0, R->F_last_inc.data(), F0.data(), dt_old, rotated, R->F_dot.data());
      R->F_last_inc = F0;
      // f_materialpoint_forward (damask_grid.cpp:33)
      for (int64_t c = 0; c < N; ++c) mp_forward(R->pts[c]);
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) R->F_aim.a[i][j] = R->F_aim_last_inc.a[i][j] + R->F_aim_dot.a[i][j] * dt;
    // P_aim update as written in mech_solver_basic.cpp:206-210 (select on the mask)
    for (int q = 0; q < 9; ++q) {
      int i = q % 3, j = q / 3;
      if (st->stress_mask[q]) {
        if (st->stress_type == SC_STRESS_P)
          R->P_aim.a[i][j] += (st->stress[q] - R->P_aim.a[i][j]) / t_remaining * dt;
        else
          R->P_aim.a[i][j] += st->stress[q] * dt;
      }
    }
    M3 rotated = rotTensor2(st->rot_q, R->F_aim, true);  // load -> field frame, active
    std::vector<double> Fn(9 * N);
    forward_field(N, R->g.wgt, dt, R->F_last_inc.data(), R->F_dot.data(), &rotated, Fn.data());
    R->F = Fn;
    for (int q = 0; q < 9; ++q) R->stress_mask[q] = st->stress_mask[q];
    for (int q = 0; q < 4; ++q) R->rot_q[q] = st->rot_q[q];
    R->delta_t = dt;
    R->stats.seconds_total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_rve_solution(sc_rve* R, sc_solution* out) {
  try {
    auto t0 = std::chrono::steady_clock::now();
    R->S = masked_compliance(R->C_volAvg, R->rot_q, R->stress_mask);
    if (R->desc.grid_num.update_gamma)
      for (int i = 0; i < 81; ++i) (&R->C_ref.a[0][0][0][0])[i] = (&R->C_minMaxAvg.a[0][0][0][0])[i] / R->g.wgt;
    int reason = snes_ngmres(R);
    out->converged = reason > 0;
    out->stag_converged = 1;
    out->required_iterations = R->total_iter;
    out->terminally_ill = R->terminally_ill;
    out->residual_evaluations = R->solve_evals;
    out->err_div = R->err_div;
    out->err_BC = R->err_BC;
    R->terminally_ill = false;
    for (int q = 0; q < 9; ++q)
      if (R->stress_mask[q]) R->P_aim.a[q % 3][q / 3] = R->P_av.a[q % 3][q / 3];
    R->stats.seconds_total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_rve_update_coords(sc_rve* R) {
  try {
    base_update_coords(R);
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_rve_run_loadcase(sc_rve* R, const sc_loadstep* steps, int n_steps, int* inc_iterations,
                        int* inc_evaluations, int max_incs) {
  try {
    const auto& gn = R->desc.grid_num;
    bool cut_back = false;
    const int sub_step_factor = 2;
    int cut_back_level = 0, total_inc = 0;
    double t = 0, delta_t = 1, delta_t_prev = 0, t_remaining = 0;
    for (int l = 0; l < n_steps; ++l) {
      const sc_loadstep& st = steps[l];
      double t_0 = t;
      bool guess = st.estimate_rate && l > 0;
      for (int inc = 0; inc < st.N; ++inc) {
        total_inc++;
        // forward_time, damask_grid.cpp:69-80
        delta_t_prev = delta_t;
        if (std::fabs(st.r - 1.0) < 1.e-9)
          delta_t = st.t / st.N;
        else
          delta_t = st.t * (std::pow(st.r, inc) - std::pow(st.r, inc + 1)) / (1.0 - std::pow(st.r, st.N));
        delta_t *= std::pow((double)sub_step_factor, -(double)cut_back_level);
        int step_fraction = 0;
        int iters = 0, evals = 0;
        sc_solution sol{};
        while (step_fraction < std::pow(sub_step_factor, cut_back_level)) {
          t_remaining = st.t + t_0 - t;
          t += delta_t;
          step_fraction++;
          if (sc_rve_forward(R, cut_back, guess, delta_t, delta_t_prev, t_remaining, &st)) return 1;
          if (sc_rve_solution(R, &sol)) return 1;
          iters += sol.required_iterations;
          evals += sol.residual_evaluations;
          bool found = sol.converged && sol.stag_converged && !sol.terminally_ill;
          if (found) {
            base_update_coords(R);
            delta_t_prev = delta_t;
            cut_back = false;
            guess = true;
          } else if (cut_back_level < gn.max_cut_back) {
            cut_back = true;
            step_fraction = (step_fraction - 1) * sub_step_factor;
            cut_back_level++;
            t = t - delta_t;
            delta_t = delta_t / sub_step_factor;
            R->stats.cutbacks++;
          } else {
            g_err = "no solution found for increment " + std::to_string(total_inc);
            return 2;
          }
        }
        // damask_grid.cpp:153 -- max(0, cut_back - 1) with a bool
        cut_back_level = std::max(0, (int)cut_back - 1);
        R->stats.increments++;
        if (inc_iterations && total_inc <= max_incs) inc_iterations[total_inc - 1] = iters;
        if (inc_evaluations && total_inc <= max_incs) inc_evaluations[total_inc - 1] = evals;
      }
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_ngmres_replay(int64_t n, int m, const double* x0, int n_evals, const double* r_rec, double* x_eval) {
  try {
    std::vector<double> X(x0, x0 + n);
    int e = 0;
    ngmres_solve(
        n, m, X,
        [&](const std::vector<double>& x, std::vector<double>& f) {
          if (e >= n_evals) throw std::runtime_error("replay ran past the recorded evaluations");
          std::memcpy(x_eval + (int64_t)e * n, x.data(), n * sizeof(double));
          std::memcpy(f.data(), r_rec + (int64_t)e * n, n * sizeof(double));
          e++;
        },
        [&](int) { return e >= n_evals ? 1 : 0; }, [](int) {});
    return e == n_evals ? 0 : 3;
  } catch (std::exception& ex) {
    g_err = ex.what();
    return 1;
  }
}

int sc_rve_form_residual(sc_rve* R, const double* F, double* r) {
  try {
    form_residual(R, F, r);
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_rve_get_field(sc_rve* R, int field, double* out, int64_t n) {
  const int64_t N = R->g.N;
  auto need = [&](int64_t k) {
    if (n < k) throw std::runtime_error("output buffer too small");
  };
  try {
    switch (field) {
      case SC_FIELD_F:
        need(9 * N);
        std::memcpy(out, R->F.data(), 9 * N * sizeof(double));
        break;
      case SC_FIELD_P:
        need(9 * N);
        std::memcpy(out, R->homP.data(), 9 * N * sizeof(double));
        break;
      case SC_FIELD_F_LAST_INC:
        need(9 * N);
        std::memcpy(out, R->F_last_inc.data(), 9 * N * sizeof(double));
        break;
      case SC_FIELD_FP:
      case SC_FIELD_LP:
      case SC_FIELD_S:
      case SC_FIELD_FE:
        need(9 * N);
        for (int64_t c = 0; c < N; ++c) {
          const PointState& s = R->pts[c];
          const M3& m = field == SC_FIELD_FP ? s.Fp : field == SC_FIELD_LP ? s.Lp : field == SC_FIELD_S ? s.S : s.Fe;
          store33(m, out + 9 * c);
        }
        break;
      case SC_FIELD_XI:
      case SC_FIELD_GAMMA:
        need(12 * N);
        for (int64_t c = 0; c < N; ++c)
          for (int i = 0; i < 12; ++i) out[12 * c + i] = field == SC_FIELD_XI ? R->pts[c].xi[i] : R->pts[c].gamma[i];
        break;
      case SC_FIELD_ITERATIONS:
        need(2 * N);
        for (int64_t c = 0; c < N; ++c) {
          out[2 * c] = R->pts[c].newton_iter;
          out[2 * c + 1] = R->pts[c].state_iter;
        }
        break;
      case SC_FIELD_X_P:
        need((int64_t)R->x_p.size());
        std::memcpy(out, R->x_p.data(), R->x_p.size() * sizeof(double));
        break;
      case SC_FIELD_X_N:
        need((int64_t)R->x_n.size());
        std::memcpy(out, R->x_n.data(), R->x_n.size() * sizeof(double));
        break;
      default:
        throw std::runtime_error("unknown field");
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_rve_set_field(sc_rve* R, int field, const double* in, int64_t n) {
  const int64_t N = R->g.N;
  try {
    if (field == SC_FIELD_F) {
      if (n < 9 * N) throw std::runtime_error("input too small");
      std::memcpy(R->F.data(), in, 9 * N * sizeof(double));
      return 0;
    }
    throw std::runtime_error("field not settable");
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_rve_get_scalars(sc_rve* R, sc_solver_scalars* o) {
  store33(R->P_av, o->P_av);
  store33(R->P_aim, o->P_aim);
  store33(R->F_aim, o->F_aim);
  store33(R->F_aim_dot, o->F_aim_dot);
  store33(R->F_aim_last_inc, o->F_aim_last_inc);
  store3333(R->C_volAvg, o->C_volAvg);
  store3333(R->C_minMaxAvg, o->C_minMaxAvg);
  store3333(R->S, o->S);
  store3333(R->C_ref, o->C_ref);
  o->err_div = R->err_div;
  o->err_BC = R->err_BC;
  o->total_iter = R->total_iter;
  return 0;
}
int sc_rve_get_stats(sc_rve* R, sc_stats* o) {
  *o = R->stats;
  return 0;
}
int sc_rve_reset_stats(sc_rve* R) {
  std::memset(&R->stats, 0, sizeof R->stats);
  return 0;
}

}  // extern "C"
