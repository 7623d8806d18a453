// ORACLE -- test infrastructure only (see oracle_internal.h).
// Material-point update restated from
//   src/phase_mechanical.f90            integrateStress :380-604, integrateStateFPI :611-693,
//                                       phase_mechanical_constitutive :1013-1100,
//                                       phase_mechanical_dPdF :1137-1258, mechanical_forward :992-1007
//   src/phase_mechanical_elastic.f90    phase_hooke_SandItsTangents :157-195
//   src/phase_mechanical_plastic.f90    plastic_LpAndItsTangents :233-292, plastic_dotState :298-339
//   src/phase_mechanical_plastic_phenopowerlaw.f90
//                                       LpAndItsTangent :298-345, dotState :351-394, kinetics_sl :444-483
//   src/phase.f90                       converged :626-637, phase_restore :490-499
//   src/homogenization.f90              homogenization_mechanical_response :219-266
// Fi = I and Li = 0 hold throughout: no eigen kinematics are active
// (phase_mechanical_eigen.f90:161 returns Li = 0) and material_V_e_0 = I.
#include <cfloat>
#include <cstring>

#include "oracle_internal.h"

namespace orc {

namespace {

constexpr int NSTATE = 2 * NSLIP;  // xi_sl(1:12), gamma_sl(13:24)

// Frobenius double contraction in Fortran element order (math_tensordot)
double tensordot(const M3& A, const M3& B) {
  double s = 0.0;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) s += A.a[i][j] * B.a[i][j];
  return s;
}

// kinetics_sl, phenopowerlaw.f90:444-483
void kinetics_sl(const Phase& ph, const M3& Mp, const double* xi, double* dot_gamma,
                 double* ddot_gamma_dtau) {
  const auto& p = ph.prm;
  for (int i = 0; i < NSLIP; ++i) {
    double tp = tensordot(Mp, ph.P_nS_pos[i]);
    double tn = tensordot(Mp, ph.P_nS_neg[i]);
    double sgn = (tp > tn) ? 1.0 : -1.0;
    double tmax = std::max(tp, tn);
    dot_gamma[i] = sgn * p.dot_gamma_0_sl * std::pow(tmax / xi[i], p.n_sl);
    if (ddot_gamma_dtau) ddot_gamma_dtau[i] = dNeq0(dot_gamma[i]) ? dot_gamma[i] * p.n_sl / tmax : 0.0;
  }
}

// phenopowerlaw_dotState, phenopowerlaw.f90:351-394 (no twins: sumF = 0)
void dot_state(const Phase& ph, const M3& S, const double* state, double* dot) {
  const auto& p = ph.prm;
  const double* xi = state;
  M3 Mp = S;  // matmul(matmul(transpose(Fi),Fi),S) with Fi = I
  double dg[NSLIP];
  kinetics_sl(ph, Mp, xi, dg, nullptr);
  for (int i = 0; i < NSLIP; ++i) dg[i] = std::fabs(dg[i]);
  const double c = 1.0 + 0.0 * std::pow(0.0, 1.0);  // 1 + c_1*sumF**c_2 with c_1 = 0
  for (int i = 0; i < NSLIP; ++i) {
    double x = 1.0 - xi[i] / (p.xi_inf_sl + 0.0);
    double left = std::copysign(std::pow(std::fabs(x), p.a_sl), x);
    double hg = 0.0;
    for (int k = 0; k < NSLIP; ++k) hg += ph.h_sl_sl[i][k] * dg[k];
    dot[i] = p.h_0_sl_sl * c * left * hg;
  }
  for (int i = 0; i < NSLIP; ++i) dot[NSLIP + i] = dg[i];
}

// phase_hooke_SandItsTangents, elastic.f90:157-195
void hooke(const Phase& ph, const M3& Fe, const M3& Fi, M3& S, T4* dS_dFe, T4* dS_dFi) {
  M3 E = scal(sub(mul(tr(Fe), Fe), eye3()), 0.5);
  double v[6], sv[6];
  m33toVoigt6_strain(mul(mul(tr(Fi), E), Fi), v);
  for (int i = 0; i < 6; ++i) {
    double s = 0.0;
    for (int j = 0; j < 6; ++j) s += ph.C66[i][j] * v[j];
    sv[i] = s;
  }
  S = voigt6to33_stress(sv);
  if (!dS_dFe && !dS_dFi) return;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      M3 Cij;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) Cij.a[k][l] = ph.C.a[i][j][k][l];
      if (dS_dFe) {
        M3 t = mul(Fe, mul(mul(Fi, Cij), tr(Fi)));
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) dS_dFe->a[i][j][k][l] = t.a[k][l];
      }
      if (dS_dFi) {
        M3 t = scal(mul(mul(E, Fi), Cij), 2.0);
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) dS_dFi->a[i][j][k][l] = t.a[k][l];
      }
    }
}

// plastic_LpAndItsTangents (plastic.f90:233-292) with phenopowerlaw_LpAndItsTangent
void lp_and_tangents(const Phase& ph, const M3& S, const M3& Fi, const double* xi, M3& Lp,
                     T4* dLp_dS, T4* dLp_dFi) {
  M3 FtF = mul(tr(Fi), Fi);
  M3 Mp = mul(FtF, S);
  double dg[NSLIP], ddg[NSLIP];
  kinetics_sl(ph, Mp, xi, dg, ddg);
  Lp = zero3();
  T4 dLp_dMp = zero4();
  for (int s = 0; s < NSLIP; ++s) {
    const M3& Pns = (dg[s] > 0.0) ? ph.P_nS_pos[s] : ph.P_nS_neg[s];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Lp.a[i][j] = Lp.a[i][j] + dg[s] * ph.P_sl[s].a[i][j];
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l)
        for (int m = 0; m < 3; ++m)
          for (int n = 0; n < 3; ++n)
            dLp_dMp.a[k][l][m][n] = dLp_dMp.a[k][l][m][n] + ddg[s] * ph.P_sl[s].a[k][l] * Pns.a[m][n];
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      M3 D;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) D.a[k][l] = dLp_dMp.a[i][j][k][l];
      if (dLp_dFi) {
        M3 t = add(mul(mul(Fi, S), tr(D)), mul(mul(Fi, D), S));
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) dLp_dFi->a[i][j][k][l] = t.a[k][l];
      }
      if (dLp_dS) {
        M3 t = mul(FtF, D);
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) dLp_dS->a[i][j][k][l] = t.a[k][l];
      }
    }
}

bool any_nan(const M3& m) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (std::isnan(m.a[i][j])) return true;
  return false;
}

// integrateStress, phase_mechanical.f90:380-604.  Returns broken.
bool integrate_stress(const Phase& ph, PointState& s, const M3& F, const M3& Fp0, double dt) {
  const auto& num = ph.num;
  const M3 I = eye3();
  M3 Lpguess = s.Lp;
  M3 invFp_current;
  if (!invert33(Fp0, invFp_current)) return true;
  M3 invFi_current = I;  // Fi0 = I
  M3 A = mul(F, invFp_current);

  // LiLoop, first pass (Liguess = 0)
  M3 invFi_new = mul(invFi_current, sub(I, scal(zero3(), dt)));
  M3 Fi_new = inv33(invFi_new);

  int jacoCounterLp = 0;
  double steplengthLp = 1.0;
  M3 residuumLp_old = zero3();
  M3 Lpguess_old = Lpguess;
  M3 deltaLp = zero3();
  M3 B, Fe, S;
  int NiterationStressLp = 0;
  for (;;) {
    NiterationStressLp++;
    if (NiterationStressLp > num.N_iter_Lp_max) return true;
    B = sub(I, scal(Lpguess, dt));
    Fe = mul(mul(A, B), invFi_new);
    T4 dS_dFe;
    hooke(ph, Fe, Fi_new, S, &dS_dFe, nullptr);
    M3 Lp_c;
    T4 dLp_dS;
    lp_and_tangents(ph, S, Fi_new, s.xi, Lp_c, &dLp_dS, nullptr);

    double atol_Lp = std::max(num.eps_rel_Lp * std::max(norm2(Lpguess), norm2(Lp_c)), num.eps_abs_Lp);
    M3 residuumLp = sub(Lpguess, Lp_c);
    if (any_nan(residuumLp)) return true;
    if (norm2(residuumLp) < atol_Lp) break;
    if (NiterationStressLp == 1 || norm2(residuumLp) < norm2(residuumLp_old)) {
      residuumLp_old = residuumLp;
      Lpguess_old = Lpguess;
      steplengthLp = 1.0;
    } else {
      steplengthLp = num.r_linesearch_Lp * steplengthLp;
      Lpguess = add(Lpguess_old, scal(deltaLp, steplengthLp));
      continue;
    }
    // calculateJacobiLp (f_update_jacobi_Lp = 1: every iteration)
    jacoCounterLp++;
    T4 dFe_dLp;
    for (int o = 0; o < 3; ++o)
      for (int p = 0; p < 3; ++p)
        for (int j = 0; j < 3; ++j)
          for (int l = 0; l < 3; ++l) dFe_dLp.a[o][j][p][l] = -dt * A.a[o][p] * invFi_new.a[l][j];
    M99 J = t3333to99(mul3333xx3333(mul3333xx3333(dLp_dS, dS_dFe), dFe_dLp));
    M99 E9 = eye99();
    double Jm[81];
    for (int a = 0; a < 9; ++a)
      for (int b = 0; b < 9; ++b) Jm[a * 9 + b] = E9.a[a][b] - J.a[a][b];
    std::array<double, 9> t9 = m33to9(residuumLp);
    if (gesv(9, Jm, t9.data()) != 0) return true;
    deltaLp = scal(m9to33(t9), -1.0);
    Lpguess = add(Lpguess, scal(deltaLp, steplengthLp));
  }
  s.newton_iter += NiterationStressLp;
  // LiLoop: Li_constitutive = 0, residuumLi = 0 -> converged at once.

  M3 invFp_new = mul(invFp_current, B);
  M3 Fp_new;
  if (!invert33(invFp_new, Fp_new)) return true;
  s.P = mul(mul(F, invFp_new), mul(S, tr(invFp_new)));
  s.S = S;
  s.Lp = Lpguess;
  double d = std::pow(det33(Fp_new), 1.0 / 3.0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) s.Fp.a[i][j] = Fp_new.a[i][j] / d;
  s.Fe = mul(mul(F, invFp_new), invFi_new);
  return false;
}

// damper, phase_mechanical.f90:674-691
double damper(const double* w0, const double* w1, const double* w2) {
  double dp12 = 0.0, dp22 = 0.0, dp01 = 0.0;
  for (int i = 0; i < NSTATE; ++i) {
    dp12 += (w0[i] - w1[i]) * (w1[i] - w2[i]);
    dp22 += (w1[i] - w2[i]) * (w1[i] - w2[i]);
    dp01 += w0[i] * w1[i];
  }
  if (std::min(dp01, dp12) < 0.0 && dp22 > 0.0) return 0.75 + 0.25 * std::tanh(2.0 + 4.0 * dp12 / dp22);
  return 1.0;
}

// integrateStateFPI, phase_mechanical.f90:611-693.  Returns broken.
bool integrate_state_fpi(const Phase& ph, PointState& s, const M3& F, const M3& Fp0,
                         const double* state0, double dt) {
  const auto& num = ph.num;
  double state[NSTATE], dotState[NSTATE], last1[NSTATE], last2[NSTATE], r[NSTATE], atol[NSTATE];
  for (int i = 0; i < NSLIP; ++i) {
    state[i] = s.xi[i];
    state[NSLIP + i] = s.gamma[i];
    atol[i] = ph.prm.atol_xi;
    atol[NSLIP + i] = ph.prm.atol_gamma;
  }
  auto store_state = [&]() {
    for (int i = 0; i < NSLIP; ++i) {
      s.xi[i] = state[i];
      s.gamma[i] = state[NSLIP + i];
    }
  };
  bool broken = true;
  dot_state(ph, s.S, state, dotState);
  for (int i = 0; i < NSTATE; ++i)
    if (std::isnan(dotState[i])) return true;
  for (int i = 0; i < NSTATE; ++i) state[i] = state0[i] + dotState[i] * dt;
  store_state();

  for (int it = 1; it <= num.N_iter_state_max; ++it) {
    s.state_iter++;
    for (int i = 0; i < NSTATE; ++i) {
      last2[i] = (it > 1) ? last1[i] : 0.0;
      last1[i] = dotState[i];
    }
    broken = integrate_stress(ph, s, F, Fp0, dt);
    if (broken) break;
    dot_state(ph, s.S, state, dotState);
    bool nan = false;
    for (int i = 0; i < NSTATE; ++i) nan = nan || std::isnan(dotState[i]);
    if (nan) break;
    double zeta = damper(dotState, last1, last2);
    for (int i = 0; i < NSTATE; ++i) dotState[i] = dotState[i] * zeta + last1[i] * (1.0 - zeta);
    for (int i = 0; i < NSTATE; ++i) r[i] = state[i] - state0[i] - dotState[i] * dt;
    for (int i = 0; i < NSTATE; ++i) state[i] = state[i] - r[i];
    store_state();
    bool conv = true;
    for (int i = 0; i < NSTATE; ++i)
      conv = conv && (std::fabs(r[i]) <= std::max(atol[i], num.eps_rel_state * std::fabs(state[i])));
    if (conv) {
      broken = false;  // plastic_deltaState: nothing to do for phenopowerlaw
      break;
    }
  }
  return broken;
}

}  // namespace

// homogenization_mechanical_response inner body for one cell (pass, one
// constituent): phase_restore(ce,.false.), mechanical_partition,
// phase_mechanical_constitutive (phase_mechanical.f90:1013-1100).
bool mp_response(const Phase& ph, PointState& s, double dt) {
  const auto& num = ph.num;
  s.newton_iter = 0;
  s.state_iter = 0;
  // mechanical_restore(ce, includeL=.false.)
  s.Fp = s.Fp0;
  s.S = s.S0;
  for (int i = 0; i < NSLIP; ++i) {
    s.xi[i] = s.xi0[i];
    s.gamma[i] = s.gamma0[i];
  }
  double state0[NSTATE];
  for (int i = 0; i < NSLIP; ++i) {
    state0[i] = s.xi0[i];
    state0[NSLIP + i] = s.gamma0[i];
  }
  M3 Lp0 = s.Lp0, Fp0 = s.Fp0, F0 = s.F0, F = s.F0;
  double stepFrac = 0.0, step = 1.0 / num.r_cutback;
  bool converged = false, todo = true;
  while (todo) {
    if (converged) {
      stepFrac = stepFrac + step;
      step = std::min(1.0 - stepFrac, num.r_increase * step);
      todo = step > 0.0;
      if (todo) {
        F0 = F;
        Lp0 = s.Lp;
        Fp0 = s.Fp;
        for (int i = 0; i < NSLIP; ++i) {
          state0[i] = s.xi[i];
          state0[NSLIP + i] = s.gamma[i];
        }
      }
    } else {
      step = num.r_cutback * step;
      s.Fp = Fp0;
      s.S = s.S0;
      if (step < 1.0) s.Lp = Lp0;
      for (int i = 0; i < NSLIP; ++i) {
        s.xi[i] = state0[i];
        s.gamma[i] = state0[NSLIP + i];
      }
      todo = step > num.r_cutback_min;
    }
    if (todo) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) F.a[i][j] = F0.a[i][j] + step * (s.F.a[i][j] - s.F0.a[i][j]);
      converged = !integrate_state_fpi(ph, s, F, Fp0, state0, step * dt);
    }
  }
  return converged;
}

// phase_mechanical_dPdF, phase_mechanical.f90:1137-1258
T4 mp_dPdF(const Phase& ph, const PointState& s, double dt) {
  const M3 Fi = eye3(), Fi0 = eye3();
  M3 devNull;
  T4 dSdFe, dSdFi;
  hooke(ph, s.Fe, Fi, devNull, &dSdFe, &dSdFi);
  T4 dLidS = zero4(), dLidFi = zero4();  // phase_LiAndItsTangents: nothing active
  M3 invFp = inv33(s.Fp), invFi = inv33(Fi), invSubFp0 = inv33(s.Fp0), invSubFi0 = inv33(Fi0);
  T4 dFidS = zero4();  // sum(abs(dLidS)) < tol_math_check
  (void)dLidFi;
  M3 Lpd;
  T4 dLpdS, dLpdFi;
  lp_and_tangents(ph, s.S, Fi, s.xi, Lpd, &dLpdS, &dLpdFi);
  {
    T4 t = mul3333xx3333(dLpdFi, dFidS);
    for (int i = 0; i < 81; ++i) (&dLpdS.a[0][0][0][0])[i] = (&t.a[0][0][0][0])[i] + (&dLpdS.a[0][0][0][0])[i];
  }
  // dSdF
  M3 t1 = tr(mul(invFp, invFi));
  M3 t2 = mul(s.F, invSubFp0);
  M3 t3 = mul(mul(s.F, invFp), invSubFi0);
  T4 rhs, tmp;
  for (int o = 0; o < 3; ++o)
    for (int p = 0; p < 3; ++p) {
      M3 dS_po, dLp_po, dLi_po;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          dS_po.a[k][l] = dSdFe.a[p][o][k][l];
          dLp_po.a[k][l] = dLpdS.a[k][l][p][o];
          dLi_po.a[k][l] = dLidS.a[k][l][p][o];
        }
      M3 r = mul(dS_po, t1);
      M3 t = add(mul(mul(t2, dLp_po), invFi), mul(t3, dLi_po));
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          rhs.a[p][o][k][l] = r.a[k][l];
          tmp.a[k][l][p][o] = t.a[k][l];
        }
    }
  T4 lhs = mul3333xx3333(dSdFe, tmp);
  T4 l2 = mul3333xx3333(dSdFi, dFidS);
  for (int i = 0; i < 81; ++i)
    (&lhs.a[0][0][0][0])[i] = (&lhs.a[0][0][0][0])[i] * dt + (&l2.a[0][0][0][0])[i];
  M99 L9 = t3333to99(lhs);
  double Am[81], Ainv[81];
  for (int a = 0; a < 9; ++a)
    for (int b = 0; b < 9; ++b) Am[a * 9 + b] = (a == b ? 1.0 : 0.0) + L9.a[a][b];
  T4 dSdF;
  if (invert(9, Am, Ainv)) {
    dSdF = rhs;
  } else {
    M99 inv9;
    for (int a = 0; a < 9; ++a)
      for (int b = 0; b < 9; ++b) inv9.a[a][b] = Ainv[a * 9 + b];
    dSdF = mul3333xx3333(t99to3333(inv9), rhs);
  }
  // dFpinvdF
  T4 t4 = mul3333xx3333(dLpdS, dSdF);
  T4 dFpinvdF;
  for (int o = 0; o < 3; ++o)
    for (int p = 0; p < 3; ++p) {
      M3 x;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) x.a[k][l] = t4.a[k][l][p][o];
      M3 y = scal(mul(invSubFp0, mul(x, invFi)), -1.0);
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) dFpinvdF.a[k][l][p][o] = y.a[k][l] * dt;
    }
  // assemble dPdF
  M3 u1 = mul(s.S, tr(invFp));
  M3 u2 = mul(s.F, invFp);
  M3 u3 = mul(u2, s.S);
  T4 dPdF = zero4();
  M3 d0 = tr(mul(invFp, u1));
  for (int p = 0; p < 3; ++p)
    for (int j = 0; j < 3; ++j)
      for (int l = 0; l < 3; ++l) dPdF.a[p][j][p][l] = d0.a[j][l];
  for (int o = 0; o < 3; ++o)
    for (int p = 0; p < 3; ++p) {
      M3 dFp, dS;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          dFp.a[k][l] = dFpinvdF.a[k][l][p][o];
          dS.a[k][l] = dSdF.a[k][l][p][o];
        }
      M3 a1 = mul(mul(s.F, dFp), u1);
      M3 a2 = mul(mul(u2, dS), tr(invFp));
      M3 a3 = mul(u3, tr(dFp));
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
          dPdF.a[k][l][p][o] = dPdF.a[k][l][p][o] + a1.a[k][l] + a2.a[k][l] + a3.a[k][l];
    }
  return dPdF;
}

// mechanical_forward, phase_mechanical.f90:992-1007
void mp_forward(PointState& s) {
  s.F0 = s.F;
  s.Fp0 = s.Fp;
  s.Lp0 = s.Lp;
  s.S0 = s.S;
  for (int i = 0; i < NSLIP; ++i) {
    s.xi0[i] = s.xi[i];
    s.gamma0[i] = s.gamma[i];
  }
}

}  // namespace orc
