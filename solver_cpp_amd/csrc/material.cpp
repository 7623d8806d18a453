// Host-side setup of the material constants used by the kernels
// (phase_mechanical_elastic.f90:74-104, phenopowerlaw.f90:137-170,
// crystal.f90:1353-1431, rotations.f90:423-447).
#include <cmath>

#include "common.h"

namespace sca {

namespace {
// e4 = 4*e when e is a multiple of 1/4 in [0, 64], else -1
int quarter(double e) {
  const double e4 = e * 4.0;
  if (e4 < 0.0 || e4 > 256.0) return -1;
  const double r = std::nearbyint(e4);
  return (r == e4) ? (int)r : -1;
}
}  // namespace

void build_matconst(const sc_material& m, const sc_phase_numerics& p, MatConst& o) {
  o.C11 = m.C_11;
  o.C12 = m.C_12;
  o.C44 = m.C_44;
  // P_s = (d/|d|) (x) (n/|n|); for <110>{111} every nonzero entry has the
  // magnitude (1/sqrt 2)(1/sqrt 3), evaluated as the oracle does
  o.v = (1.0 / std::sqrt(2.0)) * (1.0 / std::sqrt(3.0));
  for (int i = 0; i < 7; ++i) o.hv[i] = m.h_sl_sl[i];
  o.xi_inf = m.xi_inf_sl;
  o.h0 = m.h_0_sl_sl;
  o.n = m.n_sl;
  o.a = m.a_sl;
  o.gdot0 = m.dot_gamma_0_sl;
  o.atol_xi = m.atol_xi;
  o.atol_gamma = m.atol_gamma;
  o.r_cutback_min = p.r_cutback_min;
  o.r_cutback = p.r_cutback;
  o.r_increase = p.r_increase;
  o.eps_rel_state = p.eps_rel_state;
  o.N_iter_state_max = p.N_iter_state_max;
  o.N_iter_Lp_max = p.N_iter_Lp_max;
  o.r_linesearch_Lp = p.r_linesearch_Lp;
  o.eps_rel_Lp = p.eps_rel_Lp;
  o.eps_abs_Lp = p.eps_abs_Lp;
  o.n_e4 = quarter(m.n_sl);
  o.a_e4 = quarter(m.a_sl);
}

// qu2om with P = -1 (rotations.f90:423-447), (3,3) column-major out
void qu2om_host(const double q[4], double om[9]) {
  double r[3][3];
  const double qq = q[0] * q[0] - (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  r[0][0] = qq + 2.0 * q[1] * q[1];
  r[1][1] = qq + 2.0 * q[2] * q[2];
  r[2][2] = qq + 2.0 * q[3] * q[3];
  r[0][1] = 2.0 * (q[1] * q[2] - q[0] * q[3]);
  r[1][2] = 2.0 * (q[2] * q[3] - q[0] * q[1]);
  r[2][0] = 2.0 * (q[3] * q[1] - q[0] * q[2]);
  r[1][0] = 2.0 * (q[2] * q[1] + q[0] * q[3]);
  r[2][1] = 2.0 * (q[3] * q[2] + q[0] * q[1]);
  r[0][2] = 2.0 * (q[1] * q[3] + q[0] * q[2]);
  double t[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[i][j] = r[j][i];  // P < 0: transpose
  const double det = t[0][0] * (t[1][1] * t[2][2] - t[1][2] * t[2][1]) -
                     t[0][1] * (t[1][0] * t[2][2] - t[1][2] * t[2][0]) +
                     t[0][2] * (t[1][0] * t[2][1] - t[1][1] * t[2][0]);
  const double d = std::pow(det, 1.0 / 3.0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) om[i + 3 * j] = t[i][j] / d;
}

}  // namespace sca
