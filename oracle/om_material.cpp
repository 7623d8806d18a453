// ORACLE -- test infrastructure only (see oracle_internal.h).
// cF lattice tables and phase setup: src/crystal.f90 (CF_SYSTEMSLIP :41-63,
// crystal_SchmidMatrix_slip :1353-1431, crystal_interaction_SlipBySlip
// :602-913, buildInteraction :1853-1890, buildCoordinateSystem :1897-1961,
// crystal_symmetrize_C66 :1680-1719), phase_mechanical_elastic.f90:74-104 and
// phase_mechanical_plastic_phenopowerlaw.f90:137-170.
#include "oracle_internal.h"

namespace orc {

// <110>{111} systems: direction (3), plane normal (3); crystal.f90:41-63
static const int CF_SLIP[12][6] = {
    {0, 1, -1, 1, 1, 1},   {-1, 0, 1, 1, 1, 1},   {1, -1, 0, 1, 1, 1},
    {0, -1, -1, -1, -1, 1}, {1, 0, 1, -1, -1, 1},  {-1, 1, 0, -1, -1, 1},
    {0, -1, 1, 1, -1, -1}, {-1, 0, -1, 1, -1, -1}, {1, 1, 0, 1, -1, -1},
    {0, 1, 1, -1, 1, -1},  {1, 0, -1, -1, 1, -1},  {-1, -1, 0, -1, 1, -1}};

// Slip-slip interaction types for the 12 octahedral systems, row = reacting,
// column = acting (crystal.f90:623-640 upper-left block, as laid out there).
static const int CF_INTERACTION[12][12] = {
    {1, 2, 2, 4, 7, 5, 3, 5, 5, 4, 6, 7}, {2, 1, 2, 7, 4, 5, 6, 4, 7, 5, 3, 5},
    {2, 2, 1, 5, 5, 3, 6, 7, 4, 7, 6, 4}, {4, 7, 6, 1, 2, 2, 4, 6, 7, 3, 5, 5},
    {7, 4, 6, 2, 1, 2, 5, 3, 5, 6, 4, 7}, {5, 5, 3, 2, 2, 1, 7, 6, 4, 6, 7, 4},
    {3, 5, 5, 4, 6, 7, 1, 2, 2, 4, 7, 6}, {6, 4, 7, 5, 3, 5, 2, 1, 2, 7, 4, 6},
    {6, 7, 4, 7, 6, 4, 2, 2, 1, 5, 5, 3}, {4, 6, 7, 3, 5, 5, 4, 7, 6, 1, 2, 2},
    {5, 3, 5, 6, 4, 7, 7, 4, 6, 2, 1, 2}, {7, 6, 4, 6, 7, 4, 5, 5, 3, 2, 2, 1}};

void phase_setup(Phase& ph, const sc_material& m, const sc_phase_numerics& num) {
  ph.prm = m;
  ph.num = num;
  // elastic_C66 + crystal_symmetrize_C66 (cF)
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) ph.C66[i][j] = 0.0;
  ph.C66[0][0] = ph.C66[1][1] = ph.C66[2][2] = m.C_11;
  ph.C66[0][1] = ph.C66[0][2] = ph.C66[1][2] = m.C_12;
  ph.C66[3][3] = ph.C66[4][4] = ph.C66[5][5] = m.C_44;
  for (int i = 0; i < 6; ++i)
    for (int j = i + 1; j < 6; ++j) ph.C66[j][i] = ph.C66[i][j];
  ph.C = voigt66to3333_stiffness(ph.C66);

  // buildCoordinateSystem + SchmidMatrix: P = d/|d| (x) n/|n|
  for (int s = 0; s < NSLIP; ++s) {
    double d[3], n[3], nd = 0, nn = 0;
    for (int k = 0; k < 3; ++k) {
      d[k] = CF_SLIP[s][k];
      n[k] = CF_SLIP[s][3 + k];
      nd += d[k] * d[k];
      nn += n[k] * n[k];
    }
    nd = std::sqrt(nd);
    nn = std::sqrt(nn);
    for (int k = 0; k < 3; ++k) {
      d[k] = d[k] / nd;
      n[k] = n[k] / nn;
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) ph.P_sl[s].a[i][j] = d[i] * n[j];
    // cF: P_nS_pos = +P_sl, P_nS_neg = -P_sl (phenopowerlaw.f90:166-168)
    ph.P_nS_pos[s] = ph.P_sl[s];
    ph.P_nS_neg[s] = scal(ph.P_sl[s], -1.0);
  }
  // buildInteraction(l=reacting, k=acting) = values(type(reacting, acting))
  for (int l = 0; l < NSLIP; ++l)
    for (int k = 0; k < NSLIP; ++k) ph.h_sl_sl[l][k] = m.h_sl_sl[CF_INTERACTION[l][k] - 1];
}

}  // namespace orc
