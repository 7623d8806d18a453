// ORACLE -- test infrastructure only (see oracle_internal.h).
// Spectral utilities restated from src/grid/spectral.cpp (init :12-96,
// get_freq_derivative :98-130) and src/grid/mech_base.cpp (base_init
// :11-43, calculate_masked_compliance :249-300, calculate_divergence_rms
// :302-398, gamma_convolution :400-477 with the operator as intended,
// mech_base.cpp:188-205).
#include <algorithm>
#include <cstring>

#include "oracle_internal.h"

namespace orc {

// Spectral::init xi tables.  The first-derivative table is zeroed per
// component p at the Nyquist index of an even dimension p, which is what the
// reference's SpectralTestInit golden (src/grid/test/test_spectral.cpp:23-46)
// requires; spectral.cpp:72-88 zeroes all three components only at the
// all-Nyquist corner, which contradicts that golden.
void spectral_xi(const int cells[3], const double size[3], int derivative,
                 std::vector<cplx>& xi1st, std::vector<cplx>& xi2nd) {
  const int Nx = cells[0], Ny = cells[1], Nz = cells[2], nr = Nx / 2 + 1;
  const double TAU = 2.0 * M_PI;
  xi1st.assign((size_t)3 * nr * Nz * Ny, cplx(0, 0));
  xi2nd.assign((size_t)3 * nr * Nz * Ny, cplx(0, 0));
  for (int j = 0; j < Ny; ++j) {
    int ks[3];
    ks[1] = (j > Ny / 2) ? j - Ny : j;
    for (int k = 0; k < Nz; ++k) {
      ks[2] = (k > Nz / 2) ? k - Nz : k;
      for (int i = 0; i < nr; ++i) {
        ks[0] = i;
        cplx fd[3];
        for (int p = 0; p < 3; ++p) {
          if (derivative == 0) {
            fd[p] = cplx(0.0, TAU * ks[p] / size[p]);
          } else if (derivative == 1) {
            fd[p] = cplx(0.0, std::sin(TAU * ks[p] / cells[p])) / cplx(2.0 * size[p] / cells[p], 0.0);
          } else {
            int p1 = (p + 1) % 3, p2 = (p + 2) % 3;
            fd[p] = cplx(std::cos(TAU * ks[p] / cells[p]) - (p == 0 ? 1.0 : -1.0), std::sin(TAU * ks[p] / cells[p])) *
                    cplx(std::cos(TAU * ks[p1] / cells[p1]) + 1.0, std::sin(TAU * ks[p1] / cells[p1])) *
                    cplx(std::cos(TAU * ks[p2] / cells[p2]) + 1.0, std::sin(TAU * ks[p2] / cells[p2])) /
                    cplx(4.0 * size[p] / cells[p], 0.0);
          }
        }
        int idx[3] = {i, j, k};
        size_t base = 3 * (i + (size_t)nr * (k + (size_t)Nz * j));
        for (int p = 0; p < 3; ++p) {
          xi2nd[base + p] = fd[p];
          bool nyq = (cells[p] % 2 == 0) && (idx[p] == cells[p] / 2) && derivative == 0;
          xi1st[base + p] = nyq ? cplx(0.0, 0.0) : fd[p];
        }
      }
    }
  }
}

// MechBase::base_init, mech_base.cpp:11-33
void scaled_geom(const int cells[3], const double size[3], int dc, double out[3]) {
  int idx[3] = {0, 1, 2};
  if (dc == 1) {
    std::sort(idx, idx + 3, [&](int a, int b) { return size[a] < size[b]; });
    for (int j = 0; j < 3; ++j) out[j] = size[j] / size[idx[1]];
  } else if (dc == 2) {
    double ng[3] = {size[0] / cells[0], size[1] / cells[1], size[2] / cells[2]};
    std::sort(idx, idx + 3, [&](int a, int b) { return ng[a] < ng[b]; });
    for (int j = 0; j < 3; ++j) out[j] = size[j] / size[idx[1]] * cells[idx[1]];
  } else {
    for (int j = 0; j < 3; ++j) out[j] = size[j];
  }
}

void grid_setup(Grid& g, const int cells[3], const double size[3], int derivative, int dc) {
  for (int p = 0; p < 3; ++p) {
    g.cells[p] = cells[p];
    g.size[p] = size[p];
  }
  g.N = (int64_t)cells[0] * cells[1] * cells[2];
  g.nr = cells[0] / 2 + 1;
  g.K = (int64_t)g.nr * cells[1] * cells[2];
  g.wgt = std::pow((double)(cells[0] * cells[1] * cells[2]), -1);
  scaled_geom(cells, size, dc, g.scaled_size);
  spectral_xi(cells, size, derivative, g.xi1st, g.xi2nd);
}

// (3,3,N) real -> (3,3,K) complex, component-interleaved as the reference
void tensor_r2c(const Grid& g, const double* field, std::vector<cplx>& out) {
  out.assign((size_t)9 * g.K, cplx(0, 0));
  std::vector<double> comp(g.N);
  std::vector<cplx> spec(g.K);
  for (int c = 0; c < 9; ++c) {
    for (int64_t n = 0; n < g.N; ++n) comp[n] = field[9 * n + c];
    fft3d_r2c(g.cells, comp.data(), spec.data());
    for (int64_t k = 0; k < g.K; ++k) out[9 * k + c] = spec[k];
  }
}
void tensor_c2r(const Grid& g, std::vector<cplx>& in, double wgt, double* field) {
  std::vector<double> comp(g.N);
  std::vector<cplx> spec(g.K);
  for (int c = 0; c < 9; ++c) {
    for (int64_t k = 0; k < g.K; ++k) spec[k] = in[9 * k + c];
    fft3d_c2r(g.cells, spec.data(), comp.data());
    for (int64_t n = 0; n < g.N; ++n) field[9 * n + c] = comp[n] * wgt;
  }
}

// calculate_divergence_rms, mech_base.cpp:302-398 (on a given spectrum)
double divergence_rms(const Grid& g, const std::vector<cplx>& Phat) {
  const int Ny = g.cells[1], Nz = g.cells[2], nr = g.nr;
  cplx rescaled[3];
  for (int p = 0; p < 3; ++p) rescaled[p] = cplx(g.size[p] / g.scaled_size[p], 0.0);
  auto contrib = [&](int i, int k, int j) {
    size_t kk = i + (size_t)nr * (k + (size_t)Nz * j);
    const cplx* P = &Phat[9 * kk];
    const cplx* xi = &g.xi1st[3 * kk];
    cplx v[3];
    for (int n = 0; n < 3; ++n) v[n] = std::conj(xi[n]) * rescaled[n];
    double s = 0.0;
    for (int m = 0; m < 3; ++m) {
      cplx prod(0.0, 0.0);
      for (int n = 0; n < 3; ++n) prod += P[m + 3 * n] * v[n];
      s += prod.real() * prod.real();
    }
    for (int m = 0; m < 3; ++m) {
      cplx prod(0.0, 0.0);
      for (int n = 0; n < 3; ++n) prod += P[m + 3 * n] * v[n];
      s += prod.imag() * prod.imag();
    }
    return s;
  };
  double rms = 0.0;
  for (int j = 0; j < Ny; ++j)
    for (int k = 0; k < Nz; ++k) {
      for (int i = 1; i < nr - 1; ++i) rms += 2.0 * contrib(i, k, j);
      rms += contrib(0, k, j);
      rms += contrib(nr - 1, k, j);
    }
  rms = std::sqrt(rms) * g.wgt;
  if (g.cells[0] == 1) rms = rms / 2;
  return rms;
}

// gamma_convolution body (memory-efficient branch) with the intended
// operator Gamma_lmno = (A^-1)_ln conj(-xi_o) xi_m when |det Re A| > 1e-16
// (mech_base.cpp:188-205).  With C_ref = C / wgt (mech_base.cpp:157), an
// unnormalised forward transform and an unnormalised inverse (neutral_wgt = 1,
// mech_base.cpp:474-476) the non-DC modes give Gamma(C) * P exactly, and the
// DC mode is set to the aim itself (mech_base.cpp:467-473).
void gamma_apply(const Grid& g, const T4& C_ref, std::vector<cplx>& Phat, const M3& aim) {
  const int Ny = g.cells[1], Nz = g.cells[2], nr = g.nr;
  for (int j = 0; j < Ny; ++j)
    for (int k = 0; k < Nz; ++k)
      for (int i = 0; i < nr; ++i) {
        if (i == 0 && j == 0 && k == 0) continue;
        size_t kk = i + (size_t)nr * (k + (size_t)Nz * j);
        const cplx* xi = &g.xi1st[3 * kk];
        cplx dyad[3][3], Z[3][3];
        for (int l = 0; l < 3; ++l)
          for (int m = 0; m < 3; ++m) dyad[l][m] = std::conj(-xi[l]) * xi[m];
        for (int l = 0; l < 3; ++l)
          for (int m = 0; m < 3; ++m) {
            cplx s(0.0, 0.0);
            for (int n = 0; n < 3; ++n)
              for (int o = 0; o < 3; ++o) s += cplx(C_ref.a[l][n][m][o], 0.0) * dyad[n][o];
            Z[l][m] = s;
          }
        double A[36];
        for (int l = 0; l < 3; ++l)
          for (int m = 0; m < 3; ++m) {
            A[l * 6 + m] = Z[l][m].real();
            A[(l + 3) * 6 + m + 3] = Z[l][m].real();
            A[l * 6 + m + 3] = Z[l][m].imag();
            A[(l + 3) * 6 + m] = -Z[l][m].imag();
          }
        M3 Are;
        for (int l = 0; l < 3; ++l)
          for (int m = 0; m < 3; ++m) Are.a[l][m] = A[l * 6 + m];
        cplx res[3][3];
        if (std::fabs(det33(Are)) > 1e-16) {
          double Ainv[36];
          bool err = invert(6, A, Ainv);
          cplx Zi[3][3];
          for (int l = 0; l < 3; ++l)
            for (int n = 0; n < 3; ++n) Zi[l][n] = err ? cplx(0, 0) : cplx(Ainv[l * 6 + n], Ainv[l * 6 + n + 3]);
          // r_lm = sum_no Zi_ln conj(-xi_o) xi_m P_no
          for (int l = 0; l < 3; ++l)
            for (int m = 0; m < 3; ++m) {
              cplx s(0.0, 0.0);
              for (int n = 0; n < 3; ++n)
                for (int o = 0; o < 3; ++o) s += Zi[l][n] * std::conj(-xi[o]) * xi[m] * Phat[9 * kk + n + 3 * o];
              res[l][m] = s;
            }
        } else {
          for (int l = 0; l < 3; ++l)
            for (int m = 0; m < 3; ++m) res[l][m] = cplx(0.0, 0.0);
        }
        for (int l = 0; l < 3; ++l)
          for (int m = 0; m < 3; ++m) Phat[9 * kk + l + 3 * m] = res[l][m];
      }
  for (int l = 0; l < 3; ++l)
    for (int m = 0; m < 3; ++m) Phat[l + 3 * m] = cplx(aim.a[l][m], 0.0);
}

// calculate_masked_compliance, mech_base.cpp:249-300.  The stress mask is
// flattened column-major (row + 3*col) while the 9x9 matrix uses the plain
// row-major map, exactly as the reference does.
T4 masked_compliance(const T4& C, const double q[4], const int mask[9]) {
  bool m1d[9];
  int size_reduced = 0;
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row) {
      m1d[col * 3 + row] = !mask[row + 3 * col];
      size_reduced += m1d[col * 3 + row] ? 1 : 0;
    }
  M99 t99;
  std::memset(&t99, 0, sizeof t99);
  if (size_reduced > 0) {
    T4 rot = rotTensor4(q, C, false);
    M99 c99 = t3333to99(rot);
    std::vector<double> cr(size_reduced * size_reduced), sr(size_reduced * size_reduced);
    int idx = 0;
    for (int i = 0; i < 9; ++i)
      for (int j = 0; j < 9; ++j)
        if (m1d[i] && m1d[j]) {
          // c_reduced(idx / n, idx % n) in Eigen (column-major) storage
          int r = idx / size_reduced, c = idx % size_reduced;
          cr[r * size_reduced + c] = c99.a[i][j];
          ++idx;
        }
    if (invert(size_reduced, cr.data(), sr.data())) throw std::runtime_error("Matrix inversion error");
    idx = 0;
    for (int i = 0; i < 9; ++i)
      for (int j = 0; j < 9; ++j)
        if (m1d[i] && m1d[j]) {
          int r = idx / size_reduced, c = idx % size_reduced;
          t99.a[i][j] = sr[r * size_reduced + c];
          ++idx;
        }
  }
  return t99to3333(t99);
}

}  // namespace orc

namespace orc {
// Points whose tangents form C_minMaxAvg (spectral.cpp:189-207 picks the
// first cell of largest and of smallest squared norm |dPdF|^2 with exact
// comparisons).  |dPdF|^2 is invariant under rotation, so grains that differ
// only in orientation tie in exact arithmetic and the reference's choice
// among them is decided by rounding noise of its own summation.  Here norms
// within TANGENT_TIE_REL of the extreme count as tied and the first such
// cell is taken: the same cells for any implementation whose norms agree to
// far better than TANGENT_TIE_REL.
void tangent_minmax_select(const double* n2, int64_t n, int64_t& imax, int64_t& imin) {
  imax = imin = -1;
  if (n <= 0) return;
  double M = n2[0], m = n2[0];
  for (int64_t c = 1; c < n; ++c) {
    M = std::max(M, n2[c]);
    m = std::min(m, n2[c]);
  }
  const double hi = M * (1.0 - TANGENT_TIE_REL), lo = m * (1.0 + TANGENT_TIE_REL);
  for (int64_t c = 0; c < n && (imax < 0 || imin < 0); ++c) {
    if (imax < 0 && n2[c] >= hi) imax = c;
    if (imin < 0 && n2[c] <= lo) imin = c;
  }
}

}  // namespace orc
