// ORACLE -- test infrastructure only (see oracle_internal.h).
// Restates the tensor/matrix helpers of src/math.f90, src/prec.f90 and
// src/rotations.f90 that the hot path uses.
#include <cfloat>
#include <cstring>

#include "oracle_internal.h"

namespace orc {

M3 zero3() {
  M3 r;
  std::memset(&r, 0, sizeof r);
  return r;
}
M3 eye3() {
  M3 r = zero3();
  r.a[0][0] = r.a[1][1] = r.a[2][2] = 1.0;
  return r;
}
// matmul(A,B): sum over k in ascending order
M3 mul(const M3& A, const M3& B) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A.a[i][k] * B.a[k][j];
      r.a[i][j] = s;
    }
  return r;
}
M3 tr(const M3& A) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = A.a[j][i];
  return r;
}
M3 add(const M3& A, const M3& B) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = A.a[i][j] + B.a[i][j];
  return r;
}
M3 sub(const M3& A, const M3& B) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = A.a[i][j] - B.a[i][j];
  return r;
}
M3 scal(const M3& A, double s) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = A.a[i][j] * s;
  return r;
}
// norm2 over a 3x3 array, Fortran column order
double norm2(const M3& A) {
  double s = 0.0;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) s += A.a[i][j] * A.a[i][j];
  return std::sqrt(s);
}

// math.f90:689-698
double det33(const M3& m) {
  const auto& a = m.a;
  return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
         a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
         a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}

// prec.f90:113-125: |a| <= 10*tiny
bool dEq0(double a) { return std::fabs(a) <= DBL_MIN * 10.0; }
bool dNeq0(double a) { return !dEq0(a); }

// math.f90:511-545, Cramer's rule
bool invert33(const M3& Am, M3& I, double* detOut) {
  const auto& A = Am.a;
  auto& R = I.a;
  R[0][0] = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  R[1][0] = -A[1][0] * A[2][2] + A[1][2] * A[2][0];
  R[2][0] = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  double det = A[0][0] * R[0][0] + A[0][1] * R[1][0] + A[0][2] * R[2][0];
  if (dEq0(det)) {
    I = zero3();
    if (detOut) *detOut = 0.0;
    return false;
  }
  R[0][1] = -A[0][1] * A[2][2] + A[0][2] * A[2][1];
  R[1][1] = A[0][0] * A[2][2] - A[0][2] * A[2][0];
  R[2][1] = -A[0][0] * A[2][1] + A[0][1] * A[2][0];
  R[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
  R[1][2] = -A[0][0] * A[1][2] + A[0][2] * A[1][0];
  R[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = R[i][j] / det;
  if (detOut) *detOut = det;
  return true;
}
M3 inv33(const M3& A) {
  M3 r;
  invert33(A, r);
  return r;
}

T4 zero4() {
  T4 r;
  std::memset(&r, 0, sizeof r);
  return r;
}
// math.f90:440-457: C(i,j,k,l) = sum_{m,n} A(i,j,m,n) B(m,n,k,l)
T4 mul3333xx3333(const T4& A, const T4& B) {
  T4 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          double s = 0.0;
          for (int n = 0; n < 3; ++n)
            for (int m = 0; m < 3; ++m) s += A.a[i][j][m][n] * B.a[m][n][k][l];
          r.a[i][j][k][l] = s;
        }
  return r;
}
// math.f90:408-425
M3 mul3333xx33(const T4& A, const M3& B) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int n = 0; n < 3; ++n)
        for (int m = 0; m < 3; ++m) s += A.a[i][j][m][n] * B.a[m][n];
      r.a[i][j] = s;
    }
  return r;
}
// MAPPLAIN (math.f90:73-84): plain index p = 3*i + j
M99 t3333to99(const T4& m) {
  M99 r;
  for (int p = 0; p < 9; ++p)
    for (int q = 0; q < 9; ++q) r.a[p][q] = m.a[p / 3][p % 3][q / 3][q % 3];
  return r;
}
T4 t99to3333(const M99& m) {
  T4 r;
  for (int p = 0; p < 9; ++p)
    for (int q = 0; q < 9; ++q) r.a[p / 3][p % 3][q / 3][q % 3] = m.a[p][q];
  return r;
}
std::array<double, 9> m33to9(const M3& m) {
  std::array<double, 9> v;
  for (int p = 0; p < 9; ++p) v[p] = m.a[p / 3][p % 3];
  return v;
}
M3 m9to33(const std::array<double, 9>& v) {
  M3 m;
  for (int p = 0; p < 9; ++p) m.a[p / 3][p % 3] = v[p];
  return m;
}
M99 eye99() {
  M99 r;
  std::memset(&r, 0, sizeof r);
  for (int i = 0; i < 9; ++i) r.a[i][i] = 1.0;
  return r;
}

// LAPACK dgetrf semantics (unblocked, right-looking): pivot = first index of
// max |a(i,k)|, i >= k; info = first k with a zero pivot.
int lu_factor(int n, double* a, int* ipiv) {
  int info = 0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double mx = std::fabs(a[k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      double v = std::fabs(a[i * n + k]);
      if (v > mx) {
        mx = v;
        p = i;
      }
    }
    ipiv[k] = p;
    if (a[p * n + k] != 0.0) {
      if (p != k)
        for (int j = 0; j < n; ++j) std::swap(a[k * n + j], a[p * n + j]);
      double inv = 1.0 / a[k * n + k];
      for (int i = k + 1; i < n; ++i) a[i * n + k] *= inv;
    } else if (info == 0) {
      info = k + 1;
    }
    for (int i = k + 1; i < n; ++i) {
      double l = a[i * n + k];
      for (int j = k + 1; j < n; ++j) a[i * n + j] -= l * a[k * n + j];
    }
  }
  return info;
}
void lu_solve(int n, const double* lu, const int* ipiv, double* b) {
  for (int k = 0; k < n; ++k)
    if (ipiv[k] != k) std::swap(b[k], b[ipiv[k]]);
  for (int i = 1; i < n; ++i) {
    double s = b[i];
    for (int j = 0; j < i; ++j) s -= lu[i * n + j] * b[j];
    b[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= lu[i * n + j] * b[j];
    b[i] = s / lu[i * n + i];
  }
}
int gesv(int n, double* a, double* b) {
  std::vector<int> ipiv(n);
  int info = lu_factor(n, a, ipiv.data());
  if (info != 0) return info;
  lu_solve(n, a, ipiv.data(), b);
  return 0;
}
bool invert(int n, const double* a, double* inv) {
  std::vector<double> lu(a, a + n * n);
  std::vector<int> ipiv(n);
  int info = lu_factor(n, lu.data(), ipiv.data());
  if (info != 0) return true;
  std::vector<double> col(n);
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) col[i] = (i == c) ? 1.0 : 0.0;
    lu_solve(n, lu.data(), ipiv.data(), col.data());
    for (int i = 0; i < n; ++i) inv[i * n + c] = col[i];
  }
  return false;
}

// MAPVOIGT (math.f90:63-71): (1,1),(2,2),(3,3),(2,3),(1,3),(1,2)
static const int VOIGT[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
// math.f90:980-995
T4 voigt66to3333_stiffness(const double C66[6][6]) {
  T4 C = zero4();
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      int a = VOIGT[i][0], b = VOIGT[i][1], c = VOIGT[j][0], d = VOIGT[j][1];
      C.a[a][b][c][d] = C66[i][j];
      C.a[b][a][c][d] = C66[i][j];
      C.a[a][b][d][c] = C66[i][j];
      C.a[b][a][d][c] = C66[i][j];
    }
  return C;
}
// math.f90:964-973
void m33toVoigt6_strain(const M3& e, double v[6]) {
  v[0] = e.a[0][0];
  v[1] = e.a[1][1];
  v[2] = e.a[2][2];
  v[3] = 2.0 * e.a[2][1];
  v[4] = 2.0 * e.a[2][0];
  v[5] = 2.0 * e.a[0][1];
}
// math.f90:917-927
M3 voigt6to33_stress(const double s[6]) {
  M3 r;
  r.a[0][0] = s[0];
  r.a[1][0] = s[5];
  r.a[2][0] = s[4];
  r.a[0][1] = s[5];
  r.a[1][1] = s[1];
  r.a[2][1] = s[3];
  r.a[0][2] = s[4];
  r.a[1][2] = s[3];
  r.a[2][2] = s[2];
  return r;
}

// rotations.f90:423-447, convention P = -1 (:56)
M3 qu2om(const double q[4]) {
  M3 om;
  double qq = q[0] * q[0] - (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  om.a[0][0] = qq + 2.0 * q[1] * q[1];
  om.a[1][1] = qq + 2.0 * q[2] * q[2];
  om.a[2][2] = qq + 2.0 * q[3] * q[3];
  om.a[0][1] = 2.0 * (q[1] * q[2] - q[0] * q[3]);
  om.a[1][2] = 2.0 * (q[2] * q[3] - q[0] * q[1]);
  om.a[2][0] = 2.0 * (q[3] * q[1] - q[0] * q[2]);
  om.a[1][0] = 2.0 * (q[2] * q[1] + q[0] * q[3]);
  om.a[2][1] = 2.0 * (q[3] * q[2] + q[0] * q[1]);
  om.a[0][2] = 2.0 * (q[1] * q[3] + q[0] * q[2]);
  om = tr(om);  // P < 0
  double d = std::pow(det33(om), 1.0 / 3.0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) om.a[i][j] = om.a[i][j] / d;
  return om;
}
// rotations.f90:289-301: passive R T R^T, active R^T T R
M3 rotTensor2(const double q[4], const M3& T, bool active) {
  M3 R = qu2om(q);
  if (active) return mul(mul(tr(R), T), R);
  return mul(mul(R, T), tr(R));
}
// rotations.f90:308-328
T4 rotTensor4(const double q[4], const T4& T, bool active) {
  M3 R = qu2om(q);
  if (active) R = tr(R);
  T4 r = zero4();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          double s = 0.0;
          for (int m = 0; m < 3; ++m)
            for (int n = 0; n < 3; ++n)
              for (int o = 0; o < 3; ++o)
                for (int p = 0; p < 3; ++p)
                  s += R.a[i][m] * R.a[j][n] * R.a[k][o] * R.a[l][p] * T.a[m][n][o][p];
          r.a[i][j][k][l] = s;
        }
  return r;
}

static double clip(double a, double lo, double hi) { return std::min(hi, std::max(lo, a)); }
// math.f90:1117-1156
M3 rotationalPart(const M3& F) {
  M3 C = mul(tr(F), F);
  double I1 = C.a[0][0] + C.a[1][1] + C.a[2][2];
  double I2 = C.a[0][0] * C.a[1][1] + C.a[0][0] * C.a[2][2] + C.a[1][1] * C.a[2][2] -
              (C.a[0][1] * C.a[0][1] + C.a[0][2] * C.a[0][2] + C.a[1][2] * C.a[1][2]);
  double I3 = det33(C);
  M3 FF = mul(F, F);
  double trF = F.a[0][0] + F.a[1][1] + F.a[2][2];
  double IF2 = 0.5 * (trF * trF - (FF.a[0][0] + FF.a[1][1] + FF.a[2][2]));
  double x = std::pow(std::max(I1 * I1 - 3.0 * I2, 0.0), 1.5);
  double lam[3];
  if (dNeq0(x)) {
    double Phi = std::acos(clip((I1 * I1 * I1 - 4.5 * I1 * I2 + 13.5 * I3) / x, -1.0, 1.0));
    const double TAU = 2.0 * M_PI;
    for (int k = 0; k < 3; ++k) {
      double l = I1 + (2.0 * std::sqrt(std::max(I1 * I1 - 3.0 * I2, 0.0))) *
                          std::cos((Phi - TAU * (k + 1)) / 3.0);
      lam[k] = std::sqrt(std::max(l, 0.0) / 3.0);
    }
  } else {
    for (int k = 0; k < 3; ++k) lam[k] = std::sqrt(I1 / 3.0);
  }
  double IU1 = lam[0] + lam[1] + lam[2];
  double IU2 = lam[0] * lam[1] + lam[1] * lam[2] + lam[2] * lam[0];
  M3 FtFF = tr(FF), FC = mul(F, C), Ft = tr(F);
  M3 R;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      R.a[i][j] = IU1 * IF2 * (i == j ? 1.0 : 0.0) + (IU1 * IU1 - IU2) * F.a[i][j] -
                  IU1 * trF * Ft.a[i][j] + IU1 * FtFF.a[i][j] - FC.a[i][j];
  double s = std::pow(det33(R), -1.0 / 3.0);
  return scal(R, s);
}

M3 load33(const double* p) {
  M3 m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.a[i][j] = p[i + 3 * j];
  return m;
}
void store33(const M3& m, double* p) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) p[i + 3 * j] = m.a[i][j];
}
T4 load3333(const double* p) {
  T4 t;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) t.a[i][j][k][l] = p[i + 3 * j + 9 * k + 27 * l];
  return t;
}
void store3333(const T4& t, double* p) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) p[i + 3 * j + 9 * k + 27 * l] = t.a[i][j][k][l];
}

}  // namespace orc
