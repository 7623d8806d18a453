// ORACLE -- test infrastructure only.
//
// FP64 CPU restatement of the reference's spectral grid solver and
// phenopowerlaw material-point update (dmentock/solver_cpp, a DAMASK
// 3.0.0-alpha8 fork).  It is the checker the HIP product is compared with;
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it, and never as the thing measured or shipped.
//
// Each routine cites the reference file:line it restates.  Index
// conventions: M3::a[i][j] is the mathematical element (i,j); flat (3,3)
// arrays at the C-ABI are Fortran column-major (element (i,j) at i+3j).
#pragma once
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "../include/solver_cpp_amd.h"

namespace orc {

struct M3 {
  double a[3][3];
};
struct T4 {
  double a[3][3][3][3];
};
struct M99 {
  double a[9][9];
};
using cplx = std::complex<double>;

// ---- small linear algebra (math.f90) ------------------------------------
M3 zero3();
M3 eye3();
M3 mul(const M3& A, const M3& B);   // matmul
M3 tr(const M3& A);                 // transpose
M3 add(const M3& A, const M3& B);
M3 sub(const M3& A, const M3& B);
M3 scal(const M3& A, double s);
double norm2(const M3& A);          // Frobenius
double det33(const M3& m);          // math.f90:689-698
bool invert33(const M3& A, M3& inv, double* det = nullptr);  // math.f90:511-545
M3 inv33(const M3& A);              // math.f90:491-503 (zeros on failure)
bool dEq0(double a);                // prec.f90:113-125
bool dNeq0(double a);
T4 zero4();
T4 mul3333xx3333(const T4& A, const T4& B);  // math.f90:440-457
M3 mul3333xx33(const T4& A, const M3& B);    // math.f90:424-433
M99 t3333to99(const T4& m);                   // math.f90:800-816 (plain map)
T4 t99to3333(const M99& m);                   // math.f90:830-845
std::array<double, 9> m33to9(const M3& m);    // math.f90:718-728
M3 m9to33(const std::array<double, 9>& v);
M99 eye99();
// LU with partial pivoting (LAPACK dgetrf semantics); returns info (0 = ok).
int lu_factor(int n, double* a /*row-major n x n*/, int* ipiv);
void lu_solve(int n, const double* lu, const int* ipiv, double* b);
// dgesv: solve A x = b in place; returns info.
int gesv(int n, double* a, double* b);
// math_invert (math.f90:578-595): n x n inverse via LU; returns error flag.
bool invert(int n, const double* a, double* inv);

// Voigt maps (math.f90:917-995)
T4 voigt66to3333_stiffness(const double C66[6][6]);
void m33toVoigt6_strain(const M3& e, double v[6]);
M3 voigt6to33_stress(const double s[6]);

// rotations.f90
M3 qu2om(const double q[4]);                                  // :423-447
M3 rotTensor2(const double q[4], const M3& T, bool active);   // :289-301
T4 rotTensor4(const double q[4], const T4& T, bool active);   // :308-328
M3 rotationalPart(const M3& F);                               // math.f90:1117-1156

// flat (Fortran column-major) <-> M3 / T4
M3 load33(const double* p);
void store33(const M3& m, double* p);
T4 load3333(const double* p);
void store3333(const T4& t, double* p);

// ---- crystal / material ---------------------------------------------------
constexpr int NSLIP = 12;
struct Phase {
  sc_material prm;
  sc_phase_numerics num;
  double C66[6][6];
  T4 C;                       // math_Voigt66to3333_stiffness(C66)
  M3 P_sl[NSLIP];             // Schmid matrices
  M3 P_nS_pos[NSLIP], P_nS_neg[NSLIP];
  double h_sl_sl[NSLIP][NSLIP];  // (reacting, acting)
};
void phase_setup(Phase& ph, const sc_material& m, const sc_phase_numerics& num);

// ---- material point (phase_mechanical.f90 and friends) --------------------
// Per-point state; Fi == I and Li == 0 throughout (no eigen kinematics,
// material_V_e_0 = I).
struct PointState {
  // converged at end of last increment (phase_mechanical_*0, state0)
  M3 F0, Fp0, Lp0, S0;
  double xi0[NSLIP], gamma0[NSLIP];
  // current
  M3 F, Fp, Lp, S, Fe, P;
  double xi[NSLIP], gamma[NSLIP];
  int newton_iter, state_iter;   // counters of the last update
};

// phase_restore(ce,.false.) + mechanical_partition + phase_mechanical_constitutive
// (homogenization.f90:238-258).  Returns converged.
bool mp_response(const Phase& ph, PointState& s, double dt);
// phase_mechanical_dPdF, phase_mechanical.f90:1137-1258
T4 mp_dPdF(const Phase& ph, const PointState& s, double dt);
// mechanical_forward, phase_mechanical.f90:992-1007
void mp_forward(PointState& s);

// ---- FFT -------------------------------------------------------------------
// complex 1D DFT of length n (any n; mixed radix with a direct DFT fallback)
void fft1d(int n, cplx* data, int stride, int sign);
// 3D R2C on one real scalar field (Nx,Ny,Nz), out (Nx/2+1, Nz, Ny)
void fft3d_r2c(const int cells[3], const double* in, cplx* out);
void fft3d_c2r(const int cells[3], const cplx* in, double* out);  // unnormalised

// ---- spectral ----------------------------------------------------------------
struct Grid {
  int cells[3];
  double size[3];
  double scaled_size[3];
  int64_t N;          // cells
  int nr;             // Nx/2+1
  int64_t K;          // nr*Ny*Nz
  double wgt;
  std::vector<cplx> xi1st, xi2nd;  // (3, nr, Nz, Ny)
};
void grid_setup(Grid& g, const int cells[3], const double size[3], int derivative,
                int divergence_correction);
void spectral_xi(const int cells[3], const double size[3], int derivative,
                 std::vector<cplx>& xi1st, std::vector<cplx>& xi2nd);
void scaled_geom(const int cells[3], const double size[3], int dc, double out[3]);
// tensor field r2c: field (3,3,N) -> (3,3,K) complex
void tensor_r2c(const Grid& g, const double* field, std::vector<cplx>& out);
void tensor_c2r(const Grid& g, std::vector<cplx>& in, double wgt, double* field);
double divergence_rms(const Grid& g, const std::vector<cplx>& Phat);
void gamma_apply(const Grid& g, const T4& C_ref, std::vector<cplx>& Phat, const M3& aim);
T4 masked_compliance(const T4& C, const double q[4], const int mask[9]);

// selection of the C_minMaxAvg cells with a relative tie window
constexpr double TANGENT_TIE_REL = 1e-10;
void tangent_minmax_select(const double* n2, int64_t n, int64_t& imax, int64_t& imin);

}  // namespace orc
