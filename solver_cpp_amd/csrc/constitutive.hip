// Phenopowerlaw material-point update on CDNA4: one lane per material point,
// FP64, state in structure-of-arrays planes (coalesced across lanes).
//
// Follows the reference's per-cell response
//   homogenization_mechanical_response   homogenization.f90:219-266
//   phase_mechanical_constitutive        phase_mechanical.f90:1013-1100 (substep cut-back)
//   integrateStateFPI                    phase_mechanical.f90:611-693
//   integrateStress                      phase_mechanical.f90:380-604 (Lp Newton)
//   phase_hooke_SandItsTangents          phase_mechanical_elastic.f90:157-195
//   phenopowerlaw LpAndItsTangent / dotState / kinetics_sl
//                                        phase_mechanical_plastic_phenopowerlaw.f90:298-483
// with Fi = I and Li = 0 (no eigen kinematics are active).  The 9x9 Newton
// Jacobian I - dLp/dS : dS/dFe : dFe/dLp is assembled directly from its
// rank-12 structure: with G = Fe^T A and K_s = P_s : C (cubic),
//   J[(a,b),(p,l)] = delta + dt * sum_s c_s P_s(a,b) (G^T K_s)(p,l),
// c_s = d(gamma_dot_s)/d(tau_s) * sign_s, instead of three 3x3x3x3 products;
// the sum over s is taken from the slip groups' Hadamard sums
// (assemble_jacobian8), on the 8 traceless coordinates.
#include "common.h"

namespace sca {

namespace {

// cF <110>{111} slip systems (crystal.f90:41-63): direction, normal.
__device__ constexpr int SLIP[NSLIP][6] = {
    {0, 1, -1, 1, 1, 1},   {-1, 0, 1, 1, 1, 1},   {1, -1, 0, 1, 1, 1},
    {0, -1, -1, -1, -1, 1}, {1, 0, 1, -1, -1, 1},  {-1, 1, 0, -1, -1, 1},
    {0, -1, 1, 1, -1, -1}, {-1, 0, -1, 1, -1, -1}, {1, 1, 0, 1, -1, -1},
    {0, 1, 1, -1, 1, -1},  {1, 0, -1, -1, 1, -1},  {-1, -1, 0, -1, 1, -1}};

// sign pattern of P_s(i,j) = d_i n_j
__device__ __forceinline__ constexpr int psg(int s, int i, int j) { return SLIP[s][i] * SLIP[s][3 + j]; }

// cF slip-slip interaction types (reacting, acting), crystal.f90:623-640
__device__ constexpr int HTYPE[NSLIP][NSLIP] = {
    {1, 2, 2, 4, 7, 5, 3, 5, 5, 4, 6, 7}, {2, 1, 2, 7, 4, 5, 6, 4, 7, 5, 3, 5},
    {2, 2, 1, 5, 5, 3, 6, 7, 4, 7, 6, 4}, {4, 7, 6, 1, 2, 2, 4, 6, 7, 3, 5, 5},
    {7, 4, 6, 2, 1, 2, 5, 3, 5, 6, 4, 7}, {5, 5, 3, 2, 2, 1, 7, 6, 4, 6, 7, 4},
    {3, 5, 5, 4, 6, 7, 1, 2, 2, 4, 7, 6}, {6, 4, 7, 5, 3, 5, 2, 1, 2, 7, 4, 6},
    {6, 7, 4, 7, 6, 4, 2, 2, 1, 5, 5, 3}, {4, 6, 7, 3, 5, 5, 4, 7, 6, 1, 2, 2},
    {5, 3, 5, 6, 4, 7, 7, 4, 6, 2, 1, 2}, {7, 6, 4, 6, 7, 4, 5, 5, 3, 2, 2, 1}};

// Structure of the table used to share work between slip systems:
// - the systems 3p, 3p+1, 3p+2 lie on the same plane (normal), p = 0..3;
// - the systems g, g+3, g+6, g+9 (g = 0, 1, 2) have direction component g
//   zero.
// The Jacobian (assemble_jacobian8) needs, for every row r = 3i+j < 8 and
// every symmetric pair (m,l), the sum over s of f_s P_s(i,j) k_s(m,l) with
// k_s(m,l) = P_s(m,m) (m = l) or P_s(m,l) + P_s(l,m).  Restricted to one
// direction group its coefficient vector is zero or plus/minus one of the
// four Hadamard rows (1,1,1,1), (1,1,-1,-1), (1,-1,1,-1), (1,-1,-1,1), so
// the sums come from the four Hadamard sums of each group's f_s.
__device__ __forceinline__ constexpr int qcoef(int s, int r, int m, int l) {
  return psg(s, r / 3, r % 3) * (m == l ? psg(s, m, m) : psg(s, m, l) + psg(s, l, m));
}
// 0 when group g does not enter Q_r(m,l), else sign * (1 + Hadamard row)
__device__ __forceinline__ constexpr int qhad(int g, int r, int m, int l) {
  const int a = qcoef(g, r, m, l), b = qcoef(g + 3, r, m, l), c = qcoef(g + 6, r, m, l);
  if (a == 0) return 0;
  return a * (1 + (b * a > 0 ? (c * a > 0 ? 0 : 1) : (c * a > 0 ? 2 : 3)));
}
constexpr bool slip_structure_holds() {
  for (int p = 0; p < 4; ++p)
    for (int k = 1; k < 3; ++k)
      for (int j = 0; j < 3; ++j)
        if (SLIP[3 * p + k][3 + j] != SLIP[3 * p][3 + j]) return false;
  for (int s = 0; s < NSLIP; ++s)
    if (SLIP[s][s % 3] != 0) return false;
  const int had[4][4] = {{1, 1, 1, 1}, {1, 1, -1, -1}, {1, -1, 1, -1}, {1, -1, -1, 1}};
  for (int r = 0; r < 8; ++r)
    for (int m = 0; m < 3; ++m)
      for (int l = m; l < 3; ++l) {
        int groups = 0;
        for (int g = 0; g < 3; ++g) {
          const int code = qhad(g, r, m, l);
          for (int k = 0; k < 4; ++k) {
            const int want = code == 0 ? 0 : (code > 0 ? 1 : -1) * had[(code > 0 ? code : -code) - 1][k];
            if (qcoef(g + 3 * k, r, m, l) != want) return false;
          }
          groups += code != 0;
        }
        if (groups == 0) return false;
      }
  return true;
}
static_assert(slip_structure_holds(), "cF slip table: planes of three systems, Hadamard-structured Jacobian sums");

// Material constants travel in the kernel arguments (not a __constant__
// symbol) so that launches of different contexts on concurrent streams never
// share them.
struct KernelArgs {
  MpState s;
  MatConst mc;
  double dt;
  int* ill;
  int* broken;  // per point, may be null
  int* repivot; // [0] count, [1 + i] cells for the pivoting pass
  double* sub;  // substep scratch, SUB_PLANES planes of N doubles
  int64_t c_begin, c_end;  // cells [c_begin, c_end) of the N-cell planes
};


// ---- small dense helpers (row-major 3x3 in registers: m[3*i+j]) ----------
__device__ __forceinline__ void mm33(const double* A, const double* B, double* C) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

// Divisions without the compiler's range fix-ups, for operands far from the
// range limits (slip resistances, the 3x3 determinants of deformation
// gradients, the pivots of the Newton system, none of them near 2^+-1000):
// 1/x as the hardware reciprocal refined by two Newton steps (as the
// compiler's own division does before its fix-ups), and a / b as the rounded
// product a (1/b) corrected once by its exact remainder a - b q (Markstein):
// the correctly rounded quotient when 1/b is the correctly rounded
// reciprocal, within one rounding of it otherwise
__device__ __forceinline__ double rcp_refined(double x) {
  double r = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int k = 0; k < 2; ++k) r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  return r;
}
__device__ __forceinline__ double div_rcp(double a, double b, double rb) {
  const double q = a * rb;
  return __builtin_fma(__builtin_fma(-q, b, a), rb, q);
}

// math_invert33 (math.f90:511-545); returns false if |det| <= 10*tiny
__device__ __forceinline__ bool inv33(const double* A, double* R, double* detOut) {
  R[0] = A[4] * A[8] - A[5] * A[7];
  R[3] = -A[3] * A[8] + A[5] * A[6];
  R[6] = A[3] * A[7] - A[4] * A[6];
  double det = A[0] * R[0] + A[1] * R[3] + A[2] * R[6];
  if (fabs(det) <= 2.2250738585072014e-308 * 10.0) return false;
  R[1] = -A[1] * A[8] + A[2] * A[7];
  R[4] = A[0] * A[8] - A[2] * A[6];
  R[7] = -A[0] * A[7] + A[1] * A[6];
  R[2] = A[1] * A[5] - A[2] * A[4];
  R[5] = -A[0] * A[5] + A[2] * A[3];
  R[8] = A[0] * A[4] - A[1] * A[3];
  const double rd = rcp_refined(det);
#pragma unroll
  for (int k = 0; k < 9; ++k) R[k] = div_rcp(R[k], det, rd);
  if (detOut) *detOut = det;
  return true;
}

__device__ __forceinline__ double det33(const double* m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}


// x^e for x >= 0 with e = e4/4: binary powering of the integer part times
// sqrt / sqrt-sqrt for the quarter part; libm pow() when e4 < 0
__device__ __noinline__ double pow_general(double x, double e) { return pow(x, e); }
__device__ __forceinline__ double powq(double x, int e4, double e) {
  if (e4 < 0) return pow_general(x, e);
  double r = 1.0, b = x;
  int m = e4 >> 2;
  while (m) {
    if (m & 1) r *= b;
    b *= b;
    m >>= 1;
  }
  const int q = e4 & 3;
  if (q) {
    const double h = sqrt(x);
    const double f = sqrt(h);
    r *= (q == 1) ? f : (q == 2) ? h : h * f;
  }
  return r;
}
// compile-time quarter exponent E4/4 (E4 >= 0): fully unrolled
// x^(-1/4) for x in [2^-1000, 2^1000]: a start value from the hardware
// square root and reciprocal square root, then three Newton steps
// y <- y (5 - x y^4) / 4 (quadratic: full precision from any start within
// 2^-8 of the root); two correctly rounded square roots cost about three
// times as many instructions
__device__ __forceinline__ double rqrt_refined(double x) {
  double y = __builtin_amdgcn_rsq(__builtin_amdgcn_sqrt(x));
  const double mq = -0.25 * x;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double y2 = y * y;
    y = y * __builtin_fma(mq, y2 * y2, 1.25);
  }
  return y;
}

template <int E4>
__device__ __forceinline__ double powc(double x) {
  double r = 1.0, b = x;
#pragma unroll
  for (int m = E4 >> 2; m; m >>= 1) {
    if (m & 1) r *= b;
    b *= b;
  }
  constexpr int q = E4 & 3;
  if constexpr (q == 2) {
    r *= sqrt(x);
  } else if constexpr (q != 0) {
    // x^(1/4) = x y^3 and x^(3/4) = x y with y = x^(-1/4) (within 10 ulp),
    // y taken at x clamped to [2^-1000, 2^1000] and multiplied by the true x:
    // for x = 0 this gives 0, NaN propagates, and outside the range the
    // power x^(E4 >> 2) >= x^1 (E4 >= 4) underflows to 0 or overflows to inf
    // as the exact result does (E4 < 4 exponents run the runtime path)
    static_assert(E4 >= 4, "quarter exponents below 1 are not compiled in");
    const double y = rqrt_refined(fmin(fmax(x, 0x1p-1000), 0x1p1000));
    r *= (q == 1) ? x * (y * (y * y)) : x * y;
  }
  return r;
}
// exponent dispatch: E4 >= 0 compile-time, E4 == -1 runtime (mc)
template <int E4>
__device__ __forceinline__ double powe(double x, int e4rt, double e) {
  if constexpr (E4 >= 0) return powc<E4>(x);
  else return powq(x, e4rt, e);
}

// Hooke (elastic.f90:157-195) with Fi = I, cubic C66:
// E = (Fe^T Fe - I)/2, S = C66 : E (Voigt strain with engineering shear)
__device__ __forceinline__ void hooke_S(const double* Fe, double* S, const MatConst& mc) {
  double E[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      E[3 * i + j] = 0.5 * ((Fe[i] * Fe[j] + Fe[3 + i] * Fe[3 + j] + Fe[6 + i] * Fe[6 + j]) - (i == j ? 1.0 : 0.0));
  const double e11 = E[0], e22 = E[4], e33 = E[8];
  const double s11 = mc.C11 * e11 + mc.C12 * e22 + mc.C12 * e33;
  const double s22 = mc.C12 * e11 + mc.C11 * e22 + mc.C12 * e33;
  const double s33 = mc.C12 * e11 + mc.C12 * e22 + mc.C11 * e33;
  const double s23 = mc.C44 * (2.0 * E[7]);
  const double s13 = mc.C44 * (2.0 * E[6]);
  const double s12 = mc.C44 * (2.0 * E[1]);
  S[0] = s11; S[1] = s12; S[2] = s13;
  S[3] = s12; S[4] = s22; S[5] = s23;
  S[6] = s13; S[7] = s23; S[8] = s33;
}

// tau_s = Mp : P_s with Mp = S (Fi = I), for all systems: tau_s = v d_s . (S n_s),
// the three systems of a plane sharing S n (signed sums: every entry of d_s and
// n_s is 0 or +-1)
__device__ __forceinline__ void resolved_shears(const double* S, double v, double* tau) {
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    double u[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double t = SLIP[3 * p][3] > 0 ? S[3 * i] : -S[3 * i];
#pragma unroll
      for (int j = 1; j < 3; ++j) t = SLIP[3 * p][3 + j] > 0 ? t + S[3 * i + j] : t - S[3 * i + j];
      u[i] = t;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int s = 3 * p + k;
      double t = 0.0;
      bool first = true;
#pragma unroll
      for (int i = 0; i < 3; ++i)
        if (SLIP[s][i] != 0) {
          t = first ? (SLIP[s][i] > 0 ? u[i] : -u[i]) : (SLIP[s][i] > 0 ? t + u[i] : t - u[i]);
          first = false;
        }
      tau[s] = v * t;
    }
  }
}

// Lp = v sum_s g_s d_s n_s^T, the systems of a plane summed first:
// Lp = v sum_p (sum_{s on p} g_s d_s) n_p^T (row-major 3x3 out)
__device__ __forceinline__ void slip_sum(const double* g, double v, double* L) {
  double w[4][3];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double t = 0.0;
      bool first = true;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int s = 3 * p + k;
        if (SLIP[s][i] != 0) {
          t = first ? (SLIP[s][i] > 0 ? g[s] : -g[s]) : (SLIP[s][i] > 0 ? t + g[s] : t - g[s]);
          first = false;
        }
      }
      w[p][i] = t;
    }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double t = SLIP[0][3 + j] > 0 ? w[0][i] : -w[0][i];
#pragma unroll
      for (int p = 1; p < 4; ++p) t = SLIP[3 * p][3 + j] > 0 ? t + w[p][i] : t - w[p][i];
      L[3 * i + j] = v * t;
    }
}

// kinetics_sl (phenopowerlaw.f90:444-483): tau_pos = t, tau_neg = -t, taking
// the reciprocal slip resistances rxi = 1/xi (xi is fixed during the Newton
// iteration, so the divisions are done once per state iteration).
// dLp/dMp = sum_s c_s P_s (x) P_s with c_s = |gdot_s| n / |tau_s|; for a
// compile-time exponent this is gdot0 n x^(n-1) / xi, x = |tau_s| / xi, taken
// from the same powering chain without a division.
template <int NE4>
__device__ __forceinline__ void kinetics(const double* S, const double* rxi, double* dg, double* c,
                                         const MatConst& mc) {
  constexpr double TINY10 = 2.2250738585072014e-308 * 10.0;
  double tau[NSLIP];
  resolved_shears(S, mc.v, tau);
#pragma unroll
  for (int s = 0; s < NSLIP; ++s) {
    const double tp = tau[s];
    const double tn = -tp;
    const double sgn = (tp > tn) ? 1.0 : -1.0;
    const double tmax = fmax(tp, tn);
    const double x = tmax * rxi[s];
    if constexpr (NE4 >= 4) {
      const double xm1 = powc<NE4 - 4>(x);
      const double g = sgn * mc.gdot0 * (xm1 * x);
      dg[s] = g;
      if (c) c[s] = (fabs(g) <= TINY10) ? 0.0 : mc.gdot0 * mc.n * xm1 * rxi[s];
    } else {
      const double g = sgn * mc.gdot0 * powe<NE4>(x, mc.n_e4, mc.n);
      dg[s] = g;
      if (c) {
        const double dd = (fabs(g) <= TINY10) ? 0.0 : g * mc.n / tmax;
        c[s] = dd * ((g > 0.0) ? 1.0 : -1.0);
      }
    }
  }
}

// phenopowerlaw_dotState (phenopowerlaw.f90:351-394) for state (xi, gamma),
// rxi = 1/xi of the same state
template <int NE4, int AE4>
__device__ __forceinline__ void dot_state(const double* S, const double* st, const double* rxi, double* dot,
                                          const MatConst& mc) {
  double dg[NSLIP];
  kinetics<NE4>(S, rxi, dg, nullptr, mc);
  const double rxinf = 1.0 / mc.xi_inf;
#pragma unroll
  for (int s = 0; s < NSLIP; ++s) dg[s] = fabs(dg[s]);
#pragma unroll
  for (int s = 0; s < NSLIP; ++s) {
    const double x = 1.0 - st[s] * rxinf;
    const double left = copysign(powe<AE4>(fabs(x), mc.a_e4, mc.a), x);
    double hg = 0.0;
#pragma unroll
    for (int k = 0; k < NSLIP; ++k) hg += mc.hv[HTYPE[s][k] - 1] * dg[k];
    dot[s] = mc.h0 * 1.0 * left * hg;
  }
#pragma unroll
  for (int s = 0; s < NSLIP; ++s) dot[NSLIP + s] = dg[s];
}

// The Newton system J x = b of the Lp iteration lives in the traceless
// subspace: every P_s is traceless, so J maps traceless matrices to traceless
// ones, and b = Lp - Lp_constitutive is traceless.  Its solution is therefore
// traceless and solves the 8x8 system on the coordinates
//   y_k = x_k (k = 1,2,3,5,6,7, the off-diagonal row-major positions),
//   y_0 = x_00, y_4 = x_11 (x_22 = -y_0 - y_4),
// taking the rows 0..7 of J (row 8 = -(row 0 + row 4) on this subspace) and,
// for the columns 0 and 4, J[:,k] - J[:,8].  The 9x9 solution is recovered
// exactly; this is the same linear solve with an 8x8 factorisation.
constexpr int NR = 8;

// Solve J y = b (8x8, row-major) by LU with partial pivoting (first largest
// entry of the column, as LAPACK's idamax).  Used only by the pivoting pass
// (its own kernel), with J in registers.
__device__ __forceinline__ bool solve8_pivot(double* J, double* b) {
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    int p = k;
    double mx = fabs(J[NR * k + k]);
#pragma unroll
    for (int i = k + 1; i < NR; ++i) {
      const double v = fabs(J[NR * i + k]);
      if (v > mx) {
        mx = v;
        p = i;
      }
    }
    if (!(mx > 0.0)) return false;
    // exchange rows k and p by selects over the candidate rows (static
    // register indices; the columns left of k are not read again)
#pragma unroll
    for (int i = k + 1; i < NR; ++i) {
      const bool sw = p == i;
#pragma unroll
      for (int c = k; c < NR; ++c) {
        const double u = J[NR * k + c], w = J[NR * i + c];
        J[NR * k + c] = sw ? w : u;
        J[NR * i + c] = sw ? u : w;
      }
      const double u = b[k], w = b[i];
      b[k] = sw ? w : u;
      b[i] = sw ? u : w;
    }
    const double inv = 1.0 / J[NR * k + k];
#pragma unroll
    for (int i = k + 1; i < NR; ++i) {
      const double l = J[NR * i + k] * inv;
#pragma unroll
      for (int c = k + 1; c < NR; ++c) J[NR * i + c] -= l * J[NR * k + c];
      b[i] -= l * b[k];
    }
  }
#pragma unroll
  for (int i = NR - 1; i >= 0; --i) {
    double s = b[i];
#pragma unroll
    for (int c = i + 1; c < NR; ++c) s -= J[NR * i + c] * b[c];
    b[i] = s / J[NR * i + i];
  }
  return true;
}

// Column-major plane q = i+3j of a (3,3) field <-> row-major register 3*i+j
__device__ __forceinline__ int rm(int q) { return 3 * (q % 3) + q / 3; }

__device__ __forceinline__ void load33(const double* base, int64_t N, int64_t c, double* m) {
#pragma unroll
  for (int q = 0; q < 9; ++q) m[rm(q)] = base[q * N + c];
}

// Plane addressing by a uniform plane pointer plus the lane's 32-bit byte
// offset (scalar base + vector offset): one offset register serves every
// plane, instead of a 64-bit lane address per plane.  Needs (c_end) * 8 < 2^32,
// which the launcher checks.
//
// Every plane pointer given to ld/st/load33/store33 with a Lane must be the
// same in all lanes (formed from kernel arguments and uniform indices only):
// it is read from the first active lane, which keeps the compiler from
// folding the lane offset into a 64-bit per-lane address for every plane.
// A per-lane choice between two planes is made on the loaded values
// (load33_sel), never on the pointer.
struct Lane {
  uint32_t off;  // cell index * 8
};
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
typedef __attribute__((address_space(1))) double gdouble;  // global memory
__device__ __forceinline__ double ld(const double* plane, Lane l) {
  return *reinterpret_cast<const gdouble*>(uniform_u64(reinterpret_cast<uint64_t>(plane)) + l.off);
}
__device__ __forceinline__ void st(double* plane, Lane l, double v) {
  *reinterpret_cast<gdouble*>(uniform_u64(reinterpret_cast<uint64_t>(plane)) + l.off) = v;
}
__device__ __forceinline__ void load33(const double* base, int64_t N, Lane l, double* m) {
#pragma unroll
  for (int q = 0; q < 9; ++q) m[rm(q)] = ld(base + q * N, l);
}
__device__ __forceinline__ void store33(double* base, int64_t N, Lane l, const double* m) {
#pragma unroll
  for (int q = 0; q < 9; ++q) st(base + q * N, l, m[rm(q)]);
}
// load from a plane chosen per lane (the base may differ between lanes): a
// 64-bit lane address and one unconditional load, so that the loads of all
// lanes and planes issue together instead of in divergent branches
__device__ __forceinline__ double ldv(const double* plane, Lane l) {
  return *reinterpret_cast<const gdouble*>(reinterpret_cast<uint64_t>(plane) + l.off);
}
// per-lane choice of the plane set: sel ? a : b
__device__ __forceinline__ void load33_sel(bool sel, const double* a, const double* b, int64_t N, Lane l,
                                           double* m) {
  const double* base = sel ? a : b;
#pragma unroll
  for (int q = 0; q < 9; ++q) m[rm(q)] = ldv(base + q * N, l);
}

// J = I + dt sum_s c_s vec(P_s) vec(G^T K_s)^T with G = Fe^T A, restricted to
// the traceless coordinates (rows 0..7; columns 0 and 4 minus column 8), as
// an 8x8 row-major matrix.
//
// K_s = P_s : C for the cubic stiffness (tr P_s = 0) is K_s(m,l) =
// kappa(m,l) k_s(m,l) with kappa = 2 kA + kB on the diagonal and kA off it,
// kA = v C44, kB = v (C11 - C12 - 2 C44).  With f_s = dt v c_s, row r = 3i+j is
//   J(r,(p,l)) = delta + sum_m kappa(m,l) G(m,p) Q_r(m,l),
//   Q_r(m,l) = sum_s f_s P_s(i,j) k_s(m,l)   (symmetric in m,l),
// and every Q_r(m,l) is a signed sum of one or two of the twelve Hadamard sums
// of the direction groups' f_s (qhad): 24 additions for all the Q_r instead of
// a rank-one update of the 8x8 matrix per slip system.

// Row-independent part of the Jacobian: kappa G (or, LEAN, G alone, kappa
// then scaling Q in every row: 18 registers fewer for 48 more products) and
// the Hadamard sums
template <bool LEAN>
struct JacobianParts {
  double Gd[9], Go[LEAN ? 1 : 9];  // kappa G for the diagonal / off-diagonal (m,l); LEAN: G in Gd
  double H[3][4];                  // Hadamard sums of f_s over each direction group
};
template <bool LEAN>
__device__ __forceinline__ void jacobian_parts(const double* Fe, const double* A, const double* c, double dt,
                                               JacobianParts<LEAN>& jp, const MatConst& mc) {
  const double kA = mc.v * mc.C44, kB = mc.v * (mc.C11 - mc.C12 - 2.0 * mc.C44);
  const double kD = 2.0 * kA + kB;
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const double g = Fe[m] * A[p] + Fe[3 + m] * A[3 + p] + Fe[6 + m] * A[6 + p];
      if constexpr (LEAN) {
        jp.Gd[3 * m + p] = g;
      } else {
        jp.Gd[3 * m + p] = g * kD;
        jp.Go[3 * m + p] = g * kA;
      }
    }
  const double fv = dt * mc.v;
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    const double a = c[g] * fv, b = c[g + 3] * fv, x = c[g + 6] * fv, y = c[g + 9] * fv;
    const double ab = a + b, amb = a - b, xy = x + y, xmy = x - y;
    jp.H[g][0] = ab + xy;
    jp.H[g][1] = ab - xy;
    jp.H[g][2] = amb + xmy;
    jp.H[g][3] = amb - xmy;
  }
}
// row r of the 8x8 Jacobian
template <bool LEAN>
__device__ __forceinline__ void jacobian_row(const JacobianParts<LEAN>& jp, int r, double* row, const MatConst& mc) {
  double Q[3][3];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int l = m; l < 3; ++l) {
      double t = 0.0;
      bool first = true;
#pragma unroll
      for (int g = 0; g < 3; ++g) {
        const int code = qhad(g, r, m, l);
        if (code != 0) {
          const double h = jp.H[g][(code > 0 ? code : -code) - 1];
          t = first ? (code > 0 ? h : -h) : (code > 0 ? t + h : t - h);
          first = false;
        }
      }
      if constexpr (LEAN) {
        const double kA = mc.v * mc.C44, kB = mc.v * (mc.C11 - mc.C12 - 2.0 * mc.C44);
        t = t * (m == l ? 2.0 * kA + kB : kA);
      }
      Q[m][l] = t;
      Q[l][m] = t;
    }
  double Y[9];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      if constexpr (LEAN) {
        double y = jp.Gd[p] * Q[0][l];
        y += jp.Gd[3 + p] * Q[1][l];
        y += jp.Gd[6 + p] * Q[2][l];
        Y[3 * p + l] = y;
      } else {
        double y = (l == 0 ? jp.Gd : jp.Go)[p] * Q[0][l];
        y += (l == 1 ? jp.Gd : jp.Go)[3 + p] * Q[1][l];
        y += (l == 2 ? jp.Gd : jp.Go)[6 + p] * Q[2][l];
        Y[3 * p + l] = y;
      }
    }
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const double y = (q == 0 || q == 4) ? Y[q] - Y[8] : Y[q];
    row[q] = (q == r) ? 1.0 + y : y;
  }
}
__device__ __forceinline__ void assemble_jacobian8(const double* Fe, const double* A, const double* c, double dt,
                                                   double* J, const MatConst& mc) {
  JacobianParts<true> jp;
  jacobian_parts(Fe, A, c, dt, jp, mc);
#pragma unroll
  for (int r = 0; r < NR; ++r) jacobian_row(jp, r, J + NR * r, mc);
}

// The Newton system assembled and solved by LU without row exchanges, one row
// at a time: row i is eliminated by the factor rows 0..i-1 as soon as it is
// formed (the same operations on every entry, in the same order, as the
// column-by-column elimination), so only the factor rows are held.  Refuses
// (returns false; b then holds no meaning) when a multiplier exceeds 1e6 in
// magnitude or a pivot's magnitude is below 2^-1000 (zero and NaN included:
// the pivot's reciprocal is the refined hardware one, so a pivot near the
// bottom of the range goes to the pivoted pass and its exact division; the
// pivots of this system, 1 plus rate terms, are never near the top of the
// range): multipliers up to 1e6 cost at most a few of
// the 16 digits of the Newton step, far below the Newton tolerance, and the
// residual test of the next iteration uses the exact residual; only steeper
// cases (about 1 % of points had pivots below 1/100 of their column in the
// 128^3 tension case) need the pivoted pass.
__device__ __forceinline__ bool assemble_solve8_nopivot(const double* Fe, const double* A, const double* c,
                                                        double dt, double* b, const MatConst& mc) {
  JacobianParts<false> jp;
  jacobian_parts(Fe, A, c, dt, jp, mc);
  double U[NR * NR];  // factor rows (upper part), reciprocal pivots on the diagonal
  // refusals are collected and tested once at the end (what is computed
  // after one is discarded): no branches inside the unrolled elimination, and
  // the multipliers' magnitudes enter one running maximum (a NaN multiplier
  // is ignored by the maximum, as by the comparison it replaces)
  bool refused = false;
  double lmax = 0.0;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    double row[NR];
    jacobian_row(jp, i, row, mc);
#pragma unroll
    for (int k = 0; k < i; ++k) {
      const double l = row[k] * U[NR * k + k];
      lmax = fmax(lmax, fabs(l));
#pragma unroll
      for (int c2 = k + 1; c2 < NR; ++c2) row[c2] -= l * U[NR * k + c2];
      b[i] -= l * b[k];
    }
    const double piv = row[i];
    refused = refused || !(fabs(piv) >= 0x1p-1000);
    U[NR * i + i] = rcp_refined(piv);
#pragma unroll
    for (int c2 = i + 1; c2 < NR; ++c2) U[NR * i + c2] = row[c2];
  }
#pragma unroll
  for (int i = NR - 1; i >= 0; --i) {
    double s = b[i];
#pragma unroll
    for (int c2 = i + 1; c2 < NR; ++c2) s -= U[NR * i + c2] * b[c2];
    b[i] = s * U[NR * i + i];
  }
  return !(refused || lmax > 1e6);
}

// Per-point output / scratch addressing
struct PointIO {
  const MpState* s;
  int64_t N;
  Lane l;
};

// integrateStress (phase_mechanical.f90:380-604); returns 0 (integrated),
// 1 (broken) or, without PIVOT, 2 when the LU without row exchanges refused a
// pivot (the point is then redone by the PIVOT instantiation, which solves
// every Newton system with partial pivoting).  Takes the reciprocal slip
// resistances of the state and updates S and Lp in registers; Fe, P and Fp
// of the integration the state loop keeps are formed and stored by
// store_integrated.
//
// A = F invFp0 of the (sub)step and invFp0 (both in LDS, lane t) are formed
// once per substep by the caller: they do not change during the state fixed
// point.
template <int NE4, bool PIVOT>
__device__ __forceinline__ int integrate_stress(const double (*hA)[64], const double (*invFp0)[64], int t, double dt,
                                                const double* rxi, double* Lp, double* S, int& newton,
                                                const MatConst& mc) {
  // A is read from LDS where it is used, not hoisted into the registers the
  // Newton system needs: the lane index passes through an empty asm, so every
  // call's nine ds_reads are new loads (issued together, one wait), not the
  // serialised one-wait-each accesses a volatile read would be
  auto loadA = [&](double* A) {
    int tl = t;
    asm volatile("" : "+v"(tl));
#pragma unroll
    for (int k = 0; k < 9; ++k) A[k] = hA[k][tl];
  };
  double Lpg_old[9], dLp[9];
  double res_old_n2 = 0.0;  // squared norms: the tests compare non-negative values
  double step = 1.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    Lpg_old[k] = Lp[k];
    dLp[k] = 0.0;
  }
  int it = 0;
  // B, Fe and S of an iteration are not carried across the linear solve: the
  // converged ones are formed again from the converged Lp after the loop (S
  // is an output only on success)
  for (;;) {
    it++;
    if (it > mc.N_iter_Lp_max) return 1;
    double Fe[9], Sl[9];
    {
      double B[9], A[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) B[k] = (k % 4 == 0 ? 1.0 : 0.0) - dt * Lp[k];
      loadA(A);
      mm33(A, B, Fe);  // Fe = A B invFi_new, invFi_new = I
    }
    hooke_S(Fe, Sl, mc);
    double dg[NSLIP], c[NSLIP];
    kinetics<NE4>(Sl, rxi, dg, c, mc);
    double res[9];
    slip_sum(dg, mc.v, res);
    // res holds Lp_constitutive here
    double n2Lpc = 0.0, n2Lp = 0.0, n2res = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      n2Lpc += res[k] * res[k];
      n2Lp += Lp[k] * Lp[k];
      res[k] = Lp[k] - res[k];
      n2res += res[k] * res[k];
    }
    {
      // |res| < max(eps_rel max(|Lp|, |Lp_c|), eps_abs), compared squared:
      // |res|^2 < max(eps_rel^2 max(|Lp|^2, |Lp_c|^2), eps_abs^2), no square root
      const double tol2 = fmax((mc.eps_rel_Lp * mc.eps_rel_Lp) * fmax(n2Lp, n2Lpc), mc.eps_abs_Lp * mc.eps_abs_Lp);
      // a sum of squares is NaN exactly when one of its terms is
      if (isnan(n2res)) return 1;
      if (n2res < tol2) break;
      if (it == 1 || n2res < res_old_n2) {
        res_old_n2 = n2res;
#pragma unroll
        for (int k = 0; k < 9; ++k) Lpg_old[k] = Lp[k];
        step = 1.0;
      } else {
        step = mc.r_linesearch_Lp * step;
#pragma unroll
        for (int k = 0; k < 9; ++k) Lp[k] = Lpg_old[k] + dLp[k] * step;
        continue;
      }
    }
    {
      double J[NR * NR], A[9];
      loadA(A);
      if constexpr (PIVOT) {
        assemble_jacobian8(Fe, A, c, dt, J, mc);
        if (!solve8_pivot(J, res)) return 1;  // singular
      } else {
        // the point is redone by the pivoting pass
        if (!assemble_solve8_nopivot(Fe, A, c, dt, res, mc)) return 2;
      }
    }
    res[8] = -(res[0] + res[4]);  // x_22 of the traceless solution
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      dLp[k] = -res[k];
      Lp[k] = Lp[k] + dLp[k] * step;
    }
  }
  newton += it;
  double B[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) B[k] = (k % 4 == 0 ? 1.0 : 0.0) - dt * Lp[k];
  double invFp_new[9], Fpn[9];
  {
    double iF0[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) iF0[k] = invFp0[k][t];
    mm33(iF0, B, invFp_new);
  }
  if (!inv33(invFp_new, Fpn, nullptr)) return 1;
  double T1[9], A[9];
  loadA(A);
  mm33(A, B, T1);  // F invFp_new (= Fe, invFi_new = I)
  hooke_S(T1, S, mc);
  return 0;
}

// Fe, P and Fp of the last successful integrate_stress (same operations on
// the same Lp, A and invFp0, hence the same values), stored when the state
// fixed point leaves with it; integrations that a later state iteration
// supersedes store nothing
__device__ __forceinline__ void store_integrated(const double (*hA)[64], const double (*invFp0)[64], int t,
                                                 double dt, const double* Lp, const double* S, const PointIO& io) {
  const int64_t N = io.N;
  const Lane l = io.l;
  double A[9], B[9], invFp_new[9], Fpn[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    A[k] = hA[k][t];
    B[k] = (k % 4 == 0 ? 1.0 : 0.0) - dt * Lp[k];
  }
  {
    double iF0[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) iF0[k] = invFp0[k][t];
    mm33(iF0, B, invFp_new);
  }
  inv33(invFp_new, Fpn, nullptr);  // invertible: integrate_stress checked it
  double T1[9];
  mm33(A, B, T1);
  store33(io.s->Fe, N, l, T1);
  {
    double iFt[9], T2[9], P[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) iFt[3 * i + j] = invFp_new[3 * j + i];
    mm33(S, iFt, T2);
    mm33(T1, T2, P);
    store33(io.s->P, N, l, P);
  }
  const double d = pow(det33(Fpn), 1.0 / 3.0);
  const double rd = rcp_refined(d);
#pragma unroll
  for (int k = 0; k < 9; ++k) Fpn[k] = div_rcp(Fpn[k], d, rd);
  store33(io.s->Fp, N, l, Fpn);
}

}  // namespace

// Substep scratch layout (per point, planes of N doubles): F0 (9), Lp0 (9),
// Fp0 (9), state0 (24) of the current substep start, and F of the substep
// being integrated (9).  Touched only when the first full step was cut back.
constexpr int SUB_F0 = 0, SUB_LP0 = 9, SUB_FP0 = 18, SUB_ST0 = 27, SUB_FLOC = 51, SUB_PLANES = 60;

// One residual evaluation's material-point loop, in two passes.  The first
// (mp_response_kernel) carries no pivoted factorisation, so that its
// registers serve the common path only; a point whose Newton system the LU
// without row exchanges refuses leaves its loops, stores no output and is
// appended to the repivot list (its inputs are unchanged: Lp, the only plane
// that is both read and written, is stored last).  The second
// (mp_repivot_kernel) redoes exactly the listed points from the same inputs,
// every Newton system of them solved with partial pivoting (as LAPACK's
// dgesv in the reference), in a grid-stride loop over the list.  Points are
// independent, so the list order does not change any result.
//
// repivot: [0] = number of listed points, [1 + i] = cell index.

// per-lane LDS columns of one 64-lane block: FPI history
// (dotState_last(:,1), its difference to dotState_last(:,2)) and the
// accumulated-shear half of the state stay in LDS while the Lp Newton
// iteration needs the registers
struct MpLds {
  double last1[NSTATE][64];
  double d12[NSTATE][64];
  double gam[NSLIP][64];
  double iFp0[9][64];  // invFp0 of the current (sub)step
  double A[9][64];     // F_(sub)step invFp0
};

template <int NE4, int AE4, bool PIVOT>
__device__ __forceinline__ void mp_point(const KernelArgs& a, const int64_t c, MpLds& sh) {
  const MatConst& mc = a.mc;
  const MpState& s = a.s;
  const int64_t N = s.N;
  const int t = threadIdx.x;
  double (&h_last1)[NSTATE][64] = sh.last1;
  double (&h_d12)[NSTATE][64] = sh.d12;
  double (&h_gam)[NSLIP][64] = sh.gam;
  const Lane l{(uint32_t)(c * 8)};
  const PointIO io{&s, N, l};
  double* sub = a.sub;
  double Lp[9], S[9];
  load33(s.Lp, N, l, Lp);
  // S, xi and the accumulated shear are set by the restore that opens the
  // first step (phase_restore: S = S0, Fp = Fp0, state = state0)
  double xi[NSLIP];
  // substep start values live in memory: the increment's converged values
  // (F0, Lp0, Fp0, state0) until a partial substep succeeded, then the
  // substep scratch planes
  bool in_sub = false;
  int newton = 0, fpi = 0;
  auto st0 = [&](int k) -> double {
    if (k < NSLIP) return ldv((in_sub ? sub + SUB_ST0 * N : s.xi0) + k * N, l);
    return ldv((in_sub ? sub + (SUB_ST0 + NSLIP) * N : s.gamma0) + (k - NSLIP) * N, l);
  };

  // F and F0 of the increment (not written here), loaded once with Lp
  double Fi[9], F0i[9];
  load33(s.F, N, l, Fi);
  load33(s.F0, N, l, F0i);
  // F of the (sub)step: F0l + step (F - F0), phase_mechanical.f90:1093-1096
  auto load_Floc = [&](double st_, double* Fl) {
    double F[9], F0l[9], F0p[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      F[k] = Fi[k];
      F0p[k] = F0i[k];
    }
    if (in_sub) {
      load33(sub + SUB_F0 * N, N, l, F0l);
    } else {
#pragma unroll
      for (int k = 0; k < 9; ++k) F0l[k] = F0p[k];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) Fl[k] = F0l[k] + st_ * (F[k] - F0p[k]);
  };
  double stepFrac = 0.0, step = 1.0 / mc.r_cutback;
  bool converged = false, todo = true;
  bool refused = false;  // first pass only: the Newton system needs row exchanges
  double fp0v[9];        // Fp0 of the (sub)step about to be integrated (planes in order)
  // step control: the next substep after a converged one, else a cut-back
  // step from the restored start values; sets todo
  auto control = [&]() {
    if (converged) {
      stepFrac = stepFrac + step;
      step = fmin(1.0 - stepFrac, mc.r_increase * step);
      todo = step > 0.0;
      if (todo) {
        // all loads before the stores: the compiler cannot tell the planes
        // apart, and a load after a store would wait for the one before it
        double fl[9], fp[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          fl[k] = ld(sub + (SUB_FLOC + k) * N, l);
          fp[k] = ld(s.Fp + (k) * N, l);
        }
        __builtin_amdgcn_sched_barrier(0);  // all issued before the first store (one round trip)
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          st(sub + (SUB_F0 + k) * N, l, fl[k]);
          st(sub + (SUB_LP0 + k) * N, l, Lp[rm(k)]);
          st(sub + (SUB_FP0 + k) * N, l, fp[k]);
          fp0v[k] = fp[k];
        }
#pragma unroll
        for (int k = 0; k < NSLIP; ++k) {
          st(sub + (SUB_ST0 + k) * N, l, xi[k]);
          st(sub + (SUB_ST0 + NSLIP + k) * N, l, h_gam[k][t]);
        }
        in_sub = true;
      }
    } else {
      step = mc.r_cutback * step;
      // every load of the restore is issued before the first of its values
      // is stored (to LDS or memory): one memory round trip.  Left to
      // itself, the scheduler interleaves them with the LDS stores of the
      // shear column through a few staging registers, a round trip per pair.
      double stl[NSTATE];
      load33(s.S0, N, l, S);
#pragma unroll
      for (int k = 0; k < 9; ++k) fp0v[k] = ldv((in_sub ? sub + SUB_FP0 * N : s.Fp0) + k * N, l);
      if (step < 1.0) load33_sel(in_sub, sub + SUB_LP0 * N, s.Lp0, N, l, Lp);
#pragma unroll
      for (int k = 0; k < NSTATE; ++k) stl[k] = st0(k);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < NSLIP; ++k) {
        xi[k] = stl[k];
        h_gam[k][t] = stl[NSLIP + k];
      }
#pragma unroll
      for (int k = 0; k < 9; ++k) st(s.Fp + (k) * N, l, fp0v[k]);  // after the loads, as above
      todo = step > mc.r_cutback_min;
    }
  };
  // first pass: the first call restores the increment's start values; it
  // follows the entry loads directly, so that its loads and theirs share one
  // round trip.  The pivoting pass (a few points per launch, registers
  // tighter) keeps the control at the top of the loop.
  if constexpr (!PIVOT) control();
  for (;;) {
    if constexpr (PIVOT) control();
    if (!todo) break;
    {
      // A = F_(sub)step invFp0 and invFp0 are fixed during the state fixed
      // point (integrateStress, phase_mechanical.f90:425-432)
      bool fp0_singular;
      {
        double Fl[9], Fp0[9], iF0[9], A[9];
        load_Floc(step, Fl);
        if (step < 1.0) store33(sub + SUB_FLOC * N, N, l, Fl);  // start of the next substep
#pragma unroll
        for (int k = 0; k < 9; ++k) Fp0[rm(k)] = fp0v[k];
        fp0_singular = !inv33(Fp0, iF0, nullptr);
        mm33(Fl, iF0, A);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          sh.iFp0[k][t] = iF0[k];
          sh.A[k][t] = A[k];
        }
      }
      const double sdt = step * a.dt;
      // integrateStateFPI (phase_mechanical.f90:611-693)
      bool broken = true;
      double dot[NSTATE], rxi[NSLIP];
#pragma unroll
      for (int k = 0; k < NSLIP; ++k) rxi[k] = rcp_refined(xi[k]);
      {
        double stt[NSTATE];
#pragma unroll
        for (int k = 0; k < NSLIP; ++k) {
          stt[k] = xi[k];
          stt[NSLIP + k] = h_gam[k][t];
        }
        dot_state<NE4, AE4>(S, stt, rxi, dot, mc);
      }
      bool nan = false;
#pragma unroll
      for (int k = 0; k < NSTATE; ++k) nan = nan || isnan(dot[k]);
      if (!nan) {
        // start-of-step state for the whole fixed point: xi and the shear
        // column hold it here (the restore above loaded them from the start
        // planes, or a new substep stored its start planes from them)
        double s0[NSTATE];
#pragma unroll
        for (int k = 0; k < NSTATE; ++k) s0[k] = k < NSLIP ? xi[k] : h_gam[k - NSLIP][t];
#pragma unroll
        for (int k = 0; k < NSLIP; ++k) {
          xi[k] = s0[k] + dot[k] * sdt;
          h_gam[k][t] = s0[NSLIP + k] + dot[NSLIP + k] * sdt;
        }
        for (int it = 1; it <= mc.N_iter_state_max; ++it) {
          fpi++;
          // history shift of the first iteration (last2 = 0); later
          // iterations shift it at the end of the previous one, where last1
          // is already in registers
          if (it == 1) {
#pragma unroll
            for (int k = 0; k < NSTATE; ++k) {
              h_d12[k][t] = dot[k];  // last1 - last2 after the shift
              h_last1[k][t] = dot[k];
            }
          }
#pragma unroll
          for (int k = 0; k < NSLIP; ++k) rxi[k] = rcp_refined(xi[k]);
          const int rc =
              fp0_singular ? 1 : integrate_stress<NE4, PIVOT>(sh.A, sh.iFp0, t, sdt, rxi, Lp, S, newton, mc);
          refused = rc == 2;
          broken = rc != 0;
          if (broken) break;
          {
            double stt[NSTATE];
#pragma unroll
            for (int k = 0; k < NSLIP; ++k) {
              stt[k] = xi[k];
              stt[NSLIP + k] = h_gam[k][t];
            }
            dot_state<NE4, AE4>(S, stt, rxi, dot, mc);
          }
          bool nn = false;
#pragma unroll
          for (int k = 0; k < NSTATE; ++k) nn = nn || isnan(dot[k]);
          if (nn) {  // leaves with the last integration (broken stays false)
            store_integrated(sh.A, sh.iFp0, t, sdt, Lp, S, io);
            break;
          }
          // the rate history is read from LDS in one batch before the sums
          // (issued together: one wait instead of one per state variable)
          double l1v[NSTATE], d12v[NSTATE];
#pragma unroll
          for (int k = 0; k < NSTATE; ++k) {
            l1v[k] = h_last1[k][t];
            d12v[k] = h_d12[k][t];
          }
          double dp12 = 0.0, dp22 = 0.0, dp01 = 0.0;
#pragma unroll
          for (int k = 0; k < NSTATE; ++k) {
            dp12 += (dot[k] - l1v[k]) * d12v[k];
            dp22 += d12v[k] * d12v[k];
            dp01 += dot[k] * l1v[k];
          }
          const double zeta =
              (fmin(dp01, dp12) < 0.0 && dp22 > 0.0) ? 0.75 + 0.25 * tanh(2.0 + 4.0 * dp12 / dp22) : 1.0;
          double gv[NSLIP];  // the shear column, read in one batch as well
#pragma unroll
          for (int k = 0; k < NSLIP; ++k) gv[k] = h_gam[k][t];
          bool conv = true;
#pragma unroll
          for (int k = 0; k < NSTATE; ++k) {
            const double l1 = l1v[k];
            // the damped rate replaces dotState: it is dotState_last(:,1) of
            // the next iteration (phase_mechanical.f90:653-655)
            const double dk = dot[k] * zeta + l1 * (1.0 - zeta);
            dot[k] = dk;
            h_d12[k][t] = dk - l1;  // the next iteration's shift
            h_last1[k][t] = dk;
            const double cur = (k < NSLIP) ? xi[k] : gv[k - NSLIP];
            const double r = cur - s0[k] - dk * sdt;
            const double nw = cur - r;
            if (k < NSLIP) xi[k] = nw;
            else h_gam[k - NSLIP][t] = nw;
            const double at = (k < NSLIP) ? mc.atol_xi : mc.atol_gamma;
            conv = conv && (fabs(r) <= fmax(at, mc.eps_rel_state * fabs(nw)));
          }
          if (conv || it == mc.N_iter_state_max) {
            // converged, or the last iteration (integrateStateFPI then
            // returns the last integration's broken = false)
            store_integrated(sh.A, sh.iFp0, t, sdt, Lp, S, io);
            broken = false;
            break;
          }
        }
      }
      converged = !broken;
      if (refused) break;  // the pivoting pass redoes this point from its inputs
    }
    if constexpr (!PIVOT) control();
  }
  if constexpr (!PIVOT) {
    if (refused) {
      const int slot = atomicAdd(a.repivot, 1);
      a.repivot[1 + slot] = (int)c;
      return;
    }
  }
  store33(s.Lp, N, l, Lp);
  store33(s.S, N, l, S);
  double gv[NSLIP];  // the shear column, read from LDS in one batch
#pragma unroll
  for (int k = 0; k < NSLIP; ++k) gv[k] = h_gam[k][t];
#pragma unroll
  for (int k = 0; k < NSLIP; ++k) {
    st(s.xi + (k) * N, l, xi[k]);
    st(s.gamma + (k) * N, l, gv[k]);
  }
  if (s.iters) {
    s.iters[c] = newton;
    s.iters[N + c] = fpi;
  }
  if (a.broken) a.broken[c] = converged ? 0 : 1;
  if (!converged) atomicOr(a.ill, 1);
}

template <int NE4, int AE4>
__global__ void __launch_bounds__(64) mp_response_kernel(KernelArgs a) {
  __shared__ MpLds sh;
  const int64_t c = a.c_begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.c_end) return;
  mp_point<NE4, AE4, false>(a, c, sh);
}

template <int NE4, int AE4>
__global__ void __launch_bounds__(64) mp_repivot_kernel(KernelArgs a) {
  __shared__ MpLds sh;
  const int64_t cnt = a.repivot[0];  // complete: the first pass ran before this launch
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x)
    mp_point<NE4, AE4, true>(a, a.repivot[1 + i], sh);
}

// clears the repivot list and the terminally-ill flag of one response
__global__ void repivot_reset_kernel(int* repivot, int* ill) {
  repivot[0] = 0;
  *ill = 0;
}

// ---- many RVEs in one launch ---------------------------------------------------
// The material-point updates of a batch of RVEs (each with its own planes,
// material, time step and F) as one launch: block b belongs to the RVE r with
// blk_off[r] <= b < blk_off[r + 1] and takes its cells (b - blk_off[r]) * 64
// on.  Every point runs exactly the code of mp_response_kernel, so each RVE's
// results are bitwise those of its own launch.  The table is read only
// (restrict): its fields are loaded as scalars and never reloaded after the
// kernel's stores.
__device__ __forceinline__ int batch_entry(const int* __restrict__ blk_off, int n, unsigned b) {
  int lo = 0, hi = n;  // blk_off[lo] <= b < blk_off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if ((unsigned)blk_off[mid] <= b) lo = mid;
    else hi = mid;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// A table entry read through the constant address space: the loads are
// invariant for the kernel, as kernel arguments are, so the compiler may
// reload any field where it is used instead of keeping it in registers.
__device__ __forceinline__ KernelArgs load_args(const KernelArgs* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(4))) const KernelArgs ConstArgs;
  return *(const ConstArgs*)p;
#else
  return *p;
#endif
}

template <int NE4, int AE4>
__global__ void __launch_bounds__(64) mp_response_batch_kernel(const KernelArgs* __restrict__ tab,
                                                                const int* __restrict__ blk_off, int n) {
  __shared__ MpLds sh;
  const int r = batch_entry(blk_off, n, blockIdx.x);
  const KernelArgs a = load_args(tab + r);
  const int64_t c = a.c_begin + (int64_t)(blockIdx.x - (unsigned)blk_off[r]) * 64 + threadIdx.x;
  if (c >= a.c_end) return;
  mp_point<NE4, AE4, false>(a, c, sh);
}

// pivoting pass of a batch: blockIdx.y is the RVE, blockIdx.x strides over
// its list as mp_repivot_kernel does
template <int NE4, int AE4>
__global__ void __launch_bounds__(64) mp_repivot_batch_kernel(const KernelArgs* __restrict__ tab) {
  __shared__ MpLds sh;
  const KernelArgs a = load_args(tab + blockIdx.y);
  const int64_t cnt = a.repivot[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x)
    mp_point<NE4, AE4, true>(a, a.repivot[1 + i], sh);
}

__global__ void repivot_reset_batch_kernel(const KernelArgs* __restrict__ tab, int n) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  tab[r].repivot[0] = 0;
  *tab[r].ill = 0;
}


// Block partial sums of P for P_av (deterministic): block b sums cells
// [b*PAV_CELLS, (b+1)*PAV_CELLS) of each of the 9 planes, every thread its
// PAV_CELLS/256 cells at stride 256 in order, then one fixed tree over the
// 256 threads for all 9 sums at once.
constexpr int PAV_CELLS = 2048;
__device__ __forceinline__ void pav_block(const double* P, int64_t N, double* partial, int nb, unsigned blk,
                                          double (*red)[256]) {
  const int t = threadIdx.x;
  const int64_t c0 = (int64_t)blk * PAV_CELLS + t;
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = 0.0;
  if ((int64_t)(blk + 1) * PAV_CELLS <= N) {
    // interior block: no per-element test, so every load can be issued
    // ahead of the sums (same summation order)
#pragma unroll
    for (int u = 0; u < PAV_CELLS / 256; ++u) {
      const int64_t c = c0 + (int64_t)u * 256;
#pragma unroll
      for (int q = 0; q < 9; ++q) acc[q] += P[q * N + c];
    }
  } else {
#pragma unroll
    for (int u = 0; u < PAV_CELLS / 256; ++u) {
      const int64_t c = c0 + (int64_t)u * 256;
      if (c < N) {
#pragma unroll
        for (int q = 0; q < 9; ++q) acc[q] += P[q * N + c];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) red[q][t] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int q = 0; q < 9; ++q) red[q][t] += red[q][t + w];
    }
    __syncthreads();
  }
  if (t < 9) partial[(int64_t)t * nb + blk] = red[t][0];
}

__global__ void __launch_bounds__(256) pav_partial_kernel(const double* P, int64_t N, double* partial) {
  __shared__ double red[9][256];
  pav_block(P, N, partial, gridDim.x, blockIdx.x, red);
}

// P block sums of a batch: the blocks of RVE r follow blk_off[r], each as in
// its own launch (same cells, same order)
__global__ void __launch_bounds__(256) pav_partial_batch_kernel(const MpBatchPav* __restrict__ tab,
                                                                 const int* __restrict__ blk_off, int n) {
  __shared__ double red[9][256];
  const int r = batch_entry(blk_off, n, blockIdx.x);
  pav_block(tab[r].P, tab[r].N, tab[r].partial, tab[r].nb, blockIdx.x - (unsigned)blk_off[r], red);
}

// ---------------------------------------------------------------------------
// Consistent tangent dPdF (phase_mechanical_dPdF, phase_mechanical.f90:1137-1258)
// with Fi = I, dLi/dS = 0 (so dFi/dS = 0), using the rank-12 structure of
// dLp/dS = sum_s c_s P_s (x) P_s.  Output: 81 planes of N doubles, plane
// index i+3j+9k+27l for dPdF(i,j,k,l) (column-major as the reference).
// ---------------------------------------------------------------------------
namespace {
struct TangentArgs {
  MpState s;
  MatConst mc;
  double dt;
  double* dPdF;  // MODE 0: 81 planes of N (plane k+3l+9p+27o); MODE 2: 81 values per listed cell
  double* n2;    // MODE 1: per cell |dPdF|^2 (the 81 squares summed in the order they are formed)
  double* bpart; // MODE 1: per 16 cells the sums of the 81 planes, at plane * nblk + 4 * block + row
  int64_t nblk;  // MODE 1: 4 x the blocks of 64 cells
  const int64_t* cells;    // MODE 2: the listed cells
  int64_t c_begin, c_end;  // cells [c_begin, c_end); MODE 2: c_end - c_begin listed cells (at most 64)
};

// (C : X)(i,j) = sum_kl C(i,j,k,l) X(k,l) for the cubic stiffness
// (math_Voigt66to3333_stiffness of C11, C12, C44): C11 X_ii + C12 sum_{k!=i}
// X_kk on the diagonal, C44 (X_ij + X_ji) off it (row-major 3x3)
__device__ __forceinline__ void cubic_contract(const double* X, double* Y, const MatConst& mc) {
  const double t = X[0] + X[4] + X[8];
  const double d = mc.C11 - mc.C12;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      Y[3 * i + j] = (i == j) ? mc.C12 * t + d * X[4 * i] : mc.C44 * (X[3 * i + j] + X[3 * j + i]);
}
}  // namespace

// x[j] += x[j] of the lane DPP control CTRL selects, for nine doubles (the
// halves moved separately; all lanes take part)
template <int CTRL>
__device__ __forceinline__ void dpp_step(double* x) {
  int lo[9], hi[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    lo[j] = __builtin_amdgcn_update_dpp(0, __double2loint(x[j]), CTRL, 0xF, 0xF, false);
    hi[j] = __builtin_amdgcn_update_dpp(0, __double2hiint(x[j]), CTRL, 0xF, 0xF, false);
  }
#pragma unroll
  for (int j = 0; j < 9; ++j) x[j] += __hiloint2double(hi[j], lo[j]);
}

// NE4: the slip-rate exponent n_sl as a compile-time quarter (-1: runtime).
// MODE 0: dPdF of every cell into 81 planes.  MODE 1 (the solver's
// C_volAvg and C_minMaxAvg): no dPdF field, only each cell's |dPdF|^2 and
// per 16-cell row of a block the sums of the 81 planes, in a fixed order.
// MODE 2:
// dPdF of up to 64 listed cells (one block), e.g. the two extreme cells
// that MODE 1 does not keep: the same per-lane arithmetic, so the same values.
template <int NE4, int MODE>
__global__ void __launch_bounds__(64) mp_tangent_kernel(TangentArgs a) {
  const MatConst& mc = a.mc;
  const MpState& s = a.s;
  const int64_t N = s.N;
  int64_t c;
  bool live;
  if constexpr (MODE == 2) {
    if ((int64_t)threadIdx.x >= a.c_end - a.c_begin) return;
    c = a.cells[threadIdx.x];
    live = true;
  } else {
    c = a.c_begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    live = c < a.c_end;
    if constexpr (MODE == 0) {
      if (!live) return;
    } else {
      // the whole wave takes part in the block sums: a lane past the end
      // computes the last cell again and contributes nothing
      if (!live) c = a.c_end - 1;
    }
  }
  [[maybe_unused]] double n2acc = 0.0;
  const double dt = a.dt;
  double F[9], Fp[9], Fp0[9], S[9], Fe[9], xi[NSLIP];
  load33(s.F, N, c, F);
  load33(s.Fp, N, c, Fp);
  load33(s.Fp0, N, c, Fp0);
  load33(s.S, N, c, S);
  load33(s.Fe, N, c, Fe);
#pragma unroll
  for (int k = 0; k < NSLIP; ++k) xi[k] = s.xi[k * N + c];
  double invFp[9], invFp0[9];
  if (!inv33(Fp, invFp, nullptr)) {
#pragma unroll
    for (int k = 0; k < 9; ++k) invFp[k] = 0.0;
  }
  if (!inv33(Fp0, invFp0, nullptr)) {
#pragma unroll
    for (int k = 0; k < 9; ++k) invFp0[k] = 0.0;
  }
  double dg[NSLIP], cs[NSLIP], rxi[NSLIP];
#pragma unroll
  for (int k = 0; k < NSLIP; ++k) rxi[k] = 1.0 / xi[k];
  kinetics<NE4>(S, rxi, dg, cs, mc);

  // L = I + dt sum_s c_s vec(U_s) vec(P_s)^T, U_s = C : (H P_s), H = Fe^T F invFp0
  double T2[9], H[9];
  mm33(F, invFp0, T2);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) H[3 * i + j] = Fe[i] * T2[j] + Fe[3 + i] * T2[3 + j] + Fe[6 + i] * T2[6 + j];
  double L[81];
#pragma unroll
  for (int k = 0; k < 81; ++k) L[k] = (k % 10 == 0) ? 1.0 : 0.0;
#pragma unroll
  for (int sl = 0; sl < NSLIP; ++sl) {
    double Z[9];  // H P_s
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        double t = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m)
          if (psg(sl, m, l) != 0) t += H[3 * i + m] * (mc.v * (double)psg(sl, m, l));
        Z[3 * i + l] = t;
      }
    double U[9];
    cubic_contract(Z, U, mc);
    const double f = dt * cs[sl];
#pragma unroll
    for (int q = 0; q < 9; ++q)
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int o = 0; o < 3; ++o)
          if (psg(sl, p, o) != 0) L[9 * q + 3 * p + o] += f * U[q] * (mc.v * (double)psg(sl, p, o));
  }
  // dSdF = L^-1 R with R(:, (k,l)) = (Fe C(:,:,:,:) invFp^T)(k,l): L is
  // factorised once (partial pivoting, as LAPACK's dgesv: first largest entry
  // of the column; the multipliers stay where their step formed them, the row
  // exchanged at step k in piv[k]), and each column of R is formed, solved and
  // used in turn, so that R is never held whole.  Each entry of a column
  // receives the same exchanges and eliminations, in the same order, as in
  // the elimination of L and R together.
  int piv[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    int p = k;
    double mx = fabs(L[9 * k + k]);
#pragma unroll
    for (int i = k + 1; i < 9; ++i) {
      const double v = fabs(L[9 * i + k]);
      if (v > mx) {
        mx = v;
        p = i;
      }
    }
    piv[k] = p;
#pragma unroll
    for (int i = k + 1; i < 9; ++i) {
      const bool sw = (p == i);
#pragma unroll
      for (int cc = k; cc < 9; ++cc) {
        const double x = L[9 * k + cc], y = L[9 * i + cc];
        L[9 * k + cc] = sw ? y : x;
        L[9 * i + cc] = sw ? x : y;
      }
    }
    const double inv = 1.0 / L[9 * k + k];
    L[9 * k + k] = inv;  // the reciprocal pivot, for the back substitution
#pragma unroll
    for (int i = k + 1; i < 9; ++i) {
      const double l = L[9 * i + k] * inv;
      L[9 * i + k] = l;
#pragma unroll
      for (int cc = k + 1; cc < 9; ++cc) L[9 * i + cc] -= l * L[9 * k + cc];
    }
  }
  // dPdF = d(Fe S Fp^-T ...)/dF, column q = (p,o) of F at a time:
  // dFpinvdF(:,:,p,o) = -dt sum_s c_s V_s(p,o) (invFp0 P_s)
  //                   = invFp0 (sum_s w_s P_s),  w_s = -dt c_s V_s(p,o),
  // V_s(p,o) = P_s : dSdF(:,:,p,o)
  double SiT[9], iFt[9], FiFp[9], u3[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) iFt[3 * i + j] = invFp[3 * j + i];
  mm33(S, iFt, SiT);        // S invFp^T
  mm33(F, invFp, FiFp);     // F invFp
  mm33(FiFp, S, u3);        // F invFp S
  double d0[9], tmp[9];
  mm33(invFp, SiT, tmp);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) d0[3 * i + j] = tmp[3 * j + i];
  // a runtime loop over the rows p of F (three columns of R at a time): the
  // three columns' solves are independent chains that hide each other's
  // latency, and fewer columns in flight keep the registers within bounds
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    double fe[3];  // row p of Fe (chosen by value: no register array indexed at run time)
#pragma unroll
    for (int m = 0; m < 3; ++m) fe[m] = p == 0 ? Fe[m] : (p == 1 ? Fe[3 + m] : Fe[6 + m]);
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      // column (p,o) of R: (Fe C invFp^T)(i,j,p,o) = C : (Fe(p,:) (x) invFp(o,:))
      double X[9], dS[9];
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int r = 0; r < 3; ++r) X[3 * m + r] = fe[m] * invFp[3 * o + r];
      cubic_contract(X, dS, mc);
#pragma unroll
      for (int k = 0; k < 9; ++k) {
#pragma unroll
        for (int i = k + 1; i < 9; ++i) {
          const bool sw = piv[k] == i;
          const double x = dS[k], y = dS[i];
          dS[k] = sw ? y : x;
          dS[i] = sw ? x : y;
        }
#pragma unroll
        for (int i = k + 1; i < 9; ++i) dS[i] -= L[9 * i + k] * dS[k];
      }
#pragma unroll
      for (int i = 8; i >= 0; --i) {
        double t = dS[i];
#pragma unroll
        for (int j = i + 1; j < 9; ++j) t -= L[9 * i + j] * dS[j];
        dS[i] = t * L[9 * i + i];
      }
      // dS = dSdF(:,:,p,o)
      double w[NSLIP], M[9], dF[9];
      resolved_shears(dS, mc.v, w);
#pragma unroll
      for (int sl = 0; sl < NSLIP; ++sl) w[sl] = -dt * cs[sl] * w[sl];
      slip_sum(w, mc.v, M);
      mm33(invFp0, M, dF);
      double a1[9], a2[9], a3[9], t1[9], t2[9], dFt[9];
      mm33(F, dF, t1);
      mm33(t1, SiT, a1);
      mm33(FiFp, dS, t2);
      mm33(t2, iFt, a2);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) dFt[3 * i + j] = dF[3 * j + i];
      mm33(u3, dFt, a3);
      [[maybe_unused]] double col[9];  // MODE 1: this column's values, summed over the wave together
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          double v = (k == p) ? d0[3 * l + o] : 0.0;  // dPdF(p,l,p,o) = d0(l,o)
          v = v + a1[3 * k + l] + a2[3 * k + l] + a3[3 * k + l];
          // plane index of dPdF(k,l,p,o): k + 3l + 9p + 27o
          const int q = k + 3 * l + 9 * p + 27 * o;
          if constexpr (MODE == 0) {
            a.dPdF[(int64_t)q * N + c] = v;
          } else if constexpr (MODE == 2) {
            a.dPdF[81 * threadIdx.x + q] = v;
          } else {
            n2acc += v * v;
            col[3 * k + l] = live ? v : 0.0;
          }
        }
      if constexpr (MODE == 1) {
        // sums over each row of 16 lanes by DPP exchanges (plain vector
        // moves): every lane of a row ends with the same bits; the rows'
        // first lanes store them, four partials per block and plane
        dpp_step<0xB1>(col);   // quad_perm [1,0,3,2]
        dpp_step<0x4E>(col);   // quad_perm [2,3,0,1]
        dpp_step<0x141>(col);  // row_half_mirror
        dpp_step<0x140>(col);  // row_mirror
        if ((threadIdx.x & 15) == 0) {
          const int64_t at = 4 * (int64_t)blockIdx.x + (threadIdx.x >> 4);
#pragma unroll
          for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int l = 0; l < 3; ++l) a.bpart[(int64_t)(k + 3 * l + 9 * p + 27 * o) * a.nblk + at] = col[3 * k + l];
        }
      }
    }
  }
  if constexpr (MODE == 1) {
    if (live) a.n2[c] = n2acc;
  }
}

// f_materialpoint_forward: state0 <- state (phase_mechanical.f90:992-1007)
__global__ void mp_forward_kernel(MpState s) {
  const int64_t N = s.N;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    s.F0[q * N + c] = s.F[q * N + c];
    s.Fp0[q * N + c] = s.Fp[q * N + c];
    s.Lp0[q * N + c] = s.Lp[q * N + c];
    s.S0[q * N + c] = s.S[q * N + c];
  }
#pragma unroll
  for (int k = 0; k < NSLIP; ++k) {
    s.xi0[k * N + c] = s.xi[k * N + c];
    s.gamma0[k * N + c] = s.gamma[k * N + c];
  }
}

// the pieces of f_materialpoint_forward a context with swappable buffers
// needs: what = 1: F0 <- F, Lp0 <- Lp (the current values stay in use: the
// iterate and the Newton start); what = 2: Fp <- Fp0, S <- S0, xi <- xi0,
// gamma <- gamma0 (the current values again equal to the start values)
__global__ void mp_copy_kernel(MpState s, int what) {
  const int64_t N = s.N;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  if (what == 1) {
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      s.F0[q * N + c] = s.F[q * N + c];
      s.Lp0[q * N + c] = s.Lp[q * N + c];
    }
  } else {
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      s.Fp[q * N + c] = s.Fp0[q * N + c];
      s.S[q * N + c] = s.S0[q * N + c];
    }
#pragma unroll
    for (int k = 0; k < NSLIP; ++k) {
      s.xi[k * N + c] = s.xi0[k * N + c];
      s.gamma[k * N + c] = s.gamma0[k * N + c];
    }
  }
}

// mechanical_init (phase_mechanical.f90:253-262) + phenopowerlaw state init
__global__ void mp_init_kernel(MpState s, const double* Fp_init, double xi_0) {
  const int64_t N = s.N;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const double id = (q % 4 == 0) ? 1.0 : 0.0;
    const double fp = Fp_init[q * N + c];
    s.F[q * N + c] = id;
    s.F0[q * N + c] = id;
    s.Fp[q * N + c] = fp;
    s.Fp0[q * N + c] = fp;
    s.Lp[q * N + c] = 0.0;
    s.Lp0[q * N + c] = 0.0;
    s.S[q * N + c] = 0.0;
    s.S0[q * N + c] = 0.0;
    s.P[q * N + c] = 0.0;
  }
  // Fe = V_e0 Fp^T = Fp^T: plane i+3j of Fe is plane j+3i of Fp
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) s.Fe[(i + 3 * j) * N + c] = Fp_init[(j + 3 * i) * N + c];
#pragma unroll
  for (int k = 0; k < NSLIP; ++k) {
    s.xi[k * N + c] = xi_0;
    s.xi0[k * N + c] = xi_0;
    s.gamma[k * N + c] = 0.0;
    s.gamma0[k * N + c] = 0.0;
  }
}

// ---- launchers ---------------------------------------------------------------
int mp_partial_blocks(int64_t N) { return (int)((N + PAV_CELLS - 1) / PAV_CELLS); }
int mp_sub_planes() { return SUB_PLANES; }

void launch_mp_response(const MpState& s, const MatConst& mc, double dt, int* d_ill, double* d_sub, hipStream_t st,
                        int* d_broken, int64_t c_begin, int64_t c_end) {
  if (c_end < 0) c_end = s.N;
  if (c_begin < 0 || c_end > s.N || c_begin > c_end) throw std::runtime_error("material-point range out of bounds");
  if (c_begin == c_end) return;
  // lanes address planes by a 32-bit byte offset (cell index * 8)
  if (s.N > ((int64_t)1 << 29)) throw std::runtime_error("material-point planes longer than 2^29 cells");
  if (!s.repivot) throw std::runtime_error("material-point state without the repivot plane");
  KernelArgs a{s, mc, dt, d_ill, d_broken, s.repivot, d_sub, c_begin, c_end};
  const int blocks = (int)((c_end - c_begin + 63) / 64);
  // exponents of the example material (n_sl = 20, a_sl = 2.25) are compiled
  // in; any other material runs the runtime-exponent instantiation
  // second pass: a fixed grid striding over the list (it holds at most the
  // range's points)
  const int blocks2 = std::min(blocks, 1024);
  hipLaunchKernelGGL(repivot_reset_kernel, dim3(1), dim3(1), 0, st, s.repivot, d_ill);
  if (mc.n_e4 == 80 && mc.a_e4 == 9) {
    hipLaunchKernelGGL((mp_response_kernel<80, 9>), dim3(blocks), dim3(64), 0, st, a);
    hipLaunchKernelGGL((mp_repivot_kernel<80, 9>), dim3(blocks2), dim3(64), 0, st, a);
  } else {
    hipLaunchKernelGGL((mp_response_kernel<-1, -1>), dim3(blocks), dim3(64), 0, st, a);
    hipLaunchKernelGGL((mp_repivot_kernel<-1, -1>), dim3(blocks2), dim3(64), 0, st, a);
  }
  SCA_HIP(hipGetLastError());
}

// Batched update: the KernelArgs of every item grouped by exponent class
// (class 0: the example exponents, compiled in; class 1: any other), the
// block offsets of each class, and the P block-sum entries with theirs.
void launch_mp_response_batch(const MpBatchItem* items, int n, TableArena& ta, hipStream_t st) {
  if (n <= 0) return;
  KernelArgs* dka = nullptr;
  int *doff = nullptr, *dpvoff = nullptr;
  MpBatchPav* dpv = nullptr;
  KernelArgs* ka = ta.add<KernelArgs>(n, &dka);
  int* off = ta.add<int>(n + 2, &doff);
  MpBatchPav* pv = ta.add<MpBatchPav>(n, &dpv);
  int* pvoff = ta.add<int>(n + 1, &dpvoff);
  auto cls = [](const MatConst& mc) { return (mc.n_e4 == 80 && mc.a_e4 == 9) ? 0 : 1; };
  int n0 = 0;
  for (int i = 0; i < n; ++i) n0 += cls(items[i].mc) == 0 ? 1 : 0;
  int k0 = 0, k1 = n0;
  int64_t maxN = 0;
  for (int i = 0; i < n; ++i) {
    const MpBatchItem& it = items[i];
    const MpState& s = it.s;
    if (s.N > ((int64_t)1 << 29)) throw std::runtime_error("material-point planes longer than 2^29 cells");
    if (!s.repivot || !it.ill || !it.sub) throw std::runtime_error("incomplete batch item");
    const int e = cls(it.mc) == 0 ? k0++ : k1++;
    ka[e] = KernelArgs{s, it.mc, it.dt, it.ill, nullptr, s.repivot, it.sub, 0, s.N};
    // no pav_partial: the P sums come from elsewhere (no blocks)
    pv[e] = MpBatchPav{s.P, s.N, it.pav_partial, it.pav_partial ? mp_partial_blocks(s.N) : 0};
    maxN = std::max(maxN, s.N);
  }
  // block offsets: class 0 in off[0..n0], class 1 in off[n0+1..n+1]
  int64_t tot = 0;
  off[0] = 0;
  for (int i = 0; i < n0; ++i) {
    tot += (ka[i].c_end + 63) / 64;
    off[i + 1] = (int)tot;
  }
  const int64_t blocks0 = tot;
  tot = 0;
  off[n0 + 1] = 0;
  for (int i = n0; i < n; ++i) {
    tot += (ka[i].c_end + 63) / 64;
    off[i + 2] = (int)tot;
  }
  const int64_t blocks1 = tot;
  tot = 0;
  pvoff[0] = 0;
  for (int i = 0; i < n; ++i) {
    tot += pv[i].nb;
    pvoff[i + 1] = (int)tot;
  }
  const int64_t pav_blocks = tot;
  if (blocks0 + blocks1 > INT32_MAX || pav_blocks > INT32_MAX) throw std::runtime_error("batch too large");
  ta.upload(ka, n, st);
  ta.upload(off, n + 2, st);
  ta.upload(pv, n, st);
  ta.upload(pvoff, n + 1, st);
  hipLaunchKernelGGL(repivot_reset_batch_kernel, dim3((n + 63) / 64), dim3(64), 0, st, dka, n);
  // pivoting pass: enough lanes per RVE for its list (about 1 % of its points)
  const int x2 = (int)std::min<int64_t>(1024, std::max<int64_t>(16, (maxN + 1023) / 1024));
  if (n0 > 0) {
    hipLaunchKernelGGL((mp_response_batch_kernel<80, 9>), dim3((unsigned)blocks0), dim3(64), 0, st, dka, doff, n0);
    hipLaunchKernelGGL((mp_repivot_batch_kernel<80, 9>), dim3(x2, n0), dim3(64), 0, st, dka);
  }
  if (n > n0) {
    hipLaunchKernelGGL((mp_response_batch_kernel<-1, -1>), dim3((unsigned)blocks1), dim3(64), 0, st, dka + n0,
                       doff + n0 + 1, n - n0);
    hipLaunchKernelGGL((mp_repivot_batch_kernel<-1, -1>), dim3(x2, n - n0), dim3(64), 0, st, dka + n0);
  }
  if (pav_blocks > 0)
    hipLaunchKernelGGL(pav_partial_batch_kernel, dim3((unsigned)pav_blocks), dim3(256), 0, st, dpv, dpvoff, n);
  SCA_HIP(hipGetLastError());
}

void launch_pav_partial(const double* P, int64_t N, double* partial, hipStream_t st) {
  hipLaunchKernelGGL(pav_partial_kernel, dim3(mp_partial_blocks(N)), dim3(256), 0, st, P, N, partial);
  SCA_HIP(hipGetLastError());
}

// the example material's slip-rate exponent (n_sl = 20) is compiled in: the
// slip rates in the response kernel's form (x^(n-1) from one powering chain,
// c_s without a division; equal to the runtime form to rounding), and no
// pow() call path, whose register saves spill
template <int MODE>
static void launch_tangent_mode(const TangentArgs& a, int64_t blocks, hipStream_t st) {
  if (a.mc.n_e4 == 80) hipLaunchKernelGGL((mp_tangent_kernel<80, MODE>), dim3((unsigned)blocks), dim3(64), 0, st, a);
  else hipLaunchKernelGGL((mp_tangent_kernel<-1, MODE>), dim3((unsigned)blocks), dim3(64), 0, st, a);
  SCA_HIP(hipGetLastError());
}

void launch_mp_tangent(const MpState& s, const MatConst& mc, double dt, double* dPdF, hipStream_t st,
                       int64_t c_begin, int64_t c_end) {
  if (c_end < 0) c_end = s.N;
  if (c_begin < 0 || c_end > s.N || c_begin > c_end) throw std::runtime_error("material-point range out of bounds");
  if (c_begin == c_end) return;
  TangentArgs a;
  a.s = s;
  a.mc = mc;
  a.dt = dt;
  a.dPdF = dPdF;
  a.c_begin = c_begin;
  a.c_end = c_end;
  a.n2 = a.bpart = nullptr;
  a.nblk = 0;
  a.cells = nullptr;
  launch_tangent_mode<0>(a, (c_end - c_begin + 63) / 64, st);
}

int64_t tangent_blocks(int64_t N) { return (N + 63) / 64; }
int64_t tangent_partials(int64_t N) { return 4 * tangent_blocks(N); }

void launch_mp_tangent_reduce(const MpState& s, const MatConst& mc, double dt, double* n2, double* bpart,
                              hipStream_t st) {
  if (s.N < 1) return;
  TangentArgs a;
  a.s = s;
  a.mc = mc;
  a.dt = dt;
  a.dPdF = nullptr;
  a.n2 = n2;
  a.bpart = bpart;
  a.nblk = 4 * tangent_blocks(s.N);
  a.cells = nullptr;
  a.c_begin = 0;
  a.c_end = s.N;
  launch_tangent_mode<1>(a, tangent_blocks(s.N), st);
}

void launch_mp_tangent_cells(const MpState& s, const MatConst& mc, double dt, const int64_t* d_cells, int n,
                             double* out, hipStream_t st) {
  if (n < 1) return;
  if (n > 64) throw std::runtime_error("at most 64 listed cells per tangent launch");
  TangentArgs a;
  a.s = s;
  a.mc = mc;
  a.dt = dt;
  a.dPdF = out;
  a.n2 = a.bpart = nullptr;
  a.nblk = 0;
  a.cells = d_cells;
  a.c_begin = 0;
  a.c_end = n;
  launch_tangent_mode<2>(a, 1, st);
}

void launch_mp_forward(const MpState& s, hipStream_t st) {
  hipLaunchKernelGGL(mp_forward_kernel, dim3((unsigned)((s.N + 255) / 256)), dim3(256), 0, st, s);
  SCA_HIP(hipGetLastError());
}

void launch_mp_copy(const MpState& s, int what, hipStream_t st) {
  if (what != 1 && what != 2) throw std::runtime_error("material-point copy: what");
  hipLaunchKernelGGL(mp_copy_kernel, dim3((unsigned)((s.N + 255) / 256)), dim3(256), 0, st, s, what);
  SCA_HIP(hipGetLastError());
}

void launch_mp_init(const MpState& s, const double* d_Fp_init, double xi0, hipStream_t st) {
  hipLaunchKernelGGL(mp_init_kernel, dim3((unsigned)((s.N + 255) / 256)), dim3(256), 0, st, s, d_Fp_init, xi0);
  SCA_HIP(hipGetLastError());
}

}  // namespace sca
