// Spectral-stage kernels (spectral.hip): FFT plans, Gamma/divergence,
// deterministic reductions and streaming field operations.
#pragma once
#include <vector>

#include "common.h"

namespace sca {

constexpr int MAX_RED_VEC = 96;

// 3D FFT of `ncomp` real planes (N each) <-> half spectra (K complex each)
template <typename T>
struct cpx;
template <typename T>
struct FftPlan {
  int cells[3] = {0, 0, 0};
  int nr = 0;
  int64_t N = 0, K = 0;
  T* tw[3] = {nullptr, nullptr, nullptr};  // forward twiddles per axis
  // launch shape per axis pass (pencil block size, or pencils per block of
  // the generic kernels); 0: the heuristic one.  No shape changes results.
  int cfg[3] = {0, 0, 0};
  void init(const int cells[3]);
  void release();
  // times every launch shape for transforms of ncomp components and keeps
  // the fastest (measured once per precision, grid and ncomp per process)
  void tune(int ncomp, hipStream_t st);
  // real: ncomp planes of N (double or T, converted in the x pass); spec:
  // ncomp planes of K complex T (interleaved re, im).  set_comp (0: ncomp)
  // divides the components into fields of that many: the x pass pairs two
  // real rows per complex pencil only within a field, so a field's result
  // does not depend on the fields transformed with it.
  // rowsum (optional, double input and row_sums_available() only): the sum
  // of every input row (ncomp x Ny x Nz doubles, row = x line), in an order
  // that no launch shape changes
  template <typename TI>
  void forward(const TI* real, T* spec, int ncomp, hipStream_t st, int set_comp = 0,
               double* rowsum = nullptr) const;
  bool row_sums_available() const;  // the x length has a pencil kernel
  template <typename TO>
  void backward(T* spec, TO* real, int ncomp, double scale, hipStream_t st,
                int set_comp = 0) const;  // spec is clobbered
  void strided_passes(void* spec, int ncomp, int sign, hipStream_t st) const;

 private:
  int shape(int axis) const;
  void x_pass_fwd(const void* real, bool real_is_double, cpx<T>* S, int ncomp, int shp, hipStream_t st,
                  int set_comp, double* rowsum = nullptr) const;
  void strided_pass(cpx<T>* S, int ncomp, int sign, int axis, int shp, hipStream_t st) const;
};

struct GammaHost {
  int cells[3];
  double size[3];
  double rescale[3];
  double C_ref[81];  // (l,n,m,o) at 27l+9n+3m+o (row-major)
  int derivative;
  int do_gamma;
};
int gamma_blocks(int64_t K);
template <typename T>
void launch_gamma_div(T* spec, int64_t K, const GammaHost& g, const double* d_aim, const double* d_xi_tab,
                      double* div_partial, hipStream_t st);

// base_update_coords: F_hat -> u_hat (planes 0..2) and the 9 sums of F;
// then node / cell-centre coordinates from u (3 planes of N) and the sums
void launch_coords_uhat(double* spec, int64_t K, const GammaHost& g, const double* d_xi_tab, double* fsum,
                        hipStream_t st);
void launch_coords_nodes(const double* u, const double* fsum, const int cells[3], const double size[3], double* x_n,
                         hipStream_t st);
void launch_coords_points(const double* u, const double* fsum, const int cells[3], const double size[3],
                          double* x_p, hipStream_t st);

int red_blocks();
// ops[v] 0: dot(a_v, b_v); 1: sum (a_v - b_v)^2; 2: sum a_v.  partial: nvec x red_blocks()
template <typename T>
void launch_partial_reduce(const int* ops, int64_t n, int nvec, const T* const* a, const T* const* b,
                           double* partial, hipStream_t st);
// mode 0: out[2i] = <u, v_i>, out[2i+1] = <w, v_i>; mode 1: out[i] = sum (v_i - u)^2,
// for i < nv, as partials (no*nv x red_blocks()) for launch_final_reduce; each
// result bitwise equal to launch_partial_reduce's for the same pair
template <typename T>
void launch_shared_reduce(int mode, int64_t n, const T* u, const T* w, int nv, const T* const* v,
                          double* partial, hipStream_t st);
// cells of C_minMaxAvg: first cell of largest / smallest |dPdF|^2 with norms
// within TANGENT_TIE_REL of the extreme counted as tied.  n2: N doubles of
// scratch, pv: 2*red_blocks() doubles, pi: 2*red_blocks() long longs.
constexpr double TANGENT_TIE_REL = 1e-10;
void tangent_minmax_select(const double* dPdF, int64_t N, double* n2, double* pv, long long* pi, int64_t& imax,
                           int64_t& imin, hipStream_t st);
// the same selection from per-cell norms n2 already formed (pv, pi as above)
void tangent_select_norms(const double* n2, int64_t N, double* pv, long long* pi, int64_t& imax, int64_t& imin,
                          hipStream_t st);
void launch_bc_aim(const double* sums, const double S[81], const double R[9], double wgt, int rotate, double* aim,
                   hipStream_t st);
void launch_final_reduce(const double* partial, int nvec, int nb, double* out, hipStream_t st);
// the 9 sums of a field from its x-pass row sums: out[q] = sum of the rows
// q * nrow .. (q + 1) * nrow - 1, in a fixed order (one block per q)
struct RowSumItem {
  const double* rowsum;
  int64_t nrow;  // rows per component (Ny x Nz)
  double* out;
};
void launch_rowsum_reduce(const RowSumItem& item, hipStream_t st);

void launch_fwd_field(const double* last, const double* rate, double dt, double* out, int64_t n, hipStream_t st);
void launch_sub_plane_const(double* f, const double* diff, int64_t N, hipStream_t st);
void launch_fwd_diff(const double* sums, double wgt, const double* aim, double* diff, hipStream_t st);
void launch_rate(int het, const double* f0, const double* f, double dt, const double* av, double* out, int64_t N,
                 hipStream_t st);
void launch_axpby(double a, const double* x, double b, const double* y0, double* y, int64_t n, hipStream_t st);
void launch_axpby(double a, const double* x, double b, const float* y0, double* y, int64_t n, hipStream_t st);
void launch_lincomb(const double* XM, double c0, int l, const double* beta, const double* const* X, double* XA,
                    int64_t n, hipStream_t st);
// XA as launch_lincomb plus the partials (l+1 rows, red_blocks() each) of
// |XA - XM|^2 and |XA - X_i|^2, bitwise those of launch_shared_reduce(1, n,
// XA, -, l+1, {XM, X_0..}); false (nothing launched) when l + 1 exceeds one
// shared group
bool launch_lincomb_norms(const double* XM, double c0, int l, const double* beta, const double* const* X,
                          double* XA, int64_t n, double* partial, hipStream_t st);
void launch_soa_to_aos(const double* soa, double* aos, int64_t N, int ncomp, hipStream_t st);
void launch_aos_to_soa(const double* aos, double* soa, int64_t N, int ncomp, hipStream_t st);
void launch_f64_to_f32(const double* a, float* b, int64_t n, hipStream_t st);
void launch_f32_to_f64(const float* a, double* b, int64_t n, hipStream_t st);

// ---- launches over a batch of RVEs (tables in a TableArena, common.h) -----
// out[v] = sum_b partial[v*nb + b] for v < nvec, per entry (final_reduce)
struct BatchReduce {
  const double* partial;
  int nb, nvec;
  double* out;
};
void launch_final_reduce_batch(const BatchReduce* items, int n, TableArena& ta, hipStream_t st);
void launch_rowsum_reduce_batch(const RowSumItem* items, int n, TableArena& ta, hipStream_t st);
// bitwise copies of `words` 32-bit words per entry (a double is two)
struct BatchCopy {
  const void* src;
  void* dst;
  int64_t words;
};
void launch_copy_batch(const BatchCopy* items, int n, TableArena& ta, hipStream_t st);
// the DC aim S : (R P_av R^T) of each RVE (launch_bc_aim)
struct BcBatchItem {
  const double* sums;
  double S[81], R[9], wgt;
  int rotate;
  double* aim;
};
void launch_bc_aim_batch(const BcBatchItem* items, int n, TableArena& ta, hipStream_t st);
// Gamma convolution with the divergence of RVEs with the same grid (K
// wavevectors): spec of each (9 planes of K complex T), as launch_gamma_div
struct GammaBatchItem {
  void* spec;
  GammaHost gh;
  const double* d_aim;
  const double* d_xi_tab;
  double* div_partial;
};
template <typename T>
void launch_gamma_div_batch(const GammaBatchItem* items, int n, int64_t K, TableArena& ta, hipStream_t st);
// bytes of table space one RVE of a batch needs at most, all launches together
size_t batch_table_bytes_per_rve();

}  // namespace sca
