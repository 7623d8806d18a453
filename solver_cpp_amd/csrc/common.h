// Shared definitions of the HIP solver library (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/solver_cpp_amd.h"

namespace sca {

#define SCA_HIP(call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                               __FILE__ + ":" + std::to_string(__LINE__));                \
  } while (0)

// Caching allocators (pool.cpp): device and pinned host blocks the library
// returns are kept per rounded size and handed out again, and streams are
// reused, so starting and releasing RVEs calls neither hipMalloc nor hipFree
// (which synchronise the whole device) once blocks of their sizes exist.  A
// block may be returned only after the work that used it has completed.
void* dev_alloc(size_t bytes);
void dev_free(void* p);
void* host_alloc(size_t bytes);
void host_free(void* p);
void pool_trim(size_t keep_bytes);  // hipFree cached device blocks beyond keep_bytes
size_t pool_cached_bytes();
hipStream_t stream_get();  // a non-blocking stream of default priority
void stream_put(hipStream_t s);

constexpr int NSLIP = 12;
constexpr int NSTATE = 24;

// Material constants in the form the kernels use.  Slip system s has the
// Schmid matrix P_s = v * sgn_s (sgn_s: entries in {-1,0,1}, compile-time in
// the kernel), so everything built from P_s and the cubic stiffness is a
// compile-time pattern times a few runtime scalars; hv are the 7 slip-slip
// interaction values indexed by the compile-time cF interaction types.
struct MatConst {
  double C11, C12, C44;
  double v;            // |P_s| entry magnitude, (1/sqrt 2)(1/sqrt 3)
  double hv[7];
  double xi_inf, h0, n, a, gdot0;
  double atol_xi, atol_gamma;
  // phase numerics
  double r_cutback_min, r_cutback, r_increase, eps_rel_state;
  int N_iter_state_max, N_iter_Lp_max;
  double r_linesearch_Lp, eps_rel_Lp, eps_abs_Lp;
  // exponents n_sl and a_sl as multiples of 1/4 (e4 = 4*exponent) when they
  // are, so that x^e = x^(e4/4 integer part) * x^(quarter part) needs only
  // squarings and square roots; -1 means "use pow()"
  int n_e4, a_e4;
};

// Host-side setup (material.cpp)
void build_matconst(const sc_material& m, const sc_phase_numerics& p, MatConst& out);
void qu2om_host(const double q[4], double om[9]);  // column-major (3,3)

// Global state arrays of one RVE's material points, structure of arrays:
// a (3,3) tensor field is 9 planes of N doubles (component i+3j in plane
// i+3j, column-major as the reference), a slip field is 12 planes.
struct MpState {
  int64_t N = 0;
  double* F = nullptr;       // homogenization_F (input of a response)
  double* F0 = nullptr;      // phase F0 (converged at end of last increment)
  double* Fp0 = nullptr;
  double* Lp0 = nullptr;
  double* S0 = nullptr;
  double* xi0 = nullptr;
  double* gamma0 = nullptr;
  double* Lp = nullptr;      // Newton guess in, Lp out
  double* P = nullptr;
  double* Fp = nullptr;
  double* S = nullptr;
  double* Fe = nullptr;
  double* xi = nullptr;
  double* gamma = nullptr;
  int* iters = nullptr;      // (2,N): Newton iterations, state iterations of the last update
  int* repivot = nullptr;    // (1 + N): list of the points the pivoting pass redoes (count first)
};

// Argument tables of batched launches: pinned host staging and a device
// buffer.  reserve() (before a batch, may reallocate) then add() slots; each
// launcher fills its slots and uploads them on its stream.  A batch's
// tables are rewritten only after the previous batch completed.
class TableArena {
 public:
  TableArena() = default;
  TableArena(const TableArena&) = delete;
  TableArena& operator=(const TableArena&) = delete;
  ~TableArena();
  void reserve(size_t bytes);  // and empties the arena
  // n entries of E: host slot (returned) and its device address (*dev)
  template <typename E>
  E* add(int n, E** dev) {
    const size_t off = used;
    used = (used + (size_t)n * sizeof(E) + 255) & ~(size_t)255;
    if (used > cap) throw std::runtime_error("table arena too small");
    *dev = reinterpret_cast<E*>(dptr + off);
    return reinterpret_cast<E*>(hptr + off);
  }
  template <typename E>
  void upload(const E* host, int n, hipStream_t st) {
    const size_t off = reinterpret_cast<const char*>(host) - hptr;
    SCA_HIP(hipMemcpyAsync(dptr + off, host, (size_t)n * sizeof(E), hipMemcpyHostToDevice, st));
  }

 private:
  char *hptr = nullptr, *dptr = nullptr;
  size_t cap = 0, used = 0;
};

// One RVE's share of a batched material-point update
// (launch_mp_response_batch): its planes with s.F the F of this evaluation,
// its material, time step, terminally-ill flag, substep scratch and the 9 x
// mp_partial_blocks(N) doubles that receive the block sums of P.
struct MpBatchItem {
  MpState s;
  MatConst mc;
  double dt;
  int* ill;
  double* sub;
  double* pav_partial;  // null: no P block sums for this RVE
};
// P block sums of one RVE of a batch (device table entry)
struct MpBatchPav {
  const double* P;
  int64_t N;
  double* partial;
  int nb;
};

// Kernel launchers (constitutive.hip)
// cells [c_begin, c_end) of the planes (c_end < 0: all N); *d_ill is cleared
// by the launch and set when a point is terminally ill
void launch_mp_response(const MpState& s, const MatConst& mc, double dt, int* d_ill, double* d_sub, hipStream_t st,
                        int* d_broken = nullptr, int64_t c_begin = 0, int64_t c_end = -1);
int mp_partial_blocks(int64_t N);
// Material-point updates of n RVEs in one launch (and their pivoting passes
// and P block sums), each RVE's results bitwise those of launch_mp_response
// + launch_pav_partial; the argument tables go into ta.
void launch_mp_response_batch(const MpBatchItem* items, int n, TableArena& ta, hipStream_t st);
int mp_sub_planes();
void launch_pav_partial(const double* P, int64_t N, double* partial, hipStream_t st);
void launch_mp_tangent(const MpState& s, const MatConst& mc, double dt, double* dPdF /*(81,N) SoA*/,
                       hipStream_t st, int64_t c_begin = 0, int64_t c_end = -1);
// the consistent tangent of all N cells reduced instead of stored: n2 (N
// doubles) the cells' |dPdF|^2, bpart (81 x tangent_partials(N) doubles) the
// plane sums per 16 cells, at plane * tangent_partials(N) + cell / 16
int64_t tangent_blocks(int64_t N);
int64_t tangent_partials(int64_t N);
void launch_mp_tangent_reduce(const MpState& s, const MatConst& mc, double dt, double* n2, double* bpart,
                              hipStream_t st);
// dPdF of n <= 64 listed cells (device array d_cells) into out (81 per cell,
// plane k+3l+9p+27o), the values launch_mp_tangent gives for those cells
void launch_mp_tangent_cells(const MpState& s, const MatConst& mc, double dt, const int64_t* d_cells, int n,
                             double* out, hipStream_t st);
void launch_mp_forward(const MpState& s, hipStream_t st);
// what = 1: F0 <- F, Lp0 <- Lp; what = 2: Fp, S, xi, gamma <- their start values
void launch_mp_copy(const MpState& s, int what, hipStream_t st);
void launch_mp_init(const MpState& s, const double* d_Fp_init /*(9,N)*/, double xi0, hipStream_t st);

}  // namespace sca
