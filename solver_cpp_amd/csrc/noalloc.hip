// Material-point update on caller-owned device planes without any device
// allocation (sc_mp_update_noalloc).
//
// The kernels are those of sc_mp_update (launch_mp_response, launch_mp_tangent).
// The device memory they need beyond the caller's inputs and outputs -- the
// substep scratch planes, the list of points the pivoting pass redoes and the
// terminally-ill flag -- lives in `scratch`, one fixed array per GPU that the
// code object carries.  In host code `scratch` names a host-side stand-in, not
// the device array: its device address is taken once per device with
// hipGetSymbolAddress and only that address is ever handed to kernels.
//
// Every call on a GPU shares that GPU's array and leaves its contents behind.
// The pivoting list's count and the flag are cleared by a kernel at the start
// of each update (a kernel, not hipMemsetAsync, so that the update can be
// captured into a HIP graph); the substep planes of a point are written before
// they are read.  Two calls that overlap on one GPU race on the array: callers
// serialise them (one stream, or one call at a time).
#include <mutex>

#include "common.h"

namespace sca {

extern thread_local std::string g_err;

namespace {

constexpr int64_t MAX_POINTS = SC_MP_NOALLOC_MAX_POINTS;
constexpr int NOALLOC_SUB_PLANES = 60;  // planes of substep scratch per point
constexpr int64_t SUB_DOUBLES = NOALLOC_SUB_PLANES * MAX_POINTS;
// ints after the substep planes: the flag at 0, the pivoting list from 64 on
// (a count and up to MAX_POINTS cells)
constexpr int64_t INT_LIST = 64;
constexpr int64_t INTS = INT_LIST + 1 + MAX_POINTS;
constexpr int64_t SCRATCH_DOUBLES = SUB_DOUBLES + (INTS + 1) / 2;

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_scratch[MAX_DEVICES] = {};

// device address of this device's scratch array (looked up once per device)
double* device_scratch() {
  int dev = -1;
  SCA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress of the material-point scratch failed: ") +
                               hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(p);
  }
  return g_scratch[dev];
}

}  // namespace
}  // namespace sca

extern "C" int sc_mp_update_noalloc(int64_t n, double dt, const sc_material* m, const sc_phase_numerics* num,
                                    const sc_mp_device_planes* pl, void* stream) {
  using namespace sca;
  try {
    if (n > MAX_POINTS)
      throw std::runtime_error("sc_mp_update_noalloc: " + std::to_string(n) + " points exceed its capacity of " +
                               std::to_string(MAX_POINTS) + " (SC_MP_NOALLOC_MAX_POINTS)");
    if (n <= 0) return 0;
    if (!m || !num || !pl) throw std::runtime_error("sc_mp_update_noalloc: null argument");
    const double* const in[7] = {pl->F, pl->F0, pl->Fp0, pl->Lp0, pl->S0, pl->xi0, pl->gamma0};
    double* const io[7] = {pl->Lp, pl->P, pl->Fp, pl->S, pl->Fe, pl->xi, pl->gamma};
    for (const double* p : in)
      if (!p) throw std::runtime_error("sc_mp_update_noalloc: null input plane");
    for (const double* p : io)
      if (!p) throw std::runtime_error("sc_mp_update_noalloc: null output plane");
    if (mp_sub_planes() != NOALLOC_SUB_PLANES)
      throw std::runtime_error("sc_mp_update_noalloc: substep scratch layout changed");
    double* base = device_scratch();
    MatConst mc;
    build_matconst(*m, *num, mc);
    MpState s;
    s.N = n;
    s.F = const_cast<double*>(pl->F);  // read only
    s.F0 = const_cast<double*>(pl->F0);
    s.Fp0 = const_cast<double*>(pl->Fp0);
    s.Lp0 = const_cast<double*>(pl->Lp0);
    s.S0 = const_cast<double*>(pl->S0);
    s.xi0 = const_cast<double*>(pl->xi0);
    s.gamma0 = const_cast<double*>(pl->gamma0);
    s.Lp = pl->Lp;
    s.P = pl->P;
    s.Fp = pl->Fp;
    s.S = pl->S;
    s.Fe = pl->Fe;
    s.xi = pl->xi;
    s.gamma = pl->gamma;
    s.iters = pl->iterations;
    int* ints = reinterpret_cast<int*>(base + SUB_DOUBLES);
    s.repivot = ints + INT_LIST;
    hipStream_t st = static_cast<hipStream_t>(stream);
    launch_mp_response(s, mc, dt, ints, base, st, pl->broken);
    if (pl->dPdF) launch_mp_tangent(s, mc, dt, pl->dPdF, st);
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
