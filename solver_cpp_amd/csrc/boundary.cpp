// Drop-in boundary: the bind(C) entry points the reference's C++ grid solver
// calls on its Fortran core (src/grid/spectral.h:16-23, damask_grid.h:16-20,
// main.cpp:8-14, discretization_grid.h:9-22), served by the HIP material
// point.
//
// As in the reference there is one material-point population per process
// (module-global state, not reentrant).  The arrays behind the fetched
// pointers are host mirrors owned by this library, laid out as the Fortran
// arrays: homogenization_F0 / F / P (3,3,N) and dPdF (3,3,3,3,N) column-major,
// terminallyIll a 4-byte LOGICAL.  The constitutive state (Fp, Lp, S, xi,
// gamma and their converged copies) stays resident on the GPU in
// structure-of-arrays planes; each response call moves F to the device and
// response2 moves P and dPdF back, which is what the unmodified reference
// host code needs.  Hosts written against the sc_rve_* API keep every field
// on the device instead.
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.h"
#include "inputs.h"
#include "result.h"
#include "spectral.h"

namespace sca {
extern thread_local std::string g_err;

namespace {

struct Boundary {
  std::string material_path, load_path, grid_path, numerics_path;
  MaterialInput mat;
  sc_phase_numerics phase_num;
  sc_grid_numerics grid_num;
  bool have_base = false, have_discretization = false, have_damask = false;
  // discretization (discretization.f90:79-100)
  std::vector<int> material_at;  // 1-based material entries
  std::vector<double> ip0, ip, node0, node;
  int shared_nodes_begin = 0;
  int64_t N = 0;
  // homogenization arrays (host mirrors, Fortran layout)
  std::vector<double> F0, F, P, dPdF;
  int32_t terminally_ill = 0;
  std::vector<std::pair<int, double>> results;  // (increment, time) of f_materialpoint_result
  // DADF5 job file: the reference's job name unless set
  bool result_default = true, result_compression = true;
  std::string result_path;
  std::unique_ptr<ResultFile> result;
  // device state
  hipStream_t st = nullptr;
  MatConst mc;
  MpState mp;
  std::vector<double*> allocs;
  double *sub = nullptr, *stage = nullptr, *dPdF_dev = nullptr;
  int* d_ill = nullptr;
  double last_dt = 0.0;

  double* dalloc(size_t n) {
    double* p = nullptr;
    SCA_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)));
    allocs.push_back(p);
    return p;
  }
  void release_device() {
    for (double* p : allocs) (void)hipFree(p);
    allocs.clear();
    if (mp.iters) (void)hipFree(mp.iters);
    mp.iters = nullptr;
    if (mp.repivot) (void)hipFree(mp.repivot);
    mp.repivot = nullptr;
    if (d_ill) (void)hipFree(d_ill);
    d_ill = nullptr;
    if (st) (void)hipStreamDestroy(st);
    st = nullptr;
  }
  ~Boundary() { release_device(); }
};

std::unique_ptr<Boundary> g_b;

Boundary& B() {
  if (!g_b) g_b.reset(new Boundary());
  return *g_b;
}

// The bind(C) calls have no error channel; the reference's Fortran core
// stops the process on such errors (IO_error), and so does this library.
void fail(const char* where, const std::exception& e) {
  g_err = std::string(where) + ": " + e.what();
  std::fprintf(stderr, "solver_cpp_amd: %s\n", g_err.c_str());
  std::fflush(stderr);
  std::exit(1);
}

std::string c_str(const char* s, int n) {
  if (!s || n <= 0) return "";
  size_t len = 0;
  while (len < (size_t)n && s[len] != '\0') ++len;
  return std::string(s, len);
}

}  // namespace

// query hooks for tests (declared in the C header)
int boundary_result_count() { return g_b ? (int)g_b->results.size() : 0; }

}  // namespace sca

using namespace sca;

extern "C" {

// materialpoint.f90:68-117: paths are read here (material and numerics);
// the geometry and load files are read by the host solver
void f_materialpoint_initBase(const char* material, int n1, const char* load, int n2, const char* grid, int n3,
                              const char* numerics, int n4) {
  try {
    Boundary& b = B();
    b.material_path = c_str(material, n1);
    b.load_path = c_str(load, n2);
    b.grid_path = c_str(grid, n3);
    b.numerics_path = c_str(numerics, n4);
    b.mat = read_material(read_file(b.material_path));
    sc_default_phase_numerics(&b.phase_num);
    sc_default_grid_numerics(&b.grid_num);
    if (!b.numerics_path.empty()) read_numerics(read_file(b.numerics_path), b.grid_num, b.phase_num);
    b.have_base = true;
  } catch (std::exception& e) {
    fail("f_materialpoint_initBase", e);
  }
}

// discretization.f90:79-100
void f_discretization_init(int* material_at, int* n_cells, double* ip0, int* n_ip, double* node0, int* n_node,
                           int* shared_nodes_begin) {
  try {
    Boundary& b = B();
    b.N = *n_cells;
    if (*n_ip != *n_cells) throw std::runtime_error("one integration point per cell is supported");
    b.material_at.assign(material_at, material_at + b.N);
    b.ip0.assign(ip0, ip0 + 3 * (size_t)*n_ip);
    b.ip = b.ip0;
    b.node0.assign(node0, node0 + 3 * (size_t)*n_node);
    b.node = b.node0;
    b.shared_nodes_begin = shared_nodes_begin ? *shared_nodes_begin : *n_node;
    if (b.shared_nodes_begin < 0 || b.shared_nodes_begin > *n_node)
      throw std::runtime_error("shared_nodes_begin outside the node list");
    b.have_discretization = true;
  } catch (std::exception& e) {
    fail("f_discretization_init", e);
  }
}

void f_discretization_fetch_ip_node_coord_pointers(double** ip, double** node) {
  Boundary& b = B();
  *ip = b.ip.data();
  *node = b.node.data();
}

// materialpoint.f90:138-156: material, phase and homogenization set-up;
// allocates the state of every cell (mechanical_init, phase_mechanical.f90:253-262)
void f_materialpoint_initDamask(void) {
  try {
    Boundary& b = B();
    if (!b.have_base) throw std::runtime_error("f_materialpoint_initBase was not called");
    if (!b.have_discretization) throw std::runtime_error("f_discretization_init was not called");
    b.release_device();
    const int64_t N = b.N;
    const int nm = (int)b.mat.orientations.size() / 4;
    build_matconst(b.mat.material, b.phase_num, b.mc);
    SCA_HIP(hipStreamCreate(&b.st));
    MpState& mp = b.mp;
    mp = MpState{};
    mp.N = N;
    double** t9[10] = {&mp.F, &mp.F0, &mp.Fp0, &mp.Lp0, &mp.S0, &mp.Lp, &mp.P, &mp.Fp, &mp.S, &mp.Fe};
    for (double** p : t9) *p = b.dalloc(9 * (size_t)N);
    double** t12[4] = {&mp.xi0, &mp.gamma0, &mp.xi, &mp.gamma};
    for (double** p : t12) *p = b.dalloc(12 * (size_t)N);
    b.stage = b.dalloc(81 * (size_t)N);
    b.dPdF_dev = b.dalloc(81 * (size_t)N);
    b.sub = b.dalloc((size_t)mp_sub_planes() * N);
    SCA_HIP(hipMalloc(&mp.iters, 2 * N * sizeof(int)));
    SCA_HIP(hipMalloc(&mp.repivot, (N + 1) * sizeof(int)));
    SCA_HIP(hipMalloc(&b.d_ill, sizeof(int)));
    // Fp0 = orientation matrix of each cell's material (material.f90:155)
    std::vector<double> om(9 * (size_t)nm), fp(9 * (size_t)N);
    for (int m = 0; m < nm; ++m) qu2om_host(&b.mat.orientations[4 * m], &om[9 * m]);
    for (int64_t c = 0; c < N; ++c) {
      const int ma = b.material_at[c] - 1;
      if (ma < 0 || ma >= nm) throw std::runtime_error("material index out of range at cell " + std::to_string(c + 1));
      for (int q = 0; q < 9; ++q) fp[q * N + c] = om[9 * ma + q];
    }
    SCA_HIP(hipMemcpy(b.stage, fp.data(), fp.size() * sizeof(double), hipMemcpyHostToDevice));
    launch_mp_init(mp, b.stage, b.mat.material.xi_0_sl, b.st);
    SCA_HIP(hipStreamSynchronize(b.st));
    b.F0.assign(9 * (size_t)N, 0.0);
    for (int64_t c = 0; c < N; ++c)
      for (int q = 0; q < 9; q += 4) b.F0[9 * c + q] = 1.0;
    b.F = b.F0;
    b.P.assign(9 * (size_t)N, 0.0);
    b.dPdF.assign(81 * (size_t)N, 0.0);
    b.terminally_ill = 0;
    // result_init and material_init's mappings (materialpoint.f90:140-150)
    b.result.reset();
    const std::string path = b.result_default ? solver_job_name(b.grid_path, b.load_path, b.material_path,
                                                                b.numerics_path) + ".hdf5"
                                              : b.result_path;
    if (!path.empty()) {
      ResultSetup rs = result_setup_from_material(read_file(b.material_path));
      if (!b.numerics_path.empty()) rs.numerics_yaml = read_file(b.numerics_path);
      rs.compression = b.result_compression;
      b.result.reset(new ResultFile(path, rs, N));
    }
    b.have_damask = true;
  } catch (std::exception& e) {
    fail("f_materialpoint_initDamask", e);
  }
}

// homogenization.f90:491-510
void f_homogenization_fetch_tensor_pointers(double** F0, double** F, double** P, double** dPdF,
                                            void** terminally_ill) {
  Boundary& b = B();
  *F0 = b.F0.data();
  *F = b.F.data();
  *P = b.P.data();
  *dPdF = b.dPdF.data();
  *terminally_ill = &b.terminally_ill;
}

// homogenization.f90:219-266 for cells [cell_start, cell_end] (1-based):
// restore, integrate the constitutive response at homogenization_F, flag
// non-converged cells in terminallyIll
void f_homogenization_mechanical_response(double* dt, int* cell_start, int* cell_end) {
  try {
    Boundary& b = B();
    if (!b.have_damask) throw std::runtime_error("f_materialpoint_initDamask was not called");
    const int64_t c0 = *cell_start - 1, c1 = *cell_end;
    if (c0 < 0 || c1 > b.N || c0 > c1) throw std::runtime_error("cell range out of bounds");
    const int64_t n = c1 - c0;
    if (n == 0) return;
    // F of the range (3,3,n) -> planes
    SCA_HIP(hipMemcpyAsync(b.stage, b.F.data() + 9 * c0, 9 * n * sizeof(double), hipMemcpyHostToDevice, b.st));
    if (c0 == 0 && c1 == b.N) {
      launch_aos_to_soa(b.stage, b.mp.F, b.N, 9, b.st);
    } else {
      // scatter the sub-range into the N-strided planes
      std::vector<double> tmp(9 * (size_t)n);
      std::memcpy(tmp.data(), b.F.data() + 9 * c0, tmp.size() * sizeof(double));
      for (int q = 0; q < 9; ++q) {
        std::vector<double> plane(n);
        for (int64_t i = 0; i < n; ++i) plane[i] = tmp[9 * i + q];
        SCA_HIP(hipMemcpyAsync(b.mp.F + q * b.N + c0, plane.data(), n * sizeof(double), hipMemcpyHostToDevice, b.st));
        SCA_HIP(hipStreamSynchronize(b.st));
      }
    }
    launch_mp_response(b.mp, b.mc, *dt, b.d_ill, b.sub, b.st, nullptr, c0, c1);
    int ill = 0;
    SCA_HIP(hipMemcpyAsync(&ill, b.d_ill, sizeof(int), hipMemcpyDeviceToHost, b.st));
    SCA_HIP(hipStreamSynchronize(b.st));
    if (ill) b.terminally_ill = 1;
    b.last_dt = *dt;
  } catch (std::exception& e) {
    fail("f_homogenization_mechanical_response", e);
  }
}

// homogenization.f90:272-291: no thermal field is active
void f_homogenization_thermal_response(double*, int*, int*) {}

// homogenization.f90:299-324 with one integration point per element:
// homogenization_P and homogenization_dPdF of cells execElem(1)..execElem(2)
void f_homogenization_mechanical_response2(double* dt, int* execIP, int* execElem) {
  try {
    Boundary& b = B();
    if (!b.have_damask) throw std::runtime_error("f_materialpoint_initDamask was not called");
    if (execIP[0] != 1 || execIP[1] != 1) throw std::runtime_error("one integration point per cell is supported");
    const int64_t c0 = execElem[0] - 1, c1 = execElem[1];
    if (c0 < 0 || c1 > b.N || c0 > c1) throw std::runtime_error("cell range out of bounds");
    if (c1 == c0) return;
    launch_mp_tangent(b.mp, b.mc, *dt, b.dPdF_dev, b.st, c0, c1);
    const int64_t N = b.N;
    launch_soa_to_aos(b.mp.P, b.stage, N, 9, b.st);
    SCA_HIP(hipMemcpyAsync(b.P.data() + 9 * c0, b.stage + 9 * c0, 9 * (c1 - c0) * sizeof(double),
                           hipMemcpyDeviceToHost, b.st));
    SCA_HIP(hipStreamSynchronize(b.st));
    launch_soa_to_aos(b.dPdF_dev, b.stage, N, 81, b.st);
    SCA_HIP(hipMemcpyAsync(b.dPdF.data() + 81 * c0, b.stage + 81 * c0, 81 * (c1 - c0) * sizeof(double),
                           hipMemcpyDeviceToHost, b.st));
    SCA_HIP(hipStreamSynchronize(b.st));
  } catch (std::exception& e) {
    fail("f_homogenization_mechanical_response2", e);
  }
}

// materialpoint.f90:208-215: state0 <- state of every cell
void f_materialpoint_forward(void) {
  try {
    Boundary& b = B();
    if (!b.have_damask) throw std::runtime_error("f_materialpoint_initDamask was not called");
    launch_mp_forward(b.mp, b.st);
    SCA_HIP(hipStreamSynchronize(b.st));
  } catch (std::exception& e) {
    fail("f_materialpoint_forward", e);
  }
}

// materialpoint.f90:221-234 writes the increment to the DADF5 job file; no
// HDF5 output is produced here, the call records (increment, time)
void f_materialpoint_result(int* inc, double* t) {
  try {
    Boundary& b = B();
    b.results.emplace_back(*inc, *t);
    if (!b.result) return;
    // discretization_result (discretization.f90:101-116)
    std::vector<double> u_p(3 * (size_t)b.N), u_n(3 * (size_t)b.shared_nodes_begin);
    for (size_t k = 0; k < u_p.size(); ++k) u_p[k] = b.ip[k] - b.ip0[k];
    for (size_t k = 0; k < u_n.size(); ++k) u_n[k] = b.node[k] - b.node0[k];
    b.result->increment(*inc, *t, b.mp, b.st, u_p, u_n);
  } catch (std::exception& e) {
    fail("f_materialpoint_result", e);
  }
}

int sc_boundary_get_state(int field, double* out, int64_t n_values) {
  try {
    Boundary& b = B();
    if (!b.have_damask) throw std::runtime_error("f_materialpoint_initDamask was not called");
    const double* src = nullptr;
    int nc = 9;
    switch (field) {
      case SC_FIELD_FP: src = b.mp.Fp; break;
      case SC_FIELD_LP: src = b.mp.Lp; break;
      case SC_FIELD_S: src = b.mp.S; break;
      case SC_FIELD_FE: src = b.mp.Fe; break;
      case SC_FIELD_P: src = b.mp.P; break;
      case SC_FIELD_XI: src = b.mp.xi; nc = 12; break;
      case SC_FIELD_GAMMA: src = b.mp.gamma; nc = 12; break;
      default: throw std::runtime_error("field not available through the boundary");
    }
    if (n_values < nc * b.N) throw std::runtime_error("output buffer too small");
    launch_soa_to_aos(src, b.stage, b.N, nc, b.st);
    SCA_HIP(hipMemcpyAsync(out, b.stage, nc * b.N * sizeof(double), hipMemcpyDeviceToHost, b.st));
    SCA_HIP(hipStreamSynchronize(b.st));
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_boundary_result_count(void) { return boundary_result_count(); }

void sc_boundary_reset(void) { g_b.reset(); }

int sc_boundary_set_result_file(const char* path, int compression) {
  Boundary& b = B();
  b.result_default = false;
  b.result_path = path ? path : "";
  b.result_compression = compression != 0;
  return 0;
}

}  // extern "C"
