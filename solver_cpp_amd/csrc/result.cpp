// DADF5 job result file (result.h).
#include "result.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <ctime>
#include <exception>
#include <fstream>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <thread>

#include "inputs.h"

namespace sca {
namespace {

// cF <110>{111} slip systems (crystal.f90:41-63): direction, normal
constexpr int CF_SLIP[12][6] = {
    {0, 1, -1, 1, 1, 1},   {-1, 0, 1, 1, 1, 1},   {1, -1, 0, 1, 1, 1},
    {0, -1, -1, -1, -1, 1}, {1, 0, 1, -1, -1, 1},  {-1, 1, 0, -1, -1, 1},
    {0, -1, 1, 1, -1, -1}, {-1, 0, -1, 1, -1, -1}, {1, 1, 0, 1, -1, -1},
    {0, 1, 1, -1, 1, -1},  {1, 0, -1, -1, 1, -1},  {-1, -1, 0, -1, 1, -1}};

// slip system labels of crystal_labels_slip (crystal.f90:2093-2143): every
// index right aligned in two characters followed by a blank, the blanks
// after the last index of each bracket replaced by ']' and ')'
std::vector<std::string> cf_slip_labels() {
  std::vector<std::string> out;
  for (const auto& s : CF_SLIP) {
    std::string l = "[";
    char b[4];
    for (int j = 0; j < 3; ++j) {
      std::snprintf(b, sizeof b, "%2d ", s[j]);
      l += b;
    }
    l.back() = ']';
    l += '(';
    for (int j = 3; j < 6; ++j) {
      std::snprintf(b, sizeof b, "%2d ", s[j]);
      l += b;
    }
    l.back() = ')';
    out.push_back(l);
  }
  return out;
}

// now() of result.f90:816-826: YYYY-MM-DD HH:MM:SS+hhmm
std::string now_stamp() {
  std::time_t t = std::time(nullptr);
  std::tm tm{};
  localtime_r(&t, &tm);
  char b[64];
  std::strftime(b, sizeof b, "%Y-%m-%d %H:%M:%S%z", &tm);
  return b;
}

std::string command_line() {
  std::ifstream f("/proc/self/cmdline", std::ios::binary);
  std::string s((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  while (!s.empty() && s.back() == '\0') s.pop_back();
  std::replace(s.begin(), s.end(), '\0', ' ');
  return s;
}

const char* CREATOR = "solver_cpp_amd (DADF5 1.0)";

// fn(c) for c in [0, n) on up to 16 threads; the first exception thrown by
// any of them is rethrown here after all have stopped
template <typename Fn>
void parallel_for(int64_t n, Fn&& fn) {
  const int64_t nt = std::max<int64_t>(1, std::min<int64_t>({16, (int64_t)std::thread::hardware_concurrency(),
                                                             n / 4096 + 1}));
  std::vector<std::thread> pool;
  std::atomic<int64_t> next{0};
  std::mutex mu;
  std::exception_ptr err;
  const int64_t block = 4096;
  auto work = [&] {
    try {
      for (int64_t b; (b = next.fetch_add(block)) < n;) {
        const int64_t e = std::min(n, b + block);
        for (int64_t c = b; c < e; ++c) fn(c);
      }
    } catch (...) {
      std::lock_guard<std::mutex> lock(mu);
      if (!err) err = std::current_exception();
      next = n;
    }
  };
  for (int64_t t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  if (err) std::rethrow_exception(err);
}

std::vector<double> download(const double* d, int64_t n, hipStream_t st) {
  std::vector<double> h((size_t)n);
  SCA_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(double), hipMemcpyDeviceToHost, st));
  SCA_HIP(hipStreamSynchronize(st));
  return h;
}

}  // namespace

ResultSetup result_setup_from_material(const std::string& text) {
  const MaterialInput mi = read_material(text);
  ResultSetup s;
  s.phase = mi.phase_name;
  s.homogenization = mi.homogenization_name;
  s.mechanical = mi.mechanical_output;
  s.plastic = mi.plastic_output;
  s.material_yaml = text;
  return s;
}

std::string solver_job_name(const std::string& geom, const std::string& load, const std::string& material,
                            const std::string& numerics) {
  auto stem = [](const std::string& p) {
    const size_t a = p.rfind('/');
    const size_t b = p.rfind('.');
    const size_t s = (a == std::string::npos) ? 0 : a + 1;
    if (b == std::string::npos || b < s) return std::string();
    return p.substr(s, b - s);
  };
  std::string j = stem(geom) + "_" + stem(load) + "_" + stem(material);
  if (!numerics.empty()) j += "_" + stem(numerics);
  return j;
}

void ResultFile::stamp(const std::string& obj, const std::string& description, const char* unit) {
  w_.attr_str(obj, "creator", CREATOR);
  w_.attr_str(obj, "created", now_stamp());
  w_.attr_str(obj, "description", description);
  if (unit) w_.attr_str(obj, "unit", unit);
}

ResultFile::ResultFile(const std::string& path, const ResultSetup& setup, int64_t n_cells)
    : path_(path), s_(setup), N_(n_cells), w_(path, setup.compression) {
  if (N_ <= 0) throw std::runtime_error("result file of an empty grid");
  // result_init (result.f90:86-98)
  w_.attr_i32("", "DADF5_version_major", 1);
  w_.attr_i32("", "DADF5_version_minor", 0);
  w_.attr_str("", "creator", CREATOR);
  w_.attr_str("", "created", now_stamp());
  w_.attr_str("", "call", command_line());
  w_.group("cell_to");
  w_.attr_str("cell_to", "description", "mappings to place data in space");
  w_.group("setup");
  w_.attr_str("setup", "description", "input data used to run the simulation");
  // config.f90:112-127 (parse -> result_addSetupFile)
  if (!s_.material_yaml.empty()) {
    w_.dataset_str("setup/material.yaml", s_.material_yaml);
    stamp("setup/material.yaml", "material configuration", nullptr);
  }
  if (!s_.numerics_yaml.empty()) {
    w_.dataset_str("setup/numerics.yaml", s_.numerics_yaml);
    stamp("setup/numerics.yaml", "numerics configuration", nullptr);
  }
  // material_init's mappings: one phase and one homogenization, entries
  // counted per cell (0-based)
  std::vector<int64_t> entry((size_t)N_);
  for (int64_t c = 0; c < N_; ++c) entry[c] = c;
  w_.dataset_label_entry("cell_to/phase", {(uint64_t)N_, 1}, (uint32_t)s_.phase.size(),
                         std::vector<std::string>((size_t)N_, s_.phase), entry);
  stamp("cell_to/phase", "cell ID and constituent ID to phase results", nullptr);
  w_.dataset_label_entry("cell_to/homogenization", {(uint64_t)N_}, (uint32_t)s_.homogenization.size(),
                         std::vector<std::string>((size_t)N_, s_.homogenization), entry);
  stamp("cell_to/homogenization", "cell ID to homogenization results", nullptr);
  w_.flush();
}

void ResultFile::increment(int inc, double t, const MpState& mp, hipStream_t st, const std::vector<double>& u_p,
                           const std::vector<double>& u_n) {
  const int64_t N = N_;
  if (mp.N != N) throw std::runtime_error("result file and state differ in the number of cells");
  if ((int64_t)u_p.size() != 3 * N || u_n.size() % 3) throw std::runtime_error("result displacements of wrong size");
  const std::string ig = "increment_" + std::to_string(inc);
  if (w_.exists(ig)) throw std::runtime_error("result increment exists: " + ig);
  w_.group(ig);
  w_.attr_f64(ig, "t/s", t);
  const std::string pg = ig + "/phase/" + s_.phase + "/mechanical";
  w_.group(pg);
  // tensors: DADF5 stores T(c)_ij at [c][i][j] (result.f90:414-442 writes
  // the transposed Fortran array); device planes hold component i + 3 j
  std::vector<double> out;
  auto tensor = [&](const std::string& name, const double* planes, const char* desc, const char* unit) {
    const std::vector<double> h = download(planes, 9 * N, st);
    out.assign(9 * (size_t)N, 0.0);
    parallel_for(N, [&](int64_t c) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[9 * c + 3 * i + j] = h[(i + 3 * j) * N + c];
    });
    w_.dataset_f64(pg + "/" + name, {(uint64_t)N, 3, 3}, out.data());
    stamp(pg + "/" + name, desc, unit);
  };
  auto constant = [&](const std::string& name, double diag, const char* desc, const char* unit) {
    out.assign(9 * (size_t)N, 0.0);
    if (diag != 0.0)
      for (int64_t c = 0; c < N; ++c) out[9 * c] = out[9 * c + 4] = out[9 * c + 8] = diag;
    w_.dataset_f64(pg + "/" + name, {(uint64_t)N, 3, 3}, out.data());
    stamp(pg + "/" + name, desc, unit);
  };
  // phase_mechanical.f90:927-960 (labels not listed there are skipped, as
  // its select case has no default)
  for (const std::string& o : s_.mechanical) {
    if (o == "F") tensor(o, mp.F, "deformation gradient", "1");
    else if (o == "F_e") tensor(o, mp.Fe, "elastic deformation gradient", "1");
    else if (o == "F_p") tensor(o, mp.Fp, "plastic deformation gradient", "1");
    else if (o == "F_i") constant(o, 1.0, "inelastic deformation gradient", "1");
    else if (o == "L_p") tensor(o, mp.Lp, "plastic velocity gradient", "1/s");
    else if (o == "L_i") constant(o, 0.0, "inelastic velocity gradient", "1/s");
    else if (o == "P") tensor(o, mp.P, "first Piola-Kirchhoff stress", "Pa");
    else if (o == "S") tensor(o, mp.S, "second Piola-Kirchhoff stress", "Pa");
    else if (o == "O") {
      // crystallite_orientations (phase.f90:577-596): the orientation of
      // transpose(rotational part of F_e), as a quaternion (w >= 0)
      const std::vector<double> h = download(mp.Fe, 9 * N, st);
      out.assign(4 * (size_t)N, 0.0);
      parallel_for(N, [&](int64_t c) {
        double fe[9], R[9], om[9];
        for (int k = 0; k < 9; ++k) fe[k] = h[k * N + c];
        rotational_part_host(fe, R);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) om[i + 3 * j] = R[j + 3 * i];
        om2qu_host(om, &out[4 * c]);
      });
      w_.dataset_f64(pg + "/O", {(uint64_t)N, 4}, out.data());
      stamp(pg + "/O", "crystal orientation as quaternion q_0 (q_1 q_2 q_3)", "1");
      w_.attr_str(pg + "/O", "lattice", "cF");
    }
  }
  // phenopowerlaw results (phase_mechanical_plastic_phenopowerlaw.f90:409-428)
  for (const std::string& o : s_.plastic) {
    const double* src = nullptr;
    const char *desc = nullptr, *unit = nullptr;
    if (o == "xi_sl") src = mp.xi, desc = "resistance against plastic slip", unit = "Pa";
    else if (o == "gamma_sl") src = mp.gamma, desc = "plastic shear", unit = "1";
    if (!src) continue;
    const std::vector<double> h = download(src, 12 * N, st);
    out.assign(12 * (size_t)N, 0.0);
    parallel_for(N, [&](int64_t c) {
      for (int s = 0; s < 12; ++s) out[12 * c + s] = h[s * N + c];
    });
    w_.dataset_f64(pg + "/" + o, {(uint64_t)N, 12}, out.data());
    stamp(pg + "/" + o, desc, unit);
    w_.attr_str_array(pg + "/" + o, "systems", cf_slip_labels());
  }
  // homogenization pass writes no mechanical results (an empty group)
  w_.group(ig + "/homogenization/" + s_.homogenization + "/mechanical");
  // discretization_result (discretization.f90:101-116)
  w_.dataset_f64(ig + "/geometry/u_n", {(uint64_t)(u_n.size() / 3), 3}, u_n.data());
  stamp(ig + "/geometry/u_n", "displacements of the nodes", "m");
  w_.dataset_f64(ig + "/geometry/u_p", {(uint64_t)N, 3}, u_p.data());
  stamp(ig + "/geometry/u_p", "displacements of the materialpoints (cell centers)", "m");
  w_.flush();
}

}  // namespace sca
