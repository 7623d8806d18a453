// DADF5 job result file of one RVE: the groups and datasets the reference
// writes through result_init (result.f90:76-116), material_init's cell
// mappings (result.f90:492-643, 645-790) and materialpoint_result
// (materialpoint.f90:221-234: phase, homogenization and geometry results of
// one increment).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "common.h"
#include "dadf5.h"

namespace sca {

struct ResultSetup {
  std::string phase = "Aluminum", homogenization = "SX";
  std::vector<std::string> mechanical{"F", "P", "F_e", "F_p", "L_p", "O"};  // examples/grid/material.yaml:11
  std::vector<std::string> plastic{"xi_sl"};                                // examples/grid/material.yaml:22
  std::string material_yaml, numerics_yaml;  // written to setup/ when not empty
  bool compression = true;
};

// names and output lists from the text of a material configuration
ResultSetup result_setup_from_material(const std::string& material_yaml);

// stem(geom)_stem(load)_stem(material)[_stem(numerics)] (CLI.f90:311-333)
std::string solver_job_name(const std::string& geom, const std::string& load, const std::string& material,
                            const std::string& numerics);

class ResultFile {
 public:
  ResultFile(const std::string& path, const ResultSetup& setup, int64_t n_cells);
  // One increment: increment_<inc> with attribute t/s; phase/<phase>/
  // mechanical/<output> from the device planes of mp (read on stream st);
  // homogenization/<name>/mechanical; geometry/u_p (N,3) and u_n (M,3) from
  // host displacements.
  void increment(int inc, double t, const MpState& mp, hipStream_t st, const std::vector<double>& u_p,
                 const std::vector<double>& u_n);
  const std::string& path() const { return path_; }

 private:
  void stamp(const std::string& obj, const std::string& description, const char* unit);
  std::string path_;
  ResultSetup s_;
  int64_t N_;
  Dadf5Writer w_;
};

}  // namespace sca
