// Job input readers (inputs.cpp).
#pragma once
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/solver_cpp_amd.h"

namespace sca {

// YAML node of the supported subset
struct Node {
  using Ptr = std::shared_ptr<Node>;
  enum Kind { SCALAR, SEQ, MAP } kind = SCALAR;
  std::string value;                              // SCALAR
  std::vector<Ptr> seq;                           // SEQ
  std::vector<std::pair<std::string, Ptr>> map;   // MAP, in file order
  const Node* get(const std::string& key) const;  // MAP lookup, nullptr if absent
  double num() const;                             // SCALAR as a number
};

Node::Ptr parse_yaml(const std::string& text);
std::string read_file(const std::string& path);

struct MaterialInput {
  sc_material material;
  std::string phase_name;
  std::string homogenization_name;
  std::vector<std::string> mechanical_output;  // phase/<name>/mechanical/output
  std::vector<std::string> plastic_output;     // phase/<name>/mechanical/plastic/output
  std::vector<double> orientations;  // (4, n_materials), w x y z
};
MaterialInput read_material(const std::string& text);
void read_numerics(const std::string& text, sc_grid_numerics& g, sc_phase_numerics& p);
std::vector<sc_loadstep> read_load(const std::string& text);

void vti_cells_size_origin(const std::string& content, int cells[3], double size[3], double origin[3]);
std::vector<int64_t> vti_dataset_int(const std::string& content, const std::string& label);  // 0-based as stored

}  // namespace sca
