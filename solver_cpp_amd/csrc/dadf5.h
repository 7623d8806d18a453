// DADF5 result files (the reference's result layout, result.f90:76-166,
// materialpoint.f90:221-234) written by an own minimal HDF5 writer:
// superblock version 0, old-style groups (symbol tables: version-1 B-tree,
// symbol nodes, local heap), version-1 object headers, and datasets of
// little-endian doubles, fixed-length strings or (label, entry) records.
// Numeric datasets are stored as the reference stores them
// (HDF5_utilities.f90:1982-2003): chunked, every chunk with a Fletcher-32
// checksum; datasets of at least 2 * 2^17 values in chunks of 2^17 values
// along the cell dimension, shuffled and deflated (level 6) when compression
// is on.  Raw data is appended as it is added (chunks filtered in parallel);
// flush() appends the metadata of the whole tree and rewrites the
// superblock, so the file is a complete HDF5 file after every flush.
#pragma once
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace sca {

class Dadf5Writer {
 public:
  explicit Dadf5Writer(const std::string& path, bool compression = true);
  ~Dadf5Writer();
  Dadf5Writer(const Dadf5Writer&) = delete;
  Dadf5Writer& operator=(const Dadf5Writer&) = delete;

  // datasets ("a/b/name"; parent groups are created).  dims in C order, the
  // first dimension is the one that is chunked.
  void dataset_f64(const std::string& path, const std::vector<uint64_t>& dims, const double* data);
  void dataset_str(const std::string& path, const std::string& text);  // shape (1,), fixed length
  // (label, entry) records of result_mapping_phase / _homogenization
  // (result.f90:492-643): labels space padded to label_len, 64-bit entries
  void dataset_label_entry(const std::string& path, const std::vector<uint64_t>& dims, uint32_t label_len,
                           const std::vector<std::string>& labels, const std::vector<int64_t>& entries);
  // attributes of a group or dataset ("" = root)
  void attr_f64(const std::string& path, const std::string& name, double v);
  void attr_i32(const std::string& path, const std::string& name, int32_t v);
  void attr_str(const std::string& path, const std::string& name, const std::string& v);
  void attr_str_array(const std::string& path, const std::string& name, const std::vector<std::string>& v);
  void group(const std::string& path);
  bool exists(const std::string& path) const;
  void flush();

 private:
  struct Attr {
    std::string name;
    int kind;  // 0 f64, 1 i32, 2 string, 3 string array
    double f;
    int32_t i;
    std::vector<std::string> s;
  };
  struct Chunk {
    uint64_t row0, addr, size;
  };
  struct Node {
    bool is_group = true;
    std::map<std::string, std::unique_ptr<Node>> children;  // sorted by name
    std::vector<Attr> attrs;
    // dataset
    int kind = 0;  // 0 f64, 2 string, 3 label-entry record
    std::vector<uint64_t> dims;
    uint64_t str_len = 0;  // string length, or label length of a record
    uint64_t elem = 8;     // bytes per element
    bool chunked = false, shuffle = false, deflate = false;
    uint64_t chunk_rows = 0;
    std::vector<Chunk> chunks;
    uint64_t data_addr = 0, data_size = 0;  // contiguous
  };
  Node* find(const std::string& path, bool create_groups);
  Node* new_dataset(const std::string& path);
  uint64_t append(const void* data, uint64_t size);
  void write_chunked(Node& n, const char* bytes);
  uint64_t write_node(Node& n, std::vector<char>& meta, uint64_t base);
  uint64_t write_chunk_btree(const Node& n, std::vector<char>& meta, uint64_t base);

  std::FILE* f_ = nullptr;
  Node root_;
  uint64_t eof_ = 0;
  bool compression_ = true;
};

// Fletcher-32 of HDF5's checksum filter: 16-bit big-endian words, sums
// reduced modulo 65535 every 360 words, an odd last byte as a high byte
uint32_t fletcher32(const uint8_t* data, size_t n);

// quaternion (w, x, y, z), w >= 0, of a rotation matrix (column-major 3x3),
// the inverse of qu2om (rotations.f90 conventions: P = -1)
void om2qu_host(const double om[9], double q[4]);

// rotational part R of F = R U (math_rotationalPart, math.f90), column-major
void rotational_part_host(const double F[9], double R[9]);

}  // namespace sca
