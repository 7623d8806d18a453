// Minimal HDF5 writer for DADF5 result files (layout of dadf5.h).
//
// File layout: superblock (96 bytes at 0) | raw data and chunks, appended as
// added | metadata of the whole tree, appended at every flush.  A flush
// rewrites the superblock to point at the newest metadata; older metadata
// stays in the file unreferenced.  Structures follow the published HDF5 file
// format (superblock version 0; version-1 object headers with 8-byte aligned
// messages; symbol-table groups: version-1 B-tree node of type 0 over symbol
// table nodes of at most 2 * 4 entries, names in a local heap; chunk index:
// version-1 B-tree of type 1 with 2 * 32 entries per node (the default
// indexed-storage K of a version-0 superblock); dataspace version 1;
// datatype version 1; fill value version 2; layout version 3; filter
// pipeline version 1; attribute version 1).
#include "dadf5.h"

#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <thread>

namespace sca {
namespace {

constexpr uint64_t UNDEF = ~(uint64_t)0;
constexpr int LEAF_K = 4;        // symbol table node: up to 2K entries
constexpr int INTERNAL_K = 16;   // group B-tree: up to 2K children
constexpr int ISTORE_K = 32;     // chunk B-tree: up to 2K children
constexpr uint64_t SUPERBLOCK_SIZE = 96;
constexpr uint64_t CHUNK_VALUES = 1024 * 1024 / 8;  // HDF5_utilities.f90:1952

struct Buf {
  std::vector<char>& b;
  void u8(uint8_t v) { b.push_back((char)v); }
  void u16(uint16_t v) { put(&v, 2); }
  void u32(uint32_t v) { put(&v, 4); }
  void u64(uint64_t v) { put(&v, 8); }
  void put(const void* p, size_t n) {
    const char* c = (const char*)p;
    b.insert(b.end(), c, c + n);
  }
  void zeros(size_t n) { b.insert(b.end(), n, 0); }
  void align8() {
    while (b.size() % 8) b.push_back(0);
  }
};

// datatype message bodies
void dtype_f64(std::vector<char>& m) {
  Buf w{m};
  w.u8(0x11);  // version 1, class 1 (floating point)
  w.u8(0x20);  // little endian, no padding, mantissa normalised (implied msb)
  w.u8(63);    // sign bit
  w.u8(0);
  w.u32(8);
  w.u16(0);     // bit offset
  w.u16(64);    // precision
  w.u8(52);     // exponent location
  w.u8(11);     // exponent size
  w.u8(0);      // mantissa location
  w.u8(52);     // mantissa size
  w.u32(1023);  // exponent bias
}
void dtype_int(std::vector<char>& m, uint32_t bytes) {
  Buf w{m};
  w.u8(0x10);  // version 1, class 0 (fixed point)
  w.u8(0x08);  // little endian, signed
  w.u8(0);
  w.u8(0);
  w.u32(bytes);
  w.u16(0);
  w.u16((uint16_t)(8 * bytes));
}
// pad: 0 null terminated, 2 space padded (Fortran character data)
void dtype_str(std::vector<char>& m, uint64_t len, int pad = 0) {
  Buf w{m};
  w.u8(0x13);  // version 1, class 3 (string)
  w.u8((uint8_t)pad);
  w.u8(0);
  w.u8(0);
  w.u32((uint32_t)len);
}
// compound {label: string(label_len, space padded), entry: int64}
void dtype_label_entry(std::vector<char>& m, uint64_t label_len) {
  Buf w{m};
  w.u8(0x16);  // version 1, class 6 (compound)
  w.u16(2);    // members
  w.u8(0);
  w.u32((uint32_t)(label_len + 8));
  auto member = [&](const char* name, uint32_t offset, const std::vector<char>& type) {
    const size_t n = std::strlen(name) + 1;
    w.put(name, n);
    w.zeros((8 - n % 8) % 8);
    w.u32(offset);
    w.u8(0);  // dimensionality
    w.zeros(3);
    w.u32(0);  // dimension permutation
    w.u32(0);
    w.zeros(16);  // dimension sizes
    w.put(type.data(), type.size());
  };
  std::vector<char> ts, ti;
  dtype_str(ts, label_len, 2);
  dtype_int(ti, 8);
  member("label", 0, ts);
  member("entry", (uint32_t)label_len, ti);
}
void dataspace(std::vector<char>& m, const std::vector<uint64_t>& dims) {
  Buf w{m};
  w.u8(1);  // version
  w.u8((uint8_t)dims.size());
  w.u8(0);  // no maximum dimensions
  w.u8(0);
  w.u32(0);
  for (uint64_t d : dims) w.u64(d);
}

// one message of a version-1 object header
void message(std::vector<char>& oh, uint16_t type, const std::vector<char>& body) {
  Buf w{oh};
  const size_t size = (body.size() + 7) & ~(size_t)7;
  w.u16(type);
  w.u16((uint16_t)size);
  w.u8(0);  // flags
  w.zeros(3);
  w.put(body.data(), body.size());
  w.zeros(size - body.size());
}

std::vector<char> attribute_body(const std::string& name, int kind, double f, int32_t i,
                                 const std::vector<std::string>& s) {
  std::vector<char> dt, ds, data;
  if (kind == 0) {
    dtype_f64(dt);
    data.resize(8);
    std::memcpy(data.data(), &f, 8);
    dataspace(ds, {});  // scalar
  } else if (kind == 1) {
    dtype_int(dt, 4);
    data.resize(4);
    std::memcpy(data.data(), &i, 4);
    dataspace(ds, {});
  } else {
    size_t len = 1;
    for (const auto& x : s) len = std::max(len, x.size() + 1);
    dtype_str(dt, len);
    for (const auto& x : s) {
      data.insert(data.end(), x.begin(), x.end());
      data.insert(data.end(), len - x.size(), 0);
    }
    if (kind == 2)
      dataspace(ds, {});
    else
      dataspace(ds, {(uint64_t)s.size()});
  }
  std::vector<char> b;
  Buf w{b};
  w.u8(1);  // version
  w.u8(0);
  w.u16((uint16_t)(name.size() + 1));
  w.u16((uint16_t)dt.size());
  w.u16((uint16_t)ds.size());
  w.put(name.c_str(), name.size() + 1);
  w.align8();
  w.put(dt.data(), dt.size());
  w.align8();
  w.put(ds.data(), ds.size());
  w.align8();
  w.put(data.data(), data.size());
  return b;
}

// object header (version 1) from its messages
std::vector<char> object_header(const std::vector<std::pair<uint16_t, std::vector<char>>>& msgs) {
  std::vector<char> body;
  for (auto& m : msgs) message(body, m.first, m.second);
  if (body.size() > 0xFFFFFFFFu) throw std::runtime_error("result object header too large");
  std::vector<char> oh;
  Buf w{oh};
  w.u8(1);
  w.u8(0);
  w.u16((uint16_t)msgs.size());
  w.u32(1);  // reference count
  w.u32((uint32_t)body.size());
  w.u32(0);  // pad the 12-byte prefix to 16
  w.put(body.data(), body.size());
  return oh;
}

// filter pipeline message (version 1): Fletcher-32 first, then shuffle and
// deflate -- the order of H5Pset_Fletcher32, H5Pset_shuffle, H5Pset_deflate
// in initialize_write (HDF5_utilities.f90:1986-1998)
std::vector<char> filter_pipeline(bool shuffle, bool deflate, uint32_t elem) {
  std::vector<char> b;
  Buf w{b};
  w.u8(1);
  w.u8((uint8_t)(1 + shuffle + deflate));
  w.zeros(6);
  auto filter = [&](uint16_t id, const char* name, uint16_t flags, const std::vector<uint32_t>& cd) {
    const size_t n = std::strlen(name) + 1, padded = (n + 7) & ~(size_t)7;
    w.u16(id);
    w.u16((uint16_t)padded);
    w.u16(flags);
    w.u16((uint16_t)cd.size());
    w.put(name, n);
    w.zeros(padded - n);
    for (uint32_t v : cd) w.u32(v);
    if (cd.size() % 2) w.u32(0);
  };
  filter(3, "fletcher32", 0, {});
  if (shuffle) filter(2, "shuffle", 1, {elem});
  if (deflate) filter(1, "deflate", 1, {6});
  return b;
}

// filters of one chunk, in pipeline order
std::vector<char> filter_chunk(const char* raw, size_t n, bool shuffle, bool deflate, uint32_t elem) {
  std::vector<char> a(raw, raw + n);
  const uint32_t sum = fletcher32((const uint8_t*)a.data(), a.size());
  const char* sb = (const char*)&sum;
  a.insert(a.end(), sb, sb + 4);  // stored little endian
  if (shuffle && elem > 1) {
    const size_t m = a.size() / elem;
    std::vector<char> s(a.size());
    for (size_t b = 0; b < elem; ++b)
      for (size_t i = 0; i < m; ++i) s[b * m + i] = a[i * elem + b];
    std::copy(a.begin() + m * elem, a.end(), s.begin() + m * elem);
    a.swap(s);
  }
  if (deflate) {
    uLongf len = compressBound((uLong)a.size());
    std::vector<char> z(len);
    if (compress2((Bytef*)z.data(), &len, (const Bytef*)a.data(), (uLong)a.size(), 6) != Z_OK)
      throw std::runtime_error("deflate of a result chunk failed");
    z.resize(len);
    a.swap(z);
  }
  return a;
}

}  // namespace

uint32_t fletcher32(const uint8_t* d, size_t n) {
  uint32_t s1 = 0, s2 = 0;
  size_t words = n / 2;
  while (words) {
    size_t t = std::min<size_t>(words, 360);
    words -= t;
    do {
      s1 += ((uint32_t)d[0] << 8) | d[1];
      s2 += s1;
      d += 2;
    } while (--t);
    s1 = (s1 & 0xffff) + (s1 >> 16);
    s2 = (s2 & 0xffff) + (s2 >> 16);
  }
  if (n & 1) {
    s1 += (uint32_t)d[0] << 8;
    s2 += s1;
    s1 = (s1 & 0xffff) + (s1 >> 16);
    s2 = (s2 & 0xffff) + (s2 >> 16);
  }
  s1 = (s1 & 0xffff) + (s1 >> 16);
  s2 = (s2 & 0xffff) + (s2 >> 16);
  return (s2 << 16) | s1;
}

Dadf5Writer::Dadf5Writer(const std::string& path, bool compression) : compression_(compression) {
  f_ = std::fopen(path.c_str(), "wb+");
  if (!f_) throw std::runtime_error("cannot create " + path);
  std::vector<char> zero(SUPERBLOCK_SIZE, 0);
  append(zero.data(), zero.size());
}

Dadf5Writer::~Dadf5Writer() {
  if (f_) {
    try {
      flush();
    } catch (...) {
    }
    std::fclose(f_);
  }
}

uint64_t Dadf5Writer::append(const void* data, uint64_t size) {
  const uint64_t at = eof_;
  if (std::fseek(f_, (long)at, SEEK_SET) != 0 || std::fwrite(data, 1, size, f_) != size)
    throw std::runtime_error("result file write failed");
  eof_ += size;
  // keep 8-byte alignment of everything that follows
  static const char pad[8] = {0};
  const uint64_t p = (8 - eof_ % 8) % 8;
  if (p) {
    if (std::fwrite(pad, 1, p, f_) != p) throw std::runtime_error("result file write failed");
    eof_ += p;
  }
  return at;
}

Dadf5Writer::Node* Dadf5Writer::find(const std::string& path, bool create) {
  Node* n = &root_;
  size_t a = 0;
  while (a < path.size()) {
    size_t b = path.find('/', a);
    if (b == std::string::npos) b = path.size();
    const std::string part = path.substr(a, b - a);
    a = b + 1;
    if (part.empty()) continue;
    if (!n->is_group) throw std::runtime_error("result object is not a group: " + path);
    auto it = n->children.find(part);
    if (it == n->children.end()) {
      if (!create) throw std::runtime_error("no result object " + path);
      it = n->children.emplace(part, std::make_unique<Node>()).first;
    }
    n = it->second.get();
  }
  return n;
}

bool Dadf5Writer::exists(const std::string& path) const {
  try {
    const_cast<Dadf5Writer*>(this)->find(path, false);
    return true;
  } catch (const std::runtime_error&) {
    return false;
  }
}

Dadf5Writer::Node* Dadf5Writer::new_dataset(const std::string& path) {
  const size_t s = path.rfind('/');
  Node* parent = (s == std::string::npos) ? &root_ : find(path.substr(0, s), true);
  const std::string name = (s == std::string::npos) ? path : path.substr(s + 1);
  if (name.empty()) throw std::runtime_error("empty result dataset name: " + path);
  if (!parent->is_group) throw std::runtime_error("result object is not a group: " + path);
  if (parent->children.count(name)) throw std::runtime_error("result object exists: " + path);
  auto& slot = parent->children[name];
  slot = std::make_unique<Node>();
  slot->is_group = false;
  return slot.get();
}

void Dadf5Writer::group(const std::string& path) { find(path, true); }

// Split the data into chunks of chunk_rows rows (the last one zero padded to
// a full chunk, as HDF5 stores edge chunks), filter them in parallel and
// append them in order.
void Dadf5Writer::write_chunked(Node& n, const char* bytes) {
  uint64_t row = n.elem;
  for (size_t d = 1; d < n.dims.size(); ++d) row *= n.dims[d];
  const uint64_t rows = n.dims.empty() ? 1 : n.dims[0];
  const uint64_t nchunks = (rows + n.chunk_rows - 1) / n.chunk_rows;
  const uint64_t cbytes = n.chunk_rows * row;
  std::vector<std::vector<char>> out(nchunks);
  std::atomic<uint64_t> next{0};
  std::mutex mu;
  std::string err;
  auto work = [&] {
    try {
      std::vector<char> buf;
      for (uint64_t k; (k = next.fetch_add(1)) < nchunks;) {
        const uint64_t r0 = k * n.chunk_rows, r1 = std::min(rows, r0 + n.chunk_rows);
        const char* src = bytes + r0 * row;
        if (r1 - r0 < n.chunk_rows) {
          buf.assign(cbytes, 0);
          std::memcpy(buf.data(), src, (r1 - r0) * row);
          src = buf.data();
        }
        out[k] = filter_chunk(src, cbytes, n.shuffle, n.deflate, (uint32_t)n.elem);
      }
    } catch (const std::exception& e) {
      std::lock_guard<std::mutex> lock(mu);
      err = e.what();
      next = nchunks;
    }
  };
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const unsigned nt = (unsigned)std::min<uint64_t>({nchunks, hw, 16, n.deflate ? 16u : 4u});
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  if (!err.empty()) throw std::runtime_error(err);
  n.chunks.resize(nchunks);
  for (uint64_t k = 0; k < nchunks; ++k) {
    n.chunks[k] = {k * n.chunk_rows, append(out[k].data(), out[k].size()), (uint64_t)out[k].size()};
    std::vector<char>().swap(out[k]);
  }
}

void Dadf5Writer::dataset_f64(const std::string& path, const std::vector<uint64_t>& dims, const double* data) {
  if (dims.empty()) throw std::runtime_error("result dataset without dimensions: " + path);
  Node* n = new_dataset(path);
  n->kind = 0;
  n->elem = 8;
  n->dims = dims;
  uint64_t total = 1, inner = 1;
  for (size_t d = 0; d < dims.size(); ++d) {
    total *= dims[d];
    if (d) inner *= dims[d];
  }
  if (total == 0) {  // no filters, no storage (HDF5_utilities.f90:1986)
    n->data_addr = UNDEF;
    return;
  }
  n->chunked = true;
  if (total >= 2 * CHUNK_VALUES) {  // getChunks, HDF5_utilities.f90:2026-2035
    n->chunk_rows = std::max<uint64_t>(1, CHUNK_VALUES / inner);
    n->shuffle = n->deflate = compression_;
  } else {
    n->chunk_rows = dims[0];
  }
  write_chunked(*n, (const char*)data);
}

// HDF5_write_str (HDF5_utilities.f90:1453-1500): chunk (1), Fletcher-32,
// shuffle and deflate above 256 KiB
void Dadf5Writer::dataset_str(const std::string& path, const std::string& text) {
  Node* n = new_dataset(path);
  n->kind = 2;
  n->dims = {1};
  n->str_len = text.size() + 1;
  n->elem = n->str_len;
  n->chunked = true;
  n->chunk_rows = 1;
  n->shuffle = n->deflate = compression_ && text.size() > 1024 * 256;
  write_chunked(*n, text.c_str());
}

void Dadf5Writer::dataset_label_entry(const std::string& path, const std::vector<uint64_t>& dims, uint32_t label_len,
                                      const std::vector<std::string>& labels, const std::vector<int64_t>& entries) {
  uint64_t total = 1;
  for (uint64_t d : dims) total *= d;
  if (labels.size() != total || entries.size() != total)
    throw std::runtime_error("label/entry records do not match the shape of " + path);
  if (label_len == 0) throw std::runtime_error("empty labels in " + path);
  Node* n = new_dataset(path);
  n->kind = 3;
  n->dims = dims;
  n->str_len = label_len;
  n->elem = label_len + 8;
  std::vector<char> rec(total * n->elem, ' ');
  for (uint64_t c = 0; c < total; ++c) {
    if (labels[c].size() > label_len) throw std::runtime_error("label longer than its field in " + path);
    std::memcpy(&rec[c * n->elem], labels[c].data(), labels[c].size());
    std::memcpy(&rec[c * n->elem + label_len], &entries[c], 8);
  }
  n->data_size = rec.size();
  n->data_addr = append(rec.data(), rec.size());
}

void Dadf5Writer::attr_f64(const std::string& path, const std::string& name, double v) {
  find(path, true)->attrs.push_back({name, 0, v, 0, {}});
}
void Dadf5Writer::attr_i32(const std::string& path, const std::string& name, int32_t v) {
  find(path, true)->attrs.push_back({name, 1, 0.0, v, {}});
}
void Dadf5Writer::attr_str(const std::string& path, const std::string& name, const std::string& v) {
  find(path, true)->attrs.push_back({name, 2, 0.0, 0, {v}});
}
void Dadf5Writer::attr_str_array(const std::string& path, const std::string& name, const std::vector<std::string>& v) {
  find(path, true)->attrs.push_back({name, 3, 0.0, 0, v});
}

// Version-1 B-tree of type 1 over the chunks of n, built bottom up; every
// node is allocated for 2 * ISTORE_K children.  A key holds the chunk's
// stored size, its filter mask and its offset per dimension (plus a zero
// for the element dimension); the key after the last child is the last
// chunk's offset plus the chunk dimensions.  Returns the root's address.
uint64_t Dadf5Writer::write_chunk_btree(const Node& n, std::vector<char>& meta, uint64_t base) {
  const size_t rank = n.dims.size();
  struct Key {
    uint32_t size;
    std::vector<uint64_t> off;
  };
  auto chunk_key = [&](const Chunk& c) {
    Key k{(uint32_t)c.size, std::vector<uint64_t>(rank + 1, 0)};
    k.off[0] = c.row0;
    return k;
  };
  Key right{0, std::vector<uint64_t>(rank + 1, 0)};
  right.off[0] = n.chunks.back().row0 + n.chunk_rows;
  for (size_t d = 1; d < rank; ++d) right.off[d] = n.dims[d];
  const size_t key_size = 8 + 8 * (rank + 1);
  const size_t cap = 2 * ISTORE_K;
  const size_t node_size = 24 + (cap + 1) * key_size + cap * 8;
  // level 0 entries: (left key, child address)
  std::vector<std::pair<Key, uint64_t>> entries;
  for (const Chunk& c : n.chunks) entries.emplace_back(chunk_key(c), c.addr);
  for (int level = 0;; ++level) {
    const size_t nodes = (entries.size() + cap - 1) / cap;
    while (meta.size() % 8) meta.push_back(0);
    const uint64_t first = base + meta.size();
    std::vector<std::pair<Key, uint64_t>> up;
    for (size_t j = 0; j < nodes; ++j) {
      const size_t i0 = j * cap, i1 = std::min(entries.size(), i0 + cap);
      const size_t at = meta.size();
      Buf w{meta};
      w.put("TREE", 4);
      w.u8(1);
      w.u8((uint8_t)level);
      w.u16((uint16_t)(i1 - i0));
      w.u64(j ? first + (j - 1) * node_size : UNDEF);
      w.u64(j + 1 < nodes ? first + (j + 1) * node_size : UNDEF);
      auto key = [&](const Key& k) {
        w.u32(k.size);
        w.u32(0);  // filter mask: every filter applied
        for (uint64_t o : k.off) w.u64(o);
      };
      for (size_t i = i0; i < i1; ++i) {
        key(entries[i].first);
        w.u64(entries[i].second);
      }
      key(i1 < entries.size() ? entries[i1].first : right);
      meta.resize(at + node_size, 0);
      up.emplace_back(entries[i0].first, first + j * node_size);
    }
    if (nodes == 1) return first;
    entries.swap(up);
  }
}

// Serialise node n (and its subtree) into meta, whose first byte will be at
// file address base; returns the address of n's object header.
uint64_t Dadf5Writer::write_node(Node& n, std::vector<char>& meta, uint64_t base) {
  auto addr = [&] { return base + meta.size(); };
  std::vector<std::pair<uint16_t, std::vector<char>>> msgs;
  if (!n.is_group) {
    std::vector<char> ds, dt, fill, layout;
    dataspace(ds, n.dims);
    if (n.kind == 0) dtype_f64(dt);
    else if (n.kind == 2) dtype_str(dt, n.str_len);
    else dtype_label_entry(dt, n.str_len);
    {
      Buf w{fill};
      w.u8(2);  // version
      w.u8(n.chunked ? 3 : 1);  // space allocated incrementally / early
      w.u8(2);  // fill value never written
      w.u8(0);  // no fill value defined
    }
    msgs = {{0x0001, ds}, {0x0003, dt}, {0x0005, fill}};
    if (n.chunked) {
      const uint64_t bt = write_chunk_btree(n, meta, base);
      Buf w{layout};
      w.u8(3);  // version
      w.u8(2);  // chunked
      w.u8((uint8_t)(n.dims.size() + 1));
      w.u64(bt);
      w.u32((uint32_t)n.chunk_rows);
      for (size_t d = 1; d < n.dims.size(); ++d) w.u32((uint32_t)n.dims[d]);
      w.u32((uint32_t)n.elem);
      msgs.push_back({0x000B, filter_pipeline(n.shuffle, n.deflate, (uint32_t)n.elem)});
    } else {
      Buf w{layout};
      w.u8(3);
      w.u8(1);  // contiguous
      w.u64(n.data_addr);
      w.u64(n.data_size);
    }
    msgs.push_back({0x0008, layout});
  } else {
    // children first (their object header addresses go into the symbol table)
    std::vector<std::pair<std::string, uint64_t>> entries;
    for (auto& kv : n.children) entries.emplace_back(kv.first, write_node(*kv.second, meta, base));
    const size_t nsym = entries.size();
    const size_t per_node = 2 * LEAF_K;
    const size_t nnodes = (nsym + per_node - 1) / per_node;  // 0 for an empty group
    if (nnodes > 2 * (size_t)INTERNAL_K) throw std::runtime_error("result group with too many members");
    // local heap: "" at offset 0, then the names (8-byte padded)
    std::vector<char> heap_data(8, 0);
    std::vector<uint64_t> name_off(nsym);
    for (size_t i = 0; i < nsym; ++i) {
      name_off[i] = heap_data.size();
      const std::string& s = entries[i].first;
      heap_data.insert(heap_data.end(), s.begin(), s.end());
      heap_data.push_back(0);
      while (heap_data.size() % 8) heap_data.push_back(0);
    }
    while (meta.size() % 8) meta.push_back(0);
    const uint64_t heap_data_addr = addr();
    meta.insert(meta.end(), heap_data.begin(), heap_data.end());
    const uint64_t heap_addr = addr();
    {
      Buf w{meta};
      w.put("HEAP", 4);
      w.u8(0);
      w.zeros(3);
      w.u64(heap_data.size());
      w.u64(1);  // no free block (the HDF5 library's end-of-free-list value)
      w.u64(heap_data_addr);
    }
    // symbol table nodes (each allocated for 2K entries)
    std::vector<uint64_t> snod(nnodes);
    for (size_t k = 0; k < nnodes; ++k) {
      snod[k] = addr();
      const size_t i0 = k * per_node, i1 = std::min(nsym, i0 + per_node);
      Buf w{meta};
      w.put("SNOD", 4);
      w.u8(1);
      w.u8(0);
      w.u16((uint16_t)(i1 - i0));
      for (size_t i = i0; i < i1; ++i) {
        w.u64(name_off[i]);
        w.u64(entries[i].second);
        w.u32(0);  // no cached scratch data
        w.u32(0);
        w.zeros(16);
      }
      w.zeros((per_node - (i1 - i0)) * 40);
    }
    // B-tree node (type 0, leaf level) over the symbol table nodes; key k is
    // the heap offset of the last name in child k-1 (key 0: "")
    const uint64_t btree_addr = addr();
    {
      Buf w{meta};
      w.put("TREE", 4);
      w.u8(0);
      w.u8(0);
      w.u16((uint16_t)nnodes);
      w.u64(UNDEF);
      w.u64(UNDEF);
      w.u64(0);
      for (size_t k = 0; k < nnodes; ++k) {
        w.u64(snod[k]);
        const size_t last = std::min(nsym, (k + 1) * per_node);
        w.u64(name_off[last - 1]);
      }
      w.zeros((2 * INTERNAL_K - nnodes) * 16);
    }
    std::vector<char> stab;
    {
      Buf w{stab};
      w.u64(btree_addr);
      w.u64(heap_addr);
    }
    msgs = {{0x0011, stab}};
  }
  for (auto& a : n.attrs) msgs.push_back({0x000C, attribute_body(a.name, a.kind, a.f, a.i, a.s)});
  // pad meta so that the object header starts 8-aligned
  while (meta.size() % 8) meta.push_back(0);
  const uint64_t oh_addr = addr();
  const std::vector<char> oh = object_header(msgs);
  meta.insert(meta.end(), oh.begin(), oh.end());
  return oh_addr;
}

void Dadf5Writer::flush() {
  std::vector<char> meta;
  const uint64_t base = eof_;
  const uint64_t root_oh = write_node(root_, meta, base);
  // the root's B-tree and heap (cached in the superblock's root entry), read
  // back from the root object header's symbol-table message
  uint64_t root_btree = 0, root_heap = 0;
  {
    const size_t off = root_oh - base + 16 + 8;  // prefix + first message header
    std::memcpy(&root_btree, meta.data() + off, 8);
    std::memcpy(&root_heap, meta.data() + off + 8, 8);
  }
  append(meta.data(), meta.size());
  std::vector<char> sb;
  Buf w{sb};
  w.put("\x89HDF\r\n\x1a\n", 8);
  w.u8(0);  // superblock version
  w.u8(0);  // free space version
  w.u8(0);  // root group symbol table version
  w.u8(0);
  w.u8(0);  // shared header version
  w.u8(8);  // size of offsets
  w.u8(8);  // size of lengths
  w.u8(0);
  w.u16(LEAF_K);
  w.u16(INTERNAL_K);
  w.u32(0);      // consistency flags
  w.u64(0);      // base address
  w.u64(UNDEF);  // free space
  w.u64(eof_);   // end of file
  w.u64(UNDEF);  // driver information
  // root group symbol table entry
  w.u64(0);
  w.u64(root_oh);
  w.u32(1);  // cached: B-tree and heap of the root group
  w.u32(0);
  w.u64(root_btree);
  w.u64(root_heap);
  if (sb.size() != SUPERBLOCK_SIZE) throw std::runtime_error("superblock size");
  if (std::fseek(f_, 0, SEEK_SET) != 0 || std::fwrite(sb.data(), 1, sb.size(), f_) != sb.size())
    throw std::runtime_error("result file write failed");
  std::fflush(f_);
}

void om2qu_host(const double om[9], double q[4]) {
  // om is the passive matrix (transpose of the active rotation R)
  double R[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = om[j + 3 * i];
  const double tr = R[0][0] + R[1][1] + R[2][2];
  double w, x, y, z;
  if (tr > 0.0) {
    const double s = 2.0 * std::sqrt(tr + 1.0);
    w = 0.25 * s;
    x = (R[2][1] - R[1][2]) / s;
    y = (R[0][2] - R[2][0]) / s;
    z = (R[1][0] - R[0][1]) / s;
  } else if (R[0][0] > R[1][1] && R[0][0] > R[2][2]) {
    const double s = 2.0 * std::sqrt(1.0 + R[0][0] - R[1][1] - R[2][2]);
    w = (R[2][1] - R[1][2]) / s;
    x = 0.25 * s;
    y = (R[0][1] + R[1][0]) / s;
    z = (R[0][2] + R[2][0]) / s;
  } else if (R[1][1] > R[2][2]) {
    const double s = 2.0 * std::sqrt(1.0 + R[1][1] - R[0][0] - R[2][2]);
    w = (R[0][2] - R[2][0]) / s;
    x = (R[0][1] + R[1][0]) / s;
    y = 0.25 * s;
    z = (R[1][2] + R[2][1]) / s;
  } else {
    const double s = 2.0 * std::sqrt(1.0 + R[2][2] - R[0][0] - R[1][1]);
    w = (R[1][0] - R[0][1]) / s;
    x = (R[0][2] + R[2][0]) / s;
    y = (R[1][2] + R[2][1]) / s;
    z = 0.25 * s;
  }
  const double n = std::sqrt(w * w + x * x + y * y + z * z);
  const double sg = (w < 0.0) ? -1.0 : 1.0;
  q[0] = sg * w / n;
  q[1] = sg * x / n;
  q[2] = sg * y / n;
  q[3] = sg * z / n;
}

// polar decomposition F = R U by the scaled Newton iteration
// R <- (g R + R^-T / g) / 2, g = (|R^-1|_F / |R|_F)^(1/2) (Higham 1986)
void rotational_part_host(const double* F, double* R) {
  double X[9];
  std::memcpy(X, F, sizeof X);
  for (int it = 0; it < 100; ++it) {
    double inv[9];
    const double det = X[0] * (X[4] * X[8] - X[7] * X[5]) - X[3] * (X[1] * X[8] - X[7] * X[2]) +
                       X[6] * (X[1] * X[5] - X[4] * X[2]);
    if (det == 0.0) throw std::runtime_error("rotational part of a singular tensor");
    inv[0] = (X[4] * X[8] - X[7] * X[5]) / det;
    inv[3] = -(X[3] * X[8] - X[6] * X[5]) / det;
    inv[6] = (X[3] * X[7] - X[6] * X[4]) / det;
    inv[1] = -(X[1] * X[8] - X[7] * X[2]) / det;
    inv[4] = (X[0] * X[8] - X[6] * X[2]) / det;
    inv[7] = -(X[0] * X[7] - X[6] * X[1]) / det;
    inv[2] = (X[1] * X[5] - X[4] * X[2]) / det;
    inv[5] = -(X[0] * X[5] - X[3] * X[2]) / det;
    inv[8] = (X[0] * X[4] - X[3] * X[1]) / det;
    double nx = 0.0, ni = 0.0;
    for (int k = 0; k < 9; ++k) {
      nx += X[k] * X[k];
      ni += inv[k] * inv[k];
    }
    const double g = std::sqrt(std::sqrt(ni / nx));
    double Y[9], d = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        Y[i + 3 * j] = 0.5 * (g * X[i + 3 * j] + inv[j + 3 * i] / g);
        d = std::max(d, std::fabs(Y[i + 3 * j] - X[i + 3 * j]));
      }
    std::memcpy(X, Y, sizeof X);
    if (d <= 4.0 * std::numeric_limits<double>::epsilon()) break;
  }
  std::memcpy(R, X, sizeof X);
}

}  // namespace sca
