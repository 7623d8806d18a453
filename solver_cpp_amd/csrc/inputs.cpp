// Job inputs: VTK image data geometry (*.vti), material.yaml, numerics.yaml
// and load-case YAML, read on the host into the C-ABI structs.
//
//   VTI     src/grid/VTI.f90 (VTI_readCellsSizeOrigin :146-202, cellsSizeOrigin
//           :207-233, base64 + zlib blocks :326-397, readDataset_int :473-497:
//           the dataset is returned +1, i.e. 1-based material indices)
//   load    src/grid/config.cpp:240-360 (parse_load), numerics :362-460
//   material material.f90:84-205, phase_mechanical_elastic.f90:74-104,
//           phase_mechanical_plastic_phenopowerlaw.f90:75-290
//
// The YAML reader covers the subset these files use: block mappings and
// sequences by indentation, "- key: value" sequence items, flow sequences and
// flow mappings (possibly over several lines), plain and quoted scalars,
// comments and document markers.
#include <zlib.h>

#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "common.h"
#include "inputs.h"

namespace sca {
extern thread_local std::string g_err;

// ---------------------------------------------------------------------------
// YAML subset
// ---------------------------------------------------------------------------
namespace yaml {

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

std::string unquote(const std::string& s) {
  if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
    return s.substr(1, s.size() - 2);
  return s;
}

// drop a trailing comment that is not inside quotes
std::string strip_comment(const std::string& s) {
  bool sq = false, dq = false;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (c == '#' && !sq && !dq && (i == 0 || std::isspace((unsigned char)s[i - 1]))) return s.substr(0, i);
  }
  return s;
}

int bracket_balance(const std::string& s) {
  int b = 0;
  bool sq = false, dq = false;
  for (char c : s) {
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (!sq && !dq) {
      if (c == '[' || c == '{') ++b;
      if (c == ']' || c == '}') --b;
    }
  }
  return b;
}

// position of the ':' that separates a mapping key (followed by space or end)
size_t key_colon(const std::string& s) {
  bool sq = false, dq = false;
  int depth = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (!sq && !dq) {
      if (c == '[' || c == '{') ++depth;
      else if (c == ']' || c == '}') --depth;
      else if (c == ':' && depth == 0 && (i + 1 == s.size() || s[i + 1] == ' ')) return i;
    }
  }
  return std::string::npos;
}

struct Line {
  int indent;
  std::string text;
  int lineno;
};

Node::Ptr parse_flow(const std::string& s, size_t& i);

void skip_ws(const std::string& s, size_t& i) {
  while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
}

Node::Ptr scalar(const std::string& v) {
  auto n = std::make_shared<Node>();
  n->kind = Node::SCALAR;
  n->value = unquote(trim(v));
  return n;
}

Node::Ptr parse_flow(const std::string& s, size_t& i) {
  skip_ws(s, i);
  if (i < s.size() && (s[i] == '[' || s[i] == '{')) {
    const bool is_seq = s[i] == '[';
    const char close = is_seq ? ']' : '}';
    auto n = std::make_shared<Node>();
    n->kind = is_seq ? Node::SEQ : Node::MAP;
    ++i;
    for (;;) {
      skip_ws(s, i);
      if (i >= s.size()) throw std::runtime_error("YAML: unterminated flow collection");
      if (s[i] == close) {
        ++i;
        break;
      }
      if (is_seq) {
        n->seq.push_back(parse_flow(s, i));
      } else {
        size_t j = i;
        while (j < s.size() && s[j] != ':') ++j;
        if (j >= s.size()) throw std::runtime_error("YAML: flow mapping without ':'");
        const std::string key = unquote(trim(s.substr(i, j - i)));
        i = j + 1;
        n->map.emplace_back(key, parse_flow(s, i));
      }
      skip_ws(s, i);
      if (i < s.size() && s[i] == ',') ++i;
    }
    return n;
  }
  // plain or quoted scalar up to , ] }
  size_t j = i;
  bool sq = false, dq = false;
  while (j < s.size()) {
    const char c = s[j];
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (!sq && !dq && (c == ',' || c == ']' || c == '}')) break;
    ++j;
  }
  auto n = scalar(s.substr(i, j - i));
  i = j;
  return n;
}

Node::Ptr value_node(const std::string& v) {
  const std::string t = trim(v);
  if (!t.empty() && (t[0] == '[' || t[0] == '{')) {
    size_t i = 0;
    return parse_flow(t, i);
  }
  return scalar(t);
}

Node::Ptr parse_block(const std::vector<Line>& L, size_t& k, int indent);

// "key: value" items of a mapping at this indent; `first` is an inline first
// item (of a "- key: value" sequence entry)
Node::Ptr parse_map(const std::vector<Line>& L, size_t& k, int indent, const std::string* first) {
  auto n = std::make_shared<Node>();
  n->kind = Node::MAP;
  auto add_item = [&](const std::string& text, int item_indent) {
    const size_t c = key_colon(text);
    if (c == std::string::npos) throw std::runtime_error("YAML: expected 'key: value' in '" + text + "'");
    const std::string key = unquote(trim(text.substr(0, c)));
    const std::string rest = trim(text.substr(c + 1));
    if (!rest.empty()) {
      n->map.emplace_back(key, value_node(rest));
    } else if (k < L.size() && (L[k].indent > item_indent || (L[k].indent == item_indent && L[k].text.rfind("- ", 0) == 0) ||
                                (L[k].indent == item_indent && L[k].text == "-"))) {
      n->map.emplace_back(key, parse_block(L, k, L[k].indent));
    } else {
      n->map.emplace_back(key, scalar(""));
    }
  };
  if (first) add_item(*first, indent);
  while (k < L.size() && L[k].indent == indent && L[k].text.rfind("- ", 0) != 0 && L[k].text != "-") {
    const std::string text = L[k].text;
    ++k;
    add_item(text, indent);
  }
  return n;
}

Node::Ptr parse_seq(const std::vector<Line>& L, size_t& k, int indent) {
  auto n = std::make_shared<Node>();
  n->kind = Node::SEQ;
  while (k < L.size() && L[k].indent == indent && (L[k].text.rfind("- ", 0) == 0 || L[k].text == "-")) {
    const std::string rest = L[k].text == "-" ? std::string() : trim(L[k].text.substr(2));
    const int item_indent = indent + 2;
    ++k;
    if (rest.empty()) {
      n->seq.push_back(k < L.size() && L[k].indent > indent ? parse_block(L, k, L[k].indent) : scalar(""));
    } else if (rest[0] != '[' && rest[0] != '{' && key_colon(rest) != std::string::npos) {
      // "- key: value" starts a mapping whose further keys sit at indent + 2
      n->seq.push_back(parse_map(L, k, item_indent, &rest));
    } else {
      n->seq.push_back(value_node(rest));
    }
  }
  return n;
}

Node::Ptr parse_block(const std::vector<Line>& L, size_t& k, int indent) {
  if (k >= L.size()) return scalar("");
  if (L[k].text.rfind("- ", 0) == 0 || L[k].text == "-") return parse_seq(L, k, indent);
  return parse_map(L, k, indent, nullptr);
}

Node::Ptr parse(const std::string& text) {
  std::vector<Line> L;
  std::istringstream in(text);
  std::string raw;
  int lineno = 0;
  std::string pending;
  int pending_indent = 0, pending_line = 0;
  while (std::getline(in, raw)) {
    ++lineno;
    if (!raw.empty() && raw.back() == '\r') raw.pop_back();
    std::string s = strip_comment(raw);
    if (trim(s).empty()) continue;
    if (trim(s) == "---" || trim(s) == "...") continue;
    if (!pending.empty()) {  // continuation of an open flow collection
      pending += " " + trim(s);
      if (bracket_balance(pending) <= 0) {
        L.push_back({pending_indent, pending, pending_line});
        pending.clear();
      }
      continue;
    }
    int ind = 0;
    while (ind < (int)s.size() && s[ind] == ' ') ++ind;
    const std::string t = trim(s);
    if (bracket_balance(t) > 0) {
      pending = t;
      pending_indent = ind;
      pending_line = lineno;
      continue;
    }
    L.push_back({ind, t, lineno});
  }
  if (!pending.empty()) throw std::runtime_error("YAML: unterminated flow collection at line " + std::to_string(pending_line));
  size_t k = 0;
  if (L.empty()) return std::make_shared<Node>();
  auto root = parse_block(L, k, L[0].indent);
  if (k != L.size()) throw std::runtime_error("YAML: unexpected indentation at line " + std::to_string(L[k].lineno));
  return root;
}

}  // namespace yaml

const Node* Node::get(const std::string& key) const {
  if (kind != MAP) return nullptr;
  for (const auto& kv : map)
    if (kv.first == key) return kv.second.get();
  return nullptr;
}

double Node::num() const {
  if (kind != SCALAR || value.empty()) throw std::runtime_error("YAML: expected a number");
  char* end = nullptr;
  const double v = std::strtod(value.c_str(), &end);
  if (end == value.c_str() || *end != '\0') throw std::runtime_error("YAML: '" + value + "' is not a number");
  return v;
}

std::string read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::ostringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

Node::Ptr parse_yaml(const std::string& text) { return yaml::parse(text); }

// ---------------------------------------------------------------------------
// VTI
// ---------------------------------------------------------------------------
namespace {

int b64val(char c) {
  if (c >= 'A' && c <= 'Z') return c - 'A';
  if (c >= 'a' && c <= 'z') return c - 'a' + 26;
  if (c >= '0' && c <= '9') return c - '0' + 52;
  if (c == '+') return 62;
  if (c == '/') return 63;
  return -1;
}

// decode base64 characters [0, n) of s (n a multiple of 4; '=' padding allowed)
std::vector<unsigned char> b64decode(const std::string& s, size_t n) {
  std::vector<unsigned char> out;
  out.reserve(n / 4 * 3);
  for (size_t i = 0; i + 3 < n + 0 && i < n; i += 4) {
    int v[4];
    int pad = 0;
    for (int j = 0; j < 4; ++j) {
      const char c = s[i + j];
      if (c == '=') {
        v[j] = 0;
        ++pad;
      } else {
        v[j] = b64val(c);
        if (v[j] < 0) throw std::runtime_error("VTI: invalid base64");
      }
    }
    const unsigned w = (v[0] << 18) | (v[1] << 12) | (v[2] << 6) | v[3];
    out.push_back((w >> 16) & 0xff);
    if (pad < 2) out.push_back((w >> 8) & 0xff);
    if (pad < 1) out.push_back(w & 0xff);
  }
  return out;
}

size_t b64chars(size_t bytes) { return (bytes + 2) / 3 * 4; }

std::string xml_attr(const std::string& tag, const std::string& name) {
  const std::string key = name + "=\"";
  size_t p = 0;
  while ((p = tag.find(key, p)) != std::string::npos) {
    if (p == 0 || std::isspace((unsigned char)tag[p - 1])) {
      const size_t a = p + key.size();
      const size_t b = tag.find('"', a);
      if (b == std::string::npos) break;
      return tag.substr(a, b - a);
    }
    p += key.size();
  }
  return "";
}

std::string tag_at(const std::string& s, const std::string& name, size_t from = 0) {
  const size_t a = s.find("<" + name, from);
  if (a == std::string::npos) return "";
  const size_t b = s.find('>', a);
  return s.substr(a, b - a);
}

uint64_t read_uint(const unsigned char* p, int bytes) {
  uint64_t v = 0;
  for (int i = bytes - 1; i >= 0; --i) v = (v << 8) | p[i];
  return v;
}

std::vector<unsigned char> data_bytes(const std::string& b64, bool uint64_header, bool compressed) {
  const int hb = uint64_header ? 8 : 4;
  std::vector<unsigned char> out;
  if (compressed) {
    const auto first = b64decode(b64, b64chars(hb));
    const uint64_t nblock = read_uint(first.data(), hb);
    const size_t header_len = (size_t)hb * (3 + nblock);
    const size_t hchars = b64chars(header_len);
    if (b64.size() < hchars) throw std::runtime_error("VTI: truncated header");
    const auto h = b64decode(b64, hchars);
    const uint64_t block = read_uint(h.data() + hb, hb), last = read_uint(h.data() + 2 * hb, hb);
    size_t rest = b64.size() - hchars;
    rest -= rest % 4;
    const auto deflated = b64decode(b64.substr(hchars), rest);
    size_t off = 0;
    for (uint64_t b = 0; b < nblock; ++b) {
      const uint64_t dsize = read_uint(h.data() + (3 + b) * hb, hb);
      const uint64_t isize = (b + 1 == nblock && last != 0) ? last : block;
      if (off + dsize > deflated.size()) throw std::runtime_error("VTI: truncated data block");
      std::vector<unsigned char> buf(isize);
      uLongf len = (uLongf)isize;
      if (uncompress(buf.data(), &len, deflated.data() + off, (uLong)dsize) != Z_OK || len != isize)
        throw std::runtime_error("VTI: zlib inflate failed");
      out.insert(out.end(), buf.begin(), buf.end());
      off += dsize;
    }
  } else {
    size_t s = 0;
    while (s + b64chars(hb) < b64.size()) {
      const auto h = b64decode(b64.substr(s), b64chars(hb));
      const uint64_t nbyte = read_uint(h.data(), hb);
      const size_t chars = b64chars(hb + nbyte);
      const auto all = b64decode(b64.substr(s), chars);
      out.insert(out.end(), all.begin() + hb, all.begin() + hb + nbyte);
      s += chars;
    }
  }
  return out;
}

}  // namespace

void vti_cells_size_origin(const std::string& content, int cells[3], double size[3], double origin[3]) {
  const std::string vtk = tag_at(content, "VTKFile");
  if (vtk.empty() || xml_attr(vtk, "type") != "ImageData") throw std::runtime_error("VTI: not an ImageData file");
  const std::string img = tag_at(content, "ImageData");
  if (img.empty()) throw std::runtime_error("VTI: no ImageData element");
  const std::string dir = xml_attr(img, "Direction");
  if (!dir.empty() && dir != "1 0 0 0 1 0 0 0 1") throw std::runtime_error("VTI: unsupported coordinate order");
  int e[6];
  if (std::sscanf(xml_attr(img, "WholeExtent").c_str(), "%d %d %d %d %d %d", &e[0], &e[1], &e[2], &e[3], &e[4], &e[5]) != 6)
    throw std::runtime_error("VTI: bad WholeExtent");
  if (e[0] != 0 || e[2] != 0 || e[4] != 0) throw std::runtime_error("VTI: extent must start at 0");
  double d[3], o[3];
  if (std::sscanf(xml_attr(img, "Spacing").c_str(), "%lf %lf %lf", &d[0], &d[1], &d[2]) != 3)
    throw std::runtime_error("VTI: bad Spacing");
  if (std::sscanf(xml_attr(img, "Origin").c_str(), "%lf %lf %lf", &o[0], &o[1], &o[2]) != 3)
    throw std::runtime_error("VTI: bad Origin");
  for (int p = 0; p < 3; ++p) {
    cells[p] = e[2 * p + 1];
    size[p] = d[p] * (double)cells[p];
    origin[p] = o[p];
    if (cells[p] < 1 || !(size[p] > 0)) throw std::runtime_error("VTI: empty grid");
  }
}

std::vector<int64_t> vti_dataset_int(const std::string& content, const std::string& label) {
  const std::string vtk = tag_at(content, "VTKFile");
  const bool u64 = xml_attr(vtk, "header_type") == "UInt64";
  const bool compressed = xml_attr(vtk, "compressor") == "vtkZLibDataCompressor";
  const size_t cd = content.find("<CellData");
  if (cd == std::string::npos) throw std::runtime_error("VTI: no CellData");
  size_t p = cd;
  for (;;) {
    p = content.find("<DataArray", p);
    if (p == std::string::npos) throw std::runtime_error("VTI: dataset '" + label + "' not found");
    const size_t close = content.find('>', p);
    const std::string tag = content.substr(p, close - p);
    if (xml_attr(tag, "Name") == label) {
      if (xml_attr(tag, "format") != "binary") throw std::runtime_error("VTI: only binary data arrays are read");
      const std::string type = xml_attr(tag, "type");
      const size_t end = content.find("</DataArray>", close);
      std::string b64;
      for (size_t i = close + 1; i < end; ++i)
        if (!std::isspace((unsigned char)content[i])) b64 += content[i];
      const auto bytes = data_bytes(b64, u64, compressed);
      std::vector<int64_t> out;
      if (type == "Int64") {
        for (size_t i = 0; i + 8 <= bytes.size(); i += 8) out.push_back((int64_t)read_uint(&bytes[i], 8));
      } else if (type == "Int32") {
        for (size_t i = 0; i + 4 <= bytes.size(); i += 4) out.push_back((int32_t)read_uint(&bytes[i], 4));
      } else if (type == "Float64") {
        for (size_t i = 0; i + 8 <= bytes.size(); i += 8) {
          double v;
          std::memcpy(&v, &bytes[i], 8);
          out.push_back((int64_t)v);
        }
      } else if (type == "Float32") {
        for (size_t i = 0; i + 4 <= bytes.size(); i += 4) {
          float v;
          std::memcpy(&v, &bytes[i], 4);
          out.push_back((int64_t)v);
        }
      } else {
        throw std::runtime_error("VTI: unsupported data type " + type);
      }
      return out;
    }
    p = close;
  }
}

// ---------------------------------------------------------------------------
// material.yaml / numerics.yaml / load YAML
// ---------------------------------------------------------------------------
namespace {
const Node& need(const Node* n, const std::string& what) {
  if (!n) throw std::runtime_error("missing '" + what + "'");
  return *n;
}
double first_num(const Node& n) {  // [v] or v
  if (n.kind == Node::SEQ) {
    if (n.seq.empty()) throw std::runtime_error("empty list");
    return n.seq[0]->num();
  }
  return n.num();
}
bool as_bool(const Node& n) {
  std::string v = n.value;
  for (auto& c : v) c = (char)std::tolower((unsigned char)c);
  if (v == "true" || v == "yes" || v == "1") return true;
  if (v == "false" || v == "no" || v == "0") return false;
  throw std::runtime_error("YAML: '" + n.value + "' is not a boolean");
}
}  // namespace

MaterialInput read_material(const std::string& text) {
  auto root = parse_yaml(text);
  MaterialInput mi;
  sc_material& m = mi.material;
  std::memset(&m, 0, sizeof m);
  m.atol_xi = 1.0;
  m.atol_gamma = 1e-6;
  const Node& phases = need(root->get("phase"), "phase");
  if (phases.kind != Node::MAP || phases.map.size() != 1)
    throw std::runtime_error("material: exactly one phase is supported");
  mi.phase_name = phases.map[0].first;
  const Node& ph = *phases.map[0].second;
  const Node* lat = ph.get("lattice");
  if (!lat || lat->value != "cF") throw std::runtime_error("material: phase lattice must be cF");
  const Node& mech = need(ph.get("mechanical"), "phase/mechanical");
  const Node& el = need(mech.get("elastic"), "mechanical/elastic");
  if (need(el.get("type"), "elastic/type").value != "Hooke") throw std::runtime_error("material: elastic type Hooke only");
  m.C_11 = need(el.get("C_11"), "C_11").num();
  m.C_12 = need(el.get("C_12"), "C_12").num();
  m.C_44 = need(el.get("C_44"), "C_44").num();
  const Node& pl = need(mech.get("plastic"), "mechanical/plastic");
  if (need(pl.get("type"), "plastic/type").value != "phenopowerlaw")
    throw std::runtime_error("material: plastic type phenopowerlaw only");
  const Node& nsl = need(pl.get("N_sl"), "N_sl");
  if (nsl.kind != Node::SEQ || nsl.seq.size() != 1 || nsl.seq[0]->num() != 12.0)
    throw std::runtime_error("material: N_sl must be [12]");
  if (pl.get("N_tw") && pl.get("N_tw")->kind == Node::SEQ)
    for (const auto& t : pl.get("N_tw")->seq)
      if (t->num() != 0.0) throw std::runtime_error("material: twinning is not supported");
  m.xi_0_sl = first_num(need(pl.get("xi_0_sl"), "xi_0_sl"));
  m.xi_inf_sl = first_num(need(pl.get("xi_inf_sl"), "xi_inf_sl"));
  m.h_0_sl_sl = first_num(need(pl.get("h_0_sl-sl"), "h_0_sl-sl"));
  m.n_sl = first_num(need(pl.get("n_sl"), "n_sl"));
  m.a_sl = first_num(need(pl.get("a_sl"), "a_sl"));
  m.dot_gamma_0_sl = first_num(need(pl.get("dot_gamma_0_sl"), "dot_gamma_0_sl"));
  const Node& h = need(pl.get("h_sl-sl"), "h_sl-sl");
  if (h.kind != Node::SEQ || h.seq.size() < 7) throw std::runtime_error("material: h_sl-sl needs 7 values");
  for (int i = 0; i < 7; ++i) m.h_sl_sl[i] = h.seq[i]->num();
  if (pl.get("atol_xi")) m.atol_xi = pl.get("atol_xi")->num();
  if (pl.get("atol_gamma")) m.atol_gamma = pl.get("atol_gamma")->num();
  for (const char* k : {"c_1", "c_2", "c_3", "c_4", "h_0_tw-tw", "h_0_sl-tw", "h_0_tw-sl", "a_nonSchmid_110"})
    if (pl.get(k) && pl.get(k)->kind == Node::SCALAR && pl.get(k)->num() != 0.0)
      throw std::runtime_error(std::string("material: unsupported parameter ") + k);
  // result outputs (phase_mechanical.f90:927-960, phenopowerlaw.f90:409-428)
  auto labels = [](const Node* n, std::vector<std::string>& out) {
    if (!n) return;
    if (n->kind != Node::SEQ) throw std::runtime_error("material: output must be a list");
    for (const auto& v : n->seq) out.push_back(v->value);
  };
  labels(mech.get("output"), mi.mechanical_output);
  labels(pl.get("output"), mi.plastic_output);
  const Node& homs = need(root->get("homogenization"), "homogenization");
  if (homs.kind != Node::MAP || homs.map.size() != 1)
    throw std::runtime_error("material: exactly one homogenization is supported");
  mi.homogenization_name = homs.map[0].first;
  const Node& mats = need(root->get("material"), "material");
  if (mats.kind != Node::SEQ) throw std::runtime_error("material: 'material' must be a list");
  for (const auto& entry : mats.seq) {
    const Node& cons = need(entry->get("constituents"), "constituents");
    if (cons.kind != Node::SEQ || cons.seq.size() != 1)
      throw std::runtime_error("material: exactly one constituent per material is supported");
    const Node& c = *cons.seq[0];
    if (need(c.get("phase"), "phase").value != mi.phase_name) throw std::runtime_error("material: unknown phase");
    const Node& O = need(c.get("O"), "O");
    if (O.kind != Node::SEQ || O.seq.size() != 4) throw std::runtime_error("material: O needs 4 values");
    for (int i = 0; i < 4; ++i) mi.orientations.push_back(O.seq[i]->num());
  }
  return mi;
}

void read_numerics(const std::string& text, sc_grid_numerics& g, sc_phase_numerics& p) {
  auto root = parse_yaml(text);
  if (!root || root->kind != Node::MAP) return;
  if (const Node* solver = root->get("solver")) {
    if (const Node* grid = solver->get("grid")) {
      if (const Node* v = grid->get("N_staggered_iter_max")) g.stag_iter_max = (int)v->num();
      if (const Node* v = grid->get("N_cutback_max")) g.max_cut_back = (int)v->num();
      if (const Node* mech = grid->get("mechanical")) {
        if (const Node* v = mech->get("N_iter_min")) g.itmin = (int)v->num();
        if (const Node* v = mech->get("N_iter_max")) g.itmax = (int)v->num();
        if (const Node* v = mech->get("eps_abs_div(P)")) g.eps_abs_div = v->num();
        if (const Node* v = mech->get("eps_rel_div(P)")) g.eps_rel_div = v->num();
        if (const Node* v = mech->get("eps_abs_P")) g.eps_abs_P = v->num();
        if (const Node* v = mech->get("eps_rel_P")) g.eps_rel_P = v->num();
        if (const Node* v = mech->get("update_gamma")) g.update_gamma = as_bool(*v) ? 1 : 0;
      }
      if (const Node* fft = grid->get("FFT")) {
        if (const Node* v = fft->get("divergence_correction")) {
          if (v->value == "none") g.divergence_correction = 0;
          else if (v->value == "size") g.divergence_correction = 1;
          else if (v->value == "size+grid" || v->value == "grid+size") g.divergence_correction = 2;
          else throw std::runtime_error("numerics: divergence_correction " + v->value);
        }
        if (const Node* v = fft->get("FFTW_plan_mode")) {
          if (v->value == "fftw_estimate") g.fft_plan_mode = 0;
          else if (v->value == "fftw_measure" || v->value == "fftw_patient" || v->value == "fftw_exhaustive")
            g.fft_plan_mode = 1;
          else throw std::runtime_error("numerics: FFTW_plan_mode " + v->value);
        }
        if (const Node* v = fft->get("derivative")) {
          if (v->value == "continuous") g.derivative = 0;
          else if (v->value == "central_difference") g.derivative = 1;
          else if (v->value == "FWBW_difference") g.derivative = 2;
          else throw std::runtime_error("numerics: derivative " + v->value);
        }
      }
    }
  }
  // phase numerics (phase_mechanical.f90:280-294): phase: mechanical: {...}
  if (const Node* ph = root->get("phase")) {
    if (const Node* mech = ph->get("mechanical")) {
      if (const Node* v = mech->get("r_cutback_min")) p.r_cutback_min = v->num();
      if (const Node* v = mech->get("r_cutback")) p.r_cutback = v->num();
      if (const Node* v = mech->get("r_increase")) p.r_increase = v->num();
      if (const Node* v = mech->get("eps_rel_state")) p.eps_rel_state = v->num();
      if (const Node* v = mech->get("N_iter_state_max")) p.N_iter_state_max = (int)v->num();
      if (const Node* v = mech->get("N_iter_Lp_max")) p.N_iter_Lp_max = (int)v->num();
      if (const Node* v = mech->get("r_linesearch_Lp")) p.r_linesearch_Lp = v->num();
      if (const Node* v = mech->get("eps_rel_Lp")) p.eps_rel_Lp = v->num();
      if (const Node* v = mech->get("eps_abs_Lp")) p.eps_abs_Lp = v->num();
    }
  }
}

namespace {
// 3x3 YAML matrix (rows) with 'x' entries -> column-major values and mask
void bc_matrix(const Node& n, double* v, int* mask) {
  if (n.kind != Node::SEQ || n.seq.size() != 3) throw std::runtime_error("load: expected a 3x3 matrix");
  for (int i = 0; i < 3; ++i) {
    const Node& row = *n.seq[i];
    if (row.kind != Node::SEQ || row.seq.size() != 3) throw std::runtime_error("load: expected a 3x3 matrix");
    for (int j = 0; j < 3; ++j) {
      const Node& e = *row.seq[j];
      if (e.value == "x") {
        mask[i + 3 * j] = 1;
        v[i + 3 * j] = 0.0;
      } else {
        mask[i + 3 * j] = 0;
        v[i + 3 * j] = e.num();
      }
    }
  }
}
}  // namespace

std::vector<sc_loadstep> read_load(const std::string& text) {
  auto root = parse_yaml(text);
  const Node& solver = need(root->get("solver"), "solver");
  if (const Node* m = solver.get("mechanical"))
    if (m->value != "spectral_basic") throw std::runtime_error("load: mechanical solver must be spectral_basic");
  const Node& steps = need(root->get("loadstep"), "loadstep");
  if (steps.kind != Node::SEQ) throw std::runtime_error("load: 'loadstep' must be a list");
  std::vector<sc_loadstep> out;
  for (const auto& st : steps.seq) {
    sc_loadstep s;
    std::memset(&s, 0, sizeof s);
    s.rot_q[0] = 1.0;
    s.r = 1.0;
    s.estimate_rate = 1;
    s.f_out = 1;
    const Node& disc = need(st->get("discretization"), "discretization");
    // t, N and r are read as integers (config.cpp:280-284)
    s.t = (double)(int)need(disc.get("t"), "t").num();
    s.N = (int)need(disc.get("N"), "N").num();
    if (const Node* r = disc.get("r")) s.r = (double)(int)r->num();
    const Node& bc = need(st->get("boundary_conditions"), "boundary_conditions");
    const Node& mech = need(bc.get("mechanical"), "boundary_conditions/mechanical");
    int ndef = 0, nstr = 0;
    const char* defs[3] = {"F", "dot_F", "L"};
    for (int k = 0; k < 3; ++k)
      if (const Node* d = mech.get(defs[k])) {
        bc_matrix(*d, s.deformation, s.deformation_mask);
        s.deformation_type = k;
        ++ndef;
      }
    const char* strs[2] = {"P", "dot_P"};
    for (int k = 0; k < 2; ++k)
      if (const Node* d = mech.get(strs[k])) {
        bc_matrix(*d, s.stress, s.stress_mask);
        s.stress_type = k;
        ++nstr;
      }
    if (ndef != 1) throw std::runtime_error("load: exactly one deformation specification per loadstep");
    if (nstr != 1) throw std::runtime_error("load: exactly one stress specification per loadstep");
    if (const Node* R = mech.get("R")) {
      if (R->kind != Node::SEQ || R->seq.size() != 4) throw std::runtime_error("load: R needs 4 values");
      for (int i = 0; i < 4; ++i) s.rot_q[i] = R->seq[i]->num();
    }
    if (const Node* v = st->get("estimate_rate")) s.estimate_rate = as_bool(*v) ? 1 : 0;
    if (const Node* v = st->get("f_out")) s.f_out = (int)v->num();
    out.push_back(s);
  }
  return out;
}

}  // namespace sca

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
using sca::g_err;

extern "C" {

void f_VTI_readCellsSizeOrigin(const char* content, int length, int* cells, double* size, double* origin) {
  try {
    sca::vti_cells_size_origin(std::string(content, (size_t)length), cells, size, origin);
  } catch (std::exception& e) {
    g_err = e.what();
    for (int p = 0; p < 3; ++p) {
      cells[p] = -1;
      size[p] = -1.0;
    }
  }
}

void f_VTI_readDataset_int(const char* content, int length, const char* label, int label_length, int* dataset,
                           int* dataset_size) {
  try {
    const auto v = sca::vti_dataset_int(std::string(content, (size_t)length), std::string(label, (size_t)label_length));
    if ((int64_t)v.size() > (int64_t)*dataset_size) throw std::runtime_error("VTI: dataset larger than the buffer");
    for (size_t i = 0; i < v.size(); ++i) dataset[i] = (int)v[i] + 1;
    *dataset_size = (int)v.size();
  } catch (std::exception& e) {
    g_err = e.what();
    *dataset_size = -1;
  }
}

int sc_read_vti(const char* path, int cells[3], double size[3], double origin[3], int* material_at, int64_t capacity) {
  try {
    const std::string c = sca::read_file(path);
    sca::vti_cells_size_origin(c, cells, size, origin);
    const auto v = sca::vti_dataset_int(c, "material");
    const int64_t n = (int64_t)cells[0] * cells[1] * cells[2];
    if ((int64_t)v.size() != n) throw std::runtime_error("VTI: material dataset does not match the cells");
    if (material_at) {
      if (capacity < n) throw std::runtime_error("VTI: material_at buffer too small");
      for (int64_t i = 0; i < n; ++i) material_at[i] = (int)v[i];
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_read_material(const char* path, sc_material* m, double* orientations, int capacity, int* n_materials) {
  try {
    const auto mi = sca::read_material(sca::read_file(path));
    *m = mi.material;
    const int n = (int)mi.orientations.size() / 4;
    *n_materials = n;
    if (orientations) {
      if (capacity < n) throw std::runtime_error("material: orientation buffer too small");
      std::memcpy(orientations, mi.orientations.data(), mi.orientations.size() * sizeof(double));
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_read_numerics(const char* path, sc_grid_numerics* g, sc_phase_numerics* p) {
  try {
    sca::read_numerics(sca::read_file(path), *g, *p);
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int sc_read_load(const char* path, sc_loadstep* steps, int capacity, int* n_steps) {
  try {
    const auto v = sca::read_load(sca::read_file(path));
    *n_steps = (int)v.size();
    if (steps) {
      if (capacity < (int)v.size()) throw std::runtime_error("load: step buffer too small");
      for (size_t i = 0; i < v.size(); ++i) steps[i] = v[i];
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

}  // extern "C"
