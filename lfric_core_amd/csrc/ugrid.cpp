// UGRID 2D mesh files (the reference's infrastructure/source/io/ugrid_2d_mod.F90 reads them
// through netCDF): the face-node and face-face connectivity and the panel count of a mesh
// topology variable, read from a netCDF classic (CDF-1) or 64-bit-offset (CDF-2) file by a small
// parser of that published format (header of dimensions, attributes and variables, then
// big-endian data).  Host code, run once at set-up.
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "internal.h"

namespace lfric {

namespace {

enum : uint32_t { kNcDimension = 10, kNcVariable = 11, kNcAttribute = 12 };
enum : int { kNcByte = 1, kNcChar = 2, kNcShort = 3, kNcInt = 4, kNcFloat = 5, kNcDouble = 6 };

struct NcAttr {
  int type = 0;
  std::string text;              // NC_CHAR
  std::vector<double> values;    // numeric types
};

struct NcVar {
  std::vector<int> dims;
  std::map<std::string, NcAttr> attrs;
  int type = 0;
  uint64_t begin = 0;
};

struct NcFile {
  std::vector<unsigned char> bytes;
  std::vector<std::pair<std::string, uint64_t>> dims;
  std::map<std::string, NcVar> vars;
  int version = 0;
};

int type_size(int t) {
  switch (t) {
    case kNcByte:
    case kNcChar:
      return 1;
    case kNcShort:
      return 2;
    case kNcInt:
    case kNcFloat:
      return 4;
    case kNcDouble:
      return 8;
  }
  return 0;
}

class Reader {
 public:
  explicit Reader(const std::vector<unsigned char> &b) : b_(b) {}
  bool ok() const { return ok_; }
  uint32_t u32() {
    if (!need(4)) return 0;
    const uint32_t v = (uint32_t(b_[p_]) << 24) | (uint32_t(b_[p_ + 1]) << 16) | (uint32_t(b_[p_ + 2]) << 8) | b_[p_ + 3];
    p_ += 4;
    return v;
  }
  uint64_t u64() {
    const uint64_t hi = u32();
    return (hi << 32) | u32();
  }
  std::string name() {
    const uint32_t n = u32();
    if (!need(n)) return {};
    std::string s(reinterpret_cast<const char *>(&b_[p_]), n);
    p_ += pad4(n);
    return s;
  }
  void skip(uint64_t n) {
    if (need(n)) p_ += n;
  }
  const unsigned char *here() const { return b_.data() + p_; }
  bool need(uint64_t n) {
    if (!ok_ || p_ + n > b_.size()) ok_ = false;
    return ok_;
  }
  static uint64_t pad4(uint64_t n) { return (n + 3) & ~uint64_t(3); }

 private:
  const std::vector<unsigned char> &b_;
  uint64_t p_ = 0;
  bool ok_ = true;
};

double read_be(const unsigned char *q, int type) {
  uint64_t u = 0;
  const int n = type_size(type);
  for (int i = 0; i < n; ++i) u = (u << 8) | q[i];
  switch (type) {
    case kNcByte:
      return static_cast<int8_t>(u);
    case kNcShort:
      return static_cast<int16_t>(u);
    case kNcInt:
      return static_cast<int32_t>(u);
    case kNcFloat: {
      const uint32_t w = static_cast<uint32_t>(u);
      float f;
      std::memcpy(&f, &w, 4);
      return f;
    }
    case kNcDouble: {
      double d;
      std::memcpy(&d, &u, 8);
      return d;
    }
  }
  return 0.0;
}

bool read_attrs(Reader &r, std::map<std::string, NcAttr> &out) {
  const uint32_t tag = r.u32(), n = r.u32();
  if (tag == 0 && n == 0) return r.ok();
  if (tag != kNcAttribute) return false;
  for (uint32_t i = 0; i < n && r.ok(); ++i) {
    NcAttr a;
    const std::string nm = r.name();
    a.type = static_cast<int>(r.u32());
    const uint32_t ne = r.u32();
    const int ts = type_size(a.type);
    if (ts == 0) return false;
    const uint64_t nb = uint64_t(ne) * ts;
    if (!r.need(nb)) return false;
    if (a.type == kNcChar) {
      a.text.assign(reinterpret_cast<const char *>(r.here()), ne);
      while (!a.text.empty() && a.text.back() == '\0') a.text.pop_back();
    } else {
      for (uint32_t k = 0; k < ne; ++k) a.values.push_back(read_be(r.here() + uint64_t(k) * ts, a.type));
    }
    r.skip(Reader::pad4(nb));
    out[nm] = a;
  }
  return r.ok();
}

int parse(const std::string &path, NcFile &f) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return fail(LFRIC_E_ARG, "UGRID: cannot open " + path);
  f.bytes.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
  Reader r(f.bytes);
  if (f.bytes.size() < 4 || std::memcmp(f.bytes.data(), "CDF", 3) != 0)
    return fail(LFRIC_E_UNSUPPORTED, "UGRID: not a netCDF classic file (HDF5-based netCDF-4 is not read): " + path);
  f.version = f.bytes[3];
  if (f.version != 1 && f.version != 2)
    return fail(LFRIC_E_UNSUPPORTED, "UGRID: netCDF format version " + std::to_string(f.version) + " is not read");
  r.skip(4);
  r.u32();  // numrecs
  {
    const uint32_t tag = r.u32(), n = r.u32();
    if (!(tag == 0 && n == 0)) {
      if (tag != kNcDimension) return fail(LFRIC_E_ARG, "UGRID: bad dimension list in " + path);
      for (uint32_t i = 0; i < n && r.ok(); ++i) {
        const std::string nm = r.name();
        f.dims.emplace_back(nm, r.u32());
      }
    }
  }
  std::map<std::string, NcAttr> gatts;
  if (!read_attrs(r, gatts)) return fail(LFRIC_E_ARG, "UGRID: bad global attributes in " + path);
  const uint32_t tag = r.u32(), n = r.u32();
  if (!(tag == 0 && n == 0)) {
    if (tag != kNcVariable) return fail(LFRIC_E_ARG, "UGRID: bad variable list in " + path);
    for (uint32_t i = 0; i < n && r.ok(); ++i) {
      NcVar v;
      const std::string nm = r.name();
      const uint32_t nd = r.u32();
      if (nd > 64) return fail(LFRIC_E_ARG, "UGRID: bad variable " + nm + " in " + path);
      for (uint32_t k = 0; k < nd; ++k) v.dims.push_back(static_cast<int>(r.u32()));
      if (!read_attrs(r, v.attrs)) return fail(LFRIC_E_ARG, "UGRID: bad attributes of " + nm);
      v.type = static_cast<int>(r.u32());
      r.u32();  // vsize
      v.begin = f.version == 1 ? r.u32() : r.u64();
      f.vars[nm] = v;
    }
  }
  if (!r.ok()) return fail(LFRIC_E_ARG, "UGRID: truncated header in " + path);
  return LFRIC_OK;
}

// An integer connectivity variable (nface, 4) in C order, made 1-based; fill values and
// negative entries (no neighbour) become 0.
int read_connectivity(const NcFile &f, const std::string &name, std::vector<int> &out, int &rows) {
  auto it = f.vars.find(name);
  LF_CHECK(it != f.vars.end(), LFRIC_E_ARG, "UGRID: variable " + name + " not found");
  const NcVar &v = it->second;
  LF_CHECK(v.type == kNcInt && v.dims.size() == 2, LFRIC_E_UNSUPPORTED, "UGRID: " + name + " is not an int (n, 4) array");
  for (int d : v.dims)
    LF_CHECK(d >= 0 && static_cast<size_t>(d) < f.dims.size(), LFRIC_E_ARG, "UGRID: " + name + " names no dimension");
  const uint64_t n0 = f.dims[v.dims[0]].second, n1 = f.dims[v.dims[1]].second;
  LF_CHECK(n1 == 4 && n0 > 0, LFRIC_E_UNSUPPORTED, "UGRID: " + name + " is not (n, 4) (quadrilateral faces)");
  LF_CHECK(v.begin + n0 * n1 * 4 <= f.bytes.size(), LFRIC_E_ARG, "UGRID: " + name + " runs past the end of the file");
  int start = 0;
  if (auto a = v.attrs.find("start_index"); a != v.attrs.end() && !a->second.values.empty())
    start = static_cast<int>(a->second.values[0]);
  bool has_fill = false;
  int fill = 0;
  if (auto a = v.attrs.find("_FillValue"); a != v.attrs.end() && !a->second.values.empty()) {
    has_fill = true;
    fill = static_cast<int>(a->second.values[0]);
  }
  out.resize(n0 * n1);
  for (uint64_t i = 0; i < n0 * n1; ++i) {
    const int x = static_cast<int>(read_be(&f.bytes[v.begin + 4 * i], kNcInt));
    out[i] = (has_fill && x == fill) || x < start ? 0 : x - start + 1;
  }
  rows = static_cast<int>(n0);
  return LFRIC_OK;
}

}  // namespace

// The mesh topology variable (cf_role "mesh_topology"; the one named mesh_name when given),
// its face_node_connectivity and face_face_connectivity ([face][4], 1-based, 0 = no neighbour)
// and its npanels attribute (0 when absent).
int read_ugrid_mesh(const std::string &path, const std::string &mesh_name, std::vector<int> &face_nodes,
                    std::vector<int> &face_links, int &nfaces, int &npanels) {
  NcFile f;
  if (int rc = parse(path, f)) return rc;
  const NcVar *mesh = nullptr;
  std::string found;
  for (const auto &kv : f.vars) {
    auto a = kv.second.attrs.find("cf_role");
    if (a == kv.second.attrs.end() || a->second.text != "mesh_topology") continue;
    if (!mesh_name.empty() && kv.first != mesh_name) continue;
    LF_CHECK(mesh == nullptr, LFRIC_E_ARG, "UGRID: several meshes in " + path + "; name one");
    mesh = &kv.second;
    found = kv.first;
  }
  LF_CHECK(mesh, LFRIC_E_ARG, "UGRID: no mesh topology " + (mesh_name.empty() ? std::string() : mesh_name + " ") +
                                  "in " + path);
  auto att = [&](const char *k) -> std::string {
    auto a = mesh->attrs.find(k);
    return a == mesh->attrs.end() ? std::string() : a->second.text;
  };
  const std::string fn = att("face_node_connectivity"), ff = att("face_face_connectivity");
  LF_CHECK(!fn.empty() && !ff.empty(), LFRIC_E_UNSUPPORTED,
           "UGRID: mesh " + found + " lacks face_node or face_face connectivity");
  int n1 = 0, n2 = 0;
  if (int rc = read_connectivity(f, fn, face_nodes, n1)) return rc;
  if (int rc = read_connectivity(f, ff, face_links, n2)) return rc;
  LF_CHECK(n1 == n2, LFRIC_E_ARG, "UGRID: face counts of the connectivities differ");
  nfaces = n1;
  npanels = 0;
  if (auto a = mesh->attrs.find("npanels"); a != mesh->attrs.end() && !a->second.values.empty())
    npanels = static_cast<int>(a->second.values[0]);
  return LFRIC_OK;
}

}  // namespace lfric
