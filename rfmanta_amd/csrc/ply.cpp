// PLY mesh ingest, as Manta's readPlyFile builds a Mesh (Model/Readers/PlyReader.cc:136-269):
// vertex x, y, z of any scalar type read as double and stored as float, each vertex
// transformed by AffineTransform::multiply_point (Core/Geometry/AffineTransform.cc:357-366),
// faces from the "vertex_indices" list triangulated as a fan (first, previous, current;
// PlyReader.cc:84-131), then Mesh::removeDegenerateTriangles (Model/Groups/Mesh.cc:315-359).
// Formats: ascii, binary_little_endian, binary_big_endian. Host code only.

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "internal.h"
#include "mh.h"

using mh_internal::fail;

namespace {

enum class Fmt { kAscii, kLittle, kBig };

struct Prop {
  std::string name;
  bool list = false;
  int count_type = 0;  // list length type
  int type = 0;        // scalar type (or list item type)
};

struct Element {
  std::string name;
  uint64_t count = 0;
  std::vector<Prop> props;
};

// Scalar types: 1 int8, 2 uint8, 3 int16, 4 uint16, 5 int32, 6 uint32, 7 float32, 8 float64.
int type_of(const std::string& t) {
  if (t == "char" || t == "int8") return 1;
  if (t == "uchar" || t == "uint8") return 2;
  if (t == "short" || t == "int16") return 3;
  if (t == "ushort" || t == "uint16") return 4;
  if (t == "int" || t == "int32") return 5;
  if (t == "uint" || t == "uint32") return 6;
  if (t == "float" || t == "float32") return 7;
  if (t == "double" || t == "float64") return 8;
  return 0;
}
int size_of(int t) {
  static const int s[] = {0, 1, 1, 2, 2, 4, 4, 4, 8};
  return s[t];
}

class Reader {
 public:
  Reader(std::istream& in, Fmt f) : in_(in), fmt_(f) {}
  // One scalar as double (rply converts every value to double).
  bool value(int type, double& out) {
    if (fmt_ == Fmt::kAscii) {
      std::string tok;
      if (!(in_ >> tok)) return false;
      char* end = nullptr;
      out = std::strtod(tok.c_str(), &end);
      return end && *end == 0;
    }
    unsigned char b[8];
    const int n = size_of(type);
    if (!in_.read(reinterpret_cast<char*>(b), n)) return false;
    if (fmt_ == Fmt::kBig)
      for (int i = 0; i < n / 2; ++i) std::swap(b[i], b[n - 1 - i]);
    switch (type) {
      case 1: out = (double)(int8_t)b[0]; break;
      case 2: out = (double)b[0]; break;
      case 3: { int16_t v; std::memcpy(&v, b, 2); out = v; break; }
      case 4: { uint16_t v; std::memcpy(&v, b, 2); out = v; break; }
      case 5: { int32_t v; std::memcpy(&v, b, 4); out = v; break; }
      case 6: { uint32_t v; std::memcpy(&v, b, 4); out = v; break; }
      case 7: { float v; std::memcpy(&v, b, 4); out = v; break; }
      default: { double v; std::memcpy(&v, b, 8); out = v; break; }
    }
    return true;
  }

 private:
  std::istream& in_;
  Fmt fmt_;
};

struct Mesh {
  std::vector<float> v;     // 3 per vertex
  std::vector<uint32_t> i;  // 3 per triangle
};

// Cross(a - b, c - b) == (0, 0, 0) in float (Mesh.cc:327).
bool degenerate(const Mesh& m, size_t t) {
  const float* a = &m.v[3 * (size_t)m.i[t + 0]];
  const float* b = &m.v[3 * (size_t)m.i[t + 1]];
  const float* c = &m.v[3 * (size_t)m.i[t + 2]];
  const float u[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  const float w[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
  const float x = u[1] * w[2] - u[2] * w[1];
  const float y = u[2] * w[0] - u[0] * w[2];
  const float z = u[0] * w[1] - u[1] * w[0];
  return x == 0.f && y == 0.f && z == 0.f;
}

int read_ply(const char* path, const float* xf, Mesh& m) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return fail(MH_E_IO, "cannot open %s", path);
  std::string line;
  if (!std::getline(in, line) || line.substr(0, 3) != "ply")
    return fail(MH_E_INVALID, "%s: not a PLY file", path);
  Fmt fmt = Fmt::kAscii;
  bool have_fmt = false;
  std::vector<Element> els;
  for (;;) {
    if (!std::getline(in, line)) return fail(MH_E_INVALID, "%s: header ends early", path);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ls(line);
    std::string kw;
    ls >> kw;
    if (kw == "end_header") break;
    if (kw == "format") {
      std::string f;
      ls >> f;
      if (f == "ascii") fmt = Fmt::kAscii;
      else if (f == "binary_little_endian") fmt = Fmt::kLittle;
      else if (f == "binary_big_endian") fmt = Fmt::kBig;
      else return fail(MH_E_INVALID, "%s: unknown format %s", path, f.c_str());
      have_fmt = true;
    } else if (kw == "element") {
      Element e;
      ls >> e.name >> e.count;
      if (!ls) return fail(MH_E_INVALID, "%s: bad element line", path);
      els.push_back(e);
    } else if (kw == "property") {
      if (els.empty()) return fail(MH_E_INVALID, "%s: property before element", path);
      Prop p;
      std::string t;
      ls >> t;
      if (t == "list") {
        std::string ct, it;
        ls >> ct >> it >> p.name;
        p.list = true;
        p.count_type = type_of(ct);
        p.type = type_of(it);
        if (!p.count_type || !p.type) return fail(MH_E_INVALID, "%s: bad list type", path);
      } else {
        p.type = type_of(t);
        ls >> p.name;
        if (!p.type) return fail(MH_E_INVALID, "%s: bad type %s", path, t.c_str());
      }
      els.back().props.push_back(p);
    }  // comment, obj_info: ignored
  }
  if (!have_fmt) return fail(MH_E_INVALID, "%s: no format line", path);
  Reader rd(in, fmt);
  // AffineTransform::multiply_point with the reference's accumulation order.
  float mat[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}};
  if (xf) std::memcpy(mat, xf, sizeof(mat));
  uint64_t faces_seen = 0;
  bool have_vertex = false;
  for (const Element& e : els) {
    const bool is_vertex = e.name == "vertex", is_face = e.name == "face";
    int px = -1, py = -1, pz = -1, pf = -1;
    for (size_t k = 0; k < e.props.size(); ++k) {
      const Prop& p = e.props[k];
      if (is_vertex && !p.list) {
        if (p.name == "x") px = (int)k;
        if (p.name == "y") py = (int)k;
        if (p.name == "z") pz = (int)k;
      }
      if (is_face && p.list && p.name == "vertex_indices") pf = (int)k;
    }
    if (is_vertex) have_vertex = true;
    for (uint64_t r = 0; r < e.count; ++r) {
      float pt[3] = {0.f, 0.f, 0.f};
      for (size_t k = 0; k < e.props.size(); ++k) {
        const Prop& p = e.props[k];
        double val;
        if (!p.list) {
          if (!rd.value(p.type, val)) return fail(MH_E_INVALID, "%s: data ends early", path);
          if ((int)k == px) pt[0] = (float)val;
          if ((int)k == py) pt[1] = (float)val;
          if ((int)k == pz) pt[2] = (float)val;
          continue;
        }
        double len;
        if (!rd.value(p.count_type, len) || len < 0)
          return fail(MH_E_INVALID, "%s: bad list length", path);
        const uint64_t n = (uint64_t)len;
        uint32_t first = 0, prev = 0;
        for (uint64_t j = 0; j < n; ++j) {
          if (!rd.value(p.type, val)) return fail(MH_E_INVALID, "%s: data ends early", path);
          if ((int)k != pf) continue;
          if (val < 0 || val >= 4294967296.0)
            return fail(MH_E_INVALID, "%s: vertex index %g", path, val);
          const uint32_t cur = (uint32_t)val;
          if (j == 0) {
            first = cur;
          } else if (j == 1) {
            prev = cur;
          } else {  // fan: (first, previous, current)
            m.i.push_back(first);
            m.i.push_back(prev);
            m.i.push_back(cur);
            prev = cur;
          }
        }
        if ((int)k == pf) ++faces_seen;
      }
      if (is_vertex) {
        float out[3];
        for (int a = 0; a < 3; ++a) {
          float t = mat[a][0] * pt[0];
          t += mat[a][1] * pt[1];
          t += mat[a][2] * pt[2];
          t += mat[a][3];
          out[a] = t;
        }
        m.v.insert(m.v.end(), out, out + 3);
      }
    }
  }
  if (!have_vertex) return fail(MH_E_INVALID, "%s: no vertex element", path);
  // The reference adds a vertex when the next one starts and the last one when the first
  // face is read, so without faces the last vertex is never added.
  if (faces_seen == 0 && !m.v.empty()) m.v.resize(m.v.size() - 3);
  const size_t nv = m.v.size() / 3;
  for (uint32_t x : m.i)
    if (x >= nv) return fail(MH_E_INVALID, "%s: vertex index %u out of range", path, x);
  for (size_t t = 0; t < m.i.size(); t += 3) {  // removeDegenerateTriangles
    if (!degenerate(m, t)) continue;
    const size_t last = m.i.size() - 3;
    m.i[t + 0] = m.i[last + 0];
    m.i[t + 1] = m.i[last + 1];
    m.i[t + 2] = m.i[last + 2];
    m.i.resize(last);
    t -= 3;  // the triangle moved here is checked next
  }
  return MH_OK;
}

}  // namespace

int mh_ply_read(const char* path, const float* transform, float* verts, uint32_t verts_cap,
                uint32_t* idx, uint32_t tris_cap, uint32_t* n_verts, uint32_t* n_tris) {
  if (!path || !n_verts || !n_tris) return fail(MH_E_INVALID, "null argument");
  Mesh m;
  int rc;
  try {
    rc = read_ply(path, transform, m);
  } catch (const std::bad_alloc&) {
    return fail(MH_E_NOMEM, "host memory for %s", path);
  }
  if (rc) return rc;
  const uint64_t nv = m.v.size() / 3, nt = m.i.size() / 3;
  if (nv > 0xFFFFFFFFull || nt > 0xFFFFFFFFull) return fail(MH_E_UNSUPPORTED, "mesh too large");
  *n_verts = (uint32_t)nv;
  *n_tris = (uint32_t)nt;
  if (!verts && !idx) return MH_OK;  // size query
  if (!verts || !idx || verts_cap < nv || tris_cap < nt)
    return fail(MH_E_INVALID, "arrays too small: %llu vertices, %llu triangles needed",
                (unsigned long long)nv, (unsigned long long)nt);
  std::memcpy(verts, m.v.data(), m.v.size() * sizeof(float));
  std::memcpy(idx, m.i.data(), m.i.size() * sizeof(uint32_t));
  return MH_OK;
}
