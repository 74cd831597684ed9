// OBJ mesh ingest, as Manta's ObjGroup builds a Mesh from a Wavefront file read by glm
// (Model/Groups/ObjGroup.cc:184-316; Model/Readers/glm/glm.cc:182-243 groups and materials,
// 270-390 material names, 575-745 first pass, 756-960 second pass, 1599-1670 glmReadOBJ).
//
// What the Mesh receives:
//   vertices, vertex normals (vn) and texture coordinates (vt as (u, v, 1)) in file order;
//   triangles from each face as a fan (first, previous, current), grouped by glm group and
//   emitted group by group in glm's list order, which is the reverse of creation order,
//   file order inside a group;
//   a group is named "<group>_MAT_<material>" and created by g/o and usemtl statements,
//   faces before the first such statement go to the group at the head of the list (glm's
//   second pass starts there), the initial group is "No Group";
//   per corner, the normal index (or, when the file has normals and the corner none, a new
//   normal Cross(v1 - v0, v2 - v0) appended to the normal list) and the texture index (or
//   MH_NO_TEXTURE_INDEX when the file has texture coordinates and the corner none);
//   per triangle the material slot: 1 + the group's glm material when that is below glm's
//   material count, else 0 (the caller's default material); glm material 0 is
//   "NO_ASSIGNED_MATERIAL", the newmtl entries of the last mtllib follow in file order;
//   then Mesh::removeDegenerateTriangles (Model/Groups/Mesh.cc:315-359) with the normal and
//   texture indices and the material slots moved alike.
// Negative vertex indices count back from the last vertex read; negative normal and texture
// indices count as absent, as glm leaves them. Material parameters and textures are not
// read (the materials belong to the scene). Faces whose corners mix forms (v, v/t, v//n,
// v/t/n), faces with fewer than three corners and indices out of range are rejected.
// Host code only.

#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <new>
#include <string>
#include <vector>

#include "internal.h"
#include "mh.h"

using mh_internal::fail;

namespace {

// A cursor over the file text with fscanf's token rule (runs of non-space) and fgets's
// rest-of-line rule.
class Text {
 public:
  explicit Text(std::string s) : s_(std::move(s)) {}
  bool token(std::string& out) {
    skip_space();
    if (p_ >= s_.size()) return false;
    const size_t b = p_;
    while (p_ < s_.size() && !space(s_[p_])) ++p_;
    out.assign(s_, b, p_ - b);
    return true;
  }
  // The next token without consuming it ("" at the end).
  std::string peek() {
    const size_t keep = p_;
    std::string t;
    token(t);
    p_ = keep;
    return t;
  }
  std::string rest_of_line() {
    const size_t b = p_;
    while (p_ < s_.size() && s_[p_] != '\n') ++p_;
    std::string r(s_, b, p_ - b);
    if (p_ < s_.size()) ++p_;
    return r;
  }

 private:
  static bool space(char c) { return std::isspace((unsigned char)c) != 0; }
  void skip_space() {
    while (p_ < s_.size() && space(s_[p_])) ++p_;
  }
  std::string s_;
  size_t p_ = 0;
};

std::string first_word(const std::string& line) {
  size_t b = 0;
  while (b < line.size() && std::isspace((unsigned char)line[b])) ++b;
  size_t e = b;
  while (e < line.size() && !std::isspace((unsigned char)line[e])) ++e;
  return line.substr(b, e - b);
}

// Up to n floats from the start of a line (fscanf "%f"); missing ones stay 0.
void floats(const std::string& line, float* out, int n) {
  const char* c = line.c_str();
  for (int k = 0; k < n; ++k) {
    char* end = nullptr;
    const float v = std::strtof(c, &end);
    if (end == c) return;
    out[k] = v;
    c = end;
  }
}

bool read_file(const std::string& path, std::string& out) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  out.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
  return true;
}

// Corner forms, as glm tells them apart by the face's first corner.
enum Form { kV, kVT, kVTN, kVN };

// Parses one corner token in the given form: 1 when it matches, 0 when the token does not
// start with an integer (the face ends before it), -1 when it starts like one but does not
// match the form.
int corner(const std::string& tok, Form f, long& v, long& t, long& n) {
  const char* c = tok.c_str();
  char* end;
  auto integer = [&](long& out) {
    if (!(*c == '-' || *c == '+' || (*c >= '0' && *c <= '9'))) return false;
    out = std::strtol(c, &end, 10);
    if (end == c) return false;
    c = end;
    return true;
  };
  auto lit = [&](const char* s) {
    const size_t k = std::strlen(s);
    if (std::strncmp(c, s, k)) return false;
    c += k;
    return true;
  };
  t = n = 0;
  if (!integer(v)) return 0;
  bool ok = true;
  switch (f) {
    case kV: break;
    case kVT: ok = lit("/") && integer(t); break;
    case kVTN: ok = lit("/") && integer(t) && lit("/") && integer(n); break;
    case kVN: ok = lit("//") && integer(n); break;
  }
  return ok && *c == 0 ? 1 : -1;
}

Form form_of(const std::string& tok) {
  if (tok.find("//") != std::string::npos) return kVN;
  long a, b, c;
  if (std::sscanf(tok.c_str(), "%ld/%ld/%ld", &a, &b, &c) == 3) return kVTN;
  if (std::sscanf(tok.c_str(), "%ld/%ld", &a, &b) == 2) return kVT;
  return kV;
}

struct Corner {
  long v, t, n;  // as glm stores them: 1-based, 0 = absent
};

struct Group {
  std::string name;
  // Material name at the last statement that selected the group; glm looks it up in the
  // final material list (not found, as "No Material" never is: index 0).
  std::string material;
  std::vector<size_t> tris;  // indices into the triangle list, in file order
};

struct Obj {
  std::vector<float> v, vn, vt;  // vt as (u, v) pairs
  std::vector<Corner> corners;   // 3 per triangle
  std::vector<Group> groups;     // creation order
  std::vector<std::string> materials;  // glm's: [0] NO_ASSIGNED_MATERIAL, then newmtl
  bool have_mtllib = false;
};

int find_group(const Obj& o, const std::string& name) {
  for (size_t k = 0; k < o.groups.size(); ++k)
    if (o.groups[k].name == name) return (int)k;
  return -1;
}

int add_group(Obj& o, const std::string& name) {
  const int k = find_group(o, name);
  if (k >= 0) return k;
  Group g;
  g.name = name;
  o.groups.push_back(g);
  return (int)o.groups.size() - 1;
}

// Material names of a library, in newmtl order after glm's NO_ASSIGNED_MATERIAL.
int read_mtl(const std::string& path, Obj& o) {
  std::string text;
  if (!read_file(path, text)) return fail(MH_E_IO, "cannot open material library %s", path.c_str());
  Text tx(std::move(text));
  o.materials.assign(1, "NO_ASSIGNED_MATERIAL");
  std::string tok;
  while (tx.token(tok)) {
    const std::string line = tx.rest_of_line();
    if (tok[0] == 'n') o.materials.push_back(first_word(line));
  }
  o.have_mtllib = true;
  return MH_OK;
}

int parse(const char* path, Obj& o) {
  std::string text;
  if (!read_file(path, text)) return fail(MH_E_IO, "cannot open %s", path);
  const std::string p(path);
  const size_t slash = p.rfind('/');
  const std::string dir = slash == std::string::npos ? "" : p.substr(0, slash + 1);
  Text tx(std::move(text));
  std::string grp = "No Group", mtl = "No Material";
  add_group(o, grp);
  int cur = -1;  // -1: faces before the first group statement
  std::vector<size_t> early;
  std::string tok;
  while (tx.token(tok)) {
    const char c0 = tok[0];
    if (c0 == 'v' && tok.size() == 1) {
      float x[3] = {0.f, 0.f, 0.f};
      floats(tx.rest_of_line(), x, 3);
      o.v.insert(o.v.end(), x, x + 3);
    } else if (c0 == 'v' && tok[1] == 'n') {
      float x[3] = {0.f, 0.f, 0.f};
      floats(tx.rest_of_line(), x, 3);
      o.vn.insert(o.vn.end(), x, x + 3);
    } else if (c0 == 'v' && tok[1] == 't') {
      float x[2] = {0.f, 0.f};
      floats(tx.rest_of_line(), x, 2);
      o.vt.insert(o.vt.end(), x, x + 2);
    } else if (c0 == 'v') {
      // glm's first pass reports an unknown v token and goes on with the next token.
    } else if (c0 == 'm') {
      const std::string name = first_word(tx.rest_of_line());
      if (const int rc = read_mtl(dir + name, o)) return rc;
    } else if (c0 == 'u') {
      mtl = first_word(tx.rest_of_line());
      cur = add_group(o, grp + "_MAT_" + mtl);
      o.groups[cur].material = mtl;
    } else if (c0 == 'o' || c0 == 'g') {
      grp = first_word(tx.rest_of_line());
      cur = add_group(o, grp + "_MAT_" + mtl);
      o.groups[cur].material = mtl;
    } else if (c0 == 'f') {
      std::string first;
      if (!tx.token(first)) return fail(MH_E_INVALID, "%s: face without corners", path);
      const Form f = form_of(first);
      std::vector<Corner> poly;
      Corner k;
      if (corner(first, f, k.v, k.t, k.n) != 1)
        return fail(MH_E_INVALID, "%s: bad face corner %s", path, first.c_str());
      poly.push_back(k);
      for (;;) {
        const std::string nx = tx.peek();
        if (nx.empty()) break;
        const int r = corner(nx, f, k.v, k.t, k.n);
        if (r == 0) break;
        if (r < 0) return fail(MH_E_INVALID, "%s: face corner %s mixes forms", path, nx.c_str());
        tx.token(tok);
        poly.push_back(k);
      }
      if (poly.size() < 3) return fail(MH_E_INVALID, "%s: face with %zu corners", path, poly.size());
      const long nv = (long)(o.v.size() / 3) + 1;  // glm's running vertex count, 1-based
      for (Corner& q : poly)
        if (q.v < 0) q.v += nv;
      for (size_t j = 2; j < poly.size(); ++j) {
        const size_t id = o.corners.size() / 3;
        o.corners.push_back(poly[0]);
        o.corners.push_back(poly[j - 1]);
        o.corners.push_back(poly[j]);
        (cur < 0 ? early : o.groups[cur].tris).push_back(id);
      }
    } else {
      tx.rest_of_line();  // '#' comments and every other statement
    }
  }
  // glm's second pass starts on the head of its list, the group created last.
  Group& head = o.groups.back();
  head.tris.insert(head.tris.begin(), early.begin(), early.end());
  return MH_OK;
}

// The mesh handed out, with the arrays it points into.
struct Result : mh_obj_mesh {
  Result() : mh_obj_mesh() {}
  std::vector<float> v, vn, vt;
  std::vector<uint32_t> idx, nidx, tidx, mat;
  std::vector<std::string> names;
  std::vector<const char*> name_ptrs;
};

bool degenerate(const std::vector<float>& v, const uint32_t* t) {
  const float* a = &v[3 * (size_t)t[0]];
  const float* b = &v[3 * (size_t)t[1]];
  const float* c = &v[3 * (size_t)t[2]];
  const float u[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  const float w[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
  const float x = u[1] * w[2] - u[2] * w[1];
  const float y = u[2] * w[0] - u[0] * w[2];
  const float z = u[0] * w[1] - u[1] * w[0];
  return x == 0.f && y == 0.f && z == 0.f;
}

int assemble(const char* path, const Obj& o, Result& r) {
  const size_t nv = o.v.size() / 3, nn = o.vn.size() / 3, nt = o.vt.size() / 2;
  r.v = o.v;
  r.vn = o.vn;
  for (size_t k = 0; k < nt; ++k) {
    r.vt.push_back(o.vt[2 * k]);
    r.vt.push_back(o.vt[2 * k + 1]);
    r.vt.push_back(1.f);
  }
  const bool has_tex = nt > 0;
  // glm's material index of a group: the named material's position, 0 when not found.
  auto material_of = [&](const Group& g) -> uint32_t {
    for (size_t k = 0; k < o.materials.size(); ++k)
      if (o.materials[k] == g.material) return (uint32_t)k;
    return 0;
  };
  for (size_t gi = o.groups.size(); gi-- > 0;) {
    const Group& g = o.groups[gi];
    const uint32_t gm = material_of(g);
    const uint32_t slot = gm < o.materials.size() ? gm + 1 : 0;
    for (size_t id : g.tris) {
      const Corner* c = &o.corners[3 * id];
      for (int k = 0; k < 3; ++k)
        if (c[k].v < 1 || (size_t)c[k].v > nv)
          return fail(MH_E_INVALID, "%s: vertex index %ld out of range", path, c[k].v);
      for (int k = 0; k < 3; ++k) {
        r.idx.push_back((uint32_t)(c[k].v - 1));
        if (c[k].n > 0) {
          if ((size_t)c[k].n > nn)
            return fail(MH_E_INVALID, "%s: normal index %ld out of range", path, c[k].n);
          r.nidx.push_back((uint32_t)(c[k].n - 1));
        } else if (!r.vn.empty()) {
          const float* p0 = &o.v[3 * (size_t)(c[0].v - 1)];
          const float* p1 = &o.v[3 * (size_t)(c[1].v - 1)];
          const float* p2 = &o.v[3 * (size_t)(c[2].v - 1)];
          const float a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
          const float b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
          r.nidx.push_back((uint32_t)(r.vn.size() / 3));
          r.vn.push_back(a[1] * b[2] - a[2] * b[1]);
          r.vn.push_back(a[2] * b[0] - a[0] * b[2]);
          r.vn.push_back(a[0] * b[1] - a[1] * b[0]);
        }
        if (c[k].t > 0) {
          if ((size_t)c[k].t > nt)
            return fail(MH_E_INVALID, "%s: texture index %ld out of range", path, c[k].t);
          r.tidx.push_back((uint32_t)(c[k].t - 1));
        } else if (has_tex) {
          r.tidx.push_back(MH_NO_TEXTURE_INDEX);
        }
      }
      r.mat.push_back(slot);
    }
  }
  // Mesh::removeDegenerateTriangles
  for (size_t i = 0; i < r.idx.size(); i += 3) {
    if (!degenerate(r.v, &r.idx[i])) continue;
    const size_t last = r.idx.size() - 3;
    for (int k = 0; k < 3; ++k) r.idx[i + k] = r.idx[last + k];
    r.idx.resize(last);
    if (!r.nidx.empty()) {
      for (int k = 0; k < 3; ++k) r.nidx[i + k] = r.nidx[last + k];
      r.nidx.resize(last);
    }
    if (!r.tidx.empty()) {
      for (int k = 0; k < 3; ++k) r.tidx[i + k] = r.tidx[last + k];
      r.tidx.resize(last);
    }
    r.mat[i / 3] = r.mat.back();
    r.mat.pop_back();
    i -= 3;
  }
  r.names.push_back("");
  if (o.have_mtllib) r.names.insert(r.names.end(), o.materials.begin(), o.materials.end());
  for (const std::string& s : r.names) r.name_ptrs.push_back(s.c_str());
  if (r.v.size() / 3 > 0xFFFFFFFFull || r.idx.size() / 3 > 0xFFFFFFFFull ||
      r.vn.size() / 3 > 0xFFFFFFFFull)
    return fail(MH_E_UNSUPPORTED, "%s: mesh too large", path);
  mh_obj_mesh& m = r;
  m.n_verts = (uint32_t)(r.v.size() / 3);
  m.n_tris = (uint32_t)(r.idx.size() / 3);
  m.n_normals = (uint32_t)(r.vn.size() / 3);
  m.n_texcoords = (uint32_t)(r.vt.size() / 3);
  m.n_materials = (uint32_t)r.names.size();
  m.verts = r.v.data();
  m.idx = r.idx.data();
  m.normals = r.vn.empty() ? nullptr : r.vn.data();
  m.normal_idx = r.nidx.empty() ? nullptr : r.nidx.data();
  m.texcoords = r.vt.empty() ? nullptr : r.vt.data();
  m.tex_idx = r.tidx.empty() ? nullptr : r.tidx.data();
  m.face_material = r.mat.data();
  m.material_names = r.name_ptrs.data();
  return MH_OK;
}

}  // namespace

int mh_obj_read(const char* path, mh_obj_mesh** out) {
  if (!path || !out) return fail(MH_E_INVALID, "null argument");
  *out = nullptr;
  Result* r = nullptr;
  try {
    Obj o;
    if (const int rc = parse(path, o)) return rc;
    r = new Result;
    if (const int rc = assemble(path, o, *r)) {
      delete r;
      return rc;
    }
  } catch (const std::bad_alloc&) {
    delete r;
    return fail(MH_E_NOMEM, "host memory for %s", path);
  }
  *out = r;
  return MH_OK;
}

void mh_obj_free(mh_obj_mesh* mesh) {
  delete static_cast<Result*>(mesh);
}
