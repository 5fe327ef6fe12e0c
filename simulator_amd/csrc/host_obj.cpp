// host_obj.cpp -- the reference's mesh ingest on the host: the OBJ reader
// (ObjLoader::ObjLoader, ObjLoader.cpp:15-216, geometry only) and the Mesh
// set-up steps its scenes apply (Mesh::unified, get_center, scale,
// translate, rotation; Mesh.cpp:17-38, 40-92, 119-131).
//
// Vertex groups ("# object name" / "# 1234 vertices" lines, as some exporters
// write them; Meshlab's "# Object" and "# 1234 vertices," do not match, as in
// the reference) are kept for the boundary springs.
//
// The reader keeps the reference's reading of each line, including what it
// does with inputs it was not written for: a face keeps its first three
// corners; "f a b c" faces take the normal index from the vertex index and
// texture indices 0, 1, 2; in "v/t/n" faces the second and third corners
// look for their second '/' one character late (so "v//n" there yields
// texture index atoi("/n") - 1 and normal index atoi("v//n") - 1); indices
// are atoi(...) - 1 stored unsigned.  One case is rejected instead of
// reproduced: a corner without '/' in a face whose first corner has one
// leaves the reference's indices uninitialised.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "cc_internal.h"

using namespace cc;

namespace {

// ObjLoader.cpp:130-181 for one corner of a "v/t/n" face; late = 0 for the
// first corner, 1 for the others (find("/", sPos) vs find("/", sPos + 1)).
bool parse_corner(const std::string& f, size_t late, uint32_t& v, uint32_t& t, uint32_t& n) {
  size_t ePos = f.find_first_of('/');
  if (ePos == std::string::npos) return false;
  size_t sPos = 0;
  v = (uint32_t)(std::atoi(f.substr(sPos, ePos - sPos).c_str()) - 1);
  sPos = ePos + 1;
  ePos = f.find('/', sPos + late);
  t = (uint32_t)(std::atoi(f.substr(sPos, ePos == std::string::npos ? std::string::npos : ePos - sPos).c_str()) -
                 1);
  sPos = ePos + 1;   // npos + 1 == 0: the whole token again
  ePos = f.length();
  n = (uint32_t)(std::atoi(f.substr(sPos, ePos - sPos).c_str()) - 1);
  return true;
}

template <class T>
T* copy_out(const std::vector<T>& v) {
  T* p = (T*)std::malloc(sizeof(T) * (v.empty() ? 1 : v.size()));
  if (!p) throw std::bad_alloc();
  if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

}  // namespace

extern "C" {

int cc_obj_read(const char* path, cc_obj* out) {
  return guard([&] {
    CC_REQUIRE(path && out, "cc_obj_read: bad argument");
    std::memset(out, 0, sizeof(*out));
    std::ifstream input(path);
    CC_REQUIRE((bool)input, std::string("cc_obj_read: unable to open ") + path);
    std::vector<float> v4, n3, t2;
    std::vector<uint32_t> vi, ti, ni, groups;
    std::string buffer;
    uint64_t lineno = 0;
    while (std::getline(input, buffer)) {
      ++lineno;
      std::stringstream line(buffer);
      std::string f0, f1, f2, f3;
      line >> f0 >> f1 >> f2 >> f3;
      if (f0 == "mtllib") {                  // material file: rendering only
      } else if (f1 == "object") {           // a vertex group opens ("# object name")
        groups.push_back(0);
      } else if (f2 == "vertices") {         // its vertex count ("# 1234 vertices")
        CC_REQUIRE(!groups.empty(), "cc_obj_read: vertex count before any object line at line " +
                                        std::to_string(lineno));
        groups.back() = (uint32_t)std::atoi(f1.c_str());
      } else if (f0 == "v") {                // glm::vec4(atof, atof, atof, 1.0)
        v4.insert(v4.end(), {(float)std::atof(f1.c_str()), (float)std::atof(f2.c_str()),
                             (float)std::atof(f3.c_str()), 1.0f});
      } else if (f0 == "vn") {
        n3.insert(n3.end(), {(float)std::atof(f1.c_str()), (float)std::atof(f2.c_str()),
                             (float)std::atof(f3.c_str())});
      } else if (f0 == "vt") {
        t2.insert(t2.end(), {(float)std::atof(f1.c_str()), (float)std::atof(f2.c_str())});
      } else if (f0 == "f") {
        uint32_t cv[3], ct[3], cn[3];
        if (f1.find_first_of('/') == std::string::npos) {   // f 1 2 3
          const std::string* fs[3] = {&f1, &f2, &f3};
          for (int k = 0; k < 3; ++k) {
            cv[k] = (uint32_t)(std::atoi(fs[k]->c_str()) - 1);
            ct[k] = (uint32_t)k;
            cn[k] = cv[k];
          }
        } else {                                              // f v/t/n v/t/n v/t/n
          const bool ok = parse_corner(f1, 0, cv[0], ct[0], cn[0]) && parse_corner(f2, 1, cv[1], ct[1], cn[1]) &&
                          parse_corner(f3, 1, cv[2], ct[2], cn[2]);
          CC_REQUIRE(ok, "cc_obj_read: face corner without '/' in a 'v/t/n' face at line " +
                             std::to_string(lineno));
        }
        vi.insert(vi.end(), cv, cv + 3);
        ti.insert(ti.end(), ct, ct + 3);
        ni.insert(ni.end(), cn, cn + 3);
      }
    }
    CC_REQUIRE(v4.size() / 4 < (1ull << 31) && vi.size() / 3 < (1ull << 31), "cc_obj_read: file too large");
    out->nv = (uint32_t)(v4.size() / 4);
    out->nvn = (uint32_t)(n3.size() / 3);
    out->nvt = (uint32_t)(t2.size() / 2);
    out->nf = (uint32_t)(vi.size() / 3);
    out->verts4 = copy_out(v4);
    out->normals3 = copy_out(n3);
    out->tex2 = copy_out(t2);
    out->vidx = copy_out(vi);
    out->tidx = copy_out(ti);
    out->nidx = copy_out(ni);
    out->ngroups = (uint32_t)groups.size();
    out->group_counts = copy_out(groups);
  });
}

void cc_obj_free(cc_obj* obj) {
  if (!obj) return;
  std::free(obj->verts4);
  std::free(obj->normals3);
  std::free(obj->tex2);
  std::free(obj->vidx);
  std::free(obj->tidx);
  std::free(obj->nidx);
  std::free(obj->group_counts);
  std::memset(obj, 0, sizeof(*obj));
}

int cc_mesh_unify(const cc_obj* obj, float* normals3, float* tex2) {
  return guard([&] {
    CC_REQUIRE(obj && (obj->nf == 0 || (obj->vidx && obj->nidx && obj->tidx)), "cc_mesh_unify: bad argument");
    if (normals3) std::memset(normals3, 0, 3 * sizeof(float) * (size_t)obj->nv);
    if (tex2) std::memset(tex2, 0, 2 * sizeof(float) * (size_t)obj->nv);
    for (uint32_t f = 0; f < obj->nf; ++f)
      for (int j = 0; j < 3; ++j) {
        const uint32_t v = obj->vidx[3 * (size_t)f + j];
        CC_REQUIRE(v < obj->nv, "cc_mesh_unify: face vertex index out of range");
        const uint32_t n = obj->nidx[3 * (size_t)f + j], t = obj->tidx[3 * (size_t)f + j];
        // the reference reads past its arrays for indices out of range; here
        // such a corner leaves the value of an earlier corner (or zero)
        if (normals3 && n < obj->nvn) std::memcpy(normals3 + 3 * (size_t)v, obj->normals3 + 3 * (size_t)n, 12);
        if (tex2 && t < obj->nvt) std::memcpy(tex2 + 2 * (size_t)v, obj->tex2 + 2 * (size_t)t, 8);
      }
  });
}

int cc_mesh_center(const float* verts4, uint32_t n, float center[3]) {
  return guard([&] {
    CC_REQUIRE(verts4 && center && n > 0, "cc_mesh_center: bad argument");
    float sx = 0, sy = 0, sz = 0;             // float sums in vertex order (Mesh.cpp:121-131)
    for (uint32_t i = 0; i < n; ++i) {
      sx += verts4[4 * (size_t)i];
      sy += verts4[4 * (size_t)i + 1];
      sz += verts4[4 * (size_t)i + 2];
    }
    center[0] = sx / (float)n;
    center[1] = sy / (float)n;
    center[2] = sz / (float)n;
  });
}

int cc_mesh_scale(float* verts4, uint32_t n, float s) {
  float c[3];
  const int rc = cc_mesh_center(verts4, n, c);
  if (rc) return rc;
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) verts4[4 * (size_t)i + k] = (verts4[4 * (size_t)i + k] - c[k]) * s;
  return CC_OK;
}

int cc_mesh_translate(float* verts4, uint32_t n, float x, float y, float z) {
  return guard([&] {
    CC_REQUIRE(verts4 || n == 0, "cc_mesh_translate: bad argument");
    for (uint32_t i = 0; i < n; ++i) {
      verts4[4 * (size_t)i] += x;
      verts4[4 * (size_t)i + 1] += y;
      verts4[4 * (size_t)i + 2] += z;
    }
  });
}

int cc_mesh_rotate(float* verts4, uint32_t n, float degrees, int axis) {
  float c3[3];
  const int rc = cc_mesh_center(verts4, n, c3);
  if (rc) return rc;
  return guard([&] {
    CC_REQUIRE(axis >= 0 && axis <= 2, "cc_mesh_rotate: axis must be 0 (x), 1 (y) or 2 (z)");
    const float a = (float)((double)(degrees / 180.0f) * 3.1415);   // angle / 180 * 3.1415, stored as float
    const float co = std::cos(a), si = std::sin(a);
    // glm::mat4 from column braces (Mesh.cpp:66-86), times the vec4 as glm
    // 0.9.9 evaluates it: (col0 * x + col1 * y) + (col2 * z + col3 * w)
    float m[4][4] = {};
    if (axis == 0) {
      const float cols[4][4] = {{1, 0, 0, 0}, {0, co, -si, 0}, {0, si, co, 0}, {0, 0, 0, 1}};
      std::memcpy(m, cols, sizeof(m));
    } else if (axis == 1) {
      const float cols[4][4] = {{co, 0, -si, 0}, {0, 1, 0, 0}, {si, 0, co, 0}, {0, 0, 0, 1}};
      std::memcpy(m, cols, sizeof(m));
    } else {
      const float cols[4][4] = {{co, -si, 0, 0}, {si, co, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
      std::memcpy(m, cols, sizeof(m));
    }
    for (uint32_t i = 0; i < n; ++i) {
      float* p = verts4 + 4 * (size_t)i;
      const float v[4] = {p[0] - c3[0], p[1] - c3[1], p[2] - c3[2], p[3] - 0.0f};
      float r[4];
      for (int k = 0; k < 4; ++k) r[k] = (m[0][k] * v[0] + m[1][k] * v[1]) + (m[2][k] * v[2] + m[3][k] * v[3]);
      p[0] = r[0] + c3[0];
      p[1] = r[1] + c3[1];
      p[2] = r[2] + c3[2];
      p[3] = r[3] + 0.0f;
    }
  });
}

}  // extern "C"
