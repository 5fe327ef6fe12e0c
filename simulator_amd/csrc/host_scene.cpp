// host_scene.cpp -- host-side set-up: spring CSR builders (the reference's
// Springs class), Mesh::vertex_extend and seeded synthetic scene generators.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <unordered_map>

#include "cc_internal.h"

namespace cc {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

static uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

// glm::distance(vertices[i], vertices[end]) on vec4 with equal w: the
// (x+y)+(z+w) vec4 dot reduces to the vec3 order.
static float rest_length(const float* a, const float* b) { return length(ld3(b) - ld3(a)); }

static void fill_csr(const std::vector<std::vector<uint32_t>>& neigh, const float* pos, size_t stride,
                     cc_csr* out) {
  size_t n = neigh.size(), nnz = 0;
  for (auto& r : neigh) nnz += r.size();
  out->n = (uint32_t)n;
  out->nnz = (uint32_t)nnz;
  out->row = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1));
  out->col = (cc_spring*)malloc(sizeof(cc_spring) * (nnz ? nnz : 1));
  if (!out->row || !out->col) throw std::bad_alloc();
  uint32_t k = 0;
  for (size_t i = 0; i < n; ++i) {
    out->row[i] = k;
    for (uint32_t e : neigh[i]) {
      out->col[k].end = (int32_t)e;
      out->col[k].original = rest_length(pos + stride * i, pos + stride * e);
      ++k;
    }
  }
  out->row[n] = k;
}

// The nearest-neighbour search of the reference's boundary springs
// (kd_create / kd_insert3f / kd_nearest3f of its kdtree.cpp), restated for
// its results, ties included: an unbalanced 3-d tree in insertion order
// (split axis = depth mod 3, a point goes left when strictly below the
// node's coordinate), double coordinates, a bounding box grown by every
// insertion.  A search takes the root as its first best, then at each node
// visits the nearer side (left when q <= the node's coordinate), tests the
// node itself (strictly closer wins), then the farther side if the box of
// that side is strictly closer than the best.  So among equally near points
// the first one reached in that order wins, and the root wins any tie with
// it.  The walk uses an explicit stack: insertion-ordered meshes give deep
// trees.
class KdTree {
 public:
  void insert(const float* p, uint32_t data) {
    Node nd;
    for (int c = 0; c < 3; ++c) nd.p[c] = (double)p[c];
    nd.data = data;
    if (nodes_.empty()) {
      nd.dir = 0;
      nodes_.push_back(nd);
      for (int c = 0; c < 3; ++c) rmin_[c] = rmax_[c] = nd.p[c];
      return;
    }
    int32_t cur = 0;
    for (;;) {
      const Node& c = nodes_[cur];
      const bool left = nd.p[c.dir] < c.p[c.dir];
      const int32_t next = left ? c.left : c.right;
      if (next >= 0) {
        cur = next;
        continue;
      }
      nd.dir = (c.dir + 1) % 3;
      const int32_t id = (int32_t)nodes_.size();
      nodes_.push_back(nd);
      (left ? nodes_[cur].left : nodes_[cur].right) = id;
      break;
    }
    for (int c = 0; c < 3; ++c) {
      if (nd.p[c] < rmin_[c]) rmin_[c] = nd.p[c];
      if (nd.p[c] > rmax_[c]) rmax_[c] = nd.p[c];
    }
  }

  bool empty() const { return nodes_.empty(); }

  uint32_t nearest(const float* qf) const {
    const double q[3] = {(double)qf[0], (double)qf[1], (double)qf[2]};
    int32_t best = 0;
    double best_d = dist_sq(nodes_[0], q);
    double mn[3] = {rmin_[0], rmin_[1], rmin_[2]}, mx[3] = {rmax_[0], rmax_[1], rmax_[2]};
    struct Frame {
      int32_t n;
      int stage;       // 0: enter, 1: nearer side done, 2: farther side done
      double saved;    // the box coordinate the side's visit replaced
    };
    std::vector<Frame> st;
    st.push_back({0, 0, 0.0});
    while (!st.empty()) {
      Frame f = st.back();
      st.pop_back();
      const Node& nd = nodes_[f.n];
      const int dir = nd.dir;
      const bool toward_left = q[dir] - nd.p[dir] <= 0;
      const int32_t nearer = toward_left ? nd.left : nd.right;
      const int32_t farther = toward_left ? nd.right : nd.left;
      double* near_c = toward_left ? &mx[dir] : &mn[dir];
      double* far_c = toward_left ? &mn[dir] : &mx[dir];
      if (f.stage == 0) {
        if (nearer >= 0) {
          st.push_back({f.n, 1, *near_c});
          *near_c = nd.p[dir];
          st.push_back({nearer, 0, 0.0});
          continue;
        }
        f.stage = 1;
        f.saved = *near_c;        // nothing was replaced: restoring is a no-op
      }
      if (f.stage == 1) {
        *near_c = f.saved;
        const double d = dist_sq(nd, q);
        if (d < best_d) {
          best = f.n;
          best_d = d;
        }
        if (farther >= 0) {
          const double sv = *far_c;
          *far_c = nd.p[dir];
          if (box_dist_sq(mn, mx, q) < best_d) {
            st.push_back({f.n, 2, sv});
            st.push_back({farther, 0, 0.0});
            continue;
          }
          *far_c = sv;
        }
        continue;
      }
      *far_c = f.saved;            // stage 2
    }
    return nodes_[best].data;
  }

 private:
  struct Node {
    double p[3];
    uint32_t data;
    int dir;
    int32_t left = -1, right = -1;
  };
  static double dist_sq(const Node& n, const double* q) {
    double d = 0;
    for (int c = 0; c < 3; ++c) d += (n.p[c] - q[c]) * (n.p[c] - q[c]);
    return d;
  }
  static double box_dist_sq(const double* mn, const double* mx, const double* q) {
    double d = 0;
    for (int c = 0; c < 3; ++c) {
      if (q[c] < mn[c])
        d += (mn[c] - q[c]) * (mn[c] - q[c]);
      else if (q[c] > mx[c])
        d += (mx[c] - q[c]) * (mx[c] - q[c]);
    }
    return d;
  }
  std::vector<Node> nodes_;
  double rmin_[3] = {0, 0, 0}, rmax_[3] = {0, 0, 0};
};

// Springs::get_cloth_boundary_spring and get_boundary_boundary_spring
// (spring.cpp:98-209) over the OBJ vertex groups, taken in pairs {cloth
// piece, its boundary}: every boundary vertex is tied to its nearest vertex
// of its piece (cloth_boundary, in group and vertex order); every boundary
// vertex's nearest vertex on the other boundaries, when 0 < distance < 50 x
// the mean length of the first 100 of those ties, gives a pair of the two
// boundary vertices' piece vertices (the sorted set `boundary`).  Inputs the
// reference reads out of range or through a null search result are
// rejected: an odd number of groups, groups past the vertex count, an empty
// piece, fewer than 100 boundary vertices, a single boundary.
void boundary_springs(const float* verts4, uint32_t nv, const uint32_t* counts, uint32_t ngroups,
                      std::vector<std::pair<uint32_t, uint32_t>>& cloth_boundary,
                      std::set<std::pair<uint32_t, uint32_t>>& boundary) {
  CC_REQUIRE(ngroups % 2 == 0, "boundary springs: the vertex groups must come in pairs (piece, boundary)");
  uint64_t total = 0;
  for (uint32_t g = 0; g < ngroups; ++g) total += counts[g];
  CC_REQUIRE(total <= nv, "boundary springs: the vertex groups cover more vertices than the mesh has");
  uint32_t g_start = 0;
  std::vector<std::pair<uint32_t, uint32_t>> ranges;     // the boundaries' vertex ranges
  for (uint32_t g = 0; g < ngroups; g += 2) {
    CC_REQUIRE(counts[g] > 0, "boundary springs: a cloth piece without vertices");
    KdTree kd;
    for (uint32_t i = 0; i < counts[g]; ++i) kd.insert(verts4 + 4 * (size_t)(g_start + i), g_start + i);
    g_start += counts[g];
    for (uint32_t i = 0; i < counts[g + 1]; ++i)
      cloth_boundary.emplace_back(g_start + i, kd.nearest(verts4 + 4 * (size_t)(g_start + i)));
    ranges.emplace_back(g_start, g_start + counts[g + 1]);
    g_start += counts[g + 1];
  }
  CC_REQUIRE(cloth_boundary.size() >= 100,
             "boundary springs: fewer than 100 boundary vertices (the reference averages the first 100)");
  float max_dist = 0.0f;
  for (int i = 0; i < 100; ++i)
    max_dist += rest_length(verts4 + 4 * (size_t)cloth_boundary[i].first, verts4 + 4 * (size_t)cloth_boundary[i].second);
  max_dist /= 100u;
  std::map<uint32_t, uint32_t> to_piece;
  for (auto& cb : cloth_boundary) to_piece[cb.first] = cb.second;
  for (size_t i = 0; i < ranges.size(); ++i) {
    KdTree kd;
    for (size_t j = 0; j < ranges.size(); ++j) {
      if (j == i) continue;
      for (uint32_t k = ranges[j].first; k < ranges[j].second; ++k) kd.insert(verts4 + 4 * (size_t)k, k);
    }
    if (ranges[i].first == ranges[i].second) continue;
    CC_REQUIRE(!kd.empty(), "boundary springs: a single boundary has no other boundary to meet");
    for (uint32_t k = ranges[i].first; k < ranges[i].second; ++k) {
      const uint32_t r = kd.nearest(verts4 + 4 * (size_t)k);
      const float d = rest_length(verts4 + 4 * (size_t)k, verts4 + 4 * (size_t)r);
      if (d < max_dist * 50 && d > 0) boundary.insert({to_piece[k], to_piece[r]});
    }
  }
}

}  // namespace cc

using namespace cc;

extern "C" {

int cc_version(void) { return 1; }
const char* cc_last_error(void) { return g_last_error.c_str(); }

int cc_device_count(int* count) {
  return guard([&] {
    CC_REQUIRE(count, "cc_device_count: null");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
  });
}

void cc_params_default(cc_params* p) {
  if (!p) return;
  p->dt = 1.0f / 40.0f;
  p->mass = 0.3f;
  p->gravity[0] = 0.0f;
  p->gravity[1] = -0.00981f;
  p->gravity[2] = 0.0f;
  p->damp = -0.02f;
  p->kd = -0.5f;
  p->k[0] = 100.0f;
  p->k[1] = 2.0f;
  p->k[2] = 0.0f;
  p->rest_scale = 1.0f;
  p->eps = FLT_EPSILON;
  p->box_depth = 0.01f;
  p->body_extend = 0.003f;
}

void cc_csr_free(cc_csr* c) {
  if (!c) return;
  free(c->row);
  free(c->col);
  c->row = nullptr;
  c->col = nullptr;
  c->n = c->nnz = 0;
}

// Springs::Springs (spring.cpp:82-91): the boundary springs of a
// SINGLE_LAYER_BOUNDARY mesh (when groups are given), then create_neigh
// (:286-352) + CSR_spring (:374-391).
static void springs_from_mesh(const float* verts4, uint32_t nv, const uint32_t* faces, uint32_t nf,
                              const uint32_t* group_counts, uint32_t ngroups, cc_csr* out_s, cc_csr* out_b) {
  {
    CC_REQUIRE(verts4 && faces && out_s && out_b, "cc_springs_from_mesh: null argument");
    for (size_t i = 0; i < 3 * (size_t)nf; ++i)
      CC_REQUIRE(faces[i] < nv, "cc_springs_from_mesh: face index out of range");
    std::vector<std::vector<uint32_t>> n1(nv), n2(nv);
    auto add1 = [&](uint32_t a, uint32_t b) {
      auto& r = n1[a];
      if (std::find(r.begin(), r.end(), b) == r.end()) r.push_back(b);
    };
    for (uint32_t f = 0; f < nf; ++f) {
      const uint32_t* t = faces + 3 * (size_t)f;
      add1(t[0], t[1]); add1(t[0], t[2]);
      add1(t[1], t[0]); add1(t[1], t[2]);
      add1(t[2], t[0]); add1(t[2], t[1]);
    }
    if (group_counts) {      // spring.cpp:314-323: appended without the duplicate test
      std::vector<std::pair<uint32_t, uint32_t>> cloth_boundary;
      std::set<std::pair<uint32_t, uint32_t>> boundary;
      boundary_springs(verts4, nv, group_counts, ngroups, cloth_boundary, boundary);
      for (auto& cb : cloth_boundary) {
        n1[cb.first].push_back(cb.second);
        n1[cb.second].push_back(cb.first);
      }
      for (auto& b : boundary) n1[b.first].push_back(b.second);
    }
    // Matrix::Insert_Matrix (spring.cpp:30-79): the first time an edge is
    // seen (either direction) its opposite vertex is remembered; every later
    // sighting pairs that vertex with the new opposite vertex.
    std::unordered_map<uint64_t, uint32_t> first_opp;
    first_opp.reserve((size_t)nf * 2);
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    auto key = [](uint32_t a, uint32_t b) { return ((uint64_t)std::min(a, b) << 32) | std::max(a, b); };
    for (uint32_t f = 0; f < nf; ++f) {
      const uint32_t* t = faces + 3 * (size_t)f;
      const uint32_t ins[3][3] = {{t[0], t[1], t[2]}, {t[0], t[2], t[1]}, {t[1], t[2], t[0]}};
      for (auto& q : ins) {
        auto it = first_opp.find(key(q[0], q[1]));
        if (it != first_opp.end()) pairs.emplace_back(it->second, q[2]);
        else first_opp.emplace(key(q[0], q[1]), q[2]);
      }
    }
    for (auto& pr : pairs) {
      n2[pr.first].push_back(pr.second);
      n2[pr.second].push_back(pr.first);
    }
    fill_csr(n1, verts4, 4, out_s);
    fill_csr(n2, verts4, 4, out_b);
  }
}

int cc_springs_from_mesh(const float* verts4, uint32_t nv, const uint32_t* faces, uint32_t nf,
                         cc_csr* out_s, cc_csr* out_b) {
  return guard([&] { springs_from_mesh(verts4, nv, faces, nf, nullptr, 0, out_s, out_b); });
}

int cc_springs_from_mesh_groups(const float* verts4, uint32_t nv, const uint32_t* faces, uint32_t nf,
                                const uint32_t* group_counts, uint32_t ngroups, cc_csr* out_s, cc_csr* out_b) {
  return guard([&] {
    CC_REQUIRE(group_counts || ngroups == 0, "cc_springs_from_mesh_groups: null group counts");
    springs_from_mesh(verts4, nv, faces, nf, group_counts ? group_counts : &ngroups, ngroups, out_s, out_b);
  });
}

int cc_springs_grid(uint32_t nx, uint32_t nz, const float* pos3, cc_csr out[3]) {
  return guard([&] {
    CC_REQUIRE(pos3 && out && nx >= 1 && nz >= 1, "cc_springs_grid: bad argument");
    static const int off[3][4][2] = {{{1, 0}, {0, 1}, {-1, 0}, {0, -1}},
                                     {{1, 1}, {-1, 1}, {1, -1}, {-1, -1}},
                                     {{2, 0}, {0, 2}, {-2, 0}, {0, -2}}};
    size_t n = (size_t)nx * nz;
    for (int t = 0; t < 3; ++t) {
      std::vector<std::vector<uint32_t>> neigh(n);
      for (uint32_t j = 0; j < nz; ++j)
        for (uint32_t i = 0; i < nx; ++i) {
          auto& r = neigh[(size_t)j * nx + i];
          for (auto& o : off[t]) {
            long ii = (long)i + o[0], jj = (long)j + o[1];
            if (ii < 0 || jj < 0 || ii >= (long)nx || jj >= (long)nz) continue;
            r.push_back((uint32_t)(jj * nx + ii));
          }
          std::sort(r.begin(), r.end());
        }
      fill_csr(neigh, pos3, 3, &out[t]);
    }
  });
}

int cc_scene_grid_cloth(uint32_t nx, uint32_t nz, double spacing, double y0, double jitter,
                        uint64_t seed, float* pos3) {
  return guard([&] {
    CC_REQUIRE(pos3, "cc_scene_grid_cloth: null output");
    uint64_t st = seed;
    for (uint32_t j = 0; j < nz; ++j)
      for (uint32_t i = 0; i < nx; ++i) {
        size_t v = (size_t)j * nx + i;
        const double base[3] = {i * spacing, y0, -0.5 + j * spacing};
        for (int c = 0; c < 3; ++c) {
          double u = uniform01(st);
          pos3[3 * v + c] = (float)(base[c] + (2.0 * u - 1.0) * jitter);
        }
      }
  });
}

// Icosahedron refined by edge-midpoint splits projected to the unit sphere.
int cc_scene_icosphere(uint32_t level, double radius, const double center[3], float* verts3,
                       uint32_t* tris, float* normals3) {
  return guard([&] {
    CC_REQUIRE(verts3 && tris && center && level <= 10, "cc_scene_icosphere: bad argument");
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<double> V = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                             0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    for (int i = 0; i < 12; ++i) {
      double l = std::sqrt(V[3 * i] * V[3 * i] + V[3 * i + 1] * V[3 * i + 1] + V[3 * i + 2] * V[3 * i + 2]);
      for (int k = 0; k < 3; ++k) V[3 * i + k] /= l;
    }
    std::vector<uint32_t> F = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                               11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                               3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (uint32_t lv = 0; lv < level; ++lv) {
      std::unordered_map<uint64_t, uint32_t> mid;
      std::vector<uint32_t> G;
      G.reserve(F.size() * 4);
      auto midpoint = [&](uint32_t a, uint32_t b) -> uint32_t {
        uint32_t lo = std::min(a, b), hi = std::max(a, b);
        uint64_t k = ((uint64_t)lo << 32) | hi;
        auto it = mid.find(k);
        if (it != mid.end()) return it->second;
        double p[3], l = 0;
        for (int c = 0; c < 3; ++c) {
          p[c] = 0.5 * (V[3 * lo + c] + V[3 * hi + c]);
          l += p[c] * p[c];
        }
        l = std::sqrt(l);
        uint32_t idx = (uint32_t)(V.size() / 3);
        for (int c = 0; c < 3; ++c) V.push_back(p[c] / l);
        mid.emplace(k, idx);
        return idx;
      };
      for (size_t f = 0; f < F.size(); f += 3) {
        uint32_t a = F[f], b = F[f + 1], c = F[f + 2];
        uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
        uint32_t nt[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
        G.insert(G.end(), nt, nt + 12);
      }
      F.swap(G);
    }
    size_t nv = V.size() / 3;
    for (size_t i = 0; i < nv; ++i)
      for (int c = 0; c < 3; ++c) {
        verts3[3 * i + c] = (float)(center[c] + radius * V[3 * i + c]);
        if (normals3) normals3[3 * i + c] = (float)V[3 * i + c];
      }
    std::memcpy(tris, F.data(), F.size() * sizeof(uint32_t));
  });
}

// Capsule k: axis p0 -> p1, radius r.  Rings run from the p0 pole to the p1
// pole: nring hemisphere rings ending at the p0 equator, ncyl cylinder rings
// ending at the p1 equator, nring-1 hemisphere rings towards the p1 pole.
// Vertex 0 is the p0 pole, then R = 2*nring + ncyl - 1 rings of nseg
// vertices, then the p1 pole; 2*nseg*R triangles, outward CCW.
int cc_scene_capsule_body(uint32_t ncaps, double rmin, double rmax, const double lo[3],
                          const double hi[3], uint32_t nseg, uint32_t nring, uint32_t ncyl,
                          uint64_t seed, float* verts3, uint32_t* tris, float* normals3,
                          uint32_t* nv_out, uint32_t* nf_out, cc_capsules* caps) {
  return guard([&] {
    CC_REQUIRE(ncaps >= 1 && nseg >= 3 && nring >= 1 && ncyl >= 1 && lo && hi && nv_out && nf_out,
               "cc_scene_capsule_body: bad argument");
    const uint32_t R = 2 * nring + ncyl - 1;
    const uint64_t vper = 2 + (uint64_t)nseg * R, fper = 2ull * nseg * R;
    CC_REQUIRE(vper * ncaps < (1ull << 31) && fper * ncaps < (1ull << 31), "capsule body too large");
    *nv_out = (uint32_t)(vper * ncaps);
    *nf_out = (uint32_t)(fper * ncaps);
    if (!verts3) return;
    CC_REQUIRE(tris != nullptr, "cc_scene_capsule_body: tris required with verts3");
    uint64_t st = seed;
    const double PI = 3.14159265358979323846;
    for (uint32_t k = 0; k < ncaps; ++k) {
      double r = rmin + (rmax - rmin) * uniform01(st);
      double c[3];
      for (int q = 0; q < 3; ++q) c[q] = lo[q] + (hi[q] - lo[q]) * uniform01(st);
      double cz = 2.0 * uniform01(st) - 1.0, ph = 2.0 * PI * uniform01(st);
      double sz = std::sqrt(std::max(0.0, 1.0 - cz * cz));
      double d[3] = {sz * std::cos(ph), cz, sz * std::sin(ph)};
      double half = 0.05 + 0.15 * uniform01(st);
      double breath_phase = 2.0 * PI * uniform01(st);
      double p0[3], p1[3];
      for (int q = 0; q < 3; ++q) {
        p0[q] = c[q] - half * d[q];
        p1[q] = c[q] + half * d[q];
      }
      // frame e1, e2 with e1 x e2 = d
      double a[3] = {1, 0, 0};
      if (std::fabs(d[0]) > 0.9) { a[0] = 0; a[1] = 1; }
      double e1[3] = {a[1] * d[2] - a[2] * d[1], a[2] * d[0] - a[0] * d[2], a[0] * d[1] - a[1] * d[0]};
      double l1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
      for (double& x : e1) x /= l1;
      double e2[3] = {d[1] * e1[2] - d[2] * e1[1], d[2] * e1[0] - d[0] * e1[2], d[0] * e1[1] - d[1] * e1[0]};
      if (caps) {
        caps->vstart[k] = (uint32_t)(vper * k);
        for (int q = 0; q < 3; ++q) {
          caps->p0[3 * k + q] = (float)p0[q];
          caps->p1[3 * k + q] = (float)p1[q];
        }
        caps->radius[k] = (float)r;
        caps->phase[k] = (float)breath_phase;
      }
      size_t vb = vper * k;
      auto put = [&](size_t v, const double p[3], const double n[3]) {
        for (int q = 0; q < 3; ++q) {
          verts3[3 * v + q] = (float)p[q];
          if (normals3) normals3[3 * v + q] = (float)n[q];
        }
      };
      double pole[3], nrm[3];
      for (int q = 0; q < 3; ++q) { pole[q] = p0[q] - r * d[q]; nrm[q] = -d[q]; }
      put(vb, pole, nrm);
      for (uint32_t ri = 0; ri < R; ++ri) {
        double cen[3], rho, axial;   // axial: normal component along d
        if (ri < nring) {
          double phi = 0.5 * PI * (ri + 1) / nring;
          for (int q = 0; q < 3; ++q) cen[q] = p0[q] - r * std::cos(phi) * d[q];
          rho = r * std::sin(phi);
          axial = -std::cos(phi);
        } else if (ri < nring + ncyl) {
          double b = (double)(ri - nring + 1) / ncyl;
          for (int q = 0; q < 3; ++q) cen[q] = p0[q] + b * (p1[q] - p0[q]);
          rho = r;
          axial = 0.0;
        } else {
          double phi = 0.5 * PI * (nring - (ri - nring - ncyl + 1)) / nring;
          for (int q = 0; q < 3; ++q) cen[q] = p1[q] + r * std::cos(phi) * d[q];
          rho = r * std::sin(phi);
          axial = std::cos(phi);
        }
        double radial = rho / r;
        for (uint32_t s = 0; s < nseg; ++s) {
          double th = 2.0 * PI * s / nseg, cs = std::cos(th), sn = std::sin(th), p[3], n[3];
          for (int q = 0; q < 3; ++q) {
            double u = cs * e1[q] + sn * e2[q];
            p[q] = cen[q] + rho * u;
            n[q] = radial * u + axial * d[q];
          }
          put(vb + 1 + (size_t)ri * nseg + s, p, n);
        }
      }
      for (int q = 0; q < 3; ++q) { pole[q] = p1[q] + r * d[q]; nrm[q] = d[q]; }
      put(vb + 1 + (size_t)R * nseg, pole, nrm);
      size_t fb = fper * k;
      uint32_t* T = tris + 3 * fb;
      size_t f = 0;
      auto tri = [&](size_t a0, size_t a1, size_t a2) {
        T[3 * f] = (uint32_t)a0;
        T[3 * f + 1] = (uint32_t)a1;
        T[3 * f + 2] = (uint32_t)a2;
        ++f;
      };
      auto ring = [&](uint32_t ri, uint32_t s) { return vb + 1 + (size_t)ri * nseg + (s % nseg); };
      for (uint32_t s = 0; s < nseg; ++s) tri(vb, ring(0, s + 1), ring(0, s));
      for (uint32_t ri = 0; ri + 1 < R; ++ri)
        for (uint32_t s = 0; s < nseg; ++s) {
          tri(ring(ri, s), ring(ri, s + 1), ring(ri + 1, s));
          tri(ring(ri, s + 1), ring(ri + 1, s + 1), ring(ri + 1, s));
        }
      size_t np = vb + 1 + (size_t)R * nseg;
      for (uint32_t s = 0; s < nseg; ++s) tri(ring(R - 1, s), ring(R - 1, s + 1), np);
    }
    if (caps) {
      caps->ncaps = ncaps;
      caps->vstart[ncaps] = *nv_out;
    }
  });
}

// Mesh::vertex_extend (Mesh.cpp:101-117)
int cc_mesh_vertex_extend(float* verts, uint32_t stride, uint32_t nv, const uint32_t* faces,
                          uint32_t nf, const uint32_t* normal_index, const float* normals3,
                          float dist) {
  return guard([&] {
    CC_REQUIRE(verts && faces && normals3 && (stride == 3 || stride == 4),
               "cc_mesh_vertex_extend: bad argument");
    std::vector<uint8_t> visited(nv, 0);
    for (uint32_t f = 0; f < nf; ++f)
      for (int j = 0; j < 3; ++j) {
        uint32_t v = faces[3 * (size_t)f + j];
        CC_REQUIRE(v < nv, "cc_mesh_vertex_extend: face index out of range");
        if (visited[v]) continue;
        uint32_t ni = normal_index ? normal_index[3 * (size_t)f + j] : v;
        const float* n = normals3 + 3 * (size_t)ni;
        float* x = verts + (size_t)stride * v;
        x[0] = x[0] + dist * n[0];
        x[1] = x[1] + dist * n[1];
        x[2] = x[2] + dist * n[2];
        visited[v] = 1;
      }
  });
}

}  // extern "C"
