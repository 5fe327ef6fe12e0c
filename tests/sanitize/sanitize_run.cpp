// sanitize_run.cpp -- the library's host code (spring builders, scene
// generators, OBJ reader, mesh set-up) and the oracle, built for the CPU
// with AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/
// Makefile) and run on inputs that reach every branch of their loops.  Each
// library result is also compared with the oracle's, so the run is a parity
// check as well as a memory / UB check.  No GPU code is built or called.
//
// Exit status 0 when every comparison holds; a sanitizer report aborts.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>
#include <vector>

#include "../../include/clothcore.h"
#include "../../oracle/oracle.h"

static int failures = 0;
#define CHECK(cond, what)                                  \
  do {                                                     \
    if (!(cond)) {                                         \
      std::fprintf(stderr, "FAILED: %s\n", what);          \
      ++failures;                                          \
    }                                                      \
  } while (0)

static bool same_csr(const cc_csr& c, const std::vector<uint32_t>& row, const std::vector<or_spring>& col) {
  if (c.nnz != col.size() || row.size() != (size_t)c.n + 1) return false;
  if (std::memcmp(c.row, row.data(), row.size() * 4) != 0) return false;
  return col.empty() || std::memcmp(c.col, col.data(), col.size() * 8) == 0;
}

// the oracle's mesh springs (two calls: counts, then arrays)
static void oracle_mesh_springs(const std::vector<float>& v4, const std::vector<uint32_t>& f, uint32_t ncb,
                                const uint32_t* cb, uint32_t nbd, const uint32_t* bd, std::vector<uint32_t> row[2],
                                std::vector<or_spring> col[2]) {
  const uint32_t nv = (uint32_t)(v4.size() / 4), nf = (uint32_t)(f.size() / 3);
  uint32_t nnz[2] = {0, 0};
  row[0].assign(nv + 1, 0);
  row[1].assign(nv + 1, 0);
  or_springs_from_mesh_ex(nv, v4.data(), nf, f.data(), ncb, cb, nbd, bd, row[0].data(), nullptr, &nnz[0],
                          row[1].data(), nullptr, &nnz[1]);
  col[0].resize(nnz[0]);
  col[1].resize(nnz[1]);
  or_springs_from_mesh_ex(nv, v4.data(), nf, f.data(), ncb, cb, nbd, bd, row[0].data(), col[0].data(), &nnz[0],
                          row[1].data(), col[1].data(), &nnz[1]);
}

static void grid_case(uint32_t nx) {
  std::vector<float> a(3 * nx * nx), b(3 * nx * nx);
  CHECK(cc_scene_grid_cloth(nx, nx, 1.0 / (nx - 1), 1.0, 1e-4, 3, a.data()) == CC_OK, "grid cloth");
  or_grid_cloth(nx, nx, 1.0 / (nx - 1), 1.0, 1e-4, 3, b.data());
  CHECK(a == b, "grid cloth equals the oracle's");
  cc_csr c[3];
  CHECK(cc_springs_grid(nx, nx, a.data(), c) == CC_OK, "grid springs");
  uint32_t nnz[3];
  or_grid_spring_counts(nx, nx, nnz);
  std::vector<uint32_t> row[3];
  std::vector<or_spring> col[3];
  for (int t = 0; t < 3; ++t) {
    row[t].resize(nx * nx + 1);
    col[t].resize(nnz[t]);
  }
  or_grid_springs(nx, nx, a.data(), row[0].data(), col[0].data(), row[1].data(), col[1].data(), row[2].data(),
                  col[2].data());
  for (int t = 0; t < 3; ++t) {
    CHECK(same_csr(c[t], row[t], col[t]), "grid springs equal the oracle's");
    cc_csr_free(&c[t]);
  }
}

// an icosphere and a capsule body: mesh springs, vertex extend, the LBVH,
// steps and queries of the oracle over them
static void mesh_case() {
  const double center[3] = {0.5, 0.5, 0.0};
  const uint32_t level = 3, V = 10u * 64u + 2u, F = 20u * 64u;
  std::vector<float> v3(3 * V), n3(3 * V), ov3(3 * V);
  std::vector<uint32_t> t(3 * F), ot(3 * F);
  CHECK(cc_scene_icosphere(level, 0.3, center, v3.data(), t.data(), n3.data()) == CC_OK, "icosphere");
  uint32_t ont = 0;
  CHECK(or_icosphere(level, 0.3, center, ov3.data(), ot.data(), &ont) == V && ont == F, "oracle icosphere");
  CHECK(v3 == ov3 && t == ot, "icosphere equals the oracle's");
  std::vector<float> v4(4 * V);
  for (uint32_t i = 0; i < V; ++i) {
    for (int k = 0; k < 3; ++k) v4[4 * i + k] = v3[3 * i + k];
    v4[4 * i + 3] = 1.0f;
  }
  cc_csr s, b;
  CHECK(cc_springs_from_mesh(v4.data(), V, t.data(), F, &s, &b) == CC_OK, "mesh springs");
  std::vector<uint32_t> row[2];
  std::vector<or_spring> col[2];
  oracle_mesh_springs(v4, t, 0, nullptr, 0, nullptr, row, col);
  CHECK(same_csr(s, row[0], col[0]) && same_csr(b, row[1], col[1]), "mesh springs equal the oracle's");
  cc_csr_free(&s);
  cc_csr_free(&b);
  // vertex extend, library and oracle
  std::vector<float> e4 = v4, oe4 = v4;
  CHECK(cc_mesh_vertex_extend(e4.data(), 4, V, t.data(), F, nullptr, n3.data(), 0.003f) == CC_OK, "extend");
  or_vertex_extend(V, oe4.data(), F, t.data(), t.data(), n3.data(), 0.003f);
  CHECK(e4 == oe4, "vertex extend equals the oracle's");

  // capsule body through the oracle's LBVH, steps and queries
  uint32_t cv = 0, cf = 0;
  std::vector<float> bv(3 * 200000), bn(3 * 200000);
  std::vector<uint32_t> bt(3 * 400000), vstart(13);
  std::vector<float> p0(36), p1(36), rad(12), ph(12);
  cc_capsules caps{12, vstart.data(), p0.data(), p1.data(), rad.data(), ph.data()};
  const double lo[3] = {0.1, 0.2, -0.4}, hi[3] = {0.9, 0.8, 0.4};
  CHECK(cc_scene_capsule_body(12, 0.04, 0.15, lo, hi, 16, 4, 2, 7, bv.data(), bt.data(), bn.data(), &cv, &cf,
                              &caps) == CC_OK, "capsule body");
  bv.resize(3 * cv);
  bt.resize(3 * cf);
  std::vector<uint64_t> keys(cf);
  std::vector<uint32_t> order(cf), leaf_tri(3 * cf);
  std::vector<int32_t> ca(cf - 1), cb(cf - 1), pi(cf - 1), pl(cf);
  std::vector<float> bl(6 * cf), bi(6 * (cf - 1));
  const uint32_t kept = or_bvh_build(bv.data(), cf, bt.data(), 0.002f, 0, keys.data(), order.data(), leaf_tri.data(),
                                     ca.data(), cb.data(), pi.data(), pl.data(), bl.data(), bi.data());
  CHECK(kept == cf, "oracle LBVH keeps every triangle");
  or_bvh bvh{cf, bv.data(), leaf_tri.data(), ca.data(), cb.data(), bi.data(), bl.data()};
  // a 24x24 cloth falling onto it: FP32 and FP64 steps, the rollout and its adjoint
  const uint32_t nx = 24, n = nx * nx;
  std::vector<float> x(3 * n);
  or_grid_cloth(nx, nx, 1.0 / (nx - 1), 0.82, 1e-4, 1, x.data());
  uint32_t nnz[3];
  or_grid_spring_counts(nx, nx, nnz);
  std::vector<uint32_t> gr[3];
  std::vector<or_spring> gc[3];
  for (int k = 0; k < 3; ++k) {
    gr[k].resize(n + 1);
    gc[k].resize(nnz[k]);
  }
  or_grid_springs(nx, nx, x.data(), gr[0].data(), gc[0].data(), gr[1].data(), gc[1].data(), gr[2].data(),
                  gc[2].data());
  or_params prm{1e-3f, 0.3f, {0.0f, -9.8f, 0.0f}, -0.02f, -0.5f, {1000.0f, 500.0f, 100.0f}, 1.0f, 1.1920929e-7f};
  std::vector<uint8_t> pins(n, 0);
  pins[0] = pins[nx - 1] = 1;
  std::vector<float> xp = x, c(3 * n, 0.0f), ox(3 * n), oxp(3 * n), oc(3 * n), pre(3 * n);
  std::vector<int32_t> hit(n);
  std::vector<uint8_t> ins(n);
  for (int step = 0; step < 60; ++step) {
    or_step_f32(n, x.data(), xp.data(), c.data(), gr[0].data(), gc[0].data(), gr[1].data(), gc[1].data(),
                gr[2].data(), gc[2].data(), &prm, pins.data(), &bvh, ox.data(), oxp.data(), oc.data(), pre.data(),
                hit.data(), ins.data(), step % 2 ? 4 : 1);
    x.swap(ox);
    xp.swap(oxp);
    c.swap(oc);
  }
  std::vector<double> x64(x.begin(), x.end()), xp64(xp.begin(), xp.end()), c64(c.begin(), c.end()), y(3 * n),
      yp(3 * n), yc(3 * n);
  or_step_f64(n, x64.data(), xp64.data(), c64.data(), gr[0].data(), gc[0].data(), gr[1].data(), gc[1].data(),
              gr[2].data(), gc[2].data(), &prm, pins.data(), &bvh, y.data(), yp.data(), yc.data(), ins.data(),
              hit.data());
  or_params64 p64;
  or_params_widen(&prm, &p64);
  const uint32_t T = 40;
  std::vector<double> tp(3 * n * (T + 1)), tq(tp.size()), tc(tp.size()), ta(tp.size()), xf(3 * n), g(3 * n, 1.0),
      dv0(3 * n), dx0(3 * n), xz(3 * n);
  or_rollout_f64(n, x64.data(), nullptr, gr[0].data(), gc[0].data(), gr[1].data(), gc[1].data(), gr[2].data(),
                 gc[2].data(), &p64, pins.data(), &bvh, T, xf.data(), tp.data(), tq.data(), tc.data(), ta.data());
  or_rollout_f64_frozen(n, x64.data(), nullptr, gr[0].data(), gc[0].data(), gr[1].data(), gc[1].data(),
                        gr[2].data(), gc[2].data(), &p64, pins.data(), T, tc.data(), ta.data(), xz.data(), nullptr,
                        nullptr);
  CHECK(xz == xf, "frozen rollout at the recorded parameters equals the rollout");
  double dp[5];
  or_adjoint_f64(n, gr[0].data(), gc[0].data(), gr[1].data(), gc[1].data(), gr[2].data(), gc[2].data(), &p64,
                 pins.data(), T, tp.data(), tq.data(), tc.data(), g.data(), dp, dv0.data(), dx0.data());
  // queries, picking, normals, deformation, nearest neighbours
  std::vector<float> op(3 * n), on(3 * n);
  or_collide_points(n, x.data(), &bvh, hit.data(), ins.data(), op.data(), on.data());
  std::vector<uint32_t> faces;
  for (uint32_t j = 0; j + 1 < nx; ++j)
    for (uint32_t i = 0; i + 1 < nx; ++i) {
      const uint32_t a = j * nx + i;
      faces.insert(faces.end(), {a, a + nx, a + 1, a + 1, a + nx, a + nx + 1});
    }
  const float o[3] = {0.5f, 2.0f, 0.0f}, d[3] = {0.0f, -1.0f, 0.0f};
  float tt = 0.0f;
  CHECK(or_pick((uint32_t)faces.size() / 3, faces.data(), x.data(), o, d, &tt) >= 0, "pick hits the cloth");
  std::vector<float> vn(3 * n);
  or_vertex_normals(n, x.data(), (uint32_t)faces.size() / 3, faces.data(), vn.data());
  std::vector<float> def(3 * cv), br(12, 1.01f);
  const float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, Tr[3] = {0.01f, 0.0f, 0.0f};
  or_body_deform(cv, bv.data(), 12, vstart.data(), p0.data(), p1.data(), br.data(), R, Tr, def.data());
  std::vector<uint32_t> near(n);
  CHECK(or_kd_nearest(cv, bv.data(), n, x.data(), near.data()) == 0, "kd nearest");
}

// two grid pieces with boundary vertex groups: boundary springs
static void boundary_case() {
  const uint32_t nx = 17;
  const float h = 1.0f / 16.0f;
  std::vector<float> v4;
  std::vector<uint32_t> f, groups;
  for (int piece = 0; piece < 2; ++piece) {
    const uint32_t base = (uint32_t)(v4.size() / 4);
    const float x0 = piece * ((nx - 1) * h + 0.125f);
    for (uint32_t j = 0; j < nx; ++j)
      for (uint32_t i = 0; i < nx; ++i) v4.insert(v4.end(), {x0 + i * h, 0.0f, j * h, 1.0f});
    for (uint32_t j = 0; j + 1 < nx; ++j)
      for (uint32_t i = 0; i + 1 < nx; ++i) {
        const uint32_t a = base + j * nx + i;
        f.insert(f.end(), {a, a + nx, a + 1, a + 1, a + nx, a + nx + 1});
      }
    groups.push_back(nx * nx);
    uint32_t nb = 0;
    for (uint32_t k = 0; k + 1 < nx; ++k) {
      const float m = (k + 0.5f) * h;
      v4.insert(v4.end(), {x0 + m, 1.0f / 64, 0.0f, 1.0f, x0 + m, 1.0f / 64, (nx - 1) * h, 1.0f,
                           x0, 1.0f / 64, m, 1.0f, x0 + (nx - 1) * h, 1.0f / 64, m, 1.0f});
      nb += 4;
    }
    groups.push_back(nb);
  }
  const uint32_t nv = (uint32_t)(v4.size() / 4);
  cc_csr s, b;
  CHECK(cc_springs_from_mesh_groups(v4.data(), nv, f.data(), (uint32_t)f.size() / 3, groups.data(), 4, &s, &b) ==
            CC_OK, "boundary springs");
  std::vector<uint32_t> cbp(2 * nv), bd(2 * nv);
  uint32_t nbd = 0;
  const int32_t ncb = or_boundary_springs(nv, v4.data(), 4, groups.data(), cbp.data(), bd.data(), &nbd);
  CHECK(ncb > 0, "oracle boundary springs");
  std::vector<uint32_t> row[2];
  std::vector<or_spring> col[2];
  oracle_mesh_springs(v4, f, (uint32_t)ncb, cbp.data(), nbd, bd.data(), row, col);
  CHECK(same_csr(s, row[0], col[0]) && same_csr(b, row[1], col[1]), "boundary springs equal the oracle's");
  cc_csr_free(&s);
  cc_csr_free(&b);
  // a rejected layout (odd group count) reports an error and leaks nothing
  cc_csr s2, b2;
  CHECK(cc_springs_from_mesh_groups(v4.data(), nv, f.data(), (uint32_t)f.size() / 3, groups.data(), 3, &s2, &b2) !=
            CC_OK, "odd group count rejected");
}

static std::string write_tmp(const char* text) {
  char path[] = "/tmp/sanitize_objXXXXXX";
  const int fd = mkstemp(path);
  if (fd < 0) return "";
  const ssize_t w = write(fd, text, std::strlen(text));
  close(fd);
  return w == (ssize_t)std::strlen(text) ? std::string(path) : std::string();
}

static void obj_case() {
  const char* good =
      "# object piece\n# 4 vertices\nmtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\nvn 0 0 -1\n"
      "vt 0 0\nvt 1 0\nvt 1 1\nf 1 2 3\nf 1/1/1 3/3/2 4/2/1\nf 1//1 2//2 3//1\nf 2/1 3/2 4/3\nf 1 2 3 4\n";
  const std::string p = write_tmp(good);
  CHECK(!p.empty(), "temporary OBJ");
  cc_obj o;
  CHECK(cc_obj_read(p.c_str(), &o) == CC_OK && o.nv == 4 && o.nf == 5, "OBJ read");
  std::vector<float> n3(3 * o.nv), t2(2 * o.nv);
  CHECK(cc_mesh_unify(&o, n3.data(), t2.data()) == CC_OK, "unify");
  float ctr[3];
  CHECK(cc_mesh_center(o.verts4, o.nv, ctr) == CC_OK, "center");
  CHECK(cc_mesh_scale(o.verts4, o.nv, 2.0f) == CC_OK && cc_mesh_translate(o.verts4, o.nv, 0, 1, 0) == CC_OK &&
            cc_mesh_rotate(o.verts4, o.nv, 90.0f, 0) == CC_OK,
        "transforms");
  cc_obj_free(&o);
  std::remove(p.c_str());
  const std::string q = write_tmp("v 0 0 0\nf 1/1/1 2 3\n");
  cc_obj bad;
  CHECK(cc_obj_read(q.c_str(), &bad) != CC_OK, "malformed OBJ rejected");
  std::remove(q.c_str());
  CHECK(cc_obj_read("/nonexistent/none.obj", &bad) != CC_OK, "missing OBJ rejected");
}

int main() {
  grid_case(16);
  grid_case(111);
  mesh_case();
  boundary_case();
  obj_case();
  std::printf("sanitize_run: %s (%d failures)\n", failures ? "FAILED" : "ok", failures);
  return failures ? 1 : 0;
}
