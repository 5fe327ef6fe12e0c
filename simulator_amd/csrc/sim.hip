// sim.hip -- the cloth simulator: fused Verlet + CSR spring + collision step
// (verlet.cu:52-161), graph-captured per-frame sequence, state I/O, VBO
// writer and the C ABI of the simulator.
//
// Device layout (per vertex):
//   state[2]  24 B {x_cur.xyz, x_prev.xyz} ping-pong; a neighbour gather
//             reads one 24-byte record instead of two 12-byte vectors
//   contact   12 B contact normal of the last step (verlet.cu:151)
//   row       4 B merged CSR row start of all spring types; bit 31 = pinned
//   rec       8 B per directed spring {end | type << 30, rest length}
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cfloat>
#include <cstring>
#include <memory>
#include <vector>

#include <hip/hip_ext.h>

#include "cc_step.h"

namespace cc {

// One thread per vertex.  Spring forces use FMA and rsqrt (positions are
// checked against the FP64 oracle by tolerance); the collision part is exact
// order so the contact set is bit-exact for the pre-projection positions.
// WIDE: the body is walked through its four-wide form (the walk's stack in
// LDS); otherwise the kernel uses no LDS, which leaves the CUs' LDS to the
// body-update kernels that run beside it.
template <bool DEBUG, bool COUNT, int EW, bool WIDE>
__global__ __launch_bounds__(256) void k_verlet(const float* __restrict__ sin_, float* __restrict__ sout,
                                                const float* contact_in, float* contact_out, const uint32_t* __restrict__ row,
                                                const uint2* __restrict__ rec, StepArgs a, BvhView bvh,
                                                uint32_t n, int32_t* __restrict__ spill, float* __restrict__ dbg_pre,
                                                int32_t* __restrict__ dbg_hit, uint8_t* __restrict__ dbg_in,
                                                unsigned long long* __restrict__ ctr) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t visits = 0;
  const bool active = i < n;
  if (active) {
    int32_t* lds = nullptr;
    if constexpr (WIDE) {
      __shared__ int32_t wstk[WSTK_LDS * WSTK_LANES];
      lds = wstk + threadIdx.x;
    }
    step_vertex<DEBUG, EW, WIDE>(i, sin_, sout, contact_in, contact_out, row, rec, a, bvh,
                                 spill_of(spill, i, n, lds), visits, dbg_pre, dbg_hit, dbg_in);
  }
  if (COUNT) {
    // block sum into this block's own counter pair (no shared atomics)
    __shared__ unsigned long long wsum[2][4];
    unsigned long long v = visits, qn = active ? 1ull : 0ull;
    for (int off = 32; off > 0; off >>= 1) {
      v += __shfl_down(v, off, 64);
      qn += __shfl_down(qn, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      wsum[0][threadIdx.x >> 6] = v;
      wsum[1][threadIdx.x >> 6] = qn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      ctr[2 * blockIdx.x] += wsum[0][0] + wsum[0][1] + wsum[0][2] + wsum[0][3];
      ctr[2 * blockIdx.x + 1] += wsum[1][0] + wsum[1][1] + wsum[1][2] + wsum[1][3];
    }
  }
}

__global__ __launch_bounds__(256) void k_collide_points(BvhView bvh, const float* __restrict__ pts, uint32_t n,
                                                        int32_t* __restrict__ spill, int32_t* __restrict__ hit,
                                                        uint8_t* __restrict__ ins, float* __restrict__ out_pos,
                                                        float* __restrict__ out_n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  __shared__ int32_t wstk[WSTK_LDS * WSTK_LANES];
  if (i >= n) return;
  f3 p = ld3(pts + 3 * (size_t)i), nc;
  int32_t h;
  uint32_t visits = 0;
  const bool in = collide_project(bvh, p, nc, h, visits, spill_of(spill, i, n, wstk + threadIdx.x), load_wide_root(bvh));
  hit[i] = h;
  ins[i] = in ? 1 : 0;
  st3(out_pos + 3 * (size_t)i, p);
  st3(out_n + 3 * (size_t)i, nc);
}

void launch_collide_points(const BvhView& v, const float* pts, uint32_t n, int32_t* spill, int32_t* hit,
                           uint8_t* ins, float* out_pos, float* out_n, hipStream_t s) {
  if (n == 0) return;
  k_collide_points<<<grid_for(n, 256), 256, 0, s>>>(v, pts, n, spill, hit, ins, out_pos, out_n);
  CC_LAUNCHED();
}

__global__ void k_finite(const float* __restrict__ s, uint32_t n, uint32_t* __restrict__ bad) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  bool ok = true;
  for (int c = 0; c < 6; ++c) ok = ok && isfinite(s[6 * (size_t)i + c]);
  if (!ok) atomicOr(bad, 1u);
}

// compute_face_normal (verlet.cu:91-116)
__global__ void k_face_normals(const float* __restrict__ s, const uint32_t* __restrict__ faces, uint32_t nf,
                               float* __restrict__ fn) {
  const uint32_t f = blockIdx.x * 256 + threadIdx.x;
  if (f >= nf) return;
  const uint32_t* t = faces + 3 * (size_t)f;
  const f3 a = ld3(s + 6 * (size_t)t[0]), b = ld3(s + 6 * (size_t)t[1]), c = ld3(s + 6 * (size_t)t[2]);
  st3(fn + 3 * (size_t)f, normalize(cross(b - a, c - a)));
}

// update_vbo_pos + compute_vbo_normal (verlet.cu:119-127, 163-181)
__global__ void k_vbo(const float* __restrict__ s, const uint32_t* __restrict__ vrow,
                      const uint32_t* __restrict__ vcol, const float* __restrict__ fn, uint32_t n,
                      float* __restrict__ vbo_pos, float* __restrict__ vbo_nrm, float* __restrict__ keep_n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float* x = s + 6 * (size_t)i;
  reinterpret_cast<float4*>(vbo_pos)[i] = make_float4(x[0], x[1], x[2], 1.0f);
  f3 nn = mk(0.0f, 0.0f, 0.0f);
  for (uint32_t k = vrow[i]; k < vrow[i + 1]; ++k) nn = nn + ld3(fn + 3 * (size_t)vcol[k]);
  nn = normalize(nn);
  st3(vbo_nrm + 3 * (size_t)i, nn);
  st3(keep_n + 3 * (size_t)i, nn);
}

void check_state(const float* state, uint32_t n, DBuf<int32_t>& spill, size_t lanes, uint32_t* flag,
                 hipStream_t s, const char* who) {
  CC_HIP(hipMemsetAsync(flag, 0, 4, s));
  if (n) k_finite<<<grid_for(n, 256), 256, 0, s>>>(state, n, flag);
  CC_LAUNCHED();
  uint32_t bad = 0, fault = 0;
  int32_t* tail = spill.p + (size_t)STACK_SPILL * lanes;
  CC_HIP(hipMemcpyAsync(&bad, flag, 4, hipMemcpyDeviceToHost, s));
  CC_HIP(hipMemcpyAsync(&fault, tail, 4, hipMemcpyDeviceToHost, s));
  CC_HIP(hipStreamSynchronize(s));
  if (fault) {
    CC_HIP(hipMemsetAsync(tail, 0, 4, s));
    CC_HIP(hipStreamSynchronize(s));
    throw Error(CC_ERR_TRAVERSAL, std::string(who) +
                    ((fault & TRAV_OVERFLOW) ? ": traversal stack overflow: a collision query needed more deferred "
                                               "nodes than the stack holds and was taken as a miss"
                                             : ": traversal visit bound reached (corrupted tree)"));
  }
  if (bad) throw Error(CC_ERR_NONFINITE, std::string(who) + ": non-finite position found");
}

StepArgs step_args(const cc_params& prm, uint32_t ntypes) {
  StepArgs a;
  a.inv_dt = 1.0f / prm.dt;
  a.dt2 = prm.dt * prm.dt;
  a.inv_mass = 1.0f / prm.mass;
  a.gmx = prm.gravity[0] * prm.mass;
  a.gmy = prm.gravity[1] * prm.mass;
  a.gmz = prm.gravity[2] * prm.mass;
  a.damp = prm.damp;
  a.kd = prm.kd;
  a.k0 = prm.k[0];
  a.k1 = prm.k[1];
  a.k2 = prm.k[2];
  a.rest_scale = prm.rest_scale;
  a.eps = prm.eps;
  a.ntypes = ntypes;
  return a;
}

// ---------------------------------------------------------- Sim launches
// One launch of a step kernel; with events, through hipExtLaunchKernel,
// which stamps e0 / e1 with the kernel's own start and end (no marker
// packets between the kernels of the step stream).
template <class K>
static void launch_one(K kern, unsigned g, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const float* in, float* out,
                       const float* c_in, float* c_out, const uint32_t* row, const uint2* rec, const StepArgs& a,
                       const BvhView& v, uint32_t n, int32_t* spill, float* dbg_pre, int32_t* dbg_hit,
                       uint8_t* dbg_in, unsigned long long* ctr) {
  if (e0 || e1)
    hipExtLaunchKernelGGL(kern, dim3(g), dim3(256), 0, s, e0, e1, 0, in, out, c_in, c_out, row, rec, a, v, n, spill,
                          dbg_pre, dbg_hit, dbg_in, ctr);
  else
    kern<<<g, 256, 0, s>>>(in, out, c_in, c_out, row, rec, a, v, n, spill, dbg_pre, dbg_hit, dbg_in, ctr);
}

template <bool DEBUG, bool COUNT>
static void launch_k_verlet(const Sim& S, unsigned g, hipStream_t s, const float* in, float* out, const float* contact_in, float* contact_out,
                            const StepArgs& a, const BvhView& v, float* dbg_pre, int32_t* dbg_hit, uint8_t* dbg_in,
                            unsigned long long* ctr, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  const uint2* recs = S.ell_w ? S.ell.p : S.rec.p;
#define CC_STEP_ARGS g, s, e0, e1, in, out, contact_in, contact_out, S.row.p, recs, a, v, S.n, S.spill.p, dbg_pre, \
                     dbg_hit, dbg_in, ctr
  if (v.wide) {
    if (S.ell_w == ELL_W)
      launch_one(k_verlet<DEBUG, COUNT, ELL_W, true>, CC_STEP_ARGS);
    else if (S.ell_w == ELL_W2)
      launch_one(k_verlet<DEBUG, COUNT, ELL_W2, true>, CC_STEP_ARGS);
    else
      launch_one(k_verlet<DEBUG, COUNT, 0, true>, CC_STEP_ARGS);
  } else {
    if (S.ell_w == ELL_W)
      launch_one(k_verlet<DEBUG, COUNT, ELL_W, false>, CC_STEP_ARGS);
    else if (S.ell_w == ELL_W2)
      launch_one(k_verlet<DEBUG, COUNT, ELL_W2, false>, CC_STEP_ARGS);
    else
      launch_one(k_verlet<DEBUG, COUNT, 0, false>, CC_STEP_ARGS);
  }
#undef CC_STEP_ARGS
  CC_LAUNCHED();
}

void Sim::launch_verlet(int k, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  const float* in = state[cur].p;
  float* out = state[cur ^ 1].p;
  const BvhView v = body.view(k);
  const unsigned g = grid_for(n, 256);
  const StepArgs a = args();
  if (debug && counting)
    launch_k_verlet<true, true>(*this, g, s, in, out, contact.p, contact.p, a, v, dbg_pre.p, dbg_hit.p, dbg_in.p, ctr.p,
                                e0, e1);
  else if (debug)
    launch_k_verlet<true, false>(*this, g, s, in, out, contact.p, contact.p, a, v, dbg_pre.p, dbg_hit.p, dbg_in.p,
                                 ctr.p, e0, e1);
  else if (counting)
    launch_k_verlet<false, true>(*this, g, s, in, out, contact.p, contact.p, a, v, dbg_pre.p, dbg_hit.p, dbg_in.p,
                                 ctr.p, e0, e1);
  else
    launch_k_verlet<false, false>(*this, g, s, in, out, contact.p, contact.p, a, v, dbg_pre.p, dbg_hit.p, dbg_in.p,
                                  ctr.p, e0, e1);
}

void Sim::launch_step(const float* in, float* out, const float* c_in, float* c_out, const StepArgs& a, int k,
                      hipStream_t s) {
  launch_k_verlet<false, false>(*this, grid_for(n, 256), s, in, out, c_in, c_out, a, body.view(k), nullptr, nullptr,
                                nullptr, nullptr);
}

}  // namespace cc

using namespace cc;

namespace cc {

// rayTriangleIntersect (Ray.cpp:8-42, Moller-Trumbore without culling) in
// glm's operation order, for the picking kernel.
__device__ inline bool ray_triangle(f3 o, f3 d, f3 v0, f3 v1, f3 v2, float& t) {
  const f3 v0v1 = v1 - v0, v0v2 = v2 - v0;
  const f3 pvec = cross(d, v0v2);
  const float det = dot(v0v1, pvec);
  if (fabsf(det) < FLT_EPSILON) return false;
  const float invDet = 1 / det;
  const f3 tvec = o - v0;
  const float u = dot(tvec, pvec) * invDet;
  if (u < 0 || u > 1) return false;
  const f3 qvec = cross(tvec, v0v1);
  const float v = dot(d, qvec) * invDet;
  if (v < 0 || u + v > 1) return false;
  t = dot(v0v2, qvec) * invDet;
  return true;
}

// find_first_intersect_triangle (Ray.cpp:45-77) over all faces at once: key
// = (t as an ordered integer, -0 == +0) << 32 | (2^32 - 1 - face), so the
// smallest key is the smallest t and, among equal t, the later face -- the
// entry map<float, int> keeps.  One 64-bit atomic minimum per wave.
__global__ __launch_bounds__(256) void k_pick(const float* __restrict__ state, const uint32_t* __restrict__ faces,
                                              uint32_t nf, f3 o, f3 d, unsigned long long* __restrict__ out) {
  unsigned long long best = ~0ull;
  for (uint32_t f = blockIdx.x * 256 + threadIdx.x; f < nf; f += gridDim.x * 256) {
    const uint32_t* x = faces + 3 * (size_t)f;
    float t;
    if (!ray_triangle(o, d, ld3(state + 6 * (size_t)x[0]), ld3(state + 6 * (size_t)x[1]),
                      ld3(state + 6 * (size_t)x[2]), t) ||
        t != t)
      continue;
    const uint32_t b = __float_as_uint(t + 0.0f);
    const uint32_t ord = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    const unsigned long long key = ((unsigned long long)ord << 32) | (0xFFFFFFFFu - f);
    best = key < best ? key : best;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long other = __shfl_xor(best, off, 64);
    best = other < best ? other : best;
  }
  if ((threadIdx.x & 63u) == 0 && best != ~0ull) atomicMin(out, best);
}

}  // namespace cc

void cc::merge_springs(const cc_cloth_desc* c, std::vector<uint32_t>& row, std::vector<uint2>& rec) {
  const uint32_t n = c->n;
  row.assign((size_t)n + 1, 0);
  size_t total = 0;
  for (uint32_t t = 0; t < c->ntypes; ++t) {
    CC_REQUIRE(c->row[t] && (c->col[t] || c->row[t][n] == 0), "cloth: CSR arrays required for each type");
    CC_REQUIRE(c->row[t][0] == 0, "cloth: CSR row must start at 0");
    for (uint32_t i = 0; i < n; ++i) CC_REQUIRE(c->row[t][i] <= c->row[t][i + 1], "cloth: CSR row not ascending");
    total += c->row[t][n];
  }
  CC_REQUIRE(total < ROW_MASK, "cloth: too many springs");
  rec.resize(total);
  size_t k = 0;
  for (uint32_t i = 0; i < n; ++i) {
    row[i] = (uint32_t)k;
    for (uint32_t t = 0; t < c->ntypes; ++t)
      for (uint32_t e = c->row[t][i]; e < c->row[t][i + 1]; ++e) {
        const cc_spring& sp = c->col[t][e];
        CC_REQUIRE(sp.end >= 0 && (uint32_t)sp.end < n, "cloth: spring end out of range");
        uint32_t bits;
        std::memcpy(&bits, &sp.original, 4);
        rec[k++] = make_uint2((uint32_t)sp.end | (t << 30), bits);
      }
  }
  row[n] = (uint32_t)k;
}

void cc::read_vec3_from(const float* v, uint32_t stride, uint32_t n, std::vector<float>& out) {
  out.resize(3 * (size_t)n);
  for (uint32_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) out[3 * (size_t)i + c] = v[(size_t)stride * i + c];
}

extern "C" {

int cc_sim_create(const cc_cloth_desc* cloth, const cc_body_desc* bd, const cc_params* params, cc_sim** out) {
  return guard([&] {
    CC_REQUIRE(cloth && params && out, "cc_sim_create: null argument");
    CC_REQUIRE(cloth->n >= 1 && cloth->verts4, "cc_sim_create: empty cloth");
    CC_REQUIRE(cloth->ntypes >= 1 && cloth->ntypes <= 3, "cc_sim_create: ntypes must be 1..3");
    CC_REQUIRE(params->dt > 0.0f && params->mass > 0.0f, "cc_sim_create: dt and mass must be positive");
    *out = nullptr;
    int ndev = 0;
    CC_HIP(hipGetDeviceCount(&ndev));
    CC_REQUIRE(ndev > 0, "cc_sim_create: no HIP device");
    std::unique_ptr<cc_sim> h(new cc_sim);
    Sim& S = h->s;
    CC_HIP(hipGetDevice(&S.device));
    CC_HIP(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
    S.n = cloth->n;
    S.ntypes = cloth->ntypes;
    S.prm = *params;
    hipStream_t s = S.stream;
    std::vector<uint2> rec;
    merge_springs(cloth, S.h_row, rec);
    S.row.upload(S.h_row.data(), S.h_row.size(), s);
    if (rec.empty()) {
      S.rec.alloc(1);
      S.rec.zero(s);
    } else {
      S.rec.upload(rec.data(), rec.size(), s);
    }
    // x_last = x_cur, contact = 0 (simulation.cu:65-85)
    std::vector<float> x3;
    read_vec3_from(cloth->verts4, 4, S.n, x3);
    std::vector<float> packed(6 * (size_t)S.n);
    for (uint32_t i = 0; i < S.n; ++i)
      for (int c = 0; c < 3; ++c) {
        packed[6 * (size_t)i + c] = x3[3 * (size_t)i + c];
        packed[6 * (size_t)i + 3 + c] = x3[3 * (size_t)i + c];
      }
    // fixed-width copy of the records when every vertex has at most ELL_W
    // (or ELL_W2) springs (CC_SPRINGS_ELL=0 keeps the merged rows only)
    uint32_t max_deg = 0;
    for (uint32_t i = 0; i < S.n; ++i) max_deg = std::max(max_deg, S.h_row[i + 1] - S.h_row[i]);
    // the 24-wide kernel runs two waves per SIMD: only for cloths whose
    // waves all fit at once on the chip (256 CUs x 8 waves x 64 vertices)
    uint32_t w = max_deg <= (uint32_t)ELL_W ? ELL_W
                 : (max_deg <= (uint32_t)ELL_W2 && S.n <= 256u * 8u * 64u) ? ELL_W2 : 0;
    if (const char* e = std::getenv("CC_SPRINGS_ELL"))
      if (std::atoi(e) == 0) w = 0;
    if (w) {
      std::vector<uint2> er((size_t)S.n * w);
      for (uint32_t i = 0; i < S.n; ++i)
        for (uint32_t k = 0; k < w; ++k) {
          const uint32_t e = S.h_row[i] + k;
          er[(size_t)i * w + k] = e < S.h_row[i + 1] ? rec[e] : make_uint2(i | PAD_TYPE, 0u);
        }
      S.ell.upload(er.data(), er.size(), s);
      S.ell_w = (int)w;
    }
    alloc_spill(S.spill, S.n, s);
    S.state[0].upload(packed.data(), packed.size(), s);
    S.state[1].alloc(packed.size());
    S.state[1].zero(s);
    S.contact.alloc(3 * (size_t)S.n);
    S.contact.zero(s);
    S.ctr.alloc(2 * (size_t)grid_for(S.n, 256));   // per-block visit / query counts
    S.ctr.zero(s);
    S.flag.alloc(1);
    if (bd && bd->mode != CC_BODY_NONE) {
      CC_REQUIRE(bd->verts && bd->tris && (bd->stride == 3 || bd->stride == 4), "cc_sim_create: bad body");
      std::vector<float> bv;
      read_vec3_from(bd->verts, bd->stride, bd->nv, bv);
      if (bd->extend) {
        CC_REQUIRE(bd->normals3, "cc_sim_create: body normals required for extend");
        int rc = cc_mesh_vertex_extend(bv.data(), 3, bd->nv, bd->tris, bd->nf, nullptr, bd->normals3,
                                       params->body_extend);
        if (rc) throw Error(rc, cc_last_error());
      }
      MotionParams mp;
      mp.amp = bd->amp;
      mp.freq = bd->freq;
      mp.rot_amp = bd->rot_amp;
      mp.breath_amp = bd->breath_amp;
      mp.breath_freq = bd->breath_freq;
      mp.dx = bd->dir[0];
      mp.dy = bd->dir[1];
      mp.dz = bd->dir[2];
      mp.px = bd->pivot[0];
      mp.py = bd->pivot[1];
      mp.pz = bd->pivot[2];
      mp.dt = (double)params->dt;
      // Update lanes and frames per chain of launches.  A refitted body:
      // three lanes (the 501k-triangle body, A/B in one process per setting
      // on the same GPU: 44.6-44.8 us per step in graph blocks, frames
      // 50-530, against 49.3-49.5 with two lanes and 52.6-53.5 with four;
      // one call per frame 52.4-52.7 against 52.2-52.4 with two) of three
      // frames (20 single steps 33.8-34.5 us against 34.5-35.4 with two,
      // 500 steps 41.7-42.0 against 42.2-42.6, one call per frame the same).
      // A rebuilt body, whose update is a chain of 17 small launches: two
      // lanes of three frames (111 x 111 over the 50,688-triangle body:
      // 35.4-35.6 us per step over 500 steps through the graphs against
      // 37.0-37.5 with three lanes of two, 40.0-40.1 against 41.0-42.0 in
      // contact; one call per frame the same).
      const bool rebuilt = bd->mode == CC_BODY_REBUILD;
      int lanes = rebuilt ? 2 : 3;
      if (const char* e = std::getenv("CC_BODY_LANES")) lanes = std::atoi(e);
      CC_REQUIRE(lanes >= 1 && lanes <= MAX_LANES, "cc_sim_create: CC_BODY_LANES must be 1..4");
      if (const char* e = std::getenv("CC_GRAPH_STEPS")) S.graph_steps = std::atoi(e);
      CC_REQUIRE(S.graph_steps >= 2 && S.graph_steps % 2 == 0 && 2 * S.graph_steps <= MAX_SLOTS,
                 "cc_sim_create: CC_GRAPH_STEPS must be even, 2..MAX_SLOTS/2");
      int chain = 3;
      if (const char* e = std::getenv("CC_BUILD_FRAMES")) chain = std::atoi(e);
      CC_REQUIRE(chain >= 1 && chain <= MAX_FRAMES, "cc_sim_create: CC_BUILD_FRAMES must be 1..4");
      // 0: rebuilt bodies of any size take the Karras searches and the
      // refit (the path of bodies beyond 2^21 triangles)
      bool bottom_up = true;
      if (const char* e = std::getenv("CC_REBUILD_BOTTOM_UP")) bottom_up = std::atoi(e) != 0;
      S.body.init(bv.data(), bd->nv, bd->tris, bd->nf, bd->mode, params->box_depth, bd->caps, mp,
                  2 * S.graph_steps, lanes, s, chain, bottom_up);
      // frame 0 into slot 0; a refitted body keeps the topology built from
      // the frame-0 pose in every slot
      S.body.deform(0, 0, s);
      S.body.build(0, 0, s);
      S.body.share_topology(0, s);
      S.ready = 1;
      if (S.body.moving()) {
        // Plain events: an event created with hipEventDisableTiming was seen
        // to complete only with the command after it on its stream, which
        // held the lanes back until the step graph had finished.
        for (int l = 0; l < S.body.nlanes; ++l) {
          CC_HIP(hipStreamCreateWithFlags(&S.lane_s[l], hipStreamNonBlocking));
          CC_HIP(hipEventCreate(&S.ev_done[l]));
        }
        CC_HIP(hipEventCreate(&S.ev_go));
        for (int k = 0; k < S.body.nslots; ++k) CC_HIP(hipEventCreate(&S.ev_pose[k]));
        // single steps issue poses up to four frames per lane ahead
        // (measured at the 501k-triangle refitted body with two-frame
        // chains: 32.9 us per single step against 33.6 two frames per lane
        // ahead and 33.0 six)
        S.lookahead = 4 * S.body.nlanes;
        if (const char* e = std::getenv("CC_LOOKAHEAD")) S.lookahead = std::atoi(e);
        // single steps of a chained body issue up to nlanes * chain - 1
        // frames past the lookahead
        CC_REQUIRE(S.lookahead >= 1 &&
                       S.lookahead + S.body.nlanes * S.body.chain + Sim::GO_EVERY < S.body.nslots,
                   "cc_sim_create: CC_LOOKAHEAD too large for the slot ring");
      }
    }
    CC_HIP(hipStreamSynchronize(s));
    *out = h.release();
  });
}

int cc_fault_inject(int kind, int64_t count) {
  return guard([&] {
    CC_REQUIRE(kind == CC_FAULT_ALLOC || kind == CC_FAULT_LAUNCH, "cc_fault_inject: unknown kind");
    g_fault[kind == CC_FAULT_ALLOC ? FAULT_ALLOC : FAULT_LAUNCH].store(count < 0 ? -1 : count);
  });
}

int cc_sim_destroy(cc_sim* sim) {
  return guard([&] {
    if (!sim) return;
    (void)hipStreamSynchronize(sim->s.stream);
    delete sim;
  });
}

int cc_sim_step(cc_sim* sim, int nsteps, void* stream) {
  return guard([&] {
    CC_REQUIRE(sim && nsteps >= 0, "cc_sim_step: bad argument");
    sim->s.step(nsteps, (hipStream_t)stream);
  });
}

int cc_sim_synchronize(cc_sim* sim) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_sim_synchronize: null");
    CC_HIP(hipStreamSynchronize(sim->s.stream));
    CC_HIP(hipDeviceSynchronize());
  });
}

int cc_sim_get_state(cc_sim* sim, float* x_cur, float* x_prev, float* contact_n) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_sim_get_state: null");
    Sim& S = sim->s;
    CC_HIP(hipDeviceSynchronize());
    std::vector<float> packed(6 * (size_t)S.n);
    copy_on(packed.data(), S.state[S.cur].p, packed.size() * 4, hipMemcpyDeviceToHost, sim->s.stream);
    for (uint32_t i = 0; i < S.n; ++i)
      for (int c = 0; c < 3; ++c) {
        if (x_cur) x_cur[3 * (size_t)i + c] = packed[6 * (size_t)i + c];
        if (x_prev) x_prev[3 * (size_t)i + c] = packed[6 * (size_t)i + 3 + c];
      }
    if (contact_n) copy_on(contact_n, S.contact.p, 3 * (size_t)S.n * 4, hipMemcpyDeviceToHost, sim->s.stream);
  });
}

int cc_sim_set_state(cc_sim* sim, const float* x_cur, const float* x_prev, const float* contact_n) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_sim_set_state: null");
    Sim& S = sim->s;
    CC_HIP(hipDeviceSynchronize());
    std::vector<float> packed(6 * (size_t)S.n);
    copy_on(packed.data(), S.state[S.cur].p, packed.size() * 4, hipMemcpyDeviceToHost, sim->s.stream);
    for (uint32_t i = 0; i < S.n; ++i)
      for (int c = 0; c < 3; ++c) {
        if (x_cur) packed[6 * (size_t)i + c] = x_cur[3 * (size_t)i + c];
        if (x_prev) packed[6 * (size_t)i + 3 + c] = x_prev[3 * (size_t)i + c];
      }
    copy_on(S.state[S.cur].p, packed.data(), packed.size() * 4, hipMemcpyHostToDevice, sim->s.stream);
    if (contact_n) copy_on(S.contact.p, contact_n, 3 * (size_t)S.n * 4, hipMemcpyHostToDevice, sim->s.stream);
  });
}

// Checkpoint file: header, pinned vertex indices, the state as stored
// ({x_cur, x_last} per vertex), the contact normals.
namespace {
constexpr char CKPT_MAGIC[8] = {'C', 'C', 'S', 'I', 'M', 'C', 'K', 'P'};
constexpr uint32_t CKPT_VERSION = 1;
struct CkptHeader {
  char magic[8];
  uint32_t version, n, ntypes, npins;
  uint64_t frame;
  cc_params prm;
};
struct File {
  FILE* f = nullptr;
  File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};
}  // namespace

int cc_sim_save(cc_sim* sim, const char* path) {
  return guard([&] {
    CC_REQUIRE(sim && path, "cc_sim_save: null argument");
    Sim& S = sim->s;
    CC_HIP(hipDeviceSynchronize());
    std::vector<uint32_t> pins;
    for (uint32_t i = 0; i < S.n; ++i)
      if (S.h_row[i] & PIN_BIT) pins.push_back(i);
    std::vector<float> packed(6 * (size_t)S.n), contact(3 * (size_t)S.n);
    copy_on(packed.data(), S.state[S.cur].p, packed.size() * 4, hipMemcpyDeviceToHost, S.stream);
    copy_on(contact.data(), S.contact.p, contact.size() * 4, hipMemcpyDeviceToHost, S.stream);
    CkptHeader h{};
    std::memcpy(h.magic, CKPT_MAGIC, sizeof(h.magic));
    h.version = CKPT_VERSION;
    h.n = S.n;
    h.ntypes = S.ntypes;
    h.npins = (uint32_t)pins.size();
    h.frame = S.frame;
    h.prm = S.prm;
    File f(path, "wb");
    CC_REQUIRE(f.f, std::string("cc_sim_save: cannot open ") + path);
    bool ok = std::fwrite(&h, sizeof(h), 1, f.f) == 1;
    ok = ok && (pins.empty() || std::fwrite(pins.data(), 4, pins.size(), f.f) == pins.size());
    ok = ok && std::fwrite(packed.data(), 4, packed.size(), f.f) == packed.size();
    ok = ok && std::fwrite(contact.data(), 4, contact.size(), f.f) == contact.size();
    ok = ok && std::fclose(f.f) == 0;
    f.f = nullptr;
    CC_REQUIRE(ok, std::string("cc_sim_save: write failed: ") + path);
  });
}

int cc_sim_load(cc_sim* sim, const char* path) {
  return guard([&] {
    CC_REQUIRE(sim && path, "cc_sim_load: null argument");
    Sim& S = sim->s;
    File f(path, "rb");
    CC_REQUIRE(f.f, std::string("cc_sim_load: cannot open ") + path);
    CkptHeader h{};
    CC_REQUIRE(std::fread(&h, sizeof(h), 1, f.f) == 1 && std::memcmp(h.magic, CKPT_MAGIC, 8) == 0 &&
                   h.version == CKPT_VERSION,
               "cc_sim_load: not a cloth checkpoint of this version");
    CC_REQUIRE(h.n == S.n && h.ntypes == S.ntypes,
               "cc_sim_load: the checkpoint is of another cloth (vertex count or spring types differ)");
    CC_REQUIRE(h.npins <= h.n && h.prm.dt > 0.0f && h.prm.mass > 0.0f, "cc_sim_load: corrupt checkpoint");
    std::vector<uint32_t> pins(h.npins);
    std::vector<float> packed(6 * (size_t)S.n), contact(3 * (size_t)S.n);
    bool ok = pins.empty() || std::fread(pins.data(), 4, pins.size(), f.f) == pins.size();
    ok = ok && std::fread(packed.data(), 4, packed.size(), f.f) == packed.size();
    ok = ok && std::fread(contact.data(), 4, contact.size(), f.f) == contact.size();
    CC_REQUIRE(ok, "cc_sim_load: truncated checkpoint");
    for (uint32_t q : pins) CC_REQUIRE(q < S.n, "cc_sim_load: corrupt checkpoint (pin index)");
    CC_HIP(hipDeviceSynchronize());
    // parameters as cc_sim_set_params applies them, then pins, state, frame
    S.prm = h.prm;
    S.body.depth = h.prm.box_depth;
    S.body.motion.dt = (double)h.prm.dt;
    S.drop_graphs();
    for (uint32_t i = 0; i < S.n; ++i) S.h_row[i] &= ROW_MASK;
    for (uint32_t q : pins) S.h_row[q] |= PIN_BIT;
    copy_on(S.row.p, S.h_row.data(), S.h_row.size() * 4, hipMemcpyHostToDevice, S.stream);
    copy_on(S.state[S.cur].p, packed.data(), packed.size() * 4, hipMemcpyHostToDevice, S.stream);
    copy_on(S.contact.p, contact.data(), contact.size() * 4, hipMemcpyHostToDevice, S.stream);
    S.frame = h.frame;
    S.ready = h.frame;
    S.refresh_body();
  });
}

int cc_sim_frame(cc_sim* sim, uint64_t* frame) {
  return guard([&] {
    CC_REQUIRE(sim && frame, "cc_sim_frame: null");
    *frame = sim->s.frame;
  });
}

int cc_sim_block_steps(cc_sim* sim, int* steps) {
  return guard([&] {
    CC_REQUIRE(sim && steps, "cc_sim_block_steps: null");
    *steps = sim->s.graph_steps;
  });
}

int cc_sim_set_frame(cc_sim* sim, uint64_t frame) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_sim_set_frame: null");
    CC_HIP(hipDeviceSynchronize());
    sim->s.frame = frame;
    sim->s.ready = frame;
    sim->s.refresh_body();
  });
}

int cc_sim_update_vertex(cc_sim* sim, uint32_t idx, const float xyz[3]) {
  return guard([&] {
    CC_REQUIRE(sim && xyz && idx < sim->s.n, "cc_sim_update_vertex: bad argument");
    Sim& S = sim->s;
    CC_HIP(hipDeviceSynchronize());
    float rec6[6] = {xyz[0], xyz[1], xyz[2], xyz[0], xyz[1], xyz[2]};
    copy_on(S.state[S.cur].p + 6 * (size_t)idx, rec6, sizeof(rec6), hipMemcpyHostToDevice, sim->s.stream);
  });
}

int cc_sim_set_pins(cc_sim* sim, const uint32_t* idx, uint32_t k) {
  return guard([&] {
    CC_REQUIRE(sim && (idx || k == 0), "cc_sim_set_pins: bad argument");
    Sim& S = sim->s;
    for (uint32_t q = 0; q < k; ++q) CC_REQUIRE(idx[q] < S.n, "cc_sim_set_pins: index out of range");
    CC_HIP(hipDeviceSynchronize());
    for (uint32_t i = 0; i < S.n; ++i) S.h_row[i] &= ROW_MASK;
    std::vector<float> rec6(6);
    for (uint32_t q = 0; q < k; ++q) {
      S.h_row[idx[q]] |= PIN_BIT;
      float* dst = S.state[S.cur].p + 6 * (size_t)idx[q];
      copy_on(rec6.data(), dst, 24, hipMemcpyDeviceToHost, sim->s.stream);
      for (int c = 0; c < 3; ++c) rec6[3 + c] = rec6[c];
      copy_on(dst, rec6.data(), 24, hipMemcpyHostToDevice, sim->s.stream);
    }
    copy_on(S.row.p, S.h_row.data(), S.h_row.size() * 4, hipMemcpyHostToDevice, sim->s.stream);
  });
}

int cc_sim_set_params(cc_sim* sim, const cc_params* p) {
  return guard([&] {
    CC_REQUIRE(sim && p && p->dt > 0.0f && p->mass > 0.0f, "cc_sim_set_params: bad argument");
    CC_HIP(hipDeviceSynchronize());
    sim->s.prm = *p;
    sim->s.body.depth = p->box_depth;
    sim->s.body.motion.dt = (double)p->dt;
    sim->s.refresh_body();
    sim->s.drop_graphs();
  });
}

int cc_sim_set_debug(cc_sim* sim, int enable) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_sim_set_debug: null");
    Sim& S = sim->s;
    CC_HIP(hipDeviceSynchronize());
    S.debug = enable != 0;
    if (S.debug) {
      S.dbg_pre.alloc(3 * (size_t)S.n);
      S.dbg_hit.alloc(S.n);
      S.dbg_in.alloc(S.n);
    }
    S.drop_graphs();
  });
}

int cc_sim_get_debug(cc_sim* sim, float* pre3, int32_t* hit, uint8_t* inside) {
  return guard([&] {
    CC_REQUIRE(sim && sim->s.debug, "cc_sim_get_debug: debug outputs not enabled");
    Sim& S = sim->s;
    CC_HIP(hipDeviceSynchronize());
    if (pre3) copy_on(pre3, S.dbg_pre.p, 3 * (size_t)S.n * 4, hipMemcpyDeviceToHost, sim->s.stream);
    if (hit) copy_on(hit, S.dbg_hit.p, (size_t)S.n * 4, hipMemcpyDeviceToHost, sim->s.stream);
    if (inside) copy_on(inside, S.dbg_in.p, S.n, hipMemcpyDeviceToHost, sim->s.stream);
  });
}

int cc_sim_set_counting(cc_sim* sim, int enable) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_sim_set_counting: null");
    CC_HIP(hipDeviceSynchronize());
    sim->s.counting = enable != 0;
    fill_on(sim->s.ctr.p, 0, sim->s.ctr.n * sizeof(unsigned long long), sim->s.stream);
    sim->s.drop_graphs();
  });
}

int cc_sim_get_counters(cc_sim* sim, uint64_t* nodes_visited, uint64_t* queries) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_sim_get_counters: null");
    CC_HIP(hipDeviceSynchronize());
    std::vector<unsigned long long> c(sim->s.ctr.n);
    copy_on(c.data(), sim->s.ctr.p, c.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, sim->s.stream);
    unsigned long long v = 0, q = 0;
    for (size_t b = 0; b + 1 < c.size(); b += 2) {
      v += c[b];
      q += c[b + 1];
    }
    if (nodes_visited) *nodes_visited = v;
    if (queries) *queries = q;
  });
}

int cc_sim_check_finite(cc_sim* sim) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_sim_check_finite: null");
    Sim& S = sim->s;
    CC_HIP(hipDeviceSynchronize());
    check_state(S.state[S.cur].p, S.n, S.spill, S.n, S.flag.p, S.stream, "cc_sim_check_finite");
  });
}

int cc_sim_set_stack_limit(cc_sim* sim, int limit) {
  return guard([&] {
    CC_REQUIRE(sim && limit >= 1 && limit <= STACK_LIMIT, "cc_sim_set_stack_limit: limit must be 1..64");
    Sim& S = sim->s;
    CC_HIP(hipDeviceSynchronize());
    const int32_t v = limit;
    copy_on(S.spill.p + (size_t)STACK_SPILL * S.n + 1, &v, 4, hipMemcpyHostToDevice, sim->s.stream);
  });
}

int cc_body_set_vertices(cc_sim* sim, const float* verts3) {
  return guard([&] {
    CC_REQUIRE(sim && verts3, "cc_body_set_vertices: bad argument");
    Body& B = sim->s.body;
    CC_REQUIRE(B.mode != CC_BODY_NONE, "cc_body_set_vertices: no body");
    CC_HIP(hipDeviceSynchronize());
    Sim& S = sim->s;
    const int k = B.slot_of(S.frame);
    copy_on(B.rest.p, verts3, 3 * (size_t)B.nv * 4, hipMemcpyHostToDevice, sim->s.stream);
    copy_on(B.slot[k].verts.p, verts3, 3 * (size_t)B.nv * 4, hipMemcpyHostToDevice, sim->s.stream);
    // new rest pose: the topology from it as given, then the current frame's pose
    B.build(k, 0, S.stream);
    B.share_topology(k, S.stream);
    S.refresh_body();
  });
}

int cc_body_get_vertices(cc_sim* sim, float* verts3) {
  return guard([&] {
    CC_REQUIRE(sim && verts3, "cc_body_get_vertices: bad argument");
    Body& B = sim->s.body;
    CC_REQUIRE(B.mode != CC_BODY_NONE, "cc_body_get_vertices: no body");
    CC_HIP(hipDeviceSynchronize());
    copy_on(verts3, B.slot[B.slot_of(sim->s.frame)].verts.p, 3 * (size_t)B.nv * 4, hipMemcpyDeviceToHost, sim->s.stream);
  });
}

int cc_body_get_motion(cc_sim* sim, float* R, float* T, float* breath) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_body_get_motion: null");
    Body& B = sim->s.body;
    CC_REQUIRE(B.ncaps > 0, "cc_body_get_motion: body has no motion");
    CC_HIP(hipDeviceSynchronize());
    std::vector<float> m(12 + B.ncaps);
    copy_on(m.data(), B.slot[B.slot_of(sim->s.frame)].motion.p, m.size() * 4, hipMemcpyDeviceToHost, sim->s.stream);
    if (R) std::memcpy(R, m.data(), 36);
    if (T) std::memcpy(T, m.data() + 9, 12);
    if (breath) std::memcpy(breath, m.data() + 12, 4 * (size_t)B.ncaps);
  });
}

int cc_bvh_rebuild(cc_sim* sim) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_bvh_rebuild: null");
    CC_HIP(hipDeviceSynchronize());
    Sim& S = sim->s;
    const int k = S.body.slot_of(S.frame);
    S.body.build(k, 0, S.stream);
    S.body.share_topology(k, S.stream);
    S.ready = S.frame + 1;    // poses computed ahead keep the old boxes
    CC_HIP(hipStreamSynchronize(S.stream));
  });
}

int cc_bvh_refit(cc_sim* sim) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_bvh_refit: null");
    CC_HIP(hipDeviceSynchronize());
    sim->s.body.refit(sim->s.body.slot_of(sim->s.frame), 0, sim->s.stream);
    CC_HIP(hipStreamSynchronize(sim->s.stream));
  });
}

int cc_sim_set_bvh_width(cc_sim* sim, int width) {
  return guard([&] {
    CC_REQUIRE(sim && (width == 2 || width == 4), "cc_sim_set_bvh_width: width must be 2 or 4");
    CC_HIP(hipDeviceSynchronize());
    Sim& S = sim->s;
    S.body.wide_enabled = width == 4;
    S.drop_graphs();
    if (S.body.mode != CC_BODY_NONE) S.refresh_body();    // fold the current frame's boxes
  });
}

int cc_bvh_width(cc_sim* sim, int* width, uint32_t* wide_nodes, uint32_t* stack_bound) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_bvh_width: null");
    const Body& B = sim->s.body;
    const bool wide = B.nwide && B.wide_enabled;
    if (width) *width = wide ? 4 : 2;
    if (wide_nodes) *wide_nodes = B.nwide;
    if (stack_bound) *stack_bound = B.wide_stack;
  });
}

int cc_bvh_size(cc_sim* sim, uint32_t* leaves) {
  return guard([&] {
    CC_REQUIRE(sim && leaves, "cc_bvh_size: bad argument");
    *leaves = sim->s.body.mode == CC_BODY_NONE ? 0 : sim->s.body.m;
  });
}

int cc_bvh_export(cc_sim* sim, uint64_t* keys, uint32_t* order, int32_t* childA, int32_t* childB,
                  int32_t* parent_int, int32_t* parent_leaf, float* box_int, float* box_leaf, float* root6) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_bvh_export: null");
    CC_HIP(hipDeviceSynchronize());
    Sim& S = sim->s;
    const int k = S.body.slot_of(S.frame);
    // a rebuilt body's keys and leaf parents live in lane scratch that later
    // frames reuse: build this slot's tree again (same input, same tree)
    if (S.body.mode == CC_BODY_REBUILD) S.body.build(k, 0, S.stream);
    // a body walked through its four-wide form keeps no binary boxes
    else if (S.body.mode != CC_BODY_NONE) S.body.refit(k, 0, S.stream, true);
    sim->s.body.export_tree(k, keys, order, childA, childB, parent_int, parent_leaf, box_int, box_leaf, root6,
                            sim->s.stream);
  });
}

int cc_collide_points(cc_sim* sim, const float* pts3, uint32_t n, int32_t* hit, uint8_t* inside, float* out_pos,
                      float* out_n) {
  return guard([&] {
    CC_REQUIRE(sim && pts3 && hit && inside && out_pos && out_n, "cc_collide_points: null argument");
    if (n == 0) return;
    Sim& S = sim->s;
    hipStream_t s = S.stream;
    CC_HIP(hipDeviceSynchronize());
    DBuf<float> p, op, on;
    DBuf<int32_t> h;
    DBuf<uint8_t> in;
    DBuf<int32_t> sp;
    p.upload(pts3, 3 * (size_t)n, s);
    alloc_spill(sp, n, s);
    op.alloc(3 * (size_t)n);
    on.alloc(3 * (size_t)n);
    h.alloc(n);
    in.alloc(n);
    launch_collide_points(S.body.view(S.body.slot_of(S.frame)), p.p, n, sp.p, h.p, in.p, op.p, on.p, s);
    CC_HIP(hipMemcpyAsync(hit, h.p, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
    CC_HIP(hipMemcpyAsync(inside, in.p, n, hipMemcpyDeviceToHost, s));
    CC_HIP(hipMemcpyAsync(out_pos, op.p, 12 * (size_t)n, hipMemcpyDeviceToHost, s));
    CC_HIP(hipMemcpyAsync(out_n, on.p, 12 * (size_t)n, hipMemcpyDeviceToHost, s));
    uint32_t fault = 0;
    CC_HIP(hipMemcpyAsync(&fault, sp.p + (size_t)STACK_SPILL * n, 4, hipMemcpyDeviceToHost, s));
    CC_HIP(hipStreamSynchronize(s));
    if (fault) throw Error(CC_ERR_TRAVERSAL, "cc_collide_points: traversal fault (stack overflow or visit bound)");
  });
}

// Simulator::get_vertex_adjface (simulation.cu:152-180)
int cc_sim_pick(cc_sim* sim, const float origin[3], const float dir[3], int32_t* face, int32_t* vertex, float* t) {
  return guard([&] {
    CC_REQUIRE(sim && origin && dir, "cc_sim_pick: bad argument");
    Sim& S = sim->s;
    CC_REQUIRE(S.nf > 0, "cc_sim_pick: no faces (cc_sim_set_faces)");
    CC_HIP(hipDeviceSynchronize());
    DBuf<unsigned long long> res;
    res.alloc(1);
    CC_HIP(hipMemsetAsync(res.p, 0xFF, sizeof(unsigned long long), S.stream));
    const uint32_t g = std::min<uint32_t>(grid_for(S.nf, 256), 1024);
    k_pick<<<g, 256, 0, S.stream>>>(S.state[S.cur].p, S.faces.p, S.nf, ld3(origin), ld3(dir), res.p);
    CC_LAUNCHED();
    unsigned long long key = 0;
    CC_HIP(hipMemcpyAsync(&key, res.p, sizeof(key), hipMemcpyDeviceToHost, S.stream));
    CC_HIP(hipStreamSynchronize(S.stream));
    int32_t f = -1, v = -1;
    float tt = 0.0f;
    if (key != ~0ull) {
      f = (int32_t)(0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFu));
      const uint32_t ord = (uint32_t)(key >> 32);
      const uint32_t b = (ord & 0x80000000u) ? (ord & 0x7FFFFFFFu) : ~ord;
      std::memcpy(&tt, &b, 4);
      uint32_t v0 = 0;     // OnMouseDown selects the face's first vertex (scene.cpp:431-434)
      copy_on(&v0, S.faces.p + 3 * (size_t)f, 4, hipMemcpyDeviceToHost, sim->s.stream);
      v = (int32_t)v0;
    }
    if (face) *face = f;
    if (vertex) *vertex = v;
    if (t) *t = tt;
  });
}

int cc_sim_set_faces(cc_sim* sim, const uint32_t* f3i, uint32_t nf) {
  return guard([&] {
    CC_REQUIRE(sim && f3i && nf > 0, "cc_sim_set_faces: bad argument");
    Sim& S = sim->s;
    for (size_t i = 0; i < 3 * (size_t)nf; ++i) CC_REQUIRE(f3i[i] < S.n, "cc_sim_set_faces: index out of range");
    std::vector<std::vector<uint32_t>> adj(S.n);
    for (uint32_t f = 0; f < nf; ++f)
      for (int j = 0; j < 3; ++j) adj[f3i[3 * (size_t)f + j]].push_back(f);
    std::vector<uint32_t> r, c;
    r.reserve(S.n + 1);
    for (auto& a : adj) {
      r.push_back((uint32_t)c.size());
      c.insert(c.end(), a.begin(), a.end());
    }
    r.push_back((uint32_t)c.size());
    CC_HIP(hipDeviceSynchronize());
    S.nf = nf;
    S.faces.upload(f3i, 3 * (size_t)nf, S.stream);
    S.vrow.upload(r.data(), r.size(), S.stream);
    S.vcol.upload(c.data(), c.size(), S.stream);
    S.face_n.alloc(3 * (size_t)nf);
    S.vnormals.alloc(3 * (size_t)S.n);
    CC_HIP(hipStreamSynchronize(S.stream));
  });
}

int cc_sim_write_vbo(cc_sim* sim, void* dst_dev, size_t bytes, uint32_t n_tex, int lag) {
  return guard([&] {
    CC_REQUIRE(sim && dst_dev, "cc_sim_write_vbo: null");
    Sim& S = sim->s;
    CC_REQUIRE(S.nf > 0, "cc_sim_write_vbo: call cc_sim_set_faces first");
    const size_t need = 16 * (size_t)S.n + 8 * (size_t)n_tex + 12 * (size_t)S.n;
    CC_REQUIRE(bytes >= need, "cc_sim_write_vbo: buffer too small");
    float* base = (float*)dst_dev;
    float* nrm = base + 4 * (size_t)S.n + 2 * (size_t)n_tex;   // simulation.cu:210
    // positions are x_cur_out of the last step == state[cur]; the reference
    // computes face normals from x_cur_in (the step before), kept by lag=1
    const float* pos_state = S.state[S.cur].p;
    const float* fn_state = lag ? S.state[S.cur ^ 1].p : pos_state;
    k_face_normals<<<grid_for(S.nf, 256), 256, 0, S.stream>>>(fn_state, S.faces.p, S.nf, S.face_n.p);
    k_vbo<<<grid_for(S.n, 256), 256, 0, S.stream>>>(pos_state, S.vrow.p, S.vcol.p, S.face_n.p, S.n, base, nrm,
                                                    S.vnormals.p);
    CC_LAUNCHED();
    CC_HIP(hipStreamSynchronize(S.stream));
  });
}

int cc_sim_get_normals(cc_sim* sim, float* normals3) {
  return guard([&] {
    CC_REQUIRE(sim && normals3 && sim->s.nf > 0, "cc_sim_get_normals: bad argument");
    CC_HIP(hipDeviceSynchronize());
    copy_on(normals3, sim->s.vnormals.p, 12 * (size_t)sim->s.n, hipMemcpyDeviceToHost, sim->s.stream);
  });
}

int cc_sim_time_kernels(cc_sim* sim, int nsteps, cc_timing* out) {
  return guard([&] {
    CC_REQUIRE(sim && out && nsteps > 0, "cc_sim_time_kernels: bad argument");
    Sim& S = sim->s;
    hipStream_t s = S.stream;
    CC_HIP(hipDeviceSynchronize());
    std::memset(out, 0, sizeof(*out));
    const Events<5> ev;
    double t_body = 0, t_upd = 0, t_ver = 0;
    // the work of a step one kernel group at a time: the body update of the
    // next frame, then the step of this frame
    S.reset_pipeline();
    S.ready = S.frame + 1;
    for (int it = 0; it < nsteps; ++it) {
      const int nk = S.body.slot_of(S.frame + 1);
      CC_HIP(hipEventRecord(ev[0], s));
      if (S.body.moving()) S.body.deform(nk, S.frame + 1, s);
      CC_HIP(hipEventRecord(ev[1], s));
      if (S.body.mode == CC_BODY_REFIT) S.body.refit(nk, 0, s);
      else if (S.body.mode == CC_BODY_REBUILD) S.body.build(nk, 0, s);
      CC_HIP(hipEventRecord(ev[2], s));
      S.launch_verlet(S.body.slot_of(S.frame), s);
      CC_HIP(hipEventRecord(ev[3], s));
      CC_HIP(hipEventSynchronize(ev[3]));
      float a, b, c;
      CC_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
      CC_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
      CC_HIP(hipEventElapsedTime(&c, ev[2], ev[3]));
      t_body += a;
      t_upd += b;
      t_ver += c;
      S.cur ^= 1;
      ++S.frame;
      S.ready = S.frame + 1;
    }
    // whole steps from the captured graphs
    CC_HIP(hipEventRecord(ev[0], s));
    S.step(nsteps, s);
    CC_HIP(hipEventRecord(ev[4], s));
    CC_HIP(hipEventSynchronize(ev[4]));
    float w;
    CC_HIP(hipEventElapsedTime(&w, ev[0], ev[4]));
    out->steps = (uint32_t)nsteps;
    out->body_us = 1e3 * t_body / nsteps;
    if (S.body.mode == CC_BODY_REFIT) out->refit_us = 1e3 * t_upd / nsteps;
    else out->build_us = 1e3 * t_upd / nsteps;
    out->verlet_us = 1e3 * t_ver / nsteps;
    out->step_us = 1e3 * w / nsteps;
  });
}

int cc_sim_set_timing(cc_sim* sim, int enable) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_sim_set_timing: null");
    Sim& S = sim->s;
    CC_HIP(hipDeviceSynchronize());
    for (int e = 0; e < 2; ++e) S.collect(e);
    S.ring_collect_all();
    S.timing = enable != 0;
    if (S.timing && !S.tev[0][0])
      for (auto& set : S.tev)
        for (hipEvent_t& ev : set) CC_HIP(hipEventCreate(&ev));
    if (S.timing && !S.rev[0][0])
      for (auto& pr : S.rev)
        for (hipEvent_t& ev : pr) CC_HIP(hipEventCreate(&ev));
    S.t_verlet_ms = S.t_update_ms = 0;
    S.t_steps = S.t_updates = 0;
    S.tpending[0] = S.tpending[1] = false;
    S.r_head = S.r_tail = 0;
  });
}

int cc_sim_get_timing(cc_sim* sim, cc_timing* out) {
  return guard([&] {
    CC_REQUIRE(sim && out, "cc_sim_get_timing: null");
    Sim& S = sim->s;
    for (int e = 0; e < 2; ++e) S.collect(e);
    S.ring_collect_all();
    std::memset(out, 0, sizeof(*out));
    out->steps = (uint32_t)S.t_steps;
    if (S.t_steps) out->verlet_us = 1e3 * S.t_verlet_ms / (double)S.t_steps;
    if (S.t_updates) {
      const double u = 1e3 * S.t_update_ms / (double)S.t_updates;
      if (S.body.mode == CC_BODY_REBUILD) out->build_us = u;
      else out->refit_us = u;
    }
  });
}

}  // extern "C"
