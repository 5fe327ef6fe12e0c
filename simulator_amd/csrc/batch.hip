// batch.hip -- many independent cloth / body instances in one launch per step.
//
// Every instance's cloth, springs and body LBVH are concatenated in flat
// device arrays.  A workgroup covers a range of one instance's vertices, so
// the instance's step arguments and LBVH view (a per-instance descriptor)
// are uniform across the workgroup.  One launch advances all instances by
// one step with the same per-vertex code as the single-cloth kernel
// (step_vertex), and 16 steps are replayed from a captured graph.  The
// workgroup size (64, 128 or 256 vertices: one to four wavefronts per row
// segment) and the order of the workgroups (instance order, or largest
// instance first) are chosen at creation by timing each candidate on a copy
// of the initial state.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>

#include "cc_bbody.h"
#include "cc_step.h"

namespace cc {

template <int TILE, int EW>
__global__ __launch_bounds__(TILE) void k_verlet_batch(const float* __restrict__ sin_, float* __restrict__ sout,
                                                       float* __restrict__ contact, const uint32_t* __restrict__ row,
                                                       const uint2* __restrict__ rec,
                                                       const InstDesc* __restrict__ insts,
                                                       const uint4* __restrict__ blocks, uint32_t ntotal,
                                                       int32_t* __restrict__ spill) {
  const uint4 b = blocks[blockIdx.x];     // {instance, first vertex, vertex count, -}
  if (threadIdx.x >= b.z) return;
  const InstDesc d = insts[b.x];
  const uint32_t i = b.y + threadIdx.x;
  uint32_t visits = 0;
  step_vertex<false, EW, false>(i, sin_, sout, contact, contact, row, rec, d.a, d.bvh, spill_of(spill, i, ntotal), visits, nullptr,
                     nullptr, nullptr);
}

constexpr int TILES[3] = {64, 128, 256};
constexpr int GRAPH_STEPS_B = 16;

struct Batch {
  hipStream_t stream = nullptr;
  uint32_t count = 0, ntotal = 0;
  std::vector<uint32_t> v_off;               // [count+1] first vertex of each instance
  std::vector<uint32_t> h_row;
  DBuf<float> state[2], contact;
  DBuf<uint32_t> row;
  DBuf<uint2> rec;
  DBuf<uint2> ell;                           // [ELL_W * ntotal] the records, fixed width (when ell_w)
  int ell_w = 0;
  DBuf<int32_t> spill;
  DBuf<uint32_t> flag;                       // [1] finite-check scratch
  BatchBodies bodies;
  DBuf<InstDesc> insts[2];                   // per body slot (frame g uses slot g % 2 == cur)
  uint64_t frame = 0;
  // the chosen launch shape
  int tile = 256, order = 0;
  DBuf<uint4> blocks;
  uint32_t nblocks = 0;
  // moving bodies: the step runs inside the sort's launch (k_bb_sort_step)
  // with its own tile, tuned on that launch
  int fused_tile = 256;
  DBuf<uint4> fblocks;
  uint32_t nfblocks = 0;
  int cur = 0;
  hipGraphExec_t graphs[2] = {nullptr, nullptr};

  ~Batch() {
    if (stream) (void)hipStreamSynchronize(stream);   // before the buffers are freed
    for (hipGraphExec_t& g : graphs)
      if (g) (void)hipGraphExecDestroy(g);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // the workgroups for a tile size and an order (0: instance order,
  // 1: instances by decreasing vertex count)
  std::vector<uint4> block_table(int T, int ord) const {
    std::vector<uint32_t> inst(count);
    std::iota(inst.begin(), inst.end(), 0u);
    if (ord == 1)
      std::stable_sort(inst.begin(), inst.end(), [&](uint32_t a, uint32_t b) {
        return v_off[a + 1] - v_off[a] > v_off[b + 1] - v_off[b];
      });
    std::vector<uint4> t;
    for (uint32_t k : inst)
      for (uint32_t v = v_off[k]; v < v_off[k + 1]; v += (uint32_t)T)
        t.push_back(make_uint4(k, v, std::min<uint32_t>((uint32_t)T, v_off[k + 1] - v), 0u));
    return t;
  }

  template <int T>
  void launch_t(const uint4* blk, uint32_t nblk, hipStream_t s) {
    const float* in = state[cur].p;
    float* out = state[cur ^ 1].p;
    if (ell_w)
      k_verlet_batch<T, ELL_W><<<nblk, T, 0, s>>>(in, out, contact.p, row.p, ell.p, insts[cur].p, blk, ntotal,
                                                  spill.p);
    else
      k_verlet_batch<T, 0><<<nblk, T, 0, s>>>(in, out, contact.p, row.p, rec.p, insts[cur].p, blk, ntotal,
                                              spill.p);
    CC_LAUNCHED();
  }
  void launch(int T, const uint4* blk, uint32_t nblk, hipStream_t s) {
    if (T == 64)
      launch_t<64>(blk, nblk, s);
    else if (T == 128)
      launch_t<128>(blk, nblk, s);
    else
      launch_t<256>(blk, nblk, s);
  }

  StepLaunch fused_launch(const uint4* blk, uint32_t nblk, int T) const {
    return StepLaunch{state[cur].p, state[cur ^ 1].p, contact.p, row.p, ell_w ? ell.p : rec.p, insts[cur].p,
                      blk, nblk, ntotal, spill.p, (uint32_t)ell_w, (uint32_t)T};
  }

  // Time every (tile, order) candidate over a few steps from the initial
  // state on the launch the steps will use -- the plain step kernel, or,
  // with moving bodies, the sort's launch that carries the step -- keep the
  // fastest, and restore the initial state.  (The sort in the timed fused
  // launches re-sorts the set-up's boxes into the next slot, which every
  // step rebuilds before reading.)
  void autotune() {
    const size_t n6 = 6 * (size_t)ntotal, n3 = 3 * (size_t)ntotal;
    DBuf<float> s0, c0;
    s0.alloc(n6);
    c0.alloc(n3);
    CC_HIP(hipMemcpyAsync(s0.p, state[0].p, n6 * sizeof(float), hipMemcpyDeviceToDevice, stream));
    CC_HIP(hipMemcpyAsync(c0.p, contact.p, n3 * sizeof(float), hipMemcpyDeviceToDevice, stream));
    const Events<2> ev;
    const hipEvent_t e0 = ev[0], e1 = ev[1];
    const bool fused = bodies.moving && bodies.step.n_sort;
    float best = 1e30f;
    for (int T : TILES)
      for (int ord = 0; ord < 2; ++ord) {
        const std::vector<uint4> t = block_table(T, ord);
        DBuf<uint4> blk;
        blk.upload(t.data(), t.size(), stream);
        auto run = [&] {
          if (fused)
            bodies.sort_with_step(bodies.step, cur ^ 1, fused_launch(blk.p, (uint32_t)t.size(), T), stream);
          else
            launch(T, blk.p, (uint32_t)t.size(), stream);
          cur ^= 1;
        };
        for (int k = 0; k < 3; ++k) run();     // warm-up
        CC_HIP(hipEventRecord(e0, stream));
        for (int k = 0; k < 10; ++k) run();
        CC_HIP(hipEventRecord(e1, stream));
        CC_HIP(hipEventSynchronize(e1));
        float ms = 0;
        CC_HIP(hipEventElapsedTime(&ms, e0, e1));
        tuned[T == 64 ? 0 : T == 128 ? 1 : 2][ord] = ms / 10.0f;
        if (ms < best) {
          best = ms;
          tile = T;
          order = ord;
        }
      }
    cur = 0;
    CC_HIP(hipMemcpyAsync(state[0].p, s0.p, n6 * sizeof(float), hipMemcpyDeviceToDevice, stream));
    CC_HIP(hipMemcpyAsync(contact.p, c0.p, n3 * sizeof(float), hipMemcpyDeviceToDevice, stream));
    // the set-up's frame-0 bodies again in both slots the timed sorts wrote
    if (fused) {
      bodies.launch(bodies.all, 0, 0, stream);
      bodies.launch(bodies.all, 0, 1, stream);
    }
    const std::vector<uint4> t = block_table(tile, order);
    blocks.upload(t.data(), t.size(), stream);
    nblocks = (uint32_t)t.size();
    fused_tile = tile;
    fblocks.upload(t.data(), t.size(), stream);
    nfblocks = (uint32_t)t.size();
    CC_HIP(hipStreamSynchronize(stream));
  }
  float tuned[3][2] = {};                    // ms per step of each (tile, order) candidate

  // One step of frame `frame` (cur == frame % 2): the moving bodies of frame
  // + 1 are posed, sorted and built into the other slot while the cloths step
  // against the bodies of this frame; k selects frame + 1's motion.
  void one_step(int k, hipStream_t s) {
    const int nxt = cur ^ 1;
    if (bodies.moving) {
      bodies.pre(bodies.step, k, nxt, s);
      if (bodies.step.n_sort) {
        bodies.sort_with_step(bodies.step, nxt, fused_launch(fblocks.p, nfblocks, fused_tile), s);
      } else {
        launch(tile, blocks.p, nblocks, s);
      }
      bodies.post(bodies.step, nxt, s);
    } else {
      launch(tile, blocks.p, nblocks, s);
    }
    cur ^= 1;
  }

  hipGraphExec_t capture() {
    const int cur0 = cur;
    hipGraph_t graph = nullptr;
    CC_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      for (int k = 0; k < GRAPH_STEPS_B; ++k) one_step(k, stream);
    } catch (...) {
      cur = cur0;
      hipGraph_t junk = nullptr;
      (void)hipStreamEndCapture(stream, &junk);
      if (junk) (void)hipGraphDestroy(junk);
      throw;
    }
    cur = cur0;
    CC_HIP(hipStreamEndCapture(stream, &graph));
    hipGraphExec_t g = nullptr;
    hipError_t e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) throw Error(CC_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return g;
  }

  // Step g collides every cloth with its body's pose of frame g (the single
  // simulator's order); that pose was built during step g - 1 (or at set-up).
  void step(int nsteps, hipStream_t s) {
    if (!s) s = stream;
    while (nsteps > 0) {
      if (nsteps >= GRAPH_STEPS_B) {
        hipGraphExec_t& g = graphs[cur];
        if (!g) g = capture();
        bodies.upload_motion(frame + 1, GRAPH_STEPS_B, s);
        CC_HIP(hipGraphLaunch(g, s));
        nsteps -= GRAPH_STEPS_B;             // even: the parity is unchanged
        frame += GRAPH_STEPS_B;
      } else {
        bodies.upload_motion(frame + 1, 1, s);
        one_step(0, s);
        --nsteps;
        ++frame;
      }
    }
  }
};

}  // namespace cc

using namespace cc;

struct cc_batch {
  Batch b;
};

extern "C" {

int cc_batch_create(const cc_cloth_desc* cloths, const cc_body_desc* bodies, const cc_params* params, uint32_t count,
                    cc_batch** out) {
  return guard([&] {
    CC_REQUIRE(cloths && params && out && count >= 1, "cc_batch_create: bad argument");
    *out = nullptr;
    int ndev = 0;
    CC_HIP(hipGetDeviceCount(&ndev));
    CC_REQUIRE(ndev > 0, "cc_batch_create: no HIP device");
    std::unique_ptr<cc_batch> h(new cc_batch);
    Batch& B = h->b;
    CC_HIP(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking));
    hipStream_t s = B.stream;
    B.count = count;
    B.v_off.assign(count + 1, 0);
    for (uint32_t k = 0; k < count; ++k) {
      CC_REQUIRE(cloths[k].n >= 1 && cloths[k].verts4, "cc_batch_create: empty cloth");
      CC_REQUIRE(cloths[k].ntypes >= 1 && cloths[k].ntypes <= 3, "cc_batch_create: ntypes must be 1..3");
      CC_REQUIRE(params[k].dt > 0.0f && params[k].mass > 0.0f, "cc_batch_create: dt and mass must be positive");
      B.v_off[k + 1] = B.v_off[k] + cloths[k].n;
    }
    B.ntotal = B.v_off[count];
    CC_REQUIRE(B.ntotal <= END_MASK, "cc_batch_create: too many vertices");
    // cloth: state, merged springs with global indices
    std::vector<float> packed(6 * (size_t)B.ntotal);
    std::vector<uint2> rec_all;
    B.h_row.assign((size_t)B.ntotal + 1, 0);
    for (uint32_t k = 0; k < count; ++k) {
      std::vector<uint32_t> row;
      std::vector<uint2> rec;
      merge_springs(&cloths[k], row, rec);
      std::vector<float> x3;
      read_vec3_from(cloths[k].verts4, 4, cloths[k].n, x3);
      const uint32_t o = B.v_off[k];
      const uint32_t r0 = (uint32_t)rec_all.size();
      CC_REQUIRE((uint64_t)r0 + rec.size() < ROW_MASK, "cc_batch_create: too many springs");
      for (uint32_t i = 0; i < cloths[k].n; ++i) {
        B.h_row[o + i] = r0 + row[i];
        for (int c = 0; c < 3; ++c) {
          packed[6 * (size_t)(o + i) + c] = x3[3 * (size_t)i + c];
          packed[6 * (size_t)(o + i) + 3 + c] = x3[3 * (size_t)i + c];
        }
      }
      for (uint2 r : rec) rec_all.push_back(make_uint2(((r.x & END_MASK) + o) | (r.x & ~END_MASK), r.y));
    }
    B.h_row[B.ntotal] = (uint32_t)rec_all.size();
    B.state[0].upload(packed.data(), packed.size(), s);
    B.state[1].alloc(packed.size());
    B.state[1].zero(s);
    B.contact.alloc(3 * (size_t)B.ntotal);
    B.contact.zero(s);
    B.row.upload(B.h_row.data(), B.h_row.size(), s);
    if (rec_all.empty()) {
      B.rec.alloc(1);
      B.rec.zero(s);
    } else {
      B.rec.upload(rec_all.data(), rec_all.size(), s);
    }
    // fixed-width copy of the records when every vertex has at most ELL_W
    // springs (CC_SPRINGS_ELL=0 keeps the merged rows only)
    uint32_t max_deg = 0;
    for (uint32_t i = 0; i < B.ntotal; ++i) max_deg = std::max(max_deg, B.h_row[i + 1] - B.h_row[i]);
    bool ell = max_deg <= (uint32_t)ELL_W;
    if (const char* e = std::getenv("CC_SPRINGS_ELL"))
      if (std::atoi(e) == 0) ell = false;
    if (ell) {
      std::vector<uint2> er((size_t)B.ntotal * ELL_W);
      for (uint32_t i = 0; i < B.ntotal; ++i)
        for (uint32_t k = 0; k < (uint32_t)ELL_W; ++k) {
          const uint32_t e = B.h_row[i] + k;
          er[(size_t)i * ELL_W + k] = e < B.h_row[i + 1] ? rec_all[e] : make_uint2(i | PAD_TYPE, 0u);
        }
      B.ell.upload(er.data(), er.size(), s);
      B.ell_w = ELL_W;
    }
    alloc_spill(B.spill, B.ntotal, s);
    // bodies: concatenated, frame 0 built for all of them
    B.bodies.init(bodies, params, count, s);
    std::vector<InstDesc> desc(count);
    for (uint32_t k = 0; k < count; ++k) {
      desc[k].a = step_args(params[k], cloths[k].ntypes);
      desc[k].bvh = BvhView{nullptr, nullptr, nullptr, nullptr, 0};
    }
    for (int slot = 0; slot < 2; ++slot) {
      for (uint32_t j = 0; j < B.bodies.nb; ++j) desc[B.bodies.inst[j]].bvh = B.bodies.view(j, slot);
      B.insts[slot].upload(desc.data(), desc.size(), s);
    }
    CC_HIP(hipStreamSynchronize(s));
    B.autotune();
    *out = h.release();
  });
}

int cc_batch_destroy(cc_batch* b) {
  return guard([&] {
    if (!b) return;
    (void)hipStreamSynchronize(b->b.stream);
    delete b;
  });
}

int cc_batch_set_pins(cc_batch* b, uint32_t inst, const uint32_t* idx, uint32_t k) {
  return guard([&] {
    CC_REQUIRE(b && inst < b->b.count && (idx || k == 0), "cc_batch_set_pins: bad argument");
    Batch& B = b->b;
    const uint32_t o = B.v_off[inst], n = B.v_off[inst + 1] - o;
    for (uint32_t q = 0; q < k; ++q) CC_REQUIRE(idx[q] < n, "cc_batch_set_pins: index out of range");
    CC_HIP(hipDeviceSynchronize());
    for (uint32_t i = o; i < o + n; ++i) B.h_row[i] &= ROW_MASK;
    float rec6[6];
    for (uint32_t q = 0; q < k; ++q) {
      B.h_row[o + idx[q]] |= PIN_BIT;
      float* dst = B.state[B.cur].p + 6 * (size_t)(o + idx[q]);
      copy_on(rec6, dst, 24, hipMemcpyDeviceToHost, b->b.stream);
      for (int c = 0; c < 3; ++c) rec6[3 + c] = rec6[c];
      copy_on(dst, rec6, 24, hipMemcpyHostToDevice, b->b.stream);
    }
    copy_on(B.row.p + o, B.h_row.data() + o, (size_t)n * 4, hipMemcpyHostToDevice, b->b.stream);
  });
}

int cc_batch_step(cc_batch* b, int nsteps, void* stream) {
  return guard([&] {
    CC_REQUIRE(b && nsteps >= 0, "cc_batch_step: bad argument");
    b->b.step(nsteps, (hipStream_t)stream);
  });
}

int cc_batch_synchronize(cc_batch* b) {
  return guard([&] {
    CC_REQUIRE(b, "cc_batch_synchronize: null");
    CC_HIP(hipStreamSynchronize(b->b.stream));
    CC_HIP(hipDeviceSynchronize());
  });
}

int cc_batch_check(cc_batch* b) {
  return guard([&] {
    CC_REQUIRE(b, "cc_batch_check: null");
    Batch& B = b->b;
    CC_HIP(hipDeviceSynchronize());
    if (!B.flag.p) B.flag.alloc(1);
    check_state(B.state[B.cur].p, B.ntotal, B.spill, B.ntotal, B.flag.p, B.stream, "cc_batch_check");
  });
}

int cc_batch_get_state(cc_batch* b, uint32_t inst, float* x_cur, float* x_prev, float* contact_n) {
  return guard([&] {
    CC_REQUIRE(b && inst < b->b.count, "cc_batch_get_state: bad argument");
    Batch& B = b->b;
    CC_HIP(hipDeviceSynchronize());
    const uint32_t o = B.v_off[inst], n = B.v_off[inst + 1] - o;
    std::vector<float> st(6 * (size_t)n), c(3 * (size_t)n);
    copy_on(st.data(), B.state[B.cur].p + 6 * (size_t)o, st.size() * 4, hipMemcpyDeviceToHost, b->b.stream);
    copy_on(c.data(), B.contact.p + 3 * (size_t)o, c.size() * 4, hipMemcpyDeviceToHost, b->b.stream);
    for (uint32_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        if (x_cur) x_cur[3 * (size_t)i + k] = st[6 * (size_t)i + k];
        if (x_prev) x_prev[3 * (size_t)i + k] = st[6 * (size_t)i + 3 + k];
        if (contact_n) contact_n[3 * (size_t)i + k] = c[3 * (size_t)i + k];
      }
  });
}

int cc_batch_body_export(cc_batch* b, uint32_t inst, float* verts3, uint64_t* keys, uint32_t* order,
                         int32_t* childA, int32_t* childB, int32_t* parent_int, int32_t* parent_leaf,
                         float* box_int, float* box_leaf, float* root6) {
  return guard([&] {
    CC_REQUIRE(b && inst < b->b.count, "cc_batch_body_export: bad argument");
    Batch& B = b->b;
    BatchBodies& Y = B.bodies;
    uint32_t j = 0;
    while (j < Y.nb && Y.inst[j] != inst) ++j;
    CC_REQUIRE(j < Y.nb, "cc_batch_body_export: instance has no body");
    CC_HIP(hipDeviceSynchronize());
    Y.export_tree(j, B.cur, verts3, keys, order, childA, childB, parent_int, parent_leaf, box_int, box_leaf, root6,
                  B.stream);            // slot cur holds the current frame
  });
}

int cc_batch_frame(cc_batch* b, uint64_t* frame) {
  return guard([&] {
    CC_REQUIRE(b && frame, "cc_batch_frame: bad argument");
    *frame = b->b.frame;
  });
}

int cc_batch_tuning(cc_batch* b, float ms[6], uint32_t* fused) {
  return guard([&] {
    CC_REQUIRE(b, "cc_batch_tuning: null");
    const Batch& B = b->b;
    if (ms)
      for (int t = 0; t < 3; ++t)
        for (int o = 0; o < 2; ++o) ms[2 * t + o] = B.tuned[t][o];
    if (fused) *fused = (B.bodies.moving && B.bodies.step.n_sort) ? 1u : 0u;
  });
}

int cc_batch_info(cc_batch* b, uint32_t* vertices, uint32_t* tile, uint32_t* order) {
  return guard([&] {
    CC_REQUIRE(b, "cc_batch_info: null");
    if (vertices) *vertices = b->b.ntotal;
    if (tile) *tile = (uint32_t)b->b.tile;
    if (order) *order = (uint32_t)b->b.order;
  });
}

}  // extern "C"
