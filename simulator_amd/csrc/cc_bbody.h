// cc_bbody.h -- the collider bodies of a batch: every instance's body in
// concatenated arrays, posed and rebuilt (or refitted) for all instances
// together, a handful of launches per frame whatever the instance count.
#pragma once

#include <vector>

#include "cc_sim.h"

namespace cc {

constexpr int BB_MF = 12 + MAX_CAPS;   // floats per motion frame and body
constexpr int BB_FRAMES = 16;          // motion frames held on the device (one step graph)

// Work lists of one kind of pass: block tables {body, first, count, aux} of
// the grid-wide kernels, the body list of the per-body sort, and the
// workgroups of each build level.
struct BodyPass {
  DBuf<uint4> deform, boxes;
  DBuf<uint32_t> sort;
  DBuf<LvBlock> level[BB_MAX_LEVELS];
  uint32_t n_deform = 0, n_boxes = 0, n_sort = 0, n_level[BB_MAX_LEVELS] = {};
};

// What one cloth step of a batch reads and writes: per-instance step
// constants and LBVH view, and the blocks of 256 vertices of one instance.
struct InstDesc {
  StepArgs a;
  BvhView bvh;
};

struct StepLaunch {
  const float* in;
  float* out;
  float* contact;
  const uint32_t* row;
  const uint2* rec;
  const InstDesc* insts;
  const uint4* blocks;      // {instance, first vertex, count <= 256, -}
  uint32_t nblocks;
  uint32_t ntotal;
  int32_t* spill;
  uint32_t ew;              // rec holds ew records per vertex (0: merged rows at row)
  uint32_t tile;            // vertices per block entry (64, 128 or 256)
};

struct BatchBodies {
  uint32_t nb = 0;                    // bodies
  std::vector<BodyDesc> h;
  std::vector<uint32_t> inst;         // body -> instance
  std::vector<MotionParams> mp;
  std::vector<std::vector<float>> phase;
  bool moving = false;                // any REFIT / REBUILD body

  DBuf<BodyDesc> desc;
  DBuf<float> rest;                   // [3 * sum nv]
  // Two body slots: frame g lives in slot g % 2, so frame g + 1 is built
  // while the cloth steps against frame g.
  DBuf<float> verts[2];               // [3 * sum nv]
  DBuf<uint32_t> tris;                // [3 * sum m], body-local vertex indices
  DBuf<uint32_t> capv;                // [sum (ncaps + 1)]
  DBuf<float> capp0, capp1;           // [3 * sum ncaps]
  DBuf<float> boxes;                  // [6 * sum m] expanded box of each triangle
  DBuf<float> partial;                // [6 * sum ceil(m / 256)]
  DBuf<uint2> kv[2];                  // [sum m] (code, primitive) sort ping-pong
  DBuf<uint64_t> keys;                // [sum m]
  DBuf<BItem> items[2];               // [level-0 chunks * 128] unfinished nodes between levels
  DBuf<uint32_t> nitems[2];           // [level-0 chunks] their count per chunk
  DBuf<uint4> leaf_tri[2];            // [sum m]
  DBuf<int32_t> leaf_parent[2];       // [sum m]
  DBuf<Node> nodes[2];                // [sum max(m - 1, 1)]
  DBuf<float> root[2], leaf0[2];      // [6 * nb]
  DBuf<float> motion;                 // [BB_FRAMES * nb * BB_MF]
  float* h_motion[2] = {nullptr, nullptr};   // pinned staging of the motion uploads
  hipEvent_t ev_motion[2] = {nullptr, nullptr};
  int hm = 0;

  BodyPass all;    // set-up: every body built from its frame-0 pose
  BodyPass step;   // per frame: moving bodies posed, REBUILD ones re-sorted

  BatchBodies() = default;
  BatchBodies(const BatchBodies&) = delete;
  BatchBodies& operator=(const BatchBodies&) = delete;
  ~BatchBodies();

  // bodies[k] of instance k (NULL or mode NONE: no body); extends the
  // vertices when asked, checks every layout, uploads, builds frame 0.
  void init(const cc_body_desc* bodies, const cc_params* params, uint32_t count, hipStream_t s);
  BvhView view(uint32_t j, int slot) const;
  // motion of frames [frame0, frame0 + nframes) into the device table
  void upload_motion(uint64_t frame0, int nframes, hipStream_t s);
  // One pass into a slot, in three parts: pose + boxes (k selects the
  // motion frame in the table), the per-body sort, the bottom-up build.
  void pre(const BodyPass& p, int k, int slot, hipStream_t s);
  void sort(const BodyPass& p, int slot, hipStream_t s);
  void post(const BodyPass& p, int slot, hipStream_t s);
  void launch(const BodyPass& p, int k, int slot, hipStream_t s) {
    pre(p, k, slot, s);
    sort(p, slot, s);
    post(p, slot, s);
  }
  // The sort of one pass and a cloth step in one launch: the sort's
  // workgroups (largest body first) take the CUs first, the step's blocks
  // fill the CUs that the smaller sorts leave.
  void sort_with_step(const BodyPass& p, int slot, const struct StepLaunch& st, hipStream_t s);
  void export_tree(uint32_t j, int slot, float* verts3, uint64_t* keys, uint32_t* order, int32_t* childA,
                   int32_t* childB, int32_t* parent_int, int32_t* parent_leaf, float* box_int, float* box_leaf,
                   float* root6, hipStream_t s);
};

}  // namespace cc
