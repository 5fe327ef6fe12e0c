// cc_body.h -- the collider body and its Karras LBVH on the device.
#pragma once

#include "cc_internal.h"

namespace cc {

// Per-frame motion of a capsule-union body (see cc_body_desc).
struct MotionDesc {
  double amp = 0, freq = 0, dir[3] = {0, 0, 0}, rot_amp = 0, pivot[3] = {0, 0, 0};
  double breath_amp = 0, breath_freq = 0;
};

struct Body {
  uint32_t nv = 0;      // body vertices
  uint32_t m = 0;       // triangles == LBVH leaves
  uint32_t mode = CC_BODY_NONE;
  float depth = 0.01f;  // expanded box depth (primitive.cu:44)

  DBuf<float> rest;     // [3*nv] rest pose (after vertex_extend)
  DBuf<float> verts;    // [3*nv] current pose
  DBuf<uint32_t> tris;  // [3*m]

  // capsule decomposition for the deformation
  uint32_t ncaps = 0;
  DBuf<uint32_t> cap_vstart;        // [ncaps+1]
  DBuf<float> cap_p0, cap_p1, cap_phase;
  MotionDesc motion;
  DBuf<float> motion_out;           // [12 + ncaps] last float parameters

  // build scratch
  uint32_t nblk_box = 0;
  DBuf<float> boxes;                // [6*m] unsorted expanded boxes
  DBuf<float> partial;              // [6*nblk_box] per-block root boxes
  DBuf<uint32_t> code[2], idx[2];   // radix sort ping-pong
  DBuf<uint32_t> hist;              // [256 * sort blocks]

  // tree
  DBuf<uint64_t> keys;              // [m] sorted morton << 32 | primitive
  DBuf<uint4> leaf_tri;             // [m] sorted {v0, v1, v2, primitive}
  DBuf<int32_t> leaf_parent;        // [m]
  DBuf<Node> nodes;                 // [max(m-1, 1)]
  DBuf<uint32_t> counters;          // [max(m-1, 1)] refit arrival counters
  DBuf<float> root;                 // [6] root box (sentinel U all)
  DBuf<float> leaf0_box;            // [6] box of leaf 0 when m == 1

  BvhView view() const {
    BvhView v;
    v.nodes = nodes.p;
    v.leaf_tri = leaf_tri.p;
    v.verts = verts.p;
    v.leaf0_box = leaf0_box.p;
    v.m = (mode == CC_BODY_NONE) ? 0 : m;
    return v;
  }

  // set-up: host arrays, vertices stride 3 (already extended)
  void init(const float* verts3, uint32_t nv, const uint32_t* tris, uint32_t nf, uint32_t mode,
            float depth, const cc_capsules* caps, const MotionDesc& md, hipStream_t s);

  // launches (asynchronous on s)
  void deform(const uint64_t* frame_dev, double dt, hipStream_t s);  // rest -> verts
  void build(hipStream_t s);                                          // full LBVH build
  void refit(hipStream_t s);                                          // boxes only
  void frame_update(const uint64_t* frame_dev, double dt, hipStream_t s);  // per-mode update

  // host export in the reference encoding
  void export_tree(uint64_t* keys, uint32_t* order, int32_t* childA, int32_t* childB,
                   int32_t* parent_int, int32_t* parent_leaf, float* box_int, float* box_leaf,
                   float* root6, hipStream_t s);
};

}  // namespace cc
