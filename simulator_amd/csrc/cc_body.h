// cc_body.h -- the collider body and its Karras LBVH on the device.
//
// Everything the collision query reads (vertices, sorted leaf triangles,
// nodes, root box) exists twice.  The body update of frame f+1 writes one
// copy while the cloth step of frame f reads the other, so the two can run
// at the same time (see Sim::capture).
#pragma once

#include "cc_internal.h"

namespace cc {

// Per-frame motion of a capsule-union body (see cc_body_desc).
struct MotionDesc {
  double amp = 0, freq = 0, dir[3] = {0, 0, 0}, rot_amp = 0, pivot[3] = {0, 0, 0};
  double breath_amp = 0, breath_freq = 0;
};

// Motion description handed to kernels by value.  The parameters of a frame
// (R row-major, T with the pivot folded in, one breathing factor per capsule)
// are computed once per frame, in double precision with one rounding to
// float, by k_motion from the body's own frame counter.
struct MotionArgs {
  double amp, freq, dx, dy, dz, rot_amp, px, py, pz, breath_amp, breath_freq, dt;
  const float* phase;   // [ncaps]
  float* out;           // [12 + ncaps] parameters of the frame being deformed
  uint64_t* frame;      // body frame counter (read, then advanced, by k_motion)
  uint32_t ncaps;
};

// Threads of one block fill a.out for `frame`.
__device__ inline void compute_motion(const MotionArgs& a, uint64_t frame, uint32_t tid, uint32_t nthreads) {
  const double PI2 = 6.283185307179586476925;
  const double t = (double)frame * a.dt;
  if (tid == 0) {
    const double w = sin(PI2 * a.freq * t);
    const double th = a.rot_amp * w, c = cos(th), s = sin(th);
    const double R[9] = {c, 0.0, s, 0.0, 1.0, 0.0, -s, 0.0, c};   // rotation about +y
    const double P[3] = {a.px, a.py, a.pz};
    const double D[3] = {a.dx, a.dy, a.dz};
    for (int r = 0; r < 3; ++r) {
      const double rp = R[3 * r] * P[0] + R[3 * r + 1] * P[1] + R[3 * r + 2] * P[2];
      a.out[9 + r] = (float)(P[r] - rp + a.amp * w * D[r]);
    }
    for (int q = 0; q < 9; ++q) a.out[q] = (float)R[q];
  }
  for (uint32_t k = tid; k < a.ncaps; k += nthreads)
    a.out[12 + k] = (float)(1.0 + a.breath_amp * sin(PI2 * a.breath_freq * t + (double)a.phase[k]));
}

struct Body {
  uint32_t nv = 0;      // body vertices
  uint32_t m = 0;       // triangles == LBVH leaves
  uint32_t mode = CC_BODY_NONE;
  float depth = 0.01f;  // expanded box depth (primitive.cu:44)
  double dt = 0.0;      // frame time step of the motion (t = frame * dt)

  DBuf<float> rest;     // [3*nv] rest pose (after vertex_extend)
  DBuf<float> verts[2]; // [3*nv] pose of the frame held in buffer b
  DBuf<uint32_t> tris;  // [3*m]

  // capsule decomposition for the deformation
  uint32_t ncaps = 0;
  DBuf<uint32_t> cap_vstart;        // [ncaps+1]
  DBuf<float> cap_p0, cap_p1, cap_phase;
  MotionDesc motion;
  DBuf<float> motion_cur;           // [12 + ncaps] parameters of the last deformation
  DBuf<uint64_t> frame_ctr;         // [1] frame of the next body update

  // build scratch
  uint32_t nblk_box = 0;
  DBuf<float> boxes;                // [6*m] unsorted expanded boxes
  DBuf<float> partial;              // [6*nblk_box] per-block root boxes
  DBuf<uint32_t> code[2], idx[2];   // radix sort ping-pong
  DBuf<uint32_t> hist;              // [256 * sort blocks]

  // tree (per buffer where the query reads it)
  DBuf<uint64_t> keys;              // [m] sorted morton << 32 | primitive
  DBuf<uint4> leaf_tri[2];          // [m] sorted {v0, v1, v2, primitive}
  DBuf<int32_t> leaf_parent;        // [m]
  DBuf<Node> nodes[2];              // [max(m-1, 1)]
  DBuf<float> root[2];              // [6] root box (sentinel U all)
  DBuf<float> leaf0_box[2];         // [6] box of leaf 0 when m == 1
  // refit description: internal nodes whose leaf range crosses a chunk
  // boundary ("top" nodes) and per-node level information
  DBuf<int32_t> top_list;           // [max(m-1, 1)] their indices, sorted by top height
  DBuf<int32_t> top_pos;            // [max(m-1, 1)] node -> position in top_list, -1 if not top
  DBuf<uint32_t> top_count;         // [1]
  DBuf<int4> top_links;             // [max(m-1, 1)] per entry: top child A / B, parent, parent slot
  DBuf<float> top_child;            // [12 * max(m-1, 1)] boxes of the entries' non-top children
  DBuf<float> top_scratch;          // set-up work space and the top pass beyond its LDS size
  DBuf<uint8_t> top_h;              // [max(m-1, 1)] height of each entry in the top tree
  DBuf<uint32_t> top_hmax;          // [1]
  DBuf<uint32_t> top_lev;           // [66] entries of top height h are [top_lev[h], top_lev[h+1])
  DBuf<uint8_t> nflag;              // [max(m-1, 1)] in-chunk height | flags
  DBuf<uint8_t> lflag;              // [m] flags of the leaf
  DBuf<uint32_t> chunk_h;           // [chunks] largest in-chunk height

  bool moving() const { return mode == CC_BODY_REFIT || mode == CC_BODY_REBUILD; }
  // the buffer the query of a frame with ping-pong parity `par` reads
  int buf(int par) const { return moving() ? par : 0; }

  MotionArgs motion_args();
  static size_t top_scratch_floats(uint32_t m);

  BvhView view(int b) const {
    BvhView v;
    v.nodes = nodes[b].p;
    v.leaf_tri = leaf_tri[b].p;
    v.verts = verts[b].p;
    v.leaf0_box = leaf0_box[b].p;
    v.m = (mode == CC_BODY_NONE) ? 0 : m;
    return v;
  }

  // set-up: host arrays, vertices stride 3 (already extended)
  void init(const float* verts3, uint32_t nv, const uint32_t* tris, uint32_t nf, uint32_t mode,
            float depth, const cc_capsules* caps, const MotionDesc& md, hipStream_t s);

  // launches (asynchronous on s) writing buffer b
  void set_frame(uint64_t frame, hipStream_t s);   // frame of the next update
  void deform(int b, hipStream_t s);               // k_motion, then rest -> verts[b]
  void build(int b, hipStream_t s);                // full LBVH build (includes the refit)
  void top_setup(int b, hipStream_t s);            // refit description of nodes[b]'s topology
  void refit(int b, hipStream_t s);                // boxes only
  void update(int b, hipStream_t s);               // per-mode update of one frame
  void mirror_topology(int from, hipStream_t s);   // copy tree and leaves to the other buffer

  // host export in the reference encoding
  void export_tree(int b, uint64_t* keys, uint32_t* order, int32_t* childA, int32_t* childB,
                   int32_t* parent_int, int32_t* parent_leaf, float* box_int, float* box_leaf,
                   float* root6, hipStream_t s);
};

}  // namespace cc
