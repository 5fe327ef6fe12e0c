// cc_body.h -- the collider body and its Karras LBVH on the device.
//
// A static body has one copy of what the collision query reads (vertices,
// sorted leaf triangles, nodes, root box).  A moving body keeps a ring of
// frame slots: the pose of frame g lives in slot g % nslots, so the body
// poses of coming frames can be computed while the cloth steps through the
// current ones (see Sim::capture).  The updates run in "lanes", each with its
// own build scratch, so several frames can be updated at the same time.
#pragma once

#include <vector>

#include "cc_internal.h"

namespace cc {

constexpr int MAX_CAPS = 64;
// leaves per workgroup of the refit's chunk pass (refit.hip)
constexpr int REFIT_CHUNK = 2048;
constexpr int MAX_SLOTS = 64;
constexpr int MAX_LANES = 4;
// A rebuilt body can build up to MAX_FRAMES frames in one chain of launches
// (grid.y = frame): each kernel of the chain takes one pointer per frame.
constexpr int MAX_FRAMES = 4;
template <class T>
struct PF {
  T* p[MAX_FRAMES];
};
template <class T>
inline PF<T> pf1(T* p) {
  PF<T> r{};
  r.p[0] = p;
  return r;
}

// Per-frame motion of a capsule-union body (see cc_body_desc), by value.
struct MotionParams {
  double amp = 0, freq = 0, dx = 0, dy = 0, dz = 0, rot_amp = 0, px = 0, py = 0, pz = 0;
  double breath_amp = 0, breath_freq = 0, dt = 0;
  uint32_t ncaps = 0;
};

// The parameters of one frame, by value: v[0..8] = R (row-major rotation
// about +y), v[9..11] = T (with the pivot folded in), v[12+k] = breathing
// factor of capsule k.
struct MotionFrame {
  float v[12 + MAX_CAPS];
};
struct MotionFrames {
  MotionFrame f[MAX_FRAMES];
};

// Motion of frame `frame` (t = frame * dt), computed on the host in double
// precision with one rounding to float.
MotionFrame motion_of(const MotionParams& a, const float* phase_host, uint64_t frame);

// One body of the batch; offsets index the concatenated arrays.
struct BodyDesc {
  uint32_t nv, m, ncaps, mode;
  uint32_t v0;    // first vertex
  uint32_t t0;    // first triangle == first leaf / key
  uint32_t n0;    // first internal node
  uint32_t c0;    // first capsule (p0 / p1 entries)
  uint32_t cv0;   // first capsule vertex-range entry (ncaps + 1 per body)
  uint32_t p0;    // first partial root box (one per 256 triangles)
  float depth;
  uint32_t pad;
};

// A node finished by one level of the bottom-up build and handed to the
// next (its sorted-leaf range, Karras index and box).
struct alignas(16) BItem {
  int32_t l, r, idx, pad;
  float box[6];
  float pad2[2];
};

// One workgroup of a build level: chunk q of the body's level, input slots
// from in_base (n_in of them; leaves at level 0), output slots at out_base.
struct LvBlock {
  uint32_t body, q, in_base, out_base, n_in, pad[3];
};

constexpr int BB_MAX_LEVELS = 8;       // 2^30 leaves: 2^20 -> 2^17 -> ... -> 4 -> 1 chunks
constexpr uint32_t LV_MAX_LEAVES = 1u << 21;   // leaf indices of the bottom-up build carry 21 bits

// The bottom-up build (bbody.hip): work tables of its levels for the given
// bodies (level 0: chunks of 1024 leaves, level L+1: 8 chunks of level L),
// the number of level-0 chunks of m leaves, and the launches of all levels.
// keys / leaf_tri / boxes / leaf_parent / nodes are indexed through each
// body's t0 / n0; root and leaf0 hold 6 floats per body.
std::vector<std::vector<LvBlock>> build_level_tables(const std::vector<BodyDesc>& h,
                                                     const std::vector<uint32_t>& bodies);
size_t level_chunks0(uint32_t m);
// The buffers of one frame of a level build.
struct LvFrame {
  const uint64_t* keys;
  const uint4* leaf_tri;
  const float* boxes;
  BItem* items[2];
  uint32_t* nitems[2];
  Node* nodes;
  int32_t* leaf_parent;
  float* root;
  float* leaf0;
};
// nframes (1..MAX_FRAMES) frames of the same tables in each launch (grid.y)
void launch_build_levels(const BodyDesc* desc, const LvBlock* const* level, const uint32_t* n_level,
                         const LvFrame* frames, int nframes, hipStream_t s);

// One frame of the body as the collision query reads it.
struct Slot {
  DBuf<float> verts;       // [3*nv] pose
  DBuf<uint4> leaf_tri;    // [m] sorted leaves (rebuilt bodies only; else the shared topology's)
  DBuf<Node> nodes;        // [max(m-1, 1)]
  DBuf<float> root;        // [6] root box (sentinel U all)
  DBuf<float> leaf0_box;   // [6] box of leaf 0 when m == 1
  DBuf<float> motion;      // [12 + ncaps] motion parameters of the frame
  DBuf<float4> wide;       // [8 * nwide] four-wide nodes: four 32-byte entries
                           // {min xyz, max x | max yz, link, split}
};

// Scratch of one update lane.  A full lane can build a tree; a refit-only
// lane holds just what a refit writes (the topology then comes from lane 0).
struct Lane {
  bool full = false;
  uint32_t nblk_box = 0;
  DBuf<float> boxes;                // [6*m] unsorted expanded boxes
  DBuf<float> partial;              // [6*nblk_box] per-block root boxes
  DBuf<uint32_t> code[2], idx[2];   // radix sort ping-pong
  DBuf<uint32_t> hist;              // [256 * sort blocks]
  DBuf<uint32_t> digit_total;       // [1024] keys per digit of the current pass
  DBuf<uint64_t> keys;              // [m] sorted morton << 32 | primitive
  DBuf<int32_t> leaf_parent;        // [m]
  // refit description: internal nodes whose leaf range crosses a chunk
  // boundary ("top" nodes) and per-node level information
  DBuf<int32_t> top_list;           // [max(m-1, 1)] their indices, sorted by top height
  DBuf<int32_t> top_pos;            // [max(m-1, 1)] node -> position in top_list, -1 if not top
  DBuf<uint32_t> top_count;         // [1]
  DBuf<int4> top_links;             // [max(m-1, 1)] per entry: top child A / B, parent, parent slot
  DBuf<int4> links;                 // [max(m-1, 1)] per node: childA, childB, parent, other (the
                                    // refit reads these 16 bytes, not the 64-byte node records)
  DBuf<uint8_t> top_h;              // [max(m-1, 1)] height of each entry in the top tree
  DBuf<uint32_t> top_hmax;          // [1]
  DBuf<uint32_t> top_lev;           // [66] entries of top height h are [top_lev[h], top_lev[h+1])
  DBuf<uint8_t> nflag;              // [max(m-1, 1)] in-chunk height | flags
  DBuf<uint8_t> lflag;              // [m] flags of the leaf
  DBuf<uint32_t> chunk_h;           // [chunks] largest in-chunk height
  // written by every refit
  DBuf<float> top_child;            // [12 * max(m-1, 1)] boxes of the entries' non-top children
  DBuf<float> top_scratch;          // set-up work space and the top pass beyond its LDS size
  // bottom-up build (rebuilt bodies): unfinished nodes between levels
  DBuf<BItem> items[2];
  DBuf<uint32_t> nitems[2];

  void alloc(uint32_t m, bool full_lane, bool bottom_up, hipStream_t s);
};

// Host copy of a built tree in the reference encoding (see cc_bvh_export):
// sorted keys, leaf order, links, child boxes, root box.
void export_tree_raw(uint32_t m, const uint64_t* keys_d, const uint4* leaf_tri_d, const int32_t* leaf_parent_d,
                     const Node* nodes_d, const float* root_d, const float* leaf0_d, uint64_t* okeys,
                     uint32_t* order, int32_t* childA, int32_t* childB, int32_t* parent_int, int32_t* parent_leaf,
                     float* box_int, float* box_leaf, float* root6, hipStream_t s);

struct Body {
  uint32_t nv = 0;      // body vertices
  uint32_t m = 0;       // triangles == LBVH leaves
  uint32_t mode = CC_BODY_NONE;
  float depth = 0.01f;  // expanded box depth (primitive.cu:44)

  DBuf<float> rest;     // [3*nv] rest pose (after vertex_extend)
  DBuf<uint32_t> tris;  // [3*m]
  DBuf<uint4> leaf_tri; // [m] sorted leaves of the shared topology (static and refitted bodies)

  // A rebuilt body of 2..2^21 triangles builds its tree and boxes bottom-up
  // (the batched build's level kernels) instead of the Karras searches, the
  // refit description and the refit.
  bool bottom_up = false;
  DBuf<BodyDesc> bu_desc;           // [1] offsets 0
  DBuf<LvBlock> bu_level[BB_MAX_LEVELS];
  uint32_t bu_n[BB_MAX_LEVELS] = {};

  // capsule decomposition for the deformation
  uint32_t ncaps = 0;
  DBuf<uint32_t> cap_vstart;        // [ncaps+1]
  DBuf<float> cap_p0, cap_p1, cap_phase;
  std::vector<float> h_phase;       // [ncaps] host copy for the motion
  MotionParams motion;

  // Four-wide form of a static or refitted body's tree (fixed topology):
  // per wide node, the binary records holding its entries' boxes (record *
  // 2 + side, ~0 for an empty entry) and its links (wide index >= 0, ~leaf).
  // A rebuilt body (`grouped`) has one wide record per internal node,
  // folded on the device after every build (fold_frames), and no tables.
  uint32_t nwide = 0;
  uint32_t wide_stack = 0;          // most entries a traversal can defer
  bool wide_enabled = true;
  bool grouped = false;
  DBuf<uint4> w_src;
  DBuf<int4> w_link;
  DBuf<uint2> w_dst;                // [2 * (m - 1)] binary child slot -> {wide entry (record * 4 +
                                    // entry), its link}, entry ~0 for the slots of folded nodes

  int nslots = 1, nlanes = 1;
  // frames per chain of build launches (rebuilt bodies built bottom-up):
  // frame q of a chain on lane l uses the scratch of lane l + nlanes * q
  int chain = 1;
  Slot slot[MAX_SLOTS];
  Lane lane[MAX_LANES * MAX_FRAMES];

  bool moving() const { return mode == CC_BODY_REFIT || mode == CC_BODY_REBUILD; }
  // slot holding the pose of frame g
  int slot_of(uint64_t g) const { return moving() ? (int)(g % (uint64_t)nslots) : 0; }
  static size_t top_scratch_floats(uint32_t m);

  uint4* leaf_tri_of(int k) { return mode == CC_BODY_REBUILD ? slot[k].leaf_tri.p : leaf_tri.p; }
  // lane whose topology a refit on lane l reads
  Lane& topo(int l) { return mode == CC_BODY_REBUILD ? lane[l] : lane[0]; }

  BvhView view(int k) const {
    BvhView v;
    v.nodes = slot[k].nodes.p;
    v.leaf_tri = mode == CC_BODY_REBUILD ? slot[k].leaf_tri.p : leaf_tri.p;
    v.verts = slot[k].verts.p;
    v.leaf0_box = slot[k].leaf0_box.p;
    v.m = (mode == CC_BODY_NONE) ? 0 : m;
    v.wide = (nwide && wide_enabled) ? slot[k].wide.p : nullptr;
    v.grouped = grouped ? 1u : 0u;
    return v;
  }

  // set-up: host arrays, vertices stride 3 (already extended); a moving
  // body gets `slots` frame slots and `lanes` update lanes, a rebuilt one
  // built bottom-up `chain` frames per chain of launches (allow_bottom_up
  // false: a rebuilt body of any size takes the Karras searches and the
  // refit, the path of bodies beyond the bottom-up build's leaf limit)
  void init(const float* verts3, uint32_t nv, const uint32_t* tris, uint32_t nf, uint32_t mode, float depth,
            const cc_capsules* caps, const MotionParams& mp, int slots, int lanes, hipStream_t s,
            int chain = 1, bool allow_bottom_up = true);

  // launches, asynchronous on s
  void deform(int k, uint64_t frame, hipStream_t s);   // rest -> slot k pose of `frame`
  void build(int k, int l, hipStream_t s);      // full LBVH build of slot k on lane l (includes the refit)
  // F frames in one chain of launches (grid.y = frame): poses of frames[q]
  // into slots ks[q]; trees of slots ks[q] with the scratch of lanes ls[q]
  // (F > 1 only for a body built bottom-up)
  void deform_frames(int F, const int* ks, const uint64_t* frames, hipStream_t s);
  void build_frames(int F, const int* ks, const int* ls, hipStream_t s);
  void update_frames(int F, const int* ks, const int* ls, const uint64_t* frames, hipStream_t s);
  void build_levels(int F, const int* ks, const int* ls, hipStream_t s);   // bottom-up trees + boxes
  void prim_boxes_of(int k, int l, hipStream_t s);  // lane l's expanded boxes of slot k's pose
  void top_setup(int k, int l, hipStream_t s);  // refit description of slot k's topology into lane l
  // boxes of slot k: the four-wide form's only, when the body is walked
  // through it, unless `binary`
  void refit(int k, int l, hipStream_t s, bool binary = false);
  // F frames' refits in one chain of launches (grid.y = frame; a refitted
  // body: every slot shares lane 0's topology)
  void refit_frames(int F, const int* ks, const int* ls, hipStream_t s, bool binary = false);
  void update(int k, int l, uint64_t frame, hipStream_t s);   // per-mode update of slot k to `frame`
  void share_topology(int from, hipStream_t s); // copy slot `from`'s tree to every other slot
  // Four-wide form: the tables from slot k's tree and slot k's wide records
  // (host, synchronous; keeps the binary form when the stack bound fails);
  // later refits write the boxes through w_dst.
  void wide_setup(int k, hipStream_t s);
  // Rebuilt bodies: the wide records of slots ks[q] from their binary trees
  // (grid.y = frame; nothing when the binary walk is selected).
  void fold_frames(int F, const int* ks, hipStream_t s);

  // host export in the reference encoding (lane 0 holds the keys and leaf
  // parents of slot k's tree)
  void export_tree(int k, uint64_t* keys, uint32_t* order, int32_t* childA, int32_t* childB,
                   int32_t* parent_int, int32_t* parent_leaf, float* box_int, float* box_leaf,
                   float* root6, hipStream_t s);
};

}  // namespace cc
