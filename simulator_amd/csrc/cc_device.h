// cc_device.h -- device routines shared by the step, query and batched
// kernels: first-hit LBVH traversal and the triangle plane test with
// projection.  All comparisons and roundings follow the reference order so
// the colliding-vertex set is bit-exact given the same query points.
#pragma once

#include "cc_internal.h"

namespace cc {

// Deferred-node stack of one traversal beyond the four entries held in
// registers: entry k of a lane is p[k * stride] (lanes of a wave adjacent,
// so spills coalesce).  It lives in global memory that the caller provides,
// STACK_SPILL entries per lane: a private array here would make the kernel
// use scratch memory, which was seen to cap the number of resident waves.
constexpr int STACK_SPILL = 60;
// The four-wide walk keeps its first WSTK_LDS deferred entries in LDS, entry
// k of a lane at lds[k * WSTK_LANES] (the kernel's workgroup size), so every
// lane of a wave hits its own bank whatever its depth; deeper entries go to
// the global spill area.  16 entries (16 KB per workgroup) hold the deepest
// stacks the walks reach at the large workload (9), and leave room in a CU
// for the body-update kernels that run beside the step.
constexpr int WSTK_LDS = 16;
constexpr int WSTK_LANES = 256;
// the wide walk defers at most 64 entries: the ones past the LDS part fit
static_assert(64 - WSTK_LDS <= STACK_SPILL, "wide-walk stack beyond LDS exceeds the spill area");
// A spill area of L lanes is STACK_SPILL * L entries followed by a tail of
// SPILL_TAIL words: [0] the traversal faults of the launches that used it
// (TRAV_OVERFLOW: a walk needed more deferred entries than the stack limit,
// its query was then taken as a miss; TRAV_BOUND: a walk reached its visit
// bound, which only a corrupted tree can), [1] the stack limit, 64 (the
// most a walk of a valid tree defers) unless a test lowers it to force an
// overflow.  The host reads and clears the faults (cc_sim_check_finite).
constexpr int SPILL_TAIL = 2;
constexpr uint32_t TRAV_OVERFLOW = 1u, TRAV_BOUND = 2u;
constexpr int STACK_LIMIT = 64;
struct Spill {
  int32_t* p;
  uint32_t stride;
  int32_t* lds = nullptr;   // this lane's entry 0 (kernels that walk the four-wide form)
  int32_t* tail = nullptr;  // the area's tail (base + STACK_SPILL * stride)
};
// the spill area of lane i of a launch over `stride` lanes
__device__ inline Spill spill_of(int32_t* base, uint32_t i, uint32_t stride, int32_t* lds = nullptr) {
  return Spill{base + i, stride, lds, base + (size_t)STACK_SPILL * stride};
}
__device__ inline void trav_fault(const Spill& s, uint32_t bit) {
  atomicOr(reinterpret_cast<uint32_t*>(s.tail), bit);
}
__device__ inline int stack_limit(const Spill& s) { return min(s.tail[1], STACK_LIMIT); }

// host: a spill area for `lanes` lanes, its tail set to no faults and the
// default stack limit (set-up only: a synchronous copy on s)
inline void alloc_spill(DBuf<int32_t>& b, size_t lanes, hipStream_t s) {
  b.alloc((size_t)STACK_SPILL * lanes + SPILL_TAIL);
  const int32_t tail[SPILL_TAIL] = {0, STACK_LIMIT};
  copy_on(b.p + (size_t)STACK_SPILL * lanes, tail, sizeof(tail), hipMemcpyHostToDevice, s);
}

// host: the finite check of n {x_cur, x_prev} records at `state` and the
// traversal faults recorded in the spill area of `lanes` lanes (read, then
// cleared); throws CC_ERR_NONFINITE or CC_ERR_TRAVERSAL.  flag: one device
// word of scratch.  Synchronous on s.
void check_state(const float* state, uint32_t n, DBuf<int32_t>& spill, size_t lanes, uint32_t* flag,
                 hipStream_t s, const char* who);

// Deferred nodes: the top four in registers (static shifts, no indexed
// private memory), older ones in the spill area.
struct TravStack {
  int32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int sp = 0;
  int limit = STACK_LIMIT;
  Spill spill;
  __device__ bool push(int32_t v) {
    if (sp >= limit) return false;       // unreachable for a valid tree at the default limit
    if (sp >= 4) spill.p[(size_t)(sp - 4) * spill.stride] = s3;
    s3 = s2;
    s2 = s1;
    s1 = s0;
    s0 = v;
    ++sp;
    return true;
  }
  __device__ bool pop(int32_t& v) {
    if (sp == 0) return false;
    v = s0;
    s0 = s1;
    s1 = s2;
    s2 = s3;
    --sp;
    if (sp >= 4) s3 = spill.p[(size_t)(sp - 4) * spill.stride];
    return true;
  }
};

enum : int { STEP_DESCEND = -1, STEP_POP = -2, STEP_MISS = -3 };

// One internal node of D_BVH::intersect (cudaBRT.cu:239-285), given its
// record (q0..q2 = child boxes A and B, L = childA, childB, ...).  Both child
// boxes are tested; an overlapped LEAF child is returned at once (A before
// B); otherwise overlapped internal children are descended left first with
// the right one deferred on the stack.  Returns the leaf (>= 0), or
// STEP_DESCEND / STEP_POP with the next node in `next`, or STEP_MISS.
// Both overlap tests and the leaf decision are computed before the first
// branch (`&` and a select, not `&&` and two returns): otherwise the
// compiler moves the loads of the second box behind the first test, and a
// node costs two memory round trips instead of one.
__device__ inline int32_t trav_step(f3 p, float4 q0, float4 q1, float4 q2, int4 L, TravStack& st,
                                    int32_t& next) {
  // A = [q0.xyz, (q0.w, q1.x, q1.y)], B = [(q1.z, q1.w, q2.x), q2.yzw]
  const bool oL = (p.x >= q0.x) & (p.x <= q0.w) & (p.y >= q0.y) & (p.y <= q1.x) & (p.z >= q0.z) &
                  (p.z <= q1.y);
  const bool oR = (p.x >= q1.z) & (p.x <= q2.y) & (p.y >= q1.w) & (p.y <= q2.z) & (p.z >= q2.x) &
                  (p.z <= q2.w);
  const int32_t ca = L.x, cb = L.y;
  const bool hitA = oL & (ca < 0), hitB = oR & (cb < 0);
  const int32_t leaf = hitA ? ~ca : ~cb;   // A before B
  if (hitA | hitB) return leaf;
  const bool tL = oL & (ca >= 0), tR = oR & (cb >= 0);
  if (tL) {
    if (tR && !st.push(cb)) {
      trav_fault(st.spill, TRAV_OVERFLOW);
      return STEP_MISS;
    }
    next = ca;
    return STEP_DESCEND;
  }
  if (tR) {
    next = cb;
    return STEP_DESCEND;
  }
  return st.pop(next) ? STEP_POP : STEP_MISS;
}

__device__ inline bool in_box(f3 p, float lx, float ly, float lz, float hx, float hy, float hz) {
  return (p.x >= lx) & (p.x <= hx) & (p.y >= ly) & (p.y <= hy) & (p.z >= lz) & (p.z <= hz);
}

// The root record, loaded by the caller with its other first loads (the
// same address for every lane), so the walk starts without a round trip.
struct WideRoot {
  float4 q[8];
};

// The root record of the wide form for a walk (all lanes read the same
// record); unused when the view has no wide form.
template <bool WIDE = true>
__device__ inline WideRoot load_wide_root(const BvhView& b) {
  WideRoot r;
  if (WIDE && b.wide && b.m >= 2) {
#pragma unroll
    for (int k = 0; k < 8; ++k) r.q[k] = b.wide[k];
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) r.q[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  return r;
}

// The same first hit through the four-wide nodes.  A binary node X with
// children A, B considers, in this order: A if it is a leaf, B if it is a
// leaf, then the subtree of A, then the subtree of B (trav_step).  A wide
// node stores the entries of X and of its internal children in exactly that
// order of consideration (leaf children of X; then per internal child c:
// c's leaf children, then c's internal children), so the first hit is the
// first entry whose box holds p: a leaf is returned, an internal entry is
// descended with the later hits deferred behind it, and a deferred leaf is
// returned when popped.  The boxes of the folded nodes are not needed: a
// node's box is the union of its children's, so it holds p whenever a box
// below it does.  Entries past the count have an empty box (+inf, -inf).
// A record is four 32-byte entries {min xyz, max x | max yz, link, split}.
// Two layouts share the walk:
//  - static and refitted bodies (set up once per topology, on the host):
//    folded records, split 0; the set-up keeps the wide form only when no
//    root path defers more than 64 entries (sum of entries - 1 over the
//    path), so the stack cannot overflow;
//  - rebuilt bodies (folded on the device every frame): one record per
//    internal node X at X's index, holding X's leaf children and the
//    children of its internal children.  When both children A and B are
//    internal, entries 0-1 are A's children and 2-3 B's, and entry 0's
//    split word is B's index: while the first candidate lies in A's pair,
//    the walk defers B's record instead of B's two entries (or the one of
//    them that holds p).  Then every visit defers at most one node per
//    binary level it spans, as the binary walk does, so the stack stays
//    within the tree depth (below 64).
__device__ inline void wstk_put(const Spill& sp, int k, int32_t v) {
  if (k < WSTK_LDS)
    sp.lds[k * WSTK_LANES] = v;
  else
    sp.p[(size_t)(k - WSTK_LDS) * sp.stride] = v;
}
// The LDS read is unconditional (index clamped) and the global one only on
// the rare deep branch: a select between the two addresses would become one
// generic (flat) load.
__device__ inline int32_t wstk_get(const Spill& sp, int k) {
  int32_t v = sp.lds[min(k, WSTK_LDS - 1) * WSTK_LANES];
  asm volatile("" : "+v"(v));
  if (k >= WSTK_LDS) v = sp.p[(size_t)(k - WSTK_LDS) * sp.stride];
  return v;
}

// One visit of a wide record q[0..7] (entry e is q[2e], q[2e + 1]) takes
// the first entry whose box holds p: a leaf is the hit (returned, >= 0), an
// internal entry the next record (WIDE_NEXT, in `node`); with none, the
// walk pops (a deferred leaf is the hit) or misses (WIDE_MISS).  The later
// candidates of a visit (at most three) are deferred with three
// unconditional stores at the stack top, highest entry first, and the stack
// pointer then moves by their number: a slot above the new top holds
// nothing the walk reads.  This takes far fewer instructions than shifting
// a register stack entry by entry, which matters because all four waves of
// a SIMD walk at once.
enum : int { WIDE_NEXT = -1, WIDE_MISS = -2 };
template <bool GROUPED>
__device__ inline int32_t wide_visit(f3 p, float4 q0, float4 q1, float4 q2, float4 q3, float4 q4, float4 q5,
                                     float4 q6, float4 q7, const Spill& spill, int& sp, int32_t& node,
                                     int limit) {
  // Kept as booleans (lane masks): the first candidate and the deferred set
  // follow from them by selects, with no per-lane bit mask to build.
  const bool i0 = in_box(p, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y), i1 = in_box(p, q2.x, q2.y, q2.z, q2.w, q3.x, q3.y);
  bool i2 = in_box(p, q4.x, q4.y, q4.z, q4.w, q5.x, q5.y), i3 = in_box(p, q6.x, q6.y, q6.z, q6.w, q7.x, q7.y);
  const int32_t c0 = __float_as_int(q1.z), c1 = __float_as_int(q3.z), c3 = __float_as_int(q7.z);
  int32_t c2 = __float_as_int(q5.z);
  if constexpr (GROUPED) {
    // A split record whose first candidate is in entries 0-1: entries 2-3
    // become one candidate, the record of their parent (the split word;
    // the root is no node's child, so a split is never 0) when both hold
    // p, else the one that does (the same next step as that record's
    // visit, one visit fewer).
    const int32_t split = __float_as_int(q1.w);
    const bool in_a = (split != 0) & (i0 | i1);
    const int32_t b_next = (i2 & i3) ? split : (i2 ? c2 : c3);
    c2 = in_a ? b_next : c2;
    i2 = in_a ? (i2 | i3) : i2;
    i3 = i3 & !in_a;
  }
  if (!(i0 | i1 | i2 | i3)) {
    if (sp == 0) return WIDE_MISS;
    --sp;
    node = wstk_get(spill, sp);
    return node < 0 ? ~node : WIDE_NEXT;
  }
  // the first candidate's link
  const int32_t c = i0 ? c0 : (i1 ? c1 : (i2 ? c2 : c3));
  if (c < 0) return ~c;
  // the later candidates (a candidate with one before it), popped lowest
  // entry first
  const bool r1 = i1 & i0, r2 = i2 & (i0 | i1), r3 = i3 & (i0 | i1 | i2);
  const int32_t x0 = r3 ? c3 : (r2 ? c2 : c1);
  const int32_t x1 = r3 ? (r2 ? c2 : c1) : c1;
  const int k = (int)r1 + (int)r2 + (int)r3;
  if (sp + k > limit) {                 // unreachable at the default limit (both layouts
    trav_fault(spill, TRAV_OVERFLOW);   // bound the deferrals by 64)
    return WIDE_MISS;
  }
  if (sp + 3 <= WSTK_LDS) {
    int32_t* t = spill.lds + sp * WSTK_LANES;
    t[0] = x0;
    t[WSTK_LANES] = x1;
    t[2 * WSTK_LANES] = c1;
  } else {
    if (k > 0) wstk_put(spill, sp, x0);
    if (k > 1) wstk_put(spill, sp + 1, x1);
    if (k > 2) wstk_put(spill, sp + 2, c1);
  }
  sp += k;
  node = c;
  return WIDE_NEXT;
}

// GROUPED: a rebuilt body's records (split words); the walk of static and
// refitted bodies is compiled without them.
template <bool GROUPED>
__device__ inline int32_t wide_first_hit(const BvhView& b, f3 p, uint32_t& visits, Spill spill, const WideRoot& rt) {
  int sp = 0;
  int32_t node = 0;
  const int limit = stack_limit(spill);
  ++visits;
  int32_t r = wide_visit<GROUPED>(p, rt.q[0], rt.q[1], rt.q[2], rt.q[3], rt.q[4], rt.q[5], rt.q[6], rt.q[7], spill, sp, node,
                         limit);
  uint32_t guard = 0;
  for (; r == WIDE_NEXT && guard < b.m; ++guard) {
    // a 32-bit byte offset (records < 4 GiB): the loads take the base in
    // scalar registers
    const float4* nd = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(b.wide) +
                                                       ((uint32_t)node << 7));
    const float4 q0 = nd[0], q1 = nd[1], q2 = nd[2], q3 = nd[3], q4 = nd[4], q5 = nd[5], q6 = nd[6],
                 q7 = nd[7];
    // All four links arrive with the boxes: otherwise the compiler loads
    // entry 0's link apart, inside the branch that reads it, one more
    // memory round trip per descending visit.
    asm volatile("" ::"v"(q1.z), "v"(q3.z), "v"(q5.z), "v"(q7.z));
    ++visits;
    r = wide_visit<GROUPED>(p, q0, q1, q2, q3, q4, q5, q6, q7, spill, sp, node, limit);
  }
  if (r == WIDE_NEXT) trav_fault(spill, TRAV_BOUND);
  return r >= 0 ? r : -1;
}

// First-hit traversal (D_BVH::intersect, cudaBRT.cu:239-285): one node
// record (64 bytes, both child boxes and links) per memory round trip.  The
// stack never holds more entries than the node depth, and the depth of a
// Karras tree over unique 64-bit keys is below 64, so 64 entries cannot
// overflow.  Returns the sorted leaf index, or -1.  *visits counts internal
// nodes processed (wide nodes when the view has them).
// WIDE false: the caller's views never have the wide form (batched bodies),
// and the wide walk is left out of its kernel.
template <bool WIDE = true>
__device__ inline int32_t bvh_first_hit(const BvhView& b, f3 p, uint32_t& visits, Spill spill,
                                       const WideRoot& rt) {
  if (b.m == 0) return -1;
  if (b.m == 1) return box_has(b.leaf0_box, b.leaf0_box + 3, p) ? 0 : -1;
  if (WIDE && b.wide)
    return b.grouped ? wide_first_hit<true>(b, p, visits, spill, rt) : wide_first_hit<false>(b, p, visits, spill, rt);
  TravStack st;
  st.spill = spill;
  st.limit = stack_limit(spill);
  int32_t node = 0;
  // A well-formed tree visits each internal node at most once; the bound
  // only guards against a corrupted tree (the loop always terminates).
  for (uint32_t guard = 0; guard < b.m; ++guard) {
    const float4* nd = reinterpret_cast<const float4*>(b.nodes + node);
    const float4 q0 = nd[0], q1 = nd[1], q2 = nd[2];
    const int4 L = reinterpret_cast<const int4*>(nd)[3];
    ++visits;
    const int32_t r = trav_step(p, q0, q1, q2, L, st, node);
    if (r >= 0) return r;
    if (r == STEP_MISS) return -1;
  }
  trav_fault(spill, TRAV_BOUND);
  return -1;
}

// Primitive::d_intersect (primitive.cu:70-85): n = normalize(cross(v1-v0,
// v2-v0)), d = dot(p - v0, n); inside iff !(d > 0) (a NaN d is inside).
__device__ inline bool plane_test(const float* verts, uint4 t, f3 p, float& d, f3& n) {
  const f3 a = ld3(verts + 3 * (size_t)t.x), b = ld3(verts + 3 * (size_t)t.y),
           c = ld3(verts + 3 * (size_t)t.z);
  n = normalize(cross(b - a, c - a));
  d = dot(p - a, n);
  return !(d > 0.0f);
}

// collision_response_projection (verlet.cu:24-50) on a position p: when the
// first-hit leaf's plane test says inside, p += |d| n and the contact normal
// is n, else the contact normal is 0.  Returns true when projected.
template <bool WIDE = true>
__device__ inline bool collide_project(const BvhView& b, f3& p, f3& n_out, int32_t& hit,
                                       uint32_t& visits, Spill spill, const WideRoot& rt) {
  hit = bvh_first_hit<WIDE>(b, p, visits, spill, rt);
  n_out = mk(0.0f, 0.0f, 0.0f);
  if (hit < 0) return false;
  float d;
  f3 n;
  if (!plane_test(b.verts, b.leaf_tri[hit], p, d, n)) return false;
  const float dd = 1.0f * fabsf(d);
  p = p + n * dd;
  n_out = n;
  return true;
}

}  // namespace cc
