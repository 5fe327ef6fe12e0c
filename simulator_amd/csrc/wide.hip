// wide.hip -- the four-wide form of the body's LBVH.
//
// The collision query walks the Karras tree one node record per dependent
// memory round trip, and the slowest lanes of a step (points resting on the
// body) walk 50-70 nodes.  Folding every second level into its parent halves
// that chain: a wide node holds the entries of a binary node X and of X's
// internal children, four boxes and four links in one 128-byte record, and
// the traversal (wide_first_hit) tests all four at once.
//
// The entries are stored in the order the binary traversal considers them
// (D_BVH::intersect, cudaBRT.cu:239-285, returns an overlapped leaf child at
// once, A before B, before it descends into internal children), so the first
// hit, and with it the contact set, is the same leaf.  The binary tree stays
// the topology the refit walks.  The layout is computed once per topology on
// the host (Body::wide_setup), which writes the records in full from the
// binary boxes (k_fold_wide).  Later refits write each box straight into the
// wide entry that mirrors its binary child slot (w_dst) and leave the binary
// boxes alone; export and the binary walk ask for a binary refit.
//
// A rebuilt body's topology changes every frame, so its wide form is folded
// on the device after each build (k_fold_grouped): one record per internal
// node at the node's own index, no host set-up, no bound to check (the
// split records keep the walk's deferrals within the tree depth).
#include <algorithm>
#include <cstring>
#include <vector>

#include "cc_lbvh.h"

namespace cc {

// Record w of the wide form: four 32-byte entries, entry e = floats
// [8e, 8e + 8) = {min xyz, max x, max yz, link, 0}.  src[w] names for each
// entry the binary record and side (record * 2 + side, side 0 = child A)
// that holds its box, ~0 for an empty entry (box +inf / -inf, link 0).
__global__ __launch_bounds__(256) void k_fold_wide(const Node* __restrict__ nodes, const uint4* __restrict__ src,
                                                   const int4* __restrict__ link, uint32_t nw,
                                                   float4* __restrict__ wide) {
  const uint32_t w = blockIdx.x * 256 + threadIdx.x;
  if (w >= nw) return;
  const uint4 s = src[w];
  const int4 l = link[w];
  const uint32_t e4[4] = {s.x, s.y, s.z, s.w};
  const int32_t l4[4] = {l.x, l.y, l.z, l.w};
  float4* o = wide + 8 * (size_t)w;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    if (e4[e] != ~0u) {
      const float* f = reinterpret_cast<const float*>(nodes + (e4[e] >> 1)) + 6 * (e4[e] & 1u);
      for (int c = 0; c < 6; ++c) b[c] = f[c];
    }
    o[2 * e] = make_float4(b[0], b[1], b[2], b[3]);
    o[2 * e + 1] = make_float4(b[4], b[5], __int_as_float(l4[e]), 0.0f);
  }
}

// Entry e of record o: box b6 (min xyz, max xyz), link, split word.
__device__ inline void put_entry(float4* o, int e, const float* mn, const float* mx, int32_t link, int32_t split) {
  o[2 * e] = make_float4(mn[0], mn[1], mn[2], mx[0]);
  o[2 * e + 1] = make_float4(mx[1], mx[2], __int_as_float(link), __int_as_float(split));
}
__device__ inline void put_empty(float4* o, int e) {
  o[2 * e] = make_float4(INFINITY, INFINITY, INFINITY, -INFINITY);
  o[2 * e + 1] = make_float4(-INFINITY, -INFINITY, 0.0f, 0.0f);
}
// Entries e, e + 1: the children of internal node y in y's order of
// consideration (a leaf child first, A before B).
__device__ inline void put_pair(float4* o, int e, const Node& y, int32_t split) {
  if (y.childA >= 0 && y.childB < 0) {
    put_entry(o, e, y.b_min, y.b_max, y.childB, split);
    put_entry(o, e + 1, y.a_min, y.a_max, y.childA, 0);
  } else {
    put_entry(o, e, y.a_min, y.a_max, y.childA, split);
    put_entry(o, e + 1, y.b_min, y.b_max, y.childB, 0);
  }
}

// The wide form of a rebuilt body, record x from internal node x: X's leaf
// children (A before B), then the children of its internal children (A's,
// then B's).  With both children internal, entry 0's split word is B's
// index.  Reads nodes x, A and B (64 bytes each), writes 128 bytes.
__global__ __launch_bounds__(256) void k_fold_grouped(PF<const Node> nodes_f, uint32_t mi, PF<float4> wide_f) {
  const uint32_t x = blockIdx.x * 256 + threadIdx.x;
  if (x >= mi) return;
  const Node* __restrict__ N = pick(nodes_f);
  float4* o = pick(wide_f) + 8 * (size_t)x;
  const Node X = N[x];
  const bool la = X.childA < 0, lb = X.childB < 0;
  if (la && lb) {
    put_entry(o, 0, X.a_min, X.a_max, X.childA, 0);
    put_entry(o, 1, X.b_min, X.b_max, X.childB, 0);
    put_empty(o, 2);
    put_empty(o, 3);
  } else if (la || lb) {
    if (la) {
      put_entry(o, 0, X.a_min, X.a_max, X.childA, 0);
      put_pair(o, 1, N[X.childB], 0);
    } else {
      put_entry(o, 0, X.b_min, X.b_max, X.childB, 0);
      put_pair(o, 1, N[X.childA], 0);
    }
    put_empty(o, 3);
  } else {
    put_pair(o, 0, N[X.childA], X.childB);
    put_pair(o, 2, N[X.childB], 0);
  }
}

void Body::fold_frames(int F, const int* ks, hipStream_t s) {
  if (!grouped || !wide_enabled) return;
  PF<const Node> nodes{};
  PF<float4> wide{};
  for (int f = 0; f < F; ++f) {
    nodes.p[f] = slot[ks[f]].nodes.p;
    wide.p[f] = slot[ks[f]].wide.p;
  }
  k_fold_grouped<<<dim3(grid_for(m - 1, 256), (unsigned)F), 256, 0, s>>>(nodes, m - 1, wide);
  CC_LAUNCHED();
}

void Body::wide_setup(int k, hipStream_t s) {
  nwide = 0;
  wide_stack = 0;
  if (mode == CC_BODY_NONE || mode == CC_BODY_REBUILD || m < 2) return;
  const uint32_t mi = m - 1;
  std::vector<Node> nd(mi);
  CC_HIP(hipMemcpyAsync(nd.data(), slot[k].nodes.p, mi * sizeof(Node), hipMemcpyDeviceToHost, s));
  CC_HIP(hipStreamSynchronize(s));
  // Breadth-first over the kept binary nodes (the root, then the internal
  // entries of each kept node); wid maps a kept binary node to its record.
  std::vector<int32_t> wid(mi, -1), order;
  std::vector<uint4> src;
  std::vector<int4> link;
  std::vector<uint32_t> acc;     // deferrable entries on the path to each record
  order.reserve(mi / 2 + 1);
  order.push_back(0);
  wid[0] = 0;
  acc.push_back(0);
  uint32_t worst = 0;
  for (size_t h = 0; h < order.size(); ++h) {
    const int32_t x = order[h];
    uint32_t es[4], cnt = 0;
    int32_t ref[4];
    auto add = [&](int32_t rec, int side, int32_t child) {
      CC_REQUIRE(cnt < 4, "wide form: more than four entries");
      es[cnt] = (uint32_t)rec * 2u + (uint32_t)side;
      ref[cnt++] = child;
    };
    // X's children in its order of consideration (leaf children first),
    // then, while there is room, the internal entry with the largest box
    // surface is replaced in place by its own children in their order: an
    // in-place replacement keeps the whole list in the binary walk's order
    // of consideration.  (A CPU replay of a late frame of the large workload:
    // 7.44 wide visits per query and 12.4 per wave's slowest lane, against
    // 7.69 and 12.95 when every node folds exactly its children.)
    add(x, 0, nd[x].childA);
    add(x, 1, nd[x].childB);
    auto order_pair = [&](uint32_t e0) {        // leaf child of the pair first
      if (ref[e0] >= 0 && ref[e0 + 1] < 0) {
        std::swap(es[e0], es[e0 + 1]);
        std::swap(ref[e0], ref[e0 + 1]);
      }
    };
    order_pair(0);
    while (cnt < 4) {
      int best = -1;
      double best_area = -1.0;
      for (uint32_t e = 0; e < cnt; ++e) {
        if (ref[e] < 0) continue;
        const Node& P = nd[es[e] >> 1];
        const float* bmn = (es[e] & 1u) ? P.b_min : P.a_min;
        const float* bmx = (es[e] & 1u) ? P.b_max : P.a_max;
        const double dx = (double)bmx[0] - bmn[0], dy = (double)bmx[1] - bmn[1], dz = (double)bmx[2] - bmn[2];
        const double area = dx * dy + dy * dz + dz * dx;
        if (area > best_area) {
          best_area = area;
          best = (int)e;
        }
      }
      if (best < 0) break;
      const int32_t c = ref[best];
      CC_REQUIRE((uint32_t)c < mi, "wide form: link out of range");
      for (uint32_t e = cnt; e > (uint32_t)best + 1; --e) {
        es[e] = es[e - 1];
        ref[e] = ref[e - 1];
      }
      ++cnt;
      es[best] = (uint32_t)c * 2u;
      ref[best] = nd[c].childA;
      es[best + 1] = (uint32_t)c * 2u + 1u;
      ref[best + 1] = nd[c].childB;
      order_pair((uint32_t)best);
    }
    const uint32_t deferrable = acc[h] + (cnt - 1);
    worst = std::max(worst, deferrable);
    uint4 sv = make_uint4(~0u, ~0u, ~0u, ~0u);
    int4 lv = make_int4(0, 0, 0, 0);
    uint32_t* sp = &sv.x;
    int32_t* lp = &lv.x;
    for (uint32_t e = 0; e < cnt; ++e) {
      sp[e] = es[e];
      if (ref[e] < 0) {
        lp[e] = ref[e];
      } else {
        CC_REQUIRE((uint32_t)ref[e] < mi && wid[ref[e]] < 0, "wide form: node reached twice");
        wid[ref[e]] = (int32_t)order.size();
        lp[e] = (int32_t)order.size();
        order.push_back(ref[e]);
        acc.push_back(deferrable);
      }
    }
    src.push_back(sv);
    link.push_back(lv);
  }
  // The traversal stack holds 64 entries; keep the binary form otherwise.
  if (worst > 64) return;
  // Renumber the records in the order of their binary nodes (Karras
  // indices; the root stays 0): the refit, which walks binary nodes in index
  // order, then writes each chunk's entries into a compact run of records.
  {
    const size_t nw = order.size();
    std::vector<uint32_t> by(nw);
    for (size_t w = 0; w < nw; ++w) by[w] = (uint32_t)w;
    std::sort(by.begin(), by.end(), [&](uint32_t a, uint32_t b) { return order[a] < order[b]; });
    std::vector<int32_t> renum(nw);
    for (size_t r = 0; r < nw; ++r) renum[by[r]] = (int32_t)r;
    std::vector<uint4> src2(nw);
    std::vector<int4> link2(nw);
    for (size_t w = 0; w < nw; ++w) {
      int4 l = link[w];
      int32_t* lp = &l.x;
      const uint32_t* e = &src[w].x;
      for (int q = 0; q < 4; ++q)
        if (e[q] != ~0u && lp[q] >= 0) lp[q] = renum[lp[q]];
      src2[renum[w]] = src[w];
      link2[renum[w]] = l;
    }
    src.swap(src2);
    link.swap(link2);
  }
  std::vector<uint2> dst(2 * (size_t)mi, make_uint2(~0u, 0u));
  for (size_t w = 0; w < src.size(); ++w) {
    const uint32_t* e = &src[w].x;
    const int32_t* lk = &link[w].x;
    for (int q = 0; q < 4; ++q)
      if (e[q] != ~0u) dst[e[q]] = make_uint2((uint32_t)(4 * w + q), (uint32_t)lk[q]);
  }
  w_src.upload(src.data(), src.size(), s);
  w_link.upload(link.data(), link.size(), s);
  w_dst.upload(dst.data(), dst.size(), s);
  nwide = (uint32_t)src.size();
  wide_stack = worst;
  for (int j = 0; j < nslots; ++j) slot[j].wide.alloc(8 * (size_t)nwide);
  // Slot k's records in full (links, empty entries, boxes); the other slots
  // get theirs from share_topology, and every refit rewrites the boxes.
  k_fold_wide<<<grid_for(nwide, 256), 256, 0, s>>>(slot[k].nodes.p, w_src.p, w_link.p, nwide, slot[k].wide.p);
  CC_LAUNCHED();
  CC_HIP(hipStreamSynchronize(s));
}

}  // namespace cc
