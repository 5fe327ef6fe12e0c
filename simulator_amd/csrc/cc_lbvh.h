// cc_lbvh.h -- device routines of the body pose and LBVH build shared by
// the single-body kernels (lbvh.hip, refit.hip) and the batched ones
// (bbody.hip), so both produce the same poses and trees bit for bit.
#pragma once

#include "cc_body.h"

namespace cc {
// the pointer of this workgroup's frame (grid.y)
template <class T>
__device__ inline T* pick(const PF<T>& a) {
  return a.p[blockIdx.y];
}
}  // namespace cc

namespace cc {

__device__ inline void put6(float* d, const float* mn, const float* mx) {
  d[0] = mn[0]; d[1] = mn[1]; d[2] = mn[2];
  d[3] = mx[0]; d[4] = mx[1]; d[5] = mx[2];
}

__device__ inline void unite(float* mn, float* mx, const float* b6) {
  for (int c = 0; c < 3; ++c) {
    mn[c] = fminf(mn[c], b6[c]);
    mx[c] = fmaxf(mx[c], b6[3 + c]);
  }
}

// A box of the four-wide form: entry d = {record * 4 + entry, link}, ~0 for
// none.  An entry is one 32-byte sector {min xyz, max x | max yz, link, 0}
// written whole by one thread (two 16-byte stores).
__device__ inline void put_wide_at(float4* wide, uint2 d, const float* b6) {
  if (d.x == ~0u) return;
  float4* f = wide + 2 * (size_t)d.x;
  f[0] = make_float4(b6[0], b6[1], b6[2], b6[3]);
  f[1] = make_float4(b6[4], b6[5], __int_as_float((int32_t)d.y), 0.0f);
}

__device__ inline void box_sentinel(float* mn, float* mx) {   // bbox.h:17-22
  mn[0] = mn[1] = mn[2] = 100.0f;
  mx[0] = mx[1] = mx[2] = -100.0f;
}

// capsule of vertex v: the last k with vstart[k] <= v
__device__ inline uint32_t capsule_of(const uint32_t* vstart, uint32_t ncaps, uint32_t v) {
  uint32_t lo = 0, hi = ncaps - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (vstart[mid] <= v) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One body vertex of a frame: radial breathing by `breath` about the closest
// point of its capsule axis P0-P1, then R * Y + T (R row-major).
__device__ inline void deform_point(f3 X, f3 P0, f3 P1, float breath, const float* R, const float* T, float* o) {
  const f3 d = P1 - P0;
  const float dd = dot(d, d);
  float tt = 0.0f;
  if (dd > 0.0f) tt = dot(X - P0, d) / dd;
  tt = fminf(fmaxf(tt, 0.0f), 1.0f);
  const f3 A = P0 + d * tt;
  const f3 Y = A + (X - A) * breath;
  o[0] = ((R[0] * Y.x + R[1] * Y.y) + R[2] * Y.z) + T[0];
  o[1] = ((R[3] * Y.x + R[4] * Y.y) + R[5] * Y.z) + T[1];
  o[2] = ((R[6] * Y.x + R[7] * Y.y) + R[8] * Y.z) + T[2];
}

// Internal node i of the Karras tree over the m sorted 64-bit keys:
// processInternalNode / determineRange / findSplit (cudaBRT.cu:9-161) with
// delta = clz64 of the keys' xor (-1 outside [0, m)).  Writes the node's
// children and range end, and the parent link of both children.
__device__ inline void karras_node(const uint64_t* keys, int32_t m, int32_t i, Node* nodes, int32_t* leaf_parent) {
  const uint64_t ki = keys[i];
  auto delta = [&](int64_t j) -> int {
    if (j < 0 || j >= m) return -1;
    return __clzll(ki ^ keys[j]);
  };
  const int d = (delta(i + 1) - delta(i - 1)) > 0 ? 1 : -1;
  const int dmin = delta(i - d);
  int64_t lmax = 2;
  while (delta(i + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (delta(i + (l + t) * d) > dmin) l += t;
  const int32_t j = (int32_t)(i + l * d);
  const int32_t first = min(i, j), last = max(i, j);
  const uint64_t kf = keys[first];
  const int common = __clzll(kf ^ keys[last]);
  int32_t split = first, step = last - first;
  do {
    step = (step + 1) >> 1;
    const int32_t ns = split + step;
    if (ns < last && (kf >> (63 - common)) == (keys[ns] >> (63 - common))) split = ns;
  } while (step > 1);
  const int32_t ca = (split == first) ? ~split : split;
  const int32_t cb = (split + 1 == last) ? ~(split + 1) : split + 1;
  nodes[i].childA = ca;
  nodes[i].childB = cb;
  nodes[i].other = j;
  if (ca < 0) leaf_parent[split] = i; else nodes[split].parent = i;
  if (cb < 0) leaf_parent[split + 1] = i; else nodes[split + 1].parent = i;
  if (i == 0) nodes[0].parent = -1;
}

}  // namespace cc
