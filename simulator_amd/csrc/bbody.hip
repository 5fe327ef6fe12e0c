// bbody.hip -- the bodies of a batch, posed and rebuilt together.
//
// Per frame, for all bodies at once:
//   k_bb_deform  grid over body vertices      capsule breathing + rigid motion
//   k_bb_boxes   grid over triangles          expanded boxes + per-block root boxes
//   k_bb_sort    one workgroup per body       root box, Morton codes, stable
//                                             4 x 8-bit LSD sort of (code, primitive)
//   k_bb_level   chunks of 1024 leaves, then  bottom-up tree and boxes, in LDS
//                of unfinished nodes
// The one-workgroup-per-body passes keep every cross-thread exchange inside
// a workgroup (LDS counters, barriers): bodies are 2k-40k triangles, so one
// workgroup per body still spreads the batch over the chip, and nothing
// needs an agent-scope arrival protocol (on MI355X each XCD has its own L2).
// Poses, keys, topology and boxes equal the single-body path bit for bit:
// the arithmetic is the shared routines of cc_lbvh.h, the sort is stable
// from ascending primitive order (so it orders the full 64-bit key), and
// box unions are exact.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "cc_bbody.h"
#include "cc_lbvh.h"
#include "cc_step.h"

namespace cc {

constexpr int BB_BLOCK = 256;       // vertices / triangles / nodes per block of the grid kernels
constexpr int SB_THREADS = 1024;    // threads of the per-body sort
constexpr int SB_WAVES = SB_THREADS / 64;
constexpr int SB_ROWS = 8;          // rows of 64 elements per wave per sort tile
constexpr int SB_TILE = SB_THREADS * SB_ROWS;
constexpr int SB_BATCH = 4;         // elements per thread per round of the code and key loops
constexpr int SB_BITS = 10;         // digit width: three passes cover the 30-bit Morton codes
constexpr int SB_DIGITS = 1 << SB_BITS;
constexpr int SB_PASSES = 3;
constexpr uint32_t BB_MAX_M = 1u << 19;   // triangles per batched body

// Barrier ordering LDS accesses only: the global stores issued before it
// stay in flight (a full __syncthreads waits for them).
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ------------------------------------------------------------ pose
__global__ __launch_bounds__(BB_BLOCK) void k_bb_deform(const BodyDesc* __restrict__ desc,
                                                        const uint4* __restrict__ blk,
                                                        const float* __restrict__ rest, float* __restrict__ verts,
                                                        const uint32_t* __restrict__ capv,
                                                        const float* __restrict__ capp0,
                                                        const float* __restrict__ capp1,
                                                        const float* __restrict__ motion, uint32_t nb, int k) {
  __shared__ float sP0[3 * MAX_CAPS], sP1[3 * MAX_CAPS], sM[BB_MF];
  __shared__ uint32_t sV[MAX_CAPS + 1];
  const uint4 b = blk[blockIdx.x];    // {body, first vertex, count, -}
  const BodyDesc d = desc[b.x];
  const uint32_t tid = threadIdx.x;
  for (uint32_t q = tid; q < 3 * d.ncaps; q += BB_BLOCK) {
    sP0[q] = capp0[3 * (size_t)d.c0 + q];
    sP1[q] = capp1[3 * (size_t)d.c0 + q];
  }
  for (uint32_t q = tid; q <= d.ncaps; q += BB_BLOCK) sV[q] = capv[d.cv0 + q];
  const float* mf = motion + ((size_t)k * nb + b.x) * BB_MF;
  for (uint32_t q = tid; q < 12 + d.ncaps; q += BB_BLOCK) sM[q] = mf[q];
  __syncthreads();
  if (tid >= b.z) return;
  const uint32_t v = b.y + tid;
  const uint32_t c = capsule_of(sV, d.ncaps, v);
  const size_t g = (size_t)d.v0 + v;
  deform_point(ld3(rest + 3 * g), ld3(sP0 + 3 * c), ld3(sP1 + 3 * c), sM[12 + c], sM, sM + 9, verts + 3 * g);
}

// ------------------------------------------------------------ boxes
__global__ __launch_bounds__(BB_BLOCK) void k_bb_boxes(const BodyDesc* __restrict__ desc,
                                                       const uint4* __restrict__ blk,
                                                       const float* __restrict__ verts,
                                                       const uint32_t* __restrict__ tris, float* __restrict__ boxes,
                                                       float* __restrict__ partial) {
  __shared__ float red[BB_BLOCK / 64][6];
  const uint4 b = blk[blockIdx.x];    // {body, first triangle, count, partial index}
  const BodyDesc d = desc[b.x];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  float mn[3], mx[3];
  box_sentinel(mn, mx);
  if (tid < b.z) {
    const size_t i = (size_t)d.t0 + b.y + tid;
    const uint32_t* t = tris + 3 * i;
    const float* V = verts + 3 * (size_t)d.v0;
    prim_box(ld3(V + 3 * (size_t)t[0]), ld3(V + 3 * (size_t)t[1]), ld3(V + 3 * (size_t)t[2]), d.depth, mn, mx);
    put6(boxes + 6 * i, mn, mx);
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int c = 0; c < 3; ++c) {
      mn[c] = fminf(mn[c], __shfl_xor(mn[c], off, 64));
      mx[c] = fmaxf(mx[c], __shfl_xor(mx[c], off, 64));
    }
  if (lane == 0) put6(red[w], mn, mx);
  __syncthreads();
  if (tid < 6) {
    float v = red[0][tid];
    for (int q = 1; q < BB_BLOCK / 64; ++q) v = tid < 3 ? fminf(v, red[q][tid]) : fmaxf(v, red[q][tid]);
    partial[6 * (size_t)b.w + tid] = v;
  }
}

// ------------------------------------------------------------ sort
// One workgroup per body.  Three passes of 10-bit digits (the Morton codes
// have 30 bits); the digit counts of all passes come from one sweep over the
// codes.  A pass walks tiles of SB_TILE elements; wave w owns rows
// w*SB_ROWS .. w*SB_ROWS+7 of 64 consecutive elements, ranks each row's
// lanes by digit with 10 ballots, and keeps a running per-digit count of its
// rows in LDS, so a tile needs three barriers whatever its size (LDS-only
// ones: a pass's scattered stores are waited for once, at its end).
// Element order = (wave, row, lane) order, so the pass is stable.
__device__ inline void bb_sort(uint32_t j, const BodyDesc* __restrict__ desc, const float* __restrict__ boxes,
                               const float* __restrict__ partial, uint2* kv0, uint2* kv1,
                               uint64_t* __restrict__ keys, uint4* __restrict__ leaf_tri,
                               const uint32_t* __restrict__ tris, float* __restrict__ root) {
  __shared__ uint32_t hist[SB_PASSES][SB_DIGITS];
  __shared__ uint32_t wcnt[SB_WAVES][SB_DIGITS];
  __shared__ uint32_t run[SB_DIGITS];
  __shared__ uint32_t wtot[SB_WAVES];
  __shared__ float red[SB_WAVES][6];
  __shared__ float sroot[6];
  const BodyDesc d = desc[j];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const uint32_t m = d.m, t0 = d.t0;

  // root box: sentinel U the body's per-block boxes (k_morton's reduction)
  float mn[3], mx[3];
  box_sentinel(mn, mx);
  const uint32_t np = (m + BB_BLOCK - 1) / BB_BLOCK;
  for (uint32_t q = tid; q < np; q += SB_THREADS) unite(mn, mx, partial + 6 * (size_t)(d.p0 + q));
  for (int off = 32; off > 0; off >>= 1)
    for (int c = 0; c < 3; ++c) {
      mn[c] = fminf(mn[c], __shfl_xor(mn[c], off, 64));
      mx[c] = fmaxf(mx[c], __shfl_xor(mx[c], off, 64));
    }
  if (lane == 0) put6(red[w], mn, mx);
  for (uint32_t q = tid; q < SB_PASSES * SB_DIGITS; q += SB_THREADS) (&hist[0][0])[q] = 0;
  __syncthreads();
  if (tid < 6) {
    float v = red[0][tid];
    for (int q = 1; q < SB_WAVES; ++q) v = tid < 3 ? fminf(v, red[q][tid]) : fmaxf(v, red[q][tid]);
    sroot[tid] = v;
    root[6 * (size_t)j + tid] = v;
  }
  __syncthreads();
  const float rmin[3] = {sroot[0], sroot[1], sroot[2]};
  const float rext[3] = {sroot[3] - sroot[0], sroot[4] - sroot[1], sroot[5] - sroot[2]};

  // codes (compute_morton_bbox, bvh.cu:71-79) and the digit counts
  // (SB_BATCH elements per thread per round, loads issued together)
  for (uint32_t base = 0; base < m; base += SB_BATCH * SB_THREADS) {
    float bxs[SB_BATCH][6];
#pragma unroll
    for (int u = 0; u < SB_BATCH; ++u) {
      const uint32_t t = min(base + u * SB_THREADS + tid, m - 1);
      const float* bx = boxes + 6 * (size_t)(t0 + t);
#pragma unroll
      for (int c = 0; c < 6; ++c) bxs[u][c] = bx[c];
    }
#pragma unroll
    for (int u = 0; u < SB_BATCH; ++u) {
      const uint32_t t = base + u * SB_THREADS + tid;
      if (t >= m) continue;
      const uint32_t c = box_morton(bxs[u], bxs[u] + 3, rmin, rext);
      kv0[t0 + t] = make_uint2(c, t);
#pragma unroll
      for (int p = 0; p < SB_PASSES; ++p) atomicAdd(&hist[p][(c >> (SB_BITS * p)) & (SB_DIGITS - 1)], 1u);
    }
  }
  __syncthreads();
  // exclusive scans of the counts, one pass at a time: thread = digit
  static_assert(SB_DIGITS == SB_THREADS, "one digit per thread in the scans");
  for (int p = 0; p < SB_PASSES; ++p) {
    const uint32_t v = hist[p][tid];
    uint32_t incl = v;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t u = __shfl_up(incl, off, 64);
      if (lane >= (uint32_t)off) incl += u;
    }
    if (lane == 63) wtot[w] = incl;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t q = 0; q < w; ++q) base += wtot[q];
    hist[p][tid] = base + incl - v;
    __syncthreads();
  }

  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (int p = 0; p < SB_PASSES; ++p) {
    const uint2* src = ((p & 1) ? kv1 : kv0) + t0;
    uint2* dst = ((p & 1) ? kv0 : kv1) + t0;
    const int shift = SB_BITS * p;
    run[tid] = hist[p][tid];
    for (uint32_t base = 0; base < m; base += SB_TILE) {
      for (uint32_t q = lane; q < (uint32_t)SB_DIGITS; q += 64) wcnt[w][q] = 0;
      uint2 e[SB_ROWS];
      uint32_t dg[SB_ROWS], loc[SB_ROWS];
      bool ok[SB_ROWS];
#pragma unroll
      for (int r = 0; r < SB_ROWS; ++r) {
        const uint32_t idx = base + (w * SB_ROWS + r) * 64 + lane;
        ok[r] = idx < m;
        e[r] = src[ok[r] ? idx : 0];
      }
#pragma unroll
      for (int r = 0; r < SB_ROWS; ++r) {
        dg[r] = (e[r].x >> shift) & (SB_DIGITS - 1);
        uint64_t msk = __ballot(ok[r]);
#pragma unroll
        for (int bit = 0; bit < SB_BITS; ++bit) {
          const bool on = (dg[r] >> bit) & 1u;
          const uint64_t bb = __ballot(on);
          msk &= on ? bb : ~bb;
        }
        // the wave's count of this digit in its earlier rows, then this
        // row's lanes of the digit; the lowest such lane adds the row's
        // count (reads of the row precede its write within the wave)
        const uint32_t prior = wcnt[w][dg[r]];
        const uint32_t rank = __popcll(msk & lt);
        loc[r] = prior + rank;
        if (ok[r] && rank == 0) wcnt[w][dg[r]] = prior + (uint32_t)__popcll(msk);
      }
      lds_barrier();
      {                                // per digit: offsets of the waves' runs
        uint32_t s = run[tid];
        for (int q = 0; q < SB_WAVES; ++q) {
          const uint32_t c = wcnt[q][tid];
          wcnt[q][tid] = s;
          s += c;
        }
        run[tid] = s;
      }
      lds_barrier();
#pragma unroll
      for (int r = 0; r < SB_ROWS; ++r) {
        const uint32_t pos = wcnt[w][dg[r]] + loc[r];
        if (ok[r] && pos < m) dst[pos] = e[r];       // pos < m always; the test keeps stores in the body
      }
      lds_barrier();                   // wcnt is reused; the scattered stores may still be in flight
    }
    __syncthreads();                   // the pass's stores are visible to the next pass's loads
  }
  // the passes end in kv1 (odd count) or kv0: keys (morton << 32 | primitive) and leaves
  const uint2* sorted = ((SB_PASSES & 1) ? kv1 : kv0) + t0;
  for (uint32_t base = 0; base < m; base += SB_BATCH * SB_THREADS) {
    uint2 e[SB_BATCH];
    uint32_t tv[SB_BATCH][3];
#pragma unroll
    for (int u = 0; u < SB_BATCH; ++u) e[u] = sorted[min(base + u * SB_THREADS + tid, m - 1)];
#pragma unroll
    for (int u = 0; u < SB_BATCH; ++u) {
      const uint32_t* t = tris + 3 * ((size_t)t0 + min(e[u].y, m - 1));
      tv[u][0] = t[0];
      tv[u][1] = t[1];
      tv[u][2] = t[2];
    }
#pragma unroll
    for (int u = 0; u < SB_BATCH; ++u) {
      const uint32_t i = base + u * SB_THREADS + tid;
      if (i >= m) continue;
      keys[t0 + i] = ((uint64_t)e[u].x << 32) | e[u].y;
      leaf_tri[t0 + i] = make_uint4(tv[u][0], tv[u][1], tv[u][2], e[u].y);
    }
  }
}

__global__ __launch_bounds__(SB_THREADS) void k_bb_sort(const BodyDesc* __restrict__ desc,
                                                        const uint32_t* __restrict__ list,
                                                        const float* __restrict__ boxes,
                                                        const float* __restrict__ partial, uint2* kv0, uint2* kv1,
                                                        uint64_t* __restrict__ keys, uint4* __restrict__ leaf_tri,
                                                        const uint32_t* __restrict__ tris,
                                                        float* __restrict__ root) {
  bb_sort(list[blockIdx.x], desc, boxes, partial, kv0, kv1, keys, leaf_tri, tris, root);
}

// The sort's workgroups, then one cloth step of the batch in workgroups of
// SB_THREADS / TILE tiles of TILE vertices (the sort of frame g + 1 and the
// step of frame g read and write disjoint buffers).
template <int EW, int TILE>
__global__ __launch_bounds__(SB_THREADS) void k_bb_sort_step(const BodyDesc* __restrict__ desc,
                                                             const uint32_t* __restrict__ list, uint32_t nsort,
                                                             const float* __restrict__ boxes,
                                                             const float* __restrict__ partial, uint2* kv0,
                                                             uint2* kv1, uint64_t* __restrict__ keys,
                                                             uint4* __restrict__ leaf_tri,
                                                             const uint32_t* __restrict__ tris,
                                                             float* __restrict__ root, StepLaunch st) {
  if (blockIdx.x < nsort) {
    bb_sort(list[blockIdx.x], desc, boxes, partial, kv0, kv1, keys, leaf_tri, tris, root);
    return;
  }
  const uint32_t e = (blockIdx.x - nsort) * (SB_THREADS / TILE) + threadIdx.x / TILE;
  if (e >= st.nblocks) return;
  const uint4 b = st.blocks[e];
  const uint32_t lt = threadIdx.x % TILE;
  if (lt >= b.z) return;
  const InstDesc dd = st.insts[b.x];
  const uint32_t i = b.y + lt;
  uint32_t visits = 0;
  step_vertex<false, EW, false, 6>(i, st.in, st.out, st.contact, st.contact, st.row, st.rec, dd.a, dd.bvh, spill_of(st.spill, i, st.ntotal),
                     visits, nullptr, nullptr, nullptr);
}

// ------------------------------------------------------------ tree + boxes
// Bottom-up construction of the Karras tree with its boxes (after Apetrei,
// "Fast and Simple Agglomerative LBVH Construction", 2014): a node covering
// sorted leaves [l, r] has its parent on the side whose boundary key pair
// shares the longer prefix (delta(r, r+1) > delta(l-1, l): the parent's split
// is r and the node is its left child, else the split is l-1 and the node is
// the right child; with unique keys the two never tie).  A left child's
// Karras index is its r, a right child's its l, the root's 0, so every node
// lands at the index processInternalNode gives it (cudaBRT.cu:9-161).  The
// second child to arrive at a split merges the two (calculateBoudingBox's
// sentinel U A U B, cudaBRT.cu:166-206) and climbs on.
//
// A workgroup takes up to LV_C adjacent "items" (leaves at level 0, else the
// nodes its input chunks could not finish) and merges them in LDS.  A node
// whose parent split lies outside the chunk, or whose sibling is not
// finished inside it, is passed to the next level (at most 2 x 63 per chunk:
// the children of the nodes crossing the chunk's two boundaries), LV_G
// chunks of it per workgroup, until one chunk holds the whole body.  The
// items of a chunk partition its leaf range in order, so the boundary deltas
// between adjacent items are those of the leaf keys there.
constexpr int LV_C = 1024;
constexpr int LV_CAP = 128;
constexpr int LV_G = LV_C / LV_CAP;
constexpr int LV_MAX_CLIMB = 70;       // > tree depth (< 64): a bound, never reached

struct Half {                          // one child at a split: far item | far leaf << 11, and its box
  uint32_t far;
  float box[6];
};
static_assert(LV_C <= 2048, "far item index has 11 bits");

__device__ inline int key_delta(const uint64_t* K, int32_t a) {   // delta(a, a+1), both in range
  return __clzll(K[a] ^ K[a + 1]);
}

__device__ inline void link_child(int32_t c, int32_t parent, Node* N, int32_t* LP) {
  if (c < 0) LP[~c] = parent; else N[c].parent = parent;
}

// children boxes (A then B) and links of node idx; the parent field is the
// parent's to write
__device__ inline void write_node(Node* N, int32_t idx, const float* bA, const float* bB, int32_t cA, int32_t cB,
                                  int32_t other) {
  float4* q = reinterpret_cast<float4*>(N + idx);
  q[0] = make_float4(bA[0], bA[1], bA[2], bA[3]);
  q[1] = make_float4(bA[4], bA[5], bB[0], bB[1]);
  q[2] = make_float4(bB[2], bB[3], bB[4], bB[5]);
  *reinterpret_cast<int2*>(&N[idx].childA) = make_int2(cA, cB);
  N[idx].other = other;
}

template <bool LEAVES>
__global__ __launch_bounds__(LV_C) void k_bb_level(const BodyDesc* __restrict__ desc, const LvBlock* __restrict__ blk,
                                                   PF<const uint64_t> keys_f, PF<const uint4> leaf_tri_f,
                                                   PF<const float> boxes_f, PF<const BItem> items_in_f,
                                                   PF<const uint32_t> cnt_in_f, PF<BItem> items_out_f,
                                                   PF<uint32_t> cnt_out_f, PF<Node> nodes_f,
                                                   PF<int32_t> leaf_parent_f, PF<float> root_f, PF<float> leaf0_f) {
  const uint64_t* __restrict__ keys = pick(keys_f);
  const uint4* __restrict__ leaf_tri = pick(leaf_tri_f);
  const float* __restrict__ boxes = pick(boxes_f);
  const BItem* __restrict__ items_in = pick(items_in_f);
  const uint32_t* __restrict__ cnt_in = pick(cnt_in_f);
  BItem* __restrict__ items_out = pick(items_out_f);
  uint32_t* __restrict__ cnt_out = pick(cnt_out_f);
  Node* __restrict__ nodes = pick(nodes_f);
  int32_t* __restrict__ leaf_parent = pick(leaf_parent_f);
  float* __restrict__ root = pick(root_f);
  float* __restrict__ leaf0 = pick(leaf0_f);
  __shared__ int32_t sD[LV_C + 1];       // sD[k]: delta at the boundary left of item k; sD[n]: right of the last
  __shared__ uint32_t arrived[LV_C / 16];   // split k (2 bits at k % 16): +1 left child arrived, +2 right child
  __shared__ int32_t sR[LV_C];           // last leaf of item k
  __shared__ Half half[LV_C][2];         // [split][0: left child, 1: right child]
  __shared__ BItem pend[LV_CAP];
  __shared__ uint32_t npend;
  __shared__ uint32_t soff[LV_G + 1];
  const LvBlock b = blk[blockIdx.x];
  const BodyDesc d = desc[b.body];
  const int32_t m = (int32_t)d.m;
  const uint32_t tid = threadIdx.x;
  const uint64_t* K = keys + d.t0;
  Node* N = nodes + d.n0;
  int32_t* LP = leaf_parent + d.t0;
  float* R = root + 6 * (size_t)b.body;
  if (m == 1) {                          // a single leaf: no internal node
    if (LEAVES && tid == 0) {
      const float* b0 = boxes + 6 * ((size_t)d.t0 + leaf_tri[d.t0].w);
      float mn[3], mx[3];
      box_sentinel(mn, mx);
      unite(mn, mx, b0);
      for (int c = 0; c < 6; ++c) leaf0[6 * (size_t)b.body + c] = b0[c];
      put6(R, mn, mx);
    }
    return;
  }
  // this thread's item
  int32_t n, il = 0, ir = 0, iidx = 0;
  float bx[6] = {0, 0, 0, 0, 0, 0};
  if (LEAVES) {
    const int32_t c0 = (int32_t)b.q * LV_C;
    n = min(LV_C, m - c0);
    if ((int32_t)tid < n) {
      il = ir = c0 + (int32_t)tid;
      iidx = ~il;
      const float* src = boxes + 6 * ((size_t)d.t0 + leaf_tri[d.t0 + il].w);
      for (int c = 0; c < 6; ++c) bx[c] = src[c];
    }
  } else {
    if (tid == 0) {
      uint32_t acc = 0;
      for (int g = 0; g < LV_G; ++g) {
        const uint32_t slot = b.q * LV_G + g;
        soff[g] = acc;
        acc += slot < b.n_in ? min(cnt_in[b.in_base + slot], (uint32_t)LV_CAP) : 0u;
      }
      soff[LV_G] = acc;
    }
    __syncthreads();
    n = (int32_t)soff[LV_G];
    if ((int32_t)tid < n) {
      int g = 0;
      while (g + 1 < LV_G && soff[g + 1] <= tid) ++g;
      const BItem& it = items_in[((size_t)b.in_base + b.q * LV_G + g) * LV_CAP + (tid - soff[g])];
      il = it.l;
      ir = it.r;
      iidx = it.idx;
      for (int c = 0; c < 6; ++c) bx[c] = it.box[c];
    }
  }
  if (tid == 0) npend = 0;
  if ((int32_t)tid < n) {
    sR[tid] = ir;
    sD[tid + 1] = ir + 1 < m ? key_delta(K, ir) : -1;
    if (tid == 0) sD[0] = il > 0 ? key_delta(K, il - 1) : -1;
  }
  if (tid < LV_C / 16) arrived[tid] = 0;
  __syncthreads();

  if ((int32_t)tid < n) {
    int32_t kl = (int32_t)tid, kr = (int32_t)tid, gl = il, gr = ir, me = iidx;
    bool merged = false;
    int32_t cA = 0, cB = 0;
    float bA[6], bB[6];
    for (int step = 0; step < LV_MAX_CLIMB; ++step) {
      const int dl = sD[kl], dr = sD[kr + 1];
      if (dl < 0 && dr < 0) {            // [0, m): the root, node 0
        if (merged) {
          write_node(N, 0, bA, bB, cA, cB, gr);
          N[0].parent = -1;
          link_child(cA, 0, N, LP);
          link_child(cB, 0, N, LP);
        }
        put6(R, bx, bx + 3);
        break;
      }
      const bool right = dr > dl;
      if (merged) {                      // the node's index follows from its side
        me = right ? gr : gl;
        write_node(N, me, bA, bB, cA, cB, right ? gl : gr);
        link_child(cA, me, N, LP);
        link_child(cB, me, N, LP);
      }
      if (right ? kr == n - 1 : kl == 0) {   // the parent's split lies outside this chunk
        const uint32_t p = atomicAdd(&npend, 1u);
        BItem& o = pend[min(p, (uint32_t)LV_CAP - 1)];
        o.l = gl;
        o.r = gr;
        o.idx = me;
        for (int c = 0; c < 6; ++c) o.box[c] = bx[c];
        break;
      }
      const int32_t sp = right ? kr : kl - 1;     // split between items sp and sp + 1
      Half& mine = half[sp][right ? 0 : 1];
      mine.far = (uint32_t)(right ? kl : kr) | ((uint32_t)(right ? gl : gr) << 11);
      for (int c = 0; c < 6; ++c) mine.box[c] = bx[c];
      __threadfence_block();
      const uint32_t sh = 2u * ((uint32_t)sp & 15u);
      const uint32_t old = atomicAdd(&arrived[(uint32_t)sp >> 4], (right ? 1u : 2u) << sh);
      if (((old >> sh) & 3u) == 0) break;          // first child here: the sibling merges
      __threadfence_block();
      const Half& o = half[sp][right ? 1 : 0];
      const int32_t ofk = (int32_t)(o.far & 2047u), ofg = (int32_t)(o.far >> 11);
      int32_t s;
      if (right) {                       // me = left child [gl, gr], sibling = [gr + 1, far leaf]
        s = gr;
        for (int c = 0; c < 6; ++c) {
          bA[c] = bx[c];
          bB[c] = o.box[c];
        }
        gr = ofg;
        kr = ofk;
      } else {                           // sibling = left child [far leaf, gl - 1], me = right child
        s = gl - 1;
        for (int c = 0; c < 6; ++c) {
          bA[c] = o.box[c];
          bB[c] = bx[c];
        }
        gl = ofg;
        kl = ofk;
      }
      cA = gl == s ? ~s : s;
      cB = s + 1 == gr ? ~(s + 1) : s + 1;
      float mn[3], mx[3];
      box_sentinel(mn, mx);
      unite(mn, mx, bA);
      unite(mn, mx, bB);
      put6(bx, mn, mx);
      merged = true;
    }
  }
  __syncthreads();
  // A child whose sibling did not finish in this chunk waits at its split:
  // it goes to the next level too (left child [far, s], right [s + 1, far]).
  if ((int32_t)tid < n - 1) {
    const uint32_t a = (arrived[tid >> 4] >> (2u * (tid & 15u))) & 3u;
    if (a == 1u || a == 2u) {
      const int side = a == 1u ? 0 : 1;
      const Half& w = half[tid][side];
      const int32_t s = sR[tid], fg = (int32_t)(w.far >> 11);
      const int32_t l = side == 0 ? fg : s + 1, r = side == 0 ? s : fg;
      const uint32_t p = atomicAdd(&npend, 1u);
      BItem& o = pend[min(p, (uint32_t)LV_CAP - 1)];
      o.l = l;
      o.r = r;
      o.idx = l < r ? (side == 0 ? r : l) : ~l;
      for (int c = 0; c < 6; ++c) o.box[c] = w.box[c];
    }
  }
  __syncthreads();
  // the unfinished nodes, in leaf order, to this chunk's output slot
  const uint32_t np = min(npend, (uint32_t)LV_CAP);
  if (tid < np) {
    const int32_t key = pend[tid].l;
    uint32_t rank = 0;
    for (uint32_t q = 0; q < np; ++q) rank += pend[q].l < key;
    items_out[((size_t)b.out_base + b.q) * LV_CAP + rank] = pend[tid];
  }
  if (tid == 0) cnt_out[b.out_base + b.q] = np;
}

// ------------------------------------------------------------ host side
size_t level_chunks0(uint32_t m) { return (m + LV_C - 1) / LV_C; }

std::vector<std::vector<LvBlock>> build_level_tables(const std::vector<BodyDesc>& h,
                                                     const std::vector<uint32_t>& bodies) {
  // level 0 over the leaves, level L+1 over LV_G chunks of level L, until a
  // body's level has a single chunk; slot bases are per level, in list order
  std::vector<std::vector<LvBlock>> out;
  std::vector<uint32_t> chunks(h.size(), 0), prev_base(h.size(), 0), prev_n(h.size(), 0);
  for (uint32_t j : bodies) chunks[j] = (uint32_t)level_chunks0(h[j].m);
  for (int L = 0; L < BB_MAX_LEVELS; ++L) {
    std::vector<LvBlock> t;
    uint32_t base = 0;
    std::vector<uint32_t> next(h.size(), 0), base_of(h.size(), 0);
    for (uint32_t j : bodies) {
      if (chunks[j] == 0) continue;
      base_of[j] = base;
      for (uint32_t q = 0; q < chunks[j]; ++q) {
        LvBlock e{};
        e.body = j;
        e.q = q;
        e.in_base = prev_base[j];
        e.out_base = base;
        e.n_in = L == 0 ? h[j].m : prev_n[j];
        t.push_back(e);
      }
      base += chunks[j];
      next[j] = chunks[j] > 1 ? (chunks[j] + LV_G - 1) / LV_G : 0;
    }
    if (t.empty()) break;
    out.push_back(std::move(t));
    prev_base = base_of;
    prev_n = chunks;
    chunks = next;
  }
  for (uint32_t j : bodies) CC_REQUIRE(chunks[j] == 0, "bottom-up build: too many levels");
  return out;
}

void launch_build_levels(const BodyDesc* desc, const LvBlock* const* level, const uint32_t* n_level,
                         const LvFrame* frames, int nframes, hipStream_t s) {
  CC_REQUIRE(nframes >= 1 && nframes <= MAX_FRAMES, "level build: frame count");
  PF<const uint64_t> keys{};
  PF<const uint4> leaf_tri{};
  PF<const float> boxes{};
  PF<BItem> items[2] = {};
  PF<const BItem> items_c[2] = {};
  PF<uint32_t> nitems[2] = {};
  PF<const uint32_t> nitems_c[2] = {};
  PF<Node> nodes{};
  PF<int32_t> leaf_parent{};
  PF<float> root{}, leaf0{};
  for (int f = 0; f < nframes; ++f) {
    const LvFrame& F = frames[f];
    keys.p[f] = F.keys;
    leaf_tri.p[f] = F.leaf_tri;
    boxes.p[f] = F.boxes;
    for (int q = 0; q < 2; ++q) {
      items[q].p[f] = F.items[q];
      items_c[q].p[f] = F.items[q];
      nitems[q].p[f] = F.nitems[q];
      nitems_c[q].p[f] = F.nitems[q];
    }
    nodes.p[f] = F.nodes;
    leaf_parent.p[f] = F.leaf_parent;
    root.p[f] = F.root;
    leaf0.p[f] = F.leaf0;
  }
  for (int L = 0; L < BB_MAX_LEVELS; ++L) {
    if (!n_level[L]) continue;
    const int in = (L + 1) & 1, out = L == 0 ? 0 : (L & 1);
    const dim3 grid(n_level[L], (unsigned)nframes);
    if (L == 0)
      k_bb_level<true><<<grid, LV_C, 0, s>>>(desc, level[L], keys, leaf_tri, boxes, items_c[in], nitems_c[in],
                                              items[out], nitems[out], nodes, leaf_parent, root, leaf0);
    else
      k_bb_level<false><<<grid, LV_C, 0, s>>>(desc, level[L], keys, leaf_tri, boxes, items_c[in], nitems_c[in],
                                               items[out], nitems[out], nodes, leaf_parent, root, leaf0);
    CC_LAUNCHED();
  }
}

BatchBodies::~BatchBodies() {
  for (int q = 0; q < 2; ++q) {
    if (ev_motion[q]) {
      (void)hipEventSynchronize(ev_motion[q]);
      (void)hipEventDestroy(ev_motion[q]);
    }
    if (h_motion[q]) (void)hipHostFree(h_motion[q]);
  }
}

namespace {

void append_blocks(std::vector<uint4>& t, uint32_t body, uint32_t count, uint32_t aux0, bool aux) {
  for (uint32_t f = 0, q = 0; f < count; f += BB_BLOCK, ++q)
    t.push_back(make_uint4(body, f, std::min<uint32_t>(BB_BLOCK, count - f), aux ? aux0 + q : 0u));
}

void upload_list(DBuf<uint4>& d, uint32_t& n, const std::vector<uint4>& h, hipStream_t s) {
  n = (uint32_t)h.size();
  d.upload(h.data(), std::max<size_t>(h.size(), 1), s);
  if (h.empty()) d.zero(s);
}

void upload_list(DBuf<uint32_t>& d, uint32_t& n, const std::vector<uint32_t>& h, hipStream_t s) {
  n = (uint32_t)h.size();
  d.upload(h.data(), std::max<size_t>(h.size(), 1), s);
  if (h.empty()) d.zero(s);
}

}  // namespace

void BatchBodies::init(const cc_body_desc* bodies, const cc_params* params, uint32_t count, hipStream_t s) {
  std::vector<float> h_rest, h_p0, h_p1;
  std::vector<uint32_t> h_tris, h_capv;
  uint32_t nv_sum = 0, m_sum = 0, n_sum = 0, c_sum = 0, cv_sum = 0, p_sum = 0;
  for (uint32_t k = 0; k < count; ++k) {
    const cc_body_desc* bd = bodies ? &bodies[k] : nullptr;
    if (!bd || bd->mode == CC_BODY_NONE) continue;
    CC_REQUIRE(bd->mode <= CC_BODY_REBUILD, "batch body: unknown mode");
    CC_REQUIRE(bd->verts && bd->tris && (bd->stride == 3 || bd->stride == 4) && bd->nv > 0 && bd->nf > 0,
               "batch body: vertices and triangles required");
    CC_REQUIRE(bd->nf <= BB_MAX_M, "batch body: at most 524288 triangles per body");
    for (size_t i = 0; i < 3 * (size_t)bd->nf; ++i)
      CC_REQUIRE(bd->tris[i] < bd->nv, "batch body: triangle index out of range");
    std::vector<float> bv;
    read_vec3_from(bd->verts, bd->stride, bd->nv, bv);
    if (bd->extend) {
      CC_REQUIRE(bd->normals3, "batch body: normals required for extend");
      int rc = cc_mesh_vertex_extend(bv.data(), 3, bd->nv, bd->tris, bd->nf, nullptr, bd->normals3,
                                     params[k].body_extend);
      if (rc) throw Error(rc, cc_last_error());
    }
    BodyDesc d{};
    d.nv = bd->nv;
    d.m = bd->nf;
    d.mode = bd->mode;
    d.depth = params[k].box_depth;
    d.v0 = nv_sum;
    d.t0 = m_sum;
    d.n0 = n_sum;
    d.c0 = c_sum;
    d.cv0 = cv_sum;
    d.p0 = p_sum;
    MotionParams mpk;
    std::vector<float> ph;
    if (bd->mode != CC_BODY_STATIC) {
      const cc_capsules* caps = bd->caps;
      CC_REQUIRE(caps && caps->ncaps >= 1 && caps->ncaps <= MAX_CAPS, "batch body: capsule layout required (1..64)");
      CC_REQUIRE(caps->vstart[0] == 0 && caps->vstart[caps->ncaps] == bd->nv,
                 "batch body: capsule ranges must cover vertices");
      for (uint32_t q = 0; q < caps->ncaps; ++q)
        CC_REQUIRE(caps->vstart[q] <= caps->vstart[q + 1], "batch body: capsule ranges must be ascending");
      d.ncaps = caps->ncaps;
      h_capv.insert(h_capv.end(), caps->vstart, caps->vstart + caps->ncaps + 1);
      h_p0.insert(h_p0.end(), caps->p0, caps->p0 + 3 * caps->ncaps);
      h_p1.insert(h_p1.end(), caps->p1, caps->p1 + 3 * caps->ncaps);
      ph.assign(caps->phase, caps->phase + caps->ncaps);
      mpk.amp = bd->amp;
      mpk.freq = bd->freq;
      mpk.rot_amp = bd->rot_amp;
      mpk.breath_amp = bd->breath_amp;
      mpk.breath_freq = bd->breath_freq;
      mpk.dx = bd->dir[0];
      mpk.dy = bd->dir[1];
      mpk.dz = bd->dir[2];
      mpk.px = bd->pivot[0];
      mpk.py = bd->pivot[1];
      mpk.pz = bd->pivot[2];
      mpk.dt = (double)params[k].dt;
      mpk.ncaps = d.ncaps;
      moving = true;
      c_sum += d.ncaps;
      cv_sum += d.ncaps + 1;
    }
    h_rest.insert(h_rest.end(), bv.begin(), bv.end());
    h_tris.insert(h_tris.end(), bd->tris, bd->tris + 3 * (size_t)bd->nf);
    CC_REQUIRE((uint64_t)nv_sum + d.nv < (1ull << 31) && (uint64_t)m_sum + d.m < (1ull << 31),
               "batch body: too many vertices or triangles in total");
    nv_sum += d.nv;
    m_sum += d.m;
    n_sum += std::max<uint32_t>(d.m, 2) - 1;
    p_sum += (d.m + BB_BLOCK - 1) / BB_BLOCK;
    h.push_back(d);
    inst.push_back(k);
    mp.push_back(mpk);
    phase.push_back(ph);
  }
  nb = (uint32_t)h.size();
  if (nb == 0) return;
  desc.upload(h.data(), nb, s);
  rest.upload(h_rest.data(), h_rest.size(), s);
  for (auto& b : verts) b.upload(h_rest.data(), h_rest.size(), s);
  tris.upload(h_tris.data(), h_tris.size(), s);
  capv.upload(h_capv.data(), std::max<size_t>(h_capv.size(), 1), s);
  capp0.upload(h_p0.data(), std::max<size_t>(h_p0.size(), 1), s);
  capp1.upload(h_p1.data(), std::max<size_t>(h_p1.size(), 1), s);
  boxes.alloc(6 * (size_t)m_sum);
  partial.alloc(6 * (size_t)p_sum);
  for (auto& b : kv) b.alloc(m_sum);
  keys.alloc(m_sum);
  for (int q = 0; q < 2; ++q) {
    leaf_tri[q].alloc(m_sum);
    leaf_parent[q].alloc(m_sum);
    leaf_parent[q].zero(s);
    nodes[q].alloc(n_sum);
    nodes[q].zero(s);
    root[q].alloc(6 * (size_t)nb);
    root[q].zero(s);
    leaf0[q].alloc(6 * (size_t)nb);
    leaf0[q].zero(s);
  }
  motion.alloc((size_t)BB_FRAMES * nb * BB_MF);
  motion.zero(s);
  for (int q = 0; q < 2; ++q) {
    if (fault_now(FAULT_ALLOC)) throw Error(CC_ERR_ALLOC, "hipHostMalloc: injected failure (cc_fault_inject)");
    CC_HIP(hipHostMalloc((void**)&h_motion[q], (size_t)BB_FRAMES * nb * BB_MF * sizeof(float), hipHostMallocDefault));
    std::memset(h_motion[q], 0, (size_t)BB_FRAMES * nb * BB_MF * sizeof(float));
    CC_HIP(hipEventCreate(&ev_motion[q]));
  }
  // work lists: largest body first for the per-body kernels
  std::vector<uint32_t> order(nb);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return h[a].m > h[b].m; });
  size_t chunks0 = 0;
  for (uint32_t j = 0; j < nb; ++j) chunks0 += level_chunks0(h[j].m);
  for (auto& b : items) b.alloc(chunks0 * LV_CAP);
  for (auto& b : nitems) {
    b.alloc(chunks0);
    b.zero(s);
  }
  for (int which = 0; which < 2; ++which) {
    BodyPass& P = which == 0 ? all : step;
    std::vector<uint4> tdef, tbox;
    std::vector<uint32_t> lsort, lbuild;
    for (uint32_t j : order) {
      const BodyDesc& d = h[j];
      const bool mov = d.mode != CC_BODY_STATIC;
      if (!(which == 0 || mov)) continue;
      if (mov) append_blocks(tdef, j, d.nv, 0, false);
      append_blocks(tbox, j, d.m, d.p0, true);
      if (which == 0 || d.mode == CC_BODY_REBUILD) lsort.push_back(j);
      lbuild.push_back(j);
    }
    upload_list(P.deform, P.n_deform, tdef, s);
    upload_list(P.boxes, P.n_boxes, tbox, s);
    upload_list(P.sort, P.n_sort, lsort, s);
    const std::vector<std::vector<LvBlock>> t = build_level_tables(h, lbuild);
    for (int L = 0; L < BB_MAX_LEVELS; ++L) {
      P.n_level[L] = L < (int)t.size() ? (uint32_t)t[L].size() : 0u;
      if (P.n_level[L]) P.level[L].upload(t[L].data(), t[L].size(), s);
    }
  }
  // frame 0 for every body, in both slots
  upload_motion(0, 1, s);
  launch(all, 0, 0, s);
  launch(all, 0, 1, s);
}

BvhView BatchBodies::view(uint32_t j, int slot) const {
  const BodyDesc& d = h[j];
  BvhView v;
  v.nodes = nodes[slot].p + d.n0;
  v.leaf_tri = leaf_tri[slot].p + d.t0;
  v.verts = verts[slot].p + 3 * (size_t)d.v0;
  v.leaf0_box = leaf0[slot].p + 6 * (size_t)j;
  v.m = d.m;
  return v;
}

void BatchBodies::upload_motion(uint64_t frame0, int nframes, hipStream_t s) {
  if (!moving) return;
  CC_REQUIRE(nframes >= 1 && nframes <= BB_FRAMES, "batch body: motion frames");
  CC_HIP(hipEventSynchronize(ev_motion[hm]));        // the staging buffer's last copy is done
  float* hmv = h_motion[hm];
  for (int f = 0; f < nframes; ++f)
    for (uint32_t j = 0; j < nb; ++j) {
      if (h[j].mode == CC_BODY_STATIC) continue;
      const MotionFrame mf = motion_of(mp[j], phase[j].data(), frame0 + (uint64_t)f);
      std::memcpy(hmv + ((size_t)f * nb + j) * BB_MF, mf.v, (12 + h[j].ncaps) * sizeof(float));
    }
  CC_HIP(hipMemcpyAsync(motion.p, hmv, (size_t)nframes * nb * BB_MF * sizeof(float), hipMemcpyHostToDevice, s));
  CC_HIP(hipEventRecord(ev_motion[hm], s));
  hm ^= 1;
}

void BatchBodies::pre(const BodyPass& P, int k, int slot, hipStream_t s) {
  if (nb == 0) return;
  if (P.n_deform)
    k_bb_deform<<<P.n_deform, BB_BLOCK, 0, s>>>(desc.p, P.deform.p, rest.p, verts[slot].p, capv.p, capp0.p,
                                                 capp1.p, motion.p, nb, k);
  if (P.n_boxes)
    k_bb_boxes<<<P.n_boxes, BB_BLOCK, 0, s>>>(desc.p, P.boxes.p, verts[slot].p, tris.p, boxes.p, partial.p);
  CC_LAUNCHED();
}

void BatchBodies::sort(const BodyPass& P, int slot, hipStream_t s) {
  if (nb == 0 || !P.n_sort) return;
  k_bb_sort<<<P.n_sort, SB_THREADS, 0, s>>>(desc.p, P.sort.p, boxes.p, partial.p, kv[0].p, kv[1].p, keys.p,
                                             leaf_tri[slot].p, tris.p, root[slot].p);
  CC_LAUNCHED();
}

template <int EW, int TILE>
static void launch_sort_step(BatchBodies& B, const BodyPass& P, int slot, const StepLaunch& st, hipStream_t s) {
  const uint32_t per = SB_THREADS / TILE;
  const uint32_t nstep = (st.nblocks + per - 1) / per;
  k_bb_sort_step<EW, TILE><<<P.n_sort + nstep, SB_THREADS, 0, s>>>(
      B.desc.p, P.sort.p, P.n_sort, B.boxes.p, B.partial.p, B.kv[0].p, B.kv[1].p, B.keys.p, B.leaf_tri[slot].p,
      B.tris.p, B.root[slot].p, st);
}

void BatchBodies::sort_with_step(const BodyPass& P, int slot, const StepLaunch& st, hipStream_t s) {
  CC_REQUIRE(st.tile == 64 || st.tile == 128 || st.tile == 256, "batch: step tile must be 64, 128 or 256");
  if (st.ew == ELL_W) {
    if (st.tile == 64) launch_sort_step<ELL_W, 64>(*this, P, slot, st, s);
    else if (st.tile == 128) launch_sort_step<ELL_W, 128>(*this, P, slot, st, s);
    else launch_sort_step<ELL_W, 256>(*this, P, slot, st, s);
  } else {
    if (st.tile == 64) launch_sort_step<0, 64>(*this, P, slot, st, s);
    else if (st.tile == 128) launch_sort_step<0, 128>(*this, P, slot, st, s);
    else launch_sort_step<0, 256>(*this, P, slot, st, s);
  }
  CC_LAUNCHED();
}

void BatchBodies::post(const BodyPass& P, int slot, hipStream_t s) {
  if (nb == 0) return;
  const LvBlock* tabs[BB_MAX_LEVELS];
  for (int L = 0; L < BB_MAX_LEVELS; ++L) tabs[L] = P.level[L].p;
  const LvFrame f{keys.p, leaf_tri[slot].p, boxes.p, {items[0].p, items[1].p}, {nitems[0].p, nitems[1].p},
                  nodes[slot].p, leaf_parent[slot].p, root[slot].p, leaf0[slot].p};
  launch_build_levels(desc.p, tabs, P.n_level, &f, 1, s);
}

void BatchBodies::export_tree(uint32_t j, int slot, float* verts3, uint64_t* okeys, uint32_t* order,
                              int32_t* childA, int32_t* childB, int32_t* parent_int, int32_t* parent_leaf,
                              float* box_int, float* box_leaf, float* root6, hipStream_t s) {
  CC_REQUIRE(j < nb, "batch body: no such body");
  const BodyDesc& d = h[j];
  if (verts3)
    CC_HIP(hipMemcpyAsync(verts3, verts[slot].p + 3 * (size_t)d.v0, 3 * (size_t)d.nv * sizeof(float),
                          hipMemcpyDeviceToHost, s));
  export_tree_raw(d.m, keys.p + d.t0, leaf_tri[slot].p + d.t0, leaf_parent[slot].p + d.t0, nodes[slot].p + d.n0,
                  root[slot].p + 6 * (size_t)j, leaf0[slot].p + 6 * (size_t)j, okeys, order, childA, childB,
                  parent_int, parent_leaf, box_int, box_leaf, root6, s);
}

}  // namespace cc
