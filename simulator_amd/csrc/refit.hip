// refit.hip -- bottom-up LBVH box refit on gfx950.
//
// calculateBoudingBox (cudaBRT.cu:166-206) gives every internal node the box
// sentinel U A U B, computed bottom-up by the second of its two children to
// arrive.  Here the same boxes are computed level by level:
//
//  * The sorted leaves are cut into chunks of RC_CHUNK.  An internal node
//    whose leaf range lies inside one chunk is "in-chunk"; the others (their
//    range crosses a chunk boundary) are "top" nodes.
//  * Once per topology (after a build) the setup kernels record, for every
//    in-chunk node, its height inside the chunk (1 + the larger height of
//    its internal children), for every top node its height inside the top
//    tree, the list of top nodes with their links, and which nodes and
//    leaves hang directly below a top node.
//  * Per refit, k_refit_chunks has one workgroup per chunk compute leaf boxes
//    into LDS and then in-chunk boxes one height level per barrier; the
//    children of a level-h node all have smaller heights.  k_refit_top does
//    the same for the top nodes in one workgroup.  No atomics and no
//    cross-workgroup synchronisation: on MI355X an agent-scope arrival
//    protocol pays an L2 write-back / invalidate per level (each XCD has its
//    own L2); the kernel boundary between the passes gives the visibility.
//  * All boxes go to their parent's child slot (the layout the traversal
//    reads).  fmin/fmax unions are exact, so the order does not matter.
#include "cc_lbvh.h"

namespace cc {

constexpr int RC_CHUNK = REFIT_CHUNK;
// 1024 threads, two leaves and two nodes each (512 threads, four each: 51.0
// against 49.7 us per refit at the 501k-triangle body; 43.2 against 42.2 us
// per pipelined step in contact)
constexpr int RC_THREADS = 1024;
constexpr int RC_PER_THREAD = RC_CHUNK / RC_THREADS;
constexpr int TOP_THREADS = 1024;
// The per-frame top pass is ONE workgroup of 256 threads with ~100 KB of
// LDS (a single step block retiring leaves room for it on a CU).  Measured
// at the 501k-triangle refitted body: the pipelined step is the same as with
// 1024 threads and 147 KB (44.1-47.1 vs 43.5-48.4 us per step); in
// isolation the pass takes ~25 us, of which loading the chunk results ~6,
// the 30 levels ~10.5 and storing the boxes into the four-wide records ~8.5.
constexpr int TOP_REFIT_THREADS = 256;
constexpr int TOP_LOAD = 7;                                         // entries per thread per load round
constexpr int TOP_MAX_LDS = 3200;                                   // entries held in LDS (32 B each)
constexpr int TOP_PER_THREAD = (TOP_MAX_LDS + TOP_REFIT_THREADS - 1) / TOP_REFIT_THREADS;
constexpr int TOP_MAX_LEVELS = 64;                                  // top-tree height bound (depth < 64)
constexpr uint8_t BELOW_TOP = 0x80;                                 // flag: parent is a top node
constexpr uint8_t IS_A = 0x40;                                      // flag: child A of the parent
constexpr uint8_t H_MASK = 0x3f;                                    // in-chunk height (depth < 64)

// ======================================================== topology setup
// Heights are found with an arrival climb (run once per topology, so its
// latency does not matter): the second child to arrive at a node knows both
// child heights.

// top_pos[x] = list position of top node x, -1 for every other internal node
__global__ __launch_bounds__(RC_THREADS) void k_top_list(const Node* __restrict__ nodes, uint32_t m,
                                                         int32_t* __restrict__ top_list,
                                                         int32_t* __restrict__ top_pos,
                                                         uint32_t* __restrict__ top_count) {
  __shared__ uint32_t s_ntop, s_base;
  const int32_t c0 = (int32_t)(blockIdx.x * RC_CHUNK);
  const int32_t i1 = min((int32_t)m - 1, c0 + RC_CHUNK);
  const uint32_t tid = threadIdx.x;
  if (tid == 0) s_ntop = 0;
  __syncthreads();
  bool top[RC_PER_THREAD];
  uint32_t loc[RC_PER_THREAD];
  for (int q = 0; q < RC_PER_THREAD; ++q) {
    const int32_t x = c0 + q * RC_THREADS + (int32_t)tid;
    top[q] = false;
    if (x >= i1) continue;
    const int32_t o = nodes[x].other;
    top[q] = o < c0 || o >= c0 + RC_CHUNK;
    if (top[q]) loc[q] = atomicAdd(&s_ntop, 1u);
  }
  __syncthreads();
  if (tid == 0) s_base = s_ntop ? atomicAdd(top_count, s_ntop) : 0u;
  __syncthreads();
  for (int q = 0; q < RC_PER_THREAD; ++q) {
    const int32_t x = c0 + q * RC_THREADS + (int32_t)tid;
    if (x >= i1) continue;
    int32_t pos = -1;
    if (top[q]) {
      pos = (int32_t)(s_base + loc[q]);
      top_list[pos] = x;
    }
    top_pos[x] = pos;
  }
}

// in-chunk heights (low 7 bits of nflag) and the below-top flags of nodes
// and leaves; chunk_h[w] = largest in-chunk height of chunk w
__global__ __launch_bounds__(RC_THREADS) void k_chunk_heights(const Node* __restrict__ nodes,
                                                              const int32_t* __restrict__ leaf_parent, uint32_t m,
                                                              const int32_t* __restrict__ top_pos,
                                                              uint8_t* __restrict__ nflag,
                                                              uint8_t* __restrict__ lflag,
                                                              uint32_t* __restrict__ chunk_h) {
  __shared__ int32_t spar[RC_CHUNK], soth[RC_CHUNK];
  __shared__ uint32_t scnt[RC_CHUNK], shA[RC_CHUNK], shB[RC_CHUNK];
  __shared__ int32_t sca[RC_CHUNK];
  __shared__ uint32_t s_hmax;
  const int32_t c0 = (int32_t)(blockIdx.x * RC_CHUNK);
  const int32_t c1 = min((int32_t)m, c0 + RC_CHUNK);
  const int32_t i1 = min((int32_t)m - 1, c0 + RC_CHUNK);
  const uint32_t tid = threadIdx.x;
  auto inside = [&](int32_t v) { return v >= c0 && v < c0 + RC_CHUNK; };
  if (tid == 0) s_hmax = 0;
  for (int32_t x = c0 + (int32_t)tid; x < i1; x += RC_THREADS) {
    const int4 L = reinterpret_cast<const int4*>(nodes + x)[3];
    sca[x - c0] = L.x;
    spar[x - c0] = L.z;
    soth[x - c0] = L.w;
    scnt[x - c0] = 0;
  }
  __syncthreads();
  auto in_chunk = [&](int32_t x) { return inside(x) && inside(soth[x - c0]); };
  // one climb per leaf, height 0 at the leaf
  for (int32_t l = c0 + (int32_t)tid; l < c1; l += RC_THREADS) {
    int32_t x = leaf_parent[l], child = ~l;
    uint32_t h = 0;
    lflag[l] = in_chunk(x) ? 0 : (uint8_t)(BELOW_TOP | (nodes[x].childA == child ? IS_A : 0));
    while (in_chunk(x)) {
      const bool isA = sca[x - c0] == child;
      (isA ? shA : shB)[x - c0] = h;
      __threadfence_block();
      if (atomicAdd(&scnt[x - c0], 1u) == 0) break;
      __threadfence_block();
      h = 1 + max(shA[x - c0], shB[x - c0]);
      const int32_t p = spar[x - c0];
      uint8_t f = (uint8_t)h;
      if (p < 0 || !in_chunk(p)) f |= BELOW_TOP | ((p >= 0 && nodes[p].childA == x) ? IS_A : 0);
      nflag[x] = f;
      atomicMax(&s_hmax, h);
      child = x;
      x = p;
    }
  }
  __syncthreads();
  if (tid == 0) chunk_h[blockIdx.x] = s_hmax;
  // top nodes carry no in-chunk height
  for (int32_t x = c0 + (int32_t)tid; x < i1; x += RC_THREADS)
    if (!in_chunk(x)) nflag[x] = 0;
  (void)top_pos;
}

// links[i] = {position of top child A or -1, of top child B or -1, of the
// parent or -1, parent slot 2*parent + (A ? 0 : 1) or -1 at the root}
__global__ __launch_bounds__(256) void k_top_links(const Node* __restrict__ nodes,
                                                   const int32_t* __restrict__ top_list,
                                                   const int32_t* __restrict__ top_pos,
                                                   const uint32_t* __restrict__ top_count,
                                                   int4* __restrict__ links) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= *top_count) return;
  const int32_t x = top_list[i];
  const Node& n = nodes[x];
  const int32_t pa = n.childA >= 0 ? top_pos[n.childA] : -1;
  const int32_t pb = n.childB >= 0 ? top_pos[n.childB] : -1;
  int32_t pp = -1, slot = -1;
  if (n.parent >= 0) {
    pp = top_pos[n.parent];
    slot = 2 * n.parent + (nodes[n.parent].childA == x ? 0 : 1);
  }
  links[i] = make_int4(pa, pb, pp, slot);
}

// top-tree heights (1 + larger height of top children, top leaves of the top
// tree have height 1); one workgroup, arrival climb on counters in `work`
__global__ __launch_bounds__(TOP_THREADS) void k_top_heights(const int4* __restrict__ links,
                                                             const uint32_t* __restrict__ top_count,
                                                             uint32_t* __restrict__ work,
                                                             uint8_t* __restrict__ top_h,
                                                             uint32_t* __restrict__ top_hmax) {
  __shared__ uint32_t s_hmax;
  const uint32_t T = *top_count;
  uint32_t* cnt = work;
  uint32_t* hA = work + T;
  uint32_t* hB = work + 2 * (size_t)T;
  if (threadIdx.x == 0) s_hmax = 0;
  for (uint32_t i = threadIdx.x; i < T; i += TOP_THREADS) {
    const int4 L = links[i];
    cnt[i] = (L.x < 0 ? 1u : 0u) + (L.y < 0 ? 1u : 0u);
    hA[i] = 0;
    hB[i] = 0;
  }
  __syncthreads();
  for (uint32_t i0 = threadIdx.x; i0 < T; i0 += TOP_THREADS) {
    int4 L = links[i0];
    if (L.x >= 0 || L.y >= 0) continue;
    uint32_t i = i0, h = 1;
    while (true) {
      top_h[i] = (uint8_t)h;
      atomicMax(&s_hmax, h);
      const int32_t j = L.z;
      if (j < 0) break;
      const int4 P = links[j];
      (P.x == (int32_t)i ? hA : hB)[j] = h;
      __threadfence_block();
      if (atomicAdd(&cnt[j], 1u) != 1) break;
      __threadfence_block();
      h = 1 + max(hA[j], hB[j]);
      i = (uint32_t)j;
      L = P;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) *top_hmax = s_hmax;
}

// The link fields of the node records, compact: a refit reads 16 bytes per
// node instead of the 128-byte lines that hold its 64-byte record.
__global__ __launch_bounds__(256) void k_node_links(const Node* __restrict__ nodes, uint32_t mi,
                                                   int4* __restrict__ links) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < mi) links[i] = reinterpret_cast<const int4*>(nodes + i)[3];
}

// ======================================================== per-frame refit
__global__ __launch_bounds__(RC_THREADS) void k_refit_chunks(
    PF<const float> verts_f, const uint4* __restrict__ leaf_tri, const int32_t* __restrict__ leaf_parent,
    const int4* __restrict__ links, PF<Node> nodes_f, uint32_t m, float depth, const uint8_t* __restrict__ nflag,
    const uint8_t* __restrict__ lflag, const uint32_t* __restrict__ chunk_h, const int32_t* __restrict__ top_pos,
    PF<float> top_child_f, PF<float> root_f, PF<float> leaf0_f, PF<float4> wide_f,
    const uint2* __restrict__ wdst) {
  const float* __restrict__ verts = pick(verts_f);
  Node* __restrict__ nodes = pick(nodes_f);
  float* __restrict__ top_child = pick(top_child_f);
  float* __restrict__ root_out = pick(root_f);
  float* __restrict__ leaf0_box = pick(leaf0_f);
  float4* __restrict__ wide = pick(wide_f);
  __shared__ float lbox[RC_CHUNK][6];
  __shared__ float nbox[RC_CHUNK][6];
  const int32_t c0 = (int32_t)(blockIdx.x * RC_CHUNK);
  const int32_t c1 = min((int32_t)m, c0 + RC_CHUNK);            // leaves [c0, c1)
  const int32_t i1 = min((int32_t)m - 1, c0 + RC_CHUNK);        // internal nodes [c0, i1)
  const uint32_t tid = threadIdx.x;
  const uint32_t H = m > 1 ? chunk_h[blockIdx.x] : 0;

  // Loads in two batches, each issued in full before its first use (indices
  // clamped instead of guarded: a guard per entry lets the compiler sink
  // each load into its branch, one memory round trip per entry).
  // First batch: the thread's leaves (triangle, flags, parent) and internal
  // nodes (links, flags).
  const int32_t lmax = c1 - 1, xmax = max(i1 - 1, 0);
  const int32_t tmax = max((int32_t)m - 2, 0);                  // last entry of top_pos
  uint4 tri[RC_PER_THREAD];
  int4 L[RC_PER_THREAD];
  uint8_t fl[RC_PER_THREAD], lf[RC_PER_THREAD];
  int32_t lpar[RC_PER_THREAD];
  uint4 wd[RC_PER_THREAD];         // wide entries mirroring the node's two child slots
#pragma unroll
  for (int r = 0; r < RC_PER_THREAD; ++r) {
    const int32_t l = c0 + r * RC_THREADS + (int32_t)tid;
    const int32_t lc = min(l, lmax), xc = min(l, xmax);
    tri[r] = leaf_tri[lc];
    lf[r] = lflag[lc];
    lpar[r] = leaf_parent[lc];
    L[r] = links[xc];
    fl[r] = nflag[xc];
    wd[r] = wide ? reinterpret_cast<const uint4*>(wdst)[xc] : make_uint4(~0u, 0u, ~0u, 0u);
  }
#pragma unroll
  for (int r = 0; r < RC_PER_THREAD; ++r) {
    const int32_t x = c0 + r * RC_THREADS + (int32_t)tid;
    if (x >= i1) {
      L[r] = make_int4(0, 0, -1, 0);
      fl[r] = 0;
    }
    if (x >= c1) lf[r] = 0;
  }
  // Second batch: the leaves' vertices, and for every box that goes directly to
  // a top node, that node's top-pass record and (four-wide form) the wide
  // entry of its slot.
  float vtx[RC_PER_THREAD][9];
  int32_t tpos_n[RC_PER_THREAD], tpos_l[RC_PER_THREAD];
  uint2 wtop_n[RC_PER_THREAD], wtop_l[RC_PER_THREAD];
#pragma unroll
  for (int r = 0; r < RC_PER_THREAD; ++r) {
    const uint32_t t3[3] = {tri[r].x, tri[r].y, tri[r].z};
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) vtx[r][3 * k + c] = verts[3 * (size_t)t3[k] + c];
    const int32_t pn = min(max(L[r].z, 0), tmax), pl = min(max(lpar[r], 0), tmax);
    tpos_n[r] = top_pos[pn];
    tpos_l[r] = top_pos[pl];
    wtop_n[r] = wide ? wdst[2 * (size_t)pn + ((fl[r] & IS_A) ? 0 : 1)] : make_uint2(~0u, 0u);
    wtop_l[r] = wide ? wdst[2 * (size_t)pl + ((lf[r] & IS_A) ? 0 : 1)] : make_uint2(~0u, 0u);
  }
#pragma unroll
  for (int r = 0; r < RC_PER_THREAD; ++r) {
    const int32_t l = c0 + r * RC_THREADS + (int32_t)tid;
    if (l >= c1) continue;
    float mn[3], mx[3];
    prim_box(ld3(vtx[r]), ld3(vtx[r] + 3), ld3(vtx[r] + 6), depth, mn, mx);
    put6(lbox[l - c0], mn, mx);
  }
  if (m == 1) {
    if (tid == 0) {
      for (int c = 0; c < 6; ++c) leaf0_box[c] = lbox[0][c];
      float rmn[3] = {100.0f, 100.0f, 100.0f}, rmx[3] = {-100.0f, -100.0f, -100.0f};
      unite(rmn, rmx, lbox[0]);
      put6(root_out, rmn, rmx);
    }
    return;
  }
  __syncthreads();
  // in-chunk boxes, one height level per barrier
  for (uint32_t h = 1; h <= H; ++h) {
    for (int r = 0; r < RC_PER_THREAD; ++r) {
      if ((fl[r] & H_MASK) != h) continue;
      const int32_t x = c0 + r * RC_THREADS + (int32_t)tid;
      float mn[3] = {100.0f, 100.0f, 100.0f}, mx[3] = {-100.0f, -100.0f, -100.0f};
      unite(mn, mx, L[r].x < 0 ? lbox[~L[r].x - c0] : nbox[L[r].x - c0]);
      unite(mn, mx, L[r].y < 0 ? lbox[~L[r].y - c0] : nbox[L[r].y - c0]);
      put6(nbox[x - c0], mn, mx);
    }
    __syncthreads();
  }
  // stores: both child slots of every in-chunk node; boxes of nodes / leaves
  // directly below a top node to that node's slot and its top-pass record
#pragma unroll
  for (int r = 0; r < RC_PER_THREAD; ++r) {
    const int32_t x = c0 + r * RC_THREADS + (int32_t)tid;
    if (x < i1 && (fl[r] & H_MASK)) {
      const float* a = L[r].x < 0 ? lbox[~L[r].x - c0] : nbox[L[r].x - c0];
      const float* b = L[r].y < 0 ? lbox[~L[r].y - c0] : nbox[L[r].y - c0];
      if (wide) {                         // the four-wide form only
        put_wide_at(wide, make_uint2(wd[r].x, wd[r].y), a);
        put_wide_at(wide, make_uint2(wd[r].z, wd[r].w), b);
      } else {
        float4* d = reinterpret_cast<float4*>(nodes + x);
        d[0] = make_float4(a[0], a[1], a[2], a[3]);
        d[1] = make_float4(a[4], a[5], b[0], b[1]);
        d[2] = make_float4(b[2], b[3], b[4], b[5]);
      }
      if (fl[r] & BELOW_TOP) {
        const int32_t p = L[r].z;
        const float* bx = nbox[x - c0];
        if (p < 0) {                      // the whole tree lies in this chunk
          for (int c = 0; c < 6; ++c) root_out[c] = bx[c];
        } else {
          const bool isA = (fl[r] & IS_A) != 0;
          if (wide)
            put_wide_at(wide, wtop_n[r], bx);
          else
            put6(isA ? nodes[p].a_min : nodes[p].b_min, bx, bx + 3);
          put6(top_child + 12 * (size_t)tpos_n[r] + (isA ? 0 : 6), bx, bx + 3);
        }
      }
    }
    const int32_t l = c0 + r * RC_THREADS + (int32_t)tid;
    if (l < c1 && (lf[r] & BELOW_TOP)) {
      const int32_t p = lpar[r];
      const bool isA = (lf[r] & IS_A) != 0;
      if (wide)
        put_wide_at(wide, wtop_l[r], lbox[l - c0]);
      else
        put6(isA ? nodes[p].a_min : nodes[p].b_min, lbox[l - c0], lbox[l - c0] + 3);
      put6(top_child + 12 * (size_t)tpos_l[r] + (isA ? 0 : 6), lbox[l - c0], lbox[l - c0] + 3);
    }
  }
}

// Setup, one workgroup: renumber the top entries by height (counting sort)
// so that each level is a contiguous range [lev[h], lev[h+1]); positions in
// top_list / top_pos / top_links are rewritten to the sorted numbering.
__global__ __launch_bounds__(TOP_THREADS) void k_top_sort(int32_t* __restrict__ top_list,
                                                          int32_t* __restrict__ top_pos,
                                                          int4* __restrict__ links,
                                                          const uint8_t* __restrict__ top_h,
                                                          const uint32_t* __restrict__ top_count,
                                                          const uint32_t* __restrict__ top_hmax,
                                                          uint32_t* __restrict__ lev, int32_t* __restrict__ work) {
  __shared__ uint32_t cnt[TOP_MAX_LEVELS + 2], fill[TOP_MAX_LEVELS + 2];
  const uint32_t T = *top_count, H = *top_hmax, tid = threadIdx.x;
  int32_t* old_list = work;                                       // [T]
  int4* old_links = reinterpret_cast<int4*>(work + T + (T & 3 ? 4 - (T & 3) : 0));   // [T], 16 B aligned
  int32_t* newpos = reinterpret_cast<int32_t*>(old_links + T);   // [T]
  for (uint32_t h = tid; h < TOP_MAX_LEVELS + 2; h += TOP_THREADS) {
    cnt[h] = 0;
    fill[h] = 0;
  }
  __syncthreads();
  for (uint32_t i = tid; i < T; i += TOP_THREADS) {
    old_list[i] = top_list[i];
    old_links[i] = links[i];
    atomicAdd(&cnt[top_h[i]], 1u);
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (uint32_t h = 1; h <= H + 1; ++h) {
      const uint32_t c = h <= H ? cnt[h] : 0;
      cnt[h] = run;                   // start of level h
      lev[h] = run;
      run += c;
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < T; i += TOP_THREADS) {
    const uint32_t h = top_h[i];
    newpos[i] = (int32_t)(cnt[h] + atomicAdd(&fill[h], 1u));
  }
  __syncthreads();
  for (uint32_t i = tid; i < T; i += TOP_THREADS) {
    const int32_t p = newpos[i];
    const int4 L = old_links[i];
    top_list[p] = old_list[i];
    top_pos[old_list[i]] = p;
    links[p] = make_int4(L.x >= 0 ? newpos[L.x] : -1, L.y >= 0 ? newpos[L.y] : -1,
                         L.z >= 0 ? newpos[L.z] : -1, L.w);
  }
}

// Per refit, one workgroup: entry e's box = sentinel U its non-top children
// (top_child) U its top children (lower levels); one level per barrier, one
// entry per thread within a level.  LDS holds the boxes and child links;
// the slot a box goes to is read again from `links` when it is stored.
template <bool LDS>
__device__ void refit_top_levels(Node* __restrict__ nodes, uint32_t T, uint32_t H, const uint32_t* __restrict__ lev,
                                 const int4* __restrict__ links, const float* __restrict__ top_child,
                                 float* __restrict__ root_out, float* box, int2* lnk,
                                 float4* __restrict__ wide, const uint2* __restrict__ wdst) {
  const uint32_t tid = threadIdx.x;
  // Entry boxes from the non-top children's boxes.  TOP_LOAD entries per
  // thread per round, every load of a round issued before any use (indices
  // clamped, records of top children loaded and then ignored).
  for (uint32_t i0 = 0; i0 < T; i0 += TOP_REFIT_THREADS * TOP_LOAD) {
    int4 L[TOP_LOAD];
    float4 tc[TOP_LOAD][3];
#pragma unroll
    for (int u = 0; u < TOP_LOAD; ++u) {
      const uint32_t i = min(i0 + u * TOP_REFIT_THREADS + tid, T - 1);
      L[u] = links[i];
      const float4* t4 = reinterpret_cast<const float4*>(top_child + 12 * (size_t)i);
      tc[u][0] = t4[0];
      tc[u][1] = t4[1];
      tc[u][2] = t4[2];
    }
#pragma unroll
    for (int u = 0; u < TOP_LOAD; ++u) {
      const uint32_t i = i0 + u * TOP_REFIT_THREADS + tid;
      if (i >= T) continue;
      const float a[6] = {tc[u][0].x, tc[u][0].y, tc[u][0].z, tc[u][0].w, tc[u][1].x, tc[u][1].y};
      const float b[6] = {tc[u][1].z, tc[u][1].w, tc[u][2].x, tc[u][2].y, tc[u][2].z, tc[u][2].w};
      float mn[3] = {100.0f, 100.0f, 100.0f}, mx[3] = {-100.0f, -100.0f, -100.0f};
      if (L[u].x < 0) unite(mn, mx, a);
      if (L[u].y < 0) unite(mn, mx, b);
      put6(box + 6 * (size_t)i, mn, mx);
      lnk[i] = make_int2(L[u].x, L[u].y);
    }
  }
  __syncthreads();
  for (uint32_t h = 1; h <= H; ++h) {
    const uint32_t e1 = lev[h + 1];
    for (uint32_t e = lev[h] + tid; e < e1; e += TOP_REFIT_THREADS) {
      const int2 c = lnk[e];
      float* be = box + 6 * (size_t)e;
      float mn[3] = {be[0], be[1], be[2]}, mx[3] = {be[3], be[4], be[5]};
      if (c.x >= 0) unite(mn, mx, box + 6 * (size_t)c.x);
      if (c.y >= 0) unite(mn, mx, box + 6 * (size_t)c.y);
      put6(be, mn, mx);
    }
    __syncthreads();
  }
  // stores: every slot, then every wide target, of the thread's entries
  // loaded before the first store (two memory round trips in all; the LDS
  // form holds at most TOP_PER_THREAD entries per thread)
  constexpr int R = LDS ? TOP_PER_THREAD : TOP_LOAD;
  for (uint32_t i0 = 0; i0 < T; i0 += TOP_REFIT_THREADS * R) {
    int32_t sl[R];
    uint2 wd[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const uint32_t i = min(i0 + u * TOP_REFIT_THREADS + tid, T - 1);
      sl[u] = links[i].w;
    }
#pragma unroll
    for (int u = 0; u < R; ++u)
      wd[u] = (wide && sl[u] >= 0) ? wdst[sl[u]] : make_uint2(~0u, 0u);
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const uint32_t i = i0 + u * TOP_REFIT_THREADS + tid;
      if (i >= T) continue;
      const int32_t s = sl[u];
      const float* b = box + 6 * (size_t)i;
      if (wide && s >= 0) {
        put_wide_at(wide, wd[u], b);
      } else {
        float* d = s < 0 ? root_out : (s & 1 ? nodes[s >> 1].b_min : nodes[s >> 1].a_min);
        for (int c = 0; c < 6; ++c) d[c] = b[c];
      }
    }
  }
}

__global__ __launch_bounds__(TOP_REFIT_THREADS) void k_refit_top(PF<Node> nodes_f, const int4* __restrict__ links,
                                                                 const uint32_t* __restrict__ lev,
                                                                 PF<const float> top_child_f,
                                                                 const uint32_t* __restrict__ top_count,
                                                                 const uint32_t* __restrict__ top_hmax,
                                                                 PF<float> root_f, PF<float> scratch_f,
                                                                 PF<float4> wide_f,
                                                                 const uint2* __restrict__ wdst) {
  Node* __restrict__ nodes = pick(nodes_f);
  const float* __restrict__ top_child = pick(top_child_f);
  float* root_out = pick(root_f);
  float* __restrict__ scratch = pick(scratch_f);
  float4* __restrict__ wide = pick(wide_f);
  __shared__ float sbox[TOP_MAX_LDS * 6];
  __shared__ int2 slnk[TOP_MAX_LDS];
  __shared__ uint32_t slev[TOP_MAX_LEVELS + 2];
  const uint32_t T = *top_count, H = *top_hmax;
  if (T == 0) return;
  for (uint32_t h = threadIdx.x; h <= H + 1; h += TOP_REFIT_THREADS) slev[h] = lev[h];
  __syncthreads();
  if (T <= (uint32_t)TOP_MAX_LDS) {
    refit_top_levels<true>(nodes, T, H, slev, links, top_child, root_out, sbox, slnk, wide, wdst);
  } else {
    // more entries than LDS holds: the same levels on global scratch (one
    // workgroup, so the barriers order it)
    float* b = scratch;
    int2* l = reinterpret_cast<int2*>(b + 6 * (size_t)T);
    refit_top_levels<false>(nodes, T, H, slev, links, top_child, root_out, b, l, wide, wdst);
  }
}

// ======================================================== host side
void Body::top_setup(int k, int l, hipStream_t s) {
  // top_count was zeroed by k_karras (no memset node in captured graphs)
  if (mode == CC_BODY_NONE || m < 2) return;
  Lane& L = lane[l];
  const unsigned nchunk = grid_for(m, RC_CHUNK);
  const Node* nd = slot[k].nodes.p;
  k_node_links<<<grid_for(m - 1, 256), 256, 0, s>>>(nd, m - 1, L.links.p);
  CC_LAUNCHED();
  k_top_list<<<nchunk, RC_THREADS, 0, s>>>(nd, m, L.top_list.p, L.top_pos.p, L.top_count.p);
  CC_LAUNCHED();
  k_chunk_heights<<<nchunk, RC_THREADS, 0, s>>>(nd, L.leaf_parent.p, m, L.top_pos.p, L.nflag.p, L.lflag.p,
                                                L.chunk_h.p);
  CC_LAUNCHED();
  if (m > (uint32_t)RC_CHUNK) {
    k_top_links<<<grid_for(m, 256), 256, 0, s>>>(nd, L.top_list.p, L.top_pos.p, L.top_count.p, L.top_links.p);
    CC_LAUNCHED();
    k_top_heights<<<1, TOP_THREADS, 0, s>>>(L.top_links.p, L.top_count.p,
                                            reinterpret_cast<uint32_t*>(L.top_scratch.p), L.top_h.p, L.top_hmax.p);
    CC_LAUNCHED();
    k_top_sort<<<1, TOP_THREADS, 0, s>>>(L.top_list.p, L.top_pos.p, L.top_links.p, L.top_h.p, L.top_count.p,
                                         L.top_hmax.p, L.top_lev.p, reinterpret_cast<int32_t*>(L.top_scratch.p));
    CC_LAUNCHED();
  }
}

void Body::refit(int k, int l, hipStream_t s, bool binary) {
  if (mode == CC_BODY_NONE || m == 0) return;
  if (bottom_up) {                      // same keys, so the same tree; new boxes
    prim_boxes_of(k, l, s);
    build_levels(1, &k, &l, s);
  } else {
    refit_frames(1, &k, &l, s, binary);
  }
  fold_frames(1, &k, s);                // a rebuilt body's wide records
}

void Body::refit_frames(int F, const int* ks, const int* ls, hipStream_t s, bool binary) {
  CC_REQUIRE(F >= 1 && F <= MAX_FRAMES && !bottom_up, "body: frames per refit chain");
  // every frame of a chain shares the topology (that of lane 0 for a
  // refitted body, of lane ls[0] for a rebuilt one refitted in place)
  Lane& T = topo(ls[0]);
  // A body walked through its four-wide form gets only the wide boxes (the
  // refit reads just the links of the binary records); `binary` asks for
  // the binary boxes instead (export, or the binary walk).  A rebuilt body
  // refitted in place gets binary boxes, which fold_frames then folds.
  const bool W = nwide && wide_enabled && !binary && !grouped;
  PF<const float> verts{}, top_child_c{};
  PF<Node> nodes{};
  PF<float> top_child{}, root{}, leaf0{}, scratch{};
  PF<float4> wide{};
  for (int f = 0; f < F; ++f) {
    Lane& L = lane[ls[f]];
    Slot& S = slot[ks[f]];
    verts.p[f] = S.verts.p;
    nodes.p[f] = S.nodes.p;
    top_child.p[f] = L.top_child.p;
    top_child_c.p[f] = L.top_child.p;
    root.p[f] = S.root.p;
    leaf0.p[f] = S.leaf0_box.p;
    scratch.p[f] = L.top_scratch.p;
    wide.p[f] = W ? S.wide.p : nullptr;
  }
  k_refit_chunks<<<dim3(grid_for(m, RC_CHUNK), (unsigned)F), RC_THREADS, 0, s>>>(
      verts, leaf_tri_of(ks[0]), T.leaf_parent.p, T.links.p, nodes, m, depth, T.nflag.p, T.lflag.p, T.chunk_h.p,
      T.top_pos.p, top_child, root, leaf0, wide, w_dst.p);
  CC_LAUNCHED();
  if (m > (uint32_t)RC_CHUNK) {
    k_refit_top<<<dim3(1, (unsigned)F), TOP_REFIT_THREADS, 0, s>>>(nodes, T.top_links.p, T.top_lev.p, top_child_c,
                                                                   T.top_count.p, T.top_hmax.p, root, scratch, wide,
                                                                   w_dst.p);
    CC_LAUNCHED();
  }
}

size_t Body::top_scratch_floats(uint32_t m) { return 9 * (size_t)(m > 1 ? m - 1 : 1) + 64; }

}  // namespace cc
