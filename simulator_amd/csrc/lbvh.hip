// lbvh.hip -- body deformation and the Karras LBVH on gfx950: expanded
// triangle boxes, 30-bit Morton codes, LSD radix sort (LDS histogram +
// wavefront-ballot ranking) and the clz64 radix-tree build; the box refit
// lives in refit.hip.
//
// Reference path: BVHAccel (bvh.cu:154-297) computes boxes and codes on the
// GPU, sorts on the host (Utilities.h:22-67) and builds the tree with
// processInternalNode / calculateBoudingBox (cudaBRT.cu:120-206).  Here every
// stage stays on the device; keys are (morton << 32 | primitive) so all
// primitives are kept and the order is total.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cc_device.h"
#include "cc_lbvh.h"

namespace cc {

// ------------------------------------------------------------ deformation
MotionFrame motion_of(const MotionParams& a, const float* phase, uint64_t frame) {
  const double PI2 = 6.283185307179586476925;
  const double t = (double)frame * a.dt;
  MotionFrame m{};
  const double w = std::sin(PI2 * a.freq * t);
  const double th = a.rot_amp * w, c = std::cos(th), s = std::sin(th);
  const double R[9] = {c, 0.0, s, 0.0, 1.0, 0.0, -s, 0.0, c};
  const double P[3] = {a.px, a.py, a.pz};
  const double D[3] = {a.dx, a.dy, a.dz};
  for (int r = 0; r < 3; ++r) {
    const double rp = R[3 * r] * P[0] + R[3 * r + 1] * P[1] + R[3 * r + 2] * P[2];
    m.v[9 + r] = (float)(P[r] - rp + a.amp * w * D[r]);
  }
  for (int q = 0; q < 9; ++q) m.v[q] = (float)R[q];
  for (uint32_t k = 0; k < a.ncaps; ++k)
    m.v[12 + k] = (float)(1.0 + a.breath_amp * std::sin(PI2 * a.breath_freq * t + (double)phase[k]));
  return m;
}

// Every thread moves one vertex: radial breathing about the closest point of
// its capsule axis, then R * Y + T; block 0 also stores the frame's motion
// parameters for the host.
__global__ __launch_bounds__(256) void k_deform(const float* __restrict__ rest, PF<float> out_f,
                                                uint32_t nv, uint32_t ncaps, const uint32_t* __restrict__ vstart,
                                                const float* __restrict__ p0, const float* __restrict__ p1,
                                                MotionFrames mfs, PF<float> motion_f) {
  __shared__ float sP0[3 * MAX_CAPS], sP1[3 * MAX_CAPS];
  __shared__ uint32_t sV[MAX_CAPS + 1];
  const uint32_t tid = threadIdx.x;
  float* __restrict__ out = pick(out_f);
  float* __restrict__ motion_out = pick(motion_f);
  const float* sP = mfs.f[blockIdx.y].v;
  for (uint32_t k = tid; k < 3 * ncaps; k += blockDim.x) {
    sP0[k] = p0[k];
    sP1[k] = p1[k];
  }
  for (uint32_t k = tid; k <= ncaps; k += blockDim.x) sV[k] = vstart[k];
  if (blockIdx.x == 0)
    for (uint32_t k = tid; k < 12 + ncaps; k += blockDim.x) motion_out[k] = sP[k];
  __syncthreads();
  const uint32_t v = blockIdx.x * blockDim.x + tid;
  if (v >= nv) return;
  const uint32_t k = capsule_of(sV, ncaps, v);
  deform_point(ld3(rest + 3 * (size_t)v), ld3(sP0 + 3 * k), ld3(sP1 + 3 * k), sP[12 + k], sP, sP + 9,
               out + 3 * (size_t)v);
}

// ------------------------------------------------------------ boxes + codes
constexpr int BOX_THREADS = 256;

// expanded box per triangle + one partial root box per block
__global__ __launch_bounds__(BOX_THREADS) void k_prim_boxes(PF<const float> verts_f,
                                                            const uint32_t* __restrict__ tris, uint32_t m,
                                                            float depth, PF<float> boxes_f, PF<float> partial_f) {
  __shared__ float red[6][BOX_THREADS];
  const float* __restrict__ verts = pick(verts_f);
  float* __restrict__ boxes = pick(boxes_f);
  float* __restrict__ partial = pick(partial_f);
  const uint32_t i = blockIdx.x * BOX_THREADS + threadIdx.x;
  float mn[3], mx[3];
  box_sentinel(mn, mx);
  if (i < m) {
    const uint32_t* t = tris + 3 * (size_t)i;
    prim_box(ld3(verts + 3 * (size_t)t[0]), ld3(verts + 3 * (size_t)t[1]), ld3(verts + 3 * (size_t)t[2]),
             depth, mn, mx);
    float* b = boxes + 6 * (size_t)i;
    b[0] = mn[0]; b[1] = mn[1]; b[2] = mn[2];
    b[3] = mx[0]; b[4] = mx[1]; b[5] = mx[2];
  }
  for (int c = 0; c < 3; ++c) {
    red[c][threadIdx.x] = mn[c];
    red[3 + c][threadIdx.x] = mx[c];
  }
  __syncthreads();
  for (int s = BOX_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < (uint32_t)s)
      for (int c = 0; c < 3; ++c) {
        red[c][threadIdx.x] = fminf(red[c][threadIdx.x], red[c][threadIdx.x + s]);
        red[3 + c][threadIdx.x] = fmaxf(red[3 + c][threadIdx.x], red[3 + c][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[6 * (size_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// root box (sentinel U all partials, recomputed per block) then the code of
// each box centroid (compute_morton_bbox, bvh.cu:71-79)
__global__ __launch_bounds__(256) void k_morton(PF<const float> boxes_f, uint32_t m, PF<const float> partial_f,
                                                uint32_t npart, PF<uint32_t> codes_f, PF<uint32_t> idx_f,
                                                PF<float> root_f) {
  __shared__ float red[6][256];
  const float* __restrict__ boxes = pick(boxes_f);
  const float* __restrict__ partial = pick(partial_f);
  uint32_t* __restrict__ codes = pick(codes_f);
  uint32_t* __restrict__ idx = pick(idx_f);
  float* __restrict__ root_out = pick(root_f);
  float mn[3], mx[3];
  box_sentinel(mn, mx);
  for (uint32_t q = threadIdx.x; q < npart; q += 256)
    for (int c = 0; c < 3; ++c) {
      mn[c] = fminf(mn[c], partial[6 * (size_t)q + c]);
      mx[c] = fmaxf(mx[c], partial[6 * (size_t)q + 3 + c]);
    }
  for (int c = 0; c < 3; ++c) {
    red[c][threadIdx.x] = mn[c];
    red[3 + c][threadIdx.x] = mx[c];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < (uint32_t)s)
      for (int c = 0; c < 3; ++c) {
        red[c][threadIdx.x] = fminf(red[c][threadIdx.x], red[c][threadIdx.x + s]);
        red[3 + c][threadIdx.x] = fmaxf(red[3 + c][threadIdx.x], red[3 + c][threadIdx.x + s]);
      }
    __syncthreads();
  }
  const float rmin[3] = {red[0][0], red[1][0], red[2][0]};
  const float rext[3] = {red[3][0] - red[0][0], red[4][0] - red[1][0], red[5][0] - red[2][0]};
  if (blockIdx.x == 0 && threadIdx.x < 6) root_out[threadIdx.x] = red[threadIdx.x][0];
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const float* b = boxes + 6 * (size_t)i;
  codes[i] = box_morton(b, b + 3, rmin, rext);
  idx[i] = i;
}

// ------------------------------------------------------------ radix sort
// Stable LSD sort of (code, index) pairs by code, 8 bits per pass.  Because
// the indices enter in ascending order and every pass is stable, the result
// is ordered by the full 64-bit key (code << 32 | index).
constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 4;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;
// three passes of 10-bit digits cover the 30-bit Morton codes
constexpr int RADIX_BITS = 10;
constexpr int RADIX_DIGITS = 1 << RADIX_BITS;
constexpr int RADIX_PASSES = 3;
constexpr int DIGITS_PER_THREAD = RADIX_DIGITS / SORT_THREADS;

__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(PF<const uint32_t> keys_f, uint32_t n, int shift,
                                                             PF<uint32_t> hist_f, uint32_t nblk) {
  __shared__ uint32_t h[RADIX_DIGITS];
  const uint32_t* __restrict__ keys = pick(keys_f);
  uint32_t* __restrict__ hist = pick(hist_f);
  for (int q = 0; q < DIGITS_PER_THREAD; ++q) h[q * SORT_THREADS + threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * SORT_TILE;
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const uint32_t e = base + r * SORT_THREADS + threadIdx.x;
    if (e < n) atomicAdd(&h[(keys[e] >> shift) & (RADIX_DIGITS - 1)], 1u);
  }
  __syncthreads();
  for (int q = 0; q < DIGITS_PER_THREAD; ++q) {
    const uint32_t d = q * SORT_THREADS + threadIdx.x;
    hist[(size_t)d * nblk + blockIdx.x] = h[d];
  }
}

// Per digit, the exclusive prefix over blocks of the block counts (in place)
// and the digit's total: one wave per digit, 64 blocks per round.
__global__ __launch_bounds__(256) void k_radix_rowscan(PF<uint32_t> hist_f, uint32_t nblk, PF<uint32_t> total_f) {
  uint32_t* __restrict__ hist = pick(hist_f);
  uint32_t* __restrict__ total = pick(total_f);
  const uint32_t lane = threadIdx.x & 63u, d = blockIdx.x * 4 + (threadIdx.x >> 6);
  uint32_t* row = hist + (size_t)d * nblk;
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < nblk; b0 += 64) {
    const uint32_t b = b0 + lane;
    const uint32_t v = b < nblk ? row[b] : 0u;
    uint32_t incl = v;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t u = __shfl_up(incl, off, 64);
      if (lane >= (uint32_t)off) incl += u;
    }
    if (b < nblk) row[b] = carry + incl - v;
    carry += __shfl(incl, 63, 64);
  }
  if (lane == 0) total[d] = carry;
}

// Scatter with stable ranks: within a wave, lanes holding the same digit are
// matched with RADIX_BITS ballots; waves are ordered through per-digit LDS
// counts.  Thread t owns digits t, t + 256, t + 512, t + 768 in the scans.
__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(PF<const uint32_t> kin_f, PF<const uint32_t> vin_f,
                                                                uint32_t n, int shift, PF<const uint32_t> hist_f,
                                                                PF<const uint32_t> total_f, uint32_t nblk,
                                                                PF<uint32_t> kout_f, PF<uint32_t> vout_f) {
  __shared__ uint32_t goff[RADIX_DIGITS];
  const uint32_t* __restrict__ kin = pick(kin_f);
  const uint32_t* __restrict__ vin = pick(vin_f);
  const uint32_t* __restrict__ hist = pick(hist_f);
  const uint32_t* __restrict__ total = pick(total_f);
  uint32_t* __restrict__ kout = pick(kout_f);
  uint32_t* __restrict__ vout = pick(vout_f);
  __shared__ uint32_t run[RADIX_DIGITS];
  __shared__ uint32_t wsum[SORT_THREADS / 64];
  __shared__ uint32_t wcnt[SORT_THREADS / 64][RADIX_DIGITS];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  {  // digit bases: exclusive scan of the digit totals, in rounds of 256 digits
    uint32_t carry = 0;
    for (int q = 0; q < DIGITS_PER_THREAD; ++q) {
      const uint32_t d = q * SORT_THREADS + tid;
      const uint32_t v = total[d];
      uint32_t incl = v;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(incl, off, 64);
        if (lane >= (uint32_t)off) incl += u;
      }
      if (lane == 63) wsum[w] = incl;
      __syncthreads();
      uint32_t base = carry;
      for (uint32_t k = 0; k < w; ++k) base += wsum[k];
      goff[d] = base + incl - v + hist[(size_t)d * nblk + blockIdx.x];
      for (uint32_t k = 0; k < SORT_THREADS / 64; ++k) carry += wsum[k];
      __syncthreads();
    }
  }
  for (int q = 0; q < DIGITS_PER_THREAD; ++q) run[q * SORT_THREADS + tid] = 0;
  const uint32_t base = blockIdx.x * SORT_TILE;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (int r = 0; r < SORT_ITEMS; ++r) {
    for (int q = 0; q < SORT_THREADS / 64; ++q)
      for (int k = 0; k < DIGITS_PER_THREAD; ++k) wcnt[q][k * SORT_THREADS + tid] = 0;
    __syncthreads();
    const uint32_t e = base + r * SORT_THREADS + tid;
    const bool valid = e < n;
    const uint32_t key = valid ? kin[e] : 0u, val = valid ? vin[e] : 0u;
    const uint32_t d = (key >> shift) & (RADIX_DIGITS - 1);
    uint64_t msk = __ballot(valid);
    for (int b = 0; b < RADIX_BITS; ++b) {
      const uint32_t bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      msk &= bit ? bb : ~bb;
    }
    const uint32_t rank = __popcll(msk & lt);
    if (valid && rank == 0) wcnt[w][d] = __popcll(msk);
    __syncthreads();
    for (int k = 0; k < DIGITS_PER_THREAD; ++k) {
      const uint32_t dd = k * SORT_THREADS + tid;
      uint32_t s = run[dd];
      for (int q = 0; q < SORT_THREADS / 64; ++q) {
        const uint32_t c = wcnt[q][dd];
        wcnt[q][dd] = s;
        s += c;
      }
      run[dd] = s;
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = goff[d] + wcnt[w][d] + rank;
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_sort_finish(PF<const uint32_t> codes_f, PF<const uint32_t> idx_f,
                                                     const uint32_t* __restrict__ tris, uint32_t m,
                                                     PF<uint64_t> keys_f, PF<uint4> leaf_tri_f) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const uint32_t* __restrict__ codes = pick(codes_f);
  const uint32_t* __restrict__ idx = pick(idx_f);
  uint64_t* __restrict__ keys = pick(keys_f);
  uint4* __restrict__ leaf_tri = pick(leaf_tri_f);
  const uint32_t p = idx[i];
  keys[i] = ((uint64_t)codes[i] << 32) | p;
  const uint32_t* t = tris + 3 * (size_t)p;
  leaf_tri[i] = make_uint4(t[0], t[1], t[2], p);
}

// ------------------------------------------------------------ Karras build
__global__ __launch_bounds__(256) void k_karras(const uint64_t* __restrict__ keys, int32_t m,
                                                Node* __restrict__ nodes, int32_t* __restrict__ leaf_parent,
                                                uint32_t* __restrict__ top_count) {
  const int32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m - 1) return;
  karras_node(keys, m, i, nodes, leaf_parent);
  if (i == 0) *top_count = 0;      // the refit top-tree list is rebuilt after the topology
}

// ------------------------------------------------------------ host side
void Body::prim_boxes_of(int k, int l, hipStream_t s) {
  Lane& L = lane[l];
  k_prim_boxes<<<L.nblk_box, BOX_THREADS, 0, s>>>(pf1<const float>(slot[k].verts.p), tris.p, m, depth,
                                                  pf1(L.boxes.p), pf1(L.partial.p));
  CC_LAUNCHED();
}

void Body::build_levels(int F, const int* ks, const int* ls, hipStream_t s) {
  const LvBlock* tabs[BB_MAX_LEVELS];
  for (int q = 0; q < BB_MAX_LEVELS; ++q) tabs[q] = bu_level[q].p;
  LvFrame fr[MAX_FRAMES];
  for (int f = 0; f < F; ++f) {
    Lane& L = lane[ls[f]];
    Slot& S = slot[ks[f]];
    fr[f] = LvFrame{L.keys.p, leaf_tri_of(ks[f]), L.boxes.p, {L.items[0].p, L.items[1].p},
                    {L.nitems[0].p, L.nitems[1].p}, S.nodes.p, L.leaf_parent.p, S.root.p, S.leaf0_box.p};
  }
  launch_build_levels(bu_desc.p, tabs, bu_n, fr, F, s);
}

void Lane::alloc(uint32_t m, bool full_lane, bool bottom_up, hipStream_t s) {
  full = full_lane;
  const size_t mi = std::max<uint32_t>(m, 2) - 1;
  top_child.alloc(12 * mi);
  top_scratch.alloc(Body::top_scratch_floats(m));
  if (!full) return;
  nblk_box = grid_for(m, BOX_THREADS);
  boxes.alloc(6 * (size_t)m);
  partial.alloc(6 * (size_t)nblk_box);
  for (int b = 0; b < 2; ++b) {
    code[b].alloc(m);
    idx[b].alloc(m);
  }
  hist.alloc(RADIX_DIGITS * (size_t)grid_for(m, SORT_TILE));
  digit_total.alloc(RADIX_DIGITS);
  if (bottom_up) {
    const size_t c0 = level_chunks0(m);
    for (int q = 0; q < 2; ++q) {
      items[q].alloc(c0 * 128);
      nitems[q].alloc(c0);
      nitems[q].zero(s);
    }
  }
  keys.alloc(m);
  leaf_parent.alloc(m);
  leaf_parent.zero(s);
  top_list.alloc(mi);
  top_pos.alloc(mi);
  top_count.alloc(1);
  top_count.zero(s);
  top_links.alloc(mi);
  links.alloc(mi);
  links.zero(s);
  top_h.alloc(mi);
  top_hmax.alloc(1);
  top_hmax.zero(s);
  top_lev.alloc(66);
  top_lev.zero(s);
  nflag.alloc(mi);
  nflag.zero(s);
  lflag.alloc(std::max<uint32_t>(m, 1));
  lflag.zero(s);
  chunk_h.alloc(grid_for(std::max<uint32_t>(m, 1), REFIT_CHUNK));
  chunk_h.zero(s);
}

void Body::init(const float* v3, uint32_t nv_, const uint32_t* t3, uint32_t nf, uint32_t mode_, float depth_,
                const cc_capsules* caps, const MotionParams& mp, int slots, int lanes, hipStream_t s, int chain_,
                bool allow_bottom_up) {
  CC_REQUIRE(mode_ <= CC_BODY_REBUILD, "body: unknown mode");
  mode = mode_;
  nv = nv_;
  m = nf;
  depth = depth_;
  motion = mp;
  if (mode == CC_BODY_NONE) return;
  CC_REQUIRE(v3 && t3 && nv > 0, "body: vertices and triangles required");
  CC_REQUIRE(m < (1u << 30), "body: too many triangles");
  for (size_t i = 0; i < 3 * (size_t)m; ++i) CC_REQUIRE(t3[i] < nv, "body: triangle index out of range");
  nslots = moving() ? slots : 1;
  nlanes = moving() ? lanes : 1;
  CC_REQUIRE(nslots >= 1 && nslots <= MAX_SLOTS && nlanes >= 1 && nlanes <= MAX_LANES, "body: slot/lane count");
  rest.upload(v3, 3 * (size_t)nv, s);
  tris.upload(t3, 3 * (size_t)m, s);
  chain = moving() ? chain_ : 1;
  CC_REQUIRE(chain >= 1 && chain <= MAX_FRAMES, "body: frames per build chain must be 1..4");
  if (moving()) {
    CC_REQUIRE(caps && caps->ncaps >= 1 && caps->ncaps <= MAX_CAPS, "body: capsule layout required (1..64)");
    CC_REQUIRE(caps->vstart[0] == 0 && caps->vstart[caps->ncaps] == nv, "body: capsule ranges must cover vertices");
    for (uint32_t k = 0; k < caps->ncaps; ++k)
      CC_REQUIRE(caps->vstart[k] <= caps->vstart[k + 1], "body: capsule ranges must be ascending");
    ncaps = caps->ncaps;
    motion.ncaps = ncaps;
    cap_vstart.upload(caps->vstart, ncaps + 1, s);
    cap_p0.upload(caps->p0, 3 * (size_t)ncaps, s);
    cap_p1.upload(caps->p1, 3 * (size_t)ncaps, s);
    cap_phase.upload(caps->phase, ncaps, s);
    h_phase.assign(caps->phase, caps->phase + ncaps);
  }
  const size_t mi = std::max<uint32_t>(m, 2) - 1;
  if (mode != CC_BODY_REBUILD) leaf_tri.alloc(m);
  for (int k = 0; k < nslots; ++k) {
    Slot& S = slot[k];
    if (k == 0 || moving()) S.verts.upload(v3, 3 * (size_t)nv, s);
    if (mode == CC_BODY_REBUILD) S.leaf_tri.alloc(m);
    S.nodes.alloc(mi);
    S.nodes.zero(s);
    S.root.alloc(6);
    S.leaf0_box.alloc(6);
    S.leaf0_box.zero(s);
    S.motion.alloc(12 + (size_t)ncaps);
    S.motion.zero(s);
  }
  bottom_up = mode == CC_BODY_REBUILD && m >= 2 && m <= LV_MAX_LEAVES && allow_bottom_up;
  if (bottom_up) {
    BodyDesc d{};
    d.nv = nv;
    d.m = m;
    d.mode = mode;
    d.depth = depth;
    bu_desc.upload(&d, 1, s);
    const std::vector<std::vector<LvBlock>> t = build_level_tables(std::vector<BodyDesc>{d}, {0u});
    for (int q = 0; q < BB_MAX_LEVELS; ++q) {
      bu_n[q] = q < (int)t.size() ? (uint32_t)t[q].size() : 0u;
      if (bu_n[q]) bu_level[q].upload(t[q].data(), t[q].size(), s);
    }
  }
  if (mode == CC_BODY_REBUILD && !bottom_up) chain = 1;
  for (int l = 0; l < nlanes * chain; ++l) lane[l].alloc(m, l == 0 || mode == CC_BODY_REBUILD, bottom_up, s);
  // A rebuilt body is walked through wide records folded after each build,
  // one per internal node; its walk defers no more than the binary walk's
  // (the tree depth), so the bound reported is the stack limit.
  grouped = mode == CC_BODY_REBUILD && m >= 2;
  if (grouped) {
    nwide = m - 1;
    wide_stack = STACK_LIMIT;
    for (int k = 0; k < nslots; ++k) slot[k].wide.alloc(8 * (size_t)nwide);
  }
}

void Body::deform_frames(int F, const int* ks, const uint64_t* frames, hipStream_t s) {
  if (nv == 0 || ncaps == 0) return;
  MotionFrames mf;
  PF<float> out{}, mot{};
  for (int f = 0; f < F; ++f) {
    mf.f[f] = motion_of(motion, h_phase.data(), frames[f]);
    out.p[f] = slot[ks[f]].verts.p;
    mot.p[f] = slot[ks[f]].motion.p;
  }
  k_deform<<<dim3(grid_for(nv, 256), (unsigned)F), 256, 0, s>>>(rest.p, out, nv, ncaps, cap_vstart.p, cap_p0.p,
                                                               cap_p1.p, mf, mot);
  CC_LAUNCHED();
}

void Body::deform(int k, uint64_t frame, hipStream_t s) { deform_frames(1, &k, &frame, s); }

void Body::build_frames(int F, const int* ks, const int* ls, hipStream_t s) {
  if (mode == CC_BODY_NONE || m == 0) return;
  CC_REQUIRE(F >= 1 && F <= MAX_FRAMES && (F == 1 || bottom_up), "body: frames per build chain");
  PF<const float> verts{}, boxes_c{}, partial_c{};
  PF<float> boxes{}, partial{}, root{};
  PF<uint32_t> code[2] = {}, idx[2] = {}, hist{}, total{};
  PF<const uint32_t> code_c[2] = {}, idx_c[2] = {}, hist_c{}, total_c{};
  PF<uint64_t> keys{};
  PF<uint4> leaf{};
  for (int f = 0; f < F; ++f) {
    Lane& L = lane[ls[f]];
    CC_REQUIRE(L.full, "body: build on a refit-only lane");
    Slot& S = slot[ks[f]];
    verts.p[f] = S.verts.p;
    boxes.p[f] = L.boxes.p;
    boxes_c.p[f] = L.boxes.p;
    partial.p[f] = L.partial.p;
    partial_c.p[f] = L.partial.p;
    root.p[f] = S.root.p;
    for (int b = 0; b < 2; ++b) {
      code[b].p[f] = L.code[b].p;
      code_c[b].p[f] = L.code[b].p;
      idx[b].p[f] = L.idx[b].p;
      idx_c[b].p[f] = L.idx[b].p;
    }
    hist.p[f] = L.hist.p;
    hist_c.p[f] = L.hist.p;
    total.p[f] = L.digit_total.p;
    total_c.p[f] = L.digit_total.p;
    keys.p[f] = L.keys.p;
    leaf.p[f] = leaf_tri_of(ks[f]);
  }
  const unsigned nf = (unsigned)F;
  const uint32_t nblk_box = lane[ls[0]].nblk_box;
  k_prim_boxes<<<dim3(nblk_box, nf), BOX_THREADS, 0, s>>>(verts, tris.p, m, depth, boxes, partial);
  CC_LAUNCHED();
  k_morton<<<dim3(grid_for(m, 256), nf), 256, 0, s>>>(boxes_c, m, partial_c, nblk_box, code[0], idx[0], root);
  CC_LAUNCHED();
  const uint32_t nblk = grid_for(m, SORT_TILE);
  int src = 0;
  for (int pass = 0, shift = 0; pass < RADIX_PASSES; ++pass, shift += RADIX_BITS) {
    k_radix_hist<<<dim3(nblk, nf), SORT_THREADS, 0, s>>>(code_c[src], m, shift, hist, nblk);
    k_radix_rowscan<<<dim3(RADIX_DIGITS / 4, nf), 256, 0, s>>>(hist, nblk, total);
    k_radix_scatter<<<dim3(nblk, nf), SORT_THREADS, 0, s>>>(code_c[src], idx_c[src], m, shift, hist_c, total_c,
                                                            nblk, code[src ^ 1], idx[src ^ 1]);
    CC_LAUNCHED();
    src ^= 1;
  }
  k_sort_finish<<<dim3(grid_for(m, 256), nf), 256, 0, s>>>(code_c[src], idx_c[src], tris.p, m, keys, leaf);
  CC_LAUNCHED();
  if (bottom_up) {                      // tree, parent links and boxes, bottom-up
    build_levels(F, ks, ls, s);
    fold_frames(F, ks, s);
    return;
  }
  const int k = ks[0], l = ls[0];
  Lane& L = lane[l];
  if (m >= 2) {
    k_karras<<<grid_for(m - 1, 256), 256, 0, s>>>(L.keys.p, (int32_t)m, slot[k].nodes.p, L.leaf_parent.p,
                                                  L.top_count.p);
    CC_LAUNCHED();
  }
  top_setup(k, l, s);
  if (grouped) {         // binary boxes, then the wide records
    refit(k, l, s);
    return;
  }
  nwide = 0;             // a new topology: the wide table follows the refit
  refit(k, l, s);
  wide_setup(k, s);
}

void Body::build(int k, int l, hipStream_t s) { build_frames(1, &k, &l, s); }

void Body::share_topology(int from, hipStream_t s) {
  if (mode != CC_BODY_REFIT || m == 0) return;
  const Slot& F = slot[from];
  for (int k = 0; k < nslots; ++k) {
    if (k == from) continue;
    Slot& S = slot[k];
    CC_HIP(hipMemcpyAsync(S.nodes.p, F.nodes.p, F.nodes.n * sizeof(Node), hipMemcpyDeviceToDevice, s));
    CC_HIP(hipMemcpyAsync(S.root.p, F.root.p, 6 * sizeof(float), hipMemcpyDeviceToDevice, s));
    CC_HIP(hipMemcpyAsync(S.leaf0_box.p, F.leaf0_box.p, 6 * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (nwide) CC_HIP(hipMemcpyAsync(S.wide.p, F.wide.p, F.wide.n * sizeof(float4), hipMemcpyDeviceToDevice, s));
  }
}

void Body::update_frames(int F, const int* ks, const int* ls, const uint64_t* frames, hipStream_t s) {
  if (F == 1) {
    update(ks[0], ls[0], frames[0], s);
    return;
  }
  CC_REQUIRE(F <= chain, "body: frames per chain");
  deform_frames(F, ks, frames, s);
  if (mode == CC_BODY_REFIT)
    refit_frames(F, ks, ls, s);
  else
    build_frames(F, ks, ls, s);
}

void Body::update(int k, int l, uint64_t frame, hipStream_t s) {
  if (mode == CC_BODY_REFIT) {
    deform(k, frame, s);
    refit(k, l, s);
  } else if (mode == CC_BODY_REBUILD) {
    deform(k, frame, s);
    build(k, l, s);
  }
}

void export_tree_raw(uint32_t m, const uint64_t* keys_d, const uint4* leaf_tri_d, const int32_t* leaf_parent_d,
                     const Node* nodes_d, const float* root_d, const float* leaf0_d, uint64_t* okeys,
                     uint32_t* order, int32_t* childA, int32_t* childB, int32_t* parent_int, int32_t* parent_leaf,
                     float* box_int, float* box_leaf, float* root6, hipStream_t s) {
  std::vector<uint64_t> kk(m);
  std::vector<uint4> lt(m);
  std::vector<int32_t> lp(m);
  std::vector<Node> nd(m >= 2 ? m - 1 : 0);
  float r[6], l0[6];
  CC_HIP(hipMemcpyAsync(kk.data(), keys_d, m * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CC_HIP(hipMemcpyAsync(lt.data(), leaf_tri_d, m * sizeof(uint4), hipMemcpyDeviceToHost, s));
  CC_HIP(hipMemcpyAsync(lp.data(), leaf_parent_d, m * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  if (!nd.empty())
    CC_HIP(hipMemcpyAsync(nd.data(), nodes_d, nd.size() * sizeof(Node), hipMemcpyDeviceToHost, s));
  CC_HIP(hipMemcpyAsync(r, root_d, sizeof(r), hipMemcpyDeviceToHost, s));
  CC_HIP(hipMemcpyAsync(l0, leaf0_d, sizeof(l0), hipMemcpyDeviceToHost, s));
  CC_HIP(hipStreamSynchronize(s));
  for (uint32_t i = 0; i < m; ++i) {
    if (okeys) okeys[i] = kk[i];
    if (order) order[i] = lt[i].w;
  }
  if (root6) std::memcpy(root6, r, sizeof(r));
  if (m == 1) {
    if (box_leaf) std::memcpy(box_leaf, l0, sizeof(l0));
    if (parent_leaf) parent_leaf[0] = 0;
    return;
  }
  auto slot_of_child = [&](int32_t par, int32_t child, float* dst) {
    const Node& p = nd[par];
    const float* src = (p.childA == child) ? p.a_min : p.b_min;
    std::memcpy(dst, src, 6 * sizeof(float));
  };
  for (uint32_t x = 0; x + 1 < m; ++x) {
    const Node& n = nd[x];
    if (childA) childA[x] = n.childA >= 0 ? n.childA + 1 : n.childA;   // ~l == -(l+1)
    if (childB) childB[x] = n.childB >= 0 ? n.childB + 1 : n.childB;
    if (parent_int) parent_int[x] = n.parent + 1;
    if (box_int) {
      if (n.parent < 0) std::memcpy(box_int + 6 * (size_t)x, r, sizeof(r));
      else slot_of_child(n.parent, (int32_t)x, box_int + 6 * (size_t)x);
    }
  }
  for (uint32_t l = 0; l < m; ++l) {
    if (parent_leaf) parent_leaf[l] = lp[l] + 1;
    if (box_leaf) slot_of_child(lp[l], ~(int32_t)l, box_leaf + 6 * (size_t)l);
  }
}

void Body::export_tree(int k, uint64_t* okeys, uint32_t* order, int32_t* childA, int32_t* childB,
                       int32_t* parent_int, int32_t* parent_leaf, float* box_int, float* box_leaf, float* root6,
                       hipStream_t s) {
  CC_REQUIRE(mode != CC_BODY_NONE, "bvh export: no body");
  const Lane& L = lane[0];
  const Slot& S = slot[k];
  export_tree_raw(m, L.keys.p, leaf_tri_of(k), L.leaf_parent.p, S.nodes.p, S.root.p, S.leaf0_box.p, okeys, order,
                  childA, childB, parent_int, parent_leaf, box_int, box_leaf, root6, s);
}

}  // namespace cc
