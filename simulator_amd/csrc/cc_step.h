// cc_step.h -- the fused step of one cloth vertex (springs, Verlet, collision
// projection), shared by the single-cloth and batched step kernels.
#pragma once

#include "cc_device.h"
#include "cc_sim.h"

namespace cc {

// A neighbour's state {x_cur, x_prev} (24 bytes) as two 12-byte loads (two
// memory instructions per neighbour instead of three 8-byte ones).
struct NState {
  f3 p, q;
};
__device__ inline NState ld_state(const float* __restrict__ s, uint32_t j) {
  const float* sj = s + 6 * (size_t)j;
  return NState{ld3(sj), ld3(sj + 3)};
}

// One spring's term of compute_spring_force (verlet.cu:52-88) added to S:
// record rr {end | type << 30, rest}, neighbour state g.  An empty slot
// (real false) gathers the vertex itself: dP = (eps, eps, eps) > 0 and
// sc = +0, so the products are +0 and S (never -0) is unchanged.
__device__ inline void spring_term(f3 p, f3 vel, const NState& g, uint2 rr, bool real, const StepArgs& a,
                                   float ks0, float ks1, float ks2, f3& S) {
  const uint32_t t = rr.x >> 30;
  const float rest = __uint_as_float(rr.y) * a.rest_scale;
  const float ks = t == 0 ? ks0 : (t == 1 ? ks1 : ks2);
  const f3 p2 = g.p, q2 = g.q;
  const f3 dP = mk((p.x - p2.x) + a.eps, (p.y - p2.y) + a.eps, (p.z - p2.z) + a.eps);
  const f3 dV = mk(fmaf(-(p2.x - q2.x), a.inv_dt, vel.x), fmaf(-(p2.y - q2.y), a.inv_dt, vel.y),
                   fmaf(-(p2.z - q2.z), a.inv_dt, vel.z));
  const float L2 = fmaf(dP.z, dP.z, fmaf(dP.y, dP.y, dP.x * dP.x));
  const float rinv = rsqrtf(L2);
  const float L = L2 * rinv;
  const float vd = fmaf(dV.z, dP.z, fmaf(dV.y, dP.y, dV.x * dP.x));
  const float sc = real ? fmaf(-ks, L - rest, a.kd * vd * rinv) * rinv : 0.0f;
  S.x = fmaf(dP.x, sc, S.x);
  S.y = fmaf(dP.y, sc, S.y);
  S.z = fmaf(dP.z, sc, S.z);
}

// Vertex i from state `sin_` to `sout` (24 B {x_cur, x_prev} per vertex),
// its contact normal from contact_in to contact_out (the same array for an
// update in place; a thread reads and writes only its own vertex's).  Spring forces use FMA and rsqrt
// (positions are checked against the FP64 oracle by tolerance); the
// collision part is exact order so the contact set is bit-exact for the
// pre-projection positions.  verlet.cu:129-161, 52-88, 24-50.
// EW > 0: `rec` holds EW records per vertex (see ELL_W), loaded with the
// vertex's own records, and their gathers go out GB at a time; EW == 0:
// `rec` is the merged CSR and `row` its offsets, and the records load once
// the offsets have arrived.  Both sum the same springs in the same order.
// WIDE: the view may have the four-wide form (single bodies; the batched
// bodies never do).
// GB: gathers issued together on the fixed-width path (fewer registers for
// kernels that share a launch with other work).
template <bool DEBUG, int EW, bool WIDE, int GB = ELL_W>
__device__ inline void step_vertex(uint32_t i, const float* __restrict__ sin_, float* __restrict__ sout,
                                   const float* contact_in, float* contact_out, const uint32_t* __restrict__ row,
                                   const uint2* __restrict__ rec, const StepArgs& a, const BvhView& bvh,
                                   Spill spill, uint32_t& visits, float* __restrict__ dbg_pre,
                                   int32_t* __restrict__ dbg_hit, uint8_t* __restrict__ dbg_in) {
  // The vertex's own records are loaded together and waited for once: the
  // compiler otherwise issues each load at its first use, one memory round
  // trip after another.
  constexpr bool ELL = EW > 0;
  static_assert(EW % GB == 0, "fixed-width rows are whole gather batches");
  uint2 er[ELL ? EW : 1];
  if constexpr (ELL) {
#pragma unroll
    for (int k = 0; k < EW; ++k) er[k] = rec[(size_t)i * EW + k];
  }
  const WideRoot root = load_wide_root<WIDE>(bvh);
  const uint32_t r0raw = row[i], r1raw = ELL ? 0u : row[i + 1];
  const float2* s2 = reinterpret_cast<const float2*>(sin_ + 6 * (size_t)i);
  const float2 e0 = s2[0], e1 = s2[1], e2 = s2[2];
  const f3 c = ld3(contact_in + 3 * (size_t)i);
  asm volatile("" ::"v"(r0raw), "v"(r1raw), "v"(e0.x), "v"(e0.y), "v"(e1.x), "v"(e1.y), "v"(e2.x), "v"(e2.y),
               "v"(c.x), "v"(c.y), "v"(c.z));
  const uint32_t r0 = r0raw & ROW_MASK, r1 = r1raw & ROW_MASK;
  const f3 p = mk(e0.x, e0.y, e1.x), q = mk(e1.y, e2.x, e2.y);
  f3 np, nq, nc = mk(0.0f, 0.0f, 0.0f);
  int32_t hit = -1;
  bool inside = false;
  if (r0raw & PIN_BIT) {
    np = p;
    nq = p;
  } else {
    const f3 vel = mk((p.x - q.x) * a.inv_dt, (p.y - q.y) * a.inv_dt, (p.z - q.z) * a.inv_dt);
    f3 F = mk(fmaf(vel.x, a.damp, a.gmx), fmaf(vel.y, a.damp, a.gmy), fmaf(vel.z, a.damp, a.gmz));
    f3 S = mk(0.f, 0.f, 0.f);   // all spring types, in merged-row order
    // The stiffnesses as wave-uniform values (a is uniform in every caller):
    // a select between fields of `a` is otherwise compiled into a select of
    // their addresses and a load, which keeps a caller's local copy of the
    // arguments in scratch memory.
    const float ks0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(a.k0)));
    const float ks1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(a.k1)));
    const float ks2 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(a.k2)));
    if constexpr (ELL) {
      // GB neighbour gathers at a time, all issued before any use
#pragma unroll
      for (int b0 = 0; b0 < EW; b0 += GB) {
        NState g[GB];
#pragma unroll
        for (int k = 0; k < GB; ++k) g[k] = ld_state(sin_, er[b0 + k].x & END_MASK);
#pragma unroll
        for (int k = 0; k < GB; ++k)
          spring_term(p, vel, g[k], er[b0 + k], (er[b0 + k].x & PAD_TYPE) != PAD_TYPE, a, ks0, ks1, ks2, S);
      }
    } else {
      // Batches of SPRING_BATCH springs, branch-free: all record loads, then
      // all neighbour gathers, are issued before any use.  Slots past the
      // row end gather the vertex itself (with a condition per slot the
      // compiler sinks each gather into its branch, one memory round trip
      // per spring).
      for (uint32_t k0 = r0; k0 < r1; k0 += SPRING_BATCH) {
        uint2 rr[SPRING_BATCH];
        NState g[SPRING_BATCH];
#pragma unroll
        for (int q = 0; q < SPRING_BATCH; ++q) {
          rr[q] = rec[min(k0 + (uint32_t)q, r1 - 1)];
          if (k0 + q >= r1) rr[q] = make_uint2(i, 0u);
        }
#pragma unroll
        for (int q = 0; q < SPRING_BATCH; ++q) g[q] = ld_state(sin_, rr[q].x & END_MASK);
#pragma unroll
        for (int q = 0; q < SPRING_BATCH; ++q)
          spring_term(p, vel, g[q], rr[q], k0 + q < r1, a, ks0, ks1, ks2, S);
      }
    }
    F = F + S;
    const float cf = fmaf(c.z, F.z, fmaf(c.y, F.y, c.x * F.x));
    F = mk(fmaf(-c.x, cf, F.x), fmaf(-c.y, cf, F.y), fmaf(-c.z, cf, F.z));
    // pos + pos - pos_old + acc*dt*dt (verlet.cu:155) evaluated as
    // pos + ((pos - pos_old) + acc*dt^2): pos - pos_old is exact for close
    // values, so the only rounding at |pos| is the last addition.
    np = mk(p.x + fmaf(F.x * a.inv_mass, a.dt2, p.x - q.x), p.y + fmaf(F.y * a.inv_mass, a.dt2, p.y - q.y),
            p.z + fmaf(F.z * a.inv_mass, a.dt2, p.z - q.z));
    nq = p;
    if (DEBUG) st3(dbg_pre + 3 * (size_t)i, np);
    inside = collide_project<WIDE>(bvh, np, nc, hit, visits, spill, root);
    if (inside) nq = np;
  }
  if (DEBUG) {
    if (r0raw & PIN_BIT) st3(dbg_pre + 3 * (size_t)i, np);
    dbg_hit[i] = hit;
    dbg_in[i] = inside ? 1 : 0;
  }
  float* o = sout + 6 * (size_t)i;
  o[0] = np.x; o[1] = np.y; o[2] = np.z;
  o[3] = nq.x; o[4] = nq.y; o[5] = nq.z;
  // an update in place stores the normal only where it changed (bit for
  // bit): most vertices keep theirs (0 away from the body) from step to step
  if (contact_out != contact_in || __float_as_uint(nc.x) != __float_as_uint(c.x) ||
      __float_as_uint(nc.y) != __float_as_uint(c.y) || __float_as_uint(nc.z) != __float_as_uint(c.z))
    st3(contact_out + 3 * (size_t)i, nc);
}

}  // namespace cc
