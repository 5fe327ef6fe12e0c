// adjoint.hip -- reverse-mode gradient of a cloth rollout on gfx950.
//
// The forward is the plain fused step (k_verlet) with the parameters under
// study: stiffness of the three spring types, the velocity damping and the
// rest-length scale; the initial state is x_cur = x0, x_prev = x0 - v0 dt.
// The backward differentiates that step with the contacts of the forward
// held fixed: a vertex projected in a step is moved by the fixed plane
// projection (I - n n^T) u + (n . a) n, and the contact normal entering the
// next step's force projection F - (c . F) c is a constant.
//
// Per step (state x = (P, Q) -> (P', Q'), adjoints lP', lQ' -> lP, lQ):
//   u = 2P - Q + w (I - c c^T) F(P, Q),  w = dt^2 / m
//   free vertex:      P' = u, Q' = P          -> lu = lP',  lP += lQ'
//   projected vertex: P' = Q' = Pi(u)         -> lu = (I - n n^T)(lP' + lQ')
//   pinned vertex:    P' = Q' = P             -> lP += lP' + lQ'
//   lP += 2 lu, lQ -= lu, lF = w (I - c c^T) lu,
//   F = g m + damp v + sum_j f(d, dv), v = (P - Q)/dt, d = P_i - P_j + eps,
//   f = alpha d/L, alpha = -k (L - rest) + kd (dv . d)/L.
// Two kernels: k_adj_force (lu, lF and the vertex-local terms) and
// k_adj_springs (the spring terms, gathered per vertex: for a spring i->j
// the side of f_ij weighted by lF_i and, by the symmetry of the spring set
// (checked at creation), the side of f_ji weighted by lF_j).
//
// Checkpoints every `ckpt` steps; each segment is recomputed forward with
// its per-step states kept, then stepped back.  Parameter gradients: FP32
// block sums per step, accumulated per block in FP64, reduced in FP64 on
// the GPU at the end.
#include <cstring>
#include <memory>

#include "cc_device.h"
#include "cc_sim.h"

namespace cc {

constexpr int ADJ_THREADS = 256;
constexpr int NPARAM = 5;   // k0, k1, k2, damp, rest_scale

struct AdjArgs {
  float dt, inv_dt, w, damp, kd, k0, k1, k2, rest_scale, eps;
};

// Vector-Jacobian product of one spring force with the weight lam:
// gd = (df/dd)^T lam, gdv = (df/ddv)^T lam, dl = (d/L) . lam.
__device__ inline void spring_vjp(f3 d, f3 dv, float ks, float rest, float kd, f3 lam, f3& gd, f3& gdv,
                                  float& dl, float& L) {
  L = sqrtf(fmaf(d.z, d.z, fmaf(d.y, d.y, d.x * d.x)));
  const float il = 1.0f / L;
  const f3 u = mk(d.x * il, d.y * il, d.z * il);
  const float ud = fmaf(u.z, dv.z, fmaf(u.y, dv.y, u.x * dv.x));
  const float alpha = fmaf(-ks, L - rest, kd * ud);
  const float l = fmaf(u.z, lam.z, fmaf(u.y, lam.y, u.x * lam.x));
  const float kl = kd * il;
  // d alpha / d d = -ks u + kd (dv - (u . dv) u) / L;  d u / d d = (I - u u^T) / L
  const f3 da = mk(fmaf(-ks, u.x, kl * fmaf(-ud, u.x, dv.x)), fmaf(-ks, u.y, kl * fmaf(-ud, u.y, dv.y)),
                   fmaf(-ks, u.z, kl * fmaf(-ud, u.z, dv.z)));
  const float al = alpha * il;
  gd = mk(fmaf(da.x, l, al * fmaf(-u.x, l, lam.x)), fmaf(da.y, l, al * fmaf(-u.y, l, lam.y)),
          fmaf(da.z, l, al * fmaf(-u.z, l, lam.z)));
  gdv = mk(u.x * (kd * l), u.y * (kd * l), u.z * (kd * l));
  dl = l;
}

// Sum v over the block (64-wide waves, ADJ_THREADS / 64 of them); the
// result is valid in thread 0.
template <int K>
__device__ inline void block_sum(float (&v)[K]) {
  __shared__ float part[ADJ_THREADS / 64][K];
#pragma unroll
  for (int k = 0; k < K; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) part[threadIdx.x >> 6][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      float s = 0.0f;
      for (int w = 0; w < ADJ_THREADS / 64; ++w) s += part[w][k];
      v[k] = s;
    }
}

__global__ __launch_bounds__(ADJ_THREADS) void k_adj_force(
    uint32_t n, const uint32_t* __restrict__ row, const float* __restrict__ state,
    const float* __restrict__ c_in, const float* __restrict__ c_out, const float* __restrict__ lp,
    const float* __restrict__ lq, AdjArgs a, float* __restrict__ lf, float* __restrict__ locp,
    float* __restrict__ locq, double* __restrict__ acc) {
  const uint32_t i = blockIdx.x * ADJ_THREADS + threadIdx.x;
  float g[1] = {0.0f};
  if (i < n) {
    const f3 lP = ld3(lp + 3 * (size_t)i), lQ = ld3(lq + 3 * (size_t)i);
    f3 gp = mk(0.0f, 0.0f, 0.0f), gq = mk(0.0f, 0.0f, 0.0f), lF = mk(0.0f, 0.0f, 0.0f);
    if (row[i] & PIN_BIT) {
      gp = lP + lQ;
    } else {
      const f3 nn = ld3(c_out + 3 * (size_t)i);
      f3 lu;
      if (nn.x != 0.0f || nn.y != 0.0f || nn.z != 0.0f) {
        const f3 s2 = lP + lQ;
        const float ns = fmaf(nn.z, s2.z, fmaf(nn.y, s2.y, nn.x * s2.x));
        lu = mk(fmaf(-nn.x, ns, s2.x), fmaf(-nn.y, ns, s2.y), fmaf(-nn.z, ns, s2.z));
      } else {
        lu = lP;
        gp = lQ;
      }
      gp = mk(fmaf(2.0f, lu.x, gp.x), fmaf(2.0f, lu.y, gp.y), fmaf(2.0f, lu.z, gp.z));
      gq = mk(-lu.x, -lu.y, -lu.z);
      const f3 c = ld3(c_in + 3 * (size_t)i);
      const float cl = fmaf(c.z, lu.z, fmaf(c.y, lu.y, c.x * lu.x));
      lF = mk(a.w * fmaf(-c.x, cl, lu.x), a.w * fmaf(-c.y, cl, lu.y), a.w * fmaf(-c.z, cl, lu.z));
      const float* s = state + 6 * (size_t)i;
      const f3 vel = mk((s[0] - s[3]) * a.inv_dt, (s[1] - s[4]) * a.inv_dt, (s[2] - s[5]) * a.inv_dt);
      g[0] = fmaf(lF.z, vel.z, fmaf(lF.y, vel.y, lF.x * vel.x));   // F += damp * vel
      const float dd = a.damp * a.inv_dt;
      gp = mk(fmaf(dd, lF.x, gp.x), fmaf(dd, lF.y, gp.y), fmaf(dd, lF.z, gp.z));
      gq = mk(fmaf(-dd, lF.x, gq.x), fmaf(-dd, lF.y, gq.y), fmaf(-dd, lF.z, gq.z));
    }
    st3(lf + 3 * (size_t)i, lF);
    st3(locp + 3 * (size_t)i, gp);
    st3(locq + 3 * (size_t)i, gq);
  }
  block_sum(g);
  if (threadIdx.x == 0) acc[NPARAM * blockIdx.x + 3] += (double)g[0];
}

// One spring's adjoint terms (the forward's f_ij in F_i and f_ji in F_j)
// added to gp, gq and the parameter gradients g.
__device__ inline void adj_spring_term(f3 pi, f3 vi, f3 lFi, f3 pj, f3 vj, f3 lFj, uint2 r, const AdjArgs& a, f3& gp,
                                       f3& gq, float (&g)[4]) {
  const uint32_t t = r.x >> 30;
  const float rest0 = __uint_as_float(r.y), rest = rest0 * a.rest_scale;
  const float ks = t == 0 ? a.k0 : (t == 1 ? a.k1 : a.k2);
  f3 gd, gdv;
  float dl, L;
  // f_ij, in F_i
  spring_vjp(mk((pi.x - pj.x) + a.eps, (pi.y - pj.y) + a.eps, (pi.z - pj.z) + a.eps), vi - vj, ks, rest, a.kd, lFi,
             gd, gdv, dl, L);
  gp = mk(gp.x + fmaf(gdv.x, a.inv_dt, gd.x), gp.y + fmaf(gdv.y, a.inv_dt, gd.y), gp.z + fmaf(gdv.z, a.inv_dt, gd.z));
  gq = mk(fmaf(-gdv.x, a.inv_dt, gq.x), fmaf(-gdv.y, a.inv_dt, gq.y), fmaf(-gdv.z, a.inv_dt, gq.z));
  const float gk = -(L - rest) * dl;
  g[0] += t == 0 ? gk : 0.0f;
  g[1] += t == 1 ? gk : 0.0f;
  g[2] += t == 2 ? gk : 0.0f;
  g[3] += ks * rest0 * dl;
  // f_ji, in F_j (the same type and rest length)
  spring_vjp(mk((pj.x - pi.x) + a.eps, (pj.y - pi.y) + a.eps, (pj.z - pi.z) + a.eps), vj - vi, ks, rest, a.kd, lFj,
             gd, gdv, dl, L);
  gp = mk(gp.x - fmaf(gdv.x, a.inv_dt, gd.x), gp.y - fmaf(gdv.y, a.inv_dt, gd.y), gp.z - fmaf(gdv.z, a.inv_dt, gd.z));
  gq = mk(fmaf(gdv.x, a.inv_dt, gq.x), fmaf(gdv.y, a.inv_dt, gq.y), fmaf(gdv.z, a.inv_dt, gq.z));
}

// EW > 0: `rec` holds EW records per vertex (the forward's fixed-width rows,
// empty slots of type 3 skipped), loaded with the vertex's own data and
// gathered six at a time; EW == 0: the merged CSR at `row`, one spring after
// another.  The same springs in the same order either way.
template <int EW>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_springs(
    uint32_t n, const uint32_t* __restrict__ row, const uint2* __restrict__ rec, const float* __restrict__ state,
    const float* __restrict__ lf, const float* __restrict__ locp, const float* __restrict__ locq, AdjArgs a,
    float* __restrict__ lp_out, float* __restrict__ lq_out, double* __restrict__ acc) {
  const uint32_t i = blockIdx.x * ADJ_THREADS + threadIdx.x;
  float g[4] = {0.0f, 0.0f, 0.0f, 0.0f};   // k0, k1, k2, rest_scale
  if (i < n) {
    uint2 er[EW ? EW : 1];
    if constexpr (EW > 0) {
#pragma unroll
      for (int k = 0; k < EW; ++k) er[k] = rec[(size_t)i * EW + k];
    }
    const float* s = state + 6 * (size_t)i;
    const f3 pi = mk(s[0], s[1], s[2]);
    const f3 vi = mk((s[0] - s[3]) * a.inv_dt, (s[1] - s[4]) * a.inv_dt, (s[2] - s[5]) * a.inv_dt);
    const f3 lFi = ld3(lf + 3 * (size_t)i);
    f3 gp = ld3(locp + 3 * (size_t)i), gq = ld3(locq + 3 * (size_t)i);
    if constexpr (EW > 0) {
      constexpr int GB = 6;
#pragma unroll
      for (int b0 = 0; b0 < EW; b0 += GB) {
        float sj[GB][6], lj[GB][3];
#pragma unroll
        for (int k = 0; k < GB; ++k) {
          const uint32_t j = er[b0 + k].x & END_MASK;
          const float2* s2 = reinterpret_cast<const float2*>(state + 6 * (size_t)j);
          const float2 u = s2[0], v = s2[1], w = s2[2];
          sj[k][0] = u.x; sj[k][1] = u.y; sj[k][2] = v.x; sj[k][3] = v.y; sj[k][4] = w.x; sj[k][5] = w.y;
          lj[k][0] = lf[3 * (size_t)j]; lj[k][1] = lf[3 * (size_t)j + 1]; lj[k][2] = lf[3 * (size_t)j + 2];
        }
#pragma unroll
        for (int k = 0; k < GB; ++k) {
          if ((er[b0 + k].x & PAD_TYPE) == PAD_TYPE) continue;
          const f3 pj = mk(sj[k][0], sj[k][1], sj[k][2]);
          const f3 vj = mk((sj[k][0] - sj[k][3]) * a.inv_dt, (sj[k][1] - sj[k][4]) * a.inv_dt,
                           (sj[k][2] - sj[k][5]) * a.inv_dt);
          adj_spring_term(pi, vi, lFi, pj, vj, mk(lj[k][0], lj[k][1], lj[k][2]), er[b0 + k], a, gp, gq, g);
        }
      }
    } else {
      const uint32_t r0 = row[i] & ROW_MASK, r1 = row[i + 1] & ROW_MASK;
      for (uint32_t k = r0; k < r1; ++k) {
        const uint2 r = rec[k];
        const uint32_t j = r.x & END_MASK;
        const float* sj = state + 6 * (size_t)j;
        const f3 pj = mk(sj[0], sj[1], sj[2]);
        const f3 vj = mk((sj[0] - sj[3]) * a.inv_dt, (sj[1] - sj[4]) * a.inv_dt, (sj[2] - sj[5]) * a.inv_dt);
        adj_spring_term(pi, vi, lFi, pj, vj, ld3(lf + 3 * (size_t)j), r, a, gp, gq, g);
      }
    }
    st3(lp_out + 3 * (size_t)i, gp);
    st3(lq_out + 3 * (size_t)i, gq);
  }
  block_sum(g);
  if (threadIdx.x == 0) {
    double* ab = acc + NPARAM * (size_t)blockIdx.x;
    ab[0] += (double)g[0];
    ab[1] += (double)g[1];
    ab[2] += (double)g[2];
    ab[4] += (double)g[3];
  }
}

// out[p] = sum over blocks of acc[p], in FP64 (one workgroup)
__global__ __launch_bounds__(256) void k_reduce_f64(const double* __restrict__ acc, uint32_t nblk,
                                                    double* __restrict__ out) {
  __shared__ double part[256];
  for (int p = 0; p < NPARAM; ++p) {
    double s = 0.0;
    for (uint32_t b = threadIdx.x; b < nblk; b += 256) s += acc[NPARAM * (size_t)b + p];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[p] = part[0];
    __syncthreads();
  }
}

struct Adj {
  Sim* S = nullptr;
  uint32_t n = 0, nsteps = 0, ckpt = 0, nck = 0, nblk = 0;
  cc_params prm{};
  StepArgs sa{};
  AdjArgs aa{};
  bool have_forward = false;
  DBuf<float> st[2], contact;
  DBuf<float> ck_state, ck_contact;      // [nck][6n], [nck][3n]
  DBuf<float> seg_state, seg_contact;    // [ckpt+1][6n], [ckpt+1][3n]
  DBuf<float> lp[2], lq[2], lf, locp, locq;
  DBuf<double> acc, out;
  ~Adj() {
    if (S && S->stream) (void)hipStreamSynchronize(S->stream);   // before the buffers are freed
  }
};

static void check_symmetric(const Sim& S) {
  std::vector<uint2> rec(S.rec.n);
  copy_on(rec.data(), S.rec.p, rec.size() * sizeof(uint2), hipMemcpyDeviceToHost, S.stream);
  const std::vector<uint32_t>& row = S.h_row;
  for (uint32_t i = 0; i < S.n; ++i)
    for (uint32_t k = row[i] & ROW_MASK; k < (row[i + 1] & ROW_MASK); ++k) {
      const uint32_t j = rec[k].x & END_MASK, t = rec[k].x >> 30;
      bool found = false;
      for (uint32_t e = row[j] & ROW_MASK; e < (row[j + 1] & ROW_MASK) && !found; ++e)
        found = (rec[e].x & END_MASK) == i && (rec[e].x >> 30) == t && rec[e].y == rec[k].y;
      CC_REQUIRE(found, "cc_adj_create: the spring set is not symmetric (every spring i->j needs j->i "
                        "of the same type and rest length)");
    }
}

}  // namespace cc

using namespace cc;

struct cc_adj {
  Adj a;
};

extern "C" {

int cc_adj_create(cc_sim* sim, uint32_t nsteps, uint32_t ckpt_every, cc_adj** out) {
  return guard([&] {
    CC_REQUIRE(sim && out && nsteps >= 1 && ckpt_every >= 1, "cc_adj_create: bad argument");
    Sim& S = sim->s;
    CC_REQUIRE(!S.body.moving(), "cc_adj_create: a moving body is not supported (static or no body)");
    *out = nullptr;
    CC_HIP(hipDeviceSynchronize());
    check_symmetric(S);
    std::unique_ptr<cc_adj> h(new cc_adj);
    Adj& A = h->a;
    A.S = &S;
    A.n = S.n;
    A.nsteps = nsteps;
    A.ckpt = std::min(ckpt_every, nsteps);
    A.nck = (nsteps + A.ckpt - 1) / A.ckpt;
    A.nblk = grid_for(S.n, ADJ_THREADS);
    const size_t n3 = 3 * (size_t)S.n, n6 = 6 * (size_t)S.n;
    for (int b = 0; b < 2; ++b) {
      A.st[b].alloc(n6);
      A.lp[b].alloc(n3);
      A.lq[b].alloc(n3);
    }
    A.contact.alloc(n3);
    A.ck_state.alloc(A.nck * n6);
    A.ck_contact.alloc(A.nck * n3);
    A.seg_state.alloc((A.ckpt + 1) * n6);
    A.seg_contact.alloc((A.ckpt + 1) * n3);
    A.lf.alloc(n3);
    A.locp.alloc(n3);
    A.locq.alloc(n3);
    A.acc.alloc(NPARAM * (size_t)A.nblk);
    A.out.alloc(NPARAM);
    *out = h.release();
  });
}

int cc_adj_destroy(cc_adj* adj) {
  return guard([&] {
    if (!adj) return;
    (void)hipStreamSynchronize(adj->a.S->stream);
    delete adj;
  });
}

int cc_adj_forward(cc_adj* adj, const float* x0, const float* v0, const double theta[5], float* x_final) {
  return guard([&] {
    CC_REQUIRE(adj && x0 && theta, "cc_adj_forward: bad argument");
    Adj& A = adj->a;
    Sim& S = *A.S;
    hipStream_t s = S.stream;
    CC_HIP(hipDeviceSynchronize());
    A.prm = S.prm;
    for (int t = 0; t < 3; ++t) A.prm.k[t] = (float)theta[t];
    A.prm.damp = (float)theta[3];
    A.prm.rest_scale = (float)theta[4];
    A.sa = step_args(A.prm, S.ntypes);
    A.aa.dt = A.prm.dt;
    A.aa.inv_dt = A.sa.inv_dt;
    A.aa.w = A.sa.dt2 * A.sa.inv_mass;
    A.aa.damp = A.prm.damp;
    A.aa.kd = A.prm.kd;
    A.aa.k0 = A.prm.k[0];
    A.aa.k1 = A.prm.k[1];
    A.aa.k2 = A.prm.k[2];
    A.aa.rest_scale = A.prm.rest_scale;
    A.aa.eps = A.prm.eps;
    const size_t n3 = 3 * (size_t)A.n, n6 = 6 * (size_t)A.n;
    std::vector<float> packed(n6);
    for (uint32_t i = 0; i < A.n; ++i)
      for (int c = 0; c < 3; ++c) {
        const float x = x0[3 * (size_t)i + c], v = v0 ? v0[3 * (size_t)i + c] : 0.0f;
        packed[6 * (size_t)i + c] = x;
        packed[6 * (size_t)i + 3 + c] = x - v * A.prm.dt;
      }
    CC_HIP(hipMemcpyAsync(A.st[0].p, packed.data(), n6 * sizeof(float), hipMemcpyHostToDevice, s));
    A.contact.zero(s);
    int cur = 0;
    for (uint32_t t = 0; t < A.nsteps; ++t) {
      if (t % A.ckpt == 0) {
        const uint32_t c = t / A.ckpt;
        CC_HIP(hipMemcpyAsync(A.ck_state.p + c * n6, A.st[cur].p, n6 * sizeof(float), hipMemcpyDeviceToDevice, s));
        CC_HIP(hipMemcpyAsync(A.ck_contact.p + c * n3, A.contact.p, n3 * sizeof(float), hipMemcpyDeviceToDevice, s));
      }
      S.launch_step(A.st[cur].p, A.st[cur ^ 1].p, A.contact.p, A.contact.p, A.sa, 0, s);
      cur ^= 1;
    }
    if (x_final) {
      CC_HIP(hipMemcpyAsync(packed.data(), A.st[cur].p, n6 * sizeof(float), hipMemcpyDeviceToHost, s));
      CC_HIP(hipStreamSynchronize(s));
      for (uint32_t i = 0; i < A.n; ++i)
        for (int c = 0; c < 3; ++c) x_final[3 * (size_t)i + c] = packed[6 * (size_t)i + c];
    }
    CC_HIP(hipStreamSynchronize(s));
    A.have_forward = true;
  });
}

int cc_adj_backward(cc_adj* adj, const float* dl_dx_final, double dl_dtheta[5], float* dl_dv0, float* dl_dx0) {
  return guard([&] {
    CC_REQUIRE(adj && dl_dx_final, "cc_adj_backward: bad argument");
    Adj& A = adj->a;
    CC_REQUIRE(A.have_forward, "cc_adj_backward: call cc_adj_forward first");
    Sim& S = *A.S;
    hipStream_t s = S.stream;
    const size_t n3 = 3 * (size_t)A.n, n6 = 6 * (size_t)A.n;
    CC_HIP(hipMemcpyAsync(A.lp[0].p, dl_dx_final, n3 * sizeof(float), hipMemcpyHostToDevice, s));
    A.lq[0].zero(s);
    A.acc.zero(s);
    int cur = 0;
    for (uint32_t c = A.nck; c-- > 0;) {
      const uint32_t t0 = c * A.ckpt, t1 = std::min(t0 + A.ckpt, A.nsteps), L = t1 - t0;
      // the segment forward again, keeping every state
      CC_HIP(hipMemcpyAsync(A.seg_state.p, A.ck_state.p + c * n6, n6 * sizeof(float), hipMemcpyDeviceToDevice, s));
      CC_HIP(hipMemcpyAsync(A.seg_contact.p, A.ck_contact.p + c * n3, n3 * sizeof(float), hipMemcpyDeviceToDevice,
                            s));
      for (uint32_t j = 0; j < L; ++j)   // contact normals of step j in, of step j + 1 out
        S.launch_step(A.seg_state.p + j * n6, A.seg_state.p + (j + 1) * n6, A.seg_contact.p + j * n3,
                      A.seg_contact.p + (j + 1) * n3, A.sa, 0, s);
      for (uint32_t j = L; j-- > 0;) {
        const float* stj = A.seg_state.p + j * n6;
        k_adj_force<<<A.nblk, ADJ_THREADS, 0, s>>>(A.n, S.row.p, stj, A.seg_contact.p + j * n3,
                                                   A.seg_contact.p + (j + 1) * n3, A.lp[cur].p, A.lq[cur].p, A.aa,
                                                   A.lf.p, A.locp.p, A.locq.p, A.acc.p);
        CC_LAUNCHED();
        if (S.ell_w == ELL_W)
          k_adj_springs<ELL_W><<<A.nblk, ADJ_THREADS, 0, s>>>(A.n, S.row.p, S.ell.p, stj, A.lf.p, A.locp.p,
                                                              A.locq.p, A.aa, A.lp[cur ^ 1].p, A.lq[cur ^ 1].p,
                                                              A.acc.p);
        else
          k_adj_springs<0><<<A.nblk, ADJ_THREADS, 0, s>>>(A.n, S.row.p, S.rec.p, stj, A.lf.p, A.locp.p, A.locq.p,
                                                          A.aa, A.lp[cur ^ 1].p, A.lq[cur ^ 1].p, A.acc.p);
        CC_LAUNCHED();
        cur ^= 1;
      }
    }
    k_reduce_f64<<<1, 256, 0, s>>>(A.acc.p, A.nblk, A.out.p);
    CC_LAUNCHED();
    std::vector<float> lp(n3), lq(n3);
    double g[NPARAM];
    CC_HIP(hipMemcpyAsync(lp.data(), A.lp[cur].p, n3 * sizeof(float), hipMemcpyDeviceToHost, s));
    CC_HIP(hipMemcpyAsync(lq.data(), A.lq[cur].p, n3 * sizeof(float), hipMemcpyDeviceToHost, s));
    CC_HIP(hipMemcpyAsync(g, A.out.p, sizeof(g), hipMemcpyDeviceToHost, s));
    CC_HIP(hipStreamSynchronize(s));
    if (dl_dtheta) std::memcpy(dl_dtheta, g, sizeof(g));
    for (size_t k = 0; k < n3; ++k) {   // x_cur = x0, x_prev = x0 - v0 dt
      if (dl_dx0) dl_dx0[k] = lp[k] + lq[k];
      if (dl_dv0) dl_dv0[k] = -A.prm.dt * lq[k];
    }
  });
}

}  // extern "C"
