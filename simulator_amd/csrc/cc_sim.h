// cc_sim.h -- the simulator object behind cc_sim: cloth buffers, the
// collider body, the step graphs and the update lanes.
#pragma once

#include <algorithm>
#include <vector>

#include "cc_body.h"

namespace cc {

constexpr int SPRING_BATCH = 8;
constexpr uint32_t PIN_BIT = 0x80000000u;
constexpr uint32_t ROW_MASK = 0x7fffffffu;
constexpr uint32_t END_MASK = 0x3fffffffu;
// Fixed-width spring rows: a cloth whose vertices have at most ELL_W springs
// (the grid cloths: 4 structural, 4 shear, 4 bend), or at most ELL_W2 (OBJ
// meshes: their 1-ring and 2-ring springs), also keeps its merged records as
// W per vertex (record k of vertex i at i * W + k), empty slots
// {i | PAD_TYPE, 0}, so a vertex's records load with its own data.
constexpr int ELL_W = 12;
constexpr int ELL_W2 = 24;
constexpr uint32_t PAD_TYPE = 3u << 30;

struct StepArgs {
  float inv_dt, dt2, inv_mass;
  float gmx, gmy, gmz;  // gravity * mass
  float damp, kd, k0, k1, k2, rest_scale, eps;
  uint32_t ntypes;
};

StepArgs step_args(const cc_params& prm, uint32_t ntypes);

// ------------------------------------------------------------------ Sim
struct Sim {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t n = 0, ntypes = 0;
  cc_params prm{};
  Body body;
  DBuf<float> state[2];
  DBuf<float> contact;
  DBuf<uint32_t> row;
  DBuf<int32_t> spill;      // [STACK_SPILL * n] traversal stack spill of the step kernel
  DBuf<uint2> rec;
  DBuf<uint2> ell;          // [ell_w * n] the same records, fixed width (when ell_w > 0)
  int ell_w = 0;            // 0 (merged rows only), ELL_W or ELL_W2
  std::vector<uint32_t> h_row;
  int cur = 0;              // state[cur] holds frame `frame`
  uint64_t frame = 0;
  uint64_t ready = 0;       // moving body: the poses of frames [frame, ready) are in their slots
  bool debug = false, counting = false;
  DBuf<float> dbg_pre;
  DBuf<int32_t> dbg_hit;
  DBuf<uint8_t> dbg_in;
  DBuf<unsigned long long> ctr;
  DBuf<uint32_t> flag;
  // VBO
  uint32_t nf = 0;
  DBuf<uint32_t> faces, vrow, vcol;
  DBuf<float> face_n, vnormals;

  // Graphs of graph_steps steps, per slot of the first frame and state
  // parity (captured when first needed).
  // Steps per block (even; a moving body's ring holds 2 * graph_steps
  // poses).  24 = 8 updates on each of three lanes: at the 501k-triangle
  // refitted body, 52.4 us per step against 54.5 with 16 (6/5/5 updates:
  // the step stream waits for the lane with six) and 53.8 with 32.
  int graph_steps = 24;
  hipGraphExec_t gmain[MAX_SLOTS][2] = {};
  // update lanes (moving bodies): streams and the events that order them
  // against the step stream
  hipStream_t lane_s[MAX_LANES] = {};
  hipEvent_t ev_go = nullptr, ev_done[MAX_LANES] = {};
  bool lanes_inflight = false;
  // Timing (cc_sim_set_timing): event set e brackets the step graph
  // (tev[e][0..1]) and each lane's updates (tev[e][2+2l..3+2l]) on the
  // stream they run on.  Two sets alternate, so the host reads one block's events while
  // the next block runs.
  bool timing = false;
  hipEvent_t tev[2][2 + 2 * MAX_LANES] = {};
  bool tpending[2] = {false, false};
  int tset = 0;
  double t_verlet_ms = 0, t_update_ms = 0;
  uint64_t t_steps = 0, t_updates = 0;
  // Single steps (the reference's one call per frame): the pose of frame g
  // is computed on a lane up to `lookahead` frames ahead of the step (a
  // body updated one frame per chain: lane g % nlanes; `chain` frames per
  // chain: in groups, see issue_chains), and ev_pose[slot of g] marks it
  // done; pose_pending says that the step of the frame in that slot must
  // wait for it.  block_inflight: the lanes of a step graph block have not
  // been joined yet.
  int lookahead = 0;
  hipEvent_t ev_pose[MAX_SLOTS] = {};
  bool pose_pending[MAX_SLOTS] = {};
  bool block_inflight = false;
  // "go" (ev_go) is recorded on the step stream before the step of frame
  // go_frame, so it completes once the steps before that frame are done.  A
  // lane about to overwrite the slot of frame g needs the step of frame
  // g - nslots done, i.e. a go recorded at go_frame > g - nslots: single
  // steps record a new go only every GO_EVERY frames and a lane waits on it
  // only when the one it last waited on (lane_go) no longer covers its next
  // frame -- on MI355X a hipStreamWaitEvent on a pending event costs the
  // host ~15 us and an event record ~1.4 us, on the per-frame path.  (Signal
  // words with hipStreamWriteValue64 / hipStreamWaitValue64 cost the host
  // less, but the pipelined steps ran slower with them: 38.5 against 37.2 us
  // per single step, 54 against 51.5 us in contact.)
  static constexpr int GO_EVERY = 8;
  uint64_t go_frame = 0;
  bool go_valid = false;
  uint64_t lane_go[MAX_LANES] = {};
  bool lane_go_valid[MAX_LANES] = {};
  // the last event recorded on each lane, and whether lane work is pending
  // that the step stream has not joined
  hipEvent_t lane_last[MAX_LANES] = {};
  bool lane_pending[MAX_LANES] = {};
  // Timing of single steps and of their updates: a ring of event pairs on
  // the streams they bracket, read once the ring is full or on collection.
  static constexpr int TRING = 256;
  hipEvent_t rev[TRING][2] = {};
  uint8_t rkind[TRING] = {};     // 0: fused step, 1: body update
  uint8_t rcount[TRING] = {};    // frames the bracketed launches cover
  uint32_t tcount[2][MAX_LANES] = {};   // frames of each lane in a timed block
  uint64_t chains_issued = 0;    // single-step chains so far (one in four is timed)
  uint64_t r_head = 0, r_tail = 0;

  // Work still queued on the streams (a creation that failed part way, lane
  // updates issued ahead of the last step) finishes before the buffers it
  // uses are freed.
  ~Sim() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (hipStream_t ls : lane_s)
      if (ls) (void)hipStreamSynchronize(ls);
    drop_graphs();
    for (auto& set : tev)
      for (hipEvent_t e : set)
        if (e) (void)hipEventDestroy(e);
    for (auto& pr : rev)
      for (hipEvent_t e : pr)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_pose)
      if (e) (void)hipEventDestroy(e);
    if (ev_go) (void)hipEventDestroy(ev_go);
    for (int l = 0; l < MAX_LANES; ++l) {
      if (ev_done[l]) (void)hipEventDestroy(ev_done[l]);
      if (lane_s[l]) (void)hipStreamDestroy(lane_s[l]);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
  void drop_graphs() {
    for (auto& p : gmain)
      for (hipGraphExec_t& g : p) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
  }

  StepArgs args() const { return step_args(prm, ntypes); }

  // the fused step reading body slot k and state[cur], writing state[cur ^ 1];
  // e0 / e1 (optional): events stamped with the kernel's start and end
  void launch_verlet(int k, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
  // the plain fused step between caller buffers (no debug outputs, no
  // counters), with the given arguments; contact normals from c_in to c_out
  // (the same array for an update in place)
  void launch_step(const float* in, float* out, const float* c_in, float* c_out, const StepArgs& a, int k,
                   hipStream_t s);

  // Direct (not captured) update of the body pose of frame g into its slot.
  void update_direct(uint64_t g, hipStream_t s) {
    body.update(body.slot_of(g), 0, g, s);
  }

  // Forget the lane work in flight (the device has been synchronised, so
  // every event has completed): the next step starts the pipeline afresh.
  void reset_pipeline() {
    for (bool& p : pose_pending) p = false;
    for (int l = 0; l < MAX_LANES; ++l) lane_pending[l] = lane_go_valid[l] = false;
    lanes_inflight = block_inflight = go_valid = false;
  }

  void record_go(hipStream_t s) {
    CC_HIP(hipEventRecord(ev_go, s));
    go_frame = frame;
    go_valid = true;
  }

  // Moving body: drop the poses computed ahead and recompute the current
  // frame's (after a change of frame, parameters or vertices).  Synchronous.
  void refresh_body() {
    reset_pipeline();
    if (body.moving()) {
      update_direct(frame, stream);
      ready = frame + 1;
    } else if (body.mode == CC_BODY_STATIC) {
      body.refit(0, 0, stream);
    }
    CC_HIP(hipStreamSynchronize(stream));
  }

  hipGraphExec_t end_capture(int cur0) {
    cur = cur0;
    hipGraph_t graph = nullptr;
    CC_HIP(hipStreamEndCapture(stream, &graph));
    hipGraphExec_t g = nullptr;
    hipError_t e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) throw Error(CC_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return g;
  }

  template <class F>
  hipGraphExec_t capture(F&& body_of_graph) {
    const int cur0 = cur;
    CC_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      body_of_graph();
    } catch (...) {
      cur = cur0;
      hipGraph_t junk = nullptr;
      (void)hipStreamEndCapture(stream, &junk);
      if (junk) (void)hipGraphDestroy(junk);
      throw;
    }
    return end_capture(cur0);
  }

  // graph_steps fused steps from a frame in slot k0, starting at the current
  // state parity.
  hipGraphExec_t capture_main(int k0) {
    return capture([&] {
      for (int j = 0; j < graph_steps; ++j) {
        launch_verlet(body.slot_of((uint64_t)(k0 + j)), stream);
        cur ^= 1;
      }
    });
  }

  // read the events of a finished timed block
  void collect(int set) {
    if (!tpending[set]) return;
    hipEvent_t* te = tev[set];
    const int G = graph_steps;
    CC_HIP(hipEventSynchronize(te[1]));
    float ms = 0;
    CC_HIP(hipEventElapsedTime(&ms, te[0], te[1]));
    t_verlet_ms += ms;
    t_steps += (uint64_t)G;
    if (body.moving())
      for (int l = 0; l < body.nlanes; ++l) {
        CC_HIP(hipEventSynchronize(te[3 + 2 * l]));
        CC_HIP(hipEventElapsedTime(&ms, te[2 + 2 * l], te[3 + 2 * l]));
        t_update_ms += ms;
        t_updates += tcount[set][l];
      }
    tpending[set] = false;
  }

  // the step stream waits for all lane work issued so far
  void join_lanes(hipStream_t s) {
    if (!lanes_inflight) return;
    for (int l = 0; l < body.nlanes; ++l)
      if (lane_pending[l]) {
        CC_HIP(hipStreamWaitEvent(s, lane_last[l], 0));
        lane_pending[l] = false;
      }
    lanes_inflight = block_inflight = false;
    for (bool& p : pose_pending) p = false;
  }

  // ---- timing ring of single steps
  void ring_collect_one() {
    const int e = (int)(r_tail % TRING);
    CC_HIP(hipEventSynchronize(rev[e][1]));
    float ms = 0;
    CC_HIP(hipEventElapsedTime(&ms, rev[e][0], rev[e][1]));
    if (rkind[e] == 0) {
      t_verlet_ms += ms;
      ++t_steps;
    } else {
      t_update_ms += ms;
      t_updates += rcount[e];
    }
    ++r_tail;
  }
  void ring_collect_all() {
    while (r_tail < r_head) ring_collect_one();
  }
  // an event pair of the ring (the older pair in its place is read first)
  hipEvent_t* ring_next(uint8_t kind, uint8_t count = 1) {
    if (r_head - r_tail == (uint64_t)TRING) ring_collect_one();
    const int e = (int)(r_head % TRING);
    rkind[e] = kind;
    rcount[e] = count;
    ++r_head;
    return rev[e];
  }

  // Lane work of the single-step path: the poses of frames [ready, upto]
  // on lanes g % nlanes, each after a go that covers its slot (see
  // GO_EVERY; the caller has recorded one at most GO_EVERY frames ago, and
  // upto + GO_EVERY < frame + nslots).  A body built `chain` frames per
  // chain of launches takes whole groups of nlanes * chain frames from
  // `ready` instead (lane l: frames ready + l + nlanes * q), so up to
  // nlanes * chain - 1 frames past upto.
  void issue_ahead(uint64_t upto) {
    if (body.chain > 1) {
      issue_chains(upto);
      return;
    }
    for (; ready <= upto; ++ready) {
      const uint64_t g = ready;
      const int l = (int)(g % (uint64_t)body.nlanes);
      const int k = body.slot_of(g);
      hipStream_t ls = lane_s[l];
      if (!lane_go_valid[l] || lane_go[l] + (uint64_t)body.nslots <= g) {
        CC_HIP(hipStreamWaitEvent(ls, ev_go, 0));
        lane_go[l] = go_frame;
        lane_go_valid[l] = true;
      }
      // updates are timed one frame in eight: event packets on the lanes
      // cost the step several microseconds when every update carries them
      hipEvent_t* te = (timing && g % 8 == 0) ? ring_next(1) : nullptr;
      if (te) CC_HIP(hipEventRecord(te[0], ls));
      body.update(k, l, g, ls);
      if (te) CC_HIP(hipEventRecord(te[1], ls));
      CC_HIP(hipEventRecord(ev_pose[k], ls));
      lane_last[l] = ev_pose[k];
      lane_pending[l] = true;
      pose_pending[k] = true;
      lanes_inflight = true;
    }
  }

  void issue_chains(uint64_t upto) {
    const int NL = body.nlanes, C = body.chain;
    while (ready <= upto) {
      const uint64_t base = ready;
      for (int l = 0; l < NL; ++l) {
        int ks[MAX_FRAMES], lq[MAX_FRAMES];
        uint64_t fr[MAX_FRAMES];
        for (int q = 0; q < C; ++q) {
          fr[q] = base + (uint64_t)(l + NL * q);
          ks[q] = body.slot_of(fr[q]);
          lq[q] = l + NL * q;
        }
        hipStream_t ls = lane_s[l];
        if (!lane_go_valid[l] || lane_go[l] + (uint64_t)body.nslots <= fr[C - 1]) {
          CC_HIP(hipStreamWaitEvent(ls, ev_go, 0));
          lane_go[l] = go_frame;
          lane_go_valid[l] = true;
        }
        // one chain in four timed (see issue_ahead)
        hipEvent_t* te = (timing && chains_issued % 4 == 0) ? ring_next(1, (uint8_t)C) : nullptr;
        ++chains_issued;
        if (te) CC_HIP(hipEventRecord(te[0], ls));
        body.update_frames(C, ks, lq, fr, ls);
        if (te) CC_HIP(hipEventRecord(te[1], ls));
        for (int q = 0; q < C; ++q) {
          CC_HIP(hipEventRecord(ev_pose[ks[q]], ls));
          pose_pending[ks[q]] = true;
        }
        lane_last[l] = ev_pose[ks[C - 1]];
        lane_pending[l] = true;
        lanes_inflight = true;
      }
      ready = base + (uint64_t)(NL * C);
    }
  }

  // One frame (the reference's simulate(), simulation.cu:141-150).  A
  // moving body's pose of this frame was issued on a lane by earlier calls
  // (computed here on the step stream only when nothing is ahead); the
  // poses of the next `lookahead` frames are issued on the lanes, where they
  // run beside this step and the following ones.
  void step_single(hipStream_t s) {
    const int k = body.slot_of(frame);
    const bool pipe = body.moving();
    if (pipe) {
      if (block_inflight) join_lanes(s);
      if (ready <= frame) {
        update_direct(frame, s);
        ready = frame + 1;
      }
      if (pose_pending[k]) {
        CC_HIP(hipStreamWaitEvent(s, ev_pose[k], 0));
        pose_pending[k] = false;
      }
      if (!go_valid || frame >= go_frame + GO_EVERY) record_go(s);
    }
    // timed one step in four: a launch that stamps events costs the host
    // ~10 us more than a plain one (12.8 against 2.9 us on MI355X)
    hipEvent_t* te = (timing && frame % 4 == 0) ? ring_next(0) : nullptr;
    if (te)
      launch_verlet(k, s, te[0], te[1]);
    else
      launch_verlet(k, s);
    if (pipe) issue_ahead(frame + (uint64_t)lookahead);
    cur ^= 1;
    ++frame;
  }

  // One block of graph_steps steps.  A moving body's poses for these frames
  // are in their slots (the previous block's lanes computed them, or they
  // are computed here first); the lanes compute the poses of the next
  // graph_steps frames, into the slots of the last graph_steps frames (the
  // ring holds 2 * graph_steps), while the steps run:
  //
  //   step stream:  join(lanes of last block) -> go -> [step graph: G fused steps]
  //   lane l:       wait(go) -> U(f+G+l), U(f+G+l+L), ... -> done[l]
  //
  // A dependency between streams costs several microseconds, so a block has
  // two, not two per step.  "go" follows the previous step graph, so a lane
  // overwrites a slot only after the steps that read it.
  void step_graph(hipStream_t s) {
    const int G = graph_steps;
    const int k0 = body.slot_of(frame);
    hipEvent_t* te = timing ? tev[tset] : nullptr;
    if (timing) collect(tset);
    const bool pipe = body.moving();
    if (pipe) {
      // poses of this block not yet issued: on the lanes, in parallel (their
      // slots were last read by steps before this one), then joined
      if (ready < frame + G) {
        record_go(s);
        issue_ahead(frame + G - 1);
      }
      join_lanes(s);
      record_go(s);
    }
    // The step graph is submitted first: submitting a graph to a stream that
    // waits on another stream can hold the host until that wait is over,
    // and the lanes' wait ("go") is over as soon as the step graph starts.
    hipGraphExec_t& g = gmain[k0][cur];
    if (!g) g = capture_main(k0);
    if (te) CC_HIP(hipEventRecord(te[0], s));
    CC_HIP(hipGraphLaunch(g, s));
    if (te) CC_HIP(hipEventRecord(te[1], s));
    if (pipe) {
      // the lanes are launched directly: a graph launch can hold the host
      // for most of the graph's run, which delays the next lane
      for (int l = 0; l < body.nlanes; ++l) {
        hipStream_t ls = lane_s[l];
        CC_HIP(hipStreamWaitEvent(ls, ev_go, 0));
        lane_go[l] = go_frame;
        lane_go_valid[l] = true;
        if (te) CC_HIP(hipEventRecord(te[2 + 2 * l], ls));
        // the lane's frames of the next block not issued yet (single steps
        // of a chained body may have issued its first ones), body.chain at
        // a time (a rebuilt body builds them in one chain of launches, frame
        // q with lane l + nlanes * q's scratch)
        const uint64_t first = std::max(ready, frame + (uint64_t)G), end = frame + 2 * (uint64_t)G;
        uint32_t nf = 0;
        for (uint64_t f0 = first + (uint64_t)l; f0 < end; f0 += (uint64_t)(body.nlanes * body.chain)) {
          int ks[MAX_FRAMES], lq[MAX_FRAMES];
          uint64_t fr[MAX_FRAMES];
          int F = 0;
          for (uint64_t f = f0; f < end && F < body.chain; f += (uint64_t)body.nlanes, ++F) {
            fr[F] = f;
            ks[F] = body.slot_of(f);
            lq[F] = l + body.nlanes * F;
          }
          body.update_frames(F, ks, lq, fr, ls);
          nf += (uint32_t)F;
        }
        if (te) tcount[tset][l] = nf;
        if (te) CC_HIP(hipEventRecord(te[3 + 2 * l], ls));
        CC_HIP(hipEventRecord(ev_done[l], ls));
        lane_last[l] = ev_done[l];
        lane_pending[l] = true;
      }
      lanes_inflight = block_inflight = true;
      ready = frame + 2 * G;
    }
    if (te) {
      tpending[tset] = true;
      tset ^= 1;
    }
    frame += (uint64_t)G;   // G is even: the state parity is unchanged
  }

  // Runs of graph_steps steps go through the graphs; the rest are launched
  // one by one, each followed by the body update of the next frame when it
  // is not ready yet.
  void step(int nsteps, hipStream_t s) {
    if (!s) s = stream;
    while (nsteps > 0) {
      if (nsteps >= graph_steps) {
        step_graph(s);
        nsteps -= graph_steps;
      } else {
        step_single(s);
        --nsteps;
      }
    }
  }
};

// The collision query + projection of n points (collision_response_projection
// without the Verlet part) over view v; spill: a spill area of n lanes whose
// tail the caller has set.  Asynchronous on s.
void launch_collide_points(const BvhView& v, const float* pts, uint32_t n, int32_t* spill, int32_t* hit,
                           uint8_t* ins, float* out_pos, float* out_n, hipStream_t s);

// All spring types of a cloth merged into one CSR row per vertex (types in
// order, each row in its CSR order); records {end | type << 30, rest bits}.
void merge_springs(const cc_cloth_desc* c, std::vector<uint32_t>& row, std::vector<uint2>& rec);
// n vec3 read from an array of stride floats per vertex
void read_vec3_from(const float* v, uint32_t stride, uint32_t n, std::vector<float>& out);

}  // namespace cc

struct cc_sim {
  cc::Sim s;
};
