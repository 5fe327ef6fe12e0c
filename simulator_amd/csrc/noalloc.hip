// noalloc.hip -- the collision query without device allocation.
//
// cc_collide_points allocates its inputs, outputs and the traversal spill
// area on every call.  cc_collide_points_noalloc takes device inputs and
// outputs from the caller and keeps its spill area in `scratch`, a
// file-scope device array of fixed size (each GPU the library runs on holds
// one copy until the process ends), so that callers which cannot allocate
// -- code under stream capture, serving loops -- can query the body.  Calls
// that overlap on one GPU share that array and race; the array keeps the
// last call's contents, and each call sets the words that must start at a
// known value (the spill area's fault word and stack limit) with a kernel.
#include <mutex>

#include "cc_step.h"

namespace cc {

// at most this many points per call: the spill area holds STACK_SPILL
// entries per point (15.7 MB)
constexpr uint32_t NOALLOC_MAX_POINTS = 65536;
constexpr size_t NOALLOC_WORDS = (size_t)STACK_SPILL * NOALLOC_MAX_POINTS + SPILL_TAIL;

static __device__ __align__(256) int32_t scratch[NOALLOC_WORDS];

// the address of each device's copy of `scratch` (host code never passes
// the array's name itself: on the host it stands for no device address)
constexpr int NOALLOC_MAX_DEVICES = 64;
static int32_t* scratch_dev[NOALLOC_MAX_DEVICES] = {};
static std::mutex scratch_mu;

// With `dev` current: its copy of the array, looked up on the first call.
static int32_t* scratch_of(int dev) {
  CC_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "cc_collide_points_noalloc: device index out of range");
  std::lock_guard<std::mutex> g(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    CC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    CC_REQUIRE(p != nullptr, "cc_collide_points_noalloc: no address for the scratch array");
    scratch_dev[dev] = static_cast<int32_t*>(p);
  }
  return scratch_dev[dev];
}

// the spill tail of this call: no faults, the default stack limit
__global__ void k_noalloc_tail(int32_t* __restrict__ tail) {
  if (threadIdx.x == 0) {
    tail[0] = 0;
    tail[1] = STACK_LIMIT;
  }
}

// the call's traversal faults to the caller's status word
__global__ void k_noalloc_status(const int32_t* __restrict__ tail, uint32_t* __restrict__ status) {
  if (threadIdx.x == 0) *status = (uint32_t)tail[0];
}

}  // namespace cc

using namespace cc;

extern "C" int cc_collide_points_noalloc(cc_sim* sim, const float* pts3, uint32_t n, int32_t* hit,
                                         uint8_t* inside, float* out_pos, float* out_n, uint32_t* status,
                                         void* stream) {
  return guard([&] {
    CC_REQUIRE(sim, "cc_collide_points_noalloc: null simulator");
    CC_REQUIRE(n <= NOALLOC_MAX_POINTS,
               "cc_collide_points_noalloc: at most 65536 points per call (the fixed scratch size)");
    if (n == 0) return;
    CC_REQUIRE(pts3 && hit && inside && out_pos && out_n, "cc_collide_points_noalloc: null buffer");
    Sim& S = sim->s;
    hipStream_t s = stream ? (hipStream_t)stream : S.stream;
    int prev = 0;
    CC_HIP(hipGetDevice(&prev));
    if (prev != S.device) CC_HIP(hipSetDevice(S.device));
    try {
      int32_t* spill = scratch_of(S.device);
      const int k = S.body.slot_of(S.frame);
      // a pose of the current frame still being computed on an update lane
      if (S.pose_pending[k]) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        CC_HIP(hipStreamIsCapturing(s, &cs));
        CC_REQUIRE(cs == hipStreamCaptureStatusNone,
                   "cc_collide_points_noalloc: the current frame's body pose is still being computed; "
                   "synchronize the simulator before capturing the query");
        CC_HIP(hipStreamWaitEvent(s, S.ev_pose[k], 0));
      }
      int32_t* tail = spill + (size_t)STACK_SPILL * n;
      k_noalloc_tail<<<1, 64, 0, s>>>(tail);
      CC_LAUNCHED();
      launch_collide_points(S.body.view(k), pts3, n, spill, hit, inside, out_pos, out_n, s);
      if (status) {
        k_noalloc_status<<<1, 64, 0, s>>>(tail, status);
        CC_LAUNCHED();
      }
    } catch (...) {
      if (prev != S.device) (void)hipSetDevice(prev);
      throw;
    }
    if (prev != S.device) CC_HIP(hipSetDevice(prev));
  });
}
