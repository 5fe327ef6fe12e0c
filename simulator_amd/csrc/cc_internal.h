// cc_internal.h -- shared types, error handling and exact-order FP32 math for
// libclothcore.  Every translation unit is compiled with -ffp-contract=off so
// that a*b + c is two roundings, as in the reference's glm code.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/clothcore.h"

namespace cc {

// ----------------------------------------------------------------- errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

// On failure the runtime's sticky last-error is cleared as well, so a later
// hipGetLastError() after a launch reports only that launch.
#define CC_HIP(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      (void)hipGetLastError();                                                        \
      throw ::cc::Error(CC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    }                                                                                 \
  } while (0)

// ------------------------------------------------------- fault injection
// cc_fault_inject: after `count` further successful operations of a kind,
// the next one fails as the runtime would have failed it, and the switch is
// off again.  Process-wide, one counter per kind (atomic: any thread).
enum FaultKind { FAULT_ALLOC = 0, FAULT_LAUNCH = 1, FAULT_KINDS = 2 };
inline std::atomic<int64_t> g_fault[FAULT_KINDS]{{-1}, {-1}};
inline bool fault_now(int kind) {
  int64_t c = g_fault[kind].load(std::memory_order_relaxed);
  while (c >= 0)
    if (g_fault[kind].compare_exchange_weak(c, c - 1)) return c == 0;
  return false;
}

// The check after a kernel launch (and where a fault is injected).
#define CC_LAUNCHED()                                                                       \
  do {                                                                                      \
    if (::cc::fault_now(::cc::FAULT_LAUNCH))                                                \
      throw ::cc::Error(CC_ERR_HIP, "kernel launch: injected failure (cc_fault_inject)");  \
    CC_HIP(hipGetLastError());                                                              \
  } while (0)

#define CC_REQUIRE(cond, msg)                                       \
  do {                                                              \
    if (!(cond)) throw ::cc::Error(CC_ERR_INVALID, std::string(msg)); \
  } while (0)

template <class F>
int guard(F&& f) {
  try {
    f();
    return CC_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return CC_ERR_ALLOC;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return CC_ERR_INVALID;
  }
}

// Device buffer owning a hipMalloc'd allocation.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    if (fault_now(FAULT_ALLOC)) throw Error(CC_ERR_ALLOC, "hipMalloc: injected failure (cc_fault_inject)");
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(CC_ERR_ALLOC, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    n = count;
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    alloc(count);
    if (count && h) CC_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void zero(hipStream_t s) {
    if (n) CC_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
};

inline unsigned grid_for(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

// Synchronous copy / fill on the object's own stream.  The library never
// uses the null stream: created on first use, it takes one of the process's
// hardware queues (four on MI355X by default) ahead of streams created later,
// and an update lane created after it would then share a queue with the
// step stream and be serialised behind it.
inline void copy_on(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
  CC_HIP(hipMemcpyAsync(dst, src, bytes, kind, s));
  CC_HIP(hipStreamSynchronize(s));
}
inline void fill_on(void* dst, int value, size_t bytes, hipStream_t s) {
  CC_HIP(hipMemsetAsync(dst, value, bytes, s));
  CC_HIP(hipStreamSynchronize(s));
}

// N events of a measurement, destroyed on every way out of it (the HIP calls
// between their creation and the end of the measurement throw on error; if
// a creation throws, the events made before it are destroyed too).
struct Ev {
  hipEvent_t h = nullptr;
  Ev() = default;
  Ev(const Ev&) = delete;
  Ev& operator=(const Ev&) = delete;
  ~Ev() {
    if (h) (void)hipEventDestroy(h);
  }
};
template <int N>
struct Events {
  Ev e[N];
  Events() {
    for (Ev& x : e) CC_HIP(hipEventCreate(&x.h));
  }
  hipEvent_t operator[](int i) const { return e[i].h; }
};

// ------------------------------------------------------------ physics pod
struct Params {
  float dt, mass, gx, gy, gz, damp, kd, k0, k1, k2, rest_scale, eps, box_depth;
  uint32_t ntypes;
};

// ------------------------------------------------------------- LBVH node
// One internal node of the Karras tree in 64 bytes.  The node stores the
// boxes of its two children (the traversal tests children, never the node
// itself, cudaBRT.cu:252-255), so a visit is one 64-byte line.
//   child >= 0: internal node index;  child < 0: leaf ~child
//   parent: internal index, -1 for the root
//   other: the end of this node's leaf range that is not the node index
struct alignas(16) Node {
  float a_min[3];
  float a_max[3];
  float b_min[3];
  float b_max[3];
  int32_t childA, childB, parent, other;
};
static_assert(sizeof(Node) == 64, "Node must be 64 bytes");

// Read-only view of a built LBVH handed to kernels by value.
struct BvhView {
  const Node* nodes;        // [m-1]
  const uint4* leaf_tri;    // [m] sorted {v0, v1, v2, primitive}
  const float* verts;       // body vertices, float3
  const float* leaf0_box;   // [6] box of leaf 0 when m == 1
  uint32_t m;
  // The same tree with every second level folded into its parent: 128-byte
  // nodes of up to four entries (four boxes, four links); nullptr when the
  // body is traversed through `nodes`.
  const float4* wide = nullptr;
  // the wide records are a rebuilt body's (one per internal node, split
  // records): the walk takes the split words into account
  uint32_t grouped = 0;
};

// --------------------------------------------------- FP32 math, glm order
struct f3 {
  float x, y, z;
};

__host__ __device__ inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ inline f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__host__ __device__ inline float length(f3 a) { return sqrtf(dot(a, a)); }
__host__ __device__ inline f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__host__ __device__ inline f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__host__ __device__ inline f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }
__host__ __device__ inline void st3(float* p, f3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}

// Expanded triangle box, Primitive::d_get_expand_bbox (primitive.cu:30-51):
// box of the three vertices united with the three vertices moved by
// -depth * unit normal.  out: min xyz, max xyz.
__host__ __device__ inline void prim_box(f3 a, f3 b, f3 c, float depth, float* mn, float* mx) {
  mn[0] = a.x; mn[1] = a.y; mn[2] = a.z;
  mx[0] = a.x; mx[1] = a.y; mx[2] = a.z;
  auto ex = [&](f3 p) {
    mn[0] = fminf(mn[0], p.x); mn[1] = fminf(mn[1], p.y); mn[2] = fminf(mn[2], p.z);
    mx[0] = fmaxf(mx[0], p.x); mx[1] = fmaxf(mx[1], p.y); mx[2] = fmaxf(mx[2], p.z);
  };
  ex(b);
  ex(c);
  f3 n = normalize(cross(b - a, c - a));
  f3 dn = n * depth;
  ex(a - dn);
  ex(b - dn);
  ex(c - dn);
}

// bvh.cu:20-27
__host__ __device__ inline uint32_t expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// BBox::centroid + getUnitcubePosOf + d_morton3D (bbox.h:97-127, bvh.cu:31-41)
__host__ __device__ inline uint32_t box_morton(const float* mn, const float* mx, const float* rmin,
                                               const float* rext) {
  f3 c = (mk(mn[0], mn[1], mn[2]) + mk(mx[0], mx[1], mx[2])) / 2.0f;
  f3 o = c - mk(rmin[0], rmin[1], rmin[2]);
  f3 e = mk(rext[0], rext[1], rext[2]);
  f3 u = mk(0.0f, 0.0f, 0.0f);
  if (length(e) != 0.0f) u = mk(o.x / e.x, o.y / e.y, o.z / e.z);
  float x = fminf(fmaxf(u.x * 1024.0f, 0.0f), 1023.0f);
  float y = fminf(fmaxf(u.y * 1024.0f, 0.0f), 1023.0f);
  float z = fminf(fmaxf(u.z * 1024.0f, 0.0f), 1023.0f);
  return expand_bits((uint32_t)x) * 4 + expand_bits((uint32_t)y) * 2 + expand_bits((uint32_t)z);
}

// BBox::intersect, inclusive (bbox.h:84-90)
__host__ __device__ inline bool box_has(const float* mn, const float* mx, f3 p) {
  return (p.x >= mn[0] && p.x <= mx[0]) && (p.y >= mn[1] && p.y <= mx[1]) &&
         (p.z >= mn[2] && p.z <= mx[2]);
}

}  // namespace cc
