/*
 * clothcore.h -- C ABI of libclothcore, an MI355X (gfx950) cloth simulation
 * core: CSR mass-spring Verlet step, Karras LBVH over a triangle body and
 * point-vs-LBVH collision projection.
 *
 * The reference (sutongkui/simulator) has no library or FFI: its path sits
 * behind the C++ classes Simulator (src/simulator.h:9-59), Springs
 * (src/spring.h:14-36) and BVHAccel (src/bvh/bvh.h:20-101).  Each entry point
 * below names the reference member it replaces.  Plain pointers and sizes
 * only.  Every call returns CC_OK (0) or an error code; cc_last_error()
 * returns a thread-local message (the reference prints and continues,
 * src/common.h:4-10).
 *
 * Memory: host pointers unless a name says "_dev".  Arrays returned through
 * cc_csr are allocated by the library and released with cc_csr_free().
 * A cc_sim is single-threaded; separate handles may run on separate streams.
 */
#ifndef CLOTHCORE_H
#define CLOTHCORE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CC_OK 0
#define CC_ERR_INVALID 1     /* bad argument or size                       */
#define CC_ERR_HIP 2         /* HIP runtime error (no GPU, launch failure) */
#define CC_ERR_ALLOC 3       /* host or device allocation failed           */
#define CC_ERR_NONFINITE 4   /* NaN / Inf detected in the state            */
#define CC_ERR_UNSUPPORTED 5 /* option not available in this build         */
#define CC_ERR_TRAVERSAL 6   /* a collision query overflowed its traversal
                                stack (taken as a miss) or hit its visit
                                bound (corrupted tree)                      */

/* spring record, s_spring in src/spring.h:8-12 (8 bytes) */
typedef struct cc_spring {
  int32_t end;      /* other end of the spring (vertex index) */
  float original;   /* rest length                            */
} cc_spring;

/* one CSR spring set: row[n+1] offsets into col[nnz]
   (Springs::CSR_spring output, src/spring.cpp:374-391) */
typedef struct cc_csr {
  uint32_t* row;
  cc_spring* col;
  uint32_t n;
  uint32_t nnz;
} cc_csr;

/* Physics and LBVH constants.  cc_params_default() gives the reference
   values: verlet.cu:13-21 (dt=1/40, mass=0.3, g=(0,-0.00981,0), damp=-0.02,
   k = {100 structure, 2 bend}), verlet.cu:62 (kd=-0.5), verlet.cu:73
   (eps=FLT_EPSILON), primitive.cu:44 (box depth 0.01) and
   simulation.cu:132 (body extend 0.003). */
typedef struct cc_params {
  float dt;
  float mass;
  float gravity[3];
  float damp;         /* force += vel * damp                         */
  float kd;           /* spring damping coefficient                  */
  float k[3];         /* stiffness of spring type 0, 1, 2            */
  float rest_scale;   /* rest length multiplier (1 in the reference) */
  float eps;          /* bias added to every deltaP component        */
  float box_depth;    /* expanded triangle box depth                 */
  float body_extend;  /* Mesh::vertex_extend distance at body set-up */
} cc_params;

void cc_params_default(cc_params* p);

/* Library version and last error (thread-local, never NULL). */
int cc_version(void);
const char* cc_last_error(void);
/* Number of visible HIP devices (0 when no GPU); never initialises a context
   beyond the runtime's own device enumeration. */
int cc_device_count(int* count);

/* ---------------------------------------------------------------- springs */

/* Springs::Springs + CSR_structure_spring + CSR_bend_spring
   (spring.cpp:82-91, 286-391): structure = unique 1-ring neighbours from
   face edges, bend = vertices opposite each shared edge.  verts4 is
   Mesh::vertices (vec4, w ignored), faces is Mesh::vertex_indices (3 per
   face).  Rest length = distance at set-up. */
int cc_springs_from_mesh(const float* verts4, uint32_t nv, const uint32_t* faces, uint32_t nf,
                         cc_csr* out_structure, cc_csr* out_bend);
/* Springs::Springs for a SINGLE_LAYER_BOUNDARY mesh (spring.cpp:82-91,
   98-209, kdtree.cpp): the vertex groups, in pairs {cloth piece, its
   boundary}, add boundary springs to the structural rows before the CSR is
   built: each boundary vertex to its nearest piece vertex (both directions),
   and, for boundary vertices whose nearest vertex on another boundary lies
   at 0 < d < 50 x the mean of the first 100 such ties, their piece vertices
   (one direction, sorted pairs).  Nearest neighbours follow the reference's
   kd-tree, ties included.  Rejected (the reference reads out of range):
   an odd group count, groups beyond nv, an empty piece, fewer than 100
   boundary vertices, one boundary alone.  cc_springs_from_mesh is the
   SINGLE_LAYER_NOB case the reference's scene uses (main.cpp:24). */
int cc_springs_from_mesh_groups(const float* verts4, uint32_t nv, const uint32_t* faces, uint32_t nf,
                                const uint32_t* group_counts, uint32_t ngroups, cc_csr* out_s, cc_csr* out_b);

/* Grid cloth springs: type 0 structural (4-neighbour), type 1 shear (both
   diagonals), type 2 bend (skip-one).  Vertex (i, j) is index j*nx + i;
   rows list their ends in ascending index order; rest = distance. */
int cc_springs_grid(uint32_t nx, uint32_t nz, const float* pos3, cc_csr out[3]);

void cc_csr_free(cc_csr* csr);

/* --------------------------------------------------------- scene helpers */
/* Synthetic inputs (the reference loads OBJ files; configs are synthetic).
   All seeded with splitmix64 so host and GPU box produce identical data. */

/* x-z grid at height y0: x = i*spacing, z = -0.5 + j*spacing, jitter each
   coordinate by a seeded uniform in [-jitter, jitter].  pos3[3*nx*nz]. */
int cc_scene_grid_cloth(uint32_t nx, uint32_t nz, double spacing, double y0, double jitter,
                        uint64_t seed, float* pos3);

/* Icosphere (outward CCW faces).  Sizes: V = 10*4^level + 2, F = 20*4^level.
   normals3 (optional) receives unit vertex normals. */
int cc_scene_icosphere(uint32_t level, double radius, const double center[3], float* verts3,
                       uint32_t* tris, float* normals3);

/* Capsule-union body description used by the per-frame deformation. */
typedef struct cc_capsules {
  uint32_t ncaps;
  uint32_t* vstart;   /* [ncaps+1] vertex range of each capsule   */
  float* p0;          /* [3*ncaps] axis start                      */
  float* p1;          /* [3*ncaps] axis end                        */
  float* radius;      /* [ncaps]                                   */
  float* phase;       /* [ncaps] breathing phase (radians)         */
} cc_capsules;

/* Seeded union of ncaps capsules with radii in [rmin, rmax] inside the box
   [lo, hi], tessellated with nseg segments around and nring rings per
   hemisphere (+ ncyl cylinder rings).  Query sizes with verts3 == NULL:
   *nv and *nf are always written. caps (optional) receives the layout;
   arrays in caps must hold ncaps (+1) entries. normals3 optional. */
int cc_scene_capsule_body(uint32_t ncaps, double rmin, double rmax, const double lo[3],
                          const double hi[3], uint32_t nseg, uint32_t nring, uint32_t ncyl,
                          uint64_t seed, float* verts3, uint32_t* tris, float* normals3,
                          uint32_t* nv, uint32_t* nf, cc_capsules* caps);

/* Mesh::vertex_extend (Mesh.cpp:101-117): move each vertex by dist along
   the normal of the first face corner that visits it.  normal_index may be
   NULL, meaning normal index == vertex index. stride = 3 or 4 floats. */
int cc_mesh_vertex_extend(float* verts, uint32_t stride, uint32_t nv, const uint32_t* faces,
                          uint32_t nf, const uint32_t* normal_index, const float* normals3,
                          float dist);

/* ------------------------------------------------------------ mesh ingest */

/* ObjLoader::ObjLoader (ObjLoader.cpp:15-216), geometry only: v (vec4,
   w = 1), vn, vt and the first three corners of each f line, indices
   0-based as the reference stores them (atoi - 1, unsigned; "f a b c"
   faces: texture 0, 1, 2 and normal index = vertex index; the reference's
   reading of "v//n" in the second and third corners is kept).  A corner
   without '/' in a "v/t/n" face is an error (undefined in the reference).
   Arrays are malloc'ed; release with cc_obj_free. */
typedef struct cc_obj {
  uint32_t nv, nvn, nvt, nf;
  float* verts4;      /* [4 nv]  */
  float* normals3;    /* [3 nvn] */
  float* tex2;        /* [2 nvt] */
  uint32_t* vidx;     /* [3 nf] vertex index per corner  */
  uint32_t* tidx;     /* [3 nf] texture index per corner */
  uint32_t* nidx;     /* [3 nf] normal index per corner  */
  /* vertex groups (ObjLoader::vertex_object, ObjLoader.cpp:34-41): a line
     whose second token is "object" opens a group, one whose third token is
     "vertices" sets the last group's vertex count (atoi of the second) */
  uint32_t ngroups;
  uint32_t* group_counts;   /* [ngroups] */
} cc_obj;
int cc_obj_read(const char* path, cc_obj* out);
void cc_obj_free(cc_obj* obj);

/* Mesh::unified (Mesh.cpp:17-38): per-vertex normals / texture coordinates
   from the face corners, the last corner naming a vertex wins; out-of-range
   indices (the reference reads past its arrays) leave the earlier value or
   zero.  Either output may be NULL; sizes 3 nv and 2 nv floats. */
int cc_mesh_unify(const cc_obj* obj, float* normals3, float* tex2);

/* Mesh::get_center / scale / translate / rotation (Mesh.cpp:40-92,
   119-131) on vec4 vertices in place: center = float sums in vertex order
   / n; scale about the center; rotation by degrees about the center, axis
   0 = x, 1 = y, 2 = z, with the reference's matrices (angle * 3.1415 / 180,
   glm column order). */
int cc_mesh_center(const float* verts4, uint32_t n, float center[3]);
int cc_mesh_scale(float* verts4, uint32_t n, float s);
int cc_mesh_translate(float* verts4, uint32_t n, float x, float y, float z);
int cc_mesh_rotate(float* verts4, uint32_t n, float degrees, int axis);

/* --------------------------------------------------------------- simulator */

typedef struct cc_sim cc_sim;

#define CC_BODY_NONE 0     /* no collider                                   */
#define CC_BODY_STATIC 1   /* LBVH built once (the reference's behaviour)   */
#define CC_BODY_REFIT 2    /* body deformed per frame, boxes refitted       */
#define CC_BODY_REBUILD 3  /* body deformed per frame, LBVH rebuilt         */

/* Cloth: Mesh::vertices (vec4) plus 1..3 typed CSR spring sets in the
   order their forces are summed (verlet.cu:148-149). */
typedef struct cc_cloth_desc {
  uint32_t n;
  const float* verts4;
  uint32_t ntypes;
  const uint32_t* row[3];
  const cc_spring* col[3];
} cc_cloth_desc;

/* Body: vertices (stride 3 or 4 floats) and triangles.  For REFIT and
   REBUILD, caps describes the capsule decomposition that the deformation
   breathes; motion: translation amp*sin(2 pi f t)*dir, rotation about the
   y axis through pivot by rot_amp*sin(2 pi f t), radial breathing
   1 + breath_amp*sin(2 pi breath_freq t + phase_k), t = frame*dt. */
typedef struct cc_body_desc {
  uint32_t nv;
  const float* verts;
  uint32_t stride;
  uint32_t nf;
  const uint32_t* tris;
  uint32_t mode;
  uint32_t extend;         /* 1: apply vertex_extend(body_extend) first */
  const float* normals3;   /* per-vertex normals for extend (or NULL)   */
  const cc_capsules* caps; /* required for REFIT / REBUILD              */
  double amp, freq, dir[3], rot_amp, pivot[3], breath_amp, breath_freq;
} cc_body_desc;

/* Simulator::Simulator(Mesh& cloth, Mesh& body) (simulation.cu:47-52):
   uploads the cloth (x_last = x_cur, contact = 0, simulation.cu:54-100),
   the springs (init_spring, :102-125) and builds the body LBVH
   (build_bvh, :127-138).  body may be NULL (no collider). */
int cc_sim_create(const cc_cloth_desc* cloth, const cc_body_desc* body, const cc_params* params,
                  cc_sim** out);
int cc_sim_destroy(cc_sim* sim);

/* Simulator::simulate (simulation.cu:141-150) for nsteps frames: per frame
   the fused Verlet + spring + collision kernel (verlet.cu:129-161) against
   the body pose of that frame, then the ping-pong swap (:243-251).  A
   REFIT / REBUILD body's poses and LBVHs are computed on the library's own
   streams ("update lanes") ahead of the cloth: whole blocks of
   cc_sim_block_steps steps (24 by default) are replayed from captured
   hipGraphs while the lanes compute the next block's poses; the remaining
   steps -- and every call with nsteps = 1, the reference's one call per
   frame -- are launched one by one, each issuing the pose of a frame a few
   frames ahead (two per lane) on a lane, so the body update of later frames
   runs beside this frame's step.  Asynchronous on stream (NULL: the
   simulator's own stream). */
int cc_sim_step(cc_sim* sim, int nsteps, void* stream);
int cc_sim_synchronize(cc_sim* sim);
/* Steps per captured block (even).  No counterpart in the reference, whose
   loop launches every step from the host. */
int cc_sim_block_steps(cc_sim* sim, int* steps);

/* Simulator::update_vertex (simulation.cu:253-257): x_cur = x_last = xyz. */
int cc_sim_update_vertex(cc_sim* sim, uint32_t idx, const float xyz[3]);

/* Pin vertices at their current position (x_last = x_cur, zero velocity);
   pinned vertices keep that position, store a zero contact normal and skip
   collision.  Not in the reference; required by the configurations. */
int cc_sim_set_pins(cc_sim* sim, const uint32_t* idx, uint32_t k);

/* Replace physics parameters (takes effect at the next step). */
int cc_sim_set_params(cc_sim* sim, const cc_params* params);

/* Host copies of x_cur / x_last / contact normal, AoS vec3 like the
   reference buffers (simulator.h:41-50).  Any pointer may be NULL. */
int cc_sim_get_state(cc_sim* sim, float* x_cur, float* x_prev, float* contact_n);
int cc_sim_set_state(cc_sim* sim, const float* x_cur, const float* x_prev, const float* contact_n);
int cc_sim_frame(cc_sim* sim, uint64_t* frame);
/* Checkpoint of the cloth to a binary file (no counterpart in the
   reference: Simulator::save is empty, simulation.cu:233-235).
   cc_sim_save writes x_cur, x_last, the contact normals, the pinned
   vertices, the parameters and the frame; cc_sim_load restores them into a
   simulator of the same cloth (same vertex count and spring types), after
   which it steps as the saved one would have.  CC_ERR_INVALID for another
   cloth, a file that is not such a checkpoint, or one cut short.  The
   body's pose follows from the frame (the rest pose and motion are the
   simulator's creation inputs and are not saved). */
int cc_sim_save(cc_sim* sim, const char* path);
int cc_sim_load(cc_sim* sim, const char* path);
int cc_sim_set_frame(cc_sim* sim, uint64_t frame);

/* Debug outputs of the next steps: pre-projection positions (vec3), the
   sorted leaf tested (-1 none) and the inside flag of the last step.  Off by
   default; enabling costs extra bytes per step. */
int cc_sim_set_debug(cc_sim* sim, int enable);
int cc_sim_get_debug(cc_sim* sim, float* pre3, int32_t* hit, uint8_t* inside);

/* Nodes-visited counter of the collision traversal (summed since reset). */
int cc_sim_set_counting(cc_sim* sim, int enable);
int cc_sim_get_counters(cc_sim* sim, uint64_t* nodes_visited, uint64_t* queries);

/* Finite check of x_cur: CC_ERR_NONFINITE if any NaN / Inf; before that,
   CC_ERR_TRAVERSAL if a step's collision query overflowed its traversal
   stack since the last check (the reference's D_BVH::intersect,
   cudaBRT.cu:243-245, has no overflow test; here the query is taken as a
   miss and recorded, and this call reports and clears the record). */
int cc_sim_check_finite(cc_sim* sim);
/* Test hook (no counterpart in the reference): lower the traversal stack
   limit (1..64, default 64, the most a walk over a valid tree defers) so
   that the overflow record and its report can be exercised. */
int cc_sim_set_stack_limit(cc_sim* sim, int limit);

/* ------------------------------------------------------------------ body */

/* Replace the body vertices (stride 3 floats, rest pose for the deformation)
   and rebuild the LBVH from them. */
int cc_body_set_vertices(cc_sim* sim, const float* verts3);
/* Current (deformed) body vertices, vec3. */
int cc_body_get_vertices(cc_sim* sim, float* verts3);
/* The float parameters of the current frame's deformation: R[9]
   row-major, T[3] (pivot folded in), breath[ncaps]. */
int cc_body_get_motion(cc_sim* sim, float* R, float* T, float* breath);
/* BVHAccel construction (bvh.cu:264-297) / refit (calculateBoudingBox,
   cudaBRT.cu:166-206) over the current body vertices. */
int cc_bvh_rebuild(cc_sim* sim);
int cc_bvh_refit(cc_sim* sim);
int cc_bvh_size(cc_sim* sim, uint32_t* leaves);

/* Traversal form of the body (no counterpart in the reference, whose
   D_BVH::intersect, cudaBRT.cu:239-285, walks the binary tree).  Width 4
   (the default) walks a four-wide form folded from the binary tree after
   every refit or rebuild, entries in the binary walk's order of
   consideration, so first hits and contacts are the same; width 2 walks the
   binary tree.  A static or refitted body whose root paths could defer more
   than 64 entries always walks the binary tree.  cc_bvh_width reports the
   form in use, the wide node count and the deferral bound of the tree (64,
   the stack limit, for a rebuilt body: its wide walk defers no more than
   the binary walk, whose stack holds at most the tree depth). */
int cc_sim_set_bvh_width(cc_sim* sim, int width);
int cc_bvh_width(cc_sim* sim, int* width, uint32_t* wide_nodes, uint32_t* stack_bound);

/* Export the LBVH in the reference encoding (cudaBRT.h:10-33): childA/B
   > 0 internal v-1, < 0 leaf -(v+1); parent v+1, 0 = root.  keys are the
   sorted (morton << 32 | primitive) keys, order the primitive of each leaf,
   boxes are [min xyz, max xyz]; root box of the whole tree in root6.
   Sizes: leaves m, internal m-1.  Any pointer may be NULL. */
int cc_bvh_export(cc_sim* sim, uint64_t* keys, uint32_t* order, int32_t* childA, int32_t* childB,
                  int32_t* parent_int, int32_t* parent_leaf, float* box_int, float* box_leaf,
                  float* root6);

/* Collision query + projection for n points (collision_response_projection,
   verlet.cu:24-50, without the Verlet part): hit = sorted leaf tested
   (-1 none), inside = projection applied, out_pos / out_n projected point
   and normal (zero when not inside). */
int cc_collide_points(cc_sim* sim, const float* pts3, uint32_t n, int32_t* hit, uint8_t* inside,
                      float* out_pos, float* out_n);

/* cc_collide_points without any device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under stream capture, serving loops).  Every pointer is DEVICE memory the
   caller owns: pts3 (3n floats) in; hit (n), inside (n), out_pos (3n),
   out_n (3n) out, with the meanings of cc_collide_points; status (may be
   NULL) receives the query's traversal faults (0 = none; bit 0: a stack
   overflow, that query taken as a miss).  Asynchronous on stream (NULL: the
   simulator's stream); it reads the body pose of the simulator's current
   frame, so order it after the steps it follows and before the next
   cc_sim_step.  At most 65536 points per call (CC_ERR_INVALID beyond): the
   traversal spill area is a fixed 15.7 MB array per GPU, held until the
   process ends and shared by every call on that GPU -- calls that overlap
   on one GPU race. */
int cc_collide_points_noalloc(cc_sim* sim, const float* pts3, uint32_t n, int32_t* hit, uint8_t* inside,
                              float* out_pos, float* out_n, uint32_t* status, void* stream);

/* ------------------------------------------------------------------- vbo */

/* cuda_update_vbo (simulation.cu:196-231) into a device buffer with the
   reference VBO layout [vec4 pos (w=1) x n][vec2 uv x n_tex][vec3 normal x n]
   (scene.cpp:115-123).  Vertex normal = normalised sum of adjacent face
   normals (verlet.cu:91-116, 163-181); faces[3*nf] are the cloth indices.
   lag = 1 reproduces the reference's use of the previous positions for face
   normals (simulation.cu:221).  uv bytes are left untouched. */
int cc_sim_set_faces(cc_sim* sim, const uint32_t* faces, uint32_t nf);

/* Picking (Scene::OnMouseDown, scene.cpp:365-436, and
   find_first_intersect_triangle, Ray.cpp:8-77): the cloth face (the faces of
   cc_sim_set_faces) that the ray origin + t dir meets with the smallest t,
   Moller-Trumbore without culling, ties (equal t) to the later face, t not
   required to be positive -- all as the reference -- and the face's first
   vertex, the one the reference then drags with update_vertex.  -1 when no
   face is met.  The reference tests its host mesh's initial positions
   (never updated by its simulation); this tests the current positions. */
int cc_sim_pick(cc_sim* sim, const float origin[3], const float dir[3], int32_t* face, int32_t* vertex, float* t);
int cc_sim_write_vbo(cc_sim* sim, void* dst_dev, size_t bytes, uint32_t n_tex, int lag);
/* Host copy of the vertex normals of the last cc_sim_write_vbo. */
int cc_sim_get_normals(cc_sim* sim, float* normals3);

/* ----------------------------------------------------------------- timing */

/* Per-kernel timing of nsteps steps with HIP events on the simulator
   stream, one kernel group at a time (each bracketed, nothing overlapped),
   averaged per launch in microseconds; step_us is then the per-step time of
   nsteps further steps through the normal (overlapped) path.  The
   simulation advances: the call steps the cloth 2 * nsteps frames forward
   (nsteps timed one group at a time, then nsteps through cc_sim_step).
   body_us is the deformation; refit_us (REFIT body) or build_us (REBUILD
   body) the tree update after it. */
typedef struct cc_timing {
  double step_us;      /* whole step (graph replay), per step           */
  double verlet_us;    /* fused Verlet + spring + collision kernel      */
  double body_us;      /* deformation kernel                            */
  double refit_us;     /* refit kernels                                 */
  double build_us;     /* rebuild kernels (boxes, sort, tree, four-wide
                          fold) summed                                   */
  uint32_t steps;
} cc_timing;
int cc_sim_time_kernels(cc_sim* sim, int nsteps, cc_timing* out);

/* Timing inside the normal path: while enabled, cc_sim_step records HIP
   events around each graph of fused step kernels, around the single fused
   step kernels of frames divisible by four, and around each block's body
   update lanes and the single-step updates of frames divisible by eight, on
   the streams they run on, and accumulates their durations (the host reads
   the events of earlier blocks and steps while later ones run; a launch
   that stamps events costs the host ~10 us more than a plain one, so
   single steps are sampled).
   cc_sim_get_timing returns the averages since enabling: verlet_us per
   fused step kernel, refit_us (refitted body) or build_us (rebuilt body)
   per body update including its deformation, steps = steps timed.
   Calling it again resets the averages.  No counterpart in the reference
   (measurement only). */
int cc_sim_set_timing(cc_sim* sim, int enable);
int cc_sim_get_timing(cc_sim* sim, cc_timing* out);

/* Fault injection, for tests of the error paths (no counterpart in the
   reference, whose safe_cuda prints an error and continues, common.h:4-10).
   After `count` further successful operations of the kind, the next one
   fails as if the runtime had failed it, and the switch is off again:
   CC_FAULT_ALLOC a device (or pinned host) allocation -> CC_ERR_ALLOC,
   CC_FAULT_LAUNCH the check after a kernel launch -> CC_ERR_HIP (the
   kernel itself was launched and runs).  count < 0 turns the kind off.
   Process-wide.  A call that failed this way leaves its object usable
   only for destruction (a failed creation frees what it had made). */
#define CC_FAULT_ALLOC 1
#define CC_FAULT_LAUNCH 2
int cc_fault_inject(int kind, int64_t count);

/* -------------------------------------------------------------- adjoint */

/* Reverse-mode gradient of a rollout (no counterpart in the reference: the
   differentiable pattern beside the plain step).  The rollout runs nsteps
   frames of the fused step over sim's cloth, springs, pins and body (static
   or none) from x_cur = x0, x_prev = x0 - v0*dt, contact 0, with
   theta = (k0, k1, k2, damp, rest_scale) replacing those of sim's
   parameters.  The gradient holds the forward's contacts fixed: a vertex
   projected in a step is moved by that fixed plane projection, and the
   contact normal entering the next step's force projection is a constant.
   The spring set must be symmetric (each i->j with a j->i of the same type
   and rest length).  States are checkpointed every ckpt_every steps and
   recomputed in the backward pass.  sim must outlive the handle. */
typedef struct cc_adj cc_adj;
int cc_adj_create(cc_sim* sim, uint32_t nsteps, uint32_t ckpt_every, cc_adj** out);
int cc_adj_destroy(cc_adj* adj);
/* Forward rollout; x_final (3n floats, may be NULL) receives the final
   positions; v0 may be NULL (zero). */
int cc_adj_forward(cc_adj* adj, const float* x0, const float* v0, const double theta[5], float* x_final);
/* Backward of the last forward: from dL/dx_final (3n) to dL/dtheta (FP64,
   reduced in FP64 on the GPU), dL/dv0 and dL/dx0 (3n each; outputs may be
   NULL). */
int cc_adj_backward(cc_adj* adj, const float* dl_dx_final, double dl_dtheta[5], float* dl_dv0, float* dl_dx0);

/* -------------------------------------------------------------- batched */

/* Many independent instances, each a cloth with its own springs, parameters
   and body, advanced together: one launch per step covers every instance,
   runs of 16 steps replay a captured hipGraph.  The reference runs one
   Simulator per scene (simulator.h:9-59); a batch equals count separate
   cc_sim handles stepped in turn, bit for bit.  The step's tile (vertices
   of one instance per block entry: 64, 128 or 256) and the order of the
   entries (instance order, or largest instance first) are picked at
   creation by timing each candidate on the launch the steps use -- the
   step kernel, or with moving bodies the per-body sort's launch that
   carries the step (cc_batch_info reports the choice, cc_batch_tuning the
   timings).  bodies may be NULL (no colliders) or
   hold count entries of any mode: REFIT / REBUILD bodies are posed (and
   refitted / rebuilt) at every frame for all instances together before the
   cloths step, as cc_sim_step does for one body. */
typedef struct cc_batch cc_batch;
int cc_batch_create(const cc_cloth_desc* cloths, const cc_body_desc* bodies, const cc_params* params, uint32_t count,
                    cc_batch** out);
int cc_batch_destroy(cc_batch* b);
/* cc_sim_step for every instance (asynchronous on stream, NULL: the batch's
   own stream). */
int cc_batch_step(cc_batch* b, int nsteps, void* stream);
int cc_batch_synchronize(cc_batch* b);
/* cc_sim_check_finite over every instance of the batch. */
int cc_batch_check(cc_batch* b);
/* cc_sim_set_pins / cc_sim_get_state for instance inst. */
int cc_batch_set_pins(cc_batch* b, uint32_t inst, const uint32_t* idx, uint32_t k);
int cc_batch_get_state(cc_batch* b, uint32_t inst, float* x_cur, float* x_prev, float* contact_n);
/* The body of instance inst at the current frame: vertices (vec3, may be
   NULL) and the LBVH in the encoding of cc_bvh_export. */
int cc_batch_body_export(cc_batch* b, uint32_t inst, float* verts3, uint64_t* keys, uint32_t* order,
                         int32_t* childA, int32_t* childB, int32_t* parent_int, int32_t* parent_leaf,
                         float* box_int, float* box_leaf, float* root6);
int cc_batch_frame(cc_batch* b, uint64_t* frame);
/* Total vertices, the chosen workgroup size (64, 128 or 256) and order
   (0: instance order, 1: largest instance first); any pointer may be NULL. */
int cc_batch_info(cc_batch* b, uint32_t* vertices, uint32_t* tile, uint32_t* order);
/* The tuner's timings in ms per step, ms[2 * t + o] for tile 64 / 128 / 256
   (t = 0, 1, 2) and order o; fused = 1 when they timed the sort's launch
   carrying the step (moving bodies).  Either pointer may be NULL. */
int cc_batch_tuning(cc_batch* b, float ms[6], uint32_t* fused);

#ifdef __cplusplus
}
#endif
#endif
