/*
 * oracle.h -- CPU restatement of the reference cloth path (sutongkui/simulator).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, the smoke check and the CPU-baseline leg of
 * bench.py use it, and only as the checker (or as the timed CPU baseline).
 *
 * Two precisions are provided:
 *   - an FP32 exact-order transcription (built with -ffp-contract=off) that
 *     follows the reference's glm operation order, used for bit-exact checks
 *     of boxes, Morton keys, sorted order, tree topology, refit boxes, the
 *     first-hit traversal, the plane test and the FP32 step;
 *   - an FP64 single-threaded transcription of the Verlet + collision step,
 *     used for the position tolerances.
 *
 * Reference behaviour the oracle DEFINES (undefined or unportable upstream):
 *   1. root box = union of ALL expanded primitive boxes (bvh.cu:44-69 misses
 *      a tail and reads out of bounds when m % 128 == 0);
 *   2. sort keys are (uint64)morton << 32 | primitive index, every primitive
 *      is kept; delta() = clz64.  or_dedupe_sorted() reproduces the
 *      reference's de-duplication (Utilities.h:35-55) keeping the LOWEST
 *      index of each run of equal codes;
 *   3. no FMA contraction; glm 0.9.9.8 operation order
 *      dot = (x*x' + y*y') + z*z', length = sqrt(dot), normalize = v*(1/sqrt(dot));
 *   4. traversal stack depth is bounded by the key width (64) and checked;
 *   5. pinned vertices write their current position to both outputs, store a
 *      zero contact normal and skip collision (same effect as
 *      Simulator::update_vertex, simulation.cu:253-257, applied every step);
 *   6. m == 1: the single leaf box is tested directly (the reference indexes
 *      internal node 0 which does not exist, cudaBRT.cu:181-183, 290-293);
 *   7. boundary springs (spring.cpp:98-209) reject what the reference reads
 *      out of range or through a null kd-tree result: an odd group count,
 *      groups past the vertex count, an empty piece, fewer than 100
 *      boundary vertices, a boundary with no other boundary.
 *
 * Child / parent encoding follows the reference (cudaBRT.h:10-33):
 *   child > 0 -> internal node (v-1); child < 0 -> leaf -(v+1); 0 -> null;
 *   parent = v+1, 0 -> null (root).
 */
#ifndef CLOTH_ORACLE_H
#define CLOTH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* s_spring, spring.h:8-12 */
typedef struct { int32_t end; float original; } or_spring;

/* physical parameters; reference defaults in verlet.cu:13-21,62,73 */
typedef struct {
  float dt, mass;
  float gravity[3];
  float damp;        /* force += vel * damp   (verlet.cu:147, damp = -0.02) */
  float kd;          /* spring damping        (verlet.cu:62, -0.5) */
  float k[3];        /* stiffness per spring type (verlet.cu:13-14) */
  float rest_scale;  /* rest length multiplier (1.0 in the reference) */
  float eps;         /* bias added to deltaP (verlet.cu:73, FLT_EPSILON) */
} or_params;

/* The same parameters in double, for the FP64 path (finite differences
   need perturbations below float resolution). */
typedef struct {
  double dt, mass;
  double gravity[3];
  double damp, kd;
  double k[3];
  double rest_scale, eps;
} or_params64;

void or_params_widen(const or_params* p, or_params64* out);

/* A read-only view of a built LBVH, reference encoding. */
typedef struct {
  uint32_t m;               /* leaves */
  const float* verts;       /* body vertices, float3 */
  const uint32_t* leaf_tri; /* sorted triangles, 3 x u32 per leaf */
  const int32_t* childA;    /* [m-1] */
  const int32_t* childB;    /* [m-1] */
  const float* box_int;     /* [m-1][6] min xyz, max xyz */
  const float* box_leaf;    /* [m][6] */
} or_bvh;

/* --- scene generators (same PRNG as the product, written independently) --- */
uint64_t or_splitmix64(uint64_t* state);
double or_uniform01(uint64_t* state);
void or_grid_cloth(uint32_t nx, uint32_t nz, double spacing, double y0, double jitter,
                   uint64_t seed, float* pos3);
/* counts of directed springs per type for an nx*nz grid */
void or_grid_spring_counts(uint32_t nx, uint32_t nz, uint32_t nnz[3]);
/* structural / shear / bend CSR rows for a grid (rows ascending by end index) */
void or_grid_springs(uint32_t nx, uint32_t nz, const float* pos3,
                     uint32_t* row_s, or_spring* col_s,
                     uint32_t* row_sh, or_spring* col_sh,
                     uint32_t* row_b, or_spring* col_b);
/* icosphere with outward CCW faces; returns vertex count, fills tris */
uint32_t or_icosphere(uint32_t level, double radius, const double c[3],
                      float* verts3, uint32_t* tris, uint32_t* ntris);

/* --- a1-a2: Springs::create_neigh + CSR_spring (spring.cpp:286-391) --- */
/* verts4 is Mesh::vertices (vec4, w ignored).  Returns 0 on success.
   First call with col == NULL to get nnz, then with arrays allocated. */
int or_springs_from_mesh(uint32_t nv, const float* verts4, uint32_t nf, const uint32_t* faces,
                         uint32_t* row_s, or_spring* col_s, uint32_t* nnz_s,
                         uint32_t* row_b, or_spring* col_b, uint32_t* nnz_b);
/* the same with extra structural pairs appended as Springs::create_neigh
   does (spring.cpp:314-323): cloth-boundary pairs both ways, then the
   boundary set's pairs one way */
int or_springs_from_mesh_ex(uint32_t nv, const float* verts4, uint32_t nf, const uint32_t* faces,
                            uint32_t ncb, const uint32_t* cb, uint32_t nbd, const uint32_t* bd,
                            uint32_t* row_s, or_spring* col_s, uint32_t* nnz_s,
                            uint32_t* row_b, or_spring* col_b, uint32_t* nnz_b);
/* nearest of each query among pts in the reference kd-tree's order (ties
   as its search meets them); -1 for no points */
int32_t or_kd_nearest(uint32_t n, const float* pts3, uint32_t nq, const float* q3, uint32_t* out);
/* boundary springs of the vertex groups (spring.cpp:98-209): see oracle.c */
int32_t or_boundary_springs(uint32_t nv, const float* verts4, uint32_t ngroups, const uint32_t* counts,
                            uint32_t* cb, uint32_t* bd, uint32_t* nbd);

/* --- a5: Mesh::vertex_extend (Mesh.cpp:101-117) --- */
void or_vertex_extend(uint32_t nv, float* verts4, uint32_t nf, const uint32_t* faces,
                      const uint32_t* normal_index, const float* normals3, float dist);

/* --- LBVH, FP32 exact order --- */
void or_prim_box(const float* verts3, const uint32_t* tri, float depth, float box6[6]);
/* Threads for the per-item loops (boxes, deformation, refit subtrees);
   results do not depend on it.  Default 1. */
void or_set_threads(int n);
void or_prim_boxes(const float* verts3, uint32_t m, const uint32_t* tris, float depth, float* boxes6);
void or_root_box(uint32_t m, const float* boxes6, float root6[6]);
uint32_t or_expand_bits(uint32_t v);
uint32_t or_morton3d(float x, float y, float z);
uint32_t or_box_morton(const float box6[6], const float root6[6]);
void or_morton_codes(uint32_t m, const float* boxes6, const float root6[6], uint32_t* codes);
/* keys[i] sorted ascending; order[i] = primitive index of sorted position i */
void or_sort_keys(uint32_t m, const uint32_t* codes, uint64_t* keys, uint32_t* order);
/* reference de-duplication: returns kept count, compacts keys/order in place */
uint32_t or_dedupe_sorted(uint32_t m, uint64_t* keys, uint32_t* order);
/* Karras 2012 over 64-bit keys; parent_int[m-1], parent_leaf[m] */
void or_karras(uint32_t m, const uint64_t* keys, int32_t* childA, int32_t* childB,
               int32_t* parent_int, int32_t* parent_leaf);
/* bottom-up boxes: leaf boxes given, internal = sentinel U A U B */
void or_refit(uint32_t m, const int32_t* childA, const int32_t* childB, const int32_t* parent_int,
              const float* box_leaf, float* box_int);
/* full build; returns kept leaf count (== m unless compat_dedupe) */
uint32_t or_bvh_build(const float* verts3, uint32_t m, const uint32_t* tris, float depth,
                      int compat_dedupe, uint64_t* keys, uint32_t* order, uint32_t* leaf_tri,
                      int32_t* childA, int32_t* childB, int32_t* parent_int, int32_t* parent_leaf,
                      float* box_leaf, float* box_int);

/* a19: first-hit traversal (cudaBRT.cu:239-285); returns sorted leaf or -1.
   *visited (optional) receives the number of internal nodes visited. */
int32_t or_query(const or_bvh* b, const float p[3], int32_t* visited);
int32_t or_query_f64(const or_bvh* b, const double p[3]);
/* a20: Primitive::d_intersect (primitive.cu:70-85); returns 1 if !(d > 0) */
int or_plane_test(const float* verts3, const uint32_t* tri, const float p[3], float* dist, float n[3]);

/* --- a16-a18: one Verlet + spring + collision step --- */
/* rows/cols: up to 3 typed CSRs (NULL row = type absent).  pins: NULL or u8[n].
   bvh: NULL = no collider.  pre (optional) receives the pre-projection position,
   hit (optional) the sorted leaf index that was tested (-1 none) and
   inside (optional) 1 if the projection was applied. */
void or_step_f32(uint32_t n, const float* pos, const float* prev, const float* contact,
                 const uint32_t* row0, const or_spring* col0,
                 const uint32_t* row1, const or_spring* col1,
                 const uint32_t* row2, const or_spring* col2,
                 const or_params* prm, const uint8_t* pins, const or_bvh* bvh,
                 float* out_pos, float* out_prev, float* out_contact,
                 float* pre, int32_t* hit, uint8_t* inside, int nthreads);
void or_step_f64(uint32_t n, const double* pos, const double* prev, const double* contact,
                 const uint32_t* row0, const or_spring* col0,
                 const uint32_t* row1, const or_spring* col1,
                 const uint32_t* row2, const or_spring* col2,
                 const or_params* prm, const uint8_t* pins, const or_bvh* bvh,
                 double* out_pos, double* out_prev, double* out_contact, uint8_t* inside,
                 int32_t* hit);

/* or_step_f64 with double parameters (or_step_f64 widens and calls it). */
void or_step_f64p(uint32_t n, const double* pos, const double* prev, const double* contact,
                  const uint32_t* row0, const or_spring* col0,
                  const uint32_t* row1, const or_spring* col1,
                  const uint32_t* row2, const or_spring* col2,
                  const or_params64* prm, const uint8_t* pins, const or_bvh* bvh,
                  double* out_pos, double* out_prev, double* out_contact, uint8_t* inside,
                  int32_t* hit);

/* or_step_f64p that also returns, for every projected vertex, the point a
   of its plane (the tested triangle's first vertex; 0 for the others;
   plane_pt may be NULL). */
void or_step_f64_planes(uint32_t n, const double* pos, const double* prev, const double* contact,
                        const uint32_t* row0, const or_spring* col0,
                        const uint32_t* row1, const or_spring* col1,
                        const uint32_t* row2, const or_spring* col2,
                        const or_params64* prm, const uint8_t* pins, const or_bvh* bvh,
                        double* out_pos, double* out_prev, double* out_contact, uint8_t* inside,
                        int32_t* hit, double* plane_pt);

/* FP64 rollout of nsteps frames from x_cur = x0, x_prev = x0 - v0*dt,
   contact 0.  traj_p / traj_q / traj_c (optional, (nsteps+1) x 3n each)
   receive the state entering every step and the final one (traj_c[t]: the
   contact normals entering step t; traj_c[t+1]: those it produced);
   traj_a (optional, (nsteps+1) x 3n) the plane points of step t's
   projections at row t+1. */
void or_rollout_f64(uint32_t n, const double* x0, const double* v0,
                    const uint32_t* row0, const or_spring* col0,
                    const uint32_t* row1, const or_spring* col1,
                    const uint32_t* row2, const or_spring* col2,
                    const or_params64* prm, const uint8_t* pins, const or_bvh* bvh, uint32_t nsteps,
                    double* x_final, double* traj_p, double* traj_q, double* traj_c, double* traj_a);

/* The rollout with the contacts of a recorded one held fixed: a vertex with
   a contact normal n at traj_c[t+1] is projected onto the plane through
   traj_a[t+1] with normal n in step t whatever its position, the others
   never are; the force projection uses traj_c[t].  The function whose
   derivative or_adjoint_f64 computes; at the recorded parameters it equals
   the rollout bit for bit.  traj_p / traj_q (may be NULL): its positions
   and previous positions, (nsteps + 1) x n x 3. */
void or_rollout_f64_frozen(uint32_t n, const double* x0, const double* v0,
                           const uint32_t* row0, const or_spring* col0,
                           const uint32_t* row1, const or_spring* col1,
                           const uint32_t* row2, const or_spring* col2,
                           const or_params64* prm, const uint8_t* pins, uint32_t nsteps,
                           const double* traj_c, const double* traj_a, double* x_final,
                           double* traj_p, double* traj_q);

/* Adjoint of or_rollout_f64 along a recorded trajectory, with the contacts
   of the forward held fixed: a vertex projected in a step is moved by the
   fixed plane projection (I - n n^T) u + (n . a) n, and the contact normal
   that enters the next step's force projection is a constant.  Given
   dL/dx_final (3n), returns dL/d(k0, k1, k2, damp, rest_scale), dL/dv0 and
   dL/dx0 (3n each; any output may be NULL). */
void or_adjoint_f64(uint32_t n,
                    const uint32_t* row0, const or_spring* col0,
                    const uint32_t* row1, const or_spring* col1,
                    const uint32_t* row2, const or_spring* col2,
                    const or_params64* prm, const uint8_t* pins, uint32_t nsteps,
                    const double* traj_p, const double* traj_q, const double* traj_c,
                    const double* dl_dxf, double dparams[5], double* dl_dv0, double* dl_dx0);

/* collision query + projection for a batch of points (no Verlet) */
void or_collide_points(uint32_t n, const float* pts, const or_bvh* bvh,
                       int32_t* hit, uint8_t* inside, float* out_pos, float* out_n);

/* --- per-frame body deformation (capsule-union breathing + rigid motion) --- */
void or_body_deform(uint32_t nv, const float* rest3, uint32_t ncaps, const uint32_t* cap_vstart,
                    const float* cap_p0, const float* cap_p1, const float* breath,
                    const float R[9], const float T[3], float* out3);

/* find_first_intersect_triangle (Ray.cpp:45-77) over faces of verts3. */
int32_t or_pick(uint32_t nf, const uint32_t* faces, const float* verts3, const float origin[3],
                const float dir[3], float* t_out);

/* cloth vertex normals of the VBO (verlet.cu:91-116, 163-181). */
void or_vertex_normals(uint32_t n, const float* pos3, uint32_t nf, const uint32_t* faces, float* out3);

#ifdef __cplusplus
}
#endif
#endif
