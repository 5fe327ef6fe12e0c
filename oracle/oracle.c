/*
 * oracle.c -- CPU restatement of the reference cloth path.  TEST INFRASTRUCTURE
 * ONLY (see oracle.h).  Build with -ffp-contract=off: the FP32 functions follow
 * the reference's glm operation order one operation at a time.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <assert.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* small FP32 vector helpers, glm 0.9.9.8 operation order                   */
/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z; } v3f;
typedef struct { double x, y, z; } v3d;

static inline v3f f3(const float* p) { v3f r = {p[0], p[1], p[2]}; return r; }
static inline v3f vadd(v3f a, v3f b) { v3f r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static inline v3f vsub(v3f a, v3f b) { v3f r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline v3f vmul(v3f a, float s) { v3f r = {a.x * s, a.y * s, a.z * s}; return r; }
static inline v3f vdiv(v3f a, float s) { v3f r = {a.x / s, a.y / s, a.z / s}; return r; }
static inline float vdot(v3f a, v3f b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline float vlen(v3f a) { return sqrtf(vdot(a, a)); }
static inline v3f vnorm(v3f a) { return vmul(a, 1.0f / sqrtf(vdot(a, a))); }
static inline v3f vcross(v3f a, v3f b) {
  v3f r = {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
  return r;
}

static inline v3d d3(const double* p) { v3d r = {p[0], p[1], p[2]}; return r; }
static inline v3d dadd(v3d a, v3d b) { v3d r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static inline v3d dsub(v3d a, v3d b) { v3d r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline v3d dmul(v3d a, double s) { v3d r = {a.x * s, a.y * s, a.z * s}; return r; }
static inline v3d ddiv(v3d a, double s) { v3d r = {a.x / s, a.y / s, a.z / s}; return r; }
static inline double ddot(v3d a, v3d b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline v3d dcross(v3d a, v3d b) {
  v3d r = {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
  return r;
}

/* ------------------------------------------------------------------------ */
/* open-addressing hash map u64 -> u32 (edge tables)                        */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t* keys; uint32_t* vals; uint8_t* used; size_t cap; } umap;

static void umap_init(umap* m, size_t n) {
  size_t cap = 16;
  while (cap < 2 * n + 16) cap <<= 1;
  m->cap = cap;
  m->keys = (uint64_t*)calloc(cap, sizeof(uint64_t));
  m->vals = (uint32_t*)calloc(cap, sizeof(uint32_t));
  m->used = (uint8_t*)calloc(cap, 1);
}
static void umap_free(umap* m) { free(m->keys); free(m->vals); free(m->used); }
static size_t umap_slot(const umap* m, uint64_t k) {
  uint64_t h = k * 0x9E3779B97F4A7C15ull;
  size_t s = (size_t)(h >> 20) & (m->cap - 1);
  while (m->used[s] && m->keys[s] != k) s = (s + 1) & (m->cap - 1);
  return s;
}
static int umap_get(const umap* m, uint64_t k, uint32_t* v) {
  size_t s = umap_slot(m, k);
  if (!m->used[s]) return 0;
  *v = m->vals[s];
  return 1;
}
static void umap_put(umap* m, uint64_t k, uint32_t v) {
  size_t s = umap_slot(m, k);
  m->used[s] = 1; m->keys[s] = k; m->vals[s] = v;
}
static inline uint64_t ekey(uint32_t a, uint32_t b) { return ((uint64_t)a << 32) | b; }

/* ------------------------------------------------------------------------ */
/* PRNG and generators                                                      */
/* ------------------------------------------------------------------------ */
uint64_t or_splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double or_uniform01(uint64_t* s) { return (double)(or_splitmix64(s) >> 11) * 0x1.0p-53; }

/* Horizontal x-z grid at height y0; vertex (i, j) has index j*nx + i,
   x = i*spacing, z = -0.5 + j*spacing; each coordinate jittered by a seeded
   uniform in [-jitter, jitter] (three draws per vertex, x then y then z). */
void or_grid_cloth(uint32_t nx, uint32_t nz, double spacing, double y0, double jitter,
                   uint64_t seed, float* pos3) {
  uint64_t st = seed;
  for (uint32_t j = 0; j < nz; ++j)
    for (uint32_t i = 0; i < nx; ++i) {
      size_t v = (size_t)j * nx + i;
      double base[3] = {i * spacing, y0, -0.5 + j * spacing};
      for (int c = 0; c < 3; ++c) {
        double u = or_uniform01(&st);
        pos3[3 * v + c] = (float)(base[c] + (2.0 * u - 1.0) * jitter);
      }
    }
}

static const int GRID_OFF[3][4][2] = {
    {{1, 0}, {0, 1}, {-1, 0}, {0, -1}},   /* structural */
    {{1, 1}, {-1, 1}, {1, -1}, {-1, -1}}, /* shear      */
    {{2, 0}, {0, 2}, {-2, 0}, {0, -2}},   /* bend       */
};

void or_grid_spring_counts(uint32_t nx, uint32_t nz, uint32_t nnz[3]) {
  uint64_t X = nx, Z = nz;
  nnz[0] = (uint32_t)(2 * ((X - 1) * Z + X * (Z - 1)));
  nnz[1] = (uint32_t)(4 * (X - 1) * (Z - 1));
  nnz[2] = (uint32_t)(2 * ((X > 2 ? X - 2 : 0) * Z + X * (Z > 2 ? Z - 2 : 0)));
}

static int cmp_i64(const void* a, const void* b) {
  int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

/* rest = glm::distance on vec4 with w equal: length(p_end - p_i) */
static float rest_len(const float* pos3, size_t i, size_t j) {
  v3f d = vsub(f3(pos3 + 3 * j), f3(pos3 + 3 * i));
  return vlen(d);
}

void or_grid_springs(uint32_t nx, uint32_t nz, const float* pos3,
                     uint32_t* row_s, or_spring* col_s,
                     uint32_t* row_sh, or_spring* col_sh,
                     uint32_t* row_b, or_spring* col_b) {
  uint32_t* rows[3] = {row_s, row_sh, row_b};
  or_spring* cols[3] = {col_s, col_sh, col_b};
  for (int t = 0; t < 3; ++t) {
    uint32_t k = 0;
    for (uint32_t j = 0; j < nz; ++j)
      for (uint32_t i = 0; i < nx; ++i) {
        size_t v = (size_t)j * nx + i;
        rows[t][v] = k;
        int64_t nb[4]; int cnt = 0;
        for (int o = 0; o < 4; ++o) {
          int64_t ii = (int64_t)i + GRID_OFF[t][o][0], jj = (int64_t)j + GRID_OFF[t][o][1];
          if (ii < 0 || jj < 0 || ii >= nx || jj >= nz) continue;
          nb[cnt++] = jj * nx + ii;
        }
        qsort(nb, cnt, sizeof(int64_t), cmp_i64);
        for (int o = 0; o < cnt; ++o) {
          cols[t][k].end = (int32_t)nb[o];
          cols[t][k].original = rest_len(pos3, v, (size_t)nb[o]);
          ++k;
        }
      }
    rows[t][(size_t)nx * nz] = k;
  }
}

static const double ICO_T = 1.6180339887498948482;
static const int ICO_F[20][3] = {
    {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
    {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
    {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
    {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};

/* Icosahedron (12 vertices (+-1, +-t, 0) cyclic), each level splits every
   triangle in four at edge midpoints pushed to the unit sphere; vertices are
   then scaled by radius and offset by c, rounded to float at the end. */
uint32_t or_icosphere(uint32_t level, double radius, const double c[3],
                      float* verts3, uint32_t* tris, uint32_t* ntris) {
  size_t maxv = 10 * ((size_t)1 << (2 * level)) + 2, maxf = 20 * ((size_t)1 << (2 * level));
  double* V = (double*)malloc(sizeof(double) * 3 * maxv);
  uint32_t* F = (uint32_t*)malloc(sizeof(uint32_t) * 3 * maxf);
  uint32_t* G = (uint32_t*)malloc(sizeof(uint32_t) * 3 * maxf);
  const double base[12][3] = {{-1, ICO_T, 0}, {1, ICO_T, 0}, {-1, -ICO_T, 0}, {1, -ICO_T, 0},
                              {0, -1, ICO_T}, {0, 1, ICO_T}, {0, -1, -ICO_T}, {0, 1, -ICO_T},
                              {ICO_T, 0, -1}, {ICO_T, 0, 1}, {-ICO_T, 0, -1}, {-ICO_T, 0, 1}};
  uint32_t nv = 12, nf = 20;
  for (int i = 0; i < 12; ++i) {
    double l = sqrt(base[i][0] * base[i][0] + base[i][1] * base[i][1] + base[i][2] * base[i][2]);
    for (int k = 0; k < 3; ++k) V[3 * i + k] = base[i][k] / l;
  }
  for (int f = 0; f < 20; ++f)
    for (int k = 0; k < 3; ++k) F[3 * f + k] = (uint32_t)ICO_F[f][k];
  for (uint32_t lv = 0; lv < level; ++lv) {
    umap mid; umap_init(&mid, (size_t)nf * 3);
    uint32_t nf2 = 0;
    for (uint32_t f = 0; f < nf; ++f) {
      uint32_t a = F[3 * f], b = F[3 * f + 1], cc = F[3 * f + 2], m[3];
      uint32_t e[3][2] = {{a, b}, {b, cc}, {cc, a}};
      for (int q = 0; q < 3; ++q) {
        uint32_t lo = e[q][0] < e[q][1] ? e[q][0] : e[q][1];
        uint32_t hi = e[q][0] < e[q][1] ? e[q][1] : e[q][0];
        uint32_t idx;
        if (!umap_get(&mid, ekey(lo, hi), &idx)) {
          double p[3], l = 0;
          for (int k = 0; k < 3; ++k) { p[k] = 0.5 * (V[3 * lo + k] + V[3 * hi + k]); l += p[k] * p[k]; }
          l = sqrt(l);
          for (int k = 0; k < 3; ++k) V[3 * nv + k] = p[k] / l;
          idx = nv++;
          umap_put(&mid, ekey(lo, hi), idx);
        }
        m[q] = idx;
      }
      uint32_t nt[4][3] = {{a, m[0], m[2]}, {b, m[1], m[0]}, {cc, m[2], m[1]}, {m[0], m[1], m[2]}};
      for (int q = 0; q < 4; ++q) { for (int k = 0; k < 3; ++k) G[3 * nf2 + k] = nt[q][k]; ++nf2; }
    }
    umap_free(&mid);
    uint32_t* t = F; F = G; G = t; nf = nf2;
  }
  for (uint32_t i = 0; i < nv; ++i)
    for (int k = 0; k < 3; ++k) verts3[3 * i + k] = (float)(c[k] + radius * V[3 * i + k]);
  memcpy(tris, F, sizeof(uint32_t) * 3 * nf);
  *ntris = nf;
  free(V); free(F); free(G);
  return nv;
}

/* ------------------------------------------------------------------------ */
/* a1-a2: Springs::create_neigh + CSR_spring                                */
/* ------------------------------------------------------------------------ */
typedef struct { uint32_t* v; uint32_t n, cap; } ulist;
static void ul_push(ulist* l, uint32_t x) {
  if (l->n == l->cap) { l->cap = l->cap ? 2 * l->cap : 8; l->v = (uint32_t*)realloc(l->v, sizeof(uint32_t) * l->cap); }
  l->v[l->n++] = x;
}
static int ul_has(const ulist* l, uint32_t x) {
  for (uint32_t i = 0; i < l->n; ++i) if (l->v[i] == x) return 1;
  return 0;
}

static float rest_len4(const float* verts4, uint32_t i, uint32_t j) {
  v3f d = vsub(f3(verts4 + 4 * (size_t)j), f3(verts4 + 4 * (size_t)i));
  /* vec4 dot is (x+y)+(z+w); with w-w == 0 this equals the vec3 order */
  return vlen(d);
}

/* kd_create / kd_insert3f / kd_nearest3f (kdtree.cpp:91-216, 395-520,
   820-850), recursive as there: nodes in insertion order, a point goes left
   when strictly below the node's coordinate on axis depth mod 3, double
   coordinates, the bounding box grown by every insertion; a search starts
   from the root as best, recurses into the nearer side (left when q <= the
   node's coordinate), tests the node (strictly closer), then the farther
   side if its box is strictly closer than the best. */
typedef struct { double p[3]; uint32_t data; int dir; int32_t l, r; } kdn;
typedef struct { kdn* n; uint32_t cnt, cap; double mn[3], mx[3]; } kdt;
static void kd_ins(kdt* t, const float* pf, uint32_t data) {
  if (t->cnt == t->cap) { t->cap = t->cap ? 2 * t->cap : 64; t->n = (kdn*)realloc(t->n, sizeof(kdn) * t->cap); }
  kdn* x = &t->n[t->cnt];
  for (int c = 0; c < 3; ++c) x->p[c] = (double)pf[c];
  x->data = data; x->l = x->r = -1; x->dir = 0;
  const int32_t id = (int32_t)t->cnt++;
  if (id == 0) { for (int c = 0; c < 3; ++c) t->mn[c] = t->mx[c] = x->p[c]; return; }
  int32_t cur = 0;
  for (;;) {
    kdn* c = &t->n[cur];
    int32_t* link = (t->n[id].p[c->dir] < c->p[c->dir]) ? &c->l : &c->r;
    if (*link < 0) { *link = id; t->n[id].dir = (c->dir + 1) % 3; break; }
    cur = *link;
  }
  for (int c = 0; c < 3; ++c) {
    if (t->n[id].p[c] < t->mn[c]) t->mn[c] = t->n[id].p[c];
    if (t->n[id].p[c] > t->mx[c]) t->mx[c] = t->n[id].p[c];
  }
}
static double kd_d2(const kdn* x, const double* q) {
  double d = 0; for (int c = 0; c < 3; ++c) d += (x->p[c] - q[c]) * (x->p[c] - q[c]); return d;
}
static void kd_near_rec(const kdt* t, int32_t i, const double* q, int32_t* best, double* bd, double* mn, double* mx) {
  const kdn* x = &t->n[i];
  const int dir = x->dir;
  int32_t nearer, farther; double *nc, *fc;
  if (q[dir] - x->p[dir] <= 0) { nearer = x->l; farther = x->r; nc = mx + dir; fc = mn + dir; }
  else { nearer = x->r; farther = x->l; nc = mn + dir; fc = mx + dir; }
  if (nearer >= 0) { double sv = *nc; *nc = x->p[dir]; kd_near_rec(t, nearer, q, best, bd, mn, mx); *nc = sv; }
  double d = kd_d2(x, q);
  if (d < *bd) { *best = i; *bd = d; }
  if (farther >= 0) {
    double sv = *fc; *fc = x->p[dir];
    double r = 0;
    for (int c = 0; c < 3; ++c) {
      if (q[c] < mn[c]) r += (mn[c] - q[c]) * (mn[c] - q[c]);
      else if (q[c] > mx[c]) r += (mx[c] - q[c]) * (mx[c] - q[c]);
    }
    if (r < *bd) kd_near_rec(t, farther, q, best, bd, mn, mx);
    *fc = sv;
  }
}
/* data of the nearest point; the tree must not be empty */
static uint32_t kd_nearest_of(const kdt* t, const float* qf) {
  const double q[3] = {qf[0], qf[1], qf[2]};
  double mn[3], mx[3];
  memcpy(mn, t->mn, sizeof mn); memcpy(mx, t->mx, sizeof mx);
  int32_t best = 0; double bd = kd_d2(&t->n[0], q);
  kd_near_rec(t, 0, q, &best, &bd, mn, mx);
  return t->n[best].data;
}

int32_t or_kd_nearest(uint32_t n, const float* pts3, uint32_t nq, const float* q3, uint32_t* out) {
  if (n == 0) return -1;
  kdt t = {0};
  for (uint32_t i = 0; i < n; ++i) kd_ins(&t, pts3 + 3 * (size_t)i, i);
  for (uint32_t k = 0; k < nq; ++k) out[k] = kd_nearest_of(&t, q3 + 3 * (size_t)k);
  free(t.n);
  return 0;
}

/* Springs::get_cloth_boundary_spring / get_boundary_boundary_spring
   (spring.cpp:98-209).  Returns the number of cloth-boundary pairs (cb,
   2 per pair, may be NULL to count) and writes the boundary set's pairs,
   sorted and unique, to bd (count in *nbd); negative for the inputs the
   reference reads out of range: -1 odd group count, -2 groups past nv,
   -3 an empty piece, -4 fewer than 100 boundary vertices, -5 a boundary
   with no other boundary. */
static int cmp_pair(const void* a, const void* b) {
  const uint32_t* x = (const uint32_t*)a; const uint32_t* y = (const uint32_t*)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
  return 0;
}
int32_t or_boundary_springs(uint32_t nv, const float* verts4, uint32_t ngroups, const uint32_t* counts,
                            uint32_t* cb, uint32_t* bd, uint32_t* nbd) {
  if (ngroups % 2) return -1;
  uint64_t tot = 0;
  for (uint32_t g = 0; g < ngroups; ++g) tot += counts[g];
  if (tot > nv) return -2;
  float* v3 = (float*)malloc(sizeof(float) * 3 * (size_t)(nv ? nv : 1));
  for (uint32_t i = 0; i < nv; ++i) memcpy(v3 + 3 * (size_t)i, verts4 + 4 * (size_t)i, 12);
  uint32_t ncb = 0, start = 0;
  uint32_t* pairs = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (size_t)(nv + 1));
  uint32_t* rs = (uint32_t*)malloc(sizeof(uint32_t) * (ngroups + 2));
  uint32_t* re = (uint32_t*)malloc(sizeof(uint32_t) * (ngroups + 2));
  uint32_t nr = 0;
  for (uint32_t g = 0; g < ngroups; g += 2) {
    if (counts[g] == 0) { free(v3); free(pairs); free(rs); free(re); return -3; }
    kdt t = {0};
    for (uint32_t i = 0; i < counts[g]; ++i) kd_ins(&t, v3 + 3 * (size_t)(start + i), start + i);
    start += counts[g];
    for (uint32_t i = 0; i < counts[g + 1]; ++i) {
      pairs[2 * ncb] = start + i;
      pairs[2 * ncb + 1] = kd_nearest_of(&t, v3 + 3 * (size_t)(start + i));
      ++ncb;
    }
    free(t.n);
    rs[nr] = start; re[nr] = start + counts[g + 1]; ++nr;
    start += counts[g + 1];
  }
  if (ncb < 100) { free(v3); free(pairs); free(rs); free(re); return -4; }
  float md = 0.0f;
  for (int i = 0; i < 100; ++i) md += rest_len4(verts4, pairs[2 * i], pairs[2 * i + 1]);
  md /= 100u;
  /* boundary vertex -> its piece vertex */
  uint32_t* piece = (uint32_t*)calloc(nv ? nv : 1, sizeof(uint32_t));
  for (uint32_t i = 0; i < ncb; ++i) piece[pairs[2 * i]] = pairs[2 * i + 1];
  uint32_t nb = 0, capb = 64;
  uint32_t* bb = (uint32_t*)malloc(sizeof(uint32_t) * 2 * capb);
  for (uint32_t i = 0; i < nr; ++i) {
    kdt t = {0};
    for (uint32_t j = 0; j < nr; ++j) {
      if (j == i) continue;
      for (uint32_t k = rs[j]; k < re[j]; ++k) kd_ins(&t, v3 + 3 * (size_t)k, k);
    }
    if (rs[i] < re[i] && t.cnt == 0) {
      free(t.n); free(v3); free(pairs); free(rs); free(re); free(piece); free(bb); return -5;
    }
    for (uint32_t k = rs[i]; k < re[i]; ++k) {
      uint32_t r = kd_nearest_of(&t, v3 + 3 * (size_t)k);
      float d = rest_len4(verts4, k, r);
      if (d < md * 50 && d > 0) {
        if (nb == capb) { capb *= 2; bb = (uint32_t*)realloc(bb, sizeof(uint32_t) * 2 * capb); }
        bb[2 * nb] = piece[k]; bb[2 * nb + 1] = piece[r]; ++nb;
      }
    }
    free(t.n);
  }
  qsort(bb, nb, 2 * sizeof(uint32_t), cmp_pair);
  uint32_t u = 0;
  for (uint32_t i = 0; i < nb; ++i)
    if (u == 0 || bb[2 * i] != bb[2 * (u - 1)] || bb[2 * i + 1] != bb[2 * (u - 1) + 1]) {
      bb[2 * u] = bb[2 * i]; bb[2 * u + 1] = bb[2 * i + 1]; ++u;
    }
  if (cb) memcpy(cb, pairs, sizeof(uint32_t) * 2 * ncb);
  if (bd) memcpy(bd, bb, sizeof(uint32_t) * 2 * u);
  *nbd = u;
  free(v3); free(pairs); free(rs); free(re); free(piece); free(bb);
  return (int32_t)ncb;
}

int or_springs_from_mesh_ex(uint32_t nv, const float* verts4, uint32_t nf, const uint32_t* faces,
                            uint32_t ncb, const uint32_t* cb, uint32_t nbd, const uint32_t* bd,
                            uint32_t* row_s, or_spring* col_s, uint32_t* nnz_s,
                            uint32_t* row_b, or_spring* col_b, uint32_t* nnz_b) {
  ulist* n1 = (ulist*)calloc(nv, sizeof(ulist));
  ulist* n2 = (ulist*)calloc(nv, sizeof(ulist));
  for (uint32_t f = 0; f < nf; ++f) {           /* spring.cpp:290-312 */
    const uint32_t* t = faces + 3 * (size_t)f;
    const int ord[6][2] = {{0, 1}, {0, 2}, {1, 0}, {1, 2}, {2, 0}, {2, 1}};
    for (int q = 0; q < 6; ++q)
      if (!ul_has(&n1[t[ord[q][0]]], t[ord[q][1]])) ul_push(&n1[t[ord[q][0]]], t[ord[q][1]]);
  }
  for (uint32_t i = 0; i < ncb; ++i) {          /* spring.cpp:314-323 */
    ul_push(&n1[cb[2 * i]], cb[2 * i + 1]);
    ul_push(&n1[cb[2 * i + 1]], cb[2 * i]);
  }
  for (uint32_t i = 0; i < nbd; ++i) ul_push(&n1[bd[2 * i]], bd[2 * i + 1]);
  /* spring.cpp:328-351 with Matrix::Insert_Matrix (:30-79): an edge seen
     again (either direction) pairs its first opposite vertex with this one */
  umap mat; umap_init(&mat, (size_t)nf * 3);
  uint32_t* pairs = (uint32_t*)malloc(sizeof(uint32_t) * 2 * 3 * (size_t)nf + 2);
  size_t np = 0;
  for (uint32_t f = 0; f < nf; ++f) {
    const uint32_t* t = faces + 3 * (size_t)f;
    const uint32_t ins[3][3] = {{t[0], t[1], t[2]}, {t[0], t[2], t[1]}, {t[1], t[2], t[0]}};
    for (int q = 0; q < 3; ++q) {
      uint32_t i = ins[q][0], j = ins[q][1], k = ins[q][2], val;
      if (umap_get(&mat, ekey(i, j), &val) || umap_get(&mat, ekey(j, i), &val)) {
        pairs[2 * np] = val; pairs[2 * np + 1] = k; ++np;
      } else {
        umap_put(&mat, ekey(i, j), k);
      }
    }
  }
  for (size_t p = 0; p < np; ++p) {
    ul_push(&n2[pairs[2 * p]], pairs[2 * p + 1]);
    ul_push(&n2[pairs[2 * p + 1]], pairs[2 * p]);
  }
  umap_free(&mat); free(pairs);
  /* CSR_spring, spring.cpp:374-391 */
  ulist* ns[2] = {n1, n2};
  uint32_t* rows[2] = {row_s, row_b};
  or_spring* cols[2] = {col_s, col_b};
  uint32_t* nnz[2] = {nnz_s, nnz_b};
  for (int t = 0; t < 2; ++t) {
    uint32_t k = 0;
    for (uint32_t i = 0; i < nv; ++i) {
      if (rows[t]) rows[t][i] = k;
      for (uint32_t q = 0; q < ns[t][i].n; ++q) {
        if (cols[t]) {
          cols[t][k].end = (int32_t)ns[t][i].v[q];
          cols[t][k].original = rest_len4(verts4, i, ns[t][i].v[q]);
        }
        ++k;
      }
    }
    if (rows[t]) rows[t][nv] = k;
    *nnz[t] = k;
  }
  for (uint32_t i = 0; i < nv; ++i) { free(n1[i].v); free(n2[i].v); }
  free(n1); free(n2);
  return 0;
}

int or_springs_from_mesh(uint32_t nv, const float* verts4, uint32_t nf, const uint32_t* faces,
                         uint32_t* row_s, or_spring* col_s, uint32_t* nnz_s,
                         uint32_t* row_b, or_spring* col_b, uint32_t* nnz_b) {
  return or_springs_from_mesh_ex(nv, verts4, nf, faces, 0, NULL, 0, NULL, row_s, col_s, nnz_s, row_b, col_b,
                                 nnz_b);
}

/* a5: the first face corner that visits a vertex decides its normal */
void or_vertex_extend(uint32_t nv, float* verts4, uint32_t nf, const uint32_t* faces,
                      const uint32_t* normal_index, const float* normals3, float dist) {
  uint8_t* visited = (uint8_t*)calloc(nv, 1);
  for (uint32_t f = 0; f < nf; ++f)
    for (int j = 0; j < 3; ++j) {
      uint32_t v = faces[3 * (size_t)f + j];
      if (!visited[v]) {
        const float* n = normals3 + 3 * (size_t)normal_index[3 * (size_t)f + j];
        /* vertices[v] += dist * vec4(n, 0) */
        verts4[4 * (size_t)v + 0] = verts4[4 * (size_t)v + 0] + dist * n[0];
        verts4[4 * (size_t)v + 1] = verts4[4 * (size_t)v + 1] + dist * n[1];
        verts4[4 * (size_t)v + 2] = verts4[4 * (size_t)v + 2] + dist * n[2];
        verts4[4 * (size_t)v + 3] = verts4[4 * (size_t)v + 3] + dist * 0.0f;
        visited[v] = 1;
      }
    }
  free(visited);
}

/* ------------------------------------------------------------------------ */
/* LBVH, FP32 exact order                                                   */
/* ------------------------------------------------------------------------ */
static inline void box_point(float b[6], v3f p) {
  b[0] = p.x; b[1] = p.y; b[2] = p.z; b[3] = p.x; b[4] = p.y; b[5] = p.z;
}
static inline void box_expand(float b[6], const float o[6]) {   /* bbox.h:62-73 */
  b[0] = fminf(b[0], o[0]); b[1] = fminf(b[1], o[1]); b[2] = fminf(b[2], o[2]);
  b[3] = fmaxf(b[3], o[3]); b[4] = fmaxf(b[4], o[4]); b[5] = fmaxf(b[5], o[5]);
}
static inline void box_expand_pt(float b[6], v3f p) {
  float o[6]; box_point(o, p); box_expand(b, o);
}
static inline void box_sentinel(float b[6]) {                    /* bbox.h:17-22 */
  b[0] = b[1] = b[2] = 100.0f; b[3] = b[4] = b[5] = -100.0f;
}

/* Primitive::d_get_expand_bbox, primitive.cu:30-51 + d_get_normal :96-102 */
void or_prim_box(const float* verts3, const uint32_t* tri, float depth, float box6[6]) {
  v3f a = f3(verts3 + 3 * (size_t)tri[0]), b = f3(verts3 + 3 * (size_t)tri[1]),
      c = f3(verts3 + 3 * (size_t)tri[2]);
  box_point(box6, a);
  box_expand_pt(box6, b);
  box_expand_pt(box6, c);
  v3f n = vnorm(vcross(vsub(b, a), vsub(c, a)));
  v3f dn = vmul(n, depth);
  box_expand_pt(box6, vsub(a, dn));
  box_expand_pt(box6, vsub(b, dn));
  box_expand_pt(box6, vsub(c, dn));
}

/* Threads of the loops below that are independent per item (boxes,
   deformation, refit subtrees); the results do not depend on it. */
static int g_threads = 1;
void or_set_threads(int n) { g_threads = n > 1 ? n : 1; }

void or_prim_boxes(const float* verts3, uint32_t m, const uint32_t* tris, float depth, float* boxes6) {
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
  for (uint32_t i = 0; i < m; ++i) or_prim_box(verts3, tris + 3 * (size_t)i, depth, boxes6 + 6 * (size_t)i);
}

void or_root_box(uint32_t m, const float* boxes6, float root6[6]) {
  box_sentinel(root6);
  for (uint32_t i = 0; i < m; ++i) box_expand(root6, boxes6 + 6 * (size_t)i);
}

uint32_t or_expand_bits(uint32_t v) {   /* bvh.cu:20-27 */
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

uint32_t or_morton3d(float x, float y, float z) {   /* bvh.cu:31-41 */
  x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
  y = fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
  z = fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
  uint32_t xx = or_expand_bits((uint32_t)x);
  uint32_t yy = or_expand_bits((uint32_t)y);
  uint32_t zz = or_expand_bits((uint32_t)z);
  return xx * 4 + yy * 2 + zz;
}

/* BBox::centroid + getUnitcubePosOf (bbox.h:97-127) */
uint32_t or_box_morton(const float box6[6], const float root6[6]) {
  v3f mn = {box6[0], box6[1], box6[2]}, mx = {box6[3], box6[4], box6[5]};
  v3f c = vdiv(vadd(mn, mx), 2.0f);
  v3f rmin = {root6[0], root6[1], root6[2]};
  v3f ext = {root6[3] - root6[0], root6[4] - root6[1], root6[5] - root6[2]};
  v3f o2 = vsub(c, rmin);
  v3f u = {0.0f, 0.0f, 0.0f};
  if (vlen(ext) != 0.0f) { u.x = o2.x / ext.x; u.y = o2.y / ext.y; u.z = o2.z / ext.z; }
  return or_morton3d(u.x, u.y, u.z);
}

void or_morton_codes(uint32_t m, const float* boxes6, const float root6[6], uint32_t* codes) {
  for (uint32_t i = 0; i < m; ++i) codes[i] = or_box_morton(boxes6 + 6 * (size_t)i, root6);
}

static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return (x > y) - (x < y);
}

void or_sort_keys(uint32_t m, const uint32_t* codes, uint64_t* keys, uint32_t* order) {
  for (uint32_t i = 0; i < m; ++i) keys[i] = ((uint64_t)codes[i] << 32) | i;
  qsort(keys, m, sizeof(uint64_t), cmp_u64);
  for (uint32_t i = 0; i < m; ++i) order[i] = (uint32_t)(keys[i] & 0xffffffffu);
}

uint32_t or_dedupe_sorted(uint32_t m, uint64_t* keys, uint32_t* order) {
  if (m == 0) return 0;
  uint32_t n = 1;
  uint32_t cur = (uint32_t)(keys[0] >> 32);
  for (uint32_t i = 1; i < m; ++i) {
    uint32_t c = (uint32_t)(keys[i] >> 32);
    if (c != cur) { keys[n] = keys[i]; order[n] = order[i]; ++n; cur = c; }
  }
  return n;
}

static inline int clz64(uint64_t x) { return __builtin_clzll(x); }

static inline int delta64(const uint64_t* k, int64_t m, int64_t i, int64_t j) {   /* cudaBRT.cu:9-19 */
  if (j < 0 || j >= m) return -1;
  return clz64(k[i] ^ k[j]);
}

void or_karras(uint32_t m, const uint64_t* keys, int32_t* childA, int32_t* childB,
               int32_t* parent_int, int32_t* parent_leaf) {
  if (m == 0) return;
  for (uint32_t i = 0; i < m; ++i) parent_leaf[i] = 0;
  for (uint32_t i = 0; i + 1 < m; ++i) parent_int[i] = 0;
  int64_t M = m;
  for (int64_t i = 0; i + 1 < M; ++i) {
    /* determineRange, cudaBRT.cu:24-53 */
    int d = delta64(keys, M, i, i + 1) - delta64(keys, M, i, i - 1);
    d = d > 0 ? 1 : -1;
    int dmin = delta64(keys, M, i, i - d);
    int64_t lmax = 2;
    while (delta64(keys, M, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
      if (delta64(keys, M, i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int64_t first = i <= j ? i : j, last = i <= j ? j : i;
    /* findSplit, cudaBRT.cu:67-110 */
    int common = clz64(keys[first] ^ keys[last]);
    int64_t split = first, step = last - first;
    do {
      step = (step + 1) >> 1;
      int64_t ns = split + step;
      if (ns < last) {
        int diff = (keys[first] >> (63 - common)) != (keys[ns] >> (63 - common));
        if (!diff) split = ns;
      }
    } while (step > 1);
    /* processInternalNode, cudaBRT.cu:138-160 */
    if (split == first) { childA[i] = (int32_t)(-split - 1); parent_leaf[split] = (int32_t)i + 1; }
    else { childA[i] = (int32_t)(split + 1); parent_int[split] = (int32_t)i + 1; }
    if (split + 1 == last) { childB[i] = (int32_t)(-(split + 1) - 1); parent_leaf[split + 1] = (int32_t)i + 1; }
    else { childB[i] = (int32_t)(split + 1 + 1); parent_int[split + 1] = (int32_t)i + 1; }
  }
}

static const float* child_box(int32_t c, const float* box_leaf, const float* box_int) {
  return c < 0 ? box_leaf + 6 * (size_t)(-(c + 1)) : box_int + 6 * (size_t)(c - 1);
}

/* Post-order refit (children before parents) of the subtree under x.  The
   stack holds at most two entries per level of the current path, and a
   Karras tree over unique 64-bit keys is less than 64 levels deep. */
enum { REFIT_STACK = 160 };
static void refit_subtree(int32_t x0, const int32_t* childA, const int32_t* childB, const float* box_leaf,
                          float* box_int, uint8_t* state) {
  int32_t stack[REFIT_STACK];
  int sp = 0;
  stack[sp++] = x0;
  while (sp) {
    assert(sp + 2 <= REFIT_STACK);
    int32_t x = stack[sp - 1];
    if (!state[x]) {
      state[x] = 1;
      if (childB[x] > 0) stack[sp++] = childB[x] - 1;
      if (childA[x] > 0) stack[sp++] = childA[x] - 1;
    } else {
      --sp;
      float* b = box_int + 6 * (size_t)x;   /* calculateBoudingBox, cudaBRT.cu:190-198 */
      box_sentinel(b);
      box_expand(b, child_box(childA[x], box_leaf, box_int));
      box_expand(b, child_box(childB[x], box_leaf, box_int));
    }
  }
}

void or_refit(uint32_t m, const int32_t* childA, const int32_t* childB, const int32_t* parent_int,
              const float* box_leaf, float* box_int) {
  (void)parent_int;
  if (m < 2) return;
  uint8_t* state = (uint8_t*)calloc(m, 1);
  /* The internal nodes of the top six levels in breadth-first order; the
     subtrees under the internal nodes below them are refitted first (in
     parallel when threads are set: they share no node), then the top nodes
     in reverse breadth-first order.  fmin / fmax unions are exact, so the
     boxes do not depend on the order. */
  enum { TOP_DEPTH = 6 };
  int32_t top[1 << (TOP_DEPTH + 1)], sub[1 << (TOP_DEPTH + 1)];
  int ntop = 0, nsub = 0;
  int32_t lvl[1 << (TOP_DEPTH + 1)], nxt[1 << (TOP_DEPTH + 1)];
  int nl = 0;
  lvl[nl++] = 0;
  for (int d = 0; d <= TOP_DEPTH && nl; ++d) {
    int nn = 0;
    for (int q = 0; q < nl; ++q) {
      const int32_t x = lvl[q];
      if (d == TOP_DEPTH) {
        sub[nsub++] = x;
        continue;
      }
      top[ntop++] = x;
      if (childA[x] > 0) nxt[nn++] = childA[x] - 1;
      if (childB[x] > 0) nxt[nn++] = childB[x] - 1;
    }
    memcpy(lvl, nxt, sizeof(int32_t) * (size_t)nn);
    nl = nn;
  }
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(dynamic, 1)
  for (int q = 0; q < nsub; ++q) refit_subtree(sub[q], childA, childB, box_leaf, box_int, state);
  for (int q = ntop - 1; q >= 0; --q) {
    const int32_t x = top[q];
    float* b = box_int + 6 * (size_t)x;
    box_sentinel(b);
    box_expand(b, child_box(childA[x], box_leaf, box_int));
    box_expand(b, child_box(childB[x], box_leaf, box_int));
  }
  free(state);
}

uint32_t or_bvh_build(const float* verts3, uint32_t m, const uint32_t* tris, float depth,
                      int compat_dedupe, uint64_t* keys, uint32_t* order, uint32_t* leaf_tri,
                      int32_t* childA, int32_t* childB, int32_t* parent_int, int32_t* parent_leaf,
                      float* box_leaf, float* box_int) {
  if (m == 0) return 0;
  float* boxes = (float*)malloc(sizeof(float) * 6 * m);
  uint32_t* codes = (uint32_t*)malloc(sizeof(uint32_t) * m);
  float root[6];
  or_prim_boxes(verts3, m, tris, depth, boxes);
  or_root_box(m, boxes, root);
  or_morton_codes(m, boxes, root, codes);
  or_sort_keys(m, codes, keys, order);
  uint32_t mk = compat_dedupe ? or_dedupe_sorted(m, keys, order) : m;
  for (uint32_t i = 0; i < mk; ++i) {
    memcpy(box_leaf + 6 * (size_t)i, boxes + 6 * (size_t)order[i], sizeof(float) * 6);
    memcpy(leaf_tri + 3 * (size_t)i, tris + 3 * (size_t)order[i], sizeof(uint32_t) * 3);
  }
  or_karras(mk, keys, childA, childB, parent_int, parent_leaf);
  or_refit(mk, childA, childB, parent_int, box_leaf, box_int);
  free(boxes); free(codes);
  return mk;
}

static inline int box_has(const float* b, float x, float y, float z) {   /* bbox.h:84-90 */
  return (x >= b[0] && x <= b[3]) && (y >= b[1] && y <= b[4]) && (z >= b[2] && z <= b[5]);
}
static inline int box_has_d(const float* b, double x, double y, double z) {
  return (x >= (double)b[0] && x <= (double)b[3]) && (y >= (double)b[1] && y <= (double)b[4]) &&
         (z >= (double)b[2] && z <= (double)b[5]);
}

#define OR_STACK 66

/* D_BVH::intersect, cudaBRT.cu:239-285 (left-first, leaf children first) */
int32_t or_query(const or_bvh* b, const float p[3], int32_t* visited) {
  int32_t nvis = 0;
  if (visited) *visited = 0;
  if (b == NULL || b->m == 0) return -1;
  if (b->m == 1) return box_has(b->box_leaf, p[0], p[1], p[2]) ? 0 : -1;
  int32_t stack[OR_STACK];
  int sp = 0;
  stack[sp++] = -1;
  int32_t node = 0;
  do {
    ++nvis;
    int32_t a = b->childA[node], c = b->childB[node];
    int la = a < 0, lb = c < 0;
    int32_t ia = la ? -(a + 1) : a - 1, ib = lb ? -(c + 1) : c - 1;
    const float* ba = la ? b->box_leaf + 6 * (size_t)ia : b->box_int + 6 * (size_t)ia;
    const float* bb = lb ? b->box_leaf + 6 * (size_t)ib : b->box_int + 6 * (size_t)ib;
    int oL = box_has(ba, p[0], p[1], p[2]);
    int oR = box_has(bb, p[0], p[1], p[2]);
    if (oL && la) { if (visited) *visited = nvis; return ia; }
    if (oR && lb) { if (visited) *visited = nvis; return ib; }
    int tL = oL && !la, tR = oR && !lb;
    if (!tL && !tR) node = stack[--sp];
    else {
      node = tL ? ia : ib;
      if (tL && tR) { assert(sp < OR_STACK); stack[sp++] = ib; }
    }
  } while (node != -1);
  if (visited) *visited = nvis;
  return -1;
}

int32_t or_query_f64(const or_bvh* b, const double p[3]) {
  if (b == NULL || b->m == 0) return -1;
  if (b->m == 1) return box_has_d(b->box_leaf, p[0], p[1], p[2]) ? 0 : -1;
  int32_t stack[OR_STACK];
  int sp = 0;
  stack[sp++] = -1;
  int32_t node = 0;
  do {
    int32_t a = b->childA[node], c = b->childB[node];
    int la = a < 0, lb = c < 0;
    int32_t ia = la ? -(a + 1) : a - 1, ib = lb ? -(c + 1) : c - 1;
    const float* ba = la ? b->box_leaf + 6 * (size_t)ia : b->box_int + 6 * (size_t)ia;
    const float* bb = lb ? b->box_leaf + 6 * (size_t)ib : b->box_int + 6 * (size_t)ib;
    int oL = box_has_d(ba, p[0], p[1], p[2]);
    int oR = box_has_d(bb, p[0], p[1], p[2]);
    if (oL && la) return ia;
    if (oR && lb) return ib;
    int tL = oL && !la, tR = oR && !lb;
    if (!tL && !tR) node = stack[--sp];
    else {
      node = tL ? ia : ib;
      if (tL && tR) { assert(sp < OR_STACK); stack[sp++] = ib; }
    }
  } while (node != -1);
  return -1;
}

int or_plane_test(const float* verts3, const uint32_t* tri, const float p[3], float* dist, float n[3]) {
  v3f a = f3(verts3 + 3 * (size_t)tri[0]), b = f3(verts3 + 3 * (size_t)tri[1]),
      c = f3(verts3 + 3 * (size_t)tri[2]);
  v3f nn = vnorm(vcross(vsub(b, a), vsub(c, a)));
  v3f pp = {p[0], p[1], p[2]};
  float d = vdot(vsub(pp, a), nn);
  *dist = d; n[0] = nn.x; n[1] = nn.y; n[2] = nn.z;
  return !(d > 0.0f);
}

/* ------------------------------------------------------------------------ */
/* a16-a18: the fused Verlet step                                           */
/* ------------------------------------------------------------------------ */
/* compute_spring_force, verlet.cu:52-88 */
static v3f spring_row_f32(const float* pos, const float* prev, const uint32_t* row, const or_spring* col,
                          uint32_t i, v3f p, v3f v, float ks, const or_params* P) {
  v3f f = {0.0f, 0.0f, 0.0f};
  for (uint32_t k = row[i]; k < row[i + 1]; ++k) {
    uint32_t j = (uint32_t)col[k].end;
    v3f p2 = f3(pos + 3 * (size_t)j), q2 = f3(prev + 3 * (size_t)j);
    v3f v2 = vdiv(vsub(p2, q2), P->dt);
    v3f dP = vsub(p, p2);
    dP.x = dP.x + P->eps; dP.y = dP.y + P->eps; dP.z = dP.z + P->eps;
    v3f dV = vsub(v, v2);
    float dist = vlen(dP);
    float rest = col[k].original * P->rest_scale;
    float left = -ks * (dist - rest);
    float right = P->kd * (vdot(dV, dP) / dist);
    f = vadd(f, vmul(vnorm(dP), left + right));
  }
  return f;
}

static void step_one_f32(uint32_t i, const float* pos, const float* prev, const float* contact,
                         const uint32_t* rows[3], const or_spring* cols[3], const or_params* P,
                         const uint8_t* pins, const or_bvh* bvh,
                         float* out_pos, float* out_prev, float* out_contact,
                         float* pre, int32_t* hit, uint8_t* inside) {
  v3f p = f3(pos + 3 * (size_t)i), q = f3(prev + 3 * (size_t)i);
  if (pins && pins[i]) {
    for (int c = 0; c < 3; ++c) {
      out_pos[3 * (size_t)i + c] = pos[3 * (size_t)i + c];
      out_prev[3 * (size_t)i + c] = pos[3 * (size_t)i + c];
      out_contact[3 * (size_t)i + c] = 0.0f;
      if (pre) pre[3 * (size_t)i + c] = pos[3 * (size_t)i + c];
    }
    if (hit) hit[i] = -1;
    if (inside) inside[i] = 0;
    return;
  }
  v3f vel = vdiv(vsub(p, q), P->dt);                                   /* :144 */
  v3f g = {P->gravity[0], P->gravity[1], P->gravity[2]};
  v3f F = vadd(vmul(g, P->mass), vmul(vel, P->damp));                  /* :147 */
  for (int t = 0; t < 3; ++t)                                          /* :148-149 */
    if (rows[t]) F = vadd(F, spring_row_f32(pos, prev, rows[t], cols[t], i, p, vel, P->k[t], P));
  v3f c = f3(contact + 3 * (size_t)i);
  F = vsub(F, vmul(c, vdot(c, F)));                                    /* :151-152 */
  v3f acc = vdiv(F, P->mass);                                          /* :153 */
  v3f np = vadd(vsub(vadd(p, p), q), vmul(vmul(acc, P->dt), P->dt));  /* :155 */
  v3f nq = p;                                                          /* :156 */
  v3f nc = {0.0f, 0.0f, 0.0f};
  if (pre) { pre[3 * (size_t)i] = np.x; pre[3 * (size_t)i + 1] = np.y; pre[3 * (size_t)i + 2] = np.z; }
  int32_t h = -1;
  uint8_t ins = 0;
  if (bvh) {                                                           /* verlet.cu:24-50 */
    float pp[3] = {np.x, np.y, np.z};
    h = or_query(bvh, pp, NULL);
    if (h >= 0) {
      float d, n[3];
      if (or_plane_test(bvh->verts, bvh->leaf_tri + 3 * (size_t)h, pp, &d, n)) {
        float dd = 1.0f * fabsf(d);
        v3f nn = {n[0], n[1], n[2]};
        np = vadd(np, vmul(nn, dd));
        nq = np;
        nc = nn;
        ins = 1;
      }
    }
  }
  if (hit) hit[i] = h;
  if (inside) inside[i] = ins;
  out_pos[3 * (size_t)i] = np.x; out_pos[3 * (size_t)i + 1] = np.y; out_pos[3 * (size_t)i + 2] = np.z;
  out_prev[3 * (size_t)i] = nq.x; out_prev[3 * (size_t)i + 1] = nq.y; out_prev[3 * (size_t)i + 2] = nq.z;
  out_contact[3 * (size_t)i] = nc.x; out_contact[3 * (size_t)i + 1] = nc.y; out_contact[3 * (size_t)i + 2] = nc.z;
}

void or_step_f32(uint32_t n, const float* pos, const float* prev, const float* contact,
                 const uint32_t* row0, const or_spring* col0,
                 const uint32_t* row1, const or_spring* col1,
                 const uint32_t* row2, const or_spring* col2,
                 const or_params* prm, const uint8_t* pins, const or_bvh* bvh,
                 float* out_pos, float* out_prev, float* out_contact,
                 float* pre, int32_t* hit, uint8_t* inside, int nthreads) {
  const uint32_t* rows[3] = {row0, row1, row2};
  const or_spring* cols[3] = {col0, col1, col2};
#ifdef _OPENMP
  if (nthreads > 1) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
      step_one_f32((uint32_t)i, pos, prev, contact, rows, cols, prm, pins, bvh, out_pos, out_prev,
                   out_contact, pre, hit, inside);
    return;
  }
#else
  (void)nthreads;
#endif
  for (uint32_t i = 0; i < n; ++i)
    step_one_f32(i, pos, prev, contact, rows, cols, prm, pins, bvh, out_pos, out_prev, out_contact,
                 pre, hit, inside);
}

void or_params_widen(const or_params* p, or_params64* q) {
  q->dt = p->dt;
  q->mass = p->mass;
  for (int c = 0; c < 3; ++c) {
    q->gravity[c] = p->gravity[c];
    q->k[c] = p->k[c];
  }
  q->damp = p->damp;
  q->kd = p->kd;
  q->rest_scale = p->rest_scale;
  q->eps = p->eps;
}

static v3d spring_row_f64(const double* pos, const double* prev, const uint32_t* row, const or_spring* col,
                          uint32_t i, v3d p, v3d v, double ks, const or_params64* P) {
  v3d f = {0.0, 0.0, 0.0};
  const double dt = P->dt, eps = P->eps, kd = P->kd, rs = P->rest_scale;
  for (uint32_t k = row[i]; k < row[i + 1]; ++k) {
    uint32_t j = (uint32_t)col[k].end;
    v3d p2 = d3(pos + 3 * (size_t)j), q2 = d3(prev + 3 * (size_t)j);
    v3d v2 = ddiv(dsub(p2, q2), dt);
    v3d dP = dsub(p, p2);
    dP.x += eps; dP.y += eps; dP.z += eps;
    v3d dV = dsub(v, v2);
    double dist = sqrt(ddot(dP, dP));
    double rest = (double)col[k].original * rs;
    double left = -ks * (dist - rest);
    double right = kd * (ddot(dV, dP) / dist);
    f = dadd(f, dmul(dmul(dP, 1.0 / sqrt(ddot(dP, dP))), left + right));
  }
  return f;
}

void or_step_f64(uint32_t n, const double* pos, const double* prev, const double* contact,
                 const uint32_t* row0, const or_spring* col0,
                 const uint32_t* row1, const or_spring* col1,
                 const uint32_t* row2, const or_spring* col2,
                 const or_params* P32, const uint8_t* pins, const or_bvh* bvh,
                 double* out_pos, double* out_prev, double* out_contact, uint8_t* inside,
                 int32_t* hit) {
  or_params64 P;
  or_params_widen(P32, &P);
  or_step_f64p(n, pos, prev, contact, row0, col0, row1, col1, row2, col2, &P, pins, bvh, out_pos, out_prev,
               out_contact, inside, hit);
}

/* verlet.cu:129-156 in FP64: the position before the collision response */
static v3d pre_position_f64(uint32_t i, const double* pos, const double* prev, const double* contact,
                            const uint32_t* const* rows, const or_spring* const* cols, const or_params64* P) {
  const double dt = P->dt, mass = P->mass, damp = P->damp;
  v3d p = d3(pos + 3 * (size_t)i), q = d3(prev + 3 * (size_t)i);
  v3d vel = ddiv(dsub(p, q), dt);
  v3d g = {P->gravity[0], P->gravity[1], P->gravity[2]};
  v3d F = dadd(dmul(g, mass), dmul(vel, damp));
  for (int t = 0; t < 3; ++t)
    if (rows[t]) F = dadd(F, spring_row_f64(pos, prev, rows[t], cols[t], i, p, vel, P->k[t], P));
  v3d c = d3(contact + 3 * (size_t)i);
  F = dsub(F, dmul(c, ddot(c, F)));
  v3d acc = ddiv(F, mass);
  return dadd(dsub(dadd(p, p), q), dmul(dmul(acc, dt), dt));
}

void or_step_f64p(uint32_t n, const double* pos, const double* prev, const double* contact,
                  const uint32_t* row0, const or_spring* col0,
                  const uint32_t* row1, const or_spring* col1,
                  const uint32_t* row2, const or_spring* col2,
                  const or_params64* P, const uint8_t* pins, const or_bvh* bvh,
                  double* out_pos, double* out_prev, double* out_contact, uint8_t* inside,
                  int32_t* hit) {
  or_step_f64_planes(n, pos, prev, contact, row0, col0, row1, col1, row2, col2, P, pins, bvh, out_pos, out_prev,
                     out_contact, inside, hit, NULL);
}

void or_step_f64_planes(uint32_t n, const double* pos, const double* prev, const double* contact,
                        const uint32_t* row0, const or_spring* col0,
                        const uint32_t* row1, const or_spring* col1,
                        const uint32_t* row2, const or_spring* col2,
                        const or_params64* P, const uint8_t* pins, const or_bvh* bvh,
                        double* out_pos, double* out_prev, double* out_contact, uint8_t* inside,
                        int32_t* hit, double* plane_pt) {
  const uint32_t* rows[3] = {row0, row1, row2};
  const or_spring* cols[3] = {col0, col1, col2};
  for (uint32_t i = 0; i < n; ++i) {
    if (plane_pt) plane_pt[3 * (size_t)i] = plane_pt[3 * (size_t)i + 1] = plane_pt[3 * (size_t)i + 2] = 0.0;
    v3d p = d3(pos + 3 * (size_t)i), q = d3(prev + 3 * (size_t)i);
    if (pins && pins[i]) {
      for (int c = 0; c < 3; ++c) {
        out_pos[3 * (size_t)i + c] = pos[3 * (size_t)i + c];
        out_prev[3 * (size_t)i + c] = pos[3 * (size_t)i + c];
        out_contact[3 * (size_t)i + c] = 0.0;
      }
      if (inside) inside[i] = 0;
      if (hit) hit[i] = -1;
      continue;
    }
    (void)q;
    v3d np = pre_position_f64(i, pos, prev, contact, rows, cols, P);
    v3d nq = p;
    v3d nc = {0.0, 0.0, 0.0};
    uint8_t ins = 0;
    int32_t h = -1;
    if (bvh) {
      double pp[3] = {np.x, np.y, np.z};
      h = or_query_f64(bvh, pp);
      if (h >= 0) {
        const uint32_t* tri = bvh->leaf_tri + 3 * (size_t)h;
        v3d a = {bvh->verts[3 * (size_t)tri[0]], bvh->verts[3 * (size_t)tri[0] + 1], bvh->verts[3 * (size_t)tri[0] + 2]};
        v3d b = {bvh->verts[3 * (size_t)tri[1]], bvh->verts[3 * (size_t)tri[1] + 1], bvh->verts[3 * (size_t)tri[1] + 2]};
        v3d cc = {bvh->verts[3 * (size_t)tri[2]], bvh->verts[3 * (size_t)tri[2] + 1], bvh->verts[3 * (size_t)tri[2] + 2]};
        v3d nf = dcross(dsub(b, a), dsub(cc, a));
        v3d nn = dmul(nf, 1.0 / sqrt(ddot(nf, nf)));
        double d = ddot(dsub(np, a), nn);
        if (!(d > 0.0)) {
          np = dadd(np, dmul(nn, fabs(d)));
          nq = np;
          nc = nn;
          ins = 1;
          if (plane_pt) {
            plane_pt[3 * (size_t)i] = a.x;
            plane_pt[3 * (size_t)i + 1] = a.y;
            plane_pt[3 * (size_t)i + 2] = a.z;
          }
        }
      }
    }
    if (inside) inside[i] = ins;
    if (hit) hit[i] = h;
    out_pos[3 * (size_t)i] = np.x; out_pos[3 * (size_t)i + 1] = np.y; out_pos[3 * (size_t)i + 2] = np.z;
    out_prev[3 * (size_t)i] = nq.x; out_prev[3 * (size_t)i + 1] = nq.y; out_prev[3 * (size_t)i + 2] = nq.z;
    out_contact[3 * (size_t)i] = nc.x; out_contact[3 * (size_t)i + 1] = nc.y; out_contact[3 * (size_t)i + 2] = nc.z;
  }
}

void or_rollout_f64(uint32_t n, const double* x0, const double* v0,
                    const uint32_t* row0, const or_spring* col0,
                    const uint32_t* row1, const or_spring* col1,
                    const uint32_t* row2, const or_spring* col2,
                    const or_params64* P, const uint8_t* pins, const or_bvh* bvh, uint32_t nsteps,
                    double* x_final, double* traj_p, double* traj_q, double* traj_c, double* traj_a) {
  const size_t N = 3 * (size_t)n;
  double* buf = (double*)calloc(6 * N, sizeof(double));
  double *p = buf, *q = buf + N, *c = buf + 2 * N, *np = buf + 3 * N, *nq = buf + 4 * N, *nc = buf + 5 * N;
  for (size_t k = 0; k < N; ++k) {
    p[k] = x0[k];
    q[k] = x0[k] - (v0 ? v0[k] : 0.0) * P->dt;
  }
  for (uint32_t t = 0; t <= nsteps; ++t) {
    if (traj_p) memcpy(traj_p + t * N, p, N * sizeof(double));
    if (traj_q) memcpy(traj_q + t * N, q, N * sizeof(double));
    if (traj_c) memcpy(traj_c + t * N, c, N * sizeof(double));
    if (t == nsteps) break;
    or_step_f64_planes(n, p, q, c, row0, col0, row1, col1, row2, col2, P, pins, bvh, np, nq, nc, NULL, NULL,
                       traj_a ? traj_a + (t + 1) * N : NULL);
    double* tmp;
    tmp = p; p = np; np = tmp;
    tmp = q; q = nq; nq = tmp;
    tmp = c; c = nc; nc = tmp;
  }
  memcpy(x_final, p, N * sizeof(double));
  free(buf);
}

void or_rollout_f64_frozen(uint32_t n, const double* x0, const double* v0,
                           const uint32_t* row0, const or_spring* col0,
                           const uint32_t* row1, const or_spring* col1,
                           const uint32_t* row2, const or_spring* col2,
                           const or_params64* P, const uint8_t* pins, uint32_t nsteps,
                           const double* traj_c, const double* traj_a, double* x_final,
                           double* traj_p, double* traj_q) {
  const uint32_t* rows[3] = {row0, row1, row2};
  const or_spring* cols[3] = {col0, col1, col2};
  const size_t N = 3 * (size_t)n;
  double* buf = (double*)calloc(4 * N, sizeof(double));
  double *p = buf, *q = buf + N, *np = buf + 2 * N, *nq = buf + 3 * N;
  for (size_t k = 0; k < N; ++k) {
    p[k] = x0[k];
    q[k] = x0[k] - (v0 ? v0[k] : 0.0) * P->dt;
  }
  for (uint32_t t = 0; t < nsteps; ++t) {
    if (traj_p) memcpy(traj_p + t * N, p, N * sizeof(double));
    if (traj_q) memcpy(traj_q + t * N, q, N * sizeof(double));
    const double* c = traj_c + t * N;
    const double* cn = traj_c + (t + 1) * N;
    const double* an = traj_a + (t + 1) * N;
    for (uint32_t i = 0; i < n; ++i) {
      v3d xp = d3(p + 3 * (size_t)i), u, uq;
      if (pins && pins[i]) {
        u = xp;
        uq = xp;
      } else {
        u = pre_position_f64(i, p, q, c, rows, cols, P);
        uq = xp;
        const v3d nn = d3(cn + 3 * (size_t)i);
        if (nn.x != 0.0 || nn.y != 0.0 || nn.z != 0.0) {   /* the recorded plane */
          const double d = ddot(dsub(u, d3(an + 3 * (size_t)i)), nn);
          u = dadd(u, dmul(nn, -d));     /* = |d| n on the side the forward saw */
          uq = u;
        }
      }
      np[3 * (size_t)i] = u.x; np[3 * (size_t)i + 1] = u.y; np[3 * (size_t)i + 2] = u.z;
      nq[3 * (size_t)i] = uq.x; nq[3 * (size_t)i + 1] = uq.y; nq[3 * (size_t)i + 2] = uq.z;
    }
    double* tmp;
    tmp = p; p = np; np = tmp;
    tmp = q; q = nq; nq = tmp;
  }
  if (traj_p) memcpy(traj_p + (size_t)nsteps * N, p, N * sizeof(double));
  if (traj_q) memcpy(traj_q + (size_t)nsteps * N, q, N * sizeof(double));
  memcpy(x_final, p, N * sizeof(double));
  free(buf);
}

/* Vector-Jacobian product of one spring force f = alpha * d/L,
   alpha = -ks (L - rest) + kd (dv . d)/L, d = p_i - p_j + eps, with the
   weight lam: g_d = (df/dd)^T lam, g_dv = (df/ddv)^T lam; dl = (d/L) . lam. */
static void spring_vjp_f64(v3d d, v3d dv, double ks, double rest, double kd, v3d lam, v3d* gd, v3d* gdv,
                           double* dl, double* len) {
  const double L = sqrt(ddot(d, d));
  const v3d u = dmul(d, 1.0 / L);
  const double ud = ddot(u, dv);
  const double alpha = -ks * (L - rest) + kd * ud;
  const double l = ddot(u, lam);
  /* d alpha / d d = -ks u + kd (dv - (u.dv) u) / L;  d u / d d = (I - u u^T) / L */
  const v3d da = dadd(dmul(u, -ks), dmul(dsub(dv, dmul(u, ud)), kd / L));
  *gd = dadd(dmul(da, l), dmul(dsub(lam, dmul(u, l)), alpha / L));
  *gdv = dmul(u, kd * l);
  *dl = l;
  *len = L;
}

void or_adjoint_f64(uint32_t n,
                    const uint32_t* row0, const or_spring* col0,
                    const uint32_t* row1, const or_spring* col1,
                    const uint32_t* row2, const or_spring* col2,
                    const or_params64* P, const uint8_t* pins, uint32_t nsteps,
                    const double* traj_p, const double* traj_q, const double* traj_c,
                    const double* dl_dxf, double dparams[5], double* dl_dv0, double* dl_dx0) {
  const uint32_t* rows[3] = {row0, row1, row2};
  const or_spring* cols[3] = {col0, col1, col2};
  const size_t N = 3 * (size_t)n;
  const double dt = P->dt, w = dt * dt / P->mass;
  double* buf = (double*)calloc(5 * N, sizeof(double));
  double *lp = buf, *lq = buf + N, *np = buf + 2 * N, *nq = buf + 3 * N, *lf = buf + 4 * N;
  double g[5] = {0, 0, 0, 0, 0};
  memcpy(lp, dl_dxf, N * sizeof(double));
  for (uint32_t tt = nsteps; tt-- > 0;) {
    const double* p = traj_p + tt * N;
    const double* q = traj_q + tt * N;
    const double* c = traj_c + tt * N;
    const double* cn = traj_c + (tt + 1) * N;
    memset(np, 0, 2 * N * sizeof(double));
    memset(lf, 0, N * sizeof(double));
    for (uint32_t i = 0; i < n; ++i) {
      const v3d lP = d3(lp + 3 * (size_t)i), lQ = d3(lq + 3 * (size_t)i);
      v3d gp = {0, 0, 0}, gq = {0, 0, 0};
      if (pins && pins[i]) {          /* x' = x'_prev = x */
        gp = dadd(lP, lQ);
      } else {
        const v3d nn = d3(cn + 3 * (size_t)i);
        v3d lu;
        if (nn.x != 0.0 || nn.y != 0.0 || nn.z != 0.0) {   /* projected: x' = x'_prev = plane projection of u */
          const v3d s2 = dadd(lP, lQ);
          lu = dsub(s2, dmul(nn, ddot(nn, s2)));
        } else {                                         /* x' = u, x'_prev = x */
          lu = lP;
          gp = lQ;
        }
        gp = dadd(gp, dmul(lu, 2.0));                    /* u = 2x - x_prev + w (I - c c^T) F */
        gq = dmul(lu, -1.0);
        const v3d cc = d3(c + 3 * (size_t)i);
        const v3d lF = dmul(dsub(lu, dmul(cc, ddot(cc, lu))), w);
        lf[3 * (size_t)i] = lF.x; lf[3 * (size_t)i + 1] = lF.y; lf[3 * (size_t)i + 2] = lF.z;
        const v3d vel = ddiv(dsub(d3(p + 3 * (size_t)i), d3(q + 3 * (size_t)i)), dt);
        g[3] += ddot(lF, vel);                           /* F += damp * vel */
        gp = dadd(gp, dmul(lF, P->damp / dt));
        gq = dsub(gq, dmul(lF, P->damp / dt));
      }
      np[3 * (size_t)i] += gp.x; np[3 * (size_t)i + 1] += gp.y; np[3 * (size_t)i + 2] += gp.z;
      nq[3 * (size_t)i] += gq.x; nq[3 * (size_t)i + 1] += gq.y; nq[3 * (size_t)i + 2] += gq.z;
    }
    for (uint32_t i = 0; i < n; ++i) {
      if (pins && pins[i]) continue;
      const v3d lF = d3(lf + 3 * (size_t)i);
      const v3d pi = d3(p + 3 * (size_t)i), vi = ddiv(dsub(pi, d3(q + 3 * (size_t)i)), dt);
      for (int t = 0; t < 3; ++t) {
        if (!rows[t]) continue;
        for (uint32_t k = rows[t][i]; k < rows[t][i + 1]; ++k) {
          const uint32_t j = (uint32_t)cols[t][k].end;
          const v3d pj = d3(p + 3 * (size_t)j), vj = ddiv(dsub(pj, d3(q + 3 * (size_t)j)), dt);
          v3d d = dsub(pi, pj);
          d.x += P->eps; d.y += P->eps; d.z += P->eps;
          const double rest0 = (double)cols[t][k].original, rest = rest0 * P->rest_scale;
          v3d gd, gdv;
          double l, L;
          spring_vjp_f64(d, dsub(vi, vj), P->k[t], rest, P->kd, lF, &gd, &gdv, &l, &L);
          const v3d a = dadd(gd, ddiv(gdv, dt)), b = ddiv(gdv, dt);
          np[3 * (size_t)i] += a.x; np[3 * (size_t)i + 1] += a.y; np[3 * (size_t)i + 2] += a.z;
          nq[3 * (size_t)i] -= b.x; nq[3 * (size_t)i + 1] -= b.y; nq[3 * (size_t)i + 2] -= b.z;
          np[3 * (size_t)j] -= a.x; np[3 * (size_t)j + 1] -= a.y; np[3 * (size_t)j + 2] -= a.z;
          nq[3 * (size_t)j] += b.x; nq[3 * (size_t)j + 1] += b.y; nq[3 * (size_t)j + 2] += b.z;
          g[t] += -(L - rest) * l;
          g[4] += P->k[t] * rest0 * l;
        }
      }
    }
    double* tmp;
    tmp = lp; lp = np; np = tmp;
    tmp = lq; lq = nq; nq = tmp;
  }
  if (dparams) memcpy(dparams, g, sizeof(g));
  for (size_t k = 0; k < N; ++k) {   /* x_cur = x0, x_prev = x0 - v0 dt */
    if (dl_dx0) dl_dx0[k] = lp[k] + lq[k];
    if (dl_dv0) dl_dv0[k] = -dt * lq[k];
  }
  free(buf);
}

void or_collide_points(uint32_t n, const float* pts, const or_bvh* bvh,
                       int32_t* hit, uint8_t* inside, float* out_pos, float* out_n) {
  for (uint32_t i = 0; i < n; ++i) {
    v3f p = f3(pts + 3 * (size_t)i);
    v3f nc = {0.0f, 0.0f, 0.0f};
    uint8_t ins = 0;
    float pp[3] = {p.x, p.y, p.z};
    int32_t h = or_query(bvh, pp, NULL);
    if (h >= 0) {
      float d, nn[3];
      if (or_plane_test(bvh->verts, bvh->leaf_tri + 3 * (size_t)h, pp, &d, nn)) {
        v3f nv = {nn[0], nn[1], nn[2]};
        p = vadd(p, vmul(nv, 1.0f * fabsf(d)));
        nc = nv;
        ins = 1;
      }
    }
    hit[i] = h;
    inside[i] = ins;
    out_pos[3 * (size_t)i] = p.x; out_pos[3 * (size_t)i + 1] = p.y; out_pos[3 * (size_t)i + 2] = p.z;
    out_n[3 * (size_t)i] = nc.x; out_n[3 * (size_t)i + 1] = nc.y; out_n[3 * (size_t)i + 2] = nc.z;
  }
}

/* ------------------------------------------------------------------------ */
/* per-frame body deformation                                               */
/* ------------------------------------------------------------------------ */
/* Vertex v of capsule k (cap_vstart[k] <= v < cap_vstart[k+1]) is scaled
   radially about its closest point a on the capsule axis [p0, p1]:
   Y = a + (X - a) * breath[k], then moved rigidly: out = R*Y + T. */
void or_body_deform(uint32_t nv, const float* rest3, uint32_t ncaps, const uint32_t* cap_vstart,
                    const float* cap_p0, const float* cap_p1, const float* breath,
                    const float R[9], const float T[3], float* out3) {
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
  for (uint32_t v = 0; v < nv; ++v) {
    uint32_t lo = 0, hi = ncaps - 1;     /* k = the last capsule with cap_vstart[k] <= v */
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) / 2;
      if (cap_vstart[mid] <= v) lo = mid; else hi = mid - 1;
    }
    const uint32_t k = lo;
    v3f X = f3(rest3 + 3 * (size_t)v);
    v3f p0 = f3(cap_p0 + 3 * (size_t)k), p1 = f3(cap_p1 + 3 * (size_t)k);
    v3f d = vsub(p1, p0);
    float dd = vdot(d, d);
    float t = 0.0f;
    if (dd > 0.0f) t = vdot(vsub(X, p0), d) / dd;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    v3f a = vadd(p0, vmul(d, t));
    v3f Y = vadd(a, vmul(vsub(X, a), breath[k]));
    out3[3 * (size_t)v + 0] = ((R[0] * Y.x + R[1] * Y.y) + R[2] * Y.z) + T[0];
    out3[3 * (size_t)v + 1] = ((R[3] * Y.x + R[4] * Y.y) + R[5] * Y.z) + T[1];
    out3[3 * (size_t)v + 2] = ((R[6] * Y.x + R[7] * Y.y) + R[8] * Y.z) + T[2];
  }
}

/* ---------------------------------------------------------------- picking */
/* rayTriangleIntersect, Moller-Trumbore without culling (Ray.cpp:8-42), in
   glm's operation order: 1 when the ray meets the triangle (t may be
   negative: the reference does not test it). */
static int ray_triangle(v3f o, v3f d, v3f v0, v3f v1, v3f v2, float* t) {
  v3f v0v1 = vsub(v1, v0), v0v2 = vsub(v2, v0);
  v3f pvec = {d.y * v0v2.z - v0v2.y * d.z, d.z * v0v2.x - v0v2.z * d.x, d.x * v0v2.y - v0v2.x * d.y};
  float det = vdot(v0v1, pvec);
  if (fabsf(det) < FLT_EPSILON) return 0;
  float invDet = 1 / det;
  v3f tvec = vsub(o, v0);
  float u = vdot(tvec, pvec) * invDet;
  if (u < 0 || u > 1) return 0;
  v3f qvec = {tvec.y * v0v1.z - v0v1.y * tvec.z, tvec.z * v0v1.x - v0v1.z * tvec.x,
              tvec.x * v0v1.y - v0v1.x * tvec.y};
  float v = vdot(d, qvec) * invDet;
  if (v < 0 || u + v > 1) return 0;
  *t = vdot(v0v2, qvec) * invDet;
  return 1;
}

/* find_first_intersect_triangle (Ray.cpp:45-77): the face with the smallest
   t, ties (equal t, -0 == +0) to the later face -- map<float, int> with
   isect_points[t] = i -- or -1.  A NaN t is skipped (the reference's map
   would be inconsistent). */
int32_t or_pick(uint32_t nf, const uint32_t* faces, const float* verts3, const float origin[3],
                const float dir[3], float* t_out) {
  v3f o = f3(origin), d = f3(dir);
  int32_t best = -1;
  float bt = 0.0f;
  for (uint32_t f = 0; f < nf; ++f) {
    const uint32_t* x = faces + 3 * (size_t)f;
    float t;
    if (!ray_triangle(o, d, f3(verts3 + 3 * (size_t)x[0]), f3(verts3 + 3 * (size_t)x[1]),
                      f3(verts3 + 3 * (size_t)x[2]), &t))
      continue;
    if (t != t) continue;
    if (best < 0 || t <= bt) { best = (int32_t)f; bt = t; }
  }
  if (t_out) *t_out = bt;
  return best;
}

/* ------------------------------------------------------------ vbo normals */
/* compute_face_normal + get_vertex_adjface + compute_vbo_normal
   (verlet.cu:91-116, 163-181, simulation.cu:152-180): face normal =
   normalize(cross(p1 - p0, p2 - p0)); vertex normal = normalize(sum of the
   normals of the faces naming it, in face order, once per corner). */
void or_vertex_normals(uint32_t n, const float* pos3, uint32_t nf, const uint32_t* faces, float* out3) {
  float* fn = (float*)malloc(sizeof(float) * 3 * (nf ? nf : 1));
  for (uint32_t f = 0; f < nf; ++f) {
    const uint32_t* x = faces + 3 * (size_t)f;
    v3f a = f3(pos3 + 3 * (size_t)x[0]), b = f3(pos3 + 3 * (size_t)x[1]), c = f3(pos3 + 3 * (size_t)x[2]);
    v3f nn = vnorm(vcross(vsub(b, a), vsub(c, a)));
    fn[3 * (size_t)f] = nn.x; fn[3 * (size_t)f + 1] = nn.y; fn[3 * (size_t)f + 2] = nn.z;
  }
  for (uint32_t i = 0; i < n; ++i) { out3[3 * (size_t)i] = out3[3 * (size_t)i + 1] = out3[3 * (size_t)i + 2] = 0.0f; }
  /* sums in face order per vertex: walk the faces once */
  for (uint32_t f = 0; f < nf; ++f)
    for (int j = 0; j < 3; ++j) {
      float* o = out3 + 3 * (size_t)faces[3 * (size_t)f + j];
      o[0] = o[0] + fn[3 * (size_t)f]; o[1] = o[1] + fn[3 * (size_t)f + 1]; o[2] = o[2] + fn[3 * (size_t)f + 2];
    }
  for (uint32_t i = 0; i < n; ++i) {
    v3f nn = vnorm(f3(out3 + 3 * (size_t)i));
    out3[3 * (size_t)i] = nn.x; out3[3 * (size_t)i + 1] = nn.y; out3[3 * (size_t)i + 2] = nn.z;
  }
  free(fn);
}
