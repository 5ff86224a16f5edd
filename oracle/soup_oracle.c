/*
 * soup_oracle.c -- CPU restatement of the protocell-genesis soup Langevin step.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the correctness oracle for the HIP library: it is linked
 * only by tests/, by __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg.
 * No product code path calls it.
 *
 * It restates, in plain C, the reference's WGSL kernels and TS host functions for the hot path,
 * in the reference's own data layout (particle-index order, list row i*cap+s) and in the canonical
 * deterministic linearization of the reference's racy parts:
 *   - the cell fill is a stable counting sort by (cell, particle index) (engine/wgsl/neighbor.wgsl
 *     fills by atomicAdd, whose order is one legal outcome of this);
 *   - bond formation runs the per-particle body sequentially for i = 0..N-1 (bond-dispatch.wgsl
 *     races the same body through compare-and-swap; the sequential order is one legal interleaving);
 *   - the constant-pH chain is serial exactly as the reference's CPU code is.
 *
 * f32 arithmetic follows the WGSL expression order with no contraction (built with
 * -ffp-contract=off). Two documented canonical choices, shared with the HIP kernels:
 *   - wca_cut uses f32(2^(1/6)) computed once on the host instead of WGSL pow(2, 0.16666667);
 *   - s6 = (b/r)^6 is (s*s)^3 instead of pow(b/r, 6);
 *   - the Box-Muller log/cos are evaluated in double on the f32 argument and rounded to f32.
 * Forces are additionally available in FP64 (same formula, f32 inputs promoted) as the accuracy
 * reference for the f32 forces.
 *
 * Pinned by: tests/test_oracle_*.py (closed forms, brute-force sets, finite differences, exact
 * enumeration and the reference's own known-answer constants).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/soup_abi.h"

#define NONE SOUP_NONE
#define BOND_RULES 2u     /* soup/wgsl/bond-common.wgsl:41 */
#define CANDIDATE_CAP 6u  /* soup/wgsl/bond-dispatch.wgsl:12 */
#define PI_F 3.14159265f  /* engine/wgsl/forces.wgsl:35 */

/* ------------------------------------------------------------------------------------------- */
/* State in the reference layout.                                                                */
/* ------------------------------------------------------------------------------------------- */
typedef struct OrState {
  float* pos;           /* 4n */
  float* vel;           /* 4n */
  float* force;         /* 4n */
  float* charge;        /* n */
  uint32_t* bond_slots; /* 3n */
  uint32_t* center_link;
  uint32_t* center_held;
  uint32_t* frozen;
  uint32_t* bond_rng;
  uint32_t* thermo_rng;
  uint32_t* events;     /* 8 */
  uint32_t* desorb;     /* 2 */
  uint32_t* cell_of;    /* n */
  uint32_t* cell_start; /* ncells + 1 */
  uint32_t* cell_idx;   /* n */
  uint32_t* vl;         /* n * list_cap */
  uint32_t* vcount;     /* n */
  float* pos_at_rebuild;/* 4n */
  uint32_t* head_idx;   /* n_heads */
  uint32_t* es_list;    /* n_heads * es_cap */
  uint32_t* es_count;   /* n_heads */
  uint32_t nh;          /* heads found */
  uint32_t v_overflow;
  uint32_t es_overflow;
  uint32_t max_drift_bits;
  uint32_t prng;        /* protonation PCG state */
  uint32_t pad;
  uint64_t global_step;
  uint64_t next_sweep_at;
} OrState;

/* ------------------------------------------------------------------------------------------- */
/* f32 helpers: engine/wgsl/forces.wgsl:17-37, soup/wgsl/pair.wgsl:64-160.                       */
/* ------------------------------------------------------------------------------------------- */
static inline float mi_f(float d, float L) { return d - rintf(d / L) * L; } /* pair.wgsl mi3 */

static inline float len3_f(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

static inline float wca_cut_f(const SoupConfig* c, float b) { return c->wca_factor * b; }

static inline float pair_b_f(const SoupConfig* c, float ti, float tj) {
  /* pair.wgsl pairB: P.sigma * (r_i + r_j) * 0.5 */
  return c->sigma * (c->radius[(uint32_t)ti] + c->radius[(uint32_t)tj]) * 0.5f;
}

static inline float wca_dv_f(const SoupConfig* c, float r, float b) {
  /* forces.wgsl:19-23 */
  if (r >= wca_cut_f(c, b) || r <= 0.0f) return 0.0f;
  float s = b / r;
  float s2 = s * s;
  float s6 = s2 * s2 * s2;
  return -24.0f * c->epsilon / r * (2.0f * s6 * s6 - s6);
}

static inline float fene_dv_f(const SoupConfig* c, float r) {
  /* forces.wgsl:25-28 */
  float x = r / c->r_inf;
  return c->k_fene * r / (1.0f - x * x);
}

static inline float bend_dv_f(const SoupConfig* c, float r) { return c->k_bend * (r - c->r_bend); }

static inline float attr_dv_f(const SoupConfig* c, float r) {
  /* forces.wgsl:32-37 */
  float rc = wca_cut_f(c, c->b_tt);
  if (r < rc || r > rc + c->wc) return 0.0f;
  float x = PI_F * (r - rc) / (2.0f * c->wc);
  return c->epsilon * PI_F * sinf(2.0f * x) / (2.0f * c->wc);
}

static inline float es_mag_f(const SoupConfig* c, float r, float qq) {
  /* soup/wgsl/electrostatics.wgsl:65-71 */
  if (c->es_A == 0.0f || qq == 0.0f || r >= c->es_rc) return 0.0f;
  float u = c->es_A * expf(-c->es_kappa * r) / r;
  return qq * (u * (1.0f / r + c->es_kappa) - c->es_Frc);
}

static inline float acid_soap_scale_f(const SoupConfig* c, float ti, float tj, float qi, float qj) {
  /* pair.wgsl acidSoapScale */
  float s = c->acid_soap;
  if (s <= 0.0f) return 0.0f;
  if (c->cls[(uint32_t)ti] != 1u || c->cls[(uint32_t)tj] != 1u) return 0.0f;
  if ((qi != 0.0f) == (qj != 0.0f)) return 0.0f;
  return s;
}

/* nonbondedSoup (soup/wgsl/pair.wgsl:168-204): force on i from j. */
static inline void nonbonded_f(const SoupConfig* c, const float* xi, const float* xj, float qi, float qj,
                               float f[3]) {
  float dx = mi_f(xi[0] - xj[0], c->box[0]);
  float dy = mi_f(xi[1] - xj[1], c->box[1]);
  float dz = mi_f(xi[2] - xj[2], c->box[2]);
  float r = len3_f(dx, dy, dz);
  if (r < 1e-6f) return;
  float ti = xi[3], tj = xj[3];
  float b = pair_b_f(c, ti, tj);
  if (r < wca_cut_f(c, b)) {
    float dv = wca_dv_f(c, r, b);
    f[0] = f[0] - dv * dx / r;
    f[1] = f[1] - dv * dy / r;
    f[2] = f[2] - dv * dz / r;
  }
  float as = c->attr[c->cls[(uint32_t)ti]][c->cls[(uint32_t)tj]] + acid_soap_scale_f(c, ti, tj, qi, qj);
  if (as > 0.0f) {
    float m = as * attr_dv_f(c, r);
    f[0] = f[0] - m * dx / r;
    f[1] = f[1] - m * dy / r;
    f[2] = f[2] - m * dz / r;
  }
  /* esForceNear (electrostatics.wgsl:76-79, :95-99) */
  if (r < c->es_split) {
    float qq = qi * qj;
    if (qq != 0.0f) {
      float m = es_mag_f(c, r, qq);
      f[0] = f[0] + m * dx / r;
      f[1] = f[1] + m * dy / r;
      f[2] = f[2] + m * dz / r;
    }
  }
}

/* bondedForce (soup/wgsl/step.wgsl:152-193). */
static void bonded_f(const SoupConfig* c, const OrState* s, uint32_t i, float f[3]) {
  const float* xi = s->pos + 4u * i;
  for (uint32_t k = 0; k < 3; ++k) {
    uint32_t p = s->bond_slots[3u * i + k];
    if (p == NONE) continue;
    float dx = mi_f(xi[0] - s->pos[4u * p], c->box[0]);
    float dy = mi_f(xi[1] - s->pos[4u * p + 1], c->box[1]);
    float dz = mi_f(xi[2] - s->pos[4u * p + 2], c->box[2]);
    float r = fmaxf(len3_f(dx, dy, dz), 1e-6f);
    float dv = fene_dv_f(c, r);
    f[0] = f[0] - dv * dx / r;
    f[1] = f[1] - dv * dy / r;
    f[2] = f[2] - dv * dz / r;
  }
  for (uint32_t k = 0; k < 3; ++k) {
    uint32_t m = s->bond_slots[3u * i + k];
    if (m == NONE) continue;
    for (uint32_t t = 0; t < 3; ++t) {
      uint32_t o = s->bond_slots[3u * m + t];
      if (o == NONE || o == i) continue;
      float dx = mi_f(xi[0] - s->pos[4u * o], c->box[0]);
      float dy = mi_f(xi[1] - s->pos[4u * o + 1], c->box[1]);
      float dz = mi_f(xi[2] - s->pos[4u * o + 2], c->box[2]);
      float r = fmaxf(len3_f(dx, dy, dz), 1e-6f);
      float dv = bend_dv_f(c, r);
      f[0] = f[0] - dv * dx / r;
      f[1] = f[1] - dv * dy / r;
      f[2] = f[2] - dv * dz / r;
    }
  }
  uint32_t owner = s->center_link[i];
  if (owner != NONE) {
    float dx = mi_f(xi[0] - s->pos[4u * owner], c->box[0]);
    float dy = mi_f(xi[1] - s->pos[4u * owner + 1], c->box[1]);
    float dz = mi_f(xi[2] - s->pos[4u * owner + 2], c->box[2]);
    float r = fmaxf(len3_f(dx, dy, dz), 1e-6f);
    float dv = fene_dv_f(c, r);
    f[0] = f[0] - dv * dx / r;
    f[1] = f[1] - dv * dy / r;
    f[2] = f[2] - dv * dz / r;
  }
}

/* ------------------------------------------------------------------------------------------- */
/* Grid: engine/wgsl/neighbor.wgsl:19-64 with a stable fill.                                    */
/* ------------------------------------------------------------------------------------------- */
uint32_t or_cell_of(const SoupConfig* c, const float* p) {
  /* neighbor.wgsl cell_of: floor((p / box) * n), clamped to [0, dims-1] */
  uint32_t cc[3];
  for (int a = 0; a < 3; ++a) {
    float f = floorf((p[a] / c->box[a]) * (float)c->dims[a]);
    if (!(f > 0.0f)) f = 0.0f; /* max(f, 0) (NaN -> 0) */
    uint32_t u = (f >= 4294967040.0f) ? 0xFFFFFFFFu : (uint32_t)f;
    cc[a] = u > c->dims[a] - 1u ? c->dims[a] - 1u : u;
  }
  return cc[0] + c->dims[0] * (cc[1] + c->dims[1] * cc[2]);
}

void or_grid_build(const SoupConfig* c, OrState* s) {
  uint32_t n = c->n, ncells = c->dims[0] * c->dims[1] * c->dims[2];
  memset(s->cell_start, 0, sizeof(uint32_t) * (ncells + 1));
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t cell = or_cell_of(c, s->pos + 4u * i);
    s->cell_of[i] = cell;
    s->cell_start[cell + 1]++;
  }
  for (uint32_t k = 0; k < ncells; ++k) s->cell_start[k + 1] += s->cell_start[k]; /* prefix_main */
  uint32_t* cursor = (uint32_t*)malloc(sizeof(uint32_t) * ncells);
  memcpy(cursor, s->cell_start, sizeof(uint32_t) * ncells);
  for (uint32_t i = 0; i < n; ++i) s->cell_idx[cursor[s->cell_of[i]]++] = i; /* stable fill */
  free(cursor);
}

/* cell_coord (forces.wgsl:215-219) + wrap_axis (:221-225). */
static inline void cell_coord(const SoupConfig* c, const float* p, int cc[3]) {
  for (int a = 0; a < 3; ++a) {
    float f = floorf((p[a] / c->box[a]) * (float)c->dims[a]);
    int v = (f != f) ? 0 : (f < -2147483648.0f ? INT32_MIN : (f >= 2147483648.0f ? INT32_MAX : (int)f));
    if (v < 0) v = 0;
    if (v > (int)c->dims[a] - 1) v = (int)c->dims[a] - 1;
    cc[a] = v;
  }
}
static inline int wrap_axis(int v, int n) {
  int r = v % n;
  if (r < 0) r += n;
  return r;
}

/* soup_build_verlet_list_main (soup/wgsl/verlet.wgsl:53-92). Returns the overflow flag. */
uint32_t or_verlet_build(const SoupConfig* c, OrState* s) {
  uint32_t n = c->n, cap = c->list_cap, overflow = 0;
  int R = (int)c->walk_radius;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : overflow)
  for (uint32_t i = 0; i < n; ++i) {
    const float* xi = s->pos + 4u * i;
    int cc[3];
    cell_coord(c, xi, cc);
    uint32_t count = 0;
    for (int dz = -R; dz <= R; ++dz) {
      int cz = wrap_axis(cc[2] + dz, (int)c->dims[2]);
      for (int dy = -R; dy <= R; ++dy) {
        int cy = wrap_axis(cc[1] + dy, (int)c->dims[1]);
        for (int dx = -R; dx <= R; ++dx) {
          int cx = wrap_axis(cc[0] + dx, (int)c->dims[0]);
          uint32_t nc = (uint32_t)cx + c->dims[0] * ((uint32_t)cy + c->dims[1] * (uint32_t)cz);
          for (uint32_t k = s->cell_start[nc]; k < s->cell_start[nc + 1]; ++k) {
            uint32_t j = s->cell_idx[k];
            if (j == i) continue;
            float ddx = mi_f(xi[0] - s->pos[4u * j], c->box[0]);
            float ddy = mi_f(xi[1] - s->pos[4u * j + 1], c->box[1]);
            float ddz = mi_f(xi[2] - s->pos[4u * j + 2], c->box[2]);
            if (len3_f(ddx, ddy, ddz) < c->list_range) {
              if (count < cap) {
                s->vl[(size_t)i * cap + count] = j;
                count++;
              } else {
                overflow = 1;
              }
            }
          }
        }
      }
    }
    s->vcount[i] = count;
  }
  s->v_overflow |= overflow;
  return overflow;
}

/* soup_build_head_index_main + soup_build_es_list_main (electrostatics-long.wgsl:15-56). */
uint32_t or_es_build(const SoupConfig* c, OrState* s) {
  uint32_t n = c->n, nh = 0, cap = c->es_cap, overflow = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (s->pos[4u * i + 3] == (float)c->charged_kind) {
      if (nh < c->n_heads) s->head_idx[nh] = i;
      nh++;
    }
  }
  if (nh > c->n_heads) overflow = 1;
  s->nh = nh;
  uint32_t m = nh < c->n_heads ? nh : c->n_heads;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : overflow)
  for (uint32_t h = 0; h < m; ++h) {
    uint32_t i = s->head_idx[h];
    const float* xi = s->pos + 4u * i;
    uint32_t count = 0;
    for (uint32_t g = 0; g < m; ++g) {
      uint32_t j = s->head_idx[g];
      if (j == i) continue;
      float dx = mi_f(xi[0] - s->pos[4u * j], c->box[0]);
      float dy = mi_f(xi[1] - s->pos[4u * j + 1], c->box[1]);
      float dz = mi_f(xi[2] - s->pos[4u * j + 2], c->box[2]);
      if (len3_f(dx, dy, dz) < c->es_list_range) {
        if (count < cap) {
          s->es_list[(size_t)h * cap + count] = j;
          count++;
        } else {
          overflow = 1;
        }
      }
    }
    s->es_count[h] = count;
  }
  s->es_overflow |= overflow;
  return overflow;
}

/* encodeVerletRebuild (soup/src/soup-grid-verlet.ts:61-73 / :18-54): grid, DH list, Verlet list,
 * snapshot, drift reset. */
void or_rebuild(const SoupConfig* c, OrState* s) {
  or_grid_build(c, s);
  if (c->es_enabled && c->n_heads > 0) or_es_build(c, s);
  or_verlet_build(c, s);
  memcpy(s->pos_at_rebuild, s->pos, sizeof(float) * 4u * c->n);
  s->max_drift_bits = 0;
}

/* ------------------------------------------------------------------------------------------- */
/* Forces.                                                                                       */
/* ------------------------------------------------------------------------------------------- */
/* soup_es_force_far_main (electrostatics-long.wgsl:63-82), f32, adds into force. */
static void es_far_list_f(const SoupConfig* c, OrState* s) {
  uint32_t m = s->nh < c->n_heads ? s->nh : c->n_heads;
#pragma omp parallel for schedule(dynamic, 64)
  for (uint32_t h = 0; h < m; ++h) {
    uint32_t i = s->head_idx[h];
    float qi = s->charge[i];
    if (qi == 0.0f) continue;
    const float* xi = s->pos + 4u * i;
    float f[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t k = 0; k < s->es_count[h]; ++k) {
      uint32_t j = s->es_list[(size_t)h * c->es_cap + k];
      float dx = mi_f(xi[0] - s->pos[4u * j], c->box[0]);
      float dy = mi_f(xi[1] - s->pos[4u * j + 1], c->box[1]);
      float dz = mi_f(xi[2] - s->pos[4u * j + 2], c->box[2]);
      float r = len3_f(dx, dy, dz);
      if (r < c->es_split) continue; /* esForceFar */
      float qq = qi * s->charge[j];
      if (qq == 0.0f) continue;
      float mg = es_mag_f(c, r, qq);
      f[0] = f[0] + mg * dx / r;
      f[1] = f[1] + mg * dy / r;
      f[2] = f[2] + mg * dz / r;
    }
    s->force[4u * i] += f[0];
    s->force[4u * i + 1] += f[1];
    s->force[4u * i + 2] += f[2];
  }
}

/* soup_force_list_main (soup/wgsl/verlet.wgsl:132-149) + far DH (f32). */
void or_force_list(const SoupConfig* c, OrState* s) {
  uint32_t n = c->n;
#pragma omp parallel for schedule(dynamic, 256)
  for (uint32_t i = 0; i < n; ++i) {
    float f[3] = {0.0f, 0.0f, 0.0f};
    bonded_f(c, s, i, f);
    const float* xi = s->pos + 4u * i;
    float qi = s->charge[i];
    for (uint32_t k = 0; k < s->vcount[i]; ++k) {
      uint32_t j = s->vl[(size_t)i * c->list_cap + k];
      nonbonded_f(c, xi, s->pos + 4u * j, qi, s->charge[j], f);
    }
    s->force[4u * i] = f[0];
    s->force[4u * i + 1] = f[1];
    s->force[4u * i + 2] = f[2];
    s->force[4u * i + 3] = 0.0f;
  }
  if (c->es_enabled && c->n_heads > 0) es_far_list_f(c, s);
}

/* soup_force_brute_main (step.wgsl:277-292) + soup_es_force_far_brute_main (electrostatics-long.wgsl
 * :87-103), f32. */
void or_force_brute(const SoupConfig* c, OrState* s) {
  uint32_t n = c->n;
#pragma omp parallel for schedule(dynamic, 64)
  for (uint32_t i = 0; i < n; ++i) {
    float f[3] = {0.0f, 0.0f, 0.0f};
    bonded_f(c, s, i, f);
    const float* xi = s->pos + 4u * i;
    float qi = s->charge[i];
    for (uint32_t j = 0; j < n; ++j) {
      if (j == i) continue;
      nonbonded_f(c, xi, s->pos + 4u * j, qi, s->charge[j], f);
    }
    if (c->es_enabled && c->n_heads > 0 && qi != 0.0f) {
      float g[3] = {0.0f, 0.0f, 0.0f};
      for (uint32_t j = 0; j < n; ++j) {
        if (j == i) continue;
        float dx = mi_f(xi[0] - s->pos[4u * j], c->box[0]);
        float dy = mi_f(xi[1] - s->pos[4u * j + 1], c->box[1]);
        float dz = mi_f(xi[2] - s->pos[4u * j + 2], c->box[2]);
        float r = len3_f(dx, dy, dz);
        if (r < c->es_split) continue;
        float qq = qi * s->charge[j];
        if (qq == 0.0f) continue;
        float mg = es_mag_f(c, r, qq);
        g[0] = g[0] + mg * dx / r;
        g[1] = g[1] + mg * dy / r;
        g[2] = g[2] + mg * dz / r;
      }
      f[0] = f[0] + g[0];
      f[1] = f[1] + g[1];
      f[2] = f[2] + g[2];
    }
    s->force[4u * i] = f[0];
    s->force[4u * i + 1] = f[1];
    s->force[4u * i + 2] = f[2];
    s->force[4u * i + 3] = 0.0f;
  }
}

/* FP64 recomputation of the same force from the same f32 inputs (positions, charges and the f32
 * constants promoted to double). O(N^2) over all pairs when use_list == 0, over the Verlet and DH
 * lists (supersets of the interaction ranges) when use_list != 0. Energy-consistent closed forms:
 * U_wca = eps((2 s6 - 1)^2), U_attr = -eps cos^2, U_fene, U_bend, U_dh shifted-force. */
static inline double mi_d(double d, double L) { return d - nearbyint(d / L) * L; }

static void nonbonded_d(const SoupConfig* c, const float* xi, const float* xj, double qi, double qj, double f[3]) {
  double dx = mi_d((double)xi[0] - (double)xj[0], c->box[0]);
  double dy = mi_d((double)xi[1] - (double)xj[1], c->box[1]);
  double dz = mi_d((double)xi[2] - (double)xj[2], c->box[2]);
  double r = sqrt(dx * dx + dy * dy + dz * dz);
  if (r < 1e-6) return;
  uint32_t ki = (uint32_t)xi[3], kj = (uint32_t)xj[3];
  double b = (double)c->sigma * ((double)c->radius[ki] + (double)c->radius[kj]) * 0.5;
  double wf = (double)c->wca_factor;
  double m = 0.0;
  if (r < wf * b) {
    double s6 = pow(b / r, 6.0);
    m += -(-24.0 * (double)c->epsilon / r * (2.0 * s6 * s6 - s6));
  }
  double as = (double)c->attr[c->cls[ki]][c->cls[kj]];
  if (c->acid_soap > 0.0f && c->cls[ki] == 1u && c->cls[kj] == 1u && ((qi != 0.0) != (qj != 0.0)))
    as += (double)c->acid_soap;
  if (as > 0.0) {
    double rc = wf * (double)c->b_tt, wc = (double)c->wc;
    if (!(r < rc || r > rc + wc)) {
      double x = M_PI * (r - rc) / (2.0 * wc);
      m += -as * (double)c->epsilon * M_PI * sin(2.0 * x) / (2.0 * wc);
    }
  }
  double qq = qi * qj;
  if (r < (double)c->es_split && qq != 0.0 && c->es_A != 0.0f && r < (double)c->es_rc) {
    double k = (double)c->es_kappa;
    double u = (double)c->es_A * exp(-k * r) / r;
    m += qq * (u * (1.0 / r + k) - (double)c->es_Frc);
  }
  f[0] += m * dx / r;
  f[1] += m * dy / r;
  f[2] += m * dz / r;
}

static void bonded_d(const SoupConfig* c, const OrState* s, uint32_t i, double f[3]) {
  const float* xi = s->pos + 4u * i;
  double kf = c->k_fene, ri = c->r_inf, kb = c->k_bend, rb = c->r_bend;
  /* FENE over the three bond slots (k < 3) and the centerLink tether (k == 3) */
  for (uint32_t k = 0; k < 4; ++k) {
    uint32_t o = k < 3 ? s->bond_slots[3u * i + k] : s->center_link[i];
    if (o == NONE) continue;
    double dx = mi_d((double)xi[0] - s->pos[4u * o], c->box[0]);
    double dy = mi_d((double)xi[1] - s->pos[4u * o + 1], c->box[1]);
    double dz = mi_d((double)xi[2] - s->pos[4u * o + 2], c->box[2]);
    double r = fmax(sqrt(dx * dx + dy * dy + dz * dz), 1e-6);
    double x = r / ri;
    double dv = kf * r / (1.0 - x * x);
    f[0] -= dv * dx / r;
    f[1] -= dv * dy / r;
    f[2] -= dv * dz / r;
  }
  for (uint32_t k = 0; k < 3; ++k) {
    uint32_t m = s->bond_slots[3u * i + k];
    if (m == NONE) continue;
    for (uint32_t t = 0; t < 3; ++t) {
      uint32_t o = s->bond_slots[3u * m + t];
      if (o == NONE || o == i) continue;
      double dx = mi_d((double)xi[0] - s->pos[4u * o], c->box[0]);
      double dy = mi_d((double)xi[1] - s->pos[4u * o + 1], c->box[1]);
      double dz = mi_d((double)xi[2] - s->pos[4u * o + 2], c->box[2]);
      double r = fmax(sqrt(dx * dx + dy * dy + dz * dz), 1e-6);
      double dv = kb * (r - rb);
      f[0] -= dv * dx / r;
      f[1] -= dv * dy / r;
      f[2] -= dv * dz / r;
    }
  }
}

static void es_far_d(const SoupConfig* c, const OrState* s, uint32_t i, uint32_t j, double f[3]) {
  double qi = s->charge[i], qj = s->charge[j];
  double qq = qi * qj;
  if (qq == 0.0 || c->es_A == 0.0f) return;
  const float* xi = s->pos + 4u * i;
  double dx = mi_d((double)xi[0] - s->pos[4u * j], c->box[0]);
  double dy = mi_d((double)xi[1] - s->pos[4u * j + 1], c->box[1]);
  double dz = mi_d((double)xi[2] - s->pos[4u * j + 2], c->box[2]);
  double r = sqrt(dx * dx + dy * dy + dz * dz);
  if (r < (double)c->es_split || r >= (double)c->es_rc || r <= 0.0) return;
  double k = (double)c->es_kappa;
  double u = (double)c->es_A * exp(-k * r) / r;
  double m = qq * (u * (1.0 / r + k) - (double)c->es_Frc);
  f[0] += m * dx / r;
  f[1] += m * dy / r;
  f[2] += m * dz / r;
}

void or_force_f64(const SoupConfig* c, OrState* s, int use_list, double* out /* 3n */) {
  uint32_t n = c->n;
#pragma omp parallel for schedule(dynamic, 64)
  for (uint32_t i = 0; i < n; ++i) {
    double f[3] = {0.0, 0.0, 0.0};
    bonded_d(c, s, i, f);
    const float* xi = s->pos + 4u * i;
    double qi = s->charge[i];
    if (use_list) {
      for (uint32_t k = 0; k < s->vcount[i]; ++k) {
        uint32_t j = s->vl[(size_t)i * c->list_cap + k];
        nonbonded_d(c, xi, s->pos + 4u * j, qi, s->charge[j], f);
      }
    } else {
      for (uint32_t j = 0; j < n; ++j)
        if (j != i) nonbonded_d(c, xi, s->pos + 4u * j, qi, s->charge[j], f);
    }
    if (c->es_enabled && qi != 0.0) {
      if (use_list) {
        uint32_t m = s->nh < c->n_heads ? s->nh : c->n_heads;
        for (uint32_t h = 0; h < m; ++h) {
          if (s->head_idx[h] != i) continue;
          for (uint32_t k = 0; k < s->es_count[h]; ++k) es_far_d(c, s, i, s->es_list[(size_t)h * c->es_cap + k], f);
          break;
        }
      } else {
        for (uint32_t j = 0; j < n; ++j)
          if (j != i) es_far_d(c, s, i, j, f);
      }
    }
    out[3u * i] = f[0];
    out[3u * i + 1] = f[1];
    out[3u * i + 2] = f[2];
  }
}

/* FP64 potential energy of the same force field (soup/src/soup-potential.ts:145-301 restricted to
 * the terms the step integrates, with the shifted-force DH of soup/src/electrostatics.ts:277-282).
 * Used to pin the forces as -grad U by central differences. O(N^2). */
double or_energy_f64(const SoupConfig* c, const OrState* s) {
  uint32_t n = c->n;
  double u = 0.0, wf = (double)c->wca_factor;
  double rc_attr = wf * (double)c->b_tt, wc = (double)c->wc, eps = (double)c->epsilon;
  for (uint32_t i = 0; i < n; ++i) {
    const float* xi = s->pos + 4u * i;
    uint32_t ki = (uint32_t)xi[3];
    for (uint32_t j = i + 1; j < n; ++j) {
      const float* xj = s->pos + 4u * j;
      uint32_t kj = (uint32_t)xj[3];
      double dx = mi_d((double)xi[0] - xj[0], c->box[0]);
      double dy = mi_d((double)xi[1] - xj[1], c->box[1]);
      double dz = mi_d((double)xi[2] - xj[2], c->box[2]);
      double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r < 1e-6) continue;
      double b = (double)c->sigma * ((double)c->radius[ki] + (double)c->radius[kj]) * 0.5;
      if (r < wf * b) {
        double s6 = pow(b / r, 6.0), q = 2.0 * s6 - 1.0;
        u += eps * q * q;
      }
      double qi = s->charge[i], qj = s->charge[j];
      double as = (double)c->attr[c->cls[ki]][c->cls[kj]];
      if (c->acid_soap > 0.0f && c->cls[ki] == 1u && c->cls[kj] == 1u && ((qi != 0.0) != (qj != 0.0)))
        as += (double)c->acid_soap;
      if (as > 0.0) {
        if (r < rc_attr) u += -as * eps;
        else if (r <= rc_attr + wc) {
          double cc = cos(M_PI * (r - rc_attr) / (2.0 * wc));
          u += -as * eps * cc * cc;
        }
      }
      double qq = qi * qj;
      if (c->es_enabled && qq != 0.0 && c->es_A != 0.0f && r < (double)c->es_rc) {
        double A = c->es_A, k = c->es_kappa, rc = c->es_rc;
        double u1 = A * exp(-k * r) / r, urc = A * exp(-k * rc) / rc;
        u += qq * (u1 - urc + (double)c->es_Frc * (r - rc));
      }
    }
  }
  /* bonded: FENE per bond once (j > i), bend per 1-3 pair once, tether per link once */
  double kf = c->k_fene, ri = c->r_inf, kb = c->k_bend, rb = c->r_bend;
  for (uint32_t i = 0; i < n; ++i) {
    const float* xi = s->pos + 4u * i;
    for (uint32_t k = 0; k < 4; ++k) {
      uint32_t o = k < 3 ? s->bond_slots[3u * i + k] : s->center_link[i];
      if (o == NONE || o < i) continue;
      double dx = mi_d((double)xi[0] - s->pos[4u * o], c->box[0]);
      double dy = mi_d((double)xi[1] - s->pos[4u * o + 1], c->box[1]);
      double dz = mi_d((double)xi[2] - s->pos[4u * o + 2], c->box[2]);
      double r = sqrt(dx * dx + dy * dy + dz * dz), x = r / ri;
      u += -0.5 * kf * ri * ri * log(1.0 - x * x);
    }
    for (uint32_t k = 0; k < 3; ++k) {
      uint32_t m = s->bond_slots[3u * i + k];
      if (m == NONE) continue;
      for (uint32_t t = 0; t < 3; ++t) {
        uint32_t o = s->bond_slots[3u * m + t];
        if (o == NONE || o == i || o < i) continue;
        double dx = mi_d((double)xi[0] - s->pos[4u * o], c->box[0]);
        double dy = mi_d((double)xi[1] - s->pos[4u * o + 1], c->box[1]);
        double dz = mi_d((double)xi[2] - s->pos[4u * o + 2], c->box[2]);
        double r = sqrt(dx * dx + dy * dy + dz * dz), d = r - rb;
        u += 0.5 * kb * d * d;
      }
    }
  }
  return u;
}

/* ------------------------------------------------------------------------------------------- */
/* soupPotential (soup/src/soup-potential.ts:145-301) with esTotalEnergy (soup/src/electrostatics.ts */
/* :575-600), in FP64 over f32 positions, reference layout. The pair set is every unordered pair  */
/* with 0 < r2 <= cutoff^2 (the reference reaches it through a deduplicated cell list; the set is  */
/* the same, so this walks all pairs). Bonded terms from the bond slots: FENE once per bond, the   */
/* bend once per unordered partner pair of each middle bead, the tether once per linked pair.      */
/* out = {nonbonded, fene, bend, tether, total}.                                                   */
/* ------------------------------------------------------------------------------------------- */
static inline double pot_round(double x) { /* JS Math.round: floor(x + 0.5) semantics */
  double r = floor(x);
  if (x - r >= 0.5) r += 1.0;
  return r;
}
static inline double pot_mi(double d, double L) { return d - pot_round(d / L) * L; } /* electrostatics.ts:408-410 */
static double pot_dist(const double* pos, uint32_t i, uint32_t j, const double* box) {
  double dx = pot_mi(pos[4u * i] - pos[4u * j], box[0]);
  double dy = pot_mi(pos[4u * i + 1] - pos[4u * j + 1], box[1]);
  double dz = pot_mi(pos[4u * i + 2] - pos[4u * j + 2], box[2]);
  return sqrt(dx * dx + dy * dy + dz * dz);
}
static double pot_fene_v(const SoupPotentialBasis* b, double r) { /* soup-potential.ts:42-45 */
  double x = r / b->fene_r_inf;
  return -0.5 * b->fene_k * b->fene_r_inf * b->fene_r_inf * log(1.0 - x * x);
}
static double pot_attr_v(const SoupPotentialBasis* b, double r) { /* soup-potential.ts:52-57 */
  if (r < b->rc_attr) return -b->epsilon;
  if (r > b->rc_attr + b->wc) return 0.0;
  double cc = cos((M_PI * (r - b->rc_attr)) / (2.0 * b->wc));
  return -b->epsilon * cc * cc;
}

/* positions in FP64 (x, y, z, kind per particle); or_soup_potential promotes f32 positions exactly */
void or_soup_potential_d(const SoupPotentialBasis* b, uint32_t n, const double* pos, const uint32_t* bond_slots,
                         const uint32_t* center_link, const float* charges, const double* box, double out[5]) {
  double cut2 = b->cutoff * b->cutoff, nonbonded = 0.0;
  for (uint32_t i = 0; i < n; ++i) { /* soup-potential.ts:199-240 */
    uint32_t ki = (uint32_t)pos[4u * i + 3];
    for (uint32_t j = i + 1; j < n; ++j) {
      double dx = pot_mi(pos[4u * i] - pos[4u * j], box[0]);
      double dy = pot_mi(pos[4u * i + 1] - pos[4u * j + 1], box[1]);
      double dz = pot_mi(pos[4u * i + 2] - pos[4u * j + 2], box[2]);
      double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 > cut2 || r2 <= 0.0) continue;
      double r = sqrt(r2);
      uint32_t kj = (uint32_t)pos[4u * j + 3];
      double bb = (b->b_radius[ki] + b->b_radius[kj]) / 2.0;
      if (!(r >= b->wca_factor * bb)) { /* wcaV, soup-potential.ts:35-41 */
        double s6 = pow(bb / r, 6.0), q = 2.0 * s6 - 1.0;
        nonbonded += b->epsilon * q * q;
      }
      double sc = b->attr[b->cls[ki]][b->cls[kj]];
      if (b->acid_soap > 0.0 && charges && b->use_charges && b->cls[ki] == 1u && b->cls[kj] == 1u &&
          ((charges[i] == 0.0f) != (charges[j] == 0.0f)))
        sc += b->acid_soap;
      if (sc > 0.0) nonbonded += sc * pot_attr_v(b, r);
    }
  }
  double fene = 0.0, bend = 0.0, tether = 0.0;
  for (uint32_t i = 0; i < n; ++i) { /* bonds(): (i, j) with j > i, soup-potential.ts:254-256 */
    for (int k = 0; k < 3; ++k) {
      uint32_t j = bond_slots[3u * i + k];
      if (j != NONE && j > i) fene += pot_fene_v(b, pot_dist(pos, i, j, box));
    }
  }
  for (uint32_t m = 0; m < n; ++m) { /* bend over partner pairs of each middle, :258-267 */
    for (int a = 0; a < 3; ++a) {
      uint32_t pa = bond_slots[3u * m + a];
      if (pa == NONE) continue;
      for (int c2 = a + 1; c2 < 3; ++c2) {
        uint32_t pc = bond_slots[3u * m + c2];
        if (pc == NONE) continue;
        double d = pot_dist(pos, pa, pc, box) - b->bend_r0;
        bend += 0.5 * b->bend_k * d * d;
      }
    }
  }
  if (center_link && b->use_center_link) { /* :269-276 */
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t owner = center_link[i];
      if (owner == NONE || owner <= i) continue;
      tether += pot_fene_v(b, pot_dist(pos, i, owner, box));
    }
  }
  if (charges && b->use_charges && b->es_A != 0.0) { /* esTotalEnergy, electrostatics.ts:575-600 */
    double ecut2 = b->es_cutoff * b->es_cutoff;
    for (uint32_t i = 0; i < n; ++i) {
      if ((uint32_t)pos[4u * i + 3] != b->charged_kind || charges[i] == 0.0f) continue;
      for (uint32_t j = i + 1; j < n; ++j) {
        if ((uint32_t)pos[4u * j + 3] != b->charged_kind || charges[j] == 0.0f) continue;
        double dx = pot_mi(pos[4u * i] - pos[4u * j], box[0]);
        double dy = pot_mi(pos[4u * i + 1] - pos[4u * j + 1], box[1]);
        double dz = pot_mi(pos[4u * i + 2] - pos[4u * j + 2], box[2]);
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 >= ecut2) continue;
        double r = sqrt(r2);
        if (r <= 0.0) continue;
        double u1 = (b->es_A * exp(-b->es_kappa * r)) / r;
        nonbonded += (double)charges[i] * (double)charges[j] * (u1 - b->es_shiftU + b->es_shiftF * (r - b->es_cutoff));
      }
    }
  }
  out[0] = nonbonded;
  out[1] = fene;
  out[2] = bend;
  out[3] = tether;
  out[4] = nonbonded + fene + bend + tether;
}

void or_soup_potential(const SoupPotentialBasis* b, uint32_t n, const float* pos, const uint32_t* bond_slots,
                       const uint32_t* center_link, const float* charges, const double* box, double out[5]) {
  double* pd = (double*)malloc(4u * (size_t)n * sizeof(double));
  for (size_t k = 0; k < 4u * (size_t)n; ++k) pd[k] = (double)pos[k];
  or_soup_potential_d(b, n, pd, bond_slots, center_link, charges, box, out);
  free(pd);
}

/* ------------------------------------------------------------------------------------------- */
/* FP64 gradient, Hessian-vector product and parameter derivatives of soupPotential (the energy  */
/* above, soup/src/soup-potential.ts:145-301), the CPU check of the product's adjoint. Each       */
/* unordered pair term is visited once (list pairs and DH pairs with j > i, each bond from its    */
/* lower end, each 1-3 bend from its middle bead, each mutual tether from its lower end) and     */
/* applied to both of its particles. A pair term phi(r), r_vec = x_i - x_j under the minimum image, */
/* e = r_vec / r, direction dw = w_i - w_j + diag(strain) r_vec, ew = e . dw:                     */
/*   g_i += phi' e, g_j -= phi' e;                                                               */
/*   hw_i += phi'' ew e + (phi' / r)(dw - ew e), hw_j -= the same;                               */
/*   W_e -= phi' (e . diag(strain) r_vec);  d(w . F)/dtheta -= ew dphi'/dtheta.                   */
/* theta: the 10 unordered class-pair depths, acid-soap, epsilon, head charge, kappa.            */
/* Pairs come from the state's own Verlet and DH lists (or_rebuild first); the lists hold every  */
/* pair within the basis cutoffs when those are at most listRange - skin.                        */
/* ------------------------------------------------------------------------------------------- */
typedef struct {
  double* g;
  double* hw;
  double dU[14], dF[14], we;
} AdjOut;
static uint32_t adj_pidx(uint32_t a, uint32_t b) {
  uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  static const uint32_t base[4] = {0u, 4u, 7u, 9u};
  return base[lo] + (hi - lo);
}
/* one pair term (i, j) with derivatives d1, d2 at separation r_vec (length r); the caller adds the  */
/* parameter parts through the returned ew                                                         */
static double adj_pair(AdjOut* o, const double* w, const double* st, uint32_t i, uint32_t j, const double rv[3], double r,
                       double d1, double d2) {
  double e[3], dw[3], ew = 0.0, es = 0.0;
  for (int a = 0; a < 3; ++a) {
    e[a] = rv[a] / r;
    double sa = st[a] * rv[a];
    dw[a] = (w ? w[3u * i + a] - w[3u * j + a] : 0.0) + sa;
  }
  for (int a = 0; a < 3; ++a) {
    ew += e[a] * dw[a];
    es += e[a] * st[a] * rv[a];
  }
  for (int a = 0; a < 3; ++a) {
    double gi = d1 * e[a];
    double hi = d2 * ew * e[a] + (d1 / r) * (dw[a] - ew * e[a]);
    o->g[3u * i + a] += gi;
    o->g[3u * j + a] -= gi;
    o->hw[3u * i + a] += hi;
    o->hw[3u * j + a] -= hi;
  }
  o->we -= d1 * es;
  return ew;
}
static double adj_sep(const float* pos, uint32_t i, uint32_t j, const double* box, double rv[3]) {
  rv[0] = pot_mi((double)pos[4u * i] - (double)pos[4u * j], box[0]);
  rv[1] = pot_mi((double)pos[4u * i + 1] - (double)pos[4u * j + 1], box[1]);
  rv[2] = pot_mi((double)pos[4u * i + 2] - (double)pos[4u * j + 2], box[2]);
  return sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
}
static void adj_fene_pair(AdjOut* o, const SoupPotentialBasis* b, const double* w, const double* st, const float* pos,
                          const double* box, uint32_t i, uint32_t j) {
  double rv[3], r = adj_sep(pos, i, j, box, rv);
  if (!(r > 0.0)) return;
  double x2 = (r / b->fene_r_inf) * (r / b->fene_r_inf);
  adj_pair(o, w, st, i, j, rv, r, b->fene_k * r / (1.0 - x2), b->fene_k * (1.0 + x2) / ((1.0 - x2) * (1.0 - x2)));
}

void or_energy_adjoint(const OrState* s, uint32_t n, uint32_t list_cap, uint32_t es_cap, const SoupPotentialBasis* b,
                       const double* box, double q_head, const double* w, const double* strain, double* g, double* hw,
                       double* dU, double* dF, double* we) {
  const double st0[3] = {0.0, 0.0, 0.0};
  const double* st = strain ? strain : st0;
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  AdjOut* outs = (AdjOut*)calloc((size_t)nt, sizeof(AdjOut));
  for (int t = 0; t < nt; ++t) {
    outs[t].g = (double*)calloc(3u * (size_t)n, sizeof(double));
    outs[t].hw = (double*)calloc(3u * (size_t)n, sizeof(double));
  }
  const double cut2 = b->cutoff * b->cutoff, eps = b->epsilon, z1 = M_PI / (2.0 * b->wc);
  const int use_q = s->charge && b->use_charges;
  const int es = use_q && b->es_A != 0.0;
#pragma omp parallel
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    AdjOut* o = &outs[tid];
#pragma omp for schedule(dynamic, 64)
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t ki = (uint32_t)s->pos[4u * i + 3];
      const double qi = use_q ? (double)s->charge[i] : 0.0;
      for (uint32_t k = 0; k < s->vcount[i]; ++k) { /* list pairs, soup-potential.ts:199-240 */
        uint32_t j = s->vl[(size_t)i * list_cap + k];
        if (j <= i) continue;
        double rv[3], r = adj_sep(s->pos, i, j, box, rv);
        if (r * r > cut2 || !(r > 0.0)) continue;
        const uint32_t kj = (uint32_t)s->pos[4u * j + 3];
        const double bb = (b->b_radius[ki] + b->b_radius[kj]) / 2.0;
        if (r < b->wca_factor * bb) {
          double u6 = pow(bb / r, 6.0), phi = eps * (2.0 * u6 - 1.0) * (2.0 * u6 - 1.0);
          double d1 = -24.0 * eps * u6 * (2.0 * u6 - 1.0) / r, d2 = 24.0 * eps * u6 * (26.0 * u6 - 7.0) / (r * r);
          double ew = adj_pair(o, w, st, i, j, rv, r, d1, d2);
          o->dU[11] += phi / eps;
          o->dF[11] -= ew * d1 / eps;
        }
        const uint32_t ci = b->cls[ki] & 3u, cj = b->cls[kj] & 3u;
        const double qj = use_q ? (double)s->charge[j] : 0.0;
        const int acid = b->acid_soap > 0.0 && use_q && ci == 1u && cj == 1u && ((qi == 0.0) != (qj == 0.0));
        const double sc = b->attr[ci][cj] + (acid ? b->acid_soap : 0.0);
        if (sc > 0.0 && r <= b->rc_attr + b->wc) {
          double a0, a1, a2;
          if (r < b->rc_attr) {
            a0 = -eps; a1 = 0.0; a2 = 0.0;
          } else {
            double z = z1 * (r - b->rc_attr), cz = cos(z);
            a0 = -eps * cz * cz;
            a1 = eps * z1 * sin(2.0 * z);
            a2 = 2.0 * eps * z1 * z1 * cos(2.0 * z);
          }
          double ew = adj_pair(o, w, st, i, j, rv, r, sc * a1, sc * a2);
          uint32_t pi = adj_pidx(ci, cj);
          o->dU[pi] += a0;
          o->dF[pi] -= ew * a1;
          if (acid) {
            o->dU[10] += a0;
            o->dF[10] -= ew * a1;
          }
          o->dU[11] += sc * a0 / eps;
          o->dF[11] -= ew * sc * a1 / eps;
        }
      }
      if (es && qi != 0.0 && ki == b->charged_kind) { /* esTotalEnergy pairs, electrostatics.ts:575-600 */
        uint32_t h = 0, m = s->nh;
        while (h < m && s->head_idx[h] != i) ++h;
        if (h < m) {
          const double rc = b->es_cutoff, kap = b->es_kappa, erc = b->es_A * exp(-kap * rc);
          for (uint32_t k = 0; k < s->es_count[h]; ++k) {
            uint32_t j = s->es_list[(size_t)h * es_cap + k];
            if (j <= i || s->charge[j] == 0.0f) continue;
            double rv[3], r = adj_sep(s->pos, i, j, box, rv);
            if (r >= rc || !(r > 0.0)) continue;
            double qq = qi * (double)s->charge[j];
            double u1 = b->es_A * exp(-kap * r) / r, k1 = kap + 1.0 / r;
            double e0 = u1 - b->es_shiftU + b->es_shiftF * (r - rc);
            double e1 = -u1 * k1 + b->es_shiftF, e2 = u1 * k1 * k1 + u1 / (r * r);
            double ew = adj_pair(o, w, st, i, j, rv, r, qq * e1, qq * e2);
            o->dU[12] += 2.0 * qq * e0 / q_head;
            o->dF[12] -= ew * 2.0 * qq * e1 / q_head;
            o->dU[13] += qq * (-r * u1 + erc - kap * erc * (r - rc));
            o->dF[13] -= ew * qq * (u1 * r * kap - kap * erc);
          }
        }
      }
      for (int k = 0; k < 3; ++k) { /* FENE once per bond, soup-potential.ts:254-256 */
        uint32_t j = s->bond_slots[3u * i + k];
        if (j != NONE && j > i) adj_fene_pair(o, b, w, st, s->pos, box, i, j);
      }
      for (int a = 0; a < 3; ++a) { /* 1-3 bends with i as the middle bead, :258-267 */
        uint32_t pa = s->bond_slots[3u * i + a];
        if (pa == NONE) continue;
        for (int c2 = a + 1; c2 < 3; ++c2) {
          uint32_t pc = s->bond_slots[3u * i + c2];
          if (pc == NONE) continue;
          double rv[3], r = adj_sep(s->pos, pa, pc, box, rv);
          if (r > 0.0) adj_pair(o, w, st, pa, pc, rv, r, b->bend_k * (r - b->bend_r0), b->bend_k);
        }
      }
      if (s->center_link && b->use_center_link) { /* tether once per mutual link, :269-276 */
        uint32_t owner = s->center_link[i];
        if (owner != NONE && owner > i) adj_fene_pair(o, b, w, st, s->pos, box, i, owner);
      }
    }
  }
  memset(g, 0, 3u * (size_t)n * sizeof(double));
  if (hw) memset(hw, 0, 3u * (size_t)n * sizeof(double));
  for (int p = 0; p < 14; ++p) dU[p] = dF[p] = 0.0;
  *we = 0.0;
  for (int t = 0; t < nt; ++t) {
    for (size_t k = 0; k < 3u * (size_t)n; ++k) {
      g[k] += outs[t].g[k];
      if (hw) hw[k] += outs[t].hw[k];
    }
    for (int p = 0; p < 14; ++p) {
      dU[p] += outs[t].dU[p];
      dF[p] += outs[t].dF[p];
    }
    *we += outs[t].we;
    free(outs[t].g);
    free(outs[t].hw);
  }
  free(outs);
}

/* ------------------------------------------------------------------------------------------- */
/* Integrator: soup/wgsl/step.wgsl:89-142, verlet.wgsl:115-123.                                  */
/* ------------------------------------------------------------------------------------------- */
static inline uint32_t pcg_soup(uint32_t v) {
  /* step.wgsl:89-93 */
  uint32_t state = v * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
static inline float uniform01(uint32_t v) { return (float)v / 4294967296.0f; }

uint32_t or_pcg_soup(uint32_t v) { return pcg_soup(v); }

void or_kick_drift(const SoupConfig* c, OrState* s) {
  uint32_t n = c->n;
  float hdt = 0.5f * c->dt;
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < n; ++i) {
    float* v = s->vel + 4u * i;
    float* x = s->pos + 4u * i;
    const float* f = s->force + 4u * i;
    if (s->frozen[i] != 0u) {
      v[0] = v[1] = v[2] = v[3] = 0.0f;
      continue;
    }
    for (int a = 0; a < 3; ++a) {
      float va = v[a] + hdt * f[a];
      v[a] = va;
      float xa = x[a] + c->dt * va;
      xa = xa - floorf(xa / c->box[a]) * c->box[a];
      x[a] = xa;
    }
    v[3] = 0.0f;
  }
}

static inline float bm_log(float u) { return (float)log((double)u); }
static inline float bm_cos(float a) { return (float)cos((double)a); }

void or_kick_thermo(const SoupConfig* c, OrState* s) {
  uint32_t n = c->n;
  float hdt = 0.5f * c->dt;
  float sigma_v = sqrtf(2.0f * c->gamma * c->kT * c->dt);
  float gdt = c->gamma * c->dt;
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < n; ++i) {
    float* v = s->vel + 4u * i;
    if (s->frozen[i] != 0u) {
      v[0] = v[1] = v[2] = v[3] = 0.0f;
      continue;
    }
    const float* f = s->force + 4u * i;
    uint32_t st = s->thermo_rng[i];
    float noise[3];
    for (int a = 0; a < 3; ++a) {
      st = pcg_soup(st);
      float u1 = fmaxf(uniform01(st), 1e-7f);
      st = pcg_soup(st);
      float u2 = uniform01(st);
      noise[a] = sqrtf(-2.0f * bm_log(u1)) * bm_cos(6.28318531f * u2);
    }
    s->thermo_rng[i] = st;
    for (int a = 0; a < 3; ++a) {
      float va = v[a] + hdt * f[a];
      va = va - gdt * va + sigma_v * noise[a];
      v[a] = va;
    }
    v[3] = 0.0f;
  }
}

void or_max_drift(const SoupConfig* c, OrState* s) {
  uint32_t n = c->n, m = s->max_drift_bits;
  for (uint32_t i = 0; i < n; ++i) {
    float dx = mi_f(s->pos[4u * i] - s->pos_at_rebuild[4u * i], c->box[0]);
    float dy = mi_f(s->pos[4u * i + 1] - s->pos_at_rebuild[4u * i + 1], c->box[1]);
    float dz = mi_f(s->pos[4u * i + 2] - s->pos_at_rebuild[4u * i + 2], c->box[2]);
    float d2 = dx * dx + dy * dy + dz * dz;
    uint32_t b;
    memcpy(&b, &d2, 4);
    if (b > m) m = b;
  }
  s->max_drift_bits = m;
}

/* soup_scan_nonfinite_main (soup/wgsl/health.wgsl:52-68). */
void or_nonfinite(const SoupConfig* c, const OrState* s, uint32_t out[2]) {
  uint32_t bp = 0, bv = 0;
  for (uint32_t i = 0; i < c->n; ++i)
    for (int a = 0; a < 3; ++a) {
      uint32_t b;
      memcpy(&b, s->pos + 4u * i + a, 4);
      if (((b >> 23) & 0xFFu) == 0xFFu) bp++;
      memcpy(&b, s->vel + 4u * i + a, 4);
      if (((b >> 23) & 0xFFu) == 0xFFu) bv++;
    }
  out[0] = bp;
  out[1] = bv;
}

/* soup_relax_step_main (soup/wgsl/relax.wgsl:49-72). */
void or_relax_step(const SoupConfig* c, OrState* s, float d) {
  for (uint32_t i = 0; i < c->n; ++i) {
    if (s->frozen[i] != 0u) continue;
    const float* f = s->force + 4u * i;
    float a = fmaxf(fmaxf(fabsf(f[0]), fabsf(f[1])), fabsf(f[2]));
    uint32_t b;
    memcpy(&b, &a, 4);
    if (!(a > 0.0f) || ((b >> 23) & 0xFFu) == 0xFFu) continue;
    float g[3] = {f[0] / a, f[1] / a, f[2] / a};
    float l = len3_f(g[0], g[1], g[2]);
    for (int k = 0; k < 3; ++k) {
      float x = s->pos[4u * i + k] + (g[k] / l) * d;
      x = x - floorf(x / c->box[k]) * c->box[k];
      s->pos[4u * i + k] = x;
    }
  }
}

/* ------------------------------------------------------------------------------------------- */
/* Bond Monte Carlo: soup/wgsl/bond-*.wgsl, sequential canonical order.                           */
/* ------------------------------------------------------------------------------------------- */
static inline float kind_of(const OrState* s, uint32_t i) { return s->pos[4u * i + 3]; }

static inline float dist_f(const SoupConfig* c, const OrState* s, uint32_t a, uint32_t b) {
  float dx = mi_f(s->pos[4u * a] - s->pos[4u * b], c->box[0]);
  float dy = mi_f(s->pos[4u * a + 1] - s->pos[4u * b + 1], c->box[1]);
  float dz = mi_f(s->pos[4u * a + 2] - s->pos[4u * b + 2], c->box[2]);
  return len3_f(dx, dy, dz);
}

static int match_rule(const SoupConfig* c, float ti, float tj) {
  /* bond-common.wgsl:142-149 */
  for (uint32_t r = 0; r < BOND_RULES; ++r) {
    float a = c->kindA[r], b = c->kindB[r];
    if ((ti == a && tj == b) || (ti == b && tj == a)) return (int)r;
  }
  return -1;
}
static uint32_t role_of(const SoupConfig* c, uint32_t r, float kind) {
  return kind == c->kindA[r] ? (uint32_t)c->roleA[r] : (uint32_t)c->roleB[r];
}
static int has_bond_to(const OrState* s, uint32_t i, uint32_t j) {
  for (uint32_t k = 0; k < 3; ++k)
    if (s->bond_slots[3u * i + k] == j) return 1;
  return 0;
}
static inline int cas(uint32_t* p, uint32_t expected, uint32_t v) {
  if (*p != expected) return 0;
  *p = v;
  return 1;
}

/* tryClaimSlot (bond-valence.wgsl:208-266). */
static int try_claim(const SoupConfig* c, OrState* s, uint32_t particle, uint32_t role, uint32_t partner) {
  uint32_t* sl = s->bond_slots + 3u * particle;
  int hto = c->head_terminal_only > 0.5f;
  if (role == 0u) {
    if (cas(&sl[0], NONE, partner)) {
      if (hto && sl[2] != NONE) {
        sl[0] = NONE;
        return -1;
      }
      return 0;
    }
    if (cas(&sl[1], NONE, partner)) {
      if (hto && sl[2] != NONE) {
        sl[1] = NONE;
        return -1;
      }
      return 1;
    }
    return -1;
  } else if (role == 1u) {
    if (cas(&sl[2], NONE, partner)) {
      if (hto && sl[0] != NONE && sl[1] != NONE) {
        sl[2] = NONE;
        return -1;
      }
      return 2;
    }
    return -1;
  } else {
    uint32_t cap = (uint32_t)c->head_chain_cap;
    for (uint32_t k = 0; k < cap; ++k)
      if (cas(&sl[k], NONE, partner)) return (int)k;
    return -1;
  }
}
static void release_slot(OrState* s, uint32_t particle, int slot) {
  if (slot < 0) return;
  s->bond_slots[3u * particle + (uint32_t)slot] = NONE;
}
static void release_partner_slot(OrState* s, uint32_t partner, uint32_t me) {
  for (uint32_t k = 0; k < 3; ++k)
    if (cas(&s->bond_slots[3u * partner + k], me, NONE)) return;
}

/* bond-adsorption.wgsl:62-316 */
static int center_within_reach(const SoupConfig* c, const OrState* s, uint32_t a, uint32_t b) {
  return dist_f(c, s, a, b) < wca_cut_f(c, pair_b_f(c, kind_of(s, a), c->catalyst_kind)) + c->wc;
}
static int contact_catalyst(const SoupConfig* c, const OrState* s, uint32_t a, uint32_t cat) {
  return dist_f(c, s, a, cat) < wca_cut_f(c, pair_b_f(c, kind_of(s, a), c->catalyst_kind));
}

static int propagate_on_center(const SoupConfig* c, OrState* s, uint32_t i, uint32_t j, int slotI, int slotJ,
                               uint32_t catalystId) {
  uint32_t* cl = s->center_link;
  if (slotI == 0 && slotJ == 0) {
    if (catalystId == NONE) return 0;
    if (!contact_catalyst(c, s, i, catalystId)) return 0;
    if (!cas(&cl[catalystId], NONE, i)) return 0;
    if (!cas(&cl[i], NONE, catalystId)) {
      cas(&cl[catalystId], i, NONE);
      return 0;
    }
    s->center_held[catalystId] = 0u;
    return 1;
  }
  uint32_t tipOld, tipNew;
  if (slotI == 1 && slotJ == 0) {
    tipOld = i;
    tipNew = j;
  } else if (slotI == 0 && slotJ == 1) {
    tipOld = j;
    tipNew = i;
  } else {
    return 0;
  }
  uint32_t owner = cl[tipOld];
  int reclaiming = 0;
  if (owner == NONE) {
    if (catalystId == NONE) return 0;
    owner = catalystId;
    reclaiming = 1;
  }
  if (reclaiming) {
    if (!contact_catalyst(c, s, tipOld, catalystId)) return 0;
  } else {
    if (!center_within_reach(c, s, tipOld, owner)) return 0;
  }
  if (!contact_catalyst(c, s, tipNew, owner)) return 0;
  if (reclaiming) {
    if (!cas(&cl[owner], NONE, tipOld)) return 0;
    if (!cas(&cl[tipOld], NONE, owner)) {
      cas(&cl[owner], tipOld, NONE);
      return 0;
    }
  }
  if (!cas(&cl[owner], tipOld, tipNew)) {
    if (reclaiming) cas(&cl[tipOld], owner, NONE);
    return 0;
  }
  if (!cas(&cl[tipNew], NONE, owner)) {
    cas(&cl[owner], tipNew, tipOld);
    if (reclaiming) {
      cas(&cl[owner], tipOld, NONE);
      cas(&cl[tipOld], owner, NONE);
    }
    return 0;
  }
  cas(&cl[tipOld], owner, NONE);
  s->center_held[owner] = 0u;
  return 1;
}

static int terminate_on_center(const SoupConfig* c, OrState* s, uint32_t carbon) {
  uint32_t owner = s->center_link[carbon];
  if (owner == NONE) return 0;
  if (!center_within_reach(c, s, carbon, owner)) return 0;
  if (!cas(&s->center_link[carbon], owner, NONE)) return 0;
  cas(&s->center_link[owner], carbon, NONE);
  s->center_held[owner] = 0u;
  return 1;
}

static void desorb_stretch(const SoupConfig* c, OrState* s, uint32_t carbon, uint32_t owner) {
  if (dist_f(c, s, carbon, owner) >= c->r_inf) {
    if (cas(&s->center_link[carbon], owner, NONE)) {
      cas(&s->center_link[owner], carbon, NONE);
      s->center_held[owner] = 0u;
      s->desorb[0]++;
    }
  }
}

static void desorb_timeout(const SoupConfig* c, OrState* s, uint32_t cat) {
  uint32_t tip = s->center_link[cat];
  if (tip == NONE) return;
  uint32_t inc = (uint32_t)c->bond_interval_f;
  uint32_t prev = s->center_held[cat];
  s->center_held[cat] = prev + inc;
  if (prev + inc > (uint32_t)c->max_hold_steps) {
    if (cas(&s->center_link[cat], tip, NONE)) {
      cas(&s->center_link[tip], cat, NONE);
      s->center_held[cat] = 0u;
      s->desorb[1]++;
    }
  }
}

/* Draw counters of the bond Monte Carlo, per rule: [0] formation candidates that reach the attempt
 * draw, [1] those whose attempt draw passes, [2] those whose accept draw passes too, [3] bonds that
 * reach the break attempt draw, [4] break attempts that pass, [5] break accept draws that pass. Test
 * instrumentation of the Poisson thinning (soup/src/soup-buffers.ts:345-356): a pass fraction must
 * equal the uploaded probability within binomial error. */
static uint64_t g_bond_draws[SOUP_MAX_RULES][6];
void or_bond_draws(uint64_t* out) { memcpy(out, g_bond_draws, sizeof(g_bond_draws)); }
void or_bond_draws_reset(void) { memset(g_bond_draws, 0, sizeof(g_bond_draws)); }

/* bond_form_list_main = bondFormWalkList (bond-dispatch.wgsl:212-242) + bondFormDecide (:89-142),
 * for i = 0..N-1 in ascending order. */
void or_bond_form(const SoupConfig* c, OrState* s) {
  uint32_t n = c->n;
  for (uint32_t i = 0; i < n; ++i) {
    float ti = kind_of(s, i);
    uint32_t catalystNear = 0, catalystId = NONE, nCand = 0;
    uint32_t candJ[CANDIDATE_CAP], candRule[CANDIDATE_CAP];
    float rcat = wca_cut_f(c, pair_b_f(c, ti, c->catalyst_kind));
    for (uint32_t k = 0; k < s->vcount[i]; ++k) {
      uint32_t j = s->vl[(size_t)i * c->list_cap + k];
      float tj = kind_of(s, j);
      if (catalystNear == 0u && tj == c->catalyst_kind) {
        if (dist_f(c, s, i, j) < rcat) {
          catalystNear = 1u;
          catalystId = j;
        }
      }
      if (j > i && nCand < CANDIDATE_CAP) {
        int r = match_rule(c, ti, tj);
        if (r >= 0 && !has_bond_to(s, i, j)) {
          if (dist_f(c, s, i, j) < wca_cut_f(c, pair_b_f(c, ti, tj))) {
            candJ[nCand] = j;
            candRule[nCand] = (uint32_t)r;
            nCand++;
          }
        }
      }
    }
    /* bondFormDecide */
    uint32_t rng = s->bond_rng[i];
    if (ti == c->catalyst_kind) {
      desorb_timeout(c, s, i);
    } else {
      uint32_t owner = s->center_link[i];
      if (owner != NONE) desorb_stretch(c, s, i, owner);
    }
    for (uint32_t ci = 0; ci < nCand; ++ci) {
      uint32_t j = candJ[ci], r = candRule[ci];
      float tj = kind_of(s, j);
      if (c->req_cat[r] > 0.5f && catalystNear == 0u) continue;
      g_bond_draws[r][0]++;
      rng = pcg_soup(rng);
      if (uniform01(rng) >= c->att_form[r]) continue;
      g_bond_draws[r][1]++;
      rng = pcg_soup(rng);
      if (uniform01(rng) >= c->acc_form[r]) continue;
      g_bond_draws[r][2]++;
      uint32_t roleI = role_of(c, r, ti), roleJ = role_of(c, r, tj);
      int slotI = try_claim(c, s, i, roleI, j);
      if (slotI < 0) continue;
      int slotJ = try_claim(c, s, j, roleJ, i);
      if (slotJ < 0) {
        release_slot(s, i, slotI);
        continue;
      }
      int ok;
      if (roleI == 0u && roleJ == 0u) ok = propagate_on_center(c, s, i, j, slotI, slotJ, catalystId);
      else if (roleI == 1u) ok = terminate_on_center(c, s, i);
      else ok = terminate_on_center(c, s, j);
      if (!ok) {
        release_slot(s, i, slotI);
        release_slot(s, j, slotJ);
        continue;
      }
      s->events[r * 2u]++;
    }
    s->bond_rng[i] = rng;
  }
}

/* bond_break_main (bond-dispatch.wgsl:170-200), ascending i. */
void or_bond_break(const SoupConfig* c, OrState* s) {
  for (uint32_t i = 0; i < c->n; ++i) {
    float ti = kind_of(s, i);
    uint32_t rng = s->bond_rng[i];
    for (uint32_t k = 0; k < 3; ++k) {
      uint32_t p = s->bond_slots[3u * i + k];
      if (p == NONE || p < i) continue;
      int r = match_rule(c, ti, kind_of(s, p));
      if (r < 0) continue;
      g_bond_draws[r][3]++;
      rng = pcg_soup(rng);
      if (uniform01(rng) >= c->att_break[r]) continue;
      g_bond_draws[r][4]++;
      rng = pcg_soup(rng);
      if (uniform01(rng) >= c->acc_break[r]) continue;
      g_bond_draws[r][5]++;
      s->bond_slots[3u * i + k] = NONE;
      release_partner_slot(s, p, i);
      s->events[(uint32_t)r * 2u + 1u]++;
    }
    s->bond_rng[i] = rng;
  }
}

/* ------------------------------------------------------------------------------------------- */
/* Constant-pH Monte Carlo: soup/src/electrostatics.ts:308-496, soup/src/acid-soap.ts:147-198.   */
/* ------------------------------------------------------------------------------------------- */
static inline double js_round(double x) {
  /* Math.round: nearest integer, ties toward +infinity */
  double r = floor(x);
  if (x - r >= 0.5) r += 1.0;
  return r;
}
static inline double js_mod(double a, double b) { return fmod(a, b); }

double or_pcg_next(uint32_t* state) {
  /* electrostatics.ts:308-315 */
  uint32_t v = *state * 747796405u + 2891336453u;
  *state = v;
  v = ((v >> ((v >> 28) + 4u)) ^ v) * 277803737u;
  v = (v >> 22) ^ v;
  return (double)v / 4294967296.0;
}

static double es_pair_energy(const SoupConfig* c, double r, double qq) {
  /* electrostatics.ts:277-281 */
  if (c->ph_A == 0.0 || qq == 0.0 || r >= c->ph_cutoff || r <= 0.0) return 0.0;
  double u1 = (c->ph_A * exp(-c->ph_kappa * r)) / r;
  return qq * (u1 - c->ph_shiftU + c->ph_shiftF * (r - c->ph_cutoff));
}
static double acid_soap_well(const SoupConfig* c, double r) {
  /* acid-soap.ts:147-153 */
  if (c->as_scale <= 0.0) return 0.0;
  if (r < c->as_rc) return -c->as_scale * c->as_eps;
  if (r > c->as_rc + c->as_wc) return 0.0;
  double cc = cos((M_PI * (r - c->as_rc)) / (2.0 * c->as_wc));
  return -c->as_scale * c->as_eps * cc * cc;
}

typedef struct {
  uint32_t nx, ny, nz;
  double wx, wy, wz;
  uint32_t* start; /* ncell + 1 */
  uint32_t* items; /* heads sorted by cell, stable */
} HeadCells;

static uint32_t hc_key(const HeadCells* hc, const double* box, double x, double y, double z) {
  /* electrostatics.ts:336-341 */
  double X = js_mod(js_mod(x, box[0]) + box[0], box[0]);
  double Y = js_mod(js_mod(y, box[1]) + box[1], box[1]);
  double Z = js_mod(js_mod(z, box[2]) + box[2], box[2]);
  double fx = floor(X / hc->wx), fy = floor(Y / hc->wy), fz = floor(Z / hc->wz);
  uint32_t cx = (uint32_t)fmin((double)(hc->nx - 1), fx);
  uint32_t cy = (uint32_t)fmin((double)(hc->ny - 1), fy);
  uint32_t cz = (uint32_t)fmin((double)(hc->nz - 1), fz);
  return cx + hc->nx * (cy + hc->ny * cz);
}

/* protonationSweep (electrostatics.ts:435-496). heads: ascending particle indices of the charged
 * kind; charges mutated; rng advanced. */
void or_protonation_sweep(const SoupConfig* c, const float* pos, float* charges, uint32_t n, uint32_t* rng,
                          SoupSweepResult* out) {
  double box[3] = {c->box[0], c->box[1], c->box[2]};
  uint32_t nh = 0;
  uint32_t* heads = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; ++i)
    if ((int32_t)pos[4u * i + 3] == (int32_t)c->charged_kind) heads[nh++] = i;
  HeadCells hc;
  double cut = c->ph_cutoff;
  hc.nx = (uint32_t)fmax(1.0, floor(box[0] / cut));
  hc.ny = (uint32_t)fmax(1.0, floor(box[1] / cut));
  hc.nz = (uint32_t)fmax(1.0, floor(box[2] / cut));
  hc.wx = box[0] / hc.nx;
  hc.wy = box[1] / hc.ny;
  hc.wz = box[2] / hc.nz;
  uint32_t ncell = hc.nx * hc.ny * hc.nz;
  hc.start = (uint32_t*)calloc(ncell + 1, sizeof(uint32_t));
  hc.items = (uint32_t*)malloc(sizeof(uint32_t) * (nh ? nh : 1));
  uint32_t* key = (uint32_t*)malloc(sizeof(uint32_t) * (nh ? nh : 1));
  for (uint32_t h = 0; h < nh; ++h) {
    uint32_t i = heads[h];
    key[h] = hc_key(&hc, box, pos[4u * i], pos[4u * i + 1], pos[4u * i + 2]);
    hc.start[key[h] + 1]++;
  }
  for (uint32_t k = 0; k < ncell; ++k) hc.start[k + 1] += hc.start[k];
  uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * (ncell ? ncell : 1));
  memcpy(cur, hc.start, sizeof(uint32_t) * ncell);
  for (uint32_t h = 0; h < nh; ++h) hc.items[cur[key[h]]++] = heads[h];

  double kT = c->ph_kT;
  double dGintr = kT * M_LN10 * (c->ph_pKa - c->ph_pH);
  double q = c->ph_q;
  uint32_t attempts = 0, accepted = 0, dEsCount = 0;
  double dEsSum = 0.0, cut2 = cut * cut;
  uint32_t* nbuf = (uint32_t*)malloc(sizeof(uint32_t) * (nh ? nh : 1));
  uint8_t* seen = (uint8_t*)calloc(ncell ? ncell : 1, 1);
  double reach = c->as_rc + c->as_wc, reach2 = reach * reach;
  for (uint32_t a = 0; a < nh; ++a) {
    double u = or_pcg_next(rng);
    double hf = floor(u * (double)nh);
    uint32_t hsel = (uint32_t)fmin((double)(nh - 1), hf);
    uint32_t i = heads[hsel];
    attempts++;
    /* neighboursInto (electrostatics.ts:355-379) */
    double X = js_mod(js_mod((double)pos[4u * i], box[0]) + box[0], box[0]);
    double Y = js_mod(js_mod((double)pos[4u * i + 1], box[1]) + box[1], box[1]);
    double Z = js_mod(js_mod((double)pos[4u * i + 2], box[2]) + box[2], box[2]);
    int cx = (int)fmin((double)(hc.nx - 1), floor(X / hc.wx));
    int cy = (int)fmin((double)(hc.ny - 1), floor(Y / hc.wy));
    int cz = (int)fmin((double)(hc.nz - 1), floor(Z / hc.wz));
    uint32_t m = 0;
    uint32_t visited[27];
    uint32_t nvis = 0;
    for (int dz = -1; dz <= 1; ++dz) {
      int az = ((cz + dz) % (int)hc.nz + (int)hc.nz) % (int)hc.nz;
      for (int dy = -1; dy <= 1; ++dy) {
        int ay = ((cy + dy) % (int)hc.ny + (int)hc.ny) % (int)hc.ny;
        for (int dx = -1; dx <= 1; ++dx) {
          int ax = ((cx + dx) % (int)hc.nx + (int)hc.nx) % (int)hc.nx;
          uint32_t k = (uint32_t)ax + hc.nx * ((uint32_t)ay + hc.ny * (uint32_t)az);
          if (seen[k]) continue;
          seen[k] = 1;
          visited[nvis++] = k;
          for (uint32_t t = hc.start[k]; t < hc.start[k + 1]; ++t) nbuf[m++] = hc.items[t];
        }
      }
    }
    for (uint32_t t = 0; t < nvis; ++t) seen[visited[t]] = 0;
    double uEs = 0.0;
    for (uint32_t t = 0; t < m; ++t) {
      uint32_t j = nbuf[t];
      if (j == i || charges[j] == 0.0f) continue;
      double dx = (double)pos[4u * i] - (double)pos[4u * j];
      double dy = (double)pos[4u * i + 1] - (double)pos[4u * j + 1];
      double dz = (double)pos[4u * i + 2] - (double)pos[4u * j + 2];
      dx = dx - js_round(dx / box[0]) * box[0];
      dy = dy - js_round(dy / box[1]) * box[1];
      dz = dz - js_round(dz / box[2]) * box[2];
      double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 >= cut2) continue;
      uEs += es_pair_energy(c, sqrt(r2), q * (double)charges[j]);
    }
    /* acidSoapSiteWork (acid-soap.ts:170-198) */
    if (c->as_scale > 0.0) {
      double w = 0.0;
      for (uint32_t t = 0; t < m; ++t) {
        uint32_t j = nbuf[t];
        if (j == i) continue;
        double dx = (double)pos[4u * i] - (double)pos[4u * j];
        double dy = (double)pos[4u * i + 1] - (double)pos[4u * j + 1];
        double dz = (double)pos[4u * i + 2] - (double)pos[4u * j + 2];
        dx -= js_round(dx / box[0]) * box[0];
        dy -= js_round(dy / box[1]) * box[1];
        dz -= js_round(dz / box[2]) * box[2];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 >= reach2 || r2 <= 0.0) continue;
        double v = acid_soap_well(c, sqrt(r2));
        w += charges[j] == 0.0f ? v : -v;
      }
      uEs += w;
    }
    int isCharged = charges[i] != 0.0f;
    double dG = isCharged ? -(dGintr + uEs) : dGintr + uEs;
    if (dG <= 0.0 || or_pcg_next(rng) < exp(-dG / kT)) {
      charges[i] = isCharged ? 0.0f : (float)q;
      accepted++;
      if (!isCharged) {
        dEsSum += uEs / kT;
        dEsCount++;
      }
    }
  }
  uint32_t dep = 0;
  for (uint32_t h = 0; h < nh; ++h)
    if (charges[heads[h]] != 0.0f) dep++;
  out->attempts = attempts;
  out->accepted = accepted;
  out->heads = nh;
  out->deprotonated = dep;
  out->alpha = nh > 0 ? (double)dep / (double)nh : 0.0;
  out->dEsMeanKT = dEsCount > 0 ? dEsSum / (double)dEsCount : 0.0;
  free(heads);
  free(hc.start);
  free(hc.items);
  free(key);
  free(cur);
  free(nbuf);
  free(seen);
}

/* initialCharges (soup/src/soup-protonation.ts:38-49). */
uint32_t or_initial_charges(const SoupConfig* c, const float* pos, uint32_t n, uint32_t seed, float* charges) {
  uint32_t rng = seed ^ 0x9e3779b9u;
  memset(charges, 0, sizeof(float) * n);
  if (!c->es_enabled) return rng;
  double alpha = 1.0 / (1.0 + pow(10.0, c->ph_pKa - c->ph_pH));
  for (uint32_t i = 0; i < n; ++i)
    if ((int32_t)pos[4u * i + 3] == (int32_t)c->charged_kind)
      if (or_pcg_next(&rng) < alpha) charges[i] = (float)c->ph_q;
  return rng;
}

/* ------------------------------------------------------------------------------------------- */
/* Initial lattice: soup/src/soup-init-state.ts:17-31, :112-210, :257-272 (no clay, no parcel).  */
/* ------------------------------------------------------------------------------------------- */
typedef struct { uint32_t a; } Mulberry;
static double mulberry_next(Mulberry* m) {
  m->a = m->a + 0x6d2b79f5u;
  uint32_t a = m->a;
  uint32_t t = (a ^ (a >> 15)) * (1u | a);
  t = (t + ((t ^ (t >> 7)) * (61u | t))) ^ t;
  return (double)((t ^ (t >> 14))) / 4294967296.0;
}
static double gaussian(Mulberry* m) {
  double u1 = fmax(mulberry_next(m), 1e-9);
  double u2 = mulberry_next(m);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* counts: per-kind counts in kind order; writes pos/vel (4n), bond_slots (3n), bond_rng, thermo_rng. */
void or_lattice(const double box[3], const uint32_t* counts, uint32_t n_kinds, uint32_t seed, double kT, float* pos,
                float* vel, uint32_t* bond_slots, uint32_t* center_link, uint32_t* bond_rng, uint32_t* thermo_rng) {
  uint32_t n = 0;
  for (uint32_t k = 0; k < n_kinds; ++k) n += counts[k];
  Mulberry m = {seed};
  /* Math.max(1, Math.ceil(Math.cbrt(N))), then up while nx^3 < N (soup-init-state.ts:117-118): V8's
   * cbrt is exact on perfect cubes and never crosses an integer elsewhere, so this is the smallest nx
   * with nx^3 >= N. The C library's cbrt is not exact on perfect cubes (cbrt(27) = 3 + 1 ulp), so the
   * smallest such nx is searched in integers instead. */
  uint32_t nx = 1;
  while ((uint64_t)nx * nx * nx < n) nx++;
  double sp[3] = {box[0] / nx, box[1] / nx, box[2] / nx};
  uint32_t nsites = nx * nx * nx;
  uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * nsites);
  for (uint32_t i = 0; i < nsites; ++i) order[i] = i;
  for (uint32_t i = nsites - 1; i > 0; --i) {
    uint32_t j = (uint32_t)floor(mulberry_next(&m) * (double)(i + 1));
    uint32_t t = order[i];
    order[i] = order[j];
    order[j] = t;
  }
  uint32_t idx = 0;
  double jf = 0.15, sk = sqrt(kT);
  for (uint32_t k = 0; k < n_kinds; ++k) {
    for (uint32_t c = 0; c < counts[k]; ++c) {
      uint32_t site = order[idx];
      uint32_t ix = site % nx, iy = (site / nx) % nx, iz = site / (nx * nx);
      double bx = (ix + 0.5) * sp[0], by = (iy + 0.5) * sp[1], bz = (iz + 0.5) * sp[2];
      double x = bx + (mulberry_next(&m) * 2.0 - 1.0) * jf * sp[0];
      double y = by + (mulberry_next(&m) * 2.0 - 1.0) * jf * sp[1];
      double z = bz + (mulberry_next(&m) * 2.0 - 1.0) * jf * sp[2];
      pos[4u * idx] = (float)x;
      pos[4u * idx + 1] = (float)y;
      pos[4u * idx + 2] = (float)z;
      pos[4u * idx + 3] = (float)k;
      double g0 = gaussian(&m), g1 = gaussian(&m), g2 = gaussian(&m);
      vel[4u * idx] = (float)(sk * g0);
      vel[4u * idx + 1] = (float)(sk * g1);
      vel[4u * idx + 2] = (float)(sk * g2);
      vel[4u * idx + 3] = 0.0f;
      idx++;
    }
  }
  free(order);
  for (uint32_t i = 0; i < 3u * n; ++i) bond_slots[i] = NONE;
  for (uint32_t i = 0; i < n; ++i) {
    center_link[i] = NONE;
    bond_rng[i] = seed ^ ((i + 1u) * 2654435761u) ^ 0x9e3779b9u;
    thermo_rng[i] = seed ^ ((i + 1u) * 3266489917u) ^ 0x85ebca6bu;
  }
}

/* ------------------------------------------------------------------------------------------- */
/* Step orchestration: soup/src/soup-integrate.ts:83-175.                                          */
/* ------------------------------------------------------------------------------------------- */
enum { OR_OK = 0, OR_NONFINITE = -3, OR_OVERFLOW = -4, OR_DRIFT = -5 };

/* forceCandidateStatsDEBUG (soup/src/soup-readback.ts:246-263): the grid rebuilt at the current
 * positions, then soup_force_stats_main (soup/wgsl/step.wgsl:307-343): every particle's +-walkRadius
 * cell walk counts each candidate j != i, and each pair inside its WCA cut or inside the attraction
 * shell [rcAttr, rcAttr + wc] of an attracting class pair. out[0] candidates, out[1] pairs. */
void or_force_candidate_stats(const SoupConfig* c, OrState* s, uint64_t out[2]) {
  or_grid_build(c, s);
  const int R = (int)c->walk_radius;
  const float rc_attr = wca_cut_f(c, c->b_tt);
  uint64_t cand = 0, pairs = 0;
  for (uint32_t i = 0; i < c->n; ++i) {
    const float* xi = s->pos + 4u * i;
    const float ti = xi[3];
    int cc[3];
    cell_coord(c, xi, cc);
    for (int dz = -R; dz <= R; ++dz) {
      const int cz = wrap_axis(cc[2] + dz, (int)c->dims[2]);
      for (int dy = -R; dy <= R; ++dy) {
        const int cy = wrap_axis(cc[1] + dy, (int)c->dims[1]);
        for (int dx = -R; dx <= R; ++dx) {
          const int cx = wrap_axis(cc[0] + dx, (int)c->dims[0]);
          const uint32_t nc = (uint32_t)cx + c->dims[0] * ((uint32_t)cy + c->dims[1] * (uint32_t)cz);
          for (uint32_t k = s->cell_start[nc]; k < s->cell_start[nc + 1]; ++k) {
            const uint32_t j = s->cell_idx[k];
            if (j == i) continue;
            ++cand;
            const float* xj = s->pos + 4u * j;
            const float tj = xj[3];
            const float r = len3_f(mi_f(xi[0] - xj[0], c->box[0]), mi_f(xi[1] - xj[1], c->box[1]), mi_f(xi[2] - xj[2], c->box[2]));
            const float b = pair_b_f(c, ti, tj);
            const int within_wca = r < wca_cut_f(c, b);
            const int attract = c->attr[c->cls[(uint32_t)ti]][c->cls[(uint32_t)tj]] > 0.0f;
            const int within_attr = attract && r >= rc_attr && r <= rc_attr + c->wc;
            if (within_wca || within_attr) ++pairs;
          }
        }
      }
    }
  }
  out[0] = cand;
  out[1] = pairs;
}

int or_check_lists(const SoupConfig* c, const OrState* s) {
  if (s->v_overflow) return OR_OVERFLOW;
  if (c->es_enabled && c->n_heads > 0 && s->es_overflow) return OR_OVERFLOW;
  float d2;
  memcpy(&d2, &s->max_drift_bits, 4);
  double drift = sqrt(fmax(0.0, (double)d2));
  if (drift != drift) return OR_DRIFT;
  if (drift > (double)c->skin / 2.0 + 1e-6) return OR_DRIFT;
  return OR_OK;
}

void or_one_step(const SoupConfig* c, OrState* s) {
  int doBonds = (s->global_step % c->bond_interval) == 0;
  int doRebuild = (s->global_step % c->rebuild_every) == 0;
  or_kick_drift(c, s);
  if (doRebuild) or_rebuild(c, s);
  or_force_list(c, s);
  or_max_drift(c, s);
  if (doBonds) {
    or_bond_form(c, s);
    or_bond_break(c, s);
  }
  or_kick_thermo(c, s);
  s->global_step++;
}

int or_step(const SoupConfig* c, OrState* s, uint32_t nsteps, SoupSweepResult* last_sweep, uint32_t* sweeps) {
  uint32_t done = 0;
  while (done < nsteps) {
    uint32_t chunk = nsteps - done < 1000u ? nsteps - done : 1000u;
    for (uint32_t k = 0; k < chunk; ++k) or_one_step(c, s);
    uint32_t nf[2];
    or_nonfinite(c, s, nf);
    if (nf[0] || nf[1]) return OR_NONFINITE;
    int e = or_check_lists(c, s);
    if (e) return e;
    if (c->es_enabled && s->global_step >= s->next_sweep_at) {
      s->next_sweep_at = s->global_step + c->sweep_every;
      or_protonation_sweep(c, s->pos, s->charge, c->n, &s->prng, last_sweep);
      if (sweeps) (*sweeps)++;
      or_rebuild(c, s);
      or_force_list(c, s);
    }
    done += chunk;
  }
  return OR_OK;
}

int or_relax(const SoupConfig* c, OrState* s, uint32_t iterations, double d0) {
  for (uint32_t k = 0; k < iterations; ++k) {
    float d = (float)(d0 * (1.0 - (double)k / (double)iterations));
    or_rebuild(c, s);
    or_force_list(c, s);
    or_relax_step(c, s, d);
  }
  or_rebuild(c, s);
  or_force_list(c, s);
  uint32_t nf[2];
  or_nonfinite(c, s, nf);
  if (nf[0] || nf[1]) return OR_NONFINITE;
  return or_check_lists(c, s);
}

int or_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void or_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}
