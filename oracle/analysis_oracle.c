/*
 * analysis_oracle.c -- CPU restatement of the soup snapshot analysis: amphiphile recognition,
 * member clustering, the periodic encapsulated-water flood, the open-boundary cavity flood, the
 * gyration-tensor shape and the head-density peak count.
 *
 * TEST INFRASTRUCTURE ONLY (linked into liboracle.so with soup_oracle.c): called by tests/ as the
 * checker of the GPU analysis. Each function follows the reference file:line it names, in FP64 over
 * the reference's f32 inputs, with JavaScript Math.round / Math.floor semantics where the reference
 * uses them. Breadth-first floods are restated as breadth-first floods (the GPU uses union-find;
 * the visited sets are the same by definition).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define A_NONE 0xFFFFFFFFu

static inline double js_round_a(double x) { /* Math.round: round half toward +infinity */
  double r = floor(x);
  if (x - r >= 0.5) r += 1.0;
  return r;
}
static inline double mi1(double d, double box) { return d - js_round_a(d / box) * box; } /* aggregate.ts:35-38 */
static inline double wrap1(double v, double box) { return v - floor(v / box) * box; }   /* aggregate.ts:41-43 */
static inline long imod(long a, long n) { return ((a % n) + n) % n; }

/* ------------------------------------------------------------------------------------------- */
/* findAmphiphiles (soup/src/amphiphile.ts:55-141). bonds = flat (i, j) pairs in bonds() order.   */
/* Outputs, in the reference's result order (components by first particle index): head[a],       */
/* chain_off[a..a+1] into chain[], tail_len[2a], tail_len[2a+1] (0 when single-tailed).           */
/* Returns the number of amphiphiles.                                                             */
/* ------------------------------------------------------------------------------------------- */
static uint32_t uf_find_halving(int32_t* parent, uint32_t x) {
  while ((uint32_t)parent[x] != x) {
    parent[x] = parent[parent[x]];
    x = (uint32_t)parent[x];
  }
  return x;
}

uint32_t or_find_amphiphiles(uint32_t n, const float* pos, const uint32_t* bonds, uint32_t npairs,
                             const uint8_t* polar_by_kind, uint32_t nkinds, uint32_t* head, uint32_t* chain_off,
                             uint32_t* chain, uint32_t* tail_len) {
  /* adjacency in bond order (amphiphile.ts:66-73) */
  uint32_t* deg = calloc(n + 1, sizeof(uint32_t));
  for (uint32_t k = 0; k < npairs; ++k) {
    deg[bonds[2 * k]]++;
    deg[bonds[2 * k + 1]]++;
  }
  uint32_t* off = malloc((n + 1) * sizeof(uint32_t));
  off[0] = 0;
  for (uint32_t i = 0; i < n; ++i) off[i + 1] = off[i] + deg[i];
  uint32_t* cur = malloc((n + 1) * sizeof(uint32_t));
  memcpy(cur, off, (n + 1) * sizeof(uint32_t));
  uint32_t* adj = malloc((off[n] + 1) * sizeof(uint32_t));
  for (uint32_t k = 0; k < npairs; ++k) {
    uint32_t i = bonds[2 * k], j = bonds[2 * k + 1];
    adj[cur[i]++] = j;
    adj[cur[j]++] = i;
  }
  /* union-find (amphiphile.ts:75-88) */
  int32_t* parent = malloc(n * sizeof(int32_t));
  for (uint32_t i = 0; i < n; ++i) parent[i] = (int32_t)i;
  for (uint32_t k = 0; k < npairs; ++k) {
    uint32_t ra = uf_find_halving(parent, bonds[2 * k]), rb = uf_find_halving(parent, bonds[2 * k + 1]);
    if (ra != rb) parent[ra] = (int32_t)rb;
  }
  /* components in first-occurrence order (amphiphile.ts:90-96) */
  uint32_t* comp_of_root = malloc(n * sizeof(uint32_t));
  for (uint32_t i = 0; i < n; ++i) comp_of_root[i] = A_NONE;
  uint32_t* comp = malloc(n * sizeof(uint32_t));
  uint32_t ncomp = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t r = uf_find_halving(parent, i);
    if (comp_of_root[r] == A_NONE) comp_of_root[r] = ncomp++;
    comp[i] = comp_of_root[r];
  }
  uint32_t* csize = calloc(ncomp + 1, sizeof(uint32_t));
  for (uint32_t i = 0; i < n; ++i) csize[comp[i]]++;
  uint32_t* coff = malloc((ncomp + 1) * sizeof(uint32_t));
  coff[0] = 0;
  for (uint32_t c = 0; c < ncomp; ++c) coff[c + 1] = coff[c] + csize[c];
  uint32_t* ccur = malloc((ncomp + 1) * sizeof(uint32_t));
  memcpy(ccur, coff, (ncomp + 1) * sizeof(uint32_t));
  uint32_t* members = malloc((n + 1) * sizeof(uint32_t));
  for (uint32_t i = 0; i < n; ++i) members[ccur[comp[i]]++] = i;

  uint32_t na = 0, nchain = 0;
  chain_off[0] = 0;
  for (uint32_t c = 0; c < ncomp; ++c) { /* amphiphile.ts:98-138 */
    uint64_t sum_deg = 0;
    uint32_t polar_count = 0, polar_idx = A_NONE;
    int violated = 0;
    for (uint32_t k = coff[c]; k < coff[c + 1]; ++k) {
      uint32_t idx = members[k], d = deg[idx];
      uint32_t kind = (uint32_t)js_round_a((double)pos[4u * idx + 3]);
      int polar = kind < nkinds ? polar_by_kind[kind] != 0 : 0;
      sum_deg += d;
      if (polar) {
        polar_count++;
        polar_idx = idx;
        if (d < 1 || d > 2) violated = 1;
      } else if (d > 2) {
        violated = 1;
      }
    }
    if (violated || polar_count != 1) continue;
    if (sum_deg / 2 != (uint64_t)(coff[c + 1] - coff[c]) - 1u) continue;
    uint32_t t = 0;
    tail_len[2 * na] = tail_len[2 * na + 1] = 0;
    for (uint32_t s = off[polar_idx]; s < off[polar_idx + 1]; ++s) {
      uint32_t prev = polar_idx, at = adj[s], len = 0;
      for (;;) {
        chain[nchain++] = at;
        len++;
        uint32_t next = A_NONE;
        for (uint32_t q = off[at]; q < off[at + 1]; ++q)
          if (adj[q] != prev) {
            next = adj[q];
            break;
          }
        if (next == A_NONE) break;
        prev = at;
        at = next;
      }
      if (t < 2) tail_len[2 * na + t] = len;
      t++;
    }
    head[na] = polar_idx;
    chain_off[na + 1] = nchain;
    na++;
  }
  free(deg); free(off); free(cur); free(adj); free(parent); free(comp_of_root); free(comp);
  free(csize); free(coff); free(ccur); free(members);
  return na;
}

/* ------------------------------------------------------------------------------------------- */
/* clusterComponents (engine/src/aggregate.ts:65-212), z periodic or open: union of every pair with */
/* r2 <= cutoff^2 (minimum image, FP64 over the f32 inputs), found through the reference's cell   */
/* list; components in first-occurrence order, then stably sorted by size, largest first.         */
/* rank[k] = position of k's component in that order; size[r] = its size. Returns the count.      */
/* ------------------------------------------------------------------------------------------- */
static uint32_t uf_find_size(int32_t* parent, uint32_t x) {
  while ((uint32_t)parent[x] != x) {
    parent[x] = parent[parent[x]];
    x = (uint32_t)parent[x];
  }
  return x;
}
static void uf_union_size(int32_t* parent, int32_t* sz, uint32_t a, uint32_t b) { /* aggregate.ts:23-30 */
  uint32_t ra = uf_find_size(parent, a), rb = uf_find_size(parent, b);
  if (ra == rb) return;
  if (sz[ra] < sz[rb]) {
    uint32_t t = ra;
    ra = rb;
    rb = t;
  }
  parent[rb] = (int32_t)ra;
  sz[ra] += sz[rb];
}

uint32_t or_cluster_components(uint32_t m, const float* pos, const double* box, double cutoff, int periodic_z,
                               uint32_t* rank, uint32_t* size_out) {
  if (m == 0) return 0;
  long nx = (long)fmax(1.0, floor(box[0] / cutoff)), ny = (long)fmax(1.0, floor(box[1] / cutoff));
  long nz = periodic_z ? (long)fmax(1.0, floor(box[2] / cutoff)) : 0;
  double wx = box[0] / nx, wy = box[1] / ny, wz = periodic_z ? box[2] / nz : cutoff;
  if (!periodic_z) { /* aggregate.ts:101-105: open z, integer cells of width cutoff (positions in [0, box z)) */
    double zmin = 0.0, zmax = 0.0;
    for (uint32_t i = 0; i < m; ++i) {
      double z = pos[4u * i + 2];
      if (i == 0 || z < zmin) zmin = z;
      if (i == 0 || z > zmax) zmax = z;
    }
    if (zmin < 0.0) zmin = 0.0;
    nz = (long)floor(zmax / wz) + 1;
  }
  long ncell = nx * ny * nz;
  long* cell = malloc(m * sizeof(long));
  uint32_t* cnt = calloc(ncell + 1, sizeof(uint32_t));
  for (uint32_t i = 0; i < m; ++i) {
    double x = wrap1(pos[4u * i], box[0]), y = wrap1(pos[4u * i + 1], box[1]);
    long cx = (long)fmin((double)(nx - 1), floor(x / wx)), cy = (long)fmin((double)(ny - 1), floor(y / wy));
    long cz = periodic_z ? (long)fmin((double)(nz - 1), floor(wrap1(pos[4u * i + 2], box[2]) / wz))
                         : (long)fmax(0.0, floor((double)pos[4u * i + 2] / wz));
    cell[i] = cx + nx * (cy + ny * cz);
    cnt[cell[i] + 1]++;
  }
  for (long c = 0; c < ncell; ++c) cnt[c + 1] += cnt[c];
  uint32_t* cur = malloc((ncell + 1) * sizeof(uint32_t));
  memcpy(cur, cnt, (ncell + 1) * sizeof(uint32_t));
  uint32_t* bucket = malloc(m * sizeof(uint32_t));
  for (uint32_t i = 0; i < m; ++i) bucket[cur[cell[i]]++] = i;
  int32_t* parent = malloc(m * sizeof(int32_t));
  int32_t* sz = malloc(m * sizeof(int32_t));
  for (uint32_t i = 0; i < m; ++i) {
    parent[i] = (int32_t)i;
    sz[i] = 1;
  }
  double c2 = cutoff * cutoff;
  for (uint32_t i = 0; i < m; ++i) {
    long c = cell[i], cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
    double xi = pos[4u * i], yi = pos[4u * i + 1], zi = pos[4u * i + 2];
    for (long dx = -1; dx <= 1; ++dx)
      for (long dy = -1; dy <= 1; ++dy)
        for (long dz = -1; dz <= 1; ++dz) {
          if (!periodic_z && (cz + dz < 0 || cz + dz >= nz)) continue;
          long nc = imod(cx + dx, nx) + nx * (imod(cy + dy, ny) + ny * imod(cz + dz, nz));
          for (uint32_t q = cnt[nc]; q < cnt[nc + 1]; ++q) {
            uint32_t j = bucket[q];
            if (j <= i) continue;
            double ddx = mi1(xi - (double)pos[4u * j], box[0]), ddy = mi1(yi - (double)pos[4u * j + 1], box[1]),
                   ddz = periodic_z ? mi1(zi - (double)pos[4u * j + 2], box[2]) : zi - (double)pos[4u * j + 2];
            if (ddx * ddx + ddy * ddy + ddz * ddz <= c2) uf_union_size(parent, sz, i, j);
          }
        }
  }
  /* first-occurrence order, then a stable sort by size desc (aggregate.ts:201-212) */
  uint32_t* comp_of_root = malloc(m * sizeof(uint32_t));
  for (uint32_t i = 0; i < m; ++i) comp_of_root[i] = A_NONE;
  uint32_t* first = malloc(m * sizeof(uint32_t));
  uint32_t* csz = calloc(m, sizeof(uint32_t));
  uint32_t nc = 0;
  uint32_t* comp = malloc(m * sizeof(uint32_t));
  for (uint32_t i = 0; i < m; ++i) {
    uint32_t r = uf_find_size(parent, i);
    if (comp_of_root[r] == A_NONE) comp_of_root[r] = nc++;
    comp[i] = comp_of_root[r];
    csz[comp[i]]++;
  }
  for (uint32_t c = 0; c < nc; ++c) first[c] = c;
  /* insertion sort on (size desc), stable; fine for the test sizes */
  for (uint32_t a = 1; a < nc; ++a) {
    uint32_t v = first[a], b = a;
    while (b > 0 && csz[first[b - 1]] < csz[v]) {
      first[b] = first[b - 1];
      b--;
    }
    first[b] = v;
  }
  uint32_t* pos_of = malloc(nc * sizeof(uint32_t));
  for (uint32_t r = 0; r < nc; ++r) {
    pos_of[first[r]] = r;
    size_out[r] = csz[first[r]];
  }
  for (uint32_t i = 0; i < m; ++i) rank[i] = pos_of[comp[i]];
  free(cell); free(cnt); free(cur); free(bucket); free(parent); free(sz); free(comp_of_root); free(first);
  free(csz); free(comp); free(pos_of);
  return nc;
}

/* ------------------------------------------------------------------------------------------- */
/* periodicCentreOf (soup/src/water-closure.ts:56-100): circular mean per axis with the Rayleigh  */
/* test. Returns 1 and the centre, or 0 when an axis is refused.                                  */
/* ------------------------------------------------------------------------------------------- */
int or_periodic_centre(const float* pos, const uint32_t* idx, uint32_t n, const double* box, double alpha,
                       double* out) {
  if (n == 0) return 0;
  for (int a = 0; a < 3; ++a) {
    double sc = 0.0, ss = 0.0;
    for (uint32_t k = 0; k < n; ++k) {
      double th = (2.0 * M_PI * wrap1((double)pos[4u * idx[k] + a], box[a])) / box[a];
      sc += cos(th);
      ss += sin(th);
    }
    sc /= n;
    ss /= n;
    double R = sqrt(sc * sc + ss * ss), ang = atan2(ss, sc);
    if (!(R >= sqrt(-log(alpha) / n))) return 0;
    out[a] = wrap1((ang / (2.0 * M_PI)) * box[a], box[a]);
  }
  return 1;
}

/* ------------------------------------------------------------------------------------------- */
/* encapsulatedWaterVolume (soup/src/water-closure.ts:119-363) without a parcel.                  */
/* out: [0] encapsulatedCount [1] totalWater [2] reachedEmptyWater [3] totalEmptyCells            */
/* [4] unreachedCells [5] reachedEmptyCells [6] seed cell [7] dims x [8] y [9] z                  */
/* ------------------------------------------------------------------------------------------- */
static uint8_t* occupancy_periodic(const float* pos, const uint32_t* idx, uint32_t n, const double* box,
                                   double cell, double radius, long* dims) {
  long nx = (long)fmax(1.0, ceil(box[0] / cell)), ny = (long)fmax(1.0, ceil(box[1] / cell)),
       nz = (long)fmax(1.0, ceil(box[2] / cell));
  dims[0] = nx; dims[1] = ny; dims[2] = nz;
  uint8_t* occ = calloc((size_t)nx * ny * nz, 1);
  long reach = (long)fmax(1.0, ceil(radius / cell));
  double r2 = radius * radius;
  for (uint32_t k = 0; k < n; ++k) {
    uint32_t i = idx[k];
    double px = wrap1(pos[4u * i], box[0]), py = wrap1(pos[4u * i + 1], box[1]), pz = wrap1(pos[4u * i + 2], box[2]);
    long cx = (long)floor(px / cell), cy = (long)floor(py / cell), cz = (long)floor(pz / cell);
    for (long dz = -reach; dz <= reach; ++dz) {
      long gz = imod(cz + dz, nz);
      double ddz = (gz + 0.5) * cell - pz;
      ddz -= js_round_a(ddz / box[2]) * box[2];
      for (long dy = -reach; dy <= reach; ++dy) {
        long gy = imod(cy + dy, ny);
        double ddy = (gy + 0.5) * cell - py;
        ddy -= js_round_a(ddy / box[1]) * box[1];
        for (long dx = -reach; dx <= reach; ++dx) {
          long gx = imod(cx + dx, nx);
          double ddx = (gx + 0.5) * cell - px;
          ddx -= js_round_a(ddx / box[0]) * box[0];
          if (ddx * ddx + ddy * ddy + ddz * ddz <= r2) occ[gx + nx * (gy + ny * gz)] = 1;
        }
      }
    }
  }
  return occ;
}

void or_encapsulated_water(const float* pos, const uint32_t* member_idx, uint32_t nm, const uint32_t* water_idx,
                           uint32_t nw, const double* box, const double* centre, double cell, double radius,
                           double* out) {
  long d[3];
  uint8_t* occ = occupancy_periodic(pos, member_idx, nm, box, cell, radius, d);
  long nx = d[0], ny = d[1], nz = d[2], nc = nx * ny * nz;
  /* farthestEmptyCellSeed: z, y, x loops, strict '>' keeps the first maximum */
  long best = -1;
  double best_d2 = -1.0;
  for (long z = 0; z < nz; ++z)
    for (long y = 0; y < ny; ++y)
      for (long x = 0; x < nx; ++x) {
        long i = x + nx * (y + ny * z);
        if (occ[i]) continue;
        double dx = mi1((x + 0.5) * cell - centre[0], box[0]), dy = mi1((y + 0.5) * cell - centre[1], box[1]),
               dz = mi1((z + 0.5) * cell - centre[2], box[2]);
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > best_d2) {
          best_d2 = d2;
          best = i;
        }
      }
  memset(out, 0, 10 * sizeof(double));
  if (best < 0) {
    free(occ);
    out[6] = -1.0;
    return;
  }
  /* floodOutsidePeriodic: breadth-first from the seed, 6-connected, all axes wrap */
  uint8_t* vis = calloc(nc, 1);
  long* queue = malloc(nc * sizeof(long));
  long head = 0, tail = 0;
  vis[best] = 1;
  queue[tail++] = best;
  while (head < tail) {
    long i = queue[head++], z = i / (nx * ny), rem = i - z * nx * ny, y = rem / nx, x = rem - y * nx;
    long nb[6] = {imod(x + 1, nx) + nx * (y + ny * z), imod(x - 1, nx) + nx * (y + ny * z),
                  x + nx * (imod(y + 1, ny) + ny * z), x + nx * (imod(y - 1, ny) + ny * z),
                  x + nx * (y + ny * imod(z + 1, nz)), x + nx * (y + ny * imod(z - 1, nz))};
    for (int k = 0; k < 6; ++k)
      if (!occ[nb[k]] && !vis[nb[k]]) {
        vis[nb[k]] = 1;
        queue[tail++] = nb[k];
      }
  }
  double total_empty = 0, unreached = 0, reached_empty = 0;
  for (long i = 0; i < nc; ++i)
    if (!occ[i]) {
      total_empty++;
      if (!vis[i]) unreached++;
      else reached_empty++;
    }
  double enc = 0, reached_water = 0;
  for (uint32_t k = 0; k < nw; ++k) {
    uint32_t i = water_idx[k];
    double wx = wrap1(pos[4u * i], box[0]), wy = wrap1(pos[4u * i + 1], box[1]), wz = wrap1(pos[4u * i + 2], box[2]);
    long gx = (long)fmin((double)(nx - 1), floor(wx / cell)), gy = (long)fmin((double)(ny - 1), floor(wy / cell)),
         gz = (long)fmin((double)(nz - 1), floor(wz / cell));
    long c = gx + nx * (gy + ny * gz);
    if (occ[c]) continue;
    if (!vis[c]) enc++;
    else reached_water++;
  }
  out[0] = enc; out[1] = nw; out[2] = reached_water; out[3] = total_empty; out[4] = unreached;
  out[5] = reached_empty; out[6] = (double)best; out[7] = nx; out[8] = ny; out[9] = nz;
  free(occ); free(vis); free(queue);
}

/* ------------------------------------------------------------------------------------------- */
/* occupancy + enclosedVolume (engine/src/closure.ts:58-170): x, y wrap when marking, z open;     */
/* flood from the six open faces. Returns the number of unreached empty cells.                    */
/* ------------------------------------------------------------------------------------------- */
double or_enclosed_cells(const float* pos, uint32_t n, const double* box, double cell, double radius) {
  long nx = (long)fmax(1.0, ceil(box[0] / cell)), ny = (long)fmax(1.0, ceil(box[1] / cell)),
       nz = (long)fmax(1.0, ceil(box[2] / cell)), nc = nx * ny * nz;
  uint8_t* occ = calloc(nc, 1);
  long reach = (long)fmax(1.0, ceil(radius / cell));
  double r2 = radius * radius;
  for (uint32_t i = 0; i < n; ++i) {
    double px = pos[4u * i], py = pos[4u * i + 1], pz = pos[4u * i + 2];
    long cx = (long)floor(px / cell), cy = (long)floor(py / cell), cz = (long)floor(pz / cell);
    for (long dz = -reach; dz <= reach; ++dz) {
      long gz = cz + dz;
      if (gz < 0 || gz >= nz) continue;
      double ddz = (gz + 0.5) * cell - pz;
      for (long dy = -reach; dy <= reach; ++dy) {
        long gy = imod(cy + dy, ny);
        double ddy = (gy + 0.5) * cell - py;
        ddy -= js_round_a(ddy / box[1]) * box[1];
        for (long dx = -reach; dx <= reach; ++dx) {
          long gx = imod(cx + dx, nx);
          double ddx = (gx + 0.5) * cell - px;
          ddx -= js_round_a(ddx / box[0]) * box[0];
          if (ddx * ddx + ddy * ddy + ddz * ddz <= r2) occ[gx + nx * (gy + ny * gz)] = 1;
        }
      }
    }
  }
  uint8_t* vis = calloc(nc, 1);
  long* queue = malloc(nc * sizeof(long));
  long head = 0, tail = 0;
#define SEED(X, Y, Z)                                                                 \
  do {                                                                                \
    long sx_ = (X), sy_ = (Y), sz_ = (Z);                                             \
    if (sx_ >= 0 && sx_ < nx && sy_ >= 0 && sy_ < ny && sz_ >= 0 && sz_ < nz) {       \
      long si_ = sx_ + nx * (sy_ + ny * sz_);                                         \
      if (!occ[si_] && !vis[si_]) {                                                   \
        vis[si_] = 1;                                                                 \
        queue[tail++] = si_;                                                          \
      }                                                                               \
    }                                                                                 \
  } while (0)
  for (long y = 0; y < ny; ++y)
    for (long z = 0; z < nz; ++z) { SEED(0, y, z); SEED(nx - 1, y, z); }
  for (long x = 0; x < nx; ++x)
    for (long z = 0; z < nz; ++z) { SEED(x, 0, z); SEED(x, ny - 1, z); }
  for (long x = 0; x < nx; ++x)
    for (long y = 0; y < ny; ++y) { SEED(x, y, 0); SEED(x, y, nz - 1); }
  while (head < tail) {
    long i = queue[head++], z = i / (nx * ny), rem = i - z * nx * ny, y = rem / nx, x = rem - y * nx;
    SEED(x + 1, y, z); SEED(x - 1, y, z); SEED(x, y + 1, z); SEED(x, y - 1, z); SEED(x, y, z + 1); SEED(x, y, z - 1);
  }
#undef SEED
  double unreached = 0;
  for (long i = 0; i < nc; ++i)
    if (!occ[i] && !vis[i]) unreached++;
  free(occ); free(vis); free(queue);
  return unreached;
}

/* ------------------------------------------------------------------------------------------- */
/* shapeOf (engine/src/aggregate.ts:339-362) with symmetricEigenvalues3x3 / symmetricEigenvector  */
/* (:260-311). out: centre[3], radiusOfGyration, principalMoments[3], smallestAxis[3].            */
/* ------------------------------------------------------------------------------------------- */
static void sort3(double* v) {
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2 - a; ++b)
      if (v[b] > v[b + 1]) {
        double t = v[b];
        v[b] = v[b + 1];
        v[b + 1] = t;
      }
}

void or_shape_of(const float* pos, uint32_t n, double* out) {
  double cx = 0, cy = 0, cz = 0;
  for (uint32_t i = 0; i < n; ++i) {
    cx += pos[4u * i];
    cy += pos[4u * i + 1];
    cz += pos[4u * i + 2];
  }
  cx /= n; cy /= n; cz /= n;
  double Sxx = 0, Syy = 0, Szz = 0, Sxy = 0, Sxz = 0, Syz = 0;
  for (uint32_t i = 0; i < n; ++i) {
    double dx = pos[4u * i] - cx, dy = pos[4u * i + 1] - cy, dz = pos[4u * i + 2] - cz;
    Sxx += dx * dx; Syy += dy * dy; Szz += dz * dz;
    Sxy += dx * dy; Sxz += dx * dz; Syz += dy * dz;
  }
  Sxx /= n; Syy /= n; Szz /= n; Sxy /= n; Sxz /= n; Syz /= n;
  double ev[3];
  double p1 = Sxy * Sxy + Sxz * Sxz + Syz * Syz;
  if (p1 < 1e-30) {
    ev[0] = Sxx; ev[1] = Syy; ev[2] = Szz;
    sort3(ev);
  } else {
    double q = (Sxx + Syy + Szz) / 3;
    double p2 = (Sxx - q) * (Sxx - q) + (Syy - q) * (Syy - q) + (Szz - q) * (Szz - q) + 2 * p1;
    double p = sqrt(p2 / 6);
    double Bxx = (Sxx - q) / p, Byy = (Syy - q) / p, Bzz = (Szz - q) / p, Bxy = Sxy / p, Bxz = Sxz / p, Byz = Syz / p;
    double detB = Bxx * (Byy * Bzz - Byz * Byz) - Bxy * (Bxy * Bzz - Byz * Bxz) + Bxz * (Bxy * Byz - Byy * Bxz);
    double r = fmax(-1.0, fmin(1.0, detB / 2));
    double phi = acos(r) / 3;
    double largest = q + 2 * p * cos(phi), smallest = q + 2 * p * cos(phi + (2 * M_PI) / 3);
    ev[0] = smallest; ev[1] = 3 * q - largest - smallest; ev[2] = largest;
    sort3(ev);
  }
  double lam = ev[0];
  double r0[3] = {Sxx - lam, Sxy, Sxz}, r1[3] = {Sxy, Syy - lam, Syz}, r2[3] = {Sxz, Syz, Szz - lam};
  double cands[3][3];
  const double* pa[3] = {r0, r0, r1};
  const double* pb[3] = {r1, r2, r2};
  for (int k = 0; k < 3; ++k) {
    const double* a = pa[k];
    const double* b = pb[k];
    cands[k][0] = a[1] * b[2] - a[2] * b[1];
    cands[k][1] = a[2] * b[0] - a[0] * b[2];
    cands[k][2] = a[0] * b[1] - a[1] * b[0];
  }
  int best = 0;
  double bm = sqrt(cands[0][0] * cands[0][0] + cands[0][1] * cands[0][1] + cands[0][2] * cands[0][2]);
  for (int k = 1; k < 3; ++k) {
    double mg = sqrt(cands[k][0] * cands[k][0] + cands[k][1] * cands[k][1] + cands[k][2] * cands[k][2]);
    if (mg > bm) {
      bm = mg;
      best = k;
    }
  }
  out[0] = cx; out[1] = cy; out[2] = cz;
  out[3] = sqrt(fmax(0.0, ev[0] + ev[1] + ev[2]));
  out[4] = ev[0]; out[5] = ev[1]; out[6] = ev[2];
  if (bm < 1e-9) {
    out[7] = 0; out[8] = 0; out[9] = 1;
  } else {
    out[7] = cands[best][0] / bm; out[8] = cands[best][1] / bm; out[9] = cands[best][2] / bm;
  }
}

/* ------------------------------------------------------------------------------------------- */
/* computeHeadPeaks (soup/src/aggregates.ts:83-120) over densityProfileZ / bilayerPeaks /          */
/* subBinPeak (engine/src/metrics.ts:21-52, :91-107, :118-157). pos w holds kind indices.         */
/* Returns 0, 1, 2, or -1 for HEAD_PEAKS_UNAVAILABLE.                                             */
/* ------------------------------------------------------------------------------------------- */
static double sub_bin_peak(const double* v, const double* z, int nb, int i) {
  if (i == 0 || i == nb - 1 || nb < 2) return z[i];
  double dz = z[1] - z[0];
  double denom = v[i - 1] - 2 * v[i] + v[i + 1];
  if (denom >= 0) return z[i];
  double off = (0.5 * (v[i - 1] - v[i + 1])) / denom;
  return z[i] + fmax(-0.5, fmin(0.5, off)) * dz;
}

int or_head_peaks(const float* pos, uint32_t n, const double* box, const uint8_t* polar_by_kind, uint32_t nkinds,
                  uint32_t bins, double min_per_bin, double sep_min, double sep_max, double sigma) {
  uint32_t total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t k = (uint32_t)js_round_a((double)pos[4u * i + 3]);
    if (k < nkinds && polar_by_kind[k]) total++;
  }
  if (total == 0) return 0;
  double* head = calloc(bins, sizeof(double));
  double* zc = malloc(bins * sizeof(double));
  double dz = box[2] / bins;
  uint32_t kept = 0;
  for (uint32_t i = 0; i < n; ++i) {
    double z = pos[4u * i + 2];
    if (!(z >= 0 && z < box[2])) continue; /* dropEscapedZ */
    uint32_t k = (uint32_t)js_round_a((double)pos[4u * i + 3]);
    int is_head = k < nkinds && polar_by_kind[k];
    if (!is_head) continue;
    kept++;
    long b = (long)fmin((double)(bins - 1), floor(z / dz));
    head[b] += 1;
  }
  int res;
  if (kept == 0 || (double)kept / bins < min_per_bin) {
    res = -1;
  } else {
    for (uint32_t b = 0; b < bins; ++b) zc[b] = (b + 0.5) * dz;
    uint32_t a = 0;
    for (uint32_t i = 1; i < bins; ++i)
      if (head[i] > head[a]) a = i;
    double* masked = malloc(bins * sizeof(double));
    for (uint32_t i = 0; i < bins; ++i) masked[i] = fabs(zc[i] - zc[a]) <= sigma ? 0 : head[i];
    uint32_t b2 = 0;
    for (uint32_t i = 1; i < bins; ++i)
      if (masked[i] > masked[b2]) b2 = i;
    if (head[a] == 0 || masked[b2] == 0) {
      res = 1;
    } else {
      double za = sub_bin_peak(head, zc, (int)bins, (int)a), zb = sub_bin_peak(head, zc, (int)bins, (int)b2);
      double sep = za <= zb ? zb - za : za - zb;
      res = (sep >= sep_min && sep <= sep_max) ? 2 : 1;
    }
    free(masked);
  }
  free(head);
  free(zc);
  return res;
}
