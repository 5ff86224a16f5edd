// soup_analysis.inl -- HIP kernels and C-ABI of the snapshot analysis (included once by
// soup_lib.hip): amphiphile recognition over the bond graph (soup/src/amphiphile.ts:55-141),
// member clustering (engine/src/aggregate.ts:65-212), the periodic encapsulated-water flood
// (soup/src/water-closure.ts:119-363) and the open-boundary cavity flood (engine/src/closure.ts:
// 58-170).
//
// Every connected-component question is answered by one lock-free union-find: a union hooks the
// larger of two roots under the smaller with a compare-and-swap, so every pointer goes to a
// smaller index, every root is the minimum of its component, and the flattened labels do not
// depend on the order the unions ran in. A breadth-first flood from a seed visits exactly the
// seed's connected empty component, so "visited" is "has the seed's label".

// ------------------------------------------------------------------------------------------------
// Union-find
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t uf_ld(const uint32_t* p, uint32_t i) {
  return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t uf_root(uint32_t* p, uint32_t x) {
  uint32_t cur = x, nxt = uf_ld(p, cur);
  while (nxt != cur) {
    const uint32_t nn = uf_ld(p, nxt);
    if (nn != nxt) __hip_atomic_store(p + cur, nn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // path splitting
    cur = nxt;
    nxt = nn;
  }
  return cur;
}
__device__ __forceinline__ void uf_unite(uint32_t* p, uint32_t a, uint32_t b) {
  for (;;) {
    a = uf_root(p, a);
    b = uf_root(p, b);
    if (a == b) return;
    const uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
    uint32_t expected = hi;
    if (__hip_atomic_compare_exchange_strong(p + hi, &expected, lo, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return;
    a = hi;
    b = lo;
  }
}
__global__ void k_uf_init(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}
// The final labels: every entry set to its root. The walk here stores nothing along the way: a
// path-splitting store of one thread could land on an entry after that entry's own thread wrote its
// root there, and leave it pointing at a non-root ancestor. With only each thread's own final store,
// every value a walk reads is an ancestor, so every walk ends at the root.
__global__ void k_uf_flatten(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t cur = i, nxt = uf_ld(p, cur);
  while (nxt != cur) {
    cur = nxt;
    nxt = uf_ld(p, cur);
  }
  __hip_atomic_store(p + i, cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double an_round(double x) {  // Math.round
  double r = floor(x);
  if (x - r >= 0.5) r += 1.0;
  return r;
}
__device__ __forceinline__ double an_mi(double d, double L) { return d - an_round(d / L) * L; }
__device__ __forceinline__ double an_wrap(double v, double L) { return v - floor(v / L) * L; }
__device__ __forceinline__ int an_mod(int a, int n) { return ((a % n) + n) % n; }

// ------------------------------------------------------------------------------------------------
// findAmphiphiles: components of the bond graph, then the per-component checks of
// amphiphile.ts:98-118 and the tail walks of :120-136. Particles are in particle-index order.
// ------------------------------------------------------------------------------------------------
__global__ void k_an_bond_union(const uint32_t* bs, uint32_t n, uint32_t* p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < 3; ++k) {
    const uint32_t j = bs[3u * i + k];
    if (j != SOUP_NONE && j < n && j > i) uf_unite(p, i, j);
  }
}
__device__ __forceinline__ bool an_polar(const float4* pos, uint32_t i, uint32_t polar_mask) {
  const double k = an_round((double)pos[i].w);
  return k >= 0.0 && k < 32.0 && ((polar_mask >> (uint32_t)k) & 1u) != 0u;
}
__device__ __forceinline__ uint32_t an_degree(const uint32_t* bs, uint32_t i) {
  return (bs[3u * i] != SOUP_NONE) + (bs[3u * i + 1] != SOUP_NONE) + (bs[3u * i + 2] != SOUP_NONE);
}
// stat[4r]: size, degree sum, polar count, violation (a polar bead of degree 0 or > 2, or a
// non-polar bead of degree > 2)
__global__ void k_an_comp_stats(const float4* pos, const uint32_t* bs, uint32_t n, const uint32_t* p, uint32_t polar_mask,
                                uint32_t* stat, uint32_t* polar_of) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = p[i], d = an_degree(bs, i);
  const bool polar = an_polar(pos, i, polar_mask);
  atomicAdd(stat + 4u * r, 1u);
  atomicAdd(stat + 4u * r + 1u, d);
  if (polar) {
    atomicAdd(stat + 4u * r + 2u, 1u);
    polar_of[r] = i;  // read only when this component holds exactly one polar bead
  }
  if (polar ? (d < 1u || d > 2u) : d > 2u) atomicOr(stat + 4u * r + 3u, 1u);
}
// flag[r] = 1 for the root of an amphiphile component; clen[r] = its chain length (size - 1)
__global__ void k_an_amph_flag(const uint32_t* p, const uint32_t* stat, uint32_t n, uint32_t* flag, uint32_t* clen) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  bool ok = false;
  if (p[r] == r) {
    const uint32_t size = stat[4u * r], dsum = stat[4u * r + 1u], npol = stat[4u * r + 2u], viol = stat[4u * r + 3u];
    ok = viol == 0u && npol == 1u && dsum / 2u == size - 1u;
  }
  flag[r] = ok ? 1u : 0u;
  clen[r] = ok ? stat[4u * r] - 1u : 0u;
}
// adjacency[x] of amphiphile.ts:66-73 for x's bonds in bonds() order: partners below x ascending,
// then partners above x in slot order
__device__ __forceinline__ uint32_t an_adjacency(const uint32_t* bs, uint32_t x, uint32_t* out) {
  uint32_t lo[3], nlo = 0, hi[3], nhi = 0;
  for (int k = 0; k < 3; ++k) {
    const uint32_t j = bs[3u * x + k];
    if (j == SOUP_NONE) continue;
    if (j < x) lo[nlo++] = j;
    else hi[nhi++] = j;
  }
  for (uint32_t a = 1; a < nlo; ++a)
    for (uint32_t b = a; b > 0 && lo[b - 1] > lo[b]; --b) {
      const uint32_t t = lo[b];
      lo[b] = lo[b - 1];
      lo[b - 1] = t;
    }
  uint32_t m = 0;
  for (uint32_t a = 0; a < nlo; ++a) out[m++] = lo[a];
  for (uint32_t a = 0; a < nhi; ++a) out[m++] = hi[a];
  return m;
}
__global__ void k_an_amph_write(const uint32_t* bs, uint32_t n, const uint32_t* stat, const uint32_t* polar_of,
                                const uint32_t* flag, const uint32_t* rank, const uint32_t* coff, uint32_t* heads,
                                uint32_t* chain_off, uint32_t* chains, uint32_t* tails) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || !flag[r]) return;
  const uint32_t a = rank[r], h = polar_of[r], size = stat[4u * r];
  heads[a] = h;
  chain_off[a] = coff[r];
  uint32_t w = coff[r];
  uint32_t nb[3];
  const uint32_t deg = an_adjacency(bs, h, nb);
  tails[2u * a] = tails[2u * a + 1u] = 0u;
  for (uint32_t t = 0; t < deg; ++t) {
    uint32_t prev = h, cur = nb[t], len = 0;
    for (uint32_t guard = 0; guard < size; ++guard) {
      chains[w++] = cur;
      len++;
      uint32_t cn[3];
      const uint32_t cd = an_adjacency(bs, cur, cn);
      uint32_t next = SOUP_NONE;
      for (uint32_t q = 0; q < cd; ++q)
        if (cn[q] != prev) {
          next = cn[q];
          break;
        }
      if (next == SOUP_NONE) break;
      prev = cur;
      cur = next;
    }
    if (t < 2u) tails[2u * a + t] = len;
  }
}

// ------------------------------------------------------------------------------------------------
// clusterComponents over a member subset: the reference's cell list (side >= cutoff), every pair
// with r2 <= cutoff^2 (FP64 minimum image over the f32 coordinates) united.
// ------------------------------------------------------------------------------------------------
struct AnGrid {
  int nx, ny, nz, pz;  // pz = 0: z is open (cells of side cutoff from z = 0, no wrap)
  double wx, wy, wz;
};
__global__ void k_an_gather(const float4* pos, const uint32_t* idx, uint32_t m, float4* out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) out[k] = pos[idx[k]];
}
__device__ __forceinline__ uint32_t an_cell(const AnGrid& g, double3 box, float4 x) {
  const int cx = min(g.nx - 1, (int)floor(an_wrap((double)x.x, box.x) / g.wx));
  const int cy = min(g.ny - 1, (int)floor(an_wrap((double)x.y, box.y) / g.wy));
  const int cz = g.pz ? min(g.nz - 1, (int)floor(an_wrap((double)x.z, box.z) / g.wz))
                      : max(0, min(g.nz - 1, (int)floor((double)x.z / g.wz)));
  return (uint32_t)(cx + g.nx * (cy + g.ny * cz));
}
__global__ void k_an_cell_count(const float4* mp, uint32_t m, AnGrid g, double3 box, uint32_t* cell, uint32_t* cnt) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint32_t c = an_cell(g, box, mp[k]);
  cell[k] = c;
  atomicAdd(cnt + c, 1u);
}
__global__ void k_an_cell_fill(const uint32_t* cell, uint32_t m, uint32_t* cursor, uint32_t* bucket) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) bucket[atomicAdd(cursor + cell[k], 1u)] = k;
}
__global__ void k_an_member_union(const float4* mp, uint32_t m, AnGrid g, double3 box, double cut2, const uint32_t* cell,
                                  const uint32_t* start, const uint32_t* bucket, uint32_t* p) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint32_t c = cell[k];
  const int cx = (int)(c % (uint32_t)g.nx), cy = (int)((c / (uint32_t)g.nx) % (uint32_t)g.ny),
            cz = (int)(c / (uint32_t)(g.nx * g.ny));
  const float4 xi = mp[k];
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        if (!g.pz && (cz + dz < 0 || cz + dz >= g.nz)) continue;
        const uint32_t nc = (uint32_t)(an_mod(cx + dx, g.nx) + g.nx * (an_mod(cy + dy, g.ny) + g.ny * an_mod(cz + dz, g.nz)));
        for (uint32_t q = start[nc]; q < start[nc + 1]; ++q) {
          const uint32_t j = bucket[q];
          if (j <= k) continue;
          const float4 xj = mp[j];
          const double ddx = an_mi((double)xi.x - (double)xj.x, box.x), ddy = an_mi((double)xi.y - (double)xj.y, box.y),
                       ddz = g.pz ? an_mi((double)xi.z - (double)xj.z, box.z) : (double)xi.z - (double)xj.z;
          if (ddx * ddx + ddy * ddy + ddz * ddz <= cut2) uf_unite(p, k, j);
        }
      }
}

// ------------------------------------------------------------------------------------------------
// Occupancy grids and floods (cell side `cell`, dims = ceil(box / cell) per axis)
// ------------------------------------------------------------------------------------------------
struct AnOcc {
  int nx, ny, nz, reach;
  double cell, r2;
  double3 box;
};
// occupancyPeriodic (water-closure.ts:119-148): all three axes wrap
__global__ void k_an_occ_periodic(const float4* mp, uint32_t m, AnOcc o, uint8_t* occ) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const double px = an_wrap((double)mp[k].x, o.box.x), py = an_wrap((double)mp[k].y, o.box.y), pz = an_wrap((double)mp[k].z, o.box.z);
  const int cx = (int)floor(px / o.cell), cy = (int)floor(py / o.cell), cz = (int)floor(pz / o.cell);
  for (int dz = -o.reach; dz <= o.reach; ++dz) {
    const int gz = an_mod(cz + dz, o.nz);
    double ddz = (gz + 0.5) * o.cell - pz;
    ddz -= an_round(ddz / o.box.z) * o.box.z;
    for (int dy = -o.reach; dy <= o.reach; ++dy) {
      const int gy = an_mod(cy + dy, o.ny);
      double ddy = (gy + 0.5) * o.cell - py;
      ddy -= an_round(ddy / o.box.y) * o.box.y;
      for (int dx = -o.reach; dx <= o.reach; ++dx) {
        const int gx = an_mod(cx + dx, o.nx);
        double ddx = (gx + 0.5) * o.cell - px;
        ddx -= an_round(ddx / o.box.x) * o.box.x;
        if (ddx * ddx + ddy * ddy + ddz * ddz <= o.r2) occ[gx + o.nx * (gy + o.ny * gz)] = 1;
      }
    }
  }
}
// occupancy (closure.ts:58-99): x, y wrap when measuring, z open
__global__ void k_an_occ_open(const float4* mp, uint32_t m, AnOcc o, uint8_t* occ) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const double px = mp[k].x, py = mp[k].y, pz = mp[k].z;
  const int cx = (int)floor(px / o.cell), cy = (int)floor(py / o.cell), cz = (int)floor(pz / o.cell);
  for (int dz = -o.reach; dz <= o.reach; ++dz) {
    const int gz = cz + dz;
    if (gz < 0 || gz >= o.nz) continue;
    const double ddz = (gz + 0.5) * o.cell - pz;
    for (int dy = -o.reach; dy <= o.reach; ++dy) {
      const int gy = an_mod(cy + dy, o.ny);
      double ddy = (gy + 0.5) * o.cell - py;
      ddy -= an_round(ddy / o.box.y) * o.box.y;
      for (int dx = -o.reach; dx <= o.reach; ++dx) {
        const int gx = an_mod(cx + dx, o.nx);
        double ddx = (gx + 0.5) * o.cell - px;
        ddx -= an_round(ddx / o.box.x) * o.box.x;
        if (ddx * ddx + ddy * ddy + ddz * ddz <= o.r2) occ[gx + o.nx * (gy + o.ny * gz)] = 1;
      }
    }
  }
}
// farthestEmptyCellSeed (water-closure.ts:194-222): the largest minimum-image d2 from the centre,
// the first cell in x-fastest order among equals (two passes: the maximum, then the first index)
__device__ __forceinline__ double an_seed_d2(const AnOcc& o, double3 c, uint32_t i) {
  const int x = (int)(i % (uint32_t)o.nx), y = (int)((i / (uint32_t)o.nx) % (uint32_t)o.ny), z = (int)(i / (uint32_t)(o.nx * o.ny));
  const double dx = an_mi((x + 0.5) * o.cell - c.x, o.box.x), dy = an_mi((y + 0.5) * o.cell - c.y, o.box.y),
               dz = an_mi((z + 0.5) * o.cell - c.z, o.box.z);
  return dx * dx + dy * dy + dz * dz;
}
__global__ void k_an_seed_max(const uint8_t* occ, uint32_t nc, AnOcc o, double3 c, unsigned long long* best) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = 0ull;
  if (i < nc && !occ[i]) v = (unsigned long long)__double_as_longlong(an_seed_d2(o, c, i)) + 1ull;  // 0 = no empty cell
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long u = __shfl_xor(v, off, WAVE);
    v = u > v ? u : v;
  }
  if ((threadIdx.x & (WAVE - 1)) == 0 && v) atomicMax(best, v);
}
__global__ void k_an_seed_first(const uint8_t* occ, uint32_t nc, AnOcc o, double3 c, const unsigned long long* best, uint32_t* seed) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc || occ[i]) return;
  if ((unsigned long long)__double_as_longlong(an_seed_d2(o, c, i)) + 1ull == *best) atomicMin(seed, i);
}
// unions of empty face neighbours: periodic on every axis, or open (no wrap)
__global__ void k_an_grid_union(const uint8_t* occ, AnOcc o, int periodic, uint32_t* p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nc = (uint32_t)o.nx * (uint32_t)o.ny * (uint32_t)o.nz;
  if (i >= nc || occ[i]) return;
  const int x = (int)(i % (uint32_t)o.nx), y = (int)((i / (uint32_t)o.nx) % (uint32_t)o.ny), z = (int)(i / (uint32_t)(o.nx * o.ny));
  if (periodic || x + 1 < o.nx) {
    const uint32_t j = (uint32_t)((x + 1) % o.nx + o.nx * (y + o.ny * z));
    if (!occ[j]) uf_unite(p, i, j);
  }
  if (periodic || y + 1 < o.ny) {
    const uint32_t j = (uint32_t)(x + o.nx * ((y + 1) % o.ny + o.ny * z));
    if (!occ[j]) uf_unite(p, i, j);
  }
  if (periodic || z + 1 < o.nz) {
    const uint32_t j = (uint32_t)(x + o.nx * (y + o.ny * ((z + 1) % o.nz)));
    if (!occ[j]) uf_unite(p, i, j);
  }
}
// the components touching a face of an open grid are "outside" (floodOutside, closure.ts:109-140)
__global__ void k_an_mark_faces(const uint8_t* occ, AnOcc o, const uint32_t* p, uint8_t* outside) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nc = (uint32_t)o.nx * (uint32_t)o.ny * (uint32_t)o.nz;
  if (i >= nc || occ[i]) return;
  const int x = (int)(i % (uint32_t)o.nx), y = (int)((i / (uint32_t)o.nx) % (uint32_t)o.ny), z = (int)(i / (uint32_t)(o.nx * o.ny));
  if (x == 0 || x == o.nx - 1 || y == 0 || y == o.ny - 1 || z == 0 || z == o.nz - 1) outside[p[i]] = 1;
}
// cnt[0] empty cells, [1] unreached empty cells, [2] reached empty cells; reached = in the seed's
// component (periodic) or in a face-touching component (open)
__global__ void k_an_cell_counts(const uint8_t* occ, uint32_t nc, const uint32_t* p, const uint32_t* seed,
                                 const uint8_t* outside, unsigned long long* cnt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t e = 0, u = 0;
  if (i < nc && !occ[i]) {
    e = 1;
    const bool reached = outside ? outside[p[i]] != 0 : p[i] == p[*seed];
    u = reached ? 0u : 1u;
  }
  e = wave_sum_u(e);
  u = wave_sum_u(u);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    if (e) atomicAdd(cnt, (unsigned long long)e);
    if (u) atomicAdd(cnt + 1, (unsigned long long)u);
    if (e - u) atomicAdd(cnt + 2, (unsigned long long)(e - u));
  }
}
// water beads (kind == water_kind): wc[0] total, [1] encapsulated (empty, unreached cell),
// [2] reached (empty, reached cell) -- water-closure.ts:300-315
__global__ void k_an_water(const float4* pos, uint32_t n, uint32_t water_kind, AnOcc o, const uint8_t* occ,
                           const uint32_t* p, const uint32_t* seed, uint32_t* wc) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t t = 0, e = 0, r = 0;
  if (i < n && an_round((double)pos[i].w) == (double)water_kind) {
    t = 1;
    const float4 x = pos[i];
    const int gx = min(o.nx - 1, (int)floor(an_wrap((double)x.x, o.box.x) / o.cell));
    const int gy = min(o.ny - 1, (int)floor(an_wrap((double)x.y, o.box.y) / o.cell));
    const int gz = min(o.nz - 1, (int)floor(an_wrap((double)x.z, o.box.z) / o.cell));
    const uint32_t c = (uint32_t)(gx + o.nx * (gy + o.ny * gz));
    if (!occ[c]) {
      if (p[c] == p[*seed]) r = 1;
      else e = 1;
    }
  }
  t = wave_sum_u(t);
  e = wave_sum_u(e);
  r = wave_sum_u(r);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    if (t) atomicAdd(wc, t);
    if (e) atomicAdd(wc + 1, e);
    if (r) atomicAdd(wc + 2, r);
  }
}

// ------------------------------------------------------------------------------------------------
// Host side: the analysis context
// ------------------------------------------------------------------------------------------------
struct SoupAnalysis {
  hipStream_t stream = nullptr;
  std::string err;
  uint32_t n = 0, cap_n = 0;
  float4* pos = nullptr;        // particle-index order
  uint32_t* bs = nullptr;       // 3 per particle
  uint32_t* parent = nullptr;   // bond-graph union-find (n)
  uint32_t* stat = nullptr;     // 4 per particle
  uint32_t* polar_of = nullptr;
  uint32_t* flag = nullptr;     // n + 1
  uint32_t* rank = nullptr;     // n + 1
  uint32_t* clen = nullptr;     // n + 1
  uint32_t* coff = nullptr;     // n + 1
  uint32_t* heads = nullptr;
  uint32_t* chain_off = nullptr;  // n + 1
  uint32_t* chains = nullptr;
  uint32_t* tails = nullptr;    // 2 per particle
  uint32_t n_amph = 0, n_chain = 0;
  uint32_t* scan_blocks = nullptr;
  // member subsets
  uint32_t cap_m = 0;
  uint32_t* idx = nullptr;
  float4* mp = nullptr;
  uint32_t* mcell = nullptr;
  uint32_t* mparent = nullptr;
  uint32_t* bucket = nullptr;
  uint32_t cap_cells = 0;
  uint32_t* ccount = nullptr;   // cap_cells + 1
  uint32_t* cstart = nullptr;   // cap_cells + 1
  // occupancy grid
  size_t cap_grid = 0;
  uint8_t* occ = nullptr;
  uint8_t* outside = nullptr;
  uint32_t* gparent = nullptr;
  unsigned long long* best = nullptr;  // [0] seed key, [1..3] counts
  uint32_t* seed = nullptr;            // [0] seed, [1..3] water counts
};

static int afail(SoupAnalysis* a, int code, const std::string& msg) {
  if (a) a->err = msg;
  else g_create_error = msg;
  return code;
}
#define AHC(a, expr)                                                                                     \
  do {                                                                                                   \
    hipError_t e_ = (expr);                                                                              \
    if (e_ != hipSuccess) return afail(a, SOUP_E_DEVICE, fmt("HIP error %s at %s", hipGetErrorString(e_), #expr)); \
  } while (0)

template <typename T>
static int agrow(SoupAnalysis* a, T** p, size_t count) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return afail(a, SOUP_E_NOMEM, fmt("analysis: device allocation of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e)));
  *p = (T*)q;
  return 0;
}

static int an_reserve_n(SoupAnalysis* a, uint32_t n) {
  if (n <= a->cap_n && a->pos) return 0;
  int r = 0;
  if ((r = agrow(a, &a->pos, n)) || (r = agrow(a, &a->bs, 3ull * n)) || (r = agrow(a, &a->parent, n)) ||
      (r = agrow(a, &a->stat, 4ull * n)) || (r = agrow(a, &a->polar_of, n)) || (r = agrow(a, &a->flag, n + 1ull)) ||
      (r = agrow(a, &a->rank, n + 1ull)) || (r = agrow(a, &a->clen, n + 1ull)) || (r = agrow(a, &a->coff, n + 1ull)) ||
      (r = agrow(a, &a->heads, n)) || (r = agrow(a, &a->chain_off, n + 1ull)) || (r = agrow(a, &a->chains, n)) ||
      (r = agrow(a, &a->tails, 2ull * n)))
    return r;
  a->cap_n = n;
  return 0;
}
static int an_reserve_m(SoupAnalysis* a, uint32_t m) {
  if (m <= a->cap_m && a->idx) return 0;
  int r = 0;
  if ((r = agrow(a, &a->idx, m)) || (r = agrow(a, &a->mp, m)) || (r = agrow(a, &a->mcell, m)) ||
      (r = agrow(a, &a->mparent, m)) || (r = agrow(a, &a->bucket, m)))
    return r;
  a->cap_m = m;
  return 0;
}
static int an_reserve_cells(SoupAnalysis* a, uint32_t nc) {
  if (nc <= a->cap_cells && a->ccount) return 0;
  int r = 0;
  if ((r = agrow(a, &a->ccount, nc + 1ull)) || (r = agrow(a, &a->cstart, nc + 1ull))) return r;
  a->cap_cells = nc;
  return 0;
}
static int an_reserve_grid(SoupAnalysis* a, size_t nc) {
  if (nc <= a->cap_grid && a->occ) return 0;
  int r = 0;
  if ((r = agrow(a, &a->occ, nc)) || (r = agrow(a, &a->outside, nc)) || (r = agrow(a, &a->gparent, nc))) return r;
  a->cap_grid = nc;
  return 0;
}
static void an_fill(SoupAnalysis* a, uint32_t* p, uint32_t v, uint64_t count) {
  if (count) hipLaunchKernelGGL(k_fill_u32, dim3(blocks(count, 256)), dim3(256), 0, a->stream, p, v, (uint32_t)count);
}
static void an_fill8(SoupAnalysis* a, uint8_t* p, uint64_t count) {
  // a byte grid cleared as u32 words (callers size grids with 4 spare bytes)
  an_fill(a, reinterpret_cast<uint32_t*>(p), 0u, (count + 3) / 4);
}
static int an_sync(SoupAnalysis* a) {
  AHC(a, hipGetLastError());
  AHC(a, hipStreamSynchronize(a->stream));
  return 0;
}
static int an_grid_dims(SoupAnalysis* a, const double* box, double cell, double radius, AnOcc* o) {
  if (!(cell > 0.0) || !(radius >= 0.0)) return afail(a, SOUP_E_INVALID, "analysis: the grid cell and radius must be positive");
  double d[3];
  for (int k = 0; k < 3; ++k) {
    if (!(box[k] > 0.0)) return afail(a, SOUP_E_INVALID, "analysis: box lengths must be positive");
    d[k] = std::fmax(1.0, std::ceil(box[k] / cell));
  }
  if (d[0] * d[1] * d[2] > 4.0e8) return afail(a, SOUP_E_INVALID, "analysis: the occupancy grid exceeds 4e8 cells");
  o->nx = (int)d[0];
  o->ny = (int)d[1];
  o->nz = (int)d[2];
  o->reach = (int)std::fmax(1.0, std::ceil(radius / cell));
  o->cell = cell;
  o->r2 = radius * radius;
  o->box = make_double3(box[0], box[1], box[2]);
  return an_reserve_grid(a, (size_t)o->nx * o->ny * o->nz + 4);
}
// the grid flood shared by the periodic and open cases: union of empty face neighbours, flatten
static void an_flood(SoupAnalysis* a, const AnOcc& o, int periodic) {
  const uint32_t nc = (uint32_t)o.nx * (uint32_t)o.ny * (uint32_t)o.nz;
  hipLaunchKernelGGL(k_uf_init, dim3(blocks(nc, 256)), dim3(256), 0, a->stream, a->gparent, nc);
  hipLaunchKernelGGL(k_an_grid_union, dim3(blocks(nc, 256)), dim3(256), 0, a->stream, a->occ, o, periodic, a->gparent);
  hipLaunchKernelGGL(k_uf_flatten, dim3(blocks(nc, 256)), dim3(256), 0, a->stream, a->gparent, nc);
}

extern "C" {

int soup_analysis_create(SoupAnalysis** out) {
  if (!out) return afail(nullptr, SOUP_E_INVALID, "soup_analysis_create: null output");
  *out = nullptr;
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
    return afail(nullptr, SOUP_E_DEVICE, "soup_analysis_create: no HIP device is available");
  SoupAnalysis* a = new SoupAnalysis();
  if (hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking) != hipSuccess) {
    delete a;
    return afail(nullptr, SOUP_E_DEVICE, "soup_analysis_create: stream creation failed");
  }
  int r = agrow(a, &a->scan_blocks, SCAN_TILE + 1);
  if (!r) r = agrow(a, &a->best, 4);
  if (!r) r = agrow(a, &a->seed, 4);
  if (r) {
    g_create_error = a->err;
    soup_analysis_destroy(a);
    return r;
  }
  *out = a;
  return 0;
}

int soup_analysis_destroy(SoupAnalysis* a) {
  if (!a) return 0;
  if (a->stream) (void)hipStreamSynchronize(a->stream);
  void* bufs[] = {a->pos, a->bs, a->parent, a->stat, a->polar_of, a->flag, a->rank, a->clen, a->coff, a->heads,
                  a->chain_off, a->chains, a->tails, a->scan_blocks, a->idx, a->mp, a->mcell, a->mparent, a->bucket,
                  a->ccount, a->cstart, a->occ, a->outside, a->gparent, a->best, a->seed};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (a->stream) (void)hipStreamDestroy(a->stream);
  delete a;
  return 0;
}

const char* soup_analysis_last_error(const SoupAnalysis* a) { return a ? a->err.c_str() : g_create_error.c_str(); }

int soup_analysis_load(SoupAnalysis* a, uint32_t n, const float* pos4, const uint32_t* bond_slots) {
  if (!a) return SOUP_E_INVALID;
  if (!pos4 || (!bond_slots && n)) return afail(a, SOUP_E_INVALID, "soup_analysis_load: null positions or bond slots");
  int r = an_reserve_n(a, n);
  if (r) return r;
  a->n = n;
  a->n_amph = a->n_chain = 0;
  if (n) {
    AHC(a, hipMemcpyAsync(a->pos, pos4, (size_t)n * 16, hipMemcpyHostToDevice, a->stream));
    AHC(a, hipMemcpyAsync(a->bs, bond_slots, (size_t)n * 12, hipMemcpyHostToDevice, a->stream));
  }
  return an_sync(a);
}

int soup_analysis_load_system(SoupAnalysis* a, SoupHandle* h) {
  if (!a) return SOUP_E_INVALID;
  if (!h) return afail(a, SOUP_E_INVALID, "soup_analysis_load_system: null system");
  if (sync(h)) return afail(a, SOUP_E_DEVICE, std::string("soup_analysis_load_system: ") + h->err);
  int r = an_reserve_n(a, h->n);
  if (r) return r;
  a->n = h->n;
  a->n_amph = a->n_chain = 0;
  // the system's slot order translated to particle-index order on the system's own stream
  hipLaunchKernelGGL(k_to_index<float4>, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->b.pos, h->b.id, a->pos, h->n);
  hipLaunchKernelGGL(k_bs_to_index, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->b.bs, h->b.id, a->bs, h->n);
  if (sync(h)) return afail(a, SOUP_E_DEVICE, std::string("soup_analysis_load_system: ") + h->err);
  return 0;
}

int soup_analysis_amphiphiles(SoupAnalysis* a, uint32_t polar_mask, uint32_t* n_amph, uint32_t* n_chain) {
  if (!a) return SOUP_E_INVALID;
  if (!n_amph || !n_chain) return afail(a, SOUP_E_INVALID, "soup_analysis_amphiphiles: null output");
  const uint32_t n = a->n;
  *n_amph = *n_chain = 0;
  a->n_amph = a->n_chain = 0;
  if (n == 0) return 0;
  const dim3 g(blocks(n, 256)), t(256);
  hipLaunchKernelGGL(k_uf_init, g, t, 0, a->stream, a->parent, n);
  hipLaunchKernelGGL(k_an_bond_union, g, t, 0, a->stream, a->bs, n, a->parent);
  hipLaunchKernelGGL(k_uf_flatten, g, t, 0, a->stream, a->parent, n);
  an_fill(a, a->stat, 0u, 4ull * n);
  hipLaunchKernelGGL(k_an_comp_stats, g, t, 0, a->stream, a->pos, a->bs, n, a->parent, polar_mask, a->stat, a->polar_of);
  hipLaunchKernelGGL(k_an_amph_flag, g, t, 0, a->stream, a->parent, a->stat, n, a->flag, a->clen);
  scan_on(a->stream, a->scan_blocks, a->flag, n, a->rank);
  scan_on(a->stream, a->scan_blocks, a->clen, n, a->coff);
  uint32_t tot[2];
  AHC(a, hipMemcpyAsync(&tot[0], a->rank + n, 4, hipMemcpyDeviceToHost, a->stream));
  AHC(a, hipMemcpyAsync(&tot[1], a->coff + n, 4, hipMemcpyDeviceToHost, a->stream));
  hipLaunchKernelGGL(k_an_amph_write, g, t, 0, a->stream, a->bs, n, a->stat, a->polar_of, a->flag, a->rank, a->coff,
                     a->heads, a->chain_off, a->chains, a->tails);
  int r = an_sync(a);
  if (r) return r;
  a->n_amph = tot[0];
  a->n_chain = tot[1];
  AHC(a, hipMemcpy(a->chain_off + tot[0], &tot[1], 4, hipMemcpyHostToDevice));
  *n_amph = tot[0];
  *n_chain = tot[1];
  return 0;
}

int soup_analysis_amphiphiles_read(SoupAnalysis* a, uint32_t* heads, uint32_t* chain_off, uint32_t* chains,
                                   uint32_t* tail_lengths) {
  if (!a) return SOUP_E_INVALID;
  if (!heads || !chain_off || !chains || !tail_lengths)
    return afail(a, SOUP_E_INVALID, "soup_analysis_amphiphiles_read: null output");
  if (a->n_amph) {
    AHC(a, hipMemcpyAsync(heads, a->heads, 4ull * a->n_amph, hipMemcpyDeviceToHost, a->stream));
    AHC(a, hipMemcpyAsync(tail_lengths, a->tails, 8ull * a->n_amph, hipMemcpyDeviceToHost, a->stream));
  }
  AHC(a, hipMemcpyAsync(chain_off, a->chain_off, 4ull * (a->n_amph + 1), hipMemcpyDeviceToHost, a->stream));
  if (a->n_chain) AHC(a, hipMemcpyAsync(chains, a->chains, 4ull * a->n_chain, hipMemcpyDeviceToHost, a->stream));
  if (a->n_amph == 0) chain_off[0] = 0;
  return an_sync(a);
}

int soup_analysis_clusters(SoupAnalysis* a, const uint32_t* members, uint32_t m, const double box[3], double cutoff,
                           int periodic_z, uint32_t* labels) {
  if (!a) return SOUP_E_INVALID;
  if ((!members || !labels) && m) return afail(a, SOUP_E_INVALID, "soup_analysis_clusters: null members or labels");
  if (!box || !(cutoff > 0.0)) return afail(a, SOUP_E_INVALID, "soup_analysis_clusters: a box and a positive cutoff are required");
  if (m == 0) return 0;
  for (uint32_t k = 0; k < m; ++k)
    if (members[k] >= a->n)
      return afail(a, SOUP_E_INVALID, fmt("soup_analysis_clusters: member %u names particle %u (N=%u)", k, members[k], a->n));
  int r = an_reserve_m(a, m);
  if (r) return r;
  AnGrid gr;
  double dn[3];
  for (int k = 0; k < 3; ++k) {
    if (!(box[k] > 0.0)) return afail(a, SOUP_E_INVALID, "soup_analysis_clusters: box lengths must be positive");
    dn[k] = std::fmax(1.0, std::floor(box[k] / cutoff));
  }
  if (dn[0] * dn[1] * dn[2] > 2.0e8) return afail(a, SOUP_E_INVALID, "soup_analysis_clusters: the cell list exceeds 2e8 cells");
  gr.nx = (int)dn[0];
  gr.ny = (int)dn[1];
  gr.pz = periodic_z ? 1 : 0;
  gr.wx = box[0] / gr.nx;
  gr.wy = box[1] / gr.ny;
  if (periodic_z) {
    gr.nz = (int)dn[2];
    gr.wz = box[2] / gr.nz;
  } else {
    // aggregate.ts:101-105: cells of width cutoff from z = 0, no wrap; the host checks that every
    // member lies in [0, box z), so the cells 0 .. floor(box z / cutoff) cover them
    gr.wz = cutoff;
    gr.nz = (int)std::floor(box[2] / cutoff) + 1;
  }
  const uint32_t nc = (uint32_t)gr.nx * (uint32_t)gr.ny * (uint32_t)gr.nz;
  if ((r = an_reserve_cells(a, nc))) return r;
  const double3 bx = make_double3(box[0], box[1], box[2]);
  AHC(a, hipMemcpyAsync(a->idx, members, 4ull * m, hipMemcpyHostToDevice, a->stream));
  const dim3 g(blocks(m, 256)), t(256);
  hipLaunchKernelGGL(k_an_gather, g, t, 0, a->stream, a->pos, a->idx, m, a->mp);
  an_fill(a, a->ccount, 0u, nc + 1ull);
  hipLaunchKernelGGL(k_an_cell_count, g, t, 0, a->stream, a->mp, m, gr, bx, a->mcell, a->ccount);
  scan_on(a->stream, a->scan_blocks, a->ccount, nc, a->cstart);
  AHC(a, hipMemcpyAsync(a->ccount, a->cstart, 4ull * (nc + 1), hipMemcpyDeviceToDevice, a->stream));
  hipLaunchKernelGGL(k_an_cell_fill, g, t, 0, a->stream, a->mcell, m, a->ccount, a->bucket);
  hipLaunchKernelGGL(k_uf_init, g, t, 0, a->stream, a->mparent, m);
  hipLaunchKernelGGL(k_an_member_union, g, t, 0, a->stream, a->mp, m, gr, bx, cutoff * cutoff, a->mcell, a->cstart,
                     a->bucket, a->mparent);
  hipLaunchKernelGGL(k_uf_flatten, g, t, 0, a->stream, a->mparent, m);
  AHC(a, hipMemcpyAsync(labels, a->mparent, 4ull * m, hipMemcpyDeviceToHost, a->stream));
  return an_sync(a);
}

int soup_analysis_water(SoupAnalysis* a, const uint32_t* members, uint32_t m, uint32_t water_kind, const double box[3],
                        const double centre[3], double cell, double radius, SoupWaterCounts* out) {
  if (!a) return SOUP_E_INVALID;
  if (!out || !box || !centre || (!members && m)) return afail(a, SOUP_E_INVALID, "soup_analysis_water: null argument");
  for (uint32_t k = 0; k < m; ++k)
    if (members[k] >= a->n)
      return afail(a, SOUP_E_INVALID, fmt("soup_analysis_water: member %u names particle %u (N=%u)", k, members[k], a->n));
  AnOcc o;
  int r = an_grid_dims(a, box, cell, radius, &o);
  if (r) return r;
  if (m && (r = an_reserve_m(a, m))) return r;
  const uint32_t nc = (uint32_t)o.nx * (uint32_t)o.ny * (uint32_t)o.nz;
  an_fill8(a, a->occ, nc);
  if (m) {
    AHC(a, hipMemcpyAsync(a->idx, members, 4ull * m, hipMemcpyHostToDevice, a->stream));
    hipLaunchKernelGGL(k_an_gather, dim3(blocks(m, 256)), dim3(256), 0, a->stream, a->pos, a->idx, m, a->mp);
    hipLaunchKernelGGL(k_an_occ_periodic, dim3(blocks(m, 256)), dim3(256), 0, a->stream, a->mp, m, o, a->occ);
  }
  const double3 c = make_double3(centre[0], centre[1], centre[2]);
  an_fill(a, reinterpret_cast<uint32_t*>(a->best), 0u, 8);
  an_fill(a, a->seed, 0xFFFFFFFFu, 1);
  an_fill(a, a->seed + 1, 0u, 3);
  hipLaunchKernelGGL(k_an_seed_max, dim3(blocks(nc, 256)), dim3(256), 0, a->stream, a->occ, nc, o, c, a->best);
  hipLaunchKernelGGL(k_an_seed_first, dim3(blocks(nc, 256)), dim3(256), 0, a->stream, a->occ, nc, o, c, a->best, a->seed);
  uint32_t seed = 0;
  AHC(a, hipMemcpyAsync(&seed, a->seed, 4, hipMemcpyDeviceToHost, a->stream));
  if ((r = an_sync(a))) return r;
  memset(out, 0, sizeof(*out));
  out->dims[0] = (uint32_t)o.nx;
  out->dims[1] = (uint32_t)o.ny;
  out->dims[2] = (uint32_t)o.nz;
  if (seed == 0xFFFFFFFFu)
    return afail(a, SOUP_E_INVALID,
                 "farthestEmptyCellSeed: the grid is fully occupied -- there is no empty cell to seed the periodic flood fill");
  an_flood(a, o, 1);
  hipLaunchKernelGGL(k_an_cell_counts, dim3(blocks(nc, 256)), dim3(256), 0, a->stream, a->occ, nc, a->gparent, a->seed,
                     (const uint8_t*)nullptr, a->best + 1);
  hipLaunchKernelGGL(k_an_water, dim3(blocks(a->n, 256)), dim3(256), 0, a->stream, a->pos, a->n, water_kind, o, a->occ,
                     a->gparent, a->seed, a->seed + 1);
  unsigned long long cnt[3];
  uint32_t wc[3];
  AHC(a, hipMemcpyAsync(cnt, a->best + 1, 24, hipMemcpyDeviceToHost, a->stream));
  AHC(a, hipMemcpyAsync(wc, a->seed + 1, 12, hipMemcpyDeviceToHost, a->stream));
  if ((r = an_sync(a))) return r;
  out->total_empty_cells = cnt[0];
  out->unreached_cells = cnt[1];
  out->reached_empty_cells = cnt[2];
  out->total_water = wc[0];
  out->encapsulated = wc[1];
  out->reached_empty_water = wc[2];
  out->seed_cell = seed;
  return 0;
}

int soup_analysis_cavity(SoupAnalysis* a, const float* pos4, uint32_t m, const double box[3], double cell, double radius,
                         uint64_t* unreached_cells) {
  if (!a) return SOUP_E_INVALID;
  if (!unreached_cells || !box || (!pos4 && m)) return afail(a, SOUP_E_INVALID, "soup_analysis_cavity: null argument");
  AnOcc o;
  int r = an_grid_dims(a, box, cell, radius, &o);
  if (r) return r;
  if (m && (r = an_reserve_m(a, m))) return r;
  const uint32_t nc = (uint32_t)o.nx * (uint32_t)o.ny * (uint32_t)o.nz;
  an_fill8(a, a->occ, nc);
  an_fill8(a, a->outside, nc);
  if (m) {
    AHC(a, hipMemcpyAsync(a->mp, pos4, 16ull * m, hipMemcpyHostToDevice, a->stream));
    hipLaunchKernelGGL(k_an_occ_open, dim3(blocks(m, 256)), dim3(256), 0, a->stream, a->mp, m, o, a->occ);
  }
  an_flood(a, o, 0);
  hipLaunchKernelGGL(k_an_mark_faces, dim3(blocks(nc, 256)), dim3(256), 0, a->stream, a->occ, o, a->gparent, a->outside);
  an_fill(a, reinterpret_cast<uint32_t*>(a->best), 0u, 8);
  hipLaunchKernelGGL(k_an_cell_counts, dim3(blocks(nc, 256)), dim3(256), 0, a->stream, a->occ, nc, a->gparent,
                     (const uint32_t*)nullptr, a->outside, a->best + 1);
  unsigned long long cnt[3];
  AHC(a, hipMemcpyAsync(cnt, a->best + 1, 24, hipMemcpyDeviceToHost, a->stream));
  if ((r = an_sync(a))) return r;
  *unreached_cells = cnt[1];
  return 0;
}

}  // extern "C"
