// soup_kernels.inl -- HIP kernels of the soup Langevin step for gfx950 (included once by
// soup_lib.hip). Each kernel cites the reference WGSL entry point it replaces.
//
// Exact kernels (cell assignment, list builds, integrator, bond MC, constant-pH chain) use
// IEEE f32 division/sqrt and no contraction so their results are bit-identical to the oracle.
// The per-step force kernels are free to reassociate: their results are judged against an FP64
// recomputation.
#pragma once

#include "soup_device.h"

#define WAVE 64

// ================================================================================================
// Small utilities
// ================================================================================================
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ __forceinline__ float wave_sum_f(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
// FP64 sum over the 64 lanes of a wavefront (all lanes active): butterfly inside each row of 16
// lanes through DPP (no LDS round trip), then the four row sums read from lanes 0, 16, 32, 48
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_d(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l), __builtin_amdgcn_readlane(__double2loint(v), l));
}
__device__ __forceinline__ double wave_sum_d_dpp(double v) {
  v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_d<0x141>(v);  // row_half_mirror
  v += dpp_d<0x140>(v);  // row_mirror
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}
// makes this wavefront's own LDS writes visible to its other lanes
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ float uni_f(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }

__device__ __forceinline__ uint32_t wave_max_u(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, WAVE));
  return v;
}
__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, WAVE);
  return v;
}

// ---- exclusive scan of u32 (3 kernels): out[0..m] with out[m] = total -----------------------
#define SCAN_BLOCK 1024
#define SCAN_ITEMS 4
#define SCAN_TILE (SCAN_BLOCK * SCAN_ITEMS)

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
  const uint32_t lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  uint32_t x = v;
  for (int o = 1; o < WAVE; o <<= 1) {
    uint32_t y = (uint32_t)__shfl_up((int)x, o, WAVE);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == WAVE - 1) s_wave[w] = x;
  __syncthreads();
  if (w == 0) {
    uint32_t s = lane < nw ? s_wave[lane] : 0u;
    for (int o = 1; o < WAVE; o <<= 1) {
      uint32_t y = (uint32_t)__shfl_up((int)s, o, WAVE);
      if (lane >= (uint32_t)o) s += y;
    }
    if (lane < nw) s_wave[lane] = s;
  }
  __syncthreads();
  uint32_t base = w > 0 ? s_wave[w - 1] : 0u;
  total = s_wave[nw - 1];
  __syncthreads();
  return base + x - v;
}

__device__ __forceinline__ void scan_reduce_body(const uint32_t* in, uint32_t m, uint32_t* block_sums, const uint32_t blk) {
  __shared__ uint32_t s_wave[SCAN_BLOCK / WAVE];
  uint32_t base = blk * SCAN_TILE + threadIdx.x * SCAN_ITEMS, v = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k)
    if (base + k < m) v += in[base + k];
  uint32_t total;
  (void)block_excl_scan(v, s_wave, total);
  if (threadIdx.x == 0) block_sums[blk] = total;
}
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const uint32_t* in, uint32_t m, uint32_t* block_sums) { scan_reduce_body(in, m, block_sums, blockIdx.x); }

__device__ __forceinline__ void scan_blocks_body(uint32_t* block_sums, uint32_t nb, const uint32_t blk) {
  // single block, exclusive scan in place over nb <= SCAN_TILE entries; block_sums[nb] = total
  __shared__ uint32_t s_wave[SCAN_BLOCK / WAVE];
  uint32_t base = threadIdx.x * SCAN_ITEMS, v[SCAN_ITEMS], sum = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = base + k < nb ? block_sums[base + k] : 0u;
    sum += v[k];
  }
  uint32_t total;
  uint32_t off = block_excl_scan(sum, s_wave, total);
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < nb) block_sums[base + k] = off;
    off += v[k];
  }
  if (threadIdx.x == 0) block_sums[nb] = total;
}
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_blocks(uint32_t* block_sums, uint32_t nb) { scan_blocks_body(block_sums, nb, blockIdx.x); }

// the whole exclusive scan in one workgroup, for m <= SCAN_ONE_MAX: the tiles in order, each a
// coalesced block scan offset by the running total; out[m] = total
#define SCAN_ONE_MAX (SCAN_TILE * 8u)
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_one(const uint32_t* in, uint32_t m, uint32_t* out) {
  __shared__ uint32_t s_wave[SCAN_BLOCK / WAVE];
  uint32_t run = 0;
  for (uint32_t t0 = 0; t0 < m; t0 += SCAN_TILE) {
    const uint32_t base = t0 + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS], sum = 0;
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      v[k] = base + k < m ? in[base + k] : 0u;
      sum += v[k];
    }
    uint32_t total;
    uint32_t off = block_excl_scan(sum, s_wave, total) + run;
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      if (base + k < m) out[base + k] = off;
      off += v[k];
    }
    run += total;
  }
  if (threadIdx.x == 0) out[m] = run;
}

__device__ __forceinline__ void scan_apply_body(const uint32_t* in, uint32_t m, const uint32_t* block_sums, uint32_t nb, uint32_t* out, const uint32_t blk) {
  __shared__ uint32_t s_wave[SCAN_BLOCK / WAVE];
  uint32_t base = blk * SCAN_TILE + threadIdx.x * SCAN_ITEMS, v[SCAN_ITEMS], sum = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = base + k < m ? in[base + k] : 0u;
    sum += v[k];
  }
  uint32_t total;
  uint32_t off = block_excl_scan(sum, s_wave, total) + block_sums[blk];
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < m) out[base + k] = off;
    off += v[k];
  }
  if (blk == 0 && threadIdx.x == 0) out[m] = block_sums[nb];
}
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(const uint32_t* in, uint32_t m, const uint32_t* block_sums,
                                                           uint32_t nb, uint32_t* out) { scan_apply_body(in, m, block_sums, nb, out, blockIdx.x); }

// |minimum-image displacement since the last rebuild|^2 as bits, uncontracted (verlet.wgsl:115-123)
__device__ __forceinline__ uint32_t drift_sq_bits(const SoupConfig& c, float4 xi, float4 x0) {
  const float ex = mi_f(xi.x - x0.x, c.box[0]), ey = mi_f(xi.y - x0.y, c.box[1]), ez = mi_f(xi.z - x0.z, c.box[2]);
  return __float_as_uint(ex * ex + ey * ey + ez * ez);
}

// the list force's neighbour record: position and, in w, the kind with bit 3 set for a charged
// particle (every charge is 0 or chargeDeprotonated), so one 16-byte gather carries both
__device__ __forceinline__ float4 gq_pack(float4 x, float q) {
  return make_float4(x.x, x.y, x.z, (float)(((uint32_t)x.w & (SOUP_MAX_KINDS - 1u)) | (q != 0.0f ? SOUP_MAX_KINDS : 0u)));
}
__global__ __launch_bounds__(256) void k_gq_refresh(SoupConfig c, DevBuf b) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < c.n) b.gq[t] = gq_pack(b.pos[t], b.charge[t]);
}

// ================================================================================================
// Integrator: kick_drift_wrap_main (soup/wgsl/step.wgsl:96-108),
//             kick_thermostat_main (step.wgsl:110-142)
// ================================================================================================
// Also soup_max_drift_main (verlet.wgsl:115-123) on the new positions, which are the ones the force
// of this step sees: one atomicMax per workgroup. On a rebuild step the rebuild clears the maximum
// afterwards, as the reference's rebuild does before its own drift pass.
// the first half-kick, the drift and the periodic wrap of one mobile particle (kick_drift_wrap_main)
__device__ __forceinline__ float4 kick_drift_one(const SoupConfig& c, DevBuf& b, uint32_t t, float4 x, float4 v, float4 f) {
  const float hdt = 0.5f * c.dt;
  v.x = v.x + hdt * f.x;
  v.y = v.y + hdt * f.y;
  v.z = v.z + hdt * f.z;
  b.vel[t] = make_float4(v.x, v.y, v.z, 0.f);
  float px = x.x + c.dt * v.x, py = x.y + c.dt * v.y, pz = x.z + c.dt * v.z;
  px = px - floorf(px / c.box[0]) * c.box[0];
  py = py - floorf(py / c.box[1]) * c.box[1];
  pz = pz - floorf(pz / c.box[2]) * c.box[2];
  x = make_float4(px, py, pz, x.w);
  b.pos[t] = x;
  return x;
}

// the second half-kick and the Langevin thermostat of one mobile particle (kick_thermostat_main):
// the new velocity, with the particle's thermostat stream advanced by six draws
__device__ __forceinline__ float4 kick_thermo_one(const SoupConfig& c, DevBuf& b, uint32_t t, float4 v, float4 f) {
  const float hdt = 0.5f * c.dt;
  uint32_t s = b.trng[t];
  float noise[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    s = pcg_soup(s);
    float u1 = fmaxf(uniform01(s), 1e-7f);
    s = pcg_soup(s);
    float u2 = uniform01(s);
    // Box-Muller: log and cos evaluated in double on the f32 argument, rounded to f32
    float l = (float)log((double)u1);
    float cs = (float)cos((double)(6.28318531f * u2));
    noise[a] = sqrtf(-2.0f * l) * cs;
  }
  b.trng[t] = s;
  float sigma_v = sqrtf(2.0f * c.gamma * c.kT * c.dt);
  float gdt = c.gamma * c.dt;
  float vx = v.x + hdt * f.x, vy = v.y + hdt * f.y, vz = v.z + hdt * f.z;
  vx = vx - gdt * vx + sigma_v * noise[0];
  vy = vy - gdt * vy + sigma_v * noise[1];
  vz = vz - gdt * vz + sigma_v * noise[2];
  return make_float4(vx, vy, vz, 0.f);
}

// one atomicMax per workgroup of the drift maximum; the maximum only grows between rebuilds, so
// most workgroups see a larger value and skip the atomic
__device__ __forceinline__ void drift_max_block(DevBuf& b, uint32_t bits, uint32_t* s_max) {
  bits = wave_max_u(bits);
  if ((threadIdx.x & (WAVE - 1)) == 0) s_max[threadIdx.x / WAVE] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = 0;
    for (uint32_t w = 0; w < blockDim.x / WAVE; ++w) m = max(m, s_max[w]);
    if (m > __hip_atomic_load(&b.flags[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&b.flags[2], m);
  }
}

__device__ __forceinline__ void kick_drift_body(SoupConfig c, DevBuf b, const uint32_t blk) {
  __shared__ uint32_t s_max[256 / WAVE];
  const uint32_t t = blk * blockDim.x + threadIdx.x;
  uint32_t bits = 0;
  if (t < c.n) {
    float4 x = b.pos[t];
    if (b.frozen[t] != 0u) b.vel[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    else x = kick_drift_one(c, b, t, x, b.vel[t], b.force[t]);
    b.gq[t] = gq_pack(x, b.charge[t]);
    bits = drift_sq_bits(c, x, b.pos_at_rebuild[t]);
  }
  drift_max_block(b, bits, s_max);
}
__global__ __launch_bounds__(256) void k_kick_drift(SoupConfig c, DevBuf b) { kick_drift_body(c, b, blockIdx.x); }

__device__ __forceinline__ void kick_thermo_body(SoupConfig c, DevBuf b, const uint32_t blk) {
  uint32_t t = blk * blockDim.x + threadIdx.x;
  if (t >= c.n) return;
  if (b.frozen[t] != 0u) {
    b.vel[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  b.vel[t] = kick_thermo_one(c, b, t, b.vel[t], b.force[t]);
}
__global__ __launch_bounds__(256) void k_kick_thermo(SoupConfig c, DevBuf b) { kick_thermo_body(c, b, blockIdx.x); }

// kick_thermostat_main of the step that ended fused with kick_drift_wrap_main of the next, for two
// steps that follow each other with nothing in between (inside a replayed step period): both read
// the same force, and the velocity passes between them in registers, so the result is the two
// kernels' bit for bit
__device__ __forceinline__ void thermo_drift_body(SoupConfig c, DevBuf b, const uint32_t blk) {
  __shared__ uint32_t s_max[256 / WAVE];
  const uint32_t t = blk * blockDim.x + threadIdx.x;
  uint32_t bits = 0;
  if (t < c.n) {
    float4 x = b.pos[t];
    if (b.frozen[t] != 0u) {
      b.vel[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      const float4 f = b.force[t];
      x = kick_drift_one(c, b, t, x, kick_thermo_one(c, b, t, b.vel[t], f), f);
    }
    b.gq[t] = gq_pack(x, b.charge[t]);
    bits = drift_sq_bits(c, x, b.pos_at_rebuild[t]);
  }
  drift_max_block(b, bits, s_max);
}
__global__ __launch_bounds__(256) void k_thermo_drift(SoupConfig c, DevBuf b) { thermo_drift_body(c, b, blockIdx.x); }

// soup_max_drift_main (verlet.wgsl:115-123): wave max, one atomic per wave.
__global__ __launch_bounds__(256) void k_max_drift(SoupConfig c, DevBuf b) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t bits = t < c.n ? drift_sq_bits(c, b.pos[t], b.pos_at_rebuild[t]) : 0u;
  bits = wave_max_u(bits);
  if ((threadIdx.x & (WAVE - 1)) == 0 && bits != 0u) atomicMax(&b.flags[2], bits);
}

__global__ __launch_bounds__(256) void k_nonfinite(SoupConfig c, DevBuf b) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t bp = 0, bv = 0;
  if (t < c.n) {
    float4 x = b.pos[t], v = b.vel[t];
    bp = (uint32_t)nonfinite_bits(x.x) + (uint32_t)nonfinite_bits(x.y) + (uint32_t)nonfinite_bits(x.z);
    bv = (uint32_t)nonfinite_bits(v.x) + (uint32_t)nonfinite_bits(v.y) + (uint32_t)nonfinite_bits(v.z);
  }
  bp = wave_sum_u(bp);
  bv = wave_sum_u(bv);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    if (bp) atomicAdd(&b.flags[3], bp);
    if (bv) atomicAdd(&b.flags[4], bv);
  }
}

// soup_relax_step_main (soup/wgsl/relax.wgsl:49-72)
__global__ __launch_bounds__(256) void k_relax_step(SoupConfig c, DevBuf b, float d) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.n || b.frozen[t] != 0u) return;
  float4 f = b.force[t];
  float a = fmaxf(fmaxf(fabsf(f.x), fabsf(f.y)), fabsf(f.z));
  if (!(a > 0.0f) || nonfinite_bits(a)) return;
  float gx = f.x / a, gy = f.y / a, gz = f.z / a;
  float l = len3(gx, gy, gz);
  float4 x = b.pos[t];
  float px = x.x + (gx / l) * d, py = x.y + (gy / l) * d, pz = x.z + (gz / l) * d;
  px = px - floorf(px / c.box[0]) * c.box[0];
  py = py - floorf(py / c.box[1]) * c.box[1];
  pz = pz - floorf(pz / c.box[2]) * c.box[2];
  b.pos[t] = make_float4(px, py, pz, x.w);
  b.gq[t] = gq_pack(b.pos[t], b.charge[t]);
}

// ================================================================================================
// Rebuild: counting sort by (cell, x, particle index) + permutation of all state
// (replaces clear/count/prefix/fill_main, engine/wgsl/neighbor.wgsl:26-64)
// ================================================================================================
// The slots are in cell order as of the last rebuild, so the lanes of a wavefront mostly share a
// few cells: one atomic per distinct cell of the wavefront reserves that cell's ranks for all its
// lanes (the order inside a cell is fixed afterwards by x and particle index, so any distinct ranks do).
__device__ __forceinline__ void cell_count_body(SoupConfig c, DevBuf b, const uint32_t blk) {
  const uint32_t t = blk * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & (WAVE - 1);
  const bool live = t < c.n;
  const uint32_t cell = live ? cell_of(c, b.pos[t]) : 0u;
  if (live) b.cell_tmp[t] = cell;
  uint64_t todo = __builtin_amdgcn_ballot_w64(live);
  const uint64_t lower = (1ull << lane) - 1ull;
  uint32_t rank = 0;
  while (todo) {
    const uint32_t leader = (uint32_t)__builtin_ctzll(todo);
    const uint32_t lc = uni((uint32_t)__builtin_amdgcn_readlane((int)cell, (int)leader));
    const uint64_t same = __builtin_amdgcn_ballot_w64(live && cell == lc);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&b.cell_count[lc], (uint32_t)__popcll(same));
    base = uni((uint32_t)__builtin_amdgcn_readlane((int)base, (int)leader));
    if (live && cell == lc) rank = base + (uint32_t)__popcll(same & lower);
    todo &= ~same;
  }
  if (live) b.cell_rank[t] = rank;
}
__global__ __launch_bounds__(256) void k_cell_count(SoupConfig c, DevBuf b) { cell_count_body(c, b, blockIdx.x); }

// The order inside a cell: ascending x, ties by particle index (a fixed order; the Verlet build trims
// the two end cells of each stencil row to the particle's x-chord with it). The reference orders a
// cell by particle index; the list readback and the bond proposals restore that walk order by key.
// x as an unsigned key in the order of the floats (a total order: a non-finite x, which the chunk-end
// guard reports, still gets a rank of its own)
__device__ __forceinline__ uint32_t x_order_key(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
// unordered fill of the sort keys (x key, particle index) into the cell segment (keys: n entries of
// scratch), then each particle's rank among its segment's keys
__device__ __forceinline__ void cell_fill_body(SoupConfig c, DevBuf b, uint2* keys, const uint32_t blk) {
  uint32_t t = blk * blockDim.x + threadIdx.x;
  if (t >= c.n) return;
  keys[b.cell_start[b.cell_tmp[t]] + b.cell_rank[t]] = make_uint2(x_order_key(b.pos[t].x), b.id[t]);
}
__global__ __launch_bounds__(256) void k_cell_fill(SoupConfig c, DevBuf b, uint2* keys) { cell_fill_body(c, b, keys, blockIdx.x); }

__device__ __forceinline__ void cell_rank_body(SoupConfig c, DevBuf b, const uint2* keys, uint32_t* new_slot_of_old,
                                               const uint32_t blk) {
  uint32_t t = blk * blockDim.x + threadIdx.x;
  if (t >= c.n) return;
  const uint32_t cell = b.cell_tmp[t];
  const uint32_t s0 = b.cell_start[cell], s1 = b.cell_start[cell + 1];
  const uint2 me = keys[s0 + b.cell_rank[t]];
  uint32_t r = 0;
  for (uint32_t k = s0; k < s1; ++k) {
    const uint2 o = keys[k];
    r += (o.x < me.x || (o.x == me.x && o.y < me.y)) ? 1u : 0u;
  }
  new_slot_of_old[t] = s0 + r;
}
__global__ __launch_bounds__(256) void k_cell_rank(SoupConfig c, DevBuf b, const uint2* keys, uint32_t* new_slot_of_old) {
  cell_rank_body(c, b, keys, new_slot_of_old, blockIdx.x);
}

// gathers every per-particle array into the new slot order (out = the alternate buffers)
__device__ __forceinline__ void permute_body(SoupConfig c, DevBuf b, DevBuf o, const uint32_t* new_slot_of_old, const uint32_t blk) {
  uint32_t t = blk * blockDim.x + threadIdx.x;
  if (t >= c.n) return;
  uint32_t ns = new_slot_of_old[t];
  float4 x = b.pos[t];
  o.pos[ns] = x;
  o.pos_at_rebuild[ns] = x;  // soup_snapshot_positions_main
  o.vel[ns] = b.vel[t];
  o.force[ns] = b.force[t];  // a rebuild without a force pass (growBoxTo) keeps F(x) per particle
  const float q = b.charge[t];
  o.charge[ns] = q;
  o.gq[ns] = gq_pack(x, q);
  o.bs[ns] = b.bs[t];
  o.clink[ns] = b.clink[t];
  o.cheld[ns] = b.cheld[t];
  o.frozen[ns] = b.frozen[t];
  o.brng[ns] = b.brng[t];
  o.trng[ns] = b.trng[t];
  uint32_t id = b.id[t];
  o.id[ns] = id;
  o.cell_of[ns] = b.cell_tmp[t];
  b.inv[id] = ns;
  // the head marks of the new order start clear (the DH build sets them), and the drift maximum
  // restarts from zero
  o.head_of_slot[ns] = SOUP_NONE;
  o.head_flag[ns] = 0u;
  if (t == 0u) b.flags[2] = 0u;
}
__global__ __launch_bounds__(256) void k_permute(SoupConfig c, DevBuf b, DevBuf o, const uint32_t* new_slot_of_old) { permute_body(c, b, o, new_slot_of_old, blockIdx.x); }

__device__ __forceinline__ void head_slots_body(DevBuf b, uint32_t nh, const uint32_t blk) {
  uint32_t h = blk * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  uint32_t u = b.inv[b.head_id[h]];
  b.head_slot[h] = u;
  const float4 x = b.pos[u];
  b.head_q[h] = make_float4(x.x, x.y, x.z, b.charge[u]);
  b.head_of_slot[u] = h;
  b.head_flag[u] = 1u;
}
__global__ __launch_bounds__(256) void k_head_slots(DevBuf b, uint32_t nh) { head_slots_body(b, nh, blockIdx.x); }

// heads in slot order: the far-force and pair-table kernels walk rows in this spatial order
// head_q from the current positions and charges (every force evaluation)
__device__ __forceinline__ void head_refresh_body(DevBuf b, uint32_t nh, const uint32_t blk) {
  const uint32_t h = blk * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  const uint32_t u = b.head_slot[h];
  const float4 x = b.pos[u];
  b.head_q[h] = make_float4(x.x, x.y, x.z, b.charge[u]);
}
__global__ __launch_bounds__(256) void k_head_refresh(DevBuf b, uint32_t nh) { head_refresh_body(b, nh, blockIdx.x); }
__device__ __forceinline__ void head_rows_body(SoupConfig c, DevBuf b, const uint32_t blk) {
  uint32_t t = blk * blockDim.x + threadIdx.x;
  if (t < c.n && b.head_flag[t]) b.head_rows[b.head_pos[t]] = b.head_of_slot[t];
}
__global__ __launch_bounds__(256) void k_head_rows(SoupConfig c, DevBuf b) { head_rows_body(c, b, blockIdx.x); }

// largest WCA cut of any kind pair (also the largest bond-formation and catalyst distance)
__device__ __forceinline__ float wca_reach(const SoupConfig& c) {
  float wmax = 0.0f;
  for (uint32_t k = 0; k < SOUP_MAX_KINDS; ++k) wmax = fmaxf(wmax, c.radius[k]);
  return c.wca_factor * c.sigma * wmax;
}
// no term acts at or beyond max(WCA cuts, attraction shell, DH near/far split): partners there are skin
__device__ __forceinline__ float force_reach(const SoupConfig& c) {
  return fmaxf(fmaxf(wca_reach(c), c.wca_factor * c.b_tt + c.wc), c.es_split);
}
// Near/far split of a list row. A row keeps the entries closer than list_split at the rebuild in walk
// order from its front (the near part) and the others in reverse walk order from its back (the far
// part). A pair that acts now was within reach + 2 dmax at the rebuild (dmax = the largest
// displacement since), so while 2 dmax <= skin/2 the near part holds every acting pair; while the
// list is valid (dmax <= skin/2) it also holds, in walk order, every bond-formation candidate and
// catalyst (within the largest WCA cut). The 0.002 sigma cover the f32 rounding of both distances and
// of the measured drift.
__device__ __forceinline__ float list_split(const SoupConfig& c) {
  return fmaxf(force_reach(c) * 1.00005f + 0.5f * c.skin, wca_reach(c) + c.skin) + 0.002f;
}
// the force reads only the near part while the measured drift (squared, f32 bits) allows it
__device__ __forceinline__ bool near_only(const SoupConfig& c, uint32_t drift_sq_bits) {
  const float dm = sqrtf(__uint_as_float(drift_sq_bits));
  return 2.0f * dm <= list_split(c) - 0.002f - force_reach(c) * 1.00005f;
}
// entry s of a row read as near part then far part (n = near entries)
__device__ __forceinline__ uint32_t row_index(uint32_t s, uint32_t n, uint32_t cap) { return s < n ? s : cap + n - 1u - s; }
// cell of list entry u of particle t in the walk order: the cell offset (dz, dy, dx), each wrapped
// into -R..R, as one number
__device__ __forceinline__ uint32_t walk_cell(const SoupConfig& c, const uint32_t* cell_of, uint32_t ct, uint32_t u) {
  const int R = (int)c.walk_radius, side = 2 * R + 1;
  uint32_t a = ct, q = cell_of[u];
  int key = 0, scale = 1;
  for (int k = 0; k < 3; ++k) {
    const int D = (int)c.dims[k];
    int v = (int)(q % (uint32_t)D) - (int)(a % (uint32_t)D);
    if (v > R) v -= D;
    if (v < -R) v += D;
    key += (v + R) * scale;
    scale *= side;
    a /= (uint32_t)D;
    q /= (uint32_t)D;
  }
  return (uint32_t)key;
}
// position of list entry u (particle index j) in the reference's walk order: its cell, then the
// particle index (the reference orders each cell by index)
__device__ __forceinline__ uint64_t walk_key(const SoupConfig& c, const uint32_t* cell_of, uint32_t ct, uint32_t u, uint32_t j) {
  return ((uint64_t)walk_cell(c, cell_of, ct, u) << 32) | j;
}

// ================================================================================================
// Verlet list: soup_build_verlet_list_main (soup/wgsl/verlet.wgsl:53-92), one wavefront per particle.
// Canonical order: cells z-major (dz, dy, dx in -R..R), ascending x inside a cell. Lane r
// first sets up row r = (dz, dy): cells lying entirely beyond listRange + 0.01 sigma of the particle
// are dropped (they hold no entry, so the list is unchanged), which leaves one interval of dx around
// dx = 0 (from the staging, the interval's two end cells are trimmed to the x-chord of the culling
// sphere as well); a scan over the lanes lays the kept ranges end to end, and the wavefront sweeps that one
// dense candidate range, VL_BATCH chunks of 64 at a time (each lane carries its row from chunk to
// chunk, since its index only grows). Ballots compact the hits in order: the near ones (list_split)
// from the front of the row, the far ones from its back. Needs (2R + 1)^2 <= 64 rows.
//
// Two ways to reach the candidates:
// - k_verlet_cells (the rebuild): one workgroup per cell. It first stages the (2R + 1)^3 cells around
//   its cell in LDS, in canonical order, each position with its image shift already added and its
//   slot in w; then each wavefront builds the rows of the cell's particles one after the other,
//   reading candidates from LDS by the flat index. A cell whose stencil holds more than VL_STAGE
//   beads (or the configuration's smaller list_stage) builds its rows from the particle arrays
//   instead, as below.
// - verlet_row_global: candidates gathered from the particle arrays. Slots are cell-major with x
//   fastest, so the kept cells of a row are one contiguous slot range (two when the row wraps in x).
// ================================================================================================
#ifndef VL_BATCH
#define VL_BATCH 4u  // Verlet-build candidate chunks of 64 whose loads are issued together
#endif
#ifndef VL_CELLS
#define VL_CELLS 1u  // 1: the rebuild runs k_verlet_cells, 0: k_verlet_build (one wavefront per particle)
#endif
#ifndef VL_STAGE
#define VL_STAGE 4608u  // stencil beads one workgroup of k_verlet_cells stages in LDS (16 B each)
#endif
#ifndef VL_WAVES
#define VL_WAVES 16u  // wavefronts per workgroup of k_verlet_cells (two workgroups fill a CU)
#endif
#define VL_PAD (VL_BATCH * WAVE)  // staged padding the sweep may read past the stencil
static_assert(VL_STAGE >= 2u * VL_WAVES * WAVE + VL_WAVES * WAVE / 4u, "the gather path's tables alias the staging");

// Margin of the cheap list test. The squared distance taken through the stencil's known image shift
// differs from the reference's exact f32 test (minimum image by rint, then the length) by at most
// about 6 sqrt(3) L u (r + 1) + 8 u r^2 (sigma^2; u = 2^-24, L the longest box edge, r the range):
// 2.7e-4 at L = 76. Candidates nearer than r - margin or farther than r + margin are decided by the
// cheap test, the band between by the exact one; the margin (in distance) is at least 10x that bound
// for any box, and it also covers the f32 rounding of the cell gaps the row set-up culls with.
__device__ __forceinline__ float list_margin(const SoupConfig& c) {
  return fmaxf(0.001f, 4e-6f * fmaxf(c.box[0], fmaxf(c.box[1], c.box[2])));
}

// the row's counts; past the capacity (an overflow, reported) the two parts overlap
__device__ __forceinline__ void verlet_finish(const DevBuf& b, uint32_t t, uint32_t cn, uint32_t cf, uint32_t cap) {
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    const uint32_t count = cn + cf;
    b.vcount[t] = count < cap ? count : cap;
    b.vnear[t] = cn < cap ? cn : cap;
    if (count > cap) b.flags[0] = 1u;
  }
}

// the hits of one chunk of 64 candidates, compacted in lane order into the row
__device__ __forceinline__ void verlet_emit(uint32_t* row, uint32_t cap, uint64_t below, bool hit, bool near, uint32_t kk,
                                            uint32_t& cn, uint32_t& cf) {
  const uint64_t mn = __ballot(hit && near), mf = __ballot(hit && !near);
  if (hit) {
    // past the capacity (an overflow, reported) the two parts overlap; every index stays in the row
    if (near) {
      const uint32_t p = cn + (uint32_t)__popcll(mn & below);
      if (p < cap) row[p] = kk;
    } else {
      const uint32_t p = cf + (uint32_t)__popcll(mf & below);
      if (p < cap) row[cap - 1u - p] = kk;
    }
  }
  cn += (uint32_t)__popcll(mn);
  cf += (uint32_t)__popcll(mf);
}

// particle t's row from candidates gathered in the particle arrays; s_end, s_seg and s_sh are this
// wavefront's 64-entry LDS tables
__device__ __forceinline__ void verlet_row_global(const SoupConfig& c, const DevBuf& b, const uint32_t t, uint32_t* s_end,
                                                  uint4* s_seg, float4* s_sh) {
  const uint32_t lane = threadIdx.x & (WAVE - 1);
  wave_lds_sync();  // the tables may still hold the wavefront's previous row
  const float4 xv = b.pos[t];
  const float4 xi = make_float4(uni_f(xv.x), uni_f(xv.y), uni_f(xv.z), 0.0f);
  const uint32_t cell = uni(b.cell_of[t]);
  const int D0 = (int)c.dims[0], D1 = (int)c.dims[1], D2 = (int)c.dims[2];
  const int cx = (int)(cell % c.dims[0]), cy = (int)((cell / c.dims[0]) % c.dims[1]), cz = (int)(cell / (c.dims[0] * c.dims[1]));
  const float w0 = c.box[0] / (float)D0, w1 = c.box[1] / (float)D1, w2 = c.box[2] / (float)D2;
  const float lo0 = xi.x - (float)cx * w0, lo1 = xi.y - (float)cy * w1, lo2 = xi.z - (float)cz * w2;
  const float hi0 = w0 - lo0, hi1 = w1 - lo1, hi2 = w2 - lo2;
  const float lr = c.list_range;
  const float mg = list_margin(c);
  const float cull = (lr + mg) * (lr + mg);
  const float sure = (lr - mg) * (lr - mg);  // inside this the exact f32 test cannot fail
  const int R = (int)c.walk_radius, side = 2 * R + 1;
  const uint32_t nrows = (uint32_t)(side * side);
  // ---- row set-up, lane r = row (dz, dy) in canonical order ----
  uint32_t k0a = 0, lena = 0, k0b = 0, lenb = 0;
  float sxa = 0.0f, sxb = 0.0f, sy = 0.0f, sz = 0.0f;
  if (lane < nrows) {
    const int dz = (int)lane / side - R, dy = (int)lane % side - R;
    const float gz = fmaxf(dz > 0 ? (float)(dz - 1) * w2 + hi2 : (dz < 0 ? (float)(-dz - 1) * w2 + lo2 : 0.0f), 0.0f);
    const float gy = fmaxf(dy > 0 ? (float)(dy - 1) * w1 + hi1 : (dy < 0 ? (float)(-dy - 1) * w1 + lo1 : 0.0f), 0.0f);
    const float gyz = gy * gy + gz * gz;
    if (gyz <= cull) {
      int zz = cz + dz, yy = cy + dy;
      if (zz < 0) { zz += D2; sz = -c.box[2]; } else if (zz >= D2) { zz -= D2; sz = c.box[2]; }
      if (yy < 0) { yy += D1; sy = -c.box[1]; } else if (yy >= D1) { yy -= D1; sy = c.box[1]; }
      // kept x offsets form one interval [dxl, dxh] around dx = 0 (whose gap is 0)
      int dxl = 0, dxh = 0;
      for (int dx = 1; dx <= R; ++dx) {
        const float gh = (float)(dx - 1) * w0 + hi0, gl = (float)(dx - 1) * w0 + lo0;
        if (fmaxf(gh, 0.0f) * fmaxf(gh, 0.0f) + gyz <= cull) dxh = dx;
        if (fmaxf(gl, 0.0f) * fmaxf(gl, 0.0f) + gyz <= cull) dxl = -dx;
      }
      const uint32_t rowbase = c.dims[0] * ((uint32_t)yy + c.dims[1] * (uint32_t)zz);
      const int xa = cx + dxl, xb = cx + dxh;
      // at most one end of the interval wraps; the wrapped cells come first in canonical order
      int la, ha, lb = 0, hb = -1;
      if (xa < 0) {
        la = xa + D0; ha = D0 - 1; sxa = -c.box[0]; lb = 0; hb = xb;
      } else if (xb >= D0) {
        la = xa; ha = D0 - 1; lb = 0; hb = xb - D0; sxb = c.box[0];
      } else {
        la = xa; ha = xb;
      }
      k0a = b.cell_start[rowbase + (uint32_t)la];
      lena = b.cell_start[rowbase + (uint32_t)ha + 1u] - k0a;
      if (hb >= lb) {
        k0b = b.cell_start[rowbase + (uint32_t)lb];
        lenb = b.cell_start[rowbase + (uint32_t)hb + 1u] - k0b;
      }
    }
  }
  uint32_t incl = lena + lenb;
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)incl, o, WAVE);
    if (lane >= (uint32_t)o) incl += y;
  }
  s_end[lane] = incl;
  s_seg[lane] = make_uint4(k0a, lena, k0b, 0u);
  s_sh[lane] = make_float4(sxa, sxb, sy, sz);
  const uint32_t total = uni((uint32_t)__shfl((int)incl, WAVE - 1, WAVE));
  wave_lds_sync();
  // ---- dense sweep ----
  const uint32_t cap = c.list_cap;
  uint32_t* row = b.vl + (size_t)t * vl_stride(c);
  const uint64_t below = (1ull << lane) - 1ull;
  const float split = list_split(c);
  const float split2 = split * split;
  uint32_t cn = 0, cf = 0;  // near and far hits so far
  // each lane's row and that row's start: a lane's index only grows (by 64 a chunk), so the row is
  // carried from chunk to chunk and advanced past the rows that end at or before the index
  uint32_t r = 0, rstart = 0, rend = s_end[0];
  for (uint32_t base = 0; base < total; base += VL_BATCH * WAVE) {
    uint32_t kk[VL_BATCH];
    float3 sh[VL_BATCH];
    float4 xj[VL_BATCH];
#pragma unroll
    for (uint32_t k = 0; k < VL_BATCH; ++k) {
      const uint32_t f = base + k * WAVE + lane;
      kk[k] = t;
      sh[k] = make_float3(0.0f, 0.0f, 0.0f);
      if (f < total) {
        while (rend <= f) {  // f < total = s_end[nrows - 1], so r stays below nrows
          rstart = rend;
          rend = s_end[++r];
        }
        const uint32_t loc = f - rstart;
        const uint4 sg = s_seg[r];
        const float4 sv = s_sh[r];
        const bool in_a = loc < sg.y;
        kk[k] = in_a ? sg.x + loc : sg.z + (loc - sg.y);
        sh[k] = make_float3(in_a ? sv.x : sv.y, sv.z, sv.w);
      }
      xj[k] = b.pos[kk[k]];
    }
#pragma unroll
    for (uint32_t k = 0; k < VL_BATCH; ++k) {
      bool hit = false, near = false;
      if (base + k * WAVE + lane < total && kk[k] != t) {
        // cheap reject through the known image shift, then the reference's exact test
        const float ax = (xj[k].x + sh[k].x) - xi.x, ay = (xj[k].y + sh[k].y) - xi.y, az = (xj[k].z + sh[k].z) - xi.z;
        const float a2 = ax * ax + ay * ay + az * az;
        hit = a2 <= sure;
        if (!hit && a2 <= cull) hit = dist_f(c, xi, xj[k]) < lr;
        near = a2 < split2;
      }
      verlet_emit(row, cap, below, hit, near, kk[k], cn, cf);
    }
  }
  verlet_finish(b, t, cn, cf, cap);
}

__device__ __forceinline__ void verlet_build_body(SoupConfig c, DevBuf b, const uint32_t blk) {
  __shared__ uint32_t s_end[256 / WAVE][WAVE];  // inclusive prefix of kept row lengths
  __shared__ uint4 s_seg[256 / WAVE][WAVE];     // first slot of range a, its length, first slot of range b
  __shared__ float4 s_sh[256 / WAVE][WAVE];     // image shifts: x of range a, x of range b, y, z
  const uint32_t w = uni(threadIdx.x / WAVE);
  const uint32_t t = uni(blk * (blockDim.x / WAVE) + w);
  if (t >= c.n) return;
  verlet_row_global(c, b, t, s_end[w], s_seg[w], s_sh[w]);
}
__global__ __launch_bounds__(256) void k_verlet_build(SoupConfig c, DevBuf b) { verlet_build_body(c, b, xcd_block()); }

// what the rows of one cell's particles share: cell widths, thresholds, and per lane r < (2R + 1)^2
// the row's gap terms (gap along z = max(az + sz * lo_z, 0) with lo_z the particle's offset in its
// cell, likewise y)
struct VlShared {
  float w0, lo_base0, lo_base1, lo_base2;  // cell width in x; the cell's lower corner
  float cull, sure, split2, lr;
  float az, sz, ay, sy;  // this lane's row terms
  uint32_t k_row;        // this lane's first stencil cell (dx = -R)
  bool row_ok;           // lane < (2R + 1)^2
};

__device__ __forceinline__ VlShared verlet_shared(const SoupConfig& c, int cx, int cy, int cz) {
  VlShared v;
  const uint32_t lane = threadIdx.x & (WAVE - 1);
  const float w0 = c.box[0] / (float)c.dims[0], w1 = c.box[1] / (float)c.dims[1], w2 = c.box[2] / (float)c.dims[2];
  v.w0 = w0;
  v.lo_base0 = (float)cx * w0;
  v.lo_base1 = (float)cy * w1;
  v.lo_base2 = (float)cz * w2;
  v.lr = c.list_range;
  const float mg = list_margin(c);
  v.cull = (v.lr + mg) * (v.lr + mg);
  v.sure = (v.lr - mg) * (v.lr - mg);  // inside this the exact f32 test cannot fail
  const float split = list_split(c);
  v.split2 = split * split;
  const int R = (int)c.walk_radius, side = 2 * R + 1;
  // up to 32 rows (R <= 2): lanes r and r + 32 both hold row r, and split the row's two searches
  const uint32_t nrow = (uint32_t)(side * side);
  const uint32_t rl = nrow <= WAVE / 2 ? (lane & (WAVE / 2 - 1)) : lane;
  v.row_ok = rl < nrow;
  const int dz = (int)rl / side - R, dy = (int)rl % side - R;
  // gap to the row's cells: dz > 0: dz w - lo, dz < 0: lo - (dz + 1) w... written as a + s lo
  v.az = dz > 0 ? (float)dz * w2 : (dz < 0 ? (float)(-dz - 1) * w2 : 0.0f);
  v.sz = dz > 0 ? -1.0f : (dz < 0 ? 1.0f : 0.0f);
  v.ay = dy > 0 ? (float)dy * w1 : (dy < 0 ? (float)(-dy - 1) * w1 : 0.0f);
  v.sy = dy > 0 ? -1.0f : (dy < 0 ? 1.0f : 0.0f);
  v.k_row = rl * (uint32_t)side;
  return v;
}

// first staged index in [a, e) whose x exceeds thr (past_equal) or is not below it; the staged x
// ascend inside one cell
__device__ __forceinline__ uint32_t staged_x_bound(const float4* s_pos, uint32_t a, uint32_t e, float thr, bool past_equal) {
  uint32_t n = e - a;
  while (n != 0u) {
    const uint32_t half = n >> 1;
    const float xv = s_pos[a + half].x;
    if (xv < thr || (past_equal && xv == thr)) {
      a += half + 1u;
      n -= half + 1u;
    } else {
      n = half;
    }
  }
  return a;
}

// particle t's row from the stencil staged in LDS (s_pos, with s_cpre the exclusive prefix of the
// stencil cells' counts in canonical order).
// Flat candidate f of the row set lies in row r when end[r - 1] <= f < end[r] and sits at staged
// index f + beg[r] - end[r - 1]. Lane r holds end[r] and beg[r]; the wavefront walks the row
// boundaries that fall inside each chunk of 64 with uniform reads of those lanes, so a lane's staged
// index needs no LDS lookup. The cell culling only has to be conservative (a dropped cell lies
// beyond listRange + the margin), so its gaps are not computed in any particular rounding.
__device__ __forceinline__ void verlet_row_staged(const SoupConfig& c, const DevBuf& b, const uint32_t t, const VlShared& v,
                                                  const float4* s_pos, const uint32_t* s_cpre) {
  const uint32_t lane = threadIdx.x & (WAVE - 1);
  const float4 xv = b.pos[t];
  const float4 xi = make_float4(uni_f(xv.x), uni_f(xv.y), uni_f(xv.z), 0.0f);
  const float lo0 = xi.x - v.lo_base0, lo1 = xi.y - v.lo_base1, lo2 = xi.z - v.lo_base2;
  const float hi0 = v.w0 - lo0;
  const float cull = v.cull, sure = v.sure, lr = v.lr;
  const int R = (int)c.walk_radius;
  // ---- row set-up, lane r = row (dz, dy) in canonical order (and lane r + 32 too when there are at
  // most 32 rows): its kept cells dx = dxl..dxh are the stencil cells r * side + R + dxl ..
  // r * side + R + dxh, one contiguous range of the staging ----
  uint32_t beg = 0, len = 0;
  const float gz = fmaxf(v.az + v.sz * lo2, 0.0f), gy = fmaxf(v.ay + v.sy * lo1, 0.0f);
  const float gyz = gy * gy + gz * gz;
  const bool pair_lanes = (uint32_t)((2 * R + 1) * (2 * R + 1)) <= WAVE / 2;
  const bool live = v.row_ok && gyz <= cull;
  uint32_t lo_at = 0, hi_at = 0;  // the trimmed range [lo_at, hi_at) of the row's kept cells
  if (live) {
    int dxl = 0, dxh = 0;
    for (int dx = 1; dx <= R; ++dx) {
      const float gh = fmaxf((float)(dx - 1) * v.w0 + hi0, 0.0f), gl = fmaxf((float)(dx - 1) * v.w0 + lo0, 0.0f);
      if (gh * gh + gyz <= cull) dxh = dx;
      if (gl * gl + gyz <= cull) dxl = -dx;
    }
    // the first and last kept cells, each sorted by x, trimmed to the row's x-chord |x_j - x_i| <= h
    // of the culling sphere (h^2 = cull - gyz): binary searches in the staging. A trimmed bead lies
    // beyond listRange + the margin less rounding, as a culled cell does
    const uint32_t k0 = v.k_row + (uint32_t)(R + dxl), k1 = k0 + (uint32_t)(dxh - dxl);
    const float h = sqrtf(cull - gyz), xlo = xi.x - h, xhi = xi.x + h;
    if (pair_lanes) {  // lane r: the first cell's lower bound, lane r + 32: the last cell's upper bound
      const bool up = lane >= WAVE / 2;
      const uint32_t k = up ? k1 : k0;
      lo_at = staged_x_bound(s_pos, s_cpre[k], s_cpre[k + 1u], up ? xhi : xlo, up);
    } else {
      lo_at = staged_x_bound(s_pos, s_cpre[k0], s_cpre[k0 + 1u], xlo, false);
      hi_at = staged_x_bound(s_pos, s_cpre[k1], s_cpre[k1 + 1u], xhi, true);
    }
  }
  if (pair_lanes) hi_at = (uint32_t)__shfl_down((int)lo_at, WAVE / 2, WAVE);
  if (live && (!pair_lanes || lane < WAVE / 2)) {
    beg = lo_at;
    len = hi_at > lo_at ? hi_at - lo_at : 0u;
  }
  uint32_t end = len;  // inclusive prefix over the row lanes: end[r]
  const int nlanes = pair_lanes ? WAVE / 2 : WAVE;
  for (int o = 1; o < nlanes; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)end, o, WAVE);
    if (lane >= (uint32_t)o) end += y;
  }
  const uint32_t total = uni((uint32_t)__builtin_amdgcn_readlane((int)end, nlanes - 1));
  // ---- dense sweep ----
  const uint32_t cap = c.list_cap;
  uint32_t* row = b.vl + (size_t)t * vl_stride(c);
  const float split2 = v.split2;
  uint32_t cn = 0, cf = 0;  // near and far hits so far
  // uniform walk state: the row r of the next chunk's first candidate, that row's end and the
  // offset (staged index - flat index) of its candidates
  uint32_t r = 0, rend = uni((uint32_t)__builtin_amdgcn_readlane((int)end, 0));
  uint32_t off = uni((uint32_t)__builtin_amdgcn_readlane((int)beg, 0));
  for (uint32_t base = 0; base < total; base += VL_BATCH * WAVE) {
    float4 xj[VL_BATCH];
#pragma unroll
    for (uint32_t k = 0; k < VL_BATCH; ++k) {
      const uint32_t cb = base + k * WAVE, f = cb + lane;
      uint32_t o = off;
      // boundaries inside this chunk: lanes at or past end[r] belong to a later row (empty rows
      // have end[r] = end[r - 1] and are passed at once); f < total keeps r below nrows
      while (rend < cb + WAVE && rend < total) {
        const uint32_t pb = rend - cb;
        ++r;
        off = uni((uint32_t)__builtin_amdgcn_readlane((int)beg, (int)r)) - rend;
        if (lane >= pb) o = off;
        rend = uni((uint32_t)__builtin_amdgcn_readlane((int)end, (int)r));
      }
      // lanes past the row set read the culled cells after it or the far-away padding that follows
      // the staging: no hit either way
      xj[k] = s_pos[f + o];
    }
#pragma unroll
    for (uint32_t k = 0; k < VL_BATCH; ++k) {
      const uint32_t kk = __float_as_uint(xj[k].w);
      // cheap test on the shifted position, the reference's exact one on the stored position in the
      // narrow band between sure and cull (the squared distance is fused: the band is far wider than
      // its rounding)
      const float ax = xj[k].x - xi.x, ay = xj[k].y - xi.y, az = xj[k].z - xi.z;
      const float a2 = __builtin_fmaf(ax, ax, __builtin_fmaf(ay, ay, az * az));
      // the lane masks come straight from the compares (scalar registers): the sure hits, then the
      // rare band resolved exactly
      const bool sure_hit = a2 <= sure, band = !sure_hit && a2 <= cull;
      const uint64_t m_sure = __builtin_amdgcn_fcmpf(a2, sure, 5 /* ordered <= */);
      const uint64_t m_band = __builtin_amdgcn_fcmpf(a2, cull, 5 /* ordered <= */) & ~m_sure;
      bool exact = false;
      uint64_t m_exact = 0;
      if (__builtin_expect(m_band != 0, 0)) {
        if (band) exact = dist_f(c, xi, b.pos[kk]) < lr;
        m_exact = __builtin_amdgcn_ballot_w64(exact);
      }
      const bool hit = (sure_hit || exact) && kk != t;
      const bool near = a2 < split2;
      const uint64_t m_hit = (m_sure | m_exact) & __builtin_amdgcn_uicmp(kk, t, 33 /* != */);
      const uint64_t m_near = __builtin_amdgcn_fcmpf(a2, split2, 4 /* ordered < */);
      // the hits compacted in lane order: near ones from the front, far ones from the back
      const uint64_t mn = m_hit & m_near, mf = m_hit & ~m_near;
      const uint32_t pn = __builtin_amdgcn_mbcnt_hi((uint32_t)(mn >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mn, cn));
      const uint32_t pf = __builtin_amdgcn_mbcnt_hi((uint32_t)(mf >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mf, cf));
      // past the capacity (an overflow, reported) the two parts overlap; every index stays in the row.
      // While 64 more hits fit (uniform), no lane needs the capacity test.
      const uint32_t idx = near ? pn : cap - 1u - pf;
      if (cn + cf + WAVE <= cap) {
        if (hit) row[idx] = kk;
      } else if (hit && (near ? pn : pf) < cap) {
        row[idx] = kk;
      }
      cn += (uint32_t)__popcll(mn);
      cf += (uint32_t)__popcll(mf);
    }
  }
  verlet_finish(b, t, cn, cf, cap);
}

// one workgroup per cell (logical block = cell index)
__device__ __forceinline__ void verlet_cell_body(SoupConfig c, DevBuf b, const uint32_t cellid) {
  __shared__ float4 s_pos[VL_STAGE];  // stencil beads: position + image shift, w = slot bits
  __shared__ uint32_t s_cpre[344];    // exclusive prefix of the (2R + 1)^3 <= 343 stencil cells' counts
  __shared__ uint2 s_cinfo[343];      // per stencil cell: first slot, image-shift code
  __shared__ uint32_t s_wsum[VL_WAVES];
  const uint32_t ncells = c.dims[0] * c.dims[1] * c.dims[2];
  if (cellid >= ncells) return;
  const uint32_t first = uni(b.cell_start[cellid]), np = uni(b.cell_start[cellid + 1u] - first);
  if (np == 0u) return;  // the whole workgroup leaves together
  const uint32_t lane = threadIdx.x & (WAVE - 1), w = uni(threadIdx.x / WAVE);
  const int cx = (int)(cellid % c.dims[0]), cy = (int)((cellid / c.dims[0]) % c.dims[1]), cz = (int)(cellid / (c.dims[0] * c.dims[1]));
  const int R = (int)c.walk_radius, side = 2 * R + 1;
  const uint32_t nst = (uint32_t)(side * side * side);
  // ---- stencil cell k = threadIdx.x in canonical order (dz = k / side^2 - R, dy = (k / side) % side
  // - R, dx = k % side - R): its slot range and image shift (code bits 2a / 2a + 1: -L / +L on axis
  // a), then the exclusive prefix of the counts (nst <= 343, fewer than the workgroup's threads) ----
  uint32_t cnt = 0;
  if (threadIdx.x < nst) {
    const uint32_t k = threadIdx.x;
    int x = cx + (int)(k % (uint32_t)side) - R, y = cy + (int)((k / (uint32_t)side) % (uint32_t)side) - R,
        z = cz + (int)(k / (uint32_t)(side * side)) - R;
    uint32_t code = 0;
    if (x < 0) { x += (int)c.dims[0]; code |= 1u; } else if (x >= (int)c.dims[0]) { x -= (int)c.dims[0]; code |= 2u; }
    if (y < 0) { y += (int)c.dims[1]; code |= 4u; } else if (y >= (int)c.dims[1]) { y -= (int)c.dims[1]; code |= 8u; }
    if (z < 0) { z += (int)c.dims[2]; code |= 16u; } else if (z >= (int)c.dims[2]) { z -= (int)c.dims[2]; code |= 32u; }
    const uint32_t cell = (uint32_t)x + c.dims[0] * ((uint32_t)y + c.dims[1] * (uint32_t)z);
    const uint32_t cs = b.cell_start[cell];
    cnt = b.cell_start[cell + 1u] - cs;
    s_cinfo[k] = make_uint2(cs, code);
  }
  uint32_t incl = cnt;
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)incl, o, WAVE);
    if (lane >= (uint32_t)o) incl += y;
  }
  if (lane == WAVE - 1) s_wsum[w] = incl;
  __syncthreads();
  uint32_t off = 0;
  for (uint32_t v = 0; v < w; ++v) off += s_wsum[v];
  if (threadIdx.x < nst) s_cpre[threadIdx.x] = off + incl - cnt;
  if (threadIdx.x == nst - 1u) s_cpre[nst] = off + incl;
  __syncthreads();
  const uint32_t total = s_cpre[nst];
  const uint32_t stage_cap = (c.list_stage != 0u && c.list_stage < VL_STAGE) ? c.list_stage : VL_STAGE;
  if (total + VL_PAD <= stage_cap) {
    // far-away padding after the stencil: the sweep reads up to VL_PAD entries past it
    for (uint32_t e = threadIdx.x; e < VL_PAD; e += VL_WAVES * WAVE)
      s_pos[total + e] = make_float4(3.0e37f, 3.0e37f, 3.0e37f, 0.0f);
    // ---- stage: wavefront w copies stencil rows w, w + VL_WAVES, ... A row's side cells follow each
    // other in x, so each run of cells with one x image shift is one contiguous slot range (one run,
    // two where the row wraps in x, three on a grid narrower than the stencil); the lanes stream the
    // row's beads four loads at a time ----
    const uint32_t nrow = (uint32_t)(side * side);
    for (uint32_t q = w; q < nrow; q += VL_WAVES) {
      const uint32_t k = q * (uint32_t)side;
      uint32_t slot0 = 0, code = 0, pre = 0;
      if (lane <= (uint32_t)side) pre = s_cpre[k + lane];
      if (lane < (uint32_t)side) {
        const uint2 ci = s_cinfo[k + lane];
        slot0 = ci.x;
        code = ci.y;
      }
      const uint32_t cxl = (uint32_t)__shfl_up((int)(code & 3u), 1, WAVE);
      uint64_t splits = __ballot(lane > 0u && lane < (uint32_t)side && (code & 3u) != cxl);
      const uint32_t j1 = splits ? (uint32_t)__builtin_ctzll(splits) : (uint32_t)side;
      splits &= splits - 1ull;
      const uint32_t j2 = splits ? (uint32_t)__builtin_ctzll(splits) : (uint32_t)side;
      const uint32_t e0 = uni((uint32_t)__builtin_amdgcn_readlane((int)pre, 0));
      const uint32_t n = uni((uint32_t)__builtin_amdgcn_readlane((int)pre, side)) - e0;
      const uint32_t E1 = uni((uint32_t)__builtin_amdgcn_readlane((int)pre, (int)j1)) - e0;
      const uint32_t E2 = uni((uint32_t)__builtin_amdgcn_readlane((int)pre, (int)j2)) - e0;
      // slot = rel + B (rel: the bead's offset in the row), per run
      const uint32_t B0 = uni((uint32_t)__builtin_amdgcn_readlane((int)slot0, 0));
      const uint32_t B1 = uni((uint32_t)__builtin_amdgcn_readlane((int)slot0, (int)(j1 < (uint32_t)side ? j1 : 0u))) - E1;
      const uint32_t B2 = uni((uint32_t)__builtin_amdgcn_readlane((int)slot0, (int)(j2 < (uint32_t)side ? j2 : 0u))) - E2;
      const uint32_t c0 = uni((uint32_t)__builtin_amdgcn_readlane((int)code, 0));
      const uint32_t c1 = uni((uint32_t)__builtin_amdgcn_readlane((int)code, (int)(j1 < (uint32_t)side ? j1 : 0u)));
      const uint32_t c2 = uni((uint32_t)__builtin_amdgcn_readlane((int)code, (int)(j2 < (uint32_t)side ? j2 : 0u)));
      const float X0 = (c0 & 1u) ? -c.box[0] : ((c0 & 2u) ? c.box[0] : 0.0f);
      const float X1 = (c1 & 1u) ? -c.box[0] : ((c1 & 2u) ? c.box[0] : 0.0f);
      const float X2 = (c2 & 1u) ? -c.box[0] : ((c2 & 2u) ? c.box[0] : 0.0f);
      const float sy = (c0 & 4u) ? -c.box[1] : ((c0 & 8u) ? c.box[1] : 0.0f);
      const float sz = (c0 & 16u) ? -c.box[2] : ((c0 & 32u) ? c.box[2] : 0.0f);
      for (uint32_t r0 = 0; r0 < n; r0 += 4u * WAVE) {
        float4 p[4];
        uint32_t sl[4];
        float sx[4];
#pragma unroll
        for (uint32_t u = 0; u < 4u; ++u) {
          const uint32_t rel = r0 + u * WAVE + lane;
          sl[u] = rel + (rel >= E2 ? B2 : (rel >= E1 ? B1 : B0));
          sx[u] = rel >= E2 ? X2 : (rel >= E1 ? X1 : X0);
          if (rel < n) p[u] = b.pos[sl[u]];
        }
#pragma unroll
        for (uint32_t u = 0; u < 4u; ++u) {
          const uint32_t rel = r0 + u * WAVE + lane;
          if (rel < n) s_pos[e0 + rel] = make_float4(p[u].x + sx[u], p[u].y + sy, p[u].z + sz, __uint_as_float(sl[u]));
        }
      }
    }
    __syncthreads();
    const VlShared vs = verlet_shared(c, cx, cy, cz);
    for (uint32_t i = w; i < np; i += VL_WAVES) verlet_row_staged(c, b, first + i, vs, s_pos, s_cpre);
  } else {
    // stencil too dense to stage: gather from the particle arrays (the tables alias the staging)
    uint4* seg = reinterpret_cast<uint4*>(s_pos) + w * WAVE;
    float4* sh = s_pos + (VL_WAVES + w) * WAVE;
    uint32_t* ends = reinterpret_cast<uint32_t*>(s_pos + 2u * VL_WAVES * WAVE) + w * WAVE;
    for (uint32_t i = w; i < np; i += VL_WAVES) verlet_row_global(c, b, first + i, ends, seg, sh);
  }
}
__global__ __launch_bounds__(VL_WAVES * WAVE, 2 * VL_WAVES / 4) void k_verlet_cells(SoupConfig c, DevBuf b) {
  verlet_cell_body(c, b, xcd_block());
}

// ================================================================================================
// DH head list: soup_build_es_list_main (electrostatics-long.wgsl:31-56), one wavefront per head,
// candidates in ascending head order compacted with a ballot so the row keeps the reference order.
// Candidate positions come from the compact head array (L2-resident), ES_BATCH chunks of 64 per
// round of loads; entries are head indices.
// ================================================================================================
#ifndef ES_BATCH
#define ES_BATCH 8u  // DH-list candidate chunks of 64 whose loads are issued together
#endif
__device__ __forceinline__ void es_build_body(SoupConfig c, DevBuf b, uint32_t nh, const uint32_t blk) {
  const uint32_t lane = threadIdx.x & (WAVE - 1);
  const uint32_t h = uni((blk * blockDim.x + threadIdx.x) / WAVE);
  if (h >= nh) return;
  const float4 xv = b.head_q[h];
  const float4 xi = make_float4(uni_f(xv.x), uni_f(xv.y), uni_f(xv.z), 0.0f);
  const float lr = c.es_list_range, lr_hi = (lr + 0.01f) * (lr + 0.01f), lr_lo = (lr - 0.01f) * (lr - 0.01f);
  const float ib0 = 1.0f / c.box[0], ib1 = 1.0f / c.box[1], ib2 = 1.0f / c.box[2];
  const uint32_t cap = c.es_cap;
  uint32_t* row = b.es + (size_t)h * cap;
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t count = 0;
  for (uint32_t base = 0; base < nh; base += ES_BATCH * WAVE) {
    float4 xj[ES_BATCH];
#pragma unroll
    for (uint32_t k = 0; k < ES_BATCH; ++k) {
      const uint32_t g = base + k * WAVE + lane;
      xj[k] = b.head_q[g < nh ? g : h];
    }
#pragma unroll
    for (uint32_t k = 0; k < ES_BATCH; ++k) {
      const uint32_t g = base + k * WAVE + lane;
      bool hit = false;
      if (g < nh && g != h) {
        float ax = xi.x - xj[k].x, ay = xi.y - xj[k].y, az = xi.z - xj[k].z;
        ax -= c.box[0] * rintf(ax * ib0);
        ay -= c.box[1] * rintf(ay * ib1);
        az -= c.box[2] * rintf(az * ib2);
        const float a2 = ax * ax + ay * ay + az * az;
        hit = a2 <= lr_lo;
        if (!hit && a2 <= lr_hi) hit = dist_f(c, xi, xj[k]) < lr;
      }
      const uint64_t mask = __ballot(hit);
      if (hit) {
        const uint32_t p = count + (uint32_t)__popcll(mask & below);
        if (p < cap) row[p] = g;
      }
      count += (uint32_t)__popcll(mask);
    }
  }
  if (lane == 0) {
    b.es_count[h] = count < cap ? count : cap;
    if (count > cap) b.flags[1] = 1u;
  }
}
__global__ __launch_bounds__(256) void k_es_build(SoupConfig c, DevBuf b, uint32_t nh) { es_build_body(c, b, nh, blockIdx.x); }

// ---- the same rows over a head cell grid --------------------------------------------------------
// Cells of edge L / d >= (listRange + 0.02) / 2 per axis, d >= 5: every head within listRange + 0.01
// of a head (the widest the test above admits) lies in the 5^3 cells around its own cell, each of
// which the walk visits once. The heads are counting-sorted by cell (their order inside a cell is
// free: the hits are gathered in a bitmap over head indices and written out in ascending order,
// which is the row order of the all-pairs build).
struct EsGrid {
  uint32_t d[3];
  float inv_w[3];  // d / L per axis
  uint32_t* count;  // per cell
  uint32_t* start;  // exclusive prefix, ncells + 1
  float4* rec;      // heads by cell: position, w = head index bits
};
__device__ __forceinline__ uint32_t es_cell_axis(float x, float inv_w, uint32_t d) {
  const int k = (int)floorf(x * inv_w);
  return k < 0 ? 0u : (k >= (int)d ? d - 1u : (uint32_t)k);
}
__device__ __forceinline__ uint32_t es_cell_of(const EsGrid& g, float4 x) {
  return es_cell_axis(x.x, g.inv_w[0], g.d[0]) + g.d[0] * (es_cell_axis(x.y, g.inv_w[1], g.d[1]) + g.d[1] * es_cell_axis(x.z, g.inv_w[2], g.d[2]));
}
__global__ __launch_bounds__(256) void k_es_cell_count(DevBuf b, uint32_t nh, EsGrid g) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h < nh) atomicAdd(&g.count[es_cell_of(g, b.head_q[h])], 1u);
}
// count must be zero again before this (it serves as the per-cell cursor)
__global__ __launch_bounds__(256) void k_es_cell_fill(DevBuf b, uint32_t nh, EsGrid g) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  const float4 x = b.head_q[h];
  const uint32_t cell = es_cell_of(g, x);
  const uint32_t k = g.start[cell] + atomicAdd(&g.count[cell], 1u);
  g.rec[k] = make_float4(x.x, x.y, x.z, __uint_as_float(h));
}

#define ES_CELL_WAVES 4u
#ifndef ES_CELL_BATCH
#define ES_CELL_BATCH 4u  // candidate chunks of 64 whose loads are issued together
#endif
// one wavefront per head, heads taken in cell order (neighbouring wavefronts share their stencils);
// dynamic LDS: ES_CELL_WAVES bitmaps of ceil(nh / 32) words
__global__ __launch_bounds__(ES_CELL_WAVES * WAVE) void k_es_cells(SoupConfig c, DevBuf b, uint32_t nh, EsGrid g) {
  extern __shared__ uint32_t es_bits[];
  const uint32_t lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const uint32_t k = uni(blockIdx.x * ES_CELL_WAVES + w);
  if (k >= nh) return;  // no workgroup barrier below
  const uint32_t nwords = (nh + 31u) / 32u;
  uint32_t* bits = es_bits + (size_t)w * nwords;
  for (uint32_t i = lane; i < nwords; i += WAVE) bits[i] = 0u;
  const float4 xv = g.rec[k];
  const float4 xi = make_float4(uni_f(xv.x), uni_f(xv.y), uni_f(xv.z), 0.0f);
  const uint32_t hh = uni(__float_as_uint(xv.w));
  const float lr = c.es_list_range, lr_hi = (lr + 0.01f) * (lr + 0.01f), lr_lo = (lr - 0.01f) * (lr - 0.01f);
  const float ib0 = 1.0f / c.box[0], ib1 = 1.0f / c.box[1], ib2 = 1.0f / c.box[2];
  // ---- the 125 stencil cells in canonical order, rows r = lane (A) and r = 64 + lane (B): first
  // record and count; inclusive prefix of the counts over r ----
  const int cx = (int)es_cell_axis(xi.x, g.inv_w[0], g.d[0]), cy = (int)es_cell_axis(xi.y, g.inv_w[1], g.d[1]),
            cz = (int)es_cell_axis(xi.z, g.inv_w[2], g.d[2]);
  uint32_t stA = 0, nA = 0, stB = 0, nB = 0;
#pragma unroll
  for (uint32_t half = 0; half < 2u; ++half) {
    const uint32_t r = lane + half * WAVE;
    if (r < 125u) {
      int x = cx + (int)(r % 5u) - 2, y = cy + (int)((r / 5u) % 5u) - 2, z = cz + (int)(r / 25u) - 2;
      x += x < 0 ? (int)g.d[0] : (x >= (int)g.d[0] ? -(int)g.d[0] : 0);
      y += y < 0 ? (int)g.d[1] : (y >= (int)g.d[1] ? -(int)g.d[1] : 0);
      z += z < 0 ? (int)g.d[2] : (z >= (int)g.d[2] ? -(int)g.d[2] : 0);
      const uint32_t cell = (uint32_t)x + g.d[0] * ((uint32_t)y + g.d[1] * (uint32_t)z);
      const uint32_t s0 = g.start[cell], s1 = g.start[cell + 1u];
      if (half == 0u) { stA = s0; nA = s1 - s0; } else { stB = s0; nB = s1 - s0; }
    }
  }
  uint32_t eA = nA, eB = nB;
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t ya = (uint32_t)__shfl_up((int)eA, o, WAVE), yb = (uint32_t)__shfl_up((int)eB, o, WAVE);
    if (lane >= (uint32_t)o) {
      eA += ya;
      eB += yb;
    }
  }
  const uint32_t totA = uni((uint32_t)__builtin_amdgcn_readlane((int)eA, WAVE - 1));
  eB += totA;
  const uint32_t total = uni((uint32_t)__builtin_amdgcn_readlane((int)eB, WAVE - 1));
  // offset (record index - flat index) of row r: start[r] - (end[r] - count[r])
  const uint32_t oA = stA - (eA - nA), oB = stB - (eB - nB);
  auto row_end = [&](uint32_t r) -> uint32_t {
    return r < WAVE ? uni((uint32_t)__builtin_amdgcn_readlane((int)eA, (int)r)) : uni((uint32_t)__builtin_amdgcn_readlane((int)eB, (int)(r - WAVE)));
  };
  auto row_off = [&](uint32_t r) -> uint32_t {
    return r < WAVE ? uni((uint32_t)__builtin_amdgcn_readlane((int)oA, (int)r)) : uni((uint32_t)__builtin_amdgcn_readlane((int)oB, (int)(r - WAVE)));
  };
  wave_lds_sync();  // the cleared bitmap before any hit lands in it
  // ---- sweep the flat candidate range; the row boundaries inside each chunk are walked with
  // uniform reads of the lanes holding them ----
  uint32_t r = 0, rend = row_end(0), off = row_off(0);
  for (uint32_t base = 0; base < total; base += ES_CELL_BATCH * WAVE) {
    float4 xj[ES_CELL_BATCH];
#pragma unroll
    for (uint32_t q = 0; q < ES_CELL_BATCH; ++q) {
      const uint32_t cb = base + q * WAVE, f = cb + lane;
      uint32_t o = off;
      while (rend < cb + WAVE && rend < total) {  // rows ending inside this chunk (empty ones at once)
        const uint32_t pb = rend - cb;
        ++r;
        off = row_off(r);
        if (lane >= pb) o = off;
        rend = row_end(r);
      }
      xj[q] = g.rec[f < total ? f + o : k];
    }
#pragma unroll
    for (uint32_t q = 0; q < ES_CELL_BATCH; ++q) {
      const uint32_t f = base + q * WAVE + lane;
      const uint32_t gj = __float_as_uint(xj[q].w);
      if (f < total && gj != hh) {
        float ax = xi.x - xj[q].x, ay = xi.y - xj[q].y, az = xi.z - xj[q].z;
        ax -= c.box[0] * rintf(ax * ib0);
        ay -= c.box[1] * rintf(ay * ib1);
        az -= c.box[2] * rintf(az * ib2);
        const float a2 = ax * ax + ay * ay + az * az;
        bool hit = a2 <= lr_lo;
        if (!hit && a2 <= lr_hi) hit = dist_f(c, xi, make_float4(xj[q].x, xj[q].y, xj[q].z, 0.0f)) < lr;
        if (hit) atomicOr(&bits[gj >> 5], 1u << (gj & 31u));
      }
    }
  }
  wave_lds_sync();
  // ---- the row in ascending head order ----
  const uint32_t cap = c.es_cap;
  uint32_t* row = b.es + (size_t)hh * cap;
  uint32_t count = 0;
  for (uint32_t i0 = 0; i0 < nwords; i0 += WAVE) {
    const uint32_t i = i0 + lane;
    uint32_t word = i < nwords ? bits[i] : 0u;
    const uint32_t pc = (uint32_t)__popc(word);
    uint32_t incl = pc;
    for (int o = 1; o < WAVE; o <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)incl, o, WAVE);
      if (lane >= (uint32_t)o) incl += y;
    }
    uint32_t p = count + incl - pc;
    while (word) {
      const uint32_t bit = (uint32_t)__builtin_ctz(word);
      word &= word - 1u;
      if (p < cap) row[p] = i * 32u + bit;
      ++p;
    }
    count += uni((uint32_t)__builtin_amdgcn_readlane((int)incl, WAVE - 1));
  }
  if (lane == 0) {
    b.es_count[hh] = count < cap ? count : cap;
    if (count > cap) b.flags[1] = 1u;
  }
}

// ================================================================================================
// Forces: soup_force_list_main (verlet.wgsl:132-149) = bondedForce (step.wgsl:152-193) + the list
// sum of nonbondedSoup (pair.wgsl:168-204), then soup_es_force_far_main (electrostatics-long.wgsl
// :63-82). Per-kind-pair constants are staged in LDS.
// ================================================================================================
// per kind pair: x = wca_cut(pairB)^2, y = pairB^2, z = class-table depth, w = 1 if both beads are
// polar (acid-soap eligible); one 16-byte LDS read per pair
struct PairTab {
  float4 e[SOUP_MAX_KINDS * SOUP_MAX_KINDS];
};

__device__ __forceinline__ void load_pair_tab(const SoupConfig& c, PairTab& T) {
  for (uint32_t k = threadIdx.x; k < SOUP_MAX_KINDS * SOUP_MAX_KINDS; k += blockDim.x) {
    uint32_t ki = k / SOUP_MAX_KINDS, kj = k % SOUP_MAX_KINDS;
    float bb = c.sigma * (c.radius[ki] + c.radius[kj]) * 0.5f;
    float cut = c.wca_factor * bb;
    T.e[k] = make_float4(cut * cut, bb * bb, c.attr[c.cls[ki]][c.cls[kj]],
                         (c.cls[ki] == 1u && c.cls[kj] == 1u) ? 1.0f : 0.0f);
  }
  __syncthreads();
}

struct PairConst {
  float bx, by, bz, hx, hy, hz;
  float eps24, rc_attr, rc2_lo, rc2_hi, attr_k2, attr_a;
  float acid, es_A, es_kappa, es_Frc, split2, es_rc, reach2, reach, es_kappa_l2, qdep;
};

__device__ __forceinline__ PairConst make_pair_const(const SoupConfig& c) {
  PairConst p;
  p.bx = c.box[0]; p.by = c.box[1]; p.bz = c.box[2];
  p.hx = 0.5f * c.box[0]; p.hy = 0.5f * c.box[1]; p.hz = 0.5f * c.box[2];
  p.eps24 = 24.0f * c.epsilon;
  p.rc_attr = c.wca_factor * c.b_tt;
  p.rc2_lo = p.rc_attr * p.rc_attr;
  p.rc2_hi = (p.rc_attr + c.wc) * (p.rc_attr + c.wc);
  p.attr_k2 = SOUP_PI_F / c.wc;                        // sin(2x) = sin(k2 (r - rc))
  p.attr_a = c.epsilon * SOUP_PI_F / (2.0f * c.wc);
  p.acid = c.acid_soap;
  p.es_A = c.es_A; p.es_kappa = c.es_kappa; p.es_Frc = c.es_Frc;
  p.es_kappa_l2 = -c.es_kappa * 1.44269504f;  // exp(-kappa r) = exp2(es_kappa_l2 * r)
  p.split2 = c.es_split * c.es_split;
  p.es_rc = c.es_rc;
  const float reach = force_reach(c);
  p.reach2 = reach * reach * 1.0001f;
  p.reach = reach;
  p.qdep = (float)c.ph_q;
  return p;
}

// |d|^2 in one fixed order of operations (explicit fused multiply-adds, which the compiler does not
// reorder). The list force and the far DH kernel split the DH term at es_split; both decide a pair's
// side from this value of the same displacement, so every pair lands on exactly one side.
__device__ __forceinline__ float r2_fixed(float dx, float dy, float dz) {
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Minimum image with relative (not box-absolute) accuracy: for a pair straddling the periodic
// boundary the box length is taken off the coordinate near L first, which is exact (Sterbenz), so
// the small displacement carries only its own rounding. Positions lie in [0, L].
__device__ __forceinline__ float mi_acc(float a, float b, float L, float hL) {
  float d = a - b;
  if (d > hL) d = (a - L) - b;
  else if (d < -hL) d = a - (b - L);
  return d;
}

struct Acc {
  double x, y, z;
};

// sin(x) for x in [0, pi]: reflected to [0, pi/2], odd Taylor polynomial through x^11
// (truncation < 6e-8 at pi/2)
__device__ __forceinline__ float sin_0_pi(float x) {
#pragma clang fp contract(fast)
  float y = x > 1.57079633f ? 3.14159265f - x : x;
  float y2 = y * y;
  float p = -2.50521084e-8f;
  p = p * y2 + 2.75573192e-6f;
  p = p * y2 - 1.98412698e-4f;
  p = p * y2 + 8.33333333e-3f;
  p = p * y2 - 1.66666667e-1f;
  return y + y * y2 * p;
}

// Accumulates the force on i from j (nonbondedSoup). Per-pair terms in f32 with hardware reciprocal
// and square root (1 ulp), branch-free WCA and attraction shell, the rare DH-near term behind a
// branch, FP64 sum. The displacement is a - b; only a pair straddling the periodic boundary (|d| >
// L/2 on some axis) takes the exact image of mi_acc, so the result stays within 1e-5 of the FP64
// recomputation relative to max(|F_i|, median |F|).
__device__ __forceinline__ void pair_force(const PairConst& P, const PairTab& T, float4 xi, float4 xj, float qi, float qj,
                                           Acc& f) {
#pragma clang fp contract(fast)
  float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
  if (fabsf(dx) > P.hx || fabsf(dy) > P.hy || fabsf(dz) > P.hz) {
    dx = mi_acc(xi.x, xj.x, P.bx, P.hx);
    dy = mi_acc(xi.y, xj.y, P.by, P.hy);
    dz = mi_acc(xi.z, xj.z, P.bz, P.hz);
  }
  float r2 = fmaxf(r2_fixed(dx, dy, dz), 1e-30f);
  const float4 tb = T.e[((uint32_t)xi.w * SOUP_MAX_KINDS + (uint32_t)xj.w) & (SOUP_MAX_KINDS * SOUP_MAX_KINDS - 1u)];
  float ir2 = __builtin_amdgcn_rcpf(r2);
  float s2 = tb.y * ir2;
  float s6 = s2 * s2 * s2;
  float fr = (r2 < tb.x) ? P.eps24 * s6 * (2.0f * s6 - 1.0f) * ir2 : 0.0f;
  float as = tb.z + ((P.acid > 0.0f && tb.w != 0.0f && ((qi != 0.0f) != (qj != 0.0f))) ? P.acid : 0.0f);
  float r = __builtin_amdgcn_sqrtf(r2);
  float ir = __builtin_amdgcn_rcpf(r);
  float x = fminf(fmaxf(P.attr_k2 * (r - P.rc_attr), 0.0f), 3.14159265f);
  bool shell = as > 0.0f && r2 >= P.rc2_lo && r2 <= P.rc2_hi;
  fr -= shell ? as * P.attr_a * sin_0_pi(x) * ir : 0.0f;
  float qq = qi * qj;
  if (qq != 0.0f && r2 < P.split2 && r < P.es_rc) {
    float u = P.es_A * expf(-P.es_kappa * r) * ir;
    fr += qq * (u * (ir + P.es_kappa) - P.es_Frc) * ir;
  }
  if (r2 < 1e-12f) fr = 0.0f;  // r < 1e-6: no force (pair.wgsl:48)
  double g = (double)fr;
  f.x += g * (double)dx;
  f.y += g * (double)dy;
  f.z += g * (double)dz;
}

__device__ __forceinline__ void bond_term(const SoupConfig& c, const PairConst& P, float4 xi, float4 xp, bool fene, Acc& f) {
  float dx = mi_acc(xi.x, xp.x, P.bx, P.hx), dy = mi_acc(xi.y, xp.y, P.by, P.hy), dz = mi_acc(xi.z, xp.z, P.bz, P.hz);
  double r = fmax(sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz), 1e-6);
  double dv;
  if (fene) {
    double x = r / (double)c.r_inf;
    dv = (double)c.k_fene * r / (1.0 - x * x);
  } else {
    dv = (double)c.k_bend * (r - (double)c.r_bend);
  }
  double m = dv / r;
  f.x -= m * dx;
  f.y -= m * dy;
  f.z -= m * dz;
}

// bondedForce via inv[] (partners are particle indices)
__device__ __forceinline__ void bonded_force(const SoupConfig& c, const PairConst& P, const DevBuf& b, uint32_t t, float4 xi,
                                             Acc& f) {
  uint4 s = b.bs[t];
  uint32_t me = b.id[t];
  uint32_t sl[3] = {s.x, s.y, s.z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    uint32_t p = sl[k];
    if (p == SOUP_NONE) continue;
    uint32_t u = b.inv[p];
    bond_term(c, P, xi, b.pos[u], true, f);
    uint4 ms = b.bs[u];
    uint32_t ml[3] = {ms.x, ms.y, ms.z};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      uint32_t o = ml[q];
      if (o == SOUP_NONE || o == me) continue;
      bond_term(c, P, xi, b.pos[b.inv[o]], false, f);
    }
  }
  uint32_t owner = b.clink[t];
  if (owner != SOUP_NONE) bond_term(c, P, xi, b.pos[b.inv[owner]], true, f);
}

// minimum image of mi_acc written as selects (no divergent branch inside a wavefront)
__device__ __forceinline__ float mi_sel(float a, float b, float L, float hL) {
  const float d = a - b, lo = (a - L) - b, hi = a - (b - L);
  return d > hL ? lo : (d < -hL ? hi : d);
}

// nonbondedSoup (pair.wgsl:168-204) for one pair already known to lie inside the interaction reach;
// tb is the pair-table entry of the two kinds. The WCA core in FP64 from the exact f32 displacement
// (its steep terms carry the largest f32 rounding, which strong pairs that nearly cancel would
// expose); the attraction shell and the DH-near term in f32 from one hardware reciprocal square root
// (1 ulp); FP64 sum.
__device__ __forceinline__ void pair_body(const PairConst& P, float4 tb, float dx, float dy, float dz, float qi, float qj,
                                          Acc& f) {
#pragma clang fp contract(fast)
  const float r2 = r2_fixed(dx, dy, dz);
  const float ir = __builtin_amdgcn_rsqf(r2);
  const float r = r2 * ir;
  double fw = 0.0;
  if (r2 < tb.x) {
    const double r2d = (double)dx * (double)dx + (double)dy * (double)dy + (double)dz * (double)dz;
    // 1 / r2d: the hardware estimate refined by two Newton-Raphson steps (e <- e (2 - r2d e)), each
    // squaring the relative error, to within a few ulp of the correctly rounded quotient
    double ir2d = __builtin_amdgcn_rcp(r2d);
    ir2d = ir2d * (2.0 - r2d * ir2d);
    ir2d = ir2d * (2.0 - r2d * ir2d);
    const double s2 = (double)tb.y * ir2d;
    const double s6 = s2 * s2 * s2;
    fw = (double)P.eps24 * s6 * (2.0 * s6 - 1.0) * ir2d;
  }
  float fr = 0.0f;
  const float as = tb.z + ((P.acid > 0.0f && tb.w != 0.0f && ((qi != 0.0f) != (qj != 0.0f))) ? P.acid : 0.0f);
  const float x = fminf(fmaxf(P.attr_k2 * (r - P.rc_attr), 0.0f), 3.14159265f);
  const bool shell = as > 0.0f && r2 >= P.rc2_lo && r2 <= P.rc2_hi;
  fr -= shell ? as * P.attr_a * sin_0_pi(x) * ir : 0.0f;
  const float qq = qi * qj;
  if (qq != 0.0f && r2 < P.split2 && r < P.es_rc) {
    const float u = P.es_A * __builtin_amdgcn_exp2f(P.es_kappa_l2 * r) * ir;
    fr += qq * (u * (ir + P.es_kappa) - P.es_Frc) * ir;
  }
  const double g = r2 < 1e-12f ? 0.0 : (double)fr + fw;  // r < 1e-6: no force (pair.wgsl:48)
  f.x += g * (double)dx;
  f.y += g * (double)dy;
  f.z += g * (double)dz;
}

// one far DH pair of soup_es_force_far_main (electrostatics-long.wgsl:63-82): only pairs at or beyond
// the near/far split and inside the DH cutoff act; f32 terms from one reciprocal square root and one
// hardware exp2 (relative error ~1e-7 over the cutoff range), summed per lane in f32 (a lane takes at
// most a few dozen terms of a row) and over the lanes in FP64.
// The minimum image is d - L rint(d / L), three instructions and no branch. It equals the list
// force's displacement (mi_sel) for every pair that does not wrap, and is within a few 1e-6 sigma of
// it for one that does; only the side of es_split depends on the last bit, so a wrapping pair within
// 1e-4 of the split (relative, in r^2) takes mi_sel's exact displacement for that decision.
struct AccF {
  float x, y, z;
};
__device__ __forceinline__ void far_pair_lane(const PairConst& P, float ibx, float iby, float ibz, float4 xi, float4 xj,
                                              float qi, float qj, AccF& f) {
#pragma clang fp contract(fast)
  const float ax = xi.x - xj.x, ay = xi.y - xj.y, az = xi.z - xj.z;
  const float kx = __builtin_rintf(ax * ibx), ky = __builtin_rintf(ay * iby), kz = __builtin_rintf(az * ibz);
  float dx = __builtin_fmaf(-P.bx, kx, ax), dy = __builtin_fmaf(-P.by, ky, ay), dz = __builtin_fmaf(-P.bz, kz, az);
  float r2 = r2_fixed(dx, dy, dz);
  if ((kx != 0.0f || ky != 0.0f || kz != 0.0f) && fabsf(r2 - P.split2) <= 1e-4f * P.split2) {
    dx = mi_sel(xi.x, xj.x, P.bx, P.hx);
    dy = mi_sel(xi.y, xj.y, P.by, P.hy);
    dz = mi_sel(xi.z, xj.z, P.bz, P.hz);
    r2 = r2_fixed(dx, dy, dz);
  }
  const float qq = qi * qj;
  const float ir = __builtin_amdgcn_rsqf(r2);
  const float r = r2 * ir;
  const float u = P.es_A * __builtin_amdgcn_exp2f(P.es_kappa_l2 * r) * ir;
  const float fr = (qq != 0.0f && r2 >= P.split2 && r < P.es_rc) ? qq * (u * (ir + P.es_kappa) - P.es_Frc) * ir : 0.0f;
  f.x = __builtin_fmaf(fr, dx, f.x);
  f.y = __builtin_fmaf(fr, dy, f.y);
  f.z = __builtin_fmaf(fr, dz, f.z);
}

// bondedForce split over four lanes: lane k < 3 takes bond slot k (FENE + the bends through that
// partner), lane 3 the centre link
__device__ __forceinline__ void bonded_lane(const SoupConfig& c, const PairConst& P, const DevBuf& b, uint32_t t, float4 xi,
                                            uint4 own, uint32_t owner, uint32_t k, Acc& f) {
  if (k < 3u) {
    const uint32_t p = k == 0u ? own.x : (k == 1u ? own.y : own.z);
    if (p == SOUP_NONE) return;
    const uint32_t me = b.id[t], u = b.inv[p];
    bond_term(c, P, xi, b.pos[u], true, f);
    const uint4 ms = b.bs[u];
    const uint32_t ml[3] = {ms.x, ms.y, ms.z};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const uint32_t o = ml[q];
      if (o == SOUP_NONE || o == me) continue;
      bond_term(c, P, xi, b.pos[b.inv[o]], false, f);
    }
  } else if (k == 3u && owner != SOUP_NONE) {
    bond_term(c, P, xi, b.pos[b.inv[owner]], true, f);
  }
}


#ifndef FAR_BATCH
#define FAR_BATCH 16u  // far DH chunks of 64 whose loads are issued together
#endif

// FP64 sum over each aligned group of TPP lanes (TPP = 4, 8 or 16), every lane of the wavefront
// active: butterflies inside quads, half rows and rows through DPP; every lane of a group ends with
// the group's sum
template <uint32_t TPP>
__device__ __forceinline__ double group_sum_d(double v) {
  static_assert(TPP == 4u || TPP == 8u || TPP == 16u, "groups of 4, 8 or 16 lanes");
  v += dpp_d<0xB1>(v);                  // quad_perm [1,0,3,2]
  v += dpp_d<0x4E>(v);                  // quad_perm [2,3,0,1]
  if (TPP >= 8u) v += dpp_d<0x141>(v);   // row_half_mirror
  if (TPP >= 16u) v += dpp_d<0x140>(v);  // row_mirror
  return v;
}

// Far DH rows of the charged heads (soup_es_force_far_main, electrostatics-long.wgsl:63-82), one
// wavefront per head in slot order, into far_acc in FP64; the list force adds them before it rounds
// the head's force to f32.
__device__ __forceinline__ void far_dh_body(SoupConfig c, DevBuf b, uint32_t nh, const uint32_t blk) {
#pragma clang fp contract(fast)
  const uint32_t lane = threadIdx.x & (WAVE - 1), w = uni(threadIdx.x / WAVE);
  const uint32_t r = uni(blk * (blockDim.x / WAVE) + w);
  if (r >= nh) return;
  const uint32_t hh = uni(b.head_rows[r]);
  const uint32_t cap = c.es_cap;
  const uint32_t* erow = b.es + (size_t)hh * cap;
  // the first batch of row entries is requested with the head's own record and count, not after
  // them: entries at or past the count (still inside the row's capacity) are read but never used
  uint32_t v[FAR_BATCH];
#pragma unroll
  for (uint32_t k = 0; k < FAR_BATCH; ++k) {
    const uint32_t s = k * WAVE + lane;
    v[k] = s < cap ? erow[s] : hh;
  }
  const float4 xv = b.head_q[hh];
  const uint32_t ecnt = uni(b.es_count[hh]);
  const float4 xi = make_float4(uni_f(xv.x), uni_f(xv.y), uni_f(xv.z), 0.0f);
  const float qi = uni_f(xv.w);
  AccF f{0.0f, 0.0f, 0.0f};
  if (qi != 0.0f) {
    const PairConst P = make_pair_const(c);
    const float ibx = 1.0f / P.bx, iby = 1.0f / P.by, ibz = 1.0f / P.bz;
    // every chunk of a batch is loaded and summed with no branch (a branch per chunk would make
    // each load wait for the one before); entries past the count add nothing
    for (uint32_t base = 0; base < ecnt; base += FAR_BATCH * WAVE) {
      if (base > 0) {
#pragma unroll
        for (uint32_t k = 0; k < FAR_BATCH; ++k) {
          const uint32_t s = base + k * WAVE + lane;
          v[k] = s < ecnt ? erow[s] : hh;
        }
      }
      float4 xj[FAR_BATCH];
#pragma unroll
      for (uint32_t k = 0; k < FAR_BATCH; ++k) xj[k] = b.head_q[base + k * WAVE + lane < ecnt ? v[k] : hh];
#pragma unroll
      for (uint32_t k = 0; k < FAR_BATCH; ++k)
        far_pair_lane(P, ibx, iby, ibz, xi, xj[k], qi, base + k * WAVE + lane < ecnt ? xj[k].w : 0.0f, f);
    }
  }
  const double sx = wave_sum_d_dpp((double)f.x), sy = wave_sum_d_dpp((double)f.y), sz = wave_sum_d_dpp((double)f.z);
  if (lane == 0) {
    b.far_acc[3u * hh] = sx;
    b.far_acc[3u * hh + 1u] = sy;
    b.far_acc[3u * hh + 2u] = sz;
  }
}
__global__ __launch_bounds__(256) void k_far_dh(SoupConfig c, DevBuf b, uint32_t nh) { far_dh_body(c, b, nh, xcd_block()); }

#ifndef FORCE_TPP
#define FORCE_TPP 8u    // lanes per particle
#endif
#ifndef GROUP_BATCH
#define GROUP_BATCH 4u  // list entries per lane whose loads are issued together
#endif

// List force, FORCE_TPP lanes per particle: soup_force_list_main (verlet.wgsl:132-149) =
// bondedForce + the list sum of nonbondedSoup, plus the far DH sum of a charged head from far_acc.
// Lane k of a particle's group takes entries k, k + TPP, ... of its row -- only the near part while
// the measured drift allows it (list_split) -- and runs the pair terms on those inside the
// interaction reach; lane sums are FP64, summed over the group, and rounded to f32 once. A particle
// farther than reach + 0.5 sigma from every face has no pair that wraps inside the reach, so it
// takes the plain difference (equal to the minimum image for every pair that can act); the others
// take the exact image of mi_acc.
__device__ __forceinline__ void force_group_body(SoupConfig c, DevBuf b, int es, const uint32_t blk) {
#pragma clang fp contract(fast)
  __shared__ PairTab T;
  load_pair_tab(c, T);
  const uint32_t sub = threadIdx.x & (FORCE_TPP - 1u);
  const uint32_t g = (blk * blockDim.x + threadIdx.x) / FORCE_TPP;
  const bool valid = g < c.n;
  const uint32_t t = valid ? g : c.n - 1u;
  const PairConst P = make_pair_const(c);
  const uint32_t cap = c.list_cap;
  const uint32_t* row = b.vl + (size_t)t * vl_stride(c);
  const float4 xi = b.pos[t];
  const float qi = b.charge[t];
  const uint32_t cn = b.vnear[t];
  // A pair acting now was closer than reach + d_i + d_j at the rebuild (d = displacement since), and
  // d_j <= dmax, the measured maximum: while d_i + dmax stays within the split's skin the near part of
  // i's row holds every acting pair. Every row qualifies while 2 dmax does (near_only); otherwise each
  // particle tests its own displacement, so only the rows of the particles that moved far are read
  // whole.
  const uint32_t dbits = b.flags[2];
  const bool every_near = near_only(c, dbits);
  if (g == 0 && sub == 0) atomicAdd(&b.flags[every_near ? 8 : 9], 1u);
  bool all = false;
  if (!every_near && valid) {
    const float dm = sqrtf(__uint_as_float(dbits));
    const float di = sqrtf(__uint_as_float(drift_sq_bits(c, xi, b.pos_at_rebuild[t])));
    all = !(di + dm <= list_split(c) - 0.002f - force_reach(c) * 1.00005f);
  }
  {
    const uint64_t lead = __builtin_amdgcn_ballot_w64(valid && sub == 0u);
    const uint64_t whole = __builtin_amdgcn_ballot_w64(valid && sub == 0u && all);
    if ((threadIdx.x & (WAVE - 1)) == 0) {
      uint32_t* rr = b.rowreads + 2u * (blk * (blockDim.x / WAVE) + threadIdx.x / WAVE);
      rr[0] += (uint32_t)__popcll(lead & ~whole);
      rr[1] += (uint32_t)__popcll(whole);
    }
  }
  const uint32_t cnt = valid ? (all ? b.vcount[t] : cn) : 0u;
  const uint32_t cmax = uni(wave_max_u(cnt));
  const uint4 own = b.bs[t];
  const uint32_t owner = b.clink[t];
  Acc f{0.0, 0.0, 0.0};
  if (valid && (own.x & own.y & own.z & owner) != SOUP_NONE && sub < 4u) bonded_lane(c, P, b, t, xi, own, owner, sub, f);
  const float4* trow = &T.e[((uint32_t)xi.w & (SOUP_MAX_KINDS - 1u)) * SOUP_MAX_KINDS];
  const float m = P.reach + 0.5f;
  const bool interior = xi.x >= m && xi.x <= P.bx - m && xi.y >= m && xi.y <= P.by - m && xi.z >= m && xi.z <= P.bz - m;
  for (uint32_t base = 0; base < cmax; base += FORCE_TPP * GROUP_BATCH) {
    uint32_t u[GROUP_BATCH];
#pragma unroll
    for (uint32_t k = 0; k < GROUP_BATCH; ++k) {
      const uint32_t s = base + k * FORCE_TPP + sub;
      u[k] = s < cnt ? row[row_index(s, cn, cap)] : t;
    }
    float4 xj[GROUP_BATCH];
#pragma unroll
    for (uint32_t k = 0; k < GROUP_BATCH; ++k) xj[k] = b.gq[u[k]];
#pragma unroll
    for (uint32_t k = 0; k < GROUP_BATCH; ++k) {
      float dx, dy, dz;
      if (interior) {
        dx = xi.x - xj[k].x; dy = xi.y - xj[k].y; dz = xi.z - xj[k].z;
      } else {
        dx = mi_sel(xi.x, xj[k].x, P.bx, P.hx); dy = mi_sel(xi.y, xj[k].y, P.by, P.hy); dz = mi_sel(xi.z, xj[k].z, P.bz, P.hz);
      }
      const float r2 = dx * dx + dy * dy + dz * dz;
      if (base + k * FORCE_TPP + sub < cnt && r2 >= 1e-12f && r2 < P.reach2) {
        const uint32_t kw = (uint32_t)xj[k].w;
        pair_body(P, trow[kw & (SOUP_MAX_KINDS - 1u)], dx, dy, dz, qi, (kw & SOUP_MAX_KINDS) ? P.qdep : 0.0f, f);
      }
    }
  }
  // a charged head's far DH sum joins the same FP64 accumulator, so its force is rounded once
  if (es && valid && sub == 0u && qi != 0.0f) {
    const uint32_t hh = b.head_of_slot[t];
    if (hh != SOUP_NONE) {
      f.x += b.far_acc[3u * hh];
      f.y += b.far_acc[3u * hh + 1u];
      f.z += b.far_acc[3u * hh + 2u];
    }
  }
  f.x = group_sum_d<FORCE_TPP>(f.x);
  f.y = group_sum_d<FORCE_TPP>(f.y);
  f.z = group_sum_d<FORCE_TPP>(f.z);
  if (valid && sub == 0u) b.force[t] = make_float4((float)f.x, (float)f.y, (float)f.z, 0.f);
}
__global__ __launch_bounds__(256) void k_force_group(SoupConfig c, DevBuf b, int es) { force_group_body(c, b, es, xcd_block()); }

// O(N^2) reference force: soup_force_brute_main (step.wgsl:277-292), test path only
__global__ __launch_bounds__(256) void k_force_brute(SoupConfig c, DevBuf b) {
  __shared__ PairTab T;
  load_pair_tab(c, T);
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.n) return;
  const PairConst P = make_pair_const(c);
  float4 xi = b.pos[t];
  float qi = b.charge[t];
  Acc f{0.0, 0.0, 0.0};
  bonded_force(c, P, b, t, xi, f);
  for (uint32_t u = 0; u < c.n; ++u) {
    if (u == t) continue;
    pair_force(P, T, xi, b.pos[u], qi, b.charge[u], f);
  }
  b.force[t] = make_float4((float)f.x, (float)f.y, (float)f.z, 0.f);
}

__device__ __forceinline__ void far_pair(const PairConst& P, float4 xi, float4 xj, float qi, float qj, Acc& f) {
#pragma clang fp contract(fast)
  float qq = qi * qj;
  if (qq == 0.0f) return;
  float dx = mi_acc(xi.x, xj.x, P.bx, P.hx), dy = mi_acc(xi.y, xj.y, P.by, P.hy), dz = mi_acc(xi.z, xj.z, P.bz, P.hz);
  float r2 = r2_fixed(dx, dy, dz);
  if (r2 < P.split2) return;
  float r = sqrtf(r2);
  if (r >= P.es_rc) return;
  float ir = 1.0f / r;
  float u = P.es_A * expf(-P.es_kappa * r) * ir;
  double fr = (double)(qq * (u * (ir + P.es_kappa) - P.es_Frc) * ir);
  f.x += fr * (double)dx;
  f.y += fr * (double)dy;
  f.z += fr * (double)dz;
}

// far DH over all pairs: soup_es_force_far_brute_main (electrostatics-long.wgsl:87-103), test only
__global__ __launch_bounds__(256) void k_es_far_brute(SoupConfig c, DevBuf b) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.n) return;
  float qi = b.charge[t];
  if (qi == 0.0f) return;
  const PairConst P = make_pair_const(c);
  float4 xi = b.pos[t];
  Acc f{0.0, 0.0, 0.0};
  for (uint32_t u = 0; u < c.n; ++u) {
    if (u == t) continue;
    far_pair(P, xi, b.pos[u], qi, b.charge[u], f);
  }
  float4 g = b.force[t];
  b.force[t] = make_float4((float)((double)g.x + f.x), (float)((double)g.y + f.y), (float)((double)g.z + f.z), 0.f);
}

// ================================================================================================
// Bond Monte Carlo (soup/wgsl/bond-*.wgsl).
//
// Phase A (k_bond_propose, one thread per slot): bondFormWalkList (bond-dispatch.wgsl:212-242) and
// the draws of bondFormDecide (:104-112). The candidate walk reads only positions, the list and the
// particle's own slots for partners j > i, and the draws depend only on those, so this phase is
// order-independent; it records the accepted proposals.
// Phase B (k_bond_resolve, one workgroup): the state-changing part of bondFormDecide (desorption
// valves, claims, roll-backs, centre bookkeeping) applied with exactly the result of running the
// particles in ascending index order, by deterministic reservations: every pending event reserves
// the particles it may read or write (atomicMin of its rank); an event whose reservations all hold
// commits. Events that share no particle commute; an event waits for every lower-ranked event it
// conflicts with.
// ================================================================================================
__device__ __forceinline__ int match_rule(const SoupConfig& c, float ti, float tj) {
  for (uint32_t r = 0; r < SOUP_BOND_RULES; ++r) {
    float a = c.kindA[r], bb = c.kindB[r];
    if ((ti == a && tj == bb) || (ti == bb && tj == a)) return (int)r;
  }
  return -1;
}
__device__ __forceinline__ uint32_t role_of(const SoupConfig& c, uint32_t r, float kind) {
  return kind == c.kindA[r] ? (uint32_t)c.roleA[r] : (uint32_t)c.roleB[r];
}

// one wavefront per particle: the lanes test 64 list entries at a time, and ballots keep the first
// catalyst within reach and the first SOUP_CAND_CAP candidates in the reference's walk order
__device__ __forceinline__ void bond_propose_body(SoupConfig c, DevBuf b, const uint32_t blk) {
  const uint32_t lane = threadIdx.x & (WAVE - 1);
  const uint32_t t = blk * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  if (t >= c.n) return;
  const float4 xi = b.pos[t];
  const float ti = xi.w;
  // a kind that no rule names (the solvent) finds no candidate and draws nothing: the walk only
  // decides its catalyst, which no proposal of it will read
  bool bondable = false;
  for (uint32_t r = 0; r < SOUP_BOND_RULES; ++r) bondable = bondable || ti == c.kindA[r] || ti == c.kindB[r];
  if (!bondable) {
    if (lane == 0) {
      b.nacc[t] = 0u;
      b.cat_id[t] = SOUP_NONE;
    }
    return;
  }
  const uint32_t i = b.id[t];
  const uint4 own = b.bs[t];
  const float rcat = wca_cut(c, pair_b(c, ti, c.catalyst_kind));
  // Every candidate and catalyst lies in the near part of the row. The near part visits the cells in
  // walk order but orders each cell by x, where the reference orders it by particle index, so the
  // choice goes by walk key (cell, then index): the catalyst of least key, and a running list of the
  // least-key candidates (lane k holds the k-th), each new candidate inserted at its rank. Once the
  // running choice lies in cells before the last cell a chunk reached, later entries cannot change it.
  uint32_t catId = SOUP_NONE, nCand = 0;
  uint64_t catKey = ~0ull;
  uint32_t myJ = SOUP_NONE, myR = 0;
  uint64_t myKey = ~0ull;
  const uint32_t cnt = b.vnear[t], ct = b.cell_of[t];
  const uint32_t* row = b.vl + (size_t)t * vl_stride(c);
  for (uint32_t base = 0; base < cnt; base += WAVE) {
    const uint32_t s = base + lane;
    bool isCat = false, isCand = false;
    uint32_t j = 0, u = 0;
    int r = -1;
    if (s < cnt) {
      u = row[s];
      const float4 xj = b.pos[u];
      const float tj = xj.w;
      j = b.id[u];
      if (tj == c.catalyst_kind) isCat = dist_f(c, xi, xj) < rcat;
      if (j > i) {
        r = match_rule(c, ti, tj);
        if (r >= 0 && own.x != j && own.y != j && own.z != j) isCand = dist_f(c, xi, xj) < wca_cut(c, pair_b(c, ti, tj));
      }
    }
    const uint32_t kc = (isCat || isCand) ? walk_cell(c, b.cell_of, ct, u) : 0u;
    uint64_t cm = __ballot(isCat);
    while (cm != 0ull) {  // catalysts in contact: a few at most
      const int k = __ffsll((unsigned long long)cm) - 1;
      const uint64_t kk = ((uint64_t)(uint32_t)__shfl((int)kc, k, WAVE) << 32) | (uint32_t)__shfl((int)j, k, WAVE);
      if (kk < catKey) {
        catKey = kk;
        catId = (uint32_t)kk;
      }
      cm &= cm - 1ull;
    }
    uint64_t dm = __ballot(isCand);
    while (dm != 0ull) {
      const int k = __ffsll((unsigned long long)dm) - 1;
      const uint32_t cj = (uint32_t)__shfl((int)j, k, WAVE), cr = (uint32_t)__shfl(r, k, WAVE);
      const uint64_t kk = ((uint64_t)(uint32_t)__shfl((int)kc, k, WAVE) << 32) | cj;
      const uint32_t p = (uint32_t)__popcll(__ballot(lane < nCand && myKey < kk));
      if (p < SOUP_CAND_CAP) {
        const uint32_t upJ = (uint32_t)__shfl_up((int)myJ, 1, WAVE), upR = (uint32_t)__shfl_up((int)myR, 1, WAVE);
        const uint64_t upK = ((uint64_t)(uint32_t)__shfl_up((int)(myKey >> 32), 1, WAVE) << 32) |
                             (uint32_t)__shfl_up((int)(uint32_t)myKey, 1, WAVE);
        if (lane > p) {
          myJ = upJ;
          myR = upR;
          myKey = upK;
        }
        if (lane == p) {
          myJ = cj;
          myR = cr;
          myKey = kk;
        }
        nCand = nCand < SOUP_CAND_CAP ? nCand + 1u : SOUP_CAND_CAP;
      }
      dm &= dm - 1ull;
    }
    if (catId != SOUP_NONE && nCand == SOUP_CAND_CAP) {
      const uint32_t last = (base + WAVE < cnt ? base + WAVE : cnt) - 1u;
      const uint32_t lastCell = uni(walk_cell(c, b.cell_of, ct, row[last]));
      const uint32_t candCell = (uint32_t)__shfl((int)(myKey >> 32), (int)SOUP_CAND_CAP - 1, WAVE);
      if ((uint32_t)(catKey >> 32) < lastCell && candCell < lastCell) break;
    }
  }
  const uint32_t catNear = catId != SOUP_NONE ? 1u : 0u;
  // bondFormDecide draws (bond-dispatch.wgsl:104-112), serial in candidate order on every lane
  uint32_t rng = b.brng[t], nacc = 0;
  for (uint32_t k = 0; k < nCand; ++k) {
    const uint32_t cj = (uint32_t)__shfl((int)myJ, (int)k, WAVE), r = (uint32_t)__shfl((int)myR, (int)k, WAVE);
    if (c.req_cat[r] > 0.5f && catNear == 0u) continue;
    rng = pcg_soup(rng);
    if (uniform01(rng) >= c.att_form[r]) continue;
    rng = pcg_soup(rng);
    if (uniform01(rng) >= c.acc_form[r]) continue;
    if (lane == 0) {
      b.props[(size_t)t * SOUP_CAND_CAP + nacc] = make_uint2(cj, r);
      // a propagation proposal may give i, j and the catalyst a centre link: mark them
      const float tj = b.pos[b.inv[cj]].w;
      if (role_of(c, r, ti) == 0u && role_of(c, r, tj) == 0u) {
        b.flag_cc[i] = 1u;
        b.flag_cc[cj] = 1u;
        if (catId != SOUP_NONE) b.flag_cc[catId] = 1u;
      }
    }
    nacc++;
  }
  if (lane == 0) {
    b.brng[t] = rng;
    b.nacc[t] = nacc;
    b.cat_id[t] = catId;
  }
}
__global__ __launch_bounds__(256) void k_bond_propose(SoupConfig c, DevBuf b) { bond_propose_body(c, b, xcd_block()); }

// drops termination proposals that cannot succeed in any order (the carbon has no centre and no
// propagation proposal can give it one), then flags the particles that carry an event.
__device__ __forceinline__ void bond_filter_body(SoupConfig c, DevBuf b, const uint32_t blk) {
  uint32_t t = blk * blockDim.x + threadIdx.x;
  if (t >= c.n) return;
  const uint32_t i = b.id[t];
  const float ti = b.pos[t].w;
  uint32_t nacc = b.nacc[t], keep = 0;
  for (uint32_t k = 0; k < nacc; ++k) {
    uint2 p = b.props[(size_t)t * SOUP_CAND_CAP + k];
    uint32_t j = p.x, r = p.y;
    uint32_t uj = b.inv[j];
    float tj = b.pos[uj].w;
    uint32_t roleI = role_of(c, r, ti), roleJ = role_of(c, r, tj);
    bool live = true;
    if (!(roleI == 0u && roleJ == 0u)) {
      uint32_t carbon = roleI == 1u ? i : j;
      uint32_t cs = roleI == 1u ? t : uj;
      live = b.clink[cs] != SOUP_NONE || b.flag_cc[carbon] != 0u;
    }
    if (live) b.props[(size_t)t * SOUP_CAND_CAP + keep++] = p;
  }
  b.nacc[t] = keep;
  if (keep > 0u || b.clink[t] != SOUP_NONE || b.flag_cc[i] != 0u) b.ev_flag[i] = 1u;
}
__global__ __launch_bounds__(256) void k_bond_filter(SoupConfig c, DevBuf b) { bond_filter_body(c, b, blockIdx.x); }

__device__ __forceinline__ void bond_events_body(SoupConfig c, DevBuf b, const uint32_t blk) {
  uint32_t i = blk * blockDim.x + threadIdx.x;
  if (i >= c.n) return;
  if (b.ev_flag[i]) b.ev_id[b.ev_pos[i]] = i;
}
__global__ __launch_bounds__(256) void k_bond_events(SoupConfig c, DevBuf b) { bond_events_body(c, b, blockIdx.x); }

// ---- phase B helpers: all shared state accessed with device-scope atomics ----------------------
__device__ __forceinline__ uint32_t ld(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool cas_u(uint32_t* p, uint32_t expected, uint32_t v) {
  uint32_t e = expected;
  return __hip_atomic_compare_exchange_strong(p, &e, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct BondCtx {
  const SoupConfig& c;
  const DevBuf& b;
  __device__ uint32_t* slots(uint32_t id) const { return reinterpret_cast<uint32_t*>(&b.bs[b.inv[id]]); }
  __device__ uint32_t* cl(uint32_t id) const { return &b.clink[b.inv[id]]; }
  __device__ uint32_t* held(uint32_t id) const { return &b.cheld[b.inv[id]]; }
  __device__ float4 pos(uint32_t id) const { return b.pos[b.inv[id]]; }
  __device__ float dist(uint32_t a, uint32_t bb) const { return dist_f(c, pos(a), pos(bb)); }
  __device__ bool contact_cat(uint32_t a, uint32_t cat) const {
    return dist(a, cat) < wca_cut(c, pair_b(c, pos(a).w, c.catalyst_kind));
  }
  __device__ bool within_reach(uint32_t a, uint32_t owner) const {
    return dist(a, owner) < wca_cut(c, pair_b(c, pos(a).w, c.catalyst_kind)) + c.wc;
  }
};

// tryClaimSlot (bond-valence.wgsl:208-266)
__device__ int try_claim(const BondCtx& X, uint32_t particle, uint32_t role, uint32_t partner) {
  uint32_t* sl = X.slots(particle);
  bool hto = X.c.head_terminal_only > 0.5f;
  if (role == 0u) {
    if (cas_u(&sl[0], SOUP_NONE, partner)) {
      if (hto && ld(&sl[2]) != SOUP_NONE) {
        st(&sl[0], SOUP_NONE);
        return -1;
      }
      return 0;
    }
    if (cas_u(&sl[1], SOUP_NONE, partner)) {
      if (hto && ld(&sl[2]) != SOUP_NONE) {
        st(&sl[1], SOUP_NONE);
        return -1;
      }
      return 1;
    }
    return -1;
  } else if (role == 1u) {
    if (cas_u(&sl[2], SOUP_NONE, partner)) {
      if (hto && ld(&sl[0]) != SOUP_NONE && ld(&sl[1]) != SOUP_NONE) {
        st(&sl[2], SOUP_NONE);
        return -1;
      }
      return 2;
    }
    return -1;
  }
  uint32_t cap = (uint32_t)X.c.head_chain_cap;
  for (uint32_t k = 0; k < cap; ++k)
    if (cas_u(&sl[k], SOUP_NONE, partner)) return (int)k;
  return -1;
}

__device__ void release_slot(const BondCtx& X, uint32_t particle, int slot) {
  if (slot >= 0) st(&X.slots(particle)[slot], SOUP_NONE);
}

// propagateOnCenter (bond-adsorption.wgsl:87-248)
__device__ bool propagate(const BondCtx& X, uint32_t i, uint32_t j, int slotI, int slotJ, uint32_t catId) {
  if (slotI == 0 && slotJ == 0) {
    if (catId == SOUP_NONE) return false;
    if (!X.contact_cat(i, catId)) return false;
    if (!cas_u(X.cl(catId), SOUP_NONE, i)) return false;
    if (!cas_u(X.cl(i), SOUP_NONE, catId)) {
      cas_u(X.cl(catId), i, SOUP_NONE);
      return false;
    }
    st(X.held(catId), 0u);
    return true;
  }
  uint32_t tipOld, tipNew;
  if (slotI == 1 && slotJ == 0) {
    tipOld = i;
    tipNew = j;
  } else if (slotI == 0 && slotJ == 1) {
    tipOld = j;
    tipNew = i;
  } else {
    return false;
  }
  uint32_t owner = ld(X.cl(tipOld));
  bool reclaiming = false;
  if (owner == SOUP_NONE) {
    if (catId == SOUP_NONE) return false;
    owner = catId;
    reclaiming = true;
  }
  if (reclaiming) {
    if (!X.contact_cat(tipOld, catId)) return false;
  } else {
    if (!X.within_reach(tipOld, owner)) return false;
  }
  if (!X.contact_cat(tipNew, owner)) return false;
  if (reclaiming) {
    if (!cas_u(X.cl(owner), SOUP_NONE, tipOld)) return false;
    if (!cas_u(X.cl(tipOld), SOUP_NONE, owner)) {
      cas_u(X.cl(owner), tipOld, SOUP_NONE);
      return false;
    }
  }
  if (!cas_u(X.cl(owner), tipOld, tipNew)) {
    if (reclaiming) cas_u(X.cl(tipOld), owner, SOUP_NONE);
    return false;
  }
  if (!cas_u(X.cl(tipNew), SOUP_NONE, owner)) {
    cas_u(X.cl(owner), tipNew, tipOld);
    if (reclaiming) {
      cas_u(X.cl(owner), tipOld, SOUP_NONE);
      cas_u(X.cl(tipOld), owner, SOUP_NONE);
    }
    return false;
  }
  cas_u(X.cl(tipOld), owner, SOUP_NONE);
  st(X.held(owner), 0u);
  return true;
}

// terminateOnCenter (bond-adsorption.wgsl:256-268)
__device__ bool terminate(const BondCtx& X, uint32_t carbon) {
  uint32_t owner = ld(X.cl(carbon));
  if (owner == SOUP_NONE) return false;
  if (!X.within_reach(carbon, owner)) return false;
  if (!cas_u(X.cl(carbon), owner, SOUP_NONE)) return false;
  cas_u(X.cl(owner), carbon, SOUP_NONE);
  st(X.held(owner), 0u);
  return true;
}

// the whole per-particle turn of bondFormDecide after the draws
__device__ void bond_event(const BondCtx& X, uint32_t i) {
  const SoupConfig& c = X.c;
  const DevBuf& b = X.b;
  const uint32_t t = b.inv[i];
  const float ti = b.pos[t].w;
  if (ti == c.catalyst_kind) {
    // desorbTimeout (bond-adsorption.wgsl:304-316)
    uint32_t tip = ld(X.cl(i));
    if (tip != SOUP_NONE) {
      uint32_t inc = (uint32_t)c.bond_interval_f;
      uint32_t prev = ld(X.held(i));
      st(X.held(i), prev + inc);
      if (prev + inc > (uint32_t)c.max_hold_steps) {
        if (cas_u(X.cl(i), tip, SOUP_NONE)) {
          cas_u(X.cl(tip), i, SOUP_NONE);
          st(X.held(i), 0u);
          atomicAdd(&b.desorb[1], 1u);
        }
      }
    }
  } else {
    // desorbStretch (bond-adsorption.wgsl:283-292)
    uint32_t owner = ld(X.cl(i));
    if (owner != SOUP_NONE && X.dist(i, owner) >= c.r_inf) {
      if (cas_u(X.cl(i), owner, SOUP_NONE)) {
        cas_u(X.cl(owner), i, SOUP_NONE);
        st(X.held(owner), 0u);
        atomicAdd(&b.desorb[0], 1u);
      }
    }
  }
  const uint32_t nacc = b.nacc[t], catId = b.cat_id[t];
  for (uint32_t k = 0; k < nacc; ++k) {
    uint2 p = b.props[(size_t)t * SOUP_CAND_CAP + k];
    uint32_t j = p.x, r = p.y;
    float tj = X.pos(j).w;
    uint32_t roleI = role_of(c, r, ti), roleJ = role_of(c, r, tj);
    int slotI = try_claim(X, i, roleI, j);
    if (slotI < 0) continue;
    int slotJ = try_claim(X, j, roleJ, i);
    if (slotJ < 0) {
      release_slot(X, i, slotI);
      continue;
    }
    bool ok;
    if (roleI == 0u && roleJ == 0u) ok = propagate(X, i, j, slotI, slotJ, catId);
    else if (roleI == 1u) ok = terminate(X, i);
    else ok = terminate(X, j);
    if (!ok) {
      release_slot(X, i, slotI);
      release_slot(X, j, slotJ);
      continue;
    }
    atomicAdd(&b.events[r * 2u], 1u);
  }
}

// every particle index the event of particle i may read or write (superset)
__device__ uint32_t bond_event_resources(const BondCtx& X, uint32_t i, uint32_t* R) {
  const DevBuf& b = X.b;
  const uint32_t t = b.inv[i];
  uint32_t m = 0;
  auto add = [&](uint32_t x) {
    if (x == SOUP_NONE) return;
    for (uint32_t k = 0; k < m; ++k)
      if (R[k] == x) return;
    R[m++] = x;
  };
  add(i);
  add(ld(X.cl(i)));
  const uint32_t nacc = b.nacc[t];
  if (nacc > 0u) {
    add(b.cat_id[t]);
    for (uint32_t k = 0; k < nacc; ++k) {
      uint32_t j = b.props[(size_t)t * SOUP_CAND_CAP + k].x;
      add(j);
      add(ld(X.cl(j)));
    }
  }
  return m;
}

__device__ __forceinline__ void phase_barrier() {
  __threadfence();
  __syncthreads();
}

__device__ __forceinline__ void bond_resolve_body(SoupConfig c, DevBuf b, const uint32_t blk) {
  __shared__ uint32_t s_pending;
  const uint32_t nev = b.ev_pos[c.n];
  const BondCtx X{c, b};
  for (uint32_t e = threadIdx.x; e < nev; e += blockDim.x) st(&b.ev_state[e], 0u);
  phase_barrier();
  for (uint32_t round = 1;; ++round) {
    // reserve
    for (uint32_t e = threadIdx.x; e < nev; e += blockDim.x) {
      if (ld(&b.ev_state[e]) & 1u) continue;
      uint32_t* R = b.rset + (size_t)e * SOUP_RSET;
      uint32_t m = bond_event_resources(X, b.ev_id[e], R + 1);
      R[0] = m;
      for (uint32_t k = 0; k < m; ++k) atomicMin(&b.resv[R[1 + k]], e);
      st(&b.ev_state[e], round << 8);
    }
    phase_barrier();
    // commit: an event whose reservations all hold executes its whole turn
    for (uint32_t e = threadIdx.x; e < nev; e += blockDim.x) {
      uint32_t s = ld(&b.ev_state[e]);
      if ((s & 1u) || (s >> 8) != round) continue;
      const uint32_t* R = b.rset + (size_t)e * SOUP_RSET;
      bool mine = true;
      for (uint32_t k = 0; k < R[0]; ++k)
        if (ld(&b.resv[R[1 + k]]) != e) {
          mine = false;
          break;
        }
      if (mine) {
        bond_event(X, b.ev_id[e]);
        st(&b.ev_state[e], s | 1u);
      }
    }
    if (threadIdx.x == 0) s_pending = 0;
    phase_barrier();
    // release every reservation of this round, count what is left
    uint32_t pending = 0;
    for (uint32_t e = threadIdx.x; e < nev; e += blockDim.x) {
      uint32_t s = ld(&b.ev_state[e]);
      if ((s >> 8) == round) {
        const uint32_t* R = b.rset + (size_t)e * SOUP_RSET;
        for (uint32_t k = 0; k < R[0]; ++k) st(&b.resv[R[1 + k]], SOUP_NONE);
      }
      if (!(s & 1u)) pending++;
    }
    if (pending) atomicAdd(&s_pending, pending);
    phase_barrier();
    uint32_t left = s_pending;
    __syncthreads();
    if (left == 0u) break;
    // the lowest pending event always commits, so nev rounds always suffice
    if (round > nev + 2u) {
      if (threadIdx.x == 0) st(&b.flags[6], 1u);
      break;
    }
  }
}
__global__ __launch_bounds__(1024) void k_bond_resolve(SoupConfig c, DevBuf b) { bond_resolve_body(c, b, blockIdx.x); }

// bond_break_main (bond-dispatch.wgsl:170-200): each particle breaks only bonds to partners >= i,
// and releases the partner's slot holding i, which the partner's own pass never reads.
__device__ __forceinline__ void bond_break_body(SoupConfig c, DevBuf b, const uint32_t blk) {
  uint32_t t = blk * blockDim.x + threadIdx.x;
  if (t >= c.n) return;
  const uint32_t i = b.id[t];
  const float ti = b.pos[t].w;
  uint32_t* sl = reinterpret_cast<uint32_t*>(&b.bs[t]);
  uint32_t rng = b.brng[t];
  for (uint32_t k = 0; k < 3; ++k) {
    uint32_t p = ld(&sl[k]);
    if (p == SOUP_NONE || p < i) continue;
    uint32_t up = b.inv[p];
    int r = match_rule(c, ti, b.pos[up].w);
    if (r < 0) continue;
    rng = pcg_soup(rng);
    if (uniform01(rng) >= c.att_break[r]) continue;
    rng = pcg_soup(rng);
    if (uniform01(rng) >= c.acc_break[r]) continue;
    st(&sl[k], SOUP_NONE);
    uint32_t* ps = reinterpret_cast<uint32_t*>(&b.bs[up]);
    for (uint32_t q = 0; q < 3; ++q)
      if (cas_u(&ps[q], i, SOUP_NONE)) break;
    atomicAdd(&b.events[(uint32_t)r * 2u + 1u], 1u);
  }
  b.brng[t] = rng;
}
__global__ __launch_bounds__(256) void k_bond_break(SoupConfig c, DevBuf b) { bond_break_body(c, b, blockIdx.x); }

// ================================================================================================
// Constant-pH Monte Carlo (soup/src/electrostatics.ts:435-496, acid-soap.ts:170-198) on the device.
// k_proton_pairs tabulates, for every head-list entry, the FP64 unit-charge screened-Coulomb energy
// (0 beyond rc) and the acid-soap well (0 beyond its reach) at the current positions, with the
// reference's double-precision minimum image (Math.round). k_proton_chain then runs the serial
// chain in one workgroup: per attempt, the site draw, an FP64 reduction of the tabulated terms
// weighted by the current charges, the Metropolis test with the draw only when dG > 0.
// ================================================================================================
__device__ __forceinline__ double js_round(double x) {
  double r = floor(x);
  if (x - r >= 0.5) r += 1.0;
  return r;
}
__device__ __forceinline__ double pcg_next_u(uint32_t& s) {
  uint32_t v = s * 747796405u + 2891336453u;
  s = v;
  v = ((v >> ((v >> 28) + 4u)) ^ v) * 277803737u;
  v = (v >> 22) ^ v;
  return (double)v / 4294967296.0;
}

// Per head row h: tabE[h * cap + s] = the unit-charge screened-Coulomb energy of list entry s (the
// entry's head index is the DH list's own, es[h * cap + s]); the acid-soap well, nonzero only within
// its short reach, is compacted in list order into tabW / tabWg[h * cap + k], k < wcnt[h].
__global__ __launch_bounds__(256) void k_proton_pairs(SoupConfig c, DevBuf b, uint32_t nh, double* tabE, double* tabW,
                                                      uint32_t* tabWg, uint32_t* wcnt) {
  const uint32_t lane = threadIdx.x & (WAVE - 1);
  const uint32_t h = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  if (h >= nh) return;
  const float4 xi = b.pos[b.head_slot[h]];
  const uint32_t cnt = b.es_count[h];
  const double bx = c.box[0], by = c.box[1], bz = c.box[2];
  const double cut2 = c.ph_cutoff * c.ph_cutoff;
  const double reach = c.as_rc + c.as_wc, reach2 = reach * reach;
  const uint64_t lower = (1ull << lane) - 1ull;
  uint32_t nw = 0;  // wells so far in this row
  for (uint32_t s0 = 0; s0 < cnt; s0 += WAVE) {
    const uint32_t s = s0 + lane;
    double e = 0.0, w = 0.0;
    uint32_t g = 0;
    if (s < cnt) {
      g = b.es[(size_t)h * c.es_cap + s];
      const float4 xj = b.pos[b.head_slot[g]];
      double dx = (double)xi.x - (double)xj.x, dy = (double)xi.y - (double)xj.y, dz = (double)xi.z - (double)xj.z;
      dx = dx - js_round(dx / bx) * bx;
      dy = dy - js_round(dy / by) * by;
      dz = dz - js_round(dz / bz) * bz;
      double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 < cut2 && c.ph_A != 0.0) {
        double r = sqrt(r2);
        if (r > 0.0 && r < c.ph_cutoff) {
          double u1 = (c.ph_A * exp(-c.ph_kappa * r)) / r;
          e = u1 - c.ph_shiftU + c.ph_shiftF * (r - c.ph_cutoff);
        }
      }
      if (c.as_scale > 0.0 && r2 < reach2 && r2 > 0.0) {
        double r = sqrt(r2);
        if (r < c.as_rc) w = -c.as_scale * c.as_eps;
        else if (r > c.as_rc + c.as_wc) w = 0.0;
        else {
          double cc = cos((3.141592653589793 * (r - c.as_rc)) / (2.0 * c.as_wc));
          w = -c.as_scale * c.as_eps * cc * cc;
        }
      }
      tabE[(size_t)h * c.es_cap + s] = e;
    }
    // a zero well adds exactly nothing to the chain's sum, so only the nonzero ones are kept
    const uint64_t mw = __builtin_amdgcn_ballot_w64(w != 0.0);
    if (w != 0.0) {
      const uint32_t k = nw + (uint32_t)__popcll(mw & lower);
      tabW[(size_t)h * c.es_cap + k] = w;
      tabWg[(size_t)h * c.es_cap + k] = g;
    }
    nw += (uint32_t)__popcll(mw);
  }
  if (lane == 0) wcnt[h] = nw;
}

#define CHAIN_WAVES 16u
#define CHAIN_THREADS (CHAIN_WAVES * WAVE)
#ifndef CHAIN_WINDOW
#define CHAIN_WINDOW 16u  // stream positions whose sites one round evaluates (bits 0..30 of a head word)
#endif
#define CHAIN_REG 15u  // row entries per lane requested together: rows of up to 960 entries in one round trip
static_assert(CHAIN_WINDOW >= 1u && CHAIN_WINDOW <= 31u, "a head word holds the window's positions in bits 0..30");

// exp(-dG / kT) out of line: inlined into the chain's round loop, the exponential's constants are
// hoisted into vector registers for the whole loop and spilled by the row loads
__device__ __attribute__((noinline)) double chain_boltzmann(double dG, double kT) { return exp(-dG / kT); }

// OR over the 64 lanes: the same DPP butterfly inside each row of 16 lanes as wave_sum_d_dpp, then
// the four row results read from lanes 0, 16, 32, 48
__device__ __forceinline__ uint32_t wave_or_dpp(uint32_t v) {
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);  // row_half_mirror
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);  // row_mirror
  return (uint32_t)(__builtin_amdgcn_readlane((int)v, 0) | __builtin_amdgcn_readlane((int)v, 16) |
                    __builtin_amdgcn_readlane((int)v, 32) | __builtin_amdgcn_readlane((int)v, 48));
}

// The PCG state after k draws is an affine map of the state: s_k = jm[k] s_0 + ja[k] (mod 2^32),
// with jm[k] = a^k and ja[k] = c (a^(k-1) + ... + 1) for the step s -> a s + c of pcg_next_u.
__device__ __forceinline__ void chain_jump(uint32_t k, uint32_t& jm, uint32_t& ja) {
  jm = 1u;
  ja = 0u;
  for (uint32_t i = 0; i < k; ++i) {
    jm *= 747796405u;
    ja = ja * 747796405u + 2891336453u;
  }
}
// the round's stream from state st, on wavefront 0: lane l <= CHAIN_WINDOW + 1 writes the state
// after l draws (from its jump constants jm, ja for l draws), lanes <= CHAIN_WINDOW draw l, and
// lanes < CHAIN_WINDOW mark their site's head word with bit l
__device__ __forceinline__ void chain_round_stream(uint32_t st, uint32_t nh, uint32_t lane, uint32_t jm, uint32_t ja,
                                                   uint32_t* hword, uint32_t* s_site, uint32_t* s_state, double* s_u) {
  if (lane <= CHAIN_WINDOW + 1u) {
    uint32_t sr = jm * st + ja;
    s_state[lane] = sr;
    if (lane <= CHAIN_WINDOW) {
      const double u = pcg_next_u(sr);
      s_u[lane] = u;
      if (lane < CHAIN_WINDOW) {
        const double hf = floor(u * (double)nh);
        const uint32_t site = hf < (double)(nh - 1) ? (uint32_t)hf : nh - 1;
        s_site[lane] = site;
        atomicOr(&hword[site], 1u << lane);
      }
    }
  }
}

// stats: [0] attempts, [1] accepted, [2] deprotonated, [3] dEsCount; dstats: [0] dEsSum
// The serial chain of protonationSweep (electrostatics.ts:435-496) in one workgroup, in rounds.
// The PCG stream is fixed; only how many numbers an attempt consumes (one, or two when dG > 0)
// depends on the decisions. A round therefore takes the next CHAIN_WINDOW stream positions, reads
// each one's draw as a site, and evaluates all those sites' work sums at once (one wavefront per
// position) with the charges as they stand at the round start. Wavefront 0 then walks the attempts
// in stream order: an attempt's evaluation is exact unless its row (or its own head) holds a head
// whose charge an earlier attempt of this round flipped, so the walk commits attempts until the
// first such one, which the next round evaluates again from the updated charges. The decisions,
// the RNG consumption and the dEs sum are those of the attempt-by-attempt chain. Two workgroup
// barriers per round: after the stream set-up (wavefront 0), and after the evaluation.
// Head word in LDS: bit 31 = charged (chargeDeprotonated), bit r = stream position r of the round
// draws this head as its site.
__global__ __launch_bounds__(CHAIN_THREADS) void k_proton_chain(SoupConfig c, DevBuf b, uint32_t nh, const double* tabE,
                                                                const double* tabW, const uint32_t* tabWg, const uint32_t* wcnt,
                                                                uint32_t* rng_io, uint32_t* stats, double* dstats) {
  extern __shared__ uint32_t hword[];
  __shared__ uint32_t s_site[CHAIN_WINDOW], s_M[CHAIN_WINDOW], s_state[CHAIN_WINDOW + 2u], s_ctl[2];
  __shared__ double s_u[CHAIN_WINDOW + 1u], s_se[CHAIN_WINDOW], s_sa[CHAIN_WINDOW];
  const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
  for (uint32_t g = tid; g < nh; g += CHAIN_THREADS) hword[g] = b.charge[b.head_slot[g]] != 0.0f ? 0x80000000u : 0u;
  __syncthreads();
  const double kT = c.ph_kT, q = c.ph_q, qq = q * (double)(float)q;
  const double dGintr = kT * 2.302585092994046 * (c.ph_pKa - c.ph_pH);
  const bool soap = c.as_scale > 0.0;
  const uint32_t cap = c.es_cap, lim = cap < WAVE * CHAIN_REG ? cap : WAVE * CHAIN_REG;
  uint32_t accepted = 0, dEsCount = 0;  // wavefront 0's
  double dEsSum = 0.0;
  uint32_t jm = 1u, ja = 0u;
  if (w == 0) {
    chain_jump(lane <= CHAIN_WINDOW + 1u ? lane : 0u, jm, ja);
    if (lane == 0) {
      s_ctl[0] = 0;
      s_ctl[1] = *rng_io;
    }
    if (nh > 0) chain_round_stream(*rng_io, nh, lane, jm, ja, hword, s_site, s_state, s_u);
  }
  __syncthreads();
  uint32_t done = s_ctl[0];
  while (done < nh) {
    // ---- evaluate: wavefront w takes positions w, w + CHAIN_WAVES, ... ----
    for (uint32_t r = w; r < CHAIN_WINDOW; r += CHAIN_WAVES) {
      const uint32_t h = uni(s_site[r]), below = (1u << r) - 1u;  // uniform: the row bases in scalar registers
      const size_t row = (size_t)h * cap;
      // every lane requests CHAIN_REG entries and one acid-soap well with no branch (indices past
      // the register window clamped to its last entry; lim >= 1 here, since soup_create refuses
      // charged heads with a DH capacity of 0), so the loads issue back to back; entries past the
      // counts are never used
      double ee[CHAIN_REG];
      uint32_t gg[CHAIN_REG];
#pragma unroll
      for (uint32_t k = 0; k < CHAIN_REG; ++k) {
        const uint32_t s = lane + k * WAVE < lim ? lane + k * WAVE : lim - 1u;
        ee[k] = tabE[row + s];
        gg[k] = b.es[row + s];
      }
      const uint32_t sw = lane < lim ? lane : lim - 1u;
      const double wv = tabW[row + sw];
      const uint32_t wg = tabWg[row + sw];
      const uint32_t cnt = uni(b.es_count[h]), nwell = uni(wcnt[h]);
      double se = 0.0, sa = 0.0;
      uint32_t m = 0;
      // an unused entry adds exactly 0 (the sums start at +0 and take no exact -0), which leaves
      // them as the skipped entry would
#pragma unroll
      for (uint32_t k = 0; k < CHAIN_REG; ++k) {
        const bool ok = lane + k * WAVE < cnt && lane + k * WAVE < lim;
        const uint32_t wd = hword[ok ? gg[k] : h];
        se += (ok && (wd >> 31) != 0u) ? qq * ee[k] : 0.0;
        m |= ok ? (wd & below) : 0u;
      }
      for (uint32_t s = lim + lane; s < cnt; s += WAVE) {  // rows longer than the register window
        const uint32_t wd = hword[b.es[row + s]];
        if ((wd >> 31) != 0u) se += qq * tabE[row + s];
        m |= wd & below;
      }
      if (soap) {
        const bool okw = lane < nwell && lane < lim;
        const uint32_t wdw = hword[okw ? wg : h];
        sa += okw ? ((wdw >> 31) != 0u ? -wv : wv) : 0.0;
        for (uint32_t s = (lim < WAVE ? lim : WAVE) + lane; s < nwell; s += WAVE)  // more than 64 wells
          sa += (hword[tabWg[row + s]] >> 31) != 0u ? -tabW[row + s] : tabW[row + s];
      }
      se = wave_sum_d_dpp(se);
      sa = wave_sum_d_dpp(sa);
      m = wave_or_dpp(m);
      if (lane == 0) {
        s_se[r] = se;
        s_sa[r] = sa;
        s_M[r] = m | (hword[h] & below);  // earlier positions on this row's heads, or on this head
      }
    }
    __syncthreads();
    if (w == 0) {
      // ---- lane p holds position p: its site, dG, acceptance probability and conflict mask ----
      const bool live = lane < CHAIN_WINDOW;
      const uint32_t site = live ? s_site[lane] : 0u;
      const bool charged = live && (hword[site] >> 31) != 0u;
      const double uEs = live ? s_se[lane] + (soap ? s_sa[lane] : 0.0) : 0.0;
      const double dG = charged ? -(dGintr + uEs) : dGintr + uEs;
      const double pacc = (live && !(dG <= 0.0)) ? chain_boltzmann(dG, kT) : 1.0;
      const double unext = lane < CHAIN_WINDOW ? s_u[lane + 1u] : 0.0;  // the acceptance draw if p draws
      const uint32_t M = live ? s_M[lane] : 0u;
      const uint32_t drewm = (uint32_t)__builtin_amdgcn_ballot_w64(live && !(dG <= 0.0));
      const uint32_t passm = (uint32_t)__builtin_amdgcn_ballot_w64(live && unext < pacc);
      const uint32_t chgm = (uint32_t)__builtin_amdgcn_ballot_w64(charged);
      // ---- commit in stream order up to the first attempt an earlier flip of this round touches:
      // a scalar walk over the lanes' masks ----
      uint32_t p = 0, flipped = 0, nd = uni(done);
      while (nd < nh && p < CHAIN_WINDOW) {
        const uint32_t Mp = uni((uint32_t)__builtin_amdgcn_readlane((int)M, (int)p));
        if ((Mp & flipped) != 0u) break;
        const bool drew = (drewm >> p) & 1u;
        const bool acc = drew ? ((passm >> p) & 1u) != 0u : true;
        if (acc) flipped |= 1u << p;
        ++nd;
        p += drew ? 2u : 1u;
      }
      // the flips, their counts and the dEs sum in stream order
      if (live && ((flipped >> lane) & 1u) != 0u) atomicXor(&hword[site], 0x80000000u);  // lane < 32: no shift wraps
      accepted += (uint32_t)__popc(flipped);
      uint32_t dep = flipped & ~chgm;  // protonated -> deprotonated
      dEsCount += (uint32_t)__popc(dep);
      while (dep) {
        const uint32_t k = (uint32_t)__builtin_ctz(dep);
        dep &= dep - 1u;
        dEsSum += readlane_d(uEs, (int)k) / kT;
      }
      done = nd;
      const uint32_t st = s_state[p];
      wave_lds_sync();
      // ---- the next round's stream: this round's marks off, the next ones on ----
      if (lane < CHAIN_WINDOW) atomicAnd(&hword[site], ~(1u << lane));
      wave_lds_sync();
      if (lane == 0) {
        s_ctl[0] = nd;
        s_ctl[1] = st;
      }
      if (nd < nh) chain_round_stream(st, nh, lane, jm, ja, hword, s_site, s_state, s_u);
    }
    __syncthreads();
    done = s_ctl[0];
  }
  __syncthreads();
  uint32_t dep = 0;
  for (uint32_t g = tid; g < nh; g += CHAIN_THREADS) {
    const bool on = (hword[g] >> 31) != 0u;
    b.charge[b.head_slot[g]] = on ? (float)q : 0.0f;
    dep += on ? 1u : 0u;
  }
  dep = wave_sum_u(dep);
  if (lane == 0) atomicAdd(&stats[2], dep);
  if (tid == 0) {
    *rng_io = s_ctl[1];
    stats[0] = nh;
    stats[1] = accepted;
    stats[3] = dEsCount;
    dstats[0] = dEsSum;
  }
}

// ================================================================================================
// Boundary translation between slot order and particle-index order
// ================================================================================================
// ================================================================================================
// soupPotential (soup/src/soup-potential.ts:145-301) in FP64, one wavefront per particle. Each
// unordered pair is taken once, from the side with the smaller particle index (the reference's
// j > i), over the current Verlet list; the bonded terms of soup-potential.ts:254-286 are taken by
// the particle that owns them (FENE by the lower index of a bond, the bend by the middle bead, the
// tether by the lower index of a link); charged heads add esTotalEnergy's pairs (electrostatics.ts:
// 575-600) from their DH row. Per-particle sums go to part[slot] and k_potential_sum reduces them
// in a fixed order.
// ================================================================================================
__device__ __forceinline__ double pot_mi(double d, double L) { return d - js_round(d / L) * L; }
__device__ __forceinline__ double pot_dist2(const SoupConfig& c, float4 a, float4 b) {
  const double dx = pot_mi((double)a.x - (double)b.x, c.box[0]);
  const double dy = pot_mi((double)a.y - (double)b.y, c.box[1]);
  const double dz = pot_mi((double)a.z - (double)b.z, c.box[2]);
  return dx * dx + dy * dy + dz * dz;
}
__device__ __forceinline__ double pot_fene(const SoupPotentialBasis& pb, double r) {
  const double x = r / pb.fene_r_inf;
  return -0.5 * pb.fene_k * pb.fene_r_inf * pb.fene_r_inf * log(1.0 - x * x);
}
__device__ __forceinline__ double pot_attr(const SoupPotentialBasis& pb, double r) {
  if (r < pb.rc_attr) return -pb.epsilon;
  if (r > pb.rc_attr + pb.wc) return 0.0;
  const double cc = cos((3.141592653589793 * (r - pb.rc_attr)) / (2.0 * pb.wc));
  return -pb.epsilon * cc * cc;
}

__global__ __launch_bounds__(256) void k_potential(SoupConfig c, DevBuf b, SoupPotentialBasis pb, double4* part, int es) {
  const uint32_t lane = threadIdx.x & (WAVE - 1), w = uni(threadIdx.x / WAVE);
  const uint32_t t = uni(blockIdx.x * (blockDim.x / WAVE) + w);
  if (t >= c.n) return;
  const float4 xv = b.pos[t];
  const float4 xi = make_float4(uni_f(xv.x), uni_f(xv.y), uni_f(xv.z), uni_f(xv.w));
  const uint32_t idi = uni(b.id[t]);
  const uint32_t ki = (uint32_t)xi.w & (SOUP_MAX_KINDS - 1u);
  const float qi = uni_f(b.charge[t]);
  const double cut2 = pb.cutoff * pb.cutoff;
  double nb = 0.0, fene = 0.0, bend = 0.0, teth = 0.0;
  const uint32_t cnt = uni(b.vcount[t]), cn = uni(b.vnear[t]);
  const uint32_t* row = b.vl + (size_t)t * vl_stride(c);
  for (uint32_t s = lane; s < cnt; s += WAVE) {
    const uint32_t u = row[row_index(s, cn, c.list_cap)];
    if (b.id[u] <= idi) continue;
    const float4 xj = b.pos[u];
    const double r2 = pot_dist2(c, xi, xj);
    if (r2 > cut2 || r2 <= 0.0) continue;
    const double r = sqrt(r2);
    const uint32_t kj = (uint32_t)xj.w & (SOUP_MAX_KINDS - 1u);
    const double bb = (pb.b_radius[ki] + pb.b_radius[kj]) / 2.0;
    if (r < pb.wca_factor * bb) {
      const double sr = bb / r, s2 = sr * sr, s6 = s2 * s2 * s2, q = 2.0 * s6 - 1.0;
      nb += pb.epsilon * q * q;
    }
    double sc = pb.attr[pb.cls[ki] & 3u][pb.cls[kj] & 3u];
    if (pb.acid_soap > 0.0 && pb.use_charges && pb.cls[ki] == 1u && pb.cls[kj] == 1u &&
        ((qi == 0.0f) != (b.charge[u] == 0.0f)))
      sc += pb.acid_soap;
    if (sc > 0.0) nb += sc * pot_attr(pb, r);
  }
  if (lane == 0) {
    const uint4 own = b.bs[t];
    const uint32_t sl[3] = {own.x, own.y, own.z};
    for (int k = 0; k < 3; ++k) {
      if (sl[k] == SOUP_NONE) continue;
      const float4 xk = b.pos[b.inv[sl[k]]];
      if (sl[k] > idi) fene += pot_fene(pb, sqrt(pot_dist2(c, xi, xk)));
      for (int m = k + 1; m < 3; ++m) {
        if (sl[m] == SOUP_NONE) continue;
        const double d = sqrt(pot_dist2(c, xk, b.pos[b.inv[sl[m]]])) - pb.bend_r0;
        bend += 0.5 * pb.bend_k * d * d;
      }
    }
    const uint32_t owner = b.clink[t];
    if (pb.use_center_link && owner != SOUP_NONE && owner > idi) teth += pot_fene(pb, sqrt(pot_dist2(c, xi, b.pos[b.inv[owner]])));
  }
  const uint32_t hh = es ? uni(b.head_of_slot[t]) : SOUP_NONE;
  if (hh != SOUP_NONE && qi != 0.0f) {
    const uint32_t ecnt = uni(b.es_count[hh]);
    const uint32_t* erow = b.es + (size_t)hh * c.es_cap;
    const double ecut2 = pb.es_cutoff * pb.es_cutoff;
    for (uint32_t s = lane; s < ecnt; s += WAVE) {
      const uint32_t g = erow[s];
      if (b.head_id[g] <= idi) continue;
      const float4 xq = b.head_q[g];
      if (xq.w == 0.0f) continue;
      const double r2 = pot_dist2(c, xi, xq);
      if (r2 >= ecut2) continue;
      const double r = sqrt(r2);
      if (r <= 0.0) continue;
      const double u1 = (pb.es_A * exp(-pb.es_kappa * r)) / r;
      nb += (double)qi * (double)xq.w * (u1 - pb.es_shiftU + pb.es_shiftF * (r - pb.es_cutoff));
    }
  }
  nb = wave_sum_d_dpp(nb);
  if (lane == 0) part[t] = make_double4(nb, fene, bend, teth);
}

// fixed-order sum of the per-particle terms: thread k sums slots k, k + 1024, ... in order, then a
// fixed tree over the threads
__global__ __launch_bounds__(1024) void k_potential_sum(const double4* part, uint32_t n, double* out) {
  __shared__ double4 s[1024];
  double4 a = make_double4(0.0, 0.0, 0.0, 0.0);
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double4 v = part[i];
    a.x += v.x; a.y += v.y; a.z += v.z; a.w += v.w;
  }
  s[threadIdx.x] = a;
  __syncthreads();
  for (uint32_t o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const double4 v = s[threadIdx.x + o];
      s[threadIdx.x].x += v.x; s[threadIdx.x].y += v.y; s[threadIdx.x].z += v.z; s[threadIdx.x].w += v.w;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double4 v = s[0];
    out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
    out[4] = v.x + v.y + v.z + v.w;
  }
}

// ================================================================================================
// Adjoint of soupPotential (soup/src/soup-potential.ts:145-301) at the handle's configuration, FP64:
// for each particle i the gradient g_i = dU/dx_i (= -F_i of the step's force law), and for a caller's
// direction w the Hessian-vector product (H w)_i, which is -(w . dF/dx)_i, plus the partial
// derivatives of U and of w . F with respect to the parameter table: the attraction depths of the
// 10 unordered class pairs, the acid-soap depth, epsilon, the head charge and the DH screening
// kappa (at fixed coefficient and cutoff). Pair terms phi(r) contribute phi' e to g_i, with e the
// unit vector from j to i, and phi'' e (e . dw) + (phi' / r)(dw - e (e . dw)) to (H w)_i, with
// dw = w_i - w_j; parameters take each unordered pair once (the lower particle index), as the
// potential does. The pairs are those of soupPotential: list pairs within basis->cutoff, bonds,
// 1-3 bends through each centre, centre-link tethers, and DH pairs of charged heads.
//
// A diagonal strain e = (ex, ey, ez) adds diag(e) r_ij to every pair's dw, r_ij = x_i - x_j under
// the minimum image: the direction of a homogeneous deformation. Every term is a pair term, so the
// strain virial W_e = sum over pairs of (diag(e) r_ij) . f_ij (f_ij = -phi' e_ij, the force on i)
// then has dW_e/dx_i = -diag(e) g_i - (H_e)_i, where H_e is the Hessian product the strain adds to
// hw, and dW_e/dtheta is what it adds to d(w . F)/dtheta; W_e itself is reduced as one more value.
// ================================================================================================
#define ADJ_P 14u  // 10 class pairs, acid-soap, epsilon, head charge, kappa
#define ADJ_NPART (2u * ADJ_P + 1u)  // per-particle partials: dU, d(w . F), W_e
struct AdjAcc {
  double g[3], hw[3], dU[ADJ_P], dF[ADJ_P], we;
};
struct AdjStrain {
  double e[3];
};
__device__ __forceinline__ uint32_t adj_attr_index(uint32_t a, uint32_t b) {
  const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  return lo * 4u - (lo * (lo - 1u)) / 2u + (hi - lo);  // (0,0) 0 ... (0,3) 3, (1,1) 4 ... (3,3) 9
}
// accumulates the position parts of one pair term with derivatives d1 = phi'(r), d2 = phi''(r)
__device__ __forceinline__ void adj_pos(AdjAcc& A, double d1, double d2, const double e[3], double r, const double dw[3],
                                        double ew) {
  const double t = d1 / r;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    A.g[a] += d1 * e[a];
    A.hw[a] += d2 * ew * e[a] + t * (dw[a] - e[a] * ew);
  }
}
__device__ __forceinline__ double adj_fene_d1(const SoupPotentialBasis& pb, double r) {
  const double x2 = (r / pb.fene_r_inf) * (r / pb.fene_r_inf);
  return pb.fene_k * r / (1.0 - x2);
}
__device__ __forceinline__ double adj_fene_d2(const SoupPotentialBasis& pb, double r) {
  const double x2 = (r / pb.fene_r_inf) * (r / pb.fene_r_inf);
  return pb.fene_k * (1.0 + x2) / ((1.0 - x2) * (1.0 - x2));
}
// displacement x_i - x_j under the FP64 minimum image, its length and unit vector
__device__ __forceinline__ double adj_geom(const SoupConfig& c, float4 xi, float4 xj, double e[3]) {
  e[0] = pot_mi((double)xi.x - (double)xj.x, c.box[0]);
  e[1] = pot_mi((double)xi.y - (double)xj.y, c.box[1]);
  e[2] = pot_mi((double)xi.z - (double)xj.z, c.box[2]);
  return e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
}
// dw = w_i - w_j + diag(strain) r_ij (e: the unit vector from j to i, r its length); es = the strain
// part's projection on e
__device__ __forceinline__ void adj_dw(const double* w, const AdjStrain& st, uint32_t idi, uint32_t idj, const double e[3],
                                       double r, double dw[3], double& es) {
  if (w) {
    dw[0] = w[3u * idi] - w[3u * idj];
    dw[1] = w[3u * idi + 1u] - w[3u * idj + 1u];
    dw[2] = w[3u * idi + 2u] - w[3u * idj + 2u];
  } else {
    dw[0] = dw[1] = dw[2] = 0.0;
  }
  es = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double d = st.e[a] * e[a] * r;
    dw[a] += d;
    es += e[a] * d;
  }
}
// a bonded pair term (FENE: fene = true; 1-3 bend otherwise) on particle i
// (W_e takes each bonded pair term once, from its lower particle index)
__device__ __forceinline__ void adj_bonded(const SoupConfig& c, const SoupPotentialBasis& pb, float4 xi, float4 xj,
                                           uint32_t idi, uint32_t idj, const double* w, const AdjStrain& st, bool fene,
                                           AdjAcc& A) {
  double e[3];
  const double r = sqrt(adj_geom(c, xi, xj, e));
  if (!(r > 0.0)) return;
  e[0] /= r; e[1] /= r; e[2] /= r;
  double dw[3], es;
  adj_dw(w, st, idi, idj, e, r, dw, es);
  const double ew = e[0] * dw[0] + e[1] * dw[1] + e[2] * dw[2];
  const double d1 = fene ? adj_fene_d1(pb, r) : pb.bend_k * (r - pb.bend_r0);
  adj_pos(A, d1, fene ? adj_fene_d2(pb, r) : pb.bend_k, e, r, dw, ew);
  if (idj > idi) A.we -= es * d1;
}

__global__ __launch_bounds__(256) void k_energy_adjoint(SoupConfig c, DevBuf b, SoupPotentialBasis pb, int es, const double* w,
                                                        AdjStrain st, double* g_out, double* hw_out, double* part) {
  const uint32_t lane = threadIdx.x & (WAVE - 1), wv = uni(threadIdx.x / WAVE);
  const uint32_t t = uni(blockIdx.x * (blockDim.x / WAVE) + wv);
  if (t >= c.n) return;
  const float4 xv = b.pos[t];
  const float4 xi = make_float4(uni_f(xv.x), uni_f(xv.y), uni_f(xv.z), uni_f(xv.w));
  const uint32_t idi = uni(b.id[t]);
  const uint32_t ki = (uint32_t)xi.w & (SOUP_MAX_KINDS - 1u);
  const float qi = uni_f(b.charge[t]);
  const double cut2 = pb.cutoff * pb.cutoff;
  const double eps = pb.epsilon, z1 = 3.141592653589793 / (2.0 * pb.wc);
  AdjAcc A;
#pragma unroll
  for (int a = 0; a < 3; ++a) A.g[a] = A.hw[a] = 0.0;
#pragma unroll
  for (uint32_t p = 0; p < ADJ_P; ++p) A.dU[p] = A.dF[p] = 0.0;
  A.we = 0.0;
  const uint32_t cnt = uni(b.vcount[t]), cn = uni(b.vnear[t]);
  const uint32_t* row = b.vl + (size_t)t * vl_stride(c);
  for (uint32_t s = lane; s < cnt; s += WAVE) {
    const uint32_t u = row[row_index(s, cn, c.list_cap)];
    const float4 xj = b.pos[u];
    double e[3];
    const double r2 = adj_geom(c, xi, xj, e);
    if (r2 > cut2 || r2 <= 0.0) continue;
    const double r = sqrt(r2);
    e[0] /= r; e[1] /= r; e[2] /= r;
    const uint32_t idj = b.id[u];
    double dw[3], ews;
    adj_dw(w, st, idi, idj, e, r, dw, ews);
    const double ew = e[0] * dw[0] + e[1] * dw[1] + e[2] * dw[2];
    const bool once = idj > idi;  // parameters and W_e: each unordered pair once
    const uint32_t kj = (uint32_t)xj.w & (SOUP_MAX_KINDS - 1u);
    const double bb = (pb.b_radius[ki] + pb.b_radius[kj]) / 2.0;
    if (r < pb.wca_factor * bb) {
      const double sr = bb / r, s2 = sr * sr, u6 = s2 * s2 * s2;
      const double phi = eps * (2.0 * u6 - 1.0) * (2.0 * u6 - 1.0);
      const double d1 = -24.0 * eps * u6 * (2.0 * u6 - 1.0) / r;
      const double d2 = 24.0 * eps * u6 * (26.0 * u6 - 7.0) / r2;
      adj_pos(A, d1, d2, e, r, dw, ew);
      if (once) {
        A.dU[11] += phi / eps;
        A.dF[11] -= ew * d1 / eps;
        A.we -= ews * d1;
      }
    }
    const uint32_t ci = pb.cls[ki] & 3u, cj = pb.cls[kj] & 3u;
    const bool acid = pb.acid_soap > 0.0 && pb.use_charges && pb.cls[ki] == 1u && pb.cls[kj] == 1u &&
                      ((qi == 0.0f) != (b.charge[u] == 0.0f));
    const double sc = pb.attr[ci][cj] + (acid ? pb.acid_soap : 0.0);
    if (sc > 0.0 && r <= pb.rc_attr + pb.wc) {
      double a0, a1, a2;  // A(r) = pot_attr and its derivatives (epsilon included)
      if (r < pb.rc_attr) {
        a0 = -eps; a1 = 0.0; a2 = 0.0;
      } else {
        const double z = z1 * (r - pb.rc_attr), cz = cos(z);
        a0 = -eps * cz * cz;
        a1 = eps * z1 * sin(2.0 * z);
        a2 = 2.0 * eps * z1 * z1 * cos(2.0 * z);
      }
      adj_pos(A, sc * a1, sc * a2, e, r, dw, ew);
      if (once) {
        A.we -= ews * sc * a1;
        const uint32_t pi = adj_attr_index(ci, cj);
        A.dU[pi] += a0;
        A.dF[pi] -= ew * a1;
        if (acid) {
          A.dU[10] += a0;
          A.dF[10] -= ew * a1;
        }
        A.dU[11] += sc * a0 / eps;
        A.dF[11] -= ew * sc * a1 / eps;
      }
    }
  }
  if (lane == 0) {
    const uint4 own = b.bs[t];
    const uint32_t sl[3] = {own.x, own.y, own.z};
    for (int k = 0; k < 3; ++k) {
      const uint32_t p = sl[k];
      if (p == SOUP_NONE) continue;
      const uint32_t up = b.inv[p];
      adj_bonded(c, pb, xi, b.pos[up], idi, p, w, st, true, A);
      // 1-3 bends through partner p: this particle and p's other partners
      const uint4 ms = b.bs[up];
      const uint32_t ml[3] = {ms.x, ms.y, ms.z};
      for (int q = 0; q < 3; ++q) {
        const uint32_t o = ml[q];
        if (o == SOUP_NONE || o == idi) continue;
        adj_bonded(c, pb, xi, b.pos[b.inv[o]], idi, o, w, st, false, A);
      }
    }
    const uint32_t owner = b.clink[t];
    if (pb.use_center_link && owner != SOUP_NONE) adj_bonded(c, pb, xi, b.pos[b.inv[owner]], idi, owner, w, st, true, A);
  }
  const uint32_t hh = es ? uni(b.head_of_slot[t]) : SOUP_NONE;
  if (hh != SOUP_NONE && qi != 0.0f) {
    const uint32_t ecnt = uni(b.es_count[hh]);
    const uint32_t* erow = b.es + (size_t)hh * c.es_cap;
    const double ecut2 = pb.es_cutoff * pb.es_cutoff, rc = pb.es_cutoff, kap = pb.es_kappa;
    const double erc = pb.es_A * exp(-kap * rc);
    for (uint32_t s = lane; s < ecnt; s += WAVE) {
      const uint32_t gh = erow[s];
      const float4 xq = b.head_q[gh];
      if (xq.w == 0.0f) continue;
      double e[3];
      const double r2 = adj_geom(c, xi, xq, e);
      if (r2 >= ecut2) continue;
      const double r = sqrt(r2);
      if (!(r > 0.0)) continue;
      e[0] /= r; e[1] /= r; e[2] /= r;
      const uint32_t idj = b.head_id[gh];
      double dw[3], ews;
      adj_dw(w, st, idi, idj, e, r, dw, ews);
      const double ew = e[0] * dw[0] + e[1] * dw[1] + e[2] * dw[2];
      const double qq = (double)qi * (double)xq.w;
      const double u1 = (pb.es_A * exp(-kap * r)) / r, k1 = kap + 1.0 / r;
      const double e0 = u1 - pb.es_shiftU + pb.es_shiftF * (r - rc);
      const double e1 = -u1 * k1 + pb.es_shiftF;
      const double e2 = u1 * k1 * k1 + u1 / r2;
      adj_pos(A, qq * e1, qq * e2, e, r, dw, ew);
      if (idj > idi) {
        A.we -= ews * qq * e1;
        A.dU[12] += 2.0 * qq * e0 / c.ph_q;
        A.dF[12] -= ew * 2.0 * qq * e1 / c.ph_q;
        A.dU[13] += qq * (-r * u1 + erc - kap * erc * (r - rc));
        A.dF[13] -= ew * qq * (u1 * r * kap - kap * erc);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    A.g[a] = wave_sum_d_dpp(A.g[a]);
    A.hw[a] = wave_sum_d_dpp(A.hw[a]);
  }
#pragma unroll
  for (uint32_t p = 0; p < ADJ_P; ++p) {
    A.dU[p] = wave_sum_d_dpp(A.dU[p]);
    A.dF[p] = wave_sum_d_dpp(A.dF[p]);
  }
  A.we = wave_sum_d_dpp(A.we);
  if (lane == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      g_out[3u * idi + a] = A.g[a];
      if (hw_out) hw_out[3u * idi + a] = A.hw[a];
    }
#pragma unroll
    for (uint32_t p = 0; p < ADJ_P; ++p) {
      part[(size_t)t * ADJ_NPART + p] = A.dU[p];
      part[(size_t)t * ADJ_NPART + ADJ_P + p] = A.dF[p];
    }
    part[(size_t)t * ADJ_NPART + 2u * ADJ_P] = A.we;
  }
}

// fixed-order sum of the per-particle partials (k_potential_sum's scheme, ADJ_NPART values)
__global__ __launch_bounds__(1024) void k_adjoint_sum(const double* part, uint32_t n, double* out) {
  __shared__ double s[1024];
  for (uint32_t p = 0; p < ADJ_NPART; ++p) {
    double a = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) a += part[(size_t)i * ADJ_NPART + p];
    s[threadIdx.x] = a;
    __syncthreads();
    for (uint32_t o = blockDim.x / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[p] = s[0];
    __syncthreads();
  }
}

// ================================================================================================
// Trajectory-segment adjoint helpers: states move between a system's slot order and particle order
// (checkpoints and per-step records), and the reverse sweep's adjoint vectors (FP64, particle
// order, 3 values per particle) are updated on the device.
// ================================================================================================
struct StateRefs {  // particle-order arrays of one state; a null member is skipped
  float4 *pos, *vel, *force;
  float* charge;
  uint4* bs;
  uint32_t *clink, *cheld, *frozen, *brng, *trng;
};
// slot order -> particle order
__global__ __launch_bounds__(256) void k_state_gather(uint32_t n, DevBuf b, StateRefs s) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t t = b.inv[p];
  if (s.pos) s.pos[p] = b.pos[t];
  if (s.vel) s.vel[p] = b.vel[t];
  if (s.force) s.force[p] = b.force[t];
  if (s.charge) s.charge[p] = b.charge[t];
  if (s.bs) s.bs[p] = b.bs[t];
  if (s.clink) s.clink[p] = b.clink[t];
  if (s.cheld) s.cheld[p] = b.cheld[t];
  if (s.frozen) s.frozen[p] = b.frozen[t];
  if (s.brng) s.brng[p] = b.brng[t];
  if (s.trng) s.trng[p] = b.trng[t];
}
// particle order -> the system's current slot order
__global__ __launch_bounds__(256) void k_state_scatter(uint32_t n, DevBuf b, StateRefs s) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t p = b.id[t];
  if (s.pos) b.pos[t] = s.pos[p];
  if (s.vel) b.vel[t] = s.vel[p];
  if (s.force) b.force[t] = s.force[p];
  if (s.charge) b.charge[t] = s.charge[p];
  if (s.bs) b.bs[t] = s.bs[p];
  if (s.clink) b.clink[t] = s.clink[p];
  if (s.cheld) b.cheld[t] = s.cheld[p];
  if (s.frozen) b.frozen[t] = s.frozen[p];
  if (s.brng) b.brng[t] = s.brng[p];
  if (s.trng) b.trng[t] = s.trng[p];
}
// w = cw * lv
__global__ __launch_bounds__(256) void k_seg_w(uint32_t m, const double* lv, double cw, double* w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) w[i] = cw * lv[i];
}
// lx += -hw - diag(e) g: the vector-Jacobian product of the force (-H w) and, with the strain e of
// the objective (zero without), the strain virial's position gradient -diag(e) g - H_e
__global__ __launch_bounds__(256) void k_seg_lx(uint32_t n, double* lx, const double* hw, const double* g, AdjStrain st) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) lx[3u * i + a] -= hw[3u * i + a] + st.e[a] * g[3u * i + a];
}
// the velocity adjoint of one step backwards: lv = a lv + dt lx
__global__ __launch_bounds__(256) void k_seg_lv(uint32_t m, double* lv, const double* lx, double a, double dt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) lv[i] = a * lv[i] + dt * lx[i];
}
// acc[0..ADJ_P) += the d(w . F)/dtheta partials (or the dU/dtheta ones: du = 1), acc[ADJ_P] += W_e
__global__ void k_seg_acc(const double* out, double* acc, int du) {
  const uint32_t p = threadIdx.x;
  if (p < ADJ_P) acc[p] += out[(du ? 0u : ADJ_P) + p];
  if (p == ADJ_P) acc[ADJ_P] += out[2u * ADJ_P];
}

// ================================================================================================
// soupPotential of an arbitrary configuration (soup/src/soup-potential.ts:145-301), for the area
// move's trial boxes (soup/src/soup-area-move.ts:211-226): positions in particle order in their own
// periodic box, with the system's bond slots, centre links and charges. The pairs come from the
// configuration's own cell grid, as the reference builds it: cells of at least `cutoff` per axis
// (floor(L / cutoff), at least 1), each cell walking the deduplicated set of its 27 neighbours, so a
// pair is seen once from each side and taken from the lower particle index. The DH pairs of
// esTotalEnergy (electrostatics.ts:575-600) are taken over all charged heads.
// ================================================================================================
struct PaGrid {
  double L[3], w[3];
  uint32_t d[3];
};
__device__ __forceinline__ double pa_mi(double d, double L) { return d - js_round(d / L) * L; }
__device__ __forceinline__ double pa_dist2(const PaGrid& g, float4 a, float4 b) {
  const double dx = pa_mi((double)a.x - (double)b.x, g.L[0]);
  const double dy = pa_mi((double)a.y - (double)b.y, g.L[1]);
  const double dz = pa_mi((double)a.z - (double)b.z, g.L[2]);
  return dx * dx + dy * dy + dz * dz;
}
__device__ __forceinline__ uint32_t pa_axis(double x, double L, double w, uint32_t d) {
  const double v = x - floor(x / L) * L;
  const double c = floor(v / w);
  return c < (double)(d - 1u) ? (uint32_t)c : d - 1u;
}
// nonbonded pair energy of soup-potential.ts:230-252: WCA core plus the class-table attraction with
// the acid-soap increment
__device__ __forceinline__ double pot_pair(const SoupPotentialBasis& pb, uint32_t ki, uint32_t kj, float qi, float qj, double r) {
  double e = 0.0;
  const double bb = (pb.b_radius[ki] + pb.b_radius[kj]) / 2.0;
  if (r < pb.wca_factor * bb) {
    const double sr = bb / r, s2 = sr * sr, s6 = s2 * s2 * s2, q = 2.0 * s6 - 1.0;
    e += pb.epsilon * q * q;
  }
  double sc = pb.attr[pb.cls[ki] & 3u][pb.cls[kj] & 3u];
  if (pb.acid_soap > 0.0 && pb.use_charges && pb.cls[ki] == 1u && pb.cls[kj] == 1u && ((qi == 0.0f) != (qj == 0.0f)))
    sc += pb.acid_soap;
  if (sc > 0.0) e += sc * pot_attr(pb, r);
  return e;
}

__global__ __launch_bounds__(256) void k_pa_cells(const float4* xs, uint32_t n, PaGrid g, uint32_t* cell_of, uint32_t* counts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 x = xs[i];
  const uint32_t c = pa_axis(x.x, g.L[0], g.w[0], g.d[0]) +
                     g.d[0] * (pa_axis(x.y, g.L[1], g.w[1], g.d[1]) + g.d[1] * pa_axis(x.z, g.L[2], g.w[2], g.d[2]));
  cell_of[i] = c;
  atomicAdd(&counts[c], 1u);
}
// ================================================================================================
// forceCandidateStatsDEBUG (soup/src/soup-readback.ts:246-263): the step grid of the current
// positions in scratch arrays (the state's own cell order is left alone), then the cell walk of
// soup_force_stats_main (soup/wgsl/step.wgsl:307-343) counting candidates and in-range pairs.
// ================================================================================================
__global__ __launch_bounds__(256) void k_stats_cells(SoupConfig c, DevBuf b, uint32_t* cell, uint32_t* counts) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.n) return;
  const uint32_t k = cell_of(c, b.pos[t]);
  cell[t] = k;
  atomicAdd(&counts[k], 1u);
}

// one thread per particle; a pair is in range inside its WCA cut, or inside the attraction shell
// [rcAttr, rcAttr + wc] of a class pair whose table depth is positive (shouldAttract)
__global__ __launch_bounds__(256) void k_force_stats(SoupConfig c, DevBuf b, const uint32_t* cell, const uint32_t* start,
                                                     const uint32_t* order, unsigned long long* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t cand = 0, pairs = 0;
  if (t < c.n) {
    const float4 xi = b.pos[t];
    const uint32_t ci = cell[t];
    const int d0 = (int)c.dims[0], d1 = (int)c.dims[1], d2 = (int)c.dims[2];
    const int cx0 = (int)(ci % c.dims[0]), cy0 = (int)((ci / c.dims[0]) % c.dims[1]), cz0 = (int)(ci / (c.dims[0] * c.dims[1]));
    const int R = (int)c.walk_radius;
    const float rc = wca_cut(c, c.b_tt);
    const uint32_t clsi = c.cls[(uint32_t)xi.w];
    for (int dz = -R; dz <= R; ++dz) {
      const int cz = ((cz0 + dz) % d2 + d2) % d2;
      for (int dy = -R; dy <= R; ++dy) {
        const int cy = ((cy0 + dy) % d1 + d1) % d1;
        for (int dx = -R; dx <= R; ++dx) {
          const int cx = ((cx0 + dx) % d0 + d0) % d0;
          const uint32_t nc = (uint32_t)cx + c.dims[0] * ((uint32_t)cy + c.dims[1] * (uint32_t)cz);
          for (uint32_t k = start[nc]; k < start[nc + 1]; ++k) {
            const uint32_t j = order[k];
            if (j == t) continue;
            ++cand;
            const float4 xj = b.pos[j];
            const float r = dist_f(c, xi, xj);
            const bool wca = r < wca_cut(c, pair_b(c, xi.w, xj.w));
            const bool attr = c.attr[clsi][c.cls[(uint32_t)xj.w]] > 0.0f && r >= rc && r <= rc + c.wc;
            pairs += (wca || attr) ? 1u : 0u;
          }
        }
      }
    }
  }
  cand = wave_sum_u(cand);
  pairs = wave_sum_u(pairs);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    if (cand) atomicAdd(&out[0], (unsigned long long)cand);
    if (pairs) atomicAdd(&out[1], (unsigned long long)pairs);
  }
}

__global__ __launch_bounds__(256) void k_pa_fill(const uint32_t* cell_of, uint32_t n, uint32_t* cursor, uint32_t* order) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) order[atomicAdd(&cursor[cell_of[i]], 1u)] = i;
}
// each cell's entries in ascending particle index (insertion sort; cells hold tens of particles)
__global__ __launch_bounds__(256) void k_pa_sort(const uint32_t* start, uint32_t ncell, uint32_t* order) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const uint32_t a = start[c], e = start[c + 1];
  for (uint32_t k = a + 1; k < e; ++k) {
    const uint32_t v = order[k];
    uint32_t m = k;
    while (m > a && order[m - 1] > v) {
      order[m] = order[m - 1];
      --m;
    }
    order[m] = v;
  }
}

// one wavefront per particle i; part[i] = (nonbonded, fene, bend, tether)
__global__ __launch_bounds__(256) void k_potential_at(SoupConfig c, DevBuf b, SoupPotentialBasis pb, const float4* xs, PaGrid g,
                                                      const uint32_t* cell_of, const uint32_t* start, const uint32_t* order,
                                                      double4* part, uint32_t nh) {
  const uint32_t lane = threadIdx.x & (WAVE - 1), w = uni(threadIdx.x / WAVE);
  const uint32_t i = uni(blockIdx.x * (blockDim.x / WAVE) + w);
  if (i >= c.n) return;
  const float4 xv = xs[i];
  const float4 xi = make_float4(uni_f(xv.x), uni_f(xv.y), uni_f(xv.z), uni_f(xv.w));
  const uint32_t ti = uni(b.inv[i]);
  const uint32_t ki = (uint32_t)xi.w & (SOUP_MAX_KINDS - 1u);
  const float qi = uni_f(b.charge[ti]);
  const double cut2 = pb.cutoff * pb.cutoff;
  // the deduplicated neighbour cells, in the reference's dz, dy, dx order of first occurrence
  const uint32_t ci = uni(cell_of[i]);
  const uint32_t cx = ci % g.d[0], cy = (ci / g.d[0]) % g.d[1], cz = ci / (g.d[0] * g.d[1]);
  uint32_t nbr = SOUP_NONE;
  if (lane < 27u) {
    const int dz = (int)(lane / 9u) - 1, dy = (int)((lane / 3u) % 3u) - 1, dx = (int)(lane % 3u) - 1;
    const uint32_t nz = (uint32_t)(((int)cz + dz + (int)g.d[2]) % (int)g.d[2]);
    const uint32_t ny = (uint32_t)(((int)cy + dy + (int)g.d[1]) % (int)g.d[1]);
    const uint32_t nx = (uint32_t)(((int)cx + dx + (int)g.d[0]) % (int)g.d[0]);
    nbr = nx + g.d[0] * (ny + g.d[1] * nz);
  }
  bool first = lane < 27u;
  for (uint32_t k = 0; k < 27u; ++k) {
    const uint32_t v = (uint32_t)__shfl((int)nbr, (int)k, WAVE);
    if (k < lane && v == nbr) first = false;
  }
  uint64_t cells = __ballot(first);
  double nb = 0.0, fene = 0.0, bend = 0.0, teth = 0.0;
  while (cells) {
    const int k = __ffsll((unsigned long long)cells) - 1;
    cells &= cells - 1ull;
    const uint32_t cc = (uint32_t)__shfl((int)nbr, k, WAVE);
    const uint32_t e0 = start[cc], e1 = start[cc + 1];
    for (uint32_t s = e0 + lane; s < e1; s += WAVE) {
      const uint32_t j = order[s];
      if (j <= i) continue;
      const float4 xj = xs[j];
      const double r2 = pa_dist2(g, xi, xj);
      if (r2 > cut2 || r2 <= 0.0) continue;
      nb += pot_pair(pb, ki, (uint32_t)xj.w & (SOUP_MAX_KINDS - 1u), qi, b.charge[b.inv[j]], sqrt(r2));
    }
  }
  if (lane == 0) {
    const uint4 own = b.bs[ti];
    const uint32_t sl[3] = {own.x, own.y, own.z};
    for (int k = 0; k < 3; ++k) {
      if (sl[k] == SOUP_NONE) continue;
      const float4 xk = xs[sl[k]];
      if (sl[k] > i) fene += pot_fene(pb, sqrt(pa_dist2(g, xi, xk)));
      for (int m = k + 1; m < 3; ++m) {
        if (sl[m] == SOUP_NONE) continue;
        const double d = sqrt(pa_dist2(g, xk, xs[sl[m]])) - pb.bend_r0;
        bend += 0.5 * pb.bend_k * d * d;
      }
    }
    const uint32_t owner = b.clink[ti];
    if (pb.use_center_link && owner != SOUP_NONE && owner > i) teth += pot_fene(pb, sqrt(pa_dist2(g, xi, xs[owner])));
  }
  if (pb.use_charges && qi != 0.0f && ki == pb.charged_kind) {
    const double ecut2 = pb.es_cutoff * pb.es_cutoff;
    for (uint32_t h = lane; h < nh; h += WAVE) {
      const uint32_t j = b.head_id[h];
      if (j <= i) continue;
      const float qj = b.charge[b.inv[j]];
      if (qj == 0.0f) continue;
      const double r2 = pa_dist2(g, xi, xs[j]);
      if (r2 >= ecut2) continue;
      const double r = sqrt(r2);
      if (r <= 0.0) continue;
      const double u1 = (pb.es_A * exp(-pb.es_kappa * r)) / r;
      nb += (double)qi * (double)qj * (u1 - pb.es_shiftU + pb.es_shiftF * (r - pb.es_cutoff));
    }
  }
  nb = wave_sum_d_dpp(nb);
  if (lane == 0) part[i] = make_double4(nb, fene, bend, teth);
}

template <typename T>
__global__ void k_to_index(const T* src, const uint32_t* id, T* dst, uint32_t n) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[id[t]] = src[t];
}
template <typename T>
__global__ void k_from_index(const T* src, const uint32_t* id, T* dst, uint32_t n) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[id[t]];
}
// bond slots: uint4 internal -> 3 u32 per particle
__global__ void k_bs_to_index(const uint4* bs, const uint32_t* id, uint32_t* dst, uint32_t n) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  uint4 s = bs[t];
  uint32_t i = id[t];
  dst[3u * i] = s.x;
  dst[3u * i + 1] = s.y;
  dst[3u * i + 2] = s.z;
}
__global__ void k_bs_from_index(const uint32_t* src, const uint32_t* id, uint4* bs, uint32_t n) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  uint32_t i = id[t];
  bs[t] = make_uint4(src[3u * i], src[3u * i + 1], src[3u * i + 2], 0u);
}
// Verlet list in the reference layout: out[i * cap + s] = particle index, in the walk order. One
// wavefront per row: the walk keys of the row's entries (near part and far part) go to LDS, and each
// entry's rank among them is its position. Equal keys (a cell the stencil reaches at two offsets on
// a grid narrower than the stencil) keep their order in the stored row.
__global__ __launch_bounds__(WAVE) void k_vl_to_index(SoupConfig c, DevBuf b, uint32_t* out) {
  extern __shared__ uint64_t vl_keys[];
  const uint32_t t = blockIdx.x, lane = threadIdx.x;
  if (t >= c.n) return;
  const uint32_t cap = c.list_cap, i = b.id[t], ct = b.cell_of[t];
  const uint32_t* row = b.vl + (size_t)t * vl_stride(c);
  const uint32_t cnt = b.vcount[t], cn = b.vnear[t];
  for (uint32_t s = lane; s < cnt; s += WAVE) {
    const uint32_t u = row[row_index(s, cn, cap)];
    vl_keys[s] = walk_key(c, b.cell_of, ct, u, b.id[u]);
  }
  __syncthreads();
  for (uint32_t s = lane; s < cnt; s += WAVE) {
    const uint64_t me = vl_keys[s];
    uint32_t rk = 0;
    for (uint32_t q = 0; q < cnt; ++q) rk += (vl_keys[q] < me || (vl_keys[q] == me && q < s)) ? 1u : 0u;
    out[(size_t)i * cap + rk] = (uint32_t)me;
  }
  for (uint32_t s = cnt + lane; s < cap; s += WAVE) out[(size_t)i * cap + s] = SOUP_NONE;
}
// the reference's cell-sorted index list (CELL_IDX): the particle indices of each cell in ascending
// order, from the slots, which order a cell by x
__global__ __launch_bounds__(256) void k_cell_idx_to_index(SoupConfig c, DevBuf b, uint32_t* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.n) return;
  const uint32_t cell = b.cell_of[t], me = b.id[t];
  const uint32_t s0 = b.cell_start[cell], s1 = b.cell_start[cell + 1u];
  uint32_t r = 0;
  for (uint32_t k = s0; k < s1; ++k) r += b.id[k] < me ? 1u : 0u;
  out[s0 + r] = me;
}
// DH list in the reference layout: entries are head indices, exported as particle indices
__global__ void k_es_to_index(const uint32_t* es, const uint32_t* es_count, const uint32_t* head_id, uint32_t* out,
                              uint32_t nh, uint32_t cap) {
  uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  uint32_t cnt = es_count[h];
  for (uint32_t s = 0; s < cap; ++s) out[(size_t)h * cap + s] = s < cnt ? head_id[es[(size_t)h * cap + s]] : SOUP_NONE;
}
__global__ void k_vcount_to_index(const uint32_t* vcount, const uint32_t* id, uint32_t* out, uint32_t n) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[id[t]] = vcount[t];
}
// the list force's row reads summed over its per-wavefront counters: out[0] rows read as near parts,
// out[1] rows read whole; out[2] launches in which every row qualified for its near part, out[3] the
// others (flags[8], flags[9])
__global__ __launch_bounds__(1024) void k_sum_rowreads(const uint32_t* rr, uint32_t nslots, const uint32_t* flags,
                                                       unsigned long long* out) {
  __shared__ unsigned long long s_part[2][1024 / WAVE];
  unsigned long long a = 0, w = 0;
  for (uint32_t k = threadIdx.x; k < nslots; k += blockDim.x) {
    a += rr[2u * k];
    w += rr[2u * k + 1u];
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += (unsigned long long)__shfl_xor((long long)a, o, WAVE);
    w += (unsigned long long)__shfl_xor((long long)w, o, WAVE);
  }
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    s_part[0][threadIdx.x / WAVE] = a;
    s_part[1][threadIdx.x / WAVE] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sa = 0, sw = 0;
    for (uint32_t k = 0; k < blockDim.x / WAVE; ++k) {
      sa += s_part[0][k];
      sw += s_part[1][k];
    }
    out[0] = sa;
    out[1] = sw;
    out[2] = flags[8];
    out[3] = flags[9];
  }
}
// list occupancy: [0] sum verlet, [1] max verlet, [2] sum es, [3] max es (u64 sums in pairs)
__global__ __launch_bounds__(256) void k_list_stats(const uint32_t* cnt, uint32_t n, unsigned long long* sum, uint32_t* mx) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t v = t < n ? cnt[t] : 0u;
  uint32_t m = wave_max_u(v);
  uint32_t s = wave_sum_u(v);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    atomicAdd(sum, (unsigned long long)s);
    atomicMax(mx, m);
  }
}
