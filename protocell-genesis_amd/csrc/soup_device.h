// soup_device.h -- device-side state layout and the shared f32 pair/bond/RNG helpers of the soup
// Langevin step on gfx950.
//
// Layout in HBM: every per-particle array is indexed by SLOT, the particle's position in
// cell-sorted order at the last Verlet rebuild (cell-major, then ascending x inside a cell, ties by
// particle index). The rebuild permutes all per-particle state into the new slot order, so every per-step
// kernel streams its arrays with unit stride and neighbour gathers stay spatially local. id[slot] is
// the particle index (the reference's array index), inv[id] its current slot. Bond slots and
// centerLink keep holding particle indices, exactly as in the reference.
//
// The Verlet list is row-major, vl[slot * stride + s], with the stride (vl_stride) the capacity rounded
// up to 32 entries so every row starts on its own 128-byte line. Entries are SLOTS (positions in the
// same cell-sorted arrays). A row holds its near part (closer than the near/far split at the rebuild)
// from the front in the cell order of soup/wgsl/verlet.wgsl:66-88 (z-major cell walk; inside a cell
// ascending x, where the reference takes ascending particle index) and its far part from the back of
// the capacity in reverse order. The list readback and the bond proposals order entries by the
// reference's walk key (cell, then particle index).
//
// The DH list is row-major, es[h * es_cap + s]: row h is the h-th charged-kind particle in ascending
// particle-index order (electrostatics-long.wgsl:15-29); its entries are the neighbour HEAD indices,
// ascending (the reference's order). DH partners are always heads, so the force gathers them from the
// compact head_q array (about 300 KB at campaign size, resident in L2) instead of the particle arrays.
//
// Launch geometry: workgroups are dispatched round-robin over the 8 XCDs, so the gather-heavy kernels
// remap blockIdx so that each XCD owns one contiguous run of slots (a spatial slab) and its 4 MB L2
// holds that slab's neighbourhood instead of the whole box.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/soup_abi.h"

#define SOUP_CAND_CAP 6u   // soup/wgsl/bond-dispatch.wgsl:12
#define SOUP_BOND_RULES 2u // soup/wgsl/bond-common.wgsl:41
#define SOUP_PI_F 3.14159265f
#define SOUP_RSET 16u      // resources per bond event (<= 15 used)


struct DevBuf {
  float4* pos;
  float4* vel;
  float4* force;
  float4* pos_at_rebuild;
  float* charge;
  uint4* bs;          // bond slots x,y,z; w unused
  uint32_t* clink;
  uint32_t* cheld;
  uint32_t* frozen;
  uint32_t* brng;
  uint32_t* trng;
  uint32_t* id;
  uint32_t* inv;
  uint32_t* cell_of;
  // grid scratch
  uint32_t* cell_count;
  uint32_t* cell_start;
  uint32_t* cell_rank;
  uint32_t* cell_tmp;
  // lists
  uint32_t* vl;       // by slot: row of cap entries, near part from the front, far part from the back
  uint32_t* vcount;
  uint32_t* vnear;    // by slot: entries of the near part (list_split)
  uint32_t* head_id;
  uint32_t* head_slot;
  uint32_t* head_of_slot; // by slot: head index, NONE for other particles
  uint32_t* head_flag;    // by slot: 1 for a head (scan input)
  uint32_t* head_pos;     // exclusive scan of head_flag (n + 1)
  uint32_t* head_rows;    // heads in slot (spatial) order
  float4* head_q;         // by head: (x, y, z, charge), refreshed before every force evaluation
  double* far_acc;        // by head: FP64 far DH force (x, y, z), summed before the list force
  float4* gq;             // by slot: (x, y, z, kind | 8 if charged), the list force's neighbour gather
  uint32_t* es;
  uint32_t* es_count;
  // counters: [0] verlet overflow, [1] es overflow, [2] max drift bits, [3] nonfinite pos,
  // [4] nonfinite vel, [5] bond event count, [6] bond resolution did not converge, [7] spare,
  // [8] list-force launches in which every row qualified for its near part, [9] the others
  uint32_t* flags;
  // list force, per wavefront of its grid: rows it read as the near part only, rows it read whole
  // (pairs of counts, accumulated over launches)
  uint32_t* rowreads;
  uint32_t* events;   // 2 * rules
  uint32_t* desorb;   // 2
  // bond MC scratch
  uint2* props;       // 6 per slot: (partner particle index, rule)
  uint32_t* nacc;     // accepted proposals per slot
  uint32_t* cat_id;   // catalystId per slot
  uint32_t* flag_cc;  // by particle index: touched by an accepted propagation proposal
  uint32_t* ev_flag;  // by particle index: has a bond event
  uint32_t* ev_pos;   // exclusive scan of ev_flag (n + 1)
  uint32_t* ev_id;    // event -> particle index, ascending
  uint32_t* resv;     // by particle index: reservation (lowest event index), NONE when free
  uint32_t* rset;     // SOUP_RSET per event
  uint32_t* ev_state; // per event: bit0 done, bits 8.. round stamp
  uint32_t* scan_blocks;
};

// ------------------------------------------------------------------------------------------------
// RNG (step.wgsl:89-94 == bond-common.wgsl:132-137)
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pcg_soup(uint32_t v) {
  uint32_t state = v * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
// f32(v) rounds to nearest even; the division by 2^32 is an exact scaling.
__device__ __forceinline__ float uniform01(uint32_t v) { return (float)v * 2.3283064365386963e-10f; }

// ------------------------------------------------------------------------------------------------
// Exact-order f32 geometry (used where results must be bit-identical to the oracle: cell
// assignment, list membership, bond-MC distances). Built with -ffp-contract=off and IEEE division
// and square root.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ float mi_f(float d, float L) { return d - rintf(d / L) * L; }
__device__ __forceinline__ float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }
__device__ __forceinline__ float dist_f(const SoupConfig& c, float4 a, float4 b) {
  return len3(mi_f(a.x - b.x, c.box[0]), mi_f(a.y - b.y, c.box[1]), mi_f(a.z - b.z, c.box[2]));
}
__device__ __forceinline__ float pair_b(const SoupConfig& c, float ti, float tj) {
  return c.sigma * (c.radius[(uint32_t)ti] + c.radius[(uint32_t)tj]) * 0.5f;
}
__device__ __forceinline__ float wca_cut(const SoupConfig& c, float b) { return c.wca_factor * b; }

// cell_of (engine/wgsl/neighbor.wgsl:19-24); identical to cell_coord for finite positions.
__device__ __forceinline__ uint32_t cell_axis(float p, float box, uint32_t dims) {
  float f = floorf((p / box) * (float)dims);
  f = fmaxf(f, 0.0f);
  uint32_t u = (f >= 4294967040.0f) ? 0xFFFFFFFFu : (uint32_t)f;
  return u > dims - 1u ? dims - 1u : u;
}
__device__ __forceinline__ uint32_t cell_of(const SoupConfig& c, float4 p) {
  uint32_t cx = cell_axis(p.x, c.box[0], c.dims[0]);
  uint32_t cy = cell_axis(p.y, c.box[1], c.dims[1]);
  uint32_t cz = cell_axis(p.z, c.box[2], c.dims[2]);
  return cx + c.dims[0] * (cy + c.dims[1] * cz);
}

#define SOUP_XCDS 8u
// logical block of an XCD-aware launch (grid = 8 * per blocks)
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t per = gridDim.x / SOUP_XCDS;
  return (blockIdx.x % SOUP_XCDS) * per + blockIdx.x / SOUP_XCDS;
}

__device__ __forceinline__ bool nonfinite_bits(float x) { return ((__float_as_uint(x) >> 23) & 0xFFu) == 0xFFu; }

// Stored stride of a Verlet row: the capacity rounded up to 32 entries, so that every row starts on
// its own 128-byte line (the capacity itself, 1126 at the campaign, keeps the overflow semantics)
__host__ __device__ __forceinline__ uint32_t vl_stride(const SoupConfig& c) { return (c.list_cap + 31u) & ~31u; }
