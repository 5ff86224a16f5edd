// soup_packed.inl -- packed launches of the step kernels over many systems (the replica sweep).
//
// One launch covers every system of a batch: each system owns a run of the grid's workgroups, found
// from a per-launch table of block offsets (off[r] .. off[r + 1] - 1 belong to system r, off[nrep] =
// the grid), and reads its own configuration and buffers from a per-system descriptor in device
// memory. The bodies are the single-system kernels' own (soup_kernels.inl), given the workgroup's
// index inside its system, so a packed step computes exactly what the same step of each system
// computes alone.
#pragma once

#include "soup_kernels.inl"

// one system of a batch, as the kernels see it
struct RepDesc {
  SoupConfig c;
  DevBuf b;     // the current buffer assignment
  DevBuf alt;   // the alternates a rebuild permutes into
  uint32_t nh;  // charged heads whose long-range terms run (0 when the DH term is off)
  uint32_t es;  // 1 when the list force adds the far DH sums
  uint32_t pad[2];
};

// the system a workgroup belongs to: the largest r with off[r] <= blk (systems with no workgroups in
// a launch have off[r] == off[r + 1] and are never chosen)
__device__ __forceinline__ uint32_t pack_find(const uint32_t* __restrict__ off, uint32_t nrep, uint32_t blk) {
  uint32_t lo = 0, hi = nrep;
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= blk) lo = mid;
    else hi = mid;
  }
  return uni(lo);
}

#define PACK_BEGIN(OFF)                                                  \
  const uint32_t r_ = pack_find((OFF), nrep, blockIdx.x);                \
  const uint32_t lb_ = uni(blockIdx.x - (OFF)[r_]);                      \
  const RepDesc* __restrict__ d_ = reps + r_;                            \
  (void)lb_;

__global__ __launch_bounds__(256) void kp_kick_drift(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                     uint32_t nrep) {
  PACK_BEGIN(off);
  kick_drift_body(d_->c, d_->b, lb_);
}
__global__ __launch_bounds__(256) void kp_thermo_drift(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                       uint32_t nrep) {
  PACK_BEGIN(off);
  thermo_drift_body(d_->c, d_->b, lb_);
}
__global__ __launch_bounds__(256) void kp_kick_thermo(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                      uint32_t nrep) {
  PACK_BEGIN(off);
  kick_thermo_body(d_->c, d_->b, lb_);
}

// ---- rebuild ---------------------------------------------------------------------------------------
// clears cell_count over each system's cells
__global__ __launch_bounds__(256) void kp_clear_cells(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                      uint32_t nrep) {
  PACK_BEGIN(off);
  const SoupConfig& c = d_->c;
  const uint32_t i = lb_ * blockDim.x + threadIdx.x;
  if (i < c.dims[0] * c.dims[1] * c.dims[2]) d_->b.cell_count[i] = 0u;
}
__global__ __launch_bounds__(256) void kp_cell_count(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                     uint32_t nrep) {
  PACK_BEGIN(off);
  cell_count_body(d_->c, d_->b, lb_);
}
__global__ __launch_bounds__(256) void kp_cell_fill(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                    uint32_t nrep) {
  PACK_BEGIN(off);
  cell_fill_body(d_->c, d_->b, reinterpret_cast<uint2*>(d_->alt.vel), lb_);  // alt.vel is free until the permute
}
__global__ __launch_bounds__(256) void kp_cell_rank(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                    uint32_t nrep) {
  PACK_BEGIN(off);
  cell_rank_body(d_->c, d_->b, reinterpret_cast<const uint2*>(d_->alt.vel), d_->b.cell_rank, lb_);
}
__global__ __launch_bounds__(256) void kp_permute(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                  uint32_t nrep) {
  PACK_BEGIN(off);
  permute_body(d_->c, d_->b, d_->alt, d_->b.cell_rank, lb_);
}
// head_of_slot = NONE and head_flag = 0 over each system's slots (systems with the DH term)
__global__ __launch_bounds__(256) void kp_clear_heads(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                      uint32_t nrep) {
  PACK_BEGIN(off);
  const uint32_t t = lb_ * blockDim.x + threadIdx.x;
  if (t < d_->c.n) {
    d_->b.head_of_slot[t] = SOUP_NONE;
    d_->b.head_flag[t] = 0u;
  }
}
__global__ __launch_bounds__(256) void kp_head_slots(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                     uint32_t nrep) {
  PACK_BEGIN(off);
  head_slots_body(d_->b, d_->nh, lb_);
}
__global__ __launch_bounds__(256) void kp_head_rows(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                    uint32_t nrep) {
  PACK_BEGIN(off);
  head_rows_body(d_->c, d_->b, lb_);
}
__global__ __launch_bounds__(256) void kp_es_build(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                   uint32_t nrep) {
  PACK_BEGIN(off);
  es_build_body(d_->c, d_->b, d_->nh, lb_);
}
__global__ __launch_bounds__(256) void kp_verlet_build(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                       uint32_t nrep) {
  PACK_BEGIN(off);
  verlet_build_body(d_->c, d_->b, lb_);
}
// the drift maximum restarts from zero after a rebuild: thread r clears system r's
__global__ __launch_bounds__(64) void kp_clear_drift(const RepDesc* __restrict__ reps, uint32_t nrep) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nrep) reps[r].b.flags[2] = 0u;
}

// ---- exclusive scans: which = 0 cell_count -> cell_start over the cells, 1 head_flag -> head_pos
// and 2 ev_flag -> ev_pos over the slots ------------------------------------------------------------
__device__ __forceinline__ void scan_arrays(const RepDesc* d, uint32_t which, const uint32_t*& in, uint32_t*& out,
                                            uint32_t& m) {
  const SoupConfig& c = d->c;
  if (which == 0u) {
    in = d->b.cell_count;
    out = d->b.cell_start;
    m = c.dims[0] * c.dims[1] * c.dims[2];
  } else if (which == 1u) {
    in = d->b.head_flag;
    out = d->b.head_pos;
    m = c.n;
  } else {
    in = d->b.ev_flag;
    out = d->b.ev_pos;
    m = c.n;
  }
}
__global__ __launch_bounds__(SCAN_BLOCK) void kp_scan_reduce(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                             uint32_t nrep, uint32_t which) {
  PACK_BEGIN(off);
  const uint32_t* in;
  uint32_t* out;
  uint32_t m;
  scan_arrays(d_, which, in, out, m);
  scan_reduce_body(in, m, d_->b.scan_blocks, lb_);
}
// one workgroup per system that runs the scan
__global__ __launch_bounds__(SCAN_BLOCK) void kp_scan_blocks(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                             uint32_t nrep, uint32_t which) {
  PACK_BEGIN(off);
  const uint32_t* in;
  uint32_t* out;
  uint32_t m;
  scan_arrays(d_, which, in, out, m);
  uint32_t nb = (m + SCAN_TILE - 1u) / SCAN_TILE;
  if (nb == 0u) nb = 1u;
  scan_blocks_body(d_->b.scan_blocks, nb, 0u);
}
__global__ __launch_bounds__(SCAN_BLOCK) void kp_scan_apply(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                            uint32_t nrep, uint32_t which) {
  PACK_BEGIN(off);
  const uint32_t* in;
  uint32_t* out;
  uint32_t m;
  scan_arrays(d_, which, in, out, m);
  uint32_t nb = (m + SCAN_TILE - 1u) / SCAN_TILE;
  if (nb == 0u) nb = 1u;
  scan_apply_body(in, m, d_->b.scan_blocks, nb, out, lb_);
}

// ---- force -----------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void kp_head_refresh(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                       uint32_t nrep) {
  PACK_BEGIN(off);
  head_refresh_body(d_->b, d_->nh, lb_);
}
__global__ __launch_bounds__(256) void kp_far_dh(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                 uint32_t nrep) {
  PACK_BEGIN(off);
  far_dh_body(d_->c, d_->b, d_->nh, lb_);
}
__global__ __launch_bounds__(256) void kp_force_group(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                      uint32_t nrep) {
  PACK_BEGIN(off);
  force_group_body(d_->c, d_->b, (int)d_->es, lb_);
}

// ---- bond Monte Carlo --------------------------------------------------------------------------------
// flag_cc = ev_flag = 0 over each system's particles
__global__ __launch_bounds__(256) void kp_clear_bond_flags(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                           uint32_t nrep) {
  PACK_BEGIN(off);
  const uint32_t t = lb_ * blockDim.x + threadIdx.x;
  if (t < d_->c.n) {
    d_->b.flag_cc[t] = 0u;
    d_->b.ev_flag[t] = 0u;
  }
}
__global__ __launch_bounds__(256) void kp_bond_propose(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                       uint32_t nrep) {
  PACK_BEGIN(off);
  bond_propose_body(d_->c, d_->b, lb_);
}
__global__ __launch_bounds__(256) void kp_bond_filter(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                      uint32_t nrep) {
  PACK_BEGIN(off);
  bond_filter_body(d_->c, d_->b, lb_);
}
__global__ __launch_bounds__(256) void kp_bond_events(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                      uint32_t nrep) {
  PACK_BEGIN(off);
  bond_events_body(d_->c, d_->b, lb_);
}
// one workgroup per system
__global__ __launch_bounds__(1024) void kp_bond_resolve(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                        uint32_t nrep) {
  PACK_BEGIN(off);
  bond_resolve_body(d_->c, d_->b, 0u);
}
__global__ __launch_bounds__(256) void kp_bond_break(const RepDesc* __restrict__ reps, const uint32_t* __restrict__ off,
                                                     uint32_t nrep) {
  PACK_BEGIN(off);
  bond_break_body(d_->c, d_->b, lb_);
}
