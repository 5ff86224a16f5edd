// soup_lib.hip -- host side of libsoup_hip.so: the C-ABI declared in include/soup_abi.h.
//
// One handle = one periodic box on the current HIP device, one stream. The step orchestration
// mirrors soup/src/soup-integrate.ts:83-175 (per step: kick/drift/wrap, [rebuild every
// rebuildEvery], list force + far DH, drift max, [bond MC every bondAttemptInterval], kick/thermostat;
// per 1000-step chunk: the non-finite guard, the Verlet/DH guards and the due constant-pH sweep).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "soup_kernels.inl"
#include "soup_packed.inl"

#define SOUP_VERSION "libsoup_hip 0.1 (gfx950)"

static thread_local std::string g_create_error;

struct SoupHandle;
static void forget_batches_of(const SoupHandle* h);

// Per-phase device timing with HIP events on the handle's stream (soup_profile): every bracketed
// launch gets an event pair from a pool; pairs are resolved at the next synchronisation.
struct Prof {
  uint32_t mask = 0;  // bit k: phase k is bracketed
  double ms[SOUP_PROF_PHASES] = {0};
  uint64_t count[SOUP_PROF_PHASES] = {0};
  std::vector<hipEvent_t> pool;
  struct Pending {
    int phase;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
};

struct SoupHandle {
  SoupConfig c;
  uint32_t n = 0, ncells = 0, nh = 0;
  uint32_t cap = 0;              // particles every per-particle buffer holds (n <= cap)
  std::vector<uint32_t> heads;   // particle indices of the charged heads, ascending
  hipStream_t stream = nullptr;
  int device = 0;                // the HIP device current at creation, which holds every buffer
  DevBuf b{};
  DevBuf alt{};
  double2* ptab = nullptr;     // constant-pH pair table: per DH entry the energy (first nh x cap), then the wells
  uint32_t* ptab_g = nullptr;  // the wells' head indices
  uint32_t* pwcnt = nullptr;   // wells per head row
  uint32_t* pstats = nullptr;  // [0..3] chain stats, [4] rng
  double* pdstats = nullptr;
  unsigned long long* lsum = nullptr;
  uint32_t* lmax = nullptr;
  uint8_t* stage = nullptr;    // device staging for boundary translation
  size_t stage_bytes = 0;
  uint64_t global_step = 0, next_sweep_at = UINT64_MAX;
  uint32_t prng = 0;
  uint32_t sweeps = 0;
  size_t ncells_cap = 0;     // cell_count / cell_start capacity (a box change may need more)
  // head cell grid of the DH-list build (used when every axis has at least 5 cells)
  EsGrid esg{};
  bool es_grid = false;
  size_t es_cells_cap = 0;
  // soup_potential_at scratch: positions in particle order, the configuration's cell grid
  float4* pa_xs = nullptr;
  uint32_t *pa_cell = nullptr, *pa_order = nullptr, *pa_count = nullptr, *pa_start = nullptr;
  size_t pa_ncell_cap = 0;
  // soup_energy_adjoint scratch (FP64): direction, gradient, Hessian-vector product, parameter partials
  double *adj_w = nullptr, *adj_g = nullptr, *adj_hw = nullptr, *adj_part = nullptr, *adj_out = nullptr;
  SoupSweepResult last_sweep{};
  std::string err;
  std::vector<void*> allocs;
  Prof prof;
  // captured step periods (soup_step): one per assignment of the double-buffered state arrays
  struct StepGraph {
    hipGraphExec_t exec = nullptr;
    float4* key = nullptr;  // b.pos when captured
    uint32_t len = 0;       // steps
    uint32_t phase = 0;     // global_step % bondAttemptInterval at the first step
  };
  std::vector<StepGraph> graphs;
  // drift-triggered rebuilds (SOUP_REBUILD_DRIFT): the next step rebuilds; steps since the last check
  bool rebuild_due = false;
  uint64_t rebuilds = 0;  // list rebuilds made by steps (the priming and forced ones not counted)
  uint64_t gen = 0;       // bumped whenever the configuration, the particle count or a buffer changes
  uint32_t rowread_slots = 0;  // wavefronts of the list force's largest grid (b.rowreads holds two counts each)
  bool use_graphs = true;
};

static hipEvent_t prof_event(SoupHandle* h) {
  if (!h->prof.pool.empty()) {
    hipEvent_t e = h->prof.pool.back();
    h->prof.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}
// brackets the launches issued between prof_begin and prof_end
static hipEvent_t prof_begin(SoupHandle* h, int phase) {
  if (!(h->prof.mask >> phase & 1u)) return nullptr;
  hipEvent_t a = prof_event(h);
  (void)hipEventRecord(a, h->stream);
  return a;
}
static void prof_end(SoupHandle* h, int phase, hipEvent_t a) {
  if (!a) return;
  hipEvent_t b = prof_event(h);
  (void)hipEventRecord(b, h->stream);
  h->prof.pending.push_back({phase, a, b});
}
static void prof_resolve(SoupHandle* h) {
  for (auto& p : h->prof.pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      h->prof.ms[p.phase] += ms;
      h->prof.count[p.phase] += 1;
    }
    h->prof.pool.push_back(p.a);
    h->prof.pool.push_back(p.b);
  }
  h->prof.pending.clear();
}

static int fail(SoupHandle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  else g_create_error = msg;
  return code;
}

static std::string fmt(const char* f, ...) {
  char buf[2048];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof(buf), f, ap);
  va_end(ap);
  return buf;
}

#define HC(h, x)                                                                               \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) return fail((h), SOUP_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

static inline uint32_t blocks(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }
// grid of an XCD-aware launch: a multiple of the 8 XCDs (kernels map blockIdx with xcd_block())
static inline uint32_t xcd_blocks(uint64_t n, uint32_t bs) {
  uint32_t b = blocks(n, bs);
  return ((b + SOUP_XCDS - 1) / SOUP_XCDS) * SOUP_XCDS;
}

template <typename T>
static int dalloc(SoupHandle* h, T** p, size_t count) {
  size_t bytes = count * sizeof(T);
  if (bytes == 0) bytes = sizeof(T);
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess)
    return fail(h, SOUP_E_NOMEM, fmt("device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e)));
  h->allocs.push_back(q);
  *p = reinterpret_cast<T*>(q);
  return 0;
}

// frees a dalloc'd buffer (the stream must be idle)
static void dfree(SoupHandle* h, void* p) {
  if (!p) return;
  auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
  if (it != h->allocs.end()) {
    (void)hipFree(p);
    h->allocs.erase(it);
  }
}
// a buffer of at least `count` elements, reallocated (contents dropped) when `cap` is smaller
template <typename T>
static int dgrow(SoupHandle* h, T** p, size_t count, size_t& cap) {
  if (*p && count <= cap) return 0;
  dfree(h, *p);
  *p = nullptr;
  cap = 0;
  int r = dalloc(h, p, count);
  if (!r) cap = count;
  return r;
}

#define DA(h, p, cnt)                        \
  do {                                       \
    int r_ = dalloc((h), &(p), (size_t)(cnt)); \
    if (r_) return r_;                       \
  } while (0)

static void fill(SoupHandle* h, uint32_t* p, uint32_t v, uint32_t count) {
  if (count) hipLaunchKernelGGL(k_fill_u32, dim3(blocks(count, 256)), dim3(256), 0, h->stream, p, v, count);
}

// exclusive scan of m u32 -> out[0..m]; block_sums holds SCAN_TILE + 1 entries
static void scan_on(hipStream_t st, uint32_t* block_sums, const uint32_t* in, uint32_t m, uint32_t* out) {
  if (m <= SCAN_ONE_MAX) {  // one launch instead of three
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(SCAN_BLOCK), 0, st, in, m, out);
    return;
  }
  uint32_t nb = blocks(m, SCAN_TILE);
  if (nb == 0) nb = 1;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_BLOCK), 0, st, in, m, block_sums);
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(SCAN_BLOCK), 0, st, block_sums, nb);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_BLOCK), 0, st, in, m, block_sums, nb, out);
}
static void scan(SoupHandle* h, const uint32_t* in, uint32_t m, uint32_t* out) { scan_on(h->stream, h->b.scan_blocks, in, m, out); }

static bool es_on(const SoupHandle* h) { return h->c.es_enabled && h->nh > 0; }

// phases that run only at chunk ends, outside every captured step period
static const uint32_t kChunkEndPhases = (1u << SOUP_PROF_PROTON) | (1u << SOUP_PROF_SWEEP);
// whether steps may be replayed from captured graphs: not while a per-step phase is bracketed
static bool replay_ok(const SoupHandle* h) { return h->use_graphs && (h->prof.mask & ~kChunkEndPhases) == 0; }

// ---- phases ------------------------------------------------------------------------------------
static void ph_kick_drift(SoupHandle* h) {
  hipEvent_t e = prof_begin(h, SOUP_PROF_KICK_DRIFT);
  hipLaunchKernelGGL(k_kick_drift, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->c, h->b);
  prof_end(h, SOUP_PROF_KICK_DRIFT, e);
}
// the thermostat half of the previous step fused with this step's first half (see k_thermo_drift)
static void ph_thermo_drift(SoupHandle* h) {
  hipEvent_t e = prof_begin(h, SOUP_PROF_KICK_DRIFT);
  hipLaunchKernelGGL(k_thermo_drift, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->c, h->b);
  prof_end(h, SOUP_PROF_KICK_DRIFT, e);
}
static void ph_kick_thermo(SoupHandle* h) {
  hipEvent_t e = prof_begin(h, SOUP_PROF_KICK_THERMO);
  hipLaunchKernelGGL(k_kick_thermo, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->c, h->b);
  prof_end(h, SOUP_PROF_KICK_THERMO, e);
}
static void ph_drift(SoupHandle* h) {
  hipEvent_t e = prof_begin(h, SOUP_PROF_DRIFT);
  hipLaunchKernelGGL(k_max_drift, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->c, h->b);
  prof_end(h, SOUP_PROF_DRIFT, e);
}

// encodeVerletRebuild (soup/src/soup-grid-verlet.ts:61-73): grid, DH list, Verlet list, snapshot,
// drift reset -- here with the permutation of the state into the new cell order.
// The DH-list build's head grid for the current box: cells of at least (listRange + 0.02) / 2 per
// axis, so that the 5^3 cells around a head hold every head the list test can admit. Small boxes and
// few heads keep the all-pairs build. Called outside every captured period (creation, box change).
static int es_grid_setup(SoupHandle* h) {
  h->es_grid = false;
  if (!es_on(h) || h->nh < 256u) return 0;
  const double w_min = 0.5 * ((double)h->c.es_list_range + 0.02);
  size_t ncells = 1;
  for (int a = 0; a < 3; ++a) {
    const double L = (double)h->c.box[a];
    const uint32_t d = (uint32_t)std::floor(L / w_min);
    if (d < 5u) return 0;
    h->esg.d[a] = d;
    h->esg.inv_w[a] = (float)((double)d / L);
    ncells *= d;
  }
  if (ncells > (size_t)64 * 1024 * 1024) return 0;
  if (ncells > h->es_cells_cap) {
    size_t cc = h->es_cells_cap, cs = h->es_cells_cap;
    int r = dgrow(h, &h->esg.count, ncells, cc);
    if (r) return r;
    r = dgrow(h, &h->esg.start, ncells + 1, cs);
    if (r) return r;
    h->es_cells_cap = ncells;
  }
  if (!h->esg.rec) {
    int r = dalloc(h, &h->esg.rec, h->nh);
    if (r) return r;
  }
  h->es_grid = true;
  return 0;
}

static void ph_rebuild(SoupHandle* h) {
  const uint32_t n = h->n;
  hipEvent_t e0 = prof_begin(h, SOUP_PROF_SORT);
  fill(h, h->b.cell_count, 0u, h->ncells);
  hipLaunchKernelGGL(k_cell_count, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->c, h->b);
  scan(h, h->b.cell_count, h->ncells, h->b.cell_start);
  // the sort keys go to the alternate velocity buffer, which the permute overwrites next
  uint2* keys = reinterpret_cast<uint2*>(h->alt.vel);
  hipLaunchKernelGGL(k_cell_fill, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->c, h->b, keys);
  hipLaunchKernelGGL(k_cell_rank, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->c, h->b, keys, h->b.cell_rank);
  hipLaunchKernelGGL(k_permute, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->c, h->b, h->alt, h->b.cell_rank);
  std::swap(h->b.pos, h->alt.pos);
  std::swap(h->b.vel, h->alt.vel);
  std::swap(h->b.force, h->alt.force);
  std::swap(h->b.charge, h->alt.charge);
  std::swap(h->b.bs, h->alt.bs);
  std::swap(h->b.clink, h->alt.clink);
  std::swap(h->b.cheld, h->alt.cheld);
  std::swap(h->b.frozen, h->alt.frozen);
  std::swap(h->b.brng, h->alt.brng);
  std::swap(h->b.trng, h->alt.trng);
  std::swap(h->b.id, h->alt.id);
  prof_end(h, SOUP_PROF_SORT, e0);
  if (es_on(h)) {
    hipEvent_t e1 = prof_begin(h, SOUP_PROF_ES_BUILD);
    // head_of_slot and head_flag were cleared by the permutation
    hipLaunchKernelGGL(k_head_slots, dim3(blocks(h->nh, 256)), dim3(256), 0, h->stream, h->b, h->nh);
    scan(h, h->b.head_flag, n, h->b.head_pos);
    hipLaunchKernelGGL(k_head_rows, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->c, h->b);
    if (h->es_grid) {
      const uint32_t nc = h->esg.d[0] * h->esg.d[1] * h->esg.d[2];
      fill(h, h->esg.count, 0u, nc);
      hipLaunchKernelGGL(k_es_cell_count, dim3(blocks(h->nh, 256)), dim3(256), 0, h->stream, h->b, h->nh, h->esg);
      scan(h, h->esg.count, nc, h->esg.start);
      fill(h, h->esg.count, 0u, nc);
      hipLaunchKernelGGL(k_es_cell_fill, dim3(blocks(h->nh, 256)), dim3(256), 0, h->stream, h->b, h->nh, h->esg);
      const size_t lds = (size_t)ES_CELL_WAVES * ((h->nh + 31u) / 32u) * 4u;
      hipLaunchKernelGGL(k_es_cells, dim3(blocks(h->nh, ES_CELL_WAVES)), dim3(ES_CELL_WAVES * WAVE), lds, h->stream, h->c, h->b,
                         h->nh, h->esg);
    } else {
      hipLaunchKernelGGL(k_es_build, dim3(blocks((uint64_t)h->nh * WAVE, 256)), dim3(256), 0, h->stream, h->c, h->b, h->nh);
    }
    prof_end(h, SOUP_PROF_ES_BUILD, e1);
  }
  hipEvent_t e2 = prof_begin(h, SOUP_PROF_VERLET_BUILD);
  if (VL_CELLS) {
    const uint64_t ncells = (uint64_t)h->c.dims[0] * h->c.dims[1] * h->c.dims[2];
    hipLaunchKernelGGL(k_verlet_cells, dim3(xcd_blocks(ncells, 1)), dim3(VL_WAVES * WAVE), 0, h->stream, h->c, h->b);
  } else {
    hipLaunchKernelGGL(k_verlet_build, dim3(xcd_blocks((uint64_t)n * WAVE, 256)), dim3(256), 0, h->stream, h->c, h->b);
  }
  prof_end(h, SOUP_PROF_VERLET_BUILD, e2);  // the permutation restarted the drift maximum
}

// far DH rows of the charged heads (FP64 per head), then the list force, which adds them to the
// heads' sums before rounding
static void ph_force(SoupHandle* h) {
  if (es_on(h)) {
    hipEvent_t e0 = prof_begin(h, SOUP_PROF_ES_FAR);
    hipLaunchKernelGGL(k_head_refresh, dim3(blocks(h->nh, 256)), dim3(256), 0, h->stream, h->b, h->nh);
    hipLaunchKernelGGL(k_far_dh, dim3(xcd_blocks((uint64_t)h->nh * WAVE, 256)), dim3(256), 0, h->stream, h->c, h->b, h->nh);
    prof_end(h, SOUP_PROF_ES_FAR, e0);
  }
  hipEvent_t e = prof_begin(h, SOUP_PROF_FORCE);
  hipLaunchKernelGGL(k_force_group, dim3(xcd_blocks((uint64_t)h->n * FORCE_TPP, 256)), dim3(256), 0, h->stream, h->c, h->b,
                     es_on(h) ? 1 : 0);
  prof_end(h, SOUP_PROF_FORCE, e);
}

static void ph_force_brute(SoupHandle* h) {
  hipLaunchKernelGGL(k_force_brute, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->c, h->b);
  if (es_on(h)) hipLaunchKernelGGL(k_es_far_brute, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->c, h->b);
}

// encodeBondChemistry (soup/src/soup-bond-chemistry.ts:24-40): formation, then breaking
static void ph_bonds(SoupHandle* h) {
  const uint32_t n = h->n;
  hipEvent_t e = prof_begin(h, SOUP_PROF_BONDS);
  fill(h, h->b.flag_cc, 0u, n);
  fill(h, h->b.ev_flag, 0u, n);
  hipLaunchKernelGGL(k_bond_propose, dim3(xcd_blocks((uint64_t)n * WAVE, 256)), dim3(256), 0, h->stream, h->c, h->b);
  hipLaunchKernelGGL(k_bond_filter, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->c, h->b);
  scan(h, h->b.ev_flag, n, h->b.ev_pos);
  hipLaunchKernelGGL(k_bond_events, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->c, h->b);
  hipLaunchKernelGGL(k_bond_resolve, dim3(1), dim3(1024), 0, h->stream, h->c, h->b);
  hipLaunchKernelGGL(k_bond_break, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->c, h->b);
  prof_end(h, SOUP_PROF_BONDS, e);
}

// thermo_in: the previous step left its thermostat half to this step's first kernel; thermo_out:
// this step leaves its own to the next (steps that follow each other inside a replayed period)
static bool drift_mode(const SoupHandle* h) { return h->c.rebuild_mode == SOUP_REBUILD_DRIFT; }
// whether the next step rebuilds: every rebuildEvery steps (soup/src/soup-integrate.ts:149), or, in
// drift mode, when the last drift check asked for it
static bool rebuild_next(const SoupHandle* h) {
  return drift_mode(h) ? h->rebuild_due : (h->global_step % h->c.rebuild_every) == 0;
}

static void one_step(SoupHandle* h, bool thermo_in = false, bool thermo_out = false) {
  const bool do_bonds = (h->global_step % h->c.bond_interval) == 0;
  const bool do_rebuild = rebuild_next(h);
  if (do_rebuild) {
    h->rebuilds++;
    h->rebuild_due = false;
  }
  // also the drift max of this step's positions (soup-integrate.ts:96-98)
  if (thermo_in) ph_thermo_drift(h);
  else ph_kick_drift(h);
  if (do_rebuild) ph_rebuild(h);
  ph_force(h);
  if (do_bonds) ph_bonds(h);
  if (!thermo_out) ph_kick_thermo(h);
  h->global_step++;
}

// Steps are replayed from captured HIP graphs one period at a time: lcm(rebuildEvery,
// bondAttemptInterval) steps, doubled when that holds an odd number of rebuilds, so that the
// double-buffered state arrays end a period where they began and the graph's baked pointers stay
// valid. A period starts where global_step is a multiple of its length.
static uint32_t gcd_u(uint32_t a, uint32_t b) { return b ? gcd_u(b, a % b) : a; }
static uint32_t step_period(const SoupConfig& c) {
  const uint64_t p = (uint64_t)c.rebuild_every / gcd_u(c.rebuild_every, c.bond_interval) * c.bond_interval;
  const uint64_t q = (p / c.rebuild_every) % 2 ? 2 * p : p;
  return q <= 1000u ? (uint32_t)q : 0u;  // 0: no replay (a period longer than a chunk)
}
static void drop_graphs(SoupHandle* h) {
  for (auto& g : h->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  h->graphs.clear();
  h->gen++;  // every change that voids a captured step also voids a batch's view of this system
}
// P steps through the graph captured for the current buffer assignment and bond phase (captured and
// instantiated on first use); false if capture is unavailable, with the host state untouched. A
// cadence period holds an even number of rebuilds, so the buffer assignment it ends with is the one
// it began with; a drift-mode segment holds none.
static bool run_period(SoupHandle* h, uint32_t P) {
  const uint32_t phase = (uint32_t)(h->global_step % h->c.bond_interval);
  const uint64_t rebuilds = drift_mode(h) ? 0u : P / h->c.rebuild_every;
  for (auto& s : h->graphs)
    if (s.exec && s.key == h->b.pos && s.len == P && s.phase == phase) {
      if (hipGraphLaunch(s.exec, h->stream) != hipSuccess) return false;
      h->global_step += P;  // the host side of those steps: an even number of buffer swaps undoes itself
      h->rebuilds += rebuilds;
      return true;
    }
  if (h->graphs.size() >= 48) {  // a bounded cache: the oldest capture goes
    (void)hipGraphExecDestroy(h->graphs.front().exec);
    h->graphs.erase(h->graphs.begin());
  }
  float4* key = h->b.pos;
  const uint64_t step0 = h->global_step, rebuilds0 = h->rebuilds;
  const DevBuf b0 = h->b, a0 = h->alt;
  if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
  for (uint32_t k = 0; k < P; ++k) one_step(h, k > 0, k + 1 < P);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(h->stream, &graph);
  hipGraphExec_t exec = nullptr;
  const bool ok = ec == hipSuccess && graph && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (graph) (void)hipGraphDestroy(graph);
  // nothing ran while capturing: restore the host state, then run the period from the graph
  h->global_step = step0;
  h->rebuilds = rebuilds0;
  h->b = b0;
  h->alt = a0;
  (void)hipGetLastError();
  if (!ok) {
    if (exec) (void)hipGraphExecDestroy(exec);
    h->use_graphs = false;
    return false;
  }
  SoupHandle::StepGraph g;
  g.exec = exec;
  g.key = key;
  g.len = P;
  g.phase = phase;
  h->graphs.push_back(g);
  if (hipGraphLaunch(exec, h->stream) != hipSuccess) return false;
  h->global_step += P;
  h->rebuilds += rebuilds;
  return true;
}

static int sync(SoupHandle* h) {
  HC(h, hipGetLastError());
  HC(h, hipStreamSynchronize(h->stream));
  if (!h->prof.pending.empty()) prof_resolve(h);
  return 0;
}

// ---- guards --------------------------------------------------------------------------------------
static int nonfinite_counts(SoupHandle* h, uint32_t out[2]) {
  fill(h, h->b.flags + 3, 0u, 2);
  hipLaunchKernelGGL(k_nonfinite, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->c, h->b);
  HC(h, hipMemcpyAsync(out, h->b.flags + 3, 8, hipMemcpyDeviceToHost, h->stream));
  return sync(h);
}

// assertStateFinite (soup/src/soup-health.ts:55-63)
static int assert_finite(SoupHandle* h, bool has_since, uint64_t since) {
  uint32_t c[2];
  int r = nonfinite_counts(h, c);
  if (r) return r;
  if (c[0] == 0 && c[1] == 0) return 0;
  std::string window = has_since ? fmt(" (appeared in the step interval %llu..%llu)", (unsigned long long)since,
                                       (unsigned long long)h->global_step)
                                 : std::string();
  return fail(h, SOUP_E_NONFINITE,
              fmt("non-finite state at step=%llu: non-finite components of positions=%u, velocities=%u of %llu%s -- the "
                  "computation diverged (Inf/NaN), any further numbers from this run are meaningless",
                  (unsigned long long)h->global_step, c[0], c[1], (unsigned long long)h->n * 3ull, window.c_str()));
}

static std::string js_fixed(double x, int d) { return fmt("%.*f", d, x); }

// assertVerletSafety (soup/src/soup-grid-verlet.ts:91-138)
static int assert_lists(SoupHandle* h) {
  uint32_t f[8];
  HC(h, hipMemcpyAsync(f, h->b.flags, 32, hipMemcpyDeviceToHost, h->stream));
  int r = sync(h);
  if (r) return r;
  if (f[6] != 0) return fail(h, SOUP_E_DEVICE, "bond Monte Carlo: the ordered resolution of the accepted proposals did not converge");
  if (f[0] != 0)
    return fail(h, SOUP_E_OVERFLOW,
                fmt("Verlet list: verletList.listCapacity=%u was insufficient -- at least one particle found more "
                    "candidates than the list holds (data may have been silently dropped)",
                    h->c.list_cap));
  if (es_on(h) && f[1] != 0)
    return fail(h, SOUP_E_OVERFLOW,
                fmt("long-range electrostatics list: longRangeListCapacity=%u or the headIdx size (%u) was insufficient at "
                    "rc_es=%s (heads found=%u) -- data may have been silently dropped",
                    h->c.es_cap, h->c.n_heads, js_fixed(h->c.es_rc, 4).c_str(), h->nh));
  float d2;
  memcpy(&d2, &f[2], 4);
  double drift = std::sqrt(std::fmax(0.0, (double)d2));
  double bound = (double)h->c.skin / 2.0;
  if (!std::isfinite(drift))
    return fail(h, SOUP_E_DRIFT,
                "Verlet list: the measured drift is not a number (NaN) -- the state is already non-finite, see "
                "soup/src/soup-health.ts's assertStateFinite");
  if (drift > bound + 1e-6)
    return fail(h, SOUP_E_DRIFT,
                fmt("Verlet list: the measured drift %s exceeds skin/2=%s -- the analytic bound (verletList.basis) did not "
                    "hold for the real trajectory, the rebuild was not frequent enough",
                    js_fixed(drift, 4).c_str(), js_fixed(bound, 4).c_str()));
  return 0;
}

// ---- constant-pH sweep -------------------------------------------------------------------------
static int sweep(SoupHandle* h, SoupSweepResult* out) {
  hipEvent_t e_sweep = prof_begin(h, SOUP_PROF_SWEEP);
  ph_rebuild(h);  // DH list at the current positions
  double* tabE = reinterpret_cast<double*>(h->ptab);
  double* tabW = tabE + (size_t)h->nh * h->c.es_cap;
  hipLaunchKernelGGL(k_proton_pairs, dim3(blocks((uint64_t)h->nh * WAVE, 256)), dim3(256), 0, h->stream, h->c, h->b, h->nh,
                     tabE, tabW, h->ptab_g, h->pwcnt);
  fill(h, h->pstats, 0u, 4);
  HC(h, hipMemcpyAsync(h->pstats + 4, &h->prng, 4, hipMemcpyHostToDevice, h->stream));
  size_t lds = (size_t)h->nh * sizeof(float);
  hipEvent_t e = prof_begin(h, SOUP_PROF_PROTON);
  hipLaunchKernelGGL(k_proton_chain, dim3(1), dim3(CHAIN_THREADS), lds, h->stream, h->c, h->b, h->nh, tabE, tabW,
                     h->ptab_g, h->pwcnt, h->pstats + 4, h->pstats, h->pdstats);
  prof_end(h, SOUP_PROF_PROTON, e);
  uint32_t st[5];
  double ds;
  HC(h, hipMemcpyAsync(st, h->pstats, 20, hipMemcpyDeviceToHost, h->stream));
  HC(h, hipMemcpyAsync(&ds, h->pdstats, 8, hipMemcpyDeviceToHost, h->stream));
  int r = sync(h);
  if (r) return r;
  h->prng = st[4];
  SoupSweepResult res;
  res.attempts = st[0];
  res.accepted = st[1];
  res.heads = h->nh;
  res.deprotonated = st[2];
  res.alpha = h->nh > 0 ? (double)st[2] / (double)h->nh : 0.0;
  res.dEsMeanKT = st[3] > 0 ? ds / (double)st[3] : 0.0;
  if (out) *out = res;
  h->last_sweep = res;
  h->sweeps++;
  // F(x) is stale with respect to the new charges (soup-protonation.ts:259-271)
  ph_rebuild(h);
  ph_force(h);
  prof_end(h, SOUP_PROF_SWEEP, e_sweep);
  return sync(h);
}

template <typename T>
static void to_index(SoupHandle* h, const T* src) {
  hipLaunchKernelGGL(k_to_index<T>, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, src, h->b.id,
                     reinterpret_cast<T*>(h->stage), h->n);
}

// ================================================================================================
// C-ABI
// ================================================================================================
extern "C" {

const char* soup_version(void) { return SOUP_VERSION; }

const char* soup_last_error(const SoupHandle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

static int validate(const SoupConfig* c) {
  if (!c) return fail(nullptr, SOUP_E_INVALID, "soup_create: null config");
  if (c->n == 0) return fail(nullptr, SOUP_E_INVALID, "createSoup: the starting composition is empty");
  if (c->n_kinds == 0 || c->n_kinds > SOUP_MAX_KINDS)
    return fail(nullptr, SOUP_E_INVALID, fmt("data/soup.json: %u monomer kinds, but the shader holds at most 8", c->n_kinds));
  for (int a = 0; a < 3; ++a)
    if (c->dims[a] < 2 * c->walk_radius + 1)
      return fail(nullptr, SOUP_E_INVALID,
                  fmt("neighbour grid: dims=[%u,%u,%u] needs dims>=%u (2*effectiveWalkRadius+1) on all three axes",
                      c->dims[0], c->dims[1], c->dims[2], 2 * c->walk_radius + 1));
  if (c->walk_radius > 3)
    return fail(nullptr, SOUP_E_INVALID,
                fmt("soup_create: list walk radius %u exceeds 3 (the list build sets up one cell row per lane)", c->walk_radius));
  if (c->n >= (1u << 29) || c->capacity >= (1u << 29))
    return fail(nullptr, SOUP_E_INVALID,
                fmt("soup_create: N=%u (capacity %u) exceeds the 2^29 - 1 particles the kernels address", c->n, c->capacity));
  if (c->capacity != 0 && c->capacity < c->n)
    return fail(nullptr, SOUP_E_INVALID,
                fmt("createSoup: the live composition (%u) is larger than the starting one (%u) -- buffers are allocated for the "
                    "starting one",
                    c->n, c->capacity));
  if (c->rebuild_mode > SOUP_REBUILD_DRIFT)
    return fail(nullptr, SOUP_E_INVALID, fmt("soup_create: unknown rebuild mode %u", c->rebuild_mode));
  if (c->rebuild_mode == SOUP_REBUILD_DRIFT &&
      (c->drift_check_every == 0 || c->drift_check_every > 1000u || !(c->drift_trigger > 0.0f) ||
       !(c->drift_trigger <= 0.5f * c->skin)))
    return fail(nullptr, SOUP_E_INVALID,
                fmt("verletList drift mode: driftCheckEvery=%u must be 1..1000 and driftTrigger=%g must lie in (0, skin/2=%g]",
                    c->drift_check_every, (double)c->drift_trigger, 0.5 * c->skin));
  if (c->list_cap == 0 || c->rebuild_every == 0 || c->bond_interval == 0)
    return fail(nullptr, SOUP_E_INVALID, "soup_create: list capacity, rebuildEvery and bondAttemptInterval must be > 0");
  if (c->es_enabled && c->n_heads > 0 && c->es_cap == 0)
    return fail(nullptr, SOUP_E_INVALID, "soup_create: long-range list capacity must be > 0");
  if (c->es_enabled && (size_t)c->n_heads * sizeof(float) > 150u * 1024u)
    return fail(nullptr, SOUP_E_INVALID,
                fmt("soup_create: %u titratable heads exceed the constant-pH chain's LDS budget (38400)", c->n_heads));
  return 0;
}

// The initial-state arrays for n particles: every index the kernels follow must be in range before
// anything reaches the device (h is null while soup_create has no handle yet)
static int check_state(SoupHandle* h, const SoupConfig& c, uint32_t n, const SoupInitState* st, const char* who) {
  if (!st || !st->positions || !st->velocities || !st->bond_slots || !st->center_link || !st->center_held || !st->frozen ||
      !st->bond_rng || !st->thermo_rng || !st->charges)
    return fail(h, SOUP_E_INVALID, fmt("%s: every initial-state array is required", who));
  for (uint32_t i = 0; i < n; ++i) {
    float w = st->positions[4u * i + 3];
    if (!(w >= 0.0f) || w != std::floor(w) || w >= (float)c.n_kinds)
      return fail(h, SOUP_E_INVALID, fmt("%s: particle %u has kind %g outside the %u monomer kinds", who, i, w, c.n_kinds));
    for (int k = 0; k < 3; ++k) {
      uint32_t p = st->bond_slots[3u * i + k];
      if (p != SOUP_NONE && (p >= n || p == i))
        return fail(h, SOUP_E_INVALID, fmt("%s: bond slot %d of particle %u names %u (N=%u)", who, k, i, p, n));
    }
    uint32_t l = st->center_link[i];
    if (l != SOUP_NONE && (l >= n || l == i))
      return fail(h, SOUP_E_INVALID, fmt("%s: centerLink of particle %u names %u (N=%u)", who, i, l, n));
    const float q = st->charges[i];
    if (q != 0.0f && q != (float)c.ph_q)
      return fail(h, SOUP_E_INVALID,
                  fmt("%s: charge of particle %u is %g; charges are 0 or chargeDeprotonated=%g", who, i, (double)q, c.ph_q));
  }
  return 0;
}

// ascending particle indices of the charged kind (kinds never change, so this is the compaction
// soup_build_head_index_main recomputes at every rebuild)
static std::vector<uint32_t> heads_of(const SoupConfig& c, uint32_t n, const float* positions) {
  std::vector<uint32_t> heads;
  if (c.es_enabled && c.n_heads > 0)
    for (uint32_t i = 0; i < n; ++i)
      if (positions[4u * i + 3] == (float)c.charged_kind) heads.push_back(i);
  return heads;
}

// Writes n particles' state into the current buffers with slot = particle index (the identity
// order the next rebuild sorts from); the force is cleared for the priming pass to fill.
static int upload_state(SoupHandle* h, uint32_t n, const SoupInitState* st) {
  DevBuf& b = h->b;
  std::vector<uint4> bs(n);
  for (uint32_t i = 0; i < n; ++i)
    bs[i] = make_uint4(st->bond_slots[3u * i], st->bond_slots[3u * i + 1], st->bond_slots[3u * i + 2], 0u);
  std::vector<uint32_t> ids(n);
  for (uint32_t i = 0; i < n; ++i) ids[i] = i;
  HC(h, hipMemcpy(b.pos, st->positions, (size_t)n * 16, hipMemcpyHostToDevice));
  HC(h, hipMemcpy(b.vel, st->velocities, (size_t)n * 16, hipMemcpyHostToDevice));
  HC(h, hipMemcpy(b.charge, st->charges, (size_t)n * 4, hipMemcpyHostToDevice));
  HC(h, hipMemcpy(b.bs, bs.data(), (size_t)n * 16, hipMemcpyHostToDevice));
  HC(h, hipMemcpy(b.clink, st->center_link, (size_t)n * 4, hipMemcpyHostToDevice));
  HC(h, hipMemcpy(b.cheld, st->center_held, (size_t)n * 4, hipMemcpyHostToDevice));
  HC(h, hipMemcpy(b.frozen, st->frozen, (size_t)n * 4, hipMemcpyHostToDevice));
  HC(h, hipMemcpy(b.brng, st->bond_rng, (size_t)n * 4, hipMemcpyHostToDevice));
  HC(h, hipMemcpy(b.trng, st->thermo_rng, (size_t)n * 4, hipMemcpyHostToDevice));
  HC(h, hipMemcpy(b.id, ids.data(), (size_t)n * 4, hipMemcpyHostToDevice));
  HC(h, hipMemcpy(b.inv, ids.data(), (size_t)n * 4, hipMemcpyHostToDevice));
  HC(h, hipMemset(b.force, 0, (size_t)h->cap * 16));
  HC(h, hipMemset(h->alt.force, 0, (size_t)h->cap * 16));
  HC(h, hipMemset(b.resv, 0xFF, (size_t)h->cap * 4));
  return 0;
}

// The device-limit guard of createSoup (soup/src/sim.ts:296-314): every buffer that scales with N is
// checked against what this GPU can still hold before anything is allocated, and the message names
// the byte count and the limit, so a configuration that cannot fit fails loudly at creation.
static int check_device_limit(const SoupConfig& c, uint32_t cap, uint32_t nh) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;  // no figure to compare with: hipMalloc decides
  const size_t list = (size_t)cap * vl_stride(c) * 4;
  const size_t es_rows = (size_t)nh * c.es_cap;
  struct Item {
    const char* what;
    size_t bytes;
  } items[] = {
      {"verletList (N*listCapacity*4)", list},
      {"pos/vel/force (N*16)", (size_t)cap * 16},
      {"esList (heads*longRangeListCapacity*4)", es_rows * 4},
  };
  // the list, its readback staging copy, the DH rows with their FP64 pair tables, and about 330 bytes
  // of other state per particle
  const size_t all = 2 * list + es_rows * 24 + (size_t)cap * 330;
  for (const Item& it : items)
    if (it.bytes > free_b || all > free_b)
      return fail(nullptr, SOUP_E_NOMEM,
                  fmt("buffer %s needs %zu bytes at N=%u (all device buffers: %zu bytes) -- device memory free=%zu of "
                      "total=%zu bytes on this GPU; the allocation would be refused, and a run without its buffers "
                      "would return zeros that look like success",
                      it.what, it.bytes, cap, all, free_b, total_b));
  return 0;
}

int soup_create(const SoupConfig* cfg, const SoupInitState* init, SoupHandle** out) {
  if (!out) return fail(nullptr, SOUP_E_INVALID, "soup_create: null output handle");
  *out = nullptr;
  int v = validate(cfg);
  if (v) return v;
  v = check_state(nullptr, *cfg, cfg->n, init, "soup_create");
  if (v) return v;
  v = check_device_limit(*cfg, std::max(cfg->n, cfg->capacity), (uint32_t)heads_of(*cfg, cfg->n, init->positions).size());
  if (v) return v;
  SoupHandle* h = new SoupHandle();
  h->c = *cfg;
  h->n = cfg->n;
  h->cap = std::max(cfg->n, cfg->capacity);
  h->c.capacity = h->cap;
  h->ncells = cfg->dims[0] * cfg->dims[1] * cfg->dims[2];
  // every per-particle buffer holds the capacity, so a smaller live count can grow back
  const uint32_t n = h->cap;
  h->heads = heads_of(*cfg, cfg->n, init->positions);
  h->nh = (uint32_t)h->heads.size();
  const std::vector<uint32_t>& heads = h->heads;
  int r = 0;
  auto cleanup = [&](int code) {
    std::string e = h->err;
    soup_destroy(h);
    g_create_error = e;
    return code;
  };
  if ((r = [&]() -> int {
        HC(h, hipGetDevice(&h->device));
        HC(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        DevBuf& b = h->b;
        DevBuf& a = h->alt;
        DA(h, b.pos, n); DA(h, a.pos, n);
        DA(h, b.vel, n); DA(h, a.vel, n);
        DA(h, b.force, n); DA(h, a.force, n);
        DA(h, b.pos_at_rebuild, n);
        DA(h, b.charge, n); DA(h, a.charge, n);
        DA(h, b.bs, n); DA(h, a.bs, n);
        DA(h, b.clink, n); DA(h, a.clink, n);
        DA(h, b.cheld, n); DA(h, a.cheld, n);
        DA(h, b.frozen, n); DA(h, a.frozen, n);
        DA(h, b.brng, n); DA(h, a.brng, n);
        DA(h, b.trng, n); DA(h, a.trng, n);
        DA(h, b.id, n); DA(h, a.id, n);
        DA(h, b.inv, n);
        DA(h, b.cell_of, n);
        DA(h, b.cell_count, h->ncells);
        DA(h, b.cell_start, (size_t)h->ncells + 1);
        h->ncells_cap = h->ncells;
        DA(h, b.cell_rank, n);
        DA(h, b.cell_tmp, n);
        DA(h, b.vl, (size_t)n * vl_stride(*cfg));
        DA(h, b.vcount, n);
        DA(h, b.vnear, n);
        DA(h, b.flags, 10);
        h->rowread_slots = xcd_blocks((uint64_t)n * FORCE_TPP, 256) * (256u / WAVE);
        DA(h, b.rowreads, 2 * (size_t)h->rowread_slots);
        DA(h, b.events, 2 * SOUP_MAX_RULES);
        DA(h, b.desorb, 2);
        DA(h, b.props, (size_t)n * SOUP_CAND_CAP);
        DA(h, b.nacc, n);
        DA(h, b.cat_id, n);
        DA(h, b.flag_cc, n);
        DA(h, b.ev_flag, n);
        DA(h, b.ev_pos, (size_t)n + 1);
        DA(h, b.ev_id, n);
        DA(h, b.resv, n);
        DA(h, b.rset, (size_t)n * SOUP_RSET);
        DA(h, b.ev_state, n);
        DA(h, b.scan_blocks, SCAN_TILE + 1);
        uint32_t nh = h->nh > 0 ? h->nh : 1;
        DA(h, b.head_id, nh);
        DA(h, b.head_slot, nh);
        DA(h, b.head_of_slot, n);
        DA(h, b.head_flag, n);
        DA(h, b.head_pos, (size_t)n + 1);
        DA(h, b.head_rows, nh);
        DA(h, b.head_q, nh);
        DA(h, b.far_acc, (size_t)nh * 3);
        DA(h, b.gq, n);
        if (h->nh > 0) {
          DA(h, b.es, (size_t)h->nh * cfg->es_cap);
          DA(h, h->ptab, (size_t)h->nh * cfg->es_cap);
          DA(h, h->ptab_g, (size_t)h->nh * cfg->es_cap);
          DA(h, h->pwcnt, (size_t)h->nh);
        } else {
          DA(h, b.es, 1);
        }
        DA(h, b.es_count, nh);
        DA(h, h->pstats, 8);
        DA(h, h->pdstats, 2);
        DA(h, h->lsum, 2);
        DA(h, h->lmax, 2);
        h->stage_bytes = std::max<size_t>((size_t)n * cfg->list_cap * 4, (size_t)n * 16);
        if (h->nh > 0) h->stage_bytes = std::max<size_t>(h->stage_bytes, (size_t)h->nh * cfg->es_cap * 4);
        h->stage_bytes = std::max<size_t>(h->stage_bytes, (size_t)n * sizeof(double4) + 64);  // soup_potential
        DA(h, h->stage, h->stage_bytes);
        // alternates share the arrays that are rewritten in place
        a.pos_at_rebuild = b.pos_at_rebuild;
        a.cell_of = b.cell_of;
        a.inv = b.inv;
        a.head_of_slot = b.head_of_slot;
        a.head_flag = b.head_flag;
        a.head_pos = b.head_pos;
        a.head_rows = b.head_rows;
        a.head_q = b.head_q;
        a.far_acc = b.far_acc;
        a.gq = b.gq;
        a.rowreads = b.rowreads;
        int u = upload_state(h, h->n, init);
        if (u) return u;
        HC(h, hipMemset(b.flags, 0, 40));
        HC(h, hipMemset(b.rowreads, 0, 8 * (size_t)h->rowread_slots));
        uint32_t ev[2 * SOUP_MAX_RULES] = {0};
        if (init->events)
          for (uint32_t k = 0; k < 2 * cfg->n_rules && k < 2 * SOUP_MAX_RULES; ++k) ev[k] = init->events[k];
        HC(h, hipMemcpy(b.events, ev, sizeof(ev), hipMemcpyHostToDevice));
        uint32_t de[2] = {0, 0};
        if (init->desorb_events) {
          de[0] = init->desorb_events[0];
          de[1] = init->desorb_events[1];
        }
        HC(h, hipMemcpy(b.desorb, de, 8, hipMemcpyHostToDevice));
        if (h->nh > 0) HC(h, hipMemcpy(b.head_id, heads.data(), (size_t)h->nh * 4, hipMemcpyHostToDevice));
        size_t lds = (size_t)h->nh * sizeof(float);
        if (lds > 64 * 1024) HC(h, hipFuncSetAttribute((const void*)k_proton_chain, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        h->global_step = init->global_step;
        h->next_sweep_at = init->next_sweep_at;
        h->prng = init->protonation_rng;
        {
          const int rg = es_grid_setup(h);
          if (rg) return rg;
        }
        // priming pass (soup/src/sim.ts:396-408)
        ph_rebuild(h);
        ph_force(h);
        return sync(h);
      }()) != 0)
    return cleanup(r);
  *out = h;
  return 0;
}

int soup_reset_particles(SoupHandle* h, uint32_t n, const SoupInitState* st) {
  if (!h) return SOUP_E_INVALID;
  if (n == 0 || n > h->cap)
    return fail(h, SOUP_E_INVALID,
                fmt("soup_reset_particles: %u particles requested, the buffers hold 1..%u (the creation census)", n, h->cap));
  int r = check_state(h, h->c, n, st, "soup_reset_particles");
  if (r) return r;
  if (heads_of(h->c, n, st->positions) != h->heads)
    return fail(h, SOUP_E_INVALID,
                fmt("soup_reset_particles: the charged heads must keep their particle indices (the long-range list and the "
                    "constant-pH chain are sized and indexed for the %u heads of the creation census)",
                    h->nh));
  r = sync(h);
  if (r) return r;
  drop_graphs(h);  // captured for the old particle count
  r = upload_state(h, n, st);
  if (r) return r;
  h->n = n;
  h->c.n = n;
  // setActiveCount's post-conditions (soup/src/soup-evaporate.ts:58-80): lists and force of the new
  // state, then the finiteness and list guards
  ph_rebuild(h);
  ph_force(h);
  r = sync(h);
  if (r) return r;
  r = assert_finite(h, false, 0);
  if (r) return r;
  return assert_lists(h);
}

int soup_use_graphs(SoupHandle* h, int enable) {
  if (!h) return SOUP_E_INVALID;
  int r = sync(h);
  if (r) return r;
  h->use_graphs = enable != 0;
  if (!h->use_graphs) drop_graphs(h);
  return 0;
}

int soup_profile(SoupHandle* h, int enable) {
  if (!h) return SOUP_E_INVALID;
  int r = sync(h);
  if (r) return r;
  h->prof.mask = enable < 0 ? (1u << SOUP_PROF_PHASES) - 1u : (uint32_t)enable;
  for (int k = 0; k < SOUP_PROF_PHASES; ++k) {
    h->prof.ms[k] = 0.0;
    h->prof.count[k] = 0;
  }
  return 0;
}

int soup_profile_read(SoupHandle* h, double* ms, uint64_t* count) {
  if (!h || !ms || !count) return SOUP_E_INVALID;
  int r = sync(h);
  if (r) return r;
  for (int k = 0; k < SOUP_PROF_PHASES; ++k) {
    ms[k] = h->prof.ms[k];
    count[k] = h->prof.count[k];
  }
  return 0;
}

int soup_destroy(SoupHandle* h) {
  if (!h) return 0;
  forget_batches_of(h);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  h->prof.pending.clear();
  for (hipEvent_t e : h->prof.pool) (void)hipEventDestroy(e);
  drop_graphs(h);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int soup_global_step(const SoupHandle* h, uint64_t* out) {
  if (!h || !out) return SOUP_E_INVALID;
  *out = h->global_step;
  return 0;
}

// the chunk-end work of soup_step: guards, then the due constant-pH sweep
static int chunk_end(SoupHandle* h, uint64_t chunk_start) {
  int r = assert_finite(h, true, chunk_start);
  if (r) return r;
  r = assert_lists(h);
  if (r) return r;
  if (h->c.es_enabled && h->global_step >= h->next_sweep_at) {
    h->next_sweep_at = h->global_step + h->c.sweep_every;
    r = sweep(h, nullptr);
    if (r) return r;
  }
  return 0;
}

// Drift mode: the measured drift maximum since the last rebuild (flags[2]) read back after a segment
// of steps; the next step rebuilds once it exceeds the trigger, which leaves a segment's worth of
// margin below skin/2 for the steps that run before the next check. A non-finite drift never
// triggers: the chunk-end guards report that state.
static int drift_check(SoupHandle* h) {
  uint32_t bits = 0;
  HC(h, hipMemcpyAsync(&bits, h->b.flags + 2, 4, hipMemcpyDeviceToHost, h->stream));
  int r = sync(h);
  if (r) return r;
  float d2;
  memcpy(&d2, &bits, 4);
  if (std::sqrt(std::fmax(0.0f, d2)) > h->c.drift_trigger) h->rebuild_due = true;
  return 0;
}

// `count` steps of one chunk. Cadence mode: whole step periods replayed from graphs where they fit.
// Drift mode: a rebuild step when one is due, else a segment of driftCheckEvery steps without a
// rebuild (replayed from a graph when whole), each segment followed by a drift check.
static int chunk_steps(SoupHandle* h, uint32_t count) {
  if (drift_mode(h)) {
    const uint32_t E = h->c.drift_check_every;
    for (uint32_t k = 0; k < count;) {
      if (h->rebuild_due) {
        one_step(h);
        ++k;
        continue;
      }
      const uint32_t seg = std::min(E, count - k);
      if (!(seg == E && replay_ok(h) && run_period(h, E)))
        for (uint32_t j = 0; j < seg; ++j) one_step(h);
      k += seg;
      const int r = drift_check(h);
      if (r) return r;
    }
    return 0;
  }
  const uint32_t P = replay_ok(h) ? step_period(h->c) : 0u;
  for (uint32_t k = 0; k < count;) {
    if (P && h->global_step % P == 0 && count - k >= P && run_period(h, P)) {
      k += P;
      continue;
    }
    one_step(h);
    ++k;
  }
  return 0;
}

int soup_step(SoupHandle* h, uint32_t nsteps) {
  if (!h) return SOUP_E_INVALID;
  uint32_t done = 0;
  while (done < nsteps) {
    uint32_t chunk = nsteps - done < 1000u ? nsteps - done : 1000u;
    uint64_t chunk_start = h->global_step;
    int r = chunk_steps(h, chunk);
    if (r) return r;
    r = chunk_end(h, chunk_start);
    if (r) return r;
    done += chunk;
  }
  return 0;
}

int soup_step_record(SoupHandle* h, uint32_t nsteps, float* pos, uint32_t* bonds, uint32_t* links, float* charges,
                     uint8_t* swept) {
  if (!h) return SOUP_E_INVALID;
  const size_t n = h->n;
  uint32_t done = 0, s = 0;
  while (done < nsteps) {
    const uint32_t chunk = nsteps - done < 1000u ? nsteps - done : 1000u;
    const uint64_t chunk_start = h->global_step;
    for (uint32_t k = 0; k < chunk; ++k, ++s) {
      one_step(h);
      int r = drift_mode(h) ? drift_check(h) : 0;
      if (pos) r = soup_read(h, SOUP_FIELD_POS, pos + (size_t)s * 4 * n, n * 16);
      if (!r && bonds) r = soup_read(h, SOUP_FIELD_BOND_SLOTS, bonds + (size_t)s * 3 * n, n * 12);
      if (!r && links) r = soup_read(h, SOUP_FIELD_CENTER_LINK, links + (size_t)s * n, n * 4);
      if (!r && charges) r = soup_read(h, SOUP_FIELD_CHARGE, charges + (size_t)s * n, n * 4);
      if (r) return r;
      if (swept) swept[s] = 0;
    }
    const uint32_t sweeps0 = h->sweeps;
    int r = chunk_end(h, chunk_start);
    if (r) return r;
    if (h->sweeps != sweeps0) {
      // the sweep after this chunk's last step changed the charges and recomputed the force
      if (swept) swept[s - 1] = 1;
      if (charges) {
        r = soup_read(h, SOUP_FIELD_CHARGE, charges + (size_t)(s - 1) * n, n * 4);
        if (r) return r;
      }
    }
    done += chunk;
  }
  return 0;
}

int soup_set_params(SoupHandle* h, const SoupConfig* c) {
  if (!h) return SOUP_E_INVALID;
  if (!c) return fail(h, SOUP_E_INVALID, "soup_set_params: null config");
  const SoupConfig& o = h->c;
  bool same = c->n == o.n && c->n_kinds == o.n_kinds && c->walk_radius == o.walk_radius && c->list_range == o.list_range &&
              c->list_cap == o.list_cap && c->rebuild_every == o.rebuild_every && c->skin == o.skin &&
              c->es_enabled == o.es_enabled && c->es_rc == o.es_rc && c->es_split == o.es_split &&
              c->es_list_range == o.es_list_range && c->es_cap == o.es_cap && c->charged_kind == o.charged_kind &&
              c->n_heads == o.n_heads && c->bond_interval == o.bond_interval && c->ph_q == o.ph_q &&
              std::max(c->n, c->capacity) == o.capacity;
  for (int a = 0; a < 3; ++a) same = same && c->box[a] == o.box[a] && c->dims[a] == o.dims[a];
  for (uint32_t k = 0; k < SOUP_MAX_KINDS; ++k) same = same && c->radius[k] == o.radius[k] && c->cls[k] == o.cls[k];
  same = same && c->sigma == o.sigma && c->wca_factor == o.wca_factor && c->b_tt == o.b_tt && c->wc == o.wc;
  // the integrator step and the bonded geometry (the skin basis of the lists rests on dt), the bond
  // rules' identities, roles and bookkeeping, the sweep cadence and the rebuild schedule
  same = same && c->dt == o.dt && c->k_fene == o.k_fene && c->r_inf == o.r_inf && c->k_bend == o.k_bend &&
         c->r_bend == o.r_bend && c->n_rules == o.n_rules && c->catalyst_kind == o.catalyst_kind &&
         c->head_terminal_only == o.head_terminal_only && c->head_chain_cap == o.head_chain_cap &&
         c->max_hold_steps == o.max_hold_steps && c->bond_interval_f == o.bond_interval_f && c->sweep_every == o.sweep_every &&
         c->rebuild_mode == o.rebuild_mode && c->drift_check_every == o.drift_check_every && c->drift_trigger == o.drift_trigger;
  for (uint32_t k = 0; k < SOUP_MAX_RULES; ++k)
    same = same && c->kindA[k] == o.kindA[k] && c->kindB[k] == o.kindB[k] && c->roleA[k] == o.roleA[k] &&
           c->roleB[k] == o.roleB[k] && c->req_cat[k] == o.req_cat[k];
  if (!same)
    return fail(h, SOUP_E_INVALID,
                "soup_set_params: only interaction strengths may change on a live system (epsilon, the attraction and acid-soap "
                "depths, the DH coefficient, kappa and its shift force, the constant-pH energies, the thermostat and the bond "
                "attempt and accept probabilities); geometry, dt, bonded constants, list ranges, radii, capacities, bond rule "
                "identities and roles, and the rebuild and sweep schedules are fixed at creation");
  int r = sync(h);
  if (r) return r;
  drop_graphs(h);  // their kernels carry the old parameters
  h->c = *c;
  // the force the next step starts from, under the new parameters (same lists: with unchanged
  // parameters this reproduces the current force bit for bit)
  ph_force(h);
  return sync(h);
}

}  // extern "C": the packed batch's C++ helpers follow

// ================================================================================================
// Packed replica steps: one launch per kernel covers every system of a batch (soup_packed.inl)
// ================================================================================================
// Launch shapes of the packed kernels; each has a table of block offsets per system.
enum PackKind {
  PK_P256 = 0,   // 256 threads per particle block
  PK_P256_ES,    // the same, systems with the long-range DH term only
  PK_WAVE,       // one wavefront per particle
  PK_FORCE,      // FORCE_TPP lanes per particle
  PK_H256,       // one thread per charged head
  PK_HWAVE,      // one wavefront per charged head
  PK_CELLS,      // one thread per grid cell
  PK_SCAN_C,     // scan tiles over the cells
  PK_SCAN_N,     // scan tiles over the particles
  PK_SCAN_N_ES,  // the same, systems with the DH term only
  PK_ONE,        // one workgroup per system
  PK_ONE_ES,     // one workgroup per system with the DH term
  PK_KINDS
};

// The device view of a batch: descriptor tables for the systems' current buffer assignment (tab[0])
// and for the swapped one a rebuild leaves (tab[1]), the offset tables, a stream of its own and the
// step periods captured on it. Built for one list of systems and their generations; any change to a
// system's configuration, particle count or buffers (its generation) rebuilds it.
struct BatchCtx {
  std::vector<SoupHandle*> hs;
  std::vector<uint64_t> gens;
  std::vector<float4*> keys;  // each system's b.pos the tables were written for
  int device = 0;
  hipStream_t stream = nullptr;
  RepDesc* tab[2] = {nullptr, nullptr};
  uint32_t* off = nullptr;  // PK_KINDS x (nrep + 1)
  uint32_t total[PK_KINDS] = {0};
  int cur = 0;
  bool any_es = false;
  std::vector<SoupHandle::StepGraph> graphs;
  ~BatchCtx() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (tab[0]) (void)hipFree(tab[0]);
    if (tab[1]) (void)hipFree(tab[1]);
    if (off) (void)hipFree(off);
    if (stream) (void)hipStreamDestroy(stream);
  }
  uint32_t nrep() const { return (uint32_t)hs.size(); }
  const uint32_t* offs(int kind) const { return off + (size_t)kind * (hs.size() + 1); }
};

static std::mutex g_batch_mu;
static std::vector<std::unique_ptr<BatchCtx>> g_batches;  // the most recently used last

// the 11 double-buffered arrays a rebuild permutes, exchanged between two DevBufs
static void swap_permuted(DevBuf& x, DevBuf& y) {
  std::swap(x.pos, y.pos);
  std::swap(x.vel, y.vel);
  std::swap(x.force, y.force);
  std::swap(x.charge, y.charge);
  std::swap(x.bs, y.bs);
  std::swap(x.clink, y.clink);
  std::swap(x.cheld, y.cheld);
  std::swap(x.frozen, y.frozen);
  std::swap(x.brng, y.brng);
  std::swap(x.trng, y.trng);
  std::swap(x.id, y.id);
}

// writes both descriptor tables for the systems' current buffer assignment
static hipError_t pack_tables(BatchCtx& B) {
  const size_t nr = B.hs.size();
  std::vector<RepDesc> t0(nr), t1(nr);
  for (size_t r = 0; r < nr; ++r) {
    SoupHandle* h = B.hs[r];
    RepDesc d{};
    d.c = h->c;
    d.b = h->b;
    d.alt = h->alt;
    d.nh = es_on(h) ? h->nh : 0u;
    d.es = es_on(h) ? 1u : 0u;
    t0[r] = d;
    swap_permuted(d.b, d.alt);
    t1[r] = d;
    B.keys[r] = h->b.pos;
  }
  hipError_t e = hipMemcpyAsync(B.tab[0], t0.data(), nr * sizeof(RepDesc), hipMemcpyHostToDevice, B.stream);
  if (e == hipSuccess) e = hipMemcpyAsync(B.tab[1], t1.data(), nr * sizeof(RepDesc), hipMemcpyHostToDevice, B.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(B.stream);
  B.cur = 0;
  return e;
}

// the batch of these systems, built (or rebuilt after a change) on demand; null with a message in
// *why when it cannot be built
static BatchCtx* batch_of(SoupHandle* const* hs, uint32_t count, std::string* why) {
  std::lock_guard<std::mutex> lock(g_batch_mu);
  for (size_t k = 0; k < g_batches.size(); ++k) {
    BatchCtx* B = g_batches[k].get();
    if (B->hs.size() != count || !std::equal(B->hs.begin(), B->hs.end(), hs)) continue;
    bool same = true;
    for (uint32_t r = 0; r < count && same; ++r) same = B->gens[r] == hs[r]->gen;
    if (same) {
      std::unique_ptr<BatchCtx> keep = std::move(g_batches[k]);
      g_batches.erase(g_batches.begin() + k);
      g_batches.push_back(std::move(keep));
      return g_batches.back().get();
    }
    g_batches.erase(g_batches.begin() + k);  // stale: built for an older state of some system
    break;
  }
  std::unique_ptr<BatchCtx> B(new BatchCtx());
  B->hs.assign(hs, hs + count);
  B->device = hs[0]->device;
  for (uint32_t r = 0; r < count; ++r) {
    B->gens.push_back(hs[r]->gen);
    B->any_es = B->any_es || es_on(hs[r]);
  }
  B->keys.resize(count);
  const size_t nr = count;
  std::vector<uint32_t> off((size_t)PK_KINDS * (nr + 1), 0u);
  for (size_t r = 0; r < nr; ++r) {
    const SoupHandle* h = hs[r];
    const uint64_t n = h->n, nh = es_on(h) ? h->nh : 0u, ncell = (uint64_t)h->c.dims[0] * h->c.dims[1] * h->c.dims[2];
    const bool es = es_on(h);
    const uint32_t tiles_n = std::max(1u, blocks(n, SCAN_TILE)), tiles_c = std::max(1u, blocks(ncell, SCAN_TILE));
    const uint32_t nb[PK_KINDS] = {blocks(n, 256),       es ? blocks(n, 256) : 0u, blocks(n * WAVE, 256),
                                   blocks(n * FORCE_TPP, 256), blocks(nh, 256),     blocks(nh * WAVE, 256),
                                   blocks(ncell, 256),   tiles_c,                   tiles_n,
                                   es ? tiles_n : 0u,    1u,                        es ? 1u : 0u};
    for (int k = 0; k < PK_KINDS; ++k) off[(size_t)k * (nr + 1) + r + 1] = off[(size_t)k * (nr + 1) + r] + nb[k];
  }
  for (int k = 0; k < PK_KINDS; ++k) B->total[k] = off[(size_t)k * (nr + 1) + nr];
  int cur_dev = 0;
  bool ok = hipGetDevice(&cur_dev) == hipSuccess && cur_dev == B->device &&
            hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc((void**)&B->tab[0], nr * sizeof(RepDesc)) == hipSuccess &&
            hipMalloc((void**)&B->tab[1], nr * sizeof(RepDesc)) == hipSuccess &&
            hipMalloc((void**)&B->off, off.size() * sizeof(uint32_t)) == hipSuccess &&
            hipMemcpy(B->off, off.data(), off.size() * sizeof(uint32_t), hipMemcpyHostToDevice) == hipSuccess &&
            pack_tables(*B) == hipSuccess;
  if (!ok) {
    *why = "soup_step_batch: the batch's device tables could not be set up on the systems' device";
    return nullptr;
  }
  if (g_batches.size() >= 4) g_batches.erase(g_batches.begin());
  g_batches.push_back(std::move(B));
  return g_batches.back().get();
}

// drops every batch that holds this system (it is being destroyed)
static void forget_batches_of(const SoupHandle* h) {
  std::lock_guard<std::mutex> lock(g_batch_mu);
  for (size_t k = g_batches.size(); k-- > 0;)
    if (std::find(g_batches[k]->hs.begin(), g_batches[k]->hs.end(), h) != g_batches[k]->hs.end())
      g_batches.erase(g_batches.begin() + k);
}

// The packed batch path is switched off: its first run on the GPU (five uneven replicas, 230 steps)
// ended in an illegal memory access whose cause is not yet found. soup_step_batch uses the per-system
// round robin below until it is.
static constexpr bool kPackedBatch = false;

// whether these systems can step as one packed batch: one device, the reference's rebuild cadence,
// the same rebuild and bond intervals and step phase, no profile running
static bool packable(SoupHandle* const* hs, uint32_t count) {
  if (!kPackedBatch) return false;
  const SoupHandle* a = hs[0];
  const uint32_t P = step_period(a->c);
  if (!P) return false;
  for (uint32_t r = 0; r < count; ++r) {
    const SoupHandle* h = hs[r];
    if (h->device != a->device || drift_mode(h) || h->prof.mask != 0 || h->c.rebuild_every != a->c.rebuild_every ||
        h->c.bond_interval != a->c.bond_interval || h->global_step % P != a->global_step % P)
      return false;
    for (uint32_t q = 0; q < r; ++q)
      if (hs[q] == h) return false;
  }
  return true;
}

template <typename K, typename... A>
static void pk_launch(BatchCtx& B, K kern, int kind, uint32_t threads, A... extra) {
  const uint32_t g = B.total[kind];
  if (g) hipLaunchKernelGGL(kern, dim3(g), dim3(threads), 0, B.stream, B.tab[B.cur], B.offs(kind), B.nrep(), extra...);
}

static void pk_scan(BatchCtx& B, uint32_t which, int tiles, int one) {
  pk_launch(B, kp_scan_reduce, tiles, SCAN_BLOCK, which);
  pk_launch(B, kp_scan_blocks, one, SCAN_BLOCK, which);
  pk_launch(B, kp_scan_apply, tiles, SCAN_BLOCK, which);
}

// ph_rebuild of every system in one set of launches
static void pk_rebuild(BatchCtx& B) {
  pk_launch(B, kp_clear_cells, PK_CELLS, 256);
  pk_launch(B, kp_cell_count, PK_P256, 256);
  pk_scan(B, 0u, PK_SCAN_C, PK_ONE);
  pk_launch(B, kp_cell_fill, PK_P256, 256);
  pk_launch(B, kp_cell_rank, PK_P256, 256);
  pk_launch(B, kp_permute, PK_P256, 256);
  B.cur ^= 1;  // every system now reads its permuted arrays
  for (SoupHandle* h : B.hs) swap_permuted(h->b, h->alt);
  if (B.any_es) {
    pk_launch(B, kp_clear_heads, PK_P256_ES, 256);
    pk_launch(B, kp_head_slots, PK_H256, 256);
    pk_scan(B, 1u, PK_SCAN_N_ES, PK_ONE_ES);
    pk_launch(B, kp_head_rows, PK_P256_ES, 256);
    pk_launch(B, kp_es_build, PK_HWAVE, 256);
  }
  pk_launch(B, kp_verlet_build, PK_WAVE, 256);
  hipLaunchKernelGGL(kp_clear_drift, dim3(blocks(B.nrep(), 64)), dim3(64), 0, B.stream, B.tab[B.cur], B.nrep());
}

// one_step of every system: all are at the same step phase, so the rebuild and bond decisions agree
static void pk_step(BatchCtx& B, bool thermo_in, bool thermo_out) {
  const SoupHandle* a = B.hs[0];
  const bool do_bonds = (a->global_step % a->c.bond_interval) == 0;
  const bool do_rebuild = (a->global_step % a->c.rebuild_every) == 0;
  pk_launch(B, thermo_in ? kp_thermo_drift : kp_kick_drift, PK_P256, 256);
  if (do_rebuild) pk_rebuild(B);
  if (B.any_es) {
    pk_launch(B, kp_head_refresh, PK_H256, 256);
    pk_launch(B, kp_far_dh, PK_HWAVE, 256);
  }
  pk_launch(B, kp_force_group, PK_FORCE, 256);
  if (do_bonds) {
    pk_launch(B, kp_clear_bond_flags, PK_P256, 256);
    pk_launch(B, kp_bond_propose, PK_WAVE, 256);
    pk_launch(B, kp_bond_filter, PK_P256, 256);
    pk_scan(B, 2u, PK_SCAN_N, PK_ONE);
    pk_launch(B, kp_bond_events, PK_P256, 256);
    pk_launch(B, kp_bond_resolve, PK_ONE, 1024);
    pk_launch(B, kp_bond_break, PK_P256, 256);
  }
  if (!thermo_out) pk_launch(B, kp_kick_thermo, PK_P256, 256);
  for (SoupHandle* h : B.hs) {
    if (do_rebuild) h->rebuilds++;
    h->global_step++;
  }
}

// one period of P steps of the whole batch from a graph captured on the batch's stream (its
// buffer assignment returns to where it began); false when capture is unavailable
static bool pk_period(BatchCtx& B, uint32_t P) {
  const uint32_t phase = (uint32_t)(B.hs[0]->global_step % B.hs[0]->c.bond_interval);
  for (auto& g : B.graphs)
    if (g.exec && g.len == P && g.phase == phase) {
      if (hipGraphLaunch(g.exec, B.stream) != hipSuccess) return false;
      for (SoupHandle* h : B.hs) {
        h->global_step += P;
        h->rebuilds += P / h->c.rebuild_every;
      }
      return true;
    }
  std::vector<uint64_t> step0, reb0;
  std::vector<DevBuf> b0, a0;
  for (SoupHandle* h : B.hs) {
    step0.push_back(h->global_step);
    reb0.push_back(h->rebuilds);
    b0.push_back(h->b);
    a0.push_back(h->alt);
  }
  const int cur0 = B.cur;
  if (hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
  for (uint32_t k = 0; k < P; ++k) pk_step(B, k > 0, k + 1 < P);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(B.stream, &graph);
  hipGraphExec_t exec = nullptr;
  const bool ok = ec == hipSuccess && graph && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (graph) (void)hipGraphDestroy(graph);
  for (size_t r = 0; r < B.hs.size(); ++r) {
    B.hs[r]->global_step = step0[r];
    B.hs[r]->rebuilds = reb0[r];
    B.hs[r]->b = b0[r];
    B.hs[r]->alt = a0[r];
  }
  B.cur = cur0;
  (void)hipGetLastError();
  if (!ok) {
    if (exec) (void)hipGraphExecDestroy(exec);
    return false;
  }
  SoupHandle::StepGraph g;
  g.exec = exec;
  g.len = P;
  g.phase = phase;
  B.graphs.push_back(g);
  if (hipGraphLaunch(exec, B.stream) != hipSuccess) return false;
  for (SoupHandle* h : B.hs) {
    h->global_step += P;
    h->rebuilds += P / h->c.rebuild_every;
  }
  return true;
}

// soup_step_batch over a packed batch: each chunk's steps as packed launches (whole periods from
// graphs when every system replays graphs), then every system's chunk-end guards and due sweep
static int step_packed(BatchCtx& B, uint32_t nsteps, uint32_t* failed) {
  const uint32_t count = B.nrep();
  for (uint32_t r = 0; r < count; ++r) {  // the systems' own streams are idle before the batch writes
    const int e = sync(B.hs[r]);
    if (e) {
      *failed = r;
      return e;
    }
  }
  bool keys_ok = true;
  for (uint32_t r = 0; r < count; ++r) keys_ok = keys_ok && B.keys[r] == B.hs[r]->b.pos;
  if (!keys_ok && pack_tables(B) != hipSuccess) {
    *failed = 0;
    return fail(B.hs[0], SOUP_E_DEVICE, "soup_step_batch: the batch's descriptor tables could not be written");
  }
  bool graphs = true;
  for (uint32_t r = 0; r < count; ++r) graphs = graphs && B.hs[r]->use_graphs;
  const uint32_t P = step_period(B.hs[0]->c);
  std::vector<uint64_t> start(count);
  uint32_t done = 0;
  while (done < nsteps) {
    const uint32_t chunk = nsteps - done < 1000u ? nsteps - done : 1000u;
    for (uint32_t r = 0; r < count; ++r) start[r] = B.hs[r]->global_step;
    for (uint32_t k = 0; k < chunk;) {
      if (graphs && B.hs[0]->global_step % P == 0 && chunk - k >= P && pk_period(B, P)) {
        k += P;
        continue;
      }
      pk_step(B, false, false);
      ++k;
    }
    const hipError_t e = hipStreamSynchronize(B.stream);
    if (e != hipSuccess) {
      *failed = 0;
      return fail(B.hs[0], SOUP_E_DEVICE, std::string("soup_step_batch: packed steps: ") + hipGetErrorString(e));
    }
    for (uint32_t r = 0; r < count; ++r) {
      const int rc = chunk_end(B.hs[r], start[r]);
      if (rc) {
        *failed = r;
        return rc;
      }
    }
    done += chunk;
    // a sweep rebuilt some systems on their own streams, which flips their buffer assignment
    keys_ok = true;
    for (uint32_t r = 0; r < count; ++r) keys_ok = keys_ok && B.keys[r] == B.hs[r]->b.pos;
    if (!keys_ok && pack_tables(B) != hipSuccess) {
      *failed = 0;
      return fail(B.hs[0], SOUP_E_DEVICE, "soup_step_batch: the batch's descriptor tables could not be written");
    }
  }
  return 0;
}

extern "C" {

int soup_step_batch(SoupHandle* const* hs, uint32_t count, uint32_t nsteps, uint32_t* failed) {
  if (!hs || (count && !failed)) return SOUP_E_INVALID;
  for (uint32_t s = 0; s < count; ++s)
    if (!hs[s]) return SOUP_E_INVALID;
  if (count && nsteps && packable(hs, count)) {
    std::string why;
    BatchCtx* B = batch_of(hs, count, &why);
    if (B) return step_packed(*B, nsteps, failed);
    *failed = 0;
    return fail(hs[0], SOUP_E_DEVICE, why);
  }
  std::vector<uint64_t> start(count);
  std::vector<uint32_t> left(count), period(count);
  uint32_t done = 0;
  while (done < nsteps) {
    const uint32_t chunk = nsteps - done < 1000u ? nsteps - done : 1000u;
    for (uint32_t s = 0; s < count; ++s) {
      SoupHandle* h = hs[s];
      start[s] = h->global_step;
      left[s] = chunk;
      period[s] = replay_ok(h) ? step_period(h->c) : 0u;
    }
    // round robin: each system advances to its next period boundary or by one period per pass, so
    // the streams of all systems have work queued together
    for (bool busy = true; busy;) {
      busy = false;
      for (uint32_t s = 0; s < count; ++s) {
        SoupHandle* h = hs[s];
        const uint32_t P = period[s];
        if (left[s] == 0) continue;
        busy = true;
        if (drift_mode(h)) {  // its drift checks synchronise its stream: stepped through on its own
          const int r = chunk_steps(h, left[s]);
          left[s] = 0;
          if (r) {
            *failed = s;
            return r;
          }
          continue;
        }
        if (P && h->global_step % P == 0 && left[s] >= P && run_period(h, P)) {
          left[s] -= P;
          continue;
        }
        do {
          one_step(h);
          --left[s];
        } while (left[s] && !(P && h->global_step % P == 0));
      }
    }
    for (uint32_t s = 0; s < count; ++s) {
      const int r = chunk_end(hs[s], start[s]);
      if (r) {
        *failed = s;
        return r;
      }
    }
    done += chunk;
  }
  return 0;
}

size_t soup_field_bytes(const SoupHandle* h, int field) {
  if (!h) return 0;
  const size_t n = h->n;
  switch (field) {
    case SOUP_FIELD_POS:
    case SOUP_FIELD_VEL:
    case SOUP_FIELD_FORCE: return n * 16;
    case SOUP_FIELD_BOND_SLOTS: return n * 12;
    case SOUP_FIELD_CENTER_LINK:
    case SOUP_FIELD_CENTER_HELD:
    case SOUP_FIELD_FROZEN:
    case SOUP_FIELD_CHARGE:
    case SOUP_FIELD_BOND_RNG:
    case SOUP_FIELD_THERMO_RNG:
    case SOUP_FIELD_VERLET_COUNT:
    case SOUP_FIELD_VERLET_NEAR:
    case SOUP_FIELD_CELL_OF:
    case SOUP_FIELD_CELL_IDX: return n * 4;
    case SOUP_FIELD_EVENTS: return (size_t)h->c.n_rules * 2 * 4;
    case SOUP_FIELD_DESORB: return 8;
    case SOUP_FIELD_VERLET_LIST: return n * h->c.list_cap * 4;
    case SOUP_FIELD_ES_COUNT: return (size_t)h->nh * 4;
    case SOUP_FIELD_ES_LIST: return (size_t)h->nh * h->c.es_cap * 4;
    case SOUP_FIELD_CELL_START: return ((size_t)h->ncells + 1) * 4;
    case SOUP_FIELD_MAX_DRIFT_SQ: return 4;
    case SOUP_FIELD_LIST_READS: return 32;
    case SOUP_FIELD_REBUILDS: return 8;
    case SOUP_FIELD_HEAD_IDX: return (size_t)h->nh * 4;
    default: return 0;
  }
}

int soup_read(SoupHandle* h, int field, void* dst, size_t bytes) {
  if (!h || !dst) return SOUP_E_INVALID;
  size_t need = soup_field_bytes(h, field);
  if (need == 0 && field != SOUP_FIELD_ES_COUNT && field != SOUP_FIELD_ES_LIST && field != SOUP_FIELD_HEAD_IDX)
    return fail(h, SOUP_E_INVALID, fmt("soup_read: unknown field %d", field));
  if (bytes != need) return fail(h, SOUP_E_INVALID, fmt("soup_read: field %d needs %zu bytes, got %zu", field, need, bytes));
  if (need == 0) return 0;
  if (field == SOUP_FIELD_REBUILDS) {  // a host counter
    memcpy(dst, &h->rebuilds, 8);
    return 0;
  }
  const void* src = h->stage;
  const uint32_t n = h->n;
  switch (field) {
    case SOUP_FIELD_POS: to_index(h, h->b.pos); break;
    case SOUP_FIELD_VEL: to_index(h, h->b.vel); break;
    case SOUP_FIELD_FORCE: to_index(h, h->b.force); break;
    case SOUP_FIELD_CHARGE: to_index(h, h->b.charge); break;
    case SOUP_FIELD_CENTER_LINK: to_index(h, h->b.clink); break;
    case SOUP_FIELD_CENTER_HELD: to_index(h, h->b.cheld); break;
    case SOUP_FIELD_FROZEN: to_index(h, h->b.frozen); break;
    case SOUP_FIELD_BOND_RNG: to_index(h, h->b.brng); break;
    case SOUP_FIELD_THERMO_RNG: to_index(h, h->b.trng); break;
    case SOUP_FIELD_CELL_OF: to_index(h, h->b.cell_of); break;
    case SOUP_FIELD_VERLET_COUNT: to_index(h, h->b.vcount); break;
    case SOUP_FIELD_VERLET_NEAR: to_index(h, h->b.vnear); break;
    case SOUP_FIELD_BOND_SLOTS:
      hipLaunchKernelGGL(k_bs_to_index, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->b.bs, h->b.id,
                         reinterpret_cast<uint32_t*>(h->stage), n);
      break;
    case SOUP_FIELD_VERLET_LIST:
      if ((size_t)h->c.list_cap * 8u > 160u * 1024u)
        return fail(h, SOUP_E_INVALID, fmt("Verlet list readback: listCapacity=%u exceeds the 20480 entries one "
                                       "workgroup's LDS can rank", h->c.list_cap));
      if ((size_t)h->c.list_cap * 8u > 64u * 1024u)
        HC(h, hipFuncSetAttribute((const void*)k_vl_to_index, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)((size_t)h->c.list_cap * 8u)));
      hipLaunchKernelGGL(k_vl_to_index, dim3(n), dim3(WAVE), (size_t)h->c.list_cap * 8u, h->stream, h->c, h->b,
                         reinterpret_cast<uint32_t*>(h->stage));
      break;
    case SOUP_FIELD_ES_LIST:
      hipLaunchKernelGGL(k_es_to_index, dim3(blocks(h->nh, 256)), dim3(256), 0, h->stream, h->b.es, h->b.es_count,
                         h->b.head_id, reinterpret_cast<uint32_t*>(h->stage), h->nh, h->c.es_cap);
      break;
    case SOUP_FIELD_ES_COUNT: src = h->b.es_count; break;
    case SOUP_FIELD_HEAD_IDX: src = h->b.head_id; break;
    case SOUP_FIELD_CELL_IDX:
      hipLaunchKernelGGL(k_cell_idx_to_index, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->c, h->b,
                         reinterpret_cast<uint32_t*>(h->stage));
      break;
    case SOUP_FIELD_CELL_START: src = h->b.cell_start; break;
    case SOUP_FIELD_EVENTS: src = h->b.events; break;
    case SOUP_FIELD_DESORB: src = h->b.desorb; break;
    case SOUP_FIELD_MAX_DRIFT_SQ: src = h->b.flags + 2; break;
    case SOUP_FIELD_LIST_READS:
      hipLaunchKernelGGL(k_sum_rowreads, dim3(1), dim3(1024), 0, h->stream, h->b.rowreads, h->rowread_slots, h->b.flags,
                         reinterpret_cast<unsigned long long*>(h->stage));
      src = h->stage;
      break;
    default: return fail(h, SOUP_E_INVALID, fmt("soup_read: unknown field %d", field));
  }
  HC(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
  return sync(h);
}

int soup_write(SoupHandle* h, int field, const void* srcp, size_t bytes) {
  if (!h || !srcp) return SOUP_E_INVALID;
  size_t need = soup_field_bytes(h, field);
  if (bytes != need || need == 0) return fail(h, SOUP_E_INVALID, fmt("soup_write: field %d needs %zu bytes, got %zu", field, need, bytes));
  const uint32_t n = h->n;
  HC(h, hipMemcpyAsync(h->stage, srcp, bytes, hipMemcpyHostToDevice, h->stream));
  dim3 g(blocks(n, 256)), bl(256);
  switch (field) {
    case SOUP_FIELD_POS:
      hipLaunchKernelGGL(k_from_index<float4>, g, bl, 0, h->stream, (const float4*)h->stage, h->b.id, h->b.pos, n);
      // the list force reads the far part of a row only when the drift since the rebuild asks for it
      hipLaunchKernelGGL(k_max_drift, g, bl, 0, h->stream, h->c, h->b);
      hipLaunchKernelGGL(k_gq_refresh, g, bl, 0, h->stream, h->c, h->b);
      break;
    case SOUP_FIELD_VEL: hipLaunchKernelGGL(k_from_index<float4>, g, bl, 0, h->stream, (const float4*)h->stage, h->b.id, h->b.vel, n); break;
    case SOUP_FIELD_CHARGE: {
      // the list force's gather array carries a charge as a flag: every charge is 0 or chargeDeprotonated
      const float* q = static_cast<const float*>(srcp);
      for (uint32_t i = 0; i < n; ++i)
        if (q[i] != 0.0f && q[i] != (float)h->c.ph_q)
          return fail(h, SOUP_E_INVALID, fmt("soup_write: charge of particle %u is %g; charges are 0 or chargeDeprotonated=%g", i,
                                             (double)q[i], h->c.ph_q));
      hipLaunchKernelGGL(k_from_index<float>, g, bl, 0, h->stream, (const float*)h->stage, h->b.id, h->b.charge, n);
      hipLaunchKernelGGL(k_gq_refresh, g, bl, 0, h->stream, h->c, h->b);
      break;
    }
    case SOUP_FIELD_CENTER_LINK: hipLaunchKernelGGL(k_from_index<uint32_t>, g, bl, 0, h->stream, (const uint32_t*)h->stage, h->b.id, h->b.clink, n); break;
    case SOUP_FIELD_CENTER_HELD: hipLaunchKernelGGL(k_from_index<uint32_t>, g, bl, 0, h->stream, (const uint32_t*)h->stage, h->b.id, h->b.cheld, n); break;
    case SOUP_FIELD_BOND_RNG: hipLaunchKernelGGL(k_from_index<uint32_t>, g, bl, 0, h->stream, (const uint32_t*)h->stage, h->b.id, h->b.brng, n); break;
    case SOUP_FIELD_THERMO_RNG: hipLaunchKernelGGL(k_from_index<uint32_t>, g, bl, 0, h->stream, (const uint32_t*)h->stage, h->b.id, h->b.trng, n); break;
    case SOUP_FIELD_BOND_SLOTS: hipLaunchKernelGGL(k_bs_from_index, g, bl, 0, h->stream, (const uint32_t*)h->stage, h->b.id, h->b.bs, n); break;
    case SOUP_FIELD_DESORB: HC(h, hipMemcpyAsync(h->b.desorb, h->stage, 8, hipMemcpyDeviceToDevice, h->stream)); break;
    default: return fail(h, SOUP_E_INVALID, fmt("soup_write: field %d is not writable", field));
  }
  return sync(h);
}

int soup_forces(SoupHandle* h, int mode) {
  if (!h) return SOUP_E_INVALID;
  switch (mode) {
    case SOUP_FORCES_REBUILD: ph_rebuild(h); ph_force(h); break;
    case SOUP_FORCES_BRUTE: ph_force_brute(h); break;
    case SOUP_FORCES_LIST: ph_force(h); break;
    default: return fail(h, SOUP_E_INVALID, fmt("soup_forces: unknown mode %d", mode));
  }
  return sync(h);
}

int soup_run_phases(SoupHandle* h, uint32_t mask, uint32_t do_bonds_flag) {
  if (!h) return SOUP_E_INVALID;
  (void)do_bonds_flag;
  if (mask & SOUP_PHASE_KICK_DRIFT) ph_kick_drift(h);
  if (mask & SOUP_PHASE_REBUILD) ph_rebuild(h);
  if (mask & SOUP_PHASE_FORCE) ph_force(h);
  if (mask & SOUP_PHASE_DRIFT) ph_drift(h);
  if (mask & SOUP_PHASE_BONDS) ph_bonds(h);
  if (mask & SOUP_PHASE_KICK_THERMO) ph_kick_thermo(h);
  return sync(h);
}

int soup_protonation_sweep(SoupHandle* h, SoupSweepResult* out) {
  if (!h) return SOUP_E_INVALID;
  if (!h->c.es_enabled) {
    if (out) memset(out, 0, sizeof(*out));
    return 0;
  }
  h->next_sweep_at = h->global_step + h->c.sweep_every;  // maybeProtonationSweep (soup-protonation.ts:54)
  if (!es_on(h)) {
    if (out) memset(out, 0, sizeof(*out));
    return 0;
  }
  return sweep(h, out);
}

int soup_last_sweep(const SoupHandle* h, uint32_t* sweeps, SoupSweepResult* last) {
  if (!h) return SOUP_E_INVALID;
  if (sweeps) *sweeps = h->sweeps;
  if (last) *last = h->last_sweep;
  return 0;
}

int soup_protonation_rng(const SoupHandle* h, uint32_t* out) {
  if (!h || !out) return SOUP_E_INVALID;
  *out = h->prng;
  return 0;
}

int soup_nonfinite(SoupHandle* h, uint32_t out[2]) {
  if (!h || !out) return SOUP_E_INVALID;
  return nonfinite_counts(h, out);
}

int soup_check_lists(SoupHandle* h) {
  if (!h) return SOUP_E_INVALID;
  return assert_lists(h);
}

int soup_list_stats(SoupHandle* h, SoupListStats* out) {
  if (!h || !out) return SOUP_E_INVALID;
  HC(h, hipMemsetAsync(h->lsum, 0, 16, h->stream));
  HC(h, hipMemsetAsync(h->lmax, 0, 8, h->stream));
  hipLaunchKernelGGL(k_list_stats, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->b.vcount, h->n, h->lsum, h->lmax);
  if (es_on(h))
    hipLaunchKernelGGL(k_list_stats, dim3(blocks(h->nh, 256)), dim3(256), 0, h->stream, h->b.es_count, h->nh, h->lsum + 1,
                       h->lmax + 1);
  unsigned long long s[2];
  uint32_t m[2];
  HC(h, hipMemcpyAsync(s, h->lsum, 16, hipMemcpyDeviceToHost, h->stream));
  HC(h, hipMemcpyAsync(m, h->lmax, 8, hipMemcpyDeviceToHost, h->stream));
  int r = sync(h);
  if (r) return r;
  out->verlet_pairs = s[0];
  out->verlet_mean = h->n ? (double)s[0] / h->n : 0.0;
  out->verlet_max = m[0];
  out->verlet_cap = h->c.list_cap;
  out->es_pairs = es_on(h) ? s[1] : 0;
  out->es_mean = es_on(h) && h->nh ? (double)s[1] / h->nh : 0.0;
  out->es_max = es_on(h) ? m[1] : 0;
  out->es_cap = h->c.es_cap;
  return 0;
}

int soup_potential(SoupHandle* h, const SoupPotentialBasis* pb, double out[5]) {
  if (!h) return SOUP_E_INVALID;
  if (!pb || !out) return fail(h, SOUP_E_INVALID, "soup_potential: null basis or output");
  // the lists hold every pair within listRange - skin of the current positions while the drift
  // guard holds; the potential's cutoffs must lie inside that
  const double near_ok = (double)h->c.list_range - (double)h->c.skin;
  if (pb->cutoff > near_ok + 1e-9)
    return fail(h, SOUP_E_INVALID, fmt("soup_potential: pair cutoff %.6f exceeds the Verlet list's guaranteed range %.6f",
                                       pb->cutoff, near_ok));
  const bool es = es_on(h) && pb->use_charges && pb->es_A != 0.0;
  if (es && pb->es_cutoff > (double)h->c.es_list_range - (double)h->c.skin + 1e-6)
    return fail(h, SOUP_E_INVALID, fmt("soup_potential: DH cutoff %.6f exceeds the DH list's guaranteed range %.6f",
                                       pb->es_cutoff, (double)h->c.es_list_range - (double)h->c.skin));
  int r = sync(h);
  if (r) return r;
  r = assert_lists(h);
  if (r) return r;
  const size_t need = (size_t)h->n * sizeof(double4) + 64;
  if (need > h->stage_bytes) return fail(h, SOUP_E_INVALID, "soup_potential: staging buffer too small");
  double4* part = reinterpret_cast<double4*>(h->stage);
  double* dout = reinterpret_cast<double*>(h->stage + (size_t)h->n * sizeof(double4));
  if (es) hipLaunchKernelGGL(k_head_refresh, dim3(blocks(h->nh, 256)), dim3(256), 0, h->stream, h->b, h->nh);
  hipLaunchKernelGGL(k_potential, dim3(blocks((uint64_t)h->n * WAVE, 256)), dim3(256), 0, h->stream, h->c, h->b, *pb, part,
                     es ? 1 : 0);
  hipLaunchKernelGGL(k_potential_sum, dim3(1), dim3(1024), 0, h->stream, part, h->n, dout);
  HC(h, hipMemcpyAsync(out, dout, 5 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  return sync(h);
}

// the argument and list checks of an adjoint launch on h's current lists; `who` names the entry point
static int adjoint_ready(SoupHandle* h, const char* who, const SoupPotentialBasis* pb, bool& es) {
  const double near_ok = (double)h->c.list_range - (double)h->c.skin;
  if (pb->cutoff > near_ok + 1e-9)
    return fail(h, SOUP_E_INVALID, fmt("%s: pair cutoff %.6f exceeds the Verlet list's guaranteed range %.6f", who, pb->cutoff,
                                       near_ok));
  es = es_on(h) && pb->use_charges && pb->es_A != 0.0;
  if (es && pb->es_cutoff > (double)h->c.es_list_range - (double)h->c.skin + 1e-6)
    return fail(h, SOUP_E_INVALID, fmt("%s: DH cutoff %.6f exceeds the DH list's guaranteed range %.6f", who, pb->es_cutoff,
                                       (double)h->c.es_list_range - (double)h->c.skin));
  if (es && h->c.ph_q == 0.0) return fail(h, SOUP_E_INVALID, fmt("%s: chargeDeprotonated is 0", who));
  if (!h->adj_g) {  // sized for the capacity: the live count may grow back after an evaporation
    DA(h, h->adj_w, (size_t)h->cap * 3);
    DA(h, h->adj_g, (size_t)h->cap * 3);
    DA(h, h->adj_hw, (size_t)h->cap * 3);
    DA(h, h->adj_part, (size_t)h->cap * ADJ_NPART);
    DA(h, h->adj_out, ADJ_NPART);
  }
  return 0;
}

// one adjoint launch on h's current state and lists (w: device, particle order, or null): gradient
// into adj_g, the (w + strain) Hessian product into adj_hw, the reduced partials into adj_out
static void adjoint_launch(SoupHandle* h, const SoupPotentialBasis* pb, bool es, const double* w_dev, const AdjStrain& st) {
  if (es) hipLaunchKernelGGL(k_head_refresh, dim3(blocks(h->nh, 256)), dim3(256), 0, h->stream, h->b, h->nh);
  hipLaunchKernelGGL(k_energy_adjoint, dim3(blocks((uint64_t)h->n * WAVE, 256)), dim3(256), 0, h->stream, h->c, h->b, *pb, es ? 1 : 0,
                     w_dev, st, h->adj_g, h->adj_hw, h->adj_part);
  hipLaunchKernelGGL(k_adjoint_sum, dim3(1), dim3(1024), 0, h->stream, h->adj_part, h->n, h->adj_out);
}

static int energy_adjoint(SoupHandle* h, const char* who, const SoupPotentialBasis* pb, const double* strain, const double* w,
                          double* grad, double* hw, double* dU_dtheta, double* dFw_dtheta, double* W_e) {
  if (!h) return SOUP_E_INVALID;
  if (!pb || !grad) return fail(h, SOUP_E_INVALID, fmt("%s: null basis or gradient output", who));
  if ((hw || dFw_dtheta) && !w && !strain) return fail(h, SOUP_E_INVALID, fmt("%s: the w-products need a direction w", who));
  bool es = false;
  int r = adjoint_ready(h, who, pb, es);
  if (r) return r;
  r = sync(h);
  if (r) return r;
  r = assert_lists(h);
  if (r) return r;
  const uint32_t n = h->n;
  AdjStrain st{{0.0, 0.0, 0.0}};
  if (strain)
    for (int a = 0; a < 3; ++a) st.e[a] = strain[a];
  if (w) HC(h, hipMemcpyAsync(h->adj_w, w, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice, h->stream));
  adjoint_launch(h, pb, es, w ? (const double*)h->adj_w : nullptr, st);
  double sums[ADJ_NPART];
  HC(h, hipMemcpyAsync(grad, h->adj_g, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  if (hw) HC(h, hipMemcpyAsync(hw, h->adj_hw, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HC(h, hipMemcpyAsync(sums, h->adj_out, sizeof(sums), hipMemcpyDeviceToHost, h->stream));
  r = sync(h);
  if (r) return r;
  for (uint32_t p = 0; p < ADJ_P; ++p) {
    if (dU_dtheta) dU_dtheta[p] = sums[p];
    if (dFw_dtheta) dFw_dtheta[p] = sums[ADJ_P + p];
  }
  if (W_e) *W_e = sums[2 * ADJ_P];
  return 0;
}

int soup_energy_adjoint(SoupHandle* h, const SoupPotentialBasis* pb, const double* w, double* grad, double* hw,
                        double* dU_dtheta, double* dFw_dtheta) {
  return energy_adjoint(h, "soup_energy_adjoint", pb, nullptr, w, grad, hw, dU_dtheta, dFw_dtheta, nullptr);
}

int soup_energy_adjoint_strain(SoupHandle* h, const SoupPotentialBasis* pb, const double strain[3], const double* w,
                               double* grad, double* hw, double* dU_dtheta, double* dFw_dtheta, double* W_e) {
  if (h && !strain) return fail(h, SOUP_E_INVALID, "soup_energy_adjoint_strain: null strain");
  return energy_adjoint(h, "soup_energy_adjoint_strain", pb, strain, w, grad, hw, dU_dtheta, dFw_dtheta, W_e);
}

}  // extern "C": the segment adjoint's C++ helpers follow

// ---- trajectory-segment adjoint with device checkpoints ---------------------------------------------
// Device buffers of one call, freed when it returns.
struct DevScratch {
  std::vector<void*> ptrs;
  template <typename T>
  hipError_t get(T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back(q);
    *p = static_cast<T*>(q);
    return e;
  }
  ~DevScratch() {
    for (void* q : ptrs) (void)hipFree(q);
  }
};
// a full checkpoint (particle order) and the host counters the steps after it depend on
struct SegCheckpoint {
  StateRefs st{};
  uint32_t* events = nullptr;
  uint32_t* desorb = nullptr;
  uint64_t global_step = 0, next_sweep_at = 0;
  uint32_t prng = 0, sweeps = 0;
  uint64_t rebuilds = 0;
};
static StateRefs state_part(const StateRefs& s, bool topo_only) {
  StateRefs r{};
  r.pos = s.pos;
  r.charge = s.charge;
  r.bs = s.bs;
  r.clink = s.clink;
  if (!topo_only) {
    r.vel = s.vel; r.force = s.force; r.cheld = s.cheld; r.frozen = s.frozen; r.brng = s.brng; r.trng = s.trng;
  }
  return r;
}
static void seg_save(SoupHandle* h, SegCheckpoint& ck) {
  hipLaunchKernelGGL(k_state_gather, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->n, h->b, ck.st);
  const size_t ev = 2u * (size_t)h->c.n_rules;
  if (ev) (void)hipMemcpyAsync(ck.events, h->b.events, ev * 4, hipMemcpyDeviceToDevice, h->stream);
  (void)hipMemcpyAsync(ck.desorb, h->b.desorb, 8, hipMemcpyDeviceToDevice, h->stream);
  ck.global_step = h->global_step;
  ck.next_sweep_at = h->next_sweep_at;
  ck.prng = h->prng;
  ck.sweeps = h->sweeps;
  ck.rebuilds = h->rebuilds;
}
static void seg_restore(SoupHandle* r, const SegCheckpoint& ck) {
  hipLaunchKernelGGL(k_state_scatter, dim3(blocks(r->n, 256)), dim3(256), 0, r->stream, r->n, r->b, ck.st);
  const size_t ev = 2u * (size_t)r->c.n_rules;
  if (ev) (void)hipMemcpyAsync(r->b.events, ck.events, ev * 4, hipMemcpyDeviceToDevice, r->stream);
  (void)hipMemcpyAsync(r->b.desorb, ck.desorb, 8, hipMemcpyDeviceToDevice, r->stream);
  r->global_step = ck.global_step;
  r->next_sweep_at = ck.next_sweep_at;
  r->prng = ck.prng;
  r->sweeps = ck.sweeps;
  r->rebuilds = ck.rebuilds;
}
// the steps [s0, s1) of a K-step call that began at call_start, with soup_step's chunk ends (after
// every 1000th step of the call and after its last): per step (record) or in runs (forward).
// swept[s] = 1 where the chunk end after step s ran a constant-pH sweep; rec (record mode) receives
// the state after each step, rec[i] for step s0 + i - 1.
static int seg_run(SoupHandle* h, uint32_t s0, uint32_t s1, uint32_t K, uint64_t call_start, std::vector<uint8_t>& swept,
                   const std::vector<StateRefs>* rec) {
  uint32_t s = s0;
  while (s < s1) {
    const uint32_t chunk_last = std::min(K, (s / 1000u + 1u) * 1000u) - 1u;  // the chunk's last step
    const uint32_t e = std::min(s1 - 1u, chunk_last);
    if (rec) {
      for (uint32_t q = s; q <= e; ++q) {
        one_step(h);
        if (q != chunk_last)
          hipLaunchKernelGGL(k_state_gather, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->n, h->b, (*rec)[q - s0 + 1]);
      }
    } else {
      const int r = chunk_steps(h, e - s + 1u);
      if (r) return r;
    }
    if (e == chunk_last) {
      const uint32_t sw0 = h->sweeps;
      const int r = chunk_end(h, call_start + (uint64_t)(chunk_last / 1000u) * 1000u);
      if (r) return r;
      const uint8_t sw = h->sweeps != sw0 ? 1 : 0;
      if (rec) {
        if (sw != swept[e])
          return fail(h, SOUP_E_DEVICE, fmt("soup_segment_adjoint: the recomputed step %u disagrees with the forward pass "
                                            "on its constant-pH sweep", e));
        hipLaunchKernelGGL(k_state_gather, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->n, h->b, (*rec)[e - s0 + 1]);
      } else {
        swept[e] = sw;
      }
    }
    s = e + 1u;
  }
  return 0;
}
// a system holding h's current state and configuration (the replay system of the reverse sweep)
static int seg_clone(SoupHandle* h, SoupHandle** out) {
  const size_t n = h->n;
  std::vector<float> pos(4 * n), vel(4 * n), ch(n);
  std::vector<uint32_t> bs(3 * n), cl(n), held(n), fr(n), brng(n), trng(n), ev(std::max<size_t>(2u * h->c.n_rules, 1)), de(2);
  int r = soup_read(h, SOUP_FIELD_POS, pos.data(), 16 * n);
  if (!r) r = soup_read(h, SOUP_FIELD_VEL, vel.data(), 16 * n);
  if (!r) r = soup_read(h, SOUP_FIELD_CHARGE, ch.data(), 4 * n);
  if (!r) r = soup_read(h, SOUP_FIELD_BOND_SLOTS, bs.data(), 12 * n);
  if (!r) r = soup_read(h, SOUP_FIELD_CENTER_LINK, cl.data(), 4 * n);
  if (!r) r = soup_read(h, SOUP_FIELD_CENTER_HELD, held.data(), 4 * n);
  if (!r) r = soup_read(h, SOUP_FIELD_FROZEN, fr.data(), 4 * n);
  if (!r) r = soup_read(h, SOUP_FIELD_BOND_RNG, brng.data(), 4 * n);
  if (!r) r = soup_read(h, SOUP_FIELD_THERMO_RNG, trng.data(), 4 * n);
  if (!r && h->c.n_rules) r = soup_read(h, SOUP_FIELD_EVENTS, ev.data(), 8 * (size_t)h->c.n_rules);
  if (!r) r = soup_read(h, SOUP_FIELD_DESORB, de.data(), 8);
  if (r) return r;
  SoupInitState st{};
  st.positions = pos.data();
  st.velocities = vel.data();
  st.bond_slots = bs.data();
  st.center_link = cl.data();
  st.center_held = held.data();
  st.frozen = fr.data();
  st.bond_rng = brng.data();
  st.thermo_rng = trng.data();
  st.charges = ch.data();
  st.events = ev.data();
  st.desorb_events = de.data();
  st.global_step = h->global_step;
  st.next_sweep_at = h->next_sweep_at;
  st.protonation_rng = h->prng;
  SoupConfig c = h->c;
  c.n = h->n;
  r = soup_create(&c, &st, out);
  if (r) return fail(h, r, "soup_segment_adjoint: the replay system: " + std::string(soup_last_error(nullptr)));
  (*out)->use_graphs = false;
  return 0;
}

extern "C" {

int soup_segment_adjoint(SoupHandle* h, const SoupPotentialBasis* pb, const SoupSegmentSpec* sp, double* J, double* dJ_dtheta,
                         double* dJ_dx0, double* dJ_dv0) {
  static const char* who = "soup_segment_adjoint";
  if (!h) return SOUP_E_INVALID;
  if (!pb || !sp || !J || !dJ_dtheta) return fail(h, SOUP_E_INVALID, fmt("%s: null argument", who));
  const SoupConfig& c = h->c;
  const uint32_t K = sp->steps, C = sp->checkpoint_every, E = c.rebuild_every;
  if (K == 0 || C == 0) return fail(h, SOUP_E_INVALID, fmt("%s: steps and checkpoint_every must be >= 1", who));
  if (sp->objective > SOUP_OBJ_MEAN_STRAIN_VIRIAL) return fail(h, SOUP_E_INVALID, fmt("%s: unknown objective %u", who, sp->objective));
  if (drift_mode(h)) return fail(h, SOUP_E_INVALID, fmt("%s: needs the reference's rebuild cadence (rebuild = cadence)", who));
  // a recomputed stretch starts where a cadence step rebuilds anyway, so it rebuilds the lists the
  // forward pass built and reproduces it bit for bit from the checkpointed state
  if (h->global_step % E != 0 || C % E != 0)
    return fail(h, SOUP_E_INVALID, fmt("%s: the start step (%llu) and checkpoint_every (%u) must be multiples of rebuildEvery (%u)", who,
                                       (unsigned long long)h->global_step, C, E));
  if (h->global_step > 0 && (h->global_step - 1) % c.bond_interval == 0)
    return fail(h, SOUP_E_INVALID, fmt("%s: the segment starts right after a bond Monte Carlo step (its force buffer holds the "
                                       "previous topology's force); start one step later", who));
  {
    std::vector<uint32_t> fr(h->n);
    int r = soup_read(h, SOUP_FIELD_FROZEN, fr.data(), 4 * (size_t)h->n);
    if (r) return r;
    for (uint32_t v : fr)
      if (v) return fail(h, SOUP_E_INVALID, fmt("%s: frozen particles are not differentiated", who));
  }
  bool es = false;
  int r = adjoint_ready(h, who, pb, es);
  if (r) return r;
  r = sync(h);
  if (r) return r;
  const uint32_t n = h->n, nseg = (K + C - 1) / C;
  const size_t ev = 2u * (size_t)c.n_rules;
  const uint64_t call_start = h->global_step;
  DevScratch ds;
  auto alloc_state = [&](StateRefs& s, bool topo_only) -> hipError_t {
    hipError_t e = ds.get(&s.pos, n);
    if (!e) e = ds.get(&s.charge, n);
    if (!e) e = ds.get(&s.bs, n);
    if (!e) e = ds.get(&s.clink, n);
    if (!e && !topo_only) {
      e = ds.get(&s.vel, n);
      if (!e) e = ds.get(&s.force, n);
      if (!e) e = ds.get(&s.cheld, n);
      if (!e) e = ds.get(&s.frozen, n);
      if (!e) e = ds.get(&s.brng, n);
      if (!e) e = ds.get(&s.trng, n);
    }
    return e;
  };
  std::vector<SegCheckpoint> ck(nseg);
  std::vector<StateRefs> rec(C + 1u);
  double *lx = nullptr, *lv = nullptr, *acc = nullptr;
  hipError_t he = hipSuccess;
  for (auto& k : ck) {
    if (!he) he = alloc_state(k.st, false);
    if (!he) he = ds.get(&k.events, std::max<size_t>(ev, 1));
    if (!he) he = ds.get(&k.desorb, 2);
  }
  for (auto& q : rec)
    if (!he) he = alloc_state(q, true);
  if (!he) he = ds.get(&lx, 3 * (size_t)n);
  if (!he) he = ds.get(&lv, 3 * (size_t)n);
  if (!he) he = ds.get(&acc, ADJ_P + 1);
  if (he)
    return fail(h, SOUP_E_NOMEM, fmt("%s: device buffers for %u checkpoints and a %u-step record of %u particles: %s", who, nseg, C,
                                     n, hipGetErrorString(he)));
  // ---- forward pass on h (exactly soup_step(h, K)), a checkpoint before every stretch of C steps ----
  std::vector<uint8_t> swept(K, 0);
  for (uint32_t k = 0; k < nseg; ++k) {
    seg_save(h, ck[k]);
    r = seg_run(h, k * C, std::min(K, (k + 1) * C), K, call_start, swept, nullptr);
    if (r) return r;
  }
  r = sync(h);
  if (r) return r;
  HC(h, hipMemsetAsync(lx, 0, 3 * (size_t)n * sizeof(double), h->stream));
  HC(h, hipMemsetAsync(lv, 0, 3 * (size_t)n * sizeof(double), h->stream));
  HC(h, hipMemsetAsync(acc, 0, (ADJ_P + 1) * sizeof(double), h->stream));
  double Jend = 0.0;
  if (sp->objective == SOUP_OBJ_END_ENERGY) {
    // J = U(x_K) on the state the pass leaves: dJ/dx_K = g, dJ/dtheta starts at dU/dtheta
    double pot[5];
    r = soup_potential(h, pb, pot);
    if (r) return r;
    Jend = pot[4];
    r = assert_lists(h);
    if (r) return r;
    adjoint_launch(h, pb, es, nullptr, AdjStrain{{0.0, 0.0, 0.0}});
    HC(h, hipMemcpyAsync(lx, h->adj_g, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    hipLaunchKernelGGL(k_seg_acc, dim3(1), dim3(64), 0, h->stream, h->adj_out, acc, 1);
  }
  r = sync(h);
  if (r) return r;
  // ---- reverse sweep on a replay system: each stretch recomputed from its checkpoint ----
  SoupHandle* rp = nullptr;
  r = seg_clone(h, &rp);
  if (r) return r;
  std::unique_ptr<SoupHandle, int (*)(SoupHandle*)> guard(rp, soup_destroy);
  bool es_r = false;
  r = adjoint_ready(rp, who, pb, es_r);
  if (r) return fail(h, r, rp->err);
  const double dt = c.dt, a = 1.0 - (double)c.gamma * (double)c.dt;
  AdjStrain obj{{0.0, 0.0, 0.0}};
  if (sp->objective == SOUP_OBJ_MEAN_STRAIN_VIRIAL)
    for (int q = 0; q < 3; ++q) obj.e[q] = sp->strain[q] * sp->scale / (double)K;
  const AdjStrain none{{0.0, 0.0, 0.0}};
  const uint32_t n3 = 3u * n;
  // one launch at a recorded state: positions of rec[ip], topology of rec[it], direction cw * lv
  // (cw = 0: none) plus the strain st
  auto vjp = [&](const StateRefs& P, const StateRefs& T, double cw, const AdjStrain& st) {
    StateRefs ld{};
    ld.pos = P.pos;
    ld.charge = T.charge;
    ld.bs = T.bs;
    ld.clink = T.clink;
    hipLaunchKernelGGL(k_state_scatter, dim3(blocks(n, 256)), dim3(256), 0, rp->stream, n, rp->b, ld);
    ph_rebuild(rp);
    if (cw != 0.0) hipLaunchKernelGGL(k_seg_w, dim3(blocks(n3, 256)), dim3(256), 0, rp->stream, n3, lv, cw, rp->adj_w);
    adjoint_launch(rp, pb, es_r, cw != 0.0 ? (const double*)rp->adj_w : nullptr, st);
    hipLaunchKernelGGL(k_seg_lx, dim3(blocks(n, 256)), dim3(256), 0, rp->stream, n, lx, rp->adj_hw, rp->adj_g, st);
    hipLaunchKernelGGL(k_seg_acc, dim3(1), dim3(64), 0, rp->stream, rp->adj_out, acc, 0);
  };
  for (int k = (int)nseg - 1; k >= 0; --k) {
    const uint32_t s0 = (uint32_t)k * C, s1 = std::min(K, s0 + C);
    seg_restore(rp, ck[k]);
    {
      const StateRefs t0 = state_part(ck[k].st, true);
      const StateRefs& d0 = rec[0];
      HC(h, hipMemcpyAsync(d0.pos, t0.pos, 16 * (size_t)n, hipMemcpyDeviceToDevice, rp->stream));
      HC(h, hipMemcpyAsync(d0.charge, t0.charge, 4 * (size_t)n, hipMemcpyDeviceToDevice, rp->stream));
      HC(h, hipMemcpyAsync(d0.bs, t0.bs, 16 * (size_t)n, hipMemcpyDeviceToDevice, rp->stream));
      HC(h, hipMemcpyAsync(d0.clink, t0.clink, 4 * (size_t)n, hipMemcpyDeviceToDevice, rp->stream));
    }
    r = seg_run(rp, s0, s1, K, call_start, swept, &rec);
    if (r) return fail(h, r, rp->err);
    for (uint32_t s = s1; s-- > s0;) {
      const uint32_t i = s - s0 + 1u;
      const bool bond_step = (call_start + s) % c.bond_interval == 0;
      const bool has_obj = sp->objective == SOUP_OBJ_MEAN_STRAIN_VIRIAL;
      const bool last = s == K - 1u;
      // x_{s+1} is where step s's second half-kick (a dt/2, topology at the start of step s) and step
      // s + 1's first half-kick (dt/2; after a sweep, with the post-sweep topology) take their force,
      // and where the objective observes the state after step s
      double cw_start = a * dt / 2.0 + (last ? 0.0 : dt / 2.0);
      bool obj_done = !has_obj;
      if (!last && swept[s]) {
        vjp(rec[i], rec[i], dt / 2.0, has_obj ? obj : none);
        cw_start = a * dt / 2.0;
        obj_done = true;
      }
      if (!obj_done && bond_step) {  // the bonds may have changed during step s
        vjp(rec[i], rec[i], 0.0, obj);
        obj_done = true;
      }
      vjp(rec[i], rec[i - 1u], cw_start, obj_done ? none : obj);
      hipLaunchKernelGGL(k_seg_lv, dim3(blocks(n3, 256)), dim3(256), 0, rp->stream, n3, lv, lx, a, dt);
    }
    if (k == 0) vjp(rec[0], rec[0], dt / 2.0, none);  // step 0's first half-kick at x_0
    r = sync(rp);
    if (r) return fail(h, r, rp->err);
    r = assert_lists(rp);
    if (r) return fail(h, r, rp->err);
  }
  double accv[ADJ_P + 1];
  HC(h, hipMemcpy(accv, acc, sizeof(accv), hipMemcpyDeviceToHost));
  for (uint32_t p = 0; p < ADJ_P; ++p) dJ_dtheta[p] = accv[p];
  *J = sp->objective == SOUP_OBJ_END_ENERGY ? Jend : accv[ADJ_P];
  if (dJ_dx0) HC(h, hipMemcpy(dJ_dx0, lx, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  if (dJ_dv0) HC(h, hipMemcpy(dJ_dv0, lv, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

// the scratch cell grid of soup_potential_at and soup_force_candidate_stats: per-particle arrays at
// the capacity, count and start arrays grown to ncell + 1
static int pa_scratch(SoupHandle* h, size_t ncell) {
  if (!h->pa_xs) {
    DA(h, h->pa_xs, h->cap);
    DA(h, h->pa_cell, h->cap);
    DA(h, h->pa_order, h->cap);
  }
  size_t cap_c = h->pa_ncell_cap;
  int r = dgrow(h, &h->pa_count, ncell + 1, cap_c);
  if (r) return r;
  size_t cap_s = h->pa_ncell_cap;
  r = dgrow(h, &h->pa_start, ncell + 1, cap_s);
  if (r) return r;
  h->pa_ncell_cap = std::min(cap_c, cap_s);
  return 0;
}

// device buffers of one soup_potential_at evaluation: positions, the configuration's cell grid, the
// per-particle partial sums and the five totals
struct PaBufs {
  float4* xs;
  uint32_t *cell, *order, *count, *start;
  double4* part;
  double* dout;
};

// The argument checks and the configuration's cell grid (soup-potential.ts:156-158) shared by both
// forms of soup_potential_at; `who` names the entry point in messages.
static int pa_grid(SoupHandle* h, const char* who, const SoupPotentialBasis* pb, const float* positions, const double box[3],
                   const double* out, PaGrid& g, size_t& ncell) {
  if (!pb || !positions || !box || !out) return fail(h, SOUP_E_INVALID, fmt("%s: null argument", who));
  if (!(pb->cutoff > 0.0)) return fail(h, SOUP_E_INVALID, fmt("%s: the pair cutoff must be > 0", who));
  ncell = 1;
  for (int a = 0; a < 3; ++a) {
    if (!(box[a] > 0.0) || !std::isfinite(box[a])) return fail(h, SOUP_E_INVALID, fmt("%s: box edges must be finite and > 0", who));
    const double d = std::floor(box[a] / pb->cutoff);
    g.d[a] = d < 1.0 ? 1u : (uint32_t)std::min(d, 1024.0);
    g.L[a] = box[a];
    g.w[a] = box[a] / (double)g.d[a];
    ncell *= g.d[a];
  }
  for (uint32_t i = 0; i < h->n; ++i) {
    const float k = positions[4u * i + 3];
    if (!(k >= 0.0f) || k >= (float)h->c.n_kinds || !std::isfinite(positions[4u * i]) || !std::isfinite(positions[4u * i + 1]) ||
        !std::isfinite(positions[4u * i + 2]))
      return fail(h, SOUP_E_INVALID, fmt("%s: particle %u has a non-finite position or a kind outside the table", who, i));
  }
  return 0;
}

// the evaluation itself, into buffers the caller provides (no device allocation here)
static int pa_run(SoupHandle* h, const SoupPotentialBasis* pb, const float* positions, const PaGrid& g, size_t ncell,
                  const PaBufs& q, double out[5]) {
  const uint32_t n = h->n;
  HC(h, hipMemcpyAsync(q.xs, positions, (size_t)n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
  fill(h, q.count, 0u, (uint32_t)ncell);
  hipLaunchKernelGGL(k_pa_cells, dim3(blocks(n, 256)), dim3(256), 0, h->stream, q.xs, n, g, q.cell, q.count);
  scan(h, q.count, (uint32_t)ncell, q.start);
  HC(h, hipMemcpyAsync(q.count, q.start, ncell * sizeof(uint32_t), hipMemcpyDeviceToDevice, h->stream));
  hipLaunchKernelGGL(k_pa_fill, dim3(blocks(n, 256)), dim3(256), 0, h->stream, q.cell, n, q.count, q.order);
  hipLaunchKernelGGL(k_pa_sort, dim3(blocks(ncell, 256)), dim3(256), 0, h->stream, q.start, (uint32_t)ncell, q.order);
  const bool es = es_on(h) && pb->use_charges && pb->es_A != 0.0;
  hipLaunchKernelGGL(k_potential_at, dim3(blocks((uint64_t)n * WAVE, 256)), dim3(256), 0, h->stream, h->c, h->b, *pb, q.xs, g,
                     q.cell, q.start, q.order, q.part, es ? h->nh : 0u);
  hipLaunchKernelGGL(k_potential_sum, dim3(1), dim3(1024), 0, h->stream, q.part, n, q.dout);
  HC(h, hipMemcpyAsync(out, q.dout, 5 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  return sync(h);
}

int soup_potential_at(SoupHandle* h, const SoupPotentialBasis* pb, const float* positions, const double box[3], double out[5]) {
  if (!h) return SOUP_E_INVALID;
  PaGrid g;
  size_t ncell = 0;
  int r = pa_grid(h, "soup_potential_at", pb, positions, box, out, g, ncell);
  if (r) return r;
  const uint32_t n = h->n;
  const size_t need = (size_t)n * sizeof(double4) + 64;
  if (need > h->stage_bytes) return fail(h, SOUP_E_INVALID, "soup_potential_at: staging buffer too small");
  r = sync(h);
  if (r) return r;
  r = pa_scratch(h, ncell);
  if (r) return r;
  PaBufs q{h->pa_xs, h->pa_cell, h->pa_order, h->pa_count, h->pa_start, reinterpret_cast<double4*>(h->stage),
           reinterpret_cast<double*>(h->stage + (size_t)n * sizeof(double4))};
  return pa_run(h, pb, positions, g, ncell, q, out);
}

// ---- soup_potential_at_noalloc: the same evaluation in a fixed device array ---------------------
// Every buffer soup_potential_at allocates on first use lives here instead, carved at 256-byte
// boundaries: positions (16 B), cell and order (4 B each) and partial sums (32 B) per particle up to
// SOUP_NOALLOC_MAX_N, count and start arrays of SOUP_NOALLOC_MAX_CELLS + 1 entries, and the totals.
// One array per GPU, held until the process ends; calls on one GPU share it.
static constexpr size_t pa_align(size_t x) { return (x + 255u) & ~(size_t)255u; }
static constexpr size_t kNaXs = 0;
static constexpr size_t kNaCell = kNaXs + pa_align((size_t)SOUP_NOALLOC_MAX_N * 16u);
static constexpr size_t kNaOrder = kNaCell + pa_align((size_t)SOUP_NOALLOC_MAX_N * 4u);
static constexpr size_t kNaCount = kNaOrder + pa_align((size_t)SOUP_NOALLOC_MAX_N * 4u);
static constexpr size_t kNaStart = kNaCount + pa_align(((size_t)SOUP_NOALLOC_MAX_CELLS + 1u) * 4u);
static constexpr size_t kNaPart = kNaStart + pa_align(((size_t)SOUP_NOALLOC_MAX_CELLS + 1u) * 4u);
static constexpr size_t kNaOut = kNaPart + pa_align((size_t)SOUP_NOALLOC_MAX_N * 32u);
static constexpr size_t kNaBytes = kNaOut + 256u;
static __device__ __align__(256) unsigned char scratch[kNaBytes];

#define SOUP_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[SOUP_MAX_DEVICES];  // device address of `scratch` per GPU, once resolved

// The device address of this GPU's `scratch` (the array's name is a host-side stand-in that no kernel
// may be given): resolved once per device with that device current, then kept.
static int scratch_on(SoupHandle* h, unsigned char** out) {
  if (h->device < 0 || h->device >= SOUP_MAX_DEVICES)
    return fail(h, SOUP_E_INVALID, fmt("soup_potential_at_noalloc: device %d is outside the %d this library tracks", h->device,
                                       SOUP_MAX_DEVICES));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[h->device]) {
    int cur = 0;
    HC(h, hipGetDevice(&cur));
    if (cur != h->device) HC(h, hipSetDevice(h->device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != h->device) HC(h, hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return fail(h, SOUP_E_DEVICE, std::string("soup_potential_at_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[h->device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[h->device];
  return 0;
}

int soup_potential_at_noalloc(SoupHandle* h, const SoupPotentialBasis* pb, const float* positions, const double box[3],
                              double out[5]) {
  if (!h) return SOUP_E_INVALID;
  PaGrid g;
  size_t ncell = 0;
  int r = pa_grid(h, "soup_potential_at_noalloc", pb, positions, box, out, g, ncell);
  if (r) return r;
  if (h->n > SOUP_NOALLOC_MAX_N)
    return fail(h, SOUP_E_INVALID,
                fmt("soup_potential_at_noalloc: %u particles exceed the fixed scratch array's %u (use soup_potential_at)", h->n,
                    (unsigned)SOUP_NOALLOC_MAX_N));
  if (ncell > SOUP_NOALLOC_MAX_CELLS)
    return fail(h, SOUP_E_INVALID,
                fmt("soup_potential_at_noalloc: the configuration's grid of %zu cells exceeds the fixed scratch array's %u (use "
                    "soup_potential_at)",
                    ncell, (unsigned)SOUP_NOALLOC_MAX_CELLS));
  unsigned char* base = nullptr;
  r = scratch_on(h, &base);
  if (r) return r;
  r = sync(h);
  if (r) return r;
  PaBufs q{reinterpret_cast<float4*>(base + kNaXs), reinterpret_cast<uint32_t*>(base + kNaCell),
           reinterpret_cast<uint32_t*>(base + kNaOrder), reinterpret_cast<uint32_t*>(base + kNaCount),
           reinterpret_cast<uint32_t*>(base + kNaStart), reinterpret_cast<double4*>(base + kNaPart),
           reinterpret_cast<double*>(base + kNaOut)};
  return pa_run(h, pb, positions, g, ncell, q, out);
}

int soup_set_box(SoupHandle* h, const SoupConfig* geometry, const float* positions, uint32_t flags) {
  if (!h) return SOUP_E_INVALID;
  if (!geometry) return fail(h, SOUP_E_INVALID, "soup_set_box: null geometry");
  SoupConfig c = h->c;
  for (int a = 0; a < 3; ++a) {
    c.box[a] = geometry->box[a];
    c.dims[a] = geometry->dims[a];
    if (!(c.box[a] > 0.0f) || c.dims[a] == 0)
      return fail(h, SOUP_E_INVALID, "soup_set_box: box edges and grid dims must be > 0");
  }
  c.walk_radius = geometry->walk_radius;
  int v = validate(&c);
  if (v) return fail(h, v, g_create_error);
  const uint32_t n = h->n;
  if (positions) {
    for (uint32_t i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        // wrapped in FP64 and rounded to f32, a coordinate may land on L itself
        const float x = positions[4u * i + a];
        if (!(x >= 0.0f && x <= c.box[a]))
          return fail(h, SOUP_E_INVALID, fmt("soup_set_box: particle %u coordinate %d = %g lies outside [0, %g]", i, a, x, c.box[a]));
      }
  } else {
    // the particles stay where they are, which must lie inside the new box
    for (int a = 0; a < 3; ++a)
      if (c.box[a] < h->c.box[a])
        return fail(h, SOUP_E_INVALID, "soup_set_box: keeping the positions needs a box at least as large on every axis");
  }
  int r = sync(h);
  if (r) return r;
  const size_t ncells = (size_t)c.dims[0] * c.dims[1] * c.dims[2];
  if (ncells > h->ncells_cap) {
    size_t cc = h->ncells_cap, cs = h->ncells_cap;
    r = dgrow(h, &h->b.cell_count, ncells, cc);
    if (r) return r;
    r = dgrow(h, &h->b.cell_start, ncells + 1, cs);
    if (r) return r;
    h->ncells_cap = ncells;
  }
  drop_graphs(h);  // their kernels carry the old geometry
  h->c = c;
  h->ncells = (uint32_t)ncells;
  r = es_grid_setup(h);
  if (r) return r;
  if (positions) {
    r = soup_write(h, SOUP_FIELD_POS, positions, (size_t)n * 16);
    if (r) return r;
  }
  ph_rebuild(h);
  if (flags & SOUP_SET_BOX_FORCE) ph_force(h);
  r = sync(h);
  if (r) return r;
  r = assert_finite(h, false, 0);
  if (r) return r;
  return assert_lists(h);
}

// Math.round of the reference: the nearest integer, halves toward +infinity (floor(x + 0.5) would
// round 0.49999999999999994 up)
static double host_js_round(double x) {
  double r = std::floor(x);
  if (x - r >= 0.5) r += 1.0;
  return r;
}

// scaleMoleculesRigid (soup/src/soup-box-scale-math.ts:63-138): host code, FP64 in the reference's
// order. Components of the bond graph by BFS from the lowest unvisited index, neighbours in the order
// the pairs list them; each member's offset from the root accumulated one minimum-image hop at a
// time; the centre of mass (mean offset, summed in BFS order) wrapped and scaled per axis, and every
// member rebuilt around it from its unchanged offset.
int soup_scale_molecules_rigid(const float* positions, uint32_t n, const uint32_t* bond_pairs, uint32_t n_pairs,
                               const double old_box[3], const double new_box[3], float* out) {
  if (!positions || !old_box || !new_box || !out || (n_pairs && !bond_pairs)) return SOUP_E_INVALID;
  std::vector<uint32_t> deg(n + 1, 0);
  for (uint32_t k = 0; k < n_pairs; ++k) {
    const uint32_t a = bond_pairs[2u * k], b = bond_pairs[2u * k + 1];
    if (a >= n || b >= n) return SOUP_E_INVALID;
    deg[a + 1]++;
    deg[b + 1]++;
  }
  for (uint32_t i = 0; i < n; ++i) deg[i + 1] += deg[i];
  std::vector<uint32_t> adj(deg[n]), fillp(deg.begin(), deg.end() - 1);
  for (uint32_t k = 0; k < n_pairs; ++k) {
    const uint32_t a = bond_pairs[2u * k], b = bond_pairs[2u * k + 1];
    adj[fillp[a]++] = b;
    adj[fillp[b]++] = a;
  }
  auto mi1 = [](double d, double L) { return d - host_js_round(d / L) * L; };
  auto wrap1 = [](double v, double L) { return v - std::floor(v / L) * L; };
  const double s[3] = {new_box[0] / old_box[0], new_box[1] / old_box[1], new_box[2] / old_box[2]};
  std::vector<uint8_t> seen(n, 0);
  std::vector<double> off(3 * (size_t)n, 0.0);
  std::vector<uint32_t> members;
  for (uint32_t root = 0; root < n; ++root) {
    if (seen[root]) continue;
    seen[root] = 1;
    off[3u * root] = off[3u * root + 1] = off[3u * root + 2] = 0.0;
    members.clear();
    members.push_back(root);
    for (size_t q = 0; q < members.size(); ++q) {
      const uint32_t p = members[q];
      for (uint32_t e = deg[p]; e < deg[p + 1]; ++e) {
        const uint32_t c = adj[e];
        if (seen[c]) continue;
        seen[c] = 1;
        for (int a = 0; a < 3; ++a)
          off[3u * c + a] = off[3u * p + a] + mi1((double)positions[4u * c + a] - (double)positions[4u * p + a], old_box[a]);
        members.push_back(c);
      }
    }
    double sum[3] = {0.0, 0.0, 0.0};
    for (uint32_t m : members)
      for (int a = 0; a < 3; ++a) sum[a] += off[3u * m + a];
    const double cnt = (double)members.size();
    double cen[3], com[3];
    for (int a = 0; a < 3; ++a) {
      cen[a] = sum[a] / cnt;
      com[a] = wrap1((double)positions[4u * root + a] + cen[a], old_box[a]) * s[a];
    }
    for (uint32_t m : members) {
      for (int a = 0; a < 3; ++a) out[4u * m + a] = (float)wrap1(com[a] + (off[3u * m + a] - cen[a]), new_box[a]);
      out[4u * m + 3] = positions[4u * m + 3];
    }
  }
  return 0;
}

// proximityPairs (soup/src/soup-box-scale-math.ts:162-215): a cell list whose cells hold their members
// in ascending index; the FP64 arithmetic of the reference (wrap1, floor, mi1)
int soup_proximity_pairs(const float* positions, uint32_t n, const double box[3], double cutoff, const uint8_t* eligible,
                         uint32_t* pairs, uint64_t max_pairs, uint64_t* count) {
  if (!positions || !box || !eligible || !count || !(cutoff > 0.0)) return SOUP_E_INVALID;
  int nc[3];
  double w[3];
  for (int a = 0; a < 3; ++a) {
    if (!(box[a] > 0.0)) return SOUP_E_INVALID;
    nc[a] = std::max(1, (int)std::floor(box[a] / cutoff));
    w[a] = box[a] / (double)nc[a];
  }
  const size_t ncell = (size_t)nc[0] * nc[1] * nc[2];
  std::vector<std::vector<uint32_t>> bucket(ncell);
  std::vector<int> cx(n), cy(n), cz(n);
  for (uint32_t i = 0; i < n; ++i) {
    if (!eligible[i]) continue;
    int cc[3];
    for (int a = 0; a < 3; ++a) {
      const double v = (double)positions[4u * i + a];
      const double x = v - std::floor(v / box[a]) * box[a];
      cc[a] = std::min(nc[a] - 1, (int)std::floor(x / w[a]));
    }
    cx[i] = cc[0];
    cy[i] = cc[1];
    cz[i] = cc[2];
    bucket[(size_t)cc[0] + (size_t)nc[0] * ((size_t)cc[1] + (size_t)nc[1] * (size_t)cc[2])].push_back(i);
  }
  auto mi1 = [](double d, double L) { return d - host_js_round(d / L) * L; };
  const double cut2 = cutoff * cutoff;
  uint64_t k = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (!eligible[i]) continue;
    for (int dx = -1; dx <= 1; ++dx) {
      const int ncx = ((cx[i] + dx) % nc[0] + nc[0]) % nc[0];
      for (int dy = -1; dy <= 1; ++dy) {
        const int ncy = ((cy[i] + dy) % nc[1] + nc[1]) % nc[1];
        for (int dz = -1; dz <= 1; ++dz) {
          const int ncz = ((cz[i] + dz) % nc[2] + nc[2]) % nc[2];
          for (uint32_t j : bucket[(size_t)ncx + (size_t)nc[0] * ((size_t)ncy + (size_t)nc[1] * (size_t)ncz)]) {
            if (j <= i) continue;
            const double ddx = mi1((double)positions[4u * i] - (double)positions[4u * j], box[0]);
            const double ddy = mi1((double)positions[4u * i + 1] - (double)positions[4u * j + 1], box[1]);
            const double ddz = mi1((double)positions[4u * i + 2] - (double)positions[4u * j + 2], box[2]);
            if (ddx * ddx + ddy * ddy + ddz * ddz <= cut2) {
              if (pairs && k < max_pairs) {
                pairs[2 * k] = i;
                pairs[2 * k + 1] = j;
              }
              ++k;
            }
          }
        }
      }
    }
  }
  *count = k;
  return 0;
}

// mulberry32 (soup/src/soup-evaporate.ts:169-174): advances the state and returns the draw in [0, 1)
static double mulberry32_next(uint32_t& a) {
  a += 0x6d2b79f5u;
  uint32_t t = (a ^ (a >> 15)) * (1u | a);
  t = (t + (t ^ (t >> 7)) * (61u | t)) ^ t;
  return (double)(t ^ (t >> 14)) / 4294967296.0;
}

// sampleInsertionPositions (soup/src/soup-evaporate-plan.ts:202-304), host code in FP64 as the
// reference computes it: cells of side >= min_sep (floor(box / min_sep) per axis) keyed
// cx + nx (cy + ny cz), the 27 wrapped neighbour cells searched for the nearest point under the
// minimum image, candidates x, y, z drawn in that order. The search only ever looks up the keys of
// wrapped cells, so a point whose own key falls outside them (a coordinate below 0) is never found,
// as in the reference; placed beads join the search at their FP64 coordinates.
int soup_sample_insertion(const float* active, uint32_t n_active, const double box[3], uint32_t count, double min_sep,
                          uint32_t* rng, float* out, uint32_t* short_of_floor, double* min_achieved) {
  if (!box || !rng || (count && !out) || !short_of_floor || !min_achieved || (n_active && !active) ||
      !(min_sep > 0.0) || !std::isfinite(min_sep))
    return SOUP_E_INVALID;
  int64_t nc[3];
  double w[3];
  for (int a = 0; a < 3; ++a) {
    if (!(box[a] > 0.0) || !std::isfinite(box[a])) return SOUP_E_INVALID;
    nc[a] = std::max<int64_t>(1, (int64_t)std::floor(box[a] / min_sep));
    w[a] = box[a] / (double)nc[a];
  }
  const int64_t ncell = nc[0] * nc[1] * nc[2];
  if (ncell > (int64_t)1 << 27) return SOUP_E_INVALID;
  for (size_t i = 0; i < (size_t)n_active * 4; ++i)
    if (!std::isfinite(active[i])) return SOUP_E_INVALID;
  // per-cell singly linked lists: head[key] -> last point pushed, link[p] -> the one before
  std::vector<int64_t> head((size_t)ncell, -1);
  std::vector<int64_t> link;
  std::vector<double> xs, ys, zs;
  const size_t total = (size_t)n_active + count;
  link.reserve(total);
  xs.reserve(total);
  ys.reserve(total);
  zs.reserve(total);
  auto cell = [&](double v, int a) { return std::min<int64_t>(nc[a] - 1, (int64_t)std::floor(v / w[a])); };
  auto push = [&](double x, double y, double z) {
    const int64_t idx = (int64_t)xs.size();
    xs.push_back(x);
    ys.push_back(y);
    zs.push_back(z);
    const int64_t key = cell(x, 0) + nc[0] * (cell(y, 1) + nc[1] * cell(z, 2));
    if (key >= 0 && key < ncell) {
      link.push_back(head[(size_t)key]);
      head[(size_t)key] = idx;
    } else {
      link.push_back(-1);
    }
  };
  for (uint32_t i = 0; i < n_active; ++i)
    push((double)active[4u * i], (double)active[4u * i + 1], (double)active[4u * i + 2]);
  auto mi1 = [](double d, double L) { return d - host_js_round(d / L) * L; };
  auto nearest = [&](double x, double y, double z) {
    const int64_t cx = cell(x, 0), cy = cell(y, 1), cz = cell(z, 2);
    double best = INFINITY;
    for (int dx = -1; dx <= 1; ++dx) {
      const int64_t ax = ((cx + dx) % nc[0] + nc[0]) % nc[0];
      for (int dy = -1; dy <= 1; ++dy) {
        const int64_t ay = ((cy + dy) % nc[1] + nc[1]) % nc[1];
        for (int dz = -1; dz <= 1; ++dz) {
          const int64_t az = ((cz + dz) % nc[2] + nc[2]) % nc[2];
          for (int64_t j = head[(size_t)(ax + nc[0] * (ay + nc[1] * az))]; j >= 0; j = link[(size_t)j]) {
            const double mx = mi1(x - xs[(size_t)j], box[0]);
            const double my = mi1(y - ys[(size_t)j], box[1]);
            const double mz = mi1(z - zs[(size_t)j], box[2]);
            const double d2 = mx * mx + my * my + mz * mz;
            if (d2 < best) best = d2;
          }
        }
      }
    }
    return std::sqrt(best);
  };
  uint32_t shortc = 0;
  double min_a = INFINITY;
  uint32_t a = *rng;
  for (uint32_t k = 0; k < count; ++k) {
    double bx = 0.0, by = 0.0, bz = 0.0, bd = -1.0;
    for (int t = 0; t < 64; ++t) {  // INSERTION_CANDIDATES
      const double x = mulberry32_next(a) * box[0];
      const double y = mulberry32_next(a) * box[1];
      const double z = mulberry32_next(a) * box[2];
      const double d = nearest(x, y, z);
      if (d > bd) {
        bd = d;
        bx = x;
        by = y;
        bz = z;
      }
      if (d >= min_sep) break;
    }
    if (bd < min_sep) ++shortc;
    if (bd < min_a) min_a = bd;
    out[4u * k] = (float)bx;
    out[4u * k + 1] = (float)by;
    out[4u * k + 2] = (float)bz;
    out[4u * k + 3] = 0.0f;
    push(bx, by, bz);
  }
  *rng = a;
  *short_of_floor = shortc;
  *min_achieved = std::isfinite(min_a) ? min_a : 0.0;
  return 0;
}

int soup_relax(SoupHandle* h, uint32_t iterations, double d0) {
  if (!h) return SOUP_E_INVALID;
  for (uint32_t k = 0; k < iterations; ++k) {
    // the cap in FP64, rounded once to f32 (soup-relax.ts:92-93)
    float d = (float)(d0 * (1.0 - (double)k / (double)iterations));
    ph_rebuild(h);
    ph_force(h);
    hipLaunchKernelGGL(k_relax_step, dim3(blocks(h->n, 256)), dim3(256), 0, h->stream, h->c, h->b, d);
    if ((k + 1) % 25 == 0) {
      int r = sync(h);
      if (r) return r;
    }
  }
  ph_rebuild(h);
  ph_force(h);
  int r = sync(h);
  if (r) return r;
  r = assert_finite(h, false, 0);
  if (r) return r;
  return assert_lists(h);
}

int soup_force_candidate_stats(SoupHandle* h, uint64_t out[2]) {
  if (!h || !out) return SOUP_E_INVALID;
  int r = sync(h);
  if (r) return r;
  r = pa_scratch(h, h->ncells);
  if (r) return r;
  const uint32_t n = h->n;
  fill(h, h->pa_count, 0u, h->ncells);
  hipLaunchKernelGGL(k_stats_cells, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->c, h->b, h->pa_cell, h->pa_count);
  scan(h, h->pa_count, h->ncells, h->pa_start);
  HC(h, hipMemcpyAsync(h->pa_count, h->pa_start, (size_t)h->ncells * sizeof(uint32_t), hipMemcpyDeviceToDevice, h->stream));
  hipLaunchKernelGGL(k_pa_fill, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->pa_cell, n, h->pa_count, h->pa_order);
  HC(h, hipMemsetAsync(h->lsum, 0, 2 * sizeof(unsigned long long), h->stream));
  hipLaunchKernelGGL(k_force_stats, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->c, h->b, h->pa_cell, h->pa_start,
                     h->pa_order, h->lsum);
  unsigned long long v[2];
  HC(h, hipMemcpyAsync(v, h->lsum, sizeof(v), hipMemcpyDeviceToHost, h->stream));
  r = sync(h);
  if (r) return r;
  out[0] = v[0];
  out[1] = v[1];
  return 0;
}

int soup_step_phases(SoupHandle* h, uint32_t n, double ms[5]) {
  if (!h || !ms || n == 0) return SOUP_E_INVALID;
  int r = sync(h);
  if (r) return r;
  hipEvent_t a = nullptr, e = nullptr;
  HC(h, hipEventCreate(&a));
  if (hipEventCreate(&e) != hipSuccess) {
    (void)hipEventDestroy(a);
    return fail(h, SOUP_E_DEVICE, "soup_step_phases: event creation failed");
  }
  for (int g = 0; g < 5 && r == 0; ++g) {
    (void)hipEventRecord(a, h->stream);
    for (uint32_t k = 0; k < n; ++k) {
      switch (g) {
        case 0:  // full: one step with the rebuild and the bond phase
          ph_kick_drift(h);
          ph_rebuild(h);
          ph_force(h);
          ph_drift(h);
          ph_bonds(h);
          ph_kick_thermo(h);
          break;
        case 1: ph_rebuild(h); break;
        case 2: ph_force(h); break;
        case 3: ph_bonds(h); break;
        default:
          ph_kick_drift(h);
          ph_kick_thermo(h);
          break;
      }
    }
    (void)hipEventRecord(e, h->stream);
    r = sync(h);
    float t = 0.f;
    if (r == 0 && hipEventElapsedTime(&t, a, e) == hipSuccess) ms[g] = (double)t / (double)n;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(e);
  if (r) return r;
  // the lists and F(x) of the state the next step starts from
  ph_rebuild(h);
  ph_force(h);
  r = sync(h);
  if (r) return r;
  r = assert_finite(h, false, 0);
  if (r) return r;
  return assert_lists(h);
}

}  // extern "C"

#include "soup_analysis.inl"
