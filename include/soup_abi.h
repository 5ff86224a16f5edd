/*
 * soup_abi.h -- C-ABI of the MI355X soup Langevin-step library (libsoup_hip.so).
 *
 * The library owns one simulated periodic box per handle: particle state in HBM, the cell grid,
 * the Verlet and Debye-Hueckel neighbour lists, and one HIP stream. Every call blocks until the
 * device work it issued is finished. Calls on one handle must not overlap (the host keeps a
 * per-handle lock); calls on different handles are independent.
 *
 * Every entry point returns 0 on success and a negative SOUP_E* code on failure; the message of the
 * last failure on a handle is returned by soup_last_error(handle) (soup_last_error(NULL) gives the
 * message of the last failed soup_create on this thread). Messages keep the stems the reference
 * throws, so callers that match on them keep working.
 *
 * Array conventions at this boundary are the reference's: particle i is the i-th particle of the
 * creation composition; positions/velocities/forces are 4 floats per particle (x, y, z, w) with
 * pos.w = the species kind index as a float; bond slots are 3 u32 per particle, 0xFFFFFFFF = none.
 * Internally the library keeps particles in cell-sorted order and translates at this boundary.
 */
#ifndef SOUP_ABI_H
#define SOUP_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SOUP_NONE 0xFFFFFFFFu
#define SOUP_MAX_KINDS 8
#define SOUP_MAX_RULES 4
#define SOUP_CLASSES 4

enum {
  SOUP_OK = 0,
  SOUP_E_INVALID = -1,     /* bad argument / configuration */
  SOUP_E_DEVICE = -2,      /* HIP runtime failure */
  SOUP_E_NONFINITE = -3,   /* non-finite positions/velocities (soup/src/soup-health.ts:59-62) */
  SOUP_E_OVERFLOW = -4,    /* Verlet or DH list capacity exceeded (soup/src/soup-grid-verlet.ts:95-136) */
  SOUP_E_DRIFT = -5,       /* measured drift > skin/2 (soup/src/soup-grid-verlet.ts:125-136) */
  SOUP_E_NOMEM = -6        /* allocation refused (soup/src/sim.ts:296-314) */
};

/* Resolved, flat parameters. The host resolves CreateSoupOpts + data/soup.json + data/params.json
 * exactly as soup/src/sim.ts:96-278 does and fills this. f32 fields carry the values the reference
 * uploads as f32 uniforms (Params forces.wgsl:1-10, GridDims soup-buffers.ts:295-310, Species
 * pair.wgsl:61, AttrScale soup-attraction.ts:190-205, ES/ES2 electrostatics.ts:253-274,
 * VL soup-buffers.ts:231-232, BondParams bond-common.wgsl:54-81); f64 fields carry the doubles the
 * reference's CPU constant-pH Monte Carlo uses (soup/src/electrostatics.ts:435-496). */
typedef struct SoupConfig {
  uint32_t n;                 /* live particle count */
  uint32_t n_kinds;           /* monomer kinds (<= 8) */
  float box[3];               /* live box, sigma */
  uint32_t dims[3];           /* grid cells per axis = floor(box/cellSize) */
  uint32_t walk_radius;       /* GB.dims.w: +-cells walked by the list build */
  float list_range;           /* VL.x = interactionRange + skin */
  uint32_t list_cap;          /* VL.y = Verlet capacity per particle */
  uint32_t rebuild_every;     /* verletList.rebuildEvery */
  float skin;                 /* verletList.skin */
  /* Params (forces.wgsl:1-10), kT already overridden by CreateSoupOpts.kT (sim.ts:107) */
  float sigma, epsilon, k_fene, r_inf, k_bend, r_bend, wc, kT, gamma, dt;
  float b_tt;                 /* sigma * beadSizes.tail_tail: attraction rc = wca_cut(b_tt) */
  float wca_factor;           /* f32(2^(1/6)): wca_cut(b) = wca_factor * b */
  float radius[SOUP_MAX_KINDS];     /* Species.radius */
  uint32_t cls[SOUP_MAX_KINDS];     /* speciesClass: 0 apolar, 1 polar, 2 solvent, 3 mineral */
  float attr[SOUP_CLASSES][SOUP_CLASSES]; /* AttrScale.rows */
  float acid_soap;            /* AttrScale.pair.x */
  /* Debye-Hueckel, f32 as uploaded (ES, ES2) */
  uint32_t es_enabled;
  float es_A, es_kappa, es_rc, es_Frc;
  float es_split;             /* ES2.x splitRadius */
  float es_list_range;        /* ES2.y rc_es + skin */
  uint32_t es_cap;            /* ES2.z per-head capacity */
  uint32_t charged_kind;      /* ES2.w */
  uint32_t n_heads;           /* creation census of the charged kind (headIdx size) */
  /* constant-pH Monte Carlo, f64 (EsBasis) */
  double ph_A, ph_kappa, ph_cutoff, ph_shiftU, ph_shiftF;
  double ph_pKa, ph_pH, ph_kT, ph_q;   /* ph_q = chargeDeprotonated */
  double as_scale, as_rc, as_wc, as_eps; /* AcidSoapWell */
  uint32_t sweep_every;       /* electrostatics.sweepEverySteps */
  /* bond Monte Carlo (BondParams) */
  uint32_t n_rules;
  float kindA[SOUP_MAX_RULES], kindB[SOUP_MAX_RULES];
  float att_form[SOUP_MAX_RULES], att_break[SOUP_MAX_RULES];
  float acc_form[SOUP_MAX_RULES], acc_break[SOUP_MAX_RULES];
  float req_cat[SOUP_MAX_RULES], roleA[SOUP_MAX_RULES], roleB[SOUP_MAX_RULES];
  float catalyst_kind, head_terminal_only, head_chain_cap;
  float max_hold_steps;       /* adsorptionParams.x */
  float bond_interval_f;      /* adsorptionParams.y */
  uint32_t bond_interval;     /* bondAttemptInterval.steps */
  uint32_t capacity;          /* particles the buffers hold (the creation census; 0: n) */
  /* Verlet rebuild schedule. SOUP_REBUILD_CADENCE: every rebuild_every steps, as the reference
   * (soup/src/soup-integrate.ts:149), and the only mode the reference has. SOUP_REBUILD_DRIFT: the
   * measured drift maximum since the last rebuild (soup/wgsl/verlet.wgsl:115-123) is read back every
   * drift_check_every steps and the next step rebuilds once it exceeds drift_trigger (<= skin/2), so
   * a slow system rebuilds rarely and a fast one often; lists, and the bond Monte Carlo candidates
   * they feed, then differ from the reference's on the steps where the two schedules differ. */
  uint32_t rebuild_mode;
  uint32_t drift_check_every;
  float drift_trigger;
  /* The list build stages each cell's (2R + 1)^3-cell neighbourhood in LDS when it holds at most
   * this many beads (0: the library's limit, 4608 less the 256 entries of padding the sweep reads past them; larger values are capped at it) and gathers the
   * candidates from the particle arrays otherwise. Both ways give the same rows; a small value
   * forces the gather way (a launch-shape knob with no counterpart in the reference). */
  uint32_t list_stage;
  uint32_t reserved[2];
} SoupConfig;

enum { SOUP_REBUILD_CADENCE = 0, SOUP_REBUILD_DRIFT = 1 };

/* Mutable state at creation: a fresh lattice built by the host (soup/src/soup-init-state.ts:58-307)
 * or a decoded checkpoint (soup/src/checkpoint.ts:220-237). All arrays have n (or 3n/4n) entries and
 * are copied in. */
typedef struct SoupInitState {
  const float* positions;      /* 4n */
  const float* velocities;     /* 4n */
  const uint32_t* bond_slots;  /* 3n */
  const uint32_t* center_link; /* n */
  const uint32_t* center_held; /* n */
  const uint32_t* frozen;      /* n */
  const uint32_t* bond_rng;    /* n */
  const uint32_t* thermo_rng;  /* n */
  const float* charges;        /* n */
  const uint32_t* events;      /* 2 * n_rules */
  const uint32_t* desorb_events; /* 2 */
  uint64_t global_step;
  uint64_t next_sweep_at;      /* UINT64_MAX: never */
  uint32_t protonation_rng;
  uint32_t reserved;
} SoupInitState;

typedef struct SoupHandle SoupHandle;

/* Fields for soup_read / soup_write. Reads return the reference layout in particle-index order. */
enum {
  SOUP_FIELD_POS = 0,          /* f32 x4 */
  SOUP_FIELD_VEL = 1,          /* f32 x4 */
  SOUP_FIELD_FORCE = 2,        /* f32 x4: the force buffer as it stands (forcesNoRebuildDEBUG) */
  SOUP_FIELD_BOND_SLOTS = 3,   /* u32 x3 */
  SOUP_FIELD_CENTER_LINK = 4,  /* u32 */
  SOUP_FIELD_CENTER_HELD = 5,  /* u32 */
  SOUP_FIELD_FROZEN = 6,       /* u32 */
  SOUP_FIELD_CHARGE = 7,       /* f32 */
  SOUP_FIELD_BOND_RNG = 8,     /* u32 */
  SOUP_FIELD_THERMO_RNG = 9,   /* u32 */
  SOUP_FIELD_EVENTS = 10,      /* u32 x (2 * n_rules) */
  SOUP_FIELD_DESORB = 11,      /* u32 x2: stretch, timeout */
  SOUP_FIELD_VERLET_COUNT = 12,/* u32 per particle */
  SOUP_FIELD_VERLET_LIST = 13, /* u32 x list_cap per particle, entries are particle indices */
  SOUP_FIELD_ES_COUNT = 14,    /* u32 per head (head order = ascending particle index) */
  SOUP_FIELD_ES_LIST = 15,     /* u32 x es_cap per head, entries are particle indices */
  SOUP_FIELD_CELL_OF = 16,     /* u32 per particle: linear cell at the last rebuild */
  SOUP_FIELD_CELL_START = 17,  /* u32 x (ncells + 1) */
  SOUP_FIELD_CELL_IDX = 18,    /* u32 x n: particle indices in cell-sorted order */
  SOUP_FIELD_MAX_DRIFT_SQ = 19,/* f32 x1 */
  SOUP_FIELD_HEAD_IDX = 20,    /* u32 x n_heads */
  SOUP_FIELD_VERLET_NEAR = 21, /* u32 per particle: entries of the stored list row closer than the near/far
                                  split at the rebuild (no counterpart in the reference, whose row is one
                                  walk-ordered run; SOUP_FIELD_VERLET_LIST returns that run) */
  SOUP_FIELD_LIST_READS = 22,  /* u64 x4 since creation: rows the list force read as the near part only,
                                  rows it read whole (summed over its evaluations), evaluations in which
                                  every row qualified for its near part, and the others (no counterpart in
                                  the reference) */
  SOUP_FIELD_REBUILDS = 23     /* u64: Verlet rebuilds made by steps since creation, read-only (no counterpart
                                  in the reference) */
};

/* Modes for soup_forces (soup/src/soup-readback.ts:32-65). */
enum {
  SOUP_FORCES_REBUILD = 0,     /* full grid/list rebuild, then list force (+ DH far) -> forces() */
  SOUP_FORCES_BRUTE = 1,       /* O(N^2) force + O(N^2) DH far -> forcesBruteForce() */
  SOUP_FORCES_LIST = 2         /* list force with the current lists, no rebuild */
};

/* Single phases of one integration step, for parity tests (soup/src/soup-integrate.ts:83-111). */
enum {
  SOUP_PHASE_KICK_DRIFT = 1,   /* kick_drift_wrap_main */
  SOUP_PHASE_REBUILD = 2,      /* encodeVerletRebuild: grid, Verlet list, DH list, snapshot */
  SOUP_PHASE_FORCE = 4,        /* soup_force_list_main + es_force_far */
  SOUP_PHASE_DRIFT = 8,        /* soup_max_drift_main */
  SOUP_PHASE_BONDS = 16,       /* bond formation + break */
  SOUP_PHASE_KICK_THERMO = 32  /* kick_thermostat_main */
};

/* Constant-pH sweep statistics (soup/src/electrostatics.ts:317-326). */
typedef struct SoupSweepResult {
  uint32_t attempts, accepted, heads, deprotonated;
  double alpha, dEsMeanKT;
} SoupSweepResult;

/* Verlet / DH list occupancy (listOccupancyDEBUG, soup/src/soup-readback.ts:296-377). */
typedef struct SoupListStats {
  double verlet_mean; uint32_t verlet_max; uint32_t verlet_cap;
  double es_mean; uint32_t es_max; uint32_t es_cap;
  uint64_t verlet_pairs;       /* sum of verletCount */
  uint64_t es_pairs;           /* sum of esCount */
} SoupListStats;

/* createSoup (soup/src/sim.ts:95-600) below the host's planning: allocates device state, copies the
 * initial state in, and runs the priming rebuild + force pass (sim.ts:396-408). */
int soup_create(const SoupConfig* cfg, const SoupInitState* init, SoupHandle** out);
/* dispose (soup/src/soup-buffers.ts:557-596). */
int soup_destroy(SoupHandle* h);
/* Text of the last failure; never NULL. */
const char* soup_last_error(const SoupHandle* h);

/* SoupSystem.step(n) (soup/src/soup-integrate.ts:136-175): n integration steps in chunks of 1000,
 * with the non-finite and Verlet guards and the due constant-pH sweep at each chunk end. */
int soup_step(SoupHandle* h, uint32_t n);
/* n steps exactly as soup_step(h, n) advances them, with the state after every step copied out in
 * particle order (each array nullable): pos[s * 4N ...], bonds[s * 3N ...], links[s * N ...],
 * charges[s * N ...]; swept[s] = 1 where a constant-pH sweep ran after step s (its charges, and the
 * force the next step starts from, are then the post-sweep ones). The forward record of a
 * trajectory-segment adjoint. No counterpart in the reference. */
int soup_step_record(SoupHandle* h, uint32_t n, float* pos, uint32_t* bonds, uint32_t* links, float* charges,
                     uint8_t* swept);

/* setActiveCount (soup/src/soup-evaporate.ts:58-80) with the state it needs: replaces every
 * per-particle array with st's for n particles (particle order; n <= the capacity), keeps the step
 * counter, event counters and the constant-pH stream, and runs a rebuild and a force pass. Solvent
 * evaporation passes the first n particles of the current state, rehydration the current state
 * followed by the returning beads. Charged heads must keep their indices. */
int soup_reset_particles(SoupHandle* h, uint32_t n, const SoupInitState* st);

/* sampleInsertionPositions (soup/src/soup-evaporate-plan.ts:202-304), host code: `count` positions
 * in `box`, each the best of up to 64 uniform candidates by distance to everything already there
 * (the active particles, then the beads placed before it), stopping at the first candidate at least
 * min_sep away. Candidates come from a mulberry32 stream whose state *rng is advanced in place.
 * out: count x (x, y, z, 0) in f32; *short_of_floor: beads placed below min_sep; *min_achieved: the
 * smallest such distance (0 if none was placed). */
int soup_sample_insertion(const float* active, uint32_t n_active, const double box[3], uint32_t count, double min_sep,
                          uint32_t* rng, float* out, uint32_t* short_of_floor, double* min_achieved);

/* forceCandidateStatsDEBUG (soup/src/soup-readback.ts:246-263): the neighbour grid of the current
 * positions (in scratch memory: the state and its lists are not touched), then the cell walk of
 * soup_force_stats_main (soup/wgsl/step.wgsl:307-343). out[0] = candidates examined (j != i in the
 * +-walkRadius cells), out[1] = pairs within range (inside the pair's WCA cut, or inside the
 * attraction shell of an attracting class pair). 64-bit counts, where the reference's counters are
 * 32-bit atomics. */
int soup_force_candidate_stats(SoupHandle* h, uint64_t out[2]);

/* stepPhasesDEBUG (soup/src/soup-readback.ts:191-239): device milliseconds per iteration of five
 * groups, each run n times on the live state: ms[0] a full step with the rebuild and the bond
 * phase, ms[1] the rebuild (sort, DH list, Verlet list), ms[2] the list force, ms[3] the bond
 * Monte Carlo, ms[4] the two integration halves. The step counter does not move; the call ends with
 * a rebuild, a force pass and the guards, so the next step starts from consistent lists and F(x). */
int soup_step_phases(SoupHandle* h, uint32_t n, double ms[5]);

/* Replaces the interaction strengths of a live system (epsilon, the attraction table and acid-soap
 * depth, the DH coefficient, kappa and shift force, thermostat and bond-rule rates) with those of
 * *c; every geometric, range, radius and capacity field must equal the current one. The force the
 * next step starts from is recomputed under the new parameters with the current lists. For
 * parameter sweeps and finite-difference checks of the adjoint; no counterpart in the reference,
 * whose parameters are fixed at creation. */
int soup_set_params(SoupHandle* h, const SoupConfig* c);

/* `count` systems advanced by n steps together (the replica sweep): every system's steps of a
 * 1,000-step chunk are queued first, each system on its own stream, round robin by step period,
 * then each system's chunk-end guards and due constant-pH sweep run. Every system ends exactly as
 * soup_step(hs[k], n) would leave it. On failure *failed is the index of the system whose guard
 * failed (its soup_last_error holds the message). No counterpart in the reference, which steps one
 * system per page. */
int soup_step_batch(SoupHandle* const* hs, uint32_t count, uint32_t n, uint32_t* failed);
/* Global step counter (SoupSystem.steps). */
int soup_global_step(const SoupHandle* h, uint64_t* out);

/* Readbacks (soup/src/soup-readback.ts) and writes; bytes must equal the field size. */
int soup_read(SoupHandle* h, int field, void* dst, size_t bytes);
/* Writable: POS, VEL, CHARGE, BOND_SLOTS, CENTER_LINK, CENTER_HELD, BOND_RNG, THERMO_RNG, DESORB (the two
 * desorption counters, which desorbOverstretchedTethers adds to: soup/src/soup-evaporate.ts:316-325). */
int soup_write(SoupHandle* h, int field, const void* src, size_t bytes);
/* Size in bytes of a field (0 for an unknown field). */
size_t soup_field_bytes(const SoupHandle* h, int field);

/* forces() / forcesBruteForce() / list force (soup/src/soup-readback.ts:32-65); the result is in
 * SOUP_FIELD_FORCE afterwards. */
int soup_forces(SoupHandle* h, int mode);
/* One or more single phases, in step order, without advancing the step counter (test hook). */
int soup_run_phases(SoupHandle* h, uint32_t phase_mask, uint32_t do_bonds_flag);

/* protonationSweepDEBUG / maybeProtonationSweep (soup/src/soup-protonation.ts:52-77): sets the next
 * due step to now + sweepEverySteps, runs the serial constant-pH chain on the device, then the list
 * rebuild + force recompute. A no-op (zeroed result) when electrostatics is off. */
int soup_protonation_sweep(SoupHandle* h, SoupSweepResult* out);
/* Number of sweeps this handle ran and the statistics of the last one (electrostatics().sweeps/last). */
int soup_last_sweep(const SoupHandle* h, uint32_t* sweeps, SoupSweepResult* last);
/* protonationRngState. */
int soup_protonation_rng(const SoupHandle* h, uint32_t* out);

/* nonFiniteCount (soup/src/soup-health.ts:34-63): out[0] = positions, out[1] = velocities. */
int soup_nonfinite(SoupHandle* h, uint32_t out[2]);
/* assertVerletSafety (soup/src/soup-grid-verlet.ts:91-138) on demand. */
int soup_check_lists(SoupHandle* h);
/* List occupancy of the current lists. */
int soup_list_stats(SoupHandle* h, SoupListStats* out);

/* The doubles of makePotentialBasis (soup/src/soup-potential.ts:91-122): every number soupPotential
 * uses, computed by the host in FP64 from data/params.json and data/soup.json. */
typedef struct SoupPotentialBasis {
  double epsilon, fene_k, fene_r_inf, bend_k, bend_r0;
  double wc;                        /* attraction width */
  double rc_attr;                   /* wcaCut(sigma * beadSizes.tail_tail) */
  double cutoff;                    /* max(max pair wcaCut, rcAttr + wc): the pair cutoff (r2 <= cutoff^2) */
  double wca_factor;                /* 2^(1/6) in double */
  double acid_soap;                 /* acidSoapScaleOf */
  double b_radius[SOUP_MAX_KINDS];  /* sigma * radiusSigma per kind */
  double attr[4][4];                /* attractionScaleTable */
  double es_A, es_kappa, es_cutoff, es_shiftU, es_shiftF;  /* EsBasis (coeffA = 0: no DH term) */
  uint32_t cls[SOUP_MAX_KINDS];     /* speciesClasses */
  uint32_t charged_kind;
  uint32_t use_center_link;         /* soupPotential's centerLink argument given */
  uint32_t use_charges;             /* soupPotential's charges argument given */
  uint32_t reserved[5];
} SoupPotentialBasis;

/* soupPotential (soup/src/soup-potential.ts:145-301) of the handle's current configuration in FP64:
 * out = {nonbonded (incl. esTotalEnergy, electrostatics.ts:575-600), fene, bend, tether, total}.
 * Pairs come from the current Verlet and DH lists, which the drift guard keeps complete within
 * listRange - skin; the call fails if the basis cutoffs exceed that or the guard does not hold.
 * The terms are reduced in a fixed order (deterministic; equal to the reference's sum to FP64
 * rounding). */
int soup_potential(SoupHandle* h, const SoupPotentialBasis* basis, double out[5]);

/* The adjoint of soupPotential (soup/src/soup-potential.ts:145-301) at the handle's configuration,
 * in FP64, over the same pairs as soup_potential (current lists; the drift guard must hold):
 * - grad[3 N] = dU/dx per particle (particle order), which is -F of the step's force law;
 * - with a direction w[3 N] (particle order): hw[3 N] = H w, the Hessian of U times w, which is
 *   -(w . dF/dx), and dFw_dtheta = d(w . F)/dtheta;
 * - dU_dtheta = dU/dtheta.
 * theta, SOUP_ADJ_PARAMS entries: the attraction depth of the unordered class pairs (0,0), (0,1),
 * (0,2), (0,3), (1,1), (1,2), (1,3), (2,2), (2,3), (3,3) (basis->attr, both halves of the table
 * together), the acid-soap depth, epsilon, the head charge (chargeDeprotonated: every charge is 0 or
 * it) and the DH kappa at fixed coefficient and cutoff (the shifts follow kappa). w, hw and
 * dFw_dtheta may be NULL together, and dU_dtheta may be NULL. The gradient of the reference's
 * energy and forces with respect to positions and the parameter table, which the reference does
 * not compute (its soupPotential is an energy only); for sensitivities and force-matching. */
#define SOUP_ADJ_PARAMS 14
int soup_energy_adjoint(SoupHandle* h, const SoupPotentialBasis* basis, const double* w, double* grad, double* hw,
                        double* dU_dtheta, double* dFw_dtheta);

/* soup_energy_adjoint with a diagonal strain e = strain[0..2] added to the direction: every pair term
 * (list pairs, bonds, 1-3 bends, tethers, DH pairs) takes dw = w_i - w_j + diag(e) r_ij, r_ij the
 * minimum-image displacement x_i - x_j (w may be NULL: the strain alone). W_e receives the strain
 * virial sum over pair terms of (diag(e) r_ij) . f_ij, f_ij the term's force on i; hw then holds
 * H w + H_e, and dW_e/dx = -diag(e) grad - H_e, dW_e/dtheta is the strain's part of dFw_dtheta.
 * With e = (-1/2, -1/2, 1) / (Lx Ly), W_e is the configurational lateral tension of a membrane
 * normal to z, (W_zz - (W_xx + W_yy) / 2) / A. No counterpart in the reference (its area move
 * samples the area instead). */
int soup_energy_adjoint_strain(SoupHandle* h, const SoupPotentialBasis* basis, const double strain[3], const double* w,
                               double* grad, double* hw, double* dU_dtheta, double* dFw_dtheta, double* W_e);

/* Trajectory-segment adjoint (the sensitivity of a trajectory observable to the parameters, with
 * device checkpoints and a forward recompute per stretch). The system advances `steps` steps exactly
 * as soup_step(h, steps) would; a full checkpoint of the state is kept on the device before every
 * stretch of checkpoint_every steps. The reverse sweep then recomputes each stretch, last first, on
 * a replay system from its checkpoint (bit for bit, recording the stretch's states on the device)
 * and runs the pathwise adjoint of the velocity-Verlet Langevin step (soup/wgsl/step.wgsl:96-142;
 * the noise does not depend on the state) backwards through it: one force vector-Jacobian product
 * (k_energy_adjoint) per force evaluation, each at the bond slots, centre links and charges that
 * force was computed with. Bond and charge Monte Carlo moves are not differentiated.
 * objective SOUP_OBJ_END_ENERGY: J = soupPotential at the end state;
 * SOUP_OBJ_MEAN_STRAIN_VIRIAL: J = scale / steps * sum over the states after each step of W_e
 * (soup_energy_adjoint_strain's virial of the state as the step leaves it).
 * Out: J, dJ_dtheta[SOUP_ADJ_PARAMS] (theta as soup_energy_adjoint's), and (nullable) dJ_dx0 and
 * dJ_dv0 [3 N], the gradients with respect to the start positions and velocities (particle order).
 * Needs the cadence rebuild schedule, a start step and checkpoint_every that are multiples of
 * rebuildEvery, a start that does not follow a bond Monte Carlo step, and no frozen particles.
 * Device memory: a checkpoint (about 100 bytes per particle) per stretch and a record of
 * checkpoint_every + 1 states (40 bytes per particle each), freed on return. No counterpart in the
 * reference. */
enum { SOUP_OBJ_END_ENERGY = 0, SOUP_OBJ_MEAN_STRAIN_VIRIAL = 1 };
typedef struct SoupSegmentSpec {
  uint32_t steps;
  uint32_t checkpoint_every;
  uint32_t objective;
  uint32_t reserved;
  double strain[3];
  double scale;
} SoupSegmentSpec;
int soup_segment_adjoint(SoupHandle* h, const SoupPotentialBasis* basis, const SoupSegmentSpec* spec, double* J,
                         double* dJ_dtheta, double* dJ_dx0, double* dJ_dv0);

/* soupPotential (soup/src/soup-potential.ts:145-301) of an arbitrary configuration of this system:
 * positions (x, y, z, kind per particle, particle order, host memory) in the periodic box `box`,
 * with the system's current bond slots, centre links and charges. Replaces the CPU energy of the area
 * move's trials (soup/src/soup-area-move.ts:211-226): the pairs come from a cell grid of the
 * configuration itself (cells of at least basis->cutoff, the reference's deduplicated 27-cell walk),
 * the DH pairs of esTotalEnergy (electrostatics.ts:575-600) from all charged heads; FP64, fixed-order
 * sums. out: nonbonded (incl. DH), fene, bend, tether, total. */
int soup_potential_at(SoupHandle* h, const SoupPotentialBasis* basis, const float* positions, const double box[3],
                      double out[5]);

/* soup_potential_at without any device allocation: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under HIP graph capture, tight Monte Carlo loops, harnesses that
 * forbid hipMalloc). The configuration's cell grid, its positions and the partial sums live in one
 * fixed array per GPU (about 5.5 MiB, held until the process ends) instead of buffers the handle
 * allocates on first use; the result is soup_potential_at's, bit for bit. Limits: at most
 * SOUP_NOALLOC_MAX_N particles and SOUP_NOALLOC_MAX_CELLS grid cells (floor(box / cutoff) per axis),
 * else SOUP_E_INVALID with a message. Every call on one GPU uses that GPU's one array, so calls that
 * overlap on one GPU (from different threads, on any handles) race and must be serialised by the
 * caller. */
#define SOUP_NOALLOC_MAX_N 65536u
#define SOUP_NOALLOC_MAX_CELLS 262144u
int soup_potential_at_noalloc(SoupHandle* h, const SoupPotentialBasis* basis, const float* positions, const double box[3],
                              double out[5]);

/* The device half of a box change: adopts geometry->box, ->dims and ->walk_radius (the host's
 * planSoupGrid for the new box; every other field keeps its value), writes `positions` (particle
 * order, already mapped into the new box) unless it is NULL, reallocates the cell grid if it grew,
 * and rebuilds the lists.
 * - applyBoxScaleOnce (soup/src/soup-box-scale.ts:149-201): positions from
 *   soup_scale_molecules_rigid and flags SOUP_SET_BOX_FORCE, so the next step's first half-kick sees
 *   F(x) of the new geometry;
 * - growBoxTo (soup-box-scale.ts:318-373): positions NULL (nothing moves) and flags 0, a rebuild
 *   without a force pass as the reference's. */
enum { SOUP_SET_BOX_FORCE = 1 };
int soup_set_box(SoupHandle* h, const SoupConfig* geometry, const float* positions, uint32_t flags);

/* proximityPairs (soup/src/soup-box-scale-math.ts:162-215), host code: the (i, j), j > i, pairs of
 * eligible particles (eligible[i] != 0) within `cutoff` under the minimum image of `box`, found
 * through a cell list of floor(box / cutoff) cells per axis, in the reference's order (i ascending,
 * neighbour cells dx, dy, dz = -1..1, each cell's members ascending; an axis with fewer than 3 cells
 * visits its cells more than once, and those pairs repeat as they do in the reference). Writes at
 * most max_pairs pairs to pairs (2 entries each; NULL to count only) and the total to *count. */
int soup_proximity_pairs(const float* positions, uint32_t n, const double box[3], double cutoff, const uint8_t* eligible,
                         uint32_t* pairs, uint64_t max_pairs, uint64_t* count);

/* scaleMoleculesRigid (soup/src/soup-box-scale-math.ts:63-138), host code: every connected
 * component of the bond graph (`bond_pairs`, n_pairs (i, j) pairs as SoupSystem.bonds() lists them;
 * an unbonded particle is its own component) moves rigidly with its centre of mass scaled by
 * new_box/old_box per axis. out may not alias positions. No device work. */
int soup_scale_molecules_rigid(const float* positions, uint32_t n, const uint32_t* bond_pairs, uint32_t n_pairs,
                               const double old_box[3], const double new_box[3], float* out);

/* ---------------------------------------------------------------------------------------------
 * Snapshot analysis (soup/src/stages.ts:204-289's per-sample work): an analysis context holds one
 * particle snapshot in device memory -- copied from a live system or from host arrays -- and answers
 * the grid and graph questions of the reference's CPU analysis with the same results:
 * findAmphiphiles' components (soup/src/amphiphile.ts:55-141), clusterComponents' labels
 * (engine/src/aggregate.ts:65-212), encapsulatedWaterVolume's flood (soup/src/water-closure.ts:
 * 119-363) and localCavityVolume's open-boundary flood (soup/src/aggregates.ts:346-383,
 * engine/src/closure.ts:58-170). Each context has its own HIP stream; calls on one context must not
 * overlap.
 * ------------------------------------------------------------------------------------------- */
typedef struct SoupAnalysis SoupAnalysis;

/* Cell and water counts of one periodic encapsulated-water flood. */
typedef struct SoupWaterCounts {
  uint64_t total_empty_cells;    /* cells not within radius of a member */
  uint64_t unreached_cells;      /* empty cells the flood from the seed never reaches */
  uint64_t reached_empty_cells;  /* empty cells it reaches */
  uint32_t total_water;          /* beads of the water kind */
  uint32_t encapsulated;         /* water beads in empty, unreached cells */
  uint32_t reached_empty_water;  /* water beads in empty, reached cells */
  uint32_t seed_cell;            /* farthestEmptyCellSeed, x-fastest cell index */
  uint32_t dims[3];              /* ceil(box / cell) per axis */
  uint32_t reserved;
} SoupWaterCounts;

int soup_analysis_create(SoupAnalysis** out);
int soup_analysis_destroy(SoupAnalysis* a);
const char* soup_analysis_last_error(const SoupAnalysis* a);
/* The snapshot: n particles (positions x, y, z, kind as 4 floats; bond slots 3 u32 each) from host
 * memory, or the current state of a live system (particle-index order, device to device). */
int soup_analysis_load(SoupAnalysis* a, uint32_t n, const float* pos4, const uint32_t* bond_slots);
int soup_analysis_load_system(SoupAnalysis* a, SoupHandle* h);
/* findAmphiphiles (soup/src/amphiphile.ts:55-141): bit k of polar_mask = kind k is polar. Counts
 * the amphiphiles and their chain beads; soup_analysis_amphiphiles_read then returns, in the
 * reference's result order, each head, chain_offsets (n_amph + 1) into chains (tail 1 walked out
 * from the head, then tail 2) and tail_lengths (2 per amphiphile, 0 for an absent second tail). */
int soup_analysis_amphiphiles(SoupAnalysis* a, uint32_t polar_mask, uint32_t* n_amph, uint32_t* n_chain);
int soup_analysis_amphiphiles_read(SoupAnalysis* a, uint32_t* heads, uint32_t* chain_offsets, uint32_t* chains,
                                   uint32_t* tail_lengths);
/* clusterComponents' connectivity over members[0..m) (particle indices, in the caller's local
 * order): x and y periodic, z periodic when periodic_z (the soup's convention) or open (the membrane
 * convention of largestClusterCenter, positions then in [0, box z)); labels[k] = the smallest
 * local index of k's component. */
int soup_analysis_clusters(SoupAnalysis* a, const uint32_t* members, uint32_t m, const double box[3], double cutoff,
                           int periodic_z, uint32_t* labels);
/* encapsulatedWaterVolume's occupancy (members within radius, periodic), farthest empty seed from
 * centre, 6-connected periodic flood, and the water-bead counts of kind water_kind. */
int soup_analysis_water(SoupAnalysis* a, const uint32_t* members, uint32_t m, uint32_t water_kind, const double box[3],
                        const double centre[3], double cell, double radius, SoupWaterCounts* out);
/* occupancy + enclosedVolume over m given positions (x, y periodic when marking, z open; flood from
 * the six faces): the number of empty cells the flood never reaches. */
int soup_analysis_cavity(SoupAnalysis* a, const float* pos4, uint32_t m, const double box[3], double cell, double radius,
                         uint64_t* unreached_cells);

/* relaxIterations (soup/src/soup-relax.ts:87-139): normalised steepest descent with the cap
 * decaying linearly from d0 to 0, a rebuild before every iteration, and a final rebuild + force. */
int soup_relax(SoupHandle* h, uint32_t iterations, double d0);

/* Device-time profile per phase, measured with HIP events on the handle's own stream around the
 * launches of each phase. soup_profile(h, mask) resets the totals and brackets the phases whose bit
 * (1 << SOUP_PROF_x) is set; mask -1 brackets every phase, 0 stops. soup_profile_read fills
 * SOUP_PROF_PHASES accumulated milliseconds and launch counts. A mask of only the chunk-end phases
 * (SOUP_PROF_PROTON, SOUP_PROF_SWEEP) keeps step replay from captured graphs on; any per-step phase
 * turns it off while the profile runs. */
enum {
  SOUP_PROF_KICK_DRIFT = 0,
  SOUP_PROF_SORT = 1,          /* cell sort + state permutation of a rebuild */
  SOUP_PROF_ES_BUILD = 2,      /* DH head list build */
  SOUP_PROF_VERLET_BUILD = 3,  /* Verlet list build */
  SOUP_PROF_FORCE = 4,         /* list force kernel (list pairs + bonds; adds the heads' far DH sums) */
  SOUP_PROF_ES_FAR = 5,        /* far DH rows of the charged heads (head refresh + far kernel) */
  SOUP_PROF_DRIFT = 6,
  SOUP_PROF_BONDS = 7,         /* bond MC, all kernels */
  SOUP_PROF_KICK_THERMO = 8,
  SOUP_PROF_PROTON = 9,        /* constant-pH chain kernel */
  SOUP_PROF_SWEEP = 10,        /* a whole constant-pH sweep: DH list, pair tables, chain, rebuild and force */
  SOUP_PROF_PHASES = 16
};
int soup_profile(SoupHandle* h, int mask);
int soup_profile_read(SoupHandle* h, double* ms, uint64_t* count);

/* Whether soup_step replays its steps from captured HIP graphs (default on; off while a profile
 * runs): one graph per period of lcm(rebuildEvery, bondAttemptInterval) steps, so a small system
 * is not bound by launch overhead. Results are bit-identical either way. No counterpart in the
 * reference (WebGPU encodes its own command buffers); for tests and timing. */
int soup_use_graphs(SoupHandle* h, int enable);

/* Library build/version string. */
const char* soup_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SOUP_ABI_H */
