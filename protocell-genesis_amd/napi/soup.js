'use strict'
// The plain-JS host over the N-API addon (lib/soup_napi.node): the SoupSystem of
// soup/src/soup-types.ts:181-442 for the step path, over a resolved soup.
//
// A resolved soup is what createSoup's resolution (soup/src/sim.ts:96-278) produces: the packed
// SoupConfig bytes and the initial state's typed arrays. The reference's TypeScript host packs them
// from its own resolution; loadBundle reads the files protocell_soup.bundle.writeBundle writes.
// Every method is one or two addon calls; a failing call throws the library's Error, which keeps the
// reference's message stems.
const fs = require('fs')
const path = require('path')

const addon = require(path.join(__dirname, '..', 'lib', 'soup_napi.node'))

// SOUP_FIELD_* and SOUP_FORCES_* of include/soup_abi.h
const FIELD = Object.freeze({
  POS: 0, VEL: 1, FORCE: 2, BOND_SLOTS: 3, CENTER_LINK: 4, CENTER_HELD: 5, FROZEN: 6, CHARGE: 7,
  BOND_RNG: 8, THERMO_RNG: 9, EVENTS: 10, DESORB: 11, VERLET_COUNT: 12, VERLET_LIST: 13
})
const FORCES = Object.freeze({ REBUILD: 0, BRUTE: 1, LIST: 2 })

const TYPED = { f32: Float32Array, u32: Uint32Array }

// {config: ArrayBuffer, init: {...typed arrays, globalStep, nextSweepAt, protonationRng}, meta}
function loadBundle (dir) {
  const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'))
  const cfg = fs.readFileSync(path.join(dir, 'config.bin'))
  const config = cfg.buffer.slice(cfg.byteOffset, cfg.byteOffset + cfg.byteLength)
  const init = { globalStep: meta.globalStep, nextSweepAt: meta.nextSweepAt, protonationRng: meta.protonationRng }
  for (const [name, a] of Object.entries(meta.arrays)) {
    const buf = fs.readFileSync(path.join(dir, a.file))
    const T = TYPED[a.type]
    init[name] = new T(buf.buffer.slice(buf.byteOffset, buf.byteOffset + buf.byteLength))
  }
  return { config, init, meta }
}

// SoupSystem over a resolved soup
function createSoup (resolved) {
  const { config, init, meta } = resolved
  const h = addon.create(config, init)
  const N = meta.N
  let steps = meta.globalStep
  const read = (field, T, len) => addon.read(h, field, new T(len))
  const sys = {
    N,
    get steps () { return steps },
    async step (n) {
      await addon.step(h, n)
      steps = addon.globalStep(h)
    },
    particles: async () => read(FIELD.POS, Float32Array, 4 * N),
    velocities: async () => read(FIELD.VEL, Float32Array, 4 * N),
    async forces () {
      addon.forces(h, FORCES.REBUILD)
      return read(FIELD.FORCE, Float32Array, 4 * N)
    },
    forcesNoRebuildDEBUG: async () => read(FIELD.FORCE, Float32Array, 4 * N),
    bondSlots: async () => read(FIELD.BOND_SLOTS, Uint32Array, 3 * N),
    centerLinks: async () => read(FIELD.CENTER_LINK, Uint32Array, N),
    charges: async () => read(FIELD.CHARGE, Float32Array, N),
    rngState: async () => ({ bond: read(FIELD.BOND_RNG, Uint32Array, N), thermo: read(FIELD.THERMO_RNG, Uint32Array, N) }),
    // relaxColdStart (soup/src/sim.ts): only before the first step; the displacement cap in sigma
    async relaxColdStart (opts = {}) {
      const it = opts.iterations !== undefined ? opts.iterations : meta.coldStartRelax.iterations
      const dmax = opts.maxDisplacementSigma !== undefined ? opts.maxDisplacementSigma : meta.coldStartRelax.maxDisplacementSigma
      if (!Number.isInteger(it) || it < 1) throw new Error(`relaxColdStart: iterations=${it} must be an integer >= 1`)
      if (!(dmax > 0)) throw new Error(`relaxColdStart: maxDisplacementSigma=${dmax} must be positive`)
      if (steps !== 0) throw new Error(`relaxColdStart: the system is already at step ${steps} -- minimisation is allowed only before the first step, otherwise it lands inside the trajectory on which the measurements are taken`)
      addon.forces(h, FORCES.REBUILD)
      addon.relax(h, it, dmax * meta.sigma)
      addon.forces(h, FORCES.REBUILD)
      return { iterations: it, maxDisplacementStart: dmax * meta.sigma }
    },
    protonationSweepDEBUG: async () => addon.protonationSweep(h),
    dispose () { addon.destroy(h) }
  }
  return sys
}

module.exports = { addon, FIELD, FORCES, loadBundle, createSoup }
