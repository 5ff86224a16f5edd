'use strict'
// The N-API addon without a GPU: its exports, the SoupConfig size it was built against, and the
// errors soup_create raises before it touches a device, which must reach JS as Errors carrying the
// reference's message stems (soup/src/sim.ts:148-150).
// Usage: node check_addon.js <bundle dir> <offset of SoupConfig.n> <offset of SoupConfig.capacity>
const assert = require('assert')
const path = require('path')
const host = require(path.join(__dirname, '..', '..', 'protocell-genesis_amd', 'napi', 'soup.js'))

const [dir, offN, offCap] = [process.argv[2], Number(process.argv[3]), Number(process.argv[4])]
const addon = host.addon
const want = ['version', 'configBytes', 'create', 'destroy', 'step', 'stepSync', 'read', 'write', 'fieldBytes', 'forces',
  'relax', 'globalStep', 'protonationSweep']
assert.deepStrictEqual(Object.keys(addon).sort(), want.slice().sort())
for (const k of want) assert.strictEqual(typeof addon[k], 'function', k)
assert.ok(/^libsoup_hip /.test(addon.version()), addon.version())

const b = host.loadBundle(dir)
assert.strictEqual(addon.configBytes(), b.config.byteLength)
assert.strictEqual(addon.configBytes(), b.meta.configBytes)

function expectThrow (config, init, stem, code) {
  let err = null
  try {
    addon.create(config, init)
  } catch (e) {
    err = e
  }
  assert.ok(err instanceof Error, `create did not throw (${stem})`)
  assert.ok(err.message.startsWith(stem), `message ${JSON.stringify(err.message)} lacks ${JSON.stringify(stem)}`)
  if (code !== undefined) assert.strictEqual(err.code, code)
  return err
}

// an empty starting composition
const empty = b.config.slice(0)
new DataView(empty).setUint32(offN, 0, true)
expectThrow(empty, b.init, 'createSoup: the starting composition is empty', -1)
// a live count above the creation census
const over = b.config.slice(0)
new DataView(over).setUint32(offCap, b.meta.N - 1, true)
const e = expectThrow(over, b.init, `createSoup: the live composition (${b.meta.N}) is larger than the starting one (${b.meta.N - 1})`, -1)
assert.ok(e.message.endsWith('-- buffers are allocated for the starting one'), e.message)
// typed arrays shorter than the config's particle count, and a config of the wrong size
const short = Object.assign({}, b.init, { positions: b.init.positions.subarray(0, 8) })
expectThrow(b.config, short, 'soup_napi: init.positions holds 32 bytes')
expectThrow(b.config.slice(0, 16), b.init, 'soup_napi: the config holds 16 bytes')
// a handle that is not one
assert.throws(() => addon.stepSync({}, 1), /soup_napi: not a live soup handle/)
console.log('OK')
