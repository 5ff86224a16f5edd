'use strict'
// Creates a soup from a bundle through the Node host, relaxes it, steps it and writes the state.
// Usage: node run_soup.js <bundle dir> <steps> <out dir>
const fs = require('fs')
const path = require('path')
const host = require(path.join(__dirname, '..', '..', 'protocell-genesis_amd', 'napi', 'soup.js'))

async function main () {
  const [dir, steps, out] = [process.argv[2], Number(process.argv[3]), process.argv[4]]
  const sys = host.createSoup(host.loadBundle(dir))
  try {
    await sys.relaxColdStart()
    // a call made while the step runs on its worker thread is refused, and the step still completes
    const running = sys.step(steps)
    let refused = ''
    try { await sys.particles() } catch (e) { refused = e.message }
    await running
    if (!/a step is still running on this system/.test(refused)) throw new Error(`an overlapping call was not refused: '${refused}'`)
    const write = (name, a) => fs.writeFileSync(path.join(out, name), Buffer.from(a.buffer, a.byteOffset, a.byteLength))
    write('positions.f32', await sys.particles())
    write('velocities.f32', await sys.velocities())
    write('bondSlots.u32', await sys.bondSlots())
    write('charges.f32', await sys.charges())
    const f = await sys.forces()
    write('forces.f32', f)
    fs.writeFileSync(path.join(out, 'steps.json'), JSON.stringify({ steps: sys.steps }))
  } finally {
    sys.dispose()
  }
  console.log('OK')
}
main().catch(e => { console.error(e.stack || e); process.exit(1) })
