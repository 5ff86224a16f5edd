// Known-answer vectors of the initial-state and constant-pH random streams, computed in V8 (the
// reference's JavaScript runtime) from the published algorithms, written here from their
// definitions:
// - mulberry32 (T. Ettinger, 2017): a Weyl state a += 0x6D2B79F5, output mix
//   t = imul(a ^ a>>>15, a | 1); t = (t + imul(t ^ t>>>7, t | 61)) ^ t; (t ^ t>>>14) >>> 0, / 2^32;
// - Box-Muller (1958) on two such draws with u1 floored at 1e-9, as the reference's gaussian();
// - the PCG LCG step (multiplier 747796405, increment 2891336453) with the RXS-M-XS output
//   permutation (O'Neill 2014), / 2^32, as the reference's pcgNext;
// - Henderson-Hasselbalch 1 / (1 + 10^(pKa - pH)) with V8's Math.pow.
// Run: node tests/golden/make_rng_vectors.js > tests/golden/rng_vectors.json
'use strict';

function mulberry(seed) {
  var a = seed >>> 0;
  return function () {
    a = (a + 0x6d2b79f5) | 0;
    var t = Math.imul(a ^ (a >>> 15), a | 1);
    t = (t + Math.imul(t ^ (t >>> 7), t | 61)) ^ t;
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  };
}

function pcg(state) {
  var s = state >>> 0;
  return {
    next: function () {
      var v = (Math.imul(s, 747796405) + 2891336453) >>> 0;
      s = v;
      v = Math.imul((v >>> ((v >>> 28) + 4)) ^ v, 277803737) >>> 0;
      v = ((v >>> 22) ^ v) >>> 0;
      return v / 4294967296;
    },
    state: function () { return s; },
  };
}

var out = { generator: 'tests/golden/make_rng_vectors.js', node: process.version };

out.mulberry32 = [0, 1, 7, 19, 0xdeadbeef, 4294967295].map(function (seed) {
  var r = mulberry(seed);
  var u32 = [];
  for (var k = 0; k < 12; k++) u32.push(r() * 4294967296);
  return { seed: seed, u32: u32 };
});

// gaussian(): u1 = max(draw, 1e-9), u2 = draw, sqrt(-2 ln u1) cos(2 pi u2), over the first 64 pairs
out.gaussian = [19, 7].map(function (seed) {
  var r = mulberry(seed);
  var g = [];
  for (var k = 0; k < 64; k++) {
    var u1 = Math.max(r(), 1e-9);
    var u2 = r();
    g.push(Math.sqrt(-2 * Math.log(u1)) * Math.cos(2 * Math.PI * u2));
  }
  return { seed: seed, values: g };
});

// the Maxwell velocities of a lattice start: particle i's velocity components are s * gaussian()
// over draws (sites - 1) + 9 i + 3 .. + 8 of mulberry32(seed), s = sqrt(kT)
out.lattice_velocities = [{ seed: 19, n: 27, kT: 1.1 }, { seed: 7, n: 30, kT: 1.0 }].map(function (c) {
  var nx = Math.max(1, Math.ceil(Math.cbrt(c.n)));
  while (nx * nx * nx < c.n) nx++;
  var sites = nx * nx * nx;
  var r = mulberry(c.seed);
  for (var k = 0; k < sites - 1; k++) r();
  var s = Math.sqrt(c.kT);
  var v = [];
  for (var i = 0; i < c.n; i++) {
    r(); r(); r();
    for (var a = 0; a < 3; a++) {
      var u1 = Math.max(r(), 1e-9);
      var u2 = r();
      v.push(s * (Math.sqrt(-2 * Math.log(u1)) * Math.cos(2 * Math.PI * u2)));
    }
  }
  return { seed: c.seed, n: c.n, kT: c.kT, velocities: v };
});

// pcgNext from the initialCharges seed (seed ^ 0x9e3779b9) >>> 0
out.pcg_next = [19, 7, 11].map(function (seed) {
  var p = pcg((seed ^ 0x9e3779b9) >>> 0);
  var u32 = [];
  for (var k = 0; k < 16; k++) u32.push(p.next() * 4294967296);
  return { seed: seed, start: (seed ^ 0x9e3779b9) >>> 0, u32: u32, end_state: p.state() };
});

out.henderson_alpha = [[7.0, 4.9], [5.0, 4.9], [6.0, 4.9], [9.0, 4.9], [7.0, 7.0], [5.5, 4.9]].map(function (q) {
  return { pH: q[0], pKa: q[1], alpha: 1 / (1 + Math.pow(10, q[1] - q[0])) };
});

process.stdout.write(JSON.stringify(out) + '\n');
