"""soup_potential_at_noalloc against soup_potential_at, through the C-ABI.

The no-alloc entry point keeps the configuration's cell grid, positions and partial sums in one
fixed device array per GPU instead of buffers the handle allocates. Both run the same kernels with
the same fixed-order sums, so the five energies must be equal bit for bit. Two calls in a row on two
systems, the second smaller, check that what the first call left in the shared array does not leak
into the second (the cell counts are cleared on every call). Sizes past the array's limits are refused
with a message.
"""
import ctypes as C

import numpy as np
import pytest

import protocell_soup as ps
from protocell_soup import native
from soup_fixtures import charged

pytestmark = pytest.mark.gpu


def _call(lib, s, name, pos, box, basis):
    out = (C.c_double * 5)()
    bx = (C.c_double * 3)(*[float(x) for x in box])
    p = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
    code = getattr(lib, name)(s._h, C.byref(basis), p.ctypes.data_as(C.c_void_p), bx, out)
    return code, np.array(out[:], dtype=np.float64)


def _trial(s, scale, seed):
    """A trial configuration of the area move's kind: positions scaled about the origin, jittered."""
    rng = np.random.default_rng(seed)
    p = s.particles().reshape(-1, 4).copy()
    box = np.asarray(s.box, np.float64) * scale
    p[:, :3] = np.mod(p[:, :3].astype(np.float64) * scale + rng.normal(0.0, 0.01, (p.shape[0], 3)), box).astype(np.float32)
    return p.reshape(-1), box


def test_noalloc_potential_equals_the_allocating_entry_point():
    lib = native.load_library()
    big = ps.createSoup(charged(box=22.0, seed=4, pH=6.0, start={"C": 1400, "O": 900, "H": 600, "M": 40, "W": 3500}))
    small = ps.createSoup(charged(box=16.0, seed=9, pH=5.0, start={"C": 500, "O": 300, "H": 200, "M": 10, "W": 1200}))
    try:
        big.step(60)
        small.step(40)
        assert small.N < big.N <= native.NOALLOC_MAX_N
        for s, scale, seed in ((big, 1.03, 1), (small, 0.98, 2), (big, 1.0, 3)):
            basis = s.potentialBasis()
            pos, box = _trial(s, scale, seed)
            c0, ref = _call(lib, s, "soup_potential_at", pos, box, basis)
            c1, got = _call(lib, s, "soup_potential_at_noalloc", pos, box, basis)
            assert c0 == 0 and c1 == 0, (lib.soup_last_error(s._h), c0, c1)
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (got, ref)
            assert np.isfinite(ref).all()
        # a grid past the array's cell limit is refused, and the system stays usable
        basis = big.potentialBasis()
        pos, _ = _trial(big, 1.0, 4)
        code, _ = _call(lib, big, "soup_potential_at_noalloc", pos, [200.0, 200.0, 200.0], basis)
        assert code == native.SOUP_E_INVALID
        assert b"exceeds the fixed scratch array" in lib.soup_last_error(big._h)
        c1, got = _call(lib, big, "soup_potential_at_noalloc", pos, big.box, basis)
        c0, ref = _call(lib, big, "soup_potential_at", pos, big.box, basis)
        assert c0 == 0 and c1 == 0 and np.array_equal(got, ref)
    finally:
        big.dispose()
        small.dispose()


def test_noalloc_refuses_more_particles_than_the_array_holds():
    lib = native.load_library()
    # 71,942 beads of water at 0.8: more than SOUP_NOALLOC_MAX_N = 65,536
    s = ps.createSoup({"box": [44.8, 44.8, 44.8], "seed": 3, "kT": 1.0, "clay": False,
                       "start": {"C": 0, "O": 0, "H": 0, "M": 0, "W": 71_942}})
    try:
        assert s.N > native.NOALLOC_MAX_N
        basis = s.potentialBasis()
        code, _ = _call(lib, s, "soup_potential_at_noalloc", s.particles(), s.box, basis)
        assert code == native.SOUP_E_INVALID
        assert b"particles exceed the fixed scratch array's 65536" in lib.soup_last_error(s._h)
        code, ref = _call(lib, s, "soup_potential_at", s.particles(), s.box, basis)
        assert code == 0 and np.isfinite(ref).all()
    finally:
        s.dispose()
