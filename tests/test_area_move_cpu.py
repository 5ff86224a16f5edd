"""The area move's host pieces (no GPU): the rigid-molecule map (soup_scale_molecules_rigid, host
code in the library), the proposal geometry and the molecule count, pinned by the reference's own
test (tests/soup-area-move.test.ts:35-88) and by what the map must preserve."""
import math

import numpy as np

from protocell_soup.area_move import moleculeCount, proposeBox, scaleBeadsAffine, scaleMoleculesRigid


def _dist(p, b, i, j):
    d = p.reshape(-1, 4)[i, :3].astype(np.float64) - p.reshape(-1, 4)[j, :3].astype(np.float64)
    d -= np.floor(d / np.asarray(b) + 0.5) * np.asarray(b)
    return float(np.sqrt((d * d).sum()))


def test_rigid_map_reference_case():
    # tests/soup-area-move.test.ts:35-88: a 3-bead chain plus three unbonded solvent beads
    box = [10.0, 10.0, 10.0]
    pos = np.array([1, 1, 1, 0, 2, 1, 1, 1, 3, 1, 1, 1, 9.5, 0.2, 5, 4, 4, 7, 2, 4, 0.1, 9.8, 8, 4], np.float32)
    bonds = np.array([0, 1, 1, 2], np.uint32)
    assert moleculeCount(bonds, 6) == 4
    u = 0.03
    mid = proposeBox(box, u / 2, "lateral-fixed-volume")
    end = proposeBox(box, u, "lateral-fixed-volume")
    composed = scaleMoleculesRigid(scaleMoleculesRigid(pos, bonds, box, mid), bonds, mid, end)
    direct = scaleMoleculesRigid(pos, bonds, box, end)
    assert np.abs(composed - direct).max() < 1e-4
    for i, j in ((0, 1), (1, 2)):
        assert abs(_dist(direct, end, i, j) - _dist(pos, box, i, j)) < 1e-5
    for i in (3, 4, 5):
        for a in range(3):
            e = pos[4 * i + a] * (end[a] / box[a])
            assert abs(direct[4 * i + a] - (e - math.floor(e / end[a]) * end[a])) < 1e-4
    for mode in ("lateral-fixed-volume", "lateral-fixed-z"):
        b = proposeBox(box, u, mode)
        assert abs(math.log((b[0] * b[1]) / (box[0] * box[1])) - u) < 1e-12
    fv = proposeBox(box, u, "lateral-fixed-volume")
    assert abs(fv[0] * fv[1] * fv[2] / 1000.0 - 1.0) < 1e-12
    assert proposeBox(box, u, "lateral-fixed-z")[2] == box[2]
    # the kind column is carried unchanged
    assert np.array_equal(direct.reshape(-1, 4)[:, 3], pos.reshape(-1, 4)[:, 3])


def test_rigid_map_long_wrapped_chain_and_affine_singletons():
    # a 30-bead chain (bond length 0.97) that winds through a box of 8 more than once: its ends are far
    # more than half a box apart, the case the hop-by-hop offsets exist for (soup-box-scale-math.ts:
    # 50-61); every bond keeps its length, and the chain's centre moves by the box ratio
    rng = np.random.default_rng(4)
    box, new = [8.0, 8.0, 8.0], [8.4, 7.7, 8.1]
    steps = rng.normal(size=(29, 3))
    steps = 0.97 * steps / np.linalg.norm(steps, axis=1, keepdims=True)
    path = np.vstack([[1.0, 2.0, 3.0], 1.0 + np.cumsum(steps, axis=0) + [0, 1, 2]])
    path[:, 0] += np.linspace(0, 20, 30)  # wind along x
    n_chain = 30
    singles = rng.random((5, 3)) * 8.0
    xyz = np.vstack([path, singles])
    wrapped = xyz - np.floor(xyz / 8.0) * 8.0
    pos = np.zeros((35, 4), np.float32)
    pos[:, :3] = wrapped
    pairs = np.array([[k, k + 1] for k in range(n_chain - 1)], np.uint32).reshape(-1)
    assert moleculeCount(pairs, 35) == 6
    out = scaleMoleculesRigid(pos.reshape(-1), pairs, box, new).reshape(-1, 4)
    for k in range(n_chain - 1):
        assert abs(_dist(out, new, k, k + 1) - _dist(pos, box, k, k + 1)) < 2e-5
    assert np.all(out[:, :3] >= 0) and np.all(out[:, :3] <= np.asarray(new, np.float32))
    # unbonded particles are size-1 molecules: exactly the affine map of scaleBeadsAffine
    aff = scaleBeadsAffine(pos.reshape(-1), box, new).reshape(-1, 4)
    assert np.abs(out[n_chain:, :3] - aff[n_chain:, :3]).max() < 1e-5
    # the chain's unwrapped centre of mass scales by the box ratio (mod the new box)
    com_old = path.mean(axis=0)
    unw = [out[0, :3].astype(np.float64)]
    for k in range(1, n_chain):
        d = out[k, :3].astype(np.float64) - out[k - 1, :3]
        d -= np.floor(d / np.asarray(new) + 0.5) * np.asarray(new)
        unw.append(unw[-1] + d)
    com_new = np.mean(unw, axis=0)
    expect = (com_old - np.floor(com_old / 8.0) * 8.0) * (np.asarray(new) / 8.0)
    diff = com_new - expect
    diff -= np.round(diff / np.asarray(new)) * np.asarray(new)
    assert np.abs(diff).max() < 1e-4


def test_molecule_count_closed_forms():
    assert moleculeCount(np.zeros(0, np.uint32), 7) == 7
    assert moleculeCount(np.array([0, 1, 2, 3, 1, 2], np.uint32), 5) == 2  # chain 0-1-2-3 and bead 4
    assert moleculeCount(np.array([0, 1, 1, 2, 2, 0], np.uint32), 3) == 1  # a ring
