"""CPU model of the staged Verlet build's row set-up: the box cull of each stencil row's x cells,
then the trim of the row's first and last cells, sorted by x, to the particle's x-chord of the
culling sphere. Every bead within the list range of a particle (exact f32 minimum-image distance)
must stay among the row's candidates, on random positions with beads planted on the cell faces,
in f32 arithmetic as the kernel does it. The kernel itself is checked against the oracle on the GPU;
this pins the geometry of the cull and the trim for grids the GPU tests do not build."""
import numpy as np
import pytest

f32 = np.float32


def _candidates(pos, cc, cs, dims, box, lr, R, t):
    """Kept staged beads of particle t: (indices before the trim, indices after the trim)."""
    w = box / dims.astype(f32)
    mg = f32(max(0.001, 4e-6 * float(box.max())))
    cull = f32((lr + mg) * (lr + mg))
    xi = pos[t]
    cx, cy, cz = cc[t]
    lo = (xi - np.array([cx, cy, cz], f32) * w).astype(f32)
    hi0 = f32(w[0] - lo[0])
    side = 2 * R + 1
    before, after = [], []
    for lane in range(side * side):
        dz, dy = lane // side - R, lane % side - R
        az = f32(dz * w[2]) if dz > 0 else (f32((-dz - 1) * w[2]) if dz < 0 else f32(0))
        sz = f32(-1 if dz > 0 else (1 if dz < 0 else 0))
        ay = f32(dy * w[1]) if dy > 0 else (f32((-dy - 1) * w[1]) if dy < 0 else f32(0))
        sy = f32(-1 if dy > 0 else (1 if dy < 0 else 0))
        gz = max(f32(az + sz * lo[2]), f32(0))
        gy = max(f32(ay + sy * lo[1]), f32(0))
        gyz = f32(gy * gy + gz * gz)
        if gyz > cull:
            continue
        dxl = dxh = 0
        for dx in range(1, R + 1):
            gh = max(f32(f32(dx - 1) * w[0] + hi0), f32(0))
            gl = max(f32(f32(dx - 1) * w[0] + lo[0]), f32(0))
            if f32(gh * gh + gyz) <= cull:
                dxh = dx
            if f32(gl * gl + gyz) <= cull:
                dxl = -dx
        h = f32(np.sqrt(f32(cull - gyz)))
        xlo, xhi = f32(xi[0] - h), f32(xi[0] + h)
        for dx in range(dxl, dxh + 1):
            X, Y, Z, sx = cx + dx, (cy + dy) % dims[1], (cz + dz) % dims[2], f32(0)
            if X < 0:
                X, sx = X + dims[0], -box[0]
            elif X >= dims[0]:
                X, sx = X - dims[0], box[0]
            c = X + dims[0] * (Y + dims[1] * Z)
            idx = np.arange(cs[c], cs[c + 1])
            xs = (pos[idx, 0] + sx).astype(f32)  # the staged x, image shift added
            assert np.all(np.diff(xs) >= 0), "a cell's staged x must ascend"
            keep = np.ones(len(idx), bool)
            if dx == dxl:
                keep &= xs >= xlo
            if dx == dxh:
                keep &= xs <= xhi
            before.extend(idx.tolist())
            after.extend(idx[keep].tolist())
    return before, after


@pytest.mark.parametrize("box,dims,lr,n,R,seed", [
    ((24.0, 24.0, 24.0), (7, 7, 7), 3.2, 12000, 1, 3),     # one-cell stencil rows, R = 1
    ((16.0, 16.0, 16.0), (5, 5, 5), 3.1, 3500, 2, 5),      # the stencil spans the whole grid in x
    ((30.4, 30.4, 30.4), (10, 10, 10), 4.447, 30000, 2, 7),  # the campaign's cell width and list range
])
def test_chord_trim_keeps_every_neighbour(box, dims, lr, n, R, seed):
    rng = np.random.default_rng(seed)
    box, dims, lr = np.array(box, f32), np.array(dims), f32(lr)
    pos = (rng.random((n, 3)) * box).astype(f32)
    k = n // 10  # a tenth of the beads on cell faces in x, where the trim bounds are tightest
    pos[:k, 0] = (rng.integers(0, dims[0], k) * (box[0] / dims[0])).astype(f32)
    x = (pos / box) * dims.astype(f32)
    cc = np.clip(np.floor(x), 0, dims - 1).astype(np.int64)
    cell = cc[:, 0] + dims[0] * (cc[:, 1] + dims[1] * cc[:, 2])
    order = np.lexsort((np.arange(n), pos[:, 0], cell))  # slots: by cell, then x, then index
    pos, cell, cc = pos[order], cell[order], cc[order]
    ncell = int(np.prod(dims))
    cs = np.zeros(ncell + 1, np.int64)
    np.add.at(cs, cell + 1, 1)
    cs = np.cumsum(cs)
    n_before = n_after = 0
    for t in rng.choice(n, 150, replace=False):
        before, after = _candidates(pos, cc, cs, dims, box, lr, R, t)
        n_before += len(before)
        n_after += len(after)
        d = (pos - pos[t]).astype(f32)
        d = (d - box * np.rint(d / box)).astype(f32)
        r = np.sqrt((d * d).sum(1, dtype=f32)).astype(f32)
        within = set(np.nonzero(r < lr)[0].tolist()) - {int(t)}
        assert within <= set(after), (int(t), sorted(within - set(after))[:5])
    assert n_after < n_before  # the trim removes candidates
