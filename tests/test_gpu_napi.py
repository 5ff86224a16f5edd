"""The Node host on the GPU: a soup created, relaxed and stepped through the N-API addon
(napi/soup.js over lib/soup_napi.node) ends bit for bit where the same soup run through the ctypes
binding ends: positions, velocities, bonds, charges and the rebuilt force."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import protocell_soup as ps
from protocell_soup.bundle import writeBundle
from soup_fixtures import charged

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_node_host_matches_the_python_host(tmp_path):
    if shutil.which("node") is None:
        pytest.skip("no node on this machine")
    opts = charged(box=20.0, seed=6, pH=6.0)
    bundle, out = tmp_path / "bundle", tmp_path / "out"
    bundle.mkdir()
    out.mkdir()
    writeBundle(opts, str(bundle))
    K = 40
    print("NAPI node run", flush=True)
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "napi", "run_soup.js"), str(bundle), str(K), str(out)],
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
    s = ps.createSoup(opts)
    try:
        s.relaxColdStart()
        s.step(K)
        assert np.array_equal(np.fromfile(out / "positions.f32", np.float32), s.particles())
        assert np.array_equal(np.fromfile(out / "velocities.f32", np.float32), s.velocities())
        assert np.array_equal(np.fromfile(out / "bondSlots.u32", np.uint32), s.bondSlots())
        assert np.array_equal(np.fromfile(out / "charges.f32", np.float32), s.charges())
        assert np.array_equal(np.fromfile(out / "forces.f32", np.float32), s.forces())
    finally:
        s.dispose()
