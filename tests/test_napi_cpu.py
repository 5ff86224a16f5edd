"""The Node binding (lib/soup_napi.node, napi/soup.js) without a GPU: node loads the addon, which
exports every entry point of the host, was built against this SoupConfig layout, and turns the
errors soup_create raises before touching a device into JS Errors carrying the reference's
message stems (soup/src/sim.ts:148-150)."""
import os
import shutil
import subprocess

import pytest

from protocell_soup import native
from protocell_soup.bundle import writeBundle
from soup_fixtures import BROTH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "protocell-genesis_amd", "lib", "soup_napi.node")


def test_node_host_loads_and_reports_errors(tmp_path):
    if shutil.which("node") is None or not os.path.exists("/usr/include/node/node_api.h"):
        pytest.skip("no Node toolchain in this image")
    assert os.path.exists(ADDON), "lib/soup_napi.node is not built: run __graft_entry__.build()"
    meta = writeBundle(BROTH, str(tmp_path))
    assert meta["N"] == 1220 and meta["configBytes"] == 584
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "napi", "check_addon.js"), str(tmp_path),
                        str(native.SoupConfig.n.offset), str(native.SoupConfig.capacity.offset)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
