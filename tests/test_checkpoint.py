"""Checkpoint layout (soup/src/checkpoint.ts:110-237): byte format on the CPU, and the round trip
of tests/soup-checkpoint.test.ts:58-122 on the GPU."""
import json

import numpy as np
import pytest

import protocell_soup as ps
from protocell_soup.checkpoint import (decodeCheckpointResume, decodeFloat32, decodeUint32, encodeCheckpoint,
                                       encodeTyped, jsonStringify, readCheckpointFile, writeCheckpointFile)


def test_base64_is_the_typed_array_bytes():
    # IEEE-754 little-endian bytes 00 00 80 3F | 00 00 00 C0, as btoa over a Float32Array's buffer gives
    assert encodeTyped(np.array([1.0, -2.0], np.float32)) == "AACAPwAAAMA="
    assert encodeTyped(np.array([1, 0xFFFFFFFF], np.uint32)) == "AQAAAP////8="
    assert encodeTyped(np.array([1.0], dtype=">f4")) == "AACAPw=="
    x = np.random.default_rng(1).standard_normal(1001).astype(np.float32)
    assert np.array_equal(decodeFloat32(encodeTyped(x)), x)
    u = np.random.default_rng(2).integers(0, 2 ** 32, 777, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(decodeUint32(encodeTyped(u)), u)


def test_json_text_is_what_json_stringify_writes():
    assert jsonStringify({"box": [76.0, 76.0, 76.0], "kT": 1.1, "seed": 19, "x": [1e-7, 5e-5, 1e21, 0.1, -0.000001],
                          "n": None, "b": True, "s": "wet"}) == (
        '{"box":[76,76,76],"kT":1.1,"seed":19,"x":[1e-7,0.00005,1e+21,0.1,-0.000001],"n":null,"b":true,"s":"wet"}')
    assert jsonStringify([1.2345678901234568e+20, 2.5e-7, float("nan")]) == "[123456789012345680000,2.5e-7,null]"
    obj = {"a": [1.5, 2, {"c": "d"}], "e": 3.25}
    assert json.loads(jsonStringify(obj)) == obj


class _FakeSystem:
    """The CheckpointableSystem surface (checkpoint.ts:158-172) over plain arrays."""

    def __init__(self, n):
        r = np.random.default_rng(3)
        self.p = r.random(4 * n).astype(np.float32)
        self.v = r.standard_normal(4 * n).astype(np.float32)
        self.bs = r.integers(0, 2 ** 32, 3 * n, dtype=np.uint64).astype(np.uint32)
        self.cl = np.full(n, 0xFFFFFFFF, np.uint32)
        self.ch = r.integers(0, 50, n).astype(np.uint32)
        self.br = r.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
        self.tr = r.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
        self.q = -(r.random(n) < 0.3).astype(np.float32)
        self.steps, self.box, self.cyclePhase, self.cycleIndex = 1500, [18.0, 18.0, 18.0], "none", 0

    def particles(self): return self.p
    def velocities(self): return self.v
    def bondSlots(self): return self.bs
    def centerLinks(self): return self.cl
    def centerHeldSteps(self): return self.ch
    def desorbEvents(self): return {"stretch": 2, "timeout": 5}
    def charges(self): return self.q
    def protonationRngState(self): return 123456789
    def rngState(self): return {"bond": self.br, "thermo": self.tr}
    def events(self): return {"cc_bond": 3, "cc_break": 0, "co_bond": 7, "co_break": 1}
    def invariants(self): return {"monomers": {"C": 10, "O": 5}, "bonds": 4, "charge": -3.0}


def test_encode_decode_round_trip_through_a_file(tmp_path):
    sys_ = _FakeSystem(257)
    cfg = {"box": [18, 18, 18], "seed": 7, "kT": 1.1, "start": {"C": 120, "O": 40, "H": 120, "M": 10, "W": 0}}
    f = encodeCheckpoint(sys_, cfg)
    assert list(f) == ["version", "createdAt", "config", "N", "globalStep", "liveBox", "cyclePhase", "cycleIndex",
                       "activeCounts", "events", "desorbEvents", "positionsB64", "velocitiesB64", "bondSlotsB64",
                       "centerLinkB64", "centerHeldStepsB64", "bondRngB64", "thermoRngB64", "chargesB64",
                       "protonationRng"]
    assert f["N"] == 257 and f["globalStep"] == 1500 and f["version"] == 1
    path = writeCheckpointFile(str(tmp_path), "run", f)
    assert path.endswith("run-step1500.json")
    text = open(path).read()
    assert text.startswith('{"version":1,"createdAt":"') and '"liveBox":[18,18,18]' in text
    r = decodeCheckpointResume(readCheckpointFile(path))
    for key, want in (("positions", sys_.p), ("velocities", sys_.v), ("bondSlots", sys_.bs), ("centerLink", sys_.cl),
                      ("centerHeldSteps", sys_.ch), ("bondRng", sys_.br), ("thermoRng", sys_.tr), ("charges", sys_.q)):
        assert np.array_equal(r[key], want), key
    assert list(r["desorbEvents"]) == [2, 5] and r["events"]["co_bond"] == 7 and r["protonationRng"] == 123456789
    # a file without the optional charge state (pre-electrostatics checkpoints) still decodes
    g = dict(f)
    del g["chargesB64"]
    assert decodeCheckpointResume(g)["charges"] is None
    with pytest.raises(ValueError, match="unsupported version"):
        decodeCheckpointResume(dict(f, version=2))


@pytest.mark.gpu
def test_checkpoint_round_trip_reconstructs_every_field_and_the_continuation(tmp_path):
    cfg = {"box": [18, 18, 18], "seed": 7, "kT": 1.1, "clay": False, "start": {"C": 120, "O": 40, "H": 120, "M": 10, "W": 0}}
    a = ps.createSoup(cfg)
    b = None
    try:
        thermo0 = a.rngState()["thermo"].copy()
        a.step(1500)  # advanced RNG streams and bond bookkeeping, not the initial state
        f = encodeCheckpoint(a, cfg)
        path = writeCheckpointFile(str(tmp_path), "rt", f)
        b = ps.createSoup(dict(readCheckpointFile(path)["config"], resume=decodeCheckpointResume(readCheckpointFile(path))))
        assert f["globalStep"] == 1500 and b.steps == a.steps
        assert np.array_equal(a.particles(), b.particles())
        assert np.array_equal(a.velocities(), b.velocities())
        assert np.array_equal(a.bondSlots(), b.bondSlots())
        assert np.array_equal(a.centerLinks(), b.centerLinks())
        assert np.array_equal(a.centerHeldSteps(), b.centerHeldSteps())
        for k in ("bond", "thermo"):
            assert np.array_equal(a.rngState()[k], b.rngState()[k])
        assert a.desorbEvents() == b.desorbEvents() and a.events() == b.events() and a.box == b.box
        assert (a.rngState()["thermo"] != thermo0).all()
        # both continue on the same trajectory
        a.step(300)
        b.step(300)
        assert np.array_equal(a.bondSlots(), b.bondSlots())
        assert np.abs(a.particles() - b.particles()).max() < 1e-4
    finally:
        a.dispose()
        if b is not None:
            b.dispose()
