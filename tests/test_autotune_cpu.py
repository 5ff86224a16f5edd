"""The launch-shape autotuner's host side (no GPU): the variant grid around the current shapes, and
the committed tuning file the build compiles in (only the tunable macros, only allowed values)."""
from protocell_soup.autotune import DEFAULTS, TUNABLES, _name, current, variants
from protocell_soup.build import tuning_defines


def test_variant_grid_changes_one_macro_at_a_time():
    base = dict(DEFAULTS)
    vs = variants(base)
    assert vs[0] == base and len(vs) == 1 + sum(len(c) - 1 for c, _ in TUNABLES.values())
    assert len({_name(v) for v in vs}) == len(vs)
    for v in vs[1:]:
        assert sum(v[k] != base[k] for k in base) == 1


def test_committed_tuning_is_a_valid_shape():
    d = tuning_defines()
    assert set(d) <= set(TUNABLES)
    for k, v in d.items():
        assert v in TUNABLES[k][0], (k, v)
    assert current() == dict(DEFAULTS, **d)


def test_drift_check_default_comes_from_the_tuning_file(tmp_path):
    import json

    from protocell_soup.autotune import DRIFT_CHECK_CANDIDATES, tuned_drift_check_every

    p = tmp_path / "tuning.json"
    assert tuned_drift_check_every(str(p)) == 5  # nothing measured
    p.write_text(json.dumps({"defines": {}, "rebuild": {"driftCheckEvery": 10}}))
    assert tuned_drift_check_every(str(p)) == 10
    p.write_text(json.dumps({"defines": {}, "rebuild": {"driftCheckEvery": 7}}))  # not a candidate
    assert 7 not in DRIFT_CHECK_CANDIDATES and tuned_drift_check_every(str(p)) == 5


def test_tuned_drift_interval_shrinks_where_it_leaves_no_margin(monkeypatch):
    """The tuned interval (20 at kT 1, dt 0.01, skin 1.5: trigger 0.057) leaves no margin at kT 1.2;
    the resolution then takes the largest interval whose trigger stays positive, and an interval the
    caller names is refused instead of shrunk."""
    import math

    import pytest

    from protocell_soup import autotune
    from protocell_soup.sim import resolveRebuildSchedule

    monkeypatch.setattr(autotune, "tuned_drift_check_every", lambda *a, **k: 20)
    soup = {"verletList": {"skin": 1.5}}
    p = {"integrator": {"dt": 0.01}}
    r = resolveRebuildSchedule(soup, p, 1.0, {"rebuild": "drift"})
    assert r["checkEvery"] == 20 and r["trigger"] == pytest.approx(0.75 - 20 * 0.01 * 2 * math.sqrt(3.0))
    hot = resolveRebuildSchedule(soup, p, 1.2, {"rebuild": "drift"})
    vb = 2 * math.sqrt(3.6)
    assert hot["checkEvery"] == int(0.75 / (0.01 * vb)) == 19 and 0 < hot["trigger"] < 0.75
    with pytest.raises(ValueError, match="driftCheckEvery=20 the default margin leaves none"):
        resolveRebuildSchedule(soup, p, 1.2, {"rebuild": "drift", "driftCheckEvery": 20})
