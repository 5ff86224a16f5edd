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
