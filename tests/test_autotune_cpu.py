"""Shape spaces of the autotuner (no GPU)."""
from lya2pcf_amd import autotune


def test_full_space_covers_tiles_and_layouts():
    full = list(autotune.candidates("mixed", full=True))
    # forest-1 tiles of 64, 128 and 256 pixels per wave
    assert {c["px_per_lane"] for c in full} == {1, 2, 4}
    # LDS layouts: unpadded and bank-padded rows, one or two replicas
    assert {(c["row_pad"], c["replicas"]) for c in full} == {(0, 1), (1, 1), (0, 2), (1, 2)}
    # two replicas are built for 8-wave workgroups only
    assert all(c["waves"] == 8 for c in full if c["replicas"] == 2)
    # forest-1 tails walked against forest 2 in lanes: off, or up to 32, 48, 63 pixels
    assert {c["tail_max"] for c in full} == {-1, 32, 48, 63}
    assert len(full) == (3 * 2 * 3 * 3 * 2 + 3 * 2 * 3 * 2) * 4


def test_default_space_inside_full_space():
    key = lambda c: tuple(sorted(c.items()))
    for mode in ("mixed", "fp64"):
        full = {key(c) for c in autotune.candidates(mode, full=True)}
        default = [key(c) for c in autotune.candidates(mode)]
        assert set(default) <= full
    walks = [c for c in autotune.candidates("fp64") if c.get("walk")]
    assert {c["waves"] for c in walks} == {8, 16}
    assert not any(c.get("walk") for c in autotune.candidates("mixed", full=True))
