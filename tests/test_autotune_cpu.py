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


def test_measured_counters_of_the_committed_profiles():
    # the bench line copies these figures; they must match the profile summaries' own
    # arithmetic (VALU issue = VALU instructions / (256 CUs x cycles per XCD))
    import json
    import os
    for mode in ("mixed", "fp64"):
        c = autotune.measured_counters(mode)
        assert c is not None and set(c) == {"valu_issue", "lds_bank_conflict_per_cu_cycle",
                                            "lds_wait_per_wave_cycle", "valu_per_evaluated_wave_step",
                                            "cycles_per_xcd", "source"}
        with open(os.path.join(autotune.PROFILES, "counters_xi_%s_sky100k.json" % mode)) as f:
            raw = json.load(f)["raw"]
        cyc = raw["GRBM_GUI_ACTIVE"] / 8
        assert abs(c["cycles_per_xcd"] - cyc) < 1e-6 * cyc
        assert abs(c["valu_issue"] - raw["SQ_INSTS_VALU"] / (256 * cyc)) < 1e-9
        assert abs(c["lds_wait_per_wave_cycle"] - raw["SQ_WAIT_INST_LDS"] / raw["SQ_WAVE_CYCLES"]) < 1e-9
        assert 0.5 < c["valu_issue"] < 1.01
    assert autotune.measured_counters("no-such-mode") is None
