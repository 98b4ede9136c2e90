"""The C oracle against golden vectors written by the reference's own
renderer.cpp (tests/golden/make_golden.py): bit-exact field, planes and
RenderStats counters for every case."""
import numpy as np
import pytest

import golden_io
import native


@pytest.mark.parametrize("name", golden_io.case_names())
def test_oracle_matches_reference_golden(oracle, name):
    golden = golden_io.load(name)
    field, planes, stats = golden_io.render_oracle(oracle, golden["meta"])
    golden_io.assert_bit_exact(golden, field, planes, stats)


def test_golden_cases_cover_the_paths():
    """The fixture set exercises every branch family of the deep loop."""
    totals = {}
    for name in golden_io.case_names():
        s = golden_io.load(name)["stats"]
        for i, key in enumerate(("pixels", "logical", "bla", "linear", "cubic", "exact", "replay",
                                 "retries", "cycle", "tail_pixels")):
            totals[key] = totals.get(key, 0) + int(s[i])
        totals["tail_rebases"] = totals.get("tail_rebases", 0) + int(s[13])
        totals["series_pixels"] = totals.get("series_pixels", 0) + int(s[16])
        totals["glitch"] = totals.get("glitch", 0) + int(s[18])
        totals["unresolved"] = totals.get("unresolved", 0) + int(s[19])
    for key in ("linear", "cubic", "exact", "replay", "tail_pixels", "tail_rebases", "series_pixels",
                "glitch", "unresolved"):
        assert totals[key] > 0, key
    assert int(golden_io.load("replay_misiurewicz")["stats"][6]) > 0


@pytest.mark.skipif(not native.REFERENCE_SO.exists(), reason="reference build absent")
@pytest.mark.parametrize("name", ["c1_series", "bla_1e25", "c2_kalles"])
def test_golden_regenerates_from_reference(name):
    """The fixture still equals what the compiled reference renders now."""
    lib = native.load_reference()
    golden = golden_io.load(name)
    field, planes, stats = golden_io.render_abi(lib, golden["meta"], backend=0)
    golden_io.assert_bit_exact(golden, field, planes, stats)
