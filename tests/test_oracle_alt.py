"""Alternate deep formulas (Julia, burning ship, tricorn): the C oracle
against golden vectors written by the reference's renderer.cpp, and
closed-form pins of the formula-specific perturbation."""
import numpy as np
import pytest

import golden_io
import native
from alt_cases import CASES


@pytest.mark.parametrize("name", [c["name"] for c in CASES])
def test_oracle_alt_matches_reference_golden(oracle, name):
    g = golden_io.load(name)
    field, stats = golden_io.render_alt_oracle(oracle, g["meta"])
    np.testing.assert_array_equal(field.reshape(-1).view(np.uint32), g["field"].reshape(-1).view(np.uint32))
    assert int(stats[0]) == int(g["stats"][0])


def test_oracle_alt_deep_matches_direct_iteration(oracle):
    """At zoom 1e6 plain doubles still resolve the view: every pixel's
    integer escape index from the perturbation path equals a direct
    double-precision iteration of the same formula (shallow pins)."""
    for formula, julia in ((2, ("0", "0")), (3, ("0", "0")), (1, ("-0.8", "0.156"))):
        x, y = ("-1.7497", "-0.0286") if formula == 2 else (("-0.3", "0.62") if formula == 3 else ("0.29", "0.02"))
        case = dict(name="pin", formula=formula, x=x, y=y, ref_zoom="1e6", zoom="1e6", max_iter=3000, bits=128,
                    width=24, height=14, julia=list(julia), clone_zoom=None, options={})
        field, _ = golden_io.render_alt_oracle(oracle, case)
        direct = dict(width=24, height=14, zoom="1e6", x=x, y=y, max_iter=3000, formula=formula,
                      julia_re=float(julia[0]), julia_im=float(julia[1]), planes=True, options={})
        dfield, planes = native.render_direct_oracle(oracle, direct)
        deep_index = np.floor(field.reshape(-1))
        direct_index = planes.orbit_iteration.astype(np.float64) + 1
        direct_index[planes.orbit_iteration >= 3000] = 3000
        # Julia's deep path counts the initial sample as iteration 0
        if formula == 1:
            direct_index = direct_index
        agree = np.mean(np.abs(deep_index - direct_index) <= 1)
        assert agree > 0.97, (formula, agree)
