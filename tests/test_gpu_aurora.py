"""The HIP ordinary-palette colouriser and crop / atlas compositors against
the fixtures the reference's renderer.cpp wrote (every byte and float
equal) and against the C oracle on a 1920x1080 frame."""
import numpy as np
import pytest

import aurora_io
from aurora_cases import CASES
from test_aurora_cpu import OracleAurora

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip(gpu_product):
    return aurora_io.bind(gpu_product)


def assert_equal(out, expected):
    if out.dtype == np.float32:
        np.testing.assert_array_equal(out.view(np.uint32), expected.view(np.uint32))
    else:
        np.testing.assert_array_equal(out, expected)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_hip_matches_reference(hip, case):
    print(f"compositor case {case['name']}", flush=True)
    status, out = aurora_io.run(hip, case)
    assert status == 0, hip.fractal_last_error()
    assert_equal(out, aurora_io.load(case["name"])["output"])


@pytest.mark.parametrize("name", ["aurora_atlas_colourise_accents", "aurora_atlas_field_ex_centred",
                                  "aurora_atlas_rgb", "aurora_crop_colourise"])
def test_hip_frame_matches_oracle(hip, oracle, name):
    """A 1920x1080 output frame from 960x540 tiles, same controls."""
    case = dict(next(c for c in CASES if c["name"] == name))
    for key in ("pw", "cw", "sw"):
        if key in case and case[key]:
            case[key] = 960
    for key in ("ph", "ch", "sh"):
        if key in case and case[key]:
            case[key] = 540
    case["ow"], case["oh"] = 1920, 1080
    status, out = aurora_io.run(hip, case)
    assert status == 0, hip.fractal_last_error()
    expected_status, expected = aurora_io.run(OracleAurora(oracle), case)
    assert expected_status == 0
    assert_equal(out, expected)
