"""BLA validity radius eps beyond the reference's 2^-38
(`fractal_hip_create_reference_tuned`): the HIP path against the C oracle
built with the same eps, on golden-case configurations (BLA at 1e25, the
1e100 Kalles grid, the endpoint-replay field) at 2^-53 and 2^-45, and the
default 2^-38 equal to the ordinary handle.

Bar: the HIP parity bar of every golden case (escape counts, NaN mask,
test1/test2 and every RenderStats counter bit-exact; float field within
one ulp; phase / DE within 1e-9 relative)."""
import pytest

import golden_io
import native

pytestmark = pytest.mark.gpu


def oracle_render(oracle, m, eps):
    ref = native.OracleReference(oracle, m["x"], m["y"], m["ref_zoom"], m["max_iter"], m["bits"],
                                 m["image_series_order"], m["options"].get("escape_radius_mode", 0),
                                 m["options"].get("coordinate_mode", 0), bla_epsilon_exponent=eps)
    opts = native.RenderOptions.defaults(**m["options"])
    field, planes, stats = ref.render(m["width"], m["height"], m["zoom"], m["max_iter"], m["series_order"],
                                      m["series_block"], opts, planes=True, stats=True)
    return {"field": field, "orbit_iteration": planes.orbit_iteration, "phase": planes.phase,
            "de_x": planes.de_x, "de_y": planes.de_y, "test1": planes.test1, "test2": planes.test2,
            "stats": stats}


def hip_render(lib, m, eps):
    opts = native.RenderOptions.defaults(backend=2, **m["options"])
    ref = native.AbiReference(lib, m["x"], m["y"], m["ref_zoom"], m["max_iter"], m["bits"], m["image_series_order"],
                              opts, bla_epsilon_exponent=eps)
    return ref.render(m["width"], m["height"], m["zoom"], m["max_iter"], m["series_order"], m["series_block"],
                      opts, planes=True, stats=True)


@pytest.mark.parametrize("name", ["bla_1e25", "c2_kalles", "replay_misiurewicz"])
@pytest.mark.parametrize("eps", [-53, -45])
def test_bla_epsilon_against_oracle(gpu_product, oracle, name, eps):
    print(f"[gpu] BLA eps 2^{eps}, {name}", flush=True)
    m = golden_io.load(name)["meta"]
    expected = oracle_render(oracle, m, eps)
    field, planes, stats = hip_render(gpu_product, m, eps)
    golden_io.assert_hip_parity(expected, field, planes, stats)


def test_default_epsilon_is_the_reference_table(gpu_product):
    print("[gpu] BLA eps 2^-38 = the ordinary handle", flush=True)
    g = golden_io.load("bla_1e25")
    field, planes, stats = hip_render(gpu_product, g["meta"], -38)
    golden_io.assert_hip_parity(g, field, planes, stats)
