import sys
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE / "golden"))
sys.path.insert(0, str(HERE.parent / "fractal-audio-viz_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) HIP device")


@pytest.fixture(scope="session")
def oracle():
    import native
    if not native.ORACLE_SO.exists():
        pytest.fail("oracle/_build/libfav_oracle.so is not built: run __graft_entry__.build()")
    return native.load_oracle()


@pytest.fixture(scope="session")
def product():
    import native
    return native.load_product()


@pytest.fixture(scope="session")
def gpu_product(product):
    caps = product.fractal_backend_capabilities()
    assert caps & 8, f"no gfx950 device: {product.fractal_last_error()!r}"
    return product
