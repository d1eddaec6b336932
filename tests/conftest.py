import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
# The reference's own test images stored under tests/golden/images/ (their
# harness digests and quality figures are in real_images.json).
REAL_IMAGES = sorted(k for k in json.loads((GOLDEN / "real_images.json").read_text()) if k.endswith(".png"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP kernels run)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.oracle import Reference, have_reference
    if not have_reference():
        pytest.skip("oracle/_ref (reference compiled from its own source) not built")
    return Reference()


@pytest.fixture(scope="session")
def harness():
    from oracle.oracle import Harness, have_harness
    if not have_harness():
        pytest.skip("oracle/_ref/libgoofy_harness.so (reference harness compiled from its own source) not built")
    return Harness()


@pytest.fixture(scope="session")
def real_images():
    return json.loads((GOLDEN / "real_images.json").read_text())


@pytest.fixture(scope="session")
def golden_blocks():
    z = np.load(GOLDEN / "blocks.npz")
    return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden_img512():
    z = np.load(GOLDEN / "img512.npz")
    return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def digests():
    return json.loads((GOLDEN / "digests.json").read_text())


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a visible GPU (device memory through goofy_amd.hip)."""
    import goofy_amd
    from goofy_amd import hip
    assert hip.device_count() > 0, "GPU test run without a visible GPU"
    assert goofy_amd.device_count() > 0
    return goofy_amd
