import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def ref_vectors():
    """Input/output vectors written from the reference functions (see tests/golden)."""
    with np.load(GOLDEN / "reference_vectors.npz") as data:
        return {k: data[k] for k in data.files}
