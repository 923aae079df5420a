import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def golden_dir():
    return os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="session")
def leaf_sample(golden_dir):
    import numpy as np
    with np.load(os.path.join(golden_dir, "leaf_sample.npz")) as z:
        return {k: z[k] for k in z.files}


def tiles_to_image(tiles, cols):
    """Arrange (n, 4, 4, 4) tiles into an image of `cols` blocks per row (n % cols == 0)."""
    import numpy as np
    n = tiles.shape[0]
    rows = n // cols
    return np.ascontiguousarray(tiles.reshape(rows, cols, 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(rows * 4, cols * 4, 4))
