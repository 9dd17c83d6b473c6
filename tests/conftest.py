import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP kernels of libcubvh_amd.so)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.load()
    return O


@pytest.fixture(scope="session")
def dodecahedron():
    """Mesh of the reference's own test fixture (test/unsigned_distance.py:9-76)."""
    z = np.load(GOLDEN / "dodecahedron.npz")
    return z["vertices"].astype(np.float32), z["faces"].astype(np.int32)


@pytest.fixture(scope="session")
def ico4():
    from cubvh_amd import synthetic as S

    return S.icosphere(4, 0.5)


@pytest.fixture(scope="session")
def gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def record():
    """record(name, **figures): prints the measured figures of a parity test and, when
    CUBVH_PARITY_LOG names a file, appends them to it as one JSON line."""
    import json
    import os

    path = os.environ.get("CUBVH_PARITY_LOG")

    def _record(name, **figures):
        line = json.dumps({"test": name, **{k: (float(v) if isinstance(v, (np.floating, np.integer)) else v)
                                            for k, v in figures.items()}})
        print(line, flush=True)
        if path:
            with open(path, "a") as fh:
                fh.write(line + "\n")

    return _record
