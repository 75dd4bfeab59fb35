import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def pkg():
    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as o
    o.load()
    return o


@pytest.fixture(scope="session")
def synth(pkg):
    from traditional_dic_amd import synth as s
    return s


@pytest.fixture(scope="session")
def ctx(pkg):
    """A device context; GPU tests fail loudly (no fallback) if HIP is unusable."""
    c = pkg.Context(0)
    yield c
    c.close()


def read_gray(u8: np.ndarray) -> np.ndarray:
    """workflows/subset_2d.py:21-27: float32 / max, in float32."""
    a = u8.astype(np.float32)
    a /= np.float32(a.max())
    return a


@pytest.fixture(scope="session")
def stereo():
    return np.load(os.path.join(GOLDEN, "stereo_plate.npz"), allow_pickle=False)


def gradient_init(rows: np.ndarray, step: int = 4):
    """Initial affine terms from the golden u, v by finite differences on the
    step-4 propagation grid (central where both neighbours exist)."""
    key = {(int(x), int(y)): i for i, (x, y) in enumerate(rows[:, :2])}
    g = np.zeros((rows.shape[0], 4))
    for i in range(rows.shape[0]):
        x, y = int(rows[i, 0]), int(rows[i, 1])
        for c, comp in enumerate((2, 3)):
            for d, (dx, dy) in enumerate(((step, 0), (0, step))):
                ip, im = key.get((x + dx, y + dy)), key.get((x - dx, y - dy))
                if ip is not None and im is not None:
                    val = (rows[ip, comp] - rows[im, comp]) / (2 * step)
                elif ip is not None:
                    val = (rows[ip, comp] - rows[i, comp]) / step
                elif im is not None:
                    val = (rows[i, comp] - rows[im, comp]) / step
                else:
                    val = 0.0
                g[i, 2 * c + d] = val
    return g  # du_dx, du_dy, dv_dx, dv_dy
