import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs through the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    """One mupsi context on cuda:0 (GPU tests only)."""
    import torch  # noqa: F401  (initialises the HIP runtime used for tensors)
    from mupsi_amd import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def dev():
    import torch
    return torch.device("cuda:0")


def kernels():
    """Kernel parameter sets of the benchmark workloads plus edge cases."""
    return [
        {"sigma": 1.0, "length_scale": 1.0, "aniso": [1, 1, 1], "points_per_cell": 3},   # sphere_single_128
        {"sigma": 1.5, "length_scale": 2.0, "aniso": [1, 1, 4], "points_per_cell": 3},   # sphere_renewal_256, guided_sphere_quad_1024
        {"sigma": 0.3, "length_scale": 0.5, "aniso": [1, 1, 1], "points_per_cell": 4},   # icosphere_renewal_512
        {"sigma": 2.7, "length_scale": 0.37, "aniso": [1, 1, 1], "points_per_cell": 8},  # generated_batch range
        {"sigma": 0.1, "length_scale": 4.6, "aniso": [1, 1, 1], "points_per_cell": 1},   # generated_batch range
    ]


def random_points(rng: np.random.Generator, n: int, scale: float) -> np.ndarray:
    p = rng.uniform(-scale, scale, (n, 3)).astype(np.float32)
    p[: n // 8] *= 1e-3  # around the double-width cell 0
    return p
