"""Shared fixtures.  Tests marked ``gpu`` need an MI355X and the built kernel
library; everything else runs on the CPU (oracle pins, host logic, ABI)."""
import os
import sys

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU and the built HIP kernel library")


@pytest.fixture(scope="session")
def march():
    import oracle.oracle as O
    O.build(force=False)
    return O.march()


@pytest.fixture(scope="session")
def dense64():
    """(oracle Grid, origins, dirs, gt, SceneData) of the 64^3 dense workload."""
    import oracle.oracle as O
    from svox2_amd import scenes
    sc, o, d, gt = scenes.workload_dense64(seed=0)
    g = host_grid(sc, torch.float32)
    return g, o.numpy(), d.numpy(), gt.numpy(), sc


def host_grid(scene, dtype=torch.float32, links=None):
    """Oracle Grid holding exactly the values a GPU grid of storage ``dtype``
    holds (float64 copies of the rounded values)."""
    import oracle.oracle as O
    N = np.array(scene.links.shape, np.float64)
    dens = scene.density.to(dtype).double().cpu().numpy().reshape(-1)
    sh = scene.sh.to(dtype).double().cpu().numpy()
    lk = scene.links.cpu().numpy() if links is None else links
    return O.Grid(np.ascontiguousarray(lk, np.int32), dens, sh, scene.basis_dim,
                  offset=0.5 * N - 0.5, scaling=0.5 * N)


@pytest.fixture(scope="session")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from svox2_amd import _C
    _C.load_library()  # raises if the kernel library is missing
    return torch.device("cuda:0")
