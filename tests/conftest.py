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


# Gradient tolerances against the FP64 oracle.  Float32 atomics accumulate in
# arbitrary order, so results are not bit-reproducible:
#   GRAD_MAX_RTOL   max|g - g64| <= GRAD_MAX_RTOL * max|g64| over the tensor;
#   GRAD_ELEM_RTOL  |g - g64| <= GRAD_ELEM_RTOL * |g64| for every element with
#                   |g64| > GRAD_ELEM_FLOOR * max|g64|, so that a mistake
#                   confined to rows with small gradients (an edge corner's
#                   weight, a dropped term) is not hidden by the largest ones.
# Measured GPU errors are about 1e-6 (max-normalised); the per-element figure
# is bound by float32 cancellation in dsigma = ws (total T - accum).
GRAD_MAX_RTOL = 2e-5
GRAD_ELEM_RTOL = 1e-2
GRAD_ELEM_FLOOR = 1e-3


def grad_errors(g, g64):
    """(max-normalised error, worst per-element relative error over elements
    above GRAD_ELEM_FLOOR of the max) of a gradient against the FP64 oracle."""
    g = np.asarray(g, np.float64).reshape(-1)
    g64 = np.asarray(g64, np.float64).reshape(-1)
    mx = max(float(np.abs(g64).max()), 1e-30)
    err = np.abs(g - g64)
    big = np.abs(g64) > GRAD_ELEM_FLOOR * mx
    elem = float((err[big] / np.abs(g64[big])).max()) if big.any() else 0.0
    return float(err.max() / mx), elem


def assert_grad_close(g, g64, name="grad", max_rtol=GRAD_MAX_RTOL, elem_rtol=GRAD_ELEM_RTOL):
    e_max, e_elem = grad_errors(g, g64)
    print(f"{name}: max-normalised {e_max:.3e}, per-element {e_elem:.3e}")
    assert e_max <= max_rtol, f"{name}: max-normalised error {e_max:.3e} > {max_rtol}"
    assert e_elem <= elem_rtol, f"{name}: per-element error {e_elem:.3e} > {elem_rtol}"


@pytest.fixture(scope="session")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from svox2_amd import _C
    _C.load_library()  # raises if the kernel library is missing
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def sparse256(gpu):
    """(SceneData, accelerated GPU grid, its links, origins, dirs) of the
    256^3 blob grid (about 8 % occupied) with 4096 camera rays; the grid's
    thresholds are sigma 1e-8, stop 1e-7."""
    from svox2_amd import scenes
    sc, cams, gen = scenes.workload_blobs(256, scenes.SPARSE256, seed=0, n_cameras=20,
                                          device=gpu)
    grid = scenes.to_sparse_grid(sc, device=gpu)
    grid.opt.sigma_thresh, grid.opt.stop_thresh = 1e-8, 1e-7
    links_acc = grid.links.cpu().numpy()
    o, d = cams.random_batch(4096, gen)
    return sc, grid, links_acc, o, d
