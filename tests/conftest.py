"""Shared fixtures.  `-m "not gpu"` runs on any machine (oracle, host logic,
C-ABI load); `-m gpu` tests need one MI355X and the built libmtsp2.so."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "modem-gpu_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and libmtsp2.so")


@pytest.fixture(scope="session")
def gpu_ctx():
    import mtsp2
    n = mtsp2.device_count()
    if n < 1:
        raise RuntimeError("gpu test selected but no HIP device is visible")
    ctx = mtsp2.Context(0)
    yield ctx
    ctx.close()


def small_problem(seed=0, nx=10, ny=9, nzAir=4, nzEarth=7, rho=(1.0, 1000.0)):
    """A small non-uniform grid with a random LOGE model (oracle-sized)."""
    rng = np.random.default_rng(seed)
    dx = rng.uniform(100, 300, nx)
    dy = rng.uniform(100, 300, ny)
    dz = np.concatenate([np.geomspace(30000, 100, nzAir), rng.uniform(50, 200, nzEarth)])
    lc = -np.log(rng.uniform(rho[0], rho[1], (nx, ny, nzEarth)))
    return dx, dy, dz, nzAir, lc


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
