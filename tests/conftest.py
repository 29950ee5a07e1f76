import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


_GPU = None


def gpu_present() -> bool:
    """True when libcfdla can open a HIP device."""
    global _GPU
    if _GPU is None:
        from cfdrs_amd import _native as N

        _GPU = N.lib().cfdla_init(0) == N.OK
    return _GPU


def pytest_collection_modifyitems(config, items):
    gpu_items = [it for it in items if "gpu" in it.keywords]
    if not gpu_items:
        return
    if not gpu_present():
        skip = pytest.mark.skip(reason="no HIP device visible to libcfdla")
        for it in gpu_items:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def known():
    with open(os.path.join(GOLDEN, "reference_known_answers.json")) as f:
        return json.load(f)


def tridiag(n, d=2.0, o=-1.0):
    a = np.zeros((n, n))
    for i in range(n):
        a[i, i] = d
        if i > 0:
            a[i, i - 1] = o
        if i + 1 < n:
            a[i, i + 1] = o
    return a


def grid_laplacian(n):
    """n x n grid Laplacian, 4 on the diagonal, -1 to grid neighbours (row-major)."""
    m = n * n
    a = np.zeros((m, m))
    for i in range(n):
        for j in range(n):
            p = i * n + j
            a[p, p] = 4.0
            for (q, r) in ((i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)):
                if 0 <= q < n and 0 <= r < n:
                    a[p, q * n + r] = -1.0
    return a


def dense_to_csr(d):
    rp, ci, v = [0], [], []
    for row in d:
        nz = np.nonzero(row)[0]
        ci.extend(nz.tolist())
        v.extend(row[nz].tolist())
        rp.append(len(ci))
    return np.array(rp, np.int64), np.array(ci, np.int64), np.array(v, np.float64)
