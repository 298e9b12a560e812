import json
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.oracle import REF_SO, Reference
    if not REF_SO.exists():
        pytest.skip("oracle/_ref is not built (needs the reference sources)")
    return Reference()


def load_golden(name: str) -> dict:
    with np.load(GOLDEN / name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_spec(g: dict) -> list:
    return [tuple(t) for t in json.loads(str(g["spec"]))]


@pytest.fixture(scope="session")
def gpu():
    """The GPU library, initialised on device 0."""
    from tangerine_amd import _lib
    lib = _lib.load()
    _lib.check(lib.tg_device_init(0), "tg_device_init")
    return lib


def grid_points(origin, step, cells) -> np.ndarray:
    nx, ny, nz = (int(c) + 1 for c in cells)
    x = np.float32(origin[0]) + np.arange(nx, dtype=np.float32) * np.float32(step[0])
    y = np.float32(origin[1]) + np.arange(ny, dtype=np.float32) * np.float32(step[1])
    z = np.float32(origin[2]) + np.arange(nz, dtype=np.float32) * np.float32(step[2])
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return np.stack([X, Y, Z], axis=-1).reshape(-1, 3)


def canonical_faces(faces) -> np.ndarray:
    f = np.asarray(faces, dtype=np.uint64).reshape(-1, 3)
    if f.size == 0:
        return f
    k = np.argmin(f, axis=1)
    rolled = np.stack([f[np.arange(len(f)), (k + i) % 3] for i in range(3)], axis=1)
    order = np.lexsort((rolled[:, 2], rolled[:, 1], rolled[:, 0]))
    return rolled[order]


def sha(a) -> str:
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
