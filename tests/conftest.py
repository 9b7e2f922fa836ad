import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "reina-vk_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def world_triangles(arrays):
    """(n, 9) float32 world-space triangles, instance-major, computed with the
    same operation order as the library's flattening (((c0 x + c1 y) + c2 z) + c3)."""
    out = []
    for inst in arrays["instances"]:
        T = inst["transform"].reshape(4, 4).T.astype(np.float32)
        m = arrays["models"][inst["modelID"]]
        idx = arrays["indices"][m["indexOffset"]:m["indexOffset"] + 3 * m["indexCount"]]
        v = arrays["vertices"][idx][:, :3].astype(np.float32)
        c0, c1, c2, c3 = T[:3, 0], T[:3, 1], T[:3, 2], T[:3, 3]
        w = ((c0[None] * v[:, 0:1] + c1[None] * v[:, 1:2]) + c2[None] * v[:, 2:3]) + c3[None]
        out.append(w.astype(np.float32).reshape(-1, 9))
    return np.concatenate(out) if out else np.zeros((0, 9), np.float32)


@pytest.fixture(scope="session")
def cornell_arrays():
    from reina_rt import scenes
    s, cam = scenes.cornell()
    return s.build_arrays(), cam


@pytest.fixture(scope="session")
def icosphere_arrays():
    from reina_rt import scenes
    s, cam = scenes.icosphere()
    return s.build_arrays(), cam


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch
