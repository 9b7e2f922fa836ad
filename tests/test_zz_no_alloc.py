"""reina_rt_intersect_noalloc (claim counters in the library's static device
array, no allocation) against reina_rt_intersect on the same rays, through the
C ABI: hits bit for bit, on two calls in a row (the second smaller, so the
counters the first call left behind must be cleared), on the icosphere with
FP32 and 8-bit nodes. Also: the size limit is refused with a clear error and a
zero-ray call is a no-op."""
import ctypes

import numpy as np
import pytest

import reina_rt
from reina_rt import abi, scenes

pytestmark = pytest.mark.gpu


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, abi.RAY)
    rays["ox"], rays["oy"], rays["oz"] = rng.uniform(-3.0, 3.0, (n, 3)).astype(np.float32).T
    rays["dx"], rays["dy"], rays["dz"] = rng.normal(size=(n, 3)).astype(np.float32).T
    rays["tmax"] = 10000.0
    return rays


def _both(torch, dev, rays):
    lib = reina_rt.lib()
    st = torch.cuda.current_stream().cuda_stream
    r = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    ha = torch.full((len(rays), 4), -1.0, dtype=torch.float32, device="cuda")
    hb = torch.full((len(rays), 4), -2.0, dtype=torch.float32, device="cuda")
    reina_rt.check(lib.reina_rt_intersect(dev.handle, ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(ha.data_ptr()),
                                          len(rays), None, ctypes.c_void_p(st)))
    reina_rt.check(lib.reina_rt_intersect_noalloc(dev.handle, ctypes.c_void_p(r.data_ptr()),
                                                  ctypes.c_void_p(hb.data_ptr()), len(rays), ctypes.c_void_p(st)))
    torch.cuda.synchronize()
    return ha.cpu().numpy(), hb.cpu().numpy()


@pytest.mark.parametrize("flags", [0, abi.NODES_QUANTIZED])
def test_noalloc_intersect_equals_intersect(cuda, icosphere_arrays, flags):
    torch = cuda
    arrays, _ = icosphere_arrays
    dev = reina_rt.DeviceScene(arrays, flags=flags)
    try:
        for n, seed in ((300000, 1), (4099, 2)):
            print("noalloc intersect: flags %d, %d rays, seed %d" % (flags, n, seed), flush=True)
            a, b = _both(torch, dev, _rays(n, seed))
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (n, int((a != b).any(1).sum()))
            assert (a.view(np.uint32)[:, 1] != 0xFFFFFFFF).mean() > 0.05    # the rays do hit the sphere
    finally:
        dev.close()


def test_noalloc_intersect_limits(cuda, cornell_arrays):
    torch = cuda
    arrays, _ = cornell_arrays
    dev = reina_rt.DeviceScene(arrays)
    lib = reina_rt.lib()
    try:
        h = torch.zeros((1, 4), dtype=torch.float32, device="cuda")
        r = torch.zeros((1, 8), dtype=torch.float32, device="cuda")
        rc = lib.reina_rt_intersect_noalloc(dev.handle, ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(h.data_ptr()),
                                            (1 << 30) + 1, None)
        assert rc == abi.ERR_INVALID_ARG and b"at most" in lib.reina_rt_last_error()
        assert lib.reina_rt_intersect_noalloc(dev.handle, None, None, 0, None) == 0
        a, b = _both(torch, dev, _rays(1000, 3))
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    finally:
        dev.close()
