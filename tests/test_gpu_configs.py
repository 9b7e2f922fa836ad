"""The five benchmark configurations at their full sizes, through the C ABI.

C1 Cornell box 256x256, 16 spp, depth 8: GPU = oracle bit for bit, with and
   without next-event estimation.
C3 4096 instances x 2048-triangle sphere = 8,388,608 triangles (device-built
   BVH8): closest hits of camera and bounce rays against the FP64 oracle
   (>= 99.99% hit ids, every disagreement an FP64-detected tie, t within the
   FP32 error model); a 1920x1080, 32 spp, depth 12 frame over 8-bit nodes
   bit-identical to the FP32-node frame; a small frame of the dielectric /
   metal / Lambertian mix = oracle bit for bit.
C4 64 scenes of 1e3-2e6 triangles (log-uniform), 64 requests of uneven
   resolution and spp in one server: every image bit-identical to its
   reina_rt_trace image.
C5 32 resident scenes of 1e4-4e6 triangles under Poisson arrivals with a
   bounded FIFO: every completed request (a sample of them) bit-identical to
   the same request rendered alone.
"""
import numpy as np
import pytest

import oracle
import reina_rt
from reina_rt import abi, loadgen, scenes
from reina_rt.scene import make_desc, push_consts

from test_gpu_parity import camera_rays, fp32_t_error, gpu_intersect, gpu_render, secondary_rays

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _same(a, b):
    return bool(((_bits(a) == _bits(b)) | (np.isnan(a) & np.isnan(b))).all())


@pytest.mark.parametrize("nee", [False, True], ids=["bsdf", "nee"])
def test_c1_cornell_256_16spp_vs_oracle(cuda, cornell_arrays, nee):
    torch = cuda
    arrays, _ = cornell_arrays
    W = H = 256
    pc = scenes.cornell_push(W, H, spp=16, max_bounces=8, sample_batch=1)
    dev = reina_rt.DeviceScene(arrays)
    orc = oracle.OracleScene(arrays, make_desc)
    if nee:
        dev.set_trace_options(abi.TRACE_NEE)
        orc.set_nee(True)
    print("C1: cornell %dx%d 16 spp depth 8 nee=%s" % (W, H, nee), flush=True)
    got = gpu_render(torch, dev, pc, W, H)
    ref, _ = orc.render(pc, W, H, use_bvh=False, nthreads=16)
    assert _same(got, ref)
    assert np.isfinite(got[..., :3]).mean() > 0.99 and got[..., :3].mean() > 0.05
    dev.close()


@pytest.fixture(scope="module")
def c3(cuda):
    scene, cam = scenes.instanced()
    arrays = scene.build_arrays()
    print("C3: building 8.4M-triangle scene on the device (FP32 and 8-bit nodes)", flush=True)
    d32 = reina_rt.DeviceScene(arrays, flags=abi.BUILD_DEVICE)
    dq = reina_rt.DeviceScene(arrays, flags=abi.BUILD_DEVICE | abi.NODES_QUANTIZED)
    assert d32.info().num_triangles == 8388608
    yield arrays, cam, d32, dq
    d32.close()
    dq.close()


def test_c3_intersect_vs_fp64(cuda, c3):
    torch = cuda
    arrays, cam, d32, dq = c3
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_bvh(*d32.export_bvh(), 256)   # FP32 boxes cull the FP64 scan (padded by tie_rel)
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], 1920, 1080, 1, 12)
    prim = camera_rays(pc, 1920, 1080, 40000, 61)
    ph = gpu_intersect(torch, d32, prim)
    rays = np.concatenate([prim, secondary_rays(orc, prim, ph, 62)])
    print("C3: %d rays vs FP64" % len(rays), flush=True)
    hits = gpu_intersect(torch, d32, rays)
    hq = gpu_intersect(torch, dq, rays)
    assert np.array_equal(_bits(hits), _bits(hq))     # node format does not change the closest hit
    t64, p64, u64, v64, amb = orc.fp64(rays, tie_rel=1e-5, use_bvh=True, nthreads=16)
    agree = hits["prim"] == p64
    assert agree.mean() >= 0.9999, agree.mean()
    assert (~agree & ~amb).sum() == 0
    m = agree & (p64 != 0xFFFFFFFF)
    assert m[:len(prim)].sum() > 10000 and m[len(prim):].sum() > 3000
    err, scale = fp32_t_error(arrays, rays, hits, t64, p64, m)
    rel = err / t64[m]
    primary = np.arange(len(rays))[m] < len(prim)
    assert (rel[primary] <= 1e-6).mean() >= 0.998 and rel[primary].max() <= 1e-5
    # the watertight test divides by its determinant, proportional to the
    # cosine between ray and triangle normal: grazing bounce rays lose
    # precision as 1 / |cos|, so the bound is on err |cos| / scale
    from conftest import world_triangles
    wt = world_triangles(arrays).reshape(-1, 3, 3).astype(np.float64)[p64[m]]
    nrm = np.cross(wt[:, 1] - wt[:, 0], wt[:, 2] - wt[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1)[m].astype(np.float64)
    cos = np.abs(np.sum(nrm * d, 1)) / np.linalg.norm(d, axis=1)
    q = err / scale
    assert (q <= 1e-6).mean() >= 0.99 and (q * cos).max() <= 1e-5, (q.max(), (q * cos).max())


def test_c3_frame_quantized_equals_fp32(cuda, c3):
    torch = cuda
    arrays, cam, d32, dq = c3
    W, H = 1920, 1080
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 32, 12)
    print("C3: 1920x1080 32 spp depth 12, FP32 and 8-bit nodes", flush=True)
    a = gpu_render(torch, d32, pc, W, H)
    b = gpu_render(torch, dq, pc, W, H)
    assert _same(a, b)
    assert d32.stats().extension_rays >= W * H * 32   # at least the camera ray of every path
    assert np.isfinite(a[..., :3]).mean() > 0.99


def test_c3_material_mix_vs_oracle(cuda, c3):
    torch = cuda
    arrays, cam, d32, _ = c3
    W, H = 64, 36
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 2, 12, sample_batch=3)
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_bvh(*d32.export_bvh(), 256)
    got = gpu_render(torch, d32, pc, W, H)
    ref, _ = orc.render(pc, W, H, use_bvh=True, nthreads=16)
    assert _same(got, ref)
    kinds = {int(i["materialIdx"]) for i in arrays["instances"]}
    assert kinds == {0, 1, 2}


def test_c4_batched_64_scenes(cuda):
    torch = cuda
    rng = np.random.default_rng(64)
    print("C4: 64 scenes, 1e3-2e6 triangles", flush=True)
    pool = loadgen.make_pool(loadgen.log_uniform(rng, 1e3, 2e6, 64), seed=4, device_build=True)
    try:
        reqs = loadgen.make_requests(pool, 64, rng, [(128, 128), (256, 256), (512, 512), (1024, 1024)],
                                     spp_choices=(4, 16, 64, 256), depth=8, orbit=False)
        res = loadgen.batched(torch, pool, reqs, dict(path_budget_bytes=2550 << 20))
        print("C4:", res["plain"], res["batched"], flush=True)
        assert res["images_bit_identical"] is True
        assert res["batched"]["extension_rays"] == res["plain"]["extension_rays"]
    finally:
        for p in pool:
            p["dev"].close()


def test_c5_served_32_scenes(cuda):
    torch = cuda
    rng = np.random.default_rng(5)
    print("C5: 32 resident scenes, 1e4-4e6 triangles", flush=True)
    pool = loadgen.make_pool(loadgen.log_uniform(rng, 1e4, 4e6, 32), seed=5, device_build=True)
    try:
        reqs = loadgen.make_requests(pool, 200, rng, [(256, 256), (512, 512), (1280, 720), (1920, 1080)],
                                     spp_range=(16, 128))
        times = loadgen.arrivals(100.0, 2.0, np.random.default_rng(100))
        res = loadgen.served_server(torch, pool, reqs[:len(times)], times, 16, dict(path_budget_bytes=2550 << 20),
                                    keep_images=True)
        done = res.pop("images")
        print("C5: offered %d completed %d rejected %d" % (len(times), res["completed"], res["rejected"]), flush=True)
        assert res["completed"] >= 16 and res["completed"] + res["rejected"] == len(times)
        for k in sorted(done)[:: max(1, len(done) // 12)]:
            r = reqs[k]
            alone = torch.zeros((r["H"], r["W"], 4), dtype=torch.float32, device="cuda")
            pool[r["scene"]]["dev"].trace(r["pc"], alone.data_ptr(), r["W"], r["H"],
                                          stream=torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            assert torch.equal(alone.view(torch.int32), done[k].view(torch.int32)), k
    finally:
        for p in pool:
            p["dev"].close()
