"""GPU parity: the HIP kernels against the oracle, through the C ABI.

Bit-exact where both sides evaluate the same FP32 operation sequence (the
elementary functions, the watertight intersection, the full integrator, the
post-process); statistical / tolerance-based against the FP64 intersection
oracle.
"""
import numpy as np
import pytest

import oracle
import reina_rt
from reina_rt import abi, scenes
from reina_rt.scene import make_desc, push_consts

from conftest import world_triangles

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def gpu_intersect(torch, dev, rays, counters=False):
    r = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    h = torch.zeros((len(rays), 4), dtype=torch.float32, device="cuda")
    c = torch.zeros((len(rays), 2), dtype=torch.int32, device="cuda") if counters else None
    dev.intersect(r.data_ptr(), h.data_ptr(), len(rays), c.data_ptr() if counters else None,
                  stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    hits = h.cpu().numpy().view(abi.HIT).reshape(-1)
    return (hits, c.cpu().numpy().astype(np.uint32)) if counters else hits


def gpu_render(torch, dev, pc, W, H, init=None):
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda") if init is None else \
        torch.from_numpy(init.copy()).cuda()
    dev.trace(pc, img.data_ptr(), W, H, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return img.cpu().numpy()


def random_rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, abi.RAY)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    rays["ox"], rays["oy"], rays["oz"] = o.T
    rays["dx"], rays["dy"], rays["dz"] = d.T
    rays["tmax"] = 10000.0
    return rays


def camera_rays(pc, W, H, n, seed):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, abi.RAY)
    for i in range(n):
        o, d, _ = oracle.camera_ray(pc, int(rng.integers(0, W)), int(rng.integers(0, H)), W, H,
                                    int(rng.integers(0, 2 ** 32)))
        rays[i] = (o[0], o[1], o[2], 0.0, d[0], d[1], d[2], 10000.0)
    return rays


def fp32_t_error(arrays, rays, hits, t64, p64, m):
    """|t32 - t64| on the rays in mask m, and the scale the FP32 watertight
    test's error is proportional to: max(t, largest distance from the ray
    origin to a vertex of the hit triangle, in units of |d|)."""
    wt = world_triangles(arrays).reshape(-1, 3, 3).astype(np.float64)
    o = np.stack([rays["ox"], rays["oy"], rays["oz"]], 1)[m].astype(np.float64)
    dn = np.linalg.norm(np.stack([rays["dx"], rays["dy"], rays["dz"]], 1)[m].astype(np.float64), axis=1)
    D = np.linalg.norm(wt[p64[m]] - o[:, None, :], axis=2).max(1) / dn
    err = np.abs(hits["t"][m].astype(np.float64) - t64[m])
    return err, np.maximum(t64[m], D)


def secondary_rays(orc, rays, hits, seed):
    """Rays leaving the hit points in random directions (origin offset by the
    library's self-intersection rule)."""
    rng = np.random.default_rng(seed)
    m = hits["prim"] != 0xFFFFFFFF
    r, h = rays[m], hits[m]
    o = np.stack([r["ox"], r["oy"], r["oz"]], 1) + h["t"][:, None] * np.stack([r["dx"], r["dy"], r["dz"]], 1)
    d = rng.normal(size=o.shape).astype(np.float32)
    out = np.zeros(len(o), abi.RAY)
    out["ox"], out["oy"], out["oz"] = o.astype(np.float32).T
    out["dx"], out["dy"], out["dz"] = d.T
    out["tmin"] = 1e-4
    out["tmax"] = 10000.0
    return out


# ---- elementary functions ----------------------------------------------------------
@pytest.mark.parametrize("fn,lo,hi", [(0, -20.0, 20.0), (1, -20.0, 20.0), (2, 1e-30, 1e4), (3, -100.0, 90.0),
                                      (4, 0.0, 1e6), (5, 1e-6, 10.0), (6, -140.0, 130.0)])
def test_math_bitexact(cuda, fn, lo, hi):
    torch = cuda
    rng = np.random.default_rng(fn)
    x = rng.uniform(lo, hi, 200000).astype(np.float32)
    if fn == 2:
        x = np.exp(rng.uniform(np.log(1e-38), np.log(1e30), 200000)).astype(np.float32)
    y = rng.uniform(-3.0, 3.0, len(x)).astype(np.float32)
    specials = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 1e-40, 3.0e38], np.float32)
    x = np.concatenate([x, specials])
    y = np.concatenate([y, np.full(len(specials), 0.5, np.float32)])
    ref = oracle.math_eval(fn, x, y)
    tx, ty = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    out = torch.zeros_like(tx)
    reina_rt.debug_math(fn, tx.data_ptr(), ty.data_ptr(), out.data_ptr(), len(x))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    same = (_bits(got) == _bits(ref)) | (np.isnan(got) & np.isnan(ref))
    assert same.all(), "fn %d: %d mismatches, e.g. x=%r gpu=%r oracle=%r" % (
        fn, (~same).sum(), x[~same][:3], got[~same][:3], ref[~same][:3])


# ---- closest hit -------------------------------------------------------------------
def test_intersect_bitexact_cornell(cuda, cornell_arrays):
    torch = cuda
    arrays, cam = cornell_arrays
    dev = reina_rt.DeviceScene(arrays)
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_bvh(*dev.export_bvh())
    pc = scenes.cornell_push(256, 256, 1, 8)
    rays = np.concatenate([random_rays(100000, -1.5, 2.5, 1), camera_rays(pc, 256, 256, 4000, 2)])
    hits, cnt = gpu_intersect(torch, dev, rays, counters=True)
    ref_b = orc.brute32(rays)
    ref_v, ref_c = orc.bvh(rays)
    for f in ("t", "u", "v"):
        assert np.array_equal(_bits(hits[f]), _bits(ref_b[f])), f
    assert np.array_equal(hits["prim"], ref_b["prim"])
    assert np.array_equal(cnt, ref_c), "node-visit / triangle-test counters differ from the traversal restatement"
    dev.close()


def test_intersect_bitexact_icosphere(cuda, icosphere_arrays):
    torch = cuda
    arrays, cam = icosphere_arrays
    dev = reina_rt.DeviceScene(arrays)
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_bvh(*dev.export_bvh())
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], 1920, 1080, 1, 8)
    prim = camera_rays(pc, 1920, 1080, 3000, 3)
    ph = orc.brute32(prim)
    rays = np.concatenate([prim, secondary_rays(orc, prim, ph, 4), random_rays(2000, -1.5, 1.5, 5)])
    hits, cnt = gpu_intersect(torch, dev, rays, counters=True)
    ref = orc.brute32(rays)
    ref_v, ref_c = orc.bvh(rays)
    assert np.array_equal(hits["prim"], ref["prim"])
    for f in ("t", "u", "v"):
        assert np.array_equal(_bits(hits[f]), _bits(ref[f])), f
    assert np.array_equal(cnt, ref_c)
    dev.close()


def test_intersect_claims_any_batch_size(cuda, icosphere_arrays):
    """The traversal kernel's claims give every ray exactly once, whatever the
    batch size and launch shape: partitions with no rays (fewer rays than
    partitions), claims that take a partition's last rays exactly or run past
    its end, and grids far larger than the batch, whose late waves find every
    partition used up. Hits equal the full batch's (pinned to the oracle by
    test_intersect_bitexact_icosphere) and, for the small batches, the
    oracle's brute-force scan."""
    torch = cuda
    arrays, cam = icosphere_arrays
    dev = reina_rt.DeviceScene(arrays)
    orc = oracle.OracleScene(arrays, make_desc)
    rays = random_rays(70001, -1.5, 1.5, 21)
    full = gpu_intersect(torch, dev, rays)
    small = rays[:1025]
    ref = orc.brute32(small)
    assert np.array_equal(full["prim"][:1025], ref["prim"])
    assert np.array_equal(_bits(full["t"][:1025]), _bits(ref["t"]))
    assert (full["prim"] != 0xFFFFFFFF).sum() > 20000   # the batch hits the sphere
    for grid, claim in ((1, 64), (8, 512), (256, 128), (2048, 64), (4096, 512)):
        dev.set_tuning(grid_blocks=grid, claim_rays=claim)
        for m in (1, 5, 8, 9, 63, 64, 65, 1023, 1025, 8 * 512 + 3, 70001):
            h = gpu_intersect(torch, dev, rays[:m])
            assert np.array_equal(_bits(h.view(np.float32).reshape(-1, 4)),
                                  _bits(full[:m].view(np.float32).reshape(-1, 4))), (grid, claim, m)
    dev.close()


def test_intersect_quantized_bitexact(cuda, icosphere_arrays):
    """8-bit quantized nodes: the GPU hit equals brute force bit for bit, and
    the per-ray node visits / triangle tests equal the oracle's traversal of
    the same quantized nodes."""
    torch = cuda
    arrays, cam = icosphere_arrays
    dev = reina_rt.DeviceScene(arrays, flags=abi.NODES_QUANTIZED)
    info = dev.info()
    assert info.quantized == 1 and info.node_bytes == 80
    orc = oracle.OracleScene(arrays, make_desc)
    nodes, tris = dev.export_bvh()
    assert len(nodes) == 80 * info.num_nodes
    orc.set_bvh(nodes, tris, 80)
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], 1920, 1080, 1, 8)
    prim = camera_rays(pc, 1920, 1080, 3000, 13)
    ph = orc.brute32(prim)
    rays = np.concatenate([prim, secondary_rays(orc, prim, ph, 14), random_rays(2000, -1.5, 1.5, 15)])
    hits, cnt = gpu_intersect(torch, dev, rays, counters=True)
    ref = orc.brute32(rays)
    ref_v, ref_c = orc.bvh(rays)
    assert np.array_equal(hits["prim"], ref["prim"])
    for f in ("t", "u", "v"):
        assert np.array_equal(_bits(hits[f]), _bits(ref[f])), f
    assert np.array_equal(cnt, ref_c)
    dev.close()


def test_render_quantized_equals_fp32(cuda, icosphere_arrays):
    """The closest hit does not depend on the node format, so a frame traced
    over quantized nodes is bit-identical to the FP32-node frame."""
    torch = cuda
    arrays, cam = icosphere_arrays
    W, H = 160, 90
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 4, 8)
    dev32 = reina_rt.DeviceScene(arrays)
    devq = reina_rt.DeviceScene(arrays, flags=abi.NODES_QUANTIZED)
    a = gpu_render(torch, dev32, pc, W, H)
    b = gpu_render(torch, devq, pc, W, H)
    same = (_bits(a) == _bits(b)) | (np.isnan(a) & np.isnan(b))
    assert same.all(), "pixels differing: %d" % (~same).any(-1).sum()
    assert dev32.stats().extension_rays == devq.stats().extension_rays
    dev32.close()
    devq.close()


@pytest.mark.parametrize("flags", [0, abi.NODES_QUANTIZED])
def test_intersect_zero_direction_components(cuda, flags):
    """Rays inside the coordinate planes / along the axes (direction
    components exactly 0) aimed at mesh vertices and edges in those planes:
    GPU == brute force, counters == the oracle's traversal."""
    from test_oracle_geometry import plane_rays, soup_arrays
    torch = cuda
    v, f = scenes.subdivided_icosphere(3)
    v = v.astype(np.float32)
    arrays = soup_arrays(v[f].reshape(-1, 9))
    dev = reina_rt.DeviceScene(arrays, flags=flags)
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_bvh(*dev.export_bvh(), dev.info().node_bytes)
    rays = plane_rays(v, 20000, 22)
    hits, cnt = gpu_intersect(torch, dev, rays, counters=True)
    ref = orc.brute32(rays)
    ref_v, ref_c = orc.bvh(rays)
    assert np.array_equal(hits["prim"], ref["prim"])
    for fl in ("t", "u", "v"):
        assert np.array_equal(_bits(hits[fl]), _bits(ref[fl])), fl
    assert np.array_equal(cnt, ref_c)
    dev.close()


@pytest.mark.parametrize("flags", [0, abi.NODES_QUANTIZED])
@pytest.mark.parametrize("shift", [0.0, 3000.0])
def test_intersect_near_axis_far_origin(cuda, flags, shift):
    """Tiny direction components on either side of the switch between the
    two slab forms (2^-41, 2^-39, 1e-9, 1e-6) and a mesh 3000 units from the
    coordinate origin: GPU == brute force, counters == the oracle's traversal."""
    from test_oracle_geometry import near_axis_rays, shifted_icosphere
    torch = cuda
    v, arrays = shifted_icosphere(shift)
    dev = reina_rt.DeviceScene(arrays, flags=flags)
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_bvh(*dev.export_bvh(), dev.info().node_bytes)
    rays = near_axis_rays(v, 20000, 32)
    hits, cnt = gpu_intersect(torch, dev, rays, counters=True)
    ref = orc.brute32(rays)
    ref_v, ref_c = orc.bvh(rays)
    assert np.array_equal(hits["prim"], ref["prim"])
    for fl in ("t", "u", "v"):
        assert np.array_equal(_bits(hits[fl]), _bits(ref[fl])), fl
    assert np.array_equal(cnt, ref_c)
    dev.close()


def test_intersect_vs_fp64(cuda, icosphere_arrays):
    """Hit id agrees with the FP64 oracle on >= 99.99% of rays; every
    disagreement is an FP64-detected edge/vertex/equal-t tie; t within 1e-6
    relative on the agreeing primary rays."""
    torch = cuda
    arrays, cam = icosphere_arrays
    dev = reina_rt.DeviceScene(arrays)
    orc = oracle.OracleScene(arrays, make_desc)
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], 1920, 1080, 1, 8)
    prim = camera_rays(pc, 1920, 1080, 3000, 6)
    hits = gpu_intersect(torch, dev, prim)
    t64, p64, u64, v64, amb = orc.fp64(prim, tie_rel=1e-5)
    agree = hits["prim"] == p64
    assert agree.mean() >= 0.9999 or (~agree & ~amb).sum() == 0
    assert (~agree & ~amb).sum() == 0, "non-tie disagreements: %d" % (~agree & ~amb).sum()
    m = agree & (p64 != 0xFFFFFFFF)
    rel = np.abs(hits["t"][m].astype(np.float64) - t64[m]) / t64[m]
    assert rel.max() <= 1e-6, "max rel t error %g" % rel.max()
    dev.close()


@pytest.mark.parametrize("flags", [0, abi.NODES_QUANTIZED])
def test_instanced_vs_fp64(cuda, flags):
    """Instanced spheres (256 copies of the 2,048-triangle sphere; random
    scale, rotation, position): camera rays and bounce rays against the
    oracle's traversal (bit-exact) and the FP64 oracle (hit id on >= 99.99%,
    every disagreement an FP64-detected tie, t within 1e-6 relative)."""
    torch = cuda
    scene, cam = scenes.instanced(copies=256, box=40.0)
    arrays = scene.build_arrays()
    dev = reina_rt.DeviceScene(arrays, flags=flags)
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_bvh(*dev.export_bvh(), dev.info().node_bytes)
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], 640, 360, 1, 12)
    prim = camera_rays(pc, 640, 360, 10000, 41)
    ph = orc.bvh(prim)[0]
    rays = np.concatenate([prim, secondary_rays(orc, prim, ph, 42)])
    hits, cnt = gpu_intersect(torch, dev, rays, counters=True)
    ref_v, ref_c = orc.bvh(rays)
    assert np.array_equal(hits["prim"], ref_v["prim"])
    for f in ("t", "u", "v"):
        assert np.array_equal(_bits(hits[f]), _bits(ref_v[f])), f
    assert np.array_equal(cnt, ref_c)
    nodes32, tris32, _ = reina_rt.bvh_build_host(world_triangles(arrays))   # FP32 boxes cull the FP64 scan
    orc.set_bvh(nodes32, tris32, 256)
    t64, p64, u64, v64, amb = orc.fp64(rays, tie_rel=1e-5, use_bvh=True)
    agree = hits["prim"] == p64
    assert agree.mean() >= 0.9999
    assert (~agree & ~amb).sum() == 0
    m = agree & (p64 != 0xFFFFFFFF)
    assert m[:len(prim)].sum() > 500 and m[len(prim):].sum() > 300
    err, t_err_scale = fp32_t_error(arrays, rays, hits, t64, p64, m)
    rel = err / t64[m]
    primary = np.arange(len(rays))[m] < len(prim)
    # camera rays: 1e-6 relative on >= 99.8% (grazing hits are ill-conditioned)
    assert (rel[primary] <= 1e-6).mean() >= 0.998 and rel[primary].max() <= 1e-5
    # bounce rays: the FP32 error scales with the translated vertices, not t
    assert (err / t_err_scale <= 1e-6).mean() >= 0.99 and (err / t_err_scale).max() <= 1e-5
    dev.close()


# ---- integrator ------------------------------------------------------------------
TAILS = [0, 1500, 1 << 30]   # wavefront only; wavefront then tail kernel; tail kernel only


@pytest.mark.parametrize("tail", TAILS)
def test_render_cornell_bitexact(cuda, cornell_arrays, tail):
    torch = cuda
    arrays, cam = cornell_arrays
    W = H = 64
    pc = scenes.cornell_push(W, H, spp=16, max_bounces=8)
    dev = reina_rt.DeviceScene(arrays)
    dev.set_tuning(tail_paths=tail)
    got = gpu_render(torch, dev, pc, W, H)
    ref, nrays = oracle.OracleScene(arrays, make_desc).render(pc, W, H, use_bvh=False)
    assert np.array_equal(_bits(got), _bits(ref)), "pixels differing: %d" % (
        (_bits(got) != _bits(ref)).any(-1).sum())
    st = dev.stats()
    assert st.extension_rays == nrays
    assert st.tail_launches == (1 if tail else 0)
    if tail == 0:
        assert st.tail_rays == 0
    elif tail == 1500:
        assert 0 < st.tail_rays < st.extension_rays
    else:
        assert st.tail_rays == st.extension_rays
    dev.close()


@pytest.mark.parametrize("tail", TAILS)
def test_render_icosphere_bitexact(cuda, icosphere_arrays, tail):
    torch = cuda
    arrays, cam = icosphere_arrays
    W, H = 96, 54
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 4, 8)
    dev = reina_rt.DeviceScene(arrays)
    dev.set_tuning(tail_paths=tail)
    got = gpu_render(torch, dev, pc, W, H)
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_bvh(*dev.export_bvh())
    ref, nrays = orc.render(pc, W, H, use_bvh=True)
    same = _bits(got) == _bits(ref)
    same |= np.isnan(got) & np.isnan(ref)
    assert same.all(), "pixels differing: %d" % (~same).any(-1).sum()
    assert dev.stats().extension_rays == nrays
    dev.close()


@pytest.mark.parametrize("tail", [0, 1 << 30])
def test_render_partitions_bitexact(cuda, cornell_arrays, tail):
    """150 x 100 pixels = 15 queue tiles of 1024 (two per XCD partition, the
    last one partial): every partition is used and refilled (by the
    wavefront kernels, or claimed by the tail kernel)."""
    torch = cuda
    arrays, _ = cornell_arrays
    W, H = 150, 100
    pc = scenes.cornell_push(W, H, spp=3, max_bounces=8)
    dev = reina_rt.DeviceScene(arrays)
    dev.set_tuning(tail_paths=tail)
    got = gpu_render(torch, dev, pc, W, H)
    ref, nrays = oracle.OracleScene(arrays, make_desc).render(pc, W, H, use_bvh=False)
    assert np.array_equal(_bits(got), _bits(ref))
    assert dev.stats().extension_rays == nrays
    dev.close()


@pytest.mark.parametrize("tail", [0, 20000])
def test_render_readbacks_bitexact(cuda, cornell_arrays, tail):
    """320 x 256 pixels (10 queue tiles per partition), 4 spp, depth 8: up to
    32 bounce iterations, so the host's pipelined queue-length readbacks (every
    4 iterations, read one round late) decide the end of the frame and the
    switch to the tail launch; the frame equals the oracle's bit for bit."""
    torch = cuda
    arrays, _ = cornell_arrays
    W, H = 320, 256
    pc = scenes.cornell_push(W, H, spp=4, max_bounces=8)
    dev = reina_rt.DeviceScene(arrays)
    dev.set_tuning(tail_paths=tail)
    got = gpu_render(torch, dev, pc, W, H)
    ref, nrays = oracle.OracleScene(arrays, make_desc).render(pc, W, H, use_bvh=False)
    assert np.array_equal(_bits(got), _bits(ref)), "pixels differing: %d" % (
        (_bits(got) != _bits(ref)).any(-1).sum())
    st = dev.stats()
    assert st.extension_rays == nrays
    assert st.tail_launches == (1 if tail else 0)
    dev.close()


def test_running_mean_bitexact(cuda, cornell_arrays):
    """sampleBatch > 0 blends (prev * b + new) / (b + 1) (raytrace.rgen.glsl:277-282)."""
    torch = cuda
    arrays, _ = cornell_arrays
    W, H = 40, 24
    dev = reina_rt.DeviceScene(arrays)
    orc = oracle.OracleScene(arrays, make_desc)
    img_g, img_o = None, np.zeros((H, W, 4), np.float32)
    for b in range(3):
        pc = scenes.cornell_push(W, H, spp=2, max_bounces=4, sample_batch=b)
        img_g = gpu_render(torch, dev, pc, W, H, init=img_g)
        img_o, _ = orc.render(pc, W, H, img=img_o, use_bvh=False)
        assert np.array_equal(_bits(img_g), _bits(img_o)), "batch %d" % b
    dev.close()


@pytest.mark.parametrize("tail", [0, 1 << 30], ids=["wavefront", "tail_kernel"])
@pytest.mark.parametrize("defocus", [0.015, 0.2], ids=["config_default", "strong"])
def test_render_defocus_bitexact(cuda, cornell_arrays, defocus, tail):
    """Depth of field: the hexagon lens sample of raytrace.rgen.glsl:186-200,
    drawn from a copy of the pixel's RNG state (the lens sample does not
    advance the stream). config/config.toml's defocus_multiplier = 1.5 is
    0.015 after Reina.cpp's division by 100."""
    torch = cuda
    arrays, cam = cornell_arrays
    W, H = 48, 40
    pc = reina_rt.push_consts(cam["eye"], cam["center"], cam["fov"], W, H, spp=4, max_bounces=8, defocus=defocus,
                              focus_dist=2.2)
    dev = reina_rt.DeviceScene(arrays)
    dev.set_tuning(tail_paths=tail)
    got = gpu_render(torch, dev, pc, W, H)
    ref, nrays = oracle.OracleScene(arrays, make_desc).render(pc, W, H, use_bvh=False)
    assert np.array_equal(_bits(got), _bits(ref)), "pixels differing: %d" % (_bits(got) != _bits(ref)).any(-1).sum()
    pin = reina_rt.push_consts(cam["eye"], cam["center"], cam["fov"], W, H, spp=4, max_bounces=8)
    assert not np.array_equal(_bits(gpu_render(torch, dev, pin, W, H)), _bits(got))   # the lens moved the rays
    dev.close()


def _quirk_cases():
    stack, scam = scenes.dielectric_stack()
    nan, ncam = scenes.nan_floor_slab()
    box, bcam = scenes.cornell()
    return [("dielectric_distance_leak", stack.build_arrays(), scam, 24, 20, 4, 6, {}),
            ("nan_samples_dropped", nan.build_arrays(), ncam, 24, 20, 16, 12, {}),
            ("direct_clamp", box.build_arrays(), bcam, 64, 64, 4, 6,
             dict(direct_clamp=0.5, indirect_clamp=1e-3))]


@pytest.mark.parametrize("tail", [0, 1 << 30], ids=["wavefront", "tail_kernel"])
def test_render_raygen_quirks_bitexact(cuda, tail):
    """The distance carried from one sample's glass into the next, NaN
    samples dropped from the average and directClamp on the whole sample
    with indirectClamp / totalEmissiveWeight unused: the scenes
    whose closed forms pin the oracle render bit for bit the same on the GPU."""
    torch = cuda
    for name, arrays, cam, W, H, spp, bounces, kw in _quirk_cases():
        pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, spp, bounces,
                         up=cam.get("up", (0.0, 1.0, 0.0)), **kw)
        pc["totalEmissiveWeight"] = 7.0
        dev = reina_rt.DeviceScene(arrays)
        dev.set_tuning(tail_paths=tail)
        got = gpu_render(torch, dev, pc, W, H)
        ref, nrays = oracle.OracleScene(arrays, make_desc).render(pc, W, H, use_bvh=False)
        same = (_bits(got) == _bits(ref)) | (np.isnan(got) & np.isnan(ref))
        assert same.all(), "%s: pixels differing: %d" % (name, (~same).any(-1).sum())
        assert dev.stats().extension_rays == nrays, name
        if name == "nan_samples_dropped":
            assert (~np.isnan(got[..., 0])).mean() > 0.85
        dev.close()


def _fix_scenes():
    box, bcam = scenes.cornell()
    stack, scam = scenes.dielectric_stack()
    mirror, mcam = scenes.tilted_mirror()
    lobes, lcam = scenes.disney_lobes()
    return [("cornell_defocus", box.build_arrays(), bcam, dict(defocus=0.1), abi.TRACE_FIX_LENS_RNG),
            ("dielectric_stack", stack.build_arrays(), scam, {}, abi.TRACE_FIX_MEDIUM_RESET),
            ("tilted_mirror", mirror.build_arrays(), mcam, {}, abi.TRACE_FIX_NORMALS),
            ("disney_lobes", lobes.build_arrays(), lcam, {}, abi.TRACE_FIX_METAL_PDF | abi.TRACE_FIX_DISNEY_MIX)]


@pytest.mark.parametrize("fixes", [abi.TRACE_FIX_LENS_RNG, abi.TRACE_FIX_MEDIUM_RESET, abi.TRACE_FIX_NORMALS,
                                   abi.TRACE_FIX_METAL_PDF, abi.TRACE_FIX_DISNEY_MIX, abi.TRACE_FIXES],
                         ids=["lens_rng", "medium_reset", "normals", "metal_pdf", "disney_mix", "all"])
def test_render_fixes_bitexact(cuda, fixes):
    """Each REINA_RT_TRACE_FIX_* correction (and all together) renders bit
    for bit as the oracle with the same correction, through the wavefront
    and the tail kernel, and changes the frame of the scene it is about."""
    torch = cuda
    for name, arrays, cam, kw, target in _fix_scenes():
        W, H = 32, 24
        pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 4, 6, up=cam.get("up", (0.0, 1.0, 0.0)), **kw)
        orc = oracle.OracleScene(arrays, make_desc)
        orc.set_fixes(fixes)
        ref, nrays = orc.render(pc, W, H, use_bvh=False)
        dev = reina_rt.DeviceScene(arrays)
        dev.set_trace_options(fixes)
        for tail in (0, 1 << 30):
            dev.set_tuning(tail_paths=tail)
            got = gpu_render(torch, dev, pc, W, H)
            same = (_bits(got) == _bits(ref)) | (np.isnan(got) & np.isnan(ref))
            assert same.all(), "%s tail=%d: pixels differing: %d" % (name, tail, (~same).any(-1).sum())
            assert dev.stats().extension_rays == nrays
        if fixes & target:
            dev.set_trace_options(0)
            assert not np.array_equal(_bits(gpu_render(torch, dev, pc, W, H)), _bits(got)), name
        dev.close()


@pytest.mark.parametrize("log2g", [1, 3])
@pytest.mark.parametrize("nee,tail", [(False, 0), (False, 1 << 30), (True, 0)], ids=["plain", "tail_kernel", "nee"])
def test_render_sample_groups_bitexact(cuda, cornell_arrays, log2g, nee, tail):
    """REINA_RT_TRACE_SAMPLE_GROUPS: a pixel's samples in 2^n parallel
    streams, progressive (sampleBatch 1 over a given image); the frame equals
    the oracle's grouped restatement bit for bit."""
    torch = cuda
    arrays, cam = cornell_arrays
    W, H = 40, 30
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 16, 8, sample_batch=1)
    init = np.random.default_rng(3).uniform(0, 1, (H, W, 4)).astype(np.float32)
    opts = abi.TRACE_SAMPLE_GROUPS(log2g) | (abi.TRACE_NEE if nee else 0)
    dev = reina_rt.DeviceScene(arrays)
    dev.set_tuning(tail_paths=tail)
    dev.set_trace_options(opts)
    got = gpu_render(torch, dev, pc, W, H, init)
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_fixes(opts)
    orc.set_nee(nee)
    ref, nrays = orc.render(pc, W, H, img=init.copy(), use_bvh=False)
    same = (_bits(got) == _bits(ref)) | (np.isnan(got) & np.isnan(ref))
    assert same.all(), "pixels differing: %d" % (~same).any(-1).sum()
    assert dev.stats().extension_rays == nrays
    dev.close()


def test_sample_groups_errors(cuda, cornell_arrays):
    torch = cuda
    arrays, cam = cornell_arrays
    dev = reina_rt.DeviceScene(arrays)
    dev.set_trace_options(abi.TRACE_SAMPLE_GROUPS(2))
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], 8, 8, 6, 4)        # 6 spp, 4 groups
    with pytest.raises(reina_rt.ReinaError) as e:
        gpu_render(torch, dev, pc, 8, 8)
    assert e.value.code == abi.ERR_INVALID_ARG
    srv = reina_rt.Server(path_budget_bytes=64 << 20)
    img = torch.zeros((8, 8, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(reina_rt.ReinaError) as e:
        srv.submit(dev, pc, img.data_ptr(), 8, 8)
    assert e.value.code == abi.ERR_INVALID_ARG
    srv.close()
    dev.close()


def test_seed_wraps_mod_2_32(cuda, cornell_arrays):
    """The pixel seed (sampleBatch * H + y) * W + x is uint32 arithmetic
    (raytrace.rgen.glsl:259), so at 64 x 64 (W * H = 2^12) batch 2^20 reuses
    batch 0's streams. GPU = oracle bit for bit around the wrap (a 64-bit seed
    would differ), and batch 2^20 blended onto batch 0's image gives back that
    image (the same samples, to the blend's rounding)."""
    torch = cuda
    arrays, _ = cornell_arrays
    W = H = 64
    wrap = (1 << 32) // (W * H)
    dev = reina_rt.DeviceScene(arrays)
    orc = oracle.OracleScene(arrays, make_desc)
    for b in (wrap - 1, wrap, wrap + 1):
        pc = scenes.cornell_push(W, H, spp=2, max_bounces=4, sample_batch=b)
        g = gpu_render(torch, dev, pc, W, H)
        o, _ = orc.render(pc, W, H, use_bvh=False)
        assert np.array_equal(_bits(g), _bits(o)), "batch %d" % b
    first = gpu_render(torch, dev, scenes.cornell_push(W, H, spp=2, max_bounces=4, sample_batch=0), W, H)
    again = gpu_render(torch, dev, scenes.cornell_push(W, H, spp=2, max_bounces=4, sample_batch=wrap), W, H,
                       init=first)
    ok = np.isfinite(first)
    assert np.allclose(again[ok], first[ok], rtol=4e-7, atol=0)
    other = gpu_render(torch, dev, scenes.cornell_push(W, H, spp=2, max_bounces=4, sample_batch=wrap + 1), W, H,
                       init=first)
    assert not np.allclose(other[ok], first[ok], rtol=4e-7, atol=0)
    dev.close()


@pytest.mark.parametrize("spp,bounces", [(0, 8), (3, 0), (1, 1)])
def test_render_edge_cases(cuda, cornell_arrays, spp, bounces):
    torch = cuda
    arrays, _ = cornell_arrays
    W, H = 7, 5
    pc = scenes.cornell_push(W, H, spp=spp, max_bounces=bounces)
    dev = reina_rt.DeviceScene(arrays)
    got = gpu_render(torch, dev, pc, W, H)
    ref, _ = oracle.OracleScene(arrays, make_desc).render(pc, W, H, use_bvh=False)
    same = (_bits(got) == _bits(ref)) | (np.isnan(got) & np.isnan(ref))
    assert same.all()
    if spp == 0:
        assert np.isnan(got[..., :3]).all()
    if bounces == 0:
        assert (got[..., :3] == 0).all()
    dev.close()


# ---- post-process ----------------------------------------------------------------
def test_postprocess_bitexact(cuda):
    torch = cuda
    rng = np.random.default_rng(9)
    H, W = 45, 70
    img = (rng.random((H, W, 4)) ** 4 * 4.0).astype(np.float32)
    img[..., 3] = 1.0
    ref_out, ref_comb = oracle.postprocess(img, radius=5.0, threshold=1.0, intensity=0.05, exposure=1.0)
    t_in = torch.from_numpy(img).cuda()
    scratch = torch.zeros((2, H, W, 4), dtype=torch.float32, device="cuda")
    out = torch.zeros((H, W, 4), dtype=torch.uint8, device="cuda")
    reina_rt.postprocess(t_in.data_ptr(), scratch.data_ptr(), out.data_ptr(), W, H, 5.0, 1.0, 0.05, 1.0)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ref_out)
    assert np.array_equal(_bits(scratch[0].cpu().numpy()), _bits(ref_comb))


# ---- next-event estimation ---------------------------------------------------------
def _nee_scenes():
    c, ccam = scenes.cornell()
    b, bcam = scenes.light_box()
    i, icam = scenes.icosphere(levels=4)
    return [("cornell", c.build_arrays(), ccam), ("light_box", b.build_arrays(), bcam),
            ("icosphere", i.build_arrays(), icam)]


def test_emissive_tables_match_oracle(cuda):
    """The library's light tables equal the oracle's restatement byte for byte."""
    for name, arrays, cam in _nee_scenes():
        dev = reina_rt.DeviceScene(arrays)
        orc = oracle.OracleScene(arrays, make_desc)
        a, b = dev.export_emissive(), orc.export_emissive()
        assert len(a) == len(b) and np.array_equal(a, b), name
        assert dev.info().num_emissive > 0
        dev.close()


@pytest.mark.parametrize("flags,tail", [(0, 0), (0, 1 << 30), (abi.NODES_QUANTIZED, 700)])
def test_render_nee_bitexact(cuda, flags, tail):
    """NEE + MIS frames (shade writes shadow rays, the any-hit kernel adds
    the unoccluded ones) equal the oracle's sequential restatement bit for
    bit; shadow rays are counted."""
    torch = cuda
    for name, arrays, cam in _nee_scenes():
        W, H = (40, 30) if name != "icosphere" else (64, 36)
        pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 4, 6)
        dev = reina_rt.DeviceScene(arrays, flags=flags)
        dev.set_tuning(tail_paths=tail)
        dev.set_trace_options(abi.TRACE_NEE)
        got = gpu_render(torch, dev, pc, W, H)
        st = dev.stats()
        orc = oracle.OracleScene(arrays, make_desc)
        orc.set_nee(True)
        ref, nrays = orc.render(pc, W, H, use_bvh=False)
        same = (_bits(got) == _bits(ref)) | (np.isnan(got) & np.isnan(ref))
        assert same.all(), "%s: pixels differing: %d" % (name, (~same).any(-1).sum())
        assert st.extension_rays == nrays and st.shadow_rays > 0
        dev.set_trace_options(0)
        plain = gpu_render(torch, dev, pc, W, H)
        assert not np.array_equal(_bits(plain), _bits(got))
        dev.close()


@pytest.mark.parametrize("nee", [False, True], ids=["plain", "nee"])
@pytest.mark.parametrize("flags,tail", [(0, 0), (0, 1 << 30), (abi.NODES_QUANTIZED, 500)])
def test_render_disney_lobes_bitexact(cuda, nee, flags, tail):
    """Every lobe of the Disney BSDF (clearcoat, sheen, subsurface,
    anisotropic metal, glass with insideDielectric = choseGlass, and
    mixes) renders bit for bit as the oracle's restatement of
    brdfDisney.h.glsl, with and without NEE + MIS."""
    torch = cuda
    s, cam = scenes.disney_lobes()
    arrays = s.build_arrays()
    W, H = 48, 40
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, spp=4, max_bounces=8)
    dev = reina_rt.DeviceScene(arrays, flags=flags)
    dev.set_tuning(tail_paths=tail)
    dev.set_trace_options(abi.TRACE_NEE if nee else 0)
    got = gpu_render(torch, dev, pc, W, H)
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_nee(nee)
    ref, nrays = orc.render(pc, W, H, use_bvh=False)
    same = (_bits(got) == _bits(ref)) | (np.isnan(got) & np.isnan(ref))
    assert same.all(), "pixels differing: %d" % (~same).any(-1).sum()
    assert dev.stats().extension_rays == nrays
    dev.close()


def test_nee_needs_an_emitter(cuda):
    torch = cuda
    v, f = scenes.subdivided_icosphere(2)
    from test_oracle_geometry import soup_arrays
    dev = reina_rt.DeviceScene(soup_arrays(v[f].astype(np.float32).reshape(-1, 9)))
    assert dev.info().num_emissive == 0
    dev.set_trace_options(abi.TRACE_NEE)
    img = torch.zeros((8, 8, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(reina_rt.ReinaError) as e:
        dev.trace(push_consts((0, 0, 3), (0, 0, 0), 45.0, 8, 8, 1, 4), img.data_ptr(), 8, 8)
    assert e.value.code == abi.ERR_NO_EMISSIVE
    dev.close()


# ---- statistics against the converged FP64 render ----------------------------------
@pytest.mark.parametrize("nee", [False, True])
def test_rmse_vs_converged(cuda, nee):
    """Sixteen independent 64-spp GPU frames of the 24x24 Cornell box
    (sampleBatch 200000+) against the 65,536-spp FP64 converged render in
    tests/golden (made by tests/golden/make_converged.py). With no bias the
    mean squared error of a frame equals the per-pixel variance of a frame
    (plus the converged render's own, 1/1024 of a frame's).

    Standard-error model: with F frames x_f, their mean xb and unbiased
    variance v (per pixel and channel), mean_f (x_f - ref)^2 = (F-1)/F v +
    (xb - ref)^2, so measured minus predicted MSE is the pixel mean of
    D = (xb - ref)^2 - v / F - var_ref, whose expectation is 0 without bias.
    Pixels draw independent sample streams, so D's standard error is the
    spread of its per-pixel values (channels summed) over sqrt(pixels): the
    difference must lie within 2 of those standard errors. The relative RMSE
    is printed. NEE frames are compared with the same BSDF-only converged
    render (equal expectation)."""
    import os
    torch = cuda
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cornell_24x24_converged.npz"))
    ref, var_ref = g["mean"], g["var_batch64"] / 1024.0
    W = H = 24
    arrays, cam = scenes.cornell()[0].build_arrays(), None
    dev = reina_rt.DeviceScene(arrays)
    if nee:
        dev.set_trace_options(abi.TRACE_NEE)
    frames = []
    for f in range(16):
        pc = scenes.cornell_push(W, H, spp=64, max_bounces=8)
        pc = np.asarray(pc).copy()
        pc["sampleBatch"] = 200000 + f
        img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        dev.trace(pc, img.data_ptr(), W, H, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        frames.append(img.cpu().numpy()[..., :3].astype(np.float64))
    dev.close()
    # sampleBatch > 0 blends with the (zero) previous image: undo the running mean
    fr = np.array([x * (200001 + k) for k, x in enumerate(frames)])
    F = len(fr)
    v = fr.var(0, ddof=1)
    mse = ((fr - ref) ** 2).mean(axis=(1, 2, 3))
    pred = (v + var_ref).mean()
    d = (((fr.mean(0) - ref) ** 2) - v / F - var_ref).sum(-1)   # per pixel, channels summed
    assert np.isclose(d.mean() / 3.0, mse.mean() - pred, rtol=1e-6, atol=1e-12)
    se = d.std(ddof=1) / np.sqrt(d.size) / 3.0
    rrmse = np.sqrt(((fr - ref) ** 2 / (ref ** 2 + 1e-3)).mean())
    print("nee=%s mse=%.4g pred=%.4g se=%.3g z=%.2f rel_rmse=%.4f" % (
        nee, mse.mean(), pred, se, (mse.mean() - pred) / se, rrmse))
    assert abs(mse.mean() - pred) < 2.0 * se
    if not nee:
        # the frames' variance agrees with the converged render's own batch variance
        assert 0.7 < v.mean() / g["var_batch64"].mean() < 1.4


# ---- BVH8 built on the device --------------------------------------------------------
def _device_build_cases():
    from test_oracle_geometry import soup_arrays
    rng = np.random.default_rng(51)
    cases = []
    for n in (1, 2, 3, 4, 9, 100):
        cases.append(("soup%d" % n, soup_arrays(rng.uniform(-1, 1, (n, 9)).astype(np.float32))))
    same = np.tile(rng.uniform(-1, 1, (1, 9)).astype(np.float32), (50, 1))   # 50 identical triangles
    cases.append(("identical", soup_arrays(same)))
    c = rng.uniform(-1, 1, (20000, 1, 3))
    cases.append(("soup20k", soup_arrays((c + 0.05 * rng.normal(size=(20000, 3, 3))).astype(np.float32).reshape(-1, 9))))
    cases.append(("cornell", scenes.cornell()[0].build_arrays()))
    cases.append(("instanced", scenes.instanced(copies=128, box=20.0)[0].build_arrays()))
    return cases


def test_device_bvh_build(cuda):
    """REINA_RT_BUILD_DEVICE: the BVH8 built on the GPU satisfies the node
    invariants, its traversal equals brute force (oracle) and the GPU
    traversal equals the oracle's restatement, hits and counters."""
    from test_oracle_geometry import check_bvh8_structure
    torch = cuda
    for name, arrays in _device_build_cases():
        dev = reina_rt.DeviceScene(arrays, flags=abi.BUILD_DEVICE)
        nodes, tris = dev.export_bvh()
        ntri = int(dev.info().num_triangles)
        check_bvh8_structure(nodes, tris, ntri)
        orc = oracle.OracleScene(arrays, make_desc)
        orc.set_bvh(nodes, tris)
        rays = random_rays(4000, -1.5, 1.5, 52) if name != "instanced" else random_rays(4000, -2.0, 22.0, 52)
        hits, cnt = gpu_intersect(torch, dev, rays, counters=True)
        ref = orc.brute32(rays)
        ref_v, ref_c = orc.bvh(rays)
        assert np.array_equal(hits["prim"], ref["prim"]), name
        for f in ("t", "u", "v"):
            assert np.array_equal(_bits(hits[f]), _bits(ref[f])), (name, f)
        assert np.array_equal(cnt, ref_c), name
        dev.close()


@pytest.mark.parametrize("flags", [abi.BUILD_DEVICE, abi.BUILD_DEVICE | abi.NODES_QUANTIZED])
def test_device_bvh_render_equals_host(cuda, flags, icosphere_arrays):
    """A frame over the device-built BVH is bit-identical to the host-built
    one (the closest hit does not depend on the tree); the clustering build's
    SAH cost is within 5% of the host's binned-SAH build (measured 2.787 vs
    2.779 on this mesh)."""
    torch = cuda
    arrays, cam = icosphere_arrays
    W, H = 96, 54
    pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 2, 8)
    dh = reina_rt.DeviceScene(arrays, flags=abi.BUILD_HOST)
    dd = reina_rt.DeviceScene(arrays, flags=flags)
    a = gpu_render(torch, dh, pc, W, H)
    b = gpu_render(torch, dd, pc, W, H)
    same = (_bits(a) == _bits(b)) | (np.isnan(a) & np.isnan(b))
    assert same.all()
    print("sah host %.2f device %.2f depth %d/%d" % (dh.info().sah_cost, dd.info().sah_cost, dh.info().max_depth,
                                                     dd.info().max_depth))
    assert dd.info().sah_cost < 1.05 * dh.info().sah_cost
    dh.close()
    dd.close()


def test_intersect_f16_inverses_bitexact(cuda, icosphere_arrays):
    """REINA_RT_TRACE_F16_INV (ray-direction inverses rounded to FP16, slab
    intervals widened by 2^-10 |t|): the box test stays conservative, so the
    closest hit of every ray equals the FP32-inverse query and brute force,
    including rays along and near the axes (inverses beyond FP16's normal
    range keep FP32); the wider slabs cost extra node visits only."""
    torch = cuda
    arrays, _ = icosphere_arrays
    dev = reina_rt.DeviceScene(arrays)
    rays = random_rays(60000, -1.6, 1.6, 91)
    near = random_rays(6000, -1.6, 1.6, 92)
    for k, ax in enumerate(("dx", "dy", "dz")):          # axis-parallel and within 1e-6 of an axis
        sl = slice(2000 * k, 2000 * (k + 1))
        for other in ("dx", "dy", "dz"):
            if other != ax:
                near[other][sl] = np.where(np.arange(2000) % 2 == 0, 0.0, 1e-6).astype(np.float32)
    rays = np.concatenate([rays, near])
    h32, c32 = gpu_intersect(torch, dev, rays, counters=True)
    dev.set_trace_options(abi.TRACE_F16_INV)
    h16, c16 = gpu_intersect(torch, dev, rays, counters=True)
    dev.set_trace_options(0)
    assert np.array_equal(_bits(h16), _bits(h32))
    orc = oracle.OracleScene(arrays, make_desc)
    ref = orc.brute32(rays)
    assert np.array_equal(h16["prim"], ref["prim"]) and np.array_equal(_bits(h16["t"]), _bits(ref["t"]))
    assert c16[:, 0].sum() >= c32[:, 0].sum()
    print("node visits per ray: FP32 inverses %.3f, FP16 inverses %.3f" % (c32[:, 0].mean(), c16[:, 0].mean()))
    q = reina_rt.DeviceScene(arrays, flags=abi.NODES_QUANTIZED)
    q.set_trace_options(abi.TRACE_F16_INV)
    with pytest.raises(reina_rt.ReinaError):
        gpu_intersect(torch, q, rays[:10])
    q.close()
    dev.close()
