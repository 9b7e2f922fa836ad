"""Batching / serving through the C ABI (reina_rt_server_*, reina_rt_submit,
reina_rt_wait): each request's image must be bit-identical to the same
request rendered by reina_rt_trace, alone or mixed with other scenes,
resolutions, sample counts and accumulation states; admission respects the
path budget and the bounded queue."""
import numpy as np
import pytest

import oracle
import reina_rt
from reina_rt import abi, scenes
from reina_rt.scene import make_desc, push_consts

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _same(a, b):
    return bool(((_bits(a) == _bits(b)) | (np.isnan(a) & np.isnan(b))).all())


def plain(torch, dev, pc, W, H, init=None):
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda") if init is None else \
        torch.from_numpy(init.copy()).cuda()
    dev.trace(pc, img.data_ptr(), W, H, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return img.cpu().numpy()


def served(torch, srv, jobs):
    """jobs: list of (dev, pc, W, H, init). Submits all, then waits for all."""
    imgs, tickets = [], []
    for dev, pc, W, H, init in jobs:
        img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda") if init is None else \
            torch.from_numpy(init.copy()).cuda()
        torch.cuda.synchronize()
        imgs.append(img)
        tickets.append(srv.submit(dev, pc, img.data_ptr(), W, H))
    sts = [srv.wait(t) for t in tickets]
    for st in sts:
        assert st.state == abi.REQ_DONE and st.done_s >= st.admit_s >= st.submit_s
    return [i.cpu().numpy() for i in imgs]


@pytest.fixture(scope="module")
def pool(cuda):
    c_scene, c_cam = scenes.cornell()
    i_scene, i_cam = scenes.icosphere(levels=4)
    n_scene, n_cam = scenes.instanced(copies=64, box=20.0)
    out = {}
    for name, (sc, cam), flags in (("cornell", (c_scene, c_cam), 0), ("ico", (i_scene, i_cam), 0),
                                   ("inst", (n_scene, n_cam), 0)):
        arrays = sc.build_arrays()
        out[name] = (arrays, cam, reina_rt.DeviceScene(arrays, flags=flags))
    yield out
    for _, _, d in out.values():
        d.close()


def _pc(cam, W, H, spp, depth, batch=0):
    return push_consts(cam["eye"], cam["center"], cam["fov"], W, H, spp, depth, sample_batch=batch)


def test_server_request_equals_trace_and_oracle(cuda, pool):
    torch = cuda
    arrays, cam, dev = pool["cornell"]
    W = H = 48
    pc = scenes.cornell_push(W, H, spp=8, max_bounces=8)
    srv = reina_rt.Server(path_budget_bytes=64 << 20)
    got = served(torch, srv, [(dev, pc, W, H, None)])[0]
    ref = plain(torch, dev, pc, W, H)
    assert _same(got, ref)
    orc, _ = oracle.OracleScene(arrays, make_desc).render(pc, W, H, use_bvh=False)
    assert _same(got, orc)
    st = srv.stats()
    assert st.completed == 1 and st.extension_rays > 0 and st.iterations % 6 == 0
    srv.close()


def _mix_jobs(pool):
    rng = np.random.default_rng(3)
    jobs = []
    for k, (name, W, H, spp, depth, batch) in enumerate([
            ("cornell", 64, 64, 4, 8, 0), ("ico", 150, 100, 2, 8, 0), ("inst", 96, 54, 3, 12, 0),
            ("cornell", 33, 17, 5, 3, 2), ("ico", 64, 48, 1, 8, 5), ("inst", 128, 72, 2, 6, 0),
            ("cornell", 200, 120, 1, 8, 0), ("ico", 40, 40, 6, 8, 1)]):
        arrays, cam, dev = pool[name]
        pc = scenes.cornell_push(W, H, spp=spp, max_bounces=depth, sample_batch=batch) if name == "cornell" \
            else _pc(cam, W, H, spp, depth, batch)
        init = rng.uniform(0, 1, (H, W, 4)).astype(np.float32) if batch else None
        jobs.append((dev, pc, W, H, init))
    return jobs


@pytest.mark.parametrize("tail", [0, abi.TAIL_OFF, 3000], ids=["tail_default", "tail_off", "tail_3000"])
def test_server_mix_equals_alone(cuda, pool, tail):
    """Eight requests on three scenes (different resolutions including a
    partial last tile, spp, depths, accumulation with sampleBatch > 0)
    submitted together: each image equals its reina_rt_trace image and its
    image when served alone, whether the shared queues end with the tail
    launch at once (default threshold 2^20 paths), never, or after some
    graph iterations (3000 paths)."""
    torch = cuda
    jobs = _mix_jobs(pool)
    srv = reina_rt.Server(path_budget_bytes=256 << 20, tail_paths=tail)
    mixed = served(torch, srv, jobs)
    st_mix = srv.stats()
    for (dev, pc, W, H, init), got in zip(jobs, mixed):
        assert _same(got, plain(torch, dev, pc, W, H, init))
    for j in (1, 3):
        alone = served(torch, srv, [jobs[j]])[0]
        assert _same(alone, mixed[j])
    st = srv.stats()
    assert st.completed == len(jobs) + 2 and st.pending == 0 and st.running == 0 and st.slot_tiles_used == 0
    if tail == abi.TAIL_OFF:
        assert st.tail_launches == 0 and st.tail_rays == 0
    else:
        assert st_mix.tail_launches >= 1 and 0 < st_mix.tail_rays < st_mix.extension_rays
    srv.close()


@pytest.mark.parametrize("tile,iters,claim,tail", [(256, 12, 64, 3000), (4096, 24, 512, 0), (512, 18, 256, abi.TAIL_OFF)],
                         ids=["tile256_g12_c64", "tile4096_g24_c512", "tile512_g18_c256"])
def test_server_tunables_equal_trace(cuda, pool, tile, iters, claim, tail):
    """The server's slot tile, graph length and ray-claim chunk change only
    the schedule: the eight-request mix still gives every request its
    reina_rt_trace image, the path pool is counted in tiles of that size,
    and iterations advance a graph at a time."""
    torch = cuda
    jobs = _mix_jobs(pool)
    budget = 256 << 20
    srv = reina_rt.Server(path_budget_bytes=budget, slot_tile=tile, graph_iters=iters, claim_rays=claim,
                          tail_paths=tail)
    got = served(torch, srv, jobs)
    st = srv.stats()
    srv.close()
    for (dev, pc, W, H, init), g in zip(jobs, got):
        assert _same(g, plain(torch, dev, pc, W, H, init))
    tiles = budget // (abi.SERVER_BYTES_PER_PIXEL * tile)
    assert st.slot_tiles_total == tiles - tiles % 8 and st.slot_tiles_used == 0
    assert st.iterations == iters * st.graph_replays and st.completed == len(jobs)


@pytest.mark.parametrize("order,max_wait_ms,expect", [(abi.ADMIT_SHORTEST, 60000, [1, 2, 0]),
                                                      (abi.ADMIT_LONGEST, 60000, [2, 0, 1]),
                                                      (abi.ADMIT_SHORTEST, 1, [0, 1, 2])],
                         ids=["shortest", "longest", "aged_fifo"])
def test_server_admit_order(cuda, pool, order, max_wait_ms, expect):
    """REINA_RT_ADMIT_SHORTEST / _LONGEST with admit_paths = 1 (one request
    at a time): three requests submitted while a long one runs are admitted
    in order of estimated work (pixels x spp x depth), or of sample chain
    (spp x depth), when none has waited max_wait_ms; once all have waited that
    long, in submission order. Images equal reina_rt_trace's either way."""
    import time
    torch = cuda
    _, icam, idev = pool["ico"]
    jobs = []   # work: 96*96*4*8 = 294912, 32*32*2*8 = 16384, 64*48*3*12 = 110592; chain: 32, 16, 36
    for name, W, H, spp, depth in (("ico", 96, 96, 4, 8), ("cornell", 32, 32, 2, 8), ("inst", 64, 48, 3, 12)):
        _, cam, dev = pool[name]
        pc = scenes.cornell_push(W, H, spp=spp, max_bounces=depth) if name == "cornell" else _pc(cam, W, H, spp, depth)
        jobs.append((dev, pc, W, H))
    imgs = [torch.zeros((H, W, 4), dtype=torch.float32, device="cuda") for _, _, W, H in jobs]
    bimg = torch.zeros((1024, 1024, 4), dtype=torch.float32, device="cuda")
    bpc = _pc(icam, 1024, 1024, 64, 8)
    torch.cuda.synchronize()   # (a device-wide sync after this point would also wait for the server's stream)
    # tail at 2^19 paths: the long request (2^20 pixels) stays in the shared
    # queues until half its pixels are done, so the three wait behind it
    srv = reina_rt.Server(path_budget_bytes=256 << 20, admit_paths=1, admit_order=order, max_wait_ms=max_wait_ms,
                          tail_paths=1 << 19)
    btk = srv.submit(idev, bpc, bimg.data_ptr(), 1024, 1024)
    for _ in range(5000):
        if srv.poll(btk).state == abi.REQ_RUNNING:
            break
        time.sleep(0.001)
    tks = [srv.submit(dev, pc, img.data_ptr(), W, H) for (dev, pc, W, H), img in zip(jobs, imgs)]
    bst = srv.wait(btk)
    sts = [srv.wait(t) for t in tks]
    srv.close()
    assert all(s.submit_s < bst.done_s for s in sts)   # all were waiting while the long request ran
    got = sorted(range(3), key=lambda k: sts[k].admit_s)
    assert got == expect, [s.admit_s for s in sts]
    for (dev, pc, W, H), img in zip(jobs, imgs):
        assert _same(img.cpu().numpy(), plain(torch, dev, pc, W, H))


def test_server_admit_paths(cuda, pool):
    """With admit_paths = 1 a request is admitted only when nothing runs in
    the shared queues (tail launches off here: a tail launch's paths are not
    counted, so with tails the next request starts beside it): requests
    submitted together run one after another (each admitted after the
    previous one completed), and every image is still its reina_rt_trace
    image."""
    torch = cuda
    jobs = []
    for name, W, H, spp in (("ico", 64, 48, 2), ("cornell", 40, 40, 3), ("inst", 48, 27, 2), ("ico", 32, 32, 1)):
        arrays, cam, dev = pool[name]
        pc = scenes.cornell_push(W, H, spp=spp, max_bounces=8) if name == "cornell" else _pc(cam, W, H, spp, 8)
        jobs.append((dev, pc, W, H, None))
    srv = reina_rt.Server(path_budget_bytes=256 << 20, admit_paths=1, tail_paths=abi.TAIL_OFF)
    imgs, tickets = [], []
    for dev, pc, W, H, _ in jobs:
        img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        imgs.append(img)
        tickets.append(srv.submit(dev, pc, img.data_ptr(), W, H))
    sts = [srv.wait(t) for t in tickets]
    srv.close()
    for a, b in zip(sts, sts[1:]):
        assert b.admit_s >= a.done_s, (a.done_s, b.admit_s)
    for (dev, pc, W, H, _), img in zip(jobs, imgs):
        assert _same(img.cpu().numpy(), plain(torch, dev, pc, W, H))


def test_server_admission_by_budget(cuda, pool):
    """A path budget of 8 tiles (8192 pixels): a request needing more is
    rejected at submit; 6 requests of 4 tiles run at most two at a time
    and all match reina_rt_trace."""
    torch = cuda
    arrays, cam, dev = pool["ico"]
    srv = reina_rt.Server(path_budget_bytes=8 * 1024 * abi.SERVER_BYTES_PER_PIXEL)
    big = torch.zeros((100, 100, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(reina_rt.ReinaError) as e:
        srv.submit(dev, _pc(cam, 100, 100, 1, 8), big.data_ptr(), 100, 100)
    assert e.value.code == abi.ERR_OVER_BUDGET
    jobs = [(dev, _pc(cam, 64, 64, 2, 8, b), 64, 64, None) for b in range(6)]
    got = served(torch, srv, jobs)
    for (d, pc, W, H, init), g in zip(jobs, got):
        assert _same(g, plain(torch, d, pc, W, H))
    st = srv.stats()
    assert st.rejected_budget == 1 and st.completed == 6 and st.slot_tiles_total == 8
    srv.close()


def test_server_bounded_queue(cuda, pool):
    """max_pending = 2 and a budget of one request at a time: while a long
    request runs, further submits beyond two queued ones return BUSY."""
    torch = cuda
    arrays, cam, dev = pool["ico"]
    srv = reina_rt.Server(path_budget_bytes=8 * 1024 * abi.SERVER_BYTES_PER_PIXEL, max_pending=2)
    imgs = [torch.zeros((90, 90, 4), dtype=torch.float32, device="cuda") for _ in range(6)]
    tickets, busy = [], 0
    for k, img in enumerate(imgs):
        try:
            tickets.append(srv.submit(dev, _pc(cam, 90, 90, 64 if k == 0 else 1, 8), img.data_ptr(), 90, 90))
        except reina_rt.ReinaError as e:
            assert e.code == abi.ERR_BUSY
            busy += 1
    assert busy >= 2
    for t in tickets:
        assert srv.wait(t).state == abi.REQ_DONE
    assert srv.stats().rejected_busy == busy
    srv.close()


def test_server_max_requests(cuda, pool):
    """max_requests = 3 bounds queued + running requests: with one long
    request running (admit_paths = 1 keeps the others queued), two more are
    accepted and the fourth submit returns BUSY, though max_pending is 16;
    once all complete the server accepts again."""
    import time
    torch = cuda
    arrays, cam, dev = pool["ico"]
    imgs = [torch.zeros((512, 512, 4) if k == 0 else (256, 256, 4), dtype=torch.float32, device="cuda")
            for k in range(5)]
    pcs = [_pc(cam, 512, 512, 256, 8) if k == 0 else _pc(cam, 256, 256, 1, 8) for k in range(5)]
    torch.cuda.synchronize()   # (a device-wide sync after this point would also wait for the server's stream)
    srv = reina_rt.Server(path_budget_bytes=256 << 20, max_pending=16, max_requests=3, admit_paths=1)
    t0 = srv.submit(dev, pcs[0], imgs[0].data_ptr(), 512, 512)
    for _ in range(5000):
        if srv.poll(t0).state == abi.REQ_RUNNING:
            break
        time.sleep(0.001)
    tickets = [t0, srv.submit(dev, pcs[1], imgs[1].data_ptr(), 256, 256),
               srv.submit(dev, pcs[2], imgs[2].data_ptr(), 256, 256)]
    with pytest.raises(reina_rt.ReinaError) as e:
        srv.submit(dev, pcs[3], imgs[3].data_ptr(), 256, 256)
    assert e.value.code == abi.ERR_BUSY and "max_requests" in str(e.value)
    assert srv.poll(t0).state == abi.REQ_RUNNING   # the bound was hit while the long request ran
    for t in tickets:
        assert srv.wait(t).state == abi.REQ_DONE
    t4 = srv.submit(dev, pcs[4], imgs[4].data_ptr(), 256, 256)
    assert srv.wait(t4).state == abi.REQ_DONE
    assert srv.stats().rejected_busy == 1
    srv.close()


def test_server_edge_requests(cuda, pool):
    """spp = 0 and maxBounces = 0 requests complete (every sample is 0, as in
    reina_rt_trace); a quantized scene is refused by an FP32-node server."""
    torch = cuda
    arrays, cam, dev = pool["cornell"]
    srv = reina_rt.Server(path_budget_bytes=32 << 20)
    jobs = [(dev, scenes.cornell_push(32, 32, spp=0, max_bounces=8), 32, 32, None),
            (dev, scenes.cornell_push(32, 32, spp=3, max_bounces=0), 32, 32, None),
            (dev, scenes.cornell_push(32, 32, spp=2, max_bounces=1), 32, 32, None)]
    got = served(torch, srv, jobs)
    for (d, pc, W, H, init), g in zip(jobs, got):
        assert _same(g, plain(torch, d, pc, W, H))
    q = reina_rt.DeviceScene(arrays, flags=abi.NODES_QUANTIZED)
    img = torch.zeros((8, 8, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(reina_rt.ReinaError) as e:
        srv.submit(q, scenes.cornell_push(8, 8, spp=1, max_bounces=2), img.data_ptr(), 8, 8)
    assert e.value.code == abi.ERR_UNSUPPORTED
    srv.close()
    srvq = reina_rt.Server(path_budget_bytes=32 << 20, quantized=True)
    pc = scenes.cornell_push(40, 40, spp=4, max_bounces=8)
    got = served(torch, srvq, [(q, pc, 40, 40, None)])[0]
    assert _same(got, plain(torch, dev, pc, 40, 40))
    srvq.close()
    q.close()


@pytest.mark.parametrize("tail", [0, abi.TAIL_OFF], ids=["tail_default", "tail_off"])
def test_server_nee_mix(cuda, pool, tail):
    """A server created with REINA_RT_SERVER_NEE: requests on scenes whose
    trace options include REINA_RT_TRACE_NEE render with NEE + MIS, the
    others without; every image equals reina_rt_trace's on the same scene
    and options, mixed in one wavefront, with and without the tail launch."""
    torch = cuda
    rng = np.random.default_rng(11)
    c_arrays, c_cam, c_dev = pool["cornell"]
    i_arrays, i_cam, i_dev = pool["ico"]
    c_nee = reina_rt.DeviceScene(c_arrays)
    c_nee.set_trace_options(abi.TRACE_NEE)
    i_nee = reina_rt.DeviceScene(i_arrays)
    i_nee.set_trace_options(abi.TRACE_NEE)
    jobs = []
    for dev, W, H, spp, depth, batch, cornell in (
            (c_nee, 64, 64, 4, 8, 0, True), (c_dev, 48, 40, 3, 8, 0, True), (i_nee, 96, 54, 2, 8, 0, False),
            (i_dev, 64, 48, 2, 6, 0, False), (c_nee, 33, 17, 5, 3, 2, True), (i_nee, 40, 40, 3, 8, 1, False)):
        pc = scenes.cornell_push(W, H, spp=spp, max_bounces=depth, sample_batch=batch) if cornell \
            else _pc(i_cam, W, H, spp, depth, batch)
        init = rng.uniform(0, 1, (H, W, 4)).astype(np.float32) if batch else None
        jobs.append((dev, pc, W, H, init))
    srv = reina_rt.Server(path_budget_bytes=256 << 20, nee=True, tail_paths=tail)
    mixed = served(torch, srv, jobs)
    for (dev, pc, W, H, init), got in zip(jobs, mixed):
        assert _same(got, plain(torch, dev, pc, W, H, init))
    # NEE changes the image (the light sample draws extra random numbers)
    assert not _same(mixed[0], plain(torch, c_dev, jobs[0][1], 64, 64))
    st = srv.stats()
    assert st.completed == len(jobs) and st.slot_tiles_used == 0
    if tail == abi.TAIL_OFF:
        assert st.tail_launches == 0
    srv.close()
    c_nee.close()
    i_nee.close()


@pytest.mark.parametrize("tail", [0, abi.TAIL_OFF], ids=["tail_default", "tail_off"])
def test_server_fixes_mix(cuda, pool, tail):
    """Requests on scenes whose trace options carry REINA_RT_TRACE_FIX_*
    corrections, mixed with requests on the same scenes without them: every
    image equals reina_rt_trace's on the same scene and options."""
    torch = cuda
    c_arrays, c_cam, c_dev = pool["cornell"]
    i_arrays, i_cam, i_dev = pool["ico"]
    c_fix = reina_rt.DeviceScene(c_arrays)
    c_fix.set_trace_options(abi.TRACE_FIXES)
    i_fix = reina_rt.DeviceScene(i_arrays)
    i_fix.set_trace_options(abi.TRACE_FIX_METAL_PDF | abi.TRACE_FIX_LENS_RNG)
    jobs = []
    for dev, W, H, spp, cam in ((c_fix, 48, 40, 4, c_cam), (c_dev, 48, 40, 4, c_cam), (i_fix, 64, 36, 2, i_cam),
                                (i_dev, 64, 36, 2, i_cam), (c_fix, 21, 13, 3, c_cam)):
        pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, spp, 8, defocus=0.05)
        jobs.append((dev, pc, W, H, None))
    srv = reina_rt.Server(path_budget_bytes=256 << 20, tail_paths=tail)
    mixed = served(torch, srv, jobs)
    for (dev, pc, W, H, init), got in zip(jobs, mixed):
        assert _same(got, plain(torch, dev, pc, W, H, init))
    assert not _same(mixed[0], mixed[1]) and not _same(mixed[2], mixed[3])
    srv.close()
    c_fix.close()
    i_fix.close()


@pytest.mark.parametrize("tail", [0, abi.TAIL_OFF], ids=["tail_default", "tail_off"])
def test_server_sample_groups_mix(cuda, pool, tail):
    """Requests on scenes traced with REINA_RT_TRACE_SAMPLE_GROUPS (4 and 16
    streams per pixel), mixed with the same scenes without: every image
    equals reina_rt_trace's on the same scene and options."""
    torch = cuda
    rng = np.random.default_rng(21)
    c_arrays, c_cam, c_dev = pool["cornell"]
    i_arrays, i_cam, i_dev = pool["ico"]
    c_grp = reina_rt.DeviceScene(c_arrays)
    c_grp.set_trace_options(abi.TRACE_SAMPLE_GROUPS(2))
    i_grp = reina_rt.DeviceScene(i_arrays)
    i_grp.set_trace_options(abi.TRACE_SAMPLE_GROUPS(4))
    jobs = []
    for dev, W, H, spp, batch, cam in ((c_grp, 48, 40, 8, 0, c_cam), (c_dev, 48, 40, 8, 0, c_cam),
                                       (i_grp, 64, 36, 16, 1, i_cam), (i_dev, 64, 36, 16, 1, i_cam),
                                       (c_grp, 21, 13, 4, 2, c_cam)):
        pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, spp, 8, sample_batch=batch)
        init = rng.uniform(0, 1, (H, W, 4)).astype(np.float32) if batch else None
        jobs.append((dev, pc, W, H, init))
    srv = reina_rt.Server(path_budget_bytes=256 << 20, tail_paths=tail)
    mixed = served(torch, srv, jobs)
    for (dev, pc, W, H, init), got in zip(jobs, mixed):
        assert _same(got, plain(torch, dev, pc, W, H, init))
    assert not _same(mixed[0], mixed[1]) and not _same(mixed[2], mixed[3])
    st = srv.stats()
    assert st.completed == len(jobs) and st.slot_tiles_used == 0
    srv.close()
    c_grp.close()
    i_grp.close()


def test_server_nee_refusals(cuda, pool):
    """A NEE request on a server without REINA_RT_SERVER_NEE is refused
    (REINA_RT_ERR_UNSUPPORTED), and one on a scene without emissive
    instances too (REINA_RT_ERR_NO_EMISSIVE), as reina_rt_trace refuses it."""
    from test_oracle_geometry import soup_arrays
    torch = cuda
    c_arrays, _, _ = pool["cornell"]
    c_nee = reina_rt.DeviceScene(c_arrays)
    c_nee.set_trace_options(abi.TRACE_NEE)
    img = torch.zeros((8, 8, 4), dtype=torch.float32, device="cuda")
    pc = scenes.cornell_push(8, 8, spp=1, max_bounces=2)
    srv = reina_rt.Server(path_budget_bytes=32 << 20)
    with pytest.raises(reina_rt.ReinaError) as e:
        srv.submit(c_nee, pc, img.data_ptr(), 8, 8)
    assert e.value.code == abi.ERR_UNSUPPORTED
    srv.close()
    dark = reina_rt.DeviceScene(soup_arrays(np.random.default_rng(2).uniform(-1, 1, (16, 9)).astype(np.float32)))
    dark.set_trace_options(abi.TRACE_NEE)
    srvn = reina_rt.Server(path_budget_bytes=32 << 20, nee=True)
    with pytest.raises(reina_rt.ReinaError) as e:
        srvn.submit(dark, pc, img.data_ptr(), 8, 8)
    assert e.value.code == abi.ERR_NO_EMISSIVE
    srvn.close()
    dark.close()
    c_nee.close()


def test_tune_server_returns_server_arguments(cuda, pool):
    """reina_rt.autotune.tune_server: a coordinate search over the given axes
    from their first values; the winner is a valid set of Server arguments
    whose batch renders the reina_rt_trace images."""
    from reina_rt import autotune
    torch = cuda
    jobs = _mix_jobs(pool)[:4]
    batch = [(dev, pc, W, H) for dev, pc, W, H, init in jobs if init is None]
    axes = (("claim_rays", (128, 64)), ("slot_tile", (1024, 256)), ("graph_iters", (6, 12)))
    best, default, table = autotune.tune_server(torch, batch, dict(path_budget_bytes=64 << 20), axes=axes)
    assert len(table) == 4 and default["claim_rays"] == 128 and default["slot_tile"] == 1024
    assert best["ms"] == min(r["ms"] for r in table)
    assert best["stack_depth"] == autotune.server_stack_depth([int(d.info().max_depth) for d, _, _, _ in batch])
    srv = reina_rt.Server(path_budget_bytes=64 << 20, **{k: v for k, v in best.items() if k != "ms"})
    got = served(torch, srv, [(d, pc, W, H, None) for d, pc, W, H in batch])
    srv.close()
    for (d, pc, W, H), g in zip(batch, got):
        assert _same(g, plain(torch, d, pc, W, H))


@pytest.mark.parametrize("inline", [False, True], ids=["tail_beside", "tail_inline"])
def test_server_admits_while_tail_runs(cuda, pool, inline):
    """A 1024x1024, 256-spp request goes to the tail launch after its first
    graph (its 2^20 pixels are at most tail_paths). A small request submitted
    while that tail runs is admitted and completes before the long one when the
    tail runs beside the loop on half of the GPU (tail_cu_percent = 50); with
    REINA_RT_SERVER_TAIL_INLINE it waits for the tail. Both images equal
    reina_rt_trace's either way."""
    import time
    torch = cuda
    _, icam, idev = pool["ico"]
    _, ccam, cdev = pool["cornell"]
    big = torch.zeros((1024, 1024, 4), dtype=torch.float32, device="cuda")
    small = torch.zeros((64, 64, 4), dtype=torch.float32, device="cuda")
    bpc = _pc(icam, 1024, 1024, 256, 8)
    spc = scenes.cornell_push(64, 64, spp=4, max_bounces=8)
    torch.cuda.synchronize()   # (a device-wide sync after this point would also wait for the server's streams)
    srv = reina_rt.Server(path_budget_bytes=512 << 20, tail_paths=1 << 20, tail_inline=inline, tail_cu_percent=50)
    tb = srv.submit(idev, bpc, big.data_ptr(), 1024, 1024)
    for _ in range(5000):
        if srv.poll(tb).state == abi.REQ_RUNNING:
            break
        time.sleep(0.001)
    time.sleep(0.1)   # its first graph takes a few ms, its tail launch about half a second or more
    if not inline:
        assert srv.stats().tail_launches == 1
    ts = srv.submit(cdev, spc, small.data_ptr(), 64, 64)
    sst = srv.wait(ts)
    bst = srv.wait(tb)
    st = srv.stats()
    srv.close()
    if inline:
        assert sst.admit_s >= bst.done_s
    else:
        assert sst.admit_s < bst.done_s and sst.done_s < bst.done_s, (sst.admit_s, sst.done_s, bst.done_s)
    assert st.completed == 2 and st.tail_rays > 0
    assert _same(small.cpu().numpy(), plain(torch, cdev, spc, 64, 64))
    assert _same(big.cpu().numpy(), plain(torch, idev, bpc, 1024, 1024))
