"""Load generation for the batching and serving modes, with the plain
(one reina_rt_trace per request) path measured the same way beside them.

A request is one render: (scene, camera, width, height, spp, depth), i.e. one
vkCmdTraceRaysKHR of the reference. Every request's image from the server is
compared bit for bit with the plain image of the same request.

  make_pool(sizes, seed)            resident scenes (built in parallel threads)
  batched(torch, pool, reqs)        all requests at t = 0: plain one after
                                    another vs one server run
  served(torch, pool, reqs, rate)   open-loop Poisson arrivals with a bounded
                                    FIFO (overload policy: reject when full)
"""
import concurrent.futures
import time

import numpy as np

from . import DeviceScene, ReinaError, Server, abi, scenes
from .scene import push_consts


def make_pool(sizes, seed=0, quantized=False, threads=8, device_build=True):
    """Resident scenes of the given triangle counts (pool_scene), created
    concurrently; device_build False: BVH8 built on the host (REINA_RT_BUILD_HOST).
    Returns a list of dicts {dev, cam, tris}."""
    flags = (abi.NODES_QUANTIZED if quantized else 0) | (abi.BUILD_DEVICE if device_build else abi.BUILD_HOST)

    def one(k):
        sc, cam = scenes.pool_scene(int(sizes[k]), seed * 1000 + k)
        arrays = sc.build_arrays()
        dev = DeviceScene(arrays, flags=flags)
        return {"dev": dev, "cam": cam, "tris": int(dev.info().num_triangles), "depth": int(dev.info().max_depth)}
    with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, range(len(sizes))))


def log_uniform(rng, lo, hi, n):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n))


def make_requests(pool, n, rng, resolutions, spp_choices=None, spp_range=None, depth=8, orbit=True):
    reqs = []
    for _ in range(n):
        k = int(rng.integers(len(pool)))
        W, H = resolutions[int(rng.integers(len(resolutions)))]
        spp = int(spp_choices[int(rng.integers(len(spp_choices)))]) if spp_choices is not None else \
            int(rng.integers(spp_range[0], spp_range[1] + 1))
        cam = scenes.orbit_camera(pool[k]["cam"], rng) if orbit else pool[k]["cam"]
        pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, spp, depth)
        reqs.append({"scene": k, "W": W, "H": H, "spp": spp, "pc": pc})
    return reqs


def _plain_one(torch, pool, r, img):
    dev = pool[r["scene"]]["dev"]
    dev.trace(r["pc"], img.data_ptr(), r["W"], r["H"], stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return dev.stats().extension_rays


def _pcts(lat):
    if not lat:
        return {}
    a = np.asarray(lat) * 1e3
    return {"p50_ms": round(float(np.percentile(a, 50)), 2), "p99_ms": round(float(np.percentile(a, 99)), 2),
            "mean_ms": round(float(a.mean()), 2), "max_ms": round(float(a.max()), 2)}


def batched(torch, pool, reqs, server_kw, check=True, plain=True):
    """Every request submitted at once. Plain: reina_rt_trace one after
    another. Batched: one server run. Same images required. plain=False:
    the server run only (no "plain" entry, no image comparison)."""
    rays_p, t_plain, imgs_p = 0, None, None
    if plain:
        imgs_p = [torch.zeros((r["H"], r["W"], 4), dtype=torch.float32, device="cuda") for r in reqs]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r, img in zip(reqs, imgs_p):
            rays_p += _plain_one(torch, pool, r, img)
        t_plain = time.perf_counter() - t0
    imgs_b = [torch.zeros((r["H"], r["W"], 4), dtype=torch.float32, device="cuda") for r in reqs]
    torch.cuda.synchronize()
    srv = Server(**server_kw)
    st0 = srv.stats()
    t0 = time.perf_counter()
    tickets = [srv.submit(pool[r["scene"]]["dev"], r["pc"], img.data_ptr(), r["W"], r["H"])
               for r, img in zip(reqs, imgs_b)]
    sts = [srv.wait(t) for t in tickets]
    t_batch = time.perf_counter() - t0
    st1 = srv.stats()
    srv.close()
    same = None
    if check and plain:
        same = all(bool(torch.equal(a.view(torch.int32), b.view(torch.int32))) for a, b in zip(imgs_p, imgs_b))
    rays_b = st1.extension_rays - st0.extension_rays
    lat = [s.done_s - s.submit_s for s in sts]
    out = {
        "requests": len(reqs),
        "pixels": int(sum(r["W"] * r["H"] for r in reqs)),
        "paths": int(sum(r["W"] * r["H"] * r["spp"] for r in reqs)),
        "batched": {"seconds": round(t_batch, 3), "images_per_s": round(len(reqs) / t_batch, 3),
                    "mrays_per_s": round(rays_b / t_batch / 1e6, 1), "extension_rays": int(rays_b),
                    "iterations": int(st1.iterations - st0.iterations),
                    "tail_launches": int(st1.tail_launches - st0.tail_launches),
                    "tail_rays": int(st1.tail_rays - st0.tail_rays), "latency": _pcts(lat)},
    }
    if plain:
        out["plain"] = {"seconds": round(t_plain, 3), "images_per_s": round(len(reqs) / t_plain, 3),
                        "mrays_per_s": round(rays_p / t_plain / 1e6, 1), "extension_rays": int(rays_p)}
        out["speedup_images_per_s"] = round(t_plain / t_batch, 3)
        out["images_bit_identical"] = same
    return out


def arrivals(rate, duration, rng):
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append(t)


def served_plain(torch, pool, reqs, times, max_pending):
    """One worker, FIFO of at most max_pending waiting requests; an arrival
    that finds it full is rejected. The worker renders synchronously, so the
    arrivals during a render are admitted (or rejected) in order when it
    returns, exactly as at their arrival times."""
    t0 = time.perf_counter()
    q, lat, done_t, rejected, rays = [], [], [], 0, 0
    nxt = 0
    while nxt < len(reqs) or q:
        now = time.perf_counter() - t0
        while nxt < len(reqs) and times[nxt] <= now:
            if len(q) >= max_pending:
                rejected += 1
            else:
                q.append(nxt)
            nxt += 1
        if not q:
            if nxt < len(reqs):
                time.sleep(max(0.0, min(0.002, times[nxt] - now)))
            continue
        k = q.pop(0)
        r = reqs[k]
        img = torch.empty((r["H"], r["W"], 4), dtype=torch.float32, device="cuda")
        rays += _plain_one(torch, pool, r, img)
        end = time.perf_counter() - t0
        lat.append(end - times[k])
        done_t.append(end)
    span = (max(done_t) - times[0]) if done_t else 0.0
    return {"completed": len(lat), "rejected": rejected,
            "images_per_s": round(len(lat) / span, 3) if span else 0.0,
            "mrays_per_s": round(rays / span / 1e6, 1) if span else 0.0, "latency": _pcts(lat)}


def served_server(torch, pool, reqs, times, max_pending, server_kw, keep_images=False):
    """Arrivals submitted at their times to a server with a bounded FIFO of
    max_pending (BUSY = rejected); latency = completion - arrival.
    keep_images: the result also holds {request index: image} of every
    completed request under "images"."""
    srv = Server(max_pending=max_pending, **server_kw)
    st0 = srv.stats()
    t0 = time.perf_counter()
    live, lat, done_t, rejected, images = {}, [], [], 0, {}

    def reap(block):
        for tk in list(live):
            if block:
                st = srv.wait(tk)
            else:
                st = srv.poll(tk)
                if st.state != abi.REQ_DONE:
                    continue
            k, _img = live.pop(tk)
            if keep_images:
                images[k] = _img
            # server clock -> bench clock: both measure from their own start; use the request's
            # time in the server (done - submit) plus the submit delay after its arrival
            lat.append(st.done_s - st.submit_s + sub_delay[k])
            done_t.append(times[k] + lat[-1])

    sub_delay = {}
    for k, r in enumerate(reqs):
        now = time.perf_counter() - t0
        if times[k] > now:
            time.sleep(times[k] - now)
        img = torch.empty((r["H"], r["W"], 4), dtype=torch.float32, device="cuda")
        try:
            tk = srv.submit(pool[r["scene"]]["dev"], r["pc"], img.data_ptr(), r["W"], r["H"])
            sub_delay[k] = max(0.0, (time.perf_counter() - t0) - times[k])
            live[tk] = (k, img)
        except ReinaError as e:
            if e.code != abi.ERR_BUSY:
                raise
            rejected += 1
        if k % 8 == 0:
            reap(False)
    reap(True)
    st1 = srv.stats()
    srv.close()
    span = (max(done_t) - times[0]) if done_t else 0.0
    rays = st1.extension_rays - st0.extension_rays
    out = {"completed": len(lat), "rejected": rejected,
           "images_per_s": round(len(lat) / span, 3) if span else 0.0,
           "mrays_per_s": round(rays / span / 1e6, 1) if span else 0.0, "latency": _pcts(lat)}
    if keep_images:
        out["images"] = images
    return out


def served(torch, pool, reqs, rate, duration, max_pending, server_kw, seed=0, variants=None):
    """Open-loop Poisson arrivals at `rate` per second for `duration`
    seconds (requests reqs[0], reqs[1], ... in arrival order), plain and
    served; offered vs achieved rates and latency percentiles. variants:
    {name: server keyword dict}, each served the same arrivals too."""
    rng = np.random.default_rng(seed)
    times = arrivals(rate, duration, rng)
    reqs = reqs[:len(times)]
    plain = served_plain(torch, pool, reqs, times, max_pending)
    srv = served_server(torch, pool, reqs, times, max_pending, server_kw)

    def ratio(x):
        return round(x["images_per_s"] / plain["images_per_s"], 3) if plain["images_per_s"] else None
    out = {"rate_per_s": rate, "duration_s": duration, "offered": len(times), "max_pending": max_pending,
           "overload_policy": "bounded queue of max_pending waiting requests; arrivals that find it full "
                              "are rejected (reina_rt_submit returns REINA_RT_ERR_BUSY)",
           "plain": plain, "served": srv, "speedup_images_per_s": ratio(srv)}
    for name, kw in (variants or {}).items():
        v = served_server(torch, pool, reqs, times, max_pending, kw)
        v["speedup_images_per_s"] = ratio(v)
        out[name] = v
    return out
