"""Autotuner for the traversal / shading launch parameters of one scene.

Stage 1 (tune_traversal) times reina_rt_intersect on a caller-supplied ray
batch (device buffers) over a grid of (LDS stack depth, refill threshold,
persistent blocks per CU, register budget in waves per SIMD) with HIP events
on the given stream, and installs the fastest setting on the scene with
reina_rt_set_tuning. The stack depth must cover the BVH depth, so candidates
start at the scene's own default.

Stage 2 (tune_frame) times whole frames through reina_rt_trace for a few
candidate settings (the best of stage 1, the defaults, each with several
tail-launch thresholds) and installs the fastest: a frame's bounce launches
shrink as pixels finish, so the best setting for one large batch is not
always the best for a frame.

tune_server searches a server's launch parameters (the ReinaTuning fields,
the slot tile, the graph length and the admission order) for a batch of
requests over a scene pool, one parameter at a time from the defaults, by
the wall time of the whole batch through one server.
"""
import time
import itertools


def tune_traversal(dev, torch, rays_ptr, hits_ptr, n, stream=None, cus=None, reps=3,
                   refills=(8, 16, 32, 64), blocks_per_cu=(4, 8), depths=None, waves=(0, 4, 5, 6, 8)):
    stream = stream or torch.cuda.current_stream()
    sp = stream.cuda_stream
    if cus is None:
        cus = torch.cuda.get_device_properties(0).multi_processor_count
    base = dev.get_tuning()
    if depths is None:
        depths = [d for d in (8, 16, 24, 32) if d >= base["stack_depth"]][:2]
    table = []
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for sd, rf, bpc, wv in itertools.product(depths, refills, blocks_per_cu, waves):
        dev.set_tuning(stack_depth=sd, refill_min=rf, grid_blocks=cus * bpc, waves_per_simd=wv)
        dev.intersect(rays_ptr, hits_ptr, n, stream=sp)   # warm
        best = None
        for _ in range(reps):
            e0.record(stream)
            dev.intersect(rays_ptr, hits_ptr, n, stream=sp)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None else min(best, ms)
        table.append({"stack_depth": sd, "refill_min": rf, "blocks_per_cu": bpc, "waves_per_simd": wv,
                      "ms": round(best, 4)})
    win = min(table, key=lambda r: r["ms"])
    dev.set_tuning(stack_depth=win["stack_depth"], refill_min=win["refill_min"],
                   grid_blocks=cus * win["blocks_per_cu"], waves_per_simd=win["waves_per_simd"])
    return win, table


def tune_frame(dev, torch, frame, candidates, reps=2):
    """frame() renders one frame on dev's current tuning (enqueued on the
    current stream). Each candidate is a dict of reina_rt_set_tuning fields;
    the fastest (min over reps of the synchronized wall time) is installed.
    Returns (winner, table)."""
    table = []
    for c in candidates:
        dev.set_tuning(**c)
        frame()
        torch.cuda.synchronize()
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            frame()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            best = ms if best is None else min(best, ms)
        table.append(dict(c, ms=round(best, 3)))
    win = min(table, key=lambda r: r["ms"])
    dev.set_tuning(**{k: v for k, v in win.items() if k != "ms"})
    return win, table


# The first value of each axis is the library's default, where the search starts.
SERVER_AXES = (("refill_min", (32, 16)), ("grid_blocks", ("4cu", "8cu")), ("waves_per_simd", (0, 4, 8)),
               ("tail_paths", (1 << 20, 1 << 19, 1 << 21)), ("claim_rays", (128, 64, 256)),
               ("slot_tile", (1024, 256, 4096)), ("graph_iters", (6, 12, 24)), ("admit_order", (0, 2)),
               ("tail_cu_percent", (0, 50, 100)))


def server_stack_depth(depths):
    """Smallest server stack depth (8, 16, 24, 32) that serves scenes of
    these BVH depths (a scene needs depth - 1 entries)."""
    need = max(max(depths) - 1, 1)
    return next(d for d in (8, 16, 24, 32) if d >= need)


def tune_server(torch, scenes_reqs, server_kw, cus=None, reps=1, axes=SERVER_AXES):
    """scenes_reqs: list of (DeviceScene, push constants, W, H), rendered as
    one batch (all submitted at once) per candidate. server_kw: the fixed
    Server arguments (path_budget_bytes, ...). Coordinate search: start from
    the defaults, sweep one axis at a time holding the others at the best so
    far, keep a value only if the batch is faster. Returns (best, default,
    table); best and default are Server keyword dicts plus "ms"."""
    from . import Server
    if cus is None:
        cus = torch.cuda.get_device_properties(0).multi_processor_count
    sd = server_stack_depth([int(d.info().max_depth) for d, _, _, _ in scenes_reqs])
    imgs = [torch.empty((H, W, 4), dtype=torch.float32, device="cuda") for _, _, W, H in scenes_reqs]
    torch.cuda.synchronize()

    def run(cand):
        kw = dict(server_kw, **cand)
        best = None
        srv = Server(**kw)
        try:
            for _ in range(reps):
                t0 = time.perf_counter()
                tks = [srv.submit(d, pc, img.data_ptr(), W, H) for (d, pc, W, H), img in zip(scenes_reqs, imgs)]
                for t in tks:
                    srv.wait(t)
                ms = (time.perf_counter() - t0) * 1e3
                best = ms if best is None else min(best, ms)
        finally:
            srv.close()
        return round(best, 2)

    def val(k, v):
        return cus * int(v[0]) if k == "grid_blocks" else v

    cur = {"stack_depth": sd, "shade_blocks": cus * 16}
    for k, vals in axes:
        cur[k] = val(k, vals[0])
    default = dict(cur, ms=run(cur))
    table = [default]
    best = default
    for k, vals in axes:
        for v in vals[1:]:
            c = {kk: vv for kk, vv in best.items() if kk != "ms"}
            c[k] = val(k, v)
            r = dict(c, ms=run(c))
            table.append(r)
            if r["ms"] < best["ms"]:
                best = r
    return best, default, table
