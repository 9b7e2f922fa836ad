"""Benchmark: Mrays/s of the wavefront path tracer on one MI355X.

Workload (one "step" = one reina_rt_trace of a full frame): displaced
icosphere, 7 subdivision levels (327,680 triangles) + ground quad + 2 quad
lights (327,686 triangles), 1,280 Disney material groups, 1920x1080,
64 samples per pixel, max depth 8, FP32 BVH8. Frames accumulate
progressively (sampleBatch = frame index), as Reina's render loop does.

value = every traced extension ray (primary + secondary) / wall time of the
K timed frames, scene already resident in HBM.

roofline: the extend kernel (BVH8 traversal) over the frame's bounce
iterations; the last few live paths (ReinaTuning.tail_paths) finish in one
tail launch that is timed separately (tail_ms) and not part of the roofline.
traffic: HBM bytes per launch of
that kernel on this workload from the committed rocprofv3 PMC summary
(profiles/k_trace_icosphere_1080p_64spp.md: FETCH_SIZE x 2 + WRITE_SIZE, scaled by a
256 MiB copy in the same runs), or null for other sizes. Algorithmic bytes per ray =
32 (ray in) + 16 (hit out) + 208 per visited node (useful bytes of the
256-byte FP32 node: 8 x 24 B boxes + 8 B child bases + 8 B slot metadata)
+ 36 per tested triangle; visits come from the counting build of the same
kernel on the same frame, kernel time from HIP events around every extend
launch on the trace's stream.

quantized_nodes: the same frame and secondary batch over 8-bit quantized
BVH8 nodes (80 bytes per node, all useful), with a check that the frame is
bit-identical to the FP32-node frame (the closest hit does not depend on the
node format).

instanced_8m: 4096 instances of a 2048-triangle sphere flattened to
8,388,608 triangles, 1920x1080, 32 spp, depth 12: frames over FP32 and 8-bit
nodes (both BVHs built on the device, bit-identical frames), a host-built
FP32 tree beside them, and the closest hits of a sampled ray set against the
FP64 oracle.

batched: 64 scenes of 1e3-2e6 triangles (log-uniform), 64 requests of
uneven resolution and spp, all submitted to one server whose launch
parameters the server autotuner chose on the same batch, against the same
requests through reina_rt_trace one after another (images bit-identical).

served: Poisson arrivals at 50-400 requests/s for 10 s each (and 4/s, below
capacity) over 32 resident scenes of 1e4-4e6 triangles. Plain renders one
request at a time with 16 queued; the server holds at most 17 requests,
queued or running, with the tail launch beside its loop: p50 / p99 latency
and images/s of both, plus shortest-first admission, no admission bound and
the inline tail launch as variants.

cpu_baseline: the oracle's C restatement of the reference shaders (BVH
traversal over the same BVH8), multithreaded over pixels, on a bounded
sample of the same frame (whole 1-spp passes for about --cpu-seconds).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "reina-vk_amd"))

def log(*args):
    """Progress to stderr (the JSON result line is the only stdout output)."""
    print("[bench %.1fs]" % (time.perf_counter() - T0), *args, file=sys.stderr, flush=True)


T0 = time.perf_counter()
METRIC = "Mrays/s + % HBM roofline at 1080p; served images/s and p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
NODE_BYTES = 208
TRI_BYTES = 36
RAY_BYTES = 48


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline duration")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", type=str, default=None,
                    help="apply this tuning instead of autotuning: stack_depth,refill_min,grid_blocks,shade_blocks,"
                         "waves_per_simd,tail_paths (the config.tuning of a default run)")
    ap.add_argument("--no-quantized", action="store_true", help="skip the 8-bit quantized-node comparison")
    ap.add_argument("--no-device-build", action="store_true",
                    help="skip the host-built BVH beside the device-built one (build time, SAH, frame)")
    ap.add_argument("--instanced-variants", type=str, default="fp32,quantized,fp32_host_build",
                    help="node formats / builds of the instanced block, comma-separated (fp32 and quantized "
                         "are built on the device, fp32_host_build on the host)")
    ap.add_argument("--no-instanced", action="store_true",
                    help="skip the 8.4M-triangle instanced scene (FP32 vs quantized nodes, error vs the FP64 oracle)")
    ap.add_argument("--fp64-rays", type=int, default=1 << 18, help="rays checked against the FP64 oracle")
    ap.add_argument("--instanced-tuning", type=str, default=None,
                    help="instanced block: apply this tuning (as --tuning) to every variant instead of autotuning")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched block")
    ap.add_argument("--batched-groups", type=int, default=2,
                    help="log2 of the sample streams per pixel of the batched block's second pass (0: none)")
    ap.add_argument("--no-served", action="store_true", help="skip the served block")
    ap.add_argument("--small-pools", action="store_true",
                    help="batched / served blocks on small pools (16 scenes of 1e3-2e5 triangles, 32 requests; "
                         "8 scenes of 1e4-4e5 triangles) instead of the full configurations (64 scenes of 1e3-2e6 "
                         "triangles, 64 requests; 32 scenes of 1e4-4e6 triangles at 50-400 requests/s)")
    ap.add_argument("--admit-paths", type=int, default=1 << 21,
                    help="served block: the server admits while fewer paths than this are in flight")
    ap.add_argument("--max-wait-ms", type=int, default=500,
                    help="served block, shortest-first variant: age at which a request goes first")
    ap.add_argument("--batched-server", type=str, default=None,
                    help="batched block: these Server arguments (k=v,...) instead of the server autotuner's")
    ap.add_argument("--served-rates", type=str, default=None,
                    help="served block: only these arrival rates (requests/s, comma-separated), 10 s each")
    ap.add_argument("--served-variants", action="store_true",
                    help="with --served-rates: also run the shortest-first / admit-all / inline-tail variants")
    ap.add_argument("--secondary-reps", type=int, default=4, help="secondary rays per primary hit")
    ap.add_argument("--calib-copy", action="store_true", help="also run a 256 MiB device copy (PMC calibration)")
    ap.add_argument("--no-ingest", action="store_true", help="skip the OBJ / PNG loader timing")
    ap.add_argument("--no-image-quality", action="store_true",
                    help="skip the RMSE of the Cornell frame against its 65,536-spp converged render")
    ap.add_argument("--no-sample-groups", action="store_true",
                    help="skip the frame with 16 independent sample streams per pixel")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (the path tracer has no multi-GPU path here); "
              "--gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import reina_rt
    from reina_rt import abi, scenes
    from reina_rt.scene import push_consts

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    W, H = a.width, a.height
    t0 = time.perf_counter()
    scene, cam = scenes.icosphere()
    arrays = scene.build_arrays()
    dev = reina_rt.DeviceScene(arrays)
    info = dev.info()
    setup_s = time.perf_counter() - t0
    log("scene ready: %d triangles, %d nodes" % (info.num_triangles, info.num_nodes))
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda:0")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def frame(batch, d=dev):
        pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, a.spp, a.depth, sample_batch=batch)
        d.trace(pc, img.data_ptr(), W, H, stream=sp)
        return d.stats()

    sec = sh = None
    if not (a.no_secondary and a.no_autotune and a.no_quantized):
        sec, sh = build_secondary(a.secondary_reps, arrays, dev, cam, W, H, torch)
    autotune = None
    if a.tuning:
        dev.set_tuning(**parse_tuning(a.tuning, "--tuning"))
    elif not a.no_autotune:
        autotune = run_autotune(dev, sec, sh, torch, frame=lambda: frame(0),
                                profile="k_trace_icosphere_1080p_64spp.md")
        os.makedirs(os.path.join(ROOT, "gpurun_out"), exist_ok=True)
        with open(os.path.join(ROOT, "gpurun_out", "autotune_table.json"), "w") as f:
            json.dump({"intersect": autotune.pop("table"), "frame": autotune.pop("frame_table", None)}, f)

    log("timing %d frames" % a.steps)
    rays, elapsed, gpu_ms = time_frames(frame, a.warmup, a.steps, torch)
    mrays = rays / elapsed / 1e6
    log("%.1f Mrays/s" % mrays)
    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / max(a.steps, 1) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "displaced icosphere L7 (327,686 tris, 1,280 Disney groups) + ground + 2 quad lights, "
                        "%dx%d, %d spp, depth %d, FP32 BVH8, progressive frames" % (W, H, a.spp, a.depth),
            "triangles": int(info.num_triangles),
            "bvh_nodes": int(info.num_nodes),
            "bvh_depth": int(info.max_depth),
            "rays_per_frame": int(rays // max(a.steps, 1)),
            "gpu_trace_ms_per_frame": round(gpu_ms / max(a.steps, 1), 3),
            "scene_setup_s": round(setup_s, 2),
        },
    }

    if autotune is not None:
        out["config"]["autotune"] = autotune
    if autotune is not None or a.tuning:
        out["config"]["tuning"] = dev.get_tuning()
    if not a.no_roofline:
        out["roofline"] = roofline(dev, frame, NODE_BYTES, torch)
        if (W, H, a.spp, a.depth) == (1920, 1080, 64, 8):
            tr, tt = profiled_traffic("k_trace_icosphere_1080p_64spp.md", dev.get_tuning())
            out["roofline"]["traffic"] = tr
            out["roofline"]["traffic_tuning"] = tt

    if not a.no_secondary:
        out["secondary_rays"] = secondary_rays(a, dev, sec, sh, torch, NODE_BYTES)
        out["secondary_rays"]["tuning"] = dev.get_tuning()

    if not a.no_quantized:
        log("quantized nodes")
        out["quantized_nodes"] = quantized_nodes(a, arrays, dev, frame, img, sec, sh, torch)

    if not a.no_image_quality:
        log("image quality")
        out["image_quality"] = image_quality(torch)

    if not a.no_sample_groups and a.spp % 16 == 0:
        # the same frame with a pixel's samples in 16 independent streams: the
        # per-pixel sample chains (the frame's critical path) are 16x shorter
        log("sample groups")
        dev.set_trace_options(abi.TRACE_SAMPLE_GROUPS(4))
        r, el, gms = time_frames(frame, 1, a.steps, torch)
        st = dev.stats()
        dev.set_trace_options(0)
        out["sample_groups"] = {
            "groups": 16, "spp_per_group": a.spp // 16,
            "mrays_per_s": round(r / el / 1e6, 2), "ms_per_step": round(el / max(a.steps, 1) * 1e3, 3),
            "iterations": int(st.iterations), "tail_rays": int(st.tail_rays),
            "note": "REINA_RT_TRACE_SAMPLE_GROUPS(4): 16 streams per pixel seeded as sample batches "
                    "sampleBatch*16+g; not the reference's sampling (one stream per pixel), so not the headline",
        }

    if not a.no_device_build:
        log("host-built BVH beside the device-built one")
        out["bvh_build"] = bvh_build_block(a, arrays, dev, frame, img, torch)

    if not a.no_batched:
        log("batched requests")
        out["batched"] = batched_block(a, torch)
    if not a.no_served:
        log("served requests")
        out["served"] = served_block(a, torch)

    if not a.no_instanced:
        log("instanced 8.4M-triangle scene")
        out["instanced_8m"] = instanced_block(a, torch)

    if not a.no_ingest:
        log("scene ingest")
        out["ingest"] = ingest_block()

    if a.calib_copy:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    if not a.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(a, arrays, dev, cam, W, H)
    dev.close()
    print(json.dumps(out), flush=True)




TUNING_KEYS = ("stack_depth", "refill_min", "grid_blocks", "shade_blocks", "waves_per_simd", "tail_paths")


def parse_tuning(text, flag):
    vals = [int(v) for v in text.split(",")]
    if len(vals) != len(TUNING_KEYS):
        sys.exit("%s takes %d comma-separated integers: %s" % (flag, len(TUNING_KEYS), ",".join(TUNING_KEYS)))
    return dict(zip(TUNING_KEYS, vals))


def profiled_traffic(name, tuning):
    """(traffic_bytes_per_launch, tuning it was measured at) from a committed
    rocprofv3 summary: the line for this tuning (stack_depth,refill_min,
    grid_blocks,shade_blocks,waves_per_simd,tail_paths); (None, None) when
    the summary has no line measured at this tuning."""
    key = ",".join(str(tuning[k]) for k in ("stack_depth", "refill_min", "grid_blocks", "shade_blocks",
                                             "waves_per_simd", "tail_paths"))
    found = []
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith("traffic_bytes_per_launch "):
                t, v = line[len("traffic_bytes_per_launch "):].split(":")
                found.append((t.strip(), int(v)))
    except OSError:
        pass
    if tuning.get("claim_rays", 128) != 128:   # profiles are taken at the default claim chunk
        return None, None
    for t, v in found:
        if t == key:
            return v, t
    return None, None

def profiled_tunings(name):
    """The tunings (dicts of TUNING_KEYS) of the traffic lines in a committed
    rocprofv3 summary."""
    out = []
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith("traffic_bytes_per_launch "):
                vals = [int(v) for v in line[len("traffic_bytes_per_launch "):].split(":")[0].split(",")]
                out.append(dict(zip(TUNING_KEYS, vals)))
    except OSError:
        pass
    return out


def run_autotune(dev, sec, sh, torch, frame=None, profile=None):
    """Stage 1 on the secondary batch; stage 2 (when `frame` is given) on
    whole frames: the 3 fastest stage-1 settings and the defaults, each with
    tail thresholds 2^18 / 2^19 / 2^20, plus the tunings the committed
    profile `profile` was taken at. Frame times within 1.5% are ties (the
    spread between runs of one tuning): when the fastest tuning has no
    profile line, the fastest profiled one among its ties is installed, so
    the run's extend kernel has a measured HBM traffic figure."""
    from reina_rt.autotune import tune_frame, tune_traversal
    d0 = dev.get_tuning()
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    win, table = tune_traversal(dev, torch, sec.data_ptr(), sh.data_ptr(), sec.shape[0], cus=cus)
    dflt = next((r for r in table if r["stack_depth"] == d0["stack_depth"] and r["refill_min"] == d0["refill_min"]
                 and r["blocks_per_cu"] * cus == d0["grid_blocks"] and r["waves_per_simd"] == 0), None)
    out = {"best": win, "default": dflt, "candidates": len(table), "worst_ms": max(r["ms"] for r in table),
           "secondary_batch_rays": int(sec.shape[0]), "table": table}
    if frame is not None:
        keys = ("stack_depth", "refill_min", "waves_per_simd")
        top = []
        for r in sorted(table, key=lambda r: r["ms"]):
            c = {k: r[k] for k in keys}
            c["grid_blocks"] = r["blocks_per_cu"] * cus
            if c not in top:
                top.append(c)
            if len(top) == 3:
                break
        base = {k: d0[k] for k in keys}
        base["grid_blocks"] = d0["grid_blocks"]
        if base not in top:
            top.append(base)
        cands = [dict(c, tail_paths=tp) for c in top for tp in (1 << 18, 1 << 19, 1 << 20)]
        shade = d0["shade_blocks"]
        for t in (profiled_tunings(profile) if profile else []):
            c = {k: t[k] for k in ("stack_depth", "refill_min", "waves_per_simd", "grid_blocks", "tail_paths")}
            if t["shade_blocks"] == shade and t["stack_depth"] >= d0["stack_depth"] and c not in cands:
                cands.append(c)
        fwin, ftable = tune_frame(dev, torch, frame, cands)
        out["frame_fastest"] = dict(fwin)
        if profile:
            cur = dev.get_tuning()
            if profiled_traffic(profile, cur)[0] is None:
                ties = [r for r in ftable if r["ms"] <= fwin["ms"] * 1.015 and
                        profiled_traffic(profile, dict(cur, **{k: v for k, v in r.items() if k != "ms"}))[0]]
                if ties:
                    fwin = min(ties, key=lambda r: r["ms"])
                    dev.set_tuning(**{k: v for k, v in fwin.items() if k != "ms"})
        out["frame_best"] = fwin
        out["frame_candidates"] = len(ftable)
        out["frame_worst_ms"] = max(r["ms"] for r in ftable)
        out["frame_table"] = ftable
    return out


def image_quality(torch):
    """RMSE of the 24x24 Cornell frame (depth 8) against its 65,536-spp FP64
    converged render (tests/golden/cornell_24x24_converged.npz), at 64 spp (one
    frame) and 1024 spp (16 progressive frames), for the reference's sampling,
    next-event estimation, and 16 independent sample streams per pixel.
    `predicted_rmse_64spp` is sqrt(mean(per-pixel variance of a 64-spp mean)):
    an unbiased 64-spp estimator with the reference's variance lands on it."""
    import numpy as np
    import reina_rt
    from reina_rt import abi, scenes
    g = np.load(os.path.join(ROOT, "tests", "golden", "cornell_24x24_converged.npz"))
    ref, var = g["mean"], g["var_batch64"]
    sc, _ = scenes.cornell()
    dev = reina_rt.DeviceScene(sc.build_arrays())
    W = H = 24
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    sp = torch.cuda.current_stream().cuda_stream
    res = {"scene": "Cornell box 24x24, depth 8, directClamp 100", "converged_spp": int(g["spp"]),
           "predicted_rmse_64spp": round(float(np.sqrt(var.mean())), 5)}
    for name, opts in (("reference_sampling", 0), ("nee", abi.TRACE_NEE), ("sample_groups_16", abi.TRACE_SAMPLE_GROUPS(4))):
        dev.set_trace_options(opts)
        row = {}
        for b in range(16):
            dev.trace(scenes.cornell_push(W, H, spp=64, max_bounces=8, sample_batch=b), img.data_ptr(), W, H, stream=sp)
            if b in (0, 15):
                torch.cuda.synchronize()
                got = img.cpu().numpy()[..., :3].astype(np.float64)
                row["rmse_%dspp" % (64 * (b + 1))] = round(float(np.sqrt(np.mean((got - ref) ** 2))), 5)
        res[name] = row
    dev.close()
    return res


def time_frames(frame, warmup, steps, torch):
    """`warmup` untimed then `steps` timed progressive frames (sampleBatch =
    frame index); returns (extension rays, wall seconds, summed GPU ms)."""
    batch = 0
    for _ in range(warmup):
        frame(batch)
        batch += 1
    torch.cuda.synchronize()
    rays, gpu_ms = 0, 0.0
    t_start = time.perf_counter()
    for _ in range(steps):
        st = frame(batch)
        rays += st.extension_rays
        gpu_ms += st.trace_ms
        batch += 1
    torch.cuda.synchronize()
    return rays, time.perf_counter() - t_start, gpu_ms


def roofline(dev, frame, node_bytes, torch):
    """Extend-kernel roofline of one frame: algorithmic bytes from the
    counting build, average launch time from HIP events around every extend
    launch on the trace's stream."""
    from reina_rt import abi
    dev.set_trace_options(abi.TRACE_COUNT)
    stc = frame(0, dev)
    dev.set_trace_options(abi.TRACE_TIME_EXTEND)
    stt = frame(0, dev)
    dev.set_trace_options(0)
    torch.cuda.synchronize()
    assert stc.extension_rays == stt.extension_rays and stc.tail_rays == stt.tail_rays
    nrays = stt.extension_rays - stt.tail_rays   # rays traced by k_trace launches (the tail kernel is not counted)
    alg_bytes = nrays * RAY_BYTES + stc.node_visits * node_bytes + stc.tri_tests * TRI_BYTES
    if stt.extend_launches == 0 or stt.extend_ms <= 0.0:
        return {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None, "kernel": "k_trace (BVH8 traversal, extend)", "launches": 0,
                "tail_rays": int(stt.tail_rays), "tail_ms": round(float(stt.tail_ms), 3),
                "note": "every ray of the frame was traced by the tail kernel"}
    launches = stt.extend_launches
    avg_launch_ms = stt.extend_ms / launches
    achieved = (alg_bytes / launches) / (avg_launch_ms * 1e-3) / 1e9
    return {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "kernel": "k_trace (BVH8 traversal, extend)",
        "node_bytes": node_bytes,
        "bytes_per_launch": int(alg_bytes // launches),
        "avg_launch_ms": round(avg_launch_ms, 4),
        "launches": int(launches),
        "extend_ms_per_frame": round(stt.extend_ms, 3),
        "extend_share_of_trace": round(stt.extend_ms / max(stt.trace_ms, 1e-9), 3),
        "nodes_per_ray": round(stc.node_visits / max(nrays, 1), 3),
        "tris_per_ray": round(stc.tri_tests / max(nrays, 1), 3),
        "extend_grays_per_s": round(nrays / (stt.extend_ms * 1e-3) / 1e9, 3),
        "tail_rays": int(stt.tail_rays),
        "tail_ms": round(float(stt.tail_ms), 3),
    }



def bvh_build_block(a, arrays, dev, frame, img, torch):
    """The headline scene's BVH8 is built on the GPU (the library's default);
    this builds it again on the host (REINA_RT_BUILD_HOST): build times, SAH
    cost and depth of both, the frame rate over the host-built tree (its own
    default tuning), and a check that the frame is bit-identical (the closest
    hit does not depend on the tree). The headline scene was the process's
    first, so its build time includes the runtime's first-use costs; the
    device build is timed again here on a second scene."""
    import reina_rt
    from reina_rt import abi
    dd = reina_rt.DeviceScene(arrays, flags=abi.BUILD_DEVICE)
    warm_ms = float(dd.info().build_ms)
    dd.close()
    t0 = time.perf_counter()
    dh = reina_rt.DeviceScene(arrays, flags=abi.BUILD_HOST)
    setup = time.perf_counter() - t0
    hi, di = dh.info(), dev.info()
    frame(0, dev)
    ref = img.clone()
    frame(0, dh)
    same = bool(torch.equal(ref.view(torch.int32), img.view(torch.int32)))
    rays, elapsed, _ = time_frames(lambda b, d=dh: frame(b, d), a.warmup, a.steps, torch)
    res = {"device_build_ms": round(warm_ms, 1), "device_first_build_ms": round(float(di.build_ms), 1),
           "host_build_ms": round(float(hi.build_ms), 1),
           "host_scene_setup_s": round(setup, 3), "device_sah_cost": round(float(di.sah_cost), 3),
           "host_sah_cost": round(float(hi.sah_cost), 3), "device_bvh_depth": int(di.max_depth),
           "host_bvh_depth": int(hi.max_depth), "device_bvh_nodes": int(di.num_nodes),
           "host_bvh_nodes": int(hi.num_nodes), "host_tree_mrays_per_s": round(rays / elapsed / 1e6, 2),
           "frame_bit_identical_to_host_build": same}
    dh.close()
    return res

def quantized_nodes(a, arrays, dev, frame, img, sec, sh, torch):
    """The same frame and secondary batch over 8-bit quantized BVH8 nodes
    (80 bytes per node). The closest hit does not depend on the node format,
    so the frame must be bit-identical to the FP32-node frame."""
    import reina_rt
    from reina_rt import abi
    t0 = time.perf_counter()
    devq = reina_rt.DeviceScene(arrays, flags=abi.NODES_QUANTIZED)
    setup_s = time.perf_counter() - t0
    res = {"scene_setup_s": round(setup_s, 2), "node_bytes": int(devq.info().node_bytes),
           "bvh_node_MB": round(devq.info().num_nodes * devq.info().node_bytes / 1e6, 2)}
    if not a.no_autotune:
        at = run_autotune(devq, sec, sh, torch, frame=lambda: frame(0, devq))
        at.pop("table")
        at.pop("frame_table", None)
        res["autotune_best"] = at["best"]
        res["autotune_frame_best"] = at.get("frame_best")
    # bit-identity of one frame against the FP32 nodes
    frame(0, dev)
    ref = img.clone()
    frame(0, devq)
    res["frame_bit_identical_to_fp32"] = bool(torch.equal(ref.view(torch.int32), img.view(torch.int32)))
    rays, elapsed, gpu_ms = time_frames(lambda b, d=devq: frame(b, d), a.warmup, a.steps, torch)
    res["mrays_per_s"] = round(rays / elapsed / 1e6, 2)
    res["ms_per_step"] = round(elapsed / max(a.steps, 1) * 1e3, 3)
    if not a.no_roofline:
        res["roofline"] = roofline(devq, frame, 80, torch)
    if not a.no_secondary:
        res["secondary_rays"] = secondary_rays(a, devq, sec, sh, torch, 80)
    res["tuning"] = devq.get_tuning()
    devq.close()
    return res



SERVER_BUDGET = 152 * (1 << 24)   # the largest path pool: 2^24 pixels in flight (2.55 GB)


def server_tuning(torch, pool, reqs, skw, what, axes=None):
    """Server launch parameters for this pool, by reina_rt.autotune.tune_server
    on a batch of the block's own requests (axes: the searched parameters,
    default all; the others stay as skw sets them)."""
    from reina_rt import autotune
    t0 = time.perf_counter()
    best, default, table = autotune.tune_server(
        torch, [(pool[r["scene"]]["dev"], r["pc"], r["W"], r["H"]) for r in reqs], skw,
        axes=axes or autotune.SERVER_AXES)
    log("%s: server autotune %.1f s, best %.1f ms vs default %.1f ms" % (
        what, time.perf_counter() - t0, best["ms"], default["ms"]))
    return {"best": best, "default": default, "candidates": len(table), "batch_requests": len(reqs),
            "seconds": round(time.perf_counter() - t0, 1)}


def batched_block(a, torch):
    """Many requests of uneven size at once: scenes of log-uniform triangle
    count, resolutions from {128^2, 256^2, 512^2, 1024^2}, spp from
    {4, 16, 64, 256}, depth 8, Disney / dielectric materials; one server run
    vs the same requests through reina_rt_trace one after another."""
    import numpy as np

    from reina_rt import abi, loadgen
    rng = np.random.default_rng(64)
    n_sc, lo, hi, n_req = (16, 1e3, 2e5, 32) if a.small_pools else (64, 1e3, 2e6, 64)
    t0 = time.perf_counter()
    pool = loadgen.make_pool(loadgen.log_uniform(rng, lo, hi, n_sc), seed=4, device_build=True)
    setup = time.perf_counter() - t0
    log("batched: %d scenes built in %.1f s" % (n_sc, setup))
    reqs = loadgen.make_requests(pool, n_req, rng, [(128, 128), (256, 256), (512, 512), (1024, 1024)],
                                 spp_choices=(4, 16, 64, 256), depth=8, orbit=False)
    skw = dict(path_budget_bytes=SERVER_BUDGET)
    tuned = None
    if a.batched_server:
        skw.update({k: int(v) for k, v in (kv.split("=") for kv in a.batched_server.split(","))})
    elif not a.no_autotune:
        tuned = server_tuning(torch, pool, reqs, skw, "batched")
        skw = dict(skw, **{k: v for k, v in tuned["best"].items() if k != "ms"})
    res = loadgen.batched(torch, pool, reqs, skw)
    res["server"] = {k: v for k, v in skw.items() if k != "path_budget_bytes"}
    if tuned is not None:
        res["server_autotune"] = tuned
        # the same batch through a server at the default launch parameters
        res["batched_default_tuning"] = loadgen.batched(torch, pool, reqs, dict(path_budget_bytes=SERVER_BUDGET),
                                                        plain=False)["batched"]
    res.update({"scenes": n_sc, "triangles_min": min(p["tris"] for p in pool),
                "triangles_max": max(p["tris"] for p in pool),
                "triangles_total": int(sum(p["tris"] for p in pool)), "scene_setup_s": round(setup, 1),
                "note": "plain = reina_rt_trace per request, one after another; batched = all requests "
                        "submitted to one server (shared queues, graph-replayed bounce iterations)"})
    if a.batched_groups > 0:
        # the same requests with 2^n independent sample streams per pixel, plain and batched alike
        from reina_rt import abi
        g = 1 << a.batched_groups
        log("batched: %d sample streams per pixel" % g)
        for p in pool:
            p["dev"].set_trace_options(abi.TRACE_SAMPLE_GROUPS(a.batched_groups))
        rg = loadgen.batched(torch, pool, reqs, skw)
        res["sample_groups_%d" % g] = {k: rg[k] for k in ("plain", "batched", "speedup_images_per_s",
                                                          "images_bit_identical") if k in rg}
        for p in pool:
            p["dev"].set_trace_options(0)
    for p in pool:
        p["dev"].close()
    return res


def served_block(a, torch):
    """Open-loop Poisson arrivals over a resident scene pool; each request a
    random orbit camera, resolution from {256^2, 512^2, 1280x720, 1920x1080},
    spp U[16, 128], depth 8. Plain (one worker running reina_rt_trace) vs
    server, both with the same bounded FIFO."""
    import numpy as np

    from reina_rt import abi, loadgen
    rng = np.random.default_rng(5)
    if a.small_pools:
        n_sc, lo, hi, runs = 8, 1e4, 4e5, [(50.0, 4.0), (4.0, 6.0)]
    else:
        n_sc, lo, hi, runs = 32, 1e4, 4e6, [(50.0, 10.0), (100.0, 10.0), (200.0, 10.0), (400.0, 10.0), (4.0, 10.0)]
    if a.served_rates:
        runs = [(float(r), 10.0) for r in a.served_rates.split(",")]
    t0 = time.perf_counter()
    pool = loadgen.make_pool(loadgen.log_uniform(rng, lo, hi, n_sc), seed=5, device_build=True)
    setup = time.perf_counter() - t0
    log("served: %d scenes built in %.1f s" % (n_sc, setup))
    out = {"scenes": n_sc, "triangles_total": int(sum(p["tris"] for p in pool)),
           "triangles_min": min(p["tris"] for p in pool), "triangles_max": max(p["tris"] for p in pool),
           "scene_setup_s": round(setup, 1), "runs": []}
    out["admit_paths"] = a.admit_paths
    res_mix = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
    # latency choices, fixed: tail launch at 2^20 live paths beside the loop, admission points every 6
    # iterations (a sweep at 50-400 requests/s gave p50 and p99 below plain's at every rate with these)
    base = dict(path_budget_bytes=SERVER_BUDGET, tail_paths=1 << 20, graph_iters=6)
    if not a.no_autotune:
        from reina_rt import autotune
        treqs = loadgen.make_requests(pool, 12, np.random.default_rng(55), res_mix, spp_range=(16, 128))
        axes = tuple(ax for ax in autotune.SERVER_AXES
                     if ax[0] not in ("tail_paths", "graph_iters", "admit_order", "tail_cu_percent"))
        out["server_autotune"] = server_tuning(torch, pool, treqs, base, "served", axes=axes)
        base = dict(base, **{k: v for k, v in out["server_autotune"]["best"].items() if k != "ms"})
    # the server holds at most 17 requests, queued or running, as plain does (1 rendering + 16 queued)
    fifo = dict(base, admit_paths=a.admit_paths, max_requests=17)
    shortest = dict(fifo, admit_order=abi.ADMIT_SHORTEST, max_wait_ms=a.max_wait_ms)
    out["served_policy"] = ("at most 17 requests queued or running (plain: 1 rendering + 16 queued); admission in "
                            "submission order while fewer than admit_paths paths are in flight (a request is "
                            "always admitted when nothing runs)")
    out["served_shortest_policy"] = ("least estimated work first (pixels x spp x depth) with the same admit_paths; "
                                     "a request waiting %d ms or longer goes first, oldest first" % a.max_wait_ms)
    for rate, dur in runs:
        reqs = loadgen.make_requests(pool, int(rate * dur * 1.5) + 16, rng, res_mix, spp_range=(16, 128))
        variants = {}
        if not a.served_rates or a.served_variants:
            if rate in (100.0, 200.0):
                variants["served_shortest"] = shortest
            if rate == 100.0:
                # no admission target and no bound beyond the 16 queued: every request that fits the pool runs
                variants["served_admit_all"] = base
                # the tail launch as one step of the loop (arrivals wait for it) instead of beside it
                variants["served_tail_inline"] = dict(fifo, tail_inline=True)
        r = loadgen.served(torch, pool, reqs, rate, dur, 16, fifo, seed=int(rate), variants=variants)
        log("served: rate %.0f/s plain %.2f img/s served %.2f img/s, p50 %.0f / %.0f ms, p99 %.0f / %.0f ms" % (
            rate, r["plain"]["images_per_s"], r["served"]["images_per_s"],
            r["plain"]["latency"].get("p50_ms", 0), r["served"]["latency"].get("p50_ms", 0),
            r["plain"]["latency"].get("p99_ms", 0), r["served"]["latency"].get("p99_ms", 0)))
        if "served_shortest" in r:
            v = r["served_shortest"]
            log("served: rate %.0f/s shortest-first %.2f img/s, p50 %.0f ms, p99 %.0f ms" % (
                rate, v["images_per_s"], v["latency"].get("p50_ms", 0), v["latency"].get("p99_ms", 0)))
        out["runs"].append(r)
    for p in pool:
        p["dev"].close()
    return out

def instanced_block(a, torch):
    """4096 instances of a 2048-triangle sphere flattened to 8,388,608
    triangles, 1920x1080, 32 spp, depth 12: frames over FP32 and over 8-bit
    quantized BVH8 nodes, and the closest hit of a sampled ray set (camera
    rays + cosine bounce rays) against the FP64 oracle."""
    import numpy as np

    import oracle
    import reina_rt
    from reina_rt import abi, scenes
    from reina_rt.scene import make_desc, push_consts

    W, H, spp, depth = 1920, 1080, 32, 12
    scene, cam = scenes.instanced()
    arrays = scene.build_arrays()
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda:0")
    sp = torch.cuda.current_stream().cuda_stream
    res = {"workload": "4096 instances x 2048-tri sphere = 8,388,608 tris, %dx%d, %d spp, depth %d" % (W, H, spp, depth)}
    devs = {}
    wanted = set(a.instanced_variants.split(",")) | {"fp32"}   # the ray set and FP64 check use fp32
    for name, flags in (("fp32", abi.BUILD_DEVICE), ("quantized", abi.BUILD_DEVICE | abi.NODES_QUANTIZED),
                        ("fp32_host_build", abi.BUILD_HOST)):
        if name not in wanted:
            continue
        t0 = time.perf_counter()
        d = reina_rt.DeviceScene(arrays, flags=flags)
        inf = d.info()
        devs[name] = d
        log("instanced: %s scene built in %.1f s" % (name, time.perf_counter() - t0))
        res[name] = {"scene_setup_s": round(time.perf_counter() - t0, 2), "node_bytes": int(inf.node_bytes),
                     "bvh_nodes": int(inf.num_nodes), "bvh_depth": int(inf.max_depth),
                     "sah_cost": round(float(inf.sah_cost), 3),
                     "scene_MB": round(inf.device_bytes / 1e6, 1),
                     "node_MB": round(inf.num_nodes * inf.node_bytes / 1e6, 1)}

    def frame(batch, d):
        pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, spp, depth, sample_batch=batch)
        d.trace(pc, img.data_ptr(), W, H, stream=sp)
        return d.stats()

    # sampled rays: camera rays + cosine bounce rays from their hits
    prim, sec, sh = build_secondary(1, arrays, devs["fp32"], cam, W, H, torch, with_primary=True)
    g = torch.Generator(device="cuda:0")
    g.manual_seed(77)
    n = a.fp64_rays // 2
    rays = torch.cat([prim[torch.randperm(prim.shape[0], device="cuda:0", generator=g)[:n]],
                      sec[torch.randperm(sec.shape[0], device="cuda:0", generator=g)[:n]]]).contiguous()
    rays_np = rays.cpu().numpy().view(abi.RAY).reshape(-1)
    hits = {}
    for name, d in devs.items():
        h = torch.zeros((rays.shape[0], 4), device="cuda:0", dtype=torch.float32)
        d.intersect(rays.data_ptr(), h.data_ptr(), rays.shape[0], stream=sp)
        torch.cuda.synchronize()
        hits[name] = h.cpu().numpy().view(abi.HIT).reshape(-1)
    orc = oracle.OracleScene(arrays, make_desc)
    nodes, tris = devs["fp32"].export_bvh()
    orc.set_bvh(nodes, tris, 256)
    log("instanced: FP64 oracle on %d rays" % len(rays_np))
    t0 = time.perf_counter()
    t64, p64, u64, v64, amb = orc.fp64(rays_np, tie_rel=1e-5, use_bvh=True, nthreads=max(1, min(os.cpu_count() or 1, 16)))
    oracle_s = time.perf_counter() - t0
    wt = torch.tensor(world_triangles(arrays), device="cuda:0", dtype=torch.float64).reshape(-1, 3, 3)
    o64 = rays[:, 0:3].double()
    dn = rays[:, 4:7].double().norm(dim=1)
    for name in devs:
        hh = hits[name]
        agree = hh["prim"] == p64
        m = agree & (p64 != 0xFFFFFFFF)
        err = np.abs(hh["t"].astype(np.float64) - t64)
        idx = torch.from_numpy(np.nonzero(m)[0]).cuda()
        D = ((wt[torch.from_numpy(p64[m].astype(np.int64)).cuda()] - o64[idx][:, None, :]).norm(dim=2).max(1).values
             / dn[idx]).cpu().numpy()
        scale = np.zeros(len(rays_np))
        scale[m] = np.maximum(t64[m], D)
        chk = {"rays": int(len(rays_np)), "hit_id_agreement": round(float(agree.mean()), 7),
               "disagreements": int((~agree).sum()), "disagreements_not_at_ties": int((~agree & ~amb).sum()),
               "ambiguous_fraction": round(float(amb.mean()), 5),
               "hit_fraction": round(float((p64 != 0xFFFFFFFF).mean()), 4)}
        for kind, sel in (("camera", np.arange(len(rays_np)) < len(rays_np) // 2),
                          ("bounce", np.arange(len(rays_np)) >= len(rays_np) // 2)):
            mm = m & sel
            rel = err[mm] / t64[mm]
            sc = err[mm] / scale[mm]
            chk[kind] = {"hits": int(mm.sum()), "t_rel_err_max": float(rel.max()),
                         "t_rel_err_p999": float(np.quantile(rel, 0.999)),
                         "frac_t_rel_err_le_1e-6": round(float((rel <= 1e-6).mean()), 6),
                         "t_err_over_vertex_scale_max": float(sc.max()),
                         "frac_t_err_over_vertex_scale_le_1e-6": round(float((sc <= 1e-6).mean()), 6)}
        res[name]["fp64_check"] = chk
    del wt
    res["fp64_oracle_s"] = round(oracle_s, 1)
    if "fp32" in hits and "quantized" in hits:
        res["hits_bit_identical_fp32_vs_quantized"] = bool(
            np.array_equal(hits["fp32"].view(np.uint32), hits["quantized"].view(np.uint32)))
    # frames
    for name, d in devs.items():
        if a.instanced_tuning:
            d.set_tuning(**parse_tuning(a.instanced_tuning, "--instanced-tuning"))
            res[name]["tuning"] = d.get_tuning()
        elif not a.no_autotune:
            at = run_autotune(d, sec, torch.zeros((sec.shape[0], 4), device="cuda:0"), torch,
                              frame=lambda d=d: frame(0, d),
                              profile="k_trace_instanced_8m_1080p_32spp.md" if name == "fp32" else None)
            at.pop("table")
            at.pop("frame_table", None)
            res[name]["autotune_best"] = at["best"]
            res[name]["autotune_frame_best"] = at.get("frame_best")
        log("instanced: %s frames" % name)
        rays_f, elapsed, gpu_ms = time_frames(lambda b, d=d: frame(b, d), 1, 2, torch)
        res[name]["mrays_per_s"] = round(rays_f / elapsed / 1e6, 2)
        res[name]["ms_per_frame"] = round(elapsed / 2 * 1e3, 2)
        res[name]["rays_per_frame"] = int(rays_f // 2)
        if not a.no_roofline:
            res[name]["roofline"] = roofline(d, frame, 80 if name == "quantized" else 208, torch)
            if name == "fp32":
                tr, tt = profiled_traffic("k_trace_instanced_8m_1080p_32spp.md", d.get_tuning())
                res[name]["roofline"]["traffic"] = tr
                res[name]["roofline"]["traffic_tuning"] = tt
    if "fp32" in devs and "quantized" in devs:
        frame(0, devs["fp32"])
        ref = img.clone()
        frame(0, devs["quantized"])
        res["frame_bit_identical_fp32_vs_quantized"] = bool(torch.equal(ref.view(torch.int32), img.view(torch.int32)))
    for d in devs.values():
        d.close()
    return res

def ingest_block(reps=3):
    """Host-side loaders (reina_rt_model_load_obj_memory, reina_rt_image_load_memory)
    on synthetic files held in memory: the icosphere L7 mesh as OBJ text with
    positions, spherical texture coordinates and normals (v/vt/vn faces), and
    a 2048x2048 RGBA PNG (gradient + noise, Paeth-filtered rows, zlib level 6).
    Best of `reps` runs; Assimp and stb are not in this image, so no
    reference-side timing is given."""
    import zlib

    import numpy as np

    from reina_rt import scenes
    from reina_rt.scene import load_image, load_obj

    v, f = scenes.subdivided_icosphere(7)
    uv = np.stack([0.5 + np.arctan2(v[:, 2], v[:, 0]) / (2 * np.pi), 0.5 - np.arcsin(v[:, 1]) / np.pi], axis=1)
    lines = ["v %.6f %.6f %.6f" % tuple(p) for p in v]
    lines += ["vt %.6f %.6f" % tuple(t) for t in uv]
    lines += ["vn %.6f %.6f %.6f" % tuple(p) for p in v]
    lines += ["f %d/%d/%d %d/%d/%d %d/%d/%d" % (a + 1, a + 1, a + 1, b + 1, b + 1, b + 1, c + 1, c + 1, c + 1)
              for a, b, c in f]
    text = ("\n".join(lines) + "\n").encode()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        md = load_obj(text)
        best = min(best, time.perf_counter() - t0)
    ntri = len(md.indices) // 3
    res = {"obj": {"file": "icosphere L7 as OBJ (v/vt/vn)", "bytes": len(text), "triangles": ntri,
                   "vertices_out": int(len(md.vertices)), "ms": round(best * 1e3, 1),
                   "mtris_per_s": round(ntri / best / 1e6, 2), "mb_per_s": round(len(text) / best / 1e6, 1)}}

    rng = np.random.default_rng(3)
    Wt = Ht = 2048
    yy, xx = np.mgrid[0:Ht, 0:Wt]
    img = np.stack([xx % 256, yy % 256, (xx + yy) % 256, np.full_like(xx, 255)], axis=-1)
    img = (img + rng.integers(0, 8, img.shape)).clip(0, 255).astype(np.uint8)
    raw = img.reshape(Ht, Wt * 4).astype(np.int16)
    up = np.vstack([np.zeros((1, Wt * 4), np.int16), raw[:-1]])
    left = np.hstack([np.zeros((Ht, 4), np.int16), raw[:, :-4]]) if Wt > 1 else np.zeros_like(raw)
    ul = np.hstack([np.zeros((Ht, 4), np.int16), up[:, :-4]])
    p = left + up - ul
    pa, pb, pc = np.abs(p - left), np.abs(p - up), np.abs(p - ul)
    pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, up, ul))
    filt = ((raw - pred) & 0xFF).astype(np.uint8)
    rows = np.hstack([np.full((Ht, 1), 4, np.uint8), filt])
    import struct

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    png = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", Wt, Ht, 8, 6, 0, 0, 0))
           + chunk(b"IDAT", zlib.compress(rows.tobytes(), 6)) + chunk(b"IEND", b""))
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        dec = load_image(png)
        best = min(best, time.perf_counter() - t0)
    assert np.array_equal(dec, img), "PNG round trip differs"
    res["png"] = {"file": "2048x2048 RGBA8, Paeth rows", "bytes": len(png), "ms": round(best * 1e3, 1),
                  "mpixels_per_s": round(Wt * Ht / best / 1e6, 1)}
    return res


def build_secondary(reps, arrays, dev, cam, W, H, torch, with_primary=False):
    """Incoherent secondary rays of the same frame: one pinhole primary ray per
    pixel, closest hit on the GPU, then `--secondary-reps` cosine-distributed
    directions about the geometric normal of every hit (origin offset 1e-4
    along the normal). The batch is traced with reina_rt_intersect: HIP
    events on the stream around the launch give the kernel time, the
    counting build on the same batch gives node visits and triangle tests."""
    import numpy as np

    from reina_rt import abi
    from reina_rt.scene import look_at, perspective

    dev_ = torch.device("cuda:0")
    sp = torch.cuda.current_stream().cuda_stream
    inv_p = torch.tensor(np.linalg.inv(perspective(np.radians(cam["fov"]), W / H, 0.1, 100.0).astype(np.float64)),
                         device=dev_)
    inv_v = torch.tensor(np.linalg.inv(look_at(cam["eye"], cam["center"], (0, 1, 0)).astype(np.float64)), device=dev_)
    ys, xs = torch.meshgrid(torch.arange(H, device=dev_, dtype=torch.float64),
                            torch.arange(W, device=dev_, dtype=torch.float64), indexing="ij")
    ndc = torch.stack([(xs + 0.5) / W * 2 - 1, -((ys + 0.5) / H * 2 - 1), -torch.ones_like(xs), torch.ones_like(xs)], -1)
    vp = ndc.reshape(-1, 4) @ inv_p.T
    vdir = vp[:, :3] / vp[:, 3:4]
    wdir = torch.nn.functional.normalize(vdir @ inv_v[:3, :3].T, dim=1)
    eye = torch.tensor(cam["eye"], device=dev_, dtype=torch.float64)
    n1 = wdir.shape[0]
    prim = torch.zeros((n1, 8), device=dev_, dtype=torch.float32)
    prim[:, 0:3] = eye.float()
    prim[:, 4:7] = wdir.float()
    prim[:, 7] = 10000.0
    hits = torch.zeros((n1, 4), device=dev_, dtype=torch.float32)
    dev.intersect(prim.data_ptr(), hits.data_ptr(), n1, stream=sp)
    torch.cuda.synchronize()
    ids = hits[:, 1].contiguous().view(torch.int32)
    m = ids != -1
    wt = torch.tensor(world_triangles(arrays), device=dev_)
    tri = wt[ids[m].long()].reshape(-1, 3, 3)
    ng = torch.nn.functional.normalize(torch.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0], dim=1), dim=1)
    d0 = prim[m, 4:7]
    ng = torch.where(((ng * d0).sum(1) > 0)[:, None], -ng, ng)
    p = prim[m, 0:3] + hits[m, 0:1] * d0
    g = torch.Generator(device=dev_)
    g.manual_seed(1234)
    ng = ng.repeat(reps, 1)
    p = p.repeat(reps, 1)
    u1 = torch.rand(ng.shape[0], device=dev_, generator=g)
    u2 = torch.rand(ng.shape[0], device=dev_, generator=g)
    r = torch.sqrt(u1)
    phi = 2 * np.pi * u2
    helper = torch.where((ng[:, 0].abs() < 0.9)[:, None], torch.tensor([1.0, 0, 0], device=dev_),
                         torch.tensor([0, 1.0, 0], device=dev_))
    t1 = torch.nn.functional.normalize(torch.cross(ng, helper, dim=1), dim=1)
    t2 = torch.cross(ng, t1, dim=1)
    d = (t1 * (r * torch.cos(phi))[:, None] + t2 * (r * torch.sin(phi))[:, None] +
         ng * torch.sqrt(torch.clamp(1 - u1, min=0))[:, None])
    n2 = d.shape[0]
    sec = torch.zeros((n2, 8), device=dev_, dtype=torch.float32)
    sec[:, 0:3] = p + 1e-4 * ng
    sec[:, 4:7] = d
    sec[:, 7] = 10000.0
    sh = torch.zeros((n2, 4), device=dev_, dtype=torch.float32)
    if with_primary:
        return prim, sec, sh
    return sec, sh


def secondary_rays(a, dev, sec, sh, torch, node_bytes):
    import numpy as np

    n2 = sec.shape[0]
    dev_ = sec.device
    sp = torch.cuda.current_stream().cuda_stream
    cnt = torch.zeros((n2, 2), device=dev_, dtype=torch.int32)
    dev.intersect(sec.data_ptr(), sh.data_ptr(), n2, counters_ptr=cnt.data_ptr(), stream=sp)
    torch.cuda.synchronize()
    visits = int(cnt[:, 0].double().sum().item())
    tests = int(cnt[:, 1].double().sum().item())
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps_t = 5
    dev.intersect(sec.data_ptr(), sh.data_ptr(), n2, stream=sp)   # warm
    times = []
    for _ in range(reps_t):
        e0.record(stream)
        dev.intersect(sec.data_ptr(), sh.data_ptr(), n2, stream=sp)
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    ms = float(np.median(times))
    alg = n2 * RAY_BYTES + visits * node_bytes + tests * TRI_BYTES
    gbs = alg / (ms * 1e-3) / 1e9
    hit_frac = float((sh[:, 1].contiguous().view(torch.int32) != -1).float().mean().item())
    f16 = None
    if node_bytes == NODE_BYTES:
        # the same batch with FP16 inverse directions (REINA_RT_TRACE_F16_INV): conservative
        # slabs, so the hits must not change; only node visits and time can
        from reina_rt import abi
        tun = dev.get_tuning()
        dev.set_tuning(**dict(tun, waves_per_simd=0))   # the FP16 variant is built with the compiler's budget

        def timed(reps):
            out_ms = []
            for _ in range(reps):
                e0.record(stream)
                dev.intersect(sec.data_ptr(), sh16.data_ptr(), n2, stream=sp)
                e1.record(stream)
                torch.cuda.synchronize()
                out_ms.append(e0.elapsed_time(e1))
            return float(np.median(out_ms))
        sh16 = torch.zeros_like(sh)
        ms32 = timed(reps_t)
        dev.set_trace_options(abi.TRACE_F16_INV)
        cnt16 = torch.zeros_like(cnt)
        dev.intersect(sec.data_ptr(), sh16.data_ptr(), n2, counters_ptr=cnt16.data_ptr(), stream=sp)
        torch.cuda.synchronize()
        same = bool(torch.equal(sh16.view(torch.int32), sh.view(torch.int32)))
        v16 = int(cnt16[:, 0].double().sum().item())
        t16 = int(cnt16[:, 1].double().sum().item())
        ms16 = timed(reps_t)
        dev.set_trace_options(0)
        dev.set_tuning(**tun)
        f16 = {"kernel_ms": round(ms16, 3), "grays_per_s": round(n2 / (ms16 * 1e-3) / 1e9, 3),
               "fp32_inverses_same_budget_ms": round(ms32, 3),
               "nodes_per_ray": round(v16 / n2, 3), "tris_per_ray": round(t16 / n2, 3),
               "hits_bit_identical_to_fp32_inverses": same,
               "note": "ray-direction inverses rounded to FP16, slab intervals widened by 2^-10 |t| "
                       "(compiler-chosen register budget)"}
    return {
        "rays": int(n2),
        "grays_per_s": round(n2 / (ms * 1e-3) / 1e9, 3),
        "kernel_ms": round(ms, 3),
        "nodes_per_ray": round(visits / n2, 3),
        "tris_per_ray": round(tests / n2, 3),
        "bytes_per_ray": round(alg / n2, 1),
        "achieved_GBs": round(gbs, 1),
        "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
        "hit_fraction": round(hit_frac, 3),
        "note": "cosine-distributed bounce rays from the primary hits of the same 1080p frame, "
                "traced by reina_rt_intersect (same traversal kernel as the frame's extend)",
        "f16_inverses": f16,
    }


def world_triangles(arrays):
    import numpy as np
    out = []
    for inst in arrays["instances"]:
        T = inst["transform"].reshape(4, 4).T.astype(np.float32)
        m = arrays["models"][inst["modelID"]]
        idx = arrays["indices"][m["indexOffset"]:m["indexOffset"] + 3 * m["indexCount"]]
        v = arrays["vertices"][idx][:, :3].astype(np.float32)
        c0, c1, c2, c3 = T[:3, 0], T[:3, 1], T[:3, 2], T[:3, 3]
        w = ((c0[None] * v[:, 0:1] + c1[None] * v[:, 1:2]) + c2[None] * v[:, 2:3]) + c3[None]
        out.append(w.astype(np.float32).reshape(-1, 9))
    return np.concatenate(out)


def cpu_baseline(a, arrays, dev, cam, W, H):
    """Oracle restatement timed on the host cores: whole 1-spp passes of the
    same frame (sampleBatch = pass index, so every pass draws new samples)
    until the target duration is reached."""
    import numpy as np

    import oracle
    from reina_rt.scene import make_desc, push_consts

    cores = max(1, min(os.cpu_count() or 1, 16))
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_bvh(*dev.export_bvh())
    img = np.zeros((H, W, 4), np.float32)
    passes, rays, secs = 0, 0, 0.0
    while secs < a.cpu_seconds:
        pc = push_consts(cam["eye"], cam["center"], cam["fov"], W, H, 1, a.depth, sample_batch=passes)
        t = time.perf_counter()
        _, n = orc.render(pc, W, H, img=img, use_bvh=True, nthreads=cores)
        secs += time.perf_counter() - t
        rays += n
        passes += 1
    return {
        "value": round(rays / secs / 1e6, 3),
        "unit": "Mrays/s",
        "cores": cores,
        "kind": "port",
        "sample": "%d full %dx%d passes at 1 spp, depth %d (%d extension rays, %.1f s): oracle C restatement "
                  "of the reference shaders with BVH8 traversal, pthreads over pixels" % (passes, W, H, a.depth,
                                                                                        rays, secs),
    }


if __name__ == "__main__":
    main()
