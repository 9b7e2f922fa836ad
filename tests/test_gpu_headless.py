"""The C++ headless host (reina-vk_amd/host/reina_headless.cpp) drives the C
ABI like Reina's render loop: progressive frames, post-process, PNG
snapshots at save_on_samples. Its PNGs must equal, byte for byte, the same
frames rendered and post-processed through the Python binding."""
import os
import subprocess

import numpy as np
import pytest

import reina_rt
from reina_rt import scenes
from reina_rt.scene import save_scene

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "reina-vk_amd", "reina_headless")

CONFIG = """
[camera]
[camera.dof]
focus_dist = 2.2
defocus_multiplier = 0  # pinhole

[sampling]
samples_per_pixel = 4
max_bounces = 6
direct_clamp = 100
indirect_clamp = 10

[saving]
save_on_samples = [12, 4]
save_on_times = [0]

[postprocessing]
[postprocessing.bloom]
radius = 5.0
threshold = 1.0
intensity = 0.05

[postprocessing.tonemap]
exposure = 1

[headless]
width = 48
height = 40
total_samples = 24
"""

# SaveManager / Clock (src/tools/SaveManager.cpp, Clock.cpp) with 4 spp per
# frame: the clock's sample count lags (it counts frames from the second on),
# so frame k sees 4 (k - 1) samples; one save per frame, sample thresholds
# first: frame 0 saves output_0sec.png (age > 0), frame 3 (count 8 > 4)
# output_4spp.png, frame 5 (count 16 > 12) output_12spp.png. The image of
# frame k holds batches 0..k.
SAVES = {0: "output_0sec.png", 3: "output_4spp.png", 5: "output_12spp.png"}


ASSETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "assets")


@pytest.mark.parametrize("which", ["cornell", "obj_textured", "gltf"])
def test_headless_host_pngs_match_binding(cuda, tmp_path, which):
    """The scene file carries the packed arrays and (version 2) the textures:
    the OBJ scene with its PNG wall texture renders through the host as
    through the binding; a .glb is loaded and packed by the host itself and
    renders as reina_rt.gltf.load_scene's arrays do."""
    from PIL import Image
    torch = cuda
    scene_path = str(tmp_path / "cornell.rscn")
    if which == "cornell":
        sc, _ = scenes.cornell()
    elif which == "obj_textured":
        sc, _ = scenes.obj_cornell(ASSETS, subject="uv_sphere.obj")
    else:
        import test_gltf
        from reina_rt import gltf
        glb = test_gltf.build_asset().glb()
        scene_path = str(tmp_path / "scene.glb")
        (tmp_path / "scene.glb").write_bytes(glb)
        sc = gltf.load_scene(glb)
    arrays = sc.build_arrays()
    if which != "gltf":
        save_scene(arrays, scene_path)
    (tmp_path / "config.toml").write_text(CONFIG)
    W, H = 48, 40
    pc = scenes.cornell_push(W, H, spp=4, max_bounces=6)
    pc["defocusMultiplier"] = 0.0
    (tmp_path / "pc.bin").write_bytes(np.asarray(pc).tobytes())
    r = subprocess.run([BIN, str(tmp_path / "config.toml"), scene_path, str(tmp_path / "out"),
                        str(tmp_path / "pc.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert sorted(os.listdir(tmp_path / "out")) == sorted(SAVES.values())
    # the summary prints once the clock has recorded a frame time (from frame 2 on)
    assert r.stdout.count("Average category time | Ray Tracing") == 4
    dev = reina_rt.DeviceScene(arrays)
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    scratch = torch.zeros((2, H, W, 4), dtype=torch.float32, device="cuda")
    ldr = torch.zeros((H, W, 4), dtype=torch.uint8, device="cuda")
    sp = torch.cuda.current_stream().cuda_stream
    for batch in range(6):
        p = np.asarray(pc).copy()
        p["sampleBatch"] = batch
        dev.trace(p, img.data_ptr(), W, H, stream=sp)
        if batch in SAVES:
            reina_rt.postprocess(img.data_ptr(), scratch.data_ptr(), ldr.data_ptr(), W, H, stream=sp)
            torch.cuda.synchronize()
            got = np.asarray(Image.open(tmp_path / "out" / SAVES[batch]).convert("RGBA"))
            assert got.shape == (H, W, 4)
            assert np.array_equal(got, ldr.cpu().numpy()), batch
    dev.close()


def test_headless_host_reports_errors(cuda, tmp_path):
    (tmp_path / "bad.rscn").write_bytes(b"nope")
    (tmp_path / "config.toml").write_text(CONFIG)
    r = subprocess.run([BIN, str(tmp_path / "config.toml"), str(tmp_path / "bad.rscn"), str(tmp_path / "o")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "not a scene file" in r.stderr
