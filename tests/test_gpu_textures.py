"""Textured hits on the GPU (albedo textures with alpha, normal maps,
parallax bump maps on Lambertian, metal, dielectric and Disney surfaces):
frames equal the oracle's bit for bit, a white opaque texture renders like
no texture, and textured requests through the server equal reina_rt_trace."""
import numpy as np
import pytest

import oracle
import reina_rt
from reina_rt import abi, scenes
from reina_rt.scene import make_desc

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _same(a, b):
    return bool(((_bits(a) == _bits(b)) | (np.isnan(a) & np.isnan(b))).all())


def _render(torch, dev, pc, W, H):
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    dev.trace(pc, img.data_ptr(), W, H, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return img.cpu().numpy()


@pytest.fixture(scope="module")
def textured():
    s, cam = scenes.textured_cornell()
    return s.build_arrays(), cam


@pytest.mark.parametrize("tail", [0, 1 << 30], ids=["wavefront", "tail_kernel"])
def test_textured_frame_equals_oracle(cuda, textured, tail):
    torch = cuda
    arrays, _ = textured
    W = H = 48
    pc = scenes.cornell_push(W, H, spp=4, max_bounces=8)
    dev = reina_rt.DeviceScene(arrays)
    dev.set_tuning(tail_paths=tail)
    got = _render(torch, dev, pc, W, H)
    ref, nrays = oracle.OracleScene(arrays, make_desc).render(pc, W, H, use_bvh=False)
    assert _same(got, ref), "pixels differing: %d" % (_bits(got) != _bits(ref)).any(-1).sum()
    assert dev.stats().extension_rays == nrays
    dev.close()


def test_textured_frame_with_nee_equals_oracle(cuda, textured):
    torch = cuda
    arrays, _ = textured
    W = H = 40
    pc = scenes.cornell_push(W, H, spp=3, max_bounces=8)
    dev = reina_rt.DeviceScene(arrays)
    dev.set_trace_options(abi.TRACE_NEE)
    got = _render(torch, dev, pc, W, H)
    orc = oracle.OracleScene(arrays, make_desc)
    orc.set_nee(True)
    ref, _ = orc.render(pc, W, H, use_bvh=False)
    assert _same(got, ref)
    dev.close()


def test_white_opaque_texture_renders_like_no_texture(cuda):
    torch = cuda
    a_s, _ = scenes.cornell()
    b_s, _ = scenes.textured_cornell(white_opaque=True)
    pc = scenes.cornell_push(64, 64, spp=4, max_bounces=8)
    da, db = reina_rt.DeviceScene(a_s.build_arrays()), reina_rt.DeviceScene(b_s.build_arrays())
    assert _same(_render(torch, da, pc, 64, 64), _render(torch, db, pc, 64, 64))
    da.close()
    db.close()


def test_textured_requests_through_server(cuda, textured):
    torch = cuda
    arrays, _ = textured
    dev = reina_rt.DeviceScene(arrays)
    srv = reina_rt.Server(path_budget_bytes=64 << 20)
    jobs = [(32, 32, 3), (48, 24, 2)]
    imgs, tks = [], []
    for W, H, spp in jobs:
        img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        imgs.append(img)
        tks.append(srv.submit(dev, scenes.cornell_push(W, H, spp=spp, max_bounces=8), img.data_ptr(), W, H))
    for t in tks:
        assert srv.wait(t).state == abi.REQ_DONE
    for (W, H, spp), img in zip(jobs, imgs):
        assert _same(img.cpu().numpy(), _render(torch, dev, scenes.cornell_push(W, H, spp=spp, max_bounces=8), W, H))
    srv.close()
    dev.close()


def test_texture_index_validated(cuda):
    s, _ = scenes.textured_cornell()
    arrays = s.build_arrays()
    arrays["props"][0]["normalMapTexID"] = 3          # only 3 textures (0..2)
    with pytest.raises(reina_rt.ReinaError) as e:
        reina_rt.DeviceScene(arrays)
    assert e.value.code == abi.ERR_INVALID_ARG
