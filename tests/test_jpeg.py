"""JPEG textures: the library's baseline decoder (reina_rt_image_load on JPEG
data, replacing stb_image's JPEG path behind src/graphics/Image.cpp:10-41).

stb is not in this image. The Huffman decode is fixed by the JPEG standard;
the arithmetic parts restate stb's published integer IDCT, upsampling and
fixed-point YCbCr conversion. Pinned here: Pillow (libjpeg, an independent
decoder whose IDCT and colour conversion round differently) within 4 levels
per channel and a small mean, closed forms (a constant grey image at quality
100 decodes exactly: DC-only blocks give floor(D / 8 + 0.5) + 128 = the grey
value, and Cb = Cr = 128 leaves R = G = B = Y), restart intervals, all chroma
layouts, progressive files, CMYK files, and the error codes. One column of 4:2:2 files, 2 * (chroma width
- 1), is excluded from the Pillow comparison: stb's horizontal upsampler
weights that sample toward the previous chroma sample (3 * in[w-2] + in[w-1]),
libjpeg toward its own.
"""
import glob
import io
import os

import numpy as np
import pytest

import reina_rt
from reina_rt.scene import load_image

REF_JPGS = sorted(glob.glob("/root/reference/textures/*.jpg"))


def _encode(img, **kw):
    from PIL import Image
    b = io.BytesIO()
    Image.fromarray(img).save(b, format="JPEG", **kw)
    return b.getvalue()


def _pillow(data):
    from PIL import Image
    return np.asarray(Image.open(io.BytesIO(data)).convert("RGBA"))


def _close(a, b, exclude_col=None):
    d = np.abs(a.astype(int) - b.astype(int))
    if exclude_col is not None:
        d = np.delete(d, exclude_col, axis=1)
    assert d.max() <= 4, d.max()
    assert d.mean() < 0.25, d.mean()


@pytest.mark.parametrize("subsampling", [0, 1, 2], ids=["444", "422", "420"])
@pytest.mark.parametrize("shape", [(37, 53), (8, 8), (17, 200), (40, 16)])
def test_colour_jpeg_against_pillow(subsampling, shape):
    """Widths of at least 8: on narrower images libjpeg-turbo's upsampling
    near the right edge reads the decoded padding columns (differences up to
    14 levels on 3-pixel-wide images), which stb does not; those sizes are
    pinned by the closed forms below."""
    rng = np.random.default_rng(shape[0] * 7 + shape[1] + subsampling)
    # smooth content plus noise (what textures look like) keeps libjpeg and
    # stb within their rounding differences
    yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
    base = np.stack([xx * 4 % 256, yy * 3 % 256, (xx + yy) * 2 % 256], -1)
    img = np.clip(base + rng.integers(-20, 20, base.shape), 0, 255).astype(np.uint8)
    data = _encode(img, quality=92, subsampling=subsampling)
    got = load_image(data)
    assert got.shape == (shape[0], shape[1], 4) and (got[..., 3] == 255).all()
    # stb's 4:2:2 edge sample: output column 2 * (chroma width - 1)
    edge = 2 * ((shape[1] + 1) // 2 - 1) if subsampling == 1 and shape[1] > 1 else None
    _close(got, _pillow(data), exclude_col=edge)


@pytest.mark.parametrize("kw", [dict(restart_marker_blocks=3), dict(restart_marker_rows=1), dict(optimize=True)],
                         ids=["restart_blocks", "restart_rows", "optimized_huffman"])
def test_restart_intervals_and_custom_tables(kw):
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, (45, 77, 3)).astype(np.uint8)
    data = _encode(img, quality=85, **kw)
    _close(load_image(data), _pillow(data))


def test_grey_jpeg_and_constant_closed_form():
    rng = np.random.default_rng(4)
    g = rng.integers(0, 256, (33, 41)).astype(np.uint8)
    data = _encode(g, quality=90)
    got = load_image(data)
    assert (got[..., 0] == got[..., 1]).all() and (got[..., 1] == got[..., 2]).all()
    _close(got, _pillow(data))
    for v in (0, 17, 128, 200, 255):
        flat = load_image(_encode(np.full((24, 40), v, np.uint8), quality=100))
        assert (flat[..., :3] == v).all() and (flat[..., 3] == 255).all(), v
        # grey colour images: Cb = Cr = 128 in every layout and size, narrow
        # ones included (upsampling a constant is exact)
        for sub in (0, 1, 2):
            for shape in [(16, 16), (1, 1), (64, 3), (5, 7), (2, 33)]:
                col = load_image(_encode(np.full(shape + (3,), v, np.uint8), quality=100, subsampling=sub))
                assert (col[..., :3] == v).all(), (v, sub, shape)


def test_path_load_flips_rows(tmp_path):
    img = np.random.default_rng(2).integers(0, 256, (20, 30, 3)).astype(np.uint8)
    data = _encode(img)
    p = tmp_path / "x.jpg"
    p.write_bytes(data)
    assert np.array_equal(load_image(str(p)), load_image(data)[::-1])


@pytest.mark.skipif(not REF_JPGS, reason="the reference's textures/ directory is not present")
@pytest.mark.parametrize("path", REF_JPGS or ["none"], ids=[os.path.basename(p) for p in REF_JPGS] or ["none"])
def test_reference_jpegs_against_pillow(path):
    from PIL import Image
    got = load_image(path)
    want = np.asarray(Image.open(path).convert("RGBA"))[::-1]
    _close(got, want)


@pytest.mark.parametrize("subsampling", [0, 1, 2], ids=["444", "422", "420"])
@pytest.mark.parametrize("kw", [dict(), dict(optimize=True), dict(restart_marker_blocks=2), dict(quality=100)],
                         ids=["default", "optimized", "restarts", "q100"])
def test_progressive_jpeg_against_pillow(subsampling, kw):
    """Progressive files (spectral selection and successive approximation:
    DC first / refine, AC first with end-of-band runs, AC refine) decode
    like the same image's baseline file under the comparison above."""
    rng = np.random.default_rng(11 + subsampling)
    shape = (45, 67)
    yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
    base = np.stack([xx * 4 % 256, yy * 3 % 256, (xx + yy) * 2 % 256], -1)
    img = np.clip(base + rng.integers(-20, 20, base.shape), 0, 255).astype(np.uint8)
    data = _encode(img, progressive=True, subsampling=subsampling, **kw)
    assert b"\xff\xc2" in data
    got = load_image(data)
    edge = 2 * ((shape[1] + 1) // 2 - 1) if subsampling == 1 else None
    _close(got, _pillow(data), exclude_col=edge)


def test_progressive_equals_baseline_coefficients():
    """Constant images decode identically from progressive and baseline
    files (their DC-only blocks are exact in both); a grey progressive file
    is compared with Pillow."""
    for v in (0, 37, 128, 255):
        flat = np.full((24, 40, 3), v, np.uint8)
        assert np.array_equal(load_image(_encode(flat, quality=100, progressive=True)),
                              load_image(_encode(flat, quality=100)))
    g = np.random.default_rng(5).integers(0, 256, (33, 41)).astype(np.uint8)
    data = _encode(g, quality=90, progressive=True)
    _close(load_image(data), _pillow(data))


def _blinn(x, y):
    """x * y / 255 rounded as t = x * y + 128, (t + (t >> 8)) >> 8."""
    t = x * y + 128
    return (t + (t >> 8)) >> 8


def test_cmyk_jpeg():
    """4-component Adobe files (transform 0, stored inverted as Adobe
    applications and Pillow write them): RGB = stored colour times stored K
    / 255. Constant images at quality 100 decode to that closed form exactly;
    a varied image stays within the bound of Pillow's own CMYK conversion."""
    from PIL import Image
    for c, m, y, k in [(40, 120, 200, 60), (0, 0, 0, 0), (255, 255, 255, 255), (10, 250, 128, 200)]:
        b = io.BytesIO()
        Image.fromarray(np.full((16, 24, 4), (c, m, y, k), np.uint8), "CMYK").save(b, format="JPEG", quality=100)
        got = load_image(b.getvalue())
        expect = [_blinn(255 - v, 255 - k) for v in (c, m, y)]
        assert (got[..., :3] == expect).all() and (got[..., 3] == 255).all(), (c, m, y, k)
    rng = np.random.default_rng(2)
    yy, xx = np.mgrid[0:40, 0:50]
    rgb = np.clip(np.stack([xx * 4 % 256, yy * 5 % 256, (xx + yy) * 2 % 256], -1) + rng.integers(-10, 10, (40, 50, 3)),
                  0, 255).astype(np.uint8)
    b = io.BytesIO()
    Image.fromarray(rgb).convert("CMYK").save(b, format="JPEG", quality=95)
    _close(load_image(b.getvalue()), _pillow(b.getvalue()))


def test_jpeg_errors():
    img = np.random.default_rng(1).integers(0, 256, (16, 16, 3)).astype(np.uint8)
    good = _encode(img)
    for bad in [good[:30], b"\xff\xd8\xff\xe0" + b"\0" * 16, good[:2] + b"\xff\xdb\x00\x02" + good[2:4]]:
        with pytest.raises(reina_rt.ReinaError):
            load_image(bad)
