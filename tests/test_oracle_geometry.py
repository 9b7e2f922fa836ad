"""Pins the oracle's closest-hit query: watertight FP32 test vs FP64
Moller-Trumbore, BVH8 traversal restatement vs brute force (bit for bit),
BVH8 structure invariants, watertightness on closed meshes."""
import numpy as np
import pytest

import oracle
import reina_rt
from reina_rt import abi, scenes
from reina_rt.scene import Material, ModelData, Scene, make_desc

from conftest import world_triangles


def soup_arrays(tris):
    """A one-model scene from an (n, 9) triangle array (identity transform)."""
    n = len(tris)
    v = np.concatenate([tris.reshape(-1, 3), np.ones((3 * n, 1), np.float32)], 1).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.uint32)
    s = Scene()
    s.add_object(ModelData(v, idx, np.tile(np.eye(3, dtype=np.float32), (3 * n, 1, 1)), idx.copy()), np.eye(4),
                 Material())
    return s.build_arrays()


def rays_from(o, d, tmin=0.0, tmax=10000.0):
    r = np.zeros(len(o), abi.RAY)
    r["ox"], r["oy"], r["oz"] = np.asarray(o, np.float32).T
    r["dx"], r["dy"], r["dz"] = np.asarray(d, np.float32).T
    r["tmin"], r["tmax"] = tmin, tmax
    return r


def check_bvh_equals_brute(arrays, rays):
    """The BVH8 traversal, with FP32 nodes and with 8-bit quantized nodes,
    returns the brute-force closest hit bit for bit."""
    orc = oracle.OracleScene(arrays, make_desc)
    nodes, tris, depth = reina_rt.bvh_build_host(world_triangles(arrays))
    hb = orc.brute32(rays)
    cnts = []
    for nd, nbytes in ((nodes, 256), (reina_rt.bvh_quantize_host(nodes), 80)):
        orc.set_bvh(nd, tris, nbytes)
        hv, cnt = orc.bvh(rays)
        for f in ("t", "u", "v"):
            assert np.array_equal(hb[f].view(np.uint32), hv[f].view(np.uint32)), (f, nbytes)
        assert np.array_equal(hb["prim"], hv["prim"]), nbytes
        cnts.append(cnt)
    return hb, cnts[0], depth


def test_bvh_equals_brute_random_soup():
    rng = np.random.default_rng(1)
    c = rng.uniform(-1, 1, (3000, 1, 3))
    tris = (c + 0.08 * rng.normal(size=(3000, 3, 3))).astype(np.float32).reshape(-1, 9)
    # degenerate, duplicated and axis-aligned triangles
    tris[:20, 3:6] = tris[:20, 0:3]
    tris[20:40] = tris[40:60]
    tris[60:80, 1::3] = 0.25
    arrays = soup_arrays(tris)
    rays = rays_from(rng.uniform(-1.5, 1.5, (60000, 3)), rng.normal(size=(60000, 3)))
    hb, cnt, depth = check_bvh_equals_brute(arrays, rays)
    assert (hb["prim"] != 0xFFFFFFFF).mean() > 0.3
    assert depth <= 16
    assert cnt[:, 0].min() >= 1


def test_bvh_equals_brute_scenes(cornell_arrays, icosphere_arrays):
    arrays, cam = cornell_arrays
    rng = np.random.default_rng(2)
    check_bvh_equals_brute(arrays, rays_from(rng.uniform(-1.2, 2.2, (50000, 3)), rng.normal(size=(50000, 3))))
    arrays, cam = icosphere_arrays
    o = np.tile(np.array(cam["eye"], np.float32), (1500, 1))
    d = rng.normal(size=(1500, 3)) * 0.3 + (np.array(cam["center"]) - np.array(cam["eye"]))
    check_bvh_equals_brute(arrays, rays_from(o, d))


def test_bvh_structure():
    """Every triangle appears once; child boxes contain their triangles and
    their descendants' boxes; leaves hold 1..3 triangles."""
    rng = np.random.default_rng(3)
    tris = rng.uniform(-5, 5, (20000, 9)).astype(np.float32) * 0.1 + rng.uniform(-5, 5, (20000, 1)).astype(np.float32)
    nodes, tri_out, depth = reina_rt.bvh_build_host(tris)
    check_bvh8_structure(nodes, tri_out, len(tris))


def check_bvh8_structure(nodes, tri_out, ntri):
    """Invariants of a BVH8 in the library's FP32 node format."""
    N = len(nodes) // 256
    nb = nodes.reshape(N, 256)
    f = nb[:, :192].copy().view(np.float32).reshape(N, 6, 8)   # lo_x hi_x lo_y hi_y lo_z hi_z
    hdr = nb[:, 192:200].copy().view(np.uint32).reshape(N, 2)
    meta = nb[:, 200:208]
    t = tri_out.reshape(-1, 12)
    ids = t[:, 3].copy().view(np.uint32)
    assert sorted(ids.tolist()) == list(range(ntri))
    seen_nodes = {0}
    for n in range(N):
        cb, tb = hdr[n] & 0xFFFFFF
        im, lmk = hdr[n] >> 24
        assert im == sum(1 << k for k in range(8) if meta[n, k] & 0x80)
        assert lmk == sum(1 << k for k in range(8) if meta[n, k] and not meta[n, k] & 0x80)
        for k in range(8):
            m = int(meta[n, k])
            if m == 0:
                continue
            lo = np.array([f[n, 0, k], f[n, 2, k], f[n, 4, k]])
            hi = np.array([f[n, 1, k], f[n, 3, k], f[n, 5, k]])
            if m & 0x80:
                c = cb + (m & 7)
                seen_nodes.add(int(c))
                cf = f[c]
                valid = meta[c] != 0
                assert (cf[0][valid] >= lo[0]).all() and (cf[1][valid] <= hi[0]).all()
                assert (cf[2][valid] >= lo[1]).all() and (cf[3][valid] <= hi[1]).all()
                assert (cf[4][valid] >= lo[2]).all() and (cf[5][valid] <= hi[2]).all()
            else:
                cnt, off = (m >> 5) & 3, m & 31
                assert 1 <= cnt <= 3
                for j in range(cnt):
                    tv = t[tb + off + j][[0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(3, 3)
                    assert (tv >= lo - 1e-6).all() and (tv <= hi + 1e-6).all()
    assert seen_nodes == set(range(N))


def decode_quantized(q):
    """(N, 6, 8) float32 decoded child planes of 80-byte quantized nodes, in
    the FP32 node's order lo_x hi_x lo_y hi_y lo_z hi_z. q * s is exact in
    float32 (8-bit q, power-of-two s), so the float32 product plus p rounds
    once, exactly like fmaf(q, s, p)."""
    N = len(q) // 80
    qb = q.reshape(N, 80)
    p = qb[:, :12].copy().view(np.float32)
    ex = qb[:, 12:16].copy().view(np.uint32)[:, 0]
    out = np.zeros((N, 6, 8), np.float32)
    for a in range(3):
        sc = ((((ex >> (8 * a)) & 0xFF).astype(np.uint32)) << 23).view(np.float32)
        for h in range(2):
            qq = qb[:, 32 + 16 * a + 8 * h: 40 + 16 * a + 8 * h].astype(np.float32)
            out[:, 2 * a + h] = qq * sc[:, None] + p[:, a:a + 1]
    return out


@pytest.mark.parametrize("kind", ["soup", "icosphere", "flat"])
def test_quantized_nodes_contain_fp32(kind, icosphere_arrays):
    """Every decoded 8-bit child box contains the FP32 child box, the grid
    step is the smallest power of two that can cover the node (so boxes grow
    by at most a few grid steps), and topology, header and meta are the FP32
    node's."""
    rng = np.random.default_rng(11)
    if kind == "soup":
        tris = (rng.uniform(-50, 50, (30000, 1)) + rng.normal(size=(30000, 9))).astype(np.float32)
    elif kind == "icosphere":
        tris = world_triangles(icosphere_arrays[0])
    else:   # thin and far from the origin: grid steps far below the origin's ulp
        tris = rng.uniform(-1, 1, (5000, 9)).astype(np.float32)
        tris[:, 1::3] = 1e-3 * tris[:, 1::3] + 1000.0
    nodes, t, depth = reina_rt.bvh_build_host(tris)
    q = reina_rt.bvh_quantize_host(nodes)
    N = len(nodes) // 256
    assert len(q) == 80 * N
    nb, qb = nodes.reshape(N, 256), q.reshape(N, 80)
    assert np.array_equal(qb[:, 16:32], nb[:, 192:208])
    f = nb[:, :192].copy().view(np.float32).reshape(N, 6, 8)
    d = decode_quantized(q)
    valid = nb[:, 200:208] != 0
    ex = qb[:, 12:16].copy().view(np.uint32)[:, 0]
    for a in range(3):
        lo, hi = f[:, 2 * a], f[:, 2 * a + 1]
        assert (d[:, 2 * a][valid] <= lo[valid]).all()
        assert (d[:, 2 * a + 1][valid] >= hi[valid]).all()
        s = ((((ex >> (8 * a)) & 0xFF).astype(np.uint32)) << 23).view(np.float32).astype(np.float64)
        ext = np.where(valid, hi, -np.inf).max(1).astype(np.float64) - np.where(valid, lo, np.inf).min(1)
        # minimal step: half the step would not cover the node extent with 255 steps
        big = s > 2.0 ** -125
        assert (ext[big] >= 255 * s[big] / 2).all()
        grow = (d[:, 2 * a + 1] - hi).astype(np.float64) + (lo - d[:, 2 * a])
        assert (grow[valid] <= (3 * s[:, None] + 4 * np.spacing(np.abs(hi).astype(np.float32)))[valid]).all()


def test_empty_and_tiny_scenes():
    for n in (0, 1, 2, 3, 4, 9):
        rng = np.random.default_rng(n)
        tris = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
        nodes, t, depth = reina_rt.bvh_build_host(tris)
        assert len(nodes) >= 256 and len(t) == 12 * n
        if n:
            arrays = soup_arrays(tris)
            check_bvh_equals_brute(arrays, rays_from(rng.uniform(-2, 2, (2000, 3)), rng.normal(size=(2000, 3))))


def test_watertight_closed_mesh():
    """Rays from inside a closed icosphere (shared edges/vertices everywhere)
    never escape: the watertight test leaves no crack."""
    v, f = scenes.subdivided_icosphere(4)
    tris = v[f].astype(np.float32).reshape(-1, 9)
    arrays = soup_arrays(tris)
    rng = np.random.default_rng(4)
    n = 40000
    o = rng.uniform(-0.3, 0.3, (n, 3))
    d = rng.normal(size=(n, 3))
    # aim a share of rays exactly at mesh vertices and edge midpoints
    tgt = np.concatenate([v[f[:, 0]], 0.5 * (v[f[:, 0]] + v[f[:, 1]])])[: n // 2]
    d[: len(tgt)] = tgt - o[: len(tgt)]
    hb, cnt, _ = check_bvh_equals_brute(arrays, rays_from(o, d))
    assert (hb["prim"] != 0xFFFFFFFF).all()


def plane_rays(v, n, seed):
    """Rays that run inside the coordinate planes x=0, y=0, z=0 (direction
    component exactly 0, origin component exactly 0) or along the axes, aimed
    at mesh vertices in those planes: they hit shared edges and vertices
    exactly (ties) and graze the boxes' faces at zero distance."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    for i in range(n):
        a = i % 4
        if a < 3:
            o[i, a] = 0.0
            d[i, a] = 0.0
            on = v[v[:, a] == 0]
            if len(on) and i % 8 < 4:
                d[i] = on[rng.integers(len(on))] - o[i]
                d[i, a] = 0.0
        else:   # axis-parallel through a vertex's two coordinates
            ax = rng.integers(3)
            p = v[rng.integers(len(v))]
            o[i] = p
            o[i, ax] = 3.0
            d[i] = 0.0
            d[i, ax] = -1.0
    return rays_from(o, d)


def test_zero_direction_components():
    """Zero direction components (infinite inverse direction) with origins
    exactly on box planes: BVH traversal still equals brute force (a
    regression: a tiny stand-in for 1/0 made the slab end right after the
    origin when a box face sat one denormal away from it)."""
    v, f = scenes.subdivided_icosphere(3)
    v = v.astype(np.float32)
    tris = v[f].reshape(-1, 9)
    arrays = soup_arrays(tris)
    hb, cnt, _ = check_bvh_equals_brute(arrays, plane_rays(v, 20000, 21))
    assert (hb["prim"] != 0xFFFFFFFF).mean() > 0.3


def near_axis_rays(v, n, seed):
    """Rays for both slab forms of the box test: origins near the mesh
    vertices v (placed far from the coordinate origin, |o/d| is much larger
    than the hit distance: the case the 2^-16 |o/d| term of the one-FMA form
    covers), aimed at vertices; every other ray gets one direction component
    of magnitude 2^-41 (below the 2^-40 switch: two-rounding form), 2^-39
    (one-FMA form with a large inverse), 1e-9 or 1e-6."""
    rng = np.random.default_rng(seed)
    tiny = [2.0 ** -41, 2.0 ** -39, 1e-9, 1e-6]
    vs = v.astype(np.float64)
    o = vs[rng.integers(len(v), size=n)] + rng.normal(scale=0.3, size=(n, 3))
    d = vs[rng.integers(len(v), size=n)] - o
    for i in range(0, n, 2):
        d[i, (i // 2) % 3] = tiny[(i // 6) % 4] * (1.0 if i % 4 == 0 else -1.0)
    return rays_from(o.astype(np.float32), d.astype(np.float32))


def shifted_icosphere(shift):
    v, f = scenes.subdivided_icosphere(3)
    v = (v.astype(np.float64) + shift).astype(np.float32)
    return v, soup_arrays(v[f].reshape(-1, 9))


@pytest.mark.parametrize("shift", [0.0, 3000.0])
def test_near_axis_and_far_origin_rays(shift):
    """Tiny direction components on either side of the 2^-40 switch between
    the two slab forms, and a mesh 3000 units from the coordinate origin:
    BVH traversal equals brute force bit for bit."""
    v, arrays = shifted_icosphere(shift)
    hb, cnt, _ = check_bvh_equals_brute(arrays, near_axis_rays(v, 20000, 31))
    assert (hb["prim"] != 0xFFFFFFFF).mean() > 0.3


def test_fp64_bvh_culling_equals_full_scan(icosphere_arrays):
    """The FP64 oracle with BVH culling returns exactly the full scan's hit,
    t, barycentrics and ambiguity flags (camera rays, random rays, and
    plane/axis rays that hit edges exactly)."""
    v, f = scenes.subdivided_icosphere(3)
    v = v.astype(np.float32)
    cases = [(soup_arrays(v[f].reshape(-1, 9)), plane_rays(v, 3000, 31))]
    arrays, cam = icosphere_arrays
    rng = np.random.default_rng(32)
    o = np.tile(np.array(cam["eye"], np.float32), (1500, 1))
    d = rng.normal(size=(1500, 3)) * 0.3 + (np.array(cam["center"]) - np.array(cam["eye"]))
    cases.append((arrays, np.concatenate([rays_from(o, d), rays_from(rng.uniform(-1.5, 1.5, (1500, 3)),
                                                                     rng.normal(size=(1500, 3)))])))
    ambiguous = []
    for arr, rays in cases:
        orc = oracle.OracleScene(arr, make_desc)
        nodes, tris, _ = reina_rt.bvh_build_host(world_triangles(arr))
        orc.set_bvh(nodes, tris)
        full = orc.fp64(rays)
        culled = orc.fp64(rays, use_bvh=True)
        for a, b in zip(full, culled):
            assert np.array_equal(a, b)
        ambiguous.append(int(full[4].sum()))
    assert ambiguous[0] > 100          # the plane rays hit shared edges: ties are exercised


def test_invalid_rays_miss(cornell_arrays):
    arrays, _ = cornell_arrays
    r = rays_from(np.array([[np.nan, 1, 0], [0, 1, 0], [np.inf, 1, 0], [0, 1, 0]]),
                  np.array([[0, 0, -1], [0, 0, 0], [0, 0, -1], [np.nan, 0, -1]]))
    orc = oracle.OracleScene(arrays, make_desc)
    hb = orc.brute32(r)
    assert (hb["prim"] == 0xFFFFFFFF).all()


def test_fp32_vs_fp64_agreement(icosphere_arrays):
    """On camera rays the FP32 watertight hit id equals the FP64 one except at
    FP64-detected ties, and t agrees to 1e-6 relative."""
    arrays, cam = icosphere_arrays
    rng = np.random.default_rng(5)
    orc = oracle.OracleScene(arrays, make_desc)
    o = np.tile(np.array(cam["eye"], np.float32), (400, 1))
    d = rng.normal(size=(400, 3)) * 0.25 + (np.array(cam["center"]) - np.array(cam["eye"]))
    r = rays_from(o, d)
    h = orc.brute32(r)
    t64, p64, u64, v64, amb = orc.fp64(r)
    dis = h["prim"] != p64
    assert (dis & ~amb).sum() == 0
    m = ~dis & (p64 != 0xFFFFFFFF)
    assert (np.abs(h["t"][m] - t64[m]) / t64[m]).max() < 1e-6
