"""Projection, alpha blend (forward / reference-parity backward),
project_backward, loss and tonemap on the GPU against the oracle.

Bars: projection records, keys, ranges, images and project_backward outputs
bit-exact (same arithmetic on both sides); the atomically accumulated 2-D
gradients within 1e-4 norm-wise relative per attribute (float atomics sum in
a different order than the oracle's float64 accumulation)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import synth
from nanotracer_amd.api import (CameraUniform, GpuSplatBuffer, Rasterizer, TileBinner, TilingParams, loss_grad)

from .helpers import config_a_camera, config_a_scene, config_b_camera, config_b_scene, projected_record, rel_norm_err

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-4


def _camera(c, n):
    cam = CameraUniform.from_pose(c["eye"], c["target"], c["up"], c["fov_y"], c["width"], c["height"], n)
    ocam = O.camera_from_pose(c["eye"], c["target"], c["up"], c["fov_y"], c["width"], c["height"], n)
    assert bytes(cam) == bytes(ocam), "camera from_pose must agree bit for bit"
    return cam, ocam


def _forward(ctx, scene_host, c, tile=16):
    n = scene_host.n
    cam, ocam = _camera(c, n)
    g = GpuSplatBuffer.upload(ctx, scene_host)
    r = Rasterizer()
    proj = r.alloc_projected(ctx, n)
    r.project(ctx, g, cam, proj)
    params = TilingParams(c["width"], c["height"], tile, c["depth_max"])
    bins = TileBinner().bin(ctx, proj, n, params)
    img = r.alloc_image(ctx, c["width"], c["height"])
    r.composite(ctx, proj, bins, img, params)
    return dict(g=g, cam=cam, ocam=ocam, proj=proj, bins=bins, img=img, params=params, r=r)


@pytest.mark.parametrize("which", ["A", "B"])
def test_projection_bit_exact(ctx, which):
    s, c = (config_a_scene(), config_a_camera()) if which == "A" else (config_b_scene(), config_b_camera())
    f = _forward(ctx, s, c)
    gp = f["proj"].download(np.float32, s.n * 12).reshape(s.n, 12)
    op = O.project(s, f["ocam"])
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32))


def test_cpu_oracle_centers_on_axis_gpu(ctx):
    # raster.rs:1217-1265 wgsl_matches_cpu_oracle (here: bit-exact)
    s = synth.HostSplats(np.array([[0, 0, -5]], np.float32), np.array([[1, 0, 0, 0]], np.float32),
                         np.zeros((1, 3), np.float32), np.array([10.0], np.float32),
                         np.array([[0.5, 0.3, 0.1]], np.float32), np.zeros((1, 45), np.float32))
    c = dict(eye=(0, 0, 0), target=(0, 0, -1), up=(0, 1, 0), fov_y=math.pi / 2, width=256, height=256, depth_max=50.0)
    f = _forward(ctx, s, c)
    gp = f["proj"].download(np.float32, 12)
    assert gp[7] == 1.0 and abs(gp[0] - 128) < 1e-3 and abs(gp[1] - 128) < 1e-3 and abs(gp[2] - 5) < 1e-3
    assert 50 < gp[3] < 100 and abs(gp[5]) < 1e-3 and abs(gp[4] - gp[6]) < 1e-3
    assert np.array_equal(gp, O.project(s, f["ocam"])[0])


@pytest.mark.parametrize("which,tile", [("A", 16), ("A", 8), ("A", 32), ("B", 16)])
def test_forward_keys_ranges_image_bit_exact(ctx, which, tile):
    s, c = (config_a_scene(), config_a_camera()) if which == "A" else (config_b_scene(), config_b_camera())
    f = _forward(ctx, s, c, tile)
    op = O.project(s, f["ocam"])
    t = O.tiling(c["width"], c["height"], tile, c["depth_max"])
    ok, opay, orr = O.bin(op, t)
    bins = f["bins"]
    assert bins.total_pairs == ok.size
    if bins.num_tiles <= 65536:
        assert np.array_equal(bins.sorted_keys(), ok)
    assert np.array_equal(bins.sorted_payloads(), opay)
    assert np.array_equal(bins.tile_ranges(), orr)
    gimg = f["img"].download(np.float32, c["width"] * c["height"] * 4).reshape(c["height"], c["width"], 4)
    oimg = O.composite(op, opay, orr, t)
    assert np.abs(gimg - oimg).max() <= 1e-5
    assert np.array_equal(gimg.view(np.uint32), oimg.view(np.uint32)), "image must be bit-identical"


def test_composite_single_splat_renders_gaussian(ctx):
    # raster.rs:1028-1100
    proj = projected_record(32.5, 32.5, 5.0, 30.0, conic=(0.04, 0.0, 0.04))[None]
    buf = ctx.storage_buffer(proj)
    params = TilingParams(64, 64, 16, 100.0)
    bins = TileBinner().bin(ctx, buf, 1, params)
    r = Rasterizer()
    img = r.alloc_image(ctx, 64, 64)
    r.composite(ctx, buf, bins, img, params)
    px = img.download(np.float32, 64 * 64 * 4).reshape(64, 64, 4)
    centre, near, edge = px[32, 32], px[34, 34], px[0, 0]
    assert centre[0] > 0.9 and centre[3] > 0.9
    assert near[0] > 0.6 and near[0] < centre[0] + 1e-3
    assert edge[0] < 0.05 and edge[3] < 0.05
    for d in range(16):
        assert px[32 + d, 32 + d, 3] <= centre[3] + 1e-3


def _backward(ctx, f, dl_dc_host, c):
    n = f["g"].n
    r = f["r"]
    dl = ctx.storage_buffer(dl_dc_host)
    pg = r.alloc_projected_grad(ctx, n)
    r.composite_backward(ctx, f["proj"], f["bins"], f["img"], dl, pg, f["params"])
    return pg.download(np.float32, n * 12).reshape(n, 12), pg


def test_backward_accumulates_per_splat_gradients(ctx):
    # raster.rs:1102-1215 (sign invariants)
    proj = projected_record(16.0, 16.0, 5.0, 25.0, conic=(0.04, 0.0, 0.04))[None]
    buf = ctx.storage_buffer(proj)
    params = TilingParams(32, 32, 16, 100.0)
    bins = TileBinner().bin(ctx, buf, 1, params)
    r = Rasterizer()
    img = r.alloc_image(ctx, 32, 32)
    r.composite(ctx, buf, bins, img, params)
    dl = ctx.storage_buffer(np.tile(np.array([1, 1, 1, 0], np.float32), 32 * 32))
    pg = r.alloc_projected_grad(ctx, 1)
    r.composite_backward(ctx, buf, bins, img, dl, pg, params)
    g = pg.download(np.float32, 12)
    assert g[8] > 0.1 and g[9] > 0.1 and g[10] > 0.1
    assert g[2] > 0.01
    assert g[4] < -1e-4 and g[6] < -1e-4
    assert abs(g[0]) < 0.5 and abs(g[1]) < 0.5


@pytest.mark.parametrize("which,tile", [("A", 16), ("A", 8), ("A", 32), ("B", 16)])
def test_composite_backward_matches_oracle(ctx, which, tile):
    s, c = (config_a_scene(), config_a_camera()) if which == "A" else (config_b_scene(), config_b_camera())
    f = _forward(ctx, s, c, tile)
    W, H = c["width"], c["height"]
    rng = np.random.default_rng(3)
    dl = rng.normal(0, 1e-3, (H, W, 4)).astype(np.float32)
    dl[..., 3] = 0
    g, _ = _backward(ctx, f, dl, c)
    op = O.project(s, f["ocam"])
    t = O.tiling(W, H, tile, c["depth_max"])
    ok, opay, orr = O.bin(op, t)
    oimg = O.composite(op, opay, orr, t)
    og = O.composite_backward_ref(op, opay, orr, t, oimg, dl)
    for name, cols in (("dmean", [0, 1]), ("dopacity", [2]), ("dconic", [4, 5, 6]), ("dcolor", [8, 9, 10])):
        err = rel_norm_err(g[:, cols], og[:, cols])
        assert err < GRAD_TOL, f"{name}: relative error {err:.2e}"
    # element-wise floor: no single splat is far off
    scale = np.abs(og).max(axis=0) + 1e-30
    assert (np.abs(g - og) / scale).max() < 1e-3


@pytest.mark.parametrize("which", ["A", "B"])
def test_project_backward_bit_exact(ctx, which):
    s, c = (config_a_scene(), config_a_camera()) if which == "A" else (config_b_scene(), config_b_camera())
    f = _forward(ctx, s, c)
    n = s.n
    rng = np.random.default_rng(11)
    pg_host = rng.normal(0, 1e-2, (n, 12)).astype(np.float32)
    pg_host[:, [3, 7, 11]] = 0
    pg = ctx.storage_buffer(pg_host)
    f["r"].project_backward(ctx, f["g"], pg, f["cam"])
    gg = f["g"].grads()
    og = O.project_backward(s, f["ocam"], pg_host)
    for a, b in zip(gg.arrays(), og.arrays()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("kind", [0, 1])
def test_loss_grad_matches_oracle(ctx, kind):
    rng = np.random.default_rng(kind)
    H, W = 97, 131
    pred = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    tgt = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    tgt[::7, ::5, :3] = pred[::7, ::5, :3]  # exact zeros -> sign(0) = 0
    a, b = ctx.storage_buffer(pred), ctx.storage_buffer(tgt)
    dl = ctx.storage_buffer_zeroed(H * W * 16)
    loss = loss_grad(ctx, a, b, dl, W, H, kind)
    ol, odl = O.loss_grad(pred, tgt, kind)
    gdl = dl.download(np.float32, H * W * 4).reshape(H, W, 4)
    assert np.array_equal(gdl, odl)
    assert abs(loss - ol) <= 1e-9 * max(1.0, abs(ol))


def test_tonemap_matches_oracle(ctx):
    rng = np.random.default_rng(5)
    img = rng.uniform(-0.5, 4.0, (64, 80, 4)).astype(np.float32)
    img[..., 3] = rng.uniform(0, 1, (64, 80))
    buf = ctx.storage_buffer(img)
    out = ctx.storage_buffer_zeroed(64 * 80 * 4)
    Rasterizer().tonemap(ctx, buf, out, 80, 64)
    g = out.download(np.uint8, 64 * 80 * 4).reshape(64, 80, 4)
    o = O.tonemap(img)
    # the same arithmetic on both sides (nts_srgb_pow): bit for bit
    assert np.array_equal(g, o)


def _forward_ex(ctx, s, c, tile=16):
    f = _forward(ctx, s, c, tile)
    W, H = c["width"], c["height"]
    ft = ctx.storage_buffer_zeroed(W * H * 4)
    li = ctx.storage_buffer_zeroed(W * H * 4)
    f["r"].composite_ex(ctx, f["proj"], f["bins"], f["img"], f["params"], ft, li)
    f["ft"], f["li"] = ft, li
    return f


@pytest.mark.parametrize("which,tile", [("A", 16), ("A", 32), ("B", 16)])
def test_composite_ex_state_bit_exact(ctx, which, tile):
    s, c = (config_a_scene(), config_a_camera()) if which == "A" else (config_b_scene(), config_b_camera())
    f = _forward_ex(ctx, s, c, tile)
    W, H = c["width"], c["height"]
    op = O.project(s, f["ocam"])
    t = O.tiling(W, H, tile, c["depth_max"])
    ok, opay, orr = O.bin(op, t)
    oimg, oft, oli = O.composite(op, opay, orr, t, with_state=True)
    assert np.array_equal(f["ft"].download(np.float32, W * H).reshape(H, W).view(np.uint32), oft.view(np.uint32))
    assert np.array_equal(f["li"].download(np.uint32, W * H).reshape(H, W), oli)
    gimg = f["img"].download(np.float32, W * H * 4).reshape(H, W, 4)
    assert np.array_equal(gimg.view(np.uint32), oimg.view(np.uint32))


@pytest.mark.parametrize("which,tile", [("A", 16), ("A", 8), ("A", 32), ("B", 16)])
def test_recompute_backward_matches_oracle(ctx, which, tile):
    s, c = (config_a_scene(), config_a_camera()) if which == "A" else (config_b_scene(), config_b_camera())
    f = _forward_ex(ctx, s, c, tile)
    W, H = c["width"], c["height"]
    rng = np.random.default_rng(13)
    dl = rng.normal(0, 1e-3, (H, W, 4)).astype(np.float32)
    dl[..., 3] = 0
    dlb = ctx.storage_buffer(dl)
    pg = f["r"].alloc_projected_grad(ctx, s.n)
    f["r"].composite_backward_recompute(ctx, f["proj"], f["bins"], f["ft"], f["li"], dlb, pg, f["params"])
    g = pg.download(np.float32, s.n * 12).reshape(s.n, 12)
    op = O.project(s, f["ocam"])
    t = O.tiling(W, H, tile, c["depth_max"])
    ok, opay, orr = O.bin(op, t)
    oimg, oft, oli = O.composite(op, opay, orr, t, with_state=True)
    og = O.composite_backward_recompute(op, opay, orr, t, oft, oli, dl)
    for name, cols in (("dmean", [0, 1]), ("dopacity", [2]), ("dconic", [4, 5, 6]), ("dcolor", [8, 9, 10])):
        err = rel_norm_err(g[:, cols], og[:, cols])
        assert err < GRAD_TOL, f"{name}: relative error {err:.2e}"
    scale = np.abs(og).max(axis=0) + 1e-30
    assert (np.abs(g - og) / scale).max() < 1e-3


@pytest.mark.parametrize("which", ["A", "B"])
def test_project_backward_full_chain_bit_exact(ctx, which):
    from nanotracer_amd import _lib as L

    s, c = (config_a_scene(), config_a_camera()) if which == "A" else (config_b_scene(), config_b_camera())
    f = _forward(ctx, s, c)
    n = s.n
    rng = np.random.default_rng(17)
    pg_host = rng.normal(0, 1e-2, (n, 12)).astype(np.float32)
    pg_host[:, [3, 7, 11]] = 0
    pg = ctx.storage_buffer(pg_host)
    f["r"].project_backward(ctx, f["g"], pg, f["cam"], L.BWD_RECOMPUTE)
    gg = f["g"].grads()
    og = O.project_backward(s, f["ocam"], pg_host, full=True)
    for a, b in zip(gg.arrays(), og.arrays()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_recompute_and_parity_agree_without_saturation(ctx):
    """Where no pixel saturates, both backward modes are the same chain rule
    (they differ only in T = 1 - (1 - T) and where the walk starts)."""
    s = config_a_scene()
    s.logit[:] = -2.0  # low opacity: T never drops below 1e-4
    c = config_a_camera()
    f = _forward_ex(ctx, s, c)
    W, H = c["width"], c["height"]
    assert f["ft"].download(np.float32, W * H).min() > 1e-4
    dl = np.random.default_rng(19).normal(0, 1e-3, (H, W, 4)).astype(np.float32)
    dlb = ctx.storage_buffer(dl)
    a = f["r"].alloc_projected_grad(ctx, s.n)
    b = f["r"].alloc_projected_grad(ctx, s.n)
    f["r"].composite_backward(ctx, f["proj"], f["bins"], f["img"], dlb, a, f["params"])
    f["r"].composite_backward_recompute(ctx, f["proj"], f["bins"], f["ft"], f["li"], dlb, b, f["params"])
    ga = a.download(np.float32, s.n * 12).reshape(s.n, 12)
    gb = b.download(np.float32, s.n * 12).reshape(s.n, 12)
    assert rel_norm_err(ga, gb) < 1e-4


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_edge_records_blend_bit_exact(ctx, tile):
    """The blend's per-wave culling on records that stress its bounds
    (needles across the kappa limit, opacities at 1/255, NaN and infinite
    opacities, degenerate conics, off-screen means): image, final
    transmittance and last contributor equal the oracle's, which culls
    nothing (a NaN alpha clamps to 0.99 on both sides), and the recompute
    backward matches it to the usual bar."""
    from .helpers import edge_records
    W, H = 192, 128
    p = edge_records(3000, 0xED6E + tile, W, H)
    buf = ctx.storage_buffer(p)
    params = TilingParams(W, H, tile, 60.0)
    bins = TileBinner().bin(ctx, buf, len(p), params)
    r = Rasterizer()
    img = r.alloc_image(ctx, W, H)
    ft = ctx.storage_buffer_zeroed(W * H * 4)
    li = ctx.storage_buffer_zeroed(W * H * 4)
    r.composite_ex(ctx, buf, bins, img, params, ft, li)
    t = O.tiling(W, H, tile, 60.0)
    ok, opay, orr = O.bin(p, t)
    oimg, oft, oli = O.composite(p, opay, orr, t, with_state=True)
    gimg = img.download(np.float32, W * H * 4).reshape(H, W, 4)
    assert np.array_equal(gimg.view(np.uint32), oimg.view(np.uint32))
    assert np.array_equal(ft.download(np.float32, W * H).reshape(H, W).view(np.uint32), oft.view(np.uint32))
    assert np.array_equal(li.download(np.uint32, W * H).reshape(H, W), oli)
    rng = np.random.default_rng(29)
    dl = rng.normal(0, 1e-3, (H, W, 4)).astype(np.float32)
    dl[..., 3] = 0
    pg = r.alloc_projected_grad(ctx, len(p))
    r.composite_backward_recompute(ctx, buf, bins, ft, li, ctx.storage_buffer(dl), pg, params)
    g = pg.download(np.float32, len(p) * 12).reshape(len(p), 12)
    og = O.composite_backward_recompute(p, opay, orr, t, oft, oli, dl)
    rows = np.isfinite(p[:, 11])  # a NaN or infinite opacity has NaN gradients of its own
    assert np.isfinite(og[rows]).all()
    for name, cols in (("dmean", [0, 1]), ("dopacity", [2]), ("dconic", [4, 5, 6]), ("dcolor", [8, 9, 10])):
        err = rel_norm_err(g[rows][:, cols], og[rows][:, cols])
        assert err < GRAD_TOL, f"{name}: relative error {err:.2e}"
