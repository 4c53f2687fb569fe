"""The counting form of the blend kernels (a measurement aid) and the
backward's instruction-mix probe.

- A counted training step gives the same images as an uncounted one, bit
  for bit, and the same per-view losses.
- The forward's blended (pixel, splat) pairs equal the oracle forward's,
  counted pixel by pixel (orc_composite's own blend decisions: the device
  forward is bit-exact with it, so the counts agree exactly).
- In recompute mode the backward takes part at exactly the pairs the
  forward blended (it repeats the forward's tests from the last contributor
  back), so its contributing count equals the forward's blended count.
- The probe reports a positive, finite rate for every tile size and LDS
  batch of the backward.
- The FP32 FMA ceiling probe reports a consistent rate."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd import _lib as L
from nanotracer_amd import synth
from nanotracer_amd.api import CameraUniform, GpuSplatBuffer, ViewBatch, ViewSpec

pytestmark = pytest.mark.gpu

RES = ((320, 240), (480, 270), (256, 256))


def _views(n_splats):
    out = []
    for eye, w, h in synth.orbit_views(len(RES), 11, 6.0, RES):
        cam = CameraUniform.from_pose(eye, (0, 0, 0), (0, 1, 0), math.pi / 3, w, h, n_splats)
        out.append((ViewSpec(cam, w, h, 12.0), O.camera_like(cam)))
    return out


@pytest.mark.parametrize("tile", [16, 32])
def test_counted_step_equals_uncounted_and_the_oracle(ctx, tile):
    s = synth.make_scene(20_000, 0x5EED0002, 3.0)
    views = _views(s.n)
    specs = [v[0] for v in views]
    targets = []
    for spec, ocam in views:
        t = O.tiling(spec.width, spec.height, tile, spec.depth_max)
        p = O.project(synth.perturb_means(s, 5), ocam)
        _, pay, rg = O.bin(p, t)
        targets.append(O.composite(p, pay, rg, t))
    tbuf = ctx.storage_buffer(np.concatenate([x.reshape(-1) for x in targets]))
    b = ViewBatch(ctx, 8, s.n, sum(v.width * v.height for v in specs), 2_000_000)

    g = GpuSplatBuffer.upload(ctx, s)
    b.train_step(g, specs, tbuf, 1, tile, L.LOSS_L1, L.BWD_RECOMPUTE)
    plain = [b.image(k) for k in range(len(views))]
    plain_losses = b.losses()

    g2 = GpuSplatBuffer.upload(ctx, s)
    b.set_eval_counting(True)
    b.train_step(g2, specs, tbuf, 1, tile, L.LOSS_L1, L.BWD_RECOMPUTE)
    c = b.eval_counts()
    for k in range(len(views)):
        assert np.array_equal(b.image(k).view(np.uint32), plain[k].view(np.uint32)), k
    assert np.allclose(b.losses(), plain_losses, rtol=1e-12, atol=0)

    blended = 0
    for spec, ocam in views:
        t = O.tiling(spec.width, spec.height, tile, spec.depth_max)
        p = O.project(s, ocam)
        _, pay, rg = O.bin(p, t)
        blended += int(O.composite_used(p, pay, rg, t).sum(dtype=np.uint64))
    print(c, blended, flush=True)
    assert c["fwd_blended_pairs"] == blended > 0
    assert c["bwd_contributing_pairs"] == c["fwd_blended_pairs"]
    assert c["fwd_wave_evals"] > 0 and c["bwd_wave_evals"] > 0

    # counting off again: the counters are no longer written
    b.set_eval_counting(False)
    b.train_step(g2, specs, tbuf, 2, tile, L.LOSS_L1, L.BWD_RECOMPUTE)
    assert b.eval_counts() == dict(fwd_wave_evals=0, fwd_blended_pairs=0, bwd_wave_evals=0,
                                   bwd_contributing_pairs=0)
    b.free()


@pytest.mark.parametrize("tile,batch", [(8, 64), (16, 128), (32, 128), (32, 256)])
def test_backward_mix_probe(ctx, tile, batch):
    r = ctx.probe_blend_mix(tile, batch)
    assert math.isfinite(r) and r > 1e6


def test_valu_probe_reports_the_fma_ceiling(ctx):
    """nts_probe_valu: independent FP32 FMA chains at eight waves per SIMD.
    A wave64 FMA cannot issue faster than one per SIMD-cycle, and the two
    figures the probe returns agree with each other through the device's
    compute units (TFLOP/s = SIMD-cycle rate x 4 SIMDs x CUs x 64 lanes x
    2 flops x clock): an MI355X has 256 CUs."""
    v = ctx.probe_valu()
    r, t = v["fma_per_simd_cycle"], v["fma_tflops"]
    assert 0.1 < r < 1.0, v
    assert 10.0 < t < 400.0, v
    pk = ctx.probe_peaks(64 << 20)
    cus = t * 1e12 / (r * 4 * 64 * 2 * pk["clock_ghz"] * 1e9)
    assert 200 < cus < 320, (v, pk, cus)
