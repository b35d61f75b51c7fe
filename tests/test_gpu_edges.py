"""Edge cases of the shading pass on the GPU, each against the FP32 oracle
with the exact build (frames bit for bit): ragged and one-pixel viewports,
frames with no shaded pixel, visibility ids past the mesh, a single light,
the largest candidate count, many samples per pixel, and more lights than
the shade kernel stages in LDS."""
import numpy as np
import pytest

import orc
from risltc_amd import abi, synth

pytestmark = pytest.mark.gpu


def _setup(scene, ltc_table):
    import risltc_amd
    ctx = risltc_amd.Context(0)
    ctx.load_scene(scene, *ltc_table)
    return ctx, orc.Oracle(scene, *ltc_table)


def _constants(ctx, scene, w, h, camera=None):
    import risltc_amd
    return risltc_amd.write_frame_constants(camera or scene.camera, w, h, scene.mesh.dequantization_factor,
                                            scene.mesh.dequantization_summand, ctx.ltc_constants)


def _shade(ctx, consts, settings, vis_np):
    import torch
    w, h = consts.viewport_size[0], consts.viewport_size[1]
    vis = torch.from_numpy(np.ascontiguousarray(vis_np).view(np.int32)).cuda()
    out = torch.full((w * h * 4,), float("nan"), dtype=torch.float32, device="cuda")
    ctx.shade(consts, settings, vis, out)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(-1, 4)


def _bitwise(ctx, oracle, consts, settings, vis_np=None):
    if vis_np is None:
        vis_np = oracle.render_visibility(consts)
    got = _shade(ctx, consts, settings, vis_np)
    ref, _, _ = oracle.shade(consts, settings, vis_np)
    bad = np.argwhere(np.any(got.view(np.uint32) != ref.view(np.uint32), axis=1)).ravel()
    assert bad.size == 0, (bad[:5], got[bad[:2]], ref[bad[:2]])
    return got, vis_np


@pytest.fixture(scope="module")
def small(small_scene, ltc_table):
    ctx, oracle = _setup(small_scene, ltc_table)
    yield dict(ctx=ctx, oracle=oracle, scene=small_scene)
    ctx.close()


@pytest.mark.parametrize("w,h", [(1, 1), (7, 3), (65, 1), (1, 130), (33, 17)])
def test_ragged_viewports(small, w, h):
    print("viewport", w, h, flush=True)
    sc = small["scene"]
    consts = _constants(small["ctx"], sc, w, h)
    _bitwise(small["ctx"], small["oracle"], consts, sc.settings(exact_math=1, sample_count_light=2))


def test_no_shaded_pixel(small):
    """Background everywhere, ids past the mesh (treated as background) and
    light pixels (white times the exposure): no sample is shaded."""
    sc = small["scene"]
    w, h = 40, 24
    consts = _constants(small["ctx"], sc, w, h)
    rng = np.random.default_rng(3)
    vis = np.full(w * h, 0xFFFFFFFF, dtype=np.uint32)
    vis[::3] = rng.integers(sc.mesh.triangle_count, 0x7FFFFFFF, vis[::3].size, dtype=np.int64).astype(np.uint32)
    vis[1::5] = 0x80000000 | rng.integers(0, sc.mesh.triangle_count, vis[1::5].size).astype(np.uint32)
    got, _ = _bitwise(small["ctx"], small["oracle"], consts, sc.settings(exact_math=1), vis)
    light = ((vis >> 31) == 1) & (vis != 0xFFFFFFFF)
    assert np.all(got[~light, :3] == 0.0) and np.all(got[:, 3] == 1.0)
    assert np.all(got[light, :3] == np.float32(consts.exposure_factor))


def test_largest_candidate_count_and_many_samples(small):
    sc = small["scene"]
    consts = _constants(small["ctx"], sc, 16, 12)
    _bitwise(small["ctx"], small["oracle"], consts, sc.settings(exact_math=1, ris_candidate_count=4096))
    consts = _constants(small["ctx"], sc, 9, 7)
    _bitwise(small["ctx"], small["oracle"], consts,
             sc.settings(exact_math=1, sample_count_light=64, ris_candidate_count=3))


def test_one_light(ltc_table):
    sc = synth.small_scene(seed=9, width=48, height=40, light_count=1)
    ctx, oracle = _setup(sc, ltc_table)
    try:
        for M in (1, 8):
            got, vis = _bitwise(ctx, oracle, ctx.frame_constants(), sc.settings(exact_math=1, ris_candidate_count=M))
        shaded = (vis >> 31) == 0
        assert (got[shaded, :3].sum(1) > 0).mean() > 0.2
    finally:
        ctx.close()


def test_more_lights_than_lds_staging(ltc_table):
    """3000 lights: the shade kernel stages at most 2048 bounding spheres in
    LDS and reads the rest from global memory."""
    sc = synth.small_scene(seed=4, width=48, height=32, light_count=3000)
    ctx, oracle = _setup(sc, ltc_table)
    try:
        got, vis = _bitwise(ctx, oracle, ctx.frame_constants(), sc.settings(exact_math=1, sample_count_light=2))
        shaded = (vis >> 31) == 0
        assert (got[shaded, :3].sum(1) > 0).mean() > 0.2
    finally:
        ctx.close()


@pytest.mark.parametrize("exact", [1, 0])
def test_non_finite_samples_are_pink(ltc_table, exact):
    """A light of infinite radiance makes the samples whose candidates see it
    infinite or NaN; such pixels are written as (1, 0, 0.8) / exposure
    before the exposure (shading_pass.frag.glsl:764-767). Exact build: the
    frame equals the oracle's bit for bit. Fast build: the pink pixels are
    those of the oracle except on at most 0.1% of pixels (a noise stream that
    splits at a specular PSA within 1e-6 of zero draws other candidates, and
    a candidate whose LTC integral is zero in one build and a denormal in the
    other turns inf * 0 into inf * x), and every pixel is finite."""
    import copy
    sc = copy.copy(synth.small_scene(seed=0, width=96, height=64))
    recs = bytearray(sc.light_records)
    stride = 48 + 16 * sc.light_vmax
    for k in range(0, sc.light_count, 9):  # every 9th light: radiance (inf, inf, inf)
        recs[k * stride:k * stride + 12] = np.full(3, np.inf, np.float32).tobytes()
    sc.light_records = bytes(recs)
    ctx, oracle = _setup(sc, ltc_table)
    try:
        consts = _constants(ctx, sc, sc.width, sc.height)
        s = sc.settings(exact_math=exact, sample_count_light=2, ris_candidate_count=8)
        vis = oracle.render_visibility(consts)
        got = _shade(ctx, consts, s, vis)
        ref, _, _ = oracle.shade(consts, s, vis)
        e = np.float32(consts.exposure_factor)
        pink = np.array([np.float32(1.0) / e * e, np.float32(0.0) / e * e, np.float32(0.8) / e * e, 1.0],
                        np.float32)
        is_pink = np.all(got.view(np.uint32) == pink.view(np.uint32), axis=1)
        ref_pink = np.all(ref.view(np.uint32) == pink.view(np.uint32), axis=1)
        assert ref_pink.sum() > 50 and is_pink.sum() > 50
        assert np.isfinite(got).all()
        if exact:
            assert np.array_equal(is_pink, ref_pink)
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        else:
            assert (is_pink != ref_pink).mean() <= 1e-3, (is_pink.sum(), ref_pink.sum())
    finally:
        ctx.close()
