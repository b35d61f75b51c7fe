"""Files on either side of the shading pass on the GPU: a scene stored as
*.vks plus a quicksave, read back by the library, renders the same frame bit
for bit as the in-memory scene, and the accumulated frame written as .hdr
decodes to it within binary16 + RGBE precision."""
import ctypes as C

import numpy as np
import pytest

from risltc_amd import abi, formats, synth

pytestmark = pytest.mark.gpu


def _render(ctx, scene, records, vmax, mesh, camera, ltc_table, lights=None, ltc_dir=None):
    import torch
    if lights is not None:
        ctx.set_polygonal_lights(lights)
    else:
        ctx.set_lights(records, scene.light_count, vmax)
    if ltc_dir is not None:
        ctx.load_ltc_table(ltc_dir, 51)
    else:
        ctx.set_ltc_table(*ltc_table)
    ctx.set_scene(mesh, scene.materials)
    ctx.scene = scene
    consts = ctx.frame_constants(camera=camera)
    P = scene.width * scene.height
    vis = torch.empty(P, dtype=torch.int32, device="cuda")
    out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
    ctx.render_visibility(consts, vis)
    ctx.shade(consts, scene.settings(exact_math=1), vis, out)
    torch.cuda.synchronize()
    return vis.cpu().numpy(), out


def test_vks_and_quicksave_render_like_memory(ltc_table, tmp_path):
    import torch
    import risltc_amd
    print("case: small_scene seed 5 via .vks + quicksave", flush=True)
    scene = synth.small_scene(seed=5, width=192, height=128, light_count=24, spp=2)
    lights = [l.to_ctypes() for l in scene.lights]
    keep = [(l._ps, l._ws) for l in scene.lights]
    arr = (abi.PolygonalLight * len(lights))(*lights)
    size = len(lights) * abi.light_record_size(scene.light_vmax)
    buf = (C.c_uint8 * size)()
    assert risltc_amd.lib().risltc_write_lights(arr, len(lights), scene.light_vmax, buf, size) == size
    ctx = risltc_amd.Context(0)
    vis_a, out_a = _render(ctx, scene, bytes(buf), scene.light_vmax, scene.mesh, scene.camera, ltc_table)

    formats.save_vks(tmp_path / "s.vks", scene.mesh)
    formats.save_quicksave(tmp_path / "s.save", scene.camera, lights)
    vks = formats.load_vks(tmp_path / "s.vks")
    q = formats.load_quicksave(tmp_path / "s.save")
    records, vmax = q.light_records()
    vis_b, out_b = _render(ctx, scene, records, vmax, vks, q.camera, ltc_table)
    assert np.array_equal(vis_a, vis_b)
    a = out_a.cpu().numpy()
    assert np.array_equal(a.view(np.uint32), out_b.cpu().numpy().view(np.uint32))
    # the conveniences: lights straight from polygonal_light_t, the LTC
    # table from fit<i>.dat files
    from risltc_amd import ltcfit
    ltcfit.write_fit_dir(str(tmp_path / "fit"), 51)
    q2 = formats.load_quicksave(tmp_path / "s.save")
    vis_c, out_c = _render(ctx, scene, None, 0, vks, q2.camera, ltc_table, lights=q2.lights,
                           ltc_dir=tmp_path / "fit")
    assert np.array_equal(a.view(np.uint32), out_c.cpu().numpy().view(np.uint32))

    accum = torch.zeros_like(out_a)
    ctx.accumulate(accum, out_a, scene.width * scene.height, 0)
    torch.cuda.synchronize()
    img = accum.cpu().numpy().reshape(scene.height, scene.width, 4)
    formats.write_hdr(tmp_path / "f.hdr", img, half_round=True)
    rgb, _ = formats.read_hdr(tmp_path / "f.hdr")
    # negative components are written as 0
    x = np.maximum(img[..., :3].astype(np.float64), 0.0)
    m = x.max(axis=2, keepdims=True)
    # binary16 (2^-11 relative, 2^-25 absolute for subnormal halves), then
    # RGBE (< max/128 absolute)
    bound = x * 2.0 ** -11 + 2.0 ** -25 + m * (1 + 2.0 ** -11) / 128.0
    err = np.abs(rgb - x)
    bad = np.argwhere(err > bound)
    assert bad.size == 0, (bad[:4], x[tuple(bad[0][:2])], rgb[tuple(bad[0][:2])], img[tuple(bad[0][:2])])
    assert np.count_nonzero(rgb) > rgb.size // 4
    ctx.close()
    del keep
