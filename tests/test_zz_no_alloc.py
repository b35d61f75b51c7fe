"""risltc_shade_noalloc (no device allocation; the library's fixed scratch
array) against risltc_shade through the C-ABI: two calls in a row, the
second smaller, with different settings; each frame bitwise equal to
risltc_shade's. Sizes beyond the fixed scratch are refused on the host."""
import math

import numpy as np
import pytest

from risltc_amd import abi, synth

pytestmark = pytest.mark.gpu


def test_noalloc_equals_shade_on_two_calls(small_scene, ltc_table):
    import risltc_amd
    import torch
    sc = small_scene
    ctx = risltc_amd.Context(0)
    try:
        ctx.load_scene(sc, *ltc_table)
        cam2 = synth.make_camera((-4.0, 3.0, 9.0), math.pi * 1.7, math.pi / 2 - 0.9, fov=1.2)
        cases = [
            # 256 x 256 x 4 spp = 262144 pixel-samples, fast build
            (sc.camera, 256, 256, sc.settings(sample_count_light=4, ris_candidate_count=8), 3),
            # smaller and ragged, exact build, another heuristic
            (cam2, 97, 61, sc.settings(sample_count_light=2, ris_candidate_count=16, exact_math=1,
                                       mis_heuristic=abi.MIS_BALANCE), 4),
        ]
        runs = []
        for cam, w, h, s, seed in cases:
            c = risltc_amd.write_frame_constants(cam, w, h, sc.mesh.dequantization_factor,
                                                 sc.mesh.dequantization_summand, ctx.ltc_constants,
                                                 noise_seed=seed, animate=True)
            vis = torch.empty(w * h, dtype=torch.int32, device="cuda")
            ctx.render_visibility(c, vis)
            runs.append((c, s, vis, w * h))
        outs = []
        for c, s, vis, P in runs:  # two no-allocation calls in a row
            out = torch.full((P * 4,), float("nan"), dtype=torch.float32, device="cuda")
            print("noalloc %ux%u spp %u" % (c.viewport_size[0], c.viewport_size[1], s.sample_count_light),
                  flush=True)
            ctx.shade_noalloc(c, s, vis, out)
            outs.append(out)
        torch.cuda.synchronize()
        for (c, s, vis, P), out in zip(runs, outs):
            ref = torch.empty(P * 4, dtype=torch.float32, device="cuda")
            ctx.shade(c, s, vis, ref)
            torch.cuda.synchronize()
            a = out.cpu().numpy()
            b = ref.cpu().numpy()
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), \
                (c.viewport_size[0], np.nanmax(np.abs(a - b)))
            assert np.isfinite(a).all() and (a.reshape(-1, 4)[:, :3] > 0).mean() > 0.05
        # beyond the fixed scratch: refused before any GPU work
        c, s, vis, P = runs[0]
        big = sc.settings(sample_count_light=9)  # 256 * 256 * 9 > 2^19
        with pytest.raises(risltc_amd.RisltcError, match="E_UNSUPPORTED"):
            ctx.shade_noalloc(c, big, vis, outs[0])
    finally:
        ctx.close()
