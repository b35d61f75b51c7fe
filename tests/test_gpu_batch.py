"""GPU tests of batches of views, their hipGraph replays and the launch-shape
tuning.

  * a batch of views with different resolutions (ragged against every tile
    shape), cameras, settings, light sets and meshes gives, per view, the
    frame that risltc_shade gives for that view alone, bit for bit;
  * a graph replays the batch: every replay, after an update with new noise
    seeds and with other work on the context's own buffers in between, equals
    the eager batch of the same views;
  * no tuning (tile shape, cull chunk, LDS staging, trace refill) changes a
    bit, and the exact build under a non-default tuning still equals the
    oracle.
"""
import math

import numpy as np
import pytest

import orc
from risltc_amd import abi, synth

pytestmark = pytest.mark.gpu

SUBSET = 23  # lights of the second light set


def constants(ctx, scene, camera, w, h, seed=0):
    import risltc_amd
    return risltc_amd.write_frame_constants(camera, w, h, scene.mesh.dequantization_factor,
                                            scene.mesh.dequantization_summand, ctx.ltc_constants, noise_seed=seed,
                                            animate=seed != 0)


def view_specs(scene):
    """(camera, width, height, settings, light set) per view; light set 1 is
    the first SUBSET lights."""
    cam0 = scene.camera
    cam1 = synth.make_camera((6.0, -12.0, 5.0), math.pi * 0.8, math.pi / 2 - 0.3)
    cam2 = synth.make_camera((-4.0, 3.0, 9.0), math.pi * 1.7, math.pi / 2 - 0.9, fov=1.2)
    return [
        (cam0, 256, 256, scene.settings(sample_count_light=2, ris_candidate_count=8), 0),
        (cam1, 200, 120, scene.settings(sample_count_light=1, ris_candidate_count=16), 1),
        (cam2, 97, 61, scene.settings(sample_count_light=3, ris_candidate_count=4, shadow_rays=0,
                                      mis_heuristic=abi.MIS_BALANCE), 0),
        (cam0, 64, 72, scene.settings(sample_count_light=1, ris_candidate_count=8, fast_atan=1), 1),
        (cam1, 33, 5, scene.settings(sample_count_light=2, light_sampling=abi.LIGHT_UNIFORM), 0),
    ]


@pytest.fixture(scope="module")
def batch(small_scene, ltc_table):
    import risltc_amd
    import torch
    sc = small_scene
    ctx = risltc_amd.Context(0)
    ctx.load_scene(sc, *ltc_table)
    recs, vmax = synth.pack_lights(sc.lights[:SUBSET])
    set_id = ctx.add_light_set(recs, SUBSET, vmax)
    assert set_id == 1
    # a context whose main lights are the subset: single-view references
    ctx_sub = risltc_amd.Context(0)
    ctx_sub.load_scene(sc, *ltc_table)
    ctx_sub.set_lights(recs, SUBSET, vmax)
    views, refs, bufs = [], [], []
    for i, (cam, w, h, s, ls) in enumerate(view_specs(sc)):
        c = constants(ctx, sc, cam, w, h, seed=i)
        vis = torch.empty(w * h, dtype=torch.int32, device="cuda")
        ctx.render_visibility(c, vis)
        out = torch.full((w * h * 4,), float("nan"), dtype=torch.float32, device="cuda")
        ref = torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
        (ctx_sub if ls else ctx).shade(c, s, vis, ref)
        views.append(ctx.view(c, s, vis, out, ls))
        refs.append(ref)
        bufs.append((c, s, vis, out, ls, cam, w, h))
    torch.cuda.synchronize()
    yield dict(ctx=ctx, ctx_sub=ctx_sub, views=views, refs=refs, bufs=bufs, scene=sc, ltc=ltc_table)
    ctx.close()
    ctx_sub.close()


def bits(t):
    return t.cpu().numpy().view(np.uint32)


def test_batch_equals_single_views(batch):
    import torch
    ctx = batch["ctx"]
    ctx.shade_batch(batch["views"])
    torch.cuda.synchronize()
    for i, ((_, _, _, out, *_), ref) in enumerate(zip(batch["bufs"], batch["refs"])):
        assert np.array_equal(bits(out), bits(ref)), i
    # the frames are not trivial: every view has lit pixels
    for _, _, _, out, *_ in batch["bufs"]:
        o = out.cpu().numpy().reshape(-1, 4)[:, :3]
        assert np.isfinite(o).all() and (o > 0).mean() > 0.05


def test_batch_rejects_bad_views(batch):
    import risltc_amd
    import torch
    ctx = batch["ctx"]
    c, s, vis, out, *_ = batch["bufs"][0]
    bad = ctx.view(c, s, vis, out, light_set=7)
    with pytest.raises(risltc_amd.RisltcError, match="E_ARG"):
        ctx.shade_batch([bad])
    # learned views need clusters (this context has none) and cannot be mixed
    # with plain views
    learned = ctx.view(c, batch["scene"].settings(proposal=abi.PROPOSAL_LEARNED), vis, out)
    with pytest.raises(risltc_amd.RisltcError, match="E_STATE"):
        ctx.shade_batch([learned])
    with pytest.raises(risltc_amd.RisltcError, match="E_UNSUPPORTED"):
        ctx.shade_batch([batch["views"][0], learned])
    exact = ctx.view(c, batch["scene"].settings(exact_math=1), vis, out)
    with pytest.raises(risltc_amd.RisltcError, match="E_UNSUPPORTED"):
        ctx.shade_batch([batch["views"][1], exact])
    torch.cuda.synchronize()


def _reseeded(batch, seed0):
    """The batch's views with new noise seeds, and single-view references."""
    import torch
    ctx, sub = batch["ctx"], batch["ctx_sub"]
    views, refs = [], []
    for i, (_, s, vis, out, ls, cam, w, h) in enumerate(batch["bufs"]):
        c = constants(ctx, batch["scene"], cam, w, h, seed=seed0 + i)
        ref = torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
        (sub if ls else ctx).shade(c, s, vis, ref)
        views.append(ctx.view(c, s, vis, out, ls))
        refs.append(ref)
    return views, refs


def test_graph_replays_equal_eager(batch):
    import risltc_amd
    import torch
    ctx = batch["ctx"]
    sc = batch["scene"]
    g = ctx.graph(batch["views"])
    try:
        extra = torch.empty(sc.width * sc.height * 4, dtype=torch.float32, device="cuda")
        for rep in range(4):
            views, refs = _reseeded(batch, 100 * rep + 1)
            torch.cuda.synchronize()
            g.update(views)
            for _, _, _, out, *_ in batch["bufs"]:
                out.fill_(float("nan"))
            g.launch()
            torch.cuda.synchronize()
            for i, ((_, _, _, out, *_), ref) in enumerate(zip(batch["bufs"], refs)):
                assert np.array_equal(bits(out), bits(ref)), (rep, i)
            # eager work on the context's own buffers between replays
            c0, s0, vis0, *_ = batch["bufs"][0]
            ctx.shade(c0, s0, vis0, extra)
            ctx.shade_batch(batch["views"][2:])
            torch.cuda.synchronize()
        # a view of another shape does not fit the captured launch
        c, s, vis, out, ls, cam, w, h = batch["bufs"][1]
        wrong = list(batch["views"])
        wrong[1] = ctx.view(c, sc.settings(sample_count_light=2, ris_candidate_count=16), vis, out, ls)
        with pytest.raises(risltc_amd.RisltcError, match="E_ARG"):
            g.update(wrong)
        # a tuning change invalidates the captured descriptors
        t = ctx.get_tuning()
        ctx.set_tuning(**t.as_dict())
        with pytest.raises(risltc_amd.RisltcError, match="E_STATE"):
            g.launch()
    finally:
        g.close()


TUNINGS = [
    dict(wave_width=8, cull_chunk=1, lds_lights=0, trace_refill=1),
    dict(wave_width=32, cull_chunk=7, lds_lights=13, trace_refill=17),
    dict(wave_width=64, cull_chunk=32, lds_lights=2048, trace_refill=64),
    dict(wave_width=16, cull_chunk=16, lds_lights=1024, trace_refill=48),
]


def test_tunings_change_no_bit(batch):
    import torch
    ctx = batch["ctx"]
    default = ctx.get_tuning().as_dict()
    try:
        for t in TUNINGS:
            ctx.set_tuning(**t)
            for _, _, _, out, *_ in batch["bufs"]:
                out.fill_(float("nan"))
            ctx.shade_batch(batch["views"])
            torch.cuda.synchronize()
            for i, ((_, _, _, out, *_), ref) in enumerate(zip(batch["bufs"], batch["refs"])):
                assert np.array_equal(bits(out), bits(ref)), (t, i)
    finally:
        ctx.set_tuning(**default)


def test_exact_build_under_tuning_equals_oracle(batch):
    import torch
    ctx = batch["ctx"]
    sc = batch["scene"]
    default = ctx.get_tuning().as_dict()
    c, _, vis, *_ = batch["bufs"][1]
    cam, w, h = batch["bufs"][1][5:8]
    s = sc.settings(exact_math=1, sample_count_light=1, ris_candidate_count=8)
    out = torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
    try:
        ctx.set_tuning(wave_width=64, cull_chunk=5, lds_lights=0, trace_refill=24)
        ctx.shade(c, s, vis, out)
        torch.cuda.synchronize()
    finally:
        ctx.set_tuning(**default)
    o = orc.Oracle(sc, *batch["ltc"])
    ref, _, _ = o.shade(c, s, vis.cpu().numpy().view(np.uint32))
    assert np.array_equal(bits(out).reshape(-1, 4), ref.view(np.uint32))


def test_autotune_keeps_results(batch):
    import torch
    ctx = batch["ctx"]
    default = ctx.get_tuning().as_dict()
    try:
        t = ctx.autotune(batch["views"], repeats=2)
        assert t.wave_width in (8, 16, 32, 64) and 1 <= t.cull_chunk <= 32
        assert t.lds_lights <= 2048 and 1 <= t.trace_refill <= 64
        assert ctx.get_tuning().as_dict() == t.as_dict()
        ctx.shade_batch(batch["views"])
        torch.cuda.synchronize()
        for i, ((_, _, _, out, *_), ref) in enumerate(zip(batch["bufs"], batch["refs"])):
            assert np.array_equal(bits(out), bits(ref)), i
    finally:
        ctx.set_tuning(**default)


def test_batch_with_several_meshes(small_scene, ltc_table):
    """Views on two meshes (with their own hierarchies and light sets) in one
    batch, and replayed as a graph: each view equals risltc_shade on a
    context holding that view's mesh and lights, bit for bit."""
    import risltc_amd
    import torch
    a_sc = small_scene
    b_sc = synth.small_scene(seed=9, width=120, height=90, light_count=40)
    ctx = risltc_amd.Context(0)
    ref_b = risltc_amd.Context(0)
    ref_ab = risltc_amd.Context(0)
    try:
        ctx.load_scene(a_sc, *ltc_table)
        sid = ctx.add_scene(b_sc)
        lid = ctx.add_light_set(b_sc.light_records, b_sc.light_count, b_sc.light_vmax)
        assert sid == 1 and lid == 1
        ref_b.load_scene(b_sc, *ltc_table)
        ref_ab.load_scene(a_sc, *ltc_table)
        ref_ab.set_lights(b_sc.light_records, b_sc.light_count, b_sc.light_vmax)
        cam2 = synth.make_camera((5.0, -11.0, 6.0), math.pi * 0.85, math.pi / 2 - 0.4)
        specs = [(a_sc, 0, 0, a_sc.camera, 128, 96, ctx),
                 (b_sc, sid, lid, b_sc.camera, 150, 70, ref_b),
                 (a_sc, 0, lid, cam2, 77, 64, ref_ab),
                 (b_sc, sid, lid, cam2, 96, 96, ref_b)]
        views, outs, refs, keep = [], [], [], []
        for i, (sc, scene_id, light_set, cam, w, h, ref_ctx) in enumerate(specs):
            c = risltc_amd.write_frame_constants(cam, w, h, sc.mesh.dequantization_factor,
                                                 sc.mesh.dequantization_summand, ctx.ltc_constants, noise_seed=i + 1,
                                                 animate=True)
            vis = torch.empty(w * h, dtype=torch.int32, device="cuda")
            ctx.render_visibility_scene(scene_id, c, vis)
            s = sc.settings(sample_count_light=2, ris_candidate_count=8 + 8 * i)
            out = torch.full((w * h * 4,), float("nan"), dtype=torch.float32, device="cuda")
            ref = torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
            ref_ctx.shade(c, s, vis, ref)
            views.append(ctx.view(c, s, vis, out, light_set=light_set, scene=scene_id))
            outs.append(out)
            refs.append(ref)
            keep.append(vis)
        torch.cuda.synchronize()
        ctx.shade_batch(views)
        torch.cuda.synchronize()
        for i, (o, r) in enumerate(zip(outs, refs)):
            assert np.array_equal(bits(o), bits(r)), i
            assert (o.cpu().numpy().reshape(-1, 4)[:, :3] > 0).mean() > 0.05, i
        g = ctx.graph(views)
        try:
            for _ in range(3):
                for o in outs:
                    o.fill_(float("nan"))
                g.launch()
                torch.cuda.synchronize()
                for i, (o, r) in enumerate(zip(outs, refs)):
                    assert np.array_equal(bits(o), bits(r)), i
        finally:
            g.close()
    finally:
        for c_ in (ctx, ref_b, ref_ab):
            c_.close()


@pytest.mark.parametrize("K,width,outputs,per_light", [(16, 64, 256, True), (48, 128, 1024, False)])
def test_learned_batch_equals_single_views(small_scene, ltc_table, K, width, outputs, per_light):
    """Learned views (MLP source with per-view tables, and the cache-only
    mode) in one batch and in a graph: each view's frame equals its
    risltc_shade call from the same cache state bit for bit, and the batch
    adds to the shared visibility cache exactly the counts of the single
    calls together. Tables of clusters with a per-light cache; tables of
    quads of clusters (1024 outputs) with a per-cluster cache."""
    import risltc_amd
    import torch
    sc = small_scene
    ctx = risltc_amd.Context(0)
    try:
        ctx.load_scene(sc, *ltc_table)
        ctx.cluster_lights(K, 10)
        ctx.proposal_init(3, width=width, outputs=outputs)
        ctx.proposal_cache(12, 1.0, 1.0, per_light=per_light)
        specs = view_specs(sc)
        shapes = [(specs[0][0], 256, 256, 2, 8), (specs[1][0], 200, 120, 1, 16), (specs[2][0], 97, 61, 2, 4)]
        bufs = []
        for i, (cam, w, h, spp, M) in enumerate(shapes):
            c = constants(ctx, sc, cam, w, h, seed=i + 1)
            vis = torch.empty(w * h, dtype=torch.int32, device="cuda")
            ctx.render_visibility(c, vis)
            bufs.append((c, vis, w, h, spp, M))
        # warm the MLP and the cache on view 0
        c0, vis0, w0, h0, spp0, M0 = bufs[0]
        tmp = torch.empty(w0 * h0 * 4, dtype=torch.float32, device="cuda")
        s_warm = sc.settings(proposal=abi.PROPOSAL_LEARNED, sample_count_light=spp0, ris_candidate_count=M0)
        for f in range(3):
            cf = constants(ctx, sc, specs[0][0], w0, h0, seed=50 + f)
            ctx.proposal_update(cf, vis0)
            ctx.shade(cf, s_warm, vis0, tmp)
            ctx.proposal_train(cf, vis0, batch=2048, steps=1, lr=1e-2, seed=f)
        torch.cuda.synchronize()
        for mode in (abi.PROPOSAL_LEARNED, abi.PROPOSAL_CACHE):
            print("mode", mode, flush=True)
            views, outs, tables, settings = [], [], [], []
            for c, vis, w, h, spp, M in bufs:
                st = sc.settings(proposal=mode, sample_count_light=spp, ris_candidate_count=M, learned_floor_log2=3)
                tab = None
                if mode == abi.PROPOSAL_LEARNED:
                    tab = torch.empty(w * h * abi.proposal_table_bytes(outputs), dtype=torch.uint8, device="cuda")
                    ctx.proposal_update(c, vis, table=tab)
                out = torch.full((w * h * 4,), float("nan"), dtype=torch.float32, device="cuda")
                views.append(ctx.view(c, st, vis, out, proposal_table=tab))
                outs.append(out)
                tables.append(tab)
                settings.append(st)
            torch.cuda.synchronize()
            counts0 = ctx.proposal_cache_read()
            ctx.shade_batch(views)
            torch.cuda.synchronize()
            counts_b = ctx.proposal_cache_read()
            inc = np.zeros(counts0.shape, dtype=np.int64)
            refs = []
            for (c, vis, w, h, spp, M), st in zip(bufs, settings):
                ctx.proposal_cache_write(counts0)
                if mode == abi.PROPOSAL_LEARNED:
                    ctx.proposal_update(c, vis)
                ref = torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
                ctx.shade(c, st, vis, ref)
                torch.cuda.synchronize()
                inc += ctx.proposal_cache_read().astype(np.int64) - counts0
                refs.append(ref)
            for i, (out, ref) in enumerate(zip(outs, refs)):
                assert np.array_equal(bits(out), bits(ref)), (mode, i)
                o = out.cpu().numpy().reshape(-1, 4)[:, :3]
                assert np.isfinite(o).all() and (o > 0).mean() > 0.05
            assert inc.sum() > 0
            assert np.array_equal(counts_b.astype(np.int64), counts0 + inc)
            # the same views as a graph: replays from the same cache state
            g = ctx.graph(views)
            try:
                for rep in range(3):
                    ctx.proposal_cache_write(counts0)
                    for out in outs:
                        out.fill_(float("nan"))
                    g.launch()
                    torch.cuda.synchronize()
                    for i, (out, ref) in enumerate(zip(outs, refs)):
                        assert np.array_equal(bits(out), bits(ref)), (mode, rep, i)
                    assert np.array_equal(ctx.proposal_cache_read(), counts_b)
                    # eager work between replays (a plain frame, which also counts into the cache)
                    ctx.shade(bufs[0][0], sc.settings(sample_count_light=2), bufs[0][1], tmp)
                    torch.cuda.synchronize()
                # a new cache invalidates the graph
                ctx.proposal_cache(12, 1.0, 1.0, per_light=per_light)
                with pytest.raises(risltc_amd.RisltcError, match="E_STATE"):
                    g.launch()
                ctx.proposal_cache_write(counts_b)
            finally:
                g.close()
    finally:
        ctx.close()
