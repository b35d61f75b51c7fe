"""GPU parity: the HIP kernels, called through the C-ABI, against the CPU
oracle on the same seeded inputs.

  exact_math = 1 : IEEE division and square root in the kernels; the whole
                   pass (visibility, every candidate record, every sample
                   record, the output) is bitwise equal to the FP32 oracle.
  exact_math = 0 : the production build (hardware rcp/rsq/sqrt). Clipped
                   vertex counts and BVH hit flags are still bitwise equal;
                   LTC values within 2e-5 relative; the RIS choice equal
                   unless the oracle's decision was within 1e-5 of a tie;
                   the final radiance within the tolerance stated below.
"""
import numpy as np
import pytest

import orc
import parity_rules as PR
from risltc_amd import abi, synth

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


@pytest.fixture(scope="module")
def small_gpu(small_scene, ltc_table):
    import risltc_amd
    torch = _torch()
    ctx = risltc_amd.Context(0)
    ctx.load_scene(small_scene, *ltc_table)
    oracle = orc.Oracle(small_scene, *ltc_table)
    consts = ctx.frame_constants()
    P = small_scene.width * small_scene.height
    vis = torch.empty(P, dtype=torch.int32, device="cuda")
    ctx.render_visibility(consts, vis)
    torch.cuda.synchronize()
    return dict(ctx=ctx, oracle=oracle, consts=consts, vis=vis, P=P, scene=small_scene)


def _shade(g, settings, debug=True):
    torch = _torch()
    P = g["P"]
    spp = settings.sample_count_light
    M = settings.ris_candidate_count
    out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
    if debug:
        rec = torch.zeros(P * spp * abi.SAMPLE_RECORD_FLOATS, dtype=torch.float32, device="cuda")
        cand = torch.zeros(P * spp * M * abi.CANDIDATE_RECORD_WORDS, dtype=torch.int32, device="cuda")
        g["ctx"].shade_debug(g["consts"], settings, g["vis"], out, rec, cand)
    else:
        g["ctx"].shade(g["consts"], settings, g["vis"], out)
    torch.cuda.synchronize()
    o = out.cpu().numpy().reshape(P, 4)
    if not debug:
        return o, None, None
    return (o, rec.cpu().numpy().reshape(P, spp, abi.SAMPLE_RECORD_FLOATS),
            cand.cpu().numpy().view(np.uint32).reshape(P, spp, M, abi.CANDIDATE_RECORD_WORDS))


def test_primary_visibility_bitwise(small_gpu):
    g = small_gpu
    vis_o = g["oracle"].render_visibility(g["consts"])
    vis_g = g["vis"].cpu().numpy().view(np.uint32)
    assert np.array_equal(vis_g, vis_o)
    assert (vis_o == 0xFFFFFFFF).sum() > 0 and ((vis_o >> 31) == 1).sum() > (vis_o == 0xFFFFFFFF).sum()


def test_shadow_rays_bitwise(small_gpu):
    """Any-hit flags of random rays: GPU LBVH == CPU LBVH == brute force."""
    torch = _torch()
    g = small_gpu
    rng = np.random.default_rng(5)
    n = 20000
    o = rng.uniform([-10, -10, 0], [10, 10, 8], (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tmax = rng.uniform(0.1, 20, n).astype(np.float32)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 1e-3
    rays[:, 4:7] = d
    rays[:, 7] = tmax
    rt = torch.from_numpy(rays).cuda()
    hits = torch.zeros(n, dtype=torch.int32, device="cuda")
    g["ctx"].trace_shadow_rays(rt, hits)
    torch.cuda.synchronize()
    h = hits.cpu().numpy()
    L = g["oracle"].L
    ref = np.array([L.orc_bvh_any_hit(orc.C.byref(g["oracle"].bvh), orc.ptr(o[i], orc.f32p),
                                      orc.ptr(d[i], orc.f32p), 1e-3, float(tmax[i]), None) for i in range(n)])
    assert np.array_equal(h, ref)
    assert 0.05 < h.mean() < 0.95


def test_bvh_records_bitwise_sah(small_gpu):
    """The device's binned-SAH hierarchy, laid out in pre-order, equals the
    oracle's restatement word for word (boxes, escapes, leaf order)."""
    g = small_gpu
    got = g["ctx"].bvh_records()
    ref = g["oracle"].bvh_records()
    assert got.shape == ref.shape
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_lbvh_builder_bitwise(small_scene, ltc_table):
    """The Morton-code LBVH builder: records equal the oracle's LBVH, primary
    visibility and random shadow rays bitwise equal to the oracle on it, and
    any-hit flags equal to those of the SAH tree (the flags do not depend on
    the tree)."""
    import risltc_amd
    torch = _torch()
    ctx = risltc_amd.Context(0)
    ctx.set_bvh_builder(abi.BVH_LBVH)
    ctx.load_scene(small_scene, *ltc_table)
    oracle = orc.Oracle(small_scene, *ltc_table, builder=abi.BVH_LBVH)
    sah = orc.Oracle(small_scene, *ltc_table, builder=abi.BVH_SAH)
    assert np.array_equal(ctx.bvh_records().view(np.uint32), oracle.bvh_records().view(np.uint32))
    consts = ctx.frame_constants()
    P = small_scene.width * small_scene.height
    vis = torch.empty(P, dtype=torch.int32, device="cuda")
    ctx.render_visibility(consts, vis)
    torch.cuda.synchronize()
    assert np.array_equal(vis.cpu().numpy().view(np.uint32), oracle.render_visibility(consts))
    rng = np.random.default_rng(6)
    n = 5000
    o = rng.uniform([-10, -10, 0], [10, 10, 8], (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    tmax = rng.uniform(0.1, 20, n).astype(np.float32)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7], rays[:, 7] = o, 1e-3, d, tmax
    hits = torch.zeros(n, dtype=torch.int32, device="cuda")
    ctx.trace_shadow_rays(torch.from_numpy(rays).cuda(), hits)
    torch.cuda.synchronize()
    h = hits.cpu().numpy()
    for orb in (oracle, sah):
        ref = np.array([orb.L.orc_bvh_any_hit(orc.C.byref(orb.bvh), orc.ptr(o[i], orc.f32p), orc.ptr(d[i], orc.f32p),
                                              1e-3, float(tmax[i]), None) for i in range(n)])
        assert np.array_equal(h, ref)


@pytest.mark.parametrize("spp,M", [(1, 8), (2, 8)])
def test_exact_shading_bitwise(small_gpu, spp, M):
    g = small_gpu
    s = g["scene"].settings(exact_math=1, sample_count_light=spp, ris_candidate_count=M)
    out, rec, cand = _shade(g, s)
    vis = g["vis"].cpu().numpy().view(np.uint32)
    ref, rrec, rcand = g["oracle"].shade(g["consts"], s, vis, records=True, candidates=True)
    # candidates: index, clip counts, p_hat, w -- all bitwise
    assert np.array_equal(cand, rcand)
    # sample records: selected light, reservoir, clip counts, PSA, rays, estimates
    ri = rec.view(np.uint32)
    oi = rrec.view(np.uint32)
    for f in (abi.REC_LIGHT, abi.REC_WSUM, abi.REC_PHAT, abi.REC_W, abi.REC_CLIPD, abi.REC_CLIPS, abi.REC_PSAD,
              abi.REC_PSAS):
        assert np.array_equal(ri[..., f], oi[..., f]), f
    for f in list(range(abi.REC_DIR, abi.REC_DIR + 6)) + [abi.REC_TMAX, abi.REC_TMAX + 1] + \
            list(range(abi.REC_EVIS, abi.REC_EVIS + 12)):
        assert np.array_equal(ri[..., f], oi[..., f]), f
    # ray states (0 none, 1 below the shading horizon, 2 occluded, 3 visible):
    # the GPU answers traced rays in the trace kernel, the oracle inline
    gs = ri[..., abi.REC_STATE:abi.REC_STATE + 2]
    os_ = oi[..., abi.REC_STATE:abi.REC_STATE + 2]
    bad = np.argwhere(np.any(gs != os_, axis=-1))
    assert bad.size == 0, ("ray states differ", bad[:5], gs[tuple(bad[0])], os_[tuple(bad[0])])
    assert np.array_equal(ri[..., abi.REC_SEED], oi[..., abi.REC_SEED])
    cg = ri[..., abi.REC_COLOR:abi.REC_COLOR + 3]
    co = oi[..., abi.REC_COLOR:abi.REC_COLOR + 3]
    bad = np.argwhere(np.any(cg != co, axis=-1))
    assert bad.size == 0, ("sample colours differ", bad[:5])
    # the frame
    bad = np.argwhere(np.any(out.view(np.uint32) != ref.view(np.uint32), axis=-1))
    assert bad.size == 0, ("frame differs", bad.size, bad[:5].ravel(), out[bad[:3, 0]], ref[bad[:3, 0]])
    shaded = (vis >> 31) == 0
    assert (out[shaded, :3].sum(1) > 0).mean() > 0.5


@pytest.mark.parametrize("variant", [
    dict(light_sampling=abi.LIGHT_UNIFORM),
    dict(mis_heuristic=abi.MIS_BALANCE),
    dict(mis_heuristic=abi.MIS_POWER),
    dict(mis_heuristic=abi.MIS_WEIGHTED),
    dict(mis_heuristic=abi.MIS_OPTIMAL),
    dict(shadow_rays=0),
    dict(fast_atan=1),
])
def test_exact_settings_variants_bitwise(small_gpu, variant):
    g = small_gpu
    s = g["scene"].settings(exact_math=1, sample_count_light=2, **variant)
    out, _, _ = _shade(g, s, debug=False)
    vis = g["vis"].cpu().numpy().view(np.uint32)
    ref, _, _ = g["oracle"].shade(g["consts"], s, vis)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def _quantiles(x, qs=(50, 90, 99, 99.9)):
    return np.array([np.percentile(x, q) for q in qs])


def fast_build_rules(oracle, consts, scene, vis, pix, out, rec, cand, s):
    """The production build's parity rules against the oracle on the listed
    pixels (out, rec, cand: the build's frame rows and records of them):

      * noise streams: a pixel-sample starts from the FP64 oracle's noise
        state unless an earlier sample of the pixel split the streams, and
        every split starts at a final sample whose specular PSA the FP64
        oracle puts within 1e-6 of zero (a zero PSA draws no specular
        direction, shading_pass.frag.glsl:365-372);
      * clipped vertex counts of every candidate: bitwise equal to the FP32
        oracle's wherever the streams agree;
      * RIS choice: equal to the FP32 oracle's (the reference's arithmetic)
        unless a reservoir decision of the sample comes out differently
        between the two builds' normalised weights w_i / w_sum_i, on at most
        0.1% of samples (the target of a small or distant polygon carries a
        cancellation error of up to ~1e-3 relative in either FP32 build, so
        a draw within that distance of a tie can flip); against the FP64 oracle the build differs on no more samples than
        IEEE FP32 does (1.5x + 3);
      * LTC targets: within 1e-5 relative of the FP64 integrator on the same
        pixel setup, for lights wholly above the horizon that subtend at
        least 0.03 sr (smaller or horizon-clipped polygons are sums of
        nearly cancelling edge terms, which IEEE FP32 cannot hold to 1e-5).

    Returns the statistics."""
    ref, rrec, rcand = oracle.shade(consts, s, vis, pixels=pix, records=True, candidates=True)
    _, rec64, c64 = oracle.shade(consts, s, vis, pixels=pix, fp64=True, records=True, candidates=True)
    synced, cause = PR.noise_sync(consts, pix, rec, rec64)
    first = (cause >= 0)
    assert np.all(cause[first] <= 1e-6), cause[first].max()
    # streams split only after a split-causing decision or a differing choice
    assert synced[:, 0].all()
    same32 = PR.sample_start_seeds(consts, pix, rec) == PR.sample_start_seeds(consts, pix, rrec)
    cl = cand[..., 1][same32]
    assert np.array_equal(cand[..., 0][same32], rcand[..., 0][same32])
    assert np.array_equal(cl, rcand[..., 1][same32])
    seeds32 = PR.sample_start_seeds(consts, pix, rrec)
    ris = PR.ris_tie_check(seeds32, cand, rcand, rec, rrec, scene.light_count, tie=1e-6, synced=same32)
    assert ris["replay_ok"] and ris["replay_disagreements"] == 0, ris
    flips = PR.weight_flips(seeds32, cand, rcand)
    mm = ris["mismatch"]
    assert np.all(flips[mm]), int((mm & ~flips).sum())
    assert mm.sum() <= 1e-3 * same32.sum(), int(mm.sum())
    seeds = PR.sample_start_seeds(consts, pix, rec64)
    ris64 = PR.ris_tie_check(seeds, cand, c64, rec, rec64, scene.light_count, tie=1e-6, synced=synced)
    ref64 = PR.ris_tie_check(seeds, rcand, c64, rrec, rec64, scene.light_count, tie=1e-6, synced=synced)
    assert ris64["replay_ok"] and ris64["replay_disagreements"] == 0, ris64
    assert ris64["mismatches"] <= 1.5 * ref64["mismatches"] + 3, (ris64["mismatches"], ref64["mismatches"])
    errs, n = PR.ltc_bar(oracle, consts, vis, pix, cand, rec64, c64, scene, synced)
    assert n > 0 and errs.max() <= 1e-5, (n, errs.max() if n else None)
    return dict(synced=float(synced.mean()), split_causes=int(first.sum()), ris=ris,
                fp32_oracle_mismatches=int(mm.sum()),
                fp64_mismatches=int(ris64["mismatches"]), fp32_oracle_fp64_mismatches=int(ref64["mismatches"]),
                ltc_checked=n,
                ltc_max=float(errs.max()), ref=ref, rrec=rrec, rcand=rcand, c64=c64)


def test_fast_shading_parity(small_gpu):
    """The production build against the oracle under the rules of
    fast_build_rules; beyond them, on every candidate the fast build's LTC
    target error against FP64 is distributed as IEEE FP32's (the LTC target
    of a small or distant polygon is a sum of nearly cancelling edge terms,
    polygon_sampling.glsl:523-530, so one-ulp changes in the vertex
    normalisation move it by up to ~1e-4 relative, in IEEE FP32 as much as
    with the hardware approximations)."""
    g = small_gpu
    s = g["scene"].settings(exact_math=0, sample_count_light=2)
    out, rec, cand = _shade(g, s)
    vis = g["vis"].cpu().numpy().view(np.uint32)
    pix = np.arange(g["P"], dtype=np.uint32)
    shaded = (vis >> 31) == 0
    r = fast_build_rules(g["oracle"], g["consts"], g["scene"], vis, pix[shaded], out[shaded], rec[shaded],
                         cand[shaded], s)
    print({k: v for k, v in r.items() if k in ("synced", "split_causes", "ltc_checked", "ltc_max", "fp64_mismatches",
                                               "fp32_oracle_fp64_mismatches", "fp32_oracle_mismatches")})
    ref, rrec, rcand, c64 = r["ref"], r["rrec"], r["rcand"], r["c64"]
    out, rec, cand = out[shaded], rec[shaded], cand[shaded]
    assert np.array_equal(cand[..., 1], c64[..., 1])
    pf = cand[..., 2].view(np.float32).astype(np.float64)
    px = rcand[..., 2].view(np.float32).astype(np.float64)
    p64 = c64[..., 2].view(np.float32).astype(np.float64)
    live = p64 > 0
    rel_fast = np.abs(pf - p64)[live] / p64[live]
    rel_exact = np.abs(px - p64)[live] / p64[live]
    qf, qx = _quantiles(rel_fast), _quantiles(rel_exact)
    assert np.all(qf <= 4.0 * qx + 3e-7), (qf, qx)
    assert (rel_fast < 1e-5).mean() >= (rel_exact < 1e-5).mean() - 0.02
    # RIS choice: equal to the FP32 oracle's except near ties
    li_g = rec[..., abi.REC_LIGHT].view(np.int32)
    li_o = rrec[..., abi.REC_LIGHT].view(np.int32)
    frac = (li_g == li_o).mean()
    assert frac > 0.995, frac
    # output: pixels whose choices agree
    same = np.all(li_g == li_o, axis=1)
    d = np.abs(out[same, :3] - ref[same, :3])
    rel = d / (np.abs(ref[same, :3]) + 1e-3)
    assert np.percentile(rel, 99) < 1e-3, _quantiles(rel)


COMPACT_CASES = [dict(exact_math=1, sample_count_light=2, ris_candidate_count=8),
                 dict(exact_math=1, sample_count_light=1, ris_candidate_count=32, mis_heuristic=abi.MIS_BALANCE),
                 dict(exact_math=0, sample_count_light=3, ris_candidate_count=32),
                 dict(exact_math=0, sample_count_light=1, ris_candidate_count=77, shadow_rays=0),
                 dict(exact_math=1, sample_count_light=2, ris_candidate_count=16, mis_heuristic=abi.MIS_POWER),
                 dict(exact_math=0, sample_count_light=1, ris_candidate_count=16, mis_heuristic=abi.MIS_WEIGHTED),
                 dict(exact_math=1, sample_count_light=1, ris_candidate_count=16, mis_heuristic=abi.MIS_OPTIMAL)]


@pytest.mark.parametrize("case", range(len(COMPACT_CASES)))
@pytest.mark.parametrize("tuning", [None, dict(wave_width=64, cull_chunk=5, lds_lights=0, trace_refill=16)])
def test_compacted_evaluation_equals_debug(small_gpu, case, tuning):
    """The non-debug kernels evaluate the surviving candidates compacted over
    the wave; the debug kernels (bitwise equal to the oracle in the exact
    build) evaluate them per lane. The frames are equal bit for bit, in both
    builds and under a non-default launch shape. The non-debug pass also
    writes the compact 8-word sample record (occluded estimates as NaN flag
    bits) under every heuristic but the optimal one, the debug pass the full
    16-word record, so these cases also pin the compact record against the
    full one."""
    g = small_gpu
    s = g["scene"].settings(**COMPACT_CASES[case])
    ctx = g["ctx"]
    default = ctx.get_tuning().as_dict()
    try:
        if tuning:
            ctx.set_tuning(**tuning)
        a, _, _ = _shade(g, s, debug=False)
        b, _, _ = _shade(g, s, debug=True)
    finally:
        ctx.set_tuning(**default)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (a[:, :3] > 0).any()


def test_deterministic_repeat(small_gpu):
    """Two passes on the same inputs give the same bits (ray compaction order
    varies, results must not)."""
    g = small_gpu
    s = g["scene"].settings(sample_count_light=4)
    a, _, _ = _shade(g, s, debug=False)
    b, _, _ = _shade(g, s, debug=False)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_error_paths(small_gpu, ltc_table):
    import risltc_amd
    torch = _torch()
    ctx = risltc_amd.Context(0)
    g = small_gpu
    out = torch.empty(g["P"] * 4, dtype=torch.float32, device="cuda")
    s = g["scene"].settings()
    with pytest.raises(risltc_amd.RisltcError) as e:
        ctx.shade(g["consts"], s, g["vis"], out)
    assert e.value.code == abi.E_STATE
    ctx.load_scene(g["scene"], *ltc_table)
    for bad, code in ((dict(sample_count=0), abi.E_ARG), (dict(polygon_sampling_technique=3),
                                                                   abi.E_UNSUPPORTED),
                      (dict(polygon_sampling_technique=0), abi.E_UNSUPPORTED),
                      (dict(polygon_sampling_technique=abi.SAMPLE_POLYGON_AREA_TURK, proposal=abi.PROPOSAL_CACHE),
                       abi.E_UNSUPPORTED),
                      (dict(sample_count_light=0), abi.E_ARG), (dict(mis_heuristic=9), abi.E_ARG)):
        with pytest.raises(risltc_amd.RisltcError) as e:
            ctx.shade(g["consts"], g["scene"].settings(**bad), g["vis"], out)
        assert e.value.code == code
    # records with a vertex count outside [3, vmax]
    recs = bytearray(g["scene"].light_records)
    recs[32:36] = (9).to_bytes(4, "little")
    with pytest.raises(risltc_amd.RisltcError) as e:
        ctx.set_lights(bytes(recs), g["scene"].light_count, g["scene"].light_vmax)
    assert e.value.code == abi.E_FORMAT
    ctx.close()


def test_accumulate_running_mean(small_gpu):
    """accum_pass.frag.glsl:45-53: after n frames the buffer is their mean."""
    torch = _torch()
    g = small_gpu
    n = 4096
    frames = torch.rand(5, n * 4, device="cuda")
    acc = torch.zeros(n * 4, device="cuda")
    for i in range(5):
        g["ctx"].accumulate(acc, frames[i], n, i)
    torch.cuda.synchronize()
    assert torch.allclose(acc, frames.mean(0), rtol=1e-5, atol=1e-6)


TECHNIQUE_CASES = [
    dict(polygon_sampling_technique=abi.SAMPLE_POLYGON_AREA_TURK),
    dict(polygon_sampling_technique=abi.SAMPLE_POLYGON_AREA_TURK, light_sampling=abi.LIGHT_UNIFORM),
    dict(polygon_sampling_technique=abi.SAMPLE_POLYGON_AREA_TURK, shadow_rays=0, ris_candidate_count=3),
    dict(polygon_sampling_technique=abi.SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE),
    dict(polygon_sampling_technique=abi.SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE, mis_heuristic=abi.MIS_OPTIMAL,
         ris_candidate_count=5),
    dict(polygon_sampling_technique=abi.SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE, light_sampling=abi.LIGHT_UNIFORM),
]


@pytest.mark.parametrize("case", range(len(TECHNIQUE_CASES)))
def test_baseline_techniques_exact_bitwise(small_gpu, case):
    """The reference's comparison techniques (Turk area sampling and
    projected-solid-angle RIS) in the per-pixel baseline kernel: every
    candidate record, every sample record and the frame bitwise equal to the
    FP32 oracle."""
    g = small_gpu
    kw = dict(exact_math=1, sample_count_light=2, ris_candidate_count=8)
    kw.update(TECHNIQUE_CASES[case])
    print("case", kw, flush=True)
    s = g["scene"].settings(**kw)
    out, rec, cand = _shade(g, s)
    vis = g["vis"].cpu().numpy().view(np.uint32)
    ref, rrec, rcand = g["oracle"].shade(g["consts"], s, vis, records=True, candidates=True)
    assert np.array_equal(cand, rcand)
    ri, oi = rec.view(np.uint32), rrec.view(np.uint32)
    bad = np.argwhere(np.any(ri != oi, axis=-1))
    assert bad.size == 0, ("sample records differ", bad[:5], np.flatnonzero(ri[tuple(bad[0])] != oi[tuple(bad[0])]))
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    shaded = (vis >> 31) == 0
    assert (out[shaded, :3].sum(1) > 0).mean() > 0.3
    st = ri[..., abi.REC_STATE]
    if kw.get("shadow_rays", 1):
        assert (st == 2).sum() > 0 and (st == 3).sum() > 0


@pytest.mark.parametrize("technique", [abi.SAMPLE_POLYGON_AREA_TURK, abi.SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE])
def test_baseline_techniques_fast(small_gpu, technique):
    """Production math: the chosen lights equal the FP32 oracle's except near
    ties, and the frame agrees where they do."""
    g = small_gpu
    s = g["scene"].settings(exact_math=0, sample_count_light=2, ris_candidate_count=16,
                            polygon_sampling_technique=technique)
    out, rec, _ = _shade(g, s)
    vis = g["vis"].cpu().numpy().view(np.uint32)
    ref, rrec, _ = g["oracle"].shade(g["consts"], s, vis, records=True)
    li_g = rec[..., abi.REC_LIGHT].view(np.int32)
    li_o = rrec[..., abi.REC_LIGHT].view(np.int32)
    assert (li_g == li_o).mean() > 0.99
    same = np.all(li_g == li_o, axis=1)
    rel = np.abs(out[same, :3] - ref[same, :3]) / (np.abs(ref[same, :3]) + 1e-3)
    assert np.percentile(rel, 99) < 1e-3, _quantiles(rel)


def test_baseline_techniques_rejected_in_batches(small_gpu):
    import risltc_amd
    g = small_gpu
    torch = _torch()
    out = torch.empty(g["P"] * 4, dtype=torch.float32, device="cuda")
    s = g["scene"].settings(polygon_sampling_technique=abi.SAMPLE_POLYGON_AREA_TURK)
    v = risltc_amd.Context.view(g["consts"], s, g["vis"], out)
    with pytest.raises(risltc_amd.RisltcError) as e:
        g["ctx"].shade_batch([v])
    assert e.value.code == 4
