"""Benchmark of the RIS-LTC shading pass on one MI355X.

A step is one shading pass over one 1920x1080 frame of the city workload
(1024 convex 3-8-gon lights, ~1M triangles, RIS M = 32, 4 samples per
pixel): the shade kernel, the shadow-ray trace kernel and the resolve
kernel, with scene, lights, LTC table, LBVH and visibility buffer already
resident in HBM. value = samples per second of the whole pass.

Prints one JSON line. The CPU baseline is the FP32 oracle (OpenMP, the
host's cores) on a bounded sample of the same frame.

The learned_vs_plain object is the second half of the metric: GPU time for
the plain pass and for the learned pass (MLP + online visibility cache in
the RIS target) to reach relMSE 0.01 against a 65536-spp plain reference,
each pass's Msamples/s and per-kernel milliseconds per frame, ablations,
and the plain pass with the reference's two comparison techniques (Turk
area sampling, projected-solid-angle RIS; risltc_amd/convergence.py). The
learned passes first train over LEARN_PATH_FRAMES frames of a seeded camera
path (GPU time reported apart); the view is held out from that path and
from the runs that chose the learned settings, with held-out noise seeds.

The batched_views object: views of the batched configuration (each its own
city mesh of ~50k-500k triangles and its own light set of 16-8192 polygons,
resolutions from 320x240 to 1920x1080, M in {8, 16, 32}, 2 spp; 16 views by
default, --batch-views 64 for the full configuration), shaded by one call per
view, by one risltc_shade_batch, and by a hipGraph replay of that batch.

The learned_batch object: the learned pass in one batch at the large end
(16 views of 1920x1080 around a ~4M-triangle city with 16384 lights, 16 spp,
per-view proposal tables, the shared visibility cache), against the plain
batch of the same views, and as a graph replay.

The large_learned_vs_plain object (about 3 minutes; --no-large-convergence
skips it): learned vs plain time to relMSE 0.01 on one 1920x1080 view at
the large end (16384 lights, ~4M triangles, 16 spp per frame) against a
65536-spp plain reference, for the 1024-cluster width-128 network (tables
of quads of clusters) and the default 256-cluster network, each pretrained
over a camera path, on a held-out view.

The launch shape is autotuned (risltc_autotune) on the frame before timing.

  python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu] [--no-learned] [--no-batch]
                  [--no-learned-batch] [--no-large-convergence]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s @1080p; GPU time to relMSE 0.01, learned vs plain RIS-LTC"
# frames of the seeded camera path the proposal trains over before the
# learned-vs-plain runs (1080p city; and at the large end)
LEARN_PATH_FRAMES = 200
LARGE_PATH_FRAMES = 100
HBM_PEAK_GBS = 8000.0
# wave64 FP32 VALU instructions: 256 CUs x 4 SIMDs, 4 cycles each, 2.4 GHz
# (non-packed: the build disables packed FP32), in lane-instructions/s
VALU_PEAK_TLANE = 256 * 4 * 16 * 2.4e9 / 1e12
# dense BF16 MFMA peak of the MI355X
MFMA_BF16_PEAK_TF = 2500.0
PMC_FILE = os.path.join(ROOT, "profiles", "shade_city1080_traffic.json")


def algorithmic_bytes(pixels, lights, vmax, ltc_res):
    """Bytes the shading pass must move per frame: per pixel the visibility
    id (4), three vertices' packed position and normal/uv (48), the material
    index (1) and the RGBA32F output (16); the light records once; one
    Fresnel slice of the LTC table (12 bytes per texel)."""
    return pixels * (4 + 48 + 1 + 16) + lights * (48 + 16 * vmax) + ltc_res * ltc_res * 12


def batch_views(ctx, ltc_rgba, ltc_rg, count, seed=7):
    """Views shaped as the batched configuration: each view its own
    procedural city of ~50k-500k triangles and its own light set of 16-8192
    polygons (3-8 vertices, count log-uniform), resolution uniform in
    320x240 .. 1920x1080, M from {8, 16, 32}, 2 spp. Returns the views and
    the buffers they point at."""
    import torch
    import risltc_amd
    from risltc_amd import synth
    views, keep = [], []
    samples = 0
    for sc in synth.batch_view_scenes(count, seed):
        w, h = sc.width, sc.height
        sid = ctx.add_scene(sc)
        lid = ctx.add_light_set(sc.light_records, sc.light_count, sc.light_vmax)
        c = risltc_amd.write_frame_constants(sc.camera, w, h, sc.mesh.dequantization_factor,
                                             sc.mesh.dequantization_summand, ctx.ltc_constants)
        vis = torch.empty(w * h, dtype=torch.int32, device="cuda")
        ctx.render_visibility_scene(sid, c, vis)
        out = torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
        st = sc.settings()
        views.append(ctx.view(c, st, vis, out, light_set=lid, scene=sid))
        keep.append((vis, out))
        samples += w * h * st.sample_count_light
    return views, keep, samples


def bench_batch(ctx, ltc_rgba, ltc_rg, steps, warmup, count):
    """count views of the batched configuration, shaded by one risltc_shade
    call per view, by one risltc_shade_batch, and by a graph replay of that
    batch, each after its own warm-up; wall time per step over `steps`
    steps. The launch shape is autotuned on the batch."""
    import torch
    t0 = time.time()
    views, keep, samples = batch_views(ctx, ltc_rgba, ltc_rg, count)
    setup_s = time.time() - t0
    before = ctx.get_tuning().as_dict()
    tuning = ctx.autotune(views, repeats=3).as_dict()

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3 / steps

    def per_view():
        for i, v in enumerate(views):
            ctx.shade_batch([v])

    ms_views = timed(per_view)
    ms_batch = timed(lambda: ctx.shade_batch(views))
    kernel = ctx.last_timings()
    g = ctx.graph(views)
    try:
        ms_graph = timed(g.launch)
    finally:
        g.close()
    ctx.set_tuning(**before)
    rate = lambda ms: round(samples / (ms * 1e-3) / 1e6, 3)
    res = [[int(v.constants.viewport_size[0]), int(v.constants.viewport_size[1])] for v in views]
    return {"views": count, "samples_per_step": samples, "setup_s": round(setup_s, 1), "tuning": tuning,
            "resolutions": res, "candidates": sorted(set(int(v.settings.ris_candidate_count) for v in views)),
            "batch_kernel_ms": {"shade": round(kernel[0], 3), "trace": round(kernel[1], 3),
                                "resolve": round(kernel[2], 3)},
            "ms_per_step": {"per_view_calls": round(ms_views, 4), "batch": round(ms_batch, 4),
                            "graph": round(ms_graph, 4)},
            "msamples_per_s": {"per_view_calls": rate(ms_views), "batch": rate(ms_batch), "graph": rate(ms_graph)}}


def bench_learned_batch(ltc_rgba, ltc_rg, steps, warmup, views_n=16, grid=540, lights=16384, spp=16):
    """The learned pass in one batch at the large end: views_n views at
    1920x1080 orbiting a procedural city of ~4M triangles with `lights`
    3-8-gon lights, M = 32, `spp` samples per pixel, shaded as one plain batch
    and as one learned batch (MLP source with per-view tables + visibility
    cache, M = 24) and as a graph replay of the learned batch. The MLP is the
    96-128-128-128-1024 network over 1024 k-means clusters (per-pixel tables
    of 2-byte cluster ids), warmed by a few trained single-view frames. Wall
    time per step, learned steps including the refresh of every view's table
    at the learned pass's inference cadence; no convergence run (a
    65536-spp reference of 16 views is out of this bench's time)."""
    import math
    import torch
    import risltc_amd
    from risltc_amd import abi, synth, convergence
    t0 = time.time()
    sc = synth.city_scene(seed=400, light_count=lights, grid=grid, spp=spp, candidates=32)
    ctx = risltc_amd.Context(0)
    ctx.load_scene(sc, ltc_rgba, ltc_rg)
    lk = convergence.DEFAULT_LEARNED
    clusters, outputs = 1024, 1024
    ctx.cluster_lights(clusters, 20)
    ctx.proposal_init(lk["seed"], width=128, outputs=outputs)
    # a per-light cache of 16384 lights x 2^16 cells would take 8 GiB: the
    # cache keeps one slot per cluster here
    per_light = lk["cache_per_light"] and lights <= 4096
    ctx.proposal_cache(lk["cache_log2"], lk["cache_cell"], lk["cache_prior"], per_light)
    centre = np.array([0.0, 40.0, 10.0])
    W, H = sc.width, sc.height
    bufs = []
    for v in range(views_n):
        phi = 2.0 * math.pi * v / views_n
        pos = (centre[0] + 100.0 * math.sin(phi), centre[1] - 100.0 * math.cos(phi), 35.0)
        best = None
        for rz in (math.pi + phi, math.pi - phi):
            cam = synth.make_camera(pos, rz, math.pi / 2 - 0.3)
            c = risltc_amd.write_frame_constants(cam, W, H, sc.mesh.dequantization_factor,
                                                 sc.mesh.dequantization_summand, ctx.ltc_constants)
            P = np.array(c.pixel_to_ray_direction_world_space, dtype=np.float64)
            d = P[:, 0] * (W / 2) + P[:, 1] * (H / 2) + P[:, 2]
            score = d @ (centre - np.array(pos)) / np.linalg.norm(d)
            if best is None or score > best[0]:
                best = (score, cam)
        cam = best[1]
        vis = torch.empty(W * H, dtype=torch.int32, device="cuda")
        out = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
        tab = torch.empty(W * H * abi.proposal_table_bytes(outputs), dtype=torch.uint8, device="cuda")
        bufs.append([cam, vis, out, tab])
        ctx.render_visibility(risltc_amd.write_frame_constants(cam, W, H, sc.mesh.dequantization_factor,
                                                               sc.mesh.dequantization_summand, ctx.ltc_constants),
                              vis)
    plain_s = sc.settings()
    learned_s = sc.settings(proposal=abi.PROPOSAL_LEARNED, ris_candidate_count=lk["candidates"],
                            learned_floor_log2=lk["floor_log2"], skip_training_records=1)
    # warm the MLP and the cache on single-view frames of view 0
    cam0, vis0 = bufs[0][0], bufs[0][1]
    for f in range(4):
        c = risltc_amd.write_frame_constants(cam0, W, H, sc.mesh.dequantization_factor,
                                             sc.mesh.dequantization_summand, ctx.ltc_constants, noise_seed=f,
                                             animate=True)
        ctx.proposal_update(c, vis0)
        ctx.shade(c, sc.settings(proposal=abi.PROPOSAL_LEARNED, ris_candidate_count=lk["candidates"],
                                 learned_floor_log2=lk["floor_log2"]), vis0, bufs[0][2])
        ctx.proposal_train(c, vis0, batch=lk["batch"], steps=1, lr=lk["lr"], seed=f)
    plain, learned, consts = [], [], []
    for cam, vis, out, tab in bufs:
        c = risltc_amd.write_frame_constants(cam, W, H, sc.mesh.dequantization_factor,
                                             sc.mesh.dequantization_summand, ctx.ltc_constants, noise_seed=9,
                                             animate=True)
        ctx.proposal_update(c, vis, table=tab)
        plain.append(ctx.view(c, plain_s, vis, out))
        learned.append(ctx.view(c, learned_s, vis, out, proposal_table=tab))
        consts.append(c)
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    samples = views_n * W * H * spp
    # a learned step refreshes every view's proposal table at the learned
    # pass's cadence (every infer_every steps), then shades the batch
    infer = {"steps": 0, "ms": 0.0}

    def refresh_tables():
        if infer["steps"] % lk["infer_every"] == 0:
            for (cam, vis, out, tab), c in zip(bufs, consts):
                ctx.proposal_update(c, vis, table=tab)
                infer["ms"] += ctx.proposal_last_timings()[0]
        infer["steps"] += 1

    # LDS staging that keeps three work-groups per CU: 1024 spheres for the
    # plain batch; none for the learned batch, whose cache rows take LDS
    tuning = ctx.get_tuning().as_dict()

    def plain_step():
        ctx.set_tuning(**dict(tuning, lds_lights=1024))
        ctx.shade_batch(plain)

    def learned_step():
        ctx.set_tuning(**dict(tuning, lds_lights=lk["lds_lights"]))
        refresh_tables()
        ctx.shade_batch(learned)

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3 / steps

    ms_plain = timed(plain_step)
    k_plain = ctx.last_timings()
    ms_learned = timed(learned_step)
    k_learned = ctx.last_timings()
    infer_ms_per_step = infer["ms"] / max(infer["steps"], 1)
    g = ctx.graph(learned)

    def graph_step():
        refresh_tables()
        g.launch()
    try:
        ms_graph = timed(graph_step)
    finally:
        g.close()
    ctx.close()
    rate = lambda ms: round(samples / (ms * 1e-3) / 1e6, 2)
    km = lambda k: {"shade": round(k[0], 2), "trace": round(k[1], 2), "resolve": round(k[2], 2)}
    return {"views": views_n, "resolution": [W, H], "triangles": int(sc.mesh.triangle_count), "lights": lights,
            "clusters": clusters, "mlp": "96-128-128-128-%d" % outputs, "spp": spp,
            "cache_slots": "light" if per_light else "cluster",
            "candidates": {"plain": 32, "learned": lk["candidates"]}, "samples_per_step": samples,
            "setup_s": round(setup_s, 1),
            "ms_per_step": {"plain_batch": round(ms_plain, 2), "learned_batch": round(ms_learned, 2),
                            "learned_graph": round(ms_graph, 2)},
            "learned_steps_include": "proposal inference of all %d views every %d steps (%.2f ms per step "
                                     "on average) + the batch" % (views_n, lk["infer_every"], infer_ms_per_step),
            "inference_ms_per_step": round(infer_ms_per_step, 3),
            "batch_kernel_ms": {"plain": km(k_plain), "learned": km(k_learned)},
            "msamples_per_s": {"plain_batch": rate(ms_plain), "learned_batch": rate(ms_learned),
                               "learned_graph": rate(ms_graph)}}


def bench_large_convergence(ltc_rgba, ltc_rg, frames=128, ref_frames=4096):
    """Learned vs plain time to relMSE 0.01 at the large end, on one 1920x1080
    view of the ~4M-triangle city with 16384 lights and 16 spp per frame
    (M = 32 plain, 24 learned), against a 65536-spp plain reference of that
    view: the 96-128-128-128-1024 network over 1024 clusters (the large-end
    configuration's network; its tables sample quads of clusters) and the
    default 96-64-64-64-256 network over 256 clusters, each with a
    visibility cache of one slot per cluster, each pretrained over
    LARGE_PATH_FRAMES frames of a seeded camera path. The view is held out
    from the path. About 3.5 minutes of GPU time, most of it the
    reference."""
    import torch
    import risltc_amd
    from risltc_amd import synth, convergence as cv
    t0 = time.time()
    sc = synth.city_scene(seed=400, light_count=16384, grid=540, spp=16, candidates=32)
    ctx = risltc_amd.Context(0)
    try:
        ctx.load_scene(sc, ltc_rgba, ltc_rg)
        P = sc.width * sc.height
        cam = cv.held_out_camera(sc)
        path = cv.camera_path(sc, LARGE_PATH_FRAMES)
        vis = torch.empty(P, dtype=torch.int32, device="cuda")
        ctx.render_visibility(ctx.frame_constants(camera=cam), vis)
        ref = cv.reference(ctx, sc, vis, ref_frames, chunk=512, camera=cam,
                           log=lambda m: print(m, file=sys.stderr, flush=True))
        cps = (4, 8, 16, 32, 64, 128)

        def run(lk, use_path=True):
            _, _, curve = cv.accumulate(ctx, sc, vis, frames, 500000, learned=lk, checkpoints=cps, ref=ref,
                                        camera=cam, path=path if (lk and use_path) else None)
            t, ex = cv.time_to(curve, 0.01)
            km = cv._frames.kernel_ms
            r = {"time_to_target_ms": round(t, 1), "extrapolated": ex,
                 "curve": {str(k): [round(v[0], 6), round(v[1], 1)] for k, v in sorted(curve.items())},
                 "ms_per_frame": {"shade": round(km[0], 3), "trace": round(km[1], 3), "resolve": round(km[2], 3),
                                  "inference": round(km[3], 3), "training": round(km[4], 3)}}
            if lk and use_path:
                r["pretrain_frames"] = len(path)
                r["pretrain_ms"] = round(cv.accumulate.pretrain_ms, 1)
            return r

        plain = run(None)
        base = dict(cv.DEFAULT_LEARNED, cache_per_light=False)
        out = {"view": [sc.width, sc.height], "camera": "held out (convergence.held_out_camera)",
               "triangles": int(sc.mesh.triangle_count), "lights": sc.light_count,
               "spp_per_frame": sc.spp, "reference_spp": ref_frames * sc.spp, "target_relmse": 0.01,
               "plain": plain}
        k1024 = dict(base, clusters=1024, width=128, outputs=1024)
        for name, lk, use_path in (("learned_1024_clusters_w128", k1024, True),
                                   ("learned_256_clusters_w64", dict(base, clusters=256, width=64, outputs=256), True),
                                   ("learned_1024_clusters_w128_online_only", k1024, False)):
            r = run(lk, use_path)
            r["over_plain"] = round(r["time_to_target_ms"] / plain["time_to_target_ms"], 4)
            out[name] = r
            print("large end %s: %.1f ms vs plain %.1f ms" % (name, r["time_to_target_ms"],
                                                              plain["time_to_target_ms"]), file=sys.stderr, flush=True)
        out["wall_s"] = round(time.time() - t0, 1)
        return out
    finally:
        ctx.close()


def kernel_report(kernel_ms, pixels, samples, rays, shade_alg, learned):
    """Every kernel of the plain pass against the HBM roofline on its
    algorithmic bytes (live times), with the measured HBM bytes and VALU
    issue of the counter profile beside it; and the MLP inference against
    the BF16 MFMA peak. Algorithmic bytes per launch: shade as the roofline
    object; trace 40 B per shadow ray (the ray and its slot read, the hit
    word written); resolve 4 B visibility + 16 B output per pixel, the
    record of every pixel-sample (32 B: the plain pass writes the compact
    record, no light, key or occluded estimates) and the hit word of every
    ray."""
    pmc = json.load(open(PMC_FILE)) if os.path.exists(PMC_FILE) else {}
    per = dict(pmc.get("kernels", {}))
    if "bytes_per_launch" in pmc:
        per["shade_kernel"] = {k: pmc[k] for k in ("bytes_per_launch", "valu_insts_per_launch",
                                                    "valu_thread_cycles_per_launch")}
    alg = {"shade_kernel": shade_alg, "trace_kernel": 40 * rays,
           "resolve_kernel": 20 * pixels + 32 * samples + 4 * rays}
    out = {"shadow_rays": rays}
    for i, k in enumerate(("shade_kernel", "trace_kernel", "resolve_kernel")):
        ms = float(kernel_ms[i])
        r = {"ms": round(ms, 4), "algorithmic_bytes": alg[k],
             "hbm_frac": round(alg[k] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 6)}
        c = per.get(k)
        if c:
            r["measured_bytes"] = c["bytes_per_launch"]
            r["measured_hbm_frac"] = round(c["bytes_per_launch"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
            r["valu_frac"] = round(c["valu_insts_per_launch"] * 64 / (ms * 1e-3) / 1e12 / VALU_PEAK_TLANE, 4)
            r["active_lanes"] = round(c["valu_thread_cycles_per_launch"] / c["valu_insts_per_launch"] / 64, 3)
        out[k] = r
    if learned:
        lk = learned["learned_config"]
        per_call = learned["learned"]["ms_per_frame"]["inference"] * lk["infer_every"]
        h, o = 64, 256
        flops = 2.0 * pixels * (96 * h + 2 * h * h + h * o)
        if per_call > 0:
            tf = flops / (per_call * 1e-3) / 1e12
            out["proposal_kernel"] = {"ms_per_call": round(per_call, 4), "flop_per_call": flops,
                                      "tflops": round(tf, 1), "mfma_bf16_frac": round(tf / MFMA_BF16_PEAK_TF, 4),
                                      "network": "96-64-64-64-256, one 1080p view per call"}
    out["peaks"] = {"hbm_gbs": HBM_PEAK_GBS, "valu_t_lane_inst_s": round(VALU_PEAK_TLANE, 3),
                    "mfma_bf16_tflops": MFMA_BF16_PEAK_TF}
    out["counters_source"] = "profiles/city1080_counters.md"
    return out


def bench_lds_staging(ltc_rgba, ltc_rg, repeats=8):
    """Shade-kernel time (median of `repeats` frames, HIP events) of the
    1080p city at its held-out view for each number of light spheres staged
    in LDS: the plain pass (M = 32; all 1024 spheres take 16 KiB) and the
    learned pass of the learned configuration (M = 24, tables from a few
    trained frames, per-light cache rows and the cluster tables in LDS too).
    With the pixel area (28 KiB) a work-group takes 28 KiB + 16 B per sphere
    (plain) or that plus 23 KiB (learned); three work-groups fit the CU's
    160 KiB up to 53 KiB each."""
    import torch
    import risltc_amd
    from risltc_amd import abi, synth, convergence as cv
    sc = synth.city_scene()
    ctx = risltc_amd.Context(0)
    try:
        ctx.load_scene(sc, ltc_rgba, ltc_rg)
        cam = cv.held_out_camera(sc)
        c = ctx.frame_constants(camera=cam)
        P = sc.width * sc.height
        vis = torch.empty(P, dtype=torch.int32, device="cuda")
        ctx.render_visibility(c, vis)
        out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
        lk = cv.DEFAULT_LEARNED
        ctx.cluster_lights(lk["clusters"], 20)
        ctx.proposal_init(lk["seed"], width=64, outputs=256)
        ctx.proposal_cache(lk["cache_log2"], lk["cache_cell"], lk["cache_prior"], lk["cache_per_light"])
        learned = sc.settings(proposal=abi.PROPOSAL_LEARNED, ris_candidate_count=lk["candidates"],
                              learned_floor_log2=lk["floor_log2"])
        for f in range(6):
            cf = ctx.frame_constants(noise_seed=f, animate=True, camera=cam)
            ctx.proposal_update(cf, vis)
            ctx.shade(cf, learned, vis, out)
            ctx.proposal_train(cf, vis, batch=lk["batch"], steps=1, lr=lk["lr"], seed=f)
        learned.skip_training_records = 1
        tuning = ctx.get_tuning().as_dict()
        res = {"view": "held-out camera of the 1920x1080 city (1024 lights)", "plain_shade_ms": {},
               "learned_shade_ms": {}}
        for lds in (0, 256, 512, 1024):
            ctx.set_tuning(**dict(tuning, lds_lights=lds))
            for key, s in (("plain_shade_ms", sc.settings()), ("learned_shade_ms", learned)):
                ts = []
                for _ in range(repeats + 2):
                    ctx.shade(c, s, vis, out)
                    torch.cuda.synchronize()
                    ts.append(ctx.last_timings()[0])
                res[key][str(lds)] = round(float(np.median(ts[2:])), 4)
        ctx.set_tuning(**tuning)
        return res
    finally:
        ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="city")
    ap.add_argument("--exact", action="store_true", help="exact-math kernels (parity build)")
    ap.add_argument("--cpu-pixels", type=int, default=0,
                    help="pixels in the CPU-baseline sample (0 = the whole frame)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-vs-plain time-to-relMSE experiment")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-view measurement")
    ap.add_argument("--no-learned-batch", action="store_true",
                    help="skip the learned batch at the large end (16 x 1080p views, ~4M triangles, 16384 lights)")
    ap.add_argument("--no-large-convergence", action="store_true",
                    help="skip learned vs plain time to relMSE 0.01 on one view at the large end "
                         "(16384 lights, ~4M triangles; about 3 minutes)")
    ap.add_argument("--batch-views", type=int, default=16,
                    help="views in the batched measurement (the batched configuration has 64)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--learned-frames", type=int, default=0,
                    help="only run N frames of the learned pass (for profiling its kernels)")
    ap.add_argument("--lds-ab", action="store_true",
                    help="only time the shade kernel for each number of light spheres staged in LDS")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (known HBM bytes) for counter scaling")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.lds_ab:
        from risltc_amd import ltcfit
        rgba, rg, _ = ltcfit.table(51)
        print(json.dumps({"lds_staging": bench_lds_staging(rgba, rg)}))
        return
    if args.copy_calibration:
        # an elementwise kernel (not a blit) that reads 1 GiB and writes 1 GiB
        n = 1 << 28
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 3}))
        return

    import risltc_amd
    from risltc_amd import ltcfit, synth

    t0 = time.time()
    scene = synth.WORKLOADS[args.workload]["fn"]()
    rgba, rg, res = ltcfit.table(51)
    ctx = risltc_amd.Context(0)
    ctx.load_scene(scene, rgba, rg)
    consts = ctx.frame_constants()
    P = scene.width * scene.height
    spp = scene.spp
    vis = torch.empty(P, dtype=torch.int32, device="cuda")
    ctx.render_visibility(consts, vis)
    out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
    settings = scene.settings(exact_math=1 if args.exact else 0)
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    print("setup %.1f s: %d triangles, %d lights, %dx%d, spp %d, M %d" %
          (setup_s, scene.mesh.triangle_count, scene.light_count, scene.width, scene.height, spp,
           scene.candidates), file=sys.stderr, flush=True)

    if args.learned_frames > 0:
        # the learned pass's kernels as the experiment runs them, without the
        # reference and the plain pass
        from risltc_amd import convergence
        t1 = time.time()
        _, gpu_ms, _ = convergence.accumulate(ctx, scene, vis, args.learned_frames, 0,
                                              learned=dict(convergence.DEFAULT_LEARNED))
        print(json.dumps({"learned_frames": args.learned_frames, "gpu_ms": round(gpu_ms, 2),
                          "ms_per_frame": [round(x, 4) for x in convergence._frames.kernel_ms],
                          "wall_s": round(time.time() - t1, 1)}), flush=True)
        return

    tuning = None
    if not args.no_autotune:
        t1 = time.time()
        tuning = ctx.autotune([ctx.view(consts, settings, vis, out)], repeats=5).as_dict()
        print("autotune %.1f s: %s" % (time.time() - t1, tuning), file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        ctx.shade(consts, settings, vis, out)
    torch.cuda.synchronize()
    kernel_ms = np.zeros(4)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ctx.shade(consts, settings, vis, out)
        kernel_ms += np.array(ctx.last_timings())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1e3 / args.steps
    kernel_ms /= args.steps
    rays = ctx.last_ray_count()
    samples = P * spp
    value = samples / (ms_per_step * 1e-3) / 1e6
    shade_ms = kernel_ms[0]
    alg = algorithmic_bytes(P, scene.light_count, scene.light_vmax, res)
    achieved = alg / (shade_ms * 1e-3) / 1e9
    traffic = None
    valu = None
    if os.path.exists(PMC_FILE) and args.workload == "city" and not args.exact:
        pmc = json.load(open(PMC_FILE))
        traffic = pmc.get("bytes_per_launch")
        insts = pmc.get("valu_insts_per_launch")
        if insts:
            rate = insts * 64 / (shade_ms * 1e-3) / 1e12
            valu = {"achieved": round(rate, 3), "peak": round(VALU_PEAK_TLANE, 3), "unit": "T lane-inst/s",
                    "frac": round(rate / VALU_PEAK_TLANE, 4),
                    "active_lanes": round(pmc["valu_thread_cycles_per_launch"] / insts / 64, 3),
                    "insts_per_launch": insts, "source": "profiles/city1080_counters.md"}
    vis_h = vis.cpu().numpy().view(np.uint32)
    frame = out.cpu().numpy().reshape(-1, 4)
    print("kernels (ms): shade %.3f trace %.3f resolve %.3f; frame mean %s" %
          (kernel_ms[0], kernel_ms[1], kernel_ms[2], frame[:, :3].mean(0)), file=sys.stderr, flush=True)

    batch = None
    if not args.no_batch and args.workload == "city":
        batch = bench_batch(ctx, rgba, rg, args.steps, args.warmup, args.batch_views)
        print("batch: %s" % batch, file=sys.stderr, flush=True)

    learned_batch = None
    if not args.no_learned_batch and args.workload == "city" and not args.exact:
        learned_batch = bench_learned_batch(rgba, rg, max(1, min(args.steps, 3)), 1)
        print("learned batch: %s" % learned_batch, file=sys.stderr, flush=True)

    large = None
    if not args.no_large_convergence and args.workload == "city" and not args.exact:
        large = bench_large_convergence(rgba, rg)

    learned = None
    if not args.no_learned and args.workload == "city" and not args.exact:
        from risltc_amd import convergence
        t1 = time.time()
        # evaluated on a view held out from the training path and from the
        # runs that chose the learned settings, with held-out noise seeds
        cam = convergence.held_out_camera(scene)
        vis_eval = torch.empty(P, dtype=torch.int32, device="cuda")
        ctx.render_visibility(ctx.frame_constants(camera=cam), vis_eval)
        path = convergence.camera_path(scene, LEARN_PATH_FRAMES)
        learned = convergence.compare(ctx, scene, vis_eval, camera=cam, path=path, seed0=500000,
                                      log=lambda m: print(m, file=sys.stderr, flush=True))
        learned["view"] = "held-out camera (convergence.held_out_camera), noise seeds from 500000"
        learned["wall_s"] = round(time.time() - t1, 1)
        # each pass's shade kernel against the same algorithmic bytes
        for k in ("plain", "learned"):
            ms = learned[k]["ms_per_frame"]["shade"]
            learned[k]["shade_roofline_frac"] = round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 6)

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        # the host's CPU share: OMP_NUM_THREADS where the host sets it (a GPU
        # box gives one GPU 16 threads), else every logical CPU
        threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or (os.cpu_count() or 1)
        oracle = orc.Oracle(scene, rgba, rg, timing=True)
        if 0 < args.cpu_pixels < P:
            rng = np.random.default_rng(0)
            pix = np.sort(rng.choice(P, size=args.cpu_pixels, replace=False)).astype(np.uint32)
            what = "%d random pixels of the same %dx%d frame" % (pix.size, scene.width, scene.height)
        else:
            pix = None
            what = "the whole %dx%d frame" % (scene.width, scene.height)
        count = P if pix is None else pix.size
        t1 = time.perf_counter()
        oracle.shade_frame(consts, scene.settings(), vis_h, pixels=pix, threads=threads)
        cpu_s = time.perf_counter() - t1
        model = None
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": round(count * spp / cpu_s / 1e6, 4), "unit": "Msamples/s", "cores": threads,
               "kind": "port",
               "sample": "%s, %d spp, M=%d, shadow rays on the oracle's SAH tree; FP32 oracle built -O3 "
                         "-march=x86-64-v3 -ffp-contract=off, OpenMP over pixels, %.1f s"
                         % (what, spp, scene.candidates, cpu_s),
               "host_logical_cpus": os.cpu_count(), "omp_max_threads": int(oracle.L.orc_omp_max_threads()),
               "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}

    # The HBM roofline of the dominant kernel on algorithmic bytes, as the
    # record asks; beside it what the counters say limits the kernel: its
    # VALU issue rate against the CU peak (the binding resource) and its
    # measured HBM bytes per launch over its time.
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                "kernel": "shade_kernel", "algorithmic_bytes": alg}
    if traffic:
        measured = traffic / (shade_ms * 1e-3) / 1e9
        roofline["measured_traffic_gbs"] = round(measured, 2)
        roofline["measured_traffic_frac"] = round(measured / HBM_PEAK_GBS, 4)
    if valu:
        roofline["limiter"] = "valu" if valu["frac"] >= roofline.get("measured_traffic_frac", 0.0) else "hbm"
        roofline["limiter_frac"] = max(valu["frac"], roofline.get("measured_traffic_frac", 0.0))
    kernels = kernel_report(kernel_ms, P, samples, rays, alg, learned)
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural city, project-fitted LTC table)",
        "config": {"workload": "plain RIS-LTC, %dx%d, %d lights (3-8 vertices), ~%.2fM triangles, M=%d, %d spp"
                               % (scene.width, scene.height, scene.light_count, scene.mesh.triangle_count / 1e6,
                                  scene.candidates, spp),
                   "samples_per_step": samples, "exact_math": bool(args.exact), "tuning": tuning,
                   "kernel_ms": {"shade": round(kernel_ms[0], 4), "trace": round(kernel_ms[1], 4),
                                 "resolve": round(kernel_ms[2], 4)}},
        "roofline": roofline,
        "valu": valu,
        "kernels": kernels,
        "batched_views": batch,
        "learned_batch": learned_batch,
        "learned_vs_plain": learned,
        "large_learned_vs_plain": large,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
