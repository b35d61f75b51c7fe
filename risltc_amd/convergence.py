"""GPU time to a relMSE target: the plain RIS-LTC pass against the pass with
the learned target, both accumulating frames of the same scene and camera.

relMSE = mean over pixels and channels of (x - r)^2 / (r^2 + eps) with
eps = 1e-2 * mean(r)^2 (scale-free), r a high-spp plain reference. GPU time
is the sum of the HIP-event times of every kernel the method runs per frame:
shade, trace and resolve, plus proposal inference and training for the
learned method. Frames use animated noise seeds (wang_random_number of the
frame index, src/math_utilities.h:50-57), as the reference's accumulation
does (accum_pass.frag.glsl:45-53).
"""
import math

import numpy as np

from . import abi


def rel_mse(x, ref):
    """relMSE of RGBA32F frames (torch, device) over the rgb channels."""
    xs = x.view(-1, 4)[:, :3]
    r = ref.view(-1, 4)[:, :3]
    eps = 1e-2 * float(r.mean()) ** 2
    return float((((xs - r) ** 2) / (r * r + eps)).mean())


def camera_path(scene, frames, seed=11):
    """A seeded camera path for training the proposal: a smooth random walk
    of position (within 25 m of the scene camera's, at least 10 m above the
    ground) and heading (within 0.5 rad of the scene camera's, pitch within
    0.2 rad) over `frames` poses."""
    from . import synth
    rng = np.random.default_rng(seed)
    c0 = scene.camera
    p0 = np.array(c0.position_world_space, np.float64)
    pos, yaw, pitch = p0.copy(), 0.0, 0.0
    vel = np.zeros(3)
    cams = []
    for _ in range(frames):
        vel = 0.8 * vel + rng.normal(0.0, 0.6, 3) * np.array([1.0, 1.0, 0.4])
        pos = pos + vel
        d = pos - p0
        pos = p0 + d * min(1.0, 25.0 / max(np.linalg.norm(d), 1e-9))
        pos[2] = max(pos[2], 10.0)
        yaw = float(np.clip(yaw + rng.normal(0.0, 0.05), -0.5, 0.5))
        pitch = float(np.clip(pitch + rng.normal(0.0, 0.03), -0.2, 0.2))
        cams.append(synth.make_camera(pos, c0.rotation_z + yaw, c0.rotation_x + pitch, fov=c0.vertical_fov))
    return cams


def held_out_camera(scene):
    """The evaluation view of the learned experiment: 18 m right of, 6 m
    below and 0.3 rad turned from the scene camera, a pose the training path
    and the tuning runs did not use."""
    from . import synth
    c0 = scene.camera
    p = np.array(c0.position_world_space, np.float64)
    return synth.make_camera(p + np.array([18.0, 4.0, -6.0]), c0.rotation_z - 0.3, c0.rotation_x + 0.05,
                             fov=c0.vertical_fov)


def pretrain(ctx, scene, cameras, learned, seed0=700000):
    """Trains the proposal over the frames of a camera path before an
    accumulation: each frame renders the path camera's visibility, runs the
    plain pass's uniform source with the visibility cache in the target
    (RISLTC_PROPOSAL_CACHE; the cache counts the frame's shadow rays) and
    takes one Adam step on the frame's selected-light clusters. Returns the
    GPU ms of the shading and training kernels. The context must hold the
    learned pass's clusters, network and cache (accumulate's set-up)."""
    import torch
    P = scene.width * scene.height
    vis = torch.empty(P, dtype=torch.int32, device="cuda")
    out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
    mode = abi.PROPOSAL_CACHE if learned.get("cache_log2") else abi.PROPOSAL_UNIFORM
    s = scene.settings(proposal=mode, learned_floor_log2=learned.get("floor_log2", 0),
                       ris_candidate_count=learned.get("pretrain_candidates", scene.candidates))
    ms = 0.0
    for f, cam in enumerate(cameras):
        c = ctx.frame_constants(noise_seed=seed0 + f, animate=True, camera=cam)
        ctx.render_visibility(c, vis)
        ctx.shade(c, s, vis, out)
        ms += sum(ctx.last_timings()[:3])
        ctx.proposal_train(c, vis, batch=learned.get("batch", 16384), steps=1, lr=learned.get("lr", 2e-3),
                           seed=seed0 + f)
        ms += ctx.proposal_last_timings()[1]
    return ms


def accumulate(ctx, scene, vis, frames, seed0, learned=None, checkpoints=(), ref=None, candidates=None,
               technique=abi.SAMPLE_POLYGON_LTC_CP, camera=None, path=None):
    """Accumulates `frames` frames of camera (default: the scene camera)
    whose visibility buffer is vis. learned = None (plain) or a dict with
    floor_log2, infer_every, batch, lr, clusters, seed, and optionally
    cache_log2, cache_cell, cache_prior (the visibility cache), candidates
    (M), and warmup_frames / warmup_candidates (the first frames use the
    uniform source with the cache while the MLP trains on them). path: a
    list of cameras to pretrain the proposal over first (pretrain(); its GPU
    ms go to accumulate.pretrain_ms, not to the curve). technique selects
    the polygon sampling technique of a plain pass.
    Returns (image, gpu_ms, curve) where curve maps spp -> (relmse, gpu_ms)
    at checkpoints."""
    import torch
    P = scene.width * scene.height
    out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
    acc = torch.zeros(P * 4, dtype=torch.float32, device="cuda")
    lk = learned or {}
    m = lk.get("candidates", candidates)
    s = scene.settings(proposal=lk.get("proposal", abi.PROPOSAL_LEARNED) if learned else abi.PROPOSAL_UNIFORM,
                       learned_floor_log2=lk.get("floor_log2", 0), polygon_sampling_technique=technique,
                       **({"ris_candidate_count": m} if m else {}))
    ctx.proposal_cache(0)
    tuning = ctx.get_tuning().as_dict()
    if lk.get("lds_lights") is not None:
        ctx.set_tuning(**dict(tuning, lds_lights=lk["lds_lights"]))
    if learned:
        ctx.cluster_lights(learned.get("clusters", 256), 20)
        ctx.proposal_init(learned.get("seed", 1), width=learned.get("width", 64),
                          outputs=learned.get("outputs", 256))
        if learned.get("cache_log2"):
            ctx.proposal_cache(learned["cache_log2"], learned.get("cache_cell", 0.5), learned.get("cache_prior", 2.0),
                               learned.get("cache_per_light", False))
    accumulate.pretrain_ms = 0.0
    try:
        if learned and path:
            accumulate.pretrain_ms = pretrain(ctx, scene, path, learned)
        return _frames(ctx, scene, vis, frames, seed0, learned, checkpoints, ref, s, out, acc, P, camera)
    finally:
        ctx.proposal_cache(0)
        ctx.set_tuning(**tuning)
        if learned and (learned.get("width", 64), learned.get("outputs", 256)) != (64, 256):
            ctx.proposal_set_network(64, 256)


def _frames(ctx, scene, vis, frames, seed0, learned, checkpoints, ref, s, out, acc, P, camera=None):
    gpu = 0.0
    curve = {}
    kernel_ms = [0.0, 0.0, 0.0, 0.0, 0.0]  # shade, trace, resolve, inference, training
    mlp = bool(learned) and learned.get("proposal", abi.PROPOSAL_LEARNED) == abi.PROPOSAL_LEARNED
    # warm-up: the first frames draw candidates from the uniform source (the
    # cache, if any, still weights the target) while the MLP trains on them
    warm = learned.get("warmup_frames", 0) if mlp else 0
    m_learned = s.ris_candidate_count
    for f in range(frames):
        c = ctx.frame_constants(noise_seed=seed0 + f, animate=True, camera=camera)
        if mlp:
            cache_mode = abi.PROPOSAL_CACHE if learned.get("cache_log2") else abi.PROPOSAL_UNIFORM
            s.proposal = cache_mode if f < warm else abi.PROPOSAL_LEARNED
            s.ris_candidate_count = learned.get("warmup_candidates", m_learned) if f < warm else m_learned
        if mlp and f >= warm and ((f - warm) % learned.get("infer_every", 1) == 0):
            ctx.proposal_update(c, vis)
            kernel_ms[3] += ctx.proposal_last_timings()[0]
        train = mlp and f % learned.get("train_every", 1) == 0
        s.skip_training_records = 0 if train else 1
        ctx.shade(c, s, vis, out)
        t = ctx.last_timings()
        for k in range(3):
            kernel_ms[k] += t[k]
        ctx.accumulate(acc, out, P, f)
        if train:
            ctx.proposal_train(c, vis, batch=learned.get("batch", 16384), steps=1, lr=learned.get("lr", 2e-3),
                               seed=seed0 + f)
            kernel_ms[4] += ctx.proposal_last_timings()[1]
        gpu = sum(kernel_ms)
        if ref is not None and (f + 1) in checkpoints:
            curve[(f + 1) * scene.spp] = (rel_mse(acc, ref), gpu)
    _frames.kernel_ms = [k / max(frames, 1) for k in kernel_ms]
    return acc, gpu, curve


def time_to(curve, target):
    """GPU ms at which relMSE reaches target: log-log interpolation between
    checkpoints, or the 1/spp law from the last checkpoint (extrapolated)."""
    pts = sorted(curve.items())
    prev = None
    for spp, (rel, ms) in pts:
        if rel <= target:
            if prev is None:
                return ms, False
            (s0, (r0, m0)) = prev
            a = (math.log(target) - math.log(r0)) / (math.log(rel) - math.log(r0))
            return math.exp(math.log(m0) + a * (math.log(ms) - math.log(m0))), False
        prev = (spp, (rel, ms))
    spp, (rel, ms) = pts[-1]
    return ms * rel / target, True


# The learned configuration the experiment uses: learned source (MLP
# inference every 2 frames, one Adam step of 8192 samples per frame, learning
# rate 0.02), target floor 1/128 with a per-light visibility cache of 2^16
# one-metre cells, M = 24 candidates, no light spheres staged in LDS: the
# learned pass stages its cluster tables and cache rows there, and with
# spheres too a CU holds two work-groups instead of three (the spheres are
# read from L2 at almost no cost).
DEFAULT_LEARNED = dict(floor_log2=7, infer_every=2, train_every=1, batch=8192, lr=2e-2, clusters=256, seed=1,
                       cache_log2=16, cache_cell=1.0, cache_prior=1.0, cache_per_light=True, candidates=24,
                       lds_lights=0)


def reference(ctx, scene, vis, frames, chunk=2048, seed0=1000000, log=None, camera=None):
    """Plain-pass reference: the mean of frames frames (chunks of `chunk`
    accumulated on the device, their means averaged). log(msg) is called
    after each chunk."""
    import torch
    ref = None
    n = 0
    while n < frames:
        k = min(chunk, frames - n)
        img, _, _ = accumulate(ctx, scene, vis, k, seed0 + n, camera=camera)
        ref = img * (k / frames) if ref is None else ref + img * (k / frames)
        n += k
        torch.cuda.synchronize()
        if log:
            log("reference: %d of %d frames" % (n, frames))
    return ref


def compare(ctx, scene, vis, target=0.01, ref_frames=16384, plain_frames=320, frames=256, learned=None,
            ablations=True, baselines=True, log=None, camera=None, path=None, seed0=0):
    """The learned-vs-plain experiment on one scene; returns a dict for the
    bench line. The plain pass is the scene's own configuration (M and spp);
    the learned pass uses `learned` (DEFAULT_LEARNED by default). Ablations:
    the learned pass without the visibility cache (MLP only), with the MLP
    left untrained (learning rate 0), and trained in the configuration's
    stated regime (Adam, learning rate 1e-3, batches of 65,536). Baselines: the plain pass with the
    reference's comparison techniques (Turk area sampling, projected solid
    angle RIS); their time to the target is null when the curve does not
    reach it (Turk's estimate covers only each polygon's first triangle, so
    its error has a floor wherever lights have more than three vertices).

    camera (vis its visibility buffer; default the scene camera) is the view
    accumulated, from noise seeds seed0 + frame. path: cameras the learned
    passes pretrain over first (the pretraining's GPU time is reported
    apart, and in the ratio that includes it); the `online_only` ablation
    learns on the view alone."""
    learned = dict(DEFAULT_LEARNED, **(learned or {}))
    ref = reference(ctx, scene, vis, ref_frames, log=log, camera=camera)
    cps = tuple(k for k in (16, 32, 64, 128, 256, 320, 512) if k <= max(frames, plain_frames))
    spf = scene.width * scene.height * scene.spp

    def run(lk, n, technique=abi.SAMPLE_POLYGON_LTC_CP, use_path=True):
        _, _, curve = accumulate(ctx, scene, vis, n, seed0, learned=lk, checkpoints=cps, ref=ref,
                                 technique=technique, camera=camera, path=path if (lk and use_path) else None)
        km = list(_frames.kernel_ms)
        t, ex = time_to(curve, target)
        r = {"time_to_target_ms": round(t, 1), "extrapolated": ex,
             "curve": {str(k): [round(v[0], 6), round(v[1], 1)] for k, v in sorted(curve.items())},
             "ms_per_frame": {"shade": round(km[0], 4), "trace": round(km[1], 4), "resolve": round(km[2], 4),
                              "inference": round(km[3], 4), "training": round(km[4], 4)},
             "msamples_per_s": round(spf / (sum(km) * 1e-3) / 1e6, 2)}
        if lk and use_path and path:
            r["pretrain_frames"] = len(path)
            r["pretrain_ms"] = round(accumulate.pretrain_ms, 1)
        return r

    plain = run(None, plain_frames)
    learn = run(learned, frames)
    if log:
        log("plain %.1f ms, learned %.1f ms to relMSE %g" % (plain["time_to_target_ms"], learn["time_to_target_ms"],
                                                             target))
    out = {
        "target_relmse": target,
        "reference_spp": ref_frames * scene.spp,
        "learned_over_plain": round(learn["time_to_target_ms"] / plain["time_to_target_ms"], 4),
        "plain": plain,
        "learned": learn,
        "learned_config": learned,
    }
    if path:
        out["learned_over_plain_with_pretraining"] = round(
            (learn["time_to_target_ms"] + learn["pretrain_ms"]) / plain["time_to_target_ms"], 4)
    if ablations:
        out["ablations"] = {}
        runs = [("source_only", dict(learned, cache_log2=0), True),
                ("cache_only_uniform_source", dict(learned, proposal=abi.PROPOSAL_CACHE, candidates=48), True),
                ("untrained_mlp", dict(learned, lr=0.0), True),
                # the training regime the learned configuration states: Adam
                # at learning rate 1e-3 on batches of 65,536 pixel-samples
                ("adam_lr1e-3_batch65536", dict(learned, lr=1e-3, batch=65536), True)]
        if path:
            runs.append(("online_only", learned, False))
        for name, lk, use_path in runs:
            r = run(lk, frames, use_path=use_path)
            out["ablations"][name] = {"time_to_target_ms": r["time_to_target_ms"],
                                      "over_plain": round(r["time_to_target_ms"] / plain["time_to_target_ms"], 4),
                                      "relmse_at_1024spp": r["curve"].get("1024", [None])[0],
                                      "ms_per_frame": r["ms_per_frame"]}
    if baselines:
        out["baselines"] = {}
        for name, tech in (("area_turk", abi.SAMPLE_POLYGON_AREA_TURK),
                           ("projected_solid_angle", abi.SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE)):
            r = run(None, plain_frames, technique=tech)
            t = None if r["extrapolated"] else r["time_to_target_ms"]
            out["baselines"][name] = {"time_to_target_ms": t,
                                      "over_plain": None if t is None else round(t / plain["time_to_target_ms"], 4),
                                      "curve": r["curve"], "ms_per_frame": r["ms_per_frame"],
                                      "msamples_per_s": r["msamples_per_s"]}
            if log:
                log("baseline %s: %s ms to relMSE %g" % (name, t, target))
    return out
