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


def accumulate(ctx, scene, vis, frames, seed0, learned=None, checkpoints=(), ref=None):
    """Accumulates `frames` frames. learned = None (plain) or a dict with
    floor_log2, infer_every, batch, lr, clusters, seed. Returns (image,
    gpu_ms, curve) where curve maps spp -> (relmse, gpu_ms) at checkpoints."""
    import torch
    P = scene.width * scene.height
    out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
    acc = torch.zeros(P * 4, dtype=torch.float32, device="cuda")
    s = scene.settings(proposal=abi.PROPOSAL_LEARNED if learned else abi.PROPOSAL_UNIFORM,
                       learned_floor_log2=(learned or {}).get("floor_log2", 0))
    if learned:
        ctx.cluster_lights(learned.get("clusters", 256), 20)
        ctx.proposal_init(learned.get("seed", 1))
    gpu = 0.0
    curve = {}
    for f in range(frames):
        c = ctx.frame_constants(noise_seed=seed0 + f, animate=True)
        if learned and f % learned.get("infer_every", 1) == 0:
            ctx.proposal_update(c, vis)
            gpu += ctx.proposal_last_timings()[0]
        ctx.shade(c, s, vis, out)
        t = ctx.last_timings()
        gpu += t[0] + t[1] + t[2]
        ctx.accumulate(acc, out, P, f)
        if learned:
            ctx.proposal_train(c, vis, batch=learned.get("batch", 16384), steps=1, lr=learned.get("lr", 2e-3),
                               seed=seed0 + f)
            gpu += ctx.proposal_last_timings()[1]
        if ref is not None and (f + 1) in checkpoints:
            curve[(f + 1) * scene.spp] = (rel_mse(acc, ref), gpu)
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


def compare(ctx, scene, vis, target=0.01, ref_frames=2048, frames=256, learned=None):
    """The learned-vs-plain experiment on one scene; returns a dict for the
    bench line."""
    learned = learned or dict(floor_log2=5, infer_every=4, batch=16384, lr=2e-3, clusters=256, seed=1)
    ref, ref_ms, _ = accumulate(ctx, scene, vis, ref_frames, 1000000)
    cps = tuple(2 ** k for k in range(4, 20) if 2 ** k <= frames)
    _, _, plain = accumulate(ctx, scene, vis, frames, 0, checkpoints=cps, ref=ref)
    _, _, learn = accumulate(ctx, scene, vis, frames, 0, learned=learned, checkpoints=cps, ref=ref)
    tp, ep = time_to(plain, target)
    tl, el = time_to(learn, target)
    last = max(plain)
    return {
        "target_relmse": target,
        "reference_spp": ref_frames * scene.spp,
        "time_to_target_ms": {"plain": round(tp, 1), "learned": round(tl, 1)},
        "extrapolated": {"plain": ep, "learned": el},
        "learned_over_plain": round(tl / tp, 4),
        "relmse_at_%dspp" % last: {"plain": round(plain[last][0], 5), "learned": round(learn[last][0], 5)},
        "gpu_ms_at_%dspp" % last: {"plain": round(plain[last][1], 1), "learned": round(learn[last][1], 1)},
        "learned_config": learned,
    }
