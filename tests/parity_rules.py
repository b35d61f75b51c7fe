"""Helpers of the GPU parity tests that apply the parity rules of the
production (fast-math) build against the FP64 oracle:

  * the RIS-selected light must equal the FP64 oracle's wherever no
    reservoir decision of the sample was within a stated distance of a tie
    (reservoir.glsl:29-40: candidate i is taken when r_i < w_i / w_sum_i);
  * LTC target values within a relative bound on the candidates where IEEE
    FP32 itself can meet it: lights wholly above the shading point's horizon
    that subtend at least a stated solid angle (the relative error of the
    clipped-polygon integral grows as the polygon's solid angle shrinks,
    because its edge terms nearly cancel).

The reservoir draws are replayed here from the noise definition
(noise_utility.glsl:20-95: LCG 1664525 / 1013904223, u = float(state) / 2^32,
per candidate a light draw, then the reservoir draw) and checked against the
candidates' recorded light indices before use. Solid angles are computed in
FP64 numpy (Van Oosterom and Strackee's triangle formula over a fan).
"""
import numpy as np

import orc
from risltc_amd import abi

LCG_A, LCG_C = 1664525, 1013904223


def gpu_debug_subset(ctx, consts, settings, vis, pix):
    """risltc_shade_debug over the whole frame; the output, sample records and
    candidate records of the listed pixels (gathered on the device)."""
    import torch
    P = int(consts.viewport_size[0]) * int(consts.viewport_size[1])
    spp, M = settings.sample_count_light, settings.ris_candidate_count
    out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
    rec = torch.zeros(P * spp * abi.SAMPLE_RECORD_FLOATS, dtype=torch.float32, device="cuda")
    cand = torch.zeros(P * spp * M * abi.CANDIDATE_RECORD_WORDS, dtype=torch.int32, device="cuda")
    ctx.shade_debug(consts, settings, vis, out, rec, cand)
    torch.cuda.synchronize()
    idx = torch.from_numpy(np.asarray(pix, np.int64)).cuda()
    o = out.view(P, 4)[idx].cpu().numpy()
    r = rec.view(P, spp, abi.SAMPLE_RECORD_FLOATS)[idx].cpu().numpy()
    c = cand.view(P, spp, M, abi.CANDIDATE_RECORD_WORDS)[idx].cpu().numpy().view(np.uint32)
    del out, rec, cand
    torch.cuda.empty_cache()
    return dict(out=o, rec=r, cand=c)


def sample_start_seeds(consts, pix, rec):
    """Noise state at the start of each pixel-sample: the pixel's accessor
    seed for the first sample, the state recorded at the end of the previous
    sample for the others (the oracle's REC_SEED)."""
    L = orc.lib()
    W = int(consts.viewport_size[0])
    rn0 = int(consts.noise_random_numbers[0])
    pix = np.asarray(pix, np.int64)
    spp = rec.shape[1]
    seeds = np.zeros((pix.size, spp), np.uint64)
    seeds[:, 0] = [L.orc_noise_seed(int(p % W), int(p // W), W, rn0) for p in pix]
    if spp > 1:
        seeds[:, 1:] = rec[:, :-1, abi.REC_SEED].view(np.uint32)
    return seeds


def _draws(seeds, count):
    """count successive uniform draws from each state (float32)."""
    s = seeds.astype(np.uint64).copy()
    out = np.zeros(seeds.shape + (count,), np.float32)
    for k in range(count):
        s = (s * LCG_A + LCG_C) & 0xFFFFFFFF
        out[..., k] = s.astype(np.uint32).astype(np.float32) * np.float32(2.0 ** -32)
    return out


def ris_tie_check(seeds, cand, cand64, rec, rec64, light_count, tie=1e-6, synced=None):
    """Replays the uniform-source reservoir of every pixel-sample with the
    FP64 oracle's weights and compares the build's selected light (rec) with
    the FP64 oracle's (rec64). Returns counts and the decision margins:
    margin = min over candidates with w > 0 of |r_i - w_i / w_sum_i|.
    synced (bool per sample, default all): the samples whose noise stream is
    the FP64 oracle's (the others are left out; see noise_sync)."""
    M = cand64.shape[2]
    if synced is None:
        synced = np.ones(seeds.shape, bool)
    d = _draws(seeds, 2 * M)
    u_idx, r = d[..., 0::2], d[..., 1::2]
    idx = np.minimum((u_idx * np.float32(light_count)).astype(np.int64), light_count - 1)
    replay_ok = np.array_equal(idx[synced], cand64[..., 0].astype(np.int64)[synced]) and \
        np.array_equal(idx[synced], cand[..., 0][synced])
    w64 = cand64[..., 3].view(np.float32).astype(np.float64)
    wb = cand[..., 3].view(np.float32).astype(np.float64)
    ws64 = np.cumsum(w64, axis=-1)
    wsb = np.cumsum(wb, axis=-1)
    with np.errstate(invalid="ignore", divide="ignore"):
        rho64 = np.where(ws64 > 0, w64 / ws64, 0.0)
        rhob = np.where(wsb > 0, wb / wsb, 0.0)
    take64 = (w64 > 0) & (r < rho64)
    last = np.where(take64.any(-1), M - 1 - np.argmax(take64[..., ::-1], axis=-1), -1)
    chosen64 = np.where(last >= 0, np.take_along_axis(idx, np.maximum(last, 0)[..., None], -1)[..., 0], -1)
    li64 = rec64[..., abi.REC_LIGHT].view(np.int32)
    lib_ = rec[..., abi.REC_LIGHT].view(np.int32)
    replay_agrees = (chosen64 == li64) | ~synced
    margin = np.where(w64 > 0, np.abs(r - rho64), np.inf).min(-1)
    mismatch = (lib_ != li64) & synced
    drho = np.where(((w64 > 0) | (wb > 0)) & synced[..., None], np.abs(rhob - rho64), 0.0).max(-1)
    mm = margin[mismatch]
    return dict(replay_ok=bool(replay_ok), replay_disagreements=int((~replay_agrees).sum()),
                samples=int(synced.sum()), mismatches=int(mismatch.sum()),
                mismatches_beyond_tie=int((mm > tie).sum()),
                max_mismatch_margin=float(mm.max()) if mm.size else 0.0,
                samples_within_tie=int((margin <= tie).sum()),
                max_normalised_weight_error=float(drho.max()), mismatch=mismatch, margin=margin)


def weight_flips(seeds, cand_a, cand_b):
    """Per sample: whether some reservoir decision (w_i > 0 and r_i <
    w_i / w_sum_i) comes out differently with the weights of build a and of
    build b (the same candidates and draws)."""
    M = cand_a.shape[2]
    r = _draws(seeds, 2 * M)[..., 1::2]

    def takes(cand):
        w = cand[..., 3].view(np.float32).astype(np.float64)
        ws = np.cumsum(w, axis=-1)
        with np.errstate(invalid="ignore", divide="ignore"):
            rho = np.where(ws > 0, w / ws, 0.0)
        return (w > 0) & (r < rho)
    return np.any(takes(cand_a) != takes(cand_b), axis=-1)


def noise_sync(consts, pix, rec, rec64):
    """Which pixel-samples start from the FP64 oracle's noise state. The
    final sample draws its specular direction only when the specular PSA is
    positive (shading_pass.frag.glsl:365-372), so a PSA that is zero in one
    arithmetic and positive in the other shifts every later sample of the
    pixel onto other draws. Returns (synced, cause_psa): for each sample
    whether its start state equals the FP64 oracle's, and, for the first
    sample of a pixel that is not, the FP64 specular PSA of the sample before
    it (the decision that split the streams; -1 where the split came from an
    earlier split or a different selected light)."""
    sb = sample_start_seeds(consts, pix, rec)
    s64 = sample_start_seeds(consts, pix, rec64)
    synced = sb == s64
    cause = np.full(synced.shape, -1.0)
    first = ~synced & np.concatenate([np.ones((synced.shape[0], 1), bool), synced[:, :-1]], 1)
    li_b = rec[..., abi.REC_LIGHT].view(np.int32)
    li_64 = rec64[..., abi.REC_LIGHT].view(np.int32)
    for p, j in zip(*np.nonzero(first)):
        if j > 0 and li_b[p, j - 1] == li_64[p, j - 1]:
            cause[p, j] = abs(float(rec64[p, j - 1, abi.REC_PSAS]))
    return synced, cause


def light_polygons(scene):
    """World-space vertices (N, vmax, 3) of the packed light records."""
    vmax = scene.light_vmax
    stride = 48 + 16 * vmax
    raw = np.frombuffer(bytes(scene.light_records), np.uint8).reshape(scene.light_count, stride)
    f = raw[:, 48:].copy().view(np.float32).reshape(scene.light_count, vmax, 4)[..., :3]
    return f.astype(np.float64)


def _solid_angle_tri(a, b, c):
    la, lb, lc = (np.linalg.norm(x, axis=-1) for x in (a, b, c))
    num = np.einsum("...i,...i->...", a, np.cross(b, c))
    den = la * lb * lc + np.einsum("...i,...i->...", a, b) * lc + np.einsum("...i,...i->...", a, c) * lb + \
        np.einsum("...i,...i->...", b, c) * la
    return 2.0 * np.arctan2(num, den)


def candidate_solid_angles(scene, pix, vis_h, rec64, cand64, horizon_margin=1e-4):
    """For every candidate (pixel, sample, m): the solid angle its light
    subtends from the shading point, and whether the light lies wholly above
    the point's horizon (every vertex at least horizon_margin * distance
    above the tangent plane). Positions are the FP64 oracle's records;
    normals are the geometric normals of the visible triangles, turned to the
    camera."""
    polys = light_polygons(scene)
    tri = scene.mesh.dequantized().reshape(-1, 3, 3).astype(np.float64)
    prim = vis_h[np.asarray(pix, np.int64)].astype(np.int64)
    t = tri[prim]
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    x = rec64[:, 0, abi.REC_POS:abi.REC_POS + 3].astype(np.float64)
    cam = np.array(scene.camera.position_world_space, np.float64)
    n = np.where(((cam - x) * n).sum(1, keepdims=True) < 0, -n, n)
    li = cand64[..., 0].astype(np.int64)
    v = polys[li] - x[:, None, None, None, :]  # (P, spp, M, vmax, 3)
    om = np.zeros(li.shape)
    for k in range(1, v.shape[3] - 1):
        om += _solid_angle_tri(v[..., 0, :], v[..., k, :], v[..., k + 1, :])
    om = np.abs(om)
    h = np.einsum("...ki,...i->...k", v, n[:, None, None, :])
    above = np.all(h > horizon_margin * np.linalg.norm(v, axis=-1), axis=-1)
    return om, above


def pixel_alpha(scene, pix, vis_h, roughness_factor=1.0):
    """GGX alpha of the listed pixels' materials: roughness squared, scaled
    and clamped to [0.0064, 1] (shading_pass.frag.glsl:644-646)."""
    prim = vis_h[np.asarray(pix, np.int64)].astype(np.int64)
    mats = np.asarray(scene.materials, np.float64)[scene.mesh.material_index[prim].astype(np.int64)]
    return np.clip(mats[:, 4] ** 2 * roughness_factor, 0.0064, 1.0)


def ltc_bar(oracle, consts, vis_h, pix, cand, rec64, cand64, scene, synced, min_solid_angle=0.03):
    """Relative error of the build's LTC_CP targets (cand[..., 2]) against the
    FP64 integrator run on the FP32 pixel setup (the kernels' setup is the
    FP32 oracle's bit for bit), on the candidates of synced samples whose
    light lies wholly above the shading point's horizon and subtends at least
    min_solid_angle sr. Returns (errors, count of candidates checked)."""
    om, above = candidate_solid_angles(scene, pix, vis_h, rec64, cand64)
    sel = above & (om >= min_solid_angle) & synced[..., None]
    p_i, s_i, m_i = np.nonzero(sel)
    if p_i.size == 0:
        return np.zeros(0), 0
    errs = np.zeros(p_i.size)
    order = np.argsort(p_i, kind="stable")
    p_i, s_i, m_i = p_i[order], s_i[order], m_i[order]
    lights = cand[p_i, s_i, m_i, 0].astype(np.int64)
    got = cand[p_i, s_i, m_i, 2].view(np.float32).astype(np.float64)
    for p in np.unique(p_i):
        k = np.flatnonzero(p_i == p)
        _, t64, _, _ = oracle.ltc_cp_targets(consts, vis_h, [pix[p]], lights[k][None, :])
        t = t64.ravel()
        live = t > 0
        errs[k[live]] = np.abs(got[k[live]] - t[live]) / t[live]
        errs[k[~live]] = np.where(got[k[~live]] == 0, 0.0, np.inf)
    return errs, int(p_i.size)
