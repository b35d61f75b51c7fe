"""Pins of the oracle's composed estimator (CPU only).

  * the reservoir (reservoir.glsl:16-40, shading_pass.frag.glsl:723-761)
    selects each candidate light with probability w / sum(w);
  * the final sample (shading_pass.frag.glsl:292-397: PSA sampling of the
    diffuse and the LTC-space polygon, BRDF, LTC density, MIS) is an unbiased
    estimator of the polygon's reflected radiance: its mean over many noise
    seeds equals a brute-force quadrature of f * L * cos over the light's area,
    computed here in FP64 from the closed-form BRDF.
"""
import math

import numpy as np
import pytest

import orc
from risltc_amd import abi, synth


def _brdf(n, o, alb, f0, a, l):
    """GGX + Lambert as in brdfs.glsl:43-93 (closed form, FP64), for rows of
    directions l; returns one RGB row per direction."""
    h = (l + o) / np.linalg.norm(l + o, axis=1, keepdims=True)
    li, lo, ndh, odh = l @ n, o @ n, h @ n, h @ o
    D = a * a / (((a * a - 1) * ndh * ndh + 1) ** 2)
    V = 0.5 / (li * np.sqrt(lo * lo * (1 - a * a) + a * a) + lo * np.sqrt(li * li * (1 - a * a) + a * a))
    F = f0[None, :] + (1 - f0[None, :]) * ((1 - np.clip(odh, 0, 1)) ** 5)[:, None]
    lum = F @ np.array([0.21263901, 0.71516868, 0.07219232])
    return (alb[None, :] + (D * V * lum)[:, None]) / math.pi


def test_reservoir_selects_in_proportion_to_weight(small_oracle, small_scene):
    """Over all samples of the small scene, the light at candidate slot i is
    the selected one as often as sum of its weights / w_sum predicts, for every
    slot (so neither early nor late candidates are favoured)."""
    o = small_oracle
    s = small_scene.settings(sample_count_light=2, ris_candidate_count=8, shadow_rays=0)
    consts = o.frame_constants()
    vis = o.render_visibility(consts)
    _, rec, cand = o.shade(consts, s, vis, records=True, candidates=True)
    idx = cand[..., 0].astype(np.int64).reshape(-1, 8)
    w = cand[..., 3].view(np.float32).astype(np.float64).reshape(-1, 8)
    chosen = rec[..., abi.REC_LIGHT].view(np.int32).reshape(-1)
    wsum = w.sum(1)
    ok = wsum > 0
    idx, w, chosen, wsum = idx[ok], w[ok], chosen[ok], wsum[ok]
    assert idx.shape[0] > 20000
    # probability that the light of slot i is the chosen light
    same = idx[:, :, None] == idx[:, None, :]
    p = (same * w[:, None, :]).sum(2) / wsum[:, None]
    observed = (idx == chosen[:, None]).sum(0)
    expected = p.sum(0)
    sigma = np.sqrt((p * (1 - p)).sum(0))
    assert np.all(np.abs(observed - expected) < 5 * sigma + 1), (observed, expected, sigma)
    # and no sample with positive weights is left without a light
    assert np.all(chosen >= 0)


def _single_light_scene(roughness, metal):
    """A 20 m ground plane and one tilted 1 x 0.6 m light 3 m above it."""
    g = 10.0
    tris = [[(-g, -g, 0), (g, -g, 0), (g, g, 0)], [(-g, -g, 0), (g, g, 0), (-g, g, 0)]]
    nrm = [[(0, 0, 1)] * 3] * 2
    uvs = [[(0, 0), (1, 0), (1, 1)], [(0, 0), (1, 1), (0, 1)]]
    square = np.array([[-0.5, -0.5], [0.5, -0.5], [0.5, 0.5], [-0.5, 0.5]])
    light = synth.Light(square, (0.4, 0.3, 0.2), (0.5, 1.0, 3.0), (1.0, 0.6), (6.0, 4.0, 2.5))
    material = [0.6, 0.5, 0.4, 0.0, math.sqrt(roughness), metal]
    cam = synth.make_camera((0.0, -9.0, 6.0), math.pi, math.pi / 2 - 0.6)
    return synth._finish("one", tris, nrm, uvs, [0, 0], [material], [light], cam, 24, 24, 1, 4, 0)


@pytest.mark.parametrize("roughness,metal", [(0.3, 0.0), (0.05, 0.7)])
def test_estimator_is_unbiased(orc_lib, ltc_table, roughness, metal):
    scene = _single_light_scene(roughness, metal)
    o = orc.Oracle(scene, *ltc_table)
    exposure = 1.0
    consts0 = o.frame_constants(exposure=exposure)
    vis = o.render_visibility(consts0)
    ground = np.flatnonzero(vis < 2)
    pix = ground[np.linspace(0, ground.size - 1, 6).astype(int)].astype(np.uint32)
    spp = 256
    s = scene.settings(sample_count_light=spp, ris_candidate_count=2, shadow_rays=0)
    means = []
    for seed in range(10):
        c = o.frame_constants(exposure=exposure, noise_seed=seed, animate=True)
        out, rec, _ = o.shade(c, s, vis, pixels=pix, fp64=True, records=(seed == 0))
        means.append(out[:, :3])
        if seed == 0:
            pos = rec[:, 0, abi.REC_POS:abi.REC_POS + 3].astype(np.float64)
    means = np.array(means)
    est = means.mean(0)
    se = means.std(0, ddof=1) / math.sqrt(means.shape[0])
    # quadrature over the light: midpoint rule on a 240 x 240 grid
    L = scene.lights[0]
    rad = L.radiance.astype(np.float64)
    wv = L.world.astype(np.float64)
    nl = L.plane[:3].astype(np.float64)
    k = 240
    u = (np.arange(k) + 0.5) / k
    uu, vv = np.meshgrid(u, u, indexing="ij")
    pts = (wv[0] + uu[..., None] * (wv[1] - wv[0]) + vv[..., None] * (wv[3] - wv[0])).reshape(-1, 3)
    dA = np.linalg.norm(np.cross(wv[1] - wv[0], wv[3] - wv[0])) / (k * k)
    cam = np.array(scene.camera.position_world_space, np.float64)
    mat = scene.materials[0].astype(np.float64)
    base, m = mat[0:3], mat[5]
    alb = base * (1 - m)
    f0 = 0.02 * (1 - m) + base * m
    alpha = min(max(mat[4] ** 2, 0.0064), 1.0)
    n = np.array([0.0, 0.0, 1.0])
    for p in range(pix.size):
        x = pos[p]
        ov = (cam - x) / np.linalg.norm(cam - x)
        d = pts - x
        r2 = (d * d).sum(1)
        w = d / np.sqrt(r2)[:, None]
        cos_x = np.maximum(w @ n, 0.0)
        cos_y = np.abs(w @ nl)
        f = _brdf(n, ov, alb, f0, alpha, w)
        f[cos_x <= 0] = 0.0
        ref = exposure * (f * rad[None, :] * (cos_x * cos_y / r2 * dA)[:, None]).sum(0)
        assert np.all(np.abs(est[p] - ref) < 5 * se[p] + 2e-3 * ref), (p, est[p], ref, se[p])


# ---------------------------------------------------------------- learned source
def _random_tables(rng, P, K, sizes, bad=False):
    """Per-pixel tables: 16 distinct non-empty clusters with random 8-bit weights
    (some zero, as the proposal kernel writes them), on some pixels all zero.
    With bad=True, tables that no proposal kernel writes but a caller may
    pass: slots naming empty clusters, ids past the clusters, and clusters
    named twice."""
    nonempty = np.flatnonzero(sizes > 0)
    empty = np.flatnonzero(sizes == 0)
    tab = np.zeros((P, 32), np.uint8)
    for p in range(P):
        if p % 7 == 3:
            continue  # T = 0: uniform fallback
        ids = rng.choice(nonempty, min(16, nonempty.size), replace=False)
        if bad:
            pick = rng.random(ids.size)
            ids = np.where(pick < 0.15, rng.integers(K, 256, ids.size), ids)   # past the clusters
            if empty.size:
                ids = np.where((pick >= 0.15) & (pick < 0.3), rng.choice(empty, ids.size), ids)  # empty
            dup = (pick >= 0.3) & (pick < 0.4)
            ids = np.where(dup, ids[0], ids)  # the first slot's cluster again
        w = rng.integers(0, 256, ids.size)
        w[rng.random(ids.size) < 0.2] = 0
        w[0] = max(w[0], 1)
        ids[w == 0] = 0
        tab[p, :ids.size] = ids
        tab[p, 16:16 + ids.size] = w
    return tab


def _q_numpy(tab, idx, col, sizes, N):
    """q = 1/(2N) + T_bad/(2 T N) + w_g / (2 T |g|) in float32 with the
    kernel's operation order: table ids name groups (col: the group of each
    light, sizes: lights per group), T_bad is the weight of slots the draw
    sends to a uniform light (ids past the groups, empty groups), w_g the
    summed weight of the slots naming the light's group."""
    f = np.float32
    K = sizes.size
    T = int(tab[16:].astype(np.int64).sum())
    if T == 0:
        return f(1.0) / f(N)
    q = f(0.5) / f(N)
    bad = sum(int(tab[16 + s]) for s in range(16) if tab[16 + s] != 0 and (tab[s] >= K or sizes[tab[s]] == 0))
    if bad:
        q = q + (f(0.5) * f(bad)) / (f(T) * f(N))
    c = col[idx]
    wc = sum(int(tab[16 + s]) for s in range(16) if tab[16 + s] != 0 and tab[s] == c)
    if wc:
        q = q + (f(0.5) * f(wc)) / (f(T) * f(sizes[c]))
    return q


def _source_oracle(scene, ltc_table, rng, K=16, bvh=True, bad=False, group_log2=0):
    """An oracle with the learned source: K clusters (lights i -> 7 i mod K;
    with bad tables the last cluster is empty), tables of random weights over
    the groups of 1 << group_log2 consecutive clusters. Returns the oracle,
    the tables, the group of each light and the lights per group."""
    o = orc.Oracle(scene, *ltc_table, bvh=bvh)
    N = scene.light_count
    col = (np.arange(N) * 7 % (K - 1 if bad else K)).astype(np.uint32)
    groups = (K + (1 << group_log2) - 1) >> group_log2
    gcol = col >> group_log2
    sizes = np.bincount(gcol, minlength=groups)
    P = scene.width * scene.height
    tab = _random_tables(rng, P, groups, sizes, bad)
    o.set_proposal(tab, col, K, group_log2)
    return o, tab, gcol, sizes


@pytest.mark.parametrize("bad,group_log2,K", [(False, 0, 16), (True, 0, 16), (False, 2, 32), (True, 2, 30)],
                         ids=["kernel_tables", "bad_slots", "quads", "quads_bad_slots"])
def test_learned_draw_density(small_scene, ltc_table, bad, group_log2, K):
    """With the learned source: every candidate's weight is exactly p_hat / q
    with q from the table (a pdf over the lights), and the candidate lights
    are distributed as q (chi-square). Also for tables with slots of empty
    clusters, ids past the clusters and repeated clusters, and for tables
    whose slots name quads of clusters (1024-output networks; 30 clusters:
    the last quad has two)."""
    rng = np.random.default_rng(31)
    o, tab, col, sizes = _source_oracle(small_scene, ltc_table, rng, K=K, bad=bad, group_log2=group_log2)
    N = small_scene.light_count
    consts = o.frame_constants()
    vis = o.render_visibility(consts)
    shaded = np.flatnonzero((vis >> 31) == 0)
    pix = shaded[::23].astype(np.uint32)
    s = small_scene.settings(sample_count_light=4, ris_candidate_count=8, shadow_rays=0,
                             proposal=abi.PROPOSAL_LEARNED)
    _, _, cand = o.shade(consts, s, vis, pixels=pix, candidates=True)
    idx = cand[..., 0]
    phat = cand[..., 2].view(np.float32)
    w = cand[..., 3].view(np.float32)
    counts = np.zeros(N)
    expect = np.zeros(N)
    for k, p in enumerate(pix):
        qs = np.array([_q_numpy(tab[p], l, col, sizes, N) for l in range(N)], np.float32)
        assert abs(qs.astype(np.float64).sum() - 1.0) < 1e-5
        q = qs[idx[k]]
        assert np.array_equal(w[k], (phat[k] / q).astype(np.float32)), k
        counts += np.bincount(idx[k].ravel(), minlength=N)
        expect += qs.astype(np.float64) * idx[k].size
    chi2 = ((counts - expect) ** 2 / expect).sum()
    assert chi2 < N + 6 * math.sqrt(2 * N), chi2


def test_learned_target_factor_constant():
    """The cache factor of an empty entry (byte 255) is 1 to within 2 ulp."""
    f = np.float32(1.0 / 128)
    scale = (np.float32(1) - f) / np.float32(255)
    assert abs(float(f + np.float32(255) * scale) - 1.0) < 2.5e-7
    assert f + np.float32(0) * scale == f


@pytest.mark.parametrize("bad,group_log2,K", [(False, 0, 6), (True, 0, 6), (True, 2, 10)],
                         ids=["kernel_tables", "bad_slots", "quads_bad_slots"])
def test_learned_estimator_is_unbiased(orc_lib, ltc_table, bad, group_log2, K):
    """Mean of the learned-source estimator over many noise seeds equals the
    plain (uniform-source) mean on the same pixels, shadows off; also with
    tables that name empty clusters, ids past the clusters and repeated
    clusters, and with tables of quads of clusters."""
    scene = synth.small_scene(seed=3, width=48, height=48, light_count=24)
    rng = np.random.default_rng(32)
    o, _, _, _ = _source_oracle(scene, ltc_table, rng, K=K, bvh=False, bad=bad, group_log2=group_log2)
    consts = o.frame_constants()
    o2 = orc.Oracle(scene, *ltc_table)
    vis = o2.render_visibility(consts)
    pix = np.flatnonzero((vis >> 31) == 0)[::41][:10].astype(np.uint32)
    means = {}
    for prop in (abi.PROPOSAL_UNIFORM, abi.PROPOSAL_LEARNED):
        s = scene.settings(sample_count_light=256, ris_candidate_count=4, shadow_rays=0, proposal=prop)
        frames = []
        for seed in range(12):
            c = o.frame_constants(noise_seed=seed, animate=True)
            out, _, _ = o.shade(c, s, vis, pixels=pix, fp64=True)
            frames.append(out[:, :3])
        frames = np.array(frames)
        means[prop] = (frames.mean(0), frames.std(0, ddof=1) / math.sqrt(len(frames)))
    (mu_u, se_u), (mu_l, se_l) = means[abi.PROPOSAL_UNIFORM], means[abi.PROPOSAL_LEARNED]
    se = np.sqrt(se_u ** 2 + se_l ** 2)
    assert np.all(np.abs(mu_u - mu_l) < 5 * se + 1e-4 * np.abs(mu_u) + 1e-7), (mu_u, mu_l, se)


# ---------------------------------------------------------------- visibility cache
def _fmix(h):
    """MurmurHash3 x86_32 finaliser (public algorithm), uint64 arrays."""
    h = h & 0xFFFFFFFF
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & 0xFFFFFFFF
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & 0xFFFFFFFF
    h ^= h >> 16
    return h


def _mix(h, k):
    """One MurmurHash3 x86_32 block."""
    k = (k * 0xCC9E2D51) & 0xFFFFFFFF
    k = ((k << 15) | (k >> 17)) & 0xFFFFFFFF
    k = (k * 0x1B873593) & 0xFFFFFFFF
    h = (h ^ k) & 0xFFFFFFFF
    h = ((h << 13) | (h >> 19)) & 0xFFFFFFFF
    return (h * 5 + 0xE6546B64) & 0xFFFFFFFF


def _cache_oracle(scene, ltc_table, log2, cell, prior, rng, per_light=False, bvh=True):
    o, tab, col, sizes = _source_oracle(scene, ltc_table, rng, bvh=bvh)
    N = scene.light_count
    slots = N if per_light else orc.CACHE_CLUSTER_SLOTS
    used = N if per_light else 16
    counts = np.zeros((1 << log2, slots, 2), dtype=np.uint32)
    counts[:, :used] = rng.integers(0, 40, (1 << log2, used, 2))
    o.set_cache(counts, cell, prior, per_light)
    return o, tab, col, sizes, counts


def test_cache_keys_follow_cells_and_normals(small_scene, ltc_table):
    """Keys of ground pixels (normal +z, code 4) equal MurmurHash3 of their
    cell indices floor(p / cell), computed here from the recorded positions;
    the table mask keeps the low bits."""
    rng = np.random.default_rng(40)
    cell, log2 = 0.5, 12
    o = _cache_oracle(small_scene, ltc_table, log2, cell, 2.0, rng)[0]
    consts = o.frame_constants()
    vis = o.render_visibility(consts)
    ground = np.flatnonzero(vis < 2)[::23].astype(np.uint32)  # triangles 0, 1: the ground plane
    s = small_scene.settings(sample_count_light=1, ris_candidate_count=8, shadow_rays=0)
    _, rec, _ = o.shade(consts, s, vis, pixels=ground, records=True)
    rec = rec[:, 0]
    ok = rec[:, abi.REC_LIGHT].view(np.int32) >= 0
    assert ok.sum() > 100
    pos = rec[ok, abi.REC_POS:abi.REC_POS + 3].astype(np.float32)
    assert np.all(np.abs(pos[:, 2]) < 1e-3)  # the plane, up to the position quantisation
    inv = np.float32(1.0) / np.float32(cell)
    idx = np.floor(pos * inv).astype(np.int64) & 0xFFFFFFFF
    h = np.zeros(idx.shape[0], dtype=np.uint64)
    for a in range(3):
        h = _mix(h, idx[:, a].astype(np.uint64))
    h = _mix(h, np.uint64(4))
    want = _fmix(h) & ((1 << log2) - 1)
    got = rec[ok, abi.REC_CACHE_KEY].view(np.uint32)
    assert np.array_equal(got, want.astype(np.uint32))
    # neighbouring pixels in one cell share a key; the frame spans many cells
    assert 20 < np.unique(got).size < got.size


def test_cache_bytes_end_points():
    """orc_cache_bytes: 255 marks an empty entry; only unoccluded rays give
    254 (in the limit), only occluded rays 0 (in the limit); equal counts give
    127; the prior pulls small counts toward the middle; monotone in a."""
    cb = orc.cache_bytes
    c = np.array([[0, 0], [1, 0], [0, 1], [10**6, 0], [0, 10**6], [7, 7], [100, 0], [0, 100]], dtype=np.uint32)
    b = cb(c, 1.0)
    assert b[0] == 255 and b[3] == 254 and b[4] == 0 and b[5] == 127
    assert 127 < b[1] < 254 and 0 < b[2] < 127 and b[1] < b[6] and b[2] > b[7]
    assert cb(c[1:2], 8.0)[0] < cb(c[1:2], 1.0)[0]
    a = np.arange(0, 200, dtype=np.uint32)
    mono = cb(np.stack([a, np.full_like(a, 50)], 1), 1.0)
    assert np.all(np.diff(mono.astype(int)) >= 0) and mono.max() <= 254


@pytest.mark.parametrize("per_light", [False, True])
def test_cache_target_uses_entry_bytes(small_scene, ltc_table, per_light):
    """With a cache, each candidate's target is p_hat times f + v (1 - f) / 255
    with v the cache byte at the pixel's key and the light's slot (its
    cluster, or the light); 255 (about 1) for empty entries. Candidates and
    source densities do not depend on the cache."""
    rng = np.random.default_rng(41)
    log2, cell, prior = 10, 0.5, 2.0
    o, tab, col, sizes, counts = _cache_oracle(small_scene, ltc_table, log2, cell, prior, rng, per_light)
    counts[rng.random(counts.shape[:2]) < 0.25] = 0
    vis8 = o.set_cache(counts, cell, prior, per_light)
    consts = o.frame_constants()
    vis = o.render_visibility(consts)
    pix = np.flatnonzero((vis >> 31) == 0)[::19].astype(np.uint32)
    s = small_scene.settings(sample_count_light=1, ris_candidate_count=8, shadow_rays=0,
                             proposal=abi.PROPOSAL_LEARNED)
    _, rec, cached = o.shade(consts, s, vis, pixels=pix, records=True, candidates=True)
    o.set_cache(None, cell, prior)
    _, _, plain = o.shade(consts, s, vis, pixels=pix, candidates=True)
    assert np.array_equal(cached[..., 0], plain[..., 0]) and np.array_equal(cached[..., 1], plain[..., 1])
    key = rec[:, 0, abi.REC_CACHE_KEY].view(np.uint32)
    li = plain[..., 0]
    v = vis8[key[:, None, None], li if per_light else col[li]].astype(np.float32)
    assert np.any(v == 255) and np.any(v != 255)
    f = np.float32(0.125)
    fac = f + v * ((np.float32(1) - f) / np.float32(255))
    want = (plain[..., 2].view(np.float32) * fac).astype(np.float32)
    assert np.array_equal(cached[..., 2].view(np.float32), want)
    # w = target / q with the same q
    ok = plain[..., 3].view(np.float32) > 0
    q = plain[..., 2].view(np.float32)[ok] / plain[..., 3].view(np.float32)[ok]
    assert np.allclose(want[ok] / cached[..., 3].view(np.float32)[ok], q, rtol=1e-6)


def test_cached_learned_estimator_is_unbiased(orc_lib, ltc_table):
    """A random visibility cache with the learned source keeps the
    estimator's mean equal to the plain one (the target stays positive
    wherever the LTC target is)."""
    scene = synth.small_scene(seed=3, width=48, height=48, light_count=24)
    rng = np.random.default_rng(42)
    o = _cache_oracle(scene, ltc_table, 8, 0.7, 4.0, rng, per_light=True, bvh=False)[0]
    consts = o.frame_constants()
    o2 = orc.Oracle(scene, *ltc_table)
    vis = o2.render_visibility(consts)
    pix = np.flatnonzero((vis >> 31) == 0)[::41][:10].astype(np.uint32)
    means = {}
    for prop in (abi.PROPOSAL_UNIFORM, abi.PROPOSAL_LEARNED):
        s = scene.settings(sample_count_light=256, ris_candidate_count=4, shadow_rays=0, proposal=prop)
        frames = []
        for seed in range(12):
            c = o.frame_constants(noise_seed=seed, animate=True)
            out, _, _ = o.shade(c, s, vis, pixels=pix, fp64=True)
            frames.append(out[:, :3])
        frames = np.array(frames)
        means[prop] = (frames.mean(0), frames.std(0, ddof=1) / math.sqrt(len(frames)))
    (mu_u, se_u), (mu_l, se_l) = means[abi.PROPOSAL_UNIFORM], means[abi.PROPOSAL_LEARNED]
    se = np.sqrt(se_u ** 2 + se_l ** 2)
    assert np.all(np.abs(mu_u - mu_l) < 5 * se + 1e-4 * np.abs(mu_u) + 1e-7), (mu_u, mu_l, se)


# ---------------------------------------------------------------- baseline techniques


def _quadrature(scene, x, exposure, triangle_only):
    """f * L * cos_x * cos_y / r^2 over the light of _single_light_scene (a
    parallelogram v0, v1, v3 spanned), or over its triangle v0, v1, v2 only
    (u >= v in the parallelogram's coordinates; diagonal cells count half)."""
    L = scene.lights[0]
    rad = L.radiance.astype(np.float64)
    wv = L.world.astype(np.float64)
    nl = L.plane[:3].astype(np.float64)
    k = 240
    u = (np.arange(k) + 0.5) / k
    uu, vv = np.meshgrid(u, u, indexing="ij")
    wgt = np.ones_like(uu)
    if triangle_only:
        wgt = np.where(uu > vv, 1.0, np.where(uu == vv, 0.5, 0.0))
    pts = (wv[0] + uu[..., None] * (wv[1] - wv[0]) + vv[..., None] * (wv[3] - wv[0])).reshape(-1, 3)
    dA = np.linalg.norm(np.cross(wv[1] - wv[0], wv[3] - wv[0])) / (k * k)
    cam = np.array(scene.camera.position_world_space, np.float64)
    mat = scene.materials[0].astype(np.float64)
    base, m = mat[0:3], mat[5]
    alb = base * (1 - m)
    f0 = 0.02 * (1 - m) + base * m
    alpha = min(max(mat[4] ** 2, 0.0064), 1.0)
    n = np.array([0.0, 0.0, 1.0])
    ov = (cam - x) / np.linalg.norm(cam - x)
    d = pts - x
    r2 = (d * d).sum(1)
    w = d / np.sqrt(r2)[:, None]
    cos_x = np.maximum(w @ n, 0.0)
    cos_y = np.abs(w @ nl)
    f = _brdf(n, ov, alb, f0, alpha, w)
    f[cos_x <= 0] = 0.0
    return exposure * (f * rad[None, :] * (cos_x * cos_y / r2 * dA * wgt.ravel())[:, None]).sum(0)


@pytest.mark.parametrize("technique,triangle_only", [(abi.SAMPLE_POLYGON_AREA_TURK, True),
                                                     (abi.SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE, False)])
@pytest.mark.parametrize("light_sampling", [abi.LIGHT_UNIFORM, abi.LIGHT_RESERVOIR])
def test_baseline_techniques_converge(orc_lib, ltc_table, technique, triangle_only, light_sampling):
    """The reference's baselines (shading_pass.frag.glsl:410-418, 458-551)
    with shadows off: projected-solid-angle sampling converges to the light's
    reflected radiance; Turk's area sampling to the part of it due to the
    triangle of vertices 0, 1, 2 -- the reference samples and measures only
    that triangle (polygon_sampling_related_work.glsl:34-43,
    shading_pass.frag.glsl:148-152), half of this rectangular light."""
    scene = _single_light_scene(0.3, 0.0)
    o = orc.Oracle(scene, *ltc_table)
    consts0 = o.frame_constants(exposure=1.0)
    vis = o.render_visibility(consts0)
    ground = np.flatnonzero(vis < 2)
    pix = ground[np.linspace(0, ground.size - 1, 6).astype(int)].astype(np.uint32)
    s = scene.settings(sample_count_light=256, ris_candidate_count=2, shadow_rays=0, light_sampling=light_sampling,
                       polygon_sampling_technique=technique)
    means = []
    for seed in range(10):
        c = o.frame_constants(exposure=1.0, noise_seed=seed, animate=True)
        out, rec, _ = o.shade(c, s, vis, pixels=pix, fp64=True, records=(seed == 0))
        means.append(out[:, :3])
        if seed == 0:
            pos = rec[:, 0, abi.REC_POS:abi.REC_POS + 3].astype(np.float64)
    means = np.array(means)
    est, se = means.mean(0), means.std(0, ddof=1) / math.sqrt(means.shape[0])
    for p in range(pix.size):
        ref = _quadrature(scene, pos[p], 1.0, triangle_only)
        assert np.all(np.abs(est[p] - ref) < 5 * se[p] + 4e-3 * ref), (p, est[p], ref, se[p])
    if triangle_only:
        full = np.array([_quadrature(scene, pos[p], 1.0, False) for p in range(pix.size)])
        assert np.all(est < 0.8 * full)


def test_psa_uniform_light_equals_ltc_cp(small_oracle, small_scene):
    """With uniform light selection (shading_pass.frag.glsl:707-721) the
    projected-solid-angle branch (458-551) and the LTC_CP final sample
    (292-397) compute the same estimate from the same draws: the frames are
    equal bit for bit. The techniques differ only in the RIS target."""
    o = small_oracle
    consts = o.frame_constants()
    vis = o.render_visibility(consts)
    a = small_scene.settings(light_sampling=abi.LIGHT_UNIFORM, sample_count_light=2)
    b = small_scene.settings(light_sampling=abi.LIGHT_UNIFORM, sample_count_light=2,
                             polygon_sampling_technique=abi.SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE)
    oa, _, _ = o.shade(consts, a, vis)
    ob, _, _ = o.shade(consts, b, vis)
    assert np.array_equal(oa.view(np.uint32), ob.view(np.uint32))


def test_baselines_reject_learned_source(small_oracle, small_scene):
    o = small_oracle
    consts = o.frame_constants()
    vis = o.render_visibility(consts)
    s = small_scene.settings(polygon_sampling_technique=abi.SAMPLE_POLYGON_AREA_TURK, proposal=abi.PROPOSAL_CACHE)
    with pytest.raises(Exception):
        o.shade(consts, s, vis, pixels=np.arange(4, dtype=np.uint32))


# ---------------------------------------------------------------- SAMPLE_COUNT
@pytest.mark.parametrize("technique,triangle_only", [(abi.SAMPLE_POLYGON_LTC_CP, False),
                                                     (abi.SAMPLE_POLYGON_AREA_TURK, True),
                                                     (abi.SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE, False)])
@pytest.mark.parametrize("light_sampling", [abi.LIGHT_UNIFORM, abi.LIGHT_RESERVOIR])
def test_sample_count_converges(orc_lib, ltc_table, technique, triangle_only, light_sampling):
    """SAMPLE_COUNT = 3 (shading_pass.frag.glsl:365-396, 412-417, 518-559):
    every technique still converges to the quadrature of its light (a sum
    not divided by SAMPLE_COUNT would be three times too large)."""
    scene = _single_light_scene(0.3, 0.0)
    o = orc.Oracle(scene, *ltc_table)
    vis = o.render_visibility(o.frame_constants(exposure=1.0))
    ground = np.flatnonzero(vis < 2)
    pix = ground[np.linspace(0, ground.size - 1, 4).astype(int)].astype(np.uint32)
    s = scene.settings(sample_count_light=128, ris_candidate_count=2, shadow_rays=0, light_sampling=light_sampling,
                       polygon_sampling_technique=technique, sample_count=3)
    means = []
    for seed in range(8):
        c = o.frame_constants(exposure=1.0, noise_seed=seed, animate=True)
        out, rec, _ = o.shade(c, s, vis, pixels=pix, fp64=True, records=(seed == 0))
        means.append(out[:, :3])
        if seed == 0:
            pos = rec[:, 0, abi.REC_POS:abi.REC_POS + 3].astype(np.float64)
    means = np.array(means)
    est, se = means.mean(0), means.std(0, ddof=1) / math.sqrt(means.shape[0])
    for p in range(pix.size):
        ref = _quadrature(scene, pos[p], 1.0, triangle_only)
        assert np.all(np.abs(est[p] - ref) < 5 * se[p] + 4e-3 * ref), (p, est[p], ref, se[p])


def test_sample_count_averages_independent_draws(orc_lib, ltc_table):
    """With one light and uniform light selection the final sample is the
    mean of SAMPLE_COUNT independent MIS samples: the variance of a pixel over
    noise seeds falls by the factor SAMPLE_COUNT (4 against 1; draws reused
    across iterations would leave it unchanged)."""
    scene = _single_light_scene(0.3, 0.0)
    o = orc.Oracle(scene, *ltc_table)
    vis = o.render_visibility(o.frame_constants(exposure=1.0))
    ground = np.flatnonzero(vis < 2)
    pix = ground[np.linspace(0, ground.size - 1, 4).astype(int)].astype(np.uint32)
    var = {}
    for k in (1, 4):
        s = scene.settings(sample_count_light=4, shadow_rays=0, light_sampling=abi.LIGHT_UNIFORM, sample_count=k)
        outs = [o.shade(o.frame_constants(exposure=1.0, noise_seed=seed, animate=True), s, vis, pixels=pix)[0][:, :3]
                for seed in range(300)]
        var[k] = np.array(outs, np.float64).var(0, ddof=1).sum(1)
    ratio = var[4] / var[1]
    # the ratio of two variance estimates of 299 degrees of freedom each
    assert np.all((ratio > 0.25 / 1.5) & (ratio < 0.25 * 1.5)), ratio


@pytest.mark.parametrize("light_sampling", [abi.LIGHT_UNIFORM, abi.LIGHT_RESERVOIR])
def test_area_sampling_shares_one_visibility(small_oracle, small_scene, light_sampling):
    """Area sampling with SAMPLE_COUNT = 2 and one sample per pixel: the loop
    of evaluate_polygonal_light_shading keeps one visibility flag
    (shading_pass.frag.glsl:412-417, get_polygon_visibility 112-129 traces
    only while it is true) and the caller multiplies by it (720, 755), so a
    pixel is either its unshadowed value or zero -- never the mean of one
    visible and one occluded sample. The draws do not depend on visibility,
    so shadowed and unshadowed frames use the same points."""
    o = small_oracle
    consts = o.frame_constants()
    vis = o.render_visibility(consts)
    kw = dict(sample_count_light=1, ris_candidate_count=4, light_sampling=light_sampling, sample_count=2,
              polygon_sampling_technique=abi.SAMPLE_POLYGON_AREA_TURK)
    shadowed, _, _ = o.shade(consts, small_scene.settings(shadow_rays=1, **kw), vis)
    free, _, _ = o.shade(consts, small_scene.settings(shadow_rays=0, **kw), vis)
    a, b = shadowed[:, :3], free[:, :3]
    same = np.all(a.view(np.uint32) == b.view(np.uint32), axis=1)
    zero = np.all(a == 0, axis=1)
    assert np.all(same | zero)
    lit = np.any(b > 0, axis=1)
    # the scene has both outcomes among lit pixels
    assert (lit & same).sum() > 100 and (lit & zero & ~same).sum() > 100
