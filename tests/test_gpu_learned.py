"""GPU tests of the learned RIS proposal.

  * MLP inference on MFMA: the logits of risltc_proposal_update equal a numpy
    forward pass on the same features (bf16 operands and activations, FP32
    accumulation) within the accumulation-order tolerance stated below;
  * the per-pixel tables hold, per group of clusters, the four largest
    logits' clusters and round(255 exp(z - z_max));
  * learned shading (exact build: learned source, with and without a
    visibility cache) is bitwise equal to the oracle given the device's
    tables, clusters and cache;
  * the resolve kernel counts every traced shadow ray in the cache, and the
    refresh makes the cache's bytes from the counts as the oracle does;
  * one training step: the gradient of the softmax cross-entropy on the
    selected lights' clusters equals a numpy FP64 backward pass of the same
    batch (bf16 at the same points), and the Adam update equals the Adam
    formula on that gradient; repeated steps lower the loss.
"""
import math

import numpy as np
import pytest

import orc
from risltc_amd import abi, synth

pytestmark = pytest.mark.gpu

PHI = 0x9E3779B9


def bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def layer_shapes(width, outputs):
    return [(width, 96), (width, width), (width, width), (outputs, width)]


def layer_bounds(width, outputs=256):
    """(start, end) of each layer's weights + bias in the parameter array:
    every matrix [out][in] followed by its bias."""
    out, off = [], 0
    for o, i in layer_shapes(width, outputs):
        out.append((off, off + o * i + o))
        off += o * i + o
    return out


def network_of(size):
    """(width, outputs) of a parameter array's size."""
    for w in (64, 128):
        for o in (256, 1024):
            if layer_bounds(w, o)[-1][1] == size:
                return w, o
    raise ValueError(size)


def layers(params):
    p = params
    width, outputs = network_of(p.size)
    res = []
    for (a, b), (o, i) in zip(layer_bounds(width, outputs), layer_shapes(width, outputs)):
        res.append((p[a:a + o * i].reshape(o, i), p[a + o * i:b]))
    return tuple(res)


def forward(params, X):
    """Activations as the kernels form them: bf16 inputs and hidden outputs."""
    (W1, b1), (W2, b2), (W3, b3), (W4, b4) = layers(params)
    X = X.astype(np.float64)
    H1 = bf16(np.maximum(X @ bf16(W1).T.astype(np.float64) + b1, 0)).astype(np.float64)
    H2 = bf16(np.maximum(H1 @ bf16(W2).T.astype(np.float64) + b2, 0)).astype(np.float64)
    H3 = bf16(np.maximum(H2 @ bf16(W3).T.astype(np.float64) + b3, 0)).astype(np.float64)
    Z = H3 @ bf16(W4).T.astype(np.float64) + b4
    return H1, H2, H3, Z


def hash32(x):
    x = np.asarray(x, dtype=np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


@pytest.fixture(scope="module")
def learned(small_scene, ltc_table):
    return _learned_context(small_scene, ltc_table)


@pytest.fixture(scope="module")
def learned_wide(small_scene, ltc_table):
    """The 96-128-128-128-256 network."""
    return _learned_context(small_scene, ltc_table, width=128)


@pytest.fixture(scope="module")
def learned_k1024(ltc_table):
    """1024 clusters of 1500 lights and the 96-128-128-128-1024 network
    (tables of quads of clusters)."""
    sc = synth.small_scene(seed=6, width=96, height=64, light_count=1500)
    return _learned_context(sc, ltc_table, width=128, outputs=1024, K=1024)


@pytest.fixture(scope="module")
def learned_odd(ltc_table):
    """61 lights: a per-light cache row is not a multiple of 16 bytes."""
    return _learned_context(synth.small_scene(seed=5, width=96, height=80, light_count=61), ltc_table)


def _learned_context(small_scene, ltc_table, width=64, outputs=256, K=16):
    import risltc_amd
    import torch
    ctx = risltc_amd.Context(0)
    ctx.load_scene(small_scene, *ltc_table)
    col = ctx.cluster_lights(K, 20)
    ctx.proposal_init(7, width=width, outputs=outputs)
    consts = ctx.frame_constants()
    P = small_scene.width * small_scene.height
    vis = torch.empty(P, dtype=torch.int32, device="cuda")
    ctx.render_visibility(consts, vis)
    tb = abi.proposal_table_bytes(outputs)
    table = torch.zeros(P * tb, dtype=torch.uint8, device="cuda")
    logits = torch.zeros(P * outputs, dtype=torch.float32, device="cuda")
    feats = torch.zeros(P * 96, dtype=torch.int16, device="cuda")
    ctx.proposal_update(consts, vis, table, logits, feats)
    torch.cuda.synchronize()
    return dict(ctx=ctx, col=col, K=K, consts=consts, vis=vis, P=P, scene=small_scene, ltc=ltc_table,
                outputs=outputs, table=table.cpu().numpy().reshape(P, tb),
                logits=logits.cpu().numpy().reshape(P, outputs),
                feats=(feats.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32).reshape(P, 96))


def numpy_features(scene, consts, vis_h, pix):
    """The 96 inputs of the MLP per pixel, restated in FP64 numpy from the
    G-buffer inputs (the ray/triangle barycentrics of get_shading_data,
    shading_pass.frag.glsl:571-672, interpolated position and normal, view
    and reflected directions) and the encoding of the proposal kernel:
    position in the scene box q (3), sin/cos(pi 2^j q) for j = 0..6 (42),
    normal, view and reflected directions (9), GGX alpha, n.o, diffuse and
    Fresnel luminances (4), sin/cos(pi 2^j x) of normal and reflection for
    j = 0, 1 (24), sin/cos(pi o) (6), a constant 1 and seven zeros."""
    L = orc.lib()
    pix = np.asarray(pix, np.int64)
    W = int(consts.viewport_size[0])
    M = np.array(consts.pixel_to_ray_direction_world_space, np.float64)[:, :3]
    fx, fy = (pix % W).astype(np.float64), (pix // W).astype(np.float64)
    ray = M[:, 0][None] * fx[:, None] + M[:, 1][None] * fy[:, None] + M[:, 2][None]
    prim = vis_h[pix].astype(np.int64)
    tri = scene.mesh.dequantized().reshape(-1, 3, 3).astype(np.float64)[prim]
    nuv = scene.mesh.normals_uv.reshape(-1, 3, 4)[prim]
    nrm = np.zeros((pix.size, 3, 3))
    for i in range(pix.size):
        for v in range(3):
            L.orc_decode_normal_f64(int(nuv[i, v, 0]), int(nuv[i, v, 1]),
                                    nrm[i, v].ctypes.data_as(orc.f64p))
    org = np.array(consts.camera_position_world_space, np.float64)
    e0, e1 = tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]
    rce1 = np.cross(ray, e1)
    rcp_det = 1.0 / (e0 * rce1).sum(1)
    r0 = org - tri[:, 0]
    by = rcp_det * (r0 * rce1).sum(1)
    bz = -rcp_det * (ray * np.cross(e0, r0)).sum(1)
    bx = 1.0 - (by + bz)
    bc = np.stack([bx, by, bz], 1)[..., None]
    pos = (tri * bc).sum(1)
    n = (nrm * bc).sum(1)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    o = org - pos
    o /= np.linalg.norm(o, axis=1, keepdims=True)
    ndo = (n * o).sum(1)
    n = np.where(ndo[:, None] < 0, -n, n)
    ndo = np.abs(ndo)
    r = 2.0 * ndo[:, None] * n - o
    mats = np.asarray(scene.materials, np.float64)[scene.mesh.material_index[prim].astype(np.int64)]
    metal = mats[:, 5]
    lum = 0.2126 * mats[:, 0] + 0.7152 * mats[:, 1] + 0.0722 * mats[:, 2]
    alum = lum * (1.0 - metal)
    f0l = 0.02 * (1.0 - metal) + lum * metal
    alpha = np.clip(mats[:, 4] ** 2 * float(consts.roughness_factor), 0.0064, 1.0)
    # the scene box of the quantisation grid, in float as the library forms it
    f = np.asarray(scene.mesh.dequantization_factor, np.float32)
    s_ = np.asarray(scene.mesh.dequantization_summand, np.float32)
    lo, hi = s_, s_ + f * np.float32(2097151.0)
    center = (np.float32(0.5) * (hi + lo)).astype(np.float64)
    inv_half = (np.float32(1.0) / (np.float32(0.5) * (hi - lo))).astype(np.float64)
    q = (pos - center) * inv_half
    cols, scales = [q[:, a] for a in range(3)], [0.0] * 3
    for j in range(7):
        for a in range(3):
            cols += [np.sin(np.pi * 2 ** j * q[:, a]), np.cos(np.pi * 2 ** j * q[:, a])]
            scales += [0.5 * 2 ** j] * 2
    for v in (n, o, r):
        cols += [v[:, a] for a in range(3)]
        scales += [0.0] * 3
    cols += [alpha, ndo, alum, f0l]
    scales += [0.0] * 4
    for j in range(2):
        for a in range(3):
            cols += [np.sin(np.pi * 2 ** j * n[:, a]), np.cos(np.pi * 2 ** j * n[:, a]),
                     np.sin(np.pi * 2 ** j * r[:, a]), np.cos(np.pi * 2 ** j * r[:, a])]
            scales += [0.5 * 2 ** j] * 4
    for a in range(3):
        cols += [np.sin(np.pi * o[:, a]), np.cos(np.pi * o[:, a])]
        scales += [0.5] * 2
    cols.append(np.ones(pix.size))
    scales.append(0.0)
    X = np.stack(cols, 1)
    X = np.concatenate([X, np.zeros((pix.size, 96 - X.shape[1]))], 1)
    return X, np.array(scales + [0.0] * (96 - len(scales)))


def test_features_match_numpy_restatement(learned):
    """The device's features (bf16) equal the FP64 numpy restatement rounded
    to bf16, to one bf16 step (both sides round a value that sits near a
    rounding boundary) plus the sin/cos amplification of FP32 position
    rounding; almost all are equal."""
    g = learned
    vis = g["vis"].cpu().numpy().view(np.uint32)
    shaded = np.flatnonzero((vis >> 31) == 0)
    pix = np.random.default_rng(3).choice(shaded, 2000, replace=False)
    X, scale = numpy_features(g["scene"], g["consts"], vis, pix)
    assert X.shape == (pix.size, 96) and np.all(X[:, 89:] == 0)
    got = g["feats"][pix].astype(np.float64)
    want = bf16(X.astype(np.float32)).astype(np.float64)
    mag = np.maximum(np.abs(got), np.abs(want))
    ulp = np.where(mag > 0, 2.0 ** (np.floor(np.log2(np.maximum(mag, 1e-30))) - 7), 0.0)
    tol = ulp + 2 * np.pi * scale[None] * 2e-6 + 1e-6
    d = np.abs(got - want)
    assert np.all(d <= tol), (np.unravel_index(np.argmax(d - tol), d.shape), (d - tol).max())
    # columns without sin/cos amplification are nearly all equal where the
    # value is not a rounding residue near 0 (e.g. the x, y of a flat
    # ground's normal: ~1e-10 in FP32, ~1e-17 in FP64)
    big = mag > 1e-6
    flat = (scale == 0)[None] & big
    assert (d[flat] == 0).mean() > 0.97 and (d[big] == 0).mean() > 0.85, ((d[flat] == 0).mean(),
                                                                          (d[big] == 0).mean())


def table_slots(table, outputs):
    """(ids, weights) of per-pixel tables: pixels x 16 each (ids of
    clusters, or of quads of clusters for more than 256 outputs)."""
    return table[:, :16].astype(np.int64), table[:, 16:32].astype(np.int64)


def test_clusters_cover_every_light(learned):
    col = learned["col"]
    assert col.size == learned["scene"].light_count
    assert set(np.unique(col).tolist()) == set(range(learned["K"]))


def test_mlp_forward_matches_numpy(learned):
    """Logits within 2e-3 of the larger of |z| and 1 on every pixel: the two
    sides differ only in FP32 summation order and, rarely, in the bf16
    rounding of a hidden activation that this order moves across a tie."""
    _check_forward(learned)


def test_wide_mlp_forward_matches_numpy(learned_wide):
    assert learned_wide["ctx"].proposal_param_count() == abi.PROPOSAL_PARAMS_WIDE
    _check_forward(learned_wide)


def _check_forward(learned):
    params, _ = learned["ctx"].proposal_get_params()
    vis = learned["vis"].cpu().numpy().view(np.uint32)
    shaded = np.flatnonzero((vis >> 31) == 0)
    X = learned["feats"][shaded]
    assert np.all(np.isfinite(X)) and np.abs(X).max() <= 1.5
    _, _, _, Z = forward(params, X)
    got = learned["logits"][shaded].astype(np.float64)
    err = np.abs(got - Z) / np.maximum(np.abs(Z), 1.0)
    assert err.max() < 2e-3, err.max()
    assert np.mean(err < 1e-5) > 0.99


def test_tables_are_top4_per_lane_group(learned):
    _check_tables(learned)


def test_wide_tables_are_top4_per_lane_group(learned_wide):
    _check_tables(learned_wide)


def group_values(z, sizes, outputs):
    """Per table group: (value, lane group, live) -- a cluster's logit for
    256 outputs; for 1024, log(sum exp z) over the non-empty clusters of the
    quad 4 q .. 4 q + 3 (the value the proposal kernel ranks)."""
    if outputs <= 256:
        ids = np.arange(outputs)
        return z, (ids % 16) // 4, sizes > 0
    zq = z.reshape(-1, 4)
    live = sizes.reshape(-1, 4) > 0
    m = np.where(live, zq, -np.inf).max(1)
    with np.errstate(invalid="ignore", divide="ignore"):
        v = m + np.log(np.where(live, np.exp(zq - m[:, None]), 0.0).sum(1))
    ids = np.arange(outputs // 4)
    return np.where(live.any(1), v, -np.inf), ids % 4, live.any(1)


def _check_tables(learned):
    """Per lane group, the four largest group values (descending, ties by
    the smaller id) at weight round(255 exp(v - v_max)); value differences
    below 1e-5 may order either way (the kernel's __expf/__logf)."""
    vis = learned["vis"].cpu().numpy().view(np.uint32)
    shaded = np.flatnonzero((vis >> 31) == 0)
    O = learned["outputs"]
    sizes = np.bincount(learned["col"], minlength=O)
    ids, ws = table_slots(learned["table"], O)
    rng = np.random.default_rng(0)
    for p in rng.choice(shaded, 300, replace=False):
        v, lane_group, live = group_values(learned["logits"][p].astype(np.float64), sizes, O)
        vmax = v[live].max()
        for g in range(4):
            cand = np.flatnonzero((lane_group == g) & live)
            best = sorted(cand, key=lambda c: (-v[c], c))[:4]
            for s in range(4):
                c = best[s] if s < len(best) else 0
                e = math.exp(v[c] - vmax) * 255.0 + 0.5
                w = min(255, int(e)) if s < len(best) else 0
                gw = int(ws[p, 4 * g + s])
                # __expf is within a few ulps: the rounding may differ by one near .5
                assert abs(gw - w) <= (1 if abs(e - round(e)) < 1e-3 * e + 1e-3 else 0), (p, g, s, gw, w)
                if gw > 0 and ids[p, 4 * g + s] != c:
                    # a near-tie ordered the other way
                    assert abs(v[ids[p, 4 * g + s]] - v[c]) < 1e-5, (p, g, s, ids[p, 4 * g + s], c)
    assert np.all(ws[(vis >> 31) != 0] == 0)


def test_learned_shading_bitwise_vs_oracle(learned):
    """Exact build with the learned source: candidates, sample records and the
    frame equal the oracle's given the device's tables and clusters."""
    _check_learned_bitwise(learned)


def _check_learned_bitwise(learned):
    import torch
    g = learned
    sc = g["scene"]
    # the tables of the current network (an earlier test may have trained it)
    tb = g["table"].shape[1]
    table = torch.zeros(g["P"] * tb, dtype=torch.uint8, device="cuda")
    g["ctx"].proposal_update(g["consts"], g["vis"], table)
    torch.cuda.synchronize()
    g["table"] = table.cpu().numpy().reshape(g["P"], tb)
    s = sc.settings(exact_math=1, sample_count_light=2, ris_candidate_count=8, proposal=abi.PROPOSAL_LEARNED)
    P = g["P"]
    out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
    rec = torch.zeros(P * 2 * abi.SAMPLE_RECORD_FLOATS, dtype=torch.float32, device="cuda")
    cand = torch.zeros(P * 2 * 8 * abi.CANDIDATE_RECORD_WORDS, dtype=torch.int32, device="cuda")
    g["ctx"].shade_debug(g["consts"], s, g["vis"], out, rec, cand)
    torch.cuda.synchronize()
    oracle = orc.Oracle(sc, *g["ltc"])
    oracle.set_proposal(g["table"], g["col"], g["K"], abi.proposal_group_log2(g["outputs"]))
    vis = g["vis"].cpu().numpy().view(np.uint32)
    ref, rrec, rcand = oracle.shade(g["consts"], s, vis, records=True, candidates=True)
    c = cand.cpu().numpy().view(np.uint32).reshape(rcand.shape)
    assert np.array_equal(c, rcand)
    r = rec.cpu().numpy().reshape(rrec.shape)
    for f in (abi.REC_LIGHT, abi.REC_WSUM, abi.REC_PHAT, abi.REC_W):
        assert np.array_equal(r[..., f].view(np.uint32), rrec[..., f].view(np.uint32)), f
    o = out.cpu().numpy().reshape(P, 4)
    assert np.array_equal(o.view(np.uint32), ref.view(np.uint32))
    # the learned source concentrates candidates on clusters: not all uniform
    assert (table_slots(g["table"], g["outputs"])[1].sum(1) > 0).sum() > 0.3 * P


def test_training_step_gradient_and_adam(learned):
    _check_training_step(learned)


def test_wide_training_step_gradient_and_adam(learned_wide):
    _check_training_step(learned_wide)


@pytest.mark.parametrize("spp", [3, 7])
def test_keyed_records_equal_full_records(learned, spp):
    """The cached learned pass writes the 12-word keyed sample record (light,
    key, occluded estimates as flag bits); with debug records it writes the
    full 16-word one. From the same empty cache both give the same frame and
    the same cache counts, bit for bit (at 7 samples per pixel through the
    resolve that loads samples in groups of 4)."""
    import torch
    g = learned
    sc, ctx, P = g["scene"], g["ctx"], g["P"]
    res = []
    try:
        for debug in (False, True):
            ctx.proposal_cache(9, 0.75, 2.0, True)
            c = ctx.frame_constants(noise_seed=9, animate=True)
            s = sc.settings(sample_count_light=spp, ris_candidate_count=8, proposal=abi.PROPOSAL_LEARNED)
            out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
            if debug:
                rec = torch.zeros(P * spp * abi.SAMPLE_RECORD_FLOATS, dtype=torch.float32, device="cuda")
                ctx.shade_debug(c, s, g["vis"], out, rec)
            else:
                ctx.shade(c, s, g["vis"], out)
            torch.cuda.synchronize()
            res.append((out.cpu().numpy().copy(), ctx.proposal_cache_read().copy()))
    finally:
        ctx.proposal_cache(0)
    (o0, n0), (o1, n1) = res
    assert np.array_equal(o0.view(np.uint32), o1.view(np.uint32))
    assert np.array_equal(n0, n1)
    assert n0.sum() > 0.2 * P


def _check_training_step(learned):
    import torch
    g = learned
    sc = g["scene"]
    ctx = g["ctx"]
    spp = 2
    s = sc.settings(sample_count_light=spp, ris_candidate_count=8)
    P = g["P"]
    out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
    ctx.shade(g["consts"], s, g["vis"], out)
    torch.cuda.synchronize()
    cl, y = ctx.proposal_records()
    cl = cl.cpu().numpy().view(np.uint32)
    y = y.cpu().numpy()
    assert cl.size == P * spp and (cl != 0xFFFFFFFF).sum() > 0.2 * cl.size
    valid = cl != 0xFFFFFFFF
    # background and light pixels have empty records (not stale ones)
    vis_px = g["vis"].cpu().numpy().view(np.uint32)
    assert np.all(cl.reshape(P, spp)[(vis_px >> 31) != 0] == 0xFFFFFFFF)
    assert np.all(cl[valid] < g["K"]) and np.all((y[valid] >= 0) & (y[valid] <= 1))
    assert 0.05 < y[valid].mean() < 0.95
    before, _ = ctx.proposal_get_params()
    B, seed, lr = 4096, 11, 1e-3
    loss = ctx.proposal_train(g["consts"], g["vis"], batch=B, steps=1, lr=lr, seed=seed)
    after, grad = ctx.proposal_get_params()
    # the batch, as the gather kernel draws it
    step_seed = (seed * PHI + 0) & 0xFFFFFFFF
    items = hash32((step_seed * PHI + np.arange(B, dtype=np.uint64)) & 0xFFFFFFFF) % (P * spp)
    pix = (items // spp).astype(np.int64)
    X = g["feats"][pix]
    vis = g["vis"].cpu().numpy().view(np.uint32)
    ok = ((vis[pix] >> 31) == 0) & (cl[items] != 0xFFFFFFFF)
    tgt = np.where(ok, cl[items], 0).astype(np.int64)
    w = ok.astype(np.float64)
    H1, H2, H3, Z = forward(before, X)
    Zs = Z - Z.max(1, keepdims=True)
    p = np.exp(Zs) / np.exp(Zs).sum(1, keepdims=True)
    wb = w / w.sum()
    onehot = np.zeros_like(p)
    onehot[np.arange(B), tgt] = 1.0
    ref_loss = -(wb * np.log(p[np.arange(B), tgt])).sum()
    assert abs(loss - ref_loss) < 1e-3 * max(1.0, abs(ref_loss)), (loss, ref_loss)
    (W1, _), (W2, _), (W3, _), (W4, _) = layers(before)
    G4 = bf16(wb[:, None] * (p - onehot)).astype(np.float64)
    G3 = bf16((G4 @ bf16(W4).astype(np.float64)) * (H3 > 0)).astype(np.float64)
    G2 = bf16((G3 @ bf16(W3).astype(np.float64)) * (H2 > 0)).astype(np.float64)
    G1 = bf16((G2 @ bf16(W2).astype(np.float64)) * (H1 > 0)).astype(np.float64)
    ref = np.concatenate([(G1.T @ X).ravel(), G1.sum(0), (G2.T @ H1).ravel(), G2.sum(0), (G3.T @ H2).ravel(),
                          G3.sum(0), (G4.T @ H3).ravel(), G4.sum(0)])
    width, outputs = W1.shape[0], W4.shape[0]
    for name, (a, b) in zip(("W1", "W2", "W3", "W4"), layer_bounds(width, outputs)):
        rel = np.linalg.norm(grad[a:b] - ref[a:b]) / np.linalg.norm(ref[a:b])
        assert rel < 1e-2, (name, width, rel)
    # Adam, step 1: m = (1-b1) g, v = (1-b2) g^2, bias-corrected
    gd = grad.astype(np.float64)
    mhat = (0.1 * gd) / (1 - 0.9)
    vhat = (0.001 * gd * gd) / (1 - 0.999)
    want = before - lr * mhat / (np.sqrt(vhat) + 1e-8)
    assert np.allclose(after, want, rtol=1e-5, atol=1e-7)


def test_training_lowers_the_loss(learned):
    g = learned
    ctx = g["ctx"]
    first = ctx.proposal_train(g["consts"], g["vis"], batch=8192, steps=1, lr=3e-3, seed=100)
    for i in range(40):
        ctx.proposal_train(g["consts"], g["vis"], batch=8192, steps=1, lr=3e-3, seed=101 + i)
    # the same batch as the first step, after training
    last = ctx.proposal_train(g["consts"], g["vis"], batch=8192, steps=1, lr=0.0, seed=100)
    assert last < 0.9 * first, (first, last)


@pytest.mark.parametrize("per_light,odd,mode", [(False, False, abi.PROPOSAL_LEARNED),
                                                (True, False, abi.PROPOSAL_LEARNED),
                                                (True, True, abi.PROPOSAL_LEARNED),
                                                (True, False, abi.PROPOSAL_CACHE)])
def test_cached_learned_shading_bitwise_vs_oracle(learned, learned_odd, per_light, odd, mode):
    """Exact build, learned pass with a visibility cache of random counts
    (slots per cluster or per light; 64 or 61 lights; learned or uniform
    source): cache keys, candidates, sample records and the frame equal the
    oracle's given the same tables, clusters and cache."""
    import torch
    g = learned_odd if odd else learned
    sc, ctx, P = g["scene"], g["ctx"], g["P"]
    log2, cell, prior = 10, 0.5, 2.0
    rng = np.random.default_rng(77)
    N = sc.light_count
    used = N if per_light else g["K"]
    counts = np.zeros((1 << log2, N if per_light else abi.MAX_LIGHT_CLUSTERS, 2), dtype=np.uint32)
    counts[:, :used] = rng.integers(0, 60, (1 << log2, used, 2))
    ctx.proposal_cache(log2, cell, prior, per_light)
    try:
        s = sc.settings(exact_math=1, sample_count_light=2, ris_candidate_count=8, proposal=mode, shadow_rays=0)
        out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
        rec = torch.zeros(P * 2 * abi.SAMPLE_RECORD_FLOATS, dtype=torch.float32, device="cuda")
        cand = torch.zeros(P * 2 * 8 * abi.CANDIDATE_RECORD_WORDS, dtype=torch.int32, device="cuda")
        # a quarter of the entries empty (the MLP's byte is used there)
        counts[rng.random(counts.shape[:2]) < 0.25] = 0
        ctx.proposal_cache_write(counts)
        ctx.shade_debug(g["consts"], s, g["vis"], out, rec, cand)
        torch.cuda.synchronize()
        # without shadow rays nothing is traced: the counts are unchanged
        assert np.array_equal(ctx.proposal_cache_read(), counts)
        table = ctx.proposal_cache_table()
    finally:
        ctx.proposal_cache(0)
    oracle = orc.Oracle(sc, *g["ltc"])
    oracle.set_proposal(g["table"], g["col"], g["K"], abi.proposal_group_log2(g["outputs"]))
    want = oracle.set_cache(counts, cell, prior, per_light)
    # the refresh before shading made every byte from the written counts
    assert np.array_equal(table, want)
    vis = g["vis"].cpu().numpy().view(np.uint32)
    ref, rrec, rcand = oracle.shade(g["consts"], s, vis, records=True, candidates=True)
    r = rec.cpu().numpy().reshape(rrec.shape)
    assert np.array_equal(r[..., abi.REC_CACHE_KEY].view(np.uint32), rrec[..., abi.REC_CACHE_KEY].view(np.uint32))
    assert np.unique(rrec[..., abi.REC_CACHE_KEY].view(np.uint32)).size > 50
    c = cand.cpu().numpy().view(np.uint32).reshape(rcand.shape)
    assert np.array_equal(c, rcand)
    for f in (abi.REC_LIGHT, abi.REC_WSUM, abi.REC_PHAT, abi.REC_W):
        assert np.array_equal(r[..., f].view(np.uint32), rrec[..., f].view(np.uint32)), f
    o = out.cpu().numpy().reshape(P, 4)
    assert np.array_equal(o.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("spp", [2, 7])
@pytest.mark.parametrize("per_light", [False, True])
def test_resolve_counts_every_traced_ray(learned, per_light, spp):
    """Two frames with shadow rays: the cache holds, per (key, slot), the
    unoccluded and occluded shadow rays of the samples the debug records
    list, summed over both frames. At 7 samples per pixel the resolve loads
    them in groups of 4 (the last one ragged) and merges more than 4 pending
    slots."""
    import torch
    g = learned
    sc, ctx, P = g["scene"], g["ctx"], g["P"]
    ctx.proposal_cache(9, 0.75, 2.0, per_light)
    try:
        want = np.zeros((1 << 9, sc.light_count if per_light else abi.MAX_LIGHT_CLUSTERS, 2), dtype=np.int64)
        for frame in range(2):
            c = ctx.frame_constants(noise_seed=frame + 5, animate=True)
            s = sc.settings(sample_count_light=spp, ris_candidate_count=8,
                            proposal=abi.PROPOSAL_LEARNED if frame else abi.PROPOSAL_UNIFORM)
            out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
            rec = torch.zeros(P * spp * abi.SAMPLE_RECORD_FLOATS, dtype=torch.float32, device="cuda")
            ctx.shade_debug(c, s, g["vis"], out, rec)
            torch.cuda.synchronize()
            r = rec.cpu().numpy().reshape(-1, abi.SAMPLE_RECORD_FLOATS)
            li = r[:, abi.REC_LIGHT].view(np.int32)
            st = r[:, abi.REC_STATE:abi.REC_STATE + 2].view(np.uint32)
            hit = (st == 2).sum(1)
            miss = (st == 3).sum(1)
            ok = (li >= 0) & (hit + miss > 0)
            key = r[ok, abi.REC_CACHE_KEY].view(np.uint32).astype(np.int64)
            cl = (li[ok] if per_light else g["col"][li[ok]]).astype(np.int64)
            np.add.at(want, (key, cl, 0), miss[ok])
            np.add.at(want, (key, cl, 1), hit[ok])
            assert ok.sum() > 0.2 * P
        got = ctx.proposal_cache_read()
        # a third call refreshes the bytes of every cell counted in
        ctx.shade(c, sc.settings(sample_count_light=1, shadow_rays=0), g["vis"], out)
        torch.cuda.synchronize()
        table = ctx.proposal_cache_table()
    finally:
        ctx.proposal_cache(0)
    assert np.array_equal(got.astype(np.int64), want)
    assert 0.05 < want[..., 0].sum() / want.sum() < 0.95
    assert np.array_equal(table, orc.cache_bytes(got, 2.0).reshape(table.shape))
    assert (table != 255).sum() > 100


def test_kmeans_reaches_a_lloyd_fixed_point(small_scene, ltc_table):
    """risltc_cluster_lights after many iterations: every cluster is
    non-empty and every light's nearest cluster mean (the means of the
    returned assignment, over the lights' vertex centroids) is its own
    cluster, the fixed point of Lloyd's iteration; the within-cluster sum of
    squares does not grow with more iterations."""
    import risltc_amd
    ctx = risltc_amd.Context(0)
    try:
        ctx.load_scene(small_scene, *ltc_table)
        cen = np.array([np.asarray(L.world, np.float64)[:L.vertex_count].mean(0) for L in small_scene.lights])
        K = 12

        def wcss(a):
            return sum(((cen[a == k] - cen[a == k].mean(0)) ** 2).sum() for k in range(K))

        a2 = ctx.cluster_lights(K, 2).astype(np.int64)
        a200 = ctx.cluster_lights(K, 200).astype(np.int64)
        assert np.bincount(a200, minlength=K).min() > 0
        means = np.stack([cen[a200 == k].mean(0) for k in range(K)])
        d = ((cen[:, None, :] - means[None]) ** 2).sum(2)
        assert np.array_equal(d.argmin(1), a200)
        assert wcss(a200) <= wcss(a2) + 1e-9
    finally:
        ctx.close()


@pytest.mark.parametrize("exact", [1, 0])
def test_learned_compacted_equals_debug(learned, exact):
    """Learned source with a visibility cache: the non-debug (compacted)
    frame equals the debug (per-lane) frame bit for bit."""
    import torch
    g = learned
    sc, ctx, P = g["scene"], g["ctx"], g["P"]
    rng = np.random.default_rng(5)
    ctx.proposal_cache(8, 0.5, 1.0, True)
    try:
        counts = rng.integers(0, 30, (1 << 8, sc.light_count, 2)).astype(np.uint32)
        ctx.proposal_cache_write(counts)
        s = sc.settings(exact_math=exact, sample_count_light=2, ris_candidate_count=24,
                        proposal=abi.PROPOSAL_LEARNED, shadow_rays=0)
        a = torch.empty(P * 4, dtype=torch.float32, device="cuda")
        b = torch.empty(P * 4, dtype=torch.float32, device="cuda")
        rec = torch.zeros(P * 2 * abi.SAMPLE_RECORD_FLOATS, dtype=torch.float32, device="cuda")
        ctx.shade(g["consts"], s, g["vis"], a)
        ctx.shade_debug(g["consts"], s, g["vis"], b, rec)
        torch.cuda.synchronize()
    finally:
        ctx.proposal_cache(0)
    assert np.array_equal(a.cpu().numpy().view(np.uint32), b.cpu().numpy().view(np.uint32))


def test_wide_learned_pass_runs_and_trains(learned_wide):
    """The learned pass with the wide network: its tables drive the shade
    kernel (bitwise against the oracle is width-independent: the oracle takes
    the device's tables), and training lowers the loss."""
    import torch
    g = learned_wide
    ctx, sc = g["ctx"], g["scene"]
    s = sc.settings(proposal=abi.PROPOSAL_LEARNED, sample_count_light=2, ris_candidate_count=8)
    out = torch.empty(g["P"] * 4, dtype=torch.float32, device="cuda")
    ctx.shade(g["consts"], s, g["vis"], out)
    torch.cuda.synchronize()
    o = out.cpu().numpy().reshape(-1, 4)[:, :3]
    assert np.isfinite(o).all() and (o > 0).mean() > 0.3
    first = ctx.proposal_train(g["consts"], g["vis"], batch=8192, steps=1, lr=3e-3, seed=200)
    for i in range(30):
        ctx.proposal_train(g["consts"], g["vis"], batch=8192, steps=1, lr=3e-3, seed=201 + i)
    last = ctx.proposal_train(g["consts"], g["vis"], batch=8192, steps=1, lr=0.0, seed=200)
    assert last < 0.9 * first, (first, last)
    # switching the width discards the network until it is initialised again
    ctx.proposal_set_width(64)
    import risltc_amd
    with pytest.raises(risltc_amd.RisltcError, match="E_STATE"):
        ctx.proposal_update(g["consts"], g["vis"])
    with pytest.raises(risltc_amd.RisltcError, match="E_UNSUPPORTED"):
        ctx.proposal_set_width(96)


def test_k1024_forward_tables_and_training(learned_k1024):
    """The 1024-output network: logits, tables of quads of clusters (log sum
    exp of the quad's logits) and one training step's gradient against
    numpy."""
    g = learned_k1024
    assert g["ctx"].proposal_param_count() == layer_bounds(128, 1024)[-1][1]
    _check_forward(g)
    _check_tables(g)
    ids, ws = table_slots(g["table"], 1024)
    assert ids[ws > 0].max() < 256 and ids[ws > 0].max() >= 64
    _check_training_step(g)


def test_k1024_clusters_form_compact_quads(learned_k1024):
    """More than 256 clusters come in two levels: each quad of consecutive
    ids is one top-level k-means cluster, so the quads are a Lloyd fixed
    point of the lights' centroids (every light's nearest quad mean is its
    own quad) and every quad has lights."""
    sc, col = learned_k1024["scene"], learned_k1024["col"].astype(np.int64)
    cen = np.array([np.asarray(L.world, np.float64)[:L.vertex_count].mean(0) for L in sc.lights])
    quad = col // 4
    Q = 256
    assert np.array_equal(np.unique(quad), np.arange(Q))
    means = np.stack([cen[quad == q].mean(0) for q in range(Q)])
    d = ((cen[:, None, :] - means[None]) ** 2).sum(2)
    assert (d.argmin(1) == quad).mean() > 0.95


def test_k1024_learned_shading_bitwise_vs_oracle(learned_k1024):
    """The learned pass drawing from quads of clusters: candidates, records
    and frame bitwise against the oracle."""
    g = learned_k1024
    g["ctx"].proposal_update(g["consts"], g["vis"])
    _check_learned_bitwise(g)


@pytest.mark.parametrize("exact", [0, 1])
def test_k1024_compacted_equals_debug(learned_k1024, exact):
    import torch
    g = learned_k1024
    ctx, sc = g["ctx"], g["scene"]
    ctx.proposal_update(g["consts"], g["vis"])
    s = sc.settings(exact_math=exact, sample_count_light=2, ris_candidate_count=16, proposal=abi.PROPOSAL_LEARNED)
    P = g["P"]
    a = torch.empty(P * 4, dtype=torch.float32, device="cuda")
    b = torch.empty(P * 4, dtype=torch.float32, device="cuda")
    rec = torch.zeros(P * 2 * abi.SAMPLE_RECORD_FLOATS, dtype=torch.float32, device="cuda")
    ctx.shade(g["consts"], s, g["vis"], a)
    ctx.shade_debug(g["consts"], s, g["vis"], b, rec)
    torch.cuda.synchronize()
    assert np.array_equal(a.cpu().numpy().view(np.uint32), b.cpu().numpy().view(np.uint32))


def test_k1024_narrow_network_with_dead_outputs(ltc_table):
    """The 96-64-64-64-1024 network over 600 clusters (150 quads; outputs
    past the cluster count are never kept): logits, tables, and learned
    shading bitwise against the oracle."""
    sc = synth.small_scene(seed=8, width=80, height=48, light_count=1200)
    g = _learned_context(sc, ltc_table, width=64, outputs=1024, K=600)
    try:
        _check_forward(g)
        _check_tables(g)
        ids, ws = table_slots(g["table"], 1024)
        assert ids[ws > 0].max() < 150
        _check_learned_bitwise(g)
    finally:
        g["ctx"].close()


@pytest.mark.parametrize("lds_lights", [0, 13])
def test_learned_staging_layouts_bitwise_vs_oracle(learned, lds_lights):
    """The learned pass with its cluster tables and per-light cache rows in
    LDS and no spheres (0) or a part of them (13 of 64) staged before the
    tables: the exact frame equals the oracle's given the same tables,
    clusters and cache."""
    import torch
    g = learned
    sc, ctx, P = g["scene"], g["ctx"], g["P"]
    rng = np.random.default_rng(29)
    log2, cell, prior = 9, 0.5, 2.0
    counts = np.zeros((1 << log2, sc.light_count, 2), dtype=np.uint32)
    counts[:] = rng.integers(0, 40, counts.shape)
    before = ctx.get_tuning().as_dict()
    ctx.proposal_cache(log2, cell, prior, True)
    try:
        ctx.set_tuning(**dict(before, lds_lights=lds_lights))
        ctx.proposal_cache_write(counts)
        s = sc.settings(exact_math=1, sample_count_light=2, ris_candidate_count=12, proposal=abi.PROPOSAL_LEARNED,
                        shadow_rays=0)
        out = torch.empty(P * 4, dtype=torch.float32, device="cuda")
        ctx.shade(g["consts"], s, g["vis"], out)
        torch.cuda.synchronize()
    finally:
        ctx.set_tuning(**before)
        ctx.proposal_cache(0)
    oracle = orc.Oracle(sc, *g["ltc"])
    oracle.set_proposal(g["table"], g["col"], g["K"], abi.proposal_group_log2(g["outputs"]))
    oracle.set_cache(counts, cell, prior, True)
    vis = g["vis"].cpu().numpy().view(np.uint32)
    ref, _, _ = oracle.shade(g["consts"], s, vis)
    o = out.cpu().numpy().reshape(P, 4)
    assert np.array_equal(o.view(np.uint32), ref.view(np.uint32))
    assert (o[:, :3].sum(1) > 0).mean() > 0.2
