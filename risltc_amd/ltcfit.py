"""Project-generated LTC table for the GGX BRDF of brdfs.glsl:58-93.

The reference loads a fitted table from data/ggx_ltc_fit/fit<i>.dat
(src/main.c:2542, src/ltc_table.c:23-194) that is downloaded separately and
not available here. This module fits its own table in the same format:

  fit<i>.dat = uint64 resolution, then resolution^2 entries of 5 floats
               (d0, d1, d2, d3, albedo); entry j = inclination * res +
               roughness (roughness fastest, the row-major image order that
               load_ltc_table copies into the texture array).

The loader turns (d0..d3) into the matrix [[d0,0,d1],[0,d2,0],[d3,0,1]],
stores its adjugate and the shader reads that back transposed as the
shading-to-cosine transform (ltc_utility.glsl:63-72). So the file holds
M^T, where M maps the clamped cosine to the BRDF lobe in shading space
(x towards the viewer, z the normal), scaled so that M[2][2] = 1.

Grid (the texture coordinates of ltc_utility.glsl:63-66 and the constants of
ltc_table.c:184-191): roughness index i <-> sqrt(alpha) = i / (res - 1),
inclination index j <-> theta = (pi / 2) * j / (res - 1), layer k <-> F0 =
k / (fresnel_count - 1).

Fit: for every cell the 4 free entries of M minimise the symmetric error
integral of (D_ltc - f)^2 / (D_ltc + f) between the LTC density and the
normalised cosine-weighted specular lobe f (Fresnel = 1), estimated with
fixed GGX-importance samples and reparametrised LTC samples, by Adam over
all cells at once. The albedo of layer k is F0_k * I1 + (1 - F0_k) * I5 with
I1 = integral of the lobe and I5 = integral of the lobe times (1 - o.h)^5,
which is exact for the grey Schlick Fresnel of brdfs.glsl:87-90 (its
luminance weights sum to 1). The table only changes the variance of the
estimator: the final sample is weighted with the true BRDF
(shading_pass.frag.glsl:383).
"""
import math
import os
import struct

import numpy as np

DATA = os.path.join(os.path.dirname(__file__), "data")
FIT_FILE = os.path.join(DATA, "ltc_fit64.npz")


def _fit(res=64, steps=300, n_ggx=256, n_ltc=256, seed=0, lr=4e-3):
    import torch
    g = torch.Generator().manual_seed(seed)
    dt = torch.float64
    i = torch.arange(res, dtype=dt)
    alpha = torch.clamp((i / (res - 1)) ** 2, min=1.0e-3)
    theta = torch.clamp(0.5 * math.pi * i / (res - 1), max=0.5 * math.pi - 0.01)
    A, T = torch.meshgrid(alpha, theta, indexing="xy")  # A[j, i] = alpha_i, T[j, i] = theta_j
    A = A.reshape(-1, 1)
    T = T.reshape(-1, 1)
    cells = A.shape[0]
    wo = torch.stack([torch.sin(T), torch.zeros_like(T), torch.cos(T)], -1)  # (C,1,3)
    r2 = A * A

    def lobe(w):
        """cosine-weighted specular BRDF of brdfs.glsl:77-92 with F = 1"""
        li = w[..., 2]
        lo = wo[..., 2]
        h = w + wo
        h = h / torch.linalg.norm(h, dim=-1, keepdim=True)
        ndh = h[..., 2]
        den = (ndh * r2 - ndh) * ndh + 1.0
        ggx = r2 / (den * den)
        lic = torch.clamp(li, min=1e-9)
        mask = lic * torch.sqrt((-lo * r2 + lo) * lo + r2)
        shad = lo * torch.sqrt((-lic * r2 + lic) * lic + r2)
        val = ggx * 0.5 / (mask + shad) * lic / math.pi
        return torch.where(li > 0, val, torch.zeros_like(val)), h

    # GGX importance samples of the half vector (pdf = D(h) cos(theta_h))
    u1 = torch.rand(cells, n_ggx, generator=g, dtype=dt)
    u2 = torch.rand(cells, n_ggx, generator=g, dtype=dt)
    t2 = r2 * u1 / (1.0 - u1)
    ct = 1.0 / torch.sqrt(1.0 + t2)
    st = torch.sqrt(torch.clamp(1.0 - ct * ct, min=0.0))
    ph = 2.0 * math.pi * u2
    h = torch.stack([st * torch.cos(ph), st * torch.sin(ph), ct], -1)
    odh = (wo * h).sum(-1, keepdim=True)
    wi = 2.0 * odh * h - wo
    d_true = r2 / (math.pi * ((r2 - 1.0) * ct * ct + 1.0) ** 2)
    pdf_g = d_true * ct / (4.0 * torch.clamp(odh[..., 0].abs(), min=1e-12))
    f_g, hh = lobe(wi)
    w_g = torch.where(pdf_g > 0, f_g / pdf_g, torch.zeros_like(f_g))
    I1 = w_g.mean(1)
    c = torch.clamp((wo * hh).sum(-1), 0.0, 1.0)
    I5 = (w_g * (1.0 - c) ** 5).mean(1)
    fn_g = f_g / I1[:, None]

    # cosine samples for the reparametrised LTC samples
    v1 = torch.rand(cells, n_ltc, generator=g, dtype=dt)
    v2 = torch.rand(cells, n_ltc, generator=g, dtype=dt)
    rr = torch.sqrt(v1)
    wc = torch.stack([rr * torch.cos(2 * math.pi * v2), rr * torch.sin(2 * math.pi * v2),
                      torch.sqrt(torch.clamp(1.0 - v1, min=0.0))], -1)

    # moment initialisation: lobe mean direction and width
    mean = (w_g[..., None] * wi).mean(1)
    phi = torch.atan2(mean[:, 0], mean[:, 2])
    width = torch.clamp(A[:, 0], min=1e-3)
    p = torch.stack([width * torch.cos(phi), torch.sin(phi), width, -width * torch.sin(phi)], -1)
    p = p / torch.cos(phi)[:, None]
    p = p.clone().requires_grad_(True)
    opt = torch.optim.Adam([p], lr=lr)

    def matrix(p):
        M = torch.zeros(cells, 3, 3, dtype=dt)
        M[:, 0, 0] = p[:, 0]
        M[:, 0, 2] = p[:, 1]
        M[:, 1, 1] = p[:, 2]
        M[:, 2, 0] = p[:, 3]
        M[:, 2, 2] = 1.0
        return M

    def ltc_density(Minv, detinv, w):
        x = torch.einsum("cij,cnj->cni", Minv, w)
        l = torch.linalg.norm(x, dim=-1)
        z = torch.clamp(x[..., 2] / l, min=0.0)
        return z / math.pi * detinv[:, None].abs() / (l ** 3)

    def pdf_ggx(w):
        """density of the GGX-importance samples above, at direction w"""
        hh = w + wo
        hh = hh / torch.linalg.norm(hh, dim=-1, keepdim=True)
        cth = hh[..., 2]
        dtr = r2 / (math.pi * ((r2 - 1.0) * cth * cth + 1.0) ** 2)
        o = torch.clamp((wo * hh).sum(-1).abs(), min=1e-12)
        return torch.where(cth > 0, dtr * cth / (4.0 * o), torch.zeros_like(cth))

    for step in range(steps):
        opt.zero_grad()
        M = matrix(p)
        Minv = torch.linalg.inv(M)
        detinv = torch.linalg.det(Minv)
        wl = torch.einsum("cij,cnj->cni", M, wc)
        wl = wl / torch.linalg.norm(wl, dim=-1, keepdim=True)
        f_l, _ = lobe(wl)
        fn_l = f_l / I1[:, None]
        d_l = ltc_density(Minv, detinv, wl)
        d_g = ltc_density(Minv, detinv, wi)
        # balance-heuristic mixture estimate of the integral of (D - f)^2 / (D + f)
        e_g = (d_g - fn_g) ** 2 / (d_g + fn_g + 1e-12) / (0.5 * pdf_g + 0.5 * d_g + 1e-12)
        e_l = (d_l - fn_l) ** 2 / (d_l + fn_l + 1e-12) / (0.5 * pdf_ggx(wl) + 0.5 * d_l + 1e-12)
        err = 0.5 * e_g.mean(1) + 0.5 * e_l.mean(1)
        loss = err.sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            p[:, 0].clamp_(min=1e-4)
            p[:, 2].clamp_(min=1e-4)
    fit = np.zeros((res * res, 5), dtype=np.float32)
    P = p.detach().numpy()
    # file entries hold M^T: d0 = M00, d1 = M20, d2 = M11, d3 = M02
    fit[:, 0] = P[:, 0]
    fit[:, 1] = P[:, 3]
    fit[:, 2] = P[:, 2]
    fit[:, 3] = P[:, 1]
    fit[:, 4] = I1.numpy()
    return fit, I5.numpy().astype(np.float32), err.detach().numpy().astype(np.float32)


def load_fit():
    """(fit (res*res, 5), i5 (res*res,)) from the committed compact table."""
    z = np.load(FIT_FILE)
    return z["fit"], z["i5"]


def generate(res=64):
    fit, i5, err = _fit(res)
    os.makedirs(DATA, exist_ok=True)
    np.savez_compressed(FIT_FILE, fit=fit, i5=i5, err=err, res=np.array([res]))
    return fit, i5


def layer_fit(fit, i5, k, fresnel_count=51):
    """fit<k>.dat contents: the matrix is shared, albedo follows F0 = k / (count - 1)."""
    f0 = k / (fresnel_count - 1)
    out = fit.copy()
    out[:, 4] = f0 * fit[:, 4] + (1.0 - f0) * i5
    return out


def write_fit_dir(directory, fresnel_count=51, fit=None, i5=None):
    """Writes fit0.dat .. fit<count-1>.dat in the format load_ltc_table reads."""
    if fit is None:
        fit, i5 = load_fit()
    res = int(round(math.sqrt(fit.shape[0])))
    os.makedirs(directory, exist_ok=True)
    for k in range(fresnel_count):
        with open(os.path.join(directory, "fit%u.dat" % k), "wb") as fh:
            fh.write(struct.pack("<Q", res))
            fh.write(layer_fit(fit, i5, k, fresnel_count).astype("<f4").tobytes())
    return res


def quantize(fit, i5, fresnel_count=51):
    """The two UNORM16 texture arrays of load_ltc_table (src/ltc_table.c:82-116),
    vectorised: rgba16 (count, res, res, 4), rg16 (count, res, res, 2)."""
    res = int(round(math.sqrt(fit.shape[0])))
    rgba = np.zeros((fresnel_count, res * res, 4), dtype=np.uint16)
    rg = np.zeros((fresnel_count, res * res, 2), dtype=np.uint16)
    for k in range(fresnel_count):
        d = layer_fit(fit, i5, k, fresnel_count).astype(np.float32)
        z = np.float32(0.0)
        adj = np.stack([np.stack([d[:, 2], np.full_like(d[:, 2], z), -d[:, 1] * d[:, 2]], 1),
                        np.stack([np.full_like(d[:, 2], z), d[:, 0] - d[:, 1] * d[:, 3], np.full_like(d[:, 2], z)], 1),
                        np.stack([-d[:, 2] * d[:, 3], np.full_like(d[:, 2], z), d[:, 0] * d[:, 2]], 1)], 1)
        largest = np.abs(adj).reshape(-1, 9).max(1).astype(np.float32)
        adj = (adj / largest[:, None, None]).astype(np.float32)
        vals = np.stack([adj[:, 0, 0], -adj[:, 0, 2], adj[:, 1, 1], adj[:, 2, 0], adj[:, 2, 2], d[:, 4]], 1)
        vals = np.clip(vals, 0.0, 1.0).astype(np.float32)
        q = (vals * np.float32(65535.0) + np.float32(0.5)).astype(np.float32).astype(np.uint16)
        rgba[k] = q[:, 0:4]
        rg[k] = q[:, 4:6]
    return rgba.reshape(fresnel_count, res, res, 4), rg.reshape(fresnel_count, res, res, 2)


def table(fresnel_count=51):
    """(rgba16, rg16, res) ready for risltc_set_ltc_table."""
    fit, i5 = load_fit()
    rgba, rg = quantize(fit, i5, fresnel_count)
    return rgba, rg, rgba.shape[1]


if __name__ == "__main__":
    import time
    t = time.time()
    fit, i5 = generate()
    print("fitted", fit.shape, "in %.1f s" % (time.time() - t))
