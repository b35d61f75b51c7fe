"""Pins for the oracle's mesh decode and LTC texture fetch (CPU only).

decode_position_64_bit / decode_normal_32_bit (mesh_quantization.glsl:19-45)
are checked by round trips through the exporter's encoding, restated here
from tools/io_export_vulkan_blender28.py:24-46 (octahedral normals) and
486-512 (21-bit positions packed into two words). The filtered LTC fetch
(ltc_utility.glsl:56-88 through the sampler of main.c; hazard H5: clamp to
edge, layer = round-half-even of the array coordinate) is checked on tables
whose texels are an affine function of (column, row, layer): bilinear
filtering reproduces an affine function exactly, so any wrong weight, texel
centre, clamp or transposed axis shows up as an error far above rounding.
"""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

import orc
from risltc_amd import abi


def f32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def f64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ------------------------------------------------------------------ positions
def pack_position(q, light=0):
    """Two words of a 21-bit-per-axis position (exporter lines 509-512): x in
    bits 0-20 of word 0, y split as its low 11 bits in word 0 bits 21-31 and
    its high 10 bits in word 1 bits 0-9, z in word 1 bits 10-30; bit 31 of
    word 1 is the light flag the visibility pass reads."""
    qx, qy, qz = (int(v) for v in q)
    w0 = (qx | ((qy & 0x7FF) << 21)) & 0xFFFFFFFF
    w1 = ((qy >> 11) | (qz << 10) | (light << 31)) & 0xFFFFFFFF
    return w0, w1


def is_correctly_rounded_f32(value, exact):
    """value (an np.float32) is exact rounded to nearest, ties to even."""
    v = np.float32(value)
    up = np.nextafter(v, np.float32(np.inf))
    dn = np.nextafter(v, np.float32(-np.inf))
    dv = abs(Fraction(float(v)) - exact)
    du = abs(Fraction(float(up)) - exact)
    dd = abs(Fraction(float(dn)) - exact)
    if dv < du and dv < dd:
        return True
    if dv == du or dv == dd:  # a tie: the even mantissa wins
        return int(v.view(np.uint32)) % 2 == 0
    return False


def test_decode_position_round_trips_the_exporter_packing(orc_lib):
    rng = np.random.default_rng(11)
    for trial in range(400):
        q = rng.integers(0, 2 ** 21, 3)
        if trial < 8:  # the corners of the quantisation grid
            q = np.array([(trial >> a & 1) * (2 ** 21 - 1) for a in range(3)])
        light = int(rng.integers(0, 2))
        lo = rng.uniform(-500, 500, 3)
        ext = rng.uniform(0.5, 2000, 3)
        f = (ext / 2.0 ** 21).astype(np.float32)
        s = (lo + 0.5 * f).astype(np.float32)
        w0, w1 = pack_position(q, light)
        out32 = np.zeros(3, np.float32)
        out64 = np.zeros(3, np.float64)
        orc_lib.orc_decode_position_f32(w0, w1, f32p(f), f32p(s), f32p(out32))
        orc_lib.orc_decode_position_f64(w0, w1, f32p(f), f32p(s), f64p(out64))
        for a in range(3):
            exact = Fraction(int(q[a])) * Fraction(float(f[a])) + Fraction(float(s[a]))
            assert is_correctly_rounded_f32(out32[a], exact), (trial, a, q, out32[a], float(exact))
            # double has room for the exact 21 x 24-bit product, so the only
            # rounding is the final one
            assert out64[a] == float(exact), (trial, a)


def test_decode_position_fields_are_independent(orc_lib):
    """Each axis reads only its own bits: flipping every bit of the other two
    fields and the light flag leaves it unchanged."""
    f = np.ones(3, np.float32)
    s = np.zeros(3, np.float32)
    rng = np.random.default_rng(3)
    for _ in range(100):
        q = rng.integers(0, 2 ** 21, 3)
        out = np.zeros(3, np.float32)
        w0, w1 = pack_position(q, int(rng.integers(0, 2)))
        orc_lib.orc_decode_position_f32(w0, w1, f32p(f), f32p(s), f32p(out))
        assert np.array_equal(out, q.astype(np.float32))
        for a in range(3):
            other = q ^ (2 ** 21 - 1)
            other[a] = q[a]
            w0, w1 = pack_position(other, 1)
            orc_lib.orc_decode_position_f32(w0, w1, f32p(f), f32p(s), f32p(out))
            assert out[a] == np.float32(q[a])


# ------------------------------------------------------------------ normals
def encode_normal(n):
    """Octahedral 2 x 16-bit encoding of unit vectors (exporter lines 24-46):
    divide x, y by the 1-norm; for z <= 0 reflect over the diagonals,
    (1 - |swapped|) * sign (sign of 0 is +); then floor(o * 32767 + 32768.5)."""
    n = np.asarray(n, np.float64)
    o = n[:, :2] / np.abs(n).sum(axis=1, keepdims=True)
    sign = np.where(o >= 0.0, 1.0, -1.0)
    folded = (1.0 - np.abs(o[:, ::-1])) * sign
    o = np.where(n[:, 2:3] <= 0.0, folded, o)
    return np.floor(o * 32767.0 + 32768.5).astype(np.int64)


def decode_normals(orc_lib, q, fp64=False):
    out = np.zeros((len(q), 3), np.float64 if fp64 else np.float32)
    fn = orc_lib.orc_decode_normal_f64 if fp64 else orc_lib.orc_decode_normal_f32
    ptr = f64p if fp64 else f32p
    for i, (a, b) in enumerate(q):
        row = out[i]
        fn(int(a), int(b), ptr(row))
    return out


# One quantisation step of the octahedral map is 1/32767 per coordinate, and
# decode_normal_32_bit scales by 2*65534/65535^2 per step instead of 1/32767
# (a relative 1.5e-5 on the coordinate). The unfolded direction moves by at
# most ~sqrt(2) times the coordinate error, which bounds the angle by
# ~1.1e-4 rad in the upper hemisphere; the fold of the lower one adds the
# error of the reflected coordinate once more (measured maxima 1.4e-4 there).
NORMAL_ANGLE_BOUND = 2.0e-4


def test_decode_normal_round_trips_the_exporter_encoding(orc_lib):
    rng = np.random.default_rng(5)
    n = rng.normal(size=(3000, 3))
    special = [(0, 0, 1), (0, 0, -1), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0),
               (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 1, 1), (-1, -1, -1), (0.3, -0.2, -1e-9)]
    n = np.concatenate([np.asarray(special, np.float64), n])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    q = encode_normal(n)
    assert q.min() >= 1 and q.max() <= 65535
    for fp64 in (False, True):
        d = decode_normals(orc_lib, q, fp64).astype(np.float64)
        assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
        # atan2 of |cross| and dot: arccos near 1 would turn the float32
        # length error (~1e-7) into ~5e-4 rad of apparent angle
        ang = np.arctan2(np.linalg.norm(np.cross(d, n), axis=1), (d * n).sum(axis=1))
        assert ang.max() < NORMAL_ANGLE_BOUND, (fp64, ang.max(), n[ang.argmax()])
    # the hemispheres do not swap, and the fold keeps the signs of x and y
    d = decode_normals(orc_lib, q)
    assert np.all(np.sign(d[np.abs(n) > 1e-3]) == np.sign(n[np.abs(n) > 1e-3]))


def test_decode_normal_centre_is_up(orc_lib):
    """(32768, 32768) is the octahedron's centre, (0, 0, 1) up to the rounding
    of 32768/65535 and of the summand (2.3e-10 in FP32)."""
    d = decode_normals(orc_lib, [(32768, 32768)])[0]
    assert d[2] == np.float32(1.0) and np.abs(d[:2]).max() < 1e-9
    d = decode_normals(orc_lib, [(32768, 32768)], fp64=True)[0]
    assert d[2] == 1.0 and np.abs(d[:2]).max() < 1e-15


def test_decode_normal_f32_tracks_f64(orc_lib):
    rng = np.random.default_rng(9)
    q = rng.integers(1, 65536, size=(2000, 2))
    d32 = decode_normals(orc_lib, q).astype(np.float64)
    d64 = decode_normals(orc_lib, q, fp64=True)
    assert np.abs(d32 - d64).max() < 2e-6


# ------------------------------------------------------------------ LTC fetch
W_TEX, H_TEX, LAYERS = 8, 5, 4  # a non-square table catches transposed axes
# per channel: value = a + b*column + c*row + d*layer, in UNORM16 steps
AFFINE = {
    (0, 0): (1000, 700, 90, 9000),
    (0, 1): (60000, -900, -40, -4000),
    (0, 2): (5, 3, 2000, 100),
    (0, 3): (30000, 1500, -1500, 7),
    (1, 0): (20000, -50, 4000, 3),
    (1, 1): (100, 8000, 10, 1000),
}


def affine_tables():
    i = np.arange(W_TEX)[None, None, :]
    j = np.arange(H_TEX)[None, :, None]
    k = np.arange(LAYERS)[:, None, None]
    rgba = np.zeros((LAYERS, H_TEX, W_TEX, 4), np.uint16)
    rg = np.zeros((LAYERS, H_TEX, W_TEX, 2), np.uint16)
    for (t, ch), (a, b, c, d) in AFFINE.items():
        v = a + b * i + c * j + d * k
        assert v.min() >= 0 and v.max() <= 65535
        (rgba if t == 0 else rg)[..., ch] = v
    return rgba, rg


def affine_scene(rgba, rg):
    sc = orc.OrcScene()
    sc.ltc_rgba16 = rgba.ctypes.data_as(C.POINTER(C.c_uint16))
    sc.ltc_rg16 = rg.ctypes.data_as(C.POINTER(C.c_uint16))
    sc.ltc_res_roughness = W_TEX
    sc.ltc_res_inclination = H_TEX
    sc.ltc_fresnel_count = LAYERS
    return sc


def rne_layer(w):
    """Array layer of coordinate w: round half to even, clamped."""
    r = np.round(w)  # numpy rounds half to even
    return int(min(max(r, 0), LAYERS - 1))


def expected(t, ch, u, v, w):
    a, b, c, d = AFFINE[(t, ch)]
    x = min(max(u * W_TEX - 0.5, 0.0), W_TEX - 1.0)
    y = min(max(v * H_TEX - 0.5, 0.0), H_TEX - 1.0)
    return (a + b * x + c * y + d * rne_layer(w)) / 65535.0


def test_ltc_fetch_reproduces_an_affine_table(orc_lib):
    rgba, rg = affine_tables()
    sc = affine_scene(rgba, rg)
    rng = np.random.default_rng(2)
    uv = rng.uniform(-0.3, 1.3, size=(600, 2))
    ws = rng.uniform(-1.0, LAYERS + 0.5, size=600)
    err64 = err32 = 0.0
    for (u, v), w in zip(uv, ws):
        u32, v32, w32 = np.float32(u), np.float32(v), np.float32(w)
        for (t, ch) in AFFINE:
            e64 = expected(t, ch, float(u), float(v), float(w))
            g64 = orc_lib.orc_ltc_fetch_f64(C.byref(sc), t, ch, float(u), float(v), float(w))
            err64 = max(err64, abs(g64 - e64))
            e32 = expected(t, ch, float(u32), float(v32), float(w32))
            g32 = orc_lib.orc_ltc_fetch_f32(C.byref(sc), t, ch, float(u32), float(v32), float(w32))
            err32 = max(err32, abs(g32 - e32))
    assert err64 < 1e-12, err64
    # FP32: texels are rounded to float (0.5 ulp each), the filter adds two
    # fused steps, and u*W - 0.5 is rounded before the weights: a few ulp of 1
    assert err32 < 4 * 2.0 ** -23, err32


def test_ltc_fetch_clamps_to_the_edge(orc_lib):
    rgba, rg = affine_tables()
    sc = affine_scene(rgba, rg)
    for t, ch in AFFINE:
        a, b, c, d = AFFINE[(t, ch)]
        # left/top of the first texel centre and right/bottom of the last:
        # exactly the corner texel
        for u, v, i, j in [(0.0, 0.0, 0, 0), (-5.0, 0.01, 0, 0), (1.0, 1.0, W_TEX - 1, H_TEX - 1),
                           (7.0, 2.0, W_TEX - 1, H_TEX - 1), (0.5 / W_TEX, 0.5 / H_TEX, 0, 0)]:
            g = orc_lib.orc_ltc_fetch_f64(C.byref(sc), t, ch, u, v, 2.0)
            assert g == (a + b * i + c * j + d * 2) / 65535.0, (t, ch, u, v)


def test_ltc_fetch_layer_is_round_half_even(orc_lib):
    rgba, rg = affine_tables()
    sc = affine_scene(rgba, rg)
    a, b, c, d = AFFINE[(0, 0)]
    base = a + b * 3 + c * 2  # the texel centre (3, 2)
    u, v = 3.5 / W_TEX, 2.5 / H_TEX
    cases = {-0.7: 0, 0.49: 0, 0.5: 0, 0.51: 1, 1.5: 2, 2.5: 2, 2.51: 3, 3.5: 3, 9.0: 3}
    for w, layer in cases.items():
        for fn in (orc_lib.orc_ltc_fetch_f64, orc_lib.orc_ltc_fetch_f32):
            g = fn(C.byref(sc), 0, 0, u, v, w)
            assert abs(g - (base + d * layer) / 65535.0) < 1e-6, (w, layer, g)


def test_ltc_coefficients_sample_the_table_at_roughness_inclination_fresnel(orc_lib):
    """get_ltc_coefficients reads the albedo channel (rg16 channel 1) at
    u = sqrt(roughness)*rf + rs, v = acos(n.o)*if + is, layer = RNE(F0*ff + fs),
    with the constants that load_ltc_table derives for the table's size."""
    rgba, rg = affine_tables()
    sc = affine_scene(rgba, rg)
    k = abi.LtcConstants()
    orc_lib.orc_ltc_constants(C.byref(k), W_TEX, H_TEX, LAYERS)
    rng = np.random.default_rng(8)
    out = np.zeros(32, np.float32)
    nrm = np.array([0, 0, 1], np.float32)
    pos = np.zeros(3, np.float32)
    for _ in range(200):
        rough = float(rng.uniform(0.0, 1.0))
        f0 = float(rng.uniform(0.0, 1.0))
        theta = float(rng.uniform(0.0, 0.5 * math.pi))
        o = np.array([math.sin(theta), 0.0, math.cos(theta)], np.float32)
        w = f0 * k.fresnel_index_factor + k.fresnel_index_summand
        if abs(w - math.floor(w) - 0.5) < 1e-3:
            continue  # a rounding boundary: covered by the fetch test
        orc_lib.orc_ltc_coefficients_f32(C.byref(sc), C.byref(k), f0, rough, f32p(pos), f32p(nrm), f32p(o),
                                         f32p(out))
        u = math.sqrt(rough) * k.roughness_factor + k.roughness_summand
        incl = math.acos(min(max(float(o[2]), 0.0), 1.0))
        v = incl * k.inclination_factor + k.inclination_summand
        assert abs(out[30] - expected(1, 1, u, v, w)) < 2e-5, (rough, f0, theta)
