"""Pins for the CPU oracle (oracle/): every function is checked against the
reference's own table or build, against a library routine, or against a
closed form that a plausible mistake (wrong sign, dropped term, transposed
operand, wrong index) would break. CPU only."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import orc
from risltc_amd import abi, synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def f32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def f64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------- clipping rule
def test_clip_rule_matches_every_reference_case(orc_lib):
    """polygon_clipping.glsl:45-216, every case of every MAX/MIN variant."""
    data = json.load(open(os.path.join(GOLDEN, "clip_cases.json")))
    out = (C.c_int * 10)()
    checked = 0
    for variant, cases in data.items():
        max_count = int(variant.split("_")[0][3:])
        for case in cases.values():
            n, mask = case["n"], case["inside_mask"]
            vc = orc_lib.orc_clip_rule(n, mask, out)
            assert vc == case["count"], (variant, n, mask)
            for j, slot in enumerate(case["slots"]):
                if j >= max_count:
                    break
                code = out[j]
                if slot[0] == "v":
                    assert code == slot[1], (variant, n, mask, j)
                else:
                    a, b = slot[1], slot[2]
                    assert b == (a + 1) % n, "reference intersections follow edge order"
                    assert code == -(a + 1), (variant, n, mask, j)
            checked += 1
    assert checked == 1276


def test_clip_rule_extends_to_octagons(orc_lib):
    """8-gons (hazard H1): every mask gives a convex run with correct counts,
    the entry/exit intersections on the right edges, and vertex 0 padding."""
    out = (C.c_int * 10)()
    for mask in range(256):
        vc = orc_lib.orc_clip_rule(8, mask, out)
        k = bin(mask).count("1")
        runs = sum(1 for i in range(8) if (mask >> i) & 1 and not (mask >> ((i - 1) % 8)) & 1)
        if runs > 1:
            continue  # not a convex clip configuration
        expect = 0 if k == 0 else (8 if k == 8 else k + 2)
        assert vc == expect
        if 0 < k < 8:
            codes = [out[j] for j in range(vc)]
            inside = [c for c in codes if c >= 0]
            assert sorted(inside) == [i for i in range(8) if (mask >> i) & 1]
            assert sum(1 for c in codes if c < 0) == 2


def test_clip_polygon_geometry_matches_sutherland_hodgman(orc_lib):
    """Clipped polygon = brute-force Sutherland-Hodgman against z > 0 (same
    vertices as a cyclic sequence), FP64, random convex polygons."""
    rng = np.random.default_rng(3)
    for trial in range(400):
        n = int(rng.integers(3, 9))
        pts = synth.random_convex_polygon(rng, n)
        # embed in a random plane tilted through the horizon
        centre = rng.normal(size=3) * 0.4
        u = rng.normal(size=3); u /= np.linalg.norm(u)
        w = rng.normal(size=3); w -= u * (u @ w); w /= np.linalg.norm(w)
        P = centre + pts[:, :1] * u + pts[:, 1:2] * w
        v = np.zeros((9, 3))
        v[:n] = P
        v[n:] = P[0]
        vc = orc_lib.orc_clip_polygon_f64(n, f64p(v))
        # brute force
        out = []
        for i in range(n):
            a, b = P[i], P[(i + 1) % n]
            ia, ib = a[2] > 0, b[2] > 0
            if ia:
                out.append(a)
            if ia != ib:
                t = a[2] / (a[2] - b[2])
                out.append(a + t * (b - a))
        assert vc == (len(out) if len(out) >= 3 else 0)
        if vc:
            got = v[:vc]
            # same cyclic sequence up to rotation
            k = int(np.argmin(np.linalg.norm(got - out[0], axis=1)))
            rolled = np.roll(got, -k, axis=0)
            assert np.allclose(rolled, np.array(out), atol=1e-9)
            assert np.allclose(v[vc:], v[0])


# ---------------------------------------------------------------- host code vs reference build
def _light(rng, n):
    pts = synth.random_convex_polygon(rng, n) * rng.uniform(0.2, 2.0)
    L = synth.Light(pts, rng.uniform(0, 2 * math.pi, 3), rng.uniform(-5, 5, 3), rng.uniform(0.3, 2, 2),
                    rng.uniform(0.5, 10, 3))
    return L


@pytest.mark.skipif(not os.path.exists(orc.REF_LIB_PATH) and not os.path.exists("/root/reference/src"),
                    reason="reference host sources not available")
def test_update_polygonal_light_bitwise_vs_reference_build(orc_lib):
    """src/polygonal_light.c:44-98 compiled from the reference tree."""
    orc.build(reference="/root/reference")
    ref = orc.ref_lib()
    assert ref is not None
    rng = np.random.default_rng(5)
    for i in range(200):
        L = _light(rng, int(rng.integers(3, 9)))
        a, b = L.to_ctypes(), L.to_ctypes()
        keep = (L._ps, L._ws)
        L2 = synth.Light(L.plane_vertices, L.angles, L.translation, L.scaling, L.radiance)
        b = L2.to_ctypes()
        ref.update_polygonal_light(C.byref(a))
        orc_lib.orc_update_polygonal_light(C.byref(b))
        fa = np.ctypeslib.as_array(a.vertices_world_space, shape=(4 * L.vertex_count,))
        fb = np.ctypeslib.as_array(b.vertices_world_space, shape=(4 * L.vertex_count,))
        assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
        for field in ("plane", "rotation", "surface_radiance"):
            x = np.ctypeslib.as_array(getattr(a, field)).ravel()
            y = np.ctypeslib.as_array(getattr(b, field)).ravel()
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), field
        assert np.float32(a.area).view(np.uint32) == np.float32(b.area).view(np.uint32)
        del keep


def test_update_polygonal_light_closed_form(orc_lib):
    """rotation orthonormal, plane.xyz unit, plane . v = 0 for every vertex,
    area = shoelace * sx * sy."""
    rng = np.random.default_rng(6)
    for i in range(100):
        L = _light(rng, int(rng.integers(3, 9)))
        cl = L.to_ctypes()
        orc_lib.orc_update_polygonal_light(C.byref(cl))
        R = np.ctypeslib.as_array(cl.rotation)[:, :3].astype(np.float64)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)
        plane = np.ctypeslib.as_array(cl.plane).astype(np.float64)
        ws = np.ctypeslib.as_array(cl.vertices_world_space, shape=(4 * L.vertex_count,)).reshape(-1, 4)[:, :3]
        assert np.allclose(ws @ plane[:3] + plane[3], 0.0, atol=1e-5)
        area = synth.polygon_area(L.plane_vertices.astype(np.float64)) * L.scaling[0] * L.scaling[1]
        assert abs(cl.area - area) < 1e-5 * max(1.0, area)


def test_math_utilities_vs_reference_build(orc_lib):
    """matrix_inverse and wang_random_number (src/math_utilities.h:24-57)."""
    orc.build(reference="/root/reference")
    ref = orc.ref_lib()
    if ref is None:
        pytest.skip("reference tree not available")
    rng = np.random.default_rng(7)
    for i in range(100):
        m = rng.normal(size=(4, 4)).astype(np.float32)
        a = np.zeros((4, 4), np.float32)
        b = np.zeros((4, 4), np.float32)
        ref.ref_matrix_inverse(f32p(a), f32p(m))
        orc_lib.orc_matrix_inverse(f32p(b), f32p(m))
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for s in rng.integers(0, 2 ** 32, 1000, dtype=np.uint64):
        assert ref.ref_wang_random_number(int(s)) == orc_lib.orc_wang_random_number(int(s))


# ---------------------------------------------------------------- noise
def test_murmur_is_murmurhash3_x86_32(orc_lib):
    """mix + finalize with seed 0 on one 4-byte block == MurmurHash3_x86_32 of
    that block (published algorithm; known answers computed independently)."""
    def mm3(key, seed):
        c1, c2 = 0xcc9e2d51, 0x1b873593
        k = (key * c1) & 0xFFFFFFFF
        k = ((k << 15) | (k >> 17)) & 0xFFFFFFFF
        k = (k * c2) & 0xFFFFFFFF
        h = seed ^ k
        h = ((h << 13) | (h >> 19)) & 0xFFFFFFFF
        h = (h * 5 + 0xe6546b64) & 0xFFFFFFFF
        h ^= 4
        h ^= h >> 16
        h = (h * 0x85ebca6b) & 0xFFFFFFFF
        h ^= h >> 13
        h = (h * 0xc2b2ae35) & 0xFFFFFFFF
        h ^= h >> 16
        return h
    # published test vector: MurmurHash3_x86_32 of 4 zero bytes with seed 0
    assert mm3(0, 0) == 0x2362F9DE
    for key in (0, 1, 0xdeadbeef, 123456789):
        h = orc_lib.orc_murmur_mix(0, key)
        # the shader omits the length xor; restore it to compare with the published hash
        assert orc_lib.orc_murmur_finalize(h ^ 4) == mm3(key, 0)


def test_lcg_closed_form(orc_lib):
    """s_k = a^k s_0 + c (a^k - 1)/(a - 1) mod 2^32 and u = float(s) * 2^-32."""
    a, c = 1664525, 1013904223
    seed = C.c_uint32(12345)
    s0 = 12345
    for k in range(1, 200):
        u = orc_lib.orc_noise_next(C.byref(seed))
        ak = pow(a, k, 2 ** 32 * (a - 1))
        sk = (ak * s0 + c * (ak - 1) // (a - 1)) % 2 ** 32
        assert seed.value == sk
        assert np.float32(u) == np.float32(np.float32(sk) * np.float32(2.0 ** -32))


def test_noise_can_return_one_and_index_is_clamped():
    """hazard H2: float(seed) * 2^-32 rounds to 1.0 for seeds >= 0xFFFFFF80."""
    assert np.float32(np.float32(0xFFFFFF80) * np.float32(2.0 ** -32)) == 1.0
    assert np.float32(np.float32(0xFFFFFF7F) * np.float32(2.0 ** -32)) < 1.0


# ---------------------------------------------------------------- transcendentals
@pytest.mark.parametrize("name,ref,lo,hi,ulps", [
    ("prj_atanf", math.atan, -50.0, 50.0, 3),
    ("prj_acosf", math.acos, -1.0, 1.0, 4),
    ("prj_asinf", math.asin, -1.0, 1.0, 3),
    ("prj_sinf", math.sin, -8.0, 8.0, 3),
    ("prj_cosf", math.cos, -8.0, 8.0, 3),
])
def test_project_transcendentals_vs_libm(orc_lib, name, ref, lo, hi, ulps):
    fn = getattr(orc_lib, name)
    xs = np.concatenate([np.linspace(lo, hi, 20001, dtype=np.float32),
                         np.random.default_rng(1).uniform(lo, hi, 20000).astype(np.float32)])
    worst = 0.0
    for x in xs:
        y = float(np.float32(fn(float(x))))
        exact = ref(float(x))
        ulp = float(np.spacing(np.float32(abs(exact))))
        err = abs(y - exact)
        # relative bound in ulps of the result; an absolute floor of 2^-26
        # near the zeros of sin/cos, where the pi/4 reduction dominates
        if err > ulps * ulp and err > 2.0 ** -26:
            worst = max(worst, err / ulp)
    assert worst == 0.0, (name, worst)


# ---------------------------------------------------------------- LTC integral and PSA
def _spherical_psa(P):
    """Projected solid angle of a polygon above the horizon (closed form):
    1/2 sum_i theta_i n_i.z, n_i the unit normal of edge i's great circle."""
    d = P / np.linalg.norm(P, axis=1, keepdims=True)
    s = 0.0
    for i in range(len(d)):
        a, b = d[i], d[(i + 1) % len(d)]
        cr = np.cross(a, b)
        theta = math.atan2(np.linalg.norm(cr), float(a @ b))
        s += theta * cr[2] / np.linalg.norm(cr)
    return 0.5 * abs(s)


def _random_polygon_above(rng, n, clip=False):
    pts = synth.random_convex_polygon(rng, n)
    centre = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.5, 2.0) if not clip else
                       rng.uniform(-0.3, 0.5)])
    u = rng.normal(size=3); u /= np.linalg.norm(u)
    w = rng.normal(size=3); w -= u * (u @ w); w /= np.linalg.norm(w)
    scale = rng.uniform(0.1, 0.6) if not clip else rng.uniform(0.5, 2.0)
    P = centre + scale * (pts[:, :1] * u + pts[:, 1:2] * w)
    if np.cross(P[1] - P[0], P[2] - P[0]) @ P[0] < 0:  # make it clockwise as seen from the origin
        P = P[::-1].copy()
    return P


def test_ltc_integral_is_twice_the_projected_solid_angle(orc_lib):
    """calculate_ltc sums cross(a,b).z * theta / (2 pi sin(theta)) per edge
    (a rational fit of that factor: 1/(2 pi) at theta = 0), i.e. the form
    factor PSA / pi; the fit error bounds the difference."""
    rng = np.random.default_rng(11)
    worst = 0.0
    for i in range(300):
        P = _random_polygon_above(rng, int(rng.integers(3, 9)))
        if (P[:, 2] <= 0).any():
            continue
        v = np.zeros((9, 3)); v[:len(P)] = P
        got = orc_lib.orc_calculate_ltc_f64(len(P), f64p(v))
        exact = _spherical_psa(P) / math.pi
        worst = max(worst, abs(got - exact) / exact)
    assert worst < 3e-3, worst
    # the rational fit at theta -> 0 and theta = pi/2
    a0 = (0.8543985 + 0.4965155 + 0.0145206) / (3.4175940 + 4.1616724 + 1.0)
    assert abs(a0 - 1.0 / (2.0 * math.pi)) < 1e-6


def test_prepare_psa_equals_closed_form(orc_lib):
    """prepare_projected_solid_angle_polygon_sampling total == closed form,
    FP64 to 1e-9 and FP32 to 1e-4 relative; covers central and decentral cases."""
    rng = np.random.default_rng(12)
    central = 0
    for i in range(400):
        P = _random_polygon_above(rng, int(rng.integers(3, 10)) if False else int(rng.integers(3, 9)))
        if (P[:, 2] <= 0).any():
            continue
        v = np.zeros((9, 3)); v[:len(P)] = P; v[len(P):] = P[0]
        state = np.zeros(48)
        got = orc_lib.orc_prepare_psa_f64(len(P), f64p(v), f64p(state))
        central += state[36] > 0
        exact = _spherical_psa(P)
        assert abs(got - exact) <= 1e-9 * exact + 1e-12, (i, got, exact)
        v32 = v.astype(np.float32)
        g32 = orc_lib.orc_prepare_psa_f32(len(P), f32p(v32), None)
        assert abs(g32 - exact) <= 1e-4 * exact + 1e-7
    # polygons that contain the zenith (clockwise as seen from the origin)
    P = np.array([[1, 1, 1], [-1, 1, 1], [-1, -1, 1], [1, -1, 1]], dtype=np.float64)
    v = np.zeros((9, 3)); v[:4] = P; v[4:] = P[0]
    state = np.zeros(48)
    got = orc_lib.orc_prepare_psa_f64(4, f64p(v), f64p(state))
    assert state[36] > 0, "central case expected"
    assert abs(got - _spherical_psa(P)) < 1e-9
    for i in range(100):
        n = int(rng.integers(3, 9))
        pts = synth.random_convex_polygon(rng, n) * rng.uniform(0.5, 3.0)
        # inscribed in a circle of radius >= 0.5 with gaps < 0.95 pi, so an
        # offset below 0.5 cos(0.475 pi) = 0.039 keeps the zenith inside
        P = np.concatenate([pts + rng.uniform(-0.02, 0.02, 2), np.full((n, 1), rng.uniform(0.3, 2.0))], 1)
        v = np.zeros((9, 3)); v[:n] = P; v[n:] = P[0]
        got = orc_lib.orc_prepare_psa_f64(n, f64p(v), f64p(state))
        assert state[36] > 0
        assert abs(got - _spherical_psa(P)) <= 1e-9 * got


def test_psa_samples_are_uniform_in_projected_solid_angle(orc_lib):
    """Samples lie in the polygon's cone and their mean projected position
    matches the rejection-sampled mean (uniform on the projected region)."""
    rng = np.random.default_rng(13)
    for trial in range(6):
        P = _random_polygon_above(rng, int(rng.integers(3, 9)))
        if (P[:, 2] <= 0).any():
            continue
        v = np.zeros((9, 3)); v[:len(P)] = P; v[len(P):] = P[0]
        n = 20000
        uv = rng.random((n, 2))
        dirs = np.zeros((n, 3))
        orc_lib.orc_sample_psa_f64(len(P), f64p(v), f64p(np.ascontiguousarray(uv)), n, f64p(dirs))
        normals = np.array([np.cross(P[i], P[(i + 1) % len(P)]) for i in range(len(P))])
        inside = np.all((dirs @ normals.T) >= -1e-9, axis=1)
        assert inside.mean() > 0.999
        # rejection reference on the unit disk
        q = rng.uniform(-1, 1, (400000, 2))
        q = q[(q ** 2).sum(1) < 1]
        d = np.concatenate([q, np.sqrt(1 - (q ** 2).sum(1, keepdims=True))], 1)
        keep = np.all((d @ normals.T) >= 0, axis=1)
        ref_mean = q[keep].mean(0)
        spread = q[keep].std(0)
        tol = 5 * spread / math.sqrt(n) + 5 * spread / math.sqrt(keep.sum())
        assert np.all(np.abs(dirs[:, :2].mean(0) - ref_mean) < tol + 1e-4)
        # area of the projected region = PSA
        psa = orc_lib.orc_prepare_psa_f64(len(P), f64p(v), None)
        area = keep.mean() * 4.0 * (q.shape[0] / 400000)
        assert abs(area - psa) < 4 * math.sqrt(area * 4 / 400000) + 1e-3


# ---------------------------------------------------------------- BRDF
def _brdf_numpy(n, o, alb, f0, a, l):
    h = (l + o) / np.linalg.norm(l + o)
    li, lo, ndh, odh = n @ l, n @ o, n @ h, o @ h
    D = a * a / (((a * a - 1) * ndh * ndh + 1) ** 2)  # GGX without 1/pi
    V = 0.5 / (li * math.sqrt(lo * lo * (1 - a * a) + a * a) + lo * math.sqrt(li * li * (1 - a * a) + a * a))
    F = f0 + (1 - f0) * (1 - min(max(odh, 0), 1)) ** 5
    lum = F @ np.array([0.21263901, 0.71516868, 0.07219232])
    return (alb + D * V * lum) / math.pi


def test_brdf_closed_form_and_reciprocity(orc_lib):
    rng = np.random.default_rng(14)
    for i in range(300):
        n = np.array([0, 0, 1.0])
        o = rng.normal(size=3); o[2] = abs(o[2]) + 0.05; o /= np.linalg.norm(o)
        l = rng.normal(size=3); l[2] = abs(l[2]) + 0.05; l /= np.linalg.norm(l)
        alb = rng.random(3); f0 = rng.random(3) * 0.2; a = rng.uniform(0.05, 1)
        sh = np.concatenate([n, o, alb, f0, [a]])
        out = np.zeros(3)
        orc_lib.orc_evaluate_brdf_f64(f64p(sh), f64p(l), f64p(out))
        assert np.allclose(out, _brdf_numpy(n, o, alb, f0, a, l), rtol=1e-10)
        sh2 = np.concatenate([n, l, alb, f0, [a]])
        out2 = np.zeros(3)
        orc_lib.orc_evaluate_brdf_f64(f64p(sh2), f64p(o), f64p(out2))
        assert np.allclose(out, out2, rtol=1e-10)


# ---------------------------------------------------------------- LBVH
def test_lbvh_any_and_closest_hit_match_brute_force(orc_lib):
    rng = np.random.default_rng(15)
    tris = rng.uniform(-5, 5, (300, 3, 3)).astype(np.float32)
    tris[:, 1:] = tris[:, :1] + rng.normal(scale=0.7, size=(300, 2, 3)).astype(np.float32)
    flat = np.ascontiguousarray(tris.reshape(-1, 9))
    bvh = orc.OrcBvh()
    assert orc_lib.orc_bvh_build(C.byref(bvh), f32p(flat), 300) == 0
    n = 300
    # structure: every leaf reachable once, parent boxes contain children
    box = np.ctypeslib.as_array(bvh.node_box, shape=((2 * n - 1) * 6,)).reshape(-1, 6)
    left = np.ctypeslib.as_array(bvh.node_left, shape=(2 * n - 1,))
    right = np.ctypeslib.as_array(bvh.node_right, shape=(2 * n - 1,))
    seen = []
    stack = [0]
    while stack:
        x = stack.pop()
        if x >= n - 1:
            seen.append(x)
            continue
        for ch in (left[x], right[x]):
            assert np.all(box[ch, :3] >= box[x, :3]) and np.all(box[ch, 3:] <= box[x, 3:])
            stack.append(ch)
    assert sorted(seen) == list(range(n - 1, 2 * n - 1))
    hits = 0
    for i in range(3000):
        o = rng.uniform(-6, 6, 3).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        tmax = float(np.float32(rng.uniform(0.5, 12)))
        a = orc_lib.orc_bvh_any_hit(C.byref(bvh), f32p(o), f32p(d), 1e-3, tmax, None)
        b = orc_lib.orc_brute_any_hit(f32p(flat), 300, f32p(o), f32p(d), 1e-3, tmax)
        assert a == b
        hits += a
        t = C.c_float()
        p = orc_lib.orc_bvh_closest_hit(C.byref(bvh), f32p(o), f32p(d), 0.0, float("inf"), C.byref(t))
        best, bt = 0xFFFFFFFF, np.inf
        th = C.c_float()
        for k in range(300):
            if orc_lib.orc_ray_triangle(f32p(o), f32p(d), f32p(flat[k]), 0.0, float(bt), C.byref(th)):
                if th.value < bt:
                    bt, best = th.value, k
        assert p == best or (best != 0xFFFFFFFF and abs(t.value - bt) == 0)
    assert 300 < hits < 2900
    orc_lib.orc_bvh_free(C.byref(bvh))


# ---------------------------------------------------------------- binned-SAH builder
def _sah_soup(seed, n):
    rng = np.random.default_rng(seed)
    tris = rng.uniform(-5, 5, (n, 3, 3)).astype(np.float32)
    tris[:, 1:] = tris[:, :1] + rng.normal(scale=0.7, size=(n, 2, 3)).astype(np.float32)
    return np.ascontiguousarray(tris.reshape(-1, 9))


def _tree(bvh, n):
    box = np.ctypeslib.as_array(bvh.node_box, shape=((2 * n - 1) * 6,)).reshape(-1, 6).copy()
    left = np.ctypeslib.as_array(bvh.node_left, shape=(2 * n - 1,)).copy()
    right = np.ctypeslib.as_array(bvh.node_right, shape=(2 * n - 1,)).copy()
    prim = np.ctypeslib.as_array(bvh.leaf_prim, shape=(n,)).copy()
    return box, left, right, prim


def _subtree_prims(x, n, left, right, prim):
    out, stack = [], [x]
    while stack:
        y = stack.pop()
        if y >= n - 1:
            out.append(int(prim[y - (n - 1)]))
        else:
            stack += [right[y], left[y]]
    return out


def _tri_boxes(flat):
    v = flat.reshape(-1, 3, 3)
    return v.min(1), v.max(1)


@pytest.mark.parametrize("case", ["soup", "duplicates", "flat", "one", "two"])
def test_sah_tree_structure_and_hits(orc_lib, case):
    """Structure of orc_bvh_build_sah: every triangle in exactly one leaf,
    internal boxes are the exact union of their triangles' boxes, and each
    split is a bin boundary: on some axis every left centroid lies in a lower
    bin than every right centroid (or, for coincident centroids, the range is
    halved). Any-hit equals brute force over all triangles."""
    n = {"soup": 700, "duplicates": 300, "flat": 200, "one": 1, "two": 2}[case]
    flat = _sah_soup(21, n)
    if case == "duplicates":
        flat[::3] = flat[0]
    if case == "flat":
        flat[:] = flat[0]
    bvh = orc.OrcBvh()
    assert orc_lib.orc_bvh_build_sah(C.byref(bvh), f32p(flat), n) == 0
    box, left, right, prim = _tree(bvh, n)
    lo, hi = _tri_boxes(flat)
    cen = (np.float32(0.5) * (lo + hi)).astype(np.float32)
    assert sorted(prim.tolist()) == list(range(n))
    assert np.array_equal(bvh_leaf := np.ctypeslib.as_array(bvh.leaf_tri, shape=(n * 9,)).reshape(n, 9),
                          flat[prim])
    for x in range(n - 1):
        ps = _subtree_prims(x, n, left, right, prim)
        assert np.array_equal(box[x, :3], lo[ps].min(0)) and np.array_equal(box[x, 3:], hi[ps].max(0))
        pl = _subtree_prims(left[x], n, left, right, prim)
        pr = _subtree_prims(right[x], n, left, right, prim)
        c = cen[ps]
        ext = c.max(0) - c.min(0)
        if np.all(ext == 0):
            assert abs(len(pl) - len(pr)) <= 1
            continue
        ok = False
        for a in range(3):
            if ext[a] <= 0:
                continue
            scale = np.float32(16) / np.float32(ext[a])
            bins = np.minimum(((cen[:, a] - c[:, a].min()) * scale).astype(np.int64), 15)
            if bins[pl].max() < bins[pr].min():
                ok = True
        assert ok, x
    rng = np.random.default_rng(22)
    hits = 0
    for i in range(1500):
        o = rng.uniform(-6, 6, 3).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        tmax = float(np.float32(rng.uniform(0.5, 12)))
        a = orc_lib.orc_bvh_any_hit(C.byref(bvh), f32p(o), f32p(d), 1e-3, tmax, None)
        b = orc_lib.orc_brute_any_hit(f32p(flat), n, f32p(o), f32p(d), 1e-3, tmax)
        assert a == b
        hits += a
    if case == "soup":
        assert 200 < hits < 1400
    orc_lib.orc_bvh_free(C.byref(bvh))


def test_sah_root_split_minimises_binned_cost(orc_lib):
    """The root split is the first minimum of n_l A_l + n_r A_r over the 45
    binned candidates (A = dx dy + dy dz + dz dx of the exact box unions),
    evaluated here with numpy from the triangle boxes alone."""
    n = 900
    flat = _sah_soup(23, n)
    bvh = orc.OrcBvh()
    assert orc_lib.orc_bvh_build_sah(C.byref(bvh), f32p(flat), n) == 0
    box, left, right, prim = _tree(bvh, n)
    lo, hi = _tri_boxes(flat)
    cen = (np.float32(0.5) * (lo + hi)).astype(np.float32)
    pl = set(_subtree_prims(left[0], n, left, right, prim))
    costs = []
    for a in range(3):
        ext = np.float32(cen[:, a].max() - cen[:, a].min())
        scale = np.float32(16) / ext
        bins = np.minimum(((cen[:, a] - cen[:, a].min()) * scale).astype(np.int64), 15)
        for s in range(1, 16):
            L, R = bins < s, bins >= s
            if not L.any() or not R.any():
                continue
            area = lambda m: (lambda d: d[0] * d[1] + d[1] * d[2] + d[2] * d[0])(
                (hi[m].max(0) - lo[m].min(0)).astype(np.float64))
            costs.append((L.sum() * area(L) + R.sum() * area(R), a, s, set(np.flatnonzero(L).tolist())))
    best = min(c[0] for c in costs)
    chosen = [c for c in costs if c[3] == pl]
    assert chosen, "root split is not a binned split"
    # equal to the minimum up to the float32 rounding of the kernel's cost
    assert chosen[0][0] <= best * (1 + 1e-5)
    orc_lib.orc_bvh_free(C.byref(bvh))


def test_sah_records_tile_preorder(orc_lib):
    """orc_bvh_records: 5n-2 units, internal escape = first unit past the
    subtree, walking the records with escapes visits every leaf once."""
    n = 400
    flat = _sah_soup(24, n)
    bvh = orc.OrcBvh()
    assert orc_lib.orc_bvh_build_sah(C.byref(bvh), f32p(flat), n) == 0
    pos = np.zeros(6 * n, np.uint32)
    pos[1::6] = np.where(np.arange(n) % 7 == 0, 0x80000000, 0).astype(np.uint32)
    rec = np.zeros((5 * n - 2, 4), np.float32)
    orc_lib.orc_bvh_records(C.byref(bvh), pos.ctypes.data_as(C.POINTER(C.c_uint32)), f32p(rec))
    w = rec[:, 3].view(np.int32)
    leaves, x = [], 0
    while x < 5 * n - 2:
        if w[x] < 0:
            t = int(rec[x + 1, 3].view(np.uint32))
            leaves.append(t)
            assert np.array_equal(rec[x:x + 3, :3].ravel(), flat[t])
            assert rec[x + 2, 3].view(np.uint32) == (0x80000000 if t % 7 == 0 else 0)
            x += 3
        else:
            assert x + 2 < w[x] <= 5 * n - 2
            x += 2
    assert sorted(leaves) == list(range(n))
    orc_lib.orc_bvh_free(C.byref(bvh))


# ---------------------------------------------------------------- LTC table handling
def test_ltc_quantisation_three_ways(orc_lib, tmp_path):
    """load_ltc_table semantics (src/ltc_table.c:82-116): oracle entry
    quantiser == vectorised quantiser == fit-file reader of the library."""
    from risltc_amd import ltcfit
    fit, i5 = ltcfit.load_fit()
    rgba, rg = ltcfit.quantize(fit, i5, 3)
    for j in np.random.default_rng(0).integers(0, fit.shape[0], 300):
        for k in range(3):
            d = ltcfit.layer_fit(fit, i5, k, 3)[j].astype(np.float32)
            q4 = np.zeros(4, np.uint16)
            q2 = np.zeros(2, np.uint16)
            orc_lib.orc_quantize_ltc_entry(f32p(d), q4.ctypes.data_as(C.POINTER(C.c_uint16)),
                                           q2.ctypes.data_as(C.POINTER(C.c_uint16)))
            res = rgba.shape[1]
            assert np.array_equal(q4, rgba[k].reshape(-1, 4)[j])
            assert np.array_equal(q2, rg[k].reshape(-1, 2)[j])


def test_ltc_constants_match_loader(orc_lib):
    k = abi.LtcConstants()
    orc_lib.orc_ltc_constants(C.byref(k), 64, 64, 51)
    assert k.fresnel_index_factor == 50.0
    assert np.float32(k.roughness_factor) == np.float32(63.0) / np.float32(64.0)
    assert np.float32(k.inclination_summand) == np.float32(0.5) / np.float32(64.0)


def test_identity_ltc_matches_cosine(orc_lib, small_oracle):
    """A table whose matrix is the identity gives shading_to_cosine = I and
    the albedo channel; its determinant is 1."""
    fit = np.zeros((16 * 16, 5), np.float32)
    fit[:, 0] = 1.0; fit[:, 2] = 1.0; fit[:, 4] = 0.5
    from risltc_amd import ltcfit
    rgba, rg = ltcfit.quantize(fit, np.full(256, 0.5, np.float32), 2)
    sc = orc.OrcScene()
    rgba_c = np.ascontiguousarray(rgba); rg_c = np.ascontiguousarray(rg)
    sc.ltc_rgba16 = rgba_c.ctypes.data_as(C.POINTER(C.c_uint16))
    sc.ltc_rg16 = rg_c.ctypes.data_as(C.POINTER(C.c_uint16))
    sc.ltc_res_roughness = sc.ltc_res_inclination = 16
    sc.ltc_fresnel_count = 2
    k = abi.LtcConstants()
    orc_lib.orc_ltc_constants(C.byref(k), 16, 16, 2)
    out = np.zeros(32, np.float32)
    pos = np.array([1, 2, 3], np.float32)
    nrm = np.array([0, 0, 1], np.float32)
    o = np.array([0.6, 0, 0.8], np.float32)
    orc_lib.orc_ltc_coefficients_f32(C.byref(sc), C.byref(k), 0.02, 0.3, f32p(pos), f32p(nrm), f32p(o), f32p(out))
    assert np.allclose(out[12:21].reshape(3, 3), np.eye(3), atol=1e-4)
    assert abs(out[30] - 0.5) < 1e-4 and abs(out[31] - 1.0) < 1e-3
    # world_to_shading maps the normal to z, the view direction into +x
    w2s = out[:12].reshape(3, 4)
    assert np.allclose(w2s[:, :3] @ nrm, [0, 0, 1], atol=1e-6)
    v = w2s[:, :3] @ o
    assert v[0] > 0 and abs(v[1]) < 1e-6
    assert np.allclose(w2s[:, :3] @ pos + w2s[:, 3], 0, atol=1e-5)


# ---------------------------------------------------------------- frame constants
def test_pixel_rays_match_camera_trigonometry(small_oracle):
    """pixel_to_ray * (x, y, 1) points through the pixel centre of a pinhole
    camera with the given vertical field of view (direct construction)."""
    sc = small_oracle.scene
    c = small_oracle.frame_constants()
    P = np.array(c.pixel_to_ray_direction_world_space, dtype=np.float64)[:, :3]
    cam = sc.camera
    W, H = sc.width, sc.height
    cx, sx = math.cos(cam.rotation_x), math.sin(cam.rotation_x)
    cz, sz = math.cos(cam.rotation_z), math.sin(cam.rotation_z)
    rot = np.array([[cz, sz, 0], [-sz, cz, 0], [0, 0, 1]]) @ np.array([[1, 0, 0], [0, cx, sx], [0, -sx, cx]])
    top = math.tan(0.5 * cam.vertical_fov)
    right = W / H * top
    for (x, y) in [(0, 0), (W - 1, 0), (W // 2, H // 3), (W - 1, H - 1)]:
        d = P @ np.array([x, y, 1.0])
        ndc = np.array([(x + 0.5) / W * 2 - 1, (y + 0.5) / H * 2 - 1])
        view = np.array([-ndc[0] * right, ndc[1] * top, -1.0])
        expect = rot @ view
        # the reference inverts the projection matrix in FP32 (condition ~ far/near
        # = 2e4), so the directions carry ~1e-4 of rounding error
        assert np.allclose(d / np.linalg.norm(d), expect / np.linalg.norm(expect), atol=1e-3)


def test_timing_build_matches_checker_build(small_scene, ltc_table):
    """The -O3 build timed as the CPU baseline computes the same bits as the
    checker build (both IEEE, no contraction)."""
    a = orc.Oracle(small_scene, *ltc_table)
    b = orc.Oracle(small_scene, *ltc_table, timing=True)
    c = a.frame_constants()
    vis = a.render_visibility(c)
    s = small_scene.settings(sample_count_light=2)
    pix = np.flatnonzero((vis >> 31) == 0)[::7].astype(np.uint32)
    ra, _, _ = a.shade(c, s, vis, pixels=pix)
    rb, _, _ = b.shade(c, s, vis, pixels=pix)
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
    # the OpenMP frame entry point of the timing build: the same bits again
    fb = b.shade_frame(c, s, vis, pixels=pix, threads=4)
    assert np.array_equal(fb.view(np.uint32), ra.view(np.uint32))
