/*
 * SPDX-License-Identifier: GPL-3.0-or-later
 *
 * This file restates, for testing, algorithms of the risltc renderer
 * (Copyright (C) 2021, Christoph Peters, Karlsruhe Institute of Technology;
 * Copyright (C) 2023, Ishaan Shah, International Institute of Information
 * Technology, Hyderabad), which is licensed under the GNU General Public
 * License version 3 or later. As a derivative of that code, this file is
 * distributed under the same licence: see oracle/COPYING.
 */
/*
 * orc_core.c -- the per-pixel shading path of the reference, restated in C99.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h). Compiled twice: with ORC_F64 the
 * arithmetic type R is double and transcendentals come from libm; otherwise
 * R is float, GLSL fma() is fmaf, and atan/acos/sin/cos are the project
 * Cephes versions (prj_*) that the GPU kernels also use. Build with
 * -ffp-contract=off: a product followed by a sum is two roundings unless an
 * FMA is written.
 *
 * Project conventions where the GLSL leaves the result to the driver (each
 * also used by the GPU kernels):
 *   dot(a,b)      = fma(a.z,b.z, fma(a.y,b.y, a.x*b.x))
 *   M * v         = fma(M[i][2],v.z, fma(M[i][1],v.y, M[i][0]*v.x)) (+ t_i
 *                   as the innermost addend for affine 4x3 matrices)
 *   cross(a,b)    = plain products and differences (exactly antisymmetric)
 *   inversesqrt   = 1 / sqrt, normalize(v) = v * inversesqrt(dot(v,v))
 *   texture()     = bilinear in FP32 on UNORM16 / 65535, clamp to edge,
 *                   array layer = round-half-even, clamp (hazard H5)
 *   light index   = min(int(u * N), N - 1) (hazard H2: u can be 1.0)
 *   padding       = clipped polygons repeat vertex 0 in every slot past the
 *                   last vertex (hazard H3), so padding edges add exactly 0.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#ifdef ORC_F64
typedef double R;
#define SFX(name) name##_f64
#define K(x) (x)
#define FMA fma
#define SQRT sqrt
#define FABS fabs
#define FMIN fmin
#define FMAX fmax
#define FLOOR floor
#define RINT rint
#define ATAN atan
#define ACOS acos
#define SIN sin
#define COS cos
#define R_INF HUGE_VAL
#define R_PI 3.1415926535897932384626433832795
#define R_HALF_PI 1.5707963267948966192313216916398
#define R_INV_PI 0.31830988618379067153776752674503
#else
typedef float R;
#define SFX(name) name##_f32
#define K(x) (x##f)
#define FMA fmaf
#define SQRT sqrtf
#define FABS fabsf
#define FMIN fminf
#define FMAX fmaxf
#define FLOOR floorf
#define RINT rintf
#define ATAN prj_atanf
#define ACOS prj_acosf
#define SIN prj_sinf
#define COS prj_cosf
#define R_INF HUGE_VALF
#define R_PI 3.1415926535897932384626433832795f
#define R_HALF_PI 1.5707963267948966192313216916398f
#define R_INV_PI 0.31830988618379067153776752674503f
#endif

#define MAXV RISLTC_MAX_POLYGON_VERTEX_COUNT

typedef struct { R x, y; } V2;
typedef struct { R x, y, z; } V3;

static V2 v2(R x, R y) { V2 r = {x, y}; return r; }
static V3 v3(R x, R y, R z) { V3 r = {x, y, z}; return r; }
static V3 add3(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static V3 sub3(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static V3 mul3(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static V3 scl3(V3 a, R s) { return v3(a.x * s, a.y * s, a.z * s); }
static V3 sclf3(R s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static V2 add2(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }
static V2 sub2(V2 a, V2 b) { return v2(a.x - b.x, a.y - b.y); }
static V2 scl2(V2 a, R s) { return v2(a.x * s, a.y * s); }
static V2 sclf2(R s, V2 a) { return v2(s * a.x, s * a.y); }
static R dot3(V3 a, V3 b) { return FMA(a.z, b.z, FMA(a.y, b.y, a.x * b.x)); }
static R dot2(V2 a, V2 b) { return FMA(a.y, b.y, a.x * b.x); }
static V3 cross3(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static R rsq(R x) { return K(1.0) / SQRT(x); }
static V3 normalize3(V3 a) { return scl3(a, rsq(dot3(a, a))); }
static V2 normalize2(V2 a) { return scl2(a, rsq(dot2(a, a))); }
static R clampr(R x, R lo, R hi) { return FMIN(FMAX(x, lo), hi); }
static int sign_bit(R x) { return signbit(x) ? 1 : 0; }

/* affine 3x4 (row-major, column 3 = translation) times point */
typedef struct { R m[3][4]; } A34;
typedef struct { R m[3][3]; } M33;
static V3 xform(const A34* a, V3 p) {
    R r[3];
    for (int i = 0; i != 3; ++i)
        r[i] = FMA(a->m[i][2], p.z, FMA(a->m[i][1], p.y, FMA(a->m[i][0], p.x, a->m[i][3])));
    return v3(r[0], r[1], r[2]);
}
static V3 mul33(const M33* a, V3 p) {
    R r[3];
    for (int i = 0; i != 3; ++i) r[i] = FMA(a->m[i][2], p.z, FMA(a->m[i][1], p.y, a->m[i][0] * p.x));
    return v3(r[0], r[1], r[2]);
}
/* transpose(mat4x3) * vec3 keeping .xyz: R^T d */
static V3 mul_transposed(const A34* a, V3 d) {
    R r[3];
    for (int i = 0; i != 3; ++i) r[i] = FMA(a->m[2][i], d.z, FMA(a->m[1][i], d.y, a->m[0][i] * d.x));
    return v3(r[0], r[1], r[2]);
}
/* mat3 * mat4x3 */
static A34 mul33_34(const M33* s, const A34* w) {
    A34 r;
    for (int i = 0; i != 3; ++i)
        for (int k = 0; k != 4; ++k)
            r.m[i][k] = FMA(s->m[i][2], w->m[2][k], FMA(s->m[i][1], w->m[1][k], s->m[i][0] * w->m[0][k]));
    return r;
}

/* ---------------- data ---------------------------------------------------- */
typedef struct {
    V3 position, normal, outgoing;
    R lambert_outgoing;
    V3 diffuse_albedo, fresnel_0;
    R roughness;
} shading_t;

typedef struct {
    A34 world_to_shading;
    M33 shading_to_cosine, cosine_to_shading;
    R albedo, determinant;
} ltc_t;

typedef struct {
    V3 radiance;
    R plane[4];
    uint32_t vertex_count;
    V3 v[RISLTC_MAX_LIGHT_VERTEX_COUNT];
} light_t;

typedef struct {
    uint32_t vertex_count;
    V2 vertices[MAXV];
    V2 ellipses[MAXV];
    V2 inner_ellipse_0;
    R sector[MAXV];
    R psa;
} psa_t;

static void load_light(const orc_scene_t* sc, uint32_t index, light_t* l) {
    const uint8_t* rec = sc->light_records + (size_t)index * RISLTC_LIGHT_RECORD_SIZE(sc->light_max_vertex_count);
    const float* f = (const float*)rec;
    l->radiance = v3(f[0], f[1], f[2]);
    for (int i = 0; i != 4; ++i) l->plane[i] = f[4 + i];
    memcpy(&l->vertex_count, rec + 32, 4);
    for (uint32_t i = 0; i != l->vertex_count; ++i) {
        const float* p = (const float*)(rec + 48 + 16 * i);
        l->v[i] = v3(p[0], p[1], p[2]);
    }
}

/* dot(vec4(p, 1), plane) */
static R plane_side(V3 p, const R plane[4]) {
    return FMA(K(1.0), plane[3], FMA(p.z, plane[2], FMA(p.y, plane[1], p.x * plane[0])));
}

/* ---------------- clipping (polygon_clipping.glsl:19-225) ---------------- */
/* iz0 (polygon_clipping.glsl:19-25) */
static V3 iz0(V3 lhs, V3 rhs) {
    R t = lhs.z / (lhs.z - rhs.z);
    return v3(FMA(t, rhs.x, FMA(-t, lhs.x, lhs.x)), FMA(t, rhs.y, FMA(-t, lhs.y, lhs.y)), K(0.0));
}

/* clip_polygon through the case rule of orc_clip_rule; the result fills all
 * MAXV slots (slots past the count repeat slot 0). */
static uint32_t clip_polygon(uint32_t n, V3 v[MAXV]) {
    uint32_t mask = 0;
    for (uint32_t i = 0; i != n; ++i) mask |= (v[i].z > K(0.0)) ? (1u << i) : 0u;
    int codes[MAXV + 1];
    int vc = orc_clip_rule((int)n, mask, codes);
    if (vc == 0) return 0;
    V3 out[MAXV];
    for (int j = 0; j != vc; ++j) {
        int c = codes[j];
        if (c >= 0) {
            out[j] = v[c];
        } else {
            int e = -c - 1;
            out[j] = iz0(v[e], v[(e + 1) % (int)n]);
        }
    }
    for (int j = vc; j != MAXV; ++j) out[j] = out[0];
    memcpy(v, out, sizeof(out));
    return (uint32_t)vc;
}

/* ---------------- LTC integral (polygon_sampling.glsl:508-530) ----------- */
static R integrate_edge(V3 a, V3 b) {
    /* a and b are already normalized (each vertex is normalized once) */
    R x = dot3(a, b);
    R y = FABS(x);
    R A = K(0.8543985) + (K(0.4965155) + K(0.0145206) * y) * y;
    R B = K(3.4175940) + (K(4.1616724) + y) * y;
    R v = A / B;
    R ts = (x > K(0.0)) ? v : K(0.5) * (K(1.0) / SQRT(FMAX(K(1.0) - x * x, K(1e-7)))) - v;
    return (a.x * b.y - a.y * b.x) * ts;
}

static R calculate_ltc(uint32_t vc, const V3 v[MAXV]) {
    V3 nv[MAXV];
    for (uint32_t i = 0; i != vc; ++i) nv[i] = normalize3(v[i]);
    R result = K(0.0);
    for (uint32_t i = 0; i != vc; ++i) result += integrate_edge(nv[i], nv[(i + 1) % vc]);
    return FABS(result);
}

/* ---------------- projected solid angle (polygon_sampling.glsl) --------- */
static R fast_positive_atan(R y) {
    R rx = (FABS(y) > K(1.0)) ? (K(1.0) / FABS(y)) : FABS(y);
    R ry = rx * rx;
    R rz = FMA(ry, K(0.02083509974181652), K(-0.08513300120830536));
    rz = FMA(ry, rz, K(0.18014100193977356));
    rz = FMA(ry, rz, K(-0.3302994966506958));
    ry = FMA(ry, rz, K(0.9998660087585449));
    rz = FMA(K(-2.0) * ry, rx, R_HALF_PI);
    rz = (FABS(y) > K(1.0)) ? rz : K(0.0);
    rx = FMA(rx, ry, rz);
    return (y < K(0.0)) ? (R_PI - rx) : rx;
}

static int g_fast_atan = 0;
static R positive_atan(R t) {
    if (g_fast_atan) return fast_positive_atan(t);
    R offset = (t < K(0.0)) ? R_PI : K(0.0);
    return ATAN(t) + offset;
}

static R mix_fma(R x, R y, R a) { return FMA(a, y, FMA(-a, x, x)); }

static R kahan(R a, R b, R c, R d) {
    R cd = c * d;
    R err = FMA(c, d, -cd);
    R res = FMA(a, b, -cd);
    return res - err;
}
static V3 cross_stable(V3 l, V3 r) {
    return v3(kahan(l.y, r.z, l.z, r.y), kahan(l.z, r.x, l.x, r.z), kahan(l.x, r.y, l.y, r.x));
}
static V2 rotate_90(V2 v) { return v2(-v.y, v.x); }
static int is_inner_ellipse(V2 e) { return sign_bit(e.x); }

static V2 ellipse_from_edge(V3 a, V3 b) {
    V3 n = cross_stable(a, b);
    R s = K(1.0) / n.z;
    s = sign_bit(n.x) ? -s : s;
    V2 e = v2(n.x * s, n.y * s);
    e.x = (n.z != K(0.0)) ? e.x : R_INF;
    return e;
}
static V2 ellipse_transform(V2 e, V2 p) {
    R d = dot2(e, p);
    return v2(FMA(d, e.x, p.x), FMA(d, e.y, p.y));
}
static R ellipse_det(V2 e) { return FMA(e.x, e.x, FMA(e.y, e.y, K(1.0))); }
static R ellipse_rsqrt_det(V2 e) { return rsq(ellipse_det(e)); }
static R ellipse_dir_factor_rsq(V2 e, V2 d) {
    R ed = dot2(e, d);
    R dd = dot2(d, d);
    return FMA(ed, ed, dd);
}
static R ellipse_dir_factor(V2 e, V2 d) { return rsq(ellipse_dir_factor_rsq(e, d)); }
static R ellipse_normalized_dir_factor(V2 e, V2 nd) {
    R ed = dot2(e, nd);
    return rsq(FMA(ed, ed, K(1.0)));
}
static R area_from_tangents(R ird, R it, R ord, R ot) {
    R inner = ird * positive_atan(it);
    R r = FMA(ord, positive_atan(ot), -inner);
    return (r > K(0.0)) ? (K(0.5) * r) : K(0.0);
}
static R area_between(V2 ie, R ird, V2 oe, R ord, V2 d0, V2 d1) {
    R det_dirs = FMAX(K(0.0), dot2(d1, rotate_90(d0)));
    R idot = ird * dot2(d0, ellipse_transform(ie, d1));
    R odot = ord * dot2(d0, ellipse_transform(oe, d1));
    return area_from_tangents(ird, det_dirs / idot, ord, det_dirs / odot);
}
static R ellipse_area_in_sector(V2 e, V2 d0, V2 d1) {
    R rd = ellipse_rsqrt_det(e);
    R det_dirs = FMAX(K(0.0), dot2(d1, rotate_90(d0)));
    R edot = rd * dot2(d0, ellipse_transform(e, d1));
    R area = K(0.5) * rd * positive_atan(det_dirs / edot);
    return (rd > K(0.0)) ? area : K(0.0);
}

static void compare_and_swap(psa_t* p, int l, int r) {
    V2 lc = p->vertices[l];
    R nz = kahan(lc.x, -p->vertices[r].y, lc.y, -p->vertices[r].x);
    int swap = (nz == K(0.0)) ? (isinf(p->ellipses[r].x) ? 1 : 0) : (nz > K(0.0));
    if (swap) {
        p->vertices[l] = p->vertices[r];
        p->vertices[r] = lc;
        V2 e = p->ellipses[l];
        p->ellipses[l] = p->ellipses[r];
        p->ellipses[r] = e;
    }
}

/* sort_convex_polygon_vertices (polygon_sampling.glsl:441-506); nonagons
 * (only produced by clipped 8-gons, which the reference cannot handle) use a
 * complete 25-comparator network for 9 inputs before the shared tail. */
static const int g_net9[25][2] = {{0, 1}, {3, 4}, {6, 7}, {1, 2}, {4, 5}, {7, 8}, {0, 1}, {3, 4}, {6, 7},
                                  {0, 3}, {3, 6}, {0, 3}, {1, 4}, {4, 7}, {1, 4}, {2, 5}, {5, 8}, {2, 5},
                                  {1, 3}, {5, 7}, {2, 6}, {4, 6}, {2, 4}, {2, 3}, {5, 6}};
static void sort_vertices(psa_t* p) {
    uint32_t n = p->vertex_count;
    if (n == 3) {
        compare_and_swap(p, 1, 2);
    } else if (n == 4) {
        compare_and_swap(p, 1, 3);
    } else if (n == 5) {
        compare_and_swap(p, 2, 4); compare_and_swap(p, 1, 3); compare_and_swap(p, 1, 2);
        compare_and_swap(p, 0, 3); compare_and_swap(p, 3, 4);
    } else if (n == 6) {
        compare_and_swap(p, 3, 5); compare_and_swap(p, 2, 4); compare_and_swap(p, 1, 5);
        compare_and_swap(p, 0, 4); compare_and_swap(p, 4, 5); compare_and_swap(p, 1, 3);
    } else if (n == 7) {
        compare_and_swap(p, 2, 5); compare_and_swap(p, 1, 6); compare_and_swap(p, 5, 6);
        compare_and_swap(p, 3, 4); compare_and_swap(p, 0, 4); compare_and_swap(p, 4, 6);
        compare_and_swap(p, 1, 3); compare_and_swap(p, 3, 5); compare_and_swap(p, 4, 5);
    } else if (n == 8) {
        compare_and_swap(p, 2, 6); compare_and_swap(p, 3, 7); compare_and_swap(p, 1, 5);
        compare_and_swap(p, 0, 4); compare_and_swap(p, 4, 6); compare_and_swap(p, 5, 7);
        compare_and_swap(p, 6, 7); compare_and_swap(p, 4, 5); compare_and_swap(p, 1, 3);
    } else if (n == 9) {
        for (int i = 0; i != 25; ++i) compare_and_swap(p, g_net9[i][0], g_net9[i][1]);
    }
    compare_and_swap(p, 0, 2);
    if (n >= 4) compare_and_swap(p, 2, 3);
    compare_and_swap(p, 0, 1);
}

/* prepare_projected_solid_angle_polygon_sampling (polygon_sampling.glsl:545-613) */
static void prepare_psa(psa_t* p, uint32_t vc, const V3 v[MAXV]) {
    memset(p, 0, sizeof(*p));
    p->vertex_count = vc;
    p->inner_ellipse_0 = v2(K(1.0), K(0.0));
    p->vertices[0] = v2(v[0].x, v[0].y);
    p->ellipses[0] = ellipse_from_edge(v[0], v[1]);
    V2 prev = p->ellipses[0];
    for (uint32_t i = 1; i != MAXV; ++i) {
        p->vertices[i] = v2(v[i].x, v[i].y);
        if (i > 2 && i == vc) break;
        V2 e = ellipse_from_edge(v[i], v[(i + 1) % MAXV]);
        int inner = is_inner_ellipse(e);
        p->ellipses[i] = inner ? prev : e;
        p->inner_ellipse_0 = (is_inner_ellipse(prev) && !inner) ? prev : p->inner_ellipse_0;
        prev = e;
    }
    V2 e = p->ellipses[0];
    int inner = is_inner_ellipse(e);
    p->ellipses[0] = inner ? prev : e;
    p->inner_ellipse_0 = (is_inner_ellipse(prev) && !inner) ? prev : p->inner_ellipse_0;
    p->psa = K(0.0);
    if (p->inner_ellipse_0.x > K(0.0)) {
        for (uint32_t i = 0; i != MAXV; ++i) {
            if (i > 2 && i == vc) break;
            p->sector[i] = ellipse_area_in_sector(p->ellipses[i], p->vertices[i], p->vertices[(i + 1) % MAXV]);
            p->psa += p->sector[i];
        }
    } else {
        sort_vertices(p);
        V2 ie = p->inner_ellipse_0;
        R ird = ellipse_rsqrt_det(ie);
        V2 oe = v2(K(0.0), K(0.0));
        R ord = K(0.0);
        for (uint32_t i = 0; i != MAXV - 1; ++i) {
            if (i > 1 && i + 1 == vc) break;
            V2 ve = p->ellipses[i];
            int vin = is_inner_ellipse(ve);
            R vrd = ellipse_rsqrt_det(ve);
            if (i == 0) {
                oe = ve;
                ord = vrd;
            } else {
                ie = vin ? ve : ie;
                ird = vin ? vrd : ird;
                oe = vin ? oe : ve;
                ord = vin ? ord : vrd;
            }
            p->sector[i] = area_between(ie, ird, oe, ord, p->vertices[i], p->vertices[i + 1]);
            p->psa += p->sector[i];
        }
    }
}

static V2 normalize_approx_and_flip(V2 rhs, V2 semi) {
    R s = FABS(rhs.x) + FABS(rhs.y);
#ifdef ORC_F64
    s = 1.0 / s; /* the reference's exponent-flip trick is a float-only approximation of this */
#else
    uint32_t bits;
    memcpy(&bits, &s, 4);
    bits ^= 0x7F800000u;
    memcpy(&s, &bits, 4);
#endif
    s = (dot2(rhs, semi) >= K(0.0)) ? s : -s;
    return sclf2(s, rhs);
}

/* mat2 stored column-major: q[col][row] */
typedef struct { R q[2][2]; } Q2;
static Q2 outer2(V2 c, V2 r) {
    Q2 m;
    m.q[0][0] = c.x * r.x;
    m.q[0][1] = c.y * r.x;
    m.q[1][0] = c.x * r.y;
    m.q[1][1] = c.y * r.y;
    return m;
}
static Q2 sub_q(Q2 a, Q2 b) {
    Q2 m;
    for (int i = 0; i != 2; ++i)
        for (int j = 0; j != 2; ++j) m.q[i][j] = a.q[i][j] - b.q[i][j];
    return m;
}
static V2 solve_quadratic(Q2 m) {
    R cxy = K(0.5) * (m.q[0][1] + m.q[1][0]);
    R sd = SQRT(FMAX(K(0.0), cxy * cxy - m.q[0][0] * m.q[1][1]));
    R root = FABS(cxy) + sd;
    return (cxy >= K(0.0)) ? v2(root, -m.q[0][0]) : v2(m.q[1][1], root);
}

/* sample_sector_between_ellipses (polygon_sampling.glsl:668-762) */
static V2 sample_sector(V2 rn, R target_area, V2 ie, V2 oe, V2 d0, V2 d1, uint32_t iterations) {
    V2 qd[3];
    qd[0] = normalize2(d0);
    qd[2] = normalize2(d1);
    qd[1] = add2(qd[0], qd[2]);
    R nf[2][3] = {{ellipse_normalized_dir_factor(ie, qd[0]), ellipse_dir_factor(ie, qd[1]),
                   ellipse_normalized_dir_factor(ie, qd[2])},
                  {ellipse_normalized_dir_factor(oe, qd[0]), ellipse_dir_factor(oe, qd[1]),
                   ellipse_normalized_dir_factor(oe, qd[2])}};
    R sa[2] = {nf[1][0] * nf[1][1] - nf[0][0] * nf[0][1], nf[1][1] * nf[1][2] - nf[0][1] * nf[0][2]};
    R tqa = mix_fma(-sa[0], sa[1], rn.x);
    int sel = tqa <= K(0.0);
    qd[2] = sel ? qd[0] : qd[2];
    nf[0][2] = sel ? nf[0][0] : nf[0][2];
    nf[1][2] = sel ? nf[1][0] : nf[1][2];
    tqa += sel ? sa[0] : -sa[1];
    tqa *= FABS(qd[1].x * qd[2].y - qd[2].x * qd[1].y);
    V2 qn0 = add2(scl2(qd[1], nf[0][1]), scl2(qd[2], nf[0][2]));
    V2 qn1 = add2(scl2(qd[1], nf[1][1]), scl2(qd[2], nf[1][2]));
    qn0 = ellipse_transform(ie, qn0);
    qn1 = ellipse_transform(oe, qn1);
    R qo0 = dot2(qn0, qd[1]) * nf[0][1];
    R qo1 = dot2(qn1, qd[1]) * nf[1][1];
    V2 r2 = rotate_90(qd[2]);
    Q2 quad = outer2(sclf2(qo1 * nf[1][2], r2), qn0);
    quad = sub_q(quad, outer2(add2(sclf2(qo0 * nf[0][2], r2), sclf2(tqa, qn0)), qn1));
    V2 cur = solve_quadratic(quad);
    R acceptable = K(1.0e-5);
    iterations = (FABS(rn.x - K(0.5)) <= K(0.5) - acceptable) ? iterations : 0u;
    R ird = ellipse_rsqrt_det(ie);
    R ord = ellipse_rsqrt_det(oe);
    for (uint32_t i = 0; i != iterations; ++i) {
        cur = normalize_approx_and_flip(cur, qd[1]);
        V2 idir = ellipse_transform(ie, cur);
        V2 odir = ellipse_transform(oe, cur);
        R det_dirs = FMAX(K(0.0), dot2(cur, rotate_90(qd[0])));
        R err = target_area - area_from_tangents(ird, det_dirs / (ird * dot2(qd[0], idir)), ord,
                                                 det_dirs / (ord * dot2(qd[0], odir)));
        quad = sub_q(outer2(sub2(idir, odir), rotate_90(cur)), outer2(sclf2(K(2.0) * err, idir), odir));
        cur = solve_quadratic(quad);
    }
    cur = (dot2(cur, qd[1]) >= K(0.0)) ? cur : v2(-cur.x, -cur.y);
    R inner_f = K(1.0) / ellipse_dir_factor_rsq(ie, cur);
    R outer_f = K(1.0) / ellipse_dir_factor_rsq(oe, cur);
    return scl2(cur, SQRT(mix_fma(inner_f, outer_f, rn.y)));
}

/* sample_projected_solid_angle_polygon (polygon_sampling.glsl:772-828) */
static V3 sample_psa(const psa_t* p, V2 rn) {
    R target = rn.x * p->psa;
    V2 xy;
    if (p->inner_ellipse_0.x > K(0.0)) {
        V2 oe = v2(K(0.0), K(0.0)), d0 = v2(K(0.0), K(0.0));
        for (uint32_t i = 0; i != MAXV; ++i) {
            if (i > 0) target -= p->sector[i - 1];
            oe = p->ellipses[i];
            d0 = p->vertices[i];
            if ((i >= 2 && i + 1 == p->vertex_count) || target < p->sector[i]) break;
        }
        R sdet = SQRT(ellipse_det(oe));
        R angle = K(2.0) * target * sdet;
        V2 a = sclf2(COS(angle) * sdet, d0);
        V2 b = sclf2(SIN(angle), rotate_90(ellipse_transform(oe, d0)));
        xy = add2(a, b);
        xy = scl2(xy, SQRT(rn.y / ellipse_dir_factor_rsq(oe, xy)));
    } else {
        V2 ie = p->inner_ellipse_0, oe = v2(K(0.0), K(0.0)), d0 = v2(K(0.0), K(0.0)), d1 = v2(K(0.0), K(0.0));
        R spsa = K(0.0);
        for (uint32_t i = 0; i != MAXV - 1; ++i) {
            V2 ve = p->ellipses[i];
            if (i == 0) {
                oe = ve;
            } else {
                target -= p->sector[i - 1];
                int vin = is_inner_ellipse(ve);
                ie = vin ? ve : ie;
                oe = vin ? oe : ve;
            }
            d0 = p->vertices[i];
            d1 = p->vertices[i + 1];
            spsa = p->sector[i];
            if ((i >= 1 && i + 2 == p->vertex_count) || target < spsa) break;
        }
        V2 r = v2(target / spsa, rn.y);
        xy = sample_sector(r, target, ie, oe, d0, d1, 2u);
    }
    R z = SQRT(FMAX(K(0.0), FMA(-xy.x, xy.x, FMA(-xy.y, xy.y, K(1.0)))));
    return v3(xy.x, xy.y, z);
}

/* ---------------- BRDF (brdfs.glsl:43-93) ------------------------------- */
static V3 evaluate_brdf(const shading_t* sd, V3 li_dir) {
    V3 h = normalize3(add3(li_dir, sd->outgoing));
    R li = dot3(sd->normal, li_dir);
    R odh = dot3(sd->outgoing, h);
    V3 brdf = v3(K(0.0), K(0.0), K(0.0));
    brdf = add3(brdf, sd->diffuse_albedo);
    R ndh = dot3(sd->normal, h);
    R r2 = sd->roughness * sd->roughness;
    R ggx = FMA(FMA(ndh, r2, -ndh), ndh, K(1.0));
    ggx = r2 / (ggx * ggx);
    R lo = sd->lambert_outgoing;
    R masking = li * SQRT(FMA(FMA(-lo, r2, lo), lo, r2));
    R shadowing = lo * SQRT(FMA(FMA(-li, r2, li), li, r2));
    R smith = K(0.5) / (masking + shadowing);
    R c = clampr(odh, K(0.0), K(1.0));
    R fl = K(1.0) - c;
    R fs = fl * fl;
    R f5 = fs * fl * fs;
    V3 F = v3(sd->fresnel_0.x + (K(1.0) - sd->fresnel_0.x) * f5, sd->fresnel_0.y + (K(1.0) - sd->fresnel_0.y) * f5,
              sd->fresnel_0.z + (K(1.0) - sd->fresnel_0.z) * f5);
    R spec = ggx * smith * dot3(F, v3(K(0.21263901), K(0.71516868), K(0.07219232)));
    brdf = add3(brdf, v3(spec, spec, spec));
    return scl3(brdf, R_INV_PI);
}

/* ---------------- LTC table (ltc_utility.glsl:56-105) ------------------- */
static R ltc_fetch(const orc_scene_t* sc, int table, int channel, R u, R v, R w) {
    int W = (int)sc->ltc_res_roughness, H = (int)sc->ltc_res_inclination, L = (int)sc->ltc_fresnel_count;
    int C = table == 0 ? 4 : 2;
    const uint16_t* q = table == 0 ? sc->ltc_rgba16 : sc->ltc_rg16;
    R lw = RINT(w);
    int layer = lw < K(0.0) ? 0 : (lw > (R)(L - 1) ? L - 1 : (int)lw);
    R x = u * (R)W - K(0.5), y = v * (R)H - K(0.5);
    R x0 = FLOOR(x), y0 = FLOOR(y);
    R fx = x - x0, fy = y - y0;
    int ix = (int)x0, iy = (int)y0;
    int i0 = ix < 0 ? 0 : (ix > W - 1 ? W - 1 : ix), i1 = ix + 1 < 0 ? 0 : (ix + 1 > W - 1 ? W - 1 : ix + 1);
    int j0 = iy < 0 ? 0 : (iy > H - 1 ? H - 1 : iy), j1 = iy + 1 < 0 ? 0 : (iy + 1 > H - 1 ? H - 1 : iy + 1);
#define TEX(i, j) ((R)q[(((size_t)layer * H + (j)) * W + (i)) * C + channel] / K(65535.0))
    R t00 = TEX(i0, j0), t10 = TEX(i1, j0), t01 = TEX(i0, j1), t11 = TEX(i1, j1);
#undef TEX
    R a = FMA(fx, t10 - t00, t00);
    R b = FMA(fx, t11 - t01, t01);
    return FMA(fy, b - a, a);
}

static void get_ltc_coefficients(const orc_scene_t* sc, const risltc_ltc_constants_t* k, R f0, R roughness,
                                 V3 pos, V3 n, V3 o, ltc_t* ltc) {
    R ndo = dot3(n, o);
    R incl = ACOS(clampr(ndo, K(0.0), K(1.0)));
    R u = FMA(SQRT(clampr(roughness, K(0.0), K(1.0))), (R)k->roughness_factor, (R)k->roughness_summand);
    R v = FMA(incl, (R)k->inclination_factor, (R)k->inclination_summand);
    R w = FMA(clampr(f0, K(0.0), K(1.0)), (R)k->fresnel_index_factor, (R)k->fresnel_index_summand);
    R d0x = ltc_fetch(sc, 0, 0, u, v, w), d0y = ltc_fetch(sc, 0, 1, u, v, w);
    R d0z = ltc_fetch(sc, 0, 2, u, v, w), d0w = ltc_fetch(sc, 0, 3, u, v, w);
    R d1x = ltc_fetch(sc, 1, 0, u, v, w), d1y = ltc_fetch(sc, 1, 1, u, v, w);
    /* mat3(d0.x, 0, -d0.y,  0, d0.z, 0,  d0.w, 0, d1.x), column-major */
    M33 s2c = {{{d0x, K(0.0), d0w}, {K(0.0), d0z, K(0.0)}, {-d0y, K(0.0), d1x}}};
    ltc->shading_to_cosine = s2c;
    ltc->albedo = d1y;
    R det2 = d0x * d1x + d0y * d0w;
    ltc->determinant = d0z * det2;
    R inv = K(1.0) / det2;
    M33 c2s = {{{d1x * inv, K(0.0), -d0w * inv}, {K(0.0), K(1.0) / d0z, K(0.0)}, {d0y * inv, K(0.0), d0x * inv}}};
    ltc->cosine_to_shading = c2s;
    V3 xa = normalize3(v3(FMA(-ndo, n.x, o.x), FMA(-ndo, n.y, o.y), FMA(-ndo, n.z, o.z)));
    V3 ya = cross3(n, xa);
    V3 rows[3] = {xa, ya, n};
    for (int i = 0; i != 3; ++i) {
        ltc->world_to_shading.m[i][0] = (&rows[i].x)[0];
        ltc->world_to_shading.m[i][1] = (&rows[i].x)[1];
        ltc->world_to_shading.m[i][2] = (&rows[i].x)[2];
        R rp = FMA(-(&rows[i].x)[2], pos.z, FMA(-(&rows[i].x)[1], pos.y, -(&rows[i].x)[0] * pos.x));
        ltc->world_to_shading.m[i][3] = rp;
    }
}

/* evaluate_ltc_density (ltc_utility.glsl:100-105) */
static R ltc_density(const ltc_t* ltc, V3 d, R rcp_psa) {
    V3 dc = mul33(&ltc->shading_to_cosine, d);
    R l2 = dot3(dc, dc);
    R dens = FMAX(K(0.0), dc.z) * ltc->determinant / (l2 * l2);
    return dens * rcp_psa;
}

/* ---------------- G-buffer (shading_pass.frag.glsl:571-672) ------------- */
static V3 decode_normal(uint16_t qx, uint16_t qy) {
    R factor = K(2.0) * (K(65534.0) / K(65535.0));
    R summand = -(K(32768.0) / K(65535.0)) * factor;
    R ox = FMA((R)qx / K(65535.0), factor, summand);
    R oy = FMA((R)qy / K(65535.0), factor, summand);
    V3 n = v3(ox, oy, K(1.0) - FABS(ox) - FABS(oy));
    R sx = (ox >= K(0.0)) ? K(1.0) : K(-1.0);
    R sy = (oy >= K(0.0)) ? K(1.0) : K(-1.0);
    if (n.z < K(0.0)) {
        R nx = (K(1.0) - FABS(n.y)) * sx;
        R ny = (K(1.0) - FABS(n.x)) * sy;
        n.x = nx;
        n.y = ny;
    }
    return normalize3(n);
}

static V3 decode_position(uint32_t q0, uint32_t q1, const float f[3], const float s[3]) {
    uint32_t qx = q0 & 0x1FFFFFu;
    uint32_t qy = ((q0 & 0xFFE00000u) >> 21) | ((q1 & 0x3FFu) << 11);
    uint32_t qz = (q1 & 0x7FFFFC00u) >> 10;
    return v3(FMA((R)qx, (R)f[0], (R)s[0]), FMA((R)qy, (R)f[1], (R)s[1]), FMA((R)qz, (R)f[2], (R)s[2]));
}

static void get_shading_data(const orc_scene_t* sc, const risltc_frame_constants_t* c, uint32_t prim,
                             V3 ray_dir, shading_t* sd) {
    V3 p[3], nrm[3];
    R tc[3][2];
    for (int i = 0; i != 3; ++i) {
        size_t vi = (size_t)prim * 3 + i;
        p[i] = decode_position(sc->positions[2 * vi], sc->positions[2 * vi + 1], c->mesh_dequantization_factor,
                               c->mesh_dequantization_summand);
        const uint16_t* nt = sc->normals_uv + 4 * vi;
        nrm[i] = decode_normal(nt[0], nt[1]);
        tc[i][0] = FMA((R)nt[2] / K(65535.0), K(8.0), K(0.0));
        tc[i][1] = FMA((R)nt[3] / K(65535.0), K(-8.0), K(1.0));
    }
    V3 org = v3(c->camera_position_world_space[0], c->camera_position_world_space[1], c->camera_position_world_space[2]);
    V3 e0 = sub3(p[1], p[0]), e1 = sub3(p[2], p[0]);
    V3 rce1 = cross3(ray_dir, e1);
    R rcp_det = K(1.0) / dot3(e0, rce1);
    V3 r0 = sub3(org, p[0]);
    R d0 = dot3(r0, rce1);
    R by = rcp_det * d0;
    V3 e0c0 = cross3(e0, r0);
    R d1 = dot3(ray_dir, e0c0);
    R bz = -rcp_det * d1;
    R bx = K(1.0) - (by + bz);
    sd->position = v3(FMA(bx, p[0].x, FMA(by, p[1].x, bz * p[2].x)), FMA(bx, p[0].y, FMA(by, p[1].y, bz * p[2].y)),
                      FMA(bx, p[0].z, FMA(by, p[1].z, bz * p[2].z)));
    V3 ni = normalize3(v3(FMA(bx, nrm[0].x, FMA(by, nrm[1].x, bz * nrm[2].x)),
                          FMA(bx, nrm[0].y, FMA(by, nrm[1].y, bz * nrm[2].y)),
                          FMA(bx, nrm[0].z, FMA(by, nrm[1].z, bz * nrm[2].z))));
    uint32_t mi = sc->material_indices[prim];
    if (mi >= sc->material_count) mi = sc->material_count - 1;
    const risltc_material_t* m = &sc->materials[mi];
    V3 base = v3(m->base_color[0], m->base_color[1], m->base_color[2]);
    R metal = m->specular[2];
    sd->diffuse_albedo = v3(FMA(base.x, -metal, base.x), FMA(base.y, -metal, base.y), FMA(base.z, -metal, base.z));
    R om = K(1.0) - metal;
    sd->fresnel_0 = v3(K(0.02) * om + base.x * metal, K(0.02) * om + base.y * metal, K(0.02) * om + base.z * metal);
    R lr = m->specular[1];
    R rough = lr * lr;
    sd->roughness = clampr(rough * (R)c->roughness_factor, K(0.0064), K(1.0));
    /* flat normal-map texel (0.5, 0.5): tangent-space normal (0, 0, 1) */
    R nts_x = FMA(K(0.5), K(2.0), K(-1.0)), nts_y = nts_x;
    R nts_z = SQRT(FMAX(K(0.0), FMA(-nts_x, nts_x, FMA(-nts_y, nts_y, K(1.0)))));
    R te0x = tc[1][0] - tc[0][0], te0y = tc[1][1] - tc[0][1];
    R te1x = tc[2][0] - tc[0][0], te1y = tc[2][1] - tc[0][1];
    V3 nce0 = cross3(ni, e0);
    V3 e1cn = cross3(e1, ni);
    V3 tangent = add3(scl3(e1cn, te0x), scl3(nce0, te1x));
    V3 bitangent = add3(scl3(e1cn, te0y), scl3(nce0, te1y));
    R mtl = SQRT(K(0.5) * (dot3(tangent, tangent) + dot3(bitangent, bitangent)));
    nts_z *= FMAX(K(1.0e-10), mtl);
    V3 nw = v3(FMA(ni.x, nts_z, FMA(bitangent.x, nts_y, tangent.x * nts_x)),
               FMA(ni.y, nts_z, FMA(bitangent.y, nts_y, tangent.y * nts_x)),
               FMA(ni.z, nts_z, FMA(bitangent.z, nts_y, tangent.z * nts_x)));
    sd->normal = normalize3(nw);
    sd->outgoing = normalize3(sub3(org, sd->position));
    R off = FMAX(K(0.0), K(1.0e-3) - dot3(sd->normal, sd->outgoing));
    sd->normal = v3(FMA(off, sd->outgoing.x, sd->normal.x), FMA(off, sd->outgoing.y, sd->normal.y),
                    FMA(off, sd->outgoing.z, sd->normal.z));
    sd->normal = normalize3(sd->normal);
    sd->lambert_outgoing = dot3(sd->normal, sd->outgoing);
}

/* ---------------- MIS (shading_pass.frag.glsl:219-269) ------------------ */
static V3 mis_estimate(uint32_t heuristic, V3 integrand, V3 sw, R sdens, V3 ow, R odens, R ve) {
    if (heuristic == RISLTC_MIS_WEIGHTED) {
        V3 ws = add3(scl3(sw, sdens), scl3(ow, odens));
        V3 num = mul3(sw, integrand);
        return v3(num.x / ws.x, num.y / ws.y, num.z / ws.z);
    }
    if (heuristic == RISLTC_MIS_OPTIMAL_CLAMPED || heuristic == RISLTC_MIS_OPTIMAL) {
        R bw = K(1.0) / (sdens + odens);
        V3 ws = add3(scl3(sw, sdens), scl3(ow, odens));
        if (heuristic == RISLTC_MIS_OPTIMAL_CLAMPED) {
            V3 wwd = v3(sw.x / ws.x, sw.y / ws.y, sw.z / ws.z);
            R mixed0 = FMA(-ve, bw, bw);
            V3 mixed = v3(FMA(ve, wwd.x, mixed0), FMA(ve, wwd.y, mixed0), FMA(ve, wwd.z, mixed0));
            return mul3(mixed, integrand);
        }
        V3 a = scl3(sw, ve);
        V3 d = sub3(integrand, scl3(ws, ve));
        return add3(a, sclf3(bw, d));
    }
    R wod;
    if (heuristic == RISLTC_MIS_BALANCE)
        wod = K(1.0) / (sdens + odens);
    else
        wod = sdens / (sdens * sdens + odens * odens);
    return sclf3(wod, integrand);
}

/* ---------------- shading context ----------------------------------------- */
typedef struct {
    const orc_scene_t* sc;
    const risltc_frame_constants_t* c;
    const risltc_settings_t* s;
    shading_t sd;
    ltc_t ltc;
    A34 w2c;      /* shading_to_cosine * world_to_shading       */
    A34 w2s_flip; /* world_to_shading with row 1 negated          */
    A34 w2c_flip; /* shading_to_cosine * w2s_flip                */
    uint32_t seed;
} pix_t;

/* evaluate_polygonal_light_shading, SAMPLE_POLYGON_LTC_CP branch
 * (shading_pass.frag.glsl:421-456, 555-557): the RIS target */
static V3 eval_ltc_cp(const pix_t* px, const light_t* l, uint32_t* clip_counts) {
    R side = plane_side(px->sd.position, l->plane);
    const A34* w2s = (side < K(0.0)) ? &px->w2s_flip : &px->ltc.world_to_shading;
    const A34* w2c = (side < K(0.0)) ? &px->w2c_flip : &px->w2c;
    V3 vs[MAXV];
    for (uint32_t i = 0; i != l->vertex_count; ++i) vs[i] = xform(w2s, l->v[i]);
    uint32_t vcd = clip_polygon(l->vertex_count, vs);
    V3 result = v3(K(0.0), K(0.0), K(0.0));
    V3 diffuse = v3(K(0.0), K(0.0), K(0.0));
    if (vcd > 0) diffuse = mul3(sclf3(calculate_ltc(vcd, vs), px->sd.diffuse_albedo), l->radiance);
    result = add3(result, diffuse);
    for (uint32_t i = 0; i != l->vertex_count; ++i) vs[i] = xform(w2c, l->v[i]);
    uint32_t vcs = clip_polygon(l->vertex_count, vs);
    V3 ggx = v3(K(0.0), K(0.0), K(0.0));
    if (vcs > 0) ggx = sclf3(calculate_ltc(vcs, vs) * px->ltc.albedo, l->radiance);
    result = add3(result, ggx);
    if (clip_counts) *clip_counts = vcd | (vcs << 8);
    return result;
}

static V2 noise2(uint32_t* seed) {
    R a = (R)orc_noise_next(seed);
    R b = (R)orc_noise_next(seed);
    return v2(a, b);
}

/* evaluate_polygonal_light_shading_peters (shading_pass.frag.glsl:292-397),
 * with visibility through the CPU LBVH: the polygons are prepared once and
 * the SAMPLE_COUNT loop (365-396) draws one direction per technique per
 * iteration; the sum is scaled by 1 / SAMPLE_COUNT. rec (may be NULL)
 * receives the per-technique intermediate values of the first iteration.
 *
 * The same code is the SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE branch of
 * evaluate_polygonal_light_shading (shading_pass.frag.glsl:458-551): with
 * trace = 1 (visibility = true, the final sample) an occluded sample's
 * integrand is 0, exactly as above; with trace = 0 (the RIS candidates,
 * dummy_vis = false) no ray is traced and a sample above the horizon keeps
 * its unshadowed integrand (get_polygon_radiance_brdf_product, 200-208).
 * Both end with "result * (1.0f / SAMPLE_COUNT)" (396, 559). clip_counts (may be NULL) receives vcd | vcs << 8. */
static V3 eval_peters_ex(pix_t* px, const light_t* l, float* rec, int trace, uint32_t* clip_counts) {
    V3 result = v3(K(0.0), K(0.0), K(0.0));
    R side = plane_side(px->sd.position, l->plane);
    const A34* w2s = (side < K(0.0)) ? &px->w2s_flip : &px->ltc.world_to_shading;
    const A34* w2c = (side < K(0.0)) ? &px->w2c_flip : &px->w2c;
    psa_t pd, ps;
    if (rec) {
        rec[ORC_REC_POS + 0] = (float)px->sd.position.x;
        rec[ORC_REC_POS + 1] = (float)px->sd.position.y;
        rec[ORC_REC_POS + 2] = (float)px->sd.position.z;
    }
    V3 vs[MAXV];
    for (uint32_t i = 0; i != l->vertex_count; ++i) vs[i] = xform(w2s, l->v[i]);
    uint32_t vcd = clip_polygon(l->vertex_count, vs);
    if (rec) memcpy(&rec[ORC_REC_CLIPD], &vcd, 4);
    if (clip_counts) *clip_counts = vcd;
    if (vcd == 0) return result;
    prepare_psa(&pd, vcd, vs);
    for (uint32_t i = 0; i != l->vertex_count; ++i) vs[i] = xform(w2c, l->v[i]);
    uint32_t vcs = clip_polygon(l->vertex_count, vs);
    if (rec) memcpy(&rec[ORC_REC_CLIPS], &vcs, 4);
    if (clip_counts) *clip_counts = vcd | (vcs << 8);
    if (vcs == 0) {
        memset(&ps, 0, sizeof(ps));
        ps.psa = K(0.0);
    } else {
        prepare_psa(&ps, vcs, vs);
    }
    if (rec) {
        rec[ORC_REC_PSAD] = (float)pd.psa;
        rec[ORC_REC_PSAS] = (float)ps.psa;
    }
    if (pd.psa == K(0.0)) return result;
    R spec_w = px->ltc.albedo * ps.psa;
    V3 da = v3(FMAX(px->sd.diffuse_albedo.x, K(0.01)), FMAX(px->sd.diffuse_albedo.y, K(0.01)),
               FMAX(px->sd.diffuse_albedo.z, K(0.01)));
    V3 dw = scl3(da, pd.psa);
    uint32_t techniques = (ps.psa > K(0.0)) ? 2u : 1u;
    R rcp_d = K(1.0) / pd.psa;
    R rcp_s = K(1.0) / ps.psa;
    V3 sw = v3(spec_w, spec_w, spec_w);
    uint32_t heuristic = px->s->mis_heuristic;
    if (heuristic == RISLTC_MIS_OPTIMAL) {
        V3 rop = scl3(l->radiance, R_INV_PI);
        dw = mul3(dw, rop);
        sw = mul3(sw, rop);
    }
    R ve = (R)px->c->mis_visibility_estimate;
    const uint32_t iters = px->s->sample_count;
    for (uint32_t it = 0; it != iters; ++it) {
        V3 dirs[2];
        dirs[0] = sample_psa(&pd, noise2(&px->seed));
        dirs[1] = v3(K(0.0), K(0.0), K(0.0));
        if (ps.psa > K(0.0)) {
            dirs[1] = sample_psa(&ps, noise2(&px->seed));
            dirs[1] = normalize3(mul33(&px->ltc.cosine_to_shading, dirs[1]));
        }
        for (uint32_t j = 0; j != techniques; ++j) {
            V3 d = dirs[j];
            if (d.z <= K(0.0)) continue;
            R dd = d.z * rcp_d;
            R sdn = ltc_density(&px->ltc, d, rcp_s);
            V3 wdir = mul_transposed(w2s, d);
            int lambert_ok = dot3(px->sd.normal, wdir) > K(0.0);
            int vis = 0;
            uint32_t state = 1;
            R tmax = K(0.0);
            if (lambert_ok) {
                tmax = -plane_side(px->sd.position, l->plane) / dot3(wdir, v3(l->plane[0], l->plane[1], l->plane[2]));
                tmax -= K(1e-3);
                if (trace && px->s->shadow_rays && px->sc->bvh) {
                    float o[3] = {(float)px->sd.position.x, (float)px->sd.position.y, (float)px->sd.position.z};
                    float dir[3] = {(float)wdir.x, (float)wdir.y, (float)wdir.z};
                    int hit = orc_bvh_any_hit(px->sc->bvh, o, dir, 1.0e-3f, (float)tmax, NULL);
                    vis = !hit;
                    state = hit ? 2u : 3u;
                } else {
                    vis = 1;
                    state = 4u;
                }
            }
            V3 rb_vis = mul3(l->radiance, evaluate_brdf(&px->sd, wdir));
            V3 integ_vis = sclf3(d.z, rb_vis);
            V3 integ_occ = sclf3(d.z, v3(K(0.0), K(0.0), K(0.0)));
            V3 e_vis, e_occ;
            if (j == 0 && ps.psa <= K(0.0)) {
                R rd = K(1.0) / dd;
                e_vis = scl3(integ_vis, rd);
                e_occ = v3(K(0.0), K(0.0), K(0.0));
            } else if (j == 0) {
                e_vis = mis_estimate(heuristic, integ_vis, dw, dd, sw, sdn, ve);
                e_occ = mis_estimate(heuristic, integ_occ, dw, dd, sw, sdn, ve);
            } else {
                e_vis = mis_estimate(heuristic, integ_vis, sw, sdn, dw, dd, ve);
                e_occ = mis_estimate(heuristic, integ_occ, sw, sdn, dw, dd, ve);
            }
            result = add3(result, vis ? e_vis : e_occ);
            if (rec && it == 0) {
                rec[ORC_REC_DIR + 3 * j + 0] = (float)wdir.x;
                rec[ORC_REC_DIR + 3 * j + 1] = (float)wdir.y;
                rec[ORC_REC_DIR + 3 * j + 2] = (float)wdir.z;
                rec[ORC_REC_TMAX + j] = (float)tmax;
                memcpy(&rec[ORC_REC_STATE + j], &state, 4);
                rec[ORC_REC_EVIS + 3 * j + 0] = (float)e_vis.x;
                rec[ORC_REC_EVIS + 3 * j + 1] = (float)e_vis.y;
                rec[ORC_REC_EVIS + 3 * j + 2] = (float)e_vis.z;
                rec[ORC_REC_EOCC + 3 * j + 0] = (float)e_occ.x;
                rec[ORC_REC_EOCC + 3 * j + 1] = (float)e_occ.y;
                rec[ORC_REC_EOCC + 3 * j + 2] = (float)e_occ.z;
            }
        }
    }
    return scl3(result, K(1.0) / (R)iters);
}

static V3 eval_peters(pix_t* px, const light_t* l, float* rec) { return eval_peters_ex(px, l, rec, 1, NULL); }

/* sample_area_polygon_turk (polygon_sampling_related_work.glsl:34-43): a
 * uniform point on the triangle of vertices 0, 1, 2 (the reference samples
 * only this triangle of the polygon). */
static V3 turk_point(const light_t* l, V2 u) {
    R s = SQRT(u.x);
    R b0 = K(1.0) - s, b1 = s * u.y, b2 = FMA(-s, u.y, s);
    return v3(b0 * l->v[0].x + b1 * l->v[1].x + b2 * l->v[2].x, b0 * l->v[0].y + b1 * l->v[1].y + b2 * l->v[2].y,
              b0 * l->v[0].z + b1 * l->v[1].z + b2 * l->v[2].z);
}

/* get_polygon_area (shading_pass.frag.glsl:148-152): the area of the
 * triangle of vertices 0, 1, 2. */
static R turk_area(const light_t* l) {
    V3 c = cross_stable(sub3(l->v[1], l->v[0]), sub3(l->v[2], l->v[0]));
    return SQRT(dot3(c, c)) / K(2.0);
}

/* The SAMPLE_POLYGON_AREA_TURK branch of evaluate_polygonal_light_shading
 * (shading_pass.frag.glsl:410-418) for SAMPLE_COUNT = 1 at the point
 * `sample`: get_area_sample_density (polygon_sampling_related_work.glsl:
 * 57-64) and get_polygonal_light_mis_estimate (shading_pass.frag.glsl:
 * 281-287), without visibility. Returns the unshadowed estimate; *dir and
 * *tmax receive the shadow ray of get_polygon_visibility (112-129) and
 * *above whether the sample is above the shading horizon. */
static V3 eval_turk(const pix_t* px, const light_t* l, V3 sample, V3* dir, R* tmax, int* above) {
    V3 d = sub3(sample, px->sd.position);
    R d2 = dot3(d, d);
    d = scl3(d, rsq(d2));
    V3 pn = v3(l->plane[0], l->plane[1], l->plane[2]);
    R density = d2 / (FABS(dot3(pn, d)) * turk_area(l));
    R lambert = dot3(px->sd.normal, d);
    *dir = d;
    *above = lambert > K(0.0);
    *tmax = K(0.0);
    V3 rb = v3(K(0.0), K(0.0), K(0.0));
    if (*above) {
        *tmax = -plane_side(px->sd.position, l->plane) / dot3(d, pn);
        *tmax -= K(1e-3);
        rb = mul3(l->radiance, evaluate_brdf(&px->sd, d));
    }
    return (density > K(0.0)) ? scl3(rb, lambert / density) : v3(K(0.0), K(0.0), K(0.0));
}

/* Final sample of the Turk technique: the SAMPLE_COUNT loop of
 * evaluate_polygonal_light_shading (shading_pass.frag.glsl:412-417) with
 * visibility. draw = 1 (uniform light sampling, 714-720): every iteration
 * draws its own point; draw = 0 (RIS, 743-756): every iteration evaluates
 * the reservoir's point `sample`. The loop shares one visibility flag: the
 * estimate counts only when no shadow ray is occluded (visibility, once
 * false, stays false and the caller multiplies by it). SAMPLE_COUNT = 1:
 * the estimate times the visibility of its ray, zero below the horizon.
 * Returns that estimate; rec as for eval_peters (technique slot 0, the
 * first iteration). */
static V3 final_turk(pix_t* px, const light_t* l, V3 sample, float* rec, int draw) {
    const uint32_t iters = px->s->sample_count;
    V3 sum = v3(K(0.0), K(0.0), K(0.0)), first = sum;
    int all_vis = 1, first_vis = 0;
    for (uint32_t it = 0; it != iters; ++it) {
        if (draw) sample = turk_point(l, noise2(&px->seed));
        V3 dir;
        R tmax;
        int above;
        V3 col = eval_turk(px, l, sample, &dir, &tmax, &above);
        uint32_t state = 1;
        int vis = 0;
        if (above) {
            if (px->s->shadow_rays && px->sc->bvh) {
                float o[3] = {(float)px->sd.position.x, (float)px->sd.position.y, (float)px->sd.position.z};
                float fd[3] = {(float)dir.x, (float)dir.y, (float)dir.z};
                int hit = orc_bvh_any_hit(px->sc->bvh, o, fd, 1.0e-3f, (float)tmax, NULL);
                vis = !hit;
                state = hit ? 2u : 3u;
            } else {
                vis = 1;
                state = 4u;
            }
            if (!vis) all_vis = 0;
        }
        if (it == 0) {
            first = col;
            first_vis = vis;
            if (rec) {
                rec[ORC_REC_POS + 0] = (float)px->sd.position.x;
                rec[ORC_REC_POS + 1] = (float)px->sd.position.y;
                rec[ORC_REC_POS + 2] = (float)px->sd.position.z;
                rec[ORC_REC_DIR + 0] = (float)dir.x;
                rec[ORC_REC_DIR + 1] = (float)dir.y;
                rec[ORC_REC_DIR + 2] = (float)dir.z;
                rec[ORC_REC_TMAX] = (float)tmax;
                memcpy(&rec[ORC_REC_STATE], &state, 4);
                rec[ORC_REC_EVIS + 0] = (float)col.x;
                rec[ORC_REC_EVIS + 1] = (float)col.y;
                rec[ORC_REC_EVIS + 2] = (float)col.z;
            }
        }
        sum = add3(sum, col);
    }
    if (iters == 1) return first_vis ? first : v3(K(0.0), K(0.0), K(0.0));
    return all_vis ? scl3(sum, K(1.0) / (R)iters) : v3(K(0.0), K(0.0), K(0.0));
}

/* int(u * N) in float arithmetic in both builds (u is a float draw), so the
 * FP32 and FP64 oracles pick the same light; clamped to N - 1 (hazard H2) */
static uint32_t light_index_from(R u, uint32_t n) {
    int idx = (int)((float)u * (float)n);
    if (idx > (int)n - 1) idx = (int)n - 1;
    return (uint32_t)idx;
}

/* Learned RIS source (project extension; no reference code). Table of the
 * pixel: group ids t[0..15], weights t[16..31], T = their sum. Group g is the
 * run of clusters g << G .. ((g + 1) << G) - 1, G = proposal_group_log2 (0:
 * single clusters; 2: quads of the 1024-output network), whose lights are
 * the consecutive run cluster_lights[offsets[g << G] .. offsets[min((g + 1)
 * << G, count)]). Two draws a, b per candidate; every decision in float, as
 * light_index_from:
 *   T == 0 -> light_index_from(a); a < 1/2 -> light_index_from(2a);
 *   else slot s = first with (int)((2a - 1) T) (clamped to T - 1) <
 *   w_0 + .. + w_s, group id_s, light number min(size - 1, (int)(b size))
 *   of that group (an empty group or an id past the groups:
 *   light_index_from(b); learned_density counts that mass as uniform).
 * parity: bitwise against the kernel's learned shading given the same tables
 * (tests/test_gpu_learned.py); the mixture is an exact pdf of this draw
 * (tests/test_oracle_estimator.py::test_learned_draw_density). */
static uint32_t table_id(const uint8_t* t, int s) { return t[s]; }
static uint32_t table_weight(const uint8_t* t, int s) { return t[16 + s]; }

static uint32_t group_count(const orc_scene_t* sc) {
    uint32_t G = sc->proposal_group_log2;
    return (sc->cluster_count + (1u << G) - 1u) >> G;
}

/* [first, end) of group g's lights in cluster_lights (g < group_count) */
static void group_range(const orc_scene_t* sc, uint32_t g, uint32_t* first, uint32_t* end) {
    uint32_t G = sc->proposal_group_log2;
    uint32_t c1 = (g + 1u) << G;
    if (c1 > sc->cluster_count) c1 = sc->cluster_count;
    *first = sc->cluster_offsets[g << G];
    *end = sc->cluster_offsets[c1];
}

static uint32_t learned_draw(const orc_scene_t* sc, const uint8_t* t, uint32_t T, uint32_t* seed, uint32_t N) {
    float a = orc_noise_next(seed);
    float b = orc_noise_next(seed);
    if (T == 0u) return light_index_from((R)a, N);
    if (a < 0.5f) return light_index_from((R)(a + a), N);
    float u = (a + a) - 1.0f;
    uint32_t k = (uint32_t)(u * (float)T);
    if (k > T - 1u) k = T - 1u;
    uint32_t acc = 0, g = 0;
    for (int s = 0; s != 16; ++s) {
        if (k < acc + table_weight(t, s)) {
            g = table_id(t, s);
            break;
        }
        acc += table_weight(t, s);
    }
    if (g >= group_count(sc)) return light_index_from((R)b, N);
    uint32_t off, end;
    group_range(sc, g, &off, &end);
    uint32_t size = end - off;
    if (size == 0u) return light_index_from((R)b, N);
    uint32_t j = (uint32_t)(b * (float)size);
    if (j > size - 1u) j = size - 1u;
    return sc->cluster_lights[off + j];
}

/* Weight of the slots that learned_draw sends to a uniform light: ids past
 * the groups, or of empty groups. */
static uint32_t learned_fallback_weight(const orc_scene_t* sc, const uint8_t* t) {
    uint32_t bad = 0;
    for (int s = 0; s != 16; ++s) {
        uint32_t w = table_weight(t, s), g = table_id(t, s);
        if (w == 0u) continue;
        int empty = g >= group_count(sc);
        if (!empty) {
            uint32_t off, end;
            group_range(sc, g, &off, &end);
            empty = off == end;
        }
        if (empty) bad += w;
    }
    return bad;
}

/* q = 1/(2N) + T_bad / (2 T N) + w_g / (2 T |g|): w_g sums the weights of
 * every slot holding the light's group (0 when none does), T_bad is
 * learned_fallback_weight (its term is skipped when 0) */
static R learned_density(const orc_scene_t* sc, const uint8_t* t, uint32_t T, uint32_t T_bad, uint32_t idx,
                         uint32_t N) {
    if (T == 0u) return K(1.0) / (R)N;
    R q = K(0.5) / (R)N;
    if (T_bad != 0u) q = q + (K(0.5) * (R)T_bad) / ((R)T * (R)N);
    uint32_t g = sc->cluster_of_light[idx] >> sc->proposal_group_log2, wg = 0;
    for (int s = 0; s != 16; ++s)
        if (table_weight(t, s) != 0u && table_id(t, s) == g) wg += table_weight(t, s);
    if (wg != 0u) {
        uint32_t off, end;
        group_range(sc, g, &off, &end);
        q = q + (K(0.5) * (R)wg) / ((R)T * (R)(end - off));
    }
    return q;
}

/* Learned RIS target (project extension; no reference code): with a
 * visibility cache, the learned pass's target of light l is
 * p_hat * (f + v * (1 - f) / 255), f = 2^-settings.learned_floor_log2
 * (0 -> 1/8), v = the cache's byte for the pixel's cell and the light's slot
 * (its cluster, or the light itself; 255 while the entry is empty). The
 * factor is formed in float in both builds. parity: bitwise against the
 * kernel's learned shading given the same tables and cache
 * (tests/test_gpu_learned.py); unbiasedness against the plain estimator
 * (tests/test_oracle_estimator.py). */
static float cache_factor(const uint32_t* of_light, uint32_t idx, float floor_, float scale, const uint8_t* cache_row,
                          uint32_t per_light) {
    uint32_t v = cache_row[per_light ? idx : of_light[idx]];
    return floor_ + (float)v * scale;
}

/* Visibility-cache key of a shading point (project extension): the cell
 * floor(p / cell) per axis (clamped to +-1e9 first) and the dominant axis
 * and sign of the shading normal (the first axis on ties: 0/1 = +x/-x,
 * 2/3 = y, 4/5 = z), combined by MurmurHash3's mix and finaliser and masked
 * to the table. Formed in float in both builds. parity: bitwise against the
 * kernel's debug records (tests/test_gpu_learned.py); the cell and axis
 * rules by construction in tests/test_oracle_estimator.py. */
static uint32_t cache_key(const orc_scene_t* sc, V3 pos, V3 nrm) {
    const float lim = 1.0e9f, inv = sc->cache_inv_cell;
    float p[3] = {(float)pos.x, (float)pos.y, (float)pos.z};
    int32_t ix[3];
    for (int a = 0; a != 3; ++a) ix[a] = (int32_t)floorf(fminf(fmaxf(p[a] * inv, -lim), lim));
    float nx = (float)nrm.x, ny = (float)nrm.y, nz = (float)nrm.z;
    float ax = fabsf(nx), ay = fabsf(ny), az = fabsf(nz);
    uint32_t code = (ax >= ay && ax >= az) ? (nx < 0.0f ? 1u : 0u)
                                           : (ay >= az ? (ny < 0.0f ? 3u : 2u) : (nz < 0.0f ? 5u : 4u));
    uint32_t h = orc_murmur_mix(orc_murmur_mix(orc_murmur_mix(orc_murmur_mix(0u, (uint32_t)ix[0]), (uint32_t)ix[1]),
                                               (uint32_t)ix[2]),
                                code);
    return orc_murmur_finalize(h) & sc->cache_mask;
}

/* LTC_CP (the default) and, as baselines, AREA_TURK and PROJECTED_SOLID_ANGLE
 * with the uniform source; the learned and cached targets are LTC_CP only. */
static int orc_technique_supported(const risltc_settings_t* s) {
    uint32_t t = s->polygon_sampling_technique;
    if (t == RISLTC_SAMPLE_POLYGON_LTC_CP) return 1;
    return (t == RISLTC_SAMPLE_POLYGON_AREA_TURK || t == RISLTC_SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE) &&
           s->proposal == RISLTC_PROPOSAL_UNIFORM;
}

static void shade_one(const orc_scene_t* sc, const risltc_frame_constants_t* c, const risltc_settings_t* s,
                      uint32_t pixel, uint32_t prim, R out[4], float* samples, uint32_t* cands) {
    uint32_t W = c->viewport_size[0];
    uint32_t x = pixel % W, y = pixel / W;
    uint32_t spp = s->sample_count_light;
    uint32_t M = s->ris_candidate_count;
    if (samples) {
        for (uint32_t j = 0; j != spp; ++j) {
            float* r = samples + (size_t)j * RISLTC_SAMPLE_RECORD_FLOATS;
            memset(r, 0, sizeof(float) * RISLTC_SAMPLE_RECORD_FLOATS);
            int32_t none = -1;
            memcpy(&r[ORC_REC_LIGHT], &none, 4);
        }
    }
    if (cands) memset(cands, 0, sizeof(uint32_t) * (size_t)spp * M * RISLTC_CANDIDATE_RECORD_WORDS);
    /* ids past the mesh are treated as background (project rule; the
     * reference would read out of bounds) */
    if (prim == 0xFFFFFFFFu || ((prim >> 31) == 0u && (uint64_t)prim >= sc->triangle_count)) {
        out[0] = out[1] = out[2] = K(0.0);
        out[3] = K(1.0);
        return;
    }
    V3 fin = v3(K(0.0), K(0.0), K(0.0));
    R exposure = (R)c->exposure_factor;
    if ((prim >> 31) > 0) {
        fin = v3(K(1.0), K(1.0), K(1.0));
    } else {
        pix_t px;
        px.sc = sc;
        px.c = c;
        px.s = s;
        const float(*P)[4] = c->pixel_to_ray_direction_world_space;
        R fx = (R)x, fy = (R)y;
        V3 ray = v3(FMA((R)P[0][2], K(1.0), FMA((R)P[0][1], fy, (R)P[0][0] * fx)),
                    FMA((R)P[1][2], K(1.0), FMA((R)P[1][1], fy, (R)P[1][0] * fx)),
                    FMA((R)P[2][2], K(1.0), FMA((R)P[2][1], fy, (R)P[2][0] * fx)));
        get_shading_data(sc, c, prim, ray, &px.sd);
        R f0l = dot3(px.sd.fresnel_0, v3(K(0.2126), K(0.7152), K(0.0722)));
        get_ltc_coefficients(sc, &c->ltc_constants, f0l, px.sd.roughness, px.sd.position, px.sd.normal,
                             px.sd.outgoing, &px.ltc);
        px.w2c = mul33_34(&px.ltc.shading_to_cosine, &px.ltc.world_to_shading);
        px.w2s_flip = px.ltc.world_to_shading;
        for (int k = 0; k != 4; ++k) px.w2s_flip.m[1][k] = -px.w2s_flip.m[1][k];
        px.w2c_flip = mul33_34(&px.ltc.shading_to_cosine, &px.w2s_flip);
        px.seed = orc_noise_seed(x, y, W, c->noise_random_numbers[0]);
        const uint32_t ckey = sc->cache_vis ? cache_key(sc, px.sd.position, px.sd.normal) : 0u;
        if (samples)
            for (uint32_t j = 0; j != spp; ++j) memcpy(&samples[(size_t)j * RISLTC_SAMPLE_RECORD_FLOATS + ORC_REC_CACHE_KEY],
                                                      &ckey, 4);
        uint32_t N = sc->light_count;
        light_t light;
        const uint32_t tech = s->polygon_sampling_technique;
        if (s->light_sampling == RISLTC_LIGHT_UNIFORM) {
            for (uint32_t j = 0; j != spp; ++j) {
                float* rec = samples ? samples + (size_t)j * RISLTC_SAMPLE_RECORD_FLOATS : NULL;
                R u = (R)orc_noise_next(&px.seed);
                uint32_t idx = light_index_from(u, N);
                load_light(sc, idx, &light);
                V3 col;
                if (tech == RISLTC_SAMPLE_POLYGON_AREA_TURK) {
                    col = final_turk(&px, &light, v3(K(0.0), K(0.0), K(0.0)), rec, 1);
                } else {
                    col = eval_peters(&px, &light, rec);
                }
                V3 res = v3(col.x * (R)N, col.y * (R)N, col.z * (R)N);
                res = v3(res.x / (R)spp, res.y / (R)spp, res.z / (R)spp);
                fin = add3(fin, scl3(res, K(1.0)));
                if (rec) {
                    int32_t li = (int32_t)idx;
                    memcpy(&rec[ORC_REC_LIGHT], &li, 4);
                    rec[ORC_REC_W] = (float)N;
                    rec[ORC_REC_COLOR + 0] = (float)col.x;
                    rec[ORC_REC_COLOR + 1] = (float)col.y;
                    rec[ORC_REC_COLOR + 2] = (float)col.z;
                    memcpy(&rec[ORC_REC_SEED], &px.seed, 4);
                }
            }
        } else {
            for (uint32_t j = 0; j != spp; ++j) {
                float* rec = samples ? samples + (size_t)j * RISLTC_SAMPLE_RECORD_FLOATS : NULL;
                R w_sum = K(0.0), sample_value = K(0.0);
                int32_t chosen = -1;
                V3 chosen_point = v3(K(0.0), K(0.0), K(0.0));
                const int learned = s->proposal != RISLTC_PROPOSAL_UNIFORM;
                const uint8_t* tab =
                    s->proposal == RISLTC_PROPOSAL_LEARNED
                        ? sc->proposal_table + (size_t)pixel * 32u
                        : NULL;
                uint32_t T = 0, T_bad = 0;
                if (tab) {
                    for (int k2 = 0; k2 != 16; ++k2) T += table_weight(tab, k2);
                    T_bad = learned_fallback_weight(sc, tab);
                }
                const uint8_t* crow = learned && sc->cache_vis ? sc->cache_vis + (size_t)ckey * sc->cache_slots : NULL;
                const float lf = ldexpf(1.0f, -(int)(s->learned_floor_log2 ? s->learned_floor_log2 : 3u));
                const float ls = (1.0f - lf) / 255.0f;
                for (uint32_t i = 0; i != M; ++i) {
                    uint32_t idx;
                    if (learned) {
                        idx = learned_draw(sc, tab, T, &px.seed, N);
                    } else {
                        R u = (R)orc_noise_next(&px.seed);
                        idx = light_index_from(u, N);
                    }
                    load_light(sc, idx, &light);
                    uint32_t cc = 0;
                    V3 col, point = v3(K(0.0), K(0.0), K(0.0));
                    if (tech == RISLTC_SAMPLE_POLYGON_AREA_TURK) {
                        /* SAMPLE_COUNT points (shading_pass.frag.glsl:412-417);
                         * the last one is the reservoir's sample */
                        const uint32_t iters = s->sample_count;
                        col = v3(K(0.0), K(0.0), K(0.0));
                        for (uint32_t it = 0; it != iters; ++it) {
                            V3 dir;
                            R tmax;
                            int above;
                            point = turk_point(&light, noise2(&px.seed));
                            V3 e = eval_turk(&px, &light, point, &dir, &tmax, &above);
                            col = it == 0 ? e : add3(col, e);
                        }
                        if (iters > 1) col = scl3(col, K(1.0) / (R)iters);
                    } else if (tech == RISLTC_SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE) {
                        col = eval_peters_ex(&px, &light, NULL, 0, &cc);
                    } else {
                        col = eval_ltc_cp(&px, &light, &cc);
                    }
                    R p_hat = SQRT(dot3(col, col));
                    if (crow)
                        p_hat = p_hat * (R)cache_factor(sc->cluster_of_light, idx, lf, ls, crow, sc->cache_per_light);
                    R p = learned ? learned_density(sc, tab, T, T_bad, idx, N) : K(1.0) / (R)N;
                    R w = p_hat / p;
                    R r = (R)orc_noise_next(&px.seed);
                    w_sum += w;
                    if (w > K(0.0) && r < (w / w_sum)) {
                        chosen = (int32_t)idx;
                        sample_value = p_hat;
                        chosen_point = point;
                    }
                    if (cands) {
                        uint32_t* cr = cands + ((size_t)j * M + i) * RISLTC_CANDIDATE_RECORD_WORDS;
                        float fp = (float)p_hat, fw = (float)w;
                        cr[0] = idx;
                        cr[1] = cc;
                        memcpy(&cr[2], &fp, 4);
                        memcpy(&cr[3], &fw, 4);
                    }
                }
                if (rec) {
                    memcpy(&rec[ORC_REC_LIGHT], &chosen, 4);
                    rec[ORC_REC_WSUM] = (float)w_sum;
                    rec[ORC_REC_PHAT] = (float)sample_value;
                }
                if (chosen >= 0) {
                    load_light(sc, (uint32_t)chosen, &light);
                    V3 col = (tech == RISLTC_SAMPLE_POLYGON_AREA_TURK) ? final_turk(&px, &light, chosen_point, rec, 0)
                                                                      : eval_peters(&px, &light, rec);
                    R p_hat = sample_value;
                    R Wt = w_sum / ((R)M * p_hat);
                    if (p_hat == K(0.0)) Wt = K(0.0);
                    V3 res = v3((col.x * Wt) / (R)spp, (col.y * Wt) / (R)spp, (col.z * Wt) / (R)spp);
                    fin = add3(fin, res);
                    if (rec) {
                        rec[ORC_REC_W] = (float)Wt;
                        rec[ORC_REC_COLOR + 0] = (float)col.x;
                        rec[ORC_REC_COLOR + 1] = (float)col.y;
                        rec[ORC_REC_COLOR + 2] = (float)col.z;
                    }
                }
                if (rec) memcpy(&rec[ORC_REC_SEED], &px.seed, 4);
            }
        }
    }
    if (isnan(fin.x) || isnan(fin.y) || isnan(fin.z) || isinf(fin.x) || isinf(fin.y) || isinf(fin.z))
        fin = v3(K(1.0) / exposure, K(0.0) / exposure, K(0.8) / exposure);
    out[0] = fin.x * exposure;
    out[1] = fin.y * exposure;
    out[2] = fin.z * exposure;
    out[3] = K(1.0);
}

int SFX(orc_shade_pixels)(const orc_scene_t* sc, const risltc_frame_constants_t* c, const risltc_settings_t* s,
                          const uint32_t* visibility, const uint32_t* pixels, uint32_t count, R* out,
                          float* samples, uint32_t* cands) {
    if (!sc || !c || !s || !visibility || !out) return RISLTC_E_ARG;
    if (s->sample_count == 0 || s->sample_count > RISLTC_MAX_SAMPLE_COUNT) return RISLTC_E_ARG;
    if (!orc_technique_supported(s)) return RISLTC_E_UNSUPPORTED;
    if (s->proposal == RISLTC_PROPOSAL_LEARNED) {
        if (s->light_sampling != RISLTC_LIGHT_RESERVOIR) return RISLTC_E_UNSUPPORTED;
        if (!sc->proposal_table || !sc->cluster_of_light || !sc->cluster_offsets || !sc->cluster_lights ||
            sc->cluster_count == 0)
            return RISLTC_E_STATE;
    } else if (s->proposal == RISLTC_PROPOSAL_CACHE) {
        if (s->light_sampling != RISLTC_LIGHT_RESERVOIR) return RISLTC_E_UNSUPPORTED;
        if (!sc->cluster_of_light || !sc->cache_vis) return RISLTC_E_STATE;
    } else if (s->proposal != RISLTC_PROPOSAL_UNIFORM) {
        return RISLTC_E_ARG;
    }
    if (sc->light_count == 0) return RISLTC_E_STATE;
    g_fast_atan = s->fast_atan ? 1 : 0;
    uint32_t spp = s->sample_count_light, M = s->ris_candidate_count;
    for (uint32_t k = 0; k != count; ++k) {
        uint32_t p = pixels ? pixels[k] : k;
        shade_one(sc, c, s, p, visibility[p], out + 4 * (size_t)k,
                  samples ? samples + (size_t)k * spp * RISLTC_SAMPLE_RECORD_FLOATS : NULL,
                  cands ? cands + (size_t)k * spp * M * RISLTC_CANDIDATE_RECORD_WORDS : NULL);
    }
    return 0;
}

#ifndef ORC_F64
int orc_shade_frame_f32(const orc_scene_t* sc, const risltc_frame_constants_t* c, const risltc_settings_t* s,
                        const uint32_t* visibility, const uint32_t* pixels, uint32_t count, float* out,
                        int threads) {
    if (!sc || !c || !s || !visibility || !out) return RISLTC_E_ARG;
    if (s->sample_count == 0 || s->sample_count > RISLTC_MAX_SAMPLE_COUNT) return RISLTC_E_ARG;
    if (!orc_technique_supported(s)) return RISLTC_E_UNSUPPORTED;
    if (s->proposal == RISLTC_PROPOSAL_LEARNED) {
        if (s->light_sampling != RISLTC_LIGHT_RESERVOIR) return RISLTC_E_UNSUPPORTED;
        if (!sc->proposal_table || !sc->cluster_of_light || !sc->cluster_offsets || !sc->cluster_lights ||
            sc->cluster_count == 0)
            return RISLTC_E_STATE;
    } else if (s->proposal == RISLTC_PROPOSAL_CACHE) {
        if (s->light_sampling != RISLTC_LIGHT_RESERVOIR) return RISLTC_E_UNSUPPORTED;
        if (!sc->cluster_of_light || !sc->cache_vis) return RISLTC_E_STATE;
    } else if (s->proposal != RISLTC_PROPOSAL_UNIFORM) {
        return RISLTC_E_ARG;
    }
    g_fast_atan = s->fast_atan ? 1 : 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
    for (uint32_t k = 0; k < count; ++k) {
        uint32_t p = pixels ? pixels[k] : k;
        shade_one(sc, c, s, p, visibility[p], out + 4 * (size_t)k, NULL, NULL);
    }
    return 0;
}
#endif

/* ---------------- unit-level entry points ------------------------------- */
uint32_t SFX(orc_clip_polygon)(uint32_t n, R* v) {
    V3 a[MAXV];
    for (int i = 0; i != MAXV; ++i) a[i] = v3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    uint32_t vc = clip_polygon(n, a);
    for (int i = 0; i != MAXV; ++i) {
        v[3 * i] = a[i].x;
        v[3 * i + 1] = a[i].y;
        v[3 * i + 2] = a[i].z;
    }
    return vc;
}

R SFX(orc_calculate_ltc)(uint32_t vc, const R* v) {
    V3 a[MAXV];
    for (uint32_t i = 0; i != vc; ++i) a[i] = v3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    return calculate_ltc(vc, a);
}

R SFX(orc_prepare_psa)(uint32_t vc, const R* v, R* out) {
    V3 a[MAXV];
    for (int i = 0; i != MAXV; ++i) a[i] = (i < (int)vc) ? v3(v[3 * i], v[3 * i + 1], v[3 * i + 2]) : v3(v[0], v[1], v[2]);
    psa_t p;
    prepare_psa(&p, vc, a);
    if (out) {
        for (int i = 0; i != MAXV; ++i) {
            out[2 * i] = p.vertices[i].x;
            out[2 * i + 1] = p.vertices[i].y;
            out[18 + 2 * i] = p.ellipses[i].x;
            out[18 + 2 * i + 1] = p.ellipses[i].y;
            out[38 + i] = p.sector[i];
        }
        out[36] = p.inner_ellipse_0.x;
        out[37] = p.inner_ellipse_0.y;
        out[47] = p.psa;
    }
    return p.psa;
}

void SFX(orc_sample_psa)(uint32_t vc, const R* v, const R* uv, uint32_t n, R* out) {
    V3 a[MAXV];
    for (int i = 0; i != MAXV; ++i) a[i] = (i < (int)vc) ? v3(v[3 * i], v[3 * i + 1], v[3 * i + 2]) : v3(v[0], v[1], v[2]);
    psa_t p;
    prepare_psa(&p, vc, a);
    for (uint32_t k = 0; k != n; ++k) {
        V3 d = sample_psa(&p, v2(uv[2 * k], uv[2 * k + 1]));
        out[3 * k] = d.x;
        out[3 * k + 1] = d.y;
        out[3 * k + 2] = d.z;
    }
}

void SFX(orc_evaluate_brdf)(const R* sh, const R* dir, R* out) {
    shading_t sd;
    sd.normal = v3(sh[0], sh[1], sh[2]);
    sd.outgoing = v3(sh[3], sh[4], sh[5]);
    sd.diffuse_albedo = v3(sh[6], sh[7], sh[8]);
    sd.fresnel_0 = v3(sh[9], sh[10], sh[11]);
    sd.roughness = sh[12];
    sd.lambert_outgoing = dot3(sd.normal, sd.outgoing);
    sd.position = v3(K(0.0), K(0.0), K(0.0));
    V3 b = evaluate_brdf(&sd, v3(dir[0], dir[1], dir[2]));
    out[0] = b.x;
    out[1] = b.y;
    out[2] = b.z;
}

/* ---- mixed precision: an FP64 target from the FP32 pixel setup ---------------
 * orc_pixel_setup writes the setup that get_shading_data and
 * get_ltc_coefficients give the pixel (ORC_PIXEL_SETUP_FLOATS floats: position
 * 3, world_to_shading 12, w2c 12, w2s_flip 12, w2c_flip 12, diffuse albedo 3,
 * LTC albedo 1); orc_ltc_cp_target evaluates the LTC_CP target
 * |eval_ltc_cp| of one light from such a setup in the build's arithmetic. The
 * FP64 build on the FP32 setup measures the integrator (transform, clip, edge
 * integral) apart from the float rounding of the shading point itself. */
static void pack34(const A34* a, float* out) {
    for (int i = 0; i != 3; ++i)
        for (int k = 0; k != 4; ++k) out[4 * i + k] = (float)a->m[i][k];
}
static void unpack34(const float* in, A34* a) {
    for (int i = 0; i != 3; ++i)
        for (int k = 0; k != 4; ++k) a->m[i][k] = (R)in[4 * i + k];
}

int SFX(orc_pixel_setup)(const orc_scene_t* sc, const risltc_frame_constants_t* c, uint32_t pixel, uint32_t prim,
                         float* out) {
    if (!sc || !c || !out || prim == 0xFFFFFFFFu || (prim >> 31) != 0u || (uint64_t)prim >= sc->triangle_count)
        return RISLTC_E_ARG;
    const uint32_t W = c->viewport_size[0];
    pix_t px;
    const float(*P)[4] = c->pixel_to_ray_direction_world_space;
    R fx = (R)(pixel % W), fy = (R)(pixel / W);
    V3 ray = v3(FMA((R)P[0][2], K(1.0), FMA((R)P[0][1], fy, (R)P[0][0] * fx)),
                FMA((R)P[1][2], K(1.0), FMA((R)P[1][1], fy, (R)P[1][0] * fx)),
                FMA((R)P[2][2], K(1.0), FMA((R)P[2][1], fy, (R)P[2][0] * fx)));
    get_shading_data(sc, c, prim, ray, &px.sd);
    R f0l = dot3(px.sd.fresnel_0, v3(K(0.2126), K(0.7152), K(0.0722)));
    get_ltc_coefficients(sc, &c->ltc_constants, f0l, px.sd.roughness, px.sd.position, px.sd.normal, px.sd.outgoing,
                         &px.ltc);
    px.w2c = mul33_34(&px.ltc.shading_to_cosine, &px.ltc.world_to_shading);
    px.w2s_flip = px.ltc.world_to_shading;
    for (int k = 0; k != 4; ++k) px.w2s_flip.m[1][k] = -px.w2s_flip.m[1][k];
    px.w2c_flip = mul33_34(&px.ltc.shading_to_cosine, &px.w2s_flip);
    out[0] = (float)px.sd.position.x;
    out[1] = (float)px.sd.position.y;
    out[2] = (float)px.sd.position.z;
    pack34(&px.ltc.world_to_shading, out + 3);
    pack34(&px.w2c, out + 15);
    pack34(&px.w2s_flip, out + 27);
    pack34(&px.w2c_flip, out + 39);
    out[51] = (float)px.sd.diffuse_albedo.x;
    out[52] = (float)px.sd.diffuse_albedo.y;
    out[53] = (float)px.sd.diffuse_albedo.z;
    out[54] = (float)px.ltc.albedo;
    return RISLTC_OK;
}

R SFX(orc_ltc_cp_target)(const orc_scene_t* sc, const float* setup, uint32_t light, uint32_t* clip_counts) {
    pix_t px;
    memset(&px, 0, sizeof(px));
    px.sd.position = v3((R)setup[0], (R)setup[1], (R)setup[2]);
    unpack34(setup + 3, &px.ltc.world_to_shading);
    unpack34(setup + 15, &px.w2c);
    unpack34(setup + 27, &px.w2s_flip);
    unpack34(setup + 39, &px.w2c_flip);
    px.sd.diffuse_albedo = v3((R)setup[51], (R)setup[52], (R)setup[53]);
    px.ltc.albedo = (R)setup[54];
    light_t l;
    load_light(sc, light, &l);
    V3 col = eval_ltc_cp(&px, &l, clip_counts);
    return SQRT(dot3(col, col));
}

/* mesh decode (mesh_quantization.glsl:19-45): out = xyz */
void SFX(orc_decode_position)(uint32_t q0, uint32_t q1, const float f[3], const float s[3], R* out) {
    V3 p = decode_position(q0, q1, f, s);
    out[0] = p.x;
    out[1] = p.y;
    out[2] = p.z;
}

void SFX(orc_decode_normal)(uint32_t qx, uint32_t qy, R* out) {
    V3 n = decode_normal((uint16_t)qx, (uint16_t)qy);
    out[0] = n.x;
    out[1] = n.y;
    out[2] = n.z;
}

/* one filtered channel of the LTC arrays (ltc_utility.glsl:56-88, H5) */
R SFX(orc_ltc_fetch)(const orc_scene_t* sc, int table, int channel, R u, R v, R w) {
    return ltc_fetch(sc, table, channel, u, v, w);
}

#ifndef ORC_F64
void orc_ltc_coefficients_f32(const orc_scene_t* sc, const risltc_ltc_constants_t* k, float f0l, float rough,
                              const float* pos, const float* n, const float* o, float* out) {
    ltc_t ltc;
    get_ltc_coefficients(sc, k, f0l, rough, v3(pos[0], pos[1], pos[2]), v3(n[0], n[1], n[2]), v3(o[0], o[1], o[2]),
                         &ltc);
    for (int i = 0; i != 3; ++i)
        for (int j = 0; j != 4; ++j) out[4 * i + j] = ltc.world_to_shading.m[i][j];
    for (int i = 0; i != 3; ++i)
        for (int j = 0; j != 3; ++j) {
            out[12 + 3 * i + j] = ltc.shading_to_cosine.m[i][j];
            out[21 + 3 * i + j] = ltc.cosine_to_shading.m[i][j];
        }
    out[30] = ltc.albedo;
    out[31] = ltc.determinant;
}

#endif
