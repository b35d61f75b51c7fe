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
 * orc_common.c -- oracle parts that do not depend on the arithmetic type:
 * noise, the host data model, project FP32 transcendentals, the clipping
 * rule and the CPU LBVH.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h). Built with -ffp-contract=off so that
 * every a*b+c is two roundings unless fmaf is written.
 */
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

#include "orc.h"

/* ======================================================================== */
/* Noise: src/shaders/noise_utility.glsl:26-95, src/noise_table.c:24-28      */
/* ======================================================================== */

/* murmur_hash3_mix (noise_utility.glsl:26-42): one MurmurHash3 x86_32 block */
uint32_t orc_murmur_mix(uint32_t hash, uint32_t k) {
    k *= 0xcc9e2d51u;
    k = (k << 15) | (k >> 17);
    k *= 0x1b873593u;
    hash ^= k;
    hash = ((hash << 13) | (hash >> 19)) * 5u + 0xe6546b64u;
    return hash;
}

/* murmur_hash3_finalize (noise_utility.glsl:44-52): MurmurHash3 fmix32 */
uint32_t orc_murmur_finalize(uint32_t hash) {
    hash ^= hash >> 16;
    hash *= 0x85ebca6bu;
    hash ^= hash >> 13;
    hash *= 0xc2b2ae35u;
    hash ^= hash >> 16;
    return hash;
}

/* get_noise_accessor (noise_utility.glsl:77-84) */
uint32_t orc_noise_seed(uint32_t px, uint32_t py, uint32_t width, uint32_t random_number_0) {
    uint32_t index = orc_murmur_mix(0u, px + py * width);
    return orc_murmur_finalize(orc_murmur_mix(index, random_number_0));
}

/* rand_lcg + get_noise_gen (noise_utility.glsl:63-72). The product is the
 * float conversion of the 32-bit state times 2^-32 (exact power of two). */
float orc_noise_next(uint32_t* seed) {
    *seed = 1664525u * *seed + 1013904223u;
    return (float)(*seed) * (1.0f / 4294967296.0f);
}

/* threads an OpenMP region of the oracle gets (CPU-baseline reporting) */
int orc_omp_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* wang_random_number (src/math_utilities.h:50-57) */
uint32_t orc_wang_random_number(uint32_t seed) {
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed;
}

/* set_noise_constants (src/noise_table.c:24-28) */
void orc_set_noise_constants(uint32_t random_numbers[4], uint32_t* random_seed, int animate) {
    for (uint32_t i = 0; i != 4; ++i)
        random_numbers[i] = animate ? orc_wang_random_number(*random_seed * 4u + i) : (i * 0x123456u);
    ++*random_seed;
}

/* ======================================================================== */
/* Project FP32 transcendentals (Cephes single precision, S. L. Moshier).    */
/* Every polynomial is evaluated with fmaf in a fixed order; the GPU kernels */
/* use the same sequence.                                                     */
/* ======================================================================== */
float prj_atanf(float xx) {
    float x = fabsf(xx);
    float y;
    if (x > 2.414213562373095f) {
        y = 1.5707963267948966f;
        x = -1.0f / x;
    } else if (x > 0.4142135623730950f) {
        y = 0.7853981633974483f;
        x = (x - 1.0f) / (x + 1.0f);
    } else {
        y = 0.0f;
    }
    float z = x * x;
    float p = fmaf(fmaf(fmaf(8.05374449538e-2f, z, -1.38776856032e-1f), z, 1.99777106478e-1f), z,
                   -3.33329491539e-1f);
    y = y + fmaf(p * z, x, x);
    return (xx < 0.0f) ? -y : y;
}

float prj_asinf(float xx) {
    float a = fabsf(xx);
    if (!(a <= 1.0f)) return (xx - xx) / (xx - xx);
    float x, z;
    int flag;
    if (a > 0.5f) {
        z = 0.5f * (1.0f - a);
        x = sqrtf(z);
        flag = 1;
    } else {
        x = a;
        z = x * x;
        flag = 0;
    }
    float r;
    if (a < 1.0e-4f) {
        r = x;
    } else {
        float p = fmaf(fmaf(fmaf(fmaf(4.2163199048e-2f, z, 2.4181311049e-2f), z, 4.5470025998e-2f), z,
                            7.4953002686e-2f),
                       z, 1.6666752422e-1f);
        r = fmaf(p * z, x, x);
    }
    if (flag) {
        r = r + r;
        r = 1.5707963267948966f - r;
    }
    return (xx < 0.0f) ? -r : r;
}

float prj_acosf(float x) {
    if (x < -0.5f) return 3.14159265358979f - 2.0f * prj_asinf(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * prj_asinf(sqrtf(0.5f * (1.0f - x)));
    return 1.5707963267948966f - prj_asinf(x);
}

static float prj_sin_poly(float z, float x) {
    float p = fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f);
    return fmaf(p * z, x, x);
}

static float prj_cos_poly(float z) {
    float p = fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f);
    return fmaf(p * z, z, fmaf(-0.5f, z, 1.0f));
}

/* Cody-Waite reduction by pi/4 (Cephes DP1..DP3), fma form */
static float prj_reduce(float x, uint32_t* octant) {
    uint32_t j = (uint32_t)(1.27323954473516f * x);
    float y = (float)j;
    if (j & 1u) {
        j += 1u;
        y += 1.0f;
    }
    *octant = j & 7u;
    x = fmaf(-y, 0.78515625f, x);
    x = fmaf(-y, 2.4187564849853515625e-4f, x);
    x = fmaf(-y, 3.77489497744594108e-8f, x);
    return x;
}

float prj_sinf(float xx) {
    float x = fabsf(xx);
    if (!(x <= 1.0e5f)) return xx - xx;
    int negative = xx < 0.0f;
    uint32_t j;
    x = prj_reduce(x, &j);
    if (j > 3u) {
        negative = !negative;
        j -= 4u;
    }
    float z = x * x;
    float r = (j == 1u || j == 2u) ? prj_cos_poly(z) : prj_sin_poly(z, x);
    return negative ? -r : r;
}

float prj_cosf(float xx) {
    float x = fabsf(xx);
    if (!(x <= 1.0e5f)) return xx - xx;
    int negative = 0;
    uint32_t j;
    x = prj_reduce(x, &j);
    if (j > 3u) {
        j -= 4u;
        negative = !negative;
    }
    if (j > 1u) negative = !negative;
    float z = x * x;
    float r = (j == 1u || j == 2u) ? prj_sin_poly(z, x) : prj_cos_poly(z);
    return negative ? -r : r;
}

/* ======================================================================== */
/* Host data model                                                            */
/* ======================================================================== */

/* update_polygonal_light (src/polygonal_light.c:44-98). The order of every
 * float operation matches the reference so that the result is bitwise equal
 * (checked against the reference source compiled into oracle/_ref). */
void orc_update_polygonal_light(risltc_polygonal_light_t* light) {
    light->inv_scaling_x = 1.0f / light->scaling_x;
    light->inv_scaling_y = 1.0f / light->scaling_y;
    float ca = cosf(light->rotation_angles[0]), sa = sinf(light->rotation_angles[0]);
    float cb = cosf(light->rotation_angles[1]), sb = sinf(light->rotation_angles[1]);
    float cc = cosf(light->rotation_angles[2]), sc = sinf(light->rotation_angles[2]);
    float ca_sb = ca * sb, sa_sb = sa * sb;
    /* rotation about z, then y, then x */
    float r[3][4] = {
        {cb * cc, -cb * sc, -sb, 0.0f},
        {-sa_sb * cc + ca * sc, sa_sb * sc + ca * cc, -sa * cb, 0.0f},
        {ca_sb * cc + sa * sc, -ca_sb * sc + sa * cc, ca * cb, 0.0f},
    };
    memcpy(light->rotation, r, sizeof(r));
    float scale[2] = {light->scaling_x, light->scaling_y};
    const float* ps = light->vertices_plane_space;
    float* ws = light->vertices_world_space;
    for (uint32_t v = 0; v != light->vertex_count; ++v) {
        for (uint32_t row = 0; row != 3; ++row) {
            float acc = light->translation[row];
            for (uint32_t col = 0; col != 2; ++col) acc += scale[col] * r[row][col] * ps[v * 4 + col];
            ws[v * 4 + row] = acc;
        }
    }
    for (uint32_t row = 0; row != 3; ++row) light->plane[row] = r[row][2];
    light->plane[3] = -(r[0][2] * light->translation[0] + r[1][2] * light->translation[1] +
                        r[2][2] * light->translation[2]);
    /* signed area of the triangle fan around vertex 0, in plane space */
    float area = 0.0f;
    for (uint32_t t = 0; t != light->vertex_count - 2u; ++t) {
        float ax = ps[(t + 2) * 4 + 0] - ps[0], bx = ps[(t + 1) * 4 + 0] - ps[0];
        float ay = ps[(t + 2) * 4 + 1] - ps[1], by = ps[(t + 1) * 4 + 1] - ps[1];
        area += 0.5f * (ax * by - bx * ay);
    }
    area *= scale[0] * scale[1];
    float abs_area = (area < 0.0f) ? -area : area;
    light->area = abs_area;
    light->rcp_area = 1.0f / abs_area;
    for (uint32_t i = 0; i != 3; ++i) light->surface_radiance[i] = light->radiant_flux[i];
    for (uint32_t i = 0; i != 4; ++i) light->plane[i] = (area > 0.0f) ? light->plane[i] : -light->plane[i];
}

/* matrix_inverse (src/math_utilities.h:24-47): cofactor expansion, the same
 * products in the same order, then scaling by 1/det. */
void orc_matrix_inverse(float inverse[4][4], const float matrix[4][4]) {
    const float* m = &matrix[0][0];
    float* o = &inverse[0][0];
#define C3(a, b, c, d, e, f, g, h, i, j, k, l, p, q, r, s, t, u)                                         \
    (m[a] * m[b] * m[c] - m[d] * m[e] * m[f] - m[g] * m[h] * m[i] + m[j] * m[k] * m[l] +               \
     m[p] * m[q] * m[r] - m[s] * m[t] * m[u])
    o[0] = C3(5, 10, 15, 5, 11, 14, 9, 6, 15, 9, 7, 14, 13, 6, 11, 13, 7, 10);
    o[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    o[8] = C3(4, 9, 15, 4, 11, 13, 8, 5, 15, 8, 7, 13, 12, 5, 11, 12, 7, 9);
    o[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    o[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
           m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    o[5] = C3(0, 10, 15, 0, 11, 14, 8, 2, 15, 8, 3, 14, 12, 2, 11, 12, 3, 10);
    o[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
           m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    o[13] = C3(0, 9, 14, 0, 10, 13, 8, 1, 14, 8, 2, 13, 12, 1, 10, 12, 2, 9);
    o[2] = C3(1, 6, 15, 1, 7, 14, 5, 2, 15, 5, 3, 14, 13, 2, 7, 13, 3, 6);
    o[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
           m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    o[10] = C3(0, 5, 15, 0, 7, 13, 4, 1, 15, 4, 3, 13, 12, 1, 7, 12, 3, 5);
    o[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
            m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    o[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
           m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    o[7] = C3(0, 6, 11, 0, 7, 10, 4, 2, 11, 4, 3, 10, 8, 2, 7, 8, 3, 6);
    o[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
            m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    o[15] = C3(0, 5, 10, 0, 6, 9, 4, 1, 10, 4, 2, 9, 8, 1, 6, 8, 2, 5);
#undef C3
    float det = m[0] * o[0] + m[1] * o[4] + m[2] * o[8] + m[3] * o[12];
    float rcp = 1.0f / det;
    for (int i = 0; i != 16; ++i) o[i] = o[i] * rcp;
}

/* get_world_to_projection_space (src/camera.c:24-83) */
void orc_get_world_to_projection_space(float out[4][4], const risltc_camera_t* cam, float aspect) {
    float cx = cosf(cam->rotation_x), sx = sinf(cam->rotation_x);
    float cz = cosf(cam->rotation_z), sz = sinf(cam->rotation_z);
    float rx[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, cx, sx}, {0.0f, -sx, cx}};
    float rz[3][3] = {{cz, sz, 0.0f}, {-sz, cz, 0.0f}, {0.0f, 0.0f, 1.0f}};
    float rot[3][3] = {{0.0f}};
    for (int i = 0; i != 3; ++i)
        for (int j = 0; j != 3; ++j)
            for (int l = 0; l != 3; ++l) rot[i][j] += rz[i][l] * rx[l][j];
    float origin[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i != 3; ++i)
        for (int j = 0; j != 3; ++j) origin[i] -= rot[j][i] * cam->position_world_space[j];
    float view[4][4] = {{rot[0][0], rot[1][0], rot[2][0], origin[0]},
                        {rot[0][1], rot[1][1], rot[2][1], origin[1]},
                        {rot[0][2], rot[1][2], rot[2][2], origin[2]},
                        {0.0f, 0.0f, 0.0f, 1.0f}};
    float n = cam->near_plane, f = cam->far_plane;
    float top = tanf(0.5f * cam->vertical_fov);
    float right = aspect * top;
    float proj[4][4] = {{-1.0f / right, 0.0f, 0.0f, 0.0f},
                        {0.0f, 1.0f / top, 0.0f, 0.0f},
                        {0.0f, 0.0f, -(f + n) / (f - n), -2.0f * f * n / (f - n)},
                        {0.0f, 0.0f, -1.0f, 0.0f}};
    memset(out, 0, sizeof(float) * 16);
    for (int i = 0; i != 4; ++i)
        for (int j = 0; j != 4; ++j)
            for (int l = 0; l != 4; ++l) out[i][j] += proj[i][l] * view[l][j];
}

/* write_constants (src/main.c:2902-2946) */
void orc_write_frame_constants(risltc_frame_constants_t* c, const risltc_camera_t* camera, uint32_t width,
                               uint32_t height, const float deq_factor[3], const float deq_summand[3],
                               const risltc_ltc_constants_t* ltc, float exposure, float roughness_factor,
                               float mis_vis_estimate, uint32_t noise_seed, int animate) {
    memset(c, 0, sizeof(*c));
    for (int i = 0; i != 3; ++i) {
        c->mesh_dequantization_factor[i] = deq_factor[i];
        c->mesh_dequantization_summand[i] = deq_summand[i];
        c->camera_position_world_space[i] = camera->position_world_space[i];
    }
    c->mis_visibility_estimate = mis_vis_estimate;
    c->viewport_size[0] = width;
    c->viewport_size[1] = height;
    if (ltc) c->ltc_constants = *ltc;
    c->error_factor = powf(10.0f, 7.0f); /* error_min_exponent = -7 (src/main.c:230) */
    c->exposure_factor = exposure;
    c->roughness_factor = roughness_factor;
    uint32_t seed = noise_seed;
    orc_set_noise_constants(c->noise_random_numbers, &seed, animate);
    orc_get_world_to_projection_space(c->world_to_projection_space, camera,
                                      ((float)width) / ((float)height));
    float vp[4];
    vp[0] = 2.0f / width;
    vp[1] = 2.0f / height;
    vp[2] = 0.5f * vp[0] - 1.0f;
    vp[3] = 0.5f * vp[1] - 1.0f;
    float w2p[4][4], p2w[4][4];
    memcpy(w2p, c->world_to_projection_space, sizeof(w2p));
    w2p[0][3] = 0.0f;
    w2p[1][3] = 0.0f;
    w2p[2][3] = 0.0f;
    orc_matrix_inverse(p2w, w2p);
    float pix[4][3] = {{vp[0], 0.0f, vp[2]}, {0.0f, vp[1], vp[3]}, {0.0f, 0.0f, 1.0f}, {0.0f, 0.0f, 1.0f}};
    for (int i = 0; i != 3; ++i)
        for (int j = 0; j != 3; ++j)
            for (int k = 0; k != 4; ++k) c->pixel_to_ray_direction_world_space[i][j] += p2w[i][k] * pix[k][j];
}

/* load_ltc_table quantisation (src/ltc_table.c:85-115): the fit stores the
 * cosine-to-BRDF matrix [[d0,0,d1],[0,d2,0],[d3,0,1]] and the albedo d4; the
 * table stores its adjugate, scaled so the largest entry has magnitude 1,
 * with entry (0,2) negated, as UNORM16. */
void orc_quantize_ltc_entry(const float d[5], uint16_t rgba[4], uint16_t rg[2]) {
    float adj[3][3] = {{d[2], 0.0f, -d[1] * d[2]},
                       {0.0f, d[0] - d[1] * d[3], 0.0f},
                       {-d[2] * d[3], 0.0f, d[0] * d[2]}};
    float largest = fabsf(adj[0][0]);
    for (int r = 0; r != 3; ++r)
        for (int k = 0; k != 3; ++k)
            if (largest < fabsf(adj[r][k])) largest = fabsf(adj[r][k]);
    for (int r = 0; r != 3; ++r)
        for (int k = 0; k != 3; ++k) adj[r][k] /= largest;
    float vals[6] = {adj[0][0], -adj[0][2], adj[1][1], adj[2][0], adj[2][2], d[4]};
    uint16_t q[6];
    for (int i = 0; i != 6; ++i) {
        float x = vals[i];
        if (x < 0.0f) x = 0.0f;
        if (x > 1.0f) x = 1.0f;
        q[i] = (uint16_t)(x * 65535.0f + 0.5f);
    }
    rgba[0] = q[0];
    rgba[1] = q[1];
    rgba[2] = q[2];
    rgba[3] = q[3];
    rg[0] = q[4];
    rg[1] = q[5];
}

/* ltc_constants_t of load_ltc_table (src/ltc_table.c:184-191) */
void orc_ltc_constants(risltc_ltc_constants_t* o, uint32_t rr, uint32_t ri, uint32_t fc) {
    memset(o, 0, sizeof(*o));
    o->fresnel_index_factor = (float)(fc - 1);
    o->fresnel_index_summand = 0.0f;
    o->roughness_factor = (float)(rr - 1) / (float)rr;
    o->roughness_summand = 0.5f / (float)rr;
    o->inclination_factor = (float)(ri - 1) / (0.5f * 3.1415926535897932384626433832795f * ri);
    o->inclination_summand = 0.5f / (float)ri;
}

/* ======================================================================== */
/* Clipping rule (src/shaders/polygon_clipping.glsl:35-225)                  */
/* ======================================================================== */
/* The reference's autogenerated table clips a convex polygon against z > 0
 * and writes the result in place. Its output is the cyclic sequence
 * [entry point, inside vertices..., exit point], each intersection computed
 * as iz0(v[e], v[e+1]) on edge e in polygon order, rotated so that the
 * number of slots that have to be written (including the repeated first
 * vertex after the last) is smallest; among equal rotations the one that
 * keeps vertex 0 in slot 0 wins, then the smallest rotation. This rule
 * reproduces all 120 cases of the table; it also defines 8-gons, which the
 * reference cannot clip (its 3-bit vertex count overflows). */
static int clip_candidate(int n, int L, const int* cyc, int p, int* out) {
    for (int i = 0; i != L; ++i) out[i] = cyc[(i + p) % L];
    int slots = L;
    if (L < RISLTC_MAX_POLYGON_VERTEX_COUNT) {
        out[L] = out[0];
        slots = L + 1;
    }
    int cost = 0;
    for (int j = 0; j != slots; ++j)
        if (!(j < n && out[j] == j)) ++cost;
    return cost;
}

int orc_clip_rule(int n, uint32_t mask, int out[RISLTC_MAX_POLYGON_VERTEX_COUNT + 1]) {
    int inside[RISLTC_MAX_LIGHT_VERTEX_COUNT];
    int k = 0;
    for (int i = 0; i != n; ++i) {
        inside[i] = (mask >> i) & 1u;
        k += inside[i];
    }
    if (k == 0) return 0;
    if (k == n) {
        for (int i = 0; i != n; ++i) out[i] = i;
        if (n < RISLTC_MAX_POLYGON_VERTEX_COUNT) out[n] = 0;
        return n;
    }
    int s = 0;
    for (int i = 0; i != n; ++i)
        if (inside[i] && !inside[(i + n - 1) % n]) {
            s = i;
            break;
        }
    int L = k + 2;
    int cyc[RISLTC_MAX_POLYGON_VERTEX_COUNT];
    cyc[0] = -(((s + n - 1) % n) + 1);
    for (int j = 0; j != k; ++j) cyc[1 + j] = (s + j) % n;
    cyc[k + 1] = -(((s + k - 1) % n) + 1);
    int best_key = 1 << 30, best_p = 0;
    int tmp[RISLTC_MAX_POLYGON_VERTEX_COUNT + 1];
    for (int p = 0; p != L; ++p) {
        int cost = clip_candidate(n, L, cyc, p, tmp);
        int key = cost * 1024 + (tmp[0] == 0 ? 0 : 512) + p;
        if (key < best_key) {
            best_key = key;
            best_p = p;
        }
    }
    clip_candidate(n, L, cyc, best_p, out);
    return L;
}

int orc_clip_rotation(int n, uint32_t mask) {
    int inside = 0;
    for (int i = 0; i != n; ++i) inside += (mask >> i) & 1u;
    if (inside == 0 || inside == n) return 0;
    int out[RISLTC_MAX_POLYGON_VERTEX_COUNT + 1];
    int L = orc_clip_rule(n, mask, out);
    /* find p with out[0] == cyc[p]: recompute cyc */
    int s = 0;
    for (int i = 0; i != n; ++i)
        if (((mask >> i) & 1u) && !((mask >> ((i + n - 1) % n)) & 1u)) {
            s = i;
            break;
        }
    int k = L - 2;
    int cyc[RISLTC_MAX_POLYGON_VERTEX_COUNT];
    cyc[0] = -(((s + n - 1) % n) + 1);
    for (int j = 0; j != k; ++j) cyc[1 + j] = (s + j) % n;
    cyc[k + 1] = -(((s + k - 1) % n) + 1);
    for (int p = 0; p != L; ++p)
        if (cyc[p] == out[0] && cyc[(p + 1) % L] == out[1]) return p;
    return -1;
}

/* ======================================================================== */
/* LBVH: 30-bit Morton codes, Karras (2012) hierarchy, escape links.          */
/* ======================================================================== */
uint32_t orc_morton_part_1_by_2(uint32_t x) {
    x &= 0x000003ffu;
    x = (x ^ (x << 16)) & 0xff0000ffu;
    x = (x ^ (x << 8)) & 0x0300f00fu;
    x = (x ^ (x << 4)) & 0x030c30c3u;
    x = (x ^ (x << 2)) & 0x09249249u;
    return x;
}

void orc_dequantize_mesh(const uint32_t* pos, uint64_t tri_count, const float f[3], const float s[3],
                         float* out) {
    for (uint64_t v = 0; v != tri_count * 3; ++v) {
        uint32_t q0 = pos[2 * v], q1 = pos[2 * v + 1];
        uint32_t qx = q0 & 0x1FFFFFu;
        uint32_t qy = ((q0 & 0xFFE00000u) >> 21) | ((q1 & 0x3FFu) << 11);
        uint32_t qz = (q1 & 0x7FFFFC00u) >> 10;
        out[3 * v + 0] = fmaf((float)qx, f[0], s[0]);
        out[3 * v + 1] = fmaf((float)qy, f[1], s[1]);
        out[3 * v + 2] = fmaf((float)qz, f[2], s[2]);
    }
}

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

static int bvh_delta(const uint64_t* keys, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    return __builtin_clzll(keys[i] ^ keys[j]);
}

int orc_bvh_build(orc_bvh_t* b, const float* tris, uint32_t n) {
    memset(b, 0, sizeof(*b));
    if (n == 0) return 1;
    b->leaf_count = n;
    uint32_t nodes = 2 * n - 1;
    b->node_box = (float*)malloc(sizeof(float) * 6 * nodes);
    b->node_left = (int32_t*)malloc(sizeof(int32_t) * nodes);
    b->node_right = (int32_t*)malloc(sizeof(int32_t) * nodes);
    b->node_escape = (int32_t*)malloc(sizeof(int32_t) * nodes);
    b->node_parent = (int32_t*)malloc(sizeof(int32_t) * nodes);
    b->leaf_tri = (float*)malloc(sizeof(float) * 9 * n);
    b->leaf_prim = (uint32_t*)malloc(sizeof(uint32_t) * n);
    b->sorted_keys = (uint64_t*)malloc(sizeof(uint64_t) * n);
    /* centroids and their bounds */
    float* cen = (float*)malloc(sizeof(float) * 3 * n);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t t = 0; t != n; ++t) {
        for (int a = 0; a != 3; ++a) {
            float c = ((tris[9 * t + a] + tris[9 * t + 3 + a]) + tris[9 * t + 6 + a]) / 3.0f;
            cen[3 * t + a] = c;
            lo[a] = fminf(lo[a], c);
            hi[a] = fmaxf(hi[a], c);
        }
    }
    float factor[3], summand[3];
    for (int a = 0; a != 3; ++a) {
        float ext = hi[a] - lo[a];
        factor[a] = (ext > 0.0f) ? 1024.0f / ext : 0.0f;
        summand[a] = -lo[a] * factor[a];
    }
    for (uint32_t t = 0; t != n; ++t) {
        uint32_t code = 0;
        for (int a = 0; a != 3; ++a) {
            float q = cen[3 * t + a] * factor[a] + summand[a];
            q = fminf(fmaxf(q, 0.0f), 1023.0f);
            code |= orc_morton_part_1_by_2((uint32_t)q) << a;
        }
        b->sorted_keys[t] = ((uint64_t)code << 32) | t;
    }
    free(cen);
    qsort(b->sorted_keys, n, sizeof(uint64_t), cmp_u64);
    for (uint32_t k = 0; k != n; ++k) {
        uint32_t prim = (uint32_t)(b->sorted_keys[k] & 0xFFFFFFFFu);
        b->leaf_prim[k] = prim;
        memcpy(b->leaf_tri + 9 * k, tris + 9 * prim, sizeof(float) * 9);
        float* box = b->node_box + 6 * (n - 1 + k);
        for (int a = 0; a != 3; ++a) {
            float v0 = tris[9 * prim + a], v1 = tris[9 * prim + 3 + a], v2 = tris[9 * prim + 6 + a];
            box[a] = fminf(fminf(v0, v1), v2);
            box[3 + a] = fmaxf(fmaxf(v0, v1), v2);
        }
        b->node_left[n - 1 + k] = -1;
        b->node_right[n - 1 + k] = -1;
    }
    b->node_parent[0] = -1;
    const uint64_t* keys = b->sorted_keys;
    int ni = (int)n;
    for (int i = 0; i < ni - 1; ++i) {
        int d = (bvh_delta(keys, ni, i, i + 1) - bvh_delta(keys, ni, i, i - 1)) >= 0 ? 1 : -1;
        int dmin = bvh_delta(keys, ni, i, i - d);
        int lmax = 2;
        while (bvh_delta(keys, ni, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (bvh_delta(keys, ni, i, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d;
        int dnode = bvh_delta(keys, ni, i, j);
        int s = 0;
        for (int div = 2;; div *= 2) {
            int t = (l + div - 1) / div;
            if (bvh_delta(keys, ni, i, i + (s + t) * d) > dnode) s += t;
            if (t == 1) break;
        }
        int gamma = i + s * d + (d < 0 ? d : 0);
        int lo_i = i < j ? i : j, hi_i = i < j ? j : i;
        int left = (lo_i == gamma) ? (ni - 1 + gamma) : gamma;
        int right = (hi_i == gamma + 1) ? (ni - 1 + gamma + 1) : gamma + 1;
        b->node_left[i] = left;
        b->node_right[i] = right;
        b->node_parent[left] = i;
        b->node_parent[right] = i;
    }
    /* boxes bottom-up: post-order with an explicit stack; escape links
     * top-down in pre-order */
    if (n > 1) {
        int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * 2 * nodes);
        uint8_t* done = (uint8_t*)calloc(nodes, 1);
        int sp = 0;
        stack[sp++] = 0;
        while (sp) {
            int x = stack[sp - 1];
            if (x >= ni - 1) {
                --sp;
                continue;
            }
            int l = b->node_left[x], r = b->node_right[x];
            if (!done[x]) {
                done[x] = 1;
                stack[sp++] = r;
                stack[sp++] = l;
                continue;
            }
            --sp;
            float* box = b->node_box + 6 * x;
            const float* bl = b->node_box + 6 * l;
            const float* br = b->node_box + 6 * r;
            for (int a = 0; a != 3; ++a) {
                box[a] = fminf(bl[a], br[a]);
                box[3 + a] = fmaxf(bl[3 + a], br[3 + a]);
            }
        }
        free(done);
        b->node_escape[0] = -1;
        sp = 0;
        stack[sp++] = 0;
        while (sp) {
            int x = stack[--sp];
            if (x >= ni - 1) continue;
            int l = b->node_left[x], r = b->node_right[x];
            b->node_escape[l] = r;
            b->node_escape[r] = b->node_escape[x];
            stack[sp++] = r;
            stack[sp++] = l;
        }
        free(stack);
    } else {
        b->node_escape[0] = -1;
    }
    return 0;
}

/* ---- binned-SAH builder (the product's default, risltc_amd/csrc/sah.hip) --
 * Same decisions as the device builder: key-ordered min/max (so -0 < +0),
 * centroid = box centre, 16 bins per axis of non-zero centroid extent,
 * bin = min(15, (int)((c - lo) * (16 / ext))), cost n_l area_l + n_r area_r
 * with area = dx dy + dy dz + dz dx, first minimum over (axis, bin), halving
 * in index order when no axis has extent, stable partition. Internal nodes
 * are numbered in pre-order, leaves left to right. parity: the device's
 * pre-order records equal orc_bvh_records of this tree word for word
 * (tests/test_gpu_parity.py). */
#define ORC_SAH_BINS 16
static int orc_key(float f) {
    int32_t b;
    memcpy(&b, &f, 4);
    return b >= 0 ? b : (b ^ 0x7FFFFFFF);
}
static float orc_unkey(int k) {
    int32_t b = k >= 0 ? k : (k ^ 0x7FFFFFFF);
    float f;
    memcpy(&f, &b, 4);
    return f;
}
#define ORC_KEY_POS_INF 0x7F800000
#define ORC_KEY_NEG_INF ((int)(0xFF800000u ^ 0x7FFFFFFFu))
static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }
static int sah_bin(float c, float lo, float scale) {
    int b = (int)((c - lo) * scale);
    return b < ORC_SAH_BINS - 1 ? b : ORC_SAH_BINS - 1;
}
static float sah_area(const int k[6]) {
    float dx = orc_unkey(k[3]) - orc_unkey(k[0]), dy = orc_unkey(k[4]) - orc_unkey(k[1]);
    float dz = orc_unkey(k[5]) - orc_unkey(k[2]);
    return dx * dy + dy * dz + dz * dx;
}

typedef struct {
    uint32_t first, count;
    int32_t parent;
    int right; /* 1 = right child of parent */
} orc_sah_task_t;

int orc_bvh_build_sah(orc_bvh_t* b, const float* tris, uint32_t n) {
    memset(b, 0, sizeof(*b));
    if (n == 0) return 1;
    b->leaf_count = n;
    uint32_t nodes = 2 * n - 1;
    b->node_box = (float*)malloc(sizeof(float) * 6 * nodes);
    b->node_left = (int32_t*)malloc(sizeof(int32_t) * nodes);
    b->node_right = (int32_t*)malloc(sizeof(int32_t) * nodes);
    b->node_escape = (int32_t*)malloc(sizeof(int32_t) * nodes);
    b->node_parent = (int32_t*)malloc(sizeof(int32_t) * nodes);
    b->leaf_tri = (float*)malloc(sizeof(float) * 9 * n);
    b->leaf_prim = (uint32_t*)malloc(sizeof(uint32_t) * n);
    float* tbox = (float*)malloc(sizeof(float) * 6 * n);
    float* cen = (float*)malloc(sizeof(float) * 3 * n);
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * n);
    orc_sah_task_t* stack = (orc_sah_task_t*)malloc(sizeof(orc_sah_task_t) * (n + 1));
    for (uint32_t t = 0; t != n; ++t) {
        const float* v = tris + 9 * (size_t)t;
        for (int a = 0; a != 3; ++a) {
            float lo = orc_unkey(imin(imin(orc_key(v[a]), orc_key(v[3 + a])), orc_key(v[6 + a])));
            float hi = orc_unkey(imax(imax(orc_key(v[a]), orc_key(v[3 + a])), orc_key(v[6 + a])));
            tbox[6 * (size_t)t + a] = lo;
            tbox[6 * (size_t)t + 3 + a] = hi;
            cen[3 * (size_t)t + a] = 0.5f * (lo + hi);
        }
        idx[t] = t;
    }
    int32_t next_internal = 0, next_leaf = 0;
    int sp = 0;
    stack[sp].first = 0;
    stack[sp].count = n;
    stack[sp].parent = -1;
    stack[sp].right = 0;
    ++sp;
    while (sp) {
        orc_sah_task_t tk = stack[--sp];
        int32_t id;
        if (tk.count == 1) {
            uint32_t k = (uint32_t)next_leaf++;
            id = (int32_t)(n - 1 + k);
            uint32_t t = idx[tk.first];
            b->leaf_prim[k] = t;
            memcpy(b->leaf_tri + 9 * (size_t)k, tris + 9 * (size_t)t, sizeof(float) * 9);
            memcpy(b->node_box + 6 * (size_t)id, tbox + 6 * (size_t)t, sizeof(float) * 6);
            b->node_left[id] = -1;
            b->node_right[id] = -1;
        } else {
            id = next_internal++;
            int box[6] = {ORC_KEY_POS_INF, ORC_KEY_POS_INF, ORC_KEY_POS_INF, ORC_KEY_NEG_INF, ORC_KEY_NEG_INF,
                          ORC_KEY_NEG_INF};
            int cb[6] = {ORC_KEY_POS_INF, ORC_KEY_POS_INF, ORC_KEY_POS_INF, ORC_KEY_NEG_INF, ORC_KEY_NEG_INF,
                         ORC_KEY_NEG_INF};
            for (uint32_t i = tk.first; i != tk.first + tk.count; ++i) {
                uint32_t t = idx[i];
                for (int a = 0; a != 3; ++a) {
                    box[a] = imin(box[a], orc_key(tbox[6 * (size_t)t + a]));
                    box[3 + a] = imax(box[3 + a], orc_key(tbox[6 * (size_t)t + 3 + a]));
                    cb[a] = imin(cb[a], orc_key(cen[3 * (size_t)t + a]));
                    cb[3 + a] = imax(cb[3 + a], orc_key(cen[3 * (size_t)t + a]));
                }
            }
            for (int a = 0; a != 6; ++a) b->node_box[6 * (size_t)id + a] = orc_unkey(box[a]);
            float clo[3], scale[3];
            int use[3];
            for (int a = 0; a != 3; ++a) {
                clo[a] = orc_unkey(cb[a]);
                float ext = orc_unkey(cb[3 + a]) - clo[a];
                use[a] = ext > 0.0f;
                scale[a] = use[a] ? (float)ORC_SAH_BINS / ext : 0.0f;
            }
            int bin_count[3][ORC_SAH_BINS];
            int bin_box[3][ORC_SAH_BINS][6];
            for (int a = 0; a != 3; ++a)
                for (int k = 0; k != ORC_SAH_BINS; ++k) {
                    bin_count[a][k] = 0;
                    for (int c = 0; c != 6; ++c) bin_box[a][k][c] = c < 3 ? ORC_KEY_POS_INF : ORC_KEY_NEG_INF;
                }
            for (uint32_t i = tk.first; i != tk.first + tk.count; ++i) {
                uint32_t t = idx[i];
                for (int a = 0; a != 3; ++a) {
                    if (!use[a]) continue;
                    int k = sah_bin(cen[3 * (size_t)t + a], clo[a], scale[a]);
                    bin_count[a][k] += 1;
                    for (int c = 0; c != 3; ++c) {
                        bin_box[a][k][c] = imin(bin_box[a][k][c], orc_key(tbox[6 * (size_t)t + c]));
                        bin_box[a][k][3 + c] = imax(bin_box[a][k][3 + c], orc_key(tbox[6 * (size_t)t + 3 + c]));
                    }
                }
            }
            float best = INFINITY;
            int ba = -1, bb = -1;
            uint32_t nl = 0;
            for (int a = 0; a != 3; ++a) {
                if (!use[a]) continue;
                for (int sp2 = 1; sp2 != ORC_SAH_BINS; ++sp2) {
                    int lc = 0, rc = 0;
                    int l[6] = {ORC_KEY_POS_INF, ORC_KEY_POS_INF, ORC_KEY_POS_INF, ORC_KEY_NEG_INF, ORC_KEY_NEG_INF,
                                ORC_KEY_NEG_INF};
                    int r[6] = {ORC_KEY_POS_INF, ORC_KEY_POS_INF, ORC_KEY_POS_INF, ORC_KEY_NEG_INF, ORC_KEY_NEG_INF,
                                ORC_KEY_NEG_INF};
                    for (int k = 0; k != ORC_SAH_BINS; ++k) {
                        int* dst = k < sp2 ? l : r;
                        if (k < sp2) lc += bin_count[a][k];
                        else rc += bin_count[a][k];
                        for (int c = 0; c != 3; ++c) {
                            dst[c] = imin(dst[c], bin_box[a][k][c]);
                            dst[3 + c] = imax(dst[3 + c], bin_box[a][k][3 + c]);
                        }
                    }
                    if (lc == 0 || rc == 0) continue;
                    float cost = (float)lc * sah_area(l) + (float)rc * sah_area(r);
                    if (cost < best) {
                        best = cost;
                        ba = a;
                        bb = sp2;
                        nl = (uint32_t)lc;
                    }
                }
            }
            if (ba < 0) {
                nl = tk.count / 2u;
            } else {
                uint32_t w0 = 0, w1 = nl;
                for (uint32_t i = tk.first; i != tk.first + tk.count; ++i) {
                    uint32_t t = idx[i];
                    if (sah_bin(cen[3 * (size_t)t + ba], clo[ba], scale[ba]) < bb) tmp[w0++] = t;
                    else tmp[w1++] = t;
                }
                memcpy(idx + tk.first, tmp, sizeof(uint32_t) * tk.count);
            }
            /* right pushed first: the left subtree is numbered first */
            stack[sp].first = tk.first + nl;
            stack[sp].count = tk.count - nl;
            stack[sp].parent = id;
            stack[sp].right = 1;
            ++sp;
            stack[sp].first = tk.first;
            stack[sp].count = nl;
            stack[sp].parent = id;
            stack[sp].right = 0;
            ++sp;
        }
        b->node_parent[id] = tk.parent;
        if (tk.parent >= 0) {
            if (tk.right) b->node_right[tk.parent] = id;
            else b->node_left[tk.parent] = id;
        }
    }
    free(tbox);
    free(cen);
    free(idx);
    free(tmp);
    free(stack);
    /* escape links top-down in pre-order */
    int32_t* st = (int32_t*)malloc(sizeof(int32_t) * (nodes + 1));
    int top = 0;
    b->node_escape[n > 1 ? 0 : (int32_t)(n - 1)] = -1;
    if (n > 1) {
        st[top++] = 0;
        while (top) {
            int x = st[--top];
            if (x >= (int)n - 1) continue;
            int l = b->node_left[x], r = b->node_right[x];
            b->node_escape[l] = r;
            b->node_escape[r] = b->node_escape[x];
            st[top++] = r;
            st[top++] = l;
        }
    }
    free(st);
    return 0;
}

/* root node id: internal node 0, or the single leaf */
static int32_t orc_bvh_root(const orc_bvh_t* b) { return b->leaf_count > 1 ? 0 : (int32_t)(b->leaf_count - 1); }

void orc_bvh_records(const orc_bvh_t* b, const uint32_t* positions, float* out) {
    uint32_t n = b->leaf_count, nodes = 2 * n - 1;
    /* leaves below every node, by a post-order pass */
    uint32_t* leaves = (uint32_t*)calloc(nodes, sizeof(uint32_t));
    int32_t* st = (int32_t*)malloc(sizeof(int32_t) * 2 * (nodes + 1));
    uint8_t* seen = (uint8_t*)calloc(nodes, 1);
    int top = 0;
    st[top++] = orc_bvh_root(b);
    while (top) {
        int x = st[top - 1];
        if (x >= (int)n - 1) {
            leaves[x] = 1;
            --top;
            continue;
        }
        if (!seen[x]) {
            seen[x] = 1;
            st[top++] = b->node_right[x];
            st[top++] = b->node_left[x];
            continue;
        }
        --top;
        leaves[x] = leaves[b->node_left[x]] + leaves[b->node_right[x]];
    }
    /* pre-order write */
    uint32_t u = 0;
    top = 0;
    st[top++] = orc_bvh_root(b);
    while (top) {
        int x = st[--top];
        float* r = out + 4 * (size_t)u;
        if (x >= (int)n - 1) {
            uint32_t k = (uint32_t)x - (n - 1), t = b->leaf_prim[k];
            const float* v = b->leaf_tri + 9 * (size_t)k;
            int32_t m1 = -1;
            uint32_t light = positions[6 * (size_t)t + 1] & 0x80000000u;
            for (int i = 0; i != 3; ++i) {
                r[i] = v[i];
                r[4 + i] = v[3 + i];
                r[8 + i] = v[6 + i];
            }
            memcpy(&r[3], &m1, 4);
            memcpy(&r[7], &t, 4);
            memcpy(&r[11], &light, 4);
            u += 3;
        } else {
            int32_t esc = (int32_t)(u + 5 * leaves[x] - 2);
            const float* bx = b->node_box + 6 * (size_t)x;
            for (int i = 0; i != 3; ++i) {
                r[i] = bx[i];
                r[4 + i] = bx[3 + i];
            }
            memcpy(&r[3], &esc, 4);
            r[7] = 0.0f;
            u += 2;
            st[top++] = b->node_right[x];
            st[top++] = b->node_left[x];
        }
    }
    free(leaves);
    free(st);
    free(seen);
}

void orc_bvh_free(orc_bvh_t* b) {
    free(b->node_box);
    free(b->node_left);
    free(b->node_right);
    free(b->node_escape);
    free(b->node_parent);
    free(b->leaf_tri);
    free(b->leaf_prim);
    free(b->sorted_keys);
    memset(b, 0, sizeof(*b));
}

/* Moeller-Trumbore, fixed operation order, both faces, t in [t_min, t_max] */
int orc_ray_triangle(const float o[3], const float d[3], const float* v, float t_min, float t_max,
                     float* t_out) {
    float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
    float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
    float p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    float det = fmaf(e1[2], p[2], fmaf(e1[1], p[1], e1[0] * p[0]));
    if (det == 0.0f) return 0;
    float inv = 1.0f / det;
    float s[3] = {o[0] - v[0], o[1] - v[1], o[2] - v[2]};
    float u = fmaf(s[2], p[2], fmaf(s[1], p[1], s[0] * p[0])) * inv;
    if (u < 0.0f || u > 1.0f) return 0;
    float q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    float w = fmaf(d[2], q[2], fmaf(d[1], q[1], d[0] * q[0])) * inv;
    if (w < 0.0f || u + w > 1.0f) return 0;
    float t = fmaf(e2[2], q[2], fmaf(e2[1], q[1], e2[0] * q[0])) * inv;
    if (!(t >= t_min && t <= t_max)) return 0;
    *t_out = t;
    return 1;
}

/* slab test; the far distance is widened by 1 + 2*gamma(3) (Pharr, Jakob,
 * Humphreys, PBRT 3rd ed., sec. 3.9.2) so the box test is conservative */
int orc_ray_box(const float o[3], const float inv[3], const float* box, float t_min, float t_max) {
    for (int a = 0; a != 3; ++a) {
        float t0 = (box[a] - o[a]) * inv[a];
        float t1 = (box[3 + a] - o[a]) * inv[a];
        float tn = fminf(t0, t1), tf = fmaxf(t0, t1);
        t_min = fmaxf(t_min, tn);
        t_max = fminf(t_max, tf * 1.0000004f);
    }
    return t_min <= t_max;
}

int orc_bvh_any_hit(const orc_bvh_t* b, const float o[3], const float d[3], float t_min, float t_max,
                    uint32_t* visited) {
    float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
    int n = (int)b->leaf_count;
    int x = 0;
    uint32_t count = 0;
    int hit = 0;
    while (x >= 0) {
        ++count;
        if (x >= n - 1) {
            float t;
            if (orc_ray_triangle(o, d, b->leaf_tri + 9 * (x - (n - 1)), t_min, t_max, &t)) {
                hit = 1;
                break;
            }
            x = b->node_escape[x];
        } else {
            x = orc_ray_box(o, inv, b->node_box + 6 * x, t_min, t_max) ? b->node_left[x] : b->node_escape[x];
        }
    }
    if (visited) *visited = count;
    return hit;
}

uint32_t orc_bvh_closest_hit(const orc_bvh_t* b, const float o[3], const float d[3], float t_min,
                             float t_max, float* t_hit) {
    float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
    int n = (int)b->leaf_count;
    int x = 0;
    uint32_t best = 0xFFFFFFFFu;
    while (x >= 0) {
        if (x >= n - 1) {
            float t;
            if (orc_ray_triangle(o, d, b->leaf_tri + 9 * (x - (n - 1)), t_min, t_max, &t)) {
                t_max = t;
                best = b->leaf_prim[x - (n - 1)];
            }
            x = b->node_escape[x];
        } else {
            x = orc_ray_box(o, inv, b->node_box + 6 * x, t_min, t_max) ? b->node_left[x] : b->node_escape[x];
        }
    }
    if (t_hit) *t_hit = t_max;
    return best;
}

int orc_brute_any_hit(const float* tris, uint32_t count, const float o[3], const float d[3], float t_min,
                      float t_max) {
    for (uint32_t t = 0; t != count; ++t) {
        float th;
        if (orc_ray_triangle(o, d, tris + 9 * t, t_min, t_max, &th)) return 1;
    }
    return 0;
}

void orc_render_visibility(const orc_bvh_t* b, const uint32_t* positions, const risltc_frame_constants_t* c,
                           uint32_t* out, int threads) {
    uint32_t W = c->viewport_size[0], H = c->viewport_size[1];
    const float(*P)[4] = c->pixel_to_ray_direction_world_space;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
    for (uint32_t y = 0; y < H; ++y) {
        for (uint32_t x = 0; x != W; ++x) {
            float fx = (float)x, fy = (float)y;
            float d[3];
            for (int i = 0; i != 3; ++i) d[i] = fmaf(P[i][2], 1.0f, fmaf(P[i][1], fy, P[i][0] * fx));
            float t;
            uint32_t prim = orc_bvh_closest_hit(b, c->camera_position_world_space, d, 0.0f, INFINITY, &t);
            if (prim != 0xFFFFFFFFu) prim |= positions[(size_t)prim * 6 + 1] & 0x80000000u;
            out[(size_t)y * W + x] = prim;
        }
    }
}

/* Visibility-cache bytes (project extension; no reference code). parity:
 * bitwise against the library's refresh kernel (tests/test_gpu_learned.py);
 * the formula's end points by construction in tests/test_oracle_estimator.py. */
void orc_cache_bytes(const uint32_t* counts, uint64_t entries, float prior, uint8_t* out) {
    for (uint64_t i = 0; i != entries; ++i) {
        uint32_t ua = counts[2 * i], ub = counts[2 * i + 1];
        if (ua == 0u && ub == 0u) {
            out[i] = 255u;
            continue;
        }
        float a = (float)ua, b = (float)ub;
        float v = ((a + prior * 0.5f) * 254.0f) / ((a + b) + prior) + 0.5f;
        uint32_t q = (uint32_t)v;
        out[i] = (uint8_t)(q > 254u ? 254u : q);
    }
}
