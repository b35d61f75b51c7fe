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
 * orc.h -- CPU oracle for the RIS-LTC shading pass.
 *
 * TEST INFRASTRUCTURE ONLY. This is a scalar C99 restatement of the
 * reference's shading path (src/shaders/shading_pass.frag.glsl and its
 * includes, plus the host code that prepares its inputs). Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg load it, as the
 * checker. The product (risltc_amd/) never links or calls it.
 *
 * The shading path is compiled twice from orc_core.c: once with float
 * arithmetic (FP32-faithful: explicit fmaf where the GLSL writes fma, no
 * contraction, project transcendentals shared with the GPU kernels) and once
 * with double arithmetic (libm transcendentals) as the accuracy reference.
 * Functions of the two builds carry the suffixes _f32 and _f64.
 *
 * Parity pins (what checks each part) are listed per function below and in
 * the test files; "parity unpinned" marks a function no check constrains
 * beyond its restatement.
 */
#ifndef ORC_H
#define ORC_H
#include <stddef.h>
#include <stdint.h>

#include "../include/risltc/risltc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- noise (src/shaders/noise_utility.glsl:26-95) ---------- */
uint32_t orc_murmur_mix(uint32_t hash, uint32_t k);
uint32_t orc_murmur_finalize(uint32_t hash);
uint32_t orc_noise_seed(uint32_t px, uint32_t py, uint32_t width, uint32_t random_number_0);
/* rand_lcg + get_noise_gen: advances *seed, returns float(seed) * 2^-32 */
float orc_noise_next(uint32_t* seed);
/* wang_random_number (src/math_utilities.h:50-57) */
uint32_t orc_wang_random_number(uint32_t seed);
int orc_omp_max_threads(void);
void orc_set_noise_constants(uint32_t random_numbers[4], uint32_t* random_seed, int animate);

/* ---------------- project FP32 transcendentals ------------------------- */
/* Cephes (S. L. Moshier) single-precision algorithms; identical op order in
 * the GPU kernels, so FP32 results agree bit for bit. Pinned against libm
 * double precision (tests/test_oracle_math.py). */
float prj_atanf(float x);
float prj_asinf(float x);
float prj_acosf(float x);
float prj_sinf(float x);
float prj_cosf(float x);

/* ---------------- host data model ----------------------------------- */
/* update_polygonal_light (src/polygonal_light.c:44-98); pinned bitwise
 * against the reference source compiled into oracle/_ref. */
void orc_update_polygonal_light(risltc_polygonal_light_t* light);
/* write_constants (src/main.c:2902-2946) + camera.c:24-83 */
void orc_get_world_to_projection_space(float out[4][4], const risltc_camera_t* camera,
                                       float aspect_ratio);
void orc_matrix_inverse(float inverse[4][4], const float matrix[4][4]);
void orc_write_frame_constants(risltc_frame_constants_t* out, const risltc_camera_t* camera,
                               uint32_t width, uint32_t height, const float deq_factor[3],
                               const float deq_summand[3], const risltc_ltc_constants_t* ltc,
                               float exposure, float roughness_factor, float mis_vis_estimate,
                               uint32_t noise_seed, int animate);
/* load_ltc_table quantisation of one fit matrix (src/ltc_table.c:82-116):
 * data = the 5 floats of fit<i>.dat; writes 4 + 2 UNORM16 values. */
void orc_quantize_ltc_entry(const float data[5], uint16_t rgba[4], uint16_t rg[2]);
void orc_ltc_constants(risltc_ltc_constants_t* out, uint32_t res_roughness,
                       uint32_t res_inclination, uint32_t fresnel_count);

/* ---------------- clipping rule -------------------------------------- */
/* clip_polygon (src/shaders/polygon_clipping.glsl:35-225) as a rule: for n
 * input vertices with inside mask m (bit i = vertex i has z > 0), writes the
 * output slot sequence: code >= 0 is original vertex index, code < 0 encodes
 * the intersection on edge (-code - 1) -> (-code) mod n. Returns the vertex
 * count after clipping. Pinned against every case of the reference table
 * (tests/golden/clip_cases.json). */
int orc_clip_rule(int n, uint32_t mask, int out_codes[RISLTC_MAX_POLYGON_VERTEX_COUNT + 1]);
/* rotation offset p of the rule (0 when nothing is clipped) */
int orc_clip_rotation(int n, uint32_t mask);

/* ---------------- LBVH (project structure; no reference code) ------- */
typedef struct orc_bvh_s {
    uint32_t leaf_count;          /* triangles                                 */
    float* node_box;              /* (2n-1) x 6: min xyz, max xyz              */
    int32_t* node_left;           /* internal: left child node id              */
    int32_t* node_right;          /* internal: right child node id             */
    int32_t* node_escape;         /* pre-order escape link, -1 = end           */
    int32_t* node_parent;
    float* leaf_tri;              /* n x 9 vertex coordinates, sorted order    */
    uint32_t* leaf_prim;          /* original triangle id, sorted order        */
    uint64_t* sorted_keys;        /* (morton << 32) | triangle id              */
} orc_bvh_t;
/* Node ids: internal nodes 0 .. n-2, leaves n-1 .. 2n-2 (leaf k = n-1+k). */
int orc_bvh_build(orc_bvh_t* bvh, const float* triangles, uint32_t triangle_count);
/* binned-SAH tree (the product's default builder, risltc_amd/csrc/sah.hip):
 * internal nodes numbered in pre-order, leaves left to right, no keys */
int orc_bvh_build_sah(orc_bvh_t* bvh, const float* triangles, uint32_t triangle_count);
/* the pre-order records of the device layout (5 n - 2 units of 4 floats):
 * internal (min.xyz | escape unit, max.xyz | 0), leaf (v0 | -1, v1 | id,
 * v2 | light flag from the *.vks positions) */
void orc_bvh_records(const orc_bvh_t* bvh, const uint32_t* positions, float* out);
void orc_bvh_free(orc_bvh_t* bvh);
uint32_t orc_morton_part_1_by_2(uint32_t x);
/* 1 = hit in [t_min, t_max] */
int orc_bvh_any_hit(const orc_bvh_t* bvh, const float org[3], const float dir[3], float t_min,
                    float t_max, uint32_t* visited_nodes);
/* closest hit primitive id or 0xFFFFFFFF */
uint32_t orc_bvh_closest_hit(const orc_bvh_t* bvh, const float org[3], const float dir[3],
                             float t_min, float t_max, float* t_hit);
int orc_ray_triangle(const float org[3], const float dir[3], const float* v, float t_min,
                     float t_max, float* t_out);
int orc_ray_box(const float org[3], const float inv_dir[3], const float* box, float t_min,
                float t_max);
/* primary visibility: closest hit of the camera ray of every pixel, as the
 * visibility pass produces it (visibility_pass.vert.glsl:27-38): primitive
 * id, bit 31 for light triangles, 0xFFFFFFFF for background. The camera ray
 * is pixel_to_ray * (x, y, 1) with t in [0, inf). */
void orc_render_visibility(const orc_bvh_t* bvh, const uint32_t* positions,
                           const risltc_frame_constants_t* c, uint32_t* out, int threads);
/* brute force over all triangles (test reference) */
int orc_brute_any_hit(const float* triangles, uint32_t count, const float org[3],
                      const float dir[3], float t_min, float t_max);
/* dequantised triangle positions of a *.vks mesh: 9 floats per triangle */
void orc_dequantize_mesh(const uint32_t* positions, uint64_t triangle_count,
                         const float factor[3], const float summand[3], float* out_triangles);

/* ---------------- the shading path (orc_core.c, _f32 / _f64) -------- */
typedef struct orc_scene_s {
    /* mesh (src/scene.h:47-70) */
    const uint32_t* positions;
    const uint16_t* normals_uv;
    const uint8_t* material_indices;
    uint64_t triangle_count;
    const risltc_material_t* materials;
    uint32_t material_count;
    /* lights as packed by write_lights */
    const uint8_t* light_records;
    uint32_t light_count;
    uint32_t light_max_vertex_count;
    /* LTC tables (UNORM16), roughness fastest */
    const uint16_t* ltc_rgba16;
    const uint16_t* ltc_rg16;
    uint32_t ltc_res_roughness, ltc_res_inclination, ltc_fresnel_count;
    /* shadow rays */
    const orc_bvh_t* bvh;
    /* learned RIS source (project extension; read when settings.proposal is
     * RISLTC_PROPOSAL_LEARNED): per pixel table as risltc_proposal_update
     * writes it (32 bytes: 16 group ids, 16 weights) */
    const uint8_t* proposal_table;
    const uint32_t* cluster_of_light;
    const uint32_t* cluster_offsets; /* cluster_count + 1 prefix offsets      */
    const uint32_t* cluster_lights;  /* light indices sorted by cluster       */
    uint32_t cluster_count;
    /* visibility cache (project extension; NULL = none): one byte per
     * (cell, slot), cell-major, as orc_cache_bytes makes them from the
     * counts; 2^k cells, mask = 2^k - 1; a slot is a light's cluster, or the
     * light with cache_per_light */
    const uint8_t* cache_vis;
    uint32_t cache_mask;
    float cache_inv_cell;
    uint32_t cache_slots;
    uint32_t cache_per_light;
    uint32_t proposal_table_bytes; /* 32 */
    uint32_t proposal_group_log2;  /* a table slot names 1 << this clusters */
} orc_scene_t;

/* Bytes of the visibility cache from its counts (2 words per entry:
 * unoccluded a, occluded b): 255 for an empty entry, else
 * round(254 (a + prior / 2) / (a + b + prior)) in float. */
void orc_cache_bytes(const uint32_t* counts, uint64_t entries, float prior, uint8_t* out);

/* Shades the listed pixels. out_rgba: 4 floats per listed pixel. samples
 * (may be NULL): spp records of RISLTC_SAMPLE_RECORD_FLOATS per listed pixel.
 * candidates (may be NULL): spp * M records of RISLTC_CANDIDATE_RECORD_WORDS.
 * Returns 0 or an RISLTC_E_* code. */
int orc_shade_pixels_f32(const orc_scene_t* scene, const risltc_frame_constants_t* c,
                         const risltc_settings_t* s, const uint32_t* visibility,
                         const uint32_t* pixel_indices, uint32_t pixel_count, float* out_rgba,
                         float* samples, uint32_t* candidates);
int orc_shade_pixels_f64(const orc_scene_t* scene, const risltc_frame_constants_t* c,
                         const risltc_settings_t* s, const uint32_t* visibility,
                         const uint32_t* pixel_indices, uint32_t pixel_count, double* out_rgba,
                         float* samples, uint32_t* candidates);
/* OpenMP-parallel FP32 shade of all pixels (CPU baseline timing) */
int orc_shade_frame_f32(const orc_scene_t* scene, const risltc_frame_constants_t* c,
                        const risltc_settings_t* s, const uint32_t* visibility,
                        const uint32_t* pixel_indices, uint32_t pixel_count, float* out_rgba,
                        int threads);

/* Unit-level entry points (tests). Vertices are xyz triples. */
uint32_t orc_clip_polygon_f32(uint32_t vertex_count, float* v /* 9x3 in/out */);
uint32_t orc_clip_polygon_f64(uint32_t vertex_count, double* v);
float orc_calculate_ltc_f32(uint32_t vertex_count, const float* v);
double orc_calculate_ltc_f64(uint32_t vertex_count, const double* v);
/* prepare_projected_solid_angle_polygon_sampling: returns the projected solid
 * angle; out_state (may be NULL) receives 48 values: vertices 9x2,
 * ellipses 9x2, inner_ellipse_0 2, sector psa 9, psa 1 */
float orc_prepare_psa_f32(uint32_t vertex_count, const float* v, float* out_state);
double orc_prepare_psa_f64(uint32_t vertex_count, const double* v, double* out_state);
/* prepare + sample_projected_solid_angle_polygon for each (u0, u1) pair */
void orc_sample_psa_f32(uint32_t vertex_count, const float* v, const float* uv, uint32_t n,
                        float* out_dirs);
void orc_sample_psa_f64(uint32_t vertex_count, const double* v, const double* uv, uint32_t n,
                        double* out_dirs);
/* evaluate_brdf(diffuse = specular = true): shading data packed as
 * normal[3], outgoing[3], diffuse_albedo[3], fresnel_0[3], roughness */
void orc_evaluate_brdf_f32(const float* shading, const float* dir, float* out_rgb);
void orc_evaluate_brdf_f64(const double* shading, const double* dir, double* out_rgb);
/* get_ltc_coefficients: writes world_to_shading 12 (row-major 3x4),
 * shading_to_cosine 9, cosine_to_shading 9, albedo, determinant */
void orc_ltc_coefficients_f32(const orc_scene_t* scene, const risltc_ltc_constants_t* k,
                              float fresnel_luminance, float roughness, const float* position,
                              const float* normal, const float* outgoing, float* out);
/* One bilinearly filtered channel of the LTC arrays at texture coordinates
 * (u, v) and array coordinate w: clamp to edge, layer = RNE(w) clamped.
 * Pinned by affine tables (exact reproduction of an affine texel function at
 * interior points, clamping at the edges, round-half-even layers). */
float orc_ltc_fetch_f32(const orc_scene_t* scene, int table, int channel, float u, float v,
                        float w);
double orc_ltc_fetch_f64(const orc_scene_t* scene, int table, int channel, double u, double v,
                         double w);
/* Mixed precision (the integrator apart from the float shading point):
 * orc_pixel_setup writes the per-pixel setup of get_shading_data and
 * get_ltc_coefficients for pixel index `pixel` whose visible triangle is
 * `prim` (ORC_PIXEL_SETUP_FLOATS floats); orc_ltc_cp_target evaluates the
 * LTC_CP RIS target |eval_ltc_cp| of one light from a setup in the build's
 * arithmetic, clip_counts (may be NULL) = vcd | vcs << 8. The FP32 pair
 * reproduces the FP32 path's candidate targets bit for bit. */
#define ORC_PIXEL_SETUP_FLOATS 55
int orc_pixel_setup_f32(const orc_scene_t* scene, const risltc_frame_constants_t* c, uint32_t pixel, uint32_t prim,
                        float* out);
int orc_pixel_setup_f64(const orc_scene_t* scene, const risltc_frame_constants_t* c, uint32_t pixel, uint32_t prim,
                        float* out);
float orc_ltc_cp_target_f32(const orc_scene_t* scene, const float* setup, uint32_t light, uint32_t* clip_counts);
double orc_ltc_cp_target_f64(const orc_scene_t* scene, const float* setup, uint32_t light, uint32_t* clip_counts);

/* decode_position_64_bit / decode_normal_32_bit (mesh_quantization.glsl:19-45).
 * Pinned by round trips through the exporter's packing and octahedral
 * encoding (tools/io_export_vulkan_blender28.py:24-46, 486-512): position =
 * correctly rounded fma(q, f, s); normal within the 16-bit quantisation bound. */
void orc_decode_position_f32(uint32_t q0, uint32_t q1, const float f[3], const float s[3], float* out);
void orc_decode_position_f64(uint32_t q0, uint32_t q1, const float f[3], const float s[3], double* out);
void orc_decode_normal_f32(uint32_t qx, uint32_t qy, float* out);
void orc_decode_normal_f64(uint32_t qx, uint32_t qy, double* out);

/* risltc_sample_record_t layout helpers */
#define ORC_REC_LIGHT 0
#define ORC_REC_WSUM 1
#define ORC_REC_PHAT 2
#define ORC_REC_W 3
#define ORC_REC_CLIPD 4
#define ORC_REC_CLIPS 5
#define ORC_REC_PSAD 6
#define ORC_REC_PSAS 7
#define ORC_REC_DIR 8
#define ORC_REC_TMAX 14
#define ORC_REC_STATE 16
#define ORC_REC_EVIS 18
#define ORC_REC_EOCC 24
#define ORC_REC_COLOR 30
#define ORC_REC_SEED 33
#define ORC_REC_POS 34
#define ORC_REC_CACHE_KEY 37

#ifdef __cplusplus
}
#endif
#endif
