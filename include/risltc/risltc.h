/*
 * risltc.h -- C-ABI of the MI355X (gfx950) RIS-LTC shading pass.
 *
 * This header is the drop-in boundary. It replaces the Vulkan shading subpass
 * of the reference renderer: descriptor set 0 of the shading pass
 * (src/main.c:845-856) plus the full-screen draw in subpass 1
 * (src/main.c:2050-2055) become one call, risltc_shade(). The host data that
 * the reference uploads keeps its byte layout:
 *   - light records: exactly what write_lights() packs (src/main.c:456-490),
 *     i.e. polygonal_light_upload_t (src/polygonal_light.h:104-111) followed
 *     by max_vertex_count vertices of 16 bytes each;
 *   - LTC tables: the two UNORM16 texture arrays that load_ltc_table()
 *     builds (src/ltc_table.c:82-116) and ltc_constants_t
 *     (src/ltc_table.h:23-35);
 *   - per-frame constants: per_frame_constants_t (src/main.h:537-553), the
 *     std140 UBO of src/shaders/shared_constants.glsl:21-60;
 *   - mesh buffers: the quantised *.vks layout of mesh_t (src/scene.h:47-70):
 *     2 x uint32 positions per vertex, 4 x uint16 normal+uv per vertex, one
 *     uint8 material index per triangle (src/scene.c:79-83);
 *   - visibility buffer: R32_UINT primitive ids, 0xFFFFFFFF = background,
 *     bit 31 = light (src/shaders/visibility_pass.vert.glsl:30-37).
 *
 * Conventions follow the reference host code: caller-allocated structs,
 * int return with 0 = success (src/ltc_table.c:38-42), plain C types only,
 * extern "C" for C++ callers (src/polygonal_light.h:21-28). Large per-pixel
 * buffers (visibility, output, debug records) are DEVICE pointers (HIP
 * device memory) owned by the caller; everything else is host memory that
 * is copied before the call returns. A context is not thread-safe.
 */
#ifndef RISLTC_H
#define RISLTC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Error codes                                                               */
/* ------------------------------------------------------------------------ */
#define RISLTC_OK 0
#define RISLTC_E_ARG 1         /* invalid argument (null pointer, bad size)  */
#define RISLTC_E_HIP 2         /* HIP runtime error                          */
#define RISLTC_E_FORMAT 3      /* malformed input data (file, record)        */
#define RISLTC_E_UNSUPPORTED 4 /* setting outside what the kernels support   */
#define RISLTC_E_STATE 5       /* call order: e.g. shade before set_lights   */
#define RISLTC_E_IO 6          /* file could not be opened or read           */

/* Largest light polygon (vertices before clipping). The reference clipper
 * encodes the count in 3 bits and so handles at most 7
 * (src/shaders/polygon_clipping.glsl:39-42); this build extends it to 8. */
#define RISLTC_MAX_LIGHT_VERTEX_COUNT 8
/* Vertices after clipping: one more than before (src/main.c:191-204). */
#define RISLTC_MAX_POLYGON_VERTEX_COUNT 9

/* ------------------------------------------------------------------------ */
/* Host data model (byte-compatible with the reference)                      */
/* ------------------------------------------------------------------------ */

/* polygonal_light_upload_t (src/polygonal_light.h:104-111); 48 bytes. In the
 * buffer each header is followed by max_vertex_count vertices, 16 bytes each
 * (xyz + padding); write_lights() repeats vertex 0 after the last vertex
 * (src/main.c:483-488). */
typedef struct risltc_light_upload_s {
    float surface_radiance[3];
    float padding_1;
    float plane[4];
    uint32_t vertex_count;
    float padding_2[3];
} risltc_light_upload_t;

#define RISLTC_LIGHT_RECORD_SIZE(max_vertex_count) (48u + 16u * (uint32_t)(max_vertex_count))

/* polygonal_light_t (src/polygonal_light.h:73-99): host-side light in plane
 * space, turned into world space by risltc_update_polygonal_light(). */
typedef struct risltc_polygonal_light_s {
    float rotation_angles[3];
    float scaling_x;
    float translation[3];
    float scaling_y;
    float radiant_flux[3];
    float inv_scaling_x;
    float surface_radiance[3];
    float inv_scaling_y;
    float plane[4];
    uint32_t vertex_count;
    uint32_t texturing_technique;
    uint32_t texture_index;
    uint32_t padding_0;
    float rotation[3][4];
    float area, rcp_area;
    float padding_1[2];
    char* texture_file_path;
    float* vertices_plane_space; /* vertex i at [4i+0], [4i+1] */
    float* vertices_world_space; /* vertex i at [4i+0 .. 4i+2] */
} risltc_polygonal_light_t;

/* ltc_constants_t (src/ltc_table.h:23-35); 32 bytes. */
typedef struct risltc_ltc_constants_s {
    float fresnel_index_factor, fresnel_index_summand;
    float roughness_factor, roughness_summand;
    float inclination_factor, inclination_summand;
    float padding[2];
} risltc_ltc_constants_t;

/* per_frame_constants_t (src/main.h:537-553); 256 bytes, std140 offsets of
 * the UBO in src/shaders/shared_constants.glsl:21-60. VkExtent2D is two
 * uint32_t. */
typedef struct risltc_frame_constants_s {
    float mesh_dequantization_factor[3], padding_0, mesh_dequantization_summand[3];
    float error_factor;
    float world_to_projection_space[4][4];
    float pixel_to_ray_direction_world_space[3][4];
    float camera_position_world_space[3];
    float mis_visibility_estimate;
    uint32_t viewport_size[2];
    int32_t cursor_position[2];
    float exposure_factor;
    float roughness_factor;
    uint32_t noise_resolution_mask[2];
    uint32_t noise_texture_index_mask;
    uint32_t padding_3[3];
    uint32_t noise_random_numbers[4];
    risltc_ltc_constants_t ltc_constants;
} risltc_frame_constants_t;

/* first_person_camera_t (src/camera.h:29-49). */
typedef struct risltc_camera_s {
    float position_world_space[3];
    float rotation_z;
    float rotation_x;
    float vertical_fov;
    float near_plane, far_plane;
    float speed;
    int32_t rotate_camera;
    float rotation_x_0, rotation_z_0;
} risltc_camera_t;

/* Constant material. It stands in for the three material textures that
 * get_shading_data() samples (src/shaders/shading_pass.frag.glsl:629-635):
 * base_color is the base colour texel, specular is the (r, g, b) texel of
 * the specular texture (g = linear roughness, b = metalicity), and the
 * normal-map texel is the flat (0.5, 0.5). Project extension: texture
 * sampling is outside the shading-pass boundary. */
typedef struct risltc_material_s {
    float base_color[3];
    float specular[3];
} risltc_material_t;

/* mis_heuristic_t (src/main.h:58-76) */
#define RISLTC_MIS_BALANCE 0
#define RISLTC_MIS_POWER 1
#define RISLTC_MIS_WEIGHTED 2
#define RISLTC_MIS_OPTIMAL_CLAMPED 3
#define RISLTC_MIS_OPTIMAL 4

/* light_sampling_strategies_t (src/main.h:78-83) */
#define RISLTC_LIGHT_UNIFORM 0
#define RISLTC_LIGHT_RESERVOIR 1

/* sample_polygon_technique_t (src/polygonal_light.h:30-45). LTC_CP is the
 * reference default (src/main.c:227) and the production path. The two
 * baselines of the reference's comparisons (src/experiment_list.c:313-396)
 * are built with the uniform source only:
 *   AREA_TURK             -- each candidate is a uniform point on the
 *                            triangle of vertices 0, 1, 2 (Turk), target =
 *                            its unshadowed area-sampling estimate; the final
 *                            sample traces one shadow ray to that point
 *                            (shading_pass.frag.glsl:410-418, 757-759);
 *   PROJECTED_SOLID_ANGLE -- each candidate's target is the unshadowed
 *                            estimate of one projected-solid-angle sample per
 *                            technique with MIS; the final sample is the
 *                            LTC_CP final sample (shading_pass.frag.glsl:
 *                            458-551).
 * They run in a simpler per-pixel kernel and not in batches or graphs. */
#define RISLTC_SAMPLE_POLYGON_AREA_TURK 1
#define RISLTC_SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE 2
#define RISLTC_SAMPLE_POLYGON_LTC_CP 4

/* RIS source distribution */
#define RISLTC_PROPOSAL_UNIFORM 0 /* reference: p = 1/N (shading_pass.frag.glsl:733) */
#define RISLTC_PROPOSAL_LEARNED 1 /* 0.5 * MLP cluster proposal + 0.5 / N,
                                     target with the visibility cache if set */
#define RISLTC_PROPOSAL_CACHE 2   /* uniform source (three draws per
                                     candidate), target with the visibility
                                     cache: the learned pass without the MLP */

/* Render settings. The fields mirror render_settings_t (src/main.h:123-156)
 * and the shader defines built from it (src/main.c:962-991); what the
 * reference fixes at compile time is a runtime value here. */
/* SAMPLE_COUNT: samples per technique of every light evaluation that
 * samples (the final sample of every technique, the candidates of the area
 * and projected-solid-angle techniques), averaged
 * (shading_pass.frag.glsl:365-396, 412-417, 518-559). Each pixel-sample
 * holds SAMPLE_COUNT records and up to 2 * SAMPLE_COUNT shadow rays; a call
 * holds at most 2^30 records (pixels x spp x SAMPLE_COUNT). */
#define RISLTC_MAX_SAMPLE_COUNT 1024u
typedef struct risltc_settings_s {
    uint32_t sample_count;               /* SAMPLE_COUNT, 1..RISLTC_MAX_SAMPLE_COUNT */
    uint32_t sample_count_light;         /* LIGHT_SAMPLES (samples per pixel)     */
    uint32_t mis_heuristic;              /* RISLTC_MIS_*                          */
    uint32_t light_sampling;             /* RISLTC_LIGHT_*                        */
    uint32_t polygon_sampling_technique; /* RISLTC_SAMPLE_POLYGON_LTC_CP          */
    uint32_t ris_candidate_count;        /* M; the reference hard-codes 32
                                            (shading_pass.frag.glsl:726)          */
    uint32_t fast_atan;                  /* USE_FAST_ATAN                          */
    uint32_t shadow_rays;                /* 1 = trace shadow rays (reference)     */
    uint32_t proposal;                   /* RISLTC_PROPOSAL_*                     */
    uint32_t exact_math;                 /* 1 = IEEE div/sqrt everywhere: results
                                            bit-identical to the FP32 oracle      */
    uint32_t learned_floor_log2;         /* learned target: p_hat (f + (1 - f) v / 255),
                                            with a visibility cache,
                                            f = 2^-learned_floor_log2 (0 -> 3)    */
    uint32_t skip_training_records;      /* 1: no proposal training records
                                            (risltc_proposal_train then needs
                                            a later call without it)          */
    uint32_t reserved[4];
} risltc_settings_t;

/* Fills settings with the reference defaults (src/main.c:220-236): SAMPLE_COUNT
 * 1, one light sample, optimal clamped MIS, reservoir light sampling, LTC_CP,
 * M = 32, shadow rays on, exact_math off. */
void risltc_default_settings(risltc_settings_t* settings);

/* ------------------------------------------------------------------------ */
/* Host-side helpers that restate the reference's CPU code                   */
/* ------------------------------------------------------------------------ */

/* update_polygonal_light (src/polygonal_light.c:44-98). vertices_world_space
 * must be allocated (4 floats per vertex). */
void risltc_update_polygonal_light(risltc_polygonal_light_t* light);

/* write_lights (src/main.c:456-490): calls risltc_update_polygonal_light on
 * every light, then packs the records (stride RISLTC_LIGHT_RECORD_SIZE(
 * max_vertex_count)). Unlike the reference, which writes vertex 0 once after
 * the last vertex and leaves later slots untouched, every unused slot holds
 * vertex 0. Returns the number of bytes written or 0 on error. */
size_t risltc_write_lights(risltc_polygonal_light_t* lights, uint32_t light_count,
                           uint32_t max_vertex_count, void* out, size_t out_size);

/* write_constants (src/main.c:2902-2946) with get_world_to_projection_space
 * (src/camera.c:24-83) and set_noise_constants (src/noise_table.c:24-28).
 * noise_random_seed is the noise_table_t::random_seed that the reference
 * increments each frame; animate_noise selects wang_random_number() numbers
 * instead of i * 0x123456. The LTC constants are copied from ltc. */
void risltc_write_frame_constants(risltc_frame_constants_t* out, const risltc_camera_t* camera,
                                  uint32_t width, uint32_t height,
                                  const float mesh_dequantization_factor[3],
                                  const float mesh_dequantization_summand[3],
                                  const risltc_ltc_constants_t* ltc, float exposure_factor,
                                  float roughness_factor, float mis_visibility_estimate,
                                  uint32_t noise_random_seed, int animate_noise);

/* ltc_constants_t as load_ltc_table computes them (src/ltc_table.c:184-191). */
void risltc_ltc_constants_for(risltc_ltc_constants_t* out, uint32_t roughness_count,
                              uint32_t inclination_count, uint32_t fresnel_count);

/* Reads <directory>/fit<i>.dat for i < fresnel_count and quantises them the
 * way load_ltc_table does (src/ltc_table.c:31-116). rgba16 receives
 * res*res*4*fresnel_count values, rg16 res*res*2*fresnel_count values; pass
 * NULL buffers to query *out_resolution first. Returns RISLTC_E_IO when a file
 * is missing, RISLTC_E_FORMAT on inconsistent resolutions or short files. */
int risltc_load_ltc_fit(const char* directory, uint32_t fresnel_count, uint32_t* out_resolution,
                        uint16_t* rgba16, uint16_t* rg16, size_t capacity_values);

/* ------------------------------------------------------------------------ */
/* On-disk formats                                                           */
/* ------------------------------------------------------------------------ */

/* Header of a *.vks scene file (src/scene.c:419-433). */
typedef struct risltc_vks_info_s {
    uint64_t material_count;
    uint64_t triangle_count;
    float dequantization_factor[3];
    float dequantization_summand[3];
    uint64_t material_name_bytes; /* sum over materials of (name length + 1) */
} risltc_vks_info_t;

/* The file reading of load_scene (src/scene.c:409-483; written by
 * io_export_vulkan_blender28.py:481-541): uint32 marker 0xabcabc and version
 * 1, uint64 material and triangle counts, 3 + 3 float dequantisation
 * constants, per material a uint64 length, the name and a NUL, then the mesh
 * buffers in the layout risltc_set_scene takes (positions 24 B, normals and
 * texture coordinates 24 B, material index 1 B per triangle) and the uint32
 * end marker 0xE0FE0F. With positions, normals_and_tex_coords and
 * material_indices all NULL only *info is read (the whole file is still
 * checked); otherwise all three must hold triangle_capacity >= triangle_count
 * triangles. names (optional, names_capacity bytes) receives the
 * NUL-terminated names back to back. Returns RISLTC_E_IO when the file cannot
 * be opened, RISLTC_E_FORMAT on a wrong marker or version, zero triangles, a
 * short file or a missing end marker, RISLTC_E_ARG when a buffer is too
 * small. */
int risltc_load_vks(const char* path, risltc_vks_info_t* info, uint32_t* positions,
                    uint16_t* normals_and_tex_coords, uint8_t* material_indices,
                    uint64_t triangle_capacity, char* names, uint64_t names_capacity);

/* Totals of a quicksave file. */
typedef struct risltc_quicksave_info_s {
    uint32_t light_count;
    uint32_t total_vertex_count;  /* sum of vertex_count over the lights */
    uint64_t texture_path_bytes;  /* sum of the texture path sizes (with NUL) */
} risltc_quicksave_info_t;

/* quick_load (src/main.c:78-125): the camera (48 bytes), a uint32 legacy
 * count, the uint32 light count, then per light the first 88 bytes of
 * risltc_polygonal_light_t (rotation_angles .. texturing_technique), a
 * uint64 texture path size and path, 16 bytes of former pointers and
 * 4 floats per plane-space vertex. A light with scaling_y <= 0 gets
 * scaling_y = scaling_x, as in the reference. With lights == NULL only *info
 * (and *camera if not NULL) is filled. Otherwise light i's plane-space and
 * (zeroed) world-space vertex arrays are carved out of vertex_storage (8
 * floats per vertex: the plane-space vertices of a light, then its
 * world-space ones), its texture_file_path points into path_storage (or is
 * NULL) and the fields after texturing_technique are zero; run
 * risltc_write_lights (which calls risltc_update_polygonal_light) next.
 * Errors as for risltc_load_vks; a vertex count of 0 or above 65536 is
 * RISLTC_E_FORMAT. */
int risltc_load_quicksave(const char* path, risltc_camera_t* camera, risltc_quicksave_info_t* info,
                          risltc_polygonal_light_t* lights, uint32_t light_capacity,
                          float* vertex_storage, uint64_t vertex_storage_floats, char* path_storage,
                          uint64_t path_capacity);

/* quick_save (src/main.c:44-75): writes the file risltc_load_quicksave reads. */
int risltc_save_quicksave(const char* path, const risltc_camera_t* camera,
                          const risltc_polygonal_light_t* lights, uint32_t light_count);

/* Radiance RGBE image, as the reference's HDR screenshots are written with
 * stbi_write_hdr (src/main.c:2392-2401): rows top to bottom ("-Y h +X w"),
 * scanlines of 8..32767 pixels run-length coded per component, others flat.
 * The pixel bytes equal stbi_write_hdr's for finite non-negative input; the
 * header comment and EXPOSURE lines are left out. pixels holds
 * width*height*channels floats (channels 3 or 4; alpha is ignored). With
 * half_round = 1 every value is rounded to binary16 (to nearest even) first,
 * as the screenshot path stores halves (copy_pass.frag.glsl:36-46,
 * src/main.c:2339-2349). Negative and NaN components are written as 0,
 * values of 2^127 and above as the largest value below it. */
int risltc_write_hdr(const char* path, const float* pixels, uint32_t width, uint32_t height,
                     uint32_t channels, int half_round);

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */
typedef struct risltc_context_s risltc_context_t;

/* Creates a context on HIP device device_index. */
int risltc_create(risltc_context_t** out_context, int device_index);
void risltc_destroy(risltc_context_t* context);
/* Text of the last error on this context (never NULL). */
const char* risltc_last_error(const risltc_context_t* context);

/* Light records as packed by write_lights: count records of stride
 * RISLTC_LIGHT_RECORD_SIZE(max_vertex_count). max_vertex_count <= 8,
 * every vertex_count in [3, max_vertex_count]. Host pointer. */
int risltc_set_lights(risltc_context_t* context, const void* records, uint32_t count,
                      uint32_t max_vertex_count);

/* Convenience over host-side lights: risltc_write_lights (which runs
 * update_polygonal_light on every light, src/main.c:456-490) followed by
 * risltc_set_lights. max_vertex_count = 0 takes the largest vertex_count of
 * the lights. */
int risltc_set_polygonal_lights(risltc_context_t* context, risltc_polygonal_light_t* lights,
                                uint32_t count, uint32_t max_vertex_count);

/* Convenience with load_ltc_table's file semantics (src/ltc_table.c:23-194):
 * reads <directory>/fit<i>.dat for i < fresnel_count with
 * risltc_load_ltc_fit and uploads the square table with
 * risltc_set_ltc_table. Errors as for risltc_load_ltc_fit. */
int risltc_load_ltc_table(risltc_context_t* context, const char* directory, uint32_t fresnel_count);

/* LTC tables as load_ltc_table builds them: rgba16 holds fresnel_count slices
 * of roughness_count x inclination_count texels (roughness fastest), 4
 * channels; rg16 the same with 2 channels. Host pointers. */
int risltc_set_ltc_table(risltc_context_t* context, const uint16_t* rgba16, const uint16_t* rg16,
                         uint32_t roughness_count, uint32_t inclination_count,
                         uint32_t fresnel_count);

/* Shadow-ray hierarchy builders. Both lay the tree out in depth-first
 * pre-order for the same stackless traversal; hit and miss flags are the
 * same for any tree (the oracle restates both builders bit for bit).
 *   RISLTC_BVH_SAH  -- top-down binned surface area heuristic (default);
 *   RISLTC_BVH_LBVH -- 30-bit Morton codes of triangle centroids, sorted,
 *                      Karras hierarchy. */
#define RISLTC_BVH_SAH 0u
#define RISLTC_BVH_LBVH 1u
/* Selects the builder used by the next risltc_set_scene. No reference
 * counterpart: the reference's acceleration structure is built by the
 * Vulkan driver (src/scene.c:142-406). */
int risltc_set_bvh_builder(risltc_context_t* context, uint32_t builder);
/* Copies the current hierarchy (16-byte units in pre-order, see
 * risltc_set_scene) to host memory: at most `capacity` units; *units
 * receives the total. Test and tooling entry point. */
int risltc_bvh_records(risltc_context_t* context, float* out, uint64_t capacity, uint32_t* units);

/* Mesh in the *.vks layout (src/scene.h:47-70) plus a constant material per
 * material index. Uploads it and builds the hierarchy used for shadow rays
 * with the selected builder (risltc_set_bvh_builder), its nodes stored in
 * depth-first pre-order with escape links: internal node = 2 units
 * (min.xyz | escape unit, max.xyz | 0), leaf = 3 units (v0.xyz | -1,
 * v1.xyz | triangle id, v2.xyz | light flag bit 31). Light triangles
 * (bit 31 of the second position word) are part of the hierarchy, as they
 * are in the reference BLAS (src/scene.c:176-187). Host pointers. At most
 * 400,000,000 triangles (RISLTC_E_UNSUPPORTED beyond). */
int risltc_set_scene(risltc_context_t* context, const uint32_t* positions,
                     const uint16_t* normals_and_tex_coords, const uint8_t* material_indices,
                     uint64_t triangle_count, const float dequantization_factor[3],
                     const float dequantization_summand[3], const risltc_material_t* materials,
                     uint32_t material_count);

/* Primary visibility: writes the R32_UINT visibility buffer (primitive id of
 * the closest hit, bit 31 set for light triangles, 0xFFFFFFFF for
 * background) for the camera rays of constants. This is the step before the
 * shading pass (src/shaders/visibility_pass.vert.glsl:27-38), done by ray
 * casting into the LBVH instead of rasterisation. out_visibility: device
 * pointer, width*height. stream: hipStream_t or NULL. */
int risltc_render_visibility(risltc_context_t* context, const risltc_frame_constants_t* constants,
                             uint32_t* out_visibility, void* stream);

/* The shading pass: main() of src/shaders/shading_pass.frag.glsl:674-770 for
 * every pixel of constants->viewport_size. visibility: device pointer,
 * width*height. out_rgba32f: device pointer, width*height*4 floats. */
int risltc_shade(risltc_context_t* context, const risltc_frame_constants_t* constants,
                 const risltc_settings_t* settings, const uint32_t* visibility, float* out_rgba32f,
                 void* stream);

/* risltc_shade without device allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory during a frame
 * (code under stream capture, tight serving loops, hosts that forbid
 * hipMalloc). The sample records, shadow-ray list, hit words and counters
 * live in one fixed device array that the library defines, sized for at
 * most RISLTC_NOALLOC_MAX_SAMPLES sample records (width * height * spp *
 * sample_count; about 72 MiB, held by every GPU the library runs on until the process
 * ends). Larger calls fail with RISLTC_E_UNSUPPORTED. The output equals
 * risltc_shade's for the same inputs; no training records for
 * risltc_proposal_train are written. Lights, LTC table, scene, clusters,
 * proposal tables and visibility cache must be set beforehand (those calls
 * allocate). Every call on one GPU shares that GPU's array, from every
 * context: calls that overlap on one GPU (other streams, other threads)
 * race and must be serialised by the caller. */
#define RISLTC_NOALLOC_MAX_SAMPLES (1u << 19)
int risltc_shade_noalloc(risltc_context_t* context, const risltc_frame_constants_t* constants,
                         const risltc_settings_t* settings, const uint32_t* visibility, float* out_rgba32f,
                         void* stream);

/* Running mean of frames, the accumulation pass of
 * src/shaders/accum_pass.frag.glsl:45-53: accum = (accum * n + cur) / (n + 1).
 * Device pointers, pixel_count * 4 floats. */
int risltc_accumulate(risltc_context_t* context, float* accum_rgba32f, const float* frame_rgba32f,
                      uint32_t pixel_count, uint32_t accum_num, void* stream);

/* Any-hit shadow-ray queries against the scene LBVH with the semantics of
 * get_polygon_visibility (src/shaders/shading_pass.frag.glsl:112-129):
 * origin, direction, [t_min, t_max], opaque, no culling, terminate on first
 * hit. rays: device pointer, 8 floats per ray (ox, oy, oz, t_min, dx, dy,
 * dz, t_max). out_hit: device pointer, one uint32 (0/1) per ray. */
int risltc_trace_shadow_rays(risltc_context_t* context, const float* rays, uint32_t ray_count,
                             uint32_t* out_hit, void* stream);

/* Debug variant of risltc_shade for parity tests. For every pixel p and
 * light sample j it writes one record of RISLTC_SAMPLE_RECORD_FLOATS 32-bit
 * words at sample_records[(p * spp + j) * RISLTC_SAMPLE_RECORD_FLOATS]
 * (layout: risltc_sample_record_t) and, when candidate_records is not NULL,
 * M records of RISLTC_CANDIDATE_RECORD_WORDS words per pixel sample
 * (risltc_candidate_record_t). Device pointers. With sample_count > 1 the
 * ray, state and estimate fields describe the first iteration of the
 * SAMPLE_COUNT loop; color and noise_seed_end cover the whole loop. */
#define RISLTC_SAMPLE_RECORD_FLOATS 40
#define RISLTC_CANDIDATE_RECORD_WORDS 4
typedef struct risltc_sample_record_s {
    int32_t light_index;     /* RIS-selected light, -1 if none               */
    float w_sum;             /* reservoir weight sum                         */
    float p_hat;             /* target value of the selected light           */
    float W;                 /* w_sum / (M * p_hat), 0 if p_hat == 0         */
    uint32_t clip_diffuse;   /* clipped vertex count, shading space          */
    uint32_t clip_specular;  /* clipped vertex count, cosine space           */
    float psa_diffuse;       /* projected solid angle, shading space         */
    float psa_specular;      /* projected solid angle, cosine space          */
    float ray_dir[2][3];     /* world-space shadow ray directions            */
    float ray_tmax[2];       /* ray t_max (t_min = 1e-3)                     */
    uint32_t ray_state[2];   /* 0 no sample, 1 lambert <= 0, 2 traced hit,
                                3 traced miss, 4 untraced (shadow rays off)  */
    float estimate_visible[2][3];  /* MIS estimate if visible                */
    float estimate_occluded[2][3]; /* MIS estimate if occluded               */
    float color[3];          /* the light's estimate: shading_pass.frag.glsl
                                evaluate_polygonal_light_shading[_peters]()  */
    uint32_t noise_seed_end; /* LCG state after this sample                  */
    float position[3];       /* shading position (shadow ray origin)         */
    uint32_t cache_key;      /* visibility-cache cell of the pixel, 0 without */
    float pad[2];
} risltc_sample_record_t;
typedef struct risltc_candidate_record_s {
    int32_t light_index;
    uint32_t clip_counts;    /* diffuse | specular << 8                      */
    float p_hat;
    float w;
} risltc_candidate_record_t;

int risltc_shade_debug(risltc_context_t* context, const risltc_frame_constants_t* constants,
                       const risltc_settings_t* settings, const uint32_t* visibility,
                       float* out_rgba32f, float* sample_records, uint32_t* candidate_records,
                       void* stream);

/* ---- learned RIS proposal (project extension, no reference counterpart:
 * the reference draws RIS candidates uniformly and uses the unshadowed LTC
 * estimate as the target, shading_pass.frag.glsl:723-736). Two learned
 * parts, both trained online on the pass's own output:
 *   - source: an MLP maps each shading point to a distribution over light
 *     clusters (softmax of 256 or 1024 logits; per pixel the 16 kept
 *     clusters of risltc_proposal_update's table). With settings.proposal =
 *     RISLTC_PROPOSAL_LEARNED, risltc_shade draws each candidate from the
 *     mixture q = 1/(2N) + (1/2) pi_c / |c| of the uniform source and the
 *     table, so every light keeps q >= 1/(2N) and the estimator stays
 *     unbiased; the MLP is trained by softmax cross-entropy on the clusters
 *     of the lights the RIS selects.
 *   - target: with a visibility cache (risltc_proposal_cache) the target is
 *     p_hat * (f + (1 - f) v / 255), v the cache's visibility byte for the
 *     pixel's cell and the light, f = 1/8 by default
 *     (settings.learned_floor_log2): positive wherever p_hat is.
 * Sequence per frame: risltc_proposal_update (inference, any cadence),
 * risltc_shade (learned), then optionally risltc_proposal_train on the
 * frame's own samples. ------------------------------------------------- */
#define RISLTC_MAX_LIGHT_CLUSTERS 1024
/* Bytes per pixel of a proposal table: 16 one-byte group ids and 16 byte
 * weights, for every network. A group is one cluster for networks of 256
 * outputs, and the quad of clusters 4 q .. 4 q + 3 for 1024 outputs. */
#define RISLTC_PROPOSAL_TABLE_BYTES(outputs) 32u
#define RISLTC_PROPOSAL_PARAMS 31168      /* 96-64-64-64-256 MLP (default), FP32 */
#define RISLTC_PROPOSAL_PARAMS_WIDE 78464 /* 96-128-128-128-256 MLP, FP32 */
/* Explicit light clusters: cluster_of_light[i] < cluster_count <= 1024
 * (the learned source needs cluster_count <= the network's outputs). */
int risltc_set_light_clusters(risltc_context_t* context, const uint32_t* cluster_of_light,
                              uint32_t cluster_count);
/* k-means of the light centroids (deterministic start: light floor(k N / K)
 * for cluster k; empty clusters restart at the farthest light), then
 * risltc_set_light_clusters. More than 256 clusters are made in two levels:
 * ceil(K / 4) clusters of all lights, each split again into up to four
 * numbered 4 q .. 4 q + 3, so that every quad of consecutive ids is one
 * compact top-level cluster (the 1024-output network's tables sample quads).
 * out (host, may be NULL) receives the assignment. */
int risltc_cluster_lights(risltc_context_t* context, uint32_t cluster_count, uint32_t iterations,
                          uint32_t* out);
/* Hidden width of the MLP: 64 (default) or 128 (the 96-128-128-128-256
 * network). Changing it discards the current network; call
 * risltc_proposal_init or risltc_proposal_set_params next. Parameter arrays
 * then hold risltc_proposal_param_count floats (RISLTC_PROPOSAL_PARAMS or
 * RISLTC_PROPOSAL_PARAMS_WIDE), in the same order: every weight matrix
 * [out][in] followed by its bias. */
int risltc_proposal_set_width(risltc_context_t* context, uint32_t width);
/* Hidden width (64 or 128) and outputs (256 or 1024 clusters) together;
 * risltc_proposal_set_width keeps the outputs. The 1024-output network's
 * tables name quads of clusters (risltc_proposal_update). */
int risltc_proposal_set_network(risltc_context_t* context, uint32_t width, uint32_t outputs);
uint32_t risltc_proposal_param_count(const risltc_context_t* context);
/* Deterministic initialisation from a seed; resets the Adam state. */
int risltc_proposal_init(risltc_context_t* context, uint32_t seed);
/* Host copies of the FP32 parameters (risltc_proposal_param_count floats, every
 * weight matrix stored [out][in] followed by its bias) and of the last
 * training step's gradient; either pointer may be NULL. set resets Adam. */
int risltc_proposal_set_params(risltc_context_t* context, const float* params);
int risltc_proposal_get_params(risltc_context_t* context, float* params, float* gradient);
/* Inference for every pixel of the frame: writes the per-pixel tables the
 * learned shade reads, RISLTC_PROPOSAL_TABLE_BYTES each: 16 group ids
 * (bytes 0..15) and 16 weights (bytes 16..31). A group's value is the logit
 * z of a cluster (256 outputs), or log(sum exp z) over the non-empty
 * clusters of the quad 4 q .. 4 q + 3 (1024 outputs; id q). Slot 4 g + k
 * holds, among the groups of lane group g = (c % 16) / 4 with a non-empty
 * cluster, the k-th of the four with the largest values v (descending, ties
 * to the smaller id) with weight round(255 exp(v - v_max)), v_max over all
 * such groups; weights 0 for background. A candidate drawn from a slot
 * takes one light of the group's clusters uniformly. Optional device
 * outputs: a copy of the tables, the 256 (or 1024) logits and the 96 bf16
 * features per pixel. */
int risltc_proposal_update(risltc_context_t* context, const risltc_frame_constants_t* constants,
                           const uint32_t* visibility, uint8_t* table_out, float* logits_out,
                           uint16_t* features_out, void* stream);
/* Adam steps (lr, beta 0.9 / 0.999, eps 1e-8) on batches of pixel-samples of
 * the last risltc_shade: softmax cross-entropy of the logits against the
 * cluster of the sample's selected light (samples without one are left
 * out). batch is a multiple of 256. loss_out (host, may be NULL): the last
 * step's loss. */
int risltc_proposal_train(risltc_context_t* context, const risltc_frame_constants_t* constants,
                          const uint32_t* visibility, uint32_t batch, uint32_t steps,
                          float learning_rate, uint32_t seed, float* loss_out, void* stream);
/* The training records of the last risltc_shade (device outputs, may be
 * NULL): per pixel-sample the cluster of the selected light (0xFFFFFFFF =
 * none) and the fraction of its traced shadow rays that were unoccluded (0
 * without one); *count receives their number. */
int risltc_proposal_records(risltc_context_t* context, uint32_t* cluster_out, float* y_out,
                            uint64_t* count, void* stream);
/* Online visibility cache (project extension). 2^log2_entries hashed
 * world-space cells (cell = floor(position / cell_size) per axis, split by
 * the dominant axis and sign of the shading normal) x slots, where a slot is
 * a light cluster (1024 slots) or, with per_light = 1, a light (light-count
 * slots). Each entry holds two uint32 counts (a unoccluded, b occluded) and
 * one byte, 9 * slots * 2^log2_entries bytes of device memory in all
 * (log2_entries <= 20, counts at most 4 GiB). While a cache and light
 * clusters are set, every risltc_shade counts each traced shadow ray of its
 * final samples at (pixel's cell, slot of the sample's light) and, before
 * shading, recomputes the bytes of the cells counted in since:
 * 255 for an empty entry, else round(254 (a + prior_strength / 2) /
 * (a + b + prior_strength)). With settings.proposal =
 * RISLTC_PROPOSAL_LEARNED the target uses its entry's byte v (255 while
 * the entry is empty).
 * Allocates and clears; log2_entries = 0 removes the cache. New lights,
 * clusters or meshes clear it. */
int risltc_proposal_cache(risltc_context_t* context, uint32_t log2_entries, float cell_size,
                          float prior_strength, uint32_t per_light);
/* Host copies of the cache: counts as 2 words per (cell, slot), cell-major
 * (read needs capacity >= that many words, write exactly that many; a write
 * has every cell recomputed by the next shading call); bytes as one per
 * (cell, slot), as of the last shading call. */
int risltc_proposal_cache_read(risltc_context_t* context, uint32_t* counts_out, uint64_t capacity);
int risltc_proposal_cache_write(risltc_context_t* context, const uint32_t* counts, uint64_t words);
int risltc_proposal_cache_table(risltc_context_t* context, uint8_t* bytes_out, uint64_t capacity);
/* Milliseconds of the last risltc_proposal_update and risltc_proposal_train. */
int risltc_proposal_last_timings(const risltc_context_t* context, float out_ms[2]);

/* Milliseconds of the last risltc_shade call per kernel, measured with HIP
 * events on the call's stream: [0] shade, [1] shadow-ray trace, [2] resolve,
 * [3] learned-proposal inference. Entries not run are 0. */
int risltc_last_timings(const risltc_context_t* context, float out_ms[4]);

/* Shadow rays traced by the last risltc_shade or risltc_shade_noalloc call
 * (waits for it). No counterpart in the reference: the ray count of a frame,
 * for throughput and roofline figures. RISLTC_E_STATE before any such call. */
int risltc_last_ray_count(const risltc_context_t* context, uint32_t* out_rays);

/* ------------------------------------------------------------------------ */
/* Batches of views. No counterpart in the reference, which shades one view
 * per frame (main.c:2050-2055); a batch shades several views (cameras,
 * resolutions, settings, light sets, meshes) in one launch of each kernel,
 * the shadow rays of every view in one persistent trace over the views'
 * hierarchies, and a graph replays such a batch as one hipGraph. Every
 * view's frame is bitwise equal to a risltc_shade call with the same inputs
 * on a context holding its lights and mesh. */

/* A further light set (records as for risltc_set_lights) that views may
 * name; *out_id receives its id (>= 1). Id 0 names the lights of
 * risltc_set_lights. Light sets live until risltc_destroy. */
int risltc_add_light_set(risltc_context_t* context, const void* records, uint32_t count,
                         uint32_t max_vertex_count, uint32_t* out_id);

/* A further mesh (arguments as for risltc_set_scene, its hierarchy built
 * with the context's builder) that views may name; *out_id receives its id
 * (>= 1). Id 0 names the mesh of risltc_set_scene. Meshes live until
 * risltc_destroy. */
int risltc_add_scene(risltc_context_t* context, const uint32_t* positions,
                     const uint16_t* normals_and_tex_coords, const uint8_t* material_indices,
                     uint64_t triangle_count, const float dequantization_factor[3],
                     const float dequantization_summand[3], const risltc_material_t* materials,
                     uint32_t material_count, uint32_t* out_id);
/* risltc_render_visibility on the mesh with the given id. */
int risltc_render_visibility_scene(risltc_context_t* context, uint32_t scene,
                                   const risltc_frame_constants_t* constants, uint32_t* visibility_out,
                                   void* stream);

typedef struct risltc_view_s {
    risltc_frame_constants_t constants;
    risltc_settings_t settings;   /* exact_math and (proposal != UNIFORM)
                                     equal across a batch */
    const uint32_t* visibility;   /* device, R32_UINT per pixel */
    float* out;                   /* device, RGBA32F per pixel */
    uint32_t light_set;           /* 0 or an id from risltc_add_light_set */
    uint32_t scene;               /* 0 or an id from risltc_add_scene */
    const uint8_t* proposal_table; /* device, with settings.proposal =
                                      RISLTC_PROPOSAL_LEARNED: this view's
                                      table (width*height*
                                      RISLTC_PROPOSAL_TABLE_BYTES(outputs)
                                      bytes for the network's outputs) from
                                      risltc_proposal_update's table_out;
                                      NULL otherwise */
} risltc_view_t;

/* Shades count views (device pointers in each view) on stream.
 * Learned views (settings.proposal LEARNED or CACHE) run the learned pass:
 * they use light set 0 and scene 0 (the clusters, the MLP's features and
 * the visibility cache refer to them), share the context's visibility cache
 * (refreshed once before the batch; every view's shadow rays are counted
 * into it) and write no training records (train on risltc_shade frames).
 * A batch is either all learned or all plain. */
int risltc_shade_batch(risltc_context_t* context, const risltc_view_t* views, uint32_t count, void* stream);

/* Launch shape of the kernels. None of these changes any result bit.
 *   wave_width   8, 16, 32 or 64: pixels per row of a wave's tile (a wave
 *                shades wave_width x 64/wave_width pixels);
 *   cull_chunk   1..32: candidates tested against the bounding spheres per
 *                step of the cull loop;
 *   lds_lights   0..2048: bounding spheres staged in LDS per work-group
 *                (the rest are read from global memory; default 1024:
 *                more cost a work-group per CU). The learned pass also
 *                stages its cluster tables when they take at most 16 KiB;
 *                with a visibility cache it keeps three work-groups per CU
 *                at lds_lights = 0;
 *   trace_refill 1..64: idle lanes at which a wave of the persistent
 *                shadow-ray kernel fetches new rays. */
typedef struct risltc_tuning_s {
    uint32_t wave_width;
    uint32_t cull_chunk;
    uint32_t lds_lights;
    uint32_t trace_refill;
} risltc_tuning_t;

int risltc_get_tuning(const risltc_context_t* context, risltc_tuning_t* out);
int risltc_set_tuning(risltc_context_t* context, const risltc_tuning_t* tuning);
/* Coordinate descent over the tuning space on a batch: each parameter in turn
 * takes the value with the least median GPU time over `repeats` runs of the
 * batch, the others held; two sweeps. Leaves the best tuning set on the
 * context and copies it to *out (may be NULL). Synchronous: the views' output
 * buffers are overwritten. */
int risltc_autotune(risltc_context_t* context, const risltc_view_t* views, uint32_t count, uint32_t repeats,
                    risltc_tuning_t* out);

/* A batch captured as a hipGraph with buffers of its own. risltc_graph_update
 * takes new views of the same shapes (resolutions, sample counts, shadow-ray
 * and exact_math flags; other fields, pointers and light sets may change)
 * and waits for the graph's previous launch; learned-ness and whether a
 * visibility cache is used must stay too. A graph holds device pointers of
 * the context: after risltc_set_lights, risltc_set_ltc_table, risltc_set_scene,
 * risltc_set_light_clusters (or risltc_cluster_lights), risltc_proposal_cache
 * or a tuning change, risltc_graph_launch fails with RISLTC_E_STATE until
 * the graph is created again. Destroy graphs before their context. Launches
 * of one graph are ordered by their streams only: do not overlap them. */
typedef struct risltc_graph_s risltc_graph_t;
int risltc_graph_create(risltc_context_t* context, const risltc_view_t* views, uint32_t count,
                        risltc_graph_t** out_graph);
int risltc_graph_update(risltc_graph_t* graph, const risltc_view_t* views, uint32_t count);
int risltc_graph_launch(risltc_graph_t* graph, void* stream);
void risltc_graph_destroy(risltc_graph_t* graph);

#ifdef __cplusplus
}
#endif
#endif /* RISLTC_H */
