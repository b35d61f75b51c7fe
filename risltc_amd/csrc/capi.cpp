// capi.cpp -- the C-ABI of include/risltc/risltc.h.
//
// Host side of the drop-in boundary: owns the device copies of lights, LTC
// tables, mesh and LBVH, restates the reference's host code that prepares the
// shading-pass inputs, and launches the shade -> trace -> resolve kernels on
// the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/risltc/risltc.h"
#include "kernels.hpp"

static_assert(sizeof(risltc_frame_constants_t) == 256, "per_frame_constants_t is 256 bytes");
static_assert(sizeof(risltc_light_upload_t) == 48, "polygonal_light_upload_t is 48 bytes");
static_assert(sizeof(risltc_ltc_constants_t) == 32, "ltc_constants_t is 32 bytes");
static_assert(sizeof(risltc_sample_record_t) == 4 * RISLTC_SAMPLE_RECORD_FLOATS, "sample record size");
static_assert(sizeof(risltc_candidate_record_t) == 4 * RISLTC_CANDIDATE_RECORD_WORDS, "candidate record size");
static_assert(rl::kMlpParams == RISLTC_PROPOSAL_PARAMS, "MLP parameter count");
static_assert(rl::mlp_layout(rl::kMlpHidWide).params == RISLTC_PROPOSAL_PARAMS_WIDE, "wide MLP parameter count");
static_assert(rl::kMlpOutWide == RISLTC_MAX_LIGHT_CLUSTERS, "the largest network's outputs = clusters");
static_assert(rl::kCacheClusterSlots == RISLTC_MAX_LIGHT_CLUSTERS, "cache slots in cluster mode");

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, need);
        if (e == hipSuccess) bytes = need;
        return e;
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

struct LightSet {
    DevBuf lights, spheres;
    uint32_t count = 0;
    float bound = 0.0f;
};

// a further mesh with its hierarchy (id = index + 1)
struct SceneSet {
    DevBuf positions, normals_uv, material_idx, materials;
    uint64_t triangle_count = 0;
    uint32_t material_count = 0;
    rl::BvhBuild bvh;
};

// everything a batch of views writes: per-view descriptors, the hierarchy
// table of their scenes, samples, the shared shadow-ray list, hits and
// counters
struct BatchBuffers {
    DevBuf shade_args, tile_offsets, resolve_args, block_offsets, bvh_table;
    DevBuf samples, rays, ray_slots, hits, counter;
    void release() {
        DevBuf* b[] = {&shade_args, &tile_offsets, &resolve_args, &block_offsets, &bvh_table, &samples, &rays,
                       &ray_slots, &hits, &counter};
        for (DevBuf* x : b) x->release();
    }
};

struct risltc_context_s {
    int device = 0;
    std::string error;
    // lights
    DevBuf lights, spheres;
    uint32_t light_count = 0;
    float sphere_bound = 0.0f;
    // LTC table
    DevBuf ltc0, ltc1;
    uint32_t ltc_w = 0, ltc_h = 0, ltc_layers = 0;
    // mesh + LBVH
    DevBuf positions, normals_uv, material_idx, materials;
    uint64_t triangle_count = 0;
    uint32_t material_count = 0;
    rl::BvhBuild bvh;
    // per-frame work buffers
    DevBuf samples, rays, ray_slots, hits, counter;
    uint32_t bvh_builder = RISLTC_BVH_SAH;
    std::vector<float> light_centroids;  // 3 per light (host, for clustering)
    float scene_lo[3] = {0, 0, 0}, scene_hi[3] = {0, 0, 0};
    // learned proposal
    DevBuf cluster_of_light, cluster_offsets, cluster_lights;
    uint32_t cluster_count = 0;
    DevBuf mlp_params, mlp_m, mlp_v, mlp_grad, mlp_wf, mlp_wb, mlp_loss, train_scratch;
    bool mlp_ready = false;
    rl::MlpLayout mlp_L = rl::mlp_layout(rl::kMlpHid);  // hidden width and parameter offsets
    uint32_t adam_step = 0;
    DevBuf prop_table, rec_cluster, rec_y;
    uint32_t prop_width = 0, prop_height = 0;
    size_t rec_samples = 0;  // pixel-samples covered by the training records
    uint32_t rec_spp = 0;
    float prop_ms[2] = {0, 0};  // last inference, last training call
    // visibility cache: 2^vcache_log2 cells x vcache_slots x (uint2 counts,
    // byte), a dirty byte per cell
    DevBuf vcache, vcache_vis, vcache_dirty;
    uint32_t vcache_log2 = 0, vcache_slots = 0, vcache_per_light = 0;
    float vcache_cell = 1.0f, vcache_prior = 2.0f;
    // launch shape (autotuned; RISLTC_TRACE_REFILL overrides the refill at create)
    // (1024 spheres: 16 KB of LDS, so three work-groups fit a CU with the
    // pixel area; more cost a work-group per CU)
    risltc_tuning_t tuning = {16, 32, 1024, 32};
    // extra light sets and meshes for batched views (id = index + 1)
    std::vector<LightSet*> light_sets;
    std::vector<SceneSet*> scene_sets;
    // the context mesh's hierarchy as a one-entry table for the trace kernel
    DevBuf bvh_table;
    // buffers of risltc_shade_batch (graphs own theirs)
    BatchBuffers batch;
    // bumped whenever device data that view descriptors point at, or the
    // tuning they were planned with, changes: graphs of an older generation
    // refuse to launch
    uint64_t generation = 1;
    // timing
    hipEvent_t ev[8] = {};
    bool timed = false;
    // the shadow-ray counter of the last risltc_shade or risltc_shade_noalloc call
    const uint32_t* last_ray_counter = nullptr;
    bool learned_ran = false;
    // this device's copy of the fixed scratch array of risltc_shade_noalloc
    uint8_t* noalloc_scratch = nullptr;
};

static int fail(risltc_context_t* c, int code, const char* fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        c->error = buf;
    }
    return code;
}

#define HIPC(ctx, x)                                                                                 \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) return fail(ctx, RISLTC_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// ---- clip rotation table -----------------------------------------------------
// The reference clips with an autogenerated switch whose output is the cyclic
// sequence [entry point, inside vertices, exit point] rotated so that the
// fewest array slots (including the repeated first vertex) are written; ties
// keep vertex 0 in slot 0, then take the smallest rotation
// (src/shaders/polygon_clipping.glsl:45-216). This computes that rotation for
// every (n, inside mask), n = 3..8.
static int clip_rotation(int n, uint32_t mask) {
    int k = __builtin_popcount(mask);
    if (k == 0 || k == n) return 0;
    int s = 0;
    for (int i = 0; i < n; ++i)
        if (((mask >> i) & 1u) && !((mask >> ((i + n - 1) % n)) & 1u)) {
            s = i;
            break;
        }
    int L = k + 2;
    int cyc[RISLTC_MAX_POLYGON_VERTEX_COUNT];
    cyc[0] = -(((s + n - 1) % n) + 1);
    for (int j = 0; j < k; ++j) cyc[1 + j] = (s + j) % n;
    cyc[k + 1] = -(((s + k - 1) % n) + 1);
    int best = 1 << 30, best_p = 0;
    for (int p = 0; p < L; ++p) {
        int out[RISLTC_MAX_POLYGON_VERTEX_COUNT + 1];
        for (int i = 0; i < L; ++i) out[i] = cyc[(i + p) % L];
        int slots = L;
        if (L < RISLTC_MAX_POLYGON_VERTEX_COUNT) out[slots++] = out[0];
        int cost = 0;
        for (int j = 0; j < slots; ++j)
            if (!(j < n && out[j] == j)) ++cost;
        int key = cost * 1024 + (out[0] == 0 ? 0 : 512) + p;
        if (key < best) {
            best = key;
            best_p = p;
        }
    }
    return best_p;
}

static std::vector<uint8_t> clip_table() {
    std::vector<uint8_t> t(6 * 256, 0);
    for (int n = 3; n <= 8; ++n)
        for (uint32_t m = 0; m < (1u << n); ++m) t[(n - 3) * 256 + m] = (uint8_t)clip_rotation(n, m);
    return t;
}

extern "C" {

void risltc_default_settings(risltc_settings_t* s) {
    if (!s) return;
    memset(s, 0, sizeof(*s));
    s->sample_count = 1;
    s->sample_count_light = 1;
    s->mis_heuristic = RISLTC_MIS_OPTIMAL_CLAMPED;
    s->light_sampling = RISLTC_LIGHT_RESERVOIR;
    s->polygon_sampling_technique = RISLTC_SAMPLE_POLYGON_LTC_CP;
    s->ris_candidate_count = 32;
    s->shadow_rays = 1;
}

// ---- host restatements of the reference's CPU code ---------------------------

// update_polygonal_light (src/polygonal_light.c:44-98), same float operations
void risltc_update_polygonal_light(risltc_polygonal_light_t* light) {
    if (!light || !light->vertices_plane_space || !light->vertices_world_space || light->vertex_count < 3) return;
    light->inv_scaling_x = 1.0f / light->scaling_x;
    light->inv_scaling_y = 1.0f / light->scaling_y;
    const float ax = light->rotation_angles[0], ay = light->rotation_angles[1], az = light->rotation_angles[2];
    float c0 = cosf(ax), s0 = sinf(ax), c1 = cosf(ay), s1 = sinf(ay), c2 = cosf(az), s2 = sinf(az);
    float c0s1 = c0 * s1, s0s1 = s0 * s1;
    float rot[3][4] = {{c1 * c2, -c1 * s2, -s1, 0.0f},
                       {-s0s1 * c2 + c0 * s2, s0s1 * s2 + c0 * c2, -s0 * c1, 0.0f},
                       {c0s1 * c2 + s0 * s2, -c0s1 * s2 + s0 * c2, c0 * c1, 0.0f}};
    memcpy(light->rotation, rot, sizeof(rot));
    const float sc[2] = {light->scaling_x, light->scaling_y};
    const float* ps = light->vertices_plane_space;
    for (uint32_t v = 0; v < light->vertex_count; ++v)
        for (int r = 0; r < 3; ++r) {
            float acc = light->translation[r];
            for (int k = 0; k < 2; ++k) acc += sc[k] * rot[r][k] * ps[v * 4 + k];
            light->vertices_world_space[v * 4 + r] = acc;
        }
    for (int r = 0; r < 3; ++r) light->plane[r] = rot[r][2];
    light->plane[3] = -(rot[0][2] * light->translation[0] + rot[1][2] * light->translation[1] +
                        rot[2][2] * light->translation[2]);
    float area = 0.0f;
    for (uint32_t t = 0; t != light->vertex_count - 2u; ++t) {
        float a = ps[(t + 2) * 4 + 0] - ps[0], b = ps[(t + 1) * 4 + 0] - ps[0];
        float c = ps[(t + 2) * 4 + 1] - ps[1], d = ps[(t + 1) * 4 + 1] - ps[1];
        area += 0.5f * (a * d - b * c);
    }
    area *= sc[0] * sc[1];
    float aa = (area < 0.0f) ? -area : area;
    light->area = aa;
    light->rcp_area = 1.0f / aa;
    for (int i = 0; i < 3; ++i) light->surface_radiance[i] = light->radiant_flux[i];
    for (int i = 0; i < 4; ++i) light->plane[i] = (area > 0.0f) ? light->plane[i] : -light->plane[i];
}

size_t risltc_write_lights(risltc_polygonal_light_t* lights, uint32_t count, uint32_t vmax, void* out,
                           size_t out_size) {
    if (!lights || !out || vmax < 3 || vmax > RISLTC_MAX_LIGHT_VERTEX_COUNT) return 0;
    size_t stride = RISLTC_LIGHT_RECORD_SIZE(vmax);
    if (out_size < stride * count) return 0;
    auto* dst = static_cast<uint8_t*>(out);
    for (uint32_t i = 0; i < count; ++i) {
        risltc_polygonal_light_t* L = &lights[i];
        if (L->vertex_count < 3 || L->vertex_count > vmax) return 0;
        risltc_update_polygonal_light(L);
        risltc_light_upload_t up;
        memset(&up, 0, sizeof(up));
        for (int k = 0; k < 3; ++k) up.surface_radiance[k] = L->surface_radiance[k];
        for (int k = 0; k < 4; ++k) up.plane[k] = L->plane[k];
        up.vertex_count = L->vertex_count;
        uint8_t* rec = dst + stride * i;
        memcpy(rec, &up, sizeof(up));
        for (uint32_t v = 0; v < vmax; ++v) {
            uint32_t src = v < L->vertex_count ? v : 0u;
            float q[4] = {L->vertices_world_space[4 * src], L->vertices_world_space[4 * src + 1],
                          L->vertices_world_space[4 * src + 2], 0.0f};
            memcpy(rec + 48 + 16 * v, q, 16);
        }
    }
    return stride * count;
}

// matrix_inverse (src/math_utilities.h:24-47): cofactors, then times 1/det
static void matrix_inverse(float inv[4][4], const float mat[4][4]) {
    const float* m = &mat[0][0];
    float* o = &inv[0][0];
    o[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
           m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    o[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    o[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
           m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    o[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    o[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
           m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    o[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
           m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    o[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
           m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    o[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
            m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    o[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
           m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    o[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
           m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    o[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
            m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    o[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
            m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    o[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
           m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    o[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
           m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    o[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
            m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    o[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
            m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    float det = m[0] * o[0] + m[1] * o[4] + m[2] * o[8] + m[3] * o[12];
    float rcp = 1.0f / det;
    for (int i = 0; i < 16; ++i) o[i] = o[i] * rcp;
}

static uint32_t wang(uint32_t seed) {
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed;
}

void risltc_write_frame_constants(risltc_frame_constants_t* out, const risltc_camera_t* cam, uint32_t width,
                                  uint32_t height, const float deq_factor[3], const float deq_summand[3],
                                  const risltc_ltc_constants_t* ltc, float exposure, float roughness_factor,
                                  float mis_vis, uint32_t noise_seed, int animate) {
    if (!out || !cam || width == 0 || height == 0) return;
    risltc_frame_constants_t c;
    memset(&c, 0, sizeof(c));
    for (int i = 0; i < 3; ++i) {
        c.mesh_dequantization_factor[i] = deq_factor ? deq_factor[i] : 1.0f;
        c.mesh_dequantization_summand[i] = deq_summand ? deq_summand[i] : 0.0f;
        c.camera_position_world_space[i] = cam->position_world_space[i];
    }
    c.mis_visibility_estimate = mis_vis;
    c.viewport_size[0] = width;
    c.viewport_size[1] = height;
    if (ltc) c.ltc_constants = *ltc;
    c.error_factor = powf(10.0f, 7.0f);
    c.exposure_factor = exposure;
    c.roughness_factor = roughness_factor;
    for (uint32_t i = 0; i < 4; ++i) c.noise_random_numbers[i] = animate ? wang(noise_seed * 4u + i) : i * 0x123456u;
    // get_world_to_projection_space (src/camera.c:24-83)
    float cx = cosf(cam->rotation_x), sx = sinf(cam->rotation_x);
    float cz = cosf(cam->rotation_z), sz = sinf(cam->rotation_z);
    float rx[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, cx, sx}, {0.0f, -sx, cx}};
    float rz[3][3] = {{cz, sz, 0.0f}, {-sz, cz, 0.0f}, {0.0f, 0.0f, 1.0f}};
    float rot[3][3] = {{0.0f}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int l = 0; l < 3; ++l) rot[i][j] += rz[i][l] * rx[l][j];
    float org[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) org[i] -= rot[j][i] * cam->position_world_space[j];
    float view[4][4] = {{rot[0][0], rot[1][0], rot[2][0], org[0]},
                        {rot[0][1], rot[1][1], rot[2][1], org[1]},
                        {rot[0][2], rot[1][2], rot[2][2], org[2]},
                        {0.0f, 0.0f, 0.0f, 1.0f}};
    float n = cam->near_plane, f = cam->far_plane;
    float top = tanf(0.5f * cam->vertical_fov);
    float right = (((float)width) / ((float)height)) * top;
    float proj[4][4] = {{-1.0f / right, 0.0f, 0.0f, 0.0f},
                        {0.0f, 1.0f / top, 0.0f, 0.0f},
                        {0.0f, 0.0f, -(f + n) / (f - n), -2.0f * f * n / (f - n)},
                        {0.0f, 0.0f, -1.0f, 0.0f}};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int l = 0; l < 4; ++l) c.world_to_projection_space[i][j] += proj[i][l] * view[l][j];
    // pixel -> ray direction (src/main.c:2921-2944)
    float vp[4];
    vp[0] = 2.0f / width;
    vp[1] = 2.0f / height;
    vp[2] = 0.5f * vp[0] - 1.0f;
    vp[3] = 0.5f * vp[1] - 1.0f;
    float w2p[4][4], p2w[4][4];
    memcpy(w2p, c.world_to_projection_space, sizeof(w2p));
    w2p[0][3] = w2p[1][3] = w2p[2][3] = 0.0f;
    matrix_inverse(p2w, w2p);
    float pix[4][3] = {{vp[0], 0.0f, vp[2]}, {0.0f, vp[1], vp[3]}, {0.0f, 0.0f, 1.0f}, {0.0f, 0.0f, 1.0f}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 4; ++k) c.pixel_to_ray_direction_world_space[i][j] += p2w[i][k] * pix[k][j];
    *out = c;
}

void risltc_ltc_constants_for(risltc_ltc_constants_t* o, uint32_t rr, uint32_t ri, uint32_t fc) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->fresnel_index_factor = (float)(fc - 1);
    o->fresnel_index_summand = 0.0f;
    o->roughness_factor = (float)(rr - 1) / (float)rr;
    o->roughness_summand = 0.5f / (float)rr;
    o->inclination_factor = (float)(ri - 1) / (0.5f * 3.1415926535897932384626433832795f * ri);
    o->inclination_summand = 0.5f / (float)ri;
}

// load_ltc_table file handling and quantisation (src/ltc_table.c:31-116)
int risltc_load_ltc_fit(const char* dir, uint32_t fresnel_count, uint32_t* out_res, uint16_t* rgba16, uint16_t* rg16,
                        size_t capacity) {
    if (!dir || fresnel_count == 0) return RISLTC_E_ARG;
    uint64_t res = 0;
    for (uint32_t i = 0; i < fresnel_count; ++i) {
        char path[4096];
        snprintf(path, sizeof(path), "%s/fit%u.dat", dir, i);
        FILE* fh = fopen(path, "rb");
        if (!fh) {
            fprintf(stderr, "risltc_load_ltc_fit: cannot open %s\n", path);
            return RISLTC_E_IO;
        }
        uint64_t r = 0;
        if (fread(&r, sizeof(r), 1, fh) != 1 || r == 0 || r > 4096) {
            fclose(fh);
            return RISLTC_E_FORMAT;
        }
        if (i == 0) {
            res = r;
            if (out_res) *out_res = (uint32_t)res;
            if (!rgba16 && !rg16) {
                fclose(fh);
                return RISLTC_OK;
            }
            if (!rgba16 || !rg16 || capacity < res * res * 4 * fresnel_count) {
                fclose(fh);
                return RISLTC_E_ARG;
            }
        } else if (r != res) {
            fprintf(stderr, "risltc_load_ltc_fit: the fit files in %s differ in resolution\n", dir);
            fclose(fh);
            return RISLTC_E_FORMAT;
        }
        for (uint64_t j = 0; j < res * res; ++j) {
            float d[5];
            if (fread(d, sizeof(float), 5, fh) != 5) {
                fclose(fh);
                return RISLTC_E_FORMAT;
            }
            float adj[3][3] = {{d[2], 0.0f, -d[1] * d[2]}, {0.0f, d[0] - d[1] * d[3], 0.0f}, {-d[2] * d[3], 0.0f, d[0] * d[2]}};
            float big = fabsf(adj[0][0]);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    if (big < fabsf(adj[a][b])) big = fabsf(adj[a][b]);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) adj[a][b] /= big;
            float vals[6] = {adj[0][0], -adj[0][2], adj[1][1], adj[2][0], adj[2][2], d[4]};
            uint16_t q[6];
            for (int a = 0; a < 6; ++a) {
                float x = vals[a];
                if (x < 0.0f) x = 0.0f;
                if (x > 1.0f) x = 1.0f;
                q[a] = (uint16_t)(x * 65535.0f + 0.5f);
            }
            size_t base = (size_t)i * res * res + j;
            for (int a = 0; a < 4; ++a) rgba16[base * 4 + a] = q[a];
            rg16[base * 2 + 0] = q[4];
            rg16[base * 2 + 1] = q[5];
        }
        fclose(fh);
    }
    return RISLTC_OK;
}

// ---- context -----------------------------------------------------------------------
// ---- the fixed scratch of risltc_shade_noalloc --------------------------------------
// One file-scope device array holds what risltc_shade allocates per call (sample
// records, the compacted shadow-ray list, ray slots, hit words, the ray and
// trace-queue counters) for at most RISLTC_NOALLOC_MAX_SAMPLES sample records
// (pixels x spp x SAMPLE_COUNT):
// about 72 MiB, held by every GPU that loads this library until the process
// ends. Each region starts on a 256-byte boundary.
static const uint32_t kCounterWords = rl::kQueueStride * (1u + rl::kTraceQueues);
namespace noalloc {
constexpr size_t align256(size_t x) { return (x + 255u) & ~size_t(255u); }
constexpr size_t kSamples = RISLTC_NOALLOC_MAX_SAMPLES;
constexpr size_t kRecordsOff = 0;
constexpr size_t kRaysOff = align256(kRecordsOff + kSamples * rl::kSampleWords * sizeof(float));
constexpr size_t kSlotsOff = align256(kRaysOff + kSamples * 2 * sizeof(rl::ShadowRay));
constexpr size_t kHitsOff = align256(kSlotsOff + kSamples * 2 * sizeof(uint32_t));
constexpr size_t kCounterOff = align256(kHitsOff + kSamples * 2 * sizeof(uint32_t));
constexpr size_t kBytes = align256(kCounterOff + kCounterWords * sizeof(uint32_t));
}  // namespace noalloc
static __device__ __align__(256) uint8_t scratch[noalloc::kBytes];

// The array's address on each device, resolved once per device with that
// device current (the host-side name `scratch` is only a handle for the
// runtime and is never passed to a kernel).
static std::mutex g_scratch_mutex;
static uint8_t* g_scratch_of_device[256] = {};

static hipError_t resolve_scratch(int device, uint8_t** out) {
    if (device < 0 || device >= 256) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_of_device[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_of_device[device] = static_cast<uint8_t*>(p);
    }
    *out = g_scratch_of_device[device];
    return hipSuccess;
}

int risltc_create(risltc_context_t** out, int device) {
    if (!out) return RISLTC_E_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return RISLTC_E_HIP;
    if (device < 0 || device >= count) return RISLTC_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return RISLTC_E_HIP;
    auto* c = new risltc_context_t();
    c->device = device;
    std::vector<uint8_t> t = clip_table();
    if (rl::upload_clip_table(t.data(), t.size()) != hipSuccess) {
        delete c;
        return RISLTC_E_HIP;
    }
    for (auto& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return RISLTC_E_HIP;
        }
    if (resolve_scratch(device, &c->noalloc_scratch) != hipSuccess) {
        for (auto& e : c->ev)
            if (e) hipEventDestroy(e);
        delete c;
        return RISLTC_E_HIP;
    }
    if (const char* v = getenv("RISLTC_TRACE_REFILL")) {
        long r = strtol(v, nullptr, 10);
        if (r >= 1 && r <= 64) c->tuning.trace_refill = (uint32_t)r;
    }
    *out = c;
    return RISLTC_OK;
}

void risltc_destroy(risltc_context_t* c) {
    if (!c) return;
    hipSetDevice(c->device);
    for (LightSet* ls : c->light_sets) {
        ls->lights.release();
        ls->spheres.release();
        delete ls;
    }
    c->light_sets.clear();
    for (SceneSet* ss : c->scene_sets) {
        DevBuf* b[] = {&ss->positions, &ss->normals_uv, &ss->material_idx, &ss->materials};
        for (DevBuf* x : b) x->release();
        rl::free_bvh(&ss->bvh);
        delete ss;
    }
    c->scene_sets.clear();
    c->bvh_table.release();
    c->batch.release();
    DevBuf* learned_bufs[] = {&c->cluster_of_light, &c->cluster_offsets, &c->cluster_lights, &c->mlp_params,
                              &c->mlp_m, &c->mlp_v, &c->mlp_grad, &c->mlp_wf, &c->mlp_wb, &c->mlp_loss,
                              &c->train_scratch, &c->prop_table, &c->rec_cluster, &c->rec_y, &c->vcache, &c->vcache_vis,
                              &c->vcache_dirty};
    for (DevBuf* b : learned_bufs) b->release();
    DevBuf* bufs[] = {&c->lights, &c->spheres, &c->ltc0, &c->ltc1, &c->positions, &c->normals_uv, &c->material_idx,
                      &c->materials, &c->samples, &c->rays, &c->ray_slots, &c->hits, &c->counter};
    for (DevBuf* b : bufs) b->release();
    rl::free_bvh(&c->bvh);
    for (auto& e : c->ev)
        if (e) hipEventDestroy(e);
    delete c;
}

const char* risltc_last_error(const risltc_context_t* c) { return c ? c->error.c_str() : "null context"; }

static float round_up_float(double x) {
    if (!(x < 3.0e38)) return INFINITY;  // also NaN
    float f = (float)x;
    return ((double)f < x) ? nextafterf(f, INFINITY) : f;
}

// Bounding sphere of a light's vertices (xyz triples, as the kernels read
// them): centre = mean, radius = largest distance from the rounded centre,
// rounded up. Non-finite vertices give an infinite radius (never culled).
// Returns |centre| + radius.
static double bounding_sphere(const float* v, uint32_t n, float out[4]) {
    double c[3] = {0.0, 0.0, 0.0};
    for (uint32_t k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) c[a] += v[3 * k + a];
    float cf[3];
    for (int a = 0; a < 3; ++a) cf[a] = (float)(c[a] / n);
    double r2 = 0.0;
    for (uint32_t k = 0; k < n; ++k) {
        double s = 0.0;
        for (int a = 0; a < 3; ++a) {
            double t = (double)v[3 * k + a] - (double)cf[a];
            s += t * t;
        }
        r2 = std::max(r2, s);
    }
    double r = std::sqrt(r2) * (1.0 + 1e-6);
    bool finite = std::isfinite(r) && std::isfinite(cf[0]) && std::isfinite(cf[1]) && std::isfinite(cf[2]);
    if (!finite) {
        out[0] = out[1] = out[2] = 0.0f;
        out[3] = INFINITY;
        return INFINITY;
    }
    for (int a = 0; a < 3; ++a) out[a] = cf[a];
    out[3] = round_up_float(r);
    return std::sqrt((double)cf[0] * cf[0] + (double)cf[1] * cf[1] + (double)cf[2] * cf[2]) + (double)out[3];
}

struct PackedLights {
    std::vector<float> packed, spheres, centroids;
    double bound = 0.0;
};

// Repacks write_lights records (stride 48 + 16 vmax) into the kernels' 8
// float4 per light, with bounding spheres and centroids.
static int pack_lights(risltc_context_t* c, const void* records, uint32_t count, uint32_t vmax, PackedLights& out) {
    if (!records || count == 0) return fail(c, RISLTC_E_ARG, "no light records");
    if (vmax < 3 || vmax > RISLTC_MAX_LIGHT_VERTEX_COUNT)
        return fail(c, RISLTC_E_UNSUPPORTED, "max_vertex_count %u outside [3, 8]", vmax);
    const size_t stride = RISLTC_LIGHT_RECORD_SIZE(vmax);
    out.packed.assign((size_t)count * rl::kLightFloat4 * 4, 0.0f);
    out.spheres.assign((size_t)count * 4, 0.0f);
    out.centroids.assign((size_t)count * 3, 0.0f);
    out.bound = 0.0;
    const auto* src = static_cast<const uint8_t*>(records);
    for (uint32_t i = 0; i < count; ++i) {
        const uint8_t* rec = src + stride * i;
        risltc_light_upload_t up;
        memcpy(&up, rec, sizeof(up));
        if (up.vertex_count < 3 || up.vertex_count > vmax)
            return fail(c, RISLTC_E_FORMAT, "light %u has vertex_count %u outside [3, %u]", i, up.vertex_count, vmax);
        float* d = out.packed.data() + (size_t)i * rl::kLightFloat4 * 4;
        d[0] = up.surface_radiance[0];
        d[1] = up.surface_radiance[1];
        d[2] = up.surface_radiance[2];
        memcpy(&d[3], &up.vertex_count, 4);
        for (int k = 0; k < 4; ++k) d[4 + k] = up.plane[k];
        for (uint32_t v = 0; v < RISLTC_MAX_LIGHT_VERTEX_COUNT; ++v) {
            uint32_t sv = v < up.vertex_count ? v : 0u;
            float q[4];
            memcpy(q, rec + 48 + 16 * sv, 16);
            for (int a = 0; a < 3; ++a) d[8 + 3 * v + a] = q[a];
        }
        out.bound = std::max(out.bound, bounding_sphere(d + 8, up.vertex_count, &out.spheres[(size_t)i * 4]));
        for (int a = 0; a < 3; ++a) {
            double m = 0.0;
            for (uint32_t v = 0; v < up.vertex_count; ++v) m += d[8 + 3 * v + a];
            out.centroids[(size_t)i * 3 + a] = (float)(m / up.vertex_count);
        }
    }
    return RISLTC_OK;
}

static size_t vcache_words(const risltc_context_t* c) {
    return c->vcache.p ? ((size_t)2 * c->vcache_slots) << c->vcache_log2 : 0;
}

// empties the visibility cache (its cells or slots changed meaning): zero
// counts, bytes 255 (empty), no dirty cell
static int clear_vcache(risltc_context_t* c) {
    if (!c->vcache.p) return RISLTC_OK;
    HIPC(c, hipSetDevice(c->device));
    const size_t entries = vcache_words(c) / 2, cells = (size_t)1 << c->vcache_log2;
    HIPC(c, rl::launch_zero_u32(c->vcache.as<uint32_t>(), (uint32_t)vcache_words(c), nullptr));
    HIPC(c, rl::launch_fill_u32(c->vcache_vis.as<uint32_t>(), (uint32_t)((entries + 3) / 4), 0xFFFFFFFFu, nullptr));
    HIPC(c, rl::launch_zero_u32(c->vcache_dirty.as<uint32_t>(), (uint32_t)((cells + 3) / 4), nullptr));
    HIPC(c, hipDeviceSynchronize());
    return RISLTC_OK;
}

static rl::VisCacheView vcache_view(const risltc_context_t* c) {
    rl::VisCacheView v;
    v.counts = c->vcache.as<uint2>();
    v.vis = c->vcache_vis.as<uint8_t>();
    v.dirty = c->vcache_dirty.as<uint8_t>();
    v.mask = (1u << c->vcache_log2) - 1u;
    v.inv_cell = 1.0f / c->vcache_cell;
    v.prior = c->vcache_prior;
    v.slots = c->vcache_slots;
    v.per_light = c->vcache_per_light;
    return v;
}

int risltc_set_lights(risltc_context_t* c, const void* records, uint32_t count, uint32_t vmax) {
    if (!c) return RISLTC_E_ARG;
    PackedLights pl;
    int e = pack_lights(c, records, count, vmax, pl);
    if (e) return e;
    ++c->generation;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, c->lights.ensure(pl.packed.size() * sizeof(float)));
    HIPC(c, hipMemcpy(c->lights.p, pl.packed.data(), pl.packed.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPC(c, c->spheres.ensure(pl.spheres.size() * sizeof(float)));
    HIPC(c, hipMemcpy(c->spheres.p, pl.spheres.data(), pl.spheres.size() * sizeof(float), hipMemcpyHostToDevice));
    c->light_count = count;
    c->sphere_bound = round_up_float(pl.bound);
    c->light_centroids = std::move(pl.centroids);
    // clusters, and tables that name them, refer to the previous light set
    c->cluster_count = 0;
    c->prop_width = c->prop_height = 0;
    return clear_vcache(c);
}

int risltc_set_polygonal_lights(risltc_context_t* c, risltc_polygonal_light_t* lights, uint32_t count,
                                uint32_t vmax) {
    if (!c) return RISLTC_E_ARG;
    if (!lights || count == 0) return fail(c, RISLTC_E_ARG, "no lights");
    if (vmax == 0)
        for (uint32_t i = 0; i < count; ++i) vmax = std::max(vmax, lights[i].vertex_count);
    if (vmax < 3 || vmax > RISLTC_MAX_LIGHT_VERTEX_COUNT)
        return fail(c, RISLTC_E_UNSUPPORTED, "max_vertex_count %u outside [3, %d]", vmax,
                    RISLTC_MAX_LIGHT_VERTEX_COUNT);
    for (uint32_t i = 0; i < count; ++i)
        if (!lights[i].vertices_plane_space || !lights[i].vertices_world_space)
            return fail(c, RISLTC_E_ARG, "light %u: vertex arrays not allocated", i);
    std::vector<uint8_t> records((size_t)count * RISLTC_LIGHT_RECORD_SIZE(vmax));
    if (risltc_write_lights(lights, count, vmax, records.data(), records.size()) != records.size())
        return fail(c, RISLTC_E_FORMAT, "a light has a vertex count outside [3, %u]", vmax);
    return risltc_set_lights(c, records.data(), count, vmax);
}

int risltc_load_ltc_table(risltc_context_t* c, const char* dir, uint32_t fresnel_count) {
    if (!c) return RISLTC_E_ARG;
    uint32_t res = 0;
    int e = risltc_load_ltc_fit(dir, fresnel_count, &res, nullptr, nullptr, 0);
    if (e) return fail(c, e, "cannot read the LTC table in %s", dir ? dir : "(null)");
    const size_t texels = (size_t)res * res * fresnel_count;
    std::vector<uint16_t> rgba(texels * 4), rg(texels * 2);
    e = risltc_load_ltc_fit(dir, fresnel_count, &res, rgba.data(), rg.data(), rgba.size());
    if (e) return fail(c, e, "cannot read the LTC table in %s", dir);
    return risltc_set_ltc_table(c, rgba.data(), rg.data(), res, res, fresnel_count);
}

int risltc_add_light_set(risltc_context_t* c, const void* records, uint32_t count, uint32_t vmax, uint32_t* id) {
    if (!c) return RISLTC_E_ARG;
    if (!id) return fail(c, RISLTC_E_ARG, "null id");
    PackedLights pl;
    int e = pack_lights(c, records, count, vmax, pl);
    if (e) return e;
    HIPC(c, hipSetDevice(c->device));
    auto* ls = new LightSet();
    if (ls->lights.ensure(pl.packed.size() * sizeof(float)) != hipSuccess ||
        ls->spheres.ensure(pl.spheres.size() * sizeof(float)) != hipSuccess ||
        hipMemcpy(ls->lights.p, pl.packed.data(), pl.packed.size() * sizeof(float), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(ls->spheres.p, pl.spheres.data(), pl.spheres.size() * sizeof(float), hipMemcpyHostToDevice) !=
            hipSuccess) {
        ls->lights.release();
        ls->spheres.release();
        delete ls;
        return fail(c, RISLTC_E_HIP, "light set upload failed");
    }
    ls->count = count;
    ls->bound = round_up_float(pl.bound);
    c->light_sets.push_back(ls);
    *id = (uint32_t)c->light_sets.size();
    return RISLTC_OK;
}

int risltc_set_ltc_table(risltc_context_t* c, const uint16_t* rgba16, const uint16_t* rg16, uint32_t rr, uint32_t ri,
                         uint32_t fc) {
    if (!c) return RISLTC_E_ARG;
    if (!rgba16 || !rg16 || rr < 2 || ri < 2 || fc < 1) return fail(c, RISLTC_E_ARG, "bad LTC table arguments");
    size_t texels = (size_t)rr * ri * fc;
    ++c->generation;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, c->ltc0.ensure(texels * 4 * sizeof(uint16_t)));
    HIPC(c, c->ltc1.ensure(texels * 2 * sizeof(uint16_t)));
    HIPC(c, hipMemcpy(c->ltc0.p, rgba16, texels * 4 * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->ltc1.p, rg16, texels * 2 * sizeof(uint16_t), hipMemcpyHostToDevice));
    c->ltc_w = rr;
    c->ltc_h = ri;
    c->ltc_layers = fc;
    return RISLTC_OK;
}

// ---- learned proposal --------------------------------------------------------------
int risltc_set_light_clusters(risltc_context_t* c, const uint32_t* cluster_of_light, uint32_t cluster_count) {
    if (!c) return RISLTC_E_ARG;
    if (!cluster_of_light || cluster_count == 0) return fail(c, RISLTC_E_ARG, "null clusters");
    if (cluster_count > RISLTC_MAX_LIGHT_CLUSTERS)
        return fail(c, RISLTC_E_UNSUPPORTED, "at most %d clusters", RISLTC_MAX_LIGHT_CLUSTERS);
    // graphs of learned batches hold the cluster tables' device pointers
    ++c->generation;
    const uint32_t n = c->light_count;
    if (n == 0) return fail(c, RISLTC_E_STATE, "risltc_set_lights has not been called");
    std::vector<uint32_t> offsets(cluster_count + 1, 0), sorted(n);
    std::vector<uint16_t> of(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (cluster_of_light[i] >= cluster_count)
            return fail(c, RISLTC_E_ARG, "light %u in cluster %u of %u", i, cluster_of_light[i], cluster_count);
        of[i] = (uint16_t)cluster_of_light[i];
        offsets[cluster_of_light[i] + 1]++;
    }
    for (uint32_t k = 0; k < cluster_count; ++k) offsets[k + 1] += offsets[k];
    std::vector<uint32_t> fill(offsets.begin(), offsets.end() - 1);
    for (uint32_t i = 0; i < n; ++i) sorted[fill[cluster_of_light[i]]++] = i;  // stable by light index
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, c->cluster_of_light.ensure(n * sizeof(uint16_t)));
    HIPC(c, c->cluster_offsets.ensure((cluster_count + 1) * sizeof(uint32_t)));
    HIPC(c, c->cluster_lights.ensure(n * sizeof(uint32_t)));
    HIPC(c, hipMemcpy(c->cluster_of_light.p, of.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->cluster_offsets.p, offsets.data(), offsets.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->cluster_lights.p, sorted.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    c->cluster_count = cluster_count;
    c->prop_width = c->prop_height = 0;  // tables that name the previous clusters
    return clear_vcache(c);
}

int risltc_proposal_cache(risltc_context_t* c, uint32_t log2_entries, float cell_size, float prior,
                          uint32_t per_light) {
    if (!c) return RISLTC_E_ARG;
    // graphs of learned batches hold the cache's device pointers
    ++c->generation;
    if (log2_entries == 0) {
        c->vcache.release();
        c->vcache_vis.release();
        c->vcache_dirty.release();
        c->vcache_log2 = 0;
        return RISLTC_OK;
    }
    if (log2_entries > 20) return fail(c, RISLTC_E_UNSUPPORTED, "at most 2^20 cache cells");
    if (!(cell_size > 0.0f) || !std::isfinite(1.0f / cell_size))
        return fail(c, RISLTC_E_ARG, "cell size %g", (double)cell_size);
    if (!(prior > 0.0f) || !std::isfinite(prior)) return fail(c, RISLTC_E_ARG, "prior strength %g", (double)prior);
    if (per_light > 1) return fail(c, RISLTC_E_ARG, "per_light %u", per_light);
    if (per_light && c->light_count == 0) return fail(c, RISLTC_E_STATE, "risltc_set_lights has not been called");
    const uint32_t slots = per_light ? c->light_count : (uint32_t)rl::kCacheClusterSlots;
    const size_t words = ((size_t)2 * slots) << log2_entries;
    if (words > ((size_t)1 << 30)) return fail(c, RISLTC_E_UNSUPPORTED, "cache larger than 4 GiB");
    HIPC(c, hipSetDevice(c->device));
    c->vcache_log2 = 0;
    HIPC(c, c->vcache.ensure(words * sizeof(uint32_t)));
    HIPC(c, c->vcache_vis.ensure((words / 2 + 3) / 4 * 4));
    HIPC(c, c->vcache_dirty.ensure((((size_t)1 << log2_entries) + 3) / 4 * 4));
    c->vcache_log2 = log2_entries;
    c->vcache_slots = slots;
    c->vcache_per_light = per_light;
    c->vcache_cell = cell_size;
    c->vcache_prior = prior;
    return clear_vcache(c);
}

int risltc_proposal_cache_read(risltc_context_t* c, uint32_t* out, uint64_t capacity) {
    if (!c) return RISLTC_E_ARG;
    if (!c->vcache.p) return fail(c, RISLTC_E_STATE, "no visibility cache");
    if (!out || capacity < vcache_words(c)) return fail(c, RISLTC_E_ARG, "output smaller than the cache");
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipMemcpy(out, c->vcache.p, vcache_words(c) * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return RISLTC_OK;
}

int risltc_proposal_cache_write(risltc_context_t* c, const uint32_t* counts, uint64_t words) {
    if (!c) return RISLTC_E_ARG;
    if (!c->vcache.p) return fail(c, RISLTC_E_STATE, "no visibility cache");
    if (!counts || words != vcache_words(c)) return fail(c, RISLTC_E_ARG, "expected %zu words", vcache_words(c));
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipMemcpy(c->vcache.p, counts, words * sizeof(uint32_t), hipMemcpyHostToDevice));
    // every cell is recomputed by the next shading call
    const size_t cells = (size_t)1 << c->vcache_log2;
    HIPC(c, rl::launch_fill_u32(c->vcache_dirty.as<uint32_t>(), (uint32_t)((cells + 3) / 4), 0x01010101u, nullptr));
    HIPC(c, hipDeviceSynchronize());
    return RISLTC_OK;
}

int risltc_proposal_cache_table(risltc_context_t* c, uint8_t* out, uint64_t capacity) {
    if (!c) return RISLTC_E_ARG;
    if (!c->vcache.p) return fail(c, RISLTC_E_STATE, "no visibility cache");
    const size_t entries = vcache_words(c) / 2;
    if (!out || capacity < entries) return fail(c, RISLTC_E_ARG, "output smaller than the cache");
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipMemcpy(out, c->vcache_vis.p, entries, hipMemcpyDeviceToHost));
    return RISLTC_OK;
}

// Deterministic Lloyd k-means of the points p[3 i] of the listed members into
// K clusters: centres start at the members k * m / K, every iteration
// assigns each member to its nearest centre (the first on ties) and moves
// each centre to its members' mean; an empty cluster restarts at the member
// farthest from its centre. Returns the cluster of each member.
static std::vector<uint32_t> lloyd_kmeans(const float* p, const std::vector<uint32_t>& members, uint32_t K,
                                          uint32_t iterations) {
    const uint32_t n = (uint32_t)members.size();
    std::vector<double> ctr(3 * K);
    for (uint32_t k = 0; k < K; ++k) {
        uint32_t i = members[(uint32_t)(((uint64_t)k * n) / K)];
        for (int a = 0; a < 3; ++a) ctr[3 * k + a] = p[3 * i + a];
    }
    std::vector<uint32_t> asg(n, 0);
    auto d2 = [&](uint32_t m, uint32_t k) {
        const uint32_t i = members[m];
        double s = 0.0;
        for (int a = 0; a < 3; ++a) {
            double t = (double)p[3 * i + a] - ctr[3 * k + a];
            s += t * t;
        }
        return s;
    };
    for (uint32_t it = 0; it <= iterations; ++it) {
        for (uint32_t m = 0; m < n; ++m) {
            uint32_t best = 0;
            double bd = d2(m, 0);
            for (uint32_t k = 1; k < K; ++k) {
                double d = d2(m, k);
                if (d < bd) {
                    bd = d;
                    best = k;
                }
            }
            asg[m] = best;
        }
        if (it == iterations) break;
        std::vector<double> sum(3 * K, 0.0);
        std::vector<uint32_t> cnt(K, 0);
        for (uint32_t m = 0; m < n; ++m) {
            cnt[asg[m]]++;
            for (int a = 0; a < 3; ++a) sum[3 * asg[m] + a] += p[3 * members[m] + a];
        }
        for (uint32_t k = 0; k < K; ++k) {
            if (cnt[k] > 0) {
                for (int a = 0; a < 3; ++a) ctr[3 * k + a] = sum[3 * k + a] / cnt[k];
                continue;
            }
            uint32_t far = 0;
            double fd = -1.0;
            for (uint32_t m = 0; m < n; ++m) {
                double d = d2(m, asg[m]);
                if (d > fd) {
                    fd = d;
                    far = m;
                }
            }
            for (int a = 0; a < 3; ++a) ctr[3 * k + a] = p[3 * members[far] + a];
            asg[far] = k;
        }
    }
    return asg;
}

int risltc_cluster_lights(risltc_context_t* c, uint32_t cluster_count, uint32_t iterations, uint32_t* out) {
    if (!c) return RISLTC_E_ARG;
    const uint32_t n = c->light_count;
    if (n == 0) return fail(c, RISLTC_E_STATE, "risltc_set_lights has not been called");
    if (cluster_count == 0 || cluster_count > RISLTC_MAX_LIGHT_CLUSTERS || cluster_count > n)
        return fail(c, RISLTC_E_ARG, "cluster_count %u (lights %u, at most %d)", cluster_count, n,
                    RISLTC_MAX_LIGHT_CLUSTERS);
    const float* p = c->light_centroids.data();
    const uint32_t K = cluster_count;
    std::vector<uint32_t> all(n);
    for (uint32_t i = 0; i < n; ++i) all[i] = i;
    std::vector<uint32_t> asg;
    if (K <= 256) {
        asg = lloyd_kmeans(p, all, K, iterations);
    } else {
        // Two levels: ceil(K / 4) clusters of all lights, then each of those
        // split into (up to) four, numbered 4 q .. 4 q + 3: a quad of
        // clusters is one top-level cluster, which the proposal tables of
        // networks with more than 256 outputs sample as a group.
        const uint32_t Q = (K + 3u) / 4u;
        std::vector<uint32_t> top = lloyd_kmeans(p, all, Q, iterations);
        std::vector<std::vector<uint32_t>> members(Q);
        for (uint32_t i = 0; i < n; ++i) members[top[i]].push_back(i);
        asg.assign(n, 0);
        for (uint32_t q = 0; q < Q; ++q) {
            const uint32_t m = (uint32_t)members[q].size();
            uint32_t sub = std::min(4u, K - 4u * q);
            if (sub > m) sub = m;
            if (m == 0) continue;
            std::vector<uint32_t> a = lloyd_kmeans(p, members[q], sub, iterations);
            for (uint32_t j = 0; j < m; ++j) asg[members[q][j]] = 4u * q + a[j];
        }
    }
    if (out) memcpy(out, asg.data(), n * sizeof(uint32_t));
    return risltc_set_light_clusters(c, asg.data(), K);
}

static uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

static int mlp_alloc(risltc_context_t* c) {
    const size_t n = (size_t)c->mlp_L.params;
    const size_t pb = n * sizeof(float), hb = n * sizeof(uint16_t);
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, c->mlp_params.ensure(pb));
    HIPC(c, c->mlp_m.ensure(pb));
    HIPC(c, c->mlp_v.ensure(pb));
    HIPC(c, c->mlp_grad.ensure(pb));
    HIPC(c, c->mlp_wf.ensure(hb));
    HIPC(c, c->mlp_wb.ensure(hb));
    HIPC(c, c->mlp_loss.ensure(sizeof(float)));
    return RISLTC_OK;
}

int risltc_proposal_set_params(risltc_context_t* c, const float* params) {
    if (!c) return RISLTC_E_ARG;
    if (!params) return fail(c, RISLTC_E_ARG, "null params");
    int e = mlp_alloc(c);
    if (e) return e;
    const size_t pb = (size_t)c->mlp_L.params * sizeof(float);
    HIPC(c, hipMemcpy(c->mlp_params.p, params, pb, hipMemcpyHostToDevice));
    std::vector<float> zero((size_t)c->mlp_L.params, 0.0f);
    HIPC(c, hipMemcpy(c->mlp_m.p, zero.data(), pb, hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->mlp_v.p, zero.data(), pb, hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->mlp_grad.p, zero.data(), pb, hipMemcpyHostToDevice));
    HIPC(c, rl::launch_mlp_sync(c->mlp_L, c->mlp_params.as<float>(), c->mlp_wf.as<uint16_t>(),
                                c->mlp_wb.as<uint16_t>(), nullptr));
    HIPC(c, hipDeviceSynchronize());
    c->mlp_ready = true;
    c->adam_step = 0;
    return RISLTC_OK;
}

int risltc_proposal_init(risltc_context_t* c, uint32_t seed) {
    if (!c) return RISLTC_E_ARG;
    // uniform He initialisation, weights in [-sqrt(6/in), sqrt(6/in)], zero
    // biases; the output layer starts small so the first proposal is flat
    const rl::MlpLayout& M = c->mlp_L;
    std::vector<float> p((size_t)M.params, 0.0f);
    struct L {
        int w, nout, nin;
        float gain;
    } layers[4] = {{M.pw1, M.hid, rl::kMlpIn, 1.0f},
                   {M.pw2, M.hid, M.hid, 1.0f},
                   {M.pw3, M.hid, M.hid, 1.0f},
                   {M.pw4, M.out, M.hid, 0.1f}};
    for (const L& l : layers) {
        const float lim = l.gain * std::sqrt(6.0f / (float)l.nin);
        for (int e = 0; e < l.nout * l.nin; ++e) {
            uint32_t h = mix32(seed * 0x9E3779B9u ^ mix32((uint32_t)(l.w + e)));
            float u = (float)(h >> 8) * (1.0f / 16777216.0f);
            p[l.w + e] = (2.0f * u - 1.0f) * lim;
        }
    }
    return risltc_proposal_set_params(c, p.data());
}

int risltc_proposal_get_params(risltc_context_t* c, float* out, float* grad_out) {
    if (!c) return RISLTC_E_ARG;
    if (!c->mlp_ready) return fail(c, RISLTC_E_STATE, "the proposal MLP is not initialised");
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipDeviceSynchronize());
    const size_t pb = (size_t)c->mlp_L.params * sizeof(float);
    if (out) HIPC(c, hipMemcpy(out, c->mlp_params.p, pb, hipMemcpyDeviceToHost));
    if (grad_out) HIPC(c, hipMemcpy(grad_out, c->mlp_grad.p, pb, hipMemcpyDeviceToHost));
    return RISLTC_OK;
}

int risltc_proposal_set_network(risltc_context_t* c, uint32_t width, uint32_t outputs) {
    if (!c) return RISLTC_E_ARG;
    if (width != (uint32_t)rl::kMlpHid && width != (uint32_t)rl::kMlpHidWide)
        return fail(c, RISLTC_E_UNSUPPORTED, "hidden width %u: 64 or 128", width);
    if (outputs != (uint32_t)rl::kMlpOut && outputs != (uint32_t)rl::kMlpOutWide)
        return fail(c, RISLTC_E_UNSUPPORTED, "outputs %u: 256 or 1024", outputs);
    if ((int)width != c->mlp_L.hid || (int)outputs != c->mlp_L.out) {
        c->mlp_L = rl::mlp_layout((int)width, (int)outputs);
        c->mlp_ready = false;  // risltc_proposal_init or _set_params next
        c->prop_width = c->prop_height = 0;  // tables of the previous network
    }
    return RISLTC_OK;
}

int risltc_proposal_set_width(risltc_context_t* c, uint32_t width) {
    if (!c) return RISLTC_E_ARG;
    return risltc_proposal_set_network(c, width, (uint32_t)c->mlp_L.out);
}

uint32_t risltc_proposal_param_count(const risltc_context_t* c) { return c ? (uint32_t)c->mlp_L.params : 0u; }

static rl::FeatureView feature_view(const risltc_context_t* c, const risltc_frame_constants_t* k, const uint32_t* vis) {
    rl::FeatureView f;
    memset(&f, 0, sizeof(f));
    f.c = *k;
    f.scene.positions = c->positions.as<uint2>();
    f.scene.normals_uv = c->normals_uv.as<ushort4>();
    f.scene.material_idx = c->material_idx.as<uint8_t>();
    f.scene.materials = c->materials.as<float>();
    f.scene.material_count = c->material_count;
    f.scene.triangle_count = c->triangle_count;
    f.visibility = vis;
    for (int a = 0; a < 3; ++a) {
        float half = 0.5f * (c->scene_hi[a] - c->scene_lo[a]);
        f.center[a] = 0.5f * (c->scene_hi[a] + c->scene_lo[a]);
        f.inv_half[a] = half > 0.0f ? 1.0f / half : 0.0f;
    }
    return f;
}

int risltc_proposal_update(risltc_context_t* c, const risltc_frame_constants_t* k, const uint32_t* vis,
                           uint8_t* table_out, float* logits_out, uint16_t* features_out, void* stream) {
    if (!c) return RISLTC_E_ARG;
    if (!k || !vis) return fail(c, RISLTC_E_ARG, "null argument");
    if (!c->mlp_ready) return fail(c, RISLTC_E_STATE, "the proposal MLP is not initialised");
    if (c->cluster_count == 0) return fail(c, RISLTC_E_STATE, "no light clusters");
    if (!c->positions.p) return fail(c, RISLTC_E_STATE, "risltc_set_scene has not been called");
    const uint32_t W = k->viewport_size[0], H = k->viewport_size[1];
    if (W == 0 || H == 0) return fail(c, RISLTC_E_ARG, "empty viewport");
    const size_t P = (size_t)W * H;
    if (P > 0xFFFFFFFFull / 2) return fail(c, RISLTC_E_UNSUPPORTED, "viewport too large");
    hipStream_t st = (hipStream_t)stream;
    HIPC(c, hipSetDevice(c->device));
    if (c->cluster_count > (uint32_t)c->mlp_L.out)
        return fail(c, RISLTC_E_STATE, "%u clusters for a network of %d outputs", c->cluster_count, c->mlp_L.out);
    const size_t tb = (size_t)rl::prop_bytes(c->mlp_L.out);
    HIPC(c, c->prop_table.ensure(P * tb));
    rl::ProposalArgs a;
    memset(&a, 0, sizeof(a));
    a.L = c->mlp_L;
    a.f = feature_view(c, k, vis);
    a.wf = c->mlp_wf.as<uint16_t>();
    a.params = c->mlp_params.as<float>();
    a.cluster_offsets = c->cluster_offsets.as<uint32_t>();
    a.cluster_count = c->cluster_count;
    a.pixel_count = (uint32_t)P;
    a.out = c->prop_table.as<uint8_t>();
    a.logits = logits_out;
    a.features = features_out;
    HIPC(c, hipEventRecord(c->ev[4], st));
    HIPC(c, rl::launch_proposal(a, st));
    HIPC(c, hipEventRecord(c->ev[5], st));
    if (table_out) HIPC(c, hipMemcpyAsync(table_out, a.out, P * tb, hipMemcpyDeviceToDevice, st));
    c->prop_width = W;
    c->prop_height = H;
    HIPC(c, hipEventSynchronize(c->ev[5]));
    hipEventElapsedTime(&c->prop_ms[0], c->ev[4], c->ev[5]);
    return RISLTC_OK;
}

int risltc_proposal_train(risltc_context_t* c, const risltc_frame_constants_t* k, const uint32_t* vis,
                          uint32_t batch, uint32_t steps, float learning_rate, uint32_t seed, float* loss_out,
                          void* stream) {
    if (!c) return RISLTC_E_ARG;
    if (!k || !vis) return fail(c, RISLTC_E_ARG, "null argument");
    if (!c->mlp_ready) return fail(c, RISLTC_E_STATE, "the proposal MLP is not initialised");
    if (c->rec_samples == 0) return fail(c, RISLTC_E_STATE, "no training records: shade a frame with clusters set");
    if (batch == 0 || batch % 256u != 0u || batch > (1u << 22))
        return fail(c, RISLTC_E_ARG, "batch %u: a multiple of 256 up to 4194304", batch);
    if ((size_t)k->viewport_size[0] * k->viewport_size[1] * c->rec_spp != c->rec_samples)
        return fail(c, RISLTC_E_STATE, "the training records are for another viewport");
    hipStream_t st = (hipStream_t)stream;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, c->train_scratch.ensure(rl::train_scratch_bytes(batch, c->mlp_L.hid, c->mlp_L.out)));
    rl::TrainArgs t;
    memset(&t, 0, sizeof(t));
    t.L = c->mlp_L;
    t.f = feature_view(c, k, vis);
    t.B = batch;
    t.sample_count = (uint32_t)c->rec_samples;
    t.spp = c->rec_spp;
    t.rec_cluster = c->rec_cluster.as<uint32_t>();
    t.rec_y = c->rec_y.as<float>();
    t.lr = learning_rate;
    t.beta1 = 0.9f;
    t.beta2 = 0.999f;
    t.eps = 1e-8f;
    t.params = c->mlp_params.as<float>();
    t.m = c->mlp_m.as<float>();
    t.v = c->mlp_v.as<float>();
    t.grad = c->mlp_grad.as<float>();
    t.wf = c->mlp_wf.as<uint16_t>();
    t.wb = c->mlp_wb.as<uint16_t>();
    t.scratch = c->train_scratch.p;
    t.loss = c->mlp_loss.as<float>();
    HIPC(c, hipEventRecord(c->ev[6], st));
    for (uint32_t i = 0; i < steps; ++i) {
        t.step = ++c->adam_step;
        t.seed = seed * 0x9E3779B9u + i;
        HIPC(c, rl::launch_train_step(t, st));
    }
    HIPC(c, hipEventRecord(c->ev[7], st));
    HIPC(c, hipEventSynchronize(c->ev[7]));
    hipEventElapsedTime(&c->prop_ms[1], c->ev[6], c->ev[7]);
    if (loss_out) HIPC(c, hipMemcpy(loss_out, c->mlp_loss.p, sizeof(float), hipMemcpyDeviceToHost));
    return RISLTC_OK;
}

int risltc_proposal_records(risltc_context_t* c, uint32_t* cluster_out, float* y_out, uint64_t* count, void* stream) {
    if (!c) return RISLTC_E_ARG;
    if (!count) return fail(c, RISLTC_E_ARG, "null count");
    *count = c->rec_samples;
    if (c->rec_samples == 0) return RISLTC_OK;
    hipStream_t st = (hipStream_t)stream;
    HIPC(c, hipSetDevice(c->device));
    if (cluster_out)
        HIPC(c, hipMemcpyAsync(cluster_out, c->rec_cluster.p, c->rec_samples * 4, hipMemcpyDeviceToDevice, st));
    if (y_out) HIPC(c, hipMemcpyAsync(y_out, c->rec_y.p, c->rec_samples * 4, hipMemcpyDeviceToDevice, st));
    return RISLTC_OK;
}

int risltc_proposal_last_timings(const risltc_context_t* c, float out_ms[2]) {
    if (!c || !out_ms) return RISLTC_E_ARG;
    out_ms[0] = c->prop_ms[0];
    out_ms[1] = c->prop_ms[1];
    return RISLTC_OK;
}

int risltc_set_bvh_builder(risltc_context_t* c, uint32_t builder) {
    if (!c) return RISLTC_E_ARG;
    if (builder != RISLTC_BVH_SAH && builder != RISLTC_BVH_LBVH) return fail(c, RISLTC_E_ARG, "builder %u", builder);
    c->bvh_builder = builder;
    return RISLTC_OK;
}

int risltc_bvh_records(risltc_context_t* c, float* out, uint64_t capacity, uint32_t* units) {
    if (!c) return RISLTC_E_ARG;
    if (!units) return fail(c, RISLTC_E_ARG, "null units");
    if (!c->bvh.rec) return fail(c, RISLTC_E_STATE, "risltc_set_scene has not been called");
    *units = c->bvh.units;
    uint64_t n = capacity < c->bvh.units ? capacity : c->bvh.units;
    if (n > 0) {
        if (!out) return fail(c, RISLTC_E_ARG, "null output");
        HIPC(c, hipSetDevice(c->device));
        HIPC(c, hipMemcpy(out, c->bvh.rec, n * 16, hipMemcpyDeviceToHost));
    }
    return RISLTC_OK;
}

// uploads a mesh in the *.vks layout and builds its hierarchy
static int upload_mesh(risltc_context_t* c, const uint32_t* positions, const uint16_t* nuv, const uint8_t* mat_idx,
                       uint64_t tri_count, const float deq_factor[3], const float deq_summand[3],
                       const risltc_material_t* materials, uint32_t material_count, DevBuf& P, DevBuf& N, DevBuf& I,
                       DevBuf& Mt, rl::BvhBuild* bvh) {
    if (!positions || !nuv || !mat_idx || !deq_factor || !deq_summand || !materials || tri_count == 0 ||
        material_count == 0)
        return fail(c, RISLTC_E_ARG, "bad scene arguments");
    if (tri_count > rl::kMaxBvhTriangles) return fail(c, RISLTC_E_UNSUPPORTED, "too many triangles for the LBVH");
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, P.ensure(tri_count * 3 * 8));
    HIPC(c, N.ensure(tri_count * 3 * 8));
    HIPC(c, I.ensure(tri_count));
    HIPC(c, Mt.ensure((size_t)material_count * 6 * sizeof(float)));
    HIPC(c, hipMemcpy(P.p, positions, tri_count * 3 * 8, hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(N.p, nuv, tri_count * 3 * 8, hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(I.p, mat_idx, tri_count, hipMemcpyHostToDevice));
    std::vector<float> m((size_t)material_count * 6);
    for (uint32_t i = 0; i < material_count; ++i) {
        for (int k = 0; k < 3; ++k) m[6 * i + k] = materials[i].base_color[k];
        for (int k = 0; k < 3; ++k) m[6 * i + 3 + k] = materials[i].specular[k];
    }
    HIPC(c, hipMemcpy(Mt.p, m.data(), m.size() * sizeof(float), hipMemcpyHostToDevice));
    rl::free_bvh(bvh);
    HIPC(c, rl::build_bvh(P.as<uint2>(), tri_count, deq_factor, deq_summand, c->bvh_builder, bvh, nullptr));
    return RISLTC_OK;
}

static rl::BvhView bvh_view_of(const rl::BvhBuild& b) {
    rl::BvhView v;
    v.rec = b.rec;
    v.units = b.units;
    v.leaf_count = b.leaf_count;
    return v;
}

int risltc_set_scene(risltc_context_t* c, const uint32_t* positions, const uint16_t* nuv, const uint8_t* mat_idx,
                     uint64_t tri_count, const float deq_factor[3], const float deq_summand[3],
                     const risltc_material_t* materials, uint32_t material_count) {
    if (!c) return RISLTC_E_ARG;
    ++c->generation;
    int e = upload_mesh(c, positions, nuv, mat_idx, tri_count, deq_factor, deq_summand, materials, material_count,
                        c->positions, c->normals_uv, c->material_idx, c->materials, &c->bvh);
    if (e) return e;
    c->triangle_count = tri_count;
    c->material_count = material_count;
    // the quantisation grid spans the scene: 21 bits per axis
    for (int a = 0; a < 3; ++a) {
        c->scene_lo[a] = deq_summand[a];
        c->scene_hi[a] = deq_summand[a] + deq_factor[a] * 2097151.0f;
    }
    const rl::BvhView v = bvh_view_of(c->bvh);
    HIPC(c, c->bvh_table.ensure(sizeof(rl::BvhView)));
    HIPC(c, hipMemcpy(c->bvh_table.p, &v, sizeof(v), hipMemcpyHostToDevice));
    return clear_vcache(c);
}

int risltc_add_scene(risltc_context_t* c, const uint32_t* positions, const uint16_t* nuv, const uint8_t* mat_idx,
                     uint64_t tri_count, const float deq_factor[3], const float deq_summand[3],
                     const risltc_material_t* materials, uint32_t material_count, uint32_t* id) {
    if (!c) return RISLTC_E_ARG;
    if (!id) return fail(c, RISLTC_E_ARG, "null id");
    auto* ss = new SceneSet();
    int e = upload_mesh(c, positions, nuv, mat_idx, tri_count, deq_factor, deq_summand, materials, material_count,
                        ss->positions, ss->normals_uv, ss->material_idx, ss->materials, &ss->bvh);
    if (e) {
        DevBuf* b[] = {&ss->positions, &ss->normals_uv, &ss->material_idx, &ss->materials};
        for (DevBuf* x : b) x->release();
        rl::free_bvh(&ss->bvh);
        delete ss;
        return e;
    }
    ss->triangle_count = tri_count;
    ss->material_count = material_count;
    c->scene_sets.push_back(ss);
    *id = (uint32_t)c->scene_sets.size();
    return RISLTC_OK;
}

// the mesh and hierarchy a view shades: id 0 = the context's own
struct SceneRef {
    const uint2* positions;
    const ushort4* normals_uv;
    const uint8_t* material_idx;
    const float* materials;
    uint32_t material_count;
    uint64_t triangle_count;
    rl::BvhView bvh;
};

static int scene_ref(risltc_context_t* c, uint32_t id, SceneRef& r) {
    if (id == 0) {
        if (!c->positions.p) return fail(c, RISLTC_E_STATE, "risltc_set_scene has not been called");
        r = SceneRef{c->positions.as<uint2>(), c->normals_uv.as<ushort4>(), c->material_idx.as<uint8_t>(),
                     c->materials.as<float>(), c->material_count, c->triangle_count, bvh_view_of(c->bvh)};
        return RISLTC_OK;
    }
    if (id > c->scene_sets.size()) return fail(c, RISLTC_E_ARG, "scene %u does not exist", id);
    const SceneSet* ss = c->scene_sets[id - 1];
    r = SceneRef{ss->positions.as<uint2>(), ss->normals_uv.as<ushort4>(), ss->material_idx.as<uint8_t>(),
                 ss->materials.as<float>(), ss->material_count, ss->triangle_count, bvh_view_of(ss->bvh)};
    return RISLTC_OK;
}

int risltc_render_visibility_scene(risltc_context_t* c, uint32_t scene, const risltc_frame_constants_t* k,
                                   uint32_t* out, void* stream) {
    if (!c) return RISLTC_E_ARG;
    if (!k || !out) return fail(c, RISLTC_E_ARG, "null argument");
    SceneRef S;
    int e = scene_ref(c, scene, S);
    if (e) return e;
    if (!S.bvh.rec) return fail(c, RISLTC_E_STATE, "risltc_set_scene has not been called");
    if (k->viewport_size[0] == 0 || k->viewport_size[1] == 0) return fail(c, RISLTC_E_ARG, "empty viewport");
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, rl::launch_primary(S.bvh, S.positions, *k, out, (hipStream_t)stream));
    return RISLTC_OK;
}

int risltc_render_visibility(risltc_context_t* c, const risltc_frame_constants_t* k, uint32_t* out, void* stream) {
    return risltc_render_visibility_scene(c, 0, k, out, stream);
}

int risltc_trace_shadow_rays(risltc_context_t* c, const float* rays, uint32_t n, uint32_t* hits, void* stream) {
    if (!c) return RISLTC_E_ARG;
    if ((!rays || !hits) && n) return fail(c, RISLTC_E_ARG, "null argument");
    if (!c->bvh.rec) return fail(c, RISLTC_E_STATE, "risltc_set_scene has not been called");
    if (n == 0) return RISLTC_OK;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, rl::launch_trace_rays(bvh_view_of(c->bvh), rays, n, hits, (hipStream_t)stream));
    return RISLTC_OK;
}

// records of the final samples: one per pixel, light sample and iteration of
// the SAMPLE_COUNT loop, each answered by up to two shadow rays
static size_t sample_records(const risltc_frame_constants_t* k, const risltc_settings_t* s) {
    return (size_t)k->viewport_size[0] * k->viewport_size[1] * s->sample_count_light * s->sample_count;
}

static int check_settings(risltc_context_t* c, const risltc_frame_constants_t* k, const risltc_settings_t* s) {
    if (!c->ltc0.p) return fail(c, RISLTC_E_STATE, "risltc_set_ltc_table has not been called");
    if (!c->positions.p) return fail(c, RISLTC_E_STATE, "risltc_set_scene has not been called");
    if (s->sample_count == 0 || s->sample_count > RISLTC_MAX_SAMPLE_COUNT)
        return fail(c, RISLTC_E_ARG, "sample_count %u outside [1, %u]", s->sample_count, RISLTC_MAX_SAMPLE_COUNT);
    const uint32_t tech = s->polygon_sampling_technique;
    if (tech != RISLTC_SAMPLE_POLYGON_LTC_CP && tech != RISLTC_SAMPLE_POLYGON_AREA_TURK &&
        tech != RISLTC_SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE)
        return fail(c, RISLTC_E_UNSUPPORTED,
                    "polygon technique %u: sample_polygon_ltc_cp, _area_turk and _projected_solid_angle are built",
                    tech);
    if (tech != RISLTC_SAMPLE_POLYGON_LTC_CP && s->proposal != RISLTC_PROPOSAL_UNIFORM)
        return fail(c, RISLTC_E_UNSUPPORTED, "polygon technique %u runs with the uniform source only", tech);
    if (s->mis_heuristic > RISLTC_MIS_OPTIMAL) return fail(c, RISLTC_E_ARG, "mis_heuristic %u", s->mis_heuristic);
    if (s->light_sampling > RISLTC_LIGHT_RESERVOIR) return fail(c, RISLTC_E_ARG, "light_sampling %u", s->light_sampling);
    if (s->sample_count_light == 0 || s->sample_count_light > 1024)
        return fail(c, RISLTC_E_ARG, "sample_count_light %u", s->sample_count_light);
    if (s->light_sampling == RISLTC_LIGHT_RESERVOIR && (s->ris_candidate_count == 0 || s->ris_candidate_count > 4096))
        return fail(c, RISLTC_E_ARG, "ris_candidate_count %u", s->ris_candidate_count);
    if (s->proposal > RISLTC_PROPOSAL_CACHE) return fail(c, RISLTC_E_ARG, "proposal %u", s->proposal);
    if (s->learned_floor_log2 > 16) return fail(c, RISLTC_E_ARG, "learned_floor_log2 %u", s->learned_floor_log2);
    if (s->skip_training_records > 1)
        return fail(c, RISLTC_E_ARG, "skip_training_records %u", s->skip_training_records);
    if (s->shadow_rays && !c->bvh.rec) return fail(c, RISLTC_E_STATE, "no hierarchy for shadow rays");
    if (k->viewport_size[0] == 0 || k->viewport_size[1] == 0) return fail(c, RISLTC_E_ARG, "empty viewport");
    if (sample_records(k, s) * 2 >= 0x80000000ull)
        return fail(c, RISLTC_E_UNSUPPORTED, "more than 2^30 sample records (pixels x spp x sample_count) per call");
    return RISLTC_OK;
}

struct LightRef {
    const float4* lights;
    const float4* spheres;
    uint32_t count;
    float bound;
};

static int light_ref(risltc_context_t* c, uint32_t id, LightRef& r) {
    if (id == 0) {
        if (c->light_count == 0) return fail(c, RISLTC_E_STATE, "risltc_set_lights has not been called");
        r = LightRef{c->lights.as<float4>(), c->spheres.as<float4>(), c->light_count, c->sphere_bound};
        return RISLTC_OK;
    }
    if (id > c->light_sets.size()) return fail(c, RISLTC_E_ARG, "light set %u does not exist", id);
    const LightSet* ls = c->light_sets[id - 1];
    r = LightRef{ls->lights.as<float4>(), ls->spheres.as<float4>(), ls->count, ls->bound};
    return RISLTC_OK;
}

static uint32_t log2_of(uint32_t v) {
    uint32_t k = 0;
    while ((1u << (k + 1)) <= v) ++k;
    return k;
}

static int check_tuning(risltc_context_t* c, const risltc_tuning_t& t) {
    bool ok = (t.wave_width == 8 || t.wave_width == 16 || t.wave_width == 32 || t.wave_width == 64) &&
              t.cull_chunk >= 1 && t.cull_chunk <= 32 && t.lds_lights <= rl::kLdsSphereMax && t.trace_refill >= 1 &&
              t.trace_refill <= 64;
    return ok ? RISLTC_OK : fail(c, RISLTC_E_ARG, "tuning outside its ranges");
}

// the descriptor of one view, without its output buffers
static void fill_shade(risltc_context_t* c, const risltc_frame_constants_t* k, const risltc_settings_t* s,
                       const uint32_t* vis, const LightRef& L, const SceneRef& S, rl::ShadeArgs& a) {
    memset(&a, 0, sizeof(a));
    a.c = *k;
    a.spp = s->sample_count_light;
    a.iters = s->sample_count;
    a.rec_words = rl::kSampleWords;
    a.M = s->ris_candidate_count;
    a.mis = s->mis_heuristic;
    a.light_sampling = s->light_sampling;
    a.fast_atan = s->fast_atan;
    a.shadow_rays = s->shadow_rays;
    a.tile_w_log2 = log2_of(c->tuning.wave_width);
    a.chunk = c->tuning.cull_chunk;
    a.lds_lights = c->tuning.lds_lights;
    a.scene.positions = S.positions;
    a.scene.normals_uv = S.normals_uv;
    a.scene.material_idx = S.material_idx;
    a.scene.materials = S.materials;
    a.scene.material_count = S.material_count;
    a.scene.triangle_count = S.triangle_count;
    a.scene.lights = L.lights;
    a.scene.spheres = L.spheres;
    a.scene.sphere_bound = L.bound;
    a.scene.light_count = L.count;
    a.scene.ltc0 = c->ltc0.as<uint16_t>();
    a.scene.ltc1 = c->ltc1.as<uint16_t>();
    a.scene.ltc_w = c->ltc_w;
    a.scene.ltc_h = c->ltc_h;
    a.scene.ltc_layers = c->ltc_layers;
    a.visibility = vis;
    uint32_t k2 = s->learned_floor_log2 ? s->learned_floor_log2 : 3u;
    a.learned_floor = std::ldexp(1.0f, -(int)k2);
    a.learned_scale = (1.0f - a.learned_floor) / 255.0f;
    a.clusters.of_light = c->cluster_of_light.as<uint16_t>();
    a.clusters.offsets = c->cluster_offsets.as<uint32_t>();
    a.clusters.lights = c->cluster_lights.as<uint32_t>();
    a.clusters.count = c->cluster_count;
}

static void fill_resolve(const risltc_frame_constants_t* k, const risltc_settings_t* s, const uint32_t* vis,
                         const float* samples, const uint32_t* hits, float* out, uint64_t triangle_count,
                         rl::ResolveArgs& r) {
    memset(&r, 0, sizeof(r));
    r.width = k->viewport_size[0];
    r.height = k->viewport_size[1];
    r.spp = s->sample_count_light;
    r.iters = s->sample_count;
    r.rec_words = rl::kSampleWords;
    r.triangle_count = triangle_count;
    r.exposure = k->exposure_factor;
    r.visibility = vis;
    r.samples = samples;
    r.hits = hits;
    r.out = out;
}

static rl::TraceArgs trace_args(risltc_context_t* c, const rl::BvhView* bvhs, rl::ShadowRay* rays, uint32_t* slots,
                                uint32_t* counter, uint32_t* hits, size_t max_rays) {
    rl::TraceArgs t;
    t.bvhs = bvhs;
    t.rays = rays;
    t.ray_slots = slots;
    t.ray_count = counter;
    t.max_rays = (uint32_t)max_rays;
    t.hits = hits;
    t.queues = counter + rl::kQueueStride;
    t.refill = c->tuning.trace_refill;
    return t;
}

// The record size of a call (the buffers are sized for full records): for
// the LTC_CP technique without debug records, under every heuristic but the
// optimal one (the occluded estimates are then +0 or NaN and travel as flag
// bits), 8 words, or 12 when the resolve reads the selected light or the
// cache key; otherwise the full 16.
static uint32_t record_words(const risltc_settings_t* s, bool light_or_key, bool debug) {
    const bool compact = !debug && s->mis_heuristic != RISLTC_MIS_OPTIMAL &&
                         s->polygon_sampling_technique == RISLTC_SAMPLE_POLYGON_LTC_CP;
    if (!compact) return (uint32_t)rl::kSampleWords;
    return light_or_key ? rl::kSampleWordsKeyed : rl::kSampleWordsCompact;
}

// Per-call work buffers of one shading pass.
struct ShadeWork {
    float* samples;
    rl::ShadowRay* rays;
    uint32_t* ray_slots;
    uint32_t* hits;
    uint32_t* counter;
};

// fixed: the work buffers to use (risltc_shade_noalloc: nothing is
// allocated and no training records are written), or null to size the
// context's own
static int shade_impl(risltc_context_t* c, const risltc_frame_constants_t* k, const risltc_settings_t* s,
                      const uint32_t* vis, float* out, float* dbg_s, uint32_t* dbg_c, void* stream,
                      const ShadeWork* fixed = nullptr) {
    if (!c) return RISLTC_E_ARG;
    if (!k || !s || !vis || !out) return fail(c, RISLTC_E_ARG, "null argument");
    if (c->light_count == 0) return fail(c, RISLTC_E_STATE, "risltc_set_lights has not been called");
    int e = check_settings(c, k, s);
    if (e) return e;
    const bool learned = s->proposal != RISLTC_PROPOSAL_UNIFORM;
    if (learned) {
        if (s->light_sampling != RISLTC_LIGHT_RESERVOIR)
            return fail(c, RISLTC_E_UNSUPPORTED, "the learned pass is a RIS pass: light_sampling must be reservoir");
        if (c->cluster_count == 0) return fail(c, RISLTC_E_STATE, "no light clusters");
    }
    if (s->proposal == RISLTC_PROPOSAL_LEARNED &&
        (!c->prop_table.p || c->prop_width != k->viewport_size[0] || c->prop_height != k->viewport_size[1]))
        return fail(c, RISLTC_E_STATE, "risltc_proposal_update has not run for this viewport");
    if (s->proposal == RISLTC_PROPOSAL_CACHE && !c->vcache.p)
        return fail(c, RISLTC_E_STATE, "no visibility cache");
    const size_t S = (size_t)k->viewport_size[0] * k->viewport_size[1] * s->sample_count_light;
    const size_t R = sample_records(k, s);
    hipStream_t st = (hipStream_t)stream;
    HIPC(c, hipSetDevice(c->device));
    ShadeWork wk;
    if (fixed) {
        wk = *fixed;
    } else {
        HIPC(c, c->samples.ensure(R * rl::kSampleWords * sizeof(float)));
        HIPC(c, c->rays.ensure(R * 2 * sizeof(rl::ShadowRay)));
        HIPC(c, c->ray_slots.ensure(R * 2 * sizeof(uint32_t)));
        HIPC(c, c->hits.ensure(R * 2 * sizeof(uint32_t)));
        HIPC(c, c->counter.ensure(kCounterWords * sizeof(uint32_t)));
        wk = ShadeWork{c->samples.as<float>(), c->rays.as<rl::ShadowRay>(), c->ray_slots.as<uint32_t>(),
                       c->hits.as<uint32_t>(), c->counter.as<uint32_t>()};
    }
    LightRef L;
    light_ref(c, 0, L);
    SceneRef SR;
    e = scene_ref(c, 0, SR);
    if (e) return e;
    rl::ShadeArgs a;
    fill_shade(c, k, s, vis, L, SR, a);
    a.samples = wk.samples;
    a.rays = wk.rays;
    a.ray_slots = wk.ray_slots;
    a.ray_count = wk.counter;
    a.debug_samples = dbg_s;
    a.debug_cands = dbg_c;
    a.learned_pass = learned ? 1u : 0u;
    a.proposal = s->proposal == RISLTC_PROPOSAL_LEARNED ? c->prop_table.as<uint8_t>() : nullptr;
    a.proposal_group_log2 = rl::prop_group_log2(c->mlp_L.out);
    const bool cached = c->vcache.p && c->cluster_count > 0;
    if (cached) {
        if (c->vcache_per_light && c->vcache_slots != c->light_count)
            return fail(c, RISLTC_E_STATE, "the visibility cache was made for %u lights", c->vcache_slots);
        a.cache = vcache_view(c);
    }
    const bool records_out = c->cluster_count > 0 && !s->skip_training_records && !fixed;
    a.rec_words = record_words(s, records_out || cached, dbg_s != nullptr || dbg_c != nullptr);
    HIPC(c, rl::launch_zero_u32(wk.counter, kCounterWords, st));
    HIPC(c, hipEventRecord(c->ev[0], st));
    // the cells the previous resolve counted in get their new bytes (timed
    // with the shade kernel)
    if (cached) HIPC(c, rl::launch_refresh_cache(a.cache, st));
    if (s->polygon_sampling_technique == RISLTC_SAMPLE_POLYGON_LTC_CP)
        HIPC(c, rl::launch_shade(a, s->exact_math != 0, st));
    else
        HIPC(c, rl::launch_shade_baseline(a, s->polygon_sampling_technique, s->exact_math != 0, st));
    HIPC(c, hipEventRecord(c->ev[1], st));
    if (s->shadow_rays)
        HIPC(c, rl::launch_trace(trace_args(c, c->bvh_table.as<rl::BvhView>(), a.rays, a.ray_slots, a.ray_count,
                                            wk.hits, R * 2),
                                 st));
    HIPC(c, hipEventRecord(c->ev[2], st));
    rl::ResolveArgs r;
    fill_resolve(k, s, vis, a.samples, wk.hits, out, SR.triangle_count, r);
    r.rec_words = a.rec_words;
    r.debug_samples = dbg_s;
    c->rec_samples = 0;
    if (c->cluster_count > 0) {
        r.cluster_of_light = c->cluster_of_light.as<uint16_t>();
        if (!s->skip_training_records && !fixed) {
            // training records for the proposal: one per pixel-sample
            HIPC(c, c->rec_cluster.ensure(S * sizeof(uint32_t)));
            HIPC(c, c->rec_y.ensure(S * sizeof(float)));
            r.rec_cluster = c->rec_cluster.as<uint32_t>();
            r.rec_y = c->rec_y.as<float>();
            c->rec_samples = S;
            c->rec_spp = s->sample_count_light;
        }
        if (cached) r.cache = a.cache;
    }
    HIPC(c, rl::launch_resolve(r, st));
    HIPC(c, hipEventRecord(c->ev[3], st));
    c->timed = true;
    c->learned_ran = false;
    c->last_ray_counter = wk.counter;
    return RISLTC_OK;
}

// ---- batches of views -------------------------------------------------------------
struct BatchPlan {
    std::vector<rl::ShadeArgs> shade;
    std::vector<rl::ResolveArgs> resolve;
    std::vector<rl::BvhView> bvhs;  // the hierarchies of the batch's scenes
    std::vector<uint32_t> tile_off, block_off;
    size_t samples = 0;
    uint32_t tiles = 0, blocks = 0, trace_blocks = 0;
    size_t bvh_capacity = 0;
    size_t lds = 0;
    bool exact = false, shadows = false;
    bool general = false;  // some view has a SAMPLE_COUNT above 1
    bool grouped = false;  // the resolve loads samples in groups (rl::resolve_grouped)
    bool learned = false, cached = false;  // learned views; the visibility cache is refreshed first
    uint32_t group_log2 = 0;                // table slots name groups of 1 << group_log2 clusters
    rl::VisCacheView cache = {};
    // whether a graph captured with plan `o` can run this plan's descriptors
    bool fits(const BatchPlan& o) const {
        return tile_off == o.tile_off && block_off == o.block_off && samples == o.samples && lds <= o.lds &&
               exact == o.exact && shadows == o.shadows && general == o.general && learned == o.learned &&
               cached == o.cached &&
               group_log2 == o.group_log2 && bvhs.size() <= o.bvh_capacity;
    }
};

static int plan_batch(risltc_context_t* c, const risltc_view_t* views, uint32_t n, BatchBuffers& B, BatchPlan& p,
                      bool allocate) {
    if (!views || n == 0 || n > 65536) return fail(c, RISLTC_E_ARG, "view count %u", n);
    p = BatchPlan();
    std::vector<size_t> off(n);
    for (uint32_t v = 0; v < n; ++v) {
        const risltc_view_t& w = views[v];
        if (!w.visibility || !w.out) return fail(c, RISLTC_E_ARG, "view %u: null buffer", v);
        int e = check_settings(c, &w.constants, &w.settings);
        if (e) return e;
        if (w.settings.polygon_sampling_technique != RISLTC_SAMPLE_POLYGON_LTC_CP)
            return fail(c, RISLTC_E_UNSUPPORTED, "view %u: batches run sample_polygon_ltc_cp only", v);
        const bool learned = w.settings.proposal != RISLTC_PROPOSAL_UNIFORM;
        if (learned != (views[0].settings.proposal != RISLTC_PROPOSAL_UNIFORM))
            return fail(c, RISLTC_E_UNSUPPORTED, "a batch is either all learned or all plain");
        if (learned) {
            if (w.light_set != 0 || w.scene != 0)
                return fail(c, RISLTC_E_UNSUPPORTED, "view %u: learned views use light set 0 and scene 0", v);
            if (w.settings.light_sampling != RISLTC_LIGHT_RESERVOIR)
                return fail(c, RISLTC_E_UNSUPPORTED, "view %u: the learned pass is a RIS pass", v);
            if (c->cluster_count == 0) return fail(c, RISLTC_E_STATE, "no light clusters");
            if (w.settings.proposal == RISLTC_PROPOSAL_LEARNED && !w.proposal_table)
                return fail(c, RISLTC_E_ARG, "view %u: no proposal table", v);
            if (w.settings.proposal == RISLTC_PROPOSAL_CACHE && !c->vcache.p)
                return fail(c, RISLTC_E_STATE, "no visibility cache");
            if (c->vcache.p && c->vcache_per_light && c->vcache_slots != c->light_count)
                return fail(c, RISLTC_E_STATE, "the visibility cache was made for %u lights", c->vcache_slots);
        }
        if ((w.settings.exact_math != 0) != (views[0].settings.exact_math != 0))
            return fail(c, RISLTC_E_UNSUPPORTED, "a batch has one exact_math for all views");
        off[v] = p.samples;
        p.samples += sample_records(&w.constants, &w.settings);
        if (p.samples * 2 >= 0x80000000ull)
            return fail(c, RISLTC_E_UNSUPPORTED, "more than 2^30 sample records per batch");
        p.shadows = p.shadows || w.settings.shadow_rays != 0;
        p.general = p.general || w.settings.sample_count > 1u;
    }
    p.exact = views[0].settings.exact_math != 0;
    p.learned = views[0].settings.proposal != RISLTC_PROPOSAL_UNIFORM;
    p.group_log2 = p.learned ? rl::prop_group_log2(c->mlp_L.out) : 0u;
    if (p.learned && c->cluster_count > (uint32_t)c->mlp_L.out)
        return fail(c, RISLTC_E_STATE, "%u clusters for a network of %d outputs", c->cluster_count, c->mlp_L.out);
    p.cached = p.learned && c->vcache.p != nullptr;
    if (p.cached) p.cache = vcache_view(c);
    if (allocate) {
        HIPC(c, B.samples.ensure(p.samples * rl::kSampleWords * sizeof(float)));
        HIPC(c, B.rays.ensure(p.samples * 2 * sizeof(rl::ShadowRay)));
        HIPC(c, B.ray_slots.ensure(p.samples * 2 * sizeof(uint32_t)));
        HIPC(c, B.hits.ensure(p.samples * 2 * sizeof(uint32_t)));
        HIPC(c, B.counter.ensure(kCounterWords * sizeof(uint32_t)));
    }
    p.tile_off.push_back(0);
    p.block_off.push_back(0);
    std::vector<int64_t> table_of(c->scene_sets.size() + 1, -1);  // scene id -> hierarchy table index
    for (uint32_t v = 0; v < n; ++v) {
        const risltc_view_t& w = views[v];
        LightRef L;
        int e = light_ref(c, w.light_set, L);
        if (e) return e;
        SceneRef SR;
        e = scene_ref(c, w.scene, SR);
        if (e) return e;
        if (table_of[w.scene] < 0) {
            table_of[w.scene] = (int64_t)p.bvhs.size();
            p.bvhs.push_back(SR.bvh);
        }
        rl::ShadeArgs a;
        fill_shade(c, &w.constants, &w.settings, w.visibility, L, SR, a);
        a.bvh_index = (uint32_t)table_of[w.scene];
        a.samples = B.samples.as<float>() + off[v] * rl::kSampleWords;
        a.rays = B.rays.as<rl::ShadowRay>();
        a.ray_slots = B.ray_slots.as<uint32_t>();
        a.ray_count = B.counter.as<uint32_t>();
        a.slot_base = (uint32_t)(off[v] * 2);
        if (p.learned) {
            a.learned_pass = 1u;
            a.proposal = w.settings.proposal == RISLTC_PROPOSAL_LEARNED ? w.proposal_table : nullptr;
            a.proposal_group_log2 = p.group_log2;
            if (p.cached) a.cache = p.cache;
        }
        a.rec_words = record_words(&w.settings, p.cached, false);
        p.lds = std::max(p.lds, rl::shade_lds_bytes(a));
        p.tile_off.push_back(p.tile_off.back() + rl::shade_tiles(a));
        p.shade.push_back(a);
        rl::ResolveArgs r;
        fill_resolve(&w.constants, &w.settings, w.visibility, a.samples, B.hits.as<uint32_t>() + off[v] * 2, w.out,
                     SR.triangle_count, r);
        r.rec_words = a.rec_words;
        if (p.cached) {
            r.cluster_of_light = c->cluster_of_light.as<uint16_t>();
            r.cache = p.cache;
        }
        const uint32_t P = w.constants.viewport_size[0] * w.constants.viewport_size[1];
        p.block_off.push_back(p.block_off.back() + (P + 255u) / 256u);
        // (the batch kernel reads its views' arguments from memory: its
        // grouped form measured slower at up to 4 samples per pixel)
        p.grouped = p.grouped || (rl::resolve_grouped(r) && r.spp > 4u);
        p.resolve.push_back(r);
    }
    p.tiles = p.tile_off.back();
    p.blocks = p.block_off.back();
    p.bvh_capacity = std::max<size_t>(p.bvhs.size(), 1u + c->scene_sets.size());
    if (allocate) HIPC(c, B.bvh_table.ensure(p.bvh_capacity * sizeof(rl::BvhView)));
    HIPC(c, rl::trace_blocks(&p.trace_blocks));
    return RISLTC_OK;
}

static int upload_plan(risltc_context_t* c, BatchBuffers& B, const BatchPlan& p) {
    HIPC(c, B.shade_args.ensure(p.shade.size() * sizeof(rl::ShadeArgs)));
    HIPC(c, B.resolve_args.ensure(p.resolve.size() * sizeof(rl::ResolveArgs)));
    HIPC(c, B.tile_offsets.ensure(p.tile_off.size() * sizeof(uint32_t)));
    HIPC(c, B.block_offsets.ensure(p.block_off.size() * sizeof(uint32_t)));
    HIPC(c, hipMemcpy(B.shade_args.p, p.shade.data(), p.shade.size() * sizeof(rl::ShadeArgs), hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(B.resolve_args.p, p.resolve.data(), p.resolve.size() * sizeof(rl::ResolveArgs),
                      hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(B.tile_offsets.p, p.tile_off.data(), p.tile_off.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(B.block_offsets.p, p.block_off.data(), p.block_off.size() * sizeof(uint32_t),
                      hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(B.bvh_table.p, p.bvhs.data(), p.bvhs.size() * sizeof(rl::BvhView), hipMemcpyHostToDevice));
    return RISLTC_OK;
}

// kernels only (no memset, no host calls): safe under stream capture
static int launch_plan(risltc_context_t* c, BatchBuffers& B, const BatchPlan& p, hipStream_t st, bool timed) {
    const uint32_t n = (uint32_t)p.shade.size();
    HIPC(c, rl::launch_zero_u32(B.counter.as<uint32_t>(), kCounterWords, st));
    if (timed) HIPC(c, hipEventRecord(c->ev[0], st));
    if (p.cached) HIPC(c, rl::launch_refresh_cache(p.cache, st));
    HIPC(c, rl::launch_shade_batch(B.shade_args.as<rl::ShadeArgs>(), B.tile_offsets.as<uint32_t>(), n, p.tiles, p.lds,
                                   p.exact, p.learned ? 1 : 0, p.general, st));
    if (timed) HIPC(c, hipEventRecord(c->ev[1], st));
    if (p.shadows)
        HIPC(c, rl::launch_trace(trace_args(c, B.bvh_table.as<rl::BvhView>(), B.rays.as<rl::ShadowRay>(),
                                            B.ray_slots.as<uint32_t>(), B.counter.as<uint32_t>(),
                                            B.hits.as<uint32_t>(), p.samples * 2),
                                 st, p.trace_blocks));
    if (timed) HIPC(c, hipEventRecord(c->ev[2], st));
    HIPC(c, rl::launch_resolve_batch(B.resolve_args.as<rl::ResolveArgs>(), B.block_offsets.as<uint32_t>(), n, p.blocks,
                                     p.grouped, st));
    if (timed) HIPC(c, hipEventRecord(c->ev[3], st));
    return RISLTC_OK;
}

int risltc_shade_batch(risltc_context_t* c, const risltc_view_t* views, uint32_t n, void* stream) {
    if (!c) return RISLTC_E_ARG;
    HIPC(c, hipSetDevice(c->device));
    // the buffers and descriptors of the previous call may still be in use
    // (ev[3] follows every shading call's last kernel)
    HIPC(c, hipEventSynchronize(c->ev[3]));
    BatchPlan p;
    int e = plan_batch(c, views, n, c->batch, p, true);
    if (e) return e;
    hipStream_t st = (hipStream_t)stream;
    e = upload_plan(c, c->batch, p);
    if (e) return e;
    e = launch_plan(c, c->batch, p, st, true);
    if (e) return e;
    c->timed = true;
    c->learned_ran = false;
    c->last_ray_counter = nullptr;  // a batch's rays are not counted per call
    return RISLTC_OK;
}

int risltc_get_tuning(const risltc_context_t* c, risltc_tuning_t* out) {
    if (!c || !out) return RISLTC_E_ARG;
    *out = c->tuning;
    return RISLTC_OK;
}

int risltc_set_tuning(risltc_context_t* c, const risltc_tuning_t* t) {
    if (!c) return RISLTC_E_ARG;
    if (!t) return fail(c, RISLTC_E_ARG, "null tuning");
    int e = check_tuning(c, *t);
    if (e) return e;
    c->tuning = *t;
    ++c->generation;
    return RISLTC_OK;
}

// median GPU time (shade + trace + resolve) of `repeats` runs of a batch
static int time_batch(risltc_context_t* c, const risltc_view_t* views, uint32_t n, uint32_t repeats, float* ms) {
    BatchPlan p;
    int e = plan_batch(c, views, n, c->batch, p, true);
    if (e) return e;
    HIPC(c, hipDeviceSynchronize());
    e = upload_plan(c, c->batch, p);
    if (e) return e;
    std::vector<float> t;
    for (uint32_t r = 0; r <= repeats; ++r) {
        e = launch_plan(c, c->batch, p, nullptr, true);
        if (e) return e;
        HIPC(c, hipEventSynchronize(c->ev[3]));
        float a = 0, b = 0, d = 0;
        HIPC(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
        HIPC(c, hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
        HIPC(c, hipEventElapsedTime(&d, c->ev[2], c->ev[3]));
        if (r > 0) t.push_back(a + b + d);  // the first run warms caches and code
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return RISLTC_OK;
}

int risltc_autotune(risltc_context_t* c, const risltc_view_t* views, uint32_t n, uint32_t repeats,
                    risltc_tuning_t* out) {
    if (!c) return RISLTC_E_ARG;
    if (repeats == 0 || repeats > 1000) return fail(c, RISLTC_E_ARG, "repeats %u", repeats);
    HIPC(c, hipSetDevice(c->device));
    const std::vector<uint32_t> space[4] = {{8, 16, 32, 64}, {8, 16, 32}, {0, 512, 1024, 2048}, {16, 24, 32, 40, 48}};
    risltc_tuning_t best = c->tuning;
    auto field = [](risltc_tuning_t& t, int i) -> uint32_t& {
        return i == 0 ? t.wave_width : i == 1 ? t.cull_chunk : i == 2 ? t.lds_lights : t.trace_refill;
    };
    float best_ms = 0;
    int e = time_batch(c, views, n, repeats, &best_ms);
    if (e) return e;
    for (int sweep = 0; sweep < 2; ++sweep) {
        bool changed = false;
        for (int i = 0; i < 4; ++i) {
            for (uint32_t v : space[i]) {
                risltc_tuning_t t = best;
                if (field(t, i) == v) continue;
                field(t, i) = v;
                c->tuning = t;
                float ms = 0;
                e = time_batch(c, views, n, repeats, &ms);
                if (e) {
                    c->tuning = best;
                    return e;
                }
                if (ms < best_ms) {
                    best_ms = ms;
                    best = t;
                    changed = true;
                }
            }
            c->tuning = best;
        }
        if (!changed) break;
    }
    c->tuning = best;
    ++c->generation;
    if (out) *out = best;
    return RISLTC_OK;
}

}  // extern "C"

struct risltc_graph_s {
    risltc_context_t* ctx = nullptr;
    BatchBuffers buf;
    BatchPlan plan;
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    bool launched = false;
    uint64_t generation = 0;
    ~risltc_graph_s() {
        if (ctx) hipSetDevice(ctx->device);
        if (done) hipEventSynchronize(done);
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        if (done) hipEventDestroy(done);
        if (capture) hipStreamDestroy(capture);
        buf.release();
    }
};

extern "C" {

int risltc_graph_create(risltc_context_t* c, const risltc_view_t* views, uint32_t n, risltc_graph_t** out) {
    if (!c) return RISLTC_E_ARG;
    if (!out) return fail(c, RISLTC_E_ARG, "null graph");
    *out = nullptr;
    HIPC(c, hipSetDevice(c->device));
    auto* g = new risltc_graph_s();
    g->ctx = c;
    g->generation = c->generation;
    int e = plan_batch(c, views, n, g->buf, g->plan, true);
    if (!e) e = upload_plan(c, g->buf, g->plan);
    if (e) {
        delete g;
        return e;
    }
    hipError_t h = hipStreamCreateWithFlags(&g->capture, hipStreamNonBlocking);
    if (h == hipSuccess) h = hipEventCreateWithFlags(&g->done, hipEventDisableTiming);
    if (h == hipSuccess) h = hipStreamBeginCapture(g->capture, hipStreamCaptureModeThreadLocal);
    if (h == hipSuccess) {
        e = launch_plan(c, g->buf, g->plan, g->capture, false);
        // end the capture whether or not a launch failed
        hipGraph_t graph = nullptr;
        hipError_t h2 = hipStreamEndCapture(g->capture, &graph);
        g->graph = graph;
        if (!e && h2 != hipSuccess) h = h2;
    }
    if (h == hipSuccess && !e) h = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (h != hipSuccess || e) {
        delete g;
        return e ? e : fail(c, RISLTC_E_HIP, "graph capture: %s", hipGetErrorString(h));
    }
    *out = g;
    return RISLTC_OK;
}

int risltc_graph_update(risltc_graph_t* g, const risltc_view_t* views, uint32_t n) {
    if (!g) return RISLTC_E_ARG;
    risltc_context_t* c = g->ctx;
    if (g->generation != c->generation) return fail(c, RISLTC_E_STATE, "the context changed since the graph was made");
    HIPC(c, hipSetDevice(c->device));
    BatchPlan p;
    int e = plan_batch(c, views, n, g->buf, p, false);
    if (e) return e;
    if (!p.fits(g->plan)) return fail(c, RISLTC_E_ARG, "views of another shape than the graph's");
    if (g->launched) HIPC(c, hipEventSynchronize(g->done));
    e = upload_plan(c, g->buf, p);
    if (e) return e;
    p.lds = g->plan.lds;
    p.bvh_capacity = g->plan.bvh_capacity;
    g->plan = std::move(p);
    return RISLTC_OK;
}

int risltc_graph_launch(risltc_graph_t* g, void* stream) {
    if (!g) return RISLTC_E_ARG;
    risltc_context_t* c = g->ctx;
    if (g->generation != c->generation) return fail(c, RISLTC_E_STATE, "the context changed since the graph was made");
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipGraphLaunch(g->exec, (hipStream_t)stream));
    HIPC(c, hipEventRecord(g->done, (hipStream_t)stream));
    g->launched = true;
    return RISLTC_OK;
}

void risltc_graph_destroy(risltc_graph_t* g) { delete g; }

int risltc_shade(risltc_context_t* c, const risltc_frame_constants_t* k, const risltc_settings_t* s,
                 const uint32_t* vis, float* out, void* stream) {
    return shade_impl(c, k, s, vis, out, nullptr, nullptr, stream);
}

int risltc_shade_debug(risltc_context_t* c, const risltc_frame_constants_t* k, const risltc_settings_t* s,
                       const uint32_t* vis, float* out, float* samples, uint32_t* cands, void* stream) {
    if (c && !samples && !cands) return fail(c, RISLTC_E_ARG, "no debug buffer");
    return shade_impl(c, k, s, vis, out, samples, cands, stream);
}

int risltc_shade_noalloc(risltc_context_t* c, const risltc_frame_constants_t* k, const risltc_settings_t* s,
                         const uint32_t* vis, float* out, void* stream) {
    if (!c) return RISLTC_E_ARG;
    if (!k || !s) return fail(c, RISLTC_E_ARG, "null argument");
    const uint64_t S = (uint64_t)k->viewport_size[0] * k->viewport_size[1] * s->sample_count_light * s->sample_count;
    if (S > RISLTC_NOALLOC_MAX_SAMPLES)
        return fail(c, RISLTC_E_UNSUPPORTED,
                    "risltc_shade_noalloc: %llu sample records (%u x %u x %u spp x %u) exceed its fixed scratch of %u",
                    (unsigned long long)S, k->viewport_size[0], k->viewport_size[1], s->sample_count_light,
                    s->sample_count, (unsigned)RISLTC_NOALLOC_MAX_SAMPLES);
    if (!c->noalloc_scratch) return fail(c, RISLTC_E_STATE, "no scratch address for device %d", c->device);
    uint8_t* base = c->noalloc_scratch;
    const ShadeWork wk{reinterpret_cast<float*>(base + noalloc::kRecordsOff),
                       reinterpret_cast<rl::ShadowRay*>(base + noalloc::kRaysOff),
                       reinterpret_cast<uint32_t*>(base + noalloc::kSlotsOff),
                       reinterpret_cast<uint32_t*>(base + noalloc::kHitsOff),
                       reinterpret_cast<uint32_t*>(base + noalloc::kCounterOff)};
    return shade_impl(c, k, s, vis, out, nullptr, nullptr, stream, &wk);
}

int risltc_accumulate(risltc_context_t* c, float* accum, const float* frame, uint32_t count, uint32_t n, void* stream) {
    if (!c) return RISLTC_E_ARG;
    if (!accum || !frame) return fail(c, RISLTC_E_ARG, "null argument");
    if (count == 0) return RISLTC_OK;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, rl::launch_accumulate(accum, frame, count, n, (hipStream_t)stream));
    return RISLTC_OK;
}

int risltc_last_ray_count(const risltc_context_t* c, uint32_t* out) {
    if (!c || !out) return RISLTC_E_ARG;
    *out = 0;
    if (!c->last_ray_counter || !c->timed) return RISLTC_E_STATE;
    if (hipEventSynchronize(c->ev[3]) != hipSuccess) return RISLTC_E_HIP;
    if (hipMemcpy(out, c->last_ray_counter, sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
        return RISLTC_E_HIP;
    return RISLTC_OK;
}

int risltc_last_timings(const risltc_context_t* c, float ms[4]) {
    if (!c || !ms) return RISLTC_E_ARG;
    for (int i = 0; i < 4; ++i) ms[i] = 0.0f;
    if (!c->timed) return RISLTC_E_STATE;
    if (hipEventSynchronize(c->ev[3]) != hipSuccess) return RISLTC_E_HIP;
    hipEventElapsedTime(&ms[0], c->ev[0], c->ev[1]);
    hipEventElapsedTime(&ms[1], c->ev[1], c->ev[2]);
    hipEventElapsedTime(&ms[2], c->ev[2], c->ev[3]);
    return RISLTC_OK;
}

}  // extern "C"
