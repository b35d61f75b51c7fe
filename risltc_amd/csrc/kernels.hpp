// kernels.hpp -- launch parameters and entry points of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/risltc/risltc.h"

namespace rl {

// Light record repacked for the GPU: 8 x float4 = 128 bytes.
//   [0] radiance.xyz, vertex_count (uint bits)
//   [1] plane.xyzw
//   [2..7] vertices: 8 vertices x 3 floats, packed (v0.x v0.y v0.z v1.x | ...)
constexpr int kLightFloat4 = 8;

// Record of one iteration of the SAMPLE_COUNT loop of one pixel-sample
// (record index = pixel_sample * SAMPLE_COUNT + iteration), written by the
// shade kernel and read by the resolve kernel: 16 words.
//   [0] W  (0 when the sample contributes nothing)
//   [1] flags: bits 0-3 state of technique 0, bits 4-7 technique 1, bit 8 = has contribution,
//       bit 9 = joint visibility (area sampling with SAMPLE_COUNT > 1: the sample's
//       estimates count only when none of its shadow rays is occluded)
//   [2..7]  estimate if visible, technique 0 then 1 (xyz, xyz)
//   [8..13] estimate if occluded, technique 0 then 1
//   [14] selected light (-1 = none), [15] cache key
// The compact records serve a call whose MIS heuristic is not the optimal
// one: its occluded estimates are then the heuristic's weight times the zero
// integrand, +0 or NaN per channel, and flags bits 10 + 3 t + c mark the NaN
// ones (channel c of technique t). 8 words ([0..7] above) when the call needs
// neither the light nor the key (no training records, no visibility cache);
// 12 words ([0..7], then [8] light, [9] key, [10..11] unused) when it does.
constexpr int kSampleWords = 16;
constexpr uint32_t kSampleWordsCompact = 8, kSampleWordsKeyed = 12;
constexpr uint32_t kFlagContributes = 256u, kFlagJoint = 512u, kFlagOccShift = 10u;

// Shadow ray record (compacted list): float4 org.xyz + t_min, float4 dir.xyz + t_max,
// plus the slot (record index * 2 + technique) it answers.
struct ShadowRay {
    float4 org;
    float4 dir;
};

struct SceneView {
    const uint2* positions;      // 3 per triangle
    const ushort4* normals_uv;   // 3 per triangle
    const uint8_t* material_idx; // 1 per triangle
    const float* materials;      // 6 per material
    uint32_t material_count;
    uint64_t triangle_count;
    const float4* lights;        // kLightFloat4 per light
    const float4* spheres;       // per light: bounding sphere centre.xyz | radius
    float sphere_bound;          // max over lights of |centre| + radius
    uint32_t light_count;
    const uint16_t* ltc0;        // RGBA16 texels
    const uint16_t* ltc1;        // RG16 texels
    uint32_t ltc_w, ltc_h, ltc_layers;
};

// The LBVH as one array of 16-byte units in depth-first pre-order (left
// child first). Internal node: 2 units (min.xyz | escape unit, max.xyz | 0);
// a box hit continues at the next record, a miss at the escape unit (the
// first unit past the node's subtree). Leaf: 3 units (v0.xyz | -1,
// v1.xyz | triangle id, v2.xyz | light flag bit 31); it continues at the next
// record. Traversal ends at unit `units`; two zero units of padding follow.
struct BvhView {
    const float4* rec;
    uint32_t units;       // 5 * leaf_count - 2
    uint32_t leaf_count;
};
// largest triangle count whose pre-order units fit a signed 32-bit index
constexpr uint64_t kMaxBvhTriangles = 400000000ull;

// light bounding spheres are staged in LDS up to this many lights
constexpr uint32_t kLdsSphereMax = 2048;

// ---- learned RIS proposal ----------------------------------------------------------
// MLP 96 -> H -> H -> H -> 256 over light clusters, H = 64 (default) or 128.
// Parameters are one FP32 array; every weight matrix is stored [out][in],
// followed by its bias.
constexpr int kMlpIn = 96, kMlpHid = 64, kMlpOut = 256;
constexpr int kMlpHidWide = 128, kMlpOutWide = 1024;
struct MlpLayout {
    int hid;                                   // hidden width
    int pw1, pb1, pw2, pb2, pw3, pb3, pw4, pb4;  // offsets into the parameters
    int params;                                // parameter count
    int out;                                   // outputs (cluster capacity)
};
__host__ __device__ constexpr MlpLayout mlp_layout(int h, int o = kMlpOut) {
    return MlpLayout{h,
                     0,
                     h * kMlpIn,
                     h * kMlpIn + h,
                     h * kMlpIn + h + h * h,
                     h * kMlpIn + 2 * h + h * h,
                     h * kMlpIn + 2 * h + 2 * h * h,
                     h * kMlpIn + 3 * h + 2 * h * h,
                     h * kMlpIn + 3 * h + 2 * h * h + o * h,
                     h * kMlpIn + 3 * h + 2 * h * h + o * h + o,
                     o};
}
// the default layout (H = 64)
constexpr int kPW1 = 0, kPB1 = kPW1 + kMlpHid * kMlpIn;
constexpr int kPW2 = kPB1 + kMlpHid, kPB2 = kPW2 + kMlpHid * kMlpHid;
constexpr int kPW3 = kPB2 + kMlpHid, kPB3 = kPW3 + kMlpHid * kMlpHid;
constexpr int kPW4 = kPB3 + kMlpHid, kPB4 = kPW4 + kMlpOut * kMlpHid;
constexpr int kMlpParams = kPB4 + kMlpOut;  // 31,168
static_assert(mlp_layout(kMlpHid).params == kMlpParams && mlp_layout(kMlpHid).pw4 == kPW4, "default layout");
static_assert(mlp_layout(kMlpHidWide).params == 78464, "wide layout");
// Per pixel proposal table: 16 slots, each the id of a light group (bytes
// 0..15) and its weight (bytes 16..31). A group is a cluster for networks of
// at most 256 outputs, and a quad of clusters 4 q .. 4 q + 3 for the
// 1024-output network (whose clusters are numbered so that a quad's are
// neighbours). In each of 4 lane groups g = (c % 16) / 4 of the outputs the
// table keeps the 4 groups with the largest value -- the logit of a cluster,
// or log sum exp of the four logits of a quad -- at weight
// round(255 exp(value - max)).
constexpr int kPropTop = 16;
constexpr int kPropBytes = 2 * kPropTop;
__host__ __device__ constexpr int prop_bytes(int) { return kPropBytes; }
// log2 of the clusters a table slot addresses for a network's outputs
__host__ __device__ constexpr uint32_t prop_group_log2(int outputs) { return outputs > 256 ? 2u : 0u; }
// clusters a visibility cache in cluster mode has slots for
constexpr int kCacheClusterSlots = kMlpOutWide;
// default floor 1/8: scale = (1 - 0.125) / 255 in FP32
constexpr float kLearnedFloor = 0.125f;
constexpr float kLearnedScale = 0x1.c1c1c2p-9f;

struct ClusterView {
    const uint16_t* of_light;  // cluster of each light
    const uint32_t* offsets;   // count + 1 prefix offsets into lights
    const uint32_t* lights;    // light indices sorted by cluster
    uint32_t count;
};

// inputs of the feature encoder
struct FeatureView {
    risltc_frame_constants_t c;
    SceneView scene;
    const uint32_t* visibility;
    float center[3], inv_half[3];  // scene box: position -> [-1, 1]^3
};

struct ProposalArgs {
    MlpLayout L;
    FeatureView f;
    const uint16_t* wf;   // bf16 weights, master layout
    const float* params;  // FP32 master (biases)
    const uint32_t* cluster_offsets;
    uint32_t cluster_count;
    uint32_t pixel_count;
    uint8_t* out;         // kPropBytes per pixel
    float* logits;        // optional: kMlpOut per pixel
    uint16_t* features;   // optional: kMlpIn bf16 per pixel
};
hipError_t launch_proposal(const ProposalArgs& a, hipStream_t s);

// Online visibility cache: per hashed world-space cell (cell size `cell`,
// split by the dominant axis and sign of the shading normal) and slot, the
// counts of unoccluded and occluded shadow rays of the final samples whose
// light maps to that slot: its cluster (slots = kMlpOut) or the light itself
// (per_light, slots = light count). Entry (key, slot) is
// counts[key * slots + slot] = (a unoccluded, b occluded). The shade kernel
// reads a byte per entry, vis[key * slots + slot]: 255 while the entry is
// empty (the MLP's byte is used instead), else
// round(254 (a + prior / 2) / (a + b + prior)). Resolve marks the cells it
// counts in `dirty`; refresh_cache recomputes the bytes of dirty cells.
struct VisCacheView {
    uint2* counts;      // null = no cache
    uint8_t* vis;       // bytes, as above
    uint8_t* dirty;     // one byte per cell
    uint32_t mask;      // cells - 1 (a power of two)
    float inv_cell;     // 1 / cell size
    float prior;        // pseudo-counts of the symmetric Beta prior
    uint32_t slots;     // per cell
    uint32_t per_light; // slot = light index (1) or its cluster (0)
};
// recomputes the bytes of every dirty cell and clears its flag
hipError_t launch_refresh_cache(const VisCacheView& c, hipStream_t s);

// one optimisation step on a batch of (pixel, cluster, weight) items
struct TrainArgs {
    MlpLayout L;
    FeatureView f;
    uint32_t B;                    // batch size, multiple of 256
    uint32_t sample_count;         // pixel-samples in the training records
    uint32_t spp;
    uint32_t seed;                 // item selection
    const uint32_t* rec_cluster;   // per pixel-sample: cluster of the selected light or ~0
    const float* rec_y;            // per pixel-sample: fraction of its shadow rays unoccluded
    float lr, beta1, beta2, eps;
    uint32_t step;                 // 1-based Adam step
    float* params;                 // FP32 master
    float* m;
    float* v;
    float* grad;                   // out: the step's gradient
    uint16_t* wf;                  // bf16 [out][in]
    uint16_t* wb;                  // bf16 [in][out]
    void* scratch;                 // train_scratch_bytes(B)
    float* loss;                   // out: batch loss (device, one float)
};
size_t train_scratch_bytes(uint32_t B, int hid, int out);
hipError_t launch_train_step(const TrainArgs& a, hipStream_t s);
// bf16 copies of the master parameters (after set_params)
hipError_t launch_mlp_sync(const MlpLayout& L, const float* params, uint16_t* wf, uint16_t* wb, hipStream_t s);

struct ShadeArgs {
    risltc_frame_constants_t c;
    uint32_t spp, M, mis, light_sampling, fast_atan, shadow_rays;
    uint32_t iters;  // SAMPLE_COUNT: final-sample iterations per pixel-sample
    uint32_t rec_words;  // record size: kSampleWords, kSampleWordsKeyed or kSampleWordsCompact
    // launch shape (tuned; none changes a result bit): a wave covers
    // 2^tile_w_log2 x (64 >> tile_w_log2) pixels, a work-group 4 waves stacked
    // vertically; candidates are culled in chunks of `chunk` (<= 32); the
    // first lds_lights light spheres (<= kLdsSphereMax) are staged in LDS
    uint32_t tile_w_log2, chunk, lds_lights;
    uint32_t slot_base;        // added to every shadow-ray slot (batches)
    uint32_t bvh_index;        // hierarchy of the view's scene in the trace table
    SceneView scene;
    const uint32_t* visibility;
    float* samples;            // kSampleWords per pixel-sample and iteration
    ShadowRay* rays;           // compacted shadow rays
    uint32_t* ray_slots;       // slot of each compacted ray
    uint32_t* ray_count;       // device counter
    float* debug_samples;      // RISLTC_SAMPLE_RECORD_FLOATS per pixel-sample or null
    uint32_t* debug_cands;     // RISLTC_CANDIDATE_RECORD_WORDS per candidate or null
    // learned target (null = the reference's LTC target)
    // learned pass (1): three draws per candidate, the learned source when
    // proposal is set, the cache factor in the target when cache.vis is set
    uint32_t learned_pass;
    // learned source (null = uniform): kPropBytes per pixel, mixed half and
    // half with the uniform source; a slot's id names the clusters
    // id << proposal_group_log2 .. ((id + 1) << proposal_group_log2) - 1
    const uint8_t* proposal;
    uint32_t proposal_group_log2;
    ClusterView clusters;
    // learned target factor f + v (1 - f) / 255 of the visibility cache
    float learned_floor, learned_scale;  // f and (1 - f) / 255 (FP32)
    // cache keys are computed (and written to the sample records) whenever
    // counts is set; with a proposal the target reads the cache's bytes
    VisCacheView cache;
};

// The shadow-ray list is cut into kTraceQueues equal slices, one per XCD:
// work-group b runs on XCD b mod 8, starts on slice b mod 8 and takes rays
// from that slice's counter (queues[kQueueStride * q]) until it is empty,
// then moves on to the next slice. The counters start at zero.
constexpr uint32_t kTraceQueues = 8;
constexpr uint32_t kQueueStride = 32;  // words: one 128-byte line per counter

// shadow rays start 1e-3 from the shading point (scene.c:317-327 via
// get_polygon_visibility, shading_pass.frag.glsl:112-129)
constexpr float kShadowTMin = 1.0e-3f;

struct TraceArgs {
    const BvhView* bvhs;       // per ray: bvhs[the ray's hierarchy index]
    const ShadowRay* rays;     // org.w holds the hierarchy index (bits)
    const uint32_t* ray_slots;
    const uint32_t* ray_count;
    uint32_t max_rays;
    uint32_t* hits;     // one word per slot
    uint32_t* queues;   // kTraceQueues * kQueueStride words, zeroed
    uint32_t refill;    // a wave fetches rays once this many lanes are idle
};

struct ResolveArgs {
    uint32_t width, height, spp;
    uint32_t iters;  // SAMPLE_COUNT: records per pixel-sample
    uint32_t rec_words;  // record size: kSampleWords, kSampleWordsKeyed or kSampleWordsCompact
    uint64_t triangle_count;
    float exposure;
    const uint32_t* visibility;
    const float* samples;
    const uint32_t* hits;
    float* out;            // RGBA32F
    float* debug_samples;  // when set: ray outcomes and colours are written back
    // training records (optional): per pixel-sample cluster of the selected
    // light (~0 = no traced shadow ray) and the fraction of its traced shadow
    // rays that were unoccluded
    const uint16_t* cluster_of_light;
    uint32_t* rec_cluster;
    float* rec_y;
    // visibility cache (optional): every traced shadow ray of a sample is
    // counted at (the sample's cache key, its light's slot)
    VisCacheView cache;
};

// host-side launchers (shade.hip, lbvh.hip)
hipError_t launch_shade(const ShadeArgs& a, bool exact, hipStream_t s);
// the baseline techniques (RISLTC_SAMPLE_POLYGON_AREA_TURK or
// _PROJECTED_SOLID_ANGLE), uniform source, per-pixel kernel; same outputs as
// launch_shade (debug records when the debug pointers are set)
hipError_t launch_shade_baseline(const ShadeArgs& a, uint32_t technique, bool exact, hipStream_t s);
uint32_t shade_tiles(const ShadeArgs& a);
size_t shade_lds_bytes(const ShadeArgs& a);
// views and their tile offsets (count + 1 entries) in device memory
// mode: 0 plain, 1 learned (every view); general: some view has a
// SAMPLE_COUNT above 1 (the general instantiation of the tile shader)
hipError_t launch_shade_batch(const ShadeArgs* d_views, const uint32_t* d_tile_offsets, uint32_t count,
                              uint32_t total_tiles, size_t lds, bool exact, int mode, bool general, hipStream_t s);
// grouped: the resolve instantiation that loads a pixel's samples in groups
// of 4 for any sample count (some view beyond 4 samples per pixel for which
// resolve_grouped holds)
hipError_t launch_resolve_batch(const ResolveArgs* d_views, const uint32_t* d_block_offsets, uint32_t count,
                                uint32_t total_blocks, bool grouped, hipStream_t s);
// whether a resolve call takes the instantiation that loads a pixel's
// samples in groups of 4 (one iteration): measured faster for up to 4
// samples per pixel, and kept for the cache counts or training records; a
// plain pass with more samples streams faster one sample at a time
inline bool resolve_grouped(const ResolveArgs& a) {
    return a.iters == 1u && (a.spp <= 4u || a.cache.counts != nullptr || a.rec_cluster != nullptr);
}
hipError_t launch_trace(const TraceArgs& a, hipStream_t s);
// the resident grid size of the trace kernel, and a launch with that size
// computed beforehand (no device queries: usable under stream capture)
hipError_t trace_blocks(uint32_t* blocks);
hipError_t launch_trace(const TraceArgs& a, hipStream_t s, uint32_t blocks);
hipError_t launch_resolve(const ResolveArgs& a, hipStream_t s);
hipError_t launch_accumulate(float* accum, const float* frame, uint32_t pixel_count, uint32_t n,
                             hipStream_t s);
hipError_t upload_clip_table(const uint8_t* table, size_t bytes);
// clears device words with a kernel (no memset nodes, so it is safe under
// stream capture)
hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);

struct BvhBuild {
    float4* rec = nullptr;
    uint32_t units = 0;
    uint32_t leaf_count = 0;
};
// Device arrays of a binary tree over n triangles during a build: 2n - 1
// nodes, node 0 the root. Leaves have left = right = -1 and prim = triangle;
// count = triangles below a node; box = 6 floats (internal nodes).
struct SahTree {
    float* tbox;    // 6 per triangle (lo xyz, hi xyz)
    float* cent;    // 3 per triangle
    uint32_t* idx;  // triangle permutation, partitioned in place
    uint32_t* tmp;  // partition buffer
    int* left;
    int* right;
    int* parent;
    int* first;     // start of the node's range in idx
    int* count;
    uint32_t* prim;
    float* box;
    uint32_t* node_counter;
};
hipError_t build_sah_tree(const float* tris, uint32_t n, SahTree& T, hipStream_t s);
// Builds the hierarchy (RISLTC_BVH_SAH or RISLTC_BVH_LBVH) of the dequantised
// triangles of a *.vks mesh on the device and lays it out in pre-order.
hipError_t build_bvh(const uint2* d_positions, uint64_t triangle_count, const float deq_factor[3],
                     const float deq_summand[3], uint32_t builder, BvhBuild* out, hipStream_t s);
void free_bvh(BvhBuild* b);
hipError_t launch_trace_rays(const BvhView& bvh, const float* rays8, uint32_t n, uint32_t* hits,
                             hipStream_t s);
hipError_t launch_primary(const BvhView& bvh, const uint2* positions, const risltc_frame_constants_t& c,
                          uint32_t* out, hipStream_t s);

}  // namespace rl
