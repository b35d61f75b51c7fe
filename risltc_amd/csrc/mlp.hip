// mlp.hip -- the learned RIS proposal: feature encoder, MLP inference and
// training on BF16 MFMA with FP32 accumulation.
//
// No reference code: the reference draws RIS candidates uniformly
// (shading_pass.frag.glsl:730-733). The proposal adds a per-pixel
// distribution over light clusters from a 96 -> 64 -> 64 -> 64 -> 256 MLP
// (softmax of its logits); the shade kernel mixes it half and half with the
// uniform source, so every light keeps probability >= 0.5 / N and the RIS
// estimator stays unbiased. It is trained online, by softmax cross-entropy,
// on the clusters of the lights that the pass's own RIS selects.
//
// Matrix products use v_mfma_f32_16x16x32_bf16. Lane l of a wave holds
// A[l % 16][8 (l / 16) .. + 7], B[8 (l / 16) .. + 7][l % 16] and returns
// C[4 (l / 16) + i][l % 16], i = 0..3 (checked on the device against a CPU
// product). Every product is written "NT": both operands K-contiguous, so
// each fragment is one 16-byte load.
//
// Features (96, per shading point): position p in the scene box mapped to
// [-1, 1]^3 (3); sin and cos of pi 2^f p for f = 0..6 (42); shading normal n
// (3); direction to the camera o (3); reflected direction r (3); GGX alpha,
// n.o, luminance of the diffuse albedo and of F0 (4); sin and cos of pi 2^f n
// and of pi 2^f r for f = 0, 1 (24); sin and cos of pi o (6); a constant 1;
// 7 zeros.
#include <hip/hip_runtime.h>

#include "devmath.hpp"
#include "gbuffer.hpp"
#include "kernels.hpp"

namespace rl {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even (finite values)
__device__ __forceinline__ uint16_t bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_float(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// two values rounded to nearest even by v_cvt_pk_bf16_f32: a in the low half
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t bf16_pack2(float a, float b) {
    f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

__device__ __forceinline__ bf16x8 frag(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// makes one lane's LDS writes visible to the other lanes of its wave (the
// waves of a work-group run different trip counts, so no block barrier)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- feature encoder ----------------------------------------------------------------
// Writes the 96 features of the pixel's shading point; returns false (and
// zeros) for background, light and out-of-mesh pixels.
__device__ bool encode_features(const FeatureView& F, uint32_t pixel, float* f) {
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) f[k] = 0.0f;
    const SceneView& S = F.scene;
    const risltc_frame_constants_t& c = F.c;
    uint32_t prim = F.visibility[pixel];
    if (prim == 0xFFFFFFFFu || (prim >> 31) != 0u || (uint64_t)prim >= S.triangle_count) return false;
    const uint32_t W = c.viewport_size[0];
    float fx = (float)(pixel % W), fy = (float)(pixel / W);
    const float(*P)[4] = c.pixel_to_ray_direction_world_space;
    f3 ray = mk3(fma_(P[0][2], 1.0f, fma_(P[0][1], fy, P[0][0] * fx)),
                 fma_(P[1][2], 1.0f, fma_(P[1][1], fy, P[1][0] * fx)),
                 fma_(P[2][2], 1.0f, fma_(P[2][1], fy, P[2][0] * fx)));
    f3 p[3], nrm[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        size_t vi = (size_t)prim * 3 + i;
        p[i] = decode_position(S.positions[vi], c.mesh_dequantization_factor, c.mesh_dequantization_summand);
        ushort4 nt = S.normals_uv[vi];
        nrm[i] = decode_normal(nt.x, nt.y);
    }
    f3 org = mk3(c.camera_position_world_space[0], c.camera_position_world_space[1], c.camera_position_world_space[2]);
    f3 e0 = sub3(p[1], p[0]), e1 = sub3(p[2], p[0]);
    f3 rce1 = cross3(ray, e1);
    float rcp_det = 1.0f / dot3(e0, rce1);
    f3 r0 = sub3(org, p[0]);
    float by = rcp_det * dot3(r0, rce1);
    float bz = -rcp_det * dot3(ray, cross3(e0, r0));
    float bx = 1.0f - (by + bz);
    f3 pos = add3(add3(scl3(p[0], bx), scl3(p[1], by)), scl3(p[2], bz));
    f3 n = normalize3<false>(add3(add3(scl3(nrm[0], bx), scl3(nrm[1], by)), scl3(nrm[2], bz)));
    f3 o = normalize3<false>(sub3(org, pos));
    float ndo = dot3(n, o);
    if (ndo < 0.0f) {
        n = scl3(n, -1.0f);
        ndo = -ndo;
    }
    f3 r = sub3(scl3(n, 2.0f * ndo), o);
    uint32_t mi = S.material_idx[prim];
    if (mi >= S.material_count) mi = S.material_count - 1u;
    const float* m = S.materials + 6 * mi;
    float metal = m[5];
    float alum = (0.2126f * m[0] + 0.7152f * m[1] + 0.0722f * m[2]) * (1.0f - metal);
    float f0l = 0.02f * (1.0f - metal) + (0.2126f * m[0] + 0.7152f * m[1] + 0.0722f * m[2]) * metal;
    float alpha = clamp_(m[4] * m[4] * c.roughness_factor, 0.0064f, 1.0f);
    float q[3] = {(pos.x - F.center[0]) * F.inv_half[0], (pos.y - F.center[1]) * F.inv_half[1],
                  (pos.z - F.center[2]) * F.inv_half[2]};
    int k = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) f[k++] = q[a];
    // __builtin_amdgcn_sinf(x) = sin(2 pi x): sin(pi 2^j t) = sin(2 pi 2^(j-1) t)
    float scale = 0.5f;
#pragma unroll
    for (int j = 0; j < 7; ++j, scale *= 2.0f)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            f[k++] = __builtin_amdgcn_sinf(q[a] * scale);
            f[k++] = __builtin_amdgcn_cosf(q[a] * scale);
        }
    const float nv[3] = {n.x, n.y, n.z}, ov[3] = {o.x, o.y, o.z}, rv[3] = {r.x, r.y, r.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) f[k++] = nv[a];
#pragma unroll
    for (int a = 0; a < 3; ++a) f[k++] = ov[a];
#pragma unroll
    for (int a = 0; a < 3; ++a) f[k++] = rv[a];
    f[k++] = alpha;
    f[k++] = ndo;
    f[k++] = alum;
    f[k++] = f0l;
    scale = 0.5f;
#pragma unroll
    for (int j = 0; j < 2; ++j, scale *= 2.0f)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            f[k++] = __builtin_amdgcn_sinf(nv[a] * scale);
            f[k++] = __builtin_amdgcn_cosf(nv[a] * scale);
            f[k++] = __builtin_amdgcn_sinf(rv[a] * scale);
            f[k++] = __builtin_amdgcn_cosf(rv[a] * scale);
        }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        f[k++] = __builtin_amdgcn_sinf(ov[a] * 0.5f);
        f[k++] = __builtin_amdgcn_cosf(ov[a] * 0.5f);
    }
    f[k++] = 1.0f;
    return true;
}

// ---- fused inference: features -> MLP -> per-pixel top-16 table -------------------
constexpr int kLdW1 = kMlpIn + 8;  // bf16 row strides (16-byte multiples)
template <int HID>
constexpr int ld_h() { return HID + 8; }

// LDS of the inference kernel. Width 64: all weights, 8 waves (135 KB).
// Width 128: the hidden layers' weights (96 KB) and 4 waves; the output
// layer's weights (64 KB) are read from global memory, where they stay in L2.
template <int HID, int OUT, int WAVES, bool W4LDS>
struct PropShared {
    uint16_t w1[HID * kLdW1];
    uint16_t w2[HID * ld_h<HID>()];
    uint16_t w3[HID * ld_h<HID>()];
    uint16_t w4[W4LDS ? OUT * ld_h<HID>() : 8];
    float bias[3 * HID + OUT];
    uint32_t csize[OUT];
    uint16_t act_in[WAVES][16 * kLdW1];
    uint16_t act_h[WAVES][2][16 * ld_h<HID>()];
};
static_assert(sizeof(PropShared<64, 256, 8, true>) <= 160 * 1024, "width-64 LDS");
static_assert(sizeof(PropShared<128, 256, 4, false>) <= 160 * 1024, "width-128 LDS");
static_assert(sizeof(PropShared<64, 1024, 8, false>) <= 160 * 1024, "width-64, 1024 outputs LDS");
static_assert(sizeof(PropShared<128, 1024, 4, false>) <= 160 * 1024, "width-128, 1024 outputs LDS");

// one output tile's logits z[i] of clusters c0 + i (i = 0..3) for one pixel:
// written to `logits` (when not null) and merged into the sorted best-4 bv/bi
// of the live clusters, in order of i (ties keep the smaller id). Dead
// clusters read as -inf, and the tile is skipped when its largest logit does
// not beat the fourth best, which leaves the same best-4 as testing each.
__device__ __forceinline__ void merge_top4(const float (&z)[4], const bool (&live)[4], int c0, float* logits,
                                           float (&bv)[4], int (&bi)[4]) {
    if (logits) {
#pragma unroll
        for (int i = 0; i < 4; ++i) logits[i] = z[i];
    }
    float zc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) zc[i] = live[i] ? z[i] : -__builtin_inff();
    if (!(fmax_(fmax_(zc[0], zc[1]), fmax_(zc[2], zc[3])) > bv[3])) return;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (zc[i] > bv[3]) {
            bv[3] = zc[i];
            bi[3] = c0 + i;
#pragma unroll
            for (int q = 3; q > 0; --q)
                if (bv[q] > bv[q - 1]) {
                    float tv = bv[q];
                    bv[q] = bv[q - 1];
                    bv[q - 1] = tv;
                    int ti = bi[q];
                    bi[q] = bi[q - 1];
                    bi[q - 1] = ti;
                }
        }
    }
}

// The same for a network of more than 256 outputs, whose table slots name
// quads of clusters: the tile's four logits are the quad c0 / 4 of this
// lane, merged as one value, log sum exp over its live clusters (-inf when
// none is live). Ties keep the smaller quad.
__device__ __forceinline__ void merge_quad(const float (&z)[4], const bool (&live)[4], int c0, float* logits,
                                           float (&bv)[4], int (&bi)[4]) {
    if (logits) {
#pragma unroll
        for (int i = 0; i < 4; ++i) logits[i] = z[i];
    }
    float m = -__builtin_inff();
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (live[i]) m = fmax_(m, z[i]);
    if (!(m > -__builtin_inff())) return;
    float sum = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (live[i]) sum += __expf(z[i] - m);
    const float v = m + __logf(sum);
    if (!(v > bv[3])) return;
    bv[3] = v;
    bi[3] = c0 >> 2;
#pragma unroll
    for (int q = 3; q > 0; --q)
        if (bv[q] > bv[q - 1]) {
            float tv = bv[q];
            bv[q] = bv[q - 1];
            bv[q - 1] = tv;
            int ti = bi[q];
            bi[q] = bi[q - 1];
            bi[q - 1] = ti;
        }
}

// hidden layer on a 16-pixel tile: out[pixel][o] = relu(W[o] . act[pixel] + b[o])
template <int HID, int K, int LDW, int LDA>
__device__ __forceinline__ void hidden_layer(const uint16_t* W, const float* bias, const uint16_t* act, uint16_t* out,
                                             int lane) {
    const int g = lane >> 4, r = lane & 15;
#pragma unroll
    for (int ot = 0; ot < HID / 16; ++ot) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < K / 32; ++ks)
            acc = mfma(frag(W + (ot * 16 + r) * LDW + ks * 32 + 8 * g), frag(act + r * LDA + ks * 32 + 8 * g), acc);
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = fmax_(acc[i] + bias[ot * 16 + 4 * g + i], 0.0f);
        *reinterpret_cast<uint2*>(out + r * ld_h<HID>() + ot * 16 + 4 * g) =
            make_uint2(bf16_pack2(v[0], v[1]), bf16_pack2(v[2], v[3]));
    }
}

template <int HID, int OUT, int WAVES, bool W4LDS>
__global__ void __launch_bounds__(64 * WAVES) proposal_kernel(ProposalArgs A) {
    constexpr int kLdH = ld_h<HID>();
    constexpr MlpLayout L = mlp_layout(HID, OUT);
    __shared__ PropShared<HID, OUT, WAVES, W4LDS> sh;
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // stage the bf16 weights (padded rows) and biases
    for (int e = tid; e < HID * kMlpIn; e += blockDim.x) sh.w1[(e / kMlpIn) * kLdW1 + e % kMlpIn] = A.wf[L.pw1 + e];
    for (int e = tid; e < HID * HID; e += blockDim.x) {
        sh.w2[(e / HID) * kLdH + e % HID] = A.wf[L.pw2 + e];
        sh.w3[(e / HID) * kLdH + e % HID] = A.wf[L.pw3 + e];
    }
    if constexpr (W4LDS)
        for (int e = tid; e < OUT * HID; e += blockDim.x) sh.w4[(e / HID) * kLdH + e % HID] = A.wf[L.pw4 + e];
    // the output layer's weights: [out][in] rows of HID bf16
    const uint16_t* w4 = W4LDS ? sh.w4 : A.wf + L.pw4;
    constexpr int kLdW4 = W4LDS ? kLdH : HID;
    for (int e = tid; e < HID; e += blockDim.x) {
        sh.bias[e] = A.params[L.pb1 + e];
        sh.bias[HID + e] = A.params[L.pb2 + e];
        sh.bias[2 * HID + e] = A.params[L.pb3 + e];
    }
    for (int e = tid; e < OUT; e += blockDim.x) {
        sh.bias[3 * HID + e] = A.params[L.pb4 + e];
        sh.csize[e] = (uint32_t)e < A.cluster_count ? A.cluster_offsets[e + 1] - A.cluster_offsets[e] : 0u;
    }
    __syncthreads();
    const int g = lane >> 4, r = lane & 15;
    uint16_t* act = sh.act_in[wave];
    uint16_t* h0 = sh.act_h[wave][0];
    uint16_t* h1 = sh.act_h[wave][1];
    // a wave takes 64 pixels at a time: every lane encodes the features of
    // one (packed to bf16 pairs in registers), then the MLP runs on the four
    // 16-pixel sub-tiles in turn, each staged in LDS by its 16 lanes
    const uint32_t groups = (A.pixel_count + 63u) / 64u;
    for (uint32_t grp = blockIdx.x * WAVES + wave; grp < groups; grp += gridDim.x * WAVES) {
        const uint32_t my_pixel = grp * 64u + (uint32_t)lane;
        const bool my_in = my_pixel < A.pixel_count;
        uint32_t packed[kMlpIn / 2];
        bool my_shaded;
        {
            float f[kMlpIn];
            my_shaded = my_in && encode_features(A.f, my_pixel, f);
            if (!my_in) {
#pragma unroll
                for (int k = 0; k < kMlpIn; ++k) f[k] = 0.0f;
            }
#pragma unroll
            for (int k = 0; k < kMlpIn; k += 2) packed[k / 2] = bf16_pack2(f[k], f[k + 1]);
        }
        if (A.features && my_in) {
            uint32_t* fo = reinterpret_cast<uint32_t*>(A.features + (size_t)my_pixel * kMlpIn);
#pragma unroll
            for (int k = 0; k < kMlpIn / 2; ++k) fo[k] = packed[k];
        }
        if constexpr (W4LDS) {
            // output weights in LDS: the MLP runs on the four 16-pixel
            // sub-tiles in turn
#pragma unroll 1
            for (int sub = 0; sub < 4; ++sub) {
                if (g == sub) {
#pragma unroll
                    for (int k = 0; k < kMlpIn / 2; k += 4)
                        *reinterpret_cast<uint4*>(act + r * kLdW1 + 2 * k) =
                            make_uint4(packed[k], packed[k + 1], packed[k + 2], packed[k + 3]);
                }
                const uint32_t pixel = grp * 64u + (uint32_t)(16 * sub + r);
                const bool in = pixel < A.pixel_count;
                // every lane of the pixel's column learns whether it is shaded
                const bool shaded = __shfl(my_shaded ? 1 : 0, 16 * sub + r) != 0;
                wave_sync();
                hidden_layer<HID, kMlpIn, kLdW1, kLdW1>(sh.w1, sh.bias, act, h0, lane);
                wave_sync();
                hidden_layer<HID, HID, kLdH, kLdH>(sh.w2, sh.bias + HID, h0, h1, lane);
                wave_sync();
                hidden_layer<HID, HID, kLdH, kLdH>(sh.w3, sh.bias + 2 * HID, h1, h0, lane);
                wave_sync();
                // output layer: this lane keeps logits of clusters 16 t + 4 g + i of pixel r,
                // and the 4 largest of them
                float bv[4] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
                int bi[4] = {0, 0, 0, 0};
#pragma unroll 2
                for (int t = 0; t < OUT / 16; ++t) {
                    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int ks = 0; ks < HID / 32; ++ks)
                        acc = mfma(frag(w4 + (t * 16 + r) * kLdW4 + ks * 32 + 8 * g),
                                   frag(h0 + r * kLdH + ks * 32 + 8 * g), acc);
                    const int c0 = t * 16 + 4 * g;
                    float z[4];
                    bool live[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        z[i] = acc[i] + sh.bias[3 * HID + c0 + i];
                        live[i] = sh.csize[c0 + i] != 0u;
                    }
                    float* lo = A.logits && in ? A.logits + (size_t)pixel * OUT + c0 : nullptr;
                    if constexpr (OUT > 256) merge_quad(z, live, c0, lo, bv, bi);
                    else merge_top4(z, live, c0, lo, bv, bi);
                }
                // the pixel's maximum over its four lanes (lanes r, r+16, r+32, r+48)
                float zmax = bv[0];
                zmax = fmax_(zmax, __shfl_xor(zmax, 16));
                zmax = fmax_(zmax, __shfl_xor(zmax, 32));
                uint32_t ws = 0;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    uint32_t w = 0;
                    if (shaded && bv[s] > -__builtin_inff()) {
                        float e = __expf(bv[s] - zmax) * 255.0f + 0.5f;
                        w = e >= 255.0f ? 255u : (uint32_t)e;
                    }
                    if (w == 0u) bi[s] = 0;
                    ws |= w << (8 * s);
                }
                if (in) {
                    // slot 4 g + s: id at byte 4 g + s, weight at 16 + 4 g + s
                    uint32_t* o = reinterpret_cast<uint32_t*>(A.out + (size_t)pixel * kPropBytes);
                    o[g] = (uint32_t)bi[0] | ((uint32_t)bi[1] << 8) | ((uint32_t)bi[2] << 16) | ((uint32_t)bi[3] << 24);
                    o[4 + g] = ws;
                }
                wave_sync();
            }
        } else {
            // hidden layers per 16-pixel sub-tile; the last layer's activations of
            // the four sub-tiles stay in registers as MFMA B fragments (sub-tile s
            // in hb[3 - s] after the loop), so that the output layer loads each of
            // its weight fragments once for all 64 pixels
            bf16x8 hb[4][HID / 32];
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int ks = 0; ks < HID / 32; ++ks) hb[q][ks] = bf16x8{};
#pragma unroll 1
            for (int sub = 0; sub < 4; ++sub) {
                if (g == sub) {
#pragma unroll
                    for (int k = 0; k < kMlpIn / 2; k += 4)
                        *reinterpret_cast<uint4*>(act + r * kLdW1 + 2 * k) =
                            make_uint4(packed[k], packed[k + 1], packed[k + 2], packed[k + 3]);
                }
                wave_sync();
                hidden_layer<HID, kMlpIn, kLdW1, kLdW1>(sh.w1, sh.bias, act, h0, lane);
                wave_sync();
                hidden_layer<HID, HID, kLdH, kLdH>(sh.w2, sh.bias + HID, h0, h1, lane);
                wave_sync();
                hidden_layer<HID, HID, kLdH, kLdH>(sh.w3, sh.bias + 2 * HID, h1, h0, lane);
                wave_sync();
#pragma unroll
                for (int q = 3; q > 0; --q)
#pragma unroll
                    for (int ks = 0; ks < HID / 32; ++ks) hb[q][ks] = hb[q - 1][ks];
#pragma unroll
                for (int ks = 0; ks < HID / 32; ++ks) hb[0][ks] = frag(h0 + r * kLdH + ks * 32 + 8 * g);
                // the next sub-tile overwrites act and h0
                wave_sync();
            }
            // output layer: for sub-tile s this lane keeps the logits of clusters
            // 16 t + 4 g + i of pixel 16 s + r, and the 4 largest of them
            float bv[4][4];
            int bi[4][4];
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    bv[s][i] = -__builtin_inff();
                    bi[s][i] = 0;
                }
            bf16x8 wa[HID / 32];
#pragma unroll
            for (int ks = 0; ks < HID / 32; ++ks) wa[ks] = frag(w4 + r * kLdW4 + ks * 32 + 8 * g);
#pragma unroll 2
            for (int t = 0; t < OUT / 16; ++t) {
                // the next tile's weight fragments load while this tile's MFMAs run
                const int tn = t + 1 < OUT / 16 ? t + 1 : t;
                bf16x8 wn[HID / 32];
#pragma unroll
                for (int ks = 0; ks < HID / 32; ++ks) wn[ks] = frag(w4 + (tn * 16 + r) * kLdW4 + ks * 32 + 8 * g);
                float b4[4];
                bool live[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    b4[i] = sh.bias[3 * HID + t * 16 + 4 * g + i];
                    live[i] = sh.csize[t * 16 + 4 * g + i] != 0u;
                }
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int ks = 0; ks < HID / 32; ++ks) acc = mfma(wa[ks], hb[3 - s][ks], acc);
                    const uint32_t pixel = grp * 64u + (uint32_t)(16 * s + r);
                    const int c0 = t * 16 + 4 * g;
                    float z[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) z[i] = acc[i] + b4[i];
                    float* lo = A.logits && pixel < A.pixel_count ? A.logits + (size_t)pixel * OUT + c0 : nullptr;
                    if constexpr (OUT > 256) merge_quad(z, live, c0, lo, bv[s], bi[s]);
                    else merge_top4(z, live, c0, lo, bv[s], bi[s]);
                }
#pragma unroll
                for (int ks = 0; ks < HID / 32; ++ks) wa[ks] = wn[ks];
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const uint32_t pixel = grp * 64u + (uint32_t)(16 * s + r);
                const bool in = pixel < A.pixel_count;
                // every lane of the pixel's column learns whether it is shaded
                const bool shaded = __shfl(my_shaded ? 1 : 0, 16 * s + r) != 0;
                // the pixel's maximum over its four lanes (lanes r, r+16, r+32, r+48)
                float zmax = bv[s][0];
                zmax = fmax_(zmax, __shfl_xor(zmax, 16));
                zmax = fmax_(zmax, __shfl_xor(zmax, 32));
                uint32_t ws = 0;
                int id[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    uint32_t w = 0;
                    if (shaded && bv[s][q] > -__builtin_inff()) {
                        float e = __expf(bv[s][q] - zmax) * 255.0f + 0.5f;
                        w = e >= 255.0f ? 255u : (uint32_t)e;
                    }
                    id[q] = w == 0u ? 0 : bi[s][q];
                    ws |= w << (8 * q);
                }
                if (in) {
                    // slot 4 g + q: id at byte 4 g + q, weight at 16 + 4 g + q
                    uint32_t* o = reinterpret_cast<uint32_t*>(A.out + (size_t)pixel * kPropBytes);
                    o[g] = (uint32_t)id[0] | ((uint32_t)id[1] << 8) | ((uint32_t)id[2] << 16) | ((uint32_t)id[3] << 24);
                    o[4 + g] = ws;
                }
            }
        }
    }
}

hipError_t launch_proposal(const ProposalArgs& a, hipStream_t s) {
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    const bool wide = a.L.hid == kMlpHidWide, many = a.L.out == kMlpOutWide;
    if ((!wide && a.L.hid != kMlpHid) || (!many && a.L.out != kMlpOut)) return hipErrorInvalidValue;
    const uint32_t waves = wide ? 4u : 8u;
    uint32_t groups = (a.pixel_count + 63u) / 64u;
    uint32_t blocks = (groups + waves - 1u) / waves;
    uint32_t cap = (uint32_t)(cus > 0 ? cus : 1);
    if (blocks > cap) blocks = cap;
    dim3 grid(blocks), block(64 * waves);
    if (wide && many)
        hipLaunchKernelGGL((proposal_kernel<kMlpHidWide, kMlpOutWide, 4, false>), grid, block, 0, s, a);
    else if (wide)
        hipLaunchKernelGGL((proposal_kernel<kMlpHidWide, kMlpOut, 4, false>), grid, block, 0, s, a);
    else if (many)
        hipLaunchKernelGGL((proposal_kernel<kMlpHid, kMlpOutWide, 8, false>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((proposal_kernel<kMlpHid, kMlpOut, 8, true>), grid, block, 0, s, a);
    return hipGetLastError();
}

// ---- training --------------------------------------------------------------------
// C[M][N] = A[M][K] . Bt[N][K]^T, bf16 operands, FP32 accumulation. One wave
// computes a 16 x 32 block; gridDim.z splits K into slices of kslice.
enum GemmEpilogue { kEpiHidden = 0, kEpiLogits = 1, kEpiBackward = 2, kEpiPartial = 3 };

struct GemmArgs {
    const uint16_t* A;
    const uint16_t* Bt;
    uint32_t M, N, K, lda, ldb, kslice;
    const float* bias;      // hidden, logits
    uint16_t* out;          // hidden, backward: [M][N] bf16
    uint16_t* out_t;        // hidden, backward: [N][M] bf16
    float* outf;            // logits: [M][N]; partial: [slice][M][N]
    const uint16_t* mask;   // backward: [M][N] bf16 activation (relu mask)
};

template <int EPI>
__global__ void __launch_bounds__(256) gemm_nt_kernel(GemmArgs g) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int gq = lane >> 4, r = lane & 15;
    const uint32_t nblk = g.N / 32u;
    const uint32_t blk = blockIdx.x * 4u + (uint32_t)wave;
    if (blk >= (g.M / 16u) * nblk) return;
    const uint32_t m0 = (blk / nblk) * 16u, n0 = (blk % nblk) * 32u;
    const uint32_t k0 = blockIdx.z * g.kslice, k1 = min(g.K, k0 + g.kslice);
    f32x4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = acc0;
    for (uint32_t k = k0; k < k1; k += 32u) {
        bf16x8 a = frag(g.A + (size_t)(m0 + r) * g.lda + k + 8 * gq);
        acc0 = mfma(a, frag(g.Bt + (size_t)(n0 + r) * g.ldb + k + 8 * gq), acc0);
        acc1 = mfma(a, frag(g.Bt + (size_t)(n0 + 16 + r) * g.ldb + k + 8 * gq), acc1);
    }
    // acc_t[i] = C[m0 + 4 gq + i][n0 + 16 t + r]
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const f32x4 acc = t ? acc1 : acc0;
        const uint32_t n = n0 + 16u * t + (uint32_t)r;
        uint32_t packed[2] = {0u, 0u};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t m = m0 + 4u * gq + i;
            float v = acc[i];
            if constexpr (EPI == kEpiHidden) {
                v = fmax_(v + g.bias[n], 0.0f);
            } else if constexpr (EPI == kEpiLogits) {
                g.outf[(size_t)m * g.N + n] = v + g.bias[n];
            } else if constexpr (EPI == kEpiBackward) {
                if (!(bf16_float(g.mask[(size_t)m * g.N + n]) > 0.0f)) v = 0.0f;
            } else {
                g.outf[((size_t)blockIdx.z * g.M + m) * g.N + n] = v;
            }
            if constexpr (EPI == kEpiHidden || EPI == kEpiBackward) {
                uint16_t b = bf16_bits(v);
                g.out[(size_t)m * g.N + n] = b;
                packed[i >> 1] |= (uint32_t)b << (16 * (i & 1));
            }
        }
        if constexpr (EPI == kEpiHidden || EPI == kEpiBackward)
            *reinterpret_cast<uint2*>(g.out_t + (size_t)n * g.M + m0 + 4u * gq) = make_uint2(packed[0], packed[1]);
    }
}

template <int EPI>
static void gemm(const GemmArgs& g, uint32_t slices, hipStream_t s) {
    uint32_t blocks = ((g.M / 16u) * (g.N / 32u) + 3u) / 4u;
    hipLaunchKernelGGL(gemm_nt_kernel<EPI>, dim3(blocks, 1, slices), dim3(256), 0, s, g);
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// batch item b: a pseudo-random pixel-sample; features (both layouts), the
// cluster of its selected light and a weight (1, or 0 without a selection)
__global__ void train_gather_kernel(TrainArgs A, uint16_t* X, uint16_t* Xt, uint32_t* target, float* y) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= A.B) return;
    const uint32_t item = hash32(A.seed * 0x9E3779B9u + b) % A.sample_count;
    const uint32_t pixel = item / A.spp;
    float f[kMlpIn];
    bool ok = encode_features(A.f, pixel, f);
    uint32_t c = A.rec_cluster[item];
    float w = 1.0f;
    if (!ok || c >= (uint32_t)A.L.out) {
        c = 0u;
        w = 0.0f;
    }
    target[b] = c;
    y[b] = w;
    for (int k = 0; k < kMlpIn; ++k) {
        uint16_t h = bf16_bits(f[k]);
        X[(size_t)b * kMlpIn + k] = h;
        Xt[(size_t)k * A.B + b] = h;
    }
}

// sum of y over the batch in a fixed order (one block)
__global__ void batch_sum_kernel(const float* y, uint32_t B, float* out) {
    __shared__ float part[256];
    float s = 0.0f;
    for (uint32_t i = threadIdx.x; i < B; i += 256u) s += y[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0];
}

// dL/dz = (y_b / Y) (softmax(z_b) - onehot(c_b)); per-row loss terms. A
// work-group takes ROWS rows, one wave per row at a time (each lane OUT / 64
// of the logits, coalesced), and writes the transposed gradient Gt through
// LDS so that each thread stores whole 16-byte pieces of its columns.
template <int OUT>
constexpr int sg_rows() { return OUT > 256 ? 16 : 64; }

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax_(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

template <int OUT>
__global__ void __launch_bounds__(256) softmax_grad_kernel(const float* Z, const uint32_t* target, const float* y,
                                                           const float* ysum, uint32_t B, uint16_t* G, uint16_t* Gt,
                                                           float* row_loss) {
    constexpr int ROWS = sg_rows<OUT>(), PER = OUT / 64, LD = OUT + 8;
    __shared__ uint16_t tile[ROWS * LD];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t b0 = blockIdx.x * ROWS;  // B is a multiple of ROWS
    const float Y = *ysum;
    for (uint32_t rr = 0; rr < ROWS / 4; ++rr) {
        const uint32_t r = wave * (ROWS / 4) + rr, b = b0 + r;
        const float* z = Z + (size_t)b * OUT;
        float v[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) v[k] = z[lane + 64u * k];
        float m = v[0];
#pragma unroll
        for (int k = 1; k < PER; ++k) m = fmax_(m, v[k]);
        const float zm = wave_max(m);
        float e[PER], part = 0.0f;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            e[k] = __expf(v[k] - zm);
            part += e[k];
        }
        const float s = wave_sum(part);
        const float wb = Y > 0.0f ? y[b] / Y : 0.0f;
        const uint32_t c = target[b];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t n = lane + 64u * k;
            float d = wb * (e[k] / s - (n == c ? 1.0f : 0.0f));
            uint16_t h = bf16_bits(d);
            G[(size_t)b * OUT + n] = h;
            tile[r * LD + n] = h;
        }
        // z[c] from the lane that holds it (c is the same for the whole wave)
        float vc = v[0];
#pragma unroll
        for (int k = 1; k < PER; ++k)
            if ((c >> 6) == (uint32_t)k) vc = v[k];
        const float zc = __shfl(vc, (int)(c & 63u));
        if (lane == 0) row_loss[b] = wb > 0.0f ? -wb * (zc - zm - __logf(s)) : 0.0f;
    }
    __syncthreads();
    // Gt[n][b0 .. b0 + ROWS - 1], 8 rows per 16-byte store
    for (uint32_t n = threadIdx.x; n < (uint32_t)OUT; n += blockDim.x) {
#pragma unroll
        for (int q = 0; q < ROWS; q += 8) {
            uint32_t w[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                w[j] = (uint32_t)tile[(q + 2 * j) * LD + n] | ((uint32_t)tile[(q + 2 * j + 1) * LD + n] << 16);
            *reinterpret_cast<uint4*>(Gt + (size_t)n * B + b0 + q) = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
}

// grad[w] = sum over slices (fixed order); bias grads = column sums of G
__global__ void reduce_grad_kernel(const float* partial, uint32_t slices, uint32_t count, float* dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    for (uint32_t k = 0; k < slices; ++k) s += partial[(size_t)k * count + i];
    dst[i] = s;
}

// one block per column n of Gt[N][B]: fixed-order tree sum
__global__ void bias_grad_kernel(const uint16_t* Gt, uint32_t B, float* dst) {
    __shared__ float part[256];
    const uint32_t n = blockIdx.x;
    float s = 0.0f;
    for (uint32_t i = threadIdx.x; i < B; i += 256u) s += bf16_float(Gt[(size_t)n * B + i]);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) dst[n] = part[0];
}

__global__ void loss_sum_kernel(const float* row_loss, uint32_t B, float* out) {
    __shared__ float part[256];
    float s = 0.0f;
    for (uint32_t i = threadIdx.x; i < B; i += 256u) s += row_loss[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0];
}

// bf16 copies of layer weights: wf = master layout [out][in], wb = [in][out]
__device__ __forceinline__ void layer_of(const MlpLayout& L, int e, int& off, int& nout, int& nin) {
    if (e < L.pb1) {
        off = L.pw1, nout = L.hid, nin = kMlpIn;
    } else if (e >= L.pw2 && e < L.pb2) {
        off = L.pw2, nout = L.hid, nin = L.hid;
    } else if (e >= L.pw3 && e < L.pb3) {
        off = L.pw3, nout = L.hid, nin = L.hid;
    } else if (e >= L.pw4 && e < L.pb4) {
        off = L.pw4, nout = L.out, nin = L.hid;
    } else {
        off = -1, nout = nin = 0;
    }
}

__device__ __forceinline__ void write_copies(const MlpLayout& L, int e, float v, uint16_t* wf, uint16_t* wb) {
    uint16_t h = bf16_bits(v);
    wf[e] = h;
    int off, nout, nin;
    layer_of(L, e, off, nout, nin);
    if (off >= 0) {
        int o = (e - off) / nin, i = (e - off) % nin;
        wb[off + i * nout + o] = h;
    } else {
        wb[e] = h;
    }
}

__global__ void adam_kernel(TrainArgs A) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= A.L.params) return;
    const float gr = A.grad[e];
    const float m = A.beta1 * A.m[e] + (1.0f - A.beta1) * gr;
    const float v = A.beta2 * A.v[e] + (1.0f - A.beta2) * gr * gr;
    A.m[e] = m;
    A.v[e] = v;
    const float bc1 = 1.0f - __powf(A.beta1, (float)A.step);
    const float bc2 = 1.0f - __powf(A.beta2, (float)A.step);
    const float p = A.params[e] - A.lr * (m / bc1) / (__builtin_sqrtf(v / bc2) + A.eps);
    A.params[e] = p;
    write_copies(A.L, e, p, A.wf, A.wb);
}

__global__ void sync_kernel(MlpLayout L, const float* params, uint16_t* wf, uint16_t* wb) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L.params) return;
    write_copies(L, e, params[e], wf, wb);
}

hipError_t launch_mlp_sync(const MlpLayout& L, const float* params, uint16_t* wf, uint16_t* wb, hipStream_t s) {
    hipLaunchKernelGGL(sync_kernel, dim3((L.params + 255) / 256), dim3(256), 0, s, L, params, wf, wb);
    return hipGetLastError();
}

namespace {
constexpr uint32_t kTrainSlice = 2048;  // split-K slice of the weight-gradient products
struct TrainScratch {
    uint16_t *X, *Xt, *H1, *H1t, *H2, *H2t, *H3, *H3t, *G4, *G4t, *G3, *G3t, *G2, *G2t, *G1, *G1t;
    float *Z, *y, *ysum, *row_loss, *partial;
    uint32_t* target;
};
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
TrainScratch carve(void* base, uint32_t B, int hid, int out, size_t* total) {
    TrainScratch t;
    size_t off = 0;
    char* p = static_cast<char*>(base);
    auto take = [&](size_t bytes) -> void* {
        void* q = p ? p + off : nullptr;
        off += align256(bytes);
        return q;
    };
    t.X = (uint16_t*)take(2ull * B * kMlpIn);
    t.Xt = (uint16_t*)take(2ull * B * kMlpIn);
    t.H1 = (uint16_t*)take(2ull * B * hid);
    t.H1t = (uint16_t*)take(2ull * B * hid);
    t.H2 = (uint16_t*)take(2ull * B * hid);
    t.H2t = (uint16_t*)take(2ull * B * hid);
    t.H3 = (uint16_t*)take(2ull * B * hid);
    t.H3t = (uint16_t*)take(2ull * B * hid);
    t.G4 = (uint16_t*)take(2ull * B * out);
    t.G4t = (uint16_t*)take(2ull * B * out);
    t.G3 = (uint16_t*)take(2ull * B * hid);
    t.G3t = (uint16_t*)take(2ull * B * hid);
    t.G2 = (uint16_t*)take(2ull * B * hid);
    t.G2t = (uint16_t*)take(2ull * B * hid);
    t.G1 = (uint16_t*)take(2ull * B * hid);
    t.G1t = (uint16_t*)take(2ull * B * hid);
    t.Z = (float*)take(4ull * B * out);
    t.y = (float*)take(4ull * B);
    t.ysum = (float*)take(4);
    t.row_loss = (float*)take(4ull * B);
    const uint32_t slices = (B + kTrainSlice - 1) / kTrainSlice;
    t.partial = (float*)take(4ull * slices * out * hid);
    t.target = (uint32_t*)take(4ull * B);
    if (total) *total = off;
    return t;
}
}  // namespace

size_t train_scratch_bytes(uint32_t B, int hid, int out) {
    size_t total = 0;
    carve(nullptr, B, hid, out, &total);
    return total;
}

hipError_t launch_train_step(const TrainArgs& a, hipStream_t s) {
    if (a.B == 0 || a.B % 256u != 0u || a.sample_count == 0) return hipErrorInvalidValue;
    const MlpLayout& L4 = a.L;
    const uint32_t H = (uint32_t)L4.hid;
    TrainScratch t = carve(a.scratch, a.B, L4.hid, L4.out, nullptr);
    const uint32_t O = (uint32_t)L4.out;
    const uint32_t B = a.B, slices = (B + kTrainSlice - 1) / kTrainSlice;
    hipLaunchKernelGGL(train_gather_kernel, dim3(B / 256u), dim3(256), 0, s, a, t.X, t.Xt, t.target, t.y);
    hipLaunchKernelGGL(batch_sum_kernel, dim3(1), dim3(256), 0, s, t.y, B, t.ysum);
    // forward
    GemmArgs g{};
    g = GemmArgs{t.X, a.wf + L4.pw1, B, H, kMlpIn, kMlpIn, kMlpIn, kMlpIn, a.params + L4.pb1, t.H1, t.H1t, nullptr, nullptr};
    gemm<kEpiHidden>(g, 1, s);
    g = GemmArgs{t.H1, a.wf + L4.pw2, B, H, H, H, H, H, a.params + L4.pb2, t.H2, t.H2t, nullptr, nullptr};
    gemm<kEpiHidden>(g, 1, s);
    g = GemmArgs{t.H2, a.wf + L4.pw3, B, H, H, H, H, H, a.params + L4.pb3, t.H3, t.H3t, nullptr, nullptr};
    gemm<kEpiHidden>(g, 1, s);
    g = GemmArgs{t.H3, a.wf + L4.pw4, B, O, H, H, H, H, a.params + L4.pb4, nullptr, nullptr, t.Z, nullptr};
    gemm<kEpiLogits>(g, 1, s);
    if (O > 256u)
        hipLaunchKernelGGL(softmax_grad_kernel<kMlpOutWide>, dim3(B / sg_rows<kMlpOutWide>()), dim3(256), 0, s, t.Z,
                           t.target, t.y, t.ysum, B, t.G4, t.G4t, t.row_loss);
    else
        hipLaunchKernelGGL(softmax_grad_kernel<kMlpOut>, dim3(B / sg_rows<kMlpOut>()), dim3(256), 0, s, t.Z, t.target,
                           t.y, t.ysum, B, t.G4, t.G4t, t.row_loss);
    hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(256), 0, s, t.row_loss, B, a.loss);
    // layer 4: dW4[o][i] = sum_b G4[b][o] H3[b][i]; dH3 = G4 . W4, masked by H3 > 0
    struct Layer {
        const uint16_t *Gt, *At;
        uint32_t nout, nin;
        int w, bias;
    };
    const Layer layers[4] = {{t.G4t, t.H3t, O, H, L4.pw4, L4.pb4}, {t.G3t, t.H2t, H, H, L4.pw3, L4.pb3},
                             {t.G2t, t.H1t, H, H, L4.pw2, L4.pb2}, {t.G1t, t.Xt, H, kMlpIn, L4.pw1, L4.pb1}};
    const uint16_t* Gs[4] = {t.G4, t.G3, t.G2, t.G1};
    uint16_t* Gn[4] = {t.G3, t.G2, t.G1, nullptr};
    uint16_t* Gnt[4] = {t.G3t, t.G2t, t.G1t, nullptr};
    const uint16_t* masks[4] = {t.H3, t.H2, t.H1, nullptr};
    for (int l = 0; l < 4; ++l) {
        const Layer& L = layers[l];
        g = GemmArgs{L.Gt, L.At, L.nout, L.nin, B, B, B, kTrainSlice, nullptr, nullptr, nullptr, t.partial, nullptr};
        gemm<kEpiPartial>(g, slices, s);
        const uint32_t cnt = L.nout * L.nin;
        hipLaunchKernelGGL(reduce_grad_kernel, dim3((cnt + 255u) / 256u), dim3(256), 0, s, t.partial, slices, cnt,
                           a.grad + L.w);
        hipLaunchKernelGGL(bias_grad_kernel, dim3(L.nout), dim3(256), 0, s, L.Gt, B, a.grad + L.bias);
        if (Gn[l]) {
            // dA[b][i] = sum_o G[b][o] W[o][i]: Bt = W^T stored [in][out] = wb
            g = GemmArgs{Gs[l], a.wb + L.w, B, L.nin, L.nout, L.nout, L.nout, L.nout, nullptr, Gn[l], Gnt[l], nullptr,
                         masks[l]};
            gemm<kEpiBackward>(g, 1, s);
        }
    }
    hipLaunchKernelGGL(adam_kernel, dim3((L4.params + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace rl
