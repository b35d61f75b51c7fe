// shade.hip -- the RIS-LTC shading pass on gfx950.
//
// One lane shades one pixel (main() of src/shaders/shading_pass.frag.glsl:
// 674-770): G-buffer reconstruction, LTC coefficients, then per light sample
// a reservoir over M candidate lights scored by the clipped-polygon LTC
// integral (the RIS target, shading_pass.frag.glsl:723-736 and 421-456) and
// a final projected-solid-angle sample pair combined with MIS
// (shading_pass.frag.glsl:292-397). Shadow rays are not traced inline: the
// kernel appends them to a compacted ray list (one atomic per wave, lane
// offsets from a ballot) and records the MIS estimate for both outcomes; the
// trace kernel answers the rays and the resolve kernel sums the estimates in
// the reference's order. Work-groups are 16x16 pixel tiles, 4 waves of 16x4.
//
// Template parameter X selects exact IEEE arithmetic (bitwise equal to the
// FP32 oracle) or the fast variant with hardware rcp/rsq/sqrt in the
// per-candidate and final-sample math. The per-pixel setup (shading frame,
// LTC lookup) is always exact, so clip decisions agree bit for bit in both.
#include <hip/hip_runtime.h>
#include <type_traits>

#include "devmath.hpp"
#include "gbuffer.hpp"
#include "kernels.hpp"

namespace rl {

// rotation of the clipped polygon for n = 3..8 input vertices and inside
// mask m: c_clip_rot[(n - 3) * 256 + m] (see the clipping rule in capi.cpp)
__constant__ uint8_t c_clip_rot[6 * 256];

constexpr int MAXV = 9;

// four floats in LDS, for reads the compiler must emit as LDS instructions
typedef float F4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(3))) F4v LdsF4;
__device__ __forceinline__ float4 lds_f4(const LdsF4* p, uint32_t i) {
    const F4v v = p[i];
    return make_float4(v.x, v.y, v.z, v.w);
}

struct Pixel {
    f3 pos, n, o;
    float lo;
    f3 albedo, f0;
    float rough;
    float w2s[12];  // world to shading, row-major 3x4
    float s2c[9];   // shading to cosine (the LTC matrix; five non-zero entries)
    float ltc_albedo, ltc_det;
};

struct Light {
    f3 rad;
    float plane[4];
    uint32_t n;
    float v[24];
};

__device__ __forceinline__ void load_light(const float4* lights, uint32_t idx, Light& L) {
    const float4* p = lights + (size_t)idx * kLightFloat4;
    float4 a = p[0], b = p[1];
    L.rad = mk3(a.x, a.y, a.z);
    L.n = __float_as_uint(a.w);
    L.plane[0] = b.x;
    L.plane[1] = b.y;
    L.plane[2] = b.z;
    L.plane[3] = b.w;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        float4 q = p[2 + k];
        L.v[4 * k + 0] = q.x;
        L.v[4 * k + 1] = q.y;
        L.v[4 * k + 2] = q.z;
        L.v[4 * k + 3] = q.w;
    }
}

__device__ __forceinline__ float plane_side(f3 p, const float* pl) {
    return fma_(1.0f, pl[3], fma_(p.z, pl[2], fma_(p.y, pl[1], p.x * pl[0])));
}

__device__ __forceinline__ f3 xform(const float* m, float x, float y, float z) {
    return mk3(fma_(m[2], z, fma_(m[1], y, fma_(m[0], x, m[3]))), fma_(m[6], z, fma_(m[5], y, fma_(m[4], x, m[7]))),
               fma_(m[10], z, fma_(m[9], y, fma_(m[8], x, m[11]))));
}

__device__ __forceinline__ f3 mul33(const float* m, f3 p) {
    return mk3(fma_(m[2], p.z, fma_(m[1], p.y, m[0] * p.x)), fma_(m[5], p.z, fma_(m[4], p.y, m[3] * p.x)),
               fma_(m[8], p.z, fma_(m[7], p.y, m[6] * p.x)));
}

// mat3 * mat4x3 with the convention of the oracle
__device__ __forceinline__ void mul33_34(const float* s, const float* w, float* r) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 4; ++k)
            r[4 * i + k] = fma_(s[3 * i + 2], w[8 + k], fma_(s[3 * i + 1], w[4 + k], s[3 * i + 0] * w[k]));
}

// shading_to_cosine * world_to_shading, with row 1 of world_to_shading
// negated for a shading point behind the light (the mirrored polygon).
// Derived on demand rather than kept per pixel: the candidate loop of the
// fast build never needs it, and 24 fewer live registers there.
__device__ __forceinline__ void world_to_cosine(const Pixel& px, bool flip, float* m) {
    float w[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) w[k] = (flip && k >= 4 && k < 8) ? -px.w2s[k] : px.w2s[k];
    mul33_34(px.s2c, w, m);
}

// inverse of shading_to_cosine (get_ltc_coefficients, ltc_utility.glsl:80-87)
__device__ __forceinline__ void cosine_to_shading(const Pixel& px, float* c2s) {
    const float d0x = px.s2c[0], d0w = px.s2c[2], d0z = px.s2c[4], d0y = -px.s2c[6], d1x = px.s2c[8];
    float det2 = d0x * d1x + d0y * d0w;
    float inv = 1.0f / det2;
    c2s[0] = d1x * inv; c2s[1] = 0.0f; c2s[2] = -d0w * inv;
    c2s[3] = 0.0f; c2s[4] = 1.0f / d0z; c2s[5] = 0.0f;
    c2s[6] = d0y * inv; c2s[7] = 0.0f; c2s[8] = d0x * inv;
}

// iz0 (polygon_clipping.glsl:19-25)
template <bool X>
__device__ __forceinline__ f3 iz0(f3 a, f3 b) {
    float t = div_<X>(a.z, a.z - b.z);
    return mk3(fma_(t, b.x, fma_(-t, a.x, a.x)), fma_(t, b.y, fma_(-t, a.y, a.y)), 0.0f);
}

// integrateEdgeVec (polygon_sampling.glsl:508-521) on normalised vertices
template <bool X>
__device__ __forceinline__ float edge_term(f3 a, f3 b) {
    float x = dot3(a, b);
    float y = fabs_(x);
    if constexpr (X) {
        float A = 0.8543985f + (0.4965155f + 0.0145206f * y) * y;
        float B = 3.4175940f + (4.1616724f + y) * y;
        float v = div_<X>(A, B);
        float ts = v;
        if (!(x > 0.0f)) ts = 0.5f * rsq_<X>(fmax_(1.0f - x * x, 1e-7f)) - v;
        return (a.x * b.y - a.y * b.x) * ts;
    } else {
        // the same rational and cross product in fused form; edges longer
        // than 90 degrees are rare, so the reciprocal square root runs only
        // when a lane of the wave has one
        float A = fma_(fma_(0.0145206f, y, 0.4965155f), y, 0.8543985f);
        float B = fma_(y + 4.1616724f, y, 3.4175940f);
        float v = div_<X>(A, B);
        float ts = v;
        const bool longer = !(x > 0.0f);
        if (__builtin_amdgcn_ballot_w64(longer) != 0ull) {
            // (an empty volatile statement keeps the compiler from
            // speculating this block onto the common path)
            asm volatile("" ::);
            if (longer) ts = 0.5f * rsq_<X>(fmax_(fma_(-x, x, 1.0f), 1e-7f)) - v;
        }
        return fma_(a.x, b.y, -(a.y * b.x)) * ts;
    }
}

// Clip (polygon_clipping.glsl:35-225) + calculate_ltc (polygon_sampling.glsl:
// 523-530) without materialising the clipped polygon: every input edge
// contributes its clipped part; the closing edge exit -> entry is added last.
// The exact build sums the terms in the order of the reference's rotated
// output (c_clip_rot), the fast build in input-edge order.
template <bool X>
__device__ __forceinline__ float polygon_ltc(const float* px, const float* py, const float* pz, uint32_t n,
                                             uint32_t& vc) {
    const uint32_t full = (1u << n) - 1u;
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) mask |= (pz[i] > 0.0f) ? (1u << i) : 0u;
    mask &= full;
    vc = 0;
    if (mask == 0) return 0.0f;
    uint32_t k = __builtin_popcount(mask);
    bool clipped = mask != full;
    uint32_t L = clipped ? k + 2u : n;
    vc = L;
    // slots n..8 of px/py/pz hold vertex 0 (the light records repeat it), so
    // edge i always ends at slot i + 1: static indices, no address selects
    f3 nv[9];
#pragma unroll
    for (int i = 0; i < 8; ++i) nv[i] = normalize3<X>(mk3(px[i], py[i], pz[i]));
    nv[8] = nv[0];
    const uint32_t mask9 = mask | ((mask & 1u) << n);
    f3 E = mk3(0.0f, 0.0f, 0.0f), Xp = E;
    float acc = 0.0f;
    float t[MAXV];
    uint32_t s = 0, p = 0;
    if constexpr (X) {
        uint32_t prev = ((mask << 1) | (mask >> (n - 1u))) & full;
        uint32_t starts = mask & ~prev;
        s = clipped ? (uint32_t)__builtin_ctz(starts) : 0u;
        p = clipped ? c_clip_rot[(n - 3u) * 256u + mask] : 0u;
#pragma unroll
        for (int i = 0; i < MAXV; ++i) t[i] = 0.0f;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if ((uint32_t)i < n) {
            f3 a = mk3(px[i], py[i], pz[i]);
            f3 b = mk3(px[i + 1], py[i + 1], pz[i + 1]);
            f3 na = nv[i];
            f3 nb = nv[i + 1];
            bool in_a = (mask9 >> i) & 1u;
            bool in_b = (mask9 >> (i + 1)) & 1u;
            f3 P = na;
            if (in_a != in_b) {
                P = normalize3<X>(iz0<X>(a, b));
                // on the horizon (0 * finite; the fast build states it)
                if constexpr (!X) P.z = 0.0f;
                if (in_b) E = P;
                else Xp = P;
            }
            if (in_a || in_b) {
                float term = edge_term<X>(in_a ? na : P, in_b ? nb : P);
                if constexpr (X) {
                    int c;
                    if (!clipped) c = i;
                    else if (in_a && in_b) {
                        int d = i - (int)s;
                        if (d < 0) d += (int)n;
                        c = d + 1;
                    } else c = in_b ? 0 : (int)k;
                    int r = c - (int)p;
                    if (r < 0) r += (int)L;
                    t[r] = term;
                } else {
                    acc += term;
                }
            }
        }
    }
    if (clipped) {
        float term = edge_term<X>(Xp, E);
        if constexpr (X) {
            int r = (int)(k + 1u) - (int)p;
            if (r < 0) r += (int)L;
            t[r] = term;
        } else {
            acc += term;
        }
    }
    if constexpr (X) {
#pragma unroll
        for (int i = 0; i < MAXV; ++i)
            if ((uint32_t)i < L) acc += t[i];
    }
    return fabs_(acc);
}

// ---- clipped polygon, materialised in the reference's output order --------
__device__ __forceinline__ f3 pick8(const float* vx, const float* vy, const float* vz, uint32_t idx) {
    f3 r = mk3(vx[0], vy[0], vz[0]);
#pragma unroll
    for (int i = 1; i < 8; ++i)
        if (idx == (uint32_t)i) r = mk3(vx[i], vy[i], vz[i]);
    return r;
}

// In place: vx/vy/vz hold n vertices on entry; on return the clipped polygon
// with every slot past the count equal to slot 0.
// With LDS, the input vertices are indexed through the lane's column of the
// wave's LDS pixel area (col[f * 64] is field f; 24 fields) instead of
// selects, which the compiler would otherwise turn into scratch memory.
template <bool X, bool LDS = false>
__device__ __forceinline__ uint32_t clip_materialize(float* vx, float* vy, float* vz, uint32_t n,
                                                     float* col = nullptr) {
    // (the count tests per slot keep the register allocation of the final
    // sample lower than masking with `full` afterwards)
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) mask |= ((uint32_t)i < n && vz[i] > 0.0f) ? (1u << i) : 0u;
    if (mask == 0) return 0;
    const uint32_t full = (1u << n) - 1u;
    if (mask == full) {
#pragma unroll
        for (int i = 1; i < MAXV; ++i)
            if ((uint32_t)i >= n) {
                vx[i] = vx[0];
                vy[i] = vy[0];
                vz[i] = vz[0];
            }
        return n;
    }
    uint32_t k = __builtin_popcount(mask);
    uint32_t L = k + 2u;
    uint32_t prev = ((mask << 1) | (mask >> (n - 1u))) & full;
    uint32_t s = (uint32_t)__builtin_ctz(mask & ~prev);
    uint32_t p = c_clip_rot[(n - 3u) * 256u + mask];
    uint32_t sm1 = (s == 0u) ? n - 1u : s - 1u;
    uint32_t e = s + k - 1u;
    if (e >= n) e -= n;
    uint32_t e1 = (e + 1u >= n) ? e + 1u - n : e + 1u;
    if constexpr (LDS) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            col[(3 * i + 0) * 64] = vx[i];
            col[(3 * i + 1) * 64] = vy[i];
            col[(3 * i + 2) * 64] = vz[i];
        }
    }
    auto pick = [&](uint32_t idx) -> f3 {
        if constexpr (LDS) {
            const float* q = col + idx * 192u;
            return mk3(q[0], q[64], q[128]);
        } else {
            return pick8(vx, vy, vz, idx);
        }
    };
    f3 E = iz0<X>(pick(sm1), pick(s));
    f3 Xp = iz0<X>(pick(e), pick(e1));
    float ox[MAXV], oy[MAXV], oz[MAXV];
#pragma unroll
    for (int j = 0; j < MAXV; ++j) {
        uint32_t c = (uint32_t)j + p;
        if (c >= L) c -= L;
        if ((uint32_t)j >= L) c = p;  // padding repeats slot 0
        f3 r;
        if (c == 0u) r = E;
        else if (c == k + 1u) r = Xp;
        else {
            uint32_t src = s + c - 1u;
            if (src >= n) src -= n;
            r = pick(src);
        }
        ox[j] = r.x;
        oy[j] = r.y;
        oz[j] = r.z;
    }
#pragma unroll
    for (int j = 0; j < MAXV; ++j) {
        vx[j] = ox[j];
        vy[j] = oy[j];
        vz[j] = oz[j];
    }
    return L;
}

// ---- projected solid angle sampling (polygon_sampling.glsl:245-828) --------
struct Psa {
    uint32_t vc;
    f2 v[MAXV];
    f2 e[MAXV];
    f2 inner0;
    float sector[MAXV];
    float psa;
};

__device__ __forceinline__ float kahan(float a, float b, float c, float d) {
    float cd = c * d;
    float err = fma_(c, d, -cd);
    float res = fma_(a, b, -cd);
    return res - err;
}
__device__ __forceinline__ f3 cross_stable(f3 l, f3 r) {
    return mk3(kahan(l.y, r.z, l.z, r.y), kahan(l.z, r.x, l.x, r.z), kahan(l.x, r.y, l.y, r.x));
}
__device__ __forceinline__ f2 rot90(f2 v) { return mk2(-v.y, v.x); }
__device__ __forceinline__ bool inner_e(f2 e) { return sign_bit(e.x); }

template <bool X>
__device__ __forceinline__ f2 ellipse_from_edge(f3 a, f3 b) {
    f3 nn = cross_stable(a, b);
    float s = rcp_<X>(nn.z);
    s = sign_bit(nn.x) ? -s : s;
    f2 e = mk2(nn.x * s, nn.y * s);
    e.x = (nn.z != 0.0f) ? e.x : __builtin_inff();
    return e;
}
__device__ __forceinline__ f2 etrans(f2 e, f2 p) {
    float d = dot2(e, p);
    return mk2(fma_(d, e.x, p.x), fma_(d, e.y, p.y));
}
__device__ __forceinline__ float edet(f2 e) { return fma_(e.x, e.x, fma_(e.y, e.y, 1.0f)); }
template <bool X>
__device__ __forceinline__ float ersd(f2 e) {
    return rsq_<X>(edet(e));
}
__device__ __forceinline__ float edfr(f2 e, f2 d) {
    float ed = dot2(e, d);
    float dd = dot2(d, d);
    return fma_(ed, ed, dd);
}
template <bool X>
__device__ __forceinline__ float endf(f2 e, f2 nd) {
    float ed = dot2(e, nd);
    return rsq_<X>(fma_(ed, ed, 1.0f));
}

template <bool X>
__device__ __forceinline__ float positive_atan(float t, bool fast_atan) {
    if (fast_atan) {
        float rx = (fabs_(t) > 1.0f) ? rcp_<X>(fabs_(t)) : fabs_(t);
        float ry = rx * rx;
        float rz = fma_(ry, 0.02083509974181652f, -0.08513300120830536f);
        rz = fma_(ry, rz, 0.18014100193977356f);
        rz = fma_(ry, rz, -0.3302994966506958f);
        ry = fma_(ry, rz, 0.9998660087585449f);
        rz = fma_(-2.0f * ry, rx, 1.5707963267948966192313216916398f);
        rz = (fabs_(t) > 1.0f) ? rz : 0.0f;
        rx = fma_(rx, ry, rz);
        return (t < 0.0f) ? (3.1415926535897932384626433832795f - rx) : rx;
    }
    float off = (t < 0.0f) ? 3.1415926535897932384626433832795f : 0.0f;
    return atan_<X>(t) + off;
}

template <bool X>
__device__ __forceinline__ float area_from_tangents(float ird, float it, float ord, float ot, bool fa) {
    float inner = ird * positive_atan<X>(it, fa);
    float r = fma_(ord, positive_atan<X>(ot, fa), -inner);
    return (r > 0.0f) ? (0.5f * r) : 0.0f;
}
template <bool X>
__device__ __forceinline__ float area_between(f2 ie, float ird, f2 oe, float ord, f2 d0, f2 d1, bool fa) {
    float dd = fmax_(0.0f, dot2(d1, rot90(d0)));
    float idot = ird * dot2(d0, etrans(ie, d1));
    float odot = ord * dot2(d0, etrans(oe, d1));
    return area_from_tangents<X>(ird, div_<X>(dd, idot), ord, div_<X>(dd, odot), fa);
}
template <bool X>
__device__ __forceinline__ float ellipse_area_in_sector(f2 e, f2 d0, f2 d1, bool fa) {
    float rd = ersd<X>(e);
    float dd = fmax_(0.0f, dot2(d1, rot90(d0)));
    float edot = rd * dot2(d0, etrans(e, d1));
    float area = 0.5f * rd * positive_atan<X>(div_<X>(dd, edot), fa);
    return (rd > 0.0f) ? area : 0.0f;
}

__device__ __forceinline__ void cas(Psa& P, int l, int r) {
    f2 lc = P.v[l];
    float nz = kahan(lc.x, -P.v[r].y, lc.y, -P.v[r].x);
    bool swap = (nz == 0.0f) ? __builtin_isinf(P.e[r].x) : (nz > 0.0f);
    P.v[l] = swap ? P.v[r] : lc;
    P.v[r] = swap ? lc : P.v[r];
    f2 le = P.e[l];
    P.e[l] = swap ? P.e[r] : le;
    P.e[r] = swap ? le : P.e[r];
}

// sort_convex_polygon_vertices (polygon_sampling.glsl:441-506); 9 vertices
// (clipped 8-gons) use a complete 25-comparator network
__device__ __forceinline__ void sort_vertices(Psa& P) {
    uint32_t n = P.vc;
    if (n == 3) {
        cas(P, 1, 2);
    } else if (n == 4) {
        cas(P, 1, 3);
    } else if (n == 5) {
        cas(P, 2, 4); cas(P, 1, 3); cas(P, 1, 2); cas(P, 0, 3); cas(P, 3, 4);
    } else if (n == 6) {
        cas(P, 3, 5); cas(P, 2, 4); cas(P, 1, 5); cas(P, 0, 4); cas(P, 4, 5); cas(P, 1, 3);
    } else if (n == 7) {
        cas(P, 2, 5); cas(P, 1, 6); cas(P, 5, 6); cas(P, 3, 4); cas(P, 0, 4); cas(P, 4, 6);
        cas(P, 1, 3); cas(P, 3, 5); cas(P, 4, 5);
    } else if (n == 8) {
        cas(P, 2, 6); cas(P, 3, 7); cas(P, 1, 5); cas(P, 0, 4); cas(P, 4, 6); cas(P, 5, 7);
        cas(P, 6, 7); cas(P, 4, 5); cas(P, 1, 3);
    } else if (n == 9) {
        cas(P, 0, 1); cas(P, 3, 4); cas(P, 6, 7); cas(P, 1, 2); cas(P, 4, 5); cas(P, 7, 8);
        cas(P, 0, 1); cas(P, 3, 4); cas(P, 6, 7); cas(P, 0, 3); cas(P, 3, 6); cas(P, 0, 3);
        cas(P, 1, 4); cas(P, 4, 7); cas(P, 1, 4); cas(P, 2, 5); cas(P, 5, 8); cas(P, 2, 5);
        cas(P, 1, 3); cas(P, 5, 7); cas(P, 2, 6); cas(P, 4, 6); cas(P, 2, 4); cas(P, 2, 3);
        cas(P, 5, 6);
    }
    cas(P, 0, 2);
    if (n >= 4) cas(P, 2, 3);
    cas(P, 0, 1);
}

// prepare_projected_solid_angle_polygon_sampling (polygon_sampling.glsl:545-613)
template <bool X>
__device__ __forceinline__ void prepare_psa(Psa& P, uint32_t vc, const float* vx, const float* vy, const float* vz,
                                            bool fa) {
    P.vc = vc;
    P.inner0 = mk2(1.0f, 0.0f);
    P.v[0] = mk2(vx[0], vy[0]);
    P.e[0] = ellipse_from_edge<X>(mk3(vx[0], vy[0], vz[0]), mk3(vx[1], vy[1], vz[1]));
    f2 prev = P.e[0];
#pragma unroll
    for (int i = 1; i < MAXV; ++i) {
        P.v[i] = mk2(vx[i], vy[i]);
        P.e[i] = mk2(0.0f, 0.0f);
        P.sector[i] = 0.0f;
    }
#pragma unroll
    for (int i = 1; i < MAXV; ++i) {
        if (!(i > 2 && (uint32_t)i >= vc)) {
            int j = (i + 1) % MAXV;
            f2 e = ellipse_from_edge<X>(mk3(vx[i], vy[i], vz[i]), mk3(vx[j], vy[j], vz[j]));
            bool in = inner_e(e);
            P.e[i] = in ? prev : e;
            P.inner0 = (inner_e(prev) && !in) ? prev : P.inner0;
            prev = e;
        }
    }
    {
        f2 e = P.e[0];
        bool in = inner_e(e);
        P.e[0] = in ? prev : e;
        P.inner0 = (inner_e(prev) && !in) ? prev : P.inner0;
    }
    P.psa = 0.0f;
    P.sector[0] = 0.0f;
    if (P.inner0.x > 0.0f) {
#pragma unroll
        for (int i = 0; i < MAXV; ++i) {
            if (!(i > 2 && (uint32_t)i >= vc)) {
                P.sector[i] = ellipse_area_in_sector<X>(P.e[i], P.v[i], P.v[(i + 1) % MAXV], fa);
                P.psa += P.sector[i];
            }
        }
    } else {
        sort_vertices(P);
        f2 ie = P.inner0;
        float ird = ersd<X>(ie);
        f2 oe = mk2(0.0f, 0.0f);
        float ord = 0.0f;
#pragma unroll
        for (int i = 0; i < MAXV - 1; ++i) {
            if (!(i > 1 && (uint32_t)i + 1u >= vc)) {
                f2 ve = P.e[i];
                bool vin = inner_e(ve);
                float vrd = ersd<X>(ve);
                if (i == 0) {
                    oe = ve;
                    ord = vrd;
                } else {
                    ie = vin ? ve : ie;
                    ird = vin ? vrd : ird;
                    oe = vin ? oe : ve;
                    ord = vin ? ord : vrd;
                }
                P.sector[i] = area_between<X>(ie, ird, oe, ord, P.v[i], P.v[i + 1], fa);
                P.psa += P.sector[i];
            }
        }
    }
}

template <bool X>
__device__ __forceinline__ f2 nrm_approx_flip(f2 rhs, f2 semi) {
    float s = fabs_(rhs.x) + fabs_(rhs.y);
    s = __uint_as_float(__float_as_uint(s) ^ 0x7F800000u);
    s = (dot2(rhs, semi) >= 0.0f) ? s : -s;
    return sclf2(s, rhs);
}

struct Q2 {
    float q00, q01, q10, q11;  // column-major: q[col][row]
};
__device__ __forceinline__ Q2 outer2(f2 c, f2 r) { return Q2{c.x * r.x, c.y * r.x, c.x * r.y, c.y * r.y}; }
__device__ __forceinline__ Q2 subq(Q2 a, Q2 b) { return Q2{a.q00 - b.q00, a.q01 - b.q01, a.q10 - b.q10, a.q11 - b.q11}; }
template <bool X>
__device__ __forceinline__ f2 solve_quadratic(Q2 m) {
    float cxy = 0.5f * (m.q01 + m.q10);
    float sd = sqrt_<X>(fmax_(0.0f, cxy * cxy - m.q00 * m.q11));
    float root = fabs_(cxy) + sd;
    return (cxy >= 0.0f) ? mk2(root, -m.q00) : mk2(m.q11, root);
}
__device__ __forceinline__ float mix_fma(float x, float y, float a) { return fma_(a, y, fma_(-a, x, x)); }

// sample_sector_between_ellipses (polygon_sampling.glsl:668-762)
template <bool X>
__device__ __forceinline__ f2 sample_sector(f2 rn, float target_area, f2 ie, f2 oe, f2 d0, f2 d1, bool fa) {
    f2 q0 = normalize2<X>(d0);
    f2 q2 = normalize2<X>(d1);
    f2 q1 = add2(q0, q2);
    float n00 = endf<X>(ie, q0), n01 = rsq_<X>(edfr(ie, q1)), n02 = endf<X>(ie, q2);
    float n10 = endf<X>(oe, q0), n11 = rsq_<X>(edfr(oe, q1)), n12 = endf<X>(oe, q2);
    float sa0 = n10 * n11 - n00 * n01;
    float sa1 = n11 * n12 - n01 * n02;
    float tqa = mix_fma(-sa0, sa1, rn.x);
    bool sel = tqa <= 0.0f;
    q2 = sel ? q0 : q2;
    n02 = sel ? n00 : n02;
    n12 = sel ? n10 : n12;
    tqa += sel ? sa0 : -sa1;
    tqa *= fabs_(q1.x * q2.y - q2.x * q1.y);
    f2 qn0 = add2(scl2(q1, n01), scl2(q2, n02));
    f2 qn1 = add2(scl2(q1, n11), scl2(q2, n12));
    qn0 = etrans(ie, qn0);
    qn1 = etrans(oe, qn1);
    float qo0 = dot2(qn0, q1) * n01;
    float qo1 = dot2(qn1, q1) * n11;
    f2 r2 = rot90(q2);
    Q2 quad = outer2(sclf2(qo1 * n12, r2), qn0);
    quad = subq(quad, outer2(add2(sclf2(qo0 * n02, r2), sclf2(tqa, qn0)), qn1));
    f2 cur = solve_quadratic<X>(quad);
    uint32_t iters = (fabs_(rn.x - 0.5f) <= 0.5f - 1.0e-5f) ? 2u : 0u;
    float ird = ersd<X>(ie);
    float ord = ersd<X>(oe);
    for (uint32_t it = 0; it < iters; ++it) {
        cur = nrm_approx_flip<X>(cur, q1);
        f2 idir = etrans(ie, cur);
        f2 odir = etrans(oe, cur);
        float dd = fmax_(0.0f, dot2(cur, rot90(q0)));
        float err = target_area - area_from_tangents<X>(ird, div_<X>(dd, ird * dot2(q0, idir)), ord,
                                                         div_<X>(dd, ord * dot2(q0, odir)), fa);
        quad = subq(outer2(sub2(idir, odir), rot90(cur)), outer2(sclf2(2.0f * err, idir), odir));
        cur = solve_quadratic<X>(quad);
    }
    cur = (dot2(cur, q1) >= 0.0f) ? cur : mk2(-cur.x, -cur.y);
    float inner_f = rcp_<X>(edfr(ie, cur));
    float outer_f = rcp_<X>(edfr(oe, cur));
    return scl2(cur, sqrt_<X>(mix_fma(inner_f, outer_f, rn.y)));
}

// sample_projected_solid_angle_polygon (polygon_sampling.glsl:772-828)
template <bool X>
__device__ __forceinline__ f3 sample_psa(const Psa& P, f2 rn, bool fa) {
    float target = rn.x * P.psa;
    f2 xy;
    if (P.inner0.x > 0.0f) {
        f2 oe = P.e[0], d0 = P.v[0];
        bool done = false;
#pragma unroll
        for (int i = 0; i < MAXV; ++i) {
            if (!done) {
                if (i > 0) target -= P.sector[i - 1];
                oe = P.e[i];
                d0 = P.v[i];
                if ((i >= 2 && (uint32_t)i + 1u == P.vc) || target < P.sector[i]) done = true;
            }
        }
        float sdet = sqrt_<X>(edet(oe));
        float angle = 2.0f * target * sdet;
        f2 a = sclf2(cos_(angle) * sdet, d0);
        f2 b = sclf2(sin_(angle), rot90(etrans(oe, d0)));
        xy = add2(a, b);
        xy = scl2(xy, sqrt_<X>(div_<X>(rn.y, edfr(oe, xy))));
    } else {
        f2 ie = P.inner0, oe = P.e[0], d0 = P.v[0], d1 = P.v[1];
        float spsa = P.sector[0];
        bool done = false;
#pragma unroll
        for (int i = 0; i < MAXV - 1; ++i) {
            if (!done) {
                f2 ve = P.e[i];
                if (i == 0) {
                    oe = ve;
                } else {
                    target -= P.sector[i - 1];
                    bool vin = inner_e(ve);
                    ie = vin ? ve : ie;
                    oe = vin ? oe : ve;
                }
                d0 = P.v[i];
                d1 = P.v[i + 1];
                spsa = P.sector[i];
                if ((i >= 1 && (uint32_t)i + 2u == P.vc) || target < spsa) done = true;
            }
        }
        f2 r = mk2(div_<X>(target, spsa), rn.y);
        xy = sample_sector<X>(r, target, ie, oe, d0, d1, fa);
    }
    float z = sqrt_<X>(fmax_(0.0f, fma_(-xy.x, xy.x, fma_(-xy.y, xy.y, 1.0f))));
    return mk3(xy.x, xy.y, z);
}

// evaluate_brdf (brdfs.glsl:58-93), both components
template <bool X>
__device__ __forceinline__ f3 evaluate_brdf(const Pixel& px, f3 li_dir) {
    f3 h = normalize3<X>(add3(li_dir, px.o));
    float li = dot3(px.n, li_dir);
    float odh = dot3(px.o, h);
    f3 brdf = add3(mk3(0.0f, 0.0f, 0.0f), px.albedo);
    float ndh = dot3(px.n, h);
    float r2 = px.rough * px.rough;
    float ggx = fma_(fma_(ndh, r2, -ndh), ndh, 1.0f);
    ggx = div_<X>(r2, ggx * ggx);
    float lo = px.lo;
    float masking = li * sqrt_<X>(fma_(fma_(-lo, r2, lo), lo, r2));
    float shadowing = lo * sqrt_<X>(fma_(fma_(-li, r2, li), li, r2));
    float smith = div_<X>(0.5f, masking + shadowing);
    float c = clamp_(odh, 0.0f, 1.0f);
    float fl = 1.0f - c;
    float fs = fl * fl;
    float f5 = fs * fl * fs;
    f3 F = mk3(px.f0.x + (1.0f - px.f0.x) * f5, px.f0.y + (1.0f - px.f0.y) * f5, px.f0.z + (1.0f - px.f0.z) * f5);
    float spec = ggx * smith * dot3(F, mk3(0.21263901f, 0.71516868f, 0.07219232f));
    brdf = add3(brdf, mk3(spec, spec, spec));
    return scl3(brdf, 0.31830988618379067153776752674503f);
}

// get_mis_estimate (shading_pass.frag.glsl:219-269)
template <bool X>
__device__ __forceinline__ f3 mis_estimate(uint32_t h, f3 integ, f3 sw, float sd, f3 ow, float od, float ve) {
    if (h == RISLTC_MIS_WEIGHTED) {
        f3 ws = add3(scl3(sw, sd), scl3(ow, od));
        f3 num = mul3(sw, integ);
        return mk3(div_<X>(num.x, ws.x), div_<X>(num.y, ws.y), div_<X>(num.z, ws.z));
    }
    if (h == RISLTC_MIS_OPTIMAL_CLAMPED || h == RISLTC_MIS_OPTIMAL) {
        float bw = rcp_<X>(sd + od);
        f3 ws = add3(scl3(sw, sd), scl3(ow, od));
        if (h == RISLTC_MIS_OPTIMAL_CLAMPED) {
            f3 wwd = mk3(div_<X>(sw.x, ws.x), div_<X>(sw.y, ws.y), div_<X>(sw.z, ws.z));
            float m0 = fma_(-ve, bw, bw);
            f3 mixed = mk3(fma_(ve, wwd.x, m0), fma_(ve, wwd.y, m0), fma_(ve, wwd.z, m0));
            return mul3(mixed, integ);
        }
        f3 a = scl3(sw, ve);
        f3 d = sub3(integ, scl3(ws, ve));
        return add3(a, sclf3(bw, d));
    }
    float wod = (h == RISLTC_MIS_BALANCE) ? rcp_<X>(sd + od) : div_<X>(sd, sd * sd + od * od);
    return sclf3(wod, integ);
}

__device__ __forceinline__ float ltc_fetch(const uint16_t* q, int C, int ch, uint32_t W, uint32_t H, int layer,
                                           float u, float v) {
    float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    float x0 = __builtin_floorf(x), y0 = __builtin_floorf(y);
    float fx = x - x0, fy = y - y0;
    int ix = (int)x0, iy = (int)y0;
    int wi = (int)W - 1, hi = (int)H - 1;
    int i0 = ix < 0 ? 0 : (ix > wi ? wi : ix), i1 = ix + 1 < 0 ? 0 : (ix + 1 > wi ? wi : ix + 1);
    int j0 = iy < 0 ? 0 : (iy > hi ? hi : iy), j1 = iy + 1 < 0 ? 0 : (iy + 1 > hi ? hi : iy + 1);
    size_t base = (size_t)layer * H;
    float t00 = (float)q[((base + j0) * W + i0) * C + ch] / 65535.0f;
    float t10 = (float)q[((base + j0) * W + i1) * C + ch] / 65535.0f;
    float t01 = (float)q[((base + j1) * W + i0) * C + ch] / 65535.0f;
    float t11 = (float)q[((base + j1) * W + i1) * C + ch] / 65535.0f;
    float a = fma_(fx, t10 - t00, t00);
    float b = fma_(fx, t11 - t01, t01);
    return fma_(fy, b - a, a);
}

__device__ __forceinline__ void setup_pixel(const ShadeArgs& A, uint32_t prim, f3 ray, Pixel& px) {
    const SceneView& S = A.scene;
    const risltc_frame_constants_t& c = A.c;
    f3 p[3], nrm[3];
    float tc[3][2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        size_t vi = (size_t)prim * 3 + i;
        p[i] = decode_position(S.positions[vi], c.mesh_dequantization_factor, c.mesh_dequantization_summand);
        ushort4 nt = S.normals_uv[vi];
        nrm[i] = decode_normal(nt.x, nt.y);
        tc[i][0] = fma_((float)nt.z / 65535.0f, 8.0f, 0.0f);
        tc[i][1] = fma_((float)nt.w / 65535.0f, -8.0f, 1.0f);
    }
    f3 org = mk3(c.camera_position_world_space[0], c.camera_position_world_space[1], c.camera_position_world_space[2]);
    f3 e0 = sub3(p[1], p[0]), e1 = sub3(p[2], p[0]);
    f3 rce1 = cross3(ray, e1);
    float rcp_det = 1.0f / dot3(e0, rce1);
    f3 r0 = sub3(org, p[0]);
    float d0 = dot3(r0, rce1);
    float by = rcp_det * d0;
    f3 e0c0 = cross3(e0, r0);
    float d1 = dot3(ray, e0c0);
    float bz = -rcp_det * d1;
    float bx = 1.0f - (by + bz);
    px.pos = mk3(fma_(bx, p[0].x, fma_(by, p[1].x, bz * p[2].x)), fma_(bx, p[0].y, fma_(by, p[1].y, bz * p[2].y)),
                 fma_(bx, p[0].z, fma_(by, p[1].z, bz * p[2].z)));
    f3 ni = normalize3<true>(mk3(fma_(bx, nrm[0].x, fma_(by, nrm[1].x, bz * nrm[2].x)),
                                 fma_(bx, nrm[0].y, fma_(by, nrm[1].y, bz * nrm[2].y)),
                                 fma_(bx, nrm[0].z, fma_(by, nrm[1].z, bz * nrm[2].z))));
    uint32_t mi = S.material_idx[prim];
    if (mi >= S.material_count) mi = S.material_count - 1u;
    const float* m = S.materials + 6 * mi;
    f3 base = mk3(m[0], m[1], m[2]);
    float metal = m[5];
    px.albedo = mk3(fma_(base.x, -metal, base.x), fma_(base.y, -metal, base.y), fma_(base.z, -metal, base.z));
    float om = 1.0f - metal;
    px.f0 = mk3(0.02f * om + base.x * metal, 0.02f * om + base.y * metal, 0.02f * om + base.z * metal);
    float lr = m[4];
    px.rough = clamp_(lr * lr * c.roughness_factor, 0.0064f, 1.0f);
    float nts_x = fma_(0.5f, 2.0f, -1.0f), nts_y = nts_x;
    float nts_z = __builtin_sqrtf(fmax_(0.0f, fma_(-nts_x, nts_x, fma_(-nts_y, nts_y, 1.0f))));
    float te0x = tc[1][0] - tc[0][0], te0y = tc[1][1] - tc[0][1];
    float te1x = tc[2][0] - tc[0][0], te1y = tc[2][1] - tc[0][1];
    f3 nce0 = cross3(ni, e0);
    f3 e1cn = cross3(e1, ni);
    f3 tangent = add3(scl3(e1cn, te0x), scl3(nce0, te1x));
    f3 bitangent = add3(scl3(e1cn, te0y), scl3(nce0, te1y));
    float mtl = __builtin_sqrtf(0.5f * (dot3(tangent, tangent) + dot3(bitangent, bitangent)));
    nts_z *= fmax_(1.0e-10f, mtl);
    f3 nw = mk3(fma_(ni.x, nts_z, fma_(bitangent.x, nts_y, tangent.x * nts_x)),
                fma_(ni.y, nts_z, fma_(bitangent.y, nts_y, tangent.y * nts_x)),
                fma_(ni.z, nts_z, fma_(bitangent.z, nts_y, tangent.z * nts_x)));
    px.n = normalize3<true>(nw);
    px.o = normalize3<true>(sub3(org, px.pos));
    float off = fmax_(0.0f, 1.0e-3f - dot3(px.n, px.o));
    px.n = mk3(fma_(off, px.o.x, px.n.x), fma_(off, px.o.y, px.n.y), fma_(off, px.o.z, px.n.z));
    px.n = normalize3<true>(px.n);
    px.lo = dot3(px.n, px.o);
    // get_ltc_coefficients (ltc_utility.glsl:56-88)
    float f0l = dot3(px.f0, mk3(0.2126f, 0.7152f, 0.0722f));
    const risltc_ltc_constants_t& k = c.ltc_constants;
    float ndo = dot3(px.n, px.o);
    float incl = acos_(clamp_(ndo, 0.0f, 1.0f));
    float u = fma_(__builtin_sqrtf(clamp_(px.rough, 0.0f, 1.0f)), k.roughness_factor, k.roughness_summand);
    float v = fma_(incl, k.inclination_factor, k.inclination_summand);
    float w = fma_(clamp_(f0l, 0.0f, 1.0f), k.fresnel_index_factor, k.fresnel_index_summand);
    float lw = __builtin_rintf(w);
    int layer = lw < 0.0f ? 0 : (lw > (float)(S.ltc_layers - 1) ? (int)S.ltc_layers - 1 : (int)lw);
    float d0x = ltc_fetch(S.ltc0, 4, 0, S.ltc_w, S.ltc_h, layer, u, v);
    float d0y = ltc_fetch(S.ltc0, 4, 1, S.ltc_w, S.ltc_h, layer, u, v);
    float d0z = ltc_fetch(S.ltc0, 4, 2, S.ltc_w, S.ltc_h, layer, u, v);
    float d0w = ltc_fetch(S.ltc0, 4, 3, S.ltc_w, S.ltc_h, layer, u, v);
    float d1x = ltc_fetch(S.ltc1, 2, 0, S.ltc_w, S.ltc_h, layer, u, v);
    float d1y = ltc_fetch(S.ltc1, 2, 1, S.ltc_w, S.ltc_h, layer, u, v);
    px.s2c[0] = d0x; px.s2c[1] = 0.0f; px.s2c[2] = d0w;
    px.s2c[3] = 0.0f; px.s2c[4] = d0z; px.s2c[5] = 0.0f;
    px.s2c[6] = -d0y; px.s2c[7] = 0.0f; px.s2c[8] = d1x;
    px.ltc_albedo = d1y;
    float det2 = d0x * d1x + d0y * d0w;
    px.ltc_det = d0z * det2;
    f3 xa = normalize3<true>(mk3(fma_(-ndo, px.n.x, px.o.x), fma_(-ndo, px.n.y, px.o.y), fma_(-ndo, px.n.z, px.o.z)));
    f3 ya = cross3(px.n, xa);
    f3 rows[3] = {xa, ya, px.n};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        px.w2s[4 * i + 0] = rows[i].x;
        px.w2s[4 * i + 1] = rows[i].y;
        px.w2s[4 * i + 2] = rows[i].z;
        px.w2s[4 * i + 3] = fma_(-rows[i].z, px.pos.z, fma_(-rows[i].y, px.pos.y, -rows[i].x * px.pos.x));
    }
}

__device__ __forceinline__ uint32_t light_index_from(float u, uint32_t n) {
    int idx = (int)(u * (float)n);
    if (idx > (int)n - 1) idx = (int)n - 1;
    return (uint32_t)idx;
}

// light_index_from(noise_next(seed), n) with kn = n * 2^-32: the noise value
// is (float)seed * 2^-32 exactly (a power of two, no underflow), so
// (float)seed * kn rounds the same real number as u * n; one product fewer.
__device__ __forceinline__ uint32_t light_index_next(uint32_t& seed, float kn, uint32_t n) {
    seed = 1664525u * seed + 1013904223u;
    int idx = (int)((float)seed * kn);
    if (idx > (int)n - 1) idx = (int)n - 1;
    return (uint32_t)idx;
}

// evaluate_polygonal_light_shading, LTC_CP branch: the RIS target
template <bool X>
__device__ __forceinline__ f3 eval_ltc_cp(const Pixel& px, const Light& L, uint32_t& clip_counts) {
    // The mirrored polygon of a shading point behind the light (y negated)
    // has the same clip masks and negated edge terms, so the same |sum|:
    // bit for bit, since negation is exact and rounding is symmetric. The
    // exact build mirrors as the reference does (its cosine-space matrix is
    // composed with the mirror); the fast build skips it.
    const bool flip = X && plane_side(px.pos, L.plane) < 0.0f;
    float vx[9], vy[9], vz[9];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        f3 t = xform(px.w2s, L.v[3 * i], L.v[3 * i + 1], L.v[3 * i + 2]);
        vx[i] = t.x;
        vy[i] = flip ? -t.y : t.y;
        vz[i] = t.z;
    }
    vx[8] = vx[0];
    vy[8] = vy[0];
    vz[8] = vz[0];
    uint32_t vcd, vcs;
    float ltc_d = polygon_ltc<X>(vx, vy, vz, L.n, vcd);
    f3 diffuse = mk3(0.0f, 0.0f, 0.0f);
    if (vcd > 0) diffuse = mul3(sclf3(ltc_d, px.albedo), L.rad);
    f3 result = add3(mk3(0.0f, 0.0f, 0.0f), diffuse);
    if constexpr (X) {
        // the reference's composed world-to-cosine matrix, as the oracle
        float m[12];
        world_to_cosine(px, flip, m);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            f3 t = xform(m, L.v[3 * i], L.v[3 * i + 1], L.v[3 * i + 2]);
            vx[i] = t.x;
            vy[i] = t.y;
            vz[i] = t.z;
        }
        vx[8] = vx[0];
        vy[8] = vy[0];
        vz[8] = vz[0];
    } else {
        // cosine space = shading_to_cosine * (the mirrored shading-space
        // polygon above); the LTC matrix has only five non-zero entries
        const float a = px.s2c[0], b = px.s2c[2], c = px.s2c[4], d = px.s2c[6], e = px.s2c[8];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            float x = vx[i], z = vz[i];
            vx[i] = fma_(a, x, b * z);
            vy[i] = c * vy[i];
            vz[i] = fma_(d, x, e * z);
        }
    }
    float ltc_s = polygon_ltc<X>(vx, vy, vz, L.n, vcs);
    f3 ggx = mk3(0.0f, 0.0f, 0.0f);
    if (vcs > 0) ggx = sclf3(ltc_s * px.ltc_albedo, L.rad);
    result = add3(result, ggx);
    clip_counts = vcd | (vcs << 8);
    return result;
}

struct FinalOut {
    f3 color;
    f3 e_vis[2], e_occ[2];
    f3 dir[2];
    float tmax[2];
    uint32_t state[2];
    uint32_t clipd, clips;
    float psad, psas;
};

// evaluate_polygonal_light_shading_peters (shading_pass.frag.glsl:292-397),
// one iteration of its SAMPLE_COUNT loop (365-396): the polygons are prepared
// and one direction per technique is drawn. Calling it k times reproduces the
// loop, since the preparation draws nothing and gives the same polygons; the
// caller scales the sum by 1 / k. One projected-solid-angle polygon is live
// at a time (diffuse prepared and sampled, then specular), which consumes the
// noise in the reference order.
template <bool X, bool LDS = false>
__device__ __forceinline__ void eval_peters(const ShadeArgs& A, const Pixel& px, const Light& L, uint32_t& seed, FinalOut& F,
                                            float* col = nullptr) {
    bool fa = A.fast_atan != 0;
    F.color = mk3(0.0f, 0.0f, 0.0f);
    F.state[0] = F.state[1] = 0u;
    F.clipd = F.clips = 0u;
    F.psad = F.psas = 0.0f;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        F.e_vis[j] = F.e_occ[j] = F.dir[j] = mk3(0.0f, 0.0f, 0.0f);
        F.tmax[j] = 0.0f;
    }
    float side = plane_side(px.pos, L.plane);
    bool flip = side < 0.0f;
    float vx[MAXV], vy[MAXV], vz[MAXV];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        f3 t = xform(px.w2s, L.v[3 * i], L.v[3 * i + 1], L.v[3 * i + 2]);
        vx[i] = t.x;
        vy[i] = flip ? -t.y : t.y;
        vz[i] = t.z;
    }
    vx[8] = vx[0];
    vy[8] = vy[0];
    vz[8] = vz[0];
    uint32_t vcd = clip_materialize<X, LDS>(vx, vy, vz, L.n, col);
    F.clipd = vcd;
    if (vcd == 0) return;
    Psa P;
    prepare_psa<X>(P, vcd, vx, vy, vz, fa);
    F.psad = P.psa;
    float psad = P.psa;
    // the diffuse direction first (its two draws precede the specular ones),
    // so that the specular vertices are not live while it is sampled
    f3 dirs[2];
    dirs[0] = mk3(0.0f, 0.0f, 0.0f);
    if (psad != 0.0f) {
        float u0 = noise_next(seed);
        float u1 = noise_next(seed);
        dirs[0] = sample_psa<X>(P, mk2(u0, u1), fa);
    }
    // transform for the specular polygon now (the diffuse vertices are dead)
    float m[12];
    world_to_cosine(px, flip, m);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        f3 t = xform(m, L.v[3 * i], L.v[3 * i + 1], L.v[3 * i + 2]);
        vx[i] = t.x;
        vy[i] = t.y;
        vz[i] = t.z;
    }
    vx[8] = vx[0];
    vy[8] = vy[0];
    vz[8] = vz[0];
    uint32_t vcs = clip_materialize<X, LDS>(vx, vy, vz, L.n, col);
    F.clips = vcs;
    if (psad == 0.0f) {
        // the reference still prepares the specular polygon; only its
        // projected solid angle is reported
        if (vcs != 0) {
            Psa S;
            prepare_psa<X>(S, vcs, vx, vy, vz, fa);
            F.psas = S.psa;
        }
        return;
    }
    float psas = 0.0f;
    dirs[1] = mk3(0.0f, 0.0f, 0.0f);
    if (vcs != 0) {
        prepare_psa<X>(P, vcs, vx, vy, vz, fa);
        psas = P.psa;
        if (psas > 0.0f) {
            float u0 = noise_next(seed);
            float u1 = noise_next(seed);
            dirs[1] = sample_psa<X>(P, mk2(u0, u1), fa);
            float c2s[9];
            cosine_to_shading(px, c2s);
            dirs[1] = normalize3<X>(mul33(c2s, dirs[1]));
        }
    }
    F.psas = psas;
    float spec_w = px.ltc_albedo * psas;
    f3 da = mk3(fmax_(px.albedo.x, 0.01f), fmax_(px.albedo.y, 0.01f), fmax_(px.albedo.z, 0.01f));
    f3 dw = scl3(da, psad);
    uint32_t techniques = (psas > 0.0f) ? 2u : 1u;
    float rcp_d = rcp_<X>(psad);
    float rcp_s = rcp_<X>(psas);
    f3 sw = mk3(spec_w, spec_w, spec_w);
    uint32_t h = A.mis;
    if (h == RISLTC_MIS_OPTIMAL) {
        f3 rop = scl3(L.rad, 0.31830988618379067153776752674503f);
        dw = mul3(dw, rop);
        sw = mul3(sw, rop);
    }
    float ve = A.c.mis_visibility_estimate;
    float w2s_r[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) w2s_r[k] = (flip && k >= 4 && k < 8) ? -px.w2s[k] : px.w2s[k];
    f3 pn = mk3(L.plane[0], L.plane[1], L.plane[2]);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if ((uint32_t)j >= techniques) continue;
        f3 d = dirs[j];
        if (d.z <= 0.0f) continue;
        float dd = d.z * rcp_d;
        f3 dc = mul33(px.s2c, d);
        float l2 = dot3(dc, dc);
        float sdn = div_<X>(fmax_(0.0f, dc.z) * px.ltc_det, l2 * l2) * rcp_s;
        f3 wdir = mk3(fma_(w2s_r[8], d.z, fma_(w2s_r[4], d.y, w2s_r[0] * d.x)),
                      fma_(w2s_r[9], d.z, fma_(w2s_r[5], d.y, w2s_r[1] * d.x)),
                      fma_(w2s_r[10], d.z, fma_(w2s_r[6], d.y, w2s_r[2] * d.x)));
        bool lambert_ok = dot3(px.n, wdir) > 0.0f;
        uint32_t state = 1u;
        float tmax = 0.0f;
        if (lambert_ok) {
            tmax = div_<X>(-plane_side(px.pos, L.plane), dot3(wdir, pn));
            tmax -= 1e-3f;
            state = A.shadow_rays ? 5u : 4u;  // 5 = traced, outcome pending
        }
        f3 rb = mul3(L.rad, evaluate_brdf<X>(px, wdir));
        f3 iv = sclf3(d.z, rb);
        f3 io = sclf3(d.z, mk3(0.0f, 0.0f, 0.0f));
        f3 ev, eo;
        if (j == 0 && psas <= 0.0f) {
            float rd = rcp_<X>(dd);
            ev = scl3(iv, rd);
            eo = mk3(0.0f, 0.0f, 0.0f);
        } else if (j == 0) {
            ev = mis_estimate<X>(h, iv, dw, dd, sw, sdn, ve);
            eo = mis_estimate<X>(h, io, dw, dd, sw, sdn, ve);
        } else {
            ev = mis_estimate<X>(h, iv, sw, sdn, dw, dd, ve);
            eo = mis_estimate<X>(h, io, sw, sdn, dw, dd, ve);
        }
        F.e_vis[j] = ev;
        F.e_occ[j] = eo;
        F.dir[j] = wdir;
        F.tmax[j] = tmax;
        F.state[j] = state;
    }
}

__device__ __forceinline__ void clear_final(FinalOut& F) {
    F.color = mk3(0.0f, 0.0f, 0.0f);
    F.state[0] = F.state[1] = 0u;
    F.clipd = F.clips = 0u;
    F.psad = F.psas = 0.0f;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        F.e_vis[t] = F.e_occ[t] = F.dir[t] = mk3(0.0f, 0.0f, 0.0f);
        F.tmax[t] = 0.0f;
    }
}

// One iteration of a pixel-sample's final sample: its shadow rays are
// appended to the compacted list (one atomic per wave and technique; every
// lane of the wave calls this) and its record is written at index `rec_idx`
// (pixel-sample * SAMPLE_COUNT + iteration).
// COMPACT: the caller may be asked for compact records (LTC_CP kernels only)
template <bool COMPACT>
__device__ __forceinline__ void put_final(const ShadeArgs& A, uint32_t lane, bool inside, bool contributes,
                                          const FinalOut& F, f3 pos, size_t rec_idx, float Wt, uint32_t extra,
                                          int32_t chosen, uint32_t ckey) {
    // a lane's two rays (the same light, the same origin) go next to each
    // other in the list, so they are traced by neighbouring lanes
    const bool emit0 = contributes && F.state[0] == 5u, emit1 = contributes && F.state[1] == 5u;
    const uint64_t b0 = __ballot(emit0), b1 = __ballot(emit1);
    const uint32_t cnt = (uint32_t)(__popcll(b0) + __popcll(b1));
    if (cnt) {
        const uint64_t lanes = b0 | b1;
        const uint32_t first = (uint32_t)__ffsll((unsigned long long)lanes) - 1u;
        uint32_t base = 0;
        if (lane == first) base = atomicAdd(A.ray_count, cnt);
        base = (uint32_t)__shfl((int)base, (int)first);
        const uint64_t below = (1ull << lane) - 1ull;
        uint32_t off = base + (uint32_t)(__popcll(b0 & below) + __popcll(b1 & below));
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            if (t == 0 ? emit0 : emit1) {
                A.rays[off].org = make_float4(pos.x, pos.y, pos.z, __uint_as_float(A.bvh_index));
                A.rays[off].dir = make_float4(F.dir[t].x, F.dir[t].y, F.dir[t].z, F.tmax[t]);
                A.ray_slots[off] = A.slot_base + (uint32_t)(rec_idx * 2 + t);
                ++off;
            }
        }
    }
    if (inside) {
        float4* rec = reinterpret_cast<float4*>(A.samples + rec_idx * A.rec_words);
        uint32_t flags = (F.state[0] & 15u) | ((F.state[1] & 15u) << 4) | (contributes ? kFlagContributes | extra : 0u);
        if (COMPACT && A.rec_words != (uint32_t)kSampleWords) {
            // the occluded estimates are +0 or NaN: their NaN channels as flags
            const float eo[6] = {F.e_occ[0].x, F.e_occ[0].y, F.e_occ[0].z, F.e_occ[1].x, F.e_occ[1].y, F.e_occ[1].z};
#pragma unroll
            for (int b = 0; b < 6; ++b) flags |= (eo[b] == 0.0f) ? 0u : (1u << (kFlagOccShift + b));
            rec[0] = make_float4(Wt, __uint_as_float(flags), F.e_vis[0].x, F.e_vis[0].y);
            rec[1] = make_float4(F.e_vis[0].z, F.e_vis[1].x, F.e_vis[1].y, F.e_vis[1].z);
            if (A.rec_words == kSampleWordsKeyed)
                rec[2] = make_float4(__int_as_float(contributes ? chosen : -1), __uint_as_float(ckey), 0.0f, 0.0f);
        } else {
            rec[0] = make_float4(Wt, __uint_as_float(flags), F.e_vis[0].x, F.e_vis[0].y);
            rec[1] = make_float4(F.e_vis[0].z, F.e_vis[1].x, F.e_vis[1].y, F.e_vis[1].z);
            rec[2] = make_float4(F.e_occ[0].x, F.e_occ[0].y, F.e_occ[0].z, F.e_occ[1].x);
            rec[3] = make_float4(F.e_occ[1].y, F.e_occ[1].z, __int_as_float(contributes ? chosen : -1),
                                 __uint_as_float(ckey));
        }
    }
}

// The debug record of a pixel-sample (RISLTC_SAMPLE_RECORD_FLOATS words):
// the reservoir, and the first iteration of the final sample.
__device__ __forceinline__ void put_debug(float* r, const FinalOut& F, int32_t chosen, float w_sum, float sample_value,
                                          float Wt, bool contributes, f3 pos, uint32_t ckey) {
    for (int k = 0; k < RISLTC_SAMPLE_RECORD_FLOATS; ++k) r[k] = 0.0f;
    r[0] = __int_as_float(chosen);
    r[1] = w_sum;
    r[2] = sample_value;
    r[3] = Wt;
    r[4] = __uint_as_float(F.clipd);
    r[5] = __uint_as_float(F.clips);
    r[6] = F.psad;
    r[7] = F.psas;
    for (int t = 0; t < 2; ++t) {
        r[8 + 3 * t] = F.dir[t].x;
        r[9 + 3 * t] = F.dir[t].y;
        r[10 + 3 * t] = F.dir[t].z;
        r[14 + t] = F.tmax[t];
        r[16 + t] = __uint_as_float(F.state[t]);
        r[18 + 3 * t] = F.e_vis[t].x;
        r[19 + 3 * t] = F.e_vis[t].y;
        r[20 + 3 * t] = F.e_vis[t].z;
        r[24 + 3 * t] = F.e_occ[t].x;
        r[25 + 3 * t] = F.e_occ[t].y;
        r[26 + 3 * t] = F.e_occ[t].z;
    }
    r[30] = F.color.x;
    r[31] = F.color.y;
    r[32] = F.color.z;
    r[37] = __uint_as_float(ckey);
    if (contributes) {
        r[34] = pos.x;
        r[35] = pos.y;
        r[36] = pos.z;
    }
}

// ---- candidate culling ----------------------------------------------------------
// A candidate whose polygon lies entirely below the shading point's horizon
// (every vertex z <= 0 in shading space) and entirely below the clipping
// plane of the LTC cosine space has clip counts 0 | 0 and p_hat = 0, so it
// changes nothing in the reservoir (w = 0 is never taken and adds exactly 0
// to w_sum). Such candidates are recognised from the light's bounding
// sphere: centre c, radius r, plane normal m and offset o give
// m.v + o <= m.c + o + |m| r for every vertex v. The margin is 1e-4 of the
// magnitudes involved, far above the few-ulp error of the vertex transform,
// so a culled candidate is one whose clip masks are zero in both builds.
struct Cull {
    float dx, dy, dz, dw, dn, dm;  // shading-space z plane, |normal|, margin
    float cx, cy, cz, cw, cn, cm;  // cosine-space z plane
};

__device__ __forceinline__ Cull make_cull(const Pixel& px, float bound) {
    Cull k;
    float plen = __builtin_sqrtf(dot3(px.pos, px.pos));
    k.dx = px.w2s[8];
    k.dy = px.w2s[9];
    k.dz = px.w2s[10];
    k.dw = px.w2s[11];
    k.dn = __builtin_sqrtf(fma_(k.dz, k.dz, fma_(k.dy, k.dy, k.dx * k.dx)));
    k.dm = 1.0e-4f * k.dn * (bound + plen);
    // row 2 of shading_to_cosine is (s2c[6], 0, s2c[8])
    const float d = px.s2c[6], e = px.s2c[8];
    k.cx = fma_(d, px.w2s[0], e * px.w2s[8]);
    k.cy = fma_(d, px.w2s[1], e * px.w2s[9]);
    k.cz = fma_(d, px.w2s[2], e * px.w2s[10]);
    k.cw = fma_(d, px.w2s[3], e * px.w2s[11]);
    k.cn = __builtin_sqrtf(fma_(k.cz, k.cz, fma_(k.cy, k.cy, k.cx * k.cx)));
    k.cm = 1.0e-4f * k.cn * (bound + plen);
    return k;
}

// NaN anywhere makes both comparisons false: never culled
__device__ __forceinline__ bool culled(const Cull& k, float4 s) {
    float hd = fma_(k.dn, s.w, fma_(k.dz, s.z, fma_(k.dy, s.y, fma_(k.dx, s.x, k.dw))));
    float hc = fma_(k.cn, s.w, fma_(k.cz, s.z, fma_(k.cy, s.y, fma_(k.cx, s.x, k.cw))));
    return hd < -k.dm && hc < -k.cm;
}

// The noise LCG advanced by d steps (d < 128) in one affine step: d steps
// take s to A_d * s + C_d, with A_0 = 1, C_0 = 0, A_{d+1} = a * A_d and
// C_{d+1} = a * C_d + c. The 128 pairs are a constant table (1 KiB).
struct LcgJumps {
    uint32_t ac[256];  // A_d at 2d, C_d at 2d + 1
    constexpr LcgJumps() : ac() {
        uint32_t a = 1u, c = 0u;
        for (int d = 0; d < 128; ++d) {
            ac[2 * d] = a;
            ac[2 * d + 1] = c;
            a = 1664525u * a;
            c = 1664525u * c + 1013904223u;
        }
    }
};
__constant__ __align__(8) LcgJumps c_lcg_jumps = LcgJumps();
__device__ __forceinline__ uint32_t lcg_jump(uint32_t s, uint32_t d) {
    const uint2 j = reinterpret_cast<const uint2*>(c_lcg_jumps.ac)[d];
    return j.x * s + j.y;
}

// ---- learned proposal: candidate draws ------------------------------------------
// Per pixel table: 16 group ids (bytes 0..15) and 16 byte weights (bytes
// 16..31), T = the sum of the weights. Group g of the table is the run of
// clusters g << G .. ((g + 1) << G) - 1 (G = 0: single clusters; G = 2: the
// quads of the 1024-output network); its lights are the consecutive run
// lights[offsets[g << G] .. offsets[min((g + 1) << G, count)]). A candidate
// takes two draws a, b:
//   T == 0        -> light_index_from(a, N);
//   a < 1/2       -> light_index_from(2 a, N)          (the uniform half);
//   otherwise     -> t = min(T - 1, int((2 a - 1) T)), slot s = first with
//                    t < w_0 + .. + w_s, group g = id_s, light = the
//                    min(size - 1, int(b size))-th light of g (an empty
//                    group, or an id past the groups, falls back to
//                    light_index_from(b, N)).
// Its source density is q = 1/(2N) + T_bad / (2 T N) + w_g / (2 T |g|):
// w_g sums the weights of the slots that hold the light's group (0 when
// none does) and T_bad the weights of the slots that fall back to a uniform
// light (the term is skipped when T_bad = 0). The oracle computes it in the
// same operation order.
struct PropTable {
    uint32_t ids[4], ws[4];
    uint32_t T;
    uint32_t T_bad;
};

__device__ __forceinline__ uint32_t byte_of(const uint32_t* v, int s) { return (v[s >> 2] >> (8 * (s & 3))) & 0xFFu; }

__device__ __forceinline__ uint32_t id_of(const PropTable& P, int s) { return byte_of(P.ids, s); }

// sum of the four bytes of v (v_sad_u8 against zero)
__device__ __forceinline__ uint32_t byte_sum(uint32_t v) { return __builtin_amdgcn_sad_u8(v, 0u, 0u); }

// 0xFF in every byte of x that is zero, 0 in the others
__device__ __forceinline__ uint32_t zero_bytes(uint32_t x) {
    uint32_t y = (x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;  // bit 7 set where the low 7 bits are not all zero
    y = ~(y | x | 0x7F7F7F7Fu);                     // 0x80 exactly in the zero bytes
    return (y >> 7) * 0xFFu;
}

// loads the pixel's table (zeros when tab is null)
__device__ __forceinline__ void load_prop_table(const uint8_t* tab, PropTable& P) {
    P.T = 0u;
    P.T_bad = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) P.ids[k] = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) P.ws[k] = 0u;
    if (!tab) return;
    const uint4* t = reinterpret_cast<const uint4*>(tab);
    uint4 t0 = t[0], t1 = t[1];
    P.ids[0] = t0.x, P.ids[1] = t0.y, P.ids[2] = t0.z, P.ids[3] = t0.w;
    P.ws[0] = t1.x, P.ws[1] = t1.y, P.ws[2] = t1.z, P.ws[3] = t1.w;
#pragma unroll
    for (int k = 0; k < 4; ++k) P.T += byte_sum(P.ws[k]);
}

struct ClusterLds {
    const uint16_t* of_light;
    const uint32_t* offsets;
    const uint32_t* lights;
    uint32_t count;
    uint32_t group_log2;  // clusters per table group: 1 << group_log2
};

// the first and one-past-last position of group g's lights (g < groups)
__device__ __forceinline__ uint2 group_range(const ClusterLds& C, uint32_t g) {
    const uint32_t c0 = g << C.group_log2;
    uint32_t c1 = (g + 1u) << C.group_log2;
    if (c1 > C.count) c1 = C.count;
    return make_uint2(C.offsets[c0], C.offsets[c1]);
}

__device__ __forceinline__ uint32_t group_count(const ClusterLds& C) {
    return (C.count + (1u << C.group_log2) - 1u) >> C.group_log2;
}

// T_bad of the pixel's table: the weight of slots whose id is past the
// groups or names an empty group (draw_learned takes a uniform light for
// them)
__device__ __forceinline__ void prop_fallback_weight(PropTable& P, const ClusterLds& C) {
    uint32_t bad = 0u;
    const uint32_t groups = group_count(C);
#pragma unroll
    for (int s = 0; s < kPropTop; ++s) {
        const uint32_t w = byte_of(P.ws, s), g = id_of(P, s);
        if (w != 0u) {
            bool empty = g >= groups;
            if (!empty) {
                const uint2 r = group_range(C, g);
                empty = r.x == r.y;
            }
            if (empty) bad += w;
        }
    }
    P.T_bad = bad;
}

__device__ __forceinline__ uint32_t draw_learned(uint32_t& seed, const PropTable& P, const ClusterLds& C, uint32_t N) {
    float a = noise_next(seed);
    float b = noise_next(seed);
    if (P.T == 0u) return light_index_from(a, N);
    if (a < 0.5f) return light_index_from(a + a, N);
    float u = (a + a) - 1.0f;
    uint32_t t = (uint32_t)(u * (float)P.T);
    if (t > P.T - 1u) t = P.T - 1u;
    // the first slot whose running weight sum passes t: first its word of
    // four slots, then the slot in the word
    const uint32_t c1 = byte_sum(P.ws[0]), c2 = c1 + byte_sum(P.ws[1]), c3 = c2 + byte_sum(P.ws[2]);
    const uint32_t wi = (t >= c1 ? 1u : 0u) + (t >= c2 ? 1u : 0u) + (t >= c3 ? 1u : 0u);
    const uint32_t ww = wi == 0u ? P.ws[0] : wi == 1u ? P.ws[1] : wi == 2u ? P.ws[2] : P.ws[3];
    const uint32_t iw = wi == 0u ? P.ids[0] : wi == 1u ? P.ids[1] : wi == 2u ? P.ids[2] : P.ids[3];
    const uint32_t rem = t - (wi == 0u ? 0u : wi == 1u ? c1 : wi == 2u ? c2 : c3);
    const uint32_t b0 = ww & 0xFFu, b1 = b0 + ((ww >> 8) & 0xFFu), b2 = b1 + ((ww >> 16) & 0xFFu);
    const uint32_t bi = (rem >= b0 ? 1u : 0u) + (rem >= b1 ? 1u : 0u) + (rem >= b2 ? 1u : 0u);
    const uint32_t g = (iw >> (8u * bi)) & 0xFFu;
    // a group id outside the groups or an empty group: uniform
    if (g >= group_count(C)) return light_index_from(b, N);
    const uint2 r = group_range(C, g);
    const uint32_t size = r.y - r.x;
    if (size == 0u) return light_index_from(b, N);
    uint32_t j = (uint32_t)(b * (float)size);
    if (j > size - 1u) j = size - 1u;
    return C.lights[r.x + j];
}

template <bool X>
__device__ __forceinline__ float learned_density(uint32_t idx, const PropTable& P, const ClusterLds& C, uint32_t N) {
    float q = 0.5f / (float)N;
    if (P.T == 0u) return 1.0f / (float)N;
    if (P.T_bad != 0u) q = q + div_<X>(0.5f * (float)P.T_bad, (float)P.T * (float)N);
    const uint32_t g = (uint32_t)C.of_light[idx] >> C.group_log2;
    // the weights of the slots whose id is g, four slots per step (ids are
    // bytes, so a group past 255 is in no slot)
    uint32_t wg = 0;
    if (g < 256u) {
        const uint32_t rep = g * 0x01010101u;
#pragma unroll
        for (int k = 0; k < 4; ++k) wg += byte_sum(P.ws[k] & zero_bytes(P.ids[k] ^ rep));
    }
    if (wg != 0u) {
        const uint2 r = group_range(C, g);
        q = q + div_<X>(0.5f * (float)wg, (float)P.T * (float)(r.y - r.x));
    }
    return q;
}

// ---- learned RIS target ----------------------------------------------------------
// With a visibility cache the learned pass's target of candidate light l is
// p_hat * (floor + v * scale), scale = (1 - floor) / 255, v = the cache's
// byte for the pixel's cell and l's slot (255, so a factor of about 1, while
// the entry is empty). The oracle evaluates the same expression.
__device__ __forceinline__ float cache_factor(const uint16_t* of_light, uint32_t idx, float floor, float scale,
                                              const uint8_t* cache_row, const uint8_t* lds, int32_t lds_row,
                                              uint32_t per_light) {
    const uint32_t slot = per_light ? idx : (uint32_t)of_light[idx];
    uint32_t v = 255u;
    if (lds_row >= 0) v = lds[(uint32_t)lds_row + slot];
    else if (cache_row) v = cache_row[slot];
    return floor + (float)v * scale;
}

// The cache byte of counts (a, b): round(254 (a + prior / 2) / (a + b + prior)),
// 255 for an empty entry. IEEE operations in this order; the oracle
// restates it.
__device__ __forceinline__ uint32_t cache_byte(uint2 c, float prior) {
    if (c.x == 0u && c.y == 0u) return 255u;
    float a = (float)c.x, b = (float)c.y;
    float v = div_<true>((a + prior * 0.5f) * 254.0f, (a + b) + prior) + 0.5f;
    uint32_t q = (uint32_t)v;
    return q > 254u ? 254u : q;
}

// The learned target reads the cache rows of a wave's cells from LDS
// (staged once per wave) instead of by scattered global loads. Up to
// kRowsPerWave rows of up to kRowLdsMax slots per wave; the row area follows
// the spheres and the cluster table, 16-byte aligned.
constexpr uint32_t kRowLdsMax = 1024, kRowsPerWave = 4;

// The learned pass stages its cluster tables after the spheres, however
// many spheres are staged: the source's draw and density read them for
// every candidate. The cluster offsets (u32, clusters + 1) always; the
// lights by cluster (u32) and the cluster of each light (u16) when the
// three take at most kLdsTableMax bytes. Layout: lights by cluster, offsets,
// cluster of each light.
constexpr uint32_t kLdsTableMax = 16384;

// lights whose per-light tables are staged (0: none)
__device__ __host__ __forceinline__ uint32_t lds_table_lights(uint32_t learned_pass, uint32_t lights,
                                                              uint32_t clusters) {
    const size_t bytes = (size_t)lights * 6u + ((size_t)clusters + 1u) * 4u;
    return (learned_pass && clusters > 0u && bytes <= kLdsTableMax) ? lights : 0u;
}

// bytes of the spheres and the staged cluster tables
__device__ __host__ __forceinline__ size_t cluster_lds_end(uint32_t n_lds, uint32_t learned_pass, uint32_t table_lights,
                                                           uint32_t clusters) {
    const bool offsets = learned_pass && clusters > 0u;
    return (size_t)n_lds * 16u + (size_t)table_lights * 6u + (offsets ? ((size_t)clusters + 1u) * 4u : 0u);
}

__device__ __host__ __forceinline__ size_t row_lds_offset(uint32_t n_lds, uint32_t learned_pass, uint32_t table_lights,
                                                          uint32_t clusters) {
    return (cluster_lds_end(n_lds, learned_pass, table_lights, clusters) + 15u) & ~(size_t)15u;
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- survivor compaction -----------------------------------------------------------
// The non-debug kernels evaluate the candidates that survive the cull
// compacted over the wave: item (k, o) is the k-th survivor of lane o, the
// items are dealt to the 64 lanes in order of k then o, and each round's
// results go back to their owners, which run the reservoir in candidate
// order. An evaluator reads its owner's pixel from LDS (one float per field
// and lane, field-major). Every value is computed as in the owner lane.
// The fields of a Pixel that eval_ltc_cp reads: position, albedo, the world
// to shading matrix, the LTC matrix and its albedo.
constexpr uint32_t kPxFloats = 28;
constexpr uint32_t kPxLdsBytes = 4u * kPxFloats * 64u * 4u;  // 4 waves
// after the pixels: each wave's deal table of a round, one word per lane
// (owner lane | k << 8 of the item the lane evaluates)
constexpr uint32_t kDealLdsBytes = 4u * 64u * 4u;
// the plain kernels' noise jumps (d < 64: two draws for each of at most 32
// candidates of a chunk) staged in LDS after the deal tables, so that the
// evaluator's jump waits on LDS rather than global memory
constexpr uint32_t kJumpLdsEntries = 64, kJumpLdsBytes = kJumpLdsEntries * 8u;
typedef uint32_t U2v __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(3))) U2v LdsU2;
__device__ __forceinline__ uint32_t lcg_jump_lds(const uint32_t* tab, uint32_t s, uint32_t d) {
    const U2v j = ((const LdsU2*)tab)[d];
    return j.x * s + j.y;
}

__device__ __forceinline__ void px_lds_store(float* b, uint32_t l, const Pixel& p) {
    b[0u * 64u + l] = p.pos.x;
    b[1u * 64u + l] = p.pos.y;
    b[2u * 64u + l] = p.pos.z;
    b[3u * 64u + l] = p.albedo.x;
    b[4u * 64u + l] = p.albedo.y;
    b[5u * 64u + l] = p.albedo.z;
#pragma unroll
    for (uint32_t i = 0; i < 12u; ++i) b[(6u + i) * 64u + l] = p.w2s[i];
#pragma unroll
    for (uint32_t i = 0; i < 9u; ++i) b[(18u + i) * 64u + l] = p.s2c[i];
    b[27u * 64u + l] = p.ltc_albedo;
}

// (fast build: the five non-zero LTC entries)
template <bool X>
__device__ __forceinline__ void px_lds_load_eval(const float* b, uint32_t l, Pixel& p) {
    p.pos.x = b[0u * 64u + l];
    p.pos.y = b[1u * 64u + l];
    p.pos.z = b[2u * 64u + l];
    p.n = p.o = mk3(0.0f, 0.0f, 0.0f);
    p.lo = p.rough = p.ltc_det = 0.0f;
    p.f0 = mk3(0.0f, 0.0f, 0.0f);
    p.albedo.x = b[3u * 64u + l];
    p.albedo.y = b[4u * 64u + l];
    p.albedo.z = b[5u * 64u + l];
#pragma unroll
    for (uint32_t i = 0; i < 12u; ++i) p.w2s[i] = b[(6u + i) * 64u + l];
#pragma unroll
    for (uint32_t i = 0; i < 9u; ++i) p.s2c[i] = (X || (i & 1u) == 0u) ? b[(18u + i) * 64u + l] : 0.0f;
    p.ltc_albedo = b[27u * 64u + l];
}

// make_cull of the lane's pixel as stored by px_lds_store
__device__ __forceinline__ Cull cull_from_lds(const float* b, uint32_t l, float bound) {
    Pixel p = {};
    p.pos.x = b[0u * 64u + l];
    p.pos.y = b[1u * 64u + l];
    p.pos.z = b[2u * 64u + l];
#pragma unroll
    for (uint32_t i = 0; i < 12u; ++i) p.w2s[i] = b[(6u + i) * 64u + l];
    p.s2c[6] = b[24u * 64u + l];
    p.s2c[8] = b[26u * 64u + l];
    return make_cull(p, bound);
}

// LDS offset of the pixel area: after the spheres, cluster tables and rows
__device__ __host__ __forceinline__ size_t px_lds_offset(uint32_t n_lds, uint32_t learned_pass, uint32_t table_lights,
                                                         uint32_t clusters, bool rows) {
    size_t b = rows ? row_lds_offset(n_lds, learned_pass, table_lights, clusters) + 4u * kRowsPerWave * kRowLdsMax
                    : cluster_lds_end(n_lds, learned_pass, table_lights, clusters);
    return (b + 15u) & ~(size_t)15u;
}

// Cache key of a shading point: its cell floor(p / cell) per axis and the
// dominant axis and sign of the shading normal (the first axis on ties),
// combined with MurmurHash3's mix and finaliser, masked to the table.
__device__ __forceinline__ uint32_t cache_key(const VisCacheView& C, f3 p, f3 n) {
    const float lim = 1.0e9f;
    int32_t ix = (int32_t)__builtin_floorf(clamp_(p.x * C.inv_cell, -lim, lim));
    int32_t iy = (int32_t)__builtin_floorf(clamp_(p.y * C.inv_cell, -lim, lim));
    int32_t iz = (int32_t)__builtin_floorf(clamp_(p.z * C.inv_cell, -lim, lim));
    float ax = fabs_(n.x), ay = fabs_(n.y), az = fabs_(n.z);
    uint32_t code = (ax >= ay && ax >= az) ? (n.x < 0.0f ? 1u : 0u)
                                           : (ay >= az ? (n.y < 0.0f ? 3u : 2u) : (n.z < 0.0f ? 5u : 4u));
    uint32_t h = murmur_mix(murmur_mix(murmur_mix(murmur_mix(0u, (uint32_t)ix), (uint32_t)iy), (uint32_t)iz), code);
    return murmur_finalize(h) & C.mask;
}

// ---- the kernel ----------------------------------------------------------------
// One 4-wave work-group shades one tile of one view. A wave covers
// 2^tile_w_log2 x (64 >> tile_w_log2) pixels; the work-group stacks its 4
// waves vertically. The first lds_lights light spheres (all of them when
// they fit) are staged in LDS; the cluster table too for the learned target
// when every sphere is staged. DEBUG = true is the general instantiation:
// candidates evaluated per lane, debug records written when their pointers
// are set, and the SAMPLE_COUNT loop of the final sample; the production
// instantiation (DEBUG = false) takes SAMPLE_COUNT = 1, so no state is held
// across a loop there.
template <bool X, bool DEBUG, int LMODE>
__device__ __forceinline__ void shade_tile(const ShadeArgs& A, uint32_t tile, float4* sh_spheres) {
    // LMODE: 0 plain, 1 learned
    constexpr bool learned = LMODE != 0;
    ClusterLds cl{A.clusters.of_light, A.clusters.offsets, A.clusters.lights, A.clusters.count,
                  A.proposal_group_log2};
    const uint32_t n_lds = A.scene.light_count < A.lds_lights ? A.scene.light_count : A.lds_lights;
    const bool all_lds = n_lds == A.scene.light_count;
    const uint32_t tab_n = learned ? lds_table_lights(A.learned_pass, A.scene.light_count, A.clusters.count) : 0u;
    const bool tab_o = learned && A.learned_pass && A.clusters.count > 0u;
    // the plain compacted kernels' jump table
    constexpr bool jt = !learned && !DEBUG;
    uint32_t* jtab = nullptr;
    if constexpr (jt)
        jtab = reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(sh_spheres) +
                                           px_lds_offset(n_lds, 0u, 0u, A.clusters.count, false) + kPxLdsBytes +
                                           kDealLdsBytes);
    if (n_lds > 0u || tab_o || jt) {
        for (uint32_t k = threadIdx.x; k < n_lds; k += blockDim.x) sh_spheres[k] = A.scene.spheres[k];
        if constexpr (jt)
            for (uint32_t k = threadIdx.x; k < 2u * kJumpLdsEntries; k += blockDim.x) jtab[k] = c_lcg_jumps.ac[k];
        if constexpr (learned) {
            if (tab_o) {
                // after the spheres: lights by cluster (u32, when staged),
                // offsets (u32), cluster of light (u16, when staged)
                uint32_t* l = reinterpret_cast<uint32_t*>(sh_spheres + n_lds);
                uint32_t* o = l + tab_n;
                uint16_t* c = reinterpret_cast<uint16_t*>(o + A.clusters.count + 1u);
                for (uint32_t k = threadIdx.x; k < tab_n; k += blockDim.x) {
                    l[k] = A.clusters.lights[k];
                    c[k] = A.clusters.of_light[k];
                }
                for (uint32_t k = threadIdx.x; k <= A.clusters.count; k += blockDim.x) o[k] = A.clusters.offsets[k];
                cl = ClusterLds{tab_n ? c : A.clusters.of_light, o, tab_n ? l : A.clusters.lights, A.clusters.count,
                                A.proposal_group_log2};
            }
        }
        __syncthreads();
    }
    const uint32_t W = A.c.viewport_size[0], H = A.c.viewport_size[1];
    const uint32_t ww_log2 = A.tile_w_log2, ww = 1u << ww_log2, wh = 64u >> ww_log2;
    const uint32_t tiles_x = (W + ww - 1u) >> ww_log2;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t x = (tile % tiles_x) * ww + (lane & (ww - 1u));
    const uint32_t y = (tile / tiles_x) * (4u * wh) + wave * wh + (lane >> ww_log2);
    const bool inside = x < W && y < H;
    const uint32_t pixel = y * W + x;
    const uint32_t spp = A.spp;
    uint32_t prim = inside ? A.visibility[pixel] : 0xFFFFFFFFu;
    // ids past the mesh are treated as background (the reference would read
    // out of bounds)
    bool shade = inside && (prim >> 31) == 0u && (uint64_t)prim < A.scene.triangle_count;

    Pixel px = {};
    Cull cull = {};
    uint32_t seed = 0;
    if (shade) {
        const float(*P)[4] = A.c.pixel_to_ray_direction_world_space;
        float fx = (float)x, fy = (float)y;
        f3 ray = mk3(fma_(P[0][2], 1.0f, fma_(P[0][1], fy, P[0][0] * fx)),
                     fma_(P[1][2], 1.0f, fma_(P[1][1], fy, P[1][0] * fx)),
                     fma_(P[2][2], 1.0f, fma_(P[2][1], fy, P[2][0] * fx)));
        setup_pixel(A, prim, ray, px);
        // (the compacted kernels derive it per sample from the pixel in LDS)
        if constexpr (DEBUG) cull = make_cull(px, A.scene.sphere_bound);
        seed = murmur_finalize(murmur_mix(murmur_mix(0u, x + y * W), A.c.noise_random_numbers[0]));
    }
    const uint32_t ckey = (shade && A.cache.counts) ? cache_key(A.cache, px.pos, px.n) : 0u;
    // the learned target's cache rows: the wave stages the rows of up to
    // kRowsPerWave distinct cells of its shaded lanes in LDS; a lane whose
    // cell did not fit reads its row from global memory
    const uint8_t* crow = nullptr;
    int32_t lds_row = -1;  // byte offset of the lane's staged row in the LDS area, or -1
    if constexpr (learned) {
        if (A.cache.vis) {
            crow = A.cache.vis + (size_t)ckey * A.cache.slots;
            if (A.cache.slots <= kRowLdsMax) {
                const uint32_t wave_base = (uint32_t)row_lds_offset(n_lds, A.learned_pass, tab_n, A.clusters.count) +
                                           wave * kRowsPerWave * kRowLdsMax;
                uint8_t* area = reinterpret_cast<uint8_t*>(sh_spheres) + wave_base;
                const uint32_t slots = A.cache.slots;
                uint64_t pending = __ballot(shade);
                for (uint32_t r = 0; r < kRowsPerWave && pending != 0ull; ++r) {
                    const int first = __ffsll((unsigned long long)pending) - 1;
                    const uint32_t kr = (uint32_t)__shfl((int)ckey, first);
                    const bool mine = shade && ckey == kr;
                    pending &= ~__ballot(mine);
                    uint8_t* dst = area + (size_t)r * kRowLdsMax;
                    const uint8_t* src = A.cache.vis + (size_t)kr * slots;
                    if ((slots & 15u) == 0u) {
                        for (uint32_t o = lane * 16u; o < slots; o += 1024u)
                            *reinterpret_cast<uint4*>(dst + o) = *reinterpret_cast<const uint4*>(src + o);
                    } else {
                        for (uint32_t o = lane; o < slots; o += 64u) dst[o] = src[o];
                    }
                    if (mine) lds_row = (int32_t)(wave_base + r * kRowLdsMax);
                }
                wave_lds_sync();
            }
        }
    }
    const uint32_t N = A.scene.light_count;
    const float pN = 1.0f / (float)N;
    const float kN = (float)N * (1.0f / 4294967296.0f);
    const uint32_t M = A.M;
    // noise draws per candidate: light (1, or 2 for the learned source) + reservoir
    const uint32_t stride = learned ? 3u : 2u;
    const bool cached = learned && A.cache.vis != nullptr;
    // the wave's pixels in LDS for the compacted evaluation (non-debug)
    float* pxa = nullptr;
    uint32_t* deal = nullptr;
    if constexpr (!DEBUG) {
        const bool rows = learned && A.cache.vis != nullptr && A.cache.slots <= kRowLdsMax;
        uint8_t* px_base = reinterpret_cast<uint8_t*>(sh_spheres) +
                           px_lds_offset(n_lds, learned ? A.learned_pass : 0u, tab_n, A.clusters.count, rows);
        pxa = reinterpret_cast<float*>(px_base) + wave * kPxFloats * 64u;
        deal = reinterpret_cast<uint32_t*>(px_base + kPxLdsBytes) + wave * 64u;
        px_lds_store(pxa, lane, px);
        wave_lds_sync();
    }
    for (uint32_t j = 0; j < spp; ++j) {
        int32_t chosen = -1;
        float w_sum = 0.0f, sample_value = 0.0f;
        Light L;
        // the pixel's proposal table, loaded per sample so that it is not
        // live across the final sample
        PropTable pt;
        load_prop_table(learned && shade && A.proposal ? A.proposal + (size_t)pixel * kPropBytes : nullptr, pt);
        if constexpr (learned) {
            if (pt.T != 0u) prop_fallback_weight(pt, cl);
        }
        if (A.light_sampling == RISLTC_LIGHT_UNIFORM) {
            if (shade) chosen = (int32_t)light_index_from(noise_next(seed), N);
        } else if constexpr (DEBUG) {
            if (shade) {
                // Candidates in chunks of 32: first every candidate's light
                // index and cull test (cheap, wave-uniform), then the full
                // target of the survivors in their original order, the noise
                // state jumped to each survivor. Culled candidates have
                // w = 0 and leave the reservoir exactly as it was.
                const uint32_t chunk = A.chunk;
                for (uint32_t base = 0; base < M; base += chunk) {
                    const uint32_t cnt = (M - base < chunk) ? M - base : chunk;
                    const uint32_t seed_chunk = seed;
                    uint32_t live = 0;
                    for (uint32_t t = 0; t < cnt; ++t) {
                        uint32_t idx = learned ? draw_learned(seed, pt, cl, N) : light_index_from(noise_next(seed), N);
                        (void)noise_next(seed);
                        float4 sp = all_lds ? sh_spheres[idx] : (idx < n_lds ? sh_spheres[idx] : A.scene.spheres[idx]);
                        bool cut = culled(cull, sp);
                        live |= cut ? 0u : (1u << t);
                        if constexpr (DEBUG) {
                            if (A.debug_cands && cut) {
                                uint32_t* cr = A.debug_cands + (((size_t)pixel * spp + j) * M + base + t) *
                                                                   RISLTC_CANDIDATE_RECORD_WORDS;
                                cr[0] = idx;
                                cr[1] = 0u;
                                cr[2] = 0u;
                                cr[3] = 0u;
                            }
                        }
                    }
                    uint32_t s = seed_chunk, at = 0;
                    while (live) {
                        const uint32_t t = (uint32_t)__builtin_ctz(live);
                        live &= live - 1u;
                        s = lcg_jump(s, stride * (t - at));
                        at = t + 1u;
                        uint32_t idx = learned ? draw_learned(s, pt, cl, N) : light_index_from(noise_next(s), N);
                        load_light(A.scene.lights, idx, L);
                        // the cache factor first: its load is in flight
                        // during the target's evaluation
                        float factor = 1.0f;
                        if (cached)
                            factor = cache_factor(cl.of_light, idx, A.learned_floor, A.learned_scale, crow,
                                                  reinterpret_cast<const uint8_t*>(sh_spheres), lds_row,
                                                  A.cache.per_light);
                        uint32_t cc;
                        f3 col = eval_ltc_cp<X>(px, L, cc);
                        float p_hat = sqrt_<X>(dot3(col, col));
                        if (cached) p_hat = p_hat * factor;
                        float w = div_<X>(p_hat, learned ? learned_density<X>(idx, pt, cl, N) : pN);
                        float r = noise_next(s);
                        w_sum += w;
                        // exact: the reference's r < w / w_sum; fast: the
                        // same test without the division (w_sum >= w > 0)
                        bool take = X ? (r < w / w_sum) : (r * w_sum < w);
                        if (w > 0.0f && take) {
                            chosen = (int32_t)idx;
                            sample_value = p_hat;
                        }
                        if constexpr (DEBUG) {
                            if (A.debug_cands) {
                                uint32_t* cr = A.debug_cands + (((size_t)pixel * spp + j) * M + base + t) *
                                                                   RISLTC_CANDIDATE_RECORD_WORDS;
                                cr[0] = idx;
                                cr[1] = cc;
                                cr[2] = __float_as_uint(p_hat);
                                cr[3] = __float_as_uint(w);
                            }
                        }
                    }
                }
            }
        } else {
            // Candidates in chunks: every lane draws its chunk's light
            // indices and cull tests, then the wave evaluates the survivors
            // compacted (see above): the evaluator jumps the owner's noise
            // state to the candidate and draws its light and its reservoir
            // number, and each owner runs its reservoir over its survivors
            // in their original order. Culled candidates have w = 0 and
            // leave the reservoir as it was.
            const uint32_t chunk = A.chunk;
            // the cull planes, from the pixel in LDS: not held in registers
            // across the final sample
            const Cull cull_s = cull_from_lds(pxa, lane, A.scene.sphere_bound);
            for (uint32_t base = 0; base < M; base += chunk) {
                const uint32_t cnt = (M - base < chunk) ? M - base : chunk;
                const uint32_t seed_chunk = seed;
                uint32_t live = 0;
                if (shade) {
                    const LdsF4* sl = (const LdsF4*)sh_spheres;
                    for (uint32_t t = 0; t < cnt; ++t) {
                        uint32_t idx = learned ? draw_learned(seed, pt, cl, N) : light_index_next(seed, kN, N);
                        (void)noise_next(seed);
                        // (every sphere in LDS: an LDS read)
                        float4 sp = all_lds ? lds_f4(sl, idx)
                                            : (idx < n_lds ? sh_spheres[idx] : A.scene.spheres[idx]);
                        live |= culled(cull_s, sp) ? 0u : (1u << t);
                    }
                }
                const uint32_t mycnt = (uint32_t)__popc(live);
                // the owner's survivors not yet dealt (they are dealt in order)
                uint32_t live_rem = live;
                // wave cursor: the next item is rank lvl_off of level lvl
                uint32_t lvl = 0, lvl_off = 0;
                uint64_t lvl_mask = __ballot(mycnt > 0u);
                while (lvl_mask != 0ull) {
                    // deal this round's items: lane e takes the e-th remaining
                    // item; each owner writes its items' (owner, candidate)
                    // into the deal table at the evaluating lane
                    uint32_t my_t = 0u, my_o = lane;
                    bool have = false;
                    {
                        uint32_t kk = lvl, oo = lvl_off, ls = 0;
                        uint64_t m = lvl_mask;
                        const uint64_t below = (1ull << lane) - 1ull;
                        while (ls < 64u && m != 0ull) {
                            const uint32_t n = (uint32_t)__popcll(m) - oo;
                            const uint32_t rank = (uint32_t)__popcll(m & below);
                            if (((m >> lane) & 1ull) != 0ull && rank >= oo && ls + (rank - oo) < 64u) {
                                const uint32_t t = (uint32_t)__builtin_ctz(live_rem);
                                live_rem &= live_rem - 1u;
                                deal[ls + (rank - oo)] = lane | (t << 8);
                            }
                            ls += n;
                            kk += 1u;
                            oo = 0u;
                            m = __ballot(mycnt > kk);
                        }
                        wave_lds_sync();
                        if (lane < ls) {
                            const uint32_t v = deal[lane];
                            my_o = v & 63u;
                            my_t = v >> 8;
                            have = true;
                        }
                        wave_lds_sync();
                    }
                    const uint32_t seed_o = (uint32_t)__shfl((int)seed_chunk, (int)my_o);
                    PropTable pt_o;
                    pt_o.T = 0u;
                    pt_o.T_bad = 0u;
#pragma unroll
                    for (int k = 0; k < 4; ++k) pt_o.ids[k] = 0u;
#pragma unroll
                    for (int k = 0; k < 4; ++k) pt_o.ws[k] = 0u;
                    if constexpr (learned) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) pt_o.ids[k] = (uint32_t)__shfl((int)pt.ids[k], (int)my_o);
#pragma unroll
                        for (int k = 0; k < 4; ++k) pt_o.ws[k] = (uint32_t)__shfl((int)pt.ws[k], (int)my_o);
                        pt_o.T = (uint32_t)__shfl((int)pt.T, (int)my_o);
                    }
                    float p_raw = 0.0f, r_e = 0.0f;
                    uint32_t idx_e = 0u;
                    if (have) {
                        uint32_t se = jt ? lcg_jump_lds(jtab, seed_o, stride * my_t) : lcg_jump(seed_o, stride * my_t);
                        idx_e = learned ? draw_learned(se, pt_o, cl, N) : light_index_next(se, kN, N);
                        // the candidate's reservoir draw, taken by its owner
                        r_e = noise_next(se);
                        Pixel q;
                        px_lds_load_eval<X>(pxa, my_o, q);
                        load_light(A.scene.lights, idx_e, L);
                        uint32_t cc;
                        f3 col = eval_ltc_cp<X>(q, L, cc);
                        p_raw = sqrt_<X>(dot3(col, col));
                    }
                    // owners take their items of this round in level order
                    {
                        uint32_t kk = lvl, oo = lvl_off, ls = 0;
                        uint64_t m = lvl_mask;
                        while (ls < 64u && m != 0ull) {
                            const uint32_t n = (uint32_t)__popcll(m) - oo;
                            const bool mine = ((m >> lane) & 1ull) != 0ull;
                            const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                            const bool here = mine && rank >= oo && rank - oo < n && ls + (rank - oo) < 64u;
                            const int src = here ? (int)(ls + rank - oo) : (int)lane;
                            const float ph = __shfl(p_raw, src);
                            const uint32_t idx = (uint32_t)__shfl((int)idx_e, src);
                            const float r = __shfl(r_e, src);
                            if (here) {
                                float p_hat = ph;
                                if (cached)
                                    p_hat = p_hat * cache_factor(cl.of_light, idx, A.learned_floor, A.learned_scale,
                                                                 crow, reinterpret_cast<const uint8_t*>(sh_spheres),
                                                                 lds_row, A.cache.per_light);
                                float w = div_<X>(p_hat, learned ? learned_density<X>(idx, pt, cl, N) : pN);
                                w_sum += w;
                                bool take = X ? (r < w / w_sum) : (r * w_sum < w);
                                if (w > 0.0f && take) {
                                    chosen = (int32_t)idx;
                                    sample_value = p_hat;
                                }
                            }
                            ls += n;
                            kk += 1u;
                            oo = 0u;
                            m = __ballot(mycnt > kk);
                        }
                    }
                    // advance the wave cursor past the round's 64 items
                    uint32_t left = 64u;
                    while (lvl_mask != 0ull) {
                        const uint32_t n = (uint32_t)__popcll(lvl_mask) - lvl_off;
                        if (left < n) {
                            lvl_off += left;
                            break;
                        }
                        left -= n;
                        lvl += 1u;
                        lvl_off = 0u;
                        lvl_mask = __ballot(mycnt > lvl);
                    }
                }
            }
        }
        float Wt = 0.0f;
        const bool contributes = shade && chosen >= 0;
        if (contributes) {
            if (A.light_sampling == RISLTC_LIGHT_UNIFORM) {
                Wt = (float)N;
            } else {
                Wt = w_sum / ((float)M * sample_value);
                if (sample_value == 0.0f) Wt = 0.0f;
            }
        }
        const size_t sidx = (size_t)pixel * spp + j;
        // the SAMPLE_COUNT loop: every iteration prepares the two polygons
        // again (the same arithmetic, no draws) and takes one direction per
        // technique, so the draws follow the reference's loop
        const uint32_t iters = DEBUG ? A.iters : 1u;
        // the compacted kernels lend the lane's column of the LDS pixel area
        // to the final sample's clipping, and store the pixel again after it
        if constexpr (!DEBUG) wave_lds_sync();
        for (uint32_t it = 0; it < iters; ++it) {
            FinalOut F;
            clear_final(F);
            // the light is read per iteration so that it is not held across the loop
            if (contributes) {
                load_light(A.scene.lights, (uint32_t)chosen, L);
                if constexpr (DEBUG) eval_peters<X>(A, px, L, seed, F);
                else eval_peters<X, true>(A, px, L, seed, F, pxa + lane);
            }
            put_final<true>(A, lane, inside, contributes, F, px.pos, sidx * iters + it, Wt, 0u, chosen, ckey);
            if constexpr (DEBUG) {
                if (it == 0u && inside && A.debug_samples)
                    put_debug(A.debug_samples + sidx * RISLTC_SAMPLE_RECORD_FLOATS, F, shade ? chosen : -1, w_sum,
                              sample_value, Wt, contributes, px.pos, ckey);
            }
        }
        if constexpr (DEBUG) {
            if (inside && A.debug_samples)
                A.debug_samples[sidx * RISLTC_SAMPLE_RECORD_FLOATS + 33] = __uint_as_float(seed);
        } else {
            if (j + 1u < spp) {
                px_lds_store(pxa, lane, px);
                wave_lds_sync();
            }
        }
    }
}

// Three waves per SIMD: the register budget is 168 VGPRs (the compacted
// kernels spill at most a few of them); the measured gain over two waves
// comes from hiding the latency of the candidate evaluation.
template <bool X, bool DEBUG, int LMODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) shade_kernel(ShadeArgs A) {
    extern __shared__ float4 sh_spheres[];
    shade_tile<X, DEBUG, LMODE>(A, blockIdx.x, sh_spheres);
}

// the last v with offsets[v] <= i (offsets[0] = 0, n + 1 entries)
// ---- baseline techniques ----------------------------------------------------------
// The two polygon sampling techniques the reference compares against
// (src/experiment_list.c:313-396), with the uniform source. One lane per
// pixel evaluates its candidates one after the other: a projected-solid-angle
// candidate consumes 0, 2 or 4 noise draws depending on the polygon, which
// rules out the compacted evaluation of shade_kernel. Lights are read from
// global memory (they stay in L2). The kernel writes the same sample records
// and compacted shadow rays as shade_kernel, so trace and resolve are shared.

// sample_area_polygon_turk (polygon_sampling_related_work.glsl:34-43): a
// uniform point on the triangle of vertices 0, 1, 2 (the reference samples
// only this triangle of the polygon)
template <bool X>
__device__ __forceinline__ f3 turk_point(const Light& L, float u0, float u1) {
    float s = sqrt_<X>(u0);
    float b0 = 1.0f - s, b1 = s * u1, b2 = fma_(-s, u1, s);
    return mk3(b0 * L.v[0] + b1 * L.v[3] + b2 * L.v[6], b0 * L.v[1] + b1 * L.v[4] + b2 * L.v[7],
               b0 * L.v[2] + b1 * L.v[5] + b2 * L.v[8]);
}

// get_polygon_area (shading_pass.frag.glsl:148-152): the triangle of
// vertices 0, 1, 2
template <bool X>
__device__ __forceinline__ float turk_area(const Light& L) {
    f3 v0 = mk3(L.v[0], L.v[1], L.v[2]);
    f3 c = cross_stable(sub3(mk3(L.v[3], L.v[4], L.v[5]), v0), sub3(mk3(L.v[6], L.v[7], L.v[8]), v0));
    return sqrt_<X>(dot3(c, c)) / 2.0f;
}

// The AREA_TURK branch of evaluate_polygonal_light_shading
// (shading_pass.frag.glsl:410-418) at one point: get_area_sample_density
// (polygon_sampling_related_work.glsl:57-64) and
// get_polygonal_light_mis_estimate (shading_pass.frag.glsl:281-287) without
// visibility. dir and tmax are the shadow ray of get_polygon_visibility
// (112-129); above = the point is above the shading horizon.
template <bool X>
__device__ __forceinline__ f3 eval_turk(const Pixel& px, const Light& L, f3 point, f3& dir, float& tmax,
                                        bool& above) {
    f3 d = sub3(point, px.pos);
    float d2 = dot3(d, d);
    d = scl3(d, rsq_<X>(d2));
    f3 pn = mk3(L.plane[0], L.plane[1], L.plane[2]);
    float density = div_<X>(d2, fabs_(dot3(pn, d)) * turk_area<X>(L));
    float lambert = dot3(px.n, d);
    dir = d;
    above = lambert > 0.0f;
    tmax = 0.0f;
    f3 rb = mk3(0.0f, 0.0f, 0.0f);
    if (above) {
        tmax = div_<X>(-plane_side(px.pos, L.plane), dot3(d, pn));
        tmax -= 1e-3f;
        rb = mul3(L.rad, evaluate_brdf<X>(px, d));
    }
    return (density > 0.0f) ? scl3(rb, div_<X>(lambert, density)) : mk3(0.0f, 0.0f, 0.0f);
}

// the final sample of AREA_TURK at the stored point: technique slot 0 holds
// the unshadowed estimate and its shadow ray; occluded means 0
// (shading_pass.frag.glsl:757-759)
template <bool X>
__device__ __forceinline__ void final_turk(const ShadeArgs& A, const Pixel& px, const Light& L, f3 point,
                                           FinalOut& F) {
    bool above;
    f3 dir;
    float tmax;
    f3 col = eval_turk<X>(px, L, point, dir, tmax, above);
    F.e_vis[0] = col;
    F.dir[0] = dir;
    F.tmax[0] = tmax;
    F.state[0] = above ? (A.shadow_rays ? 5u : 4u) : 1u;
}

template <bool X, uint32_t TECH>
__global__ void __launch_bounds__(256) shade_baseline_kernel(ShadeArgs A) {
    const uint32_t W = A.c.viewport_size[0], H = A.c.viewport_size[1];
    const uint32_t ww_log2 = A.tile_w_log2, ww = 1u << ww_log2, wh = 64u >> ww_log2;
    const uint32_t tiles_x = (W + ww - 1u) >> ww_log2;
    const uint32_t tile = blockIdx.x;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t x = (tile % tiles_x) * ww + (lane & (ww - 1u));
    const uint32_t y = (tile / tiles_x) * (4u * wh) + wave * wh + (lane >> ww_log2);
    const bool inside = x < W && y < H;
    const uint32_t pixel = y * W + x;
    const uint32_t spp = A.spp, M = A.M, N = A.scene.light_count;
    const float pN = 1.0f / (float)N;
    uint32_t prim = inside ? A.visibility[pixel] : 0xFFFFFFFFu;
    const bool shade = inside && (prim >> 31) == 0u && (uint64_t)prim < A.scene.triangle_count;
    Pixel px = {};
    uint32_t seed = 0;
    if (shade) {
        const float(*P)[4] = A.c.pixel_to_ray_direction_world_space;
        float fx = (float)x, fy = (float)y;
        f3 ray = mk3(fma_(P[0][2], 1.0f, fma_(P[0][1], fy, P[0][0] * fx)),
                     fma_(P[1][2], 1.0f, fma_(P[1][1], fy, P[1][0] * fx)),
                     fma_(P[2][2], 1.0f, fma_(P[2][1], fy, P[2][0] * fx)));
        setup_pixel(A, prim, ray, px);
        seed = murmur_finalize(murmur_mix(murmur_mix(0u, x + y * W), A.c.noise_random_numbers[0]));
    }
    const uint32_t ckey = (shade && A.cache.counts) ? cache_key(A.cache, px.pos, px.n) : 0u;
    for (uint32_t j = 0; j < spp; ++j) {
        int32_t chosen = -1;
        float w_sum = 0.0f, sample_value = 0.0f;
        f3 chosen_point = mk3(0.0f, 0.0f, 0.0f);
        Light L;
        if (A.light_sampling == RISLTC_LIGHT_UNIFORM) {
            if (shade) chosen = (int32_t)light_index_from(noise_next(seed), N);
        } else if (shade) {
            for (uint32_t i = 0; i < M; ++i) {
                uint32_t idx = light_index_from(noise_next(seed), N);
                load_light(A.scene.lights, idx, L);
                f3 col = mk3(0.0f, 0.0f, 0.0f), point = mk3(0.0f, 0.0f, 0.0f);
                uint32_t cc = 0;
                // SAMPLE_COUNT samples (shading_pass.frag.glsl:412-417, 518-559),
                // their mean is the target; the last area sample is the
                // reservoir's point
                if constexpr (TECH == RISLTC_SAMPLE_POLYGON_AREA_TURK) {
                    for (uint32_t it = 0; it < A.iters; ++it) {
                        float u0 = noise_next(seed);
                        float u1 = noise_next(seed);
                        point = turk_point<X>(L, u0, u1);
                        f3 d;
                        float tm;
                        bool ab;
                        f3 e = eval_turk<X>(px, L, point, d, tm, ab);
                        col = it == 0u ? e : add3(col, e);
                    }
                } else {
                    // the final-sample code without tracing: a sample above
                    // the horizon keeps its unshadowed estimate
                    for (uint32_t it = 0; it < A.iters; ++it) {
                        FinalOut F;
                        eval_peters<X>(A, px, L, seed, F);
#pragma unroll
                        for (int t = 0; t < 2; ++t)
                            if (F.state[t] != 0u) col = add3(col, F.state[t] >= 4u ? F.e_vis[t] : F.e_occ[t]);
                        cc = F.clipd | (F.clips << 8);
                    }
                }
                if (A.iters > 1u) col = scl3(col, 1.0f / (float)A.iters);
                float p_hat = sqrt_<X>(dot3(col, col));
                float w = div_<X>(p_hat, pN);
                float r = noise_next(seed);
                w_sum += w;
                bool take = X ? (r < w / w_sum) : (r * w_sum < w);
                if (w > 0.0f && take) {
                    chosen = (int32_t)idx;
                    sample_value = p_hat;
                    chosen_point = point;
                }
                if (A.debug_cands) {
                    uint32_t* cr = A.debug_cands + (((size_t)pixel * spp + j) * M + i) * RISLTC_CANDIDATE_RECORD_WORDS;
                    cr[0] = idx;
                    cr[1] = cc;
                    cr[2] = __float_as_uint(p_hat);
                    cr[3] = __float_as_uint(w);
                }
            }
        }
        float Wt = 0.0f;
        const bool contributes = shade && chosen >= 0;
        if (contributes) {
            if (A.light_sampling == RISLTC_LIGHT_UNIFORM) {
                Wt = (float)N;
            } else {
                Wt = w_sum / ((float)M * sample_value);
                if (sample_value == 0.0f) Wt = 0.0f;
            }
        }
        const size_t sidx = (size_t)pixel * spp + j;
        // area sampling with SAMPLE_COUNT > 1: one visibility flag for the
        // whole loop (shading_pass.frag.glsl:412-417 with 743-756 or 714-720)
        const uint32_t joint = (TECH == RISLTC_SAMPLE_POLYGON_AREA_TURK && A.iters > 1u) ? kFlagJoint : 0u;
        for (uint32_t it = 0; it < A.iters; ++it) {
            FinalOut F;
            clear_final(F);
            if (contributes) {
                load_light(A.scene.lights, (uint32_t)chosen, L);
                if constexpr (TECH == RISLTC_SAMPLE_POLYGON_AREA_TURK) {
                    if (A.light_sampling == RISLTC_LIGHT_UNIFORM) {
                        float u0 = noise_next(seed);
                        float u1 = noise_next(seed);
                        chosen_point = turk_point<X>(L, u0, u1);
                    }
                    final_turk<X>(A, px, L, chosen_point, F);
                    // every iteration evaluates the reservoir's point: its
                    // first shadow ray answers for all of them
                    if (A.light_sampling != RISLTC_LIGHT_UNIFORM && it > 0u && F.state[0] == 5u) F.state[0] = 4u;
                } else {
                    eval_peters<X>(A, px, L, seed, F);
                }
            }
            put_final<false>(A, lane, inside, contributes, F, px.pos, sidx * A.iters + it, Wt, joint, chosen, ckey);
            if (it == 0u && inside && A.debug_samples)
                put_debug(A.debug_samples + sidx * RISLTC_SAMPLE_RECORD_FLOATS, F, shade ? chosen : -1, w_sum,
                          sample_value, Wt, contributes, px.pos, ckey);
        }
        if (inside && A.debug_samples) A.debug_samples[sidx * RISLTC_SAMPLE_RECORD_FLOATS + 33] = __uint_as_float(seed);
    }
}

__device__ __forceinline__ uint32_t view_of(const uint32_t* offsets, uint32_t n, uint32_t i) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (offsets[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Many views in one launch: work-group b shades tile b - tile_offsets[v] of
// view v, with v's descriptor read from device memory (so a captured graph
// can be replayed with new cameras and noise seeds). LMODE as for
// shade_tile: every view of the launch runs the same pass. GEN: the general
// instantiation of shade_tile (batches with a SAMPLE_COUNT above 1).
template <bool X, int LMODE, bool GEN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
shade_batch_kernel(const ShadeArgs* views, const uint32_t* tile_offsets, uint32_t count) {
    extern __shared__ float4 sh_spheres[];
    const uint32_t v = view_of(tile_offsets, count, blockIdx.x);
    shade_tile<X, GEN, LMODE>(views[v], blockIdx.x - tile_offsets[v], sh_spheres);
}

// ---- resolve: sum the estimates in the reference's order -------------------
// adds the n pending (slot, count pair) of one cache cell
__device__ __forceinline__ void flush_counts(const VisCacheView& C, uint32_t key, const uint32_t (&slot)[4],
                                             const unsigned long long (&add)[4], uint32_t n) {
#pragma unroll
    for (uint32_t q = 0; q < 4u; ++q)
        if (q < n)
            atomicAdd(reinterpret_cast<unsigned long long*>(C.counts + (size_t)key * C.slots + slot[q]), add[q]);
}

// GROUPED: a pixel's samples (one iteration) are loaded in groups of 4
// whatever their count; otherwise only up to 4 samples are (one group), more
// are loaded one by one
template <bool GROUPED>
__device__ __forceinline__ void resolve_pixel(const ResolveArgs& A, uint32_t pixel) {
    if (pixel >= A.width * A.height) return;
    uint32_t prim = A.visibility[pixel];
    const bool shaded = prim != 0xFFFFFFFFu && (prim >> 31) == 0u && (uint64_t)prim < A.triangle_count;
    if (!shaded && A.rec_cluster) {
        // no sample: empty training records
        for (uint32_t j = 0; j < A.spp; ++j) {
            A.rec_cluster[(size_t)pixel * A.spp + j] = 0xFFFFFFFFu;
            A.rec_y[(size_t)pixel * A.spp + j] = 0.0f;
        }
    }
    float4 out;
    if (prim == 0xFFFFFFFFu || ((prim >> 31) == 0u && (uint64_t)prim >= A.triangle_count)) {
        out = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    } else {
        f3 fin = mk3(0.0f, 0.0f, 0.0f);
        if ((prim >> 31) > 0u) {
            fin = mk3(1.0f, 1.0f, 1.0f);
        } else {
            const uint32_t K = A.iters;
            // records of A.rec_words words (rw4 float4s)
            const uint32_t rw4 = A.rec_words / 4u;
            const bool compact = A.rec_words != (uint32_t)kSampleWords;
            // the cache counts of the pixel's samples (one cell for all of
            // them), merged per slot before the atomics: up to 4 pending
            uint32_t pslot[4] = {0u, 0u, 0u, 0u}, np = 0u, pkey = 0u;
            unsigned long long padd[4] = {0ull, 0ull, 0ull, 0ull};
            // one sample: its records and hit flags are either loaded here or
            // (PRE: one iteration) were loaded with those of up to 3 more of
            // the pixel's samples, so that their loads overlap
            auto sample = [&](uint32_t j, auto pre, const float4* Q, uint2 Hh) {
                constexpr bool PRE = decltype(pre)::value;
                size_t sidx = (size_t)pixel * A.spp + j;
                const float4* rec = reinterpret_cast<const float4*>(A.samples + sidx * K * A.rec_words);
                // the light and key (z, w): in the last float4 of a full
                // record, the third of a keyed one, none in an 8-word one
                // (Q[3] holds them in every case)
                float4 r0 = PRE ? Q[0] : rec[0];
                float4 r3 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (PRE) {
                    r3 = Q[3];
                } else if (rw4 == 4u) {
                    r3 = rec[3];
                } else if (rw4 == 3u) {
                    const float4 t = rec[2];
                    r3 = make_float4(0.0f, 0.0f, t.x, t.y);
                }
                uint32_t flags = __float_as_uint(r0.y);
                if (!(flags & kFlagContributes)) {
                    if (A.rec_cluster) {
                        A.rec_cluster[sidx] = 0xFFFFFFFFu;
                        A.rec_y[sidx] = 0.0f;
                    }
                    return;
                }
                // joint visibility: the sample counts only if none of its
                // shadow rays is occluded
                bool all_vis = true;
                if (flags & kFlagJoint) {
                    for (uint32_t it = 0; it < K; ++it) {
                        const uint32_t fl = __float_as_uint((PRE && it == 0u) ? Q[0].y : rec[rw4 * it].y);
#pragma unroll
                        for (int t = 0; t < 2; ++t) {
                            const uint32_t h = (PRE && it == 0u) ? (t == 0 ? Hh.x : Hh.y) : A.hits[(sidx * K + it) * 2 + t];
                            if (((fl >> (4 * t)) & 15u) == 5u && h != 0u) all_vis = false;
                        }
                    }
                }
                f3 result = mk3(0.0f, 0.0f, 0.0f);
                float* dbg = A.debug_samples ? A.debug_samples + sidx * RISLTC_SAMPLE_RECORD_FLOATS : nullptr;
                uint32_t traced = 0, unoccluded = 0;
                // iterations, then techniques: the reference's summation order
                for (uint32_t it = 0; it < K; ++it) {
                    const float4* ri = rec + rw4 * it;
                    const bool pq = PRE && it == 0u;
                    float4 q0 = pq ? Q[0] : ri[0], q1 = pq ? Q[1] : ri[1], q2, q3;
                    const uint32_t fl = __float_as_uint(q0.y);
                    if (compact) {
                        // the occluded estimates: +0, NaN where flagged
                        float e[6];
#pragma unroll
                        for (int b = 0; b < 6; ++b)
                            e[b] = ((fl >> (kFlagOccShift + b)) & 1u) ? __builtin_nanf("") : 0.0f;
                        q2 = make_float4(e[0], e[1], e[2], e[3]);
                        q3 = make_float4(e[4], e[5], 0.0f, 0.0f);
                    } else {
                        q2 = pq ? Q[2] : ri[2];
                        q3 = pq ? Q[3] : ri[3];
                    }
                    f3 ev[2] = {mk3(q0.z, q0.w, q1.x), mk3(q1.y, q1.z, q1.w)};
                    f3 eo[2] = {mk3(q2.x, q2.y, q2.z), mk3(q2.w, q3.x, q3.y)};
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        uint32_t st = (fl >> (4 * t)) & 15u;
                        if (st == 0u) continue;
                        bool hit = st == 5u && (pq ? (t == 0 ? Hh.x : Hh.y) : A.hits[(sidx * K + it) * 2 + t]) != 0u;
                        traced += st == 5u ? 1u : 0u;
                        unoccluded += (st == 5u && !hit) ? 1u : 0u;
                        bool vis = (fl & kFlagJoint) ? all_vis : (st == 4u || (st == 5u && !hit));
                        result = add3(result, vis ? ev[t] : eo[t]);
                        if (dbg && it == 0u && st == 5u) dbg[16 + t] = __uint_as_float(hit ? 2u : 3u);
                    }
                }
                // result * (1 / SAMPLE_COUNT) (shading_pass.frag.glsl:396, 559)
                f3 col = scl3(result, 1.0f / (float)K);
                if (dbg) {
                    dbg[30] = col.x;
                    dbg[31] = col.y;
                    dbg[32] = col.z;
                }
                float Wt = r0.x;
                float s = (float)A.spp;
                fin = add3(fin, mk3((col.x * Wt) / s, (col.y * Wt) / s, (col.z * Wt) / s));
                if (A.rec_cluster || A.cache.counts) {
                    int32_t li = __float_as_int(r3.z);
                    bool ok = li >= 0 && traced > 0u;
                    uint32_t cl = li >= 0 ? (uint32_t)A.cluster_of_light[li] : 0xFFFFFFFFu;
                    if (A.rec_cluster) {
                        A.rec_cluster[sidx] = cl;
                        A.rec_y[sidx] = ok ? (float)unoccluded / (float)traced : 0.0f;
                    }
                    if (A.cache.counts && ok) {
                        // (unoccluded, occluded) in one 64-bit add
                        const unsigned long long add =
                            (unsigned long long)unoccluded | ((unsigned long long)(traced - unoccluded) << 32);
                        const uint32_t slot = A.cache.per_light ? (uint32_t)li : cl;
                        pkey = __float_as_uint(r3.w);
                        bool merged = false;
#pragma unroll
                        for (uint32_t q = 0; q < 4u; ++q)
                            if (q < np && pslot[q] == slot) {
                                padd[q] += add;
                                merged = true;
                            }
                        if (!merged) {
                            if (np == 4u) {
                                flush_counts(A.cache, pkey, pslot, padd, np);
                                np = 0u;
                            }
#pragma unroll
                            for (uint32_t q = 0; q < 4u; ++q)
                                if (q == np) {
                                    pslot[q] = slot;
                                    padd[q] = add;
                                }
                            ++np;
                        }
                    }
                }
                        };
            if (K == 1u && (GROUPED || A.spp <= 4u)) {
                // groups of up to 4 samples, each group's loads issued
                // together; the samples are summed in order
                const uint32_t groups = GROUPED ? (A.spp + 3u) / 4u : 1u;
                for (uint32_t g = 0; g < groups; ++g) {
                    const uint32_t j0 = 4u * g;
                    float4 Q[4][4];
                    uint2 HH[4];
#pragma unroll
                    for (uint32_t jj = 0; jj < 4u; ++jj) {
                        if (j0 + jj < A.spp) {
                            const size_t sidx = (size_t)pixel * A.spp + j0 + jj;
                            const float4* rec = reinterpret_cast<const float4*>(A.samples + sidx * A.rec_words);
#pragma unroll
                            for (int k = 0; k < 4; ++k)
                                Q[jj][k] = ((uint32_t)k < rw4) ? rec[k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                            if (rw4 == 3u) Q[jj][3] = make_float4(0.0f, 0.0f, Q[jj][2].x, Q[jj][2].y);
                            HH[jj] = reinterpret_cast<const uint2*>(A.hits)[sidx];
                        }
                    }
#pragma unroll
                    for (uint32_t jj = 0; jj < 4u; ++jj)
                        if (j0 + jj < A.spp) sample(j0 + jj, std::true_type{}, Q[jj], HH[jj]);
                }
            } else {
                for (uint32_t j = 0; j < A.spp; ++j) sample(j, std::false_type{}, nullptr, make_uint2(0u, 0u));
            }
            if (np > 0u) {
                flush_counts(A.cache, pkey, pslot, padd, np);
                if (A.cache.dirty[pkey] == 0u) A.cache.dirty[pkey] = 1u;
            }
        }
        if (__builtin_isnan(fin.x) || __builtin_isnan(fin.y) || __builtin_isnan(fin.z) || __builtin_isinf(fin.x) ||
            __builtin_isinf(fin.y) || __builtin_isinf(fin.z))
            fin = mk3(1.0f / A.exposure, 0.0f / A.exposure, 0.8f / A.exposure);
        out = make_float4(fin.x * A.exposure, fin.y * A.exposure, fin.z * A.exposure, 1.0f);
    }
    reinterpret_cast<float4*>(A.out)[pixel] = out;
}

template <bool GROUPED>
__global__ void __launch_bounds__(256) resolve_kernel(ResolveArgs A) {
    resolve_pixel<GROUPED>(A, blockIdx.x * blockDim.x + threadIdx.x);
}

template <bool GROUPED>
__global__ void __launch_bounds__(256) resolve_batch_kernel(const ResolveArgs* views, const uint32_t* block_offsets,
                                                            uint32_t count) {
    const uint32_t v = view_of(block_offsets, count, blockIdx.x);
    resolve_pixel<GROUPED>(views[v], (blockIdx.x - block_offsets[v]) * blockDim.x + threadIdx.x);
}

// accum_pass.frag.glsl:45-53: out = fma(prev, n, cur) * (1 / (n + 1))
__global__ void accumulate_kernel(float* accum, const float* frame, uint32_t count, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count * 4u) return;
    float fn = (float)n;
    float r = fma_(accum[i], fn, frame[i]);
    accum[i] = fma_(r, 1.0f / (float)(n + 1u), 0.0f);
}

uint32_t shade_tiles(const ShadeArgs& a) {
    const uint32_t W = a.c.viewport_size[0], H = a.c.viewport_size[1];
    const uint32_t ww = 1u << a.tile_w_log2, th = 4u * (64u >> a.tile_w_log2);
    return ((W + ww - 1u) / ww) * ((H + th - 1u) / th);
}

size_t shade_lds_bytes(const ShadeArgs& a) {
    const uint32_t n = a.scene.light_count < a.lds_lights ? a.scene.light_count : a.lds_lights;
    const uint32_t tables = lds_table_lights(a.learned_pass, a.scene.light_count, a.clusters.count);
    // learned pass with a cache: staged rows, kRowsPerWave per wave
    const bool rows = a.learned_pass && a.cache.vis && a.cache.slots <= kRowLdsMax;
    // then the wave's pixels and deal tables for the compacted evaluation
    return px_lds_offset(n, a.learned_pass, tables, a.clusters.count, rows) + kPxLdsBytes + kDealLdsBytes +
           (a.learned_pass ? 0u : kJumpLdsBytes);
}

static bool shade_args_ok(const ShadeArgs& a) {
    return a.tile_w_log2 >= 3u && a.tile_w_log2 <= 6u && a.chunk >= 1u && a.chunk <= 32u &&
           a.lds_lights <= kLdsSphereMax;
}

hipError_t launch_shade(const ShadeArgs& a, bool exact, hipStream_t s) {
    if (!shade_args_ok(a)) return hipErrorInvalidValue;
    dim3 grid(shade_tiles(a)), block(256);
    bool dbg = a.debug_samples != nullptr || a.debug_cands != nullptr || a.iters > 1u;
    size_t lds = shade_lds_bytes(a);
    const int mode = a.learned_pass ? 1 : 0;
#define RISLTC_SHADE_LAUNCH(XX, DD, MM) hipLaunchKernelGGL((shade_kernel<XX, DD, MM>), grid, block, lds, s, a)
    if (exact) {
        if (dbg) {
            if (mode == 1) RISLTC_SHADE_LAUNCH(true, true, 1);
            else RISLTC_SHADE_LAUNCH(true, true, 0);
        } else {
            if (mode == 1) RISLTC_SHADE_LAUNCH(true, false, 1);
            else RISLTC_SHADE_LAUNCH(true, false, 0);
        }
    } else {
        if (dbg) {
            if (mode == 1) RISLTC_SHADE_LAUNCH(false, true, 1);
            else RISLTC_SHADE_LAUNCH(false, true, 0);
        } else {
            if (mode == 1) RISLTC_SHADE_LAUNCH(false, false, 1);
            else RISLTC_SHADE_LAUNCH(false, false, 0);
        }
    }
#undef RISLTC_SHADE_LAUNCH
    return hipGetLastError();
}

hipError_t launch_shade_baseline(const ShadeArgs& a, uint32_t technique, bool exact, hipStream_t s) {
    // (the baseline kernels write full records only)
    if (!shade_args_ok(a) || a.rec_words != (uint32_t)kSampleWords) return hipErrorInvalidValue;
    dim3 grid(shade_tiles(a)), block(256);
    if (technique == RISLTC_SAMPLE_POLYGON_AREA_TURK) {
        if (exact) hipLaunchKernelGGL((shade_baseline_kernel<true, RISLTC_SAMPLE_POLYGON_AREA_TURK>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((shade_baseline_kernel<false, RISLTC_SAMPLE_POLYGON_AREA_TURK>), grid, block, 0, s, a);
    } else if (technique == RISLTC_SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE) {
        if (exact)
            hipLaunchKernelGGL((shade_baseline_kernel<true, RISLTC_SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE>), grid, block, 0,
                               s, a);
        else
            hipLaunchKernelGGL((shade_baseline_kernel<false, RISLTC_SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE>), grid, block,
                               0, s, a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_shade_batch(const ShadeArgs* d_views, const uint32_t* d_tile_offsets, uint32_t count,
                              uint32_t total_tiles, size_t lds, bool exact, int mode, bool general, hipStream_t s) {
    if (count == 0 || total_tiles == 0) return hipSuccess;
    dim3 grid(total_tiles), block(256);
#define RISLTC_BATCH_LAUNCH(XX, MM) \
    do { \
        if (general) \
            hipLaunchKernelGGL((shade_batch_kernel<XX, MM, true>), grid, block, lds, s, d_views, d_tile_offsets, count); \
        else \
            hipLaunchKernelGGL((shade_batch_kernel<XX, MM, false>), grid, block, lds, s, d_views, d_tile_offsets, count); \
    } while (0)
    if (exact) {
        if (mode == 1) RISLTC_BATCH_LAUNCH(true, 1);
        else RISLTC_BATCH_LAUNCH(true, 0);
    } else {
        if (mode == 1) RISLTC_BATCH_LAUNCH(false, 1);
        else RISLTC_BATCH_LAUNCH(false, 0);
    }
#undef RISLTC_BATCH_LAUNCH
    return hipGetLastError();
}

hipError_t launch_resolve_batch(const ResolveArgs* d_views, const uint32_t* d_block_offsets, uint32_t count,
                                uint32_t total_blocks, bool grouped, hipStream_t s) {
    if (count == 0 || total_blocks == 0) return hipSuccess;
    if (grouped)
        hipLaunchKernelGGL(resolve_batch_kernel<true>, dim3(total_blocks), dim3(256), 0, s, d_views, d_block_offsets,
                           count);
    else
        hipLaunchKernelGGL(resolve_batch_kernel<false>, dim3(total_blocks), dim3(256), 0, s, d_views, d_block_offsets,
                           count);
    return hipGetLastError();
}

hipError_t launch_resolve(const ResolveArgs& a, hipStream_t s) {
    uint32_t n = a.width * a.height;
    if (resolve_grouped(a))
        hipLaunchKernelGGL(resolve_kernel<true>, dim3((n + 255u) / 256u), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(resolve_kernel<false>, dim3((n + 255u) / 256u), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_accumulate(float* accum, const float* frame, uint32_t count, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(accumulate_kernel, dim3((count * 4u + 255u) / 256u), dim3(256), 0, s, accum, frame, count, n);
    return hipGetLastError();
}

// one work-group per cell: skips clean cells, recomputes the bytes of a dirty
// one, then clears its flag
__global__ void __launch_bounds__(256) refresh_cache_kernel(VisCacheView C) {
    const uint32_t cell = blockIdx.x;
    if (C.dirty[cell] == 0u) return;
    const size_t base = (size_t)cell * C.slots;
    for (uint32_t k = threadIdx.x; k < C.slots; k += blockDim.x)
        C.vis[base + k] = (uint8_t)cache_byte(C.counts[base + k], C.prior);
    __syncthreads();
    if (threadIdx.x == 0) C.dirty[cell] = 0u;
}

hipError_t launch_refresh_cache(const VisCacheView& c, hipStream_t s) {
    if (!c.counts || !c.vis || !c.dirty) return hipErrorInvalidValue;
    hipLaunchKernelGGL(refresh_cache_kernel, dim3(c.mask + 1u), dim3(256), 0, s, c);
    return hipGetLastError();
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s) {
    if (n == 0u) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s) { return launch_fill_u32(p, n, 0u, s); }

hipError_t upload_clip_table(const uint8_t* table, size_t bytes) {
    return hipMemcpyToSymbol(HIP_SYMBOL(c_clip_rot), table, bytes, 0, hipMemcpyHostToDevice);
}

}  // namespace rl
