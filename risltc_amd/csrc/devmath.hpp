// devmath.hpp -- device arithmetic for the gfx950 kernels.
//
// Every kernel is compiled with -ffp-contract=off: a product followed by a
// sum is two roundings unless __builtin_fmaf is written. The operation order
// of each helper is a project convention shared with the CPU oracle, so the
// exact variant (X = true: IEEE division and square root) reproduces its
// FP32 results bit for bit. The fast variant (X = false) replaces division,
// reciprocal and (inverse) square root by the hardware approximations
// v_rcp_f32 / v_rsq_f32 / v_sqrt_f32 (about 1 ulp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rl {

struct f2 {
    float x, y;
};
struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f2 mk2(float x, float y) { return f2{x, y}; }
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ float fabs_(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ float fmin_(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ float fmax_(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ float clamp_(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
__device__ __forceinline__ bool sign_bit(float x) { return (__float_as_uint(x) >> 31) != 0u; }

// ---- exact / fast policy --------------------------------------------------
template <bool X>
__device__ __forceinline__ float sqrt_(float x) {
    if constexpr (X) return __builtin_sqrtf(x);
    else return __builtin_amdgcn_sqrtf(x);
}
template <bool X>
__device__ __forceinline__ float rcp_(float x) {
    if constexpr (X) return 1.0f / x;
    else return __builtin_amdgcn_rcpf(x);
}
template <bool X>
__device__ __forceinline__ float rsq_(float x) {
    if constexpr (X) return 1.0f / __builtin_sqrtf(x);
    else return __builtin_amdgcn_rsqf(x);
}
template <bool X>
__device__ __forceinline__ float div_(float a, float b) {
    if constexpr (X) return a / b;
    else return a * __builtin_amdgcn_rcpf(b);
}

// ---- vectors ----------------------------------------------------------------
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 mul3(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 scl3(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 sclf3(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ f2 add2(f2 a, f2 b) { return mk2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ f2 sub2(f2 a, f2 b) { return mk2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ f2 scl2(f2 a, float s) { return mk2(a.x * s, a.y * s); }
__device__ __forceinline__ f2 sclf2(float s, f2 a) { return mk2(s * a.x, s * a.y); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return fma_(a.z, b.z, fma_(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ float dot2(f2 a, f2 b) { return fma_(a.y, b.y, a.x * b.x); }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <bool X>
__device__ __forceinline__ f3 normalize3(f3 a) {
    return scl3(a, rsq_<X>(dot3(a, a)));
}
template <bool X>
__device__ __forceinline__ f2 normalize2(f2 a) {
    return scl2(a, rsq_<X>(dot2(a, a)));
}

// ---- project transcendentals (Cephes single precision, S. L. Moshier) -----
// Same algorithms and operation order as the oracle's prj_* functions.
template <bool X>
__device__ __forceinline__ float atan_(float xx) {
    float x = fabs_(xx);
    float y;
    if (x > 2.414213562373095f) {
        y = 1.5707963267948966f;
        x = -rcp_<X>(x);
    } else if (x > 0.4142135623730950f) {
        y = 0.7853981633974483f;
        x = div_<X>(x - 1.0f, x + 1.0f);
    } else {
        y = 0.0f;
    }
    float z = x * x;
    float p = fma_(fma_(fma_(8.05374449538e-2f, z, -1.38776856032e-1f), z, 1.99777106478e-1f), z,
                   -3.33329491539e-1f);
    y = y + fma_(p * z, x, x);
    return (xx < 0.0f) ? -y : y;
}
__device__ __forceinline__ float asin_(float xx) {
    float a = fabs_(xx);
    if (!(a <= 1.0f)) return __builtin_nanf("");
    float x, z;
    bool flag;
    if (a > 0.5f) {
        z = 0.5f * (1.0f - a);
        x = __builtin_sqrtf(z);
        flag = true;
    } else {
        x = a;
        z = x * x;
        flag = false;
    }
    float r;
    if (a < 1.0e-4f) {
        r = x;
    } else {
        float p = fma_(fma_(fma_(fma_(4.2163199048e-2f, z, 2.4181311049e-2f), z, 4.5470025998e-2f), z,
                            7.4953002686e-2f),
                       z, 1.6666752422e-1f);
        r = fma_(p * z, x, x);
    }
    if (flag) {
        r = r + r;
        r = 1.5707963267948966f - r;
    }
    return (xx < 0.0f) ? -r : r;
}

__device__ __forceinline__ float acos_(float x) {
    if (x < -0.5f) return 3.14159265358979f - 2.0f * asin_(__builtin_sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * asin_(__builtin_sqrtf(0.5f * (1.0f - x)));
    return 1.5707963267948966f - asin_(x);
}

__device__ __forceinline__ float sin_poly(float z, float x) {
    float p = fma_(fma_(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f);
    return fma_(p * z, x, x);
}
__device__ __forceinline__ float cos_poly(float z) {
    float p = fma_(fma_(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f);
    return fma_(p * z, z, fma_(-0.5f, z, 1.0f));
}
__device__ __forceinline__ float reduce_pio4(float x, uint32_t* octant) {
    uint32_t j = (uint32_t)(1.27323954473516f * x);
    float y = (float)j;
    if (j & 1u) {
        j += 1u;
        y += 1.0f;
    }
    *octant = j & 7u;
    x = fma_(-y, 0.78515625f, x);
    x = fma_(-y, 2.4187564849853515625e-4f, x);
    x = fma_(-y, 3.77489497744594108e-8f, x);
    return x;
}
__device__ __forceinline__ float sin_(float xx) {
    float x = fabs_(xx);
    if (!(x <= 1.0e5f)) return xx - xx;
    bool negative = xx < 0.0f;
    uint32_t j;
    x = reduce_pio4(x, &j);
    if (j > 3u) {
        negative = !negative;
        j -= 4u;
    }
    float z = x * x;
    float r = (j == 1u || j == 2u) ? cos_poly(z) : sin_poly(z, x);
    return negative ? -r : r;
}
__device__ __forceinline__ float cos_(float xx) {
    float x = fabs_(xx);
    if (!(x <= 1.0e5f)) return xx - xx;
    bool negative = false;
    uint32_t j;
    x = reduce_pio4(x, &j);
    if (j > 3u) {
        j -= 4u;
        negative = !negative;
    }
    if (j > 1u) negative = !negative;
    float z = x * x;
    float r = (j == 1u || j == 2u) ? sin_poly(z, x) : cos_poly(z);
    return negative ? -r : r;
}

// ---- noise (src/shaders/noise_utility.glsl:26-95) ---------------------------
__device__ __forceinline__ uint32_t murmur_mix(uint32_t hash, uint32_t k) {
    k *= 0xcc9e2d51u;
    k = (k << 15) | (k >> 17);
    k *= 0x1b873593u;
    hash ^= k;
    hash = ((hash << 13) | (hash >> 19)) * 5u + 0xe6546b64u;
    return hash;
}
__device__ __forceinline__ uint32_t murmur_finalize(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
__device__ __forceinline__ float noise_next(uint32_t& seed) {
    seed = 1664525u * seed + 1013904223u;
    return (float)seed * (1.0f / 4294967296.0f);
}

}  // namespace rl
