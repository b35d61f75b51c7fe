// gbuffer.hpp -- decoding of the quantised *.vks vertex data on the device.
#pragma once
#include "devmath.hpp"

namespace rl {

// ---- G-buffer (shading_pass.frag.glsl:571-672), always exact ----------------
__device__ __forceinline__ f3 decode_normal(uint32_t qx, uint32_t qy) {
    const float factor = 2.0f * (65534.0f / 65535.0f);
    const float summand = -(32768.0f / 65535.0f) * factor;
    float ox = fma_((float)qx / 65535.0f, factor, summand);
    float oy = fma_((float)qy / 65535.0f, factor, summand);
    f3 n = mk3(ox, oy, 1.0f - fabs_(ox) - fabs_(oy));
    float sx = (ox >= 0.0f) ? 1.0f : -1.0f;
    float sy = (oy >= 0.0f) ? 1.0f : -1.0f;
    if (n.z < 0.0f) {
        float nx = (1.0f - fabs_(n.y)) * sx;
        float ny = (1.0f - fabs_(n.x)) * sy;
        n.x = nx;
        n.y = ny;
    }
    return normalize3<true>(n);
}

__device__ __forceinline__ f3 decode_position(uint2 q, const float* f, const float* s) {
    uint32_t qx = q.x & 0x1FFFFFu;
    uint32_t qy = ((q.x & 0xFFE00000u) >> 21) | ((q.y & 0x3FFu) << 11);
    uint32_t qz = (q.y & 0x7FFFFC00u) >> 10;
    return mk3(fma_((float)qx, f[0], s[0]), fma_((float)qy, f[1], s[1]), fma_((float)qz, f[2], s[2]));
}

}  // namespace rl
