// lbvh.hip -- linear BVH over the scene triangles and its stackless traversal.
//
// The reference traces shadow rays with VK_KHR_ray_query against a hardware
// acceleration structure (src/shaders/shading_pass.frag.glsl:112-129,
// src/scene.c:142-406). gfx950 has no ray-query hardware, so the project
// builds its own hierarchy:
//   1. triangles dequantised exactly as the shader decodes positions
//      (decode_position_64_bit, mesh_quantization.glsl:38-45);
//   2. 30-bit Morton codes of the centroids in the centroid bounds (the
//      bit interleave of tools/io_export_vulkan_blender28.py:49-77), keys
//      (code << 32 | triangle id) radix-sorted -- unique keys, so the order
//      is deterministic;
//   3. Karras (2012) hierarchy: internal node i splits the sorted range at
//      the highest differing key bit; leaves are node ids n-1 .. 2n-2;
//   4. bottom-up boxes with one atomic arrival counter per internal node;
//   5. the nodes written in depth-first pre-order (left child first) with
//      escape links, for stackless traversal: a box hit continues at the
//      next record in memory, so a descent reads consecutive cache lines.
// The CPU oracle builds the same tree from the same keys; traversal uses the
// same box and Moeller-Trumbore tests in the same operation order, so hit and
// miss flags agree bit for bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "devmath.hpp"
#include "kernels.hpp"

namespace rl {

__device__ __forceinline__ uint32_t part_1_by_2(uint32_t x) {
    x &= 0x000003ffu;
    x = (x ^ (x << 16)) & 0xff0000ffu;
    x = (x ^ (x << 8)) & 0x0300f00fu;
    x = (x ^ (x << 4)) & 0x030c30c3u;
    x = (x ^ (x << 2)) & 0x09249249u;
    return x;
}

__global__ void dequant_kernel(const uint2* pos, uint64_t n, float3 f, float3 s, float* tris, float* cent) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float v[9];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint2 q = pos[t * 3 + k];
        uint32_t qx = q.x & 0x1FFFFFu;
        uint32_t qy = ((q.x & 0xFFE00000u) >> 21) | ((q.y & 0x3FFu) << 11);
        uint32_t qz = (q.y & 0x7FFFFC00u) >> 10;
        v[3 * k + 0] = fma_((float)qx, f.x, s.x);
        v[3 * k + 1] = fma_((float)qy, f.y, s.y);
        v[3 * k + 2] = fma_((float)qz, f.z, s.z);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) tris[t * 9 + k] = v[k];
#pragma unroll
    for (int a = 0; a < 3; ++a) cent[t * 3 + a] = ((v[a] + v[3 + a]) + v[6 + a]) / 3.0f;
}

// min/max of the centroids: block partials, then one block over partials.
// min and max are exact, so any reduction order gives the same bounds.
__global__ void bounds_kernel(const float* cent, uint64_t n, float* partial) {
    __shared__ float sh[6][256];
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float c = cent[t * 3 + a];
            lo[a] = fmin_(lo[a], c);
            hi[a] = fmax_(hi[a], c);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sh[a][threadIdx.x] = lo[a];
        sh[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                sh[a][threadIdx.x] = fmin_(sh[a][threadIdx.x], sh[a][threadIdx.x + w]);
                sh[3 + a][threadIdx.x] = fmax_(sh[3 + a][threadIdx.x], sh[3 + a][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int a = 0; a < 6; ++a) partial[blockIdx.x * 6 + a] = sh[a][0];
}

__global__ void bounds_final_kernel(const float* partial, uint32_t blocks, float* out) {
    if (threadIdx.x != 0) return;
    float r[6] = {__builtin_inff(), __builtin_inff(), __builtin_inff(), -__builtin_inff(), -__builtin_inff(),
                  -__builtin_inff()};
    for (uint32_t b = 0; b < blocks; ++b)
        for (int a = 0; a < 3; ++a) {
            r[a] = fmin_(r[a], partial[b * 6 + a]);
            r[3 + a] = fmax_(r[3 + a], partial[b * 6 + 3 + a]);
        }
    for (int a = 0; a < 6; ++a) out[a] = r[a];
}

__global__ void morton_kernel(const float* cent, uint64_t n, const float* bounds, uint64_t* keys) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t code = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float ext = bounds[3 + a] - bounds[a];
        float factor = (ext > 0.0f) ? 1024.0f / ext : 0.0f;
        float summand = -bounds[a] * factor;
        float q = cent[t * 3 + a] * factor + summand;
        q = fmin_(fmax_(q, 0.0f), 1023.0f);
        code |= part_1_by_2((uint32_t)q) << a;
    }
    keys[t] = ((uint64_t)code << 32) | (uint64_t)t;
}

__device__ __forceinline__ int kdelta(const uint64_t* keys, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    return __clzll((long long)(keys[i] ^ keys[j]));
}

__global__ void karras_kernel(const uint64_t* keys, int n, int* left, int* right, int* parent, int* count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = kdelta(keys, n, i, i - d);
    int lmax = 2;
    while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = kdelta(keys, n, i, j);
    int s = 0;
    for (int div = 2;; div *= 2) {
        int t = (l + div - 1) / div;
        if (kdelta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    int gamma = i + s * d + (d < 0 ? d : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    int lc = (lo == gamma) ? (n - 1 + gamma) : gamma;
    int rc = (hi == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
    left[i] = lc;
    right[i] = rc;
    parent[lc] = i;
    parent[rc] = i;
    count[i] = hi - lo + 1;
}

__global__ void leaf_kernel(const uint64_t* keys, const float* tris, int n, float* box, int* left, int* right,
                            int* count, uint32_t* prim) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t t = (uint32_t)(keys[k] & 0xFFFFFFFFull);
    const float* v = tris + (size_t)t * 9;
    const int x = n - 1 + k;
    float* b = box + (size_t)x * 6;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b[a] = fmin_(fmin_(v[a], v[3 + a]), v[6 + a]);
        b[3 + a] = fmax_(fmax_(v[a], v[3 + a]), v[6 + a]);
    }
    left[x] = -1;
    right[x] = -1;
    count[x] = 1;
    prim[x] = t;
}

__global__ void refit_kernel(int n, const int* left, const int* right, const int* parent, float* box,
                             uint32_t* arrivals) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int node = n - 1 + k;
    int p = parent[node];
    volatile float* vb = box;
    // a well-formed tree has depth < n; the cap keeps a malformed one finite
    for (int depth = 0; p >= 0 && depth < n; ++depth) {
        __threadfence();
        uint32_t old = atomicAdd(&arrivals[p], 1u);
        if (old == 0u) return;  // the sibling subtree is not done yet
        int l = left[p], r = right[p];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float lo = fmin_(vb[(size_t)l * 6 + a], vb[(size_t)r * 6 + a]);
            float hi = fmax_(vb[(size_t)l * 6 + 3 + a], vb[(size_t)r * 6 + 3 + a]);
            vb[(size_t)p * 6 + a] = lo;
            vb[(size_t)p * 6 + 3 + a] = hi;
        }
        node = p;
        p = parent[node];
    }
}

// Writes node x of a binary tree at its pre-order unit (left child first).
// Before x in pre-order come its ancestors (2 units each) and, for every
// ancestor whose right subtree holds x, that ancestor's whole left subtree
// (a subtree of m leaves spans 5 m - 2 units). The escape unit of x is the
// first unit past its own subtree.
__global__ void layout_kernel(int nodes, const int* left, const int* right, const int* parent, const int* count,
                              const uint32_t* prim, const float* box, const float* tris, const uint2* pos,
                              float4* rec) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nodes) return;
    int depth = 0, skipped = 0;
    int y = x;
    for (int steps = 0; steps < nodes; ++steps) {
        int p = parent[y];
        if (p < 0) break;
        ++depth;
        if (right[p] == y) skipped += 5 * count[left[p]] - 2;
        y = p;
    }
    const int u = 2 * depth + skipped;
    if (left[x] >= 0) {
        int esc = u + 5 * count[x] - 2;
        const float* b = box + (size_t)x * 6;
        rec[u] = make_float4(b[0], b[1], b[2], __int_as_float(esc));
        rec[u + 1] = make_float4(b[3], b[4], b[5], 0.0f);
    } else {
        uint32_t t = prim[x];
        const float* v = tris + (size_t)t * 9;
        uint32_t light = pos[(size_t)t * 3].y & 0x80000000u;
        rec[u] = make_float4(v[0], v[1], v[2], __int_as_float(-1));
        rec[u + 1] = make_float4(v[3], v[4], v[5], __uint_as_float(t));
        rec[u + 2] = make_float4(v[6], v[7], v[8], __uint_as_float(light));
    }
}

__global__ void fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

namespace {
// device allocations of one build, released on every path
struct BuildScratch {
    void* p[20] = {};
    int k = 0;
    template <class T>
    hipError_t get(T** out, size_t bytes) {
        *out = nullptr;
        hipError_t e = hipMalloc((void**)out, bytes);
        if (e == hipSuccess) p[k++] = *out;
        return e;
    }
    ~BuildScratch() {
        for (int i = 0; i < k; ++i) hipFree(p[i]);
    }
};
}  // namespace

#define BVH_TRY(x)                         \
    do {                                   \
        hipError_t e_ = (x);               \
        if (e_ != hipSuccess) return e_;   \
    } while (0)

hipError_t build_bvh(const uint2* d_pos, uint64_t tri_count, const float deq_factor[3], const float deq_summand[3],
                     uint32_t builder, BvhBuild* out, hipStream_t s) {
    if (tri_count == 0 || tri_count > kMaxBvhTriangles) return hipErrorInvalidValue;
    if (builder != RISLTC_BVH_SAH && builder != RISLTC_BVH_LBVH) return hipErrorInvalidValue;
    BuildScratch w;
    const int n = (int)tri_count;
    const uint64_t nodes = 2 * tri_count - 1;
    const uint32_t tb = 256;
    const uint32_t tri_blocks = (uint32_t)((tri_count + tb - 1) / tb);
    const uint32_t node_blocks = (uint32_t)((nodes + tb - 1) / tb);
    float *tris, *cent, *box;
    int *left, *right, *parent, *count;
    uint32_t* prim;
    BVH_TRY(w.get(&tris, sizeof(float) * 9 * tri_count));
    BVH_TRY(w.get(&cent, sizeof(float) * 3 * tri_count));
    BVH_TRY(w.get(&box, sizeof(float) * 6 * nodes));
    BVH_TRY(w.get(&left, sizeof(int) * nodes));
    BVH_TRY(w.get(&right, sizeof(int) * nodes));
    BVH_TRY(w.get(&parent, sizeof(int) * nodes));
    BVH_TRY(w.get(&count, sizeof(int) * nodes));
    BVH_TRY(w.get(&prim, sizeof(uint32_t) * nodes));
    hipLaunchKernelGGL(dequant_kernel, dim3(tri_blocks), dim3(tb), 0, s, d_pos, tri_count,
                       make_float3(deq_factor[0], deq_factor[1], deq_factor[2]),
                       make_float3(deq_summand[0], deq_summand[1], deq_summand[2]), tris, cent);
    hipLaunchKernelGGL(fill_u32, dim3(node_blocks), dim3(tb), 0, s, reinterpret_cast<uint32_t*>(parent), nodes,
                       0xFFFFFFFFu);
    BVH_TRY(hipGetLastError());
    if (builder == RISLTC_BVH_SAH) {
        SahTree T;
        T.cent = cent;
        T.left = left;
        T.right = right;
        T.parent = parent;
        T.count = count;
        T.prim = prim;
        T.box = box;
        BVH_TRY(w.get(&T.tbox, sizeof(float) * 6 * tri_count));
        BVH_TRY(w.get(&T.idx, sizeof(uint32_t) * tri_count));
        BVH_TRY(w.get(&T.tmp, sizeof(uint32_t) * tri_count));
        BVH_TRY(w.get(&T.first, sizeof(int) * nodes));
        BVH_TRY(w.get(&T.node_counter, sizeof(uint32_t)));
        BVH_TRY(build_sah_tree(tris, (uint32_t)n, T, s));
    } else {
        const uint32_t red_blocks = 1024;
        float *partial, *bounds;
        uint64_t *keys, *keys_sorted;
        uint32_t* arrivals;
        void* temp;
        size_t temp_bytes = 0;
        BVH_TRY(w.get(&partial, sizeof(float) * 6 * red_blocks));
        BVH_TRY(w.get(&bounds, sizeof(float) * 6));
        BVH_TRY(w.get(&keys, sizeof(uint64_t) * tri_count));
        BVH_TRY(w.get(&keys_sorted, sizeof(uint64_t) * tri_count));
        BVH_TRY(w.get(&arrivals, sizeof(uint32_t) * tri_count));
        hipLaunchKernelGGL(bounds_kernel, dim3(red_blocks), dim3(256), 0, s, cent, tri_count, partial);
        hipLaunchKernelGGL(bounds_final_kernel, dim3(1), dim3(64), 0, s, partial, red_blocks, bounds);
        hipLaunchKernelGGL(morton_kernel, dim3(tri_blocks), dim3(tb), 0, s, cent, tri_count, bounds, keys);
        BVH_TRY(hipGetLastError());
        BVH_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, temp_bytes, keys, keys_sorted, n, 0, 64, s));
        BVH_TRY(w.get(&temp, temp_bytes));
        BVH_TRY(hipcub::DeviceRadixSort::SortKeys(temp, temp_bytes, keys, keys_sorted, n, 0, 64, s));
        hipLaunchKernelGGL(fill_u32, dim3(tri_blocks), dim3(tb), 0, s, arrivals, tri_count, 0u);
        if (n > 1)
            hipLaunchKernelGGL(karras_kernel, dim3(tri_blocks), dim3(tb), 0, s, keys_sorted, n, left, right, parent,
                               count);
        hipLaunchKernelGGL(leaf_kernel, dim3(tri_blocks), dim3(tb), 0, s, keys_sorted, tris, n, box, left, right,
                           count, prim);
        if (n > 1)
            hipLaunchKernelGGL(refit_kernel, dim3(tri_blocks), dim3(tb), 0, s, n, left, right, parent, box, arrivals);
        BVH_TRY(hipGetLastError());
    }
    const uint64_t units = 5 * tri_count - 2;
    float4* rec = nullptr;
    BVH_TRY(hipMalloc(&rec, sizeof(float4) * (units + 2)));
    // the two padding units after the last record read as zeros
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t*>(rec + units), 8, 0u);
    hipLaunchKernelGGL(layout_kernel, dim3(node_blocks), dim3(tb), 0, s, (int)nodes, left, right, parent, count, prim,
                       box, tris, d_pos, rec);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        hipFree(rec);
        return e;
    }
    out->rec = rec;
    out->units = (uint32_t)units;
    out->leaf_count = (uint32_t)n;
    return hipSuccess;
}
#undef BVH_TRY

void free_bvh(BvhBuild* b) {
    if (b->rec) hipFree(b->rec);
    b->rec = nullptr;
    b->units = 0;
    b->leaf_count = 0;
}

// ---- traversal -------------------------------------------------------------------
// four floats in global memory, for loads the compiler must emit as global
// (not flat) instructions
typedef float F4g __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) F4g GlobalF4;
__device__ __forceinline__ float4 global_f4(const GlobalF4* p, int i) {
    const F4g v = p[i];
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ bool ray_box(float3 o, float3 inv, float4 bmin, float4 bmax, float tmin, float tmax) {
    float t0 = (bmin.x - o.x) * inv.x, t1 = (bmax.x - o.x) * inv.x;
    tmin = fmax_(tmin, fmin_(t0, t1));
    tmax = fmin_(tmax, fmax_(t0, t1) * 1.0000004f);
    t0 = (bmin.y - o.y) * inv.y;
    t1 = (bmax.y - o.y) * inv.y;
    tmin = fmax_(tmin, fmin_(t0, t1));
    tmax = fmin_(tmax, fmax_(t0, t1) * 1.0000004f);
    t0 = (bmin.z - o.z) * inv.z;
    t1 = (bmax.z - o.z) * inv.z;
    tmin = fmax_(tmin, fmin_(t0, t1));
    tmax = fmin_(tmax, fmax_(t0, t1) * 1.0000004f);
    return tmin <= tmax;
}

__device__ __forceinline__ bool ray_tri(float3 o, float3 d, float4 a, float4 b, float4 c, float tmin, float tmax,
                                        float& t_out) {
    float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
    float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
    float px = d.y * e2z - d.z * e2y, py = d.z * e2x - d.x * e2z, pz = d.x * e2y - d.y * e2x;
    float det = fma_(e1z, pz, fma_(e1y, py, e1x * px));
    if (det == 0.0f) return false;
    float inv = 1.0f / det;
    float sx = o.x - a.x, sy = o.y - a.y, sz = o.z - a.z;
    float u = fma_(sz, pz, fma_(sy, py, sx * px)) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    float qx = sy * e1z - sz * e1y, qy = sz * e1x - sx * e1z, qz = sx * e1y - sy * e1x;
    float w = fma_(d.z, qz, fma_(d.y, qy, d.x * qx)) * inv;
    if (w < 0.0f || u + w > 1.0f) return false;
    float t = fma_(e2z, qz, fma_(e2y, qy, e2x * qx)) * inv;
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t;
    return true;
}

// Every step moves forward in pre-order (by 2, by 3 or to an escape unit
// past the node's subtree), so `units` steps bound the walk; reads of a
// record at unit x < units touch x .. x + 2, inside the padded array.
__device__ __forceinline__ bool any_hit(const BvhView& B, float3 o, float3 d, float tmin, float tmax) {
    float3 inv = make_float3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    const int end = (int)B.units;
    int x = 0;
    for (int steps = 0; x < end && steps < end; ++steps) {
        float4 a = B.rec[x], b = B.rec[x + 1], c = B.rec[x + 2];
        int w = __float_as_int(a.w);
        if (w < 0) {
            float t;
            if (ray_tri(o, d, a, b, c, tmin, tmax, t)) return true;
            x += 3;
        } else {
            x = ray_box(o, inv, a, b, tmin, tmax) ? x + 2 : w;
        }
    }
    return false;
}

__device__ __forceinline__ uint32_t closest_hit(const BvhView& B, float3 o, float3 d, float tmin, float tmax) {
    float3 inv = make_float3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    const int end = (int)B.units;
    int x = 0;
    uint32_t best = 0xFFFFFFFFu;
    for (int steps = 0; x < end && steps < end; ++steps) {
        float4 a = B.rec[x], b = B.rec[x + 1], c = B.rec[x + 2];
        int w = __float_as_int(a.w);
        if (w < 0) {
            float t;
            if (ray_tri(o, d, a, b, c, tmin, tmax, t)) {
                tmax = t;
                best = __float_as_uint(b.w) | __float_as_uint(c.w);
            }
            x += 3;
        } else {
            x = ray_box(o, inv, a, b, tmin, tmax) ? x + 2 : w;
        }
    }
    return best;
}

// Persistent any-hit traversal of the shadow-ray list. Each lane holds one
// ray; a lane whose ray is done takes a new one as soon as `refill` lanes of
// its wave are idle (one atomic per wave for all of them), so long rays do
// not hold a wave's other 63 lanes idle. The walk per ray is any_hit's.
// Every wave leaves the loop: each refill either hands out rays, each of
// which ends within `units` steps, or moves to the next slice, and after
// kTraceQueues slices there is nothing left to fetch.
__global__ void __launch_bounds__(256) trace_kernel(TraceArgs A) {
    uint32_t count = *A.ray_count;
    if (count > A.max_rays) count = A.max_rays;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    const float4* rec = nullptr;
    int end = 0;
    uint32_t q = blockIdx.x % kTraceQueues;
    uint32_t tried = 0;
    bool active = false;
    float3 o = make_float3(0.0f, 0.0f, 0.0f), d = o, inv = o;
    float tmin = 0.0f, tmax = 0.0f;
    int x = 0, steps = 0;
    uint32_t slot = 0;
    for (;;) {
        uint64_t idle = __ballot(!active);
        uint32_t n_idle = (uint32_t)__popcll(idle);
        if (tried < kTraceQueues && (n_idle >= A.refill || n_idle == 64u)) {
            uint32_t q0 = (uint32_t)(((uint64_t)count * q) / kTraceQueues);
            uint32_t q1 = (uint32_t)(((uint64_t)count * (q + 1u)) / kTraceQueues);
            uint32_t len = q1 - q0;
            uint32_t leader = (uint32_t)__ffsll((long long)idle) - 1u;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&A.queues[q * kQueueStride], n_idle);
            base = (uint32_t)__shfl((int)base, (int)leader);
            if (!active) {
                uint32_t j = base + (uint32_t)__popcll(idle & below);
                if (j < len) {
                    uint32_t i = q0 + j;
                    ShadowRay r = A.rays[i];
                    slot = A.ray_slots[i];
                    o = make_float3(r.org.x, r.org.y, r.org.z);
                    d = make_float3(r.dir.x, r.dir.y, r.dir.z);
                    inv = make_float3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
                    const BvhView bv = A.bvhs[__float_as_uint(r.org.w)];
                    rec = bv.rec;
                    end = (int)bv.units;
                    tmin = kShadowTMin;
                    tmax = r.dir.w;
                    x = 0;
                    steps = 0;
                    active = true;
                }
            }
            if (base + n_idle >= len) {
                q = (q + 1u) % kTraceQueues;
                ++tried;
            }
        }
        if (__ballot(active) == 0ull) {
            if (tried >= kTraceQueues) break;
            continue;
        }
#pragma unroll 1
        for (int it = 0; it < 8; ++it) {
            if (active) {
                // (global-memory loads: the hierarchy is device memory)
                const GlobalF4* g = (const GlobalF4*)rec;
                float4 a = global_f4(g, x), b = global_f4(g, x + 1), c = global_f4(g, x + 2);
                int w = __float_as_int(a.w);
                bool hit = false;
                if (w < 0) {
                    float t;
                    hit = ray_tri(o, d, a, b, c, tmin, tmax, t);
                    x += 3;
                } else {
                    x = ray_box(o, inv, a, b, tmin, tmax) ? x + 2 : w;
                }
                ++steps;
                if (hit || x >= end || steps >= end) {
                    A.hits[slot] = hit ? 1u : 0u;
                    active = false;
                }
            }
        }
    }
}

__global__ void __launch_bounds__(256) trace_rays_kernel(BvhView B, const float* rays, uint32_t n, uint32_t* hits) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4* r = reinterpret_cast<const float4*>(rays) + 2 * i;
    float4 a = r[0], b = r[1];
    hits[i] = any_hit(B, make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z), a.w, b.w) ? 1u : 0u;
}

__global__ void __launch_bounds__(256) primary_kernel(BvhView B, risltc_frame_constants_t c, uint32_t* out) {
    uint32_t W = c.viewport_size[0], H = c.viewport_size[1];
    uint32_t tiles_x = (W + 15u) / 16u;
    uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = (blockIdx.x % tiles_x) * 16u + (lane & 15u);
    uint32_t y = (blockIdx.x / tiles_x) * 16u + wave * 4u + (lane >> 4);
    if (x >= W || y >= H) return;
    const float(*P)[4] = c.pixel_to_ray_direction_world_space;
    float fx = (float)x, fy = (float)y;
    float3 d = make_float3(fma_(P[0][2], 1.0f, fma_(P[0][1], fy, P[0][0] * fx)),
                           fma_(P[1][2], 1.0f, fma_(P[1][1], fy, P[1][0] * fx)),
                           fma_(P[2][2], 1.0f, fma_(P[2][1], fy, P[2][0] * fx)));
    float3 o = make_float3(c.camera_position_world_space[0], c.camera_position_world_space[1],
                           c.camera_position_world_space[2]);
    out[y * W + x] = closest_hit(B, o, d, 0.0f, __builtin_inff());
}

// One resident grid: as many 256-lane work-groups as fit on the device at
// once, a multiple of the XCD count.
hipError_t trace_blocks(uint32_t* blocks) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trace_kernel, 256, 0);
    if (e != hipSuccess) return e;
    uint32_t b = (uint32_t)(cus > 0 ? cus : 1) * (uint32_t)(per_cu > 0 ? per_cu : 1);
    *blocks = (b + kTraceQueues - 1u) / kTraceQueues * kTraceQueues;
    return hipSuccess;
}

hipError_t launch_trace(const TraceArgs& a, hipStream_t s, uint32_t blocks) {
    if (blocks == 0 || blocks % kTraceQueues) return hipErrorInvalidValue;
    hipLaunchKernelGGL(trace_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_trace(const TraceArgs& a, hipStream_t s) {
    uint32_t blocks = 0;
    hipError_t e = trace_blocks(&blocks);
    if (e != hipSuccess) return e;
    return launch_trace(a, s, blocks);
}

hipError_t launch_trace_rays(const BvhView& bvh, const float* rays8, uint32_t n, uint32_t* hits, hipStream_t s) {
    hipLaunchKernelGGL(trace_rays_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, bvh, rays8, n, hits);
    return hipGetLastError();
}

hipError_t launch_primary(const BvhView& bvh, const uint2* positions, const risltc_frame_constants_t& c, uint32_t* out,
                          hipStream_t s) {
    (void)positions;
    uint32_t W = c.viewport_size[0], H = c.viewport_size[1];
    dim3 grid(((W + 15u) / 16u) * ((H + 15u) / 16u));
    hipLaunchKernelGGL(primary_kernel, grid, dim3(256), 0, s, bvh, c, out);
    return hipGetLastError();
}

}  // namespace rl
