// formats.cpp -- the reference's on-disk formats on either side of the
// shading pass: the *.vks scene reader, the quicksave camera/light reader and
// writer, and the Radiance RGBE (.hdr) image writer. Host code only; every
// reader returns RISLTC_E_FORMAT on a short or inconsistent file instead of
// reading past it.
#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/risltc/risltc.h"

namespace {

constexpr uint32_t kVksMarker = 0xabcabcu;  // src/scene.c:423
constexpr uint32_t kVksVersion = 1u;        // src/scene.c:423
constexpr uint32_t kVksEndMarker = 0xE0FE0Fu;  // src/scene.c:478
constexpr uint64_t kVksMaxTriangles = 1ull << 36;
constexpr uint64_t kVksMaxMaterials = 1ull << 24;
// Bytes of one light record in a quicksave: 20 floats and 2 uint32
// (POLYGONAL_LIGHT_QUICKSAVE_SIZE, src/polygonal_light.h).
constexpr size_t kQuicksaveLightBytes = 20 * sizeof(float) + 2 * sizeof(uint32_t);
// A quicksave vertex count beyond this is taken as a corrupt file.
constexpr uint32_t kQuicksaveMaxVertices = 1u << 16;

static_assert(offsetof(risltc_polygonal_light_t, texture_index) == kQuicksaveLightBytes,
              "the first 88 bytes of a light are its quicksave record");
static_assert(sizeof(risltc_camera_t) == 48, "first_person_camera_t is 48 bytes");

struct File {
    FILE* f = nullptr;
    explicit File(const char* path, const char* mode) : f(fopen(path, mode)) {}
    ~File() {
        if (f) fclose(f);
    }
    bool read(void* dst, size_t bytes) { return bytes == 0 || fread(dst, 1, bytes, f) == bytes; }
    bool skip(uint64_t bytes) {
        // fseek past the end succeeds, so check the position against the size.
        if (bytes == 0) return true;
        long here = ftell(f);
        if (here < 0 || fseek(f, 0, SEEK_END) != 0) return false;
        long end = ftell(f);
        if (end < 0 || (uint64_t)(end - here) < bytes) return false;
        return fseek(f, here + (long)bytes, SEEK_SET) == 0;
    }
    bool write(const void* src, size_t bytes) { return bytes == 0 || fwrite(src, 1, bytes, f) == bytes; }
};

// Per-triangle bytes of the three mesh buffers of create_mesh
// (src/scene.c:56-58), in file order.
constexpr uint64_t kPositionBytes = sizeof(uint32_t) * 2 * 3;
constexpr uint64_t kNormalUvBytes = sizeof(uint16_t) * 4 * 3;
constexpr uint64_t kMaterialBytes = sizeof(uint8_t);

}  // namespace

extern "C" int risltc_load_vks(const char* path, risltc_vks_info_t* info, uint32_t* positions,
                               uint16_t* normals_and_tex_coords, uint8_t* material_indices,
                               uint64_t triangle_capacity, char* names, uint64_t names_capacity) {
    if (!path || !info) return RISLTC_E_ARG;
    memset(info, 0, sizeof(*info));
    File fh(path, "rb");
    if (!fh.f) {
        fprintf(stderr, "risltc_load_vks: cannot open %s\n", path);
        return RISLTC_E_IO;
    }
    uint32_t head[2];
    if (!fh.read(head, sizeof(head)) || head[0] != kVksMarker || head[1] != kVksVersion) {
        fprintf(stderr, "risltc_load_vks: %s has no .vks header of a supported version\n", path);
        return RISLTC_E_FORMAT;
    }
    uint64_t counts[2];
    if (!fh.read(counts, sizeof(counts)) || !fh.read(info->dequantization_factor, 3 * sizeof(float)) ||
        !fh.read(info->dequantization_summand, 3 * sizeof(float)))
        return RISLTC_E_FORMAT;
    info->material_count = counts[0];
    info->triangle_count = counts[1];
    if (info->triangle_count == 0) {
        fprintf(stderr, "risltc_load_vks: %s declares no triangles\n", path);
        return RISLTC_E_FORMAT;
    }
    // counts no file of this format can hold (the byte sizes below would
    // overflow): a corrupt header
    if (info->triangle_count > kVksMaxTriangles || info->material_count > kVksMaxMaterials) return RISLTC_E_FORMAT;
    // Material names: uint64 length, then length bytes and a terminating NUL.
    bool want_mesh = positions || normals_and_tex_coords || material_indices;
    uint64_t name_bytes = 0;
    for (uint64_t i = 0; i < info->material_count; ++i) {
        uint64_t len = 0;
        if (!fh.read(&len, sizeof(len)) || len > (1ull << 20)) return RISLTC_E_FORMAT;
        if (names && name_bytes + len + 1 <= names_capacity) {
            if (!fh.read(names + name_bytes, len + 1)) return RISLTC_E_FORMAT;
            names[name_bytes + len] = '\0';
        } else if (!fh.skip(len + 1)) {
            return RISLTC_E_FORMAT;
        }
        name_bytes += len + 1;
    }
    info->material_name_bytes = name_bytes;
    if (names && name_bytes > names_capacity) return RISLTC_E_ARG;
    const uint64_t tri = info->triangle_count;
    if (!want_mesh) {
        // Header query: still confirm that the mesh and end marker are there.
        if (!fh.skip(tri * (kPositionBytes + kNormalUvBytes + kMaterialBytes))) return RISLTC_E_FORMAT;
    } else {
        if (!positions || !normals_and_tex_coords || !material_indices || triangle_capacity < tri)
            return RISLTC_E_ARG;
        if (!fh.read(positions, tri * kPositionBytes) || !fh.read(normals_and_tex_coords, tri * kNormalUvBytes) ||
            !fh.read(material_indices, tri * kMaterialBytes))
            return RISLTC_E_FORMAT;
    }
    uint32_t eof_marker = 0;
    if (!fh.read(&eof_marker, sizeof(eof_marker)) || eof_marker != kVksEndMarker) {
        fprintf(stderr, "risltc_load_vks: %s: no end marker after the mesh data (truncated or corrupt)\n", path);
        return RISLTC_E_FORMAT;
    }
    return RISLTC_OK;
}

extern "C" int risltc_load_quicksave(const char* path, risltc_camera_t* camera, risltc_quicksave_info_t* info,
                                     risltc_polygonal_light_t* lights, uint32_t light_capacity, float* vertex_storage,
                                     uint64_t vertex_storage_floats, char* path_storage, uint64_t path_capacity) {
    if (!path || !info) return RISLTC_E_ARG;
    memset(info, 0, sizeof(*info));
    File fh(path, "rb");
    if (!fh.f) {
        fprintf(stderr, "risltc_load_quicksave: cannot open %s\n", path);
        return RISLTC_E_IO;
    }
    risltc_camera_t cam;
    uint32_t legacy_count = 0, light_count = 0;
    if (!fh.read(&cam, sizeof(cam)) || !fh.read(&legacy_count, sizeof(legacy_count)) ||
        !fh.read(&light_count, sizeof(light_count)))
        return RISLTC_E_FORMAT;
    if (lights && (light_capacity < light_count || !vertex_storage)) return RISLTC_E_ARG;
    uint64_t vertex_floats = 0, path_bytes = 0;
    for (uint32_t i = 0; i < light_count; ++i) {
        risltc_polygonal_light_t rec;
        memset(&rec, 0, sizeof(rec));
        if (!fh.read(&rec, kQuicksaveLightBytes)) return RISLTC_E_FORMAT;
        if (rec.vertex_count == 0 || rec.vertex_count > kQuicksaveMaxVertices) return RISLTC_E_FORMAT;
        // Legacy files have no scaling_y (src/main.c:101).
        if (rec.scaling_y <= 0.0f) rec.scaling_y = rec.scaling_x;
        uint64_t path_size = 0;
        if (!fh.read(&path_size, sizeof(path_size)) || path_size > (1ull << 20)) return RISLTC_E_FORMAT;
        if (path_size) {
            if (lights && path_storage && path_bytes + path_size <= path_capacity) {
                if (!fh.read(path_storage + path_bytes, path_size)) return RISLTC_E_FORMAT;
                path_storage[path_bytes + path_size - 1] = '\0';
                rec.texture_file_path = path_storage + path_bytes;
            } else if (!fh.skip(path_size)) {
                return RISLTC_E_FORMAT;
            } else if (lights) {
                return RISLTC_E_ARG;
            }
            path_bytes += path_size;
        }
        // Two pointers kept for backward compatibility, then 4 floats per
        // plane-space vertex (src/main.c:112-116).
        if (!fh.skip(2 * 8)) return RISLTC_E_FORMAT;
        const uint64_t n = 4ull * rec.vertex_count;
        if (lights) {
            if (vertex_floats + 2 * n > vertex_storage_floats) return RISLTC_E_ARG;
            float* plane = vertex_storage + vertex_floats;
            float* world = plane + n;
            if (!fh.read(plane, n * sizeof(float))) return RISLTC_E_FORMAT;
            memset(world, 0, n * sizeof(float));
            rec.vertices_plane_space = plane;
            rec.vertices_world_space = world;
            lights[i] = rec;
        } else if (!fh.skip(n * sizeof(float))) {
            return RISLTC_E_FORMAT;
        }
        vertex_floats += 2 * n;
        info->total_vertex_count += rec.vertex_count;
    }
    info->light_count = light_count;
    info->texture_path_bytes = path_bytes;
    if (camera) *camera = cam;
    return RISLTC_OK;
}

extern "C" int risltc_save_quicksave(const char* path, const risltc_camera_t* camera,
                                     const risltc_polygonal_light_t* lights, uint32_t light_count) {
    if (!path || !camera || (light_count && !lights)) return RISLTC_E_ARG;
    for (uint32_t i = 0; i < light_count; ++i)
        if (lights[i].vertex_count == 0 || lights[i].vertex_count > kQuicksaveMaxVertices ||
            !lights[i].vertices_plane_space)
            return RISLTC_E_ARG;
    File fh(path, "wb");
    if (!fh.f) {
        fprintf(stderr, "risltc_save_quicksave: cannot open %s for writing\n", path);
        return RISLTC_E_IO;
    }
    const uint32_t legacy_count = 0;
    bool ok = fh.write(camera, sizeof(*camera)) && fh.write(&legacy_count, sizeof(legacy_count)) &&
              fh.write(&light_count, sizeof(light_count));
    const uint64_t zero_pointers[2] = {0, 0};
    for (uint32_t i = 0; ok && i < light_count; ++i) {
        const risltc_polygonal_light_t& l = lights[i];
        uint64_t path_size = l.texture_file_path ? strlen(l.texture_file_path) + 1 : 0;
        ok = fh.write(&l, kQuicksaveLightBytes) && fh.write(&path_size, sizeof(path_size)) &&
             fh.write(l.texture_file_path, path_size) && fh.write(zero_pointers, sizeof(zero_pointers)) &&
             fh.write(l.vertices_plane_space, 4ull * l.vertex_count * sizeof(float));
    }
    return ok ? RISLTC_OK : RISLTC_E_IO;
}

namespace {

// Float -> shared-exponent RGBE after G. Ward, "Real Pixels" (Graphics Gems
// II, 1991): the largest component m sets the exponent e (m = f * 2^e with
// f in [0.5, 1)) and each component becomes floor(c * 256 f / m). The scale
// is formed in single precision. Components are clamped to [0, 2^127) first
// (NaN to 0) so that every exponent byte is in range.
inline void to_rgbe(const float rgb[3], uint8_t out[4]) {
    float c[3];
    for (int k = 0; k < 3; ++k) {
        float v = rgb[k];
        if (!(v > 0.0f)) v = 0.0f;
        if (v > 0x1.fffffep126f) v = 0x1.fffffep126f;
        c[k] = v;
    }
    float m = c[0] > c[1] ? c[0] : c[1];
    m = m > c[2] ? m : c[2];
    if (m < 1e-32f) {
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    int e = 0;
    float f = frexpf(m, &e);
    float scale = f * 256.0f / m;
    for (int k = 0; k < 3; ++k) out[k] = (uint8_t)(c[k] * scale);
    out[3] = (uint8_t)(e + 128);
}

// One component of one scanline in the run-length coding of Radiance's
// "new" scanline format: every maximal stretch of at least three equal bytes
// becomes run packets (count byte 128 + n, n <= 127, then the byte), the
// bytes between such stretches become literal packets (count byte n <= 128,
// then the n bytes).
void rle_component(const uint8_t* b, uint32_t n, std::vector<uint8_t>& out) {
    uint32_t lit = 0;  // start of the pending literal bytes
    auto flush_literal = [&](uint32_t end) {
        while (lit < end) {
            uint32_t len = end - lit < 128 ? end - lit : 128;
            out.push_back((uint8_t)len);
            out.insert(out.end(), b + lit, b + lit + len);
            lit += len;
        }
    };
    uint32_t i = 0;
    while (i < n) {
        uint32_t j = i + 1;
        while (j < n && b[j] == b[i]) ++j;
        if (j - i >= 3) {
            flush_literal(i);
            for (uint32_t k = i; k < j;) {
                uint32_t len = j - k < 127 ? j - k : 127;
                out.push_back((uint8_t)(128 + len));
                out.push_back(b[i]);
                k += len;
            }
            lit = j;
        }
        i = j;
    }
    flush_literal(n);
}

inline float round_to_half(float x) { return (float)(_Float16)x; }

}  // namespace

extern "C" int risltc_write_hdr(const char* path, const float* pixels, uint32_t width, uint32_t height,
                                uint32_t channels, int half_round) {
    if (!path || !pixels || width == 0 || height == 0 || (channels != 3 && channels != 4) ||
        width > (1u << 30) / channels)
        return RISLTC_E_ARG;
    File fh(path, "wb");
    if (!fh.f) {
        fprintf(stderr, "risltc_write_hdr: cannot write %s\n", path);
        return RISLTC_E_IO;
    }
    char header[160];
    int len = snprintf(header, sizeof(header), "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", height, width);
    if (!fh.write(header, (size_t)len)) return RISLTC_E_IO;
    // Scanlines of 8 to 32767 pixels are run-length coded component by
    // component behind the 4-byte marker (2, 2, width_hi, width_lo); other
    // widths are written as flat RGBE quadruples.
    const bool rle = width >= 8 && width < 32768;
    std::vector<uint8_t> quads(4ull * width), planes(rle ? 4ull * width : 0), line;
    line.reserve(8ull * width + 16);
    for (uint32_t y = 0; y < height; ++y) {
        const float* row = pixels + (size_t)y * width * channels;
        for (uint32_t x = 0; x < width; ++x) {
            float rgb[3];
            for (int k = 0; k < 3; ++k) {
                float v = row[(size_t)x * channels + k];
                rgb[k] = half_round ? round_to_half(v) : v;
            }
            to_rgbe(rgb, &quads[4ull * x]);
        }
        line.clear();
        if (!rle) {
            line.insert(line.end(), quads.begin(), quads.end());
        } else {
            line.push_back(2);
            line.push_back(2);
            line.push_back((uint8_t)(width >> 8));
            line.push_back((uint8_t)(width & 0xff));
            for (uint32_t k = 0; k < 4; ++k) {
                for (uint32_t x = 0; x < width; ++x) planes[(size_t)k * width + x] = quads[4ull * x + k];
                rle_component(&planes[(size_t)k * width], width, line);
            }
        }
        if (!fh.write(line.data(), line.size())) return RISLTC_E_IO;
    }
    return RISLTC_OK;
}
