"""ctypes mirror of include/risltc/risltc.h.

Every structure here has the byte layout of its C counterpart, which in turn
has the layout of the reference's host structure (cited in the header).
"""
import ctypes as C

MAX_LIGHT_VERTEX_COUNT = 8
MAX_POLYGON_VERTEX_COUNT = 9
SAMPLE_RECORD_FLOATS = 40
CANDIDATE_RECORD_WORDS = 4
# sample records (pixels x spp x sample_count) of risltc_shade_noalloc's fixed
# scratch (RISLTC_NOALLOC_MAX_SAMPLES)
NOALLOC_MAX_SAMPLES = 1 << 19
# SAMPLE_COUNT range (RISLTC_MAX_SAMPLE_COUNT)
MAX_SAMPLE_COUNT = 1024

OK, E_ARG, E_HIP, E_FORMAT, E_UNSUPPORTED, E_STATE, E_IO = range(7)
ERROR_NAMES = {0: "OK", 1: "E_ARG", 2: "E_HIP", 3: "E_FORMAT", 4: "E_UNSUPPORTED",
               5: "E_STATE", 6: "E_IO"}

MIS_BALANCE, MIS_POWER, MIS_WEIGHTED, MIS_OPTIMAL_CLAMPED, MIS_OPTIMAL = range(5)
LIGHT_UNIFORM, LIGHT_RESERVOIR = 0, 1
SAMPLE_POLYGON_AREA_TURK, SAMPLE_POLYGON_PROJECTED_SOLID_ANGLE, SAMPLE_POLYGON_LTC_CP = 1, 2, 4
PROPOSAL_UNIFORM, PROPOSAL_LEARNED, PROPOSAL_CACHE = 0, 1, 2
BVH_SAH, BVH_LBVH = 0, 1
MAX_LIGHT_CLUSTERS = 1024
PROPOSAL_PARAMS = 31168
PROPOSAL_PARAMS_WIDE = 78464


def proposal_table_bytes(outputs):
    """Bytes per pixel of a proposal table (RISLTC_PROPOSAL_TABLE_BYTES):
    16 one-byte group ids and 16 weights, for every network."""
    return 32


def proposal_group_log2(outputs):
    """log2 of the clusters one table slot names: a cluster for networks of
    at most 256 outputs, a quad of clusters for more."""
    return 2 if outputs > 256 else 0

# risltc_sample_record_t word offsets
REC_LIGHT, REC_WSUM, REC_PHAT, REC_W, REC_CLIPD, REC_CLIPS, REC_PSAD, REC_PSAS = range(8)
REC_DIR, REC_TMAX, REC_STATE, REC_EVIS, REC_EOCC, REC_COLOR, REC_SEED, REC_POS = 8, 14, 16, 18, 24, 30, 33, 34
REC_CACHE_KEY = 37


def light_record_size(max_vertex_count):
    """RISLTC_LIGHT_RECORD_SIZE: 48 B header + 16 B per vertex (src/main.c:456-490)."""
    return 48 + 16 * max_vertex_count


class LtcConstants(C.Structure):
    """ltc_constants_t (src/ltc_table.h:23-35)."""
    _fields_ = [("fresnel_index_factor", C.c_float), ("fresnel_index_summand", C.c_float),
                ("roughness_factor", C.c_float), ("roughness_summand", C.c_float),
                ("inclination_factor", C.c_float), ("inclination_summand", C.c_float),
                ("padding", C.c_float * 2)]


class FrameConstants(C.Structure):
    """per_frame_constants_t (src/main.h:537-553), 256 bytes."""
    _fields_ = [("mesh_dequantization_factor", C.c_float * 3), ("padding_0", C.c_float),
                ("mesh_dequantization_summand", C.c_float * 3), ("error_factor", C.c_float),
                ("world_to_projection_space", (C.c_float * 4) * 4),
                ("pixel_to_ray_direction_world_space", (C.c_float * 4) * 3),
                ("camera_position_world_space", C.c_float * 3),
                ("mis_visibility_estimate", C.c_float),
                ("viewport_size", C.c_uint32 * 2), ("cursor_position", C.c_int32 * 2),
                ("exposure_factor", C.c_float), ("roughness_factor", C.c_float),
                ("noise_resolution_mask", C.c_uint32 * 2), ("noise_texture_index_mask", C.c_uint32),
                ("padding_3", C.c_uint32 * 3), ("noise_random_numbers", C.c_uint32 * 4),
                ("ltc_constants", LtcConstants)]


class Camera(C.Structure):
    """first_person_camera_t (src/camera.h:29-49)."""
    _fields_ = [("position_world_space", C.c_float * 3), ("rotation_z", C.c_float),
                ("rotation_x", C.c_float), ("vertical_fov", C.c_float),
                ("near_plane", C.c_float), ("far_plane", C.c_float), ("speed", C.c_float),
                ("rotate_camera", C.c_int32), ("rotation_x_0", C.c_float),
                ("rotation_z_0", C.c_float)]


class Material(C.Structure):
    _fields_ = [("base_color", C.c_float * 3), ("specular", C.c_float * 3)]


class Settings(C.Structure):
    """risltc_settings_t (mirrors render_settings_t, src/main.h:123-156)."""
    _fields_ = [("sample_count", C.c_uint32), ("sample_count_light", C.c_uint32),
                ("mis_heuristic", C.c_uint32), ("light_sampling", C.c_uint32),
                ("polygon_sampling_technique", C.c_uint32),
                ("ris_candidate_count", C.c_uint32), ("fast_atan", C.c_uint32),
                ("shadow_rays", C.c_uint32), ("proposal", C.c_uint32),
                ("exact_math", C.c_uint32), ("learned_floor_log2", C.c_uint32),
                ("skip_training_records", C.c_uint32), ("reserved", C.c_uint32 * 4)]


def default_settings(**kw):
    """The reference defaults (src/main.c:220-236) with M = 32."""
    s = Settings()
    s.sample_count = 1
    s.sample_count_light = 1
    s.mis_heuristic = MIS_OPTIMAL_CLAMPED
    s.light_sampling = LIGHT_RESERVOIR
    s.polygon_sampling_technique = SAMPLE_POLYGON_LTC_CP
    s.ris_candidate_count = 32
    s.fast_atan = 0
    s.shadow_rays = 1
    s.proposal = PROPOSAL_UNIFORM
    s.exact_math = 0
    for k, v in kw.items():
        setattr(s, k, v)
    return s


class Tuning(C.Structure):
    """risltc_tuning_t: launch shape of the kernels (changes no result bit)."""
    _fields_ = [("wave_width", C.c_uint32), ("cull_chunk", C.c_uint32), ("lds_lights", C.c_uint32),
                ("trace_refill", C.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class View(C.Structure):
    """risltc_view_t: one view of a batch."""
    _fields_ = [("constants", FrameConstants), ("settings", Settings), ("visibility", C.c_void_p),
                ("out", C.c_void_p), ("light_set", C.c_uint32), ("scene", C.c_uint32),
                ("proposal_table", C.c_void_p)]


class PolygonalLight(C.Structure):
    """polygonal_light_t (src/polygonal_light.h:73-99)."""
    _fields_ = [("rotation_angles", C.c_float * 3), ("scaling_x", C.c_float),
                ("translation", C.c_float * 3), ("scaling_y", C.c_float),
                ("radiant_flux", C.c_float * 3), ("inv_scaling_x", C.c_float),
                ("surface_radiance", C.c_float * 3), ("inv_scaling_y", C.c_float),
                ("plane", C.c_float * 4), ("vertex_count", C.c_uint32),
                ("texturing_technique", C.c_uint32), ("texture_index", C.c_uint32),
                ("padding_0", C.c_uint32), ("rotation", (C.c_float * 4) * 3),
                ("area", C.c_float), ("rcp_area", C.c_float), ("padding_1", C.c_float * 2),
                ("texture_file_path", C.c_char_p),
                ("vertices_plane_space", C.POINTER(C.c_float)),
                ("vertices_world_space", C.POINTER(C.c_float))]


assert C.sizeof(FrameConstants) == 256
assert C.sizeof(LtcConstants) == 32
assert C.sizeof(Settings) == 64
assert C.sizeof(View) == 256 + 64 + 16 + 16
