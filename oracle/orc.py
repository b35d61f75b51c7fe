# SPDX-License-Identifier: GPL-3.0-or-later
# Part of the test oracle, which restates algorithms of the risltc renderer
# (Copyright (C) 2021 Christoph Peters, KIT; (C) 2023 Ishaan Shah, IIIT
# Hyderabad; GPLv3 or later) and is distributed under the same licence:
# see oracle/COPYING.
"""ctypes binding of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker. The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from risltc_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
# the same sources built -O3 for the CPU-baseline timing (identical results)
LIB_TIMING_PATH = os.path.join(HERE, "liborc_timing.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_host.so")
# Visibility-cache slots per cell when the cache keys by light cluster: one
# per cluster up to the largest cluster count the library accepts.
CACHE_CLUSTER_SLOTS = 1024

_lib = None
_lib_timing = None
_ref = None

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)
u16p = C.POINTER(C.c_uint16)
u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)


def build(reference=None):
    """make -C oracle (also builds oracle/_ref when the reference tree exists)."""
    cmd = ["make", "-s", "-C", HERE]
    if reference:
        cmd.append("REF=%s" % reference)
    subprocess.run(cmd, check=True)


class OrcBvh(C.Structure):
    _fields_ = [("leaf_count", C.c_uint32), ("node_box", f32p), ("node_left", i32p),
                ("node_right", i32p), ("node_escape", i32p), ("node_parent", i32p),
                ("leaf_tri", f32p), ("leaf_prim", u32p), ("sorted_keys", C.POINTER(C.c_uint64))]


class OrcScene(C.Structure):
    _fields_ = [("positions", u32p), ("normals_uv", u16p), ("material_indices", u8p),
                ("triangle_count", C.c_uint64), ("materials", C.POINTER(abi.Material)),
                ("material_count", C.c_uint32), ("light_records", u8p), ("light_count", C.c_uint32),
                ("light_max_vertex_count", C.c_uint32), ("ltc_rgba16", u16p), ("ltc_rg16", u16p),
                ("ltc_res_roughness", C.c_uint32), ("ltc_res_inclination", C.c_uint32),
                ("ltc_fresnel_count", C.c_uint32), ("bvh", C.POINTER(OrcBvh)),
                ("proposal_table", u8p), ("cluster_of_light", u32p), ("cluster_offsets", u32p),
                ("cluster_lights", u32p), ("cluster_count", C.c_uint32), ("cache_vis", u8p),
                ("cache_mask", C.c_uint32), ("cache_inv_cell", C.c_float), ("cache_slots", C.c_uint32),
                ("cache_per_light", C.c_uint32), ("proposal_table_bytes", C.c_uint32),
                ("proposal_group_log2", C.c_uint32)]


def lib(timing=False):
    """The oracle library; timing=True gives its -O3 build (the CPU
    baseline), which computes the same results."""
    global _lib, _lib_timing
    if timing:
        if _lib_timing is None:
            if not os.path.exists(LIB_TIMING_PATH):
                build()
            _lib_timing = _bind(C.CDLL(LIB_TIMING_PATH))
        return _lib_timing
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


def _bind(L):
    if True:
        sig = {
            "orc_murmur_mix": (C.c_uint32, [C.c_uint32, C.c_uint32]),
            "orc_murmur_finalize": (C.c_uint32, [C.c_uint32]),
            "orc_noise_seed": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
            "orc_noise_next": (C.c_float, [u32p]),
            "orc_wang_random_number": (C.c_uint32, [C.c_uint32]),
            "orc_omp_max_threads": (C.c_int, []),
            "prj_atanf": (C.c_float, [C.c_float]),
            "prj_asinf": (C.c_float, [C.c_float]),
            "prj_acosf": (C.c_float, [C.c_float]),
            "prj_sinf": (C.c_float, [C.c_float]),
            "prj_cosf": (C.c_float, [C.c_float]),
            "orc_update_polygonal_light": (None, [C.POINTER(abi.PolygonalLight)]),
            "orc_matrix_inverse": (None, [f32p, f32p]),
            "orc_get_world_to_projection_space": (None, [f32p, C.POINTER(abi.Camera), C.c_float]),
            "orc_write_frame_constants": (None, [C.POINTER(abi.FrameConstants), C.POINTER(abi.Camera),
                                                 C.c_uint32, C.c_uint32, f32p, f32p,
                                                 C.POINTER(abi.LtcConstants), C.c_float, C.c_float,
                                                 C.c_float, C.c_uint32, C.c_int]),
            "orc_quantize_ltc_entry": (None, [f32p, u16p, u16p]),
            "orc_ltc_constants": (None, [C.POINTER(abi.LtcConstants), C.c_uint32, C.c_uint32, C.c_uint32]),
            "orc_clip_rule": (C.c_int, [C.c_int, C.c_uint32, C.POINTER(C.c_int)]),
            "orc_clip_rotation": (C.c_int, [C.c_int, C.c_uint32]),
            "orc_bvh_build": (C.c_int, [C.POINTER(OrcBvh), f32p, C.c_uint32]),
            "orc_bvh_build_sah": (C.c_int, [C.POINTER(OrcBvh), f32p, C.c_uint32]),
            "orc_bvh_records": (None, [C.POINTER(OrcBvh), u32p, f32p]),
            "orc_bvh_free": (None, [C.POINTER(OrcBvh)]),
            "orc_cache_bytes": (None, [u32p, C.c_uint64, C.c_float, u8p]),
            "orc_morton_part_1_by_2": (C.c_uint32, [C.c_uint32]),
            "orc_bvh_any_hit": (C.c_int, [C.POINTER(OrcBvh), f32p, f32p, C.c_float, C.c_float, u32p]),
            "orc_bvh_closest_hit": (C.c_uint32, [C.POINTER(OrcBvh), f32p, f32p, C.c_float, C.c_float, f32p]),
            "orc_brute_any_hit": (C.c_int, [f32p, C.c_uint32, f32p, f32p, C.c_float, C.c_float]),
            "orc_ray_triangle": (C.c_int, [f32p, f32p, f32p, C.c_float, C.c_float, f32p]),
            "orc_dequantize_mesh": (None, [u32p, C.c_uint64, f32p, f32p, f32p]),
            "orc_render_visibility": (None, [C.POINTER(OrcBvh), u32p, C.POINTER(abi.FrameConstants), u32p, C.c_int]),
            "orc_shade_pixels_f32": (C.c_int, [C.POINTER(OrcScene), C.POINTER(abi.FrameConstants),
                                               C.POINTER(abi.Settings), u32p, u32p, C.c_uint32, f32p, f32p, u32p]),
            "orc_shade_pixels_f64": (C.c_int, [C.POINTER(OrcScene), C.POINTER(abi.FrameConstants),
                                               C.POINTER(abi.Settings), u32p, u32p, C.c_uint32, f64p, f32p, u32p]),
            "orc_shade_frame_f32": (C.c_int, [C.POINTER(OrcScene), C.POINTER(abi.FrameConstants),
                                              C.POINTER(abi.Settings), u32p, u32p, C.c_uint32, f32p, C.c_int]),
            "orc_clip_polygon_f32": (C.c_uint32, [C.c_uint32, f32p]),
            "orc_clip_polygon_f64": (C.c_uint32, [C.c_uint32, f64p]),
            "orc_calculate_ltc_f32": (C.c_float, [C.c_uint32, f32p]),
            "orc_calculate_ltc_f64": (C.c_double, [C.c_uint32, f64p]),
            "orc_prepare_psa_f32": (C.c_float, [C.c_uint32, f32p, f32p]),
            "orc_prepare_psa_f64": (C.c_double, [C.c_uint32, f64p, f64p]),
            "orc_sample_psa_f32": (None, [C.c_uint32, f32p, f32p, C.c_uint32, f32p]),
            "orc_sample_psa_f64": (None, [C.c_uint32, f64p, f64p, C.c_uint32, f64p]),
            "orc_evaluate_brdf_f32": (None, [f32p, f32p, f32p]),
            "orc_evaluate_brdf_f64": (None, [f64p, f64p, f64p]),
            "orc_ltc_coefficients_f32": (None, [C.POINTER(OrcScene), C.POINTER(abi.LtcConstants), C.c_float,
                                                C.c_float, f32p, f32p, f32p, f32p]),
            "orc_ltc_fetch_f32": (C.c_float, [C.POINTER(OrcScene), C.c_int, C.c_int, C.c_float, C.c_float,
                                              C.c_float]),
            "orc_ltc_fetch_f64": (C.c_double, [C.POINTER(OrcScene), C.c_int, C.c_int, C.c_double, C.c_double,
                                               C.c_double]),
            "orc_decode_position_f32": (None, [C.c_uint32, C.c_uint32, f32p, f32p, f32p]),
            "orc_decode_position_f64": (None, [C.c_uint32, C.c_uint32, f32p, f32p, f64p]),
            "orc_pixel_setup_f32": (C.c_int, [C.POINTER(OrcScene), C.POINTER(abi.FrameConstants), C.c_uint32,
                                              C.c_uint32, f32p]),
            "orc_ltc_cp_target_f32": (C.c_float, [C.POINTER(OrcScene), f32p, C.c_uint32, u32p]),
            "orc_ltc_cp_target_f64": (C.c_double, [C.POINTER(OrcScene), f32p, C.c_uint32, u32p]),
            "orc_decode_normal_f32": (None, [C.c_uint32, C.c_uint32, f32p]),
            "orc_decode_normal_f64": (None, [C.c_uint32, C.c_uint32, f64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
    return L


def ref_lib():
    """The reference's own host code compiled by oracle/Makefile, or None."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        R = C.CDLL(REF_LIB_PATH)
        R.update_polygonal_light.argtypes = [C.POINTER(abi.PolygonalLight)]
        R.update_polygonal_light.restype = None
        R.ref_matrix_inverse.argtypes = [f32p, f32p]
        R.ref_matrix_inverse.restype = None
        R.ref_wang_random_number.argtypes = [C.c_uint32]
        R.ref_wang_random_number.restype = C.c_uint32
        if hasattr(R, "ref_half_to_float"):
            R.ref_half_to_float.argtypes = [C.c_uint16]
            R.ref_half_to_float.restype = C.c_float
        _ref = R
    return _ref


def ref_stbiw_lib():
    """The reference's stb_image_write.h compiled by oracle/Makefile, or None."""
    path = os.path.join(os.path.dirname(REF_LIB_PATH), "libref_stbiw.so")
    if not os.path.exists(path):
        return None
    R = C.CDLL(path)
    R.stbi_write_hdr.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
    R.stbi_write_hdr.restype = C.c_int
    return R


def ptr(a, t):
    return a.ctypes.data_as(t)


def cache_bytes(counts, prior_strength):
    """orc_cache_bytes over a uint32 array of (unoccluded, occluded) pairs."""
    c = np.ascontiguousarray(counts, dtype=np.uint32).reshape(-1)
    out = np.zeros(c.size // 2, dtype=np.uint8)
    lib().orc_cache_bytes(ptr(c, u32p), out.size, float(np.float32(prior_strength)), ptr(out, u8p))
    return out


class Oracle:
    """The oracle bound to one synthetic scene (risltc_amd.synth.Scene)."""

    def __init__(self, scene, ltc_rgba16, ltc_rg16, bvh=True, builder=abi.BVH_SAH, timing=False):
        self.L = lib(timing)
        self.scene = scene
        m = scene.mesh
        self.positions = np.ascontiguousarray(m.positions, dtype=np.uint32)
        self.normals_uv = np.ascontiguousarray(m.normals_uv, dtype=np.uint16)
        self.material_index = np.ascontiguousarray(m.material_index, dtype=np.uint8)
        self.materials = scene.materials_ctypes()
        self.records = np.frombuffer(scene.light_records, dtype=np.uint8).copy()
        self.rgba = np.ascontiguousarray(ltc_rgba16, dtype=np.uint16)
        self.rg = np.ascontiguousarray(ltc_rg16, dtype=np.uint16)
        res = self.rgba.shape[1]
        layers = self.rgba.shape[0]
        self.ltc_constants = abi.LtcConstants()
        self.L.orc_ltc_constants(C.byref(self.ltc_constants), res, res, layers)
        s = OrcScene()
        s.positions = ptr(self.positions, u32p)
        s.normals_uv = ptr(self.normals_uv, u16p)
        s.material_indices = ptr(self.material_index, u8p)
        s.triangle_count = m.triangle_count
        s.materials = C.cast(self.materials, C.POINTER(abi.Material))
        s.material_count = len(self.materials)
        s.light_records = ptr(self.records, u8p)
        s.light_count = scene.light_count
        s.light_max_vertex_count = scene.light_vmax
        s.ltc_rgba16 = ptr(self.rgba, u16p)
        s.ltc_rg16 = ptr(self.rg, u16p)
        s.ltc_res_roughness = res
        s.ltc_res_inclination = res
        s.ltc_fresnel_count = layers
        self.bvh = None
        if bvh:
            self.triangles = self.dequantized_triangles()
            self.bvh = OrcBvh()
            build = self.L.orc_bvh_build_sah if builder == abi.BVH_SAH else self.L.orc_bvh_build
            build(C.byref(self.bvh), ptr(self.triangles, f32p), m.triangle_count)
            s.bvh = C.pointer(self.bvh)
        self.cscene = s

    def __del__(self):
        try:
            if self.bvh is not None:
                self.L.orc_bvh_free(C.byref(self.bvh))
        except Exception:
            pass

    def set_proposal(self, table, cluster_of_light, cluster_count, group_log2=0):
        """Learned-source inputs: per pixel tables (pixels x 32 uint8, as
        risltc_proposal_update writes them), the light clusters (lights are
        ordered by cluster, then by index: the library's rule) and log2 of
        the clusters a table slot names (abi.proposal_group_log2 of the
        network's outputs)."""
        t = np.asarray(table, dtype=np.uint8)
        width = t.shape[-1] if t.ndim == 2 else 32
        assert width == 32
        self.cscene.proposal_table_bytes = width
        self.cscene.proposal_group_log2 = group_log2
        self.prop_table = np.ascontiguousarray(t).reshape(-1)
        col = np.ascontiguousarray(cluster_of_light, dtype=np.uint32)
        self.prop_of = col
        counts = np.bincount(col, minlength=cluster_count).astype(np.uint32)
        self.prop_off = np.zeros(cluster_count + 1, dtype=np.uint32)
        self.prop_off[1:] = np.cumsum(counts)
        self.prop_lights = np.argsort(col, kind="stable").astype(np.uint32)
        s = self.cscene
        s.proposal_table = ptr(self.prop_table, u8p)
        s.cluster_of_light = ptr(self.prop_of, u32p)
        s.cluster_offsets = ptr(self.prop_off, u32p)
        s.cluster_lights = ptr(self.prop_lights, u32p)
        s.cluster_count = cluster_count

    def set_cache(self, counts, cell_size, prior_strength, per_light=False):
        """Visibility cache from its counts [2^k, slots, 2] uint32 (unoccluded,
        occluded), as the library's risltc_proposal_cache_read returns them
        (slots = CACHE_CLUSTER_SLOTS = 1024 clusters, or the light count with
        per_light): the bytes
        are made by orc_cache_bytes; None removes the cache. Returns the
        bytes [2^k, slots]."""
        s = self.cscene
        if counts is None:
            self.cache = None
            s.cache_vis = None
            return None
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        slots = self.scene.light_count if per_light else CACHE_CLUSTER_SLOTS
        cells = counts.size // (2 * slots)
        assert cells * 2 * slots == counts.size and cells & (cells - 1) == 0
        self.cache = cache_bytes(counts, prior_strength)
        s.cache_vis = ptr(self.cache, u8p)
        s.cache_mask = cells - 1
        s.cache_inv_cell = float(np.float32(1.0) / np.float32(cell_size))
        s.cache_slots = slots
        s.cache_per_light = 1 if per_light else 0
        return self.cache.reshape(cells, slots)

    def bvh_records(self):
        """Pre-order records of the oracle's tree, as the device lays it out."""
        n = self.bvh.leaf_count
        out = np.zeros((5 * n - 2, 4), dtype=np.float32)
        self.L.orc_bvh_records(C.byref(self.bvh), ptr(self.positions, u32p), ptr(out, f32p))
        return out

    def dequantized_triangles(self):
        m = self.scene.mesh
        out = np.empty((m.triangle_count, 9), dtype=np.float32)
        f = np.ascontiguousarray(m.dequantization_factor, dtype=np.float32)
        s = np.ascontiguousarray(m.dequantization_summand, dtype=np.float32)
        self.L.orc_dequantize_mesh(ptr(self.positions, u32p), m.triangle_count, ptr(f, f32p), ptr(s, f32p),
                                   ptr(out, f32p))
        return out

    def frame_constants(self, exposure=1.5, roughness_factor=1.0, mis_visibility_estimate=0.5,
                        noise_seed=0, animate=False, camera=None):
        sc = self.scene
        c = abi.FrameConstants()
        f = np.ascontiguousarray(sc.mesh.dequantization_factor, dtype=np.float32)
        s = np.ascontiguousarray(sc.mesh.dequantization_summand, dtype=np.float32)
        self.L.orc_write_frame_constants(C.byref(c), C.byref(camera or sc.camera), sc.width, sc.height,
                                         ptr(f, f32p), ptr(s, f32p), C.byref(self.ltc_constants), exposure,
                                         roughness_factor, mis_visibility_estimate, noise_seed,
                                         1 if animate else 0)
        return c

    def render_visibility(self, constants, threads=8):
        out = np.empty(int(constants.viewport_size[0]) * int(constants.viewport_size[1]), dtype=np.uint32)
        self.L.orc_render_visibility(C.byref(self.bvh), ptr(self.positions, u32p), C.byref(constants),
                                     ptr(out, u32p), threads)
        return out

    def shade(self, constants, settings, visibility, pixels=None, fp64=False, records=False,
              candidates=False):
        vis = np.ascontiguousarray(visibility, dtype=np.uint32)
        if pixels is None:
            count = vis.size
            pix = None
        else:
            pix = np.ascontiguousarray(pixels, dtype=np.uint32)
            count = pix.size
        out = np.zeros((count, 4), dtype=np.float64 if fp64 else np.float32)
        spp = settings.sample_count_light
        M = settings.ris_candidate_count
        rec = np.zeros((count, spp, abi.SAMPLE_RECORD_FLOATS), dtype=np.float32) if records else None
        cand = (np.zeros((count, spp, M, abi.CANDIDATE_RECORD_WORDS), dtype=np.uint32)
                if candidates else None)
        fn = self.L.orc_shade_pixels_f64 if fp64 else self.L.orc_shade_pixels_f32
        rc = fn(C.byref(self.cscene), C.byref(constants), C.byref(settings), ptr(vis, u32p),
                ptr(pix, u32p) if pix is not None else None, count,
                ptr(out, f64p if fp64 else f32p), ptr(rec, f32p) if rec is not None else None,
                ptr(cand, u32p) if cand is not None else None)
        if rc != 0:
            raise RuntimeError("oracle shade failed: %s" % abi.ERROR_NAMES.get(rc, rc))
        return out, rec, cand

    def ltc_cp_targets(self, constants, visibility, pixels, lights):
        """RIS targets of the listed (pixel, light) pairs evaluated in FP32
        and in FP64 from the pixels' FP32 setup, lights shaped (pixels, ...):
        (t32, t64, clip counts FP32, clip counts FP64)."""
        pix = np.asarray(pixels, np.int64)
        li = np.asarray(lights, np.int64)
        vis = np.asarray(visibility, np.uint32)
        t32 = np.zeros((pix.size, li.size // max(pix.size, 1)), np.float32)
        t64 = np.zeros(t32.shape, np.float64)
        cc32 = np.zeros(t32.shape, np.uint32)
        cc64 = np.zeros(t32.shape, np.uint32)
        setup = np.zeros(55, np.float32)
        one = np.zeros(1, np.uint32)
        flat_l = li.reshape(t32.shape)
        sp = ptr(setup, f32p)
        for i, p in enumerate(pix):
            rc = self.L.orc_pixel_setup_f32(C.byref(self.cscene), C.byref(constants), int(p), int(vis[p]), sp)
            if rc != 0:
                raise RuntimeError("pixel %d has no shading point" % p)
            for k, l in enumerate(flat_l[i]):
                t32[i, k] = self.L.orc_ltc_cp_target_f32(C.byref(self.cscene), sp, int(l), ptr(one, u32p))
                cc32[i, k] = one[0]
                t64[i, k] = self.L.orc_ltc_cp_target_f64(C.byref(self.cscene), sp, int(l), ptr(one, u32p))
                cc64[i, k] = one[0]
        return t32.reshape(li.shape), t64.reshape(li.shape), cc32.reshape(li.shape), cc64.reshape(li.shape)

    def shade_frame(self, constants, settings, visibility, pixels=None, threads=8):
        vis = np.ascontiguousarray(visibility, dtype=np.uint32)
        pix = None if pixels is None else np.ascontiguousarray(pixels, dtype=np.uint32)
        count = vis.size if pix is None else pix.size
        out = np.zeros((count, 4), dtype=np.float32)
        rc = self.L.orc_shade_frame_f32(C.byref(self.cscene), C.byref(constants), C.byref(settings),
                                        ptr(vis, u32p), ptr(pix, u32p) if pix is not None else None, count,
                                        ptr(out, f32p), threads)
        if rc != 0:
            raise RuntimeError("oracle shade failed: %s" % abi.ERROR_NAMES.get(rc, rc))
        return out
