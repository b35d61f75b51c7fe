"""MI355X (gfx950) RIS-LTC shading pass.

Python host mirror of the C-ABI in include/risltc/risltc.h (librisltc.so,
built in-tree by risltc_amd/build.py). The mirror keeps the reference's host
vocabulary: light records as write_lights packs them, LTC tables as
load_ltc_table builds them, per-frame constants, and the visibility buffer.
Device buffers are torch tensors (used only for device memory and streams).

There is no CPU fallback: if librisltc.so is missing or fails to load, every
call raises.
"""
import ctypes as C
import os
import weakref

import numpy as np

from . import abi

__all__ = ["Context", "Graph", "lib", "LIB_PATH", "RisltcError", "abi"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "librisltc.so")
_lib = None


class RisltcError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("%s: %s" % (abi.ERROR_NAMES.get(code, code), message))
        self.code = code


_vp = C.c_void_p
_u32 = C.c_uint32
_f32p = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_u16p = C.POINTER(C.c_uint16)
_u8p = C.POINTER(C.c_uint8)

SIGNATURES = {
    "risltc_default_settings": (None, [C.POINTER(abi.Settings)]),
    "risltc_update_polygonal_light": (None, [C.POINTER(abi.PolygonalLight)]),
    "risltc_write_lights": (C.c_size_t, [C.POINTER(abi.PolygonalLight), _u32, _u32, _vp, C.c_size_t]),
    "risltc_write_frame_constants": (None, [C.POINTER(abi.FrameConstants), C.POINTER(abi.Camera), _u32, _u32,
                                            _f32p, _f32p, C.POINTER(abi.LtcConstants), C.c_float, C.c_float,
                                            C.c_float, _u32, C.c_int]),
    "risltc_ltc_constants_for": (None, [C.POINTER(abi.LtcConstants), _u32, _u32, _u32]),
    "risltc_load_ltc_fit": (C.c_int, [C.c_char_p, _u32, _u32p, _u16p, _u16p, C.c_size_t]),
    "risltc_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "risltc_destroy": (None, [_vp]),
    "risltc_last_error": (C.c_char_p, [_vp]),
    "risltc_set_lights": (C.c_int, [_vp, _vp, _u32, _u32]),
    "risltc_set_ltc_table": (C.c_int, [_vp, _u16p, _u16p, _u32, _u32, _u32]),
    "risltc_set_scene": (C.c_int, [_vp, _u32p, _u16p, _u8p, C.c_uint64, _f32p, _f32p,
                                   C.POINTER(abi.Material), _u32]),
    "risltc_set_bvh_builder": (C.c_int, [_vp, _u32]),
    "risltc_set_light_clusters": (C.c_int, [_vp, _u32p, _u32]),
    "risltc_cluster_lights": (C.c_int, [_vp, _u32, _u32, _u32p]),
    "risltc_proposal_init": (C.c_int, [_vp, _u32]),
    "risltc_proposal_set_params": (C.c_int, [_vp, _f32p]),
    "risltc_proposal_get_params": (C.c_int, [_vp, _f32p, _f32p]),
    "risltc_proposal_update": (C.c_int, [_vp, C.POINTER(abi.FrameConstants), _vp, _vp, _vp, _vp, _vp]),
    "risltc_proposal_train": (C.c_int, [_vp, C.POINTER(abi.FrameConstants), _vp, _u32, _u32, C.c_float, _u32,
                                        C.POINTER(C.c_float), _vp]),
    "risltc_proposal_last_timings": (C.c_int, [_vp, C.POINTER(C.c_float * 2)]),
    "risltc_proposal_records": (C.c_int, [_vp, _vp, _vp, C.POINTER(C.c_uint64), _vp]),
    "risltc_bvh_records": (C.c_int, [_vp, _f32p, C.c_uint64, C.POINTER(C.c_uint32)]),
    "risltc_render_visibility": (C.c_int, [_vp, C.POINTER(abi.FrameConstants), _vp, _vp]),
    "risltc_shade": (C.c_int, [_vp, C.POINTER(abi.FrameConstants), C.POINTER(abi.Settings), _vp, _vp, _vp]),
    "risltc_shade_debug": (C.c_int, [_vp, C.POINTER(abi.FrameConstants), C.POINTER(abi.Settings), _vp, _vp,
                                     _vp, _vp, _vp]),
    "risltc_shade_noalloc": (C.c_int, [_vp, C.POINTER(abi.FrameConstants), C.POINTER(abi.Settings), _vp, _vp,
                                       _vp]),
    "risltc_accumulate": (C.c_int, [_vp, _vp, _vp, _u32, _u32, _vp]),
    "risltc_trace_shadow_rays": (C.c_int, [_vp, _vp, _u32, _vp, _vp]),
    "risltc_last_timings": (C.c_int, [_vp, _f32p]),
    "risltc_last_ray_count": (C.c_int, [_vp, C.POINTER(C.c_uint32)]),
    "risltc_proposal_cache": (C.c_int, [_vp, _u32, C.c_float, C.c_float, _u32]),
    "risltc_proposal_cache_read": (C.c_int, [_vp, _u32p, C.c_uint64]),
    "risltc_proposal_cache_write": (C.c_int, [_vp, _u32p, C.c_uint64]),
    "risltc_proposal_cache_table": (C.c_int, [_vp, _u8p, C.c_uint64]),
    "risltc_add_light_set": (C.c_int, [_vp, _vp, _u32, _u32, _u32p]),
    "risltc_set_polygonal_lights": (C.c_int, [_vp, C.POINTER(abi.PolygonalLight), _u32, _u32]),
    "risltc_load_ltc_table": (C.c_int, [_vp, C.c_char_p, _u32]),
    "risltc_proposal_set_width": (C.c_int, [_vp, _u32]),
    "risltc_proposal_set_network": (C.c_int, [_vp, _u32, _u32]),
    "risltc_proposal_param_count": (_u32, [_vp]),
    "risltc_add_scene": (C.c_int, [_vp, _u32p, _u16p, _u8p, C.c_uint64, _f32p, _f32p, C.POINTER(abi.Material), _u32,
                                   _u32p]),
    "risltc_render_visibility_scene": (C.c_int, [_vp, _u32, C.POINTER(abi.FrameConstants), _vp, _vp]),
    "risltc_shade_batch": (C.c_int, [_vp, C.POINTER(abi.View), _u32, _vp]),
    "risltc_get_tuning": (C.c_int, [_vp, C.POINTER(abi.Tuning)]),
    "risltc_set_tuning": (C.c_int, [_vp, C.POINTER(abi.Tuning)]),
    "risltc_autotune": (C.c_int, [_vp, C.POINTER(abi.View), _u32, _u32, C.POINTER(abi.Tuning)]),
    "risltc_graph_create": (C.c_int, [_vp, C.POINTER(abi.View), _u32, C.POINTER(_vp)]),
    "risltc_graph_update": (C.c_int, [_vp, C.POINTER(abi.View), _u32]),
    "risltc_graph_launch": (C.c_int, [_vp, _vp]),
    "risltc_graph_destroy": (None, [_vp]),
}


def lib():
    """Loads librisltc.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("librisltc.so is not built (run risltc_amd/build.py or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _dptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        try:
            import torch
            return C.c_void_p(torch.cuda.current_stream().cuda_stream)
        except Exception:
            return None
    return C.c_void_p(stream)


def default_settings(**kw):
    return abi.default_settings(**kw)


def write_frame_constants(camera, width, height, deq_factor, deq_summand, ltc_constants, exposure=1.5,
                          roughness_factor=1.0, mis_visibility_estimate=0.5, noise_seed=0, animate=False):
    """write_constants (src/main.c:2902-2946) through the library."""
    c = abi.FrameConstants()
    f = np.ascontiguousarray(deq_factor, dtype=np.float32)
    s = np.ascontiguousarray(deq_summand, dtype=np.float32)
    lib().risltc_write_frame_constants(C.byref(c), C.byref(camera), width, height, _ptr(f, _f32p),
                                       _ptr(s, _f32p), C.byref(ltc_constants), exposure, roughness_factor,
                                       mis_visibility_estimate, noise_seed, 1 if animate else 0)
    return c


def ltc_constants_for(res_roughness, res_inclination, fresnel_count):
    k = abi.LtcConstants()
    lib().risltc_ltc_constants_for(C.byref(k), res_roughness, res_inclination, fresnel_count)
    return k


class Graph:
    """risltc_graph_t: a batch of views captured as one hipGraph with buffers
    of its own. update() takes views of the same shapes; launch() replays."""

    def __init__(self, ctx, views):
        self.ctx = ctx
        arr, n = Context._views(views)
        h = _vp()
        ctx._check(ctx.L.risltc_graph_create(ctx.h, arr, n, C.byref(h)), "risltc_graph_create")
        self.h = h
        ctx._graphs.add(self)

    def update(self, views):
        arr, n = Context._views(views)
        self.ctx._check(self.ctx.L.risltc_graph_update(self.h, arr, n), "risltc_graph_update")

    def launch(self, stream=None):
        self.ctx._check(self.ctx.L.risltc_graph_launch(self.h, _stream(stream)), "risltc_graph_launch")

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.risltc_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """A shading context on one HIP device (risltc_context_t)."""

    def __init__(self, device=0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.risltc_create(C.byref(h), device)
        if rc != 0:
            raise RisltcError(rc, "risltc_create(%d) failed" % device)
        self.h = h
        self.device = device
        self.ltc_constants = None
        self._keep = []
        self._graphs = weakref.WeakSet()
        self._mlp = (64, 256)  # hidden width, outputs

    def close(self):
        if getattr(self, "h", None):
            # graphs hold the context's device pointers: destroy them first
            for g in list(self._graphs):
                g.close()
            self.L.risltc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise RisltcError(rc, "%s: %s" % (what, self.L.risltc_last_error(self.h).decode()))

    # ---- uploads ---------------------------------------------------------
    def set_lights(self, records, count, max_vertex_count):
        buf = np.frombuffer(records, dtype=np.uint8)
        self._check(self.L.risltc_set_lights(self.h, buf.ctypes.data_as(_vp), count, max_vertex_count),
                    "risltc_set_lights")
        self._light_count = count

    def set_polygonal_lights(self, lights, max_vertex_count=0):
        """risltc_set_polygonal_lights over a ctypes array of
        abi.PolygonalLight (update_polygonal_light + write_lights + upload)."""
        n = len(lights)
        self._check(self.L.risltc_set_polygonal_lights(self.h, lights, n, max_vertex_count),
                    "risltc_set_polygonal_lights")
        self._light_count = n

    def load_ltc_table(self, directory, fresnel_count=51):
        """risltc_load_ltc_table: fit<i>.dat files with load_ltc_table's semantics."""
        res = C.c_uint32()
        rc = self.L.risltc_load_ltc_fit(str(directory).encode(), fresnel_count, C.byref(res), None, None, 0)
        if rc != 0:
            raise RisltcError(rc, "risltc_load_ltc_fit(%s)" % directory)
        self._check(self.L.risltc_load_ltc_table(self.h, str(directory).encode(), fresnel_count),
                    "risltc_load_ltc_table")
        self.ltc_constants = ltc_constants_for(res.value, res.value, fresnel_count)

    def set_ltc_table(self, rgba16, rg16):
        rgba = np.ascontiguousarray(rgba16, dtype=np.uint16)
        rg = np.ascontiguousarray(rg16, dtype=np.uint16)
        layers, h, w = rgba.shape[0], rgba.shape[1], rgba.shape[2]
        self._check(self.L.risltc_set_ltc_table(self.h, _ptr(rgba, _u16p), _ptr(rg, _u16p), w, h, layers),
                    "risltc_set_ltc_table")
        self.ltc_constants = ltc_constants_for(w, h, layers)

    def set_scene(self, mesh, materials):
        pos = np.ascontiguousarray(mesh.positions, dtype=np.uint32)
        nuv = np.ascontiguousarray(mesh.normals_uv, dtype=np.uint16)
        mi = np.ascontiguousarray(mesh.material_index, dtype=np.uint8)
        f = np.ascontiguousarray(mesh.dequantization_factor, dtype=np.float32)
        s = np.ascontiguousarray(mesh.dequantization_summand, dtype=np.float32)
        mats = np.ascontiguousarray(materials, dtype=np.float32).reshape(-1, 6)
        arr = (abi.Material * len(mats))()
        for i, m in enumerate(mats):
            arr[i].base_color[:] = [float(v) for v in m[0:3]]
            arr[i].specular[:] = [float(v) for v in m[3:6]]
        self._check(self.L.risltc_set_scene(self.h, _ptr(pos, _u32p), _ptr(nuv, _u16p), _ptr(mi, _u8p),
                                            mesh.triangle_count, _ptr(f, _f32p), _ptr(s, _f32p), arr, len(mats)),
                    "risltc_set_scene")

    def set_bvh_builder(self, builder):
        """abi.BVH_SAH (default) or abi.BVH_LBVH, for the next set_scene."""
        self._check(self.L.risltc_set_bvh_builder(self.h, builder), "risltc_set_bvh_builder")

    def bvh_records(self):
        """The hierarchy's pre-order records as a (units, 4) float32 array."""
        units = C.c_uint32(0)
        self._check(self.L.risltc_bvh_records(self.h, None, 0, C.byref(units)), "risltc_bvh_records")
        out = np.zeros((units.value, 4), dtype=np.float32)
        self._check(self.L.risltc_bvh_records(self.h, _ptr(out, _f32p), units.value, C.byref(units)),
                    "risltc_bvh_records")
        return out

    # ---- learned proposal ----------------------------------------------------
    def set_light_clusters(self, cluster_of_light, cluster_count):
        a = np.ascontiguousarray(cluster_of_light, dtype=np.uint32)
        self._check(self.L.risltc_set_light_clusters(self.h, _ptr(a, _u32p), cluster_count),
                    "risltc_set_light_clusters")

    def cluster_lights(self, cluster_count=256, iterations=20):
        """k-means of the lights; returns the assignment (numpy uint32)."""
        n = self.scene.light_count if getattr(self, "scene", None) is not None else None
        out = np.zeros(n if n is not None else 1 << 20, dtype=np.uint32)
        self._check(self.L.risltc_cluster_lights(self.h, cluster_count, iterations, _ptr(out, _u32p)),
                    "risltc_cluster_lights")
        return out[:n] if n is not None else out

    def proposal_init(self, seed=0, width=None, outputs=None):
        """Initialises the MLP; width 64 (default) or 128 and outputs 256
        (default) or 1024 select the network first."""
        if width is not None or outputs is not None:
            self.proposal_set_network(width or self._mlp[0], outputs or self._mlp[1])
        self._check(self.L.risltc_proposal_init(self.h, seed), "risltc_proposal_init")

    def proposal_set_width(self, width):
        self.proposal_set_network(width, self._mlp[1])

    def proposal_set_network(self, width, outputs):
        self._check(self.L.risltc_proposal_set_network(self.h, width, outputs), "risltc_proposal_set_network")
        self._mlp = (width, outputs)

    def proposal_param_count(self):
        return int(self.L.risltc_proposal_param_count(self.h))

    def proposal_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.proposal_param_count()
        self._check(self.L.risltc_proposal_set_params(self.h, _ptr(p, _f32p)), "risltc_proposal_set_params")

    def proposal_get_params(self):
        """(params, gradient of the last training step) as float32 arrays."""
        p = np.zeros(self.proposal_param_count(), dtype=np.float32)
        g = np.zeros(self.proposal_param_count(), dtype=np.float32)
        self._check(self.L.risltc_proposal_get_params(self.h, _ptr(p, _f32p), _ptr(g, _f32p)),
                    "risltc_proposal_get_params")
        return p, g

    def proposal_update(self, constants, visibility, table=None, logits=None, features=None, stream=None):
        self._check(self.L.risltc_proposal_update(self.h, C.byref(constants), _dptr(visibility), _dptr(table),
                                                  _dptr(logits), _dptr(features), _stream(stream)),
                    "risltc_proposal_update")

    def proposal_train(self, constants, visibility, batch=65536, steps=1, lr=1e-3, seed=0, stream=None):
        loss = C.c_float(0.0)
        self._check(self.L.risltc_proposal_train(self.h, C.byref(constants), _dptr(visibility), batch, steps,
                                                 lr, seed, C.byref(loss), _stream(stream)),
                    "risltc_proposal_train")
        return loss.value

    def proposal_records(self):
        """(cluster, y) training records of the last shade, as torch tensors."""
        import torch
        n = C.c_uint64(0)
        self._check(self.L.risltc_proposal_records(self.h, None, None, C.byref(n), None), "risltc_proposal_records")
        cl = torch.zeros(max(n.value, 1), dtype=torch.int32, device="cuda")
        y = torch.zeros(max(n.value, 1), dtype=torch.float32, device="cuda")
        self._check(self.L.risltc_proposal_records(self.h, _dptr(cl), _dptr(y), C.byref(n), _stream(None)),
                    "risltc_proposal_records")
        torch.cuda.synchronize()
        return cl[:n.value], y[:n.value]

    def proposal_cache(self, log2_entries, cell_size=0.5, prior_strength=2.0, per_light=False):
        """Allocates and clears the visibility cache (log2_entries = 0 removes
        it); slots are light clusters, or lights with per_light."""
        self._check(self.L.risltc_proposal_cache(self.h, log2_entries, cell_size, prior_strength,
                                                 1 if per_light else 0), "risltc_proposal_cache")
        self._cache_shape = ((1 << log2_entries), self._light_count if per_light else abi.MAX_LIGHT_CLUSTERS, 2)

    def proposal_cache_read(self):
        """The cache as a uint32 array [cells, slots, 2] (unoccluded, occluded)."""
        out = np.zeros(int(np.prod(self._cache_shape)), dtype=np.uint32)
        self._check(self.L.risltc_proposal_cache_read(self.h, _ptr(out, _u32p), out.size),
                    "risltc_proposal_cache_read")
        return out.reshape(self._cache_shape)

    def proposal_cache_table(self):
        """The cache's bytes [cells, slots] as of the last shading call."""
        out = np.zeros(self._cache_shape[0] * self._cache_shape[1], dtype=np.uint8)
        self._check(self.L.risltc_proposal_cache_table(self.h, _ptr(out, _u8p), out.size),
                    "risltc_proposal_cache_table")
        return out.reshape(self._cache_shape[:2])

    def proposal_cache_write(self, counts):
        c = np.ascontiguousarray(counts, dtype=np.uint32).ravel()
        self._check(self.L.risltc_proposal_cache_write(self.h, _ptr(c, _u32p), c.size), "risltc_proposal_cache_write")

    def proposal_last_timings(self):
        ms = (C.c_float * 2)()
        self._check(self.L.risltc_proposal_last_timings(self.h, C.byref(ms)), "risltc_proposal_last_timings")
        return list(ms)

    def load_scene(self, scene, ltc_rgba16, ltc_rg16):
        """Uploads a risltc_amd.synth.Scene and an LTC table."""
        self.set_lights(scene.light_records, scene.light_count, scene.light_vmax)
        self.set_ltc_table(ltc_rgba16, ltc_rg16)
        self.set_scene(scene.mesh, scene.materials)
        self.scene = scene

    def frame_constants(self, exposure=1.5, roughness_factor=1.0, mis_visibility_estimate=0.5, noise_seed=0,
                        animate=False, camera=None):
        sc = self.scene
        return write_frame_constants(camera or sc.camera, sc.width, sc.height, sc.mesh.dequantization_factor,
                                     sc.mesh.dequantization_summand, self.ltc_constants, exposure,
                                     roughness_factor, mis_visibility_estimate, noise_seed, animate)

    # ---- passes ------------------------------------------------------------
    def render_visibility(self, constants, out, stream=None):
        self._check(self.L.risltc_render_visibility(self.h, C.byref(constants), _dptr(out), _stream(stream)),
                    "risltc_render_visibility")

    def shade(self, constants, settings, visibility, out, stream=None):
        self._check(self.L.risltc_shade(self.h, C.byref(constants), C.byref(settings), _dptr(visibility),
                                        _dptr(out), _stream(stream)), "risltc_shade")

    def shade_noalloc(self, constants, settings, visibility, out, stream=None):
        """risltc_shade on the library's fixed scratch array (no device
        allocation; at most NOALLOC_MAX_SAMPLES pixel-samples)."""
        self._check(self.L.risltc_shade_noalloc(self.h, C.byref(constants), C.byref(settings), _dptr(visibility),
                                                _dptr(out), _stream(stream)), "risltc_shade_noalloc")

    def shade_debug(self, constants, settings, visibility, out, samples, candidates=None, stream=None):
        self._check(self.L.risltc_shade_debug(self.h, C.byref(constants), C.byref(settings), _dptr(visibility),
                                              _dptr(out), _dptr(samples), _dptr(candidates), _stream(stream)),
                    "risltc_shade_debug")

    def accumulate(self, accum, frame, pixel_count, accum_num, stream=None):
        self._check(self.L.risltc_accumulate(self.h, _dptr(accum), _dptr(frame), pixel_count, accum_num,
                                             _stream(stream)), "risltc_accumulate")

    def trace_shadow_rays(self, rays, hits, stream=None):
        n = rays.numel() // 8
        self._check(self.L.risltc_trace_shadow_rays(self.h, _dptr(rays), n, _dptr(hits), _stream(stream)),
                    "risltc_trace_shadow_rays")

    def last_ray_count(self):
        """Shadow rays traced by the last shade call."""
        n = C.c_uint32(0)
        self._check(self.L.risltc_last_ray_count(self.h, C.byref(n)), "risltc_last_ray_count")
        return n.value

    def last_timings(self):
        ms = (C.c_float * 4)()
        self._check(self.L.risltc_last_timings(self.h, ms), "risltc_last_timings")
        return list(ms)

    # ---- batches of views ----------------------------------------------------
    def add_light_set(self, records, count, max_vertex_count):
        """records: bytes / uint8 array as write_lights packs them; returns the id."""
        buf = np.ascontiguousarray(np.frombuffer(bytes(records), dtype=np.uint8))
        out = C.c_uint32(0)
        self._check(self.L.risltc_add_light_set(self.h, buf.ctypes.data_as(_vp), count, max_vertex_count,
                                                C.byref(out)), "risltc_add_light_set")
        return out.value

    def add_scene(self, scene):
        """A further mesh (a risltc_amd.synth.Scene's mesh and materials); returns its id."""
        m = scene.mesh
        pos = np.ascontiguousarray(m.positions, dtype=np.uint32)
        nuv = np.ascontiguousarray(m.normals_uv, dtype=np.uint16)
        mi = np.ascontiguousarray(m.material_index, dtype=np.uint8)
        f = np.ascontiguousarray(m.dequantization_factor, dtype=np.float32)
        a = np.ascontiguousarray(m.dequantization_summand, dtype=np.float32)
        mats = scene.materials_ctypes()
        out = C.c_uint32(0)
        self._check(self.L.risltc_add_scene(self.h, _ptr(pos, _u32p), _ptr(nuv, _u16p), _ptr(mi, _u8p),
                                            m.triangle_count, _ptr(f, _f32p), _ptr(a, _f32p),
                                            C.cast(mats, C.POINTER(abi.Material)), len(mats), C.byref(out)),
                    "risltc_add_scene")
        return out.value

    def render_visibility_scene(self, scene_id, constants, out, stream=None):
        self._check(self.L.risltc_render_visibility_scene(self.h, scene_id, C.byref(constants), _dptr(out),
                                                          _stream(stream)), "risltc_render_visibility_scene")

    @staticmethod
    def view(constants, settings, visibility, out, light_set=0, scene=0, proposal_table=None):
        """A risltc_view_t over torch device tensors (keep them alive)."""
        v = abi.View()
        v.constants = constants
        v.settings = settings
        v.visibility = visibility.data_ptr()
        v.out = out.data_ptr()
        v.light_set = light_set
        v.scene = scene
        v.proposal_table = None if proposal_table is None else proposal_table.data_ptr()
        return v

    @staticmethod
    def _views(views):
        arr = (abi.View * len(views))(*views)
        return arr, len(views)

    def shade_batch(self, views, stream=None):
        arr, n = self._views(views)
        self._check(self.L.risltc_shade_batch(self.h, arr, n, _stream(stream)), "risltc_shade_batch")

    def get_tuning(self):
        t = abi.Tuning()
        self._check(self.L.risltc_get_tuning(self.h, C.byref(t)), "risltc_get_tuning")
        return t

    def set_tuning(self, wave_width=16, cull_chunk=32, lds_lights=1024, trace_refill=32):
        t = abi.Tuning(wave_width, cull_chunk, lds_lights, trace_refill)
        self._check(self.L.risltc_set_tuning(self.h, C.byref(t)), "risltc_set_tuning")

    def autotune(self, views, repeats=3):
        arr, n = self._views(views)
        t = abi.Tuning()
        self._check(self.L.risltc_autotune(self.h, arr, n, repeats, C.byref(t)), "risltc_autotune")
        return t

    def graph(self, views):
        """A hipGraph of a batch (risltc_graph_t); see Graph."""
        return Graph(self, views)
