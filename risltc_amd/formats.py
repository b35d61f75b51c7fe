"""The reference's on-disk formats around the shading pass.

Readers and the HDR writer run in librisltc.so (risltc_load_vks,
risltc_load_quicksave, risltc_save_quicksave, risltc_write_hdr); this module
binds them and adds two pure-Python helpers:

  save_vks  -- writes a synth.Mesh in the byte layout of the reference's
               Blender exporter (io_export_vulkan_blender28.py:481-541), so
               that synthetic scenes can be stored as *.vks files;
  read_hdr  -- decodes a Radiance RGBE file (flat or run-length coded
               scanlines) into float32 RGB.
"""
import ctypes as C
import struct

import numpy as np

from . import abi, lib, RisltcError


class VksInfo(C.Structure):
    """risltc_vks_info_t."""
    _fields_ = [("material_count", C.c_uint64), ("triangle_count", C.c_uint64),
                ("dequantization_factor", C.c_float * 3), ("dequantization_summand", C.c_float * 3),
                ("material_name_bytes", C.c_uint64)]


class QuicksaveInfo(C.Structure):
    """risltc_quicksave_info_t."""
    _fields_ = [("light_count", C.c_uint32), ("total_vertex_count", C.c_uint32),
                ("texture_path_bytes", C.c_uint64)]


_SIGS = {
    "risltc_load_vks": (C.c_int, [C.c_char_p, C.POINTER(VksInfo), C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                  C.c_char_p, C.c_uint64]),
    "risltc_load_quicksave": (C.c_int, [C.c_char_p, C.POINTER(abi.Camera), C.POINTER(QuicksaveInfo),
                                        C.POINTER(abi.PolygonalLight), C.c_uint32, C.POINTER(C.c_float), C.c_uint64,
                                        C.c_char_p, C.c_uint64]),
    "risltc_save_quicksave": (C.c_int, [C.c_char_p, C.POINTER(abi.Camera), C.POINTER(abi.PolygonalLight),
                                        C.c_uint32]),
    "risltc_write_hdr": (C.c_int, [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int]),
    "risltc_write_lights": (C.c_size_t, [C.POINTER(abi.PolygonalLight), C.c_uint32, C.c_uint32, C.c_void_p,
                                         C.c_size_t]),
}


def _fn(name):
    f = getattr(lib(), name)
    f.restype, f.argtypes = _SIGS[name]
    return f


def _check(code, what):
    if code != 0:
        raise RisltcError(code, what)


def _path(p):
    return str(p).encode()


class VksScene:
    """Contents of a *.vks file: the mesh buffers risltc_set_scene takes."""

    def __init__(self, info, positions, normals_uv, material_index, material_names):
        self.triangle_count = int(info.triangle_count)
        self.dequantization_factor = np.array(info.dequantization_factor[:], dtype=np.float32)
        self.dequantization_summand = np.array(info.dequantization_summand[:], dtype=np.float32)
        self.positions = positions          # (3 * triangles, 2) uint32
        self.normals_uv = normals_uv        # (3 * triangles, 4) uint16
        self.material_index = material_index  # (triangles,) uint8
        self.material_names = material_names


def load_vks(path):
    """Reads a *.vks file with risltc_load_vks (src/scene.c:409-483)."""
    f = _fn("risltc_load_vks")
    info = VksInfo()
    _check(f(_path(path), C.byref(info), None, None, None, 0, None, 0), "load_vks(%s)" % path)
    t = int(info.triangle_count)
    pos = np.empty((3 * t, 2), dtype=np.uint32)
    nuv = np.empty((3 * t, 4), dtype=np.uint16)
    mat = np.empty(t, dtype=np.uint8)
    nb = max(1, int(info.material_name_bytes))
    names = C.create_string_buffer(nb)
    _check(f(_path(path), C.byref(info), pos.ctypes.data, nuv.ctypes.data, mat.ctypes.data, t, names, nb),
           "load_vks(%s)" % path)
    raw = names.raw[:int(info.material_name_bytes)]
    material_names = [s.decode("utf-8") for s in raw.split(b"\0")[:-1]] if raw else []
    return VksScene(info, pos, nuv, mat, material_names)


def save_vks(path, mesh, material_names=None):
    """Writes a synth.Mesh (or VksScene) as a *.vks file: the byte layout of
    io_export_vulkan_blender28.py:481-541, which load_scene reads."""
    tri = int(mesh.triangle_count)
    names = list(material_names) if material_names is not None else \
        ["material_%d" % i for i in range(int(np.max(mesh.material_index)) + 1)]
    with open(path, "wb") as fh:
        fh.write(struct.pack("<II", 0xABCABC, 1))
        fh.write(struct.pack("<QQ", len(names), tri))
        fh.write(np.asarray(mesh.dequantization_factor, np.float32).tobytes())
        fh.write(np.asarray(mesh.dequantization_summand, np.float32).tobytes())
        for n in names:
            b = n.encode("utf-8")
            fh.write(struct.pack("<Q", len(b)) + b + b"\0")
        fh.write(np.ascontiguousarray(mesh.positions, np.uint32).tobytes())
        fh.write(np.ascontiguousarray(mesh.normals_uv, np.uint16).tobytes())
        fh.write(np.ascontiguousarray(mesh.material_index, np.uint8).tobytes())
        fh.write(struct.pack("<I", 0xE0FE0F))


class Quicksave:
    """Camera and lights of a quicksave file. `lights` is a ctypes array of
    polygonal_light_t whose vertex and path pointers point into buffers this
    object keeps alive."""

    def __init__(self, camera, lights, storage, paths):
        self.camera = camera
        self.lights = lights
        self._storage = storage
        self._paths = paths

    @property
    def light_count(self):
        return len(self.lights)

    def light_records(self, max_vertex_count=None):
        """Records for risltc_set_lights, packed by risltc_write_lights
        (which runs update_polygonal_light on every light first)."""
        n = len(self.lights)
        vmax = max_vertex_count or max(int(l.vertex_count) for l in self.lights)
        size = n * abi.light_record_size(vmax)
        out = (C.c_uint8 * size)()
        wrote = _fn("risltc_write_lights")(self.lights, n, vmax, out, size)
        if wrote != size:
            raise RisltcError(1, "write_lights: a light has more than %d vertices" % vmax)
        return bytes(out), vmax


def load_quicksave(path):
    """Reads a quicksave with risltc_load_quicksave (src/main.c:78-125)."""
    f = _fn("risltc_load_quicksave")
    info = QuicksaveInfo()
    cam = abi.Camera()
    _check(f(_path(path), C.byref(cam), C.byref(info), None, 0, None, 0, None, 0), "load_quicksave(%s)" % path)
    n = int(info.light_count)
    lights = (abi.PolygonalLight * max(1, n))()
    floats = 8 * int(info.total_vertex_count)
    storage = (C.c_float * max(1, floats))()
    nb = max(1, int(info.texture_path_bytes))
    paths = C.create_string_buffer(nb)
    _check(f(_path(path), C.byref(cam), C.byref(info), lights, n, storage, floats, paths, nb),
           "load_quicksave(%s)" % path)
    if n == 0:
        lights = (abi.PolygonalLight * 0)()
    return Quicksave(cam, lights, storage, paths)


def save_quicksave(path, camera, lights):
    """Writes camera and lights (a sequence of abi.PolygonalLight) with
    risltc_save_quicksave (src/main.c:44-75)."""
    n = len(lights)
    arr = (abi.PolygonalLight * max(1, n))()
    for i, l in enumerate(lights):
        arr[i] = l
    _check(_fn("risltc_save_quicksave")(_path(path), C.byref(camera), arr, n), "save_quicksave(%s)" % path)


def write_hdr(path, image, half_round=True):
    """Writes an (H, W, 3 or 4) float32 image as Radiance RGBE with
    risltc_write_hdr. half_round=True stores what the reference's HDR
    screenshots hold (values rounded to binary16)."""
    img = np.ascontiguousarray(image, dtype=np.float32)
    if img.ndim != 3 or img.shape[2] not in (3, 4):
        raise ValueError("image must be (H, W, 3) or (H, W, 4)")
    h, w, ch = img.shape
    _check(_fn("risltc_write_hdr")(_path(path), img.ctypes.data, w, h, ch, 1 if half_round else 0),
           "write_hdr(%s)" % path)


def read_hdr(path):
    """Decodes a Radiance RGBE file with a "-Y h +X w" resolution line.
    Returns (rgb, rgbe): rgb is (h, w, 3) float32 with value =
    (byte + 0.5) * 2^(e - 136) and 0 where e = 0, rgbe the (h, w, 4) bytes."""
    data = open(path, "rb").read()
    if not data.startswith(b"#?"):
        raise ValueError("not a Radiance file")
    pos = 0
    fmt = None
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end]
        pos = end + 1
        if line.startswith(b"FORMAT="):
            fmt = line[7:]
        if line == b"":
            break
    if fmt != b"32-bit_rle_rgbe":
        raise ValueError("unsupported format %r" % fmt)
    end = data.index(b"\n", pos)
    res = data[pos:end].split()
    pos = end + 1
    if len(res) != 4 or res[0] != b"-Y" or res[2] != b"+X":
        raise ValueError("unsupported resolution line")
    h, w = int(res[1]), int(res[3])
    rgbe = np.zeros((h, w, 4), dtype=np.uint8)
    buf = np.frombuffer(data, dtype=np.uint8)
    for y in range(h):
        if 8 <= w < 32768 and pos + 4 <= len(data) and data[pos] == 2 and data[pos + 1] == 2 and \
                data[pos + 2] < 128:
            if (data[pos + 2] << 8) | data[pos + 3] != w:
                raise ValueError("scanline width mismatch")
            pos += 4
            for k in range(4):
                x = 0
                while x < w:
                    c = data[pos]
                    pos += 1
                    if c > 128:
                        n = c - 128
                        rgbe[y, x:x + n, k] = data[pos]
                        pos += 1
                    else:
                        n = c
                        if n == 0:
                            raise ValueError("zero-length literal packet")
                        rgbe[y, x:x + n, k] = buf[pos:pos + n]
                        pos += n
                    x += n
                if x != w:
                    raise ValueError("scanline overrun")
        else:
            rgbe[y] = buf[pos:pos + 4 * w].reshape(w, 4)
            pos += 4 * w
    e = rgbe[..., 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    return ((rgbe[..., :3].astype(np.float64) + 0.5) * scale[..., None]).astype(np.float32), rgbe
