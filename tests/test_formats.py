"""On-disk formats in librisltc.so (no GPU): the *.vks reader, the quicksave
reader/writer and the Radiance HDR writer.

Pins: hand-assembled files laid out field by field as the reference reads
them (src/scene.c:409-483, src/main.c:78-125), the HDR writer's bytes
against stbi_write_hdr compiled from the reference tree (oracle/_ref), the
reference's half_to_float against IEEE binary16, and decode(encode(x))
against the RGBE precision bound."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

import orc
from risltc_amd import abi, formats, synth


@pytest.fixture(scope="module")
def hiplib():
    import risltc_amd
    from risltc_amd import build
    if not os.path.exists(risltc_amd.LIB_PATH):
        build.build()
    return risltc_amd.lib()


def _err(fn, *a):
    with pytest.raises(Exception) as e:
        fn(*a)
    return getattr(e.value, "code", None)


# ---------------------------------------------------------------------------
# *.vks
# ---------------------------------------------------------------------------


def _hand_vks(tri=2, names=(b"stone", b"light.glass"), marker=0xABCABC, version=1, end=0xE0FE0F, cut=0):
    """A *.vks file assembled field by field in the order load_scene reads."""
    rng = np.random.default_rng(tri)
    pos = rng.integers(0, 2 ** 32, (3 * tri, 2), dtype=np.uint64).astype(np.uint32)
    nuv = rng.integers(0, 2 ** 16, (3 * tri, 4), dtype=np.uint64).astype(np.uint16)
    mat = rng.integers(0, 256, tri, dtype=np.uint64).astype(np.uint8)
    out = struct.pack("<IIQQ", marker, version, len(names), tri)
    out += struct.pack("<6f", 0.5, 0.25, 0.125, -1.0, -2.0, -3.0)
    for n in names:
        out += struct.pack("<Q", len(n)) + n + b"\0"
    out += pos.tobytes() + nuv.tobytes() + mat.tobytes() + struct.pack("<I", end)
    if cut:
        out = out[:-cut]
    return out, pos, nuv, mat


def test_vks_hand_assembled(hiplib, tmp_path):
    data, pos, nuv, mat = _hand_vks()
    p = tmp_path / "a.vks"
    p.write_bytes(data)
    s = formats.load_vks(p)
    assert s.triangle_count == 2
    assert s.material_names == ["stone", "light.glass"]
    assert list(s.dequantization_factor) == [0.5, 0.25, 0.125]
    assert list(s.dequantization_summand) == [-1.0, -2.0, -3.0]
    assert np.array_equal(s.positions, pos)
    assert np.array_equal(s.normals_uv, nuv)
    assert np.array_equal(s.material_index, mat)


def test_vks_header_query_and_capacity(hiplib, tmp_path):
    data, pos, nuv, mat = _hand_vks(tri=5)
    p = tmp_path / "b.vks"
    p.write_bytes(data)
    f = formats._fn("risltc_load_vks")
    info = formats.VksInfo()
    assert f(str(p).encode(), C.byref(info), None, None, None, 0, None, 0) == 0
    assert (info.material_count, info.triangle_count, info.material_name_bytes) == (2, 5, 6 + 12)
    a = np.empty((15, 2), np.uint32)
    b = np.empty((15, 4), np.uint16)
    c = np.empty(5, np.uint8)
    # triangle capacity too small, a missing buffer, names buffer too small
    assert f(str(p).encode(), C.byref(info), a.ctypes.data, b.ctypes.data, c.ctypes.data, 4, None, 0) == 1
    assert f(str(p).encode(), C.byref(info), a.ctypes.data, None, c.ctypes.data, 5, None, 0) == 1
    small = C.create_string_buffer(10)
    assert f(str(p).encode(), C.byref(info), a.ctypes.data, b.ctypes.data, c.ctypes.data, 5, small, 10) == 1


@pytest.mark.parametrize("kw,code", [
    (dict(marker=0xABCABD), 3), (dict(version=2), 3), (dict(tri=0), 3),
    (dict(end=0xE0FE0E), 3), (dict(cut=1), 3), (dict(cut=5), 3), (dict(cut=200), 3)])
def test_vks_rejects_malformed(hiplib, tmp_path, kw, code):
    data = _hand_vks(**kw)[0]
    p = tmp_path / "bad.vks"
    p.write_bytes(data)
    assert _err(formats.load_vks, p) == code


def test_vks_rejects_impossible_counts(hiplib, tmp_path):
    data = bytearray(_hand_vks()[0])
    data[16:24] = struct.pack("<Q", 1 << 62)  # triangle count
    p = tmp_path / "huge.vks"
    p.write_bytes(bytes(data))
    assert _err(formats.load_vks, p) == 3


def test_vks_missing_file(hiplib, tmp_path):
    assert _err(formats.load_vks, tmp_path / "none.vks") == 6


def test_vks_round_trip_of_synthetic_scene(hiplib, tmp_path):
    """save_vks (the exporter's layout) -> risltc_load_vks returns the mesh
    buffers risltc_set_scene takes, unchanged."""
    scene = synth.small_scene(seed=3)
    p = tmp_path / "small.vks"
    names = ["m%d" % i for i in range(len(scene.materials))]
    formats.save_vks(p, scene.mesh, names)
    s = formats.load_vks(p)
    m = scene.mesh
    assert s.triangle_count == m.triangle_count and s.material_names == names
    assert np.array_equal(s.positions, m.positions)
    assert np.array_equal(s.normals_uv, m.normals_uv)
    assert np.array_equal(s.material_index, m.material_index)
    assert np.array_equal(s.dequantization_factor, m.dequantization_factor)
    assert np.array_equal(s.dequantization_summand, m.dequantization_summand)


# ---------------------------------------------------------------------------
# quicksave
# ---------------------------------------------------------------------------


def _camera():
    cam = synth.make_camera((1.0, -2.0, 3.5), 0.25, -0.5, fov=0.9)
    cam.rotate_camera = 1
    cam.rotation_x_0, cam.rotation_z_0 = 0.125, 2.0
    return cam


def _light_bytes(rng, n, scaling_y=None):
    head = rng.normal(size=20).astype(np.float32)
    if scaling_y is not None:
        head[7] = scaling_y
    return head, struct.pack("<20f", *head) + struct.pack("<II", n, 1)


def test_quicksave_hand_assembled(hiplib, tmp_path):
    """Field order of quick_load (src/main.c:78-125), with a texture path and
    a legacy light (scaling_y <= 0 takes scaling_x)."""
    rng = np.random.default_rng(1)
    cam = _camera()
    data = bytes(cam) + struct.pack("<II", 7, 2)
    h0, b0 = _light_bytes(rng, 3)
    v0 = rng.normal(size=12).astype(np.float32)
    data += b0 + struct.pack("<Q", 9) + b"tex/a.png\0"[:9] + struct.pack("<QQ", 0, 0) + v0.tobytes()
    h1, b1 = _light_bytes(rng, 5, scaling_y=-1.0)
    v1 = rng.normal(size=20).astype(np.float32)
    data += b1 + struct.pack("<Q", 0) + struct.pack("<QQ", 0xDEAD, 0xBEEF) + v1.tobytes()
    p = tmp_path / "q.save"
    p.write_bytes(data)
    q = formats.load_quicksave(p)
    assert bytes(q.camera) == bytes(cam)
    assert q.light_count == 2
    l0, l1 = q.lights[0], q.lights[1]
    assert bytes(l0)[:80] == h0.tobytes() and l0.vertex_count == 3 and l0.texturing_technique == 1
    assert l0.texture_file_path == b"tex/a.pn"  # 9 bytes, the last one becomes the terminator
    assert np.array_equal(np.ctypeslib.as_array(l0.vertices_plane_space, shape=(12,)), v0)
    assert np.all(np.ctypeslib.as_array(l0.vertices_world_space, shape=(12,)) == 0)
    assert l1.scaling_y == l1.scaling_x == h1[3] and l1.vertex_count == 5 and not l1.texture_file_path
    assert np.array_equal(np.ctypeslib.as_array(l1.vertices_plane_space, shape=(20,)), v1)
    for cut in (1, 17, len(data) - 60):
        p.write_bytes(data[:-cut])
        assert _err(formats.load_quicksave, p) == 3


def test_quicksave_rejects_bad_vertex_count(hiplib, tmp_path):
    rng = np.random.default_rng(2)
    data = bytes(_camera()) + struct.pack("<II", 0, 1) + _light_bytes(rng, 0)[1] + struct.pack("<QQQ", 0, 0, 0)
    p = tmp_path / "z.save"
    p.write_bytes(data)
    assert _err(formats.load_quicksave, p) == 3
    assert _err(formats.load_quicksave, tmp_path / "absent.save") == 6


def test_quicksave_round_trip_and_light_records(hiplib, tmp_path):
    """quick_save -> quick_load keeps the camera, the 88 saved bytes and the
    plane-space vertices of every light; the lights then pack to the same
    records as the originals."""
    scene = synth.small_scene(seed=4, light_count=9)
    lights = [l.to_ctypes() for l in scene.lights]
    keep = [(l._ps, l._ws) for l in scene.lights]
    path_buf = C.create_string_buffer(b"textures/panel.png")
    lights[2].texture_file_path = C.cast(path_buf, C.c_char_p)
    cam = _camera()
    p = tmp_path / "rt.save"
    formats.save_quicksave(p, cam, lights)
    size = 48 + 8 + sum(88 + 8 + 16 + 16 * l.vertex_count for l in lights) + len(b"textures/panel.png") + 1
    assert os.path.getsize(p) == size
    q = formats.load_quicksave(p)
    assert bytes(q.camera) == bytes(cam) and q.light_count == len(lights)
    for a, b in zip(lights, q.lights):
        assert bytes(a)[:88] == bytes(b)[:88]
        n = 4 * a.vertex_count
        assert np.array_equal(np.ctypeslib.as_array(a.vertices_plane_space, shape=(n,)),
                              np.ctypeslib.as_array(b.vertices_plane_space, shape=(n,)))
    assert q.lights[2].texture_file_path == b"textures/panel.png"
    recs, vmax = q.light_records()
    arr = (abi.PolygonalLight * len(lights))(*lights)
    out = (C.c_uint8 * len(recs))()
    assert hiplib.risltc_write_lights(arr, len(lights), vmax, out, len(recs)) == len(recs)
    assert bytes(out) == recs
    del keep


# ---------------------------------------------------------------------------
# HDR
# ---------------------------------------------------------------------------


def _image(rng, h, w):
    img = rng.gamma(0.7, 2.0, (h, w, 4)).astype(np.float32)
    img[rng.random((h, w)) < 0.1] = 0.0
    # flat stretches so that run packets (also longer than 127) occur
    img[0, : min(w, 300)] = 0.75
    if h > 1:
        img[1, ::2] = 3.0
    img[-1, -1, :3] = [1e-33, 2e-33, 0.0]
    img[h // 2, w // 2, :3] = [6.0e4, 1.0, 1e-5]
    return img


def _stb_bytes(stb, path, img3):
    a = np.ascontiguousarray(img3, np.float32)
    assert stb.stbi_write_hdr(str(path).encode(), a.shape[1], a.shape[0], 3, a.ctypes.data) == 1
    return open(path, "rb").read()


def _split_header(data):
    i = data.index(b"\n\n")
    j = data.index(b"\n", i + 2)
    lines = data[:i].split(b"\n")
    return lines, data[i + 2:j], data[j + 1:]


@pytest.mark.parametrize("h,w", [(1, 1), (3, 7), (4, 8), (5, 33), (6, 300), (2, 1031)])
@pytest.mark.parametrize("half_round", [0, 1])
def test_hdr_bytes_match_stbi_write_hdr(hiplib, tmp_path, h, w, half_round):
    orc.build(reference="/root/reference") if os.path.isdir("/root/reference/src") else None
    stb = orc.ref_stbiw_lib()
    if stb is None:
        pytest.skip("stb_image_write from the reference tree is not built")
    rng = np.random.default_rng(h * 1000 + w)
    img = _image(rng, h, w)
    ours = tmp_path / "ours.hdr"
    formats.write_hdr(ours, img, half_round=bool(half_round))
    src = img[..., :3].astype(np.float16).astype(np.float32) if half_round else img[..., :3]
    ref = _stb_bytes(stb, tmp_path / "stb.hdr", src)
    a_lines, a_res, a_px = _split_header(open(ours, "rb").read())
    b_lines, b_res, b_px = _split_header(ref)
    assert a_res == b_res == b"-Y %d +X %d" % (h, w)
    assert a_lines[0] == b_lines[0] == b"#?RADIANCE"
    assert b"FORMAT=32-bit_rle_rgbe" in a_lines and b"FORMAT=32-bit_rle_rgbe" in b_lines
    assert a_px == b_px


@pytest.mark.parametrize("w", [5, 64, 517])
def test_hdr_decodes_within_rgbe_precision(hiplib, tmp_path, w):
    rng = np.random.default_rng(w)
    img = _image(rng, 4, w)[..., :3]
    p = tmp_path / "x.hdr"
    formats.write_hdr(p, img, half_round=False)
    rgb, rgbe = formats.read_hdr(p)
    m = img.max(axis=2, keepdims=True).astype(np.float64)
    err = np.abs(rgb.astype(np.float64) - img)
    # a mantissa byte keeps 8 bits below the shared exponent: error < m/128
    assert np.all(err <= m / 128.0 + 1e-30)
    assert np.all(rgbe[img.max(axis=2) < 1e-32] == 0)


def test_hdr_reader_on_hand_coded_scanline(tmp_path):
    """A run packet (128 + n, byte) and a literal packet (n, bytes) per
    component, as the Radiance scanline format defines them."""
    w = 10
    line = bytes([2, 2, 0, w])
    line += bytes([128 + 6, 128]) + bytes([4, 1, 2, 3, 4])  # R
    line += bytes([128 + 10, 64])  # G
    line += bytes([10] + list(range(10)))  # B
    line += bytes([128 + 10, 129])  # E: 2^(129-136) * (b + 0.5)
    p = tmp_path / "hand.hdr"
    p.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 10\n" + line)
    rgb, rgbe = formats.read_hdr(p)
    assert list(rgbe[0, :, 0]) == [128] * 6 + [1, 2, 3, 4]
    assert np.allclose(rgb[0, :, 2], (np.arange(10) + 0.5) * 2.0 ** -7)
    assert np.allclose(rgb[0, :, 1], 64.5 * 2.0 ** -7)


def test_hdr_half_rounding_is_ieee(hiplib, tmp_path):
    """half_round stores binary16 round-to-nearest-even values: the file of
    x equals the file of float16(x) written without rounding."""
    rng = np.random.default_rng(9)
    img = rng.uniform(0, 10, (3, 40, 3)).astype(np.float32)
    img[0, :4, 0] = [1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11, 2049.0, 65519.0]
    a, b = tmp_path / "a.hdr", tmp_path / "b.hdr"
    formats.write_hdr(a, img, half_round=True)
    formats.write_hdr(b, img.astype(np.float16).astype(np.float32), half_round=False)
    assert open(a, "rb").read() == open(b, "rb").read()


def test_reference_half_to_float_is_ieee_binary16(orc_lib):
    """The screenshot path decodes halves with half_to_float
    (src/math_utilities.h:60-80); it agrees with IEEE binary16 for every
    finite bit pattern, so rounding to float16 and back is what the
    reference's HDR screenshots hold."""
    if os.path.isdir("/root/reference/src"):
        orc.build(reference="/root/reference")
    ref = orc.ref_lib()
    if ref is None or not hasattr(ref, "ref_half_to_float"):
        pytest.skip("reference host code not built")
    bits = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    want = bits.view(np.float16).astype(np.float32)
    finite = np.isfinite(want)
    got = np.array([ref.ref_half_to_float(int(b)) for b in bits[finite]], dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), want[finite].view(np.uint32))


def test_hdr_argument_errors(hiplib, tmp_path):
    f = formats._fn("risltc_write_hdr")
    img = np.zeros((2, 2, 3), np.float32)
    assert f(str(tmp_path / "a.hdr").encode(), img.ctypes.data, 2, 2, 2, 0) == 1
    assert f(str(tmp_path / "a.hdr").encode(), img.ctypes.data, 0, 2, 3, 0) == 1
    assert f(str(tmp_path / "nodir" / "a.hdr").encode(), img.ctypes.data, 2, 2, 3, 0) == 6
