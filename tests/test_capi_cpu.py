"""The C-ABI library without a GPU: it loads, exports every function that
include/risltc/risltc.h declares, and its host-side restatements of the
reference's C code agree bit for bit with the oracle (and so with the
reference build, see test_oracle_pins.py). Error paths that the host can
force are checked here too."""
import ctypes as C
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import orc
from risltc_amd import abi, ltcfit, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "risltc", "risltc.h")


@pytest.fixture(scope="module")
def hiplib():
    import risltc_amd
    from risltc_amd import build
    if not os.path.exists(risltc_amd.LIB_PATH):
        build.build()
    return risltc_amd.lib()


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(risltc_[a-z_0-9]+)\s*\(", text)) - {"risltc_light_record_size"})


def test_library_exports_every_declared_function(hiplib):
    import risltc_amd
    names = declared_functions()
    assert len(names) >= 18
    out = subprocess.run(["nm", "-D", "--defined-only", risltc_amd.LIB_PATH], stdout=subprocess.PIPE,
                         check=True).stdout.decode()
    exported = set(re.findall(r" T (risltc_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(hiplib, n)


def test_noalloc_limit_matches_header():
    text = open(HEADER).read()
    m = re.search(r"#define RISLTC_NOALLOC_MAX_SAMPLES \(1u << (\d+)\)", text)
    assert m and (1 << int(m.group(1))) == abi.NOALLOC_MAX_SAMPLES
    m = re.search(r"#define RISLTC_MAX_SAMPLE_COUNT (\d+)u", text)
    assert m and int(m.group(1)) == abi.MAX_SAMPLE_COUNT


def test_struct_layouts():
    assert C.sizeof(abi.FrameConstants) == 256
    offs = {f[0]: getattr(abi.FrameConstants, f[0]).offset for f in abi.FrameConstants._fields_}
    # std140 offsets of the UBO (src/shaders/shared_constants.glsl:21-60)
    assert offs["error_factor"] == 28
    assert offs["world_to_projection_space"] == 32
    assert offs["pixel_to_ray_direction_world_space"] == 96
    assert offs["camera_position_world_space"] == 144
    assert offs["mis_visibility_estimate"] == 156
    assert offs["viewport_size"] == 160
    assert offs["exposure_factor"] == 176
    assert offs["noise_random_numbers"] == 208
    assert offs["ltc_constants"] == 224
    assert abi.light_record_size(8) == 176


def test_update_polygonal_light_matches_oracle(hiplib, orc_lib):
    rng = np.random.default_rng(21)
    for i in range(200):
        n = int(rng.integers(3, 9))
        L = synth.Light(synth.random_convex_polygon(rng, n) * rng.uniform(0.1, 2), rng.uniform(0, 7, 3),
                        rng.uniform(-9, 9, 3), rng.uniform(0.2, 3, 2), rng.uniform(0, 10, 3))
        a = L.to_ctypes()
        keep_a = (L._ps, L._ws)
        b = L.to_ctypes()
        hiplib.risltc_update_polygonal_light(C.byref(a))
        orc_lib.orc_update_polygonal_light(C.byref(b))
        wa = np.ctypeslib.as_array(a.vertices_world_space, shape=(4 * n,))
        wb = np.ctypeslib.as_array(b.vertices_world_space, shape=(4 * n,))
        assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))
        assert np.array_equal(np.ctypeslib.as_array(a.plane).view(np.uint32),
                              np.ctypeslib.as_array(b.plane).view(np.uint32))
        del keep_a


def test_write_lights_layout(hiplib):
    """write_lights (src/main.c:456-490): 48-byte header + 16 bytes per vertex
    slot; every unused slot repeats vertex 0."""
    rng = np.random.default_rng(22)
    lights = [synth.Light(synth.random_convex_polygon(rng, n), rng.uniform(0, 7, 3), rng.uniform(-9, 9, 3),
                          (1.0, 1.0), rng.uniform(0, 10, 3)) for n in (3, 5, 8)]
    arr = (abi.PolygonalLight * 3)()
    keep = []
    for i, L in enumerate(lights):
        arr[i] = L.to_ctypes()
        keep.append((L._ps, L._ws))
    out = (C.c_uint8 * (3 * 176))()
    n = hiplib.risltc_write_lights(arr, 3, 8, out, C.sizeof(out))
    assert n == 3 * 176
    buf = np.frombuffer(bytes(out), dtype=np.float32).reshape(3, 44)
    for i, L in enumerate(lights):
        ws = np.ctypeslib.as_array(arr[i].vertices_world_space, shape=(4 * L.vertex_count,)).reshape(-1, 4)
        assert buf[i].view(np.uint32)[8] == L.vertex_count
        assert np.array_equal(buf[i, 0:3], np.ctypeslib.as_array(arr[i].surface_radiance))
        for k in range(8):
            src = ws[k if k < L.vertex_count else 0, :3]
            assert np.array_equal(buf[i, 12 + 4 * k:15 + 4 * k], src)
    # a vertex count above max_vertex_count is rejected
    assert hiplib.risltc_write_lights(arr, 3, 5, out, C.sizeof(out)) == 0


def test_frame_constants_match_oracle(hiplib, orc_lib):
    scene = synth.small_scene(seed=2)
    k = abi.LtcConstants()
    hiplib.risltc_ltc_constants_for(C.byref(k), 64, 64, 51)
    k2 = abi.LtcConstants()
    orc_lib.orc_ltc_constants(C.byref(k2), 64, 64, 51)
    assert bytes(k) == bytes(k2)
    f = np.ascontiguousarray(scene.mesh.dequantization_factor, np.float32)
    s = np.ascontiguousarray(scene.mesh.dequantization_summand, np.float32)
    fp = f.ctypes.data_as(C.POINTER(C.c_float))
    sp = s.ctypes.data_as(C.POINTER(C.c_float))
    for animate in (0, 1):
        for seed in (0, 7):
            a = abi.FrameConstants()
            b = abi.FrameConstants()
            hiplib.risltc_write_frame_constants(C.byref(a), C.byref(scene.camera), 1920, 1080, fp, sp, C.byref(k),
                                                1.5, 1.0, 0.5, seed, animate)
            orc_lib.orc_write_frame_constants(C.byref(b), C.byref(scene.camera), 1920, 1080, fp, sp, C.byref(k),
                                              1.5, 1.0, 0.5, seed, animate)
            assert bytes(a) == bytes(b)
    # set_noise_constants (src/noise_table.c:24-28) without animation
    assert list(a.noise_random_numbers) != [0, 0x123456, 0x2468AC, 0x369D02]
    b = abi.FrameConstants()
    hiplib.risltc_write_frame_constants(C.byref(b), C.byref(scene.camera), 8, 8, fp, sp, C.byref(k), 1.5, 1.0,
                                        0.5, 3, 0)
    assert list(b.noise_random_numbers) == [0, 0x123456, 0x2468AC, 0x369D02]


def test_ltc_fit_reader(hiplib, tmp_path):
    """load_ltc_table file handling (src/ltc_table.c:31-116) against the
    vectorised quantiser, and its error paths."""
    fit, i5 = ltcfit.load_fit()
    d = str(tmp_path / "fit")
    res = ltcfit.write_fit_dir(d, 4, fit, i5)
    r = C.c_uint32()
    assert hiplib.risltc_load_ltc_fit(d.encode(), 4, C.byref(r), None, None, 0) == 0
    assert r.value == res
    rgba = np.zeros(4 * res * res * 4, np.uint16)
    rg = np.zeros(4 * res * res * 2, np.uint16)
    assert hiplib.risltc_load_ltc_fit(d.encode(), 4, C.byref(r), rgba.ctypes.data_as(C.POINTER(C.c_uint16)),
                                      rg.ctypes.data_as(C.POINTER(C.c_uint16)), rgba.size) == 0
    q4, q2 = ltcfit.quantize(fit, i5, 4)
    assert np.array_equal(rgba, q4.ravel())
    assert np.array_equal(rg, q2.ravel())
    # missing file
    assert hiplib.risltc_load_ltc_fit((d + "_none").encode(), 4, C.byref(r), None, None, 0) == abi.E_IO
    # inconsistent resolution
    with open(os.path.join(d, "fit2.dat"), "wb") as fh:
        fh.write(struct.pack("<Q", res + 1))
    assert hiplib.risltc_load_ltc_fit(d.encode(), 4, C.byref(r), rgba.ctypes.data_as(C.POINTER(C.c_uint16)),
                                      rg.ctypes.data_as(C.POINTER(C.c_uint16)), rgba.size) == abi.E_FORMAT
    # truncated file
    with open(os.path.join(d, "fit2.dat"), "wb") as fh:
        fh.write(struct.pack("<Q", res) + b"\0" * 100)
    assert hiplib.risltc_load_ltc_fit(d.encode(), 4, C.byref(r), rgba.ctypes.data_as(C.POINTER(C.c_uint16)),
                                      rg.ctypes.data_as(C.POINTER(C.c_uint16)), rgba.size) == abi.E_FORMAT


def test_null_context_and_argument_errors(hiplib):
    assert hiplib.risltc_set_lights(None, None, 0, 4) == abi.E_ARG
    assert hiplib.risltc_shade(None, None, None, None, None, None) == abi.E_ARG
    assert hiplib.risltc_last_error(None) == b"null context"
    assert hiplib.risltc_create(None, 0) == abi.E_ARG
    s = abi.Settings()
    hiplib.risltc_default_settings(C.byref(s))
    d = abi.default_settings()
    assert bytes(s) == bytes(d)


def test_header_layouts_match_python(tmp_path):
    """The ctypes mirror (risltc_amd/abi.py, risltc_amd/formats.py) has the
    sizes and field offsets that the C compiler gives the structs of
    include/risltc/risltc.h."""
    import shutil
    from risltc_amd import formats
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    structs = {
        "risltc_frame_constants_t": abi.FrameConstants, "risltc_settings_t": abi.Settings,
        "risltc_view_t": abi.View, "risltc_tuning_t": abi.Tuning, "risltc_camera_t": abi.Camera,
        "risltc_polygonal_light_t": abi.PolygonalLight, "risltc_material_t": abi.Material,
        "risltc_ltc_constants_t": abi.LtcConstants, "risltc_vks_info_t": formats.VksInfo,
        "risltc_quicksave_info_t": formats.QuicksaveInfo,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "risltc/risltc.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, check=True).stdout.decode().split("\n")
    got = {}
    for line in out:
        if line:
            a, b, v = line.split()
            got[(a, b)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])
