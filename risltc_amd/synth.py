"""Seeded synthetic workloads for the shading pass.

The reference renders downloaded scenes (README.md:8-11) that are not
available here, so every input is generated: a quantised triangle mesh in the
*.vks layout (src/scene.h:47-70; encoding as in
tools/io_export_vulkan_blender28.py:22-46, 495-539), constant materials,
convex polygonal lights in the polygonal_light_t form (src/polygonal_light.h:
73-99) and a first-person camera (src/camera.h:29-49).

Workloads:
  small : 256x256, 20 m ground plane + 16 boxes, 64 quad lights (sides
          U[0.1, 1] m, centres in a 10 m cube above the plane), RIS M = 8, 1 spp.
  city  : 1920x1080, procedural city of ~1M triangles (boxes with heights
          log-uniform 1-50 m), 1024 convex lights with 3-8 vertices, areas
          log-uniform 0.01-4 m^2, radiance log-uniform 0.5-50, M = 32, 4 spp.
"""
import ctypes as C
import math

import numpy as np

from . import abi

# ---------------------------------------------------------------------------
# mesh encoding
# ---------------------------------------------------------------------------


def encode_normal_32_bit(normal):
    """Octahedral UNORM16 encoding (tools/io_export_vulkan_blender28.py:22-46)."""
    normal = np.asarray(normal, dtype=np.float64)
    oct_ = normal[:, 0:2] / np.abs(normal).sum(axis=1)[:, None]
    sign = np.where(oct_ >= 0.0, 1.0, -1.0)
    oct_ = np.where((normal[:, 2:3] <= 0.0).repeat(2, 1), (1.0 - np.abs(oct_[:, ::-1])) * sign, oct_)
    factor = float(2 ** 15 - 1)
    summand = factor + 1.5
    return np.asarray(oct_ * factor + summand, dtype=np.uint16)


class Mesh:
    """Triangle soup in the *.vks layout."""

    def __init__(self, triangles, normals, uvs, material_index, light_flag):
        tri = np.asarray(triangles, dtype=np.float64).reshape(-1, 3, 3)
        self.triangle_count = tri.shape[0]
        flat = tri.reshape(-1, 3)
        box_min = flat.min(axis=0)
        box_max = flat.max(axis=0)
        ext = np.maximum(box_max - box_min, 1e-6)
        qfactor = 2.0 ** 21 / ext
        qoffset = -box_min * qfactor
        q = np.minimum(2 ** 21 - 1, np.asarray(flat * qfactor + qoffset, dtype=np.uint32))
        deq = 1.0 / qfactor
        self.dequantization_factor = deq.astype(np.float32)
        self.dequantization_summand = (box_min + 0.5 * deq).astype(np.float32)
        light = np.repeat(np.asarray(light_flag, dtype=np.uint32), 3)
        pos = np.zeros((flat.shape[0], 2), dtype=np.uint32)
        pos[:, 0] = q[:, 0] + ((q[:, 1] & 0x7FF) << 21)
        pos[:, 1] = ((q[:, 1] & 0x1FF800) >> 11) + (q[:, 2] << 10) + (light << 31)
        self.positions = pos
        nuv = np.zeros((flat.shape[0], 4), dtype=np.uint16)
        nuv[:, 0:2] = encode_normal_32_bit(np.asarray(normals, dtype=np.float64).reshape(-1, 3))
        uv = np.asarray(uvs, dtype=np.float64).reshape(-1, 2)
        nuv[:, 2:4] = np.asarray(np.clip(uv * ((2.0 ** 16 - 1.0) / 8.0) + 0.5, 0, 65535), dtype=np.uint16)
        self.normals_uv = nuv
        self.material_index = np.asarray(material_index, dtype=np.uint8)

    def dequantized(self):
        """Triangle positions as the shader decodes them (fma(q, f, s) in FP32)."""
        q0 = self.positions[:, 0]
        q1 = self.positions[:, 1]
        qx = q0 & 0x1FFFFF
        qy = ((q0 & 0xFFE00000) >> 21) | ((q1 & 0x3FF) << 11)
        qz = (q1 & 0x7FFFFC00) >> 10
        out = np.empty((q0.size, 3), dtype=np.float32)
        for a, qq in enumerate((qx, qy, qz)):
            # float64 product of two floats is exact; one rounding to float32 = fmaf
            out[:, a] = (qq.astype(np.float64) * np.float64(self.dequantization_factor[a])
                         + np.float64(self.dequantization_summand[a])).astype(np.float32)
        return out.reshape(-1, 9)


def _box_triangles(lo, hi):
    """12 outward-facing triangles of an axis-aligned box, with normals and uvs."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [  # (corner order counter-clockwise seen from outside, normal)
        ([(x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)], (0, 0, 1)),
        ([(x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0)], (0, 0, -1)),
        ([(x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)], (0, -1, 0)),
        ([(x1, y1, z0), (x0, y1, z0), (x0, y1, z1), (x1, y1, z1)], (0, 1, 0)),
        ([(x0, y1, z0), (x0, y0, z0), (x0, y0, z1), (x0, y1, z1)], (-1, 0, 0)),
        ([(x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1)], (1, 0, 0)),
    ]
    tris, nrm, uvs = [], [], []
    quv = [(0, 0), (1, 0), (1, 1), (0, 1)]
    for corners, n in faces:
        for a, b, c in ((0, 1, 2), (0, 2, 3)):
            tris.append([corners[a], corners[b], corners[c]])
            nrm.append([n, n, n])
            uvs.append([quv[a], quv[b], quv[c]])
    return tris, nrm, uvs


def _boxes_vectorized(lo, hi):
    """Vectorised _box_triangles for arrays lo, hi of shape (B, 3)."""
    proto_tris, proto_nrm, proto_uv = _box_triangles((0, 0, 0), (1, 1, 1))
    proto = np.asarray(proto_tris, dtype=np.float64)  # (12, 3, 3) in unit cube
    ext = (hi - lo)[:, None, None, :]
    tris = lo[:, None, None, :] + proto[None] * ext
    nrm = np.broadcast_to(np.asarray(proto_nrm, dtype=np.float64)[None], tris.shape)
    uv = np.broadcast_to(np.asarray(proto_uv, dtype=np.float64)[None], tris.shape[:3] + (2,))
    return tris.reshape(-1, 3, 3), nrm.reshape(-1, 3, 3), uv.reshape(-1, 3, 2)


# ---------------------------------------------------------------------------
# lights
# ---------------------------------------------------------------------------


def _rotation(angles):
    """Euler rotation of update_polygonal_light (src/polygonal_light.c:48-62)."""
    cx, sx = math.cos(angles[0]), math.sin(angles[0])
    cy, sy = math.cos(angles[1]), math.sin(angles[1])
    cz, sz = math.cos(angles[2]), math.sin(angles[2])
    return np.array([[cy * cz, -cy * sz, -sy],
                     [-sx * sy * cz + cx * sz, sx * sy * sz + cx * cz, -sx * cy],
                     [cx * sy * cz + sx * sz, -cx * sy * sz + sx * cz, cx * cy]])


class Light:
    """A polygonal light in plane space (polygonal_light_t)."""

    def __init__(self, plane_vertices, angles, translation, scaling, radiance):
        self.plane_vertices = np.asarray(plane_vertices, dtype=np.float32)  # (n, 2)
        self.angles = np.asarray(angles, dtype=np.float32)
        self.translation = np.asarray(translation, dtype=np.float32)
        self.scaling = np.asarray(scaling, dtype=np.float32)
        self.radiance = np.asarray(radiance, dtype=np.float32)
        self.vertex_count = len(plane_vertices)
        self._update()

    def _update(self):
        """World vertices and plane with the semantics of update_polygonal_light."""
        r = _rotation(self.angles.astype(np.float64))
        pv = self.plane_vertices.astype(np.float64) * self.scaling.astype(np.float64)
        world = self.translation.astype(np.float64) + pv @ r[:, :2].T
        self.world = world.astype(np.float32)
        plane = np.append(r[:, 2], -float(r[:, 2] @ self.translation.astype(np.float64)))
        p = self.plane_vertices.astype(np.float64)
        area = 0.0
        for t in range(self.vertex_count - 2):
            ax, bx = p[t + 2, 0] - p[0, 0], p[t + 1, 0] - p[0, 0]
            ay, by = p[t + 2, 1] - p[0, 1], p[t + 1, 1] - p[0, 1]
            area += 0.5 * (ax * by - bx * ay)
        area *= float(self.scaling[0]) * float(self.scaling[1])
        self.area = abs(area)
        self.plane = (plane if area > 0 else -plane).astype(np.float32)

    def to_ctypes(self):
        """A polygonal_light_t that owns its vertex arrays (kept on the object)."""
        L = abi.PolygonalLight()
        L.rotation_angles[:] = [float(a) for a in self.angles]
        L.translation[:] = [float(a) for a in self.translation]
        L.scaling_x, L.scaling_y = float(self.scaling[0]), float(self.scaling[1])
        L.radiant_flux[:] = [float(a) for a in self.radiance]
        L.vertex_count = self.vertex_count
        self._ps = (C.c_float * (4 * self.vertex_count))()
        self._ws = (C.c_float * (4 * self.vertex_count))()
        for i in range(self.vertex_count):
            self._ps[4 * i] = float(self.plane_vertices[i, 0])
            self._ps[4 * i + 1] = float(self.plane_vertices[i, 1])
        L.vertices_plane_space = C.cast(self._ps, C.POINTER(C.c_float))
        L.vertices_world_space = C.cast(self._ws, C.POINTER(C.c_float))
        return L


def pack_lights(lights, max_vertex_count=None):
    """Light records as write_lights packs them (src/main.c:456-490): header
    (radiance, plane, vertex count) + max_vertex_count vertices of 16 bytes;
    unused vertex slots repeat vertex 0."""
    vmax = max_vertex_count or max(l.vertex_count for l in lights)
    stride = abi.light_record_size(vmax)
    buf = np.zeros((len(lights), stride // 4), dtype=np.float32)
    u32 = buf.view(np.uint32)
    for i, l in enumerate(lights):
        buf[i, 0:3] = l.radiance
        buf[i, 4:8] = l.plane
        u32[i, 8] = l.vertex_count
        for k in range(vmax):
            src = l.world[k] if k < l.vertex_count else l.world[0]
            buf[i, 12 + 4 * k:12 + 4 * k + 3] = src
    return buf.tobytes(), vmax


def _light_fan(lights):
    tris, nrm = [], []
    for l in lights:
        n = l.plane[:3]
        for k in range(1, l.vertex_count - 1):
            tris.append([l.world[0], l.world[k], l.world[k + 1]])
            nrm.append([n, n, n])
    return tris, nrm


def random_convex_polygon(rng, n):
    """Vertices at sorted random angles on the unit circle (convex), with a
    minimum angular gap so that no three vertices are collinear."""
    while True:
        ang = np.sort(rng.uniform(0.0, 2.0 * math.pi, n))
        gaps = np.diff(np.append(ang, ang[0] + 2.0 * math.pi))
        if gaps.min() > 0.25 * (2.0 * math.pi / n) and gaps.max() < math.pi * 0.95:
            break
    return np.stack([np.cos(ang), np.sin(ang)], axis=1)


def polygon_area(pts):
    x, y = pts[:, 0], pts[:, 1]
    return 0.5 * abs(float(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1))))


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------


class Scene:
    """Everything the shading pass consumes, as numpy arrays and ctypes structs."""

    def __init__(self, name, mesh, materials, lights, camera, width, height, spp, candidates, seed):
        self.name = name
        self.mesh = mesh
        self.materials = np.asarray(materials, dtype=np.float32).reshape(-1, 6)
        self.lights = lights
        self.light_records, self.light_vmax = pack_lights(lights)
        self.camera = camera
        self.width, self.height = width, height
        self.spp, self.candidates = spp, candidates
        self.seed = seed

    @property
    def light_count(self):
        return len(self.lights)

    def settings(self, **kw):
        kw.setdefault("sample_count_light", self.spp)
        kw.setdefault("ris_candidate_count", self.candidates)
        return abi.default_settings(**kw)

    def materials_ctypes(self):
        arr = (abi.Material * len(self.materials))()
        for i, m in enumerate(self.materials):
            arr[i].base_color[:] = [float(v) for v in m[0:3]]
            arr[i].specular[:] = [float(v) for v in m[3:6]]
        return arr


def make_camera(position, rotation_z, rotation_x, fov=math.pi / 3.0):
    cam = abi.Camera()
    cam.position_world_space[:] = [float(v) for v in position]
    cam.rotation_z = rotation_z
    cam.rotation_x = rotation_x
    cam.vertical_fov = fov
    cam.near_plane = 0.05
    cam.far_plane = 1.0e3
    cam.speed = 2.0
    return cam


def _random_material(rng):
    rough = rng.uniform(0.05, 1.0)      # GGX alpha
    base = rng.uniform(0.0, 1.0, 3)
    # specular texel: (unused, linear roughness, metalicity); alpha = g^2
    return [base[0], base[1], base[2], 0.0, math.sqrt(rough), 0.0]


def small_scene(seed=0, width=256, height=256, light_count=64, spp=1, candidates=8):
    """20 m ground plane + 16 boxes (~200 triangles), 64 quad lights."""
    rng = np.random.default_rng(seed)
    tris, nrm, uvs, mats = [], [], [], []
    materials = [_random_material(rng)]
    g = 10.0
    ground = [[(-g, -g, 0), (g, -g, 0), (g, g, 0)], [(-g, -g, 0), (g, g, 0), (-g, g, 0)]]
    tris += ground
    nrm += [[(0, 0, 1)] * 3] * 2
    uvs += [[(0, 0), (1, 0), (1, 1)], [(0, 0), (1, 1), (0, 1)]]
    mats += [0, 0]
    for b in range(16):
        c = rng.uniform(-8.0, 8.0, 2)
        size = rng.uniform(0.5, 3.0, 2)
        h = rng.uniform(0.5, 4.0)
        t, n, u = _box_triangles((c[0] - size[0] / 2, c[1] - size[1] / 2, 0.0),
                                 (c[0] + size[0] / 2, c[1] + size[1] / 2, h))
        materials.append(_random_material(rng))
        tris += t
        nrm += n
        uvs += u
        mats += [len(materials) - 1] * len(t)
    lights = []
    square = np.array([[-0.5, -0.5], [0.5, -0.5], [0.5, 0.5], [-0.5, 0.5]])
    for i in range(light_count):
        sides = rng.uniform(0.1, 1.0, 2)
        centre = rng.uniform([-5.0, -5.0, 0.5], [5.0, 5.0, 10.5])
        angles = rng.uniform(0.0, 2.0 * math.pi, 3)
        radiance = rng.uniform(1.0, 10.0, 3)
        lights.append(Light(square, angles, centre, sides, radiance))
    return _finish("small", tris, nrm, uvs, mats, materials, lights,
                   make_camera((0.0, -16.0, 7.0), math.pi, math.pi / 2 - 0.35), width, height, spp,
                   candidates, seed)


def city_scene(seed=1, width=1920, height=1080, light_count=1024, grid=270, cell=4.0, spp=4,
               candidates=32):
    """Procedural city: grid x grid boxes (12 triangles each) on a ground plane,
    heights log-uniform in [1, 50] m; lights are convex 3-8-gons."""
    rng = np.random.default_rng(seed)
    half = grid * cell / 2.0
    materials = [_random_material(rng) for _ in range(64)]
    ii, jj = np.meshgrid(np.arange(grid), np.arange(grid), indexing="ij")
    cx = (ii.ravel() + 0.5) * cell - half
    cy = (jj.ravel() + 0.5) * cell - half
    nb = cx.size
    fx = rng.uniform(0.35, 0.85, nb) * cell
    fy = rng.uniform(0.35, 0.85, nb) * cell
    h = np.exp(rng.uniform(math.log(1.0), math.log(50.0), nb))
    lo = np.stack([cx - fx / 2, cy - fy / 2, np.zeros(nb)], axis=1)
    hi = np.stack([cx + fx / 2, cy + fy / 2, h], axis=1)
    btris, bnrm, buv = _boxes_vectorized(lo, hi)
    bmat = np.repeat(rng.integers(0, 64, nb), 12)
    # the ground is tessellated into one quad per cell: a single huge ground
    # triangle would inflate every LBVH box on its path to the whole city
    gx0 = (ii.ravel() * cell - half)[:, None]
    gy0 = (jj.ravel() * cell - half)[:, None]
    zero = np.zeros_like(gx0)
    c00 = np.concatenate([gx0, gy0, zero], 1)
    c10 = np.concatenate([gx0 + cell, gy0, zero], 1)
    c11 = np.concatenate([gx0 + cell, gy0 + cell, zero], 1)
    c01 = np.concatenate([gx0, gy0 + cell, zero], 1)
    ground = np.concatenate([np.stack([c00, c10, c11], 1), np.stack([c00, c11, c01], 1)])
    ng = ground.shape[0]
    guv = np.tile(np.array([[(0, 0), (1, 0), (1, 1)]], dtype=np.float64), (ng, 1, 1))
    tris = np.concatenate([ground, btris])
    nrm = np.concatenate([np.tile(np.array([0, 0, 1.0]), (ng, 3, 1)), bnrm])
    uvs = np.concatenate([guv, buv])
    mats = np.concatenate([np.zeros(ng, dtype=np.int64), bmat])
    lights = []
    for i in range(light_count):
        n = int(rng.integers(3, 9))
        pts = random_convex_polygon(rng, n)
        area = math.exp(rng.uniform(math.log(0.01), math.log(4.0)))
        s = math.sqrt(area / polygon_area(pts))
        centre = rng.uniform([-70.0, -30.0, 1.0], [70.0, 110.0, 25.0])
        angles = rng.uniform(0.0, 2.0 * math.pi, 3)
        radiance = np.exp(rng.uniform(math.log(0.5), math.log(50.0), 3))
        lights.append(Light(pts, angles, centre, (s, s), radiance))
    cam = make_camera((0.0, -60.0, 35.0), math.pi, math.pi / 2 - 0.45)
    return _finish("city", tris, nrm, uvs, mats, materials, lights, cam, width, height, spp,
                   candidates, seed)


def _finish(name, tris, nrm, uvs, mats, materials, lights, camera, width, height, spp, candidates,
            seed):
    ltris, lnrm = _light_fan(lights)
    tris = np.concatenate([np.asarray(tris, dtype=np.float64).reshape(-1, 3, 3),
                           np.asarray(ltris, dtype=np.float64).reshape(-1, 3, 3)])
    nrm = np.concatenate([np.asarray(nrm, dtype=np.float64).reshape(-1, 3, 3),
                          np.asarray(lnrm, dtype=np.float64).reshape(-1, 3, 3)])
    uvs = np.concatenate([np.asarray(uvs, dtype=np.float64).reshape(-1, 3, 2),
                          np.zeros((len(ltris), 3, 2))])
    mats = np.concatenate([np.asarray(mats), np.zeros(len(ltris), dtype=np.int64)])
    flag = np.concatenate([np.zeros(len(mats) - len(ltris), dtype=np.uint32),
                           np.ones(len(ltris), dtype=np.uint32)])
    mesh = Mesh(tris, nrm, uvs, mats, flag)
    return Scene(name, mesh, materials, lights, camera, width, height, spp, candidates, seed)


def batch_view_scenes(count, seed=7):
    """Scenes of the batched configuration, one per view: each its own
    procedural city of ~50k-500k triangles (grid 60-190) and its own light
    set of 16-8192 polygons (3-8 vertices, count log-uniform), resolution
    uniform in 320x240 .. 1920x1080, M from {8, 16, 32}, 2 spp. Yields one
    Scene per view (its width, height and settings are the view's)."""
    rng = np.random.default_rng(seed)
    for v in range(count):
        grid = int(rng.integers(60, 191))
        lights = int(round(math.exp(rng.uniform(math.log(16), math.log(8192)))))
        w, h = int(rng.integers(320, 1921)), int(rng.integers(240, 1081))
        m = int(rng.choice([8, 16, 32]))
        yield city_scene(seed=100 + v, width=w, height=h, light_count=lights, grid=grid, spp=2, candidates=m)


WORKLOADS = {
    "small": dict(fn=small_scene, desc="256x256, 16 boxes + ground (~200 tris), 64 quad lights, M=8, 1 spp"),
    "city": dict(fn=city_scene, desc="1920x1080, ~1M-triangle city, 1024 3-8-gon lights, M=32, 4 spp"),
}
