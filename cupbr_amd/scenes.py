"""Seeded scene generators producing the flat scene description.

The description is a dict of numpy arrays (the layout of ``cupbr_object`` /
``cupbr_light`` in include/cupbr.h, flattened):

    types  int32[n]       geometry type (GeometryType)
    params float32[n,12]  per-type parameters (see cupbr_object)
    first, count uint32[n] triangle range of a MESH object in ``tris``
    mats   float32[n,21]  albedo_e albedo_d albedo_s shininess roughness eta sigma_s sigma_a g interface type
    tris   float32[m,9]   mesh triangles (v1, v2, v3)
    lights float32[k,16]  type position intensity radiance half_extend1 half_extend2
    env    float32[h,w,3] or None
    volume float32[7]     sigma_s sigma_a g

Generators:
    cornell_box()       the 34-triangle Cornell box (5 walls + 2 rotated boxes)
    million_triangles() ~1.02M triangles: 11 displaced icospheres + 10,000 boxes, 4 area lights, env map
    mixed_primitives()  spheres, planes, quads, triangles and a mesh with every material type
"""
from __future__ import annotations

import math

import numpy as np

SPHERE, PLANE, QUAD, TRIANGLE, MESH = range(5)
LAMBERT, PHONG, MIRROR, GLASS, GGX, VOLUME = range(6)
POINT, AREA = range(2)


def material(kind=LAMBERT, albedo_e=(0, 0, 0), albedo_d=(1, 1, 1), albedo_s=(0, 0, 0), shininess=None,
             roughness=1.0, eta=1.5, sigma_s=(1, 1, 1), sigma_a=(1, 1, 1), g=0.0, interface=0) -> np.ndarray:
    m = np.zeros(21, np.float32)
    m[0:3], m[3:6], m[6:9] = albedo_e, albedo_d, albedo_s
    m[9] = np.float32(128.0) * np.float32(0.4) if shininess is None else shininess  # Material.h:39
    m[10], m[11] = max(np.float32(1e-4), np.float32(roughness)), eta
    m[12:15], m[15:18], m[18], m[19], m[20] = sigma_s, sigma_a, g, interface, kind
    return m


class SceneBuilder:
    def __init__(self):
        self.types, self.params, self.first, self.count, self.mats = [], [], [], [], []
        self.tris = []
        self.ntri = 0
        self.lights = []
        self.env = None
        self.volume = np.array([1, 1, 1, 1, 1, 1, 0], np.float32)

    def _add(self, t, p, mat, first=0, count=0):
        pp = np.zeros(12, np.float32)
        pp[:len(p)] = p
        self.types.append(t)
        self.params.append(pp)
        self.first.append(first)
        self.count.append(count)
        self.mats.append(mat)
        return len(self.types) - 1

    def sphere(self, c, r, mat):
        return self._add(SPHERE, list(c) + [r], mat)

    def plane(self, p, n, mat):
        return self._add(PLANE, list(p) + list(n), mat)

    def quad(self, p, n, e1, e2, mat):
        return self._add(QUAD, list(p) + list(n) + list(e1) + list(e2), mat)

    def triangle(self, a, b, c, mat):
        return self._add(TRIANGLE, list(a) + list(b) + list(c), mat)

    def mesh(self, tris: np.ndarray, mat):
        tris = np.asarray(tris, np.float32).reshape(-1, 9)
        v = tris.reshape(-1, 3)
        box = np.concatenate([v.min(0), v.max(0)])
        i = self._add(MESH, box, mat, self.ntri, len(tris))
        self.tris.append(tris)
        self.ntri += len(tris)
        return i

    def area_light(self, pos, radiance, e1, e2):
        self.lights.append(np.array([AREA, *pos, 0, 0, 0, *radiance, *e1, *e2], np.float32))

    def point_light(self, pos, intensity):
        self.lights.append(np.array([POINT, *pos, *intensity, 0, 0, 0, 0, 0, 0, 0, 0, 0], np.float32))

    def desc(self) -> dict:
        n = len(self.types)
        return {
            "types": np.array(self.types, np.int32), "params": np.array(self.params, np.float32).reshape(n, 12),
            "first": np.array(self.first, np.uint32), "count": np.array(self.count, np.uint32),
            "mats": np.array(self.mats, np.float32).reshape(n, 21),
            "tris": np.concatenate(self.tris).astype(np.float32) if self.tris else np.zeros((0, 9), np.float32),
            "lights": np.array(self.lights, np.float32).reshape(-1, 16), "env": self.env, "volume": self.volume,
        }


def box_triangles(center, size, yaw=0.0, rot=None) -> np.ndarray:
    """12 outward-facing triangles of a box (vertices v1, v2, v3 with normal (v2-v1)x(v3-v1) outward)."""
    sx, sy, sz = np.asarray(size, np.float64) / 2
    corners = np.array([[x, y, z] for x in (-sx, sx) for y in (-sy, sy) for z in (-sz, sz)])
    if rot is None:
        c, s = math.cos(yaw), math.sin(yaw)
        rot = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    corners = corners @ np.asarray(rot).T + np.asarray(center)
    # faces as corner index quads (a, b, c, d) wound counter-clockwise seen from outside
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    out = []
    for a, b, c_, d in faces:
        out.append(np.concatenate([corners[a], corners[b], corners[c_]]))
        out.append(np.concatenate([corners[a], corners[c_], corners[d]]))
    tris = np.array(out)
    # orient each triangle outward (normal away from the box center)
    for i, t in enumerate(tris):
        v = t.reshape(3, 3)
        nrm = np.cross(v[1] - v[0], v[2] - v[0])
        if np.dot(nrm, v.mean(0) - np.asarray(center)) < 0:
            tris[i] = np.concatenate([v[0], v[2], v[1]])
    return tris.astype(np.float32)


def cornell_box(width=256, height=256, as_triangles=True):
    """Config 0: Cornell box of 34 triangles, 0.25x0.25 ceiling light of radiance 15, 40 deg camera.

    Walls are two triangles each with normals facing the interior (back faces end a path in
    the reference). Albedos: red (0.65,0.05,0.05), green (0.12,0.45,0.15), grey 0.73.
    Returns (desc, camera_args) where camera_args feed Camera.look_at."""
    red = material(LAMBERT, albedo_d=(0.65, 0.05, 0.05))
    green = material(LAMBERT, albedo_d=(0.12, 0.45, 0.15))
    grey = material(LAMBERT, albedo_d=(0.73, 0.73, 0.73))
    b = SceneBuilder()
    P = np.array
    walls = [  # (corner, edge u, edge v, material): quads spanning the unit cube, normal = u x v inward
        (P([0, 0, 0]), P([0, 0, 1]), P([1, 0, 0]), grey),   # floor y=0, normal +y
        (P([0, 1, 0]), P([1, 0, 0]), P([0, 0, 1]), grey),   # ceiling y=1, normal -y
        (P([0, 0, 1]), P([0, 1, 0]), P([1, 0, 0]), grey),   # back z=1, normal -z
        (P([0, 0, 0]), P([0, 1, 0]), P([0, 0, 1]), red),    # left x=0, normal +x
        (P([1, 0, 0]), P([0, 0, 1]), P([0, 1, 0]), green),  # right x=1, normal -x
    ]
    for c, u, v, m in walls:
        a, bb, cc, d = c, c + u, c + u + v, c + v
        t1 = (a, bb, cc)
        t2 = (a, cc, d)
        for t in (t1, t2):
            nrm = np.cross(t[1] - t[0], t[2] - t[0])
            assert np.dot(nrm, P([0.5, 0.5, 0.5]) - t[0]) > 0
            if as_triangles:
                b.triangle(*t, m)
    tall = box_triangles((0.33, 0.30, 0.62), (0.33, 0.6, 0.33), yaw=math.radians(17.0))
    short = box_triangles((0.66, 0.15, 0.32), (0.3, 0.3, 0.3), yaw=math.radians(-17.0))
    for t in np.concatenate([tall, short]):
        v = t.reshape(3, 3)
        b.triangle(v[0], v[1], v[2], grey)
    b.area_light((0.5, 0.999, 0.5), (15, 15, 15), (0.125, 0, 0), (0, 0, 0.125))
    cam = dict(eye=(0.5, 0.5, -1.4), target=(0.5, 0.5, 0.5), up=(0, 1, 0), vfov=40.0, width=width, height=height)
    return b.desc(), cam


# ------------------------------------------------------------------ noise
def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint32) * np.uint32(73856093)) ^ (iy.astype(np.uint32) * np.uint32(19349663)) ^ \
        (iz.astype(np.uint32) * np.uint32(83492791)) ^ np.uint32(seed * 2654435761 % (1 << 32))
    h ^= h >> np.uint32(13)
    h *= np.uint32(0x5bd1e995)
    h ^= h >> np.uint32(15)
    return (h & np.uint32(0xFFFFFF)).astype(np.float64) / float(1 << 24)


def value_noise3(p: np.ndarray, seed: int) -> np.ndarray:
    """Trilinear value noise in [0,1) on the integer lattice (smoothstep weights)."""
    f = np.floor(p)
    t = p - f
    t = t * t * (3 - 2 * t)
    i = f.astype(np.int64)
    acc = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (t[..., 0] if dx else 1 - t[..., 0]) * (t[..., 1] if dy else 1 - t[..., 1]) * \
                    (t[..., 2] if dz else 1 - t[..., 2])
                acc = acc + w * _hash3(i[..., 0] + dx, i[..., 1] + dy, i[..., 2] + dz, seed)
    return acc


def fbm3(p: np.ndarray, seed: int, octaves=4) -> np.ndarray:
    """Fractal sum of value noise, normalised to [0,1)."""
    s, amp, norm_ = 0.0, 1.0, 0.0
    for o in range(octaves):
        s = s + amp * value_noise3(p * (2.0 ** o), seed + 101 * o)
        norm_ += amp
        amp *= 0.5
    return s / norm_


def icosphere(subdiv: int):
    """Unit icosphere: (vertices [V,3], faces [F,3]) with outward winding."""
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(subdiv):
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        edges.sort(axis=1)
        uniq, inv = np.unique(edges, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = base + inv[:nf], base + inv[nf:2 * nf], base + inv[2 * nf:]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    return v, f


def random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def env_value_noise(w=512, h=256, seed=99) -> np.ndarray:
    """Equirectangular FP32 environment map of smooth value noise in [0, 2]."""
    yy, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    theta = (yy + 0.5) / h * np.pi
    phi = (xx + 0.5) / w * 2 * np.pi - np.pi
    d = np.stack([np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)], -1)
    out = np.stack([fbm3(d * 3.0 + 10.0 * k, seed + k, 4) for k in range(3)], -1) * 2.0
    return out.astype(np.float32)


def million_triangles(width=1920, height=1080, seed=1234, n_spheres=11, subdiv=6, n_boxes=10000):
    """Config 1: 11 icospheres (subdivision 6, 81,920 triangles each, 4-octave fBm displacement of
    amplitude 0.1 r, r ~ U[2,8]) + 10,000 boxes (sizes U[0.2,2], uniform rotation, positions U[0,100]^3)
    = 1,021,120 triangles as 10,011 MESH objects. Materials per object: 50% Lambert (albedo U[0.1,0.9]),
    30% GGX (alpha U[0.05,0.5]), 20% glass (IOR 1.5). 4 area lights (radiance U[20,50], size U[2,5])
    above the volume and a 512x256 value-noise environment map (seed 99)."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()

    def random_material():
        u = rng.random()
        if u < 0.5:
            return material(LAMBERT, albedo_d=rng.uniform(0.1, 0.9, 3))
        if u < 0.8:
            return material(GGX, albedo_s=rng.uniform(0.3, 0.95, 3), roughness=rng.uniform(0.05, 0.5))
        return material(GLASS, albedo_s=(1, 1, 1), eta=1.5)

    v0, f0 = icosphere(subdiv)
    for k in range(n_spheres):
        r = rng.uniform(2.0, 8.0)
        c = rng.uniform(10.0, 90.0, 3)
        disp = 0.1 * r * (fbm3(v0 * 2.5 + rng.uniform(0, 100, 3), seed + k, 4) * 2.0 - 1.0)
        verts = c + v0 * (r + disp)[:, None]
        tris = verts[f0].reshape(-1, 9)
        b.mesh(tris, random_material())
    for k in range(n_boxes):
        size = rng.uniform(0.2, 2.0, 3)
        rot = random_rotation(rng)
        pos = rng.uniform(0.0, 100.0, 3)
        b.mesh(box_triangles(pos, size, rot=rot), random_material())
    for k in range(4):
        s = rng.uniform(2.0, 5.0)
        L = rng.uniform(20.0, 50.0)
        pos = (rng.uniform(20, 80), rng.uniform(95, 105), rng.uniform(20, 80))
        b.area_light(pos, (L, L, L), (s / 2, 0, 0), (0, 0, s / 2))
    b.env = env_value_noise(512, 256, 99)
    cam = dict(eye=(50.0, 50.0, -25.0), target=(50.0, 50.0, 50.0), up=(0, 1, 0), vfov=70.0, width=width,
               height=height)
    return b.desc(), cam


def mixed_primitives(width=128, height=128, env=True, seed=7):
    """Small scene with every geometry type and material type (spheres, planes, quads, triangles,
    a mesh; Lambert, Phong, mirror, glass, GGX), point and area lights, optional environment map."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    b.plane((0, -1, 0), (0, 1, 0), material(LAMBERT, albedo_d=(0.7, 0.7, 0.7)))
    b.plane((0, 0, 4), (0, 0, -1), material(PHONG, albedo_d=(0.3, 0.4, 0.5), albedo_s=(0.4, 0.4, 0.4)))
    b.quad((-1.2, 0, 2.5), (1, 0, 0), (0, 0, 1), (0, 1, 0), material(LAMBERT, albedo_d=(0.8, 0.1, 0.1)))
    b.quad((1.2, 0, 2.5), (-1, 0, 0), (0, 1, 0), (0, 0, 1), material(GGX, albedo_s=(0.9, 0.8, 0.3), roughness=0.2))
    b.sphere((-0.5, -0.6, 2.5), 0.35, material(MIRROR, albedo_s=(1, 1, 1)))
    b.sphere((0.5, -0.6, 2.2), 0.35, material(GLASS, albedo_s=(1, 1, 1), eta=1.5))
    b.sphere((0.0, 0.3, 3.0), 0.3, material(GGX, albedo_s=(0.8, 0.8, 0.8), roughness=0.05))
    b.triangle((-0.8, 0.8, 3.5), (0.8, 0.8, 3.5), (0.0, 0.2, 3.2), material(PHONG, albedo_d=(0.2, 0.6, 0.2), albedo_s=(0.5, 0.5, 0.5)))
    b.mesh(box_triangles((0.1, -0.75, 3.2), (0.5, 0.5, 0.5), yaw=0.4), material(LAMBERT, albedo_d=(0.2, 0.3, 0.9)))
    b.area_light((0, 0.95, 2.5), (8, 8, 8), (0.3, 0, 0), (0, 0, 0.3))
    b.point_light((0.6, 0.6, 1.5), (1.5, 1.5, 1.5))
    if env:
        b.env = env_value_noise(64, 32, seed)
    cam = dict(eye=(0, 0, 0), target=(0, -0.1, 1), up=(0, 1, 0), vfov=60.0, width=width, height=height)
    return b.desc(), cam


def random_triangle_soup(n=2000, seed=0, extent=10.0, size=1.0):
    """n random triangles as TRIANGLE objects (Lambert), for closest-hit parity."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    c = rng.uniform(-extent, extent, (n, 3))
    for i in range(n):
        v = c[i] + rng.normal(scale=size, size=(3, 3))
        b.triangle(v[0], v[1], v[2], material(LAMBERT, albedo_d=rng.uniform(0.1, 0.9, 3)))
    b.point_light((0, extent * 2, 0), (100, 100, 100))
    return b.desc()


def random_mesh_scene(n_meshes=50, tris_per_mesh=200, seed=0, extent=10.0):
    """Meshes of random triangles (finite object boxes, lazy box checks), plus a few spheres."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    for i in range(n_meshes):
        c = rng.uniform(-extent, extent, 3)
        v = c + rng.normal(scale=1.0, size=(tris_per_mesh, 3, 3))
        b.mesh(v.reshape(-1, 9), material(LAMBERT, albedo_d=rng.uniform(0.1, 0.9, 3)))
        if i % 10 == 0:
            b.sphere(rng.uniform(-extent, extent, 3), rng.uniform(0.5, 2.0), material(GGX, albedo_s=(0.5, 0.5, 0.5), roughness=0.3))
    b.point_light((0, extent * 2, 0), (100, 100, 100))
    return b.desc()


# ------------------------------------------------------------------ heterogeneous medium (adjoint workload)
def grid_medium_scene(res=64, width=256, height=256, seed=3, target_seed=4, target_sigma=0.05):
    """Unit-cube medium in an open room of 10 Lambertian triangles, lit by one area light.

    density: fBm (seed 3, 4 octaves at 4 cells per unit) normalised to [0, 1] on a res^3 grid;
    sigma_t = 10 * density, single-scattering albedo 0.9, HG g = 0.6, majorant 10.
    target: the same grid plus N(0, 0.05^2) noise (seed 4), clipped at 0 — the density an
    inverse-rendering test tries to recover. Surface albedos are the 10 walls' (the parameters
    of the per-surface gradient). Returns a dict with desc, camera args, medium, density,
    target_density and albedo ([n_surfaces, 3]: one row per scene object)."""
    b = SceneBuilder()
    P = np.array
    lo, hi = -0.6, 1.6
    walls = [  # (corner, u, v): normal u x v points into the room
        (P([lo, lo, lo]), P([0, 0, hi - lo]), P([hi - lo, 0, 0])),   # floor
        (P([lo, hi, lo]), P([hi - lo, 0, 0]), P([0, 0, hi - lo])),   # ceiling
        (P([lo, lo, hi]), P([0, hi - lo, 0]), P([hi - lo, 0, 0])),   # back
        (P([lo, lo, lo]), P([0, hi - lo, 0]), P([0, 0, hi - lo])),   # left
        (P([hi, lo, lo]), P([0, 0, hi - lo]), P([0, hi - lo, 0])),   # right
    ]
    base = [(0.75, 0.75, 0.75), (0.75, 0.75, 0.75), (0.6, 0.6, 0.6), (0.65, 0.12, 0.1), (0.12, 0.5, 0.15)]
    albedo = []
    for (c, u, v), a in zip(walls, base):
        m = material(LAMBERT, albedo_d=a)
        b.triangle(c, c + u, c + u + v, m)
        b.triangle(c, c + u + v, c + v, m)
        albedo += [a, a]
    b.area_light((0.5, hi - 0.15, 0.5), (12, 12, 12), (0.35, 0, 0), (0, 0, 0.35))
    g = (np.arange(res) + 0.5) / res * 4.0
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    d = fbm3(np.stack([x, y, z], -1), seed)
    d = (d - d.min()) / (d.max() - d.min())
    rng = np.random.default_rng(target_seed)
    target = np.maximum(d + rng.normal(0.0, target_sigma, d.shape), 0.0)
    medium = dict(bmin=(0.0, 0.0, 0.0), bmax=(1.0, 1.0, 1.0), sigma_t_scale=10.0, majorant=10.0,
                  scattering_albedo=0.9, g=0.6, min_collision_probability=1e-3)
    cam = dict(eye=(0.5, 0.5, -1.9), target=(0.5, 0.5, 0.5), up=(0, 1, 0), vfov=45.0, width=width, height=height)
    return dict(desc=b.desc(), camera=cam, medium=medium, density=d.astype(np.float32),
                target_density=target.astype(np.float32), albedo=np.array(albedo, np.float32))


def volume_spheres():
    """The volume renderer's test scene: a box of grey quads (green and red side walls), a
    scattering sphere with a glass interface (sigma_s 2, sigma_a 0.5, g 0.5), a coloured
    scattering sphere, a glass sphere and a 0.4 x 0.4 area light, in a thin scene medium
    (sigma_s = 0.05, sigma_a = 0.02). Seen by the reference's default camera (origin, +z)."""
    b = SceneBuilder()
    grey = material(LAMBERT, albedo_d=(0.7, 0.7, 0.7))
    b.quad((0, -1, 2), (0, 1, 0), (1, 0, 0), (0, 0, 1), grey)
    b.quad((0, 1, 2), (0, -1, 0), (1, 0, 0), (0, 0, 1), grey)
    b.quad((-1, 0, 2), (1, 0, 0), (0, 1, 0), (0, 0, 1), material(LAMBERT, albedo_d=(0.1, 0.7, 0.1)))
    b.quad((1, 0, 2), (-1, 0, 0), (0, 1, 0), (0, 0, 1), material(LAMBERT, albedo_d=(0.7, 0.1, 0.1)))
    b.quad((0, 0, 3), (0, 0, -1), (1, 0, 0), (0, 1, 0), grey)
    b.sphere((-0.4, -0.5, 2.2), 0.4, material(VOLUME, sigma_s=(2, 2, 2), sigma_a=(0.5, 0.5, 0.5), g=0.5,
                                               interface=1))
    b.sphere((0.45, -0.6, 2.5), 0.35, material(VOLUME, sigma_s=(1, 3, 1), sigma_a=(0.2, 0.2, 0.2)))
    b.sphere((0.3, 0.3, 2.6), 0.25, material(GLASS, albedo_s=(1, 1, 1)))
    b.area_light((0, 0.95, 2), (10, 10, 10), (0.2, 0, 0), (0, 0, 0.2))
    b.volume = np.array([0.05, 0.05, 0.05, 0.02, 0.02, 0.02, 0.0], np.float32)
    return b.desc()
