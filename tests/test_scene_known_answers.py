"""Known answers for the golden scene summaries (tests/golden/reference_scenes.json).

The golden file is what the oracle's restatement of SceneLoader.cpp reads from the reference's
shipped scenes, and the product loader is tested against it (tests/test_loader_abi.py). The values
below were transcribed by hand from the scene files' text (res/Scenes/*.xml, res/Meshes/*.obj) and
from the reference's declared defaults (lib/include/Geometry/Material.h:33-40,
lib/include/Geometry/Volume.h:17-20, enums Geometry.h:15-22, Material.h:13-21, Light.h:11-15), so a
mistake shared by both loaders (a dropped tag, a wrong default, a swapped field) fails here.

Summary layouts: params[12] SPHERE c,r | QUAD p,n,e1,e2; mats[21] albedo_e, albedo_d, albedo_s,
shininess, roughness, eta, sigma_s, sigma_a, g, interface, type; lights[16] type, position,
intensity, radiance, halfExtend1, halfExtend2; volume[7] sigma_s, sigma_a, g."""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_scenes.json")

SPHERE, PLANE, QUAD, TRIANGLE, MESH = range(5)
LAMBERT, PHONG, MIRROR, GLASS, GGX, VOLUME = range(6)
POINT, AREA = range(2)


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def material(albedo_e=(0, 0, 0), albedo_d=(1, 1, 1), albedo_s=(0, 0, 0), shininess=128.0 * 0.4, roughness=1.0,
             eta=1.5, sigma_s=(1, 1, 1), sigma_a=(1, 1, 1), g=0.0, interface=0, type=LAMBERT):
    """A material row with the reference's defaults for every tag the scene file leaves out."""
    return np.array([*albedo_e, *albedo_d, *albedo_s, shininess, roughness, eta, *sigma_s, *sigma_a, g, interface,
                     type], np.float32)


def quad(p, n, e1, e2):
    return np.array([*p, *n, *e1, *e2], np.float32)


def sphere(c, r):
    return np.array([*c, r] + [0] * 8, np.float32)


def test_cornell_box_sphere_volume(golden):
    """res/Scenes/CornellBoxSphereVolume.xml: 11 geometries (quads 1-5, spheres 6-9, quads 10-11),
    one point light, a scene medium."""
    d = golden["CornellBoxSphereVolume.xml"]
    assert d["types"] == [QUAD] * 5 + [SPHERE] * 4 + [QUAD] * 2
    params = np.array([
        quad((0, -1, 2), (0, 1, 0), (1, 0, 0), (0, 0, 1)),
        quad((-0.75, 1, 2), (0, -1, 0), (0.25, 0, 0), (0, 0, 1)),
        quad((-1, 0, 2), (1, 0, 0), (0, 0, 1), (0, 1, 0)),
        quad((1, 0, 2), (-1, 0, 0), (0, 1, 0), (0, 0, 1)),
        quad((0, 0, 3), (0, 0, -1), (1, 0, 0), (0, 1, 0)),
        sphere((-0.65, -0.75, 2.65), 0.25),
        sphere((0.65, -0.75, 2.65), 0.25),
        sphere((0, -0.75, 1.25), 0.25),
        sphere((0, 0.0, 2), 0.25),
        quad((0.75, 1, 2), (0, -1, 0), (0.25, 0, 0), (0, 0, 1)),
        quad((0, 1, 2), (0, -1, 0), (0.25, 0, 0), (0, 0, 1)),
    ])
    np.testing.assert_array_equal(np.array(d["params"], np.float32), params)
    mats = np.array([
        material(), material(), material(albedo_d=(0, 1, 0)), material(albedo_d=(1, 0, 0)), material(),
        material(albedo_d=(0, 0, 1), albedo_s=(1, 1, 1), type=PHONG),
        material(albedo_s=(1, 1, 1), type=MIRROR),
        material(albedo_s=(1, 1, 1), type=GLASS),
        material(albedo_d=(1, 1, 1)),
        material(), material(),
    ])
    np.testing.assert_array_equal(np.array(d["mats"], np.float32), mats)
    np.testing.assert_array_equal(np.array(d["lights"], np.float32),
                                  np.array([[POINT, 0, 2, 2, 150, 150, 150] + [0] * 9], np.float32))
    np.testing.assert_array_equal(np.array(d["volume"], np.float32),
                                  np.array([0.4, 0.4, 0.4, 1.0, 1.0, 1.0, 0.0], np.float32))
    assert d["n_tris"] == 0 and d["env_path"] is None


def test_glass_sphere_with_environment(golden):
    """res/Scenes/GlassSphere.xml: one unit glass sphere at the origin (eta 1.5), no lights, an
    environment map path."""
    d = golden["GlassSphere.xml"]
    assert d["types"] == [SPHERE]
    np.testing.assert_array_equal(np.array(d["params"], np.float32), sphere((0, 0, 0), 1.0)[None])
    np.testing.assert_array_equal(np.array(d["mats"], np.float32),
                                  material(albedo_s=(1, 1, 1), eta=1.5, type=GLASS)[None])
    assert len(d["lights"]) == 0
    assert d["env_path"] == "res/hdrtest.hdr"


def test_suzanne_mesh(golden):
    """res/Scenes/Suzanne.xml: one GGX mesh (albedo_s 1,0.8,0.9, roughness 0.1) from
    res/Meshes/suzanne_blender.obj, whose 968 'f' lines are triangles (511 'v' lines)."""
    d = golden["Suzanne.xml"]
    assert d["types"] == [MESH]
    assert d["n_tris"] == 968
    assert d["first"] == [0] and d["count"] == [968]
    np.testing.assert_array_equal(np.array(d["mats"], np.float32),
                                  material(albedo_s=(1, 0.8, 0.9), roughness=0.1, type=GGX)[None])
    lo, hi = np.array(d["params"][0][0:3]), np.array(d["params"][0][3:6])
    assert np.all(hi > lo)  # a box around the mesh, scaled by 10: Suzanne is about 2.7 x 2 x 1.7 units
    assert 15 < (hi - lo).max() < 40


def test_ggx_spheres_honour_scene_size(golden):
    """res/Scenes/GGX_Test.xml declares scene_size 10 and lists 11 GGX spheres of radius 0.9 at
    x = 0, 2, ..., 20 with roughness 0.001, 0.1, ..., 1: the loader reads the first 10."""
    d = golden["GGX_Test.xml"]
    assert d["types"] == [SPHERE] * 10
    centres = np.array(d["params"], np.float32)[:, 0:4]
    np.testing.assert_array_equal(centres, np.array([[2 * i, 0, 0, 0.9] for i in range(10)], np.float32))
    rough = np.array(d["mats"], np.float32)[:, 10]
    np.testing.assert_array_equal(rough[:2], np.float32([0.001, 0.1]))
    np.testing.assert_array_equal(rough[8:], np.float32([0.8, 0.9]))
    assert all(int(m[20]) == GGX for m in d["mats"])
