"""ORACLE — test infrastructure only.

An independent restatement of cupbr's scene loading, used to check the
product's C++ loader and to feed the CPU oracle:

* ``load_scene_xml`` follows ``SceneLoader::loadFromFile`` and
  ``detail::loadMaterial`` (lib/src/Scene/SceneLoader.cpp:20-454);
* ``load_obj`` follows ``ObjLoader::loadObj`` (lib/src/Scene/ObjLoader.cpp:13-87).

The result is the flat scene description shared by the oracle and the
product's Python front end: a dict of numpy arrays

    types  int32[n]        geometry type (SPHERE=0 PLANE=1 QUAD=2 TRIANGLE=3 MESH=4)
    params float32[n,12]   SPHERE c,r | PLANE p,n | QUAD p,n,e1,e2 | TRIANGLE v1,v2,v3 | MESH bmin,bmax
    first, count uint32[n] mesh triangle range in ``tris``
    mats   float32[n,21]   albedo_e albedo_d albedo_s shininess roughness eta sigma_s sigma_a g interface type
    tris   float32[m,9]    mesh triangles v1 v2 v3 (after scale and translation)
    lights float32[k,16]   type position intensity radiance halfExtend1 halfExtend2
    env    float32[h,w,3] or None
    volume float32[7]      sigma_s sigma_a g
"""
from __future__ import annotations

import os
import xml.etree.ElementTree as ET

import numpy as np

GEOM = {"SPHERE": 0, "PLANE": 1, "QUAD": 2, "TRIANGLE": 3, "MESH": 4}
MAT = {"LAMBERT": 0, "PHONG": 1, "MIRROR": 2, "GLASS": 3, "GGX": 4, "VOLUME": 5}
LIGHT = {"POINT": 0, "AREA": 1}


def default_material() -> np.ndarray:
    """Material.h:35-43 and Volume.h:15-21 defaults."""
    m = np.zeros(21, np.float32)
    m[3:6] = 1.0                      # albedo_d
    m[9] = np.float32(128.0) * np.float32(0.4)  # shininess = 128.0f * 0.4f
    m[10] = 1.0                       # roughness
    m[11] = 1.5                       # eta
    m[12:15] = 1.0                    # sigma_s
    m[15:18] = 1.0                    # sigma_a
    return m


def _vec(text: str) -> np.ndarray:
    # string2vector, SceneLoader.cpp:20-33 (comma separated std::stof)
    parts = text.split(",")
    return np.array([np.float32(p.strip()) for p in parts[:3]], np.float32)


def _text(el, name):
    c = el.find(name)
    return None if c is None else (c.text or "")


def load_material(el) -> np.ndarray:
    """detail::loadMaterial, SceneLoader.cpp:35-137."""
    m = default_material()
    t = _text(el, "type").strip()
    if t not in MAT:
        raise ValueError(f"Error while loading material: {t} is not a valid material type")
    m[20] = MAT[t]
    for name, sl in (("albedo_e", slice(0, 3)), ("albedo_d", slice(3, 6)), ("albedo_s", slice(6, 9)),
                     ("sigma_s", slice(12, 15)), ("sigma_a", slice(15, 18))):
        v = _text(el, name)
        if v is not None:
            m[sl] = _vec(v)
    v = _text(el, "shininess")
    if v is not None:
        m[9] = np.float32(v.strip())
    v = _text(el, "roughness")
    if v is not None:
        m[10] = max(np.float32(1e-4), np.float32(v.strip()))  # :103-106
    v = _text(el, "eta")
    if v is not None:
        m[11] = np.float32(v.strip())
    v = _text(el, "g")
    if v is not None:
        m[18] = np.float32(v.strip())
    v = _text(el, "interface")
    if v is not None and v.strip() == "GLASS":
        m[19] = 1
    return m


def load_obj(path: str, position: np.ndarray, scale: np.ndarray):
    """ObjLoader::loadObj, ObjLoader.cpp:13-87: 'v x y z' and 'f a/b/c d/e/f g/h/i' (slots 0,3,6)."""
    verts = []
    faces = []
    mn = np.full(3, np.inf, np.float32)
    mx = np.full(3, -np.inf, np.float32)
    with open(path, "r") as f:
        for line in f:
            line = line.rstrip("\n").rstrip("\r")
            if line.startswith("v "):
                r = [np.float32(x) for x in line[2:].split(" ") if x != ""]
                v = (np.array([scale[0] * r[0], scale[1] * r[1], scale[2] * r[2]], np.float32) + position).astype(np.float32)
                verts.append(v)
                mx = np.maximum(mx, v)
                mn = np.minimum(mn, v)
            elif line.startswith("f "):
                items = line[2:].replace(" ", "/").split("/")
                idx = [int(x) for x in items if x != ""]
                faces.append((idx[0] - 1, idx[3] - 1, idx[6] - 1))
    verts = np.array(verts, np.float32).reshape(-1, 3)
    tris = np.stack([np.concatenate([verts[a], verts[b], verts[c]]) for a, b, c in faces]).astype(np.float32) \
        if faces else np.zeros((0, 9), np.float32)
    return tris, mn, mx


def load_scene_xml(path: str, cwd: str | None = None) -> dict:
    """SceneLoader::loadFromFile, SceneLoader.cpp:140-454. Mesh paths resolve against ``cwd``."""
    with open(path, "r") as f:
        text = f.read()
    root = ET.fromstring("<cupbr_root>" + text + "</cupbr_root>")
    header = root.find("header")
    n = int(_text(header, "scene_size").strip())
    nl = int(_text(header, "light_count").strip())
    types = np.zeros(n, np.int32)
    params = np.zeros((n, 12), np.float32)
    first = np.zeros(n, np.uint32)
    count = np.zeros(n, np.uint32)
    mats = np.zeros((n, 21), np.float32)
    tris = []
    ntri = 0
    geo = root.find("geometry")
    for i in range(n):
        g = geo.find(f"geometry{i + 1}")
        t = _text(g, "type").strip()
        if t not in GEOM:
            raise ValueError(f"Error while loading scene: {t} is not a valid geometry type!")
        types[i] = GEOM[t]
        if t == "PLANE":
            params[i, 0:3] = _vec(_text(g, "position"))
            params[i, 3:6] = _vec(_text(g, "normal"))
        elif t == "SPHERE":
            params[i, 0:3] = _vec(_text(g, "position"))
            params[i, 3] = np.float32(_text(g, "radius").strip())
        elif t == "QUAD":
            params[i, 0:3] = _vec(_text(g, "position"))
            params[i, 3:6] = _vec(_text(g, "normal"))
            params[i, 6:9] = _vec(_text(g, "extend1"))
            params[i, 9:12] = _vec(_text(g, "extend2"))
        elif t == "TRIANGLE":
            params[i, 0:3] = _vec(_text(g, "vertex1"))
            params[i, 3:6] = _vec(_text(g, "vertex2"))
            params[i, 6:9] = _vec(_text(g, "vertex3"))
        else:
            mpath = _text(g, "path").strip()
            pos = np.zeros(3, np.float32)
            sc = np.ones(3, np.float32)
            if g.find("position") is not None:
                pos = _vec(_text(g, "position"))
            if g.find("scale") is not None:
                sc = _vec(_text(g, "scale"))
            full = mpath if os.path.isabs(mpath) else os.path.join(cwd or os.getcwd(), mpath)
            mt, mn, mx = load_obj(full, pos, sc)
            first[i] = ntri
            count[i] = len(mt)
            ntri += len(mt)
            tris.append(mt)
            params[i, 0:3] = mn
            params[i, 3:6] = mx
        mats[i] = load_material(g.find("material"))
    lights = np.zeros((nl, 16), np.float32)
    lh = root.find("lights")
    for i in range(nl):
        el = lh.find(f"light{i + 1}")
        t = _text(el, "type").strip()
        if t not in LIGHT:
            raise ValueError(f"Error while loading scene: {t} is not a valid light type!")
        lights[i, 0] = LIGHT[t]
        for name, sl in (("position", slice(1, 4)), ("intensity", slice(4, 7)), ("radiance", slice(7, 10)),
                         ("extend1", slice(10, 13)), ("extend2", slice(13, 16))):
            v = _text(el, name)
            if v is not None:
                lights[i, sl] = _vec(v)
    volume = np.array([1, 1, 1, 1, 1, 1, 0], np.float32)
    vh = root.find("volume")
    if vh is not None:
        volume[0:3] = _vec(_text(vh, "sigma_s"))
        volume[3:6] = _vec(_text(vh, "sigma_a"))
        volume[6] = np.float32(_text(vh, "g").strip())
    env_path = None
    eh = root.find("environment")
    if eh is not None:
        env_path = _text(eh, "path").strip()
    return {
        "types": types, "params": params, "first": first, "count": count, "mats": mats,
        "tris": np.concatenate(tris).astype(np.float32) if tris else np.zeros((0, 9), np.float32),
        "lights": lights, "env": None, "env_path": env_path, "volume": volume,
    }
