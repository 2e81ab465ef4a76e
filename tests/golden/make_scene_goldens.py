"""Writes tests/golden/reference_scenes.json: what the reference's 12 shipped scenes
(res/Scenes/*.xml, with res/Meshes) contain once loaded, as read by the oracle's independent
restatement of SceneLoader.cpp / ObjLoader.cpp (oracle/scene_xml.py). Run from the repository
root in a checkout next to the reference: python tests/golden/make_scene_goldens.py

The missing res/hdrtest.hdr (.MISSING_LARGE_BLOBS) is not needed: only the environment's path is
recorded. Mesh triangles are summarised by count and sums (the meshes are 12 and 968 triangles)."""
import json
import os
import shutil
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

REF = "/root/reference"


def summarise(d):
    return {
        "types": d["types"].tolist(),
        "params": np.asarray(d["params"], np.float32).tolist(),
        "first": np.asarray(d["first"]).tolist(),
        "count": np.asarray(d["count"]).tolist(),
        "mats": np.asarray(d["mats"], np.float32).tolist(),
        "lights": np.asarray(d["lights"], np.float32).tolist(),
        "volume": np.asarray(d["volume"], np.float32).tolist(),
        "n_tris": int(len(d["tris"])),
        "tris_sum": float(np.asarray(d["tris"], np.float64).sum()),
        "tris_abs_sum": float(np.abs(np.asarray(d["tris"], np.float64)).sum()),
        "env_path": d.get("env_path"),
    }


def main():
    scenes = sorted(f for f in os.listdir(os.path.join(REF, "res", "Scenes")) if f.endswith(".xml"))
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        shutil.copytree(os.path.join(REF, "res", "Scenes"), os.path.join(tmp, "res", "Scenes"))
        shutil.copytree(os.path.join(REF, "res", "Meshes"), os.path.join(tmp, "res", "Meshes"))
        cwd = os.getcwd()
        os.chdir(tmp)
        try:
            for name in scenes:
                d = oracle.load_scene_xml(f"res/Scenes/{name}", cwd=tmp)
                out[name] = summarise(d)
        finally:
            os.chdir(cwd)
    with open(os.path.join(ROOT, "tests", "golden", "reference_scenes.json"), "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print(f"wrote {len(out)} scenes")


if __name__ == "__main__":
    main()
