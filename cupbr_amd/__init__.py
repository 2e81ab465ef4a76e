"""cupbr_amd — MI355X renderer core behind cupbr's interface.

Python front end over the C ABI in ``include/cupbr.h`` (``lib/libcupbr.so``,
built by ``python -m cupbr_amd.build``). Class and method names follow the
reference's C++ interface:

    SceneLoader.load_from_file / destroy_scene  lib/include/Scene/SceneLoader.h:15-21
    PBRenderer                                   lib/include/Renderer/PBRenderer.h:27-116
    PostProcessor                                lib/include/PostProcessing/PostProcessor.h:12-52
    ToneMapper                                   lib/include/Renderer/ToneMapper.h:22-80
    Camera                                       lib/include/DataStructure/Camera.h:14-86

Every call goes to the HIP library; there is no CPU fallback. A missing
library or a missing GPU raises ``CupbrError``.
"""
from __future__ import annotations

import ctypes as C
import enum
import os

import numpy as np

__all__ = [
    "CupbrError", "RenderingMethod", "ToneMappingType", "GeometryType", "MaterialType", "LightType", "Image",
    "Camera", "Scene", "SceneLoader", "PBRenderer", "PostProcessor", "ToneMapper", "threshold_curve", "lib",
    "device_name", "OBJECT_DTYPE", "LIGHT_DTYPE",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
# CUPBR_LIBRARY names another build of the library (same ABI) to load instead of the in-tree one
LIB_PATH = os.environ.get("CUPBR_LIBRARY") or os.path.join(_HERE, "lib", "libcupbr.so")


class CupbrError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[cupbr {code}] {message}")
        self.code = code


GDPT_TRIMMED_MEAN, GDPT_SCREENED_POISSON = 0, 1
RNG_TEA_LCG, RNG_PCG = 0, 1  # cupbr_renderer_set_rng; PCG is not in the reference


class RenderingMethod(enum.IntEnum):  # PBRenderer.h:14-22
    RAYTRACER = 0
    WHITTED = 1
    PATHTRACER = 2
    METROPOLIS = 3
    GRADIENTDOMAIN = 4
    VOLUME = 5


class ToneMappingType(enum.IntEnum):  # ToneMapper.h:13-17
    REINHARD = 0
    GAMMA = 1
    ACES = 2  # not in the reference: Narkowicz's ACES filmic fit, then gamma 1/2.2


class BloomPyramid(enum.IntEnum):  # cupbr_bloom_pyramid; GAUSSIAN5 is not in the reference
    REFERENCE = 0
    GAUSSIAN5 = 1


class GeometryType(enum.IntEnum):  # Geometry.h:15-23
    SPHERE = 0
    PLANE = 1
    QUAD = 2
    TRIANGLE = 3
    MESH = 4


class MaterialType(enum.IntEnum):  # Material.h:14-22
    LAMBERT = 0
    PHONG = 1
    MIRROR = 2
    GLASS = 3
    GGX = 4
    VOLUME = 5


class LightType(enum.IntEnum):  # Light.h:11-15
    POINT = 0
    AREA = 1


class Image(enum.IntEnum):  # cupbr_image
    OUTPUT = 0
    GRADIENT_X = 1
    GRADIENT_Y = 2
    BASE = 3
    GRADIENT_X_FORWARD = 4
    GRADIENT_X_BACKWARD = 5
    GRADIENT_Y_FORWARD = 6
    GRADIENT_Y_BACKWARD = 7


# ------------------------------------------------------------------ C structs
class _Volume(C.Structure):
    _fields_ = [("sigma_s", C.c_float * 3), ("sigma_a", C.c_float * 3), ("g", C.c_float), ("interface_type", C.c_int32)]


class _Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("x_axis", C.c_float * 3), ("y_axis", C.c_float * 3),
                ("z_axis", C.c_float * 3), ("moved", C.c_int32)]


class _CameraController(C.Structure):
    _fields_ = [("aspect_ratio", C.c_float), ("pitch", C.c_float), ("yaw", C.c_float), ("first_mouse", C.c_int32),
                ("last_x", C.c_float), ("last_y", C.c_float)]


class _InputState(C.Structure):
    _fields_ = [("keys", C.c_uint32), ("cursor_x", C.c_double), ("cursor_y", C.c_double)]


KEY_W, KEY_A, KEY_S, KEY_D, KEY_LEFT_SHIFT, KEY_LEFT_CONTROL = 1, 2, 4, 8, 16, 32


class _BvhStats(C.Structure):
    _fields_ = [("nodes", C.c_uint32), ("leaves", C.c_uint32), ("primitives", C.c_uint32), ("unbounded", C.c_uint32),
                ("depth", C.c_uint32), ("stack_entries", C.c_uint32), ("build_ms", C.c_double),
                ("upload_ms", C.c_double), ("sah_cost", C.c_double)]


class _RenderStats(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("milliseconds", C.c_double), ("stage_ms", C.c_double * 4), ("stage_launches", C.c_uint32 * 4),
                ("extension_node_visits", C.c_uint64), ("extension_prim_tests", C.c_uint64),
                ("shadow_node_visits", C.c_uint64), ("shadow_prim_tests", C.c_uint64),
                ("cache_queries", C.c_uint64), ("cache_training_rows", C.c_uint64),
                ("refill_launches", C.c_uint64)]


class _GridMedium(C.Structure):
    _fields_ = [("bmin", C.c_float * 3), ("bmax", C.c_float * 3), ("nx", C.c_uint32), ("ny", C.c_uint32),
                ("nz", C.c_uint32), ("sigma_t_scale", C.c_float), ("majorant", C.c_float),
                ("scattering_albedo", C.c_float), ("g", C.c_float), ("min_collision_probability", C.c_float)]


# cupbr_object / cupbr_light as numpy records (144 and 64 bytes, no padding)
OBJECT_DTYPE = np.dtype([
    ("type", "<i4"), ("params", "<f4", 12), ("first_triangle", "<u4"), ("triangle_count", "<u4"),
    ("albedo_e", "<f4", 3), ("albedo_d", "<f4", 3), ("albedo_s", "<f4", 3), ("shininess", "<f4"),
    ("roughness", "<f4"), ("eta", "<f4"), ("sigma_s", "<f4", 3), ("sigma_a", "<f4", 3), ("g", "<f4"),
    ("interface_type", "<i4"), ("material_type", "<i4")])
# cupbr_material (84 bytes, Material.h:35-43)
MATERIAL_DTYPE = np.dtype([
    ("albedo_e", "<f4", 3), ("albedo_d", "<f4", 3), ("albedo_s", "<f4", 3), ("shininess", "<f4"),
    ("roughness", "<f4"), ("eta", "<f4"), ("sigma_s", "<f4", 3), ("sigma_a", "<f4", 3), ("g", "<f4"),
    ("interface_type", "<i4"), ("material_type", "<i4")])
LIGHT_DTYPE = np.dtype([("type", "<i4"), ("position", "<f4", 3), ("intensity", "<f4", 3), ("radiance", "<f4", 3),
                        ("half_extend1", "<f4", 3), ("half_extend2", "<f4", 3)])
assert OBJECT_DTYPE.itemsize == 144 and LIGHT_DTYPE.itemsize == 64

_lib = None


def lib():
    """The loaded libcupbr.so (raises CupbrError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CupbrError(-100, f"{LIB_PATH} is missing: build it with `python -m cupbr_amd.build`")
        # PyTorch ships its own HIP runtime under the same soname (libamdhip64.so.7). Loading torch
        # first lets libcupbr bind to that copy, so a process that uses both has one HIP runtime
        # (two runtimes in one process cannot share the GPU).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        P, U32, I32, SZ = C.c_void_p, C.c_uint32, C.c_int32, C.c_size_t
        sig = {
            "cupbr_last_error": (C.c_char_p, []),
            "cupbr_device_name": (C.c_int, [C.c_char_p, SZ]),
            "cupbr_measure_valu_peak": (C.c_int, [U32, P, P]),
            "cupbr_measure_atomic_peak": (C.c_int, [U32, U32, P, P]),
            "cupbr_camera_init": (C.c_int, [P, U32, U32]),
            "cupbr_camera_look_at": (C.c_int, [P, P, P, P, C.c_float, U32, U32]),
            "cupbr_camera_controller_init": (C.c_int, [P, U32, U32]),
            "cupbr_camera_process_input": (C.c_int, [P, P, P, C.c_float]),
            "cupbr_scene_load_xml": (C.c_int, [C.c_char_p, P]),
            "cupbr_scene_create": (C.c_int, [P, U32, P, U32, P, U32, P, U32, U32, P, P]),
            "cupbr_scene_destroy": (C.c_int, [P]),
            "cupbr_scene_upload": (C.c_int, [P]),
            "cupbr_scene_get_counts": (C.c_int, [P, P, P, P, P]),
            "cupbr_scene_get_objects": (C.c_int, [P, P, U32]),
            "cupbr_scene_get_triangles": (C.c_int, [P, P, U32]),
            "cupbr_scene_get_lights": (C.c_int, [P, P, U32]),
            "cupbr_scene_get_volume": (C.c_int, [P, P]),
            "cupbr_scene_get_bvh_stats": (C.c_int, [P, P]),
            "cupbr_scene_get_bvh_nodes": (C.c_int, [P, P, P, C.c_uint32]),
            "cupbr_scene_get_environment": (C.c_int, [P, P, SZ, P, P]),
            "cupbr_scene_get_reference_order": (C.c_int, [P, P, U32]),
            "cupbr_trace_closest": (C.c_int, [P, P, SZ, P, P, P]),
            "cupbr_trace_closest_device": (C.c_int, [P, P, SZ, P, P, P, P]),
            "cupbr_trace_closest_device_noalloc": (C.c_int, [P, P, SZ, P, P, P, P]),
            "cupbr_trace_visibility": (C.c_int, [P, P, P, SZ, P]),
            "cupbr_trace_visibility_device": (C.c_int, [P, P, P, SZ, P, P]),
            "cupbr_renderer_create": (C.c_int, [I32, P]),
            "cupbr_renderer_destroy": (C.c_int, [P]),
            "cupbr_renderer_set_output_size": (C.c_int, [P, U32, U32]),
            "cupbr_renderer_register_scene": (C.c_int, [P, P]),
            "cupbr_renderer_set_method": (C.c_int, [P, I32]),
            "cupbr_renderer_get_method": (C.c_int, [P, P]),
            "cupbr_renderer_reset": (C.c_int, [P]),
            "cupbr_renderer_set_russian_roulette": (C.c_int, [P, I32]),
            "cupbr_renderer_set_max_trace_depth": (C.c_int, [P, U32]),
            "cupbr_renderer_get_max_trace_depth": (C.c_int, [P, P]),
            "cupbr_renderer_get_frame_index": (C.c_int, [P, P]),
            "cupbr_renderer_get_output_size": (C.c_int, [P, P, P]),
            "cupbr_render": (C.c_int, [P, P, U32]),
            "cupbr_renderer_get_image": (C.c_int, [P, I32, P, SZ]),
            "cupbr_renderer_get_image_device": (C.c_int, [P, I32, P]),
            "cupbr_renderer_get_stats": (C.c_int, [P, P]),
            "cupbr_renderer_synchronize": (C.c_int, [P]),
            "cupbr_renderer_set_profiling": (C.c_int, [P, I32]),
            "cupbr_renderer_set_gdpt_reconstruction": (C.c_int, [P, I32, C.c_float, U32]),
            "cupbr_renderer_set_traversal_counting": (C.c_int, [P, I32]),
            "cupbr_post_create": (C.c_int, [P]),
            "cupbr_post_destroy": (C.c_int, [P]),
            "cupbr_post_register_image": (C.c_int, [P, P, U32, U32]),
            "cupbr_post_bloom": (C.c_int, [P, P]),
            "cupbr_post_set_bloom_pyramid": (C.c_int, [P, C.c_int32]),
            "cupbr_post_get_buffer_device": (C.c_int, [P, P]),
            "cupbr_post_get_buffer": (C.c_int, [P, P, SZ]),
            "cupbr_threshold_curve": (C.c_int, [C.c_float, C.c_float, P]),
            "cupbr_tonemapper_create": (C.c_int, [I32, P]),
            "cupbr_tonemapper_destroy": (C.c_int, [P]),
            "cupbr_tonemapper_register_image": (C.c_int, [P, P, U32, U32]),
            "cupbr_tonemapper_tonemap": (C.c_int, [P]),
            "cupbr_tonemapper_tonemap_bloom": (C.c_int, [P, P, P]),
            "cupbr_tonemapper_get_render_buffer": (C.c_int, [P, P, SZ]),
            "cupbr_tonemapper_get_render_buffer_device": (C.c_int, [P, P]),
            "cupbr_tonemapper_set_type": (C.c_int, [P, I32]),
            "cupbr_tonemapper_get_type": (C.c_int, [P, P]),
            "cupbr_tonemapper_set_exposure": (C.c_int, [P, C.c_float]),
            "cupbr_tonemapper_save_to_file": (C.c_int, [P, C.c_char_p]),
            "cupbr_grid_surface_count": (C.c_int, [P, P]),
            "cupbr_nrc_create": (C.c_int, [P, P, U32, P]),
            "cupbr_nrc_destroy": (C.c_int, [P]),
            "cupbr_nrc_get_weights": (C.c_int, [P, P]),
            "cupbr_nrc_set_weights": (C.c_int, [P, P]),
            "cupbr_nrc_infer_device": (C.c_int, [P, P, U32, P, P]),
            "cupbr_nrc_train_device": (C.c_int, [P, P, U32, U32, U32, C.c_float, P, P]),
            "cupbr_nrc_matrix_core_product": (C.c_int, [P, P, P]),
            "cupbr_renderer_set_radiance_cache": (C.c_int, [P, P, U32, U32, U32, C.c_float]),
            "cupbr_renderer_set_radiance_cache_camera_hit": (C.c_int, [P, C.c_float]),
            "cupbr_renderer_set_rng": (C.c_int, [P, C.c_int32]),
            "cupbr_renderer_save_pfm": (C.c_int, [P, I32, C.c_char_p]),
            "cupbr_pick": (C.c_int, [P, P, U32, U32, U32, U32, P, P]),
            "cupbr_scene_update_material": (C.c_int, [P, I32, P]),
            "cupbr_renderer_resume": (C.c_int, [P, P, SZ, U32]),
            "cupbr_grid_set_counting": (C.c_int, [P, I32]),
            "cupbr_grid_get_counts": (C.c_int, [P, P]),
            "cupbr_grid_render_device": (C.c_int, [P, P, P, P, P, P, U32, U32, U32, U32, U32, P, P]),
            "cupbr_grid_adjoint_device": (C.c_int, [P, P, P, P, P, P, U32, U32, U32, U32, U32, P, P, P, P]),
            "cupbr_grid_adjoint_from_forward_device": (C.c_int, [P, P, P, P, P, P, U32, U32, U32, U32, U32, P, P, P, P]),
            "cupbr_grid_render": (C.c_int, [P, P, P, P, P, P, U32, U32, U32, U32, U32, P]),
            "cupbr_grid_adjoint": (C.c_int, [P, P, P, P, P, P, U32, U32, U32, U32, U32, P, P, P]),
        }
        for name, (res, args) in sig.items():
            if os.environ.get("CUPBR_LIBRARY") and not hasattr(L, name):
                continue  # another build, from before this entry point: calling it fails by name
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise CupbrError(rc, lib().cupbr_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def device_name() -> str:
    buf = C.create_string_buffer(256)
    _check(lib().cupbr_device_name(buf, 256))
    return buf.value.decode()


def measure_valu_peak(iterations: int = 20000) -> dict:
    """cupbr_measure_valu_peak: time and wave64 FMA instruction count of a VALU-only kernel with
    every lane active (the device's VALU issue ceiling)."""
    ms = C.c_double()
    n = C.c_uint64()
    _check(lib().cupbr_measure_valu_peak(iterations, C.byref(ms), C.byref(n)))
    return {"ms": ms.value, "wave_instructions": n.value,
            "wave_instructions_per_s": n.value / (ms.value * 1e-3) if ms.value > 0 else float("nan")}


def measure_atomic_peak(iterations: int = 256, words_per_copy: int = 64 ** 3) -> dict:
    """cupbr_measure_atomic_peak: time and count of FP32 atomic adds (no return) to pseudo-random
    words of 8 per-XCD copies of words_per_copy floats (the grid adjoint's gradient layout)."""
    ms = C.c_double()
    n = C.c_uint64()
    _check(lib().cupbr_measure_atomic_peak(iterations, words_per_copy, C.byref(ms), C.byref(n)))
    return {"ms": ms.value, "atomics": n.value,
            "atomics_per_s": n.value / (ms.value * 1e-3) if ms.value > 0 else float("nan")}


def threshold_curve(threshold: float = 1.0, knee: float = 1.0) -> np.ndarray:
    """Interactor::Impl::compute_threshold (lib/src/Interaction/Interactor.cpp:436-440)."""
    out = (C.c_float * 4)()
    _check(lib().cupbr_threshold_curve(threshold, knee, out))
    return np.array(out[:], np.float32)


# ------------------------------------------------------------------ Camera
class Camera:
    """Pinhole camera: position plus image-plane axes (Camera.h:72-76)."""

    def __init__(self, width: int | None = None, height: int | None = None):
        self._c = _Camera()
        if width is not None:
            _check(lib().cupbr_camera_init(C.byref(self._c), width, height))
        else:  # Camera() = default (Camera.h:72-76)
            self._c.x_axis[:] = [1, 0, 0]
            self._c.y_axis[:] = [0, 1, 0]
            self._c.z_axis[:] = [0, 0, 1]

    @classmethod
    def look_at(cls, eye, target, up, vfov_degrees: float, width: int, height: int) -> "Camera":
        cam = cls()
        f = lambda v: (C.c_float * 3)(*[float(x) for x in v])  # noqa: E731
        _check(lib().cupbr_camera_look_at(C.byref(cam._c), f(eye), f(target), f(up), vfov_degrees, width, height))
        return cam

    def _get(self, k):
        return np.array(getattr(self._c, k)[:], np.float32)

    position = property(lambda s: s._get("position"))
    x_axis = property(lambda s: s._get("x_axis"))
    y_axis = property(lambda s: s._get("y_axis"))
    z_axis = property(lambda s: s._get("z_axis"))

    def moved(self) -> bool:
        return bool(self._c.moved)

    def set_moved(self, moved: bool):
        self._c.moved = int(moved)

    def as_dict(self) -> dict:
        return {k: self._get(k) for k in ("position", "x_axis", "y_axis", "z_axis")}

    def process_input(self, controller: "CameraController", keys: int = 0, cursor=(0.0, 0.0),
                      delta_time: float = 0.0):
        """Camera::processInput (Camera.cpp:13-90) for one frame: `keys` is a mask of KEY_* held
        down, `cursor` the window's cursor position."""
        st = _InputState(int(keys), float(cursor[0]), float(cursor[1]))
        _check(lib().cupbr_camera_process_input(C.byref(self._c), C.byref(controller._c), C.byref(st),
                                                float(delta_time)))


class CameraController:
    """The input state Camera keeps privately (Camera.h:78-85): aspect ratio, pitch, yaw, and the
    last cursor position."""

    def __init__(self, width: int, height: int):
        self._c = _CameraController()
        _check(lib().cupbr_camera_controller_init(C.byref(self._c), width, height))

    yaw = property(lambda s: s._c.yaw)
    pitch = property(lambda s: s._c.pitch)


# ------------------------------------------------------------------ Scene
def _desc_to_arrays(desc: dict):
    types = np.asarray(desc["types"], np.int32)
    n = len(types)
    objs = np.zeros(n, OBJECT_DTYPE)
    objs["type"] = types
    objs["params"] = np.asarray(desc["params"], np.float32).reshape(n, 12)
    objs["first_triangle"] = np.asarray(desc["first"], np.uint32)
    objs["triangle_count"] = np.asarray(desc["count"], np.uint32)
    m = np.asarray(desc["mats"], np.float32).reshape(n, 21)
    objs["albedo_e"], objs["albedo_d"], objs["albedo_s"] = m[:, 0:3], m[:, 3:6], m[:, 6:9]
    objs["shininess"], objs["roughness"], objs["eta"] = m[:, 9], m[:, 10], m[:, 11]
    objs["sigma_s"], objs["sigma_a"], objs["g"] = m[:, 12:15], m[:, 15:18], m[:, 18]
    objs["interface_type"] = m[:, 19].astype(np.int32)
    objs["material_type"] = m[:, 20].astype(np.int32)
    lt = np.asarray(desc["lights"], np.float32).reshape(-1, 16)
    lights = np.zeros(len(lt), LIGHT_DTYPE)
    lights["type"] = lt[:, 0].astype(np.int32)
    lights["position"], lights["intensity"], lights["radiance"] = lt[:, 1:4], lt[:, 4:7], lt[:, 7:10]
    lights["half_extend1"], lights["half_extend2"] = lt[:, 10:13], lt[:, 13:16]
    tris = np.ascontiguousarray(np.asarray(desc["tris"], np.float32).reshape(-1, 9))
    env = desc.get("env")
    env = None if env is None else np.ascontiguousarray(np.asarray(env, np.float32))
    vol = _Volume()
    v = np.asarray(desc.get("volume", [1, 1, 1, 1, 1, 1, 0]), np.float32)
    vol.sigma_s[:] = v[0:3].tolist()
    vol.sigma_a[:] = v[3:6].tolist()
    vol.g = float(v[6])
    return objs, tris, lights, env, vol


class Scene:
    """Scene handle (the reference's Scene POD, lib/include/Scene/Scene.h:15-34)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def from_desc(cls, desc: dict) -> "Scene":
        """Builds a scene from the flat description (see cupbr_amd.scenes)."""
        objs, tris, lights, env, vol = _desc_to_arrays(desc)
        h = C.c_void_p()
        eh, ew = (0, 0) if env is None else env.shape[:2]
        _check(lib().cupbr_scene_create(_ptr(objs), len(objs), _ptr(tris), len(tris), _ptr(lights), len(lights),
                                        _ptr(env), ew, eh, C.byref(vol), C.byref(h)))
        return cls(h)

    def destroy(self):
        if self._h:
            _check(lib().cupbr_scene_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @property
    def handle(self):
        if not self._h:
            raise CupbrError(-1, "scene was destroyed")
        return self._h

    def upload(self):
        """Copies the scene to the GPU now (otherwise the first trace/render does it)."""
        _check(lib().cupbr_scene_upload(self.handle))

    def counts(self):
        n, t, l, e = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_int32()
        _check(lib().cupbr_scene_get_counts(self.handle, C.byref(n), C.byref(t), C.byref(l), C.byref(e)))
        return {"objects": n.value, "triangles": t.value, "lights": l.value, "environment": bool(e.value)}

    def objects(self) -> np.ndarray:
        n = self.counts()["objects"]
        out = np.zeros(n, OBJECT_DTYPE)
        _check(lib().cupbr_scene_get_objects(self.handle, _ptr(out), n))
        return out

    def triangles(self) -> np.ndarray:
        n = self.counts()["triangles"]
        out = np.zeros((n, 9), np.float32)
        _check(lib().cupbr_scene_get_triangles(self.handle, _ptr(out), n))
        return out

    def lights(self) -> np.ndarray:
        n = self.counts()["lights"]
        out = np.zeros(n, LIGHT_DTYPE)
        _check(lib().cupbr_scene_get_lights(self.handle, _ptr(out), n))
        return out

    def volume(self) -> np.ndarray:
        v = _Volume()
        _check(lib().cupbr_scene_get_volume(self.handle, C.byref(v)))
        return np.array(list(v.sigma_s) + list(v.sigma_a) + [v.g], np.float32)

    def environment(self):
        """The environment map as float32 [h, w, 3] (row 0 first as in the file), or None."""
        w, h = C.c_uint32(), C.c_uint32()
        _check(lib().cupbr_scene_get_environment(self.handle, None, 0, C.byref(w), C.byref(h)))
        if w.value == 0:
            return None
        out = np.zeros((h.value, w.value, 3), np.float32)
        _check(lib().cupbr_scene_get_environment(self.handle, _ptr(out), out.size, None, None))
        return out

    def bvh_stats(self) -> dict:
        s = _BvhStats()
        _check(lib().cupbr_scene_get_bvh_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _BvhStats._fields_}

    def bvh_nodes(self):
        """(fp32, half): the BVH4 nodes in both device forms as raw 128-byte records (uint8 arrays of
        shape (nodes, 128)); the layouts are described at cupbr_scene_get_bvh_nodes."""
        n = self.bvh_stats()["nodes"]
        f = np.zeros((max(n, 1), 128), np.uint8)
        h = np.zeros((max(n, 1), 128), np.uint8)
        _check(lib().cupbr_scene_get_bvh_nodes(self.handle, _ptr(f), _ptr(h), max(n, 1)))
        return f[:n], h[:n]

    def reference_order(self) -> np.ndarray:
        n = self.counts()["objects"]
        out = np.zeros(max(n, 1), np.int32)
        _check(lib().cupbr_scene_get_reference_order(self.handle, _ptr(out), max(n, 1)))
        return out[:n]

    def pick(self, camera: "Camera", x: int, y: int, width: int, height: int):
        """Interaction::pickMouse: (scene_index, material record) of the object seen through pixel
        (x, y), or None on a miss."""
        m = np.zeros(1, MATERIAL_DTYPE)
        idx = C.c_int32(-1)
        _check(lib().cupbr_pick(self.handle, C.byref(camera._c), x, y, width, height, _ptr(m), C.byref(idx)))
        return None if idx.value < 0 else (idx.value, m[0])

    def update_material(self, scene_index: int, material):
        """Interaction::updateMaterial: material is a MATERIAL_DTYPE record (e.g. from pick)."""
        m = np.zeros(1, MATERIAL_DTYPE)
        m[0] = material
        _check(lib().cupbr_scene_update_material(self.handle, int(scene_index), _ptr(m)))

    def trace_closest(self, rays):
        """Closest hit for n x 6 rays (origin, normalised direction): (object_id, triangle_id, t).

        numpy input goes through host memory; a torch tensor on the GPU stays on the device."""
        if _is_torch_gpu(rays):
            import torch
            rays = rays.contiguous().float()
            n = rays.numel() // 6
            obj = torch.empty(n, dtype=torch.int32, device=rays.device)
            tri = torch.empty_like(obj)
            t = torch.empty(n, dtype=torch.float32, device=rays.device)
            stream = torch.cuda.current_stream(rays.device).cuda_stream
            _check(lib().cupbr_trace_closest_device(self.handle, rays.data_ptr(), n, obj.data_ptr(), tri.data_ptr(),
                                                    t.data_ptr(), stream))
            return obj, tri, t
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = len(rays)
        obj = np.zeros(n, np.int32)
        tri = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        _check(lib().cupbr_trace_closest(self.handle, _ptr(rays), n, _ptr(obj), _ptr(tri), _ptr(t)))
        return obj, tri, t

    def trace_visibility(self, rays, distance):
        """Tracing::traceVisibility for n rays: True where nothing is hit within distance."""
        if _is_torch_gpu(rays):
            import torch
            rays = rays.contiguous().float()
            n = rays.numel() // 6
            dist = distance.contiguous().float()
            vis = torch.empty(n, dtype=torch.uint8, device=rays.device)
            stream = torch.cuda.current_stream(rays.device).cuda_stream
            _check(lib().cupbr_trace_visibility_device(self.handle, rays.data_ptr(), dist.data_ptr(), n,
                                                       vis.data_ptr(), stream))
            return vis.bool()
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        dist = np.ascontiguousarray(distance, np.float32).reshape(-1)
        vis = np.zeros(len(rays), np.uint8)
        _check(lib().cupbr_trace_visibility(self.handle, _ptr(rays), _ptr(dist), len(rays), _ptr(vis)))
        return vis.astype(bool)


NRC_PARAMS = 16576
NRC_ROW_FLOATS = 20


class RadianceCache:
    """cupbr_nrc: the neural radiance cache (64 -> 4x64 ReLU -> 3 MLP, BF16 matrix cores).

    Inputs per query (16 floats): position[3], outgoing direction[3], normal[3], roughness,
    diffuse albedo[3], specular albedo[3]. Device arrays are torch CUDA tensors."""

    def __init__(self, box_min, box_max, seed: int = 0):
        lo = (C.c_float * 3)(*box_min)
        hi = (C.c_float * 3)(*box_max)
        h = C.c_void_p()
        _check(lib().cupbr_nrc_create(lo, hi, seed, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def destroy(self):
        if self._h:
            lib().cupbr_nrc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def get_weights(self) -> np.ndarray:
        w = np.zeros(NRC_PARAMS, np.float32)
        _check(lib().cupbr_nrc_get_weights(self._h, _ptr(w)))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32).reshape(-1)
        if w.size != NRC_PARAMS:
            raise CupbrError(-13, f"expected {NRC_PARAMS} weights")
        _check(lib().cupbr_nrc_set_weights(self._h, _ptr(w)))

    def infer(self, inputs):
        """inputs: CUDA tensor [n, 16] -> predictions [n, 3] (on torch's current stream)."""
        import torch
        x = inputs.contiguous().float()
        n = x.shape[0]
        out = torch.empty(n, 3, dtype=torch.float32, device=x.device)
        stream = torch.cuda.current_stream(x.device).cuda_stream
        _check(lib().cupbr_nrc_infer_device(self._h, x.data_ptr(), n, out.data_ptr(), stream))
        return out

    def train(self, rows, batch: int, steps: int, lr: float = 1e-3, return_loss: bool = False):
        """rows: CUDA tensor [n, 20] (16 inputs, 3 targets, 1 unused). Returns the per-step summed
        batch loss when return_loss."""
        import torch
        r = rows.contiguous().float()
        if r.shape[-1] != NRC_ROW_FLOATS:
            raise CupbrError(-13, "rows must have 20 floats each")
        loss = torch.zeros(steps, dtype=torch.float32, device=r.device) if return_loss else None
        stream = torch.cuda.current_stream(r.device).cuda_stream
        _check(lib().cupbr_nrc_train_device(self._h, r.data_ptr(), r.shape[0], batch, steps, lr,
                                            None if loss is None else loss.data_ptr(), stream))
        return loss

    @staticmethod
    def matrix_core_product(a, b) -> np.ndarray:
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        assert a.shape == (16, 32) and b.shape == (32, 16)
        d = np.zeros((16, 16), np.float32)
        _check(lib().cupbr_nrc_matrix_core_product(_ptr(a), _ptr(b), _ptr(d)))
        return d


class GridMedium:
    """cupbr_grid_medium: a box medium whose extinction is sigma_t_scale * density(x), density
    trilinear between the centres of an [nz, ny, nx] voxel grid (no reference counterpart)."""

    def __init__(self, shape, bmin=(0.0, 0.0, 0.0), bmax=(1.0, 1.0, 1.0), sigma_t_scale=10.0, majorant=10.0,
                 scattering_albedo=0.9, g=0.6, min_collision_probability=1e-3):
        nz, ny, nx = shape
        self.shape = (int(nz), int(ny), int(nx))
        self._c = _GridMedium((C.c_float * 3)(*bmin), (C.c_float * 3)(*bmax), int(nx), int(ny), int(nz),
                              sigma_t_scale, majorant, scattering_albedo, g, min_collision_probability)

    @classmethod
    def from_dict(cls, shape, d):
        return cls(shape, **d)


class GridRenderer:
    """Heterogeneous-medium path tracer and its path-replay adjoint (cupbr_grid_*).

    numpy arguments go through host memory; torch GPU tensors stay on the device and run on
    torch's current stream. Images are [h, w, 3] (row 0 = bottom), density [nz, ny, nx],
    albedo [n_surfaces, 3]."""

    def __init__(self, scene: Scene, camera: Camera, medium: GridMedium, width: int, height: int, spp: int = 1,
                 max_depth: int = 6, frame0: int = 0):
        self.scene, self.camera, self.medium = scene, camera, medium
        self.width, self.height, self.spp, self.max_depth, self.frame0 = width, height, spp, max_depth, frame0

    def set_counting(self, enable: bool):
        _check(lib().cupbr_grid_set_counting(self.scene.handle, int(enable)))

    def counts(self) -> dict:
        c = (C.c_uint64 * 4)()
        _check(lib().cupbr_grid_get_counts(self.scene.handle, c))
        return dict(closest_rays=c[0], shadow_rays=c[1], grid_reads=c[2], atomics=c[3])

    def surface_count(self) -> int:
        n = C.c_uint32()
        _check(lib().cupbr_grid_surface_count(self.scene.handle, C.byref(n)))
        return n.value

    def _args(self):
        return (self.scene.handle, C.byref(self.camera._c), C.byref(self.medium._c))

    def _check_sizes(self, density, albedo, proposal=None, d_image=None):
        """The C ABI takes plain pointers: sizes are checked here before any copy or launch."""
        nv = self.medium.shape[0] * self.medium.shape[1] * self.medium.shape[2]
        if _numel(density) != nv or (proposal is not None and _numel(proposal) != nv):
            raise CupbrError(-13, f"density/proposal must have {nv} values (grid {self.medium.shape})")
        if _numel(albedo) != 3 * self.surface_count():
            raise CupbrError(-13, f"albedo must have 3 x {self.surface_count()} values")
        if d_image is not None and _numel(d_image) != 3 * self.width * self.height:
            raise CupbrError(-13, f"d_image must have {3 * self.width * self.height} values")

    def _tail(self):
        return (self.width, self.height, self.frame0, self.spp, self.max_depth)

    def render(self, density, albedo, proposal=None):
        self._check_sizes(density, albedo, proposal)
        if _is_torch_gpu(density):
            import torch
            rho = density.detach().contiguous().float()
            alb = albedo.detach().contiguous().float()
            prop = None if proposal is None else proposal.detach().contiguous().float()
            img = torch.empty(self.height, self.width, 3, dtype=torch.float32, device=rho.device)
            stream = torch.cuda.current_stream(rho.device).cuda_stream
            _check(lib().cupbr_grid_render_device(*self._args(), rho.data_ptr(), None if prop is None else prop.data_ptr(),
                                                  alb.data_ptr(), *self._tail(), img.data_ptr(), stream))
            return img
        rho = np.ascontiguousarray(density, np.float32)
        alb = np.ascontiguousarray(albedo, np.float32)
        prop = None if proposal is None else np.ascontiguousarray(proposal, np.float32)
        img = np.zeros((self.height, self.width, 3), np.float32)
        _check(lib().cupbr_grid_render(*self._args(), _ptr(rho), _ptr(prop), _ptr(alb), *self._tail(), _ptr(img)))
        return img

    def adjoint(self, density, albedo, d_image, proposal=None, grad_density=None, grad_albedo=None,
                from_forward=False):
        """dLoss/d(density), dLoss/d(albedo) for the image cotangent d_image. With torch GPU tensors
        the gradients are added to grad_density / grad_albedo when given (else new zero tensors).
        from_forward (torch GPU tensors only): take the paths' radiance from the preceding render
        call with the same tensors (cupbr_grid_adjoint_from_forward_device); CupbrError if there
        is none."""
        self._check_sizes(density, albedo, proposal, d_image)
        if from_forward and not _is_torch_gpu(density):
            raise ValueError("from_forward needs torch GPU tensors (the forward's samples stay on the device)")
        if _is_torch_gpu(density):
            import torch
            rho = density.detach().contiguous().float()
            alb = albedo.detach().contiguous().float()
            prop = None if proposal is None else proposal.detach().contiguous().float()
            di = d_image.detach().contiguous().float()
            for name, buf, like in (("grad_density", grad_density, rho), ("grad_albedo", grad_albedo, alb)):
                if buf is not None and not (_is_torch_gpu(buf) and buf.dtype == torch.float32 and buf.is_contiguous()
                                            and buf.device == like.device and buf.numel() == like.numel()):
                    raise ValueError(f"{name} must be a contiguous float32 tensor on {like.device} with "
                                     f"{like.numel()} elements")
            gr = torch.zeros_like(rho) if grad_density is None else grad_density
            ga = torch.zeros_like(alb) if grad_albedo is None else grad_albedo
            stream = torch.cuda.current_stream(rho.device).cuda_stream
            fn = lib().cupbr_grid_adjoint_from_forward_device if from_forward else lib().cupbr_grid_adjoint_device
            _check(fn(*self._args(), rho.data_ptr(), None if prop is None else prop.data_ptr(), alb.data_ptr(),
                      *self._tail(), di.data_ptr(), gr.data_ptr(), ga.data_ptr(), stream))
            return gr, ga
        rho = np.ascontiguousarray(density, np.float32)
        alb = np.ascontiguousarray(albedo, np.float32)
        prop = None if proposal is None else np.ascontiguousarray(proposal, np.float32)
        di = np.ascontiguousarray(d_image, np.float32)
        gr = np.zeros_like(rho)
        ga = np.zeros_like(alb)
        _check(lib().cupbr_grid_adjoint(*self._args(), _ptr(rho), _ptr(prop), _ptr(alb), *self._tail(), _ptr(di),
                                        _ptr(gr), _ptr(ga)))
        return gr, ga


def read_pfm(path) -> np.ndarray:
    """Reads a colour PFM as written by PBRenderer.save_pfm: (h, w, 3) float32, row 0 = bottom."""
    with open(path, "rb") as f:
        if f.readline().strip() != b"PF":
            raise CupbrError(-5, f"{path}: not a colour PFM")
        w, h = (int(v) for v in f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(w * h * 12), dtype="<f4" if scale < 0 else ">f4")
    return data.reshape(h, w, 3).astype(np.float32)


def _numel(x) -> int:
    return int(x.numel()) if hasattr(x, "numel") and callable(x.numel) else int(np.size(x))


def _is_torch_gpu(x) -> bool:
    return type(x).__module__.startswith("torch") and getattr(x, "is_cuda", False)


class SceneLoader:
    """SceneLoader (lib/include/Scene/SceneLoader.h:15-21)."""

    @staticmethod
    def load_from_file(path: str) -> Scene:
        h = C.c_void_p()
        _check(lib().cupbr_scene_load_xml(path.encode(), C.byref(h)))
        return Scene(h)

    @staticmethod
    def destroy_scene(scene: Scene):
        scene.destroy()


# ------------------------------------------------------------------ PBRenderer
class PBRenderer:
    """PBRenderer (lib/include/Renderer/PBRenderer.h:27-116)."""

    def __init__(self, method: RenderingMethod = RenderingMethod.RAYTRACER):
        h = C.c_void_p()
        _check(lib().cupbr_renderer_create(int(method), C.byref(h)))
        self._h = h
        self._size = (0, 0)
        self._scene = None

    def __del__(self):
        try:
            if self._h:
                lib().cupbr_renderer_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def set_output_size(self, width: int, height: int):
        _check(lib().cupbr_renderer_set_output_size(self._h, width, height))
        self._size = (width, height)

    def register_scene(self, scene: Scene):
        _check(lib().cupbr_renderer_register_scene(self._h, scene.handle))
        self._scene = scene

    def set_method(self, method: RenderingMethod):
        _check(lib().cupbr_renderer_set_method(self._h, int(method)))

    def get_method(self) -> RenderingMethod:
        m = C.c_int32()
        _check(lib().cupbr_renderer_get_method(self._h, C.byref(m)))
        return RenderingMethod(m.value)

    def reset(self):
        _check(lib().cupbr_renderer_reset(self._h))

    def set_russian_roulette(self, enable: bool):
        _check(lib().cupbr_renderer_set_russian_roulette(self._h, int(bool(enable))))

    def set_max_trace_depth(self, depth: int):
        _check(lib().cupbr_renderer_set_max_trace_depth(self._h, depth))

    def get_max_trace_depth(self) -> int:
        d = C.c_uint32()
        _check(lib().cupbr_renderer_get_max_trace_depth(self._h, C.byref(d)))
        return d.value

    def frame_index(self) -> int:
        f = C.c_uint32()
        _check(lib().cupbr_renderer_get_frame_index(self._h, C.byref(f)))
        return f.value

    def render(self, camera: Camera, spp: int = 1):
        """render(Camera*) repeated spp times (asynchronous on the renderer's stream)."""
        _check(lib().cupbr_render(self._h, C.byref(camera._c), spp))

    def synchronize(self):
        _check(lib().cupbr_renderer_synchronize(self._h))

    def get_image(self, which: Image = Image.OUTPUT) -> np.ndarray:
        w, h = self._size
        out = np.zeros((h, w, 3), np.float32)
        _check(lib().cupbr_renderer_get_image(self._h, int(which), _ptr(out), out.size))
        return out

    def get_output_image(self) -> np.ndarray:
        return self.get_image(Image.OUTPUT)

    def get_gradient_x(self) -> np.ndarray:
        return self.get_image(Image.GRADIENT_X)

    def get_gradient_y(self) -> np.ndarray:
        return self.get_image(Image.GRADIENT_Y)

    def image_device_ptr(self, which: Image = Image.OUTPUT) -> int:
        p = C.c_void_p()
        _check(lib().cupbr_renderer_get_image_device(self._h, int(which), C.byref(p)))
        return p.value

    @property
    def size(self):
        return self._size

    def set_profiling(self, enable: bool):
        """Per-stage device time (HIP events on the renderer's stream), reported by stats()."""
        _check(lib().cupbr_renderer_set_profiling(self._h, int(bool(enable))))

    def save_pfm(self, path: str, which: Image = Image.OUTPUT):
        """FP32 dump of an image as a colour PFM (rows bottom-up)."""
        _check(lib().cupbr_renderer_save_pfm(self._h, int(which), str(path).encode()))

    def resume(self, image, frame_index: int):
        """Continue the running mean of a saved run: output image (h, w, 3) and its frame count."""
        img = np.ascontiguousarray(image, np.float32)
        _check(lib().cupbr_renderer_resume(self._h, _ptr(img), img.size, int(frame_index)))

    def set_radiance_cache(self, cache: "RadianceCache | None", training_paths: int = 1 << 16, train_steps: int = 16,
                           batch: int = 1 << 14, learning_rate: float = 1e-3):
        """Attach a RadianceCache (None = plain PT); see cupbr_renderer_set_radiance_cache."""
        self._cache = cache
        _check(lib().cupbr_renderer_set_radiance_cache(self._h, None if cache is None else cache.handle,
                                                       training_paths, train_steps, batch, learning_rate))

    def set_radiance_cache_camera_hit(self, min_roughness: float = float("inf")):
        """Paths also end into the cache at the camera hit on Lambertian surfaces (min_roughness <= 1)
        and GGX surfaces of roughness >= min_roughness; inf = never (the default). See
        cupbr_renderer_set_radiance_cache_camera_hit."""
        _check(lib().cupbr_renderer_set_radiance_cache_camera_hit(self._h, float(min_roughness)))

    def set_rng(self, rng: int = 0):
        """RNG_TEA_LCG (the reference's, the default) or RNG_PCG for PATHTRACER frames; see
        cupbr_renderer_set_rng."""
        _check(lib().cupbr_renderer_set_rng(self._h, int(rng)))

    def set_gdpt_reconstruction(self, mode: int = 0, alpha: float = 0.2, iterations: int = 50):
        """GRADIENTDOMAIN reconstruction: GDPT_TRIMMED_MEAN (the reference's) or
        GDPT_SCREENED_POISSON (L2, weight alpha on the primal image, Jacobi iterations)."""
        _check(lib().cupbr_renderer_set_gdpt_reconstruction(self._h, int(mode), float(alpha), int(iterations)))

    def set_traversal_counting(self, enable: bool):
        """Counts BVH4 node visits and primitive tests (reported by stats())."""
        _check(lib().cupbr_renderer_set_traversal_counting(self._h, int(bool(enable))))

    def stats(self) -> dict:
        s = _RenderStats()
        _check(lib().cupbr_renderer_get_stats(self._h, C.byref(s)))
        out = {k: getattr(s, k) for k, _ in _RenderStats._fields_}
        out["stage_ms"] = list(s.stage_ms)
        out["stage_launches"] = list(s.stage_launches)
        return out


# ------------------------------------------------------------------ post chain
class PostProcessor:
    """PostProcessor (lib/include/PostProcessing/PostProcessor.h:12-52)."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().cupbr_post_create(C.byref(h)))
        self._h = h
        self._size = (0, 0)

    def __del__(self):
        try:
            if self._h:
                lib().cupbr_post_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def register_image(self, device_ptr: int, width: int, height: int):
        _check(lib().cupbr_post_register_image(self._h, C.c_void_p(device_ptr), width, height))
        self._size = (width, height)

    def register_renderer(self, renderer: PBRenderer):
        """registerImage(pbrenderer.getOutputImage())"""
        w, h = renderer.size
        self.register_image(renderer.image_device_ptr(Image.OUTPUT), w, h)

    def bloom(self, threshold):
        c = (C.c_float * 4)(*[float(x) for x in threshold])
        _check(lib().cupbr_post_bloom(self._h, c))

    def set_bloom_pyramid(self, pyramid: "BloomPyramid"):
        """BloomPyramid.REFERENCE (the default) or GAUSSIAN5; see cupbr_post_set_bloom_pyramid."""
        _check(lib().cupbr_post_set_bloom_pyramid(self._h, int(pyramid)))

    def get_post_process_buffer(self) -> np.ndarray:
        w, h = self._size
        out = np.zeros((h, w, 3), np.float32)
        _check(lib().cupbr_post_get_buffer(self._h, _ptr(out), out.size))
        return out

    def buffer_device_ptr(self) -> int:
        p = C.c_void_p()
        _check(lib().cupbr_post_get_buffer_device(self._h, C.byref(p)))
        return p.value


class ToneMapper:
    """ToneMapper (lib/include/Renderer/ToneMapper.h:22-80)."""

    def __init__(self, type: ToneMappingType = ToneMappingType.REINHARD):
        h = C.c_void_p()
        _check(lib().cupbr_tonemapper_create(int(type), C.byref(h)))
        self._h = h
        self._size = (0, 0)

    def __del__(self):
        try:
            if self._h:
                lib().cupbr_tonemapper_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def register_image(self, device_ptr: int, width: int, height: int):
        _check(lib().cupbr_tonemapper_register_image(self._h, C.c_void_p(device_ptr), width, height))
        self._size = (width, height)

    def register_renderer(self, renderer: PBRenderer):
        w, h = renderer.size
        self.register_image(renderer.image_device_ptr(Image.OUTPUT), w, h)

    def tone_map(self):
        _check(lib().cupbr_tonemapper_tonemap(self._h))

    def tone_map_bloom(self, post: PostProcessor, threshold):
        c = (C.c_float * 4)(*[float(x) for x in threshold])
        _check(lib().cupbr_tonemapper_tonemap_bloom(self._h, post._h, c))

    def get_render_buffer(self) -> np.ndarray:
        w, h = self._size
        out = np.zeros((h, w, 4), np.uint8)
        _check(lib().cupbr_tonemapper_get_render_buffer(self._h, _ptr(out), out.size))
        return out

    def set_type(self, type: ToneMappingType):
        _check(lib().cupbr_tonemapper_set_type(self._h, int(type)))

    def get_type(self) -> ToneMappingType:
        t = C.c_int32()
        _check(lib().cupbr_tonemapper_get_type(self._h, C.byref(t)))
        return ToneMappingType(t.value)

    def set_exposure(self, exposure: float):
        _check(lib().cupbr_tonemapper_set_exposure(self._h, exposure))

    def save_to_file(self, path: str):
        _check(lib().cupbr_tonemapper_save_to_file(self._h, path.encode()))
