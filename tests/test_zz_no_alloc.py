"""cupbr_trace_closest_device_noalloc against cupbr_trace_closest_device, through the C ABI.

The no-allocation variant keeps the traversal-stack spill in the library's fixed per-device buffer
(16 MiB) and cuts the grid to what fits; its ids and depths must equal the allocating entry point's
bit for bit. Two calls in a row, the second with a smaller input, on the 1M-triangle scene (49
stack entries per lane, so the spill is in use) and on the Cornell box (no spill).
"""
import ctypes as C

import numpy as np
import pytest

import cupbr_amd as cb
from cupbr_amd import scenes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _noalloc():
    L = cb.lib()
    fn = L.cupbr_trace_closest_device_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return fn


def _rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return torch.from_numpy(np.concatenate([o, d], 1).astype(np.float32)).cuda()


def _both(scene, rays):
    n = rays.shape[0]
    outs = []
    for use_noalloc in (False, True):
        obj = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        tri = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        t = torch.full((n,), -7.0, dtype=torch.float32, device="cuda")
        stream = torch.cuda.current_stream().cuda_stream
        if use_noalloc:
            rc = _noalloc()(scene.handle, rays.data_ptr(), n, obj.data_ptr(), tri.data_ptr(), t.data_ptr(), stream)
        else:
            rc = cb.lib().cupbr_trace_closest_device(scene.handle, rays.data_ptr(), n, obj.data_ptr(), tri.data_ptr(),
                                                     t.data_ptr(), stream)
        assert rc == 0, cb.lib().cupbr_last_error().decode()
        torch.cuda.synchronize()
        outs.append((obj.cpu().numpy(), tri.cpu().numpy(), t.cpu().numpy()))
    return outs


@pytest.mark.parametrize("name", ["million", "cornell"])
def test_noalloc_equals_allocating_entry_point(name):
    if name == "million":
        desc, _ = scenes.million_triangles()
        lo, hi = 0.0, 100.0
    else:
        desc, _ = scenes.cornell_box()
        lo, hi = 0.0, 1.0
    scene = cb.Scene.from_desc(desc)
    stats_entries = scene.bvh_stats()["stack_entries"]
    print(f"{name}: {stats_entries} stack entries per lane", flush=True)
    # not resident yet: the variant refuses instead of allocating
    r0 = _rays(16, lo, hi, 1)
    z = torch.zeros(16, dtype=torch.int32, device="cuda")
    zf = torch.zeros(16, dtype=torch.float32, device="cuda")
    rc = _noalloc()(scene.handle, r0.data_ptr(), 16, z.data_ptr(), z.data_ptr(), zf.data_ptr(), None)
    assert rc == -1  # CUPBR_ERR_INVALID_ARGUMENT
    scene.upload()
    for n, seed in ((1_500_000, 2), (70_001, 3)):  # second call smaller, ragged
        rays = _rays(n, lo, hi, seed)
        (a_obj, a_tri, a_t), (b_obj, b_tri, b_t) = _both(scene, rays)
        print(f"{name}: {n} rays, {(a_obj >= 0).sum()} hits", flush=True)
        assert (a_obj >= 0).sum() > n // 10
        assert np.array_equal(a_obj, b_obj) and np.array_equal(a_tri, b_tri)
        assert np.array_equal(a_t.view(np.uint32), b_t.view(np.uint32))
