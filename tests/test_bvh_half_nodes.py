"""The per-ray traversal's compressed BVH4 nodes (binary16 plane offsets from a reference point), checked on
the CPU against the FP32 nodes they are built from: every decoded plane lies outside its FP32
plane by the stated margin and not much farther, empty slots and NaN bounds decode to planes that
never cull, and an emulation of the kernel's FP32 decode and slab test in numpy accepts every
child box that an exact (float64) slab test of its FP32 box finds on a ray."""
import numpy as np
import pytest

import cupbr_amd as cb
from cupbr_amd import scenes

EMPTY = 0xFFFFFFFF


def _fp32_view(f):
    lo = f[:, :48].view(np.float32).reshape(-1, 3, 4)      # [node, axis, child]
    hi = f[:, 48:96].view(np.float32).reshape(-1, 3, 4)
    child = f[:, 96:112].view(np.uint32)
    return lo, hi, child


def _half_view(h):
    ref = h[:, :12].view(np.float32)                      # [node, axis]
    k = h[:, 12:15].view(np.int8).astype(np.int64)          # [node, axis]
    blocks = h[:, 16:112].view(np.float16).reshape(-1, 3, 2, 8).astype(np.float64)
    child = h[:, 112:128].view(np.uint32)
    return ref, k, blocks, child


def _decode(ref, k, blocks):
    """Planes c + offset 2^k in float64 (exact): lo and hi, [node, axis, child]."""
    s = np.ldexp(1.0, k)[:, :, None]
    c = ref.astype(np.float64)[:, :, None]
    with np.errstate(invalid="ignore"):
        lo = c + blocks[:, :, 0, :4] * s
        hi = c + blocks[:, :, 0, 4:] * s
    return lo, hi


def _mixed_nan_scene():
    b = scenes.SceneBuilder()
    m = scenes.material(scenes.LAMBERT, albedo_d=(0.5,) * 3)
    rng = np.random.default_rng(3)
    for _ in range(300):
        v = rng.uniform(-3, 3, (3, 3))
        b.triangle(v[0], v[1], v[2], m)
    b.triangle((np.nan, 0, 0), (1, 0, 0), (0, 1, 0), m)
    b.triangle((0, 0, 0), (np.inf, 0, 0), (0, 1, 0), m)
    b.sphere((0.5, 0.5, 0.5), 0.25, m)
    return b.desc()


SCENES = {
    "cornell": lambda: scenes.cornell_box()[0],
    "mixed": lambda: scenes.mixed_primitives()[0],
    "meshes": lambda: scenes.random_mesh_scene(40, 300, seed=2),
    "large_coordinates": lambda: scenes.random_triangle_soup(2000, seed=5, extent=5e4, size=30.0),
    "nan_and_inf": _mixed_nan_scene,
}


@pytest.mark.parametrize("name", list(SCENES))
def test_half_nodes_enclose_fp32_boxes(name):
    s = cb.Scene.from_desc(SCENES[name]())
    f, h = s.bvh_nodes()
    lo32, hi32, ch32 = _fp32_view(f)
    ref, k, blocks, ch16 = _half_view(h)
    assert np.array_equal(ch32, ch16)
    assert np.all(k <= 0) and np.all(k >= -100)
    # the second block of each axis is the first with its halves swapped
    assert np.array_equal(blocks[:, :, 1, :4], blocks[:, :, 0, 4:], equal_nan=True)
    assert np.array_equal(blocks[:, :, 1, 4:], blocks[:, :, 0, :4], equal_nan=True)
    lo, hi = _decode(ref, k, blocks)
    empty = (ch32 == EMPTY)[:, None, :].repeat(3, 1)
    assert np.all(np.isposinf(lo[empty])) and np.all(np.isneginf(hi[empty]))
    full = ~empty
    # the node's extent per axis, as the builder measures it
    lo64, hi64 = lo32.astype(np.float64), hi32.astype(np.float64)
    fin_lo = np.where(np.isfinite(lo64), lo64, np.inf).min(2)
    fin_hi = np.where(np.isfinite(hi64), hi64, -np.inf).max(2)
    fin_lo = np.where(np.isfinite(fin_lo), fin_lo, 0.0)
    ext = np.where(np.isfinite(fin_hi), np.maximum(fin_hi - fin_lo, 0.0), 0.0)
    margin = (16.0 * 2.0 ** -24 * ext)[:, :, None].repeat(4, 2)
    slack = (2.0 ** -10 * ext + 1e-30)[:, :, None].repeat(4, 2)  # 11 bits of offsets up to ext / 2
    for q, v, sign in ((lo, lo64, -1.0), (hi, hi64, 1.0)):
        nan = full & np.isnan(v)
        assert np.all(np.isinf(q[nan]) & (np.sign(q[nan]) == sign))  # never culls
        inf = full & np.isinf(v)
        assert np.array_equal(q[inf], v[inf])
        fin = full & np.isfinite(v)
        d = (q[fin] - v[fin]) * sign  # outward distance of the decoded plane
        # offsets beyond binary16's range (a node wider than 2^15 with k capped at 0): the plane
        # moves to the largest offset (lo) or to infinity (hi), looser but still outside
        cf = ref.astype(np.float64)[:, :, None].repeat(4, 2)[fin]
        big = np.abs(v[fin] - cf) * np.ldexp(1.0, -k)[:, :, None].repeat(4, 2)[fin] > 6.5e4
        assert np.all(d >= margin[fin] * (1 - 1e-9)), name
        assert np.all((d <= margin[fin] + slack[fin]) | big | np.isinf(q[fin]))


def _emulated_half_test(ref, k, blocks, o, d):
    """The kernel's decode and slab test in FP32 (fma as one float64 operation rounded to FP32):
    for rays o, d [r, 3] and nodes [n], True where a child is entered, [r, n, 4]."""
    f32 = np.float32
    inv = (f32(1) / d.astype(f32)).astype(f32)                          # [r, 3]
    ne = inv >= 0
    A = np.ldexp(inv[:, None, :].astype(np.float64), k[None]).astype(f32)  # [r, n, 3]
    B = ((ref[None].astype(f32) - o[:, None, :].astype(f32)).astype(f32) * inv[:, None, :]).astype(f32)
    blk = np.where(ne[:, None, :, None], blocks[None, :, :, 0, :], blocks[None, :, :, 1, :])  # [r, n, 3, 8]
    with np.errstate(invalid="ignore", over="ignore"):
        t = (blk * A[..., None].astype(np.float64) + B[..., None].astype(np.float64)).astype(f32)
    near, far = t[..., :4], t[..., 4:]
    tn = np.fmax(np.fmax(np.fmax(near[:, :, 0], near[:, :, 1]), near[:, :, 2]), f32(-1e-5))
    tf = np.fmin(np.fmin(far[:, :, 0], far[:, :, 1]), far[:, :, 2])  # lim = +inf: no hit yet
    return tn <= (tf * f32(1.000002)).astype(f32)


def _exact_fp32_box_test(lo32, hi32, o, d):
    """Exact slab test (float64) of the FP32 child boxes against the rays for t >= 0."""
    o64, d64 = o.astype(np.float64), d.astype(np.float64)
    t1 = (lo32[None].astype(np.float64) - o64[:, None, :, None]) / d64[:, None, :, None]
    t2 = (hi32[None].astype(np.float64) - o64[:, None, :, None]) / d64[:, None, :, None]
    tn = np.minimum(t1, t2).max(2)
    tf = np.maximum(t1, t2).min(2)
    return (tn <= tf) & (tf >= 0)


@pytest.mark.parametrize("name", ["mixed", "meshes", "large_coordinates"])
def test_emulated_half_decode_is_conservative(name):
    s = cb.Scene.from_desc(SCENES[name]())
    f, h = s.bvh_nodes()
    lo32, hi32, ch = _fp32_view(f)
    ref, k, blocks, _ = _half_view(h)
    rng = np.random.default_rng(7)
    used = np.broadcast_to((ch != EMPTY)[:, None, :], lo32.shape)
    assert np.isfinite(lo32[used]).all() and np.isfinite(hi32[used]).all()
    span = float(np.abs(ref).max()) + 1.0
    # rays aimed at random children: at a point inside the box, at one of its corners (a grazing
    # ray, where the exact test finds entry = exit) or at the middle of an edge
    idx = np.argwhere(ch != EMPTY)
    entered = 0
    for _ in range(4):
        o = rng.uniform(-span, span, (96, 3))
        pick = idx[rng.integers(0, len(idx), 96)]
        blo = lo32[pick[:, 0], :, pick[:, 1]].astype(np.float64)
        bhi = hi32[pick[:, 0], :, pick[:, 1]].astype(np.float64)
        u = rng.uniform(0, 1, (96, 3))
        u[32:64] = rng.integers(0, 2, (32, 3))
        u[64:, 0] = 0.5
        u[64:, 1:] = rng.integers(0, 2, (32, 2))
        target = blo + u * (bhi - blo)
        d = target - o
        d[np.abs(d) < 1e-3 * np.abs(d).max(1, keepdims=True)] = 1e-3
        d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        o = o.astype(np.float32)
        ex = _exact_fp32_box_test(lo32, hi32, o, d) & (ch != EMPTY)[None]
        em = _emulated_half_test(ref, k, blocks, o, d)
        assert not np.any(ex & ~em), np.argwhere(ex & ~em)[:5]
        entered += int(ex.sum())
    assert entered > 1000
