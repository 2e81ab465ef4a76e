"""GPU parity on the benchmark workloads at reduced size.

The 1,021,120-triangle scene of the headline benchmark (`scenes.million_triangles`: 11 displaced
icospheres and 10,000 boxes, Lambert / GGX / glass materials, 4 area lights and a 512x256
environment map) rendered by PT+NEE at depth 8, against the CPU oracle with the same seeds:

  FP32 oracle, same seeds      rel-MSE <= 1e-6 (240x135, 8 spp)
  FP64 oracle, same seeds      rel-MSE <= 1e-4 (128x72, 64 spp; the FP32 and FP64 estimators take
                               different branches on a few paths, e.g. glass reflect/refract
                               decisions at the rounding edge, which converges as 1/spp)
  FP64 oracle, other seeds     per-pixel means within 3 sigma (Welch) for >= 99.5% of pixel channels
  bloom + tone map             on the GPU's own 1920x1080 frame: bloom bit-exact, bytes within 1

and the radiance cache on the same scene (estimator preservation, bit for bit), and the
heterogeneous-medium adjoint at 64^3 against central differences (h = 1e-3) of the GPU's forward
renders. rel-MSE = mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2).
"""
import numpy as np
import pytest

import cupbr_amd as cb
import oracle
from cupbr_amd import scenes

pytestmark = pytest.mark.gpu

DEPTH = 8


def rel_mse(x, r):
    return float(np.mean((np.asarray(x, np.float64) - r) ** 2 / (np.asarray(r, np.float64) ** 2 + 1e-2)))


@pytest.fixture(scope="module")
def million():
    desc, ca = scenes.million_triangles()
    return desc, ca, cb.Scene.from_desc(desc), oracle.OracleScene(desc, precisions=(32, 64))


def _camera(ca, w, h):
    return cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], w, h)


def _renderer(scene, w, h, depth=DEPTH):
    r = cb.PBRenderer(cb.RenderingMethod.PATHTRACER)
    r.set_output_size(w, h)
    r.register_scene(scene)
    r.set_max_trace_depth(depth)
    return r


def test_million_triangles_same_seed_fp32(million):
    desc, ca, scene, ref = million
    w, h, spp = 240, 135, 8
    cam = _camera(ca, w, h)
    r = _renderer(scene, w, h)
    r.render(cam, spp)
    img = r.get_output_image()
    want = ref.render(cam.as_dict(), w, h, 0, spp, depth=DEPTH, precision=32)
    e = rel_mse(img, want)
    print(f"1M tris {w}x{h} {spp} spp depth {DEPTH}: rel-MSE vs same-seed FP32 oracle {e:.2e}; "
          f"exact pixels {np.mean(img == want):.4f}; mean {img.mean():.4f}", flush=True)
    assert np.isfinite(img).all() and float(img.mean()) > 0
    assert e <= 1e-6


def test_million_triangles_against_fp64_oracle(million):
    desc, ca, scene, ref = million
    w, h = 128, 72
    cam = _camera(ca, w, h)
    n = 64
    r = _renderer(scene, w, h)
    r.render(cam, n)
    e = rel_mse(r.get_output_image(), ref.render(cam.as_dict(), w, h, 0, n, depth=DEPTH, precision=64))

    # independent seeds: GPU frames 0..255 as 128 batch means of 2 frames (recovered from the
    # running mean after each call), FP64 oracle frames 256..511 with per-sample second moments
    batches, m = 128, 2
    r = _renderer(scene, w, h)
    means = []
    for _ in range(batches):
        r.render(cam, m)
        means.append(r.get_output_image().astype(np.float64))
    assert r.frame_index() == batches * m
    b = np.stack([(k + 1) * means[k] - k * means[k - 1] if k else means[0] for k in range(batches)])
    mg, vg = b.mean(0), b.var(0, ddof=1) / batches
    no = batches * m
    _, s, s2 = ref.render(cam.as_dict(), w, h, no, no, depth=DEPTH, precision=64, stats=True)
    mo = s / no
    vo = np.maximum(s2 / no - mo ** 2, 0) / (no - 1)
    z = np.abs(mg - mo) / np.sqrt(vg + vo + 1e-30)
    frac = float(np.mean(z > 3))
    print(f"1M tris {w}x{h}: rel-MSE vs same-seed FP64 oracle at {n} spp {e:.2e}; 3-sigma outliers "
          f"{frac:.4f} ({no} samples per side)", flush=True)
    assert e <= 1e-4
    assert frac <= 0.005


def test_post_chain_on_a_1080p_frame(million):
    """Bloom (curve of threshold 1, knee 1: Interactor.cpp:53-54,436-440) and both tone curves on
    the GPU's own 1920x1080 render of the scene, against the oracle run on the same HDR image."""
    torch = pytest.importorskip("torch")
    desc, ca, scene, _ = million
    w, h = 1920, 1080
    r = _renderer(scene, w, h)
    r.render(_camera(ca, w, h), 1)
    hdr = r.get_output_image()
    assert np.isfinite(hdr).all() and float(hdr.max()) > 1.0  # the lights and env exceed the threshold
    curve = cb.threshold_curve(1.0, 1.0)
    post = cb.PostProcessor()
    post.register_renderer(r)
    post.bloom(curve)
    got = post.get_post_process_buffer()
    want = oracle.bloom(hdr, curve)
    assert np.array_equal(got, want), float(np.abs(got - want).max())
    for kind in (cb.ToneMappingType.REINHARD, cb.ToneMappingType.GAMMA, cb.ToneMappingType.ACES):
        tm = cb.ToneMapper(kind)
        tm.set_exposure(1.0)
        tm.register_image(post.buffer_device_ptr(), w, h)
        tm.tone_map()
        a = tm.get_render_buffer().astype(np.int32)
        b = oracle.tonemap(want, int(kind), 1.0).astype(np.int32)
        print(f"1080p tone map {kind.name}: {np.mean(a != b):.2e} of bytes differ by 1", flush=True)
        assert np.abs(a - b).max() <= 1 and np.mean(a != b) < 1e-3
        tm2 = cb.ToneMapper(kind)
        tm2.set_exposure(1.0)
        tm2.register_renderer(r)
        tm2.tone_map_bloom(post, curve)
        assert np.array_equal(tm2.get_render_buffer(), tm.get_render_buffer())
    # the 5-level Gaussian pyramid that the headline configuration names, as the bench runs it
    post.set_bloom_pyramid(cb.BloomPyramid.GAUSSIAN5)
    post.bloom(curve)
    got = post.get_post_process_buffer()
    want = oracle.bloom_gaussian(hdr, curve)
    assert np.array_equal(got, want), float(np.abs(got - want).max())
    tm = cb.ToneMapper(cb.ToneMappingType.REINHARD)
    tm.register_renderer(r)
    tm.tone_map_bloom(post, curve)
    a = tm.get_render_buffer().astype(np.int32)
    b = oracle.tonemap(want, int(cb.ToneMappingType.REINHARD), 1.0).astype(np.int32)
    print(f"1080p Gaussian bloom + Reinhard: {np.mean(a != b):.2e} of bytes differ by 1", flush=True)
    assert np.abs(a - b).max() <= 1 and np.mean(a != b) < 1e-3
    del torch


def test_radiance_cache_on_the_million_triangle_scene(million):
    """The cache on the benchmark scene keeps the estimator. With every path a training path the
    image is plain PT at depth 8, bit for bit, while the cache trains on rows of that scene. A zero
    cache ends every rendering path at its second vertex: on the same geometry with the glass
    objects made GGX and without the environment map (glass vertices pass the path on, and an
    escaped second ray would add the environment), that image is plain PT of depth 1."""
    desc, ca, scene, _ = million
    w, h, spp = 240, 135, 2
    cam = _camera(ca, w, h)
    lo, hi = (-5.0, -5.0, -30.0), (105.0, 110.0, 105.0)

    opaque = dict(desc)
    opaque["mats"] = desc["mats"].copy()
    glass = opaque["mats"][:, 20] == scenes.GLASS
    assert glass.any()
    opaque["mats"][glass, 20] = scenes.GGX
    opaque["mats"][glass, 10] = 0.1
    opaque["env"] = None
    s2 = cb.Scene.from_desc(opaque)
    zero = cb.RadianceCache(lo, hi)
    zero.set_weights(np.zeros(cb.NRC_PARAMS, np.float32))
    a = _renderer(s2, w, h)
    a.set_radiance_cache(zero, training_paths=0, train_steps=0)
    a.render(cam, spp)
    b = _renderer(s2, w, h, depth=1)
    b.render(cam, spp)
    assert a.stats()["cache_queries"] > 0
    assert np.array_equal(a.get_output_image(), b.get_output_image())
    del a, b, s2

    cache = cb.RadianceCache(lo, hi, seed=3)
    w0 = cache.get_weights()
    t = _renderer(scene, w, h)
    t.set_radiance_cache(cache, training_paths=w * h, train_steps=2, batch=4096)
    t.render(cam, spp)
    st = t.stats()
    p = _renderer(scene, w, h)
    p.render(cam, spp)
    print(f"cache on 1M tris: {st['cache_training_rows']} training rows", flush=True)
    assert st["cache_queries"] == 0 and st["cache_training_rows"] > 0
    assert np.array_equal(t.get_output_image(), p.get_output_image())
    w1 = cache.get_weights()
    assert np.isfinite(w1).all() and not np.array_equal(w1, w0)


def test_grid_adjoint_64_against_central_differences():
    """64^3 density grid (1 MiB, the differentiable-volume workload's size) at 96x96, 8 spp: the
    adjoint's gradient at 32 voxels drawn at random (seeded) from the decile with the largest
    gradients against central differences with h = 1e-3 of the GPU's own FP32 forward renders
    (fixed proposal, so every random decision stays fixed and the loss is smooth). The FP32
    rounding of the rendered images limits the difference quotient for voxels with small
    gradients, hence the decile."""
    G = scenes.grid_medium_scene(res=64)
    w = h = 96
    spp = 8
    ca = G["camera"]
    cam = cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], w, h)
    R = cb.GridRenderer(cb.Scene.from_desc(G["desc"]), cam, cb.GridMedium.from_dict(G["density"].shape, G["medium"]),
                        w, h, spp=spp, max_depth=6)
    rho, alb = G["density"], G["albedo"]
    target = R.render(G["target_density"], alb).astype(np.float64)

    def loss(r):
        return float(np.sum((R.render(r, alb, proposal=rho).astype(np.float64) - target) ** 2))

    img = R.render(rho, alb, proposal=rho)
    gr, _ = R.adjoint(rho, alb, (2 * (img - target)).astype(np.float32), proposal=rho)
    g = np.abs(gr.ravel())
    cand = np.flatnonzero(g >= np.quantile(g, 0.9))
    vox = np.random.default_rng(7).choice(cand, 32, replace=False)
    hh = 1e-3
    errs = []
    for v in vox:
        rp = rho.copy().ravel()
        rm = rho.copy().ravel()
        rp[v] += hh
        rm[v] -= hh
        fd = (loss(rp.reshape(rho.shape)) - loss(rm.reshape(rho.shape))) / (2 * hh)
        errs.append(abs(fd - gr.ravel()[v]) / abs(gr.ravel()[v]))
    errs = np.array(errs)
    print(f"64^3 adjoint vs central differences (h = 1e-3, 32 voxels): median rel error {np.median(errs):.2e}, "
          f"max {errs.max():.2e}", flush=True)
    assert np.median(errs) <= 5e-3
    assert errs.max() <= 5e-2
