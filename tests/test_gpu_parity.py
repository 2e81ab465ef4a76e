"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Bars (from the renderer core's correctness contract):
  closest hit     object / triangle ids and hit depth bit-exact against the reference traversal
  images          same seeds as the FP32 oracle: rel-MSE <= 1e-6 (path-for-path identical up to
                  last-ulp differences of the device math library); against the FP64 oracle of the
                  same estimator: rel-MSE <= 1e-4, and per-pixel means within 3 sigma for >= 99.5%
                  of pixels on independent seeds
  post chain      bloom bit-exact (no transcendentals); tone-mapped bytes within 1 at truncation
                  boundaries (device expf/powf differ from glibc in the last ulp)
rel-MSE = mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2).
"""
import math

import numpy as np
import pytest

import cupbr_amd as cb
import oracle
from cupbr_amd import scenes

pytestmark = pytest.mark.gpu


def rel_mse(x, r):
    return float(np.mean((np.asarray(x, np.float64) - r) ** 2 / (np.asarray(r, np.float64) ** 2 + 1e-2)))


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = _unit(rng.normal(size=(n, 3)))
    k = n // 10
    d[:k] = np.eye(3)[rng.integers(0, 3, k)] * rng.choice([-1, 1], (k, 1))  # axis-aligned
    d[k:2 * k, 0] = rng.choice([-5e-6, 5e-6, -1e-5, 0.0], k)  # components near the 1e-5 bias
    d[k:2 * k] = _unit(d[k:2 * k])
    return np.concatenate([o, d], 1).astype(np.float32)


def _surface_rays(S, rays, seed):
    obj, _, t = S.trace(rays, precision=32, mode=2)
    hit = obj >= 0
    o = rays[hit, :3] + t[hit, None] * rays[hit, 3:]
    d = _unit(np.random.default_rng(seed).normal(size=(hit.sum(), 3)))
    return np.concatenate([o, d], 1).astype(np.float32)


def _degenerate_scene():
    """Inputs that defeat the SAH build: 600 copies of one triangle in a mesh (every hit ties on t
    and is decided by the reference rank), 400 triangles around one common centroid (no centroid
    split exists) and 150 standalone triangles placed twice."""
    rng = np.random.default_rng(5)
    b = scenes.SceneBuilder()
    m = scenes.material(scenes.LAMBERT, albedo_d=(0.5,) * 3)
    t = np.array([[-1, -1, 0.3], [1, -0.5, 0.2], [0, 1, 0.4]], np.float32)
    b.mesh(np.repeat(t.reshape(1, 9), 600, 0), m)
    ang = rng.uniform(0, 2 * np.pi, (400, 3))
    r = rng.uniform(0.05, 1.5, (400, 1))
    v = np.stack([np.cos(ang), np.sin(ang), rng.uniform(-0.2, 0.2, (400, 3))], -1) * r[:, :, None]
    v = v - v.mean(1, keepdims=True) + np.array([0.2, -0.1, -0.5])
    b.mesh(v.reshape(400, 9), m)
    for _ in range(150):
        tri = rng.uniform(-1.5, 1.5, (3, 3))
        b.triangle(tri[0], tri[1], tri[2], m)
        b.triangle(tri[0], tri[1], tri[2], m)
    return b.desc()


SCENES = {
    "cornell": (lambda: scenes.cornell_box()[0], 0.0, 1.0),
    "mixed": (lambda: scenes.mixed_primitives()[0], -1.5, 4.0),
    "soup": (lambda: scenes.random_triangle_soup(3000, seed=1), -12, 12),
    "meshes": (lambda: scenes.random_mesh_scene(40, 300, seed=2), -12, 12),
    "degenerate": (_degenerate_scene, -2.0, 2.0),
    # nodes far wider than 2^16: the traversal's binary16 plane offsets overflow at the top of the
    # tree (clamped outward), and small ones far from the origin
    "large_coordinates": (lambda: scenes.random_triangle_soup(3000, seed=5, extent=5e4, size=3e3), -6e4, 6e4),
}


@pytest.mark.parametrize("name", list(SCENES))
def test_closest_hit_bit_exact(name):
    make, lo, hi = SCENES[name]
    desc = make()
    g = cb.Scene.from_desc(desc)
    ref = oracle.OracleScene(desc, precisions=(32,))
    rays = _rays(40000, lo, hi, 11)
    rays = np.concatenate([rays, _surface_rays(ref, rays[:10000], 12)])
    go, gt, gd = g.trace_closest(rays)
    ro, rt, rd = ref.trace(rays, precision=32, mode=0)
    print(f"{name}: {len(rays)} rays, {(ro >= 0).sum()} hits", flush=True)
    assert np.array_equal(go, ro) and np.array_equal(gt, rt)
    assert np.array_equal(gd, rd)
    dist = np.random.default_rng(13).uniform(0, 8, len(rays)).astype(np.float32)
    dist[:1000] = np.inf
    assert np.array_equal(g.trace_visibility(rays, dist), ref.visibility(rays, dist, precision=32, mode=0))


def test_closest_hit_million_triangles():
    desc, _ = scenes.million_triangles()
    g = cb.Scene.from_desc(desc)
    ref = oracle.OracleScene(desc, precisions=(32,))
    rays = _rays(100000, 0.0, 100.0, 21)
    go, gt, gd = g.trace_closest(rays)
    ro, rt, rd = ref.trace(rays, precision=32, mode=2)
    assert np.array_equal(go, ro) and np.array_equal(gt, rt) and np.array_equal(gd, rd)
    # the literal reference traversal (object BVH, mesh loops) on a subset
    sub = rays[:2000]
    r0 = ref.trace(sub, precision=32, mode=0)
    assert np.array_equal(go[:2000], r0[0]) and np.array_equal(gt[:2000], r0[1]) and np.array_equal(gd[:2000], r0[2])
    print(f"1M: {(go >= 0).sum()} hits of {len(rays)}; mesh triangle ids checked", flush=True)


def test_torch_device_entry_points_match_host():
    torch = pytest.importorskip("torch")
    desc = scenes.random_mesh_scene(20, 200, seed=5)
    g = cb.Scene.from_desc(desc)
    rays = _rays(20000, -12, 12, 31)
    host = g.trace_closest(rays)
    dev = g.trace_closest(torch.from_numpy(rays).cuda())
    torch.cuda.synchronize()
    for a, b in zip(host, dev):
        assert np.array_equal(a, b.cpu().numpy())
    dist = np.full(len(rays), 3.0, np.float32)
    vh = g.trace_visibility(rays, dist)
    vd = g.trace_visibility(torch.from_numpy(rays).cuda(), torch.from_numpy(dist).cuda())
    assert np.array_equal(vh, vd.cpu().numpy())


# --------------------------------------------------------------------------- integrators
def _camera(cam_args, w, h):
    return cb.Camera.look_at(cam_args["eye"], cam_args["target"], cam_args["up"], cam_args["vfov"], w, h)


def _render(scene, cam, w, h, spp, method=cb.RenderingMethod.PATHTRACER, depth=5, rr=False, calls=1):
    r = cb.PBRenderer(method)
    r.set_output_size(w, h)
    r.register_scene(scene)
    r.set_method(method)
    r.set_max_trace_depth(depth)
    r.set_russian_roulette(rr)
    for _ in range(calls):
        r.render(cam, spp // calls)
    return r


@pytest.mark.parametrize("rr", [False, True])
def test_pathtracer_same_seed_parity(rr):
    desc, ca = scenes.cornell_box()
    w = h = 64
    cam = _camera(ca, w, h)
    r = _render(cb.Scene.from_desc(desc), cam, w, h, 16, rr=rr)
    img = r.get_output_image()
    ref = oracle.OracleScene(desc).render(cam.as_dict(), w, h, 0, 16, depth=5, rr=rr, precision=32)
    e = rel_mse(img, ref)
    print(f"PT rr={rr}: rel-MSE vs same-seed FP32 oracle {e:.2e}; exact pixels {np.mean(img == ref):.4f}", flush=True)
    assert e <= 1e-6


@pytest.mark.parametrize("w,h", [(1, 1), (3, 5), (8, 8), (13, 7)])
def test_pathtracer_tiny_images(w, h):
    """Queues of 1 to 91 rays: most of the trace kernel's 8 work groups are empty and the last
    64-ray chunk is partial; the image must still match the same-seed oracle."""
    desc, ca = scenes.cornell_box()
    cam = _camera(ca, w, h)
    r = _render(cb.Scene.from_desc(desc), cam, w, h, 8)
    img = r.get_output_image()
    ref = oracle.OracleScene(desc).render(cam.as_dict(), w, h, 0, 8, depth=5, precision=32)
    assert img.shape == (h, w, 3) and np.isfinite(img).all()
    assert rel_mse(img, ref) <= 1e-6


def test_pathtracer_mixed_scene_with_environment():
    desc, ca = scenes.mixed_primitives(env=True)
    w = h = 48
    cam = _camera(ca, w, h)
    img = _render(cb.Scene.from_desc(desc), cam, w, h, 8, depth=6).get_output_image()
    ref = oracle.OracleScene(desc).render(cam.as_dict(), w, h, 0, 8, depth=6, precision=32)
    assert np.isfinite(img).all()
    assert rel_mse(img, ref) <= 1e-6


@pytest.mark.parametrize("make,depth", [(lambda: scenes.cornell_box(), 5), (lambda: scenes.mixed_primitives(env=True), 6)])
def test_pathtracer_pcg_same_seed_parity(make, depth):
    """The PCG generator (an addition): the GPU against the oracle in its PCG mode with the same
    seeds, and an image different from the reference generator's."""
    desc, ca = make()
    w = h = 48
    cam = _camera(ca, w, h)
    scene = cb.Scene.from_desc(desc)
    r = cb.PBRenderer(cb.RenderingMethod.PATHTRACER)
    r.set_output_size(w, h)
    r.register_scene(scene)
    r.set_max_trace_depth(depth)
    r.set_rng(cb.RNG_PCG)
    r.render(cam, 16)
    img = r.get_output_image()
    o = oracle.OracleScene(desc)
    oracle.set_rng(True)
    try:
        ref = o.render(cam.as_dict(), w, h, 0, 16, depth=depth, precision=32)
    finally:
        oracle.set_rng(False)
    e = rel_mse(img, ref)
    print(f"PT (PCG): rel-MSE vs same-seed FP32 oracle {e:.2e}; exact pixels {np.mean(img == ref):.4f}", flush=True)
    assert e <= 1e-6
    assert not np.array_equal(img, _render(scene, cam, w, h, 16, depth=depth).get_output_image())


def test_pathtracer_pcg_estimates_the_same_image():
    """Both generators estimate one image: on the Cornell box with the light lowered to y = 0.9
    (at 0.999 rare direct-light samples of order 1e5 dominate any two-render comparison), 256 PCG
    frames differ from 256 reference frames by no more than two independent sets of 256 reference
    frames (frames 0..255 and 256..511) differ from each other, up to a factor 2."""
    desc, ca = scenes.cornell_box()
    desc["lights"][0, 2] = 0.9
    w = h = 48
    cam = _camera(ca, w, h)
    scene = cb.Scene.from_desc(desc)

    def frames(n, pcg):
        r = cb.PBRenderer(cb.RenderingMethod.PATHTRACER)
        r.set_output_size(w, h)
        r.register_scene(scene)
        r.set_max_trace_depth(5)
        if pcg:
            r.set_rng(cb.RNG_PCG)
        r.render(cam, n)
        return r.get_output_image().astype(np.float64)

    b1 = frames(256, False)
    b2 = 2.0 * frames(512, False) - b1  # the mean of frames 256..511
    pa = frames(256, True)
    e_pcg, e_ref = rel_mse(pa, b1), rel_mse(b2, b1)
    print(f"256 frames: PCG vs reference generator rel-MSE {e_pcg:.3e}; two reference halves {e_ref:.3e}", flush=True)
    assert e_pcg <= 2.0 * e_ref


def test_pathtracer_against_fp64_oracle():
    """rel-MSE <= 1e-4 against the FP64 oracle with the same seeds, and a per-pixel 3-sigma test on
    independent seeds: GPU frames 0..4095 against oracle frames 4096..8191.

    The two-sided area light sits 0.001 below the Cornell ceiling, so the NEE estimate on the
    ceiling above it has rare samples of order 1e5: a variance taken from one side only misses
    them. Each side therefore brings its own variance of the mean (Welch): the GPU's from 128
    batch means of 32 frames (recovered from the running mean after each call), the oracle's
    from its per-sample second moments. At most 0.5% of pixel channels may fall outside 3 sigma
    (0.27% for a normal statistic)."""
    desc, ca = scenes.cornell_box()
    w = h = 48
    n = 256
    cam = _camera(ca, w, h)
    scene = cb.Scene.from_desc(desc)
    ref = oracle.OracleScene(desc, precisions=(64,))
    img = _render(scene, cam, w, h, n).get_output_image()
    r64 = ref.render(cam.as_dict(), w, h, 0, n, depth=5, precision=64)
    e = rel_mse(img, r64)

    batches, m = 128, 32
    r = _render(scene, cam, w, h, m)
    means = [r.get_output_image().astype(np.float64)]
    for _ in range(batches - 1):
        r.render(cam, m)
        means.append(r.get_output_image().astype(np.float64))
    assert r.frame_index() == batches * m
    b = np.stack([(k + 1) * means[k] - k * means[k - 1] if k else means[0] for k in range(batches)])
    mg, vg = b.mean(0), b.var(0, ddof=1) / batches
    no = batches * m
    _, s, s2 = ref.render(cam.as_dict(), w, h, no, no, depth=5, precision=64, stats=True)
    mo = s / no
    vo = np.maximum(s2 / no - mo ** 2, 0) / (no - 1)
    z = np.abs(mg - mo) / np.sqrt(vg + vo + 1e-30)
    frac = float(np.mean(z > 3))
    print(f"PT vs FP64: rel-MSE {e:.2e}; 3-sigma outliers {frac:.4f}", flush=True)
    assert e <= 1e-4
    assert frac <= 0.005


def test_running_mean_over_calls_equals_one_call():
    """render(Camera*) per frame (PBRenderer.cu:172-186): two calls of 4 spp equal one call of 8."""
    desc, ca = scenes.cornell_box()
    cam = _camera(ca, 40, 30)
    scene = cb.Scene.from_desc(desc)
    a = _render(scene, cam, 40, 30, 8, calls=2)
    b = _render(scene, cam, 40, 30, 8, calls=1)
    assert a.frame_index() == 8
    assert np.array_equal(a.get_output_image(), b.get_output_image())
    cam.set_moved(True)  # Camera::moved() restarts the mean (PBRenderer.cu:174-177)
    a.render(cam, 1)
    assert a.frame_index() == 1


def test_batches_of_frames_equal_frame_by_frame_calls():
    """One 40-spp call at 2048x2048 (168M paths) runs as batches of 32 and 8 frames (at most 128M
    paths in flight); its running mean must equal forty one-frame calls bit for bit (frames are
    accumulated in order across batches, PathTracer.cu:162-168)."""
    desc, ca = scenes.cornell_box()
    w = h = 2048
    cam = _camera(ca, w, h)
    scene = cb.Scene.from_desc(desc)
    a = _render(scene, cam, w, h, 40, calls=1)
    ia = a.get_output_image()
    del a
    b = _render(scene, cam, w, h, 40, calls=40)
    assert b.frame_index() == 40
    ib = b.get_output_image()
    assert np.isfinite(ia).all() and float(ia.mean()) > 0
    assert np.array_equal(ia, ib)


@pytest.mark.parametrize("name,w,h,spp,refills", [("mixed_env", 67, 45, 4, False), ("million", 203, 117, 4, False),
                                                  ("mixed_env", 1536, 1024, 4, True), ("million", 640, 480, 16, True)])
def test_fast_paths_equal_per_ray_traversal(name, w, h, spp, refills):
    """Bounce 0 traces each 64-ray chunk of camera rays as one packet (wave-uniform node and
    primitive loads, a child entered when any lane's interval overlaps it), and launches of at
    least 4 rays per lane of the persistent trace grid (about 2 M rays; the last two cases) hand
    fresh rays to idle lanes as soon as half of a wave is idle. The visit-counting mode traces every
    ray one per lane in 64-ray chunks. The closest hit is a minimum over candidates and visibility
    an "or" over them, independent of which lane traces a ray when, so the images must agree bit
    for bit, also where the last chunk is partial (odd image sizes)."""
    desc, ca = scenes.million_triangles() if name == "million" else scenes.mixed_primitives(env=True)
    scene = cb.Scene.from_desc(desc)
    cam = _camera(ca, w, h)
    imgs = []
    for counting in (False, True):
        r = cb.PBRenderer(cb.RenderingMethod.PATHTRACER)
        r.set_output_size(w, h)
        r.register_scene(scene)
        r.set_max_trace_depth(5)
        r.set_traversal_counting(counting)
        r.render(cam, spp)
        imgs.append(r.get_output_image())
        st = r.stats()
        if counting:
            rays = st["extension_rays"]
            assert st["refill_launches"] == 0
        else:
            refill_launches = st["refill_launches"]
    print(f"{name} {w}x{h} {spp} spp: {rays} extension rays, {refill_launches} refill launches", flush=True)
    # the large cases must exercise the refill branch (launches of >= 4 rays per lane), the small
    # ones the 64-ray chunks
    assert (refill_launches > 0) == refills
    assert np.isfinite(imgs[0]).all() and float(imgs[0].mean()) > 0
    assert np.array_equal(imgs[0], imgs[1])


def test_furnace_on_gpu():
    """Closed emissive box, no lights: each pixel is a partial sum E * sum_k rho'^k, as on the CPU."""
    b = scenes.SceneBuilder()
    m = scenes.material(scenes.LAMBERT, albedo_e=(0.5,) * 3, albedo_d=(0.6,) * 3)
    for t in scenes.box_triangles((0, 0, 0), (40, 40, 40)):
        v = t.reshape(3, 3)
        b.triangle(v[0], v[2], v[1], m)
    desc = b.desc()
    cam = cb.Camera(24, 24)
    img = _render(cb.Scene.from_desc(desc), cam, 24, 24, 1).get_output_image()
    rp = float(np.float32(0.6)) * float(np.float32(3.14159)) / float(np.float32(math.pi))
    full = 0.5 * sum(rp ** k for k in range(5))
    assert np.mean(np.abs(img - full) < 2e-6) > 0.99


def _volume_scene():
    return scenes.volume_spheres()


def test_volume_renderer_same_seed_parity():
    desc = _volume_scene()
    w = h = 48
    cam = cb.Camera(w, h)
    img = _render(cb.Scene.from_desc(desc), cam, w, h, 8, method=cb.RenderingMethod.VOLUME, depth=6).get_output_image()
    ref = oracle.OracleScene(desc).render(cam.as_dict(), w, h, 0, 8, depth=6, method=oracle.VOLUME, precision=32)
    e = rel_mse(img, ref)
    print(f"volume: rel-MSE vs same-seed FP32 oracle {e:.2e}", flush=True)
    assert np.isfinite(img).all() and e <= 1e-6


def test_gradient_domain_same_seed_parity():
    desc, ca = scenes.cornell_box()
    w = h = 40
    cam = _camera(ca, w, h)
    r = _render(cb.Scene.from_desc(desc), cam, w, h, 4, method=cb.RenderingMethod.GRADIENTDOMAIN)
    ref = oracle.OracleScene(desc)
    bufs = ref.gdpt(cam.as_dict(), w, h, 0, 4, depth=5, precision=32)
    names = [cb.Image.BASE, cb.Image.GRADIENT_X_FORWARD, cb.Image.GRADIENT_X_BACKWARD, cb.Image.GRADIENT_Y_FORWARD,
             cb.Image.GRADIENT_Y_BACKWARD]
    for k, im in enumerate(names):
        e = float(np.mean((r.get_image(im) - bufs[k]) ** 2)) / (float(np.mean(bufs[0] ** 2)) + 1e-2)
        assert e <= 1e-6, (im, e)
    # the reconstruction on the GPU's own five buffers is deterministic FP32 work: compare it
    # elementwise; against the oracle's buffers, rare path divergences leave only a tiny rel-MSE
    own = np.stack([r.get_image(im) for im in names]).astype(np.float32)
    recon_own, gx, gy = oracle.gdpt_reconstruct(own, 50)
    np.testing.assert_allclose(r.get_gradient_x(), gx, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(r.get_gradient_y(), gy, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(r.get_output_image(), recon_own, rtol=1e-5, atol=1e-6)
    recon, _, _ = oracle.gdpt_reconstruct(bufs.astype(np.float32), 50)
    assert rel_mse(r.get_output_image(), recon) <= 1e-6


@pytest.mark.parametrize("method", ["volume", "gdpt"])
@pytest.mark.parametrize("w,h", [(67, 45), (64, 48)])
def test_camera_kernels_equal_separate_ray_generation(method, w, h):
    """With camera packets the volume renderer and the gradient domain make their camera rays in
    the camera trace kernel, write the paths' state there and store one origin (queue slot 0) for
    the first shade; the visit-counting mode writes every ray to the queue in a separate
    ray-generation pass and traces it one per lane. Each ray comes from the same code either way
    and the closest hit is a minimum over candidates, so every buffer must agree bit for bit (odd
    sizes leave a partial last chunk; at 64x48 a wave holds 4 frames of 16 pixels)."""
    if method == "volume":
        desc, cam, m, depth = _volume_scene(), cb.Camera(w, h), cb.RenderingMethod.VOLUME, 6
        names = [cb.Image.OUTPUT]
    else:
        desc, ca = scenes.cornell_box()
        cam, m, depth = _camera(ca, w, h), cb.RenderingMethod.GRADIENTDOMAIN, 5
        names = [cb.Image.BASE, cb.Image.GRADIENT_X_FORWARD, cb.Image.GRADIENT_X_BACKWARD,
                 cb.Image.GRADIENT_Y_FORWARD, cb.Image.GRADIENT_Y_BACKWARD, cb.Image.OUTPUT]
    scene = cb.Scene.from_desc(desc)
    out = []
    for counting in (False, True):
        r = cb.PBRenderer(m)
        r.set_output_size(w, h)
        r.register_scene(scene)
        r.set_method(m)
        r.set_max_trace_depth(depth)
        r.set_traversal_counting(counting)
        r.render(cam, 4)
        out.append([r.get_image(im) for im in names])
    for im, a, b in zip(names, out[0], out[1]):
        assert np.isfinite(a).all() and float(np.abs(a).mean()) > 0, im
        assert np.array_equal(a, b), (im, float(np.abs(a - b).max()))


@pytest.mark.parametrize("alpha,iters", [(0.2, 50), (0.5, 7)])
def test_gradient_domain_screened_poisson_reconstruction(alpha, iters):
    """The screened-Poisson (L2) mode on the GPU's own five buffers equals the oracle's Jacobi
    iterations (deterministic FP32; an odd count ends in the ping-pong buffer and is copied to the
    output); the buffers themselves are those of the trimmed-mean mode."""
    desc, ca = scenes.cornell_box()
    w, h = 40, 32
    cam = _camera(ca, w, h)
    scene = cb.Scene.from_desc(desc)
    r = cb.PBRenderer(cb.RenderingMethod.GRADIENTDOMAIN)
    r.set_output_size(w, h)
    r.register_scene(scene)
    r.set_method(cb.RenderingMethod.GRADIENTDOMAIN)
    r.set_gdpt_reconstruction(cb.GDPT_SCREENED_POISSON, alpha, iters)
    r.render(cam, 4)
    names = [cb.Image.BASE, cb.Image.GRADIENT_X_FORWARD, cb.Image.GRADIENT_X_BACKWARD, cb.Image.GRADIENT_Y_FORWARD,
             cb.Image.GRADIENT_Y_BACKWARD]
    own = np.stack([r.get_image(im) for im in names]).astype(np.float32)
    want = oracle.gdpt_reconstruct_l2(own, alpha, iters)
    got = r.get_output_image()
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)
    assert (got[0] == 0).all() and (got[:, 0] == 0).all()
    t = _render(scene, cam, w, h, 4, method=cb.RenderingMethod.GRADIENTDOMAIN)
    for im in names:
        assert np.array_equal(t.get_image(im), r.get_image(im))


def test_raytracer_point_light_closed_form():
    """raytracing_kernel (RayTracer.cu:11-54): brdf * I / d^2 * cos for every light, shadowed."""
    b = scenes.SceneBuilder()
    b.plane((0, -1, 0), (0, 1, 0), scenes.material(scenes.LAMBERT, albedo_d=(0.5, 0.5, 0.5)))
    b.point_light((0, 1, 3), (2, 2, 2))
    desc = b.desc()
    w = h = 16
    cam = cb.Camera(w, h)
    img = _render(cb.Scene.from_desc(desc), cam, w, h, 1, method=cb.RenderingMethod.RAYTRACER).get_output_image()
    ref = oracle.OracleScene(desc, precisions=(64,))
    rays = []
    for py in range(h):
        for px in range(w):
            rx, ry = 2 * (px / w - 0.5), 2 * (py / h - 0.5)
            rays.append([0, 0, 0, *_unit([rx, ry, 1.0])])
    obj, _, _, P, _ = ref.trace(np.array(rays, np.float32), precision=64, mode=0, want_geometry=True)
    L = np.array([0, 1, 3.0])
    d = np.linalg.norm(L - P, axis=1)
    cos = np.maximum(0, (L - P)[:, 1] / d)
    want = np.where(obj >= 0, float(np.float32(0.5)) / math.pi * 2 / d ** 2 * cos, 0).reshape(h, w)
    np.testing.assert_allclose(img[..., 0], want, rtol=1e-4, atol=1e-6)


def test_xml_scene_renders_like_the_same_scene_from_arrays(tmp_path):
    """cupbr_scene_load_xml and cupbr_scene_create give the same closest hits and image."""
    desc, ca = scenes.cornell_box()
    lines = ["<header><scene_size>34</scene_size><light_count>1</light_count></header><geometry>"]
    for i in range(34):
        p = desc["params"][i]
        m = desc["mats"][i]
        lines.append(f"<geometry{i + 1}><type>TRIANGLE</type>" +
                     "".join(f"<vertex{k + 1}>{','.join(repr(float(x)) for x in p[3 * k:3 * k + 3])}</vertex{k + 1}>"
                             for k in range(3)) +
                     f"<material><type>LAMBERT</type><albedo_d>{','.join(repr(float(x)) for x in m[3:6])}</albedo_d>"
                     f"</material></geometry{i + 1}>")
    L = desc["lights"][0]
    lines.append("</geometry><lights><light1><type>AREA</type>"
                 f"<position>{','.join(repr(float(x)) for x in L[1:4])}</position>"
                 f"<radiance>{','.join(repr(float(x)) for x in L[7:10])}</radiance>"
                 f"<extend1>{','.join(repr(float(x)) for x in L[10:13])}</extend1>"
                 f"<extend2>{','.join(repr(float(x)) for x in L[13:16])}</extend2></light1></lights>")
    (tmp_path / "cornell.xml").write_text("\n".join(lines))
    a = cb.SceneLoader.load_from_file(str(tmp_path / "cornell.xml"))
    b = cb.Scene.from_desc(desc)
    cam = _camera(ca, 32, 32)
    ia = _render(a, cam, 32, 32, 4).get_output_image()
    ib = _render(b, cam, 32, 32, 4).get_output_image()
    assert np.array_equal(ia, ib)


# --------------------------------------------------------------------------- post chain
def test_bloom_and_tonemap_parity():
    rng = np.random.default_rng(40)
    for (h, w) in [(37, 53), (64, 64), (270, 480)]:
        hdr = (rng.exponential(0.7, (h, w, 3)) * rng.choice([0.2, 1.0, 6.0], (h, w, 1))).astype(np.float32)
        import torch
        t = torch.from_numpy(hdr).cuda()
        torch.cuda.synchronize()
        post = cb.PostProcessor()
        post.register_image(t.data_ptr(), w, h)
        curve = cb.threshold_curve(1.0, 0.5)
        post.bloom(curve)
        out = post.get_post_process_buffer()
        want = oracle.bloom(hdr, curve)
        assert np.array_equal(out, want), (h, w, float(np.abs(out - want).max()))
        for kind in (cb.ToneMappingType.REINHARD, cb.ToneMappingType.GAMMA, cb.ToneMappingType.ACES):
            tm = cb.ToneMapper(kind)
            tm.set_exposure(0.8)
            tm.register_image(post.buffer_device_ptr(), w, h)
            tm.tone_map()
            got = tm.get_render_buffer().astype(np.int32)
            ref = oracle.tonemap(want, int(kind), 0.8).astype(np.int32)
            assert np.abs(got - ref).max() <= 1 and np.mean(got != ref) < 1e-3
            # fused bloom + tone map writes the same buffers
            tm2 = cb.ToneMapper(kind)
            tm2.set_exposure(0.8)
            tm2.register_image(t.data_ptr(), w, h)
            tm2.tone_map_bloom(post, curve)
            assert np.array_equal(tm2.get_render_buffer(), tm.get_render_buffer())
            assert np.array_equal(post.get_post_process_buffer(), want)


def test_gaussian_bloom_parity():
    """The Gaussian-pyramid bloom mode (an addition) against the oracle, bit for bit, at sizes with
    odd and even levels; the fused tone map writes the same buffers."""
    import torch
    rng = np.random.default_rng(41)
    for (h, w) in [(2, 2), (37, 53), (64, 64), (270, 480), (1080, 1920)]:
        hdr = (rng.exponential(0.7, (h, w, 3)) * rng.choice([0.2, 1.0, 6.0], (h, w, 1))).astype(np.float32)
        t = torch.from_numpy(hdr).cuda()
        torch.cuda.synchronize()
        post = cb.PostProcessor()
        post.set_bloom_pyramid(cb.BloomPyramid.GAUSSIAN5)
        post.register_image(t.data_ptr(), w, h)
        curve = cb.threshold_curve(1.0, 0.5)
        post.bloom(curve)
        want = oracle.bloom_gaussian(hdr, curve)
        out = post.get_post_process_buffer()
        assert np.array_equal(out, want), (h, w, float(np.abs(out - want).max()))
        tm = cb.ToneMapper(cb.ToneMappingType.REINHARD)
        tm.register_image(post.buffer_device_ptr(), w, h)
        tm.tone_map()
        tm2 = cb.ToneMapper(cb.ToneMappingType.REINHARD)
        tm2.register_image(t.data_ptr(), w, h)
        tm2.tone_map_bloom(post, curve)
        assert np.array_equal(tm2.get_render_buffer(), tm.get_render_buffer())
        post.set_bloom_pyramid(cb.BloomPyramid.REFERENCE)  # and back to the reference's pyramid
        post.bloom(curve)
        assert np.array_equal(post.get_post_process_buffer(), oracle.bloom(hdr, curve))


def test_save_to_file_writes_a_bottom_up_bmp(tmp_path):
    import torch
    hdr = np.zeros((4, 6, 3), np.float32)
    hdr[0, :, 0] = 10.0  # bottom row red
    t = torch.from_numpy(hdr).cuda()
    torch.cuda.synchronize()
    tm = cb.ToneMapper(cb.ToneMappingType.GAMMA)
    tm.register_image(t.data_ptr(), 6, 4)
    tm.tone_map()
    tm.save_to_file(str(tmp_path / "o.bmp"))
    data = (tmp_path / "o.bmp").read_bytes()
    assert data[:2] == b"BM" and int.from_bytes(data[10:14], "little") == 54
    first_row = np.frombuffer(data[54:54 + 24], np.uint8).reshape(6, 4)
    assert (first_row[:, 2] == 255).all() and (first_row[:, 0] == 0).all()  # BGRA, row 0 first


# --------------------------------------------------------------------------- error behaviour
def test_error_paths():
    desc, ca = scenes.cornell_box()
    cam = _camera(ca, 8, 8)
    r = cb.PBRenderer(cb.RenderingMethod.PATHTRACER)
    with pytest.raises(cb.CupbrError) as e:
        r.render(cam, 1)
    assert e.value.code == -2  # PBRenderer.cu:140-144
    r.set_output_size(8, 8)
    with pytest.raises(cb.CupbrError) as e:
        r.render(cam, 1)
    assert e.value.code == -3  # :146-150
    r.register_scene(cb.Scene.from_desc(desc))
    for m in (cb.RenderingMethod.WHITTED, cb.RenderingMethod.METROPOLIS):
        r.set_method(m)
        with pytest.raises(cb.CupbrError) as e:
            r.render(cam, 1)
        assert e.value.code == -10
    r.set_method(cb.RenderingMethod.GRADIENTDOMAIN)
    r.set_max_trace_depth(11)
    with pytest.raises(cb.CupbrError):
        r.render(cam, 1)
    tm = cb.ToneMapper()
    with pytest.raises(cb.CupbrError) as e:
        tm.tone_map()
    assert e.value.code == -9  # ToneMapper.cu:154-157


def test_cpp_class_interface_renders_like_the_c_abi(tmp_path):
    """tests/cpp/render_example.cpp, written against the reference's class interface
    (include/cupbr.hpp), renders 16 frames, blooms, tone-maps and saves a BMP; its HDR mean
    equals the same render driven through the Python binding."""
    import subprocess

    from test_loader_abi import build_cpp_example

    exe, xml = build_cpp_example(tmp_path)
    bmp = tmp_path / "o.bmp"
    out = subprocess.run([str(exe), str(xml), str(bmp)], capture_output=True, text=True, timeout=120)
    print(out.stdout, flush=True)
    assert out.returncode == 0, out.stderr
    mean = float(out.stdout.strip().split("mean ")[1].split(",")[0])
    scene = cb.SceneLoader.load_from_file(str(xml))
    r = _render(scene, cb.Camera(96, 64), 96, 64, 16, depth=5)
    assert abs(r.get_output_image().astype(np.float64).mean() - mean) <= 1e-6 * max(1.0, abs(mean))
    data = bmp.read_bytes()
    assert data[:2] == b"BM" and len(data) == 54 + 96 * 64 * 4


def test_pfm_dump_and_resume_continue_the_running_mean(tmp_path):
    """save_pfm writes the exact FP32 image; resume(image, 8) followed by 8 frames equals one
    16-frame render (the running mean continues by frame index, PBRenderer.cu:172-186)."""
    desc, ca = scenes.cornell_box()
    cam = _camera(ca, 48, 32)
    scene = cb.Scene.from_desc(desc)
    a = _render(scene, cam, 48, 32, 8)
    path = tmp_path / "acc.pfm"
    a.save_pfm(str(path))
    saved = cb.read_pfm(str(path))
    assert np.array_equal(saved, a.get_output_image())
    b = _render(scene, cam, 48, 32, 0)
    b.resume(saved, 8)
    b.render(cam, 8)
    assert b.frame_index() == 16
    assert np.array_equal(b.get_output_image(), _render(scene, cam, 48, 32, 16).get_output_image())


def test_pick_and_update_material():
    """pickMouse returns the object the unjittered centre ray of a pixel hits (the same ray through
    trace_closest), and updateMaterial changes what the renderer sees."""
    desc, ca = scenes.cornell_box()
    w, h = 64, 64
    cam = _camera(ca, w, h)
    scene = cb.Scene.from_desc(desc)
    picks = 0
    for (x, y) in [(32, 32), (5, 40), (60, 10), (32, 60), (20, 5)]:
        got = scene.pick(cam, x, y, w, h)
        rx, ry = 2.0 * (x / w - 0.5), 2.0 * (y / h - 0.5)
        d = cam.z_axis + rx * cam.x_axis + ry * cam.y_axis
        d = d / np.linalg.norm(d)
        obj, _, t = scene.trace_closest(np.concatenate([cam.position, d])[None].astype(np.float32))
        if obj[0] < 0:
            assert got is None
            continue
        picks += 1
        assert got[0] == obj[0]
        assert np.allclose(got[1]["albedo_d"], desc["mats"][obj[0], 3:6])
    assert picks >= 4
    before = _render(scene, cam, w, h, 4).get_output_image()
    idx, mat = scene.pick(cam, 32, 32, w, h)
    mat["albedo_e"] = (2.0, 2.0, 2.0)
    scene.update_material(idx, mat)
    assert scene.pick(cam, 32, 32, w, h)[1]["albedo_e"].tolist() == [2.0, 2.0, 2.0]
    after = _render(scene, cam, w, h, 4).get_output_image()
    assert after.mean() > before.mean()
