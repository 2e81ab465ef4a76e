"""GPU tests of the neural radiance cache: the matrix-core fragment layout, inference and training
against a plain PyTorch FP32 reference of the same network (with the kernel's BF16 roundings
emulated), and the cache inside the path tracer."""
import math
import time

import numpy as np
import pytest

import cupbr_amd as cb
from cupbr_amd import scenes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

PI = np.float32(3.14159265358979)


def bf16(x):
    return x.to(torch.bfloat16).to(torch.float32)


def ste(x):  # BF16 rounding in the forward pass, identity in the backward pass
    return x + (bf16(x) - x).detach()


def encode(raw, lo, hi):
    """The cache's 64-feature encoding (cupbr_amd/csrc/nrc.hip) in FP32."""
    lo = torch.tensor(lo, dtype=torch.float32, device=raw.device)
    inv = 1.0 / (torch.tensor(hi, dtype=torch.float32, device=raw.device) - lo)
    d = torch.zeros(raw.shape[0], 14, device=raw.device)
    d[:, 0:3] = (raw[:, 0:3] - lo) * inv
    d[:, 3] = torch.acos(raw[:, 5].clamp(-1, 1)) / PI
    d[:, 4] = (torch.atan2(raw[:, 4], raw[:, 3]) / PI + 1.0) * 0.5
    d[:, 5] = torch.acos(raw[:, 8].clamp(-1, 1)) / PI
    d[:, 6] = (torch.atan2(raw[:, 7], raw[:, 6]) / PI + 1.0) * 0.5
    d[:, 7] = raw[:, 9]
    d[:, 8:14] = raw[:, 10:16]
    f = torch.zeros(raw.shape[0], 64, device=raw.device)
    for i in range(52):
        if i < 36:
            a, k, sc = i // 12, (i % 12) // 2, i % 2
        else:
            r = (i - 36) & 7
            a, k, sc = (3 if i < 44 else 5) + (r >> 2), (r & 3) >> 1, r & 1
        arg = (float(2 ** k) * d[:, a]).double() * math.pi  # sin/cos(2^k pi x), evaluated in FP64
        f[:, i] = (torch.cos(arg) if sc else torch.sin(arg)).float()
    f[:, 52] = 1.0 - torch.exp(-d[:, 7])
    f[:, 53:59] = d[:, 8:14]
    f[:, 59] = 1.0
    return f


def mlp(w, f, emulate_bf16=True):
    w = torch.as_tensor(w, device=f.device)
    Ws = [w[k * 4096:(k + 1) * 4096].view(64, 64) for k in range(4)] + [w[16384:].view(3, 64)]
    r = ste if emulate_bf16 else (lambda x: x)
    h = r(f)
    for W in Ws[:4]:
        h = r(torch.relu(h @ r(W).T))
    return h @ r(Ws[4]).T


def random_inputs(n, rng, lo=(0, 0, 0), hi=(1, 1, 1)):
    x = np.zeros((n, 16), np.float32)
    x[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    for s in (3, 6):
        v = rng.normal(size=(n, 3))
        x[:, s:s + 3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    x[:, 9] = rng.uniform(0, 1, n)
    x[:, 10:16] = rng.uniform(0, 1, (n, 6))
    return x


def test_matrix_core_fragment_layout():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(16, 32)).astype(np.float32)
    b = rng.normal(size=(32, 16)).astype(np.float32)
    d = cb.RadianceCache.matrix_core_product(a, b)
    ab = bf16(torch.from_numpy(a)).double().numpy()
    bb = bf16(torch.from_numpy(b)).double().numpy()
    ref = ab @ bb
    err = np.abs(d - ref).max()
    print(f"matrix-core product: max abs error {err:.2e}", flush=True)
    assert err <= 1e-4 * np.abs(ref).max()


def test_inference_matches_torch_reference():
    rng = np.random.default_rng(1)
    cache = cb.RadianceCache((-1, -1, -1), (2, 2, 2), seed=3)
    w = cache.get_weights()
    assert np.isfinite(w).all() and abs(w[:16384].std() - math.sqrt(2 / 64)) < 0.02
    w = (rng.normal(size=w.shape) * 0.2).astype(np.float32)  # weights that keep activations O(1)
    cache.set_weights(w)
    np.testing.assert_array_equal(cache.get_weights(), w)
    x = random_inputs(1000, rng, (-1, -1, -1), (2, 2, 2))  # not a multiple of 16
    out = cache.infer(torch.from_numpy(x).cuda()).cpu().numpy()
    f = encode(torch.from_numpy(x).cuda(), (-1, -1, -1), (2, 2, 2))
    ref_bf = mlp(torch.from_numpy(w).cuda(), f).cpu().numpy()
    ref_32 = mlp(torch.from_numpy(w).cuda(), f, emulate_bf16=False).cpu().numpy()
    e_bf = np.linalg.norm(out - ref_bf) / np.linalg.norm(ref_bf)
    e_32 = np.linalg.norm(out - ref_32) / np.linalg.norm(ref_32)
    print(f"inference: rel-L2 vs BF16-emulating reference {e_bf:.2e}, vs FP32 reference {e_32:.2e}", flush=True)
    assert e_bf <= 1e-3 and e_32 <= 5e-2


def _hash32(x):
    x = np.asarray(x, np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7feb352d) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846ca68b) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def test_training_step_matches_torch_gradient():
    """One Adam step from fresh moments moves each weight by -lr * sign(gradient) (for gradients
    well above eps); the batch is the rows the kernel draws: rows[hash(seed + q) % n]."""
    rng = np.random.default_rng(2)
    cache = cb.RadianceCache((0, 0, 0), (1, 1, 1), seed=4)
    w0 = cache.get_weights()
    n, B, lr = 4096, 1024, 1e-3
    x = random_inputs(n, rng)
    rows = np.zeros((n, 20), np.float32)
    rows[:, :16] = x
    rows[:, 16:19] = 0.5 + 0.3 * np.sin(3 * x[:, 0:3])
    loss = cache.train(torch.from_numpy(rows).cuda(), batch=B, steps=1, lr=lr, return_loss=True)
    w1 = cache.get_weights()
    seed = (1 * 0x9E3779B9 + 0x51ED270B) & 0xFFFFFFFF
    idx = (_hash32(seed + np.arange(B, dtype=np.uint64)) % n).astype(np.int64)
    xb = torch.from_numpy(rows[idx]).cuda()
    W = torch.from_numpy(w0).cuda().requires_grad_(True)
    o = mlp(W, encode(xb[:, :16], (0, 0, 0), (1, 1, 1)))
    l_ref = ((o - xb[:, 16:19]) ** 2 / (o.detach() ** 2 + 0.01)).sum()
    (l_ref / B).backward()
    g = W.grad.cpu().numpy()
    step = w1 - w0
    big = np.abs(g) > 1e-2 * np.abs(g).max()
    agree = np.mean(np.sign(step[big]) == -np.sign(g[big]))
    lr_ok = np.mean(np.abs(np.abs(step[big]) - lr) <= 0.05 * lr)
    print(f"train step: loss {float(loss[0]):.4f} vs reference {float(l_ref.detach()):.4f}; sign agreement {agree:.4f} "
          f"over {big.sum()} weights; |step| = lr for {lr_ok:.4f}", flush=True)
    assert abs(float(loss[0]) - float(l_ref.detach())) <= 1e-2 * float(l_ref.detach())
    assert agree >= 0.98 and lr_ok >= 0.98


def test_training_fits_a_smooth_function():
    rng = np.random.default_rng(3)
    cache = cb.RadianceCache((0, 0, 0), (1, 1, 1), seed=5)
    n = 1 << 15
    x = random_inputs(n, rng)
    rows = np.zeros((n, 20), np.float32)
    rows[:, :16] = x
    rows[:, 16:19] = 0.5 + 0.3 * np.sin(4 * x[:, 0:3]) * x[:, 10:13]
    r = torch.from_numpy(rows).cuda()
    steps, lr = 400, 1e-2
    losses = cache.train(r, batch=4096, steps=steps, lr=lr, return_loss=True).cpu().numpy() / 4096
    print(f"fit: relative L2 loss {losses[0]:.4f} -> {losses[-20:].mean():.5f}", flush=True)
    assert losses[-20:].mean() < 0.02 * losses[0]
    pred = cache.infer(r[:, :16]).cpu().numpy()
    y = rows[:, 16:19]
    rel = np.sqrt(np.mean((pred - y) ** 2)) / np.sqrt(np.mean(y ** 2))
    rel_const = np.sqrt(np.mean((y.mean(0) - y) ** 2)) / np.sqrt(np.mean(y ** 2))  # best constant predictor
    # the same optimisation in PyTorch FP32 (Adam, same loss, batch and step count)
    f_all = encode(r[:, :16], (0, 0, 0), (1, 1, 1))
    g = torch.Generator(device="cuda").manual_seed(0)
    W = torch.from_numpy(cb.RadianceCache((0, 0, 0), (1, 1, 1), seed=5).get_weights()).cuda().requires_grad_(True)
    opt = torch.optim.Adam([W], lr=lr, betas=(0.9, 0.99), eps=1e-10)
    for _ in range(steps):
        idx = torch.randint(0, n, (4096,), device="cuda", generator=g)
        o = mlp(W, f_all[idx], emulate_bf16=False)
        loss = ((o - r[idx, 16:19]) ** 2 / (o.detach() ** 2 + 0.01)).sum() / 4096
        opt.zero_grad()
        loss.backward()
        opt.step()
    with torch.no_grad():
        p_ref = mlp(W, f_all, emulate_bf16=False).cpu().numpy()
    rel_ref = np.sqrt(np.mean((p_ref - y) ** 2)) / np.sqrt(np.mean(y ** 2))
    print(f"fit: rms relative error {rel:.4f}; PyTorch FP32 {rel_ref:.4f}; best constant {rel_const:.4f}", flush=True)
    assert rel < 0.5 * rel_const and rel < 1.5 * rel_ref


def _cornell(w, h):
    desc, ca = scenes.cornell_box()
    cam = cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], w, h)
    return cb.Scene.from_desc(desc), cam


def _pt(scene, cam, w, h, spp, depth=5, cache=None, **kw):
    r = cb.PBRenderer(cb.RenderingMethod.PATHTRACER)
    r.set_output_size(w, h)
    r.register_scene(scene)
    r.set_max_trace_depth(depth)
    if cache is not None:
        r.set_radiance_cache(cache, **kw)
    r.render(cam, spp)
    return r


def test_cache_paths_keep_the_estimator():
    """A zero cache ends every path at its second vertex: the image equals plain PT with depth 1.
    With every path a training path the image equals plain PT (same seeds) and training runs."""
    w = h = 64
    scene, cam = _cornell(w, h)
    cache = cb.RadianceCache((0, 0, 0), (1, 1, 1))
    cache.set_weights(np.zeros(cb.NRC_PARAMS, np.float32))
    a = _pt(scene, cam, w, h, 4, cache=cache, training_paths=0, train_steps=0).get_output_image()
    b = _pt(scene, cam, w, h, 4, depth=1).get_output_image()
    assert np.array_equal(a, b)
    cache2 = cb.RadianceCache((0, 0, 0), (1, 1, 1))
    w0 = cache2.get_weights()
    r = _pt(scene, cam, w, h, 4, cache=cache2, training_paths=w * h, train_steps=2, batch=512)
    st = r.stats()
    assert np.array_equal(r.get_output_image(), _pt(scene, cam, w, h, 4).get_output_image())
    assert st["cache_queries"] == 0 and st["cache_training_rows"] > 0
    assert not np.array_equal(cache2.get_weights(), w0)


def test_camera_hit_cache_vertices_keep_the_estimator():
    """With the camera hit a cache vertex, a zero cache leaves each pixel its camera hit's emission:
    on a Cornell box whose only light is an emissive ceiling (no light list, so next-event estimation
    adds nothing) and whose floor is GGX of roughness 0.3, the image equals plain PT with depth 1.
    With a threshold above 0.3 the floor is no camera-hit vertex: its pixels take what the second
    vertex emits, every other pixel stays. With every path a training path the image equals plain
    PT, no path queries the cache, and the camera hits add training rows."""
    w = h = 64
    desc, ca = scenes.cornell_box()
    desc["lights"] = desc["lights"][:0]
    emissive = np.array(desc["mats"], np.float32)
    emissive[2:4, 0:3] = 4.0  # the ceiling's two triangles
    emissive[0:2] = scenes.material(scenes.GGX, albedo_s=(0.8, 0.8, 0.8), roughness=0.3)  # the floor
    desc["mats"] = emissive
    scene = cb.Scene.from_desc(desc)
    cam = cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], w, h)

    def cached(min_roughness, **kw):
        cache = cb.RadianceCache((0, 0, 0), (1, 1, 1))
        cache.set_weights(np.zeros(cb.NRC_PARAMS, np.float32))
        r = cb.PBRenderer(cb.RenderingMethod.PATHTRACER)
        r.set_output_size(w, h)
        r.register_scene(scene)
        r.set_max_trace_depth(5)
        r.set_radiance_cache(cache, **kw)
        r.set_radiance_cache_camera_hit(min_roughness)
        r.render(cam, 4)
        return r

    depth1 = _pt(scene, cam, w, h, 4, depth=1).get_output_image()
    assert depth1.max() > 0 and np.count_nonzero(depth1) < depth1.size  # the ceiling seen, the rest black
    a = cached(0.0, training_paths=0, train_steps=0)
    assert np.array_equal(a.get_output_image(), depth1)
    assert 0 < a.stats()["cache_queries"] <= w * h * 4  # one per camera ray that hits the box
    floor = cached(0.5, training_paths=0, train_steps=0).get_output_image()
    changed = np.any(floor != depth1, axis=2)
    assert changed.any() and np.all(floor[changed] >= depth1[changed])
    assert np.count_nonzero(changed) < w * h // 2  # the floor's pixels only
    # Lambertian camera hits are vertices only for thresholds <= 1: above it the second vertex is the
    # first cache vertex, as without the option
    second = cached(float("inf"), training_paths=0, train_steps=0).get_output_image()
    assert np.array_equal(cached(1.5, training_paths=0, train_steps=0).get_output_image(), second)
    assert not np.array_equal(second, depth1)
    full = cached(0.0, training_paths=w * h, train_steps=1, batch=512)
    st = full.stats()
    assert np.array_equal(full.get_output_image(), _pt(scene, cam, w, h, 4).get_output_image())
    assert st["cache_queries"] == 0
    rows_second = cached(float("inf"), training_paths=w * h, train_steps=1, batch=512).stats()["cache_training_rows"]
    assert st["cache_training_rows"] > rows_second


def rel_mse(x, ref):
    return float(np.mean((x.astype(np.float64) - ref) ** 2 / (ref ** 2 + 1e-2)))


def test_cache_lowers_equal_time_error_on_the_cornell_box():
    """After online training, a 1-spp frame with the cache has a lower rel-MSE against a converged
    image than plain PT given the same GPU time. The light is lowered to y = 0.9: at 0.999 the
    two-sided light next to the ceiling makes direct-lighting fireflies dominate every estimator's
    error, which a cache of indirect light cannot change."""
    w = h = 256
    desc, ca = scenes.cornell_box()
    desc["lights"][0, 2] = 0.9
    scene = cb.Scene.from_desc(desc)
    cam = cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], w, h)
    ref = _pt(scene, cam, w, h, 2048).get_output_image().astype(np.float64)
    cache = cb.RadianceCache((0, 0, 0), (1, 1, 1), seed=1)
    r = cb.PBRenderer(cb.RenderingMethod.PATHTRACER)
    r.set_output_size(w, h)
    r.register_scene(scene)
    r.set_max_trace_depth(5)
    r.set_radiance_cache(cache, training_paths=8192, train_steps=4, batch=4096, learning_rate=2e-3)
    for _ in range(300):  # online training: successive frames (fresh seeds), one at a time
        r.render(cam, 1)

    def frame_ms(rd, spp):
        ts = []
        for _ in range(5):
            rd.reset()
            rd.synchronize()
            t0 = time.perf_counter()
            rd.render(cam, spp)
            rd.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts))

    t_nrc = frame_ms(r, 1)
    r.reset()
    r.render(cam, 1)
    e_nrc = rel_mse(r.get_output_image(), ref)
    p = cb.PBRenderer(cb.RenderingMethod.PATHTRACER)
    p.set_output_size(w, h)
    p.register_scene(scene)
    p.set_max_trace_depth(5)
    t_pt = frame_ms(p, 1)
    k = max(1, round(t_nrc / t_pt))
    p.reset()
    p.render(cam, k)  # frames 0..k-1: the spp plain PT affords in the NRC frame's time
    e_pt = rel_mse(p.get_output_image(), ref)
    print(f"cornell 256^2: NRC frame {t_nrc:.3f} ms rel-MSE {e_nrc:.4f}; PT 1 spp {t_pt:.3f} ms -> {k} spp at equal "
          f"time, rel-MSE {e_pt:.4f}; ratio {e_pt / e_nrc:.2f}", flush=True)
    assert e_nrc * 1.5 < e_pt
