"""Pins of the oracle's heterogeneous-medium integrator and its path-replay adjoint (CPU).

The integrator has no reference counterpart; its oracle restates cupbr_amd/csrc/grid.hip and is
pinned here by what the mathematics fixes:
  - transmittance: a pure absorber between the camera and a wall lit by a point light gives
    exp(-sigma (l_cam + l_shadow)) * a/pi * I cos / d^2, for any proposal grid (the weighted
    tracking is unbiased whatever drives its decisions);
  - single scattering: with max_depth 1 the image is the 1-D integral of
    sigma_s exp(-sigma t) p_HG(theta) I/d^2 exp(-sigma l_shadow) along the camera ray plus the
    attenuated wall term, evaluated by quadrature;
  - the adjoint: for fixed seeds the estimate is smooth in the parameters, so central finite
    differences of the FP64 oracle's loss must equal its path-replay gradient.
"""
import math

import numpy as np
import pytest

import cupbr_amd as cb
import oracle
from cupbr_amd import scenes


def _wall_scene(light):
    b = scenes.SceneBuilder()
    a = 0.6
    b.quad((0.5, 0.5, 2.0), (0, 0, -1), (5, 0, 0), (0, 5, 0), scenes.material(scenes.LAMBERT, albedo_d=(a, a, a)))
    b.point_light(tuple(light), (10.0, 10.0, 10.0))
    return b.desc(), np.array([[a, a, a]])


def _medium(alpha, g):
    return dict(bmin=(0, 0, 0), bmax=(1, 1, 1), sigma_t_scale=10.0, majorant=10.0, scattering_albedo=alpha, g=g,
                min_collision_probability=1e-3)


# a narrow camera on the cube's axis: every ray crosses the unit cube along z
W = H = 16
CAM = cb.Camera.look_at((0.5, 0.5, -50), (0.5, 0.5, 0.5), (0, 1, 0), 0.02, W, H)


def _z(img, expect):
    return (img.mean() - expect) / (img.std(ddof=1) / math.sqrt(img.size))


@pytest.mark.parametrize("proposal", [None, 0.2, 0.05])
def test_transmittance_closed_form_for_any_proposal(proposal):
    desc, alb = _wall_scene((0.5, 0.5, -0.5))
    S = oracle.OracleScene(desc, precisions=(64,))
    sig = 1.3
    rho = np.full((4, 4, 4), sig / 10.0)
    prop = None if proposal is None else np.full_like(rho, proposal)
    img = S.grid(CAM.as_dict(), _medium(0.0, 0.0), rho, alb, W, H, spp=2000, depth=1, proposal=prop)[..., 0]
    expect = math.exp(-2 * sig) * 0.6 / math.pi * 10 / 2.5 ** 2  # l_cam = l_shadow = 1, cos = 1, d = 2.5
    assert abs(_z(img, expect)) < 4.0, (img.mean(), expect)
    assert abs(img.mean() / expect - 1) < 0.02


def test_single_scattering_closed_form():
    light = np.array([2.5, 0.5, 0.5])
    desc, alb = _wall_scene(light)
    S = oracle.OracleScene(desc, precisions=(64,))
    sig, alpha, g, I = 1.3, 0.8, 0.5, 10.0
    rho = np.full((4, 4, 4), sig / 10.0)
    img = S.grid(CAM.as_dict(), _medium(alpha, g), rho, alb, W, H, spp=4000, depth=1)[..., 0]
    n = 20000
    t = (np.arange(n) + 0.5) / n
    x = np.stack([np.full(n, 0.5), np.full(n, 0.5), t], 1)
    v = light - x
    d = np.linalg.norm(v, axis=1)
    ldir = v / d[:, None]
    l_shadow = (1.0 - x[:, 0]) / ldir[:, 0]  # to the cube face x = 1
    ph = np.array([oracle.hg_phase(g, c) for c in ldir[:, 2]])
    single = np.sum(sig * alpha * np.exp(-sig * t) * ph * I / d ** 2 * np.exp(-sig * l_shadow)) / n
    wall = math.exp(-sig) * 0.6 / math.pi * I * 0.6 / 2.5 ** 2  # wall point (0.5,0.5,2): cos 0.6, d 2.5, no medium
    expect = single + wall
    assert abs(_z(img, expect)) < 4.0, (img.mean(), expect)
    assert abs(img.mean() / expect - 1) < 0.01


def _grid_case(res=16, w=24, h=24):
    G = scenes.grid_medium_scene(res=res)
    ca = G["camera"]
    cam = cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], w, h)
    return G, cam


def test_adjoint_equals_central_differences_fp64():
    """Path-replay gradient of L = sum (I - T)^2 against central differences (h = 1e-3) on 32
    voxels and 6 albedo entries, same seeds, proposal held at the unperturbed density."""
    G, cam = _grid_case()
    w = h = 24
    spp, depth = 4, 6
    S = oracle.OracleScene(G["desc"], precisions=(64,))
    rho = G["density"].astype(np.float64)
    alb = G["albedo"].astype(np.float64)
    med = G["medium"]
    target = S.grid(cam.as_dict(), med, G["target_density"], alb, w, h, spp=spp, depth=depth)
    img = S.grid(cam.as_dict(), med, rho, alb, w, h, spp=spp, depth=depth)

    def loss(r, a):
        im = S.grid(cam.as_dict(), med, r, a, w, h, spp=spp, depth=depth, proposal=rho)
        return float(np.sum((im - target) ** 2))

    g_rho, g_alb = S.grid(cam.as_dict(), med, rho, alb, w, h, spp=spp, depth=depth, proposal=rho,
                          d_image=2 * (img - target))
    assert np.isfinite(g_rho).all() and np.isfinite(g_alb).all()
    rng = np.random.default_rng(5)
    flat = g_rho.ravel()
    candidates = np.argsort(-np.abs(flat))[:1000]  # voxels the paths actually touch
    hh = 1e-3
    for v in rng.choice(candidates, 32, replace=False):
        rp = rho.copy().ravel()
        rm = rho.copy().ravel()
        rp[v] += hh
        rm[v] -= hh
        fd = (loss(rp.reshape(rho.shape), alb) - loss(rm.reshape(rho.shape), alb)) / (2 * hh)
        assert abs(fd - flat[v]) <= 1e-4 * abs(fd) + 1e-9, (v, flat[v], fd)
    for k, c in [(0, 0), (2, 1), (5, 2), (6, 0), (8, 1), (9, 2)]:
        ap = alb.copy()
        am = alb.copy()
        ap[k, c] += hh
        am[k, c] -= hh
        fd = (loss(rho, ap) - loss(rho, am)) / (2 * hh)
        assert abs(fd - g_alb[k, c]) <= 1e-4 * abs(fd) + 1e-9, (k, c, g_alb[k, c], fd)


def test_adjoint_is_linear_in_the_cotangent_and_zero_for_zero():
    G, cam = _grid_case(res=8, w=12, h=12)
    S = oracle.OracleScene(G["desc"], precisions=(64,))
    rho, alb = G["density"].astype(np.float64), G["albedo"].astype(np.float64)
    rng = np.random.default_rng(1)
    d1, d2 = rng.normal(size=(12, 12, 3)), rng.normal(size=(12, 12, 3))
    run = lambda d: S.grid(cam.as_dict(), G["medium"], rho, alb, 12, 12, spp=2, depth=4, d_image=d)  # noqa: E731
    g1, g2, g12 = run(d1), run(d2), run(2 * d1 - 3 * d2)
    for a, b, c in zip(g1, g2, g12):
        np.testing.assert_allclose(c, 2 * a - 3 * b, rtol=1e-9, atol=1e-12)
    z = run(np.zeros((12, 12, 3)))
    assert not z[0].any() and not z[1].any()


def test_fp32_and_fp64_oracles_agree():
    G, cam = _grid_case(res=16, w=24, h=24)
    S = oracle.OracleScene(G["desc"])
    rho, alb = G["density"], G["albedo"]
    i32 = S.grid(cam.as_dict(), G["medium"], rho, alb, 24, 24, spp=8, depth=6, precision=32)
    i64 = S.grid(cam.as_dict(), G["medium"], rho, alb, 24, 24, spp=8, depth=6, precision=64)
    rel = float(np.mean((i32 - i64) ** 2 / (i64 ** 2 + 1e-2)))
    assert rel < 1e-6, rel


def test_grid_entry_points_validate_before_touching_the_gpu():
    G, cam = _grid_case(res=8, w=8, h=8)
    scene = cb.Scene.from_desc(G["desc"])
    good = G["medium"]
    R = cb.GridRenderer(scene, cam, cb.GridMedium.from_dict((8, 8, 8), good), 8, 8)
    assert R.surface_count() == 10
    with pytest.raises(cb.CupbrError):
        R.render(np.zeros((4, 4, 4), np.float32), G["albedo"])  # wrong grid size
    with pytest.raises(cb.CupbrError):
        R.render(G["density"], G["albedo"][:5])  # wrong albedo size
    for bad in (dict(majorant=float("inf")), dict(majorant=0.0), dict(majorant=1e6), dict(g=1.0),
                dict(min_collision_probability=0.0), dict(bmax=(1.0, 0.0, 1.0)), dict(bmin=(float("nan"), 0, 0))):
        m = dict(good, **bad)
        R = cb.GridRenderer(scene, cam, cb.GridMedium.from_dict((8, 8, 8), m), 8, 8)
        with pytest.raises(cb.CupbrError) as e:
            R.render(G["density"], G["albedo"])
        assert e.value.code == -1, bad
