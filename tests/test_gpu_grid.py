"""GPU parity of the heterogeneous-medium integrator and its path-replay adjoint.

Bars: forward images same-seed against the FP32 oracle (rel-MSE <= 1e-6); gradients against the
FP64 oracle's path-replay gradient (relative L2 error <= 1e-3) and against central differences of
the GPU's own forward renders; the torch.autograd wrapper returns the C-ABI adjoint's gradients.
"""
import time

import numpy as np
import pytest

import cupbr_amd as cb
import oracle
from cupbr_amd import scenes

pytestmark = pytest.mark.gpu


def rel_mse(x, r):
    return float(np.mean((np.asarray(x, np.float64) - r) ** 2 / (np.asarray(r, np.float64) ** 2 + 1e-2)))


def _case(res, w, h, spp=8, depth=6):
    G = scenes.grid_medium_scene(res=res)
    ca = G["camera"]
    cam = cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], w, h)
    scene = cb.Scene.from_desc(G["desc"])
    med = cb.GridMedium.from_dict(G["density"].shape, G["medium"])
    return G, cam, cb.GridRenderer(scene, cam, med, w, h, spp=spp, max_depth=depth)


def test_grid_forward_same_seed_parity():
    G, cam, R = _case(32, 48, 48)
    img = R.render(G["density"], G["albedo"])
    ref = oracle.OracleScene(G["desc"]).grid(cam.as_dict(), G["medium"], G["density"], G["albedo"], 48, 48, spp=8,
                                             depth=6, precision=32)
    e = rel_mse(img, ref)
    print(f"grid forward: rel-MSE vs same-seed FP32 oracle {e:.2e}; exact {np.mean(img == ref):.3f}", flush=True)
    assert np.isfinite(img).all() and e <= 1e-6
    # a fixed proposal other than the density keeps the estimate unbiased: compare its image with
    # the FP32 oracle for the same proposal
    prop = np.full_like(G["density"], 0.3)
    img2 = R.render(G["density"], G["albedo"], proposal=prop)
    ref2 = oracle.OracleScene(G["desc"]).grid(cam.as_dict(), G["medium"], G["density"], G["albedo"], 48, 48, spp=8,
                                              depth=6, precision=32, proposal=prop)
    assert rel_mse(img2, ref2) <= 1e-6


def test_grid_adjoint_against_fp64_oracle():
    G, cam, R = _case(32, 48, 48)
    S = oracle.OracleScene(G["desc"])
    target = S.grid(cam.as_dict(), G["medium"], G["target_density"], G["albedo"], 48, 48, spp=8, depth=6)
    img = R.render(G["density"], G["albedo"])
    d_img = (2 * (img - target)).astype(np.float32)
    gr, ga = R.adjoint(G["density"], G["albedo"], d_img)
    rr, ra = S.grid(cam.as_dict(), G["medium"], G["density"], G["albedo"], 48, 48, spp=8, depth=6, d_image=d_img,
                    precision=64)
    er = np.linalg.norm(gr - rr) / np.linalg.norm(rr)
    ea = np.linalg.norm(ga - ra) / np.linalg.norm(ra)
    print(f"grid adjoint vs FP64 oracle: density rel-L2 {er:.2e}, albedo rel-L2 {ea:.2e}", flush=True)
    assert np.isfinite(gr).all() and np.isfinite(ga).all()
    assert er <= 1e-3 and ea <= 1e-3


def test_grid_adjoint_ragged_grid_against_fp64_oracle():
    """A 33 x 31 x 29 grid (voxel count neither a multiple of 4 nor of 256): the private gradient
    copies are padded, and the sum back into the caller's gradient must cover exactly the grid."""
    G = scenes.grid_medium_scene(res=33)
    rho = np.ascontiguousarray(G["density"][:29, :31, :])
    tgt = np.ascontiguousarray(G["target_density"][:29, :31, :])
    ca = G["camera"]
    cam = cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], 48, 48)
    scene = cb.Scene.from_desc(G["desc"])
    R = cb.GridRenderer(scene, cam, cb.GridMedium.from_dict(rho.shape, G["medium"]), 48, 48, spp=8, max_depth=6)
    S = oracle.OracleScene(G["desc"])
    target = S.grid(cam.as_dict(), G["medium"], tgt, G["albedo"], 48, 48, spp=8, depth=6)
    d_img = (2 * (R.render(rho, G["albedo"]) - target)).astype(np.float32)
    gr, ga = R.adjoint(rho, G["albedo"], d_img)
    rr, ra = S.grid(cam.as_dict(), G["medium"], rho, G["albedo"], 48, 48, spp=8, depth=6, d_image=d_img,
                    precision=64)
    er = np.linalg.norm(gr - rr) / np.linalg.norm(rr)
    ea = np.linalg.norm(ga - ra) / np.linalg.norm(ra)
    print(f"ragged grid adjoint vs FP64 oracle: density rel-L2 {er:.2e}, albedo rel-L2 {ea:.2e}", flush=True)
    assert gr.shape == rho.shape and np.isfinite(gr).all()
    assert er <= 1e-3 and ea <= 1e-3


def test_grid_adjoint_against_gpu_central_differences():
    G, cam, R = _case(16, 32, 32, spp=4)
    S = oracle.OracleScene(G["desc"], precisions=(32,))
    target = S.grid(cam.as_dict(), G["medium"], G["target_density"], G["albedo"], 32, 32, spp=4, depth=6,
                    precision=32)
    rho, alb = G["density"], G["albedo"]

    def loss(r, a):
        im = R.render(r, a, proposal=rho).astype(np.float64)
        return float(np.sum((im - target) ** 2))

    img = R.render(rho, alb)
    gr, ga = R.adjoint(rho, alb, (2 * (img - target)).astype(np.float32), proposal=rho)
    hh = 1e-2
    worst = 0.0
    for v in np.argsort(-np.abs(gr.ravel()))[:16]:
        rp = rho.copy().ravel()
        rm = rho.copy().ravel()
        rp[v] += hh
        rm[v] -= hh
        fd = (loss(rp.reshape(rho.shape), alb) - loss(rm.reshape(rho.shape), alb)) / (2 * hh)
        worst = max(worst, abs(fd - gr.ravel()[v]) / abs(fd))
    for k in range(0, 10, 3):
        ap = alb.copy()
        am = alb.copy()
        ap[k, 1] += hh
        am[k, 1] -= hh
        fd = (loss(rho, ap) - loss(rho, am)) / (2 * hh)
        worst = max(worst, abs(fd - ga[k, 1]) / abs(fd))
    print(f"grid adjoint vs GPU central differences (h = 1e-2): worst rel error {worst:.2e}", flush=True)
    assert worst <= 2e-2


def test_adjoint_from_forward_matches_the_full_replay():
    """The adjoint that reuses the forward render's per-path radiance equals the one that replays
    every path for it (up to the order of float atomics), and refuses arguments that differ from
    the last render's."""
    torch = pytest.importorskip("torch")
    G, cam, R = _case(32, 48, 48)
    rho = torch.from_numpy(G["density"]).cuda()
    alb = torch.from_numpy(G["albedo"]).cuda()
    target = R.render(torch.from_numpy(G["target_density"]).cuda(), alb)
    img = R.render(rho, alb)
    d_img = 2 * (img - target)
    g1, a1 = R.adjoint(rho, alb, d_img, from_forward=True)
    g0, a0 = R.adjoint(rho, alb, d_img)
    torch.testing.assert_close(g1, g0, rtol=1e-4, atol=1e-5 * float(g0.abs().max()))
    torch.testing.assert_close(a1, a0, rtol=1e-4, atol=1e-5 * float(a0.abs().max()))
    with pytest.raises(cb.CupbrError):  # the last render used other densities
        R.render(torch.from_numpy(G["target_density"]).cuda(), alb)
        R.adjoint(rho, alb, d_img, from_forward=True)
    R2 = cb.GridRenderer(R.scene, cam, R.medium, 48, 48, spp=4, max_depth=6)
    R.render(rho, alb)
    with pytest.raises(cb.CupbrError):  # other sample count
        R2.adjoint(rho, alb, d_img, from_forward=True)
    with pytest.raises(ValueError):  # host arrays: the forward's samples stay on the device
        R.adjoint(G["density"], G["albedo"], d_img.cpu().numpy(), from_forward=True)


def test_torch_autograd_wrapper():
    torch = pytest.importorskip("torch")
    from cupbr_amd.autograd import GridRender

    G, cam, R = _case(32, 48, 48)
    S = oracle.OracleScene(G["desc"], precisions=(32,))
    target = torch.from_numpy(S.grid(cam.as_dict(), G["medium"], G["target_density"], G["albedo"], 48, 48, spp=8,
                                     depth=6, precision=32).astype(np.float32)).cuda()
    render = GridRender(R.scene, cam, R.medium, 48, 48, spp=8, max_depth=6)
    density = torch.from_numpy(G["density"]).cuda().requires_grad_(True)
    albedo = torch.from_numpy(G["albedo"]).cuda().requires_grad_(True)
    img = render(density, albedo)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    d_img = (2 * (img.detach() - target)).cpu().numpy()
    gr, ga = R.adjoint(G["density"], G["albedo"], d_img)
    np.testing.assert_allclose(density.grad.cpu().numpy(), gr, rtol=1e-4, atol=1e-4 * np.abs(gr).max())
    np.testing.assert_allclose(albedo.grad.cpu().numpy(), ga, rtol=1e-4, atol=1e-4 * np.abs(ga).max())

    # a few projected gradient steps on the density lower the loss (inverse rendering)
    losses = [float(loss.detach())]
    with torch.no_grad():
        step = 0.02 / float(density.grad.abs().max())
    for _ in range(4):
        with torch.no_grad():
            density -= step * density.grad
            density.clamp_(min=0.0)
        density.grad = None
        albedo.grad = None
        loss = ((render(density, albedo) - target) ** 2).sum()
        loss.backward()
        losses.append(float(loss.detach()))
    print("inverse rendering loss:", ", ".join(f"{v:.4f}" for v in losses), flush=True)
    assert losses[-1] < losses[0]


def test_grid_full_size_workload():
    """64^3 grid, 256x256, 64 spp: forward and adjoint finish and stay finite."""
    torch = pytest.importorskip("torch")
    G, cam, R = _case(64, 256, 256, spp=64)
    rho = torch.from_numpy(G["density"]).cuda()
    alb = torch.from_numpy(G["albedo"]).cuda()
    R.render(rho, alb)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    img = R.render(rho, alb)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    d_img = 2 * img
    gr, ga = R.adjoint(rho, alb, d_img)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    n = 256 * 256 * 64
    print(f"grid 64^3 256x256x64spp: forward {t1 - t0:.3f}s ({n / (t1 - t0) / 1e6:.1f} Msamples/s), "
          f"adjoint {t2 - t1:.3f}s ({n / (t2 - t1) / 1e6:.1f} Msamples/s)", flush=True)
    assert torch.isfinite(img).all() and torch.isfinite(gr).all() and torch.isfinite(ga).all()
    assert float(gr.abs().sum()) > 0
    # size-independent properties at full size: the device adjoint adds to the buffers it is given,
    # and it is linear in the image cotangent (equal up to the order of float atomics)
    gr2, ga2 = R.adjoint(rho, alb, d_img, grad_density=gr.clone(), grad_albedo=ga.clone())
    torch.testing.assert_close(gr2, 2 * gr, rtol=1e-4, atol=1e-5 * float(gr.abs().max()))
    torch.testing.assert_close(ga2, 2 * ga, rtol=1e-4, atol=1e-5 * float(ga.abs().max()))
    gr3, ga3 = R.adjoint(rho, alb, -0.5 * d_img)
    torch.testing.assert_close(gr3, -0.5 * gr, rtol=1e-4, atol=1e-5 * float(gr.abs().max()))
    torch.testing.assert_close(ga3, -0.5 * ga, rtol=1e-4, atol=1e-5 * float(ga.abs().max()))
