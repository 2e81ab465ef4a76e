"""torch.autograd wrapper of the heterogeneous-medium renderer and its path-replay adjoint.

    render = GridRender(scene, camera, medium, width, height, spp=16, max_depth=6)
    image = render(density, albedo)           # density [nz, ny, nx], albedo [n_surfaces, 3], CUDA
    loss = ((image - target) ** 2).sum()
    loss.backward()                            # density.grad, albedo.grad from the path-replay adjoint

The forward and backward run the HIP kernels of libcupbr on torch's current stream; nothing
is computed by torch except the loss. The proposal grid that drives the real/null collision
decisions is the density's value at the call (detached), or a fixed grid given to the
constructor, so the gradient is that of the same-seed estimate.
"""
from __future__ import annotations

import torch

from . import Camera, CupbrError, GridMedium, GridRenderer, Scene


class _GridRenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, density, albedo, renderer: GridRenderer, proposal):
        prop = proposal if proposal is not None else density.detach().clone()
        image = renderer.render(density, albedo, proposal=prop)
        ctx.renderer = renderer
        ctx.save_for_backward(density, albedo, prop)
        return image

    @staticmethod
    def backward(ctx, grad_image):
        density, albedo, prop = ctx.saved_tensors
        # The saved tensors cannot have changed since the forward (autograd checks their versions),
        # so the adjoint may reuse the forward pass's per-path radiance when that render was the
        # last one on the scene; otherwise every path is replayed once more.
        try:
            gd, ga = ctx.renderer.adjoint(density, albedo, grad_image, proposal=prop, from_forward=True)
        except CupbrError:
            gd, ga = ctx.renderer.adjoint(density, albedo, grad_image, proposal=prop)
        return gd, ga, None, None


class GridRender(torch.nn.Module):
    """Differentiable render of a voxel-density medium and per-surface albedos."""

    def __init__(self, scene: Scene, camera: Camera, medium: GridMedium, width: int, height: int, spp: int = 16,
                 max_depth: int = 6, frame0: int = 0, proposal: torch.Tensor | None = None):
        super().__init__()
        self.renderer = GridRenderer(scene, camera, medium, width, height, spp=spp, max_depth=max_depth,
                                     frame0=frame0)
        self.proposal = proposal

    def forward(self, density: torch.Tensor, albedo: torch.Tensor) -> torch.Tensor:
        if not (density.is_cuda and albedo.is_cuda):
            raise ValueError("GridRender runs on the GPU: density and albedo must be CUDA tensors")
        if tuple(density.shape) != self.renderer.medium.shape:
            raise ValueError(f"density shape {tuple(density.shape)} != medium grid {self.renderer.medium.shape}")
        return _GridRenderFn.apply(density, albedo, self.renderer, self.proposal)
