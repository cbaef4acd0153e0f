"""Differentiable sparse-brick-set rendering on top of the C ABI.

SBSRender is an optional torch.autograd.Function: colour = render(params) with the adjoint
kernel as its backward. Ray helpers build the ring of views used by the differentiable
workloads.
"""
import math

import numpy as np

from . import camera_inverse, eye_rays


def ring_cameras(count=8, radius=3.0, ring_count=1, amplitude_deg=45.0, look_at=(0.0, 0.0, 0.0)):
    """Camera positions of render_CLI's `cameras_distribution = 1` ring
    (render_CLI/offscreen_render.cpp:82-95), float32 arithmetic as there."""
    per_ring = max(1, count // ring_count)
    amp = np.float32(min(max(amplitude_deg, 0.0), 89.9))
    out = []
    for i in range(count):
        phi = np.float32(2.0 * 3.14159) * np.float32(i) / np.float32(per_ring)
        theta = (np.float32(3.14159) * ((np.float32(90) - amp) + np.float32(2) * (np.float32(i) / np.float32(count)) * amp)
                 / np.float32(180))
        x = np.float32(math.cos(phi)) * np.float32(math.sin(theta))
        y = np.float32(math.cos(theta))
        z = np.float32(math.sin(phi)) * np.float32(math.sin(theta))
        out.append(np.array([x * radius + look_at[0], y * radius + look_at[1], z * radius + look_at[2]], np.float32))
    return out


def ring_rays(torch, width, height, count=8, radius=3.0, fov=60.0, stream=0):
    """Primary rays of `count` ring views (row-major per view, views concatenated) on the device."""
    n = width * height
    pos = torch.zeros((count * n, 4), dtype=torch.float32, device="cuda")
    dirs = torch.zeros_like(pos)
    for v, eye in enumerate(ring_cameras(count, radius)):
        pi, vi = camera_inverse(list(eye), [0, 0, 0], [0, 1, 0], fov, width / height, 0.01, 100.0)
        eye_rays(pi, vi, width, height, pos[v * n:(v + 1) * n], dirs[v * n:(v + 1) * n], order=0, stream=stream)
    return pos, dirs


def _function():
    import torch

    class SBSRender(torch.autograd.Function):
        """colour (n, 4) = render(params) for one brick-set geometry; backward = adjoint march."""

        @staticmethod
        def forward(ctx, params, scene, geom_id, pos, dirs):
            colour = torch.zeros((pos.shape[0], 4), dtype=torch.float32, device=pos.device)
            stream = torch.cuda.current_stream().cuda_stream
            scene.sbs_forward(geom_id, params.detach().contiguous(), pos, dirs, colour, stream=stream)
            ctx.save_for_backward(params, pos, dirs)
            ctx.scene, ctx.geom_id = scene, geom_id
            return colour

        @staticmethod
        def backward(ctx, grad_colour):
            params, pos, dirs = ctx.saved_tensors
            grad = torch.zeros_like(params)
            stream = torch.cuda.current_stream().cuda_stream
            ctx.scene.sbs_backward(ctx.geom_id, params.detach().contiguous(), pos, dirs,
                                   grad_colour.contiguous(), grad, stream=stream)
            return grad, None, None, None, None

    return SBSRender


_SBS_RENDER = None


def sbs_render(params, scene, geom_id, pos, dirs):
    """Differentiable render: torch tensor params (requires_grad allowed) -> colour (n, 4)."""
    global _SBS_RENDER
    if _SBS_RENDER is None:
        _SBS_RENDER = _function()
    return _SBS_RENDER.apply(params, scene, geom_id, pos, dirs)
