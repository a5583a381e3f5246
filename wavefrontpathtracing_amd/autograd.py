"""torch.autograd binding of the differentiable renderer.

`render(renderer, texels, strengths, texture)` renders one frame of the
renderer's current scene(s) with `texels` as the RGBA float contents of
base-color texture `texture` and `strengths` as the per-instance emissive
strengths, and returns the frame's linear energy image (pixels, 4) as a
CUDA tensor.  Its backward runs the library's path-replay adjoint of that
frame (wf_renderer_render_adjoint_device), so a backward must follow its
forward before the renderer renders again.

The gradient with respect to a texel whose value is exactly 0 is 0: path
replay divides a vertex's remaining radiance by its albedo, and both vanish
there.  Optimisers should keep texels at or above a small positive value
(inverse.TEXEL_MIN), where the gradient is the derivative of the estimate.

Both directions synchronise the caller's current torch stream before the
library reads torch memory, and the library synchronises its own stream
before returning, so no stream ordering is left to the caller.
"""
from __future__ import annotations

import torch

from . import _lib as L


class _RenderFrame(torch.autograd.Function):
    @staticmethod
    def forward(ctx, texels, strengths, renderer, texture, n_pixels):
        if not texels.is_cuda:
            raise ValueError("texels must be a CUDA tensor")
        tx = texels.detach().contiguous().float()
        torch.cuda.current_stream().synchronize()
        renderer.update_texture_device(texture, tx.data_ptr())
        host = strengths.detach().float().cpu().tolist()
        cache = getattr(renderer, "_emission_cache", None)
        if cache is None or len(cache) != len(host):
            cache = [None] * len(host)
        for k, v in enumerate(host):
            if cache[k] != v:
                renderer.set_emission_strength(k, v)
                cache[k] = v
        renderer._emission_cache = cache
        renderer.render()
        img = torch.empty((n_pixels, 4), dtype=torch.float32, device=texels.device)
        renderer.read_energy_device(img.data_ptr())
        ctx.renderer, ctx.texture = renderer, texture
        ctx.tex_shape, ctx.n_inst = tx.shape, len(host)
        ctx.str_device = strengths.device
        return img

    @staticmethod
    def backward(ctx, grad_img):
        g = grad_img.detach().contiguous().float()
        dev = g.device
        d_tex = torch.empty(ctx.tex_shape, dtype=torch.float32, device=dev)
        d_em = torch.empty(ctx.n_inst, dtype=torch.float32, device=dev)
        torch.cuda.current_stream().synchronize()
        ctx.renderer.render_adjoint_device(g.data_ptr(), ctx.texture, d_tex.data_ptr(), d_em.data_ptr())
        return d_tex, d_em.to(ctx.str_device), None, None, None


def render(renderer, texels: torch.Tensor, strengths: torch.Tensor, texture: int) -> torch.Tensor:
    """Differentiable frame: (pixels, 4) energy image of every scene of the
    renderer (batch scenes concatenated in order)."""
    if texture == L.INVALID_HANDLE:
        raise ValueError("a base-color texture handle is required")
    n_pixels = renderer.image_pixels()
    return _RenderFrame.apply(texels, strengths, renderer, texture, n_pixels)
