"""torch.autograd.Function over dgpu_render (forward) and dgpu_adjoint (backward), for
the tests: the image of a scene as a differentiable function of its albedo texture 0
and of every GGX conductor's roughness.

    image = RenderFunction.apply(texels, roughness, renderer)

texels: (H, W, 3) float32 tensor of texture 0 (empty for a scene without one); roughness:
one float32 per GGX conductor
(in material order); the result is the film mean (y_res, x_res, 3).  The backward pass
is the adjoint path replay: it re-traces every path from its counter-based random numbers
and returns dL/dtexels and dL/droughness for the incoming dL/dimage.
"""
import numpy as np
import torch

from dort_amd import abi


class Renderer:
    """A scene on the GPU with the render options; the parameters are set per call."""

    def __init__(self, gpu_scene, scene, opts):
        self.gsc = gpu_scene
        self.scene = scene
        self.opts = opts
        self.ggx = [i for i, m in enumerate(scene.materials) if m.type == abi.MAT_GGX_CONDUCTOR]

    def set_parameters(self, texels, roughness):
        if texels.numel():
            self.gsc.update_texture(0, np.ascontiguousarray(texels.detach().cpu().numpy(), dtype=np.float32))
        for i, r in zip(self.ggx, roughness.detach().cpu().numpy().tolist()):
            m = abi.Material.from_buffer_copy(self.scene.materials[i])
            m.roughness = float(r)
            self.gsc.update_material(i, m)


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, texels, roughness, renderer):
        renderer.set_parameters(texels, roughness)
        ctx.renderer = renderer
        img = renderer.gsc.image(renderer.opts)
        return torch.from_numpy(np.ascontiguousarray(img, dtype=np.float32)).to(texels.device)

    @staticmethod
    def backward(ctx, grad_image):
        r = ctx.renderer
        dldi = np.ascontiguousarray(grad_image.detach().cpu().numpy(), dtype=np.float32)
        g_tex, g_rough, _, _ = r.gsc.adjoint(r.opts, dldi)
        if g_tex is None:
            g_tex = np.zeros((0, 0, 3), np.float32)
        g_t = torch.from_numpy(np.ascontiguousarray(g_tex, dtype=np.float32)).to(grad_image.device)
        g_r = torch.tensor([float(g_rough[i]) for i in r.ggx], dtype=torch.float32, device=grad_image.device)
        return g_t, g_r, None
