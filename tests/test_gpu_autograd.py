"""The adjoint as a torch.autograd.Function (tests/torch_adjoint.py): gradients that
torch's backward pass delivers for the texels and GGX roughness match central finite
differences of the GPU forward renderer (h = 1e-3, common random numbers), and a few
optimiser steps of inverse rendering lower the image loss."""
import numpy as np
import pytest
import torch

from dort_amd import abi, scenes
from tests.test_gpu_adjoint import ROUGH, ggx_plane, tiny_scene
from tests.torch_adjoint import Renderer, RenderFunction

pytestmark = pytest.mark.gpu


def _weights(res, seed):
    return torch.from_numpy(np.random.default_rng(seed).uniform(-1.0, 1.0, (res, res, 3)).astype(np.float32))


def _loss(renderer, texels, rough, w):
    with torch.no_grad():
        return float((RenderFunction.apply(texels, rough, renderer).double() * w.double()).sum())


def test_backward_texels_match_central_fd(gpu_ctx):
    sc = tiny_scene()
    opts = abi.render_opts(48, 48, spp=8, seed=4, max_depth=6, rr_depth=100)  # no RR: linear in each albedo
    r = Renderer(gpu_ctx.scene(sc), sc, opts)
    tex = torch.tensor(sc.textures[0], requires_grad=True)
    rough = torch.tensor(ROUGH, dtype=torch.float32, requires_grad=True)
    w = _weights(48, 1)
    (RenderFunction.apply(tex, rough, r) * w).sum().backward()
    g = tex.grad.numpy()
    assert rough.grad.shape == (8,) and np.isfinite(rough.grad.numpy()).all()
    h = 1e-3
    for k in np.argsort(-np.abs(g).ravel())[:6]:
        idx = np.unravel_index(k, g.shape)
        tp, tm = tex.detach().clone(), tex.detach().clone()
        tp[idx] += h
        tm[idx] -= h
        fd = (_loss(r, tp, rough, w) - _loss(r, tm, rough, w)) / (2 * h)
        print(f"texel {idx}: autograd {g[idx]:.6f} fd {fd:.6f}")
        assert abs(fd - g[idx]) <= 0.02 * abs(g[idx]) + 1e-4


@pytest.mark.parametrize("alpha", [0.2, 0.5])
def test_backward_roughness_matches_central_fd(gpu_ctx, alpha):
    sc = ggx_plane(alpha)
    opts = abi.render_opts(48, 48, spp=2048, seed=6, max_depth=2, rr_depth=100)
    r = Renderer(gpu_ctx.scene(sc), sc, opts)
    tex = torch.zeros((0, 0, 3), dtype=torch.float32, requires_grad=True)
    rough = torch.tensor([alpha], dtype=torch.float32, requires_grad=True)
    w = _weights(48, 2).abs()
    (RenderFunction.apply(tex, rough, r) * w).sum().backward()
    h = 1e-3
    fd = (_loss(r, tex, rough.detach() + h, w) - _loss(r, tex, rough.detach() - h, w)) / (2 * h)
    print(f"roughness {alpha}: autograd {rough.grad.item():.4f} fd {fd:.4f}")
    assert abs(fd - rough.grad.item()) <= 0.03 * abs(fd)


def test_inverse_rendering_steps_lower_the_loss(gpu_ctx):
    """Adam on the albedo texels and roughness toward a target rendered from hidden values."""
    hidden = tiny_scene()
    target_sc = gpu_ctx.scene(hidden)
    opts = abi.render_opts(32, 32, spp=8, seed=12, max_depth=4, rr_depth=3)
    target = torch.from_numpy(target_sc.image(opts))
    guess = tiny_scene()
    guess.textures[0] = np.full_like(guess.textures[0], 0.5)
    r = Renderer(gpu_ctx.scene(guess), guess, opts)
    tex = torch.tensor(guess.textures[0], requires_grad=True)
    rough = torch.full((8,), 0.3, dtype=torch.float32, requires_grad=True)
    opt = torch.optim.Adam([tex, rough], lr=0.02)
    losses = []
    for _ in range(12):
        opt.zero_grad()
        loss = ((RenderFunction.apply(tex, rough, r) - target) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            tex.clamp_(0.0, 1.0)
            rough.clamp_(0.02, 1.0)
        losses.append(loss.item())
    print("losses", [round(v, 3) for v in losses])
    assert losses[-1] < 0.7 * losses[0]
