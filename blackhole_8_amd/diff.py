"""Differentiable hole + disc render on the GPU and its adjoint, wrapped as a
torch.autograd.Function, and the inverse fit of (r_s, inclination, camera
distance, disc inner radius) to a target image.

The reference renders only (include/blackhole/blackhole_solution_test.cc:
161-298 has no derivative path); the smooth variant of its march is
described in include/bh_api.h (bh_diff_params).  Device memory and the
stream are torch's; the kernels are the HIP library's
(bh_diff_render_device / bh_diff_adjoint_device).  There is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import torch

from . import abi
from .render import BHError, Renderer

PARAM_NAMES = ("r_s", "incl", "dist", "r_inner")

# inverse-fit workload: 512x512, true parameters and the perturbed start
FIT_TRUE = (1.2, math.radians(72.0), 18.0, 3.5)
FIT_START = (1.32, math.radians(68.0), 19.5, 3.15)
FIT_SIZE = 512


@dataclass
class DiffConfig:
    width: int = FIT_SIZE
    height: int = FIT_SIZE
    r_outer: float = 12.0
    azim: float = 0.0
    fov: float = math.radians(70.0)
    edge_width: float = 0.25
    texture: int = 1
    precision: int = abi.PREC_FP32

    def params(self, theta) -> abi.DiffParams:
        t = [float(v) for v in theta]
        return abi.DiffParams(t[0], t[1], t[2], t[3], self.r_outer, self.azim, self.fov,
                              self.edge_width, self.width, self.height, self.texture, 0)


class DiffRenderer:
    """GPU forward and adjoint of the differentiable render, ordered after and
    before torch's current stream.  theta = (r_s, incl, dist, r_inner).

    The kernels run on a stream of this object (default priority), bound to
    the context at each call; each call waits for torch's current stream and
    makes it wait in turn, so results are ready for torch ops that follow,
    also when the current stream is the null stream."""

    def __init__(self, renderer: Renderer, cfg: DiffConfig | None = None):
        if not torch.cuda.is_available():
            raise BHError(abi.BH_ERR_NO_DEVICE, "the differentiable render needs a GPU")
        self.r = renderer
        self.cfg = cfg or DiffConfig()
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self.stream = torch.cuda.Stream(self.dev)
        self._ovf = torch.zeros(1, dtype=torch.int32, device=self.dev)

    def _bind(self):
        self.r._check(self.r.lib.bh_ctx_set_stream(self.r.ctx, C.c_void_p(self.stream.cuda_stream)))

    def _run(self, fn, tensors):
        self._bind()
        cur = torch.cuda.current_stream(self.dev)
        self.stream.wait_stream(cur)
        for t in tensors:
            if t is not None:
                t.record_stream(self.stream)
        self.r._check(fn())
        cur.wait_stream(self.stream)

    def forward(self, theta, with_hit: bool = False):
        c = self.cfg
        p = c.params(theta)
        rgb = torch.empty((c.height, c.width, 3), dtype=torch.float32, device=self.dev)
        hit = torch.empty((c.height, c.width), dtype=torch.uint8, device=self.dev) if with_hit else None
        self._run(lambda: self.r.lib.bh_diff_render_device(
            self.r.ctx, C.byref(p), c.precision, C.c_void_p(rgb.data_ptr()),
            C.c_void_p(hit.data_ptr()) if hit is not None else None), (rgb, hit))
        return (rgb, hit) if with_hit else rgb

    def adjoint(self, theta, grad_rgb: torch.Tensor, want_rgb: bool = False):
        """dL/dtheta (float64 tensor of 4 on the device) for dL/drgb; with
        want_rgb also the forward colour.  overflow() then counts the rays
        whose disc crossings exceeded the adjoint's record."""
        c = self.cfg
        g = grad_rgb.detach().to(device=self.dev, dtype=torch.float32).contiguous()
        if g.shape != (c.height, c.width, 3):
            raise ValueError(f"grad_rgb must be ({c.height}, {c.width}, 3), got {tuple(g.shape)}")
        p = c.params(theta)
        rgb = torch.empty((c.height, c.width, 3), dtype=torch.float32, device=self.dev) if want_rgb else None
        grad = torch.empty(4, dtype=torch.float64, device=self.dev)
        self._run(lambda: self.r.lib.bh_diff_adjoint_device(
            self.r.ctx, C.byref(p), c.precision, C.c_void_p(g.data_ptr()),
            C.c_void_p(rgb.data_ptr()) if rgb is not None else None, None,
            C.c_void_p(grad.data_ptr()), C.c_void_p(self._ovf.data_ptr())),
            (g, rgb, grad, self._ovf))
        return (grad, rgb) if want_rgb else grad

    def overflow(self) -> int:
        """Rays of the last adjoint whose disc crossings exceeded the record."""
        return int(self._ovf.item())

    def time(self, theta, iters: int = 20):
        """Average ms of the forward and of the forward+adjoint launch."""
        f, a = C.c_double(), C.c_double()
        p = self.cfg.params(theta)
        torch.cuda.synchronize(self.dev)
        self._bind()
        self.r._check(self.r.lib.bh_diff_time(self.r.ctx, C.byref(p), self.cfg.precision, iters,
                                              C.byref(f), C.byref(a)))
        return f.value, a.value


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, dr: DiffRenderer):
        ctx.dr = dr
        ctx.theta = theta.detach().cpu().tolist()
        ctx.theta_meta = (theta.dtype, theta.device)
        return dr.forward(ctx.theta)

    @staticmethod
    def backward(ctx, grad_rgb: torch.Tensor):
        g = ctx.dr.adjoint(ctx.theta, grad_rgb)
        dtype, device = ctx.theta_meta
        return g.to(dtype=dtype, device=device), None


def render(theta: torch.Tensor, dr: DiffRenderer) -> torch.Tensor:
    """rgb (H, W, 3) float32 in [0, 255] as a differentiable function of theta."""
    return _Render.apply(theta, dr)


# per-parameter step scale of the fit (units: r_s, rad, r_s, r_s)
FIT_SCALE = (0.025, math.radians(1.0), 0.25, 0.1)


def _pool(img: torch.Tensor, k: int) -> torch.Tensor:
    """k x k box average of an (H, W, 3) image."""
    if k == 1:
        return img
    return torch.nn.functional.avg_pool2d(img.permute(2, 0, 1).unsqueeze(0), k).squeeze(0)


def fit(dr: DiffRenderer, target: torch.Tensor, theta0, iters: int = 200, lr: float = 0.2,
        pyramid=(32, 16, 8, 4, 2, 1), log=None):
    """Adam on the mean squared error to target, coarse to fine: the
    iterations are split evenly over the box-filter sizes in `pyramid`, and
    the loss of stage i is the sum over pyramid[:i + 1] of the mean squared
    error between k x k averages of the render and of the target.  The noisy
    disc texture makes the full-resolution loss rough in r_s and the
    distance: the coarse levels find the basin and, kept in the sum, hold
    the finer stages in it while they find the minimum.  The learning rate
    is annealed on a cosine to 5% over the iterations; theta = theta0 +
    scale * x, x from 0.  Returns (theta, history of (full-resolution mse,
    theta))."""
    scale = torch.tensor(FIT_SCALE, dtype=torch.float64)
    base = torch.tensor([float(v) for v in theta0], dtype=torch.float64)
    x = torch.zeros(4, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([x], lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=iters, eta_min=0.05 * lr)
    targets = {k: _pool(target, k) for k in pyramid}
    hist = []
    for it in range(iters):
        stage = min(len(pyramid) - 1, it * len(pyramid) // iters)
        opt.zero_grad()
        theta = base + scale * x
        img = render(theta, dr)
        loss = sum(((_pool(img, k) - targets[k]) ** 2).mean() for k in pyramid[:stage + 1])
        loss.backward()
        mse = float(((img.detach() - target) ** 2).mean())
        hist.append((mse, theta.detach().tolist()))
        if log is not None and (it % 20 == 0 or it == iters - 1):
            log(it, hist[-1])
        opt.step()
        sched.step()
    theta = (base + scale * x).detach()
    return theta.tolist(), hist


def load_target_textures(r: Renderer):
    """Bind the disc texture of the 1080p hole + disc workload to slot 1."""
    from . import scenes

    tex = scenes.gen_disc_texture()
    r.set_texture(1, tex)
    return {1: tex}

