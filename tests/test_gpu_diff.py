"""GPU differentiable render: forward parity with the oracle's differentiable
mode, the adjoint against the oracle's central finite differences, the
torch.autograd.Function, and the inverse fit.

Bars:
  forward   FP64: classes identical, colour within 1e-4 (0..255 scale);
            FP32: classes identical except <= 1e-4 of pixels, colour within
            1/255 of full scale.
  gradient  dL/d(r_s, incl, dist, r_inner) within 1e-4 relative of the FP64
            oracle's central differences (h = 1e-6) for L = sum(w * rgb),
            w ~ U[0.5, 1.5], over pixels whose structure signature is the same
            at theta +- h (the render is only piecewise smooth: texel cells,
            crossing counts and classes change discretely).  FP64 at 64x64
            (the gradcheck), FP32 at 64x64 and 256x256.
  fit       200 Adam iterations (coarse-to-fine loss) at 512x512 FP32 from the perturbed start
            reduce the loss twenty-fold and bring every parameter to within a
            tenth of its starting error.
"""
import os

import numpy as np
import pytest
import torch

from blackhole_8_amd import abi, diff
from blackhole_8_amd.render import BHError, Renderer
import oracle

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def dctx():
    """A context of its own: the differentiable renderer binds a torch stream."""
    r = Renderer(0)
    tex = diff.load_target_textures(r)
    yield r, tex
    r.close()


def _dr(dctx, W, prec, **kw):
    r, _ = dctx
    return diff.DiffRenderer(r, diff.DiffConfig(width=W, height=W, precision=prec, **kw))


@pytest.mark.parametrize("W,prec,theta", [
    (64, abi.PREC_FP64, diff.FIT_TRUE), (128, abi.PREC_FP64, diff.FIT_START),
    (512, abi.PREC_FP32, diff.FIT_TRUE), (256, abi.PREC_FP32, diff.FIT_START)])
def test_forward_matches_oracle(dctx, W, prec, theta):
    print(f"[case] diff forward {W}x{W} prec={prec} theta={theta}", flush=True)
    dr = _dr(dctx, W, prec)
    rgb, hit = dr.forward(theta, with_hit=True)
    torch.cuda.synchronize()
    rgb = rgb.cpu().numpy().astype(np.float64)
    hit = hit.cpu().numpy()
    orgb, ohit, _ = oracle.render_diff(dr.cfg.params(theta), dctx[1], THREADS)
    d = np.abs(rgb - orgb).max(axis=2)
    print(f"[diff] class mismatch {(hit != ohit).sum()} max|drgb| {d.max():.3e}", flush=True)
    if prec == abi.PREC_FP64:
        assert (hit == ohit).all()
        assert d.max() <= 1e-4
    else:
        assert (hit != ohit).mean() <= 1e-4
        assert (d[hit == ohit] <= 1.0).all()
    assert (ohit == abi.HIT_DISC).mean() > 0.2 and (ohit == abi.HIT_HORIZON).sum() > 100


def _fd_and_adjoint(dctx, W, prec, h=1e-6, seed=5):
    dr = _dr(dctx, W, prec)
    th = list(diff.FIT_TRUE)
    tex = dctx[1]
    rgb0, _, sig0 = oracle.render_diff(dr.cfg.params(th), tex, THREADS)
    w = np.random.default_rng(seed).uniform(0.5, 1.5, size=rgb0.shape)
    out = []
    for i in range(4):
        tp, tm = list(th), list(th)
        tp[i] += h
        tm[i] -= h
        rp, _, sp = oracle.render_diff(dr.cfg.params(tp), tex, THREADS)
        rm, _, sm = oracle.render_diff(dr.cfg.params(tm), tex, THREADS)
        mask = (sp == sig0) & (sm == sig0)
        wm = w * mask[..., None]
        fd = float((wm * (rp - rm)).sum() / (2 * h))
        g = dr.adjoint(th, torch.from_numpy(wm.astype(np.float32)))
        torch.cuda.synchronize()
        out.append((diff.PARAM_NAMES[i], float(g[i].item()), fd, int((~mask).sum()), dr.overflow()))
    return out


@pytest.mark.parametrize("W,prec", [(64, abi.PREC_FP64), (64, abi.PREC_FP32), (256, abi.PREC_FP32)])
def test_adjoint_matches_oracle_central_differences(dctx, W, prec):
    print(f"[case] diff gradient {W}x{W} prec={prec}", flush=True)
    for name, g, fd, masked, ovf in _fd_and_adjoint(dctx, W, prec):
        rel = abs(g - fd) / abs(fd)
        print(f"[grad] {name}: adjoint {g:.9e} fd {fd:.9e} rel {rel:.2e} masked {masked}",
              flush=True)
        assert ovf == 0
        assert masked <= 0.01 * W * W
        assert rel <= 1e-4


def test_adjoint_is_deterministic_and_returns_forward(dctx):
    print("[case] diff adjoint determinism 128x128 FP32", flush=True)
    dr = _dr(dctx, 128, abi.PREC_FP32)
    g_rgb = torch.rand((128, 128, 3), generator=torch.Generator().manual_seed(3)).cuda()
    g1, rgb1 = dr.adjoint(diff.FIT_TRUE, g_rgb, want_rgb=True)
    g2 = dr.adjoint(diff.FIT_TRUE, g_rgb)
    rgb0 = dr.forward(diff.FIT_TRUE)
    torch.cuda.synchronize()
    assert torch.equal(g1, g2)
    assert torch.equal(rgb1, rgb0)


def test_autograd_function_calls_the_adjoint(dctx):
    print("[case] diff autograd function 64x64 FP64", flush=True)
    dr = _dr(dctx, 64, abi.PREC_FP64)
    theta = torch.tensor(diff.FIT_TRUE, dtype=torch.float64, requires_grad=True)
    w = torch.rand((64, 64, 3), generator=torch.Generator().manual_seed(7)).cuda()
    img = diff.render(theta, dr)
    assert img.shape == (64, 64, 3) and img.dtype == torch.float32 and img.is_cuda
    (img * w).sum().backward()
    ref = dr.adjoint(diff.FIT_TRUE, w).cpu()
    assert theta.grad is not None and theta.grad.dtype == torch.float64
    torch.testing.assert_close(theta.grad, ref, rtol=0, atol=0)


def test_inverse_fit_moves_toward_truth(dctx):
    print("[case] diff inverse fit 512x512 FP32, 200 Adam iterations", flush=True)
    dr = _dr(dctx, 512, abi.PREC_FP32)
    target = dr.forward(diff.FIT_TRUE)
    theta, hist = diff.fit(dr, target, diff.FIT_START, iters=200)
    loss0, loss1 = hist[0][0], hist[-1][0]
    print(f"[fit] loss {loss0:.1f} -> {loss1:.1f}, theta {theta} (true {diff.FIT_TRUE})",
          flush=True)
    assert dr.overflow() == 0
    assert loss1 < 0.05 * loss0
    for t1, t0, tt in zip(theta, diff.FIT_START, diff.FIT_TRUE):
        assert abs(t1 - tt) < 0.1 * abs(t0 - tt)


def test_invalid_parameters_are_rejected(dctx):
    r, _ = dctx
    bad = [dict(r_s=0.0), dict(dist=0.5), dict(r_inner=13.0), dict(edge_width=-1.0),
           dict(width=0)]
    base = dict(zip(diff.PARAM_NAMES, diff.FIT_TRUE))
    for b in bad:
        cfg = diff.DiffConfig(width=b.pop("width", 32), height=32, edge_width=b.pop("edge_width", 0.25))
        th = dict(base, **b)
        dr = diff.DiffRenderer(r, cfg)
        with pytest.raises(BHError) as e:
            dr.forward([th[n] for n in diff.PARAM_NAMES])
        assert e.value.code == abi.BH_ERR_INVALID_ARG
    cfg = diff.DiffConfig(width=32, height=32, texture=9)
    with pytest.raises(BHError) as e:
        diff.DiffRenderer(r, cfg).forward(diff.FIT_TRUE)
    assert e.value.code == abi.BH_ERR_EMPTY_TEXTURE
