"""Differentiable rendering on the HIP rasterizer.

`DiffRenderer` drives sp_diff_forward / sp_diff_backward (include/splat_hip.h)
on torch device tensors; `render` wraps them in a torch.autograd.Function, so
a loss built on the returned images back-propagates into the raw gaussian
parameters:

    means [n,3], log_scales [n,3], quats [n,4] (x,y,z,w, normalised inside),
    opacity_logits [n], colors [n,3] (base colour = SH DC term),
    sh_rest [n,(deg+1)^2-1,3]

The images are the float values C + T*bg that splat-tui's to_u8 would
quantise (src/render/mod.rs:282-285), rendered with the `cpu` backend's
blend rules. There is no CPU path: the HIP library must be built.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

PARAM_KEYS = ("means", "log_scales", "quats", "opacity_logits", "colors", "sh_rest")


def params_from_scene(scene) -> dict:
    """Raw parameters (numpy float32) of an activated Scene: log of the
    scales, logit of the opacities, the unit quaternions as stored."""
    op = np.clip(scene.opacities.astype(np.float64), 1e-7, 1 - 1e-7)
    rest = scene.sh_rest.reshape(len(scene), -1, 3) if scene.sh_rest is not None else np.zeros((len(scene), 0, 3))
    return {
        "means": np.ascontiguousarray(scene.positions, np.float32),
        "log_scales": np.log(scene.scales).astype(np.float32),
        "quats": np.ascontiguousarray(scene.rotations, np.float32),
        "opacity_logits": np.log(op / (1 - op)).astype(np.float32),
        "colors": np.ascontiguousarray(scene.colors, np.float32),
        "sh_rest": np.ascontiguousarray(rest, np.float32),
    }


def _torch():
    import torch

    return torch


def _ptr(t) -> int | None:
    return t.data_ptr() if t is not None and t.numel() > 0 else None


class DiffRenderer:
    """A differentiable-render handle for n gaussians of SH degree sh_degree."""

    def __init__(self, n: int, sh_degree: int, device: int = 0, initial_pairs: int = 0):
        """initial_pairs: the pair buffer's first capacity (0: automatic). A
        forward whose pairs exceed it grows the buffer and renders again."""
        lib = N.hip()
        opt = N.SpOptions()
        lib.sp_default_options(C.byref(opt))
        opt.device = device
        opt.initial_pairs = initial_pairs
        h = C.c_void_p()
        N.check(lib.sp_diff_create(n, sh_degree, C.byref(opt), C.byref(h)))
        self._h = h
        self.n = n
        self.sh_degree = sh_degree
        self.device = device
        self.token = 0  # increments with every forward; the backward must follow its forward
        self.last_stats = None
        self.last_backward_ms = None
        self._nviews = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.hip().sp_destroy(h)
            self._h = None

    def _params(self, params: dict):
        torch = _torch()
        rest = (self.sh_degree + 1) ** 2 - 1
        want = {"means": (self.n, 3), "log_scales": (self.n, 3), "quats": (self.n, 4),
                "opacity_logits": (self.n,), "colors": (self.n, 3), "sh_rest": (self.n, rest, 3)}
        out = {}
        for k in PARAM_KEYS:
            t = params[k]
            if tuple(t.shape) != want[k]:
                raise ValueError(f"{k}: shape {tuple(t.shape)}, expected {want[k]}")
            if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
                raise ValueError(f"{k}: need a contiguous float32 device tensor")
            out[k] = t
        sp = N.SpParams(*[_ptr(out[k]) for k in PARAM_KEYS])
        return out, sp

    def forward(self, params: dict, views, sizes, background=(0.0, 0.0, 0.0)) -> list:
        """Render len(views) float images (list of [h, w, 3] tensors)."""
        torch = _torch()
        _, sp = self._params(params)
        nv = len(views)
        sizes = [(int(w), int(h)) for w, h in sizes]
        if len(sizes) != nv:
            raise ValueError("one (width, height) per view")
        dev = params["means"].device
        imgs = [torch.empty((h, w, 3), dtype=torch.float32, device=dev) for w, h in sizes]
        varr = (N.SpView * max(nv, 1))(*views)
        wh = np.array([x for s in sizes for x in s], np.uint32)
        outs = (C.c_void_p * max(nv, 1))(*[_ptr(t) for t in imgs])
        bg = np.ascontiguousarray(background, np.float32)
        stats = (N.SpStats * max(nv, 1))()
        torch.cuda.synchronize(dev)
        N.check(N.hip().sp_diff_forward(self._h, C.byref(sp), varr, wh.ctypes.data_as(N.u32p), nv,
                                        bg.ctypes.data_as(N.f32p), outs, stats))
        self.token += 1
        self._nviews = nv
        self.last_stats = [stats[v] for v in range(nv)]
        return imgs

    def backward(self, params: dict, grad_images) -> dict:
        """dL/dparameters for dL/dimage of the last forward (same params)."""
        torch = _torch()
        p, sp = self._params(params)
        if len(grad_images) != self._nviews:
            raise ValueError("one gradient image per view of the last forward")
        gimgs = [g.to(torch.float32).contiguous() for g in grad_images]
        grads = {k: torch.empty_like(p[k]) for k in PARAM_KEYS}
        sg = N.SpParamGrads(*[_ptr(grads[k]) for k in PARAM_KEYS])
        gptr = (C.c_void_p * max(len(gimgs), 1))(*[_ptr(g) for g in gimgs])
        ms = C.c_double(0.0)
        torch.cuda.synchronize(p["means"].device)
        N.check(N.hip().sp_diff_backward(self._h, C.byref(sp), gptr, C.byref(sg), C.byref(ms)))
        self.last_backward_ms = ms.value
        return grads


def _function():
    torch = _torch()

    class RenderFunction(torch.autograd.Function):
        """images = render(params); backward through sp_diff_backward."""

        @staticmethod
        def forward(ctx, renderer, views, sizes, background, *tensors):
            params = dict(zip(PARAM_KEYS, tensors))
            imgs = renderer.forward(params, views, sizes, background)
            ctx.renderer = renderer
            ctx.views = views
            ctx.sizes = sizes
            ctx.background = background
            ctx.token = renderer.token
            ctx.save_for_backward(*tensors)
            if len({(int(w), int(h)) for w, h in sizes}) != 1:
                raise ValueError("render() stacks the views: all sizes must be equal")
            return torch.stack(imgs)

        @staticmethod
        def backward(ctx, gout):
            tensors = ctx.saved_tensors
            params = dict(zip(PARAM_KEYS, tensors))
            r = ctx.renderer
            if r.token != ctx.token:  # another forward ran since: restore this one's state
                r.forward(params, ctx.views, ctx.sizes, ctx.background)
            g = r.backward(params, [gout[v] for v in range(gout.shape[0])])
            return (None, None, None, None) + tuple(g[k] for k in PARAM_KEYS)

    return RenderFunction


_FN = None


def render(renderer: DiffRenderer, views, sizes, background, means, log_scales, quats, opacity_logits, colors,
           sh_rest):
    """Differentiable render of equal-size views: a [V, h, w, 3] float tensor."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(renderer, views, sizes, background, means, log_scales, quats, opacity_logits, colors, sh_rest)
