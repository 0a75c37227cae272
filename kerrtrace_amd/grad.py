"""Differentiable renderer over the C ABI (kt_grad_*): the observer scene as
a function of theta = (spin, inclination [rad], r_obs [M], mdot), with the
adjoint pass exposed as a torch.autograd.Function.

The gradient is that of the rendered linear HDR frame with the committed
step sequence and every event (disk crossing, horizon, escape, lookup-table
cell) held fixed: the derivative of the integrator the forward pass ran.
The reference renders no gradients; this is an addition of this library.
"""
import ctypes as C

from ._abi import KT_ADJOINT_FP64, KtGradOpts, KtGradStats
from ._lib import check, lib
from .config import GRAD_FOV_DEG


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("kerrtrace_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


class GradRenderer:
    """One differentiable renderer handle: forward(theta) then backward(dL/dhdr)."""

    def __init__(self, cfg, fov_deg=GRAD_FOV_DEG, checkpoint_interval=32, log_steps_per_ray=512,
                 adjoint_precision=KT_ADJOINT_FP64):
        """cfg.precision chooses the forward tracer (FP64 or mixed);
        adjoint_precision the reverse pass (KT_ADJOINT_FP64 or _FP32)."""
        _torch()
        self.cfg = cfg.copy()
        self.opts = KtGradOpts(float(fov_deg), int(checkpoint_interval), int(log_steps_per_ray),
                               int(adjoint_precision), 0)
        h = C.c_void_p()
        check(lib().kt_grad_create(C.byref(self.cfg), C.byref(self.opts), C.byref(h)))
        self._h = h
        self.frame = 0  # forward passes so far; a backward belongs to the last one

    def close(self):
        if getattr(self, "_h", None):
            lib().kt_grad_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        return (self.cfg.height, self.cfg.width)

    def forward(self, theta, hdr=None, cls=None, stream=None):
        """Trace the frame at theta (4 floats); returns (hdr float32 [H, W, 4],
        cls uint8 [H, W] or None) on the current torch stream."""
        from .renderer import _check_tensor
        torch = _torch()
        th = (C.c_double * 4)(*[float(v) for v in theta])
        H, W = self.shape
        s = stream if stream is not None else torch.cuda.current_stream()
        if hdr is None:
            with torch.cuda.stream(s):
                hdr = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
        _check_tensor(hdr, (H, W, 4), torch.float32)
        if cls is not None:
            _check_tensor(cls, (H, W), torch.uint8)
        check(lib().kt_grad_forward(self._h, th, C.c_void_p(hdr.data_ptr()),
                                    C.c_void_p(cls.data_ptr() if cls is not None else 0),
                                    C.c_void_p(s.cuda_stream)))
        self.frame += 1
        return hdr, cls

    def backward(self, dl_dhdr, per_pixel=False, stream=None):
        """d(sum dl_dhdr * hdr)/d theta for the last forward frame: float64 [4]
        on the device (and the per-pixel contributions [H, W, 4] if asked)."""
        torch = _torch()
        H, W = self.shape
        s = stream if stream is not None else torch.cuda.current_stream()
        cur = torch.cuda.current_stream()
        dl = dl_dhdr.detach()
        if dl.numel() != H * W * 4:
            raise ValueError(f"dl_dhdr must hold {H}x{W}x4 values, got shape {tuple(dl.shape)}")
        if s != cur:
            s.wait_stream(cur)  # dl_dhdr is produced on the current stream
        with torch.cuda.stream(s):
            # conversions and outputs are made on s, so their memory is
            # ordered with the adjoint kernels that use it
            dl = dl.to(device="cuda", dtype=torch.float32).reshape(H, W, 4).contiguous()
            grad = torch.empty(4, dtype=torch.float64, device="cuda")
            gpx = torch.empty((H, W, 4), dtype=torch.float64, device="cuda") if per_pixel else None
        if dl.data_ptr() == dl_dhdr.data_ptr() and s != cur:
            dl_dhdr.record_stream(s)  # the caller's own buffer is read on s
        check(lib().kt_grad_backward(self._h, C.c_void_p(dl.data_ptr()), C.c_void_p(grad.data_ptr()),
                                     C.c_void_p(gpx.data_ptr() if gpx is not None else 0),
                                     C.c_void_p(s.cuda_stream)))
        return (grad, gpx) if per_pixel else grad

    def stats(self):
        st = KtGradStats()
        check(lib().kt_grad_get_stats(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}


def _function():
    torch = _torch()

    class KerrRender(torch.autograd.Function):
        """hdr = render(theta): float32 [H, W, 4] from float64 theta [4]."""

        @staticmethod
        def forward(ctx, theta, renderer):
            values = theta.detach().to("cpu", torch.float64).tolist()
            hdr, _ = renderer.forward(values)
            ctx.renderer = renderer
            ctx.frame = renderer.frame
            ctx.theta_meta = (theta.device, theta.dtype)
            return hdr

        @staticmethod
        def backward(ctx, grad_hdr):
            r = ctx.renderer
            if r.frame != ctx.frame:
                raise RuntimeError("the renderer traced another frame after this one: run backward before the "
                                   "next forward (the adjoint replays the last frame's step log)")
            g = r.backward(grad_hdr)
            device, dtype = ctx.theta_meta
            return g.to(device=device, dtype=dtype), None

    return KerrRender


_KERR_RENDER = None


def render(theta, renderer):
    """Differentiable render: theta (tensor [4], requires_grad as wanted) ->
    hdr float32 [H, W, 4]."""
    global _KERR_RENDER
    if _KERR_RENDER is None:
        _KERR_RENDER = _function()
    return _KERR_RENDER.apply(theta, renderer)


def _gauss_blur(img, sigma):
    """Separable Gaussian blur (radius 3 sigma, edge pixels repeated) of an
    [H, W, C] float64 image, differentiable through torch."""
    torch = _torch()
    F = torch.nn.functional
    r = max(1, int(3.0 * sigma))
    x = torch.arange(-r, r + 1, dtype=torch.float64, device=img.device)
    w = torch.exp(-0.5 * (x / sigma) ** 2)
    w = w / w.sum()
    t = img.permute(2, 0, 1).unsqueeze(1)
    t = F.conv2d(F.pad(t, (r, r, 0, 0), mode="replicate"), w.view(1, 1, 1, -1))
    t = F.conv2d(F.pad(t, (0, 0, r, r), mode="replicate"), w.view(1, 1, -1, 1))
    return t.squeeze(1).permute(1, 2, 0)


def fit_adam(renderer, target_hdr, theta0, iters=200, step=(0.01, 0.01, 0.5, 0.01), hold=0.6, final_lr=0.1,
             blur_px=8.0, clip=3.0, betas=(0.9, 0.99), log_every=0):
    """Fit theta to a target frame with Adam on the L2 loss over the linear
    HDR colour channels.  Adam's step is about its learning rate in each
    coordinate whatever the gradient's scale, so theta is optimised as
    theta = step * z on z: `step` is the per-iteration step size in spin,
    inclination [rad], r_obs [M] and mdot.  The learning rate stays 1 for
    the first `hold` fraction of the iterations, then falls linearly to
    `final_lr`.

    Coarse to fine: iteration k descends the L2 loss of both frames blurred
    by a Gaussian of sigma = blur_px * (1 - k / (iters - 1)) pixels (the
    plain L2 loss once sigma < 0.25 px, so the last iterations descend the
    loss as stated).  The loss is only piecewise smooth: pixels that change
    class (disk, shadow, sky) make it jump, and the frozen-event gradient
    does not see those jumps; blurring spreads each jump over a few pixels
    and damps it against the smooth part.

    Spikes: a pixel whose orbit is chaotic in theta (near the photon ring)
    now and then gives a frozen-event gradient ~100x the usual; one such
    spike inflates Adam's second moment and the steps collapse for the
    rest of the fit.  From the 6th iteration on, each gradient component
    (of z) is clipped to `clip` times the median of its magnitudes over
    the previous 10 iterations (clip=0: off), and the second moment's
    decay is 0.99.

    On the 512x512 scene, without blur and clipping the fit ends at
    different points for starts 1e-7 apart (spin off by 0.02 to 0.24);
    with them, starts 0, +-1e-7 and 2e-7 away end within 0.001 in spin,
    0.02 deg, 0.02 M and 0.002 in mdot of the target, with the FP64 pair
    and with the mixed forward + FP32 adjoint.  The iterate of lowest plain
    L2 loss is returned.

    Returns (theta [4] list of the best iterate, history of (plain L2 loss,
    theta) before each update)."""
    torch = _torch()
    scale = torch.tensor(step, dtype=torch.float64)
    z = (torch.tensor([float(v) for v in theta0], dtype=torch.float64) / scale).requires_grad_(True)
    opt = torch.optim.Adam([z], lr=1.0, betas=betas)
    n_hold = int(hold * iters)
    target = target_hdr[..., :3].detach().to(torch.float64)
    history = []
    grads = []  # per-iteration |gradient| of z, for the clipping
    best = (float("inf"), [float(v) for v in theta0])
    for it in range(iters):
        frac = 0.0 if it < n_hold else (it - n_hold) / max(iters - 1 - n_hold, 1)
        for group in opt.param_groups:
            group["lr"] = 1.0 + (final_lr - 1.0) * frac
        sigma = blur_px * (1.0 - it / max(iters - 1, 1))
        opt.zero_grad()
        theta = z * scale
        hdr = render(theta, renderer)[..., :3].to(torch.float64)
        plain = ((hdr - target) ** 2).sum()
        loss = ((_gauss_blur(hdr, sigma) - _gauss_blur(target, sigma)) ** 2).sum() if sigma >= 0.25 else plain
        loss.backward()
        gabs = z.grad.detach().abs()
        if clip > 0 and len(grads) >= 5:
            lim = clip * torch.stack(grads[-10:]).median(dim=0).values
            z.grad.copy_(torch.maximum(torch.minimum(z.grad, lim), -lim))
        grads.append(gabs)
        history.append((float(plain.detach()), [float(v) for v in theta.detach()]))
        if history[-1][0] < best[0]:
            best = history[-1]
        opt.step()
        with torch.no_grad():
            z[0].clamp_(0.0, 0.999 / step[0])
            z[3].clamp_(min=1e-3 / step[3])
        if log_every and it % log_every == 0:
            print(f"[adam] it {it} loss {history[-1][0]:.6e} theta {history[-1][1]}", flush=True)
    return list(best[1]), history


def autotune_checkpoint(cfg, theta, intervals=(8, 16, 32, 64), fov_deg=GRAD_FOV_DEG, log_steps_per_ray=512, reps=2,
                        adjoint_precision=KT_ADJOINT_FP64):
    """Checkpoint interval whose adjoint pass is fastest for this scene:
    for each interval a renderer traces theta and runs the backward pass
    (unit loss weights) `reps` times.  Returns (best interval, {interval:
    best backward ms}).  Gradients do not depend on the interval."""
    torch = _torch()
    times = {}
    for k in intervals:
        with GradRenderer(cfg, fov_deg=fov_deg, checkpoint_interval=k, log_steps_per_ray=log_steps_per_ray,
                          adjoint_precision=adjoint_precision) as g:
            hdr, _ = g.forward(theta)
            w = torch.ones_like(hdr)
            best = float("inf")
            for _ in range(reps):
                g.backward(w)
                best = min(best, g.stats()["backward_ms"])
            times[k] = best
    return min(times, key=times.get), times
