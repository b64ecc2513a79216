"""Differentiable seepage window over the C-ABI (sf_diff_*): numpy and torch.

    win = DiffWindow(scene, n_steps=200, checkpoint_every=20)
    pos, sat = win.forward(theta, vel0)          # theta = (sf_cd, sf_c0, sf_cmc, tan friction angle)
    g_theta, g_pos0, g_vel0 = win.backward(pos_bar, sat_bar)

    pos, sat = seepage_window(win, theta_t, vel0_t)   # torch.autograd.Function on cuda tensors

The forward and the reverse passes both run in libsf_seepage.so on the GPU;
this module only moves arrays (host numpy or device torch pointers) across
the boundary.  Arrays are in the scene's particle order.  The physical
controls map to theta as cd = g rho0 / D (drag D), c0 = 0.7 sigma and
cmc = sigma (capillary sigma), tan friction angle = mu; `physical_grad`
applies that chain rule.
"""
from __future__ import annotations

import copy
import ctypes as C

import numpy as np

from . import _abi
from ._abi import FIELDS, check, lib, ptr
from .system import SFSystem

THETA = ("sf_cd", "sf_c0", "sf_cmc", "dem_tan_friction_angle")
GRAV_RHO0 = 9.81 * 1000.0


def theta_of(D: float, sigma: float, mu: float) -> np.ndarray:
    return np.array([GRAV_RHO0 / D, 0.7 * sigma, sigma, mu], np.float32)


def physical_grad(theta_bar, D: float) -> dict:
    """d loss / d (drag D, capillary sigma, friction mu) from d loss / d theta."""
    tb = np.asarray(theta_bar, np.float64)
    return {"drag_D": float(tb[0] * (-GRAV_RHO0 / (D * D))), "capillary_sigma": float(0.7 * tb[1] + tb[2]),
            "friction_mu": float(tb[3])}


class DiffWindow:
    """A fixed-dt window of `n_steps` steps of one scene with gradients."""

    def __init__(self, scene, n_steps: int, dt: float = 1e-4, checkpoint_every: int = 20, device: int = -1):
        self.scene = scene
        self.n = scene.n
        self.n_steps = int(n_steps)
        self.dt = float(dt)
        self.every = int(checkpoint_every)
        self.sys = SFSystem.from_scene(scene, device=device)
        self._p0 = copy.copy(self.sys._p)
        self._ran = False

    @property
    def handle(self):
        return self.sys._h

    def set_theta(self, theta) -> None:
        t = np.asarray(theta, np.float64).reshape(4)
        p = copy.copy(self._p0)
        for name, v in zip(THETA, t):
            setattr(p, name, float(v))
        self.sys.set_params(p)

    def set_stream(self, stream: int | None) -> None:
        check(lib().sf_set_stream(self.handle, stream))

    def run(self, theta=None, vel0=None, vel0_device: int | None = None) -> None:
        """Runs the window from the scene's initial state (vel0: host (n,3)
        array, or a device pointer to n*3 floats)."""
        if theta is not None:
            self.set_theta(theta)
        if self._ran:
            if vel0_device is not None:
                check(lib().sf_diff_reset(self.handle, vel0_device, 1))
            else:
                v = None if vel0 is None else np.ascontiguousarray(vel0, np.float32)
                check(lib().sf_diff_reset(self.handle, ptr(v), 0))
        elif vel0 is not None or vel0_device is not None:
            v = vel0 if vel0 is not None else None
            if v is None:  # first run with a device array: stage through the host once
                raise ValueError("the first run takes host velocities (or none)")
            self.sys.set_field("vel", np.ascontiguousarray(v, np.float32))
        check(lib().sf_diff_forward(self.handle, self.dt, self.n_steps, self.every))
        self._ran = True

    def forward(self, theta=None, vel0=None):
        self.run(theta, vel0)
        pos = np.zeros((self.n, 3), np.float32)
        sat = np.zeros(self.n, np.float32)
        check(lib().sf_diff_get(self.handle, FIELDS["pos"][0], ptr(pos), 0))
        check(lib().sf_diff_get(self.handle, FIELDS["saturation"][0], ptr(sat), 0))
        return pos, sat

    def backward(self, pos_bar=None, sat_bar=None, vel_bar=None):
        f = lambda a, shape: None if a is None else np.ascontiguousarray(np.broadcast_to(a, shape), np.float32)
        pb, sb, vb = f(pos_bar, (self.n, 3)), f(sat_bar, (self.n,)), f(vel_bar, (self.n, 3))
        th = np.zeros(4, np.float32)
        gp = np.zeros((self.n, 3), np.float32)
        gv = np.zeros((self.n, 3), np.float32)
        check(lib().sf_diff_backward(self.handle, ptr(pb), ptr(vb), ptr(sb), ptr(th), ptr(gp), ptr(gv), 0))
        return th, gp, gv


def msd_loss(pos, sat, target_pos, target_sat, sand):
    """Mean squared deviation of the sand's final positions and saturations,
    and its gradients (zero on water)."""
    ns = max(int(np.count_nonzero(sand)), 1)
    dp = np.where(sand[:, None], pos.astype(np.float64) - target_pos, 0.0)
    ds = np.where(sand, sat.astype(np.float64) - target_sat, 0.0)
    loss = float((dp * dp).sum() / ns + (ds * ds).sum() / ns)
    return loss, (2.0 * dp / ns).astype(np.float32), (2.0 * ds / ns).astype(np.float32)


try:  # torch is optional for the C-ABI, required only for the autograd wrapper
    import torch

    class SeepageWindowFunction(torch.autograd.Function):
        """(theta[4], vel0[n,3]) -> (pos[n,3], sat[n]) after the window, on cuda
        tensors; backward runs the library's reverse passes."""

        @staticmethod
        def forward(ctx, window: DiffWindow, theta, vel0):
            dev = vel0.device
            window.set_stream(torch.cuda.current_stream(dev).cuda_stream)
            v = vel0.detach().to(torch.float32).contiguous()
            if window._ran:
                window.run(theta.detach().cpu().numpy(), vel0_device=v.data_ptr())
            else:
                window.run(theta.detach().cpu().numpy(), vel0=v.cpu().numpy())
            pos = torch.empty((window.n, 3), dtype=torch.float32, device=dev)
            sat = torch.empty((window.n,), dtype=torch.float32, device=dev)
            check(lib().sf_diff_get(window.handle, FIELDS["pos"][0], pos.data_ptr(), 1))
            check(lib().sf_diff_get(window.handle, FIELDS["saturation"][0], sat.data_ptr(), 1))
            ctx.window = window
            ctx.theta_device = theta.device
            return pos, sat

        @staticmethod
        def backward(ctx, gpos, gsat):
            w = ctx.window
            dev = gpos.device if gpos is not None else gsat.device
            w.set_stream(torch.cuda.current_stream(dev).cuda_stream)
            gp = None if gpos is None else gpos.detach().to(torch.float32).contiguous()
            gs = None if gsat is None else gsat.detach().to(torch.float32).contiguous()
            th = np.zeros(4, np.float32)
            gv = torch.empty((w.n, 3), dtype=torch.float32, device=dev)
            check(lib().sf_diff_backward(w.handle, None if gp is None else gp.data_ptr(), None,
                                         None if gs is None else gs.data_ptr(), ptr(th), None, gv.data_ptr(), 1))
            return None, torch.from_numpy(th).to(ctx.theta_device), gv

    def seepage_window(window: DiffWindow, theta, vel0):
        return SeepageWindowFunction.apply(window, theta, vel0)

except ImportError:  # pragma: no cover
    torch = None
