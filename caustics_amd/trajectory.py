"""Source trajectories with annual parallax on the GPU (reference:
src/caustics/trajectory.py).

``AnnualParallaxTrajectory(t, coords)`` keeps the reference's constructor and
``compute(t, parametrization="cartesian", **params)`` (trajectory.py:12-157).  At
construction the Sun's position and velocity relative to the Earth are tabulated on the
daily grid ``t_jpl = arange(t[0], t[-1] + 1, 1)`` (:36) and projected east / north on
the plane of the sky of ``coords`` (:46-102), as the reference does in NumPy.  Evaluating
the trajectory is one HIP kernel (``caus_trajectory_annual_parallax``) for any number of
parameter sets: ``compute`` takes scalars (one trajectory of shape ``t.shape``, as the
reference) or arrays of L parameter sets (shape ``(L,) + t.shape``), and is
differentiable in every parameter (the kernel's reverse mode), so gradients flow on
into ``mag_point_source`` / ``mag`` of the resulting ``w``.

The reference takes the Earth's barycentric position and velocity from astropy
(``get_body_barycentric_posvel("earth", Time(t_jpl + 2450000, format="jd",
scale="tdb"))``, :85-93).  The same call is made here when astropy is importable;
otherwise the caller passes that table as ``earth_pos`` / ``earth_vel`` ((M, 3) arrays
in AU and AU/day on ``t_jpl``).
"""
import numpy as np
import torch

from . import _lib

__all__ = ["AnnualParallaxTrajectory"]

_PARAM_ORDER = {"cartesian": ("t0", "u0", "tE", "piEE", "piEN"),
                "polar": ("t0", "u0", "tE", "piE", "psi")}


def _direction(coords):
    """Unit vector to the source: an astropy SkyCoord (its ICRS cartesian xyz, as the
    reference reads it at :62), an (ra, dec) pair in degrees, or a 3-vector."""
    if hasattr(coords, "cartesian"):
        d = np.asarray(coords.cartesian.xyz.value, np.float64)
    else:
        c = np.asarray(coords, np.float64).reshape(-1)
        if c.size == 2:
            ra, dec = np.deg2rad(c)
            d = np.array([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)])
        elif c.size == 3:
            d = c
        else:
            raise ValueError("coords: a SkyCoord, (ra_deg, dec_deg) or a 3-vector")
    n = np.linalg.norm(d)
    if not np.isfinite(n) or n == 0:
        raise ValueError("coords: direction of zero length")
    return d / n


def _earth_posvel_astropy(t_jpl):
    try:
        from astropy.coordinates import get_body_barycentric_posvel
        from astropy.time import Time
    except ImportError as e:
        raise ImportError("astropy is not importable: pass the Earth's barycentric "
                          "position and velocity on t_jpl as earth_pos / earth_vel") from e
    pos, vel = get_body_barycentric_posvel("earth", Time(t_jpl + 2450000, format="jd",
                                                         scale="tdb"))
    return pos.xyz.value.T, vel.xyz.value.T


class _Parallax(torch.autograd.Function):
    @staticmethod
    def forward(ctx, P, t, xp, tab, polar):
        L, T = P.shape[0], t.numel()
        w = torch.empty((L, T), dtype=torch.complex128, device=P.device)
        rc = _lib.lib().caus_trajectory_annual_parallax(
            _lib.stream_ptr(), _lib.ptr(t), T, _lib.ptr(xp), _lib.ptr(tab), xp.numel(),
            _lib.ptr(P), L, polar, _lib.ptr(w), None, None)
        _lib.check(rc, "caus_trajectory_annual_parallax")
        ctx.save_for_backward(P, t, xp, tab)
        ctx.polar = polar
        return w

    @staticmethod
    def backward(ctx, g_w):
        P, t, xp, tab = ctx.saved_tensors
        g = g_w.to(torch.complex128).contiguous()
        gP = torch.empty_like(P)
        rc = _lib.lib().caus_trajectory_annual_parallax(
            _lib.stream_ptr(), _lib.ptr(t), t.numel(), _lib.ptr(xp), _lib.ptr(tab), xp.numel(),
            _lib.ptr(P), P.shape[0], ctx.polar, None, _lib.ptr(torch.view_as_real(g)),
            _lib.ptr(gP))
        _lib.check(rc, "caus_trajectory_annual_parallax")
        return gP, None, None, None, None


class AnnualParallaxTrajectory:
    def __init__(self, t, coords, earth_pos=None, earth_vel=None):
        """t: observation times (HJD - 2450000); coords: the source direction."""
        self.t = np.asarray(t, np.float64)
        self.t_jpl = np.arange(self.t[0], self.t[-1] + 1, 1)
        if self.t_jpl.size < 2:
            raise ValueError("the time grid needs at least two days")
        self.coords = coords
        if earth_pos is None or earth_vel is None:
            earth_pos, earth_vel = _earth_posvel_astropy(self.t_jpl)
        earth_pos = np.asarray(earth_pos, np.float64)
        earth_vel = np.asarray(earth_vel, np.float64)
        if earth_pos.shape != (self.t_jpl.size, 3) or earth_vel.shape != earth_pos.shape:
            raise ValueError(f"earth_pos / earth_vel must be ({self.t_jpl.size}, 3) on t_jpl")
        d = _direction(coords)
        # plane-of-sky basis at the source (:60-73): east = z x d, north = d x east
        e_east = np.cross([0.0, 0.0, 1.0], d)
        e_north = np.cross(d, e_east)
        e_east /= np.linalg.norm(e_east)
        e_north /= np.linalg.norm(e_north)
        # the Sun relative to the Earth: minus the Earth's barycentric vectors (:92-93)
        self.s_e, self.s_n = -earth_pos @ e_east, -earth_pos @ e_north
        self.s_e_dot, self.s_n_dot = -earth_vel @ e_east, -earth_vel @ e_north
        self._xp = torch.as_tensor(self.t_jpl, device="cuda")
        self._tab = torch.as_tensor(np.stack([self.s_e, self.s_n, self.s_e_dot, self.s_n_dot]),
                                    device="cuda").contiguous()

    def compute(self, t, parametrization="cartesian", **params):
        """Source positions u_e + i u_n at times t (trajectory.py:120-157).  params:
        t0, u0, tE and piEE, piEN ("cartesian") or piE, psi ("polar"); each a scalar or
        an array / tensor of L parameter sets."""
        if parametrization not in _PARAM_ORDER:
            raise ValueError("Invalid parametrization. Choose from 'polar' (piE, psi) or "
                             "'cartesian' (piEE, piEN).")
        names = _PARAM_ORDER[parametrization]
        missing = [k for k in names if k not in params]
        if missing:
            raise ValueError(f"missing parameters {missing}")
        vals = [v if torch.is_tensor(v) else torch.as_tensor(np.asarray(v, np.float64))
                for v in (params[k] for k in names)]
        vals = [v.to(device="cuda", dtype=torch.float64).reshape(-1) for v in vals]
        L = max(v.numel() for v in vals)
        if any(v.numel() not in (1, L) for v in vals):
            raise ValueError("parameter arrays must share one length")
        batched = any(v.numel() == L and L > 1 for v in vals) or any(
            np.ndim(params[k]) > 0 for k in names)
        P = torch.stack([v.expand(L) for v in vals], -1).contiguous()
        tt = t if torch.is_tensor(t) else torch.as_tensor(np.asarray(t, np.float64))
        shape = tuple(tt.shape)
        tt = tt.to(device="cuda", dtype=torch.float64).reshape(-1).contiguous()
        w = _Parallax.apply(P, tt, self._xp, self._tab, 1 if parametrization == "polar" else 0)
        return w.reshape((L,) + shape) if batched else w.reshape(shape)
