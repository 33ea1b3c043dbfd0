"""NumPy FP64 restatement of the annual-parallax source trajectory.
TEST INFRASTRUCTURE ONLY.

Follows (fbartolic/caustics, src/caustics/trajectory.py):
  sky_basis ............ _project_vector_onto_sky :46-77 (east = z x d, north = d x east,
                         both normalised; d the unit vector to the source)
  sun_table ............ _compute_sun_position_and_velocity :80-102 (the Earth's
                         barycentric position / velocity negated, then projected)
  interp ............... jax.numpy.interp as the reference calls it (:106-112).  JAX is a
                         dependency of the reference without a pinned version
                         (setup.py:113); its published implementation is restated:
                         i = clip(searchsorted(xp, x, 'right'), 1, M-1),
                         f = fp[i-1] + ((x - xp[i-1]) / dx) * df (f = fp[i-1] when
                         |dx| <= spacing(eps)), fp[0] below xp[0], fp[-1] above xp[-1]
  delta_sun ............ _compute_delta_sun_position_and_velocity :104-118
  trajectory ........... compute :120-157 (cartesian: psi = atan2(piEE, piEN),
                         piE = sqrt(piEN^2 + piEE^2); polar: piE, psi given)
The ephemeris itself (astropy's get_body_barycentric_posvel, :88-93) is an input here:
astropy is absent, so the tables the tests use are analytic orbits.

Pins (tests/test_oracle_trajectory.py): the sky basis in closed form for sources on
the equator and its orthonormality elsewhere; interp against numpy.interp inside the
grid and the clamped ends outside; delta_sun against the closed form of a circular
Earth orbit (exact on grid epochs, within the linear-interpolation bound between);
delta_sun(t0) = 0 and the rectilinear limit pi_E = 0.
"""
import numpy as np


def direction_from_radec(ra_deg, dec_deg):
    ra, dec = np.deg2rad(ra_deg), np.deg2rad(dec_deg)
    return np.array([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)])


def sky_basis(direction):
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    e_north = np.array([0.0, 0.0, 1.0])
    e_east_sky = np.cross(e_north, d)
    e_north_sky = np.cross(d, e_east_sky)
    return e_east_sky / np.linalg.norm(e_east_sky), e_north_sky / np.linalg.norm(e_north_sky)


def sun_table(earth_pos, earth_vel, direction):
    """(4, M) rows s_e, s_n, sdot_e, sdot_n from the Earth's barycentric position (M, 3)
    and velocity (M, 3)."""
    ee, en = sky_basis(direction)
    s_t = -np.asarray(earth_pos, np.float64)
    v_t = -np.asarray(earth_vel, np.float64)
    return np.stack([s_t @ ee, s_t @ en, v_t @ ee, v_t @ en])


def interp(x, xp, fp):
    x = np.asarray(x, np.float64)
    xp = np.asarray(xp, np.float64)
    fp = np.asarray(fp, np.float64)
    i = np.clip(np.searchsorted(xp, x, side="right"), 1, len(xp) - 1)
    df = fp[i] - fp[i - 1]
    dx = xp[i] - xp[i - 1]
    delta = x - xp[i - 1]
    dx0 = np.abs(dx) <= np.spacing(np.finfo(np.float64).eps)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = np.where(dx0, fp[i - 1], fp[i - 1] + (delta / np.where(dx0, 1.0, dx)) * df)
    f = np.where(x < xp[0], fp[0], f)
    f = np.where(x > xp[-1], fp[-1], f)
    return f


def delta_sun(t, t0, xp, tab):
    s_e_t = interp(t, xp, tab[0])
    s_n_t = interp(t, xp, tab[1])
    s_e_t0 = interp(t0, xp, tab[0])
    s_n_t0 = interp(t0, xp, tab[1])
    s_e_dot_t0 = interp(t0, xp, tab[2])
    s_n_dot_t0 = interp(t0, xp, tab[3])
    delta_s_e = s_e_t - s_e_t0 - (t - t0) * s_e_dot_t0
    delta_s_n = s_n_t - s_n_t0 - (t - t0) * s_n_dot_t0
    return delta_s_e, delta_s_n


def trajectory(t, xp, tab, parametrization="cartesian", **p):
    """w = u_e + i u_n at epochs t for one parameter set (t0, u0, tE and piEE, piEN or
    piE, psi)."""
    t = np.asarray(t, np.float64)
    if parametrization == "polar":
        piE, psi = p["piE"], p["psi"]
    elif parametrization == "cartesian":
        psi = np.arctan2(p["piEE"], p["piEN"])
        piE = np.sqrt(p["piEN"] ** 2 + p["piEE"] ** 2)
    else:
        raise ValueError(parametrization)
    dse, dsn = delta_sun(t, p["t0"], xp, tab)
    u_e = p["u0"] * np.cos(psi) + (t - p["t0"]) / p["tE"] * np.sin(psi) + piE * dse
    u_n = -p["u0"] * np.sin(psi) + (t - p["t0"]) / p["tE"] * np.cos(psi) + piE * dsn
    return u_e + 1j * u_n


def circular_orbit(times, radius=1.0, period=365.25, phase=0.3, incl=0.4090928):
    """Analytic Earth orbit for the tests: a circle of `radius` AU in a plane tilted by
    `incl` about x; position (M, 3) AU and velocity (M, 3) AU/day at `times`."""
    om = 2 * np.pi / period
    a = om * np.asarray(times, np.float64) + phase
    c, s = np.cos(a), np.sin(a)
    ci, si = np.cos(incl), np.sin(incl)
    pos = radius * np.stack([c, s * ci, s * si], -1)
    vel = radius * om * np.stack([-s, c * ci, c * si], -1)
    return pos, vel
