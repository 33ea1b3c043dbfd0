"""Pins of the annual-parallax trajectory oracle (src/caustics/trajectory.py) by closed
forms that do not reuse its code: the sky basis of a source at (ra, dec), a circular
Earth orbit whose projected Sun position is a sine, numpy.interp, and the rectilinear
limit.  The reference's own test compares with MulensModel and astropy ephemerides
(tests/test_trajectory.py:10-45), both absent here."""
import numpy as np
import pytest

from oracle import trajectory as T


def closed_form_sun(times, ra_deg, dec_deg, R=1.0, period=365.25, phase=0.3, incl=0.4090928):
    """Sun relative to the Earth, projected east / north, for the circular orbit of
    T.circular_orbit: east = (-sin a, cos a, 0), north = (-sin d cos a, -sin d sin a,
    cos d), r = R (cos f, sin f cos i, sin f sin i)."""
    a, d = np.deg2rad(ra_deg), np.deg2rad(dec_deg)
    om = 2 * np.pi / period
    f = om * np.asarray(times, float) + phase
    ci, si = np.cos(incl), np.sin(incl)
    rx, ry, rz = np.cos(f), np.sin(f) * ci, np.sin(f) * si
    vx, vy, vz = -np.sin(f) * om, np.cos(f) * ci * om, np.cos(f) * si * om
    se = -R * (-rx * np.sin(a) + ry * np.cos(a))
    sn = -R * (-rx * np.sin(d) * np.cos(a) - ry * np.sin(d) * np.sin(a) + rz * np.cos(d))
    sde = -R * (-vx * np.sin(a) + vy * np.cos(a))
    sdn = -R * (-vx * np.sin(d) * np.cos(a) - vy * np.sin(d) * np.sin(a) + vz * np.cos(d))
    return se, sn, sde, sdn


def test_sky_basis_on_the_equator():
    for ra in (0.0, 37.0, 200.0):
        ee, en = T.sky_basis(T.direction_from_radec(ra, 0.0))
        a = np.deg2rad(ra)
        np.testing.assert_allclose(ee, [-np.sin(a), np.cos(a), 0.0], atol=1e-15)
        np.testing.assert_allclose(en, [0.0, 0.0, 1.0], atol=1e-15)


def test_sky_basis_is_orthonormal_and_points_north():
    g = np.random.default_rng(0)
    for _ in range(20):
        d = T.direction_from_radec(g.uniform(0, 360), g.uniform(-80, 80))
        ee, en = T.sky_basis(d)
        assert abs(ee @ d) < 1e-15 and abs(en @ d) < 1e-15 and abs(ee @ en) < 1e-15
        np.testing.assert_allclose([ee @ ee, en @ en], 1.0, atol=1e-15)
        assert en[2] > 0
        # (east, north, d) is left-handed on the sky as seen from inside: east x north = d
        np.testing.assert_allclose(np.cross(ee, en), d, atol=1e-15)


def test_interp_matches_numpy_inside_and_clamps_outside():
    g = np.random.default_rng(1)
    xp = np.arange(8000.37, 8300.37, 1.0)
    fp = g.normal(size=xp.size)
    x = g.uniform(xp[0], xp[-1], 1000)
    np.testing.assert_allclose(T.interp(x, xp, fp), np.interp(x, xp, fp), rtol=1e-14, atol=1e-14)
    np.testing.assert_array_equal(T.interp(xp, xp, fp)[:-1], fp[:-1])
    np.testing.assert_array_equal(T.interp([xp[0] - 3.0, xp[-1] + 5.0], xp, fp), [fp[0], fp[-1]])


@pytest.mark.parametrize("ra,dec", [(271.19, -26.99), (30.0, 10.0), (120.0, -60.0)])
def test_delta_sun_matches_circular_orbit(ra, dec):
    xp = np.arange(8000.25, 8401.25, 1.0)
    pos, vel = T.circular_orbit(xp)
    tab = T.sun_table(pos, vel, T.direction_from_radec(ra, dec))
    ref = np.stack(closed_form_sun(xp, ra, dec))
    np.testing.assert_allclose(tab, ref, atol=1e-14)
    # on grid epochs the interpolation is exact: delta_sun equals the closed form
    t0 = xp[150]
    t = xp[::7]
    se, sn, sde, sdn = closed_form_sun(t, ra, dec)
    se0, sn0, sde0, sdn0 = closed_form_sun(t0, ra, dec)
    de, dn = T.delta_sun(t, t0, xp, tab)
    np.testing.assert_allclose(de, se - se0 - (t - t0) * sde0, atol=1e-12)
    np.testing.assert_allclose(dn, sn - sn0 - (t - t0) * sdn0, atol=1e-12)
    # between grid epochs: within the linear-interpolation error R w^2 h^2 / 8 of the
    # positions and of the velocity at t0 times |t - t0|
    t0 = 8150.6
    t = np.linspace(8001.0, 8399.0, 517)
    se, sn, sde, sdn = closed_form_sun(t, ra, dec)
    se0, sn0, sde0, sdn0 = closed_form_sun(t0, ra, dec)
    de, dn = T.delta_sun(t, t0, xp, tab)
    om = 2 * np.pi / 365.25
    bound = 2 * om ** 2 / 8 + np.abs(t - t0) * om ** 3 / 8 + 1e-12
    assert np.all(np.abs(de - (se - se0 - (t - t0) * sde0)) <= bound)
    assert np.all(np.abs(dn - (sn - sn0 - (t - t0) * sdn0)) <= bound)
    assert np.abs(de).max() > 0.1   # the parallax offset is not trivially small here


def test_trajectory_limits():
    xp = np.arange(8000.0, 8200.0, 1.0)
    pos, vel = T.circular_orbit(xp)
    tab = T.sun_table(pos, vel, T.direction_from_radec(270.0, -30.0))
    t = np.linspace(7990.0, 8210.0, 301)
    g = np.random.default_rng(2)
    for _ in range(5):
        u0, tE, psi, t0 = g.uniform(-1, 1), g.uniform(10, 100), g.uniform(-3, 3), g.uniform(8010, 8190)
        # pi_E = 0: straight line, |u|^2 = u0^2 + tau^2
        w = T.trajectory(t, xp, tab, "polar", t0=t0, u0=u0, tE=tE, piE=0.0, psi=psi)
        np.testing.assert_allclose(np.abs(w) ** 2, u0 ** 2 + ((t - t0) / tE) ** 2, rtol=1e-13)
        # at t0 the parallax offset vanishes for any pi_E: w(t0) = u0 (cos psi - i sin psi)
        w0 = T.trajectory([t0], xp, tab, "polar", t0=t0, u0=u0, tE=tE, piE=0.7, psi=psi)
        np.testing.assert_allclose(w0, u0 * np.exp(-1j * psi), atol=1e-15)
        # the two parametrizations agree
        piE = g.uniform(0.05, 1.0)
        wc = T.trajectory(t, xp, tab, "cartesian", t0=t0, u0=u0, tE=tE,
                          piEE=piE * np.sin(psi), piEN=piE * np.cos(psi))
        wp = T.trajectory(t, xp, tab, "polar", t0=t0, u0=u0, tE=tE, piE=piE, psi=psi)
        np.testing.assert_allclose(wc, wp, rtol=1e-13, atol=1e-13)
