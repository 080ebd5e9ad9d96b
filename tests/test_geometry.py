"""CPU checks of the FP64 host geometry (orbit interpolation, WGS84 geolocation, incidence angle, coordinate
conversions, processing velocity, slant-to-ground fit, scene centre) through the C ABI.

The reference has no tests or fixtures for these functions (sar/orbit_state_vector.cc, util/src/geo_tools.cc,
sar/processing_velocity_estimation.cc, sar/srgr.cu:73-107, main_flow.cc:499-522), so each one is pinned by geometry
with a closed-form answer: polynomial and straight-line trajectories (exact under Lagrange interpolation), an
equatorial sensor moving along the polar axis (the zero-Doppler plane cuts the ellipsoid in a circle), a sensor on
the ellipsoid normal (incidence 0) and the law of cosines. The library and the numpy restatement in the oracle are
also compared on an inclined orbit over a rotating Earth."""
import math

import numpy as np
import pytest

import pyoracle as po
from asarf import _lib as L
from asarf import focus

A, B = po.WGS84_A, po.WGS84_B


# ----------------------------------------------------------------------------------------------- orbits ----
def rotating_orbit(times, radius=7.159e6, incl_deg=98.55, phase0=0.75):
    """Circular inclined orbit seen from an Earth-fixed frame (Earth rotation 7.2921159e-5 rad/s)."""
    gm, we = 3.986004418e14, 7.2921159e-5
    w = math.sqrt(gm / radius ** 3)
    ci, si = math.cos(math.radians(incl_deg)), math.sin(math.radians(incl_deg))
    pos, vel = [], []
    for t in times:
        u = phase0 + w * t
        pi = radius * np.array([math.cos(u), math.sin(u) * ci, math.sin(u) * si])
        vi = radius * w * np.array([-math.sin(u), math.cos(u) * ci, math.cos(u) * si])
        c, s = math.cos(we * t), math.sin(we * t)
        rot = np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])
        drot = we * np.array([[-s, c, 0.0], [-c, -s, 0.0], [0.0, 0.0, 0.0]])
        pos.append(rot @ pi)
        vel.append(rot @ vi + drot @ pi)
    return np.array(pos), np.array(vel)


def right_looking_guess(p, v, offset=330e3):
    up = p / np.linalg.norm(p)
    right = np.cross(v / np.linalg.norm(v), up)
    g = up * A + right * offset
    return g * (A / np.linalg.norm(g))


def scene_geometry(ts, pos, vel, **kw):
    g = dict(osv_t=list(ts), osv_pos=pos, osv_vel=vel, first_line_time=0.0, line_time_interval=1.0 / 1652.4156,
             azimuth_size=30000, range_size=5616, chirp_samples=522, slant_range_first=826e3,
             range_spacing=7.8, carrier_frequency=5.331e9, prf=1652.4156, az_bandwidth_fraction=0.8,
             vr_poly=[0.0, 0.0, 7100.0], dc_poly=[0.0, 0.0, 120.0], center_point=[A, 0.0, 0.0])
    g.update(kw)
    return g


def inclined_scene():
    ts = np.arange(-40.0, 61.0, 10.0)
    pos, vel = rotating_orbit(ts)
    g = scene_geometry(ts, pos, vel)
    mid_t = po._line_time(g, g["azimuth_size"] // 2)
    p, v = po.interpolate_orbit(ts, pos, vel, mid_t)
    g["center_point"] = list(po.geolocate(p, v, right_looking_guess(p, v), 826e3 + 2547 * 7.8))
    return g


# ------------------------------------------------------------------------------------ orbit interpolation ----
def test_interpolation_is_exact_for_polynomial_trajectories():
    # degree-5 trajectory through 6 state vectors: Lagrange interpolation reproduces it at any time
    coef = np.array([[7.0e6, 1.2e3, -3.1, 2e-3, -4e-6, 1e-9],
                     [-1.1e6, -7.4e3, 2.2, -1e-3, 3e-6, -2e-9],
                     [2.0e5, 3.3e2, 4.5, 5e-4, -1e-6, 5e-10]])
    f = lambda t: np.array([np.polyval(c[::-1], t) for c in coef])
    df = lambda t: np.array([np.polyval(np.polyder(c[::-1]), t) for c in coef])
    ts = np.array([100.0, 110.0, 120.0, 130.0, 140.0, 150.0])
    pos = [f(t - 100.0) for t in ts]
    vel = [df(t - 100.0) for t in ts]
    for t in (100.0, 104.123456, 127.5, 149.999999, 155.25):
        p, v = focus.interpolate_orbit(ts, pos, vel, t)
        assert np.abs(p - f(t - 100.0)).max() < 1e-6, t
        assert np.abs(v - df(t - 100.0)).max() < 1e-9, t
        po_p, po_v = po.interpolate_orbit(ts, pos, vel, t)
        assert np.abs(p - po_p).max() < 1e-6 and np.abs(v - po_v).max() < 1e-9
    # at a node the node itself comes back
    p, v = focus.interpolate_orbit(ts, pos, vel, 120.0)
    assert np.array_equal(p, pos[2]) and np.array_equal(v, vel[2])


def test_interpolation_uses_whole_microseconds():
    # times are used at microsecond resolution (sar/orbit_state_vector.cc:24): 0.4 us later is the same instant
    ts = [0.0, 10.0]
    pos = [np.zeros(3), np.array([10.0, 0.0, 0.0])]
    vel = [np.ones(3), np.ones(3)]
    p0, _ = focus.interpolate_orbit(ts, pos, vel, 3.000001)
    p1, _ = focus.interpolate_orbit(ts, pos, vel, 3.0000014)
    assert p0[0] == p1[0] and abs(p0[0] - 3.000001) < 1e-12


def test_interpolation_rejects_unordered_vectors():
    with pytest.raises(L.AsarfError):
        focus.interpolate_orbit([0.0, 0.0], [np.zeros(3)] * 2, [np.zeros(3)] * 2, 0.0)


# ------------------------------------------------------------------------------------------ geolocation ----
@pytest.mark.parametrize("r", [826e3, 870e3, 930e3])
def test_geolocation_equatorial_closed_form(r):
    # sensor on the x axis moving along +z: the zero-Doppler plane is z = 0, which cuts the ellipsoid in the
    # equator circle |p| = a; the range circle then fixes x = (rs^2 + a^2 - R^2) / (2 rs)
    rs = A + 790e3
    x = (rs * rs + A * A - r * r) / (2 * rs)
    want = np.array([x, math.sqrt(A * A - x * x), 0.0])
    got = focus.geolocate([rs, 0, 0], [0, 0, 7450.0], [A, 2e5, 1e3], r)
    assert np.abs(got - want).max() < 1e-6
    assert np.abs(po.geolocate([rs, 0, 0], [0, 0, 7450.0], [A, 2e5, 1e3], r) - want).max() < 1e-6
    # the other side of the track from a guess on the other side
    got = focus.geolocate([rs, 0, 0], [0, 0, 7450.0], [A, -2e5, 1e3], r)
    assert np.abs(got - want * [1, -1, 1]).max() < 1e-6


def test_geolocation_satisfies_its_three_equations_on_an_inclined_orbit():
    ts = np.arange(0.0, 41.0, 10.0)
    pos, vel = rotating_orbit(ts)
    for t in (3.0, 17.5, 33.25):
        p, v = focus.interpolate_orbit(ts, pos, vel, t)
        for r in (830e3, 905e3):
            q = focus.geolocate(p, v, right_looking_guess(p, v), r)
            d = q - p
            assert abs(v @ d) / (np.linalg.norm(v) * r) < 1e-12  # zero Doppler
            assert abs(np.linalg.norm(d) - r) < 1e-6  # on the range sphere
            assert abs((q[0] ** 2 + q[1] ** 2) / A ** 2 + q[2] ** 2 / B ** 2 - 1.0) < 1e-14  # on the ellipsoid
            assert np.abs(q - po.geolocate(p, v, right_looking_guess(p, v), r)).max() < 1e-6


def test_geolocation_errors():
    rs = A + 790e3
    with pytest.raises(L.AsarfError):
        focus.geolocate([rs, 0, 0], [0, 0, 7450.0], [A, 2e5, 0], -1.0)
    with pytest.raises(L.AsarfError):  # shorter than the altitude: no point on the ellipsoid
        focus.geolocate([rs, 0, 0], [0, 0, 7450.0], [A, 2e5, 0], 500e3)


# -------------------------------------------------------------------------------------- incidence angle ----
@pytest.mark.parametrize("r", [826e3, 900e3])
def test_incidence_angle_law_of_cosines(r):
    # on the equator the ellipsoid normal is radial: rs^2 = a^2 + R^2 + 2 a R cos(theta)
    rs = A + 790e3
    q = focus.geolocate([rs, 0, 0], [0, 0, 7450.0], [A, 2e5, 0], r)
    want = math.degrees(math.acos((rs * rs - A * A - r * r) / (2 * A * r)))
    assert abs(focus.incidence_angle(q, [rs, 0, 0]) - want) < 1e-9
    assert abs(po.incidence_angle(q, [rs, 0, 0]) - want) < 1e-9


def test_incidence_angle_is_zero_along_the_normal():
    for lat, lon in ((45.0, 10.0), (-63.5, 170.0), (12.0, -100.0)):
        ground, sensor = focus.geo_to_xyz(lat, lon, 0.0), focus.geo_to_xyz(lat, lon, 800e3)
        assert focus.incidence_angle(ground, sensor) < 1e-5


# ------------------------------------------------------------------------------ coordinate conversions ----
def test_geo_to_xyz_closed_forms():
    assert np.abs(focus.geo_to_xyz(0.0, 0.0, 0.0) - [A, 0, 0]).max() < 1e-9
    assert np.abs(focus.geo_to_xyz(90.0, 0.0, 0.0) - [0, 0, po.WGS84_B_FULL]).max() < 1e-8
    assert np.abs(focus.geo_to_xyz(0.0, 90.0, 100.0) - [0, A + 100.0, 0]).max() < 1e-8
    for lat, lon in ((45.0, 10.0), (-33.3, -71.1), (80.0, 179.0)):
        g0, g1 = focus.geo_to_xyz(lat, lon, 0.0), focus.geo_to_xyz(lat, lon, 1000.0)
        bf = po.WGS84_B_FULL
        assert abs((g0[0] ** 2 + g0[1] ** 2) / A ** 2 + g0[2] ** 2 / bf ** 2 - 1.0) < 1e-15
        # the height moves along the ellipsoid normal (the gradient) by exactly that height
        n = np.array([g0[0] / A ** 2, g0[1] / A ** 2, g0[2] / bf ** 2])
        d = g1 - g0
        assert abs(np.linalg.norm(d) - 1000.0) < 1e-6
        assert np.linalg.norm(np.cross(d / 1000.0, n / np.linalg.norm(n))) < 1e-12
        assert np.abs(g0 - po.geo_to_xyz(lat, lon, 0.0)).max() < 1e-8


def test_xyz_to_geo_round_trip_on_the_ellipsoid():
    # Bowring's formula is exact for points on the ellipsoid (height 0)
    for lat in (-89.0, -45.5, -0.25, 0.0, 12.5, 60.0, 89.9):
        for lon in (-179.5, -120.0, -10.0, 0.0, 45.0, 100.0, 179.0):
            if lat == 0.0 and lon == 0.0:
                continue
            got = focus.xyz_to_geo(focus.geo_to_xyz(lat, lon, 0.0))
            assert abs(got[0] - lat) < 1e-9 and abs(got[1] - lon) < 1e-9, (lat, lon, got)
            assert got == pytest.approx(po.xyz_to_geo(po.geo_to_xyz(lat, lon, 0.0)), abs=1e-9)


def test_xyz_to_geo_longitude_on_the_negative_x_axis():
    # atan(y / x) with y = +0 and x < 0 is -0, which neither branch moves (util/include/geo_tools.h:83-88): the
    # reference reports longitude 0 there
    assert focus.xyz_to_geo([-A, 0.0, 0.0])[1] == 0.0
    assert po.xyz_to_geo([-A, 0.0, 0.0])[1] == 0.0
    assert abs(focus.xyz_to_geo([-A, 1e-3, 0.0])[1] - 180.0) < 1e-6


# ----------------------------------------------------------------------------------- processing velocity ----
def straight_line_scene():
    # constant velocity along +z past the equator: Lagrange is exact on a line, and at the zero-Doppler point
    # R(t)^2 = R0^2 + |v|^2 (t - t0)^2, so every Vr sample is |v|
    rs, v = A + 790e3, np.array([120.0, -40.0, 7450.0])
    ts = np.array([-20.0, -10.0, 0.0, 10.0, 20.0])
    p0 = np.array([rs, 0.0, 0.0])
    pos = [p0 + v * t for t in ts]
    vel = [v for _ in ts]
    g = scene_geometry(ts, pos, vel, first_line_time=-15000 * 0.0006, line_time_interval=0.0006,
                       center_point=[A * 0.99, 3e5, 0.0])
    return g, v


def test_processing_velocity_straight_line_is_speed():
    g, v = straight_line_scene()
    poly, vrs, xs = focus.estimate_processing_velocity(g)
    range_end = g["range_size"] - g["chirp_samples"]
    assert list(xs) == list(range(0, range_end, range_end // 10))
    assert len(xs) == 11  # 5094 // 10 = 509 and 10 * 509 < 5094
    assert np.abs(vrs / np.linalg.norm(v) - 1.0).max() < 1e-9
    assert np.abs(np.polyval(poly, xs) / np.linalg.norm(v) - 1.0).max() < 1e-9


def test_processing_velocity_matches_oracle_on_an_inclined_orbit():
    g = inclined_scene()
    poly, vrs, xs = focus.estimate_processing_velocity(g)
    rpoly, rvrs, rxs = po.estimate_processing_velocity(g)
    assert np.array_equal(xs, rxs)
    assert np.abs(vrs / rvrs - 1.0).max() < 1e-9
    assert np.abs(np.polyval(poly, xs) - np.polyval(rpoly, xs)).max() < 1e-5
    # the effective velocity lies between the ground-track speed and the sensor speed
    p, v = po.interpolate_orbit(g["osv_t"], g["osv_pos"], g["osv_vel"], 9.0)
    assert np.all(vrs < np.linalg.norm(v)) and np.all(vrs > np.linalg.norm(v) * A / np.linalg.norm(p))


def test_processing_velocity_rejects_bad_geometry():
    g, _ = straight_line_scene()
    with pytest.raises(L.AsarfError):
        focus.estimate_processing_velocity(dict(g, range_size=500, chirp_samples=522))
    with pytest.raises(L.AsarfError):
        focus.estimate_processing_velocity(dict(g, osv_t=[0.0, -1.0, 2.0, 3.0, 4.0]))


# ------------------------------------------------------------------------------------------ SRGR fit ----
def equatorial_circle_scene():
    # circular orbit in the x-z plane, at x = rs on the middle line moving along +z (a state vector sits on
    # that instant): every zero-Doppler point of that line is on the equator circle
    rs, w = A + 790e3, 1.05e-3
    ts = np.arange(-30.0, 31.0, 10.0)
    pos = [rs * np.array([math.cos(w * t), 0.0, math.sin(w * t)]) for t in ts]
    vel = [rs * w * np.array([-math.sin(w * t), 0.0, math.cos(w * t)]) for t in ts]
    g = scene_geometry(ts, pos, vel, first_line_time=-1.0, line_time_interval=0.001, azimuth_size=2000,
                       center_point=[A * 0.99, 3e5, 0.0])
    return g, rs


def test_srgr_fit_equatorial_closed_form():
    g, rs = equatorial_circle_scene()
    grsr, width = focus.srgr_fit(g, 12.5)
    r = g["slant_range_first"] + np.arange(100) * (g["range_size"] // 100) * g["range_spacing"]
    phi = np.arccos((rs * rs + A * A - r * r) / (2 * rs * A))  # angle at the Earth centre
    chord = 2 * A * np.sin((phi - phi[0]) / 2)
    # the library's fit equals numpy's least-squares quartic through the closed-form (chord, slant) pairs
    want = np.polyfit(chord, r - r[0], 4)
    assert np.abs(np.polyval(grsr, chord) - np.polyval(want, chord)).max() < 1e-6
    assert np.abs(np.polyval(grsr, chord) - (r - r[0])).max() < 0.05  # quartic fit residual, cm level
    assert width == int(chord[-1] / 12.5)
    rg, rw = po.srgr_fit(g, 12.5)
    assert rw == width and np.abs(np.polyval(grsr, chord) - np.polyval(rg, chord)).max() < 1e-6


def test_srgr_fit_matches_oracle_on_an_inclined_orbit():
    g = inclined_scene()
    grsr, width = focus.srgr_fit(g, 12.5)
    rg, rw = po.srgr_fit(g, 12.5)
    assert width == rw and 8000 < width < 10000  # ~112 km of ground under 43.8 km of slant range
    gd = np.linspace(0.0, width * 12.5, 50)
    assert np.abs(np.polyval(grsr, gd) - np.polyval(rg, gd)).max() < 1e-6
    with pytest.raises(L.AsarfError):
        focus.srgr_fit(dict(g, range_size=99, chirp_samples=0), 12.5)


# --------------------------------------------------------------------------------------- scene centre ----
def test_scene_center_straight_line_closed_form():
    # sensor (rs, 0, v t) moving along +z: at time t the zero-Doppler plane z = v t cuts the ellipsoid in a circle
    # of radius a' = a sqrt(1 - z^2 / b^2), and the range circle fixes x = (rs^2 + a'^2 - R^2) / (2 rs)
    rs, vz = A + 790e3, 7450.0
    ts = np.array([-10.0, 10.0])
    pos = [np.array([rs, 0.0, vz * t]) for t in ts]
    vel = [np.array([0.0, 0.0, vz]) for _ in ts]
    g = scene_geometry(ts, pos, vel, first_line_time=-1.0, line_time_interval=0.001, azimuth_size=2000,
                       center_point=[0.0, 0.0, 0.0])
    r = g["slant_range_first"] + ((g["range_size"] - g["chirp_samples"]) // 2) * g["range_spacing"]

    def point(z):
        ap = A * math.sqrt(1.0 - z * z / B ** 2)
        x = (rs * rs + ap * ap - r * r) / (2 * rs)
        return np.array([x, math.sqrt(ap * ap - x * x), z])

    c, spacing = focus.scene_center(g, [A * 0.99, 3e5, 0.0])
    assert np.abs(c - point(0.0)).max() < 1e-6
    assert abs(spacing - np.linalg.norm(point(vz * 0.001) - point(0.0))) < 1e-6
    rc, rsp = po.scene_center(g, [A * 0.99, 3e5, 0.0])
    assert np.abs(c - rc).max() < 1e-6 and abs(spacing - rsp) < 1e-9


# ----------------------------------------------------------------------------- aperture and results window ----
def test_aperture_pixels_match_the_doppler_history():
    # straight-line pass at speed V: the Doppler of a target at closest range R0 is f(t) = -(2 / lambda) V^2 t / R(t)
    # with R(t) = sqrt(R0^2 + V^2 t^2); the aperture is the time (in lines) for |f| to reach fDC + PRF * B / 2,
    # which to first order in V t / R0 is fmax / Ka
    g, v = straight_line_scene()
    speed = np.linalg.norm(v)
    g = dict(g, vr_poly=[0.0, 0.0, speed], dc_poly=[0.0, 0.0, 95.0])
    got = focus.aperture_pixels(g)
    assert got == po.aperture_pixels(g)
    range_end = g["range_size"] - g["chirp_samples"]
    xs = list(range(0, range_end, range_end // 10))
    assert len(got) == len(xs) == 11
    lam = po.SPEED_OF_LIGHT / g["carrier_frequency"]
    fmax = 95.0 + g["prf"] * g["az_bandwidth_fraction"] / 2.0
    for x, n in zip(xs, got):
        r0 = g["slant_range_first"] + x * g["range_spacing"]
        # invert |f(t)| = fmax exactly: (2 V^2 t / lambda)^2 = fmax^2 (R0^2 + V^2 t^2)
        t = fmax * r0 / (speed * math.sqrt((2 * speed / lam) ** 2 - fmax ** 2))
        assert abs(n / (t * g["prf"]) - 1.0) < 2e-3, (x, n, t * g["prf"])


def test_aperture_pixels_capacity_and_geometry_errors():
    g, _ = straight_line_scene()
    s, keep = focus.geometry(g)
    import ctypes
    out, n = (ctypes.c_int * 4)(), ctypes.c_int()
    with pytest.raises(L.AsarfError):
        L.check(L.lib().asarf_aperture_pixels(ctypes.byref(s), out, 4, ctypes.byref(n)))
    with pytest.raises(L.AsarfError):
        focus.aperture_pixels(dict(g, range_size=530))  # fewer than 10 range samples after the chirp


@pytest.mark.parametrize("dc,before,after,aperture,want_after_start,want_before_stop", [
    (120.0, 3000, 3000, 700, 0, 0),      # enough lines cached either side
    (120.0, 500, 3000, 700, 200, 0),     # positive centroid: shortfall at the start
    (-80.0, 3000, 100, 700, 0, 600),     # negative centroid: shortfall at the stop
    (0.0, 0, 0, 700, 0, 700),            # zero counts as not positive
    (1e-9, 0, 0, 700, 700, 0),
])
def test_results_window_cases(dc, before, after, aperture, want_after_start, want_before_stop):
    w = focus.results_window(dc, before, after, aperture, 16)
    assert w == po.results_window(dc, before, after, aperture, 16)
    assert w["lines_before_sensing_start"] == before and w["lines_after_sensing_stop"] == after
    assert w["lines_after_sensing_start"] == want_after_start
    assert w["lines_before_sensing_stop"] == want_before_stop
    assert w["near_range_pixels"] == 8 and w["far_range_pixels"] == 8
    with pytest.raises(L.AsarfError):
        focus.results_window(dc, -1, after, aperture, 16)


# ------------------------------------------------------------------------------- geolocation grid tie points ----
def test_geolocation_tie_points_equatorial_closed_form():
    # line 1000 of the straight-line scene is sensed at t = 0 with the sensor at (rs, 0, 0) moving along +z: every
    # tie point lies on the equator (latitude 0) at x = (rs^2 + a^2 - R^2) / (2 rs), y > 0
    import ctypes
    rs, vz = A + 790e3, 7450.0
    ts = np.array([-10.0, 10.0])
    g = scene_geometry(ts, [np.array([rs, 0.0, vz * t]) for t in ts], [np.array([0.0, 0.0, vz])] * 2,
                       first_line_time=-1.0, line_time_interval=0.001, azimuth_size=2000,
                       center_point=[A * 0.99, 3e5, 0.0])
    s, keep = focus.geometry(g)
    for grsr, scale in ((None, 1.0), ([0.0, 0.0, 0.0, 1.0, 0.0], 1.0), ([0.0, 0.0, 0.0, 2.0, 0.0], 2.0)):
        tp = L.TiePoints()
        gp = (ctypes.c_double * 5)(*grsr) if grsr else None
        L.check(L.lib().asarf_geolocation_tie_points(ctypes.byref(s), 1000, gp, ctypes.byref(tp)))
        for i in range(11):
            samp = i * g["range_size"] // 10
            r = g["slant_range_first"] + scale * samp * g["range_spacing"]
            x = (rs * rs + A * A - r * r) / (2 * rs)
            y = math.sqrt(A * A - x * x)
            assert tp.samp_numbers[i] == samp + 1
            assert tp.lats[i] == 0
            assert abs(tp.longs[i] - int(math.degrees(math.atan2(y, x)) * 1e6)) <= 1
            assert tp.slant_range_times[i] == np.float32(2 * r / po.SPEED_OF_LIGHT * 1e9)
            inc = math.degrees(math.acos((rs * rs - A * A - r * r) / (2 * A * r)))
            assert abs(tp.angles[i] - inc) < 1e-4
    del keep
