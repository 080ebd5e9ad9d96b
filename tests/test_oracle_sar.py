"""Pins the oracle's SAR stages with closed forms, independent library routines and invariants.

Each test names what fixes the answer; none re-derives the oracle's own formula.
"""
import math

import numpy as np
import pytest

import pyoracle as po
from asarf import analysis, configs, scene
from conftest import golden


def seven_smooth_above(n, limit=131072):
    best = None
    for v in range(n + 1, limit + 1):
        m = v
        for f in (2, 3, 5, 7):
            while m % f == 0:
                m //= f
        if m == 1:
            best = v
            break
    return -1 if best is None else best


# ------------------------------------------------------------------------------------------ padding rule ----
@pytest.mark.parametrize("n", [1, 2, 100, 4096, 5616 + 704, 5616 + 521, 16384, 30000, 4116, 6400, 131071, 131072])
def test_padding_rule_is_strictly_greater_smooth_number(n):
    assert po.next_fft_size(n) == seven_smooth_above(n)


def test_padding_rule_known_shapes():
    # the shapes of the five workloads (4096 -> 4116 = 2^2 3 7^3 because the bound is strict)
    assert po.next_fft_size(4096) == 4116
    assert po.next_fft_size(16384) == 16464
    assert po.next_fft_size(30000) == 30240
    assert po.next_fft_size(5616 + 704) == 6400
    assert po.next_fft_size(5616 + 521) == 6144


# ------------------------------------------------------------------------------------------ chirp ----
@pytest.mark.parametrize("which", ["ers", "asar"])
@pytest.mark.parametrize("window", [False, True])
def test_chirp_matches_closed_form(which, window):
    g = golden("chirp_vectors.npz")
    fs, T, B, n, kr = g[f"{which}_params"]
    n = int(n)
    ref = g[f"{which}_windowed" if window else f"{which}_plain"]
    pad = po.next_fft_size(n + 5616)
    c = po.generate_chirp(kr, fs, n, window, pad)
    assert np.all(c[n:] == 0)
    assert np.abs(c[:n] - ref).max() < 2e-7  # float rounding of unit-energy samples
    assert abs(np.sum(np.abs(c.astype(np.complex128)) ** 2) - 1.0) < 1e-5


def test_chirp_window_is_periodic_hamming():
    g = golden("windows.npz")
    c_w = po.generate_chirp(4.189e11, 18.96e6, 704, True, 6400)[:704].astype(np.complex128)
    c_p = po.generate_chirp(4.189e11, 18.96e6, 704, False, 6400)[:704].astype(np.complex128)
    ratio = np.abs(c_w) / np.abs(c_p)
    w = g["hamming_periodic_704"]
    assert np.abs(ratio / ratio[352] - w / w[352]).max() < 1e-5


# ------------------------------------------------------------------------------------------ polyfit ----
def test_polyfit_matches_numpy():
    g = golden("polyfit_vectors.npz")
    for key, deg in (("dc", 2), ("sr", 4), ("r0", 0), ("r1", 1), ("r3", 3), ("r5", 5)):
        x, y, ref = g[f"{key}_x"], g[f"{key}_y"], g[f"{key}_c"]
        c = po.polyfit(x, y, deg)
        # compare the fitted curves (coefficients of ill-conditioned fits may differ in low bits)
        fit = np.polyval(c, x)
        fit_ref = np.polyval(ref, x)
        assert np.abs(fit - fit_ref).max() <= 1e-9 * max(1.0, np.abs(y).max()), key
        assert np.abs(c[-1] - ref[-1]) <= 1e-6 * max(1.0, abs(ref[-1])) + 1e-9, key


def test_polyfit_recovers_exact_polynomial_highest_power_first():
    x = np.arange(0, 5775, 175, dtype=float)
    p = np.array([-3e-6, 0.05, 150.0])
    c = po.polyfit(x, np.polyval(p, x), 2)
    assert np.allclose(c, p, rtol=1e-10, atol=1e-12)


# ------------------------------------------------------------------------------------------ I/Q correction ----
def imbalanced_block(na=64, nr=96, a=3.0, b=2.0, eps_deg=7.0, dci=15.5, dcq=14.0):
    """I = a cos(t) + dci, Q = b sin(t + eps) + dcq on an exact angle grid: all sums are trig identities."""
    t = 2 * np.pi * np.arange(na * nr).reshape(na, nr) / (na * nr) * 7.0
    e = np.deg2rad(eps_deg)
    return a * np.cos(t) + dci + 1j * (b * np.sin(t + e) + dcq)


def test_iq_recovers_dc_gain_phase():
    na, nr = 64, 96
    raw = imbalanced_block(na, nr)
    img = np.zeros((na + 4, nr + 8), np.complex128)
    img[:na, :nr] = raw
    res = po.iq_correction(img, nr, na, quirks=0)
    assert abs(res["dc_i"] - 15.5) < 1e-12 and abs(res["dc_q"] - 14.0) < 1e-12
    assert abs(res["iq_gain"] - 1.5) < 1e-12
    assert abs(res["phase_error"] - 7.0) < 1e-10
    assert abs(res["stddev_i"] - 3.0 / math.sqrt(2)) < 1e-12
    out = img[:na, :nr]
    # corrected output is balanced and orthogonal: (b cos t, b sin t)
    assert abs(np.sum(out.real ** 2) - np.sum(out.imag ** 2)) < 1e-9 * np.sum(out.real ** 2)
    assert abs(np.sum(out.real * out.imag)) < 1e-9 * np.sum(out.real ** 2)
    # padding untouched without the padded-width quirk
    assert np.all(img[:, nr:] == 0) and np.all(img[na:, :] == 0)


def test_iq_padded_width_quirk_closed_form():
    na, nr, nrp = 64, 96, 120
    raw = imbalanced_block(na, nr, a=2.0, b=2.0, eps_deg=0.0, dci=15.5, dcq=15.5)
    img = np.zeros((na + 4, nrp), np.complex128)
    img[:na, :nr] = raw
    res = po.iq_correction(img, nr, na, quirks=po.Q1_IQ_PADDED_WIDTH | po.Q2_COS_NO_SQRT)
    dci, dcq = raw.real.mean(), raw.imag.mean()
    s_ii = np.sum((raw.real - dci) ** 2)
    s_qq = np.sum((raw.imag - dcq) ** 2)
    s_iq = np.sum((raw.real - dci) * (raw.imag - dcq))
    P = na * (nrp - nr)
    g = math.sqrt((s_ii + P * dci ** 2) / (s_qq + P * dcq ** 2))
    d = (s_iq / g + P * dci * dcq / g) / ((s_ii + P * dci ** 2) / g ** 2)
    assert abs(res["iq_gain"] - g) < 1e-12
    assert abs(res["phase_error"] - math.degrees(math.asin(d))) < 1e-9
    assert d > 0.3  # the padding makes the phase estimate large for offset-binary data
    # padding columns now hold the corrected image of raw 0: (-dci/g, (-dcq + dci d/g)/(1 - d^2))
    pad = img[:na, nr:]
    assert np.allclose(pad.real, -dci / g, atol=1e-12)
    assert np.allclose(pad.imag, (-dcq + dci * d / g) / (1 - d * d), atol=1e-12)


def test_iq_nan_samples_are_skipped_and_zeroed():
    na, nr = 32, 48
    raw = imbalanced_block(na, nr)
    raw[3, 5] = np.nan + 1j * np.nan
    raw[7, 11] = np.nan + 0j
    img = np.zeros((na, nr), np.complex128)
    img[:, :] = raw
    ntot = na * nr
    res = po.iq_correction(img, nr, na, n_total=ntot, quirks=0)
    valid = ~(np.isnan(raw.real) | np.isnan(raw.imag))
    assert abs(res["dc_i"] - raw.real[valid].sum() / ntot) < 1e-12  # divides by n_total, not the valid count
    assert img[3, 5] == 0 and img[7, 11] == 0
    assert not np.isnan(img).any()


# ------------------------------------------------------------------------------------------ Doppler centroid ----
def test_dc_of_pure_phase_ramp():
    na, nr, prf, f = 40, 5616, 1679.9, 123.4
    y = np.arange(na)[:, None]
    img = np.zeros((na, 6400), np.complex128)
    img[:, :nr] = np.exp(2j * np.pi * f * y / prf) * np.ones((1, nr))
    poly, bx, bf = po.doppler_centroid(img, nr, na, prf, quirks=0)
    assert len(bf) == 32
    assert np.abs(bf - f).max() < 1e-9
    assert abs(poly[2] - f) < 1e-8 and abs(poly[0]) < 1e-12 and abs(poly[1]) < 1e-10


def test_dc_blockwise_frequencies_fit_matches_numpy():
    na, nr, prf = 24, 5616, 1652.0
    step = nr // 32
    nb = (nr + step - 1) // step
    fb = -200.0 + 450.0 * (np.arange(nb) / (nb - 1)) ** 1.5
    img = np.zeros((na, 6144), np.complex128)
    y = np.arange(na)[:, None]
    for b in range(nb):
        x0, x1 = b * step, min((b + 1) * step, 6144)
        img[:, x0:x1] = np.exp(2j * np.pi * fb[b] * y / prf)
    poly, bx, bf = po.doppler_centroid(img, nr, na, prf)
    assert len(bf) == 33  # 5616 % 32 != 0: the extra block reads the padding columns
    assert bx[-1] == (5600 + 5775) // 2
    assert np.abs(bf - fb).max() < 1e-9
    assert np.allclose(poly, np.polyfit(bx, bf, 2), rtol=1e-8, atol=1e-10)


def test_dc_extra_block_sees_padding():
    na, nr, prf = 16, 5616, 1652.0
    img = np.zeros((na, 6144), np.complex128)
    y = np.arange(na)[:, None]
    img[:, :nr] = np.exp(2j * np.pi * 100.0 * y / prf)
    img[:, nr:] = 5.0  # constant padding: lag product is real positive
    _, _, bf = po.doppler_centroid(img, nr, na, prf)
    assert abs(bf[0] - 100.0) < 1e-9
    assert abs(bf[-1]) < 5.0  # dominated by the padding, pulled towards 0 Hz
    _, _, bf_fixed = po.doppler_centroid(img, nr, na, prf, quirks=0)
    assert len(bf_fixed) == 32


# ------------------------------------------------------------------------------------------ range compression ----
def test_range_compression_equals_circular_correlation():
    rng = np.random.default_rng(5)
    na, nr, nc, nrp = 6, 200, 31, 252
    img = np.zeros((na, nrp), np.complex128)
    img[:, :nr] = rng.standard_normal((na, nr)) + 1j * rng.standard_normal((na, nr))
    chirp = np.zeros(nrp, np.complex64)
    chirp[:nc] = (rng.standard_normal(nc) + 1j * rng.standard_normal(nc)).astype(np.complex64)
    src = img.copy()
    nrv = po.range_compression(img, nr, na, chirp, nc)
    assert nrv == nr - nc
    c = chirp.astype(np.complex128)
    for yy in range(na):
        for k in (0, 1, 57, nrv - 1):
            ref = np.sum(src[yy, k:k + nc] * np.conj(c[:nc]))
            assert abs(img[yy, k] - ref) < 1e-10
    assert np.all(img[:, nrv:] == 0)


@pytest.mark.parametrize("window,pslr_lo,pslr_hi", [(False, -13.8, -12.8), (True, -46.0, -38.0)])
def test_compressed_pulse_sidelobes(window, pslr_lo, pslr_hi):
    fs, T, B = 18.96e6, 37.12e-6, 15.55e6
    nc = int(math.floor(T * fs + 0.5))
    kr = B / T
    nr, nrp = 2048, po.next_fft_size(2048 + nc)
    chirp = po.generate_chirp(kr, fs, nc, window, nrp)
    tx = po.generate_chirp(kr, fs, nc, False, nrp)[:nc]  # transmitted pulse, unweighted
    k0 = 600
    img = np.zeros((40, nrp), np.complex128)
    img[:, k0:k0 + nc] = tx
    nrv = po.range_compression(img, nr, 40, chirp, nc)
    cut = img[20, :nrv]
    assert int(np.argmax(np.abs(cut))) == k0
    m = analysis.cut_metrics(analysis.upsample1d(cut[k0 - 64:k0 + 64], 16), 16)
    assert pslr_lo < m["pslr"] < pslr_hi


# ------------------------------------------------------------------------------------------ RCMC / matched filter ----
def rcmc_params(prf=1679.9, dc=150.0, wavelength=0.05657, dr=7.905919, r0=830e3, vr=7050.0):
    return po.rda_params(wavelength, dr, r0, prf, 0.8, [0.0, 0.0, dc], [0.0, 0.0, vr])


def test_rcmc_window_is_numpy_blackman16():
    w = po.rcmc_window()
    assert np.array_equal(w, golden("windows.npz")["blackman16"].astype(np.float32))


def test_rcmc_zero_frequency_row_is_identity():
    nrv, nbins = 300, 64
    rng = np.random.default_rng(1)
    f = rng.standard_normal((nbins, nrv)) + 1j * rng.standard_normal((nbins, nrv))
    out = po.rcmc(f, nrv, rcmc_params(dc=0.0))
    assert np.allclose(out[0, 8:nrv - 8], f[0, 8:nrv - 8], atol=1e-12)
    assert np.all(out[0, :8] == 0) and np.all(out[0, nrv - 8:] == 0)


def test_rcmc_shift_of_a_slow_exponential_matches_migration_formula():
    """A slowly varying range exponential is interpolated almost exactly, so the phase of out/in recovers the
    applied shift; it must equal the hyperbolic range migration lambda^2 f^2 R0 / (8 Vr^2) in pixels."""
    nrv, nbins, prf = 400, 16, 1679.9
    lam, dr, r_first, vr = 0.05657, 7.905919, 830e3, 7050.0
    w = 0.05
    x = np.arange(nrv)
    f = np.tile(np.exp(1j * w * x), (nbins, 1)).astype(np.complex128)
    p = rcmc_params(prf=prf, dc=0.0, wavelength=lam, dr=dr, r0=r_first, vr=vr)
    out = po.rcmc(f, nrv, p)
    for k in (1, 3, 5, 7, 9, 13):
        fk = k * prf / nbins if k < nbins // 2 else (k - nbins) * prf / nbins
        xs = np.arange(40, nrv - 40)
        shift = lam ** 2 * fk ** 2 * (r_first + xs * dr) / (8 * vr ** 2) / dr
        meas = np.angle(out[k, xs] / f[k, xs]) / w
        assert np.abs(meas - shift).max() < 2e-3, k
        assert np.abs(np.abs(out[k, xs]) - 1).max() < 2e-3


def test_matched_filter_is_unit_modulus_inside_band():
    nrv, nbins, prf, dc = 50, 256, 1679.9, -300.0
    f = np.ones((nbins, nrv), np.complex128)
    po.az_matched_filter(f, nrv, rcmc_params(prf=prf, dc=dc))
    mag = np.abs(f[:, 10])
    inside = mag > 0.5
    assert np.allclose(mag[inside], 1.0, atol=1e-12)
    # kept bandwidth is the azimuth bandwidth fraction of the PRF
    assert abs(inside.sum() - 0.8 * nbins) <= 2


@pytest.mark.parametrize("nbins,beta,dc", [(1120, 0.8, -300.0), (1120, 0.8, 211.7), (1125, 0.8, 35.0),
                                           (1008, 0.8, -300.0), (2520, 0.8, 640.0), (1000, 0.7, 0.0)])
def test_matched_filter_band_edges(nbins, beta, dc):
    # bins are whole multiples of prf / nbins and the kept band is centred on the bin-quantised centroid, so
    # exactly 2 floor(beta nbins / 2) + 1 bins survive in every column: when beta nbins / 2 is an integer the
    # two edge bins lie exactly on the band limits and are kept
    nrv = 8
    f = np.ones((nbins, nrv), np.complex128)
    p = rcmc_params(prf=1679.9, dc=dc)
    p.az_bandwidth_fraction = beta
    po.az_matched_filter(f, nrv, p)
    kept = (np.abs(f) > 0.5).sum(axis=0)
    assert np.all(kept == 2 * int(np.floor(beta * nbins / 2 + 1e-9)) + 1), kept


# ------------------------------------------------------------------------------------------ whole path ----
@pytest.mark.parametrize("dc", [0.0, 150.0, -150.0])
def test_point_target_focuses_at_zero_doppler_position(dc):
    # weak Doppler-correlated clutter gives the ACCC estimator a signature in every range block
    cfg = configs.ers_block().scaled(na=2048, nr=1500, n_targets_az=1, n_targets_rg=1, tgt_rg=(0.5, 0.5),
                                     tgt_az=(0.5, 0.5), snr_db=-999, dc_near=dc, dc_far=dc, clutter_power=1e-4)
    _, flt = scene.generate(cfg, want_float=True)
    tx, ty = scene.targets(cfg)
    slc, res = po.focus_slc(flt, cfg, quirks=0)
    assert abs(np.polyval(res["dc_poly"], 400) - dc) < 1.0
    m = analysis.point_target(slc, tx[0], ty[0])
    assert abs(m["peak_x"] - tx[0]) < 0.1 and abs(m["peak_y"] - ty[0]) < 0.1
    # unweighted azimuth: sinc response, IRW 0.886/beta lines, PSLR -13.3 dB
    assert abs(m["azimuth"]["irw"] - 0.886 / 0.8) < 0.05
    assert -14.0 < m["azimuth"]["pslr"] < -12.5
    # Hamming-weighted range: IRW ~ 1.30 fs/B samples (0.886 for rect x 1.47), PSLR near -42 dB
    assert abs(m["range"]["irw"] - 1.30 * cfg.fs / cfg.bandwidth) < 0.05
    assert m["range"]["pslr"] < -38.0


def test_dc_estimate_of_clutter_scene():
    """ACCC on Doppler-correlated clutter recovers a range-linear centroid (the reference's DC model)."""
    cfg = configs.asar_scene().scaled(na=1024, nr=1400, n_targets_random=0, clutter_power=1.0)
    _, flt = scene.generate(cfg, want_float=True)
    img = np.zeros((1024, po.next_fft_size(1400 + cfg.nc)), np.complex128)
    img[:, :1400] = flt
    poly, bx, bf = po.doppler_centroid(img, 1400, 1024, cfg.prf, quirks=0)
    nrv = 1400 - cfg.nc
    truth = cfg.dc_near + (cfg.dc_far - cfg.dc_near) * bx / (nrv - 1)
    assert np.abs(bf - truth).max() < 8.0
    assert np.abs(np.polyval(poly, bx) - truth).max() < 3.0


# ------------------------------------------------------------------------------------------ IMP tail ----
def test_multilook_constant_intensity_is_preserved():
    slc = np.full((300, 40), 2.0 + 1.0j)
    out = po.az_multilook(slc, 4.298)
    assert out.shape == (int(np.round(np.float32(300) / np.float32(12.5 / 4.298))), 40)
    assert np.allclose(out, 5.0, atol=1e-12)


def test_multilook_interpolates_linear_intensity_ramp():
    rows = np.arange(400, dtype=float)
    slc = np.sqrt(1.0 + 0.5 * rows)[:, None] * np.ones((1, 8))
    out = po.az_multilook(slc, 12.5 / 2.5)
    # intensity ramp: the two 4-line means are the ramp at centres y r - 0.5 and y r + 0.5, so the lerp equals
    # the ramp at y r + mu - 0.5 away from the edges
    r = np.float32(2.5)
    for y in (10, 50, 120):
        pos = np.float32(y) * r
        mu = pos - np.floor(pos)
        expect = 1.0 + 0.5 * (np.floor(pos) - 0.5 + mu)
        assert abs(out[y, 3] - expect) < 1e-9


def test_rsl_closed_form():
    img = np.ones((3, 100))
    po.rsl(img, 850e3, 7.8, 800e3)
    r = 850e3 + np.arange(100) * 7.8
    assert np.allclose(img[1], (r / 800e3) ** 1.5, rtol=1e-6)


def test_srgr_identity_and_linear_ramp():
    img = (3.0 + 0.25 * np.arange(500))[None, :] * np.ones((4, 1))
    ident = po.srgr(img, [0, 0, 0, 1.0, 0], 12.5, 12.5, 400)
    assert np.allclose(ident, img[:, :400], atol=1e-12)
    scaled = po.srgr(img, [0, 0, 0, 0.37, 0], 12.5, 12.5, 400)
    ri = 0.37 * np.arange(400)
    assert np.allclose(scaled[2], 3.0 + 0.25 * ri, atol=1e-9)
