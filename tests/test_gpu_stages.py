"""GPU parity of every stage entry point against the FP64 oracle on the same seeded inputs (HIP path through the
C ABI). Tolerances: FP32 kernels vs FP64 oracle, relative RMS over the compared region."""
import numpy as np
import pytest
import torch

import pyoracle as po
from asarf import _lib as L
from asarf import configs, focus, scene

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / max(np.sum(np.abs(b) ** 2), 1e-300)))


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rand_c(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


# ------------------------------------------------------------------------------------------------ FFT ----
@pytest.mark.parametrize("n", [16, 60, 84, 343, 1000, 2401, 4116, 6144, 6400, 8192])
def test_row_fft_matches_oracle(ctx, n):
    x = rand_c((9, n), n)
    t = cuda(x)
    focus.fft(ctx, t, L.FFT_ROWS)
    ref = x.astype(np.complex128)
    po.fft_rows(ref)
    got = t.cpu().numpy()
    assert rel_rms(got, ref) < 2e-6, n
    focus.fft(ctx, t, L.FFT_ROWS, inverse=True, scale=1.0 / n)
    assert rel_rms(t.cpu().numpy(), x) < 4e-6


@pytest.mark.parametrize("n,cols", [(64, 16), (168, 37), (180, 21), (84, 16), (49, 3), (112, 40), (147, 17),
                                    (128, 16), (256, 33), (210, 5), (625, 33), (4116, 40), (16464, 17), (30240, 19),
                                    (32768, 9)])
def test_column_fft_matches_oracle(ctx, n, cols):
    x = rand_c((n, cols), n + cols)
    t = cuda(x)
    focus.fft(ctx, t, L.FFT_COLUMNS)
    ref = x.astype(np.complex128)
    po.fft_cols(ref)
    assert rel_rms(t.cpu().numpy(), ref) < 2e-6, n
    focus.fft(ctx, t, L.FFT_COLUMNS, inverse=True, scale=1.0 / n)
    assert rel_rms(t.cpu().numpy(), x) < 4e-6


# ------------------------------------------------------------------------------------------ I/Q correction ----
@pytest.mark.parametrize("quirks", [po.QUIRKS_REFERENCE, 0])
def test_iq_correction_matches_oracle(ctx, quirks):
    cfg = configs.ers_block().scaled(na=300, nr=1100, n_targets_az=1, n_targets_rg=1)
    raw, _ = scene.generate(cfg)
    x = scene.raw_to_complex(raw)
    nrp, nap = po.next_fft_size(1100 + cfg.nc), po.next_fft_size(300)
    img = np.zeros((nap, nrp), np.complex64)
    img[:300, :1100] = x
    img[5, 7] = np.nan + 1j * np.nan
    img[9, 3] = np.nan + 2.0j
    ref = img.astype(np.complex128)
    rres = po.iq_correction(ref, 1100, 300, n_total=300 * 1100, quirks=quirks)
    t = cuda(img)
    gres = focus.iq_correction(ctx, t, 1100, 300, n_total=300 * 1100, quirks=quirks)
    for k in rres:
        assert abs(gres[k] - rres[k]) <= 1e-9 * max(1.0, abs(rres[k])), k
    assert rel_rms(t.cpu().numpy(), ref) < 1e-6


def test_doppler_centroid_matches_oracle(ctx):
    cfg = configs.asar_scene().scaled(na=512, nr=1400, n_targets_random=0, clutter_power=1.0)
    _, flt = scene.generate(cfg, want_float=True)
    nrp = po.next_fft_size(1400 + cfg.nc)
    img = np.zeros((512, nrp), np.complex64)
    img[:, :1400] = flt
    img[:, 1400:] = 0.25 - 0.5j  # padding content enters the last block
    for quirks in (po.QUIRKS_REFERENCE, 0):
        poly_r, bx_r, bf_r = po.doppler_centroid(img.astype(np.complex128), 1400, 512, cfg.prf, quirks=quirks)
        poly_g, bx_g, bf_g = focus.doppler_centroid(ctx, cuda(img), 1400, 512, cfg.prf, quirks=quirks)
        assert np.array_equal(bx_g, bx_r)
        assert np.abs(bf_g - bf_r).max() < 1e-6
        assert np.abs(np.polyval(poly_g, bx_r) - np.polyval(poly_r, bx_r)).max() < 1e-6


# ------------------------------------------------------------------------------------- range compression ----
@pytest.mark.parametrize("nr,na", [(5616, 48), (1300, 64)])
def test_range_compression_matches_oracle(ctx, nr, na):
    cfg = configs.ers_block()
    nrp = po.next_fft_size(nr + cfg.nc)
    x = np.zeros((na, nrp), np.complex64)
    x[:, :nr] = rand_c((na, nr), nr)
    x[:, nr:] = 3.0 + 1.0j  # padding values do not reach the kept columns
    chirp = focus.generate_chirp(cfg, nrp)
    ref = x.astype(np.complex128)
    nrv = po.range_compression(ref, nr, na, chirp, cfg.nc)
    t = cuda(x)
    nrv_g = focus.range_compression(ctx, t, nr, na, chirp, cfg.nc)
    assert nrv_g == nrv
    got = t.cpu().numpy()
    assert rel_rms(got[:, :nrv], ref[:, :nrv]) < 2e-6
    assert np.all(got[:, nrv:] == 0)


# ------------------------------------------------------------------------------------- range-Doppler ----
# nap <= 1204: one column transform per pass; nap = 1500: four-step split, every pass out of place through the
# consumed source buffer, column-tiled RCMC, or (8x the wavelength: migration beyond the tile halo) the row kernel
@pytest.mark.parametrize("na,nrv,dc,wl", [(200, 300, 150.0, 1), (600, 700, -300.0, 1), (1100, 500, 40.0, 1),
                                          (1500, 400, 80.0, 1), (1500, 400, -120.0, 8)])
def test_range_doppler_matches_oracle(ctx, na, nrv, dc, wl):
    cfg = configs.ers_block()
    nap = po.next_fft_size(na)
    stride = nrv + 13
    src = np.zeros((nap, stride), np.complex64)
    src[:na, :nrv] = rand_c((na, nrv), na + nrv)
    dcp = [2e-5, -0.01, dc]
    p = focus.rda_params(cfg, dcp)
    p.wavelength = cfg.wavelength * wl
    op = po.rda_params(cfg.wavelength * wl, cfg.range_spacing, cfg.r_first, cfg.prf, cfg.az_bandwidth_fraction, dcp,
                       cfg.vr_poly)
    ref = po.range_doppler(src.astype(np.complex128), nrv, na, op)
    out = focus.range_doppler(ctx, p, cuda(src), nrv, na)
    got = out.cpu().numpy()
    assert rel_rms(got[:na, :nrv], ref[:na, :nrv]) < 1e-5
    assert np.all(got[:, nrv:] == 0)


def test_range_doppler_adopts_and_trades_workspace(ctx):
    """RangeDopplerAlgorithm's buffer trading: out adopts ws, ws receives src's memory."""
    cfg = configs.ers_block()
    na, nrv = 120, 200
    nap = po.next_fft_size(na)
    src = torch.zeros((nap, nrv), dtype=torch.complex64, device="cuda")
    src[:na] = cuda(rand_c((na, nrv), 3))
    wsbuf = torch.empty_like(src)
    import ctypes
    s = focus.image(src, nrv, na)
    o = L.Image()
    ws = L.Workspace(wsbuf.data_ptr(), wsbuf.numel() * 8)
    p = focus.rda_params(cfg, [0, 0, 100.0])
    L.check(L.lib().asarf_range_doppler(ctx.h, ctypes.byref(p), ctypes.byref(s), ctypes.byref(o), ctypes.byref(ws)))
    assert o.data == wsbuf.data_ptr() and ws.data == src.data_ptr() and s.data is None
    assert o.x_size == nrv and o.y_size == na
    ctx.sync()


# ------------------------------------------------------------------------------------------- IMP tail ----
@pytest.mark.parametrize("width", [400, 399])  # two columns per lane (16-byte loads) / one column per lane
def test_multilook_rsl_srgr_match_oracle(ctx, width):
    rng = np.random.default_rng(11)
    slc = rand_c((900, width), 5)
    ml_r = po.az_multilook(slc.astype(np.complex128), 4.298)
    ml_g, sp = focus.az_multilook(ctx, cuda(slc), width, 900, 4.298)
    assert ml_g.shape == ml_r.shape
    assert rel_rms(ml_g.cpu().numpy(), ml_r) < 1e-6
    assert abs(sp - 4.298 * np.float32(12.5 / 4.298)) < 1e-9
    po.rsl(ml_r, 850e3, 7.8, 800e3)
    focus.range_spreading_loss(ctx, ml_g, 850e3, 7.8, 800e3)
    assert rel_rms(ml_g.cpu().numpy(), ml_r) < 1e-6
    grsr = np.array([1e-19, -2e-13, 1.1e-7, 0.39, 0.0]) * np.array([1, 1, 1, 1, 1])
    out_nr = 700
    sr_r = po.srgr(ml_r, grsr, 7.8, 12.5, out_nr)
    sr_g = focus.srgr(ctx, ml_g, grsr, 7.8, 12.5, out_nr)
    assert rel_rms(sr_g.cpu().numpy(), sr_r) < 1e-6
    del rng


# ------------------------------------------------------------------------- the reference's stage sequence ----
@pytest.mark.parametrize("which", ["ers", "asar", "ers1500"])
def test_stage_sequence_matches_oracle_and_plan(ctx, which):
    """main.cc:243-283 through the stage entry points (RawDataCorrection, CalculateDopplerCentroid,
    RangeCompression, RangeDopplerAlgorithm) on one padded image equals the FP64 oracle and the fused plan
    (ers1500: 1500 lines, a four-step azimuth split)."""
    if which == "ers":
        cfg = configs.ers_block().scaled(na=700, nr=1500, n_targets_az=2, n_targets_rg=2, tgt_az=(0.4, 0.6))
    elif which == "ers1500":
        cfg = configs.ers_block().scaled(na=1500, nr=1500, n_targets_az=2, n_targets_rg=2, tgt_az=(0.4, 0.6))
    else:
        cfg = configs.asar_scene().scaled(na=640, nr=1400, n_targets_random=6)
    raw, _ = scene.generate(cfg)
    na, nr = cfg.na, cfg.nr
    nrp, nap = po.next_fft_size(nr + cfg.nc), po.next_fft_size(na)
    img = np.zeros((nap, nrp), np.complex64)
    img[:na, :nr] = scene.raw_to_complex(raw)
    t = cuda(img)
    iq = focus.iq_correction(ctx, t, nr, na)
    dc_poly, _, _ = focus.doppler_centroid(ctx, t, nr, na, cfg.prf)
    chirp = focus.generate_chirp(cfg, nrp)
    nrv = focus.range_compression(ctx, t, nr, na, chirp, cfg.nc)
    out = focus.range_doppler(ctx, focus.rda_params(cfg, dc_poly), t, nrv, na)
    slc = out[:na, :nrv].cpu().numpy()

    ref, rres = po.focus_slc(scene.raw_to_complex(raw), cfg)
    assert nrv == rres["nrv"]
    err = rel_rms(slc, ref)
    print(f"{which}: stage sequence vs oracle rel rms {err:.3e}")
    assert err < 1e-4
    for k, v in rres["iq"].items():
        assert abs(iq[k] - v) <= 1e-9 * max(1.0, abs(v)), k

    plan = focus.Plan(ctx, cfg)
    plan.execute(torch.from_numpy(raw).cuda())
    pslc = plan.slc_tensor().cpu().numpy()
    pres = plan.results()
    plan.close()
    # the stage sequence estimates the Doppler centroid on the FP32-corrected image, as the reference does; the
    # plan (and the oracle) from exact moments of the raw bytes: they agree to FP32 rounding of the samples
    xs = np.arange(0, nrv)
    assert np.abs(np.polyval(pres["dc_poly"], xs) - np.polyval(dc_poly, xs)).max() < 1e-3
    assert np.abs(np.polyval(rres["dc_poly"], xs) - np.polyval(dc_poly, xs)).max() < 1e-3
    assert pslc.shape == slc.shape
    assert rel_rms(slc, pslc) < 1e-5
