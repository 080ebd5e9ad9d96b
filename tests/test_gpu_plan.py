"""GPU parity of the fused focusing plan (raw bytes -> SLC [-> IMP]) against the FP64 oracle, and point-target
quality parity (PSLR/ISLR within 0.05 dB, IRW within 1 %)."""
import numpy as np
import pytest
import torch

import pyoracle as po
from asarf import analysis, configs, focus, scene
from parity import oracle_crop, oracle_imp, rel_rms_real

pytestmark = pytest.mark.gpu

SLC_TOL = 1e-4  # relative RMS, FP32 GPU vs FP64 oracle


def rel_rms(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def run_plan(ctx, cfg, quirks=po.QUIRKS_REFERENCE, imp=False):
    raw, _ = scene.generate(cfg)
    traw = torch.from_numpy(raw).cuda()
    plan = focus.Plan(ctx, cfg, quirks=quirks, imp=imp)
    plan.execute(traw)
    slc = plan.slc_tensor().cpu().numpy()
    res = plan.results()
    imp_img = plan.imp_out.cpu().numpy() if imp else None
    plan.close()
    return raw, slc, res, imp_img


@pytest.mark.parametrize("quirks", [po.QUIRKS_REFERENCE, 0])
@pytest.mark.parametrize("which", ["ers", "asar"])
def test_plan_matches_oracle_small(ctx, which, quirks):
    if which == "ers":
        cfg = configs.ers_block().scaled(na=700, nr=1500, n_targets_az=2, n_targets_rg=2, tgt_az=(0.4, 0.6))
    else:
        cfg = configs.asar_scene().scaled(na=640, nr=1400, n_targets_random=6)
    raw, slc, res, _ = run_plan(ctx, cfg, quirks)
    ref, rres = po.focus_slc(scene.raw_to_complex(raw), cfg, quirks=quirks)
    err = rel_rms(slc, ref)
    print(f"{which} quirks={quirks} slc rel rms {err:.3e}")
    assert err < SLC_TOL
    for k, v in rres["iq"].items():
        assert abs(res["iq"][k] - v) <= 1e-9 * max(1.0, abs(v)), k
    xs = np.arange(0, cfg.nr - cfg.nc)
    assert np.abs(np.polyval(res["dc_poly"], xs) - np.polyval(rres["dc_poly"], xs)).max() < 1e-6


def test_plan_ers_block_full_size_parity_and_quality(ctx):
    """4096 x 5616 ERS block: SLC within 1e-4 of the oracle; point-target PSLR/ISLR within 0.05 dB, IRW 1 %."""
    cfg = configs.ers_block()
    raw, slc, res, _ = run_plan(ctx, cfg)
    ref, rres = po.focus_slc(scene.raw_to_complex(raw), cfg)
    err = rel_rms(slc, ref)
    print(f"ers block slc rel rms {err:.3e}")
    assert err < SLC_TOL
    tx, ty = scene.targets(cfg)
    worst = {"pslr": 0.0, "islr": 0.0, "irw": 0.0}
    for x, y in zip(tx, ty):
        mg = analysis.point_target(slc, x, y)
        mr = analysis.point_target(ref, x, y)
        for ax in ("range", "azimuth"):
            worst["pslr"] = max(worst["pslr"], abs(mg[ax]["pslr"] - mr[ax]["pslr"]))
            worst["islr"] = max(worst["islr"], abs(mg[ax]["islr"] - mr[ax]["islr"]))
            worst["irw"] = max(worst["irw"], abs(mg[ax]["irw"] / mr[ax]["irw"] - 1.0))
    print("worst target metric deltas", worst)
    assert worst["pslr"] < 0.05 and worst["islr"] < 0.05 and worst["irw"] < 0.01


@pytest.mark.parametrize("nr", [1400, 1401])  # even / odd multilook width (two columns per lane / one)
def test_plan_imp_matches_oracle_chain(ctx, nr):
    cfg = configs.asar_scene().scaled(na=1200, nr=nr, n_targets_random=8)
    raw, slc, res, imp = run_plan(ctx, cfg, imp=True)
    ref, rres = po.focus_slc(scene.raw_to_complex(raw), cfg)
    crop = oracle_crop(cfg, rres["dc_poly"])  # from the oracle alone, not from the plan's report
    assert tuple(res["crop"]) == crop, (res["crop"], crop)
    gr = oracle_imp(cfg, ref, crop)
    assert res["imp_rows"] == gr.shape[0]
    got = imp[: res["imp_rows"], :gr.shape[1]]
    err = rel_rms_real(got, gr)
    print(f"imp rel rms {err:.3e}")
    assert err < 2e-4


def run_plan_env(ctx, cfg, env, **kw):
    """run_plan with environment overrides that the plan reads at creation."""
    import os
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return run_plan(ctx, cfg, **kw)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("which", ["ers", "asar"])
def test_plan_fused_azimuth_core_matches_separate_kernels(ctx, which):
    """The fused pass-B/RCMC/matched-filter/pass-B' kernel (the default where the azimuth length splits with a
    compiled second pass: 720 = 8 x 90, 1512 = 18 x 84) against the separate kernels (ASARF_AZ_FUSED=0, balanced
    split) on the same raw data, and both against the oracle."""
    if which == "ers":
        cfg = configs.ers_block().scaled(na=700, nr=1500, n_targets_az=2, n_targets_rg=2, tgt_az=(0.4, 0.6))
    else:
        cfg = configs.asar_scene().scaled(na=1500, nr=1400, n_targets_random=6)
    raw, slc_f, res_f, _ = run_plan(ctx, cfg)
    _, slc_s, res_s, _ = run_plan_env(ctx, cfg, {"ASARF_AZ_FUSED": "0"})
    assert res_f["azimuth_core_fused"] and not res_s["azimuth_core_fused"]
    assert 0 < res_f["rcmc_shift_bound"] <= 8
    err = rel_rms(slc_f, slc_s)
    ref, _ = po.focus_slc(scene.raw_to_complex(raw), cfg)
    err_f, err_s = rel_rms(slc_f, ref), rel_rms(slc_s, ref)
    print(f"{which} fused vs separate rel rms {err:.3e}; vs oracle fused {err_f:.3e} separate {err_s:.3e}")
    assert err < 1e-6
    assert err_f < SLC_TOL and err_s < SLC_TOL


def test_plan_large_migration_takes_separate_path(ctx):
    """L-band wavelength: the RCMC shift bound exceeds the fused tile's halo, the device-side gate selects the
    separate kernels, and the SLC still matches the oracle."""
    cfg = configs.ers_block().scaled(na=700, nr=1500, n_targets_az=2, n_targets_rg=2, tgt_az=(0.4, 0.6),
                                     wavelength=0.236)
    raw, slc, res, _ = run_plan(ctx, cfg)
    assert not res["azimuth_core_fused"] and res["rcmc_shift_bound"] > 8
    ref, _ = po.focus_slc(scene.raw_to_complex(raw), cfg)
    err = rel_rms(slc, ref)
    print(f"large migration slc rel rms {err:.3e} (shift bound {res['rcmc_shift_bound']})")
    assert err < SLC_TOL


@pytest.mark.parametrize("wavelength", [0.236, 0.32])
def test_plan_tiled_and_row_rcmc_kernels(ctx, wavelength):
    """Default path: the column-tiled RCMC kernel covers shifts up to 24 pixels; beyond that the device-side
    bound hands the work to the row kernel. Both must reproduce the oracle."""
    cfg = configs.ers_block().scaled(na=700, nr=1500, n_targets_az=2, n_targets_rg=2, tgt_az=(0.4, 0.6),
                                     wavelength=wavelength)
    raw, slc, res, _ = run_plan(ctx, cfg)
    ref, _ = po.focus_slc(scene.raw_to_complex(raw), cfg)
    err = rel_rms(slc, ref)
    print(f"wavelength {wavelength}: shift bound {res['rcmc_shift_bound']}, slc rel rms {err:.3e}")
    assert (res["rcmc_shift_bound"] <= 24) == (wavelength < 0.3)
    assert err < SLC_TOL


@pytest.mark.parametrize("nr,offset", [(1400, 0), (1400, 8), (1400, 2), (1402, 0), (1404, 0)])
def test_plan_statistics_load_width(ctx, nr, offset):
    """The raw statistics pass reads 8 samples per lane with 16-byte loads and packed dot products when
    nr % 8 == 0 and the raw buffer is 16-byte aligned (nr = 1400: Doppler blocks of 43 columns, so every other
    block starts inside a 2-sample word), 4 samples with 8-byte loads when only nr % 4 == 0 or 8-byte alignment
    holds, one sample otherwise: each gives the oracle's I/Q statistics, Doppler centroid and SLC."""
    cfg = configs.asar_scene().scaled(na=640, nr=nr, n_targets_random=4)
    raw, _ = scene.generate(cfg)
    buf = torch.empty(raw.size + 16, dtype=torch.uint8, device="cuda")
    traw = buf[offset:offset + raw.size]
    traw.copy_(torch.from_numpy(raw.reshape(-1).view(np.uint8)))
    plan = focus.Plan(ctx, cfg)
    plan.execute(traw)
    slc = plan.slc_tensor().cpu().numpy()
    res = plan.results()
    plan.close()
    ref, rres = po.focus_slc(scene.raw_to_complex(raw), cfg)
    err = rel_rms(slc, ref)
    print(f"nr={nr} offset={offset} slc rel rms {err:.3e}")
    assert err < SLC_TOL
    for k, v in rres["iq"].items():
        assert abs(res["iq"][k] - v) <= 1e-9 * max(1.0, abs(v)), k
    xs = np.arange(0, cfg.nr - cfg.nc)
    assert np.abs(np.polyval(res["dc_poly"], xs) - np.polyval(rres["dc_poly"], xs)).max() < 1e-6


@pytest.mark.parametrize("na", [1100, 1120, 2500])
def test_plan_band_edge_sizes(ctx, na):
    """Azimuth lengths (1120, 1125, 2520) for which 0.8 x nbins / 2 is an integer put bins exactly on the band
    limits of the matched filter: the column-tiled RCMC kernel (which skips out-of-band pixels) and the oracle
    must keep the same bins."""
    cfg = configs.ers_block().scaled(na=na, nr=900, n_targets_az=1, n_targets_rg=1)
    raw, slc, res, _ = run_plan(ctx, cfg)
    ref, _ = po.focus_slc(scene.raw_to_complex(raw), cfg)
    err = rel_rms(slc, ref)
    print(f"na={na}: slc rel rms {err:.3e}, shift bound {res['rcmc_shift_bound']}")
    assert res["rcmc_shift_bound"] <= 24  # the tiled kernel ran
    assert err < SLC_TOL


# scene sizes spread over generic four-step splits, both formats, single- and two-pass azimuth transforms
@pytest.mark.parametrize("which,na,nr", [("ers", 2770, 1204), ("asar", 376, 1915), ("ers", 2342, 2438),
                                         ("asar", 2710, 2201), ("ers", 1119, 1980), ("asar", 2052, 2130),
                                         ("asar", 565, 1518), ("asar", 2707, 724)])
def test_plan_size_sweep(ctx, which, na, nr):
    if which == "ers":
        cfg = configs.ers_block().scaled(na=na, nr=nr, n_targets_az=1, n_targets_rg=2)
    else:
        cfg = configs.asar_scene().scaled(na=na, nr=nr, n_targets_random=4)
    raw, slc, res, _ = run_plan(ctx, cfg)
    ref, rres = po.focus_slc(scene.raw_to_complex(raw), cfg)
    err = rel_rms(slc, ref)
    print(f"{which} {na}x{nr}: slc rel rms {err:.3e}")
    assert err < SLC_TOL
    xs = np.arange(0, cfg.nr - cfg.nc)
    assert np.abs(np.polyval(res["dc_poly"], xs) - np.polyval(rres["dc_poly"], xs)).max() < 1e-6


@pytest.mark.parametrize("which", ["asar_1400", "asar_1401", "ers"])
def test_plan_imp_fused_multilook_srgr_is_bit_identical(ctx, which):
    """The fused multilook + SRGR kernel (the detected slant-range image kept in LDS) against the two kernels with
    that image in HBM between them (ASARF_IMP_FUSED=0): the same IMP product bit for bit, over the device-side
    crop and row count."""
    if which == "ers":
        cfg = configs.ers_block().scaled(na=1400, nr=1500, n_targets_az=2, n_targets_rg=2)
    else:
        cfg = configs.asar_scene().scaled(na=1200, nr=int(which[-4:]), n_targets_random=8)
    _, _, res_f, imp_f = run_plan(ctx, cfg, imp=True)
    _, _, res_s, imp_s = run_plan_env(ctx, cfg, {"ASARF_IMP_FUSED": "0"}, imp=True)
    assert res_f["imp_rows"] == res_s["imp_rows"] and tuple(res_f["crop"]) == tuple(res_s["crop"])
    rows = res_f["imp_rows"]
    assert rows > 0
    np.testing.assert_array_equal(imp_f[:rows], imp_s[:rows])



@pytest.mark.parametrize("chunks", ["2", "3", "4"])
def test_plan_azimuth_chunks_are_bit_identical(ctx, chunks):
    """The azimuth tail in column chunks over two streams (each chunk's core waits for the previous chunk's, and
    runs beside its in-place pass A') against one unsplit chunk: the same SLC and IMP bit for bit."""
    cfg = configs.asar_scene().scaled(na=1200, nr=3200, n_targets_random=8)
    _, slc_1, res_1, imp_1 = run_plan_env(ctx, cfg, {"ASARF_AZ_CHUNKS": "1"}, imp=True)
    _, slc_k, res_k, imp_k = run_plan_env(ctx, cfg, {"ASARF_AZ_CHUNKS": chunks}, imp=True)
    assert res_1["azimuth_core_fused"] and res_k["azimuth_core_fused"]
    np.testing.assert_array_equal(slc_k, slc_1)
    rows = res_1["imp_rows"]
    np.testing.assert_array_equal(imp_k[:rows], imp_1[:rows])
