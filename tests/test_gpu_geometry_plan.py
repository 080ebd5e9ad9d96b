"""The host geometry wired into focusing, as main.cc:251-252 and sar/srgr.cu:73-121 do it: the processing velocity
Vr(x) estimated from an orbit (EstimateProcessingVelocity, processing_velocity_estimation.cc:98-148), the
azimuth spacing from the scene centre (main_flow.cc:499-522) and the slant-to-ground polynomial fitted against the
same orbit (srgr.cu:73-107) drive the fused plan with the IMP tail. The plan's SLC and IMP are compared with the
FP64 oracle run on the numpy restatement of the same geometry (pyoracle), so the library's geometry, the plan and
the oracle are independent of one another."""
import numpy as np
import pytest
import torch

import pyoracle as po
from asarf import configs, focus, scene
from parity import oracle_crop, rel_rms, rel_rms_real, target_deltas
from test_geometry import right_looking_guess, rotating_orbit, scene_geometry

pytestmark = pytest.mark.gpu

SOL = 299792458.0


def test_orbit_geometry_drives_the_plan(ctx):
    cfg0 = configs.asar_scene().scaled(na=2400, nr=1800, n_targets_random=6)
    ts = np.arange(-40.0, 61.0, 10.0)
    pos, vel = rotating_orbit(ts)
    g = scene_geometry(ts, pos, vel, azimuth_size=cfg0.na, range_size=cfg0.nr, chirp_samples=cfg0.nc,
                       slant_range_first=cfg0.r_first, range_spacing=cfg0.range_spacing,
                       carrier_frequency=SOL / cfg0.wavelength, prf=cfg0.prf, line_time_interval=1.0 / cfg0.prf,
                       az_bandwidth_fraction=cfg0.az_bandwidth_fraction, vr_poly=cfg0.vr_poly,
                       dc_poly=[0.0, 0.0, 0.0])
    mid_t = po._line_time(g, g["azimuth_size"] // 2)
    p, v = po.interpolate_orbit(ts, pos, vel, mid_t)
    r_mid = cfg0.r_first + (cfg0.nr - cfg0.nc) // 2 * cfg0.range_spacing
    g["center_point"] = list(po.geolocate(p, v, right_looking_guess(p, v), r_mid))

    # library geometry -> plan; numpy restatement -> oracle
    vr_lib, _, xs = focus.estimate_processing_velocity(g)
    vr_ref, _, _ = po.estimate_processing_velocity(g)
    assert np.abs(np.polyval(vr_lib, xs) / np.polyval(vr_ref, xs) - 1.0).max() < 1e-9
    vr_span = np.polyval(vr_lib, xs).max() - np.polyval(vr_lib, xs).min()
    assert vr_span > 0.05, vr_span  # the orbit makes Vr vary across the swath
    _, az_lib = focus.scene_center(g, g["center_point"])
    _, az_ref = po.scene_center(g, g["center_point"])
    assert abs(az_lib / az_ref - 1.0) < 1e-9
    # the SRGR fit runs on the cropped swath (near = 8 pixels removed, nr_s = nrv - 16), as the plan's IMP tail
    near, nrv = 8, cfg0.nr - cfg0.nc
    g_crop = dict(g, slant_range_first=cfg0.r_first + near * cfg0.range_spacing, range_size=nrv - 16,
                  chirp_samples=0, vr_poly=list(vr_lib))
    grsr_lib, ng_lib = focus.srgr_fit(g_crop, 12.5)
    grsr_ref, ng_ref = po.srgr_fit(dict(g_crop, vr_poly=list(vr_ref)), 12.5)
    assert ng_lib == ng_ref

    cfg = cfg0.scaled(vr_coeffs=tuple(float(c) for c in vr_lib))
    raw, _ = scene.generate(cfg)
    plan = focus.Plan(ctx, cfg, imp=True, srgr=(grsr_lib, ng_lib), azimuth_spacing=az_lib)
    plan.execute(torch.from_numpy(raw).cuda())
    slc = plan.slc_tensor().cpu().numpy()
    res = plan.results()
    imp = plan.imp_out.cpu().numpy()
    plan.close()

    cfg_ref = cfg0.scaled(vr_coeffs=tuple(float(c) for c in vr_ref))
    ref, rres = po.focus_slc(scene.raw_to_complex(raw), cfg_ref)
    err = rel_rms(slc, ref)
    worst, n = target_deltas(cfg, slc, ref)
    print(f"Vr {np.polyval(vr_lib, 0):.2f}..{np.polyval(vr_lib, nrv):.2f} m/s, azimuth spacing {az_lib:.4f} m, "
          f"ground width {ng_lib}: slc rel rms {err:.3e}, {n} targets, worst {worst}", flush=True)
    assert err < 1e-4
    assert worst["pslr"] < 0.05 and worst["islr"] < 0.05 and worst["irw"] < 0.01
    # IMP: the oracle's crop, multilook at the scene-centre spacing, RSL and SRGR with the numpy fit
    crop = oracle_crop(cfg_ref, rres["dc_poly"])
    assert tuple(res["crop"]) == crop
    front, back, cn, cf = crop
    sub = ref[front:cfg.na - back, cn:nrv - cf]
    ml = po.az_multilook(sub, az_ref)
    po.rsl(ml, cfg.r_first + cn * cfg.range_spacing, cfg.range_spacing, 800e3)
    gr = po.srgr(ml, grsr_ref, cfg.range_spacing, 12.5, ng_ref)
    assert res["imp_rows"] == gr.shape[0]
    ierr = rel_rms_real(imp[: res["imp_rows"], : gr.shape[1]], gr)
    print(f"imp {gr.shape} rel rms {ierr:.3e}")
    assert ierr < 2e-4
