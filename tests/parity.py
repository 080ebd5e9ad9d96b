"""Shared checks of the GPU parity tests: error measures, the oracle's results window (the crop before the IMP
tail), the oracle's IMP chain and point-target quality deltas. Everything here is computed from the FP64 oracle
and the scene configuration alone, never from what the HIP path reports."""
import numpy as np

import pyoracle as po
from asarf import analysis, scene


def rel_rms(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def rel_rms_real(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def oracle_max_aperture(cfg):
    """Largest aperture in lines over the ten range steps, with the Doppler centroid still 0 and the initial Vr
    (CalcAperturePixels, sar/sar_metadata.h:117-141, called before the DC estimate at main.cc:251)."""
    nrv = cfg.nr - cfg.nc
    step = nrv // 10
    return max(po.aperture_pixels_at(cfg.wavelength, cfg.prf, cfg.az_bandwidth_fraction, cfg.vr_poly, [],
                                     cfg.r_first, cfg.range_spacing, x) for x in range(0, nrv, step))


def oracle_crop(cfg, dc_poly, lines_before=0, lines_after=0):
    """(front, back, near, far) lines and range pixels removed before output: CalcResultsWindow
    (main_flow.cc:155-201) on the estimated Doppler constant term (main.cc:296-298, the last coefficient of the
    highest-power-first polynomial) and SubsetResultsAndReassembleMeta's sums (main_flow.cc:599-602)."""
    w = po.results_window(dc_poly[-1], lines_before, lines_after, oracle_max_aperture(cfg))
    front = w["lines_before_sensing_start"] + w["lines_after_sensing_start"]
    back = w["lines_before_sensing_stop"] + w["lines_after_sensing_stop"]
    return front, back, w["near_range_pixels"], w["far_range_pixels"]


def spherical_grsr(r_first, dr, range_size, rs, re, gs=12.5):
    """Slant-range-to-ground fit of sar/srgr.cu:73-121 on a spherical earth, in numpy: 100 slant ranges
    r_first + i*(range_size // 100)*dr; the ground point of slant range R lies at central angle
    acos((rs^2 + re^2 - R^2) / (2 rs re)); ground distance = chord from the first point; degree-4 least squares of
    R - r_first on that distance (highest power first); output width floor(g_99 / gs)."""
    r = r_first + np.arange(100) * (range_size // 100) * dr
    a = np.arccos((rs * rs + re * re - r * r) / (2.0 * rs * re))
    g = 2.0 * re * np.sin(0.5 * (a - a[0]))
    return np.polyfit(g, r - r_first, 4), int(g[99] / gs)


def oracle_imp(cfg, slc_ref, crop):
    """Crop, 4-line boxcar multilook, range spreading loss and SRGR on the oracle SLC (main.cc:303-327)."""
    front, back, near, far = crop
    sub = slc_ref[front:cfg.na - back, near:slc_ref.shape[1] - far]
    ml = po.az_multilook(sub, cfg.az_spacing)
    po.rsl(ml, cfg.r_first + near * cfg.range_spacing, cfg.range_spacing, 800e3)
    nr_s = slc_ref.shape[1] - near - far
    grsr, ng = spherical_grsr(cfg.r_first + near * cfg.range_spacing, cfg.range_spacing, nr_s,
                              cfg.earth_radius + cfg.orbit_height, cfg.earth_radius, 12.5)
    return po.srgr(ml, grsr, cfg.range_spacing, 12.5, ng)


def target_deltas(cfg, slc, ref):
    """Worst |PSLR|, |ISLR| (dB) and relative IRW differences over every point target of the scene, range and
    azimuth cuts. Returns (worst dict, number of targets measured)."""
    tx, ty = scene.targets(cfg)
    worst = {"pslr": 0.0, "islr": 0.0, "irw": 0.0}
    n = 0
    for x, y in zip(tx, ty):
        try:
            mr = analysis.point_target(ref, x, y)
        except ValueError:
            continue  # window would leave the image
        mg = analysis.point_target(slc, x, y)
        n += 1
        for ax in ("range", "azimuth"):
            worst["pslr"] = max(worst["pslr"], abs(mg[ax]["pslr"] - mr[ax]["pslr"]))
            worst["islr"] = max(worst["islr"], abs(mg[ax]["islr"] - mr[ax]["islr"]))
            worst["irw"] = max(worst["irw"], abs(mg[ax]["irw"] / mr[ax]["irw"] - 1.0))
    return worst, n
