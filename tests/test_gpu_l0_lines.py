"""ERS level-0 lines focused without an intermediate image (ASARF_RAW_ERS_LINES): the plan reads each line where
its SWST code places it, treats the range samples no line sample lands on as the NaN the reference's decoder
writes there (ConvertErsImSamplesToComplex, envisat_format_kernels.cu:116-133: skipped by the I/Q statistics,
zero after ZeroNaNs), and pads as main.cc:233-239 does. The oracle decodes the same lines into the NaN-filled
image (its restatement of the decoder) and focuses that. SLC, I/Q estimates and Doppler polynomial are compared,
for line pitches that do and do not allow the 4-sample loads, and for the IMP product."""
import numpy as np
import pytest
import torch

import pyoracle as po
from asarf import configs, focus, scene
from parity import oracle_crop, oracle_imp, rel_rms, rel_rms_real

pytestmark = pytest.mark.gpu


def _lines(cfg, spl, line_bytes, codes, rng):
    """The scene's raw samples as level-0 lines of line_bytes (spl I/Q pairs, then filler) and per-line SWST codes
    drawn from `codes` in runs of lines, as a scene's sampling window moves."""
    raw, _ = scene.generate(cfg.scaled(nr=spl))
    na = cfg.na
    lines = np.zeros((na, line_bytes), np.uint8)
    lines[:, : 2 * spl] = raw.reshape(na, 2 * spl).view(np.uint8)
    lines[:, 2 * spl:] = rng.integers(0, 256, (na, line_bytes - 2 * spl), dtype=np.uint8)
    runs = np.repeat(rng.choice(codes, size=na // 100 + 1), 100)[:na]
    return lines, runs.astype(np.int16)


# The last two cases reach the compile-time range kernels: 5616 + 4 * 19 range samples (+ a 704-sample chirp) pad
# to 6400 (k_rc_static), 5424 + 4 * 3 to 6144 (k_rc_pad).
@pytest.mark.parametrize("spl,pad,imp,steps", [(1400, 32, False, (0, 22, 44)), (1400, 6, False, (0, 22, 44)),
                                               (1402, 8, False, (0, 22, 44)), (1400, 32, True, (0, 22, 44)),
                                               (5616, 32, False, (0, 9, 19)), (5424, 16, False, (0, 2, 3))])
def test_ers_lines_plan_matches_oracle(ctx, spl, pad, imp, steps):
    rng = np.random.default_rng(spl + pad)
    codes = 700 + np.array(steps)
    base = configs.ers_block().scaled(na=1500 if spl < 5000 else 1100, nr=spl, n_targets_az=2, n_targets_rg=2,
                                      tgt_az=(0.4, 0.6))
    lines, swst = _lines(base, spl, 2 * spl + pad, codes, rng)
    min_swst = int(swst.min())
    range_size = spl + 4 * (int(swst.max()) - min_swst)
    cfg = base.scaled(nr=range_size)
    plan = focus.Plan(ctx, cfg, imp=imp, lines=(spl, lines.shape[1], 4))
    if spl > 5000:
        assert plan.params.nrp in (6144, 6400)
    plan.execute_lines(torch.from_numpy(lines).cuda(), torch.from_numpy(swst).cuda(), min_swst)
    slc = plan.slc_tensor().cpu().numpy()
    res = plan.results()
    imp_img = plan.imp_out.cpu().numpy() if imp else None
    plan.close()
    nrp = focus.padding_size(range_size + cfg.nc)
    img = po.ingest_ers_im(lines, spl, swst.view(np.uint16), min_swst, range_size, nrp, cfg.na)
    assert np.isnan(img[:, :range_size]).any()
    ref, rres = po.focus_slc(img[: cfg.na, :range_size], cfg)
    err = rel_rms(slc, ref)
    print(f"spl {spl} pitch {lines.shape[1]} range size {range_size}: slc rel rms {err:.3e}", flush=True)
    assert err < 1e-4
    for k, v in rres["iq"].items():
        assert abs(res["iq"][k] - v) <= 1e-9 * max(1.0, abs(v)), k
    xs = np.arange(0, cfg.nr - cfg.nc)
    assert np.abs(np.polyval(res["dc_poly"], xs) - np.polyval(rres["dc_poly"], xs)).max() < 1e-6
    if imp:
        crop = oracle_crop(cfg, rres["dc_poly"])
        assert tuple(res["crop"]) == crop
        gr = oracle_imp(cfg, ref, crop)
        assert rel_rms_real(imp_img[: res["imp_rows"], : gr.shape[1]], gr) < 2e-4


def test_ers_lines_plan_rejects_plain_execute(ctx):
    cfg = configs.ers_block().scaled(na=300, nr=1400)
    plan = focus.Plan(ctx, cfg, lines=(1400, 2800, 4))
    with pytest.raises(Exception):
        plan.execute(torch.zeros(300 * 1400 * 2, dtype=torch.uint8, device="cuda"))
    plan.close()
