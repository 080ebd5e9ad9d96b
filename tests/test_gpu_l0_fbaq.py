"""ASAR IM level-0 packets focused without an intermediate image (ASARF_RAW_ASAR_FBAQ4): data set records built to
the reference's field offsets carry a scene's samples as FBAQ-4 blocks (a quantiser id and 63 code bytes each),
the host packet parser (envisat_im_parse.cc:433-600) turns them into records, and the plan decodes the records in
its statistics and range-compression loads (ConvertAsarImBlocksToComplex, envisat_format_kernels.cu:160-210, with
the NaN fill and padded copy of main_flow.cc:537-540 and main.cc:233-239). The oracle decodes the same records with
its restatement of the decoder into the NaN-filled image and focuses that. SLC, I/Q estimates (FP64 sums on both
sides), Doppler polynomial and the IMP product are compared; some records are short, so their tails are NaN."""
import ctypes
import os
import sys

import numpy as np
import pytest
import torch

import pyoracle as po
from asarf import _lib as L
from asarf import configs, focus, scene
from parity import oracle_crop, oracle_imp, rel_rms, rel_rms_real

sys.path.insert(0, os.path.dirname(__file__))
from test_l0_util import asar_packet  # noqa: E402

pytestmark = pytest.mark.gpu


def fbaq_tables():
    """Reconstruction tables indexed as Fbaq4IndexCalc (256 converted code + block id): block b has step
    0.5 + 0.1 b and 16 symmetric levels; Q's levels are 2 % larger (a gain imbalance for the I/Q correction)."""
    conv = np.arange(16)[:, None]
    step = 0.5 + 0.1 * np.arange(256)[None, :]
    i_lut = ((conv - 7.5) * step).astype(np.float32).ravel()
    return i_lut, (1.02 * i_lut).astype(np.float32)


def encode_line(i, q):
    """FBAQ-4 data bytes of one line: per 63 samples a block id (the step that spans the block's largest sample)
    and one byte per sample, I code in the high nibble; code c encodes converted level 15 - c (c > 7) or 8 + c."""
    out = bytearray()
    for b0 in range(0, i.size, 63):
        bi, bq = i[b0:b0 + 63], q[b0:b0 + 63]
        amp = max(np.abs(bi).max(), np.abs(bq).max(), 1e-3)
        blk = int(np.clip(np.ceil((amp / 7.5 - 0.5) / 0.1), 0, 255))
        step = 0.5 + 0.1 * blk

        def codes(x):
            conv = np.clip(np.rint(x / step + 7.5), 0, 15).astype(np.int64)
            return np.where(conv >= 8, conv - 8, 15 - conv)

        out.append(blk)
        out += ((codes(bi) << 4) | codes(bq)).astype(np.uint8).tobytes()
    return bytes(out)


def records_from_scene(cfg, spl, swst, rec_bytes, rng):
    """The scene's int8 samples as parsed FBAQ-4 records: packets with the encoded blocks go through
    asarf_parse_asar_im_packet; every seventh line carries only spl - 97 samples."""
    raw, _ = scene.generate(cfg.scaled(nr=spl))
    iq = raw.reshape(cfg.na, spl, 2).astype(np.float64)
    recs = np.zeros((cfg.na, rec_bytes), np.uint8)
    for y in range(cfg.na):
        n = spl - 97 if y % 7 == 3 else spl
        pkt, _ = asar_packet(rng, 0, 0, swst=int(swst[y]), data=encode_line(iq[y, :n, 0], iq[y, :n, 1]))
        m = L.AsarEcho()
        L.check(L.lib().asarf_parse_asar_im_packet(pkt, len(pkt), ctypes.byref(m), recs[y].ctypes.data, rec_bytes))
        assert m.swst_code == swst[y]
    return recs


# 5604 and 5770 samples (+ up to 11 SWST steps and the 521-sample chirp) pad to 6144 and 6400, the lengths with
# compile-time range kernels (k_rc_pad, k_rc_static); the others run the runtime-radix kernel.
@pytest.mark.parametrize("spl,imp", [(1400, False), (1400, True), (2017, False), (5604, False), (5770, False)])
def test_asar_fbaq4_plan_matches_oracle(ctx, spl, imp):
    rng = np.random.default_rng(spl + imp)
    base = configs.asar_scene().scaled(na=1200, nr=spl, n_targets_random=4)
    swst = np.repeat(rng.choice(np.array([700, 705, 711]), size=base.na // 100 + 1), 100)[: base.na].astype(np.uint16)
    min_swst = int(swst.min())
    range_size = spl + int(swst.max()) - min_swst
    rec_bytes = 2 + 64 * (spl // 63 + 1) + 5
    recs = records_from_scene(base, spl, swst, rec_bytes, rng)
    i_lut, q_lut = fbaq_tables()
    cfg = base.scaled(nr=range_size)
    plan = focus.Plan(ctx, cfg, imp=imp, fbaq=(rec_bytes, 1))
    if spl > 5000:
        assert plan.params.nrp in (6144, 6400)
    plan.execute_fbaq4(torch.from_numpy(recs).cuda(), torch.from_numpy(swst.view(np.int16)).cuda(), min_swst,
                       torch.from_numpy(i_lut).cuda(), torch.from_numpy(q_lut).cuda())
    slc = plan.slc_tensor().cpu().numpy()
    res = plan.results()
    imp_img = plan.imp_out.cpu().numpy() if imp else None
    plan.close()
    nrp = focus.padding_size(range_size + cfg.nc)
    img = po.ingest_asar_fbaq4(recs, swst, min_swst, i_lut, q_lut, range_size, nrp, cfg.na)
    assert np.isnan(img[:, :range_size]).any()
    ref, rres = po.focus_slc(img[: cfg.na, :range_size], cfg)
    err = rel_rms(slc, ref)
    print(f"spl {spl} range size {range_size}: slc rel rms {err:.3e}", flush=True)
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


def test_asar_fbaq4_plan_argument_checks(ctx):
    cfg = configs.asar_scene().scaled(na=300, nr=1400)
    plan = focus.Plan(ctx, cfg, fbaq=(2 + 64 * 23, 1))
    recs = torch.zeros((300, 2 + 64 * 23), dtype=torch.uint8, device="cuda")
    swst = torch.zeros(300, dtype=torch.int16, device="cuda")
    with pytest.raises(L.AsarfError):
        plan.execute(torch.zeros(300 * 1400 * 2, dtype=torch.uint8, device="cuda"))
    short = torch.zeros(1024, dtype=torch.float32, device="cuda")
    with pytest.raises(L.AsarfError, match="4096"):
        plan.execute_fbaq4(recs, swst, 0, short, short)
    plan.close()
