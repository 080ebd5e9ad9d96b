"""GPU parity at the full sizes of the measured workloads, against the FP64 oracle:

* the ASAR IM scene (30000 x 5616 int8, 200 targets in clutter) focused to SLC and IMP by the fused plan: SLC
  relative RMS, I/Q estimates, Doppler polynomial, every target's PSLR/ISLR/IRW, the results window computed from
  the oracle alone, and the IMP product of the oracle's crop, multilook, RSL and SRGR chain;
* the high-migration scene (16384 x 5616, Doppler swept +-0.4 PRF across range) with the 16-tap sinc RCMC;
* the 64-block heterogeneous batch: one captured graph bit-identical to sequential launches, and blocks of both
  kinds (the smallest and the largest ERS and swath-varied ASAR blocks) against the oracle.

The reference's own flow is main.cc:233-327 (and main_flow.cc:155-201, 585-632 for the crop)."""
import numpy as np
import pytest
import torch

import pyoracle as po
from asarf import configs, focus, scene
from parity import oracle_crop, oracle_imp, rel_rms, rel_rms_real, target_deltas

pytestmark = pytest.mark.gpu

SLC_TOL = 1e-4


def _focus(ctx, cfg, imp=False, **kw):
    raw, _ = scene.generate(cfg)
    plan = focus.Plan(ctx, cfg, imp=imp, **kw)
    plan.execute(torch.from_numpy(raw).cuda())
    slc = plan.slc_tensor().cpu().numpy()
    res = plan.results()
    imp_img = plan.imp_out.cpu().numpy() if imp else None
    plan.close()
    return raw, slc, res, imp_img


def _check_estimates(cfg, res, rres):
    for k, v in rres["iq"].items():
        assert abs(res["iq"][k] - v) <= 1e-9 * max(1.0, abs(v)), k
    xs = np.arange(0, cfg.nr - cfg.nc)
    assert np.abs(np.polyval(res["dc_poly"], xs) - np.polyval(rres["dc_poly"], xs)).max() < 1e-6


def test_asar_scene_full_size_slc_imp_and_targets(ctx):
    """The benchmarked workload end to end at 30000 x 5616 with the IMP tail."""
    cfg = configs.asar_scene()
    print(f"case {cfg.name} {cfg.na}x{cfg.nr}", flush=True)
    raw, slc, res, imp = _focus(ctx, cfg, imp=True)
    ref, rres = po.focus_slc(scene.raw_to_complex(raw), cfg)
    del raw
    err = rel_rms(slc, ref)
    print(f"slc rel rms {err:.3e}", flush=True)
    assert err < SLC_TOL
    _check_estimates(cfg, res, rres)
    worst, n = target_deltas(cfg, slc, ref)
    print(f"{n} targets, worst deltas {worst}", flush=True)
    assert n >= 190
    assert worst["pslr"] < 0.05 and worst["islr"] < 0.05 and worst["irw"] < 0.01
    crop = oracle_crop(cfg, rres["dc_poly"])
    assert tuple(res["crop"]) == crop, (res["crop"], crop)
    gr = oracle_imp(cfg, ref, crop)
    assert res["imp_rows"] == gr.shape[0]
    got = imp[: res["imp_rows"], : gr.shape[1]]
    ierr = rel_rms_real(got, gr)
    print(f"imp {gr.shape} rel rms {ierr:.3e}", flush=True)
    assert ierr < 2e-4


@pytest.mark.parametrize("sign", [+1, -1])
def test_results_window_both_doppler_signs_with_cached_lines(ctx, sign):
    """The crop comes from the oracle's results window for the estimated Doppler constant term: a positive
    constant term removes the aperture shortfall after the sensing start, a negative one before the stop."""
    dc = (120.0, 260.0) if sign > 0 else (-260.0, -80.0)
    cfg = configs.asar_scene().scaled(na=1600, nr=1500, n_targets_random=6, dc_near=dc[0], dc_far=dc[1])
    lb, la = 37, 51
    raw, slc, res, imp = _focus(ctx, cfg, imp=True, lines_before=lb, lines_after=la)
    ref, rres = po.focus_slc(scene.raw_to_complex(raw), cfg)
    assert np.sign(rres["dc_poly"][-1]) == sign
    crop = oracle_crop(cfg, rres["dc_poly"], lb, la)
    print(f"sign {sign}: crop {res['crop']} oracle {crop}")
    assert tuple(res["crop"]) == crop
    if sign > 0:
        assert crop[0] > lb and crop[1] == la
    else:
        assert crop[0] == lb and crop[1] > la
    gr = oracle_imp(cfg, ref, crop)
    assert res["imp_rows"] == gr.shape[0]
    assert rel_rms_real(imp[: res["imp_rows"], : gr.shape[1]], gr) < 2e-4


def test_high_migration_full_size_sinc16(ctx):
    """16384 x 5616 with the Doppler centroid swept +-0.4 PRF across range (16-tap sinc RCMC)."""
    cfg = configs.high_migration()
    print(f"case {cfg.name} {cfg.na}x{cfg.nr}", flush=True)
    raw, slc, res, _ = _focus(ctx, cfg)
    ref, rres = po.focus_slc(scene.raw_to_complex(raw), cfg)
    err = rel_rms(slc, ref)
    print(f"slc rel rms {err:.3e}, shift bound {res['rcmc_shift_bound']}", flush=True)
    assert err < SLC_TOL
    _check_estimates(cfg, res, rres)
    worst, n = target_deltas(cfg, slc, ref)
    print(f"{n} targets, worst deltas {worst}", flush=True)
    assert n == 25
    assert worst["pslr"] < 0.05 and worst["islr"] < 0.05 and worst["irw"] < 0.01


@pytest.mark.timeout(400)
def test_batch_64_blocks_full_size(ctx):
    """All 64 blocks of the heterogeneous batch: the captured graph equals sequential launches bit for bit on two
    replays, and the smallest and largest block of each kind match the oracle."""
    cfgs = configs.batch_blocks(64)
    b = focus.Batch(ctx, n_streams=4)
    for i, c in enumerate(cfgs):
        b.add(c, torch.from_numpy(scene.generate(c)[0]).cuda())
    print("64 blocks added", flush=True)
    b.execute(focus.BATCH_SEQUENTIAL)
    b.sync()
    seq = [b.slc_tensor(i) for i in range(len(cfgs))]
    for rep in range(2):
        b.execute(focus.BATCH_GRAPH)
        b.sync()
        for i in range(len(cfgs)):
            assert torch.equal(b.slc_tensor(i), seq[i]), (rep, i, cfgs[i].name)
        print(f"graph replay {rep} equals sequential", flush=True)
    kinds = {"ers": [i for i, c in enumerate(cfgs) if c.raw_format == "u8_ers"],
             "asar": [i for i, c in enumerate(cfgs) if c.raw_format == "i8"]}
    picks = []
    for k, idx in kinds.items():
        idx = sorted(idx, key=lambda i: cfgs[i].na * cfgs[i].nr)
        picks += [idx[0], idx[-1]]
    seq_host = {i: seq[i].cpu().numpy() for i in picks}
    del seq
    b.close()
    torch.cuda.empty_cache()
    for i in picks:
        c = cfgs[i]
        ref, _ = po.focus_slc(scene.raw_to_complex(scene.generate(c)[0]), c)
        err = rel_rms(seq_host[i], ref)
        print(f"block {i} {c.name} (nc {c.nc}): slc rel rms {err:.2e}", flush=True)
        assert err < SLC_TOL


@pytest.mark.parametrize("dtype", ["cf32", "f32"])
def test_image_subset_copy_matches_numpy_slice(ctx, dtype):
    """asarf_image_subset_copy (the strided copy of SubsetResultsAndReassembleMeta, main_flow.cc:621-627) against
    a numpy slice, into a destination with its own pitch."""
    from asarf import _lib as L
    rng = np.random.default_rng(5)
    ys, xs, stride = 301, 517, 528
    if dtype == "cf32":
        host = (rng.standard_normal((ys, stride)) + 1j * rng.standard_normal((ys, stride))).astype(np.complex64)
    else:
        host = rng.standard_normal((ys, stride)).astype(np.float32)
    dev = torch.from_numpy(host).cuda()
    img = focus.image(dev, xs, ys)
    for (x0, y0, w, h) in [(8, 40, xs - 16, ys - 40 - 23), (0, 0, xs, ys), (xs - 1, ys - 1, 1, 1)]:
        dst = torch.full((h, w + 5), -7, dtype=dev.dtype, device="cuda")
        focus.subset_copy(ctx, img, x0, y0, w, h, dst)
        got = dst.cpu().numpy()
        assert np.array_equal(got[:, :w], host[y0:y0 + h, x0:x0 + w])
        assert np.all(got[:, w:] == -7)  # the pitch padding of the destination is not written
    # a window past the logical image is rejected before any copy
    dst = torch.zeros((4, 4), dtype=dev.dtype, device="cuda")
    with pytest.raises(L.AsarfError):
        focus.subset_copy(ctx, img, xs - 2, 0, 4, 4, dst)
