"""Throughput of the fused focusing plan on one MI355X.

One step = one pass of the GPU focusing stage over one synthetic raw scene already resident in HBM: I/Q and
Doppler statistics, range compression, azimuth FFT, RCMC + azimuth matched filter, azimuth IFFT, and the IMP tail
(crop, multilook x range-spreading loss, slant-to-ground resampling). Default workload: the full ENVISAT ASAR IM
scene, 30000 x 5616 complex int8 samples.

Prints one JSON line. value = raw input samples focused per second (Msamples/s). The roofline object is for the
kernel with the largest average launch time, whose duration is measured with HIP events on the plan's stream.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "asar-focus_amd"))

PEAK_HBM_GBS = 8000.0
METRIC = "focused SLC Msamples/s per GPU + % HBM roofline; learned/BF16 vs plain at equal PSLR"


def kernel_bytes(name, cfg, plan, res):
    """Compulsory HBM bytes of one launch (logical sizes; zero padding synthesised in registers counts 0)."""
    na, nr = cfg.na, cfg.nr
    nrv = cfg.nr - cfg.nc
    nap = plan.params.nap
    if name == "raw_moments":
        return 2 * na * nr
    if name == "range_compression":
        return 2 * na * nr + 8 * na * nrv
    if name in ("azimuth_fft_a",):
        return 8 * na * nrv + 8 * nap * nrv
    if name in ("azimuth_fft_b", "rcmc_matched_filter", "azimuth_ifft_b", "azimuth_core_fused"):
        return 16 * nap * nrv
    if name == "azimuth_ifft_a":
        return 8 * nap * nrv + 8 * na * nrv
    if name == "multilook_rsl":
        front, back, near, far = res["crop"]
        return 8 * (na - front - back) * (nrv - near - far) + 4 * res["imp_rows"] * (nrv - near - far)
    if name == "srgr":
        front, back, near, far = res["crop"]
        return 4 * res["imp_rows"] * (nrv - near - far) + 4 * res["imp_rows"] * plan.ng
    if name == "multilook_srgr":
        front, back, near, far = res["crop"]
        return 8 * (na - front - back) * (nrv - near - far) + 4 * res["imp_rows"] * plan.ng
    return None


def kernel_sources_sha1():
    """SHA-1 over the library's kernel and host sources (asar-focus_amd/csrc, sorted by name): the key under which
    the committed counter profiles were measured."""
    import hashlib

    h = hashlib.sha1()
    d = os.path.join(ROOT, "asar-focus_amd", "csrc")
    for name in sorted(os.listdir(d)):
        if name.endswith((".hip", ".h", ".cpp")):
            h.update(name.encode())
            with open(os.path.join(d, name), "rb") as f:
                h.update(f.read())
    return h.hexdigest()


def cpu_baseline(cfg, raw, min_seconds=10.0, max_reps=3):
    """FP64 oracle (the timed CPU baseline) on the same scene: the whole SLC chain (I/Q, Doppler, range
    compression, range-Doppler), repeated until about min_seconds of CPU work have been timed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from asarf import scene

    x = scene.raw_to_complex(raw)
    threads = po.get_threads()
    times = []
    while len(times) < max_reps and sum(times) < min_seconds:
        t0 = time.perf_counter()
        po.focus_slc(x, cfg)
        times.append(time.perf_counter() - t0)
    dt = sum(times) / len(times)
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
    except OSError:
        pass
    return {
        "value": round(cfg.na * cfg.nr / dt / 1e6, 2),
        "unit": "Msamples/s",
        "cores": threads,
        "host_nproc": os.cpu_count(),
        "host_cpu": model,
        "kind": "port",
        "sample": f"the same {cfg.na} x {cfg.nr} scene, whole SLC chain (I/Q, DC, RC, RDA) in the FP64 oracle, "
                  f"{len(times)} run(s), {dt:.2f} s each on {threads} threads",
    }


def _rel_rms(a, b):
    import numpy as np
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def _quality(cfg, slc, ref):
    """SLC relative RMS error and worst point-target PSLR/ISLR (dB) and IRW (relative) deltas against the FP64
    oracle over every target of the scene, range and azimuth cuts."""
    from asarf import analysis, scene
    tx, ty = scene.targets(cfg)
    worst = {"pslr_db": 0.0, "islr_db": 0.0, "irw_rel": 0.0}
    n = 0
    for x, y in zip(tx, ty):
        try:
            mr = analysis.point_target(ref, x, y)
        except ValueError:
            continue
        mg = analysis.point_target(slc, x, y)
        n += 1
        for ax in ("range", "azimuth"):
            worst["pslr_db"] = max(worst["pslr_db"], abs(mg[ax]["pslr"] - mr[ax]["pslr"]))
            worst["islr_db"] = max(worst["islr_db"], abs(mg[ax]["islr"] - mr[ax]["islr"]))
            worst["irw_rel"] = max(worst["irw_rel"], abs(mg[ax]["irw"] / mr[ax]["irw"] - 1.0))
    return {"slc_rel_rms": _rel_rms(slc, ref), "targets": n, **{k: round(v, 5) for k, v in worst.items()}}


def _time_plan(plan, traw, steps, warmup=2):
    import torch
    for _ in range(warmup):
        plan.execute(traw)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        plan.execute(traw)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def forms_probe(steps=10, train_pairs=10_000_000, train_steps=2400, threshold=1e-3):
    """The two alternative forms against the plain FP32 path, each on its own workload, with quality measured
    against the FP64 oracle on the same scene (outside the headline value):
    * split-bf16 matrix-core transforms on the ERS block at range 8192 / azimuth 32768;
    * the learned RCMC interpolator (trained on the plan's own 16-tap pairs, judged per cell at `threshold`,
      16-tap sinc elsewhere) on the high-migration scene.
    speedup = plain ms / form ms; matched = every target within 0.1 dB PSLR/ISLR of the oracle."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from asarf import configs, focus, scene

    def entry(cfg, plain_ms, plain_q, form_ms, form_q, **extra):
        matched = form_q["pslr_db"] <= 0.1 and form_q["islr_db"] <= 0.1
        return {"workload": cfg.name, "plain_fp32_ms": round(plain_ms, 4), "form_ms": round(form_ms, 4),
                "speedup": round(plain_ms / form_ms, 3), "matched_quality": matched,
                "plain_vs_oracle": plain_q, "form_vs_oracle": form_q, **extra}

    out = {}
    ctx = focus.Context(0)
    # split-precision transforms
    cfg = configs.ers_block_split()
    raw, _ = scene.generate(cfg)
    traw = torch.from_numpy(raw).cuda()
    ref, _ = po.focus_slc(scene.raw_to_complex(raw), cfg, nrp=cfg.nrp, nap=cfg.nap)
    plan = focus.Plan(ctx, cfg)
    ms_fp32 = _time_plan(plan, traw, steps)
    q_fp32 = _quality(cfg, plan.slc_tensor().cpu().numpy(), ref)
    plan.set_fft(True)
    ms_bf16 = _time_plan(plan, traw, steps)
    bf = plan.slc_tensor().cpu().numpy()
    q_bf16 = _quality(cfg, bf, ref)
    mask = np.abs(ref) > 0.1 * np.abs(ref).max()
    phase = float(np.sqrt(np.mean(np.angle(bf[mask] * np.conj(ref[mask])) ** 2)))
    plan.close()
    out["split_bf16"] = entry(cfg, ms_fp32, q_fp32, ms_bf16, q_bf16, fft_sizes=[cfg.nrp, cfg.nap],
                              phase_rms_rad=phase)
    del raw, traw, ref, bf
    torch.cuda.empty_cache()
    # learned RCMC
    cfg = configs.high_migration()
    raw, _ = scene.generate(cfg)
    traw = torch.from_numpy(raw).cuda()
    ref, _ = po.focus_slc(scene.raw_to_complex(raw), cfg)
    plan = focus.Plan(ctx, cfg)
    ms_sinc = _time_plan(plan, traw, steps)
    q_sinc = _quality(cfg, plan.slc_tensor().cpu().numpy(), ref)
    t0 = time.perf_counter()
    pairs = plan.rcmc_pairs(traw, train_pairs, seed=0x7A1)
    val = plan.rcmc_pairs(traw, 1 << 18, seed=0x7A2)
    net = focus.RcmcNet(ctx, seed=3)
    net.train(pairs, batch=8192, steps=train_steps, lr=2e-3)
    table = net.build_table(val, nd=8192, ns=2, judge_nd=256, threshold=threshold)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del pairs, val
    plan.set_rcmc(focus.RCMC_LEARNED, net)
    ms_learned = _time_plan(plan, traw, steps)
    q_learned = _quality(cfg, plan.slc_tensor().cpu().numpy(), ref)
    plan.close()
    net.close()
    out["learned_rcmc"] = entry(cfg, ms_sinc, q_sinc, ms_learned, q_learned, judge_threshold=threshold,
                                judge_cells_accepted=int(table["n_ok"]), judge_cells=int(table["rms"].size),
                                train_pairs=train_pairs, train_s=round(train_s, 2))
    ctx.close()
    return out


def ingest_probe(reps=20):
    """Level-0 decoders at full scene size (synthetic bytes): ERS IM 30000 lines of 5616 u8 I/Q pairs and ASAR IM
    30000 FBAQ-4 records of 5616 samples, each into a 6400 x 30240 padded cf32 image (NaN fill and zero padding
    included). GB/s counts the input bytes plus every padded output element once."""
    import torch
    from asarf import focus
    ctx = focus.Context(0)
    lines, spl = 30000, 5616
    x_size, x_stride, y_stride = spl + 4 * 66, 6400, 30240
    g = torch.Generator(device="cuda").manual_seed(5)
    out = torch.empty((y_stride, x_stride), dtype=torch.complex64, device="cuda")
    swst = (700 + 22 * torch.randint(0, 4, (lines,), device="cuda", generator=g)).to(torch.int16)
    ers = torch.randint(0, 32, (lines, 2 * spl), device="cuda", generator=g, dtype=torch.uint8)
    blocks = (spl + 62) // 63
    rec = torch.randint(0, 256, (lines, 64 * blocks + 2), device="cuda", generator=g, dtype=torch.uint8)
    length = 64 * blocks
    rec[:, 0] = length & 0xFF
    rec[:, 1] = length >> 8
    i_lut = torch.randn(4096, device="cuda", generator=g)
    q_lut = torch.randn(4096, device="cuda", generator=g)
    res = {}
    for name, run, in_bytes in (
            ("ers_im_u8", lambda: focus.ingest_ers_im(ctx, ers, spl, swst, 700, out, x_size), ers.numel()),
            ("asar_im_fbaq4", lambda: focus.ingest_asar_fbaq4(ctx, rec, swst, 700, i_lut, q_lut, out, x_size),
             rec.numel())):
        run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        nbytes = in_bytes + 8 * x_stride * y_stride
        res[name] = {"ms": round(ms, 4), "gb_per_s": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "msamples_per_s": round(lines * spl / (ms * 1e-3) / 1e6, 1)}
    del ers, rec
    # output conditioning at the ASAR scene's product sizes: SLC 30000 x 5095 (pitch 5104) to int16 records,
    # detected 10128 x 7033 to uint16 records, 17-byte record headers
    slc = torch.randn((30240, 5104), dtype=torch.complex64, device="cuda", generator=g) * 1000
    det = torch.rand((10128, 7033), device="cuda", generator=g) * 1e4
    for name, run, nbytes in (
            ("condition_slc", lambda: focus.condition_slc(ctx, slc, 5095, 30000, 1.0, 17),
             30000 * 5095 * 8 + 30000 * (17 + 4 * 5095)),
            ("condition_detected", lambda: focus.condition_detected(ctx, det, 1.0, 17),
             det.numel() * 4 + 10128 * (17 + 2 * 7033))):
        run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[name] = {"ms": round(ms, 4), "gb_per_s": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    ctx.close()
    return {"ingest_probe": res, "lines": lines, "samples_per_line": spl, "padded_image": [x_stride, y_stride]}


def l0_plan_probe(steps=10):
    """The plan fed with level-0 input at full scene size against the same plan on raw I/Q byte pairs:
    * ASAR IM scene 30000 x 5616 as FBAQ-4 records (the scene's int8 samples quantised on the device: per block of
      63 samples the quantiser step that spans its largest sample, 16 levels per step), SWST runs over 0-6 codes
      (range size 5622, so the range transform stays 6144 long as in the raw-pair plan);
    * ERS block 4096 x 5616 as u8 level-0 lines with SWST runs over 0-19 codes (range size 5692).
    ms = one plan execution (SLC + IMP); Msamples/s counts the raw samples of the raw-pair plan."""
    import numpy as np
    import torch
    from asarf import configs, focus, scene
    ctx = focus.Context(0)
    out = {}
    g = torch.Generator(device="cuda").manual_seed(9)
    # ASAR FBAQ-4
    cfg = configs.asar_scene()
    raw, _ = scene.generate(cfg)
    traw = torch.from_numpy(raw).cuda()
    plan = focus.Plan(ctx, cfg, imp=True)
    ms_raw = _time_plan(plan, traw, steps)
    plan.close()
    na, spl = cfg.na, cfg.nr
    nblk = (spl + 62) // 63
    iq = traw.view(torch.int8).view(na, spl, 2).float()
    iq = torch.nn.functional.pad(iq, (0, 0, 0, nblk * 63 - spl)).view(na, nblk, 63, 2)
    amp = iq.abs().amax(dim=(2, 3)).clamp_min(1e-3)
    blk = torch.clamp(torch.ceil((amp / 7.5 - 0.5) / 0.1), 0, 255)
    step = (0.5 + 0.1 * blk)[:, :, None, None]
    conv = torch.clamp(torch.round(iq / step + 7.5), 0, 15).to(torch.int32)
    code = torch.where(conv >= 8, conv - 8, 15 - conv)
    byte = ((code[..., 0] << 4) | code[..., 1]).to(torch.uint8)
    data_len = 64 * (nblk - 1) + 1 + (spl - 63 * (nblk - 1))
    rec_bytes = (2 + 64 * nblk + 15) // 16 * 16
    rec = torch.zeros((na, rec_bytes), dtype=torch.uint8, device="cuda")
    body = torch.cat([blk.to(torch.uint8)[:, :, None], byte], dim=2).reshape(na, 64 * nblk)
    rec[:, 2:2 + 64 * nblk] = body
    rec[:, 2 + data_len:] = 0
    rec[:, 0] = data_len & 0xFF
    rec[:, 1] = data_len >> 8
    swst = (700 + torch.tensor([0, 3, 6], device="cuda")[torch.randint(0, 3, (na // 100 + 1,), device="cuda",
                                                                        generator=g)]).repeat_interleave(100)[:na]
    swst = swst.to(torch.int16).contiguous()
    conv_l = torch.arange(16, device="cuda", dtype=torch.float32)[:, None]
    i_lut = ((conv_l - 7.5) * (0.5 + 0.1 * torch.arange(256, device="cuda", dtype=torch.float32))[None, :]).reshape(-1)
    q_lut = (1.02 * i_lut).contiguous()
    fcfg = cfg.scaled(nr=spl + 6)
    plan = focus.Plan(ctx, fcfg, imp=True, fbaq=(rec_bytes, 1))
    run = lambda: plan.execute_fbaq4(rec, swst, 700, i_lut, q_lut)  # noqa: E731
    for _ in range(2):
        run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        run()
    e1.record()
    torch.cuda.synchronize()
    ms_fbaq = e0.elapsed_time(e1) / steps
    plan.set_timing(True)
    run()
    st = {k: round(v, 4) for k, v in plan.kernel_times().items()}
    plan.set_timing(False)
    plan.close()
    out["asar_fbaq4_records"] = {"workload": cfg.name, "record_bytes": rec_bytes,
                                 "fft_sizes": [focus.padding_size(fcfg.nr + fcfg.nc), focus.padding_size(na)],
                                 "raw_pairs_plan_ms": round(ms_raw, 4),
                                 "fbaq4_plan_ms": round(ms_fbaq, 4),
                                 "fbaq4_msamples_per_s": round(na * spl / (ms_fbaq * 1e-3) / 1e6, 1),
                                 "fbaq4_stages_ms": st}
    del traw, iq, rec, body, byte, code, conv
    torch.cuda.empty_cache()
    # ERS lines
    cfg = configs.ers_block()
    raw, _ = scene.generate(cfg)
    traw = torch.from_numpy(raw).cuda()
    plan = focus.Plan(ctx, cfg, imp=True)
    ms_raw = _time_plan(plan, traw, steps)
    plan.close()
    na, spl = cfg.na, cfg.nr
    line_bytes = 2 * spl
    lines = traw.view(na, line_bytes)
    swst = (700 + torch.tensor([0, 9, 19], device="cuda")[torch.randint(0, 3, (na // 100 + 1,), device="cuda",
                                                                        generator=g)]).repeat_interleave(100)[:na]
    swst = swst.to(torch.int16).contiguous()
    lcfg = cfg.scaled(nr=spl + 4 * 19)
    plan = focus.Plan(ctx, lcfg, imp=True, lines=(spl, line_bytes, 4))
    run = lambda: plan.execute_lines(lines, swst, 700)  # noqa: E731
    for _ in range(2):
        run()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        run()
    e1.record()
    torch.cuda.synchronize()
    ms_lines = e0.elapsed_time(e1) / steps
    plan.set_timing(True)
    run()
    st = {k: round(v, 4) for k, v in plan.kernel_times().items()}
    plan.set_timing(False)
    plan.close()
    out["ers_lines"] = {"workload": cfg.name, "fft_sizes": [lcfg.nrp or focus.padding_size(lcfg.nr + lcfg.nc),
                                                             focus.padding_size(na)],
                        "raw_pairs_plan_ms": round(ms_raw, 4), "lines_plan_ms": round(ms_lines, 4),
                        "lines_msamples_per_s": round(na * spl / (ms_lines * 1e-3) / 1e6, 1), "lines_stages_ms": st}
    ctx.close()
    return {"l0_plan_probe": out}


def stages_probe(reps=5):
    """The reference's call sequence (main.cc:243-283) through the stage entry points at the ASAR scene size:
    RawDataCorrection, CalculateDopplerCentroid, RangeCompression and RangeDopplerAlgorithm on one padded cf32
    image (30240 x 6144) made from the scene's raw bytes. Each stage is timed with events on the context's stream,
    the host synchronisations of the statistics stages included; the padded image is restored (untimed) before
    each repetition because the stages work in place."""
    import numpy as np
    import torch
    from asarf import configs, focus, scene
    cfg = configs.asar_scene()
    raw, _ = scene.generate(cfg)
    na, nr = cfg.na, cfg.nr
    nrp, nap = focus.padding_size(nr + cfg.nc), focus.padding_size(na)
    ctx = focus.Context(0)
    img0 = torch.zeros((nap, nrp), dtype=torch.complex64, device="cuda")
    r = torch.from_numpy(raw).cuda().float()
    img0[:na, :nr] = torch.complex(r[..., 0], r[..., 1])
    del r
    t = torch.empty_like(img0)
    out = torch.empty_like(img0)
    chirp = focus.generate_chirp(cfg, nrp)
    names = ("iq_correction", "doppler_centroid", "range_compression", "range_doppler")
    acc = {k: [] for k in names}
    for rep in range(reps + 1):
        t.copy_(img0)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        ev[0].record()
        focus.iq_correction(ctx, t, nr, na)
        ev[1].record()
        dc_poly, _, _ = focus.doppler_centroid(ctx, t, nr, na, cfg.prf)
        ev[2].record()
        nrv = focus.range_compression(ctx, t, nr, na, chirp, cfg.nc)
        ev[3].record()
        focus.range_doppler(ctx, focus.rda_params(cfg, dc_poly), t, nrv, na, out=out)
        ev[4].record()
        torch.cuda.synchronize()
        if rep:  # the first repetition warms up
            for i, k in enumerate(names):
                acc[k].append(ev[i].elapsed_time(ev[i + 1]))
    ctx.close()
    ms = {k: round(float(np.median(v)), 4) for k, v in acc.items()}
    total = sum(ms.values())
    return {"stages_probe": {"workload": "asar_im_30000x5616", "padded_image": [nrp, nap], "stages_ms": ms,
                             "total_ms": round(total, 4),
                             "msamples_per_s": round(na * nr / (total * 1e-3) / 1e6, 1)}}


def run_batch(args):
    """Heterogeneous batch of 64 ERS/ASAR blocks: per-block sequential launches against one captured graph."""
    import numpy as np
    import torch

    from asarf import configs, focus, scene

    cfgs = configs.batch_blocks(64)
    if args.range_pad == "fast":
        # range lines at the compile-time transform lengths (the same lines up to FP32 rounding)
        cfgs = [c.scaled(nrp=focus.fast_range_size(c.nr + c.nc)) for c in cfgs]
    t0 = time.perf_counter()
    raws = [torch.from_numpy(scene.generate(c)[0]).cuda() for c in cfgs]
    gen_s = time.perf_counter() - t0
    ctx = focus.Context(0)
    t1 = time.perf_counter()
    rows = sorted({c.nrp or focus.padding_size(c.nr + c.nc) for c in cfgs})
    cols = sorted({focus.padding_size(c.na) for c in cfgs})
    tuned = {"rows": {}, "columns": {}}
    if args.tune_file and os.path.exists(args.tune_file):
        tuned = {"loaded": focus.tune_load(ctx, args.tune_file), "file": args.tune_file}
    else:
        for n in rows:
            tuned["rows"][n] = focus.autotune(ctx, n, columns=False, batch=512)["choice"]
        for n in cols:
            tuned["columns"][n] = focus.autotune(ctx, n, columns=True, batch=1024)["choice"]
        if args.tune_file:
            focus.tune_save(ctx, args.tune_file)
    tune_s = time.perf_counter() - t1
    b = focus.Batch(ctx, n_streams=args.streams)
    for c, r in zip(cfgs, raws):
        b.add(c, r)
    samples = sum(c.na * c.nr for c in cfgs)

    def timed(mode, k):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(k):
            b.execute(mode)
        b.sync()
        return (time.perf_counter() - t) / k

    for _ in range(args.warmup):
        b.execute(focus.BATCH_SEQUENTIAL)
        b.execute(focus.BATCH_GRAPH)
    b.sync()
    steps = max(1, args.steps // 4)
    seq_s = timed(focus.BATCH_SEQUENTIAL, steps)
    graph_s = timed(focus.BATCH_GRAPH, steps)
    b.close()
    ctx.close()
    line = {
        "metric": METRIC,
        "value": round(samples / graph_s / 1e6, 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(graph_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox blocks: point targets + 10 dB noise)",
        "config": {"workload": "batch64_ers_asar_blocks", "blocks": len(cfgs), "raw_samples": samples,
                   "azimuth_lines": [c.na for c in cfgs], "range_samples": [c.nr for c in cfgs],
                   "streams": args.streams, "form": f"one captured graph (fork-join over {args.streams} streams)",
                   "range_transform": ("compile-time lengths >= the reference padding (asarf_fft_fast_range_size)"
                                       if args.range_pad == "fast" else "reference padding")},
        "sequential_msamples_per_s": round(samples / seq_s / 1e6, 2),
        "sequential_ms": round(seq_s * 1e3, 3),
        "graph_speedup": round(seq_s / graph_s, 3),
        "autotune": {"row_lengths": len(rows), "column_lengths": len(cols), "tune_s": round(tune_s, 2),
                     "choices": tuned},
        "scene_generation_s": round(gen_s, 2),
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="asar_scene")
    ap.add_argument("--no-imp", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=5)
    ap.add_argument("--rcmc", choices=["sinc16", "learned", "sinc8"], default="sinc16",
                    help="RCMC interpolator; 'learned' first trains the network on pairs from this scene")
    ap.add_argument("--fft", choices=["fp32", "split_bf16"], default="fp32",
                    help="transform precision (split_bf16: matrix-core DFT passes, nrp 8192 / nap 32768 configs)")
    ap.add_argument("--train-pairs", type=int, default=10_000_000)
    ap.add_argument("--train-steps", type=int, default=2400)
    ap.add_argument("--judge-threshold", type=float, default=1e-3)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (range and azimuth transform schedules) before the plan is built")
    ap.add_argument("--ingest-probe", action="store_true",
                    help="time the level-0 decoders (ERS u8, ASAR FBAQ-4) and the output conditioning at full scene "
                         "size and print their rates")
    ap.add_argument("--l0-probe", action="store_true",
                    help="the plan fed with FBAQ-4 records / ERS level-0 lines at full size against raw I/Q pairs")
    ap.add_argument("--stages-probe", action="store_true",
                    help="time the reference's stage sequence (I/Q correction, Doppler centroid, range compression, "
                         "range-Doppler) through the stage entry points at the ASAR scene size")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the forms object (split-bf16 and learned RCMC against the plain path on their own "
                         "workloads, about half a minute)")
    ap.add_argument("--tune-file", default=None,
                    help="batch64: autotuner cache file; loaded when it exists (no tuning launches), else written "
                         "after tuning")
    ap.add_argument("--streams", type=int, default=4, help="batch64: streams the blocks are spread over")
    ap.add_argument("--range-pad", choices=["fast", "reference"], default="fast",
                    help="batch64: range transform lengths (fast: the compile-time length at most 1.5 times the "
                         "reference's padding, same lines up to FP32 rounding)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (5 times): the known-size reference that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters of the plan's kernels")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by whole azimuth blocks per GPU, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.config == "batch64":
        run_batch(args)
        return
    if args.ingest_probe:
        print(json.dumps(ingest_probe()))
        return
    if args.stages_probe:
        print(json.dumps(stages_probe()))
        return
    if args.l0_probe:
        print(json.dumps(l0_plan_probe()))
        return
    if args.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": src.numel() * 4, "copy_probe_bytes_written": dst.numel() * 4}))
        return

    from asarf import configs, focus, scene

    cfg = configs.CONFIGS[args.config]()
    imp = not args.no_imp
    t0 = time.perf_counter()
    raw, _ = scene.generate(cfg)
    gen_s = time.perf_counter() - t0
    traw = torch.from_numpy(raw).cuda()
    ctx = focus.Context(0)
    tune = None
    if not args.no_autotune and args.fft == "fp32":
        # schedules measured on this device for this scene's transform lengths (outside the timed region)
        t1 = time.perf_counter()
        nrp = cfg.nrp or focus.padding_size(cfg.nr + cfg.nc)
        nap = cfg.nap or focus.padding_size(cfg.na)
        tune = {"range": focus.autotune(ctx, nrp, columns=False, batch=1024),
                "azimuth": focus.autotune(ctx, nap, columns=True, batch=cfg.nr - cfg.nc)}
        tune["tune_s"] = round(time.perf_counter() - t1, 2)
    plan = focus.Plan(ctx, cfg, imp=imp)
    if args.fft == "split_bf16":
        plan.set_fft(True)
    rcmc_info = {"interpolator": args.rcmc}
    net = None
    if args.rcmc == "learned":
        t1 = time.perf_counter()
        pairs = plan.rcmc_pairs(traw, args.train_pairs, seed=0x7A1)
        val = plan.rcmc_pairs(traw, 1 << 18, seed=0x7A2)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        net = focus.RcmcNet(ctx, seed=3)
        hist = net.train(pairs, batch=8192, steps=args.train_steps, lr=2e-3)
        t3 = time.perf_counter()
        table = net.build_table(val, nd=8192, ns=2, judge_nd=256, threshold=args.judge_threshold)
        ok = np.isfinite(table["rms"]) & np.isfinite(table["sinc8_rms"])
        rcmc_info.update({
            "train_pairs": args.train_pairs, "pair_generation_s": round(t2 - t1, 3), "train_steps": args.train_steps,
            "batch": 8192, "train_s": round(t3 - t2, 3), "final_loss": float(np.mean(hist[-50:])),
            "judge_threshold": args.judge_threshold, "judge_cells_accepted": table["n_ok"],
            "judge_cells": int(table["rms"].size),
            "median_rms_learned": float(np.median(table["rms"][ok])),
            "median_rms_sinc8": float(np.median(table["sinc8_rms"][ok])),
        })
        del pairs, val
        plan.set_rcmc(focus.RCMC_LEARNED, net)
    elif args.rcmc == "sinc8":
        plan.set_rcmc(focus.RCMC_SINC8)

    for _ in range(args.warmup):
        plan.execute(traw)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        plan.execute(traw)
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps
    res = plan.results()

    # per-kernel launch durations (HIP events between launches on the plan's stream)
    plan.set_timing(True)
    acc = {}
    for _ in range(args.timing_steps):
        plan.execute(traw)
        for k, v in plan.kernel_times().items():
            acc.setdefault(k, []).append(v)
    plan.set_timing(False)
    stages = {k: float(np.mean(v)) for k, v in acc.items()}
    gbs = {}
    for k, v in stages.items():
        b = kernel_bytes(k, cfg, plan, res)
        if b:
            gbs[k] = b / (v * 1e-3) / 1e9
    if not stages:  # --timing-steps 0 (profiler passes): throughput only
        print(json.dumps({"metric": METRIC, "value": round(cfg.na * cfg.nr / (ms * 1e-3) / 1e6, 2),
                          "ms_per_step": round(ms, 4)}))
        plan.close()
        ctx.close()
        return
    dom = max((k for k in stages if kernel_bytes(k, cfg, plan, res)), key=lambda k: stages[k])
    dom_bytes = kernel_bytes(dom, cfg, plan, res)
    achieved = dom_bytes / (stages[dom] * 1e-3) / 1e9
    # HBM traffic and VALU issue occupancy come from the committed counter passes (profiles/traffic.json,
    # profiles/issue.json); each records the hash of the kernel sources it was measured on, and a figure whose
    # sources have changed since is reported as null (stale) rather than replayed
    src_sha = kernel_sources_sha1()
    traffic, traffic_current = None, None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        traffic_current = tj.get("_sources_sha1") == src_sha
        if traffic_current:
            traffic = tj.get(cfg.name, {}).get(dom)
    # VALU issue occupancy per stage (SQ_ACTIVE_INST_VALU against the SIMD cycles of the dispatch): near 1 means
    # the kernel is bound by its instruction count rather than by HBM
    valu_busy, issue_current = {}, None
    ifile = os.path.join(ROOT, "profiles", "issue.json")
    if os.path.exists(ifile):
        with open(ifile) as f:
            ij = json.load(f)
        issue_current = ij.get("_sources_sha1") == src_sha
        if issue_current:
            valu_busy = {k: v["valu_busy"] for k, v in ij.get(cfg.name, {}).items()}

    # measured device-to-device copy rate over a buffer the size of the plan's image: the second denominator of
    # the memory-bound stages (what a plain read + write of the same bytes achieves on this device)
    nbytes = plan.params.nap * (cfg.nr - cfg.nc) * 8
    csrc = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    cdst = torch.empty_like(csrc)
    cdst.copy_(csrc)
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record()
    for _ in range(5):
        cdst.copy_(csrc)
    c1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * nbytes / (c0.elapsed_time(c1) / 5 * 1e-3) / 1e9
    del csrc, cdst

    value = cfg.na * cfg.nr / (ms * 1e-3) / 1e6
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x2 (hi/lo) MFMA, f32 accumulate" if args.fft == "split_bf16" else "f32",
        "data": "synthetic (seeded Philox scene: point targets in gamma-textured clutter)",
        "config": {
            "workload": cfg.name,
            "azimuth_lines": cfg.na,
            "range_samples": cfg.nr,
            "raw_format": cfg.raw_format,
            "fft_sizes": [plan.params.nrp, plan.params.nap],
            "imp": imp,
            "form": ("split-bf16 MFMA DFT passes, " if args.fft == "split_bf16" else "plain FP32 Stockham, ") + {"sinc16": "16-tap sinc RCMC", "sinc8": "8-tap sinc RCMC",
                                                 "learned": "learned 8-tap RCMC (16-tap sinc where judged)"}[args.rcmc],
            "unit_of_work": "one raw input sample (azimuth line x range sample) taken from the device-resident raw "
                            "buffer to the focused SLC and IMP; slc_output_msamples_per_s counts output samples",
        },
        "rcmc": rcmc_info,
        "autotune": tune,
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes": dom_bytes,
            "ms": round(stages[dom], 4),
            "copy_gbs_measured": round(copy_gbs, 1),
            "frac_of_copy": round(achieved / copy_gbs, 4),
            "valu_busy": valu_busy.get(dom),
            "valu_busy_stages": valu_busy or None,
            "profiles_current": {"traffic": traffic_current, "issue": issue_current},
        },
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "stages_gbs": {k: round(v, 1) for k, v in gbs.items()},
        "stages_frac_of_copy": {k: round(v / copy_gbs, 3) for k, v in gbs.items()},
        "slc_output_msamples_per_s": round(cfg.na * (cfg.nr - cfg.nc) / (ms * 1e-3) / 1e6, 2),
        "plan_device_gb": round(plan.device_bytes / 1e9, 3),
        "scene_generation_s": round(gen_s, 2),
        "estimates": {"dc_poly": res["dc_poly"], "iq_gain": res["iq"]["iq_gain"],
                      "phase_error_deg": res["iq"]["phase_error"], "imp_rows": res["imp_rows"]},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(cfg, raw)
    plan.close()
    if net is not None:
        net.close()
    ctx.close()
    if not args.no_forms and args.fft == "fp32" and args.rcmc == "sinc16":
        del traw
        torch.cuda.empty_cache()
        line["forms"] = forms_probe()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
