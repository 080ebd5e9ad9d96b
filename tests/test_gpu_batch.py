"""Autotuner and batched focusing on the GPU: tuned schedules keep the results, a heterogeneous batch gives the
same SLCs sequentially and as one captured graph, and repeated graph replays (with eager work between them)
reproduce the eager result every time."""
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
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def test_autotuned_schedules_keep_parity():
    ctx = focus.Context(0)
    cfg = configs.asar_scene().scaled(na=700, nr=1400, n_targets_random=4)
    nrp = focus.padding_size(cfg.nr + cfg.nc)
    nap = focus.padding_size(cfg.na)
    rows = focus.autotune(ctx, nrp, columns=False, batch=256)
    cols = focus.autotune(ctx, nap, columns=True, batch=512)
    print(f"rows {nrp}: {rows}; columns {nap}: {cols}")
    assert rows["candidates"] >= 1 and rows["ms"] > 0
    assert cols["candidates"] >= 1 and cols["ms"] > 0 and nap % cols["choice"] == 0
    raw, _ = scene.generate(cfg)
    plan = focus.Plan(ctx, cfg)
    plan.execute(torch.from_numpy(raw).cuda())
    slc = plan.slc_tensor().cpu().numpy()
    plan.close()
    ref, _ = po.focus_slc(scene.raw_to_complex(raw), cfg)
    assert rel_rms(slc, ref) < 1e-4
    # a length with a compile-time kernel: the tuner weighs it against runtime radix orders
    r6400 = focus.autotune(ctx, 6400, columns=False, batch=128)
    assert r6400["candidates"] >= 2
    ctx.close()


def test_tuner_cache_file_round_trip(tmp_path):
    """Choices saved from one context load into another and drive its plans; the file names the device
    architecture and every choice of the first context."""
    ctx = focus.Context(0)
    cfg = configs.asar_scene().scaled(na=700, nr=1400, n_targets_random=4)
    nrp = focus.padding_size(cfg.nr + cfg.nc)
    nap = focus.padding_size(cfg.na)
    rows = focus.autotune(ctx, nrp, columns=False, batch=256)
    cols = focus.autotune(ctx, nap, columns=True, batch=512)
    path = tmp_path / "tune.txt"
    focus.tune_save(ctx, str(path))
    text = path.read_text().splitlines()
    assert text[0].startswith("asarf-tune 1 gfx950")
    assert f"col {nap} {cols['choice']}" in text
    assert any(t.startswith(f"row {nrp} ") for t in text)
    ctx.close()
    ctx2 = focus.Context(0)
    assert focus.tune_load(ctx2, str(path)) == 2
    # saving the loaded context reproduces the file
    path2 = tmp_path / "tune2.txt"
    focus.tune_save(ctx2, str(path2))
    assert path2.read_text() == path.read_text()
    raw, _ = scene.generate(cfg)
    plan = focus.Plan(ctx2, cfg)
    plan.execute(torch.from_numpy(raw).cuda())
    slc = plan.slc_tensor().cpu().numpy()
    plan.close()
    ref, _ = po.focus_slc(scene.raw_to_complex(raw), cfg)
    assert rel_rms(slc, ref) < 1e-4
    # another architecture's file is refused and changes nothing
    other = tmp_path / "other.txt"
    other.write_text("asarf-tune 1 gfx942\ncol 30240 168\n")
    with pytest.raises(RuntimeError, match="tuned on gfx942"):
        focus.tune_load(ctx2, str(other))
    ctx2.close()


def test_batch_graph_equals_sequential_and_replays(ctx):
    cfgs = configs.batch_blocks(6, na_range=(300, 900))
    cfgs = [c.scaled(nr=min(c.nr, 1500)) for c in cfgs]
    raws = [torch.from_numpy(scene.generate(c)[0]).cuda() for c in cfgs]
    b = focus.Batch(ctx, n_streams=3)
    for c, r in zip(cfgs, raws):
        b.add(c, r)
    b.execute(focus.BATCH_SEQUENTIAL)
    seq = [b.slc_tensor(i).cpu().numpy() for i in range(len(cfgs))]
    for rep in range(3):
        b.execute(focus.BATCH_GRAPH)
        b.sync()
        got = [b.slc_tensor(i).cpu().numpy() for i in range(len(cfgs))]
        for i in range(len(cfgs)):
            assert np.array_equal(got[i], seq[i]), (rep, i)
        # eager work on the same buffers between replays
        b.execute(focus.BATCH_SEQUENTIAL)
        torch.cuda.synchronize()
    ref, _ = po.focus_slc(scene.raw_to_complex(scene.generate(cfgs[1])[0]), cfgs[1])
    err = rel_rms(seq[1], ref)
    print(f"batch block 1 ({cfgs[1].name}) slc rel rms {err:.2e}")
    assert err < 1e-4
    b.close()


@pytest.mark.parametrize("which,nvar", [("asar", 10), ("ers", 6), ("ers8192", 7)])
def test_range_compression_variants_match_oracle(ctx, which, nvar):
    """Every compile-time range-compression variant the autotuner may pick gives the oracle's SLC."""
    import os
    if which == "asar":
        cfg = configs.asar_scene().scaled(na=256, n_targets_random=3)
    elif which == "ers":
        cfg = configs.ers_block().scaled(na=256, n_targets_az=1, n_targets_rg=3, tgt_az=(0.5, 0.5))
    else:
        cfg = configs.ers_block().scaled(na=256, n_targets_az=1, n_targets_rg=3, tgt_az=(0.5, 0.5), nrp=8192)
    raw, _ = scene.generate(cfg)
    traw = torch.from_numpy(raw).cuda()
    ref, _ = po.focus_slc(scene.raw_to_complex(raw), cfg)
    for v in range(nvar):
        os.environ["ASARF_RC_VARIANT"] = str(v)
        try:
            plan = focus.Plan(ctx, cfg)
            plan.execute(traw)
            slc = plan.slc_tensor().cpu().numpy()
            plan.close()
        finally:
            del os.environ["ASARF_RC_VARIANT"]
        err = rel_rms(slc, ref)
        print(f"{which} variant {v}: slc rel rms {err:.2e}")
        assert err < 1e-4, v
