"""Two contexts on one device, created with different column-transform settings and driven from two host threads
at once: each keeps its own settings (they live in the context, not in the process) and both reproduce the
FP64 oracle's SLC."""
import os
import threading

import numpy as np
import pytest
import torch

import pyoracle as po
from asarf import configs, focus, scene
from parity import rel_rms

pytestmark = pytest.mark.gpu


def _make(cfg, env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        ctx = focus.Context(0, torch.cuda.Stream())
        plan = focus.Plan(ctx, cfg)  # the context reads its settings here, at initialisation
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return ctx, plan


def test_two_contexts_two_threads_keep_their_own_settings():
    # 2400 lines pad to 2401 = 49 x 49: the column passes have a compile-time kernel that ASARF_COL_STATIC=0 turns
    # off, so the two contexts run different kernels
    cfg = configs.asar_scene().scaled(na=2400, nr=1400, n_targets_random=4)
    raw, _ = scene.generate(cfg)
    traw = torch.from_numpy(raw).cuda()
    ref, _ = po.focus_slc(scene.raw_to_complex(raw), cfg)
    made = {"width16": _make(cfg, {"ASARF_COL_WIDTH": "16"}),
            "runtime_radix": _make(cfg, {"ASARF_COL_STATIC": "0"}),
            "default": _make(cfg, {})}
    out, errs = {}, []

    def work(name):
        try:
            ctx, plan = made[name]
            for _ in range(3):
                plan.execute(traw)
            slc = plan.slc_tensor()
            ctx.sync()
            out[name] = slc.cpu().numpy()
        except Exception as e:  # reported in the main thread
            errs.append((name, e))

    torch.cuda.synchronize()
    threads = [threading.Thread(target=work, args=(k,)) for k in made]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errs, errs
    for name, slc in out.items():
        err = rel_rms(slc, ref)
        print(f"{name}: slc rel rms {err:.3e}")
        assert err < 1e-4
    # the runtime-radix kernels round differently from the compile-time ones: had the settings leaked between
    # contexts the two results would be identical
    assert not np.array_equal(out["runtime_radix"], out["default"])
    for ctx, plan in made.values():
        plan.close()
        ctx.close()
