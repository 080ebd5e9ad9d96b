"""Split-precision (bfloat16 hi/lo on the matrix cores) transforms: accuracy of the FFT kernels against numpy's
FP64 FFT, and the split-precision form of the ERS block (range 8192, azimuth 32768) against the FP64 oracle at
the same padded sizes: SLC relative RMS error, phase RMS error and point-target PSLR/ISLR."""
import numpy as np
import pytest
import torch

import pyoracle as po
from asarf import _lib as L
from asarf import analysis, configs, focus, scene

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("inverse", [False, True])
def test_rows_8192(ctx, inverse):
    rng = np.random.default_rng(1)
    x = (rng.standard_normal((37, 8192)) + 1j * rng.standard_normal((37, 8192))).astype(np.complex64)
    t = torch.from_numpy(x.copy()).cuda()
    focus.fft(ctx, t, L.FFT_ROWS, inverse=inverse, split_bf16=True)
    ref = np.fft.ifft(x.astype(np.complex128), axis=1) * 8192 if inverse else np.fft.fft(x.astype(np.complex128), axis=1)
    err = rel(t.cpu().numpy(), ref)
    print(f"rows 8192 inverse={inverse}: rel err {err:.2e}")
    assert err < 3e-5


@pytest.mark.parametrize("n", [128, 256, 32768])
def test_columns(ctx, n):
    rng = np.random.default_rng(n)
    cols = 37 if n < 1000 else 21
    x = (rng.standard_normal((n, cols)) + 1j * rng.standard_normal((n, cols))).astype(np.complex64)
    for inverse in (False, True):
        t = torch.from_numpy(x.copy()).cuda()
        focus.fft(ctx, t, L.FFT_COLUMNS, inverse=inverse, split_bf16=True)
        xd = x.astype(np.complex128)
        ref = np.fft.ifft(xd, axis=0) * n if inverse else np.fft.fft(xd, axis=0)
        err = rel(t.cpu().numpy(), ref)
        print(f"columns {n} inverse={inverse}: rel err {err:.2e}")
        assert err < 3e-5


def test_unsupported_lengths_are_rejected(ctx):
    t = torch.zeros((4, 6144), dtype=torch.complex64, device="cuda")
    with pytest.raises(RuntimeError):
        focus.fft(ctx, t, L.FFT_ROWS, split_bf16=True)


def test_split_bf16_plan_of_the_ers_block(ctx):
    """ERS block with range 8192 and azimuth 32768 transforms, split-bf16 against the FP64 oracle at the same
    sizes and against the FP32 plan."""
    cfg = configs.ers_block_split()
    raw, _ = scene.generate(cfg)
    traw = torch.from_numpy(raw).cuda()
    plan = focus.Plan(ctx, cfg)
    plan.execute(traw)
    fp32 = plan.slc_tensor().cpu().numpy()
    plan.set_fft(True)
    plan.execute(traw)
    bf = plan.slc_tensor().cpu().numpy()
    plan.close()
    ref, _ = po.focus_slc(scene.raw_to_complex(raw), cfg, nrp=cfg.nrp, nap=cfg.nap)
    e_bf, e_fp32 = rel(bf, ref), rel(fp32, ref)
    mask = np.abs(ref) > 0.1 * np.abs(ref).max()
    phase = float(np.sqrt(np.mean(np.angle(bf[mask] * np.conj(ref[mask])) ** 2)))
    print(f"ERS split-bf16: SLC rel RMS {e_bf:.2e} (FP32 {e_fp32:.2e}), phase RMS {phase:.2e} rad on {mask.sum()} px")
    worst = 0.0
    tx, ty = scene.targets(cfg)
    for x, y in zip(tx, ty):
        mg = analysis.point_target(bf, x, y)
        mr = analysis.point_target(ref, x, y)
        for ax in ("range", "azimuth"):
            worst = max(worst, abs(mg[ax]["pslr"] - mr[ax]["pslr"]), abs(mg[ax]["islr"] - mr[ax]["islr"]))
    print(f"split-bf16 worst PSLR/ISLR delta vs oracle {worst:.4f} dB")
    assert e_bf < 1e-4
    assert phase < 1e-3
    assert worst < 0.05
