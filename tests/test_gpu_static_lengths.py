"""The compile-time kernels that serve arbitrary scene sizes: column transforms of every 2^a 3^b 5^c 7^d length from
16 to 400 (16- and 32-column tiles, and the 48-column tiles of the long strided first passes) against numpy's FFT in
FP64, the fused azimuth core at every compiled second-pass length against the separate kernels and the FP64
oracle, and range transforms padded past the reference's length (asarf_fft_fast_range_size) against the oracle at
the reference's padding."""
import os

import numpy as np
import pytest
import torch

import pyoracle as po
from asarf import _lib as L
from asarf import configs, focus, scene
from parity import rel_rms

pytestmark = pytest.mark.gpu


def smooth(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


STATIC_LENGTHS = [n for n in range(16, 401) if smooth(n)]
CORE_LENGTHS = [84, 90, 98, 96, 80, 100, 72, 105, 81, 112, 64, 120, 75, 126, 128, 125]


class _env:
    def __init__(self, **kv):
        self.kv = kv

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kv}
        os.environ.update(self.kv)

    def __exit__(self, *exc):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("width", ["16", "32"])
def test_every_static_column_length(width):
    with _env(ASARF_COL_WIDTH=width):
        ctx = focus.Context(0)  # the width is read when the context is created
    try:
        for n in STATIC_LENGTHS:
            assert focus.static_column_length(n), n
            cols = 37
            x = (np.random.default_rng(n).standard_normal((n, cols, 2)) @ [1, 1j]).astype(np.complex64)
            t = torch.from_numpy(x).cuda()
            focus.fft(ctx, t, L.FFT_COLUMNS)
            ref = np.fft.fft(x.astype(np.complex128), axis=0)
            err = rel_rms(t.cpu().numpy(), ref)
            assert err < 2e-6, (n, err)
            focus.fft(ctx, t, L.FFT_COLUMNS, inverse=True, scale=1.0 / n)
            err = rel_rms(t.cpu().numpy(), x)
            assert err < 4e-6, (n, err)
    finally:
        ctx.close()


def _run(ctx, cfg, raw, env):
    with _env(**env):
        plan = focus.Plan(ctx, cfg)
    try:
        plan.execute(torch.from_numpy(raw).cuda())
        return plan.slc_tensor().cpu().numpy(), plan.results(), plan.params.nap
    finally:
        plan.close()


@pytest.mark.parametrize("n2", CORE_LENGTHS)
def test_core_second_pass_length(ctx, n2):
    """ERS block of nap - 1 lines (padded to nap = n2 x m): the core forced to n2 against the separate kernels on
    the same raw data, and both against the oracle."""
    m = next(k for k in (6, 7, 8, 9, 10, 12) if n2 * k >= 600)
    nap = n2 * m
    cfg = configs.ers_block().scaled(na=nap - 1, nr=1500, n_targets_az=2, n_targets_rg=2, tgt_az=(0.4, 0.6))
    assert focus.padding_size(cfg.na) == nap
    raw, _ = scene.generate(cfg)
    slc_f, res_f, _ = _run(ctx, cfg, raw, {"ASARF_AZ_CORE_N2": str(n2)})
    slc_s, res_s, _ = _run(ctx, cfg, raw, {"ASARF_AZ_FUSED": "0"})
    assert res_f["azimuth_core_fused"] and not res_s["azimuth_core_fused"], n2
    ref, _ = po.focus_slc(scene.raw_to_complex(raw), cfg)
    err, err_f = rel_rms(slc_f, slc_s), rel_rms(slc_f, ref)
    print(f"n2 {n2} nap {nap}: fused vs separate {err:.3e}, vs oracle {err_f:.3e}", flush=True)
    assert err < 1e-6
    assert err_f < 1e-4


@pytest.mark.parametrize("nr,nc", [(4100, 420), (5200, 500), (5616, 704)])
def test_fast_range_length_matches_reference_padding(ctx, nr, nc):
    """Range lines transformed at the compile-time length asarf_fft_fast_range_size(nr + nc) instead of the
    reference's padding: the SLC matches the oracle run at the reference's padding."""
    fs = 19.208e6
    cfg = configs.asar_scene().scaled(na=700, nr=nr, n_targets_random=4, pulse_len=nc / fs, fs=fs)
    pad, fast = focus.padding_size(cfg.nr + cfg.nc), focus.fast_range_size(cfg.nr + cfg.nc)
    assert fast in (6144, 6400) and fast >= pad
    raw, _ = scene.generate(cfg)
    slc, res, _ = _run(ctx, cfg.scaled(nrp=fast), raw, {})
    ref, _ = po.focus_slc(scene.raw_to_complex(raw), cfg)
    err = rel_rms(slc, ref)
    print(f"nr {nr} nc {cfg.nc}: padding {pad} -> {fast}, slc rel rms {err:.3e}", flush=True)
    assert err < 1e-4


def test_strided_first_pass_of_every_long_length(tmp_path):
    """Two-pass column transforms n1 x 16 whose strided first pass has a long compile-time length (n1 > 310: the
    48-column tiles), the split forced through the tuner cache file, against numpy's FFT in FP64."""
    longs = [n for n in STATIC_LENGTHS if n > 310]
    ctx = focus.Context(0)
    path = str(tmp_path / "tune.txt")
    focus.tune_save(ctx, path)  # the header with this device's architecture
    ctx.close()
    with open(path, "a") as f:
        for n1 in longs:
            f.write(f"col {n1 * 16} {n1}\n")
    ctx = focus.Context(0)
    try:
        assert focus.tune_load(ctx, path) == len(longs)
        for n1 in longs:
            n = n1 * 16
            x = (np.random.default_rng(n1).standard_normal((n, 37, 2)) @ [1, 1j]).astype(np.complex64)
            t = torch.from_numpy(x).cuda()
            focus.fft(ctx, t, L.FFT_COLUMNS)
            err = rel_rms(t.cpu().numpy(), np.fft.fft(x.astype(np.complex128), axis=0))
            assert err < 2e-6, (n1, err)
            focus.fft(ctx, t, L.FFT_COLUMNS, inverse=True, scale=1.0 / n)
            err = rel_rms(t.cpu().numpy(), x)
            assert err < 4e-6, (n1, err)
    finally:
        ctx.close()
