"""CPU checks of the C ABI: the library loads, exports every declared entry point, its host helpers agree with
the oracle, and argument validation rejects bad calls before any device work (no GPU needed)."""
import ctypes
import re

import numpy as np
import pytest

import pyoracle as po
from asarf import _lib as L
from asarf import configs
from conftest import golden


def declared_functions():
    text = open(L.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(asarf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)


def test_version_and_window_size():
    assert L.lib().asarf_version().decode().startswith("asarf")
    assert L.lib().asarf_rcmc_window_size() == 16


@pytest.mark.parametrize("n", [1, 4096, 5616 + 704, 5616 + 521, 16384, 30000, 131072])
def test_padding_rule_matches_oracle(n):
    assert L.lib().asarf_fft_padding_size(n) == po.next_fft_size(n)


@pytest.mark.parametrize("window", [0, 1])
def test_chirp_is_bit_identical_to_oracle(window):
    c = configs.ers_block()
    ci = L.ChirpInfo(c.fs, c.pulse_len, c.nc, c.bandwidth, c.kr, window)
    out = np.zeros(6400, np.complex64)
    L.check(L.lib().asarf_generate_chirp(ctypes.byref(ci), 6400, out.ctypes.data))
    ref = po.generate_chirp(c.kr, c.fs, c.nc, window, 6400)
    assert np.array_equal(out, ref)


def test_polyfit_matches_numpy_golden():
    g = golden("polyfit_vectors.npz")
    for key, deg in (("dc", 2), ("sr", 4), ("r3", 3)):
        x, y, ref = g[f"{key}_x"], g[f"{key}_y"], g[f"{key}_c"]
        c = np.zeros(deg + 1)
        L.check(L.lib().asarf_polyfit(x.ctypes.data, y.ctypes.data, x.size, deg, c.ctypes.data))
        assert np.abs(np.polyval(c, x) - np.polyval(ref, x)).max() <= 1e-9 * max(1.0, np.abs(y).max())
        assert abs(L.lib().asarf_polyval(c.ctypes.data, deg + 1, x[3]) - np.polyval(c, x[3])) < 1e-9 * abs(y).max()


def test_srgr_spherical_fit_is_monotone_and_consistent():
    g = np.zeros(5)
    n = ctypes.c_int()
    r0, dr = 850e3, 7.804
    L.check(L.lib().asarf_srgr_fit_spherical(r0, dr, 5079, 6371e3 + 790e3, 6371e3, 12.5, g.ctypes.data,
                                             ctypes.byref(n)))
    assert 7000 < n.value < 11000
    gd = np.linspace(0, n.value * 12.5, 200)
    s = np.polyval(g, gd)
    assert abs(s[0]) < 0.5 and np.all(np.diff(s) > 0)
    # far end: slant offset of the last fitted sample (99 * (5079 // 100) samples) is reached at g_99
    assert abs(np.polyval(g, n.value * 12.5) - 99 * (5079 // 100) * dr) < 20.0


def test_validation_without_device():
    lib = L.lib()
    ctx = ctypes.c_void_p()
    assert lib.asarf_ctx_create(-1, None, ctypes.byref(ctx)) == L.EINVAL
    L.check(lib.asarf_ctx_create(0, None, ctypes.byref(ctx)))  # deferred initialisation: no device touched
    try:
        img = L.Image(None, 10, 10, 16, 16, L.CF32)
        r = L.IqResults()
        assert lib.asarf_iq_correction(ctx, ctypes.byref(img), 0, 7, ctypes.byref(r)) == L.EINVAL
        img = L.Image(ctypes.c_void_p(0x1000), 20, 10, 16, 16, L.CF32)  # x_size > x_stride
        assert lib.asarf_iq_correction(ctx, ctypes.byref(img), 0, 7, ctypes.byref(r)) == L.ESHAPE
        assert "Stride smaller" in L.last_error()
        img = L.Image(ctypes.c_void_p(0x1000), 10, 10, 16, 16, L.CF32)
        assert lib.asarf_iq_correction(ctx, ctypes.byref(img), 0, 64, ctypes.byref(r)) == L.EINVAL  # unknown quirk
        chirp = np.zeros(32, np.complex64)
        assert lib.asarf_range_compression(ctx, ctypes.byref(img), chirp.ctypes.data, 32, 4, None) == L.ESHAPE
        assert "chirp size mismatch" in L.last_error()
        assert lib.asarf_range_compression(ctx, ctypes.byref(img), chirp.ctypes.data, 16, 10, None) == L.EINVAL
        p = L.RdaParams()
        p.wavelength, p.range_spacing, p.slant_range_first, p.prf, p.az_bandwidth_fraction = 0.05, 7.9, 8e5, 0.0, 0.8
        out = L.Image()
        assert lib.asarf_range_doppler(ctx, ctypes.byref(p), ctypes.byref(img), ctypes.byref(out), None) == L.EINVAL
        p.prf = 1600.0
        p.vr_poly[2] = -1.0
        assert lib.asarf_range_doppler(ctx, ctypes.byref(p), ctypes.byref(img), ctypes.byref(out), None) == L.EINVAL
        p.vr_poly[2] = 7000.0
        ws = L.Workspace(ctypes.c_void_p(0x2000), 100)  # too small to adopt
        assert lib.asarf_range_doppler(ctx, ctypes.byref(p), ctypes.byref(img), ctypes.byref(out),
                                       ctypes.byref(ws)) == L.EWORKSPACE
        poly = np.zeros(3)
        assert lib.asarf_doppler_centroid(ctx, ctypes.byref(img), 1600.0, 7, poly.ctypes.data, None, None, 0,
                                          None) == L.ESHAPE  # fewer than 32 range samples
        assert lib.asarf_doppler_centroid(ctx, ctypes.byref(img), -1.0, 7, poly.ctypes.data, None, None, 0,
                                          None) == L.EINVAL
        fp = L.FocusParams()
        fp.na, fp.nr, fp.nc, fp.raw_format = 100, 200, 300, L.RAW_U8
        h = ctypes.c_void_p()
        chirp = np.zeros(512, np.complex64)
        assert lib.asarf_plan_create(ctx, ctypes.byref(fp), chirp.ctypes.data, ctypes.byref(h)) == L.ESHAPE
        fp.nc = 50
        fp.raw_format = 9
        assert lib.asarf_plan_create(ctx, ctypes.byref(fp), chirp.ctypes.data, ctypes.byref(h)) == L.EINVAL
        assert lib.asarf_fft_c2c(ctx, None, 64, 1, 64, 0, 0, 1.0) == L.EINVAL
        assert lib.asarf_fft_c2c(ctx, ctypes.c_void_p(0x1000), 64, 1, 32, 0, 0, 1.0) == L.ESHAPE
        assert lib.asarf_srgr(ctx, ctypes.c_void_p(0x1000), 1, 4, poly.ctypes.data, 7.9, 12.5, 10,
                              ctypes.c_void_p(0x2000)) == L.EINVAL
        assert lib.asarf_az_multilook(ctx, ctypes.byref(img), 0.0, ctypes.byref(ws), ctypes.byref(ctypes.c_int()),
                                      None) == L.EINVAL
    finally:
        lib.asarf_ctx_destroy(ctx)


@pytest.mark.parametrize("text, what", [
    ("", "empty"),
    ("# only a comment\n", "empty"),
    ("asarf-tune 2 gfx950\n", "version 1"),
    ("asarf-tune 1\n", "architecture"),
    ("asarf-tune 1 gfx950\ncol 30240 11\n", "col:"),           # 11 does not divide 30240
    ("asarf-tune 1 gfx950\ncol 30240\n", "col:"),
    ("asarf-tune 1 gfx950\nrow 6144 static 99\n", "variant"),
    ("asarf-tune 1 gfx950\nrow 6144 radix 8 8 8 8\n", "radix order"),   # product 4096
    ("asarf-tune 1 gfx950\nrow 6144 radix 8 8 x\n", "radix order"),
    ("asarf-tune 1 gfx950\nrow 6144 fast\n", "'static' or 'radix'"),
    ("asarf-tune 1 gfx950\nrows 6144 static 0\n", "unknown keyword"),
    ("asarf-tune 1 gfx950\ncol -5 1\n", "col:"),
])
def test_tuner_cache_rejects_malformed_files_without_device(tmp_path, text, what):
    """A malformed autotuner cache file is refused while parsing, before any device call."""
    lib = L.lib()
    ctx = ctypes.c_void_p()
    L.check(lib.asarf_ctx_create(0, None, ctypes.byref(ctx)))
    try:
        path = tmp_path / "tune.txt"
        path.write_text(text)
        n = ctypes.c_int(-1)
        assert lib.asarf_tune_load(ctx, str(path).encode(), ctypes.byref(n)) == L.EINVAL
        assert what in L.last_error()
        assert n.value == 0
        assert lib.asarf_tune_load(ctx, str(tmp_path / "missing.txt").encode(), ctypes.byref(n)) == L.EINVAL
    finally:
        lib.asarf_ctx_destroy(ctx)


def test_memory_model_trading_without_device():
    """ReleaseMemory / InitExtPtr semantics (device_padded_image.cuh:49-61,104-111) are pure bookkeeping."""
    lib = L.lib()
    img = L.Image(ctypes.c_void_p(0x4000), 100, 50, 128, 64, L.CF32)
    assert lib.asarf_image_bytes(ctypes.byref(img)) == 128 * 64 * 8
    ws = L.Workspace()
    L.check(lib.asarf_image_release(ctypes.byref(img), ctypes.byref(ws)))
    assert ws.data == 0x4000 and ws.bytes == 128 * 64 * 8 and img.data is None
    out = L.Image()
    assert lib.asarf_image_adopt(ctypes.byref(out), 100, 50, 128, 65, L.CF32, ctypes.byref(ws)) == L.EWORKSPACE
    L.check(lib.asarf_image_adopt(ctypes.byref(out), 100, 50, 128, 64, L.CF32, ctypes.byref(ws)))
    assert out.data == 0x4000 and ws.data is None and ws.bytes == 0
    assert lib.asarf_image_adopt(ctypes.byref(out), 1, 1, 1, 1, L.CF32, ctypes.byref(ws)) == L.EINVAL


def test_static_length_queries_without_device():
    """Column lengths with compile-time kernels are exactly the 2^a 3^b 5^c 7^d from 16 to 400; the fast range
    length is the reference's padding or a compile-time length (6144, 6400, 8192) at most 1.5 times it."""
    def smooth(n):
        for p in (2, 3, 5, 7):
            while n % p == 0:
                n //= p
        return n == 1

    lib = L.lib()
    for n in range(1, 1200):
        assert lib.asarf_fft_static_column_length(n) == int(16 <= n <= 400 and smooth(n)), n
    for n in range(100, 9000, 7):
        pad, fast = lib.asarf_fft_padding_size(n), lib.asarf_fft_fast_range_size(n)
        assert fast >= pad > n
        if fast != pad:
            assert fast in (6144, 6400, 8192) and 2 * fast <= 3 * pad, (n, pad, fast)
        elif pad not in (6144, 6400, 8192):
            assert not any(pad <= s <= 1.5 * pad for s in (6144, 6400, 8192)), (n, pad)
    assert lib.asarf_fft_fast_range_size(5616 + 521) == 6144 and lib.asarf_fft_fast_range_size(4500) == 6144
    assert lib.asarf_fft_fast_range_size(6200) == 6400 and lib.asarf_fft_fast_range_size(100) == 105
