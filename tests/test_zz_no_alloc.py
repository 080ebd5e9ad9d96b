"""asarf_doppler_centroid_noalloc, the allocation-free variant of asarf_doppler_centroid (its lag sums live in a
fixed device array instead of the context's scratch), against that entry point through the C ABI: two calls in
a row, the second on a smaller image, so lag sums left over from the first call in the fixed array must not leak
into the second. Same kernel and summation order, so the results are bit-identical."""
import ctypes

import numpy as np
import pytest
import torch

from asarf import _lib as L
from asarf import focus

pytestmark = pytest.mark.gpu


def _ramp_image(ys, xs, stride, f_near, f_far, prf, seed):
    """Lines with an azimuth phase ramp whose frequency runs linearly across range, plus noise."""
    rng = np.random.default_rng(seed)
    x = np.arange(stride)
    f = f_near + (f_far - f_near) * x / max(1, xs - 1)
    y = np.arange(ys)[:, None]
    img = np.exp(2j * np.pi * f[None, :] * y / prf) * (1.0 + 0.1 * rng.standard_normal((ys, stride)))
    img = img + 0.2 * (rng.standard_normal((ys, stride)) + 1j * rng.standard_normal((ys, stride)))
    return img.astype(np.complex64)


def _call(fn, ctx, t, xs, ys, prf, quirks):
    img = focus.image(t, xs, ys)
    poly = np.zeros(3)
    bf = np.zeros(64)
    bx = np.zeros(64)
    nb = ctypes.c_int()
    st = fn(ctx.h, ctypes.byref(img), prf, quirks, poly.ctypes.data, bf.ctypes.data, bx.ctypes.data, 64,
            ctypes.byref(nb))
    L.check(st)
    return poly, bf[:nb.value].copy(), bx[:nb.value].copy()


def test_noalloc_doppler_centroid_equals_allocating_entry_point():
    lib = L.lib()
    ctx = focus.Context(0)
    prf = 1652.0
    try:
        # (lines, range size, row pitch, quirks): 1024 lines fill 62 strips x 33 blocks of lag sums, 40 lines
        # only 39 strips, so the second call leaves the first call's sums in the upper part of the fixed array
        for ys, xs, stride, quirks, seed in [(1024, 1400, 1536, L.QUIRKS_REFERENCE, 1), (40, 700, 704, 0, 2)]:
            host = _ramp_image(ys, xs, stride, -180.0, 240.0, prf, seed)
            t = torch.from_numpy(host).cuda()
            ref = _call(lib.asarf_doppler_centroid, ctx, t, xs, ys, prf, quirks)
            got = _call(lib.asarf_doppler_centroid_noalloc, ctx, t, xs, ys, prf, quirks)
            print(f"{ys}x{xs}: {len(ref[1])} blocks, poly {got[0]}", flush=True)
            for a, b in zip(ref, got):
                assert np.array_equal(a, b)
            # the estimate is the ramp's: within a few Hz of the centroid at the middle of the range
            mid = np.polyval(got[0], xs / 2)
            assert abs(mid - 30.0) < 5.0, mid
        # validation happens before any launch, with the entry point named in the error
        img = focus.image(t, 20, 40)
        poly = np.zeros(3)
        st = lib.asarf_doppler_centroid_noalloc(ctx.h, ctypes.byref(img), prf, 0, poly.ctypes.data, None, None, 0,
                                                None)
        assert st != 0 and b"asarf_doppler_centroid_noalloc" in lib.asarf_last_error()
        st = lib.asarf_doppler_centroid_noalloc(ctx.h, ctypes.byref(img), prf, 0, None, None, None, 0, None)
        assert st != 0
    finally:
        ctx.close()
