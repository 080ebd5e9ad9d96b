"""GPU parity of the level-0 sample decoders (ERS IM u8 I/Q, ASAR IM FBAQ-4) against the oracle's restatement of
the reference's scatter kernels: bit-identical values, NaN where no sample lands, zero padding; and the decoded
ERS image through the I/Q correction stage."""
import numpy as np
import pytest
import torch

import pyoracle as po
from asarf import focus

pytestmark = pytest.mark.gpu


def same(a, b):
    """Bitwise-equal complex64 arrays with NaN == NaN."""
    a, b = np.asarray(a, np.complex64), np.asarray(b, np.complex64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def ers_lines(rng, lines, spl, line_bytes):
    samples = np.zeros((lines, line_bytes), np.uint8)
    samples[:, : 2 * spl] = rng.integers(0, 32, (lines, 2 * spl), dtype=np.uint8)
    # ERS SWST codes move in steps of 22 and stay within 500..1500 (main_flow.cc:551-559)
    swst = (700 + 22 * rng.integers(0, 4, lines)).astype(np.uint16)
    return samples, swst


# line pitch 2 spl + pad: a multiple of 4 takes the 4-byte sample-pair loads (odd spl: the last pair is split)
@pytest.mark.parametrize("lines,spl,range_size,x_stride,y_stride,pad",
                         [(37, 300, 340, 352, 48, 6), (37, 300, 340, 352, 48, 8), (37, 301, 341, 352, 48, 6),
                          (200, 5616, 5616 + 4 * 66, 6400, 210, 6), (200, 5616, 5616 + 4 * 66, 6400, 210, 8)])
def test_ers_decode_matches_oracle(ctx, lines, spl, range_size, x_stride, y_stride, pad):
    rng = np.random.default_rng(lines + pad)
    samples, swst = ers_lines(rng, lines, spl, line_bytes=2 * spl + pad)
    min_swst = int(swst.min())
    out = torch.full((y_stride, x_stride), 7 + 7j, dtype=torch.complex64, device="cuda")
    focus.ingest_ers_im(ctx, torch.from_numpy(samples).cuda(), spl, torch.from_numpy(swst.view(np.int16)).cuda(),
                        min_swst, out, range_size)
    ref = po.ingest_ers_im(samples, spl, swst, min_swst, range_size, x_stride, y_stride)
    got = out.cpu().numpy()
    print(f"ers {lines}x{spl}: NaN {np.isnan(ref).sum()} of {ref.size}")
    assert same(got, ref)


def fbaq_records(rng, n, max_len, record_bytes, short_every=5):
    recs = rng.integers(0, 256, (n, record_bytes), dtype=np.uint8)
    lengths = np.full(n, max_len)
    lengths[::short_every] = rng.integers(1, max_len, lengths[::short_every].shape)  # short records: NaN tail
    recs[:, 0] = lengths & 0xFF
    recs[:, 1] = lengths >> 8
    return recs


@pytest.mark.parametrize("n,blocks,range_size,x_stride,y_stride,lut_items",
                         [(23, 3, 200, 208, 32, 4096), (120, 90, 90 * 63 + 40, 6144, 128, 4096),
                          (40, 5, 300, 301, 41, 3000)])  # odd pitch; short tables: codes past them stay NaN
def test_asar_fbaq4_decode_matches_oracle(ctx, n, blocks, range_size, x_stride, y_stride, lut_items):
    rng = np.random.default_rng(n)
    max_len = 64 * blocks
    recs = fbaq_records(rng, n, max_len, record_bytes=max_len + 2 + 5)
    swst = rng.integers(900, 940, n).astype(np.uint16)
    min_swst = int(swst.min())
    i_lut = rng.standard_normal(lut_items).astype(np.float32)
    q_lut = rng.standard_normal(lut_items).astype(np.float32)
    out = torch.full((y_stride, x_stride), 7 + 7j, dtype=torch.complex64, device="cuda")
    focus.ingest_asar_fbaq4(ctx, torch.from_numpy(recs).cuda(), torch.from_numpy(swst.view(np.int16)).cuda(),
                            min_swst, torch.from_numpy(i_lut).cuda(), torch.from_numpy(q_lut).cuda(), out,
                            range_size)
    ref = po.ingest_asar_fbaq4(recs, swst, min_swst, i_lut, q_lut, range_size, x_stride, y_stride)
    got = out.cpu().numpy()
    print(f"fbaq4 {n} records x {blocks} blocks: NaN {np.isnan(ref).sum()} of {ref.size}")
    assert same(got, ref)


def test_fbaq4_index_from_library_matches_oracle():
    for block in (0, 5, 255):
        for code in range(16):
            assert focus.fbaq4_index(block, code) == po.fbaq4_index(block, code)


def test_decoded_ers_image_through_iq_correction(ctx):
    """Decode (NaN where the SWST shift leaves no sample) -> RawDataCorrection + ZeroNaNs, against the oracle on the
    oracle's decode of the same bytes."""
    rng = np.random.default_rng(9)
    lines, spl, range_size, x_stride, y_stride = 64, 500, 500 + 4 * 66, 800, 72
    samples, swst = ers_lines(rng, lines, spl, line_bytes=2 * spl)
    min_swst = int(swst.min())
    out = torch.zeros((y_stride, x_stride), dtype=torch.complex64, device="cuda")
    focus.ingest_ers_im(ctx, torch.from_numpy(samples).cuda(), spl, torch.from_numpy(swst.view(np.int16)).cuda(),
                        min_swst, out, range_size)
    ref = po.ingest_ers_im(samples, spl, swst, min_swst, range_size, x_stride, y_stride).astype(np.complex128)
    n_total = lines * spl
    gres = focus.iq_correction(ctx, out, range_size, lines, n_total=n_total)
    rres = po.iq_correction(ref, range_size, lines, n_total=n_total)  # in place on ref, NaN -> 0
    for k in rres:
        assert abs(gres[k] - rres[k]) <= 1e-9 * max(1.0, abs(rres[k])), k
    got = out.cpu().numpy()
    assert not np.isnan(got).any()
    err = np.sqrt(np.sum(np.abs(got - ref) ** 2) / np.sum(np.abs(ref) ** 2))
    print(f"decoded ERS through I/Q correction: rel rms {err:.2e}, gain {gres['iq_gain']:.6f}")
    assert err < 1e-6


def test_parsed_asar_packets_decode_on_the_device(ctx):
    """ASAR IM data set records through the host packet parser into the device FBAQ-4 decoder, bit-identical to
    the oracle's decoder on the same parsed records."""
    import ctypes
    import sys
    import os
    sys.path.insert(0, os.path.dirname(__file__))
    from asarf import _lib as L
    from test_l0_util import asar_packet
    rng = np.random.default_rng(21)
    n, rec_bytes = 40, 2 + 64 * 90
    swst = np.array([700 + 11 * (i % 4) for i in range(n)], np.uint16)
    recs = np.zeros((n, rec_bytes), np.uint8)
    for y in range(n):
        pkt, _ = asar_packet(rng, 80 + y % 10, (7 * y) % 64, swst=int(swst[y]))
        m = L.AsarEcho()
        L.check(L.lib().asarf_parse_asar_im_packet(pkt, len(pkt), ctypes.byref(m), recs[y].ctypes.data, rec_bytes))
        assert m.swst_code == swst[y]
    i_lut = rng.standard_normal(4096).astype(np.float32)
    q_lut = rng.standard_normal(4096).astype(np.float32)
    range_size, x_stride, y_stride = 5616 + 33, 5664, 48
    out = torch.full((y_stride, x_stride), 3 - 3j, dtype=torch.complex64, device="cuda")
    focus.ingest_asar_fbaq4(ctx, torch.from_numpy(recs).cuda(), torch.from_numpy(swst.view(np.int16)).cuda(), 700,
                            torch.from_numpy(i_lut).cuda(), torch.from_numpy(q_lut).cuda(), out, range_size)
    ref = po.ingest_asar_fbaq4(recs, swst, 700, i_lut, q_lut, range_size, x_stride, y_stride)
    assert same(out.cpu().numpy(), ref)
