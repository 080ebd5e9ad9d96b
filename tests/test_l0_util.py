"""CPU checks of the level-0 host utilities through the C ABI against the reference's own unit-test vectors
(envisat_format/unit-test/envisat_types_test.cc, envisat_utils_test.cc, parse_util_test.cc, ers_sbt_test.cc,
asar_constants_test.cc), the time arithmetic against Python's datetime on random values, and the ASAR IM packet
parser on records laid out as the reference's parser reads them (envisat_im_parse.cc:433-600; the reference ships
no level-0 files), whose sample records must decode through the oracle's FBAQ-4 decoder."""
import ctypes
import datetime as dt
import struct

import numpy as np
import pytest

import pyoracle as po
from asarf import _lib as L


def mjd(d, s, u):
    return L.Mjd(d, s, u)


def cmp(a, b):
    return L.lib().asarf_mjd_compare(ctypes.byref(a), ctypes.byref(b))


def add_micros(m, us):
    out = L.Mjd()
    L.check(L.lib().asarf_mjd_add_micros(ctypes.byref(m), us, ctypes.byref(out)))
    return out.astuple()


# ------------------------------------------------------------------------------------------ MJD ordering ----
def test_mjd_ordering_reference_cases():
    # envisat_types_test.cc:29-103
    a, b = mjd(10, 10, 10), mjd(10, 10, 11)
    assert cmp(a, b) != 0 and cmp(a, a) == 0
    b.micros = 10
    assert cmp(a, b) == 0
    for lo, hi in (((10, 10, 9), (10, 10, 10)), ((10, 10, 10), (11, 10, 10)), ((10, 0, 10), (10, 10, 10))):
        assert cmp(mjd(*lo), mjd(*hi)) == -1 and cmp(mjd(*hi), mjd(*lo)) == 1


def test_mjd_add_micros_reference_cases():
    # envisat_utils_test.cc:27-43
    assert add_micros(mjd(10, 10, 10), 180) == (10, 10, 190)
    assert add_micros(mjd(1234, 42, 1000), 1000000) == (1234, 43, 1000)


def test_mjd_add_micros_matches_datetime():
    epoch = dt.datetime(2000, 1, 1)
    rng = np.random.default_rng(5)
    for _ in range(300):
        d, s, u = int(rng.integers(-8000, 9000)), int(rng.integers(0, 86400)), int(rng.integers(0, 10 ** 6))
        add = int(rng.integers(-10 ** 12, 10 ** 12))
        t = epoch + dt.timedelta(days=d, seconds=s, microseconds=u) + dt.timedelta(microseconds=add)
        delta = t - epoch  # timedelta normalises to days (floor), seconds of day, microseconds
        assert add_micros(mjd(d, s, u), add) == (delta.days, delta.seconds, delta.microseconds)


def test_mjd_arithmetic_out_of_range_is_rejected():
    # day counts beyond +-1e8 or sums past the 64-bit microsecond range are errors, not wrapped times
    out = L.Mjd()
    for m, us in ((mjd(200_000_000, 0, 0), 0), (mjd(-200_000_000, 0, 0), 1), (mjd(99_000_000, 0, 0), 2 ** 62),
                  (mjd(0, 0, 0), -(2 ** 63) + 5)):
        assert L.lib().asarf_mjd_add_micros(ctypes.byref(m), us, ctypes.byref(out)) != 0
    assert add_micros(mjd(99_000_000, 5, 6), 1) == (99_000_000, 5, 7)


# ----------------------------------------------------------------------------------------- counter gap ----
def test_counter_gap_reference_cases():
    # parse_util_test.cc:27-49 (MaxValueForBits<uint16_t, 14> = 16383)
    gap = L.lib().asarf_counter_gap
    m = (1 << 14) - 1
    assert m == 16383 and (1 << 40) - 1 == 1099511627775
    assert gap(m, m, m) == 0 and gap(15, 15, m) == 0 and gap(0, 0, m) == 0
    assert gap(0, 1, m) == 1 and gap(0, m, m) == m
    assert gap(m, 0, m) == 1 and gap(m, 2, m) == 3


def test_counter_gap_is_modular_distance():
    gap = L.lib().asarf_counter_gap
    rng = np.random.default_rng(2)
    for bits in (14, 16, 32):
        m = (1 << bits) - 1
        for _ in range(200):
            a, b = int(rng.integers(0, m + 1)), int(rng.integers(0, m + 1))
            assert gap(a, b, m) == (b - a) % (m + 1)


# ----------------------------------------------------------------------------------- ERS SBT correction ----
def register(epoch_1950, ms, sbt=0, period=0):
    p = L.Patc(720, b"E1", 929, epoch_1950, ms, sbt, period)
    ref = L.PatcReference()
    L.check(L.lib().asarf_patc_register(ctypes.byref(p), ctypes.byref(ref)))
    return ref


def test_register_patc_rejects_bad_epochs():
    # ers_sbt_test.cc:36-56
    for epoch, ms in ((365 * 50, -1023), (-1, 1000), (14974, 0)):
        with pytest.raises(L.AsarfError):
            register(epoch, ms)
    register(14975, 0)  # 1991-01-01 is accepted


def adjust(ref, start, sbt, overflow=False):
    t = L.Mjd(*start)
    L.check(L.lib().asarf_adjust_isp_sensing_time(ctypes.byref(ref) if ref is not None else None, ctypes.byref(t),
                                                  sbt, int(overflow)))
    return t.astuple()


def test_adjust_isp_sensing_time_reference_cases():
    # ers_sbt_test.cc:58-124 (values from the TDS01 test data set)
    assert adjust(None, (1, 1, 1), 1000) == (1, 1, 1)  # no PATC registered: unchanged
    ref = register(15236, 77966950, 1440813490, 3906249)
    assert adjust(ref, (-3026, 77935, 870958), 1440805534) == (-3026, 77935, 871883)  # counter before the PATC's
    assert adjust(ref, (0, 0, 0), 1440813490) == (-3026, 77966, 950000)  # equal
    assert adjust(ref, (-3026, 77966, 960794), 1440813493) == (-3026, 77966, 961718)  # after
    assert adjust(ref, (-3026, 77935, 870958), 1440805534, overflow=True) == (-2831, 7147, 576915)  # wrapped


def test_parse_patc_record():
    rec = bytearray(b" " * 54)
    rec[15:19] = b"0720"
    rec[20:22] = b"E1"
    rec[30:35] = b"00929"
    rec[38:54] = struct.pack("<iiII", 15236, 77966950, 1440813490, 3906249)
    p = L.Patc()
    L.check(L.lib().asarf_parse_patc(bytes(rec), 54, ctypes.byref(p)))
    assert (p.cyclic_counter, p.mission, p.orbit_number) == (720, b"E1", 929)
    assert (p.epoch_1950, p.milliseconds, p.sbt_counter, p.sbt_period) == (15236, 77966950, 1440813490, 3906249)
    with pytest.raises(L.AsarfError):
        L.check(L.lib().asarf_parse_patc(bytes(rec), 53, ctypes.byref(p)))
    bad = bytearray(rec)
    bad[30:35] = b"  x12"
    with pytest.raises(L.AsarfError):
        L.check(L.lib().asarf_parse_patc(bytes(bad), 54, ctypes.byref(p)))


# ------------------------------------------------------------------------------------------ product types ----
def test_product_types_reference_cases():
    # asar_constants_test.cc:27-64
    f = L.lib().asarf_product_type_from_name
    assert f(b"SAR_IM__0PWDSI19920904_101949_00000018C087_00294_05950_0000.E1") == L.PRODUCT_SAR_IM0
    assert f(b"ASA_IM__0PNPDK20040109_194924_000000182023_00157_09730_1479.N1") == L.PRODUCT_ASA_IM0
    assert f(b"ASA_AM__0PNPDK20040109_194924_000000182023_00157_09730_1479.N1") == L.PRODUCT_UNIDENTIFIED
    out = ctypes.c_int()
    for src, tgt, want in ((L.PRODUCT_SAR_IM0, b"IMS", L.PRODUCT_SAR_IMS), (L.PRODUCT_ASA_IM0, b"IMS", L.PRODUCT_ASA_IMS),
                           (L.PRODUCT_SAR_IM0, b"imp", L.PRODUCT_SAR_IMP), (L.PRODUCT_ASA_IM0, b"Imp", L.PRODUCT_ASA_IMP)):
        L.check(L.lib().asarf_target_product(src, tgt, ctypes.byref(out)))
        assert out.value == want
    with pytest.raises(L.AsarfError):
        L.check(L.lib().asarf_target_product(L.PRODUCT_SAR_IMS, b"IMS", ctypes.byref(out)))
    with pytest.raises(L.AsarfError):
        L.check(L.lib().asarf_target_product(L.PRODUCT_ASA_IM0, b"IMM", ctypes.byref(out)))
    for prod, inst in ((L.PRODUCT_SAR_IM0, L.INSTRUMENT_SAR), (L.PRODUCT_SAR_IMS, L.INSTRUMENT_SAR),
                       (L.PRODUCT_ASA_IM0, L.INSTRUMENT_ASAR), (L.PRODUCT_ASA_IMS, L.INSTRUMENT_ASAR)):
        L.check(L.lib().asarf_instrument(prod, ctypes.byref(out)))
        assert out.value == inst
    for prod in (L.PRODUCT_SAR_IMP, L.PRODUCT_UNIDENTIFIED):
        with pytest.raises(L.AsarfError):
            L.check(L.lib().asarf_instrument(prod, ctypes.byref(out)))


# ------------------------------------------------------------------------------------- ASAR IM packets ----
def asar_packet(rng, blocks, tail, swst=700, mode=0x54, dfh=29, flags=0x80, extra=0, sensing=(-3026, 77935, 870958),
                data=None):
    """A data set record laid out as envisat_im_parse.cc:433-600 reads it (big-endian fields); `data` (block ids and
    code bytes) replaces the random sample blocks."""
    if data is None:
        data_len = 64 * blocks + tail
        data = bytearray(rng.integers(0, 256, data_len, dtype=np.uint8).tobytes())
        for b in range(blocks + (1 if tail else 0)):
            data[64 * b] = int(rng.integers(0, 256))  # block id
    data_len = len(data)
    hdr = struct.pack(">iII", *sensing)                                       # ISP sensing time
    hdr += struct.pack(">iIIHHHH", -3026, 77935, 870000, data_len + 29, 0, 0, 0)  # FEP (ISP length)
    hdr += struct.pack(">HHHH", 0x0A11, (1 << 14) | 513, data_len + 29, dfh)  # packet id, sequence, lengths
    hdr += bytes([0, mode]) + (0x0102030405).to_bytes(5, "big") + b"\x00"      # mode id, on-board time, spare
    hdr += (0x0A0B0C).to_bytes(3, "big") + bytes([(17 << 2) | 2, flags | 0x3, 0x45])  # mode count, beam, flags
    hdr += struct.pack(">HHH", 2112, swst, 5616)                              # PRI, SWST, echo window codes
    hdr += struct.pack(">HH", (9 << 12) | (21 << 7) | 0x40 | 13, (700 << 6) | 33)  # converters, pulse code
    hdr += bytes([7, 200]) + struct.pack(">H", 7777)                          # chirp bw, aux tx, resampling
    assert len(hdr) == 68
    return bytes(hdr) + bytes(data) + bytes(extra), bytes(data)


def parse(pkt, record_bytes=0):
    m = L.AsarEcho()
    rec = np.full(record_bytes, 0xAB, np.uint8) if record_bytes else None
    L.check(L.lib().asarf_parse_asar_im_packet(pkt, len(pkt), ctypes.byref(m),
                                               rec.ctypes.data if rec is not None else None, record_bytes))
    return m, rec


def test_asar_packet_fields_and_record():
    rng = np.random.default_rng(4)
    pkt, data = asar_packet(rng, 3, 10, extra=5)
    m, rec = parse(pkt, record_bytes=300)
    assert m.isp_sensing_time.astuple() == (-3026, 77935, 870958) and m.fep_time.astuple() == (-3026, 77935, 870000)
    assert (m.isp_length, m.packet_id, m.segmentation_flags, m.sequence_count) == (len(data) + 29, 0x0A11, 1, 513)
    assert m.dsr_bytes == 68 + len(data)  # extra bytes after the blocks belong to the next record
    assert m.packet_length == len(data) + 29 and m.mode_id == 0x54 and m.onboard_time == 0x0102030405
    assert (m.mode_count, m.antenna_beam_set_no, m.comp_ratio) == (0x0A0B0C, 17, 2)
    assert (m.echo_flag, m.noise_flag, m.cal_flag, m.cal_type, m.cycle_count) == (1, 0, 0, 0, 0x345)
    assert (m.pri_code, m.swst_code, m.echo_window_code) == (2112, 700, 5616)
    assert (m.upconverter_raw, m.downconverter_raw, m.tx_pol, m.rx_pol, m.cal_row_number) == (9, 21, 1, 0, 13)
    assert (m.tx_pulse_code, m.beam_adj_delta, m.chirp_pulse_bw_code, m.aux_tx_monitor_level) == (700, 33, 7, 200)
    assert m.resampling_factor == 7777
    assert m.sample_bytes == 202 and m.samples == 3 * 63 + 9
    assert rec[0] | (rec[1] << 8) == 202 and bytes(rec[2:204]) == data and not rec[204:].any()


def test_asar_packet_records_feed_the_decoder():
    # parsed records are the decoder's input: sample s of line y comes from code byte 64 (s / 63) + 1 + s % 63
    rng = np.random.default_rng(9)
    i_lut, q_lut = rng.standard_normal(4096).astype(np.float32), rng.standard_normal(4096).astype(np.float32)
    swsts = [700, 722, 700, 744]
    recs, datas = np.zeros((4, 2 + 64 * 6), np.uint8), []
    for y, sw in enumerate(swsts):
        pkt, data = asar_packet(rng, 5 + (y % 2), 0 if y % 2 else 20, swst=sw)
        m, rec = parse(pkt, record_bytes=recs.shape[1])
        assert m.swst_code == sw
        recs[y], _ = rec, datas.append(data)
    img = po.ingest_asar_fbaq4(recs, np.array(swsts, np.uint16), 700, i_lut, q_lut, 500, 512, 8)
    for y, (sw, data) in enumerate(zip(swsts, datas)):
        for s in (0, 1, 62, 63, 64, 125, 200, 300):
            k = 64 * (s // 63) + 1 + s % 63
            x = s + sw - 700
            if k >= len(data) or x >= 500:
                continue
            bid, code = data[64 * (s // 63)], data[k]
            want = (i_lut[po.fbaq4_index(bid, code >> 4)], q_lut[po.fbaq4_index(bid, code & 15)])
            assert (img[y, x].real, img[y, x].imag) == want, (y, s)


def test_asar_packet_errors_and_non_echo():
    rng = np.random.default_rng(5)
    for kw in (dict(dfh=28), dict(mode=0x55)):
        pkt, _ = asar_packet(rng, 2, 0, **kw)
        with pytest.raises(L.AsarfError):
            parse(pkt)
    pkt, _ = asar_packet(rng, 2, 0)
    m = L.AsarEcho()
    assert L.lib().asarf_parse_asar_im_packet(pkt[:-1], len(pkt) - 1, ctypes.byref(m), None, 0) == L.ESHAPE
    assert L.lib().asarf_parse_asar_im_packet(pkt[:60], 60, ctypes.byref(m), None, 0) == L.ESHAPE
    small = np.zeros(100, np.uint8)
    assert L.lib().asarf_parse_asar_im_packet(pkt, len(pkt), ctypes.byref(m), small.ctypes.data, 100) == L.EWORKSPACE
    # a calibration packet: metadata parsed, no record written
    pkt, _ = asar_packet(rng, 2, 0, flags=0x20)
    m, rec = parse(pkt, record_bytes=200)
    assert (m.echo_flag, m.cal_flag) == (0, 1) and (rec == 0xAB).all()


# -------------------------------------------------------------------------------------- ERS IM packets ----
def ers_packet(rng, sensing, sbt, swst=700, isp_length=11465, aux=0xAA, dr_no=987654):
    """An 11498-byte data set record laid out as ers_im_parse.cc:366-548 reads it (big-endian fields)."""
    samples = rng.integers(0, 32, 11232, dtype=np.uint8).tobytes()
    p = bytearray(11498)
    p[0:12] = struct.pack(">iII", *sensing)
    p[24:26] = struct.pack(">H", isp_length)
    p[32:36] = struct.pack(">I", dr_no)
    p[36], p[37], p[46], p[47] = 17, 23, aux, 0x21
    p[48:52] = struct.pack(">I", sbt)
    p[52], p[53] = 0xAA, 0x11
    p[54:62] = struct.pack(">IHH", 424242, swst, 2800)
    p[266:] = samples
    return bytes(p), samples


def parse_ers(pkt):
    m = L.ErsEcho()
    line = np.zeros(11232, np.uint8)
    L.check(L.lib().asarf_parse_ers_im_packet(pkt, len(pkt), ctypes.byref(m), line.ctypes.data))
    return m, line


def test_ers_packet_fields_line_and_sbt_time():
    rng = np.random.default_rng(3)
    # the TDS01 values of ers_sbt_test.cc:58-124 carried through packets: the recorded ISP time and SBT of each
    # packet, corrected with the registered PATC, give the reference test's expected times
    ref = register(15236, 77966950, 1440813490, 3906249)
    last_sbt = None
    for sensing, sbt, want in (((-3026, 77935, 870958), 1440805534, (-3026, 77935, 871883)),
                               ((-3026, 77966, 960794), 1440813493, (-3026, 77966, 961718))):
        pkt, samples = ers_packet(rng, sensing, sbt)
        m, line = parse_ers(pkt)
        assert m.isp_sensing_time.astuple() == sensing and m.sbt == sbt
        assert (m.isp_length, m.data_record_number, m.packet_counter, m.subcommutation_counter) == (11465, 987654, 17, 23)
        assert (m.obrc_orbit, m.activity_task, m.sample_flags, m.image_format_counter) == (0x21, 0xAA, 0x11, 424242)
        assert (m.swst_code, m.pri_code) == (700, 2800) and bytes(line) == samples
        t = m.isp_sensing_time
        overflow = last_sbt is not None and sbt < last_sbt
        L.check(L.lib().asarf_adjust_isp_sensing_time(ctypes.byref(ref), ctypes.byref(t), sbt, int(overflow)))
        assert t.astuple() == want
        last_sbt = sbt


def test_ers_packet_lines_feed_the_decoder_and_errors():
    rng = np.random.default_rng(8)
    swsts = [700, 704, 700, 712]
    lines = np.zeros((4, 11232), np.uint8)
    for y, sw in enumerate(swsts):
        pkt, _ = ers_packet(rng, (0, 0, 0), 1, swst=sw)
        m, lines[y] = parse_ers(pkt)
        assert m.swst_code == sw
    img = po.ingest_ers_im(lines, 5616, np.array(swsts, np.uint16), 700, 5616 + 48, 5680, 8)
    for y, sw in enumerate(swsts):
        for s in (0, 1, 100, 5615):
            x = s + 4 * (sw - 700)
            assert (img[y, x].real, img[y, x].imag) == (lines[y, 2 * s], lines[y, 2 * s + 1])
    m = L.ErsEcho()
    for kw in (dict(isp_length=11464), dict(aux=0xAB)):
        pkt, _ = ers_packet(rng, (0, 0, 0), 1, **kw)
        assert L.lib().asarf_parse_ers_im_packet(pkt, len(pkt), ctypes.byref(m), None) == L.EINVAL
    pkt, _ = ers_packet(rng, (0, 0, 0), 1)
    assert L.lib().asarf_parse_ers_im_packet(pkt, 11497, ctypes.byref(m), None) == L.ESHAPE


# ------------------------------------------------------------------------------- product header and MDSR ----
def product_file(rng, n_packets, dsr_size=b"-0000000001", num_dsr_field=None):
    mph = (b'PRODUCT="ASA_IM__0PNPDK20040109_194924_000000182023_00157_09730_1479.N1"\n'
           b'SENSING_START="09-JAN-2004 19:49:24.000000"\nREL_ORBIT=+00157\nPHASE=2\n')
    mph = mph.ljust(1246, b" ") + b"\n"
    sph = b'SPH_DESCRIPTOR="ASAR Image Mode Level 0"\nSWATH="IS2"\nTX_RX_POLAR="V/V"\n'.ljust(1956 - 4 * 280, b" ")
    packets, datas = [], []
    for i in range(n_packets):
        pkt, data = asar_packet(rng, 3 + i % 3, 5 * i, swst=700 + 11 * i)
        packets.append(pkt)
        datas.append(data)
    body = b"".join(packets)
    offset = 1247 + 1956
    dsd = (b'DS_NAME="ASAR SOURCE PACKETS         "\nDS_TYPE=M\nFILENAME="' + b" " * 62 + b'"\n'
           + b"DS_OFFSET=+%020d<bytes>\nDS_SIZE=+%020d<bytes>\n" % (offset, len(body))
           + (b"NUM_DSR=+%010d\n" % n_packets if num_dsr_field is None else b"NUM_DSR=" + num_dsr_field + b"\n")
           + b"DSR_SIZE=" + dsr_size + b"<bytes>\n")
    dsd = dsd.ljust(279, b" ") + b"\n"
    spare = (b"DS_NAME=\n" * 7).ljust(279, b" ") + b"\n"
    sph += dsd + spare * 3
    assert len(mph) == 1247 and len(sph) == 1956
    return mph + sph + body, datas


def test_header_lookup_and_mdsr_descriptor():
    rng = np.random.default_rng(12)
    f, _ = product_file(rng, 4)
    v = ctypes.create_string_buffer(128)
    L.check(L.lib().asarf_header_get(f, 1247, b"PRODUCT", v, 128))
    assert v.value == b"ASA_IM__0PNPDK20040109_194924_000000182023_00157_09730_1479.N1"
    assert L.lib().asarf_product_type_from_name(v.value) == L.PRODUCT_ASA_IM0
    L.check(L.lib().asarf_header_get(f, 1247, b"REL_ORBIT", v, 128))
    assert v.value == b"+00157"
    L.check(L.lib().asarf_header_get(f[1247:], 1956, b"SWATH", v, 128))
    assert v.value == b"IS2"
    with pytest.raises(L.AsarfError):
        L.check(L.lib().asarf_header_get(f, 1247, b"ABS_ORBIT", v, 128))
    d = L.Dsd()
    L.check(L.lib().asarf_level0_mdsr(f, len(f), ctypes.byref(d)))
    assert d.ds_name == b'DS_NAME="ASAR SOURCE PACKETS         "' and d.ds_type == b"DS_TYPE=M"
    assert (d.ds_offset, d.num_dsr) == (1247 + 1956, 4) and d.ds_size == len(f) - 3203
    assert d.dsr_size == 1  # "-0000000001" (variable-size records) reads as 1, as the reference reads it
    f2, _ = product_file(rng, 1, dsr_size=b"+0000011498")
    L.check(L.lib().asarf_level0_mdsr(f2, len(f2), ctypes.byref(d)))
    assert d.dsr_size == 11498
    assert L.lib().asarf_level0_mdsr(f, 3000, ctypes.byref(d)) == L.ESHAPE


def test_mdsr_numbers_that_do_not_fit_64_bits_are_errors():
    # the reference reads the counts with std::stoul, which throws past 64 bits; a wrapped count would let the
    # record walk run past the data set
    rng = np.random.default_rng(18)
    d = L.Dsd()
    f, _ = product_file(rng, 1, num_dsr_field=b"+18446744073709551615")  # 2^64 - 1 still fits
    L.check(L.lib().asarf_level0_mdsr(f, len(f), ctypes.byref(d)))
    assert d.num_dsr == 2 ** 64 - 1
    for field in (b"+18446744073709551616", b"+99999999999999999999999"):
        f, _ = product_file(rng, 1, num_dsr_field=field)
        assert L.lib().asarf_level0_mdsr(f, len(f), ctypes.byref(d)) == L.EINVAL
    f, _ = product_file(rng, 1, dsr_size=b"+000000000000000000000000011498")  # leading zeros do not count
    L.check(L.lib().asarf_level0_mdsr(f, len(f), ctypes.byref(d)))
    assert d.dsr_size == 11498


def test_level0_file_walk_to_decoder_records():
    # MPH -> MDSR -> packets stepped by 39 + ISP length -> decoder records
    rng = np.random.default_rng(13)
    f, datas = product_file(rng, 6)
    d = L.Dsd()
    L.check(L.lib().asarf_level0_mdsr(f, len(f), ctypes.byref(d)))
    off, swsts = d.ds_offset, []
    recs = np.zeros((d.num_dsr, 2 + 64 * 6), np.uint8)
    for i in range(d.num_dsr):
        m = L.AsarEcho()
        L.check(L.lib().asarf_parse_asar_im_packet(f[off:], len(f) - off, ctypes.byref(m), recs[i].ctypes.data,
                                                   recs.shape[1]))
        assert bytes(recs[i, 2:2 + m.sample_bytes]) == datas[i]
        swsts.append(m.swst_code)
        off += m.dsr_bytes
    assert off == len(f) and swsts == [700 + 11 * i for i in range(6)]


# ------------------------------------------------------------------------------------ ASAR packet forecast ----
def forecast(mds, n, start, stop, before, after, cap=64):
    b, a, cnt = ctypes.c_long(before), ctypes.c_long(after), ctypes.c_long()
    offs = (ctypes.c_long * cap)()
    times = (L.Mjd * cap)()
    L.check(L.lib().asarf_asar_forecast(mds, len(mds), n, ctypes.byref(L.Mjd(*start)), ctypes.byref(L.Mjd(*stop)),
                                        ctypes.byref(b), ctypes.byref(a), offs, times, cap, ctypes.byref(cnt)))
    return list(offs[:cnt.value]), [t.astuple()[2] for t in times[:cnt.value]], b.value, a.value


NOISE, CAL, ECHO = 0x40, 0x20, 0x80


def build_mds(rng, flags):
    pkts, offs, pos = [], [], 0
    for i, f in enumerate(flags):
        pkt, _ = asar_packet(rng, 1, 0, flags=f, sensing=(100, 5, 1000 * i))
        pkts.append(pkt)
        offs.append(pos)
        pos += len(pkt)
    return b"".join(pkts), offs


def test_asar_forecast_window_and_trimming():
    # envisat_im_parse.cc:297-384 on ten records sensed 1 ms apart: leading noise and calibration records are
    # skipped until the first echo, a noise record later on is skipped, a calibration record is kept; the window
    # is [4 ms, 7 ms]; one line may be kept before the start and one after the stop
    rng = np.random.default_rng(14)
    flags = [NOISE, CAL, ECHO, ECHO, ECHO, CAL, NOISE, ECHO, ECHO, ECHO]
    mds, offs = build_mds(rng, flags)
    o, t, before, after = forecast(mds, 10, (100, 5, 4000), (100, 5, 7000), 1, 1)
    # kept before trimming: 2, 3 (before the start), 4, 5 (cal), 7, 8 (after the stop); 9 ends the walk
    assert t == [3000, 4000, 5000, 7000, 8000] and o == [offs[i] for i in (3, 4, 5, 7, 8)]
    assert (before, after) == (1, 1)
    # no lines allowed outside the window: both edges trimmed
    o, t, before, after = forecast(mds, 10, (100, 5, 4000), (100, 5, 7000), 0, 0)
    assert t == [4000, 5000, 7000] and (before, after) == (0, 0)
    # generous allowances keep every echo and calibration record from the first echo on
    o, t, before, after = forecast(mds, 10, (100, 5, 4000), (100, 5, 7000), 10, 10)
    assert t == [2000, 3000, 4000, 5000, 7000, 8000, 9000] and (before, after) == (2, 2)


def test_asar_forecast_drops_leading_calibration_after_trim():
    rng = np.random.default_rng(15)
    flags = [ECHO, CAL, ECHO, ECHO, ECHO]
    mds, offs = build_mds(rng, flags)
    # before the start: records 0 (echo) and 1 (cal); one allowed -> record 0 goes, then the leading calibration
    # record 1 goes too and the kept count before the start drops to 0
    o, t, before, after = forecast(mds, 5, (100, 5, 2000), (100, 5, 9000), 1, 0)
    assert t == [2000, 3000, 4000] and o == offs[2:] and (before, after) == (0, 0)
    with pytest.raises(L.AsarfError):
        forecast(mds, 5, (100, 5, 2000), (100, 5, 9000), 9, 9, cap=2)
    with pytest.raises(L.AsarfError):
        forecast(mds[:-10], 6, (100, 5, 2000), (100, 5, 9000), 9, 9)


def test_asar_forecast_calibration_after_the_start_keeps_the_count():
    # no allowance before the start: both echoes before it go; the calibration records that now lead the list were
    # sensed at and after the start, so dropping them leaves the count before the start at 0 (the reference's
    # unsigned count would wrap below zero here)
    rng = np.random.default_rng(19)
    mds, offs = build_mds(rng, [ECHO, ECHO, CAL, CAL, ECHO, ECHO])
    o, t, before, after = forecast(mds, 6, (100, 5, 2000), (100, 5, 9000), 0, 0)
    assert t == [4000, 5000] and o == offs[4:] and (before, after) == (0, 0)
    # with one line allowed before the start, the echo at 1 ms is kept and nothing is dropped after it
    o, t, before, after = forecast(mds, 6, (100, 5, 2000), (100, 5, 9000), 1, 0)
    assert t == [1000, 2000, 3000, 4000, 5000] and (before, after) == (1, 0)


# ------------------------------------------------------------------------------------- ERS packet forecast ----
def ers_forecast(mds, n, start, stop, before, after, patc=None, cap=64):
    b, a, cnt = ctypes.c_long(before), ctypes.c_long(after), ctypes.c_long()
    offs = (ctypes.c_long * cap)()
    times = (L.Mjd * cap)()
    L.check(L.lib().asarf_ers_forecast(mds, len(mds), n, ctypes.byref(L.Mjd(*start)), ctypes.byref(L.Mjd(*stop)),
                                       ctypes.byref(patc) if patc is not None else None, ctypes.byref(b),
                                       ctypes.byref(a), offs, times, cap, ctypes.byref(cnt)))
    return [o // 11498 for o in offs[:cnt.value]], [t.astuple() for t in times[:cnt.value]], b.value, a.value


def test_ers_forecast_fills_missing_records_and_trims():
    # ers_im_parse.cc:228-331 on six records 1 ms apart whose record numbers jump from 12 to 15: the two missing
    # records repeat record 2 (counted by its time); record 3, which follows the gap, is not counted against the
    # window (the reference counts only on the no-gap branch); window [2 ms, 4 ms]
    rng = np.random.default_rng(16)
    drs = [10, 11, 12, 15, 16, 17]
    mds = b"".join(ers_packet(rng, (100, 5, 1000 * i), 7, dr_no=d)[0] for i, d in enumerate(drs))
    idx, t, before, after = ers_forecast(mds, 6, (100, 5, 2000), (100, 5, 4000), 10, 10)
    assert idx == [0, 1, 2, 2, 2, 3, 4, 5] and [x[2] for x in t] == [0, 1000, 2000, 2000, 2000, 3000, 4000, 5000]
    assert (before, after) == (2, 1)
    idx, t, before, after = ers_forecast(mds, 6, (100, 5, 2000), (100, 5, 4000), 0, 0)
    assert idx == [2, 2, 2, 3, 4] and (before, after) == (0, 0)
    # more than 900 missing records is an error
    mds2 = b"".join(ers_packet(rng, (100, 5, 1000 * i), 7, dr_no=d)[0] for i, d in enumerate([1, 2, 905]))
    with pytest.raises(L.AsarfError):
        ers_forecast(mds2, 3, (100, 5, 0), (100, 5, 9000), 9, 9)
    # a record count whose byte size overflows 64-bit arithmetic is rejected before the walk
    for n in (2 ** 63 // 11498 + 1, 2 ** 62, 7):
        with pytest.raises(L.AsarfError):
            ers_forecast(mds, n, (100, 5, 0), (100, 5, 9000), 9, 9)


def test_ers_forecast_sbt_correction():
    # the TDS01 PATC of ers_sbt_test.cc: the recorded times are replaced by the SBT-derived ones
    rng = np.random.default_rng(17)
    ref = register(15236, 77966950, 1440813490, 3906249)
    mds = b"".join(ers_packet(rng, (1, 2, 3), sbt, dr_no=d)[0] for d, sbt in ((5, 1440813490), (6, 1440813493)))
    idx, t, _, _ = ers_forecast(mds, 2, (-4000, 0, 0), (0, 0, 0), 9, 9, patc=ref)
    assert idx == [0, 1] and t == [(-3026, 77966, 950000), (-3026, 77966, 961718)]


# ------------------------------------------------------------------------------------- instrument file ----
def test_ins_fbaq4_tables():
    # offsets from the reference's layout (envisat_aux_file.h): MPH 1247 + SPH 378, FixedLifetimeParameters 72608,
    # then in FixedBaselineParameters i_LUT_8bit[256], q_LUT_8bit[256], i_LUT_fbaq4[4096], i_LUT_fbaq3[2048],
    # i_LUT_fbaq2[1024], q_LUT_fbaq4[4096]; big-endian floats
    rng = np.random.default_rng(18)
    buf = bytearray(rng.integers(0, 256, 173273, dtype=np.uint8).tobytes())
    i_want = rng.standard_normal(4096).astype(np.float32)
    q_want = rng.standard_normal(4096).astype(np.float32)
    fbp = 1247 + 378 + 72608
    i_at = fbp + 4 * 512
    q_at = i_at + 4 * (4096 + 2048 + 1024)
    buf[i_at:i_at + 16384] = i_want.astype(">f4").tobytes()
    buf[q_at:q_at + 16384] = q_want.astype(">f4").tobytes()
    assert q_at + 16384 <= 1247 + 378 + 72608 + 89408  # inside FixedBaselineParameters
    i_lut, q_lut = np.zeros(4096, np.float32), np.zeros(4096, np.float32)
    L.check(L.lib().asarf_ins_fbaq4_tables(bytes(buf), len(buf), i_lut.ctypes.data, q_lut.ctypes.data))
    assert np.array_equal(i_lut, i_want) and np.array_equal(q_lut, q_want)
    assert L.lib().asarf_ins_fbaq4_tables(bytes(buf[:-1]), len(buf) - 1, i_lut.ctypes.data,
                                          q_lut.ctypes.data) == L.ESHAPE
