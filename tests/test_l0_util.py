"""CPU checks of the level-0 host utilities through the C ABI against the reference's own unit-test vectors
(envisat_format/unit-test/envisat_types_test.cc, envisat_utils_test.cc, parse_util_test.cc, ers_sbt_test.cc,
asar_constants_test.cc) and, for the time arithmetic, against Python's datetime on random values."""
import ctypes
import datetime as dt
import struct

import numpy as np
import pytest

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
