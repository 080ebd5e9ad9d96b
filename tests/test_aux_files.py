"""Auxiliary inputs of a real product (csrc/aux_files.cpp): DORIS orbit files (doris_orbit.cc:82-304), the CON and
XCA calibration files (envisat_aux_file.cc:41-131), the instrument file's swath parameters and the radar timing of
the level-0 flow (main_flow.cc:389-460). The reference ships no such files and no tests for these readers, so the
files are built here to the layouts the reference reads (sizes, offsets, text records) and every value is checked
against what was written or against the closed form of the timing."""
import ctypes
import datetime as dt
import math
import struct

import numpy as np
import pytest

from asarf import _lib as L

C = 299792458.0
MONTHS = ["JAN", "FEB", "MAR", "APR", "MAY", "JUN", "JUL", "AUG", "SEP", "OCT", "NOV", "DEC"]
EPOCH = dt.datetime(2000, 1, 1)


Mjd, Osv, DorisMeta, InsSwath, Timing = L.Mjd, L.Osv, L.DorisMeta, L.InsSwath, L.Timing


def lib():
    return L.lib()


def mjd(t: dt.datetime) -> Mjd:
    d = t - EPOCH
    return Mjd(d.days, d.seconds, d.microseconds)


def header(pairs, size):
    text = "".join(f'{k}="{v}"\n' if isinstance(v, str) else f"{k}={v}\n" for k, v in pairs)
    assert len(text) <= size
    return (text + " " * (size - len(text) - 1) + "\n").encode()


def orbit_state(k):
    """Record k of the synthetic orbit: position and velocity printed to fixed decimals (what the parser reads)."""
    a = 2 * math.pi * k / 100.8
    pos = (7.16e6 * math.cos(a), 7.16e6 * math.sin(a), 1.0e5 * math.sin(3 * a))
    vel = (-7450.0 * math.sin(a), 7450.0 * math.cos(a), 30.0 * math.cos(3 * a))
    return [float(f"{v:.3f}") for v in pos], [float(f"{v:.6f}") for v in vel]


def doris_file(t0, n_records=1589, ds_name="DORIS PRECISE ORBIT STATE VECTORS", bad_record=None, bad_text=None):
    """A DOR_VOR_AX file: 1247-byte MPH, 378-byte SPH, n_records text records of 128 characters + newline, one
    minute apart from t0."""
    mph = header([("PRODUCT", "DOR_VOR_AXVF-P20120424_120200_20040101_215528_20040103_002328"), ("PROC_STAGE", "V"),
                  ("PHASE", "3"), ("CYCLE", "+024"), ("REL_ORBIT", "+00123"), ("ABS_ORBIT", "+09656")], 1247)
    sph = header([("SPH_DESCRIPTOR", "DORIS PRECISE ORBIT"), ("DS_NAME", ds_name), ("DS_TYPE", "M")], 378)
    recs = []
    for k in range(n_records):
        t = t0 + dt.timedelta(minutes=k)
        pos, vel = orbit_state(k)
        items = [f"{t.day:02d}-{MONTHS[t.month - 1]}-{t.year:04d}", t.strftime("%H:%M:%S") + f".{t.microsecond:06d}",
                 f"{-0.398765 + 1e-6 * k:.6f}", f"{9656 + k // 101:05d}", *[f"{v:.3f}" for v in pos],
                 *[f"{v:.6f}" for v in vel], "0"]
        line = " ".join(items)
        line = line.replace(" ", " " * (1 + 128 - len(line)), 1)  # pad inside, after the date: 128 characters
        if k == bad_record:
            line = bad_text
        assert len(line) == 128
        recs.append(line + "\n")
    data = mph + sph + "".join(recs).encode()
    return data


def parse_doris(data, start, stop, capacity=64):
    l = lib()
    osv = (Osv * max(capacity, 1))()
    times = (Mjd * max(capacity, 1))()
    n = ctypes.c_int(-1)
    meta = DorisMeta()
    st = l.asarf_doris_orbit(data, len(data), ctypes.byref(mjd(start)), ctypes.byref(mjd(stop)), osv, times, capacity,
                             ctypes.byref(n), ctypes.byref(meta))
    return st, osv[: max(n.value, 0)], times[: max(n.value, 0)], meta


def test_doris_orbit_window_vectors_and_metadata():
    t0 = dt.datetime(2004, 1, 1, 21, 56, 0)
    data = doris_file(t0)
    assert len(data) == 206606
    start = t0 + dt.timedelta(minutes=600)
    stop = start + dt.timedelta(seconds=10)
    st, osv, times, meta = parse_doris(data, start, stop)
    assert st == L.OK, L.last_error()
    # records from start - 5 min to stop + 5 min: minutes 595 .. 605
    assert len(osv) == 11
    for i, (v, tm) in enumerate(zip(osv, times)):
        k = 595 + i
        t = t0 + dt.timedelta(minutes=k)
        assert (tm.days, tm.seconds, tm.micros) == (mjd(t).days, mjd(t).seconds, 0)
        assert v.t == pytest.approx((t - EPOCH).total_seconds(), abs=1e-6)
        pos, vel = orbit_state(k)
        assert list(v.pos) == pos and list(v.vel) == vel
    # metadata: the first kept vector later than start - 1 min is the one at start
    assert meta.orbit_name.decode().startswith("DOR_VOR_AXVF-P2012")
    assert meta.vector_source == b"DP" and meta.phase == b"3" and meta.cycle == 24
    assert meta.abs_orbit == 9656 + 600 // 101 and meta.delta_ut1 == pytest.approx(-0.398765 + 600e-6, abs=1e-12)
    assert list(meta.pos) == orbit_state(600)[0]
    assert (meta.state_vector_time.days, meta.state_vector_time.seconds) == (mjd(start).days, mjd(start).seconds)


@pytest.mark.parametrize("name,source", [("DORIS PRECISE ORBIT STATE VECTORS", b"DP"), ("REAPER ORBIT", b"RO"),
                                         ("DORIS PRELIMINARY", b"??")])
def test_doris_vector_source(name, source):
    t0 = dt.datetime(1995, 5, 3, 21, 0, 0)
    data = doris_file(t0, n_records=1801, ds_name=name)
    assert len(data) == 233954  # the ERS size
    st, osv, _, meta = parse_doris(data, t0 + dt.timedelta(minutes=30), t0 + dt.timedelta(minutes=31))
    assert st == L.OK, L.last_error()
    assert meta.vector_source == source and len(osv) == 12
    assert osv[0].t == pytest.approx((t0 + dt.timedelta(minutes=25) - EPOCH).total_seconds())


def test_doris_errors():
    t0 = dt.datetime(2004, 1, 1, 21, 56, 0)
    data = doris_file(t0)
    start = t0 + dt.timedelta(minutes=600)
    l = lib()
    n = ctypes.c_int()
    osv = (Osv * 64)()
    assert l.asarf_doris_orbit(data[:-1], len(data) - 1, ctypes.byref(mjd(start)), ctypes.byref(mjd(start)), osv,
                               None, 64, ctypes.byref(n), None) == L.EINVAL
    # a window outside the file: fewer than 10 vectors
    far = t0 - dt.timedelta(days=3)
    st, *_ = parse_doris(data, far, far)
    assert st == L.EINVAL and "10 needed" in L.last_error()
    # too small a capacity
    st, *_ = parse_doris(data, start, start, capacity=5)
    assert st == L.EWORKSPACE
    # a record with 10 items, a trailing separator (an empty 12th item), an impossible date, a bad month
    good = doris_file(t0)[1625:1625 + 128].decode()
    items = good.split()
    for bad, what in [(" ".join(items[:10]).ljust(128, "0"), "items"),
                      (" ".join(items).replace(" ", " " * (128 - len(" ".join(items))), 1) + " ", "items"),
                      (good.replace(items[0], "32-JAN-2004", 1), "date"),
                      (good.replace(items[0], items[0][:3] + "XYZ" + items[0][6:], 1), "date")]:
        assert len(bad) == 128
        st, *_ = parse_doris(doris_file(t0, bad_record=3, bad_text=bad), start, start)
        assert st == L.EINVAL and what in L.last_error(), (bad, L.last_error())


def test_doris_time_forms():
    """Month names in either case and fractions of fewer than six digits parse."""
    t0 = dt.datetime(2004, 2, 28, 23, 50, 0)
    data = bytearray(doris_file(t0))
    rec = 1625
    line = data[rec:rec + 128].decode()
    new = line.replace("FEB", "feb", 1).replace(".000000", ".5     ", 1)
    data[rec:rec + 128] = new.encode()
    st, osv, times, _ = parse_doris(bytes(data), t0 + dt.timedelta(minutes=5), t0 + dt.timedelta(minutes=5))
    assert st == L.OK, L.last_error()
    assert (times[0].seconds, times[0].micros) == ((23 * 3600 + 50 * 60), 500000)
    # the leap day and the next month follow
    assert (t0 + dt.timedelta(minutes=10) - EPOCH).days == times[10].days


@pytest.mark.parametrize("name,ok,start,stop", [
    ("DOR_VOR_AXVF-P20120424_120200_20040101_215528_20040103_002328", True,
     dt.datetime(2004, 1, 1, 21, 55, 28), dt.datetime(2004, 1, 3, 0, 23, 28)),
    ("DOR_VOR_AXVF-P19950503_210000_19950503_210000_210000_19950505_030000", True,
     dt.datetime(1995, 5, 3, 21, 0, 0), dt.datetime(1995, 5, 5, 3, 0, 0)),
    ("DOR_VOR_AXVF-P20120424_120200_20040101_215528_20040103_00232", False, None, None),
    ("DOR_XYZ_AXVF-P20120424_120200_20040101_215528_20040103_002328", False, None, None),
    ("DOR_VOR_AXVF-P20120424_120200_20040132_215528_20040103_002328", False, None, None),
])
def test_doris_filename_window(name, ok, start, stop):
    a, b = Mjd(), Mjd()
    st = lib().asarf_doris_filename_window(name.encode(), ctypes.byref(a), ctypes.byref(b))
    assert (st == L.OK) == ok, L.last_error()
    if ok:
        assert (a.days, a.seconds, a.micros) == (mjd(start).days, mjd(start).seconds, 0)
        assert (b.days, b.seconds) == (mjd(stop).days, mjd(stop).seconds)


def put_floats(buf, off, vals):
    buf[off:off + 4 * len(vals)] = struct.pack(f">{len(vals)}f", *vals)


def test_con_and_xca_and_calibration_constant():
    rng = np.random.default_rng(2)
    con = bytearray(rng.integers(0, 256, 5721, dtype=np.uint8).tobytes())
    imp, ims = rng.uniform(1, 100, 7).astype(np.float32), rng.uniform(1, 100, 7).astype(np.float32)
    put_floats(con, 1625 + 1020, imp)
    put_floats(con, 1625 + 1048, ims)
    g_ims, g_imp = (ctypes.c_float * 7)(), (ctypes.c_float * 7)()
    l = lib()
    assert l.asarf_con_process_gains(bytes(con), len(con), g_ims, g_imp) == L.OK
    assert np.array_equal(np.array(g_ims[:], np.float32), ims) and np.array_equal(np.array(g_imp[:], np.float32), imp)
    assert l.asarf_con_process_gains(bytes(con), 5720, g_ims, g_imp) == L.ESHAPE
    xca = bytearray(rng.integers(0, 256, 28177, dtype=np.uint8).tobytes())
    slc, prec = rng.uniform(1e3, 1e5, 7).astype(np.float32), rng.uniform(1e3, 1e5, 7).astype(np.float32)
    put_floats(xca, 1625 + 16 + 28, slc)
    put_floats(xca, 1625 + 16 + 84, prec)
    s_slc, s_prec = (ctypes.c_float * 7)(), (ctypes.c_float * 7)()
    for dsr, want in [(26552, L.OK), (0, L.OK), (26551, L.EINVAL)]:
        xca[1625 + 12:1625 + 16] = struct.pack(">I", dsr)
        assert l.asarf_xca_scaling_factors(bytes(xca), len(xca), s_slc, s_prec) == want
    xca[1625 + 12:1625 + 16] = struct.pack(">I", 26552)
    assert l.asarf_xca_scaling_factors(bytes(xca), len(xca), s_slc, s_prec) == L.OK
    assert np.array_equal(np.array(s_slc[:], np.float32), slc) and np.array_equal(np.array(s_prec[:], np.float32), prec)
    k = ctypes.c_double()
    for detected, ers, gain, sf, div in [(1, 0, imp[2], prec[2], 1), (0, 0, ims[2], slc[2], 1), (0, 1, ims[1], slc[1], 8),
                                         (1, 1, imp[1], prec[1], 1)]:
        assert l.asarf_calibration_constant(detected, ers, float(gain), float(sf), ctypes.byref(k)) == L.OK
        assert k.value == pytest.approx(float(gain) / math.sqrt(float(sf)) / div, rel=1e-15)
    assert l.asarf_calibration_constant(1, 0, 1.0, 0.0, ctypes.byref(k)) == L.EINVAL


def ins_file(rng, swath, fs, f0, bias, amp, phase, dur, rank):
    ins = bytearray(rng.integers(0, 256, 173273, dtype=np.uint8).tobytes())
    flp = 1625
    put_floats(ins, flp + 16, [f0, fs])
    put_floats(ins, flp + 65564 + 36 * swath, [*amp, *phase, dur])
    put_floats(ins, flp + 69912, [bias])
    ins[flp + 72608 + 88986 + 28 + 2 * swath: flp + 72608 + 88986 + 30 + 2 * swath] = struct.pack(">H", rank)
    return bytes(ins)


def test_ins_swath_params_and_radar_timing():
    rng = np.random.default_rng(5)
    fs, f0, bias = np.float32(19.207680e6), np.float32(5.331004e9), np.float32(5.4e-7)
    amp = np.float32([1.0, 0.1, 0.01, 0.0])
    phase = np.float32([0.0, 0.0, 2.9522e11, 0.0])
    dur = np.float32(2.71e-5)
    l = lib()
    for swath in (0, 1, 6):
        ins = ins_file(rng, swath, fs, f0, bias, amp, phase, dur, 9 + swath)
        p = InsSwath()
        assert l.asarf_ins_swath_params(ins, len(ins), swath, ctypes.byref(p)) == L.OK
        assert (p.radar_sampling_rate, p.radar_frequency, p.range_gate_bias) == (fs, f0, bias)
        assert list(p.chirp_amplitude) == list(amp) and list(p.chirp_phase) == list(phase)
        assert p.chirp_duration == dur and p.swst_rank == 9 + swath
        # the FBAQ-4 table reader sees the same file layout (its tables follow the swath parameters)
        i_lut, q_lut = (ctypes.c_float * 4096)(), (ctypes.c_float * 4096)()
        assert L.lib().asarf_ins_fbaq4_tables(ins, len(ins), i_lut, q_lut) == L.OK
    assert l.asarf_ins_swath_params(ins, len(ins), 7, ctypes.byref(p)) == L.EINVAL
    assert l.asarf_ins_swath_params(ins, len(ins) - 1, 0, ctypes.byref(p)) == L.ESHAPE
    t = Timing()
    # ASAR: PRF = fs / PRI code; two-way time (min SWST + rank x PRI code) / fs - bias
    assert l.asarf_radar_timing(ctypes.byref(p), 0, 2876, 1002, 1, ctypes.byref(t)) == L.OK
    fsd = float(fs)
    assert t.prf == pytest.approx(fsd / 2876, rel=1e-15)
    assert t.two_way_slant_range_time == pytest.approx((1002 + p.swst_rank * 2876) / fsd - float(bias), rel=1e-14)
    assert t.slant_range_first == pytest.approx(t.two_way_slant_range_time * C / 2, rel=1e-15)
    assert t.chirp_kr == float(phase[2]) * 2 and t.chirp_samples == round(float(dur) * fsd)
    assert t.chirp_bandwidth == pytest.approx(t.chirp_kr * float(dur), rel=1e-15)
    assert t.wavelength == pytest.approx(C / float(f0), rel=1e-15) and t.range_spacing == pytest.approx(C / (2 * fsd))
    assert t.line_time_interval == pytest.approx(1 / t.prf, rel=1e-15)
    # ERS: PRI = (code + 2) x 210.94 ns; two-way time rank x PRI + min SWST x 4 / fs - bias
    assert l.asarf_radar_timing(ctypes.byref(p), 1, 2820, 220, 4, ctypes.byref(t)) == L.OK
    pri = (2820 + 2.0) * 210.94e-9
    assert t.prf == pytest.approx(1 / pri, rel=1e-15) and 1640 < t.prf < 1720
    assert t.two_way_slant_range_time == pytest.approx(p.swst_rank * pri + 220 * 4 / fsd - float(bias), rel=1e-14)
    assert l.asarf_radar_timing(ctypes.byref(p), 0, 0, 1, 1, ctypes.byref(t)) == L.EINVAL
