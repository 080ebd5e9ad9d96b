"""Level-1 product header (csrc/l1_writer.cpp, ConstructEnvisatFileHeader envisat_lvl1_writer.cc:326-570): MPH and
SPH text fields, the 18 data set descriptors and the big-endian annotation data sets. The reference has no tests or
sample products for the writer, so the header is read back here field by field: text through the MPH/SPH key
lookup, binary records at the offsets of the ENVISAT format, and every value compared with what went in or with
its closed form (corner geolocation, orbit interpolation, the azimuth FM rate fit, the DC coefficient conversion)."""
import ctypes
import datetime as dt
import math
import struct

import numpy as np
import pytest

from asarf import _lib as L
from asarf import focus
from test_geometry import right_looking_guess, rotating_orbit, scene_geometry

EPOCH = dt.datetime(2000, 1, 1)
T0 = (dt.datetime(2004, 1, 6, 21, 55, 28) - EPOCH).total_seconds()  # first line, s since 2000-01-01
C = 299792458.0


def mjd(t):
    """Seconds since 2000-01-01 as asarf_mjd."""
    d = dt.timedelta(seconds=t)
    return L.Mjd(d.days, d.seconds, d.microseconds)


def text_time(t):
    x = EPOCH + dt.timedelta(seconds=t)
    return x.strftime("%d-%b-%Y %H:%M:%S").upper() + f".{x.microsecond:06d}"


def make_inputs(detected, az=3000, rg=5100):
    """A product 3000 lines long on an inclined orbit over a rotating Earth; all inputs kept alive on the object."""
    class Keep:
        pass
    k = Keep()
    ts = np.arange(-40.0, 61.0, 10.0)
    pos, vel = rotating_orbit(ts)
    g = scene_geometry(ts + T0, pos, vel, first_line_time=T0, azimuth_size=az, range_size=rg,
                       range_spacing=12.5 if detected else 7.8)
    mid_t = T0 + 0.5 * az / 1652.4156
    p, v = rotating_orbit([mid_t - T0])
    g["center_point"] = list(focus.geolocate(p[0], v[0], right_looking_guess(p[0], v[0]), 826e3 + rg / 2 * 7.8))
    g["vr_poly"] = [-1.5e-6, 0.02, 7080.0]
    g["dc_poly"] = [2e-6, -0.03, 180.0]
    k.geom, k.arr = focus.geometry(g)
    k.g = g
    k.meta = L.DorisMeta()
    k.meta.orbit_name = b"DOR_VOR_AXVF-P20120424_120200_20040101_215528_20040103_002328"
    k.meta.vector_source = b"DP"
    k.meta.phase = b"3"
    k.meta.cycle = 23
    k.meta.delta_ut1 = -0.398765
    k.meta.abs_orbit = 9656
    k.meta.state_vector_time = mjd(T0 - 20)
    k.meta.pos[:] = [-2361847.123, 6612345.5, 1234567.875]
    k.meta.vel[:] = [1553.123456, -12.25, -7380.5]
    k.grsr = (ctypes.c_double * 5)(1.2e-13, -3e-9, 5e-5, 0.55, 0.0)
    i = L.L1Inputs()
    i.l0_product_name = b"ASA_IM__0PNPDE20040106_215528_000000162023_00172_09656_0000.N1"
    i.detected = int(detected)
    i.acquisition_station, i.processing_station, i.software_version = b"PDHS-E", b"CGI-EE", b"asarf/0.2"
    i.processing_time = mjd(T0 + 86400 * 3000)
    i.sensing_start, i.sensing_stop = mjd(T0 + 0.5), mjd(T0 + 1.5)
    i.first_line_time = mjd(T0)
    i.last_line_time = mjd(T0 + (az - 1) / 1652.4156)
    i.orbit = ctypes.pointer(k.meta)
    i.rel_orbit = 172
    i.utc_sbt_time, i.sat_binary_time, i.clock_step = mjd(T0 - 3600), 123456789, 3906249
    i.leap_utc, i.leap_sign, i.leap_err, i.product_err = mjd(0), 0, b"0", 0
    i.swath, i.ascending, i.polarization = b"IS2", 0, b"V/V"
    i.echo_comp_method, i.echo_comp_ratio = b"FBAQ", b"4/8"
    i.configuration_file = b"ASA_CON_AXVIEC20030909_000000_20020815_000000_20021017_130000"
    i.instrument_file = b"ASA_INS_AXVIEC20031209_113421_20030211_000000_20041231_000000"
    i.external_characterization_file = b""
    i.external_calibration_file = b"ASA_XCA_AXVIEC20040108_094553_20031201_000000_20041231_000000"
    i.orbit_dataset_name = k.meta.orbit_name
    i.geometry = ctypes.pointer(k.geom)
    i.azimuth_spacing = 4.05
    i.two_way_slant_range_time = 2 * 826e3 / C
    i.range_sampling_rate = 19.20768e6
    i.chirp_window, i.chirp_alpha, i.chirp_kr, i.azimuth_window_alpha = 1, 0.75, 5.9044e11, 0.75
    i.grsr5 = k.grsr if detected else None
    i.first_sbt, i.first_mjd = 0x1234ABCD, mjd(T0)
    (i.first_swst_code, i.last_swst_code, i.pri_code, i.tx_pulse_len_code, i.tx_bw_code, i.echo_win_len_code,
     i.up_code, i.down_code, i.resamp_code, i.beam_adj_code, i.beam_set_num_code,
     i.tx_monitor_code) = 1002, 1013, 2876, 521, 255, 5616, 9, 33, 7777, 4, 17, 200
    i.swst_changes, i.swst_rank, i.tx_monitor_value = 3, 9, -16.064
    i.iq = L.IqResults(0.0952, 0.0714, 0.97846, -0.066, 11.98, 12.24)
    i.statistics[:] = [1200.5, 3.25, 640.0, 641.5] if not detected else [830.25, 0.0, 410.75, 0.0]
    th = L.SqThresholds()
    for n, v in zip([f for f, _ in L.SqThresholds._fields_],
                    [0.1, 0.2, 0.3, 0.5, 0.0, 2.0, 12.0, 30.0, 0.4, 50.0, 1000.0, 100.0, 600.0, 0.01, 0.02, 8]):
        setattr(th, n, v)
    i.sq = th
    i.mds_records = az
    i.mds_record_size = 17 + rg * (2 if detected else 4)
    k.inputs = i
    return k


def header(k):
    n = L.lib().asarf_l1_header_size(k.inputs.detected)
    buf = (ctypes.c_ubyte * n)()
    got = ctypes.c_long()
    L.check(L.lib().asarf_l1_header(ctypes.byref(k.inputs), buf, n, ctypes.byref(got)))
    assert got.value == n
    return bytes(buf)


def get(text, key):
    v = ctypes.create_string_buffer(256)
    L.check(L.lib().asarf_header_get(text, len(text), key.encode(), v, 256))
    return v.value.decode()


def be(fmt, b, off):
    return struct.unpack_from(">" + fmt, b, off)


@pytest.mark.parametrize("detected", [False, True])
def test_mph_and_sph_text(detected):
    k = make_inputs(detected)
    h = header(k)
    mph, sph = h[:1247], h[1247:1247 + 6099]
    # every header line is KEY=value, or blank spare, and the blocks end on newlines
    for block in (mph, sph):
        assert block.endswith(b"\n")
        for line in block.decode().split("\n")[:-1]:
            assert "=" in line or line.strip() == "", line
    name = get(mph, "PRODUCT").rstrip()
    l0 = k.inputs.l0_product_name.decode()
    assert name == l0[:6] + ("P" if detected else "S") + l0[7] + "1" + l0[9:]
    assert get(mph, "REF_DOC").rstrip() == "PO-RS-MDA-GS-2009_4/C" and get(mph, "PROC_STAGE") == "N"
    assert get(mph, "SENSING_START").rstrip() == text_time(T0 + 0.5)
    assert get(mph, "SENSING_STOP").rstrip() == text_time(T0 + 1.5)
    assert get(mph, "PHASE") == "3" and get(mph, "CYCLE") == "+023" and get(mph, "REL_ORBIT") == "+00172"
    assert get(mph, "ABS_ORBIT") == "+09656" and get(mph, "DELTA_UT1") == "-.398765<s>"
    assert get(mph, "X_POSITION") == "-2361847.123<m>" and get(mph, "Z_VELOCITY") == "-7380.500000<m/s>"
    assert get(mph, "VECTOR_SOURCE").rstrip() == "DP" and get(mph, "CLOCK_STEP") == "+0003906249<ps>"
    hdr = L.lib().asarf_l1_header_size(int(detected))
    rec = k.inputs.mds_record_size
    assert get(mph, "TOT_SIZE") == f"{hdr + 3000 * rec:+021d}<bytes>"
    assert get(mph, "SPH_SIZE") == "+0000006099<bytes>" and get(mph, "NUM_DSD") == "+0000000018"
    assert get(mph, "NUM_DATA_SETS") == ("+0000000006" if detected else "+0000000005")
    # SPH
    assert get(sph, "SPH_DESCRIPTOR").rstrip() == ("Image Mode Precision Image" if detected else "Image Mode SLC Image")
    assert get(sph, "FIRST_LINE_TIME").rstrip() == text_time(T0)
    assert get(sph, "SAMPLE_TYPE").rstrip() == ("DETECTED" if detected else "COMPLEX")
    assert get(sph, "COMPRESSION").rstrip() == "FBAQ4" and get(sph, "PASS").rstrip() == "DESCENDING"
    assert get(sph, "AZIMUTH_LOOKS") == ("+004" if detected else "+001")
    assert get(sph, "LINE_LENGTH") == f"{5100:+06d}<samples>"
    assert float(get(sph, "RANGE_SPACING")[:-3]) == np.float32(12.5 if detected else 7.8)
    assert get(sph, "DATA_TYPE").rstrip() == ("UWORD" if detected else "SWORD")
    # corners: the zero-Doppler points of the first and last lines at near, mid and far range
    g = k.g
    for w, t in (("FIRST", T0), ("LAST", T0 + 2999 / 1652.4156)):
        p, v = rotating_orbit([t - T0])
        for where, r in (("NEAR", 826e3), ("MID", 826e3 + g["range_spacing"] * 5100 / 2),
                         ("FAR", 826e3 + g["range_spacing"] * 5100)):
            osv = L.Osv()
            L.check(L.lib().asarf_interpolate_orbit(k.geom.osv, k.geom.n_osv, t, ctypes.byref(osv)))
            xyz = focus.geolocate(osv.pos[:], osv.vel[:], g["center_point"], r)
            ll = (ctypes.c_double * 2)()
            L.check(L.lib().asarf_xyz_to_geo((ctypes.c_double * 3)(*xyz), ll))
            assert get(sph, f"{w}_{where}_LAT") == f"{int(ll[0] * 1e6):+011d}<10-6degN>"
            assert get(sph, f"{w}_{where}_LONG") == f"{int(ll[1] * 1e6):+011d}<10-6degE>"


def dsds(h):
    out = []
    base = 1247 + 6099 - 18 * 280
    for d in range(18):
        t = h[base + 280 * d: base + 280 * (d + 1)]
        out.append({key: get(t, key) for key in ("DS_NAME", "DS_TYPE", "FILENAME", "DS_OFFSET", "DS_SIZE", "NUM_DSR",
                                                  "DSR_SIZE")})
    return out


@pytest.mark.parametrize("detected", [False, True])
def test_data_set_descriptors_point_at_the_annotations(detected):
    k = make_inputs(detected)
    h = header(k)
    ds = dsds(h)
    names = [d["DS_NAME"].rstrip() for d in ds]
    assert names == ["MDS1 SQ ADS", "MDS2 SQ ADS", "MAIN PROCESSING PARAMS ADS", "DOP CENTROID COEFFS ADS",
                     "SR GR ADS", "CHIRP PARAMS ADS", "MDS1 ANTENNA ELEV PATT ADS", "MDS2 ANTENNA ELEV PATT ADS",
                     "GEOLOCATION GRID ADS", "MAP PROJECTION GADS", "MDS1", "MDS2", "LEVEL 0 PRODUCT",
                     "ASAR PROCESSOR CONFIG", "INSTRUMENT CHARACTERIZATION", "EXTERNAL CHARACTERIZATION",
                     "EXTERNAL CALIBRATION", "ORBIT STATE VECTOR 1"]
    num = lambda s: int(s.split("<")[0])  # noqa: E731
    # internal data sets follow each other from the end of the SPH; the MDS starts where the header ends
    pos = 1247 + 6099
    for d in ds:
        if d["FILENAME"].rstrip() in ("NOT USED",) or d["DS_TYPE"] == "R":
            assert num(d["DS_OFFSET"]) == 0 and num(d["DS_SIZE"]) == 0
            continue
        assert num(d["DS_OFFSET"]) == pos
        assert num(d["DS_SIZE"]) == num(d["NUM_DSR"]) * num(d["DSR_SIZE"])
        pos += num(d["DS_SIZE"])
    assert num(ds[10]["DS_OFFSET"]) == len(h)
    assert num(ds[10]["DSR_SIZE"]) == k.inputs.mds_record_size and num(ds[10]["NUM_DSR"]) == 3000
    assert (ds[4]["FILENAME"].rstrip() == "NOT USED") == (not detected)
    assert ds[12]["FILENAME"].rstrip() == k.inputs.l0_product_name.decode()
    assert ds[15]["FILENAME"].rstrip() == "NOT USED" and ds[15]["DS_TYPE"] == "R"
    assert ds[8]["NUM_DSR"] == "+0000000012" and ds[8]["DSR_SIZE"] == "+0000000521<bytes>"


@pytest.mark.parametrize("detected", [False, True])
def test_annotation_records(detected):
    k = make_inputs(detected)
    h = header(k)
    i, g = k.inputs, k.g
    ds = dsds(h)
    off = {d["DS_NAME"].rstrip(): int(d["DS_OFFSET"].split("<")[0]) for d in ds}
    fs = i.range_sampling_rate
    dt_ = 1 / fs
    first = mjd(T0)
    # summary quality
    sq = off["MDS1 SQ ADS"]
    assert be("iII", h, sq) == (first.days, first.seconds, first.micros)
    assert h[sq + 154: sq + 157] == b"IS2"
    assert be("2f", h, sq + 110) == pytest.approx((0.0952, 0.0714), rel=1e-7)
    assert be("I", h, sq + 91)[0] == 8
    out_mean = i.statistics[0]
    # flags: |value - expected| > threshold, both channels for the SLC
    assert h[sq + 19] == int(abs(np.float32(out_mean) - np.float32(1000.0)) > 50.0)
    assert h[sq + 13] == int(abs(np.float32(0.0952) - 0.0) > 0.5 or (not detected and abs(np.float32(0.0714)) > 0.5))
    assert h[sq + 14] == int(abs(np.float32(11.98) - 12.0) > 2.0)
    # main processing parameters
    mp = off["MAIN PROCESSING PARAMS ADS"]
    s0 = mjd(T0 + 0.5)
    assert be("iII", h, mp) == (s0.days, s0.seconds, s0.micros)
    assert h[mp + 25: mp + 28] == b"123" and h[mp + 41: mp + 44] == b"IS2"
    assert be("3f", h, mp + 44) == pytest.approx((g["range_spacing"], 4.05, 1 / 1652.4156), rel=1e-7)
    assert be("2I", h, mp + 56) == (3000, 5100)
    assert h[mp + 64: mp + 69] == (b"UWORD" if detected else b"SWORD")
    assert h[mp + 124] == 1 and h[mp + 123] == int(detected) and h[mp + 127] == int(detected)
    assert be("f", h, mp + 141 + 32)[0] == pytest.approx(0.97846, rel=1e-7)
    dl = mp + 357
    assert be("2I", h, dl) == (0x1234AB, 0xCD)
    assert be("H", h, dl + 60)[0] == 2876 and be("H", h, dl + 120)[0] == 7777
    assert be("f", h, dl + 286)[0] == pytest.approx(1002 * dt_, rel=1e-7)
    assert be("f", h, dl + 306)[0] == pytest.approx(1002 * dt_, rel=1e-7)  # the first SWST code, as the reference
    assert be("f", h, dl + 386)[0] == pytest.approx(16e6, rel=1e-7)
    assert be("f", h, dl + 426)[0] == pytest.approx(25.1625 + 0.52 * 9, rel=1e-7)
    assert be("I", h, dl + 536)[0] == 9
    rp = mp + 975
    assert be("3f", h, rp + 4) == pytest.approx((800000.0, fs, 5.331e9), rel=1e-7)
    assert h[rp + 18: rp + 25] == b"HAMMING" and be("f", h, rp + 25)[0] == pytest.approx(0.75)
    assert be("f", h, rp + 69 + 24)[0] == pytest.approx(5.9044e11 / 2, rel=1e-7)
    # azimuth FM rate: quadratic fit of Ka = -2 Vr^2 / (lambda R0) over two-way time from the first sample
    ap = mp + 1264
    idx = np.array([(5100 * j) // 10 for j in range(10)], float)
    r0 = 826e3 + idx * g["range_spacing"]
    vr = np.polyval(g["vr_poly"], idx)
    ka = -2 * vr ** 2 / (C / 5.331e9 * r0)
    t = 2 * r0 / C - i.two_way_slant_range_time
    fit = np.polyfit(t, ka, 2)
    assert be("3f", h, ap + 25) == pytest.approx((fit[2], fit[1], fit[0]), rel=1e-5, abs=1e-3)
    assert be("f", h, ap + 37)[0] == pytest.approx(i.two_way_slant_range_time * 1e9, rel=1e-7)
    assert be("f", h, ap + 10)[0] == pytest.approx(1652.4156 * 0.8, rel=1e-7)
    assert h[mp + 1593: mp + 1600] == b"FBAQ4/8"
    # orbit state vectors at lines 0, n/4, ...: cm and 1e-5 m/s of the interpolated orbit
    for j in range(5):
        tt = T0 + math.trunc((3000 * j // 4) / 1652.4156 * 1e6) * 1e-6
        osv = L.Osv()
        L.check(L.lib().asarf_interpolate_orbit(k.geom.osv, k.geom.n_osv, tt, ctypes.byref(osv)))
        got = be("iII6i", h, mp + 1765 + 36 * j)
        m = mjd(tt)
        assert got[:3] == (m.days, m.seconds, m.micros)
        assert got[3:6] == tuple(int(v * 1e2) for v in osv.pos) and got[6:] == tuple(int(v * 1e5) for v in osv.vel)
    # Doppler centroid: range-index polynomial to two-way slant time
    dc = off["DOP CENTROID COEFFS ADS"]
    assert be("3f", h, dc + 17) == pytest.approx((180.0, -0.03 / dt_, 2e-6 / dt_ ** 2), rel=1e-7)
    # slant to ground (detected)
    if detected:
        sr = off["SR GR ADS"]
        assert be("5f", h, sr + 21) == pytest.approx((826e3, 0.55, 5e-5, -3e-9, 1.2e-13), rel=1e-7)
    # geolocation grid: 12 records, tie points of their first and last lines
    geo = off["GEOLOCATION GRID ADS"]
    step = 3000 // 11
    for r in (0, 5, 11):
        b = geo + 521 * r
        line = r * 3000 // 11
        assert be("2I", h, b + 13) == (line, step - 1)
        for w, ln, t0 in ((0, line, b + 25), (1, line + step - 1, b + 279)):
            tp = L.TiePoints()
            L.check(L.lib().asarf_geolocation_tie_points(ctypes.byref(k.geom), ln, k.grsr if detected else None,
                                                         ctypes.byref(tp)))
            assert be("11I", h, t0) == tuple(tp.samp_numbers)
            assert be("11i", h, t0 + 132) == tuple(tp.lats) and be("11i", h, t0 + 176) == tuple(tp.longs)


def test_header_errors():
    k = make_inputs(False)
    n = L.lib().asarf_l1_header_size(0)
    buf = (ctypes.c_ubyte * n)()
    got = ctypes.c_long()
    assert L.lib().asarf_l1_header(ctypes.byref(k.inputs), buf, n - 1, ctypes.byref(got)) == L.EWORKSPACE
    k.inputs.mds_record_size += 1
    assert L.lib().asarf_l1_header(ctypes.byref(k.inputs), buf, n, ctypes.byref(got)) == L.ESHAPE
    k.inputs.mds_record_size -= 1
    k.inputs.swath = b"IS12"
    assert L.lib().asarf_l1_header(ctypes.byref(k.inputs), buf, n, ctypes.byref(got)) == L.EINVAL
    k.inputs.swath = b"IS2"
    k.meta.pos[0] = 1.2e8  # a position wider than its MPH field
    assert L.lib().asarf_l1_header(ctypes.byref(k.inputs), buf, n, ctypes.byref(got)) == L.EINVAL
    assert "X_POSITION" in L.last_error()
    k.meta.pos[0] = 1.0
    d = make_inputs(True)
    d.inputs.grsr5 = None
    m = L.lib().asarf_l1_header_size(1)
    dbuf = (ctypes.c_ubyte * m)()
    assert L.lib().asarf_l1_header(ctypes.byref(d.inputs), dbuf, m, ctypes.byref(got)) == L.EINVAL


def test_con_summary_quality_offsets():
    rng = np.random.default_rng(8)
    con = bytearray(rng.integers(0, 256, 5721, dtype=np.uint8).tobytes())
    r = 1625
    vals = rng.uniform(0, 10, 11).astype(np.float32)
    con[r + 16: r + 60] = struct.pack(">11f", *vals)
    con[r + 124: r + 128] = struct.pack(">I", 77)
    con[r + 128: r + 136] = struct.pack(">2f", 1.5, 2.5)
    con[r + 144: r + 152] = struct.pack(">2f", 3.5, 4.5)
    con[r + 192: r + 200] = struct.pack(">2f", 5.5, 6.5)
    for detected, exp in ((0, (1.5, 2.5)), (1, (3.5, 4.5))):
        t = L.SqThresholds()
        L.check(L.lib().asarf_con_summary_quality(bytes(con), len(con), detected, ctypes.byref(t)))
        assert (t.thresh_chirp_broadening, t.thresh_input_gaps) == (vals[0], vals[10])
        assert (t.thresh_output_mean, t.thresh_output_std_dev) == (vals[7], vals[8])
        assert (t.exp_output_mean, t.exp_output_std_dev) == exp and (t.exp_input_mean, t.exp_input_std_dev) == (5.5, 6.5)
        assert t.lines_per_gap == 77
