"""ctypes binding of the FP64 oracle (build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product path. Arrays are numpy complex128 images with an explicit row stride (the padded image model).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

Q1_IQ_PADDED_WIDTH = 1
Q2_COS_NO_SQRT = 2
Q4_DC_EXTRA_BLOCK = 4
QUIRKS_REFERENCE = Q1_IQ_PADDED_WIDTH | Q2_COS_NO_SQRT | Q4_DC_EXTRA_BLOCK

_dp = ctypes.POINTER(ctypes.c_double)


class IqResults(ctypes.Structure):
    _fields_ = [("dc_i", ctypes.c_double), ("dc_q", ctypes.c_double), ("iq_gain", ctypes.c_double),
                ("phase_error", ctypes.c_double), ("stddev_i", ctypes.c_double), ("stddev_q", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RdaParams(ctypes.Structure):
    _fields_ = [("wavelength", ctypes.c_double), ("range_spacing", ctypes.c_double),
                ("slant_range_first", ctypes.c_double), ("prf", ctypes.c_double),
                ("az_bandwidth_fraction", ctypes.c_double), ("dc_poly", ctypes.c_double * 3),
                ("vr_poly", ctypes.c_double * 3)]


class FocusParams(ctypes.Structure):
    _fields_ = [("na", ctypes.c_int), ("nr", ctypes.c_int), ("nc", ctypes.c_int), ("nrp", ctypes.c_int),
                ("nap", ctypes.c_int), ("prf", ctypes.c_double), ("wavelength", ctypes.c_double),
                ("range_spacing", ctypes.c_double), ("slant_range_first", ctypes.c_double),
                ("az_bandwidth_fraction", ctypes.c_double), ("vr_poly", ctypes.c_double * 3),
                ("chirp_kr", ctypes.c_double), ("range_sampling_rate", ctypes.c_double),
                ("apply_range_window", ctypes.c_int), ("quirks", ctypes.c_int),
                ("n_total_samples", ctypes.c_size_t)]


class FocusResults(ctypes.Structure):
    _fields_ = [("iq", IqResults), ("dc_poly", ctypes.c_double * 3), ("nrv", ctypes.c_int)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build the oracle (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        sig = {
            "or_fft_rows": (None, [vp, ci, ci, ci, ci]),
            "or_fft_cols": (None, [vp, ci, ci, ci, ci]),
            "or_dft": (None, [vp, vp, ci, ci]),
            "or_next_fft_size": (ci, [ci]),
            "or_generate_chirp": (None, [cd, cd, ci, ci, ci, vp]),
            "or_polyfit": (ci, [vp, vp, ci, ci, vp]),
            "or_iq_correction": (None, [vp, ci, ci, ci, ci, ctypes.c_size_t, ci, ctypes.POINTER(IqResults)]),
            "or_doppler_centroid": (ci, [vp, ci, ci, ci, cd, ci, vp, vp, ci, vp]),
            "or_range_compression": (ci, [vp, ci, ci, ci, vp, ci]),
            "or_range_doppler": (None, [vp, vp, ci, ci, ci, ci, ctypes.POINTER(RdaParams)]),
            "or_rcmc": (None, [vp, vp, ci, ci, ci, ctypes.POINTER(RdaParams)]),
            "or_az_matched_filter": (None, [vp, ci, ci, ci, ctypes.POINTER(RdaParams)]),
            "or_rcmc_window": (None, [vp]),
            "or_az_multilook_rows": (ci, [ci, cd]),
            "or_az_multilook": (None, [vp, ci, ci, ci, cd, vp]),
            "or_rsl": (None, [vp, ci, ci, cd, cd, cd]),
            "or_srgr": (None, [vp, ci, ci, vp, cd, cd, ci, vp]),
            "or_aperture_pixels": (ci, [cd, cd, cd, vp, vp, ci, cd, cd, ci]),
            "or_focus_slc": (ci, [vp, ctypes.POINTER(FocusParams), vp, ctypes.POINTER(FocusResults)]),
            "or_set_threads": (None, [ci]),
            "or_get_threads": (ci, []),
            "or_mlp_num_params": (ci, []),
            "or_mlp_forward": (None, [vp, cd, cd, vp]),
            "or_mlp_loss_grad": (cd, [vp, vp, ctypes.c_long, vp]),
            "or_sinc8_weights": (None, [cd, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _c128(a):
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.ctypes.data


def set_threads(n: int):
    lib().or_set_threads(int(n))


def get_threads() -> int:
    return lib().or_get_threads()


def fft_rows(a: np.ndarray, inverse=False, n=None):
    """In place on a 2-D complex128 array: FFT of length n (default a.shape[1]) over every row."""
    n = a.shape[1] if n is None else n
    lib().or_fft_rows(_c128(a), n, a.shape[0], a.shape[1], int(inverse))
    return a


def fft_cols(a: np.ndarray, inverse=False, cols=None):
    cols = a.shape[1] if cols is None else cols
    lib().or_fft_cols(_c128(a), a.shape[0], cols, a.shape[1], int(inverse))
    return a


def dft(x: np.ndarray, inverse=False):
    x = np.ascontiguousarray(x, np.complex128)
    out = np.empty_like(x)
    lib().or_dft(x.ctypes.data, out.ctypes.data, x.size, int(inverse))
    return out


def next_fft_size(n: int) -> int:
    return lib().or_next_fft_size(int(n))


def generate_chirp(kr, fs, n_samples, apply_window, padding) -> np.ndarray:
    out = np.zeros(padding, np.complex64)
    lib().or_generate_chirp(kr, fs, n_samples, int(apply_window), padding, out.ctypes.data)
    return out


def polyfit(x, y, degree):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    c = np.zeros(degree + 1)
    rc = lib().or_polyfit(x.ctypes.data, y.ctypes.data, x.size, degree, c.ctypes.data)
    if rc != 0:
        raise ValueError("polyfit failed")
    return c


def iq_correction(img, x_size, y_size, n_total=0, quirks=QUIRKS_REFERENCE):
    r = IqResults()
    lib().or_iq_correction(_c128(img), x_size, y_size, img.shape[1], img.shape[0], n_total, quirks, ctypes.byref(r))
    return r.as_dict()


def doppler_centroid(img, x_size, y_size, prf, quirks=QUIRKS_REFERENCE):
    bf = np.zeros(64)
    bx = np.zeros(64)
    poly = np.zeros(3)
    nb = lib().or_doppler_centroid(_c128(img), x_size, y_size, img.shape[1], prf, quirks, bf.ctypes.data,
                                   bx.ctypes.data, 64, poly.ctypes.data)
    if nb < 0:
        raise ValueError("doppler centroid failed")
    return poly, bx[:nb].copy(), bf[:nb].copy()


def range_compression(img, x_size, y_size, chirp: np.ndarray, nc):
    ch = np.ascontiguousarray(chirp, np.complex64)
    assert ch.size == img.shape[1]
    return lib().or_range_compression(_c128(img), x_size, y_size, img.shape[1], ch.ctypes.data, nc)


def rda_params(wavelength, range_spacing, r_first, prf, beta, dc_poly, vr_poly) -> RdaParams:
    p = RdaParams()
    p.wavelength, p.range_spacing, p.slant_range_first, p.prf, p.az_bandwidth_fraction = (
        wavelength, range_spacing, r_first, prf, beta)
    for i in range(3):
        p.dc_poly[i] = dc_poly[i]
        p.vr_poly[i] = vr_poly[i]
    return p


def range_doppler(src, x_size, y_size, p: RdaParams):
    out = np.zeros_like(src)
    lib().or_range_doppler(_c128(src), _c128(out), x_size, y_size, src.shape[1], src.shape[0], ctypes.byref(p))
    return out


def rcmc(freq_img, x_size, p: RdaParams):
    out = np.zeros_like(freq_img)
    lib().or_rcmc(_c128(freq_img), _c128(out), x_size, freq_img.shape[0], freq_img.shape[1], ctypes.byref(p))
    return out


def az_matched_filter(freq_img, x_size, p: RdaParams):
    lib().or_az_matched_filter(_c128(freq_img), x_size, freq_img.shape[0], freq_img.shape[1], ctypes.byref(p))
    return freq_img


def rcmc_window() -> np.ndarray:
    w = np.zeros(16, np.float32)
    lib().or_rcmc_window(w.ctypes.data)
    return w


def az_multilook(slc: np.ndarray, az_spacing: float) -> np.ndarray:
    slc = np.ascontiguousarray(slc, np.complex128)
    rows = lib().or_az_multilook_rows(slc.shape[0], az_spacing)
    out = np.zeros((rows, slc.shape[1]))
    lib().or_az_multilook(slc.ctypes.data, slc.shape[1], slc.shape[0], slc.shape[1], az_spacing, out.ctypes.data)
    return out


def rsl(img: np.ndarray, r_first, dr, r_ref=800000.0):
    assert img.dtype == np.float64 and img.flags.c_contiguous
    lib().or_rsl(img.ctypes.data, img.shape[1], img.shape[0], r_first, dr, r_ref)
    return img


def srgr(img: np.ndarray, grsr5, dr, gs, out_nr):
    img = np.ascontiguousarray(img, np.float64)
    g = np.ascontiguousarray(grsr5, np.float64)
    out = np.zeros((img.shape[0], out_nr))
    lib().or_srgr(img.ctypes.data, img.shape[1], img.shape[0], g.ctypes.data, dr, gs, out_nr, out.ctypes.data)
    return out


def aperture_pixels_at(wavelength, prf, beta, vr_poly, dc_poly, r_first, dr, x):
    """CalcAperturePixels at one range index (sar/sar_metadata.h:117-125), C oracle."""
    v = np.ascontiguousarray(vr_poly, np.float64)
    d = np.ascontiguousarray(dc_poly if len(dc_poly) else [0.0], np.float64)
    return lib().or_aperture_pixels(wavelength, prf, beta, v.ctypes.data, d.ctypes.data, len(dc_poly), r_first, dr,
                                    int(x))


def focus_slc(raw_c64: np.ndarray, cfg, quirks=QUIRKS_REFERENCE, nrp=None, nap=None):
    """Whole path on a (na, nr) complex64 raw block. Returns (slc complex128 (na, nrv), results dict)."""
    raw = np.ascontiguousarray(raw_c64, np.complex64)
    na, nr = raw.shape
    fp = FocusParams()
    fp.na, fp.nr, fp.nc = na, nr, cfg.nc
    fp.nrp = cfg.nrp if nrp is None else nrp
    fp.nap = cfg.nap if nap is None else nap
    fp.prf, fp.wavelength, fp.range_spacing = cfg.prf, cfg.wavelength, cfg.range_spacing
    fp.slant_range_first, fp.az_bandwidth_fraction = cfg.r_first, cfg.az_bandwidth_fraction
    for i, v in enumerate(cfg.vr_poly):
        fp.vr_poly[i] = v
    fp.chirp_kr, fp.range_sampling_rate = cfg.kr, cfg.fs
    fp.apply_range_window = int(cfg.apply_range_window)
    fp.quirks = quirks
    fp.n_total_samples = 0
    nrv = nr - cfg.nc
    slc = np.zeros((na, nrv), np.complex128)
    res = FocusResults()
    rc = lib().or_focus_slc(raw.ctypes.data, ctypes.byref(fp), slc.ctypes.data, ctypes.byref(res))
    if rc != 0:
        raise RuntimeError(f"oracle focus failed ({rc})")
    return slc, {"iq": res.iq.as_dict(), "dc_poly": list(res.dc_poly), "nrv": res.nrv}


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even of float32 values to bfloat16, returned as float32 (finite inputs)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32)


def mlp_loss_grad(params: np.ndarray, pairs: np.ndarray, with_grad=True):
    """FP64 loss sum and gradient sum of the learned-RCMC network over float32 pairs (n x 20)."""
    w = np.ascontiguousarray(params, np.float64)
    pr = np.ascontiguousarray(pairs, np.float32)
    g = np.zeros(w.size, np.float64) if with_grad else None
    loss = lib().or_mlp_loss_grad(w.ctypes.data, pr.ctypes.data, pr.shape[0], g.ctypes.data if with_grad else None)
    return loss, g


def mlp_forward(params: np.ndarray, u0: float, u1: float) -> np.ndarray:
    w = np.ascontiguousarray(params, np.float64)
    out = np.zeros(16, np.float64)
    lib().or_mlp_forward(w.ctypes.data, float(u0), float(u1), out.ctypes.data)
    return out


def sinc8_weights(d: float) -> np.ndarray:
    out = np.zeros(8, np.float64)
    lib().or_sinc8_weights(float(d), out.ctypes.data)
    return out


# ------------------------------------------------------------------- level-0 sample decode (numpy restatement)
# These follow the reference's scatter kernels literally: a NaN-filled range_size-wide buffer receives every input
# sample at its shifted column (main_flow.cc:537-540 and the kernels below), and is then copied into a zeroed padded
# image (main.cc:233-239). Samples whose column is outside [0, range_size) are dropped (the reference does not bound
# that write).

def fbaq4_index(block: int, code: int) -> int:
    """Fbaq4IndexCalc (envisat_format/include/envisat_parse_util.cuh:75-84)."""
    conv = 15 - code if code > 7 else 8 + code
    return 256 * conv + block


def _padded(buf: np.ndarray, x_stride: int, y_stride: int) -> np.ndarray:
    out = np.zeros((y_stride, x_stride), np.complex64)
    out[: buf.shape[0], : buf.shape[1]] = buf
    return out


def ingest_ers_im(samples: np.ndarray, samples_per_line: int, swst: np.ndarray, min_swst: int, range_size: int,
                  x_stride: int, y_stride: int) -> np.ndarray:
    """ConvertErsImSamplesToComplexKernel (envisat_format/src/envisat_format_kernels.cu:116-133): byte x of line y
    is component x % 2 of column x / 2 + 4 * (swst[y] - min_swst)."""
    lines = samples.shape[0]
    buf = np.full((lines, range_size), np.nan + 1j * np.nan, np.complex64)
    for y in range(lines):
        off = 4 * (int(swst[y]) - min_swst)
        b = samples[y, : 2 * samples_per_line].astype(np.float32)
        cols = np.arange(samples_per_line) + off
        ok = (cols >= 0) & (cols < range_size)
        buf[y, cols[ok]] = (b[0::2] + 1j * b[1::2])[ok]
    return _padded(buf, x_stride, y_stride)


def ingest_asar_fbaq4(records: np.ndarray, swst: np.ndarray, min_swst: int, i_lut: np.ndarray, q_lut: np.ndarray,
                      range_size: int, x_stride: int, y_stride: int) -> np.ndarray:
    """ConvertAsarImBlocksToComplexKernel (envisat_format/src/envisat_format_kernels.cu:160-210): record = 2-byte
    little-endian length L, then blocks of 64 bytes (block id, 63 codes); byte x >= 2 of record y with k = x - 2 <
    L and k % 64 != 0 goes to column k - (k / 64 + 1) + swst[y] - min_swst, I and Q from the tables."""
    n, item = records.shape
    lut_items = i_lut.size
    buf = np.full((n, range_size), np.nan + 1j * np.nan, np.complex64)
    for y in range(n):
        rec = records[y]
        length = int(rec[0]) | (int(rec[1]) << 8)
        k = np.arange(0, item - 2)
        k = k[(k < length) & (k % 64 != 0)]
        blk = k // 64
        block_id = rec[2 + 64 * blk].astype(np.int64)
        code = rec[2 + k].astype(np.int64)
        ci, cq = code >> 4, code & 15
        ii = 256 * np.where(ci > 7, 15 - ci, 8 + ci) + block_id
        qi = 256 * np.where(cq > 7, 15 - cq, 8 + cq) + block_id
        cols = k - (blk + 1) + (int(swst[y]) - min_swst)
        ok = (ii < lut_items) & (qi < lut_items) & (cols >= 0) & (cols < range_size)
        buf[y, cols[ok]] = i_lut[ii[ok]] + 1j * q_lut[qi[ok]]
    return _padded(buf, x_stride, y_stride)


# ---------------------------------------------------------------- output conditioning (numpy restatement)
def _sat_trunc(v: np.ndarray, lo: float, hi: float) -> np.ndarray:
    """float32 -> integer as the reference's device casts behave: clamp, truncate toward zero, NaN -> 0."""
    v = np.asarray(v, np.float32)
    out = np.trunc(np.clip(np.nan_to_num(v, nan=0.0, posinf=hi, neginf=lo), lo, hi))
    return out.astype(np.int64)


def condition_slc(slc: np.ndarray, x_size: int, y_size: int, calibration: float, header: int) -> np.ndarray:
    """CalibrateClampToBigEndian (envisat_format/src/envisat_format_kernels.cu:28-68): per row, header zero bytes then
    big-endian int16 I and Q of clamp(value * calibration) (float32 product)."""
    v = np.asarray(slc[:y_size, :x_size], np.complex64)
    cal = np.float32(calibration)
    iq = np.empty((y_size, 2 * x_size), ">i2")  # big-endian storage
    iq[:, 0::2] = _sat_trunc(v.real * cal, -32768.0, 32767.0)
    iq[:, 1::2] = _sat_trunc(v.imag * cal, -32768.0, 32767.0)
    out = np.zeros((y_size, header + 4 * x_size), np.uint8)
    out[:, header:] = iq.view(np.uint8).reshape(y_size, 4 * x_size)
    return out


def condition_detected(img: np.ndarray, calibration: float, header: int) -> np.ndarray:
    """CalibrateClampToBigEndianDetected (envisat_format_kernels.cu:70-98): per row, header zero bytes then uint16 of
    10 log10(value * calibration) in float32, saturated to [0, 65535], low byte first (the kernel does not swap)."""
    rows, cols = img.shape
    with np.errstate(divide="ignore", invalid="ignore"):
        val = np.float32(10.0) * np.log10(np.asarray(img, np.float32) * np.float32(calibration))
    u = np.empty((rows, cols), "<u2")  # little-endian storage
    u[:] = _sat_trunc(val, 0.0, 65535.0)
    out = np.zeros((rows, header + 2 * cols), np.uint8)
    out[:, header:] = u.view(np.uint8).reshape(rows, 2 * cols)
    return out


def image_statistics(img: np.ndarray, x_size: int, y_size: int, imaged: bool = False) -> list:
    """CalculateStatistics(Imaged) (main_flow.cc:666-676) over CalcMean / CalcStdDev
    (cuda_util/device_padded_image.cu:331-414): a sample with a NaN in either component is skipped whole (:179,
    :230), sums are divided by x_size * y_size."""
    v = np.asarray(img[:y_size, :x_size], np.complex128)
    n = float(x_size * y_size)
    ok = ~(np.isnan(v.real) | np.isnan(v.imag))
    re, im = v.real[ok], v.imag[ok]
    mi, mq = re.sum() / n, im.sum() / n
    si = np.sqrt(np.sum((re - mi) ** 2) / n)
    sq = np.sqrt(np.sum((im - mq) ** 2) / n)
    if imaged:
        return [(mi + mq) / 2, 0.0, (si + sq) / 2, 0.0]
    return [mi, mq, si, sq]


# ------------------------------------------------------------------- FP64 host geometry (numpy restatement)
WGS84_A, WGS84_B, SPEED_OF_LIGHT = 6378137.0, 6356752.314245, 299792458.0


def interpolate_orbit(ts, pos, vel, t):
    """InterpolateOrbit (sar/orbit_state_vector.cc:23-48): Lagrange through every state vector, times relative to the
    first in whole microseconds."""
    ts = np.asarray(ts, np.float64)
    rel = lambda x: np.round((x - ts[0]) * 1e6) * 1e-6
    x, xs = rel(t), rel(ts)
    p, v = np.zeros(3), np.zeros(3)
    for i in range(len(ts)):
        m = 1.0
        for j in range(len(ts)):
            if i != j:
                m *= (x - xs[j]) / (xs[i] - xs[j])
        p += m * np.asarray(pos[i], np.float64)
        v += m * np.asarray(vel[i], np.float64)
    return p, v


def geolocate(sensor_pos, sensor_vel, guess, slant_range):
    """RangeDopplerGeoLocate (util/src/geo_tools.cc:27-85): Newton on the Doppler, range and WGS84 equations."""
    s, v, p = (np.asarray(a, np.float64) for a in (sensor_pos, sensor_vel, guess))
    p = p.copy()
    a2, b2 = WGS84_A ** 2, WGS84_B ** 2
    for _ in range(101):
        d = p - s
        A = np.array([v, 2.0 * d, [2 * p[0] / a2, 2 * p[1] / a2, 2 * p[2] / b2]])
        B = np.array([v @ d, d @ d - slant_range ** 2, p[0] ** 2 / a2 + p[1] ** 2 / a2 + p[2] ** 2 / b2 - 1.0])
        dx = np.linalg.solve(A, -B)
        p += dx
        if np.linalg.norm(dx) < 1e-6:
            return p
    raise RuntimeError("geolocation did not converge")


def incidence_angle(earth_point, sensor_pos):
    """CalcIncidenceAngle (util/src/geo_tools.cc:87-106), degrees (full-precision b, util/include/geo_tools.h:36)."""
    e = np.asarray(earth_point, np.float64)
    n = np.array([e[0] / WGS84_A ** 2, e[1] / WGS84_A ** 2, e[2] / WGS84_B_FULL ** 2])
    d = np.asarray(sensor_pos, np.float64) - e
    return float(np.degrees(np.arccos(n @ d / (np.linalg.norm(n) * np.linalg.norm(d)))))


def _line_time(g, idx):
    """CalcAzimuthTime (sar/sar_metadata.h:90-93): whole microseconds after the first line."""
    return g["first_line_time"] + np.trunc(g["line_time_interval"] * idx * 1e6) * 1e-6


def estimate_processing_velocity(g):
    """EstimateProcessingVelocity (sar/processing_velocity_estimation.cc:56-121) on a geometry dict."""
    az = g["azimuth_size"] // 2
    range_end = g["range_size"] - g["chirp_samples"]
    step = range_end // 10
    xs, vrs = [], []
    for x in range(0, range_end, step):
        r0 = g["slant_range_first"] + x * g["range_spacing"]
        vr = np.polyval(g["vr_poly"], x)
        ka = 2.0 * g["carrier_frequency"] * vr * vr / (SPEED_OF_LIGHT * r0)
        fmax = np.polyval(g["dc_poly"], x) + g["prf"] * g["az_bandwidth_fraction"] / 2.0
        aperture = int(np.round(fmax / ka * g["prf"]))
        tc = _line_time(g, az)
        cp, cv = interpolate_orbit(g["osv_t"], g["osv_pos"], g["osv_vel"], tc)
        ep = geolocate(cp, cv, g["center_point"], r0)
        astep = aperture // 8
        ts, ys = [], []
        for i in range(9):
            ti = _line_time(g, az + (i - 4) * astep)
            op, _ = interpolate_orbit(g["osv_t"], g["osv_pos"], g["osv_vel"], ti)
            ys.append(np.sum((op - ep) ** 2) - r0 * r0)
            ts.append(np.round((tc - ti) * 1e6) * 1e-6)
        A = np.stack([np.ones(9), np.asarray(ts) ** 2], axis=1)
        sol = np.linalg.lstsq(A, np.asarray(ys), rcond=None)[0]
        xs.append(float(x))
        vrs.append(float(np.sqrt(sol[1])))
    return np.polyfit(xs, vrs, 2), np.asarray(vrs), np.asarray(xs)


def srgr_fit(g, ground_spacing):
    """SlantRangeToGroundRangeInterpolation host part (sar/srgr.cu:73-107) on a geometry dict."""
    p, v = interpolate_orbit(g["osv_t"], g["osv_pos"], g["osv_vel"], _line_time(g, g["azimuth_size"] // 2))
    first = geolocate(p, v, g["center_point"], g["slant_range_first"])
    gd, sr = [], []
    for i in range(100):
        r = g["slant_range_first"] + (i * (g["range_size"] // 100)) * g["range_spacing"]
        q = geolocate(p, v, g["center_point"], r)
        gd.append(np.linalg.norm(q - first))
        sr.append(r - g["slant_range_first"])
    return np.polyfit(gd, sr, 4), int(gd[-1] / ground_spacing)


WGS84_B_FULL = 6356752.3142451794975639665996337
_E2 = 2.0 * ((WGS84_A - WGS84_B_FULL) / WGS84_A) - ((WGS84_A - WGS84_B_FULL) / WGS84_A) ** 2
_EP2 = _E2 / (1.0 - _E2)


def geo_to_xyz(lat, lon, alt=0.0):
    """Geo2xyzWgs84 (util/include/geo_tools.h:52-74)."""
    la, lo = np.radians(lat), np.radians(lon)
    n = WGS84_A / np.sqrt(1.0 - _E2 * np.sin(la) ** 2)
    return np.array([(n + alt) * np.cos(la) * np.cos(lo), (n + alt) * np.cos(la) * np.sin(lo),
                     (n * (1.0 - _E2) + alt) * np.sin(la)])


def xyz_to_geo(xyz):
    """xyz2geoWGS84 (util/include/geo_tools.h:76-95): Bowring's closed form; (latitude, longitude) degrees."""
    x, y, z = (float(v) for v in xyz)
    s = np.hypot(x, y)
    th = np.arctan(z * WGS84_A / (s * WGS84_B_FULL))
    lon = np.degrees(np.arctan(y / x))
    if lon < 0 and y >= 0:
        lon += 180.0
    elif lon > 0 and y < 0:
        lon -= 180.0
    lat = np.degrees(np.arctan((z + _EP2 * WGS84_B_FULL * np.sin(th) ** 3) / (s - _E2 * WGS84_A * np.cos(th) ** 3)))
    return float(lat), float(lon)


def scene_center(g, guess):
    """Scene centre and azimuth spacing (main_flow.cc:499-522)."""
    mid = g["azimuth_size"] // 2
    r = g["slant_range_first"] + ((g["range_size"] - g["chirp_samples"]) // 2) * g["range_spacing"]
    p, v = interpolate_orbit(g["osv_t"], g["osv_pos"], g["osv_vel"], _line_time(g, mid))
    c = geolocate(p, v, guess, r)
    p1, v1 = interpolate_orbit(g["osv_t"], g["osv_pos"], g["osv_vel"], _line_time(g, mid + 1))
    c1 = geolocate(p1, v1, c, r)
    return c, float(np.linalg.norm(c1 - c))


def aperture_pixels(g):
    """CalcAperturePixels (sar/sar_metadata.h:117-141): aperture in lines at range indices 0, step, ..."""
    range_end = g["range_size"] - g["chirp_samples"]
    out = []
    for x in range(0, range_end, range_end // 10):
        r0 = g["slant_range_first"] + x * g["range_spacing"]
        vr = np.polyval(g["vr_poly"], x)
        ka = 2.0 * g["carrier_frequency"] * vr * vr / (SPEED_OF_LIGHT * r0)
        fmax = np.polyval(g["dc_poly"], x) + g["prf"] * g["az_bandwidth_fraction"] / 2.0
        out.append(int(np.round(fmax / ka * g["prf"])))
    return out


def results_window(dc_constant, lines_before_start, lines_after_stop, max_aperture, rcmc_window=16):
    """CalcResultsWindow (main_flow.cc:155-201)."""
    w = dict(lines_before_sensing_start=lines_before_start, lines_after_sensing_start=0,
             lines_before_sensing_stop=0, lines_after_sensing_stop=lines_after_stop,
             near_range_pixels=rcmc_window // 2, far_range_pixels=rcmc_window // 2)
    if dc_constant > 0:
        if lines_before_start < max_aperture:
            w["lines_after_sensing_start"] = max_aperture - lines_before_start
    elif lines_after_stop < max_aperture:
        w["lines_before_sensing_stop"] = max_aperture - lines_after_stop
    return w
