"""Python-side mirror of the reference's focusing interfaces over the C ABI, with torch tensors as device memory.

Each function names the reference function it mirrors; arguments keep the reference's meaning. Device work runs
on the torch stream the Context was created with, so torch events and synchronisation order correctly with it.
"""
from __future__ import annotations

import ctypes
import os
import math

import numpy as np
import torch

from . import _lib as L
from .configs import SceneConfig


class Context:
    """asarf_ctx bound to a CUDA (HIP) device and a torch stream (CudaInit, cuda_device_init.h:26-43)."""

    def __init__(self, device: int = 0, stream: torch.cuda.Stream | None = None):
        self.device = device
        self.stream = stream if stream is not None else torch.cuda.current_stream(device)
        h = ctypes.c_void_p()
        L.check(L.lib().asarf_ctx_create(device, ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h)))
        self.h = h

    def sync(self):
        L.check(L.lib().asarf_sync(self.h))

    def close(self):
        if self.h:
            L.lib().asarf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def image(t: torch.Tensor, x_size: int, y_size: int) -> L.Image:
    """asarf_image view of a 2-D torch tensor [y_stride, x_stride] (complex64 or float32)."""
    assert t.is_cuda and t.is_contiguous() and t.dim() == 2
    dtype = L.CF32 if t.dtype == torch.complex64 else L.F32
    assert t.dtype in (torch.complex64, torch.float32)
    return L.Image(t.data_ptr(), x_size, y_size, t.shape[1], t.shape[0], dtype)


def iq_correction(ctx: Context, t: torch.Tensor, x_size: int, y_size: int, n_total: int = 0,
                  quirks: int = L.QUIRKS_REFERENCE) -> dict:
    """RawDataCorrection + ZeroNaNs (sar/iq_correction.cu:188-273)."""
    img = image(t, x_size, y_size)
    r = L.IqResults()
    L.check(L.lib().asarf_iq_correction(ctx.h, ctypes.byref(img), n_total, quirks, ctypes.byref(r)))
    return r.as_dict()


def doppler_centroid(ctx: Context, t: torch.Tensor, x_size: int, y_size: int, prf: float,
                     quirks: int = L.QUIRKS_REFERENCE):
    """CalculateDopplerCentroid (sar/fractional_doppler_centroid.cu:98-162) -> (poly[3], block_x, block_freq)."""
    img = image(t, x_size, y_size)
    poly = np.zeros(3)
    bf = np.zeros(64)
    bx = np.zeros(64)
    nb = ctypes.c_int()
    L.check(L.lib().asarf_doppler_centroid(ctx.h, ctypes.byref(img), prf, quirks, poly.ctypes.data, bf.ctypes.data,
                                           bx.ctypes.data, 64, ctypes.byref(nb)))
    return poly, bx[:nb.value].copy(), bf[:nb.value].copy()


def range_compression(ctx: Context, t: torch.Tensor, x_size: int, y_size: int, chirp: np.ndarray, nc: int) -> int:
    """RangeCompression (sar/range_compression.cu:57-111), in place; returns the new x_size."""
    img = image(t, x_size, y_size)
    ch = np.ascontiguousarray(chirp, np.complex64)
    L.check(L.lib().asarf_range_compression(ctx.h, ctypes.byref(img), ch.ctypes.data, ch.size, nc, None))
    return img.x_size


def rda_params(cfg: SceneConfig, dc_poly) -> L.RdaParams:
    p = L.RdaParams()
    p.wavelength, p.range_spacing, p.slant_range_first = cfg.wavelength, cfg.range_spacing, cfg.r_first
    p.prf, p.az_bandwidth_fraction = cfg.prf, cfg.az_bandwidth_fraction
    for i in range(3):
        p.dc_poly[i] = dc_poly[i]
        p.vr_poly[i] = cfg.vr_poly[i]
    return p


def range_doppler(ctx: Context, params: L.RdaParams, src: torch.Tensor, x_size: int, y_size: int,
                  out: torch.Tensor | None = None) -> torch.Tensor:
    """RangeDopplerAlgorithm (sar/range_doppler_algorithm.cu:312-407); src is consumed as scratch."""
    s = image(src, x_size, y_size)
    if out is None:
        out = torch.empty_like(src)
    o = image(out, x_size, y_size)
    L.check(L.lib().asarf_range_doppler(ctx.h, ctypes.byref(params), ctypes.byref(s), ctypes.byref(o), None))
    return out


def ingest_ers_im(ctx: Context, samples: torch.Tensor, samples_per_line: int, swst: torch.Tensor, min_swst: int,
                  out: torch.Tensor, x_size: int) -> L.Image:
    """ConvertErsImSamplesToComplex + NaN fill + padded copy (envisat_format_kernels.cu:116-158,
    main_flow.cc:537-540, main.cc:233-239): samples [lines, line_bytes] uint8, swst [lines] int16/uint16 codes,
    out [y_stride, x_stride] complex64, all on the device; out's first `lines` rows are the logical region."""
    assert samples.is_cuda and samples.dtype == torch.uint8 and samples.dim() == 2 and samples.is_contiguous()
    assert swst.is_cuda and swst.dtype in (torch.int16, torch.uint16) and swst.numel() == samples.shape[0]
    img = image(out, x_size, samples.shape[0])
    L.check(L.lib().asarf_ingest_ers_im(ctx.h, samples.data_ptr(), samples.shape[1], samples_per_line,
                                        samples.shape[0], swst.data_ptr(), min_swst, ctypes.byref(img)))
    return img


def ingest_asar_fbaq4(ctx: Context, records: torch.Tensor, swst: torch.Tensor, min_swst: int, i_lut: torch.Tensor,
                      q_lut: torch.Tensor, out: torch.Tensor, x_size: int) -> L.Image:
    """ConvertAsarImBlocksToComplex + Fbaq4IndexCalc + NaN fill + padded copy (envisat_format_kernels.cu:160-247):
    records [n, record_bytes] uint8, swst [n] codes, i_lut / q_lut float32 tables, out complex64 [y_stride,
    x_stride], all on the device."""
    assert records.is_cuda and records.dtype == torch.uint8 and records.dim() == 2 and records.is_contiguous()
    assert swst.is_cuda and swst.dtype in (torch.int16, torch.uint16) and swst.numel() == records.shape[0]
    assert i_lut.is_cuda and q_lut.is_cuda and i_lut.dtype == torch.float32 and i_lut.numel() == q_lut.numel()
    img = image(out, x_size, records.shape[0])
    L.check(L.lib().asarf_ingest_asar_fbaq4(ctx.h, records.data_ptr(), records.shape[1], records.shape[0],
                                            swst.data_ptr(), min_swst, i_lut.data_ptr(), q_lut.data_ptr(),
                                            i_lut.numel(), ctypes.byref(img)))
    return img


def condition_slc(ctx: Context, slc: torch.Tensor, x_size: int, y_size: int, calibration: float,
                  header_bytes: int) -> torch.Tensor:
    """ConditionResultsSLC (envisat_format_kernels.cu:28-68): uint8 tensor [y_size, header + 4 x_size]."""
    img = image(slc, x_size, y_size)
    out = torch.empty((y_size, header_bytes + 4 * x_size), dtype=torch.uint8, device=slc.device)
    L.check(L.lib().asarf_condition_slc(ctx.h, ctypes.byref(img), calibration, header_bytes, out.data_ptr(),
                                        out.numel()))
    return out


def condition_detected(ctx: Context, img: torch.Tensor, calibration: float, header_bytes: int) -> torch.Tensor:
    """ConditionResultsDetected (envisat_format_kernels.cu:70-98): uint8 tensor [rows, header + 2 cols]."""
    assert img.is_cuda and img.dtype == torch.float32 and img.dim() == 2 and img.is_contiguous()
    rows, cols = img.shape
    out = torch.empty((rows, header_bytes + 2 * cols), dtype=torch.uint8, device=img.device)
    L.check(L.lib().asarf_condition_detected(ctx.h, img.data_ptr(), cols, rows, calibration, header_bytes,
                                             out.data_ptr(), out.numel()))
    return out


def image_statistics(ctx: Context, t: torch.Tensor, x_size: int, y_size: int, imaged: bool = False) -> list:
    """CalculateStatistics / CalculateStatisticsImaged (main_flow.cc:666-676)."""
    img = image(t, x_size, y_size)
    out = (ctypes.c_double * 4)()
    L.check(L.lib().asarf_image_statistics(ctx.h, ctypes.byref(img), int(imaged), out))
    return list(out)


def fbaq4_index(block: int, code: int) -> int:
    """Fbaq4IndexCalc (envisat_parse_util.cuh:75-84), from the library."""
    return int(L.lib().asarf_fbaq4_index(block, code))


def az_multilook(ctx: Context, slc: torch.Tensor, x_size: int, y_size: int, az_spacing: float):
    """AzimuthMultiLook (sar/az_multilook.cu:82-101) -> (float tensor [rows, x_size], new azimuth spacing)."""
    img = image(slc, x_size, y_size)
    ratio = np.float32(12.5 / az_spacing)
    rows = int(np.round(np.float32(y_size) / ratio))
    out = torch.empty((max(rows, 1), x_size), dtype=torch.float32, device=slc.device)
    ws = L.Workspace(out.data_ptr(), out.numel() * 4)
    nrow = ctypes.c_int()
    sp = ctypes.c_double()
    L.check(L.lib().asarf_az_multilook(ctx.h, ctypes.byref(img), az_spacing, ctypes.byref(ws), ctypes.byref(nrow),
                                       ctypes.byref(sp)))
    return out[: nrow.value], sp.value


def range_spreading_loss(ctx: Context, img: torch.Tensor, r_first: float, dr: float, r_ref: float = 800000.0):
    """RangeSpreadingLossCorrection (sar/range_spreading_loss.cu:42-69), in place."""
    assert img.dtype == torch.float32 and img.is_contiguous()
    L.check(L.lib().asarf_range_spreading_loss(ctx.h, img.data_ptr(), img.shape[1], img.shape[0], r_first, dr, r_ref))
    return img


def srgr(ctx: Context, img: torch.Tensor, grsr5, dr: float, gs: float, out_nr: int) -> torch.Tensor:
    """SlantRangeToGroundRangeInterpolation kernel part (sar/srgr.cu:49-71)."""
    g = np.ascontiguousarray(grsr5, np.float64)
    out = torch.empty((img.shape[0], out_nr), dtype=torch.float32, device=img.device)
    L.check(L.lib().asarf_srgr(ctx.h, img.data_ptr(), img.shape[1], img.shape[0], g.ctypes.data, dr, gs, out_nr,
                               out.data_ptr()))
    return out


def srgr_fit_spherical(r_first, dr, range_size, orbit_radius, earth_radius, gs=12.5):
    g = np.zeros(5)
    n = ctypes.c_int()
    L.check(L.lib().asarf_srgr_fit_spherical(r_first, dr, range_size, orbit_radius, earth_radius, gs, g.ctypes.data,
                                             ctypes.byref(n)))
    return g, n.value


# ----------------------------------------------------------------------------------- FP64 host geometry ----
def _osv_array(ts, pos, vel):
    arr = (L.Osv * len(ts))()
    for i, t in enumerate(ts):
        arr[i].t = float(t)
        for k in range(3):
            arr[i].pos[k] = float(pos[i][k])
            arr[i].vel[k] = float(vel[i][k])
    return arr


def _vec3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


def interpolate_orbit(ts, pos, vel, t: float):
    """InterpolateOrbit (sar/orbit_state_vector.cc:23-48): (position, velocity) at time t."""
    arr = _osv_array(ts, pos, vel)
    o = L.Osv()
    L.check(L.lib().asarf_interpolate_orbit(arr, len(ts), float(t), ctypes.byref(o)))
    return np.array(o.pos[:]), np.array(o.vel[:])


def geolocate(sensor_pos, sensor_vel, guess, slant_range: float) -> np.ndarray:
    """RangeDopplerGeoLocate (util/src/geo_tools.cc:27-85): zero-Doppler WGS84 point at the slant range."""
    out = _vec3([0, 0, 0])
    L.check(L.lib().asarf_geolocate(_vec3(sensor_pos), _vec3(sensor_vel), _vec3(guess), float(slant_range), out))
    return np.array(out[:])


def incidence_angle(earth_point, sensor_pos) -> float:
    """CalcIncidenceAngle (util/src/geo_tools.cc:87-106), degrees."""
    return float(L.lib().asarf_incidence_angle(_vec3(earth_point), _vec3(sensor_pos)))


def geometry(g: dict):
    """asarf_geometry from a dict with the keys of the struct (orbit as osv_t / osv_pos / osv_vel); returns the
    struct and the state-vector array it points to (keep both alive)."""
    arr = _osv_array(g["osv_t"], g["osv_pos"], g["osv_vel"])
    s = L.Geometry()
    s.osv = ctypes.cast(arr, ctypes.POINTER(L.Osv))
    s.n_osv = len(g["osv_t"])
    for k in ("first_line_time", "line_time_interval", "slant_range_first", "range_spacing", "carrier_frequency",
              "prf", "az_bandwidth_fraction"):
        setattr(s, k, float(g[k]))
    for k in ("azimuth_size", "range_size", "chirp_samples"):
        setattr(s, k, int(g[k]))
    for k in ("vr_poly", "dc_poly", "center_point"):
        getattr(s, k)[:] = [float(v) for v in g[k]]
    return s, arr


def estimate_processing_velocity(g: dict):
    """EstimateProcessingVelocity (sar/processing_velocity_estimation.cc:98-121): (Vr polynomial, Vr samples,
    range indices of the samples)."""
    s, keep = geometry(g)
    poly, vrs, xs, n = np.zeros(3), np.zeros(16), np.zeros(16), ctypes.c_int()
    L.check(L.lib().asarf_estimate_processing_velocity(ctypes.byref(s), poly.ctypes.data_as(L._P(ctypes.c_double)),
                                                        vrs.ctypes.data_as(L._P(ctypes.c_double)),
                                                        xs.ctypes.data_as(L._P(ctypes.c_double)), ctypes.byref(n)))
    del keep
    return poly, vrs[:n.value], xs[:n.value]


def srgr_fit(g: dict, ground_spacing: float = 12.5):
    """Slant-to-ground polynomial and output width from the orbit (sar/srgr.cu:73-107)."""
    s, keep = geometry(g)
    poly, n = np.zeros(5), ctypes.c_int()
    L.check(L.lib().asarf_srgr_fit(ctypes.byref(s), float(ground_spacing), poly.ctypes.data_as(L._P(ctypes.c_double)),
                                   ctypes.byref(n)))
    del keep
    return poly, n.value


def fft(ctx: Context, t: torch.Tensor, layout: int = L.FFT_ROWS, inverse: bool = False, scale: float = 1.0,
        n: int | None = None, batch: int | None = None, split_bf16: bool = False):
    """Batched C2C FFT in place (rows: transforms along the last axis; columns: along the first). split_bf16
    selects the matrix-core transforms with bfloat16 hi/lo operands."""
    assert t.dtype == torch.complex64 and t.is_contiguous() and t.is_cuda and t.dim() == 2
    if layout == L.FFT_ROWS:
        n = t.shape[1] if n is None else n
        batch = t.shape[0] if batch is None else batch
    else:
        n = t.shape[0] if n is None else n
        batch = t.shape[1] if batch is None else batch
    f = L.lib().asarf_fft_c2c_split_bf16 if split_bf16 else L.lib().asarf_fft_c2c
    L.check(f(ctx.h, t.data_ptr(), n, batch, t.shape[1], layout, L.INVERSE if inverse else L.FORWARD, scale))
    return t


def generate_chirp(cfg: SceneConfig, padding: int) -> np.ndarray:
    """GenerateChirpData (sar/sar_chirp.cc:70-104) via the library's FP64 host implementation."""
    ci = L.ChirpInfo(cfg.fs, cfg.pulse_len, cfg.nc, cfg.bandwidth, cfg.kr, int(cfg.apply_range_window))
    out = np.zeros(padding, np.complex64)
    L.check(L.lib().asarf_generate_chirp(ctypes.byref(ci), padding, out.ctypes.data))
    return out


def padding_size(n: int) -> int:
    return L.lib().asarf_fft_padding_size(int(n))


def fast_range_size(n: int) -> int:
    """A range transform length for n = nr + nc with a compile-time range kernel (same lines up to FP32 rounding
    as the reference's padding_size(n)); padding_size(n) when none is within 1.5 times it."""
    return L.lib().asarf_fft_fast_range_size(int(n))


def static_column_length(n: int) -> bool:
    """True when azimuth sub-transforms of length n run a compile-time column kernel."""
    return bool(L.lib().asarf_fft_static_column_length(int(n)))


def subset_copy(ctx: Context, img: L.Image, x0: int, y0: int, w: int, h: int, dst: torch.Tensor):
    eb = 8 if img.dtype == L.CF32 else 4
    L.check(L.lib().asarf_image_subset_copy(ctx.h, ctypes.byref(img), x0, y0, w, h, dst.data_ptr(),
                                            dst.stride(0) * eb))
    return dst


def focus_params(cfg: SceneConfig, quirks: int = L.QUIRKS_REFERENCE, imp: bool = False,
                 orbit_height: float | None = None, lines_before: int = 0, lines_after: int = 0,
                 srgr: tuple | None = None, azimuth_spacing: float | None = None, lines: tuple | None = None,
                 fbaq: tuple | None = None):
    """asarf_focus_params of a scene configuration (and the SRGR fit when imp). lines_before / lines_after are the
    lines cached before the sensing start and after the sensing stop (main.cc:296-298). srgr = (grsr5, ground
    range size) replaces the spherical-earth fit (e.g. asarf_srgr_fit from an orbit); azimuth_spacing replaces
    vr / prf. lines = (samples_per_line, line_bytes, swst_mult) makes the input ERS level-0 lines
    (ASARF_RAW_ERS_LINES), fbaq = (record_bytes, swst_mult) ASAR FBAQ-4 records (ASARF_RAW_ASAR_FBAQ4); cfg.nr is
    then the range size after the SWST shifts."""
    p = L.FocusParams()
    p.na, p.nr, p.nc = cfg.na, cfg.nr, cfg.nc
    p.nrp = cfg.nrp or padding_size(cfg.nr + cfg.nc)
    p.nap = cfg.nap or padding_size(cfg.na)
    p.raw_format = L.RAW_U8 if cfg.raw_format == "u8_ers" else L.RAW_I8
    if lines is not None:
        p.raw_format = L.RAW_ERS_LINES
        p.samples_per_line, p.line_bytes, p.swst_mult = int(lines[0]), int(lines[1]), int(lines[2])
    if fbaq is not None:
        p.raw_format = L.RAW_ASAR_FBAQ4
        p.samples_per_line, p.line_bytes, p.swst_mult = 0, int(fbaq[0]), int(fbaq[1])
    p.prf, p.wavelength, p.range_spacing = cfg.prf, cfg.wavelength, cfg.range_spacing
    p.slant_range_first, p.az_bandwidth_fraction = cfg.r_first, cfg.az_bandwidth_fraction
    for i, v in enumerate(cfg.vr_poly):
        p.vr_poly[i] = v
    p.n_total_samples = 0
    p.quirks = quirks
    p.imp = int(imp)
    p.lines_before, p.lines_after = int(lines_before), int(lines_after)
    grsr, ng = None, None
    if imp:
        near = 8
        nr_s = cfg.nr - cfg.nc - 16
        h = cfg.orbit_height if orbit_height is None else orbit_height
        if srgr is not None:
            grsr, ng = [float(v) for v in srgr[0]], int(srgr[1])
        else:
            grsr, ng = srgr_fit_spherical(cfg.r_first + near * cfg.range_spacing, cfg.range_spacing, nr_s,
                                          cfg.earth_radius + h, cfg.earth_radius, 12.5)
        p.azimuth_spacing = cfg.az_spacing if azimuth_spacing is None else azimuth_spacing
        p.reference_range = 800000.0
        for i in range(5):
            p.grsr_poly[i] = grsr[i]
        p.gr_range_size = ng
        p.ground_spacing = 12.5
    return p, grsr, ng


class Plan:
    """Fused focusing plan (the GPU stage of main.cc:233-327) for one scene configuration."""

    def __init__(self, ctx: Context, cfg: SceneConfig, quirks: int = L.QUIRKS_REFERENCE, imp: bool = False,
                 chirp: np.ndarray | None = None, orbit_height: float | None = None, lines_before: int = 0,
                 lines_after: int = 0, srgr: tuple | None = None, azimuth_spacing: float | None = None,
                 lines: tuple | None = None, fbaq: tuple | None = None):
        self.ctx = ctx
        self.cfg = cfg
        p, self.grsr, self.ng = focus_params(cfg, quirks, imp, orbit_height, lines_before, lines_after, srgr,
                                             azimuth_spacing, lines, fbaq)
        self.nrv = cfg.nr - cfg.nc
        self.params = p
        self.chirp = generate_chirp(cfg, p.nrp) if chirp is None else np.ascontiguousarray(chirp, np.complex64)
        h = ctypes.c_void_p()
        L.check(L.lib().asarf_plan_create(ctx.h, ctypes.byref(p), self.chirp.ctypes.data, ctypes.byref(h)))
        self.h = h
        rows, cols = ctypes.c_int(), ctypes.c_int()
        L.check(L.lib().asarf_plan_imp_capacity(h, ctypes.byref(rows), ctypes.byref(cols)))
        self.imp_shape = (rows.value, cols.value)
        self.imp_out = None
        if imp:
            self.imp_out = torch.zeros(self.imp_shape, dtype=torch.float32, device=f"cuda:{ctx.device}")
        self.slc = L.Image()

    @property
    def device_bytes(self) -> int:
        return int(L.lib().asarf_plan_device_bytes(self.h))

    def execute(self, raw: torch.Tensor):
        assert raw.is_cuda and raw.is_contiguous() and raw.numel() == self.cfg.na * self.cfg.nr * 2
        L.check(L.lib().asarf_plan_execute(self.h, raw.data_ptr(), ctypes.byref(self.slc),
                                           self.imp_out.data_ptr() if self.imp_out is not None else None))
        return self.slc

    def execute_lines(self, lines: torch.Tensor, swst: torch.Tensor, min_swst: int):
        """asarf_plan_execute_lines: ERS level-0 lines [na, line_bytes] uint8 and their SWST codes [na] int16, on
        the device."""
        assert lines.is_cuda and lines.dtype == torch.uint8 and lines.is_contiguous()
        assert lines.shape[0] == self.cfg.na and lines.shape[1] == self.params.line_bytes
        assert swst.is_cuda and swst.dtype == torch.int16 and swst.numel() == self.cfg.na
        L.check(L.lib().asarf_plan_execute_lines(self.h, lines.data_ptr(), swst.data_ptr(), int(min_swst),
                                                 ctypes.byref(self.slc),
                                                 self.imp_out.data_ptr() if self.imp_out is not None else None))
        return self.slc

    def execute_fbaq4(self, records: torch.Tensor, swst: torch.Tensor, min_swst: int, i_lut: torch.Tensor,
                      q_lut: torch.Tensor):
        """asarf_plan_execute_fbaq4: ASAR FBAQ-4 records [na, record_bytes] uint8, their SWST codes [na] int16 and
        the two 4096-entry float32 reconstruction tables, on the device."""
        assert records.is_cuda and records.dtype == torch.uint8 and records.is_contiguous()
        assert records.shape[0] == self.cfg.na and records.shape[1] == self.params.line_bytes
        assert swst.is_cuda and swst.dtype == torch.int16 and swst.numel() == self.cfg.na
        assert i_lut.is_cuda and q_lut.is_cuda and i_lut.dtype == torch.float32 and q_lut.dtype == torch.float32
        assert i_lut.is_contiguous() and q_lut.is_contiguous() and i_lut.numel() == q_lut.numel()
        L.check(L.lib().asarf_plan_execute_fbaq4(self.h, records.data_ptr(), swst.data_ptr(), int(min_swst),
                                                 i_lut.data_ptr(), q_lut.data_ptr(), i_lut.numel(),
                                                 ctypes.byref(self.slc),
                                                 self.imp_out.data_ptr() if self.imp_out is not None else None))
        return self.slc

    def rcmc_pairs(self, raw: torch.Tensor, n: int, seed: int) -> torch.Tensor:
        """n training pairs (n x 20 float32, device) for the learned RCMC interpolator from this scene."""
        assert raw.is_cuda and raw.is_contiguous() and raw.numel() == self.cfg.na * self.cfg.nr * 2
        out = torch.empty((n, RCMC_PAIR_FLOATS), dtype=torch.float32, device=f"cuda:{self.ctx.device}")
        L.check(L.lib().asarf_plan_rcmc_pairs(self.h, raw.data_ptr(), n, seed, out.data_ptr()))
        return out

    def set_fft(self, split_bf16: bool):
        L.check(L.lib().asarf_plan_set_fft(self.h, FFT_SPLIT_BF16 if split_bf16 else FFT_FP32))

    def set_rcmc(self, interpolator: int, net: "RcmcNet | None" = None):
        L.check(L.lib().asarf_plan_set_rcmc(self.h, interpolator, net.h if net is not None else None))
        self._net = net  # keep the table alive while the plan uses it

    def slc_tensor(self) -> torch.Tensor:
        """Copy of the focused SLC (na x nrv complex64)."""
        out = torch.empty((self.slc.y_size, self.slc.x_size), dtype=torch.complex64,
                          device=f"cuda:{self.ctx.device}")
        subset_copy(self.ctx, self.slc, 0, 0, self.slc.x_size, self.slc.y_size, out)
        return out

    def results(self) -> dict:
        r = L.FocusResults()
        L.check(L.lib().asarf_plan_results(self.h, ctypes.byref(r)))
        nb = r.n_blocks
        return {"iq": r.iq.as_dict(), "dc_poly": list(r.dc_poly), "n_blocks": nb,
                "block_freq": list(r.block_freq[:nb]), "block_x": list(r.block_x[:nb]), "nrv": r.nrv,
                "crop": (r.crop_front, r.crop_back, r.crop_near, r.crop_far), "imp_rows": r.imp_rows,
                "azimuth_core_fused": bool(r.azimuth_core_fused), "rcmc_shift_bound": r.rcmc_shift_bound}

    def set_timing(self, on: bool):
        L.check(L.lib().asarf_plan_set_timing(self.h, int(on)))

    def kernel_times(self) -> dict:
        names = (ctypes.c_char_p * 32)()
        ms = (ctypes.c_float * 32)()
        n = ctypes.c_int()
        L.check(L.lib().asarf_plan_kernel_times(self.h, names, ms, 32, ctypes.byref(n)))
        return {names[i].decode(): ms[i] for i in range(min(n.value, 32))}

    def close(self):
        if self.h:
            L.lib().asarf_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


RCMC_PAIR_FLOATS = 20
FFT_FP32, FFT_SPLIT_BF16 = 0, 1
RCMC_SINC16, RCMC_LEARNED, RCMC_SINC8 = 0, 1, 2


class RcmcNet:
    """Learned RCMC interpolator (network, trainer, tap table) on one context."""

    def __init__(self, ctx: Context, seed: int = 1):
        self.ctx = ctx
        h = ctypes.c_void_p()
        L.check(L.lib().asarf_rcmc_net_create(ctx.h, seed, ctypes.byref(h)))
        self.h = h
        self.n_params = L.lib().asarf_rcmc_net_num_params()

    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        L.check(L.lib().asarf_rcmc_net_get_params(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                  self.n_params))
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        L.check(L.lib().asarf_rcmc_net_set_params(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                  self.n_params))

    def grad(self, pairs: torch.Tensor):
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_double()
        L.check(L.lib().asarf_rcmc_net_grad(self.h, pairs.data_ptr(), pairs.shape[0],
                                            g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(loss)))
        return loss.value, g

    def train(self, pairs: torch.Tensor, batch: int, steps: int, lr: float) -> np.ndarray:
        hist = np.zeros(steps, np.float32)
        L.check(L.lib().asarf_rcmc_net_train(self.h, pairs.data_ptr(), pairs.shape[0], batch, steps, lr,
                                             hist.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return hist

    def build_table(self, val_pairs: torch.Tensor, nd=8192, ns=2, judge_nd=256, threshold=1e-3):
        nj = judge_nd * ns
        rms = np.zeros(nj, np.float32)
        rms8 = np.zeros(nj, np.float32)
        nok = ctypes.c_int()
        fp = ctypes.POINTER(ctypes.c_float)
        L.check(L.lib().asarf_rcmc_net_build_table(self.h, val_pairs.data_ptr(), val_pairs.shape[0], nd, ns, judge_nd,
                                                   threshold, rms.ctypes.data_as(fp), rms8.ctypes.data_as(fp),
                                                   ctypes.byref(nok)))
        return {"rms": rms.reshape(ns, judge_nd), "sinc8_rms": rms8.reshape(ns, judge_nd), "n_ok": nok.value}

    def close(self):
        if self.h:
            L.lib().asarf_rcmc_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(ctx: Context, n: int, columns: bool, batch: int, reps: int = 3) -> dict:
    """Runs the library's autotuner for one transform length; later plans on ctx use its choice."""
    choice, n_cand = ctypes.c_int(), ctypes.c_int()
    ms = ctypes.c_float()
    L.check(L.lib().asarf_autotune_fft(ctx.h, n, L.FFT_COLUMNS if columns else L.FFT_ROWS, batch, reps,
                                       ctypes.byref(choice), ctypes.byref(ms), ctypes.byref(n_cand)))
    return {"choice": choice.value, "ms": ms.value, "candidates": n_cand.value}


def tune_save(ctx: Context, path: str) -> None:
    """Writes ctx's autotuner choices to a cache file."""
    L.check(L.lib().asarf_tune_save(ctx.h, os.fsencode(path)))


def tune_load(ctx: Context, path: str) -> int:
    """Reads an autotuner cache file into ctx (same device architecture); returns the number of choices."""
    n = ctypes.c_int()
    L.check(L.lib().asarf_tune_load(ctx.h, os.fsencode(path), ctypes.byref(n)))
    return n.value


BATCH_SEQUENTIAL, BATCH_GRAPH = 0, 1


class Batch:
    """Blocks of uneven size focused as one batch (sequential launches or one captured graph)."""

    def __init__(self, ctx: Context, n_streams: int = 4):
        self.ctx = ctx
        h = ctypes.c_void_p()
        L.check(L.lib().asarf_batch_create(ctx.h, n_streams, ctypes.byref(h)))
        self.h = h
        self._keep = []
        self.cfgs = []

    def add(self, cfg: SceneConfig, raw: torch.Tensor, quirks: int = L.QUIRKS_REFERENCE) -> int:
        assert raw.is_cuda and raw.is_contiguous() and raw.numel() == cfg.na * cfg.nr * 2
        p, _, _ = focus_params(cfg, quirks, False)
        chirp = generate_chirp(cfg, p.nrp)
        idx = ctypes.c_int()
        L.check(L.lib().asarf_batch_add(self.h, ctypes.byref(p), chirp.ctypes.data, raw.data_ptr(), None,
                                        ctypes.byref(idx)))
        self._keep.append((raw, chirp))
        self.cfgs.append(cfg)
        return idx.value

    def execute(self, mode: int):
        L.check(L.lib().asarf_batch_execute(self.h, mode))

    def sync(self):
        L.check(L.lib().asarf_batch_sync(self.h))

    def slc_tensor(self, i: int) -> torch.Tensor:
        img = L.Image()
        L.check(L.lib().asarf_batch_slc(self.h, i, ctypes.byref(img)))
        out = torch.empty((img.y_size, img.x_size), dtype=torch.complex64, device=f"cuda:{self.ctx.device}")
        subset_copy(self.ctx, img, 0, 0, img.x_size, img.y_size, out)
        return out

    def close(self):
        if self.h:
            L.lib().asarf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def geo_to_xyz(latitude: float, longitude: float, altitude: float = 0.0) -> np.ndarray:
    """Geo2xyzWgs84 (util/include/geo_tools.h:52-74)."""
    out = _vec3([0, 0, 0])
    L.check(L.lib().asarf_geo_to_xyz(float(latitude), float(longitude), float(altitude), out))
    return np.array(out[:])


def xyz_to_geo(xyz):
    """xyz2geoWGS84 (util/include/geo_tools.h:76-95): (latitude, longitude) in degrees."""
    out = (ctypes.c_double * 2)()
    L.check(L.lib().asarf_xyz_to_geo(_vec3(xyz), out))
    return out[0], out[1]


def scene_center(g: dict, guess_xyz):
    """Scene centre and azimuth spacing (main_flow.cc:499-522)."""
    s, keep = geometry(g)
    c, sp = _vec3([0, 0, 0]), ctypes.c_double()
    L.check(L.lib().asarf_scene_center(ctypes.byref(s), _vec3(guess_xyz), c, ctypes.byref(sp)))
    del keep
    return np.array(c[:]), sp.value


def aperture_pixels(g: dict) -> list:
    """CalcAperturePixels (sar/sar_metadata.h:117-141)."""
    s, keep = geometry(g)
    out, n = (ctypes.c_int * 16)(), ctypes.c_int()
    L.check(L.lib().asarf_aperture_pixels(ctypes.byref(s), out, 16, ctypes.byref(n)))
    del keep
    return list(out[:n.value])


def results_window(dc_constant: float, lines_before_start: int, lines_after_stop: int, max_aperture: int,
                   rcmc_window: int = 16) -> dict:
    """CalcResultsWindow (main_flow.cc:155-201)."""
    w = L.ResultsWindow()
    L.check(L.lib().asarf_results_window(float(dc_constant), int(lines_before_start), int(lines_after_stop),
                                         int(max_aperture), int(rcmc_window), ctypes.byref(w)))
    return {k: getattr(w, k) for k, _ in L.ResultsWindow._fields_}
