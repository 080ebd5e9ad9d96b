"""Synthetic scene configurations: the five workload shapes the focusser is measured and checked on.

Each entry carries the radar parameters a real Level-0 product would supply through the host flow (the reference
reads them from the L0 headers and aux files, main_flow.cc:285-528), plus the synthetic-scene recipe.
Range spacing is c / (2 fs); the chirp FM rate is bandwidth / pulse length (sar/sar_metadata.h:34-44).
"""
from __future__ import annotations

import dataclasses
import math

SOL = 299792458.0
SEED_BASE = 0x5EED0000


@dataclasses.dataclass
class SceneConfig:
    name: str
    na: int
    nr: int
    fs: float
    pulse_len: float
    bandwidth: float
    prf: float
    wavelength: float
    vr: float
    r_first: float
    dc_near: float
    dc_far: float
    raw_format: str  # "u8_ers" | "i8"
    n_targets_az: int = 0
    n_targets_rg: int = 0
    n_targets_random: int = 0
    tgt_rg: tuple = (0.15, 0.75)
    tgt_az: tuple = (0.25, 0.75)
    snr_db: float = -999.0
    clutter_power: float = 0.0
    clutter_shape: float = 4.0
    clutter_rho: float = 0.9  # lag-1 azimuth correlation of the clutter (Doppler signature)
    quant_std_lsb: float = 5.0
    az_band_frac: float = 0.8
    az_bandwidth_fraction: float = 0.8  # processing bandwidth, main_flow.cc:525
    apply_range_window: bool = True
    nrp: int = 0  # 0 -> reference padding rule
    nap: int = 0
    seed: int = SEED_BASE
    orbit_height: float = 790e3  # spherical-earth SRGR geometry
    earth_radius: float = 6371e3
    # processing velocity over the range index (highest power first), e.g. estimated from an orbit; empty -> the
    # constant vr the scene was simulated with
    vr_coeffs: tuple = ()

    @property
    def nc(self) -> int:
        # round half away from zero, as std::round in the L0 metadata assembly
        v = self.pulse_len * self.fs
        return int(math.floor(v + 0.5))

    @property
    def kr(self) -> float:
        return self.bandwidth / self.pulse_len

    @property
    def range_spacing(self) -> float:
        return SOL / (2.0 * self.fs)

    @property
    def az_spacing(self) -> float:
        return self.vr / self.prf

    @property
    def vr_poly(self):
        return [float(v) for v in self.vr_coeffs] if self.vr_coeffs else [0.0, 0.0, self.vr]

    def scaled(self, na: int | None = None, nr: int | None = None, name: str | None = None, **kw):
        """A smaller scene with the same radar (for oracle-sized parity tests)."""
        d = dataclasses.asdict(self)
        if na is not None:
            d["na"] = na
        if nr is not None:
            d["nr"] = nr
        if name is not None:
            d["name"] = name
        d.update(kw)
        return SceneConfig(**d)


def ers_block() -> SceneConfig:
    """Plain method, small ERS-like block: 4096 x 5616 uint8 5-bit I/Q, 25 targets, 10 dB SNR."""
    return SceneConfig(
        name="ers_block_4096x5616",
        na=4096, nr=5616, fs=18.96e6, pulse_len=37.12e-6, bandwidth=15.55e6, prf=1679.9,
        wavelength=0.05657, vr=7050.0, r_first=830e3, dc_near=150.0, dc_far=150.0, raw_format="u8_ers",
        n_targets_az=5, n_targets_rg=5, tgt_rg=(0.15, 0.75), tgt_az=(0.25, 0.75), snr_db=10.0,
        quant_std_lsb=5.0, seed=SEED_BASE + 1,
    )


def asar_scene() -> SceneConfig:
    """Plain method, full ENVISAT ASAR IM scene: 30000 x 5616 complex int8, 200 targets in gamma-textured clutter."""
    return SceneConfig(
        name="asar_im_30000x5616",
        na=30000, nr=5616, fs=19.208e6, pulse_len=27.1e-6, bandwidth=16e6, prf=1652.0,
        wavelength=0.05623, vr=7100.0, r_first=850e3, dc_near=-200.0, dc_far=250.0, raw_format="i8",
        n_targets_random=200, tgt_rg=(0.1, 0.9), tgt_az=(0.15, 0.85), clutter_power=1.0, clutter_shape=4.0,
        quant_std_lsb=12.0, seed=SEED_BASE + 2,
    )


def ers_block_split() -> SceneConfig:
    """Split-precision form of the ERS block: same bytes, range 8192-point and azimuth 32768-point FFTs."""
    c = ers_block()
    c.name = "ers_block_split_8192x32768"
    c.nrp = 8192
    c.nap = 32768
    return c


def high_migration() -> SceneConfig:
    """Learned-RCMC form: 16384 x 5616, Doppler centroid swept +-0.4 PRF across range."""
    return SceneConfig(
        name="high_migration_16384x5616",
        na=16384, nr=5616, fs=18.96e6, pulse_len=37.12e-6, bandwidth=15.55e6, prf=1652.0,
        wavelength=0.05657, vr=7050.0, r_first=830e3, dc_near=-0.4 * 1652.0, dc_far=0.4 * 1652.0,
        raw_format="u8_ers", n_targets_az=5, n_targets_rg=5, snr_db=10.0, quant_std_lsb=5.0,
        seed=SEED_BASE + 4,
    )


def batch_blocks(n_blocks: int = 64, na_range=(2048, 16384), seed: int = SEED_BASE + 5) -> list:
    """Heterogeneous batch: half ERS-like blocks (5616 samples, 704-sample chirp, u8), half ASAR swath-varied
    blocks (range 4000-6000 samples, chirp 400-560 samples, i8); azimuth lengths log-uniform in na_range; point
    targets plus 10 dB noise as in the ERS block. Block k is seeded with seed + k."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_blocks):
        na = int(round(float(np.exp(rng.uniform(np.log(na_range[0]), np.log(na_range[1]))))))
        if k % 2 == 0:
            c = ers_block().scaled(na=na, name=f"batch{k:02d}_ers_{na}x5616", n_targets_az=3, n_targets_rg=3,
                                   seed=seed + k)
        else:
            fs = 19.208e6
            nr = int(rng.integers(4000, 6001))
            nc = int(rng.integers(400, 561))
            c = SceneConfig(
                name=f"batch{k:02d}_asar_{na}x{nr}", na=na, nr=nr, fs=fs, pulse_len=nc / fs,
                bandwidth=16e6, prf=1652.0, wavelength=0.05623, vr=7100.0, r_first=850e3, dc_near=-200.0,
                dc_far=250.0, raw_format="i8", n_targets_az=3, n_targets_rg=3, snr_db=10.0, quant_std_lsb=12.0,
                seed=seed + k)
        out.append(c)
    return out


CONFIGS = {
    "ers_block": ers_block,
    "asar_scene": asar_scene,
    "ers_block_split": ers_block_split,
    "high_migration": high_migration,
}
