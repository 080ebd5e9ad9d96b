"""Point-target impulse-response analysis: 3 dB resolution (IRW), PSLR and ISLR along range and azimuth cuts.

A window around the peak is upsampled by zero-padding its 2-D spectrum (the focused image is band-limited), the
peak is re-located on the fine grid and 1-D cuts through it are measured:
  IRW  : width between the -3 dB points, in original samples;
  PSLR : highest sidelobe outside the main lobe (first nulls) relative to the peak, dB;
  ISLR : sidelobe energy within the window / main-lobe energy (null to null), dB.
"""
from __future__ import annotations

import numpy as np


def _upsample2d(win: np.ndarray, up: int) -> np.ndarray:
    ny, nx = win.shape
    f = np.fft.fftshift(np.fft.fft2(win))
    py, px = ny * (up - 1), nx * (up - 1)
    f = np.pad(f, ((py // 2, py - py // 2), (px // 2, px - px // 2)))
    return np.fft.ifft2(np.fft.ifftshift(f)) * (up * up)


def upsample1d(x: np.ndarray, up: int) -> np.ndarray:
    """Band-limited interpolation of a 1-D complex signal by zero-padding its centred spectrum."""
    n = len(x)
    f = np.fft.fftshift(np.fft.fft(x))
    pad = n * (up - 1)
    f = np.pad(f, (pad // 2, pad - pad // 2))
    return np.fft.ifft(np.fft.ifftshift(f)) * up


def cut_metrics(cut: np.ndarray, up: int) -> dict:
    """IRW / PSLR / ISLR of an already upsampled 1-D cut (up = upsampling factor)."""
    return _cut_metrics(cut, up)


def _cut_metrics(cut: np.ndarray, up: int) -> dict:
    p = np.abs(cut) ** 2
    k = int(np.argmax(p))
    pk = p[k]
    half = pk / 2.0
    # -3 dB points by linear interpolation on the fine grid
    i = k
    while i > 0 and p[i] > half:
        i -= 1
    left = i + (half - p[i]) / (p[i + 1] - p[i]) if p[i + 1] != p[i] else i
    j = k
    while j < len(p) - 1 and p[j] > half:
        j += 1
    right = j - 1 + (p[j - 1] - half) / (p[j - 1] - p[j]) if p[j - 1] != p[j] else j
    irw = (right - left) / up
    # first nulls: walk outwards while power decreases
    a = k
    while a > 0 and p[a - 1] < p[a]:
        a -= 1
    b = k
    while b < len(p) - 1 and p[b + 1] < p[b]:
        b += 1
    main = p[a:b + 1].sum()
    side = p[:a].sum() + p[b + 1:].sum()
    side_peak = max(p[:a].max() if a > 0 else 0.0, p[b + 1:].max() if b < len(p) - 1 else 0.0)
    pslr = 10 * np.log10(side_peak / pk) if side_peak > 0 else -np.inf
    islr = 10 * np.log10(side / main) if side > 0 else -np.inf
    return {"irw": float(irw), "pslr": float(pslr), "islr": float(islr)}


def point_target(img: np.ndarray, x: float, y: float, win: int = 32, up: int = 16, search: int = 3) -> dict:
    """Metrics of the target nearest (x, y) (x = range column, y = azimuth row) in a complex image."""
    img = np.asarray(img)
    ny, nx = img.shape
    xi, yi = int(round(x)), int(round(y))
    y0, y1 = max(0, yi - search), min(ny, yi + search + 1)
    x0, x1 = max(0, xi - search), min(nx, xi + search + 1)
    sub = np.abs(img[y0:y1, x0:x1])
    dy, dx = np.unravel_index(int(np.argmax(sub)), sub.shape)
    py, px = y0 + dy, x0 + dx
    h = win // 2
    if py - h < 0 or px - h < 0 or py + h > ny or px + h > nx:
        raise ValueError("target too close to the image edge")
    w = img[py - h:py + h, px - h:px + h].astype(np.complex128)
    u = _upsample2d(w, up)
    fy, fx = np.unravel_index(int(np.argmax(np.abs(u))), u.shape)
    rng = _cut_metrics(u[fy, :], up)
    azi = _cut_metrics(u[:, fx], up)
    return {
        "peak_x": px - h + fx / up, "peak_y": py - h + fy / up, "peak": float(np.abs(u[fy, fx])),
        "range": rng, "azimuth": azi,
    }
