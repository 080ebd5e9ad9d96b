"""The oracle's output conditioning against closed forms: saturating truncation, byte order, dB values."""
import numpy as np

import pyoracle as po


def test_slc_records_big_endian_saturating_truncation():
    slc = np.array([[1.5 - 1.5j, 40000 - 40000j, -0.9 + 2.99j]], np.complex64)
    out = po.condition_slc(slc, 3, 1, 1.0, header=5)
    assert out.shape == (1, 5 + 12) and (out[0, :5] == 0).all()
    vals = out[0, 5:].view(">i2").astype(int).tolist()
    assert vals == [1, -1, 32767, -32768, 0, 2]
    assert out[0, 5:7].tolist() == [0x00, 0x01] and out[0, 7:9].tolist() == [0xFF, 0xFF]


def test_detected_records_db_low_byte_first():
    img = np.array([[10.0, 1e3, 1e4, 0.0, -5.0, np.nan, 3.4e38]], np.float32)
    out = po.condition_detected(img, 1.0, header=17)
    assert out.shape == (1, 17 + 14) and (out[0, :17] == 0).all()
    vals = out[0, 17:].view("<u2").astype(int).tolist()
    # 10 log10: 10, 30, 40 dB; zero -> -inf -> 0; negative and NaN -> 0; 3.4e38 -> 385.3 dB -> 385
    assert vals == [10, 30, 40, 0, 0, 0, 385]
    assert out[0, 17:19].tolist() == [10, 0]


def test_statistics_skip_nan_and_divide_by_total():
    img = np.array([[1 + 2j, 3 + 4j], [np.nan + 0j, 5 + 6j]], np.complex64)
    m = po.image_statistics(img, 2, 2)
    # sums over the three finite samples, divided by 4
    assert np.isclose(m[0], 9 / 4) and np.isclose(m[1], 12 / 4)
    re, im = np.array([1, 3, 5.0]), np.array([2, 4, 6.0])
    assert np.isclose(m[2], np.sqrt(np.sum((re - 9 / 4) ** 2) / 4))
    # the NaN sample is skipped whole, so its zero Q component does not enter the Q spread
    assert np.isclose(m[3], np.sqrt(np.sum((im - 12 / 4) ** 2) / 4))
    mi = po.image_statistics(img, 2, 2, imaged=True)
    assert np.isclose(mi[0], (m[0] + m[1]) / 2) and mi[1] == 0 and np.isclose(mi[2], (m[2] + m[3]) / 2)
