"""The level-0 sample decoders of the oracle against the reference's own known-answer test and hand-built records.

The FBAQ-4 index table below is the one envisat_format/unit-test/envisat_parse_util_test.cc:24-74 spells out code
by code (sign-magnitude order: 1111 -> 0 ... 1000 -> 7, 0000 -> 8 ... 0111 -> 15)."""
import numpy as np

import pyoracle as po

KAT = {0b1111: 0, 0b1110: 1, 0b1101: 2, 0b1100: 3, 0b1011: 4, 0b1010: 5, 0b1001: 6, 0b1000: 7,
       0b0000: 8, 0b0001: 9, 0b0010: 10, 0b0011: 11, 0b0100: 12, 0b0101: 13, 0b0110: 14, 0b0111: 15}


def test_fbaq4_index_matches_reference_table():
    for block in (0, 1, 77, 255):
        for code, idx in KAT.items():
            assert po.fbaq4_index(block, code) == 256 * idx + block


def test_ers_decode_places_each_line_at_its_swst_shift():
    samples = np.array([[1, 2, 3, 4, 5, 6, 0, 0], [7, 8, 9, 10, 11, 12, 0, 0]], np.uint8)  # 3 pairs, 8-byte lines
    swst = np.array([5, 6], np.uint16)
    out = po.ingest_ers_im(samples, 3, swst, 5, range_size=9, x_stride=12, y_stride=4)
    assert out.shape == (4, 12)
    # line 0: no shift; line 1: shifted by 4 columns (ERS SWST step)
    np.testing.assert_array_equal(out[0, :3], [1 + 2j, 3 + 4j, 5 + 6j])
    np.testing.assert_array_equal(out[1, 4:7], [7 + 8j, 9 + 10j, 11 + 12j])
    assert np.isnan(out[0, 3:9]).all() and np.isnan(out[1, :4]).all() and np.isnan(out[1, 7:9]).all()
    assert (out[:2, 9:] == 0).all() and (out[2:] == 0).all()


def test_asar_fbaq4_decode_of_a_hand_built_record():
    lut_i = np.arange(4096, dtype=np.float32)
    lut_q = -np.arange(4096, dtype=np.float32)
    # one record: L = 66 bytes = block (id 3, 63 codes) + block (id 9, 1 code); record buffer 80 bytes
    rec = np.zeros(80, np.uint8)
    rec[0], rec[1] = 66, 0
    rec[2] = 3
    rec[3:66] = np.arange(63, dtype=np.uint8) * 4 + 0x1F  # I code = high nibble, Q code = low nibble
    rec[66] = 9
    rec[67] = 0x70
    out = po.ingest_asar_fbaq4(rec[None, :], np.array([2], np.uint16), 1, lut_i, lut_q, range_size=70, x_stride=72,
                               y_stride=2)
    # sample s of the record sits at column s + 1 (SWST 2 - min 1)
    for s in range(63):
        code = int(rec[3 + s])
        want = po.fbaq4_index(3, code >> 4) - 1j * po.fbaq4_index(3, code & 15)
        assert out[0, s + 1] == want
    assert out[0, 64] == po.fbaq4_index(9, 7) - 1j * po.fbaq4_index(9, 0)
    assert np.isnan(out[0, 0]) and np.isnan(out[0, 65:70]).all()
    assert (out[0, 70:] == 0).all() and (out[1] == 0).all()
