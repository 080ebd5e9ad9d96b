"""Synthetic scene generator: Philox-4x32-10 known-answer vectors (Salmon et al., SC'11, Random123 KAT), and the
scene's determinism and quantisation."""
import numpy as np

from asarf import configs, scene


def test_philox_known_answers():
    assert scene.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert scene.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                                      0x6D5451FD]
    assert scene.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_scene_is_deterministic_and_quantised():
    cfg = configs.ers_block().scaled(na=256, nr=900, n_targets_az=1, n_targets_rg=1)
    a, fa = scene.generate(cfg, want_float=True)
    b, _ = scene.generate(cfg)
    assert np.array_equal(a, b)
    assert a.dtype == np.uint8 and a.max() <= 31
    f = a.astype(np.float64) - 15.5
    assert 4.0 < f[..., 0].std() < 6.0  # std(I) about 5 LSB
    c = configs.asar_scene().scaled(na=128, nr=900, n_targets_random=3)
    r, _ = scene.generate(c)
    assert r.dtype == np.int8 and 10.0 < r.astype(float)[..., 0].std() < 14.0


def test_targets_on_grid():
    tx, ty = scene.targets(configs.ers_block())
    assert len(tx) == 25
    assert tx.min() == round(0.15 * 4912) and tx.max() == round(0.75 * 4912)
    assert ty.min() == 1024 and ty.max() == 3072
