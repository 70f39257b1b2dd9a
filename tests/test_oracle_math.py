"""Pins the oracle's arithmetic helpers: Philox4x32-10 against the published
known-answer vectors (Salmon et al., SC'11, Random123 kat_vectors) and the
deterministic cos/sin against correctly rounded float64 libm values."""
import numpy as np


def test_philox_known_answers(oracle_mod):
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, expect in kat:
        assert oracle_mod.philox(ctr, key).tolist() == expect


def test_cos_sin_correctly_rounded(oracle_mod):
    rng = np.random.default_rng(0)
    # the sampled azimuths are float(2 pi) * u, u in [0, 1); also cover [0, 4 pi)
    x = np.concatenate([np.float32(2 * np.pi) * rng.random(500000, dtype=np.float32),
                        rng.uniform(0, 4 * np.pi, 200000).astype(np.float32),
                        np.array([0.0, np.pi / 4, np.pi / 2, np.pi, 1.5 * np.pi], np.float32)])
    x = x.astype(np.float32)
    c, s = oracle_mod.cos_sin(x)
    rc = np.cos(x.astype(np.float64)).astype(np.float32)
    rs = np.sin(x.astype(np.float64)).astype(np.float32)
    assert np.array_equal(c, rc)
    assert np.array_equal(s, rs)
