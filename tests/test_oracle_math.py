"""Pins the oracle's arithmetic helpers: Philox4x32-10 against the published
known-answer vectors (Salmon et al., SC'11, Random123 kat_vectors) and the
deterministic cos/sin against correctly rounded float64 libm values."""
import numpy as np

from pow_cases import pow_inputs as _pow_inputs, pow_special as _pow_special


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


def test_pow_matches_float64_pow_rounded(oracle_mod):
    """The deterministic pow equals float64 pow rounded to float wherever the
    float64 value is not within 1e-11 (relative) of a float rounding midpoint,
    and is within one float ulp elsewhere (its own double evaluation errs by
    less than 1e-13 relative)."""
    rng = np.random.default_rng(7)
    x, y = _pow_inputs(rng)
    ours = oracle_mod.pow32(x, y)
    with np.errstate(all="ignore"):
        v = np.power(x.astype(np.float64), y.astype(np.float64))
        r = v.astype(np.float32)
    fin = np.isfinite(v) & np.isfinite(r) & (r != 0)
    up = np.nextafter(r, np.float32(np.inf)).astype(np.float64)
    dn = np.nextafter(r, np.float32(-np.inf)).astype(np.float64)
    rd = r.astype(np.float64)
    with np.errstate(all="ignore"):
        dist = np.minimum(np.abs(v - (rd + up) / 2), np.abs(v - (rd + dn) / 2))
    clear = fin & (dist > 1e-11 * np.abs(v))
    assert clear.sum() > 0.9 * fin.sum()
    assert np.array_equal(ours[clear], r[clear])
    amb = fin & ~clear
    assert np.all(np.abs(ours[amb].view(np.int32).astype(np.int64) -
                         r[amb].view(np.int32).astype(np.int64)) <= 1)
    # zeros, infinities and NaN as float64 pow gives them
    rest = ~fin
    assert np.array_equal(np.isnan(ours[rest]), np.isnan(r[rest]))
    ok = rest & ~np.isnan(r)
    assert np.array_equal(ours[ok], r[ok])


def test_pow_special_cases_follow_powf(oracle_mod):
    """Zeros, signed zeros, infinities, NaN and negative bases as C's pow
    (float64 pow rounded to float): pow(x, 0) = pow(1, y) = 1, odd integral
    powers keep the sign, non-integral powers of finite negative bases are
    NaN, pow(-inf, y) is +-inf or +-0."""
    x, y = _pow_special()
    ours = oracle_mod.pow32(x, y)
    with np.errstate(all="ignore"):
        ref = np.power(x.astype(np.float64), y.astype(np.float64)).astype(np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(ours), nan)
    assert np.array_equal(ours[~nan].view(np.uint32), ref[~nan].view(np.uint32))
