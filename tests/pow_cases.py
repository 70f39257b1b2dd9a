"""Inputs for the checks of the Metal lobe's deterministic pow."""
import numpy as np


def pow_inputs(rng):
    """Inputs of the Metal lobe's pow: the sampler's xi^(1/s) (xi in [0, 1)),
    the density's nh^s (nh in [-1, 1], s integral and not), and general
    positive bases down to subnormals with exponents of either sign."""
    n = 300000
    s = np.exp(rng.uniform(np.log(1e-2), np.log(1e6), n)).astype(np.float32)
    xi = rng.random(n, dtype=np.float32)
    nh = rng.uniform(-1, 1, n).astype(np.float32)
    si = np.round(rng.uniform(0, 200, n)).astype(np.float32)
    xg = np.exp2(rng.uniform(-149, 128, n)).astype(np.float32)
    yg = rng.uniform(-40, 40, n).astype(np.float32)
    x = np.concatenate([xi, nh, nh, xg])
    y = np.concatenate([(np.float32(1) / s).astype(np.float32), s, si, yg])
    return x.astype(np.float32), y.astype(np.float32)


def pow_special():
    v = [0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 2.0, -2.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45,
         3.4e38, 7.0, -3.0, 1.5, 2.5]
    x, y = np.meshgrid(np.array(v, np.float32), np.array(v, np.float32))
    return x.ravel().astype(np.float32), y.ravel().astype(np.float32)
