"""Numerical shortcuts of the device code checked against the operation they replace."""
import numpy as np


def test_schlick_fifth_power_equals_rounded_double_pow():
    """The shading code evaluates Schlick's (1 - cos)^5 as x2 * x2 * x in double (x2 exact for a
    float x) instead of the reference's std::pow(float, int), which computes in double; after
    rounding to float the two agree (here on 4.2M random and near-grazing arguments)."""
    rng = np.random.default_rng(0)
    c = np.concatenate([rng.random(4_000_000, dtype=np.float32),
                        np.float32(1) - rng.random(200_000, dtype=np.float32) * np.float32(1e-3),
                        np.array([0.0, 1.0, 0.5, np.nextafter(np.float32(1), np.float32(0))], np.float32)])
    x = (np.float32(1) - c).astype(np.float64)
    ref = np.power(x, 5.0).astype(np.float32)
    x2 = x * x
    dev = (x2 * x2 * x).astype(np.float32)
    assert np.array_equal(ref, dev)
