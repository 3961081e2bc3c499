"""The stochastic rounding of the 16-bit stores, restated with numpy: for a
value v, the 256 dither levels the hash can pick round v to one of its two
neighbours on the FP16 grid, and their mean is v to within 1/512 of the grid
spacing (unbiased rounding), for normal and subnormal FP16 values of either
sign.  The spacing computed from the float exponent bits must be numpy's FP16
spacing.  Pure host arithmetic."""
import numpy as np


def spacing_bits(v, mant=10, min_exp=0x38800000):
    b = np.float32(v).view(np.uint32)
    eb = max(int(b) & 0x7F800000, min_exp)
    return np.uint32(eb - (mant << 23)).view(np.float32)


def stochastic(v, d):
    """v + (d + 0.5)/256 spacing - spacing/2, rounded to nearest FP16 (d = 0..255)."""
    v = np.float32(v)
    ulp = spacing_bits(v)
    c = np.float32(np.float64(v) - 0.5 * np.float64(ulp))
    y = np.float32(np.float64(c) + (d + 0.5) * np.float64(ulp) / 256.0)
    return np.float16(y)


def test_spacing_is_the_fp16_grid_spacing():
    rng = np.random.default_rng(0)
    for v in np.concatenate([rng.uniform(-0.3, 0.3, 200), rng.uniform(-1e-4, 1e-4, 200), [1e-7, -3e-6, 6.2e-5]]):
        h = np.float16(v)
        if h == 0:
            continue
        assert spacing_bits(v) == np.float32(np.spacing(np.abs(h))), v


def test_dither_levels_round_to_the_neighbours_without_bias():
    rng = np.random.default_rng(1)
    for v in np.concatenate([rng.uniform(-0.3, 0.3, 300), rng.uniform(-2e-5, 2e-5, 100)]).astype(np.float32):
        r = np.array([stochastic(v, d) for d in range(256)], dtype=np.float64)
        h = np.float16(v)
        if np.float64(h) <= v:
            below, above = h, np.nextafter(h, np.float16(np.inf))
        else:
            below, above = np.nextafter(h, np.float16(-np.inf)), h
        assert set(r.tolist()) <= {float(below), float(above)}, v
        ulp = np.float64(spacing_bits(v))
        assert abs(r.mean() - np.float64(v)) <= ulp / 512 + 1e-12, (v, r.mean() - v)
