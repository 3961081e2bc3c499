"""The stochastic rounding of the 16-bit stores, restated on the host.

gfx950's v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32, as measured on the device for
every seed: the magnitude's float32 bits below the 16-bit mantissa (13 bits for
FP16, 16 for BF16) are added to the top 13 / 16 bits of the seed and the sum
is truncated; for an FP16 result below 2^-14 (subnormal) the discarded
fraction of the 2^-24 spacing is taken to 13 bits, floored, and compared the
same way.  `sr_f16` / `sr_bf16` restate that rule; tests/test_gpu_rounding.py
checks the device against them bit for bit.  Here the rule itself is checked
against what stochastic rounding must do, with numpy's float16 grid as the
yardstick: over all seed levels a value lands only on its two grid neighbours
and its mean is the value itself (exactly, for normal values, including values
beside binade edges and exact powers of two of either sign); the seeds the
step kernel draws have uniform top bits.  Pure host arithmetic."""
import numpy as np

SR_WEYL = 0x9E3779B9


def _f32_bits(v):
    return np.asarray(v, dtype=np.float32).view(np.uint32).astype(np.uint64)


def sr_bf16(v, seed):
    """float32 -> bfloat16 value (as float32), stochastic with 32-bit seed(s)."""
    b = _f32_bits(v)
    sign, mag = b & 0x80000000, b & 0x7FFFFFFF
    s = np.asarray(seed, dtype=np.uint64) >> np.uint64(16)
    m = (mag + s) & ~np.uint64(0xFFFF)
    return (m | sign).astype(np.uint32).view(np.float32)


def sr_f16(v, seed):
    """float32 -> float16, stochastic with 32-bit seed(s)."""
    v = np.asarray(v, dtype=np.float32)
    seed = np.broadcast_to(np.asarray(seed, dtype=np.uint64), v.shape)
    out = np.empty(v.shape, dtype=np.float16)
    a = np.abs(v.astype(np.float64))
    sub = a < 2.0 ** -14
    # normal FP16 range: add the top 13 seed bits to the 13 discarded bits
    b = _f32_bits(v)
    m = ((b & 0x7FFFFFFF) + (seed >> np.uint64(19))) & ~np.uint64(0x1FFF)
    with np.errstate(over="ignore"):
        norm = ((m | (b & 0x80000000)).astype(np.uint32).view(np.float32)).astype(np.float16)
    # subnormal range: the 2^-24 grid, discarded fraction floored to 13 bits
    q = a * 2.0 ** 24
    fl = np.floor(q)
    frac13 = np.floor((q - fl) * 8192.0)
    up = (seed >> np.uint64(19)).astype(np.float64) + frac13 >= 8192.0
    subv = np.copysign(np.where(up, fl + 1.0, fl) * 2.0 ** -24, v).astype(np.float16)  # keeps -0.0
    out[...] = np.where(sub, subv, norm)
    return out


def _neighbours16(v):
    h = np.float16(v)
    if np.float64(h) <= np.float64(v):
        return h, np.nextafter(h, np.float16(np.inf))
    return np.nextafter(h, np.float16(-np.inf)), h


def _levels(bits):
    return (np.arange(1 << bits, dtype=np.uint64) << np.uint64(32 - bits)).astype(np.uint32)


def _probe_values(rng):
    edges = [2.0 ** -7, 2.0 ** -5, 2.0 ** -3, 0.5, 2.0 ** -14]
    vals = [0.0, 1.0, -1.0, 0.0078125, -0.0078125, 1.0 / 18, -1.0 / 36, 1.0 / 3]
    for e in edges:  # exact powers of two of both signs, and their float32 neighbours
        for s in (1.0, -1.0):
            p = np.float32(s * e)
            vals += [p, np.nextafter(p, np.float32(0)), np.nextafter(p, np.float32(s * np.inf))]
    vals += list(rng.uniform(-0.3, 0.3, 400)) + list(rng.uniform(-0.02, 0.02, 200))
    return np.array(vals, dtype=np.float32)


def test_fp16_rule_lands_on_the_grid_neighbours_and_is_exactly_unbiased():
    lev = _levels(13)
    vals = _probe_values(np.random.default_rng(0))
    for v in vals[(np.abs(vals) >= 2.0 ** -14) | (vals == 0)]:  # below 2^-14: the subnormal test
        r = sr_f16(np.full(lev.shape, v), lev).astype(np.float64)
        lo, hi = _neighbours16(v)
        assert set(r.tolist()) <= {float(lo), float(hi)}, v
        # exact: the mean over the 8192 levels is v itself (sums of dyadic
        # rationals are exact in float64 here)
        assert r.mean() == np.float64(v), (v, r.mean() - np.float64(v))
        if np.float64(np.float16(v)) == np.float64(v):
            assert np.all(r == np.float64(v)), v  # grid points never move


def test_bf16_rule_lands_on_the_grid_neighbours_and_is_exactly_unbiased():
    lev = _levels(16)
    for v in _probe_values(np.random.default_rng(1))[:150]:
        r = sr_bf16(np.full(lev.shape, v), lev).astype(np.float64)
        b = int(np.float32(v).view(np.uint32))
        lo = np.uint32(b & 0xFFFF0000).view(np.float32)
        hi = np.uint32((b & 0xFFFF0000) + 0x10000).view(np.float32)
        assert set(r.tolist()) <= {float(lo), float(hi)}, v
        assert r.mean() == np.float64(v), (v, r.mean() - np.float64(v))


def test_fp16_subnormal_bias_is_below_a_2_to_13_part_of_the_spacing():
    lev = _levels(13)
    rng = np.random.default_rng(2)
    below_edge = np.nextafter(np.float32(2.0 ** -14), np.float32(0))
    for v in np.concatenate([rng.uniform(-6e-5, 6e-5, 200),
                             [2.0 ** -25, 1.5 * 2.0 ** -24, 3e-8, below_edge, -below_edge]]).astype(np.float32):
        r = sr_f16(np.full(lev.shape, v), lev).astype(np.float64)
        lo, hi = _neighbours16(v)
        assert set(r.tolist()) <= {float(lo), float(hi)}, v
        assert abs(r.mean() - np.float64(v)) <= 2.0 ** -24 / 8192 + 1e-300, (v, r.mean() - v)


def test_signed_zeros_stay_signed():
    z = np.array([0.0, -0.0], dtype=np.float32)
    for seed in (0, 0x42881309, 0xFFFFFFFF):
        assert sr_f16(z, seed).view(np.uint16).tolist() == [0x0000, 0x8000]
        assert (sr_bf16(z, seed).view(np.uint32) >> 16).tolist() == [0x0000, 0x8000]


def test_extreme_seeds_truncate_or_round_away_from_zero():
    v = np.random.default_rng(3).uniform(-0.1, 0.1, 1000).astype(np.float32)
    v = v[np.abs(v) >= 2.0 ** -14]
    lo_mag = np.array([min(_neighbours16(x), key=lambda h: abs(float(h))) for x in v], dtype=np.float16)
    hi_mag = np.array([max(_neighbours16(x), key=lambda h: abs(float(h))) for x in v], dtype=np.float16)
    exact = v.astype(np.float16).astype(np.float32) == v
    # seed 0 adds nothing: truncation toward zero
    assert np.array_equal(sr_f16(v, np.zeros(v.shape, np.uint32)), lo_mag)
    # the largest seed rounds every inexact value away from zero
    top = sr_f16(v, np.full(v.shape, 0xFFFFFFFF, np.uint32))
    assert np.array_equal(top[~exact], hi_mag[~exact]) and np.array_equal(top[exact], v[exact].astype(np.float16))


def test_step_seeds_have_uniform_top_bits():
    # seed of store q = hash + q * 0x9E3779B9: for uniformly distributed hashes
    # the top 13 bits are uniform for every q
    rng = np.random.default_rng(4)
    h = rng.integers(0, 1 << 32, 200_000, dtype=np.uint64)
    for q in (0, 1, 7, 18):
        top = ((h + np.uint64(q * SR_WEYL)) & np.uint64(0xFFFFFFFF)) >> np.uint64(19)
        counts = np.bincount(top.astype(np.int64), minlength=8192)
        # chi-square with 8191 degrees of freedom: mean 8191, sd ~128
        e = len(h) / 8192
        chi2 = float(((counts - e) ** 2 / e).sum())
        assert abs(chi2 - 8191) < 6 * 128, (q, chi2)
