"""Seeded synthetic workloads for the batched and served runs.

All draws come from raw splitmix64 outputs (never library distributions), so
a seed gives the same lattices and request stream on any platform.
"""
import math

import nflow

MASK = (1 << 64) - 1


def splitmix64(x):
    z = (x + 0x9E3779B97F4A7C15) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


class Stream:
    """Counter-based uniform draws: the k-th draw is splitmix64(seed + k)."""

    def __init__(self, seed):
        self.seed = seed & MASK
        self.k = 0

    def u64(self):
        self.k += 1
        return splitmix64((self.seed + self.k * 0x632BE59BD9B4E019) & MASK)

    def uniform(self):  # [0, 1)
        return (self.u64() >> 11) * (1.0 / (1 << 53))

    def integer(self, lo, hi):  # inclusive
        return lo + self.u64() % (hi - lo + 1)

    def log_uniform(self, lo, hi):
        return math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * self.uniform())

    def exponential(self, mean):
        return -mean * math.log(1.0 - self.uniform())


def cavity_input(edge, re, mrt, storage, u=0.05, cs=0.17):
    """Lid-driven cavity of edge L at Reynolds number Re: nu = U L / Re, tau = 3 nu + 0.5."""
    nu = u * edge / re
    return nflow.defaults(nx=edge, ny=edge, nz=edge, casetype=nflow.CAVITY, storage=storage,
                          enable_mrt=1 if mrt else 0, enable_les=1 if mrt else 0, smagorinsky=cs if mrt else 0.0,
                          tau=3.0 * nu + 0.5, ref_vel=(u, 0.0, 0.0), init_lid_velocity=0)


def channel_input(length, storage, ny=128, nz=128, tau=0.56, seed=0):
    return nflow.defaults(nx=length, ny=ny, nz=nz, casetype=nflow.CHANNEL, storage=storage, enable_mrt=1,
                          enable_les=1, smagorinsky=0.17, tau=tau, inlet_peak=0.05, perturb=0.01, seed=seed)


def batched_cavities(seed=1, count=64, storage=nflow.F16):
    """64 cavities: edges uniform in {32, 48, ..., 160}, Re log-uniform in
    [100, 5000] (nu = U L / Re, U = 0.05), SRT or MRT-LES with p = 1/2."""
    s = Stream(seed)
    out = []
    for _ in range(count):
        edge = 32 + 16 * s.integer(0, 8)
        re = s.log_uniform(100.0, 5000.0)
        mrt = s.uniform() < 0.5
        out.append(dict(edge=edge, re=re, mrt=mrt, input=cavity_input(edge, re, mrt, storage)))
    return out


def request_stream(seed=7, count=200, mean_gap_s=0.05, channel_fraction=0.5, step_scale=1.0, storage=nflow.F16,
                   edge_range=(48, 256), length_range=(256, 1024), steps_range=(500, 20000)):
    """Poisson arrivals (exponential gaps), a mix of cavities (edge uniform in
    edge_range, Re log-uniform in [100, 5000], MRT-LES) and 128x128 channels
    (length uniform in length_range), steps uniform in steps_range scaled by
    step_scale."""
    s = Stream(seed)
    t = 0.0
    reqs = []
    for i in range(count):
        t += s.exponential(mean_gap_s)
        steps = max(1, int(round(s.integer(*steps_range) * step_scale)))
        if s.uniform() < channel_fraction:
            length = s.integer(*length_range)
            inp = channel_input(length, storage, seed=seed * 1000 + i)
            kind = f"channel {length}x128x128"
        else:
            edge = s.integer(*edge_range)
            re = s.log_uniform(100.0, 5000.0)
            inp = cavity_input(edge, re, True, storage)
            kind = f"cavity {edge}^3 Re {re:.0f}"
        reqs.append(dict(tag=i + 1, arrival=t, steps=steps, input=inp, kind=kind,
                         nodes=inp.nx * inp.ny * inp.nz))
    return reqs
