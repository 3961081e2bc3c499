#!/usr/bin/env python3
"""Writes the full-length FP64-oracle fixtures of the GPU accuracy tests
(tests/test_gpu_long.py): the 256^3 MRT-LES cavity at 5000 steps, the
128^3 cavity at tau 0.51 over 5000 steps and the 512 x 128 x 128 channel at
10000 steps, each as rho - 1 and u on a regular subsample of the box; and
the 128^3 case once more with the oracle's algorithm in float32 (the
round-off yardstick of the FP32 parity bar).

The oracle (oracle/liboracle.so, FP64, OpenMP) needs hours for the two large
cases, far longer than a GPU test may take, so it runs here once and the GPU
tests compare against what it saved.  Output: tests/golden/long_<case>.npz
with, per checkpoint step s, the arrays s<steps>_rho_minus_1, _ux, _uy, _uz
(float32, [ny/k][nz/k][nx/k], every k-th node per axis starting at k/2) and
s<steps>_diag (mass, momentum x/y/z, sum |rho - rho_prev| of the whole box,
FP64), plus a JSON description of the case.

Usage: python tests/golden/make_long_fixtures.py [case ...]  (default: all)
       OMP_NUM_THREADS sets the oracle's worker count.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402

MRT_LES = dict(enable_mrt=1, enable_les=1, smagorinsky=0.17)
CASES = {
    # the benchmark workload at its full length: C2 at 256^3, tau 0.6
    "cavity256_tau0.6": dict(
        params=dict(nx=256, ny=256, nz=256, casetype=oracle.CAVITY, tau=0.6, ref_vel=(0.05, 0.0, 0.0),
                    init_lid_velocity=0, **MRT_LES),
        checkpoints={1000: 8, 5000: 4}),
    # C2 at the low end of its tau range
    "cavity128_tau0.51": dict(
        params=dict(nx=128, ny=128, nz=128, casetype=oracle.CAVITY, tau=0.51, ref_vel=(0.05, 0.0, 0.0),
                    init_lid_velocity=0, **MRT_LES),
        checkpoints={5000: 4}),
    # the same case through the float32 build of the oracle (liboracle_f32.so)
    "cavity128_tau0.51_f32": dict(
        params=dict(nx=128, ny=128, nz=128, casetype=oracle.CAVITY, tau=0.51, ref_vel=(0.05, 0.0, 0.0),
                    init_lid_velocity=0, **MRT_LES),
        checkpoints={5000: 4}, fp32=True),
    # C3: square duct, parabolic inlet with a seeded 1% perturbation, outlet at rho = 1
    "channel512x128x128": dict(
        params=dict(nx=512, ny=128, nz=128, casetype=oracle.CHANNEL, tau=0.56, inlet_peak=0.05, perturb=0.01,
                    seed=2024, **MRT_LES),
        checkpoints={2000: 8, 10000: 4}),
}


def subsample(a, k):
    return np.ascontiguousarray(a[k // 2::k, k // 2::k, k // 2::k])


def make(name):
    c = CASES[name]
    p = oracle.default_params(**c["params"])
    fp32 = c.get("fp32", False)
    L = oracle.Lattice(p, fp32=fp32)
    L.initialize()
    out = {"case": np.array(json.dumps({"name": name, "params": c["params"],
                                        "checkpoints": {str(k): v for k, v in c["checkpoints"].items()},
                                        "sample": "every k-th node per axis from k/2, [y][z][x]",
                                        "oracle": "oracle/liboracle_f32.so (FP32)" if fp32
                                        else "oracle/liboracle.so (FP64)"}))}
    t0 = time.time()
    done = 0
    for steps in sorted(c["checkpoints"]):
        while done < steps:
            k = min(100, steps - done)
            L.run(k)
            done += k
            el = time.time() - t0
            print(f"{name}: {done}/{max(c['checkpoints'])} steps, {el:.0f} s", flush=True)
        k = c["checkpoints"][steps]
        r, ux, uy, uz = L.read_macro()
        m, mom, d = L.diagnostics()
        if not np.isfinite(m):
            raise RuntimeError(f"{name}: oracle diverged by step {steps}")
        pre = f"s{steps}"
        out[pre + "_rho_minus_1"] = subsample(r - 1.0, k).astype(np.float32)
        for nm, a in (("ux", ux), ("uy", uy), ("uz", uz)):
            out[f"{pre}_{nm}"] = subsample(a, k).astype(np.float32)
        out[pre + "_diag"] = np.array([m, mom[0], mom[1], mom[2], d])
    path = os.path.join(HERE, f"long_{name}.npz")
    np.savez(path + ".tmp.npz", **out)
    os.replace(path + ".tmp.npz", path)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.1f} MB) in {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES)
    for n in names:
        if n not in CASES:
            sys.exit(f"unknown case {n}; known: {', '.join(CASES)}")
    for n in names:
        make(n)
