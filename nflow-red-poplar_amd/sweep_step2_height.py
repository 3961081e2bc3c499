"""Rates of one and two steps per launch over chunk heights, for refitting the
automatic chunk-height rule of Batch::step2_height on another device.

For each lattice (FP32, MRT-LES Cs 0.17; cavities tau 0.6, channels tau 0.56
with a 0.05 inlet peak) it prints one JSON line: the rate in G node updates
per second at temporal = 0 (one step per launch), -1 (the library's choice)
and each fixed height, as [height used, rate], from the library's HIP events
over 200 steps after 20.

    python nflow-red-poplar_amd/sweep_step2_height.py [--sizes 256x256x256,512x128x128c ...]

A trailing "c" marks a channel."""
import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import nflow  # noqa: E402

DEFAULT_SIZES = ("384x256x256 320x256x256 448x256x256 384x128x128c 640x128x128c 768x128x128c 896x128x128c "
                 "1024x128x128c 256x192x256 512x128x128c 256x256x256 160x160x160 192x192x192 224x224x224 "
                 "128x128x128 96x96x96 112x112x112 144x144x144 176x176x176 208x208x208 240x240x240 288x288x288 "
                 "384x384x384 512x512x512 256x128x128c 128x128x128c 320x128x128c 256x64x256 512x256x512 "
                 "256x96x256 200x300x100 160x64x160 64x64x64 48x48x48").split()
HEIGHTS = (0, -1, 8, 12, 16, 24, 32, 48, 64)


def rate(nx, ny, nz, channel, temporal, steps=200, warm=20):
    case = nflow.CHANNEL if channel else nflow.CAVITY
    inp = nflow.defaults(nx=nx, ny=ny, nz=nz, casetype=case, storage=nflow.F32, enable_mrt=1, enable_les=1,
                         smagorinsky=0.17, tau=0.56 if channel else 0.6, inlet_peak=0.05, perturb=0.01, seed=3)
    b = nflow.Batch([inp])
    try:
        b.set_tuning(512, 0, 0, 0, temporal)
        h = b.step2_height()
        b.initialize()
        b.run(warm)
        b.synchronize()
        b.run(steps)
        ms, _ = b.timing()
    finally:
        b.close()
    return h, round(nx * ny * nz / (ms * 1e-3) / 1e9, 2)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default=",".join(DEFAULT_SIZES))
    args = ap.parse_args()
    for s in args.sizes.split(","):
        channel = s.endswith("c")
        nx, ny, nz = map(int, s.rstrip("c").split("x"))
        out = {"size": f"{nx}x{ny}x{nz}", "case": "channel" if channel else "cavity"}
        for t in HEIGHTS:
            if t <= ny:
                out[str(t)] = list(rate(nx, ny, nz, channel, t))
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
