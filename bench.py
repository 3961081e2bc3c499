#!/usr/bin/env python3
"""Benchmark: D3Q19 MRT-LES MLUPS and % of HBM peak on one MI355X.

Default workload (the configuration the headline metric is quoted on):
3-D lid-driven cavity, 256^3 interior nodes, D3Q19 MRT collision with
Smagorinsky LES (Cs = 0.17), tau = 0.6, lid u_x = 0.05, FP32 populations.
One step = one fused stream+collide launch over all 16.8 M nodes.  Inputs are
resident in HBM before the timed region; the timed region is exactly K steps
bracketed by device synchronisation.

Prints ONE JSON line.  Extra workloads: --workload cavity128 | cavity256 |
channel | cavity256_f16 | cavity256_bf16.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nflow-red-poplar_amd"))

METRIC = "D3Q19 MRT-LES MLUPS & % HBM peak; served-batch p50/p99 latency at ≤1e-3 err"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "cavity256": dict(nx=256, ny=256, nz=256, casetype=1, storage=0,
                      desc="3D lid-driven cavity 256^3, D3Q19 MRT + Smagorinsky LES (Cs 0.17), tau 0.6, "
                           "lid u_x 0.05, FP32 populations"),
    "cavity128": dict(nx=128, ny=128, nz=128, casetype=1, storage=0,
                      desc="3D lid-driven cavity 128^3, D3Q19 MRT + Smagorinsky LES (Cs 0.17), tau 0.6, "
                           "lid u_x 0.05, FP32 populations"),
    "cavity256_f16": dict(nx=256, ny=256, nz=256, casetype=1, storage=1,
                          desc="3D lid-driven cavity 256^3, D3Q19 MRT + LES (Cs 0.17), tau 0.6, FP16 storage "
                               "of f - w_q, FP32 arithmetic"),
    "cavity256_bf16": dict(nx=256, ny=256, nz=256, casetype=1, storage=2,
                           desc="3D lid-driven cavity 256^3, D3Q19 MRT + LES (Cs 0.17), tau 0.6, BF16 storage "
                                "of f - w_q, FP32 arithmetic"),
    "channel": dict(nx=512, ny=128, nz=128, casetype=0, storage=0,
                    desc="square-duct channel 512x128x128, D3Q19 MRT + LES (Cs 0.17), tau 0.56, parabolic "
                         "inlet peak 0.05 with seeded 1% perturbation, pressure outlet, FP32 populations"),
}
DTYPE = {0: "f32", 1: "f16-storage/f32-compute", 2: "bf16-storage/f32-compute"}
BYTES_PER_UPDATE = {0: 152, 1: 76, 2: 76}


def make_input(nflow, w):
    tau = 0.56 if w["casetype"] == 0 else 0.6
    return nflow.defaults(nx=w["nx"], ny=w["ny"], nz=w["nz"], casetype=w["casetype"], storage=w["storage"],
                          enable_mrt=1, enable_les=1, smagorinsky=0.17, tau=tau, ref_vel=(0.05, 0.0, 0.0),
                          init_lid_velocity=0, inlet_peak=0.05, perturb=0.01, seed=2024)


def traffic_from_profiles(name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), if recorded."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("bytes_per_launch")


def cpu_baseline(w, budget_s=15.0):
    """The FP64 oracle timed on this host on a bounded sample of the same workload."""
    threads = min(os.cpu_count() or 1, 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    p = oracle.default_params(nx=w["nx"], ny=w["ny"], nz=w["nz"], casetype=w["casetype"], enable_mrt=1,
                              enable_les=1, smagorinsky=0.17, tau=0.56 if w["casetype"] == 0 else 0.6,
                              ref_vel=(0.05, 0.0, 0.0), init_lid_velocity=0, inlet_peak=0.05, perturb=0.01,
                              seed=2024)
    L = oracle.Lattice(p)
    L.initialize()
    t = time.perf_counter()
    L.run(1)
    t1 = time.perf_counter() - t
    steps = max(1, min(20, int(budget_s / max(t1, 1e-3))))
    t = time.perf_counter()
    L.run(steps)
    dt = time.perf_counter() - t
    n = w["nx"] * w["ny"] * w["nz"]
    return {"value": round(n * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": oracle.lib().orc_num_threads(),
            "kind": "port",
            "sample": f"FP64 oracle (C, OpenMP), same workload, {steps} timed steps after 1 warm-up step "
                      f"({dt:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="cavity256", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--autotune", action="store_true", help="autotune the launch shape before timing")
    ap.add_argument("--graph-steps", type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU slab decomposition is "
                                   "a design (DESIGN.md), not code", "n_gpus": args.gpus}))
        sys.exit(2)

    import nflow

    w = WORKLOADS[args.workload]
    name, cus, hbm = nflow.device_info()
    inp = make_input(nflow, w)
    # the single-lattice solver is a batch of one; use the batch handle so the
    # launch shape can be tuned
    b = nflow.Batch([inp])
    if args.autotune:
        tuning = b.autotune(steps_per_trial=10)
        b.set_tuning(tuning["block"], tuning["xcd_order"], tuning["persistent"], args.graph_steps)
    else:
        b.set_tuning(256, 1, 0, args.graph_steps)
    b.initialize()
    b.run(args.warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.run(args.steps)
    b.synchronize()
    wall = time.perf_counter() - t0
    dev_ms, _ = b.timing()
    m, mom, _ = b.diagnostics(0)
    n = w["nx"] * w["ny"] * w["nz"]
    mlups = n * args.steps / wall / 1e6
    bpu = BYTES_PER_UPDATE[w["storage"]]
    achieved = bpu * n / (dev_ms * 1e-3) / 1e9
    copy_gbs, _ = nflow.copy_bandwidth(2 * 19 * 4 * 4_000_000, streams=19, reps=20)
    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[w["storage"]],
        "data": "synthetic (lattice initialised at rest, rho = 1)",
        "config": {"workload": w["desc"], "nx": w["nx"], "ny": w["ny"], "nz": w["nz"], "nodes": n,
                   "tuning": b.get_tuning(), "device": name, "cus": cus},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(args.workload),
                     "bytes_per_update": bpu, "kernel_ms": round(dev_ms, 5)},
        "hbm_frac_of_peak": round(achieved / HBM_PEAK_GBS, 4),
        "copy_ceiling_gbs": round(copy_gbs, 1),
        "mass_drift": abs(m - n) / n,
    }
    b.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
