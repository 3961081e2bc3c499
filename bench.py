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
import statistics
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nflow-red-poplar_amd"))

METRIC = "D3Q19 MRT-LES MLUPS & % HBM peak; served-batch p50/p99 latency at ≤1e-3 err"
HBM_PEAK_GBS = 8000.0
CAVITY, CHANNEL = 1, 2  # nflow_input.casetype (include/nflow.h)

WORKLOADS = {
    "cavity256": dict(nx=256, ny=256, nz=256, casetype=CAVITY, storage=0,
                      desc="3D lid-driven cavity 256^3, D3Q19 MRT + Smagorinsky LES (Cs 0.17), tau 0.6, "
                           "lid u_x 0.05, FP32 populations"),
    "cavity128": dict(nx=128, ny=128, nz=128, casetype=CAVITY, storage=0,
                      desc="3D lid-driven cavity 128^3, D3Q19 MRT + Smagorinsky LES (Cs 0.17), tau 0.6, "
                           "lid u_x 0.05, FP32 populations"),
    "cavity256_f16": dict(nx=256, ny=256, nz=256, casetype=CAVITY, storage=1, sr=1,
                          desc="3D lid-driven cavity 256^3, D3Q19 MRT + LES (Cs 0.17), tau 0.6, FP16 storage "
                               "of f - w_q with stochastic rounding (the default), FP32 arithmetic"),
    "cavity256_f16rn": dict(nx=256, ny=256, nz=256, casetype=CAVITY, storage=1, sr=0,
                            desc="3D lid-driven cavity 256^3, D3Q19 MRT + LES (Cs 0.17), tau 0.6, FP16 storage "
                                 "of f - w_q rounded to nearest, FP32 arithmetic"),
    "cavity256_bf16": dict(nx=256, ny=256, nz=256, casetype=CAVITY, storage=2, sr=1,
                           desc="3D lid-driven cavity 256^3, D3Q19 MRT + LES (Cs 0.17), tau 0.6, BF16 storage "
                                "of f - w_q with stochastic rounding (the default), FP32 arithmetic"),
    "cavity256_bf16rn": dict(nx=256, ny=256, nz=256, casetype=CAVITY, storage=2, sr=0,
                             desc="3D lid-driven cavity 256^3, D3Q19 MRT + LES (Cs 0.17), tau 0.6, BF16 storage "
                                  "of f - w_q rounded to nearest, FP32 arithmetic"),
    "slab256x32": dict(nx=256, ny=32, nz=256, casetype=CAVITY, storage=0,
                       desc="one GPU's share of a 256^3 cavity cut into 8 y-slabs: 256x32x256, D3Q19 MRT + LES "
                            "(Cs 0.17), tau 0.6, FP32 populations (slab faces run the wall rules)"),
    "slab256x32_f16": dict(nx=256, ny=32, nz=256, casetype=CAVITY, storage=1,
                           desc="one GPU's share of a 256^3 cavity cut into 8 y-slabs: 256x32x256, D3Q19 MRT + "
                                "LES (Cs 0.17), tau 0.6, FP16 storage of f - w_q (slab faces run the wall rules)"),
    "channel": dict(nx=512, ny=128, nz=128, casetype=CHANNEL, storage=0,
                    desc="square-duct channel 512x128x128, D3Q19 MRT + LES (Cs 0.17), tau 0.56, parabolic "
                         "inlet peak 0.05 with seeded 1% perturbation, pressure outlet, FP32 populations"),
}
DTYPE = {0: "f32", 1: "f16-storage/f32-compute", 2: "bf16-storage/f32-compute"}
BYTES_PER_UPDATE = {0: 152, 1: 76, 2: 76}


def make_input(nflow, w):
    tau = 0.56 if w["casetype"] == CHANNEL else 0.6
    p = nflow.defaults(nx=w["nx"], ny=w["ny"], nz=w["nz"], casetype=w["casetype"], storage=w["storage"],
                       enable_mrt=1, enable_les=1, smagorinsky=0.17, tau=tau, ref_vel=(0.05, 0.0, 0.0),
                       init_lid_velocity=0, inlet_peak=0.05, perturb=0.01, seed=2024)
    p.stochastic_rounding = w.get("sr", 1)
    return p


def timed_segments(k):
    """Split k timed steps into runs of even length but the last when k is
    odd: with two steps per launch an even run is whole two-step launches
    (with no read-back between runs the library ends it with a pair) and an
    odd one ends with a single step.  Four near-equal runs from 100 steps on,
    for a median; one below that, since reading a run's device time waits for
    it and leaves the GPU idle until the next run is launched."""
    if k < 100:
        return [k] if k else []
    pairs, odd = divmod(k, 2)
    segs = [2 * (pairs // 4 + (1 if i < pairs % 4 else 0)) for i in range(4)]
    segs[-1] += odd
    return segs


def traffic_from_profiles(name, kernel="k_step"):
    """HBM bytes per launch of the kernel measured by rocprofv3 --pmc, if recorded."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}{'_step2' if kernel == 'k_step2' else ''}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("bytes_per_launch")


def valu_per_node_from_profiles(name, kernel="k_step"):
    """VALU instructions of one launch of the kernel per 64 nodes of the
    lattice (SQ_INSTS_VALU x 64 / nodes) measured by rocprofv3 --pmc, if
    recorded."""
    path = os.path.join(ROOT, "profiles", f"instructions_{name}{'_step2' if kernel == 'k_step2' else ''}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get(kernel, {}).get("valu_per_node")


def accuracy_leg(nflow, w):
    """A fresh lattice of the benchmark workload run for 1,000 and 5,000 steps
    and compared with the FP64 oracle's fixture of the same run (every 8th /
    4th node per axis; tests/golden/make_long_fixtures.py wrote it): rel-L2 of
    u and of rho - 1 at each checkpoint.  Only the 256^3 cavity has one."""
    import numpy as np

    path = os.path.join(ROOT, "tests", "golden", "long_cavity256_tau0.6.npz")
    if not (w["nx"] == w["ny"] == w["nz"] == 256 and w["casetype"] == CAVITY):
        return None
    if not os.path.exists(path):
        return {"note": f"{os.path.relpath(path, ROOT)} not generated"}
    fx = np.load(path, allow_pickle=False)
    cps = {int(k): v for k, v in json.loads(str(fx["case"]))["checkpoints"].items()}
    s = nflow.Solver(make_input(nflow, w))
    s.initialize()
    out = {"reference": "FP64 oracle, " + os.path.relpath(path, ROOT), "steps": [], "rel_l2_u": [],
           "rel_l2_rho_minus_1": []}
    done = 0
    for steps in sorted(cps):
        s.run(steps - done)
        done = steps
        k = cps[steps]
        r, ux, uy, uz = (a[k // 2::k, k // 2::k, k // 2::k].astype(np.float64) for a in s.read_macro())
        uo = np.stack([fx[f"s{steps}_{c}"] for c in ("ux", "uy", "uz")]).astype(np.float64)
        ro = fx[f"s{steps}_rho_minus_1"].astype(np.float64)
        out["steps"].append(steps)
        out["rel_l2_u"].append(float(np.linalg.norm(np.stack([ux, uy, uz]) - uo) / np.linalg.norm(uo)))
        out["rel_l2_rho_minus_1"].append(float(np.linalg.norm(r - 1.0 - ro) / np.linalg.norm(ro)))
    s.close()
    return out


def cpu_model():
    """The host CPU's model name with its family / model / stepping numbers
    (virtual machines often report a bare "AMD EPYC")."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = line.split(":", 1)
                info.setdefault(k.strip(), v.strip())
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    ids = [f"{k} {info[k]}" for k in ("cpu family", "model", "stepping") if k in info]
    return f"{name} ({', '.join(ids)})" if ids else name


def cpu_baseline(w, budget_s=15.0, single_budget_s=8.0):
    """The FP64 oracle timed on this host on a bounded sample of the same
    workload: all worker threads (the reported value) and one thread."""
    threads = min(os.cpu_count() or 1, 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    p = oracle.default_params(nx=w["nx"], ny=w["ny"], nz=w["nz"], casetype=w["casetype"], enable_mrt=1,
                              enable_les=1, smagorinsky=0.17, tau=0.56 if w["casetype"] == CHANNEL else 0.6,
                              ref_vel=(0.05, 0.0, 0.0), init_lid_velocity=0, inlet_peak=0.05, perturb=0.01,
                              seed=2024)
    L = oracle.Lattice(p)
    L.initialize()
    n = w["nx"] * w["ny"] * w["nz"]

    def timed(budget, max_steps):
        t = time.perf_counter()
        L.run(1)
        t1 = time.perf_counter() - t
        steps = max(1, min(max_steps, int(budget / max(t1, 1e-3))))
        t = time.perf_counter()
        L.run(steps)
        return steps, time.perf_counter() - t

    cores = oracle.lib().orc_num_threads()
    steps, dt = timed(budget_s, 60)
    oracle.lib().orc_set_num_threads(1)
    steps1, dt1 = timed(single_budget_s, 20)
    oracle.lib().orc_set_num_threads(cores)
    return {"value": round(n * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle (C, OpenMP), same workload, {steps} timed steps after 1 warm-up step "
                      f"({dt:.1f} s) on {cores} threads; single thread {steps1} steps ({dt1:.1f} s)",
            "single_thread_value": round(n * steps1 / dt1 / 1e6, 3), "cpu_model": cpu_model(),
            "host_cpus": os.cpu_count()}


def solo_run(nflow, inp, steps):
    """One request solved alone: create, initialise, run, diagnostics, destroy.
    Returns the seconds taken and the mass (NaN for a lattice that diverged)."""
    t = time.perf_counter()
    s = nflow.Solver(inp)
    s.initialize()
    s.run(steps)
    try:
        m, _, _ = s.diagnostics()
    except nflow.NflowError as e:
        if e.code != nflow.EDIVERGED:
            raise
        m = float("nan")
    s.close()
    return time.perf_counter() - t, m


def bench_batch(args):
    """64 seeded cavities (edges 32..160, Re 100..5000, SRT or MRT-LES), FP16
    storage, one launch per step over all of them, graph-captured steps."""
    import nflow
    import workloads

    lat = workloads.batched_cavities(seed=args.seed, count=64, storage=nflow.F16)
    inputs = [d["input"] for d in lat]
    b = nflow.Batch(inputs)
    tuning = b.autotune(steps_per_trial=10)
    b.set_tuning(tuning["block"], tuning["xcd_order"], tuning["persistent"], args.graph_steps)
    b.initialize()
    b.run(args.warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.run(args.steps)
    b.synchronize()
    wall = time.perf_counter() - t0
    dev_ms, _ = b.timing()
    nodes = b.nodes()
    diverged = b.diverged()  # the step kernel's instability flags
    b.close()
    # the same lattices one at a time with the single-lattice solver
    solo = {}
    for storage, name in ((nflow.F16, "f16"), (nflow.F32, "f32")):
        tot = 0.0
        for d in lat:
            inp = workloads.cavity_input(d["edge"], d["re"], d["mrt"], storage)
            dt, _ = solo_run(nflow, inp, args.steps + args.warmup)
            tot += dt
        solo[name] = nodes * (args.steps + args.warmup) / tot / 1e6
    mlups = nodes * args.steps / wall / 1e6
    print(json.dumps({
        "metric": "batched D3Q19 aggregate MLUPS (64 cavities, one launch per step)", "value": round(mlups, 1),
        "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4), "higher_is_better": True,
        "dtype": "f16-storage/f32-compute", "data": f"synthetic, seed {args.seed}",
        "config": {"workload": "64 cavities, edges uniform in {32,48,...,160}, Re log-uniform [100,5000], "
                               "nu = U L/Re, U = 0.05, SRT or MRT-LES (Cs 0.17) with p = 1/2",
                   "nodes": nodes, "tuning": tuning, "graph_steps": args.graph_steps,
                   "srt": sum(1 for d in lat if not d["mrt"])},
        "device_ms_per_step": round(dev_ms, 4),
        "roofline": {"bound": "hbm", "achieved": round(76 * nodes / (dev_ms * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(76 * nodes / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "solo_f16_mlups": round(solo["f16"], 1), "solo_f32_mlups": round(solo["f32"], 1),
        "speedup_vs_solo_f32": round(mlups / solo["f32"], 3),
        "diverged_lattices": diverged,
    }))


def serve_stream(seed, requests, step_scale, channel_fraction=0.5, tick=100, stochastic_rounding=True,
                 tau_floor=0.0, oracle_check=True, progress=True):
    """A seeded Poisson request stream served by the continuous-batching loop
    (FP16 storage), against the same requests solved one at a time with the
    plain FP32 solver.  Returns a dict: aggregate MLUPS, p50/p99 latency, the
    served/serial ratio and the velocity errors checked (the four smallest and
    the four lowest-tau 16-bit requests against their FP32 solo runs; the
    smallest cavity also against the FP64 oracle)."""
    import threading

    import numpy as np

    import nflow
    import workloads

    reqs = workloads.request_stream(seed=seed, count=requests, mean_gap_s=0.05, channel_fraction=channel_fraction,
                                    step_scale=step_scale, storage=nflow.F16)
    for r in reqs:
        r["input"].stochastic_rounding = 1 if stochastic_rounding else 0
    _, _, hbm = nflow.device_info()
    budget = min(200 << 30, int(hbm * 0.75))
    srv = nflow.Server(budget, storage=nflow.F16, max_resident=64, steps_per_tick=tick)
    srv.set_tau_floor(tau_floor)
    check = sorted(reqs, key=lambda r: r["nodes"])[:4]
    check += [r for r in sorted(reqs, key=lambda r: r["input"].tau)
              if r["input"].tau >= tau_floor and r not in check][:4]
    check_tags = {r["tag"] for r in check}
    cavities = [r for r in reqs if "cavity" in r["kind"]]
    smallest_cavity = min(cavities, key=lambda r: r["nodes"] * r["steps"]) if cavities else None
    if smallest_cavity is not None:
        check_tags.add(smallest_cavity["tag"])
    t_start = time.perf_counter()

    def submitter():
        for r in reqs:
            delay = r["arrival"] - (time.perf_counter() - t_start)
            if delay > 0:
                time.sleep(delay)
            srv.submit(r["input"], r["steps"], r["tag"], want_fields=r["tag"] in check_tags)

    th = threading.Thread(target=submitter)
    th.start()
    results = {}
    while len(results) < len(reqs):
        for r in srv.poll(64, wait_ms=200):
            results[r.tag] = dict(status=r.status, submit=r.submit_s, admit=r.admit_s, done=r.done_s)
            if progress and len(results) % 8 == 0:
                print(f"served {len(results)}/{len(reqs)}", file=sys.stderr, flush=True)
    th.join()
    wall = max(v["done"] for v in results.values()) - min(v["submit"] for v in results.values())
    st = srv.stats()
    fields = {}
    for r in reqs:
        if r["tag"] in check_tags and results[r["tag"]]["status"] == 0:
            fields[r["tag"]] = srv.fetch_macro(r["tag"], (r["input"].ny, r["input"].nz, r["input"].nx))
    srv.close()
    lat = np.array([v["done"] - v["submit"] for v in results.values()])
    updates = sum(r["nodes"] * r["steps"] for r in reqs)
    ok = sum(1 for v in results.values() if v["status"] == 0)
    diverged = sum(1 for v in results.values() if v["status"] == nflow.EDIVERGED)
    # the same requests, one at a time, plain FP32 solver
    serial = 0.0
    errs = []
    for k, r in enumerate(reqs):
        if progress and k % 8 == 0:
            print(f"serial {k}/{len(reqs)}", file=sys.stderr, flush=True)
        inp = r["input"]
        plain = nflow.defaults(**{k: getattr(inp, k) for k, _ in inp._fields_ if k != "ref_vel"},
                               ref_vel=tuple(inp.ref_vel))
        plain.storage = nflow.F32
        plain.stochastic_rounding = 0
        dt, _ = solo_run(nflow, plain, r["steps"])
        serial += dt
        if r["tag"] in fields:
            s = nflow.Solver(plain)
            s.initialize()
            s.run(r["steps"])
            _, ux, uy, uz = s.read_macro()
            s.close()
            g = fields[r["tag"]]
            u32 = np.stack([ux, uy, uz]).astype(np.float64)
            u16 = np.stack(g[1:]).astype(np.float64)
            errs.append({"request": r["kind"], "tau": round(inp.tau, 5), "steps": r["steps"],
                         "fp32_storage": inp.tau < tau_floor, "reference": "FP32 solo run",
                         "rel_l2_u": float(np.linalg.norm(u16 - u32) / np.linalg.norm(u32))})
    if oracle_check and smallest_cavity is not None and smallest_cavity["tag"] in fields:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        inp = smallest_cavity["input"]
        L = oracle.Lattice(oracle.default_params(
            nx=inp.nx, ny=inp.ny, nz=inp.nz, casetype=inp.casetype, tau=inp.tau, enable_mrt=inp.enable_mrt,
            enable_les=inp.enable_les, smagorinsky=inp.smagorinsky, ref_vel=tuple(inp.ref_vel),
            init_lid_velocity=inp.init_lid_velocity))
        L.initialize()
        L.run(smallest_cavity["steps"])
        _, xo, yo, zo = L.read_macro()
        uo = np.stack([xo, yo, zo])
        u16 = np.stack(fields[smallest_cavity["tag"]][1:]).astype(np.float64)
        errs.append({"request": smallest_cavity["kind"], "tau": round(inp.tau, 5), "steps": smallest_cavity["steps"],
                     "fp32_storage": False, "reference": "FP64 oracle",
                     "rel_l2_u": float(np.linalg.norm(u16 - uo) / np.linalg.norm(uo))})
    mlups = updates / wall / 1e6
    serial_mlups = updates / serial / 1e6
    return {
        "value": round(mlups, 1), "unit": "MLUPS", "dtype": "f16-storage/f32-compute" +
        (", stochastic rounding" if stochastic_rounding else ""),
        "data": f"synthetic Poisson stream, seed {seed}",
        "workload": f"{len(reqs)} requests, Poisson arrivals mean 50 ms; {1 - channel_fraction:g} cavity edge "
                    f"U[48,256] Re logU[100,5000] MRT-LES, {channel_fraction:g} channel length U[256,1024] x 128 "
                    f"x 128 MRT-LES; steps U[500,20000] x {step_scale}; {budget / 2**30:.0f} GiB budget",
        "steps_per_tick": tick,
        "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
        "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
        "completed_ok": ok, "diverged": diverged, "wall_s": round(wall, 3), "node_updates": updates,
        "serial_plain_f32_mlups": round(serial_mlups, 1), "speedup_vs_serial_plain": round(mlups / serial_mlups, 3),
        "peak_resident": st["peak_resident"], "peak_bytes": st["peak_bytes"],
        "server_busy_s": round(st["busy_s"], 3), "server_step_s": round(st["step_s"], 3), "ticks": st["ticks"],
        "step_mlups": round(st["node_updates"] / max(st["step_s"], 1e-9) / 1e6, 1),
        "tau_floor": tau_floor, "fp32_requests": sum(1 for r in reqs if r["input"].tau < tau_floor),
        "rel_l2_u_checked": errs,
        "max_rel_l2_u_checked": max(e["rel_l2_u"] for e in errs) if errs else None,
        "latency_includes": "queueing: requests arrive faster than they are served",
    }


def bench_serve(args):
    """Poisson request stream served by the continuous-batching loop (FP16
    storage), against the same requests solved one at a time with the plain
    FP32 solver."""
    r = serve_stream(args.seed, args.requests, args.step_scale, args.channel_fraction, args.tick,
                     args.rounding == "stochastic", args.tau_floor)
    print(json.dumps({"metric": "served-batch aggregate MLUPS and p50/p99 latency", "n_gpus": 1,
                      "higher_is_better": True, **r}))


def bench_precision(args):
    """Storage precision against the FP64 oracle on the channel and cavity
    configurations: rel-L2 of rho and u per storage format, and MLUPS."""
    import numpy as np

    import nflow

    os.environ.setdefault("OMP_NUM_THREADS", str(min(os.cpu_count() or 1, 16)))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    cases = []
    want = args.cases.split(",")
    if "channel" in want:
        cases.append(("channel 512x128x128 MRT-LES Cs 0.17 tau 0.56, inlet peak 0.05 + 1% seeded", args.channel_steps,
                      dict(nx=512, ny=128, nz=128, casetype=CHANNEL, enable_mrt=1, enable_les=1, smagorinsky=0.17,
                           tau=0.56, inlet_peak=0.05, perturb=0.01, seed=2024, ref_vel=(0.05, 0.0, 0.0),
                           init_lid_velocity=0)))
    if "cavity" in want:
        cases.append((f"cavity 128^3 MRT-LES Cs 0.17 tau {args.cavity_tau:g}, lid 0.05", args.cavity_steps,
                      dict(nx=128, ny=128, nz=128, casetype=CAVITY, enable_mrt=1, enable_les=1, smagorinsky=0.17,
                           tau=args.cavity_tau, ref_vel=(0.05, 0.0, 0.0), init_lid_velocity=0)))

    def oracle_macro(c, steps, fp32):
        L = oracle.Lattice(oracle.default_params(**c), fp32=fp32)
        L.initialize()
        done = 0
        while done < steps:  # in chunks, with a progress line each
            k = min(250, steps - done)
            L.run(k)
            done += k
            print(f"oracle{'-f32' if fp32 else ''} {c['nx']}x{c['ny']}x{c['nz']}: {done}/{steps} steps",
                  file=sys.stderr, flush=True)
        m = [np.asarray(a, dtype=np.float64) for a in L.read_macro()]
        del L
        return m

    def errors(r, u, ro, uo):
        return {"rel_l2_u": float(np.linalg.norm(u - uo) / np.linalg.norm(uo)),
                "rel_l2_rho": float(np.linalg.norm(r - ro) / np.linalg.norm(ro)),
                "rel_l2_rho_minus_1": float(np.linalg.norm(r - ro) / np.linalg.norm(ro - 1.0))}

    table = []
    for name, steps, c in cases:
        ro, xo, yo, zo = oracle_macro(c, steps, False)
        uo = np.stack([xo, yo, zo])
        if args.f32_oracle:  # the FP32 round-off floor: the oracle's own algorithm in FP32
            r3, x3, y3, z3 = oracle_macro(c, steps, True)
            table.append({"case": name, "steps": steps, "storage": "oracle-f32",
                          **errors(r3, np.stack([x3, y3, z3]), ro, uo)})
        for storage, sr, sname in ((nflow.F32, 0, "f32"), (nflow.F16, 1, "f16-shifted stochastic"),
                                   (nflow.F16, 0, "f16-shifted nearest"), (nflow.BF16, 1, "bf16-shifted stochastic"),
                                   (nflow.BF16, 0, "bf16-shifted nearest")):
            inp = nflow.defaults(storage=storage, **c)
            inp.stochastic_rounding = sr
            b = nflow.Batch([inp])
            b.set_tuning(256 if storage else 512, 0, 0, 10)
            b.initialize()
            b.run(steps)
            ms, _ = b.timing()
            r, ux, uy, uz = b.read_macro(0)
            b.close()
            u = np.stack([ux, uy, uz]).astype(np.float64)
            r = r.astype(np.float64)
            table.append({"case": name, "steps": steps, "storage": sname, **errors(r, u, ro, uo),
                          "mlups": round(c["nx"] * c["ny"] * c["nz"] / (ms * 1e-3) / 1e6, 1)})
    print(json.dumps({"metric": "rel-L2 error of rho and u against the FP64 oracle per storage format",
                      "unit": "relative L2", "n_gpus": 1, "data": "synthetic", "results": table}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="cavity256", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="single", choices=("single", "batch", "serve", "precision"))
    ap.add_argument("--channel-steps", type=int, default=200, help="precision mode: channel steps")
    ap.add_argument("--cavity-steps", type=int, default=1000, help="precision mode: cavity steps")
    ap.add_argument("--cavity-tau", type=float, default=0.6, help="precision mode: cavity relaxation time")
    ap.add_argument("--cases", default="channel,cavity", help="precision mode: comma-separated channel,cavity")
    ap.add_argument("--f32-oracle", action="store_true", help="precision mode: also the FP32 build of the oracle")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the comparison with the FP64 fixture")
    ap.add_argument("--no-serve", action="store_true", help="default workload: skip the served-stream leg")
    ap.add_argument("--serve-requests", type=int, default=16, help="served-stream leg: requests")
    ap.add_argument("--serve-step-scale", type=float, default=0.1, help="served-stream leg: step scale")
    ap.add_argument("--serve-seed", type=int, default=7, help="served-stream leg: seed")
    ap.add_argument("--no-autotune", action="store_true", help="fixed launch shape instead of the autotuner's")
    ap.add_argument("--graph-steps", type=int, default=10)
    ap.add_argument("--temporal", type=int, default=None,
                    help="FP32 cavity or channel: two steps per launch, y swept in chunks of this many planes (0: one "
                         "step per launch, -1: the library's choice; default: the autotuner's)")
    ap.add_argument("--block", type=int, default=512, help="workgroup size with --no-autotune")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--requests", type=int, default=48, help="serve mode: length of the request stream")
    ap.add_argument("--step-scale", type=float, default=0.1, help="serve mode: scale of the per-request steps")
    ap.add_argument("--tick", type=int, default=100, help="serve mode: steps per tick")
    ap.add_argument("--channel-fraction", type=float, default=0.5, help="serve mode: share of channel requests")
    ap.add_argument("--rounding", default="stochastic", choices=("stochastic", "nearest"),
                    help="serve mode: rounding of the 16-bit stores of every request")
    ap.add_argument("--tau-floor", type=float, default=0.0,
                    help="serve mode: requests with tau below this are kept in FP32 storage")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU slab decomposition is "
                                   "designed but not built", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.mode == "batch":
        return bench_batch(args)
    if args.mode == "serve":
        return bench_serve(args)
    if args.mode == "precision":
        return bench_precision(args)

    import nflow

    w = WORKLOADS[args.workload]
    name, cus, hbm = nflow.device_info()
    inp = make_input(nflow, w)
    # the single-lattice solver is a batch of one; use the batch handle so the
    # launch shape can be tuned
    b = nflow.Batch([inp])
    tuned_ms = None
    if not args.no_autotune:
        tuning = b.autotune(steps_per_trial=10)
        tuned_ms = tuning["ms_per_step"]
        b.set_tuning(tuning["block"], tuning["xcd_order"], tuning["persistent"], args.graph_steps,
                     tuning["temporal"] if args.temporal is None else args.temporal)
    else:
        b.set_tuning(args.block, 0, 0, args.graph_steps, -1 if args.temporal is None else args.temporal)
    h2 = b.step2_height()
    b.initialize()
    b.run(args.warmup)
    b.synchronize()
    # K timed steps in segments, each timed with HIP events on the solver's
    # stream; the kernel time reported is the median segment's
    segs = timed_segments(args.steps)
    seg_ms = []
    t0 = time.perf_counter()
    for k in segs:
        if k == 0:
            continue
        b.run(k)
        seg_ms.append(b.timing()[0])
    b.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = statistics.median(seg_ms)
    n = w["nx"] * w["ny"] * w["nz"]
    mlups = n * args.steps / wall / 1e6
    bpu = BYTES_PER_UPDATE[w["storage"]]
    if h2 > 0:
        # the dominant kernel is k_step2 (two steps per launch, one read and one
        # write of the populations); its launch time from two runs of 2m + 1
        # steps (m two-step launches and one single step each)
        seg2 = []
        for _ in range(3):
            tot = []
            for steps in (201, 401):
                b.run(steps)
                tot.append(b.timing()[0] * steps)
            seg2.append((tot[1] - tot[0]) / 100)
        kernel = {"name": "k_step2", "steps_per_launch": 2, "launch_ms": statistics.median(seg2),
                  "launch_ms_samples": [round(x, 5) for x in seg2]}
    else:
        kernel = {"name": "k_step", "steps_per_launch": 1, "launch_ms": dev_ms}
    m, mom, _ = b.diagnostics(0)
    # algorithmic bytes per launch: one read and one write of the populations
    # (bpu per node), whether the launch advances one step or two
    bpl = bpu * n
    achieved = bpl / (kernel["launch_ms"] * 1e-3) / 1e9
    # streaming ceilings of the same access width: one array, and 19 arrays
    # side by side (the population streams of a plane-per-population layout)
    copy1_gbs, _ = nflow.copy_bandwidth(2 * 19 * 4 * 4_000_000, streams=1, reps=20)
    copy_gbs, _ = nflow.copy_bandwidth(2 * 19 * 4 * 4_000_000, streams=19, reps=20)
    copy_runs_gbs, _ = nflow.copy_bandwidth(2 * 19 * 4 * 4_000_000, streams=0, reps=20)
    fma_g, _ = nflow.valu_rate(reps=5)
    kn = kernel["name"]
    vpn = valu_per_node_from_profiles(args.workload, kn)
    valu = {"fma_chain_g_wave_per_s": round(fma_g, 1), "kernel": kn, "valu_per_64_nodes": vpn,
            "g_wave_per_s": None if vpn is None else round(vpn * n / 64 / (kernel["launch_ms"] * 1e-3) / 1e9, 1)}
    if vpn is not None:
        valu["over_fma_chain"] = round(valu["g_wave_per_s"] / fma_g, 3)
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
        "dtype": DTYPE[w["storage"]] + ("" if w["storage"] == 0 else
                                        ", stochastic rounding" if w.get("sr", 1) else ", rounded to nearest"),
        "data": "synthetic (lattice initialised at rest, rho = 1)",
        "config": {"workload": w["desc"], "nx": w["nx"], "ny": w["ny"], "nz": w["nz"], "nodes": n,
                   "tuning": dict(b.get_tuning(), ms_per_step=tuned_ms, autotuned=not args.no_autotune, step2_height=h2,
                                  graph_replay=h2 == 0 and args.graph_steps > 0 and min(segs, default=0) >= args.graph_steps),
                   "device": name, "cus": cus},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(args.workload, kn),
                     "kernel": kn, "steps_per_launch": kernel["steps_per_launch"], "bytes_per_launch": bpl,
                     "bytes_per_update": bpu // kernel["steps_per_launch"], "kernel_ms": round(kernel["launch_ms"], 5),
                     "kernel_ms_samples": kernel.get("launch_ms_samples"),
                     "step_ms_segments": [round(x, 5) for x in seg_ms]},
        "hbm_frac_of_peak": round(achieved / HBM_PEAK_GBS, 4),
        "copy_ceiling_gbs": {"1_stream": round(copy1_gbs, 1), "19_streams": round(copy_gbs, 1),
                             "population_runs": round(copy_runs_gbs, 1)},
        "valu_issue": valu,
        "mass_drift": abs(m - n) / n,
    }
    b.close()
    if not args.no_accuracy:
        acc = accuracy_leg(nflow, w)
        if acc is not None:
            out["accuracy"] = acc
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
    if args.workload == "cavity256" and not args.no_serve:
        # the metric's second half: a short served stream at real sizes
        out["serve"] = serve_stream(args.serve_seed, args.serve_requests, args.serve_step_scale, progress=False)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
