"""Full-length runs of the benchmark configurations against the FP64 oracle.

The oracle needs hours for these (256^3 x 5000 steps, 512 x 128 x 128 x
10000 steps), so tests/golden/make_long_fixtures.py ran it once and saved
rho - 1 and u on a regular subsample of the box (every k-th node per axis);
the GPU runs the full length here (seconds) and is compared on the same nodes.

Bars (rel-L2 of u over the sampled nodes):
  FP32 storage: within 4x the float32 oracle's own drift where that fixture
    exists (128^3), otherwise below 1e-4 (five times the largest FP32 figure
    measured on these runs, 1.9e-5);
  FP16 storage with stochastic rounding: <= 1e-3, the precision target;
  FP16 rounding to nearest and BF16: reported, not asserted (to nearest,
    FP16 exceeds 1e-3 on the channel at 10000 steps)."""
import json
import os

import numpy as np
import pytest

import nflow

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FP32_FACTOR = 4.0
TOL_U_F32_ABS = 1e-4
TOL_U_F16 = 1e-3


def fixture(name):
    path = os.path.join(GOLDEN, f"long_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"{os.path.basename(path)} not generated yet (tests/golden/make_long_fixtures.py {name})")
    f = np.load(path, allow_pickle=False)
    return {k: f[k] for k in f.files}, json.loads(str(f["case"]))


def sample(a, k):
    return a[k // 2::k, k // 2::k, k // 2::k]


def run(params, storage, sr, checkpoints):
    """rho - 1 and u of a GPU run at each checkpoint step (subsampled as the fixture)."""
    kw = dict(params)
    kw["ref_vel"] = tuple(kw.get("ref_vel", (0.05, 0.0, 0.0)))
    inp = nflow.defaults(storage=storage, **kw)
    inp.stochastic_rounding = sr
    s = nflow.Solver(inp)
    s.initialize()
    out = {}
    done = 0
    for steps in sorted(checkpoints):
        s.run(steps - done)
        done = steps
        r, ux, uy, uz = s.read_macro()
        k = checkpoints[steps]
        out[steps] = (sample(r, k) - 1.0, [sample(a, k) for a in (ux, uy, uz)])
    s.close()
    return out


def errors(got, fx, steps):
    rm1, u = got
    uo = np.stack([fx[f"s{steps}_{c}"] for c in ("ux", "uy", "uz")]).astype(np.float64)
    ro = fx[f"s{steps}_rho_minus_1"].astype(np.float64)
    eu = float(np.linalg.norm(np.stack(u).astype(np.float64) - uo) / np.linalg.norm(uo))
    er = float(np.linalg.norm(rm1.astype(np.float64) - ro) / np.linalg.norm(ro))
    return eu, er


VARIANTS = [("FP32", nflow.F32, 0), ("FP16 stochastic", nflow.F16, 1), ("FP16 nearest", nflow.F16, 0),
            ("BF16 stochastic", nflow.BF16, 1), ("BF16 nearest", nflow.BF16, 0)]


@pytest.mark.parametrize("name", ["cavity128_tau0.51", "cavity256_tau0.6", "channel512x128x128"])
def test_full_length_against_the_fp64_oracle(name):
    fx, case = fixture(name)
    cps = {int(k): v for k, v in case["checkpoints"].items()}
    floor = None
    if os.path.exists(os.path.join(GOLDEN, f"long_{name}_f32.npz")):
        f32, _ = fixture(name + "_f32")
        last = max(cps)
        k = cps[last]
        floor = errors((f32[f"s{last}_rho_minus_1"], [f32[f"s{last}_{c}"] for c in ("ux", "uy", "uz")]), fx, last)
        print(f"{name}: float32 oracle after {last} steps: rel-L2 u {floor[0]:.2e}, rho-1 {floor[1]:.2e}")
    res = {}
    for label, storage, sr in VARIANTS:
        got = run(case["params"], storage, sr, cps)
        for steps in sorted(cps):
            res[(label, steps)] = errors(got[steps], fx, steps)
            print(f"{name} {label:16s} {steps:6d} steps: rel-L2 u {res[(label, steps)][0]:.2e}, "
                  f"rho-1 {res[(label, steps)][1]:.2e}")
    last = max(cps)
    eu32 = res[("FP32", last)][0]
    if floor is not None:
        assert eu32 <= FP32_FACTOR * floor[0] + 2e-7, (eu32, floor)
    else:
        assert eu32 <= TOL_U_F32_ABS, eu32
    for steps in cps:
        assert res[("FP16 stochastic", steps)][0] <= TOL_U_F16, (steps, res[("FP16 stochastic", steps)])
