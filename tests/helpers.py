"""Shared helpers: run the same configuration through the HIP library (C ABI)
and through the FP64 oracle, and compare."""
import numpy as np

import nflow
import oracle

# FP32 parity is measured against the FP64 oracle, with the allowance set by
# round-off itself: the oracle's own algorithm evaluated in float
# (liboracle_f32.so) drifts from FP64 by err32; the GPU, which evaluates the
# same maths in float in another order (fused multiply-adds, the orthogonal
# form of the MRT transform), must stay within FP32_FACTOR * err32 + floor.
FP32_FACTOR = 4.0
FLOOR_U = 2e-7
FLOOR_RHO = 2e-6
# FP16 / BF16 storage of f - w with FP32 arithmetic: absolute bound on the
# relative L2 velocity error against the FP64 oracle
TOL_U_F16 = 1e-3


def cfg(**kw):
    """Keyword set accepted by both nflow.defaults() and oracle.default_params()."""
    base = dict(nx=16, ny=16, nz=16, tau=0.6, smagorinsky=0.0, enable_les=0, enable_mrt=1,
                ref_vel=(0.05, 0.0, 0.0), init_lid_velocity=0, casetype=1)
    base.update(kw)
    return base


def gpu_input(c, storage=nflow.F32):
    return nflow.defaults(storage=storage, **c)


def oracle_params(c):
    return oracle.default_params(**c)


def run_oracle(c, steps, slices=None, fp32=False, mask=None):
    L = oracle.Lattice(oracle_params(c), slices, fp32=fp32)
    if mask is not None:
        L.set_mask(mask)
    L.initialize()
    L.run(steps)
    return L


def run_gpu(c, steps, storage=nflow.F32, slices=None, mask=None):
    s = nflow.Solver(gpu_input(c, storage), slices)
    if mask is not None:
        s.set_solid_mask(mask)
    s.initialize()
    if steps:
        s.run(steps)
    return s


def rel_l2(a, b, shift=0.0):
    a = np.asarray(a, dtype=np.float64) - shift
    b = np.asarray(b, dtype=np.float64) - shift
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def macro_errors(gpu_macro, orc_macro):
    r, ux, uy, uz = gpu_macro
    ro, xo, yo, zo = orc_macro
    eu = rel_l2(np.stack([ux, uy, uz]), np.stack([xo, yo, zo]))
    er = rel_l2(r, ro, shift=1.0)
    return eu, er


def fp32_parity(c, steps, gpu=None, mask=None):
    """Run GPU (FP32), FP64 oracle and FP32 oracle; return a dict of errors
    and assert the GPU is within the round-off allowance."""
    s = gpu if gpu is not None else run_gpu(c, steps, mask=mask)
    o64 = run_oracle(c, steps, mask=mask)
    o32 = run_oracle(c, steps, fp32=True, mask=mask)
    m64 = o64.read_macro()
    eu, er = macro_errors(s.read_macro(), m64)
    eu32, er32 = macro_errors(o32.read_macro(), m64)
    res = dict(eu=eu, er=er, eu32=eu32, er32=er32)
    print(f"rel-L2 vs FP64: GPU u {eu:.2e} rho-1 {er:.2e} | FP32 oracle u {eu32:.2e} rho-1 {er32:.2e}")
    assert eu <= FP32_FACTOR * eu32 + FLOOR_U, res
    assert er <= FP32_FACTOR * er32 + FLOOR_RHO, res
    return res, s, o64, o32
