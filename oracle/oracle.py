"""ctypes access to the FP64 oracle (liboracle.so) and to oracle/_ref/libref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference that GPU results are compared
with (and, in bench.py, timed beside).  The product library never imports this
module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
Q = 19
PIPE, CAVITY, CHANNEL = 0, 1, 2  # POISEUILLE = 0 and CAVITY = 1 as in common.h:26-27
NODE_FLUID, NODE_FIXEDWALL, NODE_MOVINGWALL = 1 << 1, 1 << 2, 1 << 3
NODE_INLET, NODE_OUTLET, NODE_STRUCT = 1 << 4, 1 << 5, 1 << 6


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("casetype", ctypes.c_int), ("enable_mrt", ctypes.c_int), ("enable_les", ctypes.c_int),
        ("init_lid_velocity", ctypes.c_int),
        ("tau", ctypes.c_double), ("smagorinsky", ctypes.c_double), ("ref_den", ctypes.c_double),
        ("ref_vel", ctypes.c_double * 3),
        ("w_eps", ctypes.c_double), ("w_epsj", ctypes.c_double), ("w_xx", ctypes.c_double),
        ("inlet_peak", ctypes.c_double), ("perturb", ctypes.c_double),
        ("seed", ctypes.c_ulonglong),
        ("s_rates", ctypes.c_double * Q),
        ("pipe_radius", ctypes.c_double), ("pipe_cy", ctypes.c_double), ("pipe_cz", ctypes.c_double),
    ]


_libs = {}
_ref = None


def lib(fp32=False):
    """liboracle.so (FP64), or with fp32=True liboracle_f32.so: the same
    algorithm evaluated in single precision, used to measure FP32 round-off."""
    name = "liboracle_f32.so" if fp32 else "liboracle.so"
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        vp = ctypes.c_void_p
        L.orc_default_params.argtypes = [ctypes.POINTER(Params)]
        L.orc_tables.argtypes = [dp, ip, dp]
        L.orc_mrt_matrix.argtypes = [dp]
        L.orc_mrt_inverse_exact.argtypes = [dp, dp]
        L.orc_mrt_inverse_exact.restype = ctypes.c_int
        L.orc_decomposition.argtypes = [ctypes.c_int] * 6 + [ip, ctypes.c_int]
        L.orc_decomposition.restype = ctypes.c_int
        L.orc_slice_chunk.argtypes = [ctypes.c_int, ip, ip]
        L.orc_slice_chunk.restype = ctypes.c_int
        L.orc_set_move_decomp.argtypes = [ctypes.c_int, ip, ip]
        L.orc_set_move_decomp.restype = ctypes.c_int
        L.orc_splitmix64.argtypes = [ctypes.c_ulonglong]
        L.orc_splitmix64.restype = ctypes.c_ulonglong
        L.orc_inlet_xi.argtypes = [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int]
        L.orc_inlet_xi.restype = ctypes.c_double
        L.orc_inlet_profile.argtypes = [ctypes.POINTER(Params), dp]
        L.orc_cell_feq.argtypes = [dp, dp, ctypes.c_double, dp, dp]
        L.orc_cell_tau_eddy.argtypes = [ctypes.POINTER(Params), dp, ctypes.c_double, dp]
        L.orc_cell_tau_eddy.restype = ctypes.c_double
        for fn in (L.orc_cell_mrt, L.orc_cell_srt):
            fn.argtypes = [ctypes.POINTER(Params), dp, ctypes.c_double, dp, ctypes.c_double, dp]
        L.orc_create.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ip]
        L.orc_create.restype = vp
        L.orc_destroy.argtypes = [vp]
        L.orc_set_mask.argtypes = [vp, vp]
        L.orc_initialize.argtypes = [vp]
        L.orc_step.argtypes = [vp]
        L.orc_run.argtypes = [vp, ctypes.c_int]
        L.orc_steps.argtypes = [vp]
        L.orc_steps.restype = ctypes.c_long
        L.orc_read_arF_padded.argtypes = [vp, dp]
        L.orc_write_arF_padded.argtypes = [vp, dp]
        L.orc_read_macro.argtypes = [vp, dp, dp, dp, dp]
        L.orc_read_node_type.argtypes = [vp, ip, ip]
        L.orc_diagnostics.argtypes = [vp, dp, dp, dp]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        L.orc_precision_bytes.restype = ctypes.c_int
        _libs[name] = L
    return _libs[name]


def ref():
    """The reference's own decomposition.cpp / inverseMRT.cpp, or None if not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        ip = ctypes.POINTER(ctypes.c_int)
        fp = ctypes.POINTER(ctypes.c_float)
        R.ref_decomp_counts.argtypes = [ip, ip, ip]
        R.ref_decomposition.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ctypes.c_int]
        R.ref_decomposition.restype = ctypes.c_int
        R.ref_set_move_decomp.argtypes = [ctypes.c_int, ip, ip]
        R.ref_set_move_decomp.restype = ctypes.c_int
        R.ref_slice_chunk.argtypes = [ctypes.c_int, ip, ip]
        R.ref_slice_chunk.restype = ctypes.c_int
        R.ref_inverse_mrt.argtypes = [fp, fp]
        _ref = R
    return _ref


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def default_params(**kw):
    p = Params()
    lib().orc_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k in ("ref_vel", "s_rates"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


def tables():
    d = np.zeros((3, Q))
    opp = np.zeros(Q, dtype=np.int32)
    w = np.zeros(Q)
    lib().orc_tables(_dp(d), _ip(opp), _dp(w))
    return d, opp, w


def mrt_matrix():
    m = np.zeros((Q, Q))
    lib().orc_mrt_matrix(_dp(m))
    return m


def mrt_inverse_exact(m):
    m = np.ascontiguousarray(m, dtype=np.float64)
    out = np.zeros((Q, Q))
    rc = lib().orc_mrt_inverse_exact(_dp(m), _dp(out))
    return rc, out


def decomposition(nx, ny, nz, xd, yd, zd, max_tiles=4096):
    s = np.zeros((max_tiles, 3, 2), dtype=np.int32)
    nt = lib().orc_decomposition(nx, ny, nz, xd, yd, zd, _ip(s), max_tiles)
    if nt < 0:
        raise ValueError("bad decomposition")
    return s[:nt].copy()


def slice_chunk(tile, slices):
    s = np.ascontiguousarray(slices, dtype=np.int32)
    out = np.zeros(10, dtype=np.int32)
    rc = lib().orc_slice_chunk(tile, _ip(s), _ip(out))
    return rc, out


def set_move_decomp(slices):
    s = np.ascontiguousarray(slices, dtype=np.int32)
    mv = np.zeros((len(s), Q), dtype=np.int32)
    lib().orc_set_move_decomp(len(s), _ip(s), _ip(mv))
    return mv


def inlet_profile(p):
    out = np.zeros((p.ny, p.nz))
    lib().orc_inlet_profile(ctypes.byref(p), _dp(out))
    return out


def cell_feq(rho, u):
    d, _, w = tables()
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(Q)
    lib().orc_cell_feq(_dp(d), _dp(w), rho, _dp(u), _dp(out))
    return out


def cell_tau_eddy(p, f, rho, u):
    f = np.ascontiguousarray(f, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_cell_tau_eddy(ctypes.byref(p), _dp(f), rho, _dp(u))


def cell_collide(p, f, rho, u, tau_eddy, mrt=True):
    f = np.ascontiguousarray(f, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(Q)
    fn = lib().orc_cell_mrt if mrt else lib().orc_cell_srt
    fn(ctypes.byref(p), _dp(f), rho, _dp(u), tau_eddy, _dp(out))
    return out


class Lattice:
    """One oracle lattice (orc_create .. orc_destroy); fp32=True runs the
    single-precision build of the same algorithm."""

    def __init__(self, p, slices=None, fp32=False):
        if slices is None:
            slices = decomposition(p.nx, p.ny, p.nz, 1, 1, 1)
        self.p = p
        self.L = lib(fp32)
        self.slices = np.ascontiguousarray(slices, dtype=np.int32)
        self.h = self.L.orc_create(ctypes.byref(p), len(self.slices), _ip(self.slices))
        if not self.h:
            raise ValueError("orc_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    @property
    def shape(self):
        return (self.p.ny, self.p.nz, self.p.nx)

    def initialize(self):
        self.L.orc_initialize(self.h)

    def set_mask(self, mask):
        """Solid nodes [ny][nz][nx] (None: the case's own geometry); before initialize()."""
        if mask is None:
            self.L.orc_set_mask(self.h, None)
            return
        m = np.ascontiguousarray(np.asarray(mask) != 0, dtype=np.uint8)
        assert m.shape == self.shape
        self.L.orc_set_mask(self.h, m.ctypes.data)

    def run(self, n):
        self.L.orc_run(self.h, n)

    def steps(self):
        return self.L.orc_steps(self.h)

    def read_arF(self):
        p = self.p
        out = np.zeros((p.ny + 2, p.nz + 2, p.nx + 2, Q))
        self.L.orc_read_arF_padded(self.h, _dp(out))
        return out

    def write_arF(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.p.ny + 2, self.p.nz + 2, self.p.nx + 2, Q)
        self.L.orc_write_arF_padded(self.h, _dp(a))

    def read_macro(self):
        s = self.shape
        r, x, y, z = (np.zeros(s) for _ in range(4))
        self.L.orc_read_macro(self.h, _dp(r), _dp(x), _dp(y), _dp(z))
        return r, x, y, z

    def node_type(self):
        t = np.zeros(self.shape, dtype=np.int32)
        d = np.zeros(self.shape, dtype=np.int32)
        self.L.orc_read_node_type(self.h, _ip(t), _ip(d))
        return t, d

    def diagnostics(self):
        m = np.zeros(1)
        mom = np.zeros(3)
        d = np.zeros(1)
        self.L.orc_diagnostics(self.h, _dp(m), _dp(mom), _dp(d))
        return float(m[0]), mom, float(d[0])
