"""ctypes binding of include/nflow.h (lib/libnflow.so).

The solver itself is the C++/HIP library; this module only marshals numpy
arrays across the C ABI for the tests, bench.py and Python callers.  It never
computes anything itself and raises if the native library is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnflow.so")
Q = 19
PIPE, CAVITY, CHANNEL = 0, 1, 2  # POISEUILLE = 0 and CAVITY = 1 as in common.h:26-27
F32, F16, BF16 = 0, 1, 2
NODE_FLUID, NODE_FIXEDWALL, NODE_MOVINGWALL = 1 << 1, 1 << 2, 1 << 3
NODE_INLET, NODE_OUTLET, NODE_STRUCT = 1 << 4, 1 << 5, 1 << 6
ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "ESTATE", -5: "EBUDGET", -6: "ENOTFOUND", -7: "EDIVERGED"}
EDIVERGED = -7


class NflowError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Input(ctypes.Structure):
    _fields_ = [
        ("h", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("maxitr", ctypes.c_int), ("nwrite", ctypes.c_int), ("casetype", ctypes.c_int),
        ("enable_mrt", ctypes.c_int), ("enable_les", ctypes.c_int), ("storage", ctypes.c_int),
        ("init_lid_velocity", ctypes.c_int),
        ("tau", ctypes.c_double), ("smagorinsky", ctypes.c_double), ("ref_den", ctypes.c_double),
        ("ref_vel", ctypes.c_double * 3),
        ("w_eps", ctypes.c_double), ("w_epsj", ctypes.c_double), ("w_xx", ctypes.c_double),
        ("inlet_peak", ctypes.c_double), ("perturb", ctypes.c_double), ("seed", ctypes.c_ulonglong),
        ("pipe_radius", ctypes.c_double), ("pipe_cy", ctypes.c_double), ("pipe_cz", ctypes.c_double),
        ("stochastic_rounding", ctypes.c_int),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int), ("xcd_order", ctypes.c_int), ("persistent", ctypes.c_int),
                ("graph_steps", ctypes.c_int), ("temporal", ctypes.c_int), ("ms_per_step", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Request(ctypes.Structure):
    _fields_ = [("input", Input), ("nsteps", ctypes.c_int), ("want_fields", ctypes.c_int),
                ("tag", ctypes.c_ulonglong)]


class Result(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_ulonglong), ("status", ctypes.c_int),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("steps", ctypes.c_longlong),
                ("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("mass", ctypes.c_double), ("mom", ctypes.c_double * 3), ("drho_l1", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_longlong), ("admitted", ctypes.c_longlong),
                ("completed", ctypes.c_longlong), ("rejected", ctypes.c_longlong),
                ("node_updates", ctypes.c_longlong), ("launches", ctypes.c_longlong),
                ("busy_s", ctypes.c_double), ("peak_bytes", ctypes.c_size_t), ("peak_resident", ctypes.c_int),
                ("ticks", ctypes.c_longlong), ("step_s", ctypes.c_double)]


# every function declared in include/nflow.h with its ctypes signature
_I, _D, _VP, _SZ, _LL = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_longlong
_IP, _FP, _DP = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "nflow_input_defaults": (None, [ctypes.POINTER(Input)]),
    "nflow_decomposition": (_I, [ctypes.POINTER(Input), _I, _I, _I, _IP, _I]),
    "nflow_slice_chunk": (_I, [ctypes.c_uint, _IP] + [_IP] * 10),
    "nflow_set_move_decomp": (_I, [_I, _IP, _IP]),
    "nflow_mrt_matrix": (None, [_FP]),
    "nflow_inverse_mrt": (_I, [_FP, _FP]),
    "nflow_create": (_I, [ctypes.POINTER(Input), _I, _IP, ctypes.POINTER(_VP)]),
    "nflow_initialize": (_I, [_VP]),
    "nflow_run": (_I, [_VP, _I]),
    "nflow_synchronize": (_I, [_VP]),
    "nflow_steps_done": (_LL, [_VP]),
    "nflow_read_arF": (_I, [_VP, _FP]),
    "nflow_write_arF": (_I, [_VP, _FP]),
    "nflow_read_macro": (_I, [_VP, _FP, _FP, _FP, _FP]),
    "nflow_read_node_type": (_I, [_VP, _IP, _IP]),
    "nflow_set_solid_mask": (_I, [_VP, ctypes.c_void_p]),
    "nflow_diagnostics": (_I, [_VP, _DP, _DP, _DP]),
    "nflow_diagnostics_noalloc": (_I, [_VP, _DP, _DP, _DP]),
    "nflow_timing": (_I, [_VP, _DP, _IP]),
    "nflow_destroy": (None, [_VP]),
    "nflow_last_error": (ctypes.c_char_p, []),
    "nflow_save_checkpoint": (_I, [_VP, ctypes.c_char_p]),
    "nflow_load_checkpoint": (_I, [_VP, ctypes.c_char_p]),
    "nflow_checkpoint_info": (_I, [ctypes.c_char_p, _IP, _IP, _IP, ctypes.POINTER(_LL)]),
    "nflow_write_vtu": (_I, [ctypes.c_char_p, _I, _I, _I, _FP, _FP, _FP, _FP, _IP]),
    "nflow_write_csv": (_I, [ctypes.c_char_p, _I, _I, _I, _FP, _FP, _FP, _FP]),
    "nflow_batch_create": (_I, [_I, ctypes.POINTER(Input), ctypes.POINTER(_VP)]),
    "nflow_batch_initialize": (_I, [_VP]),
    "nflow_batch_autotune": (_I, [_VP, _I, ctypes.POINTER(Tuning)]),
    "nflow_batch_set_tuning": (_I, [_VP, ctypes.POINTER(Tuning)]),
    "nflow_batch_get_tuning": (_I, [_VP, ctypes.POINTER(Tuning)]),
    "nflow_batch_step2_height": (_I, [_VP]),
    "nflow_set_tuning": (_I, [_VP, ctypes.POINTER(Tuning)]),
    "nflow_batch_run": (_I, [_VP, _I]),
    "nflow_batch_synchronize": (_I, [_VP]),
    "nflow_batch_read_arF": (_I, [_VP, _I, _FP]),
    "nflow_batch_read_macro": (_I, [_VP, _I, _FP, _FP, _FP, _FP]),
    "nflow_batch_diagnostics": (_I, [_VP, _I, _DP, _DP, _DP]),
    "nflow_batch_timing": (_I, [_VP, _DP, _IP]),
    "nflow_batch_diverged": (_I, [_VP, _IP]),
    "nflow_batch_nodes": (_LL, [_VP]),
    "nflow_batch_device_bytes": (_SZ, [_VP]),
    "nflow_batch_destroy": (None, [_VP]),
    "nflow_server_create": (_I, [_SZ, _I, _I, _I, ctypes.POINTER(_VP)]),
    "nflow_server_set_tau_floor": (_I, [_VP, ctypes.c_double]),
    "nflow_server_submit": (_I, [_VP, ctypes.POINTER(Request)]),
    "nflow_server_poll": (_I, [_VP, ctypes.POINTER(Result), _I, _I]),
    "nflow_server_fetch_macro": (_I, [_VP, ctypes.c_ulonglong, _FP, _FP, _FP, _FP]),
    "nflow_server_drain": (_I, [_VP]),
    "nflow_server_stats_get": (_I, [_VP, ctypes.POINTER(ServerStats)]),
    "nflow_server_destroy": (None, [_VP]),
    "nflow_copy_bandwidth": (_I, [_SZ, _I, _I, _DP, _DP]),
    "nflow_valu_rate": (_I, [_I, _DP, _DP]),
    "nflow_round16": (_I, [_FP, ctypes.POINTER(ctypes.c_uint), _SZ, _I, _I, ctypes.POINTER(ctypes.c_ushort)]),
    "nflow_device_info": (_I, [ctypes.c_char_p, _I, _IP, ctypes.POINTER(ctypes.c_size_t)]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise NflowError(rc, lib().nflow_last_error().decode())
    return rc


def _fp(a):
    return a.ctypes.data_as(_FP)


def _ip(a):
    return a.ctypes.data_as(_IP)


def defaults(**kw):
    """nflow_input_defaults() with keyword overrides (ref_vel accepts a 3-sequence)."""
    p = Input()
    lib().nflow_input_defaults(ctypes.byref(p))
    for k, v in kw.items():
        if k == "ref_vel":
            for i, x in enumerate(v):
                p.ref_vel[i] = x
        else:
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
    if "h" not in kw:
        p.h = p.nx
    return p


def decomposition(inp, xd, yd, zd, max_tiles=None):
    max_tiles = max_tiles or xd * yd * zd
    s = np.zeros((max_tiles, 3, 2), dtype=np.int32)
    nt = _check(lib().nflow_decomposition(ctypes.byref(inp), xd, yd, zd, _ip(s), max_tiles))
    return s[:nt].copy()


def slice_chunk(tile, slices):
    s = np.ascontiguousarray(slices, dtype=np.int32)
    outs = [ctypes.c_int() for _ in range(10)]
    rc = lib().nflow_slice_chunk(tile, _ip(s), *[ctypes.byref(o) for o in outs])
    return rc, np.array([o.value for o in outs], dtype=np.int32)


def set_move_decomp(slices):
    s = np.ascontiguousarray(slices, dtype=np.int32)
    mv = np.zeros((len(s), Q), dtype=np.int32)
    _check(lib().nflow_set_move_decomp(len(s), _ip(s), _ip(mv)))
    return mv


def mrt_matrix():
    m = np.zeros((Q, Q), dtype=np.float32)
    lib().nflow_mrt_matrix(_fp(m))
    return m


def inverse_mrt(m):
    m = np.ascontiguousarray(m, dtype=np.float32)
    out = np.zeros((Q, Q), dtype=np.float32)
    _check(lib().nflow_inverse_mrt(_fp(m), _fp(out)))
    return out


class Solver:
    """One lattice: nflow_create .. nflow_destroy."""

    def __init__(self, inp, slices=None):
        self.inp = inp
        h = _VP()
        if slices is None:
            _check(lib().nflow_create(ctypes.byref(inp), 0, None, ctypes.byref(h)))
        else:
            s = np.ascontiguousarray(slices, dtype=np.int32)
            _check(lib().nflow_create(ctypes.byref(inp), len(s), _ip(s), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().nflow_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def shape(self):
        return (self.inp.ny, self.inp.nz, self.inp.nx)

    def set_tuning(self, block=256, xcd_order=0, persistent=0, graph_steps=10, temporal=-1):
        t = Tuning(block, xcd_order, persistent, graph_steps, temporal, 0.0)
        _check(lib().nflow_set_tuning(self.h, ctypes.byref(t)))

    def initialize(self):
        _check(lib().nflow_initialize(self.h))

    def run(self, n):
        _check(lib().nflow_run(self.h, n))

    def synchronize(self):
        _check(lib().nflow_synchronize(self.h))

    def steps_done(self):
        return lib().nflow_steps_done(self.h)

    def read_arF(self):
        i = self.inp
        out = np.empty((i.ny + 2, i.nz + 2, i.nx + 2, Q), dtype=np.float32)
        _check(lib().nflow_read_arF(self.h, _fp(out)))
        return out

    def write_arF(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        i = self.inp
        if a.shape != (i.ny + 2, i.nz + 2, i.nx + 2, Q):
            raise ValueError(f"arF must have shape {(i.ny + 2, i.nz + 2, i.nx + 2, Q)}")
        _check(lib().nflow_write_arF(self.h, _fp(a)))

    def read_macro(self):
        r, x, y, z = (np.empty(self.shape, dtype=np.float32) for _ in range(4))
        _check(lib().nflow_read_macro(self.h, _fp(r), _fp(x), _fp(y), _fp(z)))
        return r, x, y, z

    def node_type(self):
        t = np.empty(self.shape, dtype=np.int32)
        d = np.empty(self.shape, dtype=np.int32)
        _check(lib().nflow_read_node_type(self.h, _ip(t), _ip(d)))
        return t, d

    def set_solid_mask(self, mask):
        """Solid nodes, [ny][nz][nx] booleans (None: the case's own geometry).
        Call before initialize()."""
        if mask is None:
            _check(lib().nflow_set_solid_mask(self.h, None))
            return
        m = np.ascontiguousarray(np.asarray(mask) != 0, dtype=np.uint8)
        if m.shape != self.shape:
            raise ValueError(f"mask must have shape {self.shape}")
        _check(lib().nflow_set_solid_mask(self.h, m.ctypes.data))

    def diagnostics(self):
        m, d = ctypes.c_double(), ctypes.c_double()
        mom = (ctypes.c_double * 3)()
        _check(lib().nflow_diagnostics(self.h, ctypes.byref(m), mom, ctypes.byref(d)))
        return m.value, np.array(list(mom)), d.value

    def diagnostics_noalloc(self):
        """diagnostics() through nflow_diagnostics_noalloc (static device scratch)."""
        m, d = ctypes.c_double(), ctypes.c_double()
        mom = (ctypes.c_double * 3)()
        _check(lib().nflow_diagnostics_noalloc(self.h, ctypes.byref(m), mom, ctypes.byref(d)))
        return m.value, np.array(list(mom)), d.value

    def timing(self):
        ms, n = ctypes.c_double(), ctypes.c_int()
        _check(lib().nflow_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def save_checkpoint(self, path):
        _check(lib().nflow_save_checkpoint(self.h, os.fsencode(path)))

    def load_checkpoint(self, path):
        _check(lib().nflow_load_checkpoint(self.h, os.fsencode(path)))


def checkpoint_info(path):
    nx, ny, nz, steps = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
    _check(lib().nflow_checkpoint_info(os.fsencode(path), ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz),
                                       ctypes.byref(steps)))
    return nx.value, ny.value, nz.value, steps.value


def write_vtu(path, rho, ux, uy, uz, node_type=None):
    ny, nz, nx = rho.shape
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (rho, ux, uy, uz)]
    t = None if node_type is None else np.ascontiguousarray(node_type, dtype=np.int32)
    _check(lib().nflow_write_vtu(os.fsencode(path), nx, ny, nz, *[_fp(a) for a in arrs],
                                 None if t is None else _ip(t)))


def write_csv(path, rho, ux, uy, uz):
    ny, nz, nx = rho.shape
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (rho, ux, uy, uz)]
    _check(lib().nflow_write_csv(os.fsencode(path), nx, ny, nz, *[_fp(a) for a in arrs]))


class Batch:
    """Many independent lattices advanced together (nflow_batch_*)."""

    def __init__(self, inputs):
        self.inputs = list(inputs)
        arr = (Input * len(self.inputs))(*self.inputs)
        h = _VP()
        _check(lib().nflow_batch_create(len(self.inputs), arr, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().nflow_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def initialize(self):
        _check(lib().nflow_batch_initialize(self.h))

    def autotune(self, steps_per_trial=10):
        t = Tuning()
        _check(lib().nflow_batch_autotune(self.h, steps_per_trial, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, block=256, xcd_order=0, persistent=0, graph_steps=10, temporal=-1):
        t = Tuning(block, xcd_order, persistent, graph_steps, temporal, 0.0)
        _check(lib().nflow_batch_set_tuning(self.h, ctypes.byref(t)))

    def get_tuning(self):
        t = Tuning()
        _check(lib().nflow_batch_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def step2_height(self):
        """y-chunk height of the two-step launches its runs use (0: one step per launch)."""
        return _check(lib().nflow_batch_step2_height(self.h))

    def run(self, n):
        _check(lib().nflow_batch_run(self.h, n))

    def synchronize(self):
        _check(lib().nflow_batch_synchronize(self.h))

    def read_macro(self, idx):
        i = self.inputs[idx]
        shape = (i.ny, i.nz, i.nx)
        r, x, y, z = (np.empty(shape, dtype=np.float32) for _ in range(4))
        _check(lib().nflow_batch_read_macro(self.h, idx, _fp(r), _fp(x), _fp(y), _fp(z)))
        return r, x, y, z

    def read_arF(self, idx):
        i = self.inputs[idx]
        out = np.empty((i.ny + 2, i.nz + 2, i.nx + 2, Q), dtype=np.float32)
        _check(lib().nflow_batch_read_arF(self.h, idx, _fp(out)))
        return out

    def diagnostics(self, idx):
        m, d = ctypes.c_double(), ctypes.c_double()
        mom = (ctypes.c_double * 3)()
        _check(lib().nflow_batch_diagnostics(self.h, idx, ctypes.byref(m), mom, ctypes.byref(d)))
        return m.value, np.array(list(mom)), d.value

    def diverged(self):
        """Indices of the lattices that went unstable since initialisation."""
        f = np.zeros(len(self.inputs), dtype=np.int32)
        _check(lib().nflow_batch_diverged(self.h, _ip(f)))
        return [int(i) for i in np.nonzero(f)[0]]

    def timing(self):
        ms, n = ctypes.c_double(), ctypes.c_int()
        _check(lib().nflow_batch_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def nodes(self):
        return lib().nflow_batch_nodes(self.h)

    def device_bytes(self):
        return lib().nflow_batch_device_bytes(self.h)


class Server:
    """Continuous-batching request loop (nflow_server_*)."""

    def __init__(self, budget_bytes, storage=F32, max_resident=64, steps_per_tick=100):
        h = _VP()
        _check(lib().nflow_server_create(int(budget_bytes), storage, max_resident, steps_per_tick, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().nflow_server_destroy(self.h)
            self.h = None

    __del__ = close

    def set_tau_floor(self, tau_min):
        """Requests with tau below tau_min are kept in FP32 storage (0: off)."""
        _check(lib().nflow_server_set_tau_floor(self.h, float(tau_min)))

    def submit(self, inp, nsteps, tag, want_fields=False):
        r = Request(inp, nsteps, 1 if want_fields else 0, tag)
        _check(lib().nflow_server_submit(self.h, ctypes.byref(r)))

    def poll(self, max_results=64, wait_ms=0):
        buf = (Result * max_results)()
        n = _check(lib().nflow_server_poll(self.h, buf, max_results, wait_ms))
        return [buf[i] for i in range(n)]

    def fetch_macro(self, tag, shape):
        r, x, y, z = (np.empty(shape, dtype=np.float32) for _ in range(4))
        _check(lib().nflow_server_fetch_macro(self.h, tag, _fp(r), _fp(x), _fp(y), _fp(z)))
        return r, x, y, z

    def drain(self):
        _check(lib().nflow_server_drain(self.h))

    def stats(self):
        s = ServerStats()
        _check(lib().nflow_server_stats_get(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}


def copy_bandwidth(nbytes, streams=19, reps=10):
    g, ms = ctypes.c_double(), ctypes.c_double()
    _check(lib().nflow_copy_bandwidth(int(nbytes), streams, reps, ctypes.byref(g), ctypes.byref(ms)))
    return g.value, ms.value


def valu_rate(reps=5):
    """Wave-level FP32 FMA instructions per second (1e9) of independent FMA
    chains, and the milliseconds per probe launch."""
    g, ms = ctypes.c_double(), ctypes.c_double()
    _check(lib().nflow_valu_rate(reps, ctypes.byref(g), ctypes.byref(ms)))
    return g.value, ms.value


def round16(values, storage, seeds=None):
    """The step kernel's 16-bit store conversion of float32 `values`: to
    nearest even, or stochastically with uint32 `seeds` (one per value).
    Returns the uint16 bit patterns."""
    import numpy as np

    v = np.ascontiguousarray(values, dtype=np.float32)
    out = np.empty(v.shape, dtype=np.uint16)
    sp = None
    if seeds is not None:
        sd = np.ascontiguousarray(seeds, dtype=np.uint32)
        if sd.shape != v.shape:
            raise ValueError("one seed per value")
        sp = sd.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))
    _check(lib().nflow_round16(_fp(v), sp, v.size, storage, 0 if seeds is None else 1,
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_ushort))))
    return out


def device_info():
    name = ctypes.create_string_buffer(256)
    cu = ctypes.c_int()
    mem = ctypes.c_size_t()
    _check(lib().nflow_device_info(name, 256, ctypes.byref(cu), ctypes.byref(mem)))
    return name.value.decode(), cu.value, mem.value
